// png_io.cpp — the reference's on-disk image formats (util/ImageHelpers.java):
// palette loading from the first row of a PNG with the default-palette fallback,
// and saving an RGBA image as a PNG.  The PNG chunk/filter layer is written
// here against the PNG specification (ISO/IEC 15948); zlib does the deflate.
//
// Decoded: colour types 0 (grey), 2 (RGB), 3 (palette, with tRNS), 4 (grey +
// alpha) and 6 (RGBA) at every bit depth the specification allows, plain or
// Adam7-interlaced.  Only the first row is reconstructed — that is all
// loadColorPaletteFromFile reads (image.getRGB(0, 0, w, 1, ...)).
#include <sys/stat.h>
#include <zlib.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace cu {
namespace {

uint32_t be32(const uint8_t *p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
    std::fclose(f);
    return true;
}

// Bounds on what a palette file may ask the reader to allocate: the first row
// holds at most kMaxWidth entries, and an interlaced file is inflated up to its
// sixth pass (where row 0 ends) only while that stays below kMaxInflate bytes.
constexpr uint32_t kMaxWidth = 1u << 24;
constexpr size_t kMaxInflate = size_t(1) << 30;

struct Png {
    uint32_t width = 0, height = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
};

int channels(int ctype) {
    switch (ctype) {
        case 0: return 1;
        case 2: return 3;
        case 3: return 1;
        case 4: return 2;
        case 6: return 4;
        default: return 0;
    }
}

bool depth_allowed(int ctype, int depth) {
    switch (ctype) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

// Chunk walk; false on a structural error or a chunk whose CRC does not match
// (a check ImageIO's reader does not make: such a file falls back here).
bool parse(const std::vector<uint8_t> &b, Png &png) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (b.size() < 8 || std::memcmp(b.data(), sig, 8) != 0) return false;
    size_t i = 8;
    bool ihdr = false, iend = false;
    while (i + 12 <= b.size() && !iend) {
        const uint32_t n = be32(&b[i]);
        if (n > b.size() - i - 12) return false;
        const uint8_t *type = &b[i + 4], *data = &b[i + 8];
        const uint32_t crc = be32(&b[i + 8 + n]);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, n + 4) != crc) return false;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (n != 13 || ihdr) return false;
            png.width = be32(data);
            png.height = be32(data + 4);
            png.depth = data[8];
            png.ctype = data[9];
            png.interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || png.interlace > 1) return false;
            ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            png.plte.assign(data, data + n);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            png.trns.assign(data, data + n);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            png.idat.insert(png.idat.end(), data, data + n);
        } else if (!std::memcmp(type, "IEND", 4)) {
            iend = true;
        } else if (!(type[0] & 0x20)) {
            return false;  // unknown critical chunk
        }
        i += 12 + n;
    }
    // like ImageIO's reader, a stream that ends after its image data needs no IEND
    if (!ihdr || png.idat.empty() || png.width == 0 || png.height == 0) return false;
    if (png.width > kMaxWidth) return false;  // no palette is this wide; refuse the allocation
    if (!depth_allowed(png.ctype, png.depth)) return false;
    if (png.ctype == 3 && (png.plte.empty() || png.plte.size() % 3)) return false;
    return true;
}

bool inflate_all(const std::vector<uint8_t> &in, std::vector<uint8_t> &out, size_t limit) {
    z_stream z;
    std::memset(&z, 0, sizeof z);
    if (inflateInit(&z) != Z_OK) return false;
    z.next_in = const_cast<Bytef *>(in.data());
    z.avail_in = (uInt)in.size();
    out.resize(limit);
    z.next_out = out.data();
    z.avail_out = (uInt)limit;
    const int rc = inflate(&z, Z_FINISH);
    const size_t got = limit - z.avail_out;
    inflateEnd(&z);
    // the first row may be complete before the stream ends (we stop at `limit`)
    if (rc != Z_STREAM_END && rc != Z_BUF_ERROR && rc != Z_OK) return false;
    out.resize(got);
    return true;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Reverse the filter of a pass's first scanline (the previous line is zero).
bool unfilter_first(const uint8_t *line, size_t bytes, size_t bpp, std::vector<uint8_t> &out) {
    const int ft = line[0];
    out.assign(line + 1, line + 1 + bytes);
    for (size_t k = 0; k < bytes; ++k) {
        const int a = k >= bpp ? out[k - bpp] : 0;  // left; up and up-left are 0
        switch (ft) {
            case 0: case 2: break;
            case 1: out[k] = (uint8_t)(out[k] + a); break;
            case 3: out[k] = (uint8_t)(out[k] + (a >> 1)); break;
            case 4: out[k] = (uint8_t)(out[k] + paeth(a, 0, 0)); break;
            default: return false;
        }
    }
    return true;
}

// ARGB (java.awt.image.BufferedImage.getRGB's default sRGB model) of sample x
// of an unfiltered scanline.  Grey is replicated and 16-bit samples keep their
// high byte.
uint32_t argb_of(const Png &png, const std::vector<uint8_t> &row, uint32_t x) {
    const int d = png.depth, ch = channels(png.ctype);
    auto sample = [&](int c) -> uint32_t {  // 8-bit value of channel c
        if (d == 16) return row[((size_t)x * ch + c) * 2];
        if (d == 8) return row[(size_t)x * ch + c];
        const size_t bit = (size_t)x * d;  // d < 8: one channel (grey or index)
        const uint32_t v = (row[bit >> 3] >> (8 - d - (bit & 7))) & ((1u << d) - 1u);
        return png.ctype == 3 ? v : v * 255u / ((1u << d) - 1u);
    };
    auto raw16 = [&](int c) -> uint32_t {
        return d == 16 ? (uint32_t)row[((size_t)x * ch + c) * 2] << 8 | row[((size_t)x * ch + c) * 2 + 1]
                       : 0u;
    };
    uint32_t r, g, b, a = 255;
    switch (png.ctype) {
        case 0: {
            r = g = b = sample(0);
            if (png.trns.size() >= 2) {
                const uint32_t key = (uint32_t)png.trns[0] << 8 | png.trns[1];
                const uint32_t v = d == 16 ? raw16(0)
                                   : d == 8 ? row[x]
                                            : ((row[((size_t)x * d) >> 3] >> (8 - d - (((size_t)x * d) & 7))) &
                                               ((1u << d) - 1u));
                if (v == key) a = 0;
            }
            break;
        }
        case 2:
            r = sample(0);
            g = sample(1);
            b = sample(2);
            if (png.trns.size() >= 6) {
                const uint32_t kr = (uint32_t)png.trns[0] << 8 | png.trns[1];
                const uint32_t kg = (uint32_t)png.trns[2] << 8 | png.trns[3];
                const uint32_t kb = (uint32_t)png.trns[4] << 8 | png.trns[5];
                const bool hit = d == 16 ? (raw16(0) == kr && raw16(1) == kg && raw16(2) == kb)
                                         : (r == kr && g == kg && b == kb);
                if (hit) a = 0;
            }
            break;
        case 3: {
            const uint32_t idx = sample(0);
            if ((size_t)idx * 3 + 2 >= png.plte.size()) return 0xff000000u;  // out of the palette: black
            r = png.plte[idx * 3];
            g = png.plte[idx * 3 + 1];
            b = png.plte[idx * 3 + 2];
            if (idx < png.trns.size()) a = png.trns[idx];
            break;
        }
        case 4: r = g = b = sample(0); a = sample(1); break;
        default: r = sample(0); g = sample(1); b = sample(2); a = sample(3); break;
    }
    return a << 24 | r << 16 | g << 8 | b;
}

// First image row as ARGB words; false if the file is not a PNG we can read.
bool first_row_argb(const std::string &path, std::vector<uint32_t> &argb) {
    std::vector<uint8_t> file;
    Png png;
    if (!read_file(path, file) || !parse(file, png)) return false;
    const int ch = channels(png.ctype);
    const size_t bits = (size_t)ch * png.depth, bpp = std::max<size_t>(1, bits / 8);
    auto line_bytes = [&](uint32_t w) { return (w * bits + 7) / 8; };
    argb.assign(png.width, 0u);
    if (!png.interlace) {
        const size_t need = 1 + line_bytes(png.width);
        std::vector<uint8_t> raw, row;
        if (!inflate_all(png.idat, raw, need) || raw.size() < need) return false;
        if (!unfilter_first(raw.data(), need - 1, bpp, row)) return false;
        for (uint32_t x = 0; x < png.width; ++x) argb[x] = argb_of(png, row, x);
        return true;
    }
    // Adam7: row 0 holds samples of passes 1 (x = 8k), 2 (8k + 4), 4 (4k + 2) and 6 (2k + 1).
    static const uint32_t x0[7] = {0, 4, 0, 2, 0, 1, 0}, dx[7] = {8, 8, 4, 4, 2, 2, 1};
    static const uint32_t y0[7] = {0, 0, 4, 0, 2, 0, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    size_t total = 0, offset[7];
    uint32_t pw[7];
    for (int p = 0; p < 7; ++p) {
        pw[p] = png.width > x0[p] ? (png.width - x0[p] + dx[p] - 1) / dx[p] : 0;
        const uint32_t ph = png.height > y0[p] ? (png.height - y0[p] + dy[p] - 1) / dy[p] : 0;
        offset[p] = total;
        if (pw[p] && ph) total += ph * (1 + line_bytes(pw[p]));
    }
    const size_t need = pw[5] ? offset[5] + 1 + line_bytes(pw[5])
                      : pw[3] ? offset[3] + 1 + line_bytes(pw[3])
                      : pw[1] ? offset[1] + 1 + line_bytes(pw[1])
                              : 1 + line_bytes(pw[0]);
    if (need > kMaxInflate) return false;
    std::vector<uint8_t> raw, row;
    if (!inflate_all(png.idat, raw, need)) return false;
    for (int p : {0, 1, 3, 5}) {
        if (!pw[p]) continue;
        if (raw.size() < offset[p] + 1 + line_bytes(pw[p])) return false;
        if (!unfilter_first(raw.data() + offset[p], line_bytes(pw[p]), bpp, row)) return false;
        for (uint32_t k = 0; k < pw[p]; ++k) argb[x0[p] + k * dx[p]] = argb_of(png, row, k);
    }
    return true;
}

void put32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
    put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), &out[start], (uInt)(data.size() + 4)));
}

// mkdir -p of the parent directory of `path` (Files.createDirectories).
void make_parents(const std::string &path) {
    for (size_t i = 1; i < path.size(); ++i) {
        if (path[i] != '/') continue;
        const std::string dir = path.substr(0, i);
        if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "cannot create directory " + dir);
    }
}

}  // namespace
}  // namespace cu

using namespace cu;

extern "C" {

int cu_palette_load_or_default(const char *path, uint32_t *out, size_t cap, size_t *count,
                               int *loaded) {
    return guarded([&] {
        if (!out || !count) fail(CU_ERR_ILLEGAL_ARGUMENT, "out/count is NULL");
        std::vector<uint32_t> argb;
        const bool ok = path && *path && first_row_argb(path, argb);
        if (loaded) *loaded = ok ? 1 : 0;
        if (!ok) {
            if (cap < 1536) fail(CU_ERR_ILLEGAL_ARGUMENT, "need 1536 entries for the default palette");
            *count = 1536;
            const int rc = cu_default_palette(out, 1536);
            if (rc != CU_OK) fail(rc, cu_last_error());
            return;
        }
        if (cap < argb.size())
            fail(CU_ERR_ILLEGAL_ARGUMENT, "palette has " + std::to_string(argb.size()) + " entries");
        const int rc = cu_palette_from_argb(argb.data(), out, argb.size());
        if (rc != CU_OK) fail(rc, cu_last_error());
        *count = argb.size();
    });
}

int cu_image_save_png(const char *path, const uint32_t *rgba, int width, int height) {
    return guarded([&] {
        if (!path || !*path || !rgba) fail(CU_ERR_ILLEGAL_ARGUMENT, "path/rgba is NULL");
        if (width <= 0 || height <= 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "image size must be positive");
        // 8-bit RGB scanlines, filter 0: R, G, B from the low three bytes (alpha dropped,
        // the reference's DirectColorModel(24, 0xff, 0xff00, 0xff0000))
        const size_t w = (size_t)width, h = (size_t)height;
        std::vector<uint8_t> raw(h * (1 + 3 * w));
        for (size_t y = 0; y < h; ++y) {
            uint8_t *line = &raw[y * (1 + 3 * w)];
            line[0] = 0;
            for (size_t x = 0; x < w; ++x) {
                const uint32_t v = rgba[y * w + x];
                line[1 + 3 * x] = (uint8_t)v;
                line[2 + 3 * x] = (uint8_t)(v >> 8);
                line[3 + 3 * x] = (uint8_t)(v >> 16);
            }
        }
        uLongf zlen = compressBound((uLong)raw.size());
        std::vector<uint8_t> z(zlen);
        if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "deflate failed");
        z.resize(zlen);
        std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        std::vector<uint8_t> ihdr;
        put32(ihdr, (uint32_t)width);
        put32(ihdr, (uint32_t)height);
        ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
        chunk(out, "IHDR", ihdr);
        chunk(out, "IDAT", z);
        chunk(out, "IEND", {});
        make_parents(path);
        FILE *f = std::fopen(path, "wb");
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, std::string("cannot open ") + path);
        const bool written = std::fwrite(out.data(), 1, out.size(), f) == out.size();
        if (std::fclose(f) != 0 || !written) fail(CU_ERR_ILLEGAL_ARGUMENT, std::string("cannot write ") + path);
    });
}

}  // extern "C"
