// host_sanitize.cpp — drives the host-only entry points of the C ABI under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_host_sanitizers.py with g++; no device is touched):
//   * the palette PNG reader on every truncation of each input file, on every
//     single-byte corruption with the chunk CRCs recomputed (so the damage
//     reaches IHDR validation, the filters and inflate), and on IHDR headers of
//     every colour type, depth and interlace with extreme sizes;
//   * PNG output round trips of odd sizes;
//   * palettes, plane mapping, the precision switch, launch geometry and tile
//     set-up with edge arguments.
// Exit status 0 and no sanitizer report is a pass.
//   usage: host_sanitize <png>...
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "chaosultra.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s (%s)\n", what, cu_last_error());
        ++failures;
    }
}

std::vector<uint8_t> read_all(const char *path) {
    std::vector<uint8_t> b;
    FILE *f = std::fopen(path, "rb");
    if (!f) return b;
    int c;
    while ((c = std::fgetc(f)) != EOF) b.push_back((uint8_t)c);
    std::fclose(f);
    return b;
}

void write_all(const std::string &path, const std::vector<uint8_t> &b) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

// Recompute every chunk CRC that lies inside the buffer.
void fix_crcs(std::vector<uint8_t> &b) {
    size_t i = 8;
    while (i + 12 <= b.size()) {
        const uint32_t n = (uint32_t)b[i] << 24 | (uint32_t)b[i + 1] << 16 | (uint32_t)b[i + 2] << 8 | b[i + 3];
        if (n > b.size() - i - 12) return;
        const uint32_t c = (uint32_t)crc32(crc32(0L, Z_NULL, 0), &b[i + 4], n + 4);
        b[i + 8 + n] = (uint8_t)(c >> 24);
        b[i + 9 + n] = (uint8_t)(c >> 16);
        b[i + 10 + n] = (uint8_t)(c >> 8);
        b[i + 11 + n] = (uint8_t)c;
        i += 12 + n;
    }
}

// Load `bytes` as a palette; the result must be the file's first row or the default.
void load(const std::string &tmp, const std::vector<uint8_t> &bytes) {
    write_all(tmp, bytes);
    static std::vector<uint32_t> out(1 << 16);
    size_t n = 0;
    int loaded = -1;
    const int rc = cu_palette_load_or_default(tmp.c_str(), out.data(), out.size(), &n, &loaded);
    expect(rc == CU_OK || rc == CU_ERR_ILLEGAL_ARGUMENT, "load status");
    if (rc == CU_OK) expect(n > 0 && n <= out.size() && (loaded == 0 || loaded == 1), "load result");
    if (rc == CU_OK && loaded == 0) expect(n == 1536, "default palette size");
}

std::vector<uint8_t> chunk(const char *type, const std::vector<uint8_t> &data) {
    std::vector<uint8_t> c = {(uint8_t)(data.size() >> 24), (uint8_t)(data.size() >> 16),
                              (uint8_t)(data.size() >> 8), (uint8_t)data.size()};
    c.insert(c.end(), type, type + 4);
    c.insert(c.end(), data.begin(), data.end());
    c.insert(c.end(), 4, 0);
    return c;
}

}  // namespace

int main(int argc, char **argv) {
    const std::string tmp = "host_sanitize_tmp.png";
    std::mt19937 rng(7);
    // ---- PNG reader on damaged inputs
    for (int a = 1; a < argc; ++a) {
        const std::vector<uint8_t> good = read_all(argv[a]);
        expect(!good.empty(), "read fixture");
        for (size_t n = 0; n <= good.size(); ++n)
            load(tmp, std::vector<uint8_t>(good.begin(), good.begin() + n));
        for (size_t i = 8; i < good.size(); ++i) {
            std::vector<uint8_t> b = good;
            b[i] ^= (uint8_t)(1u + rng() % 255u);
            fix_crcs(b);
            load(tmp, b);
        }
    }
    // ---- IHDR of every colour type / depth / interlace with edge sizes
    const uint32_t sizes[] = {1, 2, 7, 8, 9, 1000, 1u << 16, 1u << 24, (1u << 24) + 1, 0x7fffffffu,
                              0xffffffffu, 0};
    const uint8_t depths[] = {1, 2, 4, 8, 16, 3, 0};
    const uint8_t ctypes[] = {0, 2, 3, 4, 6, 1, 7};
    for (uint32_t w : sizes)
        for (uint32_t h : {1u, 3u, 0x7fffffffu})
            for (uint8_t d : depths)
                for (uint8_t ct : ctypes)
                    for (uint8_t il : {0, 1}) {
                        std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
                        const std::vector<uint8_t> ihdr = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8),
                                                           (uint8_t)w, (uint8_t)(h >> 24), (uint8_t)(h >> 16),
                                                           (uint8_t)(h >> 8), (uint8_t)h, d, ct, 0, 0, il};
                        const auto c1 = chunk("IHDR", ihdr);
                        png.insert(png.end(), c1.begin(), c1.end());
                        if (ct == 3) {
                            const auto p = chunk("PLTE", std::vector<uint8_t>(9, 200));
                            png.insert(png.end(), p.begin(), p.end());
                            const auto t = chunk("tRNS", std::vector<uint8_t>(2, 17));
                            png.insert(png.end(), t.begin(), t.end());
                        }
                        std::vector<uint8_t> raw(1 + (size_t)std::min<uint32_t>(w, 64) * 8, 0x5a);
                        raw[0] = (uint8_t)(rng() % 6);  // filter types 0..4 and an invalid 5
                        uLongf zl = compressBound((uLong)raw.size());
                        std::vector<uint8_t> z(zl);
                        compress(z.data(), &zl, raw.data(), (uLong)raw.size());
                        z.resize(zl);
                        const auto c2 = chunk("IDAT", z);
                        png.insert(png.end(), c2.begin(), c2.end());
                        const auto c3 = chunk("IEND", {});
                        png.insert(png.end(), c3.begin(), c3.end());
                        fix_crcs(png);
                        load(tmp, png);
                    }
    // ---- PNG writer round trips
    for (int k = 0; k < 40; ++k) {
        const int w = 1 + (int)(rng() % 300), h = 1 + (int)(rng() % 7);
        std::vector<uint32_t> img((size_t)w * h);
        for (auto &v : img) v = rng();
        expect(cu_image_save_png(tmp.c_str(), img.data(), w, h) == CU_OK, "save png");
        std::vector<uint32_t> pal(1 << 16);
        size_t n = 0;
        int loaded = 0;
        expect(cu_palette_load_or_default(tmp.c_str(), pal.data(), pal.size(), &n, &loaded) == CU_OK && loaded &&
                   n == (size_t)w,
               "png round trip");
        for (int x = 0; x < w && x < (int)n; ++x)
            expect(pal[x] == ((img[x] & 0xffffffu) | 0xff000000u), "round-trip pixel");
    }
    expect(cu_image_save_png(tmp.c_str(), nullptr, 4, 4) == CU_ERR_ILLEGAL_ARGUMENT, "save null");
    expect(cu_image_save_png("/proc/no/such/dir/x.png", std::vector<uint32_t>(4).data(), 2, 2) ==
               CU_ERR_ILLEGAL_ARGUMENT,
           "save unwritable");
    std::remove(tmp.c_str());
    // ---- palettes, mapping, precision, geometry, tiles
    std::vector<uint32_t> pal(1536), rgba(1536);
    expect(cu_default_palette(pal.data(), pal.size()) == CU_OK, "default palette");
    expect(cu_default_palette(pal.data(), 10) == CU_ERR_ILLEGAL_ARGUMENT, "short palette");
    expect(cu_palette_from_argb(pal.data(), rgba.data(), pal.size()) == CU_OK, "argb");
    const double inf = std::numeric_limits<double>::infinity(), nan = std::nan("");
    double seg[4];
    for (double z : {1e-300, 1e-20, 1.0, 3.0, 1e300, inf, nan, 0.0, -1.0})
        for (int w : {1, 7, 4096, 0, -3})
            (void)cu_plane_segment_from_center(-0.5, 0.0, z, w, 100, seg);
    int prec = 0;
    for (double s : {1e-300, 1e-17, 1e-9, 1e-3, 3.0, inf, nan})
        (void)cu_precision_for_segment(-0.75, 0.1, -0.75 + s, 0.1 + s, 1920, 1080, &prec);
    unsigned gx = 0, gy = 0;
    for (int w : {0, 1, 31, 32, 33, 65535 * 32, 65536 * 32, -1, std::numeric_limits<int>::max()})
        (void)cu_launch_geometry(w, 100, &gx, &gy);
    cu_tile_desc t;
    int safe = 0;
    for (double s : {1e-300, 1e-10, 1.0, inf, nan})
        for (uint32_t w : {0u, 1u, 4096u, 0xffffffffu})
            for (uint32_t spp : {0u, 1u, 4u, 64u, 65u})
                if (cu_tile_init(&t, -2.0, -1.5, -2.0 + s, -1.5 + s, w, 64, 1000, 4.0, CU_FRACTAL_MANDELBROT,
                                 0.0, 0.0, spp, CU_SAMPLE_JITTER, 9) == CU_OK)
                    (void)cu_tile_fused_safe(&t, 64, &safe);
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("host_sanitize: ok\n");
    return 0;
}
