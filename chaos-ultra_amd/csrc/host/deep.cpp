// deep.cpp — the deep-zoom render form (cu_deep_*): a double-double reference
// orbit on the GPU, a perturbation (or direct FP64) pass over every pixel, and
// glitch rounds around secondary references.  Kernels in perturb.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine_helpers.hpp"
#include "handles.hpp"

namespace {

// Device-side argument block of perturb{Float,Double} / deepDirectDouble*.
struct DeepArgs {
    const void *orbit;
    const uint32_t *orbit_len;
    uint32_t *counts;
    double *smooth;
    unsigned char *glitch;
    uint32_t *rgba;
    const uint32_t *palette;
    unsigned long long *iters;
    double centre_x, centre_y;
    double pixel_w, pixel_h;
    double ref_ox, ref_oy;
    double bailout_sq;
    double smooth_c;
    uint32_t width, height, max_iter, flags;
    uint32_t palette_length;
    float glitch_eps;
    uint32_t only_glitched;
    uint32_t pad;
};
static_assert(sizeof(DeepArgs) == 160, "DeepArgs layout");

constexpr uint32_t kMaxReferences = 16;

// (a_hi + a_lo) + b in double-double (the same operation sequence as the
// kernels' dd_add with b_lo = 0).
void dd_add_double(double a_hi, double a_lo, double b, double &hi, double &lo) {
    auto two_sum = [](double x, double y, double &s, double &e) {
        s = x + y;
        const double bb = s - x;
        e = (x - (s - bb)) + (y - bb);
    };
    auto quick_two_sum = [](double x, double y, double &s, double &e) {
        s = x + y;
        e = y - (s - x);
    };
    double s, e, t, f;
    two_sum(a_hi, b, s, e);
    two_sum(a_lo, 0.0, t, f);
    e = e + t;
    quick_two_sum(s, e, s, e);
    e = e + f;
    quick_two_sum(s, e, hi, lo);
}

}  // namespace

struct cu_deep_s {
    cu_engine engine = nullptr;
    cu_deep_desc desc{};
    uint32_t capacity_iter = 0;  // max_iter the orbit buffers were sized for
    // primary orbit held in d_orbit[0]: its centre, bailout and iteration limit
    bool orbit_valid = false;
    double orbit_key[5] = {0, 0, 0, 0, 0};
    uint32_t orbit_mi = 0;
    int method = CU_DEEP_PERTURB_F64;
    bool fused = false;  // direct form: fused step variant
    size_t entry_bytes = 32;
    void *d_orbit[2] = {nullptr, nullptr};  // primary, secondary
    void *d_len = nullptr;                  // u32[2]
    void *d_counts = nullptr, *d_smooth = nullptr, *d_glitch = nullptr;
    void *d_iters = nullptr;                // CU_ITER_SLOTS u64
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t last_stream = nullptr;
    uint32_t references = 0, glitched = 0;
    bool rendered = false;
};

namespace cu {
namespace {

void free_deep(cu_deep d) {
    for (void *p : {d->d_orbit[0], d->d_orbit[1], d->d_len, d->d_counts, d->d_smooth, d->d_glitch,
                    d->d_iters})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : d->ev)
        if (e) (void)hipEventDestroy(e);
    delete d;
}

}  // namespace

void validate_deep(const cu_deep_desc &d, int method) {
    if (method != CU_DEEP_DIRECT_F64 && method != CU_DEEP_PERTURB_F32 &&
        method != CU_DEEP_PERTURB_F64)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "method must be CU_DEEP_DIRECT_F64, _PERTURB_F32 or _PERTURB_F64");
    if (d.width == 0 || d.height == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty frame");
    if ((uint64_t)d.width * d.height >= (1ull << 32)) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame too large");
    if (!(d.pixel_w > 0) || !(d.pixel_h > 0) || !std::isfinite(d.pixel_w) || !std::isfinite(d.pixel_h))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "pixel sizes must be positive and finite");
    if (!(d.bailout_sq >= 4.0) || !std::isfinite(d.bailout_sq))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "bailout_sq must be finite and >= 4");
    if (!std::isfinite(d.centre_x) || !std::isfinite(d.centre_y) || !std::isfinite(d.centre_x_lo) ||
        !std::isfinite(d.centre_y_lo))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "centre must be finite");
    // every sample point within |c| <= 2: escape stays permanent for the unchecked blocks
    const double hx = std::fabs(d.centre_x) + 0.5 * d.width * d.pixel_w;
    const double hy = std::fabs(d.centre_y) + 0.5 * d.height * d.pixel_h;
    if (!(hx * hx + hy * hy <= 4.0)) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame must lie within |c| <= 2");
    if (d.max_iter >= 0xFFFFFFF0u) fail(CU_ERR_ILLEGAL_ARGUMENT, "max_iter too large");
    if (d.max_references > kMaxReferences)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "max_references must be <= 16");
    if (!(d.flags & CU_DEEP_REBASE) && !(d.glitch_eps > 0.0f && d.glitch_eps < 1.0f))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "glitch_eps must be in (0, 1) without rebasing");
    if (d.flags & ~(CU_DEEP_REBASE | CU_DEEP_REUSE_ORBIT))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown flags");
}

namespace {

const char *orbit_kernel(const cu_deep_s *d) {
    return d->method == CU_DEEP_PERTURB_F32 ? "referenceOrbitFloat" : "referenceOrbitDouble";
}

void launch_orbit(cu_deep d, cu_stream s, int slot, double cx_hi, double cx_lo, double cy_hi,
                  double cy_lo, uint32_t mi) {
    cu_function f = function(d->engine, orbit_kernel(d));
    double bsq = d->desc.bailout_sq;
    void *orbit = d->d_orbit[slot];
    uint32_t *len = (uint32_t *)d->d_len + slot;
    void *params[] = {&cx_hi, &cx_lo, &cy_hi, &cy_lo, &mi, &bsq, &orbit, &len};
    launch(d->engine, f, dim3(1), dim3(64), params, s, false);
}

DeepArgs base_args(cu_deep d, const uint32_t *pal, uint32_t plen, uint32_t *rgba) {
    const cu_deep_desc &c = d->desc;
    DeepArgs a{};
    a.orbit = d->d_orbit[0];
    a.orbit_len = (const uint32_t *)d->d_len;
    a.counts = (uint32_t *)d->d_counts;
    a.smooth = (double *)d->d_smooth;
    a.glitch = (unsigned char *)d->d_glitch;
    a.rgba = rgba;
    a.palette = pal;
    a.iters = (unsigned long long *)d->d_iters;
    a.centre_x = c.centre_x;
    a.centre_y = c.centre_y;
    a.pixel_w = c.pixel_w;
    a.pixel_h = c.pixel_h;
    a.bailout_sq = c.bailout_sq;
    a.smooth_c = smooth_constant(c.bailout_sq);
    a.width = c.width;
    a.height = c.height;
    a.max_iter = c.max_iter;
    a.flags = c.flags;
    a.palette_length = pal ? plen : 0u;
    a.glitch_eps = c.glitch_eps;
    return a;
}

void launch_pixels(cu_deep d, cu_stream s, DeepArgs &a) {
    std::string name;
    if (d->method == CU_DEEP_DIRECT_F64)
        name = d->fused ? "deepDirectDoubleFused" : "deepDirectDoubleExact";
    else
        name = d->method == CU_DEEP_PERTURB_F32 ? "perturbFloat" : "perturbDouble";
    cu_function f = function(d->engine, name);
    const void *orbit = a.orbit;
    void *params[] = {&a, &orbit};  // direct kernels take only the first
    // direct and FP32 perturbation: 256 threads = 16 x 16 pixels; FP64
    // perturbation: 512 threads = 32 x 16 pixels (a perturbation block shares
    // one LDS copy of the reference orbit)
    const uint32_t bw = d->method == CU_DEEP_PERTURB_F64 ? 32 : 16, bh = 16;
    dim3 grid((d->desc.width + bw - 1) / bw, (d->desc.height + bh - 1) / bh, 1);
    launch(d->engine, f, grid, dim3(bw * bh), params, s, false);
}

// The glitched pixel nearest the centroid of all glitched pixels.
bool pick_secondary(const std::vector<unsigned char> &flags, uint32_t W, uint32_t H, uint32_t &px,
                    uint32_t &py, uint32_t &count) {
    double sx = 0, sy = 0;
    count = 0;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x)
            if (flags[(size_t)y * W + x]) {
                sx += x;
                sy += y;
                ++count;
            }
    if (count == 0) return false;
    const double cx = sx / count, cy = sy / count;
    double best = -1;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x)
            if (flags[(size_t)y * W + x]) {
                const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
                if (best < 0 || d2 < best) {
                    best = d2;
                    px = x;
                    py = y;
                }
            }
    return true;
}

}  // namespace
}  // namespace cu

using namespace cu;

extern "C" {

int cu_deep_create(cu_engine e, const cu_deep_desc *desc, int method, cu_deep *out) {
    cu_deep d = nullptr;
    int rc = guarded([&] {
        if (!e || !desc || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        validate_deep(*desc, method);
        d = new cu_deep_s();
        d->engine = e;
        d->desc = *desc;
        d->capacity_iter = desc->max_iter;
        d->method = method;
        d->entry_bytes = method == CU_DEEP_PERTURB_F32 ? 16 : 32;
        if (method == CU_DEEP_DIRECT_F64) {
            cu_tile_desc t{};
            t.left = desc->centre_x - 0.5 * desc->width * desc->pixel_w;
            t.top = desc->centre_y + 0.5 * desc->height * desc->pixel_h;
            t.pixel_w = desc->pixel_w;
            t.pixel_h = desc->pixel_h;
            t.kind = CU_FRACTAL_MANDELBROT;
            d->fused = fused_safe(t, 64) && std::fabs(desc->centre_x) >= std::ldexp(1.0, -300) &&
                       std::fabs(desc->centre_y) >= std::ldexp(1.0, -300);
        }
        const size_t pixels = (size_t)desc->width * desc->height;
        ensure_device();
        if (method != CU_DEEP_DIRECT_F64) {
            const size_t entries = (size_t)desc->max_iter + 1;
            d->d_orbit[0] = device_alloc(entries * d->entry_bytes);
            if (!(desc->flags & CU_DEEP_REBASE) && desc->max_references > 0)
                d->d_orbit[1] = device_alloc(entries * d->entry_bytes);
        }
        d->d_len = device_alloc(2 * sizeof(uint32_t));
        d->d_counts = device_alloc(pixels * 4);
        d->d_smooth = device_alloc(pixels * 8);
        d->d_glitch = device_alloc(pixels);
        d->d_iters = device_alloc(CU_ITER_SLOTS * 8);
        for (hipEvent_t &ev : d->ev) check(hipEventCreate(&ev), "hipEventCreate");
        *out = d;
    });
    if (rc != CU_OK && d) free_deep(d);
    return rc;
}

int cu_deep_retarget(cu_deep d, const cu_deep_desc *desc) {
    return guarded([&] {
        if (!d || !desc) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        validate_deep(*desc, d->method);
        if (desc->width != d->desc.width || desc->height != d->desc.height)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "retarget keeps the frame size");
        if (desc->max_iter > d->capacity_iter)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "max_iter above the capacity the frame was created with");
        const bool rounds_before = !(d->desc.flags & CU_DEEP_REBASE) && d->desc.max_references > 0;
        const bool rounds_after = !(desc->flags & CU_DEEP_REBASE) && desc->max_references > 0;
        if (rounds_after && !rounds_before)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "glitch rounds need a frame created without rebasing");
        std::lock_guard<std::mutex> lock(d->engine->mutex);
        d->desc = *desc;
    });
}

int cu_deep_destroy(cu_deep d) {
    return guarded([&] {
        if (!d) return;
        ensure_device();
        free_deep(d);
    });
}

int cu_deep_render(cu_deep d, const uint32_t *d_palette, uint32_t palette_length,
                   uint32_t *d_rgba, cu_stream stream) {
    return guarded([&] {
        if (!d) fail(CU_ERR_ILLEGAL_ARGUMENT, "deep frame is NULL");
        if ((d_rgba == nullptr) != (d_palette == nullptr))
            fail(CU_ERR_ILLEGAL_ARGUMENT, "palette and rgba go together");
        if (d_palette && palette_length == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty palette");
        std::lock_guard<std::mutex> lock(d->engine->mutex);
        ensure_device();
        hipStream_t st = stream ? stream->stream : nullptr;
        const cu_deep_desc &c = d->desc;
        // The primary orbit is kept when the caller allows it and the centre, bailout
        // and a large enough iteration limit match: an orbit computed to more
        // iterations than a frame uses gives that frame the same result (its pixels
        // stop at max_iter before reaching the orbit's end).
        const double key[5] = {c.centre_x, c.centre_x_lo, c.centre_y, c.centre_y_lo, c.bailout_sq};
        const bool keep = (c.flags & CU_DEEP_REUSE_ORBIT) && d->orbit_valid &&
                          std::memcmp(key, d->orbit_key, sizeof(key)) == 0 &&
                          c.max_iter <= d->orbit_mi;
        fill_u32(d->engine, stream, d->d_iters, 0, CU_ITER_SLOTS * 2);
        if (keep)
            fill_u32(d->engine, stream, (uint32_t *)d->d_len + 1, 0, 1);
        else
            fill_u32(d->engine, stream, d->d_len, 0, 2);
        check(hipEventRecord(d->ev[0], st), "hipEventRecord");
        d->references = 0;
        d->glitched = 0;
        if (d->method != CU_DEEP_DIRECT_F64 && !keep) {
            const uint32_t mi = (c.flags & CU_DEEP_REUSE_ORBIT) ? d->capacity_iter : c.max_iter;
            launch_orbit(d, stream, 0, c.centre_x, c.centre_x_lo, c.centre_y, c.centre_y_lo, mi);
            d->references = 1;
            d->orbit_valid = true;
            std::memcpy(d->orbit_key, key, sizeof(key));
            d->orbit_mi = mi;
        }
        check(hipEventRecord(d->ev[1], st), "hipEventRecord");
        DeepArgs a = base_args(d, d_palette, palette_length, d_rgba);
        launch_pixels(d, stream, a);
        check(hipEventRecord(d->ev[2], st), "hipEventRecord");
        const bool rounds = d->method != CU_DEEP_DIRECT_F64 && !(c.flags & CU_DEEP_REBASE);
        if (rounds) {
            const size_t pixels = (size_t)c.width * c.height;
            std::vector<unsigned char> flags(pixels);
            for (uint32_t r = 0;; ++r) {
                check(hipStreamSynchronize(st), "hipStreamSynchronize");
                check(hipMemcpy(flags.data(), d->d_glitch, pixels, hipMemcpyDeviceToHost),
                      "hipMemcpy D2H");
                uint32_t px = 0, py = 0, count = 0;
                if (!pick_secondary(flags, c.width, c.height, px, py, count)) break;
                d->glitched = count;
                if (r >= c.max_references) break;
                const double ox = ((double)px + 0.5 - (double)c.width * 0.5) * c.pixel_w;
                const double oy = ((double)c.height * 0.5 - ((double)py + 0.5)) * c.pixel_h;
                double rx_hi, rx_lo, ry_hi, ry_lo;
                dd_add_double(c.centre_x, c.centre_x_lo, ox, rx_hi, rx_lo);
                dd_add_double(c.centre_y, c.centre_y_lo, oy, ry_hi, ry_lo);
                launch_orbit(d, stream, 1, rx_hi, rx_lo, ry_hi, ry_lo, c.max_iter);
                ++d->references;
                DeepArgs g = a;
                g.orbit = d->d_orbit[1];
                g.orbit_len = (const uint32_t *)d->d_len + 1;
                g.ref_ox = ox;
                g.ref_oy = oy;
                g.only_glitched = 1;
                launch_pixels(d, stream, g);
                d->glitched = 0;
            }
        }
        check(hipEventRecord(d->ev[3], st), "hipEventRecord");
        d->last_stream = st;
        d->rendered = true;
    });
}

int cu_deep_outputs(cu_deep d, uint32_t **d_counts, double **d_smooth, uint8_t **d_glitch) {
    return guarded([&] {
        if (!d) fail(CU_ERR_ILLEGAL_ARGUMENT, "deep frame is NULL");
        if (d_counts) *d_counts = (uint32_t *)d->d_counts;
        if (d_smooth) *d_smooth = (double *)d->d_smooth;
        if (d_glitch) *d_glitch = (uint8_t *)d->d_glitch;
    });
}

int cu_deep_stats(cu_deep d, cu_deep_report *out) {
    return guarded([&] {
        if (!d || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (!d->rendered) fail(CU_ERR_ILLEGAL_STATE, "deep frame has not been rendered");
        ensure_device();
        check(hipEventSynchronize(d->ev[3]), "hipEventSynchronize");
        cu_deep_report s{};
        if (cu_iters_total((const unsigned long long *)d->d_iters, &s.iterations) != CU_OK)
            fail(CU_ERR_DEVICE, cu_last_error());
        uint32_t len[2] = {0, 0};
        check(hipMemcpy(len, d->d_len, sizeof(len), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        s.reference_length = len[0];
        s.references = d->references;
        s.glitched = d->glitched;
        check(hipEventElapsedTime(&s.orbit_ms, d->ev[0], d->ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&s.render_ms, d->ev[1], d->ev[2]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&s.total_ms, d->ev[0], d->ev[3]), "hipEventElapsedTime");
        *out = s;
    });
}

int cu_deep_orbit(cu_deep d, void *host_out, uint32_t capacity_entries, uint32_t *length) {
    return guarded([&] {
        if (!d || !host_out || !length) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (d->method == CU_DEEP_DIRECT_F64) fail(CU_ERR_ILLEGAL_STATE, "direct frames have no orbit");
        if (!d->rendered) fail(CU_ERR_ILLEGAL_STATE, "deep frame has not been rendered");
        ensure_device();
        check(hipEventSynchronize(d->ev[3]), "hipEventSynchronize");
        uint32_t M = 0;
        check(hipMemcpy(&M, d->d_len, sizeof(M), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if ((uint64_t)M + 1 > capacity_entries) fail(CU_ERR_ILLEGAL_ARGUMENT, "capacity too small");
        check(hipMemcpy(host_out, d->d_orbit[0], ((size_t)M + 1) * d->entry_bytes,
                        hipMemcpyDeviceToHost), "hipMemcpy D2H");
        *length = M;
    });
}

}  // extern "C"
