// chaos_oracle.cpp — CPU oracle of the escape-time render path.
//
// TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg run beside the HIP path.  Nothing in the
// product links or loads it.
//
// Built with g++ -O2 -ffp-contract=off -fno-fast-math (x86-64 SSE: float and
// double arithmetic round to their own formats, no excess precision), so every
// expression below rounds exactly as written, left to right.
//
// Pinning (what checks each function; tests/oracle_pins.py holds the checks,
// tests/test_oracle_mutants.py shows each one failing on a copy of this file
// with a planted mistake).  The reference's computeFractal cannot be compiled
// here (its sources need CUDA toolkit headers this image lacks), so nothing is
// checked against a build of it:
//   or_mandelbrot_* / or_julia_*   FP32 and FP64: numpy float32/float64 brute
//                                  force of the reference loop on grids and near
//                                  the boundary, and on the whole C1 frame;
//                                  dyadic special points (c = -2 stops at
//                                  |z|^2 = 4); mirror and point symmetry;
//                                  Julia = Mandelbrot - 1; seahorse and cusp
//                                  asymptotics.
//   or_escape_*, or_tile_point_*   the FP32 tile path against the float32 brute
//                                  force; sample positions exact on dyadic grids
//                                  and equal to float32 evaluation elsewhere;
//                                  300-digit mpmath orbits at other bailouts.
//   or_smooth                      mpmath on tiny inputs; continuity across
//                                  count bands.
//   or_jitter                      published SplitMix64 vectors.
//   or_render_main / _advanced     special cases (1 spp = corner sample; reuse of
//                                  an identical rectangle copies the value).
//                                  The __activemask() membership of the votes is
//                                  not pinned by anything the reference holds.
//   or_foveation                   float64 closed form at distances from the
//                                  fovea to beyond 60 degrees (CUDA's atanf is
//                                  taken as the double arctangent rounded once).
//   or_compose*                    known palette slots (half-integers, values at
//                                  and beyond L, negatives) and grey levels;
//                                  byte order.
//   or_render_undersampled         the set of pixels written and their values.
//   or_default_palette             the reference's palettes/defaultPalette.png.
#include "chaos_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kMaxSuperSampling = 64;  // fractalRendererGeneric.cu:12
constexpr uint32_t kAdaptiveThreshold = 10; // fractalRendererGeneric.cu:93
constexpr float kPiF = 3.14159265358979f;   // fractalRendererGeneric.cu:11

// ---------------------------------------------------------------------------
// escape loop: mandelbrot.cu:16-21 with the bailout constant as a parameter
// ---------------------------------------------------------------------------
template <class Real>
uint32_t escape_loop(uint32_t max_iter, Real x, Real y, Real cx, Real cy, Real bailout_sq,
                     Real *mag2, int *escaped) {
    uint32_t i = 0;
    *escaped = 0;
    *mag2 = 0;
    while (i < max_iter) {
        Real m = x * x + y * y;
        if (!(m < bailout_sq)) {
            *escaped = 1;
            *mag2 = m;
            break;
        }
        Real xn = x * x - y * y + cx;  // (x*x - y*y) + cx
        y = 2 * x * y + cy;            // ((2*x)*y) + cy
        x = xn;
        ++i;
    }
    return i;
}

// Iterations executed by compute_fractal (loop bodies run: the escape step, or
// max_iter for an orbit that never escapes), summed over all calls since the
// last or_render_iterations().  The renderer kernels count the same quantity.
std::atomic<uint64_t> g_render_iterations{0};

// computeFractal<Real> of the two in-scope modules, bailout 4.
template <class Real>
float compute_fractal(int kind, uint32_t max_iter, Real px, Real py, double jcx, double jcy) {
    Real m;
    int esc;
    if (kind == CU_FRACTAL_JULIA) {
        // julia.cu:7 — c is the __constant__ double pair cast to Real; returns i.
        uint32_t i = escape_loop<Real>(max_iter, px, py, (Real)jcx, (Real)jcy, (Real)4, &m, &esc);
        g_render_iterations.fetch_add(i, std::memory_order_relaxed);
        return (float)i;
    }
    uint32_t i = escape_loop<Real>(max_iter, (Real)0, (Real)0, px, py, (Real)4, &m, &esc);
    g_render_iterations.fetch_add(i, std::memory_order_relaxed);
    if (i == max_iter) return 0;  // mandelbrot.cu:22-23
    return (float)i;
}

// ---------------------------------------------------------------------------
// planar tiles (extension: no reference counterpart)
// ---------------------------------------------------------------------------
void sample_offset(const cu_tile_desc &t, uint32_t px, uint32_t py, uint32_t s, double *dx,
                   double *dy) {
    switch (t.pattern) {
        case CU_SAMPLE_CENTRE: *dx = 0.5; *dy = 0.5; break;
        case CU_SAMPLE_CORNER: *dx = 0.0; *dy = 0.0; break;
        default: {
            uint64_t id = ((uint64_t)py * t.width + px) * t.spp + s;
            or_jitter(t.seed, id, dx, dy);
        }
    }
}

template <class Real>
void tile_point(const cu_tile_desc &t, uint32_t px, uint32_t py, uint32_t s, Real *x, Real *y);

template <>
void tile_point<double>(const cu_tile_desc &t, uint32_t px, uint32_t py, uint32_t s, double *x,
                        double *y) {
    double dx, dy;
    sample_offset(t, px, py, s, &dx, &dy);
    *x = t.left + ((double)px + dx) * t.pixel_w;
    *y = t.top - ((double)py + dy) * t.pixel_h;
}

template <>
void tile_point<float>(const cu_tile_desc &t, uint32_t px, uint32_t py, uint32_t s, float *x,
                       float *y) {
    double dx, dy;
    sample_offset(t, px, py, s, &dx, &dy);
    *x = t.left_f + ((float)px + (float)dx) * t.pixel_w_f;
    *y = t.top_f - ((float)py + (float)dy) * t.pixel_h_f;
}

template <class Real>
uint64_t tile_rows(const cu_tile_desc &t, uint32_t row0, uint32_t row1, uint32_t *counts,
                   double *smooth, int threads) {
    uint64_t iters = 0;
    if (threads <= 0) {
#ifdef _OPENMP
        threads = omp_get_max_threads();
#else
        threads = 1;
#endif
    }
    const Real bsq = (Real)t.bailout_sq;
#pragma omp parallel for schedule(dynamic) reduction(+ : iters) num_threads(threads)
    for (int64_t py = row0; py < (int64_t)row1; ++py) {
        for (uint32_t px = 0; px < t.width; ++px) {
            uint32_t count_sum = 0;
            double smooth_sum = 0;
            for (uint32_t s = 0; s < t.spp; ++s) {
                Real X, Y;
                tile_point<Real>(t, px, (uint32_t)py, s, &X, &Y);
                Real zx = 0, zy = 0, cx = X, cy = Y;
                if (t.kind == CU_FRACTAL_JULIA) {
                    zx = X;
                    zy = Y;
                    cx = (Real)t.julia_cx;
                    cy = (Real)t.julia_cy;
                }
                Real m;
                int esc;
                uint32_t n = escape_loop<Real>(t.max_iter, zx, zy, cx, cy, bsq, &m, &esc);
                iters += n;
                uint32_t c = esc ? n : (t.kind == CU_FRACTAL_JULIA ? t.max_iter : 0u);
                count_sum += c;
                smooth_sum += esc ? or_smooth(n, (double)m, t.bailout_sq) : 0.0;
            }
            uint64_t o = t.out_offset + (uint64_t)py * t.out_pitch + px;
            counts[o] = count_sum / t.spp;
            if (smooth) smooth[o] = smooth_sum / (double)t.spp;
        }
    }
    return iters;
}

// ---------------------------------------------------------------------------
// sampleTheFractal<Real> in lock step over one 8x4 vote group
// ---------------------------------------------------------------------------
struct Lane {
    uint32_t px = 0, py = 0;
    int site = 0;              // which call site of sampleTheFractal the lane reaches (0: none)
    float sample_count_f = 0;  // in/out parameter sampleCountF
    uint32_t sample_count = 0;
    uint32_t escape_sum = 0;
    float samples[kAdaptiveThreshold] = {};
    bool in_loop = false;
    uint32_t result = 0;
};

// computeDispersion (fractalRendererGeneric.cu:24-34).  For data_length > 10
// the reference reads past its 10-entry array (undefined); only the stored
// entries are summed here, the divisor keeps n = data_length.
float compute_dispersion(const float *data, uint32_t data_length, float mean) {
    uint32_t n = data_length;
    float variance = 0;
    uint32_t stored = std::min(data_length, kAdaptiveThreshold);
    for (uint32_t i = 0; i < stored; i++) variance += (data[i] - mean) * (data[i] - mean);
    variance /= (n - 1);
    return variance / mean;
}

template <class Real>
void sample_group(std::vector<Lane *> &lanes, uint32_t width, uint32_t height, Real lbx, Real lby,
                  Real rtx, Real rty, uint32_t max_iter, bool adaptive, int kind, double jcx,
                  double jcy) {
    if (lanes.empty()) return;
    // pixelSize = image.size() / gridSize.cast<Real>()  (:97)
    const Real psx = (rtx - lbx) / (Real)width;
    const Real psy = (rty - lby) / (Real)height;
    bool any = false;
    for (Lane *l : lanes) {
        if (l->sample_count_f < 1) {  // :87-90
            l->sample_count_f = 0;
            l->result = 0;
            l->in_loop = false;
            continue;
        }
        l->sample_count = std::min(kMaxSuperSampling, (uint32_t)std::round(l->sample_count_f));
        l->escape_sum = 0;
        l->in_loop = l->sample_count > 0;
        if (!l->in_loop) {  // round() gave 0 samples: the mean divides by zero in the reference
            l->result = 0;
            l->sample_count_f = 0;
        }
        any = any || l->in_loop;
    }
    const uint32_t skip_first = 2;
    for (uint32_t i = 0; any; ++i) {
        for (Lane *l : lanes) {
            if (!l->in_loop) continue;
            Real dx, dy;
            if (i <= skip_first) {
                dx = dy = (Real)i / (Real)(skip_first + 1);  // :107-108
            } else {                                          // :110-116
                float per_row_f = std::sqrt(l->sample_count_f - skip_first);
                uint32_t per_row_i = (uint32_t)std::round(per_row_f);
                float fx = ((i - skip_first) % per_row_i) / per_row_f;
                float fy = ((i - skip_first) / per_row_i) / per_row_f;
                dx = fx;
                dy = fy;
            }
            // c = leftTop + (1,-1) * (pixel + delta) * pixelSize  (:119-124)
            Real X = lbx + ((Real)l->px + dx) * psx;
            Real Y = rty - ((Real)l->py + dy) * psy;
            uint32_t escape_time = (uint32_t)compute_fractal<Real>(kind, max_iter, X, Y, jcx, jcy);
            l->escape_sum += escape_time;
            if (i < kAdaptiveThreshold) l->samples[i] = (float)escape_time;
        }
        // lanes entering the adaptive block (:133); the three votes run over them
        std::vector<Lane *> part;
        for (Lane *l : lanes)
            if (l->in_loop && ((adaptive && (i > 0 && i < kAdaptiveThreshold)) ||
                               (i == l->sample_count / 2)))
                part.push_back(l);
        if (!part.empty()) {
            std::vector<float> disp(part.size());
            for (size_t k = 0; k < part.size(); ++k) {
                float mean = part[k]->escape_sum / (i + 1);  // integer mean (:135)
                disp[k] = compute_dispersion(part[k]->samples, i, mean);
            }
            const float threshold = 0.01f;
            bool vote1 = false;
            if (i == 1) {
                vote1 = true;
                for (Lane *l : part)
                    vote1 = vote1 && (std::fabs(l->samples[0] - l->samples[1]) < FLT_EPSILON);
            }
            if (i == 1 && vote1) {
                for (Lane *l : part) l->sample_count = i + 1;
            } else {
                bool vote2 = true;
                for (size_t k = 0; k < part.size(); ++k) vote2 = vote2 && (disp[k] < threshold);
                if (vote2) {
                    for (Lane *l : part) l->sample_count = i + 1;
                } else {
                    bool vote3 = true;
                    for (size_t k = 0; k < part.size(); ++k) vote3 = vote3 && (disp[k] <= 1);
                    if (vote3)
                        for (Lane *l : part)
                            if (i >= l->sample_count / 2) l->sample_count = i + 1;
                }
            }
        }
        any = false;
        for (Lane *l : lanes) {
            if (!l->in_loop) continue;
            if (!(i + 1 < l->sample_count)) {
                l->in_loop = false;
                l->result = l->escape_sum / l->sample_count;  // :152
                l->sample_count_f = (float)l->sample_count;   // :153
            } else {
                any = true;
            }
        }
    }
}

inline or_pixel_info *pixel_ptr(or_pixel_info *base, int64_t pitch, uint32_t x, uint32_t y) {
    return (or_pixel_info *)((char *)base + (int64_t)y * pitch) + x;
}
inline const or_pixel_info *pixel_ptr(const or_pixel_info *base, int64_t pitch, uint32_t x,
                                      uint32_t y) {
    return (const or_pixel_info *)((const char *)base + (int64_t)y * pitch) + x;
}

template <class Real>
void render_main(int kind, double jcx, double jcy, or_pixel_info *out, int64_t pitch, uint32_t W,
                 uint32_t H, double lbx, double lby, double rtx, double rty, uint32_t max_iter,
                 float max_ss, uint32_t flags) {
    const uint32_t gw = (W + 7) / 8, gh = (H + 3) / 4;
#pragma omp parallel for schedule(dynamic)
    for (int64_t g = 0; g < (int64_t)gw * gh; ++g) {
        uint32_t gx = (uint32_t)(g % gw), gy = (uint32_t)(g / gw);
        Lane lanes[32];
        std::vector<Lane *> group;
        for (uint32_t k = 0; k < 32; ++k) {
            uint32_t x = gx * 8 + k % 8, y = gy * 4 + k / 8;
            if (x >= W || y >= H) continue;  // :172 — out-of-image threads leave
            lanes[k].px = x;
            lanes[k].py = y;
            lanes[k].sample_count_f = max_ss;
            group.push_back(&lanes[k]);
        }
        sample_group<Real>(group, W, H, (Real)lbx, (Real)lby, (Real)rtx, (Real)rty, max_iter,
                           (flags & CU_FLAG_ADAPTIVE_SS) != 0, kind, jcx, jcy);
        for (Lane *l : group) {
            or_pixel_info p;
            std::memset(&p, 0, sizeof p);
            p.value = (float)l->result;  // pixel_info_t(uint value, float weight)
            p.weight = l->sample_count_f;
            *pixel_ptr(out, pitch, l->px, l->py) = p;
        }
    }
}

// getWarpingOriginOfSampleReuse<Real> (:194-203)
template <class Real>
void warping_origin(uint32_t px, uint32_t py, uint32_t W, uint32_t H, Real lbx, Real lby, Real rtx,
                    Real rty, Real olbx, Real olby, Real ortx, Real orty, float *ox, float *oy) {
    py = H - py;
    Real plane_x = (Real)px / (Real)W * (rtx - lbx) + lbx;
    Real plane_y = (Real)py / (Real)H * (rty - lby) + lby;
    Real rel_x = (plane_x - olbx) / (ortx - olbx);
    Real rel_y = (plane_y - olby) / (orty - olby);
    float fx = (float)rel_x * (float)W;
    float fy = (float)rel_y * (float)H;
    fy = H - fy;
    *ox = fx;
    *oy = fy;
}

// readFromArrayUsingFiltering<float> (:259-302)
or_pixel_info read_filtered(const or_pixel_info *arr, int64_t pitch, float xb, float yb) {
    uint32_t i = (uint32_t)std::floor(xb);
    uint32_t j = (uint32_t)std::floor(yb);
    float alpha = xb - i;
    float beta = yb - j;
    or_pixel_info t00 = *pixel_ptr(arr, pitch, i, j);
    or_pixel_info t10 = *pixel_ptr(arr, pitch, i + 1, j);
    or_pixel_info t01 = *pixel_ptr(arr, pitch, i, j + 1);
    or_pixel_info t11 = *pixel_ptr(arr, pitch, i + 1, j + 1);
    float r = (1 - alpha) * (1 - beta) * t00.value + alpha * (1 - beta) * t10.value +
              (1 - alpha) * beta * t01.value + alpha * beta * t11.value;
    float w = (1 - alpha) * (1 - beta) * t00.weight + alpha * (1 - beta) * t10.weight +
              (1 - alpha) * beta * t01.weight + alpha * beta * t11.weight;
    or_pixel_info res;
    std::memset(&res, 0, sizeof res);
    res.value = r;
    res.weight = w;
    return res;
}

template <class Real>
void render_advanced(int kind, double jcx, double jcy, or_pixel_info *out, int64_t pitch,
                     uint32_t W, uint32_t H, double lbx, double lby, double rtx, double rty,
                     uint32_t max_iter, float max_ss, uint32_t flags, double olbx, double olby,
                     double ortx, double orty, const or_pixel_info *in, int64_t in_pitch,
                     uint32_t fx, uint32_t fy) {
    const uint32_t gw = (W + 7) / 8, gh = (H + 3) / 4;
    const bool adaptive = (flags & CU_FLAG_ADAPTIVE_SS) != 0;
#pragma omp parallel for schedule(dynamic)
    for (int64_t g = 0; g < (int64_t)gw * gh; ++g) {
        uint32_t gx = (uint32_t)(g % gw), gy = (uint32_t)(g / gw);
        Lane lanes[32];
        or_pixel_info results[32], reused[32];
        bool inside[32];
        std::vector<Lane *> site1, site2;
        std::vector<uint32_t> members;
        for (uint32_t k = 0; k < 32; ++k) {
            uint32_t x = gx * 8 + k % 8, y = gy * 4 + k / 8;
            if (x >= W || y >= H) continue;
            members.push_back(k);
            Lane &l = lanes[k];
            l.px = x;
            l.py = y;
            float advised = max_ss;  // fov_result_t fovResult(maxSuperSampling, false)
            int ins = 0;
            if ((flags & CU_FLAG_FOVEATION) && (flags & CU_FLAG_ZOOMING) &&
                (flags & CU_FLAG_ZOOMING_IN))
                or_foveation(x, y, fx, fy, max_ss, &advised, &ins);
            inside[k] = ins != 0;
            bool reusing = false;
            std::memset(&reused[k], 0, sizeof(or_pixel_info));
            // images 1 pixel wide or high never reuse (the reference would read outside)
            if ((flags & CU_FLAG_SAMPLE_REUSE) && W >= 2 && H >= 2) {
                float ox, oy;
                warping_origin<Real>(x, y, W, H, (Real)lbx, (Real)lby, (Real)rtx, (Real)rty,
                                     (Real)olbx, (Real)olby, (Real)ortx, (Real)orty, &ox, &oy);
                int oix = (int)std::round(ox), oiy = (int)std::round(oy);
                if (oix < 2 || (uint32_t)oix >= W - 2 || oiy < 2 || (uint32_t)oiy >= H - 2) {
                    reusing = false;
                } else {
                    reused[k] = read_filtered(in, in_pitch, ox, oy);
                    reusing = !((double)reused[k].weight < 0.1);
                }
            }
            or_pixel_info &res = results[k];
            std::memset(&res, 0, sizeof res);
            l.sample_count_f = advised;
            l.site = 0;
            if (reusing) {
                res = reused[k];
                res.is_reused = 1;
                res.weight_of_new_samples = 0;
                if ((flags & CU_FLAG_ZOOMING_IN) && inside[k]) {
                    l.site = 1;
                    site1.push_back(&l);
                }
            } else {
                if (l.sample_count_f < 1) l.sample_count_f = 1;
                l.site = 2;
                site2.push_back(&l);
            }
        }
        sample_group<Real>(site1, W, H, (Real)lbx, (Real)lby, (Real)rtx, (Real)rty, max_iter,
                           adaptive, kind, jcx, jcy);
        sample_group<Real>(site2, W, H, (Real)lbx, (Real)lby, (Real)rtx, (Real)rty, max_iter,
                           adaptive, kind, jcx, jcy);
        for (uint32_t k : members) {
            Lane &l = lanes[k];
            or_pixel_info &res = results[k];
            if (l.site == 1) {  // :349-356
                float samples = (float)l.result;
                float sample_count = l.sample_count_f;
                reused[k].weight = (float)((double)reused[k].weight * 0.75);
                res.weight_of_new_samples = sample_count;
                res.weight = reused[k].weight + sample_count;
                res.value = (reused[k].value * reused[k].weight + samples * sample_count) / res.weight;
            } else if (l.site == 2) {  // :358-362
                res.value = (float)l.result;
                res.weight = l.sample_count_f;
            }
            *pixel_ptr(out, pitch, l.px, l.py) = res;
        }
    }
}

// colorizeSampleCount (:67-77).  The reference converts 255*rel (float) to a
// signed char; that conversion is defined here as truncation to int followed by
// the low byte.
uint32_t colorize_sample_count(uint32_t sample_count, uint32_t full) {
    sample_count = std::min(sample_count, full);
    float rel = sample_count / (float)full;
    uint32_t v = (uint32_t)(int)(255 * rel) & 0xffu;
    return v | (v << 8) | (v << 16) | (0xffu << 24);
}

// Palette slot L - (round(v) mod L) - 1 of colorize (mandelbrot.cu:27-35),
// round half away from zero.  A value that is not above zero takes the slot
// of 0 (the reference's values never are; a negative smooth value of an orbit
// starting far outside would make the unsigned conversion undefined).
uint32_t palette_slot(double v, uint32_t length) {
    uint32_t vi = v > 0.0 ? (uint32_t)std::round(v) : 0u;
    return length - (vi % length) - 1;
}

// colorize (mandelbrot.cu:27-35); the float value rounds as a float (roundf).
uint32_t colorize(const uint32_t *palette, uint32_t length, float value) {
    uint32_t vi = value > 0.0f ? (uint32_t)std::round(value) : 0u;
    return palette[length - (vi % length) - 1];
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

float or_mandelbrot_d(uint32_t max_iter, double cx, double cy) {
    return compute_fractal<double>(CU_FRACTAL_MANDELBROT, max_iter, cx, cy, 0, 0);
}
float or_mandelbrot_f(uint32_t max_iter, float cx, float cy) {
    return compute_fractal<float>(CU_FRACTAL_MANDELBROT, max_iter, cx, cy, 0, 0);
}
float or_julia_d(uint32_t max_iter, double zx, double zy, double jcx, double jcy) {
    return compute_fractal<double>(CU_FRACTAL_JULIA, max_iter, zx, zy, jcx, jcy);
}
float or_julia_f(uint32_t max_iter, float zx, float zy, double jcx, double jcy) {
    return compute_fractal<float>(CU_FRACTAL_JULIA, max_iter, zx, zy, jcx, jcy);
}

uint32_t or_escape_d(uint32_t max_iter, double zx, double zy, double cx, double cy,
                     double bailout_sq, double *mag2, int *escaped) {
    return escape_loop<double>(max_iter, zx, zy, cx, cy, bailout_sq, mag2, escaped);
}
uint32_t or_escape_f(uint32_t max_iter, float zx, float zy, float cx, float cy, float bailout_sq,
                     float *mag2, int *escaped) {
    return escape_loop<float>(max_iter, zx, zy, cx, cy, bailout_sq, mag2, escaped);
}

double or_smooth(uint32_t n, double mag2, double bailout_sq) {
    return (double)n + 1.0 - std::log2(std::log(mag2) / std::log(bailout_sq));
}

uint64_t or_splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Sample j uses output j of SplitMix64 seeded with `seed` (state seed + j*gamma
// before the generator's own increment): high 32 bits -> dx, low 32 -> dy.
void or_jitter(uint64_t seed, uint64_t sample_id, double *dx, double *dy) {
    uint64_t h = or_splitmix64(seed + sample_id * 0x9E3779B97F4A7C15ull);
    *dx = (double)(uint32_t)(h >> 32) * 0x1.0p-32;
    *dy = (double)(uint32_t)h * 0x1.0p-32;
}

void or_tile_point_d(const cu_tile_desc *t, uint32_t px, uint32_t py, uint32_t s, double *x,
                     double *y) {
    tile_point<double>(*t, px, py, s, x, y);
}
void or_tile_point_f(const cu_tile_desc *t, uint32_t px, uint32_t py, uint32_t s, float *x,
                     float *y) {
    tile_point<float>(*t, px, py, s, x, y);
}

uint64_t or_escape_tile_rows(const cu_tile_desc *t, int precision, uint32_t row0, uint32_t row1,
                             uint32_t *counts, double *smooth, int threads) {
    if (row1 > t->height) row1 = t->height;
    if (row0 >= row1) return 0;
    if (precision == 32) return tile_rows<float>(*t, row0, row1, counts, smooth, threads);
    return tile_rows<double>(*t, row0, row1, counts, smooth, threads);
}

uint64_t or_render_iterations(void) { return g_render_iterations.exchange(0); }

void or_render_main(int precision, int kind, double jcx, double jcy, or_pixel_info *out,
                    int64_t pitch, uint32_t W, uint32_t H, double lbx, double lby, double rtx,
                    double rty, uint32_t max_iter, float max_ss, uint32_t flags) {
    if (precision == 32)
        render_main<float>(kind, jcx, jcy, out, pitch, W, H, lbx, lby, rtx, rty, max_iter, max_ss,
                           flags);
    else
        render_main<double>(kind, jcx, jcy, out, pitch, W, H, lbx, lby, rtx, rty, max_iter, max_ss,
                            flags);
}

void or_render_advanced(int precision, int kind, double jcx, double jcy, or_pixel_info *out,
                        int64_t pitch, uint32_t W, uint32_t H, double lbx, double lby, double rtx,
                        double rty, uint32_t max_iter, float max_ss, uint32_t flags, double olbx,
                        double olby, double ortx, double orty, const or_pixel_info *in,
                        int64_t in_pitch, uint32_t fx, uint32_t fy) {
    if (precision == 32)
        render_advanced<float>(kind, jcx, jcy, out, pitch, W, H, lbx, lby, rtx, rty, max_iter,
                               max_ss, flags, olbx, olby, ortx, orty, in, in_pitch, fx, fy);
    else
        render_advanced<double>(kind, jcx, jcy, out, pitch, W, H, lbx, lby, rtx, rty, max_iter,
                                max_ss, flags, olbx, olby, ortx, orty, in, in_pitch, fx, fy);
}

void or_foveation(uint32_t px, uint32_t py, uint32_t fx, uint32_t fy, float max_ss,
                  float *advised, int *inside) {
    // per-warp normalisation to the 8x4 rectangle (:231)
    uint32_t nx = px - px % 8, ny = py - py % 4;
    const float foveal = max_ss >= 1 ? 5.5 : 5.5 * max_ss;  // double expression → float
    const float peripheral = 60;
    const float screen_distance = 60;          // :216
    const float pixel_width_cm = 0.02652;      // :217
    float ddx = (float)fx - (float)nx, ddy = (float)fy - (float)ny;
    float focus_distance = std::sqrt(ddx * ddx + ddy * ddy) * pixel_width_cm;
    // atan: the reference's CUDA atanf is not available; the arctangent is taken
    // in double and rounded once to float (the GPU kernel does the same).
    float angle = (float)std::atan((double)(focus_distance / screen_distance)) * 180 / kPiF;
    // linearMapping<float>(angle, foveal, peripheral, 1, 0) (:209-214)
    float a = 1, b = 0, r = foveal, s = peripheral;
    float k = (b - a) / (s - r);
    float q = (s * a - b * r) / (s - r);
    float lm = k * angle + q;
    float rel = (float)std::fmin(1.0, (double)lm);
    float adv = max_ss * rel;
    int ins = 0;
    if (angle <= foveal) {
        adv = (float)std::fmax(1.0, (double)adv);
        ins = 1;
    }
    *advised = adv;
    *inside = ins;
}

void or_compose(const or_pixel_info *in, int64_t in_pitch, uint32_t *out, uint32_t out_pitch,
                uint32_t W, uint32_t H, const uint32_t *palette, uint32_t L, float max_ss,
                int visualise) {
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const or_pixel_info &p = *pixel_ptr(in, in_pitch, x, y);
            uint32_t color;
            if (visualise) {
                uint32_t full = (uint32_t)std::fmax(1.0, (double)max_ss);
                color = colorize_sample_count((uint32_t)p.weight, full);
                if (p.is_reused) color = colorize_sample_count((uint32_t)p.weight_of_new_samples, full);
            } else {
                color = colorize(palette, L, p.value);
            }
            out[(size_t)y * out_pitch + x] = color;
        }
}

void or_compose_smooth(const double *smooth, uint32_t smooth_pitch, uint32_t *out,
                       uint32_t out_pitch, uint32_t W, uint32_t H, const uint32_t *palette,
                       uint32_t L) {
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            out[(size_t)y * out_pitch + x] = palette[palette_slot(smooth[(size_t)y * smooth_pitch + x], L)];
        }
}

void or_render_undersampled(or_pixel_info *out, int64_t pitch, uint32_t W, uint32_t H, float lbx,
                            float lby, float rtx, float rty, uint32_t max_iter, uint32_t level) {
    if (level == 0 || level > W || level > H) return;  // would write outside the image
    float pw = (rtx - lbx) / (float)W;
    float ph = (rty - lby) / (float)H;
    for (uint64_t y = 0; (uint32_t)y < H - level; y += level)
        for (uint64_t x = 0; (uint32_t)x < W - level; x += level) {
            float cx = lbx + (uint32_t)x * pw;
            float cy = rty - (uint32_t)y * ph;
            uint32_t e = (uint32_t)compute_fractal<float>(CU_FRACTAL_MANDELBROT, max_iter, cx, cy, 0, 0);
            or_pixel_info p;
            std::memset(&p, 0, sizeof p);
            p.value = (float)e;
            p.weight = 1 / (float)level;
            *pixel_ptr(out, pitch, (uint32_t)x, (uint32_t)y) = p;
        }
}

static uint32_t rgb_to_rgba(int r, int g, int b) {
    return ((uint32_t)r & 0xffu) | (((uint32_t)g & 0xffu) << 8) | (((uint32_t)b & 0xffu) << 16) |
           0xff000000u;
}

void or_default_palette(uint32_t *p) {
    const int max = 256, full = 255;
    for (int i = 0; i < max; i++) p[i] = rgb_to_rgba(i, 0, std::min(full, full / 2 + i));
    for (int i = max; i < 2 * max; i++) p[i] = rgb_to_rgba(full, 0, full - i);
    for (int i = 2 * max; i < 3 * max; i++) p[i] = rgb_to_rgba(full, i, 0);
    for (int i = 3 * max; i < 4 * max; i++) p[i] = rgb_to_rgba(full - i, full, 0);
    for (int i = 4 * max; i < 5 * max; i++) p[i] = rgb_to_rgba(0, full, i);
    for (int i = 5 * max; i < 6 * max; i++) p[i] = rgb_to_rgba(0, full - i, full);
}

uint32_t or_argb_to_rgba(uint32_t argb) {
    uint32_t a = argb >> 24, r = (argb >> 16) & 0xff, g = (argb >> 8) & 0xff, b = argb & 0xff;
    // Integer.reverseBytes then an arithmetic >>8: the top byte fills with the
    // sign of b, then the alpha byte is OR-ed in.
    uint32_t top = ((b & 0x80) ? 0xffu : 0u) | a;
    return (top << 24) | (b << 16) | (g << 8) | r;
}

}  // extern "C"
