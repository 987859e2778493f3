// deep_oracle.cpp — CPU oracle of the deep-zoom (perturbation) render form.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's checks, never by
// the product path.  Built into liboracle.so with FP contraction off; every
// fused multiply-add below is an explicit std::fma / std::fmaf, so the
// arithmetic is one IEEE-754 operation sequence that the gfx950 kernels repeat
// op for op.
//
// The form has no reference counterpart: the reference's double kernel is the
// deepest it goes (cudarenderer/RenderingKernel.java:124-138 selects
// double, and cudarenderer/CudaFractalRenderer.java:217-220 only logs when the
// pixel is below the double ulp).  Its iteration is the one of
// src/main/cuda/fractals/mandelbrot.cu:16-20 (escape at |z|^2 >= bailout_sq),
// taken relative to a reference orbit:
//
//   z_n = Z_m + d_n,   d_{n+1} = (2 Z_m + d_n) d_n + dc,   m <- m + 1
//
// Pixel (px, py) of a W x H frame centred on C samples the plane at
//   C + (ox, oy),  ox = (px + 0.5 - W/2) * pw,  oy = (H/2 - (py + 0.5)) * ph
// (double offsets; C itself may be a double-double).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "chaos_oracle.h"

namespace {

// ---- double-double (Dekker / Knuth error-free transformations) ----
struct DD {
    double hi, lo;
};

inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
inline DD quick_two_sum(double a, double b) {
    const double s = a + b;
    const double e = b - (s - a);
    return {s, e};
}
inline DD two_prod(double a, double b) {
    const double p = a * b;
    const double e = std::fma(a, b, -p);
    return {p, e};
}
inline DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    double e = s.lo + t.hi;
    s = quick_two_sum(s.hi, e);
    e = s.lo + t.lo;
    return quick_two_sum(s.hi, e);
}
inline DD dd_neg(DD a) { return {-a.hi, -a.lo}; }
inline DD dd_mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    const double cross = a.hi * b.lo + a.lo * b.hi;
    const double e = p.lo + cross;
    return quick_two_sum(p.hi, e);
}
inline DD dd_twice(DD a) { return {a.hi + a.hi, a.lo + a.lo}; }

// One step Z <- Z^2 + C of the reference orbit in double-double, compensated
// rather than composed from dd_mul and dd_add: the leading products are split
// exactly (two_prod), the sums of leading terms are exact (two_sum), and the
// low-order terms (the products' errors, the cross terms 2 xh xl - 2 yh yl and
// 2 (xh yl + xl yh), C's low parts) are summed in plain double before one
// renormalisation.  Terms of order xl^2 are dropped.  The absolute error per step
// is a few units of 2^-104 |Z|^2, the accuracy of the composed form, in about
// half its operations (47 against 95).
inline void orbit_step(DD &x, DD &y, DD cx, DD cy) {
    const DD px = two_prod(x.hi, x.hi), py = two_prod(y.hi, y.hi), pxy = two_prod(x.hi, y.hi);
    // x' = xh^2 - yh^2 + cxh + [ex - ey + 2 (xh xl - yh yl) + cxl]
    const DD s1 = two_sum(px.hi, -py.hi);
    const DD s2 = two_sum(s1.hi, cx.hi);
    const double u = std::fma(x.hi, x.lo, -(y.hi * y.lo));
    const double xlo = ((s1.lo + s2.lo) + (px.lo - py.lo)) + ((u + u) + cx.lo);
    // y' = 2 xh yh + cyh + [2 exy + 2 (xh yl + xl yh) + cyl]
    const DD s3 = two_sum(pxy.hi + pxy.hi, cy.hi);
    const double w = std::fma(x.hi, y.lo, x.lo * y.hi);
    const double ylo = (s3.lo + (pxy.lo + pxy.lo)) + ((w + w) + cy.lo);
    x = quick_two_sum(s2.hi, xlo);
    y = quick_two_sum(s3.hi, ylo);
}

// Reference orbit Z_0 = 0, Z_{n+1} = Z_n^2 + C in double-double, stored rounded
// to Real: entry n = {Real(X.hi), Real(Y.hi), Real(0.5*(X.hi^2 + Y.hi^2)), 0}.
// Stops after storing the first Z_n with X.hi^2 + Y.hi^2 >= bailout_sq, or Z_max_iter.
// Returns M, the index of the last stored entry.
template <class Real>
uint32_t reference_orbit(double cx_hi, double cx_lo, double cy_hi, double cy_lo, uint32_t max_iter,
                         double bailout_sq, Real *orbit) {
    const DD cx{cx_hi, cx_lo}, cy{cy_hi, cy_lo};
    DD x{0.0, 0.0}, y{0.0, 0.0};
    uint32_t n = 0;
    for (;;) {
        const double x2 = x.hi * x.hi;
        const double y2 = y.hi * y.hi;
        const double mag = x2 + y2;
        orbit[4 * n + 0] = (Real)x.hi;
        orbit[4 * n + 1] = (Real)y.hi;
        orbit[4 * n + 2] = (Real)(0.5 * mag);
        orbit[4 * n + 3] = (Real)0;
        if (!(mag < bailout_sq) || n == max_iter) return n;
        orbit_step(x, y, cx, cy);
        ++n;
    }
}

inline float fma_r(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fma_r(double a, double b, double c) { return std::fma(a, b, c); }

// One pixel of the perturbation (the checked form; the gfx950 kernel's
// unchecked blocks replay to exactly this sequence).  flags bit 0 = rebase
// (Zhuoran: restart the reference when |z| < |d|, tested as
// fma(Zx, dx, Zy * dy) < -|Z|^2/2); without it, Pauldelbrot's test
// |z|^2 < glitch_eps * |Z|^2 (m > 0) stops the pixel as glitched.  Either way
// the reference restarts from Z_0 when the pixel outlives it (m == M).
// Returns the iterations executed; *state = 1 escaped, 2 max_iter, 3 glitched.
template <class Real>
uint32_t perturb_pixel(const Real *orbit, uint32_t M, Real dcx, Real dcy, uint32_t max_iter,
                       Real bailout_sq, uint32_t flags, Real glitch_eps, Real *mag_out,
                       int *state) {
    const bool rebase = (flags & 1u) != 0;
    Real dx = 0, dy = 0;
    uint32_t m = 0, n = 0;
    for (;;) {
        Real Zx = orbit[4 * m + 0], Zy = orbit[4 * m + 1];
        const Real h = orbit[4 * m + 2];
        const Real zx = Zx + dx;
        const Real zy = Zy + dy;
        const Real mag = fma_r(zx, zx, zy * zy);
        if (!(n < max_iter)) {
            *state = 2;
            *mag_out = 0;
            return n;
        }
        if (!(mag < bailout_sq)) {
            *state = 1;
            *mag_out = mag;
            return n;
        }
        if (rebase) {
            if (fma_r(Zx, dx, Zy * dy) < -h || m == M) {
                dx = zx;
                dy = zy;
                m = 0;
                Zx = 0;
                Zy = 0;
            }
        } else if (m == M) {
            dx = zx;
            dy = zy;
            m = 0;
            Zx = 0;
            Zy = 0;
        } else if (m > 0 && mag < glitch_eps * (h + h)) {
            *state = 3;
            *mag_out = mag;
            return n;
        }
        const Real ax = fma_r((Real)2, Zx, dx);
        const Real ay = fma_r((Real)2, Zy, dy);
        const Real ndx = fma_r(ax, dx, fma_r(-ay, dy, dcx));
        const Real ndy = fma_r(ax, dy, fma_r(ay, dx, dcy));
        dx = ndx;
        dy = ndy;
        ++m;
        ++n;
    }
}

}  // namespace

extern "C" {

// (a_hi + a_lo) + (b_hi + b_lo) in double-double.
void or_dd_add(double a_hi, double a_lo, double b_hi, double b_lo, double *hi, double *lo) {
    const DD r = dd_add({a_hi, a_lo}, {b_hi, b_lo});
    *hi = r.hi;
    *lo = r.lo;
}

uint32_t or_reference_orbit_f(double cx_hi, double cx_lo, double cy_hi, double cy_lo,
                              uint32_t max_iter, double bailout_sq, float *orbit) {
    return reference_orbit<float>(cx_hi, cx_lo, cy_hi, cy_lo, max_iter, bailout_sq, orbit);
}
uint32_t or_reference_orbit_d(double cx_hi, double cx_lo, double cy_hi, double cy_lo,
                              uint32_t max_iter, double bailout_sq, double *orbit) {
    return reference_orbit<double>(cx_hi, cx_lo, cy_hi, cy_lo, max_iter, bailout_sq, orbit);
}
uint32_t or_perturb_pixel_f(const float *orbit, uint32_t M, float dcx, float dcy,
                            uint32_t max_iter, float bailout_sq, uint32_t flags, float glitch_eps,
                            float *mag_out, int *state) {
    return perturb_pixel<float>(orbit, M, dcx, dcy, max_iter, bailout_sq, flags, glitch_eps,
                                mag_out, state);
}
uint32_t or_perturb_pixel_d(const double *orbit, uint32_t M, double dcx, double dcy,
                            uint32_t max_iter, double bailout_sq, uint32_t flags,
                            double glitch_eps, double *mag_out, int *state) {
    return perturb_pixel<double>(orbit, M, dcx, dcy, max_iter, bailout_sq, flags, glitch_eps,
                                 mag_out, state);
}

// Plane offset of pixel (px, py) from the frame centre.
void or_deep_offset(uint32_t px, uint32_t py, uint32_t width, uint32_t height, double pixel_w,
                    double pixel_h, double *ox, double *oy) {
    *ox = ((double)px + 0.5 - (double)width * 0.5) * pixel_w;
    *oy = ((double)height * 0.5 - ((double)py + 0.5)) * pixel_h;
}

// Direct FP64 iteration at c = C + offset (the reference loop, bailout_sq).
uint32_t or_deep_direct_d(double cx, double cy, double ox, double oy, uint32_t max_iter,
                          double bailout_sq, double *mag_out, int *escaped) {
    return or_escape_d(max_iter, 0.0, 0.0, cx + ox, cy + oy, bailout_sq, mag_out, escaped);
}

// Direct double-double iteration at c = (cx_hi + cx_lo) + ox exactly (likewise y):
// the oracle beyond FP64 range.  Escape test on the high parts.
uint32_t or_deep_direct_dd(double cx_hi, double cx_lo, double cy_hi, double cy_lo, double ox,
                           double oy, uint32_t max_iter, double bailout_sq, double *mag_out,
                           int *escaped) {
    const DD cx = dd_add({cx_hi, cx_lo}, {ox, 0.0});
    const DD cy = dd_add({cy_hi, cy_lo}, {oy, 0.0});
    DD x{0.0, 0.0}, y{0.0, 0.0};
    uint32_t n = 0;
    *escaped = 0;
    *mag_out = 0.0;
    while (n < max_iter) {
        const double mag = x.hi * x.hi + y.hi * y.hi;
        if (!(mag < bailout_sq)) {
            *escaped = 1;
            *mag_out = mag;
            break;
        }
        const DD xx = dd_mul(x, x);
        const DD yy = dd_mul(y, y);
        const DD xy = dd_mul(x, y);
        const DD xn = dd_add(dd_add(xx, dd_neg(yy)), cx);
        y = dd_add(dd_twice(xy), cy);
        x = xn;
        ++n;
    }
    return n;
}

// A set of pixels of a deep frame by each method (OpenMP over pixels).
// method 0: FP32 perturbation against the float `orbit` (reference at centre +
// (ref_ox, ref_oy)); 3: FP64 perturbation against the double `orbit`;
// 1: direct FP64; 2: direct double-double.  counts follow the
// Mandelbrot convention (max_iter -> 0); mag = |z|^2 at escape (0 otherwise);
// state as or_perturb_pixel (methods 1/2: 1 escaped, 2 max_iter).
uint64_t or_deep_pixels(int method, const uint32_t *px, const uint32_t *py, uint64_t n_pixels,
                        uint32_t width, uint32_t height, double cx_hi, double cx_lo, double cy_hi,
                        double cy_lo, double pixel_w, double pixel_h, uint32_t max_iter,
                        double bailout_sq, const void *orbit, uint32_t M, double ref_ox,
                        double ref_oy, uint32_t flags, float glitch_eps, uint32_t *counts,
                        double *mag, int32_t *state, int threads) {
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total) num_threads(threads)
    for (int64_t k = 0; k < (int64_t)n_pixels; ++k) {
        double ox, oy;
        or_deep_offset(px[k], py[k], width, height, pixel_w, pixel_h, &ox, &oy);
        uint32_t it = 0;
        int st = 0;
        double mg = 0.0;
        if (method == 0) {
            const float dcx = (float)(ox - ref_ox), dcy = (float)(oy - ref_oy);
            float mf = 0.0f;
            it = perturb_pixel<float>((const float *)orbit, M, dcx, dcy, max_iter,
                                      (float)bailout_sq, flags, glitch_eps, &mf, &st);
            mg = mf;
        } else if (method == 3) {
            it = perturb_pixel<double>((const double *)orbit, M, ox - ref_ox, oy - ref_oy, max_iter,
                                       bailout_sq, flags, (double)glitch_eps, &mg, &st);
        } else {
            int esc = 0;
            it = method == 1 ? or_deep_direct_d(cx_hi, cy_hi, ox, oy, max_iter, bailout_sq, &mg, &esc)
                             : or_deep_direct_dd(cx_hi, cx_lo, cy_hi, cy_lo, ox, oy, max_iter,
                                                 bailout_sq, &mg, &esc);
            st = esc ? 1 : 2;
        }
        total += it;
        counts[k] = st == 2 ? 0u : it;
        mag[k] = st == 2 ? 0.0 : mg;
        state[k] = st;
    }
    return total;
}

}  // extern "C"
