// perturb_core.hpp — the deep-zoom perturbation form's device code: the
// double-double reference orbit, the perturbation step with rebasing and
// glitch detection, and the per-pixel frame kernels' bodies.  Included by
// perturb.hip (the escape_family module's entry points) and by the library's
// allocation-free deep render (csrc/host/deep_static.hip).
//
// Below a pixel of ~1e-13 of the plane the reference's double kernel cannot
// tell pixels apart (cudarenderer/RenderingKernel.java:124-138 and
// cudarenderer/CudaFractalRenderer.java:217-220 stop there).  This form
// iterates every pixel as a small delta around one reference orbit:
//
//   z_n = Z_m + d_n,   d_{n+1} = (2 Z_m + d_n) d_n + dc,   m <- m + 1
//
// The reference orbit Z is computed once per frame on the GPU by one lane in
// double-double (FP64 error-free transformations, ~106-bit significand): an
// orbit computed in plain FP64 drifts from the true orbit long before the
// pixel scale of a 1e-20 frame.  It is stored rounded to the delta type,
// each entry {Zx, Zy, |Z|^2/2, 0}.  Deltas are FP32 (twice the FP64 issue rate
// on gfx950) or FP64 (far more precise).
//
// Rebasing (Zhuoran): when |z| < |d| — tested as fma(Zx, dx, Zy dy) <
// -|Z|^2/2 — or when the pixel outlives the reference (m == M), the pixel
// restarts the reference from Z_0 = 0 with d = z.  Without rebasing,
// Pauldelbrot's test |z|^2 < eps |Z|^2 marks a pixel glitched and the host
// re-renders glitched pixels around a secondary reference.
//
// Hot loop: unchecked blocks of K steps (one orbit load and 8 FMA-class
// VALU ops per step).  A block ends early at the first step where some lane's
// rebase test fires, and those lanes take the rebase there; escape is tested
// once, at the block's end (escape is permanent for bailout_sq >= 4, |c| <= 2),
// and only a block with an escape is replayed with checked steps — the same
// operation sequence as the oracle's checked loop.
#pragma once
#include "cu_types.hpp"
#include "escape_core.hpp"

#pragma clang fp contract(off)

namespace {

// ---- double-double (Dekker / Knuth error-free transformations) ----
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_quick_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}

__device__ __forceinline__ DD dd_two_prod(double a, double b) {
    const double p = a * b;
    return {p, __builtin_fma(a, b, -p)};  // double operands
}

// One step Z <- Z^2 + C of the reference orbit, compensated rather than composed
// from dd_mul and dd_add: the leading products are split exactly, the sums of
// leading terms are exact (two-sums), and the low-order terms (the products'
// errors, the cross terms 2 xh xl - 2 yh yl and 2 (xh yl + xl yh), C's low parts)
// are summed in plain double before one renormalisation; terms of order xl^2 are
// dropped.  The absolute error per step is a few units of 2^-104 |Z|^2, as for
// the composed form, in about half its instructions: the orbit is one lane's
// serial chain, issued one instruction at a time.
__device__ __forceinline__ void dd_orbit_step(DD &x, DD &y, DD cx, DD cy) {
    const DD px = dd_two_prod(x.hi, x.hi), py = dd_two_prod(y.hi, y.hi);
    const DD pxy = dd_two_prod(x.hi, y.hi);
    const DD s1 = dd_two_sum(px.hi, -py.hi);
    const DD s2 = dd_two_sum(s1.hi, cx.hi);
    const double u = __builtin_fma(x.hi, x.lo, -(y.hi * y.lo));
    const double xlo = ((s1.lo + s2.lo) + (px.lo - py.lo)) + ((u + u) + cx.lo);
    const DD s3 = dd_two_sum(pxy.hi + pxy.hi, cy.hi);
    const double w = __builtin_fma(x.hi, y.lo, x.lo * y.hi);
    const double ylo = (s3.lo + (pxy.lo + pxy.lo)) + ((w + w) + cy.lo);
    x = dd_quick_two_sum(s2.hi, xlo);
    y = dd_quick_two_sum(s3.hi, ylo);
}

template <class Real>
struct alignas(4 * sizeof(Real)) OrbitEntry {
    Real x, y, h, pad;  // Z_m and |Z_m|^2 / 2
};

// Pointers into LDS and global memory typed with their address spaces, so that
// loads through them stay LDS or global (or scalar) loads: through untyped
// pointers the compiler may merge reads from different branches into one
// generic (flat) load of a selected address.
template <class T>
using lds_ptr = const T __attribute__((address_space(3))) *;
template <class T>
using global_ptr = const T __attribute__((address_space(1))) *;

template <class Real, class Ptr>
__device__ __forceinline__ OrbitEntry<Real> load_entry(Ptr p) {
    OrbitEntry<Real> e;
    e.x = p->x;
    e.y = p->y;
    e.h = p->h;
    e.pad = 0;
    return e;
}

// The reference orbit as the pixel kernels read it: entries below `staged`
// from the block's LDS copy, the rest (an orbit longer than the LDS budget)
// from global memory.
template <class Real>
struct OrbitView {
    global_ptr<OrbitEntry<Real>> global;
    lds_ptr<OrbitEntry<Real>> lds;
    uint staged;
    __device__ __forceinline__ OrbitEntry<Real> at(uint m) const {
        OrbitEntry<Real> e = load_entry<Real>(lds + (m < staged ? m : 0u));
        if (m >= staged) e = load_entry<Real>(global + m);
        return e;
    }
};

template <class Real>
__device__ void reference_orbit(double cx_hi, double cx_lo, double cy_hi, double cy_lo,
                                uint max_iter, double bailout_sq, OrbitEntry<Real> *orbit,
                                uint *length) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const DD cx{cx_hi, cx_lo}, cy{cy_hi, cy_lo};
    DD x{0.0, 0.0}, y{0.0, 0.0};
    uint n = 0;
    for (;;) {
        const double x2 = x.hi * x.hi;
        const double y2 = y.hi * y.hi;
        const double mag = x2 + y2;
        OrbitEntry<Real> e;
        e.x = (Real)x.hi;
        e.y = (Real)y.hi;
        e.h = (Real)(0.5 * mag);
        e.pad = 0;
        orbit[n] = e;
        if (!(mag < bailout_sq) || n == max_iter) break;
        dd_orbit_step(x, y, cx, cy);
        ++n;
    }
    *length = n;
}

struct DeepArgs {
    const void *orbit;         // OrbitEntry<Real>[M + 1]
    const uint *orbit_len;     // M
    uint *counts;              // W x H, row-major
    double *smooth;
    unsigned char *glitch;     // 1 = glitched (Pauldelbrot) in the last pass over the pixel
    uint *rgba;                // optional palette colour
    const uint *palette;
    unsigned long long *iters; // CU_ITER_SLOTS slots
    double centre_x, centre_y; // direct form only: c = centre + offset
    double pixel_w, pixel_h;
    double ref_ox, ref_oy;     // reference point - frame centre (plane units)
    double bailout_sq;
    double smooth_c;           // log2(log2(bailout_sq))
    uint width, height, max_iter, flags;
    uint palette_length;
    float glitch_eps;
    uint only_glitched;        // 1: pixels whose glitch flag is set, others untouched
    uint pad;
};
static_assert(sizeof(DeepArgs) == 160, "DeepArgs layout");

constexpr uint kDeepRebase = 1u;

__device__ __forceinline__ void deep_offset(const DeepArgs &a, uint px, uint py, double &ox,
                                            double &oy) {
    ox = ((double)px + 0.5 - (double)a.width * 0.5) * a.pixel_w;
    oy = ((double)a.height * 0.5 - ((double)py + 0.5)) * a.pixel_h;
}

// The rebase test |Z + d| < |d|, i.e. |Z|^2 + 2 Z.d < 0, evaluated as
// Zx dx + Zy dy < -h with h = |Z|^2 / 2 from the orbit entry: one product, one
// FMA and a compare.  Every operand taken from the orbit enters a different
// instruction, so a wave that reads the orbit through scalar registers needs no
// copies into vector registers.  At Z = 0 (m = 0, h = 0) the sum is a zero and
// the test does not fire for finite d.
template <class Real>
__device__ __forceinline__ bool rebase_test(Real Zx, Real Zy, Real h, Real dx, Real dy) {
    return cu::fma_r(Zx, dx, Zy * dy) < -h;
}

// One checked step of the perturbation loop (the oracle's or_perturb_pixel
// body).  Returns 0 to continue, 1 escaped (mag set), 2 max_iter, 3 glitched.
template <class Real>
__device__ __forceinline__ int checked_step(const OrbitView<Real> &Z, uint M, Real &dx, Real &dy,
                                            uint &m, uint &n, Real dcx, Real dcy, uint max_iter,
                                            Real bsq, bool rebase, Real geps, Real &mag_out) {
    const OrbitEntry<Real> e = Z.at(m);
    Real Zx = e.x, Zy = e.y;
    const Real zx = Zx + dx;
    const Real zy = Zy + dy;
    const Real mag = cu::fma_r(zx, zx, zy * zy);
    if (!(n < max_iter)) return 2;
    if (!(mag < bsq)) {
        mag_out = mag;
        return 1;
    }
    if (rebase) {
        if (rebase_test(Zx, Zy, e.h, dx, dy) || m == M) {
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
    } else if (m > 0 && mag < geps * (e.h + e.h)) {
        mag_out = mag;
        return 3;
    }
    const Real ax = cu::fma_r((Real)2, Zx, dx);
    const Real ay = cu::fma_r((Real)2, Zy, dy);
    const Real ndx = cu::fma_r(ax, dx, cu::fma_r(-ay, dy, dcx));
    const Real ndy = cu::fma_r(ax, dy, cu::fma_r(ay, dx, dcy));
    dx = ndx;
    dy = ndy;
    ++m;
    ++n;
    return 0;
}

// K unchecked steps from orbit entry zp[0]; returns the ballot of lanes whose
// rebase test fired at any of the K states.  No branches: the orbit loads of the
// whole block can be issued ahead of the arithmetic.
template <class Real, int K, class Ptr>
__device__ __forceinline__ uint64_t unchecked_block(Ptr zp, Real &dx, Real &dy, Real dcx,
                                                    Real dcy) {
    uint64_t flagged = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const OrbitEntry<Real> e = load_entry<Real>(zp + k);
        flagged |= cu::ballot(rebase_test(e.x, e.y, e.h, dx, dy));
        const Real ax = cu::fma_r((Real)2, e.x, dx);
        const Real ay = cu::fma_r((Real)2, e.y, dy);
        const Real ndx = cu::fma_r(ax, dx, cu::fma_r(-ay, dy, dcx));
        const Real ndy = cu::fma_r(ax, dy, cu::fma_r(ay, dx, dcy));
        dx = ndx;
        dy = ndy;
    }
    return flagged;
}

// The same steps, stopping at the first step k0 at which some lane's rebase test
// fires: every lane is left in its state at step k0, `fired` holds the lanes whose
// test fired there and `end` is orbit entry zp[k0].  Returns k0 (K, with `end`
// = zp[K], when no test fires).  Only lanes in `running` count.  Its branches keep
// the orbit reads from being issued ahead, so it runs after unchecked_block
// reported a firing, or directly while rebases come in consecutive blocks.
template <class Real, int K, class Ptr>
__device__ __forceinline__ uint stopping_block(Ptr zp, Real &dx, Real &dy, Real dcx, Real dcy,
                                               uint64_t running, uint64_t &fired,
                                               OrbitEntry<Real> &end) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const OrbitEntry<Real> e = load_entry<Real>(zp + k);
        fired = cu::ballot(rebase_test(e.x, e.y, e.h, dx, dy)) & running;
        if (fired != 0) {
            end = e;
            return (uint)k;
        }
        const Real ax = cu::fma_r((Real)2, e.x, dx);
        const Real ay = cu::fma_r((Real)2, e.y, dy);
        const Real ndx = cu::fma_r(ax, dx, cu::fma_r(-ay, dy, dcx));
        const Real ndy = cu::fma_r(ax, dy, cu::fma_r(ay, dx, dcy));
        dx = ndx;
        dy = ndy;
    }
    end = load_entry<Real>(zp + K);
    return (uint)K;
}

// Whole pixel; lanes whose `active` is false only take part in the ballots.
//
// While every running lane can take an unchecked block (its iteration budget and
// the orbit both have K steps left), the whole wave takes one, in wave-uniform
// control flow: lanes that have finished compute along with their m held at 0,
// so every read stays in bounds, and change nothing (ballots are masked to the
// running lanes).  When the running lanes all read the same orbit entries (same
// m), the block reads them through the scalar cache; otherwise each lane reads
// its own entries from the LDS copy, or from global memory past it.  When some
// running lane cannot take a block, every running lane takes one checked step.
//
// A block ends at step k0 for the whole wave: K, or the first step at which a
// running lane's rebase test fired (found by rerunning the block in the form that
// stops there, or by running that form directly after two blocks that stopped).
// There the escape test runs on z = Z_m + d.  Escaped (at step k0, hence at some
// step of the block since escape is permanent): the lane replays the block with
// checked steps from its start.  Otherwise the lanes whose test fired take the
// rebase d = z, m = 0 with n counting the k0 steps taken, and the others simply
// stand at step k0.  The checked loop would take the same rebase at the same
// state and then step from Z_0 = 0; the next block starts there, and its first
// rebase test, at m = 0 with Z_0 = 0 and |Z_0|^2 = 0, compares a zero with -0
// and does not fire, while lanes that did not fire repeat a test that passed —
// so every block advances.
template <class Real, int K>
__device__ int perturb_pixel(const OrbitView<Real> &Z, uint M, Real dcx, Real dcy, uint max_iter,
                             Real bsq, bool rebase, Real geps, bool active, uint &n_out,
                             Real &mag_out) {
    Real dx = 0, dy = 0;
    uint m = 0, n = 0;
    int state = active ? 0 : 4;
    mag_out = 0;
    uint streak = 0;  // wave-uniform: consecutive blocks that stopped at a rebase
    for (;;) {
        const uint64_t running = cu::ballot(state == 0);
        if (running == 0) break;
        const bool can = rebase && max_iter - n >= (uint)K && M - m >= (uint)K;
        const bool block = (cu::ballot(can) & running) == running;
        bool replay = false;
        if (block) {
            const Real sx = dx, sy = dy;
            const uint mu = __builtin_amdgcn_readlane(m, (int)__builtin_ctzll(running));
            const bool same = (cu::ballot(m != mu) & running) == 0;
            const bool lds = (cu::ballot(m + (uint)K >= Z.staged) & running) == 0;
            OrbitEntry<Real> e;
            uint64_t fired = 0;
            uint k0 = K;
            // after two blocks in a row that stopped at a rebase, rebases come often
            // here: go straight to the form that stops at the first one
            bool stop = streak >= 2;
            if (!stop) {
                uint64_t flagged;
                if (same) {
                    e = load_entry<Real>(Z.global + (mu + (uint)K));
                    flagged = unchecked_block<Real, K>(Z.global + mu, dx, dy, dcx, dcy);
                } else if (lds) {
                    e = load_entry<Real>(Z.lds + (m + (uint)K));
                    flagged = unchecked_block<Real, K>(Z.lds + m, dx, dy, dcx, dcy);
                } else {
                    e = load_entry<Real>(Z.global + (m + (uint)K));
                    flagged = unchecked_block<Real, K>(Z.global + m, dx, dy, dcx, dcy);
                }
                if ((flagged & running) != 0) {  // a rebase test fired: rerun up to there
                    dx = sx;
                    dy = sy;
                    stop = true;
                }
            }
            if (stop) {
                if (same)
                    k0 = stopping_block<Real, K>(Z.global + mu, dx, dy, dcx, dcy, running, fired, e);
                else if (lds)
                    k0 = stopping_block<Real, K>(Z.lds + m, dx, dy, dcx, dcy, running, fired, e);
                else
                    k0 = stopping_block<Real, K>(Z.global + m, dx, dy, dcx, dcy, running, fired,
                                                 e);
            }
            streak = k0 < (uint)K ? streak + 1 : 0;
            if (state == 0) {
                const Real zx = e.x + dx;
                const Real zy = e.y + dy;
                const Real mag = cu::fma_r(zx, zx, zy * zy);
                if (!(mag < bsq)) {
                    dx = sx;
                    dy = sy;
                    replay = true;
                } else {
                    n += k0;
                    m += k0;
                    if (cu::in_mask(fired)) {
                        dx = zx;
                        dy = zy;
                        m = 0;
                    }
                }
            }
        }
        if (state == 0 && (!block || replay)) {
            const int steps = block ? K : 1;
            for (int s = 0; s < steps && state == 0; ++s)
                state = checked_step<Real>(Z, M, dx, dy, m, n, dcx, dcy, max_iter, bsq, rebase,
                                           geps, mag_out);
            if (state != 0) m = 0;  // finished: later blocks read from entry 0
        }
    }
    n_out = n;
    return state;
}

__device__ __forceinline__ void deep_write(const DeepArgs &a, uint64_t o, int state, uint n,
                                           double mag) {
    const uint count = (state == 1 || state == 3) ? n : 0u;  // Mandelbrot: max_iter -> 0
    const double sm = state == 1 ? cu::smooth_value(n, mag, a.smooth_c) : 0.0;
    a.counts[o] = count;
    if (a.smooth) a.smooth[o] = sm;
    if (a.glitch) a.glitch[o] = state == 3 ? 1 : 0;
    if (a.rgba) {
        a.rgba[o] = a.palette[cu::palette_slot(sm, a.palette_length)];
    }
}

// A block of WX x WY waves, each wave 8 x 8 pixels.
template <uint WX, uint WY>
__device__ __forceinline__ void deep_pixel_coords(uint &px, uint &py) {
    const uint lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    px = (blockIdx.x * WX + (wave % WX)) * 8u + (lane & 7u);
    py = (blockIdx.y * WY + (wave / WX)) * 8u + (lane >> 3);
}

// Perturbation blocks stage the first LdsBytes of the orbit in LDS and share it
// among their WX x WY waves.  A CU holds 160 KB of LDS and 32 waves, and a
// block's wave slots only free up when its slowest wave ends, so smaller blocks
// keep more waves resident; larger copies keep more of the orbit in LDS.
//   FP32: 2 x 2 waves (16 x 16 pixels, 256 threads), 20 KB = 1280 entries, 8 blocks per CU
//   FP64: 4 x 2 waves (32 x 16 pixels, 512 threads), 40 KB = 1280 entries, 4 blocks per CU
template <class Real, int K, uint WX, uint WY, uint LdsBytes>
__device__ void perturb_frame(const DeepArgs a, const OrbitEntry<Real> *__restrict__ orbit) {
    __shared__ OrbitEntry<Real> lds[LdsBytes / sizeof(OrbitEntry<Real>)];
    const uint M = *a.orbit_len;
    const uint staged = min(M + 1u, LdsBytes / (uint)sizeof(OrbitEntry<Real>));
    for (uint k = threadIdx.x; k < staged; k += blockDim.x) lds[k] = orbit[k];
    __syncthreads();
    const OrbitView<Real> view{(global_ptr<OrbitEntry<Real>>)orbit,
                               (lds_ptr<OrbitEntry<Real>>)lds, staged};
    uint px, py;
    deep_pixel_coords<WX, WY>(px, py);
    const uint64_t o = (uint64_t)py * a.width + px;
    bool active = px < a.width && py < a.height;
    if (active && a.only_glitched) active = a.glitch[o] != 0;
    Real dcx = 0, dcy = 0;
    if (active) {
        double ox, oy;
        deep_offset(a, px, py, ox, oy);
        dcx = (Real)(ox - a.ref_ox);
        dcy = (Real)(oy - a.ref_oy);
    }
    uint n = 0;
    Real mag = 0;
    const int state = perturb_pixel<Real, K>(
        view, M, dcx, dcy, a.max_iter, (Real)a.bailout_sq,
        (a.flags & kDeepRebase) != 0, (Real)a.glitch_eps, active, n, mag);
    if (active) deep_write(a, o, state, n, (double)mag);
    cu::add_iterations(a.iters, active ? (uint64_t)n : 0u);
}

// Direct FP64 at c = centre + offset: the reference loop with bailout_sq.
template <int K, bool kFused>
__device__ void deep_direct(const DeepArgs a) {
    uint px, py;
    deep_pixel_coords<2, 2>(px, py);
    const bool active = px < a.width && py < a.height;
    uint64_t executed = 0;
    if (active) {
        double ox, oy;
        deep_offset(a, px, py, ox, oy);
        const double cx = a.centre_x + ox, cy = a.centre_y + oy;
        const cu::Orbit<double> r = cu::escape_orbit<double, K, kFused>(
            0.0, 0.0, cx, cy, a.max_iter, a.bailout_sq, a.bailout_sq >= 4.0);
        executed = r.n;
        deep_write(a, (uint64_t)py * a.width + px, r.escaped ? 1 : 2, r.n, r.mag2);
    }
    cu::add_iterations(a.iters, executed);
}

}  // namespace
