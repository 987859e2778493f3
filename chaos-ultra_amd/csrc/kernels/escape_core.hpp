// escape_core.hpp — the quadratic-family orbit z <- z^2 + c for gfx950.
//
// Rounding follows the reference's loop (src/main/cuda/fractals/mandelbrot.cu:16-20)
// with contraction off (the whole tree is built with -ffp-contract=off):
//     test   x*x + y*y < bailout_sq
//     x'   = (x*x - y*y) + cx
//     y'   = ((2*x)*y) + cy
// The orbit runs in unchecked blocks of K steps.  Once |z|^2 >= R^2 with R >= 2
// and |c| <= R the orbit stays outside (|z'| >= |z|^2 - |c| >= |z|(|z|-1) >= |z|),
// and for Mandelbrot (z0 = 0) a c with |c| > R leaves at step 1 and stays out,
// so one test at the end of a block tells whether the orbit left inside it.  A
// block whose end test fails is replayed step by step, which yields the exact
// escape step and the |z|^2 of the failing test; inf/NaN produced after the
// escape inside the discarded block fail the test as well.  The block keeps its
// start and its midpoint: by the same permanence, a midpoint that is still
// inside means the escape lies in the second half, so a replay runs at most K/2
// checked steps.
#pragma once
#include "cu_types.hpp"

#pragma clang fp contract(off)

// Wave-wide sum over the active lanes (ROCm device library, DPP based).
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(
    unsigned long long);

namespace cu {

// Fused multiply-add with one rounding in the operands' own precision
// (__builtin_fma alone is the double builtin: float operands would be promoted).
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One step.  kFused computes y' as fma(x*y, 2, cy): 2*(x*y) is exact, so this
// equals ((2x)*y)+cy whenever x*y is zero or at least the smallest normal —
// one VALU op fewer (6 per step instead of 7).
template <class Real, bool kFused>
__device__ __forceinline__ void quad_step(Real &x, Real &y, Real cx, Real cy) {
    const Real x2 = x * x;
    const Real y2 = y * y;
    const Real xn = (x2 - y2) + cx;
    if constexpr (kFused) {
        y = fma_r(x * y, (Real)2, cy);
    } else {
        y = ((x + x) * y) + cy;
    }
    x = xn;
}

// One step with the squares carried: (x2, y2) hold x*x and y*y of the current
// state on entry and of the new state on exit, so a block-end test reads them
// instead of recomputing them.  Same operations and roundings as quad_step.
template <class Real, bool kFused>
__device__ __forceinline__ void quad_step_sq(Real &x, Real &y, Real &x2, Real &y2, Real cx,
                                             Real cy) {
    const Real xn = (x2 - y2) + cx;
    if constexpr (kFused) {
        y = fma_r(x * y, (Real)2, cy);
    } else {
        y = ((x + x) * y) + cy;
    }
    x = xn;
    x2 = x * x;
    y2 = y * y;
}

// Where the replay of a failed block starts: at its midpoint (iteration
// i + half) when that state is still inside, else at its start.
template <class Real>
__device__ __forceinline__ void replay_start(Real &x, Real &y, uint &i, Real sx, Real sy, Real mx,
                                             Real my, Real bailout_sq, uint half) {
    if (mx * mx + my * my < bailout_sq) {
        x = mx;
        y = my;
        i += half;
    } else {
        x = sx;
        y = sy;
    }
}

// Checked steps from (x, y, i) until the bailout test fails or i reaches
// max_iter: the reference loop, for lanes that cannot use an unchecked block.
template <class Real, bool kFused>
__device__ __forceinline__ void checked_steps(Real x, Real y, Real cx, Real cy, uint &i,
                                              uint max_iter, Real bsq, bool &escaped, Real &mag2) {
    while (i < max_iter) {
        const Real x2 = x * x;
        const Real y2 = y * y;
        const Real m = x2 + y2;
        if (!(m < bsq)) {
            escaped = true;
            mag2 = m;
            return;
        }
        const Real xn = (x2 - y2) + cx;
        if constexpr (kFused) {
            y = fma_r(x * y, (Real)2, cy);
        } else {
            y = ((x + x) * y) + cy;
        }
        x = xn;
        ++i;
    }
}

// Exact escape of a lane whose unchecked block from (sx, sy) at iteration i
// failed its end test, (x, y) being the block's end state and (mx, my) its
// midpoint.  By escape permanence the first failing test lies in the half whose
// start is inside: at iteration i0 + k for some k in [0, K/2), or at i0 + K/2
// itself, where the failing |z|^2 is the one already known (end or midpoint).
// The K/2 tests run unrolled without branches; every lane steps the same
// number of times, so the wave spends one straight pass on its pending lanes.
// Returns n, the iteration of the failing test, and its |z|^2.
template <class Real, int K, bool kFused>
__device__ __forceinline__ uint locate_escape(Real x, Real y, Real sx, Real sy, Real mx, Real my,
                                              Real cx, Real cy, uint i, Real bsq, Real &mag2) {
    const Real m_end = x * x + y * y;
    const Real m_mid = mx * mx + my * my;
    const bool mid_in = m_mid < bsq;
    Real zx = mid_in ? mx : sx;
    Real zy = mid_in ? my : sy;
    const uint i0 = mid_in ? i + (uint)K / 2 : i;
    uint n = i0 + (uint)K / 2;
    mag2 = mid_in ? m_end : m_mid;
    bool found = false;
#pragma unroll
    for (int k = 0; k < K / 2; ++k) {
        const Real x2 = zx * zx;
        const Real y2 = zy * zy;
        const Real m = x2 + y2;
        const bool first = !found && !(m < bsq);
        n = first ? i0 + (uint)k : n;
        mag2 = first ? m : mag2;
        found = found || first;
        if (k + 1 < K / 2) {
            const Real xn = (x2 - y2) + cx;
            if constexpr (kFused) {
                zy = fma_r(zx * zy, (Real)2, cy);
            } else {
                zy = ((zx + zx) * zy) + cy;
            }
            zx = xn;
        }
    }
    return n;
}

// Result of one orbit.
template <class Real>
struct Orbit {
    uint n;        // iterations executed
    bool escaped;  // stopped on the bailout test (else n == max_iter)
    Real mag2;     // x*x + y*y of the failing test when escaped
};

// The reference loop  while (i < maxIter && x*x+y*y < R^2) { step; ++i; }
// with unchecked blocks of K steps when `blocks` holds (escape is permanent for
// these inputs) and a checked tail/replay.
template <class Real, int K, bool kFused>
__device__ __forceinline__ Orbit<Real> escape_orbit(Real x, Real y, Real cx, Real cy,
                                                     uint max_iter, Real bailout_sq,
                                                     bool blocks) {
    uint i = 0;
    if (blocks) {
        while (max_iter - i >= (uint)K) {
            const Real sx = x, sy = y;
#pragma unroll
            for (int k = 0; k < K / 2; ++k) quad_step<Real, kFused>(x, y, cx, cy);
            const Real mx = x, my = y;
#pragma unroll
            for (int k = K / 2; k < K; ++k) quad_step<Real, kFused>(x, y, cx, cy);
            if (!(x * x + y * y < bailout_sq)) {
                replay_start(x, y, i, sx, sy, mx, my, bailout_sq, (uint)K / 2);
                break;
            }
            i += K;
        }
    }
    Orbit<Real> r;
    r.escaped = false;
    r.mag2 = 0;
    while (i < max_iter) {
        const Real x2 = x * x;
        const Real y2 = y * y;
        const Real m = x2 + y2;
        if (!(m < bailout_sq)) {
            r.escaped = true;
            r.mag2 = m;
            break;
        }
        const Real xn = (x2 - y2) + cx;
        if constexpr (kFused) {
            y = fma_r(x * y, (Real)2, cy);
        } else {
            y = ((x + x) * y) + cy;
        }
        x = xn;
        ++i;
    }
    r.n = i;
    return r;
}

// Whether unchecked blocks are valid: R >= 2 and, when c is fixed (Julia), |c| <= R.
__device__ __forceinline__ bool blocks_valid(double bailout_sq, bool fixed_c, double cx,
                                             double cy) {
    if (!(bailout_sq >= 4.0)) return false;
    if (!fixed_c) return true;
    return cx * cx + cy * cy <= bailout_sq * 0.999;
}

// log2 of a positive normal double in about 35 VALU operations.  v = m * 2^e
// with m in [sqrt(1/2), sqrt(2)); with s = (m - 1) / (m + 1),
//     ln m = 2 atanh(s) = 2s + sum_{k>=1} 2 s^(2k+1) / (2k + 1),
// and |s| <= 0.1716 (s^2 <= 0.0295), so the ten terms kept leave a truncation
// error below 1e-17 relative.  m - 1 is exact; m + 1 and the Newton-refined
// quotient round once each, so the result is within a few ulp of log2(v).
// Zero, subnormal, infinite, NaN and negative inputs take the library log2.
__device__ __forceinline__ double log2_positive(double v) {
    int e = __builtin_amdgcn_frexp_exp(v);  // v = mant * 2^e, mant in [0.5, 1)
    double m = __builtin_amdgcn_frexp_mant(v);
    const bool low = m < 0.70710678118654752440;
    m = low ? m + m : m;
    e = low ? e - 1 : e;
    const double f = m - 1.0;
    const double d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    r = fma_r(r, fma_r(-d, r, 1.0), r);
    r = fma_r(r, fma_r(-d, r, 1.0), r);
    double s = f * r;
    s = fma_r(fma_r(-s, d, f), r, s);
    const double z = s * s;
    double p = 2.0 / 21.0;
    p = fma_r(p, z, 2.0 / 19.0);
    p = fma_r(p, z, 2.0 / 17.0);
    p = fma_r(p, z, 2.0 / 15.0);
    p = fma_r(p, z, 2.0 / 13.0);
    p = fma_r(p, z, 2.0 / 11.0);
    p = fma_r(p, z, 2.0 / 9.0);
    p = fma_r(p, z, 2.0 / 7.0);
    p = fma_r(p, z, 2.0 / 5.0);
    p = fma_r(p, z, 2.0 / 3.0);
    const double ln_m = fma_r(s * z, p, s + s);
    double out = fma_r(ln_m, 1.44269504088896340736, (double)e);  // log2(e)
    if (!(v >= 0x1p-1022 && v <= 0x1.fffffffffffffp+1023)) out = log2(v);
    return out;
}

// n + 1 - log2(ln(mag2) / ln(R^2)) — the smooth (continuous) escape value —
// evaluated as n + 1 - (log2(log2(mag2)) - C) with the per-tile constant
// C = log2(log2(R^2)) (smooth_constant): two log2 and no division per sample.
// The two forms differ only in rounding, by about 1e-15 before n + 1 is added.
__device__ __forceinline__ double smooth_value(uint n, double mag2, double smooth_c) {
    return (double)n + 1.0 - (log2_positive(log2_positive(mag2)) - smooth_c);
}

__device__ __forceinline__ double smooth_constant(double bailout_sq) {
    return log2(log2(bailout_sq));
}

// Palette slot of a colour value: L - (round(v) mod L) - 1, round half away
// from zero (colorize, mandelbrot.cu:27-35).  A value that is not above zero
// takes the slot of 0: the reference's counts never are, but a smooth value is
// negative for an orbit that starts far outside (|z0| or |c| far beyond R), and
// converting a negative value to an unsigned index is undefined.
__device__ __forceinline__ uint palette_slot(double v, uint length) {
    const uint vi = v > 0.0 ? (uint)round(v) : 0u;
    return length - (vi % length) - 1;
}
__device__ __forceinline__ uint palette_slot(float v, uint length) {
    const uint vi = v > 0.0f ? (uint)roundf(v) : 0u;
    return length - (vi % length) - 1;
}

// SplitMix64 (Steele, Lea, Flood 2014): output j of the generator seeded with
// `seed` is mix(seed + (j+1) * 0x9E3779B97F4A7C15).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Jitter of sample j: output j of SplitMix64(seed); dx from its high 32 bits,
// dy from its low 32 bits, each scaled by 2^-32 into [0, 1).
__device__ __forceinline__ void jitter(uint64_t seed, uint64_t sample_id, double &dx, double &dy) {
    const uint64_t h = splitmix64(seed + sample_id * 0x9E3779B97F4A7C15ull);
    dx = (double)(uint32_t)(h >> 32) * 0x1.0p-32;
    dy = (double)(uint32_t)h * 0x1.0p-32;
}

// Offset of the sample numbered `sample_id` = (py * width + px) * spp + s in its tile.
__device__ __forceinline__ void sample_offset_of(const cu_tile_desc &t, uint64_t sample_id,
                                                 double &dx, double &dy) {
    if (t.pattern == CU_SAMPLE_CENTRE) {
        dx = 0.5;
        dy = 0.5;
    } else if (t.pattern == CU_SAMPLE_CORNER) {
        dx = 0.0;
        dy = 0.0;
    } else {
        jitter(t.seed, sample_id, dx, dy);
    }
}

__device__ __forceinline__ void sample_offset(const cu_tile_desc &t, uint px, uint py, uint s,
                                              double &dx, double &dy) {
    sample_offset_of(t, ((uint64_t)py * t.width + px) * t.spp + s, dx, dy);
}

// X = left + (px + dx) * pixel_w,  Y = top - (py + dy) * pixel_h
template <class Real>
__device__ __forceinline__ void tile_point(const cu_tile_desc &t, uint px, uint py, double dx,
                                           double dy, Real &X, Real &Y);

template <>
__device__ __forceinline__ void tile_point<double>(const cu_tile_desc &t, uint px, uint py,
                                                   double dx, double dy, double &X, double &Y) {
    X = t.left + ((double)px + dx) * t.pixel_w;
    Y = t.top - ((double)py + dy) * t.pixel_h;
}

template <>
__device__ __forceinline__ void tile_point<float>(const cu_tile_desc &t, uint px, uint py,
                                                  double dx, double dy, float &X, float &Y) {
    X = t.left_f + ((float)px + (float)dx) * t.pixel_w_f;
    Y = t.top_f - ((float)py + (float)dy) * t.pixel_h_f;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint lane_id() { return __lane_id(); }
// The iteration counter is CU_ITER_SLOTS u64 slots; each wave adds its total
// to one slot chosen by its global wave number, so a launch of a million waves
// does not serialise on one address.  The caller sums the slots.  All 64 lanes
// must be active: the wave sum is the device library's DPP reduction.
__device__ __forceinline__ void add_iterations(unsigned long long *iters, uint64_t v) {
    v = __ockl_wfred_add_u64(v);
    const uint block = blockIdx.x + blockIdx.y * gridDim.x;
    const uint wave = block * ((blockDim.x * blockDim.y + 63u) >> 6) +
                      ((threadIdx.x + threadIdx.y * blockDim.x) >> 6);
    if (iters && __lane_id() == 0 && v)
        atomicAdd(iters + (wave & (CU_ITER_SLOTS - 1u)), (unsigned long long)v);
}
// Column of pixel blocks this workgroup covers in a 2-D grid over a frame.
// Consecutive workgroups go round-robin to the 8 XCDs and then to their CUs, so
// with a grid 64 blocks wide every CU's workgroups share blockIdx.x: a CU works
// one column of the frame, and the CUs over the set's interior (orbits that run
// all max_iter steps) end last while the rest idle.  Each row of blocks is
// rotated by a golden-ratio fraction of the width, so a CU's workgroups land on
// columns spread over the frame.  A rotation per row maps blocks one to one.
__device__ __forceinline__ uint skewed_block_x() {
    const uint gx = gridDim.x;
    const uint skew = (uint)(((uint64_t)gx * 40503u) >> 16);  // 0.618 * gx
    return (uint)(((uint64_t)blockIdx.x + (uint64_t)blockIdx.y * skew) % gx);
}
// This lane's bit of a wave-uniform lane mask (the inverse of ballot): the mask
// becomes the execution mask itself, with no per-lane shift and compare.
__device__ __forceinline__ bool in_mask(uint64_t mask) {
    return __builtin_amdgcn_inverse_ballot_w64(mask);
}

}  // namespace cu
