// render_generic.hip — the kernels every fractal module exports, for gfx950.
//
// Built once per fractal with -DCU_FRACTAL_HEADER="fractals/<name>.hip" naming
// the fractal's file (it defines `struct Fractal` and debugFractal()), into
// <name>.hsaco, which the host
// loads with hipModuleLoad.  Kernel names, parameter lists and the module
// global VISUALIZE_SAMPLE_COUNT are the module ABI of
// src/main/cuda/fractalRendererGeneric.cu, so a host written against the
// reference binds them unchanged; the one difference is compose, whose two GL
// surface handles are plain device pointers here (a linear RGBA8 image and the
// palette), each still one 8-byte argument.
//
// Launch geometry is the reference's: blocks of 32x32 threads over
// ceil(W/32) x ceil(H/32) blocks.  A wave64 covers two of the reference's
// 32-lane warps, i.e. two 8x4-pixel rectangles side by side; each half-wave is
// one vote group of the adaptive supersampling, exactly the 8x4 rectangles the
// reference votes over.  No lane returns early: out-of-image lanes stay in the
// wave with no work, so every ballot runs in wave-uniform control flow and the
// vote membership is explicit (same half-wave, same call site, still sampling).
// A wave's orbits are pooled across its 64 pixels and their samples (see
// sample_the_fractal), so a lane whose orbit ends takes another instead of
// idling until the wave's slowest pixel is done.
#include "cu_types.hpp"
#include "escape_core.hpp"

#include <cfloat>

#ifndef CU_FRACTAL_HEADER
#error "build with -DCU_FRACTAL_HEADER=\"fractals/<name>.hip\""
#endif
#include CU_FRACTAL_HEADER

#pragma clang fp contract(off)

__constant__ bool VISUALIZE_SAMPLE_COUNT = false;

// Iterations executed by this module's render kernels, CU_ITER_SLOTS u64 slots
// (each wave adds its total to slot `wave mod CU_ITER_SLOTS`).  A host that wants
// totals zeroes the symbol, renders and sums it; kernel parameter lists are
// unchanged.  Not part of the reference's module ABI.
__device__ unsigned long long cu_iteration_slots[CU_ITER_SLOTS];

namespace {

constexpr uint kBlockSteps = 16;       // unchecked iterations per block
constexpr float kPiF = 3.14159265358979f;

// Flag bits (fractalRendererGeneric.cu:157-161)
constexpr uint kAdaptiveFlag = 1u << 0;
constexpr uint kFoveationFlag = 1u << 2;
constexpr uint kReuseFlag = 1u << 3;
constexpr uint kZoomingFlag = 1u << 4;
constexpr uint kZoomingInFlag = 1u << 5;

// getImageIndexes (fractalRendererGeneric.cu:36-53): thread -> pixel so that
// each 32-thread group covers an 8x4 rectangle of the 32x32 block tile.  The
// tile of a block is its row of the grid rotated (cu::skewed_block_x), which
// spreads the blocks a CU receives over the frame; a tile and its vote groups
// are the reference's, only the block that computes them differs.  A block of
// 32 x 4 threads (the *Strip kernels) covers 32 x 4 pixels the same way: the
// same 8x4 rectangles, and the same two of them in each wave.
__device__ __forceinline__ Point<uint> image_indexes() {
    const uint t = threadIdx.x + threadIdx.y * 32u;
    const uint warp = t >> 5, lane = t & 31u;
    Point<uint> p;
    p.x = cu::skewed_block_x() * 32u + (warp & 3u) * 8u + (lane >> 4) * 4u + (lane & 3u);
    p.y = blockIdx.y * blockDim.y + (warp >> 2) * 4u + ((lane >> 2) & 3u);
    return p;
}

__device__ __forceinline__ pixel_info_t *pixel_at(pixel_info_t *base, long pitch, uint x, uint y) {
    return (pixel_info_t *)((char *)base + (long)y * pitch) + x;
}
__device__ __forceinline__ const pixel_info_t *pixel_at(const pixel_info_t *base, long pitch,
                                                        uint x, uint y) {
    return (const pixel_info_t *)((const char *)base + (long)y * pitch) + x;
}

// One 16-byte store with the padding bytes cleared.
__device__ __forceinline__ void store_pixel(pixel_info_t *dst, const pixel_info_t &p) {
    uint4 v;
    v.x = __float_as_uint(p.value);
    v.y = __float_as_uint(p.weight);
    v.z = p.isReused ? 1u : 0u;
    v.w = __float_as_uint(p.weightOfNewSamples);
    *reinterpret_cast<uint4 *>(dst) = v;
}

__device__ __forceinline__ pixel_info_t load_pixel(const pixel_info_t *src) {
    const uint4 v = *reinterpret_cast<const uint4 *>(src);
    pixel_info_t p;
    p.value = __uint_as_float(v.x);
    p.weight = __uint_as_float(v.y);
    p.isReused = (v.z & 0xffu) != 0;
    p.weightOfNewSamples = __uint_as_float(v.w);
    return p;
}

// computeFractal<Real> of the module, bailout 4 (the reference's constant).
template <class Real>
__device__ __forceinline__ float compute_fractal(uint max_iter, Real X, Real Y, bool blocks,
                                                 uint64_t &executed) {
    Real zx, zy, cx, cy;
    Fractal::start<Real>(X, Y, zx, zy, cx, cy);
    const cu::Orbit<Real> o =
        cu::escape_orbit<Real, kBlockSteps, false>(zx, zy, cx, cy, max_iter, (Real)4, blocks);
    executed += o.n;
    return Fractal::value(o.n, max_iter);
}

// Add a wave's executed iterations to cu_iteration_slots (whole wave calls it).
__device__ __forceinline__ void count_iterations(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const uint tid = threadIdx.x + threadIdx.y * blockDim.x;
    const uint waves = (blockDim.x * blockDim.y + 63u) >> 6;
    const uint wave = (blockIdx.y * gridDim.x + blockIdx.x) * waves + (tid >> 6);
    if (cu::lane_id() == 0 && v)
        atomicAdd(cu_iteration_slots + (wave & (CU_ITER_SLOTS - 1u)), (unsigned long long)v);
}

// computeDispersion (fractalRendererGeneric.cu:24-34) over the stored samples of
// one lane's pixel (the first kAdaptiveThreshold; the reference's array holds no
// more), kept in the wave's LDS as [round][lane] so that they take no registers.
__device__ __forceinline__ float dispersion(const float (&samples)[kAdaptiveThreshold][64],
                                            uint lane, uint n, float mean) {
    float variance = 0;
    const uint stored = n < kAdaptiveThreshold ? n : kAdaptiveThreshold;
    for (uint j = 0; j < stored; ++j) {
        const float v = samples[j][lane];
        variance += (v - mean) * (v - mean);
    }
    variance /= (n - 1);
    return variance / mean;
}

// Pixel of flat thread index t of the block (getImageIndexes, :36-53).
__device__ __forceinline__ Point<uint> image_indexes_at(uint t) {
    const uint warp = t >> 5, lane = t & 31u;
    Point<uint> p;
    p.x = cu::skewed_block_x() * 32u + (warp & 3u) * 8u + (lane >> 4) * 4u + (lane & 3u);
    p.y = blockIdx.y * blockDim.y + (warp >> 2) * 4u + ((lane >> 2) & 3u);
    return p;
}

// Sample offsets of sample i of a pixel asked for sampleCountF samples (:105-117).
// The first three lie on the diagonal at i / 3, the three quotients rounded once
// at compile time; the rest use the pixel's row constants (sample_rows).
constexpr uint kSkipFirst = 2;

__device__ __forceinline__ void sample_rows(float sample_count_f, float &per_row_f,
                                           uint &per_row_i) {
    per_row_f = sample_count_f > (float)kSkipFirst ? sqrtf(sample_count_f - kSkipFirst) : 1.f;
    per_row_i = (uint)roundf(per_row_f);
}

template <class Real>
__device__ __forceinline__ void sample_offset(uint i, float per_row_f, uint per_row_i, Real &dx,
                                              Real &dy) {
    if (i <= kSkipFirst) {
        constexpr Real third = (Real)1 / (Real)(kSkipFirst + 1);
        constexpr Real two_thirds = (Real)2 / (Real)(kSkipFirst + 1);
        dx = dy = i == 0 ? (Real)0 : (i == 1 ? third : two_thirds);
    } else {
        dx = (float)((i - kSkipFirst) % per_row_i) / per_row_f;
        dy = (float)((i - kSkipFirst) / per_row_i) / per_row_f;
    }
}

// Whether round i of a pixel taking n samples holds the votes (:133, with the
// reference's && / || precedence).
__device__ __forceinline__ bool vote_round(uint i, uint n, bool adaptive) {
    return (adaptive && (i > 0 && i < kAdaptiveThreshold)) || (i == n / 2);
}

// The first round >= c (c < n) whose votes may stop the pixel: every sample up to
// it is needed whatever the votes say.
__device__ __forceinline__ uint next_vote_round(uint c, uint n, bool adaptive) {
    if (adaptive && c < kAdaptiveThreshold) return c == 0 ? (n / 2 == 0 ? 0u : 1u) : c;
    return c <= n / 2 ? n / 2 : n - 1;
}

// Rounds of results a pixel may have waiting, and the number of waiting lanes at
// which a wave stops its unchecked blocks to hand out new samples.
constexpr uint kRing = 4;
constexpr uint kRefill = 16;
constexpr unsigned long long kAbsent = ~0ull;  // above any iteration count

// One wave's LDS: each pixel's results of the rounds not yet voted on
// (iterations executed, kAbsent until posted) and the hand-out slots.
struct WaveLds {
    unsigned long long result[kRing][64];   // [sample % kRing][owner lane]
    float samples[kAdaptiveThreshold][64];  // the owner's first sample values, [round][lane]
    uint handout[64];                       // owner lane | sample << 8, by rank
};

// Orders this wave's LDS accesses across the phases below (the LDS serves a
// wave's instructions in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// sampleTheFractal<Real> (fractalRendererGeneric.cu:85-155) for the lanes with
// site != 0, whole wave, with the wave's orbits pooled.
//
// Lane k owns the pixel of lane k: its sample count, escape sum and first 10
// sample values, and its votes with the lanes sharing its half-wave and site
// (the reference's 32-lane warp and call site).  Orbits are pooled apart from
// that: a pixel's samples are independent orbits, so any lane of the wave may
// iterate sample j of any pixel of the wave.  Lanes run unchecked blocks of K
// steps; once kRefill lanes wait, the finished orbits post their iteration
// counts to the owner's LDS ring, every round whose vote group has all its
// results is voted on exactly as the round-by-round loop does, and idle lanes
// take new samples — first those every outcome of the next vote needs, then
// samples past it, up to kRing rounds ahead of the votes.  A sample that a vote
// makes unnecessary is dropped (or its posted result cleared) and never added
// to the sums or to the iteration count, so pixels, sample counts and
// iterations equal the round-by-round loop's.  Must be called by the whole wave.
template <class Real, bool kFused>
__device__ uint sample_the_fractal(WaveLds &lds, int site, Point<uint> grid, Rectangle<Real> image,
                                   uint max_iter, float &sample_count_f, bool adaptive,
                                   uint64_t &executed) {
    constexpr int K = (int)kBlockSteps;
    const uint lane = cu::lane_id();
    const uint wave_base = (threadIdx.x + threadIdx.y * blockDim.x) & ~63u;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    const uint64_t half = (lane < 32) ? 0x00000000ffffffffull : 0xffffffff00000000ull;
    const uint64_t site1 = cu::ballot(site == 1), site2 = cu::ballot(site == 2);
    const uint64_t group = half & (site == 1 ? site1 : (site == 2 ? site2 : 0ull));
    const bool blocks = Fractal::blocks_valid(4.0) && max_iter >= (uint)K;
    const Real bsq = (Real)4;

    // ---- the pixel this lane owns ----
    bool in_loop = site != 0;
    uint sample_count = 0, escape_sum = 0, result = 0;
    uint confirmed = 0, issued = 0;  // rounds voted on; samples handed out
    if (in_loop) {
        if (sample_count_f < 1) {  // :87-90
            sample_count_f = 0;
            in_loop = false;
        } else {
            sample_count = min(kMaxSuperSampling, (uint)roundf(sample_count_f));
        }
    }
    float per_row_f;  // the offsets use the caller's sampleCountF
    uint per_row_i;
    sample_rows(sample_count_f, per_row_f, per_row_i);
#pragma unroll
    for (uint r = 0; r < kRing; ++r) lds.result[r][lane] = kAbsent;
    // pixelSize = image.size() / gridSize.cast<Real>()  (:97)
    const Real psx = (image.right_top.x - image.left_bottom.x) / (Real)grid.x;
    const Real psy = (image.right_top.y - image.left_bottom.y) / (Real)grid.y;

    // ---- the orbit this lane iterates ----
    int mode = 0;  // 0 idle, 1 unchecked blocks, 2 pending (locate the escape / checked steps)
    bool replay = false;
    uint owner = lane, sample = 0, i = 0;
    Real x = 0, y = 0, cx = 0, cy = 0, sx = 0, sy = 0, mx = 0, my = 0;
    wave_lds_fence();

    for (;;) {
        // ---- vote on every round whose group has posted all its results ----
        for (;;) {
            const uint slot = confirmed % kRing;
            const unsigned long long posted = in_loop ? lds.result[slot][lane] : kAbsent;
            const bool have = in_loop && posted != kAbsent;
            const uint n_here = (uint)posted;
            // every ballot runs in wave-uniform control flow (never inside && or ?:)
            const uint64_t missing = cu::ballot(in_loop && !have);
            const bool ready = have && (missing & group) == 0;
            if (cu::ballot(ready) == 0) break;
            const uint ri = confirmed;
            if (ready) {
                lds.result[slot][lane] = kAbsent;
                const uint escape_time = (uint)Fractal::value(n_here, max_iter);
                executed += n_here;
                escape_sum += escape_time;
                if (ri < kAdaptiveThreshold) lds.samples[ri][lane] = (float)escape_time;
            }
            const bool part = ready && vote_round(ri, sample_count, adaptive);
            float disp = 0;
            if (part) {
                const float mean = escape_sum / (ri + 1);  // integer mean (:135)
                disp = dispersion(lds.samples, lane, ri, mean);
            }
            const uint64_t P = cu::ballot(part) & group;
            const bool v1 =
                (cu::ballot(part && fabsf(lds.samples[0][lane] - lds.samples[1][lane]) <
                                        FLT_EPSILON) &
                 group) == P;
            const bool v2 = (cu::ballot(part && disp < 0.01f) & group) == P;
            const bool v3 = (cu::ballot(part && disp <= 1.f) & group) == P;
            if (part) {
                if (ri == 1 && v1) {
                    sample_count = ri + 1;
                } else if (v2) {
                    sample_count = ri + 1;
                } else if (v3 && ri >= sample_count / 2) {
                    sample_count = ri + 1;
                }
            }
            if (ready) {
                confirmed = ri + 1;
                if (!(confirmed < sample_count)) {
                    in_loop = false;
                    result = escape_sum / sample_count;  // :152
                    sample_count_f = (float)sample_count;
#pragma unroll
                    for (uint r = 0; r < kRing; ++r) lds.result[r][lane] = kAbsent;
                }
            }
            wave_lds_fence();
        }
        // ---- drop orbits whose pixel has stopped before them ----
        {
            const bool owner_in = __shfl(in_loop ? 1 : 0, (int)owner, 64) != 0;
            const uint owner_count = (uint)__shfl((int)sample_count, (int)owner, 64);
            if (mode != 0 && !(owner_in && sample < owner_count)) mode = 0;
        }
        if (cu::ballot(in_loop) == 0) break;

        // ---- hand out samples: those the next vote needs, then later ones ----
        bool more = false;
        for (int pass = 0; pass < 2; ++pass) {
            for (;;) {
                const uint64_t idle = cu::ballot(mode == 0);
                const bool can = in_loop && issued < sample_count && issued < confirmed + kRing &&
                                 (pass == 1 || issued <= next_vote_round(confirmed, sample_count,
                                                                         adaptive));
                const uint64_t avail = cu::ballot(can);
                more = avail != 0;
                if (idle == 0 || avail == 0) break;
                const uint n_take = min((uint)__popcll(idle), (uint)__popcll(avail));
                const uint orank = (uint)__popcll(avail & lanes_below);
                if (can && orank < n_take) {
                    lds.handout[orank] = lane | (issued << 8);
                    ++issued;
                }
                wave_lds_fence();
                const bool take = mode == 0 && (uint)__popcll(idle & lanes_below) < n_take;
                uint h = lane;
                if (take) h = lds.handout[__popcll(idle & lanes_below)];
                wave_lds_fence();
                const uint src = h & 63u;
                const float orow_f = __shfl(per_row_f, (int)src, 64);
                const uint orow_i = (uint)__shfl((int)per_row_i, (int)src, 64);
                if (take) {
                    owner = src;
                    sample = h >> 8;
                    const Point<uint> op = image_indexes_at(wave_base + owner);
                    Real dx, dy;
                    sample_offset<Real>(sample, orow_f, orow_i, dx, dy);
                    // c = leftTop + (1,-1) * (pixel + delta) * pixelSize  (:119-124)
                    const Real X = image.left_bottom.x + ((Real)op.x + dx) * psx;
                    const Real Y = image.right_top.y - ((Real)op.y + dy) * psy;
                    Fractal::start<Real>(X, Y, x, y, cx, cy);
                    i = 0;
                    replay = false;
                    mode = blocks ? 1 : 2;
                }
            }
        }

        // ---- unchecked blocks until kRefill lanes wait (idle ones count while work is left) ----
        Real x2 = x * x, y2 = y * y;  // squares of the current state, carried
        for (;;) {
            if (mode == 1) {
                sx = x;
                sy = y;
#pragma unroll
                for (int k = 0; k < K / 2; ++k)
                    cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
                mx = x;
                my = y;
#pragma unroll
                for (int k = K / 2; k < K; ++k)
                    cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
                if (x2 + y2 < bsq) {
                    i += K;
                    if (max_iter - i < (uint)K) mode = 2;
                } else {
                    mode = 2;
                    replay = true;
                }
            }
            const uint64_t run = cu::ballot(mode == 1);
            const uint64_t pending = cu::ballot(mode == 2);
            const uint64_t idle = cu::ballot(mode == 0);
            const uint waiting = (uint)__popcll(pending) + (more ? (uint)__popcll(idle) : 0u);
            if (run == 0 || waiting >= kRefill) break;
        }

        // ---- finish the pending orbits and post their iteration counts ----
        if (mode == 2) {
            Real mag2 = 0;
            if (replay) {
                // a failing test at max_iter itself is never reached by the reference loop:
                // n == max_iter then reads as an orbit that did not escape
                i = cu::locate_escape<Real, K, kFused>(x, y, sx, sy, mx, my, cx, cy, i, bsq, mag2);
                replay = false;
            } else {
                bool escaped = false;
                cu::checked_steps<Real, kFused>(x, y, cx, cy, i, max_iter, bsq, escaped, mag2);
            }
            lds.result[sample % kRing][owner] = (unsigned long long)i;
            mode = 0;
        }
        wave_lds_fence();
    }
    return result;
}

// Whether the fused step (fma(x*y, 2, cy)) may replace ((2x)*y)+cy for this
// launch: the rule the planar tiles use (the origin, pixel sizes and a Julia c
// each zero or far from the subnormal range; in FP32 an origin of at least
// 2^-11), so that x*y is never a nonzero subnormal and both forms round alike.
template <class Real>
__device__ __forceinline__ bool fused_valid(Rectangle<Real> image, Point<uint> grid) {
    const Real left = image.left_bottom.x, top = image.right_top.y;
    const Real psx = (image.right_top.x - image.left_bottom.x) / (Real)grid.x;
    const Real psy = (image.right_top.y - image.left_bottom.y) / (Real)grid.y;
    if constexpr (sizeof(Real) == 8) {
        const double T = 0x1.0p-300;
        return (left == 0 || fabs(left) >= T) && (top == 0 || fabs(top) >= T) && psx >= T &&
               psy >= T && Fractal::template c_fused_valid<Real>(T);
    } else {
        const double T = 0x1.0p-11;
        return fabs((double)left) >= T && fabs((double)top) >= T &&
               Fractal::template c_fused_valid<Real>(T);
    }
}

template <class Real, uint kWaves>
__device__ void render_main(pixel_info_t *out, long out_pitch, Point<uint> size,
                            Rectangle<Real> image, uint max_iter, float max_ss, uint flags) {
    const Point<uint> idx = image_indexes();
    const bool inside = idx.x < size.x && idx.y < size.y;
    float sample_count = max_ss;
    uint64_t executed = 0;
    __shared__ WaveLds lds_waves[kWaves];
    WaveLds &lds = lds_waves[(threadIdx.x + threadIdx.y * blockDim.x) >> 6];
    const bool adaptive = (flags & kAdaptiveFlag) != 0;
    const uint r = fused_valid<Real>(image, size)
                       ? sample_the_fractal<Real, true>(lds, inside ? 1 : 0, size, image, max_iter,
                                                        sample_count, adaptive, executed)
                       : sample_the_fractal<Real, false>(lds, inside ? 1 : 0, size, image, max_iter,
                                                         sample_count, adaptive, executed);
    count_iterations(executed);
    if (inside) {
        pixel_info_t p;
        p.value = (float)r;
        p.weight = sample_count;
        p.isReused = false;
        p.weightOfNewSamples = 0;
        store_pixel(pixel_at(out, out_pitch, idx.x, idx.y), p);
    }
}

// getWarpingOriginOfSampleReuse<Real> (fractalRendererGeneric.cu:194-203)
template <class Real>
__device__ __forceinline__ Point<float> warping_origin(Point<uint> p, Point<uint> grid,
                                                       Rectangle<Real> cur, Rectangle<Real> old) {
    p.y = grid.y - p.y;
    const Real plane_x =
        (Real)p.x / (Real)grid.x * (cur.right_top.x - cur.left_bottom.x) + cur.left_bottom.x;
    const Real plane_y =
        (Real)p.y / (Real)grid.y * (cur.right_top.y - cur.left_bottom.y) + cur.left_bottom.y;
    const Real rel_x = (plane_x - old.left_bottom.x) / (old.right_top.x - old.left_bottom.x);
    const Real rel_y = (plane_y - old.left_bottom.y) / (old.right_top.y - old.left_bottom.y);
    Point<float> o;
    o.x = (float)rel_x * (float)grid.x;
    o.y = (float)rel_y * (float)grid.y;
    o.y = grid.y - o.y;
    return o;
}

// readFromArrayUsingFiltering<float> (fractalRendererGeneric.cu:259-302)
__device__ __forceinline__ pixel_info_t read_filtered(const pixel_info_t *arr, long pitch,
                                                      Point<float> c) {
    const uint i = (uint)floorf(c.x);
    const uint j = (uint)floorf(c.y);
    const float alpha = c.x - i;
    const float beta = c.y - j;
    const pixel_info_t t00 = load_pixel(pixel_at(arr, pitch, i, j));
    const pixel_info_t t10 = load_pixel(pixel_at(arr, pitch, i + 1, j));
    const pixel_info_t t01 = load_pixel(pixel_at(arr, pitch, i, j + 1));
    const pixel_info_t t11 = load_pixel(pixel_at(arr, pitch, i + 1, j + 1));
    pixel_info_t r;
    r.value = (1 - alpha) * (1 - beta) * t00.value + alpha * (1 - beta) * t10.value +
              (1 - alpha) * beta * t01.value + alpha * beta * t11.value;
    r.weight = (1 - alpha) * (1 - beta) * t00.weight + alpha * (1 - beta) * t10.weight +
               (1 - alpha) * beta * t01.weight + alpha * beta * t11.weight;
    r.isReused = false;
    r.weightOfNewSamples = 0;
    return r;
}

// getFoveationAdvisedSampleCount + linearMapping (fractalRendererGeneric.cu:209-257).
// The arctangent is taken in double and rounded once to float.
__device__ __forceinline__ void foveation(Point<uint> pixel, Point<uint> focus, float max_ss,
                                          float &advised, bool &inside) {
    const uint nx = pixel.x - pixel.x % kVoteGroupW, ny = pixel.y - pixel.y % kVoteGroupH;
    const float foveal = max_ss >= 1 ? 5.5 : 5.5 * (double)max_ss;
    const float peripheral = 60;
    const float screen_distance = 60;
    const float pixel_width_cm = 0.02652;
    const float ddx = (float)focus.x - (float)nx, ddy = (float)focus.y - (float)ny;
    const float focus_distance = sqrtf(ddx * ddx + ddy * ddy) * pixel_width_cm;
    const float angle = (float)atan((double)(focus_distance / screen_distance)) * 180 / kPiF;
    const float a = 1, b = 0, r = foveal, s = peripheral;
    const float k = (b - a) / (s - r);
    const float q = (s * a - b * r) / (s - r);
    const float lm = k * angle + q;
    const float rel = (float)fmin(1.0, (double)lm);
    advised = max_ss * rel;
    inside = false;
    if (angle <= foveal) {
        advised = (float)fmax(1.0, (double)advised);
        inside = true;
    }
}

template <class Real, uint kWaves>
__device__ void render_advanced(pixel_info_t *out, long out_pitch, Point<uint> size,
                                Rectangle<Real> image, uint max_iter, float max_ss, uint flags,
                                Rectangle<Real> image_reused, const pixel_info_t *in,
                                long in_pitch, Point<uint> focus) {
    const Point<uint> idx = image_indexes();
    const bool in_image = idx.x < size.x && idx.y < size.y;

    float advised = max_ss;
    bool in_focus = false;
    if (in_image && (flags & kFoveationFlag) && (flags & kZoomingFlag) && (flags & kZoomingInFlag))
        foveation(idx, focus, max_ss, advised, in_focus);

    bool reusing = false;
    pixel_info_t reused;
    reused.value = 0;
    reused.weight = 0;
    reused.isReused = false;
    reused.weightOfNewSamples = 0;
    // (a 1-pixel-wide or -high image would wrap size-2 below and read outside the
    // previous frame; such images never reuse)
    if (in_image && (flags & kReuseFlag) && size.x >= 2 && size.y >= 2) {
        const Point<float> origin = warping_origin<Real>(idx, size, image, image_reused);
        const int ox = (int)roundf(origin.x), oy = (int)roundf(origin.y);
        if (ox < 2 || (uint)ox >= size.x - 2 || oy < 2 || (uint)oy >= size.y - 2) {
            reusing = false;
        } else {
            reused = read_filtered(in, in_pitch, origin);
            reusing = !((double)reused.weight < 0.1);
        }
    }

    float sample_count = advised;
    pixel_info_t result;
    result.value = 0;
    result.weight = 0;
    result.isReused = false;
    result.weightOfNewSamples = 0;
    int site = 0;
    if (in_image) {
        if (reusing) {
            result = reused;
            result.isReused = true;
            result.weightOfNewSamples = 0;
            if ((flags & kZoomingInFlag) && in_focus) site = 1;
        } else {
            if (sample_count < 1) sample_count = 1;
            site = 2;
        }
    }
    uint64_t executed = 0;
    __shared__ WaveLds lds_waves[kWaves];
    WaveLds &lds = lds_waves[(threadIdx.x + threadIdx.y * blockDim.x) >> 6];
    const bool adaptive = (flags & kAdaptiveFlag) != 0;
    const uint r = fused_valid<Real>(image, size)
                       ? sample_the_fractal<Real, true>(lds, site, size, image, max_iter,
                                                        sample_count, adaptive, executed)
                       : sample_the_fractal<Real, false>(lds, site, size, image, max_iter,
                                                         sample_count, adaptive, executed);
    count_iterations(executed);
    if (site == 1) {  // :349-356
        const float samples = (float)r;
        reused.weight = (float)((double)reused.weight * 0.75);
        result.weightOfNewSamples = sample_count;
        result.weight = reused.weight + sample_count;
        result.value = (reused.value * reused.weight + samples * sample_count) / result.weight;
    } else if (site == 2) {  // :358-362
        result.value = (float)r;
        result.weight = sample_count;
    }
    if (in_image) store_pixel(pixel_at(out, out_pitch, idx.x, idx.y), result);
}

// colorizeSampleCount (fractalRendererGeneric.cu:67-77); the float -> char
// conversion of 255*rel is truncation followed by the low byte.
__device__ __forceinline__ uint colorize_sample_count(uint sample_count, uint full) {
    sample_count = min(sample_count, full);
    const float rel = sample_count / (float)full;
    const uint v = (uint)(int)(255 * rel) & 0xffu;
    return v | (v << 8) | (v << 16) | (0xffu << 24);
}

}  // namespace

extern "C" __global__ void init() {}

#define CU_RENDER_KERNELS(SUFFIX, BOUNDS, BOUNDS_ADV_DOUBLE, WAVES)                               \
    extern "C" __global__ BOUNDS void fractalRenderMainFloat##SUFFIX(                               \
        pixel_info_t *output, long outputPitch, Point<uint> outputSize, Rectangle<float> image,     \
        uint maxIterations, float maxSuperSampling, uint flags) {                                   \
        render_main<float, WAVES>(output, outputPitch, outputSize, image, maxIterations,            \
                                  maxSuperSampling, flags);                                         \
    }                                                                                               \
    extern "C" __global__ BOUNDS void fractalRenderMainDouble##SUFFIX(                              \
        pixel_info_t *output, long outputPitch, Point<uint> outputSize, Rectangle<double> image,    \
        uint maxIterations, float maxSuperSampling, uint flags) {                                   \
        render_main<double, WAVES>(output, outputPitch, outputSize, image, maxIterations,           \
                                   maxSuperSampling, flags);                                        \
    }                                                                                               \
    extern "C" __global__ BOUNDS void fractalRenderAdvancedFloat##SUFFIX(                           \
        pixel_info_t *output, long outputPitch, Point<uint> outputSize, Rectangle<float> image,     \
        uint maxIterations, float maxSuperSampling, uint flags, Rectangle<float> imageReused,       \
        const pixel_info_t *input, long inputPitch, Point<uint> focus) {                            \
        render_advanced<float, WAVES>(output, outputPitch, outputSize, image, maxIterations,        \
                                      maxSuperSampling, flags, imageReused, input, inputPitch,      \
                                      focus);                                                       \
    }                                                                                               \
    extern "C" __global__ BOUNDS_ADV_DOUBLE void fractalRenderAdvancedDouble##SUFFIX(               \
        pixel_info_t *output, long outputPitch, Point<uint> outputSize, Rectangle<double> image,    \
        uint maxIterations, float maxSuperSampling, uint flags, Rectangle<double> imageReused,      \
        const pixel_info_t *input, long inputPitch, Point<uint> focus) {                            \
        render_advanced<double, WAVES>(output, outputPitch, outputSize, image, maxIterations,       \
                                       maxSuperSampling, flags, imageReused, input, inputPitch,     \
                                       focus);                                                      \
    }

// The reference's kernels: 32 x 32-thread blocks over ceil(W/32) x ceil(H/32)
// (CudaFractalRenderer.java:242-258), 16 waves and their LDS per block.
CU_RENDER_KERNELS(, __launch_bounds__(1024), __launch_bounds__(1024), 16u)

// The same kernels for 32 x 4-thread blocks over ceil(W/32) x ceil(H/4), which
// the renderer layer and the serving loop launch: the same pixels, vote groups
// and results.  A 1024-thread block's LDS and wave slots free only when its last
// wave ends, and with 70 VGPRs a CU holds 28 waves, so the next such block waits
// for 4 of them: about 3 waves per SIMD stay resident.  Blocks of 2 waves keep
// the wave slots filled (14 blocks per CU, 136 KB of LDS).  The FP64 advanced
// kernel needs 78 VGPRs: it is held to 6 waves per SIMD rather than spill.
CU_RENDER_KERNELS(Strip, __launch_bounds__(128, 7), __launch_bounds__(128, 6), 2u)

// compose (fractalRendererGeneric.cu:455-475).  The palette is staged in LDS
// (up to kPaletteLds entries; longer palettes are read from global memory) and
// the block strides over the image, so any grid covers it; the reference's
// ceil(W/32) x ceil(H/32) grid of 32x32 blocks writes one pixel per thread.
constexpr uint kPaletteLds = 4096;

extern "C" __global__ __launch_bounds__(1024) void compose(
    const pixel_info_t *inputMain, long inputMainPitch, const pixel_info_t *inputBcg,
    long inputBcgPitch, uint *output, uint width, uint height, const uint *palette,
    uint paletteLength, float maxSuperSampling) {
    (void)inputBcg;
    (void)inputBcgPitch;
    __shared__ uint lds_palette[kPaletteLds];
    const bool staged = paletteLength <= kPaletteLds;
    const uint nthreads = blockDim.x * blockDim.y;
    const uint tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (staged)
        for (uint k = tid; k < paletteLength; k += nthreads) lds_palette[k] = palette[k];
    __syncthreads();
    const uint *pal = staged ? lds_palette : palette;
    const bool visualise = VISUALIZE_SAMPLE_COUNT;
    const uint full = (uint)fmax(1.0, (double)maxSuperSampling);
    for (uint y = blockIdx.y * blockDim.y + threadIdx.y; y < height; y += gridDim.y * blockDim.y)
        for (uint x = blockIdx.x * blockDim.x + threadIdx.x; x < width;
             x += gridDim.x * blockDim.x) {
            const pixel_info_t p = load_pixel(pixel_at(inputMain, inputMainPitch, x, y));
            uint color;
            if (visualise) {
                color = colorize_sample_count((uint)p.weight, full);
                if (p.isReused) color = colorize_sample_count((uint)p.weightOfNewSamples, full);
            } else {
                // colorize (mandelbrot.cu:27-35)
                color = pal[cu::palette_slot(p.value, paletteLength)];
            }
            output[(size_t)y * width + x] = color;
        }
}

// fractalRenderUnderSampled (fractalRendererGeneric.cu:478-503): never launched
// by the reference's host; kept with its CUDA parameter list.  Thread (X, Y)
// computes the pixel (X*level, Y*level) in FP32 and writes it once.
extern "C" __global__ void fractalRenderUnderSampled(pixel_info_t *output, long outputPitch,
                                                     uint width, uint height, float left_bottom_x,
                                                     float left_bottom_y, float right_top_x,
                                                     float right_top_y, uint maxIterations,
                                                     uint underSamplingLevel) {
    // a level of 0 or above the image size would wrap `width - level` and write
    // outside the image; nothing is written then
    if (underSamplingLevel == 0 || underSamplingLevel > width || underSamplingLevel > height) return;
    const uint idx_x = (blockDim.x * blockIdx.x + threadIdx.x) * underSamplingLevel;
    const uint idx_y = (blockDim.y * blockIdx.y + threadIdx.y) * underSamplingLevel;
    if (idx_x >= width - underSamplingLevel || idx_y >= height - underSamplingLevel) return;
    const float pw = (right_top_x - left_bottom_x) / (float)width;
    const float ph = (right_top_y - left_bottom_y) / (float)height;
    const float cx = left_bottom_x + idx_x * pw;
    const float cy = right_top_y - idx_y * ph;
    uint64_t executed = 0;
    const uint e =
        (uint)compute_fractal<float>(maxIterations, cx, cy, Fractal::blocks_valid(4.0), executed);
    pixel_info_t p;
    p.value = (float)e;
    p.weight = 1 / (float)underSamplingLevel;
    p.isReused = false;
    p.weightOfNewSamples = 0;
    store_pixel(pixel_at(output, outputPitch, idx_x, idx_y), p);
}

// debug (fractalRendererGeneric.cu:505-513): thread (0,0) calls debugFractal().
extern "C" __global__ void debug() {
    const uint x = blockDim.x * blockIdx.x + threadIdx.x;
    const uint y = blockDim.y * blockIdx.y + threadIdx.y;
    if (x == 0 && y == 0) debugFractal();
}
