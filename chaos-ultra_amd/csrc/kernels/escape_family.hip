// escape_family.hip — planar escape-time kernels of the quadratic family for gfx950.
//
// These serve the throughput forms that have no counterpart in the reference:
// tiles of pixel-centre or seeded-jitter samples with any bailout radius and a
// smooth value, a persistent launch over many tiles of uneven depth, and a
// palette compose of the smooth plane.  Mandelbrot and Julia tiles mix freely
// (per-tile `kind`); the orbit rounding is the reference's (escape_core.hpp).
//
// Kernels (all extern "C", bound by name from escape_family.hsaco):
//   escapeTile{Double,Float}{Exact,Fused}      thread per pixel, one tile
//   escapeTilesPersistent{Double,Float}{Exact,Fused}
//                                              lane-refill work queue, many tiles
//   composeSmooth                              smooth plane -> RGBA8, LDS palette
//   fillU32                                    device fill (graph-capture safe)
#include "cu_types.hpp"
#include "escape_core.hpp"

#pragma clang fp contract(off)

namespace {

constexpr uint kMaxLdsTiles = 1024;
constexpr uint kPaletteLds = 4096;

// Start sample `sample_id` = (py * width + px) * spp + s of tile t (pixel (px, py)):
// z0 and c for the tile's kind.
template <class Real>
__device__ __forceinline__ void begin_sample_of(const cu_tile_desc &t, uint px, uint py,
                                                uint64_t sample_id, Real &x, Real &y, Real &cx,
                                                Real &cy) {
    double dx, dy;
    cu::sample_offset_of(t, sample_id, dx, dy);
    Real X, Y;
    cu::tile_point<Real>(t, px, py, dx, dy, X, Y);
    if (t.kind == CU_FRACTAL_JULIA) {
        x = X;
        y = Y;
        cx = (Real)t.julia_cx;
        cy = (Real)t.julia_cy;
    } else {
        x = 0;
        y = 0;
        cx = X;
        cy = Y;
    }
}

// Start sample s of pixel (px, py) of tile t: z0 and c for the tile's kind.
template <class Real>
__device__ __forceinline__ void begin_sample(const cu_tile_desc &t, uint px, uint py, uint s,
                                             Real &x, Real &y, Real &cx, Real &cy) {
    begin_sample_of<Real>(t, px, py, ((uint64_t)py * t.width + px) * t.spp + s, x, y, cx, cy);
}

__device__ __forceinline__ bool tile_blocks_valid(const cu_tile_desc &t) {
    return cu::blocks_valid(t.bailout_sq, t.kind == CU_FRACTAL_JULIA, t.julia_cx, t.julia_cy);
}

// ---------------------------------------------------------------------------
// thread per pixel over one tile (any spp; samples in sequence)
// ---------------------------------------------------------------------------
template <class Real, int K, bool kFused>
__device__ void escape_tile(const cu_tile_desc &t, double smooth_c, uint *counts, double *smooth,
                            unsigned long long *iters) {
    const uint px = cu::skewed_block_x() * blockDim.x + threadIdx.x;
    const uint py = blockIdx.y * blockDim.y + threadIdx.y;
    uint64_t executed = 0;
    if (px < t.width && py < t.height) {
        const bool blocks = tile_blocks_valid(t);
        const Real bsq = (Real)t.bailout_sq;
        const bool julia = t.kind == CU_FRACTAL_JULIA;
        uint count_sum = 0;
        double smooth_sum = 0;
        for (uint s = 0; s < t.spp; ++s) {
            Real x, y, cx, cy;
            begin_sample<Real>(t, px, py, s, x, y, cx, cy);
            const cu::Orbit<Real> o =
                cu::escape_orbit<Real, K, kFused>(x, y, cx, cy, t.max_iter, bsq, blocks);
            executed += o.n;
            count_sum += o.escaped ? o.n : (julia ? t.max_iter : 0u);
            smooth_sum += o.escaped ? cu::smooth_value(o.n, (double)o.mag2, smooth_c) : 0.0;
        }
        const uint64_t o = t.out_offset + (uint64_t)py * t.out_pitch + px;
        counts[o] = count_sum / t.spp;
        if (smooth) smooth[o] = smooth_sum / (double)t.spp;
    }
    cu::add_iterations(iters, executed);
}

// ---------------------------------------------------------------------------
// one sample per lane over one tile (spp a power of two <= 64)
// ---------------------------------------------------------------------------
// A wave covers 64/spp pixels as a near-square block (spp 4: 4x4 pixels x 4
// samples), so neighbouring lanes iterate neighbouring points and the wave's
// longest orbit stays close to its mean; per-sample setup and the smooth value
// run on all lanes at once.  The spp lanes of a pixel sum their counts and
// smooth values in sample order (lane s==0 adds lanes 1..spp-1 in turn), the
// same order as a sequential loop.  Block: 256 threads = 4 waves in 2x2.
// shape 0 is the near-square block; each further shape is twice as wide (capped
// at one row of 64/spp pixels).
__device__ __forceinline__ void wave_pixel_block(uint spp, uint &bw, uint &bh, uint shape = 0) {
    const uint ppw = 64u / spp;
    const uint lg = 31u - __clz(ppw);
    bw = min(ppw, 1u << (((lg + 1u) >> 1) + shape));
    bh = ppw / bw;
}

template <class Real, int K, bool kFused>
__device__ void escape_tile_samples(const cu_tile_desc &t, double smooth_c, uint *counts,
                                    double *smooth, unsigned long long *iters) {
    const uint spp = t.spp;
    const uint lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint bw, bh;
    wave_pixel_block(spp, bw, bh);
    const uint pix = lane / spp, s = lane % spp;
    const uint px = (cu::skewed_block_x() * 2u + (wave & 1u)) * bw + pix % bw;
    const uint py = (blockIdx.y * 2u + (wave >> 1)) * bh + pix / bw;
    const bool inside = px < t.width && py < t.height;
    const bool julia = t.kind == CU_FRACTAL_JULIA;
    uint count = 0;
    double sm = 0;
    uint64_t executed = 0;
    if (inside) {
        Real x, y, cx, cy;
        begin_sample<Real>(t, px, py, s, x, y, cx, cy);
        const cu::Orbit<Real> o = cu::escape_orbit<Real, K, kFused>(
            x, y, cx, cy, t.max_iter, (Real)t.bailout_sq, tile_blocks_valid(t));
        executed = o.n;
        count = o.escaped ? o.n : (julia ? t.max_iter : 0u);
        sm = o.escaped ? cu::smooth_value(o.n, (double)o.mag2, smooth_c) : 0.0;
    }
    uint count_sum = count;
    double smooth_sum = sm;
    for (uint j = 1; j < spp; ++j) {  // wave-uniform: every lane takes part in the shuffles
        const uint cj = __shfl(count, (int)(lane + j), 64);
        const double sj = __shfl(sm, (int)(lane + j), 64);
        count_sum += cj;
        smooth_sum += sj;
    }
    if (inside && s == 0) {
        const uint64_t o = t.out_offset + (uint64_t)py * t.out_pitch + px;
        counts[o] = count_sum / spp;
        if (smooth) smooth[o] = smooth_sum / (double)spp;
    }
    cu::add_iterations(iters, executed);
}

// ---------------------------------------------------------------------------
// persistent lane-refill launch over many tiles
// ---------------------------------------------------------------------------
// Every wave keeps 64 independent samples in flight.  Lanes run unchecked
// blocks of K steps in lock step; a lane whose block fails its end test (or
// whose remaining budget is below K) becomes pending.  Once `refill` lanes are
// pending or idle, the pending lanes locate the exact escape step in the half of
// their last block that holds it (cu::locate_escape), finish their sample, start the next sample of the same
// pixel or write the pixel, and idle lanes take new pixels from a wave-private
// chunk of the global pixel queue (one atomic per `chunk` pixels).  Pixels of
// one tile are numbered row-major from the tile's work_offset.  There is no
// barrier after the tile table is staged, and every wave leaves once the queue
// is exhausted and its lanes are idle.
template <class Real, int K, bool kFused>
__device__ void escape_tiles_persistent(const cu_tile_desc *__restrict__ tiles, uint n_tiles,
                                        uint64_t total, uint *counts, double *smooth,
                                        unsigned long long *iters, unsigned long long *queue,
                                        uint chunk, uint refill) {
    __shared__ uint64_t lds_offset[kMaxLdsTiles];
    __shared__ double lds_smooth_c[kMaxLdsTiles];
    for (uint k = threadIdx.x; k < n_tiles; k += blockDim.x) {
        lds_offset[k] = tiles[k].work_offset;
        lds_smooth_c[k] = cu::smooth_constant(tiles[k].bailout_sq);
    }
    __syncthreads();

    const uint lane = cu::lane_id();
    const uint64_t lanes_below = (1ull << lane) - 1ull;

    int mode = 0;          // 0 idle, 1 running unchecked blocks, 2 pending checked work
    bool replay = false;   // pending lane's last block failed: locate the escape in it
    uint tile = 0, px = 0, py = 0, s = 0, i = 0, count_sum = 0;
    double smooth_sum = 0;
    Real x = 0, y = 0, cx = 0, cy = 0, sx = 0, sy = 0, mx = 0, my = 0, bsq = 4;
    uint max_iter = 0;
    uint64_t executed = 0;

    // Wave-uniform chunk cursor.  A wave's first chunk follows from its number,
    // so the launch does not start with every wave's atomic on the one counter;
    // the counter hands out the chunks after all of the first ones.
    const uint64_t wave_global = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                                (threadIdx.x >> 6));
    const uint64_t first_items = (uint64_t)gridDim.x * (blockDim.x >> 6) * chunk;
    uint64_t cur = wave_global * chunk;
    uint64_t end = min(cur + chunk, total);
    bool exhausted = false;

    for (;;) {
        // ---- refill idle lanes ----
        uint64_t idle = cu::ballot(mode == 0);
        while (idle != 0 && !exhausted) {
            if (cur >= end) {
                unsigned long long base = 0;
                if (lane == 0 && first_items < total) base = atomicAdd(queue, (unsigned long long)chunk);
                base = first_items + __shfl(base, 0, 64);
                if (base >= total) {
                    exhausted = true;
                    break;
                }
                cur = base;
                end = min((uint64_t)base + chunk, total);
            }
            const uint n = (uint)__popcll(idle);
            const uint64_t avail = end - cur;
            const uint take = (uint)min((uint64_t)n, avail);
            const uint rank = (uint)__popcll(idle & lanes_below);
            if (mode == 0 && rank < take) {
                const uint64_t item = cur + rank;
                uint lo = 0, hi = n_tiles - 1;  // last tile with work_offset <= item
                while (lo < hi) {
                    const uint mid = (lo + hi + 1) >> 1;
                    if (lds_offset[mid] <= item) lo = mid; else hi = mid - 1;
                }
                tile = lo;
                const cu_tile_desc &t = tiles[tile];
                const uint local = (uint)(item - lds_offset[tile]);  // a tile holds < 2^32 pixels
                px = local % t.width;
                py = local / t.width;
                s = 0;
                count_sum = 0;
                smooth_sum = 0;
                max_iter = t.max_iter;
                bsq = (Real)t.bailout_sq;
                begin_sample<Real>(t, px, py, 0, x, y, cx, cy);
                i = 0;
                replay = false;
                mode = (tile_blocks_valid(t) && max_iter >= (uint)K) ? 1 : 2;
            }
            cur += take;
            idle = cu::ballot(mode == 0);
        }
        if (cu::ballot(mode != 0) == 0) break;

        // ---- unchecked blocks until enough lanes wait ----
        for (;;) {
            if (mode == 1) {
                sx = x;
                sy = y;
#pragma unroll
                for (int k = 0; k < K / 2; ++k) cu::quad_step<Real, kFused>(x, y, cx, cy);
                mx = x;
                my = y;
#pragma unroll
                for (int k = K / 2; k < K; ++k) cu::quad_step<Real, kFused>(x, y, cx, cy);
                if (x * x + y * y < bsq) {
                    i += K;
                    if (max_iter - i < (uint)K) mode = 2;
                } else {
                    mode = 2;
                    replay = true;
                }
            }
            const uint64_t run = cu::ballot(mode == 1);
            const uint waiting = (uint)__popcll(cu::ballot(mode == 2)) +
                                 (exhausted ? 0u : (uint)__popcll(cu::ballot(mode == 0)));
            if (run == 0 || waiting >= refill) break;
        }

        // ---- pending lanes: locate the escape or run checked steps, finish the sample ----
        if (mode == 2) {
            bool escaped = false;
            Real mag2 = 0;
            if (replay) {
                // a failing test at max_iter itself is never reached by the reference loop
                i = cu::locate_escape<Real, K, kFused>(x, y, sx, sy, mx, my, cx, cy, i, bsq, mag2);
                escaped = i < max_iter;
                replay = false;
            } else {
                cu::checked_steps<Real, kFused>(x, y, cx, cy, i, max_iter, bsq, escaped, mag2);
            }
            const cu_tile_desc &t = tiles[tile];
            executed += i;
            count_sum += escaped ? i : (t.kind == CU_FRACTAL_JULIA ? max_iter : 0u);
            smooth_sum += escaped ? cu::smooth_value(i, (double)mag2, lds_smooth_c[tile]) : 0.0;
            ++s;
            if (s < t.spp) {
                begin_sample<Real>(t, px, py, s, x, y, cx, cy);
                i = 0;
                mode = (tile_blocks_valid(t) && max_iter >= (uint)K) ? 1 : 2;
            } else {
                const uint64_t o = t.out_offset + (uint64_t)py * t.out_pitch + px;
                counts[o] = count_sum / t.spp;
                if (smooth) smooth[o] = smooth_sum / (double)t.spp;
                mode = 0;
            }
        }
    }
    cu::add_iterations(iters, executed);
}

}  // namespace

#define CU_TILE_KERNELS(NAME, REAL, FUSED)                                                      \
    extern "C" __global__ __launch_bounds__(256) void escapeTile##NAME(                         \
        cu_tile_desc tile, double smooth_c, uint *counts, double *smooth,                       \
        unsigned long long *iters) {                                                            \
        escape_tile<REAL, 16, FUSED>(tile, smooth_c, counts, smooth, iters);                    \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void escapeTileSamples##NAME(                  \
        cu_tile_desc tile, double smooth_c, uint *counts, double *smooth,                       \
        unsigned long long *iters) {                                                            \
        escape_tile_samples<REAL, 16, FUSED>(tile, smooth_c, counts, smooth, iters);            \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void escapeTileSamples##NAME##K8(              \
        cu_tile_desc tile, double smooth_c, uint *counts, double *smooth,                       \
        unsigned long long *iters) {                                                            \
        escape_tile_samples<REAL, 8, FUSED>(tile, smooth_c, counts, smooth, iters);             \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void escapeTilesPersistent##NAME##K8(          \
        const cu_tile_desc *tiles, uint n_tiles, uint64_t total, uint *counts, double *smooth,  \
        unsigned long long *iters, unsigned long long *queue, uint chunk, uint refill) {        \
        escape_tiles_persistent<REAL, 8, FUSED>(tiles, n_tiles, total, counts, smooth, iters,   \
                                                queue, chunk, refill);                          \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void escapeTilesPersistent##NAME##K16(         \
        const cu_tile_desc *tiles, uint n_tiles, uint64_t total, uint *counts, double *smooth,  \
        unsigned long long *iters, unsigned long long *queue, uint chunk, uint refill) {        \
        escape_tiles_persistent<REAL, 16, FUSED>(tiles, n_tiles, total, counts, smooth, iters,  \
                                                 queue, chunk, refill);                         \
    }                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void escapeTilesPersistent##NAME##K32(         \
        const cu_tile_desc *tiles, uint n_tiles, uint64_t total, uint *counts, double *smooth,  \
        unsigned long long *iters, unsigned long long *queue, uint chunk, uint refill) {        \
        escape_tiles_persistent<REAL, 32, FUSED>(tiles, n_tiles, total, counts, smooth, iters,  \
                                                 queue, chunk, refill);                         \
    }

CU_TILE_KERNELS(DoubleExact, double, false)
CU_TILE_KERNELS(DoubleFused, double, true)
CU_TILE_KERNELS(FloatExact, float, false)
CU_TILE_KERNELS(FloatFused, float, true)

// Palette compose of a smooth plane: idx = L - (round(v) mod L) - 1 (the
// colorize of mandelbrot.cu:27-35 applied to the smooth value; v <= 0 as 0).
extern "C" __global__ __launch_bounds__(256) void composeSmooth(
    const double *smooth, uint smooth_pitch, uint *rgba, uint rgba_pitch, uint width,
    uint height, const uint *palette, uint palette_length) {
    __shared__ uint lds_palette[kPaletteLds];
    const bool staged = palette_length <= kPaletteLds;
    const uint nthreads = blockDim.x * blockDim.y;
    const uint tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (staged)
        for (uint k = tid; k < palette_length; k += nthreads) lds_palette[k] = palette[k];
    __syncthreads();
    const uint *pal = staged ? lds_palette : palette;
    for (uint y = blockIdx.y * blockDim.y + threadIdx.y; y < height; y += gridDim.y * blockDim.y)
        for (uint x = blockIdx.x * blockDim.x + threadIdx.x; x < width;
             x += gridDim.x * blockDim.x) {
            rgba[(size_t)y * rgba_pitch + x] =
                pal[cu::palette_slot(smooth[(size_t)y * smooth_pitch + x], palette_length)];
        }
}

// Order-sensitive checksum of n words: sum of p[i] * (2i + 1) mod 2^64, added to
// *out (integer addition, so the result does not depend on the schedule).
extern "C" __global__ __launch_bounds__(256) void checksumU32(const uint *p, uint64_t n,
                                                              unsigned long long *out) {
    uint64_t acc = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (uint64_t)gridDim.x * blockDim.x)
        acc += (uint64_t)p[k] * (2ull * k + 1ull);
    acc = __ockl_wfred_add_u64(acc);
    if (cu::lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

extern "C" __global__ __launch_bounds__(256) void fillU32(uint *p, uint value, uint64_t n) {
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (uint64_t)gridDim.x * blockDim.x)
        p[k] = value;
}

// The straggler-compaction frame pipeline shares this module's helpers.
#include "escape_pipeline.hip"
// The deep-zoom perturbation form likewise.
#include "perturb.hip"
