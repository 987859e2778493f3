// escape_pipeline.hip — the straggler-compaction frame pipeline (part of escape_family.hsaco).
//
// A frame is a set of tiles whose spp is a power of two <= 64.  Every sample
// gets one lane; a pass-1 wave holds 64/spp whole pixels.  Escape counts of
// neighbouring samples spread widely (in the 4096^2, width-1e-4 frame the median
// orbit is ~100 steps and the 99th percentile ~500), so a wave that waited for
// its slowest lane would leave most lanes idle.  Instead, as soon as no more than
// `handoff` lanes of a wave are still iterating, those lanes push their orbit
// state (z, i, tile, sample) to a straggler queue and the wave retires; the
// dispatcher starts fresh, coherent waves in its place.  Later passes run the
// queue the same way (64 queued orbits per wave), the last one to completion.
//
// Pixel results: a pass-1 pixel none of whose samples was queued is reduced in
// registers (its spp lanes summed in sample order, the order of a sequential
// loop) and written at once — count, smooth value and palette colour.  Only a
// pixel with a queued sample goes through the per-sample planes: its finished
// samples are stored there, its `pending` flag is set, and combinePending sums
// the planes of flagged pixels once the last pass is done.
//
//   escapePass1<Real><Variant>K<K>   wave-blocks of all tiles -> pixels / stragglers
//   escapePassQ<Real><Variant>K<K>   straggler queue -> samples / stragglers
//   combinePending                   flagged pixels: per-sample planes -> pixel planes

namespace {

template <class Real>
struct Straggler {
    Real x, y;    // orbit state
    uint i;       // iterations executed so far
    uint tile;    // tile index
    uint sample;  // (py * width + px) * spp + s within the tile
    uint pad;
};
static_assert(sizeof(Straggler<double>) == 32, "FP64 straggler record is 32 bytes");
static_assert(sizeof(Straggler<float>) == 24, "FP32 straggler record is 24 bytes");

struct PipeArgs {
    const cu_tile_desc *tiles;
    const uint *block_offset;       // [n_tiles + 1] prefix of wave-blocks (pass 1)
    const uint64_t *sample_offset;  // [n_tiles] first sample of each tile in the sample planes
    const double *smooth_c;         // [n_tiles] log2(log2(bailout_sq))
    uint *sample_count;             // per-sample planes (used by pixels with queued samples)
    double *sample_smooth;
    unsigned long long *iters;      // CU_ITER_SLOTS slots
    uint *counts;                   // pixel planes (tile out_offset + row * out_pitch + column)
    double *smooth;
    uint *rgba;                     // palette colour plane, or null
    const uint *palette;
    unsigned char *pending;         // per pixel: 1 when its samples finish in later passes
    uint n_tiles;
    uint handoff;                   // hand off when running lanes <= handoff (0: run to completion)
    uint subq_capacity;             // records per sub-queue
    uint palette_length;
    uint wave_shape;                // pass-1 pixel block of a wave (wave_pixel_block)
    uint pad;
};
static_assert(sizeof(PipeArgs) == 120, "PipeArgs layout");

// The straggler queue is kSubQueues sub-queues, each a region of subq_capacity
// records with its own fill counter: the wave (or queue chunk) number picks the
// sub-queue, so appends from a million waves spread over 1024 addresses instead
// of serialising on one counter.
constexpr uint kSubQueues = 1024;
// Queue chunks a drain wave takes per atomic on the shared chunk counter.  One
// address takes every grab: at one chunk per grab, the grabs of a whole pass-1
// queue (80 K chunks for the 4096^2 4-spp frame) serialised there and doubled
// the drain's time; at 4 the last grabs still share out evenly.
constexpr uint kDrainGrab = 4;

// Result of a pending lane (its block failed, or it is within K steps of
// max_iter, or it cannot use blocks): the exact stop, the count in the
// fractal's computeFractal convention and the smooth value.  i becomes the
// iterations executed by the sample.
template <class Real, int K, bool kFused>
__device__ __forceinline__ void finish_pending(Real x, Real y, Real sx, Real sy, Real mx, Real my,
                                               Real cx, Real cy, uint &i, uint max_iter, Real bsq,
                                               double smooth_c, bool julia, bool replay,
                                               uint &count, double &smooth) {
    bool escaped = false;
    Real mag2 = 0;
    if (replay) {
        // a failing test at max_iter itself is never reached by the reference loop
        i = cu::locate_escape<Real, K, kFused>(x, y, sx, sy, mx, my, cx, cy, i, bsq, mag2);
        escaped = i < max_iter;
    } else {
        cu::checked_steps<Real, kFused>(x, y, cx, cy, i, max_iter, bsq, escaped, mag2);
    }
    count = escaped ? i : (julia ? max_iter : 0u);
    smooth = escaped ? cu::smooth_value(i, (double)mag2, smooth_c) : 0.0;
}

// Orbit of one lane with the hand-off rule; shared by all passes.  Returns true
// when the lane finished its sample (count and smooth set); a lane still
// running at hand-off is appended to out_queue instead.
template <class Real, int K, bool kFused>
__device__ __forceinline__ bool run_lanes(const PipeArgs &a, bool active, Real x, Real y, Real cx,
                                          Real cy, uint i, uint max_iter, Real bsq, double smooth_c,
                                          bool julia, bool blocks, uint tile, uint sample,
                                          Straggler<Real> *out_queue, uint *out_count, uint &count,
                                          double &smooth) {
    // mode: 0 none, 1 running blocks, 2 pending (checked finish)
    int mode = !active ? 0 : ((blocks && max_iter - i >= (uint)K) ? 1 : 2);
    const uint i0 = i;  // iterations the sample had executed before this pass
    bool replay = false;
    Real sx = x, sy = y, mx = x, my = y;
    Real x2 = x * x, y2 = y * y;          // squares of the current state, carried
    const uint last = max_iter - (uint)K;  // a block may start at i <= last
    auto block = [&]() {
        if (mode == 1) {
            sx = x;
            sy = y;
#pragma unroll
            for (int k = 0; k < K / 2; ++k) cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
            mx = x;
            my = y;
#pragma unroll
            for (int k = K / 2; k < K; ++k) cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
            if (x2 + y2 < bsq) {
                i += K;
                if (i > last) mode = 2;
            } else {
                mode = 2;
                replay = true;
            }
        }
    };
    auto keep_going = [&]() {
        const uint64_t run = cu::ballot(mode == 1);
        return run != 0 && !(a.handoff && (uint)__popcll(run) <= a.handoff);
    };
    // two blocks per trip: the state registers return to place without copies
    while (keep_going()) {
        block();
        if (!keep_going()) break;
        block();
    }
    count = 0;
    smooth = 0;
    if (mode == 2)
        finish_pending<Real, K, kFused>(x, y, sx, sy, mx, my, cx, cy, i, max_iter, bsq, smooth_c,
                                        julia, replay, count, smooth);
    // this pass's own iterations, whether the sample finished or is handed on
    const uint64_t executed = mode != 0 ? (uint64_t)(i - i0) : 0u;
    // lanes still running: append to this wave's sub-queue (one atomic per wave)
    const bool push = mode == 1;
    const uint64_t pushers = cu::ballot(push);
    if (pushers) {
        uint base = 0;
        const uint first = (uint)__ffsll((unsigned long long)pushers) - 1u;
        if (cu::lane_id() == first) base = atomicAdd(out_count, (uint)__popcll(pushers));
        base = __shfl(base, (int)first, 64);
        if (push) {
            const uint rank = (uint)__popcll(pushers & ((1ull << cu::lane_id()) - 1ull));
            Straggler<Real> s;
            s.x = x;
            s.y = y;
            s.i = i;
            s.tile = tile;
            s.sample = sample;
            s.pad = 0;
            out_queue[base + rank] = s;
        }
    }
    cu::add_iterations(a.iters, executed);
    return mode == 2;
}

// Pixel outputs from the in-order sums of its 2^lspp samples: integer mean
// count, mean smooth value and the palette colour L - (round(smooth) mod L) - 1.
// Scaling by 2^-lspp is exact, so it equals the division by spp.
__device__ __forceinline__ void write_pixel(const PipeArgs &a, uint lspp, uint64_t o, uint csum,
                                            double ssum) {
    const double sm = __builtin_amdgcn_ldexp(ssum, -(int)lspp);
    a.counts[o] = csum >> lspp;
    if (a.smooth) a.smooth[o] = sm;
    if (a.rgba) {
        a.rgba[o] = a.palette[cu::palette_slot(sm, a.palette_length)];
    }
}

__device__ __forceinline__ void store_sample(const PipeArgs &a, uint tile, uint sample, uint count,
                                             double smooth) {
    const uint64_t so = a.sample_offset[tile] + sample;
    a.sample_count[so] = count;
    a.sample_smooth[so] = smooth;
}

// Pass 1: wave w of the launch takes one wave-block of the frame.  The wave
// number, tile and block geometry are wave-uniform (scalar registers); spp and
// the block width are powers of two, so lane -> (pixel, sample) is shifts.
template <class Real, int K, bool kFused>
__device__ void pass1(const PipeArgs a, uint total_blocks, Straggler<Real> *out_queue,
                      uint *out_count) {
    const uint wave_global = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                            (threadIdx.x >> 6));
    if (wave_global >= total_blocks) return;  // whole waves only: no barrier below
    uint lo = 0, hi = a.n_tiles - 1;          // wave-uniform search of the tile
    while (lo < hi) {
        const uint mid = (lo + hi + 1) >> 1;
        if (a.block_offset[mid] <= wave_global) lo = mid; else hi = mid - 1;
    }
    const uint tile = lo;
    const cu_tile_desc &t = a.tiles[tile];
    const uint spp = t.spp, lspp = (uint)__builtin_ctz(spp), lane = cu::lane_id();
    uint bw, bh;
    wave_pixel_block(spp, bw, bh, a.wave_shape);
    const uint lbw = (uint)__builtin_ctz(bw);
    const uint b = wave_global - a.block_offset[tile];
    const uint nbx = (t.width + bw - 1) >> lbw;
    // Row `by` of wave blocks is rotated by 0.618 of its width: workgroups go
    // round-robin to the XCDs and their CUs, so without it a CU's workgroups
    // would share a column of the frame (256 workgroups per row at 4096 px, 4 spp)
    // and the CUs over the set's interior would run its long orbits.
    const uint by = b / nbx;
    const uint rot = (uint)(((uint64_t)nbx * 40503u) >> 16);
    const uint bx = (uint)(((uint64_t)(b % nbx) + (uint64_t)by * rot) % nbx);
    const uint pix = lane >> lspp, s = lane & (spp - 1u);
    const uint px = bx * bw + (pix & (bw - 1u));
    const uint py = by * bh + (pix >> lbw);
    const bool active = px < t.width && py < t.height;
    const uint sample = ((py * t.width + px) << lspp) | s;
    Real x = 0, y = 0, cx = 0, cy = 0;
    if (active) begin_sample_of<Real>(t, px, py, sample, x, y, cx, cy);
    const uint q = wave_global % kSubQueues;
    uint count;
    double sm;
    const bool done = run_lanes<Real, K, kFused>(
        a, active, x, y, cx, cy, 0u, t.max_iter, (Real)t.bailout_sq, a.smooth_c[tile],
        t.kind == CU_FRACTAL_JULIA, tile_blocks_valid(t), tile, sample,
        out_queue + (uint64_t)q * a.subq_capacity, out_count + q, count, sm);

    // the spp lanes of a pixel are lanes first .. first + spp - 1
    const uint64_t queued = cu::ballot(active && !done);
    const uint first = lane - s;
    const uint64_t group = spp == 64 ? ~0ull : (((1ull << spp) - 1ull) << first);
    const bool pending = (queued & group) != 0;
    uint csum = count;
    double ssum = sm;
    for (uint j = 1; j < spp; ++j) {  // wave-uniform: every lane takes part in the shuffles
        csum += __shfl(count, (int)(lane + j), 64);
        ssum += __shfl(sm, (int)(lane + j), 64);
    }
    if (active) {
        const uint64_t o = t.out_offset + (uint64_t)py * t.out_pitch + px;
        if (pending) {
            if (done) store_sample(a, tile, sample, count, sm);
        } else if (s == 0) {
            write_pixel(a, lspp, o, csum, ssum);
        }
        if (s == 0) a.pending[o] = pending ? 1 : 0;
    }
}

// Queue chunks of 64 records numbered densely over the sub-queues: every block
// copies the fill counts to LDS and scans ceil(fill / 64), so chunk c lies in the
// last sub-queue q with prefix[q] <= c, as its chunk c - prefix[q].  Needs
// blockDim.x == 256 (thread t owns sub-queues 4t .. 4t+3); returns the total.
struct QueueIndex {
    uint fill[kSubQueues];
    uint prefix[kSubQueues + 1];
    uint wave_total[4];
};

__device__ __forceinline__ uint build_queue_index(QueueIndex &ix, const uint *fill) {
    const uint t = threadIdx.x, q0 = 4u * t;
    uint c[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint f = fill[q0 + j];
        ix.fill[q0 + j] = f;
        c[j] = (f + 63u) >> 6;
        sum += c[j];
    }
    uint incl = sum;  // inclusive scan within the wave
    const uint ln = cu::lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint v = __shfl_up(incl, off, 64);
        if (ln >= (uint)off) incl += v;
    }
    if (ln == 63) ix.wave_total[t >> 6] = incl;
    __syncthreads();
    uint run = incl - sum;
    for (uint w = 0; w < (t >> 6); ++w) run += ix.wave_total[w];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ix.prefix[q0 + j] = run;
        run += c[j];
    }
    if (t == blockDim.x - 1) ix.prefix[kSubQueues] = run;
    __syncthreads();
    return ix.prefix[kSubQueues];
}

// Sub-queue of dense chunk c (wave-uniform search in LDS).
__device__ __forceinline__ uint chunk_queue(const QueueIndex &ix, uint c) {
    uint lo = 0, hi = kSubQueues - 1;
    while (lo < hi) {
        const uint mid = (lo + hi + 1) >> 1;
        if (ix.prefix[mid] <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Later passes: a fixed grid strides over the chunks of 64 queued orbits.  Chunk
// c is chunk c / kSubQueues of sub-queue c % kSubQueues, so the low chunks of
// every sub-queue come first; a chunk past its sub-queue's fill is skipped.
// Stragglers of sub-queue q go back to sub-queue q of the output queue, which
// therefore never holds more than the input sub-queue did.
// Lane state of a queued orbit: its z, i, tile and sample from the record, c
// recomputed from the tile and the sample number.
template <class Real>
__device__ __forceinline__ void load_straggler(const PipeArgs &a, const Straggler<Real> &e, Real &x,
                                               Real &y, Real &cx, Real &cy, uint &i, uint &tile,
                                               uint &sample, uint &max_iter, Real &bsq,
                                               double &smooth_c, bool &julia, bool &blocks) {
    tile = e.tile;
    sample = e.sample;
    i = e.i;
    const cu_tile_desc &t = a.tiles[tile];
    const uint pixel = sample >> __builtin_ctz(t.spp);
    Real zx, zy;
    begin_sample_of<Real>(t, pixel % t.width, pixel / t.width, sample, zx, zy, cx, cy);
    x = e.x;
    y = e.y;
    max_iter = t.max_iter;
    bsq = (Real)t.bailout_sq;
    smooth_c = a.smooth_c[tile];
    julia = t.kind == CU_FRACTAL_JULIA;
    blocks = tile_blocks_valid(t);
}

template <class Real, int K, bool kFused>
__device__ void passq_chunk(const PipeArgs &a, const Straggler<Real> *in_queue, uint n, uint chunk,
                            Straggler<Real> *out_queue, uint *out_count) {
    const uint idx = chunk * 64u + cu::lane_id();
    const bool active = idx < n;
    Real x = 0, y = 0, cx = 0, cy = 0, bsq = 4;
    uint i = 0, tile = 0, sample = 0, max_iter = 0;
    double smooth_c = 1;
    bool julia = false, blocks = false;
    if (active)
        load_straggler<Real>(a, in_queue[idx], x, y, cx, cy, i, tile, sample, max_iter, bsq,
                             smooth_c, julia, blocks);
    uint count;
    double sm;
    if (run_lanes<Real, K, kFused>(a, active, x, y, cx, cy, i, max_iter, bsq, smooth_c, julia,
                                   blocks, tile, sample, out_queue, out_count, count, sm))
        store_sample(a, tile, sample, count, sm);
}

template <class Real, int K, bool kFused>
__device__ void passq(const PipeArgs a, const Straggler<Real> *in_queue, const uint *in_count,
                      Straggler<Real> *out_queue, uint *out_count) {
    __shared__ QueueIndex ix;
    const uint chunks = build_queue_index(ix, in_count);
    const uint wave_global = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                            (threadIdx.x >> 6));
    const uint total_waves = gridDim.x * (blockDim.x >> 6);
    for (uint c = wave_global; c < chunks; c += total_waves) {
        const uint q = chunk_queue(ix, c), k = c - ix.prefix[q];
        const uint64_t region = (uint64_t)q * a.subq_capacity;
        passq_chunk<Real, K, kFused>(a, in_queue + region, ix.fill[q], k, out_queue + region,
                                     out_count + q);
    }
}

// Drain: one persistent launch runs every orbit left in the pass-1 queue to its
// end, refilling idle lanes from the queue (the lane-refill scheme of
// escape_tiles_persistent applied to queued orbits), so the longest orbits
// start at once instead of after several hand-offs.  A wave takes kDrainGrab
// dense queue chunks (build_queue_index) per atomic on `grab` and hands their
// records to idle lanes.  Once `refill` lanes wait, the
// pending lanes finish (exact stop, smooth value, per-sample store) and idle
// lanes take new records.  Every wave leaves when the chunks are exhausted and
// its lanes are idle; there is no barrier after the prefix.
template <class Real, int K, bool kFused>
__device__ void drain(const PipeArgs a, const Straggler<Real> *queue, const uint *fill,
                      uint *grab, uint refill) {
    __shared__ QueueIndex ix;
    const uint total_chunks = build_queue_index(ix, fill);
    const uint lane = cu::lane_id();
    const uint64_t lanes_below = (1ull << lane) - 1ull;

    int mode = 0;  // 0 idle, 1 running unchecked blocks, 2 pending
    bool replay = false;
    Real x = 0, y = 0, cx = 0, cy = 0, sx = 0, sy = 0, mx = 0, my = 0, bsq = 4;
    Real x2 = 0, y2 = 0;  // squares of the current state, carried from step to step
    uint i = 0, i0 = 0, tile = 0, sample = 0, max_iter = 0;
    double smooth_c = 1;
    bool julia = false, blocks = false;
    uint64_t executed = 0;

    uint cq = 0, cur = 0, end = 0;  // wave-uniform record cursor in sub-queue cq
    // Wave-uniform: chunks [next, last) taken by this wave.  A wave's first
    // kDrainGrab chunks follow from its number, so the launch does not start
    // with every wave's atomic on the one counter; later grabs come after all
    // of the first ones.
    const uint wave_global = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                            (threadIdx.x >> 6));
    const uint first_chunks = gridDim.x * (blockDim.x >> 6) * kDrainGrab;
    uint next = wave_global * kDrainGrab;
    uint last = min(next + kDrainGrab, total_chunks);
    bool exhausted = false;

    for (;;) {
        // ---- refill idle lanes ----
        uint64_t idle = cu::ballot(mode == 0);
        while (idle != 0 && !exhausted) {
            if (cur >= end) {
                if (next >= last) {
                    uint c0 = 0;
                    if (lane == 0 && first_chunks < total_chunks) c0 = atomicAdd(grab, kDrainGrab);
                    next = first_chunks + __builtin_amdgcn_readfirstlane(__shfl(c0, 0, 64));
                    last = min(next + kDrainGrab, total_chunks);
                }
                if (next >= total_chunks) {
                    exhausted = true;
                    break;
                }
                const uint c = next++;
                cq = chunk_queue(ix, c);
                cur = (c - ix.prefix[cq]) * 64u;
                end = min(cur + 64u, ix.fill[cq]);
            }
            const uint n = (uint)__popcll(idle);
            const uint take = min(n, end - cur);
            const uint rank = (uint)__popcll(idle & lanes_below);
            if (mode == 0 && rank < take) {
                const Straggler<Real> e = queue[(uint64_t)cq * a.subq_capacity + cur + rank];
                load_straggler<Real>(a, e, x, y, cx, cy, i, tile, sample, max_iter, bsq, smooth_c,
                                     julia, blocks);
                i0 = i;
                replay = false;
                mode = (blocks && max_iter - i >= (uint)K) ? 1 : 2;
                x2 = x * x;
                y2 = y * y;
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
                for (int k = 0; k < K / 2; ++k) cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
                mx = x;
                my = y;
#pragma unroll
                for (int k = K / 2; k < K; ++k) cu::quad_step_sq<Real, kFused>(x, y, x2, y2, cx, cy);
                if (x2 + y2 < bsq) {
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

        // ---- pending lanes finish their sample ----
        if (mode == 2) {
            uint count;
            double sm;
            finish_pending<Real, K, kFused>(x, y, sx, sy, mx, my, cx, cy, i, max_iter, bsq,
                                            smooth_c, julia, replay, count, sm);
            executed += i - i0;  // the iterations run here; earlier passes counted theirs
            store_sample(a, tile, sample, count, sm);
            mode = 0;
        }
    }
    cu::add_iterations(a.iters, executed);
}

}  // namespace

#define CU_PIPE_KERNELS(NAME, REAL, FUSED, STEPS)                                                 \
    extern "C" __global__ __launch_bounds__(256) void escapePass1##NAME##K##STEPS(                \
        PipeArgs a, uint total_blocks, Straggler<REAL> *out_queue, uint *out_count) {             \
        pass1<REAL, STEPS, FUSED>(a, total_blocks, out_queue, out_count);                         \
    }                                                                                             \
    extern "C" __global__ __launch_bounds__(256) void escapePassQ##NAME##K##STEPS(                \
        PipeArgs a, const Straggler<REAL> *in_queue, const uint *in_count,                        \
        Straggler<REAL> *out_queue, uint *out_count) {                                            \
        passq<REAL, STEPS, FUSED>(a, in_queue, in_count, out_queue, out_count);                   \
    }                                                                                             \
    extern "C" __global__ __launch_bounds__(256) void escapeDrain##NAME##K##STEPS(                \
        PipeArgs a, const Straggler<REAL> *queue, const uint *fill, uint *grab, uint refill) {    \
        drain<REAL, STEPS, FUSED>(a, queue, fill, grab, refill);                                  \
    }

CU_PIPE_KERNELS(DoubleExact, double, false, 8)
CU_PIPE_KERNELS(DoubleExact, double, false, 16)
CU_PIPE_KERNELS(DoubleFused, double, true, 8)
CU_PIPE_KERNELS(DoubleFused, double, true, 16)
CU_PIPE_KERNELS(FloatExact, float, false, 8)
CU_PIPE_KERNELS(FloatExact, float, false, 16)
CU_PIPE_KERNELS(FloatFused, float, true, 8)
CU_PIPE_KERNELS(FloatFused, float, true, 16)

namespace {

// combinePending for a frame of one tile with 4 samples per pixel (the planar
// frame): a thread takes 4 consecutive pixels, reads their 4 pending flags as
// one word and, for the pending ones, issues every per-sample load (one 16-byte
// load of counts, two of smooth values per pixel) before using any, so the
// kernel waits for one round of loads per 4 pixels instead of one per pixel.
// Flag words are 4-byte aligned (the plane starts on an allocation boundary),
// and so are the 16-byte sample groups (sample_offset[0] = 0, 4 per pixel).
__device__ void combine_one_tile_spp4(const PipeArgs &a, uint64_t total_pixels) {
    const cu_tile_desc &t = a.tiles[0];
    const uint64_t soff = a.sample_offset[0];
    const uint *flags4 = reinterpret_cast<const uint *>(a.pending);
    const uint4 *counts4 = reinterpret_cast<const uint4 *>(a.sample_count + soff);
    const double2 *smooth2 = reinterpret_cast<const double2 *>(a.sample_smooth + soff);
    const uint64_t groups = total_pixels / 4;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint f = flags4[g];
        if (f == 0) continue;
        uint4 c[4];
        double2 s0[4], s1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((f >> (8 * j)) & 0xffu) {
                const uint64_t p = g * 4 + j;
                c[j] = counts4[p];
                s0[j] = smooth2[2 * p];
                s1[j] = smooth2[2 * p + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if ((f >> (8 * j)) & 0xffu) {
                const uint64_t local = g * 4 + j;
                const uint csum = c[j].x + c[j].y + c[j].z + c[j].w;
                const double ssum = ((s0[j].x + s0[j].y) + s1[j].x) + s1[j].y;
                uint64_t o = t.out_offset + local;
                if (t.out_pitch != t.width)
                    o = t.out_offset + (uint64_t)(local / t.width) * t.out_pitch + local % t.width;
                write_pixel(a, 2u, o, csum, ssum);
            }
        }
    }
}

}  // namespace

// Pixels whose samples finished in several passes: sum the per-sample planes in
// sample order and write the pixel outputs.  Thread per frame pixel; pixel p lies
// in the last tile whose work_offset (pixel prefix) is <= p.
extern "C" __global__ __launch_bounds__(256) void combinePending(PipeArgs a,
                                                                 uint64_t total_pixels) {
    if (a.n_tiles == 1 && a.tiles[0].spp == 4 && total_pixels % 4 == 0) {
        combine_one_tile_spp4(a, total_pixels);  // wave-uniform: every thread of the launch
        return;
    }
    __shared__ uint64_t lds_offset[kMaxLdsTiles];
    for (uint k = threadIdx.x; k < a.n_tiles; k += blockDim.x)
        lds_offset[k] = a.tiles[k].work_offset;
    __syncthreads();
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total_pixels;
         p += (uint64_t)gridDim.x * blockDim.x) {
        if (!a.pending[p]) continue;
        uint lo = 0, hi = a.n_tiles - 1;
        while (lo < hi) {
            const uint mid = (lo + hi + 1) >> 1;
            if (lds_offset[mid] <= p) lo = mid; else hi = mid - 1;
        }
        const cu_tile_desc &t = a.tiles[lo];
        const uint local = (uint)(p - lds_offset[lo]);
        const uint64_t so = a.sample_offset[lo] + (uint64_t)local * t.spp;
        uint csum = 0;
        double ssum = 0;
        for (uint s = 0; s < t.spp; ++s) {
            csum += a.sample_count[so + s];
            ssum += a.sample_smooth[so + s];
        }
        uint64_t o = t.out_offset + local;  // rows packed: pitch == width
        if (t.out_pitch != t.width)
            o = t.out_offset + (uint64_t)(local / t.width) * t.out_pitch + local % t.width;
        write_pixel(a, (uint)__builtin_ctz(t.spp), o, csum, ssum);
    }
}
