// frame.cpp — the straggler-compaction frame pipeline (pass 1, tail passes,
// combine) over the escape-family module.  See escape_pipeline.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "engine_helpers.hpp"
#include "handles.hpp"

namespace {

// Device-side argument block of escapePass1 / escapePassQ (same layout as the
// kernel's PipeArgs).
struct PipeArgs {
    const cu_tile_desc *tiles;
    const uint32_t *block_offset;
    const uint64_t *sample_offset;
    const double *smooth_c;
    uint32_t *sample_count;
    double *sample_smooth;
    unsigned long long *iters;
    uint32_t *counts;
    double *smooth;
    uint32_t *rgba;
    const uint32_t *palette;
    unsigned char *pending;
    uint32_t n_tiles;
    uint32_t handoff;
    uint32_t subq_capacity;
    uint32_t palette_length;
    uint32_t wave_shape;
    uint32_t pad;
};
static_assert(sizeof(PipeArgs) == 120, "PipeArgs layout");

constexpr uint32_t kMaxHandoff = 32;   // records a wave may push
constexpr uint32_t kWaveShapes = 4;    // pixel-block shapes of a pass-1 wave
constexpr uint32_t kSubQueues = 1024;  // sub-queues per straggler queue (kernel constant)
constexpr uint32_t kMaxTailPasses = 8;
// iteration slots (2 x CU_ITER_SLOTS u64), a counter set for pass 1 and each tail
// pass, and the drain's chunk counter
constexpr uint64_t kControlWords = 4ull * CU_ITER_SLOTS + (kMaxTailPasses + 1ull) * kSubQueues + 64;

// The bw x bh pixel block of a pass-1 wave (bw * bh = 64 / spp): shape 0 is the
// near-square block, each further shape twice as wide (capped at one row).
void wave_pixel_block(uint32_t spp, uint32_t shape, uint32_t &bw, uint32_t &bh) {
    uint32_t ppw = 64 / spp, lg = 0;
    while ((1u << (lg + 1)) <= ppw) ++lg;
    bw = std::min(ppw, 1u << (((lg + 1) >> 1) + shape));
    bh = ppw / bw;
}

}  // namespace

struct cu_frame_s {
    cu_engine engine = nullptr;
    int precision = 64;
    bool fused = false;
    std::vector<cu_tile_desc> tiles;
    uint64_t total_pixels = 0, total_samples = 0;
    uint32_t total_blocks[kWaveShapes] = {0, 0, 0, 0};
    uint32_t subq_cap = 0;  // records per sub-queue
    void *d_tiles = nullptr, *d_block_offset = nullptr, *d_sample_offset = nullptr;
    void *d_smooth_c = nullptr;  // f64 per tile: log2(log2(bailout_sq))
    void *d_sample_count = nullptr, *d_sample_smooth = nullptr;
    void *d_pending = nullptr;   // u8 per pixel
    void *d_queue[2] = {nullptr, nullptr};
    void *d_counts = nullptr, *d_smooth = nullptr;
    // Control block, cleared by one fill per render: 2 x CU_ITER_SLOTS u64
    // iteration slots (pass 1, later passes), then one u32[kSubQueues] set of
    // sub-queue fill counts per pass (pass 1 writes set 0, tail pass p set p).
    void *d_control = nullptr;
    void *d_iters = nullptr;     // = d_control
    void *d_counters = nullptr;  // = d_control + the iteration slots
    hipEvent_t ev_start = nullptr, ev_pass1 = nullptr, ev_end = nullptr;
    hipStream_t last_stream = nullptr;
    bool rendered = false;
    bool replayed = false;                 // last render was a graph replay (no pass-1 event)
    hipGraph_t graph = nullptr;            // captured render (cu_frame_capture)
    hipGraphExec_t graph_exec = nullptr;
    cu_function drain_fn = nullptr;        // drain kernel whose occupancy is cached below
    int drain_blocks_per_cu = 0;           // resident 256-thread drain blocks per CU
};

namespace cu {
namespace {

void *dalloc(size_t bytes) { return device_alloc(bytes); }

void free_frame(cu_frame f) {
    for (void *p : {f->d_tiles, f->d_block_offset, f->d_sample_offset, f->d_smooth_c,
                    f->d_sample_count, f->d_sample_smooth, f->d_pending, f->d_queue[0],
                    f->d_queue[1], f->d_counts, f->d_smooth, f->d_control})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {f->ev_start, f->ev_pass1, f->ev_end})
        if (e) (void)hipEventDestroy(e);
    if (f->graph_exec) (void)hipGraphExecDestroy(f->graph_exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    delete f;
}

void fill(cu_frame f, cu_stream s, void *p, uint32_t value, uint64_t words) {
    fill_u32(f->engine, s, p, value, words);
}

cu_frame_tune checked_tune(const cu_frame_tune *tune, const uint32_t *d_palette,
                           uint32_t palette_length, const uint32_t *d_rgba) {
    cu_frame_tune tn{16, 24, 0, 0, 40};  // pass 1, then one drain of its whole queue
    if (tune) tn = *tune;
    if (tn.block_steps != 8 && tn.block_steps != 16)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "block_steps must be 8 or 16");
    if (tn.handoff_lanes == 0 || tn.handoff_lanes > kMaxHandoff)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "handoff_lanes must be in [1, 32]");
    if (tn.drain_refill > 64) fail(CU_ERR_ILLEGAL_ARGUMENT, "drain_refill must be in [0, 64]");
    if (tn.drain_refill == 0 && (tn.tail_passes == 0 || tn.tail_passes > kMaxTailPasses))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "tail_passes must be in [1, 8]");
    if (tn.drain_refill != 0 && tn.tail_passes >= kMaxTailPasses)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "tail_passes must be in [0, 7] before a drain");
    if (tn.wave_shape >= kWaveShapes) fail(CU_ERR_ILLEGAL_ARGUMENT, "wave_shape must be in [0, 3]");
    if ((d_rgba == nullptr) != (d_palette == nullptr))
        fail(CU_ERR_ILLEGAL_ARGUMENT, "palette and rgba go together");
    if (d_palette && palette_length == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty palette");
    return tn;
}

// All launches of one render on `stream`: clears, pass 1, the tail passes and
// the combine of pixels with queued samples.  `events` records the timing
// events (not while capturing a graph).
void enqueue_render(cu_frame f, const cu_frame_tune &tn, const uint32_t *d_palette,
                    uint32_t palette_length, uint32_t *d_rgba, cu_stream stream, bool events) {
    hipStream_t st = stream ? stream->stream : nullptr;
    const std::string suffix = std::string(real_name(f->precision)) +
                               (f->fused ? "Fused" : "Exact") + "K" +
                               std::to_string(tn.block_steps);
    cu_function pass1 = function(f->engine, "escapePass1" + suffix);
    cu_function passq = function(f->engine, "escapePassQ" + suffix);
    cu_function drain = function(f->engine, "escapeDrain" + suffix);
    cu_function combine = function(f->engine, "combinePending");
    uint32_t *counters = (uint32_t *)f->d_counters;
    const uint32_t n_tiles = (uint32_t)f->tiles.size();

    fill(f, stream, f->d_control, 0, kControlWords);
    if (events) check(hipEventRecord(f->ev_start, st), "hipEventRecord");

    PipeArgs a{(const cu_tile_desc *)f->d_tiles,
               (const uint32_t *)f->d_block_offset + tn.wave_shape * (n_tiles + 1),
               (const uint64_t *)f->d_sample_offset,
               (const double *)f->d_smooth_c,
               (uint32_t *)f->d_sample_count,
               (double *)f->d_sample_smooth,
               (unsigned long long *)f->d_iters,
               (uint32_t *)f->d_counts,
               (double *)f->d_smooth,
               d_rgba,
               d_palette,
               (unsigned char *)f->d_pending,
               n_tiles,
               tn.handoff_lanes,
               f->subq_cap,
               d_palette ? palette_length : 0u,
               tn.wave_shape,
               0u};
    {
        uint32_t total_blocks = f->total_blocks[tn.wave_shape];
        void *q = f->d_queue[0];
        uint32_t *c = counters;
        void *params[] = {&a, &total_blocks, &q, &c};
        launch(f->engine, pass1, dim3((total_blocks + 3) / 4), dim3(256), params, stream, false);
    }
    if (events) check(hipEventRecord(f->ev_pass1, st), "hipEventRecord");
    const uint32_t cus = (uint32_t)std::max(1, compute_units());
    // tail passes p = 1..tail_passes: pass p reads queue (p-1) & 1 with fill set
    // p-1 and writes queue p & 1 with fill set p.  Without a drain the last of
    // them runs every orbit to its end; with one, all hand off and the drain
    // takes what the last left.
    for (uint32_t p = 1; p <= tn.tail_passes; ++p) {
        const uint32_t in = (p - 1) & 1u, outq = p & 1u;
        PipeArgs ap = a;
        ap.handoff = (!tn.drain_refill && p == tn.tail_passes) ? 0u : tn.handoff_lanes;
        ap.iters = (unsigned long long *)f->d_iters + CU_ITER_SLOTS;
        const void *qin = f->d_queue[in];
        const uint32_t *cin = counters + (p - 1) * kSubQueues;
        void *qout = f->d_queue[outq];
        uint32_t *cout = counters + p * kSubQueues;
        void *params[] = {&ap, &qin, &cin, &qout, &cout};
        launch(f->engine, passq, dim3(cus * 8), dim3(256), params, stream, false);
    }
    if (tn.drain_refill) {
        PipeArgs ap = a;
        ap.iters = (unsigned long long *)f->d_iters + CU_ITER_SLOTS;
        const void *q = f->d_queue[tn.tail_passes & 1u];
        const uint32_t *fills = counters + tn.tail_passes * kSubQueues;
        uint32_t *grab = counters + (kMaxTailPasses + 1) * kSubQueues;
        uint32_t refill = tn.drain_refill;
        void *params[] = {&ap, &q, &fills, &grab, &refill};
        // 256 threads per block: the kernel's queue prefix assumes it.  One
        // resident set of blocks: as many per CU as its registers allow, so every
        // wave slot holds a refilling wave and none waits behind the grid.
        if (f->drain_fn != drain) {
            int per_cu = 0;
            check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, drain->function, 256, 0),
                  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
            f->drain_fn = drain;
            f->drain_blocks_per_cu = std::max(1, per_cu);
        }
        launch(f->engine, drain, dim3(cus * (uint32_t)f->drain_blocks_per_cu), dim3(256), params,
               stream, false);
    }
    {
        uint64_t total = f->total_pixels;
        void *params[] = {&a, &total};
        unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, (uint64_t)cus * 128);
        launch(f->engine, combine, dim3(blocks), dim3(256), params, stream, false);
    }
    if (events) check(hipEventRecord(f->ev_end, st), "hipEventRecord");
}

}  // namespace
}  // namespace cu

using namespace cu;

extern "C" {

int cu_frame_create(cu_engine e, const cu_tile_desc *tiles, uint32_t n_tiles, int precision,
                    int variant, cu_frame *out) {
    cu_frame f = nullptr;
    int rc = guarded([&] {
        if (!e || !tiles || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (n_tiles == 0 || n_tiles > kMaxPersistentTiles)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "n_tiles must be in [1, 1024]");
        real_name(precision);
        if (variant < -1 || variant > 1) fail(CU_ERR_ILLEGAL_ARGUMENT, "variant must be -1, 0 or 1");
        f = new cu_frame_s();
        f->engine = e;
        f->precision = precision;
        f->tiles.assign(tiles, tiles + n_tiles);
        std::vector<uint32_t> block_offset(kWaveShapes * (n_tiles + 1), 0);
        std::vector<uint64_t> sample_offset(n_tiles, 0);
        std::vector<double> smooth_c(n_tiles, 0.0);
        bool all_safe = true;
        uint64_t blocks[kWaveShapes] = {0, 0, 0, 0}, pixels = 0, samples = 0;
        for (uint32_t k = 0; k < n_tiles; ++k) {
            cu_tile_desc &t = f->tiles[k];
            t.out_pitch = t.width;
            validate_tile(t);
            if (t.spp > 64 || (t.spp & (t.spp - 1)))
                fail(CU_ERR_ILLEGAL_ARGUMENT, "frame tiles need spp a power of two <= 64");
            if ((uint64_t)t.width * t.height * t.spp >= (1ull << 32))
                fail(CU_ERR_ILLEGAL_ARGUMENT, "tile holds 2^32 samples or more");
            for (uint32_t w = 0; w < kWaveShapes; ++w) {
                uint32_t bw, bh;
                wave_pixel_block(t.spp, w, bw, bh);
                block_offset[w * (n_tiles + 1) + k] = (uint32_t)blocks[w];
                blocks[w] += (uint64_t)((t.width + bw - 1) / bw) * ((t.height + bh - 1) / bh);
            }
            sample_offset[k] = samples;
            t.work_offset = pixels;
            t.out_offset = pixels;
            pixels += (uint64_t)t.width * t.height;
            samples += (uint64_t)t.width * t.height * t.spp;
            smooth_c[k] = smooth_constant(t.bailout_sq);
            all_safe = all_safe && fused_safe(t, precision);
        }
        uint64_t max_blocks = 0;
        for (uint32_t w = 0; w < kWaveShapes; ++w) {
            if (blocks[w] >= (1ull << 31)) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame too large");
            block_offset[w * (n_tiles + 1) + n_tiles] = (uint32_t)blocks[w];
            f->total_blocks[w] = (uint32_t)blocks[w];
            max_blocks = std::max(max_blocks, blocks[w]);
        }
        f->fused = variant == 1 || (variant == -1 && all_safe);
        f->total_pixels = pixels;
        f->total_samples = samples;
        // pass 1 sends wave w to sub-queue w % kSubQueues; a later pass sends a
        // sub-queue's stragglers back to the same sub-queue, never more than it read
        f->subq_cap = (uint32_t)((max_blocks + kSubQueues - 1) / kSubQueues) * kMaxHandoff;
        const size_t queue_records = (size_t)kSubQueues * f->subq_cap;
        const size_t rec = precision == 64 ? 32 : 24;
        ensure_device();
        f->d_tiles = dalloc(n_tiles * sizeof(cu_tile_desc));
        f->d_block_offset = dalloc(block_offset.size() * sizeof(uint32_t));
        f->d_sample_offset = dalloc(n_tiles * sizeof(uint64_t));
        f->d_smooth_c = dalloc(n_tiles * sizeof(double));
        f->d_sample_count = dalloc(samples * 4);
        f->d_sample_smooth = dalloc(samples * 8);
        f->d_pending = dalloc(pixels);
        f->d_queue[0] = dalloc(queue_records * rec);
        f->d_queue[1] = dalloc(queue_records * rec);
        f->d_counts = dalloc(pixels * 4);
        f->d_smooth = dalloc(pixels * 8);
        f->d_control = dalloc(kControlWords * sizeof(uint32_t));
        f->d_iters = f->d_control;
        f->d_counters = (char *)f->d_control + 2 * CU_ITER_SLOTS * 8;
        check(hipMemcpy(f->d_tiles, f->tiles.data(), n_tiles * sizeof(cu_tile_desc),
                        hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipMemcpy(f->d_block_offset, block_offset.data(), block_offset.size() * 4,
                        hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipMemcpy(f->d_sample_offset, sample_offset.data(), sample_offset.size() * 8,
                        hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipMemcpy(f->d_smooth_c, smooth_c.data(), smooth_c.size() * 8,
                        hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipEventCreate(&f->ev_start), "hipEventCreate");
        check(hipEventCreate(&f->ev_pass1), "hipEventCreate");
        check(hipEventCreate(&f->ev_end), "hipEventCreate");
        *out = f;
    });
    if (rc != CU_OK && f) free_frame(f);
    return rc;
}

int cu_frame_destroy(cu_frame f) {
    return guarded([&] {
        if (!f) return;
        ensure_device();
        free_frame(f);
    });
}

int cu_frame_render(cu_frame f, const cu_frame_tune *tune, const uint32_t *d_palette,
                    uint32_t palette_length, uint32_t *d_rgba, cu_stream stream) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame is NULL");
        const cu_frame_tune tn = checked_tune(tune, d_palette, palette_length, d_rgba);
        std::lock_guard<std::mutex> lock(f->engine->mutex);
        ensure_device();
        enqueue_render(f, tn, d_palette, palette_length, d_rgba, stream, true);
        f->last_stream = stream ? stream->stream : nullptr;
        f->rendered = true;
        f->replayed = false;
    });
}

int cu_frame_capture(cu_frame f, const cu_frame_tune *tune, const uint32_t *d_palette,
                     uint32_t palette_length, uint32_t *d_rgba) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame is NULL");
        const cu_frame_tune tn = checked_tune(tune, d_palette, palette_length, d_rgba);
        std::lock_guard<std::mutex> lock(f->engine->mutex);
        ensure_device();
        hipStream_t cs = nullptr;
        check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        cu_stream_s wrap{cs};
        hipGraph_t g = nullptr;
        try {
            check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_render(f, tn, d_palette, palette_length, d_rgba, &wrap, false);
            } catch (...) {
                (void)hipStreamEndCapture(cs, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            check(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
            hipGraphExec_t exec = nullptr;
            check(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
            if (f->graph_exec) (void)hipGraphExecDestroy(f->graph_exec);
            if (f->graph) (void)hipGraphDestroy(f->graph);
            f->graph = g;
            f->graph_exec = exec;
        } catch (...) {
            if (g && g != f->graph) (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cs);
            throw;
        }
        check(hipStreamDestroy(cs), "hipStreamDestroy");
    });
}

int cu_frame_replay(cu_frame f, cu_stream stream) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame is NULL");
        if (!f->graph_exec) fail(CU_ERR_ILLEGAL_STATE, "frame has no captured render");
        std::lock_guard<std::mutex> lock(f->engine->mutex);
        ensure_device();
        hipStream_t st = stream ? stream->stream : nullptr;
        check(hipEventRecord(f->ev_start, st), "hipEventRecord");
        check(hipGraphLaunch(f->graph_exec, st), "hipGraphLaunch");
        check(hipEventRecord(f->ev_end, st), "hipEventRecord");
        f->last_stream = st;
        f->rendered = true;
        f->replayed = true;
    });
}

int cu_frame_outputs(cu_frame f, uint32_t **d_counts, double **d_smooth, uint64_t *total_pixels) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame is NULL");
        if (d_counts) *d_counts = (uint32_t *)f->d_counts;
        if (d_smooth) *d_smooth = (double *)f->d_smooth;
        if (total_pixels) *total_pixels = f->total_pixels;
    });
}

int cu_frame_iterations(cu_frame f, unsigned long long *total, unsigned long long *pass1) {
    return guarded([&] {
        if (!f || !total) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (!f->rendered) fail(CU_ERR_ILLEGAL_STATE, "frame has not been rendered");
        ensure_device();
        check(hipStreamSynchronize(f->last_stream), "hipStreamSynchronize");
        const unsigned long long *slots = (const unsigned long long *)f->d_iters;
        unsigned long long first = 0, rest = 0;
        if (cu_iters_total(slots, &first) != CU_OK || cu_iters_total(slots + CU_ITER_SLOTS, &rest) != CU_OK)
            fail(CU_ERR_DEVICE, cu_last_error());
        *total = first + rest;
        if (pass1) *pass1 = first;
    });
}

int cu_frame_last_ms(cu_frame f, float *pass1_ms, float *total_ms) {
    return guarded([&] {
        if (!f) fail(CU_ERR_ILLEGAL_ARGUMENT, "frame is NULL");
        if (!f->rendered) fail(CU_ERR_ILLEGAL_STATE, "frame has not been rendered");
        ensure_device();
        check(hipEventSynchronize(f->ev_end), "hipEventSynchronize");
        if (pass1_ms) {
            if (f->replayed) *pass1_ms = 0.0f;  // a graph replay records no pass-1 event
            else check(hipEventElapsedTime(pass1_ms, f->ev_start, f->ev_pass1), "hipEventElapsedTime");
        }
        if (total_ms) check(hipEventElapsedTime(total_ms, f->ev_start, f->ev_end), "hipEventElapsedTime");
    });
}

}  // extern "C"
