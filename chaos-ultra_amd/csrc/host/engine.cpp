// engine.cpp — layer 3 of the C ABI: launches of the escape-family module
// (escape_family.hsaco) for the planar tile forms, the persistent launch and
// the smooth compose, with HIP-event timing on the launch stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "engine_helpers.hpp"
#include "handles.hpp"

namespace cu {

cu_function function(cu_engine e, const std::string &name) {
    cu_function f = nullptr;
    int rc = cu_module_get_function(e->module, name.c_str(), &f);
    if (rc != CU_OK) fail(rc, cu_last_error());
    return f;
}

hipStream_t stream_of(cu_stream s) { return s ? s->stream : nullptr; }

void launch(cu_engine e, cu_function f, dim3 grid, dim3 block, void **params, cu_stream s,
            bool timed) {
    hipStream_t st = stream_of(s);
    if (timed) check(hipEventRecord(e->start, st), "hipEventRecord");
    check(hipModuleLaunchKernel(f->function, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0,
                                st, params, nullptr),
          (std::string("hipModuleLaunchKernel ") + f->name).c_str());
    if (timed) {
        check(hipEventRecord(e->stop, st), "hipEventRecord");
        e->timed = true;
    }
}

const char *real_name(int precision) {
    if (precision == 64) return "Double";
    if (precision == 32) return "Float";
    fail(CU_ERR_ILLEGAL_ARGUMENT, "precision must be 32 or 64");
}

bool ok_magnitude(double v, double threshold) { return v == 0 || std::fabs(v) >= threshold; }

// Sufficient condition for the fused step: the plane origin, pixel sizes and a
// Julia c are each zero or far from the subnormal range, so every sample point
// and orbit component is zero or at least ~2^-360 (FP64) / ~2^-35 (FP32) in
// magnitude and no product x*y of the orbit reaches the subnormal range.
bool fused_safe(const cu_tile_desc &t, int precision) {
    if (precision == 64) {
        const double T = std::ldexp(1.0, -300);
        bool ok = ok_magnitude(t.left, T) && ok_magnitude(t.top, T) && t.pixel_w >= T &&
                  t.pixel_h >= T;
        if (t.kind == CU_FRACTAL_JULIA) ok = ok && ok_magnitude(t.julia_cx, T) && ok_magnitude(t.julia_cy, T);
        return ok;
    }
    const double T = std::ldexp(1.0, -11);
    bool ok = std::fabs((double)t.left_f) >= T && std::fabs((double)t.top_f) >= T;
    if (t.kind == CU_FRACTAL_JULIA)
        ok = ok && ok_magnitude((double)(float)t.julia_cx, T) && ok_magnitude((double)(float)t.julia_cy, T);
    return ok;
}

void validate_tile(const cu_tile_desc &t) {
    if (t.width == 0 || t.height == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "tile has no pixels");
    if (t.spp == 0 || t.spp > 4096) fail(CU_ERR_ILLEGAL_ARGUMENT, "spp must be in [1, 4096]");
    if (t.max_iter == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "max_iter must be positive");
    if (t.kind > CU_FRACTAL_JULIA) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown fractal kind");
    if (t.pattern > CU_SAMPLE_CORNER) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown sample pattern");
    if (t.out_pitch < t.width) fail(CU_ERR_ILLEGAL_ARGUMENT, "out_pitch smaller than width");
    for (double v : {t.left, t.top, t.pixel_w, t.pixel_h, t.julia_cx, t.julia_cy, t.bailout_sq})
        if (!std::isfinite(v)) fail(CU_ERR_ILLEGAL_ARGUMENT, "tile parameters must be finite");
    if (!(t.bailout_sq > 0)) fail(CU_ERR_ILLEGAL_ARGUMENT, "bailout_sq must be positive");
}

void validate_tune(const cu_tune &tn) {
    if (tn.block_steps != 8 && tn.block_steps != 16 && tn.block_steps != 32)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "block_steps must be 8, 16 or 32");
    if (tn.waves_per_cu < 4 || tn.waves_per_cu > 32 || tn.waves_per_cu % 4)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "waves_per_cu must be a multiple of 4 in [4, 32]");
    if (tn.chunk == 0 || tn.chunk > (1u << 20)) fail(CU_ERR_ILLEGAL_ARGUMENT, "chunk must be in [1, 2^20]");
    if (tn.refill_lanes == 0 || tn.refill_lanes > 64)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "refill_lanes must be in [1, 64]");
}

void *device_alloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) return nullptr;
    check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
}

void fill_u32(cu_engine e, cu_stream s, void *p, uint32_t value, uint64_t words) {
    cu_function fn = function(e, "fillU32");
    void *params[] = {&p, &value, &words};
    const unsigned blocks = (unsigned)std::min<uint64_t>((words + 255) / 256, 1024);
    launch(e, fn, dim3(blocks), dim3(256), params, s, false);
}

bool pick_fused(int variant, const cu_tile_desc &t, int precision) {
    if (variant == 0) return false;
    if (variant == 1) return true;
    if (variant == -1) return fused_safe(t, precision);
    fail(CU_ERR_ILLEGAL_ARGUMENT, "variant must be -1, 0 or 1");
}

// Arguments of the persistent launch, checked; the tuning it runs with.
cu_tune persistent_args(cu_engine e, int variant, const cu_tile_desc *d_tiles, uint32_t n_tiles,
                        uint32_t *d_counts, unsigned long long *d_queue, const cu_tune *tune) {
    if (!e || !d_tiles || !d_counts || !d_queue) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
    if (n_tiles == 0 || n_tiles > kMaxPersistentTiles)
        fail(CU_ERR_ILLEGAL_ARGUMENT, "n_tiles must be in [1, 1024]");
    if (variant < -1 || variant > 1) fail(CU_ERR_ILLEGAL_ARGUMENT, "variant must be -1, 0 or 1");
    if (variant == -1)
        fail(CU_ERR_ILLEGAL_ARGUMENT,
             "the persistent launch takes an explicit variant (check each tile with cu_tile_fused_safe)");
    cu_tune tn = e->tune;
    if (tune) {
        validate_tune(*tune);
        tn = *tune;
    }
    return tn;
}

// The persistent launch on `stream`: the queue counter cleared by a kernel (safe
// under graph capture), with clear_iters the iteration slots too, then the
// kernel itself.
void enqueue_persistent(cu_engine e, int precision, int variant, const cu_tile_desc *d_tiles,
                        uint32_t n_tiles, uint64_t total, uint32_t *d_counts, double *d_smooth,
                        unsigned long long *d_iters, unsigned long long *d_queue, const cu_tune &tn,
                        cu_stream stream, bool clear_iters, bool timed) {
    std::string name = std::string("escapeTilesPersistent") + real_name(precision) +
                       (variant == 1 ? "Fused" : "Exact") + "K" + std::to_string(tn.block_steps);
    cu_function f = function(e, name);
    fill_u32(e, stream, d_queue, 0, 2);
    if (clear_iters && d_iters) fill_u32(e, stream, d_iters, 0, (uint64_t)CU_ITER_SLOTS * 2);
    const uint32_t cus = (uint32_t)std::max(1, compute_units());
    dim3 block(256, 1, 1);
    dim3 grid(cus * tn.waves_per_cu / 4, 1, 1);
    uint32_t chunk = tn.chunk, refill = tn.refill_lanes;
    void *params[] = {&d_tiles, &n_tiles, &total, &d_counts, &d_smooth, &d_iters,
                      &d_queue, &chunk, &refill};
    launch(e, f, grid, block, params, stream, timed);
}

}  // namespace cu

using namespace cu;

extern "C" {

int cu_tile_init(cu_tile_desc *t, double lbx, double lby, double rtx, double rty, uint32_t w,
                 uint32_t h, uint32_t max_iter, double bailout_sq, uint32_t kind, double jcx,
                 double jcy, uint32_t spp, uint32_t pattern, uint64_t seed) {
    return guarded([&] {
        if (!t) fail(CU_ERR_ILLEGAL_ARGUMENT, "tile is NULL");
        if (w == 0 || h == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "tile has no pixels");
        std::memset(t, 0, sizeof *t);
        t->left = lbx;
        t->top = rty;
        t->pixel_w = (rtx - lbx) / (double)w;  // image.size() / gridSize (fractalRendererGeneric.cu:97)
        t->pixel_h = (rty - lby) / (double)h;
        const float lf = (float)lbx, bf = (float)lby, rf = (float)rtx, tf = (float)rty;
        t->left_f = lf;
        t->top_f = tf;
        t->pixel_w_f = (rf - lf) / (float)w;
        t->pixel_h_f = (tf - bf) / (float)h;
        t->julia_cx = jcx;
        t->julia_cy = jcy;
        t->bailout_sq = bailout_sq;
        t->seed = seed;
        t->width = w;
        t->height = h;
        t->out_pitch = w;
        t->max_iter = max_iter;
        t->kind = kind;
        t->spp = spp;
        t->pattern = pattern;
        validate_tile(*t);
    });
}

int cu_tile_fused_safe(const cu_tile_desc *t, int precision, int *safe) {
    return guarded([&] {
        if (!t || !safe) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        real_name(precision);
        *safe = fused_safe(*t, precision) ? 1 : 0;
    });
}

int cu_engine_create(const char *dir, cu_engine *out) {
    return guarded([&] {
        if (!dir || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "dir/out is NULL");
        std::string path = std::string(dir) + "/escape_family.hsaco";
        cu_module m = nullptr;
        int rc = cu_module_load(path.c_str(), &m);
        if (rc != CU_OK) fail(rc, cu_last_error());
        auto *e = new cu_engine_s();
        e->module = m;
        check(hipEventCreate(&e->start), "hipEventCreate");
        check(hipEventCreate(&e->stop), "hipEventCreate");
        *out = e;
    });
}

int cu_engine_destroy(cu_engine e) {
    return guarded([&] {
        if (!e) return;
        ensure_device();
        if (e->start) (void)hipEventDestroy(e->start);
        if (e->stop) (void)hipEventDestroy(e->stop);
        cu_module m = e->module;
        delete e;
        if (m && cu_module_unload(m) != CU_OK) fail(CU_ERR_DEVICE, cu_last_error());
    });
}

int cu_escape_tile(cu_engine e, int precision, int variant, const cu_tile_desc *tile,
                   uint32_t *d_counts, double *d_smooth, unsigned long long *d_iters,
                   cu_stream stream) {
    return guarded([&] {
        if (!e || !tile || !d_counts) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(e->mutex);
        validate_tile(*tile);
        ensure_device();
        const bool fused = pick_fused(variant, *tile, precision);
        cu_tile_desc t = *tile;
        double smooth_c = smooth_constant(t.bailout_sq);
        void *params[] = {&t, &smooth_c, &d_counts, &d_smooth, &d_iters};
        const bool lane_per_sample = t.spp <= 64 && (t.spp & (t.spp - 1)) == 0;
        std::string name = std::string(lane_per_sample ? "escapeTileSamples" : "escapeTile") +
                           real_name(precision) + (fused ? "Fused" : "Exact") +
                           (lane_per_sample && e->tune.block_steps == 8 ? "K8" : "");
        cu_function f = function(e, name);
        if (lane_per_sample) {
            // a wave covers bw x bh pixels (bw*bh = 64/spp), a block 2x2 waves
            uint32_t ppw = 64 / t.spp, lg = 0;
            while ((1u << (lg + 1)) <= ppw) ++lg;
            uint32_t bw = 1u << ((lg + 1) >> 1), bh = ppw / bw;
            dim3 grid((t.width + 2 * bw - 1) / (2 * bw), (t.height + 2 * bh - 1) / (2 * bh), 1);
            launch(e, f, grid, dim3(256, 1, 1), params, stream, true);
        } else {
            dim3 grid((t.width + 15) / 16, (t.height + 15) / 16, 1);
            launch(e, f, grid, dim3(16, 16, 1), params, stream, true);
        }
    });
}

int cu_escape_tiles_persistent(cu_engine e, int precision, int variant, const cu_tile_desc *d_tiles,
                               uint32_t n_tiles, uint64_t total, uint32_t *d_counts,
                               double *d_smooth, unsigned long long *d_iters,
                               unsigned long long *d_queue, const cu_tune *tune, cu_stream stream) {
    return guarded([&] {
        const cu_tune tn = persistent_args(e, variant, d_tiles, n_tiles, d_counts, d_queue, tune);
        std::lock_guard<std::mutex> lock(e->mutex);
        ensure_device();
        if (total == 0) return;
        enqueue_persistent(e, precision, variant, d_tiles, n_tiles, total, d_counts, d_smooth,
                           d_iters, d_queue, tn, stream, false, true);
    });
}

int cu_escape_tiles_persistent_capture(cu_engine e, int precision, int variant,
                                       const cu_tile_desc *d_tiles, uint32_t n_tiles,
                                       uint64_t total, uint32_t *d_counts, double *d_smooth,
                                       unsigned long long *d_iters, unsigned long long *d_queue,
                                       const cu_tune *tune, cu_graph *out) {
    cu_graph g = nullptr;
    int rc = guarded([&] {
        if (!out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        const cu_tune tn = persistent_args(e, variant, d_tiles, n_tiles, d_counts, d_queue, tune);
        if (total == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "no pixels to capture");
        (void)real_name(precision);
        std::lock_guard<std::mutex> lock(e->mutex);
        ensure_device();
        g = new cu_graph_s();
        check(hipEventCreate(&g->start), "hipEventCreate");
        check(hipEventCreate(&g->stop), "hipEventCreate");
        hipStream_t cs = nullptr;
        check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        cu_stream_s wrap{cs};
        try {
            check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueue_persistent(e, precision, variant, d_tiles, n_tiles, total, d_counts, d_smooth,
                                   d_iters, d_queue, tn, &wrap, true, false);
            } catch (...) {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(cs, &partial);
                if (partial) (void)hipGraphDestroy(partial);
                throw;
            }
            check(hipStreamEndCapture(cs, &g->graph), "hipStreamEndCapture");
            check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        } catch (...) {
            (void)hipStreamDestroy(cs);
            throw;
        }
        check(hipStreamDestroy(cs), "hipStreamDestroy");
        *out = g;
    });
    if (rc != CU_OK && g) (void)cu_graph_destroy(g);
    return rc;
}

int cu_graph_launch(cu_graph g, cu_stream stream) {
    return guarded([&] {
        if (!g || !g->exec) fail(CU_ERR_ILLEGAL_ARGUMENT, "graph is NULL");
        ensure_device();
        hipStream_t st = stream ? stream->stream : nullptr;
        check(hipEventRecord(g->start, st), "hipEventRecord");
        check(hipGraphLaunch(g->exec, st), "hipGraphLaunch");
        check(hipEventRecord(g->stop, st), "hipEventRecord");
        g->launched = true;
    });
}

int cu_graph_last_ms(cu_graph g, float *ms) {
    return guarded([&] {
        if (!g || !ms) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (!g->launched) fail(CU_ERR_ILLEGAL_STATE, "graph has not been launched");
        ensure_device();
        check(hipEventSynchronize(g->stop), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, g->start, g->stop), "hipEventElapsedTime");
    });
}

int cu_graph_destroy(cu_graph g) {
    return guarded([&] {
        if (!g) return;
        ensure_device();
        if (g->launched) (void)hipEventSynchronize(g->stop);  // the last replay has ended
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->start) (void)hipEventDestroy(g->start);
        if (g->stop) (void)hipEventDestroy(g->stop);
        delete g;
    });
}

int cu_compose_smooth(cu_engine e, const double *d_smooth, uint32_t smooth_pitch, uint32_t *d_rgba,
                      uint32_t rgba_pitch, uint32_t w, uint32_t h, const uint32_t *d_palette,
                      uint32_t palette_length, cu_stream stream) {
    return guarded([&] {
        if (!e || !d_smooth || !d_rgba || !d_palette) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (palette_length == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty palette");
        if (smooth_pitch < w || rgba_pitch < w) fail(CU_ERR_ILLEGAL_ARGUMENT, "pitch smaller than width");
        if (w == 0 || h == 0) return;
        std::lock_guard<std::mutex> lock(e->mutex);
        ensure_device();
        cu_function f = function(e, "composeSmooth");
        void *params[] = {&d_smooth, &smooth_pitch, &d_rgba, &rgba_pitch, &w, &h, &d_palette,
                          &palette_length};
        dim3 block(64, 4, 1);
        dim3 grid((w + 63) / 64, std::min<uint32_t>((h + 3) / 4, 4096), 1);
        launch(e, f, grid, block, params, stream, true);
    });
}

int cu_engine_configure(cu_engine e, const cu_tune *tune) {
    return guarded([&] {
        if (!e || !tune) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        validate_tune(*tune);
        std::lock_guard<std::mutex> lock(e->mutex);
        e->tune = *tune;
    });
}

int cu_iters_total(const unsigned long long *d_iters, unsigned long long *total) {
    return guarded([&] {
        if (!d_iters || !total) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        ensure_device();
        unsigned long long slots[CU_ITER_SLOTS];
        check(hipMemcpy(slots, d_iters, sizeof slots, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        unsigned long long s = 0;
        for (unsigned long long v : slots) s += v;
        *total = s;
    });
}

int cu_engine_last_kernel_ms(cu_engine e, float *ms) {
    return guarded([&] {
        if (!e || !ms) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (!e->timed) fail(CU_ERR_ILLEGAL_STATE, "no timed launch yet");
        ensure_device();
        check(hipEventSynchronize(e->stop), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, e->start, e->stop), "hipEventElapsedTime");
    });
}

}  // extern "C"
