// serve.cpp — a continuous-batching request loop over zoom sessions (cu_serve_*).
//
// Counterpart of the reference's single-session display loop
// (rendering/GLRenderer.java:113-162 driving CudaFractalRenderer.renderFast /
// renderQuality, cudarenderer/CudaFractalRenderer.java:158-206), for many
// sessions at once on one GPU:
//
//   * a session owns the reference's ping-pong pixel_info_t buffers
//     (cudarenderer/DeviceMemoryDoubleBuffer2D.java:17-152), an RGBA8 image, a
//     palette and a deep-zoom frame; they stay resident between its frames so
//     each frame reuses the previous one's samples;
//   * sessions are admitted against a fixed HBM budget (their whole footprint
//     is reserved at admission and released at close); a session that does not
//     fit waits, first come first served;
//   * frame requests join their session's queue; whenever one of `slots` HIP
//     streams is free, the oldest request of an idle admitted session starts on
//     it, so frames of different sessions overlap on the device while each
//     session's frames run in order (a session's results do not depend on what
//     else runs);
//   * a frame at a plane width >= deep_width runs the module's
//     fractalRenderAdvanced kernel with sample reuse and foveation around the
//     view centre (fractalRenderMain for a session's first frame, or after a
//     deep frame), then compose; a deeper frame runs the perturbation form;
//   * a frame's completion is an event on its stream; its checksum (of the
//     RGBA image) comes back through pinned host memory with it;
//   * a request that passed cu_serve_submit's checks but fails later (a device
//     allocation at admission, a launch) completes with a non-zero status: a
//     failure belongs to its request and never to another session's call;
//   * cu_serve_poll takes the server's lock only for each round of retire /
//     admit / dispatch, so a submit from another thread waits for one round at
//     most, not for the poll's whole wait.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "engine_helpers.hpp"
#include "handles.hpp"

namespace {

constexpr uint32_t kMaxSlots = 16;
constexpr int kBlockDim = 32;  // reference launch geometry (CudaFractalRenderer.java:36-37)
constexpr int kStripRows = 4;  // rows of a block of the render kernels' 32 x 4-thread form

template <class T>
struct Pt {
    T x, y;
};
template <class T>
struct Rect {
    T lbx, lby, rtx, rty;
};

struct Request {
    cu_frame_request req;
    double submit_s;
};

struct Session {
    uint32_t id = 0;
    uint32_t width = 0, height = 0, deep_max_iter = 0;
    int state = 0;  // 0 waiting for admission, 1 admitted, 2 closing, 3 closed
    size_t bytes = 0;
    std::vector<uint32_t> host_palette;
    void *buf[2] = {nullptr, nullptr};
    size_t pitch[2] = {0, 0};
    int primary = 0;
    void *rgba = nullptr, *palette = nullptr;
    cu_deep deep = nullptr;
    bool has_prev = false;
    Rect<double> prev{};
    std::deque<Request> queue;
    bool busy = false;
    uint64_t frames = 0;
    double open_s = 0, admit_s = -1;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    cu_stream_s wrap{};
    bool busy = false;
    uint32_t session = 0;
    cu_frame_done info{};
    void *d_sum = nullptr;
    unsigned long long *h_sum = nullptr;  // pinned
};

}  // namespace

struct cu_serve_s {
    cu_engine engine = nullptr;
    cu_module module = nullptr;
    cu_function main_f = nullptr, main_d = nullptr, adv_f = nullptr, adv_d = nullptr,
                compose = nullptr, checksum = nullptr;
    // the render kernels for 32 x 4-thread blocks (null if the module lacks them)
    cu_function main_fs = nullptr, main_ds = nullptr, adv_fs = nullptr, adv_ds = nullptr;
    void *d_module_iters = nullptr;  // module symbol cu_iteration_slots
    cu_serve_config cfg{};
    std::vector<Slot> slots;
    std::vector<Session> sessions;   // index = session id - 1
    size_t used = 0, peak = 0;
    unsigned long long deep_iterations = 0;
    uint64_t frames_done = 0;
    std::deque<cu_frame_done> failed;  // completions of failed requests, reported by poll
    std::string last_failure;
    std::chrono::steady_clock::time_point t0;
    std::mutex mutex;
};

namespace cu {
namespace {

double now_s(const cu_serve_s *s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

size_t pitch_of(uint32_t width) {  // cuMemAllocPitch-like: rows aligned to 256 bytes
    const size_t row = (size_t)width * 16;
    return (row + 255) / 256 * 256;
}

size_t session_bytes(uint32_t w, uint32_t h, uint32_t palette_len, uint32_t deep_max_iter) {
    size_t b = 2 * pitch_of(w) * h + (size_t)w * h * 4 + (size_t)palette_len * 4;
    if (deep_max_iter)  // deep frame: counts, smooth, glitch planes, orbit, counters
        b += (size_t)w * h * 13 + ((size_t)deep_max_iter + 1) * 32 + CU_ITER_SLOTS * 8 + 8;
    return b;
}

// The deep-frame description of request r on session ss (bailout 4, rebasing,
// the orbit kept while the centre stays).
cu_deep_desc deep_desc(const Session &ss, const cu_frame_request &r) {
    const double pw = r.width / ss.width;
    return cu_deep_desc{r.centre_x, r.centre_y, 0, 0, pw, pw, 4.0, ss.width, ss.height, r.max_iter,
                        CU_DEEP_REBASE | CU_DEEP_REUSE_ORBIT, 0, 1e-6f};
}

// Complete request rq of session ss as failed with `status`.
void fail_request(cu_serve_s *s, const Session &ss, const Request &rq, int status,
                  const std::string &message) {
    cu_frame_done d{};
    d.tag = rq.req.tag;
    d.session = ss.id;
    d.path = rq.req.width < s->cfg.deep_width ? 2u : (ss.has_prev ? 0u : 1u);
    d.submit_s = rq.submit_s;
    d.start_s = d.done_s = now_s(s);
    d.status = status;
    s->failed.push_back(d);
    s->last_failure = message;
}

void release(cu_serve_s *s, Session &ss) {
    for (void *p : {ss.buf[0], ss.buf[1], ss.rgba, ss.palette})
        if (p) (void)hipFree(p);
    ss.buf[0] = ss.buf[1] = ss.rgba = ss.palette = nullptr;
    if (ss.deep) (void)cu_deep_destroy(ss.deep);
    ss.deep = nullptr;
    if (ss.state == 1 || ss.state == 2) s->used -= ss.bytes;
    ss.state = 3;
}

// Allocate a waiting session's buffers and reserve its footprint (throws CuError).
void admit_one(cu_serve_s *s, Session &ss) {
    for (int k = 0; k < 2; ++k) {
        ss.pitch[k] = pitch_of(ss.width);
        ss.buf[k] = device_alloc(ss.pitch[k] * ss.height);
        check(hipMemset(ss.buf[k], 0, ss.pitch[k] * ss.height), "hipMemset");
    }
    ss.rgba = device_alloc((size_t)ss.width * ss.height * 4);
    ss.palette = device_alloc(ss.host_palette.size() * 4);
    check(hipMemcpy(ss.palette, ss.host_palette.data(), ss.host_palette.size() * 4,
                    hipMemcpyHostToDevice), "hipMemcpy H2D");
    if (ss.deep_max_iter) {
        const double pw = 1e-13 / ss.width;
        cu_deep_desc d{-0.75, 0.0, 0, 0, pw, pw, 4.0, ss.width, ss.height, ss.deep_max_iter,
                       CU_DEEP_REBASE | CU_DEEP_REUSE_ORBIT, 0, 1e-6f};
        if (cu_deep_create(s->engine, &d, (int)s->cfg.deep_method, &ss.deep) != CU_OK)
            fail(CU_ERR_DEVICE, cu_last_error());
    }
    ss.state = 1;
    ss.admit_s = now_s(s);
    s->used += ss.bytes;
    s->peak = std::max(s->peak, s->used);
}

// Admit waiting sessions in order while they fit the budget.  A session whose
// allocation fails is closed and its queued requests complete as failed.
void admit(cu_serve_s *s) {
    for (Session &ss : s->sessions) {
        if (ss.state != 0) continue;
        if (s->used + ss.bytes > s->cfg.hbm_budget) break;  // first come first served
        try {
            admit_one(s, ss);
        } catch (const CuError &e) {
            release(s, ss);  // frees what was allocated; state 0 -> 3, nothing was reserved
            for (const Request &rq : ss.queue) fail_request(s, ss, rq, e.status, e.message);
            ss.queue.clear();
        }
    }
}

void launch_module(cu_function f, unsigned gx, unsigned gy, void **params, hipStream_t st) {
    check(hipModuleLaunchKernel(f->function, gx, gy, 1, kBlockDim, kBlockDim, 1, 0, st, params,
                                nullptr),
          (std::string("hipModuleLaunchKernel ") + f->name).c_str());
}

// A render kernel over a W x H frame: its 32 x 4-thread form over ceil(W/32) x
// ceil(H/4) blocks when the module has one (the same pixels and results), else
// the reference's 32 x 32 blocks.
void launch_render(cu_function f, cu_function strip, uint32_t W, uint32_t H, void **params,
                   hipStream_t st) {
    const unsigned gx = (W + kBlockDim - 1) / kBlockDim;
    if (strip) {
        check(hipModuleLaunchKernel(strip->function, gx, (H + kStripRows - 1) / kStripRows, 1,
                                    kBlockDim, kStripRows, 1, 0, st, params, nullptr),
              (std::string("hipModuleLaunchKernel ") + strip->name).c_str());
    } else {
        launch_module(f, gx, (H + kBlockDim - 1) / kBlockDim, params, st);
    }
}

// One frame of session ss on slot sl (asynchronous on the slot's stream).
void dispatch(cu_serve_s *s, Session &ss, Slot &sl, const Request &rq) {
    const cu_frame_request &r = rq.req;
    const uint32_t W = ss.width, H = ss.height;
    hipStream_t st = sl.stream;
    const double w = r.width, h = r.width * (double)H / (double)W;
    const Rect<double> rect{r.centre_x - 0.5 * w, r.centre_y - 0.5 * h, r.centre_x + 0.5 * w,
                            r.centre_y + 0.5 * h};
    uint32_t path = 0;
    if (r.width < s->cfg.deep_width) {
        if (!ss.deep) fail(CU_ERR_ILLEGAL_ARGUMENT, "session opened without deep frames");
        const cu_deep_desc d = deep_desc(ss, r);
        if (cu_deep_retarget(ss.deep, &d) != CU_OK) fail(CU_ERR_ILLEGAL_ARGUMENT, cu_last_error());
        if (cu_deep_render(ss.deep, (const uint32_t *)ss.palette, (uint32_t)ss.host_palette.size(),
                           (uint32_t *)ss.rgba, &sl.wrap) != CU_OK)
            fail(CU_ERR_DEVICE, cu_last_error());
        ss.has_prev = false;  // the reuse buffers do not hold this frame
        path = 2;
    } else {
        int precision = CU_PRECISION_DOUBLE;
        if (s->cfg.shallow_policy == CU_SERVE_SHALLOW_REFERENCE_SWITCH &&
            cu_precision_for_segment(rect.lbx, rect.lby, rect.rtx, rect.rty, (int)W, (int)H,
                                     &precision) != CU_OK)
            fail(CU_ERR_ILLEGAL_ARGUMENT, cu_last_error());
        const bool fp32 = precision == CU_PRECISION_SINGLE;
        Pt<uint32_t> size{W, H};
        Rect<float> rf{(float)rect.lbx, (float)rect.lby, (float)rect.rtx, (float)rect.rty};
        Rect<double> rd = rect;
        uint32_t maxit = r.max_iter;
        float max_ss = s->cfg.max_super_sampling;
        const unsigned gx = (W + kBlockDim - 1) / kBlockDim, gy = (H + kBlockDim - 1) / kBlockDim;
        if (!ss.has_prev) {  // renderQuality: main kernel into the primary buffer
            void *out = ss.buf[ss.primary];
            long pitch = (long)ss.pitch[ss.primary];
            uint32_t flags = CU_FLAG_ADAPTIVE_SS;
            void *params[] = {&out, &pitch, &size, fp32 ? (void *)&rf : (void *)&rd, &maxit,
                              &max_ss, &flags};
            launch_render(fp32 ? s->main_f : s->main_d, fp32 ? s->main_fs : s->main_ds, W, H, params,
                          st);
            path = 1;
        } else {  // renderFast: advanced kernel primary -> secondary, then swap
            void *in = ss.buf[ss.primary];
            long in_pitch = (long)ss.pitch[ss.primary];
            void *out = ss.buf[1 - ss.primary];
            long out_pitch = (long)ss.pitch[1 - ss.primary];
            uint32_t flags = CU_FLAG_ADAPTIVE_SS | CU_FLAG_SAMPLE_REUSE | CU_FLAG_ZOOMING |
                             CU_FLAG_ZOOMING_IN | CU_FLAG_FOVEATION;
            Rect<float> of{(float)ss.prev.lbx, (float)ss.prev.lby, (float)ss.prev.rtx,
                           (float)ss.prev.rty};
            Rect<double> od = ss.prev;
            Pt<uint32_t> focus{W / 2, H / 2};
            void *params[] = {&out, &out_pitch, &size, fp32 ? (void *)&rf : (void *)&rd, &maxit,
                              &max_ss, &flags, fp32 ? (void *)&of : (void *)&od, &in, &in_pitch,
                              &focus};
            launch_render(fp32 ? s->adv_f : s->adv_d, fp32 ? s->adv_fs : s->adv_ds, W, H, params, st);
            ss.primary = 1 - ss.primary;
            path = 0;
        }
        void *main = ss.buf[ss.primary];
        long main_pitch = (long)ss.pitch[ss.primary];
        void *bcg = ss.buf[1 - ss.primary];
        long bcg_pitch = (long)ss.pitch[1 - ss.primary];
        void *rgba = ss.rgba;
        void *pal = ss.palette;
        uint32_t plen = (uint32_t)ss.host_palette.size();
        uint32_t w32 = W, h32 = H;
        void *cparams[] = {&main, &main_pitch, &bcg, &bcg_pitch, &rgba, &w32, &h32, &pal, &plen,
                           &max_ss};
        launch_module(s->compose, gx, gy, cparams, st);
        ss.has_prev = true;
        ss.prev = rect;
    }
    // checksum of the image, copied back with the completion
    fill_u32(s->engine, &sl.wrap, sl.d_sum, 0, 2);
    {
        const void *p = ss.rgba;
        uint64_t n = (uint64_t)W * H;
        void *out = sl.d_sum;
        void *params[] = {&p, &n, &out};
        const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
        launch(s->engine, s->checksum, dim3(blocks), dim3(256), params, &sl.wrap, false);
    }
    check(hipMemcpyAsync(sl.h_sum, sl.d_sum, 8, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipEventRecord(sl.done, st), "hipEventRecord");
    sl.busy = true;
    sl.session = ss.id;
    sl.info = cu_frame_done{};
    sl.info.tag = r.tag;
    sl.info.session = ss.id;
    sl.info.path = path;
    sl.info.submit_s = rq.submit_s;
    sl.info.start_s = now_s(s);
    ss.busy = true;
}

// Retire failed requests and finished slots into out[]; returns how many were
// written.  Never throws: a frame whose completion reports an error is returned
// with that status.
uint32_t retire(cu_serve_s *s, cu_frame_done *out, uint32_t capacity, uint32_t n) {
    while (!s->failed.empty() && n < capacity) {
        out[n++] = s->failed.front();
        s->failed.pop_front();
    }
    for (Slot &sl : s->slots) {
        if (!sl.busy || n >= capacity) continue;
        hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        Session &ss = s->sessions[sl.session - 1];
        sl.info.done_s = now_s(s);
        if (q != hipSuccess) {
            (void)hipGetLastError();
            sl.info.status = CU_ERR_DEVICE;
            s->last_failure = std::string("frame: ") + hipGetErrorString(q);
            ss.has_prev = false;
        } else {
            sl.info.checksum = *sl.h_sum;
            if (sl.info.path == 2) {
                cu_deep_report rep{};
                if (cu_deep_stats(ss.deep, &rep) != CU_OK) {
                    sl.info.status = CU_ERR_DEVICE;
                    s->last_failure = cu_last_error();
                } else {
                    sl.info.iterations = rep.iterations;
                    s->deep_iterations += rep.iterations;
                }
            }
        }
        out[n++] = sl.info;
        sl.busy = false;
        ss.busy = false;
        ss.frames++;
        s->frames_done++;
        if (ss.state == 2 && ss.queue.empty()) release(s, ss);
    }
    return n;
}

void dispatch_ready(cu_serve_s *s) {
    for (Slot &sl : s->slots) {
        if (sl.busy) continue;
        Session *best = nullptr;
        for (Session &ss : s->sessions)
            if ((ss.state == 1 || ss.state == 2) && !ss.busy && !ss.queue.empty() &&
                (!best || ss.queue.front().submit_s < best->queue.front().submit_s))
                best = &ss;
        if (!best) return;
        const Request rq = best->queue.front();
        best->queue.pop_front();
        try {
            dispatch(s, *best, sl, rq);
        } catch (const CuError &e) {
            // the session's buffers may hold part of this frame: its next frame starts afresh
            best->has_prev = false;
            best->busy = false;
            sl.busy = false;
            fail_request(s, *best, rq, e.status, e.message);
        }
    }
}

bool any_busy(const cu_serve_s *s) {
    for (const Slot &sl : s->slots)
        if (sl.busy) return true;
    return false;
}

void free_server(cu_serve_s *s) {
    for (Session &ss : s->sessions)
        if (ss.state != 3) release(s, ss);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.d_sum) (void)hipFree(sl.d_sum);
        if (sl.h_sum) (void)hipHostFree(sl.h_sum);
    }
    if (s->engine) (void)cu_engine_destroy(s->engine);
    if (s->module) (void)cu_module_unload(s->module);
    delete s;
}

cu_function module_fn(cu_module m, const char *name) {
    cu_function f = nullptr;
    if (cu_module_get_function(m, name, &f) != CU_OK) fail(CU_ERR_INITIALIZATION, cu_last_error());
    return f;
}
cu_function module_fn_if(cu_module m, const char *name) {
    cu_function f = nullptr;
    return cu_module_get_function(m, name, &f) == CU_OK ? f : nullptr;
}

}  // namespace
}  // namespace cu

using namespace cu;

extern "C" {

int cu_serve_create(const char *kernels_dir, const cu_serve_config *cfg, cu_serve *out) {
    cu_serve s = nullptr;
    int rc = guarded([&] {
        if (!kernels_dir || !cfg || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (cfg->slots == 0 || cfg->slots > kMaxSlots)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "slots must be in [1, 16]");
        if (cfg->deep_method != CU_DEEP_PERTURB_F32 && cfg->deep_method != CU_DEEP_PERTURB_F64)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "deep_method must be CU_DEEP_PERTURB_F32 or _F64");
        if (!(cfg->max_super_sampling >= 1.0f && cfg->max_super_sampling <= 64.0f))
            fail(CU_ERR_ILLEGAL_ARGUMENT, "max_super_sampling must be in [1, 64]");
        if (!(cfg->deep_width >= 0.0)) fail(CU_ERR_ILLEGAL_ARGUMENT, "deep_width must be >= 0");
        if (cfg->shallow_policy != CU_SERVE_SHALLOW_FP64 &&
            cfg->shallow_policy != CU_SERVE_SHALLOW_REFERENCE_SWITCH)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown shallow_policy");
        s = new cu_serve_s();
        s->cfg = *cfg;
        s->t0 = std::chrono::steady_clock::now();
        ensure_device();
        if (cu_engine_create(kernels_dir, &s->engine) != CU_OK) fail(CU_ERR_INITIALIZATION, cu_last_error());
        const std::string path = std::string(kernels_dir) + "/mandelbrot.hsaco";
        if (cu_module_load(path.c_str(), &s->module) != CU_OK) fail(CU_ERR_INITIALIZATION, cu_last_error());
        s->main_f = module_fn(s->module, "fractalRenderMainFloat");
        s->main_d = module_fn(s->module, "fractalRenderMainDouble");
        s->adv_f = module_fn(s->module, "fractalRenderAdvancedFloat");
        s->adv_d = module_fn(s->module, "fractalRenderAdvancedDouble");
        s->main_fs = module_fn_if(s->module, "fractalRenderMainFloatStrip");
        s->main_ds = module_fn_if(s->module, "fractalRenderMainDoubleStrip");
        s->adv_fs = module_fn_if(s->module, "fractalRenderAdvancedFloatStrip");
        s->adv_ds = module_fn_if(s->module, "fractalRenderAdvancedDoubleStrip");
        s->compose = module_fn(s->module, "compose");
        s->checksum = function(s->engine, "checksumU32");
        size_t bytes = 0;
        if (cu_module_get_global(s->module, "cu_iteration_slots", &s->d_module_iters, &bytes) != CU_OK)
            fail(CU_ERR_INITIALIZATION, cu_last_error());
        check(hipMemset(s->d_module_iters, 0, bytes), "hipMemset");
        s->slots.resize(cfg->slots);
        for (Slot &sl : s->slots) {
            check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
            sl.wrap.stream = sl.stream;
            sl.d_sum = device_alloc(8);
            check(hipHostMalloc((void **)&sl.h_sum, 8, hipHostMallocDefault), "hipHostMalloc");
        }
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *out = s;
    });
    if (rc != CU_OK && s) free_server(s);
    return rc;
}

int cu_serve_destroy(cu_serve s) {
    return guarded([&] {
        if (!s) return;
        ensure_device();
        free_server(s);
    });
}

int cu_serve_open(cu_serve s, uint32_t width, uint32_t height, const uint32_t *palette,
                  uint32_t palette_length, uint32_t deep_max_iter, uint32_t *session) {
    return guarded([&] {
        if (!s || !palette || !session) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        if (width < 2 || height < 2 || width > 16384 || height > 16384)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "session size must be in [2, 16384]^2");
        if (palette_length == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "empty palette");
        std::lock_guard<std::mutex> lock(s->mutex);
        Session ss;
        ss.id = (uint32_t)s->sessions.size() + 1;
        ss.width = width;
        ss.height = height;
        ss.deep_max_iter = deep_max_iter;
        ss.host_palette.assign(palette, palette + palette_length);
        ss.bytes = session_bytes(width, height, palette_length, deep_max_iter);
        if (ss.bytes > s->cfg.hbm_budget)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "session needs more than the whole HBM budget");
        ss.open_s = now_s(s);
        s->sessions.push_back(std::move(ss));
        *session = s->sessions.back().id;
        ensure_device();
        admit(s);
    });
}

int cu_serve_submit(cu_serve s, uint32_t session, const cu_frame_request *req) {
    return guarded([&] {
        if (!s || !req) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (session == 0 || session > s->sessions.size()) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown session");
        Session &ss = s->sessions[session - 1];
        if (ss.state >= 2) fail(CU_ERR_ILLEGAL_STATE, "session is closed");
        if (!(req->width > 0) || !std::isfinite(req->width) || !std::isfinite(req->centre_x) ||
            !std::isfinite(req->centre_y) || req->max_iter == 0)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "bad frame request");
        if (req->width < s->cfg.deep_width) {
            if (ss.deep_max_iter == 0)
                fail(CU_ERR_ILLEGAL_ARGUMENT, "deep frame on a session opened without deep frames");
            if (req->max_iter > ss.deep_max_iter)
                fail(CU_ERR_ILLEGAL_ARGUMENT,
                     "deep frame with max_iter above the session's deep_max_iter");
            // the checks cu_deep_retarget would run at dispatch (frame within |c| <= 2,
            // positive pixel size), so a bad request is refused here and not later
            validate_deep(deep_desc(ss, *req), (int)s->cfg.deep_method);
        }
        ss.queue.push_back(Request{*req, now_s(s)});
        ensure_device();
        dispatch_ready(s);
    });
}

int cu_serve_close(cu_serve s, uint32_t session) {
    return guarded([&] {
        if (!s) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (session == 0 || session > s->sessions.size()) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown session");
        Session &ss = s->sessions[session - 1];
        if (ss.state == 3 || ss.state == 2) return;
        ensure_device();
        if (ss.state == 0) {
            ss.state = 3;  // never admitted: nothing to free
            ss.queue.clear();
            return;
        }
        ss.state = 2;
        if (!ss.busy && ss.queue.empty()) release(s, ss);
        admit(s);
    });
}

int cu_serve_poll(cu_serve s, cu_frame_done *out, uint32_t capacity, uint32_t *count,
                  int wait_ms) {
    return guarded([&] {
        if (!s || !count || (capacity && !out)) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        ensure_device();
        const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(wait_ms);
        uint32_t n = 0;
        std::unique_lock<std::mutex> lock(s->mutex);
        for (;;) {
            n = retire(s, out, capacity, n);
            admit(s);
            dispatch_ready(s);
            if (n > 0 || wait_ms <= 0 || !any_busy(s) || n >= capacity) break;
            if (std::chrono::steady_clock::now() >= deadline) break;
            lock.unlock();  // let other threads submit or close between the rounds
            std::this_thread::yield();
            lock.lock();
        }
        *count = n;
    });
}

int cu_serve_last_failure(cu_serve s, char *buf, size_t buf_len) {
    return guarded([&] {
        if (!s || !buf || buf_len == 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        const size_t n = std::min(buf_len - 1, s->last_failure.size());
        std::memcpy(buf, s->last_failure.data(), n);
        buf[n] = 0;
    });
}

int cu_serve_read_rgba(cu_serve s, uint32_t session, uint32_t *dst, size_t count) {
    return guarded([&] {
        if (!s || !dst) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (session == 0 || session > s->sessions.size()) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown session");
        Session &ss = s->sessions[session - 1];
        if (ss.state != 1 && ss.state != 2) fail(CU_ERR_ILLEGAL_STATE, "session holds no image");
        if (ss.busy) fail(CU_ERR_ILLEGAL_STATE, "session has a frame in flight");
        const size_t n = (size_t)ss.width * ss.height;
        if (count < n) fail(CU_ERR_ILLEGAL_ARGUMENT, "destination too small");
        ensure_device();
        check(hipMemcpy(dst, ss.rgba, n * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
}

int cu_serve_stats(cu_serve s, cu_serve_report *out) {
    return guarded([&] {
        if (!s || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        ensure_device();
        for (Slot &sl : s->slots)
            if (sl.busy) check(hipStreamSynchronize(sl.stream), "hipStreamSynchronize");
        unsigned long long module_iters = 0;
        if (cu_iters_total((const unsigned long long *)s->d_module_iters, &module_iters) != CU_OK)
            fail(CU_ERR_DEVICE, cu_last_error());
        cu_serve_report r{};
        r.iterations = module_iters + s->deep_iterations;
        r.frames = s->frames_done;
        r.bytes_in_use = s->used;
        r.peak_bytes = s->peak;
        uint32_t admitted = 0, waiting = 0;
        for (const Session &ss : s->sessions) {
            if (ss.admit_s >= 0) ++admitted;
            if (ss.state == 0) ++waiting;
        }
        r.sessions_admitted = admitted;
        r.sessions_waiting = waiting;
        *out = r;
    });
}

int cu_serve_session_times(cu_serve s, uint32_t session, double *open_s, double *admit_s) {
    return guarded([&] {
        if (!s || !open_s || !admit_s) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL argument");
        std::lock_guard<std::mutex> lock(s->mutex);
        if (session == 0 || session > s->sessions.size()) fail(CU_ERR_ILLEGAL_ARGUMENT, "unknown session");
        *open_s = s->sessions[session - 1].open_s;
        *admit_s = s->sessions[session - 1].admit_s;
    });
}

}  // extern "C"
