// renderer.cpp — layer 2 of the C ABI: the reference's plugin surface
// (FractalRendererProvider / FractalRenderer / RenderingModel) implemented in
// C++ over the driver layer, one class per reference class:
//
//   Provider        cudarenderer/CudaFractalRendererProvider.java:14-91
//   Renderer        cudarenderer/CudaFractalRenderer.java:32-430
//   Module(+Julia)  cudarenderer/FractalRenderingModule.java:31-280, modules/Module*.java
//   RenderKernel    cudarenderer/RenderingKernel.java, KernelMain*.java, KernelAdvanced*.java
//   DoubleBuffer    cudarenderer/DeviceMemoryDoubleBuffer2D.java:17-152
//
// The GL textures of GLParams become a linear RGBA8 device image and a device
// palette; everything else (state machine, precision switch, reuse bookkeeping,
// launch geometry, error classes and messages) follows the reference.  Calls
// are synchronous, as the reference's are; one mutex serialises the layer.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "handles.hpp"

namespace cu {
namespace {

std::mutex g_mutex;
bool g_debug_mode = false;  // -Ddebug=true (util/JavaHelpers.java:12,25-27)

constexpr int kBlockDim = 32;          // CudaFractalRenderer.java:36-37
constexpr int kMaxGridDim = 65536 - 1;  // CudaFractalRenderer.java:34
constexpr int kDefaultUnderSamplingLevel = 4;
constexpr size_t kPixelInfoBytes = 16;  // SIZEOF_PIXEL_INFO_T (4) x 4 bytes

// Math.ulp(double) / Math.ulp(float)
double java_ulp(double v) {
    v = std::fabs(v);
    if (!std::isfinite(v)) return std::isnan(v) ? v : INFINITY;
    if (v == 0) return std::nextafter(0.0, 1.0);
    return std::nextafter(v, INFINITY) - v;
}
double java_ulp_f(float v) {
    v = std::fabs(v);
    if (!std::isfinite(v)) return std::isnan(v) ? v : INFINITY;
    if (v == 0) return (double)std::nextafter(0.0f, 1.0f);
    // ulp of the binade: distance to the next float above
    return (double)(std::nextafter(v, INFINITY) - v);
}

struct Segment {
    double lbx = 0, lby = 0, rtx = 0, rty = 0;
};

// RenderingKernel.isSegmentBoundsAt{Float,Double}Limit (cudarenderer/RenderingKernel.java:124-138)
bool at_float_limit(const Segment &s, int w, int h) {
    double dx = java_ulp_f((float)s.lbx), dy = java_ulp_f((float)s.lby);
    double pw = std::fabs(s.rtx - s.lbx) / (double)w;
    double ph = std::fabs(s.rty - s.lby) / (double)h;
    return (pw < dx) || (ph < dy);
}
bool at_double_limit(const Segment &s, int w, int h) {
    double dx = java_ulp(s.lbx), dy = java_ulp(s.lby);
    double pw = std::fabs(s.rtx - s.lbx) / (double)w;
    double ph = std::fabs(s.rty - s.lby) / (double)h;
    return (pw < dx) || (ph < dy);
}

// CudaFractalRenderer.updateFloatPrecision (:409-419)
int precision_for(const Segment &s, int w, int h) {
    int p = CU_PRECISION_SINGLE;
    if (at_float_limit(s, w, h)) p = CU_PRECISION_DOUBLE;
    if (at_double_limit(s, w, h)) p = CU_PRECISION_TOO_BIG;
    return p;
}

void check_finite(double v, const char *name) {  // RenderingKernel.checkArgument (:143-146)
    if (!std::isfinite(v)) {
        std::ostringstream o;
        o << "Argument " << name << " must be a finite float but is " << v << ".";
        fail(CU_ERR_ILLEGAL_ARGUMENT, o.str());
    }
}

Segment model_segment(const cu_rendering_model &m) {
    return Segment{m.left_bottom_x, m.left_bottom_y, m.right_top_x, m.right_top_y};
}

void log_error(cu_rendering_model *m, const std::string &msg) {  // PublicErrorLogger.logError
    if (!m) return;
    m->error_count++;
    std::snprintf(m->last_error, sizeof m->last_error, "%s", msg.c_str());
}

// Model.setPlaneSegmentFromCenter (rendering/Model.java:247-256)
void plane_from_center(double cx, double cy, double zoom, int w, int h, double out[4]) {
    double rel_h = 1;
    double rel_w = rel_h / (double)h * w;
    out[0] = cx - rel_w * zoom / 2;
    out[1] = cy - rel_h * zoom / 2;
    out[2] = cx + rel_w * zoom / 2;
    out[3] = cy + rel_h * zoom / 2;
}

// ---------------------------------------------------------------------------
// FractalRenderingModule
// ---------------------------------------------------------------------------
const char *const kKernelNames[] = {"fractalRenderMainFloat",     "fractalRenderMainDouble",
                                    "fractalRenderAdvancedFloat", "fractalRenderAdvancedDouble",
                                    "fractalRenderUnderSampled",  "compose",
                                    "debug"};
// The same four render kernels for 32 x 4-thread blocks (this build's modules
// export them; a module without them renders through the kernels above).
const char *const kStripKernelNames[] = {"fractalRenderMainFloatStrip",
                                         "fractalRenderMainDoubleStrip",
                                         "fractalRenderAdvancedFloatStrip",
                                         "fractalRenderAdvancedDoubleStrip"};
constexpr int kStripRows = 4;  // rows of a strip block

class Module {
  public:
    Module(std::string path, std::string name) : path_(std::move(path)), name_(std::move(name)) {}
    virtual ~Module() {
        if (module_) cu_module_unload(module_);
    }

    // FractalRenderingModule.initialize (:73-100)
    virtual void initialize() {
        if (initialized_) fail(CU_ERR_ILLEGAL_STATE, "Module already initialized.");
        cu_module m = nullptr;
        int rc = cu_module_load(path_.c_str(), &m);
        if (rc != CU_OK) fail(rc, cu_last_error());
        module_ = m;
        kernels_.clear();
        for (const char *k : kKernelNames) {
            cu_function f = nullptr;
            rc = cu_module_get_function(module_, k, &f);
            if (rc != CU_OK) {
                std::string msg = cu_last_error();
                cu_module_unload(module_);
                module_ = nullptr;
                fail(rc, msg);
            }
            kernels_[k] = f;
        }
        for (const char *k : kStripKernelNames) {
            cu_function f = nullptr;
            if (cu_module_get_function(module_, k, &f) == CU_OK) kernels_[k] = f;
        }
        initialized_ = true;
    }

    bool initialized() const { return initialized_; }
    const std::string &name() const { return name_; }

    cu_function kernel(const std::string &k) const {  // getKernel (:127-133)
        if (!initialized_)
            fail(CU_ERR_ILLEGAL_STATE, "Module has not been initialized or has been closed.");
        auto it = kernels_.find(k);
        if (it == kernels_.end()) fail(CU_ERR_ILLEGAL_ARGUMENT, "No such kernel available: " + k);
        return it->second;
    }
    // An optional kernel of the module, or null.
    cu_function kernel_if(const std::string &k) const {
        auto it = kernels_.find(k);
        return it == kernels_.end() ? nullptr : it->second;
    }

    void close() {  // (:140-147)
        if (module_) {
            cu_module_unload(module_);
            module_ = nullptr;
        }
        kernels_.clear();
        initialized_ = false;
    }

    // Sum of the module's cu_iteration_slots (iterations executed since load).
    unsigned long long iterations() const {
        if (!initialized_) fail(CU_ERR_ILLEGAL_STATE, "Module has not been initialized.");
        void *d = nullptr;
        size_t bytes = 0;
        int rc = cu_module_get_global(module_, "cu_iteration_slots", &d, &bytes);
        if (rc != CU_OK) fail(rc, cu_last_error());
        std::vector<unsigned long long> slots(bytes / sizeof(unsigned long long));
        rc = cu_memcpy_dtoh(slots.data(), d, slots.size() * sizeof(unsigned long long));
        if (rc != CU_OK) fail(rc, cu_last_error());
        unsigned long long total = 0;
        for (unsigned long long v : slots) total += v;
        return total;
    }

    void write_constant(const std::string &sym, const void *v, size_t bytes) {
        if (!initialized_) fail(CU_ERR_ILLEGAL_STATE, "Module has not been initialized.");
        int rc = cu_module_write_global(module_, sym.c_str(), v, bytes);
        if (rc != CU_OK) fail(rc, cu_last_error());
    }

    virtual void set_custom_params(const std::string &) {}

    // FractalRenderingModule.supplyDefaultValues (:277-279)
    virtual void supply_default_values(cu_rendering_model &m, int, int) {
        m.fractal_custom_params[0] = 0;
    }

  protected:
    // parseParamsAsDoubles (:233-241): tokens split on ',' or ';', Double.parseDouble each.
    static std::vector<double> parse_doubles(const std::string &text) {
        std::vector<double> out;
        std::string tok;
        auto flush = [&](bool last) {
            if (last && tok.empty() && !out.empty()) return;  // String.split drops trailing empties
            std::string t = tok;
            t.erase(0, t.find_first_not_of(" \t\n\r"));
            t.erase(t.find_last_not_of(" \t\n\r") + 1);
            char *end = nullptr;
            double v = std::strtod(t.c_str(), &end);
            if (t.empty() || !end || *end != 0)
                fail(CU_ERR_ILLEGAL_ARGUMENT, "NumberFormatException: For input string: \"" + tok + "\"");
            out.push_back(v);
            tok.clear();
        };
        for (char c : text) {
            if (c == ',' || c == ';') flush(false);
            else tok.push_back(c);
        }
        flush(true);
        return out;
    }

  private:
    std::string path_, name_;
    cu_module module_ = nullptr;
    std::map<std::string, cu_function> kernels_;
    bool initialized_ = false;
};

// modules/ModuleMandelbrot.java:9-25
class ModuleMandelbrot : public Module {
  public:
    explicit ModuleMandelbrot(const std::string &dir) : Module(dir + "/mandelbrot.hsaco", "mandelbrot") {}
    void supply_default_values(cu_rendering_model &m, int w, int h) override {
        Module::supply_default_values(m, w, h);
        double seg[4];
        plane_from_center(-0.5, 0, 2, w, h, seg);
        m.left_bottom_x = seg[0];
        m.left_bottom_y = seg[1];
        m.right_top_x = seg[2];
        m.right_top_y = seg[3];
        m.max_iterations = 1600;
        m.max_super_sampling = 5;
    }
};

// modules/ModuleJulia.java:9-49
class ModuleJulia : public Module {
  public:
    explicit ModuleJulia(const std::string &dir) : Module(dir + "/julia.hsaco", "julia") {}
    void initialize() override {
        Module::initialize();
        set_c(0, 0);
    }
    void set_c(double re, double im) {
        const double c[2] = {re, im};
        write_constant("julia_c", c, sizeof c);
    }
    void set_custom_params(const std::string &text) override {
        std::vector<double> v = parse_doubles(text);
        if (v.size() < 2)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "ArrayIndexOutOfBoundsException: julia parameters need \"re;im\"");
        set_c(v[0], v[1]);
    }
    void supply_default_values(cu_rendering_model &m, int w, int h) override {
        Module::supply_default_values(m, w, h);
        m.max_iterations = 900;
        std::snprintf(m.fractal_custom_params, sizeof m.fractal_custom_params, "%s", "-0.4;0.6");
    }
};

// ---------------------------------------------------------------------------
// DeviceMemoryDoubleBuffer2D
// ---------------------------------------------------------------------------
class DoubleBuffer {
  public:
    ~DoubleBuffer() { free_all(); }
    void reallocate(int w, int h) {
        free_all();
        alloc(w, h, &a1_, &p1_);
        primary_dirty_ = true;
        alloc(w, h, &a2_, &p2_);
    }
    void free_all() {
        if (a1_) cu_mem_free(a1_);
        if (a2_) cu_mem_free(a2_);
        a1_ = a2_ = nullptr;
    }
    void *primary() const { return a1_; }
    void *secondary() const { return a2_; }
    long primary_pitch() const { return (long)p1_; }
    long secondary_pitch() const { return (long)p2_; }
    void switch_buffers() {
        switched_ = !switched_;
        std::swap(a1_, a2_);
        std::swap(p1_, p2_);
    }
    void reset_order() {
        if (switched_) switch_buffers();
    }
    bool primary_dirty() const { return primary_dirty_; }
    void set_primary_dirty(bool v) { primary_dirty_ = v; }

  private:
    static void alloc(int w, int h, void **p, size_t *pitch) {
        *p = nullptr;
        *pitch = 0;
        if ((long)w * h == 0) return;
        int rc = cu_mem_alloc_pitch(p, pitch, (size_t)w * kPixelInfoBytes, (size_t)h, kPixelInfoBytes);
        if (rc != CU_OK) fail(rc, cu_last_error());
    }
    void *a1_ = nullptr, *a2_ = nullptr;
    size_t p1_ = 0, p2_ = 0;
    bool switched_ = false;
    bool primary_dirty_ = true;
};

// ---------------------------------------------------------------------------
// rendering kernels (parameter marshalling of KernelMain / KernelAdvanced)
// ---------------------------------------------------------------------------
template <class T>
struct PointT {
    T x, y;
};
template <class T>
struct RectT {
    T lbx, lby, rtx, rty;
};

struct RenderKernel {
    cu_function fn = nullptr;
    cu_function strip = nullptr;  // the same kernel for 32 x 4-thread blocks, if the module has it
    bool is_float = false;
    bool advanced = false;
    int width = 0, height = 0;
    int max_iterations = 1;
    Segment seg;
    void *out = nullptr;
    long out_pitch = 0;
    float max_ss = 1;
    uint32_t flags = CU_FLAG_ADAPTIVE_SS;  // KernelMain: setUseAdaptiveSuperSampling(true)
    Segment origin;
    void *in = nullptr;
    long in_pitch = 0;
    PointT<uint32_t> focus{0, 0};

    void set_bit(uint32_t bit, bool v) { flags = v ? (flags | bit) : (flags & ~bit); }

    void set_max_iterations(int n) {  // RenderingKernel.setMaxIterations (:87-91)
        if (n < 1)
            fail(CU_ERR_ILLEGAL_ARGUMENT,
                 "maxIterations must be a positive number, but is : " + std::to_string(n));
        max_iterations = n;
    }
    void set_segment(const Segment &s) {  // setPlaneSegment (:128-141)
        check_finite(s.lbx, "segment left_bottom_x");
        check_finite(s.lby, "segment left_bottom_y");
        check_finite(s.rtx, "segment right_top_x");
        check_finite(s.rty, "segment right_top_y");
        seg = s;
    }
    void set_origin(const Segment &s) {  // KernelAdvanced.setOriginSegment (:49-55)
        check_finite(s.lbx, "origin segment left_bottom_x");
        check_finite(s.lby, "origin segment left_bottom_y");
        check_finite(s.rtx, "origin segment right_top_x");
        check_finite(s.rty, "origin segment right_top_y");
        origin = s;
    }
    // setParamsFromModel of RenderingKernel + KernelMain (+ KernelAdvanced)
    void set_params_from_model(const cu_rendering_model &m) {
        set_max_iterations(m.max_iterations);
        set_segment(model_segment(m));
        set_bit(CU_FLAG_VISUALISE_SAMPLE_COUNT, m.visualise_sample_count != 0);
        set_bit(CU_FLAG_ADAPTIVE_SS, m.use_adaptive_super_sampling != 0);
        max_ss = m.max_super_sampling;
        if (advanced) {
            focus = PointT<uint32_t>{(uint32_t)m.mouse_focus_x, (uint32_t)m.mouse_focus_y};
            set_bit(CU_FLAG_ZOOMING, m.zooming != 0);
            set_bit(CU_FLAG_ZOOMING_IN, m.zooming_in != 0);
            set_bit(CU_FLAG_FOVEATION, m.use_foveated_rendering != 0);
            set_bit(CU_FLAG_SAMPLE_REUSE, m.use_sample_reuse != 0);
        }
    }

    // Launch and synchronise: the strip kernel over ceil(W/32) x ceil(H/4) blocks
    // of 32 x 4 threads when the module has it (the same pixels, vote groups and
    // results), else the reference's geometry.
    void launch() {
        PointT<uint32_t> size{(uint32_t)width, (uint32_t)height};
        RectT<double> rd{seg.lbx, seg.lby, seg.rtx, seg.rty};
        RectT<float> rf{(float)seg.lbx, (float)seg.lby, (float)seg.rtx, (float)seg.rty};
        RectT<double> od{origin.lbx, origin.lby, origin.rtx, origin.rty};
        RectT<float> of{(float)origin.lbx, (float)origin.lby, (float)origin.rtx, (float)origin.rty};
        uint32_t maxit = (uint32_t)max_iterations;
        void *image = is_float ? (void *)&rf : (void *)&rd;
        void *reused = is_float ? (void *)&of : (void *)&od;
        void *params[] = {&out, &out_pitch, &size, image, &maxit, &max_ss, &flags,
                          reused, &in, &in_pitch, &focus};
        unsigned gx = (unsigned)((width + kBlockDim - 1) / kBlockDim);
        unsigned gy = (unsigned)((height + kBlockDim - 1) / kBlockDim);
        const int64_t strip_rows = ((int64_t)height + kStripRows - 1) / kStripRows;
        int rc;
        if (strip && strip_rows <= kMaxGridDim)
            rc = cu_launch_kernel(strip, gx, (unsigned)strip_rows, 1, kBlockDim, kStripRows, 1, 0,
                                  params);
        else
            rc = cu_launch_kernel(fn, gx, gy, 1, kBlockDim, kBlockDim, 1, 0, params);
        if (rc == CU_OK) rc = cu_ctx_synchronize();
        if (rc != CU_OK) fail(rc, cu_last_error());
    }
};

// grid checks of launchRenderingKernel / launchDrawingKernel (:242-253, :305-331);
// returns false for an empty image (nothing is launched).
bool grid_for(int w, int h, unsigned *gx, unsigned *gy) {
    // in 64 bits: a width near INT_MAX must not overflow on the way to the check
    const int64_t x = ((int64_t)w + kBlockDim - 1) / kBlockDim;
    const int64_t y = ((int64_t)h + kBlockDim - 1) / kBlockDim;
    if (x <= 0 || y <= 0) return false;
    if (x > kMaxGridDim)
        fail(CU_ERR_RENDERER, "Unsupported input parameter: width must be smaller than " +
                                  std::to_string(kMaxGridDim * kBlockDim));
    if (y > kMaxGridDim)
        fail(CU_ERR_RENDERER, "Unsupported input parameter: height must be smaller than " +
                                  std::to_string(kMaxGridDim * kBlockDim));
    *gx = (unsigned)x;
    *gy = (unsigned)y;
    return true;
}

// ---------------------------------------------------------------------------
// CudaFractalRenderer
// ---------------------------------------------------------------------------
class Renderer {
  public:
    explicit Renderer(Module *module) : module_(module) {
        main_f_.fn = module->kernel("fractalRenderMainFloat");
        main_f_.strip = module->kernel_if("fractalRenderMainFloatStrip");
        main_f_.is_float = true;
        main_d_.fn = module->kernel("fractalRenderMainDouble");
        main_d_.strip = module->kernel_if("fractalRenderMainDoubleStrip");
        adv_f_.fn = module->kernel("fractalRenderAdvancedFloat");
        adv_f_.strip = module->kernel_if("fractalRenderAdvancedFloatStrip");
        adv_f_.is_float = true;
        adv_f_.advanced = true;
        adv_d_.fn = module->kernel("fractalRenderAdvancedDouble");
        adv_d_.strip = module->kernel_if("fractalRenderAdvancedDoubleStrip");
        adv_d_.advanced = true;
        compose_ = module->kernel("compose");
        debug_ = module->kernel("debug");
        (void)kDefaultUnderSamplingLevel;  // the under-sampled kernel is never launched
        memory_.reallocate(width(), height());
    }
    ~Renderer() { release_output(); }

    int width() const { return main_f_.width; }
    int height() const { return main_f_.height; }
    int state() const { return state_; }
    const std::string &fractal_name() const { return module_->name(); }
    Module *module() const { return module_; }

    // initializeRendering (:84-101)
    void initialize_rendering(int w, int h, const uint32_t *palette, int palette_length) {
        if (state_ == CU_STATE_READY_TO_RENDER) fail(CU_ERR_ILLEGAL_STATE, "Already initialized.");
        if (w < 0 || h < 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "negative output size");
        if (!palette || palette_length <= 0)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "palette must hold at least one colour");
        for (RenderKernel *k : {&main_f_, &main_d_, &adv_f_, &adv_d_}) {
            k->width = w;
            k->height = h;
        }
        memory_.reallocate(w, h);
        release_output();
        int rc = cu_mem_alloc(&output_, (size_t)w * h * 4);
        if (rc == CU_OK) rc = cu_mem_alloc(&palette_, (size_t)palette_length * 4);
        if (rc == CU_OK) rc = cu_memcpy_htod(palette_, palette, (size_t)palette_length * 4);
        if (rc != CU_OK) fail(rc, cu_last_error());
        palette_length_ = palette_length;
        state_ = CU_STATE_READY_TO_RENDER;
    }

    // freeRenderingResources (:103-112)
    void free_rendering_resources() {
        if (state_ == CU_STATE_NOT_INITIALIZED) fail(CU_ERR_ILLEGAL_STATE, "Already free.");
        memory_.free_all();
        release_output();
        state_ = CU_STATE_NOT_INITIALIZED;
    }

    // launchDebugKernel (:146-154)
    void launch_debug_kernel() {
        int rc = cu_launch_kernel(debug_, 1, 1, 1, kBlockDim, kBlockDim, 1, 0, nullptr);
        if (rc == CU_OK) rc = cu_ctx_synchronize();
        if (rc != CU_OK) fail(rc, cu_last_error());
    }

    // renderFast (:158-184)
    void render_fast(cu_rendering_model *model) {
        if (state_ != CU_STATE_READY_TO_RENDER)
            fail(CU_ERR_ILLEGAL_STATE, "Renderer has to be initialized first");
        set_module_constants(*model);
        if (model->sample_reuse_cache_dirty || memory_.primary_dirty() || !last_) {
            render_quality(model);
            return;
        }
        update_float_precision(model);
        RenderKernel &k = pick(model->floating_point_precision, adv_f_, adv_d_);
        k.set_origin(model_segment(*last_));
        k.set_params_from_model(*model);
        k.in = memory_.primary();
        k.in_pitch = memory_.primary_pitch();
        k.out = memory_.secondary();
        k.out_pitch = memory_.secondary_pitch();
        launch_rendering_kernel(k, model);
        memory_.switch_buffers();
        launch_drawing_kernel(model);
        last_ = *model;
    }

    // renderQuality (:186-206)
    void render_quality(cu_rendering_model *model) {
        if (state_ != CU_STATE_READY_TO_RENDER)
            fail(CU_ERR_ILLEGAL_STATE, "Renderer has to be initialized first");
        set_module_constants(*model);
        update_float_precision(model);
        RenderKernel &k = pick(model->floating_point_precision, main_f_, main_d_);
        memory_.reset_order();
        k.set_params_from_model(*model);
        k.out = memory_.primary();
        k.out_pitch = memory_.primary_pitch();
        launch_rendering_kernel(k, model);
        launch_drawing_kernel(model);
        last_ = *model;
        memory_.set_primary_dirty(false);
        model->sample_reuse_cache_dirty = 0;
    }

    // close (:380-386): frees resources and closes the module
    void close() {
        if (state_ == CU_STATE_READY_TO_RENDER) free_rendering_resources();
        memory_.free_all();
        module_->close();
    }

    void read_output(uint32_t *dst, size_t count) {
        if (state_ != CU_STATE_READY_TO_RENDER)
            fail(CU_ERR_ILLEGAL_STATE, "Renderer has to be initialized first");
        size_t n = (size_t)width() * height();
        if (count < n)
            fail(CU_ERR_ILLEGAL_ARGUMENT, "Output buffer must be at least width * height words long");
        int rc = cu_memcpy_dtoh(dst, output_, n * 4);
        if (rc != CU_OK) fail(rc, cu_last_error());
    }

    void read_pixel_info(bool primary, void *dst, size_t bytes) {
        if (state_ != CU_STATE_READY_TO_RENDER)
            fail(CU_ERR_ILLEGAL_STATE, "Renderer has to be initialized first");
        size_t row = (size_t)width() * kPixelInfoBytes;
        if (bytes < row * height()) fail(CU_ERR_ILLEGAL_ARGUMENT, "destination too small");
        const void *src = primary ? memory_.primary() : memory_.secondary();
        long pitch = primary ? memory_.primary_pitch() : memory_.secondary_pitch();
        int rc = cu_memcpy_2d_dtoh(dst, row, src, (size_t)pitch, row, (size_t)height());
        if (rc != CU_OK) fail(rc, cu_last_error());
    }

  private:
    RenderKernel &pick(int precision, RenderKernel &f, RenderKernel &d) {
        switch (precision) {  // getFloatOrDoubleKernel (:208-225)
            case CU_PRECISION_DOUBLE: return d;
            case CU_PRECISION_SINGLE: return f;
            case CU_PRECISION_TOO_BIG: return d;
            default: fail(CU_ERR_ILLEGAL_STATE, "Precision not supported: " + std::to_string(precision));
        }
    }

    // setModuleConstants (:227-233)
    void set_module_constants(const cu_rendering_model &m) {
        if (!last_ || (m.visualise_sample_count != 0) != (last_->visualise_sample_count != 0)) {
            const uint32_t v = m.visualise_sample_count ? 1u : 0u;  // int[]{1|0}, 1-byte symbol
            module_->write_constant("VISUALIZE_SAMPLE_COUNT", &v, 1);
        }
    }

    // updateFloatPrecision (:409-419)
    void update_float_precision(cu_rendering_model *m) {
        main_d_.set_params_from_model(*m);
        m->floating_point_precision = precision_for(main_d_.seg, main_d_.width, main_d_.height);
    }

    // launchRenderingKernel (:235-269)
    void launch_rendering_kernel(RenderKernel &k, cu_rendering_model *logger) {
        unsigned gx, gy;
        if (!grid_for(k.width, k.height, &gx, &gy)) return;
        try {
            k.launch();
        } catch (const CuError &e) {
            log_error(logger, "Error just after launching a kernel:" + e.message);
            if (g_debug_mode) throw;
        }
    }

    // launchDrawingKernel (:275-364)
    void launch_drawing_kernel(cu_rendering_model *logger) {
        try {
            unsigned gx, gy;
            if (!grid_for(width(), height(), &gx, &gy)) return;
            void *main = memory_.primary();
            long main_pitch = memory_.primary_pitch();
            void *bcg = memory_.secondary();
            long bcg_pitch = memory_.secondary_pitch();
            uint32_t w = (uint32_t)width(), h = (uint32_t)height();
            uint32_t plen = (uint32_t)palette_length_;
            float max_ss = logger ? logger->max_super_sampling : 1.f;
            void *params[] = {&main, &main_pitch, &bcg, &bcg_pitch, &output_, &w, &h,
                              &palette_, &plen, &max_ss};
            int rc = cu_launch_kernel(compose_, gx, gy, 1, kBlockDim, kBlockDim, 1, 0, params);
            if (rc == CU_OK) rc = cu_ctx_synchronize();
            if (rc != CU_OK) {
                log_error(logger, std::string("Error just after launching a kernel:") + cu_last_error());
                if (g_debug_mode) fail(rc, cu_last_error());
            }
        } catch (const CuError &e) {
            log_error(logger, "Error during kernel launch preparation:" + e.message);
            if (g_debug_mode) throw;
        }
    }

    void release_output() {
        if (output_) cu_mem_free(output_);
        if (palette_) cu_mem_free(palette_);
        output_ = palette_ = nullptr;
    }

    Module *module_;
    RenderKernel main_f_, main_d_, adv_f_, adv_d_;
    cu_function compose_ = nullptr, debug_ = nullptr;
    DoubleBuffer memory_;
    void *output_ = nullptr, *palette_ = nullptr;
    int palette_length_ = 0;
    int state_ = CU_STATE_NOT_INITIALIZED;
    std::optional<cu_rendering_model> last_;
};

// ---------------------------------------------------------------------------
// CudaFractalRendererProvider
// ---------------------------------------------------------------------------
class Provider {
  public:
    explicit Provider(const std::string &dir) {
        add(std::make_unique<ModuleJulia>(dir));
        add(std::make_unique<ModuleMandelbrot>(dir));
    }
    ~Provider() {
        renderers_.clear();
        modules_.clear();
    }

    std::string available() const {
        std::string s;
        for (auto &kv : modules_) s += (s.empty() ? "" : ",") + kv.first;
        return s;
    }

    // getRenderer (:47-66)
    Renderer *get_renderer(const std::string &name, bool force) {
        if (active_) {
            if (active_->fractal_name() == name && !force) return active_;
            active_->close();
        }
        auto it = modules_.find(name);
        if (it == modules_.end()) fail(CU_ERR_ILLEGAL_ARGUMENT, "Unknown fractal: " + name);
        Module *m = it->second.get();
        if (!m->initialized()) m->initialize();
        renderers_.push_back(std::make_unique<Renderer>(m));
        active_ = renderers_.back().get();
        return active_;
    }

  private:
    void add(std::unique_ptr<Module> m) { modules_[m->name()] = std::move(m); }
    std::map<std::string, std::unique_ptr<Module>> modules_;
    std::vector<std::unique_ptr<Renderer>> renderers_;  // every renderer handed out stays valid
    Renderer *active_ = nullptr;
};

// handle tables
std::map<int32_t, std::unique_ptr<Provider>> g_providers;
std::map<int32_t, std::pair<Provider *, Renderer *>> g_renderers;
std::map<const Renderer *, int32_t> g_renderer_ids;
int32_t g_next_handle = 1;

Provider *provider(cu_provider h) {
    auto it = g_providers.find(h);
    if (it == g_providers.end()) fail(CU_ERR_ILLEGAL_STATE, "stale or unknown provider handle");
    return it->second.get();
}
Renderer *renderer(cu_renderer h) {
    auto it = g_renderers.find(h);
    if (it == g_renderers.end()) fail(CU_ERR_ILLEGAL_STATE, "stale or unknown renderer handle");
    return it->second.second;
}
cu_renderer handle_for(Provider *p, Renderer *r) {
    auto it = g_renderer_ids.find(r);
    if (it != g_renderer_ids.end()) return it->second;
    int32_t id = g_next_handle++;
    g_renderers[id] = {p, r};
    g_renderer_ids[r] = id;
    return id;
}

template <class F>
int locked(F &&body) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        body();
    });
}

}  // namespace
}  // namespace cu

using namespace cu;

extern "C" {

int cu_set_debug_mode(int enabled) {
    return locked([&] { g_debug_mode = enabled != 0; });
}

int cu_provider_create(const char *kernels_dir, cu_provider *out) {
    return locked([&] {
        if (!kernels_dir || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "kernels_dir/out is NULL");
        int32_t id = g_next_handle++;
        g_providers[id] = std::make_unique<Provider>(kernels_dir);
        *out = id;
    });
}

int cu_provider_destroy(cu_provider p) {
    return locked([&] {
        Provider *pp = provider(p);
        for (auto it = g_renderers.begin(); it != g_renderers.end();) {
            if (it->second.first == pp) {
                g_renderer_ids.erase(it->second.second);
                it = g_renderers.erase(it);
            } else {
                ++it;
            }
        }
        g_providers.erase(p);
    });
}

int cu_provider_available_fractals(cu_provider p, char *buf, size_t len) {
    return locked([&] {
        std::string s = provider(p)->available();
        if (!buf || len < s.size() + 1) fail(CU_ERR_ILLEGAL_ARGUMENT, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

int cu_provider_get_default_renderer(cu_provider p, cu_renderer *out) {
    return cu_provider_get_renderer(p, "mandelbrot", 0, out);
}

int cu_provider_get_renderer(cu_provider p, const char *name, int force, cu_renderer *out) {
    return locked([&] {
        if (!name || !out) fail(CU_ERR_ILLEGAL_ARGUMENT, "name/out is NULL");
        Provider *pp = provider(p);
        Renderer *r = pp->get_renderer(name, force != 0);
        *out = handle_for(pp, r);
    });
}

int cu_renderer_initialize_rendering(cu_renderer r, int w, int h, const uint32_t *palette,
                                     int palette_length) {
    return locked([&] { renderer(r)->initialize_rendering(w, h, palette, palette_length); });
}
int cu_renderer_free_rendering_resources(cu_renderer r) {
    return locked([&] { renderer(r)->free_rendering_resources(); });
}
int cu_renderer_get_state(cu_renderer r, int *state) {
    return locked([&] {
        if (!state) fail(CU_ERR_ILLEGAL_ARGUMENT, "state is NULL");
        *state = renderer(r)->state();
    });
}
int cu_renderer_get_width(cu_renderer r, int *w) {
    return locked([&] {
        if (!w) fail(CU_ERR_ILLEGAL_ARGUMENT, "width is NULL");
        *w = renderer(r)->width();
    });
}
int cu_renderer_get_height(cu_renderer r, int *h) {
    return locked([&] {
        if (!h) fail(CU_ERR_ILLEGAL_ARGUMENT, "height is NULL");
        *h = renderer(r)->height();
    });
}
int cu_renderer_launch_debug_kernel(cu_renderer r) {
    return locked([&] { renderer(r)->launch_debug_kernel(); });
}
int cu_renderer_render_fast(cu_renderer r, cu_rendering_model *m) {
    return locked([&] {
        if (!m) fail(CU_ERR_ILLEGAL_ARGUMENT, "model is NULL");
        renderer(r)->render_fast(m);
    });
}
int cu_renderer_render_quality(cu_renderer r, cu_rendering_model *m) {
    return locked([&] {
        if (!m) fail(CU_ERR_ILLEGAL_ARGUMENT, "model is NULL");
        renderer(r)->render_quality(m);
    });
}
int cu_renderer_close(cu_renderer r) {
    return locked([&] { renderer(r)->close(); });
}
int cu_renderer_set_fractal_custom_params(cu_renderer r, const char *text) {
    return locked([&] {
        if (!text) fail(CU_ERR_ILLEGAL_ARGUMENT, "text is NULL");
        renderer(r)->module()->set_custom_params(text);
    });
}
int cu_renderer_get_fractal_name(cu_renderer r, char *buf, size_t len) {
    return locked([&] {
        const std::string &s = renderer(r)->fractal_name();
        if (!buf || len < s.size() + 1) fail(CU_ERR_ILLEGAL_ARGUMENT, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}
int cu_renderer_supply_default_values(cu_renderer r, cu_rendering_model *m, int w, int h) {
    return locked([&] {
        if (!m) fail(CU_ERR_ILLEGAL_ARGUMENT, "model is NULL");
        if (w <= 0 || h <= 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "canvas size must be positive");
        renderer(r)->module()->supply_default_values(*m, w, h);
    });
}
int cu_renderer_iterations(cu_renderer r, unsigned long long *total) {
    return locked([&] {
        if (!total) fail(CU_ERR_ILLEGAL_ARGUMENT, "total is NULL");
        *total = renderer(r)->module()->iterations();
    });
}

int cu_renderer_read_output(cu_renderer r, uint32_t *rgba, size_t count) {
    return locked([&] {
        if (!rgba) fail(CU_ERR_ILLEGAL_ARGUMENT, "rgba is NULL");
        renderer(r)->read_output(rgba, count);
    });
}
int cu_renderer_read_pixel_info(cu_renderer r, int primary, void *dst, size_t bytes) {
    return locked([&] {
        if (!dst) fail(CU_ERR_ILLEGAL_ARGUMENT, "dst is NULL");
        renderer(r)->read_pixel_info(primary != 0, dst, bytes);
    });
}

int cu_plane_segment_from_center(double cx, double cy, double zoom, int w, int h, double out[4]) {
    return guarded([&] {
        if (!out) fail(CU_ERR_ILLEGAL_ARGUMENT, "out is NULL");
        if (w <= 0 || h <= 0) fail(CU_ERR_ILLEGAL_ARGUMENT, "canvas size must be positive");
        plane_from_center(cx, cy, zoom, w, h, out);
    });
}

int cu_precision_for_segment(double lbx, double lby, double rtx, double rty, int w, int h,
                             int *precision) {
    return guarded([&] {
        if (!precision) fail(CU_ERR_ILLEGAL_ARGUMENT, "precision is NULL");
        *precision = precision_for(Segment{lbx, lby, rtx, rty}, w, h);
    });
}

int cu_launch_geometry(int w, int h, unsigned *gx, unsigned *gy) {
    return guarded([&] {
        if (!gx || !gy) fail(CU_ERR_ILLEGAL_ARGUMENT, "gx/gy is NULL");
        *gx = *gy = 0;
        grid_for(w, h, gx, gy);
    });
}

// ImageHelpers.createDefaultColorPalette (util/ImageHelpers.java:77-111)
int cu_default_palette(uint32_t *out, size_t count) {
    return guarded([&] {
        const int max = 256, full = 255;
        if (!out || count < (size_t)max * 6) fail(CU_ERR_ILLEGAL_ARGUMENT, "need 1536 entries");
        auto rgba = [](int r, int g, int b) {
            return ((uint32_t)r & 0xffu) | (((uint32_t)g & 0xffu) << 8) |
                   (((uint32_t)b & 0xffu) << 16) | 0xff000000u;
        };
        for (int i = 0; i < max * 6; ++i) {
            switch (i / max) {
                case 0: out[i] = rgba(i, 0, std::min(full, full / 2 + i)); break;
                case 1: out[i] = rgba(full, 0, full - i); break;
                case 2: out[i] = rgba(full, i, 0); break;
                case 3: out[i] = rgba(full - i, full, 0); break;
                case 4: out[i] = rgba(0, full, i); break;
                default: out[i] = rgba(0, full - i, full); break;
            }
        }
    });
}

// ImageHelpers.fromBGRAtoRGBA (util/ImageHelpers.java:118-130): reverse the
// bytes, arithmetic shift right by 8, OR the alpha byte back in.
int cu_palette_from_argb(const uint32_t *argb, uint32_t *rgba, size_t count) {
    return guarded([&] {
        if ((!argb || !rgba) && count) fail(CU_ERR_ILLEGAL_ARGUMENT, "NULL buffer");
        for (size_t i = 0; i < count; ++i) {
            int32_t original = (int32_t)argb[i];
            int32_t alpha = original >> 24;
            int32_t updated = (int32_t)__builtin_bswap32((uint32_t)original);
            updated >>= 8;
            updated |= (int32_t)(((uint32_t)alpha << 24) & 0xff000000u);
            rgba[i] = (uint32_t)updated;
        }
    });
}

}  // extern "C"
