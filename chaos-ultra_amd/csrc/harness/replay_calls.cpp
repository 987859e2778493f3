// replay_calls.cpp — a native host that drives the C ABI through the reference
// host's call sequences, as the Java host would through a JNI shim:
//
//   (A) init / module load   Main -> GLRenderer.init -> CudaFractalRendererProvider
//                            -> getDefaultRenderer -> initializeRendering
//                            (gui/Main.java:15-17, rendering/GLRenderer.java:55-97,248-278,
//                             cudarenderer/CudaFractalRendererProvider.java:33-66,
//                             cudarenderer/CudaFractalRenderer.java:65-101)
//   (B) renderQuality        GLRenderer.display -> renderQuality -> compose
//                            (rendering/GLRenderer.java:113-185,
//                             cudarenderer/CudaFractalRenderer.java:187-206)
//   (C) renderFast           zoom frames with sample reuse and foveation
//                            (cudarenderer/CudaFractalRenderer.java:159-184,
//                             rendering/RenderingController.java:130-150)
//   (D) reload / debug       getRenderer(name, forceReload) closes the active renderer
//                            and reloads the module; launchDebugKernel; fractal switch
//                            (cudarenderer/CudaFractalRendererProvider.java:43-66,
//                             rendering/RenderingController.java:232-234,279-283)
//   (E) custom parameters    setFractalCustomParams("re;im") -> julia_c
//                            (cudarenderer/modules/ModuleJulia.java:28-37)
//
// Each rendered step writes its primary pixel_info_t buffer (rows packed,
// width*16 bytes) and composed RGBA8 image to <out_dir>/<step>.info and
// <step>.rgba, and a line of <out_dir>/manifest.jsonl with the model it
// rendered, so the outputs can be checked against an independent restatement.
// Every status code is checked; an unexpected one ends the run with exit 1.
//
//   usage: chaosultra_replay <kernels_dir> <out_dir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/chaosultra.h"

namespace {

const char *g_out = nullptr;
FILE *g_manifest = nullptr;

[[noreturn]] void die(const char *what, int rc) {
    std::fprintf(stderr, "replay: %s -> status %d (%s)\n", what, rc, cu_last_error());
    std::exit(1);
}

void ok(int rc, const char *what) {
    if (rc != CU_OK) die(what, rc);
}

void expect(int rc, int want, const char *what) {
    if (rc != want) die(what, rc);
    std::printf("[replay] %s -> status %d as expected\n", what, rc);
}

void write_file(const std::string &name, const void *p, size_t bytes) {
    const std::string path = std::string(g_out) + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "replay: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::fclose(f);
}

// The model of a renderQuality / renderFast call (the Java model fields).
cu_rendering_model make_model(const double seg[4], int max_iter, float max_ss) {
    cu_rendering_model m;
    std::memset(&m, 0, sizeof m);
    m.left_bottom_x = seg[0];
    m.left_bottom_y = seg[1];
    m.right_top_x = seg[2];
    m.right_top_y = seg[3];
    m.max_iterations = max_iter;
    m.max_super_sampling = max_ss;
    m.use_adaptive_super_sampling = 1;
    return m;
}

// Zoom the segment about pixel `focus` by `coeff` (RenderingController.zoomAt's geometry).
void zoom_about(double seg[4], int fx_px, int fy_px, int w, int h, double coeff) {
    const double fx = seg[0] + (seg[2] - seg[0]) * fx_px / w;
    const double fy = seg[3] - (seg[3] - seg[1]) * fy_px / h;
    const double lbx = fx - (fx - seg[0]) * coeff, lby = fy - (fy - seg[1]) * coeff;
    const double rtx = fx + (seg[2] - fx) * coeff, rty = fy + (seg[3] - fy) * coeff;
    seg[0] = lbx;
    seg[1] = lby;
    seg[2] = rtx;
    seg[3] = rty;
}

// Read back the step's outputs and record it.
void record(const char *step, cu_renderer r, const char *fractal, const cu_rendering_model &m,
            const double *prev_seg, const char *prev_step, int w, int h, uint32_t flags,
            const char *julia_c) {
    if (m.error_count != 0) {
        std::fprintf(stderr, "replay: step %s logged %d errors: %s\n", step, m.error_count,
                     m.last_error);
        std::exit(1);
    }
    std::vector<uint8_t> info((size_t)w * h * 16);
    std::vector<uint32_t> rgba((size_t)w * h);
    ok(cu_renderer_read_pixel_info(r, 1, info.data(), info.size()), "read_pixel_info");
    ok(cu_renderer_read_output(r, rgba.data(), rgba.size()), "read_output");
    write_file(std::string(step) + ".info", info.data(), info.size());
    write_file(std::string(step) + ".rgba", rgba.data(), rgba.size() * 4);
    std::fprintf(g_manifest,
                 "{\"step\": \"%s\", \"fractal\": \"%s\", \"width\": %d, \"height\": %d, "
                 "\"segment\": [%.17g, %.17g, %.17g, %.17g], \"max_iter\": %d, "
                 "\"max_ss\": %.9g, \"flags\": %u, \"precision\": %d, \"julia_c\": \"%s\", "
                 "\"focus\": [%d, %d]",
                 step, fractal, w, h, m.left_bottom_x, m.left_bottom_y, m.right_top_x,
                 m.right_top_y, m.max_iterations, (double)m.max_super_sampling, flags,
                 m.floating_point_precision, julia_c, m.mouse_focus_x, m.mouse_focus_y);
    if (prev_seg)
        std::fprintf(g_manifest, ", \"prev_step\": \"%s\", \"prev_segment\": [%.17g, %.17g, %.17g, %.17g]",
                     prev_step, prev_seg[0], prev_seg[1], prev_seg[2], prev_seg[3]);
    std::fprintf(g_manifest, "}\n");
    std::fflush(g_manifest);
    std::printf("[replay] step %s rendered (%s %dx%d, precision %d)\n", step, fractal, w, h,
                m.floating_point_precision);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <kernels_dir> <out_dir>\n", argv[0]);
        return 2;
    }
    const char *kernels = argv[1];
    g_out = argv[2];
    g_manifest = std::fopen((std::string(g_out) + "/manifest.jsonl").c_str(), "w");
    if (!g_manifest) {
        std::fprintf(stderr, "replay: cannot open the manifest in %s\n", g_out);
        return 1;
    }
    const int W = 320, H = 200;

    // ---- (A) init / module load ----
    ok(cu_init(0), "cu_init");
    cu_provider prov = 0, missing = 0;
    cu_renderer none = 0;
    // modules load on first use: a missing code object fails there, as cuModuleLoad does
    ok(cu_provider_create("/nonexistent/kernels", &missing), "provider on a missing dir");
    expect(cu_provider_get_default_renderer(missing, &none), CU_ERR_ILLEGAL_ARGUMENT,
           "getDefaultRenderer with no code object");
    ok(cu_provider_destroy(missing), "provider destroy");
    ok(cu_provider_create(kernels, &prov), "cu_provider_create");
    char names[256];
    ok(cu_provider_available_fractals(prov, names, sizeof names), "available_fractals");
    std::printf("[replay] available fractals: %s\n", names);
    cu_renderer r = 0;
    ok(cu_provider_get_default_renderer(prov, &r), "get_default_renderer");
    char name[64];
    ok(cu_renderer_get_fractal_name(r, name, sizeof name), "get_fractal_name");
    int state = -1;
    ok(cu_renderer_get_state(r, &state), "get_state");
    if (state != CU_STATE_NOT_INITIALIZED) die("state before initializeRendering", state);
    std::vector<uint32_t> palette(1536);
    ok(cu_default_palette(palette.data(), palette.size()), "default_palette");
    write_file("palette.rgba", palette.data(), palette.size() * 4);
    ok(cu_renderer_initialize_rendering(r, W, H, palette.data(), (int)palette.size()),
       "initializeRendering");

    // ---- (B) renderQuality ----
    double seg[4];
    ok(cu_plane_segment_from_center(-0.7453, 0.1127, 0.01, W, H, seg), "segment");
    cu_rendering_model m = make_model(seg, 800, 5.0f);
    m.use_sample_reuse = 1;
    ok(cu_renderer_render_quality(r, &m), "renderQuality");
    record("B", r, name, m, nullptr, nullptr, W, H, CU_FLAG_ADAPTIVE_SS, "");

    // ---- (C) renderFast: three zoom-in frames about a focus, reuse + foveation ----
    const int focus_x = 200, focus_y = 70;
    const char *prev_step = "B";
    const char *steps[3] = {"C1", "C2", "C3"};
    for (int k = 0; k < 3; ++k) {
        double old_seg[4] = {m.left_bottom_x, m.left_bottom_y, m.right_top_x, m.right_top_y};
        double s2[4] = {old_seg[0], old_seg[1], old_seg[2], old_seg[3]};
        zoom_about(s2, focus_x, focus_y, W, H, 0.977);
        m.left_bottom_x = s2[0];
        m.left_bottom_y = s2[1];
        m.right_top_x = s2[2];
        m.right_top_y = s2[3];
        m.max_super_sampling = 4.0f;
        m.zooming = 1;
        m.zooming_in = 1;
        m.use_foveated_rendering = 1;
        m.mouse_focus_x = focus_x;
        m.mouse_focus_y = focus_y;
        ok(cu_renderer_render_fast(r, &m), "renderFast");
        const uint32_t flags = CU_FLAG_ADAPTIVE_SS | CU_FLAG_SAMPLE_REUSE | CU_FLAG_ZOOMING |
                               CU_FLAG_ZOOMING_IN | CU_FLAG_FOVEATION;
        record(steps[k], r, name, m, old_seg, prev_step, W, H, flags, "");
        prev_step = steps[k];
    }

    // ---- (D) reload, debug kernel, fractal switch ----
    cu_renderer same = 0;
    ok(cu_provider_get_renderer(prov, "mandelbrot", 0, &same), "getRenderer(mandelbrot)");
    if (same != r) die("getRenderer without forceReload returned another renderer", same);
    cu_renderer r2 = 0;
    ok(cu_provider_get_renderer(prov, "mandelbrot", 1, &r2), "getRenderer(mandelbrot, force)");
    if (r2 == r) die("forceReload returned the old renderer", r2);
    ok(cu_renderer_get_state(r, &state), "get_state of the reloaded-away renderer");
    if (state != CU_STATE_NOT_INITIALIZED) die("forceReload left the old renderer ready", state);
    cu_rendering_model stale = make_model(seg, 100, 1.0f);
    expect(cu_renderer_render_quality(r, &stale), CU_ERR_ILLEGAL_STATE,
           "renderQuality on the closed renderer");
    ok(cu_renderer_initialize_rendering(r2, W, H, palette.data(), (int)palette.size()),
       "initializeRendering after reload");
    ok(cu_renderer_launch_debug_kernel(r2), "launchDebugKernel");
    cu_rendering_model m2 = make_model(seg, 800, 5.0f);
    ok(cu_renderer_render_quality(r2, &m2), "renderQuality after reload");
    record("D", r2, name, m2, nullptr, nullptr, W, H, CU_FLAG_ADAPTIVE_SS, "");
    cu_renderer rj = 0;
    ok(cu_provider_get_renderer(prov, "julia", 0, &rj), "getRenderer(julia)");
    ok(cu_renderer_get_state(r2, &state), "get_state of the switched-away renderer");
    if (state != CU_STATE_NOT_INITIALIZED) die("switching fractals left the old renderer ready", state);

    // ---- (E) custom parameters on julia ----
    const int WJ = 256, HJ = 160;
    cu_rendering_model mj;
    std::memset(&mj, 0, sizeof mj);
    ok(cu_renderer_supply_default_values(rj, &mj, WJ, HJ), "supplyDefaultValues");
    ok(cu_renderer_initialize_rendering(rj, WJ, HJ, palette.data(), (int)palette.size()),
       "initializeRendering(julia)");
    ok(cu_renderer_set_fractal_custom_params(rj, "-0.8;0.156"), "setFractalCustomParams");
    double segj[4];
    ok(cu_plane_segment_from_center(0.0, 0.0, 2.5, WJ, HJ, segj), "segment");
    cu_rendering_model me = make_model(segj, 500, 4.0f);
    ok(cu_renderer_render_quality(rj, &me), "renderQuality(julia)");
    record("E", rj, "julia", me, nullptr, nullptr, WJ, HJ, CU_FLAG_ADAPTIVE_SS, "-0.8;0.156");
    expect(cu_renderer_set_fractal_custom_params(rj, "abc;1"), CU_ERR_ILLEGAL_ARGUMENT,
           "malformed custom parameters");

    ok(cu_renderer_free_rendering_resources(rj), "freeRenderingResources");
    expect(cu_renderer_free_rendering_resources(rj), CU_ERR_ILLEGAL_STATE, "second free");
    ok(cu_renderer_close(rj), "close");
    ok(cu_provider_destroy(prov), "provider destroy");
    std::fclose(g_manifest);
    std::printf("[replay] done\n");
    return 0;
}
