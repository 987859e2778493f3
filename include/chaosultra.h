/*
 * chaosultra.h — C ABI of the MI355X escape-time render path.
 *
 * Three layers, all plain C (pointers, sizes, status codes):
 *
 *  1. Driver layer.  The calls the reference's JCuda binding makes on this path,
 *     one entry point each (context, module, function, global, pitched memory,
 *     launch, synchronize).  A Java/JNI or Panama shim binds these where the
 *     reference bound jcuda.driver.JCudaDriver / jcuda.runtime.JCuda.
 *  2. Renderer layer.  The reference's plugin surface (FractalRendererProvider,
 *     FractalRenderer, RenderingModel) implemented natively in C++ on top of
 *     layer 1 and exported as handles, so a thin Java shim can forward each
 *     interface method to one call.
 *  3. Escape-family layer.  Throughput forms with no counterpart in the
 *     reference: planar escape-time tiles (pixel centres or seeded jitter,
 *     any bailout radius, smooth value), the persistent multi-tile launch,
 *     palette compose of smooth values, and the FP32 perturbation path.
 *
 * Status codes map 1:1 onto the reference's exception classes; the message of
 * the last failing call on the calling thread is returned by cu_last_error().
 * Every call is synchronous unless it takes a stream argument.
 */
#ifndef CHAOSULTRA_H
#define CHAOSULTRA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* status codes                                                              */
/* ------------------------------------------------------------------------- */
enum cu_status {
    CU_OK = 0,
    CU_ERR_ILLEGAL_ARGUMENT = 1, /* java.lang.IllegalArgumentException        */
    CU_ERR_INITIALIZATION = 2,   /* cudarenderer.CudaInitializationException   */
    CU_ERR_RENDERER = 3,         /* rendering.FractalRendererException         */
    CU_ERR_ILLEGAL_STATE = 4,    /* java.lang.IllegalStateException            */
    CU_ERR_DEVICE = 5            /* jcuda.CudaException (any other HIP error)  */
};

/* Message of the last non-CU_OK return on this thread ("" if none). */
const char *cu_last_error(void);

/* ------------------------------------------------------------------------- */
/* 1. driver layer                                                           */
/* ------------------------------------------------------------------------- */

/* Replaces CudaHelpers.cudaInit (cudarenderer/CudaHelpers.java:27-45):
 * cuInit + cuDeviceGet(device) + cuCtxCreate.  Idempotent. */
int cu_init(int device);
/* Number of visible devices; 0 (and CU_OK) on a machine without a GPU. */
int cu_device_count(int *count);

typedef struct cu_module_s *cu_module;
typedef struct cu_function_s *cu_function;

/* cuModuleLoad (cudarenderer/FractalRenderingModule.java:73-88).  A missing
 * file is CU_ERR_ILLEGAL_ARGUMENT and is reported before any device call;
 * other load failures are CU_ERR_INITIALIZATION. */
int cu_module_load(const char *path, cu_module *out);
/* cuModuleUnload (cudarenderer/FractalRenderingModule.java:140-147). */
int cu_module_unload(cu_module module);
/* cuModuleGetFunction (cudarenderer/CudaKernel.java:28-43); an unknown name is
 * CU_ERR_ILLEGAL_ARGUMENT, as the reference's CUDA_ERROR_NOT_FOUND mapping. */
int cu_module_get_function(cu_module module, const char *name, cu_function *out);
/* cuModuleGetGlobal (cudarenderer/FractalRenderingModule.java:163-169). */
int cu_module_get_global(cu_module module, const char *name, void **dptr, size_t *bytes);
/* writeToConstantMemory (cudarenderer/FractalRenderingModule.java:171-227):
 * fails with CU_ERR_ILLEGAL_ARGUMENT if the symbol is smaller than `bytes`;
 * otherwise copies min(symbol size, bytes) bytes and zero-fills the rest of the
 * symbol (the reference copies the full symbol size from a host array of the
 * value's size). */
int cu_module_write_global(cu_module module, const char *name, const void *value, size_t bytes);

/* cuMemAlloc / cuMemAllocPitch (cudarenderer/DeviceMemoryDoubleBuffer2D.java:116-146) */
int cu_mem_alloc(void **dptr, size_t bytes);
int cu_mem_alloc_pitch(void **dptr, size_t *pitch, size_t width_bytes, size_t height,
                       unsigned element_bytes);
int cu_mem_free(void *dptr);
int cu_memcpy_htod(void *dst, const void *src, size_t bytes);
int cu_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int cu_memcpy_dtod(void *dst, const void *src, size_t bytes);
/* cuMemcpy2D device→host (cudarenderer/CudaFractalRenderer.java:119-134) */
int cu_memcpy_2d_dtoh(void *dst, size_t dst_pitch, const void *src, size_t src_pitch,
                      size_t width_bytes, size_t height);
/* Synchronous fill of `count` 32-bit words. */
int cu_memset_u32(void *dptr, uint32_t value, size_t count);
/* Page-locked host memory (hipHostMalloc) for reading frames back: device-to-host
 * copies into it run at several times the rate of copies into pageable memory.
 * No counterpart in the reference, whose frames reach the host through GL
 * interop; for callers that read frames back, such as a serving front end. */
int cu_host_alloc(void **ptr, size_t bytes);
int cu_host_free(void *ptr);

/* cuLaunchKernel on the null stream (cudarenderer/CudaFractalRenderer.java:371-378);
 * `params` is the JCuda-style array of pointers to each argument. */
int cu_launch_kernel(cu_function f, unsigned grid_x, unsigned grid_y, unsigned grid_z,
                     unsigned block_x, unsigned block_y, unsigned block_z,
                     unsigned shared_bytes, void **params);
/* cuCtxSynchronize */
int cu_ctx_synchronize(void);

/* Streams and events (HIP-native additions used by the batched request loop). */
typedef struct cu_stream_s *cu_stream;
int cu_stream_create(cu_stream *out);
int cu_stream_destroy(cu_stream s);
int cu_stream_synchronize(cu_stream s);
/* Non-blocking completion query: *done = 1 when all work on s has finished. */
int cu_stream_query(cu_stream s, int *done);

/* ------------------------------------------------------------------------- */
/* 2. renderer layer (FractalRenderer / FractalRendererProvider)             */
/* ------------------------------------------------------------------------- */

/* FractalRendererState (rendering/FractalRendererState.java) */
enum { CU_STATE_NOT_INITIALIZED = 0, CU_STATE_READY_TO_RENDER = 1 };
/* FloatPrecision (util/FloatPrecision.java) */
enum { CU_PRECISION_SINGLE = 0, CU_PRECISION_DOUBLE = 1, CU_PRECISION_TOO_BIG = 2 };

/* Flag bits of the kernel flag word (fractalRendererGeneric.cu:157-161,
 * cudarenderer/KernelMain.java:12-13, cudarenderer/KernelAdvanced.java:17-20). */
enum {
    CU_FLAG_ADAPTIVE_SS = 1u << 0,
    CU_FLAG_VISUALISE_SAMPLE_COUNT = 1u << 1,
    CU_FLAG_FOVEATION = 1u << 2,
    CU_FLAG_SAMPLE_REUSE = 1u << 3,
    CU_FLAG_ZOOMING = 1u << 4,
    CU_FLAG_ZOOMING_IN = 1u << 5
};

/* RenderingModel (rendering/model/RenderingModel.java:6-20 and the seven
 * interfaces it extends) as a plain struct.  Fields marked in/out are written
 * back by the renderer as the Java model setters would be. */
typedef struct cu_rendering_model {
    double left_bottom_x, left_bottom_y, right_top_x, right_top_y; /* PlaneSegment */
    int32_t max_iterations;                                        /* IterationLimitModel */
    float max_super_sampling;                                      /* SuperSamplingModel */
    int32_t use_adaptive_super_sampling;
    int32_t visualise_sample_count;
    int32_t use_foveated_rendering;                                /* FoveatedRenderingModel */
    int32_t mouse_focus_x, mouse_focus_y;
    int32_t zooming, zooming_in;
    int32_t use_sample_reuse;                                      /* SampleReuseModel */
    int32_t sample_reuse_cache_dirty;                              /* in/out */
    int32_t floating_point_precision;                              /* out: CU_PRECISION_* */
    int32_t error_count;   /* out: PublicErrorLogger.logError calls made by this render */
    char last_error[256];  /* out: text of the last logged error                      */
    char fractal_custom_params[256]; /* DefaultFractalModel */
} cu_rendering_model;

typedef int32_t cu_provider; /* handle, > 0 */
typedef int32_t cu_renderer; /* handle, > 0; stale handles are CU_ERR_ILLEGAL_STATE */

/* CudaFractalRendererProvider() — registers the fractal modules found as
 * <kernels_dir>/<name>.hsaco (the reference's cudaKernelsDir property). */
int cu_provider_create(const char *kernels_dir, cu_provider *out);
int cu_provider_destroy(cu_provider p);
/* getAvailableFractals(): comma-separated, sorted. */
int cu_provider_available_fractals(cu_provider p, char *buf, size_t buf_len);
int cu_provider_get_default_renderer(cu_provider p, cu_renderer *out);
int cu_provider_get_renderer(cu_provider p, const char *fractal_name, int force_reload,
                             cu_renderer *out);

/* FractalRenderer (rendering/FractalRenderer.java:29-76).  The GL textures of
 * GLParams are replaced by an RGBA8 device image of width×height and a palette
 * of palette_length RGBA words (R in the least significant byte). */
int cu_renderer_initialize_rendering(cu_renderer r, int width, int height,
                                     const uint32_t *palette, int palette_length);
int cu_renderer_free_rendering_resources(cu_renderer r);
int cu_renderer_get_state(cu_renderer r, int *state);
int cu_renderer_get_width(cu_renderer r, int *width);
int cu_renderer_get_height(cu_renderer r, int *height);
int cu_renderer_launch_debug_kernel(cu_renderer r);
int cu_renderer_render_fast(cu_renderer r, cu_rendering_model *model);
int cu_renderer_render_quality(cu_renderer r, cu_rendering_model *model);
int cu_renderer_close(cu_renderer r);
int cu_renderer_set_fractal_custom_params(cu_renderer r, const char *text);
int cu_renderer_get_fractal_name(cu_renderer r, char *buf, size_t buf_len);
int cu_renderer_supply_default_values(cu_renderer r, cu_rendering_model *model, int canvas_width,
                                      int canvas_height);
/* Read back the composed RGBA8 image (width*height words, row-major). */
int cu_renderer_read_output(cu_renderer r, uint32_t *rgba, size_t count);
/* Iterations executed by the renderer's fractal module since it was loaded: the
 * sum of its cu_iteration_slots (each sample adds the loop bodies computeFractal
 * ran — the escape step, or maxIterations for an orbit that never escapes).  The
 * reference has no counterpart; a caller takes differences around a render to
 * get its Giga-iterations/s. */
int cu_renderer_iterations(cu_renderer r, unsigned long long *total);
/* Read back the primary (1) or secondary (0) pixel_info_t buffer, rows packed
 * (width*16 bytes each). */
int cu_renderer_read_pixel_info(cu_renderer r, int primary, void *dst, size_t bytes);
/* Debug mode (-Ddebug=true): launch errors are returned instead of only logged. */
int cu_set_debug_mode(int enabled);

/* ---- host-side model logic (no device needed) ---- */

/* Model.setPlaneSegmentFromCenter (rendering/Model.java:247-256); out = lbx, lby, rtx, rty */
int cu_plane_segment_from_center(double center_x, double center_y, double zoom, int canvas_width,
                                 int canvas_height, double out[4]);
/* RenderingKernel.isSegmentBoundsAt{Float,Double}Limit + CudaFractalRenderer.updateFloatPrecision
 * (cudarenderer/RenderingKernel.java:124-138, cudarenderer/CudaFractalRenderer.java:409-419) */
int cu_precision_for_segment(double lbx, double lby, double rtx, double rty, int width, int height,
                             int *precision);
/* Grid of ⌈w/32⌉×⌈h/32⌉ blocks; CU_ERR_RENDERER beyond 65535 per dimension
 * (cudarenderer/CudaFractalRenderer.java:242-253). */
int cu_launch_geometry(int width, int height, unsigned *grid_x, unsigned *grid_y);
/* ImageHelpers.createDefaultColorPalette (util/ImageHelpers.java:77-111): 1536 entries. */
int cu_default_palette(uint32_t *out, size_t count);
/* ImageHelpers.fromBGRAtoRGBA (util/ImageHelpers.java:118-130) */
int cu_palette_from_argb(const uint32_t *argb, uint32_t *rgba, size_t count);
/* ImageHelpers.loadColorPaletteOrDefault / loadColorPaletteFromFile
 * (util/ImageHelpers.java:43-73): the first row of the PNG at `path` as RGBA words
 * (getRGB's ARGB through fromBGRAtoRGBA), or the 1536-entry default palette when
 * `path` is NULL or empty, does not exist, or is not a readable PNG (the reference
 * falls back on an IOException; a file ImageIO cannot identify, on which the
 * reference would fail with a NullPointerException, also falls back here).
 * *count = entries written; *loaded (may be NULL) = 1 if the file was used.
 * CU_ERR_ILLEGAL_ARGUMENT if `cap` is too small. */
int cu_palette_load_or_default(const char *path, uint32_t *out, size_t cap, size_t *count,
                               int *loaded);
/* ImageHelpers.saveImageToFile(rgbs, w, h, name, "png") (util/ImageHelpers.java:192-214):
 * an 8-bit RGB PNG of width*height RGBA words, R from the least significant byte,
 * alpha dropped; missing parent directories are created. */
int cu_image_save_png(const char *path, const uint32_t *rgba, int width, int height);

/* ------------------------------------------------------------------------- */
/* 3. escape-family layer                                                    */
/* ------------------------------------------------------------------------- */

enum { CU_FRACTAL_MANDELBROT = 0, CU_FRACTAL_JULIA = 1 };
enum { CU_SAMPLE_CENTRE = 0, CU_SAMPLE_JITTER = 1, CU_SAMPLE_CORNER = 2 };

/* One rectangular tile of samples.  Sample s of pixel (px, py) is taken at
 *   X = left + (px + dx) * pixel_w,   Y = top - (py + dy) * pixel_h
 * (the reference's op order, fractalRendererGeneric.cu:97,119-124) in FP64, or
 * with the *_f fields in FP32.  (dx, dy) is (0.5, 0.5) for CU_SAMPLE_CENTRE,
 * (0, 0) for CU_SAMPLE_CORNER, and the seeded SplitMix64 jitter for
 * CU_SAMPLE_JITTER.  Mandelbrot: z0 = 0, c = (X, Y); Julia: z0 = (X, Y),
 * c = (julia_cx, julia_cy).  Outputs per pixel: count = the fractal's
 * computeFractal value (Mandelbrot interior → 0, Julia interior → max_iter),
 * integer mean over samples; smooth = mean over samples of
 * n + 1 - log2(ln|z_n|^2 / ln bailout_sq) (0 for an interior sample). */
typedef struct cu_tile_desc {
    double left, top, pixel_w, pixel_h;         /* FP64 mapping */
    float left_f, top_f, pixel_w_f, pixel_h_f;  /* FP32 mapping (float-cast segment) */
    double julia_cx, julia_cy;
    double bailout_sq;
    uint64_t seed;
    uint64_t out_offset;  /* element offset of pixel (0,0) in the output planes */
    uint64_t work_offset; /* prefix sum of width*height over earlier tiles */
    uint32_t width, height;
    uint32_t out_pitch;   /* elements per output row */
    uint32_t max_iter;
    uint32_t kind;        /* CU_FRACTAL_* */
    uint32_t spp;         /* samples per pixel, >= 1 */
    uint32_t pattern;     /* CU_SAMPLE_* */
    uint32_t reserved;
} cu_tile_desc;

/* Fill a tile from a plane segment: pixel sizes are (rt - lb) / size computed in
 * FP64 and, for the *_f fields, in FP32 from the float-cast segment (the
 * reference's KernelMainFloat cast, cudarenderer/KernelMainFloat.java:19-21). */
int cu_tile_init(cu_tile_desc *tile, double lbx, double lby, double rtx, double rty,
                 uint32_t width, uint32_t height, uint32_t max_iter, double bailout_sq,
                 uint32_t kind, double julia_cx, double julia_cy, uint32_t spp, uint32_t pattern,
                 uint64_t seed);

/* Tuning of the persistent launch (chosen by the autotuner). */
typedef struct cu_tune {
    uint32_t block_steps;   /* K: unchecked iterations per block (8, 16, 32)       */
    uint32_t waves_per_cu;  /* resident waves per CU in the persistent grid         */
    uint32_t chunk;         /* pixels taken per global-queue grab                    */
    uint32_t refill_lanes;  /* retire/refill when at least this many lanes are idle  */
} cu_tune;

typedef struct cu_engine_s *cu_engine;
/* Default tuning of the engine's launches (the grid launch uses block_steps 8
 * or 16; the persistent launch uses all four fields when called with NULL). */
int cu_engine_configure(cu_engine e, const cu_tune *tune);
/* Loads <kernels_dir>/escape_family.hsaco. */
int cu_engine_create(const char *kernels_dir, cu_engine *out);
int cu_engine_destroy(cu_engine e);

/* Variant of the FP64/FP32 step: 0 = exact form ((2x)*y + cy, 7 VALU ops),
 * 1 = fused form (fma(x*y, 2, cy), 6 VALU ops; equal to the exact form unless
 * |x*y| falls below the smallest normal), -1 = choose (fused unless the tile's
 * inputs can reach that range; see cu_tile_fused_safe). */
int cu_tile_fused_safe(const cu_tile_desc *tile, int precision, int *safe);

/* Iterations executed (an interior sample counts max_iter) are accumulated in
 * CU_ITER_SLOTS device u64 slots: each wave adds to one slot, the total is the
 * sum of the slots (cu_iters_total). */
#define CU_ITER_SLOTS 1024u
int cu_iters_total(const unsigned long long *d_iters, unsigned long long *total);

/* Grid launch over one tile.  precision: 32 or 64.  When spp is a power of two
 * <= 64 each lane takes one sample and a wave covers 64/spp pixels; otherwise
 * each lane takes one pixel and its samples in sequence.  d_counts / d_smooth
 * are device planes addressed through out_offset/out_pitch (d_smooth may be
 * NULL); d_iters (CU_ITER_SLOTS zeroed u64, may be NULL) receives the
 * iterations executed. */
int cu_escape_tile(cu_engine e, int precision, int variant, const cu_tile_desc *tile,
                   uint32_t *d_counts, double *d_smooth, unsigned long long *d_iters,
                   cu_stream stream);

/* Persistent lane-refill launch over n_tiles (<= 1024) tiles whose descriptors
 * are already in device memory (d_tiles, work_offset = prefix sum of pixel
 * counts, total_pixels = their sum).  d_queue is a device u64 work counter that
 * the call clears with a kernel.  variant must be 0 or 1 here.  tune may be
 * NULL (K=16, 32 waves/CU, chunk 256, refill at 16 lanes). */
int cu_escape_tiles_persistent(cu_engine e, int precision, int variant,
                               const cu_tile_desc *d_tiles, uint32_t n_tiles,
                               uint64_t total_pixels, uint32_t *d_counts, double *d_smooth,
                               unsigned long long *d_iters, unsigned long long *d_queue,
                               const cu_tune *tune, cu_stream stream);

/* The persistent launch (same arguments as cu_escape_tiles_persistent, its clears
 * included, and d_iters' slots cleared as well) captured into a HIP graph, so a
 * batch of tiles runs as one replayable launch; every node is a kernel.  The
 * graph keeps the device pointers it was captured with. */
typedef struct cu_graph_s *cu_graph;
int cu_escape_tiles_persistent_capture(cu_engine e, int precision, int variant,
                                       const cu_tile_desc *d_tiles, uint32_t n_tiles,
                                       uint64_t total_pixels, uint32_t *d_counts, double *d_smooth,
                                       unsigned long long *d_iters, unsigned long long *d_queue,
                                       const cu_tune *tune, cu_graph *out);
/* One replay of the graph on `stream` (HIP events around it). */
int cu_graph_launch(cu_graph g, cu_stream stream);
/* Milliseconds of the last replay (synchronises on it). */
int cu_graph_last_ms(cu_graph g, float *ms);
/* Waits for the last replay, then frees the graph. */
int cu_graph_destroy(cu_graph g);

/* Palette compose of a smooth plane: idx = L - (round(v) mod L) - 1
 * (mandelbrot.cu:27-35), RGBA8 out.  The palette is staged in LDS. */
int cu_compose_smooth(cu_engine e, const double *d_smooth, uint32_t smooth_pitch,
                      uint32_t *d_rgba, uint32_t rgba_pitch, uint32_t width, uint32_t height,
                      const uint32_t *d_palette, uint32_t palette_length, cu_stream stream);

/* Elapsed milliseconds of the last engine launch on `stream`, measured with HIP
 * events recorded on that stream around the kernel. */
int cu_engine_last_kernel_ms(cu_engine e, float *ms);

/* ---- frame pipeline (straggler compaction) ----
 * A frame owns its tiles (each spp a power of two <= 64) and all device
 * workspace.  Rendering runs pass 1 over every sample (one lane each), hands the
 * orbits still running in a wave off to a queue once no more than
 * handoff_lanes of its 64 lanes iterate, drains the queue in tail_passes
 * further passes (the last one to completion, unless a persistent drain launch
 * takes what they leave), and writes the pixel planes
 * (count: integer mean; smooth: mean summed in sample order) and, when a
 * palette is given, RGBA8.  Pixels whose samples all finish in pass 1 are
 * written there; the others are combined after the last pass.  Tiles are laid
 * out back to back in the pixel planes (out_offset/out_pitch/work_offset of the
 * given tiles are replaced). */
typedef struct cu_frame_tune {
    uint32_t block_steps;   /* 8 or 16 */
    uint32_t handoff_lanes; /* 1..32 */
    uint32_t tail_passes;   /* 1..8; with a drain 0..7, the hand-off passes run before it */
    uint32_t wave_shape;    /* 0..3: pass-1 wave covers a bw x bh pixel block, bw * bh = 64/spp;
                               0 near-square (spp 1: 8x8), each step twice as wide (16x4, 32x2, 64x1) */
    uint32_t drain_refill;  /* 0: the queue runs through tail_passes launches, the last to the end;
                               1..64: every tail pass hands off and one persistent drain launch
                               runs what is left to the end, refilling a wave's idle lanes once
                               this many wait */
} cu_frame_tune;
typedef struct cu_frame_s *cu_frame;
int cu_frame_create(cu_engine e, const cu_tile_desc *tiles, uint32_t n_tiles, int precision,
                    int variant, cu_frame *out);
int cu_frame_destroy(cu_frame f);
/* d_palette / d_rgba may be NULL (no compose).  tune may be NULL: block_steps 16, handoff_lanes 24,
   tail_passes 0, wave_shape 0, drain_refill 40 (pass 1, then one drain of its whole queue). */
int cu_frame_render(cu_frame f, const cu_frame_tune *tune, const uint32_t *d_palette,
                    uint32_t palette_length, uint32_t *d_rgba, cu_stream stream);
/* Device pointers of the pixel planes (total_pixels elements each). */
int cu_frame_outputs(cu_frame f, uint32_t **d_counts, double **d_smooth, uint64_t *total_pixels);
/* Iterations executed by the last render, in all passes and in pass 1 alone
 * (pass1 may be NULL).  Synchronises the frame's stream. */
int cu_frame_iterations(cu_frame f, unsigned long long *total, unsigned long long *pass1);
/* Milliseconds of the last render: pass 1 and the whole render, HIP events on
 * its stream (pass1_ms is 0 after a graph replay). */
int cu_frame_last_ms(cu_frame f, float *pass1_ms, float *total_ms);
/* Capture one render (same arguments as cu_frame_render) into a HIP graph owned
 * by the frame; every launch in it is a kernel (clears included), so the graph
 * can be replayed any number of times.  Recapturing replaces the graph. */
int cu_frame_capture(cu_frame f, const cu_frame_tune *tune, const uint32_t *d_palette,
                     uint32_t palette_length, uint32_t *d_rgba);
/* One launch of the captured graph on `stream`. */
int cu_frame_replay(cu_frame f, cu_stream stream);

/* ---- deep zoom: perturbation around a double-double reference orbit ----
 * No reference counterpart: the reference stops at its double kernel
 * (cudarenderer/RenderingKernel.java:124-138; cudarenderer/CudaFractalRenderer.java:217-220
 * only logs below the double ulp).  A W x H Mandelbrot frame centred on
 * C = (centre_x + centre_x_lo, centre_y + centre_y_lo): pixel (px, py) samples
 * C + ((px + 0.5 - W/2) * pixel_w, (H/2 - (py + 0.5)) * pixel_h), offsets in
 * double.  Counts follow mandelbrot.cu:22-23 (max_iter -> 0); bailout |z|^2 >=
 * bailout_sq.  Methods:
 *   CU_DEEP_DIRECT_F64   direct FP64 at the double-rounded c (the reference
 *                        rounding; bit-exact with the oracle)
 *   CU_DEEP_PERTURB_F32  FP32 deltas around one reference orbit at C computed on
 *   CU_DEEP_PERTURB_F64  FP64 deltas    the GPU in double-double, stored rounded
 * flags CU_DEEP_REBASE: Zhuoran rebasing (restart the reference when |z| < |delta|).
 * Without it, Pauldelbrot's test |z|^2 < glitch_eps |Z|^2 marks glitched pixels and
 * cu_deep_render re-renders them around up to max_references secondary
 * reference orbits (the glitched pixel nearest the glitched centroid). */
#define CU_DEEP_DIRECT_F64 1
#define CU_DEEP_PERTURB_F32 32
#define CU_DEEP_PERTURB_F64 64
#define CU_DEEP_REBASE 1u
/* flags CU_DEEP_REUSE_ORBIT: keep the primary orbit between renders while the
 * centre and bailout stay the same (it is then computed to the frame's creation
 * max_iter); results are those of a fresh orbit. */
#define CU_DEEP_REUSE_ORBIT 2u
typedef struct cu_deep_desc {
    double centre_x, centre_y;
    double centre_x_lo, centre_y_lo;
    double pixel_w, pixel_h;
    double bailout_sq;        /* >= 4 */
    uint32_t width, height;
    uint32_t max_iter;
    uint32_t flags;
    uint32_t max_references;  /* secondary references per render, <= 16 */
    float glitch_eps;         /* Pauldelbrot threshold on |z|^2 / |Z|^2, e.g. 1e-6 */
} cu_deep_desc;
typedef struct cu_deep_report {
    unsigned long long iterations; /* executed by the last render, all passes */
    uint32_t reference_length;     /* M: last index of the primary reference orbit */
    uint32_t references;           /* reference orbits computed by the last render (0: kept) */
    uint32_t glitched;             /* pixels still glitched after the last render */
    uint32_t reserved;
    float orbit_ms;                /* primary reference orbit */
    float render_ms;               /* first pass over all pixels */
    float total_ms;                /* whole render, glitch rounds included */
    float reserved2;
} cu_deep_report;
typedef struct cu_deep_s *cu_deep;
int cu_deep_create(cu_engine e, const cu_deep_desc *d, int method, cu_deep *out);
int cu_deep_destroy(cu_deep d);
/* Point a deep frame at another view of the same size and method (centre, pixel
 * sizes, bailout, flags; max_iter no larger than at creation).  Device buffers
 * are kept. */
int cu_deep_retarget(cu_deep d, const cu_deep_desc *desc);
/* d_palette / d_rgba may be NULL.  Synchronous when glitch rounds run. */
int cu_deep_render(cu_deep d, const uint32_t *d_palette, uint32_t palette_length,
                   uint32_t *d_rgba, cu_stream stream);
/* Device planes, W*H each (row-major): counts, smooth values, glitch flags (u8). */
int cu_deep_outputs(cu_deep d, uint32_t **d_counts, double **d_smooth, uint8_t **d_glitch);
/* Statistics of the last render (synchronises). */
int cu_deep_stats(cu_deep d, cu_deep_report *out);
/* Copy of the primary reference orbit of the last render: entries of 4 values
 * {Zx, Zy, |Z|^2/2, 0} in the delta precision; *length = M (entries M + 1). */
int cu_deep_orbit(cu_deep d, void *host_out, uint32_t capacity_entries, uint32_t *length);

/* ---- deep zoom without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Renders the frame cu_deep_render renders for a cu_deep
 * created with the same desc and method (perturbation with CU_DEEP_REBASE, so
 * no glitch rounds) into the caller's planes: counts (W*H u32, required),
 * smooth (W*H f64, may be NULL), RGBA8 through the palette (both or neither),
 * iterations into CU_ITER_SLOTS zeroed u64 slots (may be NULL).  The reference
 * orbit and its length live in a static device array of CU_DEEP_STATIC_BYTES
 * per GPU, held until the process ends; max_iter above CU_DEEP_STATIC_MAX_ITER is
 * CU_ERR_ILLEGAL_ARGUMENT.  The call enqueues two kernels on `stream` and
 * allocates nothing.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race.  CU_DEEP_REUSE_ORBIT is
 * accepted and has no effect (the orbit is computed on every call). */
#define CU_DEEP_STATIC_MAX_ITER 65535u
#define CU_DEEP_STATIC_BYTES ((CU_DEEP_STATIC_MAX_ITER + 1u) * 32u + 256u)
int cu_deep_render_static(int method, const cu_deep_desc *desc, uint32_t *d_counts,
                          double *d_smooth, const uint32_t *d_palette, uint32_t palette_length,
                          uint32_t *d_rgba, unsigned long long *d_iters, cu_stream stream);

/* ---- serving: continuous batching of zoom-session frame requests ----
 * Many sessions, one GPU.  A session owns the reference's ping-pong
 * pixel_info_t buffers, an RGBA8 image, its palette and a deep-zoom frame, all
 * resident so each frame reuses the previous frame's samples.  Sessions are
 * admitted against hbm_budget (footprint reserved at admission, released at
 * close; first come, first served).  Requests queue per session; whenever one
 * of `slots` streams is free the oldest request of an idle admitted session
 * starts on it, so sessions overlap on the device while each session's frames
 * run in order.  A frame with width >= deep_width runs the mandelbrot module's
 * fractalRenderAdvanced kernel (reuse + foveation around the view centre,
 * zooming in; fractalRenderMain for a first frame) and compose
 * (cudarenderer/CudaFractalRenderer.java:158-206), in FP64 unless
 * shallow_policy asks for the reference's float/double switch; a narrower one
 * runs the deep-zoom form (bailout 4, rebasing).  Replaces the single-session
 * display loop of rendering/GLRenderer.java:113-162.
 * A request that passes cu_serve_submit's checks but fails later (device
 * memory, a launch) completes with a non-zero cu_frame_done.status instead of
 * failing another call. */
enum { CU_SERVE_SHALLOW_FP64 = 0, CU_SERVE_SHALLOW_REFERENCE_SWITCH = 1 };
typedef struct cu_serve_config {
    uint64_t hbm_budget;       /* bytes for admitted sessions */
    double deep_width;         /* frames narrower than this use perturbation (e.g. 1e-12) */
    uint32_t slots;            /* frames in flight, one HIP stream each: 1..16 */
    uint32_t deep_method;      /* CU_DEEP_PERTURB_F32 or CU_DEEP_PERTURB_F64 */
    float max_super_sampling;  /* maxSuperSampling of fresh pixels, 1..64 */
    uint32_t shallow_policy;   /* CU_SERVE_SHALLOW_FP64 (default): the module kernels in FP64;
                                  CU_SERVE_SHALLOW_REFERENCE_SWITCH: FP32 unless the pixel is
                                  below the float ulp (cudarenderer/RenderingKernel.java:124-138) */
} cu_serve_config;
typedef struct cu_frame_request {
    double centre_x, centre_y; /* view centre */
    double width;              /* real-axis width of the view; height = width * H / W */
    uint64_t tag;              /* returned with the completion */
    uint32_t max_iter;
    uint32_t reserved;
} cu_frame_request;
typedef struct cu_frame_done {
    uint64_t tag;
    uint32_t session;
    uint32_t path;             /* 0 reuse (advanced kernel), 1 full (main kernel), 2 deep */
    double submit_s, start_s, done_s; /* seconds since cu_serve_create */
    unsigned long long iterations;    /* deep frames (module-kernel frames count in the total) */
    unsigned long long checksum;      /* sum of rgba[i] * (2i + 1) mod 2^64 of the frame */
    int32_t status;                   /* CU_OK, or the CU_ERR_* the frame failed with (its
                                         message: cu_serve_last_failure) */
    uint32_t reserved;
} cu_frame_done;
typedef struct cu_serve_report {
    unsigned long long iterations;    /* every frame so far */
    uint64_t frames;
    uint64_t bytes_in_use, peak_bytes;
    uint32_t sessions_admitted, sessions_waiting;
} cu_serve_report;
typedef struct cu_serve_s *cu_serve;
int cu_serve_create(const char *kernels_dir, const cu_serve_config *cfg, cu_serve *out);
int cu_serve_destroy(cu_serve s);
/* deep_max_iter: largest max_iter of the session's deep frames (0: no deep frames). */
int cu_serve_open(cu_serve s, uint32_t width, uint32_t height, const uint32_t *palette,
                  uint32_t palette_length, uint32_t deep_max_iter, uint32_t *session);
/* Checks the request (a finite positive width, max_iter > 0; a deep frame must lie
 * within |c| <= 2, have a positive pixel size and max_iter <= deep_max_iter) and
 * queues it; CU_ERR_ILLEGAL_ARGUMENT and nothing queued when a check fails. */
int cu_serve_submit(cu_serve s, uint32_t session, const cu_frame_request *req);
/* Release the session's memory once its queued frames are done. */
int cu_serve_close(cu_serve s, uint32_t session);
/* Progress: retire finished frames into out[], admit waiting sessions, start
 * ready frames.  Waits up to wait_ms for a completion while frames are in flight,
 * releasing the server's lock between its checks so other threads can submit
 * meanwhile. */
int cu_serve_poll(cu_serve s, cu_frame_done *out, uint32_t capacity, uint32_t *count,
                  int wait_ms);
/* Message of the last frame that completed with a non-zero status ("" if none). */
int cu_serve_last_failure(cu_serve s, char *buf, size_t buf_len);
/* The session's last completed image (no frame may be in flight). */
int cu_serve_read_rgba(cu_serve s, uint32_t session, uint32_t *dst, size_t count);
int cu_serve_stats(cu_serve s, cu_serve_report *out);
int cu_serve_session_times(cu_serve s, uint32_t session, double *open_s, double *admit_s);

#ifdef __cplusplus
}
#endif

#endif /* CHAOSULTRA_H */
