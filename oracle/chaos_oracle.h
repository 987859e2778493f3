/*
 * chaos_oracle.h — CPU oracle of the escape-time render path.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg as the checker; never by the product path.
 *
 * Each function restates one reference function op for op (citations are
 * reference-relative file:line) with FP contraction off; the planar-tile
 * extensions that have no reference counterpart (bailout radius, smooth value,
 * pixel-centre / jittered samples) say so.
 */
#ifndef CHAOS_ORACLE_H
#define CHAOS_ORACLE_H

#include <stdint.h>
#include "../include/chaosultra.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pixel_info_t (src/main/cuda/helpers.cuh:106-130): 16 bytes. */
typedef struct or_pixel_info {
    float value;
    float weight;
    uint8_t is_reused;
    uint8_t pad[3];
    float weight_of_new_samples;
} or_pixel_info;

/* computeFractal<Real> of mandelbrot.cu:11-25 (interior -> 0) and
 * julia.cu:5-17 (c = (Real)julia_c, interior -> maxIter). */
float or_mandelbrot_d(uint32_t max_iter, double cx, double cy);
float or_mandelbrot_f(uint32_t max_iter, float cx, float cy);
float or_julia_d(uint32_t max_iter, double zx, double zy, double julia_cx, double julia_cy);
float or_julia_f(uint32_t max_iter, float zx, float zy, double julia_cx, double julia_cy);

/* The same loop with the bailout constant 4 generalised to bailout_sq
 * (planar-tile extension).  Returns the iterations executed; *escaped = 1 when the loop
 * stopped on the bailout test, with *mag2 = that test's x*x+y*y. */
uint32_t or_escape_d(uint32_t max_iter, double zx, double zy, double cx, double cy, double bailout_sq,
                     double *mag2, int *escaped);
uint32_t or_escape_f(uint32_t max_iter, float zx, float zy, float cx, float cy, float bailout_sq,
                     float *mag2, int *escaped);
/* n + 1 - log2(ln(mag2) / ln(bailout_sq))  (smooth value, planar-tile extension) */
double or_smooth(uint32_t n, double mag2, double bailout_sq);

/* SplitMix64 finaliser and the per-sample jitter (planar-tile extension): sample j of a
 * tile takes output j of SplitMix64(seed); dx = high 32 bits * 2^-32,
 * dy = low 32 bits * 2^-32. */
uint64_t or_splitmix64(uint64_t z);
void or_jitter(uint64_t seed, uint64_t sample_id, double *dx, double *dy);
/* Sample position of (px, py, s) in a tile, FP64 / FP32 mapping. */
void or_tile_point_d(const cu_tile_desc *t, uint32_t px, uint32_t py, uint32_t s, double *x, double *y);
void or_tile_point_f(const cu_tile_desc *t, uint32_t px, uint32_t py, uint32_t s, float *x, float *y);

/* Rows [row0, row1) of a tile into the count / smooth planes (addressed with
 * the tile's out_offset / out_pitch; smooth may be NULL).  precision 32 or 64.
 * OpenMP over rows with `threads` threads (<= 0: all cores).  Returns the
 * iterations executed (an interior sample counts max_iter). */
uint64_t or_escape_tile_rows(const cu_tile_desc *t, int precision, uint32_t row0, uint32_t row1,
                             uint32_t *counts, double *smooth, int threads);

/* fractalRenderMain<Real> (fractalRendererGeneric.cu:168-181) over a whole
 * image, with sampleTheFractal (:85-155) run in lock step per 8x4 vote group
 * (getImageIndexes, :36-53).  precision 32/64; kind CU_FRACTAL_*. */
/* Iterations executed by computeFractal calls since the last call (render_main,
 * render_advanced, render_undersampled, the point queries); resets the count. */
uint64_t or_render_iterations(void);
void or_render_main(int precision, int kind, double julia_cx, double julia_cy, or_pixel_info *out,
                    int64_t out_pitch_bytes, uint32_t width, uint32_t height, double lbx, double lby,
                    double rtx, double rty, uint32_t max_iter, float max_ss, uint32_t flags);

/* fractalRenderAdvanced<Real> (fractalRendererGeneric.cu:307-368). */
void or_render_advanced(int precision, int kind, double julia_cx, double julia_cy,
                        or_pixel_info *out, int64_t out_pitch_bytes, uint32_t width, uint32_t height,
                        double lbx, double lby, double rtx, double rty, uint32_t max_iter,
                        float max_ss, uint32_t flags, double old_lbx, double old_lby, double old_rtx,
                        double old_rty, const or_pixel_info *in, int64_t in_pitch_bytes,
                        uint32_t focus_x, uint32_t focus_y);

/* getFoveationAdvisedSampleCount (fractalRendererGeneric.cu:226-257). */
void or_foveation(uint32_t px, uint32_t py, uint32_t focus_x, uint32_t focus_y, float max_ss,
                  float *advised, int *inside);

/* compose + colorize + colorizeSampleCount (fractalRendererGeneric.cu:455-475,
 * 67-77; mandelbrot.cu:27-35) into a linear RGBA8 image. */
void or_compose(const or_pixel_info *in, int64_t in_pitch_bytes, uint32_t *out, uint32_t out_pitch,
                uint32_t width, uint32_t height, const uint32_t *palette, uint32_t palette_length,
                float max_ss, int visualise_sample_count);

/* Palette compose of a smooth plane (colorize of mandelbrot.cu:27-35 applied to
 * round(smooth)). */
void or_compose_smooth(const double *smooth, uint32_t smooth_pitch, uint32_t *out,
                       uint32_t out_pitch, uint32_t width, uint32_t height,
                       const uint32_t *palette, uint32_t palette_length);

/* fractalRenderUnderSampled (fractalRendererGeneric.cu:478-503), float only. */
void or_render_undersampled(or_pixel_info *out, int64_t out_pitch_bytes, uint32_t width,
                            uint32_t height, float lbx, float lby, float rtx, float rty,
                            uint32_t max_iter, uint32_t level);

/* ImageHelpers.createDefaultColorPalette / fromBGRAtoRGBA (util/ImageHelpers.java:77-130). */
void or_default_palette(uint32_t *out /* 1536 */);
uint32_t or_argb_to_rgba(uint32_t argb);

/* ---- deep zoom (deep_oracle.cpp): FP32 perturbation, direct FP64 and
 * double-double oracles; pixel (px, py) of a W x H frame centred on C samples
 * C + ((px + 0.5 - W/2) * pw, (H/2 - (py + 0.5)) * ph). ---- */
void or_dd_add(double a_hi, double a_lo, double b_hi, double b_lo, double *hi, double *lo);
uint32_t or_reference_orbit_f(double cx_hi, double cx_lo, double cy_hi, double cy_lo,
                              uint32_t max_iter, double bailout_sq, float *orbit);
uint32_t or_reference_orbit_d(double cx_hi, double cx_lo, double cy_hi, double cy_lo,
                              uint32_t max_iter, double bailout_sq, double *orbit);
uint32_t or_perturb_pixel_f(const float *orbit, uint32_t M, float dcx, float dcy,
                            uint32_t max_iter, float bailout_sq, uint32_t flags, float glitch_eps,
                            float *mag_out, int *state);
uint32_t or_perturb_pixel_d(const double *orbit, uint32_t M, double dcx, double dcy,
                            uint32_t max_iter, double bailout_sq, uint32_t flags,
                            double glitch_eps, double *mag_out, int *state);
void or_deep_offset(uint32_t px, uint32_t py, uint32_t width, uint32_t height, double pixel_w,
                    double pixel_h, double *ox, double *oy);
uint32_t or_deep_direct_d(double cx, double cy, double ox, double oy, uint32_t max_iter,
                          double bailout_sq, double *mag_out, int *escaped);
uint32_t or_deep_direct_dd(double cx_hi, double cx_lo, double cy_hi, double cy_lo, double ox,
                           double oy, uint32_t max_iter, double bailout_sq, double *mag_out,
                           int *escaped);
uint64_t or_deep_pixels(int method, const uint32_t *px, const uint32_t *py, uint64_t n_pixels,
                        uint32_t width, uint32_t height, double cx_hi, double cx_lo, double cy_hi,
                        double cy_lo, double pixel_w, double pixel_h, uint32_t max_iter,
                        double bailout_sq, const void *orbit, uint32_t M, double ref_ox,
                        double ref_oy, uint32_t flags, float glitch_eps, uint32_t *counts,
                        double *mag, int32_t *state, int threads);

#ifdef __cplusplus
}
#endif

#endif
