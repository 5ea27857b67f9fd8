/*
 * oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a faithful C++20 CPU restatement of the
 * hdRestir per-pixel hot path (RaycastPass -> IntegrationPass with PathIntegrator and
 * MIS / RIS direct lighting -> AccumulationPass). Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg load it, to check results against and to time the CPU
 * baseline; the product library (libhdrgpu.so) never links, loads or calls it.
 *
 * Inputs use the same POD descriptors as include/hdrgpu.h so both sides see identical
 * scenes, frames and pipeline settings.
 */
#ifndef HDR_ORACLE_H
#define HDR_ORACLE_H

#include "../include/hdrgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* hoist_light_set = 0 reproduces the reference exactly, including a fresh
 * UniformLightSampler (with its unordered_set of all lights) per path
 * (path_integrator.cpp:99); 1 builds it once per frame, with identical output. */
oracle_ctx *oracle_create(const hdr_scene_desc *scene, int32_t width, int32_t height,
                          int32_t hoist_light_set, int32_t threads);
void        oracle_destroy(oracle_ctx *ctx);
/* RenderPipeline::ClearPersistentBuffers: drop reservoirs and accumulation */
void        oracle_reset(oracle_ctx *ctx);
/* hdr_view_reset_accumulation: drop the accumulation, keep reservoirs and the previous frame's
 * records (reprojected temporal reuse, an extension) */
void        oracle_reset_accumulation(oracle_ctx *ctx);
/* Per pixel, the previous frame's pixel the last frame reprojected it to (-1: none). */
void        oracle_read_reprojection(oracle_ctx *ctx, int32_t *out);
/* firefly_in_place = 1: the CPU AccumulationPass (raster-sequential, in place);
 * 0: out-of-place 3x3 clamp on the unclamped frame (the reference's GPU shader). */
void        oracle_set_firefly_in_place(oracle_ctx *ctx, int32_t in_place);

/* Renders one frame restricted to the pixel rectangle [x0,x1) x [y0,y1) (pixel indices
 * and seeds are those of the full width x height image). color_out: width*height*4
 * floats (only the rectangle is written); debug_out: width*height records or NULL.
 * Returns the wall time of the integration pass in seconds. */
double oracle_render_frame(oracle_ctx *ctx, const hdr_frame_desc *frame, const hdr_pipeline_desc *pipe,
                           int32_t x0, int32_t y0, int32_t x1, int32_t y1, float *color_out,
                           hdr_pixel_debug *debug_out);

void oracle_read_reservoirs(oracle_ctx *ctx, hdr_reservoir *out);
/* AccumulationPass's Variance metrics over a pixel rectangle: out[3] = mean and max estimator
 * variance, mean luminance; 0 below 2 samples. lum_sum / lum_sum_sq (optional) get the per-pixel
 * luminance sums of the samples and of their squares. */
int32_t oracle_read_variance(oracle_ctx *ctx, int32_t x0, int32_t y0, int32_t x1, int32_t y1, double *out,
                             double *lum_sum, double *lum_sum_sq);
void oracle_read_aovs(oracle_ctx *ctx, float *depth, float *albedo, float *normal);
void oracle_read_counters(oracle_ctx *ctx, hdr_frame_counters *out);

/* Scene::IntersectScene on arbitrary rays; triangle index is within its mesh. */
int32_t oracle_dome_cdf(oracle_ctx *ctx, int32_t light, float *row_cdf, float *col_cdf, float *total_lum);
void oracle_intersect(oracle_ctx *ctx, const float *origins, const float *dirs, int32_t n,
                      int32_t *instance_out, int32_t *triangle_out, float *t_out);

/* Raw engine outputs / Rng::NextFloat values of std::mt19937_64(seed). */
void oracle_rng_raw(const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out);
void oracle_rng_float(const uint32_t *seeds, int32_t n, int32_t n_draws, float *out);

/* Unit-level exports used by the pin tests (declared in the tests' ctypes bindings):
 * oracle_reservoir_run, oracle_random_index, oracle_spectrum_roundtrip, oracle_spectrum_to_rgb,
 * oracle_ggx, oracle_power_heuristic, oracle_sample_bounce, oracle_sample_light,
 * oracle_light_index, oracle_closure, oracle_inverse4. */

/* Jittered and pixel-centre camera rays (GenerateCameraRay) for a whole frame:
 * origins/dirs are width*height*3 floats. */
void oracle_camera_rays(const hdr_frame_desc *frame, int32_t width, int32_t height, int32_t centre,
                        float *origins, float *dirs);

/* The passes after accumulation (post_process.cpp): DenoiserPass's prefilter (the chromaticity
 * blur; the reference's denoiser output without OIDN) when chroma_blur, then PostProcessPass's lens
 * flare and chromatic aberration (ca > 0), on width x height RGBA floats in place (alpha kept).
 * Returns -1 for an empty image. */
int32_t oracle_post_process(float *rgba, int32_t width, int32_t height, int32_t chroma_blur, int32_t lens_flare,
                            float ca);

#ifdef __cplusplus
}
#endif

#endif
