/*
 * hdrgpu.h -- C-ABI of the MI355X (gfx950) per-pixel hot path of the hdRestir
 * spectral renderer: camera rays -> BVH closest-hit / shadow traversal ->
 * UsdPreviewSurface/GGX over 4 hero wavelengths -> MIS / RIS / ReSTIR direct
 * lighting -> spectral->RGB -> accumulation.
 *
 * Pure C, POD structs with fixed-width fields, no torch / HIP / USD types.
 * Every call returns an hdr_status and never throws; hdr_last_error() returns
 * a thread-local message for the last failing call on the calling thread.
 *
 * A context (and every scene / view created from it) is used from one host
 * thread at a time.  Device state (scene BVHs, light tables, per-view
 * reservoirs and accumulation) stays resident in HBM between calls.
 *
 * Reference interfaces replaced (paths relative to the hdRestir checkout):
 *   - RenderPass{Inputs,Outputs,_execute}            source/passes/render_pass.h:13-42
 *   - RaycastPass::_execute  (G-buffer, AOVs)        source/passes/raycast_pass.cpp:57-162
 *   - IntegrationPass::_execute (per-pixel Li)       source/passes/integration_pass.cpp:157-252
 *   - AccumulationPass / GpuAccumulationPass         source/passes/accumulation_pass.cpp:64-139,
 *                                                    source/gpu_passes/gpu_accumulation_pass.cpp:24-85
 *   - IDirectLightIntegrator::Li (MIS / RIS)         source/integrators/direct_light_integrator_interface.h:44-55
 *   - RisDirectLightIntegrator::MakeFactory(M, useReservoir, skipVisibility)
 *                                                    source/restir/ris_direct_light_integrator.h:73-74
 *   - PathTracePassSettings / RISPipelineSettings    source/integrators/path_integrator.h:12-19,
 *                                                    source/restir_pipeline/ris_pipeline.h:15-21
 *   - IScene (IntersectScene, GetMaterial, GetLights, GetLightAtHit, GetEnvironment)
 *                                                    source/sceneInterface/scene_interface.h:36-62
 *   - RenderPipeline::ClearPersistentBuffers         source/pipeline/render_pipeline.cpp:42-45
 */
#ifndef HDRGPU_H
#define HDRGPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDRGPU_ABI_VERSION 3u

typedef enum hdr_status {
    HDR_OK = 0,
    HDR_ERR_INVALID_ARGUMENT = 1,
    HDR_ERR_OUT_OF_MEMORY = 2,
    HDR_ERR_DEVICE = 3,
    HDR_ERR_UNSUPPORTED = 4,
    HDR_ERR_CANCELLED = 5,
    HDR_ERR_CAPACITY = 6
} hdr_status;

/* ---- scene description (host memory; copied to the device by hdr_scene_upload) ---- */

/* One triangle mesh subset = one BLAS (reference: MeshSubset, source/sceneInterface/mesh.h:19-25,
 * built by BVH::Build, source/renderContext/bvh.cpp:66-182). Attribute arrays are optional
 * (NULL / 0) and are either per-vertex (count == num_points) or face-varying
 * (count == 3 * num_triangles), as in BVH::Build. */
typedef struct hdr_mesh_desc {
    uint32_t       struct_size; /* sizeof(hdr_mesh_desc) */
    const float   *points;     /* xyz, num_points */
    int32_t        num_points;
    const int32_t *indices;    /* 3 per triangle */
    int32_t        num_triangles;
    const float   *normals;    /* xyz */
    int32_t        num_normals;
    const float   *uvs;        /* uv */
    int32_t        num_uvs;
    const float   *colors;     /* rgb (displayColor) */
    int32_t        num_colors;
} hdr_mesh_desc;

/* UsdPreviewSurface parameters (reference: PreviewSurfaceParams,
 * source/materials/preview_surface_params.h:12-52), textures excluded. */
typedef struct hdr_material_desc {
    uint32_t struct_size;  /* sizeof(hdr_material_desc) */
    float diffuse_color[3];
    float metallic;
    float roughness;
    float specular_color[3];
    float specular;
    float opacity;
    float ior;
    float transmission;
    float transmission_color[3];
    float transmission_depth;
    float emission_color[3];
    float emission;
    float coat;
    float coat_color[3];
    float coat_roughness;
    float coat_ior;
    float sheen;
    float sheen_color[3];
    float sheen_roughness;
    float subsurface;
    float subsurface_color[3];
    int32_t thin_walled;
} hdr_material_desc;

/* One mesh instance (reference: Scene::MeshInstance, source/scene/scene.h:112-122).
 * transform is a GfMatrix4f in row-vector convention (p' = p * M), row-major. */
typedef struct hdr_instance_desc {
    uint32_t struct_size;  /* sizeof(hdr_instance_desc) */
    float   transform[16];
    int32_t mesh;       /* index into hdr_scene_desc::meshes */
    int32_t material;   /* index into materials, or -1 for DefaultMaterial */
    int32_t light;      /* light this mesh prim is (Scene::GetLightAtHit), or -1 */
    int32_t _pad;
} hdr_instance_desc;

typedef enum hdr_light_type {
    HDR_LIGHT_DISTANT = 0,
    HDR_LIGHT_DOME = 1,
    HDR_LIGHT_RECT = 2,
    HDR_LIGHT_POINT = 3
} hdr_light_type;

/* One light (reference: LightParams source/sceneInterface/light_params.h:15-31 and the
 * LightType switch in source/scene/light_factory.cpp:11-25). transform is a GfMatrix4d,
 * row-vector convention, row-major. For HDR_LIGHT_DOME, dome_pixels holds
 * dome_width*dome_height linear RGB float texels, row 0 first. */
typedef struct hdr_light_desc {
    uint32_t     struct_size;  /* sizeof(hdr_light_desc) */
    int32_t      type;
    float        color[3];
    float        intensity;
    float        exposure;
    double       transform[16];
    float        width;
    float        height;
    float        shaping_cone_angle;     /* degrees, 180 = no shaping */
    float        shaping_cone_softness;
    const float *dome_pixels;
    int32_t      dome_width;
    int32_t      dome_height;
} hdr_light_desc;

typedef struct hdr_scene_desc {
    uint32_t                 struct_size;   /* sizeof(hdr_scene_desc) */
    const hdr_mesh_desc     *meshes;
    int32_t                  num_meshes;
    const hdr_material_desc *materials;
    int32_t                  num_materials;
    const hdr_instance_desc *instances;
    int32_t                  num_instances;
    const hdr_light_desc    *lights;        /* order == Scene::GetLights() order */
    int32_t                  num_lights;
    int32_t                  environment;   /* index of the dome light that is the environment, or -1 */
} hdr_scene_desc;

/* ---- per-frame inputs ---- */

/* reference: CameraParams, source/scene/camera_params.h:8-30 */
typedef struct hdr_camera_params {
    int32_t enable_dof;
    int32_t enable_exposure;
    float   focal_length;
    float   f_stop;
    float   focus_distance;
    int32_t bokeh_blades;
    float   lens_distortion;
    float   iso;
    float   shutter_speed;
} hdr_camera_params;

/* reference: RenderContext, source/renderContext/render_context.h:21-37 */
typedef struct hdr_frame_desc {
    uint32_t          struct_size;        /* sizeof(hdr_frame_desc) */
    double            view[16];           /* GfMatrix4d world->camera, row-vector convention */
    double            proj[16];           /* GfMatrix4d projection */
    int32_t           frame_index;        /* drives the per-pixel seeds */
    int32_t           has_camera_params;
    hdr_camera_params camera;
} hdr_frame_desc;

typedef enum hdr_direct_light_kind {
    HDR_DIRECT_MIS = 0,   /* PathTracer pipeline: MisDirectLightIntegrator */
    HDR_DIRECT_RIS = 1    /* RIS pipeline: RisDirectLightIntegrator */
} hdr_direct_light_kind;

/* reference: PathTracePassSettings + maxDepth (source/integrators/path_integrator.h:12-19),
 * RISPipelineSettings (source/restir_pipeline/ris_pipeline.h:15-21),
 * Denoiser::Config::EnableFireflyFilter (accumulation pass). */
typedef struct hdr_pipeline_desc {
    uint32_t struct_size;               /* sizeof(hdr_pipeline_desc) */
    int32_t  direct_light;              /* hdr_direct_light_kind */
    int32_t  candidate_count;           /* RIS M */
    int32_t  use_reservoir;
    int32_t  skip_visibility;
    int32_t  max_depth;
    int32_t  max_reflection_bounces;
    int32_t  max_refraction_bounces;
    int32_t  enable_subsurface;
    int32_t  render_ibl_background;
    float    roulette_aggressiveness;
    int32_t  enable_firefly_filter;
    /* Extension beyond the reference (0 = reference behaviour): after every pixel's temporal
     * merge, resample among its reservoir and those of spatial_count (<= 8) neighbours of this
     * frame picked in a disc of spatial_radius pixels (depth / normal similarity test), with
     * generalized balance-heuristic weights; the winner gets the shadow ray and is stored for
     * the next frame. At max_depth > 1 the reuse replaces the primary vertex's direct light;
     * the later vertices are rendered as without it and their radiance is added after it.
     * Supported for ReSTIR (use_reservoir) with skip_visibility, opaque materials and no mesh
     * light, at any depth, with or without an environment or emissive materials; otherwise
     * hdr_render_frame returns HDR_ERR_UNSUPPORTED. */
    int32_t  spatial_count;
    float    spatial_radius;
    /* Extension beyond the reference (0 = reference behaviour), with spatial reuse (spatial_count
     * 1 .. 7): reprojected temporal reuse under camera motion. The reference clears the reservoirs
     * on every camera move (renderer.cpp:401-406) and merges the same pixel's reservoir otherwise
     * (ris_direct_light_integrator.cpp:127-155). With this set, a frame whose view or projection
     * matrix differs from the view's previous frame skips that same-pixel merge (its sample belongs
     * to another shading point) and instead offers the previous frame's final reservoir at the
     * pixel where this pixel's primary hit was seen -- the hit projected with the previous camera,
     * depth within 10 % of the previous G-buffer's, geometric normals within 25 degrees -- as one
     * more domain of the spatial resampling, shifted to this pixel's vertex and weighted with the
     * generalized balance heuristic like the spatial neighbours. Frames with an unchanged camera
     * merge as the reference does, so their results are the same as without the flag. A caller
     * keeps the reservoirs on a camera move by clearing the accumulation with
     * hdr_view_reset_accumulation instead of hdr_view_reset. The history is the previous frame of
     * the view rendered with hdr_render_frame with this flag set; lens distortion turns the
     * reprojection off (the same-pixel merge is still skipped on a move). Not supported by
     * hdr_render_batch (HDR_ERR_UNSUPPORTED). */
    int32_t  temporal_reprojection;
} hdr_pipeline_desc;

/* Per-pixel persistent reservoir (reference: WeightedReservoir<RISLightCandidate>,
 * source/restir/reservoir.h:69-161 + RISLightCandidate ris_direct_light_integrator.h:26-43),
 * stored as an 80-byte POD. */
#define HDR_RES_HAS_SAMPLE      0x01u   /* ChosenSample engaged */
#define HDR_RES_HAS_CANDIDATE   0x02u   /* ChosenSample->Candidate engaged */
#define HDR_RES_VIS_TESTED      0x04u   /* VisibilityTested */
#define HDR_RES_PDF_SOLID_ANGLE 0x08u   /* Ls.Pdf.space == SolidAngle */
#define HDR_RES_LIGHT_DELTA     0x10u   /* Light->IsDeltaLight() */
typedef struct hdr_reservoir {
    float    dir[3];
    float    dist;
    float    color[3];
    float    pdf;
    float    light_normal[3];
    float    target;            /* targetFunction p-hat (always float-representable) */
    float    throughput[4];
    float    wsum;
    int32_t  count;             /* CandidateCount (M) */
    float    w;
    uint32_t flags_light;       /* HDR_RES_* flags in bits 24..31, light index + 1 in bits 0..23 */
} hdr_reservoir;

/* Per-pixel parity record (optional output of hdr_render_frame): the same record is
 * written by the CPU oracle so integer streams can be compared bit-exactly. */
typedef struct hdr_pixel_debug {
    uint32_t draws;          /* RNG draws consumed by the pixel this frame */
    uint32_t light_hash;     /* FNV-1a over every uniform light-selection index, in order */
    uint32_t select_hash;    /* FNV-1a over every RIS / reservoir selection result, in order */
    int32_t  n_light_picks;  /* number of light-selection draws */
    int32_t  prim_instance;  /* G-buffer instance id (PrimId), -1 on miss */
    int32_t  prim_triangle;  /* G-buffer triangle index within its mesh, -1 on miss */
    float    prim_t;         /* G-buffer hit distance */
    float    sample_rgb[3];  /* this frame's Color before accumulation */
} hdr_pixel_debug;

/* ---- opaque handles ---- */
typedef struct hdr_context hdr_context;
typedef struct hdr_scene   hdr_scene;
typedef struct hdr_view    hdr_view;

const char *hdr_last_error(void);
uint32_t    hdr_abi_version(void);

hdr_status hdr_context_create(int32_t device, hdr_context **out);
hdr_status hdr_context_destroy(hdr_context *ctx);
/* cooperative cancellation (reference: IRenderJob::IsCancelled, source/rendererInterface/render_job.h:9-35):
 * a set flag makes the next hdr_render_frame return HDR_ERR_CANCELLED before any launch. */
hdr_status hdr_context_set_cancel(hdr_context *ctx, int32_t cancelled);

/* Builds BLAS/TLAS (midpoint-split, as Scene::BuildRenderState / BVH::Build), light and
 * material tables on the host and uploads them (reference: Scene::BuildRenderState,
 * source/scene/scene.cpp:118-271); a dome light's importance-sampling CDFs
 * (DomeLight::Prepare, source/scene/dome_light.cpp:49-76) are built on the device from the
 * uploaded texels. The scene stays resident until hdr_scene_destroy. */
hdr_status hdr_scene_upload(hdr_context *ctx, const hdr_scene_desc *desc, hdr_scene **out);
hdr_status hdr_scene_destroy(hdr_scene *scene);
/* Scene change between frames (reference: the Hydra prim syncs (HdRestirMesh / HdRestirInstancer
 * transforms, HdRestirMaterial, HdRestirLight in source/hydra/) followed by
 * Scene::BuildRenderState, source/scene/scene.cpp:118-271, which rebuilds the instance list, the
 * TLAS and the lights every frame and keeps each mesh's BVH). The instances (transforms, bound
 * materials and lights, their number), materials and lights of `desc` replace the scene's; the
 * meshes and their BLAS stay as uploaded: desc->num_meshes must equal the uploaded count and
 * desc->meshes is not read. The TLAS is rebuilt, not refitted, so its tree -- and the choice among
 * equal-distance hits -- is the reference's for the new transforms. A dome light whose texels are
 * unchanged keeps its device CDFs; others are rebuilt. The call waits for the device's queued work
 * first; no frame of the scene may be issued from another thread during it. The scene gets a new
 * identity (a view's G-buffer of the old scene is not reused). As in the reference, whose render
 * thread restarts on a scene change, accumulating views are reset by the caller (hdr_view_reset,
 * frame index 0). HDR_ERR_INVALID_ARGUMENT leaves the scene unchanged; after HDR_ERR_DEVICE or
 * HDR_ERR_OUT_OF_MEMORY the scene can only be destroyed. */
hdr_status hdr_scene_update(hdr_scene *scene, const hdr_scene_desc *desc);
/* device bytes held by the scene */
hdr_status hdr_scene_device_bytes(const hdr_scene *scene, uint64_t *out);
/* Reads back the dome CDFs of light `light` as built on the device (DomeLight::Prepare,
 * dome_light.cpp:49-76): row_cdf h+1 floats, col_cdf h*(w+1) floats (row-major), the total
 * sin-weighted luminance. HDR_ERR_INVALID_ARGUMENT for a light without texels. */
hdr_status hdr_scene_read_dome(const hdr_scene *scene, int32_t light, float *row_cdf, float *col_cdf,
                               float *total_lum);

/* A view owns the device-resident persistent buffers of one pipeline instance:
 * reservoirs (__ris_reservoir) and accumulation (__accum_color), plus the G-buffer. */
hdr_status hdr_view_create(hdr_context *ctx, int32_t width, int32_t height, hdr_view **out);
hdr_status hdr_view_destroy(hdr_view *view);
/* == RenderPipeline::ClearPersistentBuffers: zero reservoirs and accumulation */
hdr_status hdr_view_reset(hdr_view *view);
/* Zero the accumulation, the Variance sums and the sample count, and keep the reservoirs and the
 * previous frame's records: what a caller does on a camera move with temporal_reprojection set
 * (hdr_pipeline_desc). An addition with no counterpart in the reference, whose camera moves clear
 * both (RenderPipeline::ClearPersistentBuffers via renderer.cpp:401-406). */
hdr_status hdr_view_reset_accumulation(hdr_view *view);
hdr_status hdr_view_device_bytes(const hdr_view *view, uint64_t *out);

/* Device bytes a view of width x height holds to render frames of `pipeline` on `scene`: the buffers
 * hdr_view_create allocates plus those hdr_render_frame allocates for this pipeline (RNG spill slots,
 * the deferred or wavefront path's buffers, spatial reuse's, the firefly filter's frame). Validates
 * the pipeline on the scene as hdr_render_frame does (HDR_ERR_INVALID_ARGUMENT, HDR_ERR_UNSUPPORTED)
 * and makes no device call. No counterpart in the reference (its buffers live in host memory); used
 * for admission by memory footprint. */
hdr_status hdr_view_plan_bytes(const hdr_scene *scene, int32_t width, int32_t height,
                               const hdr_pipeline_desc *pipeline, uint64_t *bytes);

/* One frame of RaycastPass -> {RIS,}PathTracePass -> AccumulationPass on the device.
 * Asynchronous with respect to the host unless an output below is requested.
 * color_out (optional, host, width*height*4 floats): accumulated "Color" (RGBA32F).
 * debug_out (optional, host, width*height records): parity records. */
hdr_status hdr_render_frame(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                            const hdr_pipeline_desc *pipeline, float *color_out,
                            hdr_pixel_debug *debug_out);

/* Copies of the view's persistent state (host buffers of width*height elements). */
hdr_status hdr_view_read_color(hdr_view *view, float *rgba_out);
hdr_status hdr_view_read_reservoirs(hdr_view *view, hdr_reservoir *out);
/* RaycastPass AOVs: Depth (1 float), Albedo (3), Normal (3) per pixel; misses write 1 / 0 / 0. */
hdr_status hdr_view_read_aovs(hdr_view *view, float *depth, float *albedo, float *normal);
/* Device pointer of the accumulated colour (RGBA32F) for zero-copy consumers. */
hdr_status hdr_view_color_device_ptr(hdr_view *view, void **out);

/* AccumulationPass's Variance metrics (reference: accumulation_pass.cpp:84-137, compiled there with
 * METRICS_ENABLED). With the metrics on, every accumulated frame also adds its sample's luminance
 * (0.2126 R + 0.7152 G + 0.0722 B in float, after the firefly clamp) and its square to per-pixel
 * sums in double (16 B per pixel; cleared by hdr_view_reset). Turning them on clears the sums;
 * turning them off frees them. */
hdr_status hdr_view_set_metrics(hdr_view *view, int32_t enable);
typedef struct hdr_variance_stats {
    uint32_t struct_size;            /* sizeof(hdr_variance_stats) */
    int32_t  sample_count;           /* frame_index + 1 of the view's last frame */
    int32_t  valid;                  /* 0 below 2 samples (the reference computes nothing there) */
    int32_t  _pad;
    double   mean_estimator_variance; /* mean over pixels of max(0, (S2 - n mean^2) / (n - 1)) / n */
    double   max_estimator_variance;
    double   mean_luminance;          /* mean over pixels of S / n */
} hdr_variance_stats;
/* The statistics the reference logs and stores as its Variance output (VarianceStats {mean, max}),
 * over the whole view, reduced on the device in a fixed order; lum_sum / lum_sum_sq (optional, host,
 * width*height doubles) receive the per-pixel sums. HDR_ERR_INVALID_ARGUMENT when the metrics are
 * off. Waits for the view's frames. */
hdr_status hdr_view_read_variance(hdr_view *view, hdr_variance_stats *out, double *lum_sum, double *lum_sum_sq);

/* ---- resolve, upscale and the progressive render loop ---- */

/* Outputs of a view (reference: the AOV output specs Color / Depth / Albedo / Normal,
 * source/passes/output_names.h) and the render-buffer component formats the reference writes
 * (HdFormatFloat32, Float16, UNorm8, SNorm8, Int32; hd_restir_render_buffer.cpp:46-74). */
enum { HDR_AOV_COLOR = 0, HDR_AOV_DEPTH = 1, HDR_AOV_ALBEDO = 2, HDR_AOV_NORMAL = 3 };
enum { HDR_FORMAT_FLOAT32 = 0, HDR_FORMAT_FLOAT16 = 1, HDR_FORMAT_UNORM8 = 2, HDR_FORMAT_SNORM8 = 3,
       HDR_FORMAT_INT32 = 4 };

/* The size a window renders at for a resolution level (reference: CameraFrame::RenderedWidth /
 * RenderedHeight, source/renderContext/camera_frame.h:34-41,72-76): ceil(window / 2^level) with
 * the level clamped to 0 .. 4, at least 1. Create the view at this size; hdr_view_resolve upscales
 * to the window. Host-only. */
hdr_status hdr_rendered_size(int32_t window_width, int32_t window_height, int32_t level, int32_t *width,
                             int32_t *height);

/* Renderer::ResolveTargets -> HdRestirRenderBuffer::CopyFromFrameBuffer (renderer.cpp:536-563,
 * hd_restir_render_buffer.cpp:113-145) with UpscalePass (upscale_pass.cpp:17-94) on the device:
 * output `aov` of the view, nearest-texel resampled from the view's size to window_width x
 * window_height (equal sizes copy), converted to `components` (1-4) components of `format` per pixel
 * as _WriteOutput does (missing components 0; Float16 rounded to nearest even; UNorm8 / SNorm8
 * clamp(v * 255 | 127) truncated; Int32 truncated; NaN converts to 0 in the integer formats), and
 * copied to dst (host, window_width * window_height * components elements, rows packed). Waits for
 * the view's frames. Depth / albedo / normal are enabled by their first request, as with
 * hdr_view_read_aovs, and hold zeros until the next frame. */
hdr_status hdr_view_resolve(hdr_view *view, int32_t aov, int32_t window_width, int32_t window_height,
                            int32_t format, int32_t components, void *dst);

/* ---- after accumulation: the denoiser's prefilter and the post-process ---- */

/* Denoiser::Config (denoiser.h:14-19) and PostProcess::Config (post_process.h:13-17). */
typedef struct hdr_post_desc {
    uint32_t struct_size;       /* sizeof(hdr_post_desc) */
    int32_t chromaticity_blur;  /* EnableChromaticityBlur: the denoiser's 3x3 chroma average */
    int32_t oidn;               /* EnableDenoiser's OIDN network: not part of this library, must be 0 */
    int32_t lens_flare;         /* EnableLensFlare */
    float chromatic_aberration; /* ChromaticAberration: channel shift in pixels at the image corners (> 0: on,
                                   at most 1e6) */
} hdr_post_desc;

/* DenoiserPass then PostProcessPass on the view's Color output, in place and in the order of the
 * reference's PathTracerPost pipeline (path_tracer_pipeline.h:224-229): the chromaticity blur
 * (denoiser_pass.cpp:13-38, denoiser.cpp:54-94,154-201 -- the reference denoiser's whole output when
 * it is built without OIDN), then the lens flare and the chromatic aberration
 * (post_process_pass.cpp:10-21, post_process.cpp:23-160). Colour channels change, alpha is kept.
 * Color holds the mean of the view's accumulated frames; the accumulation is not touched, so the
 * next frame writes the plain mean again (AccumulationPass rewrites Color from its persistent sum,
 * accumulation_pass.cpp:102-103). Queued on the view's stream; hdr_view_read_color /
 * hdr_view_resolve read the result. HDR_ERR_UNSUPPORTED when oidn is set. */
hdr_status hdr_view_post_process(hdr_view *view, const hdr_post_desc *desc);

/* Accumulated samples of the view: frame_index + 1 of the last frame issued to it, 0 after
 * hdr_view_reset (reference: Camera::FrameCount, renderer.cpp:401-406,477). */
hdr_status hdr_view_sample_count(const hdr_view *view, int32_t *samples);

/* The render thread's loop (reference: _RenderCallback, hd_restir_render_delegate.cpp:45-54, with
 * Renderer::IsConverged, renderer.cpp:565-569, and RendererPipelineState::IsConverged,
 * renderer_pipeline_state.cpp:184-190): while the view holds fewer than target_samples samples and
 * the context is not cancelled, renders `frame` with frame_index = the view's sample count, at most
 * max_frames frames per call (so a caller can resolve between calls, as Hydra resolves its render
 * buffers while the thread runs). *converged = sample count >= target_samples. HDR_ERR_CANCELLED
 * when the context's cancel flag stops it. */
hdr_status hdr_render_progressive(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                                  const hdr_pipeline_desc *pipeline, int32_t target_samples, int32_t max_frames,
                                  int32_t *converged);
/* Frames frame->frame_index .. frame->frame_index + count - 1 of `frame`'s camera, with the same
 * results as `count` hdr_render_frame calls (accumulation, Variance sums and sample count; no
 * colour copy). An addition with no counterpart in the reference, which renders one frame per
 * Renderer::Render (renderer.cpp:390-478): frames that keep no state from one to the next -- MIS,
 * or RIS without the reservoir; no spatial reuse, no firefly filter, no depth of field -- and run
 * on the integration megakernel or the lean path tracer without RNG spill slots share launches
 * (up to 64 frames, with their samples staged in at most 256 MiB and accumulated in frame order);
 * other frames are rendered one by one. hdr_render_progressive renders its frames this way. */
hdr_status hdr_render_frames(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                             const hdr_pipeline_desc *pipeline, int32_t count);
/* Blocks until all work queued on the view's stream has finished. */
hdr_status hdr_view_synchronize(hdr_view *view);
/* Device time (ms) of the last hdr_render_frame's kernels: G-buffer (RaycastPass), integration,
 * and the whole frame (G-buffer through accumulation). */
hdr_status hdr_view_last_timing(hdr_view *view, float *gbuffer_ms, float *integrate_ms, float *frame_ms);

/* Kernel stages of one frame, timed by HIP events on the view's stream. */
enum {
    HDR_STAGE_GBUFFER = 0,    /* primary rays -> G-buffer (RaycastPass)                           */
    HDR_STAGE_INTEGRATE = 1,  /* integration megakernel, or the deferred path's shading kernel    */
    HDR_STAGE_SHADOW = 2,     /* deferred path: any-hit shadow rays of the winners               */
    HDR_STAGE_ACCUMULATE = 3, /* deferred path: visibility pick + accumulation                   */
    HDR_STAGE_FIREFLY = 4,    /* firefly filter + accumulation (enable_firefly_filter)           */
    HDR_STAGE_SPATIAL = 5,    /* spatial reuse: neighbour resampling + colour (on the wavefront
                                 path this launch lies inside the HDR_STAGE_INTEGRATE span)      */
    HDR_STAGE_COUNT = 6
};
/* Per-stage device time (ms) summed over the frames rendered since the previous call, and the
 * number of launches of each stage; clears the sums. Timing is read back lazily, so frames are
 * never serialised by it. Plays the role of the reference's per-pass timing metrics
 * (core/metrics_on_buffers.h, RESTIR_BUFFER_METRIC_SCOPE). */
hdr_status hdr_view_read_stage_times(hdr_view *view, double *stage_ms, int32_t *launches);

/* ---- batches of views ----
 * No counterpart in the reference, which renders one view per Renderer::Render call
 * (renderer.cpp:390-478). A batch renders one frame for each of up to 255 views in merged
 * launches: one G-buffer launch over every view's 8x8 tiles, one shading launch per kernel path
 * (deferred shading, spatial reuse, integration megakernel, inlined path tracer; the wavefront
 * path's per-vertex shading and traversal launches), one shadow launch over the combined queue and
 * one accumulation launch. Each view gets exactly the frame hdr_render_frame would
 * give it (same streams of random numbers, same results); views may differ in scene, size,
 * pipeline and frame. Not available with counting views or the firefly filter
 * (HDR_ERR_UNSUPPORTED). A view's later calls (reads, resets, single frames) are ordered after
 * the batch. A batch is used from one host thread at a time. */
typedef struct hdr_batch hdr_batch;
hdr_status hdr_batch_create(hdr_context *ctx, hdr_batch **out);
hdr_status hdr_batch_destroy(hdr_batch *batch);
hdr_status hdr_render_batch(hdr_batch *batch, int32_t n, hdr_view *const *views, const hdr_scene *const *scenes,
                            const hdr_frame_desc *frames, const hdr_pipeline_desc *pipes);
/* Blocks until the batch's queued steps have finished. */
hdr_status hdr_batch_synchronize(hdr_batch *batch);
/* Device time (ms) of the last step, from its first launch to its last. */
hdr_status hdr_batch_last_timing(hdr_batch *batch, float *step_ms);
/* Steps queued so far (numbered 1, 2, ...) and how many of them have finished on the device,
 * without waiting. */
hdr_status hdr_batch_poll(hdr_batch *batch, uint64_t *issued, uint64_t *completed);

/* ---- launch shape and autotuning ----
 * No counterpart in the reference (its passes run on CPU threads). The queue-driven launches --
 * the wavefront path's shading and traversal launches and the deferred path's shadow launch --
 * loop over their queue with a fixed number of 64-lane waves; the traversal and wavefront
 * shading kernels exist in variants compiled for a minimum number of waves per SIMD (fewer
 * registers, more waves in flight). A launch shape never changes a result: it only moves paths
 * between waves. Batch steps are captured into a hipGraph per argument slot and launch shape and
 * replayed (use_graphs, default 1). */
typedef struct hdr_launch_config {
    uint32_t struct_size;       /* sizeof(hdr_launch_config) */
    int32_t  queue_waves;       /* waves per queue-driven launch: 0 = one per 64 queue slots, at most
                                   32768 (the persistent shadow launch: 4 per SIMD of the device);
                                   else 64 .. 65536 */
    int32_t  trace_occupancy;   /* traversal kernels: 4, 6, 8 = at least that many waves per SIMD, 1 = the
                                   compiler's register budget; 0 = 8 */
    int32_t  shade_occupancy;   /* wavefront shading kernels: 3 = at least 3 waves per SIMD, 1 = the
                                   compiler's register budget; 0 = 3 */
    int32_t  use_graphs;        /* batches: 1 = replay captured graphs, 0 = launch every kernel */
    int32_t  lds_stack;         /* traversal kernels (G-buffer, shadow, wavefront traversal): 1 = the top
                                   4 TLAS / 8 BLAS stack entries of each lane in LDS, 0 = private memory
                                   (shadow rays then run on persistent lanes that take the next query as
                                   their ray is decided) */
    int32_t  _reserved[2];
} hdr_launch_config;
hdr_status hdr_view_set_launch_config(hdr_view *view, const hdr_launch_config *cfg);
hdr_status hdr_batch_set_launch_config(hdr_batch *batch, const hdr_launch_config *cfg);
hdr_status hdr_batch_get_launch_config(hdr_batch *batch, hdr_launch_config *cfg);
/* Autotuner: renders `reps` + 4 steps of the given views per candidate launch shape (stack home,
 * queue waves, traversal occupancy, shading occupancy: a coordinate search), keeps the fastest in the
 * batch and reports it. The views' reservoirs and accumulation advance like any rendered frames
 * (reset them before the frames that count). reps: 1 .. 64. */
hdr_status hdr_batch_autotune(hdr_batch *batch, int32_t n, hdr_view *const *views, const hdr_scene *const *scenes,
                              const hdr_frame_desc *frames, const hdr_pipeline_desc *pipes, int32_t reps,
                              hdr_launch_config *chosen, float *best_step_ms);

/* ---- serving ----
 * No counterpart in the reference, whose renderer serves one Hydra render pass at a time
 * (renderer.cpp:390-478). A server keeps scenes resident (the caller's hdr_scene handles), gives
 * each admitted request its own view (reservoirs, accumulation) on the device, admits queued
 * requests first-in first-out while their views fit hbm_budget_bytes and fewer than max_active are
 * running, and renders either
 *   batched = 1: continuous batching -- every step renders one frame of every running request in
 *                one merged launch set (hdr_render_batch); requests join when memory and a slot free
 *                up and leave when their last frame is done; steps run ahead on the device and a
 *                request is retired when the step that carries its last frame has finished (the
 *                host waits only when nothing is left to queue); or
 *   batched = 0: the plain per-request path -- one request at a time, all its frames through
 *                hdr_render_frame.
 * Admission takes queued requests in arrival order; one whose view does not fit the free memory is
 * passed over by later ones that fit, at most 8 times, after which admission waits for it.
 * A request renders frames frame.frame_index .. + spp - 1 of its own view, so its image is the
 * same alone and in any mix. Finished views return to a pool (reused for the same size, freed
 * oldest-first when memory is needed). The caller steps the server (hdr_server_step), or with
 * own_thread the server runs on a host thread of its own and the caller only submits and collects
 * (hdr_server_wait_result). The calls are serialised by the server's lock. */
typedef struct hdr_server hdr_server;
typedef struct hdr_server_desc {
    uint32_t struct_size;
    int32_t  max_active;        /* requests on the device at once, 1..255 (1 when batched = 0) */
    uint64_t hbm_budget_bytes;  /* device memory the server's views may hold */
    int32_t  batched;
    int32_t  own_thread;        /* 1: the server admits and renders on a host thread of its own (the caller
                                   submits and collects; hdr_server_step is then HDR_ERR_UNSUPPORTED) */
} hdr_server_desc;
typedef struct hdr_request_desc {
    uint32_t struct_size;
    int32_t  width, height;
    int32_t  spp;               /* frames to accumulate */
    const hdr_scene *scene;
    hdr_frame_desc frame;       /* camera; frame_index of the first frame */
    hdr_pipeline_desc pipe;
} hdr_request_desc;
typedef struct hdr_request_result {
    uint64_t id;
    int32_t  width, height, spp;
    int32_t  status;            /* hdr_status of the request: HDR_OK, or the error its frames failed with */
    double   queued_ms;         /* submission -> admission */
    double   latency_ms;        /* submission -> last frame finished */
} hdr_request_result;
typedef struct hdr_server_stats {
    int32_t  queued, active, finished, pooled_views;
    uint64_t hbm_in_use;
    uint64_t steps;
} hdr_server_stats;
hdr_status hdr_server_create(hdr_context *ctx, const hdr_server_desc *desc, hdr_server **out);
hdr_status hdr_server_destroy(hdr_server *server);
/* Validates the request's pipeline on its scene (hdr_view_plan_bytes: HDR_ERR_INVALID_ARGUMENT,
 * HDR_ERR_UNSUPPORTED, as hdr_render_frame would) and queues it; HDR_ERR_OUT_OF_MEMORY when its view
 * alone exceeds the budget. */
hdr_status hdr_server_submit(hdr_server *server, const hdr_request_desc *req, uint64_t *id);
/* Admits what fits, renders (one merged frame for every running request, or one whole request) and
 * retires the requests the device has finished (*n_finished of them). A request whose frames fail
 * finishes with that status in its result; the others go on. */
hdr_status hdr_server_step(hdr_server *server, int32_t *n_finished);
/* The oldest finished request not yet taken (*available = 0 if none). */
hdr_status hdr_server_peek_result(hdr_server *server, hdr_request_result *out, int32_t *available);
/* As hdr_server_peek_result, waiting up to timeout_ms for a result when the server runs its own
 * thread; returns the worker's error if it stopped on one. */
hdr_status hdr_server_wait_result(hdr_server *server, int32_t timeout_ms, hdr_request_result *out,
                                  int32_t *available);
/* Copies a finished request's image (RGBA32F, width*height*4; rgba may be null) and releases it. */
hdr_status hdr_server_take_result(hdr_server *server, uint64_t id, float *rgba);
hdr_status hdr_server_get_stats(hdr_server *server, hdr_server_stats *out);

/* ---- parity / measurement entry points ---- */

/* Closest hit for n world-space rays (reference: Scene::IntersectScene, scene.cpp:515-535).
 * Outputs per ray: instance id (-1 miss), triangle index within its mesh, t. */
hdr_status hdr_scene_intersect(hdr_scene *scene, const float *origins, const float *dirs, int32_t n,
                               int32_t *instance_out, int32_t *triangle_out, float *t_out);

/* First n_draws raw 64-bit outputs of mt19937_64(seed) for each of n seeds, computed on the
 * device by the same cursor engine the integrator uses (reference: Rng, source/core/rng.h:10-56). */
hdr_status hdr_rng_raw(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out);
/* Same, mapped through uniform_real_distribution<float> (Rng::NextFloat). */
hdr_status hdr_rng_float(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, float *out);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc): hdr_rng_raw on device buffers.
 * seeds (n uint32) and out (n * n_draws uint64, row per seed) are device pointers; the launch is
 * queued on `stream` (a hipStream_t, NULL = the null stream) and the call returns without waiting.
 * Streams drawn past output 311 keep their state in a static device scratch of the library, 32 MiB
 * per device (312 words per seed): with n_draws > 311, n may be at most 13443, else
 * HDR_ERR_CAPACITY before any launch. Allocates nothing; the scratch is set up by
 * hdr_context_create. Calls that overlap on one device share that scratch and race. */
hdr_status hdr_rng_raw_noalloc(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out,
                               void *stream);

/* Work counters of the last frame (rays by class, BVH node visits, triangle tests), filled
 * only when counting is enabled on the view (costs atomics; off for timing runs). */
typedef struct hdr_frame_counters {
    uint64_t primary_rays;
    uint64_t bounce_rays;
    uint64_t shadow_rays;
    uint64_t candidate_rays;        /* BSDF-candidate rays actually traced */
    uint64_t candidate_rays_skipped;/* BSDF-candidate rays proven unable to contribute */
    uint64_t node_visits;
    uint64_t triangle_tests;
    uint64_t instance_visits;
    uint64_t rng_draws;
    uint64_t primary_node_visits;   /* node visits / triangle tests of the G-buffer (RaycastPass) rays */
    uint64_t primary_triangle_tests;
} hdr_frame_counters;
hdr_status hdr_view_set_counting(hdr_view *view, int32_t enable);
hdr_status hdr_view_read_counters(hdr_view *view, hdr_frame_counters *out);

#ifdef __cplusplus
}
#endif

#endif /* HDRGPU_H */
