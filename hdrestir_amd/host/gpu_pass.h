// gpu_pass.h -- C++20 adapter over include/hdrgpu.h with the shape of the renderer's pass API.
//
// The reference's pixel loop lives behind RenderPass{Inputs, Outputs, _execute}
// (source/passes/render_pass.h:13-42): RaycastPass produces "GBuffer" (+ Depth / Albedo / Normal),
// {RIS,}PathTracePass (IntegrationPass) turns it into "Color", AccumulationPass averages "Color"
// over frames (source/restir_pipeline/ris_pipeline.h:23-46). GpuPathTracePass replaces those three
// passes with one device-resident pass: same output names, same settings
// (PathTracePassSettings + RISPipelineSettings), ClearPersistentBuffers semantics, and C++
// exceptions for failures the way DBG_ASSERT / GSL Expects report them in the reference.
// No USD types: the host fills hdr_scene_desc / hdr_frame_desc from its own scene state.
#pragma once

#include <array>
#include <cstring>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/hdrgpu.h"

namespace hdrgpu {

class Error : public std::runtime_error {
  public:
    Error(hdr_status s, const std::string &what) : std::runtime_error(what), status(s) {}
    hdr_status status;
};

inline void Check(hdr_status s, std::string_view call) {
    if (s != HDR_OK) throw Error(s, std::string(call) + ": " + hdr_last_error());
}

// Output names (source/passes/output_names.h).
inline constexpr std::string_view kColorOutputName{"Color"};
inline constexpr std::string_view kDepthOutputName{"Depth"};
inline constexpr std::string_view kAlbedoOutputName{"Albedo"};
inline constexpr std::string_view kNormalOutputName{"Normal"};

// PathTracePassSettings (source/integrators/path_integrator.h:12-19) + maxDepth.
struct PathTracePassSettings {
    bool EnableSubsurface{true};
    int MaxReflectionBounces{8};
    int MaxRefractionBounces{8};
    bool RenderIblBackground{true};
    float RouletteAggressiveness{1.0f};
};

// Denoiser::Config (source/passes/denoiser.h:14-19; its firefly filter is AccumulationPass's, in
// PipelineSettings) and PostProcess::Config (source/passes/post_process.h:13-17). EnableDenoiser asks
// for the OIDN network, which the library does not carry (Execute then throws HDR_ERR_UNSUPPORTED).
struct DenoiserConfig {
    bool EnableDenoiser{false};
    bool EnableChromaticityBlur{false};
};
struct PostProcessConfig {
    bool EnableLensFlare{false};
    float ChromaticAberration{0.0f};
};

// Pipeline selection: PathTracer (MIS direct light) or RIS (source/restir_pipeline/ris_pipeline.h:15-21).
// With Denoiser / PostProcess set, the passes after accumulation run as in the reference's
// PathTracerPost pipeline (source/pipeline/path_tracer_pipeline.h:224-229).
struct PipelineSettings {
    enum class Kind { PathTracer, RIS } kind{Kind::RIS};
    PathTracePassSettings PathTrace{};
    int MaxDepth{32};
    int CandidateCount{16};
    bool UseReservoir{true};
    bool SkipVisibility{false};
    bool EnableFireflyFilter{false};
    DenoiserConfig Denoiser{};
    PostProcessConfig PostProcess{};
    // Extensions beyond the reference (hdr_pipeline_desc): spatial reuse among SpatialCount
    // neighbours within SpatialRadius pixels, and reprojected temporal reuse under camera motion.
    int SpatialCount{0};
    float SpatialRadius{30.0f};
    bool TemporalReprojection{false};

    // DenoiserPass / PostProcessPass run when one of their switches is on (denoiser_pass.cpp:17-20,
    // post_process_pass.cpp:14-17).
    [[nodiscard]] bool HasPostPasses() const {
        return Denoiser.EnableDenoiser || Denoiser.EnableChromaticityBlur || PostProcess.EnableLensFlare ||
               PostProcess.ChromaticAberration > 0.0f;
    }
    [[nodiscard]] hdr_post_desc PostDesc() const {
        hdr_post_desc d{};
        d.struct_size = sizeof(d);
        d.chromaticity_blur = Denoiser.EnableChromaticityBlur;
        d.oidn = Denoiser.EnableDenoiser;
        d.lens_flare = PostProcess.EnableLensFlare;
        d.chromatic_aberration = PostProcess.ChromaticAberration;
        return d;
    }

    [[nodiscard]] hdr_pipeline_desc Desc() const {
        hdr_pipeline_desc d{};
        d.struct_size = sizeof(d);
        d.direct_light = kind == Kind::RIS ? HDR_DIRECT_RIS : HDR_DIRECT_MIS;
        d.candidate_count = CandidateCount;
        d.use_reservoir = UseReservoir;
        d.skip_visibility = SkipVisibility;
        d.max_depth = MaxDepth;
        d.max_reflection_bounces = PathTrace.MaxReflectionBounces;
        d.max_refraction_bounces = PathTrace.MaxRefractionBounces;
        d.enable_subsurface = PathTrace.EnableSubsurface;
        d.render_ibl_background = PathTrace.RenderIblBackground;
        d.roulette_aggressiveness = PathTrace.RouletteAggressiveness;
        d.enable_firefly_filter = EnableFireflyFilter;
        d.spatial_count = SpatialCount;
        d.spatial_radius = SpatialRadius;
        d.temporal_reprojection = TemporalReprojection;
        return d;
    }
};

// The frame inputs IntegrationPass reads from RenderContext (source/renderContext/render_context.h:21-37).
struct RenderContext {
    std::array<double, 16> viewMatrix{};   // GfMatrix4d, row-vector convention
    std::array<double, 16> projMatrix{};
    int frameIndex{0};
    std::optional<hdr_camera_params> cameraParams{};
};

class Context {
  public:
    explicit Context(int device = 0) { Check(hdr_context_create(device, &_h), "hdr_context_create"); }
    ~Context() { hdr_context_destroy(_h); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    // IRenderJob::IsCancelled analogue: a cancelled context refuses further frames.
    void SetCancelled(bool c) { Check(hdr_context_set_cancel(_h, c), "hdr_context_set_cancel"); }
    [[nodiscard]] hdr_context *Get() const { return _h; }

  private:
    hdr_context *_h{nullptr};
};

// Scene::BuildRenderState equivalent: BVHs, lights and materials uploaded once and kept resident.
class Scene {
  public:
    Scene(Context &ctx, const hdr_scene_desc &desc) { Check(hdr_scene_upload(ctx.Get(), &desc, &_h), "hdr_scene_upload"); }
    ~Scene() { hdr_scene_destroy(_h); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    [[nodiscard]] hdr_scene *Get() const { return _h; }

  private:
    hdr_scene *_h{nullptr};
};

// RaycastPass + {RIS,}PathTracePass + AccumulationPass on the device.
class GpuPathTracePass {
  public:
    GpuPathTracePass(Context &ctx, Scene &scene, int width, int height, PipelineSettings settings)
        : _scene{scene}, _settings{settings}, _width{width}, _height{height} {
        Check(hdr_view_create(ctx.Get(), width, height, &_view), "hdr_view_create");
    }
    ~GpuPathTracePass() { hdr_view_destroy(_view); }
    GpuPathTracePass(const GpuPathTracePass &) = delete;
    GpuPathTracePass &operator=(const GpuPathTracePass &) = delete;

    // RenderPass::Inputs / Outputs: the pass consumes nothing from other passes and produces the
    // accumulated colour plus the RaycastPass AOVs.
    [[nodiscard]] static std::vector<std::string> Inputs() { return {}; }
    [[nodiscard]] static std::vector<std::string> Outputs() {
        return {std::string{kColorOutputName}, std::string{kDepthOutputName}, std::string{kAlbedoOutputName},
                std::string{kNormalOutputName}};
    }

    // One frame; `color` receives width*height RGBA32F (the "Color" buffer after accumulation).
    void Execute(const RenderContext &ctx, std::span<float> color) {
        if (color.size() < size_t(_width) * _height * 4) throw Error(HDR_ERR_INVALID_ARGUMENT, "color buffer too small");
        hdr_frame_desc f{};
        f.struct_size = sizeof(f);
        std::memcpy(f.view, ctx.viewMatrix.data(), sizeof f.view);
        std::memcpy(f.proj, ctx.projMatrix.data(), sizeof f.proj);
        f.frame_index = ctx.frameIndex;
        f.has_camera_params = ctx.cameraParams.has_value();
        if (ctx.cameraParams) f.camera = *ctx.cameraParams;
        const hdr_pipeline_desc p = _settings.Desc();
        if (!_settings.HasPostPasses()) {
            Check(hdr_render_frame(_view, _scene.Get(), &f, &p, color.data(), nullptr), "hdr_render_frame");
            return;
        }
        Check(hdr_render_frame(_view, _scene.Get(), &f, &p, nullptr, nullptr), "hdr_render_frame");
        const hdr_post_desc post = _settings.PostDesc();
        Check(hdr_view_post_process(_view, &post), "hdr_view_post_process");
        Check(hdr_view_read_color(_view, color.data()), "hdr_view_read_color");
    }

    // RenderPipeline::ClearPersistentBuffers (camera / settings / AOV change).
    void ClearPersistentBuffers() { Check(hdr_view_reset(_view), "hdr_view_reset"); }

    // The camera moved (renderer.cpp:401-406 clears the persistent buffers here). With
    // TemporalReprojection the reservoirs stay for the next frame's reprojected reuse and only the
    // accumulation restarts.
    void CameraMoved() {
        if (_settings.TemporalReprojection)
            Check(hdr_view_reset_accumulation(_view), "hdr_view_reset_accumulation");
        else
            ClearPersistentBuffers();
    }

    void ReadAovs(std::span<float> depth, std::span<float> albedo, std::span<float> normal) {
        Check(hdr_view_read_aovs(_view, depth.data(), albedo.data(), normal.data()), "hdr_view_read_aovs");
    }

  private:
    Scene &_scene;
    PipelineSettings _settings;
    int _width, _height;
    hdr_view *_view{nullptr};
};

}  // namespace hdrgpu
