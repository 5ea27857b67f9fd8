// adapter_smoke.cpp -- drives libhdrgpu.so through the C++20 pass adapter the way a host
// pipeline would: build a scene description, upload it, execute frames, reset, cancel.
//
//   adapter_smoke            render a lit box on device 0 with the RIS and PathTracer pipelines
//   adapter_smoke --no-device   expect context creation to fail cleanly (machines without a GPU)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpu_pass.h"

namespace {

struct BoxScene {
    std::vector<float> points;
    std::vector<int32_t> indices;
    hdr_mesh_desc mesh{};
    hdr_material_desc mat{};
    hdr_instance_desc inst{};
    hdr_light_desc light{};
    hdr_scene_desc desc{};

    BoxScene() {
        const float h = 1.0f;
        const float p[8][3] = {{-h, -h, -h}, {h, -h, -h}, {h, h, -h}, {-h, h, -h},
                               {-h, -h, h},  {h, -h, h},  {h, h, h},  {-h, h, h}};
        for (auto &v : p) points.insert(points.end(), v, v + 3);
        const int q[6][4] = {{0, 1, 2, 3}, {5, 4, 7, 6}, {4, 0, 3, 7}, {1, 5, 6, 2}, {4, 5, 1, 0}, {3, 2, 6, 7}};
        for (auto &f : q) {
            const int t[6] = {f[0], f[1], f[2], f[0], f[2], f[3]};
            indices.insert(indices.end(), t, t + 6);
        }
        mesh.struct_size = sizeof mesh;
        mat.struct_size = sizeof mat;
        inst.struct_size = sizeof inst;
        light.struct_size = sizeof light;
        mesh.points = points.data();
        mesh.num_points = 8;
        mesh.indices = indices.data();
        mesh.num_triangles = 12;
        const float base[3] = {0.7f, 0.6f, 0.5f}, one[3] = {1, 1, 1};
        std::memcpy(mat.diffuse_color, base, sizeof base);
        mat.roughness = 0.5f;
        std::memcpy(mat.specular_color, one, sizeof one);
        mat.specular = 1.0f;
        mat.opacity = 1.0f;
        mat.ior = 1.5f;
        std::memcpy(mat.transmission_color, one, sizeof one);
        std::memcpy(mat.emission_color, one, sizeof one);
        std::memcpy(mat.coat_color, one, sizeof one);
        mat.coat_roughness = 0.1f;
        mat.coat_ior = 1.5f;
        std::memcpy(mat.sheen_color, one, sizeof one);
        mat.sheen_roughness = 0.3f;
        std::memcpy(mat.subsurface_color, one, sizeof one);
        for (int i = 0; i < 16; ++i) inst.transform[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        inst.mesh = 0;
        inst.material = 0;
        inst.light = -1;
        // rect light 0.5 x 0.5 below the ceiling, facing down: local -Z -> world -Y
        light.type = HDR_LIGHT_RECT;
        light.color[0] = light.color[1] = light.color[2] = 1.0f;
        light.intensity = 20.0f;
        const double t[16] = {1, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, 0.98, 0, 1};
        std::memcpy(light.transform, t, sizeof t);
        light.width = light.height = 0.5f;
        light.shaping_cone_angle = 180.0f;
        desc.struct_size = sizeof desc;
        desc.meshes = &mesh;
        desc.num_meshes = 1;
        desc.materials = &mat;
        desc.num_materials = 1;
        desc.instances = &inst;
        desc.num_instances = 1;
        desc.lights = &light;
        desc.num_lights = 1;
        desc.environment = -1;
    }
};

hdrgpu::RenderContext Camera(int frame, double z = 0.9) {
    hdrgpu::RenderContext c;
    // camera at (0, 0, z) looking down -Z: view = translate(0, 0, -z)
    for (int i = 0; i < 16; ++i) c.viewMatrix[i] = (i % 5 == 0) ? 1.0 : 0.0;
    c.viewMatrix[14] = -z;
    const double f = 1.0 / std::tan(30.0 * 3.14159265358979 / 180.0), n = 0.01, fa = 100.0;
    c.projMatrix = {f, 0, 0, 0, 0, f, 0, 0, 0, 0, -(fa + n) / (fa - n), -1, 0, 0, -2 * fa * n / (fa - n), 0};
    c.frameIndex = frame;
    return c;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::strcmp(argv[1], "--no-device") == 0) {
        try {
            hdrgpu::Context ctx(0);
        } catch (const hdrgpu::Error &e) {
            std::printf("no device: %s (status %d)\n", e.what(), int(e.status));
            return e.status == HDR_ERR_INVALID_ARGUMENT || e.status == HDR_ERR_DEVICE ? 0 : 1;
        }
        std::printf("unexpected: a device is present\n");
        return 1;
    }
    const int W = 64, H = 64;
    BoxScene box;
    hdrgpu::Context ctx(0);
    hdrgpu::Scene scene(ctx, box.desc);
    std::vector<float> color(size_t(W) * H * 4);
    double means[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        hdrgpu::PipelineSettings s;
        s.kind = k == 0 ? hdrgpu::PipelineSettings::Kind::RIS : hdrgpu::PipelineSettings::Kind::PathTracer;
        s.MaxDepth = 3;
        hdrgpu::GpuPathTracePass pass(ctx, scene, W, H, s);
        for (int f = 0; f < 8; ++f) pass.Execute(Camera(f), color);
        double sum = 0;
        for (size_t i = 0; i < color.size(); i += 4) {
            for (int c = 0; c < 3; ++c) {
                if (!std::isfinite(color[i + c])) {
                    std::printf("non-finite pixel\n");
                    return 1;
                }
                sum += color[i + c];
            }
        }
        means[k] = sum / (3.0 * W * H);
        pass.ClearPersistentBuffers();
        pass.Execute(Camera(0), color);
    }
    std::printf("mean RIS %.5f PathTracer %.5f\n", means[0], means[1]);
    if (!(means[0] > 0.0 && means[1] > 0.0) || std::fabs(means[0] - means[1]) > 0.1 * means[1]) return 1;
    // reprojected temporal reuse (an extension): a dolly of four frames where each move keeps the
    // reservoirs (CameraMoved clears only the accumulation); every frame stays finite and lit
    {
        hdrgpu::PipelineSettings s;
        s.kind = hdrgpu::PipelineSettings::Kind::RIS;
        s.MaxDepth = 1;
        s.SkipVisibility = true;
        s.SpatialCount = 2;
        s.SpatialRadius = 8.0f;
        s.TemporalReprojection = true;
        hdrgpu::GpuPathTracePass dolly(ctx, scene, W, H, s);
        for (int f = 0; f < 4; ++f) {
            if (f > 0) dolly.CameraMoved();
            dolly.Execute(Camera(0, 0.9 - 0.02 * f), color);
            double sum = 0.0;
            for (size_t i = 0; i < color.size(); i += 4) {
                if (!std::isfinite(color[i]) || !std::isfinite(color[i + 1]) || !std::isfinite(color[i + 2])) {
                    std::printf("non-finite pixel with reprojection\n");
                    return 1;
                }
                sum += color[i] + color[i + 1] + color[i + 2];
            }
            const double mean = sum / (3.0 * W * H);
            std::printf("reprojected dolly frame %d mean %.5f\n", f, mean);
            if (!(mean > 0.5 * means[1] && mean < 2.0 * means[1])) return 1;
        }
    }
    // the PathTracerPost passes: the chromaticity blur keeps the mean luma of a frame up to the
    // clamp at 0 and the edge pixels; the OIDN switch is refused
    {
        hdrgpu::PipelineSettings s;
        s.kind = hdrgpu::PipelineSettings::Kind::PathTracer;
        s.MaxDepth = 3;
        hdrgpu::GpuPathTracePass plain(ctx, scene, W, H, s);
        s.Denoiser.EnableChromaticityBlur = true;
        s.PostProcess.ChromaticAberration = 2.0f;
        hdrgpu::GpuPathTracePass post(ctx, scene, W, H, s);
        std::vector<float> a(color.size()), b(color.size());
        plain.Execute(Camera(0), a);
        post.Execute(Camera(0), b);
        double la = 0, lb = 0;
        for (size_t i = 0; i < a.size(); i += 4) {
            la += 0.25 * a[i] + 0.5 * a[i + 1] + 0.25 * a[i + 2];
            lb += 0.25 * b[i] + 0.5 * b[i + 1] + 0.25 * b[i + 2];
        }
        std::printf("post: mean luma %.5f -> %.5f\n", la / (W * H), lb / (W * H));
        if (a == b || std::fabs(la - lb) > 0.05 * la) return 1;
        s.Denoiser.EnableDenoiser = true;
        hdrgpu::GpuPathTracePass oidn(ctx, scene, W, H, s);
        try {
            oidn.Execute(Camera(0), b);
            std::printf("OIDN request accepted\n");
            return 1;
        } catch (const hdrgpu::Error &e) {
            if (e.status != HDR_ERR_UNSUPPORTED) return 1;
        }
    }
    // cancellation: a cancelled context refuses the next frame with HDR_ERR_CANCELLED
    hdrgpu::GpuPathTracePass pass(ctx, scene, W, H, hdrgpu::PipelineSettings{});
    ctx.SetCancelled(true);
    try {
        pass.Execute(Camera(0), color);
        std::printf("cancel ignored\n");
        return 1;
    } catch (const hdrgpu::Error &e) {
        if (e.status != HDR_ERR_CANCELLED) return 1;
    }
    std::printf("adapter smoke ok\n");
    return 0;
}
