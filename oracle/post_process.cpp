// post_process.cpp -- CPU oracle of the passes after accumulation: DenoiserPass's prefilter and
// PostProcessPass.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). A line-by-line restatement of the reference's CPU loops,
// in their order of operations (compiled with -ffp-contract=off, so every product and sum rounds as
// written):
//   ChromaticityBlur ....... Denoiser::_applyChromaticityBlur (source/passes/denoiser.cpp:54-94)
//   LensFlare .............. PostProcess::_applyLensFlare (source/passes/post_process.cpp:23-109),
//                            its scatter loop over source pixels in raster order
//   ChromaticAberration .... PostProcess::_applyChromaticAberration (post_process.cpp:111-142)
//   oracle_post_process .... DenoiserPass::_execute -> Denoiser::Run without OIDN
//                            (denoiser_pass.cpp:13-38, denoiser.cpp:154-201), then
//                            PostProcessPass::_execute -> PostProcess::Run (post_process_pass.cpp:10-21,
//                            post_process.cpp:144-160, 197-210): the order of the reference's
//                            PathTracerPost pipeline (path_tracer_pipeline.h:224-229)
// What pins each part (tests/test_oracle_pins.py; a planted mistake fails its pin): the blur of a
// constant image returns the image (the YCoCg round trip) and of a grey image its luminance; a lone
// bright pixel spreads the flare's cross and diagonals with weights 0.1 / (|d| + 1) (x 0.5 on the
// diagonals) at every other offset; an image under the threshold is unchanged; the aberration moves
// red inward and blue outward along the radius by ca * r / r_max and keeps green and the centre.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace {

std::vector<float> PackRgb(const float *rgba, std::size_t n) {
    std::vector<float> c(n * 3);
    for (std::size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) c[i * 3 + k] = rgba[i * 4 + k];
    return c;
}

// denoiser.cpp:54-94
void ChromaticityBlur(std::vector<float> &out, const std::vector<float> &color, unsigned width, unsigned height) {
    for (int y = 0; y < int(height); ++y) {
        for (int x = 0; x < int(width); ++x) {
            const std::size_t idx = (std::size_t(y) * width + std::size_t(x)) * 3;
            const float r = color[idx], g = color[idx + 1], b = color[idx + 2];
            const float Y = 0.25f * r + 0.5f * g + 0.25f * b;
            float sumCo = 0.0f, sumCg = 0.0f, weightSum = 0.0f;
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    const int nx = std::clamp(x + dx, 0, int(width) - 1);
                    const int ny = std::clamp(y + dy, 0, int(height) - 1);
                    const std::size_t n = (std::size_t(ny) * width + std::size_t(nx)) * 3;
                    const float nr = color[n], ng = color[n + 1], nb = color[n + 2];
                    sumCo += 0.50f * nr - 0.5f * nb;
                    sumCg += -0.25f * nr + 0.5f * ng - 0.25f * nb;
                    weightSum += 1.0f;
                }
            }
            const float Co = sumCo / weightSum;
            const float Cg = sumCg / weightSum;
            out[idx] = std::max(0.0f, Y + Co - Cg);
            out[idx + 1] = std::max(0.0f, Y + Cg);
            out[idx + 2] = std::max(0.0f, Y - Co - Cg);
        }
    }
}

// post_process.cpp:23-109: threshold 2 on Rec.709 luminance, then each bright pixel scatters its
// excess to every other offset d in [-blurSize, blurSize] along the row, the column and both
// diagonals; 0.1 of the scattered sum is added to the colour.
void LensFlare(std::vector<float> &finalColor, const std::vector<float> &color, unsigned width, unsigned height) {
    std::vector<float> bloom(std::size_t(width) * height * 3, 0.0f);
    constexpr float threshold = 2.0f;
    for (unsigned i = 0; i < width * height; ++i) {
        const float r = color[i * 3], g = color[i * 3 + 1], b = color[i * 3 + 2];
        const float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
        if (lum > threshold) {
            bloom[i * 3] = r - threshold;
            bloom[i * 3 + 1] = g - threshold;
            bloom[i * 3 + 2] = b - threshold;
        }
    }
    const int blurSize = std::max(1, int(std::min(width, height) / 20));
    std::vector<float> blurred(std::size_t(width) * height * 3, 0.0f);
    auto add = [&](std::size_t t, std::size_t s, float w, float extra) {
        for (int k = 0; k < 3; ++k) {
            if (extra == 1.0f) blurred[t + k] += bloom[s + k] * w;
            else blurred[t + k] += bloom[s + k] * w * extra;
        }
    };
    for (unsigned y = 0; y < height; ++y) {
        for (unsigned x = 0; x < width; ++x) {
            const std::size_t idx = (std::size_t(y) * width + x) * 3;
            if (bloom[idx] == 0.0f && bloom[idx + 1] == 0.0f && bloom[idx + 2] == 0.0f) continue;
            for (int d = -blurSize; d <= blurSize; d += 2) {
                if (d == 0) continue;
                const float weight = 1.0f / (float(std::abs(d)) + 1.0f);
                const int xd = int(x) + d, yd = int(y) + d, ym = int(y) - d;
                if (xd >= 0 && xd < int(width)) add((std::size_t(y) * width + unsigned(xd)) * 3, idx, weight, 1.0f);
                if (yd >= 0 && yd < int(height)) add((std::size_t(yd) * width + x) * 3, idx, weight, 1.0f);
                if (xd >= 0 && xd < int(width) && yd >= 0 && yd < int(height))
                    add((std::size_t(yd) * width + unsigned(xd)) * 3, idx, weight, 0.5f);
                if (xd >= 0 && xd < int(width) && ym >= 0 && ym < int(height))
                    add((std::size_t(ym) * width + unsigned(xd)) * 3, idx, weight, 0.5f);
            }
        }
    }
    for (unsigned i = 0; i < width * height; ++i)
        for (int k = 0; k < 3; ++k) finalColor[i * 3 + k] += blurred[i * 3 + k] * 0.1f;
}

// post_process.cpp:111-142: red from ca * r / r_max pixels further in along the radius, blue from as
// far out, green kept.
void ChromaticAberration(std::vector<float> &finalColor, unsigned width, unsigned height, float ca) {
    std::vector<float> caColor = finalColor;
    const float maxDist = std::sqrt(float(width * width + height * height)) * 0.5f;
    for (unsigned y = 0; y < height; ++y) {
        for (unsigned x = 0; x < width; ++x) {
            const float dx = float(x) - float(width) * 0.5f;
            const float dy = float(y) - float(height) * 0.5f;
            const float dist = std::sqrt(dx * dx + dy * dy);
            const float dirX = dist > 0.0f ? dx / dist : 0.0f;
            const float dirY = dist > 0.0f ? dy / dist : 0.0f;
            const float shift = ca * (dist / maxDist);
            const int rx = std::clamp(int(float(x) - dirX * shift), 0, int(width) - 1);
            const int ry = std::clamp(int(float(y) - dirY * shift), 0, int(height) - 1);
            const int bx = std::clamp(int(float(x) + dirX * shift), 0, int(width) - 1);
            const int by = std::clamp(int(float(y) + dirY * shift), 0, int(height) - 1);
            const std::size_t idx = (std::size_t(y) * width + x) * 3;
            const std::size_t rIdx = (std::size_t(ry) * width + unsigned(rx)) * 3;
            const std::size_t bIdx = (std::size_t(by) * width + unsigned(bx)) * 3;
            caColor[idx] = finalColor[rIdx];
            caColor[idx + 2] = finalColor[bIdx + 2];
        }
    }
    finalColor = std::move(caColor);
}

}  // namespace

extern "C" {

// rgba: width * height RGBA floats, rewritten in place (alpha kept). chroma_blur:
// Denoiser::Config::EnableChromaticityBlur (the denoiser's output without OIDN); lens_flare and ca:
// PostProcess::Config. Returns -1 for an empty image.
int32_t oracle_post_process(float *rgba, int32_t width, int32_t height, int32_t chroma_blur, int32_t lens_flare,
                            float ca) {
    if (width <= 0 || height <= 0) return -1;
    const unsigned W = unsigned(width), H = unsigned(height);
    const std::size_t n = std::size_t(W) * H;
    if (chroma_blur) {  // denoiser.cpp:154-176 (no OIDN: the output is the prefiltered colour)
        const std::vector<float> color = PackRgb(rgba, n);
        std::vector<float> prefiltered(n * 3);
        ChromaticityBlur(prefiltered, color, W, H);
        for (std::size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) rgba[i * 4 + k] = prefiltered[i * 3 + k];
    }
    if (lens_flare || ca > 0.0f) {  // post_process.cpp:144-160
        const std::vector<float> color = PackRgb(rgba, n);
        std::vector<float> finalColor = color;
        if (lens_flare) LensFlare(finalColor, color, W, H);
        if (ca > 0.0f) ChromaticAberration(finalColor, W, H, ca);
        for (std::size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) rgba[i * 4 + k] = finalColor[i * 3 + k];
    }
    return 0;
}

}  // extern "C"
