// k_post.hip -- The passes after accumulation, on the device: DenoiserPass's prefilter (the
// chromaticity blur, which is the reference denoiser's whole output when it is built without OIDN)
// and PostProcessPass (lens flare, chromatic aberration). Image kernels over a view's RGBA float
// colour, HBM / L2 bound; every sum is formed in the reference loop's order, so the results equal
// the CPU passes bit for bit.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace hdr {

namespace {
// std::max(0.0f, v): v when 0 < v, else +0 (also for NaN and -0).
__device__ __forceinline__ float pos(float v) { return 0.0f < v ? v : 0.0f; }
}  // namespace

// Denoiser::_applyChromaticityBlur (denoiser.cpp:54-94): luma Y of the pixel, chroma (Co, Cg) averaged
// over its 3x3 neighbourhood with clamped edges, back to RGB clamped at 0. One pixel per lane of a
// 16x16 block; the nine neighbours are L1 / L2 hits of the block's own rows.
__global__ void __launch_bounds__(256) chroma_blur_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          int W, int H) {
    const int x = int(blockIdx.x) * 16 + int(threadIdx.x), y = int(blockIdx.y) * 16 + int(threadIdx.y);
    if (x >= W || y >= H) return;
    const float4 c = src[size_t(y) * W + x];
    const float Y = 0.25f * c.x + 0.5f * c.y + 0.25f * c.z;
    float sumCo = 0.0f, sumCg = 0.0f, weightSum = 0.0f;
    for (int dy = -1; dy <= 1; ++dy) {
        const int ny = min(max(y + dy, 0), H - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = min(max(x + dx, 0), W - 1);
            const float4 n = src[size_t(ny) * W + nx];
            sumCo += 0.50f * n.x - 0.5f * n.z;
            sumCg += -0.25f * n.x + 0.5f * n.y - 0.25f * n.z;
            weightSum += 1.0f;
        }
    }
    const float Co = sumCo / weightSum, Cg = sumCg / weightSum;
    dst[size_t(y) * W + x] = make_float4(pos(Y + Co - Cg), pos(Y + Cg), pos(Y - Co - Cg), c.w);
}

// The lens flare's bright pass (post_process.cpp:26-41): a pixel whose Rec.709 luminance exceeds 2
// scatters its excess over 2. One block per image row; w = 1 marks a pixel that scatters (some
// component of its excess is nonzero: the CPU loop skips the others), and row_any[y] = 1 when any
// pixel of row y does, so the gather skips the rows that scatter nothing.
__global__ void __launch_bounds__(256) flare_bright_kernel(const float4 *__restrict__ src, float4 *__restrict__ bloom,
                                                           int *__restrict__ row_any, int W) {
    const int y = int(blockIdx.x);
    int any = 0;
    for (int x = int(threadIdx.x); x < W; x += int(blockDim.x)) {
        const size_t i = size_t(y) * W + x;
        const float4 c = src[i];
        const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
        float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (lum > 2.0f) b = make_float4(c.x - 2.0f, c.y - 2.0f, c.z - 2.0f, 0.0f);
        const bool scatters = b.x != 0.0f || b.y != 0.0f || b.z != 0.0f;
        b.w = scatters ? 1.0f : 0.0f;
        bloom[i] = b;
        any |= scatters ? 1 : 0;
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) row_any[y] = any;
}

namespace {
__device__ __forceinline__ void flare_add(float3 &acc, const float4 &b, float w, bool half) {
    if (b.w == 0.0f) return;
    if (half) {
        acc.x += b.x * w * 0.5f;
        acc.y += b.y * w * 0.5f;
        acc.z += b.z * w * 0.5f;
    } else {
        acc.x += b.x * w;
        acc.y += b.y * w;
        acc.z += b.z * w;
    }
}
}  // namespace

// The lens flare's blur and composite (post_process.cpp:43-108) as a gather. The CPU loop scatters:
// each bright source (x, y) adds bloom * 1/(|d|+1) to (x+d, y) and (x, y+d), and half of that to
// (x+d, y+d) and (x+d, y-d), for d = -B, -B+2, ..., B (d != 0, B = max(1, min(W, H) / 20)). A source
// reaches a target along at most one of the four directions (they differ in dy = 0, dx = 0, dx = dy,
// dx = -dy), so each target sums one term per source, in the scatter's order: sources in raster
// order. Here a lane walks its target's sources in that order -- rows ty - B .. ty + B; in a row at
// offset e = ty - sy != 0 (with e an offset of the set) the sources tx - |e|, tx, tx + |e| (a
// diagonal, the column, the other diagonal); in its own row tx - d for d from B down to -B -- and
// adds 0.1 of the sum to the colour: out = colour + blurred * 0.1 on every pixel, as the CPU does.
__global__ void __launch_bounds__(256) flare_gather_kernel(const float4 *__restrict__ src,
                                                           const float4 *__restrict__ bloom,
                                                           const int *__restrict__ row_any, float4 *__restrict__ dst,
                                                           int W, int H, int B) {
    const int tx = int(blockIdx.x) * 16 + int(threadIdx.x), ty = int(blockIdx.y) * 16 + int(threadIdx.y);
    if (tx >= W || ty >= H) return;
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    const int y0 = max(0, ty - B), y1 = min(H - 1, ty + B);
    for (int sy = y0; sy <= y1; ++sy) {
        if (!row_any[sy]) continue;
        const float4 *row = bloom + size_t(sy) * W;
        const int e = ty - sy;
        if (e == 0) {
            for (int d = B; d >= -B; d -= 2) {
                if (d == 0) continue;
                const int sx = tx - d;
                if (sx < 0 || sx >= W) continue;
                flare_add(acc, row[sx], 1.0f / (float(abs(d)) + 1.0f), false);
            }
            continue;
        }
        const int ae = abs(e);
        if (((e + B) & 1) != 0) continue;  // not an offset of the set (ae <= B by the row range)
        const float w = 1.0f / (float(ae) + 1.0f);
        if (tx - ae >= 0) flare_add(acc, row[tx - ae], w, true);
        flare_add(acc, row[tx], w, false);
        if (tx + ae < W) flare_add(acc, row[tx + ae], w, true);
    }
    const size_t i = size_t(ty) * W + tx;
    const float4 c = src[i];
    dst[i] = make_float4(c.x + acc.x * 0.1f, c.y + acc.y * 0.1f, c.z + acc.z * 0.1f, c.w);
}

// PostProcess::_applyChromaticAberration (post_process.cpp:111-142): red read ca * r / r_max pixels
// further in along the pixel's radius from the image centre, blue as far out (truncated, clamped to
// the image), green and alpha kept.
__global__ void __launch_bounds__(256) chromatic_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        int W, int H, float ca) {
    const int x = int(blockIdx.x) * 16 + int(threadIdx.x), y = int(blockIdx.y) * 16 + int(threadIdx.y);
    if (x >= W || y >= H) return;
    const unsigned w = unsigned(W), h = unsigned(H);
    const float maxDist = sqrtf(float(w * w + h * h)) * 0.5f;
    const float dx = float(x) - float(w) * 0.5f;
    const float dy = float(y) - float(h) * 0.5f;
    const float dist = sqrtf(dx * dx + dy * dy);
    const float dirX = dist > 0.0f ? dx / dist : 0.0f;
    const float dirY = dist > 0.0f ? dy / dist : 0.0f;
    const float shift = ca * (dist / maxDist);
    const int rx = min(max(int(float(x) - dirX * shift), 0), W - 1);
    const int ry = min(max(int(float(y) - dirY * shift), 0), H - 1);
    const int bx = min(max(int(float(x) + dirX * shift), 0), W - 1);
    const int by = min(max(int(float(y) + dirY * shift), 0), H - 1);
    const size_t i = size_t(y) * W + x;
    const float4 c = src[i];
    dst[i] = make_float4(src[size_t(ry) * W + rx].x, c.y, src[size_t(by) * W + bx].z, c.w);
}

}  // namespace hdr
