// hdrgpu.hip -- host side and C-ABI of libhdrgpu.so (declared in include/hdrgpu.h), plus the small
// kernels of scene upload, queue clears, the firefly filter and the RNG parity entry point. The
// rendering kernels live in k_*.hip (declared in kernels.h).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hdrgpu.h"
#include "host_scene.h"
#include "kernels.h"
#include "wavefront.h"

using namespace hdr;


// DomeLight::Prepare (dome_light.cpp:49-76) at scene upload: per row y, the luminance of each texel
// weighted by sin(theta_y), its running sum as the column CDF and the row's total, summed texel by
// texel in the reference's order (one thread per row keeps that order), then normalised by the
// row total; the row totals go to row[y + 1] for dome_cdf_total_kernel.
__global__ void dome_cdf_rows_kernel(const float *__restrict__ pix, const float *__restrict__ sin_theta, int w, int h,
                                     float *row, float *col) {
    const int y = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (y >= h) return;
    const float sinTheta = sin_theta[y];
    float *c = col + size_t(y) * size_t(w + 1);
    const float *p = pix + size_t(y) * size_t(w) * 3;
    float rowLum = 0.0f;
    c[0] = 0.0f;
    for (int x = 0; x < w; ++x) {
        float lum = 0.2126f * p[3 * x] + 0.7152f * p[3 * x + 1] + 0.0722f * p[3 * x + 2];
        lum *= sinTheta;
        rowLum += lum;
        c[x + 1] = c[x] + lum;
    }
    if (rowLum > 0)
        for (int x = 0; x <= w; ++x) c[x] /= rowLum;
    row[y + 1] = rowLum;
}
// The row CDF (running sum of the row totals, normalised by their sum) and the dome's total
// luminance and pdf scale (totalLum * (pi/h) * (2pi/w)), in the reference's order on one thread.
__global__ void dome_cdf_total_kernel(float *row, int w, int h, GDome *rec) {
    float total = 0.0f;
    row[0] = 0.0f;
    for (int y = 0; y < h; ++y) {
        const float rowLum = row[y + 1];
        total += rowLum;
        row[y + 1] = row[y] + rowLum;
    }
    if (total > 0)
        for (int y = 0; y <= h; ++y) row[y] /= total;
    rec->total_lum = total;
    rec->pdf_scale = total * (float(M_PI) / float(h)) * (2.0f * float(M_PI) / float(w));
}

// Clears the shadow-query count and the persistent shadow launch's claim counter (a kernel rather
// than a memset, so the step can sit in a captured graph).
__global__ void zero_u32x2_kernel(unsigned int *p) {
    p[0] = 0u;
    p[1] = 0u;
}

// ---------------------------------------------------------------------------- wavefront path
// RIS / ReSTIR at every vertex of paths of any depth (wavefront.h). Per frame: the G-buffer, then
// wave_shade0 (primary vertex, one wave per 8x8 tile), wave_trace (bounce rays and shadow rays of
// vertex 0), then wave_shade / wave_trace for vertices 1 .. depth-1 over the live paths only, then
// wave_resolve (spectrum -> RGB and accumulation). Queue counts are indexed by vertex parity; each
// launch clears the counts the launch after it appends to.


// Clears the wavefront queue counts (a kernel, so a step can sit in a captured graph).
__global__ void wave_zero_counts_kernel(unsigned int *c) {
    if (threadIdx.x < kWaveCounts) c[threadIdx.x] = 0u;
}

// Out-of-place 3x3 firefly clamp (the reference's GPU accumulation shader semantics,
// accumulation.slang:55-81; threshold 4 * max neighbour luminance + 0.5) + accumulation.
// AccumulationPass (accumulation_pass.cpp:98-106) once per frame over `frames` staged samples of
// consecutive frames, in frame order: the sums, the metrics and the final colour of that many
// frame-by-frame accumulations.
__global__ void accumulate_frames_kernel(const float4 *stage, int frames, size_t P, int first_frame, float4 *accum,
                                         float4 *color, double *lum_sum, double *lum_sq) {
    const size_t pixel = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (pixel >= P) return;
    for (int f = 0; f < frames; ++f)
        accumulate_sample(accum, color, lum_sum, lum_sq, int(pixel), stage[size_t(f) * P + pixel],
                          1.0f / float(first_frame + f + 1));
}

__global__ void firefly_accumulate_kernel(const float4 *frame, float4 *accum, float4 *color, double *lum_sum,
                                          double *lum_sq, int W, int H, float invN) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    auto lum = [](float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; };
    const int i = y * W + x;
    float4 c = frame[i];
    const float l = lum(c);
    float maxN = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (dx == 0 && dy == 0) continue;
            const int nx = iclamp(x + dx, 0, W - 1), ny = iclamp(y + dy, 0, H - 1);
            maxN = smax(maxN, lum(frame[ny * W + nx]));
        }
    const float thr = maxN * 4.0f + 0.5f;
    if (l > thr && l > 0.0f) {
        const float s = thr / l;
        c.x *= s;
        c.y *= s;
        c.z *= s;
    }
    accumulate_sample(accum, color, lum_sum, lum_sq, i, c, invN);
}

// The Variance metrics of AccumulationPass (accumulation_pass.cpp:108-136), first pass: per block
// of 256 pixels, the sum and maximum of the per-pixel estimator variance and the sum of the
// per-pixel mean luminance (n = accumulated samples >= 2), reduced in a fixed order.
__global__ void variance_partial_kernel(const double *lum_sum, const double *lum_sq, int P, int n, double *partial) {
    __shared__ double ssum[256], smax[256], slum[256];
    const int t = threadIdx.x, i = blockIdx.x * 256 + t;
    double ev = 0.0, mean = 0.0;
    if (i < P) {
        const double nn = double(n);
        mean = lum_sum[i] / nn;
        const double centered = lum_sq[i] - nn * mean * mean;
        const double sampleVariance = fmax(0.0, centered / (nn - 1.0));
        ev = sampleVariance / nn;
    }
    ssum[t] = ev;
    smax[t] = ev;
    slum[t] = mean;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            ssum[t] += ssum[t + h];
            smax[t] = fmax(smax[t], smax[t + h]);
            slum[t] += slum[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[3 * blockIdx.x] = ssum[0];
        partial[3 * blockIdx.x + 1] = smax[0];
        partial[3 * blockIdx.x + 2] = slum[0];
    }
}
// Second pass: the blocks' partials in order; out = {mean, max estimator variance, mean luminance}.
__global__ void variance_final_kernel(const double *partial, int blocks, int P, double *out) {
    if (threadIdx.x != 0) return;
    double sum = 0.0, mx = 0.0, lum = 0.0;
    for (int b = 0; b < blocks; ++b) {
        sum += partial[3 * b];
        mx = fmax(mx, partial[3 * b + 1]);
        lum += partial[3 * b + 2];
    }
    out[0] = sum / double(P);
    out[1] = mx;
    out[2] = lum / double(P);
}

// One render-buffer write of a resolved output: UpscalePass's nearest texel (upscale_pass.cpp:17-31,
// source x = min(sw - 1, x * sw / dw) in integer arithmetic, likewise y) of an sc-component source,
// stored as `comps` components of `fmt` the way HdRestirRenderBuffer::_WriteOutput converts
// (hd_restir_render_buffer.cpp:46-74): missing components 0, half rounded to nearest even, UNorm8
// clamp(v * 255, 0, 255) and SNorm8 clamp(v * 127, -128, 127) truncated, Int32 truncated (NaN: 0).
__global__ void resolve_kernel(const float *__restrict__ src, int sc, int sw, int sh, int dw, int dh, int fmt,
                               int comps, unsigned char *__restrict__ dst) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= dw || y >= dh) return;
    const int sx = min(sw - 1, int((long long)x * sw / dw));
    const int sy = min(sh - 1, int((long long)y * sh / dh));
    const float *v = src + (size_t(sy) * sw + sx) * sc;
    const size_t o = size_t(y) * dw + x;
    for (int c = 0; c < comps; ++c) {
        const float f = c < sc ? v[c] : 0.0f;
        switch (fmt) {
        case HDR_FORMAT_FLOAT32: reinterpret_cast<float *>(dst)[o * comps + c] = f; break;
        case HDR_FORMAT_FLOAT16:
            reinterpret_cast<unsigned short *>(dst)[o * comps + c] = __half_as_ushort(__float2half_rn(f));
            break;
        case HDR_FORMAT_UNORM8: {
            float t = f * 255.0f;
            t = t < 0.0f ? 0.0f : (255.0f < t ? 255.0f : t);
            dst[o * comps + c] = t == t ? (unsigned char)(unsigned int)t : 0;
            break;
        }
        case HDR_FORMAT_SNORM8: {
            float t = f * 127.0f;
            t = t < -128.0f ? -128.0f : (127.0f < t ? 127.0f : t);
            reinterpret_cast<signed char *>(dst)[o * comps + c] = t == t ? (signed char)(int)t : 0;
            break;
        }
        default: reinterpret_cast<int *>(dst)[o * comps + c] = f == f ? int(f) : 0; break;
        }
    }
}

__global__ void rng_kernel(const uint32_t *seeds, int n, int nd, uint64_t *raw, float *flt, uint64_t *spill) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mt64 r;
    r.init(seeds[i], spill ? spill + size_t(i) * kMtN : nullptr);
    for (int k = 0; k < nd; ++k) {
        const uint64_t v = r.next_raw();
        if (raw) raw[size_t(i) * nd + k] = v;
        if (flt) flt[size_t(i) * nd + k] = mt_canonical(v);
    }
}

// Scratch of the allocation-free entry point hdr_rng_raw_noalloc: the 312-word spill slots of the
// streams drawn past output 311, one slot per seed (at most kRngScratchBytes / 2496 = 13443 seeds).
// Each device holds its own copy until the process ends. Host code reaches it only through the
// address hipGetSymbolAddress returns on that device (g_scratch, set by hdr_context_create), never
// through this name.
constexpr size_t kRngScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint64_t scratch[kRngScratchBytes / sizeof(uint64_t)];


// ============================================================================ host side

namespace {

thread_local std::string g_err;

// Per-device address of `scratch` (set once per device, with that device current).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::atomic<uint64_t *> g_scratch[kMaxDevices];

hdr_status fail(hdr_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail(e_ == hipErrorOutOfMemory ? HDR_ERR_OUT_OF_MEMORY : HDR_ERR_DEVICE,              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                              \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t bytes) {
        n = bytes;
        return hipMalloc(&p, bytes ? bytes : 1);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Pinned host memory (hipHostMalloc'd; freed on destruction).
template <class T>
struct PinnedBuf {
    T *p = nullptr;
    size_t n = 0;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t alloc(size_t count) {
        n = count;
        return hipHostMalloc(reinterpret_cast<void **>(&p), (count ? count : 1) * sizeof(T), hipHostMallocDefault);
    }
};

}  // namespace

struct hdr_context {
    int device = 0;
    int cancelled = 0;
    int cus = 0;  // compute units of the device (persistent launches fill them)
};

struct hdr_scene {
    hdr_context *ctx = nullptr;
    uint64_t id = 0;  // unique per upload and per update (G-buffer reuse keys on it)
    DevBuf nodes;     // BLAS nodes of every mesh, then room for tlas_cap TLAS node records
    DevBuf geo;       // triangles, then their shading attributes
    DevBuf state;     // TLAS slot table, instances, materials, lights, dome records
    std::vector<std::unique_ptr<DevBuf>> dome_mem;  // per dome: texels, row CDF, column CDFs
    size_t tlas_cap = 0;
    DeviceScene ds{};
    HostScene hs;     // the host build (triangles and attributes dropped after the upload)
    uint64_t bytes = 0;
    int max_sample_draws_per_vertex = 3;  // SampleBounce draws: 3, +3 per coat / sheen layer present
    bool any_translucent = false;         // some material has opacity < 0.999 (unbounded draws)
};

// Launch shape of the queue-driven launches (hdr_launch_config, validated by the caller):
// waves per launch and the minimum-waves-per-SIMD variant of the traversal / shading kernels.
struct LaunchShape {
    int queue_waves = 0;  // 0: one wave per 64 queue slots, at most kDefaultQueueWaves
    int trace_occ = 8;    // at least 8 waves per SIMD (1: the compiler's register budget)
    int shade_occ = 3;
    bool lds_stack = false;  // traversal stacks with LDS rings (RingStack) instead of private memory
};
constexpr int kDefaultQueueWaves = 32768;
// Entries of a lane's any-hit stack in LDS (8 KB per wave; 4 waves per SIMD hold 128 of a CU's 160 KB).
constexpr int kLdsStack = 32;
// The persistent any-hit lanes of a scene fit their stack in kLdsStack LDS entries: a tree whose
// deepest leaf is at depth D (root at 1) stacks at most D - 1 entries, the TLAS's below the BLAS's.
static bool lds_stack_fits(const hdr_scene *s) {
    return s->hs.max_tlas_depth + s->hs.max_blas_depth - 2 <= kLdsStack;
}
static int queue_blocks(const LaunchShape &ls, size_t cap) {
    const size_t need = std::max<size_t>(1, (cap + kLanes - 1) / kLanes);
    return int(std::min<size_t>(need, size_t(ls.queue_waves > 0 ? ls.queue_waves : kDefaultQueueWaves)));
}
template <bool C, bool L, bool T>
static void launch_wave_trace_l(const LaunchShape &ls, int blocks, hipStream_t st, const FrameArgs *a, int b) {
#define HDR_WT(O)                                                                                          \
    if (T) hipLaunchKernelGGL((wave_trace_traced_kernel<C, O, L>), dim3(blocks), dim3(kLanes), 0, st, a, b); \
    else hipLaunchKernelGGL((wave_trace_kernel<C, O, L>), dim3(blocks), dim3(kLanes), 0, st, a, b);
    switch (ls.trace_occ) {
    case 4: HDR_WT(4) break;
    case 6: HDR_WT(6) break;
    case 8: HDR_WT(8) break;
    default: HDR_WT(1) break;
    }
#undef HDR_WT
}
// Traversal launch b: bounce and shadow rays one per lane (wave_trace_kernel); on traced frames the
// bounce rays (wave_trace_traced_kernel), then the escape and shadow rays on persistent lanes
// (wave_anyhit_kernel, 4 waves per SIMD of the device, or queue_waves). one_scene: every view of
// the launch renders a[0]'s scene.
template <bool C>
static void launch_wave_trace(const LaunchShape &ls, int blocks, hipStream_t st, const FrameArgs *a, int b,
                              bool traced, int cus, bool one_scene, bool lds_fit) {
    if (traced) {
        if (ls.lds_stack) launch_wave_trace_l<C, true, true>(ls, blocks, st, a, b);
        else launch_wave_trace_l<C, false, true>(ls, blocks, st, a, b);
    } else {
        if (ls.lds_stack) launch_wave_trace_l<C, true, false>(ls, blocks, st, a, b);
        else launch_wave_trace_l<C, false, false>(ls, blocks, st, a, b);
    }
    if (traced) {
        const dim3 g(unsigned(ls.queue_waves > 0 ? ls.queue_waves : cus * 4 * 4)), l(kLanes);
        if (lds_fit) {
            if (one_scene) hipLaunchKernelGGL((wave_anyhit_kernel<C, true, true, kLdsStack>), g, l, 0, st, a, b);
            else hipLaunchKernelGGL((wave_anyhit_kernel<C, true, false, kLdsStack>), g, l, 0, st, a, b);
        } else {
            if (one_scene) hipLaunchKernelGGL((wave_anyhit_kernel<C, true, true, 0>), g, l, 0, st, a, b);
            else hipLaunchKernelGGL((wave_anyhit_kernel<C, true, false, 0>), g, l, 0, st, a, b);
        }
    }
}
// Shading launch b >= 1; traced frames: wave_shade_traced_kernel.
template <bool C, bool D>
static void launch_wave_shade(const LaunchShape &ls, int blocks, size_t lds, hipStream_t st, const FrameArgs *a, int b,
                              bool traced, bool gen = false) {
    const dim3 g(blocks), l(kLanes);
    if (traced && gen) {  // a view with mesh lights: the build that traces generated candidates
        if (ls.shade_occ == 3) hipLaunchKernelGGL((wave_shade_traced_kernel<C, D, 3, true>), g, l, lds, st, a, b);
        else hipLaunchKernelGGL((wave_shade_traced_kernel<C, D, 1, true>), g, l, lds, st, a, b);
    } else if (traced) {
        if (ls.shade_occ == 3) hipLaunchKernelGGL((wave_shade_traced_kernel<C, D, 3, false>), g, l, lds, st, a, b);
        else hipLaunchKernelGGL((wave_shade_traced_kernel<C, D, 1, false>), g, l, lds, st, a, b);
    } else {
        if (ls.shade_occ == 3) hipLaunchKernelGGL((wave_shade_kernel<C, D, 3>), g, l, lds, st, a, b);
        else hipLaunchKernelGGL((wave_shade_kernel<C, D, 1>), g, l, lds, st, a, b);
    }
}
// The primary-vertex launch: lean RNG stream when no pixel can pass output 310 there; traced
// candidates when BSDF-sampled rays can see an emitter.
template <bool C, bool D>
static void launch_wave_shade0(bool lean, bool traced, unsigned tiles, size_t lds, hipStream_t st, const FrameArgs *a,
                               const int *t, int n) {
    const dim3 g(tiles), b(kLanes);
    if (lean && traced) hipLaunchKernelGGL((wave_shade0_kernel<C, D, true, true>), g, b, lds, st, a, t, n);
    else if (lean) hipLaunchKernelGGL((wave_shade0_kernel<C, D, true, false>), g, b, lds, st, a, t, n);
    else if (traced) hipLaunchKernelGGL((wave_shade0_kernel<C, D, false, true>), g, b, lds, st, a, t, n);
    else hipLaunchKernelGGL((wave_shade0_kernel<C, D, false, false>), g, b, lds, st, a, t, n);
}
template <bool C, bool L>
static void launch_shadow_l(const LaunchShape &ls, int blocks, hipStream_t st, const FrameArgs *a) {
    switch (ls.trace_occ) {
    case 4: hipLaunchKernelGGL((shadow_kernel<C, 4, L>), dim3(blocks), dim3(kLanes), 0, st, a); break;
    case 6: hipLaunchKernelGGL((shadow_kernel<C, 6, L>), dim3(blocks), dim3(kLanes), 0, st, a); break;
    case 8: hipLaunchKernelGGL((shadow_kernel<C, 8, L>), dim3(blocks), dim3(kLanes), 0, st, a); break;
    default: hipLaunchKernelGGL((shadow_kernel<C, 1, L>), dim3(blocks), dim3(kLanes), 0, st, a); break;
    }
}
// Shadow rays of a launch's queue (at most `cap` queries). Without the launch shape's LDS rings the
// persistent kernel runs (4 waves per SIMD at most 128 VGPRs, lanes refilled in groups of 16): one
// wave per SIMD slot of the device, or queue_waves of the launch shape. one_scene: every view of the
// launch renders args[0]'s scene. lds_fit: every scene's stacks fit kLdsStack LDS entries per lane
// (lds_stack_fits), so the persistent kernel keeps them in LDS instead of private memory.
template <bool C>
static void launch_shadow(const LaunchShape &ls, int cus, size_t cap, hipStream_t st, const FrameArgs *a,
                          bool one_scene, bool lds_fit) {
    if (!ls.lds_stack) {
        const size_t slots = ls.queue_waves > 0 ? size_t(ls.queue_waves) : size_t(cus) * 4 * 4;
        const dim3 g(unsigned(std::min(slots, std::max<size_t>(1, (cap + kLanes - 1) / kLanes)))), l(kLanes);
        if (lds_fit) {
            if (one_scene) hipLaunchKernelGGL((shadow_persistent_kernel<C, 4, 16, true, kLdsStack>), g, l, 0, st, a);
            else hipLaunchKernelGGL((shadow_persistent_kernel<C, 4, 16, false, kLdsStack>), g, l, 0, st, a);
        } else {
            if (one_scene) hipLaunchKernelGGL((shadow_persistent_kernel<C, 4, 16, true, 0>), g, l, 0, st, a);
            else hipLaunchKernelGGL((shadow_persistent_kernel<C, 4, 16, false, 0>), g, l, 0, st, a);
        }
        return;
    }
    launch_shadow_l<C, true>(ls, queue_blocks(ls, cap), st, a);
}
template <bool C>
static void launch_gbuffer(const LaunchShape &ls, unsigned tiles, hipStream_t st, const FrameArgs *a, const int *t,
                           int n) {
    if (ls.lds_stack) hipLaunchKernelGGL((gbuffer_kernel<C, true>), dim3(tiles), dim3(kLanes), 0, st, a, t, n);
    else hipLaunchKernelGGL((gbuffer_kernel<C, false>), dim3(tiles), dim3(kLanes), 0, st, a, t, n);
}

// The lean path tracer with registers for 3 waves per SIMD once the launch holds at least four
// waves per SIMD of the GPU (16-frame 256 x 256 launches 0.210 -> 0.175 ms per frame); smaller
// launches keep 2 (a single 256 x 256 frame: 0.545 ms, 0.581 with 3).
template <bool C, bool D>
static void launch_pt_lean(int cus, unsigned tiles, unsigned frames, hipStream_t st, const FrameArgs *a,
                           const int *t, int n) {
    const dim3 grid(tiles, frames);
    if (size_t(tiles) * frames >= size_t(16) * size_t(cus))
        hipLaunchKernelGGL((pt_lean_kernel<C, D, 3>), grid, dim3(kLanes), 0, st, a, t, n);
    else
        hipLaunchKernelGGL((pt_lean_kernel<C, D, 2>), grid, dim3(kLanes), 0, st, a, t, n);
}

// Spatial-reuse launches, compiled for 4, 6 or 9 domains per pixel (spatial_domains).
template <bool C, bool D>
static void launch_spatial(int nd, dim3 grid, hipStream_t st, const FrameArgs *a, const int *t, int n) {
    if (nd <= 4) hipLaunchKernelGGL((spatial_kernel<C, D, 4>), grid, dim3(kLanes), 0, st, a, t, n);
    else if (nd <= 6) hipLaunchKernelGGL((spatial_kernel<C, D, 6>), grid, dim3(kLanes), 0, st, a, t, n);
    else hipLaunchKernelGGL((spatial_kernel<C, D, 9>), grid, dim3(kLanes), 0, st, a, t, n);
}
template <bool C>
static void launch_wave_spatial(int nd, dim3 grid, hipStream_t st, const FrameArgs *a, const int *t, int n) {
    if (nd <= 4) hipLaunchKernelGGL((wave_spatial_kernel<C, 4>), grid, dim3(kLanes), 0, st, a, t, n);
    else if (nd <= 6) hipLaunchKernelGGL((wave_spatial_kernel<C, 6>), grid, dim3(kLanes), 0, st, a, t, n);
    else hipLaunchKernelGGL((wave_spatial_kernel<C, 9>), grid, dim3(kLanes), 0, st, a, t, n);
}

struct hdr_view {
    hdr_context *ctx = nullptr;
    int W = 0, H = 0;
    hipStream_t stream = nullptr;
    DevBuf res, accum, color, frame, dbg, depth, albedo, normal, counters, spill, gbuf, rgb2, flags, queue, qcount;
    DevBuf status;  // device status word (kStatusRngOverrun, ...), read at the view's synchronisations
    DevBuf res_t, saux;  // spatial reuse: post-temporal reservoirs, per-pixel {ray dir, u}, {throughput}
    DevBuf wstate, whit, wray, wcq, walive, wcnt;  // wavefront path: path records, hits, queues
    DevBuf wesc, wesq;                             // traced frames: escape masks, escape-ray queue
    DevBuf dargs, dtile; // per-frame FrameArgs records (one per timing-ring slot); launch tile table {0, tiles}
    PinnedBuf<FrameArgs> hargs;  // pinned staging of dargs
    LaunchShape shape;           // queue-driven launches (hdr_view_set_launch_config)
    int counting = 0;
    int frames = 0;
    int32_t samples = 0;  // accumulated samples: frame_index + 1 of the last frame issued, 0 after a reset
    DevBuf resolved;      // staging of hdr_view_resolve
    DevBuf post;          // hdr_view_post_process: two RGBA planes, the bright pass, per-row flags
    DevBuf lum_sum, lum_sq, variance;  // Variance metrics (hdr_view_set_metrics): per-pixel sums, reduction
    // Reprojected temporal reuse: the previous frame's camera (hist_cam) and whether that frame left
    // what reproject_kernel reads (hist_records: a single-view spatial frame with
    // temporal_reprojection, whose final reservoirs, {ray, u} records and kept G-buffer records stay
    // until the next frame's gather); rp holds the kept G-buffer records and the gathered records
    // (GRec, GRec, float4, hdr_reservoir planes of P entries).
    bool hist_cam = false, hist_records = false;
    double hist_view[16] = {}, hist_proj[16] = {};
    float hist_origin[3] = {};
    uint64_t hist_scene = 0;
    DevBuf rp;
    // Frames overlap: the G-buffer is double-buffered by frame parity, and the G-buffer kernel of
    // frame N+1 runs on `gstream` as soon as frame N-1's shading has read that buffer (event
    // `shade_done[parity]`), beside frame N's shading, shadow rays and accumulation on `stream`;
    // frame N+1's shading waits for its G-buffer (event `gbuf_done`).
    hipStream_t gstream = nullptr;
    hipEvent_t shade_done[2] = {nullptr, nullptr}, gbuf_done = nullptr, serial = nullptr;
    bool shade_recorded[2] = {false, false};
    // G-buffer reuse: the pixel-centre rays (RaycastPass) depend only on the scene, the camera and
    // the view size, so while they stay the same (and the camera has no depth of field, whose lens
    // draws change every frame) the G-buffer of an earlier frame is this frame's; gcur is the half
    // holding the latest G-buffer and gkey what it was computed for.
    struct GbufKey {
        uint64_t scene_id;
        double view[16], proj[16];
        int32_t has_cam;
        hdr_camera_params cam;
        int32_t aovs;
    };
    GbufKey gkey;
    bool gvalid = false;
    bool gnext_valid = false;        // the G-buffer being written is reusable (no depth of field)
    int gcur = 1;                    // the first frame writes half 0
    bool single_since_batch = true;  // a single frame used the G-buffer after the last batch step
    // Per-kernel timing: one slot of events per frame in a ring, read back lazily (when the slot is
    // reused or the sums are requested), so timing never stalls the streams.
    // beg[k] .. end[k] brackets stage k (HDR_STAGE_*) on the stream it runs on; bit k of `ran`
    // says stage k launched; fin marks the end of the frame.
    struct StageSlot {
        hipEvent_t beg[HDR_STAGE_COUNT] = {};
        hipEvent_t end[HDR_STAGE_COUNT] = {};
        hipEvent_t fin = nullptr;
        uint32_t ran = 0;
        bool pending = false;
    };
    static constexpr int kRing = 16;
    StageSlot ring[kRing];
    int ring_next = 0, ring_last = -1;
    double stage_ms[HDR_STAGE_COUNT] = {};
    int32_t stage_n[HDR_STAGE_COUNT] = {};
};

// The view's device status word, read after its streams have been synchronised: a set
// kStatusRngOverrun bit means some frame since the last hdr_view_reset drew past the RNG outputs its
// stream was set up for, so its results are wrong.
static hdr_status view_status(hdr_view *v) {
    unsigned int word = 0;
    HIP_TRY(hipMemcpy(&word, v->status.p, sizeof word, hipMemcpyDeviceToHost));
    if (word & kStatusRngOverrun)
        return fail(HDR_ERR_CAPACITY, "a pixel drew past the random numbers its stream was set up for (RNG spill "
                                      "slot missing); the view's frames since its last reset are wrong");
    if (word & kStatusStackOverflow)
        return fail(HDR_ERR_CAPACITY, "a shadow ray's traversal stack in LDS overflowed; the view's frames since its "
                                      "last reset are wrong");
    return HDR_OK;
}

static hdr_status harvest(hdr_view *v, hdr_view::StageSlot &s) {
    if (!s.pending) return HDR_OK;
    HIP_TRY(hipEventSynchronize(s.fin));
    for (int k = 0; k < HDR_STAGE_COUNT; ++k) {
        if (!(s.ran & (1u << k))) continue;
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, s.beg[k], s.end[k]));
        v->stage_ms[k] += ms;
        v->stage_n[k] += 1;
    }
    s.pending = false;
    return HDR_OK;
}

extern "C" {

const char *hdr_last_error(void) { return g_err.c_str(); }
uint32_t hdr_abi_version(void) { return HDRGPU_ABI_VERSION; }

hdr_status hdr_context_create(int32_t device, hdr_context **out) {
    if (!out) return fail(HDR_ERR_INVALID_ARGUMENT, "out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(HDR_ERR_INVALID_ARGUMENT, "invalid device index");
    HIP_TRY(hipSetDevice(device));
    if (device < kMaxDevices) {  // the device address of hdr_rng_raw_noalloc's scratch, once per device
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch[device].load()) {
            void *p = nullptr;
            HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) return fail(HDR_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
            g_scratch[device].store(static_cast<uint64_t *>(p));
        }
    }
    auto *c = new (std::nothrow) hdr_context();
    if (!c) return fail(HDR_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->device = device;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus <= 0)
        c->cus = 256;
    *out = c;
    return HDR_OK;
}

hdr_status hdr_context_destroy(hdr_context *ctx) {
    delete ctx;
    return HDR_OK;
}

hdr_status hdr_context_set_cancel(hdr_context *ctx, int32_t cancelled) {
    if (!ctx) return fail(HDR_ERR_INVALID_ARGUMENT, "ctx is null");
    ctx->cancelled = cancelled;
    return HDR_OK;
}

namespace {

void swap_buf(DevBuf &a, DevBuf &b) {
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

// Per-material draw bound and translucency of a scene (plan_pipeline reads them per frame).
void material_traits(const hdr_scene_desc &d, hdr_scene *sc) {
    int draws = 3;
    bool translucent = false;
    for (int i = 0; i < d.num_materials; ++i) {
        if (d.materials[i].coat > 0.0f) draws += 3;
        if (d.materials[i].sheen > 0.0f) draws += 3;
        if (d.materials[i].opacity < 0.999f) translucent = true;
    }
    sc->max_sample_draws_per_vertex = std::min(draws, 9);
    sc->any_translucent = translucent;
}

// Places the host build's TLAS, instances, materials and lights in device memory: the TLAS nodes
// after the BLAS nodes (the node allocation grows when they do not fit its room), everything
// else in the state allocation (reused when large enough). Dome k keeps the texels and CDFs it
// has on the device when keep[k] (its record then comes from kept_rec[k]); the others are
// uploaded and their CDFs (DomeLight::Prepare) built on the device. The device is idle.
hdr_status place_state(hdr_scene *sc, const std::vector<char> &keep, const std::vector<GDome> &kept_rec,
                       std::vector<std::unique_ptr<DevBuf>> &old_dome_mem) {
    HostScene &h = sc->hs;
    const size_t tlasNodes = h.nodes.size() - size_t(h.blas_nodes);
    if (tlasNodes > sc->tlas_cap || !sc->nodes.p) {
        // TLAS of n instances: at most 4n - 1 records plus the odd-base pad (children are
        // allocated for leaves too), so this room holds every rebuild with as many instances
        const size_t cap = std::max(tlasNodes, size_t(4) * h.instances.size() + 2);
        DevBuf nn;
        HIP_TRY(nn.alloc((size_t(h.blas_nodes) + cap) * sizeof(GNode)));
        if (sc->nodes.p && h.blas_nodes > 0)
            HIP_TRY(hipMemcpy(nn.p, sc->nodes.p, size_t(h.blas_nodes) * sizeof(GNode), hipMemcpyDeviceToDevice));
        else if (h.blas_nodes > 0)
            HIP_TRY(hipMemcpy(nn.p, h.nodes.data(), size_t(h.blas_nodes) * sizeof(GNode), hipMemcpyHostToDevice));
        swap_buf(sc->nodes, nn);
        sc->tlas_cap = cap;
    }
    if (tlasNodes)
        HIP_TRY(hipMemcpy(static_cast<GNode *>(sc->nodes.p) + h.blas_nodes, h.nodes.data() + h.blas_nodes,
                          tlasNodes * sizeof(GNode), hipMemcpyHostToDevice));

    size_t off = 0;
    auto place = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t oTidx = place(h.tlas_idx.size() * sizeof(int));
    const size_t oTrank = place(h.tlas_rank.size() * sizeof(int));
    const size_t oInst = place(h.instances.size() * sizeof(GInstance));
    const size_t oMat = place(h.materials.size() * sizeof(GMaterial));
    const size_t oLights = place(h.lights.size() * sizeof(GLight));
    const size_t oDomes = place(h.domes.size() * sizeof(GDome));
    if (off > sc->state.n || !sc->state.p) {
        DevBuf ns;
        HIP_TRY(ns.alloc(off));
        swap_buf(sc->state, ns);
    }
    char *base = static_cast<char *>(sc->state.p);
    auto up = [&](size_t o, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    HIP_TRY(up(oTidx, h.tlas_idx.data(), h.tlas_idx.size() * sizeof(int)));
    HIP_TRY(up(oTrank, h.tlas_rank.data(), h.tlas_rank.size() * sizeof(int)));
    HIP_TRY(up(oInst, h.instances.data(), h.instances.size() * sizeof(GInstance)));
    HIP_TRY(up(oMat, h.materials.data(), h.materials.size() * sizeof(GMaterial)));
    HIP_TRY(up(oLights, h.lights.data(), h.lights.size() * sizeof(GLight)));

    std::vector<std::unique_ptr<DevBuf>> domeMem(h.domes.size());
    std::vector<GDome> gd(h.domes.size());
    std::vector<const float *> sinT(h.domes.size(), nullptr);  // per-row weights of new domes
    GDome *recs = reinterpret_cast<GDome *>(base + oDomes);
    for (size_t k = 0; k < h.domes.size(); ++k) {
        const HostScene::Dome &dm = h.domes[k];
        if (k < keep.size() && keep[k]) {
            domeMem[k] = std::move(old_dome_mem[k]);
            gd[k] = kept_rec[k];
            continue;
        }
        size_t doff = 0;
        auto dplace = [&](size_t bytes) {
            const size_t o = doff;
            doff += (bytes + 255) & ~size_t(255);
            return o;
        };
        const size_t oPix = dplace(dm.pixels.size() * sizeof(float));
        const size_t oRow = dplace(size_t(dm.h + 1) * sizeof(float));
        const size_t oCol = dplace(size_t(dm.h) * size_t(dm.w + 1) * sizeof(float));
        const size_t oSin = dplace(size_t(dm.h) * sizeof(float));
        domeMem[k].reset(new DevBuf());
        HIP_TRY(domeMem[k]->alloc(doff));
        char *db = static_cast<char *>(domeMem[k]->p);
        HIP_TRY(hipMemcpy(db + oPix, dm.pixels.data(), dm.pixels.size() * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(db + oSin, dm.sin_theta.data(), size_t(dm.h) * sizeof(float), hipMemcpyHostToDevice));
        sinT[k] = reinterpret_cast<const float *>(db + oSin);
        gd[k].pixels = reinterpret_cast<const float *>(db + oPix);
        gd[k].row_cdf = reinterpret_cast<const float *>(db + oRow);
        gd[k].col_cdf = reinterpret_cast<const float *>(db + oCol);
        gd[k].w = dm.w;
        gd[k].h = dm.h;
        gd[k].total_lum = 0.0f;  // set by dome_cdf_total_kernel
        gd[k].pdf_scale = 0.0f;
    }
    if (!gd.empty()) HIP_TRY(up(oDomes, gd.data(), gd.size() * sizeof(GDome)));
    // DomeLight::Prepare on the device: column CDFs one row per thread, then the row CDF
    for (size_t k = 0; k < h.domes.size(); ++k) {
        if (k < keep.size() && keep[k]) continue;
        const HostScene::Dome &dm = h.domes[k];
        float *row = const_cast<float *>(gd[k].row_cdf);
        float *col = const_cast<float *>(gd[k].col_cdf);
        hipLaunchKernelGGL(dome_cdf_rows_kernel, dim3((dm.h + 63) / 64), dim3(64), 0, 0, gd[k].pixels, sinT[k],
                           dm.w, dm.h, row, col);
        hipLaunchKernelGGL(dome_cdf_total_kernel, dim3(1), dim3(1), 0, 0, row, dm.w, dm.h, recs + k);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    old_dome_mem.clear();
    sc->dome_mem = std::move(domeMem);

    DeviceScene &ds = sc->ds;
    ds.nodes = static_cast<const GNode *>(sc->nodes.p);
    ds.tlas_idx = reinterpret_cast<const int *>(base + oTidx);
    ds.tlas_rank = reinterpret_cast<const int *>(base + oTrank);
    ds.instances = reinterpret_cast<const GInstance *>(base + oInst);
    ds.materials = reinterpret_cast<const GMaterial *>(base + oMat);
    ds.lights = reinterpret_cast<const GLight *>(base + oLights);
    ds.domes = recs;
    ds.tlas_root = h.tlas_root;
    ds.num_instances = int(h.instances.size());
    ds.num_materials = h.num_materials;
    ds.num_lights = int(h.lights.size());
    ds.environment = h.environment;
    ds.env_dome = h.env_dome;
    ds.has_mesh_lights = h.has_mesh_lights;
    ds.has_emission = h.has_emission;
    uint64_t bytes = sc->nodes.n + sc->geo.n + sc->state.n;
    for (const auto &m : sc->dome_mem) bytes += m->n;
    sc->bytes = bytes;
    static std::atomic<uint64_t> nextSceneId{1};
    sc->id = nextSceneId.fetch_add(1);
    return HDR_OK;
}

}  // namespace

hdr_status hdr_scene_upload(hdr_context *ctx, const hdr_scene_desc *desc, hdr_scene **out) {
    if (!ctx || !desc || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_ptr<hdr_scene> sc(new (std::nothrow) hdr_scene());
    if (!sc) return fail(HDR_ERR_OUT_OF_MEMORY, "host allocation failed");
    sc->ctx = ctx;
    const std::string err = build_host_scene(*desc, sc->hs);
    if (!err.empty()) return fail(HDR_ERR_INVALID_ARGUMENT, err);
    material_traits(*desc, sc.get());
    HostScene &h = sc->hs;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t triBytes = (h.tris.size() * sizeof(GTri) + 255) & ~size_t(255);
    HIP_TRY(sc->geo.alloc(triBytes + h.attrs.size() * sizeof(GTriAttr)));
    char *g = static_cast<char *>(sc->geo.p);
    if (!h.tris.empty()) {
        HIP_TRY(hipMemcpy(g, h.tris.data(), h.tris.size() * sizeof(GTri), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g + triBytes, h.attrs.data(), h.attrs.size() * sizeof(GTriAttr), hipMemcpyHostToDevice));
    }
    sc->ds.tris = reinterpret_cast<const GTri *>(g);
    sc->ds.attrs = reinterpret_cast<const GTriAttr *>(g + triBytes);
    std::vector<std::unique_ptr<DevBuf>> none;
    const hdr_status st = place_state(sc.get(), {}, {}, none);
    if (st != HDR_OK) return st;
    std::vector<GTri>().swap(h.tris);
    std::vector<GTriAttr>().swap(h.attrs);
    *out = sc.release();
    return HDR_OK;
}

hdr_status hdr_scene_update(hdr_scene *scene, const hdr_scene_desc *desc) {
    if (!scene || !desc) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(scene->ctx->device));
    // frames in flight read the state this call rewrites
    HIP_TRY(hipDeviceSynchronize());
    std::vector<HostScene::Dome> oldDomes;
    oldDomes.swap(scene->hs.domes);
    const std::string err = rebuild_host_state(*desc, scene->hs);
    if (!err.empty()) {
        scene->hs.domes.swap(oldDomes);
        return fail(HDR_ERR_INVALID_ARGUMENT, err);
    }
    // a dome whose texels are unchanged keeps its device texels, CDFs and record
    const std::vector<HostScene::Dome> &nd = scene->hs.domes;
    std::vector<char> keep(nd.size(), 0);
    std::vector<GDome> keptRec(nd.size());
    for (size_t k = 0; k < nd.size() && k < oldDomes.size(); ++k) {
        keep[k] = nd[k].w == oldDomes[k].w && nd[k].h == oldDomes[k].h && nd[k].pixels == oldDomes[k].pixels;
        if (keep[k]) HIP_TRY(hipMemcpy(&keptRec[k], scene->ds.domes + k, sizeof(GDome), hipMemcpyDeviceToHost));
    }
    material_traits(*desc, scene);
    return place_state(scene, keep, keptRec, scene->dome_mem);
}

hdr_status hdr_scene_destroy(hdr_scene *scene) {
    if (scene) (void)hipSetDevice(scene->ctx->device);
    delete scene;
    return HDR_OK;
}

hdr_status hdr_scene_read_dome(const hdr_scene *scene, int32_t light, float *row_cdf, float *col_cdf,
                                float *total_lum) {
    if (!scene || !row_cdf || !col_cdf || !total_lum) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (light < 0 || light >= int32_t(scene->hs.lights.size()) || scene->hs.lights[size_t(light)].dome < 0)
        return fail(HDR_ERR_INVALID_ARGUMENT, "not a dome light with texels");
    HIP_TRY(hipSetDevice(scene->ctx->device));
    const int k = scene->hs.lights[size_t(light)].dome;
    GDome gd;
    HIP_TRY(hipMemcpy(&gd, scene->ds.domes + k, sizeof(GDome), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(row_cdf, gd.row_cdf, size_t(gd.h + 1) * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(col_cdf, gd.col_cdf, size_t(gd.h) * size_t(gd.w + 1) * sizeof(float), hipMemcpyDeviceToHost));
    *total_lum = gd.total_lum;
    return HDR_OK;
}

hdr_status hdr_scene_device_bytes(const hdr_scene *scene, uint64_t *out) {
    if (!scene || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    *out = scene->bytes;
    return HDR_OK;
}

hdr_status hdr_view_create(hdr_context *ctx, int32_t width, int32_t height, hdr_view **out) {
    if (!ctx || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (width <= 0 || height <= 0 || int64_t(width) * height >= (int64_t(1) << 24))
        return fail(HDR_ERR_INVALID_ARGUMENT, "invalid view size (at most 2^24 - 1 pixels)");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<hdr_view> v(new (std::nothrow) hdr_view());
    if (!v) return fail(HDR_ERR_OUT_OF_MEMORY, "host allocation failed");
    v->ctx = ctx;
    v->W = width;
    v->H = height;
    const size_t P = size_t(width) * height;
    HIP_TRY(v->res.alloc(P * sizeof(hdr_reservoir)));
    HIP_TRY(v->accum.alloc(P * sizeof(float4)));
    HIP_TRY(v->color.alloc(P * sizeof(float4)));
    HIP_TRY(v->counters.alloc(sizeof(hdr_frame_counters)));
    HIP_TRY(v->status.alloc(sizeof(unsigned int)));
    HIP_TRY(v->gbuf.alloc(2 * P * sizeof(GRec)));  // two halves, by frame parity
    HIP_TRY(v->qcount.alloc(2 * sizeof(unsigned int)));  // query count, persistent claim counter
    HIP_TRY(v->dargs.alloc(hdr_view::kRing * sizeof(FrameArgs)));
    HIP_TRY(v->dtile.alloc(2 * sizeof(int)));
    HIP_TRY(v->hargs.alloc(hdr_view::kRing));
    {
        const int tiles[2] = {0, ((width + 7) / 8) * ((height + 7) / 8)};
        HIP_TRY(hipMemcpy(v->dtile.p, tiles, sizeof tiles, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&v->gstream, hipStreamNonBlocking));
    for (hipEvent_t &e : v->shade_done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&v->gbuf_done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&v->serial, hipEventDisableTiming));
    for (auto &slot : v->ring) {
        for (auto &e : slot.beg) HIP_TRY(hipEventCreate(&e));
        for (auto &e : slot.end) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventCreate(&slot.fin));
    }
    HIP_TRY(hipMemsetAsync(v->res.p, 0, v->res.n, v->stream));
    HIP_TRY(hipMemsetAsync(v->accum.p, 0, v->accum.n, v->stream));
    HIP_TRY(hipMemsetAsync(v->color.p, 0, v->color.n, v->stream));
    HIP_TRY(hipMemsetAsync(v->counters.p, 0, v->counters.n, v->stream));
    HIP_TRY(hipMemsetAsync(v->status.p, 0, v->status.n, v->stream));
    HIP_TRY(hipStreamSynchronize(v->stream));
    *out = v.release();
    return HDR_OK;
}

hdr_status hdr_view_destroy(hdr_view *view) {
    if (!view) return HDR_OK;
    (void)hipSetDevice(view->ctx->device);
    if (view->stream) (void)hipStreamSynchronize(view->stream);
    if (view->gstream) (void)hipStreamSynchronize(view->gstream);
    for (auto &slot : view->ring) {
        for (auto &e : slot.beg)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : slot.end)
            if (e) (void)hipEventDestroy(e);
        if (slot.fin) (void)hipEventDestroy(slot.fin);
    }
    for (hipEvent_t e : {view->shade_done[0], view->shade_done[1], view->gbuf_done, view->serial})
        if (e) (void)hipEventDestroy(e);
    if (view->gstream) (void)hipStreamDestroy(view->gstream);
    if (view->stream) (void)hipStreamDestroy(view->stream);
    delete view;
    return HDR_OK;
}

hdr_status hdr_view_reset(hdr_view *view) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipMemsetAsync(view->res.p, 0, view->res.n, view->stream));
    HIP_TRY(hipMemsetAsync(view->accum.p, 0, view->accum.n, view->stream));
    HIP_TRY(hipMemsetAsync(view->status.p, 0, view->status.n, view->stream));
    if (view->lum_sum.p) {
        HIP_TRY(hipMemsetAsync(view->lum_sum.p, 0, view->lum_sum.n, view->stream));
        HIP_TRY(hipMemsetAsync(view->lum_sq.p, 0, view->lum_sq.n, view->stream));
    }
    view->frames = 0;
    view->samples = 0;
    view->hist_cam = false;
    view->hist_records = false;
    return HDR_OK;
}

hdr_status hdr_view_reset_accumulation(hdr_view *view) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipMemsetAsync(view->accum.p, 0, view->accum.n, view->stream));
    if (view->lum_sum.p) {
        HIP_TRY(hipMemsetAsync(view->lum_sum.p, 0, view->lum_sum.n, view->stream));
        HIP_TRY(hipMemsetAsync(view->lum_sq.p, 0, view->lum_sq.n, view->stream));
    }
    view->samples = 0;
    return HDR_OK;
}

hdr_status hdr_view_device_bytes(const hdr_view *v, uint64_t *out) {
    if (!v || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    *out = v->lum_sum.n + v->lum_sq.n + v->variance.n + v->resolved.n + v->res.n + v->accum.n + v->color.n + v->frame.n + v->dbg.n + v->depth.n + v->albedo.n + v->normal.n +
           v->counters.n + v->spill.n + v->gbuf.n + v->rgb2.n + v->flags.n + v->queue.n + v->qcount.n + v->res_t.n +
           v->saux.n + v->wstate.n + v->whit.n + v->wray.n + v->wcq.n + v->walive.n + v->wcnt.n + v->wesc.n +
           v->wesq.n + v->post.n + v->rp.n;
    return HDR_OK;
}

hdr_status hdr_view_set_counting(hdr_view *view, int32_t enable) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    view->counting = enable != 0;
    return HDR_OK;
}

static hdr_status ensure(DevBuf &b, size_t bytes) {
    if (b.n >= bytes && b.p) return HDR_OK;
    if (b.p) {
        (void)hipFree(b.p);
        b.p = nullptr;
    }
    HIP_TRY(b.alloc(bytes));
    return HDR_OK;
}

// Launch plan of one view's frame.
struct FramePlan {
    bool deferred = false;  // the deferred-shadow path (else the integration megakernel)
    bool lean_pt = false;   // the inlined path tracer (pathlean.h) instead of the megakernel
    bool spatial = false;   // spatial reuse after the temporal pass (deferred and wavefront paths)
    int spatial_nd = 0;     // its domains per pixel, rounded up to a compiled size (spatial_domains)
    bool wave = false;      // the wavefront path (RIS at every vertex of deep paths, or the path tracer; wavefront.h)
    bool wave_lean0 = false;  // its primary vertex cannot draw past RNG output 310 (lean stream there)
    bool wave_traced = false; // BSDF-sampled candidates can see an emitter: bounce and escape rays per vertex
    bool spill = false;       // RNG spill slots (a pixel can draw past output 311)
    size_t escapes = 0;       // traced frames: escape-queue capacity (rays)
    int depth = 0;          // path vertices (max_depth)
    size_t lds = 0;         // dynamic LDS of the shading / integration kernel
    unsigned tiles = 0;     // 8x8 tiles of the view
    int prep_window = 0;      // wavefront: draws a shading launch can make on a path's stream (0: no
                              // stream reaches output 624, no slot preparation)
    int prep_vertex = 0;      // the first shaded vertex whose launch can reach output 624
    bool gbuf_reuse = false;  // the view's latest G-buffer is this frame's (no G-buffer launch)
    int gbuf_half = 0;        // the G-buffer half this frame reads (and, without reuse, writes)
};

// Device bytes of the buffers a frame's plan needs beyond those every view holds (0: not needed).
struct FrameNeeds {
    size_t spill = 0, frame = 0, dbg = 0, res_t = 0, saux = 0, rgb2 = 0, flags = 0, queue = 0;
    size_t wstate = 0, whit = 0, wesc = 0;            // the wavefront's per-view state
    size_t wray = 0, wcq = 0, walive = 0, wesq = 0;   // and its queues (a batch's are shared)
    size_t rp = 0;                                     // reprojected temporal reuse: gathered records
    size_t sum() const {
        return spill + frame + dbg + res_t + saux + rgb2 + flags + queue + wstate + whit + wesc + wray + wcq +
               walive + wesq + rp;
    }
};

// Smallest view that renders RIS without skipVisibility on the wavefront (plan_pipeline).
constexpr size_t kVisWavefrontPixels = size_t(1) << 17;
// Smallest view that renders the path tracer on the wavefront (plan_pipeline). Measured on the
// many-lights scene, lean path tracer against wavefront, ms per frame at depth 4 / depth 8:
// 256^2 1.72 / 2.45 against 2.26 / 4.54, 512^2 2.04 / 3.06 against 2.63 / 5.10, 1280x720
// 3.93 / 5.49 against 3.58 / 6.20, 1920x1080 6.84 / 9.28 against 5.47 / 8.36 (identical images).
constexpr size_t kPtWavefrontPixels = size_t(1) << 20;
// and from which it also beats the lean path tracer's multi-frame launches (hdr_render_frames):
// 1280x720 3.11 ms per frame in 16-frame lean launches against 3.58 on the wavefront, 1920x1080
// 6.01 against 5.47 (depth 4).
constexpr size_t kPtWavefrontMultiPixels = size_t(3) << 19;
// The path tracer's wavefront is off: a serving run of 1080p path-traced requests beside batched
// ReSTIR views ended in a GPU memory fault whose cause reading the kernels did not show.
constexpr bool kPtWavefront = false;

// The escape queue of `cap` entries: two float4 each, then a float (tmax) each.
static size_t escape_queue_bytes(size_t cap) {
    cap = std::max<size_t>(1, cap);
    return cap * (2 * sizeof(float4) + sizeof(float));
}

// Device bytes every view holds from hdr_view_create on: reservoirs, accumulation, colour, the two
// G-buffer halves, counters, status word, argument ring and tile table.
static size_t view_base_bytes(size_t P) {
    return P * (sizeof(hdr_reservoir) + 2 * sizeof(float4) + 2 * sizeof(GRec)) + sizeof(hdr_frame_counters) +
           2 * sizeof(unsigned int) + 16 * sizeof(FrameArgs) + 2 * sizeof(int);
}

// Validates a pipeline on a scene and decides the frame's kernel path and the buffers it needs.
// Host-only: no device call; hdr_view_plan_bytes and the server's admission use it too.
// multi: the frame could share a launch with others (hdr_render_frames): a path-traced frame then
// keeps the lean path tracer, which takes several frames per launch, below
// kPtWavefrontMultiPixels, and takes the wavefront frame by frame from there.
static hdr_status plan_pipeline(const hdr_scene *scene, int W, int H, const hdr_pipeline_desc *pipe, bool debug,
                                FramePlan &plan, FrameNeeds &need, bool multi = false) {
    if (!scene || !pipe) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (pipe->struct_size != sizeof(hdr_pipeline_desc))
        return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_pipeline_desc size");
    if (pipe->direct_light != HDR_DIRECT_MIS && pipe->direct_light != HDR_DIRECT_RIS)
        return fail(HDR_ERR_INVALID_ARGUMENT, "unknown direct_light kind");
    if (pipe->max_depth < 1 || pipe->max_depth > 1024) return fail(HDR_ERR_INVALID_ARGUMENT, "max_depth");
    if (pipe->candidate_count < 1 || pipe->candidate_count > 64)
        return fail(HDR_ERR_INVALID_ARGUMENT, "candidate_count must be in [1, 64]");
    if (pipe->spatial_count < 0 || pipe->spatial_count > kMaxSpatial)
        return fail(HDR_ERR_INVALID_ARGUMENT, "spatial_count must be in [0, 8]");
    if (pipe->spatial_count > 0 && !(pipe->spatial_radius >= 1.0f && pipe->spatial_radius <= 4096.0f))
        return fail(HDR_ERR_INVALID_ARGUMENT, "spatial_radius must be in [1, 4096] pixels");
    const size_t P = size_t(W) * H;
    plan = FramePlan();
    need = FrameNeeds();
    // technique split (ris_direct_light_integrator.cpp:93-100); the environment is always one of the
    // lights, so the light sampler covers it and no separate sky technique exists
    const bool ris = pipe->direct_light == HDR_DIRECT_RIS;
    const int M = pipe->candidate_count;
    const int nNee = std::max(1, M / 2), nBsdf = std::max(0, M - nNee), nCand = nNee + nBsdf;
    const bool traceGenerated = scene->ds.has_mesh_lights || scene->ds.environment >= 0;
    const bool traceLastConnection = scene->ds.has_emission || scene->ds.environment >= 0;
    // upper bound of RNG draws per pixel; past 311 the engine needs its spill slots
    int bound = 5, bound0 = 5;  // bound0: through the primary vertex only
    int perVertex = 0;
    {
        const int sb = scene->max_sample_draws_per_vertex;
        perVertex = 1 + sb + 1;  // opacity, bounce sample, roulette
        if (!ris) perVertex += 3;
        else perVertex += nNee * 3 + std::max(0, nBsdf - 1) * sb + nCand + 1;
        bound += pipe->max_depth * perVertex;
        bound0 += perVertex;
    }
    plan.spill = scene->any_translucent || bound >= kMtN;
    if (plan.spill) need.spill = P * 2 * kMtN * sizeof(uint64_t);
    if (pipe->enable_firefly_filter) need.frame = P * sizeof(float4);
    if (debug) need.dbg = P * sizeof(hdr_pixel_debug);
    // The deferred-shadow path applies when the winner's shadow ray is the only ray after the
    // G-buffer: RIS at depth 1, skipVisibility, no translucent material (no pass-through rays),
    // nothing BSDF-sampled can reach an emitter (no emissive material, mesh light or environment),
    // and no pixel draws past RNG output 310 (the lean stream of that path has no spill slot).
    const bool skipVis = pipe->skip_visibility != 0;
    const bool deferred = ris && pipe->max_depth == 1 && skipVis && !scene->any_translucent && !traceGenerated &&
                          !traceLastConnection && bound < kMtN - 1;
    // The wavefront path: the same conditions at any depth and RNG range (frames the deferred path
    // does not take). With an environment or emissive materials, BSDF-sampled candidates can see an
    // emitter: the wavefront then emits each vertex's bounce ray and generated candidates' escape
    // rays before its RIS (traced frames; at most 32 candidates, one bit each in the escape mask).
    // Without skipVisibility the NEE candidates' shadow rays are emitted the same way (their bits
    // are the NEE candidates' indices, below the generated candidates').
    // Views under kVisWavefrontPixels keep visibility frames on the megakernel: a 256 x 256 frame of
    // RIS M = 16 at depth 4 takes 2.8 ms there and 3.4 ms as a wavefront (its ~1000 waves leave the
    // GPU's SIMDs underfilled in every one of the wavefront's launches), while 384 x 384 takes 4.7 ms
    // as a wavefront and 1080p 20 ms (187 ms on the megakernel). Both give the same results.
    const bool needTraced = traceGenerated || traceLastConnection || !skipVis;
    // With mesh lights a generated candidate's closest hit decides it: the traced shading launches
    // trace those rays themselves (gen_traced_weight) instead of emitting escape rays.
    // Visibility frames with mesh lights stay on the megakernel, which measured faster there (1080p
    // ReSTIR M = 8 at depth 2: 20.3 ms against 21.5 ms as a wavefront).
    // The path tracer (MIS) takes the untraced wavefront on views from kPtWavefrontPixels when its
    // BSDF rays can reach no emitter (mis_nee_wave: the connection then adds nothing).
    // Switched off (kPtWavefront): a serving run with it on ended in a GPU memory fault whose cause
    // is not shown; path-traced frames take the lean path tracer and the megakernel meanwhile.
    const bool ptWave = kPtWavefront && !ris && !scene->any_translucent && !traceLastConnection &&
                        !scene->ds.has_mesh_lights && P >= (multi ? kPtWavefrontMultiPixels : kPtWavefrontPixels);
    const bool wave = (!deferred && ris && !scene->any_translucent && !(needTraced && nCand > 32) &&
                       (skipVis || (P >= kVisWavefrontPixels && !scene->ds.has_mesh_lights))) ||
                      ptWave;
    const bool waveTraced = wave && ris && needTraced;
    // Spatial reuse (an extension) runs on the deferred and wavefront paths with temporal reservoirs.
    const bool spatial = pipe->spatial_count > 0;
    if (spatial && !(ris && (deferred || wave) && pipe->use_reservoir && skipVis && !scene->ds.has_mesh_lights))
        return fail(HDR_ERR_UNSUPPORTED,
                    "spatial reuse needs ReSTIR (use_reservoir) with skip_visibility, opaque materials and no mesh "
                    "light");
    if (pipe->temporal_reprojection) {
        if (!spatial)
            return fail(HDR_ERR_UNSUPPORTED, "temporal_reprojection runs in the spatial pass (spatial_count >= 1)");
        if (pipe->spatial_count > kMaxSpatial - 1)
            return fail(HDR_ERR_INVALID_ARGUMENT, "spatial_count must be in [1, 7] with temporal_reprojection");
        need.rp = P * (2 * sizeof(GRec) + sizeof(float4) + sizeof(hdr_reservoir));
    }
    if (spatial) {
        need.res_t = P * sizeof(hdr_reservoir);
        need.saux = P * 2 * sizeof(float4);
    }
    if (deferred) {
        need.rgb2 = P * 2 * sizeof(float4);
        need.flags = P * sizeof(uint32_t);
        need.queue = P * sizeof(ShadowQuery);
    }
    if (wave) {
        need.wstate = P * kWaveFields * sizeof(float4);
        // traced frames: hit records of odd and even vertices, escape masks and the escape queue
        need.whit = (waveTraced ? 2 : 1) * P * sizeof(GRec);
        if (waveTraced) {  // per vertex: the generated candidates' escape rays, the NEE candidates' shadow rays
            const bool genEscapes = traceGenerated && !scene->ds.has_mesh_lights;  // else traced inline
            plan.escapes = (genEscapes ? P * size_t(nCand - nNee) : 0) + (skipVis ? 0 : P * size_t(nNee));
            need.wesc = P * sizeof(unsigned int);
            need.wesq = escape_queue_bytes(plan.escapes);
        }
        need.wray = P * 2 * sizeof(float4);
        need.queue = P * sizeof(ShadowQuery);
        need.wcq = P * sizeof(float4);
        need.walive = 2 * P * sizeof(unsigned int) + kWaveCounts * sizeof(unsigned int);  // two parities + counts
        if (spatial) {  // per pixel: emission + reused direct, visible / occluded; visibility + slot
            need.rgb2 = P * 2 * sizeof(float4);
            need.flags = P * sizeof(uint32_t);
        }
    }
    // candidate weights in LDS: the NEE candidates on the deferred path (its BSDF candidates carry
    // no weight), every candidate on the general path
    plan.lds = size_t(!ris ? 1 : ((deferred || (wave && !waveTraced)) ? nNee : nCand)) * kLanes * sizeof(double);
    plan.deferred = deferred;
    plan.wave = wave;
    plan.wave_lean0 = wave && bound0 < kMtN - 1;
    if (wave && bound >= 2 * kMtN) {  // streams may reach output 624: slots prepared ahead of it
        // a shading launch draws a vertex on the stream and, on traced frames, the next vertex's
        // generation on its copy; vertex u's stream starts at most at 5 + u * perVertex
        plan.prep_window = (waveTraced ? 2 : 1) * perVertex;
        plan.prep_vertex = std::max(0, (2 * kMtN - 5 - plan.prep_window + perVertex - 1) / perVertex);
    }
    plan.wave_traced = waveTraced;
    plan.depth = pipe->max_depth;
    plan.lean_pt = !deferred && !wave && !ris && !scene->any_translucent && bound < kMtN - 1;
    plan.spatial = spatial;
    plan.spatial_nd = spatial ? spatial_domains(pipe->spatial_count + (pipe->temporal_reprojection ? 1 : 0)) : 0;
    plan.tiles = unsigned(((W + 7) / 8) * ((H + 7) / 8));
    return HDR_OK;
}

// Validates a frame, makes sure the view's buffers for it exist and fills its FrameArgs.
// shared_queues: a batch substitutes its own wavefront queues, so the view's are not allocated.
static hdr_status prepare_frame(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                                const hdr_pipeline_desc *pipe, bool debug, FrameArgs &a, FramePlan &plan,
                                bool shared_queues = false, bool multi = false) {
    if (!view || !scene || !frame || !pipe) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (frame->struct_size != sizeof(hdr_frame_desc)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_frame_desc size");
    if (view->ctx->cancelled) return fail(HDR_ERR_CANCELLED, "render cancelled");
    const int W = view->W, H = view->H;
    const size_t P = size_t(W) * H;
    FrameNeeds need;
    if (hdr_status st = plan_pipeline(scene, W, H, pipe, debug, plan, need, multi)) return st;
    HIP_TRY(hipSetDevice(view->ctx->device));

    std::memset(&a, 0, sizeof a);
    a.s = scene->ds;
    double m[4][4], inv[4][4];
    for (int i = 0; i < 16; ++i) m[i / 4][i % 4] = frame->view[i];
    invert4(m, inv);
    for (int i = 0; i < 16; ++i) a.inv_view[i] = inv[i / 4][i % 4];
    for (int i = 0; i < 16; ++i) m[i / 4][i % 4] = frame->proj[i];
    invert4(m, inv);
    for (int i = 0; i < 16; ++i) a.inv_proj[i] = inv[i / 4][i % 4];
    a.W = W;
    a.H = H;
    a.frame_index = frame->frame_index;
    a.has_cam = frame->has_camera_params != 0;
    a.cam = frame->camera;
    if (a.has_cam && a.cam.enable_exposure) {  // GetExposureMultiplier, camera_params.h:23-30
        const float safeF = std::max(a.cam.f_stop, 0.1f);
        a.exposure_mul = (a.cam.iso / 100.0f) * a.cam.shutter_speed / (safeF * safeF) * 100.0f;
    } else {
        a.exposure_mul = 1.0f;
    }
    a.direct = pipe->direct_light;
    a.use_res = (pipe->direct_light == HDR_DIRECT_RIS && pipe->use_reservoir != 0) ? 1 : 0;  // reservoirs are RIS state
    a.skip_vis = pipe->skip_visibility != 0;
    a.max_depth = pipe->max_depth;
    a.max_refl = pipe->max_reflection_bounces;
    a.max_refr = pipe->max_refraction_bounces;
    a.enable_ss = pipe->enable_subsurface != 0;
    a.ibl_bg = pipe->render_ibl_background != 0;
    a.roulette = pipe->roulette_aggressiveness;
    const int M = pipe->candidate_count;
    a.n_nee = std::max(1, M / 2);
    a.n_sky = 0;
    a.n_bsdf = std::max(0, M - a.n_nee - a.n_sky);
    a.n_cand = a.n_nee + a.n_sky + a.n_bsdf;
    a.trace_generated = (scene->ds.has_mesh_lights || scene->ds.environment >= 0) ? 1 : 0;
    a.trace_last_connection = (scene->ds.has_emission || scene->ds.environment >= 0) ? 1 : 0;
    a.gen_inline = (plan.wave && scene->ds.has_mesh_lights) ? 1 : 0;
    a.light_select_pdf = scene->ds.num_lights > 0 ? 1.0f / float(scene->ds.num_lights) : 0.0f;
    a.firefly = pipe->enable_firefly_filter != 0;
    a.inv_n = 1.0f / float(frame->frame_index + 1);
    a.frames = 1;
    a.res = static_cast<hdr_reservoir *>(view->res.p);
    a.status = static_cast<unsigned int *>(view->status.p);
    a.accum = static_cast<float4 *>(view->accum.p);
    a.color = static_cast<float4 *>(view->color.p);
    a.lum_sum = static_cast<double *>(view->lum_sum.p);  // null unless the metrics are on
    a.lum_sq = static_cast<double *>(view->lum_sq.p);

    // HDRGPU_TEST_WITHHOLD_SPILL=1: a test hook that withholds the spill slots a frame needs, to check
    // that the kernels report the overrun (their draws past the slot return 0; no memory is touched)
    const char *withholdEnv = std::getenv("HDRGPU_TEST_WITHHOLD_SPILL");
    const bool withholdSpill = withholdEnv && withholdEnv[0] == '1';
    if (need.spill && !withholdSpill) {
        if (hdr_status s = ensure(view->spill, need.spill)) return s;
        a.spill = static_cast<uint64_t *>(view->spill.p);
    }
    if (need.frame) {
        if (hdr_status s = ensure(view->frame, need.frame)) return s;
        a.frame_rgb = static_cast<float4 *>(view->frame.p);
    }
    if (need.dbg) {
        if (hdr_status s = ensure(view->dbg, need.dbg)) return s;
        a.dbg = static_cast<hdr_pixel_debug *>(view->dbg.p);
    }
    // AOVs are written every frame once allocated (first read allocates them)
    if (view->depth.p) {
        a.aov_depth = static_cast<float *>(view->depth.p);
        a.aov_albedo = static_cast<float *>(view->albedo.p);
        a.aov_normal = static_cast<float *>(view->normal.p);
    }
    {  // this frame's G-buffer: the latest one if it was computed for the same rays, else a new one
        hdr_view::GbufKey key;
        std::memset(&key, 0, sizeof key);
        key.scene_id = scene->id;
        std::memcpy(key.view, frame->view, sizeof key.view);
        std::memcpy(key.proj, frame->proj, sizeof key.proj);
        key.has_cam = frame->has_camera_params != 0;
        if (key.has_cam) key.cam = frame->camera;
        key.aovs = view->depth.p != nullptr;
        const bool dof = key.has_cam && frame->camera.enable_dof;
        plan.gbuf_reuse = view->gvalid && !dof && !view->counting && std::memcmp(&key, &view->gkey, sizeof key) == 0;
        plan.gbuf_half = plan.gbuf_reuse ? view->gcur : (view->gcur ^ 1);
        if (!plan.gbuf_reuse) {  // recorded here, committed once the frame is queued
            view->gkey = key;
            view->gvalid = false;
            view->gnext_valid = !dof;
        }
    }
    a.gbuf = static_cast<GRec *>(view->gbuf.p) + size_t(plan.gbuf_half) * P;
    if (plan.spatial) {
        if (hdr_status st = ensure(view->res_t, need.res_t)) return st;
        if (hdr_status st = ensure(view->saux, need.saux)) return st;
        a.res_t = static_cast<hdr_reservoir *>(view->res_t.p);
        a.spatial_aux = static_cast<float4 *>(view->saux.p);
        a.spatial_count = pipe->spatial_count;
        a.spatial_radius = int(pipe->spatial_radius);
    }
    if (plan.spatial && pipe->temporal_reprojection) {  // reprojected temporal reuse on a camera move
        const bool moved = view->hist_cam && (std::memcmp(frame->view, view->hist_view, sizeof view->hist_view) != 0 ||
                                              std::memcmp(frame->proj, view->hist_proj, sizeof view->hist_proj) != 0);
        const bool distortion = a.has_cam && a.cam.lens_distortion != 0.0f;
        a.temporal_skip = moved ? 1 : 0;
        a.reproject = (moved && view->hist_records && view->hist_scene == scene->id && !distortion) ? 1 : 0;
        if (hdr_status st = ensure(view->rp, need.rp)) return st;
        a.hist_g = static_cast<GRec *>(view->rp.p);  // kept by every spatial pass of the pipeline
        a.rp_g = a.hist_g + P;
        a.rp_aux = reinterpret_cast<float4 *>(a.rp_g + P);
        a.rp_res = reinterpret_cast<hdr_reservoir *>(a.rp_aux + P);
        if (a.reproject) {
            std::memcpy(a.prev_view, view->hist_view, sizeof a.prev_view);
            std::memcpy(a.prev_proj, view->hist_proj, sizeof a.prev_proj);
            for (int k = 0; k < 3; ++k) a.prev_origin[k] = view->hist_origin[k];
        }
    }
    if (need.rgb2) {
        if (hdr_status st = ensure(view->rgb2, need.rgb2)) return st;
        if (hdr_status st = ensure(view->flags, need.flags)) return st;
        a.rgb2 = static_cast<float4 *>(view->rgb2.p);
        a.flags = static_cast<uint32_t *>(view->flags.p);
    }
    if (plan.deferred) {
        if (hdr_status st = ensure(view->queue, need.queue)) return st;
        a.queue = static_cast<ShadowQuery *>(view->queue.p);
        a.qcount = static_cast<unsigned int *>(view->qcount.p);
    }
    if (plan.wave) {
        if (hdr_status st = ensure(view->wstate, need.wstate)) return st;
        if (hdr_status st = ensure(view->whit, need.whit)) return st;
        if (plan.wave_traced) {
            if (hdr_status st = ensure(view->wesc, need.wesc)) return st;
            a.whit2 = static_cast<GRec *>(view->whit.p) + P;
            a.wesc = static_cast<unsigned int *>(view->wesc.p);
        }
        a.wstate = static_cast<float4 *>(view->wstate.p);
        a.whit = static_cast<GRec *>(view->whit.p);
        if (!shared_queues) {  // the view's own queues
            if (hdr_status st = ensure(view->wray, need.wray)) return st;
            if (hdr_status st = ensure(view->queue, need.queue)) return st;
            if (hdr_status st = ensure(view->wcq, need.wcq)) return st;
            if (hdr_status st = ensure(view->walive, 2 * P * sizeof(unsigned int))) return st;  // two parities
            if (hdr_status st = ensure(view->wcnt, kWaveCounts * sizeof(unsigned int))) return st;
            if (plan.wave_traced) {
                if (hdr_status st = ensure(view->wesq, need.wesq)) return st;
                a.wesq = static_cast<float4 *>(view->wesq.p);
                a.wesq_t = reinterpret_cast<float *>(a.wesq + 2 * std::max<size_t>(1, plan.escapes));
            }
            a.wray = static_cast<float4 *>(view->wray.p);
            a.queue = static_cast<ShadowQuery *>(view->queue.p);
            a.wcq = static_cast<float4 *>(view->wcq.p);
            a.walive = static_cast<unsigned int *>(view->walive.p);
            a.walive2 = a.walive + P;
            a.wcnt = static_cast<unsigned int *>(view->wcnt.p);
        }
    }
    a.view = 0;
    return HDR_OK;
}

hdr_status hdr_view_plan_bytes(const hdr_scene *scene, int32_t width, int32_t height, const hdr_pipeline_desc *pipe,
                               uint64_t *bytes) {
    if (!bytes) return fail(HDR_ERR_INVALID_ARGUMENT, "bytes is null");
    if (width <= 0 || height <= 0 || int64_t(width) * height >= (int64_t(1) << 24))
        return fail(HDR_ERR_INVALID_ARGUMENT, "invalid view size (at most 2^24 - 1 pixels)");
    FramePlan plan;
    FrameNeeds need;
    if (hdr_status st = plan_pipeline(scene, width, height, pipe, false, plan, need)) return st;
    *bytes = uint64_t(view_base_bytes(size_t(width) * size_t(height)) + need.sum());
    return HDR_OK;
}

// The frame's camera, for the next frame's move test and reprojection; `records`: the frame left
// reservoirs, {ray, u} records and a G-buffer for reprojected temporal reuse.
static void remember_camera(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame, const FrameArgs &a,
                            bool records) {
    view->hist_cam = true;
    view->hist_records = records;
    view->hist_scene = scene->id;
    std::memcpy(view->hist_view, frame->view, sizeof view->hist_view);
    std::memcpy(view->hist_proj, frame->proj, sizeof view->hist_proj);
    const V3D o = xform_point_d(a.inv_view, V3D{0.0, 0.0, 0.0});  // camera_ray's origin
    view->hist_origin[0] = float(o.x);
    view->hist_origin[1] = float(o.y);
    view->hist_origin[2] = float(o.z);
}

// One frame, or `frames` consecutive frames of a static view in one launch (multi_frame_ok:
// megakernel or lean path tracer frames that keep no state from frame to frame).
static hdr_status render_frames_impl(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                                     const hdr_pipeline_desc *pipe, float *color_out, hdr_pixel_debug *debug_out,
                                     int frames) {
    FrameArgs a;
    FramePlan plan;
    if (hdr_status st = prepare_frame(view, scene, frame, pipe, debug_out != nullptr, a, plan, false, frames > 1))
        return st;
    const int W = view->W, H = view->H;
    const size_t P = size_t(W) * H;
    if (frames > 1) {  // the frames' samples are staged, then accumulated in frame order
        if (hdr_status st = ensure(view->frame, size_t(frames) * P * sizeof(float4))) return st;
        a.frame_rgb = static_cast<float4 *>(view->frame.p);
        a.frames = frames;
    }
    const bool deferred = plan.deferred, spatial = plan.spatial;
    const size_t lds = plan.lds;
    const int gp = plan.gbuf_half;  // this frame's G-buffer half (prepare_frame chose it)
    // LDS stacks for the any-hit lanes only when no G-buffer launch runs beside them on the second
    // stream: that launch's waves need the LDS too (dolly headline frame 3.04 -> 3.12 ms with them)
    const bool ldsFit = lds_stack_fits(scene) && plan.gbuf_reuse;
    const dim3 grid(plan.tiles);  // one block per 8x8 tile
    if (view->counting) {
        HIP_TRY(hipMemsetAsync(view->counters.p, 0, view->counters.n, view->stream));
        a.counters = static_cast<unsigned long long *>(view->counters.p);
    }
    hdr_view::StageSlot &ts = view->ring[view->ring_next];
    if (hdr_status st = harvest(view, ts)) return st;  // the slot's previous frame has finished
    ts.ran = 0;
    // this frame's arguments: pinned staging -> the slot's device record, ahead of the G-buffer
    view->hargs.p[view->ring_next] = a;
    const FrameArgs *da = static_cast<const FrameArgs *>(view->dargs.p) + view->ring_next;
    const int *dt = static_cast<const int *>(view->dtile.p);
    // G-buffer on the second stream: after the previous frame's shading has read the G-buffer, or,
    // when counting, after everything queued so far (counters are per frame). A reused G-buffer
    // needs no launch (only the argument copy runs there).
    if (view->counting) {
        HIP_TRY(hipEventRecord(view->serial, view->stream));
        HIP_TRY(hipStreamWaitEvent(view->gstream, view->serial, 0));
    } else if (!plan.gbuf_reuse && view->shade_recorded[gp]) {
        HIP_TRY(hipStreamWaitEvent(view->gstream, view->shade_done[gp], 0));
    }
    HIP_TRY(hipMemcpyAsync(const_cast<FrameArgs *>(da), &view->hargs.p[view->ring_next], sizeof(FrameArgs),
                           hipMemcpyHostToDevice, view->gstream));
    if (!plan.gbuf_reuse) {
        HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_GBUFFER], view->gstream));
        if (view->counting) launch_gbuffer<true>(view->shape, plan.tiles, view->gstream, da, dt, 1);
        else launch_gbuffer<false>(view->shape, plan.tiles, view->gstream, da, dt, 1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_GBUFFER], view->gstream));
        ts.ran |= 1u << HDR_STAGE_GBUFFER;
    }
    HIP_TRY(hipEventRecord(view->gbuf_done, view->gstream));
    HIP_TRY(hipStreamWaitEvent(view->stream, view->gbuf_done, 0));
    HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_INTEGRATE], view->stream));
    if (a.reproject) {  // the previous frame's records, gathered before this frame overwrites them
        hipLaunchKernelGGL(reproject_kernel, grid, dim3(kLanes), 0, view->stream, da, dt, 1);
        HIP_TRY(hipGetLastError());
    }
    if (deferred) {
        HIP_TRY(hipMemsetAsync(a.qcount, 0, 2 * sizeof(unsigned int), view->stream));
#define HDR_SHADE(C, D) \
    hipLaunchKernelGGL((shade_direct_kernel<C, D>), grid, dim3(kLanes), lds, view->stream, da, dt, 1)
        if (view->counting) {
            if (debug_out) HDR_SHADE(true, true);
            else HDR_SHADE(true, false);
        } else {
            if (debug_out) HDR_SHADE(false, true);
            else HDR_SHADE(false, false);
        }
#undef HDR_SHADE
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_INTEGRATE], view->stream));
        if (spatial) {
            HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_SPATIAL], view->stream));
#define HDR_SPATIAL(C, D) launch_spatial<C, D>(plan.spatial_nd, grid, view->stream, da, dt, 1)
            if (view->counting) {
                if (debug_out) HDR_SPATIAL(true, true);
                else HDR_SPATIAL(true, false);
            } else {
                if (debug_out) HDR_SPATIAL(false, true);
                else HDR_SPATIAL(false, false);
            }
#undef HDR_SPATIAL
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_SPATIAL], view->stream));
            ts.ran |= 1u << HDR_STAGE_SPATIAL;
        }
        HIP_TRY(hipEventRecord(view->shade_done[gp], view->stream));
        view->shade_recorded[gp] = true;
        HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_SHADOW], view->stream));
        if (view->counting) launch_shadow<true>(view->shape, view->ctx->cus, P, view->stream, da, true, ldsFit);
        else launch_shadow<false>(view->shape, view->ctx->cus, P, view->stream, da, true, ldsFit);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_SHADOW], view->stream));
        HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_ACCUMULATE], view->stream));
        if (debug_out)
            hipLaunchKernelGGL((accumulate_deferred_kernel<true>), grid, dim3(kLanes), 0, view->stream, da, dt, 1);
        else
            hipLaunchKernelGGL((accumulate_deferred_kernel<false>), grid, dim3(kLanes), 0, view->stream, da, dt, 1);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_ACCUMULATE], view->stream));
        ts.ran |= (1u << HDR_STAGE_INTEGRATE) | (1u << HDR_STAGE_SHADOW) | (1u << HDR_STAGE_ACCUMULATE);
    } else if (plan.wave) {
        const int qblocks = queue_blocks(view->shape, P);
        hipLaunchKernelGGL(wave_zero_counts_kernel, dim3(1), dim3(kLanes), 0, view->stream, a.wcnt);
#define HDR_W0(C, D) \
    launch_wave_shade0<C, D>(plan.wave_lean0, plan.wave_traced, plan.tiles, lds, view->stream, da, dt, 1)
#define HDR_WS(C, D) \
    launch_wave_shade<C, D>(view->shape, qblocks, lds, view->stream, da, b, plan.wave_traced, a.gen_inline != 0)
        if (view->counting) {
            if (debug_out) HDR_W0(true, true);
            else HDR_W0(true, false);
        } else {
            if (debug_out) HDR_W0(false, true);
            else HDR_W0(false, false);
        }
        HIP_TRY(hipGetLastError());
        // spatial reuse after the launch that stores the post-temporal reservoirs: wave_shade0, or
        // on traced frames the shading launch of vertex 0 (launch 1); inside the integration
        // stage's span on this path
        auto spatial_launch = [&]() -> hdr_status {
            HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_SPATIAL], view->stream));
            if (view->counting) launch_wave_spatial<true>(plan.spatial_nd, grid, view->stream, da, dt, 1);
            else launch_wave_spatial<false>(plan.spatial_nd, grid, view->stream, da, dt, 1);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_SPATIAL], view->stream));
            ts.ran |= 1u << HDR_STAGE_SPATIAL;
            return HDR_OK;
        };
        // traced frames: shading launch v finishes vertex v-1, so there is one more launch pair
        const int launches = plan.wave_traced ? plan.depth + 1 : plan.depth;
        for (int b = 0; b < launches; ++b) {
            if (b > 0) {
                const int u = plan.wave_traced ? b - 1 : b;  // the vertex the launch shades
                if (plan.prep_window > 0 && u >= plan.prep_vertex)
                    hipLaunchKernelGGL(wave_prepare_slots_kernel, dim3(qblocks), dim3(kLanes), 0, view->stream, da, u,
                                       plan.prep_window);
                if (view->counting) {
                    if (debug_out) HDR_WS(true, true);
                    else HDR_WS(true, false);
                } else {
                    if (debug_out) HDR_WS(false, true);
                    else HDR_WS(false, false);
                }
                HIP_TRY(hipGetLastError());
            }
            if (spatial && b == (plan.wave_traced ? 1 : 0))
                if (hdr_status st = spatial_launch()) return st;
            if (b == (plan.wave_traced ? 1 : 0)) {  // the G-buffer (vertex 0's hits) has been read
                HIP_TRY(hipEventRecord(view->shade_done[gp], view->stream));
                view->shade_recorded[gp] = true;
            }
            if (view->counting)
                launch_wave_trace<true>(view->shape, qblocks, view->stream, da, b, plan.wave_traced, view->ctx->cus,
                                        true, ldsFit);
            else
                launch_wave_trace<false>(view->shape, qblocks, view->stream, da, b, plan.wave_traced, view->ctx->cus,
                                         true, ldsFit);
            HIP_TRY(hipGetLastError());
        }
#undef HDR_W0
#undef HDR_WS
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_INTEGRATE], view->stream));
        HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_ACCUMULATE], view->stream));
#define HDR_WR(C, D) \
    hipLaunchKernelGGL((wave_resolve_kernel<C, D>), grid, dim3(kLanes), 0, view->stream, da, dt, 1)
        if (view->counting) {
            if (debug_out) HDR_WR(true, true);
            else HDR_WR(true, false);
        } else {
            if (debug_out) HDR_WR(false, true);
            else HDR_WR(false, false);
        }
#undef HDR_WR
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_ACCUMULATE], view->stream));
        ts.ran |= (1u << HDR_STAGE_INTEGRATE) | (1u << HDR_STAGE_ACCUMULATE);
    } else if (plan.lean_pt) {
#define HDR_PT(C, D) launch_pt_lean<C, D>(view->ctx->cus, plan.tiles, unsigned(frames), view->stream, da, dt, 1)
        if (view->counting) {
            if (debug_out) HDR_PT(true, true);
            else HDR_PT(true, false);
        } else {
            if (debug_out) HDR_PT(false, true);
            else HDR_PT(false, false);
        }
#undef HDR_PT
    } else if (view->counting) {
        if (debug_out) hipLaunchKernelGGL((integrate_kernel<true, true>), grid, dim3(kLanes), lds, view->stream, da, dt, 1);
        else hipLaunchKernelGGL((integrate_kernel<true, false>), grid, dim3(kLanes), lds, view->stream, da, dt, 1);
    } else {
        if (debug_out) hipLaunchKernelGGL((integrate_kernel<false, true>), grid, dim3(kLanes), lds, view->stream, da, dt, 1);
        else hipLaunchKernelGGL((integrate_kernel<false, false>), dim3(plan.tiles, frames), dim3(kLanes), lds, view->stream,
                                da, dt, 1);
    }
    HIP_TRY(hipGetLastError());
    if (frames > 1) {
        hipLaunchKernelGGL(accumulate_frames_kernel, dim3(unsigned((P + 255) / 256)), dim3(256), 0, view->stream,
                           a.frame_rgb, frames, P, frame->frame_index, a.accum, a.color, a.lum_sum, a.lum_sq);
        HIP_TRY(hipGetLastError());
    }
    if (!deferred && !plan.wave) {
        ts.ran |= 1u << HDR_STAGE_INTEGRATE;
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_INTEGRATE], view->stream));
        HIP_TRY(hipEventRecord(view->shade_done[gp], view->stream));
        view->shade_recorded[gp] = true;
    }
    if (a.firefly) {
        HIP_TRY(hipEventRecord(ts.beg[HDR_STAGE_FIREFLY], view->stream));
        const dim3 b2(16, 16), g2((W + 15) / 16, (H + 15) / 16);
        hipLaunchKernelGGL(firefly_accumulate_kernel, g2, b2, 0, view->stream, a.frame_rgb, a.accum, a.color,
                           a.lum_sum, a.lum_sq, W, H,
                           a.inv_n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(ts.end[HDR_STAGE_FIREFLY], view->stream));
        ts.ran |= 1u << HDR_STAGE_FIREFLY;
    }
    HIP_TRY(hipEventRecord(ts.fin, view->stream));
    ts.pending = true;
    view->ring_last = view->ring_next;
    view->ring_next = (view->ring_next + 1) % hdr_view::kRing;
    view->frames++;
    view->samples = frame->frame_index + frames;
    if (!plan.gbuf_reuse) {
        view->gcur = gp;
        view->gvalid = view->gnext_valid;
    }
    view->single_since_batch = true;
    remember_camera(view, scene, frame, a, spatial && pipe->temporal_reprojection);
    if (color_out) {
        HIP_TRY(hipMemcpyAsync(color_out, view->color.p, P * sizeof(float4), hipMemcpyDeviceToHost, view->stream));
    }
    if (debug_out) {
        HIP_TRY(hipMemcpyAsync(debug_out, view->dbg.p, P * sizeof(hdr_pixel_debug), hipMemcpyDeviceToHost,
                               view->stream));
    }
    if (color_out || debug_out) {
        HIP_TRY(hipStreamSynchronize(view->stream));
        return view_status(view);
    }
    return HDR_OK;
}

hdr_status hdr_render_frame(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                            const hdr_pipeline_desc *pipe, float *color_out, hdr_pixel_debug *debug_out) {
    return render_frames_impl(view, scene, frame, pipe, color_out, debug_out, 1);
}

// Frames that can share a launch: the megakernel or the lean path tracer without spill slots (each
// frame's pixels then use nothing per pixel but their own sample), no state carried from frame to
// frame (MIS, or RIS without the reservoir; no spatial reuse), no firefly filter (its clamp reads
// each frame's neighbourhood), no depth of field (each frame's lens draws make its own G-buffer),
// and no counting.
static bool multi_frame_ok(const hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                           const hdr_pipeline_desc *pipe) {
    FramePlan plan;
    FrameNeeds need;
    if (plan_pipeline(scene, view->W, view->H, pipe, false, plan, need, true) != HDR_OK) return false;
    if (plan.deferred || plan.wave || plan.spill) return false;
    if (pipe->enable_firefly_filter || pipe->spatial_count > 0) return false;
    if (pipe->direct_light == HDR_DIRECT_RIS && pipe->use_reservoir) return false;
    if (frame->has_camera_params && frame->camera.enable_dof) return false;
    return !view->counting;
}

hdr_status hdr_render_frames(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                             const hdr_pipeline_desc *pipe, int32_t count) {
    if (!view || !scene || !frame || !pipe) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (count < 0) return fail(HDR_ERR_INVALID_ARGUMENT, "count must be >= 0");
    if (frame->struct_size != sizeof(hdr_frame_desc)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_frame_desc.struct_size mismatch");
    if (pipe->struct_size != sizeof(hdr_pipeline_desc)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_pipeline_desc size");
    const size_t P = size_t(view->W) * view->H;
    // frames per launch: the staged samples within 256 MiB, at most 64
    const int maxF = int(std::max<size_t>(1, std::min<size_t>(64, (size_t(256) << 20) / (P * sizeof(float4)))));
    const bool multi = count > 1 && maxF > 1 && multi_frame_ok(view, scene, frame, pipe);
    hdr_frame_desc f = *frame;
    for (int done = 0; done < count;) {
        const int n = multi ? std::min(count - done, maxF) : 1;
        f.frame_index = frame->frame_index + done;
        if (hdr_status st = render_frames_impl(view, scene, &f, pipe, nullptr, nullptr, n)) return st;
        done += n;
    }
    return HDR_OK;
}

hdr_status hdr_view_synchronize(hdr_view *view) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipStreamSynchronize(view->stream));
    HIP_TRY(hipStreamSynchronize(view->gstream));
    return view_status(view);
}

hdr_status hdr_view_last_timing(hdr_view *view, float *gbuffer_ms, float *integrate_ms, float *frame_ms) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    float g = 0.0f, i = 0.0f, f = 0.0f;
    if (view->ring_last >= 0) {
        HIP_TRY(hipSetDevice(view->ctx->device));
        hdr_view::StageSlot &s = view->ring[view->ring_last];
        HIP_TRY(hipEventSynchronize(s.fin));
        const int last = (s.ran & (1u << HDR_STAGE_ACCUMULATE)) ? HDR_STAGE_ACCUMULATE : HDR_STAGE_INTEGRATE;
        const bool gb = (s.ran & (1u << HDR_STAGE_GBUFFER)) != 0;  // not when the G-buffer was reused
        if (gb) HIP_TRY(hipEventElapsedTime(&g, s.beg[HDR_STAGE_GBUFFER], s.end[HDR_STAGE_GBUFFER]));
        HIP_TRY(hipEventElapsedTime(&i, s.beg[HDR_STAGE_INTEGRATE], s.end[last]));
        HIP_TRY(hipEventElapsedTime(&f, s.beg[gb ? HDR_STAGE_GBUFFER : HDR_STAGE_INTEGRATE], s.fin));
    }
    if (gbuffer_ms) *gbuffer_ms = g;
    if (integrate_ms) *integrate_ms = i;
    if (frame_ms) *frame_ms = f;
    return HDR_OK;
}

hdr_status hdr_view_read_stage_times(hdr_view *view, double *stage_ms, int32_t *launches) {
    if (!view || !stage_ms || !launches) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(view->ctx->device));
    for (int k = 1; k <= hdr_view::kRing; ++k) {  // oldest first
        if (hdr_status st = harvest(view, view->ring[(view->ring_next + k - 1) % hdr_view::kRing])) return st;
    }
    for (int k = 0; k < HDR_STAGE_COUNT; ++k) {
        stage_ms[k] = view->stage_ms[k];
        launches[k] = view->stage_n[k];
        view->stage_ms[k] = 0.0;
        view->stage_n[k] = 0;
    }
    return HDR_OK;
}

hdr_status hdr_view_read_color(hdr_view *view, float *rgba_out) {
    if (!view || !rgba_out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipMemcpyAsync(rgba_out, view->color.p, view->color.n, hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return view_status(view);
}

hdr_status hdr_view_read_reservoirs(hdr_view *view, hdr_reservoir *out) {
    if (!view || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipMemcpyAsync(out, view->res.p, view->res.n, hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return view_status(view);
}

// First request of the AOVs: allocate them (zeroed); the G-buffer kernel writes them from the next
// frame on.
static hdr_status enable_aovs(hdr_view *view) {
    if (view->depth.p) return HDR_OK;
    const size_t P = size_t(view->W) * view->H;
    view->gvalid = false;  // the next frame runs the G-buffer kernel, which writes the AOVs
    if (hdr_status s = ensure(view->depth, P * sizeof(float))) return s;
    if (hdr_status s = ensure(view->albedo, P * 3 * sizeof(float))) return s;
    if (hdr_status s = ensure(view->normal, P * 3 * sizeof(float))) return s;
    HIP_TRY(hipMemsetAsync(view->depth.p, 0, view->depth.n, view->stream));
    HIP_TRY(hipMemsetAsync(view->albedo.p, 0, view->albedo.n, view->stream));
    HIP_TRY(hipMemsetAsync(view->normal.p, 0, view->normal.n, view->stream));
    return HDR_OK;
}

hdr_status hdr_view_read_aovs(hdr_view *view, float *depth, float *albedo, float *normal) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    HIP_TRY(hipSetDevice(view->ctx->device));
    const size_t P = size_t(view->W) * view->H;
    if (hdr_status s = enable_aovs(view)) return s;
    if (depth) HIP_TRY(hipMemcpyAsync(depth, view->depth.p, P * sizeof(float), hipMemcpyDeviceToHost, view->stream));
    if (albedo)
        HIP_TRY(hipMemcpyAsync(albedo, view->albedo.p, P * 3 * sizeof(float), hipMemcpyDeviceToHost, view->stream));
    if (normal)
        HIP_TRY(hipMemcpyAsync(normal, view->normal.p, P * 3 * sizeof(float), hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return HDR_OK;
}

hdr_status hdr_rendered_size(int32_t window_width, int32_t window_height, int32_t level, int32_t *width,
                             int32_t *height) {
    if (!width || !height) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (window_width <= 0 || window_height <= 0) return fail(HDR_ERR_INVALID_ARGUMENT, "window size must be positive");
    const int div = 1 << std::clamp(level, 0, 4);
    *width = std::max(1, (window_width + div - 1) / div);
    *height = std::max(1, (window_height + div - 1) / div);
    return HDR_OK;
}

hdr_status hdr_view_sample_count(const hdr_view *view, int32_t *samples) {
    if (!view || !samples) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    *samples = view->samples;
    return HDR_OK;
}

hdr_status hdr_view_resolve(hdr_view *view, int32_t aov, int32_t window_width, int32_t window_height,
                            int32_t format, int32_t components, void *dst) {
    if (!view || !dst) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (aov < HDR_AOV_COLOR || aov > HDR_AOV_NORMAL) return fail(HDR_ERR_INVALID_ARGUMENT, "unknown aov");
    if (format < HDR_FORMAT_FLOAT32 || format > HDR_FORMAT_INT32) return fail(HDR_ERR_INVALID_ARGUMENT, "unknown format");
    if (components < 1 || components > 4) return fail(HDR_ERR_INVALID_ARGUMENT, "components must be 1 .. 4");
    if (window_width <= 0 || window_height <= 0 || int64_t(window_width) * window_height >= (int64_t(1) << 28))
        return fail(HDR_ERR_INVALID_ARGUMENT, "window size out of range");
    HIP_TRY(hipSetDevice(view->ctx->device));
    if (aov != HDR_AOV_COLOR)
        if (hdr_status s = enable_aovs(view)) return s;
    const float *src = nullptr;
    int sc = 0;
    switch (aov) {
    case HDR_AOV_COLOR: src = static_cast<const float *>(view->color.p), sc = 4; break;
    case HDR_AOV_DEPTH: src = static_cast<const float *>(view->depth.p), sc = 1; break;
    case HDR_AOV_ALBEDO: src = static_cast<const float *>(view->albedo.p), sc = 3; break;
    default: src = static_cast<const float *>(view->normal.p), sc = 3; break;
    }
    const size_t esize = (format == HDR_FORMAT_FLOAT32 || format == HDR_FORMAT_INT32) ? 4
                         : format == HDR_FORMAT_FLOAT16                               ? 2
                                                                                      : 1;
    const size_t bytes = size_t(window_width) * window_height * size_t(components) * esize;
    if (hdr_status s = ensure(view->resolved, bytes)) return s;
    const dim3 b(16, 16), g((window_width + 15) / 16, (window_height + 15) / 16);
    hipLaunchKernelGGL(resolve_kernel, g, b, 0, view->stream, src, sc, view->W, view->H, window_width, window_height,
                       format, components, static_cast<unsigned char *>(view->resolved.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dst, view->resolved.p, bytes, hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return view_status(view);
}

hdr_status hdr_view_post_process(hdr_view *view, const hdr_post_desc *desc) {
    if (!view || !desc) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (desc->struct_size != sizeof(hdr_post_desc)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_post_desc size");
    if (desc->oidn) return fail(HDR_ERR_UNSUPPORTED, "the OIDN denoiser is not part of this library");
    const float ca = desc->chromatic_aberration;
    // shifts beyond the int range have no defined conversion in the reference (post_process.cpp:128-131)
    if (!(ca <= 1.0e6f)) return fail(HDR_ERR_INVALID_ARGUMENT, "chromatic_aberration must be a number <= 1e6");
    if (!desc->chromaticity_blur && !desc->lens_flare && !(ca > 0.0f)) return HDR_OK;  // both passes return early
    HIP_TRY(hipSetDevice(view->ctx->device));
    const int W = view->W, H = view->H;
    const size_t P = size_t(W) * H;
    if (hdr_status s = ensure(view->post, 3 * P * sizeof(float4) + size_t(H) * sizeof(int))) return s;
    float4 *color = static_cast<float4 *>(view->color.p);
    float4 *plane[2] = {static_cast<float4 *>(view->post.p), static_cast<float4 *>(view->post.p) + P};
    float4 *bloom = plane[1] + P;
    int *rowAny = reinterpret_cast<int *>(bloom + P);
    const dim3 b(16, 16), g((W + 15) / 16, (H + 15) / 16);
    // each pass reads `cur` and writes the other plane; the last result is copied back into Color
    const float4 *cur = color;
    int next = 0;
    if (desc->chromaticity_blur) {
        hipLaunchKernelGGL(chroma_blur_kernel, g, b, 0, view->stream, cur, plane[next], W, H);
        cur = plane[next];
        next ^= 1;
    }
    if (desc->lens_flare) {
        const int blur = std::max(1, std::min(W, H) / 20);  // post_process.cpp:43
        hipLaunchKernelGGL(flare_bright_kernel, dim3(H), dim3(256), 0, view->stream, cur, bloom, rowAny, W);
        hipLaunchKernelGGL(flare_gather_kernel, g, b, 0, view->stream, cur, bloom, rowAny, plane[next], W, H, blur);
        cur = plane[next];
        next ^= 1;
    }
    if (ca > 0.0f) {
        hipLaunchKernelGGL(chromatic_kernel, g, b, 0, view->stream, cur, plane[next], W, H, ca);
        cur = plane[next];
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(color, cur, P * sizeof(float4), hipMemcpyDeviceToDevice, view->stream));
    return HDR_OK;
}

hdr_status hdr_render_progressive(hdr_view *view, const hdr_scene *scene, const hdr_frame_desc *frame,
                                  const hdr_pipeline_desc *pipe, int32_t target_samples, int32_t max_frames,
                                  int32_t *converged) {
    if (!view || !scene || !frame || !pipe || !converged) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (max_frames < 0) return fail(HDR_ERR_INVALID_ARGUMENT, "max_frames must be >= 0");
    if (frame->struct_size != sizeof(hdr_frame_desc)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_frame_desc.struct_size mismatch");
    if (view->ctx->cancelled && max_frames > 0 && view->samples < target_samples) {
        *converged = 0;
        return fail(HDR_ERR_CANCELLED, "cancelled");
    }
    // the call's frames go through hdr_render_frames, which shares launches among frames that keep
    // no state from one to the next (the cancel flag is read before each launch)
    hdr_frame_desc f = *frame;
    f.frame_index = view->samples;
    const int n = int(std::min<int64_t>(max_frames, std::max<int64_t>(0, int64_t(target_samples) - view->samples)));
    const hdr_status s = hdr_render_frames(view, scene, &f, pipe, n);
    *converged = view->samples >= target_samples ? 1 : 0;
    return s;
}

hdr_status hdr_view_set_metrics(hdr_view *view, int32_t enable) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipStreamSynchronize(view->stream));
    HIP_TRY(hipStreamSynchronize(view->gstream));
    if (!enable) {
        view->lum_sum.release();
        view->lum_sq.release();
        view->variance.release();
        return HDR_OK;
    }
    if (view->lum_sum.p) return HDR_OK;
    const size_t P = size_t(view->W) * view->H;
    if (hdr_status s = ensure(view->lum_sum, P * sizeof(double))) return s;
    if (hdr_status s = ensure(view->lum_sq, P * sizeof(double))) return s;
    HIP_TRY(hipMemsetAsync(view->lum_sum.p, 0, view->lum_sum.n, view->stream));
    HIP_TRY(hipMemsetAsync(view->lum_sq.p, 0, view->lum_sq.n, view->stream));
    return HDR_OK;
}

hdr_status hdr_view_read_variance(hdr_view *view, hdr_variance_stats *out, double *lum_sum, double *lum_sum_sq) {
    if (!view || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (out->struct_size != sizeof(hdr_variance_stats))
        return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_variance_stats.struct_size mismatch");
    if (!view->lum_sum.p) return fail(HDR_ERR_INVALID_ARGUMENT, "the view's metrics are off (hdr_view_set_metrics)");
    HIP_TRY(hipSetDevice(view->ctx->device));
    const int P = view->W * view->H;
    out->sample_count = view->samples;
    out->valid = view->samples >= 2 ? 1 : 0;
    out->mean_estimator_variance = out->max_estimator_variance = out->mean_luminance = 0.0;
    if (out->valid) {
        const int blocks = (P + 255) / 256;
        if (hdr_status s = ensure(view->variance, (size_t(blocks) * 3 + 3) * sizeof(double))) return s;
        double *part = static_cast<double *>(view->variance.p);
        hipLaunchKernelGGL(variance_partial_kernel, dim3(blocks), dim3(256), 0, view->stream,
                           static_cast<const double *>(view->lum_sum.p), static_cast<const double *>(view->lum_sq.p),
                           P, view->samples, part + 3);
        hipLaunchKernelGGL(variance_final_kernel, dim3(1), dim3(64), 0, view->stream, part + 3, blocks, P, part);
        HIP_TRY(hipGetLastError());
        double r[3];
        HIP_TRY(hipMemcpyAsync(r, part, sizeof r, hipMemcpyDeviceToHost, view->stream));
        HIP_TRY(hipStreamSynchronize(view->stream));
        out->mean_estimator_variance = r[0];
        out->max_estimator_variance = r[1];
        out->mean_luminance = r[2];
    }
    if (lum_sum)
        HIP_TRY(hipMemcpyAsync(lum_sum, view->lum_sum.p, view->lum_sum.n, hipMemcpyDeviceToHost, view->stream));
    if (lum_sum_sq)
        HIP_TRY(hipMemcpyAsync(lum_sum_sq, view->lum_sq.p, view->lum_sq.n, hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return view_status(view);
}

hdr_status hdr_view_color_device_ptr(hdr_view *view, void **out) {
    if (!view || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    *out = view->color.p;
    return HDR_OK;
}

hdr_status hdr_view_read_counters(hdr_view *view, hdr_frame_counters *out) {
    if (!view || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(view->ctx->device));
    HIP_TRY(hipMemcpyAsync(out, view->counters.p, sizeof(hdr_frame_counters), hipMemcpyDeviceToHost, view->stream));
    HIP_TRY(hipStreamSynchronize(view->stream));
    return HDR_OK;
}

hdr_status hdr_scene_intersect(hdr_scene *scene, const float *origins, const float *dirs, int32_t n,
                               int32_t *instance_out, int32_t *triangle_out, float *t_out) {
    if (!scene || n < 0 || (n && (!origins || !dirs || !instance_out || !triangle_out || !t_out)))
        return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return HDR_OK;
    HIP_TRY(hipSetDevice(scene->ctx->device));
    DevBuf dO, dD, dI, dT, dt;
    HIP_TRY(dO.alloc(size_t(n) * 12));
    HIP_TRY(dD.alloc(size_t(n) * 12));
    HIP_TRY(dI.alloc(size_t(n) * 4));
    HIP_TRY(dT.alloc(size_t(n) * 4));
    HIP_TRY(dt.alloc(size_t(n) * 4));
    HIP_TRY(hipMemcpy(dO.p, origins, size_t(n) * 12, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dD.p, dirs, size_t(n) * 12, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(intersect_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, scene->ds,
                       static_cast<const float *>(dO.p), static_cast<const float *>(dD.p), n,
                       static_cast<int *>(dI.p), static_cast<int *>(dT.p), static_cast<float *>(dt.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(instance_out, dI.p, size_t(n) * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(triangle_out, dT.p, size_t(n) * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(t_out, dt.p, size_t(n) * 4, hipMemcpyDeviceToHost));
    return HDR_OK;
}

static hdr_status rng_common(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t nd, uint64_t *raw,
                             float *flt) {
    if (!ctx || !seeds || n < 0 || nd < 0) return fail(HDR_ERR_INVALID_ARGUMENT, "invalid argument");
    if (n == 0 || nd == 0) return HDR_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf dS, dOut, dSpill;
    HIP_TRY(dS.alloc(size_t(n) * 4));
    HIP_TRY(dOut.alloc(size_t(n) * nd * (raw ? 8 : 4)));
    if (nd > 311) HIP_TRY(dSpill.alloc(size_t(n) * kMtN * 8));
    HIP_TRY(hipMemcpy(dS.p, seeds, size_t(n) * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(rng_kernel, dim3((n + 63) / 64), dim3(64), 0, 0, static_cast<const uint32_t *>(dS.p), n, nd,
                       raw ? static_cast<uint64_t *>(dOut.p) : nullptr, raw ? nullptr : static_cast<float *>(dOut.p),
                       static_cast<uint64_t *>(dSpill.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(raw ? static_cast<void *>(raw) : static_cast<void *>(flt), dOut.p, dOut.n,
                      hipMemcpyDeviceToHost));
    return HDR_OK;
}

hdr_status hdr_rng_raw(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out) {
    if (!out) return fail(HDR_ERR_INVALID_ARGUMENT, "out is null");
    return rng_common(ctx, seeds, n, n_draws, out, nullptr);
}

hdr_status hdr_rng_float(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, float *out) {
    if (!out) return fail(HDR_ERR_INVALID_ARGUMENT, "out is null");
    return rng_common(ctx, seeds, n, n_draws, nullptr, out);
}

// hdr_rng_raw on device buffers, with the spill slots in the static scratch: no allocation, no
// copy, no wait (the launch is queued on `stream`).
hdr_status hdr_rng_raw_noalloc(hdr_context *ctx, const uint32_t *seeds, int32_t n, int32_t n_draws, uint64_t *out,
                               void *stream) {
    if (!ctx || n < 0 || n_draws < 0) return fail(HDR_ERR_INVALID_ARGUMENT, "invalid argument");
    if (n == 0 || n_draws == 0) return HDR_OK;
    if (!seeds || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null device pointer");
    uint64_t *spill = ctx->device < kMaxDevices ? g_scratch[ctx->device].load() : nullptr;
    if (!spill) return fail(HDR_ERR_UNSUPPORTED, "no scratch was set up for this device");
    const bool needSpill = n_draws > 311;  // outputs from 312 on come from a materialised state
    if (needSpill && size_t(n) * kMtN * sizeof(uint64_t) > kRngScratchBytes)
        return fail(HDR_ERR_CAPACITY, "more than " + std::to_string(kRngScratchBytes / (kMtN * sizeof(uint64_t))) +
                                          " seeds with n_draws > 311 exceed the 32 MiB scratch");
    HIP_TRY(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(rng_kernel, dim3((n + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream), seeds, n,
                       n_draws, out, nullptr, needSpill ? spill : nullptr);
    HIP_TRY(hipGetLastError());
    return HDR_OK;
}


// ---------------------------------------------------------------------------- batches of views
struct hdr_batch {
    hdr_context *ctx = nullptr;
    hipStream_t stream = nullptr;
    // A step's argument copy and G-buffer launch run on `gstream`, beside the previous step's
    // shading on `stream` (the views' G-buffers are double-buffered by frame parity); the step's
    // shading waits for them (event `gdone`).
    hipStream_t gstream = nullptr;
    hipEvent_t gdone = nullptr;
    hipEvent_t beg = nullptr, end = nullptr;
    bool timed = false;
    DevBuf queue, qcount;  // shadow queries of every view of a step, and their number
    DevBuf wray, wqueue, wcq, walive, wcnt;  // wavefront-path queues of a step (shared by its views)
    DevBuf wesq;                             // and the traced frames' escape-ray queue
    // steps in flight: arguments and tile tables staged in pinned memory, copied to the slot's
    // device buffer; a slot is reused once its step has finished
    static constexpr int kSlots = 4;
    struct Slot {
        DevBuf dev;
        PinnedBuf<char> host;
        hipEvent_t done = nullptr;
        bool pending = false;
        uint64_t step = 0;  // number of the step the slot last carried (1, 2, ...)
    };
    Slot slots[kSlots];
    int next = 0;
    uint64_t issued = 0;  // steps queued so far
    LaunchShape shape;    // queue-driven launches (hdr_batch_set_launch_config / hdr_batch_autotune)
    bool graphs = true;   // steps replay captured graphs
    struct Graph {
        std::vector<uint64_t> key;  // launch shape the graph was captured for
        hipGraphExec_t exec = nullptr;
    };
    Graph graph[kSlots];  // one per argument slot (its buffers are baked into the captured copy)
};

hdr_status hdr_batch_create(hdr_context *ctx, hdr_batch **out) {
    if (!ctx || !out) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<hdr_batch> b(new (std::nothrow) hdr_batch());
    if (!b) return fail(HDR_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->ctx = ctx;
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&b->gdone, hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&b->beg));
    HIP_TRY(hipEventCreate(&b->end));
    for (auto &sl : b->slots) HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    HIP_TRY(b->qcount.alloc(2 * sizeof(unsigned int)));  // query count, persistent claim counter
    HIP_TRY(b->wcnt.alloc(kWaveCounts * sizeof(unsigned int)));
    *out = b.release();
    return HDR_OK;
}

hdr_status hdr_batch_destroy(hdr_batch *b) {
    if (!b) return HDR_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->gstream) (void)hipStreamSynchronize(b->gstream);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->graph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto &sl : b->slots)
        if (sl.done) (void)hipEventDestroy(sl.done);
    if (b->beg) (void)hipEventDestroy(b->beg);
    if (b->end) (void)hipEventDestroy(b->end);
    if (b->gdone) (void)hipEventDestroy(b->gdone);
    if (b->gstream) (void)hipStreamDestroy(b->gstream);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return HDR_OK;
}

hdr_status hdr_batch_synchronize(hdr_batch *b) {
    if (!b) return fail(HDR_ERR_INVALID_ARGUMENT, "batch is null");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->gstream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return HDR_OK;
}

hdr_status hdr_batch_poll(hdr_batch *b, uint64_t *issued, uint64_t *completed) {
    if (!b || !issued || !completed) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipSetDevice(b->ctx->device));
    uint64_t done = b->issued;  // steps run in order on the batch's stream
    for (const hdr_batch::Slot &sl : b->slots) {
        if (!sl.pending) continue;
        const hipError_t e = hipEventQuery(sl.done);
        if (e == hipSuccess) continue;
        if (e != hipErrorNotReady) HIP_TRY(e);
        done = std::min(done, sl.step - 1);
    }
    *issued = b->issued;
    *completed = done;
    return HDR_OK;
}

hdr_status hdr_batch_last_timing(hdr_batch *b, float *step_ms) {
    if (!b || !step_ms) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    *step_ms = 0.0f;
    if (!b->timed) return HDR_OK;
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipEventSynchronize(b->end));
    HIP_TRY(hipEventElapsedTime(step_ms, b->beg, b->end));
    return HDR_OK;
}

static hdr_status shape_of(const hdr_launch_config *c, LaunchShape &ls) {
    if (!c || c->struct_size != sizeof(hdr_launch_config)) return fail(HDR_ERR_INVALID_ARGUMENT, "hdr_launch_config size");
    if (c->queue_waves != 0 && (c->queue_waves < 64 || c->queue_waves > 65536))
        return fail(HDR_ERR_INVALID_ARGUMENT, "queue_waves must be 0 or in [64, 65536]");
    const int t = c->trace_occupancy == 0 ? 8 : c->trace_occupancy;
    if (t != 1 && t != 4 && t != 6 && t != 8)
        return fail(HDR_ERR_INVALID_ARGUMENT, "trace_occupancy must be 0, 1, 4, 6 or 8");
    const int sh = c->shade_occupancy == 0 ? 3 : c->shade_occupancy;
    if (sh != 1 && sh != 3) return fail(HDR_ERR_INVALID_ARGUMENT, "shade_occupancy must be 0, 1 or 3");
    if (c->lds_stack != 0 && c->lds_stack != 1) return fail(HDR_ERR_INVALID_ARGUMENT, "lds_stack must be 0 or 1");
    ls.queue_waves = c->queue_waves;
    ls.trace_occ = t;
    ls.shade_occ = sh;
    ls.lds_stack = c->lds_stack != 0;
    return HDR_OK;
}
static void config_of(const LaunchShape &ls, bool graphs, hdr_launch_config *c) {
    std::memset(c, 0, sizeof *c);
    c->struct_size = sizeof(hdr_launch_config);
    c->queue_waves = ls.queue_waves;
    c->trace_occupancy = ls.trace_occ;
    c->shade_occupancy = ls.shade_occ;
    c->use_graphs = graphs ? 1 : 0;
    c->lds_stack = ls.lds_stack ? 1 : 0;
}

hdr_status hdr_view_set_launch_config(hdr_view *view, const hdr_launch_config *cfg) {
    if (!view) return fail(HDR_ERR_INVALID_ARGUMENT, "view is null");
    LaunchShape ls;
    if (hdr_status st = shape_of(cfg, ls)) return st;
    view->shape = ls;
    return HDR_OK;
}

hdr_status hdr_batch_set_launch_config(hdr_batch *b, const hdr_launch_config *cfg) {
    if (!b) return fail(HDR_ERR_INVALID_ARGUMENT, "batch is null");
    LaunchShape ls;
    if (hdr_status st = shape_of(cfg, ls)) return st;
    b->shape = ls;
    b->graphs = cfg->use_graphs != 0;
    return HDR_OK;
}

hdr_status hdr_batch_get_launch_config(hdr_batch *b, hdr_launch_config *cfg) {
    if (!b || !cfg) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    config_of(b->shape, b->graphs, cfg);
    return HDR_OK;
}

hdr_status hdr_render_batch(hdr_batch *b, int32_t n, hdr_view *const *views, const hdr_scene *const *scenes,
                            const hdr_frame_desc *frames, const hdr_pipeline_desc *pipes) {
    if (!b || !views || !scenes || !frames || !pipes) return fail(HDR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 1 || n > 255) return fail(HDR_ERR_INVALID_ARGUMENT, "a batch holds 1 to 255 views");
    for (int i = 0; i < n; ++i) {
        if (!views[i] || views[i]->ctx != b->ctx) return fail(HDR_ERR_INVALID_ARGUMENT, "view of another context");
        if (views[i]->counting) return fail(HDR_ERR_UNSUPPORTED, "counting views render one at a time");
        if (pipes[i].struct_size == sizeof(hdr_pipeline_desc) && pipes[i].enable_firefly_filter)
            return fail(HDR_ERR_UNSUPPORTED, "the firefly filter renders one view at a time");
        if (pipes[i].struct_size == sizeof(hdr_pipeline_desc) && pipes[i].temporal_reprojection)
            return fail(HDR_ERR_UNSUPPORTED, "temporal_reprojection renders one view at a time");
        for (int j = 0; j < i; ++j)
            if (views[j] == views[i]) return fail(HDR_ERR_INVALID_ARGUMENT, "a view appears twice in a batch");
    }
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::vector<FrameArgs> args(static_cast<size_t>(n));
    std::vector<FramePlan> plans(static_cast<size_t>(n));
    size_t queueCap = 0, waveCap = 0, escCap = 0;
    // wavefront views in two groups (bounce rays traced by the traversal launch / by the shading
    // launches), each run as its own launch sequence over the shared queues
    int waveDepth[2] = {0, 0};
    bool waveLean0[2] = {true, true};  // every view's primary vertex stays below RNG output 311
    bool waveSpatial[2] = {false, false};  // some view of the group reuses spatially
    bool waveGen[2] = {false, false};      // some view of the group traces generated candidates (mesh lights)
    int wavePrepWindow[2] = {0, 0};        // slot preparation (FramePlan::prep_window): the largest window
    int wavePrepVertex[2] = {INT_MAX, INT_MAX};  // and the first vertex that needs it, over the group
    bool oneScene = true;                  // every view renders the first view's scene
    bool ldsFit = true;  // every scene's any-hit stacks fit in LDS and no view launches a G-buffer
    for (int i = 0; i < n; ++i) {
        oneScene = oneScene && scenes[i] == scenes[0];
        if (hdr_status st =
                prepare_frame(views[i], scenes[i], &frames[i], &pipes[i], false, args[i], plans[i], true))
            return st;
        ldsFit = ldsFit && lds_stack_fits(scenes[i]) && plans[i].gbuf_reuse;  // as in hdr_render_frame
        if (plans[i].deferred) queueCap += size_t(views[i]->W) * views[i]->H;
        if (plans[i].wave) {
            const int t = plans[i].wave_traced ? 1 : 0;
            waveCap += size_t(views[i]->W) * views[i]->H;
            escCap += plans[i].escapes;
            waveDepth[t] = std::max(waveDepth[t], plans[i].depth);
            waveLean0[t] = waveLean0[t] && plans[i].wave_lean0;
            if (plans[i].prep_window > 0) {
                wavePrepWindow[t] = std::max(wavePrepWindow[t], plans[i].prep_window);
                wavePrepVertex[t] = std::min(wavePrepVertex[t], plans[i].prep_vertex);
            }
            waveSpatial[t] = waveSpatial[t] || plans[i].spatial;
            waveGen[t] = waveGen[t] || args[i].gen_inline != 0;
        }
    }
    if (queueCap * sizeof(ShadowQuery) > b->queue.n || waveCap * sizeof(ShadowQuery) > b->wqueue.n ||
        (escCap > 0 && escape_queue_bytes(escCap) > b->wesq.n)) {
        HIP_TRY(hipStreamSynchronize(b->stream));  // earlier steps may still read the old queues
        if (hdr_status st = ensure(b->queue, queueCap * sizeof(ShadowQuery))) return st;
        if (waveCap > 0) {
            if (hdr_status st = ensure(b->wqueue, waveCap * sizeof(ShadowQuery))) return st;
            if (hdr_status st = ensure(b->wray, waveCap * 2 * sizeof(float4))) return st;
            if (hdr_status st = ensure(b->wcq, waveCap * sizeof(float4))) return st;
            if (hdr_status st = ensure(b->walive, 2 * waveCap * sizeof(unsigned int))) return st;
        }
        if (escCap > 0)
            if (hdr_status st = ensure(b->wesq, escape_queue_bytes(escCap))) return st;
    }
    for (int i = 0; i < n; ++i) {
        args[i].view = i;
        args[i].qcount = static_cast<unsigned int *>(b->qcount.p);
        if (plans[i].wave) {
            args[i].queue = static_cast<ShadowQuery *>(b->wqueue.p);
            args[i].wray = static_cast<float4 *>(b->wray.p);
            args[i].wcq = static_cast<float4 *>(b->wcq.p);
            args[i].walive = static_cast<unsigned int *>(b->walive.p);
            args[i].walive2 = args[i].walive + waveCap;
            args[i].wcnt = static_cast<unsigned int *>(b->wcnt.p);
            args[i].wesq = static_cast<float4 *>(b->wesq.p);
            args[i].wesq_t = reinterpret_cast<float *>(args[i].wesq + 2 * std::max<size_t>(1, escCap));
        } else {
            args[i].queue = static_cast<ShadowQuery *>(b->queue.p);
        }
    }
    // launch groups: every view (shadow), the views that need a new G-buffer, deferred views
    // (shading, accumulation), their spatial-reuse subset, and megakernel views; each gets its own
    // argument array and tile table
    enum { kAll, kGb, kDef, kSpa, kMega, kLean, kWave, kWaveT, kGroups };
    std::vector<int> members[kGroups];
    for (int i = 0; i < n; ++i) {
        members[kAll].push_back(i);
        if (!plans[i].gbuf_reuse) members[kGb].push_back(i);
        if (plans[i].deferred) members[kDef].push_back(i);
        if (plans[i].deferred && plans[i].spatial) members[kSpa].push_back(i);
        if (!plans[i].deferred && !plans[i].lean_pt && !plans[i].wave) members[kMega].push_back(i);
        if (plans[i].lean_pt) members[kLean].push_back(i);
        if (plans[i].wave) members[plans[i].wave_traced ? kWaveT : kWave].push_back(i);
    }
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t off = 0, argOff[kGroups], tileOff[kGroups];
    unsigned tiles[kGroups];
    size_t lds[kGroups];
    int spatialNd[kGroups];  // spatial-reuse launches: the group's largest domain count
    for (int g = 0; g < kGroups; ++g) {
        argOff[g] = off;
        off = align(off + members[g].size() * sizeof(FrameArgs));
        tileOff[g] = off;
        off = align(off + (members[g].size() + 1) * sizeof(int));
        tiles[g] = 0;
        lds[g] = 0;
        spatialNd[g] = 0;
        for (int i : members[g]) {
            tiles[g] += plans[i].tiles;
            lds[g] = std::max(lds[g], plans[i].lds);
            spatialNd[g] = std::max(spatialNd[g], plans[i].spatial_nd);
        }
    }
    hdr_batch::Slot &sl = b->slots[b->next];
    if (sl.pending) {
        HIP_TRY(hipEventSynchronize(sl.done));  // the slot's previous step has finished
        sl.pending = false;
    }
    if (sl.host.n < off) {
        sl.host = PinnedBuf<char>();
        HIP_TRY(sl.host.alloc(off));
    }
    if (hdr_status st = ensure(sl.dev, off)) return st;
    char *hp = sl.host.p;
    const char *dp = static_cast<const char *>(sl.dev.p);
    for (int g = 0; g < kGroups; ++g) {
        FrameArgs *ga = reinterpret_cast<FrameArgs *>(hp + argOff[g]);
        int *tb = reinterpret_cast<int *>(hp + tileOff[g]);
        int base = 0;
        for (size_t k = 0; k < members[g].size(); ++k) {
            ga[k] = args[size_t(members[g][k])];
            if (g == kWave || g == kWaveT) ga[k].view = int(k);  // wavefront queue codes index the group's array
            tb[k] = base;
            base += int(plans[size_t(members[g][k])].tiles);
        }
        tb[members[g].size()] = base;
    }
    // order after each view's earlier work; stage the arguments. The G-buffer halves this step
    // writes were last read two steps ago, or by a single frame of the view since its last step.
    for (int i = 0; i < n; ++i) {
        HIP_TRY(hipEventRecord(views[i]->serial, views[i]->stream));
        HIP_TRY(hipStreamWaitEvent(b->stream, views[i]->serial, 0));
        if (views[i]->single_since_batch) HIP_TRY(hipStreamWaitEvent(b->gstream, views[i]->serial, 0));
    }
    {
        const hdr_batch::Slot &prev2 = b->slots[(b->next + hdr_batch::kSlots - 2) % hdr_batch::kSlots];
        if (prev2.pending) HIP_TRY(hipStreamWaitEvent(b->gstream, prev2.done, 0));
    }
    auto A = [&](int g) { return reinterpret_cast<const FrameArgs *>(dp + argOff[g]); };
    auto T = [&](int g) { return reinterpret_cast<const int *>(dp + tileOff[g]); };
    auto N = [&](int g) { return int(members[g].size()); };
    const LaunchShape ls = b->shape;
    // the step: argument copy and G-buffer on gstream, then every other launch on stream (captured
    // once per slot and launch shape when graphs are on, so later steps replay it)
    auto enqueue_g = [&]() -> hipError_t {
        const hipError_t e = hipMemcpyAsync(sl.dev.p, sl.host.p, off, hipMemcpyHostToDevice, b->gstream);
        if (e != hipSuccess) return e;
        if (N(kGb) > 0) launch_gbuffer<false>(ls, tiles[kGb], b->gstream, A(kGb), T(kGb), N(kGb));
        return hipGetLastError();
    };
    auto enqueue = [&]() -> hipError_t {
        hipLaunchKernelGGL(zero_u32x2_kernel, dim3(1), dim3(1), 0, b->stream, static_cast<unsigned int *>(b->qcount.p));
        if (N(kDef) > 0)
            hipLaunchKernelGGL((shade_direct_kernel<false, false>), dim3(tiles[kDef]), dim3(kLanes), lds[kDef],
                               b->stream, A(kDef), T(kDef), N(kDef));
        if (N(kSpa) > 0) launch_spatial<false, false>(spatialNd[kSpa], dim3(tiles[kSpa]), b->stream, A(kSpa), T(kSpa), N(kSpa));
        if (N(kMega) > 0)
            hipLaunchKernelGGL((integrate_kernel<false, false>), dim3(tiles[kMega]), dim3(kLanes), lds[kMega],
                               b->stream, A(kMega), T(kMega), N(kMega));
        if (N(kLean) > 0)
            launch_pt_lean<false, false>(b->ctx->cus, tiles[kLean], 1u, b->stream, A(kLean), T(kLean), N(kLean));
        for (int wg : {int(kWave), int(kWaveT)}) {
            if (N(wg) == 0) continue;
            const int t = wg == kWaveT ? 1 : 0;
            const int qblocks = queue_blocks(ls, waveCap);
            hipLaunchKernelGGL(wave_zero_counts_kernel, dim3(1), dim3(kLanes), 0, b->stream,
                               static_cast<unsigned int *>(b->wcnt.p));
            launch_wave_shade0<false, false>(waveLean0[t], t == 1, tiles[wg], lds[wg], b->stream, A(wg), T(wg), N(wg));
            // traced frames: shading launch v finishes vertex v-1 (one more launch pair), and the
            // post-temporal reservoirs are stored by launch 1
            const int launches = t == 1 ? waveDepth[t] + 1 : waveDepth[t];
            for (int v = 0; v < launches; ++v) {
                if (v > 0) {
                    const int u = t == 1 ? v - 1 : v;  // the vertex the launch shades
                    if (wavePrepWindow[t] > 0 && u >= wavePrepVertex[t])
                        hipLaunchKernelGGL(wave_prepare_slots_kernel, dim3(qblocks), dim3(kLanes), 0, b->stream, A(wg),
                                           u, wavePrepWindow[t]);
                    launch_wave_shade<false, false>(ls, qblocks, lds[wg], b->stream, A(wg), v, t == 1, waveGen[t]);
                }
                if (waveSpatial[t] && v == t)
                    launch_wave_spatial<false>(spatialNd[wg], dim3(tiles[wg]), b->stream, A(wg), T(wg), N(wg));
                launch_wave_trace<false>(ls, qblocks, b->stream, A(wg), v, t == 1, b->ctx->cus, oneScene, ldsFit);
            }
            hipLaunchKernelGGL((wave_resolve_kernel<false, false>), dim3(tiles[wg]), dim3(kLanes), 0, b->stream,
                               A(wg), T(wg), N(wg));
        }
        if (N(kDef) > 0) {
            launch_shadow<false>(ls, b->ctx->cus, queueCap, b->stream, A(kAll), oneScene, ldsFit);
            hipLaunchKernelGGL((accumulate_deferred_kernel<false>), dim3(tiles[kDef]), dim3(kLanes), 0, b->stream,
                               A(kDef), T(kDef), N(kDef));
        }
        return hipGetLastError();
    };
    HIP_TRY(enqueue_g());
    HIP_TRY(hipEventRecord(b->gdone, b->gstream));
    HIP_TRY(hipStreamWaitEvent(b->stream, b->gdone, 0));
    if (b->graphs) {
        // launch shape key: everything baked into the captured launches (buffers, grid sizes,
        // argument layout); the arguments themselves are copied before the graph runs
        std::vector<uint64_t> key = {uint64_t(n), uint64_t(off), uint64_t(uintptr_t(sl.dev.p)),
                                     uint64_t(uintptr_t(sl.host.p)), uint64_t(uintptr_t(b->queue.p)),
                                     uint64_t(uintptr_t(b->wqueue.p)), uint64_t(uintptr_t(b->wray.p)),
                                     uint64_t(uintptr_t(b->wesq.p)), uint64_t(escCap),
                                     uint64_t(uintptr_t(b->wcq.p)), uint64_t(uintptr_t(b->walive.p)),
                                     uint64_t(queueCap), uint64_t(waveCap), uint64_t(waveDepth[0]),
                                     uint64_t(waveDepth[1]), uint64_t(waveLean0[0]), uint64_t(waveLean0[1]),
                                     uint64_t(waveSpatial[0]), uint64_t(waveSpatial[1]), uint64_t(waveGen[1]),
                                     uint64_t(wavePrepWindow[0]), uint64_t(wavePrepWindow[1]),
                                     uint64_t(uint32_t(wavePrepVertex[0])), uint64_t(uint32_t(wavePrepVertex[1])),
                                     uint64_t(ls.queue_waves), uint64_t(ls.trace_occ), uint64_t(ls.shade_occ),
                                     uint64_t(ls.lds_stack), uint64_t(oneScene), uint64_t(ldsFit)};
        for (int g = 0; g < kGroups; ++g) {
            key.push_back(uint64_t(tiles[g]));
            key.push_back(uint64_t(lds[g]));
            key.push_back(uint64_t(members[g].size()));
            key.push_back(uint64_t(spatialNd[g]));
            for (int i : members[g]) key.push_back(uint64_t(i));
        }
        hdr_batch::Graph &gr = b->graph[b->next];
        if (!gr.exec || gr.key != key) {
            if (gr.exec) {
                HIP_TRY(hipGraphExecDestroy(gr.exec));
                gr.exec = nullptr;
            }
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
            const hipError_t le = enqueue();
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                HIP_TRY(le != hipSuccess ? le : ce);
            }
            const hipError_t ie = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(ie);
            gr.key = std::move(key);
        }
        HIP_TRY(hipEventRecord(b->beg, b->stream));
        HIP_TRY(hipGraphLaunch(gr.exec, b->stream));
    } else {
        HIP_TRY(hipEventRecord(b->beg, b->stream));
        HIP_TRY(enqueue());
    }
    HIP_TRY(hipEventRecord(b->end, b->stream));
    HIP_TRY(hipEventRecord(sl.done, b->stream));
    sl.pending = true;
    sl.step = ++b->issued;
    b->timed = true;
    b->next = (b->next + 1) % hdr_batch::kSlots;
    // later work of each view (reads, single frames, resets) follows this step
    for (int i = 0; i < n; ++i) {
        HIP_TRY(hipStreamWaitEvent(views[i]->stream, sl.done, 0));
        HIP_TRY(hipStreamWaitEvent(views[i]->gstream, sl.done, 0));
        views[i]->frames++;
        views[i]->samples = frames[i].frame_index + 1;
        if (!plans[i].gbuf_reuse) {
            views[i]->gcur = plans[i].gbuf_half;
            views[i]->gvalid = views[i]->gnext_valid;
        }
        views[i]->single_since_batch = false;
        remember_camera(views[i], scenes[i], &frames[i], args[i], false);
    }
    return HDR_OK;
}

// Coordinate search over the launch shape: the traversal stacks' home (private memory or LDS
// rings), queue waves, then the traversal kernels' occupancy variant, then the wavefront shading
// kernel's; each candidate is timed (device time, HIP events)
// over `reps` steps after one untimed step, and the fastest is kept. Every candidate renders the
// same frames exactly (the launch shape moves work between waves, never changes a path's result).
hdr_status hdr_batch_autotune(hdr_batch *b, int32_t n, hdr_view *const *views, const hdr_scene *const *scenes,
                              const hdr_frame_desc *frames, const hdr_pipeline_desc *pipes, int32_t reps,
                              hdr_launch_config *chosen, float *best_step_ms) {
    if (!b || reps < 1 || reps > 64) return fail(HDR_ERR_INVALID_ARGUMENT, "reps must be in [1, 64]");
    HIP_TRY(hipSetDevice(b->ctx->device));
    auto time_shape = [&](const LaunchShape &ls, float &ms) -> hdr_status {
        b->shape = ls;
        for (int w = 0; w < hdr_batch::kSlots; ++w)  // capture every slot's graph outside the timing
            if (hdr_status st = hdr_render_batch(b, n, views, scenes, frames, pipes)) return st;
        HIP_TRY(hipStreamSynchronize(b->stream));
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipEventRecord(e0, b->stream));
        hdr_status st = HDR_OK;
        for (int r = 0; r < reps && st == HDR_OK; ++r) st = hdr_render_batch(b, n, views, scenes, frames, pipes);
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ms = t / float(reps);
        return st;
    };
    LaunchShape best = b->shape;
    float bestMs = 0.0f;
    if (hdr_status st = time_shape(best, bestMs)) return st;
    {
        LaunchShape c = best;
        c.lds_stack = !c.lds_stack;
        float ms;
        if (hdr_status st = time_shape(c, ms)) return st;
        if (ms < bestMs) bestMs = ms, best = c;
    }
    for (int q : {4096, 8192, 16384, 65536}) {
        LaunchShape c = best;
        c.queue_waves = q;
        float ms;
        if (hdr_status st = time_shape(c, ms)) return st;
        if (ms < bestMs) bestMs = ms, best = c;
    }
    for (int t : {1, 4, 6, 8}) {
        LaunchShape c = best;
        c.trace_occ = t;
        float ms;
        if (hdr_status st = time_shape(c, ms)) return st;
        if (ms < bestMs) bestMs = ms, best = c;
    }
    {
        LaunchShape c = best;
        c.shade_occ = c.shade_occ == 3 ? 1 : 3;
        float ms;
        if (hdr_status st = time_shape(c, ms)) return st;
        if (ms < bestMs) bestMs = ms, best = c;
    }
    b->shape = best;
    if (chosen) config_of(best, b->graphs, chosen);
    if (best_step_ms) *best_step_ms = bestMs;
    return HDR_OK;
}

}  // extern "C"

