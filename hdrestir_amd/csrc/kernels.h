// kernels.h -- declarations of the kernels defined in the k_*.hip translation units, for the host
// code in hdrgpu.hip that launches them. Each kernel template is explicitly instantiated, for every
// variant launched here, at the end of its k_*.hip file; the kernel TUs compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "integrator.h"

namespace hdr {

// k_trace.hip
template <bool kCount, bool kLds>
__global__ void gbuffer_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views);
template <bool kCount, int kOcc, bool kLds>
__global__ void shadow_kernel(const FrameArgs *__restrict__ args);
template <bool kCount, int kOcc, int kRefill, bool kOneScene, int kStk>
__global__ void shadow_persistent_kernel(const FrameArgs *__restrict__ args);
template <bool kCount, int kOcc, bool kLds>
__global__ void wave_trace_kernel(const FrameArgs *__restrict__ args, int bounce);
template <bool kCount, int kOcc, bool kLds>
__global__ void wave_trace_traced_kernel(const FrameArgs *__restrict__ args, int v);
template <bool kCount, bool kTraced, bool kOneScene, int kStk>
__global__ void wave_anyhit_kernel(const FrameArgs *__restrict__ args, int v);
__global__ void intersect_kernel(DeviceScene s, const float *o, const float *d, int n, int *inst, int *tri, float *t);

// k_mega.hip
template <bool kCount, bool kDebug>
__global__ void integrate_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views);
template <bool kCount, bool kDebug, int kOcc>
__global__ void pt_lean_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views);

// k_deferred.hip
template <bool kCount, bool kDebug>
__global__ void shade_direct_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base,
                                    int n_views);
template <bool kCount, bool kDebug, int kND>
__global__ void spatial_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views);
template <bool kDebug>
__global__ void accumulate_deferred_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base,
                                           int n_views);
template <bool kCount, int kND>
__global__ void wave_spatial_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base,
                                    int n_views);
template <bool kCount, bool kDebug>
__global__ void wave_resolve_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base,
                                    int n_views);
__global__ void reproject_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base, int n_views);

// k_wave0.hip
template <bool kCount, bool kDebug, bool kLean, bool kTraced>
__global__ void wave_shade0_kernel(const FrameArgs *__restrict__ args, const int *__restrict__ tile_base,
                                   int n_views);

// k_wshade.hip
template <bool kCount, bool kDebug, int kOcc>
__global__ void wave_shade_kernel(const FrameArgs *__restrict__ args, int bounce);
template <bool kCount, bool kDebug, int kOcc, bool kGen>
__global__ void wave_shade_traced_kernel(const FrameArgs *__restrict__ args, int v);
__global__ void wave_prepare_slots_kernel(const FrameArgs *__restrict__ args, int u, int window);

// k_post.hip
__global__ void chroma_blur_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, int W, int H);
__global__ void flare_bright_kernel(const float4 *__restrict__ src, float4 *__restrict__ bloom,
                                    int *__restrict__ row_any, int W);
__global__ void flare_gather_kernel(const float4 *__restrict__ src, const float4 *__restrict__ bloom,
                                    const int *__restrict__ row_any, float4 *__restrict__ dst, int W, int H, int B);
__global__ void chromatic_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, int W, int H, float ca);

}  // namespace hdr
