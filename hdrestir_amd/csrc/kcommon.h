// kcommon.h -- device helpers shared by the kernel translation units (k_*.hip): wave sums, the
// traversal-stack homes, the G-buffer ray and record, the launch's view / tile lookup, the shadow
// queue append and the wavefront live lists.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/hdrgpu.h"
#include "pathlean.h"
#include "wavefront.h"

namespace hdr {

// ============================================================================ kernels

// RaycastPass (raycast_pass.cpp:57-162): closest hit of the pixel-centre ray, one 8x8 tile per wave.
// Sum over the 64 lanes of the wave (every lane active), added once by lane 0.
__device__ __forceinline__ void wave_add(unsigned long long *dst, unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

// LDS rings of the traversal stacks (RingStack): the top 4 TLAS and 8 BLAS entries of each lane.
constexpr int kRingT = 4, kRingB = 8;
#define HDR_RING_LDS                                \
    __shared__ int ring_tr[kRingT * kLanes];        \
    __shared__ float ring_tt[kRingT * kLanes];      \
    __shared__ int ring_br[kRingB * kLanes];        \
    __shared__ float ring_bt[kRingB * kLanes];

// Closest hit with the stacks in private memory or with LDS rings (the lane's columns).
template <bool kCount, bool kLds>
__device__ __forceinline__ bool closest_hit(const DeviceScene &s, V3 o, V3 d, RawHit &rh, TraceStats &ts, int *tr,
                                            float *tt, int *br, float *bt) {
    if (kLds) {
        int ptr[kStackDepth], pbr[kStackDepth];
        float ptt[kStackDepth], pbt[kStackDepth];
        RingStack<kRingT, true> t;
        RingStack<kRingB, true> b;
        t.bind(tr, tt, ptr, ptt);
        b.bind(br, bt, pbr, pbt);
        return trace_closest_s<kCount>(s, o, d, rh, ts, t, b);
    }
    return trace_closest<kCount>(s, o, d, rh, ts);
}
template <bool kCount, bool kLds>
__device__ __forceinline__ bool any_hit(const DeviceScene &s, V3 o, V3 d, float tmax, TraceStats &ts, int *tr,
                                        int *br) {
    if (kLds) {
        int ptr[kStackDepth], pbr[kStackDepth];
        RingStack<kRingT, false> t;
        RingStack<kRingB, false> b;
        t.bind(tr, nullptr, ptr, nullptr);
        b.bind(br, nullptr, pbr, nullptr);
        return trace_any_s<kCount>(s, o, d, tmax, ts, t, b);
    }
    return trace_any<kCount>(s, o, d, tmax, ts);
}

// The G-buffer ray of a pixel (RaycastPass: pixel centre; the camera's lens draws when it has
// camera parameters).
__device__ __forceinline__ void gbuffer_ray(const FrameArgs &a, int x, int y, V3 &co, V3 &cd) {
    const int pixel = y * a.W + x;
    if (a.has_cam) {
        const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)a.frame_index * 12345ULL);
        Mt64 crng;
        crng.init(seed, nullptr);
        camera_ray(a, float(x) + 0.5f, float(y) + 0.5f, &crng, co, cd);
    } else {
        camera_ray(a, float(x) + 0.5f, float(y) + 0.5f, static_cast<Mt64 *>(nullptr), co, cd);
    }
}
// The hit record and AOVs of a pixel.
__device__ __forceinline__ void gbuffer_store(const FrameArgs &a, GRec *gbuf, int pixel, V3 co, V3 cd, bool hit,
                                              const RawHit &rh) {
    HitRec g = {};
    if (hit) g = make_hit_record(a.s, co, cd, rh);
    store_grec(gbuf + pixel, hit, g);
    if (a.aov_depth) {
        a.aov_depth[pixel] = hit ? g.depth : 1.0f;
        const V3 al = hit ? g.albedo : v3s(0.0f);
        const V3 nn = hit ? g.ng : v3s(0.0f);
        a.aov_albedo[3 * pixel + 0] = al.x;
        a.aov_albedo[3 * pixel + 1] = al.y;
        a.aov_albedo[3 * pixel + 2] = al.z;
        a.aov_normal[3 * pixel + 0] = nn.x;
        a.aov_normal[3 * pixel + 1] = nn.y;
        a.aov_normal[3 * pixel + 2] = nn.z;
    }
}
template <bool kCount, bool kLds = false>
__device__ __forceinline__ void gbuffer_pixel(const FrameArgs &a, GRec *gbuf, int x, int y, TraceStats &ts, int *tr,
                                              float *tt, int *br, float *bt) {
    V3 co, cd;
    gbuffer_ray(a, x, y, co, cd);
    RawHit rh;
    const bool hit = closest_hit<kCount, kLds>(a.s, co, cd, rh, ts, tr, tt, br, bt);
    gbuffer_store(a, gbuf, y * a.W + x, co, cd, hit, rh);
}

// Launch layout: one wave64 per 8x8 pixel tile; the blocks of a launch enumerate the tiles of
// every view in it, view v owning blocks tile_base[v] .. tile_base[v + 1) (a 1-D grid).
__device__ __forceinline__ int launch_view(const int *__restrict__ tile_base, int n_views) {
    const int b = int(blockIdx.x);
    int lo = 0, hi = n_views - 1;
    while (lo < hi) {  // last view whose first block is <= b (uniform across the block)
        const int mid = (lo + hi + 1) >> 1;
        if (tile_base[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define HDR_TILE(args, tile_base, n_views)                                   \
    const int view_ = launch_view(tile_base, n_views);                        \
    const FrameArgs &a = (args)[view_];                                       \
    const int tile_ = int(blockIdx.x) - (tile_base)[view_];                   \
    const int tw_ = (a.W + 7) >> 3;                                           \
    const int lane = threadIdx.x;                                             \
    const int x = (tile_ % tw_) * 8 + (lane & 7);                             \
    const int y = (tile_ / tw_) * 8 + (lane >> 3);                            \
    const bool valid = x < a.W && y < a.H;


// Wave-aggregated append of this lane's shadow query (if any) to the launch's queue.
__device__ __forceinline__ void append_query(const FrameArgs &a, int lane, bool query, const ShadowQuery &q) {
    const unsigned long long mask = __ballot(query);
    const int n = __popcll(mask);
    if (n > 0) {
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(a.qcount, unsigned(n));
        base = __shfl(base, 0);
        if (query) {
            const int rank = __popcll(mask & ((1ULL << lane) - 1ULL));
            a.queue[base + rank] = q;
        }
    }
}

// Live-path list of vertex v (two lists, by vertex parity, so a launch can read one and append to
// the other).
__device__ __forceinline__ unsigned int *alive_list(const FrameArgs &a0, int v) {
    return (v & 1) ? a0.walive2 : a0.walive;
}

}  // namespace hdr
