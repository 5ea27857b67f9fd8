// k_trace.hip -- Traversal kernels: the G-buffer (RaycastPass), the deferred path's shadow rays, the
// wavefront's bounce / escape / shadow rays, and closest hits for the parity entry point.
#include "kcommon.h"

namespace hdr {

template <bool kCount, bool kLds = false>
__global__ void __launch_bounds__(kLanes) gbuffer_kernel(const FrameArgs *__restrict__ args,
                                                         const int *__restrict__ tile_base, int n_views) {
    HDR_RING_LDS
    HDR_TILE(args, tile_base, n_views)
    TraceStats ts = {};
    if (kCount || kLds) {  // per-lane traversal (the counting pass counts each ray's own work)
        if (valid)
            gbuffer_pixel<kCount, kLds>(a, a.gbuf, x, y, ts, ring_tr + lane, ring_tt + lane, ring_br + lane,
                                        ring_bt + lane);
    } else {  // the tile's 64 primary rays traverse together
        __shared__ WaveStack wt, wb;
        V3 co = v3s(0.0f), cd = v3(0.0f, 0.0f, 1.0f);
        if (valid) gbuffer_ray(a, x, y, co, cd);
        RawHit rh;
        const bool hit = trace_closest_wave(a.s, valid, co, cd, rh, wt, wb);
        if (valid) gbuffer_store(a, a.gbuf, y * a.W + x, co, cd, hit, rh);
    }
    if (kCount) {
        unsigned long long *c = a.counters;
        wave_add(c + 0, valid ? 1ULL : 0ULL);
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
        wave_add(c + 9, ts.nodes);
        wave_add(c + 10, ts.tris);
    }
}

// Any-hit shadow rays of the queue; an occluded winner keeps zero throughput in its reservoir
// (the visibility fix-up of ris_direct_light_integrator.cpp:163-175).
template <bool kCount, int kOcc = 1, bool kLds = false>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
shadow_kernel(const FrameArgs *__restrict__ args) {
    HDR_RING_LDS
    const int lane = threadIdx.x;
    const FrameArgs &a0 = args[0];  // the queue and its count are shared by the views of the launch
    const unsigned int n = *a0.qcount;
    TraceStats ts = {};
    for (unsigned int i = blockIdx.x * kLanes + threadIdx.x; i < n; i += gridDim.x * kLanes) {
        const ShadowQuery q = a0.queue[i];
        const uint32_t code = __float_as_uint(q.d.w);
        const FrameArgs &a = args[code >> 24];
        const int pixel = int(code & 0xffffffu);
        if (any_hit<kCount, kLds>(a.s, v3(q.o.x, q.o.y, q.o.z), v3(q.d.x, q.d.y, q.d.z), q.o.w, ts, ring_tr + lane,
                                  ring_br + lane)) {
            a.flags[pixel] = 3u;
            hdr_reservoir &r = a.res[pixel];
            r.throughput[0] = r.throughput[1] = r.throughput[2] = r.throughput[3] = 0.0f;
        }
    }
    if (kCount) {
        unsigned long long *c = a0.counters;
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
    }
}

// Shadow ray i of the queue (vertex `vtx`): an unoccluded winner's contribution joins the path
// total (a spatial winner's colour is already staged); an occluded primary winner keeps zero
// throughput in its reservoir (ris_direct_light_integrator.cpp:163-175).
template <bool kCount, bool kLds>
__device__ __forceinline__ void wave_shadow(const FrameArgs *__restrict__ args, unsigned int i, int vtx,
                                            TraceStats &ts, int *ring_tr, int *ring_br) {
    const ShadowQuery q = args[0].queue[i];
    const uint32_t code = __float_as_uint(q.d.w);
    const FrameArgs &a = args[code >> 24];
    const int pixel = int(code & 0xffffffu);
    if (!any_hit<kCount, kLds>(a.s, v3(q.o.x, q.o.y, q.o.z), v3(q.d.x, q.d.y, q.d.z), q.o.w, ts, ring_tr, ring_br)) {
        if (!(vtx == 0 && a.spatial_count > 0)) {
            float4 *t = a.wstate + 4 * wave_pixels(a) + pixel;
            const Spec tot = spec_of(*t) + spec_of(args[0].wcq[i]);
            *t = spec4(tot);
        }
    } else if (vtx == 0 && a.use_res) {
        if (a.spatial_count > 0) a.flags[pixel] |= 1u;  // the spatial winner is occluded
        hdr_reservoir &r = a.res[pixel];
        r.throughput[0] = r.throughput[1] = r.throughput[2] = r.throughput[3] = 0.0f;
    }
}

// The bounce rays (closest hit: the next vertex, appended to the live list when there is one) and
// the shadow rays (any hit: an unoccluded winner's contribution joins the path total; an occluded
// primary winner keeps zero throughput in its reservoir, ris_direct_light_integrator.cpp:163-175)
// of vertex `bounce`.
template <bool kCount, int kOcc = 1, bool kLds = false>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
wave_trace_kernel(const FrameArgs *__restrict__ args, int bounce) {
    HDR_RING_LDS
    const int lane = threadIdx.x;
    const int par = bounce & 1;
    unsigned int *cnt = args[0].wcnt;
    const unsigned int nr = cnt[par], ns = cnt[2 + par];
    if (blockIdx.x == 0 && lane == 0) {
        cnt[par ^ 1] = 0u;        // rays of the next vertex
        cnt[2 + (par ^ 1)] = 0u;  // shadow queries of the next vertex
    }
    TraceStats ts = {};
    unsigned int nb = 0;
    for (unsigned int b0 = blockIdx.x * kLanes; b0 < nr; b0 += gridDim.x * kLanes) {
        const unsigned int i = b0 + unsigned(lane);
        bool hit = false;
        uint32_t code = 0;
        if (i < nr) {
            const float4 o4 = args[0].wray[2 * i], d4 = args[0].wray[2 * i + 1];
            code = __float_as_uint(o4.w);
            const FrameArgs &a = args[code >> 24];
            const int pixel = int(code & 0xffffffu);
            const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
            RawHit rh;
            hit = closest_hit<kCount, kLds>(a.s, o, d, rh, ts, ring_tr + lane, ring_tt + lane, ring_br + lane,
                                            ring_bt + lane);
            if (hit) store_grec(a.whit + pixel, true, make_hit_record(a.s, o, d, rh));
            ++nb;
        }
        const unsigned int ai = wave_append(cnt + 4 + (par ^ 1), hit, lane);
        if (hit) alive_list(args[0], bounce + 1)[ai] = code;
    }
    for (unsigned int i = blockIdx.x * kLanes + unsigned(lane); i < ns; i += gridDim.x * kLanes)
        wave_shadow<kCount, kLds>(args, i, bounce, ts, ring_tr + lane, ring_br + lane);
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 1, nb);
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
    }
}

// Traced frames: traversal launch v (0 .. depth) after shading launch v -- vertex v's bounce rays
// (closest hit: vertex v+1's hit record, or a miss record) and escape rays (any hit up to 1e30:
// the candidate's bit of the pixel's escape mask when none), and vertex v-1's shadow rays.
template <bool kCount, int kOcc = 1, bool kLds = false>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
wave_trace_traced_kernel(const FrameArgs *__restrict__ args, int v) {
    HDR_RING_LDS
    const int lane = threadIdx.x;
    const int par = v & 1;
    unsigned int *cnt = args[0].wcnt;
    const unsigned int nr = cnt[par], ne = cnt[6 + par], ns = v > 0 ? cnt[2 + (par ^ 1)] : 0u;
    if (blockIdx.x == 0 && lane == 0) {  // the counts the next shading launch appends to
        cnt[par ^ 1] = 0u;        // bounce rays of vertex v+1
        cnt[6 + (par ^ 1)] = 0u;  // escape rays of vertex v+1
        cnt[2 + par] = 0u;        // shadow queries of vertex v
        cnt[4 + (par ^ 1)] = 0u;  // live paths of vertex v+1
    }
    TraceStats ts = {};
    unsigned int nb = 0;
    for (unsigned int i = blockIdx.x * kLanes + unsigned(lane); i < nr; i += gridDim.x * kLanes) {
        const float4 o4 = args[0].wray[2 * i], d4 = args[0].wray[2 * i + 1];
        const uint32_t code = __float_as_uint(o4.w);
        const FrameArgs &a = args[code >> 24];
        const int pixel = int(code & 0xffffffu);
        const V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
        RawHit rh;
        GRec *dst = traced_hits(a, v + 1) + pixel;
        if (closest_hit<kCount, kLds>(a.s, o, d, rh, ts, ring_tr + lane, ring_tt + lane, ring_br + lane,
                                      ring_bt + lane))
            store_grec(dst, true, make_hit_record(a.s, o, d, rh));
        else
            store_grec_miss(dst);
        ++nb;
    }
    for (unsigned int i = blockIdx.x * kLanes + unsigned(lane); i < ne; i += gridDim.x * kLanes) {
        const float4 o4 = args[0].wesq[2 * i], d4 = args[0].wesq[2 * i + 1];
        const uint32_t code = __float_as_uint(o4.w);
        const FrameArgs &a = args[code >> 24];
        const int pixel = int(code & 0xffffffu);
        if (!any_hit<kCount, kLds>(a.s, v3(o4.x, o4.y, o4.z), v3(d4.x, d4.y, d4.z), 1e30f, ts, ring_tr + lane,
                                   ring_br + lane))
            atomicOr(a.wesc + pixel, 1u << __float_as_int(d4.w));
    }
    for (unsigned int i = blockIdx.x * kLanes + unsigned(lane); i < ns; i += gridDim.x * kLanes)
        wave_shadow<kCount, kLds>(args, i, v - 1, ts, ring_tr + lane, ring_br + lane);
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 1, nb);
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
    }
}

__global__ void intersect_kernel(DeviceScene s, const float *o, const float *d, int n, int *inst, int *tri,
                                 float *t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    TraceStats ts = {};
    RawHit h;
    const V3 oo = v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
    const V3 dd = v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    if (trace_closest<false>(s, oo, dd, h, ts)) {
        const HitRec r = make_hit_record(s, oo, dd, h);
        inst[i] = r.inst;
        tri[i] = r.tri;
        t[i] = r.depth;
    } else {
        inst[i] = -1;
        tri[i] = -1;
        t[i] = 0.0f;
    }
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
template __global__ void gbuffer_kernel<true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void gbuffer_kernel<true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void gbuffer_kernel<false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void gbuffer_kernel<false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void shadow_kernel<true, 1, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 1, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 4, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 4, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 6, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 6, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 8, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 8, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 1, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 1, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 4, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 4, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 6, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 6, false>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 8, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 8, false>(const FrameArgs *__restrict__);
template __global__ void wave_trace_kernel<true, 1, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 1, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 4, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 4, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 6, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 6, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 8, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<true, 8, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 1, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 1, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 4, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 4, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 6, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 6, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 8, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_kernel<false, 8, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 1, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 1, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 4, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 4, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 6, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 6, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 8, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<true, 8, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 1, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 1, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 4, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 4, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 6, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 6, false>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 8, true>(const FrameArgs *__restrict__, int);
template __global__ void wave_trace_traced_kernel<false, 8, false>(const FrameArgs *__restrict__, int);

}  // namespace hdr
