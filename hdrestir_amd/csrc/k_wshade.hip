// k_wshade.hip -- The wavefront's shading launches of vertices >= 1 (untraced frames).
#include "kcommon.h"

namespace hdr {

// Vertex `bounce` >= 1 of every live path (the codes wave_trace appended).
template <bool kCount, bool kDebug, int kOcc = 1>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
wave_shade_kernel(const FrameArgs *__restrict__ args, int bounce) {
    extern __shared__ double wl[];
    const int lane = threadIdx.x;
    const int par = bounce & 1;
    unsigned int *cnt = args[0].wcnt;
    const unsigned int n = cnt[4 + par];
    if (blockIdx.x == 0 && lane == 0) cnt[4 + (par ^ 1)] = 0u;  // live paths of the next vertex: appended by wave_trace
    PixelStats st = {};
    unsigned int nq = 0;
    for (unsigned int b0 = blockIdx.x * kLanes; b0 < n; b0 += gridDim.x * kLanes) {
        const unsigned int i = b0 + unsigned(lane);
        bool query = false, ray = false;
        ShadowQuery q;
        float4 cq = make_float4(0.0f, 0.0f, 0.0f, 0.0f), ro = cq, rd = cq;
        if (i < n) {
            const uint32_t code = alive_list(args[0], bounce)[i];
            const FrameArgs &a = args[code >> 24];
            const int pixel = int(code & 0xffffffu);
            WavePath<Mt64Deep> p;
            load_path(a, pixel, p);
            const SpecTab tab = make_tab(p.utab);
            HitRec h, next;
            (void)load_grec(a.whit + pixel, h);
            wave_vertex<kCount, kDebug, false, false>(a, tab, bounce, h, p, wl + lane, pixel, st, query, q, cq, ray, ro, rd,
                                               next);
            store_path(a, pixel, p, false);
        }
        const unsigned int qi = wave_append(cnt + 2 + par, query, lane);
        if (query) {
            args[0].queue[qi] = q;
            args[0].wcq[qi] = cq;
        }
        const unsigned int ri = wave_append(cnt + par, ray, lane);
        if (ray) {
            args[0].wray[2 * ri] = ro;
            args[0].wray[2 * ri + 1] = rd;
        }
        nq += query ? 1u : 0u;
    }
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 1, st.bounce);
        wave_add(c + 2, nq);
        wave_add(c + 3, st.cand);
        wave_add(c + 4, st.cand_skipped);
        wave_add(c + 5, st.ts.nodes);
        wave_add(c + 6, st.ts.tris);
        wave_add(c + 7, st.ts.instances);
    }
}

// Before a shading launch: each live path of vertex u whose stream can reach output 624 within
// `window` draws, and whose spill slot is not prepared yet, gets the slot's state at output 624 (the
// third twist, mt_prepare_third) now. Built lazily at output 624, that state took one pass of ~1250
// dependent global-memory steps per lane group reaching it, and the lanes of a wave reach it at
// different draws, so a few waves ran for milliseconds alone at the end of deep frames. Here every
// lane of a wave builds its slot in the same pass, with out-of-place twists whose chunks of loads
// are in flight together; the record's ready bit tells the stream. Paths whose streams never come
// close are only read (two fields of the record).
__global__ void __launch_bounds__(kLanes) wave_prepare_slots_kernel(const FrameArgs *__restrict__ args, int u,
                                                                    int window) {
    const unsigned int n = args[0].wcnt[4 + (u & 1)];
    for (unsigned int i = blockIdx.x * kLanes + threadIdx.x; i < n; i += gridDim.x * kLanes) {
        const uint32_t code = alive_list(args[0], u)[i];
        const FrameArgs &a = args[code >> 24];
        const int pixel = int(code & 0xffffffu);
        uint64_t *slot = wave_slot(a, pixel);
        if (!slot) continue;
        const size_t P = wave_pixels(a);
        const uint32_t k = __float_as_uint(a.wstate[2 * P + pixel].z);
        float *bitsp = &a.wstate[7 * P + pixel].x;
        const uint32_t bits = __float_as_uint(*bitsp);
        if ((bits & kWaveSlotReady) || k >= 624u || k + uint32_t(window) < 624u) continue;
        const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)a.frame_index * 12345ULL);
        mt_prepare_third(slot, slot + kMtN, seed);
        *bitsp = __uint_as_float(bits | kWaveSlotReady);
    }
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
template __global__ void wave_shade_kernel<true, true, 1>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<true, true, 3>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<true, false, 1>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<true, false, 3>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<false, true, 1>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<false, true, 3>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<false, false, 1>(const FrameArgs *__restrict__, int);
template __global__ void wave_shade_kernel<false, false, 3>(const FrameArgs *__restrict__, int);

}  // namespace hdr
