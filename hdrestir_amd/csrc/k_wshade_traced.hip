// k_wshade_traced.hip -- The wavefront's shading launches of traced frames.
#include "kcommon.h"

namespace hdr {

// Traced frames (wavefront.h: wave_gen): launch v (1 .. depth) shades vertex u = v - 1 of every
// path in vertex u's live list -- its ray queries were answered by the traversal launch before --
// and, for the paths that go on, emits vertex v's ray queries (bounce ray, the NEE candidates'
// shadow rays without skipVisibility, escape rays) and appends the path to vertex v's live list. Vertex u's shadow query joins the queue of parity u.
// kGen: some view of the launch has mesh lights, so its generated candidates are traced here
// (gen_traced_weight); the other builds hold no traversal code.
template <bool kCount, bool kDebug, int kOcc = 1, bool kGen = false>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
wave_shade_traced_kernel(const FrameArgs *__restrict__ args, int v) {
    extern __shared__ double wl[];
    const int lane = threadIdx.x;
    const int u = v - 1;
    unsigned int *cnt = args[0].wcnt;
    const unsigned int n = cnt[4 + (u & 1)];
    PixelStats st = {};
    unsigned int nq = 0;
    for (unsigned int b0 = blockIdx.x * kLanes; b0 < n; b0 += gridDim.x * kLanes) {
        const unsigned int i = b0 + unsigned(lane);
        bool query = false, cont = false, bray = false;
        ShadowQuery q;
        float4 cq = make_float4(0.0f, 0.0f, 0.0f, 0.0f), ro = cq, rd = cq, bo = cq, bd = cq;
        uint32_t code = 0;
        Mt64Deep gr;
        Vertex gv;
        SpecTab tab;
        int k0 = 0, nc = 0, nv = 0;  // nv: NEE candidates whose shadow rays wave_emit_visibility emits
        const FrameArgs *ga = nullptr;
        if (i < n) {
            code = alive_list(args[0], u)[i];
            const FrameArgs &a = args[code >> 24];
            const int pixel = int(code & 0xffffffu);
            WavePath<Mt64Deep> p;
            load_path(a, pixel, p, u == 0);
            tab = make_tab(p.utab);
            HitRec h, next;
            (void)load_grec((u == 0 ? a.gbuf : traced_hits(a, u)) + pixel, h);
            wave_vertex<kCount, kDebug, true, kGen>(a, tab, u, h, p, wl + lane, pixel, st, query, q, cq, cont, ro, rd,
                                              next);
            if (u == 0 && a.spatial_count > 0) a.spatial_aux[2 * pixel + 1] = spec4(p.outer);
            if (cont) {
                uint64_t *slot = wave_slot(a, pixel);
                gr = rng_fork(p.rng, slot ? slot + kMtN : nullptr);
                wave_gen(a, tab, v, next, p, pixel, gr, gv, bray, bo, bd, k0);
                nc = a.n_cand;
                nv = a.skip_vis == 0 && a.s.num_lights > 0 ? a.n_nee : 0;
                ga = &a;
            }
            store_path(a, pixel, p, u == 0);
        }
        const unsigned int qi = wave_append(cnt + 2 + (u & 1), query, lane);
        if (query) {
            args[0].queue[qi] = q;
            args[0].wcq[qi] = cq;
        }
        nq += wave_emit_visibility(cnt + 6 + (v & 1), args[0].wesq, args[0].wesq_t, ga, gv, gr, nv, code, lane);
        wave_emit_escapes(cnt + 6 + (v & 1), args[0].wesq, tab, gv, gr, k0, nc, code, lane);
        const unsigned int ri = wave_append(cnt + (v & 1), bray, lane);
        if (bray) {
            args[0].wray[2 * ri] = bo;
            args[0].wray[2 * ri + 1] = bd;
        }
        const unsigned int ai = wave_append(cnt + 4 + (v & 1), cont, lane);
        if (cont) alive_list(args[0], v)[ai] = code;
        nq += query ? 1u : 0u;
    }
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 2, nq);
        wave_add(c + 3, st.cand);
        wave_add(c + 4, st.cand_skipped);
        if (kGen) {  // the generated candidates' traversals
            wave_add(c + 5, st.ts.nodes);
            wave_add(c + 6, st.ts.tris);
            wave_add(c + 7, st.ts.instances);
        }
    }
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
#define HDR_WST(C, D)                                                                              \
    template __global__ void wave_shade_traced_kernel<C, D, 1, false>(const FrameArgs *__restrict__, int); \
    template __global__ void wave_shade_traced_kernel<C, D, 3, false>(const FrameArgs *__restrict__, int); \
    template __global__ void wave_shade_traced_kernel<C, D, 1, true>(const FrameArgs *__restrict__, int);  \
    template __global__ void wave_shade_traced_kernel<C, D, 3, true>(const FrameArgs *__restrict__, int);
HDR_WST(true, true)
HDR_WST(true, false)
HDR_WST(false, true)
HDR_WST(false, false)
#undef HDR_WST

}  // namespace hdr
