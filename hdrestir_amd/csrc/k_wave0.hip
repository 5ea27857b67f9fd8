// k_wave0.hip -- The wavefront's primary-vertex launch.
#include "kcommon.h"

namespace hdr {

template <bool kCount, bool kDebug, bool kLean = false, bool kTraced = false>
__global__ void __launch_bounds__(kLanes) wave_shade0_kernel(const FrameArgs *__restrict__ args,
                                                             const int *__restrict__ tile_base, int n_views) {
    extern __shared__ double wl[];
    HDR_TILE(args, tile_base, n_views)
    unsigned int *cnt = args[0].wcnt;  // the queues of a launch are shared by its views
    // live paths of vertex 1: appended by wave_trace (traced frames: this launch appends vertex
    // 0's live list after the frame-start clear)
    if (!kTraced && blockIdx.x == 0 && lane == 0) cnt[5] = 0u;
    PixelStats st = {};
    bool query = false, ray = false, bray = false;
    ShadowQuery q;
    float4 cq = make_float4(0.0f, 0.0f, 0.0f, 0.0f), ro = cq, rd = cq;
    HitRec next;
    uint32_t code = 0;
    using R = typename std::conditional<kLean, Mt64Lean, Mt64Deep>::type;
    R gr;           // traced: the stream copy wave_gen leaves at the first generated candidate
    Vertex gv;      // and the vertex
    SpecTab tab;
    int k0 = 0, nc = 0, nv = 0;  // nv: NEE candidates whose shadow rays wave_emit_visibility emits
    if (valid) {
        const int pixel = y * a.W + x;
        const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)a.frame_index * 12345ULL);
        WavePath<R> p;
        p.rng.init(seed, wave_slot(a, pixel));
        const float px = float(x) + p.rng.next_float();
        const float py = float(y) + p.rng.next_float();
        V3 o0, d0;
        camera_ray(a, px, py, &p.rng, o0, d0);
        p.utab = p.rng.next_float();
        tab = make_tab(p.utab);
        p.dbg = Trace{2166136261u, 2166136261u, 0};
        p.tp = spec(1.0f);
        p.outer = spec(1.0f);
        p.total = spec(0.0f);
        p.refl = p.refr = 0;
        p.dir = d0;
        HitRec g;
        p.primary_hit = load_grec(a.gbuf + pixel, g);
        if (kTraced) {  // vertex 0's ray queries; its shading runs in the next shading launch
            ray = p.primary_hit;
            if (ray) {
                gr = rng_fork(p.rng, wave_slot(a, pixel) ? wave_slot(a, pixel) + kMtN : nullptr);
                wave_gen(a, tab, 0, g, p, pixel, gr, gv, bray, ro, rd, k0);
                nc = a.n_cand;
                nv = a.skip_vis == 0 && a.s.num_lights > 0 ? a.n_nee : 0;
            }
        } else if (p.primary_hit) {
            wave_vertex<kCount, kDebug, kTraced, false>(a, tab, 0, g, p, wl + lane, pixel, st, query, q, cq, ray, ro, rd,
                                                 next);
        }
        if (!p.primary_hit && a.s.environment >= 0 && a.ibl_bg)  // a miss sees the environment (PathIntegrator::Li, path_integrator.cpp:48-209)
            p.total = rgb2spec(tab, env_radiance(a.s, d0));
        if (a.spatial_count > 0) {  // the spatial launch's record of the primary vertex
            a.spatial_aux[2 * pixel] = make_float4(d0.x, d0.y, d0.z, p.primary_hit ? p.utab : -1.0f);
            // traced frames: the throughput is recorded where vertex 0 is shaded
            if (!kTraced && p.primary_hit) a.spatial_aux[2 * pixel + 1] = spec4(p.outer);
        }
        if (kLean) flag_rng_overrun(a, p.rng.k > 311u);  // the lean stream stops at output 310
        store_path(a, pixel, p, true);
        code = query_code(a, pixel);
        if (kDebug) {
            a.dbg[pixel].prim_instance = p.primary_hit ? g.inst : -1;
            a.dbg[pixel].prim_triangle = p.primary_hit ? g.tri : -1;
            a.dbg[pixel].prim_t = p.primary_hit ? g.depth : 0.0f;
        }
    }
    unsigned int nvis = 0;  // traced frames without skipVisibility: the NEE candidates' shadow rays
    const unsigned int qi = wave_append(cnt + 2, query, lane);
    if (query) {
        args[0].queue[qi] = q;
        args[0].wcq[qi] = cq;
    }
    if (kTraced) {  // vertex 0's live list, bounce rays and escape rays
        const unsigned int ai = wave_append(cnt + 4, ray, lane);
        if (ray) alive_list(args[0], 0)[ai] = code;
        nvis = wave_emit_visibility(cnt + 6, args[0].wesq, args[0].wesq_t, &a, gv, gr, nv, code, lane);
        wave_emit_escapes(cnt + 6, args[0].wesq, tab, gv, gr, k0, nc, code, lane);
        if (kLean && valid && ray) flag_rng_overrun(a, gr.k > 311u);  // the generation's lean stream too
        ray = bray;
    }
    const unsigned int ri = wave_append(cnt + 0, ray, lane);
    if (ray) {
        args[0].wray[2 * ri] = ro;
        args[0].wray[2 * ri + 1] = rd;
    }
    if (kCount) {
        unsigned long long *c = a.counters;
        wave_add(c + 1, st.bounce);
        wave_add(c + 2, (query ? 1ULL : 0ULL) + nvis);
        wave_add(c + 3, st.cand);
        wave_add(c + 4, st.cand_skipped);
        wave_add(c + 5, st.ts.nodes);
        wave_add(c + 6, st.ts.tris);
        wave_add(c + 7, st.ts.instances);
    }
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
template __global__ void wave_shade0_kernel<true, true, true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, true, true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, true, false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, true, false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, false, true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, false, true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, false, false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<true, false, false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, true, true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, true, true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, true, false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, true, false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, false, true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, false, true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, false, false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_shade0_kernel<false, false, false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);

}  // namespace hdr
