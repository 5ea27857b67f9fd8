// k_deferred.hip -- The deferred-shadow path (shading, spatial reuse, accumulation) and the wavefront's spatial
// and resolve launches.
#include "kcommon.h"

namespace hdr {

// ---------------------------------------------------------------------------- deferred-shadow path
// RIS at depth 1 with skipVisibility in a scene where nothing BSDF-sampled can reach an emitter:
// the only ray after the G-buffer is the RIS winner's shadow ray. Shading runs without traversal
// code (more waves in flight), the shadow rays are compacted into a queue for an any-hit kernel,
// and a last pass picks the visible / occluded colour and accumulates.


template <bool kCount, bool kDebug>
__device__ __forceinline__ void shade_direct_pixel(const FrameArgs &a, const GRec *gbuf, int x, int y, double *wlane,
                                                   PixelStats &st, uint32_t &draws, float4 *rgb2, uint32_t *flags,
                                                   bool &query, ShadowQuery &q) {
    const int pixel = y * a.W + x;
    const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)a.frame_index * 12345ULL);
    HitRec g;
    const bool gHit = load_grec(gbuf + pixel, g);
    Mt64Lean rng;
    rng.init(seed, nullptr);
    const float px = float(x) + rng.next_float();
    const float py = float(y) + rng.next_float();
    V3 ro, rd;
    camera_ray(a, px, py, &rng, ro, rd);
    const float utab = rng.next_float();
    const SpecTab tab = make_tab(utab);
    Trace dbg = {2166136261u, 2166136261u, 0};
    V3 qo = v3s(0.0f), qd = v3s(0.0f);
    float qtmax = 0.0f;
    query = false;
    Spec lVis = spec(0.0f), lOcc = spec(0.0f);
    if (gHit) {  // PathIntegrator::Li / _li at depth 1 (path_integrator.cpp:112-177); opacity >= 0.999
        const Clo c = get_closure(a.s, g, a.enable_ss != 0);
        Spec throughput = spec(1.0f);
        const Vertex v0 = make_vertex(a, tab, g, c, rd, throughput);
        const Spec total = rgb2spec(tab, mv3(c.m->emission));
        // the path's BSDF connection: sampled, never traced (it can reach no emitter)
        int refl = 0, refr = 0;
        Bounce first;
        const bool ok = sample_bounce_inl(tab, v0.c, v0.sn, v0.inside, v0.h.pos, v0.rayDir, a.max_refl, a.max_refr,
                                          refl, refr, rng, first);
        if (kCount && ok) ++st.cand_skipped;
        const Spec direct = ris_direct_lean<kCount, kDebug>(a, tab, v0, ok && first.pdf > 0.0f, rng, wlane, pixel, st,
                                                           dbg, a.use_res != 0, nullptr, query, qo, qd, qtmax);
        lVis = throughput * (total + direct);
        lOcc = throughput * (total + spec(0.0f));
        if (a.spatial_count > 0) {  // the spatial pass finishes this pixel
            a.spatial_aux[2 * pixel] = make_float4(rd.x, rd.y, rd.z, utab);
            a.spatial_aux[2 * pixel + 1] =
                make_float4(throughput.v[0], throughput.v[1], throughput.v[2], throughput.v[3]);
        }
    } else if (a.s.environment >= 0 && a.ibl_bg) {
        lVis = lOcc = rgb2spec(tab, env_radiance(a.s, rd));
    }
    if (a.spatial_count > 0 && !gHit) a.spatial_aux[2 * pixel] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    V3 cv = spec2rgb(tab, lVis), co = spec2rgb(tab, lOcc);
    if (a.has_cam && a.cam.enable_exposure) {
        cv = cv * a.exposure_mul;
        co = co * a.exposure_mul;
    }
    if (!(a.spatial_count > 0 && gHit)) {
        rgb2[pixel] = make_float4(cv.x, cv.y, cv.z, 1.0f);
        rgb2[size_t(a.W) * a.H + pixel] = make_float4(co.x, co.y, co.z, 1.0f);
        flags[pixel] = query ? 1u : 0u;
    }
    if (query) {
        q.o = make_float4(qo.x, qo.y, qo.z, qtmax);
        q.d = make_float4(qd.x, qd.y, qd.z, __uint_as_float(query_code(a, pixel)));
    }
    if (kDebug) {
        hdr_pixel_debug d;
        d.draws = rng.k;
        d.light_hash = dbg.lhash;
        d.select_hash = dbg.shash;
        d.n_light_picks = dbg.picks;
        d.prim_instance = gHit ? g.inst : -1;
        d.prim_triangle = gHit ? g.tri : -1;
        d.prim_t = gHit ? g.depth : 0.0f;
        d.sample_rgb[0] = cv.x;
        d.sample_rgb[1] = cv.y;
        d.sample_rgb[2] = cv.z;
        a.dbg[pixel] = d;
    }
    draws = rng.k;
    flag_rng_overrun(a, rng.k > 311u);  // the lean stream stops at output 310
}

template <bool kCount, bool kDebug>
__global__ void __launch_bounds__(kLanes) shade_direct_kernel(const FrameArgs *__restrict__ args,
                                                              const int *__restrict__ tile_base, int n_views) {
    extern __shared__ double wl[];
    HDR_TILE(args, tile_base, n_views)
    PixelStats st = {};
    uint32_t draws = 0;
    bool query = false;
    ShadowQuery q;
    if (valid)
        shade_direct_pixel<kCount, kDebug>(a, a.gbuf, x, y, wl + lane, st, draws, a.rgb2, a.flags, query, q);
    append_query(a, lane, query, q);
    if (kCount) {
        unsigned long long *c = a.counters;
        wave_add(c + 2, query ? 1ULL : 0ULL);
        wave_add(c + 4, st.cand_skipped);
        wave_add(c + 8, draws);
    }
}

// Spatial reuse (deferred.h: spatial_reuse_pixel) and the final colours of the shaded pixels; the
// winner's shadow query joins the queue as in shade_direct_kernel.
template <bool kCount, bool kDebug, int kND>
__global__ void __launch_bounds__(kLanes) spatial_kernel(const FrameArgs *__restrict__ args,
                                                         const int *__restrict__ tile_base, int n_views) {
    __shared__ int dl[(kMaxSpatial + 1) * kLanes];
    __shared__ float cl[(kMaxSpatial + 1) * kLanes];
    HDR_TILE(args, tile_base, n_views)
    const GRec *gbuf = a.gbuf;
    float4 *rgb2 = a.rgb2;
    uint32_t *flags = a.flags;
    bool query = false;
    ShadowQuery q;
    if (valid) {
        const int pixel = y * a.W + x;
        if (a.hist_g) a.hist_g[pixel] = gbuf[pixel];  // for the next frame's reprojection
        Vertex vq;
        SpecTab tab;
        if (load_spatial_vertex(a, gbuf, pixel, vq, tab)) {
            V3 qo = v3s(0.0f), qd = v3s(0.0f);
            float qtmax = 0.0f;
            int ssel = -1;
            const Spec direct = spatial_reuse_pixel<kDebug, kND>(a, gbuf, x, y, vq, tab, dl + lane, cl + lane, query, qo,
                                                            qd, qtmax, ssel);
            const float4 t4 = a.spatial_aux[2 * pixel + 1];
            const Spec throughput = Spec{{t4.x, t4.y, t4.z, t4.w}};
            const Spec total = rgb2spec(tab, mv3(vq.c.m->emission));
            const Spec lVis = throughput * (total + direct);
            const Spec lOcc = throughput * (total + spec(0.0f));
            V3 cv = spec2rgb(tab, lVis), co = spec2rgb(tab, lOcc);
            if (a.has_cam && a.cam.enable_exposure) {
                cv = cv * a.exposure_mul;
                co = co * a.exposure_mul;
            }
            rgb2[pixel] = make_float4(cv.x, cv.y, cv.z, 1.0f);
            rgb2[size_t(a.W) * a.H + pixel] = make_float4(co.x, co.y, co.z, 1.0f);
            flags[pixel] = query ? 1u : 0u;
            if (query) {
                q.o = make_float4(qo.x, qo.y, qo.z, qtmax);
                q.d = make_float4(qd.x, qd.y, qd.z, __uint_as_float(query_code(a, pixel)));
            }
            if (kDebug) a.dbg[pixel].select_hash = fnv(a.dbg[pixel].select_hash, uint32_t(ssel));
        }
    }
    append_query(a, lane, query, q);
    if (kCount) wave_add(a.counters + 2, query ? 1ULL : 0ULL);
}

// Picks the visible or occluded colour and accumulates (AccumulationPass, accumulation_pass.cpp:98-106).
template <bool kDebug>
__global__ void __launch_bounds__(kLanes) accumulate_deferred_kernel(const FrameArgs *__restrict__ args,
                                                                     const int *__restrict__ tile_base,
                                                                     int n_views) {
    HDR_TILE(args, tile_base, n_views)
    if (!valid) return;
    const int i = y * a.W + x;
    const float4 c4 = a.rgb2[(a.flags[i] == 3u ? size_t(a.W) * a.H : size_t(0)) + i];
    if (kDebug) {
        a.dbg[i].sample_rgb[0] = c4.x;
        a.dbg[i].sample_rgb[1] = c4.y;
        a.dbg[i].sample_rgb[2] = c4.z;
    }
    if (a.firefly) {
        a.frame_rgb[i] = c4;
        return;
    }
    accumulate_sample(a.accum, a.color, a.lum_sum, a.lum_sq, i, c4, a.inv_n);
}

// Spatial reuse of the primary vertex (deferred.h: spatial_reuse_pixel), after wave_shade0 has
// stored every pixel's post-temporal reservoir. Per shaded pixel: the primary vertex's emission
// plus the reused direct light (rgb2[p]) and plus nothing (rgb2[P + p]), as spectra, and flags[p]
// = (selected domain slot + 1) << 1; the winner's shadow ray (none when it was tested in an
// earlier frame) joins vertex 0's shadow queue, whose traversal sets bit 0 of the flags when it
// is occluded.
template <bool kCount, int kND>
__global__ void __launch_bounds__(kLanes) wave_spatial_kernel(const FrameArgs *__restrict__ args,
                                                              const int *__restrict__ tile_base, int n_views) {
    __shared__ int dl[(kMaxSpatial + 1) * kLanes];
    __shared__ float cl[(kMaxSpatial + 1) * kLanes];
    HDR_TILE(args, tile_base, n_views)
    bool query = false;
    ShadowQuery q;
    if (valid && a.spatial_count > 0) {
        const int pixel = y * a.W + x;
        if (a.hist_g) a.hist_g[pixel] = a.gbuf[pixel];  // for the next frame's reprojection
        Vertex vq;
        SpecTab tab;
        if (load_spatial_vertex(a, a.gbuf, pixel, vq, tab)) {
            V3 qo = v3s(0.0f), qd = v3s(0.0f);
            float qtmax = 0.0f;
            int ssel = -1;
            const Spec direct = spatial_reuse_pixel<false, kND>(a, a.gbuf, x, y, vq, tab, dl + lane, cl + lane, query, qo,
                                                           qd, qtmax, ssel);
            const Spec e = rgb2spec(tab, mv3(vq.c.m->emission));
            a.rgb2[pixel] = spec4(e + direct);
            a.rgb2[size_t(a.W) * a.H + pixel] = spec4(e + spec(0.0f));
            a.flags[pixel] = uint32_t(ssel + 1) << 1;
            if (query) {
                q.o = make_float4(qo.x, qo.y, qo.z, qtmax);
                q.d = make_float4(qd.x, qd.y, qd.z, __uint_as_float(query_code(a, pixel)));
            }
        }
    }
    const unsigned int qi = wave_append(args[0].wcnt + 2, query, lane);
    if (query) args[0].queue[qi] = q;
    if (kCount) wave_add(a.counters + 2, query ? 1ULL : 0ULL);
}

// Reprojected temporal reuse (an extension): before this frame's shading and spatial passes
// overwrite them, copies the previous frame's records at the pixel that saw each pixel's primary
// hit (reproject_pixel) -- its G-buffer record (as the previous spatial pass kept it), its {ray dir,
// wavelength u} and its final reservoir -- into the pixel's rp_* slots, which the spatial pass reads
// as one more domain; rp_aux.w = -1 when there is none. Per pixel 64 + 16 + 80 B read and written,
// coalesced on the write side. Everything it reads is written on the view's stream, so the next
// frame's G-buffer launch on the second stream does not wait for it.
__global__ void __launch_bounds__(kLanes) reproject_kernel(const FrameArgs *__restrict__ args,
                                                           const int *__restrict__ tile_base, int n_views) {
    HDR_TILE(args, tile_base, n_views)
    if (!valid) return;
    const int pixel = y * a.W + x;
    HitRec g;
    const int j = load_grec(a.gbuf + pixel, g) ? reproject_pixel(a, g) : -1;
    if (j < 0) {
        a.rp_aux[pixel] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        return;
    }
    a.rp_g[pixel] = a.hist_g[j];
    a.rp_aux[pixel] = a.spatial_aux[2 * j];
    a.rp_res[pixel] = a.res[j];
}

// The path's radiance (outer throughput x total) to RGB, exposure, accumulation
// (integration_pass.cpp:236-243, accumulation_pass.cpp:98-106).
template <bool kCount, bool kDebug>
__global__ void __launch_bounds__(kLanes) wave_resolve_kernel(const FrameArgs *__restrict__ args,
                                                              const int *__restrict__ tile_base, int n_views) {
    HDR_TILE(args, tile_base, n_views)
    uint32_t draws = 0;
    if (valid) {
        const int pixel = y * a.W + x;
        const size_t P = wave_pixels(a);
        const float4 *w = a.wstate;
        const float4 r2 = w[2 * P + pixel], r5 = w[5 * P + pixel], r7 = w[7 * P + pixel];
        const SpecTab tab = make_tab(r5.w);
        const bool hit = (__float_as_uint(r7.x) & kWavePrimaryHit) != 0;
        uint32_t shash = __float_as_uint(r7.z);
        Spec L;
        if (hit && a.spatial_count > 0) {  // outer * ((emission + reused direct) + later vertices)
            const uint32_t f = a.flags[pixel];
            const Spec e = spec_of(a.rgb2[((f & 1u) ? P : size_t(0)) + pixel]);
            L = spec_of(w[6 * P + pixel]) * (e + spec_of(w[4 * P + pixel]));
            if (kDebug) shash = fnv(shash, uint32_t(int(f >> 1) - 1));
        } else {
            L = hit ? spec_of(w[6 * P + pixel]) * spec_of(w[4 * P + pixel]) : spec_of(w[4 * P + pixel]);
        }
        V3 rgb = spec2rgb(tab, L);
        if (a.has_cam && a.cam.enable_exposure) rgb = rgb * a.exposure_mul;
        const float4 c4 = make_float4(rgb.x, rgb.y, rgb.z, 1.0f);
        if (a.firefly) {
            a.frame_rgb[pixel] = c4;
        } else {
            accumulate_sample(a.accum, a.color, a.lum_sum, a.lum_sq, pixel, c4, a.inv_n);
        }
        draws = __float_as_uint(r2.z);
        flag_rng_overrun(a, !wave_slot(a, pixel) && draws > 624u);  // outputs from 624 on need the slot
        if (kDebug) {
            hdr_pixel_debug &d = a.dbg[pixel];
            d.draws = draws;
            d.light_hash = __float_as_uint(r7.y);
            d.select_hash = shash;
            d.n_light_picks = __float_as_int(r7.w);
            d.sample_rgb[0] = rgb.x;
            d.sample_rgb[1] = rgb.y;
            d.sample_rgb[2] = rgb.z;
        }
    }
    if (kCount) wave_add(a.counters + 8, draws);
}

// ---------------------------------------------------------------------------- instantiations
// (launched from hdrgpu.hip, which sees the declarations in kernels.h)
template __global__ void shade_direct_kernel<true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void shade_direct_kernel<true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void shade_direct_kernel<false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void shade_direct_kernel<false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, true, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, true, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, true, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, false, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, false, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<true, false, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, true, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, true, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, true, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, false, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, false, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void spatial_kernel<false, false, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void accumulate_deferred_kernel<true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void accumulate_deferred_kernel<false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<true, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<true, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<true, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<false, 4>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<false, 6>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_spatial_kernel<false, 9>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_resolve_kernel<true, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_resolve_kernel<true, false>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_resolve_kernel<false, true>(const FrameArgs *__restrict__, const int *__restrict__, int);
template __global__ void wave_resolve_kernel<false, false>(const FrameArgs *__restrict__, const int *__restrict__, int);

}  // namespace hdr
