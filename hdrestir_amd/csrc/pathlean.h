// pathlean.h -- the path tracer (MIS direct lighting, any depth) for frames whose every material
// is opaque and whose pixels cannot draw past RNG output 310 (the host checks both): one thread
// per pixel as in the integration megakernel (PathIntegrator::Li / _li, path_integrator.cpp:48-209,
// with MisDirectLightIntegrator::Li, mis_direct_light_integrator.cpp:70-271), but every function
// is inlined -- traversal included -- so there are no call frames; the RNG is the lean stream.
// Results are identical to the megakernel's on such frames.
#pragma once

#include "deferred.h"

namespace hdr {

// MisDirectLightIntegrator::Li at one vertex: light-sampler NEE with its shadow ray, then the BSDF
// technique on the path's own connection (hasConn) and its hit.
template <bool kCount>
__device__ __forceinline__ Spec mis_direct_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                Mt64Lean &rng, bool hasConn, const Bounce &cb, bool cHit,
                                                const HitRec &ch, PixelStats &st, Trace &dbg) {
    const DeviceScene &s = a.s;
    Spec total = spec(0.0f);
    if (s.num_lights > 0) {  // _integrateNEE with the light sampler
        const uint32_t n = uint32_t(s.num_lights);
        const float u = rng.next_float();
        uint32_t idx = uint32_t(u * float(n));
        if (idx > n - 1u) idx = n - 1u;
        dbg.lhash = fnv(dbg.lhash, idx);
        ++dbg.picks;
        const GLight &L = s.lights[idx];
        LSample ls;
        if (sample_light_inl(s, L, v.h.pos, rng, ls)) {
            ls.pdf = ls.pdf * a.light_select_pdf;
            const bool delta = L.is_delta != 0;
            const float nDotL = dot(v.sn, ls.dir);
            const float dist2 = ls.dist * ls.dist;
            const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
            if (nDotL > 0.0f && !(cosY <= 0.0f && !delta)) {
                const float pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
                if (pNee > 0.0f) {
                    if (kCount) ++st.shadow;
                    const bool occ = trace_any<kCount>(s, v.h.pos + v.sn * 1e-4f, ls.dir, ls.dist - 1e-3f, st.ts);
                    if (!occ) {
                        const V3 wo = -v.rayDir;
                        const float bsdfPdf = ggx_pdf(v.c, v.sn, wo, ls.dir);
                        const V3 bsdf = ggx_eval(v.c, v.sn, wo, ls.dir);
                        const Spec radiance = rgb2spec(tab, bsdf) * rgb2spec(tab, ls.color) * (nDotL / pNee);
                        const float w = (delta || !hasConn) ? 1.0f : power2(pNee, bsdfPdf);
                        total = total + radiance * w;
                    }
                }
            }
        }
    }
    if (hasConn) {  // _integrateBSDFConnection, :181-255
        Spec part = spec(0.0f);
        if (cb.impossible || cb.pdf <= 0.0f) {
            if (!cHit) {
                if (s.environment >= 0) part = rgb2spec(tab, env_radiance(s, cb.d)) * cb.tmul;
                else part = spec(0.0f) * cb.tmul;
            } else {
                part = rgb2spec(tab, hit_emission(s, ch)) * cb.tmul;
            }
        } else {
            float pNeeSampler = 0.0f, pNeeEnv = 0.0f;
            Spec radiance = spec(0.0f);
            bool early = false;
            if (!cHit) {
                if (s.environment < 0) {
                    early = true;
                } else {
                    radiance = rgb2spec(tab, env_radiance(s, cb.d));
                    const float envSel = a.light_select_pdf;
                    const float envPdf = env_pdf(s, cb.d);
                    pNeeSampler = envSel * envPdf;
                    pNeeEnv = envSel <= 0.0f ? envPdf : 0.0f;
                }
            } else {
                radiance = rgb2spec(tab, hit_emission(s, ch));
                const int hl = s.instances[ch.inst].light;
                if (hl >= 0) {
                    const float sel = a.light_select_pdf;
                    if (sel > 0.0f) {
                        const V3 off = ch.pos - cb.o;
                        const float dist2 = dot(off, off);
                        const float cosY = smax(0.0f, dot(-cb.d, ch.ng));
                        if (cosY > 0.0f) {
                            bool solid;
                            const float ap = light_eval_pdf(s, s.lights[hl], cb.d, solid);
                            pNeeSampler = sel * to_solid_angle(ap, solid, dist2, cosY);
                        }
                    }
                }
            }
            if (!early) {
                if (pNeeSampler + pNeeEnv <= 0.0f) {
                    part = radiance * cb.tmul;
                } else {
                    const float w = power3(cb.pdf, 1, pNeeSampler, 1, pNeeEnv, 1);
                    part = radiance * w * cb.tmul;
                }
            }
        }
        total = total + part;
    }
    return total;
}

// MisDirectLightIntegrator::Li at one wavefront vertex of a frame whose BSDF rays can reach no
// emitter (no emissive material, environment or mesh light): the light-sampler NEE sample with the
// draws of mis_direct_lean, its shadow ray recorded as a query for the traversal launch instead of
// traced; the returned contribution counts when that ray is unoccluded. The path's own BSDF
// connection (_integrateBSDFConnection) can only reach an emitter-free surface or nothing, so it
// adds nothing and only its existence (hasConn) enters the NEE sample's MIS weight.
template <class R>
__device__ __forceinline__ Spec mis_nee_wave(const FrameArgs &a, const SpecTab &tab, const Vertex &v, R &rng,
                                             bool hasConn, Trace &dbg, bool &query, V3 &qo, V3 &qd, float &qtmax) {
    const DeviceScene &s = a.s;
    query = false;
    if (s.num_lights <= 0) return spec(0.0f);
    const uint32_t n = uint32_t(s.num_lights);
    const float u = rng.next_float();
    uint32_t idx = uint32_t(u * float(n));
    if (idx > n - 1u) idx = n - 1u;
    dbg.lhash = fnv(dbg.lhash, idx);
    ++dbg.picks;
    const GLight &L = s.lights[idx];
    LSample ls;
    if (!sample_light_inl(s, L, v.h.pos, rng, ls)) return spec(0.0f);
    ls.pdf = ls.pdf * a.light_select_pdf;
    const bool delta = L.is_delta != 0;
    const float nDotL = dot(v.sn, ls.dir);
    const float dist2 = ls.dist * ls.dist;
    const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
    if (!(nDotL > 0.0f && !(cosY <= 0.0f && !delta))) return spec(0.0f);
    const float pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
    if (!(pNee > 0.0f)) return spec(0.0f);
    query = true;
    qo = v.h.pos + v.sn * 1e-4f;
    qd = ls.dir;
    qtmax = ls.dist - 1e-3f;
    const V3 wo = -v.rayDir;
    const float bsdfPdf = ggx_pdf(v.c, v.sn, wo, ls.dir);
    const V3 bsdf = ggx_eval(v.c, v.sn, wo, ls.dir);
    const Spec radiance = rgb2spec(tab, bsdf) * rgb2spec(tab, ls.color) * (nDotL / pNee);
    const float w = (delta || !hasConn) ? 1.0f : power2(pNee, bsdfPdf);
    return spec(0.0f) + radiance * w;
}

// PathIntegrator::Li for one pixel, as path_li (integrator.h) with MIS direct lighting; the first
// vertex is iteration 0 of the bounce loop (its roulette test can never fire, and 1 * x = x).
// fidx: the frame index (seed); stage: the frame's sample goes there instead of into the
// accumulation (a launch of several frames, accumulated in frame order afterwards), or null.
template <bool kCount, bool kDebug>
__device__ __forceinline__ void pt_lean_pixel(const FrameArgs &a, int x, int y, PixelStats &st, uint32_t &draws,
                                              int fidx, float4 *stage) {
    const DeviceScene &s = a.s;
    const int pixel = y * a.W + x;
    const uint32_t seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)fidx * 12345ULL);
    HitRec g;
    const bool gHit = load_grec(a.gbuf + pixel, g);
    Mt64Lean rng;
    rng.init(seed, nullptr);
    const float px = float(x) + rng.next_float();
    const float py = float(y) + rng.next_float();
    V3 ro, rd;
    camera_ray(a, px, py, &rng, ro, rd);
    const SpecTab tab = make_tab(rng.next_float());
    Trace dbg = {2166136261u, 2166136261u, 0};
    Spec L = spec(0.0f);
    if (gHit) {
        Spec outer = spec(1.0f), tp = spec(1.0f), total = spec(0.0f);
        int refl = 0, refr = 0;
        bool has = true;
        HitRec cur = g;
        V3 curD = rd;
#pragma unroll 1
        for (int bounce = 0; bounce < a.max_depth; ++bounce) {
            if (!has) break;
            const Clo c = get_closure(s, cur, a.enable_ss != 0);  // opaque: no pass-through draw
            const Vertex v = make_vertex(a, tab, cur, c, curD, bounce == 0 ? outer : tp);
            if (bounce == 0) total = rgb2spec(tab, mv3(c.m->emission));
            Bounce b;
            const bool ok = sample_bounce_inl(tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, a.max_refl, a.max_refr,
                                              refl, refr, rng, b);
            bool cHit = false;
            HitRec ch;
            if (ok) {
                if (bounce == a.max_depth - 1 && !a.trace_last_connection) {
                    if (kCount) ++st.cand_skipped;
                } else {
                    if (kCount) ++st.bounce;
                    RawHit rh;
                    cHit = trace_closest<kCount>(s, b.o, b.d, rh, st.ts);
                    if (cHit) ch = make_hit_record(s, b.o, b.d, rh);
                }
            }
            const Spec d = mis_direct_lean<kCount>(a, tab, v, rng, ok, b, cHit, ch, st, dbg);
            total = total + tp * d;
            if (!ok) break;
            tp = tp * b.tmul;
            curD = b.d;
            has = cHit;
            cur = ch;
            if (b.skip_roulette) continue;
            if (bounce > 3) {
                const float pr = sclamp(spec_max(tp) * a.roulette, 0.0f, 1.0f);
                if (rng.next_float() > pr) break;
                tp = tp * (1.0f / pr);
            }
        }
        L = outer * total;
    } else if (s.environment >= 0 && a.ibl_bg) {
        L = rgb2spec(tab, env_radiance(s, rd));
    }
    V3 rgb = spec2rgb(tab, L);
    if (a.has_cam && a.cam.enable_exposure) rgb = rgb * a.exposure_mul;
    const float4 c4 = make_float4(rgb.x, rgb.y, rgb.z, 1.0f);
    if (stage) {
        stage[pixel] = c4;
    } else if (a.firefly) {
        a.frame_rgb[pixel] = c4;
    } else {  // AccumulationPass (accumulation_pass.cpp:98-106)
        accumulate_sample(a.accum, a.color, a.lum_sum, a.lum_sq, pixel, c4, a.inv_n);
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
        d.sample_rgb[0] = rgb.x;
        d.sample_rgb[1] = rgb.y;
        d.sample_rgb[2] = rgb.z;
        a.dbg[pixel] = d;
    }
    draws = rng.k;
    flag_rng_overrun(a, rng.k > 311u);  // the lean stream stops at output 310
}

}  // namespace hdr
