// deferred.h -- shading body of the deferred-shadow path: RIS / ReSTIR direct lighting at the
// primary hit (PathIntegrator::_li at depth 1, path_integrator.cpp:92-121, with
// RisDirectLightIntegrator::Li, ris_direct_light_integrator.cpp:84-215) for frames where
//   - skipVisibility is set, so candidates are ranked without shadow rays,
//   - no material is translucent (no pass-through draws), no material emits, no mesh is a light
//     and there is no environment, so every BSDF-sampled candidate and the path's own BSDF
//     connection have zero target (their rays can reach no emitter), and
//   - no pixel can draw past RNG output 310.
// The host checks all three. Under them the per-pixel work is straight-line code with no ray
// queries except the winner's shadow ray, which is recorded for the shadow kernel: everything
// here is inlined into the kernel, so no call frames and no scratch memory are needed, and the
// BSDF candidates reduce to their RNG draws (they still advance the stream exactly as the
// reference does). Results are identical to the general path (integrator.h) on such frames.
#pragma once

#include "integrator.h"

namespace hdr {

// G-buffer record of one pixel: 4 x float4 = 64 B, written and read coalesced.
struct GRec {
    float4 a;  // pos.xyz, depth
    float4 b;  // ng.xyz, instance (bits; -1 = miss)
    float4 c;  // ns.xyz, triangle (bits)
    float4 d;  // albedo.xyz, unused
};

__device__ __forceinline__ void store_grec(GRec *dst, bool hit, const HitRec &g) {
    float4 *p = reinterpret_cast<float4 *>(dst);
    gstore4(p, make_float4(g.pos.x, g.pos.y, g.pos.z, g.depth));
    gstore4(p + 1, make_float4(g.ng.x, g.ng.y, g.ng.z, __int_as_float(hit ? g.inst : -1)));
    gstore4(p + 2, make_float4(g.ns.x, g.ns.y, g.ns.z, __int_as_float(g.tri)));
    gstore4(p + 3, make_float4(g.albedo.x, g.albedo.y, g.albedo.z, 0.0f));
}
// A record that says "no hit" (load_grec returns false for it).
__device__ __forceinline__ void store_grec_miss(GRec *dst) {
    gstore4(reinterpret_cast<float4 *>(dst) + 1, make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1)));
}
__device__ __forceinline__ bool load_grec(const GRec *src, HitRec &g) {
    const float4 *p = reinterpret_cast<const float4 *>(src);
    GRec r;
    r.a = gload4(p);
    r.b = gload4(p + 1);
    r.c = gload4(p + 2);
    r.d = gload4(p + 3);
    g.pos = v3(r.a.x, r.a.y, r.a.z);
    g.depth = r.a.w;
    g.ng = v3(r.b.x, r.b.y, r.b.z);
    g.inst = __float_as_int(r.b.w);
    g.ns = v3(r.c.x, r.c.y, r.c.z);
    g.tri = __float_as_int(r.c.w);
    g.albedo = v3(r.d.x, r.d.y, r.d.z);
    return g.inst >= 0;
}

// A shadow query of the deferred path: the winner's shadow ray, answered by the any-hit kernel.
struct ShadowQuery {
    float4 o;  // origin.xyz, tmax
    float4 d;  // dir.xyz, view << 24 | pixel (bits)
};
// Launch view and pixel of a shadow query (the host keeps views below 2^24 pixels, batches below
// 256 views).
__device__ __forceinline__ uint32_t query_code(const FrameArgs &a, int pixel) {
    return (uint32_t(a.view) << 24) | uint32_t(pixel);
}


// UniformLightSampler::ProposeCandidate light pick (uniform_light_sampler.cpp:22-47).
template <bool kDebug, class R>
__device__ __forceinline__ int pick_light_lean(const FrameArgs &a, R &rng, Trace &dbg, bool record) {
    const uint32_t n = uint32_t(a.s.num_lights);
    const float u = rng.next_float();
    uint32_t idx = uint32_t(u * float(n));
    if (idx > n - 1u) idx = n - 1u;
    if (kDebug && record) {
        dbg.lhash = fnv(dbg.lhash, idx);
        ++dbg.picks;
    }
    return int(idx);
}

// One light-sampler NEE candidate (_generateNEECandidates :407-443 with _evaluateLightSample
// :482-527): returns the RIS weight, fills `out` when given. vis < 0: skipVisibility (no shadow ray,
// VisibilityTested false); 0 / 1: the candidate's shadow ray, traced beforehand, was occluded /
// unoccluded (wave_emit_visibility) -- an occluded candidate keeps zero contribution.
template <class R>
__device__ __forceinline__ double nee_candidate_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                     R &rng, int li, Cand *out, int vis = -1) {
    const DeviceScene &s = a.s;
    const GLight &L = s.lights[li];
    LSample ls;
    if (!sample_light_inl(s, L, v.h.pos, rng, ls)) return 0.0;
    ls.pdf = ls.pdf * a.light_select_pdf;
    const bool delta = L.is_delta != 0;
    const V3 wo = -v.rayDir;
    const float nDotL = dot(v.sn, ls.dir);
    const float dist2 = ls.dist * ls.dist;
    const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
    const float pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
    Spec contribution = spec(0.0f);
    float pBsdf = 0.0f;
    bool visTested = true;
    if (!(nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !delta))) {
        pBsdf = ggx_pdf(v.c, v.sn, wo, ls.dir);
        const V3 bsdf = ggx_eval(v.c, v.sn, wo, ls.dir);
        const Spec tmul = rgb2spec(tab, bsdf * nDotL);
        if (vis != 0) contribution = rgb2spec(tab, ls.color) * tmul;
        visTested = vis >= 0;
    }
    const float target = spec_lum(tab, contribution);
    const float mis = !delta ? power3(pNee, a.n_nee, 0.0f, a.n_sky, pBsdf, a.n_bsdf) : 1.0f;
    if (out) {
        out->has_cand = true;
        out->delta = delta;
        out->solid = ls.solid_angle;
        out->vis_tested = visTested;
        out->light = li;
        out->dir = ls.dir;
        out->color = ls.color;
        out->lnormal = ls.lnormal;
        out->dist = ls.dist;
        out->pdf = ls.pdf;
        out->tp = contribution;
        out->target = target;
    }
    return double(mis * target) / (double(pNee) + 1e-9);
}

// Replays the candidate stream from `rng` (a copy taken before candidate 0) and rebuilds NEE
// candidate `kk` (vis: as nee_candidate_lean); earlier candidates only consume their draws.
template <bool kDebug, class R>
__device__ __forceinline__ void rebuild_nee_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                 R rng, int kk, Trace &dbg, Cand &out, int vis = -1) {
    const DeviceScene &s = a.s;
    for (int k = 0; k < kk; ++k) {
        const int li = pick_light_lean<kDebug>(a, rng, dbg, false);
        const int nd = light_draws(s, s.lights[li]);
        for (int q = 0; q < nd; ++q) rng.discard();
    }
    const int li = pick_light_lean<kDebug>(a, rng, dbg, false);
    (void)nee_candidate_lean(a, tab, v, rng, li, &out, vis);
}

// The winner's deferred shadow ray (the visibility fix-up of :159-175 through _evaluateLightSample
// with a shadow ray): the query is recorded and the contribution is the unoccluded one; the
// shadow kernel zeroes it when the ray is blocked.
__device__ __forceinline__ Spec winner_contribution_lean(const SpecTab &tab, const Vertex &v, const Cand &ch,
                                                         bool &query, V3 &qo, V3 &qd, float &qtmax) {
    const V3 wo = -v.rayDir;
    const float nDotL = dot(v.sn, ch.dir);
    const float dist2 = ch.dist * ch.dist;
    const float cosY = smax(0.0f, dot(-ch.dir, ch.lnormal));
    const float pNee = to_solid_angle(ch.pdf, ch.solid, dist2, smax(cosY, 1e-6f));
    if (nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !ch.delta)) return spec(0.0f);
    const V3 bsdf = ggx_eval(v.c, v.sn, wo, ch.dir);
    const Spec tmul = rgb2spec(tab, bsdf * nDotL);
    query = true;
    qo = v.h.pos + v.sn * 1e-4f;
    qd = ch.dir;
    qtmax = ch.dist - 1e-3f;
    return rgb2spec(tab, ch.color) * tmul;
}

// A copy of the stream for the winner's replay (the lean stream has no spill slot to share).
__device__ __forceinline__ Mt64Lean rng_fork(const Mt64Lean &r, uint64_t *) { return r; }

// RisDirectLightIntegrator::Li on a vertex whose BSDF candidates can reach no emitter. With
// `reservoir` (use_reservoir at the primary vertex) the streaming reservoir and the temporal merge
// run (:121-185), otherwise RandomIndexFromWeights (:187-214). wl: this lane's NEE candidate weights
// in LDS (wl[k * kLanes]); slot2: the spare spill slot for the replay copy of a deep stream.
template <bool kCount, bool kDebug, class R>
__device__ __forceinline__ Spec ris_direct_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                bool connCand, R &rng, double *wl, int pixel,
                                                PixelStats &st, Trace &dbg, bool reservoir, uint64_t *slot2,
                                                bool &query, V3 &qo, V3 &qd, float &qtmax) {
    const DeviceScene &s = a.s;
    const int nNee = a.n_nee;
    const R start = rng_fork(rng, slot2);
    // NEE candidates
    for (int k = 0; k < nNee; ++k) {
        double w = 0.0;
        if (s.num_lights > 0) {
            const int li = pick_light_lean<kDebug>(a, rng, dbg, true);
            w = nee_candidate_lean(a, tab, v, rng, li, nullptr);
        }
        wl[k * kLanes] = w;
    }
    // BSDF candidates: the path's connection (zero target) and the generated ones, whose rays can
    // reach no emitter on this path (std::nullopt, :341-368); only their draws remain.
    const int nGen = a.n_bsdf - (connCand ? 1 : 0);
    for (int k = 0; k < nGen; ++k) {
        skip_bounce_draws_inl(tab, v.c, v.sn, v.inside, v.rayDir, rng);
        if (kCount) ++st.cand_skipped;
    }
    const int zeroCands = a.n_cand - nNee;  // candidates with weight 0 (counted by the reservoir)

    if (reservoir) {
        // streaming WRS over the candidates (WeightedReservoir::Update, reservoir.h:76-84)
        float wsum = 0.0f;
        int count = 0;
        int sel = -1;
        for (int q = 0; q < nNee; ++q) {
            ++count;
            const float w = float(wl[q * kLanes]);
            if (w <= 0.0f) continue;
            wsum += w;
            if (rng.next_float() <= w / wsum) sel = q;
        }
        count += zeroCands;
        // temporal merge with the stored reservoir of this pixel (:127-155)
        const hdr_reservoir prev = temporal_prev(a, pixel);
        const uint32_t pflags = prev.flags_light >> 24;
        Cand prevC;
        if ((pflags & HDR_RES_HAS_SAMPLE) && (pflags & HDR_RES_HAS_CANDIDATE)) {
            prevC.has_cand = true;
            prevC.delta = (pflags & HDR_RES_LIGHT_DELTA) != 0;
            prevC.solid = (pflags & HDR_RES_PDF_SOLID_ANGLE) != 0;
            prevC.vis_tested = (pflags & HDR_RES_VIS_TESTED) != 0;
            prevC.light = int(prev.flags_light & 0xffffffu) - 1;
            prevC.dir = v3(prev.dir[0], prev.dir[1], prev.dir[2]);
            prevC.color = v3(prev.color[0], prev.color[1], prev.color[2]);
            prevC.lnormal = v3(prev.light_normal[0], prev.light_normal[1], prev.light_normal[2]);
            prevC.dist = prev.dist;
            prevC.pdf = prev.pdf;
            // _reEvaluateForTemporalReuse (:529-545): stored direction / distance, no shadow ray
            Spec reeval = spec(0.0f);
            {
                const V3 wo = -v.rayDir;
                const float nDotL = dot(v.sn, prevC.dir);
                const float dist2 = prevC.dist * prevC.dist;
                const float cosY = smax(0.0f, dot(-prevC.dir, prevC.lnormal));
                const float pNee = to_solid_angle(prevC.pdf, prevC.solid, dist2, smax(cosY, 1e-6f));
                if (!(nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !prevC.delta))) {
                    const V3 bsdf = ggx_eval(v.c, v.sn, wo, prevC.dir);
                    reeval = rgb2spec(tab, prevC.color) * rgb2spec(tab, bsdf * nDotL);
                }
            }
            const float phatCurr = spec_lum(tab, reeval);
            const float phatGen = prev.target;
            if (phatGen > 0.0f) {
                prevC.target = phatCurr;
                prevC.tp = reeval;
                const float mergeWeight = prev.wsum * (phatCurr / phatGen);
                count += prev.count;
                if (mergeWeight > 0.0f) {
                    wsum += mergeWeight;
                    if (rng.next_float() <= mergeWeight / wsum) sel = -2;
                }
            }
        }
        if (kDebug) dbg.shash = fnv(dbg.shash, uint32_t(sel));
        Cand ch;
        bool has = false;
        if (sel >= 0) {
            rebuild_nee_lean<kDebug>(a, tab, v, start, sel, dbg, ch);
            has = true;
        } else if (sel == -2) {
            ch = prevC;
            has = true;
        }
        // Finalize (reservoir.h:108-125)
        float W = 0.0f;
        if (has && count > 0) {
            const float phat = ch.target;
            if (phat > 0.0f) {
                const int perFrame = a.n_cand > 1 ? a.n_cand : 1;
                int frameCount = count / perFrame;
                if (frameCount < 1) frameCount = 1;
                W = wsum / (float(frameCount) * phat);
                if (!isfinite(W)) W = 0.0f;
            }
        }
        if (a.spatial_count > 0) {  // spatial reuse follows: it tests the final winner
            store_reservoir(&a.res_t[pixel], has, ch, wsum, count, W);
            return spec(0.0f);
        }
        if (has && !ch.vis_tested && ch.has_cand) {
            ch.tp = winner_contribution_lean(tab, v, ch, query, qo, qd, qtmax);
            ch.vis_tested = true;
        }
        store_reservoir(&a.res[pixel], has, ch, wsum, count, W);
        if (!has || W <= 0.0f) return spec(0.0f);
        return ch.tp * W;
    }

    // RIS::RandomIndexFromWeights (random_index.h:27-63)
    double sum = 0.0;
    for (int q = 0; q < nNee; ++q) {
        const double w = wl[q * kLanes];
        if (w > 0.0) sum += w;
    }
    int chosen = -1;
    if (sum > 0.0) {
        double rem = double(rng.next_float()) * sum;
        for (int q = 0; q < nNee; ++q) {
            const double w = wl[q * kLanes];
            if (w <= 0.0) continue;
            rem -= w;
            if (rem <= 0.0) {
                chosen = q;
                break;
            }
        }
    }
    if (kDebug) dbg.shash = fnv(dbg.shash, uint32_t(chosen));
    if (chosen < 0) return spec(0.0f);
    Cand ch;
    rebuild_nee_lean<kDebug>(a, tab, v, start, chosen, dbg, ch);
    const float W = float(sum / double(ch.target));
    Spec tp = ch.tp;
    if (!ch.vis_tested && ch.has_cand) tp = winner_contribution_lean(tab, v, ch, query, qo, qd, qtmax);
    return tp * W;
}

// ---------------------------------------------------------------- traced candidates
// RisDirectLightIntegrator::Li on a vertex whose BSDF-sampled candidates can see an emitter: the
// environment (escaped generated candidates) or an emissive material (the path's own connection).
// Without mesh lights a generated candidate that hits geometry has zero weight (:341-368) and only
// needs to know whether its ray escapes -- an any-hit query with no bound, which accepts exactly the
// hits the reference's closest-hit query would find. With mesh lights (gen_inline) its closest hit
// decides it, traced in the shading launch (gen_traced_weight). The connection candidate uses the
// bounce ray's closest hit, traced by the caller before this call.

// _evaluateBSDFConnection (:547-634) via the connection's hit, with the candidate's RIS weight
// (:263-294).
__device__ __forceinline__ double conn_weight_lean(const FrameArgs &a, const SpecTab &tab, const Bounce &b, bool hasHit,
                                                   const HitRec &hit, Cand *out) {
    const DeviceScene &s = a.s;
    Spec emission = spec(0.0f);
    float pNee = 0.0f;
    if (!hasHit) {
        if (s.environment >= 0) {
            const float sel = a.light_select_pdf;
            const float envPdf = env_pdf(s, b.d);
            pNee = sel > 0.0f ? sel * envPdf : (a.n_sky > 0 ? envPdf : 0.0f);
            emission = rgb2spec(tab, env_radiance(s, b.d));
        }
    } else {
        emission = rgb2spec(tab, hit_emission(s, hit));
        const int hl = s.instances[hit.inst].light;
        if (hl >= 0) {
            const float sel = a.light_select_pdf;
            const V3 off = hit.pos - b.o;
            const float dist2 = dot(off, off);
            const float cosY = smax(0.0f, dot(-b.d, hit.ng));
            if (cosY > 0.0f && sel > 0.0f) {
                bool solid;
                const float ap = light_eval_pdf(s, s.lights[hl], b.d, solid);
                pNee = sel * to_solid_angle(ap, solid, dist2, cosY);
            }
        }
    }
    const Spec contribution = emission * b.imul;
    const float target = spec_lum(tab, contribution);
    float pNeeSampler = pNee, pNeeSky = 0.0f;
    if (!hasHit && (s.environment < 0 || a.light_select_pdf <= 0.0f)) {
        pNeeSampler = 0.0f;
        pNeeSky = pNee;
    }
    const double mis = double(power3(b.pdf, a.n_bsdf, pNeeSampler, a.n_nee, pNeeSky, a.n_sky));
    if (out) {
        out->has_cand = false;
        out->vis_tested = true;
        out->light = -1;
        out->tp = contribution;
        out->target = target;
    }
    return mis * double(target) / (double(b.pdf) + 1e-9);
}

// A generated BSDF candidate whose ray escaped to the environment (:296-340).
__device__ __forceinline__ double gen_env_weight(const FrameArgs &a, const SpecTab &tab, const Bounce &b, Cand *out) {
    const DeviceScene &s = a.s;
    if (s.environment < 0) return 0.0;
    const V3 wi = b.d;
    const V3 envRGB = env_radiance(s, wi);
    const Spec envEm = rgb2spec(tab, envRGB);
    const float pBsdf = b.pdf;
    const float sel = a.light_select_pdf;
    const float envPdf = env_pdf(s, wi);
    const float pNeeSampler = sel * envPdf;
    const float pNeeSky = (sel <= 0.0f && a.n_sky > 0) ? envPdf : 0.0f;
    const float pNee = pNeeSampler + pNeeSky;
    const Spec tp = envEm * b.imul;
    const float target = spec_lum(tab, tp);
    const double mis = double(power3(pBsdf, a.n_bsdf, pNeeSampler, a.n_nee, pNeeSky, a.n_sky));
    if (out) {
        out->has_cand = pNee > 0.0f;
        out->delta = false;
        out->solid = true;
        out->vis_tested = true;
        out->light = s.environment;
        out->dir = wi;
        out->color = envRGB;
        out->lnormal = -wi;
        out->dist = 1e30f;
        out->pdf = pNee;
        out->tp = tp;
        out->target = target;
    }
    return mis * double(target) / (double(pBsdf) + 1e-9);
}

// A generated BSDF candidate in a scene with mesh lights (_generateBSDFCandidates :296-402, as the
// integration megakernel's gen_bsdf_candidate): its ray's closest hit, traced here out of line,
// decides it. An escape sees the environment (gen_env_weight); a hit on an instance registered as a
// light sees that light's emission (_evaluateBSDFConnection :547-634, weighted against the light's
// pdf); any other hit is no candidate. `escaped` reports the escape.
template <bool kCount>
__device__ __forceinline__ double gen_traced_weight(const FrameArgs &a, const SpecTab &tab, const Bounce &b,
                                                    TraceStats &ts, bool &escaped, Cand *out) {
    const DeviceScene &s = a.s;
    RawHit rh;
    escaped = !trace_closest_call<kCount>(s, b.o, b.d, rh, ts);
    if (escaped) return gen_env_weight(a, tab, b, out);
    const HitRec hit = make_hit_record_call(s, b.o, b.d, rh);
    const int hl = s.instances[hit.inst].light;
    if (hl < 0) return 0.0;
    const V3 emissionRGB = hit_emission(s, hit);
    const Spec emission = rgb2spec(tab, emissionRGB);
    const float sel = a.light_select_pdf;
    float pNee = 0.0f;
    {
        const V3 off = hit.pos - b.o;
        const float dist2 = dot(off, off);
        const float cosY = smax(0.0f, dot(-b.d, hit.ng));
        if (cosY > 0.0f && sel > 0.0f) {
            bool solid;
            const float ap = light_eval_pdf(s, s.lights[hl], b.d, solid);
            pNee = sel * to_solid_angle(ap, solid, dist2, cosY);
        }
    }
    const Spec contribution = emission * b.imul;
    const float target = spec_lum(tab, contribution);
    const double mis = double(power3(b.pdf, a.n_bsdf, pNee, a.n_nee, 0.0f, a.n_sky));
    if (out) {
        bool solid;
        float lp = light_eval_pdf(s, s.lights[hl], b.d, solid);
        lp *= sel;
        out->has_cand = true;
        out->delta = s.lights[hl].is_delta != 0;
        out->solid = solid;
        out->vis_tested = true;
        out->light = hl;
        out->dir = b.d;
        out->color = emissionRGB;
        out->lnormal = hit.ng;
        out->dist = hit.depth;
        out->pdf = lp;
        out->tp = contribution;
        out->target = target;
    }
    return mis * double(target) / (double(b.pdf) + 1e-9);
}

// Rebuilds candidate kk by replaying the stream from `rng` (taken before candidate 0): NEE
// candidates as rebuild_nee_lean, the connection from its hit, a generated candidate by replaying
// the draws before it and sampling it again (esc: which generated candidates escaped and, without
// skipVisibility, which NEE candidates' shadow rays were unoccluded).
template <bool kDebug, bool kGen, class R>
__device__ __forceinline__ void rebuild_traced(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                               const Bounce *conn, bool connHit, const HitRec &connRec,
                                               bool connCand, R rng, uint32_t esc, int kk, Trace &dbg, Cand &out) {
    const DeviceScene &s = a.s;
    const int nNee = a.n_nee;
    if (kk < nNee) {
        rebuild_nee_lean<kDebug>(a, tab, v, rng, kk, dbg, out, a.skip_vis ? -1 : int((esc >> kk) & 1u));
        return;
    }
    for (int k = 0; k < nNee; ++k) {
        const int li = pick_light_lean<kDebug>(a, rng, dbg, false);
        const int nd = light_draws(s, s.lights[li]);
        for (int q = 0; q < nd; ++q) rng.discard();
    }
    int k = nNee + a.n_sky;
    if (connCand) {
        if (k == kk) {
            (void)conn_weight_lean(a, tab, *conn, connHit, connRec, &out);
            return;
        }
        ++k;
    }
    for (; k < kk; ++k) skip_bounce_draws_inl(tab, v.c, v.sn, v.inside, v.rayDir, rng);
    Bounce b;
    int refl = 0, refr = 0;
    const bool sampled = sample_bounce_inl(tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, 1, 1, refl, refr, rng, b);
    if (kGen && a.gen_inline) {  // the winner's ray again (not counted: the weight pass counted it)
        TraceStats unused = {};
        bool escaped;
        if (sampled) (void)gen_traced_weight<false>(a, tab, b, unused, escaped, &out);
    } else if ((esc >> kk) & 1u) {
        (void)gen_env_weight(a, tab, b, &out);
    }
}

// RisDirectLightIntegrator::Li (:84-215) with traced BSDF candidates; weights of every candidate in
// LDS (wl[k * kLanes], k < n_cand). Shadow query and return value as ris_direct_lean. The
// generated candidates' escape tests (any hit up to 1e30, from the sampled ray) were answered
// beforehand: bit k of `escm` is set when candidate k's ray escapes. Without skipVisibility the NEE
// candidates' shadow rays were answered too: bit k < n_nee is set when candidate k's is unoccluded.
template <bool kCount, bool kDebug, bool kGen, class R>
__device__ __forceinline__ Spec ris_direct_traced(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                  const Bounce *conn, bool connHit, const HitRec &connRec, R &rng,
                                                  double *wl, int pixel, PixelStats &st, Trace &dbg, bool reservoir,
                                                  uint64_t *slot2, uint32_t escm, bool &query, V3 &qo, V3 &qd,
                                                  float &qtmax) {
    const DeviceScene &s = a.s;
    const int nNee = a.n_nee, nCand = a.n_cand;
    const R start = rng_fork(rng, slot2);
    for (int k = 0; k < nNee; ++k) {
        double w = 0.0;
        if (s.num_lights > 0) {
            const int li = pick_light_lean<kDebug>(a, rng, dbg, true);
            w = nee_candidate_lean(a, tab, v, rng, li, nullptr, a.skip_vis ? -1 : int((escm >> k) & 1u));
        }
        wl[k * kLanes] = w;
    }
    int k = nNee + a.n_sky;
    const bool connCand = conn != nullptr && conn->pdf > 0.0f;
    if (connCand) {
        wl[k * kLanes] = conn_weight_lean(a, tab, *conn, connHit, connRec, nullptr);
        ++k;
    }
    // the NEE candidates' visibility travels with the escape bits to the winner's rebuild
    uint32_t esc = a.skip_vis ? 0u : (escm & ((nNee < 32 ? (1u << nNee) : 0u) - 1u));
    for (; k < nCand; ++k) {
        double w = 0.0;
        if (!a.trace_generated) {  // nothing a generated candidate can reach: only its draws
            skip_bounce_draws_inl(tab, v.c, v.sn, v.inside, v.rayDir, rng);
            if (kCount) ++st.cand_skipped;
        } else {
            Bounce b;
            int refl = 0, refr = 0;
            if (sample_bounce_inl(tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, 1, 1, refl, refr, rng, b)) {
                if (kCount) ++st.cand;
                if (kGen && a.gen_inline) {
                    bool escaped;
                    w = gen_traced_weight<kCount>(a, tab, b, st.ts, escaped, nullptr);
                    if (escaped) esc |= 1u << k;
                } else if ((escm >> k) & 1u) {
                    esc |= 1u << k;
                    w = gen_env_weight(a, tab, b, nullptr);
                }
            }
        }
        wl[k * kLanes] = w;
    }

    Cand ch;
    bool has = false;
    if (reservoir) {
        float wsum = 0.0f;
        int count = 0;
        int sel = -1;
        for (int q = 0; q < nCand; ++q) {  // WeightedReservoir::Update (reservoir.h:76-84)
            ++count;
            const float w = float(wl[q * kLanes]);
            if (w <= 0.0f) continue;
            wsum += w;
            if (rng.next_float() <= w / wsum) sel = q;
        }
        const hdr_reservoir prev = temporal_prev(a, pixel);  // temporal merge (:127-155)
        const uint32_t pflags = prev.flags_light >> 24;
        Cand prevC;
        if ((pflags & HDR_RES_HAS_SAMPLE) && (pflags & HDR_RES_HAS_CANDIDATE)) {
            prevC.has_cand = true;
            prevC.delta = (pflags & HDR_RES_LIGHT_DELTA) != 0;
            prevC.solid = (pflags & HDR_RES_PDF_SOLID_ANGLE) != 0;
            prevC.vis_tested = (pflags & HDR_RES_VIS_TESTED) != 0;
            prevC.light = int(prev.flags_light & 0xffffffu) - 1;
            prevC.dir = v3(prev.dir[0], prev.dir[1], prev.dir[2]);
            prevC.color = v3(prev.color[0], prev.color[1], prev.color[2]);
            prevC.lnormal = v3(prev.light_normal[0], prev.light_normal[1], prev.light_normal[2]);
            prevC.dist = prev.dist;
            prevC.pdf = prev.pdf;
            Spec reeval = spec(0.0f);  // _reEvaluateForTemporalReuse (:529-545)
            {
                const V3 wo = -v.rayDir;
                const float nDotL = dot(v.sn, prevC.dir);
                const float dist2 = prevC.dist * prevC.dist;
                const float cosY = smax(0.0f, dot(-prevC.dir, prevC.lnormal));
                const float pNee = to_solid_angle(prevC.pdf, prevC.solid, dist2, smax(cosY, 1e-6f));
                if (!(nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !prevC.delta))) {
                    const V3 bsdf = ggx_eval(v.c, v.sn, wo, prevC.dir);
                    reeval = rgb2spec(tab, prevC.color) * rgb2spec(tab, bsdf * nDotL);
                }
            }
            const float phatCurr = spec_lum(tab, reeval);
            const float phatGen = prev.target;
            if (phatGen > 0.0f) {
                prevC.target = phatCurr;
                prevC.tp = reeval;
                const float mergeWeight = prev.wsum * (phatCurr / phatGen);
                count += prev.count;
                if (mergeWeight > 0.0f) {
                    wsum += mergeWeight;
                    if (rng.next_float() <= mergeWeight / wsum) sel = -2;
                }
            }
        }
        if (kDebug) dbg.shash = fnv(dbg.shash, uint32_t(sel));
        if (sel >= 0) {
            rebuild_traced<kDebug, kGen>(a, tab, v, conn, connHit, connRec, connCand, start, esc, sel, dbg, ch);
            has = true;
        } else if (sel == -2) {
            ch = prevC;
            has = true;
        }
        float W = 0.0f;  // Finalize (reservoir.h:108-125)
        if (has && count > 0) {
            const float phat = ch.target;
            if (phat > 0.0f) {
                const int perFrame = nCand > 1 ? nCand : 1;
                int frameCount = count / perFrame;
                if (frameCount < 1) frameCount = 1;
                W = wsum / (float(frameCount) * phat);
                if (!isfinite(W)) W = 0.0f;
            }
        }
        if (a.spatial_count > 0) {  // spatial reuse follows: it tests the final winner
            store_reservoir(&a.res_t[pixel], has, ch, wsum, count, W);
            return spec(0.0f);
        }
        if (has && !ch.vis_tested && ch.has_cand) {
            ch.tp = winner_contribution_lean(tab, v, ch, query, qo, qd, qtmax);
            ch.vis_tested = true;
        }
        store_reservoir(&a.res[pixel], has, ch, wsum, count, W);
        if (!has || W <= 0.0f) return spec(0.0f);
        return ch.tp * W;
    }
    double sum = 0.0;  // RIS::RandomIndexFromWeights (random_index.h:27-63)
    for (int q = 0; q < nCand; ++q) {
        const double w = wl[q * kLanes];
        if (w > 0.0) sum += w;
    }
    int chosen = -1;
    if (sum > 0.0) {
        double rem = double(rng.next_float()) * sum;
        for (int q = 0; q < nCand; ++q) {
            const double w = wl[q * kLanes];
            if (w <= 0.0) continue;
            rem -= w;
            if (rem <= 0.0) {
                chosen = q;
                break;
            }
        }
    }
    if (kDebug) dbg.shash = fnv(dbg.shash, uint32_t(chosen));
    if (chosen < 0) return spec(0.0f);
    rebuild_traced<kDebug, kGen>(a, tab, v, conn, connHit, connRec, connCand, start, esc, chosen, dbg, ch);
    const float W = float(sum / double(ch.target));
    Spec tp = ch.tp;
    if (!ch.vis_tested && ch.has_cand) tp = winner_contribution_lean(tab, v, ch, query, qo, qd, qtmax);
    return tp * W;
}

// ---------------------------------------------------------------- spatial reuse
// An extension beyond the reference (which reuses temporally only), run after every pixel's
// temporal pass. Pixel q resamples among its own post-temporal reservoir and those of up to
// spatial_count neighbours picked within spatial_radius pixels that pass a similarity test
// (depth within 10 %, geometric normals within 25 degrees). Reservoir j (confidence c_j = its
// candidate count, capped at 20 frames) offers its sample y_j shifted to q, weighted with the
// generalized balance heuristic
//     w_j = c_j W_j p_j(y_j) p_q(y_j) / sum_k c_k p_k(y_j -> k) |J_{q->k}|,
// whose weights sum to one over the domains, so the combination stays unbiased for any BSDF.
// The winner gets W = sum w / p_q(y) and a (deferred) shadow ray at q, and is stored for the next
// frame in the reservoir form of reservoir.h (WSum = W * frames * p_q). Neighbour picks and
// selections draw from a splitmix64 stream seeded by (pixel, frame): integer-only, so results are
// identical on every target, and the reference's per-pixel mt19937_64 stream is untouched.
// With reprojected temporal reuse on a frame whose camera moved, domain 1 (before the neighbours)
// is the previous frame's final reservoir at the pixel that saw q's hit (reproject_pixel), with
// that pixel's vertex of the previous frame; its slot in the debug record is kMaxSpatial.

constexpr int kMaxSpatial = 8;
// Domains of a pixel's spatial reuse (itself + up to spatial_count neighbours), rounded up to the
// register-array sizes the spatial kernels are compiled for.
constexpr int spatial_domains(int spatial_count) { return spatial_count <= 3 ? 4 : (spatial_count <= 5 ? 6 : 9); }

struct SpatialRng {
    uint64_t s;
    __device__ __forceinline__ SpatialRng(uint32_t pixel, int32_t frame)
        : s(uint64_t(pixel) * 0x9E3779B97F4A7C15ULL ^ uint64_t(uint32_t(frame)) * 0xC2B2AE3D27D4EB4FULL) {}
    __device__ __forceinline__ uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    __device__ __forceinline__ float next_float() { return float(next() >> 40) * 0x1p-24f; }
};

__device__ __forceinline__ bool pick_neighbour(SpatialRng &r, int x, int y, int W, int H, int R, int &nx, int &ny) {
    const uint64_t D = uint64_t(2 * R + 1);
    for (int t = 0; t < 8; ++t) {
        const uint64_t z = r.next();
        const int dx = int((z & 0xffffULL) % D) - R;
        const int dy = int(((z >> 16) & 0xffffULL) % D) - R;
        if ((dx == 0 && dy == 0) || dx * dx + dy * dy > R * R) continue;
        nx = x + dx;
        ny = y + dy;
        return nx >= 0 && ny >= 0 && nx < W && ny < H;
    }
    return false;
}

__device__ __forceinline__ bool spatially_similar(const HitRec &q, const HitRec &n) {
    if (fabsf(n.depth - q.depth) > 0.1f * q.depth) return false;
    return dot(q.ng, n.ng) >= 0.90630779f;  // cos 25 degrees
}

// A light sample of a reservoir, as seen from one shading point.
struct RSample {
    V3 dir, color, lnormal;
    float dist, pdf;
    bool solid, delta;
    int light;
};

__device__ __forceinline__ RSample rsample_of(const hdr_reservoir &r) {
    const uint32_t f = r.flags_light >> 24;
    RSample y;
    y.dir = v3(r.dir[0], r.dir[1], r.dir[2]);
    y.color = v3(r.color[0], r.color[1], r.color[2]);
    y.lnormal = v3(r.light_normal[0], r.light_normal[1], r.light_normal[2]);
    y.dist = r.dist;
    y.pdf = r.pdf;
    y.solid = (f & HDR_RES_PDF_SOLID_ANGLE) != 0;
    y.delta = (f & HDR_RES_LIGHT_DELTA) != 0;
    y.light = int(r.flags_light & 0xffffffu) - 1;
    return y;
}

// The sample seen from another shading point: distant and dome samples are directions; a point
// light is sampled again there (it draws nothing); an area sample keeps its point on the light.
__device__ __forceinline__ bool retarget(const FrameArgs &a, const RSample &y, V3 from, V3 to, RSample &out) {
    if (y.light < 0 || y.light >= a.s.num_lights) return false;
    const GLight &L = a.s.lights[y.light];
    out = y;
    if (L.type == HDR_LIGHT_DISTANT || L.type == HDR_LIGHT_DOME) return true;
    if (L.type == HDR_LIGHT_POINT) {
        LSample ls;
        Mt64Lean none;  // point lights draw nothing
        none.init(0u, nullptr);
        if (!sample_light_inl(a.s, L, to, none, ls)) return false;
        out.dir = ls.dir;
        out.color = ls.color;
        out.lnormal = ls.lnormal;
        out.dist = ls.dist;
        out.pdf = ls.pdf * a.light_select_pdf;
        out.solid = ls.solid_angle;
        return true;
    }
    const V3 P = from + y.dir * y.dist;
    const V3 toLight = P - to;
    const float dist = length(toLight);
    out.dir = divd(toLight, double(dist));
    out.dist = dist;
    return true;
}

// |d omega_to / d omega_from| of moving a sample: 1 for directions and point lights; for a point on
// an area light, (cos_to / d_to^2) / (cos_from / d_from^2).
__device__ __forceinline__ float shift_jacobian(const RSample &from, const RSample &to) {
    if (from.solid) return 1.0f;
    const float cosN = smax(0.0f, dot(-from.dir, from.lnormal));
    const float cosQ = smax(0.0f, dot(-to.dir, to.lnormal));
    if (cosN <= 0.0f) return 0.0f;
    return (cosQ / (to.dist * to.dist)) / (cosN / (from.dist * from.dist));
}

// The unshadowed contribution of a sample at a vertex (_reEvaluateForTemporalReuse, :529-545).
__device__ __forceinline__ Spec reeval_sample(const SpecTab &tab, const Vertex &v, const RSample &y) {
    const V3 wo = -v.rayDir;
    const float nDotL = dot(v.sn, y.dir);
    const float dist2 = y.dist * y.dist;
    const float cosY = smax(0.0f, dot(-y.dir, y.lnormal));
    const float pNee = to_solid_angle(y.pdf, y.solid, dist2, smax(cosY, 1e-6f));
    if (nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && !y.delta)) return spec(0.0f);
    const V3 bsdf = ggx_eval(v.c, v.sn, wo, y.dir);
    return rgb2spec(tab, y.color) * rgb2spec(tab, bsdf * nDotL);
}

// Domain d of a pixel's spatial reuse: pixel d of this frame (d < P), or for d = P + q the previous
// frame's records that reproject_kernel gathered for pixel q (reprojected temporal reuse).
__device__ __forceinline__ const GRec *domain_grec(const FrameArgs &a, const GRec *gbuf, int d, int P) {
    return d < P ? gbuf + d : a.rp_g + (d - P);
}
__device__ __forceinline__ const hdr_reservoir &domain_res(const FrameArgs &a, int d, int P) {
    return d < P ? a.res_t[d] : a.rp_res[d - P];
}
__device__ __forceinline__ float4 domain_aux(const FrameArgs &a, int d, int P) {
    return d < P ? a.spatial_aux[2 * d] : a.rp_aux[d - P];
}

// The shading vertex of a domain for the spatial pass, from its G-buffer record and the temporal
// pass's record {ray direction, wavelength u} (u < 0: the pixel was not shaded).
__device__ __forceinline__ bool load_domain_vertex(const FrameArgs &a, const GRec *gbuf, int d, int P, Vertex &v,
                                                   SpecTab &tab) {
    const float4 x0 = domain_aux(a, d, P);
    if (x0.w < 0.0f) return false;
    HitRec g;
    if (!load_grec(domain_grec(a, gbuf, d, P), g)) return false;
    tab = make_tab(x0.w);
    const Clo c = get_closure(a.s, g, a.enable_ss != 0);
    Spec unused = spec(1.0f);
    v = make_vertex(a, tab, g, c, v3(x0.x, x0.y, x0.z), unused);
    return true;
}
__device__ __forceinline__ bool load_spatial_vertex(const FrameArgs &a, const GRec *gbuf, int pixel, Vertex &v,
                                                    SpecTab &tab) {
    return load_domain_vertex(a, gbuf, pixel, a.W * a.H, v, tab);
}

// Reprojected temporal reuse (an extension): the pixel of the previous frame that saw hit g, or -1.
// The hit point goes through the previous frame's view and projection matrices (row vectors, in
// double; clip w > 0), its NDC [-1, 1] maps to [0, W) x [0, H) as camera_ray's inverse, and the
// pixel containing it qualifies if it was shaded then (its {ray, u} record has u >= 0) and its
// G-buffer record (the copy the previous spatial pass kept) is similar: depth within 10 % of the hit's distance from the previous camera,
// geometric normals within 25 degrees (spatially_similar's thresholds).
__device__ __forceinline__ int reproject_pixel(const FrameArgs &a, const HitRec &g) {
    const double X = double(g.pos.x), Y = double(g.pos.y), Z = double(g.pos.z);
    const double *V = a.prev_view, *Q = a.prev_proj;
    const double cx = X * V[0] + Y * V[4] + Z * V[8] + V[12];
    const double cy = X * V[1] + Y * V[5] + Z * V[9] + V[13];
    const double cz = X * V[2] + Y * V[6] + Z * V[10] + V[14];
    const double cw = X * V[3] + Y * V[7] + Z * V[11] + V[15];
    const double hx = cx * Q[0] + cy * Q[4] + cz * Q[8] + cw * Q[12];
    const double hy = cx * Q[1] + cy * Q[5] + cz * Q[9] + cw * Q[13];
    const double hw = cx * Q[3] + cy * Q[7] + cz * Q[11] + cw * Q[15];
    if (!(hw > 0.0)) return -1;
    const double px = (hx / hw + 1.0) * (0.5 * double(a.W));
    const double py = (hy / hw + 1.0) * (0.5 * double(a.H));
    if (!(px >= 0.0 && py >= 0.0 && px < double(a.W) && py < double(a.H))) return -1;
    const int j = int(py) * a.W + int(px);
    if (a.spatial_aux[2 * j].w < 0.0f) return -1;
    HitRec gj;
    if (!load_grec(a.hist_g + j, gj)) return -1;
    const float dist = length(g.pos - v3(a.prev_origin[0], a.prev_origin[1], a.prev_origin[2]));
    if (fabsf(gj.depth - dist) > 0.1f * dist) return -1;
    return dot(g.ng, gj.ng) >= 0.90630779f ? j : -1;  // cos 25 degrees
}

// Spatial reuse at pixel (x, y). dl / cl: this lane's domain pixels and confidences in LDS
// (stride kLanes). Returns the direct light at q; records the winner's shadow query. kND: the most
// domains the launch's pixels can have (1 + spatial_count, rounded up to 4, 6 or 9).
//
// Domain-major evaluation of the balance-heuristic weights: pass 1 moves every domain's sample to q
// (kept in registers), pass 2 visits each neighbour domain once -- its G-buffer record, closure and
// wavelength table -- and adds its term to every sample's denominator, pass 3 forms the weights and
// resamples. Each denominator is summed over the domains in the same order k = 0 .. nd-1 as the
// sample-by-sample formulation, so every weight, draw and selection is the same.
template <bool kDebug, int kND>
__device__ __forceinline__ Spec spatial_reuse_pixel(const FrameArgs &a, const GRec *gbuf, int x, int y,
                                                    const Vertex &vq, const SpecTab &tabq, int *dl, float *cl,
                                                    bool &query, V3 &qo, V3 &qd, float &qtmax, int &ssel) {
    const int pixel = y * a.W + x;
    const int P = a.W * a.H;
    const int maxK = kND - 1 - (a.reproject ? 1 : 0);
    const int K = a.spatial_count < maxK ? a.spatial_count : maxK;
    const int R = a.spatial_radius;
    const int M = a.n_cand > 1 ? a.n_cand : 1;
    const int cap = 20 * M;
    int nd = 0;
    dl[0] = pixel;
    ++nd;
    int slotBits = 0;  // slot of domain d (>= 1) in bits 4(d-1) .. 4(d-1)+3
    if (a.reproject && a.rp_aux[pixel].w >= 0.0f) {  // the previous frame's reservoir where q was seen
        dl[nd * kLanes] = P + pixel;
        slotBits |= kMaxSpatial << (4 * (nd - 1));
        ++nd;
    }
    SpatialRng srng(uint32_t(pixel), a.frame_index);
#pragma unroll 1
    for (int k = 0; k < K; ++k) {
        int nx, ny;
        if (!pick_neighbour(srng, x, y, a.W, a.H, R, nx, ny)) continue;
        const int n = ny * a.W + nx;
        if (a.spatial_aux[2 * n].w < 0.0f) continue;
        HitRec gn;
        if (!load_grec(gbuf + n, gn) || !spatially_similar(vq.h, gn)) continue;
        dl[nd * kLanes] = n;
        slotBits |= k << (4 * (nd - 1));
        ++nd;
    }
    int countSum = 0;
#pragma unroll 1
    for (int d = 0; d < nd; ++d) {
        const int c = domain_res(a, dl[d * kLanes], P).count;
        const int cc = c < cap ? c : cap;
        cl[d * kLanes] = float(cc);
        countSum += cc;
    }
    // pass 1: each domain's sample moved to q
    RSample yq[kND];
    float phJ[kND], jac[kND], phQ[kND], wJ[kND], den[kND];
    bool ok[kND];
    bool vis0 = false;
    const float cl0 = cl[0];
#pragma unroll
    for (int j = 0; j < kND; ++j) {
        ok[j] = false;
        phJ[j] = jac[j] = phQ[j] = wJ[j] = den[j] = 0.0f;
        if (j >= nd) continue;
        const int dj = dl[j * kLanes];
        const hdr_reservoir rj = domain_res(a, dj, P);
        const uint32_t fj = rj.flags_light >> 24;
        if (j == 0) vis0 = (fj & HDR_RES_VIS_TESTED) != 0;
        if (!(fj & HDR_RES_HAS_SAMPLE) || !(fj & HDR_RES_HAS_CANDIDATE) || rj.w <= 0.0f) continue;
        const float phatJ = rj.target;
        if (phatJ <= 0.0f) continue;
        const RSample atJ = rsample_of(rj);
        const float4 pj = domain_grec(a, gbuf, dj, P)->a;
        const V3 posJ = v3(pj.x, pj.y, pj.z);
        if (!retarget(a, atJ, posJ, vq.h.pos, yq[j])) continue;
        const float phatQ = spec_lum(tabq, reeval_sample(tabq, vq, yq[j]));
        if (phatQ <= 0.0f) continue;
        const float jacJQ = j == 0 ? 1.0f : shift_jacobian(atJ, yq[j]);
        if (jacJQ <= 0.0f) continue;
        ok[j] = true;
        phJ[j] = phatJ;
        jac[j] = jacJQ;
        phQ[j] = phatQ;
        wJ[j] = rj.w;
        den[j] = cl0 * (j == 0 ? phatJ / jacJQ : phatQ);  // the k = 0 term of its denominator
    }
    // pass 2: the neighbour domains' terms, domain by domain
#pragma unroll
    for (int k = 1; k < kND; ++k) {
        if (k >= nd) continue;
        bool any = false;
#pragma unroll
        for (int j = 0; j < kND; ++j) any = any || (ok[j] && j != k);
        Vertex vk;
        SpecTab tabk;
        const bool loaded = any && load_domain_vertex(a, gbuf, dl[k * kLanes], P, vk, tabk);
        const float ck = cl[k * kLanes];
#pragma unroll
        for (int j = 0; j < kND; ++j) {
            if (!ok[j]) continue;
            float pk;
            if (k == j) {
                pk = phJ[j] / jac[j];
            } else {
                pk = 0.0f;
                RSample atK;
                if (loaded && retarget(a, yq[j], vq.h.pos, vk.h.pos, atK))
                    pk = spec_lum(tabk, reeval_sample(tabk, vk, atK)) * shift_jacobian(yq[j], atK);
            }
            den[j] += ck * pk;
        }
    }
    // pass 3: weights and resampling, in domain order
    float wsum = 0.0f, bestPhQ = 0.0f;
    int chosen = -1;
    RSample best = yq[0];
#pragma unroll
    for (int j = 0; j < kND; ++j) {
        if (!ok[j] || !(den[j] > 0.0f)) continue;
        const float w = cl[j * kLanes] * wJ[j] * phJ[j] * phQ[j] / den[j];
        if (!(w > 0.0f)) continue;
        wsum += w;
        if (srng.next_float() <= w / wsum) {
            chosen = j;
            best = yq[j];
            bestPhQ = phQ[j];
        }
    }
    const bool has = chosen >= 0;
    Cand ch;
    ssel = -1;
    if (has) {
        ch.has_cand = true;
        ch.delta = best.delta;
        ch.solid = best.solid;
        ch.vis_tested = chosen == 0 ? vis0 : false;
        ch.light = best.light;
        ch.dir = best.dir;
        ch.color = best.color;
        ch.lnormal = best.lnormal;
        ch.dist = best.dist;
        ch.pdf = best.pdf;
        ch.tp = reeval_sample(tabq, vq, best);  // the same evaluation as in pass 1
        ch.target = bestPhQ;
        ssel = chosen == 0 ? -1 : ((slotBits >> (4 * (chosen - 1))) & 0xf);
    }
    const int count = countSum < cap ? countSum : cap;
    float W = 0.0f, wstore = 0.0f;
    if (has) {
        W = wsum / ch.target;
        if (!isfinite(W)) W = 0.0f;
        int frameCount = count / M;
        if (frameCount < 1) frameCount = 1;
        wstore = W * float(frameCount) * ch.target;
    }
    if (has && !ch.vis_tested && ch.has_cand) {
        ch.tp = winner_contribution_lean(tabq, vq, ch, query, qo, qd, qtmax);
        ch.vis_tested = true;
    }
    store_reservoir(&a.res[pixel], has, ch, wstore, count, W);
    if (!has || W <= 0.0f) return spec(0.0f);
    return ch.tp * W;
}

}  // namespace hdr
