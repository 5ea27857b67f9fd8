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

// UniformLightSampler::ProposeCandidate light pick (uniform_light_sampler.cpp:22-47).
template <bool kDebug>
__device__ __forceinline__ int pick_light_lean(const FrameArgs &a, Mt64Lean &rng, Trace &dbg, bool record) {
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

// One light-sampler NEE candidate evaluated under skipVisibility (_generateNEECandidates :407-443
// with _evaluateLightSample :482-527): returns the RIS weight, fills `out` when given.
__device__ __forceinline__ double nee_candidate_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                     Mt64Lean &rng, int li, Cand *out) {
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
        contribution = rgb2spec(tab, ls.color) * tmul;
        visTested = false;
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
// candidate `kk`; earlier candidates only consume their draws.
template <bool kDebug>
__device__ __forceinline__ void rebuild_nee_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                 Mt64Lean rng, int kk, Trace &dbg, Cand &out) {
    const DeviceScene &s = a.s;
    for (int k = 0; k < kk; ++k) {
        const int li = pick_light_lean<kDebug>(a, rng, dbg, false);
        const int nd = light_draws(s, s.lights[li]);
        for (int q = 0; q < nd; ++q) rng.discard();
    }
    const int li = pick_light_lean<kDebug>(a, rng, dbg, false);
    (void)nee_candidate_lean(a, tab, v, rng, li, &out);
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

// RisDirectLightIntegrator::Li at the primary vertex on the deferred path. wl: this lane's NEE
// candidate weights in LDS (wl[k * kLanes]).
template <bool kCount, bool kDebug>
__device__ __forceinline__ Spec ris_direct_lean(const FrameArgs &a, const SpecTab &tab, const Vertex &v,
                                                bool connCand, Mt64Lean &rng, double *wl, int pixel,
                                                PixelStats &st, Trace &dbg, bool &query, V3 &qo, V3 &qd,
                                                float &qtmax) {
    const DeviceScene &s = a.s;
    const int nNee = a.n_nee;
    const Mt64Lean start = rng;
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

    if (a.use_res) {
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
        const hdr_reservoir prev = a.res[pixel];
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

}  // namespace hdr
