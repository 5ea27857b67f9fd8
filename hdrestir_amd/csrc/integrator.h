// integrator.h -- the per-pixel estimator on the device: RaycastPass (G-buffer hit of the pixel
// centre), IntegrationPass (jittered ray, hero wavelengths, PathIntegrator::Li with MIS or RIS /
// ReSTIR direct lighting) and the spectral->RGB write-back, one thread per pixel.
#pragma once

#include "../../include/hdrgpu.h"
#include "shade.h"

namespace hdr {

struct GRec;
struct ShadowQuery;

// Frame-constant arguments of one view's frame. Kernels read them from a device array (one
// record per view of a launch), so one launch can cover the tiles of many views.
struct FrameArgs {
    DeviceScene s;
    double inv_view[16];
    double inv_proj[16];
    int W, H, frame_index;
    int has_cam;
    hdr_camera_params cam;
    float exposure_mul;
    int direct;          // HDR_DIRECT_MIS / HDR_DIRECT_RIS
    int use_res;
    int skip_vis;
    int max_depth;
    int max_refl, max_refr;
    int enable_ss;
    int ibl_bg;
    float roulette;
    int n_nee, n_sky, n_bsdf, n_cand;  // RIS technique split (ris_direct_light_integrator.cpp:93-100)
    int trace_generated;               // 0: generated BSDF candidates provably cannot contribute
    int trace_last_connection;         // 0: the last vertex's BSDF connection ray provably cannot contribute
    int gen_inline;                    // wavefront with mesh lights: a generated candidate's closest hit is
                                       //   traced in the shading launch (an escape ray cannot decide it)
    float light_select_pdf;            // 1 / number of lights
    int firefly;
    float inv_n;                       // 1 / (frame_index + 1)
    int frames;                        // frames of this launch (megakernel / lean path tracer; else 1)
    hdr_reservoir *res;                // persistent reservoirs (read by the temporal merge, written last)
    hdr_reservoir *res_t;              // spatial reuse: this frame's post-temporal reservoirs
    float4 *spatial_aux;               // spatial reuse: per pixel {ray dir, wavelength u}, {throughput}
    int spatial_count;                 // spatial reuse: neighbours per pixel (0 = off)
    int spatial_radius;                // spatial reuse: pick radius in pixels
    // reprojected temporal reuse (an extension): temporal_skip = the camera moved, so the
    // same-pixel temporal merge is skipped; reproject = the previous frame's records are usable,
    // so reproject_kernel gathers them and the spatial pass adds them as a domain
    int temporal_skip;
    int reproject;
    double prev_view[16];              // the previous frame's view and projection matrices
    double prev_proj[16];
    float prev_origin[3];              // the previous camera's position (camera_ray's origin)
    GRec *hist_g;                      // per pixel, the G-buffer record of the last spatial pass (kept by
                                       //   it for the next frame's gather; null: not kept)
    GRec *rp_g;                        // per pixel: the previous G-buffer record where its hit was seen,
    float4 *rp_aux;                    //   that pixel's {ray dir, wavelength u} (u < 0: none),
    hdr_reservoir *rp_res;             //   and its final reservoir of the previous frame
    float4 *accum;
    float4 *frame_rgb;                 // this frame's colour (firefly path / debugging)
    float4 *color;                     // accumulated colour
    double *lum_sum, *lum_sq;          // Variance metrics: per-pixel luminance sums (null when off)
    hdr_pixel_debug *dbg;
    float *aov_depth, *aov_albedo, *aov_normal;
    uint64_t *spill;                   // 2 x 312 words per pixel, or null
    unsigned long long *counters;      // hdr_frame_counters as 9 x u64, or null
    unsigned int *status;              // the view's status word: kStatusRngOverrun, ... (atomicOr)
    // per-view work buffers
    GRec *gbuf;                        // G-buffer, one record per pixel
    float4 *rgb2;                      // deferred path: visible / occluded colour per pixel
    uint32_t *flags;                   // deferred path: 0 no shadow ray, 1 visible, 3 occluded
    ShadowQuery *queue;                // deferred path: shadow queries (shared by the views of a batch)
    unsigned int *qcount;              // deferred path: number of queries (shared likewise)
    int view;                          // index of this record in the launch (carried by shadow queries)
    // wavefront path (paths of any depth, one launch per vertex): per-view path records and the
    // bounce rays' hits; queues shared by the views of a launch
    float4 *wstate;                    // kWaveFields float4 per pixel, field-major
    GRec *whit;                        // hit of each path's last bounce ray (traced frames: of odd vertices)
    GRec *whit2;                       // traced frames: hit of the vertex, even vertices >= 2
    float4 *wray;                      // bounce-ray queue: {origin, code}, {direction, -}
    float4 *wcq;                       // contribution of each shadow query (index-aligned with queue)
    float4 *wesq;                      // traced frames: escape-ray queue {origin, code}, {direction, k}
    float *wesq_t;                     // and the tmax of its entries whose k carries kEscTmax
    unsigned int *wesc;                // traced frames: per pixel, bit k = candidate k's ray escaped
    unsigned int *walive;              // codes of the live paths of even vertices (bounce ray hit)
    unsigned int *walive2;             // and of odd vertices
    unsigned int *wcnt;                // queue counts by vertex parity: rays[2], shadows[2], alive[2], escapes[2],
                                       // then the any-hit claim counters: escapes[2], shadows[2] (kWaveCounts)
};

// The stored reservoir the temporal merge reads (ris_direct_light_integrator.cpp:127-155): none on
// a frame whose camera moved under reprojected temporal reuse, where the pixel's stored sample
// belongs to another shading point (the spatial pass reuses the reprojected one instead).
__device__ __forceinline__ hdr_reservoir temporal_prev(const FrameArgs &a, int pixel) {
    if (a.temporal_skip) {
        hdr_reservoir none;
        none.flags_light = 0u;
        none.count = 0;
        none.wsum = 0.0f;
        none.target = 0.0f;
        return none;
    }
    return a.res[pixel];
}

struct PixelStats {
    uint32_t primary, bounce, shadow, cand, cand_skipped;
    TraceStats ts;
};

// Ray queries of the integrator. With kDeferred the shading kernel holds no traversal code: the
// only query it can meet (the RIS winner's shadow ray, when skipVisibility leaves the candidates
// untested and nothing BSDF-sampled can reach an emitter) is recorded and answered by the
// shadow kernel; closest-hit queries cannot occur on that path (the host checks eligibility).
template <bool kCount, bool kDeferred>
struct Tracer {
    const DeviceScene &s;
    PixelStats &st;
    bool has_query = false;
    V3 qo, qd;
    float qtmax = 0.0f;
    __device__ bool trace(V3 o, V3 d, HitRec &h) {
        if constexpr (kDeferred) {
            return false;
        } else {
            RawHit rh;
            if (!trace_closest_call<kCount>(s, o, d, rh, st.ts)) return false;
            h = make_hit_record_call(s, o, d, rh);
            return true;
        }
    }
    __device__ bool occluded(V3 o, V3 d, float tmax) {
        if constexpr (kDeferred) {
            has_query = true;
            qo = o;
            qd = d;
            qtmax = tmax;
            return false;
        } else {
            return trace_any_call<kCount>(s, o, d, tmax, st.ts);
        }
    }
};

// GenerateCameraRay (source/scene/camera_ray.cpp:10-76)
template <class R>
__device__ __forceinline__ void camera_ray(const FrameArgs &a, float px, float py, R *rng, V3 &o, V3 &d) {
    float ndcX = (2.0f * px / float(a.W)) - 1.0f;
    float ndcY = (2.0f * py / float(a.H)) - 1.0f;
    if (a.has_cam && a.cam.lens_distortion != 0.0f) {
        const float r2 = ndcX * ndcX + ndcY * ndcY;
        const float f = 1.0f + a.cam.lens_distortion * r2;
        ndcX *= f;
        ndcY *= f;
    }
    const V3D od = xform_point_d(a.inv_view, V3D{0.0, 0.0, 0.0});
    const V3 origin = v3(float(od.x), float(od.y), float(od.z));
    const V3D nc = xform_point_d(a.inv_proj, V3D{double(ndcX), double(ndcY), -1.0});
    const V3 nearCam = v3(float(nc.x), float(nc.y), float(nc.z));
    if (!a.has_cam || !a.cam.enable_dof) {
        const V3D nw = xform_point_d(a.inv_view, V3D{double(nearCam.x), double(nearCam.y), double(nearCam.z)});
        const V3 nearWorld = v3(float(nw.x), float(nw.y), float(nw.z));
        o = origin;
        d = normalized(nearWorld - origin);
        return;
    }
    const float aperture = (a.cam.focal_length / 10.0f) / (2.0f * a.cam.f_stop);
    float lensU, lensV;
    if (a.cam.bokeh_blades < 3) {
        const float r = sqrtf(rng->next_float());
        const float theta = float(2.0 * kPiD * double(rng->next_float()));
        lensU = r * cos_f(theta);
        lensV = r * sin_f(theta);
    } else {
        const float theta = float(2.0 * kPiD * double(rng->next_float()));
        const float r = sqrtf(rng->next_float());
        const float sectorAngle = float(2.0 * kPiD / double(float(a.cam.bokeh_blades)));
        const float sector = floorf(theta / sectorAngle);
        const float angleIn = theta - sector * sectorAngle;
        const float dd = cos_f(sectorAngle / 2.0f) / cos_f(sectorAngle / 2.0f - angleIn);
        lensU = r * dd * cos_f(theta);
        lensV = r * dd * sin_f(theta);
    }
    lensU *= aperture;
    lensV *= aperture;
    const V3 focalCam = nearCam * a.cam.focus_distance;
    const V3D lw = xform_point_d(a.inv_view, V3D{double(lensU), double(lensV), 0.0});
    const V3D fw = xform_point_d(a.inv_view, V3D{double(focalCam.x), double(focalCam.y), double(focalCam.z)});
    const V3 lensWorld = v3(float(lw.x), float(lw.y), float(lw.z));
    const V3 focalWorld = v3(float(fw.x), float(fw.y), float(fw.z));
    o = lensWorld;
    d = normalized(focalWorld - lensWorld);
}

__device__ __forceinline__ uint32_t fnv(uint32_t h, uint32_t v) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 16777619u;
    }
    return h;
}

struct Trace {  // parity instrumentation
    uint32_t lhash, shash;
    int picks;
};

// A shading vertex: hit record, closure, shading normal and the incoming direction.
struct Vertex {
    HitRec h;
    Clo c;
    V3 sn;
    V3 rayDir;
    bool inside;
};

// A BSDF-sampled connection (BsdfBounceConnection): the bounce and the hit of its ray.
struct Conn {
    Bounce b;
    bool has_hit;
    HitRec hit;
};

// One RIS candidate as reservoir sample (RISLightCandidate).
struct Cand {
    bool has_cand;
    bool delta;
    bool solid;
    bool vis_tested;
    int light;
    V3 dir, color, lnormal;
    float dist, pdf;
    Spec tp;
    float target;
};

template <bool kCountT, bool kDebugT, bool kDeferredT = false>
struct PixelCtx {
    static constexpr bool kCount = kCountT;
    static constexpr bool kDebug = kDebugT;
    static constexpr bool kDeferred = kDeferredT;
    const FrameArgs &a;
    const SpecTab &tab;
    Mt64 &rng;
    Tracer<kCountT, kDeferredT> &tr;
    PixelStats &st;
    Trace &dbg;
    double *wlds;      // this lane's candidate weights: wlds[k * kLanes]
    uint64_t *slot2;   // RNG fork spill slot
};

// Bits of a view's device status word. kStatusRngOverrun: a pixel drew past the RNG outputs its
// stream was set up for (the lean stream past output 310, a stream without a spill slot past
// 311 or 623); such draws return 0 instead of the engine's value, and the host reports the view as
// failed (HDR_ERR_CAPACITY) at its next synchronisation.
// AccumulationPass (accumulation_pass.cpp:64-106) for one pixel's frame sample c (after the firefly
// clamp): accum += c, colour = accum / (frame_index + 1); with the Variance metrics on
// (METRICS_ENABLED, :95-98) the sample's luminance -- in float, as GetLuminance -- and its square
// are added to the pixel's sums in double.
__device__ __forceinline__ void accumulate_sample(float4 *accum, float4 *color, double *lum_sum, double *lum_sq,
                                                  int pixel, float4 c, float inv_n) {
    if (lum_sum) {
        const double l = double(0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z);
        lum_sum[pixel] += l;
        lum_sq[pixel] += l * l;
    }
    float4 acc = accum[pixel];
    acc.x += c.x;
    acc.y += c.y;
    acc.z += c.z;
    acc.w += c.w;
    accum[pixel] = acc;
    color[pixel] = make_float4(acc.x * inv_n, acc.y * inv_n, acc.z * inv_n, acc.w * inv_n);
}

constexpr unsigned int kStatusRngOverrun = 1u;
// kStatusStackOverflow: an any-hit traversal with its stack in LDS would have pushed past the
// stack's kStk entries. The host admits that variant only for scenes whose tree depths fit it, so
// this does not happen; if it did, the entry is dropped (no write outside the stack) and the view
// reports HDR_ERR_CAPACITY at its next synchronisation.
constexpr unsigned int kStatusStackOverflow = 2u;
__device__ __forceinline__ void flag_rng_overrun(const FrameArgs &a, bool overrun) {
    if (overrun && a.status) atomicOr(a.status, kStatusRngOverrun);
}

constexpr int kWaveCounts = 12;  // words of FrameArgs::wcnt
// Flag in an escape-queue entry's k: the ray ends at wesq_t[entry] (an NEE candidate's shadow ray)
// instead of 1e30 (a generated candidate's escape ray).
constexpr int kEscTmax = 0x100;
constexpr int kLanes = 64;  // threads per block; candidate weights live at wlds[k * kLanes + lane]

// Evaluation of a light sample at a vertex (RisDirectLightIntegrator::_evaluateLightSample,
// ris_direct_light_integrator.cpp:482-527). occl: in BUILD mode, the recorded shadow result.
struct NeeEval {
    Spec contribution;
    float pNee, pBsdf;
    bool useMis;
    bool vis_tested;
    bool occluded;
};

template <class Ctx>
__device__ NeeEval eval_light_sample(Ctx &p, const Vertex &v, bool delta, const LSample &ls,
                                     bool skipVis, int occlMode /*0 trace, 1 occluded, 2 visible*/) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    NeeEval r;
    r.contribution = spec(0.0f);
    r.vis_tested = true;
    r.occluded = false;
    const V3 wo = -v.rayDir;
    const float nDotL = dot(v.sn, ls.dir);
    const float dist2 = ls.dist * ls.dist;
    const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
    r.pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
    r.useMis = !delta;
    if (nDotL <= 0.0f || r.pNee <= 0.0f || (cosY <= 0.0f && !delta)) {
        r.pBsdf = 0.0f;  // the candidate's contribution (and so its RIS weight) is zero
        return r;
    }
    r.pBsdf = ggx_pdf(v.c, v.sn, wo, ls.dir);
    const V3 bsdf = ggx_eval(v.c, v.sn, wo, ls.dir);
    const Spec tmul = rgb2spec(p.tab, bsdf * nDotL);
    if (skipVis) {
        r.contribution = rgb2spec(p.tab, ls.color) * tmul;
        r.vis_tested = false;
        return r;
    }
    bool occ;
    if (occlMode == 0) {
        const V3 so = v.h.pos + v.sn * 1e-4f;
        if (kCount) ++p.st.shadow;
        occ = p.tr.occluded(so, ls.dir, ls.dist - 1e-3f);
    } else {
        occ = occlMode == 1;
    }
    r.occluded = occ;
    if (!occ) r.contribution = rgb2spec(p.tab, ls.color) * tmul;
    return r;
}

// Light pick of UniformLightSampler::ProposeCandidate (uniform_light_sampler.cpp:22-47).
template <class Ctx>
__device__ __forceinline__ int pick_light(Ctx &p, bool record) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const uint32_t n = uint32_t(p.a.s.num_lights);
    const float u = p.rng.next_float();
    uint32_t idx = uint32_t(u * float(n));
    if (idx > n - 1u) idx = n - 1u;
    if (kDebug && record) {
        p.dbg.lhash = fnv(p.dbg.lhash, idx);
        ++p.dbg.picks;
    }
    return int(idx);
}

template <class Ctx>
__device__ __forceinline__ int pick_light_with(Ctx &p, Mt64 &rng, bool record) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const uint32_t n = uint32_t(p.a.s.num_lights);
    const float u = rng.next_float();
    uint32_t idx = uint32_t(u * float(n));
    if (idx > n - 1u) idx = n - 1u;
    if (kDebug && record) {
        p.dbg.lhash = fnv(p.dbg.lhash, idx);
        ++p.dbg.picks;
    }
    return int(idx);
}

// Emission of the material at a hit (GetClosure(hit).Emission): zero for the default material.
__device__ __forceinline__ V3 hit_emission(const DeviceScene &s, const HitRec &h) {
    const GMaterial &m = s.materials[s.instances[h.inst].material];
    return mv3(m.emission);
}

// Copy of the RNG for the candidate replay. Past output 312 the second-twist state lives in
// the spill slot; the copy gets its own slot so the main stream can move on.
__device__ __forceinline__ Mt64 fork_rng(const Mt64 &r, uint64_t *slot2) {
    Mt64 c = r;
    if (r.k > 312u && slot2) {
        for (int q = 0; q < kMtN; ++q) slot2[q] = r.slot[q];
    }
    c.slot = slot2;
    return c;
}

// RisDirectLightIntegrator::_evaluateBSDFConnection (ris_direct_light_integrator.cpp:547-634) for a
// connection with BsdfPdf > 0 and no impossible flag: emission seen along the ray and the NEE pdf
// of reaching it.
struct ConnEval {
    Spec emission;
    float pNee;
    bool has_light;
    int light;
};
template <class Ctx>
__device__ ConnEval eval_connection(Ctx &p, const Conn &cn) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const DeviceScene &s = p.a.s;
    ConnEval r;
    r.emission = spec(0.0f);
    r.pNee = 0.0f;
    r.has_light = false;
    r.light = -1;
    if (!cn.has_hit) {
        if (s.environment < 0) return r;
        const float sel = p.a.light_select_pdf;
        const float envPdf = env_pdf(s, cn.b.d);
        r.pNee = sel > 0.0f ? sel * envPdf : (p.a.n_sky > 0 ? envPdf : 0.0f);
        r.emission = rgb2spec(p.tab, env_radiance(s, cn.b.d));
        r.light = s.environment;
        return r;
    }
    r.emission = rgb2spec(p.tab, hit_emission(s, cn.hit));
    const int hl = s.instances[cn.hit.inst].light;
    if (hl >= 0) {
        r.has_light = true;
        r.light = hl;
        const float sel = p.a.light_select_pdf;
        const V3 off = cn.hit.pos - cn.b.o;
        const float dist2 = dot(off, off);
        const float cosY = smax(0.0f, dot(-cn.b.d, cn.hit.ng));
        if (cosY > 0.0f && sel > 0.0f) {
            bool solid;
            const float ap = light_eval_pdf(s, s.lights[hl], cn.b.d, solid);
            r.pNee = sel * to_solid_angle(ap, solid, dist2, cosY);
        }
    }
    return r;
}

// RIS weight of the path's own BSDF connection (first BSDF candidate, :263-294).
template <class Ctx>
__device__ double conn_candidate(Ctx &p, const Conn &cn, Cand *out) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const ConnEval e = eval_connection(p, cn);
    const Spec contribution = e.emission * cn.b.imul;
    const float target = spec_lum(p.tab, contribution);
    float pNeeSampler = e.pNee, pNeeSky = 0.0f;
    if (!cn.has_hit && (p.a.s.environment < 0 || p.a.light_select_pdf <= 0.0f)) {
        pNeeSampler = 0.0f;
        pNeeSky = e.pNee;
    }
    const double mis = double(power3(cn.b.pdf, p.a.n_bsdf, pNeeSampler, p.a.n_nee, pNeeSky, p.a.n_sky));
    if (out) {
        out->has_cand = false;
        out->vis_tested = true;
        out->light = -1;
        out->tp = contribution;
        out->target = target;
    }
    return mis * double(target) / (double(cn.b.pdf) + 1e-9);
}

// One generated BSDF candidate (:296-402). mode 0: weight, 1: draws only, 2: full candidate.
template <class Ctx>
__device__ double gen_bsdf_candidate(Ctx &p, const Vertex &v, Mt64 &rng, int mode, Cand *out) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const DeviceScene &s = p.a.s;
    if (mode == 1 || !p.a.trace_generated) {
        // Only the draws matter: either a replay up to the winner, or a scene without environment
        // and mesh lights, where the candidate is std::nullopt whatever it samples (:341-368).
        skip_bounce_draws(p.tab, v.c, v.sn, v.inside, v.rayDir, rng);
        if (kCount && mode == 0) ++p.st.cand_skipped;
        return 0.0;
    }
    Conn cn;
    int refl = 0, refr = 0;
    if (!sample_bounce(p.tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, 1, 1, refl, refr, rng, cn.b)) return 0.0;
    if (kCount) ++p.st.cand;
    cn.has_hit = p.tr.trace(cn.b.o, cn.b.d, cn.hit);
    if (!cn.has_hit) {
        if (s.environment < 0) return 0.0;
        const V3 wi = cn.b.d;
        const V3 envRGB = env_radiance(s, wi);
        const Spec envEm = rgb2spec(p.tab, envRGB);
        const float pBsdf = cn.b.pdf;
        const float sel = p.a.light_select_pdf;
        const float envPdf = env_pdf(s, wi);
        const float pNeeSampler = sel * envPdf;
        const float pNeeSky = (sel <= 0.0f && p.a.n_sky > 0) ? envPdf : 0.0f;
        const float pNee = pNeeSampler + pNeeSky;
        const Spec tp = envEm * cn.b.imul;
        const float target = spec_lum(p.tab, tp);
        const double mis = double(power3(pBsdf, p.a.n_bsdf, pNeeSampler, p.a.n_nee, pNeeSky, p.a.n_sky));
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
    const int hl = s.instances[cn.hit.inst].light;
    if (hl < 0) return 0.0;
    const ConnEval e = eval_connection(p, cn);
    const Spec contribution = e.emission * cn.b.imul;
    const float target = spec_lum(p.tab, contribution);
    const double mis = double(power3(cn.b.pdf, p.a.n_bsdf, e.pNee, p.a.n_nee, 0.0f, p.a.n_sky));
    if (out) {
        bool solid;
        float lp = light_eval_pdf(s, s.lights[hl], cn.b.d, solid);
        lp *= p.a.light_select_pdf;
        out->has_cand = true;
        out->delta = s.lights[hl].is_delta != 0;
        out->solid = solid;
        out->vis_tested = true;
        out->light = hl;
        out->dir = cn.b.d;
        out->color = hit_emission(s, cn.hit);
        out->lnormal = cn.hit.ng;
        out->dist = cn.hit.depth;
        out->pdf = lp;
        out->tp = contribution;
        out->target = target;
    }
    return mis * double(target) / (double(cn.b.pdf) + 1e-9);
}

// One NEE candidate (:407-443). mode 0: weight (+ occlusion bit), 1: draws only, 2: full candidate.
template <class Ctx>
__device__ double nee_candidate(Ctx &p, const Vertex &v, Mt64 &rng, int mode, bool occlBit,
                                bool &occludedOut, Cand *out) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const DeviceScene &s = p.a.s;
    occludedOut = false;
    if (s.num_lights <= 0) return 0.0;
    const int li = pick_light_with(p, rng, mode == 0);
    const GLight &L = s.lights[li];
    if (mode == 1) {
        const int nd = light_draws(s, L);
        for (int q = 0; q < nd; ++q) rng.discard();
        return 0.0;
    }
    LSample ls;
    if (!sample_light(s, L, v.h.pos, rng, ls)) return 0.0;
    ls.pdf = ls.pdf * p.a.light_select_pdf;
    const bool delta = L.is_delta != 0;
    const NeeEval e = eval_light_sample(p, v, delta, ls, p.a.skip_vis != 0, mode == 0 ? 0 : (occlBit ? 1 : 2));
    occludedOut = e.occluded;
    const float target = spec_lum(p.tab, e.contribution);
    const float mis = e.useMis ? power3(e.pNee, p.a.n_nee, 0.0f, p.a.n_sky, e.pBsdf, p.a.n_bsdf) : 1.0f;
    if (out) {
        out->has_cand = true;
        out->delta = delta;
        out->solid = ls.solid_angle;
        out->vis_tested = e.vis_tested;
        out->light = li;
        out->dir = ls.dir;
        out->color = ls.color;
        out->lnormal = ls.lnormal;
        out->dist = ls.dist;
        out->pdf = ls.pdf;
        out->tp = e.contribution;
        out->target = target;
    }
    return double(mis * target) / (double(e.pNee) + 1e-9);
}

// Rebuilds candidate `kk` by replaying the generation from `start` (a copy of the RNG taken before
// candidate 0). occl: per-candidate shadow results recorded during generation.
template <class Ctx>
__device__ void rebuild_candidate(Ctx &p, const Vertex &v, const Conn *conn, bool connCand,
                                  Mt64 rng, uint64_t occl, int kk, Cand &out) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const int nNee = p.a.n_nee;
    bool dummy;
    for (int k = 0; k < nNee; ++k) {
        if (k < kk) {
            nee_candidate(p, v, rng, 1, false, dummy, nullptr);
        } else {
            nee_candidate(p, v, rng, 2, ((occl >> k) & 1ULL) != 0, dummy, &out);
            return;
        }
    }
    int k = nNee + p.a.n_sky;
    if (connCand) {
        if (k == kk) {
            conn_candidate(p, *conn, &out);
            return;
        }
        ++k;
    }
    for (; k < p.a.n_cand; ++k) {
        if (k < kk) {
            gen_bsdf_candidate(p, v, rng, 1, nullptr);
        } else {
            gen_bsdf_candidate(p, v, rng, 2, &out);
            return;
        }
    }
}

__device__ __forceinline__ void store_reservoir(hdr_reservoir *dst, bool has, const Cand &c, float wsum, int count,
                                                float w) {
    hdr_reservoir r;
    r.dir[0] = c.dir.x;
    r.dir[1] = c.dir.y;
    r.dir[2] = c.dir.z;
    r.dist = c.dist;
    r.color[0] = c.color.x;
    r.color[1] = c.color.y;
    r.color[2] = c.color.z;
    r.pdf = c.pdf;
    r.light_normal[0] = c.lnormal.x;
    r.light_normal[1] = c.lnormal.y;
    r.light_normal[2] = c.lnormal.z;
    r.target = c.target;
    for (int q = 0; q < 4; ++q) r.throughput[q] = c.tp.v[q];
    r.wsum = wsum;
    r.count = count;
    r.w = w;
    uint32_t flags = 0, light = 0;
    if (has) {
        flags |= HDR_RES_HAS_SAMPLE;
        if (c.vis_tested) flags |= HDR_RES_VIS_TESTED;
        if (c.has_cand) {
            flags |= HDR_RES_HAS_CANDIDATE;
            if (c.solid) flags |= HDR_RES_PDF_SOLID_ANGLE;
            if (c.delta) flags |= HDR_RES_LIGHT_DELTA;
            light = uint32_t(c.light + 1);
        }
    } else {
        r.target = 0.0f;
        for (int q = 0; q < 4; ++q) r.throughput[q] = 0.0f;
        r.dir[0] = r.dir[1] = r.dir[2] = 0.0f;
        r.dist = 0.0f;
        r.color[0] = r.color[1] = r.color[2] = 0.0f;
        r.pdf = 0.0f;
        r.light_normal[0] = r.light_normal[1] = r.light_normal[2] = 0.0f;
    }
    if (has && !c.has_cand) {
        r.dir[0] = r.dir[1] = r.dir[2] = 0.0f;
        r.dist = 0.0f;
        r.color[0] = r.color[1] = r.color[2] = 0.0f;
        r.pdf = 0.0f;
        r.light_normal[0] = r.light_normal[1] = r.light_normal[2] = 0.0f;
    }
    r.flags_light = (flags << 24) | (light & 0xffffffu);
    *dst = r;
}

// RisDirectLightIntegrator::Li (ris_direct_light_integrator.cpp:84-215).
template <class Ctx>
__device__ Spec ris_direct(Ctx &p, const Vertex &v, const Conn *conn, bool primary, int pixel) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const FrameArgs &a = p.a;
    const int nNee = a.n_nee;
    const Mt64 start = fork_rng(p.rng, p.slot2);
    uint64_t occl = 0;
    // candidate generation: NEE, then BSDF (the path connection first when it has a pdf)
    for (int k = 0; k < nNee; ++k) {
        bool occ;
        const double w = nee_candidate(p, v, p.rng, 0, false, occ, nullptr);
        if (occ) occl |= (1ULL << k);
        p.wlds[k * kLanes] = w;
    }
    int k = nNee + a.n_sky;
    const bool connCand = conn && conn->b.pdf > 0.0f;
    if (connCand) {
        p.wlds[k * kLanes] = conn_candidate(p, *conn, nullptr);
        ++k;
    }
    for (; k < a.n_cand; ++k) p.wlds[k * kLanes] = gen_bsdf_candidate(p, v, p.rng, 0, nullptr);

    if (a.use_res && primary) {
        // streaming WRS over the candidates (WeightedReservoir::Update, reservoir.h:76-84)
        float wsum = 0.0f;
        int count = 0;
        int sel = -1;
        for (int q = 0; q < a.n_cand; ++q) {
            ++count;
            const float w = float(p.wlds[q * kLanes]);
            if (w <= 0.0f) continue;
            wsum += w;
            if (p.rng.next_float() <= w / wsum) sel = q;
        }
        // temporal merge with the stored reservoir of this pixel (:127-155)
        hdr_reservoir prev = temporal_prev(a, pixel);
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
                    reeval = rgb2spec(p.tab, prevC.color) * rgb2spec(p.tab, bsdf * nDotL);
                }
            }
            const float phatCurr = spec_lum(p.tab, reeval);
            const float phatGen = prev.target;
            if (phatGen > 0.0f) {
                prevC.target = phatCurr;
                prevC.tp = reeval;
                const float mergeWeight = prev.wsum * (phatCurr / phatGen);
                count += prev.count;
                if (mergeWeight > 0.0f) {
                    wsum += mergeWeight;
                    if (p.rng.next_float() <= mergeWeight / wsum) sel = -2;
                }
            }
        }
        if (kDebug) p.dbg.shash = fnv(p.dbg.shash, uint32_t(sel));
        Cand ch;
        bool has = false;
        if (sel >= 0) {
            rebuild_candidate(p, v, conn, connCand, start, occl, sel, ch);
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
        // winner visibility fix-up (:159-175)
        if (has && !ch.vis_tested && ch.has_cand) {
            LSample ls;
            ls.dir = ch.dir;
            ls.color = ch.color;
            ls.lnormal = ch.lnormal;
            ls.dist = ch.dist;
            ls.pdf = ch.pdf;
            ls.solid_angle = ch.solid;
            const NeeEval e = eval_light_sample(p, v, ch.delta, ls, false, 0);
            ch.tp = e.contribution;
            ch.vis_tested = true;
        }
        store_reservoir(&a.res[pixel], has, ch, wsum, count, W);
        if (!has || W <= 0.0f) return spec(0.0f);
        return ch.tp * W;
    }

    // RIS::RandomIndexFromWeights (random_index.h:27-63)
    double sum = 0.0;
    for (int q = 0; q < a.n_cand; ++q) {
        const double w = p.wlds[q * kLanes];
        if (w > 0.0) sum += w;
    }
    int chosen = -1;
    if (sum > 0.0) {
        double rem = double(p.rng.next_float()) * sum;
        for (int q = 0; q < a.n_cand; ++q) {
            const double w = p.wlds[q * kLanes];
            if (w <= 0.0) continue;
            rem -= w;
            if (rem <= 0.0) {
                chosen = q;
                break;
            }
        }
    }
    if (kDebug) p.dbg.shash = fnv(p.dbg.shash, uint32_t(chosen));
    if (chosen < 0) return spec(0.0f);
    Cand ch;
    rebuild_candidate(p, v, conn, connCand, start, occl, chosen, ch);
    const float W = float(sum / double(ch.target));
    Spec tp = ch.tp;
    if (!ch.vis_tested && ch.has_cand) {
        LSample ls;
        ls.dir = ch.dir;
        ls.color = ch.color;
        ls.lnormal = ch.lnormal;
        ls.dist = ch.dist;
        ls.pdf = ch.pdf;
        ls.solid_angle = ch.solid;
        tp = eval_light_sample(p, v, ch.delta, ls, false, 0).contribution;
    }
    return tp * W;
}

// MisDirectLightIntegrator::Li (mis_direct_light_integrator.cpp:70-271).
template <class Ctx>
__device__ Spec mis_direct(Ctx &p, const Vertex &v, const Conn *conn) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const FrameArgs &a = p.a;
    const DeviceScene &s = a.s;
    const bool useBsdf = conn != nullptr;
    Spec total = spec(0.0f);
    if (s.num_lights > 0) {  // _integrateNEE with the light sampler
        const int li = pick_light_with(p, p.rng, true);
        const GLight &L = s.lights[li];
        LSample ls;
        if (sample_light(s, L, v.h.pos, p.rng, ls)) {
            ls.pdf = ls.pdf * a.light_select_pdf;
            const bool delta = L.is_delta != 0;
            const float nDotL = dot(v.sn, ls.dir);
            const float dist2 = ls.dist * ls.dist;
            const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
            if (nDotL > 0.0f && !(cosY <= 0.0f && !delta)) {
                const float pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
                if (pNee > 0.0f) {
                    if (kCount) ++p.st.shadow;
                    const bool occ = p.tr.occluded(v.h.pos + v.sn * 1e-4f, ls.dir, ls.dist - 1e-3f);
                    if (!occ) {
                        const V3 wo = -v.rayDir;
                        const float bsdfPdf = ggx_pdf(v.c, v.sn, wo, ls.dir);
                        const V3 bsdf = ggx_eval(v.c, v.sn, wo, ls.dir);
                        const Spec radiance =
                            rgb2spec(p.tab, bsdf) * rgb2spec(p.tab, ls.color) * (nDotL / pNee);
                        const float w = (delta || !useBsdf) ? 1.0f : power2(pNee, bsdfPdf);
                        total = total + radiance * w;
                    }
                }
            }
        }
    }
    if (conn) {  // _integrateBSDFConnection, :181-255
        const Conn &cn = *conn;
        Spec part = spec(0.0f);
        if (cn.b.impossible || cn.b.pdf <= 0.0f) {
            if (!cn.has_hit) {
                if (s.environment >= 0) part = rgb2spec(p.tab, env_radiance(s, cn.b.d)) * cn.b.tmul;
                else part = spec(0.0f) * cn.b.tmul;
            } else {
                part = rgb2spec(p.tab, hit_emission(s, cn.hit)) * cn.b.tmul;
            }
        } else {
            float pNeeSampler = 0.0f, pNeeEnv = 0.0f;
            Spec radiance = spec(0.0f);
            bool early = false;
            if (!cn.has_hit) {
                if (s.environment < 0) {
                    early = true;
                } else {
                    radiance = rgb2spec(p.tab, env_radiance(s, cn.b.d));
                    const float envSel = a.light_select_pdf;
                    const float envPdf = env_pdf(s, cn.b.d);
                    pNeeSampler = envSel * envPdf;
                    pNeeEnv = envSel <= 0.0f ? envPdf : 0.0f;
                }
            } else {
                radiance = rgb2spec(p.tab, hit_emission(s, cn.hit));
                const int hl = s.instances[cn.hit.inst].light;
                if (hl >= 0) {
                    const float sel = a.light_select_pdf;
                    if (sel > 0.0f) {
                        const V3 off = cn.hit.pos - cn.b.o;
                        const float dist2 = dot(off, off);
                        const float cosY = smax(0.0f, dot(-cn.b.d, cn.hit.ng));
                        if (cosY > 0.0f) {
                            bool solid;
                            const float ap = light_eval_pdf(s, s.lights[hl], cn.b.d, solid);
                            pNeeSampler = sel * to_solid_angle(ap, solid, dist2, cosY);
                        }
                    }
                }
            }
            if (!early) {
                if (pNeeSampler + pNeeEnv <= 0.0f) {
                    part = radiance * cn.b.tmul;
                } else {
                    const float w = power3(cn.b.pdf, 1, pNeeSampler, 1, pNeeEnv, 1);
                    part = radiance * w * cn.b.tmul;
                }
            }
        }
        total = total + part;
    }
    return total;
}

template <class Ctx>
__device__ __forceinline__ Spec direct_light(Ctx &p, const Vertex &v, const Conn *conn,
                                             bool primary, int pixel) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    if (p.a.direct == HDR_DIRECT_RIS) return ris_direct(p, v, conn, primary, pixel);
    return mis_direct(p, v, conn);
}

// Samples the bounce at a vertex and traces it (Detail::SampleBounceWithConnection,
// direct_light_integrator_interface.h:18-40). When the ray provably cannot change the result
// (last vertex, no emissive material, no environment) it is not traced and reads as a miss.
template <class Ctx>
__device__ bool bounce_with_connection(Ctx &p, const Vertex &v, int &refl, int &refr,
                                       bool last, Conn &cn) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    if (!sample_bounce(p.tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, p.a.max_refl, p.a.max_refr, refl, refr, p.rng,
                       cn.b))
        return false;
    if (last && !p.a.trace_last_connection) {
        cn.has_hit = false;
        if (kCount) ++p.st.cand_skipped;
        return true;
    }
    if (kCount) ++p.st.bounce;
    cn.has_hit = p.tr.trace(cn.b.o, cn.b.d, cn.hit);
    return true;
}

__device__ __forceinline__ Vertex make_vertex(const FrameArgs &a, const SpecTab &tab, const HitRec &h, const Clo &c,
                                              V3 rayDir, Spec &throughput) {
    Vertex v;
    v.h = h;
    v.c = c;
    v.rayDir = rayDir;
    v.inside = dot(c.n, rayDir) > 0.0f;
    beer(tab, throughput, *c.m, h.depth, v.inside);
    v.sn = v.inside ? -c.n : c.n;
    return v;
}

// PathIntegrator::Li / _li (source/integrators/path_integrator.cpp:48-209).
template <class Ctx>
__device__ Spec path_li(Ctx &p, V3 rayDir, bool hasHit, HitRec hit, int pixel) {
    constexpr bool kCount = Ctx::kCount;
    constexpr bool kDebug = Ctx::kDebug;
    (void)kDebug;
    const FrameArgs &a = p.a;
    const DeviceScene &s = a.s;
    Spec throughput = spec(1.0f);
    while (hasHit) {
        const Clo c = get_closure(s, hit, a.enable_ss != 0);
        if (c.m->opacity >= 0.999f || p.rng.next_float() <= c.m->opacity) {
            const Vertex v0 = make_vertex(a, p.tab, hit, c, rayDir, throughput);
            // _li
            Spec total = rgb2spec(p.tab, mv3(c.m->emission));
            Spec tp = spec(1.0f);
            int refl = 0, refr = 0;
            Conn first;
            const bool ok = bounce_with_connection(p, v0, refl, refr, a.max_depth <= 1, first);
            total = total + direct_light(p, v0, ok ? &first : nullptr, true, pixel);
            if (ok) {
                tp = tp * first.b.tmul;
                V3 curO = first.b.o, curD = first.b.d;
                bool hasNext = first.has_hit;
                HitRec next = first.hit;
                for (int bounce = 1; bounce < a.max_depth; ++bounce) {
                    if (!hasNext) break;
                    const HitRec bh = next;
                    const Clo bc = get_closure(s, bh, a.enable_ss != 0);
                    if (bc.m->opacity < 0.999f && p.rng.next_float() > bc.m->opacity) {
                        curO = bh.pos + curD * 1e-4f;
                        if (kCount) ++p.st.bounce;
                        hasNext = p.tr.trace(curO, curD, next);
                        --bounce;
                        continue;
                    }
                    const Vertex v = make_vertex(a, p.tab, bh, bc, curD, tp);
                    Conn cn;
                    const bool bok = bounce_with_connection(p, v, refl, refr, bounce == a.max_depth - 1, cn);
                    total = total + tp * direct_light(p, v, bok ? &cn : nullptr, false, pixel);
                    if (!bok) break;
                    tp = tp * cn.b.tmul;
                    curO = cn.b.o;
                    curD = cn.b.d;
                    hasNext = cn.has_hit;
                    next = cn.hit;
                    if (cn.b.skip_roulette) continue;
                    if (bounce > 3) {
                        const float pr = sclamp(spec_max(tp) * a.roulette, 0.0f, 1.0f);
                        if (p.rng.next_float() > pr) break;
                        tp = tp * (1.0f / pr);
                    }
                }
            }
            return throughput * total;
        }
        const V3 o = hit.pos + rayDir * 1e-4f;
        if (kCount) ++p.st.bounce;
        hasHit = p.tr.trace(o, rayDir, hit);
    }
    if (s.environment >= 0 && a.ibl_bg) return rgb2spec(p.tab, env_radiance(s, rayDir));
    return spec(0.0f);
}

}  // namespace hdr
