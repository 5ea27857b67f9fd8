// wavefront.h -- PathIntegrator::Li (path_integrator.cpp:48-209) with RIS / ReSTIR direct lighting
// (ris_direct_light_integrator.cpp:84-215), or MIS direct lighting (the path tracer,
// mis_direct_light_integrator.cpp:70-271), at every vertex, for paths of any depth, run as a
// wavefront: one shading launch per path vertex and one traversal launch between them, with the
// path's state in HBM. A shading launch samples the vertex's bounce, evaluates its direct light and
// records two ray queries: the bounce ray (closest hit, the next vertex) and the direct light's
// shadow ray (any hit: the RIS winner's or the MIS light sample's). The traversal launch answers
// both queues and compacts the paths that go on into the next vertex's work list, so later
// (sparser) vertices launch only over live paths.
//
// Spatial reuse (an extension, deferred.h: spatial_reuse_pixel): the primary vertex stores its
// post-temporal reservoir and the spatial launch after wave_shade0 resamples it with the
// neighbours'; its winner's shadow ray joins vertex 0's shadow queue and only sets the pixel's
// occlusion flag. The path total then holds the later vertices from zero, and the pixel's
// radiance is outer * ((emission + reused direct) + total).
//
// Frames the host admits here have opaque materials (no pass-through rays). Untraced frames: RIS
// with skipVisibility, or the path tracer, where nothing BSDF-sampled can reach an emitter, so the
// BSDF candidates reduce to their RNG draws. Traced frames (wave_gen below): an environment,
// emissive materials or mesh lights, or RIS without skipVisibility; the vertex's ray queries are
// emitted and answered before it is shaded. Streams may run past RNG output 310: the second-twist
// machinery (spill slots) stays out of line.
//
// Results are identical to the megakernel's: every vertex consumes the same draws in the same order,
// and the radiance sum is formed in the same order -- vertex b's term `tp * d` is added to the path
// total by the vertex's shading launch, or by the traversal launch right after it when d waits for
// its shadow ray; the next vertex adds its term only after that launch.
#pragma once

#include <type_traits>

#include "pathlean.h"

namespace hdr {

// The deep stream (Mt64Deep) is in mt64.h.

// The winner's replay copy: once the stream has materialised its spill slot (output 625 on), the
// copy works on the spare slot.
__device__ __forceinline__ Mt64Deep rng_fork(const Mt64Deep &r, uint64_t *slot2) {
    Mt64Deep c = r;
    if ((r.k > 624u || r.ready) && slot2 && r.slot) mt_copy_slot(slot2, r.slot);  // a prepared slot goes along
    c.slot = slot2;
    return c;
}

// ---------------------------------------------------------------- path record
// Field-major in HBM (field f of pixel i at wstate[f * P + i]), 9 x 16 B per pixel, so each field
// of a wave's paths is one coalesced 1 KB access.
constexpr int kWaveFields = 9;
constexpr uint32_t kWavePrimaryHit = 1u << 31;  // in field 7.x: the G-buffer ray hit
constexpr uint32_t kWaveSlotReady = 1u << 30;   // in field 7.x: the spill slot holds the third twist

// R: Mt64Deep, or Mt64Lean on the primary vertex when the host has checked that no pixel can draw
// past output 310 there (the two share Mt64's cursor state, so the record is the same).
template <class R = Mt64Deep>
struct WavePath {
    R rng;
    Spec tp, total, outer;
    V3 dir;      // direction of the ray that reached the current vertex
    float utab;  // hero-wavelength draw
    int refl, refr;
    bool primary_hit;
    Trace dbg;
};

__device__ __forceinline__ float4 pack_u64x2(uint64_t a, uint64_t b) {
    return make_float4(__uint_as_float(uint32_t(a)), __uint_as_float(uint32_t(a >> 32)),
                       __uint_as_float(uint32_t(b)), __uint_as_float(uint32_t(b >> 32)));
}
__device__ __forceinline__ uint64_t unpack_lo(float4 v) {
    return uint64_t(__float_as_uint(v.x)) | (uint64_t(__float_as_uint(v.y)) << 32);
}
__device__ __forceinline__ uint64_t unpack_hi(float4 v) {
    return uint64_t(__float_as_uint(v.z)) | (uint64_t(__float_as_uint(v.w)) << 32);
}
__device__ __forceinline__ float4 spec4(const Spec &s) { return make_float4(s.v[0], s.v[1], s.v[2], s.v[3]); }
__device__ __forceinline__ Spec spec_of(float4 v) { return Spec{{v.x, v.y, v.z, v.w}}; }

__device__ __forceinline__ size_t wave_pixels(const FrameArgs &a) { return size_t(a.W) * size_t(a.H); }
__device__ __forceinline__ uint64_t *wave_slot(const FrameArgs &a, int pixel) {
    return a.spill ? a.spill + size_t(pixel) * (2 * kMtN) : nullptr;
}

// outer: only at the primary vertex (later vertices leave field 6 alone).
template <class R>
__device__ __forceinline__ void store_path(const FrameArgs &a, int pixel, const WavePath<R> &p, bool outer) {
    float4 *w = a.wstate;
    const size_t P = wave_pixels(a);
    w[pixel] = pack_u64x2(p.rng.lo, p.rng.lo1);
    w[P + pixel] = pack_u64x2(p.rng.hi, p.rng.hi1);
    w[2 * P + pixel] = make_float4(__uint_as_float(uint32_t(p.rng.s156)), __uint_as_float(uint32_t(p.rng.s156 >> 32)),
                                   __uint_as_float(p.rng.k), __uint_as_float(p.rng.j));
    w[3 * P + pixel] = spec4(p.tp);
    w[4 * P + pixel] = spec4(p.total);
    w[5 * P + pixel] = make_float4(p.dir.x, p.dir.y, p.dir.z, p.utab);
    if (outer) w[6 * P + pixel] = spec4(p.outer);
    const uint32_t bits = (uint32_t(p.refl) & 0x7fffu) | ((uint32_t(p.refr) & 0x7fffu) << 15) |
                          (p.primary_hit ? kWavePrimaryHit : 0u) | (p.rng.ready ? kWaveSlotReady : 0u);
    w[7 * P + pixel] = make_float4(__uint_as_float(bits), __uint_as_float(p.dbg.lhash), __uint_as_float(p.dbg.shash),
                                   __int_as_float(p.dbg.picks));
    if constexpr (std::is_same<R, Mt64Deep>::value)
        if (p.rng.k > 467u) w[8 * P + pixel] = pack_u64x2(p.rng.b155, p.rng.b311);  // set once pass A ends
}

// outer: also the outer throughput (field 6), for a launch that shades the primary vertex.
__device__ __forceinline__ void load_path(const FrameArgs &a, int pixel, WavePath<Mt64Deep> &p, bool outer = false) {
    const float4 *w = a.wstate;
    const size_t P = wave_pixels(a);
    if (outer) p.outer = spec_of(w[6 * P + pixel]);
    const float4 r0 = w[pixel], r1 = w[P + pixel], r2 = w[2 * P + pixel];
    p.rng.lo = unpack_lo(r0);
    p.rng.lo1 = unpack_hi(r0);
    p.rng.hi = unpack_lo(r1);
    p.rng.hi1 = unpack_hi(r1);
    p.rng.s156 = unpack_lo(r2);
    p.rng.k = __float_as_uint(r2.z);
    p.rng.j = __float_as_uint(r2.w);
    p.rng.seed = uint32_t(pixel) ^ uint32_t((unsigned long long)(long long)a.frame_index * 12345ULL);
    p.rng.slot = wave_slot(a, pixel);
    p.tp = spec_of(w[3 * P + pixel]);
    p.total = spec_of(w[4 * P + pixel]);
    const float4 r5 = w[5 * P + pixel];
    p.dir = v3(r5.x, r5.y, r5.z);
    p.utab = r5.w;
    const float4 r7 = w[7 * P + pixel];
    const uint32_t bits = __float_as_uint(r7.x);
    p.refl = int(bits & 0x7fffu);
    p.refr = int((bits >> 15) & 0x7fffu);
    p.primary_hit = (bits & kWavePrimaryHit) != 0;
    p.rng.ready = (bits & kWaveSlotReady) != 0;
    p.dbg.lhash = __float_as_uint(r7.y);
    p.dbg.shash = __float_as_uint(r7.z);
    p.dbg.picks = __float_as_int(r7.w);
    if (p.rng.k > 467u) {
        const float4 r8 = w[8 * P + pixel];
        p.rng.b155 = unpack_lo(r8);
        p.rng.b311 = unpack_hi(r8);
    }
}

// Traced frames: hit records of vertex v >= 1, double-buffered by parity (the shading launch of
// vertex v reads vertex v's and vertex v+1's records; the traversal launch after it writes vertex
// v+2's into the first buffer). Vertex 0's record is the G-buffer's.
__device__ __forceinline__ GRec *traced_hits(const FrameArgs &a, int v) { return (v & 1) ? a.whit : a.whit2; }

// ---------------------------------------------------------------- one path vertex
// The body of one iteration of _li's bounce loop (path_integrator.cpp:112-177) at vertex `bounce`
// (0 = the primary hit, whose throughput is the outer one and whose direct term is not scaled by
// tp): closure, bounce sample (its ray is traced by the next launch, as the path's BSDF
// connection), RIS direct lighting, path throughput and Russian roulette. Outputs the vertex's
// shadow query with its contribution `tp * d` (added to the total if the ray is unoccluded) and the
// bounce ray when the path goes on.
// kTraced: frames whose BSDF-sampled candidates can see an emitter (environment, emissive material;
// no mesh light): the vertex's ray queries were emitted by wave_gen and answered by a traversal
// launch -- the bounce ray's hit (the connection candidate's hit and the next vertex) is read from
// the hit buffer of vertex bounce + 1, the generated candidates' escape tests from the pixel's
// escape mask -- and the next vertex's hit is returned in `next` (ray = a next vertex exists)
// instead of a bounce ray.
template <bool kCount, bool kDebug, bool kTraced, bool kGen, class R>
__device__ __forceinline__ void wave_vertex(const FrameArgs &a, const SpecTab &tab, int bounce, const HitRec &h,
                                            WavePath<R> &p, double *wl, int pixel, PixelStats &st, bool &query,
                                            ShadowQuery &q, float4 &cq, bool &ray, float4 &ro, float4 &rd,
                                            HitRec &next) {
    const DeviceScene &s = a.s;
    const bool first = bounce == 0;
    const Clo c = get_closure(s, h, a.enable_ss != 0);
    Vertex v;
    if (first) v = make_vertex(a, tab, h, c, p.dir, p.outer);
    else v = make_vertex(a, tab, h, c, p.dir, p.tp);
    // with spatial reuse the primary vertex's emission and direct light join in the spatial pass;
    // the total then holds the later vertices only
    if (first) p.total = a.spatial_count > 0 ? spec(0.0f) : rgb2spec(tab, mv3(c.m->emission));
    Bounce b;
    const bool ok = sample_bounce_inl(tab, v.c, v.sn, v.inside, v.h.pos, v.rayDir, a.max_refl, a.max_refr, p.refl,
                                      p.refr, p.rng, b);
    const bool last = bounce == a.max_depth - 1;
    bool cHit = false;
    if (kTraced) {  // SampleBounceWithConnection (direct_light_integrator_interface.h:18-40)
        if (ok) {
            if (last && !a.trace_last_connection) {
                if (kCount) ++st.cand_skipped;
            } else {  // traced (and counted) by the traversal launch
                cHit = load_grec(traced_hits(a, bounce + 1) + pixel, next);
            }
        }
    } else if (kCount && ok && last) {
        ++st.cand_skipped;  // the last connection ray is never traced here
    }
    V3 qo = v3s(0.0f), qd = v3s(0.0f);
    float qtmax = 0.0f;
    bool qq = false;
    uint64_t *slot = wave_slot(a, pixel);
    Spec d;
    if (kTraced)
        d = ris_direct_traced<kCount, kDebug, kGen>(a, tab, v, ok ? &b : nullptr, cHit, next, p.rng, wl, pixel, st, p.dbg,
                                              first && a.use_res, slot ? slot + kMtN : nullptr,
                                              (a.trace_generated || a.skip_vis == 0) ? a.wesc[pixel] : 0u, qq,
                                              qo, qd, qtmax);
    else if (a.direct == HDR_DIRECT_MIS)  // the path tracer (frames whose BSDF rays reach no emitter)
        d = mis_nee_wave(a, tab, v, p.rng, ok, p.dbg, qq, qo, qd, qtmax);
    else
        d = ris_direct_lean<kCount, kDebug>(a, tab, v, ok && b.pdf > 0.0f, p.rng, wl, pixel, st, p.dbg,
                                            first && a.use_res, slot ? slot + kMtN : nullptr, qq, qo, qd, qtmax);
    const Spec td = first ? d : p.tp * d;
    if (qq) {
        query = true;
        cq = spec4(td);
        q.o = make_float4(qo.x, qo.y, qo.z, qtmax);
        q.d = make_float4(qd.x, qd.y, qd.z, __uint_as_float(query_code(a, pixel)));
    } else {
        p.total = p.total + td;
    }
    ray = false;
    if (!ok) return;
    p.tp = p.tp * b.tmul;
    if (!b.skip_roulette && bounce > 3) {
        const float pr = sclamp(spec_max(p.tp) * a.roulette, 0.0f, 1.0f);
        if (p.rng.next_float() > pr) return;
        p.tp = p.tp * (1.0f / pr);
    }
    if (last) return;
    p.dir = b.d;
    if (kTraced) {
        ray = cHit;
        return;
    }
    ray = true;
    ro = make_float4(b.o.x, b.o.y, b.o.z, __uint_as_float(query_code(a, pixel)));
    rd = make_float4(b.d.x, b.d.y, b.d.z, 0.0f);
}

// ---------------------------------------------------------------- traced frames: ray queries
// Before wave_vertex runs vertex `bounce` of a traced frame, wave_gen draws what it will draw up to
// its generated candidates -- the bounce sample and the NEE candidates' draws -- on a copy of the
// path's stream (the record keeps the position before the vertex) and emits the rays whose
// answers the vertex needs: the bounce ray (closest hit; not at the last vertex unless the last
// connection is traced), without skipVisibility each NEE candidate's shadow ray (through
// wave_emit_visibility: wave_gen then leaves r before the NEE candidates) and, through
// wave_emit_escapes, each sampled generated candidate's escape ray (any hit up to 1e30). r and vx
// are left at the first generated candidate k0 (k0 = n_cand when nothing generated can see an
// emitter). Clears the pixel's escape mask.
template <class R>
__device__ __forceinline__ void wave_gen(const FrameArgs &a, const SpecTab &tab, int bounce, const HitRec &h,
                                         const WavePath<R> &p, int pixel, R &r, Vertex &vx, bool &bray, float4 &bo,
                                         float4 &bd, int &k0) {
    const DeviceScene &s = a.s;
    const Clo c = get_closure(s, h, a.enable_ss != 0);
    Spec thr = bounce == 0 ? p.outer : p.tp;
    vx = make_vertex(a, tab, h, c, p.dir, thr);
    int refl = p.refl, refr = p.refr;
    Bounce b;
    const bool ok = sample_bounce_inl(tab, vx.c, vx.sn, vx.inside, vx.h.pos, vx.rayDir, a.max_refl, a.max_refr, refl,
                                      refr, r, b);
    const bool last = bounce == a.max_depth - 1;
    bray = ok && !(last && !a.trace_last_connection);
    const uint32_t code = query_code(a, pixel);
    if (bray) {
        bo = make_float4(b.o.x, b.o.y, b.o.z, __uint_as_float(code));
        bd = make_float4(b.d.x, b.d.y, b.d.z, 0.0f);
    }
    k0 = a.n_cand;
    const bool vis = a.skip_vis == 0;  // the NEE candidates' shadow rays follow (wave_emit_visibility)
    if (!a.trace_generated && !vis) return;
    a.wesc[pixel] = 0u;
    if (!vis && s.num_lights > 0) {  // the NEE candidates' draws (ris_direct_traced's order)
        Trace none;
        for (int k = 0; k < a.n_nee; ++k) {
            const int li = pick_light_lean<false>(a, r, none, false);
            const int nd = light_draws(s, s.lights[li]);
            for (int q = 0; q < nd; ++q) r.discard();
        }
    }
    // with mesh lights the shading launch traces the generated candidates itself (gen_traced_weight)
    if (!a.trace_generated || a.gen_inline) return;
    k0 = a.n_nee + a.n_sky + ((ok && b.pdf > 0.0f) ? 1 : 0);  // after the connection candidate
}

// Wave-aggregated append: one atomic per wave, each flagged lane gets its slot. All 64 lanes of
// the wave must call it together.
__device__ __forceinline__ unsigned int wave_append(unsigned int *count, bool want, int lane) {
    const unsigned long long mask = __ballot(want);
    const int n = __popcll(mask);
    unsigned int base = 0;
    if (n > 0) {
        if (lane == 0) base = atomicAdd(count, unsigned(n));
        base = __shfl(base, 0);
    }
    return base + unsigned(__popcll(mask & ((1ULL << lane) - 1ULL)));
}

// Without skipVisibility: the shadow rays of this lane's n NEE candidates (wave_gen's state, r
// before candidate 0; n = 0 for a lane with none), appended to the escape queue {origin, code},
// {direction, k | kEscTmax} with their tmax in wesq_t. Each candidate makes the draws and takes the
// decisions of nee_candidate_lean -- light pick, light sample, then the ray only when the candidate
// can contribute (n . l > 0, a positive pdf, the light's front face) -- and its ray is
// _evaluateLightSample's: from pos + n_s 1e-4 toward the sample, ending 1e-3 short of it
// (ris_direct_light_integrator.cpp:482-527). Leaves r after the NEE candidates. All 64 lanes of the
// wave call it together; it returns the lane's ray count.
template <class R>
__device__ __forceinline__ unsigned int wave_emit_visibility(unsigned int *count, float4 *queue, float *tq,
                                                             const FrameArgs *a, const Vertex &vx, R &r, int n,
                                                             uint32_t code, int lane) {
    unsigned int emitted = 0;
    for (int k = 0;; ++k) {
        const bool live = k < n;
        if (__ballot(live) == 0ULL) break;
        bool want = false;
        V3 so = v3s(0.0f), dir = v3s(0.0f);
        float tmax = 0.0f;
        if (live) {
            const DeviceScene &s = a->s;
            Trace none;
            const int li = pick_light_lean<false>(*a, r, none, false);
            const GLight &L = s.lights[li];
            LSample ls;
            if (sample_light_inl(s, L, vx.h.pos, r, ls)) {
                ls.pdf = ls.pdf * a->light_select_pdf;
                const float nDotL = dot(vx.sn, ls.dir);
                const float dist2 = ls.dist * ls.dist;
                const float cosY = smax(0.0f, dot(-ls.dir, ls.lnormal));
                const float pNee = to_solid_angle(ls.pdf, ls.solid_angle, dist2, smax(cosY, 1e-6f));
                if (!(nDotL <= 0.0f || pNee <= 0.0f || (cosY <= 0.0f && L.is_delta == 0))) {
                    want = true;
                    so = vx.h.pos + vx.sn * 1e-4f;
                    dir = ls.dir;
                    tmax = ls.dist - 1e-3f;
                }
            }
        }
        const unsigned int qi = wave_append(count, want, lane);
        if (want) {
            queue[2 * qi] = make_float4(so.x, so.y, so.z, __uint_as_float(code));
            queue[2 * qi + 1] = make_float4(dir.x, dir.y, dir.z, __int_as_float(k | kEscTmax));
            tq[qi] = tmax;
            ++emitted;
        }
    }
    return emitted;
}

// Escape rays of this lane's generated candidates k0 .. n_cand-1 (wave_gen's state), appended to
// the escape queue {origin, code}, {direction, k}. All 64 lanes of the wave call it together
// (a lane with nothing to emit passes k0 >= n); it loops until every lane is done.
template <class R>
__device__ __forceinline__ void wave_emit_escapes(unsigned int *count, float4 *queue, const SpecTab &tab,
                                                  const Vertex &vx, R &r, int k0, int n, uint32_t code, int lane) {
    int k = k0;
    for (;;) {
        const bool live = k < n;
        if (__ballot(live) == 0ULL) break;
        bool want = false;
        Bounce b;
        const int kc = k;
        if (live) {
            int refl = 0, refr = 0;
            want = sample_bounce_inl(tab, vx.c, vx.sn, vx.inside, vx.h.pos, vx.rayDir, 1, 1, refl, refr, r, b);
            ++k;
        }
        const unsigned int qi = wave_append(count, want, lane);
        if (want) {
            queue[2 * qi] = make_float4(b.o.x, b.o.y, b.o.z, __uint_as_float(code));
            queue[2 * qi + 1] = make_float4(b.d.x, b.d.y, b.d.z, __int_as_float(kc));
        }
    }
}

}  // namespace hdr
