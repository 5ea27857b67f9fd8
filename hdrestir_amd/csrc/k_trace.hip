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

// Any-hit rays of a queue on persistent lanes. With one ray per lane, a wave runs until its longest
// traversal ends (the counters show about a sixth of the lanes active per instruction on shadow
// rays). Here a lane whose ray is decided takes the next unclaimed ray of the queue -- rays are
// claimed in groups through *cursor, zero at launch -- so lanes stay busy until the queue is
// drained. Lanes refill between leaves once at least kRefill of them are idle, or all are.
// query(i, view, o, d, tmax) decodes ray i and the view (args index) it belongs to; done(i, view,
// hit) records its verdict. kOneScene: every view of the launch renders args[0]'s scene, whose tables
// are then wave-uniform. Each ray's traversal is trace_any_s's, step for step, so the verdicts are
// the same.
// kStk = 0: the TLAS and BLAS stacks are private arrays of kStackDepth entries (scratch memory).
// kStk > 0: one stack of kStk entries per lane in LDS -- the lane's column of the wave's array `lds`
// (kStk * kLanes ints), entry k at lds[k * kLanes + lane] -- holding the TLAS entries with the current
// BLAS's above them (bsp is then the index where the BLAS's entries start). A tree whose deepest leaf
// is at depth D (root at 1) puts at most D - 1 entries on the stack, one per level below its root, so
// the host takes this variant only for scenes whose TLAS and BLAS depths add up to at most kStk + 2.
template <bool kCount, int kRefill, bool kOneScene, int kStk, class Query, class Done>
__device__ __forceinline__ void anyhit_queue(const FrameArgs *__restrict__ args, unsigned int n,
                                             unsigned int *cursor, TraceStats &st, int *lds, Query query, Done done) {
    const int lane = threadIdx.x;
    const FrameArgs &a0 = args[0];
    int tref[kStk > 0 ? 1 : kStackDepth], bref[kStk > 0 ? 1 : kStackDepth];
    int *const col = lds + lane;
    int tsp = 0, bsp = 0;  // kStk > 0: tsp counts every entry of the lane's stack
    RayBox R;
    V3 o = v3s(0.0f), d = v3s(0.0f), oo = o, od = d;
    float tmax = 0.0f;
    int cur = kNoNode, lnext = 0, lend = 0;
    bool blas = false, world = true;
    unsigned int ray = 0, view = 0;  // the lane's ray (queue index) and view
    bool live = false, drained = false;  // drained: every ray is claimed (wave-uniform)
    for (;;) {
        if (!drained) {
            const unsigned long long idle = __ballot(!live);
            const int nidle = __popcll(idle);
            if (nidle >= kRefill || nidle == kLanes) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned int base = 0;
                if (lane == leader) base = atomicAdd(cursor, unsigned(nidle));
                base = __shfl(base, leader);
                if (base + unsigned(nidle) >= n) drained = true;
                if (!live) {
                    const unsigned int i = base + unsigned(__popcll(idle & ((1ULL << lane) - 1ULL)));
                    if (i < n) {
                        ray = i;
                        query(i, view, o, d, tmax);
                        const DeviceScene &s = kOneScene ? a0.s : args[view].s;
                        bool miss = true;
                        if (s.tlas_root >= 0) {  // the root test of trace_any_s
                            R = make_raybox<true>(o, d);
                            const GNode nr = load_node(s.nodes, s.tlas_root);
                            if (kCount) ++st.nodes;
                            float ta;
                            if (box_hit<true>(R, nr, ta) && !(ta > tmax)) {
                                cur = node_ref(nr);
                                tsp = bsp = 0;
                                blas = false;
                                world = true;
                                lnext = lend = 0;
                                oo = o;
                                od = d;
                                live = true;
                                miss = false;
                            }
                        }
                        if (miss) done(i, view, false);
                    }
                }
            }
        }
        if (__ballot(live) == 0ULL) {
            if (drained) break;
            continue;
        }
        if (!live) continue;
        const DeviceScene &s = kOneScene ? a0.s : args[view].s;
        // trace_any_s's iterations, grouped as descent to a leaf (the inner-node loop), the leaf, then
        // the pop -- one group per pass, so the wave reaches the refill between leaves
        while (cur >= 0) {
            const GNode l = load_node(s.nodes, cur), r = load_node(s.nodes, cur + 1);
            if (kCount) st.nodes += 2;
            float tl, tr;
            const bool hl = box_hit<true>(R, l, tl) && !(tl > tmax);
            const bool hr = box_hit<true>(R, r, tr) && !(tr > tmax);
            if (!(hl || hr)) {
                cur = kNoNode;
                break;
            }
            const bool leftFirst = hl && (!hr || !(tr < tl));
            const int first = leftFirst ? node_ref(l) : node_ref(r);
            if (hl && hr) {
                const int second = leftFirst ? node_ref(r) : node_ref(l);
                if constexpr (kStk > 0) {
                    if (tsp < kStk) {
                        col[kLanes * tsp++] = second;
                    } else {  // beyond the depth bound the host checked: dropped, and reported
                        const FrameArgs &av = args[view];
                        if (av.status) atomicOr(av.status, kStatusStackOverflow);
                    }
                } else if (blas) bref[bsp++] = second;
                else tref[tsp++] = second;
            }
            cur = first;
        }
        if (cur != kNoNode) {
            const int code = -cur - 1;
            const int start = code >> 3, count = code & 7;
            bool hit = false;
            if (blas) {
                for (int k = 0; k < count; ++k) {
                    const GTri tri = load_tri(s.tris, start + k);
                    if (kCount) ++st.tris;
                    float tt, uu, vv;
                    if (tri_hit(tri, oo, od, tt, uu, vv) && tt < tmax) {
                        hit = true;
                        break;
                    }
                }
            } else {
                lnext = start;
                lend = start + count;
            }
            if (hit) {
                done(ray, view, true);
                live = false;
                continue;
            }
        }
        cur = kNoNode;
        if (blas) {
            if constexpr (kStk > 0) {
                if (tsp > bsp) {
                    cur = col[kLanes * --tsp];
                    continue;
                }
            } else if (bsp > 0) {
                cur = bref[--bsp];
                continue;
            }
            blas = false;
        }
        while (lnext < lend) {
            float minv[16];
            const int root = load_inst_inv(s.instances, gload_i(s.tlas_idx + lnext++), minv);
            if (kCount) ++st.instances;
            if (root < 0) continue;
            oo = xform_point(minv, o);
            od = xform_dir(minv, d);
            R = make_raybox<false>(oo, od);
            world = false;
            bsp = kStk > 0 ? tsp : 0;
            const GNode b = load_node(s.nodes, root);
            if (kCount) ++st.nodes;
            float ta;
            if (box_hit<true>(R, b, ta) && !(ta > tmax)) {
                cur = node_ref(b);
                blas = true;
                break;
            }
        }
        if (blas) continue;
        if (!world) {
            R = make_raybox<true>(o, d);
            world = true;
        }
        if (tsp == 0) {
            done(ray, view, false);  // no occluder
            live = false;
            continue;
        }
        if constexpr (kStk > 0) cur = col[kLanes * --tsp];
        else cur = tref[--tsp];
    }
}

// The deferred path's shadow rays on persistent lanes (anyhit_queue over the launch's query queue;
// qcount[1] is its claim counter): an occluded winner keeps zero throughput in its reservoir (the
// visibility fix-up of ris_direct_light_integrator.cpp:163-175).
template <bool kCount, int kOcc, int kRefill, bool kOneScene, int kStk>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
shadow_persistent_kernel(const FrameArgs *__restrict__ args) {
    const FrameArgs &a0 = args[0];
    __shared__ int stk[kStk > 0 ? kStk * kLanes : 1];
    TraceStats st = {};
    anyhit_queue<kCount, kRefill, kOneScene, kStk>(
        args, a0.qcount[0], a0.qcount + 1, st, stk,
        [&](unsigned int i, unsigned int &view, V3 &o, V3 &d, float &tmax) {
            const ShadowQuery q = a0.queue[i];
            view = __float_as_uint(q.d.w) >> 24;
            o = v3(q.o.x, q.o.y, q.o.z);
            d = v3(q.d.x, q.d.y, q.d.z);
            tmax = q.o.w;
        },
        [&](unsigned int i, unsigned int view, bool hit) {
            if (!hit) return;
            const FrameArgs &a = args[view];
            const int pixel = int(__float_as_uint(a0.queue[i].d.w) & 0xffffffu);
            a.flags[pixel] = 3u;
            hdr_reservoir &rv = a.res[pixel];
            rv.throughput[0] = rv.throughput[1] = rv.throughput[2] = rv.throughput[3] = 0.0f;
        });
    if (kCount) {
        unsigned long long *c = a0.counters;
        wave_add(c + 5, st.nodes);
        wave_add(c + 6, st.tris);
        wave_add(c + 7, st.instances);
    }
}

// Verdict of shadow ray i of the wavefront queue (vertex `vtx`): an unoccluded winner's contribution
// joins the path total (a spatial winner's colour is already staged); an occluded primary winner
// keeps zero throughput in its reservoir (ris_direct_light_integrator.cpp:163-175).
__device__ __forceinline__ void wave_shadow_done(const FrameArgs *__restrict__ args, unsigned int i, int vtx,
                                                 bool occluded) {
    const uint32_t code = __float_as_uint(args[0].queue[i].d.w);
    const FrameArgs &a = args[code >> 24];
    const int pixel = int(code & 0xffffffu);
    if (!occluded) {
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
// the shadow rays (any hit, one per lane: wave_shadow_done) of vertex `bounce`.
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
    for (unsigned int i = blockIdx.x * kLanes + unsigned(lane); i < ns; i += gridDim.x * kLanes) {
        const ShadowQuery q = args[0].queue[i];
        const FrameArgs &a = args[__float_as_uint(q.d.w) >> 24];
        wave_shadow_done(args, i, bounce,
                         any_hit<kCount, kLds>(a.s, v3(q.o.x, q.o.y, q.o.z), v3(q.d.x, q.d.y, q.d.z), q.o.w, ts,
                                               ring_tr + lane, ring_br + lane));
    }
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 1, nb);
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
    }
}

// Traced frames: traversal launch v (0 .. depth) after shading launch v -- vertex v's bounce rays
// (closest hit: vertex v+1's hit record, or a miss record); vertex v's escape rays and vertex v-1's
// shadow rays follow in wave_anyhit_kernel.
template <bool kCount, int kOcc = 1, bool kLds = false>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(kOcc)))
wave_trace_traced_kernel(const FrameArgs *__restrict__ args, int v) {
    HDR_RING_LDS
    const int lane = threadIdx.x;
    const int par = v & 1;
    unsigned int *cnt = args[0].wcnt;
    const unsigned int nr = cnt[par];
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
    if (kCount) {
        unsigned long long *c = args[0].counters;
        wave_add(c + 1, nb);
        wave_add(c + 5, ts.nodes);
        wave_add(c + 6, ts.tris);
        wave_add(c + 7, ts.instances);
    }
}

// The any-hit rays of traced wavefront traversal launch v, on persistent lanes: vertex v's escape
// rays (any hit up to 1e30: the candidate's bit of the pixel's escape mask when none) and vertex
// v-1's shadow rays. The queues' claim counters are wcnt[8 + parity] (escape rays) and
// wcnt[10 + parity] (shadow rays); this launch clears the other parity's for the next one (the
// frame's first pair is cleared with the counts). (Untraced frames trace their shadow rays in
// wave_trace_kernel, one per lane: persistent lanes measured slower on those shorter queues.)
template <bool kCount, bool kTraced, bool kOneScene, int kStk>
__global__ void __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(4)))
wave_anyhit_kernel(const FrameArgs *__restrict__ args, int v) {
    const FrameArgs &a0 = args[0];
    __shared__ int stk[kStk > 0 ? kStk * kLanes : 1];
    const int par = v & 1;
    unsigned int *cnt = a0.wcnt;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[8 + (par ^ 1)] = 0u;
        cnt[10 + (par ^ 1)] = 0u;
    }
    TraceStats st = {};
    if (kTraced) {
        anyhit_queue<kCount, 16, kOneScene, kStk>(
            args, cnt[6 + par], cnt + 8 + par, st, stk,
            [&](unsigned int i, unsigned int &view, V3 &o, V3 &d, float &tmax) {
                const float4 o4 = a0.wesq[2 * i], d4 = a0.wesq[2 * i + 1];
                view = __float_as_uint(o4.w) >> 24;
                o = v3(o4.x, o4.y, o4.z);
                d = v3(d4.x, d4.y, d4.z);
                tmax = (__float_as_int(d4.w) & kEscTmax) ? a0.wesq_t[i] : 1e30f;
            },
            [&](unsigned int i, unsigned int view, bool hit) {  // escaped / unoccluded: bit k
                if (hit) return;
                const uint32_t code = __float_as_uint(a0.wesq[2 * i].w);
                atomicOr(args[view].wesc + int(code & 0xffffffu), 1u << (__float_as_int(a0.wesq[2 * i + 1].w) & 31));
            });
    }
    const int vtx = kTraced ? v - 1 : v;
    const unsigned int ns = kTraced ? (v > 0 ? cnt[2 + (par ^ 1)] : 0u) : cnt[2 + par];
    anyhit_queue<kCount, 16, kOneScene, kStk>(
        args, ns, cnt + 10 + par, st, stk,
        [&](unsigned int i, unsigned int &view, V3 &o, V3 &d, float &tmax) {
            const ShadowQuery q = a0.queue[i];
            view = __float_as_uint(q.d.w) >> 24;
            o = v3(q.o.x, q.o.y, q.o.z);
            d = v3(q.d.x, q.d.y, q.d.z);
            tmax = q.o.w;
        },
        [&](unsigned int i, unsigned int, bool hit) { wave_shadow_done(args, i, vtx, hit); });
    if (kCount) {
        unsigned long long *c = a0.counters;
        wave_add(c + 5, st.nodes);
        wave_add(c + 6, st.tris);
        wave_add(c + 7, st.instances);
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
template __global__ void shadow_kernel<true, 4, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 6, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<true, 8, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 1, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 4, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 6, true>(const FrameArgs *__restrict__);
template __global__ void shadow_kernel<false, 8, true>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<true, 4, 16, true, 0>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<true, 4, 16, true, 32>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<true, 4, 16, false, 0>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<true, 4, 16, false, 32>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<false, 4, 16, true, 0>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<false, 4, 16, true, 32>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<false, 4, 16, false, 0>(const FrameArgs *__restrict__);
template __global__ void shadow_persistent_kernel<false, 4, 16, false, 32>(const FrameArgs *__restrict__);
template __global__ void wave_anyhit_kernel<true, true, true, 0>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<true, true, true, 32>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<true, true, false, 0>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<true, true, false, 32>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<false, true, true, 0>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<false, true, true, 32>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<false, true, false, 0>(const FrameArgs *__restrict__, int);
template __global__ void wave_anyhit_kernel<false, true, false, 32>(const FrameArgs *__restrict__, int);
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
