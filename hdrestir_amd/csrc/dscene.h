// dscene.h -- device-resident scene (HBM layout) and BVH traversal.
//
// Layout (all arrays in one device allocation, 256-byte aligned sub-ranges):
//   nodes      GNode[ ]   32 B: BLAS nodes of every mesh, then TLAS nodes. Internal node:
//                         left = index of the left child (right = left + 1); leaf:
//                         left = -(first triangle) - 1, count = triangles (<= 4) or instances (<= 2).
//                         Every tree starts at an odd index (a padding record precedes it when
//                         needed), so each sibling pair sits in one 64-byte aligned block.
//   tris       GTri[ ]    48 B: v0, v1, v2 (+ original triangle index, mesh-local) -- the
//                         36 bytes the Moller-Trumbore test reads, one 16-B aligned record.
//   attrs      GTriAttr[] 72 B: shading normals and vertex colours, read once per ray for the
//                         winning triangle only.
//   tlas_idx   int[ ]         TLAS leaf slots -> instance index.
//   tlas_rank  int[ ]         TLAS leaf slots -> the instance's slot in the reference's TLAS.
//   instances  GInstance[] 144 B: inverse / forward GfMatrix4f, BLAS root, material, light.
// The BLAS are the reference's midpoint-split BVH2 (leaf <= 4 triangles), built on the host in the
// reference's order (source/renderContext/bvh.cpp:184-241), so a triangle's record index ranks it
// as the reference's traversal would. The TLAS is a binned-SAH tree (one instance per leaf); the
// reference's midpoint TLAS (source/scene/scene.cpp:352-407) is built too, only to give each
// instance its rank, the slot it has there. Closest-hit ties are broken on (rank, record), which is
// the reference's winner whatever tree is traversed (trace_closest_s).
#pragma once

#include <climits>

#include "gfm.h"

namespace hdr {

struct alignas(16) GNode {
    float mnx, mny, mnz;
    int left;
    float mxx, mxy, mxz;
    int count;
};
struct alignas(16) GTri {
    float v0x, v0y, v0z;
    int orig;
    float v1x, v1y, v1z;
    int _p1;
    float v2x, v2y, v2z;
    int _p2;
};
struct GTriAttr {
    float n0[3], n1[3], n2[3];
    float c0[3], c1[3], c2[3];
};
struct GInstance {
    float inv[16];
    float xf[16];
    int root;      // BLAS root node (global index), -1 for an empty mesh
    int material;  // material record (default material = num_materials)
    int light;     // light index this mesh prim is (GetLightAtHit), or -1
    int has_params;
};

// Closure parameters of one material (PreviewSurfaceParams -> BSDFClosure, preview_surface.cpp:61-124).
struct GMaterial {
    float base[3];
    float emission[3];    // EmissionColor * Emission
    float spec_color[3];
    float coat_color[3];
    float sheen_color[3];
    float ss_color[3];
    float trans_color[3];
    float metallic, roughness, ior, specular, transmission, trans_depth;
    float coat, coat_roughness, coat_ior, sheen, sheen_roughness, subsurface, opacity;
    int thin_walled;
    int is_default;       // DefaultMaterial: base = hit albedo, normal = geometric, Lambert sampling
    int _pad[3];
};

// One light with the per-sample constants of SampleLight precomputed on the host with the same
// float operations (rect_light.cpp:14-37, point_light.cpp:16-55, distant_light.cpp:10-17).
struct GLight {
    int type;
    int is_delta;
    float width, height;
    float color[3];       // Color * EffectiveIntensity()
    float area;           // Width * Height
    float m[16];          // GfMatrix4f(Transform)
    float normal[3];      // rect: Normalize(TransformDir(0,0,-1)); distant: its direction
    float inv_area;       // 1 / area
    float pos[3];         // point: ExtractTranslation
    float cone_angle;
    float cos_cone, cos_inner, cone_rad, softness;
    int dome;             // index into the dome table, -1
    float eff_int;        // EffectiveIntensity() (dome: texel * Color * this, dome_light.cpp:98,147)
    int _pad[2];
    float raw_color[3];   // Color
    float _pad2;
};

struct GDome {
    const float *pixels;  // w*h*3
    const float *row_cdf; // h+1
    const float *col_cdf; // h*(w+1)
    int w, h;
    float total_lum;
    float pdf_scale;      // totalLum * (pi/h) * (2pi/w)
};

struct DeviceScene {
    const GNode *nodes;
    const GTri *tris;
    const GTriAttr *attrs;
    const int *tlas_idx;
    const int *tlas_rank;  // slot -> the instance's slot in the reference's TLAS (closest-hit tie key)
    const GInstance *instances;
    const GMaterial *materials;
    const GLight *lights;
    const GDome *domes;
    int tlas_root;        // -1 when the scene has no instances
    int num_instances;
    int num_materials;
    int num_lights;
    int environment;      // light index of the environment dome, -1
    int env_dome;         // dome index of the environment, -1
    int has_mesh_lights;  // some instance maps to a light (GetLightAtHit can return non-null)
    int has_emission;     // some material has non-zero emission
};

// Traversal statistics (only accumulated when counting is enabled).
struct TraceStats {
    uint32_t nodes, tris, instances;
};

// IntersectAABB (bvh.cpp:32-64 with double thresholds; scene.cpp:13-46 with float thresholds)
// is box_hit below, with the per-ray reciprocals hoisted.

// IntersectTriangle, bvh.cpp:8-30 (Moller-Trumbore; reject |det| < 1e-8, accept t > 1e-4).
__device__ __forceinline__ bool tri_hit(const GTri &t, V3 o, V3 d, float &tt, float &uu, float &vv) {
    const V3 v0 = v3(t.v0x, t.v0y, t.v0z);
    const V3 e1 = v3(t.v1x, t.v1y, t.v1z) - v0;
    const V3 e2 = v3(t.v2x, t.v2y, t.v2z) - v0;
    const V3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    // double(|det|) < 1e-8 and double(tt) > 1e-4 as float compares: 1e-8f and 1e-4f are the
    // largest floats below 1e-8 and 1e-4 (no float lies between), so the outcomes are the same
    if (fabsf(det) <= 1e-8f) return false;
    const float invDet = 1.0f / det;
    const V3 tvec = o - v0;
    const float u = dot(tvec, pvec) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 qvec = cross(tvec, e1);
    const float v = dot(d, qvec) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    tt = dot(e2, qvec) * invDet;
    uu = u;
    vv = v;
    return tt > 1e-4f;
}

// Loads and stores through the global address space: the scene's and the views' arrays are device
// allocations, so global_load / global_store (a generic pointer would become flat_load, whose
// completion the wave must also wait for with the LDS and scalar-load counter).
typedef float GF4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload4(const void *p) {
    const GF4 v = *(const __attribute__((address_space(1))) GF4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore4(void *p, float4 v) {
    GF4 w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    w.w = v.w;
    *(__attribute__((address_space(1))) GF4 *)p = w;
}
__device__ __forceinline__ int gload_i(const int *p) { return *(const __attribute__((address_space(1))) int *)p; }

__device__ __forceinline__ GTri load_tri(const GTri *tris, int i) {
    const float4 *p = reinterpret_cast<const float4 *>(tris + i);
    const float4 a = gload4(p), b = gload4(p + 1), c = gload4(p + 2);
    GTri t;
    t.v0x = a.x;
    t.v0y = a.y;
    t.v0z = a.z;
    t.orig = __float_as_int(a.w);
    t.v1x = b.x;
    t.v1y = b.y;
    t.v1z = b.z;
    t._p1 = __float_as_int(b.w);
    t.v2x = c.x;
    t.v2y = c.y;
    t.v2z = c.z;
    t._p2 = __float_as_int(c.w);
    return t;
}
// An instance's inverse transform and BLAS root (the traversal's view of a GInstance).
__device__ __forceinline__ int load_inst_inv(const GInstance *insts, int ii, float m[16]) {
    const float4 *p = reinterpret_cast<const float4 *>(insts + ii);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = gload4(p + q);
        m[4 * q] = v.x;
        m[4 * q + 1] = v.y;
        m[4 * q + 2] = v.z;
        m[4 * q + 3] = v.w;
    }
    return __float_as_int(gload4(p + 8).x);  // root: the first int after inv[16] and xf[16]
}

__device__ __forceinline__ GNode load_node(const GNode *nodes, int i) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    const float4 a = gload4(p), b = gload4(p + 1);
    GNode n;
    n.mnx = a.x;
    n.mny = a.y;
    n.mnz = a.z;
    n.left = __float_as_int(a.w);
    n.mxx = b.x;
    n.mxy = b.y;
    n.mxz = b.z;
    n.count = __float_as_int(b.w);
    return n;
}

struct RawHit {
    float t;
    int inst;
    int tri;   // global triangle record index
    float u, v;
};

constexpr int kStackDepth = 64;

// Traversal stacks. An entry is a node reference and, for closest-hit queries, the entry distance
// of its box (kTa). PrivStack keeps every entry in private (scratch) memory. RingStack keeps the
// top kRing entries of each lane in LDS -- the lane's column of a wave-shared array, slot i at
// lds[i * 64] -- and moves older entries to private memory only when the ring overflows, so the
// pops on the traversal's critical path read LDS.
template <bool kTa>
struct PrivStack {
    // The entries live in arrays of the caller (private memory); the cursor is kept apart from them
    // so it stays in a register (a struct holding both would keep it in scratch beside the
    // dynamically indexed array).
    int *ref;
    float *ta;  // entry distances (kTa); any-hit queries keep none
    int sp;
    __device__ __forceinline__ PrivStack(int *r, float *t) : ref(r), ta(t), sp(0) {}
    __device__ __forceinline__ void reset() { sp = 0; }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ void push(int r, float t) {
        ref[sp] = r;
        if (kTa) ta[sp] = t;
        ++sp;
    }
    __device__ __forceinline__ void pop(int &r, float &t) {
        --sp;
        r = ref[sp];
        t = kTa ? ta[sp] : 0.0f;
    }
};

template <int kRing, bool kTa>
struct RingStack {
    static_assert((kRing & (kRing - 1)) == 0, "ring size must be a power of two");
    int *lr;    // this lane's LDS column of references (slot i at lr[i * 64])
    float *lt;  // and of entry distances (kTa)
    int *ref;   // older entries: arrays of the caller in private memory (kept apart from the cursors)
    float *ta;
    int sp, valid;  // entries on the stack; how many of the top ones are in the ring
    __device__ __forceinline__ void bind(int *r, float *t, int *pr, float *pt) {
        lr = r;
        lt = t;
        ref = pr;
        ta = pt;
        sp = valid = 0;
    }
    __device__ __forceinline__ void reset() { sp = valid = 0; }
    __device__ __forceinline__ bool empty() const { return sp == 0; }
    __device__ __forceinline__ void push(int r, float t) {
        const int slot = (sp & (kRing - 1)) * 64;
        if (valid == kRing) {  // evict the oldest ring entry (stack position sp - kRing)
            ref[sp - kRing] = lr[slot];
            if (kTa) ta[sp - kRing] = lt[slot];
        } else {
            ++valid;
        }
        lr[slot] = r;
        if (kTa) lt[slot] = t;
        ++sp;
    }
    __device__ __forceinline__ void pop(int &r, float &t) {
        --sp;
        if (valid > 0) {
            const int slot = (sp & (kRing - 1)) * 64;
            r = lr[slot];
            t = kTa ? lt[slot] : 0.0f;
            --valid;
        } else {
            r = ref[sp];
            t = kTa ? ta[sp] : 0.0f;
        }
    }
};

// Per-ray constants of IntersectAABB: origin, reciprocal direction and parallel-axis flags,
// computed once with the same float operations the per-node test performs, so every box test
// sees the values the reference computes per node.
struct RayBox {
    float o[3], inv[3];
    bool par[3];
    bool anypar;  // some axis is parallel (the general slab test is needed)
};
template <bool kFloatThresholds>
__device__ __forceinline__ RayBox make_raybox(V3 o, V3 d) {
    RayBox r;
    const float oo[3] = {o.x, o.y, o.z};
    const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.o[i] = oo[i];
        r.par[i] = kFloatThresholds ? (fabsf(dd[i]) < 1e-8f) : (double(fabsf(dd[i])) < 1e-8);
        r.inv[i] = r.par[i] ? 0.0f : 1.0f / dd[i];
    }
    r.anypar = r.par[0] || r.par[1] || r.par[2];
    return r;
}
// Branch-free form of the reference's test, with the same outcome: the per-axis early exits
// (tmin > tmax after an axis; a parallel axis outside its slab) are folded into one verdict at the
// end -- tmin only grows and tmax only shrinks from axis to axis (NaN operands are ignored by
// smax / smin), so a box rejected after some axis is still rejected after the last one. The
// reference's double threshold double(tmax) > 1e-4 equals tmax > 1e-4f (1e-4f is the largest
// float below 1e-4, and no float lies between it and 1e-4), so both variants share it.
//
// Rays without a parallel axis (nearly all) take a shorter form: the slab ends are finite there
// (|1/d| <= 1e8 on finite boxes), so the swap is a min / max pair, and v_max / v_min stand for the
// reference's std::max / std::min (they differ only for NaN operands, which finite slabs do not
// produce, and in the sign of a zero result, which no comparison sees).
template <bool kFloatThresholds>
__device__ __forceinline__ bool box_hit(const RayBox &r, const GNode &n, float &tmin_out) {
    bool miss = n.mnx > n.mxx || n.mny > n.mxy || n.mnz > n.mxz;  // GfRange3f::IsEmpty
    const float mn[3] = {n.mnx, n.mny, n.mnz};
    const float mx[3] = {n.mxx, n.mxy, n.mxz};
    if (!r.anypar) {
        float tmin = -1e30f, tmax = 1e30f;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float t1 = (mn[i] - r.o[i]) * r.inv[i];
            const float t2 = (mx[i] - r.o[i]) * r.inv[i];
            tmin = fmaxf(tmin, fminf(t1, t2));
            tmax = fminf(tmax, fmaxf(t1, t2));
        }
        tmin_out = tmin;
        return !miss && !(tmin > tmax) && tmax > 1e-4f;
    }
    float tmin = -1e30f, tmax = 1e30f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float t1 = (mn[i] - r.o[i]) * r.inv[i];
        const float t2 = (mx[i] - r.o[i]) * r.inv[i];
        const bool sw = t1 > t2;
        const float lo = sw ? t2 : t1, hi = sw ? t1 : t2;
        miss = miss || (r.par[i] && (r.o[i] < mn[i] || r.o[i] > mx[i]));
        tmin = r.par[i] ? tmin : smax(tmin, lo);
        tmax = r.par[i] ? tmax : smin(tmax, hi);
    }
    tmin_out = tmin;
    return !miss && !(tmin > tmax) && tmax > 0.0f && tmax > 1e-4f;
}

// Stack / cursor form of a node: inner -> index of its left child (the right one follows it);
// leaf -> -((first << 3) | count) - 1 (count <= 4 triangles or <= 2 instances).
__device__ __forceinline__ int node_ref(const GNode &n) {
    return n.left >= 0 ? n.left : -((((-n.left) - 1) << 3) | n.count) - 1;
}

// No node to process: the stack is popped next (leaf codes stay above it: fewer than 2^28
// triangles and instance slots, checked on the host).
constexpr int kNoNode = INT_MIN;

// Scene::Intersect (scene.cpp:409-513) + BVH::_IntersectNode (bvh.cpp:251-305) return the closest
// hit; among hits at the same t the reference keeps the one it meets first in its left-first
// depth-first order (strict t < best at every level, instances in TLAS-leaf order). That order is
// the order of the arrays: both of its builds partition a node's range into [left child, right
// child), so its TLAS slots and each BLAS's triangle records are laid out in left-first DFS
// order. This traversal visits the nearer child first (a closer hit is found sooner and prunes
// more) and keeps the reference's winner by breaking t ties on (the instance's slot in the
// reference's TLAS -- tlas_rank --, triangle record): the earliest in the reference's order wins,
// whatever order the candidates are met in and whatever tree holds the instances. A box is pruned
// when its entry distance exceeds the best t (ties are still visited), at push and again at pop;
// a parent's box contains its children's and the slab test is monotone in the box, so every
// triangle at or below the best t is tested.
//
// One loop serves both levels: a lane is either in the TLAS (world ray) or in an instance's BLAS
// (object ray), and an inner node of either level runs the same code, so the lanes of a wave do not
// wait at level boundaries for each other.
template <bool kCount, class TS, class BS>
__device__ __forceinline__ bool trace_closest_s(const DeviceScene &s, V3 o, V3 d, RawHit &hit, TraceStats &st,
                                                TS &tstk, BS &bstk) {
    hit.t = 1e30f;
    hit.inst = -1;
    if (s.tlas_root < 0) return false;
    RayBox R = make_raybox<true>(o, d);  // the current level's ray
    tstk.reset();
    int cur;
    {
        const GNode n = load_node(s.nodes, s.tlas_root);
        if (kCount) ++st.nodes;
        float ta;
        if (!box_hit<true>(R, n, ta) || ta > hit.t) return false;
        cur = node_ref(n);
    }
    bool was = false, blas = false, world = true;
    int lnext = 0, lend = 0;  // instance slots of the current TLAS leaf still to traverse
    int ii = -1, slot = -1;   // current instance and its rank (tlas_rank of its TLAS slot)
    int bslot = 0x7fffffff;   // rank of the best hit's instance (tie-break key with hit.tri)
    V3 oo = o, od = d;
    for (;;) {
        if (cur >= 0) {  // inner node of either level
            const GNode l = load_node(s.nodes, cur), r = load_node(s.nodes, cur + 1);
            if (kCount) st.nodes += 2;
            float tl, tr;
            const bool hl = box_hit<true>(R, l, tl) && !(tl > hit.t);
            const bool hr = box_hit<true>(R, r, tr) && !(tr > hit.t);
            if (hl || hr) {
                const bool leftFirst = hl && (!hr || !(tr < tl));
                if (hl && hr) {
                    if (blas) bstk.push(leftFirst ? node_ref(r) : node_ref(l), leftFirst ? tr : tl);
                    else tstk.push(leftFirst ? node_ref(r) : node_ref(l), leftFirst ? tr : tl);
                }
                cur = leftFirst ? node_ref(l) : node_ref(r);
                continue;
            }
        } else if (cur != kNoNode) {  // leaf
            const int code = -cur - 1;
            const int start = code >> 3, count = code & 7;
            if (blas) {
                for (int q = 0; q < count; ++q) {
                    const GTri tri = load_tri(s.tris, start + q);
                    if (kCount) ++st.tris;
                    float tt, uu, vv;
                    if (tri_hit(tri, oo, od, tt, uu, vv) &&
                        (tt < hit.t || (tt == hit.t && (slot < bslot || (slot == bslot && start + q < hit.tri))))) {
                        hit.t = tt;
                        hit.inst = ii;
                        hit.tri = start + q;
                        hit.u = uu;
                        hit.v = vv;
                        bslot = slot;
                        was = true;
                    }
                }
            } else {
                lnext = start;
                lend = start + count;
            }
        }
        cur = kNoNode;
        if (blas) {
            bool more = false;
            while (!bstk.empty()) {
                int rr;
                float ta;
                bstk.pop(rr, ta);
                if (ta > hit.t) continue;
                cur = rr;
                more = true;
                break;
            }
            if (more) continue;
            blas = false;  // the instance is done
        }
        while (lnext < lend) {  // the leaf's next instance
            ii = gload_i(s.tlas_idx + lnext);
            slot = gload_i(s.tlas_rank + lnext++);
            float minv[16];
            const int root = load_inst_inv(s.instances, ii, minv);
            if (kCount) ++st.instances;
            if (root < 0) continue;
            oo = xform_point(minv, o);
            od = xform_dir(minv, d);
            R = make_raybox<false>(oo, od);
            world = false;
            bstk.reset();
            const GNode b = load_node(s.nodes, root);
            if (kCount) ++st.nodes;
            float ta;
            if (box_hit<true>(R, b, ta) && !(ta > hit.t)) {
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
        bool more = false;
        while (!tstk.empty()) {
            int rr;
            float ta;
            tstk.pop(rr, ta);
            if (ta > hit.t) continue;
            cur = rr;
            more = true;
            break;
        }
        if (!more) break;
    }
    return was;
}
// ---------------------------------------------------------------- wave-coherent closest hit
// The 64 rays of one wave (the primary rays of an 8x8 tile, coherent) traverse together: the wave
// walks the union of the nodes its rays visit, with one stack of {node record, lane mask} per level
// in LDS, and loads every node, triangle and instance once for all lanes, from uniform addresses
// (scalar loads). Each lane keeps its own decisions: the box test and the prune against its own
// best t at every node, again when an entry is popped (the box is tested again: same operands,
// so the same entry distance as at the push), and the tie rule of trace_closest_s in the leaves.
// So each lane tests the triangles it would test alone and returns the same hit; only the visiting
// order differs (the child that more lanes meet first goes first), and the tie rule makes the
// winner independent of the order. All 64 lanes of the wave call it together; lanes without a ray
// pass act = false.
struct WaveStack {
    int rec[kStackDepth];                  // node record index (the box to test again at the pop)
    unsigned long long mask[kStackDepth];  // lanes that entered it
};

__device__ __forceinline__ float4 sload4(const void *p) {  // uniform address: a scalar load
    const GF4 v = *(const __attribute__((address_space(4))) GF4 *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ GNode load_node_u(const GNode *nodes, int i) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    const float4 a = sload4(p), b = sload4(p + 1);
    GNode n;
    n.mnx = a.x;
    n.mny = a.y;
    n.mnz = a.z;
    n.left = __float_as_int(a.w);
    n.mxx = b.x;
    n.mxy = b.y;
    n.mxz = b.z;
    n.count = __float_as_int(b.w);
    return n;
}
__device__ __forceinline__ GTri load_tri_u(const GTri *tris, int i) {
    const float4 *p = reinterpret_cast<const float4 *>(tris + i);
    const float4 a = sload4(p), b = sload4(p + 1), c = sload4(p + 2);
    GTri t;
    t.v0x = a.x;
    t.v0y = a.y;
    t.v0z = a.z;
    t.orig = __float_as_int(a.w);
    t.v1x = b.x;
    t.v1y = b.y;
    t.v1z = b.z;
    t._p1 = __float_as_int(b.w);
    t.v2x = c.x;
    t.v2y = c.y;
    t.v2z = c.z;
    t._p2 = __float_as_int(c.w);
    return t;
}
__device__ __forceinline__ int load_inst_inv_u(const GInstance *insts, int ii, float m[16]) {
    const float4 *p = reinterpret_cast<const float4 *>(insts + ii);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = sload4(p + q);
        m[4 * q] = v.x;
        m[4 * q + 1] = v.y;
        m[4 * q + 2] = v.z;
        m[4 * q + 3] = v.w;
    }
    return __float_as_int(sload4(p + 8).x);
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v)))),
                   hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v >> 32))));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Pops the top entries of a wave stack until one still holds a lane whose box test passes against
// its current best t; returns false when the stack empties.
__device__ __forceinline__ bool wave_pop(const DeviceScene &s, WaveStack &S, int &sp, const RayBox &R, float best,
                                         int lane, int &cur, unsigned long long &cm) {
    while (sp > 0) {
        --sp;
        const int rec = uniform_i(S.rec[sp]);
        const unsigned long long m = uniform_u64(S.mask[sp]);
        const GNode n = load_node_u(s.nodes, rec);
        float ta;
        const bool h = ((m >> lane) & 1ULL) && box_hit<true>(R, n, ta) && !(ta > best);
        const unsigned long long mm = __ballot(h);
        if (mm != 0ULL) {
            cur = node_ref(n);
            cm = mm;
            return true;
        }
    }
    return false;
}
__device__ __forceinline__ void wave_push(WaveStack &S, int &sp, int rec, unsigned long long m) {
    S.rec[sp] = rec;  // every lane writes the same value
    S.mask[sp] = m;
    ++sp;
}

__device__ __forceinline__ bool trace_closest_wave(const DeviceScene &s, bool act, V3 o, V3 d, RawHit &hit,
                                                   WaveStack &tstk, WaveStack &bstk) {
    const int lane = int(threadIdx.x & 63u);
    hit.t = 1e30f;
    hit.inst = -1;
    hit.tri = -1;
    hit.u = hit.v = 0.0f;
    if (s.tlas_root < 0) return false;
    RayBox R = make_raybox<true>(o, d);
    int tsp = 0, bsp = 0;
    unsigned long long cm;  // lanes in the current node
    int cur;
    {
        const GNode n = load_node_u(s.nodes, s.tlas_root);
        float ta;
        const bool h = act && box_hit<true>(R, n, ta) && !(ta > hit.t);
        cm = __ballot(h);
        if (cm == 0ULL) return false;
        cur = node_ref(n);
    }
    bool was = false, blas = false, world = true;
    int lnext = 0, lend = 0;    // instance slots of the current TLAS leaf still to traverse
    unsigned long long lm = 0;  // and the lanes that reached that leaf
    int ii = -1, slot = -1;
    int bslot = 0x7fffffff;
    V3 oo = o, od = d;
    for (;;) {
        if (cur >= 0) {  // inner node of either level
            const GNode l = load_node_u(s.nodes, cur), r = load_node_u(s.nodes, cur + 1);
            const bool in = (cm >> lane) & 1ULL;
            float tl = 0.0f, tr = 0.0f;
            const bool hl = in && box_hit<true>(R, l, tl) && !(tl > hit.t);
            const bool hr = in && box_hit<true>(R, r, tr) && !(tr > hit.t);
            const unsigned long long ml = __ballot(hl), mr = __ballot(hr);
            if ((ml | mr) != 0ULL) {
                const unsigned long long mf = __ballot(hl && (!hr || !(tr < tl)));
                const bool leftFirst = ml != 0ULL && (mr == 0ULL || 2 * __popcll(mf) >= __popcll(ml | mr));
                if (ml != 0ULL && mr != 0ULL) {
                    if (blas) wave_push(bstk, bsp, leftFirst ? cur + 1 : cur, leftFirst ? mr : ml);
                    else wave_push(tstk, tsp, leftFirst ? cur + 1 : cur, leftFirst ? mr : ml);
                }
                cur = node_ref(leftFirst ? l : r);
                cm = leftFirst ? ml : mr;
                continue;
            }
        } else if (cur != kNoNode) {  // leaf
            const int code = -cur - 1;
            const int start = code >> 3, count = code & 7;
            if (blas) {
                const bool in = (cm >> lane) & 1ULL;
                for (int q = 0; q < count; ++q) {
                    const GTri tri = load_tri_u(s.tris, start + q);
                    float tt, uu, vv;
                    if (in && tri_hit(tri, oo, od, tt, uu, vv) &&
                        (tt < hit.t || (tt == hit.t && (slot < bslot || (slot == bslot && start + q < hit.tri))))) {
                        hit.t = tt;
                        hit.inst = ii;
                        hit.tri = start + q;
                        hit.u = uu;
                        hit.v = vv;
                        bslot = slot;
                        was = true;
                    }
                }
            } else {
                lnext = start;
                lend = start + count;
                lm = cm;
            }
        }
        cur = kNoNode;
        if (blas) {
            if (wave_pop(s, bstk, bsp, R, hit.t, lane, cur, cm)) continue;
            blas = false;  // the instance is done
        }
        while (lnext < lend) {  // the leaf's next instance
            ii = *(const __attribute__((address_space(4))) int *)(s.tlas_idx + lnext);
            slot = *(const __attribute__((address_space(4))) int *)(s.tlas_rank + lnext++);
            float minv[16];
            const int root = load_inst_inv_u(s.instances, ii, minv);
            if (root < 0) continue;
            oo = xform_point(minv, o);
            od = xform_dir(minv, d);
            R = make_raybox<false>(oo, od);
            world = false;
            bsp = 0;
            const GNode b = load_node_u(s.nodes, root);
            float ta;
            const bool h = ((lm >> lane) & 1ULL) && box_hit<true>(R, b, ta) && !(ta > hit.t);
            const unsigned long long m = __ballot(h);
            if (m != 0ULL) {
                cur = node_ref(b);
                cm = m;
                blas = true;
                break;
            }
        }
        if (blas) continue;
        if (!world) {
            R = make_raybox<true>(o, d);
            world = true;
        }
        if (!wave_pop(s, tstk, tsp, R, hit.t, lane, cur, cm)) break;
    }
    return was;
}

template <bool kCount>
__device__ __forceinline__ bool trace_closest(const DeviceScene &s, V3 o, V3 d, RawHit &hit, TraceStats &st) {
    int tr[kStackDepth], br[kStackDepth];
    float tt[kStackDepth], bt[kStackDepth];
    PrivStack<true> t(tr, tt), b(br, bt);
    return trace_closest_s<kCount>(s, o, d, hit, st, t, b);
}

// Shadow query: is there a hit with 1e-4 < t < tmax? The reference fires a closest-hit ray and
// tests Depth < Dist - 1e-3 (ris_direct_light_integrator.cpp:513-519, mis_direct_light_integrator.cpp:100-105);
// the same traversal with the bound fixed at tmax, stopped at the first accepted triangle,
// answers the same question (a box entered before the closest hit cannot be pruned by a
// larger bound), so subtrees are pruned as soon as their box is tested. The answer does not depend
// on the order subtrees are visited in, so the nearer child goes first (occluders are met sooner).
// One loop for both levels, as in trace_closest_s.
template <bool kCount, class TS, class BS>
__device__ __forceinline__ bool trace_any_s(const DeviceScene &s, V3 o, V3 d, float tmax, TraceStats &st, TS &tstk,
                                            BS &bstk) {
    if (s.tlas_root < 0) return false;
    RayBox R = make_raybox<true>(o, d);
    tstk.reset();
    int cur;
    {
        const GNode n = load_node(s.nodes, s.tlas_root);
        if (kCount) ++st.nodes;
        float ta;
        if (!box_hit<true>(R, n, ta) || ta > tmax) return false;
        cur = node_ref(n);
    }
    bool blas = false, world = true;
    int lnext = 0, lend = 0;
    V3 oo = o, od = d;
    for (;;) {
        if (cur >= 0) {  // nearer child first: any occluder ends the query, so order is free
            const GNode l = load_node(s.nodes, cur), r = load_node(s.nodes, cur + 1);
            if (kCount) st.nodes += 2;
            float tl, tr;
            const bool hl = box_hit<true>(R, l, tl) && !(tl > tmax);
            const bool hr = box_hit<true>(R, r, tr) && !(tr > tmax);
            if (hl || hr) {
                const bool leftFirst = hl && (!hr || !(tr < tl));
                const int first = leftFirst ? node_ref(l) : node_ref(r);
                if (hl && hr) {
                    const int second = leftFirst ? node_ref(r) : node_ref(l);
                    if (blas) bstk.push(second, 0.0f);
                    else tstk.push(second, 0.0f);
                }
                cur = first;
                continue;
            }
        } else if (cur != kNoNode) {
            const int code = -cur - 1;
            const int start = code >> 3, count = code & 7;
            if (blas) {
                for (int q = 0; q < count; ++q) {
                    const GTri tri = load_tri(s.tris, start + q);
                    if (kCount) ++st.tris;
                    float tt, uu, vv;
                    if (tri_hit(tri, oo, od, tt, uu, vv) && tt < tmax) return true;
                }
            } else {
                lnext = start;
                lend = start + count;
            }
        }
        cur = kNoNode;
        if (blas) {
            if (!bstk.empty()) {
                float unused;
                bstk.pop(cur, unused);
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
            bstk.reset();
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
        if (tstk.empty()) break;
        float unused;
        tstk.pop(cur, unused);
    }
    return false;
}
template <bool kCount>
__device__ __forceinline__ bool trace_any(const DeviceScene &s, V3 o, V3 d, float tmax, TraceStats &st) {
    int tr[kStackDepth], br[kStackDepth];
    PrivStack<false> t(tr, nullptr), b(br, nullptr);
    return trace_any_s<kCount>(s, o, d, tmax, st, t, b);
}

// HitRecord of the winner (Scene::Intersect :456-497 and IntersectScene :515-535): world
// normals through TransformDir + GetNormalized, albedo = material diffuse colour or the
// interpolated vertex colour.
struct HitRec {
    V3 pos, ng, ns, albedo;
    float depth;
    int inst;
    int tri;   // mesh-local original triangle index
};

__device__ __forceinline__ HitRec make_hit_record(const DeviceScene &s, V3 o, V3 d, const RawHit &h) {
    const GInstance &inst = s.instances[h.inst];
    const GTri tr = s.tris[h.tri];
    const GTriAttr &at = s.attrs[h.tri];
    const V3 v0 = v3(tr.v0x, tr.v0y, tr.v0z), v1 = v3(tr.v1x, tr.v1y, tr.v1z), v2 = v3(tr.v2x, tr.v2y, tr.v2z);
    const V3 nObj = normalized(cross(v1 - v0, v2 - v0));
    const float w = 1.0f - h.u - h.v;
    const V3 n0 = v3(at.n0[0], at.n0[1], at.n0[2]);
    V3 sObj;
    if (dot(n0, n0) > 1e-6f) {
        const V3 n1 = v3(at.n1[0], at.n1[1], at.n1[2]), n2 = v3(at.n2[0], at.n2[1], at.n2[2]);
        sObj = normalized(n0 * w + n1 * h.u + n2 * h.v);
    } else {
        sObj = nObj;
    }
    HitRec r;
    r.pos = o + h.t * d;
    r.ng = normalized(xform_dir(inst.xf, nObj));
    r.ns = normalized(xform_dir(inst.xf, sObj));
    if (inst.has_params) {
        const GMaterial &m = s.materials[inst.material];
        r.albedo = v3(m.base[0], m.base[1], m.base[2]);
    } else {
        const V3 c0 = v3(at.c0[0], at.c0[1], at.c0[2]), c1 = v3(at.c1[0], at.c1[1], at.c1[2]),
                 c2 = v3(at.c2[0], at.c2[1], at.c2[2]);
        r.albedo = c0 * w + c1 * h.u + c2 * h.v;
    }
    r.depth = h.t;
    r.inst = h.inst;
    r.tri = tr.orig;
    return r;
}

// Out-of-line copies for the shading megakernel: one traversal body (and one stack frame)
// instead of one inlined copy per call site.
template <bool kCount>
__device__ __noinline__ bool trace_closest_call(const DeviceScene &s, V3 o, V3 d, RawHit &hit, TraceStats &st) {
    return trace_closest<kCount>(s, o, d, hit, st);
}
template <bool kCount>
__device__ __noinline__ bool trace_any_call(const DeviceScene &s, V3 o, V3 d, float tmax, TraceStats &st) {
    return trace_any<kCount>(s, o, d, tmax, st);
}
__device__ __noinline__ HitRec make_hit_record_call(const DeviceScene &s, V3 o, V3 d, const RawHit &h) {
    return make_hit_record(s, o, d, h);
}

}  // namespace hdr
