// host_scene.cpp -- builds the device scene layout on the host.
//
// BLAS: BVH::Build / BVH::_Subdivide (source/renderContext/bvh.cpp:66-241): midpoint split on
// the longest axis of the node bounds, leaf <= 4 triangles, in-place partition by centroid with
// the same swap sequence, fallback to a median split. TLAS: Scene::_BuildTLAS /
// _SubdivideTLAS (source/scene/scene.cpp:333-407), leaf <= 2 instances. Reproducing the trees
// node for node keeps the traversal order -- and the choice among equal-t hits -- identical.
// Light, material and dome constants are evaluated here once with the float operations the
// reference performs per sample (rect_light.cpp, point_light.cpp, distant_light.cpp,
// dome_light.cpp:21-77, preview_surface.cpp:61-124).
#include "host_scene.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace hdr {
namespace {

struct F3 {
    float x, y, z;
};
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 mulf(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 muld(F3 a, double s) { return {float(double(a.x) * s), float(double(a.y) * s), float(double(a.z) * s)}; }
inline float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 normalizedf(F3 a) {
    const float len = std::sqrt(dotf(a, a));
    const float eps = float(1e-10);
    return muld(a, 1.0 / double(len > eps ? len : eps));
}
inline float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Box {  // GfRange3f
    F3 mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void extend(F3 p) {
        if (p.x < mn.x) mn.x = p.x;
        if (p.y < mn.y) mn.y = p.y;
        if (p.z < mn.z) mn.z = p.z;
        if (p.x > mx.x) mx.x = p.x;
        if (p.y > mx.y) mx.y = p.y;
        if (p.z > mx.z) mx.z = p.z;
    }
    bool empty() const { return mn.x > mx.x || mn.y > mx.y || mn.z > mx.z; }
    F3 size() const { return sub(mx, mn); }
};

GNode make_node(const Box &b, int left, int count) {
    GNode n;
    n.mnx = b.mn.x;
    n.mny = b.mn.y;
    n.mnz = b.mn.z;
    n.mxx = b.mx.x;
    n.mxy = b.mx.y;
    n.mxz = b.mx.z;
    n.left = left;
    n.count = count;
    return n;
}

// Trees are appended at odd indices: a tree's sibling pairs sit at (odd, even) local offsets
// 1-2, 3-4, ..., so with an odd base every pair starts on a 64-byte boundary of the 256-byte
// aligned node array and both children arrive in one fetch.
void pad_to_odd(std::vector<GNode> &nodes) {
    if (nodes.size() % 2 == 0) {
        GNode pad{};
        pad.mnx = pad.mny = pad.mnz = 1.0f;  // empty box; the record is never referenced
        pad.mxx = pad.mxy = pad.mxz = -1.0f;
        nodes.push_back(pad);
    }
}

struct BuildTri {
    F3 v0, v1, v2, centroid;
    int orig;
    GTriAttr attr;
};

struct BlasBuilder {
    std::vector<BuildTri> tris;
    std::vector<GNode> nodes;  // local indices; leaf left = -(local start) - 1
    std::vector<Box> bounds;
    int depth = 0;

    void subdivide(int nodeIdx, int start, int end, int d) {
        depth = std::max(depth, d);
        Box b;
        for (int i = start; i < end; ++i) {
            b.extend(tris[i].v0);
            b.extend(tris[i].v1);
            b.extend(tris[i].v2);
        }
        const int count = end - start;
        if (count <= 4) {
            nodes[nodeIdx] = make_node(b, -start - 1, count);
            return;
        }
        const F3 size = b.size();
        int axis = 0;
        if (size.y > size.x) axis = 1;
        if (size.z > comp(size, axis)) axis = 2;
        const float splitPos = comp(b.mn, axis) + comp(size, axis) * 0.5f;
        int i = start, j = end - 1;
        while (i <= j) {
            if (comp(tris[i].centroid, axis) < splitPos) {
                i++;
            } else {
                std::swap(tris[i], tris[j]);
                j--;
            }
        }
        const int leftCount = i - start;
        if (leftCount == 0 || leftCount == count) i = start + count / 2;
        const int leftIdx = int(nodes.size());
        nodes.push_back(GNode{});
        nodes.push_back(GNode{});
        nodes[nodeIdx] = make_node(b, leftIdx, 0);
        subdivide(leftIdx, start, i, d + 1);
        subdivide(leftIdx + 1, i, end, d + 1);
    }
};

}  // namespace

void invert4(const double a[4][4], double out[4][4]) {
    const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det == 0.0) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) out[i][j] = (i == j) ? double(FLT_MAX) : 0.0;
        return;
    }
    const double r = 1.0 / det;
    out[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * r;
    out[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * r;
    out[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * r;
    out[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * r;
    out[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * r;
    out[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * r;
    out[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * r;
    out[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * r;
    out[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * r;
    out[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * r;
    out[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * r;
    out[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * r;
    out[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * r;
    out[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * r;
    out[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * r;
    out[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * r;
}

namespace {

// GfMatrix4f::Transform (float, with GfProject)
F3 xform_point_h(const float *m, F3 p) {
    const float x = p.x * m[0] + p.y * m[4] + p.z * m[8] + m[12];
    const float y = p.x * m[1] + p.y * m[5] + p.z * m[9] + m[13];
    const float z = p.x * m[2] + p.y * m[6] + p.z * m[10] + m[14];
    const float w = p.x * m[3] + p.y * m[7] + p.z * m[11] + m[15];
    const float inv = (w != 0.0f) ? 1.0f / w : 1.0f;
    return {inv * x, inv * y, inv * z};
}
F3 xform_dir_h(const float *m, F3 p) {
    return {p.x * m[0] + p.y * m[4] + p.z * m[8], p.x * m[1] + p.y * m[5] + p.z * m[9],
            p.x * m[2] + p.y * m[6] + p.z * m[10]};
}

bool finite_all(const float *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

}  // namespace

namespace {

// Counts, arrays and record sizes of everything but the meshes.
std::string check_state_arrays(const hdr_scene_desc &d) {
    if (d.struct_size != sizeof(hdr_scene_desc)) return "hdr_scene_desc.struct_size mismatch";
    if (d.num_meshes < 0 || d.num_materials < 0 || d.num_instances < 0 || d.num_lights < 0)
        return "negative element count in hdr_scene_desc";
    if ((d.num_materials && !d.materials) || (d.num_instances && !d.instances) || (d.num_lights && !d.lights))
        return "null array with a non-zero count in hdr_scene_desc";
    for (int i = 0; i < d.num_materials; ++i)
        if (d.materials[i].struct_size != sizeof(hdr_material_desc)) return "hdr_material_desc.struct_size mismatch";
    for (int i = 0; i < d.num_instances; ++i)
        if (d.instances[i].struct_size != sizeof(hdr_instance_desc)) return "hdr_instance_desc.struct_size mismatch";
    for (int i = 0; i < d.num_lights; ++i)
        if (d.lights[i].struct_size != sizeof(hdr_light_desc)) return "hdr_light_desc.struct_size mismatch";
    if (d.environment >= d.num_lights || (d.environment >= 0 && d.lights[d.environment].type != HDR_LIGHT_DOME))
        return "environment must index a dome light";
    return std::string();
}

// Materials, instances, TLAS and lights of a scene, built against its meshes' BLAS.
struct StateBuild {
    std::vector<GNode> tlas;  // TLAS nodes with global indices (placed from tlas_base on)
    int tlas_base = 0;
    HostScene s;              // every field but the geometry
};
std::string build_state(const hdr_scene_desc &d, const HostScene &geo, StateBuild &sb);
// Replaces the state of `out` with the built one; the BLAS nodes stay in front of the TLAS.
void commit_state(StateBuild &sb, HostScene &out);

}  // namespace

std::string build_host_scene(const hdr_scene_desc &d, HostScene &out) {
    if (d.struct_size != sizeof(hdr_scene_desc)) return "hdr_scene_desc.struct_size mismatch";
    if (d.num_meshes > 0 && !d.meshes) return "null array with a non-zero count in hdr_scene_desc";
    for (int i = 0; i < d.num_meshes; ++i)
        if (d.meshes[i].struct_size != sizeof(hdr_mesh_desc)) return "hdr_mesh_desc.struct_size mismatch";
    {
        const std::string e = check_state_arrays(d);
        if (!e.empty()) return e;
    }

    // ---- BLAS per mesh (BVH::Build, bvh.cpp:66-182)
    std::vector<int> &meshRoot = out.mesh_root;
    meshRoot.assign(size_t(d.num_meshes), -1);
    out.mesh_range.assign(size_t(d.num_meshes) * 6, 0.0f);
    for (int mi = 0; mi < d.num_meshes; ++mi) {
        const hdr_mesh_desc &m = d.meshes[mi];
        if (m.num_points < 0 || m.num_triangles < 0) return "negative mesh counts";
        if ((m.num_points && !m.points) || (m.num_triangles && !m.indices)) return "mesh arrays missing";
        if (m.num_points && !finite_all(m.points, size_t(m.num_points) * 3)) return "non-finite mesh point";
        const int nIdx = m.num_triangles;
        BlasBuilder bb;
        bb.tris.reserve(nIdx);
        auto P = [&](int k) { return F3{m.points[3 * k], m.points[3 * k + 1], m.points[3 * k + 2]}; };
        for (int i = 0; i < nIdx; ++i) {
            const int ids[3] = {m.indices[3 * i], m.indices[3 * i + 1], m.indices[3 * i + 2]};
            bool ok = true;
            for (int k = 0; k < 3; ++k) ok = ok && ids[k] >= 0 && ids[k] < m.num_points;
            if (!ok) continue;
            BuildTri t;
            t.v0 = P(ids[0]);
            t.v1 = P(ids[1]);
            t.v2 = P(ids[2]);
            for (int k = 0; k < 3; ++k) {
                F3 n{0.0f, 0.0f, 0.0f}, c{1.0f, 1.0f, 1.0f};
                if (m.num_normals > 0 && m.normals) {
                    const int src = (m.num_normals == nIdx * 3) ? i * 3 + k : ids[k];
                    if (m.num_normals == nIdx * 3 || ids[k] < m.num_normals)
                        n = F3{m.normals[3 * src], m.normals[3 * src + 1], m.normals[3 * src + 2]};
                    else
                        n = F3{0.0f, 1.0f, 0.0f};
                }
                if (m.num_colors > 0 && m.colors) {
                    const int src = (m.num_colors == nIdx * 3) ? i * 3 + k : ids[k];
                    if (m.num_colors == nIdx * 3 || ids[k] < m.num_colors)
                        c = F3{m.colors[3 * src], m.colors[3 * src + 1], m.colors[3 * src + 2]};
                }
                float *nn = k == 0 ? t.attr.n0 : (k == 1 ? t.attr.n1 : t.attr.n2);
                float *cc = k == 0 ? t.attr.c0 : (k == 1 ? t.attr.c1 : t.attr.c2);
                nn[0] = n.x;
                nn[1] = n.y;
                nn[2] = n.z;
                cc[0] = c.x;
                cc[1] = c.y;
                cc[2] = c.z;
            }
            t.centroid = muld(add(add(t.v0, t.v1), t.v2), 1.0 / double(3.0f));
            t.orig = i;
            bb.tris.push_back(t);
        }
        Box range;
        for (int i = 0; i < nIdx; ++i)
            for (int k = 0; k < 3; ++k) {
                const int id = m.indices[3 * i + k];
                if (id >= 0 && id < m.num_points) range.extend(P(id));
            }
        const float rv[6] = {range.mn.x, range.mn.y, range.mn.z, range.mx.x, range.mx.y, range.mx.z};
        std::memcpy(&out.mesh_range[size_t(mi) * 6], rv, sizeof rv);
        if (bb.tris.empty()) continue;
        bb.nodes.reserve(size_t(nIdx) * 2);
        bb.nodes.push_back(GNode{});
        bb.subdivide(0, 0, int(bb.tris.size()), 1);
        out.max_blas_depth = std::max(out.max_blas_depth, bb.depth);
        pad_to_odd(out.nodes);
        const int nodeBase = int(out.nodes.size());
        const int triBase = int(out.tris.size());
        for (GNode n : bb.nodes) {
            if (n.left < 0) n.left = -(triBase + (-n.left - 1)) - 1;
            else n.left += nodeBase;
            out.nodes.push_back(n);
        }
        for (const BuildTri &t : bb.tris) {
            GTri g{};
            g.v0x = t.v0.x;
            g.v0y = t.v0.y;
            g.v0z = t.v0.z;
            g.v1x = t.v1.x;
            g.v1y = t.v1.y;
            g.v1z = t.v1.z;
            g.v2x = t.v2.x;
            g.v2y = t.v2.y;
            g.v2z = t.v2.z;
            g.orig = t.orig;
            out.tris.push_back(g);
            out.attrs.push_back(t.attr);
        }
        meshRoot[mi] = nodeBase;
    }

    out.num_meshes = d.num_meshes;
    out.blas_nodes = int(out.nodes.size());
    if (out.max_blas_depth > kStackDepth - 2) return "BVH deeper than the traversal stack";
    if (out.tris.size() >= (size_t(1) << 28))
        return "more than 2^28 triangles or instances (leaf reference range of the traversal)";
    StateBuild sb;
    const std::string e = build_state(d, out, sb);
    if (!e.empty()) return e;
    commit_state(sb, out);
    return std::string();
}

std::string rebuild_host_state(const hdr_scene_desc &d, HostScene &out) {
    {
        const std::string e = check_state_arrays(d);
        if (!e.empty()) return e;
    }
    if (d.num_meshes != out.num_meshes) return "an update must keep the scene's mesh count";
    StateBuild sb;
    const std::string e = build_state(d, out, sb);
    if (!e.empty()) return e;
    commit_state(sb, out);
    return std::string();
}

namespace {

float half_area(const Box &b) {
    if (b.empty()) return 0.0f;
    const F3 e = b.size();
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

// Binned SAH over the instances idx[start, end) (16 bins per axis over the centroid bounds; cost of
// a split = half-area(left) * n_left + half-area(right) * n_right), down to one instance per leaf;
// where every centroid coincides, the range is halved. Children are allocated as a pair (right =
// left + 1) before the recursion, as the midpoint builds do, so the node layout is the same.
void sah_tlas(std::vector<GNode> &tl, std::vector<int> &idx, const std::vector<Box> &bounds,
              const std::vector<F3> &cen, int nodeIdx, int start, int end, int dpt, int &maxd) {
    maxd = std::max(maxd, dpt);
    Box b, cb;
    for (int i = start; i < end; ++i) {
        b.extend(bounds[size_t(idx[size_t(i)])].mn);
        b.extend(bounds[size_t(idx[size_t(i)])].mx);
        cb.extend(cen[size_t(idx[size_t(i)])]);
    }
    const int count = end - start;
    if (count <= 1) {
        tl[size_t(nodeIdx)] = make_node(b, -start - 1, count);
        return;
    }
    constexpr int kBins = 16;
    int bestAxis = -1, bestBin = 0;
    float bestCost = FLT_MAX;
    const F3 ce = cb.size();
    for (int axis = 0; axis < 3; ++axis) {
        const float lo = comp(cb.mn, axis), ext = comp(ce, axis);
        if (!(ext > 0.0f)) continue;
        Box bb[kBins];
        int bn[kBins] = {};
        for (int i = start; i < end; ++i) {
            const int k = idx[size_t(i)];
            int bin = int(float(kBins) * ((comp(cen[size_t(k)], axis) - lo) / ext));
            bin = std::min(kBins - 1, std::max(0, bin));
            ++bn[bin];
            bb[bin].extend(bounds[size_t(k)].mn);
            bb[bin].extend(bounds[size_t(k)].mx);
        }
        float rightArea[kBins];
        int rightN[kBins];
        Box acc;
        int n = 0;
        for (int bin = kBins - 1; bin > 0; --bin) {
            if (bn[bin]) {  // an empty bin's box is the empty sentinel: no corners to add
                acc.extend(bb[bin].mn);
                acc.extend(bb[bin].mx);
            }
            n += bn[bin];
            rightArea[bin] = half_area(acc);
            rightN[bin] = n;
        }
        acc = Box();
        n = 0;
        for (int bin = 1; bin < kBins; ++bin) {  // split between bin - 1 and bin
            if (bn[bin - 1]) {
                acc.extend(bb[bin - 1].mn);
                acc.extend(bb[bin - 1].mx);
            }
            n += bn[bin - 1];
            if (n == 0 || rightN[bin] == 0) continue;
            const float cost = half_area(acc) * float(n) + rightArea[bin] * float(rightN[bin]);
            if (cost < bestCost) {
                bestCost = cost;
                bestAxis = axis;
                bestBin = bin;
            }
        }
    }
    int mid;
    if (bestAxis < 0) {
        mid = start + count / 2;
    } else {
        const float lo = comp(cb.mn, bestAxis), ext = comp(ce, bestAxis);
        auto left = [&](int k) {
            int bin = int(float(kBins) * ((comp(cen[size_t(k)], bestAxis) - lo) / ext));
            bin = std::min(kBins - 1, std::max(0, bin));
            return bin < bestBin;
        };
        mid = int(std::partition(idx.begin() + start, idx.begin() + end, left) - idx.begin());
        if (mid == start || mid == end) mid = start + count / 2;
    }
    const int leftIdx = int(tl.size());
    tl.push_back(GNode{});
    tl.push_back(GNode{});
    tl[size_t(nodeIdx)] = make_node(b, leftIdx, 0);
    sah_tlas(tl, idx, bounds, cen, leftIdx, start, mid, dpt + 1, maxd);
    sah_tlas(tl, idx, bounds, cen, leftIdx + 1, mid, end, dpt + 1, maxd);
}

std::string build_state(const hdr_scene_desc &d, const HostScene &geo, StateBuild &sb) {
    HostScene &out = sb.s;
    // ---- materials (+ default material record at index num_materials)
    for (int i = 0; i < d.num_materials; ++i) {
        const hdr_material_desc &s = d.materials[i];
        GMaterial g{};
        auto cp = [](float *dst, const float *src) { dst[0] = src[0], dst[1] = src[1], dst[2] = src[2]; };
        cp(g.base, s.diffuse_color);
        g.emission[0] = s.emission_color[0] * s.emission;
        g.emission[1] = s.emission_color[1] * s.emission;
        g.emission[2] = s.emission_color[2] * s.emission;
        cp(g.spec_color, s.specular_color);
        cp(g.coat_color, s.coat_color);
        cp(g.sheen_color, s.sheen_color);
        cp(g.ss_color, s.subsurface_color);
        cp(g.trans_color, s.transmission_color);
        g.metallic = s.metallic;
        g.roughness = std::max(0.001f, s.roughness);
        g.ior = s.ior;
        g.specular = s.specular;
        g.transmission = s.transmission;
        g.trans_depth = s.transmission_depth;
        g.coat = s.coat;
        g.coat_roughness = s.coat_roughness;
        g.coat_ior = s.coat_ior;
        g.sheen = s.sheen;
        g.sheen_roughness = s.sheen_roughness;
        g.subsurface = s.subsurface;
        g.opacity = s.opacity;
        g.thin_walled = s.thin_walled;
        g.is_default = 0;
        if (g.emission[0] != 0.0f || g.emission[1] != 0.0f || g.emission[2] != 0.0f) out.has_emission = 1;
        out.materials.push_back(g);
    }
    {
        GMaterial g{};  // BSDFClosure defaults (shading_types.h:18-42) with DefaultMaterial's overrides
        const float one[3] = {1.0f, 1.0f, 1.0f};
        std::memcpy(g.spec_color, one, sizeof one);
        std::memcpy(g.coat_color, one, sizeof one);
        std::memcpy(g.sheen_color, one, sizeof one);
        std::memcpy(g.trans_color, one, sizeof one);
        g.ss_color[0] = 0.7f;
        g.ss_color[1] = 0.2f;
        g.ss_color[2] = 0.2f;
        g.metallic = 0.0f;
        g.roughness = 1.0f;
        g.ior = 1.5f;
        g.specular = 0.5f;
        g.coat_ior = 1.5f;
        g.sheen_roughness = 0.3f;
        g.opacity = 1.0f;
        g.is_default = 1;
        out.materials.push_back(g);
    }

    // ---- instances (appendMeshInstance, scene.cpp:178-207) + TLAS
    std::vector<Box> instBounds(d.num_instances);
    std::vector<F3> instCentroid(d.num_instances);
    for (int i = 0; i < d.num_instances; ++i) {
        const hdr_instance_desc &s = d.instances[i];
        if (s.mesh < 0 || s.mesh >= d.num_meshes) return "instance references an invalid mesh";
        if (s.material >= d.num_materials) return "instance references an invalid material";
        if (s.light >= d.num_lights) return "instance references an invalid light";
        if (!finite_all(s.transform, 16)) return "non-finite instance transform";
        GInstance g{};
        std::memcpy(g.xf, s.transform, sizeof g.xf);
        double a[4][4], o[4][4];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) a[r][c] = s.transform[r * 4 + c];
        invert4(a, o);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) g.inv[r * 4 + c] = float(o[r][c]);
        g.root = geo.mesh_root[size_t(s.mesh)];
        g.has_params = s.material >= 0 ? 1 : 0;
        g.material = s.material >= 0 ? s.material : d.num_materials;
        g.light = s.light;
        if (s.light >= 0) out.has_mesh_lights = 1;
        out.instances.push_back(g);
        // _TransformBounds, scene.cpp:100-116
        const float *rv = &geo.mesh_range[size_t(s.mesh) * 6];
        Box r;
        r.mn = F3{rv[0], rv[1], rv[2]};
        r.mx = F3{rv[3], rv[4], rv[5]};
        Box b;
        if (r.empty()) {
            b = r;
        } else {
            const F3 mn = r.mn, mx = r.mx;
            const F3 corners[8] = {{mn.x, mn.y, mn.z}, {mx.x, mn.y, mn.z}, {mn.x, mx.y, mn.z}, {mx.x, mx.y, mn.z},
                                   {mn.x, mn.y, mx.z}, {mx.x, mn.y, mx.z}, {mn.x, mx.y, mx.z}, {mx.x, mx.y, mx.z}};
            for (const F3 &c : corners) b.extend(xform_point_h(s.transform, c));
        }
        instBounds[i] = b;
        instCentroid[i] = mulf(add(b.mn, b.mx), 0.5f);
    }
    if (d.num_instances > 0) {
        std::vector<int> idx(d.num_instances);
        for (int i = 0; i < d.num_instances; ++i) idx[i] = i;
        std::vector<GNode> tl;
        tl.reserve(size_t(d.num_instances) * 2);
        tl.push_back(GNode{});
        int tdepth = 0;
        // The reference's TLAS (below) fixes which of two equally distant hits on different
        // instances wins: the one in the earlier leaf slot. Its slots become the instances' tie
        // ranks; the tree the device traverses is a binned-SAH one (sah_tlas).
        // Scene::_SubdivideTLAS, scene.cpp:352-407 (children allocated before the leaf test)
        struct Rec {
            static void go(std::vector<GNode> &tl, std::vector<int> &idx, const std::vector<Box> &bounds,
                           const std::vector<F3> &cen, int nodeIdx, int start, int end, int dpt, int &maxd) {
                maxd = std::max(maxd, dpt);
                const int leftIdx = int(tl.size());
                tl.push_back(GNode{});
                tl.push_back(GNode{});
                Box b;
                for (int i = start; i < end; ++i) {
                    b.extend(bounds[idx[i]].mn);
                    b.extend(bounds[idx[i]].mx);
                }
                const int count = end - start;
                if (count <= 2) {
                    tl[nodeIdx] = make_node(b, -start - 1, count);
                    return;
                }
                tl[nodeIdx] = make_node(b, leftIdx, 0);
                const F3 size = b.size();
                int axis = 0;
                if (size.y > size.x) axis = 1;
                if (size.z > comp(size, axis)) axis = 2;
                const float splitPos = comp(b.mn, axis) + comp(size, axis) * 0.5f;
                int i = start, j = end - 1;
                while (i <= j) {
                    if (comp(cen[idx[i]], axis) < splitPos) {
                        ++i;
                    } else {
                        std::swap(idx[i], idx[j]);
                        --j;
                    }
                }
                if (i == start || i == end) i = start + count / 2;
                go(tl, idx, bounds, cen, leftIdx, start, i, dpt + 1, maxd);
                go(tl, idx, bounds, cen, leftIdx + 1, i, end, dpt + 1, maxd);
            }
        };
        Rec::go(tl, idx, instBounds, instCentroid, 0, 0, d.num_instances, 1, tdepth);
        std::vector<int> rank(size_t(d.num_instances));
        for (int slot = 0; slot < d.num_instances; ++slot) rank[size_t(idx[size_t(slot)])] = slot;
        for (int i = 0; i < d.num_instances; ++i) idx[size_t(i)] = i;
        tl.clear();
        tl.push_back(GNode{});
        tdepth = 0;
        sah_tlas(tl, idx, instBounds, instCentroid, 0, 0, d.num_instances, 1, tdepth);
        out.max_tlas_depth = tdepth;
        // odd base, as pad_to_odd places every tree
        const int base = geo.blas_nodes % 2 == 0 ? geo.blas_nodes + 1 : geo.blas_nodes;
        for (GNode n : tl) {
            if (n.left >= 0 && n.count == 0) n.left += base;
            sb.tlas.push_back(n);
        }
        sb.tlas_base = base;
        out.tlas_idx = idx;
        out.tlas_rank.resize(idx.size());
        for (size_t slot = 0; slot < idx.size(); ++slot) out.tlas_rank[slot] = rank[size_t(idx[slot])];
        out.tlas_root = base;
    }
    if (out.max_tlas_depth > kStackDepth - 2) return "BVH deeper than the traversal stack";
    if (out.tlas_idx.size() >= (size_t(1) << 28))
        return "more than 2^28 triangles or instances (leaf reference range of the traversal)";

    // ---- lights (per-sample constants of SampleLight)
    for (int i = 0; i < d.num_lights; ++i) {
        const hdr_light_desc &s = d.lights[i];
        if (s.type < HDR_LIGHT_DISTANT || s.type > HDR_LIGHT_POINT) return "unknown light type";
        GLight g{};
        g.type = s.type;
        g.is_delta = s.type == HDR_LIGHT_DISTANT ? 1 : 0;
        g.width = s.width;
        g.height = s.height;
        for (int k = 0; k < 16; ++k) g.m[k] = float(s.transform[k]);
        const float eff = s.intensity * std::exp2(s.exposure);  // LightParams::EffectiveIntensity
        g.eff_int = eff;
        for (int k = 0; k < 3; ++k) {
            g.raw_color[k] = s.color[k];
            g.color[k] = s.color[k] * eff;
        }
        g.area = s.width * s.height;
        g.inv_area = 1.f / g.area;
        const F3 down = normalizedf(xform_dir_h(g.m, F3{0.0f, 0.0f, -1.0f}));
        g.normal[0] = down.x;
        g.normal[1] = down.y;
        g.normal[2] = down.z;
        g.pos[0] = g.m[12];
        g.pos[1] = g.m[13];
        g.pos[2] = g.m[14];
        g.cone_angle = s.shaping_cone_angle;
        g.softness = s.shaping_cone_softness;
        g.cone_rad = s.shaping_cone_angle * float(M_PI / 180.0);
        g.cos_cone = std::cos(g.cone_rad);
        g.cos_inner = std::cos(g.cone_rad * (1.0f - s.shaping_cone_softness));
        g.dome = -1;
        if (s.type == HDR_LIGHT_DOME && s.dome_pixels && s.dome_width > 0 && s.dome_height > 0) {
            HostScene::Dome dm;
            dm.w = s.dome_width;
            dm.h = s.dome_height;
            const int w = dm.w, h = dm.h;
            dm.pixels.assign(s.dome_pixels, s.dome_pixels + size_t(w) * h * 3);
            // DomeLight::Prepare (dome_light.cpp:49-76) runs on the device at upload (the row and
            // column CDFs over every texel, dome_cdf_rows_kernel / dome_cdf_total_kernel); the host
            // keeps the per-row sin(theta) weights, evaluated with the host libm as the reference does.
            dm.sin_theta.resize(size_t(h));
            for (int y = 0; y < h; ++y)
                dm.sin_theta[size_t(y)] = float(std::sin(M_PI * double(float(y + 0.5f)) / double(float(h))));
            g.dome = int(out.domes.size());
            out.domes.push_back(std::move(dm));
        }
        out.lights.push_back(g);
    }
    out.environment = d.environment;
    out.env_dome = d.environment >= 0 ? out.lights[d.environment].dome : -1;
    out.num_materials = d.num_materials;
    return std::string();
}

void commit_state(StateBuild &sb, HostScene &out) {
    out.nodes.resize(size_t(out.blas_nodes));
    if (!sb.tlas.empty()) {
        pad_to_odd(out.nodes);
        out.nodes.insert(out.nodes.end(), sb.tlas.begin(), sb.tlas.end());
    }
    HostScene &s = sb.s;
    out.tlas_idx = std::move(s.tlas_idx);
    out.tlas_rank = std::move(s.tlas_rank);
    out.instances = std::move(s.instances);
    out.materials = std::move(s.materials);
    out.lights = std::move(s.lights);
    out.domes = std::move(s.domes);
    out.tlas_root = s.tlas_root;
    out.environment = s.environment;
    out.env_dome = s.env_dome;
    out.has_mesh_lights = s.has_mesh_lights;
    out.has_emission = s.has_emission;
    out.max_tlas_depth = s.max_tlas_depth;
    out.num_materials = s.num_materials;
}

}  // namespace
}  // namespace hdr
