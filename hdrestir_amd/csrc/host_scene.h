// host_scene.h -- host-side scene build: BLAS/TLAS construction and packing into the device layout.
#pragma once

#include <string>
#include <vector>

#include "../../include/hdrgpu.h"
#include "dscene.h"

namespace hdr {

struct HostScene {
    std::vector<GNode> nodes;        // BLAS nodes of all meshes, then TLAS nodes (from blas_nodes on)
    std::vector<GTri> tris;
    std::vector<GTriAttr> attrs;
    std::vector<int> tlas_idx;       // TLAS leaf slot -> instance (in the traversed tree's leaf order)
    std::vector<int> tlas_rank;      // TLAS leaf slot -> the instance's slot in the reference's tree
    std::vector<GInstance> instances;
    std::vector<GMaterial> materials; // num_materials records + the default material
    std::vector<GLight> lights;
    struct Dome {
        std::vector<float> pixels;     // w*h*3 texels
        std::vector<float> sin_theta;  // per-row weight of DomeLight::Prepare
        int w = 0, h = 0;              // the CDFs (h+1 rows, h*(w+1) columns) are built on the device
    };
    std::vector<Dome> domes;
    int tlas_root = -1;
    int environment = -1;
    int env_dome = -1;
    int has_mesh_lights = 0;
    int has_emission = 0;
    int max_blas_depth = 0;
    int max_tlas_depth = 0;
    // kept from the geometry build for rebuilding the instances and the TLAS
    int num_meshes = 0;
    int blas_nodes = 0;             // nodes.size() after the BLAS of every mesh
    std::vector<int> mesh_root;     // BLAS root node per mesh (-1: no valid triangle)
    std::vector<float> mesh_range;  // object-space bounds per mesh: min xyz, max xyz
    int num_materials = 0;          // records of the scene's own materials (the default follows)
};

// Validates the descriptor and builds everything; returns an empty string on success or an
// error message.
std::string build_host_scene(const hdr_scene_desc &desc, HostScene &out);

// Scene::BuildRenderState (scene.cpp:118-271) over a scene whose meshes are built: replaces the
// materials, instances, TLAS and lights of `out` with those of `desc` and keeps the BLAS of
// every mesh (desc.meshes is not read; desc.num_meshes must equal the built count). On an
// error `out` is left unchanged.
std::string rebuild_host_state(const hdr_scene_desc &desc, HostScene &out);

// GfMatrix4d / GfMatrix4f inverse by 2x2 sub-determinants in double (singular -> FLT_MAX scale).
void invert4(const double a[4][4], double out[4][4]);

}  // namespace hdr
