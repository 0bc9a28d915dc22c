// Host-side SAH wide-BVH builder: binned SAH binary tree, greedy collapse to
// 8-wide nodes with octant-ordered child slots, 8-bit quantized child boxes.
// Replaces optixAccelBuild/optixAccelCompact of the reference
// (world/gas_manager.cpp:190-238, world/ias_manager.cpp:54-113).
#pragma once

#include <stdint.h>
#include <string>
#include <vector>

#include "pt_bvh.h"

namespace pupil {

struct BvhBuildOptions {
    int num_bins = 32;
    int threads = 0;  // 0 = hardware concurrency
};

struct BvhBuildStats {
    uint32_t binary_nodes = 0;
    uint32_t wide_nodes = 0;
    uint32_t leaf_tris = 0;
    uint32_t depth = 0;
    double sah_cost = 0.0;
    double build_seconds = 0.0;
};

// World center and radius of a sphere given as the unit sphere mapped by the
// 3x4 rows `m` (SetSphereAreaEmitter, world/emitter.cpp:224-241: center =
// M (0,0,0), radius = |M (1,0,0) - center|).  The same values feed the BVH
// leaves, the emitter table and the CPU checks.
inline void sphere_center_radius(const float *m, float out4[4]) {
    const f3 c = mk3(m[3], m[7], m[11]);
    const f3 px = mk3(m[0] + m[3], m[4] + m[7], m[8] + m[11]);
    out4[0] = c.x;
    out4[1] = c.y;
    out4[2] = c.z;
    out4[3] = length(c - px);
}

// positions: 4 floats per vertex (xyz used); indices: 4 uint32 per triangle
// (first three used); spheres4: world-space center xyz + radius per sphere
// (may be null when num_spheres is 0).  Triangle t gets primitive id t,
// sphere k gets kPrimSphere | k.  Returns false with `err` set on invalid
// input.
bool build_wide_bvh(const float *positions4, uint32_t num_vertices, const uint32_t *indices4, uint32_t num_tris,
                    const float *spheres4, uint32_t num_spheres, const BvhBuildOptions &opt,
                    std::vector<WideNode> &nodes, std::vector<LeafTri> &tris, BvhBuildStats &stats,
                    std::string &err);

}  // namespace pupil
