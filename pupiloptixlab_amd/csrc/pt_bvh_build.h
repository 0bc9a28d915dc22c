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

// positions: 4 floats per vertex (xyz used); indices: 4 uint32 per triangle
// (first three used).  Returns false with `err` set on invalid input.
bool build_wide_bvh(const float *positions4, uint32_t num_vertices, const uint32_t *indices4, uint32_t num_tris,
                    const BvhBuildOptions &opt, std::vector<WideNode> &nodes, std::vector<LeafTri> &tris,
                    BvhBuildStats &stats, std::string &err);

}  // namespace pupil
