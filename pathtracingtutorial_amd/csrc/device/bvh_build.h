// GPU binned-SAH BVH build (host entry point).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace ptd {

struct BvhBuildOptions {
    int bins = 24;            // SAH buckets (reference: 24)
    int leaf_threshold = 4;   // leaf when count < leaf_threshold (reference: 4)
    int max_leaf = 0;         // >0: no SAH early-out leaf above this size
};

// Topology of a built BVH, in build (breadth-first) order with root 0 and
// slot 1 unused; children of a node are adjacent (left, left+1).
struct BvhHostTopology {
    std::vector<float> lo, hi;        // 3 floats per node
    std::vector<uint32_t> left;       // 0 for a leaf
    std::vector<uint32_t> start, count;
    std::vector<uint32_t> prim_order; // BVH slot -> original face
    uint32_t num_nodes = 0, depth = 0, max_leaf = 0;
};

struct BvhDeviceResult {
    float4* nodes = nullptr;      // 2 float4 per node, device order
    uint32_t num_nodes = 0;
    uint32_t* prim_order = nullptr;  // BVH slot -> original face
    uint32_t depth = 0, max_leaf = 0;
    float build_ms = 0.0f;
    std::vector<uint32_t> level_begin, level_count;  // nodes [begin, begin + count) per level, root first
};

// Builds over nfaces triangles given device vertex positions and original
// faces.  Allocates result->nodes and result->prim_order (hipMalloc, owned by
// the caller).  Optionally downloads the topology.  Returns false + *err.
bool gpu_build_bvh(const float4* d_vpos, const uint4* d_faces, uint32_t nfaces, const BvhBuildOptions& opt,
                   hipStream_t stream, BvhDeviceResult* result, BvhHostTopology* topo, std::string* err);

// Device-wide exclusive prefix sum of n uint32 (n <= 4,194,304); *d_total (may
// be null) receives the sum.  tmp must hold 2048 + 32 uint32.
void scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint32_t* d_tmp, uint32_t* d_total,
                        hipStream_t stream);

}  // namespace ptd
