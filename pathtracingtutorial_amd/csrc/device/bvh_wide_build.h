// Binary -> 4-wide BVH collapse on the GPU (host entry point).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace ptd {

struct WideDeviceResult {
    float4* nodes = nullptr;  // 4 float4 per wide node (hipMalloc, owned by the caller)
    uint32_t num_nodes = 0;
    uint32_t max_stack = 0;   // traversal stack entries a ray can need
    uint32_t depth = 0;       // wide levels
    bool finite = true;       // false: a child box had non-finite bounds (nodes is null)
};

// Collapses the binary tree the GPU builder made (2 float4 per node in device
// order: {lo.xyz, word0}, {hi.xyz, word1}; root 0, slot 1 unused; an inner
// node's children are word0 and word1, a leaf's words are kLeafFlag | first
// and the count) into the 4-wide layout of host/bvh_wide.h.  level_begin /
// level_count list the nodes of each binary level ([begin, begin + count),
// root level first).  The cut the surface-area program picks, the
// breadth-first node numbering and the quantized boxes are those of the host
// collapse (ptamd::collapse_wide_bvh) on the same tree, bit for bit.
// Synchronous on `stream`.  Returns false + *err on a device error or a
// malformed tree; non-finite bounds return true with finite = false.
bool gpu_collapse_wide(const float4* nodes, uint32_t num_nodes, const std::vector<uint32_t>& level_begin,
                       const std::vector<uint32_t>& level_count, hipStream_t stream, WideDeviceResult* out,
                       std::string* err);

}  // namespace ptd
