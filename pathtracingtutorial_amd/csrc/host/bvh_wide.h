// 4-wide BVH with quantized child boxes, collapsed from the binary BVH.
//
// A wide node is 64 B (four float4), one cache line, holding up to four
// children:
//   n[0] = {P.x, P.y, P.z, S.x}
//   n[1] = {S.y, S.z, qlo.x, qlo.y}
//   n[2] = {qlo.z, qhi.x, qhi.y, qhi.z}
//   n[3] = {child word 0, 1, 2, 3}
// P is the node's box corner, S a power of two per axis; qlo.a / qhi.a hold
// one byte per child (byte i = child i).  Child i's box on axis a is
// [fma(qlo_i, S.a, P.a), fma(qhi_i, S.a, P.a)] in single precision, and the
// bytes are chosen so that this decoded box contains the child's exact box:
// a ray-box test on it never rejects a box that the exact box's test would
// accept (the slab arithmetic is monotone in the planes).  Child words:
// a wide-node index, kLeafFlag | first triangle for a leaf (the triangle
// records mark a leaf's last triangle), or kWideEmpty for an unused slot.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace ptamd {

constexpr uint32_t kWideEmpty = 0xFFFFFFFFu;

struct WideBvh {
    std::vector<float> words;  // 16 per node
    uint32_t num_nodes = 0;
    uint32_t max_stack = 0;    // traversal stack entries a ray can need
    uint32_t depth = 0;        // wide levels
};

// Collapses a binary BVH given in the reference's node layout
// (accel/bvh_common.h:16-24: bmin[3], data0, bmax[3], data1; internal nodes
// hold both child indices, leaves kLeafFlag | first and the count; root at
// 0).  Which binary nodes become wide nodes is chosen by a dynamic program
// that minimises the summed surface area of the wide nodes (the surface area
// heuristic; binary leaves are kept as they are); each wide node's children
// are the up to four descendants of its binary node that the program picked,
// in the binary tree's left-to-right order.  Nodes are numbered breadth
// first.  Returns false with *err on a malformed tree or non-finite bounds.
bool collapse_wide_bvh(const float* nodes8, uint32_t num_nodes, WideBvh* out, std::string* err);

}  // namespace ptamd
