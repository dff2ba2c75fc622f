// Host-side launchers of the render and probe kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_scene.h"

namespace ptd {

// Device scratch of the plain megakernel: the tile queue's counters (one per
// XCD band, cleared by the launch) and the traversal stack entries beyond
// the LDS part of every thread of its persistent grid.
struct PlainScratch {
    uint32_t* tile_ctr;
    unsigned long long* spill;
    size_t spill_entries;
};
// The scratch a launch on this device needs (wide: the 4-wide BVH's stack).
hipError_t render_plain_scratch(bool wide, size_t* spill_entries, int* tile_counters);
// wide: walk the 4-wide BVH when the scene has one (else the binary tree).
hipError_t launch_render_plain(const DevScene& s, const RenderParams& p, float4* film_new, float4* film_acc,
                               DevCounters* ctr, bool count, bool wide, const PlainScratch& scratch,
                               hipStream_t stream);
hipError_t launch_probe_bsdf(const pt_material* d_mats, int n, const float* d_dirs, const int32_t* d_entering,
                             const float* d_u, float* d_out, hipStream_t stream);
hipError_t launch_probe_sampler(int type, uint32_t seed, const int32_t* d_xyi, int n, int dims, uint32_t* d_u,
                                float* d_f, hipStream_t stream);
// wide: traverse the 4-wide BVH (s.wnodes) instead of the binary one.
hipError_t launch_trace_rays(const DevScene& s, const float* d_rays8, int n, int any, bool wide, float* d_t,
                             int32_t* d_face, float* d_bary, hipStream_t stream);
// Prepares the triangle records of the traversal layout (dev_scene.h) and the
// BVH-ordered faces from the original faces and the build's slot order.
hipError_t launch_emit_tris(const float4* d_vpos, const uint4* d_faces_orig, const uint32_t* d_prim_order,
                            const float4* d_nodes, uint32_t num_nodes, uint32_t nfaces, float4* d_tris,
                            uint4* d_faces_bvh, hipStream_t stream);

hipError_t launch_copy_probe(const float4* src, float4* dst, size_t n, hipStream_t stream);
// DevScene::shade_rec (4 float4 per triangle) from the BVH-ordered faces and
// the vertex normals and uvs.
hipError_t launch_shade_records(const uint4* d_faces_bvh, const float4* d_vnrm, const float2* d_vuv, uint32_t nfaces,
                                float4* d_out, hipStream_t stream);
// Counts (into *d_out, which the caller zeroes) the floats of magnitude in
// [2^-126, 2^126) whose recip_rn differs from IEEE 1.0f / x.
hipError_t launch_probe_reciprocal(unsigned long long* d_out, hipStream_t stream);

}  // namespace ptd
