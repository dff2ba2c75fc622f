// Device-resident scene layout (HBM) and the per-launch parameter block.
//
//  nodes   : BVH, 32 B per node as two float4 {bmin.xyz, data0}, {bmax.xyz, data1}
//            (the reference's BVHNode layout, accel/bvh_common.h:14-65), in the
//            device order: root at 0, slot 1 unused, siblings adjacent from 2 on,
//            so one 64 B load fetches both children of a node.
//            internal: data0 = index of the left child (right = left + 1)
//            leaf    : data0 = 0x80000000 | first triangle, data1 = count
//  wnodes  : the same tree collapsed to 4 children per node with quantized
//            child boxes, 64 B per node (layout in dev_traverse.h).
//  tris    : 48 B per triangle in BVH order, three float4
//            {p0.xyz, face | last<<31}, {p1-p0, 0}, {p2-p0, 0}; bit 31 of the
//            first word marks the last triangle of a leaf; face is the original
//            face index.
//  faces   : uint4 {v0, v1, v2, material} per triangle in BVH order.
//  vpos/vnrm/vuv : vertex position (w=0), normal (w=0) and uv.
//  vtan    : vertex tangent xyz + handedness (only when a normal map is used).
#pragma once

#include <stdint.h>

#include "pt_amd.h"

namespace ptd {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kStackSize = 32;  // traversal stack entries; the builder caps depth at this
constexpr int kWideStack = 64;  // the same for the 4-wide BVH; wider trees that could need more are not used

struct TexDesc {
    uint32_t offset;  // first texel in the texel pool
    int32_t w, h;
    int32_t wrap_s, wrap_t;  // 0 wrap, 1 clamp, 2 mirror
    int32_t linear;
    int32_t _pad[2];
};

struct DevScene {
    const float4* nodes;
    const float4* wnodes;        // 4-wide BVH, 4 float4 per node (null when not built)
    const float4* tris;
    const uint4* faces;
    const float4* vpos;
    const float4* vnrm;
    const float2* vuv;
    const float4* vtan;
    const pt_material* materials;
    const TexDesc* textures;
    const float4* texels;
    // lights
    const uint4* emissive;       // {tri (BVH order), pmf bits, cdf bits, 0}
    const float* tri_pmf;        // per BVH-ordered face
    const float2* tri_alias;     // {prob, alias bits}
    const float* env_pmf;
    const float2* env_alias;
    const float4* env;           // env texels, RGBA
    const pt_punctual_light* lights;
    // per BVH-ordered triangle, the vertex normals, uvs and material in one
    // 64 B record: {n0.xyz, n1.x}, {n1.yz, n2.xy}, {n2.z, t0.xy, t1.x},
    // {t1.y, t2.xy, material bits} (null: read the face and the vertices)
    const float4* shade_rec;
    uint32_t num_nodes, num_faces, num_materials, num_textures;
    uint32_t num_emissive, num_env_texels, num_lights, num_wnodes;
    int32_t env_w, env_h, env_wrap_s, env_wrap_t;
};

struct RenderParams {
    float cam_pos[3], cam_fwd[3], cam_right[3], cam_up[3];
    float tan_half_fov, aspect, aperture, focal_distance;
    int32_t width, height, spf, max_depth;
    int32_t sampler;  // 0 PCG, 1 Sobol
    uint32_t seed;
    uint32_t frame_count;
    int32_t tiles_x;  // 16x16 tiles per image row
};

struct DevCounters {
    unsigned long long closest_rays, shadow_rays, node_visits, tri_tests, paths, stack_overflows;
    unsigned long long shadow_node_visits, shadow_tri_tests;  // the any-hit share of node_visits / tri_tests
    // the wavefront shading kernel's events (pt_counters)
    unsigned long long shade_items, shade_env_lookups, shade_hits, shade_light_samples, shade_shadow_entries,
        shade_continuations, shade_radiance_adds;
};

}  // namespace ptd
