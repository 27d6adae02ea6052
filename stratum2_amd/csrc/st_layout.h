// st_layout.h — HBM layouts shared by the host builder and the gfx950 kernels.
//
// BVH: binary tree, one 64-byte node per inner node holding BOTH child boxes,
// so one node fetch (4 x 16-byte loads of one 64-byte line) tests two boxes:
//   f[0] = c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   f[1] = c0.lo.z c0.hi.z c1.lo.x c1.hi.x
//   f[2] = c1.lo.y c1.hi.y c1.lo.z c1.hi.z
//   f[3] = (int) child0, child1, count0, count1
// count == 0: child is an inner node index; count > 0: child is a leaf of
// `count` triangles starting at triangle index `child`; count < 0: empty child
// (only the root of a scene with at most 4 triangles has one).
//
// Triangles: 48 bytes, world space, in leaf order:
//   t[0] = v0.xyz, id (instance | primitive << 16, as InstanceData ids)
//   t[1] = v1.xyz, 0
//   t[2] = v2.xyz, 0
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/stratum2.h"

namespace st {

constexpr int kNodeBytes = 64;
constexpr int kTriBytes = 48;
constexpr int kMaxLeafTris = 4;
constexpr int kMaxDepth = 64;      // builder guarantee (median-split fallback beyond it)
constexpr int kLdsStack = 16;      // per-lane stack entries kept in LDS; the rest spill to HBM
constexpr int kEnvMaxLevels = 16;
constexpr int kQuant = 16384;      // gRadianceQuantizationFactor, testrenderer.slang:25

struct DevScene {
    const float4* nodes;
    const float4* tris;
    uint32_t node_count, tri_count;
    // reference-format scene buffers (shading, lights)
    const st_instance* inst;
    const st_transform* xf;
    const uint32_t* mat;
    const st_mesh_vertex_info* vi;
    const uint8_t* const* vb;      // device pointers of the vertex/index byte buffers
    const uint64_t* vb_size;
    uint32_t inst_count, mat_words, vi_count, vb_count;
    const uint32_t* light_inst;
    uint32_t light_count;
    // environment
    int has_env, env_has_image, env_levels;
    float env_emission[3];
    const float4* env_level[kEnvMaxLevels];
    int env_w[kEnvMaxLevels], env_h[kEnvMaxLevels];
    const float2* env_alias;       // per level-0 texel: {prob, alias index bits}
    const float* env_pdf;          // per level-0 texel: uv-space pdf
    float sphere[4];               // scene bounding sphere, TestRenderer.cpp:458-461
};

}  // namespace st
