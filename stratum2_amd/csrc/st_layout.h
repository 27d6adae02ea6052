// st_layout.h — HBM layouts shared by the host builder and the gfx950 kernels.
//
// BVH: 4-wide tree of 48-byte nodes (three 16-byte loads test four boxes):
//   f[0] = origin.xyz (float), ex | ey << 8 | ez << 16 | types(0..7) << 24
//   f[1] = qlo.x[4], qlo.y[4], qlo.z[4], qhi.x[4]          (uint8 each)
//   f[2] = qhi.y[4], qhi.z[4], inner | types(8..11) << 28, first triangle
// child box = origin + q * 2^(e-127) per axis (exactly conservative);
// child type 0 empty, 1..4 leaf of that many triangles (contiguous from the
// node's first triangle in child order), 7 inner (contiguous from `inner`).
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

constexpr int kTriBytes = 48;
constexpr int kMaxLeafTris = 4;
constexpr int kMaxDepth = 64;      // builder guarantee (median-split fallback beyond it)
constexpr int kMaxStack = 3 * kMaxDepth + 1;  // a 4-wide step pushes at most 3 entries
constexpr int kLdsStack = 16;      // per-lane stack entries kept in LDS; the rest spill to HBM
constexpr int kEnvMaxLevels = 16;
constexpr int kQuant = 16384;      // gRadianceQuantizationFactor, testrenderer.slang:25
constexpr int kMaxVolumes = 8;     // gVolumeCount, src/Shaders/compat/scene.h:13-15
constexpr int kMaxVolumeInstances = 8;
constexpr uint32_t kNoMedium = 0xffffu;  // INVALID_INSTANCE in the 16-bit medium field
constexpr int kMaxMediaSegments = 256;   // traceRay calls per scattering / visibility ray with media

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
    // emitter table: per light, 4 float4 {transform rows 0-2, {triangles, first
    // triangle, emission word, emission scale}}; per emissive triangle, 5 float4
    // {v0 | has-normals, v1 | n0.x, v2 | n0.y, n0.z n1, n2} in object space
    const float4* light_rec;
    const float4* light_tri;
    // environment
    int has_env, env_has_image, env_levels;
    float env_emission[3];
    const float4* env_level[kEnvMaxLevels];
    int env_w[kEnvMaxLevels], env_h[kEnvMaxLevels];
    const float2* env_alias;       // per level-0 texel: {prob, alias index bits}
    const float* env_pdf;          // per level-0 texel: uv-space pdf
    float env_jmax;                // largest in-texel jitter of the alias sampler (see st_scene_create)
    float sphere[4];               // scene bounding sphere, TestRenderer.cpp:458-461
    // media: volume instances (procedural boxes) and their NanoVDB float grids
    const st_transform* inv_xf;    // per instance, TransformData::inverse
    uint32_t vol_count, vinst_count;
    const uint8_t* vol[kMaxVolumes];
    uint64_t vol_root[kMaxVolumes];  // byte offset of each grid's root node
    uint32_t vinst[kMaxVolumeInstances];
    const st_volume_info* vinfo;   // VolumeInfo records, in the scene's order
    uint32_t vinfo_count;
    // differentiable rendering: the per-triangle albedo (3 floats per
    // flattened triangle, null: none) of Lambertian camera paths, and each
    // instance's first flattened triangle
    const float* tri_alb;
    const uint32_t* tri_base;
    // procedural sphere instances (SphereInstanceData, compat/scene.h:60-69): a
    // second 4-wide tree over their world boxes, whose leaf records carry the
    // instance index in the first record's w
    const float4* sph_nodes;
    const float4* sph_tris;
    uint32_t sph_node_count;
    // material images (Material::store's image indices 0, 2, 3): every level
    // of every used image in tex_data; per level {first float4, w, h, 0};
    // per image {first level, level count} (count 0: no material uses it)
    const float4* tex_data;
    const uint4* tex_lv;
    const uint2* tex_img;
    uint32_t tex_count;
    int has_tex;
    int has_alpha;  // some mesh material has an alpha mask (its triangles' records carry 1 in the second w)
};

}  // namespace st
