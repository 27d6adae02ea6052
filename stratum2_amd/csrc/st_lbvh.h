// st_lbvh.h — device BVH builder: linear BVH (Morton order + radix tree)
// collapsed on the GPU into the 4-wide 48-byte traversal format of st_layout.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

struct LbvhResult {
    float4* nodes = nullptr;  // node_count 4-wide nodes (3 float4 each), hipMalloc'ed
    float4* tris = nullptr;   // tri_count triangles (3 float4 each) in leaf order, hipMalloc'ed
    uint32_t node_count = 0, tri_count = 0;
    int depth4 = 0;           // levels of the 4-wide tree
    double ms = 0;            // device time of the build
};

// tris: n world-space triangle records (v0.xyz|id, v1, v2; 48 bytes each) in
// device memory.  Returns freshly allocated node and triangle buffers (the
// caller frees them with hipFree).  Throws std::runtime_error on failure,
// including a tree deeper than the traversal stack allows.
LbvhResult build_lbvh(const float4* tris, uint32_t n, hipStream_t stream);

}  // namespace st
