// st_bvh.h — host BVH builder interface (see st_bvh.cpp).
#pragma once

#include <stdint.h>

#include <vector>

namespace st {

constexpr int kBvhMaxLeaf = 4;
constexpr int kBvhMaxDepth = 64;

struct NodeOut {  // 64 bytes, the layout described in st_layout.h
    float f[12];
    int32_t i[4];
};
struct TriOut {  // 48 bytes
    float f[12];
};
static_assert(sizeof(NodeOut) == 64, "node must be one 64-byte line");
static_assert(sizeof(TriOut) == 48, "triangle record is 48 bytes");

struct BvhOut {
    std::vector<NodeOut> nodes;
    std::vector<TriOut> tris;
    int depth = 0;
};

// verts: 9 floats (world-space v0, v1, v2) per triangle; ids: instance|primitive<<16.
BvhOut build_bvh(const float* verts, const uint32_t* ids, uint32_t n);

}  // namespace st
