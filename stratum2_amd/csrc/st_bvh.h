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

// 4-wide node, 48 bytes = three 16-byte loads for four child boxes:
//   o[3]        quantisation origin (the node box minimum), float
//   x0          exponent bytes ex | ey << 8 | ez << 16 (scale_a = 2^(e_a - 127)),
//               child types 0..7 in bits 24..31 (children 0, 1 and bits 0-1 of 2)
//   q[24]       qlo.x[4] qlo.y[4] qlo.z[4] qhi.x[4] qhi.y[4] qhi.z[4] (uint8)
//   inner       first inner child node (bits 0..27), child type bits 8..11 (bits 28..31)
//   tri         first triangle of the leaf children (leaf triangles are contiguous
//               in child order)
// child type: 0 empty, 1..4 leaf with that many triangles, 7 inner node.
// Decoded boxes o + q * scale are conservative in float arithmetic.
struct Node4Out {
    float o[3];
    uint32_t x0;
    uint8_t q[24];
    uint32_t inner;
    uint32_t tri;
};
static_assert(sizeof(Node4Out) == 48, "4-wide node is 48 bytes");

struct BvhOut {
    std::vector<NodeOut> nodes;     // binary tree (build intermediate)
    std::vector<TriOut> tris;       // triangles in binary-leaf order
    std::vector<Node4Out> nodes4;   // traversal tree
    std::vector<TriOut> tris4;      // triangles in 4-wide leaf order
    int depth = 0;
    int depth4 = 0;
};

// verts: 9 floats (world-space v0, v1, v2) per triangle; ids: instance|primitive<<16.
BvhOut build_bvh(const float* verts, const uint32_t* ids, uint32_t n);

}  // namespace st
