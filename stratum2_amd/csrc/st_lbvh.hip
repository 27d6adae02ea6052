// st_lbvh.hip — GPU BVH builder for scene updates.
//
// The reference rebuilds its top-level acceleration structure on every scene
// update (src/App/Scene.cpp:726-754, driver builds with ePreferFastTrace).
// Here the whole triangle BVH is rebuilt on the device:
//   1. centroid bounds (per-wave reduction, ordered-int atomics)
//   2. 30-bit Morton codes of the triangle centroids, radix-sorted (rocPRIM)
//   3. binary radix tree over the sorted codes (Karras, HPG 2012), ties in
//      the codes broken by the sorted index
//   4. bottom-up boxes (one thread per leaf, the second arrival at a node
//      merges its children)
//   5. top-down collapse into the 4-wide 48-byte node format of st_layout.h,
//      level by level: each node takes its binary children and repeatedly
//      opens the largest-area child that holds more than 4 triangles, as the
//      host collapse does; subtrees of at most 4 triangles become leaves.
// Child boxes are quantised exactly as the host builder does (conservative in
// float arithmetic), so the traversal kernels run unchanged on either tree.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "st_layout.h"
#include "st_lbvh.h"

namespace st {
namespace {

#define LB_CHECK(x)                                                                                           \
    do {                                                                                                      \
        hipError_t e_ = (x);                                                                                  \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));       \
    } while (0)

struct Buf {
    void* p = nullptr;
    Buf() = default;
    Buf(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t bytes) {
        LB_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
    void* release() {
        void* r = p;
        p = nullptr;
        return r;
    }
};

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNone = 0xffffffffu;

// float <-> order-preserving uint (for atomic min/max on floats)
__device__ __forceinline__ uint32_t f2o(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) { return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o); }

__global__ void k_init(uint32_t* bounds, uint32_t* counters) {
    if (threadIdx.x < 3) {
        bounds[threadIdx.x] = 0xffffffffu;  // min
        bounds[3 + threadIdx.x] = 0;        // max
    }
    if (threadIdx.x < 4) counters[threadIdx.x] = 0;
}

__device__ __forceinline__ void tri_box(const float4* t, uint32_t i, float lo[3], float hi[3]) {
    const float4 a = t[3 * (size_t)i], b = t[3 * (size_t)i + 1], c = t[3 * (size_t)i + 2];
    lo[0] = fminf(fminf(a.x, b.x), c.x);
    lo[1] = fminf(fminf(a.y, b.y), c.y);
    lo[2] = fminf(fminf(a.z, b.z), c.z);
    hi[0] = fmaxf(fmaxf(a.x, b.x), c.x);
    hi[1] = fmaxf(fmaxf(a.y, b.y), c.y);
    hi[2] = fmaxf(fmaxf(a.z, b.z), c.z);
}

__global__ void k_centroid_bounds(const float4* tris, uint32_t n, uint32_t* bounds) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float lo[3], hi[3];
        tri_box(tris, i, lo, hi);
        for (int k = 0; k < 3; k++) {
            const uint32_t o = f2o(0.5f * (lo[k] + hi[k]));
            mn[k] = o < mn[k] ? o : mn[k];
            mx[k] = o > mx[k] ? o : mx[k];
        }
    }
    for (int k = 0; k < 3; k++) {
        uint32_t a = mn[k], b = mx[k];
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t oa = __shfl_xor(a, off), ob = __shfl_xor(b, off);
            a = oa < a ? oa : a;
            b = ob > b ? ob : b;
        }
        if (__lane_id() == 0) {
            atomicMin(&bounds[k], a);
            atomicMax(&bounds[3 + k], b);
        }
    }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {  // 10 bits -> every third bit
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(const float4* tris, uint32_t n, const uint32_t* bounds, uint32_t* keys, uint32_t* vals) {
    float lo[3], ext[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = o2f(bounds[k]);
        ext[k] = o2f(bounds[3 + k]) - lo[k];
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float tlo[3], thi[3];
        tri_box(tris, i, tlo, thi);
        uint32_t q[3];
        for (int k = 0; k < 3; k++) {
            const float c = 0.5f * (tlo[k] + thi[k]);
            const float u = ext[k] > 0 ? (c - lo[k]) / ext[k] : 0.0f;
            q[k] = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
        }
        keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
        vals[i] = i;
    }
}

// Karras 2012: internal node i of the radix tree over n sorted keys.
__device__ __forceinline__ int lb_delta(const uint32_t* keys, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    const uint32_t ka = keys[a], kb = keys[b];
    if (ka == kb) return 32 + __clz((uint32_t)a ^ (uint32_t)b);
    return __clz(ka ^ kb);
}

__global__ void k_radix_tree(const uint32_t* keys, int n, uint32_t* left, uint32_t* right, uint32_t* first, uint32_t* count,
                             uint32_t* node_parent, uint32_t* leaf_parent) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const int d = (lb_delta(keys, n, i, i + 1) - lb_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = lb_delta(keys, n, i, i - d);
        int lmax = 2;
        while (lb_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (lb_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = lb_delta(keys, n, i, j);
        int s = 0, t = l;
        do {
            t = (t + 1) / 2;
            if (lb_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + (d < 0 ? -1 : 0);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        if (lo == gamma) {
            left[i] = (uint32_t)gamma | kLeafBit;
            leaf_parent[gamma] = (uint32_t)i;
        } else {
            left[i] = (uint32_t)gamma;
            node_parent[gamma] = (uint32_t)i;
        }
        if (hi == gamma + 1) {
            right[i] = (uint32_t)(gamma + 1) | kLeafBit;
            leaf_parent[gamma + 1] = (uint32_t)i;
        } else {
            right[i] = (uint32_t)(gamma + 1);
            node_parent[gamma + 1] = (uint32_t)i;
        }
        first[i] = (uint32_t)lo;
        count[i] = (uint32_t)(hi - lo + 1);
    }
}

// Leaf boxes (sorted order), rounded outward by one ulp as the host builder does.
__global__ void k_leaf_boxes(const float4* tris, const uint32_t* vals, uint32_t n, float* leaf_box) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        float lo[3], hi[3];
        tri_box(tris, vals[k], lo, hi);
        for (int a = 0; a < 3; a++) {
            leaf_box[6 * (size_t)k + a] = nextafterf(lo[a], -INFINITY);
            leaf_box[6 * (size_t)k + 3 + a] = nextafterf(hi[a], INFINITY);
        }
    }
}

__device__ __forceinline__ float ld_coherent(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Bottom-up boxes: the second thread to reach a node merges both children.
__global__ void k_node_boxes(uint32_t n, const uint32_t* left, const uint32_t* right, const uint32_t* node_parent,
                             const uint32_t* leaf_parent, const float* leaf_box, float* node_box, uint32_t* flags,
                             uint32_t* err) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t p = leaf_parent[k];
        // a radix tree over 30-bit codes plus index bits is at most 62 levels deep;
        // the bound only guards against a malformed tree
        for (int steps = 0; p != kNone; steps++) {
            if (steps > 96 || p >= n - 1) {
                atomicOr(err, 8u);
                break;
            }
            __threadfence();
            if (atomicAdd(&flags[p], 1u) == 0) break;  // the sibling is not done yet
            __threadfence();
            float b[6];
            for (int a = 0; a < 3; a++) {
                b[a] = INFINITY;
                b[3 + a] = -INFINITY;
            }
            const uint32_t ch[2] = {left[p], right[p]};
            if ((ch[0] & ~kLeafBit) >= n || (ch[1] & ~kLeafBit) >= n) {
                atomicOr(err, 8u);
                break;
            }
            for (int c = 0; c < 2; c++) {
                const float* src = (ch[c] & kLeafBit) ? leaf_box + 6 * (size_t)(ch[c] & ~kLeafBit) : node_box + 6 * (size_t)ch[c];
                for (int a = 0; a < 3; a++) {
                    b[a] = fminf(b[a], ld_coherent(src + a));
                    b[3 + a] = fmaxf(b[3 + a], ld_coherent(src + 3 + a));
                }
            }
            for (int a = 0; a < 6; a++) node_box[6 * (size_t)p + a] = b[a];
            p = node_parent[p];
        }
    }
}

// decoded coordinate o + q * 2^(e-127), as the traversal computes it
__device__ __forceinline__ float lb_decode(float o, uint32_t q, int e) { return o + (float)q * __uint_as_float((uint32_t)e << 23); }

struct Kid {
    float lo[3], hi[3];
    uint32_t ref;    // internal node index, or kLeafBit | sorted position
    uint32_t first;  // first sorted triangle
    uint32_t count;  // triangles below
};

__device__ __forceinline__ float kid_area(const Kid& k) {
    float d[3];
    for (int a = 0; a < 3; a++) d[a] = fmaxf(0.0f, k.hi[a] - k.lo[a]);
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

__device__ __forceinline__ Kid make_kid(uint32_t ref, const uint32_t* first, const uint32_t* count, const float* leaf_box,
                                        const float* node_box, uint32_t n) {
    Kid k;
    if ((ref & ~kLeafBit) >= n) ref = kLeafBit;  // malformed tree: flagged by the caller's checks
    k.ref = ref;
    const float* b;
    if (ref & kLeafBit) {
        k.first = ref & ~kLeafBit;
        k.count = 1;
        b = leaf_box + 6 * (size_t)k.first;
    } else {
        k.first = first[ref];
        k.count = count[ref];
        b = node_box + 6 * (size_t)ref;
    }
    for (int a = 0; a < 3; a++) {
        k.lo[a] = b[a];
        k.hi[a] = b[3 + a];
    }
    return k;
}

struct Item {
    uint32_t ref, slot;
};

// One level of the top-down collapse.  counters: [0] nodes, [1] triangles,
// [2] next-level items, [3] error flags.
__global__ void k_collapse(const Item* items, uint32_t nitems, Item* next, const uint32_t* left, const uint32_t* right,
                           const uint32_t* first, const uint32_t* count, const float* leaf_box, const float* node_box,
                           const uint32_t* vals, const float4* tris_in, uint32_t n, float4* nodes_out, float4* tris_out,
                           uint32_t* counters) {
    for (uint32_t it = blockIdx.x * blockDim.x + threadIdx.x; it < nitems; it += gridDim.x * blockDim.x) {
        const Item job = items[it];
        // guards against a malformed tree: every index below stays in its buffer
        if (job.slot > n || (job.ref != kNone && job.ref >= n - 1)) {
            atomicOr(&counters[3], 16u);
            continue;
        }
        Kid kids[4];
        int nk = 0;
        if (job.ref == kNone) {  // root of a tree of at most 4 triangles: one leaf child
            Kid k;
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t t = 0; t < n; t++)
                for (int a = 0; a < 3; a++) {
                    lo[a] = fminf(lo[a], leaf_box[6 * (size_t)t + a]);
                    hi[a] = fmaxf(hi[a], leaf_box[6 * (size_t)t + 3 + a]);
                }
            for (int a = 0; a < 3; a++) {
                k.lo[a] = lo[a];
                k.hi[a] = hi[a];
            }
            k.ref = kNone;
            k.first = 0;
            k.count = n;
            kids[nk++] = k;
        } else {
            kids[nk++] = make_kid(left[job.ref], first, count, leaf_box, node_box, n);
            kids[nk++] = make_kid(right[job.ref], first, count, leaf_box, node_box, n);
        }
        // open the largest-area child with more than kMaxLeafTris triangles until there are four
        while (nk < 4) {
            int best = -1;
            float ba = 0;
            for (int i = 0; i < nk; i++) {
                if ((kids[i].ref & kLeafBit) || kids[i].ref == kNone || kids[i].count <= (uint32_t)kMaxLeafTris) continue;
                const float a = kid_area(kids[i]);
                if (best < 0 || a > ba) {
                    best = i;
                    ba = a;
                }
            }
            if (best < 0) break;
            const uint32_t r = kids[best].ref;
            if (r >= n - 1) {
                atomicOr(&counters[3], 16u);
                break;
            }
            for (int i = nk - 1; i > best; i--) kids[i + 1] = kids[i];
            kids[best] = make_kid(left[r], first, count, leaf_box, node_box, n);
            kids[best + 1] = make_kid(right[r], first, count, leaf_box, node_box, n);
            nk++;
        }
        // quantisation frame (same arithmetic as the host collapse)
        float L[3], U[3];
        int e[3];
        for (int a = 0; a < 3; a++) {
            L[a] = INFINITY;
            U[a] = -INFINITY;
            for (int i = 0; i < nk; i++) {
                L[a] = fminf(L[a], kids[i].lo[a]);
                U[a] = fmaxf(U[a], kids[i].hi[a]);
            }
            const double ext = (double)U[a] - (double)L[a];
            int k = ext > 0 ? (int)ceil(log2(ext / 255.0)) : -100;
            k = k < -126 ? -126 : (k > 127 ? 127 : k);
            while (k < 127 && lb_decode(L[a], 255, k + 127) < U[a]) k++;
            e[a] = k + 127;
            if (lb_decode(L[a], 255, e[a]) < U[a]) atomicOr(&counters[3], 1u);
        }
        uint32_t n_inner = 0, n_leaf_tris = 0;
        for (int i = 0; i < nk; i++) {
            const bool inner = !(kids[i].ref & kLeafBit) && kids[i].ref != kNone && kids[i].count > (uint32_t)kMaxLeafTris;
            if (inner) n_inner++;
            else n_leaf_tris += kids[i].count;
        }
        const uint32_t base_inner = n_inner ? atomicAdd(&counters[0], n_inner) : 0;
        const uint32_t base_tri = atomicAdd(&counters[1], n_leaf_tris);
        if (base_inner + n_inner >= (1u << 28) || base_inner + n_inner > n + 1 || base_tri + n_leaf_tris > n) {
            atomicOr(&counters[3], 2u);
            continue;
        }
        uint8_t q[24];
        uint32_t types = 0, r_inner = 0, r_tri = 0;
        for (int c = 0; c < 4; c++) {
            uint32_t qlo[3] = {255, 255, 255}, qhi[3] = {0, 0, 0}, type = 0;
            if (c < nk) {
                const Kid& k = kids[c];
                for (int a = 0; a < 3; a++) {
                    const float sc = __uint_as_float((uint32_t)e[a] << 23);
                    const double ql = floor(((double)k.lo[a] - L[a]) / sc), qh = ceil(((double)k.hi[a] - L[a]) / sc);
                    qlo[a] = (uint32_t)fmax(0.0, fmin(255.0, ql));
                    qhi[a] = (uint32_t)fmax(0.0, fmin(255.0, qh));
                    while (qlo[a] > 0 && lb_decode(L[a], qlo[a], e[a]) > k.lo[a]) qlo[a]--;
                    while (qhi[a] < 255 && lb_decode(L[a], qhi[a], e[a]) < k.hi[a]) qhi[a]++;
                    if (lb_decode(L[a], qlo[a], e[a]) > k.lo[a] || lb_decode(L[a], qhi[a], e[a]) < k.hi[a])
                        atomicOr(&counters[3], 4u);
                }
                const bool inner = !(k.ref & kLeafBit) && k.ref != kNone && k.count > (uint32_t)kMaxLeafTris;
                if (inner) {
                    type = 7;
                    const uint32_t slot = base_inner + r_inner++;
                    next[atomicAdd(&counters[2], 1u)] = Item{k.ref, slot};
                } else {
                    type = k.count;
                    for (uint32_t t = 0; t < k.count; t++) {
                        const uint32_t src = vals[(k.first + t) % n] % n, dst = base_tri + r_tri++;
                        tris_out[3 * (size_t)dst] = tris_in[3 * (size_t)src];
                        tris_out[3 * (size_t)dst + 1] = tris_in[3 * (size_t)src + 1];
                        tris_out[3 * (size_t)dst + 2] = tris_in[3 * (size_t)src + 2];
                    }
                }
            }
            for (int a = 0; a < 3; a++) {
                q[4 * a + c] = (uint8_t)qlo[a];
                q[12 + 4 * a + c] = (uint8_t)qhi[a];
            }
            types |= type << (3 * c);
        }
        uint32_t qw[6];
        for (int w = 0; w < 6; w++)
            qw[w] = (uint32_t)q[4 * w] | ((uint32_t)q[4 * w + 1] << 8) | ((uint32_t)q[4 * w + 2] << 16) | ((uint32_t)q[4 * w + 3] << 24);
        const uint32_t x0 = (uint32_t)e[0] | ((uint32_t)e[1] << 8) | ((uint32_t)e[2] << 16) | ((types & 0xffu) << 24);
        const uint32_t inner_w = base_inner | ((types >> 8) << 28);
        float4* o = nodes_out + 3 * (size_t)job.slot;
        o[0] = make_float4(L[0], L[1], L[2], __uint_as_float(x0));
        o[1] = make_float4(__uint_as_float(qw[0]), __uint_as_float(qw[1]), __uint_as_float(qw[2]), __uint_as_float(qw[3]));
        o[2] = make_float4(__uint_as_float(qw[4]), __uint_as_float(qw[5]), __uint_as_float(inner_w), __uint_as_float(base_tri));
    }
}

int grid_for(uint32_t n, int block) { return (int)std::min<uint64_t>((n + block - 1) / block, 4096); }

}  // namespace

LbvhResult build_lbvh(const float4* tris, uint32_t n, hipStream_t s) {
    LbvhResult r;
    if (n == 0) return r;
    if (n >= (1u << 30)) throw std::runtime_error("too many triangles for the device BVH builder");
    hipEvent_t e0, e1;
    LB_CHECK(hipEventCreate(&e0));
    LB_CHECK(hipEventCreate(&e1));
    LB_CHECK(hipEventRecord(e0, s));
    Buf bounds, counters, keys, keys2, vals, vals2, left, right, first, count, nparent, lparent, lbox, nbox, flags, items0, items1;
    Buf nodes, tout, temp;
    bounds.alloc(6 * 4);
    counters.alloc(4 * 4);
    keys.alloc(4ull * n);
    keys2.alloc(4ull * n);
    vals.alloc(4ull * n);
    vals2.alloc(4ull * n);
    left.alloc(4ull * n);
    right.alloc(4ull * n);
    first.alloc(4ull * n);
    count.alloc(4ull * n);
    nparent.alloc(4ull * n);
    lparent.alloc(4ull * n);
    lbox.alloc(24ull * n);
    nbox.alloc(24ull * n);
    flags.alloc(4ull * n);
    items0.alloc(8ull * n + 8);
    items1.alloc(8ull * n + 8);
    nodes.alloc(48ull * (n + 1));
    tout.alloc(48ull * n);
    const int B = 256;
    hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, bounds.as<uint32_t>(), counters.as<uint32_t>());
    hipLaunchKernelGGL(k_centroid_bounds, dim3(grid_for(n, B)), dim3(B), 0, s, tris, n, bounds.as<uint32_t>());
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n, B)), dim3(B), 0, s, tris, n, (const uint32_t*)bounds.as<uint32_t>(),
                       keys.as<uint32_t>(), vals.as<uint32_t>());
    size_t temp_bytes = 0;
    LB_CHECK(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<uint32_t>(),
                                       vals2.as<uint32_t>(), n, 0, 30, s));
    temp.alloc(temp_bytes);
    LB_CHECK(rocprim::radix_sort_pairs(temp.p, temp_bytes, keys.as<uint32_t>(), keys2.as<uint32_t>(), vals.as<uint32_t>(),
                                       vals2.as<uint32_t>(), n, 0, 30, s));
    LB_CHECK(hipMemsetAsync(nparent.p, 0xff, 4ull * n, s));
    LB_CHECK(hipMemsetAsync(lparent.p, 0xff, 4ull * n, s));
    LB_CHECK(hipMemsetAsync(flags.p, 0, 4ull * n, s));
    if (n > 1)
        hipLaunchKernelGGL(k_radix_tree, dim3(grid_for(n, B)), dim3(B), 0, s, (const uint32_t*)keys2.as<uint32_t>(), (int)n,
                           left.as<uint32_t>(), right.as<uint32_t>(), first.as<uint32_t>(), count.as<uint32_t>(),
                           nparent.as<uint32_t>(), lparent.as<uint32_t>());
    hipLaunchKernelGGL(k_leaf_boxes, dim3(grid_for(n, B)), dim3(B), 0, s, tris, (const uint32_t*)vals2.as<uint32_t>(), n,
                       lbox.as<float>());
    if (n > 1)
        hipLaunchKernelGGL(k_node_boxes, dim3(grid_for(n, B)), dim3(B), 0, s, n, (const uint32_t*)left.as<uint32_t>(),
                           (const uint32_t*)right.as<uint32_t>(), (const uint32_t*)nparent.as<uint32_t>(),
                           (const uint32_t*)lparent.as<uint32_t>(), (const float*)lbox.as<float>(), nbox.as<float>(),
                           flags.as<uint32_t>(), counters.as<uint32_t>() + 3);
    LB_CHECK(hipGetLastError());
    // top-down collapse, one launch per level of the 4-wide tree
    const Item root{n > (uint32_t)kMaxLeafTris ? 0u : kNone, 0u};
    LB_CHECK(hipMemcpyAsync(items0.p, &root, sizeof root, hipMemcpyHostToDevice, s));
    const uint32_t one = 1;
    LB_CHECK(hipMemcpyAsync(counters.p, &one, 4, hipMemcpyHostToDevice, s));  // the root occupies node 0
    uint32_t nitems = 1;
    int level = 0;
    Item* cur = items0.as<Item>();
    Item* nxt = items1.as<Item>();
    while (nitems) {
        LB_CHECK(hipMemsetAsync(counters.as<uint32_t>() + 2, 0, 4, s));
        hipLaunchKernelGGL(k_collapse, dim3(grid_for(nitems, 64)), dim3(64), 0, s, (const Item*)cur, nitems, nxt,
                           (const uint32_t*)left.as<uint32_t>(), (const uint32_t*)right.as<uint32_t>(),
                           (const uint32_t*)first.as<uint32_t>(), (const uint32_t*)count.as<uint32_t>(),
                           (const float*)lbox.as<float>(), (const float*)nbox.as<float>(), (const uint32_t*)vals2.as<uint32_t>(),
                           tris, n, nodes.as<float4>(), tout.as<float4>(), counters.as<uint32_t>());
        LB_CHECK(hipGetLastError());
        uint32_t c[4];
        LB_CHECK(hipMemcpyAsync(c, counters.p, 16, hipMemcpyDeviceToHost, s));
        LB_CHECK(hipStreamSynchronize(s));
        if (c[3]) throw std::runtime_error("device BVH quantisation failed (flags " + std::to_string(c[3]) + ")");
        level++;
        if (level > kMaxDepth) throw std::runtime_error("device BVH deeper than the traversal stack");
        nitems = c[2];
        std::swap(cur, nxt);
        r.node_count = c[0];
        r.tri_count = c[1];
    }
    if (r.tri_count != n) throw std::runtime_error("device BVH collapse lost triangles");
    LB_CHECK(hipEventRecord(e1, s));
    LB_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    LB_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    r.ms = ms;
    r.depth4 = level;
    r.nodes = reinterpret_cast<float4*>(nodes.release());
    r.tris = reinterpret_cast<float4*>(tout.release());
    return r;
}

}  // namespace st
