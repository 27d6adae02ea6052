// st_trace.h — gfx950 BVH traversal with the watertight ray/triangle test.
//
// One lane per ray.  Each step loads one 48-byte 4-wide node (three 16-byte
// loads), tests its four quantised child boxes, intersects hit leaf children
// immediately and descends into the nearest hit inner child, pushing the
// others.  The per-lane stack lives in LDS (kLdsStack entries, lane-interleaved
// so a wave's pushes hit distinct banks) and spills to HBM past that depth.
// The cost of a divergent vector load is paid per instruction in the texture
// address unit, so the node format minimises loads per box tested.
//
// Closest-hit ties (equal t, e.g. a ray through a shared edge) resolve to the
// smaller instance|primitive id, so the result does not depend on BVH shape or
// visiting order.
#pragma once

#include "st_device.h"
#include "st_shade.h"

namespace st {

struct Hit {
    float t, b1, b2;
    uint32_t id;
    bool hit;
};

struct WRay {  // Woop-Benthin-Wald shear constants, JCGT 2(1) 2013
    v3 o;
    uint32_t k;  // kx | ky << 2 | kz << 4
    float Sx, Sy, Sz;
    __device__ __forceinline__ int kx() const { return (int)(k & 3u); }
    __device__ __forceinline__ int ky() const { return (int)((k >> 2) & 3u); }
    __device__ __forceinline__ int kz() const { return (int)(k >> 4); }
};

__device__ __forceinline__ WRay wray_make(v3 o, v3 d) {
    WRay r;
    r.o = o;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (comp(d, kz) < 0) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    float dz = comp(d, kz);
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.0f / dz;
    r.k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
    return r;
}

// Watertight test; t > 0 required; b1/b2 are the weights of v1/v2.
__device__ __forceinline__ bool wt_intersect(const WRay& r, v3 v0, v3 v1, v3 v2, float& t, float& b1, float& b2) {
    v3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const int kx = r.kx(), ky = r.ky(), kz = r.kz();
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Ax = comp(A, kx) - r.Sx * Akz, Ay = comp(A, ky) - r.Sy * Akz;
    float Bx = comp(B, kx) - r.Sx * Bkz, By = comp(B, ky) - r.Sy * Bkz;
    float Cx = comp(C, kx) - r.Sx * Ckz, Cy = comp(C, ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt > 0.0f)) return false;
    t = tt;
    b1 = V * rcp;
    b2 = W * rcp;
    return true;
}

// Conservative slab test: boxes are rounded outward by the builder and the
// exit distance is widened by 2 ulp, so a box that contains a hit is entered.
__device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz, v3 o, v3 inv, float tmax,
                                     float& tn) {
    float t0x = (lx - o.x) * inv.x, t1x = (hx - o.x) * inv.x;
    float t0y = (ly - o.y) * inv.y, t1y = (hy - o.y) * inv.y;
    float t0z = (lz - o.z) * inv.z, t1z = (hz - o.z) * inv.z;
    float a = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.0f));
    float b = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    tn = a;
    return a <= b * 1.00000024f;
}

struct TraverseStats {
    uint32_t nodes, tris;
};

// The alpha test of a candidate hit on a triangle whose material has an
// alpha mask (intersection.hlsli:125-147): commit when the mask, sampled at
// level 0 at the hit's texture coordinate, reaches the material's cutoff.
// The mask is channel 0 of the material image (the reference's mImage1s).
__device__ inline bool alpha_pass(const DevScene& S, uint32_t id, float b1, float b2) {
    const uint32_t ii = id & 0xffffu, prim = id >> 16;
    const st_instance in = S.inst[ii];
    const uint32_t* mw = S.mat + (in.type_material_address >> 4) / 4;
    const uint32_t mask = mw[12];
    const float cutoff = __uint_as_float(mw[13]);
    const st_mesh_vertex_info vi = S.vi[in.data & 0xffffu];
    uint32_t tri[3];
    load_tri(S, vi, prim, tri);
    float t0[2] = {0, 0}, t1[2] = {0, 0}, t2[2] = {0, 0};
    const uint32_t tb = vi.packed_buffer_indices[1] >> 16, ts = vi.packed_strides >> 24;
    if (tb < ST_IMAGE_LIMIT) {
        const uint8_t* Tp = S.vb[tb] + vi.packed_offsets[3];
        __builtin_memcpy(t0, Tp + (size_t)ts * tri[0], 8);
        __builtin_memcpy(t1, Tp + (size_t)ts * tri[1], 8);
        __builtin_memcpy(t2, Tp + (size_t)ts * tri[2], 8);
    }
    const float u = t0[0] + (t1[0] - t0[0]) * b1 + (t2[0] - t0[0]) * b2;
    const float v = t0[1] + (t1[1] - t0[1]) * b1 + (t2[1] - t0[1]) * b2;
    return tex_bilinear(S, S.tex_img[mask].x, u, v).x >= cutoff;
}

// raySphere (compat/common.h:124-135) about the origin: the entry and exit
// parameters of o + t d, or (0, 0) when the ray misses
__device__ __forceinline__ void ray_sphere(v3 o, v3 d, float r, float& t0, float& t1) {
    const float a = dot(d, d), b = dot(o, d);
    const v3 l = o * a - d * b;
    float det = sq(a * r) - dot(l, l);
    if (det < 0) {
        t0 = t1 = 0;
        return;
    }
    const float inv_a = 1 / a;
    det = sqrtf(det * inv_a) * inv_a;
    const float m = -b * inv_a;
    t0 = m - det;
    t1 = m + det;
}
// The object-space ray of instance ii (CandidateObjectRayOrigin/Direction)
__device__ __forceinline__ void object_ray(const DevScene& S, uint32_t ii, v3 o, v3 d, v3& oo, v3& od) {
    const st_transform& T = S.inv_xf[ii];
    oo = mk(T.m[0][0] * o.x + T.m[0][1] * o.y + T.m[0][2] * o.z + T.m[0][3],
            T.m[1][0] * o.x + T.m[1][1] * o.y + T.m[1][2] * o.z + T.m[1][3],
            T.m[2][0] * o.x + T.m[2][1] * o.y + T.m[2][2] * o.z + T.m[2][3]);
    od = mk(T.m[0][0] * d.x + T.m[0][1] * d.y + T.m[0][2] * d.z, T.m[1][0] * d.x + T.m[1][1] * d.y + T.m[1][2] * d.z,
            T.m[2][0] * d.x + T.m[2][1] * d.y + T.m[2][2] * d.z);
}

// Traversal state of one lane.  step() visits one 4-wide node: its leaf
// children are intersected at once, the nearest hit inner child is next and
// the others are pushed.  Returns false when the ray is finished.
struct Lane {
    WRay wr;
    v3 inv;
    v3 dir;  // read by the sphere tests only
    Hit h;
    uint32_t node;
    int sp;

    __device__ __forceinline__ void start(v3 o, v3 d, float tmax) {
        wr = wray_make(o, d);
        dir = d;
        inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        h.hit = false;
        h.t = tmax;
        h.b1 = h.b2 = 0;
        h.id = 0xffffffffu;
        node = 0;
        sp = 0;
    }

    template <bool ALPHA = false>
    __device__ __forceinline__ bool tri_test(const DevScene& S, uint32_t ti) {
        const float4* tp = S.tris + 3 * (size_t)ti;
        const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        float t, b1, b2;
        if (!wt_intersect(wr, mk(t0.x, t0.y, t0.z), mk(t1.x, t1.y, t1.z), mk(t2.x, t2.y, t2.z), t, b1, b2)) return false;
        const uint32_t id = __float_as_uint(t0.w);
        if (ALPHA && __float_as_uint(t1.w) && (t < h.t || (t == h.t && h.id != 0xffffffffu && id < h.id)) &&
            !alpha_pass(S, id, b1, b2))
            return false;
        if (t < h.t || (t == h.t && h.id != 0xffffffffu && id < h.id)) {
            h.t = t;
            h.b1 = b1;
            h.b2 = b2;
            h.id = id;
            h.hit = true;
            return true;
        }
        return false;
    }

    // a sphere instance's procedural hit (intersection.hlsli:82-90): the entry
    // distance, or the exit one from inside, when it is in (0, t); ties go to
    // the smaller id as for triangles
    __device__ __forceinline__ bool sph_test(const DevScene& S, uint32_t k) {
        const uint32_t ii = __float_as_uint(S.sph_tris[3 * (size_t)k].w);
        v3 oo, od;
        object_ray(S, ii, wr.o, dir, oo, od);
        float t0, t1;
        ray_sphere(oo, od, __uint_as_float(S.inst[ii].data), t0, t1);
        if (!(t0 < t1)) return false;
        const float t = t0 > 0 ? t0 : t1;
        const uint32_t id = ii | 0xffff0000u;
        if (!(t > 0) || !(t < h.t || (t == h.t && h.id != 0xffffffffu && id < h.id))) return false;
        h.t = t;
        h.b1 = h.b2 = 0;
        h.id = id;
        h.hit = true;
        return true;
    }

    __device__ __forceinline__ void push(uint32_t v, uint32_t* lds, uint32_t* spill, uint32_t stride, uint32_t gtid) {
        if (sp < kLdsStack) lds[sp * (int)blockDim.x + (int)threadIdx.x] = v;
        else spill[(size_t)(sp - kLdsStack) * stride + gtid] = v;
        sp++;
    }

    // One 4-wide node: decode four quantised child boxes (o + q * 2^e, exact
    // in float), slab-test them, intersect hit leaf children at once, continue
    // with the nearest hit inner child and push the others farthest first.
    // SPH: a step in the sphere instances' tree (leaf records are spheres);
    // ALPHA: triangles of alpha-masked materials take the alpha test
    template <bool ANY, bool COUNT, bool SPH = false, bool ALPHA = false>
    __device__ __forceinline__ bool step(const DevScene& S, uint32_t* lds, uint32_t* spill, uint32_t spill_stride,
                                         uint32_t gtid, TraverseStats& st) {
        const float4* np = (SPH ? S.sph_nodes : S.nodes) + 3 * (size_t)node;
        const float4 w0 = np[0], w1 = np[1], w2 = np[2];
        if (COUNT) st.nodes++;
        const uint32_t x0 = __float_as_uint(w0.w), wi = __float_as_uint(w2.z);
        const float sx = __uint_as_float((x0 & 0xffu) << 23), sy = __uint_as_float(((x0 >> 8) & 0xffu) << 23),
                    sz = __uint_as_float(((x0 >> 16) & 0xffu) << 23);
        const uint32_t types = (x0 >> 24) | ((wi >> 28) << 8);
        const uint32_t base_inner = wi & 0x0fffffffu, base_tri = __float_as_uint(w2.w);
        const uint32_t qlx = __float_as_uint(w1.x), qly = __float_as_uint(w1.y), qlz = __float_as_uint(w1.z);
        const uint32_t qhx = __float_as_uint(w1.w), qhy = __float_as_uint(w2.x), qhz = __float_as_uint(w2.y);
        const v3 o = wr.o;
        // sort keys: entry distance bits (non-negative floats order as integers)
        // with the child slot in the low 2 bits; empty = all ones
        uint32_t key[4];
        uint32_t tri_off = 0, innermask = 0, trimask = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t type = (types >> (3 * c)) & 7u;
            const int sh = 8 * c;
            const float lx = w0.x + (float)((qlx >> sh) & 0xffu) * sx, hx = w0.x + (float)((qhx >> sh) & 0xffu) * sx;
            const float ly = w0.y + (float)((qly >> sh) & 0xffu) * sy, hy = w0.y + (float)((qhy >> sh) & 0xffu) * sy;
            const float lz = w0.z + (float)((qlz >> sh) & 0xffu) * sz, hz = w0.z + (float)((qhz >> sh) & 0xffu) * sz;
            float tn;
            const bool hit = type != 0 && slab(lx, hx, ly, hy, lz, hz, o, inv, h.t, tn);
            key[c] = 0xffffffffu;
            if (type == 7) {
                if (hit) key[c] = (__float_as_uint(tn) & ~3u) | (uint32_t)c;
                innermask |= 1u << c;
            } else {
                // leaf: mark its triangles (offsets tri_off .. tri_off + type - 1, at most 16 per node)
                if (hit) trimask |= ((1u << type) - 1u) << tri_off;
                tri_off += type;
            }
        }
        // one loop over all marked triangles keeps the lanes of a wavefront
        // that have leaf work active together
        while (trimask) {
            const uint32_t j = (uint32_t)__builtin_ctz(trimask);
            trimask &= trimask - 1u;
            if (COUNT) st.tris++;
            if ((SPH ? sph_test(S, base_tri + j) : tri_test<ALPHA>(S, base_tri + j)) && ANY) return false;
        }
        // sort the (at most 4) inner candidates by entry distance
#define ST_CSWAP(i, j)                                   \
    {                                                    \
        const uint32_t lo = min(key[i], key[j]);         \
        key[j] = max(key[i], key[j]);                    \
        key[i] = lo;                                     \
    }
        ST_CSWAP(0, 1) ST_CSWAP(2, 3) ST_CSWAP(0, 2) ST_CSWAP(1, 3) ST_CSWAP(1, 2)
#undef ST_CSWAP
        auto child = [&](uint32_t k) { return base_inner + (uint32_t)__popc(innermask & ((1u << (k & 3u)) - 1u)); };
        if (key[3] != 0xffffffffu) push(child(key[3]), lds, spill, spill_stride, gtid);
        if (key[2] != 0xffffffffu) push(child(key[2]), lds, spill, spill_stride, gtid);
        if (key[1] != 0xffffffffu) push(child(key[1]), lds, spill, spill_stride, gtid);
        if (key[0] != 0xffffffffu) {
            node = child(key[0]);
            return true;
        }
        if (sp == 0) return false;
        sp--;
        node = sp < kLdsStack ? lds[sp * (int)blockDim.x + (int)threadIdx.x]
                              : spill[(size_t)(sp - kLdsStack) * spill_stride + gtid];
        return true;
    }
};

// ANY: stop at the first hit in (0, tmax).  lds: this block's stack,
// [kLdsStack][blockDim.x]; spill: HBM stack region, [kMaxDepth - kLdsStack][stride].
template <bool ANY, bool COUNT>
__device__ __forceinline__ Hit traverse(const DevScene& S, v3 o, v3 d, float tmax, uint32_t* lds, uint32_t* spill,
                                        uint32_t spill_stride, uint32_t gtid, TraverseStats& st) {
    Lane L;
    L.start(o, d, tmax);
    if (S.node_count != 0)
        while (L.step<ANY, COUNT>(S, lds, spill, spill_stride, gtid, st)) {
        }
    if (S.sph_node_count != 0 && !(ANY && L.h.hit)) {  // then the sphere instances, up to the nearest hit
        L.node = 0;
        L.sp = 0;
        while (L.step<ANY, COUNT, true>(S, lds, spill, spill_stride, gtid, st)) {
        }
    }
    return L.h;
}

// Persistent traversal over a queue of n rays with per-lane dynamic fetch:
// a lane whose ray is finished takes the next ray index (one atomic per
// wavefront per refill), so lanes do not idle behind the slowest ray of their
// wavefront.  Src supplies load(i, o, d, tmax) and done(i, hit).
//
// XCD-aware: workgroups are dispatched to the XCDs round robin, so block b
// runs on XCD b % kXcds.  The queue is dealt out in chunks of kChunk rays,
// chunk c to XCD c % kXcds, and each XCD has its own cursor over its chunks
// (fetch[x * kFetchStride], one 128-byte line each).  Rays that are
// neighbours in the queue (neighbouring pixels, for camera rays) thus meet in
// one XCD's L2, all XCDs still sweep the queue together (the chunks of one
// round are spread over the whole machine), and the ray-fetch atomics spread
// over eight addresses instead of one.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kFetchStride = 32;  // words between the XCDs' cursors
constexpr uint32_t kChunk = 2048;
// SPH: the scene has sphere instances; a lane done with the triangles walks
// the spheres' tree next, up to its nearest hit (any hit: only when it found
// none).
template <bool ANY, bool COUNT, class Src, bool SPH = false, bool ALPHA = false>
__device__ __forceinline__ void trace_queue(const DevScene& S, uint32_t n, uint32_t* fetch, Src& src, uint32_t* lds,
                                            uint32_t* spill, TraverseStats& st) {
    const uint32_t stride = gridDim.x * blockDim.x, gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane_id = __lane_id();
    const uint32_t home = blockIdx.x % kXcds;
    // rays dealt to this XCD: whole rounds of kXcds chunks, then its share of the last round
    const uint32_t round = kChunk * kXcds, rem = n % round;
    const uint32_t mine = (n / round) * kChunk + (rem > home * kChunk ? min(rem - home * kChunk, kChunk) : 0u);
    uint32_t* cursor = fetch + home * kFetchStride;
    Lane L;
    bool active = false, drained = false, sph = false;
    uint32_t idx = 0;
    while (true) {
        if (!drained) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                const uint32_t cnt = (uint32_t)__popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                uint32_t base = 0;
                if ((int)lane_id == leader) base = atomicAdd(cursor, cnt);
                base = __shfl(base, leader);
                if ((uint64_t)base + cnt >= mine) drained = true;
                if (!active) {
                    const uint32_t j = base + (uint32_t)__popcll(need & ((1ull << lane_id) - 1ull));
                    if (j < mine) {
                        const uint32_t i = ((j / kChunk) * kXcds + home) * kChunk + j % kChunk;
                        v3 o, d;
                        float tmax;
                        src.load(i, o, d, tmax);
                        L.start(o, d, tmax);
                        idx = i;
                        sph = SPH && S.node_count == 0;
                        active = S.node_count != 0 || SPH;
                        if (!active) src.done(i, L.h);
                    }
                }
            }
        }
        if (!__any(active)) break;
        if (SPH) {
            if (active && !(sph ? L.step<ANY, COUNT, true>(S, lds, spill, stride, gtid, st)
                                : L.step<ANY, COUNT, false, ALPHA>(S, lds, spill, stride, gtid, st))) {
                if (!sph && !(ANY && L.h.hit)) {
                    sph = true;
                    L.node = 0;
                    L.sp = 0;
                } else {
                    src.done(idx, L.h);
                    active = false;
                }
            }
        } else if (active && !L.step<ANY, COUNT, false, ALPHA>(S, lds, spill, stride, gtid, st)) {
            src.done(idx, L.h);
            active = false;
        }
    }
}

}  // namespace st
