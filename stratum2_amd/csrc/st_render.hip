// st_render.hip — the wavefront bounce loop on gfx950 and the C ABI.
//
// The reference runs one thread per pixel for every bounce and keeps a 64-byte
// PathState per pixel (testrenderer.slang:1087-1144, gMultiDispatch).  Here a
// wave of W*H*k paths (k samples at once) lives in SoA arrays in HBM and only
// live paths are re-launched: each bounce is
//   trace   — closest hit of every queued ray                   (k_trace_closest)
//   shade   — emission MIS, NEE sampling, RR, BSDF sampling;
//             appends shadow rays and surviving paths to compacted
//             queues with one atomic per wavefront (ballot+popcount)  (k_shade)
//   shadow  — any-hit of the appended shadow rays, fixed-point NEE  (k_trace_shadow)
// and a final resolve sums each pixel's samples in sample order (k_resolve).
// Scenes with media swap the two traversal kernels for k_trace_media (delta
// tracking through the current medium, boundary crossings, scattering
// vertices) and k_shadow_media (ratio tracking), and shade medium vertices
// with the Henyey-Greenstein phase function.
// All kernels are persistent grid-stride loops whose item counts are read on
// the device, so the host never waits inside a frame.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/stratum2.h"
#include "st_bvh.h"
#include "st_cache.h"
#include "st_layout.h"
#include "st_lbvh.h"
#include "st_shade.h"
#include "st_svgf.h"
#include "st_tonemap.h"
#include "st_trace.h"
#include "st_volume.h"

using namespace st;

namespace {

thread_local std::string g_error;

int set_error(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_CHECK(x)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess)                                                                              \
            throw HipError(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +        \
                           std::to_string(__LINE__) + ")");                                                \
    } while (0)

constexpr int kTraceBlock = 128;
constexpr int kShadeBlock = 256;
constexpr int kTraceBlocksPerCU = 16;
// WaveState::counters (kCounterWords words): [0..1] queue counts (ping-pong),
// [2] shadow count, [3] radiance-cache queries, [4..7] u64 ray counts,
// [8..9] u64 radiance-cache queries of the call, then the traversal fetch cursors:
// kXcds per traversal kernel, kFetchStride words (one 128-byte line) apart
constexpr int kFetchClosest = 256, kFetchShadow = kFetchClosest + (int)(kXcds * kFetchStride),
              kFetchRays = kFetchShadow + (int)(kXcds * kFetchStride);
constexpr int kCounterWords = kFetchRays + (int)(kXcds * kFetchStride);  // 32 wavefronts per CU (7-8 per SIMD at <= 72 VGPRs), 8 KB LDS each

// ----------------------------------------------------------------- device state
struct WaveState {
    float4* ray;        // per path: o.xyz, packed direction bits
    float4* thr;        // per path: throughput.xyz, unused
    uint4* st;          // per path: rng counter, path length, t16|pdfA16<<16, dirPdf16
    float4* em;         // per path: emission accumulated in fp32 (AddRadiance)
    uint4* nee;         // per path: NEE in 1/16384 fixed point (AddRadianceInterlocked)
    uint4* hit;         // per queue slot: t, id, b1, b2
    uint32_t* queue[2]; // live path ids, ping-pong
    float4* shadow;     // per shadow record: {o, tmax}, {d, path}, {contrib, weight}
    uint32_t* counters; // [0..1] queue counts (ping-pong), [2] shadow count, [4..5] u64 rays
    // ReSTIR DI screen-space reuse (one sample per wave), per pixel
    uint4* pri;         // primary hit: id, b1, b2, packed ray direction
    float4* r_init;     // 2 per pixel: light-sample seed (float4), {target, W, M, depth}
    float4* r_temp;     // after temporal reuse
    float4* r_prev;     // after spatial reuse: this frame's result, next frame's history
    float4* psd;        // 2 per pixel: the primary hit's shading data (pos, material, gn, sn, tangent, area)
    uint32_t* med;      // per path: current medium instance (kNoMedium outside media)
    // bidirectional mode (gUseVC)
    float2* vcm;        // per path: dVC, dVCM (fp16-rounded unless ST_FLAG_VC_FP32_MIS)
    float4* lv_raw;     // light vertices, lv_stride slots per light subpath: {local pos, id}, {thr, dir in}, {length, dVC, dVCM}
    uint32_t* lv_cnt;   // per light subpath: vertices stored
    uint32_t* lv_off;   // exclusive prefix sums of lv_cnt
    float4* lv;         // the cache: vertices in (subpath, depth) order
    uint32_t* lv_total; // vertices in the cache
    // radiance cache (st_render_cached): queries of paths ending in the cache
    // ({pos, thr.x}, {wo, thr.y}, {normal, thr.z}, {albedo, path}), counted in
    // counters[3]; training rows (12-float query, {throughput, 0}, {gathered
    // radiance, weight} -> {target, weight})
    float4* cq;
    float* tq;
    float4* tthr;
    float4* tt;
    // hash-grid reservoir reuse: this sample's entries ({local pos, id},
    // seed, {target, W, M, checksum}, {local dir in, 0}) at slot
    // vertex * W * H + pixel, and the previous sample's grid (its entries,
    // and their slots sorted by cell key)
    float4* hg_rec;
    const float4* hg_prev_rec;
    const unsigned long long* hg_prev_keys;
    const uint32_t* hg_prev_slots;
};

struct RenderConsts {
    uint32_t W, H, npix, wave_paths, k, sample_base, seed, max_depth, min_len, flags, rr_start, cand, tiled, spatial;
    float env_prob;
    float max_m, radius;
    uint32_t view_medium, max_null;
    uint32_t cache_depth, train_stride, train_frame;  // radiance cache: query depth, training-path selection
    float hg_cell;                                    // hash-grid reservoir reuse: cell size,
    uint32_t hg_dmax, hg_prev, hg_nslots;             // reservoir vertices kept per path, previous grid valid, slots
    uint32_t light_paths;  // light tracing and the bidirectional mode: light subpaths per sample (mLightSubpathCount)
    uint32_t lv_stride;    // bidirectional mode: light vertex slots per subpath (max_depth - 1)
    float ipd;             // light tracing: ViewData::imagePlaneDist (compat/scene.h:147)
    st_view view;
    st_transform inv_cam;  // world to camera (TransformData::inverse of the camera transform)
};

// The second pass of st_adjoint (path replay): the first pass's fp32 radiance
// of each path, dLoss/dHDR, and the gradient sinks (null: not wanted).
struct AdjArgs {
    const float4* ltot;  // per path
    const float4* d_hdr; // per pixel, RGBA
    float* grad_alb;     // 3 per flattened triangle
    float* grad_grid;    // per 4-byte word of grid grad_vol
    uint32_t grad_vol, spp;
};

// Per-block accumulation of the albedo gradient: a 512-entry open-addressing
// table in LDS (key = albedo entry), flushed to the gradient once at the end
// of the kernel, so the global atomics per launch scale with the blocks and
// the distinct triangles they saw, not with the vertices (thousands of paths
// land on each wall triangle).
constexpr int kAdjSlots = 512;
struct AdjTable {
    uint32_t key[kAdjSlots];
    float val[3 * kAdjSlots];
};
__device__ inline void adj_table_clear(AdjTable& t) {
    for (int i = threadIdx.x; i < kAdjSlots; i += blockDim.x) {
        t.key[i] = 0xffffffffu;
        t.val[3 * i] = t.val[3 * i + 1] = t.val[3 * i + 2] = 0.0f;
    }
}
// into the table (8 probes), else straight to the gradient
__device__ inline void adj_table_add(AdjTable& t, float* base, uint32_t key, float x, float y, float z) {
    const uint32_t h = (key * 2654435761u) >> 23;
    for (uint32_t probe = 0; probe < 8; probe++) {
        const uint32_t sl = (h + probe) & (kAdjSlots - 1);
        const uint32_t old = atomicCAS(&t.key[sl], 0xffffffffu, key);
        if (old == 0xffffffffu || old == key) {
            atomicAdd(&t.val[3 * sl], x);
            atomicAdd(&t.val[3 * sl + 1], y);
            atomicAdd(&t.val[3 * sl + 2], z);
            return;
        }
    }
    atomicAdd(base + 3 * (size_t)key, x);
    atomicAdd(base + 3 * (size_t)key + 1, y);
    atomicAdd(base + 3 * (size_t)key + 2, z);
}
__device__ inline void adj_table_flush(const AdjTable& t, float* base) {
    for (int i = threadIdx.x; i < kAdjSlots; i += blockDim.x) {
        const uint32_t k = t.key[i];
        if (k == 0xffffffffu) continue;
        atomicAdd(base + 3 * (size_t)k, t.val[3 * i]);
        atomicAdd(base + 3 * (size_t)k + 1, t.val[3 * i + 1]);
        atomicAdd(base + 3 * (size_t)k + 2, t.val[3 * i + 2]);
    }
}
// Adds v at key for every lane whose key is not ~0u: the lanes that share the
// first lane's key (a tile of primary rays on one wall triangle) are summed
// across the wavefront first, the others go to the table one by one.  Every
// lane of the wavefront must call it.
__device__ inline void wave_add_keyed3(AdjTable& t, float* base, uint32_t key, v3 v) {
    bool want = key != 0xffffffffu;
    const unsigned long long m = __ballot(want);
    if (m) {
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t k = __shfl(key, leader);
        const bool mine = want && key == k;
        float x = mine ? v.x : 0.0f, y = mine ? v.y : 0.0f, z = mine ? v.z : 0.0f;
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_xor(x, off);
            y += __shfl_xor(y, off);
            z += __shfl_xor(z, off);
        }
        if ((int)__lane_id() == leader) adj_table_add(t, base, k, x, y, z);
        if (mine) want = false;
    }
    if (want) adj_table_add(t, base, key, v.x, v.y, v.z);
}

// dLoss/dL of the path through pixel pix: dLoss/dHDR / spp
__device__ __forceinline__ v3 adj_pixel_weight(const AdjArgs& a, uint32_t pix) {
    const float4 d = a.d_hdr[pix];
    return mk(d.x / (float)a.spp, d.y / (float)a.spp, d.z / (float)a.spp);
}

// Auxiliary outputs of sample 0's first vertex (StoreAuxillaryData); null = not wanted.
struct AuxOut {
    uint32_t* visibility;
    float* albedo;
    float* depth;
    float* prev_uv;
};

// A dVC / dVCM store: fp16, as the reference's PathState keeps them
// (testrenderer.slang:75-82), unless ST_FLAG_VC_FP32_MIS.
__device__ __forceinline__ float qvc(const RenderConsts& c, float x) {
    return (c.flags & ST_FLAG_VC_FP32_MIS) ? x : q16(x);
}

// -------------------------------------------------------------------- kernels
// Path order inside one sample: 8x8 pixel tiles (one wavefront of primary rays
// covers one tile) when the image is a multiple of 8 in both dimensions,
// scanline order otherwise.  q is the path index within its sample.
__device__ __forceinline__ void pixel_of(const RenderConsts& c, uint32_t q, uint32_t& px, uint32_t& py) {
    if (c.tiled) {
        const uint32_t t = q >> 6, j = q & 63u, tx = c.W >> 3;
        px = (t % tx) * 8 + (j & 7u);
        py = (t / tx) * 8 + (j >> 3);
    } else {
        px = q % c.W;
        py = q / c.W;
    }
}
__device__ __forceinline__ uint32_t path_of_pixel(const RenderConsts& c, uint32_t px, uint32_t py) {
    if (!c.tiled) return py * c.W + px;
    return (((py >> 3) * (c.W >> 3) + (px >> 3)) << 6) | ((py & 7u) << 3) | (px & 7u);
}

// ViewData::toRaster (compat/scene.h:158-165) with ProjectionData::projectPoint
// (compat/transform.h:139-155); false, with ox, oy unchanged, when p is not visible.
__device__ inline bool to_raster(const st_view& V, v3 p, float& ox, float& oy) {
    const st_projection& pr = V.projection;
    const float sz = p.z > 0 ? 1.0f : (p.z < 0 ? -1.0f : 0.0f);
    const float sn = pr.near_plane > 0 ? 1.0f : (pr.near_plane < 0 ? -1.0f : 0.0f);
    if (sz != sn) return false;
    float r0, r1, r2, r3;
    if (pr.vertical_fov < 0) {
        r0 = p.x * pr.scale[0] + pr.offset[0];
        r1 = p.y * pr.scale[1] + pr.offset[1];
        r2 = (p.z - pr.far_plane) / (pr.near_plane - pr.far_plane);
        r3 = 1;
    } else {
        r0 = p.x * pr.scale[0] + p.z * pr.offset[0];
        r1 = p.y * pr.scale[1] + p.z * pr.offset[1];
        r2 = fabsf(pr.near_plane);
        r3 = p.z * sn;
    }
    const float aw = fabsf(r3);
    if (fabsf(r0) >= aw || fabsf(r1) >= aw || fabsf(r2) >= aw) return false;
    const float ew = (float)(V.image_max[0] - V.image_min[0]), eh = (float)(V.image_max[1] - V.image_min[1]);
    ox = (float)V.image_min[0] + ew * ((r0 / r3) * 0.5f + 0.5f);
    oy = (float)V.image_min[1] + eh * ((-r1 / r3) * 0.5f + 0.5f);
    return true;
}

// StoreAuxillaryData (testrenderer.slang:195-218) for a static scene in steady
// state: identity motion transform, previous view = this view.
__device__ inline void store_aux(const RenderConsts& c, const AuxOut& aux, uint32_t pix, uint32_t px, uint32_t py,
                                 uint32_t id, uint32_t sn, bool has_instance, v3 pos, v3 albedo, float t16) {
    if (aux.visibility) {
        aux.visibility[2 * pix] = id;
        aux.visibility[2 * pix + 1] = sn;
    }
    if (!aux.albedo && !aux.depth && !aux.prev_uv) return;
    float pcx = (float)px + 0.5f, pcy = (float)py + 0.5f;
    float prev_depth = t16;
    if (has_instance) {
        const v3 pc = xf_point(c.inv_cam, pos);
        prev_depth = len(pc);
        to_raster(c.view, pc, pcx, pcy);
    }
    if (aux.albedo) reinterpret_cast<float4*>(aux.albedo)[pix] = make_float4(albedo.x, albedo.y, albedo.z, 1.0f);
    if (aux.depth) reinterpret_cast<float4*>(aux.depth)[pix] = make_float4(t16, prev_depth, 1.0f, 1.0f);
    if (aux.prev_uv) reinterpret_cast<float2*>(aux.prev_uv)[pix] = make_float2(pcx / (float)c.W, pcy / (float)c.H);
}

__global__ void __launch_bounds__(kShadeBlock) k_camera(RenderConsts c, WaveState w) {
    const uint32_t n = c.wave_paths;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const uint32_t k = p / c.npix;
        uint32_t px, py;
        pixel_of(c, p % c.npix, px, py);
        Rng rng{px, py, c.seed + c.sample_base + k, 0};
        float ox = 0.5f, oy = 0.5f;
        if (c.flags & ST_FLAG_PIXEL_JITTER) {
            float4 j = rng.nextf();
            ox = j.x;
            oy = j.y;
        }
        // InitializeCameraPath, testrenderer.slang:938-982; ViewData::toWorld,
        // compat/scene.h:152-157; ProjectionData::backProject, transform.h:159-170
        const st_view& V = c.view;
        const float ew = (float)(V.image_max[0] - V.image_min[0]), eh = (float)(V.image_max[1] - V.image_min[1]);
        const float u = (((float)px + ox) - (float)V.image_min[0]) / ew;
        const float v = (((float)py + oy) - (float)V.image_min[1]) / eh;
        const float cx = 2 * u - 1, cy = -(2 * v - 1);
        const st_projection& pr = V.projection;
        const float sg = pr.near_plane > 0 ? 1.0f : (pr.near_plane < 0 ? -1.0f : 0.0f);
        v3 bp = mk(pr.near_plane * (cx * sg - pr.offset[0]) / pr.scale[0],
                   pr.near_plane * (cy * sg - pr.offset[1]) / pr.scale[1], pr.near_plane);
        if (pr.vertical_fov < 0)  // orthographic projection (mVerticalFoV < 0)
            bp = mk((cx - pr.offset[0]) / pr.scale[0], (cy - pr.offset[1]) / pr.scale[1], pr.near_plane);
        const v3 ld = normalize(bp);
        const st_transform& T = V.camera_to_world;
        const uint32_t dir = pack_normal(normalize(xf_vector(T, ld)));
        w.ray[p] = make_float4(T.m[0][3], T.m[1][3], T.m[2][3], __uint_as_float(dir));
        w.thr[p] = make_float4(1, 1, 1, 0);
        w.em[p] = make_float4(0, 0, 0, 0);
        if (c.flags & ST_FLAG_BDPT_LVC) {
            // eye sub-path MIS quantities (testrenderer.slang:958-977); the specular
            // flag sits in bit 31 of the dirPdf word; the light pass has already
            // added its camera connections to this path's accumulator
            const float cos_cam = fabsf(ld.z);
            const float ipcd = c.ipd / cos_cam;
            const float cam_pdfw = sq(ipcd) / cos_cam;
            w.st[p] = make_uint4(rng.w, 0, 0, f2h(cam_pdfw) | 0x80000000u);
            w.vcm[p] = make_float2(qvc(c, 0.0f), qvc(c, mis1((float)c.light_paths / cam_pdfw)));
        } else {
            w.st[p] = make_uint4(rng.w, 0, 0, 0);
            w.nee[p] = make_uint4(0, 0, 0, 0);
        }
        w.med[p] = c.view_medium;
        w.queue[0][p] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.counters[0] = n;
        w.counters[1] = 0;
        w.counters[2] = 0;
        w.counters[3] = 0;
        for (int i = 0; i < 2 * (int)kXcds; i++) w.counters[kFetchClosest + i * kFetchStride] = 0;
    }
}

struct ClosestSrc {
    WaveState w;
    const uint32_t* q;
    __device__ __forceinline__ void load(uint32_t i, v3& o, v3& d, float& tmax) const {
        const float4 r = w.ray[q[i]];
        o = mk(r.x, r.y, r.z);
        d = unpack_normal(__float_as_uint(r.w));
        tmax = __int_as_float(0x7f800000);
    }
    __device__ __forceinline__ void done(uint32_t i, const Hit& h) const {
        w.hit[i] = make_uint4(__float_as_uint(h.t), h.hit ? h.id : 0xffffffffu, __float_as_uint(h.b1), __float_as_uint(h.b2));
    }
};

template <bool COUNT, bool SPH = false, bool ALPHA = false>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7))) k_trace_closest(DevScene S, WaveState w, int qsel, uint32_t* spill,
                                                               unsigned long long* stats) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    ClosestSrc src{w, w.queue[qsel]};
    TraverseStats ts{0, 0};
    trace_queue<false, COUNT, ClosestSrc, SPH, ALPHA>(S, w.counters[qsel], &w.counters[kFetchClosest], src, lds, spill, ts);
    if (COUNT) {
        atomicAdd(&stats[0], (unsigned long long)ts.nodes);
        atomicAdd(&stats[1], (unsigned long long)ts.tris);
    }
}

// Shadow rays: any-hit, then AddRadianceInterlocked (testrenderer.slang:237-261)
// into the path's own fixed-point accumulator.
// AddRadianceInterlocked (testrenderer.slang:237-261): p is the path (camera
// paths) or the pixel slot (light tracing).  max(0, x) maps NaN to 0, as the
// GPU max of the reference does.
__device__ __forceinline__ float clamp0(float x) { return x > 0 ? x : 0.0f; }
__device__ __forceinline__ void add_shadow_contribution(const WaveState& w, uint32_t flags, uint32_t p, v3 ct, float wgt) {
    if (wgt <= 0 || all_nonpos(ct)) return;
    const v3 rw = ct * wgt;
    if (flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC)) {
        // light paths of one launch meet in the same pixels (and a bidirectional
        // camera vertex makes two shadow rays): 32-bit atomics, whose sums do
        // not depend on the order
        const float q0 = clamp0(rw.x) * (float)kQuant, q1 = clamp0(rw.y) * (float)kQuant, q2 = clamp0(rw.z) * (float)kQuant;
        uint32_t* a = reinterpret_cast<uint32_t*>(w.nee + p);
        const uint32_t i0 = (uint32_t)(q0 >= 2147483647.0f ? 2147483647 : (int32_t)q0);
        const uint32_t i1 = (uint32_t)(q1 >= 2147483647.0f ? 2147483647 : (int32_t)q1);
        const uint32_t i2 = (uint32_t)(q2 >= 2147483647.0f ? 2147483647 : (int32_t)q2);
        if (i0) atomicAdd(a, i0);
        if (i1) atomicAdd(a + 1, i1);
        if (i2) atomicAdd(a + 2, i2);
        return;
    }
    if (flags & ST_FLAG_FLOAT_NEE) {
        float4 e = w.em[p];
        e.x += fmax_(0.0f, rw.x);
        e.y += fmax_(0.0f, rw.y);
        e.z += fmax_(0.0f, rw.z);
        w.em[p] = e;
        return;
    }
    const float q0 = clamp0(rw.x) * (float)kQuant, q1 = clamp0(rw.y) * (float)kQuant, q2 = clamp0(rw.z) * (float)kQuant;
    // one shadow ray per path per bounce: a plain read-modify-write is race free
    uint4 acc = w.nee[p];
    acc.x += (uint32_t)(q0 >= 2147483647.0f ? 2147483647 : (int32_t)q0);
    acc.y += (uint32_t)(q1 >= 2147483647.0f ? 2147483647 : (int32_t)q1);
    acc.z += (uint32_t)(q2 >= 2147483647.0f ? 2147483647 : (int32_t)q2);
    w.nee[p] = acc;
}

struct ShadowSrc {
    WaveState w;
    uint32_t flags;
    __device__ __forceinline__ void load(uint32_t i, v3& o, v3& d, float& tmax) const {
        const float4 a = w.shadow[3 * (size_t)i], b = w.shadow[3 * (size_t)i + 1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmax = a.w;
    }
    __device__ __forceinline__ void done(uint32_t i, const Hit& h) const {
        if (h.hit) return;
        const float4 c = w.shadow[3 * (size_t)i + 2];
        const uint32_t p = __float_as_uint(w.shadow[3 * (size_t)i + 1].w);
        // light tracing keeps the light path in .w; its weight is 1 (no MIS)
        const bool lt_only = (flags & ST_FLAG_TRACE_FROM_LIGHT) && !(flags & ST_FLAG_BDPT_LVC);
        add_shadow_contribution(w, flags, p, mk(c.x, c.y, c.z), lt_only ? 1.0f : c.w);
    }
};

template <bool COUNT, bool SPH = false, bool ALPHA = false>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7))) k_trace_shadow(DevScene S, WaveState w, uint32_t flags, uint32_t* spill,
                                                              unsigned long long* stats) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    ShadowSrc src{w, flags};
    TraverseStats ts{0, 0};
    trace_queue<true, COUNT, ShadowSrc, SPH, ALPHA>(S, w.counters[2], &w.counters[kFetchShadow], src, lds, spill, ts);
    if (COUNT) {
        atomicAdd(&stats[2], (unsigned long long)ts.nodes);
        atomicAdd(&stats[3], (unsigned long long)ts.tris);
    }
}

// The trace kernels of a scene: with sphere instances, the lanes walk the
// spheres' tree after the triangles'.
using ClosestFn = void (*)(DevScene, WaveState, int, uint32_t*, unsigned long long*);
using ShadowFn = void (*)(DevScene, WaveState, uint32_t, uint32_t*, unsigned long long*);
static ClosestFn closest_kernel(bool count, const DevScene& S) {
    const bool sph = S.sph_node_count != 0;
    if (S.has_alpha)
        return count ? (sph ? k_trace_closest<true, true, true> : k_trace_closest<true, false, true>)
                     : (sph ? k_trace_closest<false, true, true> : k_trace_closest<false, false, true>);
    return count ? (sph ? k_trace_closest<true, true> : k_trace_closest<true, false>)
                 : (sph ? k_trace_closest<false, true> : k_trace_closest<false, false>);
}
static ShadowFn shadow_kernel(bool count, const DevScene& S) {
    const bool sph = S.sph_node_count != 0;
    if (S.has_alpha)
        return count ? (sph ? k_trace_shadow<true, true, true> : k_trace_shadow<true, false, true>)
                     : (sph ? k_trace_shadow<false, true, true> : k_trace_shadow<false, false, true>);
    return count ? (sph ? k_trace_shadow<true, true> : k_trace_shadow<true, false>)
                 : (sph ? k_trace_shadow<false, true> : k_trace_shadow<false, false>);
}

// Sum of a per-lane count over the wavefront, added once per wave.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// Random stream of path p at counter `counter`: camera paths are keyed by their
// pixel and sample; light paths (light_paths > 0) by their dispatch index
// (q % W, q / W) and sample (testrenderer.slang:983-984, 1092-1094).
__device__ __forceinline__ Rng path_rng(const RenderConsts& c, uint32_t p, uint32_t counter) {
    if (c.light_paths) {
        const uint32_t q = p % c.light_paths;
        return Rng{q % c.W, q / c.W, c.seed + c.sample_base + p / c.light_paths, counter};
    }
    uint32_t px, py;
    pixel_of(c, p % c.npix, px, py);
    return Rng{px, py, c.seed + c.sample_base + p / c.npix, counter};
}

// SceneParameters::GetMediumIndex (scene.hlsli:32-40): the first volume whose
// object-space bounds hold the point.
__device__ inline uint32_t medium_at(const DevScene& S, v3 pos) {
    for (uint32_t i = 0; i < S.vinfo_count; i++) {
        const st_volume_info vi = S.vinfo[i];
        const v3 lp = xf_point(S.inv_xf[vi.instance_index], pos);
        if (lp.x >= vi.min[0] && lp.y >= vi.min[1] && lp.z >= vi.min[2] && lp.x <= vi.max[0] && lp.y <= vi.max[1] &&
            lp.z <= vi.max[2])
            return vi.instance_index;
    }
    return kNoMedium;
}

// traceScatteringRay (intersection.hlsli:195-260) for every queued path of a
// scene with media: closest hit over triangles and volume boxes, delta
// tracking through the path's current medium up to that hit, medium
// boundaries crossed in place.  Updates the path's throughput (divided by the
// average free-flight pdf, TraceRay testrenderer.slang:1029-1030), random
// counter and medium.  A scattering vertex is reported as id medium | 0xffff
// << 16 with t = its distance from the ray origin, and its position replaces
// the ray origin; surface hits keep t of the last segment (the shading kernel
// measures the distance from the origin, intersection.hlsli:244).  One lane
// per path, grid-stride.
// ADJ: the path replay of st_adjoint; each collision in the gradient's grid
// adds its term with the radiance still to come on the path as weight.
template <bool ADJ = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace_media(DevScene S, RenderConsts c, WaveState w, int qsel, uint32_t* spill,
                                                             unsigned long long* stats, AdjArgs adj) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    const uint32_t n = w.counters[qsel];
    const uint32_t* q = w.queue[qsel];
    const uint32_t stride = gridDim.x * blockDim.x, gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraverseStats ts{0, 0};
    uint32_t segs = 0, reads = 0;
    for (uint32_t i = gtid; i < n; i += stride) {
        const uint32_t p = q[i];
        const float4 r = w.ray[p];
        const v3 origin = mk(r.x, r.y, r.z);
        const v3 dir = unpack_normal(__float_as_uint(r.w));
        const float4 t4 = w.thr[p];
        v3 thr = mk(t4.x, t4.y, t4.z);
        uint4 s4 = w.st[p];
        uint32_t med = w.med[p];
        Rng rng = path_rng(c, p, s4.x);
        float dirpdf = 1;
        v3 o = origin;
        float tmax = __int_as_float(0x7f800000);
        uint4 out = make_uint4(0, 0xffffffffu, 0, 0);
        bool scattered = false, done = false;
        v3 spos = mk(0, 0, 0);
        DensitySink sink{nullptr, 0, mk(0, 0, 0)};
        if (ADJ && adj.grad_grid) {  // every later contribution carries this ray's tracking factors
            uint32_t px, py;
            pixel_of(c, p % c.npix, px, py);
            const float4 lt = adj.ltot[p], e = w.em[p];
            sink = DensitySink{adj.grad_grid, adj.grad_vol,
                               adj_pixel_weight(adj, py * c.W + px) * mk(lt.x - e.x, lt.y - e.y, lt.z - e.z)};
        }
        // bounded: a ray still crossing boundaries after kMaxMediaSegments
        // traceRay calls ends as a miss (the reference loop is unbounded)
        for (int seg = 0; tmax > 1e-6f && seg < kMaxMediaSegments; seg++) {
            Hit h = traverse<false, false>(S, o, dir, tmax, lds, spill, stride, gtid, ts);
            segs++;
            bool is_vol;
            uint32_t vgn = 0;
            volume_boxes(S, o, dir, h, is_vol, vgn);
            if (med != kNoMedium) {
                const Medium m = load_medium(S, S.inst[med].type_material_address >> 4);
                const st_transform& iT = S.inv_xf[med];
                v3 dp3, np3;
                bool sc;
                const v3 pp = delta_track<true, ADJ>(S, m, rng, xf_point(iT, o), xf_vector(iT, dir), h.hit ? h.t : tmax, thr,
                                                     dp3, np3, sc, c.max_null, reads, &sink);
                dirpdf *= avg3(dp3);
                if (sc) {
                    spos = xf_point(S.xf[med], pp);
                    out = make_uint4(__float_as_uint(len(spos - origin)), med | (0xffffu << 16), 0, 0);
                    scattered = done = true;
                    break;
                }
            }
            if (!h.hit) {
                out = make_uint4(__float_as_uint(tmax), 0xffffffffu, 0, 0);
                done = true;
                break;
            }
            if (!is_vol) {
                out = make_uint4(__float_as_uint(h.t), h.id, __float_as_uint(h.b1), __float_as_uint(h.b2));
                done = true;
                break;
            }
            const v3 bp = o + dir * h.t;
            const v3 ng = unpack_normal(vgn);
            med = dot(ng, dir) > 0 ? kNoMedium : (h.id & 0xffffu);
            o = ray_offset(bp, ng, dir);
            tmax -= h.t;
        }
        if (!done) out = make_uint4(__float_as_uint(tmax), 0xffffffffu, 0, 0);
        thr = thr / dirpdf;
        w.hit[i] = out;
        w.thr[p] = make_float4(thr.x, thr.y, thr.z, t4.w);
        s4.x = rng.w;
        w.st[p] = s4;
        w.med[p] = med;
        if (scattered) w.ray[p] = make_float4(spos.x, spos.y, spos.z, r.w);
    }
    wave_add_u64(reinterpret_cast<unsigned long long*>(w.counters + 4), segs);
    if (c.flags & ST_FLAG_COUNT_TRAVERSAL) wave_add_u64(&stats[0], reads);
}

// One visibility ray through surfaces and media: beta is multiplied by the
// ratio-tracking transmittance (0 when a surface blocks), neepdf by the
// tracking pdf.  With ADJ, each collision in the sink's grid adds its term.
template <bool ADJ>
__device__ inline void shadow_media_walk(const DevScene& S, const RenderConsts& c, Rng& rng, uint32_t cur, v3 o, v3 d, float tmax,
                                         v3& beta, float& neepdf, uint32_t* lds, uint32_t* spill, uint32_t stride,
                                         uint32_t gtid, TraverseStats& ts, uint32_t& segs, uint32_t& reads,
                                         const DensitySink* sink) {
    neepdf = 1;
    for (int seg = 0; tmax > 1e-6f; seg++) {
        if (seg == kMaxMediaSegments) {  // bounded like the scattering ray: counts as blocked
            beta = mk(0, 0, 0);
            neepdf = 0;
            return;
        }
        Hit h = traverse<false, false>(S, o, d, tmax, lds, spill, stride, gtid, ts);
        segs++;
        bool is_vol;
        uint32_t vgn = 0;
        volume_boxes(S, o, d, h, is_vol, vgn);
        if (h.hit && !is_vol) {
            beta = mk(0, 0, 0);
            neepdf = 0;
            return;
        }
        if (cur != kNoMedium) {
            const Medium m = load_medium(S, S.inst[cur].type_material_address >> 4);
            const st_transform& iT = S.inv_xf[cur];
            v3 dp3, np3;
            bool sc;
            delta_track<false, ADJ>(S, m, rng, xf_point(iT, o), xf_vector(iT, d), h.hit ? h.t : tmax, beta, dp3, np3, sc,
                                    c.max_null, reads, sink);
            neepdf *= avg3(np3);
            if (!(beta.x != 0 && beta.y != 0 && beta.z != 0)) return;
        }
        if (!h.hit) return;
        const v3 bp = o + d * h.t;
        const v3 ng = unpack_normal(vgn);
        if (dot(ng, d) < 0) {
            cur = h.id & 0xffffu;
            o = ray_offset(bp, -ng);
        } else {
            cur = kNoMedium;
            o = ray_offset(bp, ng);
        }
        tmax -= h.t;
    }
}

// traceVisibilityRay with media (intersection.hlsli:262-338): closest hits,
// any surface blocks, ratio tracking through each medium crossed; the NEE
// contribution is divided by the average tracking pdf (testrenderer.slang:
// 1153-1162).  The reference reads an uninitialised static path state here;
// this uses the vertex's medium and its own random stream keyed like the
// path's (pixel or light dispatch index, sample) and by path length (offset
// 0x40000000 | length << 16).  Light tracing keeps the light path in the
// record's weight slot and adds with weight 1 to the pixel slot.  ADJ: the
// path replay of st_adjoint walks a contributing ray a second time, from the
// same random stream, with the fp32 contribution as the sink's weight.
template <bool ADJ = false>
__global__ void __launch_bounds__(kTraceBlock) k_shadow_media(DevScene S, RenderConsts c, WaveState w, uint32_t* spill,
                                                              unsigned long long* stats, AdjArgs adj) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    const uint32_t n = w.counters[2];
    const uint32_t stride = gridDim.x * blockDim.x, gtid = blockIdx.x * blockDim.x + threadIdx.x;
    TraverseStats ts{0, 0};
    uint32_t segs = 0, reads = 0;
    for (uint32_t i = gtid; i < n; i += stride) {
        const float4 a = w.shadow[3 * (size_t)i], b = w.shadow[3 * (size_t)i + 1], cc = w.shadow[3 * (size_t)i + 2];
        const uint32_t p = __float_as_uint(b.w);
        const uint32_t path = c.light_paths ? __float_as_uint(cc.w) : p;
        const uint32_t plen = w.st[path].y;
        const uint32_t key = 0x40000000u | (plen << 16);
        Rng rng = path_rng(c, path, key);
        v3 beta = mk(cc.x, cc.y, cc.z);
        float neepdf = 1;
        shadow_media_walk<false>(S, c, rng, w.med[path], mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, beta, neepdf, lds, spill,
                                 stride, gtid, ts, segs, reads, nullptr);
        if (!(all_nonpos(beta) || neepdf <= 0)) {
            const float wgt = c.light_paths ? 1.0f : cc.w;
            add_shadow_contribution(w, c.flags, p, beta / neepdf, wgt);
            if (ADJ && adj.grad_grid && wgt > 0) {
                const v3 rw = (beta / neepdf) * wgt;
                const v3 add3 = mk(fmax_(0.0f, rw.x), fmax_(0.0f, rw.y), fmax_(0.0f, rw.z));
                if (!all_nonpos(add3)) {
                    uint32_t px, py;
                    pixel_of(c, p % c.npix, px, py);
                    const DensitySink sink{adj.grad_grid, adj.grad_vol, adj_pixel_weight(adj, py * c.W + px) * add3};
                    Rng rng2 = path_rng(c, path, key);
                    v3 beta2 = mk(cc.x, cc.y, cc.z);
                    float np2 = 1;
                    uint32_t segs2 = 0, reads2 = 0;
                    shadow_media_walk<true>(S, c, rng2, w.med[path], mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, beta2, np2,
                                            lds, spill, stride, gtid, ts, segs2, reads2, &sink);
                }
            }
        }
    }
    wave_add_u64(reinterpret_cast<unsigned long long*>(w.counters + 4) + 1, segs);
    if (c.flags & ST_FLAG_COUNT_TRAVERSAL) wave_add_u64(&stats[2], reads);
}


// Wavefront-aggregated append: one atomic per wave, lane order preserved.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    const uint32_t lane = __lane_id();
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (cnt && (int)lane == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// One light candidate evaluated at a shading point: ReservoirSampleDataDI
// (testrenderer.slang:568-606) and the plain NEE record of ConnectToLight (:751-779).
struct Cand {
    v3 contrib;      // f * Le (ReSTIR) or throughput * Le * f / pdfW (plain NEE)
    v3 dir;
    float dist, src_pdf, wtoa, mis_w;
};

template <bool LAMBERT, bool ALIAS>
__device__ __forceinline__ bool eval_candidate(const DevScene& S, const RenderConsts& c, const Material& m, const Frame& f,
                                               v3 lgn, v3 ldin, v3 pos, float4 rnd, float cp_eff, Cand& x, float& lum_rad) {
    const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rnd, pos);
    x.src_pdf = L.pdf;
    x.dir = L.dir;
    x.dist = L.dist;
    x.contrib = mk(0, 0, 0);
    x.wtoa = 0;
    x.mis_w = 0;
    lum_rad = lum(L.radiance);
    if (L.pdf <= 0) return false;
    const float AtoW = L.finite ? sq(L.dist) / L.cos_light : 1.0f;
    const v3 ldout = f.to_local(L.dir);
    Eval fe = bsdf_eval<LAMBERT>(m, ldin, ldout);
    fe.f = fe.f * shading_normal_correction(f, lgn, ldin, ldout);
    fe.fwd *= cp_eff;
    fe.rev *= cp_eff;
    x.contrib = fe.f * L.radiance;
    x.wtoa = L.finite ? L.cos_light / sq(L.dist) : 1.0f;
    x.mis_w = (c.flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(L.pdf, fe.fwd / AtoW);  // ComputeMIS_DI, :530-566
    x.src_pdf = L.pdf;
    return true;
}

__device__ __forceinline__ float mis_pair(float a, float b) {  // mis(a, b), testrenderer.slang:287-290
    const float a2 = a * a, d = a2 + b * b;
    return d > 0 ? a2 / d : 0.0f;
}

// ------------------------------------------------ hash-grid reservoir reuse
// HashGrid::FindCellIndex's checksum of the cell holding p (hashgrid.hlsli:
// 51-58): the cell's identity in the grid
__device__ __forceinline__ uint32_t hg_checksum(v3 p, float cs) {
    const int x = (int)floorf(p.x / cs), y = (int)floorf(p.y / cs), z = (int)floorf(p.z / cs);
    const uint32_t h = xxhash32(xxhash32(__float_as_uint(cs)) + xxhash32((uint32_t)z + xxhash32((uint32_t)y + xxhash32((uint32_t)x))));
    return h > 1u ? h : 1u;
}
// first index in keys[0, n) whose key is >= k (upper: > k)
__device__ __forceinline__ uint32_t hg_bound(const unsigned long long* keys, uint32_t n, unsigned long long k, bool upper) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const unsigned long long v = keys[mid];
        if (upper ? v <= k : v < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
// The vertex of a hash-grid reuse: where its entry goes and what it holds
struct HgVertex {
    const WaveState* w;
    uint32_t pix, path_len, id;  // linear pixel, path length, instance | primitive << 16
};
// Reuse from the previous sample's grid and append to this sample's
// (SampleIlluminationReSTIR's reuse and store, testrenderer.slang:686-723;
// ReservoirReuse, :608-663).  sel / sel_target / sel_seed / W: the vertex's
// reservoir after RIS, updated in place.
template <bool LAMBERT, bool ALIAS>
__device__ inline void hash_grid_reuse(const DevScene& S, const RenderConsts& c, const HgVertex& h, Rng& rng, const Material& m,
                                       const Frame& f, v3 lgn, v3 ldin, v3 pos, float cp_eff, Cand& sel, float& sel_target,
                                       float4& sel_seed, float& W) {
    float M = (float)c.cand;
    if (c.hg_prev) {
        // a point jittered over one cell on the tangent plane, one entry of its cell
        const float4 dr = rng.nextf();
        float dx, dy;
        concentric_disc(dr.x, dr.y, dx, dy);
        const v3 qp = pos + f.to_world(mk(dx * c.hg_cell, dy * c.hg_cell, 0.0f));
        const unsigned long long key = hg_checksum(qp, c.hg_cell);
        const uint32_t lo = hg_bound(h.w->hg_prev_keys, c.hg_nslots, key, false);
        const uint32_t hi = hg_bound(h.w->hg_prev_keys, c.hg_nslots, key, true);
        if (hi > lo) {
            const uint32_t slot = h.w->hg_prev_slots[lo + rng.next().x % (hi - lo)];
            const float4* e = h.w->hg_prev_rec + 4 * (size_t)slot;
            const float4 e0 = e[0], seed = e[1], e2 = e[2], e3 = e[3];
            // the entry's light sample here
            Cand xn;
            float lrn;
            const bool okn = eval_candidate<LAMBERT, ALIAS>(S, c, m, f, lgn, ldin, pos, seed, cp_eff, xn, lrn);
            const float tn = (okn && lrn > 0) ? lum(xn.contrib) * xn.wtoa : 0.0f;
            // this vertex's light sample at the entry's point (HashGridData::getShadingData)
            float phq = 0;
            {
                const uint32_t id = __float_as_uint(e0.w);
                const v3 lp = mk(e0.x, e0.y, e0.z);
                Shading sq;
                tri_shading(S, id & 0xffffu, id >> 16, 0.0f, 0.0f, sq, &lp);
                const Material mq = load_material<false>(S, sq, c.flags);
                const Frame fq = make_frame(sq);
                Cand xq;
                float lrq;
                const bool okq = eval_candidate<LAMBERT, ALIAS>(S, c, mq, fq, fq.to_local(unpack_normal(sq.gn)), mk(e3.x, e3.y, e3.z),
                                                                sq.pos, sel_seed, 1.0f, xq, lrq);
                if (okq && lrq > 0) phq = lum(xq.contrib) * xq.wtoa;
            }
            // pairwise MIS (a current reservoir of target 0 weighs 0 here, where the
            // reference's 0/0 makes the sum NaN and drops the entry)
            const float qM = e2.z;
            const float m0 = mis_pair(M * sel_target, qM * phq), m1 = mis_pair(qM * e2.x, M * tn);
            const float w0 = m0 * sel_target * W, w1 = m1 * tn * e2.y;
            M += qM;
            const float total = w0 + w1;
            if (rng.nextf().x * total <= w1) {
                sel = xn;
                sel_target = tn;
                sel_seed = seed;
            }
            W = sel_target > 0 ? total / sel_target : 0.0f;
        }
    }
    if (W > 0 && h.path_len - 1 < c.hg_dmax) {
        float4* e = h.w->hg_rec + 4 * ((size_t)(h.path_len - 1) * c.npix + h.pix);
        const v3 lp = xf_point(S.inv_xf[h.id & 0xffffu], pos);
        e[0] = make_float4(lp.x, lp.y, lp.z, __uint_as_float(h.id));
        e[1] = sel_seed;
        e[2] = make_float4(sel_target, W, fmin_(M, c.max_m * (float)c.cand), __uint_as_float(hg_checksum(pos, c.hg_cell)));
        e[3] = make_float4(ldin.x, ldin.y, ldin.z, 0.0f);
    }
}

// Next-event estimation at a surface vertex.  Plain: one light sample
// (ConnectToLight, testrenderer.slang:751-779).  RESTIR: resampled importance
// sampling over c.cand candidates (SampleIlluminationReSTIR, :665-684; the
// ReSTIR ConnectToLight, :730-747).  Returns false when no shadow ray is made.
template <bool LAMBERT, bool ALIAS, bool RESTIR>
__device__ __forceinline__ bool next_event(const DevScene& S, const RenderConsts& c, Rng& rng, const Material& m,
                                           const Frame& f, v3 lgn, v3 ldin, v3 pos, uint32_t gn, v3 thr, float cp_eff,
                                           float4& sh0, float4& sh1, float4& sh2, uint32_t p, const HgVertex* hg = nullptr) {
    Cand sel;
    float wgt_out;
    v3 contrib;
    if (!RESTIR) {
        const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rng.nextf(), pos);
        if (all_nonpos(L.radiance) || L.pdf <= 0) return false;
        const float AtoW = L.finite ? sq(L.dist) / L.cos_light : 1.0f;
        const float dpw = L.pdf * AtoW;
        const v3 ldout = f.to_local(L.dir);
        Eval fe = bsdf_eval<LAMBERT>(m, ldin, ldout);
        fe.f = fe.f * shading_normal_correction(f, lgn, ldin, ldout);
        fe.fwd *= cp_eff;
        fe.rev *= cp_eff;
        contrib = ((thr * L.radiance) * fe.f) / dpw;
        wgt_out = (c.flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(L.pdf, fe.fwd / AtoW);
        sel.dir = L.dir;
        sel.dist = L.dist;
    } else {
        float wsum = 0, sel_target = 0;
        float4 sel_seed = make_float4(0, 0, 0, 0);
        sel.contrib = mk(0, 0, 0);
        sel.dir = mk(0, 0, 0);
        sel.dist = 0;
        sel.src_pdf = 0;
        sel.wtoa = 0;
        sel.mis_w = 0;
        for (uint32_t i = 0; i < c.cand; i++) {
            Cand x;
            float lr;
            const float4 rnd = rng.nextf();
            const bool ok = eval_candidate<LAMBERT, ALIAS>(S, c, m, f, lgn, ldin, pos, rnd, cp_eff, x, lr);
            const float target = (ok && lr > 0) ? lum(x.contrib) * x.wtoa : 0.0f;
            if (target <= 0) continue;
            wsum += target / x.src_pdf;
            if (rng.nextf().x <= (target / x.src_pdf) / wsum) {
                sel = x;
                sel_target = target;
                sel_seed = rnd;
            }
        }
        float W = sel_target > 0 ? wsum / ((float)c.cand * sel_target) : 0.0f;
        if (hg) hash_grid_reuse<LAMBERT, ALIAS>(S, c, *hg, rng, m, f, lgn, ldin, pos, cp_eff, sel, sel_target, sel_seed, W);
        if (W <= 0) return false;
        W *= sel.wtoa;
        contrib = (thr * sel.contrib) * W;
        wgt_out = sel.mis_w;
    }
    const v3 o = ray_offset(pos, unpack_normal(gn), sel.dir);
    const float tmax = fmax_(sel.dist * 0.99f, sel.dist - 1e-2f);  // AdjustIntersectionDistance
    sh0 = make_float4(o.x, o.y, o.z, tmax);
    sh1 = make_float4(sel.dir.x, sel.dir.y, sel.dir.z, __uint_as_float(p));
    sh2 = make_float4(contrib.x, contrib.y, contrib.z, wgt_out);
    return true;
}

// Resampled importance sampling over c.cand candidates (SampleIlluminationReSTIR,
// testrenderer.slang:665-684) keeping the winning seed for reuse.
template <bool LAMBERT, bool ALIAS>
__device__ __forceinline__ void ris_reservoir(const DevScene& S, const RenderConsts& c, Rng& rng, const Material& m,
                                              const Frame& f, v3 lgn, v3 ldin, v3 pos, float cp_eff, float4& seed,
                                              float& target, float& W) {
    float wsum = 0;
    target = 0;
    seed = make_float4(0, 0, 0, 0);
    for (uint32_t i = 0; i < c.cand; i++) {
        const float4 rnd = rng.nextf();
        Cand x;
        float lr;
        const bool ok = eval_candidate<LAMBERT, ALIAS>(S, c, m, f, lgn, ldin, pos, rnd, cp_eff, x, lr);
        const float t = (ok && lr > 0) ? lum(x.contrib) * x.wtoa : 0.0f;
        if (t <= 0) continue;
        wsum += t / x.src_pdf;
        if (rng.nextf().x <= (t / x.src_pdf) / wsum) {
            seed = rnd;
            target = t;
        }
    }
    W = target > 0 ? wsum / ((float)c.cand * target) : 0.0f;
}

// Shading point of a pixel's primary hit, rebuilt from its stored record.
struct Prim {
    Shading sd;
    Material m;
    Frame f;
    v3 lgn, ldin;
};
// The triangle's shading data was computed once, by k_ris, and stored bit for
// bit, so the reuse passes read one 32-byte record instead of the mesh buffers.
__device__ __forceinline__ void rebuild_primary(const DevScene& S, const WaveState& w, size_t idx, Prim& P, uint32_t flags) {
    const uint4 rec = w.pri[idx];
    const float4 a = w.psd[2 * idx], b = w.psd[2 * idx + 1];
    P.sd.pos = mk(a.x, a.y, a.z);
    P.sd.material = __float_as_uint(a.w);
    P.sd.gn = __float_as_uint(b.x);
    P.sd.sn = __float_as_uint(b.y);
    P.sd.tan = __float_as_uint(b.z);
    P.sd.area = b.w;
    P.m = load_material(S, P.sd.material, flags);
    P.f = make_frame(P.sd);
    P.lgn = P.f.to_local(unpack_normal(P.sd.gn));
    P.ldin = P.f.to_local(-unpack_normal(rec.w));
}
// target function of a light-sample seed at a shading point (0 when unusable)
template <bool LAMBERT, bool ALIAS>
__device__ __forceinline__ float seed_target(const DevScene& S, const RenderConsts& c, const Prim& P, float4 seed, Cand& x) {
    float lr;
    const float cp = 1 >= c.rr_start ? m_cont_prob(P.m) : 1.0f;  // primary vertex: path length 1
    const bool ok = eval_candidate<LAMBERT, ALIAS>(S, c, P.m, P.f, P.lgn, P.ldin, P.sd.pos, seed, cp, x, lr);
    return (ok && lr > 0) ? lum(x.contrib) * x.wtoa : 0.0f;
}

// RIS at the primary vertices of the screen-space reuse mode
// (SampleIlluminationReSTIR, testrenderer.slang:665-684), as a lean pass before
// k_shade: it keeps each pixel's reservoir and primary record, and advances
// the path's random counter past the candidates, where k_shade continues the
// path (emission, RR and BSDF sampling draw no numbers before NEE).
// SPH: the scene has sphere instances (their shading inlined only then)
template <bool LAMBERT, bool ALIAS, bool SPH = false>
__global__ void __launch_bounds__(kShadeBlock) k_ris(DevScene S, RenderConsts c, WaveState w, int qsel) {
    const uint32_t n = w.counters[qsel];
    const uint32_t* q = w.queue[qsel];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = q[i];
        const uint4 hit = w.hit[i];
        const uint32_t k = p / c.npix;
        uint32_t px, py;
        pixel_of(c, p % c.npix, px, py);
        const size_t rp = (size_t)k * c.npix + py * c.W + px;  // sample k of the wave
        if (hit.y == 0xffffffffu) {
            w.r_init[2 * rp + 1] = make_float4(0, 0, 0, 0);
            continue;
        }
        const float4 r = w.ray[p];
        uint4 s4 = w.st[p];
        Rng rng{px, py, c.seed + c.sample_base + k, s4.x};
        Shading sd;
        if (SPH) hit_shading(S, hit, mk(r.x, r.y, r.z), unpack_normal(__float_as_uint(r.w)), sd);
        else tri_shading(S, hit.y & 0xffffu, hit.y >> 16, __uint_as_float(hit.z), __uint_as_float(hit.w), sd);
        const Material m = load_material<false>(S, sd, c.flags);
        const Frame f = make_frame(sd);
        const v3 ldin = f.to_local(-unpack_normal(__float_as_uint(r.w)));
        const v3 lgn = f.to_local(unpack_normal(sd.gn));
        const uint32_t path_len = s4.y + 1;
        const bool alive = path_len < c.max_depth && any_pos(m.base());
        const float cp_eff = path_len >= c.rr_start ? m_cont_prob(m) : 1.0f;
        float4 seed = make_float4(0, 0, 0, 0);
        float target = 0, W = 0;
        if (alive && !m_singular<LAMBERT>(m) && (c.flags & ST_FLAG_SAMPLE_DIRECT))
            ris_reservoir<LAMBERT, ALIAS>(S, c, rng, m, f, lgn, ldin, sd.pos, cp_eff, seed, target, W);
        w.pri[rp] = make_uint4(hit.y, hit.z, hit.w, __float_as_uint(r.w));
        w.psd[2 * rp] = make_float4(sd.pos.x, sd.pos.y, sd.pos.z, __uint_as_float(sd.material));
        w.psd[2 * rp + 1] = make_float4(__uint_as_float(sd.gn), __uint_as_float(sd.sn), __uint_as_float(sd.tan), sd.area);
        w.r_init[2 * rp] = seed;
        w.r_init[2 * rp + 1] = make_float4(target, W, target > 0 ? (float)c.cand : 0.0f, __uint_as_float(hit.x));
        s4.x = rng.w;
        w.st[p] = s4;
    }
}

// ADJ: the path replay of st_adjoint, which adds the albedo gradient of each
// Lambertian vertex with a per-triangle albedo (every contribution after its
// emission carries one factor albedo_c: dLoss/dL * (the path's first-pass
// radiance - what it has gathered so far) / albedo_c).
// Training-path selection of st_render_cached: of each run of train_stride
// consecutive paths (g = sample * W * H + pixel) one continues unchanged and
// gives training row g / train_stride.
__device__ __forceinline__ bool train_path(const RenderConsts& c, uint32_t g, uint32_t& row) {
    row = g / c.train_stride;
    return g % c.train_stride == rc_pcg(row ^ rc_pcg(c.train_frame)) % c.train_stride;
}

// The cache-depth vertex of st_render_cached: a training path writes its
// row (query, throughput, radiance gathered so far incl. this vertex's
// emission) and goes on; any other path fills a cache query and stops.
__device__ __forceinline__ bool cache_vertex(const RenderConsts& c, const WaveState& w, uint32_t p, uint32_t g, v3 pos, v3 wo,
                                             v3 nrm, v3 alb, v3 thr, float4& em, bool& em_dirty, float4 cqr[4]) {
    uint32_t row;
    if (train_path(c, g, row)) {
        if (!em_dirty) {
            em = w.em[p];
            em_dirty = true;
        }
        const uint4 nq = w.nee[p];
        float* q = w.tq + 12 * (size_t)row;
        const float v[12] = {pos.x, pos.y, pos.z, wo.x, wo.y, wo.z, nrm.x, nrm.y, nrm.z, alb.x, alb.y, alb.z};
        for (int i = 0; i < 12; i++) q[i] = v[i];
        w.tthr[row] = make_float4(thr.x, thr.y, thr.z, 0.0f);
        w.tt[row] = make_float4(em.x + (float)nq.x / (float)kQuant, em.y + (float)nq.y / (float)kQuant,
                                em.z + (float)nq.z / (float)kQuant, 1.0f);
        return false;
    }
    cqr[0] = make_float4(pos.x, pos.y, pos.z, thr.x);
    cqr[1] = make_float4(wo.x, wo.y, wo.z, thr.y);
    cqr[2] = make_float4(nrm.x, nrm.y, nrm.z, thr.z);
    cqr[3] = make_float4(alb.x, alb.y, alb.z, __uint_as_float(p));
    return true;
}

// TEX: the scene has material images (their sampling compiled only here)
template <bool LAMBERT, bool ALIAS, bool RESTIR, bool MEDIA, bool ADJ = false, bool CACHE = false, bool TEX = false>
__global__ void __launch_bounds__(kShadeBlock) k_shade(DevScene S, RenderConsts c, WaveState w, int qsel,
                                                       AuxOut aux, int first_wave, AdjArgs adj) {
    __shared__ AdjTable adj_tab;  // ADJ only
    if (ADJ) {
        adj_table_clear(adj_tab);
        __syncthreads();
    }
    const uint32_t n = w.counters[qsel];
    const uint32_t* q = w.queue[qsel];
    uint32_t* qout = w.queue[qsel ^ 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    // every lane of a wave must reach the ballots: iterate to a wave-uniform bound
    const uint32_t base0 = blockIdx.x * blockDim.x + threadIdx.x - __lane_id();
    for (uint32_t i0 = base0; i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        const bool valid = i < n;
        bool emit_shadow = false, cont = false;
        float4 sh0 = make_float4(0, 0, 0, 0), sh1 = sh0, sh2 = sh0;
        uint32_t p = 0;
        float4 ray_out = sh0, thr_out = sh0;
        uint32_t adj_key = 0xffffffffu;  // ADJ: this vertex's albedo entry and gradient
        v3 adj_val = mk(0, 0, 0);
        bool to_cache = false;  // CACHE: this path ends in a cache query
        float4 cqr[4];
        if (valid) {
            p = q[i];
            const uint4 hit = w.hit[i];
            const float4 r = w.ray[p];
            float4 thr4 = w.thr[p];
            uint4 s4 = w.st[p];
            const uint32_t k = p / c.npix;
            uint32_t px, py;
            pixel_of(c, p % c.npix, px, py);
            const uint32_t pix = py * c.W + px;
            Rng rng{px, py, c.seed + c.sample_base + k, s4.x};
            uint32_t path_len = s4.y + 1;  // TraceRay, testrenderer.slang:1026-1058
            const v3 dir = unpack_normal(__float_as_uint(r.w));
            v3 thr = mk(thr4.x, thr4.y, thr4.z);
            const bool hit_any = hit.y != 0xffffffffu;
            float dirpdf = h2f(s4.w);
            float4 em = make_float4(0, 0, 0, 0);
            bool em_dirty = false;
            // AddEmission, testrenderer.slang:788-808 (+ --minPathLength gate, vcm.slang:960)
            auto add_emission = [&](v3 e, float dpa, float WtoA) {
                float wgt = 1;
                if (path_len > 1 && (c.flags & ST_FLAG_SAMPLE_DIRECT) && dirpdf > 0) wgt = mis2(dirpdf * WtoA, dpa);
                if (path_len < c.min_len) return;
                if (wgt <= 0 || all_nonpos(e)) return;
                if (!em_dirty) {
                    em = w.em[p];
                    em_dirty = true;
                }
                const v3 a = e * wgt;
                em.x += a.x;
                em.y += a.y;
                em.z += a.z;
            };
            if (!hit_any) {
                if (path_len == 1 && first_wave && k == 0)  // a miss stores albedo 1 (TraceRay, :1040-1041)
                    store_aux(c, aux, pix, px, py, 0xffffffffu, pack_normal(dir), false, mk(0, 0, 0), mk(1, 1, 1),
                              q16(__int_as_float(0x7f800000)));
                float dpw;
                const v3 e = thr * evaluate_environment<ALIAS>(S, c.env_prob, dir, dpw);
                if (any_pos(e)) {
                    float dpa = (1.0f / ST_PI) / sq(S.sphere[3]);  // concentricDiscPdfA()/R^2
                    if (c.flags & ST_FLAG_ENV_MIS_BALANCED) dpa = dpw;
                    add_emission(e, dpa, 1);
                }
                s4.y = path_len;
            } else if (MEDIA && (hit.y >> 16) == 0xffffu) {
                // medium scattering vertex: ProcessInteraction<false>(Medium) (testrenderer.slang:816-935)
                const uint32_t ii = hit.y & 0xffffu;
                const v3 pos = mk(r.x, r.y, r.z);  // written by k_trace_media
                float pdfa = 0;  // pickPdf 0 over shape area 0 (intersection.hlsli:232-236)
                if (S.light_count > 0) {
                    pdfa = 0.0f / (0.0f * (float)S.light_count);
                    if (S.has_env) pdfa *= 1 - c.env_prob;
                }
                const float t16 = q16(__uint_as_float(hit.x)), pdfa16 = q16(pdfa);
                if (!all_nonpos(thr)) {  // ExtendPath, testrenderer.slang:1062-1063
                    const Medium m = load_medium(S, S.inst[ii].type_material_address >> 4);
                    if (path_len == 1 && first_wave && k == 0)
                        store_aux(c, aux, pix, px, py, hit.y, 0, true, pos, m.albedo(), t16);
                    const v3 ldin = -dir;
                    bool alive = path_len < c.max_depth && m.can_eval();
                    if (CACHE && alive && path_len == c.cache_depth) {
                        to_cache = cache_vertex(c, w, p, (c.sample_base + k) * c.npix + pix, pos, ldin, mk(0, 0, 0), m.albedo(),
                                                thr, em, em_dirty, cqr);
                        if (to_cache) alive = false;
                    }
                    const bool rr_on = path_len >= c.rr_start;
                    if (RESTIR && alive && !m.singular() && (c.flags & ST_FLAG_SAMPLE_DIRECT)) {
                        // SampleIlluminationReSTIR at a medium vertex (testrenderer.slang:568-606,
                        // 665-684, 730-747): the phase function, cosOut 1, no offset
                        const float cpm = rr_on ? m.cont_prob() : 1.0f;
                        float wsum = 0, sel_t = 0, sel_pdf = 0, sel_wtoa = 0, sel_mis = 0, sel_dist = 0;
                        v3 sel_c = mk(0, 0, 0), sel_dir = mk(0, 0, 0);
                        for (uint32_t ci = 0; ci < c.cand; ci++) {
                            const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rng.nextf(), pos);
                            if (!(lum(L.radiance) > 0 && L.pdf > 0)) continue;
                            const float v = hg_eval(m.g, ldin, L.dir);
                            const float AtoW = L.finite ? sq(L.dist) / L.cos_light : 1.0f;
                            const v3 xc = mk(v, v, v) * L.radiance;
                            const float xw = L.finite ? L.cos_light / sq(L.dist) : 1.0f;
                            const float target = lum(xc) * xw;
                            if (target <= 0) continue;
                            wsum += target / L.pdf;
                            if (rng.nextf().x <= (target / L.pdf) / wsum) {
                                sel_t = target;
                                sel_pdf = L.pdf;
                                sel_c = xc;
                                sel_wtoa = xw;
                                sel_dir = L.dir;
                                sel_dist = L.dist;
                                sel_mis = (c.flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(L.pdf, (v * cpm) / AtoW);
                            }
                        }
                        const float Wr = sel_t > 0 ? wsum / ((float)c.cand * sel_t) : 0.0f;
                        if (Wr > 0 && path_len + 1 >= c.min_len) {
                            const v3 contrib = (thr * sel_c) * (Wr * sel_wtoa);
                            sh0 = make_float4(pos.x, pos.y, pos.z, fmax_(sel_dist * 0.99f, sel_dist - 1e-2f));
                            sh1 = make_float4(sel_dir.x, sel_dir.y, sel_dir.z, __uint_as_float(p));
                            sh2 = make_float4(contrib.x, contrib.y, contrib.z, sel_mis);
                            emit_shadow = true;
                        }
                        (void)sel_pdf;
                    } else if (alive && !m.singular() && (c.flags & ST_FLAG_SAMPLE_DIRECT)) {
                        // ConnectToLight (:751-779) with the phase function; cosOut = 1
                        const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rng.nextf(), pos);
                        if (!(all_nonpos(L.radiance) || L.pdf <= 0)) {
                            const float AtoW = L.finite ? sq(L.dist) / L.cos_light : 1.0f;
                            const float dpw = L.pdf * AtoW;
                            const float v = hg_eval(m.g, ldin, L.dir);
                            const float fwd = v * (rr_on ? m.cont_prob() : 1.0f);
                            const v3 contrib = ((thr * L.radiance) * v) / dpw;
                            const float wgt = (c.flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(L.pdf, fwd / AtoW);
                            if (path_len + 1 >= c.min_len) {
                                // rayOffset of a medium vertex is the vertex itself (intersection.hlsli:33-35)
                                sh0 = make_float4(pos.x, pos.y, pos.z, fmax_(L.dist * 0.99f, L.dist - 1e-2f));
                                sh1 = make_float4(L.dir.x, L.dir.y, L.dir.z, __uint_as_float(p));
                                sh2 = make_float4(contrib.x, contrib.y, contrib.z, wgt);
                                emit_shadow = true;
                            }
                        }
                    }
                    if (c.flags & ST_FLAG_DIRECT_ONLY) alive = false;
                    if (alive) {
                        const float cp = fmin_(1.0f, m.cont_prob());
                        const float rr = rng.nextf().x;
                        if (rr_on && rr > cp) alive = false;
                        if (alive) {
                            const float4 rb = rng.nextf();
                            const v3 sdir = hg_sample(m.g, rb.x, rb.y, ldin);
                            const float v = hg_eval(m.g, ldin, sdir);
                            const float fwd = v * (rr_on ? cp : 1.0f);
                            thr = thr * (v / fwd);
                            dirpdf = (1 - fabsf(m.g)) == 0 ? 0.0f : fwd;
                            const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
                            cont = !nan && !all_nonpos(thr) && path_len + 1 <= c.max_depth;
                            ray_out = make_float4(pos.x, pos.y, pos.z, __uint_as_float(pack_normal(sdir)));
                            thr_out = make_float4(thr.x, thr.y, thr.z, 0);
                            s4.w = f2h(dirpdf);
                        }
                    }
                }
                s4.y = path_len;
                s4.z = f2h(t16) | (f2h(pdfa16) << 16);
            } else {
                const uint32_t ii = hit.y & 0xffffu, prim = hit.y >> 16;
                Shading sd;
                hit_shading(S, hit, mk(r.x, r.y, r.z), dir, sd);
                const float pick = prim_pick(S, hit.y);
                float pdfa = 0;  // LightSamplePdfA, lights.hlsli:264-271
                if (S.light_count > 0) {
                    pdfa = pick / (sd.area * (float)S.light_count);
                    if (S.has_env) pdfa *= 1 - c.env_prob;
                }
                // with media, the distance is measured from the ray origin (intersection.hlsli:244)
                const float thit = MEDIA ? len(sd.pos - mk(r.x, r.y, r.z)) : __uint_as_float(hit.x);
                const float t16 = q16(thit), pdfa16 = q16(pdfa);
                const bool live = !MEDIA || !all_nonpos(thr);  // ExtendPath, testrenderer.slang:1062-1063
                Material m = load_material<TEX>(S, sd, c.flags);
                uint32_t alb_entry = 0xffffffffu;
                if (LAMBERT && S.tri_alb) apply_triangle_albedo(S, m, ii, prim, alb_entry);
                if (live && path_len == 1 && first_wave && k == 0)
                    store_aux(c, aux, pix, px, py, hit.y, sd.sn, true, sd.pos, m.base(), t16);
                const Frame f = make_frame(sd);
                const v3 ldin = f.to_local(-dir);
                const v3 lgn = f.to_local(unpack_normal(sd.gn));
                const float cos_in = fabsf(ldin.z);
                if (live && ldin.z > 0) {
                    const v3 e = thr * m.emission();
                    if (any_pos(e)) add_emission(e, pdfa16, cos_in / sq(t16));
                }
                if (ADJ && live && alb_entry != 0xffffffffu && adj.grad_alb) {
                    if (!em_dirty) {
                        em = w.em[p];
                        em_dirty = true;
                    }
                    const float4 lt = adj.ltot[p];
                    const v3 wg = adj_pixel_weight(adj, pix) * mk(lt.x - em.x, lt.y - em.y, lt.z - em.z);
                    adj_key = alb_entry;
                    adj_val = mk(m.bo.x > 0 ? wg.x / m.bo.x : 0.0f, m.bo.y > 0 ? wg.y / m.bo.y : 0.0f,
                                 m.bo.z > 0 ? wg.z / m.bo.z : 0.0f);
                }
                bool alive = live && path_len < c.max_depth && any_pos(m.base());
                if (CACHE && alive && path_len == c.cache_depth) {
                    to_cache = cache_vertex(c, w, p, (c.sample_base + k) * c.npix + pix, sd.pos, -dir, unpack_normal(sd.sn),
                                            m.base(), thr, em, em_dirty, cqr);
                    if (to_cache) alive = false;
                }
                // Russian roulette from path length rr_start (0: every bounce, testrenderer.slang:890-892)
                const bool rr_on = path_len >= c.rr_start;
                const float cp_eff = rr_on ? m_cont_prob(m) : 1.0f;
                // primary vertex of the screen-space reuse mode: k_ris kept the RIS
                // reservoir (and advanced the random counter); the shadow ray is
                // made after temporal + spatial reuse
                const bool reuse = (c.flags & ST_FLAG_RESTIR_DI_REUSE) && path_len == 1;
                if (!reuse && alive && !m_singular<LAMBERT>(m) && (c.flags & ST_FLAG_SAMPLE_DIRECT)) {
                    // the screen-space reuse mode keeps reservoirs at primary vertices;
                    // deeper vertices run RIS with ST_FLAG_RESTIR_DI_DEEP, else plain NEE
                    const bool ris_here = RESTIR && (!(c.flags & ST_FLAG_RESTIR_DI_REUSE) || (c.flags & ST_FLAG_RESTIR_DI_DEEP));
                    const HgVertex hgv{&w, pix, path_len, hit.y};
                    const HgVertex* hgp = (RESTIR && (c.flags & ST_FLAG_RESTIR_DI_HASHGRID)) ? &hgv : nullptr;
                    emit_shadow = (ris_here ? next_event<LAMBERT, ALIAS, true>(S, c, rng, m, f, lgn, ldin, sd.pos, sd.gn, thr,
                                                                              cp_eff, sh0, sh1, sh2, p, hgp)
                                            : next_event<LAMBERT, ALIAS, false>(S, c, rng, m, f, lgn, ldin, sd.pos, sd.gn, thr,
                                                                               cp_eff, sh0, sh1, sh2, p)) &&
                                  path_len + 1 >= c.min_len;
                }
                if (c.flags & ST_FLAG_DIRECT_ONLY) alive = false;  // gSampleDirectIlluminationOnly (:886-888)
                if (alive) {
                    const float cp = fmin_(1.0f, m_cont_prob(m));
                    const float rr = rng.nextf().x;
                    if (rr_on && rr > cp) alive = false;
                    if (alive) {
                        const float cpx = rr_on ? cp : 1.0f;
                        const float4 rb = rng.nextf();
                        Sample s = bsdf_sample<LAMBERT>(m, rb.x, rb.y, rb.z, ldin);
                        s.fwd *= cpx;
                        s.rev *= cpx;
                        thr = thr * (s.f / s.fwd);
                        dirpdf = s.rough == 0 ? 0.0f : s.fwd;
                        thr = thr * shading_normal_correction(f, lgn, ldin, s.dir);
                        const v3 wd = f.to_world(s.dir);
                        const uint32_t pd = pack_normal(wd);
                        const v3 o = ray_offset(sd.pos, unpack_normal(sd.gn), wd);
                        const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
                        // the next RenderIteration skips zero-throughput paths (testrenderer.slang:1137)
                        cont = !nan && !all_nonpos(thr) && path_len + 1 <= c.max_depth;
                        ray_out = make_float4(o.x, o.y, o.z, __uint_as_float(pd));
                        thr_out = make_float4(thr.x, thr.y, thr.z, 0);
                        s4.w = f2h(dirpdf);
                    }
                }
                s4.y = path_len;
                s4.z = f2h(t16) | (f2h(pdfa16) << 16);
            }
            s4.x = rng.w;
            w.st[p] = s4;
            if (em_dirty) w.em[p] = em;
            if (cont) {
                w.ray[p] = ray_out;
                w.thr[p] = thr_out;
            }
        }
        if (ADJ) wave_add_keyed3(adj_tab, adj.grad_alb, adj_key, adj_val);
        if (CACHE) {
            const uint32_t ci = wave_append(&w.counters[3], to_cache);
            if (to_cache)
                for (int j = 0; j < 4; j++) w.cq[4 * (size_t)ci + j] = cqr[j];
        }
        const uint32_t si = wave_append(&w.counters[2], emit_shadow);
        if (emit_shadow) {
            w.shadow[3 * (size_t)si] = sh0;
            w.shadow[3 * (size_t)si + 1] = sh1;
            w.shadow[3 * (size_t)si + 2] = sh2;
        }
        const uint32_t qi = wave_append(&w.counters[qsel ^ 1], cont);
        if (cont) qout[qi] = p;
    }
    if (ADJ) {
        __syncthreads();
        if (adj.grad_alb) adj_table_flush(adj_tab, adj.grad_alb);
    }
}

// Temporal reuse: combine this frame's RIS reservoir with the previous frame's
// final reservoir of the same pixel (ReservoirReuse, testrenderer.slang:608-663).
// The camera is static and unjittered, so both live in the same domain.
template <bool LAMBERT, bool ALIAS>
// j: the sample of the wave (its reservoirs at j * npix); r_prev holds the
// final reservoirs of the sample before it.
__global__ void __launch_bounds__(kShadeBlock) k_restir_temporal(DevScene S, RenderConsts c, WaveState w, int has_prev,
                                                                 uint32_t j) {
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < c.npix; pix += gridDim.x * blockDim.x) {
        const size_t rp = (size_t)j * c.npix + pix;
        float4 seed = w.r_init[2 * rp], b = w.r_init[2 * rp + 1];
        if (b.z > 0 && has_prev) {
            const float4 qs = w.r_prev[2 * pix], qb = w.r_prev[2 * pix + 1];
            if (qb.z > 0) {
                Prim P;
                rebuild_primary(S, w, rp, P, c.flags);
                Cand x;
                const float tq = seed_target<LAMBERT, ALIAS>(S, c, P, qs, x);  // q's seed here
                const uint32_t px = pix % c.W, py = pix / c.W;
                Rng rng{px, py, c.seed + c.sample_base + j, 1u << 20};
                const float m0 = mis_pair(b.z * b.x, qb.z * b.x), m1 = mis_pair(qb.z * qb.x, b.z * tq);
                const float w0 = m0 * b.x * b.y, w1 = m1 * tq * qb.y;
                const float total = w0 + w1;
                float M = b.z + qb.z, target = b.x;
                if (rng.nextf().x * total <= w1 && tq > 0) {
                    seed = qs;
                    target = tq;
                }
                const float Wn = target > 0 ? total / target : 0.0f;
                b = make_float4(target, Wn, fmin_(M, c.max_m * (float)c.cand), b.w);
            }
        }
        w.r_temp[2 * rp] = seed;
        w.r_temp[2 * rp + 1] = b;
    }
}

// Spatial reuse over c.spatial neighbours in a disc of c.radius pixels, then the
// shadow ray of the final sample (ConnectToLight with ReSTIR, :730-747).
template <bool LAMBERT, bool ALIAS>
__global__ void __launch_bounds__(kShadeBlock) k_restir_spatial(DevScene S, RenderConsts c, WaveState w, uint32_t j) {
    const size_t jb = (size_t)j * c.npix;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t base0 = blockIdx.x * blockDim.x + threadIdx.x - __lane_id();
    for (uint32_t i0 = base0; i0 < c.npix; i0 += stride) {
        const uint32_t pix = i0 + __lane_id();
        bool emit = false;
        float4 sh0 = make_float4(0, 0, 0, 0), sh1 = sh0, sh2 = sh0;
        if (pix < c.npix) {
            float4 seed = w.r_temp[2 * (jb + pix)], b = w.r_temp[2 * (jb + pix) + 1];
            if (b.z > 0) {
                Prim P;
                rebuild_primary(S, w, jb + pix, P, c.flags);
                const v3 nc = unpack_normal(P.sd.sn);
                Cand cur;
                seed_target<LAMBERT, ALIAS>(S, c, P, seed, cur);
                const uint32_t px = pix % c.W, py = pix / c.W;
                Rng rng{px, py, c.seed + c.sample_base + j, (1u << 20) + 64};
                for (uint32_t n = 0; n < c.spatial; n++) {
                    const float4 u = rng.nextf();
                    float dx, dy;
                    concentric_disc(u.x, u.y, dx, dy);
                    const int qx = (int)px + (int)floorf(dx * c.radius + 0.5f), qy = (int)py + (int)floorf(dy * c.radius + 0.5f);
                    if (qx < 0 || qy < 0 || qx >= (int)c.W || qy >= (int)c.H || (qx == (int)px && qy == (int)py)) continue;
                    const uint32_t qpix = (uint32_t)qy * c.W + (uint32_t)qx;
                    const float4 qs = w.r_temp[2 * (jb + qpix)], qb = w.r_temp[2 * (jb + qpix) + 1];
                    if (!(qb.z > 0)) continue;
                    // similar surfaces only: depth within 10%, normals within ~25 degrees
                    if (fabsf(qb.w - b.w) > 0.1f * b.w) continue;
                    Prim Q;
                    rebuild_primary(S, w, jb + qpix, Q, c.flags);
                    if (dot(unpack_normal(Q.sd.sn), nc) < 0.9f) continue;
                    Cand xn, xq;
                    const float tn = seed_target<LAMBERT, ALIAS>(S, c, P, qs, xn);     // neighbour's seed here
                    const float tc_q = seed_target<LAMBERT, ALIAS>(S, c, Q, seed, xq); // our seed there
                    const float m0 = mis_pair(b.z * b.x, qb.z * tc_q), m1 = mis_pair(qb.z * qb.x, b.z * tn);
                    const float w0 = m0 * b.x * b.y, w1 = m1 * tn * qb.y;
                    const float total = w0 + w1;
                    if (u.z * total <= w1 && tn > 0) {
                        seed = qs;
                        b.x = tn;
                        cur = xn;
                    }
                    b.y = b.x > 0 ? total / b.x : 0.0f;
                    b.z = fmin_(b.z + qb.z, c.max_m * (float)c.cand);
                }
                if (b.y > 0 && b.x > 0 && 2 >= c.min_len) {
                    const float Wp = b.y * cur.wtoa;  // convert 1/pdfA to 1/pdfW
                    const v3 contrib = cur.contrib * Wp;  // primary throughput is 1
                    const v3 o = ray_offset(P.sd.pos, unpack_normal(P.sd.gn), cur.dir);
                    const float tmax = fmax_(cur.dist * 0.99f, cur.dist - 1e-2f);
                    const uint32_t p = j * c.npix + path_of_pixel(c, px, py);
                    sh0 = make_float4(o.x, o.y, o.z, tmax);
                    sh1 = make_float4(cur.dir.x, cur.dir.y, cur.dir.z, __uint_as_float(p));
                    sh2 = make_float4(contrib.x, contrib.y, contrib.z, cur.mis_w);
                    emit = true;
                }
            } else {
                b = make_float4(0, 0, 0, 0);
            }
            w.r_prev[2 * pix] = seed;
            w.r_prev[2 * pix + 1] = b;
        }
        const uint32_t si = wave_append(&w.counters[2], emit);
        if (emit) {
            w.shadow[3 * (size_t)si] = sh0;
            w.shadow[3 * (size_t)si + 1] = sh1;
            w.shadow[3 * (size_t)si + 2] = sh2;
        }
    }
}

// ------------------------------------------------------------ light tracing
// Light path p of a wave is light subpath p % light_paths of sample
// p / light_paths; its random stream starts at offset 0xFFFF (path_rng).

// InitializeLightPath (testrenderer.slang:983-1024): SampleIllumination(rnd, 0)
// picks the emitter; a triangle emits cosine-distributed about its geometry
// normal from the offset sample point, the environment emits along -d from a
// disc of the scene sphere's radius facing d (throughput Le / pdfW, no disc
// pdf, as the reference leaves it).  A sample with zero pdf (a one-sided light
// facing away from the origin, or no light) or a NaN throughput starts no path:
// the reference carries such a path on with a NaN throughput whose
// contributions clamp to zero.  Also clears the wave's pixel accumulators.
template <bool ALIAS>
__global__ void __launch_bounds__(kShadeBlock) k_light_init(DevScene S, RenderConsts c, WaveState w) {
    const uint32_t n = c.wave_paths, nacc = c.k * c.npix, stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nacc; i += stride) {
        w.em[i] = make_float4(0, 0, 0, 0);
        w.nee[i] = make_uint4(0, 0, 0, 0);
    }
    const uint32_t base0 = blockIdx.x * blockDim.x + threadIdx.x - __lane_id();
    for (uint32_t i0 = base0; i0 < n; i0 += stride) {
        const uint32_t p = i0 + __lane_id();
        bool live = false;
        if (p < n) {
            Rng rng = path_rng(c, p, 0xFFFFu);
            const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rng.nextf(), mk(0, 0, 0));
            const float4 dr = rng.nextf();
            v3 thr = L.radiance / L.pdf, o, d;
            float dirpdf, direct_pdfa = 1, emission_pdfw = 1, cos_light = 1;
            if (L.finite) {
                const v3 ld = sample_cos_hemisphere(dr.x, dr.y);
                const v3 nrm = unpack_normal(L.gn);
                v3 t0, t1;
                make_onb(nrm, t0, t1);
                o = ray_offset(L.pos, nrm);
                d = normalize(t0 * ld.x + t1 * ld.y + nrm * ld.z);
                dirpdf = q16(cos_hemisphere_pdfw(ld.z));  // mDirPdfW is stored in fp16
                thr = thr * (ld.z / dirpdf);
                direct_pdfa = L.pdf;
                emission_pdfw = L.pdf * dirpdf;
                cos_light = ld.z;
            } else {
                float x, y;
                concentric_disc(dr.x, dr.y, x, y);
                const v3 nd = -L.dir;
                v3 t0, t1;
                make_onb(nd, t0, t1);
                const float R = S.sphere[3];
                o = mk(S.sphere[0], S.sphere[1], S.sphere[2]) + (t0 * x + t1 * y + L.dir) * R;
                d = nd;
                dirpdf = q16(L.pdf);
            }
            const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
            live = L.pdf > 0 && !nan && any_pos(thr) && c.max_depth > 0;
            if (c.flags & ST_FLAG_BDPT_LVC) w.lv_cnt[p] = 0;
            if (live) {
                w.ray[p] = make_float4(o.x, o.y, o.z, __uint_as_float(pack_normal(d)));
                w.thr[p] = make_float4(thr.x, thr.y, thr.z, 0);
                // bidirectional mode: light sub-path MIS quantities (testrenderer.slang:1016-1019;
                // light samples are never singular, lights.hlsli:168) and isFiniteLight in bit 31
                if (c.flags & ST_FLAG_BDPT_LVC)
                    w.vcm[p] = make_float2(qvc(c, mis1(cos_light / emission_pdfw)), qvc(c, mis1(direct_pdfa / emission_pdfw)));
                w.st[p] = make_uint4(rng.w, 0, 0, f2h(dirpdf) | (L.finite ? 0x80000000u : 0u));
                // a finite light's medium (GetMediumIndex); environment paths start outside
                w.med[p] = (L.finite && S.vinst_count > 0) ? medium_at(S, L.pos) : kNoMedium;
            }
        }
        const uint32_t qi = wave_append(&w.counters[0], live);
        if (live) w.queue[0][qi] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int i = 0; i < 2 * (int)kXcds; i++) w.counters[kFetchClosest + i * kFetchStride] = 0;
}

// Sub-path MIS quantities after a BSDF sample (ProcessInteraction,
// testrenderer.slang:909-922): [tech. rep. (53)-(55)] for a singular sample,
// else (34)-(36).  vcm = {dVC, dVCM}.
__device__ __forceinline__ void vc_after_sample(const RenderConsts& c, float2& vcm, const Sample& sm, float cos_out) {
    if (sm.rough == 0) {
        vcm.y = qvc(c, 0.0f);
        vcm.x = qvc(c, vcm.x * mis1(cos_out));
    } else {
        vcm.x = qvc(c, mis1(cos_out / sm.fwd) * (vcm.x * mis1(sm.rev) + vcm.y + 0));
        vcm.y = qvc(c, mis1(1.0f / sm.fwd));
    }
}

// ConnectToCamera (testrenderer.slang:441-524).  camera_link projects a vertex
// (false when the camera does not see it); camera_record takes the BSDF value f
// toward the camera and |cos| cos_out there, converts image-plane to surface
// density, divides by the number of light paths and fills the shadow record:
// it ends at the camera, unshortened, and adds to the pixel the vertex projects
// to, int(y) * W + int(x) (one beyond the image is dropped).
struct CamLink {
    float cos_cam;  // |cos| at the camera of the direction to the vertex
    float pcx, pcy; // raster position
};
__device__ __forceinline__ bool camera_link(const RenderConsts& c, v3 pos, CamLink& l) {
    const v3 pc = xf_point(c.inv_cam, pos);
    l.pcx = l.pcy = 0;
    if (!to_raster(c.view, pc, l.pcx, l.pcy)) return false;
    l.cos_cam = fabsf(normalize(pc).z);
    return true;
}
// In the bidirectional mode (vc) the record carries the MIS weight [tech. rep.
// (46)] from the vertex's dVC, dVCM and the reverse pdf frev; else the light path.
__device__ __forceinline__ bool camera_record(const RenderConsts& c, uint32_t p, v3 thr, v3 f, float cos_cam, float cos_out,
                                              float dist, v3 origin, v3 tc, float pcx, float pcy, float4& sh0, float4& sh1,
                                              float4& sh2, bool vc = false, float2 vcm = make_float2(0, 0),
                                              float frev = 0) {
    const float ipcd = c.ipd / cos_cam;
    const float i2s = (sq(ipcd) / cos_cam) * cos_out / sq(dist);
    const float s2i = 1.0f / i2s;
    const v3 contrib = (thr * f) / ((float)c.light_paths * s2i * cos_out);
    const uint32_t pix = (uint32_t)(int32_t)pcy * c.W + (uint32_t)(int32_t)pcx;
    if (pix >= c.npix) return false;
    const uint32_t slot = (p / c.light_paths) * c.npix + path_of_pixel(c, pix % c.W, pix / c.W);
    float tag = __uint_as_float(p);
    if (vc) {
        const float wl = mis1(i2s / (float)c.light_paths) * (0 + vcm.y + vcm.x * mis1(frev));
        tag = 1.f / (wl + 1.f);
    }
    sh0 = make_float4(origin.x, origin.y, origin.z, dist);
    sh1 = make_float4(tc.x, tc.y, tc.z, __uint_as_float(slot));
    sh2 = make_float4(contrib.x, contrib.y, contrib.z, tag);
    return true;
}

// ProcessInteraction<true> (testrenderer.slang:810-935) at the vertices of the
// queued light paths: a vertex that can scatter and is not singular connects
// to the camera (the adjoint BSDF, or the phase function in a medium, toward
// it), then Russian roulette and an adjoint BSDF (or phase) sample extend the
// path.  A miss ends a light path without emission.  MEDIA: scenes with
// volumes, whose scattering vertices k_trace_media reports as id
// medium | 0xffff << 16 at the position it wrote into the ray origin.
// VC: the light pass of the bidirectional mode (surfaces only): the sub-path MIS
// quantities are completed at each hit (testrenderer.slang:812-824), each
// non-singular vertex short of max_depth - 1 segments is stored in its
// subpath's slots (StoreLightVertex, :329-341), and the camera connection is
// MIS-weighted.
template <bool LAMBERT, bool MEDIA, bool VC>
__global__ void __launch_bounds__(kShadeBlock) k_shade_light(DevScene S, RenderConsts c, WaveState w, int qsel) {
    const uint32_t n = w.counters[qsel];
    const uint32_t* q = w.queue[qsel];
    uint32_t* qout = w.queue[qsel ^ 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t base0 = blockIdx.x * blockDim.x + threadIdx.x - __lane_id();
    const st_transform& T = c.view.camera_to_world;
    const v3 cam = mk(T.m[0][3], T.m[1][3], T.m[2][3]);  // transformPoint(0) of the view transform
    for (uint32_t i0 = base0; i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        bool emit_shadow = false, cont = false;
        float4 sh0 = make_float4(0, 0, 0, 0), sh1 = sh0, sh2 = sh0, ray_out = sh0, thr_out = sh0;
        float2 vcm_out = make_float2(0, 0);
        uint32_t p = 0;
        if (i < n) {
            p = q[i];
            const uint4 hit = w.hit[i];
            const float4 r = w.ray[p];
            const float4 thr4 = w.thr[p];
            v3 thr = mk(thr4.x, thr4.y, thr4.z);
            // ExtendPath (testrenderer.slang:1062-1063): tracking through media
            // may have zeroed the throughput
            if (hit.y != 0xffffffffu && (!MEDIA || !all_nonpos(thr))) {
                uint4 s4 = w.st[p];
                Rng rng = path_rng(c, p, s4.x);
                const uint32_t path_len = s4.y + 1;
                const v3 dir = unpack_normal(__float_as_uint(r.w));
                const bool rr_on = path_len >= c.rr_start;
                if (MEDIA && (hit.y >> 16) == 0xffffu) {
                    const Medium m = load_medium(S, S.inst[hit.y & 0xffffu].type_material_address >> 4);
                    const v3 pos = mk(r.x, r.y, r.z);
                    const v3 ldin = -dir;  // media work in world space
                    bool alive = path_len < c.max_depth && m.can_eval();
                    CamLink l;
                    if (alive && !m.singular() && path_len + 1 >= c.min_len && camera_link(c, pos, l)) {
                        v3 tc = cam - pos;
                        const float dist = len(tc);
                        tc = tc / dist;
                        const float v = hg_eval(m.g, ldin, tc);
                        // rayOffset of a medium vertex is the vertex itself (intersection.hlsli:33-35)
                        emit_shadow = camera_record(c, p, thr, mk(v, v, v), l.cos_cam, 1.0f, dist, pos, tc, l.pcx, l.pcy,
                                                    sh0, sh1, sh2);
                    }
                    if (alive) {
                        const float cp = fmin_(1.0f, m.cont_prob());
                        const float rr = rng.nextf().x;
                        if (rr_on && rr > cp) alive = false;
                        if (alive) {
                            const float4 rb = rng.nextf();
                            const v3 sdir = hg_sample(m.g, rb.x, rb.y, ldin);
                            const float v = hg_eval(m.g, ldin, sdir);
                            const float fwd = v * (rr_on ? cp : 1.0f);
                            thr = thr * (v / fwd);
                            const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
                            cont = !nan && !all_nonpos(thr) && path_len + 1 <= c.max_depth;
                            ray_out = make_float4(pos.x, pos.y, pos.z, __uint_as_float(pack_normal(sdir)));
                            thr_out = make_float4(thr.x, thr.y, thr.z, 0);
                        }
                    }
                } else {
                    Shading sd;
                    hit_shading(S, hit, mk(r.x, r.y, r.z), dir, sd);
                    const Material m = load_material<false>(S, sd, c.flags);
                    const Frame f = make_frame(sd);
                    const v3 ldin = f.to_local(-dir);
                    const v3 lgn = f.to_local(unpack_normal(sd.gn));
                    const v3 gn = unpack_normal(sd.gn);
                    float2 vcm = make_float2(0, 0);
                    if (VC) {
                        const float cos_in = fabsf(ldin.z), t16 = q16(__uint_as_float(hit.x));
                        vcm = w.vcm[p];
                        if (path_len > 1 || (s4.w >> 31)) vcm.y = qvc(c, vcm.y * mis1(sq(t16)));
                        vcm.y = qvc(c, vcm.y / mis1(cos_in));
                        vcm.x = qvc(c, vcm.x / mis1(cos_in));
                    }
                    bool alive = path_len < c.max_depth && any_pos(m.base());
                    if (VC && alive && !m_singular<LAMBERT>(m) && path_len + 2 <= c.max_depth) {
                        const uint32_t j = w.lv_cnt[p];
                        float4* v = w.lv_raw + 3 * ((size_t)p * c.lv_stride + j);
                        const v3 lp = xf_point(S.inv_xf[hit.y & 0xffffu], sd.pos);
                        v[0] = make_float4(lp.x, lp.y, lp.z, __uint_as_float(hit.y));
                        v[1] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(pack_normal(ldin)));
                        v[2] = make_float4(__uint_as_float(path_len), vcm.x, vcm.y, 0);
                        w.lv_cnt[p] = j + 1;
                    }
                    CamLink l;
                    if (alive && !m_singular<LAMBERT>(m) && path_len + 1 >= c.min_len && camera_link(c, sd.pos, l)) {
                        v3 tc = cam - sd.pos;
                        const float dist = len(tc);
                        tc = tc / dist;
                        const v3 ldout = f.to_local(tc);
                        Eval fe = bsdf_eval<LAMBERT, true>(m, ldin, ldout);
                        fe.f = fe.f * shading_normal_correction<true>(f, lgn, ldin, ldout);
                        emit_shadow = camera_record(c, p, thr, fe.f, l.cos_cam, fabsf(ldout.z), dist,
                                                    ray_offset(sd.pos, gn, tc), tc, l.pcx, l.pcy, sh0, sh1, sh2, VC, vcm,
                                                    fe.rev * m_cont_prob(m));
                    }
                    if (alive) {
                        const float cp = fmin_(1.0f, m_cont_prob(m));
                        const float rr = rng.nextf().x;
                        if (rr_on && rr > cp) alive = false;
                        if (alive) {
                            const float cpx = rr_on ? cp : 1.0f;
                            const float4 rb = rng.nextf();
                            Sample sm = bsdf_sample<LAMBERT, true>(m, rb.x, rb.y, rb.z, ldin);
                            sm.fwd *= cpx;
                            sm.rev *= cpx;
                            thr = thr * (sm.f / sm.fwd);
                            if (VC) {
                                vc_after_sample(c, vcm, sm, fabsf(sm.dir.z));
                                vcm_out = vcm;
                            }
                            thr = thr * shading_normal_correction<true>(f, lgn, ldin, sm.dir);
                            const v3 wd = f.to_world(sm.dir);
                            const v3 o = ray_offset(sd.pos, gn, wd);
                            const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
                            cont = !nan && !all_nonpos(thr) && path_len + 1 <= c.max_depth;
                            ray_out = make_float4(o.x, o.y, o.z, __uint_as_float(pack_normal(wd)));
                            thr_out = make_float4(thr.x, thr.y, thr.z, 0);
                        }
                    }
                }
                s4.x = rng.w;
                s4.y = path_len;
                w.st[p] = s4;
                if (cont) {
                    w.ray[p] = ray_out;
                    w.thr[p] = thr_out;
                    if (VC) w.vcm[p] = vcm_out;
                }
            }
        }
        const uint32_t si = wave_append(&w.counters[2], emit_shadow);
        if (emit_shadow) {
            w.shadow[3 * (size_t)si] = sh0;
            w.shadow[3 * (size_t)si + 1] = sh1;
            w.shadow[3 * (size_t)si + 2] = sh2;
        }
        const uint32_t qi = wave_append(&w.counters[qsel ^ 1], cont);
        if (cont) qout[qi] = p;
    }
}

// The camera pass of the bidirectional mode (gUseVC, surfaces only):
// ProcessInteraction<false> (testrenderer.slang:826-935) with the eye sub-path
// MIS quantities completed at each hit, emission weighted by [tech. rep. (43)]
// (AddEmission, :788-806), NEE weighted by ComputeMIS_DI's VC branch
// (:530-566), and one connection to a uniformly chosen vertex of this sample's
// light vertex cache, weighted by cache size / light subpaths
// (SampleLightVertex + ConnectToLightSubpath, :355-437).  Up to two shadow
// records per vertex, both added to the path's own accumulator.
template <bool LAMBERT, bool ALIAS>
__global__ void __launch_bounds__(kShadeBlock) k_shade_vc(DevScene S, RenderConsts c, WaveState w, int qsel, AuxOut aux,
                                                          int first_wave) {
    const uint32_t n = w.counters[qsel];
    const uint32_t* q = w.queue[qsel];
    uint32_t* qout = w.queue[qsel ^ 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t base0 = blockIdx.x * blockDim.x + threadIdx.x - __lane_id();
    const uint32_t lv_count = *w.lv_total;
    for (uint32_t i0 = base0; i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        bool sh_a = false, sh_b = false, cont = false;
        float4 a0 = make_float4(0, 0, 0, 0), a1 = a0, a2 = a0, b0 = a0, b1 = a0, b2 = a0, ray_out = a0, thr_out = a0;
        float2 vcm_out = make_float2(0, 0);
        uint32_t p = 0;
        if (i < n) {
            p = q[i];
            const uint4 hit = w.hit[i];
            const float4 r = w.ray[p];
            const float4 thr4 = w.thr[p];
            uint4 s4 = w.st[p];
            float2 vcm = w.vcm[p];
            const uint32_t k = p / c.npix;
            uint32_t px, py;
            pixel_of(c, p % c.npix, px, py);
            const uint32_t pix = py * c.W + px;
            Rng rng{px, py, c.seed + c.sample_base + k, s4.x};
            const uint32_t path_len = s4.y + 1;
            const v3 dir = unpack_normal(__float_as_uint(r.w));
            v3 thr = mk(thr4.x, thr4.y, thr4.z);
            bool specular = (s4.w >> 31) != 0;
            float4 em = make_float4(0, 0, 0, 0);
            bool em_dirty = false;
            // partial eye sub-path MIS weight [tech. rep. (43)]
            auto add_emission = [&](v3 e, float dpa, float epw) {
                float wgt = 1;
                if (path_len > 1 && !specular) wgt = 1 / (1 + (mis1(dpa) * vcm.y + mis1(epw) * vcm.x));
                if (wgt <= 0 || all_nonpos(e)) return;
                if (!em_dirty) {
                    em = w.em[p];
                    em_dirty = true;
                }
                const v3 a = e * wgt;
                em.x += a.x;
                em.y += a.y;
                em.z += a.z;
            };
            if (hit.y == 0xffffffffu) {
                if (path_len == 1 && first_wave && k == 0)
                    store_aux(c, aux, pix, px, py, 0xffffffffu, pack_normal(dir), false, mk(0, 0, 0), mk(1, 1, 1),
                              q16(__int_as_float(0x7f800000)));
                float dpw;
                const v3 e = thr * evaluate_environment<ALIAS>(S, c.env_prob, dir, dpw);
                if (any_pos(e)) {
                    float dpa = (1.0f / ST_PI) / sq(S.sphere[3]);  // concentricDiscPdfA()/R^2
                    if (c.flags & ST_FLAG_ENV_MIS_BALANCED) dpa = dpw;
                    add_emission(e, dpa, dpw);
                }
                s4.y = path_len;
            } else {
                const uint32_t ii = hit.y & 0xffffu, prim = hit.y >> 16;
                Shading sd;
                hit_shading(S, hit, mk(r.x, r.y, r.z), dir, sd);
                const float pick = prim_pick(S, hit.y);
                float pdfa = 0;  // LightSamplePdfA, lights.hlsli:264-271
                if (S.light_count > 0) {
                    pdfa = pick / (sd.area * (float)S.light_count);
                    if (S.has_env) pdfa *= 1 - c.env_prob;
                }
                const float t16 = q16(__uint_as_float(hit.x)), pdfa16 = q16(pdfa);
                const Material m = load_material<false>(S, sd, c.flags);
                const Frame f = make_frame(sd);
                const v3 ldin = f.to_local(-dir);
                const v3 lgn = f.to_local(unpack_normal(sd.gn));
                const v3 gn = unpack_normal(sd.gn);
                const float cos_in = fabsf(ldin.z);
                vcm.y = qvc(c, vcm.y * mis1(sq(t16)));  // [tech. rep. (31)-(33)], :826-834
                vcm.y = qvc(c, vcm.y / mis1(cos_in));
                vcm.x = qvc(c, vcm.x / mis1(cos_in));
                if (path_len == 1 && first_wave && k == 0)
                    store_aux(c, aux, pix, px, py, hit.y, sd.sn, true, sd.pos, m.base(), t16);
                if (ldin.z > 0) {
                    const v3 e = thr * m.emission();
                    if (any_pos(e)) add_emission(e, pdfa16, pdfa16 * cos_hemisphere_pdfw(ldin.z));
                }
                bool alive = path_len < c.max_depth && any_pos(m.base());
                const bool rr_on = path_len >= c.rr_start;
                const float cp_eff = rr_on ? m_cont_prob(m) : 1.0f;
                if (alive && !m_singular<LAMBERT>(m)) {
                    // ConnectToLight (:751-779) with the VC weight [tech. rep. (44), (45)]
                    const LightSample L = sample_illumination<ALIAS>(S, c.env_prob, rng.nextf(), sd.pos);
                    if (!(all_nonpos(L.radiance) || L.pdf <= 0)) {
                        const float AtoW = L.finite ? sq(L.dist) / L.cos_light : 1.0f;
                        const float dpw = L.pdf * AtoW;
                        const v3 ldout = f.to_local(L.dir);
                        Eval fe = bsdf_eval<LAMBERT>(m, ldin, ldout);
                        fe.f = fe.f * shading_normal_correction(f, lgn, ldin, ldout);
                        fe.fwd *= cp_eff;
                        fe.rev *= cp_eff;
                        const v3 contrib = ((thr * L.radiance) * fe.f) / dpw;
                        const float epw = L.pdf * cos_hemisphere_pdfw(L.cos_light);
                        const float wl = mis1(fe.fwd / dpw);
                        const float wc = mis1(epw * fabsf(ldout.z) / (dpw * L.cos_light)) * (0 + vcm.y + vcm.x * mis1(fe.rev));
                        const float wgt = 1 / (wl + 1 + wc);
                        const v3 o = ray_offset(sd.pos, gn, L.dir);
                        a0 = make_float4(o.x, o.y, o.z, fmax_(L.dist * 0.99f, L.dist - 1e-2f));
                        a1 = make_float4(L.dir.x, L.dir.y, L.dir.z, __uint_as_float(p));
                        a2 = make_float4(contrib.x, contrib.y, contrib.z, wgt);
                        sh_a = true;
                    }
                    // SampleLightVertex: next().x picks a cached vertex uniformly
                    const uint32_t rv = rng.next().x;
                    if (lv_count > 0) {
                        const float4* v = w.lv + 3 * (size_t)(rv % lv_count);
                        const float4 v0 = v[0], v1 = v[1], v2 = v[2];
                        if (path_len + 1 + __float_as_uint(v2.x) <= c.max_depth) {
                            const uint32_t vid = __float_as_uint(v0.w);
                            const v3 lp = mk(v0.x, v0.y, v0.z);
                            Shading sv;
                            tri_shading(S, vid & 0xffffu, vid >> 16, 0.0f, 0.0f, sv, &lp);
                            v3 d = sv.pos - sd.pos;
                            const float dist = len(d);
                            d = d / dist;
                            const Material mv = load_material<false>(S, sv, c.flags);
                            const Frame fv = make_frame(sv);
                            const v3 vin = unpack_normal(__float_as_uint(v1.w));
                            const v3 vout = fv.to_local(-d);
                            const v3 vlgn = fv.to_local(unpack_normal(sv.gn));
                            const float cos_light = fabsf(vout.z);
                            Eval fl = bsdf_eval<LAMBERT, true>(mv, vin, vout);
                            fl.f = fl.f * shading_normal_correction<true>(fv, vlgn, vin, vout);
                            const float cpv = m_cont_prob(mv);
                            fl.fwd *= cpv;
                            fl.rev *= cpv;
                            const v3 vthr = mk(v1.x, v1.y, v1.z);
                            if (!all_nonpos(vthr) && !all_nonpos(fl.f)) {
                                const v3 ldout = f.to_local(d);
                                Eval fe = bsdf_eval<LAMBERT>(m, ldin, ldout);
                                fe.f = fe.f * shading_normal_correction(f, lgn, ldin, ldout);
                                fe.fwd *= cp_eff;
                                fe.rev *= cp_eff;
                                const float wn = (float)lv_count / (float)c.light_paths;
                                const float inv_d2 = 1 / sq(dist);
                                const v3 contrib = (thr * vthr) * (((fe.f * fl.f) * inv_d2) * wn);
                                const float cam_pdfa = fe.fwd * cos_light * inv_d2;
                                const float light_pdfa = fl.fwd * fabsf(ldout.z) * inv_d2;
                                const float wl = mis1(cam_pdfa) * (0 + v2.z + v2.y * mis1(fl.rev));   // [tech. rep. (40)]
                                const float wc = mis1(light_pdfa) * (0 + vcm.y + vcm.x * mis1(fe.rev));  // [tech. rep. (41)]
                                const float wgt = 1 / (wl + 1 + wc);
                                const v3 o = ray_offset(sd.pos, gn, d);
                                b0 = make_float4(o.x, o.y, o.z, fmax_(dist * 0.99f, dist - 1e-2f));
                                b1 = make_float4(d.x, d.y, d.z, __uint_as_float(p));
                                b2 = make_float4(contrib.x, contrib.y, contrib.z, wgt);
                                sh_b = true;
                            }
                        }
                    }
                }
                if (alive) {
                    const float cp = fmin_(1.0f, m_cont_prob(m));
                    const float rr = rng.nextf().x;
                    if (rr_on && rr > cp) alive = false;
                    if (alive) {
                        const float cpx = rr_on ? cp : 1.0f;
                        const float4 rb = rng.nextf();
                        Sample sm = bsdf_sample<LAMBERT>(m, rb.x, rb.y, rb.z, ldin);
                        sm.fwd *= cpx;
                        sm.rev *= cpx;
                        thr = thr * (sm.f / sm.fwd);
                        const float dirpdf = sm.rough == 0 ? 0.0f : sm.fwd;
                        vc_after_sample(c, vcm, sm, fabsf(sm.dir.z));
                        if (sm.rough != 0) specular = false;
                        thr = thr * shading_normal_correction(f, lgn, ldin, sm.dir);
                        const v3 wd = f.to_world(sm.dir);
                        const v3 o = ray_offset(sd.pos, gn, wd);
                        const bool nan = thr.x != thr.x || thr.y != thr.y || thr.z != thr.z;
                        cont = !nan && !all_nonpos(thr) && path_len + 1 <= c.max_depth;
                        ray_out = make_float4(o.x, o.y, o.z, __uint_as_float(pack_normal(wd)));
                        thr_out = make_float4(thr.x, thr.y, thr.z, 0);
                        vcm_out = vcm;
                        s4.w = f2h(dirpdf) | (specular ? 0x80000000u : 0u);
                    }
                }
                s4.y = path_len;
                s4.z = f2h(t16) | (f2h(pdfa16) << 16);
            }
            s4.x = rng.w;
            w.st[p] = s4;
            if (em_dirty) w.em[p] = em;
            if (cont) {
                w.ray[p] = ray_out;
                w.thr[p] = thr_out;
                w.vcm[p] = vcm_out;
            }
        }
        const uint32_t ia = wave_append(&w.counters[2], sh_a);
        if (sh_a) {
            w.shadow[3 * (size_t)ia] = a0;
            w.shadow[3 * (size_t)ia + 1] = a1;
            w.shadow[3 * (size_t)ia + 2] = a2;
        }
        const uint32_t ib = wave_append(&w.counters[2], sh_b);
        if (sh_b) {
            w.shadow[3 * (size_t)ib] = b0;
            w.shadow[3 * (size_t)ib + 1] = b1;
            w.shadow[3 * (size_t)ib + 2] = b2;
        }
        const uint32_t qi = wave_append(&w.counters[qsel ^ 1], cont);
        if (cont) qout[qi] = p;
    }
}

// The light vertex cache of one sample: each subpath's stored vertices moved to
// its exclusive prefix sum, so the cache is ordered by subpath and depth.
__global__ void k_lvc_compact(RenderConsts c, WaveState w, uint32_t n) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const uint32_t cnt = w.lv_cnt[q], off = w.lv_off[q];
        for (uint32_t j = 0; j < cnt; j++)
            for (int e = 0; e < 3; e++) w.lv[3 * (size_t)(off + j) + e] = w.lv_raw[3 * ((size_t)q * c.lv_stride + j) + e];
        if (q == n - 1) *w.lv_total = off + cnt;
    }
}

// Sum each pixel's samples in sample order (ProcessAtomicOutput + running mean).
__global__ void k_resolve(RenderConsts c, WaveState w, float4* acc, float* hdr, uint32_t* path_length, float inv_spp,
                          int last) {
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < c.npix; pix += gridDim.x * blockDim.x) {
        float4 a = acc[pix];
        const uint32_t q = path_of_pixel(c, pix % c.W, pix / c.W);
        for (uint32_t k = 0; k < c.k; k++) {
            const uint32_t p = k * c.npix + q;
            const float4 e = w.em[p];
            const uint4 q = w.nee[p];
            a.x += e.x + (float)q.x / (float)kQuant;
            a.y += e.y + (float)q.y / (float)kQuant;
            a.z += e.z + (float)q.z / (float)kQuant;
            if (path_length) path_length[(size_t)(c.sample_base + k) * c.npix + pix] = w.st[p].y;
        }
        acc[pix] = a;
        if (last) {
            hdr[4 * (size_t)pix + 0] = a.x * inv_spp;
            hdr[4 * (size_t)pix + 1] = a.y * inv_spp;
            hdr[4 * (size_t)pix + 2] = a.z * inv_spp;
            hdr[4 * (size_t)pix + 3] = 1.0f;
        }
    }
}

// Training targets of st_render_cached for the rows whose training path is
// in this wave (rows [row_lo, row_hi)): per channel, (the path's final
// radiance - the radiance it had gathered at the cache vertex) / its
// throughput there (0 where that throughput is 0).  Rows without a record
// keep weight 0.
__global__ void k_rc_targets(RenderConsts c, WaveState w, uint32_t row_lo, uint32_t row_hi) {
    const uint64_t g0 = (uint64_t)c.sample_base * c.npix, g1 = g0 + (uint64_t)c.k * c.npix;
    for (uint32_t row = row_lo + blockIdx.x * blockDim.x + threadIdx.x; row < row_hi; row += gridDim.x * blockDim.x) {
        const uint64_t g = (uint64_t)row * c.train_stride + rc_pcg(row ^ rc_pcg(c.train_frame)) % c.train_stride;
        if (g < g0 || g >= g1) continue;
        const float4 before = w.tt[row];
        if (!(before.w > 0)) continue;
        const uint32_t pix = (uint32_t)(g % c.npix);
        const uint32_t p = (uint32_t)(g / c.npix - c.sample_base) * c.npix + path_of_pixel(c, pix % c.W, pix / c.W);
        const float4 e = w.em[p];
        const uint4 nq = w.nee[p];
        const float4 th = w.tthr[row];
        const float tx = e.x + (float)nq.x / (float)kQuant, ty = e.y + (float)nq.y / (float)kQuant,
                    tz = e.z + (float)nq.z / (float)kQuant;
        w.tt[row] = make_float4(th.x > 0 ? (tx - before.x) / th.x : 0.0f, th.y > 0 ? (ty - before.y) / th.y : 0.0f,
                                th.z > 0 ? (tz - before.z) / th.z : 0.0f, 1.0f);
    }
}

// Sort keys of a sample's hash-grid entries: the cell checksum of each entry
// with weight, 2^32 for empty slots (sorted last); values: the slots.
__global__ void k_hg_keys(const float4* rec, uint32_t n, unsigned long long* keys, uint32_t* slots) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 e2 = rec[4 * (size_t)i + 2];
        keys[i] = e2.y > 0 ? (unsigned long long)__float_as_uint(e2.w) : (1ull << 32);
        slots[i] = i;
    }
}

// st_render_cached's training rows with weight, moved in row order to the
// front of dense arrays (queries, target rgb): flags for the scan, then the
// move; the last thread writes the count.
__global__ void k_rc_row_flags(const float4* tt, uint32_t rows, uint32_t* flag) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) flag[i] = tt[i].w > 0 ? 1u : 0u;
}
__global__ void k_rc_row_compact(const float* tq, const float4* tt, const uint32_t* flag, const uint32_t* off, uint32_t rows,
                                 float* q_out, float* t_out, uint32_t* count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x) {
        if (i == rows - 1) *count = off[i] + flag[i];
        if (!flag[i]) continue;
        const uint32_t d = off[i];
        for (int a = 0; a < 12; a++) q_out[12 * (size_t)d + a] = tq[12 * (size_t)i + a];
        const float4 t = tt[i];
        t_out[3 * (size_t)d] = t.x;
        t_out[3 * (size_t)d + 1] = t.y;
        t_out[3 * (size_t)d + 2] = t.z;
    }
}

__global__ void k_count(WaveState w, int qsel, int media) {
    // rays traced this bounce: extension queue + shadow queue (the media
    // kernels count their traceRay segments themselves)
    unsigned long long* r = reinterpret_cast<unsigned long long*>(w.counters + 4);
    if (!media) {
        r[0] += w.counters[qsel];
        r[1] += w.counters[2];
    }
    r[2] += w.counters[3];  // radiance-cache queries
    w.counters[2] = 0;
    w.counters[3] = 0;
    w.counters[qsel] = 0;
    for (int i = 0; i < 2 * (int)kXcds; i++) w.counters[kFetchClosest + i * kFetchStride] = 0;  // traversal fetch cursors
}

struct RaySrc {
    const float4* rays;
    uint32_t* out;
    int closest;
    __device__ __forceinline__ void load(uint32_t i, v3& o, v3& d, float& tmax) const {
        const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
        tmax = a.w;
    }
    __device__ __forceinline__ void done(uint32_t i, const Hit& h) const {
        if (closest)
            reinterpret_cast<uint4*>(out)[i] =
                make_uint4(__float_as_uint(h.t), h.hit ? h.id : 0xffffffffu, __float_as_uint(h.b1), __float_as_uint(h.b2));
        else
            out[i] = h.hit ? 1u : 0u;
    }
};

template <bool SPH, bool ALPHA = false>
__global__ void __launch_bounds__(kTraceBlock) k_trace_rays(DevScene S, const float4* rays, uint32_t n, int closest,
                                                            uint32_t* out, uint32_t* spill, uint32_t* fetch) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    RaySrc src{rays, out, closest};
    TraverseStats ts{0, 0};
    if (closest) trace_queue<false, false, RaySrc, SPH, ALPHA>(S, n, fetch, src, lds, spill, ts);
    else trace_queue<true, false, RaySrc, SPH, ALPHA>(S, n, fetch, src, lds, spill, ts);
}

// World-space triangle records (v0.xyz|id, v1, v2) of every mesh instance
// under the current transforms, for the device BVH builder.  tri_inst[t] is
// the instance of flattened triangle t, inst_first[i] its first triangle.
__global__ void k_flatten(DevScene S, const uint32_t* tri_inst, const uint32_t* inst_first, uint32_t ntri, float4* out) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntri; t += gridDim.x * blockDim.x) {
        const uint32_t i = tri_inst[t], p = t - inst_first[i];
        const st_instance in = S.inst[i];
        const st_transform T = S.xf[i];
        const st_mesh_vertex_info vi = S.vi[in.data & 0xffffu];
        uint32_t tri[3];
        load_tri(S, vi, p, tri);
        const uint32_t pb = vi.packed_buffer_indices[0] >> 16, ps = (vi.packed_strides >> 8) & 0xffu;
        const uint8_t* P = S.vb[pb] + vi.packed_offsets[1];
        const v3 a = xf_point(T, ld_v3(P + (size_t)ps * tri[0])), b = xf_point(T, ld_v3(P + (size_t)ps * tri[1])),
                 c = xf_point(T, ld_v3(P + (size_t)ps * tri[2]));
        out[3 * (size_t)t] = make_float4(a.x, a.y, a.z, __uint_as_float(i | (p << 16)));
        const uint32_t alpha = S.mat[(in.type_material_address >> 4) / 4 + 12] < ST_IMAGE_LIMIT ? 1u : 0u;
        out[3 * (size_t)t + 1] = make_float4(b.x, b.y, b.z, __uint_as_float(alpha));
        out[3 * (size_t)t + 2] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

// Tonemapper::render's two passes (tonemap.slang:124-192).  maxReduce: the
// maximum of rgb and luminance in 1/16384 fixed point (one atomic per wave).
// Emitter table (DevScene::light_rec, light_tri) from the scene buffers: one
// block per light; tri_shading's loaders copy each triangle's object-space
// vertices and normals bit for bit, and the light's transform and emission
// words are copied as they are.
__global__ void k_light_table(DevScene S, const uint32_t* tri_base, float4* rec, float4* tri) {
    const uint32_t li = blockIdx.x;
    if (li >= S.light_count) return;
    const uint32_t ii = S.light_inst[li];
    const st_instance in = S.inst[ii];
    const bool sphere = (in.type_material_address & 0xfu) == 1;
    const uint32_t pc = sphere ? 0u : in.data >> 16, base = tri_base[li];
    if (threadIdx.x == 0) {
        const st_transform& T = S.xf[ii];
        for (int r = 0; r < 3; r++) rec[4 * (size_t)li + r] = make_float4(T.m[r][0], T.m[r][1], T.m[r][2], T.m[r][3]);
        const uint32_t addr = in.type_material_address >> 4;
        rec[4 * (size_t)li + 3] = make_float4(__uint_as_float(pc), sphere ? __uint_as_float(in.data) : __uint_as_float(base),
                                              __uint_as_float(S.mat[addr / 4 + 1]), __uint_as_float(S.mat[addr / 4 + 8]));
    }
    if (sphere) return;
    const st_mesh_vertex_info vi = S.vi[in.data & 0xffffu];
    const uint32_t pb = vi.packed_buffer_indices[0] >> 16, ps = (vi.packed_strides >> 8) & 0xffu;
    const uint32_t nb = vi.packed_buffer_indices[1] & 0xffffu, ns = (vi.packed_strides >> 16) & 0xffu;
    const bool has_n = nb < ST_IMAGE_LIMIT;
    for (uint32_t prim = threadIdx.x; prim < pc; prim += blockDim.x) {
        uint32_t t[3];
        load_tri(S, vi, prim, t);
        const uint8_t* P = S.vb[pb] + vi.packed_offsets[1];
        const v3 v0 = ld_v3(P + (size_t)ps * t[0]), v1 = ld_v3(P + (size_t)ps * t[1]), v2 = ld_v3(P + (size_t)ps * t[2]);
        v3 n0 = mk(0, 0, 0), n1 = n0, n2 = n0;
        if (has_n) {
            const uint8_t* N = S.vb[nb] + vi.packed_offsets[2];
            n0 = ld_v3(N + (size_t)ns * t[0]);
            n1 = ld_v3(N + (size_t)ns * t[1]);
            n2 = ld_v3(N + (size_t)ns * t[2]);
        }
        float4* o = tri + 5 * ((size_t)base + prim);
        o[0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float(has_n ? 1u : 0u));
        o[1] = make_float4(v1.x, v1.y, v1.z, n0.x);
        o[2] = make_float4(v2.x, v2.y, v2.z, n0.y);
        o[3] = make_float4(n0.z, n1.x, n1.y, n1.z);
        o[4] = make_float4(n2.x, n2.y, n2.z, 0.0f);
    }
}

__global__ void k_tonemap_max(const float4* in, const float4* albedo, uint32_t n, uint32_t* mx) {
    uint32_t m[4] = {0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 c = in[i];
        v3 v = mk(c.x, c.y, c.z);
        if (albedo) {
            const float4 a = albedo[i];
            v = v * mk(a.x, a.y, a.z);
        }
        const float l = lum(v);
        if (v.x != v.x || v.y != v.y || v.z != v.z || l != l || l <= 0) continue;
        const float q[4] = {v.x * kTonemapQuant, v.y * kTonemapQuant, v.z * kTonemapQuant, l * kTonemapQuant};
        for (int k = 0; k < 4; k++) {
            const float cl = fmin_(fmax_(q[k], 0.0f), 4294967296.0f);  // clamp(v * Q, 0, float(0xFFFFFFFF))
            const uint32_t u = cl >= 4294967296.0f ? 0xffffffffu : (uint32_t)cl;
            m[k] = m[k] > u ? m[k] : u;
        }
    }
    for (int k = 0; k < 4; k++) {
        uint32_t v = m[k];
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_xor(v, off);
            v = v > o ? v : o;
        }
        if (__lane_id() == 0 && v) atomicMax(&mx[k], v);
    }
}

// main: exposure, operator, optional sRGB encode
__global__ void k_tonemap(const float4* in, const float4* albedo, uint32_t n, const uint32_t* mx, float exposure, int mode,
                          int gamma, float4* out) {
    const float4 mv = make_float4((float)mx[0] / kTonemapQuant, (float)mx[1] / kTonemapQuant, (float)mx[2] / kTonemapQuant,
                                  (float)mx[3] / kTonemapQuant);
    const float scale = powf(2.0f, exposure);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 c = in[i];
        v3 r = mk(c.x, c.y, c.z);
        if (albedo) {
            const float4 a = albedo[i];
            r = r * (mk(1e-2f, 1e-2f, 1e-2f) + mk(a.x, a.y, a.z));
        }
        r = r * scale;
        r = tonemap(mode, r, mv);
        if (gamma) r = mk(tm_srgb1(r.x), tm_srgb1(r.y), tm_srgb1(r.z));
        out[i] = make_float4(r.x, r.y, r.z, 1.0f);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// -------------------------------------------------------------------- host side
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void reserve(size_t bytes) {
        if (bytes <= n) return;
        if (p) HIP_CHECK(hipFree(p));
        p = nullptr;
        HIP_CHECK(hipMalloc(&p, bytes));
        n = bytes;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

}  // namespace

// Device state of one wave of paths.  Up to kWaveSlots waves are in flight at
// once, each in its own slot and on its own stream, so that one wave's last
// bounces (few rays, bound by the slowest ray) overlap the next wave's first
// bounces.  Four slots match the four hardware queues a process gets; with
// waves of about 2M paths (one 1080p sample) this measured fastest at 1080p
// (more slots share queues, larger waves were slower).
constexpr int kWaveSlots = 4;
constexpr uint32_t kWavePaths = 2u << 20;

// Fixed device scratch of st_trace_rays_noalloc: a 256-byte header holding the
// traversal's ray-fetch counter, then the HBM part of the per-lane traversal
// stacks.  One copy per GPU, held until the process ends; calls on one GPU
// share it.  Host code never passes `scratch` itself to a kernel (on the host
// it names a stand-in, not device memory): it uses the device address that
// hipGetSymbolAddress returns, kept per device in g_scratch.
constexpr size_t kScratchHeader = 4 * kXcds * kFetchStride;  // the ray-fetch cursors, one 128-byte line each
constexpr size_t kScratchSpill = 16u << 20;
constexpr size_t kScratchBytes = kScratchHeader + kScratchSpill;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};
struct WaveBufs {
    DevBuf ray, thr, stt, em, nee, hit, q0, q1, shadow, counters, spill, pri, psd, rinit, rtemp, med;
    DevBuf vcm, lv_raw, lv, lv_cnt, lv_off, lv_total, scan_tmp;  // bidirectional mode
    DevBuf ltot;  // st_adjoint: each path's first-pass radiance
    DevBuf cq;    // st_render_cached: queries of the paths that end in the cache
    hipStream_t stream = nullptr;        // slot 0: the context's stream
    hipEvent_t reuse_done = nullptr;     // this slot's last wave finished its screen-space reuse
    hipEvent_t resolve_done = nullptr;   // this slot's last wave finished its resolve
    hipEvent_t done = nullptr;           // this slot's part of an ST_FLAG_ASYNC call is enqueued up to here
};

struct st_context {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    WaveBufs slot[kWaveSlots];
    hipEvent_t start_done = nullptr;     // the frame's clears on the context stream
    hipEvent_t entry = nullptr;          // the caller's null-stream work at a call's entry
    uint8_t* scratch = nullptr;          // this device's fixed scratch (device address of `scratch`)
    DevBuf acc, stats, raysbuf, outbuf, rprev, tmax;
    // hash-grid reservoir reuse: two samples' entries and sorted keys / slots
    DevBuf hg_rec[2], hg_keys[2], hg_slots[2], hg_keys_tmp, hg_slots_tmp, hg_sort_tmp;
    // the grid the last sample of the previous hash-grid st_render built: its
    // buffer (hg_last), scene and entry count; valid until a st_render without
    // the flag, a failed call or ST_FLAG_HISTORY_RESET
    bool hg_valid = false;
    int hg_last = 0;
    uint64_t hg_scene = 0, hg_ns = 0;
    double timings[12] = {0};
    double cache_queries = 0;            // paths that ended in the radiance cache in the last call
    double cache_rows = 0;               // its training rows with weight 1 (when its loss was asked for)
    // HIP-event timed kernel classes: 0 closest-hit trace, 1 shading, 2 shadow
    // trace, 3 primary-vertex RIS (k_ris), 4 screen-space reuse passes,
    // 5 radiance-cache inference + targets, 6 radiance-cache training
    static constexpr int kKinds = 7;
    static constexpr int kRenderKinds = 5;  // the classes st_last_kernel_times reports
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[kKinds];
    size_t ev_used[kKinds] = {0, 0, 0, 0, 0, 0, 0};
    double kernel_ms[kKinds] = {0, 0, 0, 0, 0, 0, 0};
};

struct st_scene {
    st_context* ctx = nullptr;
    DevScene d{};
    std::vector<DevBuf*> bufs;
    // host copies for the camera's medium (TestRenderer.cpp:431-446)
    std::vector<st_volume_info> volume_info;
    std::vector<st_transform> inv_xf;
    uint64_t stats[4] = {0, 0, 0, 0};
    uint64_t generation = 0;  // bumped by st_scene_update: device pointers and constants may have changed
    uint64_t serial = 0;      // unique per created scene (a kept hash grid names its scene by it)
    // for scene updates: per-instance object-space bounds (the scene sphere),
    // the flattened triangles' owners, and the BVH buffers (replaced on update)
    struct InstBox {
        int kind;  // 0 none, 1 mesh, 2 volume
        float lo[3], hi[3];
    };
    std::vector<InstBox> boxes;
    std::vector<uint32_t> inst_first;
    uint64_t ntri = 0;
    const uint32_t* tri_inst = nullptr;
    const uint32_t* inst_first_dev = nullptr;
    void* bvh_nodes = nullptr;  // hipMalloc'ed by the device builder, else null
    void* bvh_tris = nullptr;
    std::vector<uint64_t> vol_bytes;  // each grid's size
    DevBuf tri_alb;                   // st_scene_set_triangle_albedo's copy
    // emitter table, rebuilt when the instances move
    const uint32_t* light_tri_base = nullptr;
    float4* light_rec = nullptr;
    float4* light_tri = nullptr;
    // sphere instances (index, radius) and their tree, rebuilt when the instances move
    std::vector<uint32_t> sph_inst;
    std::vector<float> sph_rad;
    DevBuf sph_nodes, sph_tris;
    // The tree over the sphere instances' world boxes (the transformed corners
    // of [-r, r]^3, widened by 2^-20 of their magnitude): the triangle builder
    // over records whose first two vertices are the box corners.
    void build_sphere_tree(const st_transform* xf) {
        const uint32_t n = (uint32_t)sph_inst.size();
        d.sph_node_count = 0;
        if (n == 0) return;
        std::vector<float> verts(9 * (size_t)n);
        std::vector<uint32_t> ids(n);
        for (uint32_t k = 0; k < n; k++) {
            const st_transform& T = xf[sph_inst[k]];
            const float r = sph_rad[k];
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c = 0; c < 8; c++) {
                const float v[3] = {(c & 1) ? r : -r, (c & 2) ? r : -r, (c & 4) ? r : -r};
                for (int a = 0; a < 3; a++) {
                    const float w = T.m[a][0] * v[0] + T.m[a][1] * v[1] + T.m[a][2] * v[2] + T.m[a][3];
                    lo[a] = std::min(lo[a], w);
                    hi[a] = std::max(hi[a], w);
                }
            }
            for (int a = 0; a < 3; a++) {
                const float m = std::max(std::fabs(lo[a]), std::fabs(hi[a])) * 0x1p-20f + 1e-30f;
                verts[9 * (size_t)k + a] = lo[a] - m;
                verts[9 * (size_t)k + 3 + a] = hi[a] + m;
                verts[9 * (size_t)k + 6 + a] = lo[a] - m;
            }
            ids[k] = sph_inst[k];
        }
        BvhOut b = build_bvh(verts.data(), ids.data(), n);
        if (3 * b.depth4 + 1 > kMaxStack) throw std::runtime_error("sphere tree deeper than the traversal stack");
        sph_nodes.reserve(b.nodes4.size() * sizeof(Node4Out));
        sph_tris.reserve(b.tris4.size() * sizeof(TriOut));
        HIP_CHECK(hipMemcpy(sph_nodes.p, b.nodes4.data(), b.nodes4.size() * sizeof(Node4Out), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(sph_tris.p, b.tris4.data(), b.tris4.size() * sizeof(TriOut), hipMemcpyHostToDevice));
        d.sph_nodes = sph_nodes.as<float4>();
        d.sph_tris = sph_tris.as<float4>();
        d.sph_node_count = (uint32_t)b.nodes4.size();
    }
    ~st_scene() {
        for (auto* b : bufs) delete b;
        if (bvh_nodes) (void)hipFree(bvh_nodes);
        if (bvh_tris) (void)hipFree(bvh_tris);
    }
    void build_light_table(hipStream_t stream) {
        if (d.light_count == 0) return;
        hipLaunchKernelGGL(k_light_table, dim3(d.light_count), dim3(64), 0, stream, d, light_tri_base, light_rec, light_tri);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(stream));
    }
    template <class T>
    T* alloc(size_t bytes) {
        DevBuf* b = new DevBuf();
        bufs.push_back(b);
        b->reserve(std::max<size_t>(bytes, 16));
        return b->as<T>();
    }
    template <class T>
    T* upload(const void* src, size_t bytes) {
        DevBuf* b = new DevBuf();
        bufs.push_back(b);
        b->reserve(std::max<size_t>(bytes, 16));
        if (bytes) HIP_CHECK(hipMemcpy(b->p, src, bytes, hipMemcpyHostToDevice));
        return b->as<T>();
    }
};

namespace {

hipEvent_t get_event(st_context* ctx, int kind, bool start) {
    auto& v = ctx->ev[kind];
    size_t i = ctx->ev_used[kind];
    if (i >= v.size()) {
        hipEvent_t a, b;
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        v.push_back({a, b});
    }
    return start ? v[i].first : v[i].second;
}

// HIP events around one launch on its stream; a null context times nothing
// (work being captured into a graph).
struct Timed {
    st_context* ctx;
    int kind;
    hipStream_t stream;
    Timed(st_context* c, int k, hipStream_t s) : ctx(c), kind(k), stream(s) {
        if (ctx) HIP_CHECK(hipEventRecord(get_event(c, k, true), s));
    }
    ~Timed() {
        if (!ctx) return;
        (void)hipEventRecord(get_event(ctx, kind, false), stream);
        ctx->ev_used[kind]++;
    }
};

// The context's streams do not synchronise with the null stream, so every
// call that reads caller memory first makes the context stream wait for the
// work the caller has queued on the device's null stream (where torch's
// default stream puts its kernels and copies).
void order_after_caller(st_context* ctx) {
    HIP_CHECK(hipEventRecord(ctx->entry, nullptr));
    HIP_CHECK(hipStreamWaitEvent(ctx->stream, ctx->entry, 0));
}

void validate_desc(const st_scene_desc* d) {
    if (!d) throw std::invalid_argument("null scene description");
    if (d->instance_count && (!d->instances || !d->transforms)) throw std::invalid_argument("instances/transforms missing");
    if (d->instance_count > 0xFFFF) throw std::invalid_argument("more than 65535 instances (16-bit instance ids)");
    if (d->volume_count > (uint32_t)kMaxVolumes) throw std::invalid_argument("more than 8 volumes (gVolumeCount)");
    if (d->volume_count && !d->volumes) throw std::invalid_argument("volumes missing");
    if (d->volume_info_count && !d->volume_info) throw std::invalid_argument("volume info missing");
    for (uint32_t i = 0; i < d->volume_info_count; i++) {
        const uint32_t ii = d->volume_info[i].instance_index;
        if (ii >= d->instance_count || (d->instances[ii].type_material_address & 0xfu) != 2)
            throw std::invalid_argument("volume info must name a volume instance");
    }
    uint32_t nvol = 0, nsph = 0;
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const st_instance& in = d->instances[i];
        const uint32_t type = in.type_material_address & 0xfu;
        if (type == 2) {  // VolumeInstanceData (compat/scene.h:71-80) with a 28-byte medium record
            if (++nvol > (uint32_t)kMaxVolumeInstances) throw std::domain_error("more than 8 volume instances");
            const uint32_t addr = in.type_material_address >> 4;
            if (addr % 4 || (uint64_t)addr / 4 + 7 > d->material_word_count) throw std::invalid_argument("medium address out of range");
            const uint32_t dv = d->material_data[addr / 4 + 4];
            if (in.data != 0xFFFFFFFFu && in.data >= d->volume_count) throw std::invalid_argument("volume index out of range");
            if (dv != 0xFFFFFFFFu && dv >= d->volume_count) throw std::invalid_argument("density volume index out of range");
            if (d->material_data[addr / 4 + 5] != 0xFFFFFFFFu) throw std::domain_error("albedo volumes are not supported");
            continue;
        }
        if (type == 1) {  // SphereInstanceData (compat/scene.h:60-69): the radius in data
            float r;
            std::memcpy(&r, &in.data, 4);
            if (!(r > 0) || !std::isfinite(r)) throw std::invalid_argument("sphere radius must be positive and finite");
        } else if (type != 0) {
            throw std::domain_error("only mesh, sphere and volume instances are supported");
        }
        const uint32_t addr = in.type_material_address >> 4;
        if ((uint64_t)addr / 4 + 16 > d->material_word_count) throw std::invalid_argument("material address out of range");
        // images: base colour and packed words 2, 3; emission images, normal
        // maps and alpha masks are refused
        const uint32_t* mw = d->material_data + addr / 4;
        if (mw[5] < ST_IMAGE_LIMIT || mw[9] < ST_IMAGE_LIMIT)
            throw std::domain_error("emission images and normal maps are not supported");
        if (mw[12] < ST_IMAGE_LIMIT && type != 0) throw std::domain_error("alpha masks apply to mesh materials only");
        for (int k : {4, 6, 7, 12})
            if (mw[k] < ST_IMAGE_LIMIT && (mw[k] >= d->image_count || !d->images || !d->images[mw[k]].rgba ||
                                           d->images[mw[k]].width == 0 || d->images[mw[k]].height == 0))
                throw std::invalid_argument("material image index out of range or empty");
        if (d->material_data[addr / 4 + 9] < ST_IMAGE_LIMIT) throw std::domain_error("normal maps are not supported");
        if (type == 0 && (in.data & 0xffffu) >= d->mesh_vertex_info_count) throw std::invalid_argument("vertex info index out of range");
        if (type == 1) nsph++;
    }
    // a medium's scattering vertices take the primitive id 0xffff that marks sphere hits
    if (nsph && nvol) throw std::domain_error("sphere instances together with media are not supported");
}

// The material images' mip chains (2x2 box, as Image::generateMipMaps blits,
// Image.cpp:282-308, in the same float operations as the environment's), all
// levels of all used images in one buffer.
void build_material_images(st_scene* s, const st_scene_desc* d) {
    DevScene& D = s->d;
    D.has_tex = 0;
    D.has_alpha = 0;
    D.tex_count = 0;
    std::vector<uint2> img(d->image_count + 1, make_uint2(0, 0));
    std::vector<uint4> lvt;
    std::vector<float> data;
    bool any = false;
    for (uint32_t i = 0; i < d->instance_count; i++) {
        const uint32_t type = d->instances[i].type_material_address & 0xfu;
        if (type > 1) continue;
        const uint32_t* mw = d->material_data + (d->instances[i].type_material_address >> 4) / 4;
        if (type == 0 && mw[12] < ST_IMAGE_LIMIT) D.has_alpha = 1;
        for (int k : {4, 6, 7, 12}) {
            const uint32_t id = mw[k];
            if (id >= ST_IMAGE_LIMIT || img[id].y) continue;
            const st_image_desc& im = d->images[id];
            std::vector<float> cur(im.rgba, im.rgba + 4ull * im.width * im.height);
            int w0 = (int)im.width, h0 = (int)im.height;
            img[id].x = (uint32_t)lvt.size();
            for (int l = 0; l < 16; l++) {
                lvt.push_back(make_uint4((uint32_t)(data.size() / 4), (uint32_t)w0, (uint32_t)h0, 0u));
                data.insert(data.end(), cur.begin(), cur.end());
                img[id].y++;
                if (!(w0 > 1 || h0 > 1) || l == 15) break;
                const int w1 = w0 > 1 ? w0 / 2 : 1, h1 = h0 > 1 ? h0 / 2 : 1;
                std::vector<float> nx(4ull * w1 * h1);
                for (int y = 0; y < h1; y++)
                    for (int x = 0; x < w1; x++) {
                        const int x0 = w0 > 1 ? 2 * x : 0, x1 = w0 > 1 ? 2 * x + 1 : 0;
                        const int y0 = h0 > 1 ? 2 * y : 0, y1 = h0 > 1 ? 2 * y + 1 : 0;
                        for (int ch = 0; ch < 4; ch++) {
                            auto T = [&](int xx, int yy) { return cur[4 * ((size_t)yy * w0 + xx) + ch]; };
                            nx[4 * ((size_t)y * w1 + x) + ch] = ((T(x0, y0) + T(x1, y0)) + (T(x0, y1) + T(x1, y1))) * 0.25f;
                        }
                    }
                cur.swap(nx);
                w0 = w1;
                h0 = h1;
            }
            any = true;
        }
    }
    if (!any) return;
    if (data.size() / 4 >= (1ull << 32)) throw std::domain_error("material images too large");
    D.tex_data = s->upload<float4>(data.data(), data.size() * sizeof(float));
    D.tex_lv = s->upload<uint4>(lvt.data(), lvt.size() * sizeof(uint4));
    D.tex_img = s->upload<uint2>(img.data(), img.size() * sizeof(uint2));
    D.tex_count = d->image_count;
    D.has_tex = 1;
}

// Scene AABB (Scene.cpp:622-631 mesh instances, :672-711 volumes, where the
// reference lets the running min/max feed the later corners) and the bounding
// sphere of TestRenderer.cpp:458-461.
template <class Box>
void scene_sphere(const std::vector<Box>& boxes, const st_transform* xf, float sphere[4]) {
    auto xfp = [](const st_transform& T, const float v[3], float o[3]) {
        for (int r = 0; r < 3; r++) o[r] = T.m[r][0] * v[0] + T.m[r][1] * v[1] + T.m[r][2] * v[2] + T.m[r][3];
    };
    float amin[3] = {INFINITY, INFINITY, INFINITY}, amax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < boxes.size(); i++) {
        const Box& b = boxes[i];
        if (b.kind == 2) {
            float mn[3] = {b.lo[0], b.lo[1], b.lo[2]}, mx[3] = {b.hi[0], b.hi[1], b.hi[2]};
            for (int cidx = 0; cidx < 8; cidx++) {
                const float corner[3] = {(cidx % 2) == 0 ? mn[0] : mx[0], ((cidx % 4) / 2) == 0 ? mn[1] : mx[1],
                                         (cidx / 4) == 0 ? mn[2] : mx[2]};
                float wc[3];
                xfp(xf[i], corner, wc);
                for (int k = 0; k < 3; k++) {
                    mn[k] = std::min(mn[k], wc[k]);
                    mx[k] = std::max(mx[k], wc[k]);
                }
            }
            for (int k = 0; k < 3; k++) {
                amin[k] = std::min(amin[k], mn[k]);
                amax[k] = std::max(amax[k], mx[k]);
            }
        } else if (b.kind == 1) {
            for (int cidx = 0; cidx < 8; cidx++) {
                const float corner[3] = {(cidx % 2) == 0 ? b.lo[0] : b.hi[0], ((cidx % 4) / 2) == 0 ? b.lo[1] : b.hi[1],
                                         (cidx / 4) == 0 ? b.lo[2] : b.hi[2]};
                float w[3];
                xfp(xf[i], corner, w);
                for (int k = 0; k < 3; k++) {
                    amin[k] = std::min(amin[k], w[k]);
                    amax[k] = std::max(amax[k], w[k]);
                }
            }
        }
    }
    float cen[3], dd = 0;
    for (int k = 0; k < 3; k++) cen[k] = (amax[k] + amin[k]) / 2;
    for (int k = 0; k < 3; k++) dd += (amax[k] - cen[k]) * (amax[k] - cen[k]);
    sphere[0] = cen[0];
    sphere[1] = cen[1];
    sphere[2] = cen[2];
    sphere[3] = std::sqrt(dd);
}

// Structural check of a 4-wide tree before the traversal kernels may use it:
// every inner child index in range and referenced once, every triangle in
// exactly one leaf, depth within the traversal stack.  Returns "" when valid.
std::string check_bvh4(const std::vector<Node4Out>& nodes, uint32_t tri_count) {
    const size_t nn = nodes.size();
    if (nn == 0) return "empty tree";
    std::vector<uint8_t> seen_node(nn, 0), seen_tri(tri_count, 0);
    std::vector<uint32_t> level{0}, next;
    seen_node[0] = 1;
    int depth = 0;
    size_t reached = 0;
    while (!level.empty()) {
        if (++depth > kMaxDepth) return "tree deeper than the traversal stack";
        next.clear();
        for (uint32_t ni : level) {
            reached++;
            const Node4Out& n = nodes[ni];
            const uint32_t types = (n.x0 >> 24) | ((n.inner >> 28) << 8);
            const uint32_t base_inner = n.inner & 0x0fffffffu;
            uint32_t r = 0, off = 0;
            for (int c = 0; c < 4; c++) {
                const uint32_t t = (types >> (3 * c)) & 7u;
                if (t == 7) {
                    const uint64_t ci = (uint64_t)base_inner + r++;
                    if (ci >= nn || seen_node[ci]) return "bad inner child index";
                    seen_node[ci] = 1;
                    next.push_back((uint32_t)ci);
                } else if (t >= 1 && t <= 4) {
                    for (uint32_t k = 0; k < t; k++) {
                        const uint64_t ti = (uint64_t)n.tri + off + k;
                        if (ti >= tri_count || seen_tri[ti]) return "bad leaf triangle range";
                        seen_tri[ti] = 1;
                    }
                    off += t;
                } else if (t != 0) {
                    return "bad child type";
                }
            }
        }
        level.swap(next);
    }
    if (reached != nn) return "unreachable nodes";
    for (uint32_t t = 0; t < tri_count; t++)
        if (!seen_tri[t]) return "triangle missing from the leaves";
    return "";
}

// NanoVDB 32.3 float grid: header checks and a walk of every child offset
// (root tiles -> upper -> lower -> leaf), so the kernels' reads stay inside
// the buffer.  Returns the byte offset of the root node.
uint64_t validate_nanovdb(const uint8_t* b, uint64_t n) {
    auto need = [&](uint64_t off, uint64_t len) {
        if (off > n || len > n - off) throw std::invalid_argument("NanoVDB grid: offset out of range");
    };
    auto r32 = [&](uint64_t o) { need(o, 4); uint32_t v; std::memcpy(&v, b + o, 4); return v; };
    auto r64 = [&](uint64_t o) { need(o, 8); uint64_t v; std::memcpy(&v, b + o, 8); return v; };
    need(0, 672 + 64 + 64);
    if (r64(0) != 0x304244566F6E614Eull) throw std::invalid_argument("NanoVDB grid: bad magic");
    if ((r32(16) >> 21) != 32) throw std::domain_error("NanoVDB grid: major version must be 32");
    if (r32(636) != 1) throw std::domain_error("NanoVDB grid: only float grids are supported");
    const uint64_t root = 672 + r64(672 + 24);
    if (root % 8) throw std::invalid_argument("NanoVDB grid: misaligned root");
    need(root, 64);
    const uint32_t ntiles = r32(root + 24);
    need(root + 64, 32ull * ntiles);
    for (uint32_t t = 0; t < ntiles; t++) {
        const int64_t child = (int64_t)r64(root + 64 + 32ull * t + 8);
        if (child == 0) continue;
        if (child < 0 || child % 8) throw std::invalid_argument("NanoVDB grid: bad upper offset");
        const uint64_t up = root + (uint64_t)child;
        need(up, 270400);
        for (uint32_t n2 = 0; n2 < 32768; n2++) {
            if (!((r32(up + 4128 + 4 * (n2 >> 5)) >> (n2 & 31)) & 1)) continue;
            const int64_t lo_off = (int64_t)r64(up + 8256 + 8ull * n2);
            if (lo_off <= 0 || lo_off % 8) throw std::invalid_argument("NanoVDB grid: bad lower offset");
            const uint64_t lo = up + (uint64_t)lo_off;
            need(lo, 33856);
            for (uint32_t n1 = 0; n1 < 4096; n1++) {
                if (!((r32(lo + 544 + 4 * (n1 >> 5)) >> (n1 & 31)) & 1)) continue;
                const int64_t lf = (int64_t)r64(lo + 1088 + 8ull * n1);
                if (lf <= 0 || lf % 8) throw std::invalid_argument("NanoVDB grid: bad leaf offset");
                need(lo + (uint64_t)lf, 2144);
            }
        }
    }
    return root;
}

// TransformData::inverse of a 3x4 affine transform, in double.
st_transform inverse_transform(const st_transform& T) {
    st_transform r{};
    double a[3][3], inv[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = T.m[i][j];
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0) return r;
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int i = 0; i < 3; i++) {
        double tr = 0;
        for (int j = 0; j < 3; j++) {
            r.m[i][j] = (float)inv[i][j];
            tr -= inv[i][j] * T.m[j][3];
        }
        r.m[i][3] = (float)tr;
    }
    return r;
}

}  // namespace

extern "C" {

const char* st_last_error(void) { return g_error.c_str(); }
int st_version(void) { return 1; }

int st_context_create(int device, st_context** out) {
    if (!out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null output");
    try {
        if (device < 0 || device >= kMaxDevices) return set_error(ST_ERROR_INVALID_ARGUMENT, "device index out of range");
        auto ctx = std::make_unique<st_context>();
        ctx->device = device;
        HIP_CHECK(hipSetDevice(device));
        {
            // the fixed scratch's device address, once per device, with the device current
            std::lock_guard<std::mutex> lock(g_scratch_mu);
            if (!g_scratch[device]) {
                void* ptr = nullptr;
                HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
                if (!ptr) throw HipError("hipGetSymbolAddress(scratch) returned no address");
                g_scratch[device] = ptr;
            }
            ctx->scratch = static_cast<uint8_t*>(g_scratch[device]);
        }
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->cus = prop.multiProcessorCount;
        HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        // more streams of the same (default) priority for the other wave slots
        ctx->slot[0].stream = ctx->stream;
        for (int sl = 1; sl < kWaveSlots; sl++) HIP_CHECK(hipStreamCreateWithFlags(&ctx->slot[sl].stream, hipStreamNonBlocking));
        for (WaveBufs& b : ctx->slot) {
            HIP_CHECK(hipEventCreateWithFlags(&b.reuse_done, hipEventDisableTiming));
            HIP_CHECK(hipEventCreateWithFlags(&b.resolve_done, hipEventDisableTiming));
            HIP_CHECK(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
        }
        HIP_CHECK(hipEventCreateWithFlags(&ctx->start_done, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ctx->entry, hipEventDisableTiming));
        *out = ctx.release();
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_context_destroy(st_context* ctx) {
    if (!ctx) return ST_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int sl = 1; sl < kWaveSlots; sl++)
        if (ctx->slot[sl].stream) (void)hipStreamSynchronize(ctx->slot[sl].stream);
    for (auto& v : ctx->ev)
        for (auto& e : v) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for (WaveBufs& b : ctx->slot) {
        if (b.reuse_done) (void)hipEventDestroy(b.reuse_done);
        if (b.resolve_done) (void)hipEventDestroy(b.resolve_done);
        if (b.done) (void)hipEventDestroy(b.done);
    }
    if (ctx->start_done) (void)hipEventDestroy(ctx->start_done);
    if (ctx->entry) (void)hipEventDestroy(ctx->entry);
    for (int sl = 1; sl < kWaveSlots; sl++)
        if (ctx->slot[sl].stream) (void)hipStreamDestroy(ctx->slot[sl].stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ST_OK;
}

int st_sync(st_context* ctx) {
    if (!ctx) return set_error(ST_ERROR_INVALID_ARGUMENT, "null context");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return set_error(ST_ERROR_HIP, hipGetErrorString(e));
    return ST_OK;
}

int st_scene_create(st_context* ctx, const st_scene_desc* d, st_scene** out) {
    if (!ctx || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    try {
        validate_desc(d);
        HIP_CHECK(hipSetDevice(ctx->device));
        auto s = std::make_unique<st_scene>();
        static std::atomic<uint64_t> next_serial{1};
        s->serial = next_serial++;
        s->ctx = ctx;
        DevScene& D = s->d;
        // reference-format buffers
        D.inst = s->upload<st_instance>(d->instances, sizeof(st_instance) * d->instance_count);
        D.xf = s->upload<st_transform>(d->transforms, sizeof(st_transform) * d->instance_count);
        D.mat = s->upload<uint32_t>(d->material_data, 4ull * d->material_word_count + 64);
        D.vi = s->upload<st_mesh_vertex_info>(d->mesh_vertex_info, sizeof(st_mesh_vertex_info) * d->mesh_vertex_info_count);
        D.inst_count = d->instance_count;
        D.mat_words = d->material_word_count;
        D.vi_count = d->mesh_vertex_info_count;
        std::vector<const uint8_t*> vbp(d->vertex_buffer_count + 1, nullptr);
        std::vector<uint64_t> vbs(d->vertex_buffer_count + 1, 0);
        for (uint32_t i = 0; i < d->vertex_buffer_count; i++) {
            vbp[i] = s->upload<uint8_t>(d->vertex_buffers[i].data, d->vertex_buffers[i].size_bytes);
            vbs[i] = d->vertex_buffers[i].size_bytes;
        }
        D.vb = s->upload<const uint8_t*>(vbp.data(), sizeof(void*) * vbp.size());
        D.vb_size = s->upload<uint64_t>(vbs.data(), 8 * vbs.size());
        D.vb_count = d->vertex_buffer_count;
        // flatten mesh instances into world-space triangles; lights; scene AABB
        auto u32at = [&](uint32_t buf, uint64_t off) {
            if (buf >= d->vertex_buffer_count || off + 4 > d->vertex_buffers[buf].size_bytes)
                throw std::invalid_argument("vertex/index buffer access out of range");
            uint32_t v;
            std::memcpy(&v, (const uint8_t*)d->vertex_buffers[buf].data + off, 4);
            return v;
        };
        auto u16at = [&](uint32_t buf, uint64_t off) {
            if (buf >= d->vertex_buffer_count || off + 2 > d->vertex_buffers[buf].size_bytes)
                throw std::invalid_argument("index buffer access out of range");
            uint16_t v;
            std::memcpy(&v, (const uint8_t*)d->vertex_buffers[buf].data + off, 2);
            return (uint32_t)v;
        };
        auto f3at = [&](uint32_t buf, uint64_t off, float o[3]) {
            if (buf >= d->vertex_buffer_count || off + 12 > d->vertex_buffers[buf].size_bytes)
                throw std::invalid_argument("position buffer access out of range");
            std::memcpy(o, (const uint8_t*)d->vertex_buffers[buf].data + off, 12);
        };
        auto xfp = [](const st_transform& T, const float v[3], float o[3]) {
            for (int r = 0; r < 3; r++) o[r] = T.m[r][0] * v[0] + T.m[r][1] * v[1] + T.m[r][2] * v[2] + T.m[r][3];
        };
        uint64_t ntri = 0;
        for (uint32_t i = 0; i < d->instance_count; i++)
            if ((d->instances[i].type_material_address & 0xfu) == 0) ntri += d->instances[i].data >> 16;
        // media: grids, volume instances, inverse transforms
        std::vector<uint64_t> roots(d->volume_count);
        for (uint32_t v = 0; v < d->volume_count; v++) {
            const st_buffer_view& g = d->volumes[v];
            if (!g.data) throw std::invalid_argument("volume data missing");
            roots[v] = validate_nanovdb((const uint8_t*)g.data, g.size_bytes);
            D.vol[v] = s->upload<uint8_t>(g.data, g.size_bytes);
            D.vol_root[v] = roots[v];
            s->vol_bytes.push_back(g.size_bytes);
        }
        D.vol_count = d->volume_count;
        s->inv_xf.resize(d->instance_count + 1);
        for (uint32_t i = 0; i < d->instance_count; i++) s->inv_xf[i] = inverse_transform(d->transforms[i]);
        D.inv_xf = s->upload<st_transform>(s->inv_xf.data(), sizeof(st_transform) * s->inv_xf.size());
        D.vinst_count = 0;
        for (uint32_t i = 0; i < d->instance_count; i++)
            if ((d->instances[i].type_material_address & 0xfu) == 2) D.vinst[D.vinst_count++] = i;
        for (uint32_t i = 0; i < d->volume_info_count; i++) s->volume_info.push_back(d->volume_info[i]);
        D.vinfo = s->upload<st_volume_info>(s->volume_info.data(), sizeof(st_volume_info) * s->volume_info.size());
        D.vinfo_count = (uint32_t)s->volume_info.size();
        std::vector<float> verts(9 * ntri + 9);
        std::vector<uint32_t> ids(ntri + 1), owner(ntri + 1, 0);
        s->boxes.resize(d->instance_count);
        s->inst_first.assign(d->instance_count + 1, 0);
        std::vector<uint32_t> lights;
        uint64_t t = 0;
        for (uint32_t i = 0; i < d->instance_count; i++) {
            const st_instance& in = d->instances[i];
            st_scene::InstBox& ib0 = s->boxes[i];
            ib0.kind = 0;
            s->inst_first[i] = (uint32_t)t;
            if ((in.type_material_address & 0xfu) == 2) {
                // volume box: the grid's world bbox, or the [-1, 1] cube (Scene.cpp:672-680)
                float mn[3] = {-1, -1, -1}, mx[3] = {1, 1, 1};
                if (in.data < d->volume_count) {
                    double bb[6];
                    std::memcpy(bb, (const uint8_t*)d->volumes[in.data].data + 560, 48);
                    for (int k = 0; k < 3; k++) {
                        mn[k] = (float)bb[k];
                        mx[k] = (float)bb[3 + k];
                    }
                }
                ib0.kind = 2;
                for (int k = 0; k < 3; k++) {
                    ib0.lo[k] = mn[k];
                    ib0.hi[k] = mx[k];
                }
                continue;
            }
            const uint32_t addr = in.type_material_address >> 4;
            const uint32_t ew = d->material_data[addr / 4 + 1];
            if ((ew & 0xffu) || ((ew >> 8) & 0xffu) || ((ew >> 16) & 0xffu)) lights.push_back(i);  // Scene.cpp:487-491
            if ((in.type_material_address & 0xfu) == 1) {
                // sphere: the scene box takes the transformed corners of [-r, r]^3
                // (the reference's center +- r, its sphere transforms being translations)
                float r;
                std::memcpy(&r, &in.data, 4);
                s->sph_inst.push_back(i);
                s->sph_rad.push_back(r);
                ib0.kind = 1;
                for (int k = 0; k < 3; k++) {
                    ib0.lo[k] = -r;
                    ib0.hi[k] = r;
                }
                continue;
            }
            const st_mesh_vertex_info& vi = d->mesh_vertex_info[in.data & 0xffffu];
            const uint32_t pc = in.data >> 16;
            const uint32_t ib = vi.packed_buffer_indices[0] & 0xffffu, pb = vi.packed_buffer_indices[0] >> 16;
            const uint32_t is = vi.packed_strides & 0xffu, ps = (vi.packed_strides >> 8) & 0xffu;
            if (is != 2 && is != 4) throw std::invalid_argument("index stride must be 2 or 4");
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t p = 0; p < pc; p++, t++) {
                for (int j = 0; j < 3; j++) {
                    const uint64_t io = (uint64_t)vi.packed_offsets[0] + ((uint64_t)p * 3 + j) * is;
                    const uint32_t idx = is == 2 ? u16at(ib, io) : u32at(ib, io);
                    float v[3];
                    f3at(pb, (uint64_t)vi.packed_offsets[1] + (uint64_t)ps * idx, v);
                    for (int k = 0; k < 3; k++) {
                        lo[k] = std::min(lo[k], v[k]);
                        hi[k] = std::max(hi[k], v[k]);
                    }
                    xfp(d->transforms[i], v, &verts[9 * t + 3 * j]);
                }
                ids[t] = i | (p << 16);
                owner[t] = i;
            }
            if (pc == 0) continue;
            ib0.kind = 1;  // transformed AABB corners, Scene.cpp:622-631
            for (int k = 0; k < 3; k++) {
                ib0.lo[k] = lo[k];
                ib0.hi[k] = hi[k];
            }
        }
        scene_sphere(s->boxes, d->transforms, D.sphere);
        s->ntri = ntri;
        s->tri_inst = s->upload<uint32_t>(owner.data(), 4 * owner.size());
        s->inst_first_dev = s->upload<uint32_t>(s->inst_first.data(), 4 * s->inst_first.size());
        D.tri_base = s->inst_first_dev;  // the per-triangle albedo's index base of each instance
        D.tri_alb = nullptr;
        D.light_inst = s->upload<uint32_t>(lights.data(), 4 * lights.size());
        D.light_count = (uint32_t)lights.size();
        if (D.light_count > 0) {
            // the emitter table: per light its first triangle in the table
            std::vector<uint32_t> base(lights.size());
            uint64_t total = 0;
            for (size_t li = 0; li < lights.size(); li++) {
                base[li] = (uint32_t)total;
                if ((d->instances[lights[li]].type_material_address & 0xfu) == 0) total += d->instances[lights[li]].data >> 16;
            }
            if (total >= (1ull << 31)) throw std::domain_error("too many emissive triangles");
            s->light_tri_base = s->upload<uint32_t>(base.data(), 4 * base.size());
            s->light_rec = s->alloc<float4>(64 * lights.size());
            s->light_tri = s->alloc<float4>(80 * std::max<uint64_t>(1, total));
            D.light_rec = s->light_rec;
            D.light_tri = s->light_tri;
            s->build_light_table(ctx->stream);
        }
        s->build_sphere_tree(d->transforms);
        build_material_images(s.get(), d);
        BvhOut bvh = build_bvh(verts.data(), ids.data(), (uint32_t)ntri);
        if (D.has_alpha)
            for (TriOut& t : bvh.tris4) {
                uint32_t id;
                std::memcpy(&id, &t.f[3], 4);
                const uint32_t alpha = d->material_data[(d->instances[id & 0xffffu].type_material_address >> 4) / 4 + 12] < ST_IMAGE_LIMIT;
                std::memcpy(&t.f[7], &alpha, 4);
            }
        D.nodes = s->upload<float4>(bvh.nodes4.data(), bvh.nodes4.size() * sizeof(Node4Out));
        D.tris = s->upload<float4>(bvh.tris4.data(), bvh.tris4.size() * sizeof(TriOut));
        D.node_count = ntri ? (uint32_t)bvh.nodes4.size() : 0;
        D.tri_count = (uint32_t)bvh.tris4.size();
        if (3 * bvh.depth4 + 1 > kMaxStack) throw std::runtime_error("BVH deeper than the traversal stack");
        s->stats[0] = D.node_count;
        s->stats[1] = D.tri_count;
        s->stats[2] = (uint64_t)bvh.depth4;
        s->stats[3] = bvh.nodes4.size() * sizeof(Node4Out) + bvh.tris4.size() * sizeof(TriOut);
        // environment: record (EnvironmentMap::store, Scene.hpp:42-45), mip chain, alias table
        D.has_env = d->environment_material_address != 0xFFFFFFFFu;
        if (D.has_env) {
            const uint32_t a = d->environment_material_address / 4;
            if ((uint64_t)a + 4 > d->material_word_count) throw std::invalid_argument("environment address out of range");
            for (int k = 0; k < 3; k++) std::memcpy(&D.env_emission[k], &d->material_data[a + k], 4);
            const uint32_t img = d->material_data[a + 3];
            if (img < ST_IMAGE_LIMIT) {
                if (img >= d->image_count) throw std::invalid_argument("environment image index out of range");
                const st_image_desc& im = d->images[img];
                if (im.width == 0 || im.height == 0 || !im.rgba) throw std::invalid_argument("empty environment image");
                D.env_has_image = 1;
                std::vector<std::vector<float>> lv(1);
                std::vector<int> lw{(int)im.width}, lh{(int)im.height};
                lv[0].assign(im.rgba, im.rgba + 4ull * im.width * im.height);
                while (lw.back() > 1 || lh.back() > 1) {  // 2x2 box mips, Image.cpp:282-308
                    const int w0 = lw.back(), h0 = lh.back(), w1 = w0 > 1 ? w0 / 2 : 1, h1 = h0 > 1 ? h0 / 2 : 1;
                    std::vector<float> nx(4ull * w1 * h1);
                    const std::vector<float>& src = lv.back();
                    for (int y = 0; y < h1; y++)
                        for (int x = 0; x < w1; x++) {
                            const int x0 = w0 > 1 ? 2 * x : 0, x1 = w0 > 1 ? 2 * x + 1 : 0;
                            const int y0 = h0 > 1 ? 2 * y : 0, y1 = h0 > 1 ? 2 * y + 1 : 0;
                            for (int ch = 0; ch < 4; ch++) {
                                auto T = [&](int xx, int yy) { return src[4 * ((size_t)yy * w0 + xx) + ch]; };
                                nx[4 * ((size_t)y * w1 + x) + ch] = ((T(x0, y0) + T(x1, y0)) + (T(x0, y1) + T(x1, y1))) * 0.25f;
                            }
                        }
                    lv.push_back(std::move(nx));
                    lw.push_back(w1);
                    lh.push_back(h1);
                }
                if ((int)lv.size() > kEnvMaxLevels) throw std::invalid_argument("environment image too large");
                D.env_levels = (int)lv.size();
                for (size_t l = 0; l < lv.size(); l++) {
                    D.env_level[l] = s->upload<float4>(lv[l].data(), lv[l].size() * 4);
                    D.env_w[l] = lw[l];
                    D.env_h[l] = lh[l];
                }
                // alias table over luminance * sin(theta) of level 0 (Walker/Vose)
                const int W = lw[0], H = lh[0];
                const size_t N = (size_t)W * H;
                std::vector<double> wgt(N);
                double total = 0;
                for (int y = 0; y < H; y++) {
                    const double st = std::sin(M_PI * (y + 0.5) / H);
                    for (int x = 0; x < W; x++) {
                        const float* px = &lv[0][4 * ((size_t)y * W + x)];
                        const double l = 0.2126 * px[0] + 0.7152 * px[1] + 0.0722 * px[2];
                        wgt[(size_t)y * W + x] = std::max(0.0, l) * st;
                        total += wgt[(size_t)y * W + x];
                    }
                }
                std::vector<float2> alias(N);
                std::vector<float> pdf(N);
                if (total <= 0) {
                    for (size_t i = 0; i < N; i++) wgt[i] = 1;
                    total = (double)N;
                }
                std::vector<double> sc(N);
                std::vector<uint32_t> small, large;
                for (size_t i = 0; i < N; i++) {
                    sc[i] = wgt[i] / total * N;
                    pdf[i] = (float)(wgt[i] / total * N);  // uv-space density of texel i
                    (sc[i] < 1 ? small : large).push_back((uint32_t)i);
                }
                while (!small.empty() && !large.empty()) {
                    uint32_t sidx = small.back(), lidx = large.back();
                    small.pop_back();
                    alias[sidx] = make_float2((float)sc[sidx], 0);
                    uint32_t al = lidx;
                    std::memcpy(&alias[sidx].y, &al, 4);
                    sc[lidx] = (sc[lidx] + sc[sidx]) - 1;
                    if (sc[lidx] < 1) {
                        large.pop_back();
                        small.push_back(lidx);
                    }
                }
                for (uint32_t i : large) {
                    alias[i] = make_float2(1.0f, 0);
                    std::memcpy(&alias[i].y, &i, 4);
                }
                for (uint32_t i : small) {
                    alias[i] = make_float2(1.0f, 0);
                    std::memcpy(&alias[i].y, &i, 4);
                }
                // the alias sampler's in-texel jitter stops short of 1 by a margin
                // that keeps x + j / W and its evaluation (int)(u * W) inside texel
                // x at this map size: 1 - 2^(ceil(log2(max(W, H))) - 21)
                int lg = 0;
                while ((1 << lg) < std::max(W, H)) lg++;
                D.env_jmax = 1.0f - std::ldexp(1.0f, lg - 21);
                D.env_alias = s->upload<float2>(alias.data(), N * sizeof(float2));
                D.env_pdf = s->upload<float>(pdf.data(), N * sizeof(float));
            }
        }
        *out = s.release();
        return ST_OK;
    } catch (const std::domain_error& e) {
        return set_error(ST_ERROR_UNSUPPORTED, e.what());
    } catch (const std::invalid_argument& e) {
        return set_error(ST_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc& e) {
        return set_error(ST_ERROR_OUT_OF_MEMORY, e.what());
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_scene_update(st_scene* s, const st_transform* transforms, double* build_ms) {
    if (s && s->d.has_alpha) return set_error(ST_ERROR_UNSUPPORTED, "st_scene_update with alpha masks is not supported");
    if (!s || !transforms) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        s->generation++;
        DevScene& D = s->d;
        const uint32_t ni = D.inst_count;
        // transforms, their inverses and the scene sphere
        HIP_CHECK(hipMemcpy(const_cast<st_transform*>(D.xf), transforms, sizeof(st_transform) * ni, hipMemcpyHostToDevice));
        for (uint32_t i = 0; i < ni; i++) s->inv_xf[i] = inverse_transform(transforms[i]);
        HIP_CHECK(hipMemcpy(const_cast<st_transform*>(D.inv_xf), s->inv_xf.data(), sizeof(st_transform) * ni, hipMemcpyHostToDevice));
        scene_sphere(s->boxes, transforms, D.sphere);
        s->build_light_table(ctx->stream);  // the emitters' transforms moved too
        s->build_sphere_tree(transforms);
        double ms = 0;
        if (s->ntri > 0) {
            // flatten on the device and rebuild the BVH there
            DevBuf flat;
            flat.reserve(48 * s->ntri);
            const uint32_t n = (uint32_t)s->ntri;
            const int grid = std::min<int>((n + 255) / 256, ctx->cus * 16);
            hipLaunchKernelGGL(k_flatten, dim3(grid), dim3(256), 0, ctx->stream, D, s->tri_inst, s->inst_first_dev, n,
                               flat.as<float4>());
            HIP_CHECK(hipGetLastError());
            LbvhResult r = build_lbvh(flat.as<float4>(), n, ctx->stream);
            std::vector<Node4Out> hn(r.node_count);
            HIP_CHECK(hipMemcpy(hn.data(), r.nodes, sizeof(Node4Out) * hn.size(), hipMemcpyDeviceToHost));
            const std::string bad = (3 * r.depth4 + 1 > kMaxStack) ? std::string("tree deeper than the traversal stack")
                                                                     : check_bvh4(hn, r.tri_count);
            if (!bad.empty() || r.tri_count != n) {
                (void)hipFree(r.nodes);
                (void)hipFree(r.tris);
                return set_error(ST_ERROR_LIMIT, "device BVH rejected: " + (bad.empty() ? std::string("triangle count") : bad));
            }
            if (s->bvh_nodes) (void)hipFree(s->bvh_nodes);
            if (s->bvh_tris) (void)hipFree(s->bvh_tris);
            s->bvh_nodes = r.nodes;
            s->bvh_tris = r.tris;
            D.nodes = r.nodes;
            D.tris = r.tris;
            D.node_count = r.node_count;
            D.tri_count = r.tri_count;
            s->stats[0] = r.node_count;
            s->stats[1] = r.tri_count;
            s->stats[2] = (uint64_t)r.depth4;
            s->stats[3] = (uint64_t)r.node_count * sizeof(Node4Out) + (uint64_t)r.tri_count * sizeof(TriOut);
            ms = r.ms;
        }
        if (build_ms) *build_ms = ms;
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_scene_destroy(st_scene* s) {
    if (!s) return ST_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return ST_OK;
}

int st_scene_stats(const st_scene* s, uint64_t out[4]) {
    if (!s || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 4; i++) out[i] = s->stats[i];
    return ST_OK;
}

using ShadeFn = void (*)(DevScene, RenderConsts, WaveState, int, AuxOut, int, AdjArgs);
// the path replay of st_adjoint (plain NEE)
static ShadeFn shade_kernel_adj(bool lambert, bool alias, bool media) {
    static const ShadeFn t[8] = {
        k_shade<false, false, false, false, true>, k_shade<false, true, false, false, true>,
        k_shade<true, false, false, false, true>,  k_shade<true, true, false, false, true>,
        k_shade<false, false, false, true, true>,  k_shade<false, true, false, true, true>,
        k_shade<true, false, false, true, true>,   k_shade<true, true, false, true, true>};
    return t[(media ? 4 : 0) | (lambert ? 2 : 0) | (alias ? 1 : 0)];
}
static ShadeFn shade_kernel(bool lambert, bool alias, bool restir, bool media, bool tex = false) {
    if (tex) {  // material images: no media (refused with them)
        static const ShadeFn tt[8] = {
            k_shade<false, false, false, false, false, false, true>, k_shade<false, false, true, false, false, false, true>,
            k_shade<false, true, false, false, false, false, true>,  k_shade<false, true, true, false, false, false, true>,
            k_shade<true, false, false, false, false, false, true>,  k_shade<true, false, true, false, false, false, true>,
            k_shade<true, true, false, false, false, false, true>,   k_shade<true, true, true, false, false, false, true>};
        return tt[(lambert ? 4 : 0) | (alias ? 2 : 0) | (restir ? 1 : 0)];
    }
    if (media) {  // RIS at every vertex (no reuse) with media
        static const ShadeFn mt[8] = {k_shade<false, false, false, true>, k_shade<false, false, true, true>,
                                      k_shade<false, true, false, true>,  k_shade<false, true, true, true>,
                                      k_shade<true, false, false, true>,  k_shade<true, false, true, true>,
                                      k_shade<true, true, false, true>,   k_shade<true, true, true, true>};
        return mt[(lambert ? 4 : 0) | (alias ? 2 : 0) | (restir ? 1 : 0)];
    }
    const int i = (lambert ? 4 : 0) | (alias ? 2 : 0) | (restir ? 1 : 0);
    static const ShadeFn table[8] = {
        k_shade<false, false, false, false>, k_shade<false, false, true, false>, k_shade<false, true, false, false>,
        k_shade<false, true, true, false>,   k_shade<true, false, false, false>, k_shade<true, false, true, false>,
        k_shade<true, true, false, false>,   k_shade<true, true, true, false>};
    return table[i];
}

static ShadeFn shade_kernel_cache(bool lambert, bool alias, bool media) {
    static const ShadeFn t[8] = {
        k_shade<false, false, false, false, false, true>, k_shade<false, true, false, false, false, true>,
        k_shade<true, false, false, false, false, true>,  k_shade<true, true, false, false, false, true>,
        k_shade<false, false, false, true, false, true>,  k_shade<false, true, false, true, false, true>,
        k_shade<true, false, false, true, false, true>,   k_shade<true, true, false, true, false, true>};
    return t[(media ? 4 : 0) | (lambert ? 2 : 0) | (alias ? 1 : 0)];
}

static void ensure_spill(WaveBufs& b, int grid, int block) {
    b.spill.reserve((size_t)(kMaxStack - kLdsStack + 1) * grid * block * 4);
}

// Buffers of one wave slot for `paths` paths (reuse: k samples of npix pixels).
static void reserve_wave(WaveBufs& b, uint64_t paths, bool reuse, uint64_t npix_k, int tgrid) {
    b.ray.reserve(paths * 16);
    b.thr.reserve(paths * 16);
    b.stt.reserve(paths * 16);
    b.em.reserve(paths * 16);
    b.nee.reserve(paths * 16);
    b.hit.reserve(paths * 16);
    b.q0.reserve(paths * 4);
    b.q1.reserve(paths * 4);
    b.shadow.reserve(paths * 48);
    b.med.reserve(paths * 4);
    b.counters.reserve(4 * kCounterWords);
    if (reuse) {
        b.pri.reserve(npix_k * 16);
        b.psd.reserve(npix_k * 32);
        b.rinit.reserve(npix_k * 32);
        b.rtemp.reserve(npix_k * 32);
    }
    ensure_spill(b, tgrid, kTraceBlock);
}

static WaveState wave_state(WaveBufs& b, st_context* ctx) {
    WaveState w;
    std::memset(&w, 0, sizeof w);
    w.ray = b.ray.as<float4>();
    w.thr = b.thr.as<float4>();
    w.st = b.stt.as<uint4>();
    w.em = b.em.as<float4>();
    w.nee = b.nee.as<uint4>();
    w.hit = b.hit.as<uint4>();
    w.queue[0] = b.q0.as<uint32_t>();
    w.queue[1] = b.q1.as<uint32_t>();
    w.shadow = b.shadow.as<float4>();
    w.counters = b.counters.as<uint32_t>();
    w.pri = b.pri.as<uint4>();
    w.psd = b.psd.as<float4>();
    w.r_init = b.rinit.as<float4>();
    w.r_temp = b.rtemp.as<float4>();
    w.r_prev = ctx->rprev.as<float4>();
    w.med = b.med.as<uint32_t>();
    w.vcm = b.vcm.as<float2>();
    w.lv_raw = b.lv_raw.as<float4>();
    w.lv = b.lv.as<float4>();
    w.lv_cnt = b.lv_cnt.as<uint32_t>();
    w.lv_off = b.lv_off.as<uint32_t>();
    w.lv_total = b.lv_total.as<uint32_t>();
    return w;
}

// Per-class device times of the last render (HIP events), traversal counters
// and the rays of both wave slots; writes the ray counts to `ray_counts`.
static void begin_timings(st_context* ctx) {
    for (int kk = 0; kk < st_context::kKinds; kk++) ctx->ev_used[kk] = 0;
}
static void finish_timings(st_context* ctx, int slots, uint64_t* ray_counts) {
    for (int kind = 0; kind < st_context::kKinds; kind++) {
        double ms = 0;
        for (size_t i = 0; i < ctx->ev_used[kind]; i++) {
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, ctx->ev[kind][i].first, ctx->ev[kind][i].second));
            ms += t;
        }
        ctx->kernel_ms[kind] = ms;
        if (kind < 3) {
            ctx->timings[kind] = ms;
            ctx->timings[3 + kind] = (double)ctx->ev_used[kind];
        }
    }
    unsigned long long hs[4], rc[2] = {0, 0};
    HIP_CHECK(hipMemcpy(hs, ctx->stats.p, 32, hipMemcpyDeviceToHost));
    ctx->cache_queries = 0;
    for (int sl = 0; sl < slots; sl++) {
        unsigned long long r[3];
        HIP_CHECK(hipMemcpy(r, ctx->slot[sl].counters.as<uint8_t>() + 16, 24, hipMemcpyDeviceToHost));
        rc[0] += r[0];
        rc[1] += r[1];
        ctx->cache_queries += (double)r[2];
    }
    for (int i = 0; i < 4; i++) ctx->timings[6 + i] = (double)hs[i];
    ctx->timings[10] = (double)rc[0];
    ctx->timings[11] = (double)rc[1];
    if (ray_counts) HIP_CHECK(hipMemcpy(ray_counts, rc, 16, hipMemcpyHostToDevice));
}

static int render_light(st_scene* s, const st_view* view, const st_render_params* p, const st_outputs* o);
static int render_vc(st_scene* s, const st_view* view, const st_render_params* p, const st_outputs* o);

// Argument checks of a camera-path render (the path tracer of
// TestRenderer::render); ST_OK when it can run.
// A per-triangle albedo serves Lambertian camera paths with plain NEE.
static int check_triangle_albedo(const st_scene* s, uint32_t flags) {
    if (!s->d.tri_alb) return ST_OK;
    if (!(flags & ST_FLAG_LAMBERTIAN))
        return set_error(ST_ERROR_UNSUPPORTED, "a per-triangle albedo needs ST_FLAG_LAMBERTIAN");
    if (flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE))
        return set_error(ST_ERROR_UNSUPPORTED, "a per-triangle albedo serves camera paths with plain NEE");
    return ST_OK;
}

static int check_camera(st_scene* s, const st_render_params* p) {
    if ((p->flags & ST_FLAG_RESTIR_DI_REUSE) && !(p->flags & ST_FLAG_RESTIR_DI))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "ST_FLAG_RESTIR_DI_REUSE needs ST_FLAG_RESTIR_DI");
    if ((p->flags & ST_FLAG_RESTIR_DI_REUSE) && (p->spatial_neighbours > 64 || !(p->spatial_radius >= 0) || p->spatial_radius > 4096))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "spatial reuse: at most 64 neighbours in a radius of at most 4096 px");
    if ((p->flags & ST_FLAG_RESTIR_DI) && (p->candidate_samples == 0 || p->candidate_samples > 1024))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "candidate_samples must be in [1, 1024]");
    if ((uint64_t)p->width * p->height > (1ull << 28)) return set_error(ST_ERROR_LIMIT, "image too large");
    if (p->max_null_collisions > (1u << 20)) return set_error(ST_ERROR_INVALID_ARGUMENT, "max_null_collisions above 2^20");

    if (s->d.vinst_count > 0 && (p->flags & (ST_FLAG_RESTIR_DI_REUSE | ST_FLAG_RESTIR_DI_HASHGRID)))
        return set_error(ST_ERROR_UNSUPPORTED, "ReSTIR DI reuse is not supported in scenes with media (RIS is)");
    if (p->flags & ST_FLAG_RESTIR_DI_HASHGRID) {
        if (!(p->flags & ST_FLAG_RESTIR_DI) || (p->flags & ST_FLAG_RESTIR_DI_REUSE))
            return set_error(ST_ERROR_INVALID_ARGUMENT, "ST_FLAG_RESTIR_DI_HASHGRID needs ST_FLAG_RESTIR_DI and excludes ST_FLAG_RESTIR_DI_REUSE");
        if (!(p->hashgrid_cell_size >= 0) || p->hashgrid_cell_size > 1e30f)
            return set_error(ST_ERROR_INVALID_ARGUMENT, "hashgrid_cell_size must be a finite value >= 0");
        if (p->max_depth < 2) return set_error(ST_ERROR_INVALID_ARGUMENT, "hash-grid reuse needs max_depth >= 2");
        // two entries of 16 words per pixel, cleared by one 32-bit count
        if ((uint64_t)p->width * p->height * 32 >= (1ull << 32)) return set_error(ST_ERROR_LIMIT, "image too large for the hash grid");
    }
    return check_triangle_albedo(s, p->flags);
}

// Samples per wave of a camera-path render: params.samples_per_wave, or as
// many as fill about kWavePaths paths; 0 when a wave would hold 2^31 paths.
static uint32_t wave_samples(const st_render_params* p, uint32_t want) {
    const uint32_t npix = p->width * p->height;
    uint32_t k = want;
    if (k == 0) k = std::max<uint32_t>(1, std::min<uint32_t>(p->spp, kWavePaths / npix));
    k = std::min(k, p->spp);
    return (uint64_t)npix * k >= (1ull << 31) ? 0 : k;
}

// The radiance cache of an st_render_cached call: the network the queries
// read and the training rows the training paths write.
struct CacheJob {
    CacheGrid grid;
    const _Float16* table;  // fp16 table and weights
    const _Float16* w16;
    float* tq;              // rows x 12
    float4* tthr;           // rows
    float4* tt;             // rows
    uint32_t rows, depth, stride, frame;
};

// One camera-path render (TestRenderer::render's dispatches, testrenderer.slang:
// 1087-1183) enqueued on given wave slots and accumulators.  st_render runs
// it on the context's four slots; st_batch_render on one slot per request,
// inside a captured graph.  Clears are kernels, not memsets, so the sequence
// can be captured.
struct CameraJob {
    st_scene* s;
    st_view view;
    st_render_params p;
    st_outputs o;
    WaveBufs* slot[kWaveSlots];      // wave slots: buffers and streams
    uint32_t* counters[kWaveSlots];  // the slots' counter blocks (kCounterWords words)
    int slots;
    uint32_t k;                      // samples per wave
    float4* acc;                     // per pixel: the sum over samples
    float4* rprev;                   // per pixel: final reservoirs of the last sample (reuse)
    unsigned long long* stats;       // traversal statistics (COUNT_TRAVERSAL)
    int tgrid, sgrid;                // trace and shade grids
    bool timed;                      // HIP-event timings of the kernel classes
    bool store_rays;                 // write the ray counts to o.ray_counts on the device
    const CacheJob* cache;           // st_render_cached, else null
    bool hashgrid;                   // ST_FLAG_RESTIR_DI_HASHGRID: one sample per wave, one slot
    int hg_first;                    // grid buffer of the call's first sample (the other holds the previous grid)
    bool hg_prev_valid;              // the previous grid (from the previous call) may be reused by the first sample
};


__global__ void k_store_rays(const uint32_t* counters, unsigned long long* out) {
    if (threadIdx.x < 2) out[threadIdx.x] = reinterpret_cast<const unsigned long long*>(counters + 4)[threadIdx.x];
}
// the ray counts of an ST_FLAG_ASYNC call: the sum over its wave slots' counters
struct SlotCounters {
    const uint32_t* c[kWaveSlots];
};
__global__ void k_sum_rays(SlotCounters sc, int slots, unsigned long long* out) {
    if (threadIdx.x < 2) {
        unsigned long long v = 0;
        for (int sl = 0; sl < slots; sl++) v += reinterpret_cast<const unsigned long long*>(sc.c[sl] + 4)[threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// The constants of a camera-path render shared by every wave.
static RenderConsts camera_consts(const st_scene* s, const st_view& view, const st_render_params* p) {
    RenderConsts c;
    std::memset(&c, 0, sizeof c);
    c.W = p->width;
    c.H = p->height;
    c.npix = p->width * p->height;
    c.seed = p->seed;
    c.max_depth = p->max_depth;
    c.min_len = p->min_path_length;
    c.flags = p->flags;
    c.rr_start = p->rr_start_depth;
    c.env_prob = p->environment_sample_probability;
    c.view = view;
    c.inv_cam = inverse_transform(view.camera_to_world);
    c.max_null = p->max_null_collisions;
    // the camera's medium: the last volume whose object-space bounds hold the
    // camera position (TestRenderer.cpp:431-446)
    c.view_medium = kNoMedium;
    const st_transform& T = view.camera_to_world;
    const float cam[3] = {T.m[0][3], T.m[1][3], T.m[2][3]};
    for (const st_volume_info& vi : s->volume_info) {
        if (vi.instance_index >= s->d.inst_count) continue;
        const st_transform& iT = s->inv_xf[vi.instance_index];
        float lp[3];
        for (int r = 0; r < 3; r++) lp[r] = iT.m[r][0] * cam[0] + iT.m[r][1] * cam[1] + iT.m[r][2] * cam[2] + iT.m[r][3];
        if (lp[0] >= vi.min[0] && lp[1] >= vi.min[1] && lp[2] >= vi.min[2] && lp[0] <= vi.max[0] && lp[1] <= vi.max[1] &&
            lp[2] <= vi.max[2])
            c.view_medium = vi.instance_index;
    }
    c.tiled = (p->width % 8 == 0 && p->height % 8 == 0) ? 1 : 0;
    c.cand = p->candidate_samples;
    c.max_m = p->max_m;
    c.spatial = p->spatial_neighbours;
    c.radius = p->spatial_radius;
    return c;
}

static void enqueue_camera(st_context* ctx, const CameraJob& j) {
    st_scene* s = j.s;
    const st_render_params* p = &j.p;
    const st_outputs* o = &j.o;
    const uint32_t npix = p->width * p->height, k = j.k;
    const bool reuse = (p->flags & ST_FLAG_RESTIR_DI_REUSE) != 0;
    const bool media = s->d.vinst_count > 0;
    const int slots = j.slots, tgrid = j.tgrid, sgrid = j.sgrid;
    hipStream_t s0 = j.slot[0]->stream;
    const int fgrid = std::min<int>((int)((npix * 4 + 255) / 256), ctx->cus * 8);
    hipLaunchKernelGGL(k_fill_u32, dim3(fgrid), dim3(256), 0, s0, reinterpret_cast<uint32_t*>(j.acc), npix * 4, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s0, reinterpret_cast<uint32_t*>(j.stats), 16u, 0u);
    for (int sl = 0; sl < slots; sl++)
        hipLaunchKernelGGL(k_fill_u32, dim3(kCounterWords / 256), dim3(256), 0, s0, j.counters[sl], (uint32_t)kCounterWords, 0u);
    if (slots > 1) {
        HIP_CHECK(hipEventRecord(ctx->start_done, s0));
        for (int sl = 1; sl < slots; sl++) HIP_CHECK(hipStreamWaitEvent(j.slot[sl]->stream, ctx->start_done, 0));
    }
    RenderConsts c = camera_consts(s, j.view, p);
    const CacheJob* cj = j.cache;
    if (cj) {
        c.cache_depth = cj->depth;
        c.train_stride = cj->stride;
        c.train_frame = cj->frame;
    }
    if (j.hashgrid) {
        c.hg_cell = p->hashgrid_cell_size > 0 ? p->hashgrid_cell_size : 0.5f;
        c.hg_dmax = std::min<uint32_t>(2, p->max_depth - 1);
        c.hg_nslots = npix * c.hg_dmax;
    }
    const bool lambert = (p->flags & (ST_FLAG_LAMBERTIAN | ST_FLAG_LAMBERT_GGX)) != 0, alias = (p->flags & ST_FLAG_ENV_ALIAS) != 0;
    const bool restir = (p->flags & ST_FLAG_RESTIR_DI) != 0;
    const bool count = (p->flags & ST_FLAG_COUNT_TRAVERSAL) != 0;
    unsigned long long* stats = j.stats;
    // Wave i runs in slot i % slots on that slot's stream; the slot's
    // previous wave (i - slots) finished before it on the same stream.  Across slots,
    // two orders matter: the screen-space reuse of wave i reads the final
    // reservoirs of wave i - 1, and resolves add samples in sample order.
    uint32_t wi = 0;
    for (uint32_t base = 0; base < p->spp; base += k, wi++) {
        WaveBufs& B = *j.slot[wi % slots];
        WaveBufs& prev = *j.slot[(wi + slots - 1) % slots];
        hipStream_t sm = B.stream;
        WaveState w = wave_state(B, ctx);
        w.counters = j.counters[wi % slots];
        w.r_prev = j.rprev;
        if (cj) {
            w.cq = B.cq.as<float4>();
            w.tq = cj->tq;
            w.tthr = cj->tthr;
            w.tt = cj->tt;
        }
        const int hg_cur = (j.hg_first + wi) & 1;
        if (j.hashgrid) {  // this sample's grid, and the previous sample's to reuse from
            c.hg_prev = (base > 0 || j.hg_prev_valid) ? 1u : 0u;
            w.hg_rec = ctx->hg_rec[hg_cur].as<float4>();
            w.hg_prev_rec = ctx->hg_rec[hg_cur ^ 1].as<float4>();
            w.hg_prev_keys = ctx->hg_keys[hg_cur ^ 1].as<unsigned long long>();
            w.hg_prev_slots = ctx->hg_slots[hg_cur ^ 1].as<uint32_t>();
            hipLaunchKernelGGL(k_fill_u32, dim3(ctx->cus * 8), dim3(256), 0, sm, ctx->hg_rec[hg_cur].as<uint32_t>(), c.hg_nslots * 16, 0u);
        }
        uint32_t* spill = B.spill.as<uint32_t>();
        c.k = std::min(k, p->spp - base);
        c.sample_base = base;
        c.wave_paths = npix * c.k;
        const int cgrid = std::min<int>((c.wave_paths + kShadeBlock - 1) / kShadeBlock, ctx->cus * 16);
        hipLaunchKernelGGL(k_camera, dim3(cgrid), dim3(kShadeBlock), 0, sm, c, w);
        int qsel = 0;
        for (uint32_t b = 0; b < std::max<uint32_t>(1, p->max_depth); b++) {
            {
                Timed tm(j.timed ? ctx : nullptr, 0, sm);
                if (media) hipLaunchKernelGGL(k_trace_media<false>, dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, c, w, qsel, spill, stats, AdjArgs{});
                else if (count) hipLaunchKernelGGL(closest_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
                else hipLaunchKernelGGL(closest_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
            }
            // with screen-space reuse, k_ris runs the primary vertices' RIS
            // first and every vertex after takes plain NEE: the shading
            // kernel without the candidate loop
            if (reuse && b == 0) {
                Timed tm(j.timed ? ctx : nullptr, 3, sm);
                const bool sph = s->d.sph_node_count != 0;
                auto kr = sph ? (lambert ? (alias ? k_ris<true, true, true> : k_ris<true, false, true>)
                                         : (alias ? k_ris<false, true, true> : k_ris<false, false, true>))
                              : (lambert ? (alias ? k_ris<true, true> : k_ris<true, false>)
                                         : (alias ? k_ris<false, true> : k_ris<false, false>));
                hipLaunchKernelGGL(kr, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, c, w, qsel);
            }
            {
                Timed tm(j.timed ? ctx : nullptr, 1, sm);
                AuxOut aux{nullptr, nullptr, nullptr, nullptr};
                if (base == 0) aux = AuxOut{o->visibility, o->albedo, o->depth, o->prev_uv};
                auto* kfn = cj ? shade_kernel_cache(lambert, alias, media) : shade_kernel(lambert, alias, restir && (!reuse || (p->flags & ST_FLAG_RESTIR_DI_DEEP)), media, s->d.has_tex != 0);
                hipLaunchKernelGGL(kfn, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, c, w, qsel, aux, base == 0 ? 1 : 0,
                                   AdjArgs{});
            }
            if (cj && b + 1 == cj->depth) {  // the paths that ended in the cache take its radiance
                Timed tm(j.timed ? ctx : nullptr, 5, sm);
                const int igrid = std::min<int>((int)((c.wave_paths + 63) / 64), ctx->cus * 4);
                hipLaunchKernelGGL(k_rc_infer_paths, dim3(igrid), dim3(256), 0, sm, cj->grid, cj->table, cj->w16, w.cq,
                                   w.counters + 3, w.em);
            }
            if (reuse && b == 0) {
                const int rgrid2 = std::min<int>((npix + kShadeBlock - 1) / kShadeBlock, ctx->cus * 8);
                auto kt = lambert ? (alias ? k_restir_temporal<true, true> : k_restir_temporal<true, false>)
                                  : (alias ? k_restir_temporal<false, true> : k_restir_temporal<false, false>);
                auto ks = lambert ? (alias ? k_restir_spatial<true, true> : k_restir_spatial<true, false>)
                                  : (alias ? k_restir_spatial<false, true> : k_restir_spatial<false, false>);
                if (wi > 0 && slots > 1) HIP_CHECK(hipStreamWaitEvent(sm, prev.reuse_done, 0));
                // temporal reuse chains consecutive samples: sample j of the wave
                // reuses sample j - 1's final reservoirs
                for (uint32_t jj = 0; jj < c.k; jj++) {
                    Timed tm(j.timed ? ctx : nullptr, 4, sm);
                    hipLaunchKernelGGL(kt, dim3(rgrid2), dim3(kShadeBlock), 0, sm, s->d, c, w, (base + jj) > 0 ? 1 : 0, jj);
                    hipLaunchKernelGGL(ks, dim3(rgrid2), dim3(kShadeBlock), 0, sm, s->d, c, w, jj);
                }
                if (slots > 1) HIP_CHECK(hipEventRecord(B.reuse_done, sm));
            }
            {
                Timed tm(j.timed ? ctx : nullptr, 2, sm);
                if (media) hipLaunchKernelGGL(k_shadow_media<false>, dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, c, w, spill, stats, AdjArgs{});
                else if (count) hipLaunchKernelGGL(shadow_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, c.flags, spill, stats);
                else hipLaunchKernelGGL(shadow_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, c.flags, spill, stats);
            }
            hipLaunchKernelGGL(k_count, dim3(1), dim3(1), 0, sm, w, qsel, media ? 1 : 0);
            qsel ^= 1;
        }
        if (j.hashgrid) {  // this sample's grid: entries sorted by cell, stable in slot order
            const int hgrid = std::min<int>((int)((c.hg_nslots + 255) / 256), ctx->cus * 8);
            hipLaunchKernelGGL(k_hg_keys, dim3(hgrid), dim3(256), 0, sm, w.hg_rec, c.hg_nslots, ctx->hg_keys_tmp.as<unsigned long long>(),
                               ctx->hg_slots_tmp.as<uint32_t>());
            size_t sb = ctx->hg_sort_tmp.n;
            HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(ctx->hg_sort_tmp.p, sb, ctx->hg_keys_tmp.as<unsigned long long>(),
                                                         ctx->hg_keys[hg_cur].as<unsigned long long>(), ctx->hg_slots_tmp.as<uint32_t>(),
                                                         ctx->hg_slots[hg_cur].as<uint32_t>(), (int)c.hg_nslots, 0, 33, sm));
        }
        if (cj) {  // training targets of the rows whose path is in this wave
            const uint64_t g0 = (uint64_t)base * npix, g1 = g0 + (uint64_t)c.k * npix;
            const uint32_t lo = (uint32_t)(g0 / cj->stride), hi = (uint32_t)std::min<uint64_t>(cj->rows, (g1 - 1) / cj->stride + 1);
            if (hi > lo) {
                Timed tm(j.timed ? ctx : nullptr, 5, sm);
                hipLaunchKernelGGL(k_rc_targets, dim3(std::min<int>((int)((hi - lo + 255) / 256), ctx->cus * 4)), dim3(256), 0, sm, c, w,
                                   lo, hi);
            }
        }
        const bool last = base + c.k >= p->spp;
        const int rgrid = std::min<int>((npix + 255) / 256, ctx->cus * 8);
        if (wi > 0 && slots > 1) HIP_CHECK(hipStreamWaitEvent(sm, prev.resolve_done, 0));
        hipLaunchKernelGGL(k_resolve, dim3(rgrid), dim3(256), 0, sm, c, w, j.acc, o->hdr, o->path_length,
                           1.0f / (float)p->spp, last ? 1 : 0);
        if (slots > 1) HIP_CHECK(hipEventRecord(B.resolve_done, sm));
        HIP_CHECK(hipGetLastError());
    }
    if (j.store_rays && o->ray_counts)  // one slot: its counters hold every wave's rays
        hipLaunchKernelGGL(k_store_rays, dim3(1), dim3(64), 0, s0, j.counters[0], (unsigned long long*)o->ray_counts);
}

// Material images serve camera paths with plain NEE or RIS, in scenes without media.
static int check_images(const st_scene* s, uint32_t flags) {
    if (!s->d.has_tex) return ST_OK;
    // (alpha masks are material images too: the same modes, and no media)
    if (s->d.vinst_count > 0 || (flags & (ST_FLAG_RESTIR_DI_REUSE | ST_FLAG_RESTIR_DI_HASHGRID | ST_FLAG_TRACE_FROM_LIGHT |
                                           ST_FLAG_BDPT_LVC)))
        return set_error(ST_ERROR_UNSUPPORTED, "material images serve camera paths with plain NEE or RIS, without media or reuse");
    return ST_OK;
}

int st_render(st_scene* s, const st_view* view, const st_render_params* p, const st_outputs* o) {
    if (!s || !view || !p || !o || !o->hdr) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (check_images(s, p->flags) != ST_OK) return ST_ERROR_UNSUPPORTED;
    if (p->width == 0 || p->height == 0 || p->spp == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty render");
    if (p->flags & ST_FLAG_TRACE_FROM_LIGHT) return render_light(s, view, p, o);
    if (p->flags & ST_FLAG_BDPT_LVC) return render_vc(s, view, p, o);
    const int rc = check_camera(s, p);
    if (rc != ST_OK) return rc;
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        const uint32_t npix = p->width * p->height;
        const bool reuse = (p->flags & ST_FLAG_RESTIR_DI_REUSE) != 0;
        const bool hashgrid = (p->flags & ST_FLAG_RESTIR_DI_HASHGRID) != 0;
        // hash-grid reuse chains the samples: one per wave, one after another
        const uint32_t k = hashgrid ? 1 : wave_samples(p, p->samples_per_wave);
        if (k == 0) return set_error(ST_ERROR_LIMIT, "too many paths per wave");
        const uint64_t paths = (uint64_t)npix * k;
        const uint32_t waves = (p->spp + k - 1) / k;
        const int slots = (p->flags & ST_FLAG_SERIAL_WAVES) || hashgrid ? 1 : (int)std::min<uint32_t>(kWaveSlots, waves);
        const int tgrid = ctx->cus * kTraceBlocksPerCU, sgrid = ctx->cus * 8;
        order_after_caller(ctx);
        for (int sl = 0; sl < slots; sl++) reserve_wave(ctx->slot[sl], paths, reuse, paths, tgrid);
        // the previous call's last grid is reused when it was built for this
        // scene at this entry count (TestRenderer keeps mPrevHashGrid across frames)
        const uint64_t ns = hashgrid ? (uint64_t)npix * std::min<uint32_t>(2, p->max_depth - 1) : 0;
        const bool hg_prev_valid = hashgrid && ctx->hg_valid && !(p->flags & ST_FLAG_HISTORY_RESET) && ctx->hg_scene == s->serial &&
                                   ctx->hg_ns == ns;
        ctx->hg_valid = false;
        if (hashgrid) {
            for (int i = 0; i < 2; i++) {
                ctx->hg_rec[i].reserve(ns * 64);
                ctx->hg_keys[i].reserve(ns * 8);
                ctx->hg_slots[i].reserve(ns * 4);
            }
            ctx->hg_keys_tmp.reserve(ns * 8);
            ctx->hg_slots_tmp.reserve(ns * 4);
            size_t sb = 0;
            HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                                         (uint32_t*)nullptr, (uint32_t*)nullptr, (int)ns, 0, 33));
            ctx->hg_sort_tmp.reserve(std::max<size_t>(sb, 16));
        }
        ctx->acc.reserve((size_t)npix * 16);
        if (reuse) ctx->rprev.reserve((size_t)npix * 32);
        ctx->stats.reserve(64);
        begin_timings(ctx);
        CameraJob j;
        std::memset(&j, 0, sizeof j);
        j.s = s;
        j.view = *view;
        j.p = *p;
        j.o = *o;
        for (int sl = 0; sl < slots; sl++) {
            j.slot[sl] = &ctx->slot[sl];
            j.counters[sl] = ctx->slot[sl].counters.as<uint32_t>();
        }
        j.slots = slots;
        j.k = k;
        j.acc = ctx->acc.as<float4>();
        j.rprev = ctx->rprev.as<float4>();
        j.stats = ctx->stats.as<unsigned long long>();
        j.tgrid = tgrid;
        j.sgrid = sgrid;
        j.timed = true;
        j.hashgrid = hashgrid;
        j.hg_first = ctx->hg_last ^ 1;
        j.hg_prev_valid = hg_prev_valid;
        const bool async = (p->flags & ST_FLAG_ASYNC) != 0;
        j.timed = !async;
        enqueue_camera(ctx, j);
        if (hashgrid) {
            ctx->hg_last = (j.hg_first + (int)waves - 1) & 1;
            ctx->hg_scene = s->serial;
            ctx->hg_ns = ns;
            ctx->hg_valid = true;
        }
        if (async) {  // join the wave slots into the context stream and return
            SlotCounters sc{};
            for (int sl = 0; sl < slots; sl++) sc.c[sl] = j.counters[sl];
            for (int sl = 1; sl < slots; sl++) {
                HIP_CHECK(hipEventRecord(ctx->slot[sl].done, ctx->slot[sl].stream));
                HIP_CHECK(hipStreamWaitEvent(ctx->stream, ctx->slot[sl].done, 0));
            }
            if (o->ray_counts)
                hipLaunchKernelGGL(k_sum_rays, dim3(1), dim3(64), 0, ctx->stream, sc, slots, (unsigned long long*)o->ray_counts);
            HIP_CHECK(hipGetLastError());
            return ST_OK;
        }
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipStreamSynchronize(ctx->slot[sl].stream));
        finish_timings(ctx, slots, o->ray_counts);
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

// ------------------------------------------------------------ differentiable rendering
int st_scene_set_triangle_albedo(st_scene* s, const float* rgb, uint32_t count) {
    if (!s) return set_error(ST_ERROR_INVALID_ARGUMENT, "null scene");
    if (rgb && !s->sph_inst.empty()) return set_error(ST_ERROR_UNSUPPORTED, "a per-triangle albedo in a scene with sphere instances");
    try {
        HIP_CHECK(hipSetDevice(s->ctx->device));
        if (!rgb) {
            if (s->d.tri_alb) s->generation++;  // captured graphs hold the scene's constants
            s->d.tri_alb = nullptr;
            return ST_OK;
        }
        if ((uint64_t)count != s->ntri)
            return set_error(ST_ERROR_INVALID_ARGUMENT, "the albedo table holds 3 floats per flattened triangle");
        order_after_caller(s->ctx);
        s->tri_alb.reserve(12ull * count + 16);
        HIP_CHECK(hipMemcpyAsync(s->tri_alb.p, rgb, 12ull * count, hipMemcpyDefault, s->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        if (s->d.tri_alb != s->tri_alb.as<float>()) s->generation++;
        s->d.tri_alb = s->tri_alb.as<float>();
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_scene_set_volume_grid(st_scene* s, uint32_t volume, const void* grid, uint64_t bytes) {
    if (!s || !grid) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (volume >= s->vol_bytes.size()) return set_error(ST_ERROR_INVALID_ARGUMENT, "no such volume");
    if (bytes != s->vol_bytes[volume]) return set_error(ST_ERROR_INVALID_ARGUMENT, "the grid must keep its size");
    try {
        HIP_CHECK(hipSetDevice(s->ctx->device));
        order_after_caller(s->ctx);
        std::vector<uint8_t> host(bytes);
        HIP_CHECK(hipMemcpyAsync(host.data(), grid, bytes, hipMemcpyDefault, s->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        // the same tree: the root (and with it every node offset) where it was
        if (validate_nanovdb(host.data(), bytes) != s->d.vol_root[volume])
            return set_error(ST_ERROR_INVALID_ARGUMENT, "the grid must keep its tree layout");
        HIP_CHECK(hipMemcpy(const_cast<uint8_t*>(s->d.vol[volume]), host.data(), bytes, hipMemcpyHostToDevice));
        return ST_OK;
    } catch (const std::invalid_argument& e) {
        return set_error(ST_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

// A wave's fp32 radiance per path, kept for the replay
__global__ void k_adj_total(uint32_t n, const float4* em, float4* ltot) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) ltot[p] = em[p];
}

// Path-replay adjoint.  Every wave is rendered twice from the same seeds with
// fp32 NEE: the first pass keeps each path's radiance, the second replays the
// path and, at each differentiable event, adds dLoss/dL times the radiance
// still to come times the event's log-derivative (k_shade<.., ADJ>,
// k_trace_media<ADJ>, k_shadow_media<ADJ>).  Waves are independent (the
// gradients are sums), so they run in the context's four slots with no order
// between them (one slot with ST_FLAG_SERIAL_WAVES).  Kernel classes are
// timed as in st_render (st_last_kernel_times; st_last_timings' ray counts
// cover both passes).
int st_adjoint(st_scene* s, const st_view* view, const st_render_params* p, const float* d_hdr, float* grad_albedo,
               float* grad_grid, uint32_t grad_volume) {
    if (!s || !view || !p || !d_hdr) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (p->width == 0 || p->height == 0 || p->spp == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty render");
    if (p->flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE))
        return set_error(ST_ERROR_UNSUPPORTED, "the adjoint replays camera paths with plain NEE");
    if (s->d.has_tex) return set_error(ST_ERROR_UNSUPPORTED, "the adjoint in a scene with material images");
    if (grad_albedo && (!s->d.tri_alb || !(p->flags & ST_FLAG_LAMBERTIAN)))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "albedo gradients need a per-triangle albedo and ST_FLAG_LAMBERTIAN");
    if (grad_grid && grad_volume >= s->d.vol_count) return set_error(ST_ERROR_INVALID_ARGUMENT, "no such volume");
    const int rc = check_camera(s, p);
    if (rc != ST_OK) return rc;
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        const uint32_t npix = p->width * p->height;
        const uint32_t k = wave_samples(p, p->samples_per_wave);
        if (k == 0) return set_error(ST_ERROR_LIMIT, "too many paths per wave");
        const uint64_t paths = (uint64_t)npix * k;
        const uint32_t waves = (p->spp + k - 1) / k;
        const int slots = (p->flags & ST_FLAG_SERIAL_WAVES) ? 1 : (int)std::min<uint32_t>(kWaveSlots, waves);
        const int tgrid = ctx->cus * kTraceBlocksPerCU, sgrid = ctx->cus * 8;
        const bool media = s->d.vinst_count > 0, lambert = (p->flags & (ST_FLAG_LAMBERTIAN | ST_FLAG_LAMBERT_GGX)) != 0;
        const bool alias = (p->flags & ST_FLAG_ENV_ALIAS) != 0;
        order_after_caller(ctx);
        for (int sl = 0; sl < slots; sl++) {
            reserve_wave(ctx->slot[sl], paths, false, paths, tgrid);
            ctx->slot[sl].ltot.reserve(paths * 16);
        }
        ctx->stats.reserve(64);
        begin_timings(ctx);
        hipStream_t s0 = ctx->slot[0].stream;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s0, ctx->stats.as<uint32_t>(), 16u, 0u);
        for (int sl = 0; sl < slots; sl++)
            hipLaunchKernelGGL(k_fill_u32, dim3(kCounterWords / 256), dim3(256), 0, s0, ctx->slot[sl].counters.as<uint32_t>(),
                               (uint32_t)kCounterWords, 0u);
        if (slots > 1) {
            HIP_CHECK(hipEventRecord(ctx->start_done, s0));
            for (int sl = 1; sl < slots; sl++) HIP_CHECK(hipStreamWaitEvent(ctx->slot[sl].stream, ctx->start_done, 0));
        }
        RenderConsts c = camera_consts(s, *view, p);
        c.flags |= ST_FLAG_FLOAT_NEE;
        unsigned long long* stats = ctx->stats.as<unsigned long long>();
        const AuxOut aux{nullptr, nullptr, nullptr, nullptr};
        auto* shade = shade_kernel(lambert, alias, false, media);
        auto* shade_adj = shade_kernel_adj(lambert, alias, media);
        uint32_t wi = 0;
        for (uint32_t base = 0; base < p->spp; base += k, wi++) {
            WaveBufs& B = ctx->slot[wi % slots];
            hipStream_t sm = B.stream;
            WaveState w = wave_state(B, ctx);
            uint32_t* spill = B.spill.as<uint32_t>();
            c.k = std::min(k, p->spp - base);
            c.sample_base = base;
            c.wave_paths = npix * c.k;
            const AdjArgs adj{B.ltot.as<float4>(), reinterpret_cast<const float4*>(d_hdr), grad_albedo, grad_grid, grad_volume,
                              p->spp};
            const int cgrid = std::min<int>((c.wave_paths + kShadeBlock - 1) / kShadeBlock, ctx->cus * 16);
            for (int pass = 0; pass < 2; pass++) {
                const bool replay = pass == 1;
                hipLaunchKernelGGL(k_camera, dim3(cgrid), dim3(kShadeBlock), 0, sm, c, w);
                int qsel = 0;
                for (uint32_t b = 0; b < std::max<uint32_t>(1, p->max_depth); b++) {
                    {
                        Timed tm(ctx, 0, sm);
                        if (media)
                            hipLaunchKernelGGL(replay ? k_trace_media<true> : k_trace_media<false>, dim3(tgrid), dim3(kTraceBlock),
                                               0, sm, s->d, c, w, qsel, spill, stats, adj);
                        else
                            hipLaunchKernelGGL(closest_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill,
                                               stats);
                    }
                    {
                        Timed tm(ctx, 1, sm);
                        hipLaunchKernelGGL(replay ? shade_adj : shade, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, c, w, qsel, aux,
                                           0, adj);
                    }
                    {
                        Timed tm(ctx, 2, sm);
                        if (media)
                            hipLaunchKernelGGL(replay ? k_shadow_media<true> : k_shadow_media<false>, dim3(tgrid),
                                               dim3(kTraceBlock), 0, sm, s->d, c, w, spill, stats, adj);
                        else
                            hipLaunchKernelGGL(shadow_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, c.flags,
                                               spill, stats);
                    }
                    hipLaunchKernelGGL(k_count, dim3(1), dim3(1), 0, sm, w, qsel, media ? 1 : 0);
                    qsel ^= 1;
                }
                if (!replay) {
                    const int g2 = std::min<int>((c.wave_paths + 255) / 256, ctx->cus * 8);
                    hipLaunchKernelGGL(k_adj_total, dim3(g2), dim3(256), 0, sm, c.wave_paths, w.em, B.ltot.as<float4>());
                }
            }
            HIP_CHECK(hipGetLastError());
        }
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipStreamSynchronize(ctx->slot[sl].stream));
        finish_timings(ctx, slots, nullptr);
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

// ------------------------------------------------------------ radiance cache
struct st_cache {
    st_context* ctx = nullptr;
    st_cache_config cfg{};
    CacheGrid grid{};
    uint32_t table_floats = 0;  // levels x entries x 2
    DevBuf table32, table16, tm, tv, tgrad;
    DevBuf w32, w16, w16t, wm, wv, wgrad;  // 4 layers of 64 x 64
    DevBuf table_ema, w_ema, table16e, w16e;  // the parameters' moving average (config.ema_decay > 0)
    bool ema() const { return cfg.ema_decay > 0; }
    // the fp16 parameters inference reads: the moving average's when kept
    const _Float16* infer_table() const { return (ema() ? table16e : table16).as<_Float16>(); }
    const _Float16* infer_weights() const { return (ema() ? w16e : w16).as<_Float16>(); }
    DevBuf X0, H1, H2, H3, Y, dA, dB;      // rows x 64 fp16
    DevBuf wpart;                          // weight-gradient partials, one 64 x 64 block per wavefront
    DevBuf loss;
    DevBuf trq, trthr, trt;  // st_render_cached's training rows: queries, throughputs, targets
    DevBuf flag, off, scan_tmp, cq, ct, count;  // and the rows with weight, compacted
    uint32_t rows = 0;
    uint32_t step = 0;
    void reserve_rows(uint32_t n) {
        if (n <= rows) return;
        const size_t b = (size_t)n * kMlpWidth * 2;
        for (DevBuf* d : {&X0, &H1, &H2, &H3, &Y, &dA, &dB}) d->reserve(b);
        rows = n;
    }
};

namespace {
constexpr uint32_t kMlpParams = kMlpLayers * kMlpWidth * kMlpWidth;
constexpr uint32_t kCacheMaxRows = 1u << 24;

int grid_for(int n, int per, int cap) { return std::max(1, std::min((n + per - 1) / per, cap)); }

// X0 -> H1 -> H2 -> H3 -> Y on the context stream over n rows (at most n
// when n_dev, a device count, is given), with the training parameters or
// (infer) the ones inference reads
void cache_forward(st_cache* c, const float* queries, uint32_t n, const uint32_t* n_dev = nullptr, bool infer = false) {
    hipStream_t st = c->ctx->stream;
    const int cus = c->ctx->cus;
    const _Float16* w = infer ? c->infer_weights() : c->w16.as<_Float16>();
    hipLaunchKernelGGL(k_rc_encode, dim3(grid_for((int)n, 256, cus * 16)), dim3(256), 0, st, c->grid,
                       infer ? c->infer_table() : c->table16.as<_Float16>(), queries, n, n_dev, c->X0.as<_Float16>());
    const int lg = grid_for((int)n, 256, cus * 8);  // 4 wavefronts of 64 rows per block
    _Float16* in[4] = {c->X0.as<_Float16>(), c->H1.as<_Float16>(), c->H2.as<_Float16>(), c->H3.as<_Float16>()};
    _Float16* out[4] = {c->H1.as<_Float16>(), c->H2.as<_Float16>(), c->H3.as<_Float16>(), c->Y.as<_Float16>()};
    for (int l = 0; l < kMlpLayers; l++)
        hipLaunchKernelGGL(k_mlp_layer, dim3(lg), dim3(256), 0, st, w + (size_t)l * kMlpWidth * kMlpWidth, in[l], out[l],
                           (const _Float16*)nullptr, n, n_dev, l < kMlpLayers - 1 ? 1 : 0);
}

void cache_copy16(st_cache* c) {
    hipStream_t st = c->ctx->stream;
    hipLaunchKernelGGL(k_rc_copy16, dim3(grid_for((int)c->table_floats, 256, c->ctx->cus * 16)), dim3(256), 0, st,
                       c->table32.as<float>(), c->table_floats, c->table16.as<_Float16>(), (_Float16*)nullptr);
    hipLaunchKernelGGL(k_rc_copy16, dim3(grid_for((int)kMlpParams, 256, 64)), dim3(256), 0, st, c->w32.as<float>(), kMlpParams,
                       c->w16.as<_Float16>(), c->w16t.as<_Float16>());
    if (c->ema()) {  // the moving average restarts at the parameters
        HIP_CHECK(hipMemcpyAsync(c->table_ema.p, c->table32.p, 4ull * c->table_floats, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpyAsync(c->w_ema.p, c->w32.p, 4ull * kMlpParams, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpyAsync(c->table16e.p, c->table16.p, 2ull * c->table_floats, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpyAsync(c->w16e.p, c->w16.p, 2ull * kMlpParams, hipMemcpyDeviceToDevice, st));
    }
}
}  // namespace

int st_cache_create(st_context* ctx, const st_cache_config* cfg, const float bbox_min[3], const float bbox_max[3], st_cache** out) {
    if (!ctx || !cfg || !bbox_min || !bbox_max || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (cfg->levels < 1 || cfg->levels > kCacheMaxLevels || cfg->log2_table < 8 || cfg->log2_table > 24 ||
        cfg->base_resolution < 2 || cfg->max_resolution < cfg->base_resolution || cfg->max_resolution > (1u << 20))
        return set_error(ST_ERROR_INVALID_ARGUMENT,
                         "cache: 1-16 levels, 2^8 to 2^24 entries per level, 2 <= base resolution <= max resolution <= 2^20");
    if (!(cfg->learning_rate > 0) || !(cfg->beta1 >= 0 && cfg->beta1 < 1) || !(cfg->beta2 >= 0 && cfg->beta2 < 1) ||
        !(cfg->epsilon > 0) || !(cfg->ema_decay >= 0 && cfg->ema_decay < 1))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "cache: learning rate > 0, betas and ema_decay in [0, 1), epsilon > 0");
    for (int a = 0; a < 3; a++)
        if (!(bbox_max[a] > bbox_min[a])) return set_error(ST_ERROR_INVALID_ARGUMENT, "cache: empty bounding box");
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        auto c = std::make_unique<st_cache>();
        c->ctx = ctx;
        c->cfg = *cfg;
        CacheGrid& g = c->grid;
        for (int a = 0; a < 3; a++) {
            g.lo[a] = bbox_min[a];
            g.inv_ext[a] = 1.0f / (bbox_max[a] - bbox_min[a]);
        }
        g.levels = cfg->levels;
        g.table_mask = (1u << cfg->log2_table) - 1;
        // geometric progression of resolutions (Mueller et al., "Instant
        // Neural Graphics Primitives", 2022, eq. 2-3)
        const double b = cfg->levels > 1 ? std::exp((std::log((double)cfg->max_resolution) - std::log((double)cfg->base_resolution)) /
                                                    (double)(cfg->levels - 1))
                                         : 1.0;
        for (uint32_t l = 0; l < kCacheMaxLevels; l++) {
            g.res[l] = l < cfg->levels ? (float)std::floor((double)cfg->base_resolution * std::pow(b, (double)l)) : 0.0f;
            // levels whose (res + 2)^3 corner points fit the table index them directly
            const double n = (double)g.res[l] + 2;
            if (l < cfg->levels && n * n * n <= (double)(1u << cfg->log2_table)) g.dense |= 1u << l;
        }
        c->table_floats = cfg->levels * (1u << cfg->log2_table) * 2;
        // initial parameters: table entries uniform in +-1e-4, weights uniform
        // in +-sqrt(6 / (fan in + fan out)); the output matrix's rows 3..63 zero
        std::vector<float> t(c->table_floats), w(kMlpParams, 0.0f);
        uint64_t state = 0x9E3779B97F4A7C15ull ^ cfg->seed;
        auto uni = [&]() {
            state = state * 6364136223846793005ull + 1442695040888963407ull;
            return (float)((state >> 40) * (1.0 / 16777216.0));
        };
        for (float& x : t) x = (uni() * 2 - 1) * 1e-4f;
        for (int l = 0; l < kMlpLayers; l++) {
            const int fo = l == kMlpLayers - 1 ? 3 : kMlpWidth;
            const float a = std::sqrt(6.0f / (float)(kMlpWidth + fo));
            for (int o = 0; o < fo; o++)
                for (int i = 0; i < kMlpWidth; i++) w[(size_t)l * kMlpWidth * kMlpWidth + o * kMlpWidth + i] = (uni() * 2 - 1) * a;
        }
        const size_t tb = 4ull * c->table_floats;
        c->table32.reserve(tb);
        c->tm.reserve(tb);
        c->tv.reserve(tb);
        c->tgrad.reserve(tb);
        c->table16.reserve(tb / 2);
        c->w32.reserve(4ull * kMlpParams);
        c->wm.reserve(4ull * kMlpParams);
        c->wv.reserve(4ull * kMlpParams);
        c->wgrad.reserve(4ull * kMlpParams);
        c->w16.reserve(2ull * kMlpParams);
        c->w16t.reserve(2ull * kMlpParams);
        if (c->ema()) {
            c->table_ema.reserve(tb);
            c->table16e.reserve(tb / 2);
            c->w_ema.reserve(4ull * kMlpParams);
            c->w16e.reserve(2ull * kMlpParams);
        }
        c->loss.reserve(64);
        HIP_CHECK(hipMemcpy(c->table32.p, t.data(), tb, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(c->w32.p, w.data(), 4ull * kMlpParams, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemset(c->tm.p, 0, tb));
        HIP_CHECK(hipMemset(c->tv.p, 0, tb));
        HIP_CHECK(hipMemset(c->tgrad.p, 0, tb));
        HIP_CHECK(hipMemset(c->wm.p, 0, 4ull * kMlpParams));
        HIP_CHECK(hipMemset(c->wv.p, 0, 4ull * kMlpParams));
        HIP_CHECK(hipMemset(c->wgrad.p, 0, 4ull * kMlpParams));
        cache_copy16(c.get());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        *out = c.release();
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_cache_destroy(st_cache* c) {
    if (!c) return ST_OK;
    (void)hipSetDevice(c->ctx->device);
    (void)hipStreamSynchronize(c->ctx->stream);
    delete c;
    return ST_OK;
}

int st_cache_infer(st_cache* c, const float* queries, uint32_t n, float* radiance) {
    if (!c || (n && (!queries || !radiance))) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (n > kCacheMaxRows) return set_error(ST_ERROR_LIMIT, "at most 2^24 queries per call");
    if (n == 0) return ST_OK;
    try {
        HIP_CHECK(hipSetDevice(c->ctx->device));
        order_after_caller(c->ctx);
        c->reserve_rows(n);
        cache_forward(c, queries, n, nullptr, true);
        hipLaunchKernelGGL(k_rc_output, dim3(grid_for((int)n, 256, c->ctx->cus * 16)), dim3(256), 0, c->ctx->stream, c->Y.as<_Float16>(),
                           n, radiance);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

namespace {
// One Adam step on n samples already on the device (at most n when n_dev, a
// device count, is given), on the context stream; targets: tstride (3, or 4
// with a weight) floats per sample.  The gradient is scaled by 1 / n
// (Adam's step does not depend on the scale).
void cache_train_step(st_cache* c, const float* queries, const float* targets, uint32_t tstride, uint32_t n,
                      const uint32_t* n_dev = nullptr) {
    st_context* ctx = c->ctx;
    hipStream_t st = ctx->stream;
    const int cus = ctx->cus;
    c->reserve_rows(n);
    cache_forward(c, queries, n, n_dev);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int)c->table_floats, 256, cus * 16)), dim3(256), 0, st, c->tgrad.as<uint32_t>(),
                       c->table_floats, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((int)kMlpParams, 256, 64)), dim3(256), 0, st, c->wgrad.as<uint32_t>(), kMlpParams, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, c->loss.as<uint32_t>(), 16u, 0u);
    hipLaunchKernelGGL(k_rc_loss, dim3(grid_for((int)n, 256, cus * 8)), dim3(256), 0, st, c->Y.as<_Float16>(), targets, tstride,
                       n, n_dev, c->dA.as<_Float16>(), c->loss.as<float>());
    const _Float16* wt = c->w16t.as<_Float16>();
    float* wg = c->wgrad.as<float>();
    // rows per wavefront of the weight-gradient product: 128, more when
    // there would be over 2048 wavefronts (the partials' buffer)
    const uint32_t chunk = std::max<uint32_t>(128, ((n + 2047) / 2048 + 15) / 16 * 16);
    const uint32_t nw = (n + chunk - 1) / chunk;
    c->wpart.reserve((size_t)nw * kMlpWidth * kMlpWidth * 4);
    const int wgrid = grid_for((int)nw, 4, 1 << 20);  // 4 wavefronts of `chunk` rows per block
    const int lg = grid_for((int)n, 256, cus * 8);
    _Float16* acts[4] = {c->X0.as<_Float16>(), c->H1.as<_Float16>(), c->H2.as<_Float16>(), c->H3.as<_Float16>()};
    _Float16* d_out = c->dA.as<_Float16>();  // gradient at layer l's output
    _Float16* d_in = c->dB.as<_Float16>();
    for (int l = kMlpLayers - 1; l >= 0; l--) {
        const size_t off = (size_t)l * kMlpWidth * kMlpWidth;
        hipLaunchKernelGGL(k_mlp_wgrad, dim3(wgrid), dim3(256), 0, st, d_out, acts[l], c->wpart.as<float>(), n, n_dev, chunk);
        hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(kMlpWidth * kMlpWidth / 256, std::min<uint32_t>(nw, 32)), dim3(256), 0, st,
                           c->wpart.as<float>(), nw, wg + off, 1.0f);
        // through the layer (transposed weights) and the ReLU that made its input
        hipLaunchKernelGGL(k_mlp_layer, dim3(lg), dim3(256), 0, st, wt + off, d_out, d_in, l > 0 ? acts[l] : (const _Float16*)nullptr,
                           n, n_dev, 0);
        std::swap(d_out, d_in);
    }
    // d_out now holds dLoss/dX0
    hipLaunchKernelGGL(k_rc_encode_bwd, dim3(grid_for((int)n, 256, cus * 16)), dim3(256), 0, st, c->grid, queries, n, n_dev, d_out,
                       c->tgrad.as<float>());
    c->step++;
    const float b1 = c->cfg.beta1, b2 = c->cfg.beta2;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)c->step), c2 = 1.0f - (float)std::pow((double)b2, (double)c->step);
    const float gs = 1.0f / (float)n;
    hipLaunchKernelGGL(k_rc_adam, dim3(grid_for((int)kMlpParams, 256, 64)), dim3(256), 0, st, c->w32.as<float>(), c->wm.as<float>(),
                       c->wv.as<float>(), (const float*)wg, kMlpParams, gs, c->cfg.learning_rate, b1, b2, c->cfg.epsilon, c1, c2,
                       c->w16.as<_Float16>(), c->w16t.as<_Float16>(), 0, c->ema() ? c->w_ema.as<float>() : nullptr,
                       c->w16e.as<_Float16>(), c->cfg.ema_decay);
    hipLaunchKernelGGL(k_rc_adam, dim3(grid_for((int)c->table_floats, 256, cus * 16)), dim3(256), 0, st, c->table32.as<float>(),
                       c->tm.as<float>(), c->tv.as<float>(), (const float*)c->tgrad.as<float>(), c->table_floats, gs,
                       c->cfg.learning_rate, b1, b2, c->cfg.epsilon, c1, c2, c->table16.as<_Float16>(), (_Float16*)nullptr, 1,
                       c->ema() ? c->table_ema.as<float>() : nullptr, c->table16e.as<_Float16>(), c->cfg.ema_decay);
    HIP_CHECK(hipGetLastError());
}
}  // namespace

int st_cache_train(st_cache* c, const float* queries, const float* targets, uint32_t n, float* loss) {
    if (!c || (n && (!queries || !targets))) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (n > kCacheMaxRows) return set_error(ST_ERROR_LIMIT, "at most 2^24 samples per step");
    if (n == 0) return ST_OK;
    try {
        HIP_CHECK(hipSetDevice(c->ctx->device));
        order_after_caller(c->ctx);
        cache_train_step(c, queries, targets, 3, n);
        float l[2] = {0, 0};
        HIP_CHECK(hipMemcpyAsync(l, c->loss.p, 8, hipMemcpyDeviceToHost, c->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
        if (loss) *loss = l[0] / (float)n;
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_cache_param_counts(const st_cache* c, uint32_t* table_floats, uint32_t* weight_floats) {
    if (!c) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (table_floats) *table_floats = c->table_floats;
    if (weight_floats) *weight_floats = kMlpParams;
    return ST_OK;
}

int st_cache_get_params(st_cache* c, float* table, float* weights, float* table_grad, float* weight_grad) {
    if (!c) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    try {
        HIP_CHECK(hipSetDevice(c->ctx->device));
        HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
        if (table) HIP_CHECK(hipMemcpy(table, c->table32.p, 4ull * c->table_floats, hipMemcpyDeviceToHost));
        if (weights) HIP_CHECK(hipMemcpy(weights, c->w32.p, 4ull * kMlpParams, hipMemcpyDeviceToHost));
        if (table_grad) HIP_CHECK(hipMemcpy(table_grad, c->tgrad.p, 4ull * c->table_floats, hipMemcpyDeviceToHost));
        if (weight_grad) HIP_CHECK(hipMemcpy(weight_grad, c->wgrad.p, 4ull * kMlpParams, hipMemcpyDeviceToHost));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_cache_set_params(st_cache* c, const float* table, const float* weights) {
    if (!c) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    try {
        HIP_CHECK(hipSetDevice(c->ctx->device));
        HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
        if (table) HIP_CHECK(hipMemcpy(c->table32.p, table, 4ull * c->table_floats, hipMemcpyHostToDevice));
        if (weights) HIP_CHECK(hipMemcpy(c->w32.p, weights, 4ull * kMlpParams, hipMemcpyHostToDevice));
        cache_copy16(c);
        HIP_CHECK(hipStreamSynchronize(c->ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_render_cached(st_scene* s, st_cache* cache, const st_view* view, const st_render_params* p,
                     const st_cache_render_params* cp, const st_outputs* o) {
    if (!s || !cache || !view || !p || !cp || !o || !o->hdr) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (p->width == 0 || p->height == 0 || p->spp == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty render");
    if (cache->ctx != s->ctx) return set_error(ST_ERROR_INVALID_ARGUMENT, "the cache belongs to another context");
    if (p->flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC))
        return set_error(ST_ERROR_UNSUPPORTED, "the radiance cache ends camera paths of the path tracer");
    if (s->d.has_tex) return set_error(ST_ERROR_UNSUPPORTED, "the radiance cache in a scene with material images");
    if ((p->flags & ST_FLAG_RESTIR_DI) && !(p->flags & ST_FLAG_RESTIR_DI_REUSE))
        return set_error(ST_ERROR_UNSUPPORTED, "with the radiance cache, ReSTIR DI runs at primary vertices (ST_FLAG_RESTIR_DI_REUSE)");
    if (p->flags & ST_FLAG_RESTIR_DI_DEEP)
        return set_error(ST_ERROR_UNSUPPORTED, "with the radiance cache, deeper vertices take one-sample NEE (no ST_FLAG_RESTIR_DI_DEEP)");
    if (cp->cache_depth < 1 || cp->cache_depth >= p->max_depth)
        return set_error(ST_ERROR_INVALID_ARGUMENT, "cache_depth must be in [1, max_depth)");
    if ((p->flags & ST_FLAG_RESTIR_DI_REUSE) && cp->cache_depth < 2)
        return set_error(ST_ERROR_INVALID_ARGUMENT, "with screen-space reuse the primary vertex keeps its NEE: cache_depth >= 2");
    if (cp->train_stride < 1) return set_error(ST_ERROR_INVALID_ARGUMENT, "train_stride must be at least 1");
    if (cp->train_steps > 1024) return set_error(ST_ERROR_INVALID_ARGUMENT, "at most 1024 training steps per call");
    const int rc = check_camera(s, p);
    if (rc != ST_OK) return rc;
    const uint64_t npix = (uint64_t)p->width * p->height, total = npix * p->spp;
    if (total >= (1ull << 32)) return set_error(ST_ERROR_LIMIT, "at most 2^32 - 1 paths per call");
    const uint64_t rows = (total + cp->train_stride - 1) / cp->train_stride;
    if (rows > kCacheMaxRows) return set_error(ST_ERROR_LIMIT, "at most 2^24 training rows per call (raise train_stride)");
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        const bool reuse = (p->flags & ST_FLAG_RESTIR_DI_REUSE) != 0;
        const uint32_t k = wave_samples(p, p->samples_per_wave);
        if (k == 0) return set_error(ST_ERROR_LIMIT, "too many paths per wave");
        const uint64_t paths = npix * k;
        const uint32_t waves = (p->spp + k - 1) / k;
        const int slots = (p->flags & ST_FLAG_SERIAL_WAVES) ? 1 : (int)std::min<uint32_t>(kWaveSlots, waves);
        const int tgrid = ctx->cus * kTraceBlocksPerCU, sgrid = ctx->cus * 8;
        order_after_caller(ctx);
        for (int sl = 0; sl < slots; sl++) {
            reserve_wave(ctx->slot[sl], paths, reuse, paths, tgrid);
            ctx->slot[sl].cq.reserve(paths * 64);
        }
        ctx->acc.reserve((size_t)npix * 16);
        if (reuse) ctx->rprev.reserve((size_t)npix * 32);
        ctx->stats.reserve(64);
        cache->trq.reserve(rows * 48);
        cache->trthr.reserve(rows * 16);
        cache->trt.reserve(rows * 16);
        cache->reserve_rows((uint32_t)rows);
        begin_timings(ctx);
        hipStream_t s0 = ctx->stream;
        // rows start empty: zero queries (finite network inputs) and weights
        const int fg = ctx->cus * 8;
        hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, s0, cache->trq.as<uint32_t>(), (uint32_t)(rows * 12), 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(fg), dim3(256), 0, s0, cache->trt.as<uint32_t>(), (uint32_t)(rows * 4), 0u);
        CacheJob cj;
        cj.grid = cache->grid;
        cj.table = cache->infer_table();
        cj.w16 = cache->infer_weights();
        cj.tq = cache->trq.as<float>();
        cj.tthr = cache->trthr.as<float4>();
        cj.tt = cache->trt.as<float4>();
        cj.rows = (uint32_t)rows;
        cj.depth = cp->cache_depth;
        cj.stride = cp->train_stride;
        cj.frame = cp->frame;
        CameraJob j;
        std::memset(&j, 0, sizeof j);
        j.s = s;
        j.view = *view;
        j.p = *p;
        j.o = *o;
        for (int sl = 0; sl < slots; sl++) {
            j.slot[sl] = &ctx->slot[sl];
            j.counters[sl] = ctx->slot[sl].counters.as<uint32_t>();
        }
        j.slots = slots;
        j.k = k;
        j.acc = ctx->acc.as<float4>();
        j.rprev = ctx->rprev.as<float4>();
        j.stats = ctx->stats.as<unsigned long long>();
        j.tgrid = tgrid;
        j.sgrid = sgrid;
        j.timed = true;
        j.cache = &cj;
        enqueue_camera(ctx, j);
        // every wave's targets are in before the training steps
        for (int sl = 1; sl < slots; sl++) HIP_CHECK(hipStreamWaitEvent(s0, ctx->slot[sl].resolve_done, 0));
        if (cp->train_queries)
            HIP_CHECK(hipMemcpyAsync(cp->train_queries, cache->trq.p, rows * 48, hipMemcpyDeviceToDevice, s0));
        if (cp->train_targets)
            HIP_CHECK(hipMemcpyAsync(cp->train_targets, cache->trt.p, rows * 16, hipMemcpyDeviceToDevice, s0));
        float l[2] = {0, 0};
        if (cp->train_steps > 0) {
            // the rows with weight to the front, in row order: training covers them only
            Timed tm(ctx, 5, s0);
            const int cg = std::min<int>((int)((rows + 255) / 256), ctx->cus * 8);
            cache->flag.reserve(rows * 4);
            cache->off.reserve(rows * 4);
            cache->cq.reserve(rows * 48);
            cache->ct.reserve(rows * 12);
            cache->count.reserve(64);
            size_t sb = 0;
            HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)rows, s0));
            cache->scan_tmp.reserve(std::max<size_t>(sb, 16));
            sb = cache->scan_tmp.n;
            hipLaunchKernelGGL(k_rc_row_flags, dim3(cg), dim3(256), 0, s0, cache->trt.as<float4>(), (uint32_t)rows,
                               cache->flag.as<uint32_t>());
            HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cache->scan_tmp.p, sb, cache->flag.as<uint32_t>(), cache->off.as<uint32_t>(),
                                                       (int)rows, s0));
            hipLaunchKernelGGL(k_rc_row_compact, dim3(cg), dim3(256), 0, s0, cache->trq.as<float>(), cache->trt.as<float4>(),
                               cache->flag.as<uint32_t>(), cache->off.as<uint32_t>(), (uint32_t)rows, cache->cq.as<float>(),
                               cache->ct.as<float>(), cache->count.as<uint32_t>());
        }
        for (uint32_t st = 0; st < cp->train_steps; st++) {
            Timed tm(ctx, 6, s0);
            cache_train_step(cache, cache->cq.as<float>(), cache->ct.as<float>(), 3, (uint32_t)rows, cache->count.as<uint32_t>());
            if (st == 0 && cp->loss) HIP_CHECK(hipMemcpyAsync(l, cache->loss.p, 8, hipMemcpyDeviceToHost, s0));
        }
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipStreamSynchronize(ctx->slot[sl].stream));
        if (cp->loss) *cp->loss = l[1] > 0 ? l[0] / l[1] : 0.0f;
        finish_timings(ctx, slots, o->ray_counts);
        ctx->cache_rows = cp->loss ? (double)l[1] : 0.0;
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_cache_last_times(st_context* ctx, double out[6]) {
    if (!ctx || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    out[0] = ctx->kernel_ms[5];
    out[1] = (double)ctx->ev_used[5];
    out[2] = ctx->kernel_ms[6];
    out[3] = (double)ctx->ev_used[6];
    out[4] = ctx->cache_queries;
    out[5] = ctx->cache_rows;
    return ST_OK;
}

// ------------------------------------------------------------ batches
// A batch renders many requests (scene, view, params, outputs), each exactly
// as one st_render call would, in one graph launch.  Requests are spread over
// the context's streams by longest-processing-time-first; each runs its waves
// one after another in its stream's wave buffers, with accumulators, counters
// and reuse history of its own, and the streams run side by side.  The whole
// sequence of launches is captured once into a hipGraph and replayed on every
// st_batch_render: the host issues one launch per batch instead of a few
// thousand.  st_batch_autotune times candidate configurations (trace and
// shade grid sizes, paths per wave, streams) and keeps the fastest.
struct st_batch {
    st_context* ctx = nullptr;
    std::vector<st_render_request> req;
    st_batch_config cfg{};
    struct Per {
        DevBuf counters, acc, rprev, stats;
        uint32_t k = 1;
        int stream = 0;
        double cost = 0;
    };
    std::vector<std::unique_ptr<Per>> per;
    WaveBufs bufs[kWaveSlots];
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> generation;  // of each request's scene when the graph was captured
    hipEvent_t fork = nullptr, join[kWaveSlots] = {};
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~st_batch() {
        drop_graph();
        if (fork) (void)hipEventDestroy(fork);
        for (int i = 0; i < kWaveSlots; i++)
            if (join[i]) (void)hipEventDestroy(join[i]);
    }
};

// Defaults: st_render's grids, four streams, and waves four times st_render's
// (a request's waves run one after another on its stream, so fewer, larger
// waves leave fewer half-empty tail bounces; on a 32-request mix of 128^2 to
// 1024^2 px this measured 460 ms per batch against 670 ms at 2^21 paths).
static st_batch_config default_batch_config() {
    st_batch_config c{};
    c.trace_blocks_per_cu = kTraceBlocksPerCU;
    c.shade_blocks_per_cu = 8;
    c.wave_paths = 4 * kWavePaths;
    c.streams = kWaveSlots;
    c.ms = 0;
    return c;
}

// Buffers and stream assignment for a configuration (no launches).
static void batch_plan(st_batch* b, const st_batch_config& cfg) {
    st_context* ctx = b->ctx;
    b->drop_graph();
    b->cfg = cfg;
    const int streams = (int)cfg.streams;
    const int tgrid = ctx->cus * (int)cfg.trace_blocks_per_cu;
    std::vector<size_t> order(b->req.size());
    for (size_t r = 0; r < order.size(); r++) {
        const st_render_params& p = b->req[r].params;
        st_batch::Per& q = *b->per[r];
        const uint32_t npix = p.width * p.height;
        uint32_t k = p.samples_per_wave;
        if (k == 0) k = std::max<uint32_t>(1, std::min<uint32_t>(p.spp, cfg.wave_paths / npix));
        q.k = std::min(k, p.spp);
        if ((uint64_t)npix * q.k >= (1ull << 31)) throw std::invalid_argument("too many paths per wave");
        // estimated cost: path vertices, with the candidate loop of RIS counted twice
        q.cost = (double)npix * p.spp * std::max<uint32_t>(1, p.max_depth) * ((p.flags & ST_FLAG_RESTIR_DI) ? 2.0 : 1.0);
        order[r] = r;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return b->per[a]->cost > b->per[c]->cost; });
    std::vector<double> load(streams, 0.0);
    std::vector<uint64_t> paths(streams, 0);
    std::vector<bool> reuse(streams, false);
    for (size_t r : order) {
        const int i = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        st_batch::Per& q = *b->per[r];
        const st_render_params& p = b->req[r].params;
        q.stream = i;
        load[i] += q.cost;
        paths[i] = std::max<uint64_t>(paths[i], (uint64_t)p.width * p.height * q.k);
        reuse[i] = reuse[i] || (p.flags & ST_FLAG_RESTIR_DI_REUSE);
        const size_t npix = (size_t)p.width * p.height;
        q.counters.reserve(4 * kCounterWords);
        q.acc.reserve(npix * 16);
        if (p.flags & ST_FLAG_RESTIR_DI_REUSE) q.rprev.reserve(npix * 32);
        q.stats.reserve(64);
    }
    for (int i = 0; i < streams; i++) {
        b->bufs[i].stream = ctx->slot[i].stream;
        if (paths[i]) reserve_wave(b->bufs[i], paths[i], reuse[i], paths[i], tgrid);
    }
}

// The batch's launches: forked from the first stream, every request on its
// stream, joined back (enqueued inside the capture).
static void batch_enqueue(st_batch* b) {
    st_context* ctx = b->ctx;
    const int streams = (int)b->cfg.streams;
    hipStream_t s0 = ctx->slot[0].stream;
    HIP_CHECK(hipEventRecord(b->fork, s0));
    for (int i = 1; i < streams; i++) HIP_CHECK(hipStreamWaitEvent(ctx->slot[i].stream, b->fork, 0));
    for (int i = 0; i < streams; i++)
        for (size_t r = 0; r < b->req.size(); r++) {
            st_batch::Per& q = *b->per[r];
            if (q.stream != i) continue;
            const st_render_request& R = b->req[r];
            CameraJob j;
            std::memset(&j, 0, sizeof j);
            j.s = R.scene;
            j.view = R.view;
            j.p = R.params;
            j.o = R.outputs;
            j.slot[0] = &b->bufs[i];
            j.counters[0] = q.counters.as<uint32_t>();
            j.slots = 1;
            j.k = q.k;
            j.acc = q.acc.as<float4>();
            j.rprev = q.rprev.as<float4>();
            j.stats = q.stats.as<unsigned long long>();
            j.tgrid = ctx->cus * (int)b->cfg.trace_blocks_per_cu;
            j.sgrid = ctx->cus * (int)b->cfg.shade_blocks_per_cu;
            j.timed = false;
            j.store_rays = true;
            enqueue_camera(ctx, j);
        }
    for (int i = 1; i < streams; i++) {
        HIP_CHECK(hipEventRecord(b->join[i], ctx->slot[i].stream));
        HIP_CHECK(hipStreamWaitEvent(s0, b->join[i], 0));
    }
}

// Capture the batch's launches into a graph and instantiate it.
static void batch_capture(st_batch* b) {
    st_context* ctx = b->ctx;
    b->drop_graph();
    hipStream_t s0 = ctx->slot[0].stream;
    HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
    try {
        batch_enqueue(b);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s0, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(s0, &b->graph));
    HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

static void batch_launch(st_batch* b) {
    st_context* ctx = b->ctx;
    // a scene rebuilt since the capture has new buffers: capture again
    for (size_t r = 0; r < b->req.size() && b->exec; r++)
        if (b->generation[r] != b->req[r].scene->generation) b->drop_graph();
    if (!b->exec) {
        batch_capture(b);
        for (size_t r = 0; r < b->req.size(); r++) b->generation[r] = b->req[r].scene->generation;
    }
    order_after_caller(ctx);
    HIP_CHECK(hipGraphLaunch(b->exec, ctx->slot[0].stream));
    HIP_CHECK(hipStreamSynchronize(ctx->slot[0].stream));
}

int st_batch_create(st_context* ctx, const st_render_request* requests, uint32_t count, st_batch** out) {
    if (!ctx || !out || (count && !requests)) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (count == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty batch");
    if (count > 4096) return set_error(ST_ERROR_LIMIT, "at most 4096 requests per batch");
    for (uint32_t r = 0; r < count; r++) {
        const st_render_request& R = requests[r];
        if (!R.scene || !R.outputs.hdr) return set_error(ST_ERROR_INVALID_ARGUMENT, "request without scene or hdr output");
        if (R.scene->ctx != ctx) return set_error(ST_ERROR_INVALID_ARGUMENT, "request scene belongs to another context");
        const st_render_params& p = R.params;
        if (p.width == 0 || p.height == 0 || p.spp == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty render in a request");
        if (p.flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC))
            return set_error(ST_ERROR_UNSUPPORTED, "batches run camera paths (no light tracing or bidirectional mode)");
        if (p.flags & ST_FLAG_RESTIR_DI_HASHGRID)
            return set_error(ST_ERROR_UNSUPPORTED, "hash-grid reuse chains a request's samples through the context's grids: not in batches");
        if (check_images(R.scene, p.flags) != ST_OK) return ST_ERROR_UNSUPPORTED;
        const int rc = check_camera(R.scene, &p);
        if (rc != ST_OK) return rc;
    }
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        auto b = std::make_unique<st_batch>();
        b->ctx = ctx;
        b->req.assign(requests, requests + count);
        b->generation.assign(count, 0);
        for (uint32_t r = 0; r < count; r++) b->per.push_back(std::make_unique<st_batch::Per>());
        HIP_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        for (int i = 0; i < kWaveSlots; i++) HIP_CHECK(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
        batch_plan(b.get(), default_batch_config());
        *out = b.release();
        return ST_OK;
    } catch (const std::invalid_argument& e) {
        return set_error(ST_ERROR_LIMIT, e.what());
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_batch_render(st_batch* b) {
    if (!b) return set_error(ST_ERROR_INVALID_ARGUMENT, "null batch");
    try {
        HIP_CHECK(hipSetDevice(b->ctx->device));
        batch_launch(b);
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_batch_get_config(const st_batch* b, st_batch_config* out) {
    if (!b || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    *out = b->cfg;
    return ST_OK;
}

int st_batch_set_config(st_batch* b, const st_batch_config* cfg) {
    if (!b || !cfg) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (cfg->trace_blocks_per_cu < 1 || cfg->trace_blocks_per_cu > 16 || cfg->shade_blocks_per_cu < 1 ||
        cfg->shade_blocks_per_cu > 32 || cfg->streams < 1 || cfg->streams > kWaveSlots || cfg->wave_paths < 1 ||
        cfg->wave_paths > (1u << 26))
        return set_error(ST_ERROR_INVALID_ARGUMENT,
                         "batch config: 1-16 trace and 1-32 shade blocks per CU, 1-4 streams, 1 to 2^26 paths per wave");
    try {
        HIP_CHECK(hipSetDevice(b->ctx->device));
        batch_plan(b, *cfg);
        return ST_OK;
    } catch (const std::invalid_argument& e) {
        return set_error(ST_ERROR_LIMIT, e.what());
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

// The in-library autotuner (timed coordinate descent over the four
// configuration fields) is switched off: its one run aborted the process on
// the host (glibc "free(): invalid pointer") with no cause found in this
// library's code.  Configurations are set with st_batch_set_config.
int st_batch_autotune(st_batch* b, uint32_t repeats, st_batch_config* chosen) {
    (void)repeats;
    (void)chosen;
    if (!b) return set_error(ST_ERROR_INVALID_ARGUMENT, "null batch");
    return set_error(ST_ERROR_UNSUPPORTED, "st_batch_autotune is switched off; set the configuration with st_batch_set_config");
}

int st_batch_destroy(st_batch* b) {
    if (!b) return ST_OK;
    (void)hipSetDevice(b->ctx->device);
    for (int i = 0; i < kWaveSlots; i++) (void)hipStreamSynchronize(b->ctx->slot[i].stream);
    delete b;
    return ST_OK;
}

int st_render_batch(st_context* ctx, const st_render_request* requests, uint32_t count) {
    st_batch* b = nullptr;
    int rc = st_batch_create(ctx, requests, count, &b);
    if (rc != ST_OK) return rc;
    rc = st_batch_render(b);
    st_batch_destroy(b);
    return rc;
}

static int render_light(st_scene* s, const st_view* view, const st_render_params* p, const st_outputs* o) {
    if ((uint64_t)p->width * p->height > (1ull << 28)) return set_error(ST_ERROR_LIMIT, "image too large");
    if (check_triangle_albedo(s, p->flags) != ST_OK) return ST_ERROR_UNSUPPORTED;
    if (p->max_null_collisions > (1u << 20)) return set_error(ST_ERROR_INVALID_ARGUMENT, "max_null_collisions above 2^20");
    if (!(p->light_subpath_factor >= 0) || !(p->light_subpath_factor <= 64))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "light_subpath_factor must be in [0, 64]");
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        order_after_caller(ctx);
        const uint32_t npix = p->width * p->height;
        // mLightSubpathCount = max(1, uint(W * H * factor)), TestRenderer.cpp:452;
        // the product is bounded first, so the conversion stays defined
        const float Lf = (float)npix * p->light_subpath_factor;
        if (!(Lf < 2147483648.0f)) return set_error(ST_ERROR_LIMIT, "too many light subpaths");
        const uint32_t L = std::max<uint32_t>(1u, (uint32_t)Lf);
        const uint32_t span = std::max(L, npix);
        uint32_t k = p->samples_per_wave;
        if (k == 0) k = std::max<uint32_t>(1, std::min<uint32_t>(p->spp, kWavePaths / span));
        k = std::min(k, p->spp);
        const uint64_t paths = (uint64_t)span * k;
        if (paths >= (1ull << 31)) return set_error(ST_ERROR_LIMIT, "too many paths per wave");
        const uint32_t waves = (p->spp + k - 1) / k;
        const int slots = (p->flags & ST_FLAG_SERIAL_WAVES) ? 1 : (int)std::min<uint32_t>(kWaveSlots, waves);
        const int tgrid = ctx->cus * kTraceBlocksPerCU, sgrid = ctx->cus * 8;
        for (int sl = 0; sl < slots; sl++) reserve_wave(ctx->slot[sl], paths, false, 0, tgrid);
        begin_timings(ctx);
        ctx->acc.reserve((size_t)npix * 16);
        ctx->stats.reserve(64);
        hipStream_t s0 = ctx->stream;
        HIP_CHECK(hipMemsetAsync(ctx->acc.p, 0, (size_t)npix * 16, s0));
        HIP_CHECK(hipMemsetAsync(ctx->stats.p, 0, 64, s0));
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipMemsetAsync(ctx->slot[sl].counters.p, 0, 4 * kCounterWords, s0));
        HIP_CHECK(hipEventRecord(ctx->start_done, s0));
        for (int sl = 1; sl < slots; sl++) HIP_CHECK(hipStreamWaitEvent(ctx->slot[sl].stream, ctx->start_done, 0));
        RenderConsts c;
        std::memset(&c, 0, sizeof c);
        c.W = p->width;
        c.H = p->height;
        c.npix = npix;
        c.seed = p->seed;
        c.max_depth = p->max_depth;
        c.min_len = p->min_path_length;
        // the reference turns NEE, ReSTIR DI and the direct-only mode off when
        // light tracing (TestRenderer.cpp:119-124)
        c.flags = (p->flags & ~(ST_FLAG_SAMPLE_DIRECT | ST_FLAG_DIRECT_ONLY | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE |
                                ST_FLAG_FLOAT_NEE));
        c.rr_start = p->rr_start_depth;
        c.env_prob = p->environment_sample_probability;
        c.view = *view;
        c.inv_cam = inverse_transform(view->camera_to_world);
        c.view_medium = kNoMedium;
        c.max_null = p->max_null_collisions;
        c.light_paths = L;
        // ViewData::imagePlaneDist, compat/scene.h:147
        c.ipd = (float)std::abs(view->image_max[1] - view->image_min[1]) / (2 * std::tan(view->projection.vertical_fov / 2));
        c.tiled = (p->width % 8 == 0 && p->height % 8 == 0) ? 1 : 0;
        const bool lambert = (p->flags & (ST_FLAG_LAMBERTIAN | ST_FLAG_LAMBERT_GGX)) != 0, alias = (p->flags & ST_FLAG_ENV_ALIAS) != 0;
        const bool count = (p->flags & ST_FLAG_COUNT_TRAVERSAL) != 0;
        const bool media = s->d.vinst_count > 0;
        auto* stats = ctx->stats.as<unsigned long long>();
        // waves rotate through the slots and their streams; resolves keep sample order
        uint32_t wi = 0;
        for (uint32_t base = 0; base < p->spp; base += k, wi++) {
            WaveBufs& B = ctx->slot[wi % slots];
            WaveBufs& prev = ctx->slot[(wi + slots - 1) % slots];
            hipStream_t sm = B.stream;
            WaveState w = wave_state(B, ctx);
            uint32_t* spill = B.spill.as<uint32_t>();
            c.k = std::min(k, p->spp - base);
            c.sample_base = base;
            c.wave_paths = L * c.k;
            HIP_CHECK(hipMemsetAsync(B.counters.p, 0, 16, sm));  // queue and shadow counts
            const int igrid = std::min<int>((std::max(c.wave_paths, c.k * npix) + kShadeBlock - 1) / kShadeBlock, ctx->cus * 16);
            if (alias) hipLaunchKernelGGL(k_light_init<true>, dim3(igrid), dim3(kShadeBlock), 0, sm, s->d, c, w);
            else hipLaunchKernelGGL(k_light_init<false>, dim3(igrid), dim3(kShadeBlock), 0, sm, s->d, c, w);
            int qsel = 0;
            for (uint32_t b = 0; b < std::max<uint32_t>(1, p->max_depth); b++) {
                {
                    Timed tm(ctx, 0, sm);
                    if (media) hipLaunchKernelGGL(k_trace_media<false>, dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, c, w, qsel, spill, stats, AdjArgs{});
                    else if (count) hipLaunchKernelGGL(closest_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
                    else hipLaunchKernelGGL(closest_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
                }
                {
                    Timed tm(ctx, 1, sm);
                    auto kfn = media ? (lambert ? k_shade_light<true, true, false> : k_shade_light<false, true, false>)
                                     : (lambert ? k_shade_light<true, false, false> : k_shade_light<false, false, false>);
                    hipLaunchKernelGGL(kfn, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, c, w, qsel);
                }
                {
                    Timed tm(ctx, 2, sm);
                    if (media) hipLaunchKernelGGL(k_shadow_media<false>, dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, c, w, spill, stats, AdjArgs{});
                    else if (count) hipLaunchKernelGGL(shadow_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, c.flags, spill, stats);
                    else hipLaunchKernelGGL(shadow_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, c.flags, spill, stats);
                }
                hipLaunchKernelGGL(k_count, dim3(1), dim3(1), 0, sm, w, qsel, media ? 1 : 0);
                qsel ^= 1;
            }
            const bool last = base + c.k >= p->spp;
            const int rgrid = std::min<int>((npix + 255) / 256, ctx->cus * 8);
            if (wi > 0 && slots > 1) HIP_CHECK(hipStreamWaitEvent(sm, prev.resolve_done, 0));
            hipLaunchKernelGGL(k_resolve, dim3(rgrid), dim3(256), 0, sm, c, w, ctx->acc.as<float4>(), o->hdr,
                               (uint32_t*)nullptr, 1.0f / (float)p->spp, last ? 1 : 0);
            HIP_CHECK(hipEventRecord(B.resolve_done, sm));
            HIP_CHECK(hipGetLastError());
        }
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipStreamSynchronize(ctx->slot[sl].stream));
        finish_timings(ctx, slots, o->ray_counts);
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

// TestRenderer with gUseVC (TestRenderer.cpp:478-482, 559-629), one sample per
// wave: the light pass (light subpaths store their vertices and connect to the
// camera, into the pixels' accumulators), the cache compaction, then the camera
// pass (emission, NEE and a cached-vertex connection at every vertex), then the
// resolve.  Both passes of a sample run on its slot's stream.
static int render_vc(st_scene* s, const st_view* view, const st_render_params* p, const st_outputs* o) {
    if ((uint64_t)p->width * p->height > (1ull << 28)) return set_error(ST_ERROR_LIMIT, "image too large");
    if (check_triangle_albedo(s, p->flags) != ST_OK) return ST_ERROR_UNSUPPORTED;
    if (s->d.vinst_count > 0 || (p->flags & (ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE | ST_FLAG_FLOAT_NEE | ST_FLAG_DIRECT_ONLY)) ||
        p->min_path_length)
        return set_error(ST_ERROR_UNSUPPORTED, "the bidirectional mode supports surfaces with plain NEE only");
    // its MIS weights fold the continuation probability in at every vertex
    if (p->rr_start_depth > 1)
        return set_error(ST_ERROR_UNSUPPORTED, "the bidirectional mode runs Russian roulette at every vertex");
    if (!(p->light_subpath_factor >= 0) || !(p->light_subpath_factor <= 64))
        return set_error(ST_ERROR_INVALID_ARGUMENT, "light_subpath_factor must be in [0, 64]");
    if (p->max_depth < 1 || p->max_depth > 256) return set_error(ST_ERROR_INVALID_ARGUMENT, "max_depth must be in [1, 256]");
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        order_after_caller(ctx);
        const uint32_t npix = p->width * p->height;
        const float Lf = (float)npix * p->light_subpath_factor;  // bounded before the conversion
        if (!(Lf < 2147483648.0f)) return set_error(ST_ERROR_LIMIT, "too many light subpaths");
        const uint32_t L = std::max<uint32_t>(1u, (uint32_t)Lf);
        const uint32_t stride = std::max<uint32_t>(1u, p->max_depth - 1);
        const uint64_t paths = std::max<uint64_t>(L, npix);
        if (paths >= (1ull << 31) || (uint64_t)L * stride >= (1ull << 31))
            return set_error(ST_ERROR_LIMIT, "too many paths or light vertices per sample");
        const int slots = (p->flags & ST_FLAG_SERIAL_WAVES) ? 1 : (int)std::min<uint32_t>(kWaveSlots, p->spp);
        const int tgrid = ctx->cus * kTraceBlocksPerCU, sgrid = ctx->cus * 8;
        size_t scan_bytes = 0;
        HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)L));
        for (int sl = 0; sl < slots; sl++) {
            WaveBufs& b = ctx->slot[sl];
            reserve_wave(b, paths, false, 0, tgrid);
            b.shadow.reserve(paths * 96);  // NEE and a light-vertex connection per vertex
            b.vcm.reserve(paths * 8);
            b.lv_raw.reserve((size_t)L * stride * 48);
            b.lv.reserve((size_t)L * stride * 48);
            b.lv_cnt.reserve((size_t)L * 4);
            b.lv_off.reserve((size_t)L * 4);
            b.lv_total.reserve(16);
            b.scan_tmp.reserve(std::max<size_t>(scan_bytes, 16));
        }
        ctx->acc.reserve((size_t)npix * 16);
        ctx->stats.reserve(64);
        hipStream_t s0 = ctx->stream;
        HIP_CHECK(hipMemsetAsync(ctx->acc.p, 0, (size_t)npix * 16, s0));
        HIP_CHECK(hipMemsetAsync(ctx->stats.p, 0, 64, s0));
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipMemsetAsync(ctx->slot[sl].counters.p, 0, 4 * kCounterWords, s0));
        HIP_CHECK(hipEventRecord(ctx->start_done, s0));
        for (int sl = 1; sl < slots; sl++) HIP_CHECK(hipStreamWaitEvent(ctx->slot[sl].stream, ctx->start_done, 0));
        begin_timings(ctx);
        RenderConsts c;
        std::memset(&c, 0, sizeof c);
        c.W = p->width;
        c.H = p->height;
        c.npix = npix;
        c.seed = p->seed;
        c.max_depth = p->max_depth;
        c.flags = p->flags & ~(uint32_t)ST_FLAG_SAMPLE_DIRECT;  // NEE runs as part of the bidirectional mode
        c.rr_start = p->rr_start_depth;
        c.env_prob = p->environment_sample_probability;
        c.view = *view;
        c.inv_cam = inverse_transform(view->camera_to_world);
        c.view_medium = kNoMedium;
        c.max_null = p->max_null_collisions;
        c.light_paths = L;
        c.lv_stride = stride;
        c.ipd = (float)std::abs(view->image_max[1] - view->image_min[1]) / (2 * std::tan(view->projection.vertical_fov / 2));
        c.tiled = (p->width % 8 == 0 && p->height % 8 == 0) ? 1 : 0;
        c.k = 1;
        const bool lambert = (p->flags & (ST_FLAG_LAMBERTIAN | ST_FLAG_LAMBERT_GGX)) != 0, alias = (p->flags & ST_FLAG_ENV_ALIAS) != 0;
        const bool count = (p->flags & ST_FLAG_COUNT_TRAVERSAL) != 0;
        auto* stats = ctx->stats.as<unsigned long long>();
        const uint32_t depth = std::max<uint32_t>(1, p->max_depth);
        auto bounces = [&](hipStream_t sm, const RenderConsts& cc, WaveState& w, uint32_t* spill, bool light,
                           const AuxOut& aux, int first) {
            int qsel = 0;
            for (uint32_t b = 0; b < depth; b++) {
                {
                    Timed tm(ctx, 0, sm);
                    if (count) hipLaunchKernelGGL(closest_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
                    else hipLaunchKernelGGL(closest_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, qsel, spill, stats);
                }
                {
                    Timed tm(ctx, 1, sm);
                    if (light) {
                        auto kfn = lambert ? k_shade_light<true, false, true> : k_shade_light<false, false, true>;
                        hipLaunchKernelGGL(kfn, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, cc, w, qsel);
                    } else {
                        auto kfn = lambert ? (alias ? k_shade_vc<true, true> : k_shade_vc<true, false>)
                                           : (alias ? k_shade_vc<false, true> : k_shade_vc<false, false>);
                        hipLaunchKernelGGL(kfn, dim3(sgrid), dim3(kShadeBlock), 0, sm, s->d, cc, w, qsel, aux, first);
                    }
                }
                {
                    Timed tm(ctx, 2, sm);
                    if (count) hipLaunchKernelGGL(shadow_kernel(true, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, cc.flags, spill, stats);
                    else hipLaunchKernelGGL(shadow_kernel(false, s->d), dim3(tgrid), dim3(kTraceBlock), 0, sm, s->d, w, cc.flags, spill, stats);
                }
                hipLaunchKernelGGL(k_count, dim3(1), dim3(1), 0, sm, w, qsel, 0);
                qsel ^= 1;
            }
        };
        uint32_t wi = 0;
        for (uint32_t base = 0; base < p->spp; base++, wi++) {
            WaveBufs& B = ctx->slot[wi % slots];
            WaveBufs& prev = ctx->slot[(wi + slots - 1) % slots];
            hipStream_t sm = B.stream;
            WaveState w = wave_state(B, ctx);
            uint32_t* spill = B.spill.as<uint32_t>();
            c.sample_base = base;
            // light pass: light paths keyed by their dispatch index (path_rng)
            RenderConsts cl = c;
            cl.flags |= ST_FLAG_TRACE_FROM_LIGHT;
            cl.wave_paths = L;
            HIP_CHECK(hipMemsetAsync(B.counters.p, 0, 16, sm));
            const int igrid = std::min<int>((uint32_t)((paths + kShadeBlock - 1) / kShadeBlock), ctx->cus * 16);
            if (alias) hipLaunchKernelGGL(k_light_init<true>, dim3(igrid), dim3(kShadeBlock), 0, sm, s->d, cl, w);
            else hipLaunchKernelGGL(k_light_init<false>, dim3(igrid), dim3(kShadeBlock), 0, sm, s->d, cl, w);
            bounces(sm, cl, w, spill, true, AuxOut{nullptr, nullptr, nullptr, nullptr}, 0);
            // the cache, in (subpath, depth) order
            size_t sb = B.scan_tmp.n;
            HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(B.scan_tmp.p, sb, w.lv_cnt, w.lv_off, (int)L, sm));
            const int lgrid = std::min<int>((L + kShadeBlock - 1) / kShadeBlock, ctx->cus * 16);
            hipLaunchKernelGGL(k_lvc_compact, dim3(lgrid), dim3(kShadeBlock), 0, sm, c, w, L);
            // camera pass
            RenderConsts cc = c;
            cc.wave_paths = npix;
            const int cgrid = std::min<int>((npix + kShadeBlock - 1) / kShadeBlock, ctx->cus * 16);
            hipLaunchKernelGGL(k_camera, dim3(cgrid), dim3(kShadeBlock), 0, sm, cc, w);
            AuxOut aux{nullptr, nullptr, nullptr, nullptr};
            if (base == 0) aux = AuxOut{o->visibility, o->albedo, o->depth, o->prev_uv};
            bounces(sm, cc, w, spill, false, aux, base == 0 ? 1 : 0);
            const bool last = base + 1 >= p->spp;
            const int rgrid = std::min<int>((npix + 255) / 256, ctx->cus * 8);
            if (wi > 0 && slots > 1) HIP_CHECK(hipStreamWaitEvent(sm, prev.resolve_done, 0));
            hipLaunchKernelGGL(k_resolve, dim3(rgrid), dim3(256), 0, sm, cc, w, ctx->acc.as<float4>(), o->hdr,
                               o->path_length, 1.0f / (float)p->spp, last ? 1 : 0);
            HIP_CHECK(hipEventRecord(B.resolve_done, sm));
            HIP_CHECK(hipGetLastError());
        }
        for (int sl = 0; sl < slots; sl++) HIP_CHECK(hipStreamSynchronize(ctx->slot[sl].stream));
        finish_timings(ctx, slots, o->ray_counts);
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_last_timings(st_context* ctx, double out[12]) {
    if (!ctx || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 12; i++) out[i] = ctx->timings[i];
    return ST_OK;
}

int st_last_kernel_times(st_context* ctx, double out[10]) {
    if (!ctx || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < st_context::kRenderKinds; i++) {
        out[i] = ctx->kernel_ms[i];
        out[st_context::kRenderKinds + i] = (double)ctx->ev_used[i];
    }
    return ST_OK;
}

int st_trace_rays(st_scene* s, const float* rays, uint32_t count, int closest, uint32_t* hits) {
    if (!s || (count && (!rays || !hits))) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    st_context* ctx = s->ctx;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        if (count == 0) return ST_OK;
        order_after_caller(ctx);
        const int grid = std::min<int>((count + kTraceBlock - 1) / kTraceBlock, ctx->cus * kTraceBlocksPerCU);
        WaveBufs& B = ctx->slot[0];
        ensure_spill(B, ctx->cus * kTraceBlocksPerCU, kTraceBlock);
        B.counters.reserve(4 * kCounterWords);
        HIP_CHECK(hipMemsetAsync(B.counters.as<uint32_t>() + kFetchRays, 0, 4 * kXcds * kFetchStride, ctx->stream));
        hipLaunchKernelGGL(s->d.has_alpha ? (s->d.sph_node_count ? k_trace_rays<true, true> : k_trace_rays<false, true>)
                                             : (s->d.sph_node_count ? k_trace_rays<true> : k_trace_rays<false>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream, s->d, (const float4*)rays, count,
                           closest, hits, B.spill.as<uint32_t>(), B.counters.as<uint32_t>() + kFetchRays);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_trace_rays_noalloc(st_scene* s, const float* rays, uint32_t count, int closest, uint32_t* hits) {
    if (!s || (count && (!rays || !hits))) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (count > 0x7fffffffu) return set_error(ST_ERROR_LIMIT, "at most 2^31 - 1 rays per call");
    st_context* ctx = s->ctx;
    // Per lane, nearest-first traversal holds at most 3 stack entries per tree
    // level; past the LDS part they go to the scratch's spill region, laid out
    // [entry][lane] over the grid's lanes, so the grid is sized to fit it.
    const uint64_t depth = s->stats[2];
    const uint64_t entries = 3 * depth + 1 > (uint64_t)kLdsStack ? 3 * depth + 1 - kLdsStack : 1;
    const uint64_t fit_blocks = kScratchSpill / (entries * 4 * kTraceBlock);
    if (fit_blocks < 1)
        return set_error(ST_ERROR_LIMIT, "BVH too deep for the fixed traversal scratch of st_trace_rays_noalloc");
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        if (count == 0) return ST_OK;
        order_after_caller(ctx);
        const uint64_t want = ((uint64_t)count + kTraceBlock - 1) / kTraceBlock;
        const int grid = (int)std::min<uint64_t>({want, fit_blocks, (uint64_t)ctx->cus * kTraceBlocksPerCU});
        uint32_t* fetch = reinterpret_cast<uint32_t*>(ctx->scratch);
        uint32_t* spill = reinterpret_cast<uint32_t*>(ctx->scratch + kScratchHeader);
        // the fetch counter starts at zero (a fill kernel: also right inside a captured graph)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, fetch, kXcds * kFetchStride, 0u);
        hipLaunchKernelGGL(s->d.has_alpha ? (s->d.sph_node_count ? k_trace_rays<true, true> : k_trace_rays<false, true>)
                                             : (s->d.sph_node_count ? k_trace_rays<true> : k_trace_rays<false>), dim3(grid), dim3(kTraceBlock), 0, ctx->stream, s->d, (const float4*)rays, count,
                           closest, hits, spill, fetch);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_trace_primary(st_scene* s, const st_view* view, uint32_t width, uint32_t height, uint32_t* visibility, float* t_out) {
    if (!s || !view || !visibility) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    st_render_params p;
    std::memset(&p, 0, sizeof p);
    p.width = width;
    p.height = height;
    p.max_depth = 1;
    p.spp = 1;
    p.flags = 0;
    p.environment_sample_probability = 0.9f;
    st_context* ctx = s->ctx;
    try {
        const size_t npix = (size_t)width * height;
        ctx->outbuf.reserve(npix * 32);
        float* hdr = ctx->outbuf.as<float>();
        float* depth = t_out ? hdr + 4 * npix : nullptr;
        st_outputs o{hdr, visibility, nullptr, nullptr, nullptr, depth, nullptr};
        const int rc = st_render(s, view, &p, &o);
        if (rc != ST_OK || !t_out) return rc;
        // t_out[i] = depth[i].x: the first hit's fp16-stored distance (inf on a miss)
        HIP_CHECK(hipMemcpy2DAsync(t_out, 4, depth, 16, 4, npix, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_tonemap(st_context* ctx, const float* hdr, const float* albedo, uint32_t width, uint32_t height, float exposure,
               int mode, int gamma_correct, float* out) {
    if (!ctx || !hdr || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (mode < ST_TONEMAP_RAW || mode > ST_TONEMAP_VIRIDIS_LENGTH_RGB) return set_error(ST_ERROR_INVALID_ARGUMENT, "unknown tonemap mode");
    if ((uint64_t)width * height > (1ull << 30)) return set_error(ST_ERROR_LIMIT, "image too large");
    const uint32_t n = width * height;
    if (n == 0) return ST_OK;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        order_after_caller(ctx);
        ctx->tmax.reserve(16);
        uint32_t* mx = ctx->tmax.as<uint32_t>();
        const int grid = std::min<int>((n + 255) / 256, ctx->cus * 8);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, mx, 4u, 0u);
        hipLaunchKernelGGL(k_tonemap_max, dim3(grid), dim3(256), 0, ctx->stream, reinterpret_cast<const float4*>(hdr),
                           reinterpret_cast<const float4*>(albedo), n, mx);
        hipLaunchKernelGGL(k_tonemap, dim3(grid), dim3(256), 0, ctx->stream, reinterpret_cast<const float4*>(hdr),
                           reinterpret_cast<const float4*>(albedo), n, (const uint32_t*)mx, exposure, mode, gamma_correct,
                           reinterpret_cast<float4*>(out));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

/* ---- SVGF denoiser ------------------------------------------------------ */

}  // extern "C"

// History of one extent: two slots of accumulated colour and moments and of
// the kept visibility and depth (slot `cur` holds the last frame's), and the
// two filter images of the a-trous passes.
// Kernel classes timed per call: accumulation (or restart), variance, filter
// passes, history copy.
constexpr int kSvgfClasses = 4, kSvgfMaxEvents = 20;

struct st_denoiser {
    st_context* ctx = nullptr;
    uint32_t W = 0, H = 0;
    DevBuf color[2], moments[2], vis[2], depth[2], temp[2], geo;
    int cur = 0;
    bool have_history = false, reset = false;
    uint32_t frames = 0;
    hipEvent_t ev[2 * kSvgfMaxEvents] = {};
    int ev_class[kSvgfMaxEvents] = {};
    int nev = 0;
    double timings[2 * kSvgfClasses] = {};
    ~st_denoiser() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    // Brackets one kernel launch on `s` with the next event pair.
    template <class F>
    void timed(int cls, hipStream_t s, F&& launch) {
        const int k = nev++;
        ev_class[k] = cls;
        HIP_CHECK(hipEventRecord(ev[2 * k], s));
        launch();
        HIP_CHECK(hipEventRecord(ev[2 * k + 1], s));
    }
};

extern "C" {

int st_denoiser_create(st_context* ctx, uint32_t width, uint32_t height, st_denoiser** out) {
    if (!ctx || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (width == 0 || height == 0) return set_error(ST_ERROR_INVALID_ARGUMENT, "empty extent");
    if ((uint64_t)width * height > (1ull << 28)) return set_error(ST_ERROR_LIMIT, "image too large");
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        auto d = std::make_unique<st_denoiser>();
        d->ctx = ctx;
        d->W = width;
        d->H = height;
        const size_t n = (size_t)width * height;
        for (int k = 0; k < 2; k++) {
            d->color[k].reserve(n * 16);
            d->moments[k].reserve(n * 8);
            d->vis[k].reserve(n * 8);
            d->depth[k].reserve(n * 4);
            d->temp[k].reserve(n * 16);
        }
        d->geo.reserve(n * 16);
        for (hipEvent_t& e : d->ev) HIP_CHECK(hipEventCreate(&e));
        *out = d.release();
        return ST_OK;
    } catch (const std::bad_alloc&) {
        return set_error(ST_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

int st_denoiser_destroy(st_denoiser* d) {
    if (!d) return ST_OK;
    (void)hipSetDevice(d->ctx->device);
    delete d;
    return ST_OK;
}

int st_denoiser_reset(st_denoiser* d) {
    if (!d) return set_error(ST_ERROR_INVALID_ARGUMENT, "null denoiser");
    d->reset = true;
    d->frames = 0;
    return ST_OK;
}

uint32_t st_denoiser_accumulated_frames(const st_denoiser* d) { return d ? d->frames : 0; }

int st_denoiser_last_timings(const st_denoiser* d, double out[8]) {
    if (!d || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 2 * kSvgfClasses; i++) out[i] = d->timings[i];
    return ST_OK;
}

int st_denoise(st_denoiser* d, const st_view* view, const st_denoise_inputs* in, const st_denoise_params* p,
               float* out) {
    if (!d || !view || !in || !p || !out) return set_error(ST_ERROR_INVALID_ARGUMENT, "null argument");
    if (!in->radiance || !in->prev_uv || !in->visibility || !in->depth)
        return set_error(ST_ERROR_INVALID_ARGUMENT, "radiance, prev_uv, visibility and depth are required");
    if ((p->flags & ST_DENOISE_DEMODULATE_ALBEDO) && !in->albedo)
        return set_error(ST_ERROR_INVALID_ARGUMENT, "ST_DENOISE_DEMODULATE_ALBEDO needs the albedo image");
    if (in->instance_index_map && in->instance_count == 0)
        return set_error(ST_ERROR_INVALID_ARGUMENT, "instance_index_map without instance_count");
    if (view->image_min[0] < 0 || view->image_min[1] < 0 || view->image_max[0] > (int32_t)d->W ||
        view->image_max[1] > (int32_t)d->H || view->image_min[0] >= view->image_max[0] ||
        view->image_min[1] >= view->image_max[1])
        return set_error(ST_ERROR_INVALID_ARGUMENT, "view bounds must lie inside the denoiser's extent");
    if (p->filter_type > ST_FILTER_BOX5_SUBSAMPLED) return set_error(ST_ERROR_INVALID_ARGUMENT, "unknown filter type");
    if (p->atrous_iterations > 15) return set_error(ST_ERROR_LIMIT, "at most 15 filter passes (step 2^14)");
    try {
        st_context* ctx = d->ctx;
        HIP_CHECK(hipSetDevice(ctx->device));
        order_after_caller(ctx);
        hipStream_t st = ctx->stream;
        const SvgfTiles t = svgf_tiles(d->W, d->H);
        const dim3 grid(t.per_xcd * kSvgfXcds), block(kSvgfTileX * kSvgfTileY);
        const SvgfView v{view->image_min[0], view->image_min[1], view->image_max[0], view->image_max[1]};
        const int c = d->cur ^ 1, pv = d->cur;
        d->nev = 0;
        float4* out4 = reinterpret_cast<float4*>(out);
        const float4* radiance = reinterpret_cast<const float4*>(in->radiance);
        const uint2* vis = reinterpret_cast<const uint2*>(in->visibility);
        const float4* depth = reinterpret_cast<const float4*>(in->depth);
        if (!d->reset && d->have_history) {
            const uint32_t iters = p->atrous_iterations;
            SvgfAccumArgs a{};
            a.t = t;
            a.v = v;
            a.flags = p->flags;
            a.history_limit = p->history_limit;
            a.cos_normal = std::cos(2.0f * (180.0f / 3.14159265358979f));  // cos(degrees(2)), a radian argument
            a.shadow_offset = p->shadow_preserve_offset;
            a.shadow_scale = p->shadow_preserve_scale;
            a.map = in->instance_index_map;
            a.map_count = in->instance_count;
            a.vis = vis;
            a.depth = depth;
            a.prev_uv = reinterpret_cast<const float2*>(in->prev_uv);
            a.input = radiance;
            a.albedo = reinterpret_cast<const float4*>(in->albedo);
            a.prev_vis = d->vis[pv].as<uint2>();
            a.prev_depth = d->depth[pv].as<float>();
            a.prev_color = d->color[pv].as<float4>();
            a.prev_moments = d->moments[pv].as<float2>();
            a.color = d->color[c].as<float4>();
            a.moments = d->moments[c].as<float2>();
            a.keep_vis = d->vis[c].as<uint2>();
            a.keep_depth = d->depth[c].as<float>();
            a.geo = iters == 0 ? nullptr : d->geo.as<float4>();
            a.out = iters == 0 ? out4 : nullptr;
            d->timed(0, st, [&] { hipLaunchKernelGGL(k_svgf_accumulate, grid, block, 0, st, a); });
            if (iters > 0) {
                SvgfVarianceArgs va{};
                va.t = t;
                va.v = v;
                va.flags = p->flags;
                va.history_limit = p->history_limit;
                va.boost_length = p->variance_boost_length;
                va.map = in->instance_index_map;
                va.map_count = in->instance_count;
                va.vis = d->vis[c].as<uint2>();
                va.depth = depth;
                va.geo = d->geo.as<float4>();
                va.color = d->color[c].as<float4>();
                va.moments = d->moments[c].as<float2>();
                va.out = d->temp[0].as<float4>();
                d->timed(1, st, [&] { hipLaunchKernelGGL(k_svgf_variance, grid, block, 0, st, va); });
                float4* temp[2] = {d->temp[0].as<float4>(), d->temp[1].as<float4>()};
                for (uint32_t i = 0; i < iters; i++) {
                    // the last pass writes the caller's image instead of filter image (i + 1) % 2
                    float4* dst = i + 1 == iters ? out4 : temp[(i + 1) % 2];
                    SvgfAtrousArgs aa{};
                    aa.t = t;
                    aa.v = v;
                    aa.flags = p->flags;
                    aa.filter_type = p->filter_type;
                    aa.iteration = i;
                    aa.step = 1 << i;
                    aa.sigma_boost = p->sigma_luminance_boost;
                    aa.geo = d->geo.as<float4>();
                    aa.depth = depth;
                    aa.in = temp[i % 2];
                    aa.out = dst;
                    d->timed(2, st, [&] { hipLaunchKernelGGL(k_svgf_atrous, grid, block, 0, st, aa); });
                    if (i + 1 == p->history_tap) {
                        // copy_rgb reads filter image 0: this pass's output when (i + 1) is even
                        const float4* f0 = (i + 1) % 2 == 0 ? dst : temp[0];
                        d->timed(3, st, [&] {
                            hipLaunchKernelGGL(k_svgf_copy_rgb, grid, block, 0, st, t, f0, d->color[c].as<float4>());
                        });
                    }
                }
            }
            d->frames++;
        } else {
            d->timed(0, st, [&] {
                hipLaunchKernelGGL(k_svgf_restart, grid, block, 0, st, t, vis, depth, radiance, d->color[c].as<float4>(),
                                   d->moments[c].as<float2>(), d->vis[c].as<uint2>(), d->depth[c].as<float>(),
                                   out == in->radiance ? nullptr : out4);
            });
            d->reset = false;
            d->frames = 0;
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(st));
        for (double& x : d->timings) x = 0;
        for (int k = 0; k < d->nev; k++) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, d->ev[2 * k], d->ev[2 * k + 1]));
            d->timings[d->ev_class[k]] += ms;
            d->timings[kSvgfClasses + d->ev_class[k]] += 1;
        }
        d->cur = c;
        d->have_history = true;
        return ST_OK;
    } catch (const std::exception& e) {
        return set_error(ST_ERROR_HIP, e.what());
    }
}

}  // extern "C"
