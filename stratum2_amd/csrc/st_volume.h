// st_volume.h — participating media on gfx950: NanoVDB float-grid reads,
// procedural volume boxes, delta tracking (free flight + spectral MIS) and
// ratio tracking, Henyey-Greenstein phase function.
//
// Reference: src/Shaders/materials/medium.hlsli:17-182 (Medium, deltaTrack),
// src/Shaders/common/intersection.hlsli:92-122 (volume AABB candidates),
// :195-260 (traceScatteringRay), :262-338 (traceVisibilityRay), and the
// NanoVDB 32.3 read path of extern/nanovdb/PNanoVDB.h (grid -> tree -> root
// tiles -> upper 32^3 -> lower 16^3 -> leaf 8^3; byte offsets below).
// Grids are checked on the host at scene creation (every child offset stays
// inside its buffer), so these reads need no bounds checks.
#pragma once

#include "st_device.h"
#include "st_layout.h"
#include "st_shade.h"

namespace st {

__device__ __forceinline__ uint32_t g32(const uint8_t* b, uint64_t o) { return *reinterpret_cast<const uint32_t*>(b + o); }
__device__ __forceinline__ uint64_t g64(const uint8_t* b, uint64_t o) {
    return *reinterpret_cast<const unsigned long long*>(b + o);
}
__device__ __forceinline__ float gf(const uint8_t* b, uint64_t o) { return *reinterpret_cast<const float*>(b + o); }

// Byte address of the value at integer index (i, j, k):
// pnanovdb_readaccessor_get_value_address for a float grid (a leaf voxel, a
// tile of an internal node or of the root, or the background).  root = byte
// offset of the root node.
__device__ inline uint64_t nvdb_addr(const uint8_t* b, uint64_t root, int i, int j, int k) {
    const uint32_t ntiles = g32(b, root + 24);
    const uint64_t key = (uint64_t)((uint32_t)k >> 12) | ((uint64_t)((uint32_t)j >> 12) << 21) |
                         ((uint64_t)((uint32_t)i >> 12) << 42);
    for (uint32_t t = 0; t < ntiles; t++) {
        const uint64_t tile = root + 64 + 32 * (uint64_t)t;
        if (g64(b, tile) != key) continue;
        const int64_t child = (int64_t)g64(b, tile + 8);
        if (child == 0) return tile + 20;
        const uint64_t up = root + (uint64_t)child;
        const uint32_t n2 = (uint32_t)(((i & 4095) >> 7) << 10) + (uint32_t)(((j & 4095) >> 7) << 5) + (uint32_t)((k & 4095) >> 7);
        if (!((g32(b, up + 4128 + 4 * (n2 >> 5)) >> (n2 & 31)) & 1)) return up + 8256 + 8 * (uint64_t)n2;
        const uint64_t lo = up + g64(b, up + 8256 + 8 * (uint64_t)n2);
        const uint32_t n1 = (uint32_t)(((i & 127) >> 3) << 8) + (uint32_t)(((j & 127) >> 3) << 4) + (uint32_t)((k & 127) >> 3);
        if (!((g32(b, lo + 544 + 4 * (n1 >> 5)) >> (n1 & 31)) & 1)) return lo + 1088 + 8 * (uint64_t)n1;
        const uint64_t leaf = lo + g64(b, lo + 1088 + 8 * (uint64_t)n1);
        const uint32_t n0 = (uint32_t)((i & 7) << 6) + (uint32_t)((j & 7) << 3) + (uint32_t)(k & 7);
        return leaf + 96 + 4 * (uint64_t)n0;
    }
    return root + 28;  // background
}
// pnanovdb_read_float of that address
__device__ inline float nvdb_read(const uint8_t* b, uint64_t root, int i, int j, int k) {
    return gf(b, nvdb_addr(b, root, i, j, k));
}

// Grid map (GridData::mMap at byte 296: matf 0, invmatf 36, vecf 72),
// pnanovdb_map_apply[_inverse][_jacobi]: index = invmat (world - vec).
__device__ inline v3 nvdb_map(const uint8_t* b, int off, v3 v, bool translate, bool inverse) {
    const uint64_t m = 296 + off;
    if (inverse && translate) v = mk(v.x - gf(b, 296 + 72), v.y - gf(b, 296 + 76), v.z - gf(b, 296 + 80));
    v3 r = mk(v.x * gf(b, m) + v.y * gf(b, m + 4) + v.z * gf(b, m + 8),
              v.x * gf(b, m + 12) + v.y * gf(b, m + 16) + v.z * gf(b, m + 20),
              v.x * gf(b, m + 24) + v.y * gf(b, m + 28) + v.z * gf(b, m + 32));
    if (!inverse && translate) r = mk(r.x + gf(b, 296 + 72), r.y + gf(b, 296 + 76), r.z + gf(b, 296 + 80));
    return r;
}
__device__ __forceinline__ v3 world_to_index(const uint8_t* b, v3 v) { return nvdb_map(b, 36, v, true, true); }
__device__ __forceinline__ v3 world_to_index_dir(const uint8_t* b, v3 v) { return nvdb_map(b, 36, v, false, true); }
__device__ __forceinline__ v3 index_to_world(const uint8_t* b, v3 v) { return nvdb_map(b, 0, v, true, false); }
__device__ __forceinline__ v3 index_to_world_dir(const uint8_t* b, v3 v) { return nvdb_map(b, 0, v, false, false); }

// Medium record (Material.hpp:49-66 store; medium.hlsli:17-40 load)
struct Medium {
    v3 density;
    uint32_t albedo_packed, dens_vol, alb_vol;
    float g;
    __device__ __forceinline__ v3 albedo() const { return mk(un8(albedo_packed, 0), un8(albedo_packed, 8), un8(albedo_packed, 16)); }
    __device__ __forceinline__ bool can_eval() const { return any_pos(density) && albedo_packed > 0; }
    __device__ __forceinline__ bool singular() const { return fabsf(g) > 0.999f; }
    __device__ __forceinline__ float cont_prob() const { return sat(lum(albedo() * density)); }
};
__device__ __forceinline__ Medium load_medium(const DevScene& S, uint32_t addr) {
    Medium m;
    const uint32_t* w = S.mat + addr / 4;
    m.density = mk(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]));
    m.albedo_packed = w[3];
    m.dens_vol = w[4];
    m.alb_vol = w[5];
    m.g = __uint_as_float(w[6]);
    return m;
}

// Henyey-Greenstein evaluateReflectance / sampleDirection (medium.hlsli:44-78);
// din points back along the incoming ray.
// (pow(d, 1.5) is taken as d * sqrt(d): within 1.5 ulp, and the same operations
// in the oracle)
__device__ __forceinline__ float hg_eval(float g, v3 din, v3 dout) {
    const float d = 1 + g * g + 2 * g * dot(din, dout);
    return 1 / (4 * ST_PI) * (1 - g * g) / (d * sqrtf(d));
}
__device__ inline v3 hg_sample(float g, float r0, float r1, v3 din) {
    if (fabsf(g) < 1e-4f) return sample_uniform_sphere(r0, r1);
    const float g2 = g * g;
    const float tmp = (g2 - 1) / (2 * r0 * g - (g + 1));
    const float ce = (tmp * tmp - (1 + g2)) / (2 * g);
    const float se = sqrtf(fmax_(1 - ce * ce, 0));
    const float az = 2 * ST_PI * r1;
    v3 f0, f1;
    make_onb(din, f0, f1);
    return (f0 * (se * cos_(az)) + f1 * (se * sin_(az))) + din * ce;
}

__device__ __forceinline__ v3 expv(v3 a) { return mk(exp_(a.x), exp_(a.y), exp_(a.z)); }
__device__ __forceinline__ float avg3(v3 a) { return (a.x + a.y + a.z) / 3; }
__device__ __forceinline__ float max3(v3 a) { return fmax_(fmax_(a.x, a.y), a.z); }

// Density-gradient sink of the path-replay adjoint: at each collision in
// grid `vol`, grad[value address / 4] += sum_c wgt_c * dlog(beta_c)/d(value),
// with wgt = dLoss/dL times the radiance that the collision's factor
// multiplies.  Tracking decisions and pdfs are not differentiated
// (detached sampling).
struct DensitySink {
    float* grad;
    uint32_t vol;
    v3 wgt;
};

// deltaTrack<SCATTER> (medium.hlsli:104-181): origin/dir in the medium
// instance's object space.  Multiplies beta by the transmittance (times
// sigma_t * sigma_s at a real collision); returns the scattering position in
// object space when *scattered.
// *reads counts the grid reads (collisions) for the traffic model.  With
// ADJ, each collision in the sink's grid adds its gradient term.
template <bool SCATTER, bool ADJ = false>
__device__ inline v3 delta_track(const DevScene& S, const Medium& m, Rng& rng, v3 origin, v3 dir, float tmax, v3& beta,
                                 v3& dir_pdf, v3& nee_pdf, bool& scattered, uint32_t max_null, uint32_t& reads,
                                 const DensitySink* sink = nullptr) {
    scattered = false;
    dir_pdf = mk(1, 1, 1);
    nee_pdf = mk(1, 1, 1);
    const int ch = (int)(rng.next().x % 3u);
    if (m.dens_vol >= S.vol_count) {
        // homogeneous
        const v3 st = m.density;
        const float sc = comp(st, ch);
        if (sc < 1e-6f) return mk(0, 0, 0);
        const float t = -log_(1 - rng.nextf().x) / sc;
        if (t < tmax && SCATTER) {
            scattered = true;
            beta = beta * ((expv((-st) * t) * st) * m.albedo());
            dir_pdf = dir_pdf * (exp_(-sc * t) * sc);
            return origin + dir * t;
        }
        beta = beta * expv((-st) * tmax);
        dir_pdf = dir_pdf * exp_(-sc * tmax);
        return mk(0, 0, 0);
    }
    const uint8_t* g = S.vol[m.dens_vol];
    const uint64_t root = S.vol_root[m.dens_vol];
    const v3 maj = m.density * gf(g, root + 36);
    if (comp(maj, ch) < 1e-6f) return mk(0, 0, 0);
    origin = world_to_index(g, origin);
    dir = world_to_index_dir(g, dir);
    const float invc = 1 / comp(maj, ch);
    const float invmax = 1 / max3(maj);
    for (uint32_t it = 0; it < max_null && any_pos(beta); it++) {
        const float t = -log_(1 - rng.nextf().x) * invc;
        if (t >= tmax) {
            const v3 tr = expv((-maj) * tmax);
            beta = beta * tr;
            nee_pdf = nee_pdf * tr;
            dir_pdf = dir_pdf * tr;
            break;
        }
        origin = origin + dir * t;
        tmax -= t;
        reads++;
        const uint64_t va = nvdb_addr(g, root, (int)floorf(origin.x), (int)floorf(origin.y), (int)floorf(origin.z));
        const float dv = gf(g, va);
        const v3 sig_t = m.density * dv;
        const v3 sig_s = m.albedo();  // readAlbedo is 1 without an albedo volume
        const v3 tr = expv((-maj) * t) * invmax;
        bool real = false;
        if (SCATTER) real = rng.nextf().x < comp(sig_t, ch) * invc;
        const bool sink_here = ADJ && sink->grad && sink->vol == m.dens_vol;
        if (real) {
            // beta_c gains sigma_t,c = density_c * value: dlog = 1 / value
            if (sink_here && dv > 0) atomicAdd(sink->grad + va / 4, (sink->wgt.x + sink->wgt.y + sink->wgt.z) / dv);
            beta = beta * ((tr * sig_t) * sig_s);
            dir_pdf = dir_pdf * (tr * sig_t);
            scattered = true;
            return index_to_world(g, origin);
        }
        const v3 d = maj - sig_t;
        if (sink_here) {  // beta_c gains maj_c - sigma_t,c: dlog = -density_c / (maj_c - sigma_t,c)
            float gs = 0;
            if (d.x != 0) gs -= sink->wgt.x * m.density.x / d.x;
            if (d.y != 0) gs -= sink->wgt.y * m.density.y / d.y;
            if (d.z != 0) gs -= sink->wgt.z * m.density.z / d.z;
            atomicAdd(sink->grad + va / 4, gs);
        }
        beta = beta * (tr * d);
        dir_pdf = dir_pdf * (tr * d);
        nee_pdf = nee_pdf * (tr * maj);
    }
    return mk(0, 0, 0);
}

// Procedural volume candidates of traceRay (intersection.hlsli:92-122): the
// index-space box [bbox_min, bbox_max + 1] of the grid (or the [-1, 1] cube
// without one), tested after the triangles; the closest candidate in (0, h.t)
// replaces the hit.  Only boxes the ray actually crosses are candidates, as
// the acceleration structure reports an AABB only when the ray meets it.
__device__ inline void volume_boxes(const DevScene& S, v3 o, v3 d, Hit& h, bool& is_vol, uint32_t& vol_gn) {
    is_vol = false;
    for (uint32_t v = 0; v < S.vinst_count; v++) {
        const uint32_t ii = S.vinst[v];
        const st_transform& iT = S.inv_xf[ii];
        const v3 oo = xf_point(iT, o), od = xf_vector(iT, d);
        const uint32_t vidx = S.inst[ii].data;
        v3 org, dir, bmin, bmax;
        const uint8_t* g = nullptr;
        if (vidx >= S.vol_count) {
            org = oo;
            dir = od;
            bmin = mk(-1, -1, -1);
            bmax = mk(1, 1, 1);
        } else {
            g = S.vol[vidx];
            const uint64_t root = S.vol_root[vidx];
            org = world_to_index(g, oo);
            dir = world_to_index_dir(g, od);
            bmin = mk((float)(int32_t)g32(g, root), (float)(int32_t)g32(g, root + 4), (float)(int32_t)g32(g, root + 8));
            bmax = mk((float)(int32_t)g32(g, root + 12), (float)(int32_t)g32(g, root + 16), (float)(int32_t)g32(g, root + 20)) +
                   mk(1, 1, 1);
        }
        const v3 t0 = (bmin - org) / dir, t1 = (bmax - org) / dir;
        const float sx = fmax_(fmax_(fmin_(t0.x, t1.x), fmin_(t0.y, t1.y)), fmin_(t0.z, t1.z));
        const float sy = fmin_(fmin_(fmax_(t0.x, t1.x), fmax_(t0.y, t1.y)), fmax_(t0.z, t1.z));
        if (!(sx <= sy)) continue;
        const float t = sx > 0 ? sx : sy;
        if (!(t < h.t && t > 0)) continue;
        v3 ln = mk(t1.x == t, t1.y == t, t1.z == t) - mk(t0.x == t, t0.y == t, t0.z == t);
        if (g) ln = index_to_world_dir(g, ln);
        vol_gn = pack_normal(normalize(xf_vector(S.xf[ii], ln)));
        h.t = t;
        h.id = ii | (0xffffu << 16);
        h.hit = true;
        h.b1 = h.b2 = 0;
        is_vol = true;
    }
}

}  // namespace st
