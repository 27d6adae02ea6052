// st_shade.h — scene access on the device: triangle shading frames, materials,
// light and environment sampling.  Restates src/Shaders/common/scene.hlsli,
// shading_data.hlsli, material.hlsli and lights.hlsli over the flat HBM copies
// of the reference's scene buffers.
#pragma once

#include "st_device.h"

namespace st {

struct Shading {  // ShadingData fields used on camera paths (compat/scene.h:185-198)
    v3 pos;
    uint32_t material;  // byte address
    uint32_t gn, sn, tan;
    float area;
    float u, v, uvs;    // mTexcoord, mTexcoordScreenSize (read by the textured kernels only)
};

__device__ __forceinline__ v3 xf_point(const st_transform& T, v3 v) {
    return mk(T.m[0][0] * v.x + T.m[0][1] * v.y + T.m[0][2] * v.z + T.m[0][3],
              T.m[1][0] * v.x + T.m[1][1] * v.y + T.m[1][2] * v.z + T.m[1][3],
              T.m[2][0] * v.x + T.m[2][1] * v.y + T.m[2][2] * v.z + T.m[2][3]);
}
__device__ __forceinline__ v3 xf_vector(const st_transform& T, v3 v) {
    return mk(T.m[0][0] * v.x + T.m[0][1] * v.y + T.m[0][2] * v.z, T.m[1][0] * v.x + T.m[1][1] * v.y + T.m[1][2] * v.z,
              T.m[2][0] * v.x + T.m[2][1] * v.y + T.m[2][2] * v.z);
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) {
    uint32_t v;
    __builtin_memcpy(&v, p, 4);
    return v;
}
__device__ __forceinline__ v3 ld_v3(const uint8_t* p) {
    float f[3];
    __builtin_memcpy(f, p, 12);
    return mk(f[0], f[1], f[2]);
}

// LoadTriangleIndices, scene.hlsli:43-61 (16- or 32-bit indices)
__device__ __forceinline__ void load_tri(const DevScene& S, const st_mesh_vertex_info& vi, uint32_t prim, uint32_t tri[3]) {
    const uint32_t ib = vi.packed_buffer_indices[0] & 0xffffu, stride = vi.packed_strides & 0xffu;
    const uint8_t* b = S.vb[ib] + vi.packed_offsets[0] + (size_t)prim * 3 * stride;
    if (stride == 2) {
        uint16_t v[3];
        __builtin_memcpy(v, b, 6);
        tri[0] = v[0];
        tri[1] = v[1];
        tri[2] = v[2];
    } else {
        tri[0] = ld_u32(b);
        tri[1] = ld_u32(b + 4);
        tri[2] = ld_u32(b + 8);
    }
}

// makeSphereShadingData, shading_data.hlsli:174-189: the sphere instance ii
// at the object-space point lp.  The area is the object-space 4 pi r^2 (the
// reference's scene build leaves sphere transforms a translation), and the
// tangent comes from the spherical uv taken as angles, as the reference
// writes it.
__device__ __attribute__((noinline)) void sphere_shading(const DevScene& S, uint32_t ii, v3 lp, Shading& r) {
    const st_instance in = S.inst[ii];
    const st_transform T = S.xf[ii];
    const v3 n = normalize(xf_vector(T, lp));
    r.pos = xf_point(T, lp);
    r.material = in.type_material_address >> 4;
    r.gn = r.sn = pack_normal(n);
    const float rad = __uint_as_float(in.data);
    r.area = 4 * ST_PI * rad * rad;
    float u, v;
    cart_to_uv(normalize(lp), u, v);
    const float su = sin_(u), cu = cos_(u), sv = sin_(v);
    const v3 dpdu = xf_vector(T, mk(-su * sv, 0.0f, cu * sv));
    r.tan = pack_normal(normalize(dpdu - n * dot(n, dpdu)));
    r.u = u;
    r.v = v;
    const float cv = cos_(v);
    const v3 dpdv = xf_vector(T, mk(cu * cv, -sv, su * cv));
    r.uvs = 1 / fmax_(len(dpdu), len(dpdv));
}

// makeTriangleShadingData(WithoutPosition), shading_data.hlsli:59-151
// lp != nullptr: the shading data at an object-space point on the triangle
// (makeTriangleShadingData from a local position, shading_data.hlsli:153-174):
// barycentrics solved from the point, world position T * lp.  A sphere
// instance (prim 0xffff) takes its shading data at lp.
__device__ inline void tri_shading(const DevScene& S, uint32_t ii, uint32_t prim, float bx, float by, Shading& r,
                                   const v3* lp = nullptr) {
    if (prim == 0xffffu) {
        sphere_shading(S, ii, *lp, r);
        return;
    }
    const st_instance in = S.inst[ii];
    const st_transform T = S.xf[ii];
    const st_mesh_vertex_info vi = S.vi[in.data & 0xffffu];
    uint32_t tri[3];
    load_tri(S, vi, prim, tri);
    const uint32_t pb = vi.packed_buffer_indices[0] >> 16, ps = (vi.packed_strides >> 8) & 0xffu;
    const uint8_t* P = S.vb[pb] + vi.packed_offsets[1];
    const v3 v0 = ld_v3(P + (size_t)ps * tri[0]), v1 = ld_v3(P + (size_t)ps * tri[1]), v2 = ld_v3(P + (size_t)ps * tri[2]);
    if (lp) {
        const v3 v1v0 = v1 - v0, v2v0 = v2 - v0, p_v0 = *lp - v0;
        const float d00 = dot(v1v0, v1v0), d01 = dot(v1v0, v2v0), d11 = dot(v2v0, v2v0);
        const float d20 = dot(p_v0, v1v0), d21 = dot(p_v0, v2v0);
        const float den = d00 * d11 - d01 * d01;
        bx = (d11 * d20 - d01 * d21) / den;
        by = (d00 * d21 - d01 * d20) / den;
    }
    r.material = in.type_material_address >> 4;
    const v3 dPds = xf_vector(T, v0 - v2), dPdt = xf_vector(T, v1 - v2);
    v3 gn = cross(dPds, dPdt);
    const float area2 = len(gn);
    gn = gn / area2;
    r.gn = pack_normal(gn);
    r.area = area2 / 2;
    float t0[2] = {0, 0}, t1[2] = {0, 0}, t2[2] = {0, 0};
    const uint32_t tb = vi.packed_buffer_indices[1] >> 16, ts = vi.packed_strides >> 24;
    if (tb < ST_IMAGE_LIMIT) {
        const uint8_t* Tp = S.vb[tb] + vi.packed_offsets[3];
        __builtin_memcpy(t0, Tp + (size_t)ts * tri[0], 8);
        __builtin_memcpy(t1, Tp + (size_t)ts * tri[1], 8);
        __builtin_memcpy(t2, Tp + (size_t)ts * tri[2], 8);
    }
    const float duvds0 = t2[0] - t0[0], duvds1 = t2[1] - t0[1], duvdt0 = t2[0] - t1[0], duvdt1 = t2[1] - t1[1];
    r.u = t0[0] + (t1[0] - t0[0]) * bx + duvds0 * by;
    r.v = t0[1] + (t1[1] - t0[1]) * bx + duvds1 * by;
    r.uvs = 1;
    const float det = duvds0 * duvdt1 - duvdt0 * duvds1;
    v3 dPdu;
    if (det != 0) {
        const float invDet = 1 / det;
        const float dsdu = duvdt1 * invDet, dtdu = -duvds1 * invDet;
        dPdu = -(dPds * dsdu + dPdt * dtdu);
        const float dsdv = duvdt0 * invDet, dtdv = -duvds0 * invDet;  // as the reference writes them
        const v3 dPdv = -(dPds * dsdv + dPdt * dtdv);
        r.uvs = 1 / fmax_(len(dPdu), len(dPdv));
    } else {
        v3 o1;
        make_onb(gn, dPdu, o1);
    }
    bool snvalid = false;
    v3 sn = mk(0, 0, 0);
    const uint32_t nb = vi.packed_buffer_indices[1] & 0xffffu, ns = (vi.packed_strides >> 16) & 0xffu;
    if (nb < ST_IMAGE_LIMIT) {  // gShadingNormals is on by default (TestRenderer.cpp:22)
        const uint8_t* N = S.vb[nb] + vi.packed_offsets[2];
        const v3 n0 = ld_v3(N + (size_t)ns * tri[0]), n1 = ld_v3(N + (size_t)ns * tri[1]), n2 = ld_v3(N + (size_t)ns * tri[2]);
        sn = (n0 + (n1 - n0) * bx) + (n2 - n0) * by;
        snvalid = !((sn.x == 0 && sn.y == 0 && sn.z == 0) || sn.x != sn.x || sn.y != sn.y || sn.z != sn.z);
    }
    if (!snvalid) {
        r.sn = r.gn;
        r.tan = pack_normal(normalize(dPdu));
    } else {
        sn = normalize(xf_vector(T, sn));
        const v3 tg = normalize(dPdu - sn * dot(sn, dPdu));
        r.sn = pack_normal(sn);
        r.tan = pack_normal(tg);
        if (dot(sn, gn) < 0) r.gn = pack_normal(-gn);
    }
    r.pos = lp ? xf_point(T, *lp) : xf_point(T, (v0 + (v1 - v0) * bx) + (v2 - v0) * by);
}

// The shading data of a traversal hit {t, id, b1, b2} of the ray (o, d): a
// triangle at its barycentrics, or a sphere at the object-space point o' + d' t
// (CommittedObjectRayOrigin + CommittedObjectRayDirection * t,
// intersection.hlsli:178-181).  Its primitive pick pdf is 1 for a sphere.
__device__ inline void hit_shading(const DevScene& S, uint4 hit, v3 o, v3 d, Shading& sd) {
    const uint32_t ii = hit.y & 0xffffu, prim = hit.y >> 16;
    if (prim == 0xffffu) {
        const st_transform& T = S.inv_xf[ii];
        const float t = __uint_as_float(hit.x);
        const v3 oo = mk(T.m[0][0] * o.x + T.m[0][1] * o.y + T.m[0][2] * o.z + T.m[0][3],
                         T.m[1][0] * o.x + T.m[1][1] * o.y + T.m[1][2] * o.z + T.m[1][3],
                         T.m[2][0] * o.x + T.m[2][1] * o.y + T.m[2][2] * o.z + T.m[2][3]);
        const v3 od = mk(T.m[0][0] * d.x + T.m[0][1] * d.y + T.m[0][2] * d.z, T.m[1][0] * d.x + T.m[1][1] * d.y + T.m[1][2] * d.z,
                         T.m[2][0] * d.x + T.m[2][1] * d.y + T.m[2][2] * d.z);
        sphere_shading(S, ii, oo + od * t, sd);
        return;
    }
    tri_shading(S, ii, prim, __uint_as_float(hit.z), __uint_as_float(hit.w), sd);
}
__device__ __forceinline__ float prim_pick(const DevScene& S, uint32_t id) {
    return (id >> 16) == 0xffffu ? 1.0f : 1.0f / (float)(S.inst[id & 0xffffu].data >> 16);
}

// SampleImage4 (scene.hlsli:101-109) of material image img: SampleLevel at
// lod = log2(max(uvs * max(w, h), 1e-6)) (0 without a screen size), linear
// between the two nearest levels as the static sampler filters
// (TestRenderer.cpp:54-57), bilinear with repeat addressing in each; the lod
// is clamped to the chain and log2 is the renderer's log over ln 2.
__device__ inline float4 tex_bilinear(const DevScene& S, uint32_t lvi, float u, float v) {
    const uint4 L = S.tex_lv[lvi];
    const int w = (int)L.y, h = (int)L.z;
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), ax = x - fx0, ay = y - fy0;
    const int x0 = (int)fx0, y0 = (int)fy0;
    const int xa = ((x0 % w) + w) % w, xb = (((x0 + 1) % w) + w) % w;
    const int ya = ((y0 % h) + h) % h, yb = (((y0 + 1) % h) + h) % h;
    const float4* P = S.tex_data + L.x;
    const float4 t00 = P[(size_t)ya * w + xa], t10 = P[(size_t)ya * w + xb], t01 = P[(size_t)yb * w + xa],
                 t11 = P[(size_t)yb * w + xb];
    const float4 top = make_float4(t00.x * (1 - ax) + t10.x * ax, t00.y * (1 - ax) + t10.y * ax, t00.z * (1 - ax) + t10.z * ax,
                                   t00.w * (1 - ax) + t10.w * ax);
    const float4 bot = make_float4(t01.x * (1 - ax) + t11.x * ax, t01.y * (1 - ax) + t11.y * ax, t01.z * (1 - ax) + t11.z * ax,
                                   t01.w * (1 - ax) + t11.w * ax);
    return make_float4(top.x * (1 - ay) + bot.x * ay, top.y * (1 - ay) + bot.y * ay, top.z * (1 - ay) + bot.z * ay,
                       top.w * (1 - ay) + bot.w * ay);
}
__device__ inline float4 tex_sample(const DevScene& S, uint32_t img, float u, float v, float uvs) {
    const uint2 I = S.tex_img[img];
    const uint4 L0 = S.tex_lv[I.x];
    float lod = 0;
    if (uvs > 0) lod = log_(fmax_(uvs * (float)(L0.y > L0.z ? L0.y : L0.z), 1e-6f)) * 1.44269504088896340736f;
    if (!(lod > 0)) lod = 0;
    if (lod > (float)(I.y - 1)) lod = (float)(I.y - 1);
    const int l0 = (int)floorf(lod), l1 = l0 + 1 < (int)I.y ? l0 + 1 : l0;
    const float f = lod - (float)l0;
    const float4 a = tex_bilinear(S, I.x + (uint32_t)l0, u, v);
    if (!(f > 0) || l1 == l0) return a;
    const float4 b = tex_bilinear(S, I.x + (uint32_t)l1, u, v);
    return make_float4(a.x * (1 - f) + b.x * f, a.y * (1 - f) + b.y * f, a.z * (1 - f) + b.z * f, a.w * (1 - f) + b.w * f);
}
// D3DX_FLOAT4_to_R8G8B8A8_UNORM and back
__device__ __forceinline__ uint32_t pack_unorm4(float4 v) {
    return (uint32_t)(sat(v.x) * 255 + 0.5f) | ((uint32_t)(sat(v.y) * 255 + 0.5f) << 8) |
           ((uint32_t)(sat(v.z) * 255 + 0.5f) << 16) | ((uint32_t)(sat(v.w) * 255 + 0.5f) << 24);
}
__device__ __forceinline__ float4 unpack_unorm4(uint32_t w) {
    return make_float4((float)(w & 0xffu) / 255.0f, (float)((w >> 8) & 0xffu) / 255.0f, (float)((w >> 16) & 0xffu) / 255.0f,
                       (float)(w >> 24) / 255.0f);
}
// LoadMaterial with its images (material.hlsli:22-45): base colour times the
// image (setBaseColor re-quantises to 8 bits), packed word 2 modulated with
// roughness through smoothness, packed word 3 modulated.  Emission images are
// refused at scene creation.
__device__ inline void apply_textures(const DevScene& S, Material& m, uint32_t addr, float u, float v,
                                                         float uvs) {
    const uint4 ii = *reinterpret_cast<const uint4*>(S.mat + addr / 4 + 4);
    if (ii.x < S.tex_count && S.tex_img[ii.x].y) {
        const float4 t = tex_sample(S, ii.x, u, v, uvs);
        const v3 b = m.base();
        const float bb[3] = {b.x * t.x, b.y * t.y, b.z * t.z};
        uint32_t w = m.p[0] & 0xff000000u;
        for (int c = 0; c < 3; c++) w |= (uint32_t)floorf(sat(bb[c]) * 255 + 0.5f) << (8 * c);
        m.p[0] = w;
    }
    if (ii.z < S.tex_count && S.tex_img[ii.z].y) {
        float4 x = unpack_unorm4(m.p[2]);
        const float4 t = tex_sample(S, ii.z, u, v, uvs);
        x.x *= t.x;
        x.z *= t.z;
        x.w *= t.w;
        x.y = 1 - (1 - x.y) * (1 - t.y);
        m.p[2] = pack_unorm4(x);
    }
    if (ii.w < S.tex_count && S.tex_img[ii.w].y) {
        const float4 x = unpack_unorm4(m.p[3]), t = tex_sample(S, ii.w, u, v, uvs);
        m.p[3] = pack_unorm4(make_float4(x.x * t.x, x.y * t.y, x.z * t.z, x.w * t.w));
    }
}

// ShadingData::toLocal / toWorld / shadingNormalCorrection, shading_data.hlsli:24-54
struct Frame {
    v3 t, b, n;
    __device__ __forceinline__ v3 to_local(v3 v) const { return mk(dot(v, t), dot(v, b), dot(v, n)); }
    __device__ __forceinline__ v3 to_world(v3 v) const { return (t * v.x + b * v.y) + n * v.z; }
};
__device__ __forceinline__ Frame make_frame(const Shading& sd) {
    Frame f;
    f.n = unpack_normal(sd.sn);
    f.t = unpack_normal(sd.tan);
    f.b = cross(f.n, f.t);
    return f;
}
__device__ __forceinline__ float sgn(float x) { return x > 0 ? 1.0f : (x < 0 ? -1.0f : 0.0f); }
// ShadingData::shadingNormalCorrection, shading_data.hlsli:35-54.  ADJ: light
// paths, whose throughput also takes the ratio of geometric to shading cosines.
template <bool ADJ = false>
__device__ __forceinline__ float shading_normal_correction(const Frame& f, v3 lgn, v3 din, v3 dout) {
    const float ngin = dot(lgn, din), ngout = dot(lgn, dout);
    if (sgn(ngout) != sgn(dout.z) || sgn(ngin) != sgn(din.z)) return 0;
    float G = 1;
    if (ADJ) {
        const float num = ngout * din.z, denom = dout.z * ngin;
        if (fabsf(denom) > 1e-5f) G *= fabsf(num / denom);
    }
    return G;
}

// LoadMaterial, material.hlsli:15-21 (configurations carry no material textures)
// flags: the render's ST_FLAG_*; with ST_FLAG_LAMBERT_GGX (and not
// ST_FLAG_LAMBERTIAN) a material of metallic at least 0.5 is a GGX conductor
__device__ __forceinline__ Material load_material(const DevScene& S, uint32_t addr, uint32_t flags) {
    Material m;
    const uint4 w = *reinterpret_cast<const uint4*>(S.mat + addr / 4);
    m.p[0] = w.x;
    m.p[1] = w.y;
    m.p[2] = w.z;
    m.p[3] = w.w;
    m.escale = __uint_as_float(S.mat[addr / 4 + 8]);
    m.has_bo = false;
    m.bo = mk(0, 0, 0);
    m.model = ((flags & (ST_FLAG_LAMBERT_GGX | ST_FLAG_LAMBERTIAN)) == ST_FLAG_LAMBERT_GGX && m.metal() >= 0.5f) ? 2u : 0u;
    return m;
}
// load_material at a shading point; TEX: with its images (the kernels
// instantiated for scenes with material images)
template <bool TEX>
__device__ __forceinline__ Material load_material(const DevScene& S, const Shading& sd, uint32_t flags) {
    Material m = load_material(S, sd.material, flags);
    if (TEX) apply_textures(S, m, sd.material, sd.u, sd.v, sd.uvs);
    return m;
}

// The per-triangle albedo (st_scene_set_triangle_albedo) of a triangle hit,
// replacing the material's base colour; entry = the triangle's index in the
// flattened order (mesh instances in order, primitives in order).
__device__ __forceinline__ void apply_triangle_albedo(const DevScene& S, Material& m, uint32_t inst, uint32_t prim,
                                                      uint32_t& entry) {
    entry = S.tri_base[inst] + prim;
    const float* a = S.tri_alb + 3 * (size_t)entry;
    m.bo = mk(a[0], a[1], a[2]);
    m.has_bo = true;
}

// ------------------------------------------------------------- environment
__device__ __forceinline__ float texel_lum(const DevScene& S, int level, int x, int y) {
    const float4 t = S.env_level[level][(size_t)y * S.env_w[level] + x];
    return lum(mk(t.x, t.y, t.z));
}
// bilinear SampleLevel(uv, 0), repeat addressing (TestRenderer.cpp:54-57)
__device__ inline v3 env_bilinear(const DevScene& S, float u, float v) {
    const int w = S.env_w[0], h = S.env_h[0];
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), ax = x - fx0, ay = y - fy0;
    const int x0 = (int)fx0, y0 = (int)fy0;
    const int xa = ((x0 % w) + w) % w, xb = (((x0 + 1) % w) + w) % w;
    const int ya = ((y0 % h) + h) % h, yb = (((y0 + 1) % h) + h) % h;
    const float4* L = S.env_level[0];
    const float4 t00 = L[(size_t)ya * w + xa], t10 = L[(size_t)ya * w + xb], t01 = L[(size_t)yb * w + xa],
                 t11 = L[(size_t)yb * w + xb];
    const float r0 = t00.x * (1 - ax) + t10.x * ax, r1 = t01.x * (1 - ax) + t11.x * ax;
    const float g0 = t00.y * (1 - ax) + t10.y * ax, g1 = t01.y * (1 - ax) + t11.y * ax;
    const float b0 = t00.z * (1 - ax) + t10.z * ax, b1 = t01.z * (1 - ax) + t11.z * ax;
    return mk(r0 * (1 - ay) + r1 * ay, g0 * (1 - ay) + g1 * ay, b0 * (1 - ay) + b1 * ay);
}
// SampleTexel, lights.hlsli:49-97 (hierarchical warp down the mip pyramid)
__device__ inline void sample_texel(const DevScene& S, float rx, float ry, float& u, float& v, float& pdf) {
    const int lc = S.env_levels;
    int cx = 0, cy = 0, lw = 1, lh = 1;
    pdf = 1;
    const int iend = lc - 1 < 11 ? lc - 1 : 11;
    for (int i = 1; i < iend; i++) {
        const int L = lc - 1 - i, W = S.env_w[L], H = S.env_h[L];
        const float inv_h = 1 / (float)H;
        cx *= W / lw;
        cy *= H / lh;
        float p0 = 0, p1 = 0, p2 = 0, p3 = 0;
        if ((uint32_t)W - (uint32_t)cx > 1) {
            const float sy = sin_(ST_PI * ((float)cy + 0.5f) * inv_h);
            p0 = texel_lum(S, L, cx, cy) * sy;
            p1 = texel_lum(S, L, cx + 1, cy) * sy;
        }
        if ((uint32_t)H - (uint32_t)cy > 1) {
            const float sy = sin_(ST_PI * ((float)cy + 1.5f) * inv_h);
            p2 = texel_lum(S, L, cx, cy + 1) * sy;
            p3 = texel_lum(S, L, cx + 1, cy + 1) * sy;
        }
        const float sum = p0 + p1 + p2 + p3;
        if (sum < 1e-6f) continue;
        const float q[4] = {p0 / sum, p1 / sum, p2 / sum, p3 / sum};
        for (int j = 0; j < 4; j++) {
            if (j == 3 || rx < q[j]) {
                cx += j % 2;
                cy += j / 2;
                pdf *= q[j];
                rx /= q[j];
                break;
            }
            rx -= q[j];
        }
        lw = W;
        lh = H;
    }
    pdf *= (float)((uint32_t)lw * (uint32_t)lh);
    u = ((float)cx + rx) / (float)lw;
    v = ((float)cy + ry) / (float)lh;
}
// SampleTexelPdf, lights.hlsli:98-139
__device__ inline float sample_texel_pdf(const DevScene& S, float u, float v) {
    const int lc = S.env_levels;
    int lw = 1, lh = 1;
    float pdf = 1;
    const int iend = lc - 1 < 11 ? lc - 1 : 11;
    for (int i = 1; i < iend; i++) {
        const int L = lc - 1 - i, W = S.env_w[L], H = S.env_h[L];
        const float inv_h = 1 / (float)H;
        const int ux = (int)((float)W * u), uy = (int)((float)H * v);
        const int cx = (ux / 2) * 2, cy = (uy / 2) * 2;
        const int ox = ux - cx < 1 ? ux - cx : 1, oy = uy - cy < 1 ? uy - cy : 1;
        float p[4] = {0, 0, 0, 0};
        if ((uint32_t)W - (uint32_t)cx > 1) {
            const float sy = sin_(ST_PI * ((float)cy + 0.5f) * inv_h);
            p[0] = texel_lum(S, L, cx, cy) * sy;
            p[1] = texel_lum(S, L, cx + 1, cy) * sy;
        }
        if ((uint32_t)H - (uint32_t)cy > 1) {
            const float sy = sin_(ST_PI * ((float)cy + 1.5f) * inv_h);
            p[2] = texel_lum(S, L, cx, cy + 1) * sy;
            p[3] = texel_lum(S, L, cx + 1, cy + 1) * sy;
        }
        const float sum = p[0] + p[1] + p[2] + p[3];
        if (sum < 1e-6f) continue;
        pdf *= p[oy * 2 + ox] / sum;
        lw = W;
        lh = H;
    }
    return pdf * (float)((uint32_t)lw * (uint32_t)lh);
}
// Alias-table environment sampling (ENV_ALIAS): one texel by Walker's method
// over luminance*sin(theta) of the full-resolution map, uniform inside it.
// Four independent uniforms: rx picks the table column, ra is the accept test
// against the column's probability, ju and jv place the point in the texel.
// (Taking the accept test and the jitter from the fraction of rx * N would
// leave them 23 - log2(N) bits: 4 bits for a 1024x512 map.)
__device__ inline void sample_alias(const DevScene& S, float rx, float ra, float ju, float jv, float& u, float& v,
                                    float& pdf) {
    const int W = S.env_w[0], H = S.env_h[0];
    const uint32_t N = (uint32_t)W * (uint32_t)H;
    uint32_t i = (uint32_t)(rx * (float)N);
    if (i >= N) i = N - 1;
    const float2 e = S.env_alias[i];
    if (!(ra < e.x)) i = __float_as_uint(e.y);
    pdf = S.env_pdf[i];
    u = ((float)(i % (uint32_t)W) + fmin_(ju, S.env_jmax)) / (float)W;
    v = ((float)(i / (uint32_t)W) + fmin_(jv, S.env_jmax)) / (float)H;
}
__device__ __forceinline__ float alias_pdf(const DevScene& S, float u, float v) {
    const int W = S.env_w[0], H = S.env_h[0];
    int x = (int)(u * (float)W), y = (int)(v * (float)H);
    x = x < 0 ? 0 : (x >= W ? W - 1 : x);
    y = y < 0 ? 0 : (y >= H ? H - 1 : y);
    return S.env_pdf[(size_t)y * W + x];
}

// EvaluateEnvironment, lights.hlsli:143-163
template <bool ALIAS>
__device__ inline v3 evaluate_environment(const DevScene& S, float env_prob, v3 d, float& pdfw) {
    if (!S.has_env) {
        pdfw = 0;
        return mk(0, 0, 0);
    }
    v3 e = mk(S.env_emission[0], S.env_emission[1], S.env_emission[2]);
    if (S.env_has_image) {
        float u, v;
        cart_to_uv(d, u, v);
        e = e * env_bilinear(S, u, v);
        const float p = ALIAS ? alias_pdf(S, u, v) : sample_texel_pdf(S, u, v);
        pdfw = p / (2 * ST_PI * ST_PI * sqrtf(1 - d.y * d.y));
    } else {
        pdfw = 1 / (4 * ST_PI);
    }
    if (S.light_count > 0) pdfw *= env_prob;
    return e;
}

// The position, geometry normal and area of a point on an emissive triangle,
// as tri_shading computes them (makeTriangleShadingData, shading_data.hlsli:
// 59-151: the geometry normal flips to the side of an interpolated vertex
// normal), from the emitter table: two loads instead of the chain light ->
// instance -> vertex info -> indices -> vertices.
__device__ inline void light_point(const DevScene& S, uint32_t li, uint32_t prim, float bx, float by, v3& pos,
                                   uint32_t& gnp, float& area) {
    const float4* L = S.light_rec + 4 * (size_t)li;
    const float4 r0 = L[0], r1 = L[1], r2 = L[2];
    const st_transform T = {{{r0.x, r0.y, r0.z, r0.w}, {r1.x, r1.y, r1.z, r1.w}, {r2.x, r2.y, r2.z, r2.w}}};
    const float4* R = S.light_tri + 5 * ((size_t)__float_as_uint(L[3].y) + prim);
    const float4 f0 = R[0], f1 = R[1], f2 = R[2];
    const v3 v0 = mk(f0.x, f0.y, f0.z), v1 = mk(f1.x, f1.y, f1.z), v2 = mk(f2.x, f2.y, f2.z);
    const v3 dPds = xf_vector(T, v0 - v2), dPdt = xf_vector(T, v1 - v2);
    v3 gn = cross(dPds, dPdt);
    const float area2 = len(gn);
    gn = gn / area2;
    gnp = pack_normal(gn);
    area = area2 / 2;
    if (__float_as_uint(f0.w) & 1u) {
        const float4 f3 = R[3], f4 = R[4];
        const v3 n0 = mk(f1.w, f2.w, f3.x), n1 = mk(f3.y, f3.z, f3.w), n2 = mk(f4.x, f4.y, f4.z);
        v3 sn = (n0 + (n1 - n0) * bx) + (n2 - n0) * by;
        const bool snvalid = !((sn.x == 0 && sn.y == 0 && sn.z == 0) || sn.x != sn.x || sn.y != sn.y || sn.z != sn.z);
        if (snvalid) {
            sn = normalize(xf_vector(T, sn));
            if (dot(sn, gn) < 0) gnp = pack_normal(-gn);
        }
    }
    pos = xf_point(T, (v0 + (v1 - v0) * bx) + (v2 - v0) * by);
}

// A point on an emissive sphere instance (SampleEmission's sphere branch,
// lights.hlsli:215-220): makeSphereShadingData at radius * sampleUniformSphere.
__device__ inline void light_sphere(const DevScene& S, uint32_t li, float rad, float u1, float u2, v3& pos, uint32_t& gnp,
                                    float& area) {
    const float4* L = S.light_rec + 4 * (size_t)li;
    const float4 r0 = L[0], r1 = L[1], r2 = L[2];
    const st_transform T = {{{r0.x, r0.y, r0.z, r0.w}, {r1.x, r1.y, r1.z, r1.w}, {r2.x, r2.y, r2.z, r2.w}}};
    const v3 lp = sample_uniform_sphere(u1, u2) * rad;
    gnp = pack_normal(normalize(xf_vector(T, lp)));
    pos = xf_point(T, lp);
    area = 4 * ST_PI * rad * rad;
}

struct LightSample {  // IlluminationSampleRecord, lights.hlsli:32-45
    v3 radiance, dir;
    float pdf, dist, cos_light;
    bool finite;
    v3 pos;       // finite lights: the sampled point
    uint32_t gn;  // finite lights: its packed geometry normal
};

// SampleEmission + SampleIllumination, lights.hlsli:166-260.  The reference
// leaves mDirectionToLight unset for environment samples (lights.hlsli:234-244);
// here it is the sampled direction.
template <bool ALIAS>
__device__ inline LightSample sample_illumination(const DevScene& S, float env_prob, float4 rnd, v3 ref) {
    LightSample r;
    r.radiance = mk(0, 0, 0);
    r.dir = mk(0, 0, 0);
    r.pdf = 0;
    r.dist = 0;
    r.cos_light = 0;
    r.finite = false;
    r.pos = mk(0, 0, 0);
    r.gn = 0;
    if (S.has_env) {
        if (S.light_count == 0 || rnd.w < env_prob) {
            v3 d;
            float pdf, su = 0, sv = 0;
            if (S.env_has_image) {
                float u, v;
                if (ALIAS) {
                    // rnd.z is free on this branch; rnd.w, rescaled past the
                    // environment-or-emitter choice, is uniform again
                    const float ju = S.light_count > 0 ? rnd.w / env_prob : rnd.w;
                    sample_alias(S, rnd.x, rnd.z, ju, rnd.y, u, v, pdf);
                } else {
                    sample_texel(S, rnd.x, rnd.y, u, v, pdf);
                }
                su = u;
                sv = v;
                d = uv_to_cart(u, v);
                pdf /= (2 * ST_PI * ST_PI * sqrtf(1 - sq(d.y)));
            } else {
                d = sample_uniform_sphere(rnd.x, rnd.y);
                pdf = 1 / (4 * ST_PI);
            }
            if (S.light_count > 0) pdf *= env_prob;
            r.pdf = pdf;
            r.dist = __int_as_float(0x7f800000);
            r.cos_light = 1;
            r.dir = d;
            if (ALIAS && S.env_has_image) {
                // alias-table mode: the radiance at the sampled uv itself (no
                // direction -> uv round trip; equal up to float rounding)
                r.radiance = mk(S.env_emission[0], S.env_emission[1], S.env_emission[2]) * env_bilinear(S, su, sv);
            } else {
                float tmp;
                r.radiance = evaluate_environment<ALIAS>(S, env_prob, d, tmp);
            }
            return r;
        }
        if (S.light_count > 0) rnd.w = (rnd.w - env_prob) / (1 - env_prob);
    }
    if (S.light_count == 0) return r;
    const uint32_t li = (uint32_t)(rnd.z * (float)S.light_count) % S.light_count;
    float pdf = 1 / (float)S.light_count;
    if (S.has_env) pdf *= 1 - env_prob;
    // {triangles, first, emission word, scale}; a sphere instance: {0, radius, ...}
    const float4 l3 = S.light_rec[4 * (size_t)li + 3];
    const uint32_t pc = __float_as_uint(l3.x);
    v3 lpos;
    uint32_t lgn;
    float larea;
    if (pc == 0) {
        light_sphere(S, li, l3.y, rnd.x, rnd.y, lpos, lgn, larea);
    } else {
        const uint32_t prim = (uint32_t)(rnd.w * (float)pc) % pc;
        pdf /= (float)pc;
        const float a = sqrtf(rnd.x);  // sampleUniformTriangle, common.hlsli:36-39
        light_point(S, li, prim, 1 - a, a * rnd.y, lpos, lgn, larea);
    }
    pdf /= larea;
    const uint32_t ew = __float_as_uint(l3.z);  // Material::emission, as load_material's record holds it
    r.radiance = mk(un8(ew, 0), un8(ew, 8), un8(ew, 16)) * l3.w;
    r.pdf = pdf;
    r.finite = true;
    r.pos = lpos;
    r.gn = lgn;
    v3 dd = lpos - ref;
    r.dist = len(dd);
    r.dir = dd / r.dist;
    r.cos_light = -dot(unpack_normal(lgn), r.dir);
    if (r.cos_light <= 0) {  // lights are one-sided
        r.radiance = mk(0, 0, 0);
        r.pdf = 0;
    }
    return r;
}

}  // namespace st
