// st_device.h — gfx950 device functions of the path-tracing hot path.
//
// Restates the per-vertex arithmetic of Stratum2's TestRenderer for HIP:
// pcg4d sampling (src/Shaders/common/rng.hlsli), octahedral snorm16 packing
// (common.hlsli, D3DX_DXGIFormatConvert.inl), triangle shading frames
// (shading_data.hlsli), Disney/Lambert BSDFs (materials/disney.hlsli,
// lambertian.hlsli, common/microfacet.hlsli) and light/environment sampling
// (common/lights.hlsli).  Everything is compiled with -ffp-contract=off so that
// each operation rounds once, in the order the reference writes it.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "st_layout.h"
#include "st_math.h"

#define ST_PI 3.14159265358979323846f
#define ST_1_PI (1.0f / ST_PI)

namespace st {

struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 operator/(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ v3 normalize(v3 a) { return a / sqrtf(dot(a, a)); }
__device__ __forceinline__ bool any_pos(v3 a) { return a.x > 0 || a.y > 0 || a.z > 0; }
__device__ __forceinline__ bool all_nonpos(v3 a) { return a.x <= 0 && a.y <= 0 && a.z <= 0; }
__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float sat(float a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }
__device__ __forceinline__ float sq(float a) { return a * a; }
__device__ __forceinline__ float lum(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ v3 lerp(v3 a, v3 b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float lerp(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// ---------------------------------------------------------------- rng
// xxhash32 and pcg, src/Shaders/common/rng.hlsli:5-21
__device__ __forceinline__ uint32_t xxhash32(uint32_t p) {
    uint32_t h = p + 374761393u;
    h = 668265263u * ((h << 17) | (h >> 15));
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t pcg1(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
// pcg4d, src/Shaders/common/rng.hlsli:23-35
__device__ __forceinline__ uint4 pcg4d(uint4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    v.x ^= v.x >> 16; v.y ^= v.y >> 16; v.z ^= v.z >> 16; v.w ^= v.w >> 16;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    return v;
}
// RandomSampler (rng.hlsli:39-60): state (px, py, seed, counter); the counter is
// the only part that has to live in the path state.
struct Rng {
    uint32_t px, py, seed, w;
    __device__ __forceinline__ uint4 next() {
        w++;
        return pcg4d(make_uint4(px, py, seed, w));
    }
    __device__ __forceinline__ float4 nextf() {
        uint4 v = next();
        return make_float4(__uint_as_float(0x3f800000u | (v.x >> 9)) - 1.0f, __uint_as_float(0x3f800000u | (v.y >> 9)) - 1.0f,
                           __uint_as_float(0x3f800000u | (v.z >> 9)) - 1.0f, __uint_as_float(0x3f800000u | (v.w >> 9)) - 1.0f);
    }
};

// ------------------------------------------------------------ packing
// D3DX_FLOAT_to_INT(SaturateSigned(v), 32767), D3DX_DXGIFormatConvert.inl:290-322
__device__ __forceinline__ int32_t snorm16(float v) {
    if (v != v) v = 0;
    v = fmin_(fmax_(v, -1.0f), 1.0f);
    return (int32_t)truncf(v * 32767.0f + (v >= 0 ? 0.5f : -0.5f));
}
__device__ __forceinline__ float unsnorm16(int32_t i) { return fmax_((float)i / 32767.0f, -1.0f); }
// packNormal, common.hlsli:3-19
__device__ __forceinline__ uint32_t pack_normal(v3 v) {
    float r = 1.0f / (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
    float px = v.x * r, py = v.y * r;
    if (v.z <= 0) {
        float nx = (1 - fabsf(py)) * (px >= 0 ? 1.0f : -1.0f);
        float ny = (1 - fabsf(px)) * (py >= 0 ? 1.0f : -1.0f);
        px = nx;
        py = ny;
    }
    return ((uint32_t)snorm16(px) & 0xffffu) | ((uint32_t)snorm16(py) << 16);
}
// unpackNormal, common.hlsli:8-22
__device__ __forceinline__ v3 unpack_normal(uint32_t p) {
    float x = unsnorm16((int32_t)(p << 16) >> 16), y = unsnorm16((int32_t)(p & 0xffff0000u) >> 16);
    v3 v = mk(x, y, 1 - (fabsf(x) + fabsf(y)));
    if (v.z < 0) {
        float nx = (1 - fabsf(v.y)) * (v.x >= 0 ? 1.0f : -1.0f);
        float ny = (1 - fabsf(v.x)) * (v.y >= 0 ? 1.0f : -1.0f);
        v.x = nx;
        v.y = ny;
    }
    return normalize(v);
}
// FP16 storage of PathState fields (testrenderer.slang:75-97): RNE conversion.
__device__ __forceinline__ uint32_t f2h(float f) { return (uint32_t)__half_as_ushort(__float2half_rn(f)); }
__device__ __forceinline__ float h2f(uint32_t h) { return __half2float(__ushort_as_half((unsigned short)(h & 0xffffu))); }
__device__ __forceinline__ float q16(float f) { return h2f(f2h(f)); }

__device__ __forceinline__ float un8(uint32_t w, int shift) { return (float)((w >> shift) & 0xffu) / 255.0f; }

// rayOffset, src/Shaders/common/intersection.hlsli:17-32
__device__ __forceinline__ float roff1(float p, float n) {
    const float origin = 1.0f / 32.0f, fs = 1.0f / 65536.0f;
    int32_t of = (int32_t)(256.0f * n);
    if (fabsf(p) < origin) return p + fs * n;
    return __int_as_float(__float_as_int(p) + ((p < 0.0f) ? -of : of));
}
__device__ __forceinline__ v3 ray_offset(v3 p, v3 n) { return mk(roff1(p.x, n.x), roff1(p.y, n.y), roff1(p.z, n.z)); }
__device__ __forceinline__ v3 ray_offset(v3 p, v3 n, v3 d) { return ray_offset(p, dot(n, d) < 0 ? -n : n); }

// makeOrthonormal, common.hlsli:24-34
__device__ __forceinline__ void make_onb(v3 N, v3& r0, v3& r1) {
    if (N.x != N.y || N.x != N.z) r0 = mk(N.z - N.y, N.x - N.z, N.y - N.x);
    else r0 = mk(N.z - N.y, N.x + N.z, -N.y - N.x);
    r0 = normalize(r0);
    r1 = cross(N, r0);
}
// warps, common.hlsli:36-77
__device__ __forceinline__ v3 sample_uniform_sphere(float u1, float u2) {
    float z = u2 * 2 - 1, r = sqrtf(fmax_(1 - z * z, 0)), phi = 2 * ST_PI * u1;
    return mk(r * cos_(phi), r * sin_(phi), z);
}
__device__ __forceinline__ void concentric_disc(float u1, float u2, float& ox, float& oy) {
    float ux = 2 * u1 - 1, uy = 2 * u2 - 1, th, r;
    if (ux == 0 && uy == 0) {
        ox = oy = 0;
        return;
    }
    if (fabsf(ux) > fabsf(uy)) {
        r = ux;
        th = (ST_PI / 4) * (uy / ux);
    } else {
        r = uy;
        th = ST_PI / 2 - (ST_PI / 4) * (ux / uy);
    }
    ox = r * cos_(th);
    oy = r * sin_(th);
}
__device__ __forceinline__ v3 sample_cos_hemisphere(float u1, float u2) {
    float x, y;
    concentric_disc(u1, u2, x, y);
    return mk(x, y, sqrtf(fmax_(0, 1 - (x * x + y * y))));
}
__device__ __forceinline__ float cos_hemisphere_pdfw(float c) { return fmax_(c, 0.0f) / ST_PI; }
__device__ __forceinline__ float mis1(float a) { return a * a; }
// mis, testrenderer.slang:284-290.  Where a pdf overflowed (the fp16 store
// of a sharp lobe's direction pdf reads back as inf, testrenderer.slang:
// 93-97) the reference's quotient is inf / inf = NaN and poisons the pixel;
// here it takes its limit: 1 for the infinite pdf, 0 against it, 1/2 for two.
__device__ __forceinline__ float mis2(float a, float b) {
    const float a2 = mis1(a), b2 = mis1(b);
    const float w = a2 / (a2 + b2);
    return w == w ? w : (a2 > b2 ? 1.0f : (a2 < b2 ? 0.0f : 0.5f));
}
// compat/common.h:137-150
__device__ __forceinline__ float stable_atan2(float y, float x) {
    return x == 0.0f ? (y == 0 ? 0 : (y < 0 ? -ST_PI / 2 : ST_PI / 2)) : atan2_(y, x);
}
__device__ __forceinline__ void cart_to_uv(v3 v, float& u, float& w) {
    float th = stable_atan2(v.z, v.x);
    u = th * ST_1_PI * 0.5f + 0.5f;
    w = acos_(fmin_(fmax_(v.y, -1.0f), 1.0f)) * ST_1_PI;
}
__device__ __forceinline__ v3 uv_to_cart(float u, float v) {
    u = u * 2 - 1;
    u *= ST_PI;
    v *= ST_PI;
    float sp = sin_(v);
    return mk(sp * cos_(u), cos_(v), sp * sin_(u));
}

// ---------------------------------------------------------- materials
// PackedMaterialData (compat/material_data.h:12-77) + emission scale.
struct Material {
    uint32_t p[4];
    float escale;
    bool has_bo;  // bo replaces the packed base colour (a per-triangle albedo)
    uint32_t model;  // 2: the GGX conductor of ST_FLAG_LAMBERT_GGX (read by the Lambert instantiations only)
    v3 bo;
    __device__ __forceinline__ v3 base() const { return has_bo ? bo : mk(un8(p[0], 0), un8(p[0], 8), un8(p[0], 16)); }
    __device__ __forceinline__ v3 emission() const { return mk(un8(p[1], 0), un8(p[1], 8), un8(p[1], 16)) * escale; }
    __device__ __forceinline__ float spec() const { return un8(p[0], 24); }
    __device__ __forceinline__ float sheen() const { return un8(p[1], 24); }
    __device__ __forceinline__ float metal() const { return un8(p[2], 0); }
    __device__ __forceinline__ float rough() const { return un8(p[2], 8); }
    __device__ __forceinline__ float aniso() const { return un8(p[2], 16); }
    __device__ __forceinline__ float subsurf() const { return un8(p[2], 24); }
    __device__ __forceinline__ float cc() const { return un8(p[3], 0); }
    __device__ __forceinline__ float ccgloss() const { return un8(p[3], 8); }
    __device__ __forceinline__ float trans() const { return un8(p[3], 16); }
    __device__ __forceinline__ float eta1() const { return un8(p[3], 24) + 1; }
};

// microfacet.hlsli:22-179
// x^5 of the Schlick Fresnel and retro-reflection terms (disney.hlsli uses
// pow(x, 5)): three multiplies, within 3 ulp of the correctly rounded power
__device__ __forceinline__ float pow5(float x) {
    const float x2 = x * x;
    return x2 * x2 * x;
}
__device__ __forceinline__ v3 schlick3(v3 F0, float c) {
    float k = pow5(fmax_(1 - c, 0.0f));
    return mk(F0.x + (1 - F0.x) * k, F0.y + (1 - F0.y) * k, F0.z + (1 - F0.z) * k);
}
__device__ __forceinline__ float schlick1(float F0, float c) { return F0 + (1 - F0) * pow5(fmax_(1 - c, 0.0f)); }
__device__ __forceinline__ float fresnel_dielectric(float ni, float eta) {
    float nt2 = 1 - (1 - ni * ni) / (eta * eta);
    if (nt2 < 0) return 1;
    float nt = sqrtf(nt2), a = fabsf(ni);
    float rs = (a - eta * nt) / (a + eta * nt);
    float rp = (eta * a - nt) / (eta * a + nt);
    return (rs * rs + rp * rp) / 2;
}
__device__ __forceinline__ float gtr1(float ndh, float a) {
    float a2 = a * a, t = 1 + (a2 - 1) * ndh * ndh;
    return (a2 - 1) / (ST_PI * log_(a2) * t);
}
__device__ __forceinline__ float gtr2(v3 h, float ax, float ay) {
    v3 hs = h / mk(ax, ay, 1);
    float l2 = dot(hs, hs);
    return 1 / (ST_PI * ax * ay * l2 * l2);
}
__device__ __forceinline__ float smith2(v3 v, float ax, float ay) {
    float ax2 = ax * ax, ay2 = ay * ay;
    v3 v2 = v * v;
    float L = (-1 + sqrtf(1 + (v2.x * ax2 + v2.y * ay2) / v2.z)) / 2;
    return 1 / (1 + L);
}
__device__ __forceinline__ v3 sample_vndf(v3 din, float ax, float ay, float r0, float r1) {
    bool flip = din.z < 0;
    if (flip) din = -din;
    v3 hin = normalize(mk(ax * din.x, ay * din.y, din.z));
    float r = sqrtf(r0), phi = 2 * ST_PI * r1;
    float t1 = r * cos_(phi), t2 = r * sin_(phi);
    float s = (1 + hin.z) / 2;
    t2 = (1 - s) * sqrtf(1 - t1 * t1) + s * t2;
    v3 dn = mk(t1, t2, sqrtf(fmax_(0.0f, 1 - t1 * t1 - t2 * t2)));
    v3 f0, f1;
    make_onb(hin, f0, f1);
    v3 hn = f0 * dn.x + f1 * dn.y + hin * dn.z;
    v3 o = normalize(mk(ax * hn.x, ay * hn.y, fmax_(0.0f, hn.z)));
    return flip ? -o : o;
}

struct Eval {
    v3 f;
    float fwd, rev;
};
struct Sample {
    v3 dir, f;
    float fwd, rev, eta, rough;
};

// Disney eval, materials/disney.hlsli:9-139.  ADJ: light paths (Adjoint), whose
// refracted radiance is not scaled by eta^2 (eta_factor, :123).
template <bool ADJ = false>
__device__ inline v3 disney_eval(const Material& m, v3 din, v3 dout) {
    float aspect = sqrtf(1 - m.aniso() * 0.9f), alpha = sq(m.rough());
    float ax = fmax_(1e-4f, alpha / aspect), ay = fmax_(1e-4f, alpha * aspect), metallic = m.metal();
    if (din.z * dout.z > 0) {
        float ndin = fabsf(din.z), ndout = dout.z;
        v3 h = normalize(din + dout);
        if (h.z * din.z < 0) h = -h;
        float ndh = h.z, hdin = dot(h, din), hdout = dot(h, dout);
        v3 c = mk(0, 0, 0);
        if (din.z >= 0 && dout.z >= 0) {
            float Fd90 = 0.5f + 2 * m.rough() * hdout * hdout;
            float sno = pow5(1 - ndout), sni = pow5(1 - ndin), sho = pow5(1 - hdout);
            float based = (1 + (Fd90 - 1) * sno) * (1 + (Fd90 - 1) * sni);
            float Fss90 = hdout * hdout * m.rough();
            float Fss = (1 + (Fss90 - 1) * sno) * (1 + (Fss90 - 1) * sni);
            float ss = 1.25f * (Fss * (1 / (ndout + ndin) - 0.5f) + 0.5f);
            v3 bc = m.base();
            c = c + ((bc * ((1 - m.trans()) * (1 - metallic))) * (lerp(based, ss, m.subsurf()) / ST_PI)) * ndout;
            float l = lum(bc);
            v3 ctint = l > 0 ? bc / l : mk(1, 1, 1);
            v3 csheen = lerp(mk(1, 1, 1), ctint, 0.5f);
            c = c + ((csheen * ((1 - metallic) * m.sheen())) * sho) * ndout;
            if (ndin > 0 && hdout > 0 && ndh > 0) {
                float eta = m.eta1();
                float f0 = (eta - 1) * (eta - 1) / ((eta + 1) * (eta + 1));
                v3 c0 = lerp(lerp(mk(1, 1, 1), ctint, 0.5f) * (m.spec() * f0), m.base(), metallic);
                v3 F = schlick3(c0, hdout);
                float D = gtr2(h, ax, ay), G = smith2(din, ax, ay) * smith2(dout, ax, ay);
                float sw = 1 - m.trans() * (1 - metallic);
                c = c + (((F * sw) * D) * G) / (4 * ndin);
            }
            if (m.cc() > 0 && ndin > 0 && ndh > 0) {  // a zero clearcoat weight adds exactly 0
                float ac = lerp(0.1f, 0.001f, m.ccgloss());
                float Fc = schlick1(0.04f, hdout), Dc = gtr1(ndh, ac);
                float Gc = smith2(din, 0.25f, 0.25f) * smith2(dout, 0.25f, 0.25f);
                float v = m.cc() * Fc * Dc * Gc / (4 * ndin);
                c = c + mk(v, v, v);
            }
        }
        float gw = (1 - metallic) * m.trans();
        if (gw > 0) {
            float eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
            float Fg = fresnel_dielectric(hdin, eta), D = gtr2(h, ax, ay);
            float G = smith2(din, ax, ay) * smith2(dout, ax, ay);
            c = c + m.base() * (gw * (Fg * D * G) / (4 * fabsf(ndin)));
        }
        return c;
    }
    float gw = (1 - metallic) * m.trans();
    if (gw <= 0) return mk(0, 0, 0);
    float eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
    v3 h = normalize(din + dout * eta);
    if (h.z * din.z < 0) h = -h;
    float hdin = dot(h, din), hdout = dot(h, dout), sd = hdin + eta * hdout;
    float Fg = fresnel_dielectric(hdin, eta), D = gtr2(h, ax, ay);
    float G = smith2(din, ax, ay) * smith2(dout, ax, ay);
    v3 bc = m.base();
    v3 sqb = mk(sqrtf(bc.x), sqrtf(bc.y), sqrtf(bc.z));
    if (ADJ) return sqb * (gw * ((1 / (eta * eta)) * (1 - Fg) * D * G * eta * eta * fabsf(hdout * hdin)) / (fabsf(din.z) * sd * sd));
    return sqb * (gw * ((1 - Fg) * D * G * eta * eta * fabsf(hdout * hdin)) / (fabsf(din.z) * sd * sd));
}

// Disney evalPdf, disney.hlsli:141-231
__device__ inline float disney_pdf(const Material& m, v3 din, v3 dout) {
    float alpha = sq(m.rough()), metallic = m.metal(), aspect = sqrtf(1 - m.aniso() * 0.9f);
    float ax = fmax_(0.0001f, alpha / aspect), ay = fmax_(0.0001f, alpha * aspect);
    float ac = (1 - m.ccgloss()) * 0.1f + m.ccgloss() * 0.001f, tr = m.trans();
    float dw = (1 - metallic) * (1 - tr), mw = (1 - tr * (1 - metallic)), gw = (1 - metallic) * tr, cw = m.cc();
    float tw = dw + mw + gw + cw;
    float dp = dw / tw, mp = mw / tw, gp = gw / tw, cp = cw / tw;
    if (din.z < 0) {
        dp = 0;
        mp = 0;
        cp = 0;
        if (gp > 0) gp = 1;
    }
    if (din.z * dout.z > 0) {
        v3 h = normalize(din + dout);
        if (h.z < 0) h = -h;
        float ndin = din.z, ndh = h.z, hdin = dot(h, din), hdout = dot(h, dout);
        if (din.z >= 0 && dout.z >= 0) {
            dp *= fmax_(dout.z, 0.0f) / ST_PI;
            if (ndin > 0) {
                float D = gtr2(h, ax, ay), Gi = smith2(din, ax, ay);
                mp *= (D * Gi / (4 * ndin));
            } else {
                mp = 0;
            }
            if (cp != 0 && ndh > 0 && hdout > 0) cp *= (gtr1(ndh, ac) * ndh / (4 * hdout));
            else cp = 0;
        }
        if (gp != 0) {  // finite factor (ndin != 0 here): a zero weight stays exactly 0
            float eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
            float Fg = fresnel_dielectric(hdin, eta), D = gtr2(h, ax, ay), Gi = smith2(din, ax, ay);
            gp *= (Fg * D * Gi / (4 * fabsf(ndin)));
        }
    } else {
        float eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
        v3 h = normalize(din + dout * eta);
        if (h.z < 0) h = -h;
        float hdin = dot(h, din), hdout = dot(h, dout);
        float D = gtr2(h, ax, ay), Gi = smith2(din, ax, ay), Fg = fresnel_dielectric(hdin, eta);
        float sd = hdin + eta * hdout, dh = eta * eta * hdout / (sd * sd);
        gp *= (1 - Fg) * D * Gi * fabsf(dh * hdin / din.z);
    }
    return dp + mp + gp + cp;
}

// Disney sample, disney.hlsli:233-380
__device__ inline Sample disney_sample(const Material& m, v3 din, float rx, float ry, float rz) {
    Sample r;
    r.dir = mk(0, 0, 0);
    r.f = mk(0, 0, 0);
    r.fwd = r.rev = r.eta = r.rough = 0;
    float aspect = sqrtf(1 - m.aniso() * 0.9f), alpha = sq(m.rough());
    float ax = fmax_(0.0001f, alpha / aspect), ay = fmax_(0.0001f, alpha * aspect);
    float ac = (1 - m.ccgloss()) * 0.1f + m.ccgloss() * 0.001f;
    float metallic = m.metal(), tr = m.trans();
    float dw = (1 - metallic) * (1 - tr), mw = (1 - tr * (1 - metallic)), gw = (1 - metallic) * tr, cw = m.cc();
    if (din.z >= 0) {
        float tw = dw + mw + gw + cw;
        float dp = dw / tw, mp = mw / tw, gp = gw / tw;
        if (rz <= dp) {
            r.dir = sample_cos_hemisphere(rx, ry);
            r.rough = 1;
            return r;
        } else if (rz <= (dp + mp)) {
            v3 h = sample_vndf(din, ax, ay, rx, ry);
            r.dir = normalize(-din + h * (2 * dot(din, h)));
            r.rough = m.rough();
            return r;
        } else if (rz <= (dp + mp + gp)) {
            if (gp <= 0) return r;
            v3 h = sample_vndf(din, ax, ay, rx, ry);
            float hdin = dot(h, din), eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
            float F = fresnel_dielectric(hdin, eta), u = (rz - (dp + mp)) / gp;
            if (u <= F) {
                r.dir = normalize(-din + h * (2 * dot(din, h)));
                return r;  // the reference returns before assigning a roughness (stays 0)
            }
            float hdo2 = 1 - (1 - hdin * hdin) / (eta * eta);
            if (hdo2 <= 0) return r;
            if (hdin < 0) h = -h;
            float hdo = sqrtf(hdo2);
            r.dir = (-din) / eta + h * (fabsf(hdin) / eta - hdo);
            r.eta = eta;
            r.rough = m.rough();
            return r;
        } else {
            float a2 = ac * ac;
            float ch = sqrtf(fmax_(0.0f, (1 - pow_(a2, 1 - rx)) / (1 - a2)));
            float sh = sqrtf(fmax_(1 - ch * ch, 0.0f)), az = 2 * ST_PI * ry;
            v3 mn = mk(sh * cos_(az), sh * sin_(az), ch);
            r.dir = normalize(-din + mn * (2 * dot(din, mn)));
            r.rough = sqrtf(ac);
            return r;
        }
    }
    v3 h = sample_vndf(din, ax, ay, rx, ry);
    if (h.z < 0) h = -h;
    float hdin = dot(h, din), eta = din.z > 0 ? m.eta1() : 1 / m.eta1();
    float F = fresnel_dielectric(hdin, eta);
    if (rz <= F) {
        r.dir = normalize(-din + h * (2 * dot(din, h)));
    } else {
        float hdo2 = 1 - (1 - hdin * hdin) / (eta * eta);
        if (hdo2 <= 0) return r;
        if (hdin < 0) h = -h;
        float hdo = sqrtf(hdo2);
        r.dir = (-din) / eta + h * (fabsf(hdin) / eta - hdo);
        r.eta = eta;
    }
    r.rough = m.rough();
    return r;
}

// Visible-normal sampling as Heitz publishes it (JCGT 7(4), 2018, "Sampling the
// GGX Distribution of Visible Normals", listing 1): the disk's first axis is
// horizontal and its second lies in the plane of the normal and the view, so
// the half-disk warp follows the visible hemisphere.  (sample_vndf, the
// reference's, builds the disk frame with makeOrthonormal instead, whose axes
// are not so aligned away from normal incidence: its directions are then not
// distributed as the pdf it reports.)
__device__ __forceinline__ v3 sample_vndf_heitz(v3 din, float ax, float ay, float r0, float r1) {
    const bool flip = din.z < 0;
    if (flip) din = -din;
    const v3 vh = normalize(mk(ax * din.x, ay * din.y, din.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const v3 e1 = lensq > 0 ? mk(-vh.y, vh.x, 0) / sqrtf(lensq) : mk(1, 0, 0);
    const v3 e2 = cross(vh, e1);
    const float r = sqrtf(r0), phi = 2 * ST_PI * r1;
    float t1 = r * cos_(phi), t2 = r * sin_(phi);
    const float s = (1 + vh.z) / 2;
    t2 = (1 - s) * sqrtf(1 - t1 * t1) + s * t2;
    const v3 nh = e1 * t1 + e2 * t2 + vh * sqrtf(fmax_(0.0f, 1 - t1 * t1 - t2 * t2));
    const v3 o = normalize(mk(ax * nh.x, ay * nh.y, fmax_(0.0f, nh.z)));
    return flip ? -o : o;
}

// The GGX conductor of ST_FLAG_LAMBERT_GGX: the metallic lobe of the Disney
// BSDF alone (disney.hlsli:56-70 and :163-169 at metallic 1, no clearcoat):
// f = F(h.o) D(h) G1(i) G1(o) / (4 |n.i|) (the cosine of o included), F
// Schlick's about the base colour, D the anisotropic GGX (GTR2) with Disney's
// (alpha / aspect, alpha * aspect), alpha = roughness^2 (at least 1e-4), G1
// Smith's; directions from the visible normals (sample_vndf_heitz), pdf =
// D G1(i) / (4 |n.i|).
// Reflection only, two-sided as the Lambert model.
__device__ inline Eval ggx_eval(const Material& m, v3 din, v3 dout) {
    Eval r;
    r.f = mk(0, 0, 0);
    r.fwd = r.rev = 0;
    if (!(din.z * dout.z > 0)) return r;
    if (din.z < 0) {
        din = -din;
        dout = -dout;
    }
    const float aspect = sqrtf(1 - m.aniso() * 0.9f), alpha = sq(m.rough());
    const float ax = fmax_(1e-4f, alpha / aspect), ay = fmax_(1e-4f, alpha * aspect);
    const v3 h = normalize(din + dout);
    const float D = gtr2(h, ax, ay), Gi = smith2(din, ax, ay), Go = smith2(dout, ax, ay);
    const v3 F = schlick3(m.base(), dot(h, dout));
    r.f = (F * (D * Gi * Go)) / (4 * din.z);
    r.fwd = D * Gi / (4 * din.z);
    r.rev = D * Go / (4 * dout.z);
    return r;
}
__device__ inline Sample ggx_sample(const Material& m, v3 din, float rx, float ry) {
    const float aspect = sqrtf(1 - m.aniso() * 0.9f), alpha = sq(m.rough());
    const float ax = fmax_(1e-4f, alpha / aspect), ay = fmax_(1e-4f, alpha * aspect);
    const v3 h = sample_vndf_heitz(din, ax, ay, rx, ry);
    Sample r;
    r.dir = normalize(-din + h * (2 * dot(din, h)));
    const Eval f = ggx_eval(m, din, r.dir);
    r.f = f.f;
    r.fwd = f.fwd;
    r.rev = f.rev;
    r.eta = 0;
    r.rough = m.rough();
    return r;
}
__device__ inline Eval lambert_eval(const Material& m, v3 din, v3 dout) {
    Eval r;
    if (din.z * dout.z < 0) {
        r.f = mk(0, 0, 0);
        r.fwd = r.rev = 0;
        return r;
    }
    r.f = (m.base() * fabsf(dout.z)) / ST_PI;
    r.fwd = cos_hemisphere_pdfw(fmax_(1e-5f, fabsf(dout.z)));
    r.rev = cos_hemisphere_pdfw(fmax_(1e-5f, fabsf(din.z)));
    return r;
}
__device__ inline Sample lambert_sample(const Material& m, v3 din, float rx, float ry) {
    Sample r;
    r.dir = sample_cos_hemisphere(rx, ry);
    if (din.z < 0) r.dir = -r.dir;
    r.f = (m.base() * fabsf(r.dir.z)) / ST_PI;
    r.fwd = cos_hemisphere_pdfw(fmax_(1e-5f, fabsf(r.dir.z)));
    r.rev = cos_hemisphere_pdfw(fmax_(1e-5f, fabsf(din.z)));
    r.rough = 1;
    r.eta = 0;
    return r;
}

// evaluateReflectance: Lambert (lambertian.hlsli:13-24) or Disney (disney.hlsli:391-403).
// The Lambert instantiations also serve ST_FLAG_LAMBERT_GGX, whose metallic
// materials are GGX conductors (model 2), so the Disney kernels carry no
// per-material branch.
template <bool LAMBERT, bool ADJ = false>
__device__ inline Eval bsdf_eval(const Material& m, v3 din, v3 dout) {
    if (LAMBERT) return m.model == 2 ? ggx_eval(m, din, dout) : lambert_eval(m, din, dout);
    Eval r;
    r.f = disney_eval<ADJ>(m, din, dout);
    r.fwd = disney_pdf(m, din, dout);
    r.rev = disney_pdf(m, dout, din);
    return r;
}
// sampleDirection: lambertian.hlsli:29-39, disney.hlsli:415-429
template <bool LAMBERT, bool ADJ = false>
__device__ inline Sample bsdf_sample(const Material& m, float rx, float ry, float rz, v3 din) {
    if (LAMBERT) return m.model == 2 ? ggx_sample(m, din, rx, ry) : lambert_sample(m, din, rx, ry);
    Sample r = disney_sample(m, din, rx, ry, rz);
    Eval f = bsdf_eval<false, ADJ>(m, din, r.dir);
    r.f = f.f;
    r.fwd = f.fwd;
    r.rev = f.rev;
    return r;
}
template <bool LAMBERT>
__device__ __forceinline__ bool m_singular(const Material& m) {
    if (LAMBERT) return m.model == 2 && (1 - m.rough()) > 0.75f;  // a GGX conductor: isSingular at metallic 1
    return fmax_(fmax_(m.spec(), m.metal()), m.trans()) * (1 - m.rough()) > 0.75f;
}
__device__ __forceinline__ float m_cont_prob(const Material& m) { return sat(lum(m.base())); }

}  // namespace st
