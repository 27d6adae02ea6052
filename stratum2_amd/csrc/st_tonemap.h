// st_tonemap.h — the reference's post-HDR display transform on gfx950:
// exposure 2^E, the TonemapMode operators and the sRGB encode of
// src/Shaders/kernels/tonemap.slang:30-192 (helpers from
// src/Shaders/compat/common.h:81-104).  Shared by the kernels in
// st_render.hip; every operator keeps the reference's operation order.
#pragma once

#include "st_device.h"

namespace st {

constexpr float kTonemapQuant = 16384.0f;  // gMaxQuantization, tonemap.slang:122

__device__ __forceinline__ v3 tm_reinhard(v3 c) {
    const float l = lum(c);
    const v3 tc = c / (mk(1, 1, 1) + c);
    const v3 a = c / (1 + l);
    return a + (tc - a) * tc;  // lerp(c / (1 + l), tc, tc)
}
__device__ __forceinline__ v3 tm_reinhard_ext(v3 c, v3 mx) {
    const v3 m = mk(mx.x == 0 ? 1.0f : mx.x, mx.y == 0 ? 1.0f : mx.y, mx.z == 0 ? 1.0f : mx.z);
    return c / (mk(1, 1, 1) + c) * (mk(1, 1, 1) + c / (m * m));
}
__device__ __forceinline__ v3 tm_reinhard_lum(v3 c) {
    const float l = lum(c);
    const float l1 = l / (1 + l);
    return c * (l1 / l);
}
__device__ __forceinline__ v3 tm_reinhard_lum_ext(v3 c, float max_l) {
    const float l = lum(c);
    const float m = max_l == 0 ? 1.0f : max_l;
    const float l1 = (l / (1 + l)) * (1 + l / (m * m));
    return c * (l1 / l);
}
__device__ __forceinline__ float tm_u2(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
__device__ __forceinline__ v3 tm_uncharted2(v3 c, float max_l) {
    const float w = tm_u2(max_l == 0 ? 1.0f : max_l);
    return mk(tm_u2(c.x), tm_u2(c.y), tm_u2(c.z)) / w;
}
__device__ __forceinline__ float tm_filmic1(float c) {
    c = fmax_(0.0f, c - 0.004f);
    return (c * (6.2f * c + 0.5f)) / (c * (6.2f * c + 1.7f) + 0.06f);
}
__device__ __forceinline__ float tm_rtt1(float v) {
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}
__device__ __forceinline__ v3 tm_aces(v3 v) {
    // input and output matrices of aces_fitted (tonemap.slang:83-101), row-major
    const v3 i = mk(0.59719f * v.x + 0.35458f * v.y + 0.04823f * v.z, 0.07600f * v.x + 0.90834f * v.y + 0.01566f * v.z,
                    0.02840f * v.x + 0.13383f * v.y + 0.83777f * v.z);
    const v3 r = mk(tm_rtt1(i.x), tm_rtt1(i.y), tm_rtt1(i.z));
    return mk(sat(1.60475f * r.x + -0.53108f * r.y + -0.07367f * r.z), sat(-0.10208f * r.x + 1.10813f * r.y + -0.00605f * r.z),
              sat(-0.00327f * r.x + -0.07276f * r.y + 1.07602f * r.z));
}
__device__ __forceinline__ float tm_aces_approx1(float v) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    v *= 0.6f;
    return sat((v * (a * v + b)) / (v * (c * v + d) + e));
}
// viridisQuintic (compat/common.h:96-104): quintic fit of the viridis colormap
__device__ __forceinline__ v3 tm_viridis(float x) {
    const float x2 = x * x, x3 = x2 * x, x4 = x * x3, x5 = x2 * x3;
    return mk((0.280268003f + -0.143510503f * x + 2.225793877f * x2 + -14.815088879f * x3) + (25.212752309f * x4 + -11.772589584f * x5),
              (-0.002117546f + 1.617109353f * x + -1.909305070f * x2 + 2.701152864f * x3) + (-1.685288385f * x4 + 0.178738871f * x5),
              (0.300805501f + 2.614650302f * x + -12.019139090f * x2 + 28.933559110f * x3) + (-33.491294770f * x4 + 13.762053843f * x5));
}
// rgbToSrgb as the reference writes it (compat/common.h:88-94): pow(rgb * 1.055, 1/2.4) - 0.055
__device__ __forceinline__ float tm_srgb1(float c) { return c <= 0.0031308f ? c * 12.92f : powf(c * 1.055f, 1 / 2.4f) - 0.055f; }

__device__ inline v3 tonemap(int mode, v3 c, float4 mx) {
    switch (mode) {
        case ST_TONEMAP_REINHARD: return tm_reinhard(c);
        case ST_TONEMAP_REINHARD_EXTENDED: return tm_reinhard_ext(c, mk(mx.x, mx.y, mx.z));
        case ST_TONEMAP_REINHARD_LUMINANCE: return tm_reinhard_lum(c);
        case ST_TONEMAP_REINHARD_LUMINANCE_EXTENDED: return tm_reinhard_lum_ext(c, mx.w);
        case ST_TONEMAP_UNCHARTED2: return tm_uncharted2(c, mx.w);
        case ST_TONEMAP_FILMIC: return mk(tm_filmic1(c.x), tm_filmic1(c.y), tm_filmic1(c.z));
        case ST_TONEMAP_ACES: return tm_aces(c);
        case ST_TONEMAP_ACES_APPROX: return mk(tm_aces_approx1(c.x), tm_aces_approx1(c.y), tm_aces_approx1(c.z));
        case ST_TONEMAP_VIRIDIS_R: return tm_viridis(sat(lum(c)));
        case ST_TONEMAP_VIRIDIS_LENGTH_RGB: return tm_viridis(sat(lum(c) / (mx.w == 0 ? 1.0f : mx.w)));
        default: return c;
    }
}

}  // namespace st
