// st_svgf.h — the reference's SVGF denoiser on gfx950: temporal accumulation
// with reprojection (src/Shaders/kernels/svgf/temporal_accumulation.slang),
// variance estimation (estimate_variance.slang) and the edge-aware filter
// passes (atrous.slang), each in the reference's operation order.
//
// Layout: W*H row-major images.  Colour RGBA32F, moments RG32F, visibility
// 2 x u32, depth RGBA32F {depth, previous depth, derivative xy}.  The history
// keeps visibility and only the depth channel of the frame (the accumulation
// reads nothing else of the previous frame's depth), double-buffered with the
// accumulated colour and moments so that accumulation writes the next history
// while it gathers the previous one.  Accumulation also writes a geometry
// image {decoded shading normal, depth}: the variance estimate and the filter
// passes read a neighbour's normal and depth in one 16-byte load instead of
// decoding the octahedral normal (five divisions and a square root) at every
// tap.  The decode is the same function, so the weights are unchanged.
//
// Launch shape: 64x4-pixel tiles (one row of 64 pixels per wavefront, so each
// image load of a wavefront is one contiguous 1 KiB span).  Tiles are handed
// to the 8 XCDs in contiguous bands: consecutive workgroups go to different
// XCDs, so workgroup b takes tile (b % 8) * per_xcd + b / 8 and the rows an
// XCD's tiles share as filter halos stay in that XCD's L2.
#pragma once

#include "../../include/stratum2.h"
#include "st_device.h"

namespace st {

constexpr int kSvgfTileX = 64, kSvgfTileY = 4, kSvgfXcds = 8;
constexpr uint32_t kSvgfInvalidInstance = 0xFFFFu;

// The edge-stopping weights' transcendentals, as shading hardware evaluates
// them: pow(x, 2^K) by K squarings (within (2^K - 1)/2 ulp of the exact power,
// about the error of exp2(y * log2 x) on a GPU) and exp(x) as the hardware
// 2^x of x * log2(e).  Both are a few instructions where the library
// functions take dozens, and these taps are ALU-bound.
template <int K>
__device__ __forceinline__ float pow_2k(float x) {
#pragma unroll
    for (int i = 0; i < K; i++) x = x * x;
    return x;
}
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }

struct SvgfView {
    int min_x, min_y, max_x, max_y;  // ViewData::mImageMin / mImageMax
};

struct SvgfTiles {
    uint32_t W, H, tiles_x, tiles, per_xcd;
};

__host__ inline SvgfTiles svgf_tiles(uint32_t W, uint32_t H) {
    SvgfTiles t;
    t.W = W;
    t.H = H;
    t.tiles_x = (W + kSvgfTileX - 1) / kSvgfTileX;
    t.tiles = t.tiles_x * ((H + kSvgfTileY - 1) / kSvgfTileY);
    t.per_xcd = (t.tiles + kSvgfXcds - 1) / kSvgfXcds;
    return t;
}

// Pixel of this thread, or false when its tile or pixel lies outside the image.
__device__ __forceinline__ bool svgf_pixel(const SvgfTiles& t, int& x, int& y) {
    const uint32_t b = blockIdx.x;
    const uint32_t tile = (b % kSvgfXcds) * t.per_xcd + b / kSvgfXcds;
    if (tile >= t.tiles) return false;
    x = (int)((tile % t.tiles_x) * kSvgfTileX + threadIdx.x % kSvgfTileX);
    y = (int)((tile / t.tiles_x) * kSvgfTileY + threadIdx.x / kSvgfTileX);
    return x < (int)t.W && y < (int)t.H;
}

// ViewData::isInside (scene.h:146); the host checks that the view lies in the image.
__device__ __forceinline__ bool svgf_inside(const SvgfView& v, int x, int y) {
    return x >= v.min_x && y >= v.min_y && x < v.max_x && y < v.max_y;
}

__device__ __forceinline__ uint32_t svgf_map(const uint32_t* map, uint32_t count, uint32_t i) {
    if (!map) return i;
    return i < count ? map[i] : kSvgfInvalidInstance;
}

__device__ __forceinline__ bool finite3(float4 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z); }
__device__ __forceinline__ bool finite4(float4 c) { return finite3(c) && isfinite(c.w); }

struct SvgfAccumArgs {
    SvgfTiles t;
    SvgfView v;
    uint32_t flags, history_limit;
    float cos_normal, shadow_offset, shadow_scale;
    const uint32_t* map;
    uint32_t map_count;
    const uint2* vis;
    const float4* depth;
    const float2* prev_uv;
    const float4* input;
    const float4* albedo;
    const uint2* prev_vis;
    const float* prev_depth;
    const float4* prev_color;
    const float2* prev_moments;
    float4* color;       // accumulated colour {rgb, sample count}
    float2* moments;     // accumulated {luminance, luminance^2}
    uint2* keep_vis;     // this frame's visibility and depth, the next call's history
    float* keep_depth;
    float4* geo;         // {decoded normal, depth} of this frame, for the passes that follow (or null)
    float4* out;         // also receives the colour when no filter pass follows (else null)
};

// temporal_accumulation.slang:34-128.
__global__ __launch_bounds__(256) void k_svgf_accumulate(SvgfAccumArgs a) {
    int x, y;
    if (!svgf_pixel(a.t, x, y)) return;
    const size_t i = (size_t)y * a.t.W + x;
    const uint2 vis = a.vis[i];
    const float4 dep = a.depth[i];
    a.keep_vis[i] = vis;
    a.keep_depth[i] = dep.x;
    const v3 n = unpack_normal(vis.y);
    if (a.geo) a.geo[i] = make_float4(n.x, n.y, n.z, dep.x);

    float4 cp = make_float4(0, 0, 0, 0);
    float mp0 = 0, mp1 = 0, sum_w = 0;
    if (a.flags & ST_DENOISE_REPROJECTION) {
        const uint32_t inst = vis.x & 0xFFFFu;
        if (inst != kSvgfInvalidInstance) {
            const float2 uv = a.prev_uv[i];
            const float px = (float)a.v.min_x + uv.x * (float)(a.v.max_x - a.v.min_x) - 0.5f;
            const float py = (float)a.v.min_y + uv.y * (float)(a.v.max_y - a.v.min_y) - 0.5f;
            // positions beyond +-2^24 (or NaN) lie outside every image: no taps
            const bool far = !(fabsf(px) < 16777216.0f && fabsf(py) < 16777216.0f);
            const int ix = far ? 0 : (int)px, iy = far ? 0 : (int)py;  // int2(): truncation
            const float wx = px - floorf(px), wy = py - floorf(py);    // frac()
            const float dd = sqrtf(dep.z * dep.z + dep.w * dep.w);
            const uint32_t want = svgf_map(a.map, a.map_count, inst);
            for (int yy = 0; yy <= 1 && !far; yy++)
                for (int xx = 0; xx <= 1; xx++) {
                    const int qx = ix + xx, qy = iy + yy;
                    if (!svgf_inside(a.v, qx, qy)) continue;
                    const size_t j = (size_t)qy * a.t.W + qx;
                    const uint2 pv = a.prev_vis[j];
                    if (want != (pv.x & 0xFFFFu)) continue;
                    if ((a.flags & ST_DENOISE_CHECK_NORMAL) && dot(n, unpack_normal(pv.y)) < a.cos_normal) continue;
                    if ((a.flags & ST_DENOISE_CHECK_DEPTH) && fabsf(dep.y - a.prev_depth[j]) >= 2 * dd) continue;
                    const float4 c = a.prev_color[j];
                    if (c.w <= 0 || !finite4(c)) continue;
                    const float wc = (xx == 0 ? (1 - wx) : wx) * (yy == 0 ? (1 - wy) : wy);
                    const float2 m = a.prev_moments[j];
                    cp.x += c.x * wc;
                    cp.y += c.y * wc;
                    cp.z += c.z * wc;
                    cp.w += c.w * wc;
                    mp0 += m.x * wc;
                    mp1 += m.y * wc;
                    sum_w += wc;
                }
        }
    } else {
        cp = a.prev_color[i];
        if (!finite3(cp)) {
            cp = make_float4(0, 0, 0, 0);
        } else {
            const float2 m = a.prev_moments[i];
            mp0 = m.x;
            mp1 = m.y;
        }
        sum_w = 1;
    }

    float4 cc = a.input[i];
    if (a.flags & ST_DENOISE_DEMODULATE_ALBEDO) {
        const float4 al = a.albedo[i];
        cc.x /= (1e-2f + al.x);
        cc.y /= (1e-2f + al.y);
        cc.z /= (1e-2f + al.z);
    }
    if (!finite3(cc)) cc = make_float4(0, 0, 0, 0);
    if (!isfinite(mp0) || !isfinite(mp1)) mp0 = mp1 = 0;
    const float l = lum(mk(cc.x, cc.y, cc.z));

    float4 o;
    float2 mo;
    if (sum_w > 0 && cp.w > 0) {
        const float inv = 1 / sum_w;
        cp.x *= inv;
        cp.y *= inv;
        cp.z *= inv;
        cp.w *= inv;
        mp0 *= inv;
        mp1 *= inv;
        float n = cp.w + cc.w;
        if (a.history_limit > 0 && n > (float)a.history_limit) n = (float)a.history_limit;
        float alpha = sat(cc.w / n);
        if (a.flags & ST_DENOISE_SHADOW_HYSTERESIS) {
            const float lp = lum(mk(cp.x, cp.y, cp.z));
            const float den = l + lp;
            const float q = sat((den == 0 ? 0.0f : (lp - l) / den) - a.shadow_offset);
            const float alpha_l = q * q * a.shadow_scale;
            alpha = 1 - (1 - alpha) * sat(1 - alpha_l);
        }
        o = make_float4(lerp(cp.x, cc.x, alpha), lerp(cp.y, cc.y, alpha), lerp(cp.z, cc.z, alpha), n);
        mo = make_float2(lerp(mp0, l, alpha), lerp(mp1, l * l, alpha));
    } else {
        o = cc;
        mo = make_float2(l, l * l);
    }
    a.color[i] = o;
    a.moments[i] = mo;
    if (a.out) a.out[i] = o;
}

// Start of a history (Denoiser.cpp:243-249): cleared colour and moments, the
// frame's visibility and depth kept, the radiance passed through.
__global__ __launch_bounds__(256) void k_svgf_restart(SvgfTiles t, const uint2* vis, const float4* depth,
                                                       const float4* input, float4* color, float2* moments,
                                                       uint2* keep_vis, float* keep_depth, float4* out) {
    int x, y;
    if (!svgf_pixel(t, x, y)) return;
    const size_t i = (size_t)y * t.W + x;
    color[i] = make_float4(0, 0, 0, 0);
    moments[i] = make_float2(0, 0);
    keep_vis[i] = vis[i];
    keep_depth[i] = depth[i].x;
    if (out) out[i] = input[i];
}

struct SvgfVarianceArgs {
    SvgfTiles t;
    SvgfView v;
    uint32_t flags, history_limit;
    float boost_length;
    const uint32_t* map;
    uint32_t map_count;
    const uint2* vis;
    const float4* depth;
    const float4* geo;
    const float4* color;
    const float2* moments;
    float4* out;
};

// estimate_variance.slang:23-74.
__global__ __launch_bounds__(256) void k_svgf_variance(SvgfVarianceArgs a) {
    int x, y;
    if (!svgf_pixel(a.t, x, y)) return;
    const size_t i = (size_t)y * a.t.W + x;
    float4 c = a.color[i];
    float2 m = a.moments[i];
    const uint2 vis = a.vis[i];
    const uint32_t inst = vis.x & 0xFFFFu;
    const float histlen = c.w;
    if (inst == kSvgfInvalidInstance || histlen >= (float)a.history_limit) {
        a.out[i] = make_float4(c.x, c.y, c.z, fabsf(m.y - m.x * m.x));
        return;
    }
    const float4 dep = a.depth[i];
    const float4 g = a.geo[i];
    const v3 n = mk(g.x, g.y, g.z);
    const uint32_t want = svgf_map(a.map, a.map_count, inst);
    float sum_w = 1;
    const int r = histlen > 1 ? 2 : 3;
    for (int yy = -r; yy <= r; yy++)
        for (int xx = -r; xx <= r; xx++) {
            if (xx == 0 && yy == 0) continue;
            const int qx = x + xx, qy = y + yy;
            if (!svgf_inside(a.v, qx, qy)) continue;
            const size_t j = (size_t)qy * a.t.W + qx;
            if (want != (a.vis[j].x & 0xFFFFu)) continue;
            const float4 gj = a.geo[j];
            float w_z = 0, w_n = 1;
            if (a.flags & ST_DENOISE_CHECK_DEPTH) {
                const float gx = dep.z * (float)xx, gy = dep.w * (float)yy;
                w_z = fabsf(gj.w - dep.x) / (sqrtf(gx * gx + gy * gy) + 1e-2f);
            }
            if (a.flags & ST_DENOISE_CHECK_NORMAL) w_n = pow_2k<7>(sat(dot(mk(gj.x, gj.y, gj.z), n)));
            const float w = exp_hw(-w_z) * w_n;
            if (!isfinite(w)) continue;
            const float2 mj = a.moments[j];
            const float4 cj = a.color[j];
            m.x += mj.x * w;
            m.y += mj.y * w;
            c.x += cj.x * w;
            c.y += cj.y * w;
            c.z += cj.z * w;
            sum_w += w;
        }
    sum_w = 1 / sum_w;
    m.x *= sum_w;
    m.y *= sum_w;
    c.x *= sum_w;
    c.y *= sum_w;
    c.z *= sum_w;
    float var = fabsf(m.y - m.x * m.x);
    if (a.boost_length > 0) var *= fmax_(1.0f, a.boost_length / (1 + c.w));
    a.out[i] = make_float4(c.x, c.y, c.z, var);
}

struct SvgfAtrousArgs {
    SvgfTiles t;
    SvgfView v;
    uint32_t flags, filter_type, iteration;
    int step;
    float sigma_boost;
    const float4* geo;
    const float4* depth;
    const float4* in;
    float4* out;
};

// TapData (atrous.slang:31-89), holding the launch's fields by value.
struct SvgfTap {
    SvgfView v;
    uint32_t W, flags, iteration;
    int step;
    const float4* geo;
    const float4* in;
    int x, y;
    v3 n;
    float z, dzx, dzy, l, sigma_l;
    float4 sum;
    float sum_w;

    __device__ __forceinline__ void tap(int ox, int oy, float kw) {
        const int qx = x + ox, qy = y + oy;
        if (!svgf_inside(v, qx, qy)) return;
        const size_t j = (size_t)qy * W + qx;
        const float4 c = in[j];
        const float lp = lum(mk(c.x, c.y, c.z));
        const float w_l = fabsf(lp - l) / fmax_(sigma_l, 1e-10f);
        float w_z = 0, w_n = 1;
        if (flags & (ST_DENOISE_CHECK_DEPTH | ST_DENOISE_CHECK_NORMAL)) {
            const float4 g = geo[j];
            if (flags & ST_DENOISE_CHECK_DEPTH) {
                // the offset already carries the step; the reference scales it again
                const float gx = dzx * (float)(ox * step), gy = dzy * (float)(oy * step);
                w_z = fabsf(g.w - z) / (sqrtf(gx * gx + gy * gy) + 1e-2f);
            }
            if (flags & ST_DENOISE_CHECK_NORMAL) w_n = pow_2k<8>(fmax_(0.0f, dot(mk(g.x, g.y, g.z), n)));
        }
        const float w = exp_hw(-(w_l * w_l) - w_z) * kw * w_n;
        if (!isfinite(w)) return;
        sum.x += c.x * w;
        sum.y += c.y * w;
        sum.z += c.z * w;
        sum.w += c.w * (w * w);
        sum_w += w;
    }
    __device__ __forceinline__ void box(int r) {
        for (int yy = -r; yy <= r; yy++)
            for (int xx = -r; xx <= r; xx++)
                if (xx != 0 || yy != 0) tap(xx * step, yy * step, 1.0f);
    }
    __device__ __forceinline__ void subsampled() {
        const int s = step;
        if ((iteration & 1) == 0) {
            tap(-2 * s, 0, 1.0f);
            tap(2 * s, 0, 1.0f);
        } else {
            tap(0, -2 * s, 1.0f);
            tap(0, 2 * s, 1.0f);
        }
        tap(-s, s, 1.0f);
        tap(s, s, 1.0f);
        tap(-s, -s, 1.0f);
        tap(s, -s, 1.0f);
    }
    // 5x5 B3-spline taps of atrous(), atrous.slang:128-158, in its order
    __device__ __forceinline__ void atrous5() {
        const int s = step;
        const float k1 = 2.0f / 3.0f, k2 = 1.0f / 6.0f, k11 = 4.0f / 9.0f, k12 = 1.0f / 9.0f, k22 = 1.0f / 36.0f;
        tap(s, 0, k1); tap(0, s, k1); tap(-s, 0, k1); tap(0, -s, k1);
        tap(2 * s, 0, k2); tap(0, 2 * s, k2); tap(-2 * s, 0, k2); tap(0, -2 * s, k2);
        tap(s, s, k11); tap(-s, s, k11); tap(-s, -s, k11); tap(s, -s, k11);
        tap(s, 2 * s, k12); tap(-s, 2 * s, k12); tap(-s, -2 * s, k12); tap(s, -2 * s, k12);
        tap(2 * s, s, k12); tap(-2 * s, s, k12); tap(-2 * s, -s, k12); tap(2 * s, -s, k12);
        tap(2 * s, 2 * s, k22); tap(-2 * s, 2 * s, k22); tap(-2 * s, -2 * s, k22); tap(2 * s, -2 * s, k22);
    }
};

// atrous.slang:162-227.
__global__ __launch_bounds__(256) void k_svgf_atrous(SvgfAtrousArgs a) {
    int x, y;
    if (!svgf_pixel(a.t, x, y)) return;
    const size_t i = (size_t)y * a.t.W + x;
    SvgfTap t;
    t.v = a.v;
    t.W = a.t.W;
    t.flags = a.flags;
    t.iteration = a.iteration;
    t.step = a.step;
    t.geo = a.geo;
    t.in = a.in;
    t.x = x;
    t.y = y;
    const float4 g = a.geo[i];
    t.n = mk(g.x, g.y, g.z);
    const float4 dep = a.depth[i];
    t.z = dep.x;
    t.dzx = dep.z;
    t.dzy = dep.w;
    t.sum_w = 1;
    t.sum = a.in[i];
    t.l = lum(mk(t.sum.x, t.sum.y, t.sum.z));
    // compute_sigma_luminance (atrous.slang:47-62): the centre takes kernel[1][1]
    float s = t.sum.w * (1.0f / 16.0f);
    for (int yy = -1; yy <= 1; yy++)
        for (int xx = -1; xx <= 1; xx++) {
            if (xx == 0 && yy == 0) continue;
            if (!svgf_inside(a.v, x + xx, y + yy)) continue;
            const float k = (xx == 0 || yy == 0) ? (1.0f / 8.0f) : (1.0f / 16.0f);
            s += a.in[(size_t)(y + yy) * a.t.W + (x + xx)].w * k;
        }
    t.sigma_l = sqrtf(fmax_(s, 0.0f)) * a.sigma_boost;
    if (!isinf(t.z)) {
        switch (a.filter_type) {
            case ST_FILTER_ATROUS: t.atrous5(); break;
            default:
            case ST_FILTER_BOX3: t.box(1); break;
            case ST_FILTER_BOX5: t.box(2); break;
            case ST_FILTER_SUBSAMPLED: t.subsampled(); break;
            case ST_FILTER_BOX3_SUBSAMPLED:
                if (a.step == 1) t.box(1); else t.subsampled();
                break;
            case ST_FILTER_BOX5_SUBSAMPLED:
                if (a.step == 1) t.box(2); else t.subsampled();
                break;
        }
    }
    const float inv = 1 / t.sum_w;
    a.out[i] = make_float4(t.sum.x * inv, t.sum.y * inv, t.sum.z * inv, t.sum.w * (inv * inv));
}

// copy_rgb (atrous.slang:229-235): the filtered colour becomes the history colour.
__global__ __launch_bounds__(256) void k_svgf_copy_rgb(SvgfTiles t, const float4* filtered, float4* color) {
    int x, y;
    if (!svgf_pixel(t, x, y)) return;
    const size_t i = (size_t)y * t.W + x;
    const float4 f = filtered[i];
    const float4 c = color[i];
    color[i] = make_float4(f.x, f.y, f.z, c.w);
}

}  // namespace st
