/* svgf_oracle.c — CPU restatement of the reference's SVGF denoiser kernels,
 * used only as the checker of the HIP denoiser in tests (never shipped, never
 * measured).  One function per kernel, each over the whole image, in the
 * reference's operation order:
 *   or_svgf_accumulate  src/Shaders/kernels/svgf/temporal_accumulation.slang:34-128
 *   or_svgf_variance    src/Shaders/kernels/svgf/estimate_variance.slang:23-74
 *   or_svgf_atrous      src/Shaders/kernels/svgf/atrous.slang:36-227
 *   or_svgf_copy_rgb    src/Shaders/kernels/svgf/atrous.slang:229-235
 * The control flow around them (Denoiser::denoise, src/App/Denoiser.cpp:113-255)
 * is restated in oracle.py.  Images are row-major W*H; colours RGBA32F,
 * moments RG32F, visibility 2 x u32, depth RGBA32F {depth, previous depth,
 * derivative xy}.  getViewIndex is 0 for every pixel
 * (src/Shaders/common/scene.hlsli:109-113), so one view bounds the taps.
 * Parity: pinned by closed forms (running means and moments, exact shifts,
 * the history cap, rejection thresholds, tap patterns as delta responses) in
 * tests/test_oracle_svgf.py. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pt_oracle.h"

#define INVALID_INSTANCE 0xFFFFu

typedef struct { float x, y, z, w; } f4;

static inline float sat(float a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }
static inline float lum3(float x, float y, float z) { return x * 0.2126f + y * 0.7152f + z * 0.0722f; }
static inline float lerpf(float a, float b, float t) { return a + (b - a) * t; }
static inline int finite4(f4 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z) && isfinite(c.w); }
static inline f4 ld4(const float* img, size_t i) { f4 c = {img[4 * i], img[4 * i + 1], img[4 * i + 2], img[4 * i + 3]}; return c; }
static inline void st4(float* img, size_t i, f4 c) { img[4 * i] = c.x; img[4 * i + 1] = c.y; img[4 * i + 2] = c.z; img[4 * i + 3] = c.w; }

/* VisibilityData::normal(): octahedral snorm16 (common/normal_packing). */
static void unpack_n(uint32_t p, float n[3]) { or_unpack_normal(p, n); }
static inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static inline int inside(const st_view* v, int x, int y) {
    return x >= v->image_min[0] && y >= v->image_min[1] && x < v->image_max[0] && y < v->image_max[1];
}
/* mInstanceIndexMap[i]; a null map is the identity, indices past the map give 0xFFFF. */
static inline uint32_t map_instance(const uint32_t* map, uint32_t count, uint32_t i) {
    if (!map) return i;
    return i < count ? map[i] : INVALID_INSTANCE;
}

/* cos(degrees(2)) of the normal test, temporal_accumulation.slang:60: degrees(2)
 * is 114.59 (a radian argument to cos), so the threshold is cos(114.59 rad). */
float or_svgf_normal_threshold(void) { return cosf(2.0f * (180.0f / 3.14159265358979f)); }

void or_svgf_accumulate(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* map,
                        uint32_t map_count, const uint32_t* vis, const uint32_t* prev_vis, const float* depth,
                        const float* prev_depth, const float* prev_uv, const float* input, const float* albedo,
                        const float* prev_color, const float* prev_moments, float* color, float* moments) {
    const int reproj = (p->flags & ST_DENOISE_REPROJECTION) != 0, demod = (p->flags & ST_DENOISE_DEMODULATE_ALBEDO) != 0;
    const int check_n = (p->flags & ST_DENOISE_CHECK_NORMAL) != 0, check_z = (p->flags & ST_DENOISE_CHECK_DEPTH) != 0;
    const int hyst = (p->flags & ST_DENOISE_SHADOW_HYSTERESIS) != 0;
    const float cos_thr = or_svgf_normal_threshold();
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            f4 cp = {0, 0, 0, 0};
            float mp0 = 0, mp1 = 0, sum_w = 0;
            if (reproj) {
                const uint32_t inst = vis[2 * i] & 0xFFFFu;
                if (inst != INVALID_INSTANCE) {
                    const float ex = (float)(view->image_max[0] - view->image_min[0]);
                    const float ey = (float)(view->image_max[1] - view->image_min[1]);
                    const float px = (float)view->image_min[0] + prev_uv[2 * i] * ex - 0.5f;
                    const float py = (float)view->image_min[1] + prev_uv[2 * i + 1] * ey - 0.5f;
                    /* positions beyond +-2^24 (or NaN) lie outside every image: no taps
                     * (keeps the integer conversion defined) */
                    const int far = !(fabsf(px) < 16777216.0f && fabsf(py) < 16777216.0f);
                    const int ix = far ? 0 : (int)px, iy = far ? 0 : (int)py;  /* int2(): truncation */
                    const float wx = px - floorf(px), wy = py - floorf(py);  /* frac() */
                    float n[3];
                    unpack_n(vis[2 * i + 1], n);
                    const float dd = sqrtf(depth[4 * i + 2] * depth[4 * i + 2] + depth[4 * i + 3] * depth[4 * i + 3]);
                    for (int yy = 0; yy <= 1 && !far; yy++)
                        for (int xx = 0; xx <= 1; xx++) {
                            const int qx = ix + xx, qy = iy + yy;
                            if (!inside(view, qx, qy)) continue;
                            if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;  /* outside the image */
                            const size_t j = (size_t)qy * W + qx;
                            if (map_instance(map, map_count, inst) != (prev_vis[2 * j] & 0xFFFFu)) continue;
                            if (check_n) {
                                float pn[3];
                                unpack_n(prev_vis[2 * j + 1], pn);
                                if (dot3(n, pn) < cos_thr) continue;
                            }
                            if (check_z && fabsf(depth[4 * i + 1] - prev_depth[4 * j]) >= 2 * dd) continue;
                            const f4 c = ld4(prev_color, j);
                            if (c.w <= 0 || !finite4(c)) continue;
                            const float wc = (xx == 0 ? (1 - wx) : wx) * (yy == 0 ? (1 - wy) : wy);
                            cp.x += c.x * wc; cp.y += c.y * wc; cp.z += c.z * wc; cp.w += c.w * wc;
                            mp0 += prev_moments[2 * j] * wc;
                            mp1 += prev_moments[2 * j + 1] * wc;
                            sum_w += wc;
                        }
                }
            } else {
                cp = ld4(prev_color, i);
                if (!isfinite(cp.x) || !isfinite(cp.y) || !isfinite(cp.z)) {
                    cp.x = cp.y = cp.z = cp.w = 0;
                } else {
                    mp0 = prev_moments[2 * i];
                    mp1 = prev_moments[2 * i + 1];
                }
                sum_w = 1;
            }
            f4 cc = ld4(input, i);
            if (demod) {
                cc.x /= (1e-2f + albedo[4 * i]);
                cc.y /= (1e-2f + albedo[4 * i + 1]);
                cc.z /= (1e-2f + albedo[4 * i + 2]);
            }
            if (!isfinite(cc.x) || !isfinite(cc.y) || !isfinite(cc.z)) cc.x = cc.y = cc.z = cc.w = 0;
            if (!isfinite(mp0) || !isfinite(mp1)) mp0 = mp1 = 0;
            const float l = lum3(cc.x, cc.y, cc.z);
            if (sum_w > 0 && cp.w > 0) {
                const float inv = 1 / sum_w;
                cp.x *= inv; cp.y *= inv; cp.z *= inv; cp.w *= inv;
                mp0 *= inv; mp1 *= inv;
                float n = cp.w + cc.w;
                if (p->history_limit > 0 && n > (float)p->history_limit) n = (float)p->history_limit;
                float alpha = sat(cc.w / n);
                if (hyst) {
                    const float lp = lum3(cp.x, cp.y, cp.z);
                    const float den = l + lp;
                    const float q = sat((den == 0 ? 0.0f : (lp - l) / den) - p->shadow_preserve_offset);
                    const float alpha_l = q * q * p->shadow_preserve_scale;
                    alpha = 1 - (1 - alpha) * sat(1 - alpha_l);
                }
                const f4 o = {lerpf(cp.x, cc.x, alpha), lerpf(cp.y, cc.y, alpha), lerpf(cp.z, cc.z, alpha), n};
                st4(color, i, o);
                moments[2 * i] = lerpf(mp0, l, alpha);
                moments[2 * i + 1] = lerpf(mp1, l * l, alpha);
            } else {
                st4(color, i, cc);
                moments[2 * i] = l;
                moments[2 * i + 1] = l * l;
            }
        }
}

void or_svgf_variance(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* map,
                      uint32_t map_count, const uint32_t* vis, const float* depth, const float* color,
                      const float* moments, float* out) {
    const int check_n = (p->flags & ST_DENOISE_CHECK_NORMAL) != 0, check_z = (p->flags & ST_DENOISE_CHECK_DEPTH) != 0;
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            f4 c = ld4(color, i);
            float m0 = moments[2 * i], m1 = moments[2 * i + 1];
            const uint32_t inst = vis[2 * i] & 0xFFFFu;
            const float histlen = c.w;
            if (inst == INVALID_INSTANCE || histlen >= (float)p->history_limit) {
                const f4 o = {c.x, c.y, c.z, fabsf(m1 - m0 * m0)};
                st4(out, i, o);
                continue;
            }
            float n[3];
            unpack_n(vis[2 * i + 1], n);
            float sum_w = 1;
            const int r = histlen > 1 ? 2 : 3;
            for (int yy = -r; yy <= r; yy++)
                for (int xx = -r; xx <= r; xx++) {
                    if (xx == 0 && yy == 0) continue;
                    const int qx = (int)x + xx, qy = (int)y + yy;
                    if (!inside(view, qx, qy)) continue;
                    if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;
                    const size_t j = (size_t)qy * W + qx;
                    if (map_instance(map, map_count, inst) != (vis[2 * j] & 0xFFFFu)) continue;
                    float w_z = 0, w_n = 1;
                    if (check_z) {
                        const float gx = depth[4 * i + 2] * (float)xx, gy = depth[4 * i + 3] * (float)yy;
                        w_z = fabsf(depth[4 * j] - depth[4 * i]) / (sqrtf(gx * gx + gy * gy) + 1e-2f);
                    }
                    if (check_n) {
                        float pn[3];
                        unpack_n(vis[2 * j + 1], pn);
                        w_n = powf(sat(dot3(pn, n)), 128.0f);
                    }
                    const float w = expf(-w_z) * w_n;
                    if (!isfinite(w)) continue;
                    m0 += moments[2 * j] * w;
                    m1 += moments[2 * j + 1] * w;
                    c.x += color[4 * j] * w;
                    c.y += color[4 * j + 1] * w;
                    c.z += color[4 * j + 2] * w;
                    sum_w += w;
                }
            sum_w = 1 / sum_w;
            m0 *= sum_w; m1 *= sum_w;
            c.x *= sum_w; c.y *= sum_w; c.z *= sum_w;
            float v = fabsf(m1 - m0 * m0);
            if (p->variance_boost_length > 0) {
                const float b = p->variance_boost_length / (1 + c.w);
                v *= b > 1 ? b : 1;
            }
            const f4 o = {c.x, c.y, c.z, v};
            st4(out, i, o);
        }
}

/* TapData of atrous.slang:31-89. */
typedef struct {
    const st_view* view;
    const float* in;
    const uint32_t* vis;
    const float* depth;
    uint32_t W, H;
    int check_n, check_z, step;
    int x, y;
    float n[3], z, dz[2], l, sigma_l;
    f4 sum;
    float sum_w;
} tap_t;

static void tap(tap_t* t, int ox, int oy, float kw) {
    const int qx = t->x + ox, qy = t->y + oy;
    if (!inside(t->view, qx, qy)) return;
    if (qx < 0 || qy < 0 || qx >= (int)t->W || qy >= (int)t->H) return;
    const size_t j = (size_t)qy * t->W + qx;
    const f4 c = ld4(t->in, j);
    const float lp = lum3(c.x, c.y, c.z);
    const float w_l = fabsf(lp - t->l) / fmaxf(t->sigma_l, 1e-10f);
    float w_z = 0, w_n = 1;
    if (t->check_z) {
        /* offset is already scaled by the step; the reference scales it again */
        const float gx = t->dz[0] * (float)(ox * t->step), gy = t->dz[1] * (float)(oy * t->step);
        w_z = fabsf(t->depth[4 * j] - t->z) / (sqrtf(gx * gx + gy * gy) + 1e-2f);
    }
    if (t->check_n) {
        float pn[3];
        unpack_n(t->vis[2 * j + 1], pn);
        w_n = powf(fmaxf(0.0f, dot3(pn, t->n)), 256.0f);
    }
    const float w = expf(-(w_l * w_l) - w_z) * kw * w_n;
    if (!isfinite(w)) return;
    t->sum.x += c.x * w; t->sum.y += c.y * w; t->sum.z += c.z * w; t->sum.w += c.w * (w * w);
    t->sum_w += w;
}

static void box(tap_t* t, int r) {
    for (int yy = -r; yy <= r; yy++)
        for (int xx = -r; xx <= r; xx++)
            if (xx != 0 || yy != 0) tap(t, xx * t->step, yy * t->step, 1.0f);
}

static void subsampled(tap_t* t, uint32_t iteration) {
    const int s = t->step;
    if ((iteration & 1) == 0) {
        tap(t, -2 * s, 0, 1.0f);
        tap(t, 2 * s, 0, 1.0f);
    } else {
        tap(t, 0, -2 * s, 1.0f);
        tap(t, 0, 2 * s, 1.0f);
    }
    tap(t, -s, s, 1.0f);
    tap(t, s, s, 1.0f);
    tap(t, -s, -s, 1.0f);
    tap(t, s, -s, 1.0f);
}

/* The 5x5 B3-spline taps of atrous(), atrous.slang:128-158, in its order. */
static void atrous5(tap_t* t) {
    static const int off[24][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {2, 0}, {0, 2}, {-2, 0}, {0, -2},
                                   {1, 1}, {-1, 1}, {-1, -1}, {1, -1}, {1, 2}, {-1, 2}, {-1, -2}, {1, -2},
                                   {2, 1}, {-2, 1}, {-2, -1}, {2, -1}, {2, 2}, {-2, 2}, {-2, -2}, {2, -2}};
    static const float kw[24] = {2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f, 1.0f / 6.0f, 1.0f / 6.0f,
                                 1.0f / 6.0f, 1.0f / 6.0f, 4.0f / 9.0f, 4.0f / 9.0f, 4.0f / 9.0f, 4.0f / 9.0f,
                                 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                                 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
    for (int k = 0; k < 24; k++) tap(t, off[k][0] * t->step, off[k][1] * t->step, kw[k]);
}

void or_svgf_atrous(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* vis,
                    const float* depth, uint32_t iteration, const float* in, float* out) {
    tap_t t;
    t.view = view; t.in = in; t.vis = vis; t.depth = depth; t.W = W; t.H = H;
    t.check_n = (p->flags & ST_DENOISE_CHECK_NORMAL) != 0;
    t.check_z = (p->flags & ST_DENOISE_CHECK_DEPTH) != 0;
    t.step = 1 << iteration;
    static const float kern[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t i = (size_t)y * W + x;
            t.x = (int)x; t.y = (int)y;
            unpack_n(vis[2 * i + 1], t.n);
            t.z = depth[4 * i];
            t.dz[0] = depth[4 * i + 2];
            t.dz[1] = depth[4 * i + 3];
            t.sum_w = 1;
            t.sum = ld4(in, i);
            t.l = lum3(t.sum.x, t.sum.y, t.sum.z);
            /* compute_sigma_luminance: the centre takes kernel[1][1] */
            float s = t.sum.w * kern[1][1];
            for (int yy = -1; yy <= 1; yy++)
                for (int xx = -1; xx <= 1; xx++) {
                    if (xx == 0 && yy == 0) continue;
                    const int qx = (int)x + xx, qy = (int)y + yy;
                    if (!inside(view, qx, qy)) continue;
                    if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;
                    s += in[4 * ((size_t)qy * W + qx) + 3] * kern[abs(xx)][abs(yy)];
                }
            t.sigma_l = sqrtf(fmaxf(s, 0.0f)) * p->sigma_luminance_boost;
            if (!isinf(t.z)) {
                switch (p->filter_type) {
                    case ST_FILTER_ATROUS: atrous5(&t); break;
                    default:
                    case ST_FILTER_BOX3: box(&t, 1); break;
                    case ST_FILTER_BOX5: box(&t, 2); break;
                    case ST_FILTER_SUBSAMPLED: subsampled(&t, iteration); break;
                    case ST_FILTER_BOX3_SUBSAMPLED:
                        if (t.step == 1) box(&t, 1); else subsampled(&t, iteration);
                        break;
                    case ST_FILTER_BOX5_SUBSAMPLED:
                        if (t.step == 1) box(&t, 2); else subsampled(&t, iteration);
                        break;
                }
            }
            const float inv = 1 / t.sum_w;
            const f4 o = {t.sum.x * inv, t.sum.y * inv, t.sum.z * inv, t.sum.w * (inv * inv)};
            st4(out, i, o);
        }
}

void or_svgf_copy_rgb(uint32_t W, uint32_t H, const float* filtered, float* color) {
    for (size_t i = 0; i < (size_t)W * H; i++) {
        color[4 * i] = filtered[4 * i];
        color[4 * i + 1] = filtered[4 * i + 1];
        color[4 * i + 2] = filtered[4 * i + 2];
    }
}
