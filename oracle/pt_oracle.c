/*
 * pt_oracle.c — CPU ORACLE for the path-tracing hot path (test infrastructure).
 *
 * Plain C99 restatement of Stratum2's TestRenderer (multi-dispatch, deferred
 * shadow rays, fixed-point NEE accumulation), built with -ffp-contract=off so
 * every float operation is a single IEEE-754 operation in the order the
 * reference writes it.  Each function cites the reference lines it restates.
 * Only tests, the smoke entry and the bench CPU-baseline leg may load it.
 *
 * Parity status of each part (the check that pins it is in tests/):
 *   rng / hash / packing / f16 / rayOffset  pinned: hand-checked constants and
 *                                           closed-form properties (tests/test_oracle_units.py)
 *   transcendentals (sin, cos, atan2, acos, pinned: float64 error bounds, 1-3.5 ulp
 *   log, exp, pow)                          (tests/test_oracle_math.py)
 *   watertight triangle test + BVH          pinned: FP64 brute force over all triangles
 *   camera                                  pinned: closed-form pinhole invariants
 *   BSDFs                                   pinned: furnace / reciprocity / pdf integration; Disney
 *                                           eval value by float64 GGX / Burley closed forms
 *                                           (tests/test_oracle_disney.py)
 *   env mip-warp sampling                   pinned: sample pdf == SampleTexelPdf, integrates to 1
 *   env alias-table sampling                pinned: table mass == float64 texel weights, chi-square
 *                                           of sampled texels, in-texel uniformity
 *                                           (tests/test_oracle_alias.py)
 *   bounce loop                             pinned: closed-form furnace scene (white Lambert
 *                                           sphere interior), emission-only analytic image
 *   Russian roulette from path length k     pinned: identity beyond the maximum depth; same
 *                                           expectation as roulette everywhere and never
 *                                           (tests/test_oracle_headline_modes.py)
 *   ReSTIR DI RIS                           pinned: unbiased against one-sample NEE (test_oracle_units.py)
 *   ReSTIR DI screen-space reuse            pinned: same expectation as RIS and one-sample NEE;
 *                                           independent of thread count (test_oracle_headline_modes.py)
 *   NanoVDB read path, delta / ratio        pinned: the reference's PNanoVDB compiled into oracle/_ref,
 *   tracking, Henyey-Greenstein             chord transmittance, furnace, HG pdf (test_oracle_media.py)
 *   light tracing, adjoint BSDF             pinned: estimates the camera-path image, adjoint
 *                                           reciprocity (test_oracle_light.py)
 *   bidirectional mode (light vertex cache) pinned: with fp32 MIS estimates the path tracer's
 *                                           image; the fp16 overflow (test_oracle_bdpt.py)
 *   tonemapper                              pinned: operator closed forms (test_oracle_units.py)
 *   driver BVH tie-breaking, f32tof16 and   parity unpinned (Vulkan driver / Slang codegen,
 *   bilinear weights of the reference        not runnable here)
 */
#include "pt_oracle.h"

#include <math.h>
#include <float.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The arithmetic type.  The default build computes in float, with the
 * renderer's operations in the renderer's order (the parity build).  Built
 * with -DOR_FP64 the same algorithm computes in double, with libm's double
 * transcendentals: the high-precision reference for converged images.  Data
 * in its stored formats (scene buffers, textures, alias tables, fp16 and
 * snorm packing, fixed-point sums, the caller's arrays) stays in those
 * formats in both builds. */
#ifdef OR_FP64
typedef double real;
#define sqrtr sqrt
#define fabsr fabs
#define floorr floor
#define fminr fmin
#define fmaxr fmax
#define tanr tan
#define powr pow
#define nextafterr nextafter
#define truncr trunc
#define ldexpr ldexp
#define OR_PI 3.14159265358979323846
#else
typedef float real;
#define sqrtr sqrtf
#define fabsr fabsf
#define floorr floorf
#define fminr fminf
#define fmaxr fmaxf
#define tanr tanf
#define powr powf
#define nextafterr nextafterf
#define truncr truncf
#define ldexpr ldexpf
#define OR_PI 3.14159265358979323846f
#endif
#define OR_1_PI ((real)1 / OR_PI)
#define OR_Q 16384 /* gRadianceQuantizationFactor, testrenderer.slang:25 */

static _Thread_local char g_err[256];
const char* or_last_error(void) { return g_err; }
static int fail(const char* m) {
    snprintf(g_err, sizeof g_err, "%s", m);
    return 1;
}

/* ------------------------------------------------------------------ math */
typedef struct { real x, y, z; } f3;
static inline f3 v3(real x, real y, real z) { f3 r = {x, y, z}; return r; }
static inline f3 add(f3 a, f3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub(f3 a, f3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 mul(f3 a, f3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline f3 mulf(f3 a, real s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline f3 divv(f3 a, f3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline f3 divf(f3 a, real s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline f3 neg(f3 a) { return v3(-a.x, -a.y, -a.z); }
static inline real dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real length3(f3 a) { return sqrtr(dot(a, a)); }
static inline f3 normalize3(f3 a) { return divf(a, sqrtr(dot(a, a))); }
static inline real fget(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline int any_gt0(f3 a) { return a.x > 0 || a.y > 0 || a.z > 0; }
static inline int all_le0(f3 a) { return a.x <= 0 && a.y <= 0 && a.z <= 0; }
static inline real minf(real a, real b) { return a < b ? a : b; }
static inline real maxf(real a, real b) { return a > b ? a : b; }
static inline real saturate(real a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }
static inline real max3f(f3 a) { return maxf(maxf(a.x, a.y), a.z); }
static inline real pow2(real a) { return a * a; }
/* max(0, x) of the reference's GPU code: NaN becomes 0 */
static inline real clamp0(real x) { return x > 0 ? x : 0.0f; }
/* luminance, src/Shaders/compat/common.h:51 */
static inline real luminance(f3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

static inline uint32_t asuint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float asfloat(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------ rng.hlsli */
/* xxhash32, src/Shaders/common/rng.hlsli:5-13 */
uint32_t or_xxhash32(uint32_t p) {
    const uint32_t P2 = 2246822519U, P3 = 3266489917U, P4 = 668265263U, P5 = 374761393U;
    uint32_t h = p + P5;
    h = P4 * ((h << 17) | (h >> 15));
    h = P2 * (h ^ (h >> 15));
    h = P3 * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
/* pcg, rng.hlsli:17-21 */
uint32_t or_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
/* pcg4d, rng.hlsli:23-35 */
void or_pcg4d(const uint32_t in[4], uint32_t o[4]) {
    uint32_t x = in[0] * 1664525u + 1013904223u, y = in[1] * 1664525u + 1013904223u;
    uint32_t z = in[2] * 1664525u + 1013904223u, w = in[3] * 1664525u + 1013904223u;
    x += y * w; y += z * x; z += x * y; w += y * z;
    x ^= x >> 16; y ^= y >> 16; z ^= z >> 16; w ^= w >> 16;
    x += y * w; y += z * x; z += x * y; w += y * z;
    o[0] = x; o[1] = y; o[2] = z; o[3] = w;
}
/* RandomSampler, rng.hlsli:39-60: state (pixel.x, pixel.y, seed, offset). */
typedef struct { uint32_t s[4]; } rng_t;
static void rng_init(rng_t* r, uint32_t seed, uint32_t px, uint32_t py, uint32_t offset) {
    r->s[0] = px; r->s[1] = py; r->s[2] = seed; r->s[3] = offset;
}
static void rng_next(rng_t* r, uint32_t o[4]) { r->s[3]++; or_pcg4d(r->s, o); }
static void rng_nextf(rng_t* r, real o[4]) {
    uint32_t v[4];
    rng_next(r, v);
    for (int i = 0; i < 4; i++) o[i] = asfloat(0x3f800000u | (v[i] >> 9)) - 1.0f;
}

/* ------------------------------------------------------------- packing */
/* D3DX_FLOAT_to_INT / D3DX_INT_to_FLOAT / SaturateSigned,
 * src/Shaders/common/D3DX_DXGIFormatConvert.inl:290-322 */
static int32_t snorm16(real v) {
    if (v != v) v = 0;
    v = minf(maxf(v, -1.0f), 1.0f);
    return (int32_t)truncr(v * 32767.0f + (v >= 0 ? 0.5f : -0.5f));
}
static real unsnorm16(int32_t i) { return maxf((real)i / 32767.0f, -1.0f); }
/* packNormal2 / packNormal, src/Shaders/common/common.hlsli:3-19 */
static uint32_t pack_normal(f3 v) {
    real r = 1.0f / (fabsr(v.x) + fabsr(v.y) + fabsr(v.z));
    real px = v.x * r, py = v.y * r;
    if (v.z <= 0) {
        real nx = (1 - fabsr(py)) * (px >= 0 ? 1.0f : -1.0f);
        real ny = (1 - fabsr(px)) * (py >= 0 ? 1.0f : -1.0f);
        px = nx; py = ny;
    }
    return ((uint32_t)snorm16(px) & 0xffffu) | ((uint32_t)snorm16(py) << 16);
}
/* unpackNormal2 / unpackNormal, common.hlsli:8-22 */
static f3 unpack_normal(uint32_t p) {
    real x = unsnorm16((int32_t)(p << 16) >> 16), y = unsnorm16((int32_t)(p & 0xffff0000u) >> 16);
    f3 v = v3(x, y, 1 - (fabsr(x) + fabsr(y)));
    if (v.z < 0) {
        real nx = (1 - fabsr(v.y)) * (v.x >= 0 ? 1.0f : -1.0f);
        real ny = (1 - fabsr(v.x)) * (v.y >= 0 ? 1.0f : -1.0f);
        v.x = nx; v.y = ny;
    }
    return normalize3(v);
}
uint32_t or_pack_normal(const float v[3]) { return pack_normal(v3(v[0], v[1], v[2])); }
void or_unpack_normal(uint32_t p, float o[3]) { f3 v = unpack_normal(p); o[0] = v.x; o[1] = v.y; o[2] = v.z; }

/* f32tof16 / f16tof32 (testrenderer.slang:75-97), round-to-nearest-even. */
uint16_t or_f32_to_f16(float f) {
    uint32_t x = asuint(f), sign = (x >> 16) & 0x8000u, ex = (x >> 23) & 0xffu, m = x & 0x7fffffu;
    if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (m ? 0x200u : 0));
    int e = (int)ex - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        m |= 0x800000u;
        int sh = 14 - e;
        uint32_t hm = m >> sh, rem = m & ((1u << sh) - 1), half = 1u << (sh - 1);
        if (rem > half || (rem == half && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t r = sign | ((uint32_t)e << 10) | (m >> 13), rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r++;
    return (uint16_t)r;
}
float or_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        real v = (real)m * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (e == 0x1f) return asfloat(sign | 0x7f800000u | (m << 13));
    return asfloat(sign | ((e + 112) << 23) | (m << 13));
}
static real q16(real f) { return or_f16_to_f32(or_f32_to_f16(f)); }

/* unorm8 material fields, src/Shaders/compat/material_data.h:43-64 */
static real un8(uint32_t w, int shift) { return (real)((w >> shift) & 0xffu) / 255.0f; }

/* rayOffset, src/Shaders/common/intersection.hlsli:17-32 */
static real roff1(real p, real n) {
    const real origin = 1.0f / 32.0f, fs = 1.0f / 65536.0f;
    int32_t of = (int32_t)(256.0f * n);
    if (fabsr(p) < origin) return p + fs * n;
    return asfloat((uint32_t)((int32_t)asuint(p) + ((p < 0.0f) ? -of : of)));
}
static f3 ray_offset(f3 p, f3 n) { return v3(roff1(p.x, n.x), roff1(p.y, n.y), roff1(p.z, n.z)); }
static f3 ray_offset_dir(f3 p, f3 n, f3 d) { return ray_offset(p, dot(n, d) < 0 ? neg(n) : n); }
void or_ray_offset(const float p[3], const float n[3], float o[3]) {
    f3 r = ray_offset(v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]));
    o[0] = r.x; o[1] = r.y; o[2] = r.z;
}

/* makeOrthonormal, common.hlsli:24-34 (rows r0, r1, r2 = N) */
/* ------------------------------------------ deterministic transcendentals */
/* The renderer's transcendental functions as fixed sequences of float
 * operations (round to nearest, no contraction), the same sequences the GPU
 * library runs, so that both sample the same directions and distances bit for
 * bit.  Method: Cody-Waite argument reduction and the single-precision
 * minimax polynomials of the Cephes library (S. L. Moshier, "Methods and
 * Programs for Mathematical Functions", 1989: sinf.c, atanf.c, asinf.c,
 * logf.c, expf.c), evaluated in float.  The reference's own precision
 * (Slang -> SPIR-V -> driver) is unpinned; these are pinned against float64
 * by their error bounds in tests/test_oracle_math.py. */
#ifndef OR_FP64
static real om_poly_sin(real x, real z) {
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
static real om_poly_cos(real z) {
    const real y = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z;
    return (y - 0.5f * z) + 1.0f;
}
static real om_reduce(real a, int* jo) {
    int j = (int)(a * 1.27323954473516f);
    real y = (real)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    *jo = j;
    return ((a - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}
static real om_sin(real x) {
    real sign = 1.0f, a = x;
    if (a < 0) { a = -a; sign = -1.0f; }
    int j;
    const real r = om_reduce(a, &j);
    if (j > 3) { sign = -sign; j -= 4; }
    const real z = r * r;
    const real y = (j == 1 || j == 2) ? om_poly_cos(z) : om_poly_sin(r, z);
    return sign < 0 ? -y : y;
}
static real om_cos(real x) {
    const real a = x < 0 ? -x : x;
    int j;
    const real r = om_reduce(a, &j);
    real sign = 1.0f;
    if (j > 3) { j -= 4; sign = -sign; }
    if (j > 1) sign = -sign;
    const real z = r * r;
    const real y = (j == 1 || j == 2) ? om_poly_sin(r, z) : om_poly_cos(z);
    return sign < 0 ? -y : y;
}
static real om_atan_pos(real x) {
    real y = 0.0f;
    if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -(1.0f / x); }
    else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
    const real z = x * x;
    return y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
}
/* atan2 for x != 0 (stable_atan2 handles x == 0, compat/common.h:137-140) */
static real om_atan2(real y, real x) {
    const real q = y / x;
    const real t = q < 0 ? -om_atan_pos(-q) : om_atan_pos(q);
    if (x > 0) return t;
    return y < 0 ? t - 3.14159265358979f : t + 3.14159265358979f;
}
static real om_asin_pos(real a, int* big) {
    real x, z;
    *big = a > 0.5f;
    if (*big) { z = 0.5f * (1.0f - a); x = sqrtr(z); }
    else { x = a; z = x * x; }
    return ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
            1.6666752422e-1f) * z * x + x;
}
static real om_acos(real x) {
    const real a = x < 0 ? -x : x;
    int big;
    const real s = om_asin_pos(a, &big);
    if (big) {
        const real r = s + s;
        return x < 0 ? 3.14159265358979f - r : r;
    }
    return x < 0 ? 1.5707963267948966f + s : 1.5707963267948966f - s;
}
static real om_log(real x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 16777216.0f; e = -24; }
    const uint32_t b = asuint(x);
    e += (int)((b >> 23) & 0xffu) - 126;
    real m = asfloat((b & 0x007fffffu) | 0x3f000000u);
    if (m < 0.707106781186547524f) { e -= 1; m = m + m - 1.0f; }
    else m = m - 1.0f;
    const real z = m * m;
    real y = ((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m +
                   1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m +
               3.3333331174e-1f) * m * z;
    const real fe = (real)e;
    y += -2.12194440e-4f * fe;
    y += -0.5f * z;
    return (m + y) + 0.693359375f * fe;
}
static real om_exp(real x) {
    if (x != x) return x;
    if (x > 88.7228391f) return INFINITY;
    if (x < -87.0f) return 0.0f;
    const real fz = floorr(1.44269504088896341f * x + 0.5f);
    real r = x - fz * 0.693359375f;
    r = r - fz * -2.12194440e-4f;
    const int n = (int)fz;
    const real z = r * r;
    const real p = ((((((1.9875691500e-4f * r + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r +
                      1.6666665459e-1f) * r + 5.0000001201e-1f) * z + r) + 1.0f;
    const int n1 = n / 2, n2 = n - n1;
    return (p * asfloat((uint32_t)(n1 + 127) << 23)) * asfloat((uint32_t)(n2 + 127) << 23);
}
static real om_pow(real x, real y) { return om_exp(y * om_log(x)); }
#else
/* the double build: libm's double functions */
static real om_sin(real x) { return sin(x); }
static real om_cos(real x) { return cos(x); }
static real om_atan2(real y, real x) { return atan2(y, x); }
static real om_acos(real x) { return acos(x); }
static real om_log(real x) { return log(x); }
static real om_exp(real x) { return exp(x); }
static real om_pow(real x, real y) { return pow(x, y); }
#endif
static real om_eval(int fn, real x, real y) {
    switch (fn) {
        case 0: return om_sin(x);
        case 1: return om_cos(x);
        case 2: return om_atan2(x, y);
        case 3: return om_acos(x);
        case 4: return om_log(x);
        case 5: return om_exp(x);
        case 6: return om_pow(x, y);
        default: return NAN;
    }
}
void or_math(int fn, uint32_t n, const float* x, const float* y, float* out) {
    for (uint32_t i = 0; i < n; i++) out[i] = om_eval(fn, x[i], y ? y[i] : 0.0f);
}

static void make_orthonormal(f3 N, f3 r[3]) {
    if (N.x != N.y || N.x != N.z) r[0] = v3(N.z - N.y, N.x - N.z, N.y - N.x);
    else r[0] = v3(N.z - N.y, N.x + N.z, -N.y - N.x);
    r[0] = normalize3(r[0]);
    r[1] = cross(N, r[0]);
    r[2] = N;
}
/* warps, common.hlsli:36-77 */
static void sample_uniform_triangle(real u1, real u2, real b[2]) {
    real a = sqrtr(u1);
    b[0] = 1 - a; b[1] = a * u2;
}
static f3 sample_uniform_sphere(real u1, real u2) {
    real z = u2 * 2 - 1, r = sqrtr(maxf(1 - z * z, 0)), phi = 2 * OR_PI * u1;
    return v3(r * om_cos(phi), r * om_sin(phi), z);
}
static void sample_concentric_disc(real u1, real u2, real o[2]) {
    real ux = 2 * u1 - 1, uy = 2 * u2 - 1, theta, r;
    if (ux == 0 && uy == 0) { o[0] = o[1] = 0; return; }
    if (fabsr(ux) > fabsr(uy)) { r = ux; theta = (OR_PI / 4) * (uy / ux); }
    else { r = uy; theta = OR_PI / 2 - (OR_PI / 4) * (ux / uy); }
    o[0] = r * om_cos(theta); o[1] = r * om_sin(theta);
}
static f3 sample_cos_hemisphere(real u1, real u2) {
    real xy[2];
    sample_concentric_disc(u1, u2, xy);
    return v3(xy[0], xy[1], sqrtr(maxf(0, 1 - (xy[0] * xy[0] + xy[1] * xy[1]))));
}
static real cos_hemisphere_pdfw(real c) { return maxf(c, 0.0f) / OR_PI; }
static real concentric_disc_pdfa(void) { return 1.0f / OR_PI; }
static real mis1(real a) { return a * a; }
/* mis, testrenderer.slang:284-290, with its limit where a pdf overflowed
 * (inf / inf: 1 for the infinite pdf, 0 against it, 1/2 for two) instead of
 * the NaN the reference's quotient gives */
static real mis2(real a, real b) {
    const real a2 = mis1(a), b2 = mis1(b);
    const real w = a2 / (a2 + b2);
    return w == w ? w : (a2 > b2 ? 1.0f : (a2 < b2 ? 0.0f : 0.5f));
}

/* stableAtan2 / cartesianToSphericalUv / sphericalUvToCartesian, compat/common.h:137-150 */
static real stable_atan2(real y, real x) {
    return x == 0.0f ? (y == 0 ? 0 : (y < 0 ? -OR_PI / 2 : OR_PI / 2)) : om_atan2(y, x);
}
static void cart_to_uv(f3 v, real uv[2]) {
    real th = stable_atan2(v.z, v.x);
    uv[0] = th * OR_1_PI * 0.5f + 0.5f;
    uv[1] = om_acos(minf(maxf(v.y, -1.0f), 1.0f)) * OR_1_PI;
}
static f3 uv_to_cart(real u, real v) {
    u = u * 2 - 1;
    u *= OR_PI; v *= OR_PI;
    real sp = om_sin(v);
    return v3(sp * om_cos(u), om_cos(v), sp * om_sin(u));
}

/* ------------------------------------------------------------- scene */
typedef struct { real lo[3], hi[3]; uint32_t a, n; } bnode; /* n>0: leaf [a, a+n); else children a, a+1 */
typedef struct { int w, h; float* rgba; } level_t;

struct or_scene {
    uint32_t ninst;
    st_instance* inst;
    st_transform* xf;
    uint32_t nmatw;
    uint32_t* mat;
    uint32_t nvi;
    st_mesh_vertex_info* vi;
    uint32_t nvb;
    uint8_t** vb;
    uint64_t* vbsz;
    /* environment */
    int has_env;
    uint32_t env_addr;
    f3 env_emission;
    int env_has_image;
    int nlevels;
    level_t lv[16];
    /* alias table over luminance * sin(theta) of level 0 (ST_FLAG_ENV_ALIAS) */
    float* alias_prob;
    uint32_t* alias_idx;
    float* alias_pdf;
    real alias_jmax; /* largest in-texel jitter: 1 - 2^(ceil(log2(max(W, H))) - 21) */
    /* lights: emissive instances in instance order (Scene.cpp:487-491) */
    uint32_t nlights;
    uint32_t* light_inst;
    real sphere[4];
    /* flattened world-space triangles */
    uint32_t ntris;
    float* tv;     /* 9 floats each */
    uint32_t* tid; /* instance | primitive << 16 */
    bnode* nodes;
    uint32_t nnodes;
    uint32_t* order;
    /* media: volume instances (InstanceType::eVolume) and their NanoVDB grids */
    uint32_t nvolinst;
    uint32_t volinst[8];
    st_transform* inv_xf;          /* inverse transforms of all instances */
    uint32_t nvols;
    uint8_t* vol[8];
    uint64_t volsz[8];
    uint32_t nvolinfo;
    st_volume_info* volinfo;
    /* differentiable rendering: the per-triangle albedo (3 floats per
     * flattened triangle, NULL: none) that replaces the base colour of
     * Lambertian camera-path vertices, each instance's first flattened
     * triangle, and "weight" copies of the grids (NULL: the grid itself): the
     * tracking decisions read vol[], the path weights read volw[] */
    float* tri_alb;
    uint32_t* tri_base;
    uint8_t* volw[8];
    /* procedural sphere instances (SphereInstanceData, compat/scene.h:60-69) */
    uint32_t nsph;
    uint32_t* sph_inst;
    /* material images (Material::store's image indices), each with its mip
     * chain; ntex = number of images, tex[i].n = 0 when no material uses it */
    uint32_t ntex;
    int has_alpha; /* a mesh material carries an alpha mask */
    struct tex_s { int n; level_t lv[16]; }* tex;
};

static f3 xf_point(const st_transform* t, f3 v) {
    return v3(t->m[0][0] * v.x + t->m[0][1] * v.y + t->m[0][2] * v.z + t->m[0][3],
              t->m[1][0] * v.x + t->m[1][1] * v.y + t->m[1][2] * v.z + t->m[1][3],
              t->m[2][0] * v.x + t->m[2][1] * v.y + t->m[2][2] * v.z + t->m[2][3]);
}
static f3 xf_vector(const st_transform* t, f3 v) {
    return v3(t->m[0][0] * v.x + t->m[0][1] * v.y + t->m[0][2] * v.z,
              t->m[1][0] * v.x + t->m[1][1] * v.y + t->m[1][2] * v.z,
              t->m[2][0] * v.x + t->m[2][1] * v.y + t->m[2][2] * v.z);
}

static uint32_t vi_field(uint32_t w, int start, int len) { return (w >> start) & ((len == 32) ? 0xffffffffu : ((1u << len) - 1)); }

/* LoadTriangleIndices, src/Shaders/common/scene.hlsli:43-61 */
static int load_tri(const or_scene* s, const st_mesh_vertex_info* vi, uint32_t prim, uint32_t tri[3]) {
    uint32_t ib = vi_field(vi->packed_buffer_indices[0], 0, 16), stride = vi_field(vi->packed_strides, 0, 8);
    if (ib >= s->nvb) return 0;
    uint64_t off = (uint64_t)vi->packed_offsets[0] + (uint64_t)prim * 3 * stride;
    if (off + 3 * stride > s->vbsz[ib]) return 0;
    const uint8_t* b = s->vb[ib] + off;
    for (int i = 0; i < 3; i++) {
        if (stride == 2) { uint16_t v; memcpy(&v, b + 2 * i, 2); tri[i] = v; }
        else { uint32_t v; memcpy(&v, b + 4 * i, 4); tri[i] = v; }
    }
    return 1;
}
static int load_f3(const or_scene* s, uint32_t buf, uint32_t offset, uint32_t stride, uint32_t idx, f3* o) {
    if (buf >= s->nvb) return 0;
    uint64_t off = (uint64_t)offset + (uint64_t)stride * idx;
    if (off + 12 > s->vbsz[buf]) return 0;
    float f[3];
    memcpy(f, s->vb[buf] + off, 12);
    *o = v3(f[0], f[1], f[2]);
    return 1;
}
static int load_f2(const or_scene* s, uint32_t buf, uint32_t offset, uint32_t stride, uint32_t idx, real o[2]) {
    if (buf >= s->nvb) return 0;
    uint64_t off = (uint64_t)offset + (uint64_t)stride * idx;
    if (off + 8 > s->vbsz[buf]) return 0;
    float f[2];
    memcpy(f, s->vb[buf] + off, 8);
    o[0] = f[0];
    o[1] = f[1];
    return 1;
}

typedef struct {
    f3 pos;
    uint32_t flags_material;
    uint32_t gn, sn, tan;
    real area;
    real uv[2];
    real uvs; /* mTexcoordScreenSize */
} shading_t; /* ShadingData, src/Shaders/compat/scene.h:185-198 (fields used here) */

/* makeTriangleShadingDataWithoutPosition + makeTriangleShadingData,
 * src/Shaders/common/shading_data.hlsli:59-151 */
/* makeTriangleShadingData, shading_data.hlsli:142-174: from barycentrics, or
 * (lp != NULL) from an object-space point on the triangle, whose barycentrics
 * are solved from the point and whose world position is T * lp. */
/* makeSphereShadingData, shading_data.hlsli:174-189, at the object-space
 * point lp: normal T lp normalised, area 4 pi r^2 (object space), the tangent
 * from the spherical uv taken as angles, as the reference writes it. */
static void make_sphere_shading(const or_scene* s, uint32_t ii, f3 lp, shading_t* r) {
    const st_instance* in = &s->inst[ii];
    const st_transform* T = &s->xf[ii];
    const f3 n = normalize3(xf_vector(T, lp));
    r->pos = xf_point(T, lp);
    r->flags_material = (in->type_material_address >> 4) << 4;
    r->gn = r->sn = pack_normal(n);
    const real rad = asfloat(in->data);
    r->area = 4 * OR_PI * rad * rad;
    real uv[2];
    cart_to_uv(normalize3(lp), uv);
    r->uv[0] = uv[0];
    r->uv[1] = uv[1];
    const real su = om_sin(uv[0]), cu = om_cos(uv[0]), sv = om_sin(uv[1]), cv = om_cos(uv[1]);
    const f3 dpdu = xf_vector(T, v3(-su * sv, 0.0f, cu * sv));
    const f3 dpdv = xf_vector(T, v3(cu * cv, -sv, su * cv));
    r->tan = pack_normal(normalize3(sub(dpdu, mulf(n, dot(n, dpdu)))));
    r->uvs = 1 / maxf(length3(dpdu), length3(dpdv));
}
static void make_tri_shading_x(const or_scene* s, uint32_t ii, uint32_t prim, real bx, real by, const f3* lp, shading_t* r) {
    if (prim == 0xFFFFu) { /* a sphere instance, at lp */
        make_sphere_shading(s, ii, *lp, r);
        return;
    }
    const st_instance* in = &s->inst[ii];
    const st_transform* T = &s->xf[ii];
    const st_mesh_vertex_info* vi = &s->vi[in->data & 0xffffu];
    uint32_t tri[3] = {0, 0, 0};
    load_tri(s, vi, prim, tri);
    uint32_t pb = vi_field(vi->packed_buffer_indices[0], 16, 16), ps = vi_field(vi->packed_strides, 8, 8);
    f3 v0 = v3(0, 0, 0), v1 = v0, v2 = v0;
    load_f3(s, pb, vi->packed_offsets[1], ps, tri[0], &v0);
    load_f3(s, pb, vi->packed_offsets[1], ps, tri[1], &v1);
    load_f3(s, pb, vi->packed_offsets[1], ps, tri[2], &v2);
    if (lp) {
        const f3 v1v0 = sub(v1, v0), v2v0 = sub(v2, v0), p_v0 = sub(*lp, v0);
        const real d00 = dot(v1v0, v1v0), d01 = dot(v1v0, v2v0), d11 = dot(v2v0, v2v0);
        const real d20 = dot(p_v0, v1v0), d21 = dot(p_v0, v2v0);
        const real den = d00 * d11 - d01 * d01;
        bx = (d11 * d20 - d01 * d21) / den;
        by = (d00 * d21 - d01 * d20) / den;
    }
    r->flags_material = (in->type_material_address >> 4) << 4;
    f3 dPds = xf_vector(T, sub(v0, v2)), dPdt = xf_vector(T, sub(v1, v2));
    f3 gn = cross(dPds, dPdt);
    real area2 = length3(gn);
    gn = divf(gn, area2);
    r->gn = pack_normal(gn);
    r->area = area2 / 2;
    real t0[2] = {0, 0}, t1[2] = {0, 0}, t2[2] = {0, 0};
    uint32_t tb = vi_field(vi->packed_buffer_indices[1], 16, 16), ts = vi_field(vi->packed_strides, 24, 8);
    if (tb < ST_IMAGE_LIMIT) {
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[0], t0);
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[1], t1);
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[2], t2);
    }
    real duvds[2] = {t2[0] - t0[0], t2[1] - t0[1]}, duvdt[2] = {t2[0] - t1[0], t2[1] - t1[1]};
    r->uv[0] = t0[0] + (t1[0] - t0[0]) * bx + duvds[0] * by;
    r->uv[1] = t0[1] + (t1[1] - t0[1]) * bx + duvds[1] * by;
    real det = duvds[0] * duvdt[1] - duvdt[0] * duvds[1];
    real invDet = 1 / det;
    real dsdu = duvdt[1] * invDet, dtdu = -duvds[1] * invDet;
    real dsdv = duvdt[0] * invDet, dtdv = -duvds[0] * invDet; /* as the reference writes them */
    f3 dPdu;
    if (det != 0) {
        dPdu = neg(add(mulf(dPds, dsdu), mulf(dPdt, dtdu)));
        const f3 dPdv = neg(add(mulf(dPds, dsdv), mulf(dPdt, dtdv)));
        r->uvs = 1 / maxf(length3(dPdu), length3(dPdv));
    } else {
        f3 onb[3];
        make_orthonormal(gn, onb);
        dPdu = onb[0];
        r->uvs = 1;
    }
    int snvalid = 0;
    f3 sn = v3(0, 0, 0);
    uint32_t nb = vi_field(vi->packed_buffer_indices[1], 0, 16), ns = vi_field(vi->packed_strides, 16, 8);
    if (nb < ST_IMAGE_LIMIT) { /* gShadingNormals defaults to true (TestRenderer.cpp:22) */
        f3 n0, n1, n2;
        if (load_f3(s, nb, vi->packed_offsets[2], ns, tri[0], &n0) && load_f3(s, nb, vi->packed_offsets[2], ns, tri[1], &n1) &&
            load_f3(s, nb, vi->packed_offsets[2], ns, tri[2], &n2)) {
            sn = add(add(n0, mulf(sub(n1, n0), bx)), mulf(sub(n2, n0), by));
            snvalid = !((sn.x == 0 && sn.y == 0 && sn.z == 0) || sn.x != sn.x || sn.y != sn.y || sn.z != sn.z);
        }
    }
    if (!snvalid) {
        r->sn = r->gn;
        r->tan = pack_normal(normalize3(dPdu));
    } else {
        sn = normalize3(xf_vector(T, sn));
        f3 tg = normalize3(sub(dPdu, mulf(sn, dot(sn, dPdu))));
        r->sn = pack_normal(sn);
        r->tan = pack_normal(tg);
        if (dot(sn, gn) < 0) r->gn = pack_normal(neg(gn));
    }
    r->pos = lp ? xf_point(T, *lp) : xf_point(T, add(add(v0, mulf(sub(v1, v0), bx)), mulf(sub(v2, v0), by)));
}
static void make_tri_shading(const or_scene* s, uint32_t ii, uint32_t prim, real bx, real by, shading_t* r) {
    make_tri_shading_x(s, ii, prim, bx, by, NULL, r);
}

/* ShadingData::toLocal / toWorld / shadingNormalCorrection, shading_data.hlsli:24-54 */
static f3 to_local(const shading_t* sd, f3 v) {
    f3 n = unpack_normal(sd->sn), t = unpack_normal(sd->tan);
    f3 b = cross(n, t); /* bitangent flip flag is never set for triangles */
    return v3(dot(v, t), dot(v, b), dot(v, n));
}
static f3 to_world(const shading_t* sd, f3 v) {
    f3 n = unpack_normal(sd->sn), t = unpack_normal(sd->tan);
    f3 b = cross(n, t);
    return add(add(mulf(t, v.x), mulf(b, v.y)), mulf(n, v.z));
}
static real signf(real x) { return x > 0 ? 1.0f : (x < 0 ? -1.0f : 0.0f); }
/* ShadingData::shadingNormalCorrection, shading_data.hlsli:35-54; adj: light paths (Adjoint) */
static real shading_normal_correction_x(const shading_t* sd, f3 din, f3 dout, int adj) {
    f3 lg = to_local(sd, unpack_normal(sd->gn));
    real ngin = dot(lg, din), ngout = dot(lg, dout);
    if (signf(ngout) != signf(dout.z) || signf(ngin) != signf(din.z)) return 0;
    real G = 1;
    if (adj) {
        const real num = ngout * din.z, denom = dout.z * ngin;
        if (fabsr(denom) > 1e-5f) G *= fabsr(num / denom);
    }
    return G;
}
static real shading_normal_correction(const shading_t* sd, f3 din, f3 dout) { return shading_normal_correction_x(sd, din, dout, 0); }

/* --------------------------------------------------------- materials */
/* has_bo: bo replaces the packed base colour (per-triangle albedo) */
typedef struct { uint32_t p[4]; real escale; int has_bo; f3 bo; } material_t;
static material_t load_material(const or_scene* s, uint32_t addr) {
    material_t m;
    memset(&m, 0, sizeof m);
    if ((uint64_t)addr / 4 + 9 > s->nmatw) return m;
    memcpy(m.p, s->mat + addr / 4, 16);
    m.escale = asfloat(s->mat[addr / 4 + 8]);
    return m;
}
static f3 m_base(const material_t* m) {
    return m->has_bo ? m->bo : v3(un8(m->p[0], 0), un8(m->p[0], 8), un8(m->p[0], 16));
}
static f3 m_emission(const material_t* m) {
    return mulf(v3(un8(m->p[1], 0), un8(m->p[1], 8), un8(m->p[1], 16)), m->escale);
}
#define M_SPEC(m) un8((m)->p[0], 24)
#define M_SHEEN(m) un8((m)->p[1], 24)
#define M_METAL(m) un8((m)->p[2], 0)
#define M_ROUGH(m) un8((m)->p[2], 8)
#define M_ANISO(m) un8((m)->p[2], 16)
#define M_SUBSURF(m) un8((m)->p[2], 24)
#define M_CC(m) un8((m)->p[3], 0)
#define M_CCGLOSS(m) un8((m)->p[3], 8)
#define M_TRANS(m) un8((m)->p[3], 16)
#define M_ETA1(m) (un8((m)->p[3], 24) + 1)

/* microfacet.hlsli:22-179 */
/* pow(x, 5) of disney.hlsli as (x^2)^2 x: within 3 ulp of the correctly
 * rounded power, the same operations as the GPU library */
static real powf5(real x) { const real x2 = x * x; return x2 * x2 * x; }
static f3 schlick3(f3 F0, real c) {
    real k = powf5(maxf(1 - c, 0.0f));
    return v3(F0.x + (1 - F0.x) * k, F0.y + (1 - F0.y) * k, F0.z + (1 - F0.z) * k);
}
static real schlick1(real F0, real c) { return F0 + (1 - F0) * powf5(maxf(1 - c, 0.0f)); }
static real fresnel_dielectric3(real ni, real nt, real eta) {
    real rs = (ni - eta * nt) / (ni + eta * nt);
    real rp = (eta * ni - nt) / (eta * ni + nt);
    return (rs * rs + rp * rp) / 2;
}
static real fresnel_dielectric(real ni, real eta) {
    real nt2 = 1 - (1 - ni * ni) / (eta * eta);
    if (nt2 < 0) return 1;
    return fresnel_dielectric3(fabsr(ni), sqrtr(nt2), eta);
}
static real GTR1(real ndh, real a) {
    real a2 = a * a, t = 1 + (a2 - 1) * ndh * ndh;
    return (a2 - 1) / (OR_PI * om_log(a2) * t);
}
static real GTR2a(f3 h, real ax, real ay) {
    f3 hs = divv(h, v3(ax, ay, 1));
    real l2 = dot(hs, hs);
    return 1 / (OR_PI * ax * ay * l2 * l2);
}
static real smith_gtr2(f3 v, real ax, real ay) {
    real ax2 = ax * ax, ay2 = ay * ay;
    f3 v2 = mul(v, v);
    real L = (-1 + sqrtr(1 + (v2.x * ax2 + v2.y * ay2) / v2.z)) / 2;
    return 1 / (1 + L);
}
static real smith_gtr1(f3 v) { return smith_gtr2(v, 0.25f, 0.25f); }
static f3 sample_vndf(f3 din, real ax, real ay, real r0, real r1) {
    int flip = din.z < 0;
    if (flip) din = neg(din);
    f3 hin = normalize3(v3(ax * din.x, ay * din.y, din.z));
    real r = sqrtr(r0), phi = 2 * OR_PI * r1;
    real t1 = r * om_cos(phi), t2 = r * om_sin(phi);
    real s = (1 + hin.z) / 2;
    t2 = (1 - s) * sqrtr(1 - t1 * t1) + s * t2;
    f3 dn = v3(t1, t2, sqrtr(maxf(0.0f, 1 - t1 * t1 - t2 * t2)));
    f3 fr[3];
    make_orthonormal(hin, fr);
    f3 hn = add(add(mulf(fr[0], dn.x), mulf(fr[1], dn.y)), mulf(hin, dn.z));
    f3 o = normalize3(v3(ax * hn.x, ay * hn.y, maxf(0.0f, hn.z)));
    return flip ? neg(o) : o;
}
static f3 lerp3(f3 a, f3 b, real t) { return add(a, mulf(sub(b, a), t)); }
static real lerp1(real a, real b, real t) { return a + (b - a) * t; }

/* Disney eval, src/Shaders/materials/disney.hlsli:9-139; adj: Adjoint (light paths),
 * refraction scaled by eta_factor = 1 / eta^2 (:123) */
static f3 disney_eval(const material_t* m, f3 din, f3 dout, int adj) {
    real aspect = sqrtr(1 - M_ANISO(m) * 0.9f), alpha = pow2(M_ROUGH(m));
    real ax = maxf(1e-4f, alpha / aspect), ay = maxf(1e-4f, alpha * aspect), metallic = M_METAL(m);
    if (din.z * dout.z > 0) {
        real ndin = fabsr(din.z), ndout = dout.z;
        f3 h = normalize3(add(din, dout));
        if (h.z * din.z < 0) h = neg(h);
        real ndh = h.z, hdin = dot(h, din), hdout = dot(h, dout);
        f3 c = v3(0, 0, 0);
        if (din.z >= 0 && dout.z >= 0) {
            real Fd90 = 0.5f + 2 * M_ROUGH(m) * hdout * hdout;
            real sno = powf5(1 - ndout), sni = powf5(1 - ndin), sho = powf5(1 - hdout);
            real based = (1 + (Fd90 - 1) * sno) * (1 + (Fd90 - 1) * sni);
            real Fss90 = hdout * hdout * M_ROUGH(m);
            real Fss = (1 + (Fss90 - 1) * sno) * (1 + (Fss90 - 1) * sni);
            real ss = 1.25f * (Fss * (1 / (ndout + ndin) - 0.5f) + 0.5f);
            f3 bc = m_base(m);
            c = add(c, mulf(mulf(mulf(bc, (1 - M_TRANS(m)) * (1 - metallic)), lerp1(based, ss, M_SUBSURF(m)) / OR_PI), ndout));
            real lum = luminance(bc);
            f3 ctint = lum > 0 ? divf(bc, lum) : v3(1, 1, 1);
            f3 csheen = lerp3(v3(1, 1, 1), ctint, 0.5f);
            c = add(c, mulf(mulf(mulf(csheen, (1 - metallic) * M_SHEEN(m)), sho), ndout));
            if (ndin > 0 && hdout > 0 && ndh > 0) {
                real eta = M_ETA1(m);
                real f0 = (eta - 1) * (eta - 1) / ((eta + 1) * (eta + 1));
                f3 scol = lerp3(v3(1, 1, 1), ctint, 0.5f);
                f3 c0 = lerp3(mulf(scol, M_SPEC(m) * f0), m_base(m), metallic);
                f3 F = schlick3(c0, hdout);
                real D = GTR2a(h, ax, ay), G = smith_gtr2(din, ax, ay) * smith_gtr2(dout, ax, ay);
                real sw = 1 - M_TRANS(m) * (1 - metallic);
                c = add(c, divf(mulf(mulf(mulf(F, sw), D), G), 4 * ndin));
            }
            if (ndin > 0 && ndh > 0) {
                real ac = lerp1(0.1f, 0.001f, M_CCGLOSS(m));
                real Fc = schlick1(0.04f, hdout), Dc = GTR1(ndh, ac), Gc = smith_gtr1(din) * smith_gtr1(dout);
                real v = M_CC(m) * Fc * Dc * Gc / (4 * ndin);
                c = add(c, v3(v, v, v));
            }
        }
        real gw = (1 - metallic) * M_TRANS(m);
        if (gw > 0) {
            real eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
            real Fg = fresnel_dielectric(hdin, eta), D = GTR2a(h, ax, ay);
            real G = smith_gtr2(din, ax, ay) * smith_gtr2(dout, ax, ay);
            c = add(c, mulf(m_base(m), gw * (Fg * D * G) / (4 * fabsr(ndin))));
        }
        return c;
    }
    real gw = (1 - metallic) * M_TRANS(m);
    if (gw <= 0) return v3(0, 0, 0);
    real eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
    f3 h = normalize3(add(din, mulf(dout, eta)));
    if (h.z * din.z < 0) h = neg(h);
    real hdin = dot(h, din), hdout = dot(h, dout), sd = hdin + eta * hdout;
    real Fg = fresnel_dielectric(hdin, eta), D = GTR2a(h, ax, ay);
    real G = smith_gtr2(din, ax, ay) * smith_gtr2(dout, ax, ay);
    f3 bc = m_base(m);
    f3 sq = v3(sqrtr(bc.x), sqrtr(bc.y), sqrtr(bc.z));
    const real eta_factor = adj ? (1 / (eta * eta)) : 1;
    return mulf(sq, gw * (eta_factor * (1 - Fg) * D * G * eta * eta * fabsr(hdout * hdin)) / (fabsr(din.z) * sd * sd));
}

/* Disney evalPdf, disney.hlsli:141-231 */
static real disney_pdf(const material_t* m, f3 din, f3 dout) {
    real alpha = pow2(M_ROUGH(m)), metallic = M_METAL(m), aspect = sqrtr(1 - M_ANISO(m) * 0.9f);
    real ax = maxf(0.0001f, alpha / aspect), ay = maxf(0.0001f, alpha * aspect);
    real ac = (1 - M_CCGLOSS(m)) * 0.1f + M_CCGLOSS(m) * 0.001f, tr = M_TRANS(m);
    real dw = (1 - metallic) * (1 - tr), mw = (1 - tr * (1 - metallic)), gw = (1 - metallic) * tr, cw = M_CC(m);
    real tw = dw + mw + gw + cw;
    real dp = dw / tw, mp = mw / tw, gp = gw / tw, cp = cw / tw;
    if (din.z < 0) {
        dp = 0; mp = 0; cp = 0;
        if (gp > 0) gp = 1;
    }
    if (din.z * dout.z > 0) {
        f3 h = normalize3(add(din, dout));
        if (h.z < 0) h = neg(h);
        real ndin = din.z, ndh = h.z, hdin = dot(h, din), hdout = dot(h, dout);
        if (din.z >= 0 && dout.z >= 0) {
            dp *= maxf(dout.z, 0.0f) / OR_PI;
            if (ndin > 0) {
                real D = GTR2a(h, ax, ay), Gi = smith_gtr2(din, ax, ay);
                mp *= (D * Gi / (4 * ndin));
            } else mp = 0;
            if (ndh > 0 && hdout > 0) {
                real Dc = GTR1(ndh, ac);
                cp *= (Dc * ndh / (4 * hdout));
            } else cp = 0;
        }
        real eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
        real Fg = fresnel_dielectric(hdin, eta), D = GTR2a(h, ax, ay), Gi = smith_gtr2(din, ax, ay);
        gp *= (Fg * D * Gi / (4 * fabsr(ndin)));
    } else {
        real eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
        f3 h = normalize3(add(din, mulf(dout, eta)));
        if (h.z < 0) h = neg(h);
        real hdin = dot(h, din), hdout = dot(h, dout);
        real D = GTR2a(h, ax, ay), Gi = smith_gtr2(din, ax, ay), Fg = fresnel_dielectric(hdin, eta);
        real sd = hdin + eta * hdout, dh = eta * eta * hdout / (sd * sd);
        gp *= (1 - Fg) * D * Gi * fabsr(dh * hdin / din.z);
    }
    return dp + mp + gp + cp;
}

typedef struct { f3 dir, f; real fwd, rev, eta, rough; } dsr_t;   /* DirectionSampleRecord */
typedef struct { f3 f; real fwd, rev; } rer_t;                    /* ReflectanceEvalRecord */

/* Disney sample, disney.hlsli:233-380 (returns direction, eta, roughness) */
static dsr_t disney_sample(const material_t* m, f3 din, f3 rnd) {
    dsr_t r;
    memset(&r, 0, sizeof r);
    real aspect = sqrtr(1 - M_ANISO(m) * 0.9f), alpha = pow2(M_ROUGH(m));
    real ax = maxf(0.0001f, alpha / aspect), ay = maxf(0.0001f, alpha * aspect);
    real ac = (1 - M_CCGLOSS(m)) * 0.1f + M_CCGLOSS(m) * 0.001f;
    real metallic = M_METAL(m), tr = M_TRANS(m);
    real dw = (1 - metallic) * (1 - tr), mw = (1 - tr * (1 - metallic)), gw = (1 - metallic) * tr, cw = M_CC(m);
    if (din.z >= 0) {
        real tw = dw + mw + gw + cw;
        real dp = dw / tw, mp = mw / tw, gp = gw / tw;
        if (rnd.z <= dp) {
            r.dir = sample_cos_hemisphere(rnd.x, rnd.y);
            r.eta = 0; r.rough = 1;
            return r;
        } else if (rnd.z <= (dp + mp)) {
            f3 h = sample_vndf(din, ax, ay, rnd.x, rnd.y);
            r.dir = normalize3(add(neg(din), mulf(h, 2 * dot(din, h))));
            r.eta = 0; r.rough = M_ROUGH(m);
            return r;
        } else if (rnd.z <= (dp + mp + gp)) {
            if (gp <= 0) return r;
            f3 h = sample_vndf(din, ax, ay, rnd.x, rnd.y);
            real hdin = dot(h, din), eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
            real F = fresnel_dielectric(hdin, eta), u = (rnd.z - (dp + mp)) / gp;
            if (u <= F) {
                r.dir = normalize3(add(neg(din), mulf(h, 2 * dot(din, h))));
                r.eta = 0;
                return r; /* roughness stays 0: reference returns before setting it */
            }
            real hdo2 = 1 - (1 - hdin * hdin) / (eta * eta);
            if (hdo2 <= 0) { memset(&r, 0, sizeof r); return r; }
            if (hdin < 0) h = neg(h);
            real hdo = sqrtr(hdo2);
            r.dir = add(divf(neg(din), eta), mulf(h, fabsr(hdin) / eta - hdo));
            r.eta = eta;
            r.rough = M_ROUGH(m);
            return r;
        } else {
            real a2 = ac * ac;
            real ch = sqrtr(maxf(0.0f, (1 - om_pow(a2, 1 - rnd.x)) / (1 - a2)));
            real sh = sqrtr(maxf(1 - ch * ch, 0.0f)), az = 2 * OR_PI * rnd.y;
            f3 mn = v3(sh * om_cos(az), sh * om_sin(az), ch);
            r.dir = normalize3(add(neg(din), mulf(mn, 2 * dot(din, mn))));
            r.eta = 0;
            r.rough = sqrtr(ac);
            return r;
        }
    }
    f3 h = sample_vndf(din, ax, ay, rnd.x, rnd.y);
    if (h.z < 0) h = neg(h);
    real hdin = dot(h, din), eta = din.z > 0 ? M_ETA1(m) : 1 / M_ETA1(m);
    real F = fresnel_dielectric(hdin, eta);
    if (rnd.z <= F) {
        r.dir = normalize3(add(neg(din), mulf(h, 2 * dot(din, h))));
        r.eta = 0;
    } else {
        real hdo2 = 1 - (1 - hdin * hdin) / (eta * eta);
        if (hdo2 <= 0) { memset(&r, 0, sizeof r); return r; }
        if (hdin < 0) h = neg(h);
        real hdo = sqrtr(hdo2);
        r.dir = add(divf(neg(din), eta), mulf(h, fabsr(hdin) / eta - hdo));
        r.eta = eta;
    }
    r.rough = M_ROUGH(m);
    return r;
}

/* Visible-normal sampling as Heitz publishes it (JCGT 7(4), 2018, listing 1):
 * the disk frame's first axis horizontal, its second in the plane of the
 * normal and the view.  sample_vndf (the reference's) builds the frame with
 * makeOrthonormal, whose axes are not so aligned away from normal incidence,
 * so its directions are not distributed as the pdf it reports. */
static f3 sample_vndf_heitz(f3 din, real ax, real ay, real r0, real r1) {
    const int flip = din.z < 0;
    if (flip) din = neg(din);
    const f3 vh = normalize3(v3(ax * din.x, ay * din.y, din.z));
    const real lensq = vh.x * vh.x + vh.y * vh.y;
    const f3 e1 = lensq > 0 ? divf(v3(-vh.y, vh.x, 0), sqrtr(lensq)) : v3(1, 0, 0);
    const f3 e2 = cross(vh, e1);
    const real r = sqrtr(r0), phi = 2 * OR_PI * r1;
    real t1 = r * om_cos(phi), t2 = r * om_sin(phi);
    const real s = (1 + vh.z) / 2;
    t2 = (1 - s) * sqrtr(1 - t1 * t1) + s * t2;
    const f3 nh = add(add(mulf(e1, t1), mulf(e2, t2)), mulf(vh, sqrtr(maxf(0.0f, 1 - t1 * t1 - t2 * t2))));
    const f3 o = normalize3(v3(ax * nh.x, ay * nh.y, maxf(0.0f, nh.z)));
    return flip ? neg(o) : o;
}
/* The GGX conductor of ST_FLAG_LAMBERT_GGX: the Disney BSDF's metallic lobe
 * alone (disney.hlsli:56-70 and :163-169 at metallic 1, no clearcoat), F
 * Schlick's about the base colour, reflection only, two-sided. */
static rer_t ggx_eval(const material_t* m, f3 din, f3 dout) {
    rer_t r;
    r.f = v3(0, 0, 0); r.fwd = r.rev = 0;
    if (!(din.z * dout.z > 0)) return r;
    if (din.z < 0) { din = neg(din); dout = neg(dout); }
    const real aspect = sqrtr(1 - M_ANISO(m) * 0.9f), alpha = pow2(M_ROUGH(m));
    const real ax = maxf(1e-4f, alpha / aspect), ay = maxf(1e-4f, alpha * aspect);
    const f3 h = normalize3(add(din, dout));
    const real D = GTR2a(h, ax, ay), Gi = smith_gtr2(din, ax, ay), Go = smith_gtr2(dout, ax, ay);
    const f3 F = schlick3(m_base(m), dot(h, dout));
    r.f = divf(mulf(F, D * Gi * Go), 4 * din.z);
    r.fwd = D * Gi / (4 * din.z);
    r.rev = D * Go / (4 * dout.z);
    return r;
}
static dsr_t ggx_sample(const material_t* m, f3 din, f3 rnd) {
    const real aspect = sqrtr(1 - M_ANISO(m) * 0.9f), alpha = pow2(M_ROUGH(m));
    const real ax = maxf(1e-4f, alpha / aspect), ay = maxf(1e-4f, alpha * aspect);
    const f3 h = sample_vndf_heitz(din, ax, ay, rnd.x, rnd.y);
    dsr_t r;
    r.dir = normalize3(add(neg(din), mulf(h, 2 * dot(din, h))));
    const rer_t f = ggx_eval(m, din, r.dir);
    r.f = f.f; r.fwd = f.fwd; r.rev = f.rev;
    r.eta = 0; r.rough = M_ROUGH(m);
    return r;
}
/* the model of a material: 0 Disney, 1 Lambert, 2 GGX conductor.  mode: the
 * configuration's (0 Disney, 1 all Lambert, 2 per material by metallic) */
static int m_model(const material_t* m, int mode) { return mode == 2 ? (M_METAL(m) >= 0.5f ? 2 : 1) : mode; }

/* evaluateReflectance for Disney (disney.hlsli:391-403) or Lambert (lambertian.hlsli:13-24),
 * or per material Lambert or the GGX conductor */
static rer_t bsdf_eval_x(const material_t* m, int mode, f3 din, f3 dout, int adj) {
    rer_t r;
    const int lambert = m_model(m, mode) == 1;
    if (m_model(m, mode) == 2) return ggx_eval(m, din, dout);
    if (lambert) {
        if (din.z * dout.z < 0) { r.f = v3(0, 0, 0); r.fwd = r.rev = 0; return r; }
        r.f = divf(mulf(m_base(m), fabsr(dout.z)), OR_PI);
        r.fwd = cos_hemisphere_pdfw(maxf(1e-5f, fabsr(dout.z)));
        r.rev = cos_hemisphere_pdfw(maxf(1e-5f, fabsr(din.z)));
        return r;
    }
    r.f = disney_eval(m, din, dout, adj);
    r.fwd = disney_pdf(m, din, dout);
    r.rev = disney_pdf(m, dout, din);
    return r;
}
/* sampleDirection (disney.hlsli:415-429, lambertian.hlsli:29-39) */
static rer_t bsdf_eval(const material_t* m, int lambert, f3 din, f3 dout) { return bsdf_eval_x(m, lambert, din, dout, 0); }
static dsr_t bsdf_sample_x(const material_t* m, int mode, f3 rnd, f3 din, int adj) {
    dsr_t r;
    const int lambert = m_model(m, mode) == 1;
    if (m_model(m, mode) == 2) return ggx_sample(m, din, rnd);
    if (lambert) {
        r.dir = sample_cos_hemisphere(rnd.x, rnd.y);
        if (din.z < 0) r.dir = neg(r.dir);
        r.f = divf(mulf(m_base(m), fabsr(r.dir.z)), OR_PI);
        r.fwd = cos_hemisphere_pdfw(maxf(1e-5f, fabsr(r.dir.z)));
        r.rev = cos_hemisphere_pdfw(maxf(1e-5f, fabsr(din.z)));
        r.rough = 1; r.eta = 0;
        return r;
    }
    r = disney_sample(m, din, rnd);
    rer_t f = bsdf_eval_x(m, 0, din, r.dir, adj);
    r.f = f.f; r.fwd = f.fwd; r.rev = f.rev;
    return r;
}
static dsr_t bsdf_sample(const material_t* m, int lambert, f3 rnd, f3 din) { return bsdf_sample_x(m, lambert, rnd, din, 0); }
static int m_singular(const material_t* m, int mode) {
    const int model = m_model(m, mode);
    if (model == 1) return 0;
    if (model == 2) return (1 - M_ROUGH(m)) > 0.75f; /* isSingular at metallic 1 */
    return maxf(maxf(M_SPEC(m), M_METAL(m)), M_TRANS(m)) * (1 - M_ROUGH(m)) > 0.75f;
}
static int m_can_eval(const material_t* m) { f3 b = m_base(m); return any_gt0(b); }
static real m_cont_prob(const material_t* m) { return saturate(luminance(m_base(m))); }
static void tex_sample(const or_scene* s, uint32_t img, real u, real v, real uvs, real out[4]);
static uint32_t pack_unorm4(const real v[4]);
static void unpack_unorm4(uint32_t w, real v[4]);
/* LoadMaterial with its images (material.hlsli:22-45) at the shading point's
 * uv and screen size: base colour times the image (setBaseColor re-quantises
 * to 8 bits), the metallic/roughness/anisotropy/subsurface word modulated
 * (roughness through smoothness), the clearcoat/transmission/eta word
 * modulated.  Emission images are refused at scene creation. */
static material_t load_material_sd(const or_scene* s, const shading_t* sd) {
    const uint32_t addr = sd->flags_material >> 4;
    material_t m = load_material(s, addr);
    if (!s->ntex || (uint64_t)addr / 4 + 8 > s->nmatw) return m;
    const uint32_t* ii = s->mat + addr / 4 + 4;
    if (ii[0] < s->ntex && s->tex[ii[0]].n) {
        real t[4];
        tex_sample(s, ii[0], sd->uv[0], sd->uv[1], sd->uvs, t);
        const f3 b = m_base(&m);
        const real bb[3] = {b.x * t[0], b.y * t[1], b.z * t[2]};
        m.p[0] &= 0xff000000u;
        for (int c = 0; c < 3; c++) m.p[0] |= (uint32_t)floorr(saturate(bb[c]) * 255 + 0.5f) << (8 * c);
    }
    if (ii[2] < s->ntex && s->tex[ii[2]].n) {
        real v[4], t[4];
        unpack_unorm4(m.p[2], v);
        tex_sample(s, ii[2], sd->uv[0], sd->uv[1], sd->uvs, t);
        v[0] *= t[0];
        v[2] *= t[2];
        v[3] *= t[3];
        v[1] = 1 - (1 - v[1]) * (1 - t[1]);
        m.p[2] = pack_unorm4(v);
    }
    if (ii[3] < s->ntex && s->tex[ii[3]].n) {
        real v[4], t[4];
        unpack_unorm4(m.p[3], v);
        tex_sample(s, ii[3], sd->uv[0], sd->uv[1], sd->uvs, t);
        for (int c = 0; c < 4; c++) v[c] *= t[c];
        m.p[3] = pack_unorm4(v);
    }
    return m;
}


void or_bsdf_eval(const uint32_t mat[5], int lambertian, const float din[3], const float dout[3], float o[5]) {
    material_t m;
    memset(&m, 0, sizeof m);
    memcpy(m.p, mat, 16);
    m.escale = asfloat(mat[4]);
    rer_t r = bsdf_eval(&m, lambertian, v3(din[0], din[1], din[2]), v3(dout[0], dout[1], dout[2]));
    o[0] = r.f.x; o[1] = r.f.y; o[2] = r.f.z; o[3] = r.fwd; o[4] = r.rev;
}
void or_bsdf_eval_adjoint(const uint32_t mat[5], int lambertian, int adjoint, const float din[3], const float dout[3],
                          float o[5]) {
    material_t m;
    memset(&m, 0, sizeof m);
    memcpy(m.p, mat, 16);
    m.escale = asfloat(mat[4]);
    rer_t r = bsdf_eval_x(&m, lambertian, v3(din[0], din[1], din[2]), v3(dout[0], dout[1], dout[2]), adjoint);
    o[0] = r.f.x; o[1] = r.f.y; o[2] = r.f.z; o[3] = r.fwd; o[4] = r.rev;
}
void or_bsdf_sample(const uint32_t mat[5], int lambertian, const float rnd[3], const float din[3], float o[10]) {
    material_t m;
    memset(&m, 0, sizeof m);
    memcpy(m.p, mat, 16);
    m.escale = asfloat(mat[4]);
    dsr_t r = bsdf_sample(&m, lambertian, v3(rnd[0], rnd[1], rnd[2]), v3(din[0], din[1], din[2]));
    o[0] = r.dir.x; o[1] = r.dir.y; o[2] = r.dir.z; o[3] = r.f.x; o[4] = r.f.y; o[5] = r.f.z;
    o[6] = r.fwd; o[7] = r.rev; o[8] = r.eta; o[9] = r.rough;
}

/* ------------------------------------------------------- environment */
static const float* texel(const level_t* l, int x, int y) { return l->rgba + 4 * ((size_t)y * l->w + x); }
static real texel_lum(const level_t* l, int x, int y) {
    const float* t = texel(l, x, y);
    return luminance(v3(t[0], t[1], t[2]));
}
/* mip chain by 2x2 box (vkCmdBlitImage with linear filter, src/Core/Image.cpp:282-308) */
static int build_level_chain(level_t* lv) {
    int l = 0;
    while ((lv[l].w > 1 || lv[l].h > 1) && l < 15) {
        level_t* a = &lv[l];
        level_t* b = &lv[l + 1];
        b->w = a->w > 1 ? a->w / 2 : 1;
        b->h = a->h > 1 ? a->h / 2 : 1;
        b->rgba = (float*)malloc(sizeof(float) * 4 * (size_t)b->w * b->h);
        for (int y = 0; y < b->h; y++)
            for (int x = 0; x < b->w; x++) {
                int x0 = a->w > 1 ? 2 * x : 0, x1 = a->w > 1 ? 2 * x + 1 : 0;
                int y0 = a->h > 1 ? 2 * y : 0, y1 = a->h > 1 ? 2 * y + 1 : 0;
                for (int c = 0; c < 4; c++) {
                    real v = ((texel(a, x0, y0)[c] + texel(a, x1, y0)[c]) + (texel(a, x0, y1)[c] + texel(a, x1, y1)[c])) * 0.25f;
                    b->rgba[4 * ((size_t)y * b->w + x) + c] = v;
                }
            }
        l++;
    }
    return l + 1;
}
static void build_mips(or_scene* s) { s->nlevels = build_level_chain(s->lv); }
/* bilinear SampleLevel(uv, 0) with repeat addressing (TestRenderer.cpp:54-57) */
static f3 env_bilinear(const level_t* l, real u, real v) {
    real x = u * (real)l->w - 0.5f, y = v * (real)l->h - 0.5f;
    real fx0 = floorr(x), fy0 = floorr(y), ax = x - fx0, ay = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    int xa = ((x0 % l->w) + l->w) % l->w, xb = (((x0 + 1) % l->w) + l->w) % l->w;
    int ya = ((y0 % l->h) + l->h) % l->h, yb = (((y0 + 1) % l->h) + l->h) % l->h;
    const float *t00 = texel(l, xa, ya), *t10 = texel(l, xb, ya), *t01 = texel(l, xa, yb), *t11 = texel(l, xb, yb);
    f3 r;
    real* rr = &r.x;
    for (int c = 0; c < 3; c++) {
        real top = t00[c] * (1 - ax) + t10[c] * ax;
        real bot = t01[c] * (1 - ax) + t11[c] * ax;
        rr[c] = top * (1 - ay) + bot * ay;
    }
    return r;
}
/* bilinear SampleLevel of all four channels at one level, repeat addressing */
static void bilinear4(const level_t* l, real u, real v, real out[4]) {
    real x = u * (real)l->w - 0.5f, y = v * (real)l->h - 0.5f;
    real fx0 = floorr(x), fy0 = floorr(y), ax = x - fx0, ay = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    int xa = ((x0 % l->w) + l->w) % l->w, xb = (((x0 + 1) % l->w) + l->w) % l->w;
    int ya = ((y0 % l->h) + l->h) % l->h, yb = (((y0 + 1) % l->h) + l->h) % l->h;
    const float *t00 = texel(l, xa, ya), *t10 = texel(l, xb, ya), *t01 = texel(l, xa, yb), *t11 = texel(l, xb, yb);
    for (int c = 0; c < 4; c++) {
        real top = t00[c] * (1 - ax) + t10[c] * ax;
        real bot = t01[c] * (1 - ax) + t11[c] * ax;
        out[c] = top * (1 - ay) + bot * ay;
    }
}
/* SampleImage4 (scene.hlsli:101-109): SampleLevel at lod = log2(max(uvScreenSize
 * * max(w, h), 1e-6)) (0 without a screen size), with the static sampler's
 * linear filter between the two nearest levels (TestRenderer.cpp:54-57); the
 * lod is clamped to the chain.  log2 is the renderer's log over ln 2. */
static void tex_sample(const or_scene* s, uint32_t img, real u, real v, real uvs, real out[4]) {
    const struct tex_s* t = &s->tex[img];
    real lod = 0;
    if (uvs > 0) {
        const real wh = (real)(t->lv[0].w > t->lv[0].h ? t->lv[0].w : t->lv[0].h);
        lod = om_log(maxf(uvs * wh, 1e-6f)) * (real)1.44269504088896340736;
    }
    if (!(lod > 0)) lod = 0;
    if (lod > (real)(t->n - 1)) lod = (real)(t->n - 1);
    const int l0 = (int)floorr(lod), l1 = l0 + 1 < t->n ? l0 + 1 : l0;
    const real f = lod - (real)l0;
    real a[4], b[4];
    bilinear4(&t->lv[l0], u, v, a);
    if (f > 0 && l1 != l0) {
        bilinear4(&t->lv[l1], u, v, b);
        for (int c = 0; c < 4; c++) out[c] = a[c] * (1 - f) + b[c] * f;
    } else {
        for (int c = 0; c < 4; c++) out[c] = a[c];
    }
}
/* D3DX_FLOAT4_to_R8G8B8A8_UNORM: saturate, scale by 255, round half up */
static uint32_t pack_unorm4(const real v[4]) {
    uint32_t w = 0;
    for (int c = 0; c < 4; c++) w |= (uint32_t)(saturate(v[c]) * 255 + 0.5f) << (8 * c);
    return w;
}
static void unpack_unorm4(uint32_t w, real v[4]) {
    for (int c = 0; c < 4; c++) v[c] = (real)((w >> (8 * c)) & 0xffu) / 255.0f;
}

/* unit entry: SampleImage4 of one RGBA32F image (its mip chain built here) */
int or_tex_sample(const float* rgba, uint32_t w, uint32_t h, float u, float v, float uvs, float out[4]) {
    or_scene tmp;
    memset(&tmp, 0, sizeof tmp);
    struct tex_s t;
    memset(&t, 0, sizeof t);
    t.lv[0].w = (int)w; t.lv[0].h = (int)h;
    t.lv[0].rgba = (float*)malloc(sizeof(float) * 4 * (size_t)w * h);
    memcpy(t.lv[0].rgba, rgba, sizeof(float) * 4 * (size_t)w * h);
    t.n = build_level_chain(t.lv);
    tmp.tex = &t;
    tmp.ntex = 1;
    real o[4];
    tex_sample(&tmp, 0, u, v, uvs, o);
    for (int c = 0; c < 4; c++) out[c] = o[c];
    for (int l = 0; l < 16; l++) free(t.lv[l].rgba);
    return 0;
}

/* SampleTexel, src/Shaders/common/lights.hlsli:49-97 */
static void sample_texel(const or_scene* s, real rx, real ry, real uv[2], real* pdf) {
    int levelCount = s->nlevels, cx = 0, cy = 0, lw = 1, lh = 1;
    *pdf = 1;
    int iend = levelCount - 1 < 11 ? levelCount - 1 : 11;
    for (int i = 1; i < iend; i++) {
        const level_t* L = &s->lv[levelCount - 1 - i];
        real inv_h = 1 / (real)L->h;
        cx *= L->w / lw; cy *= L->h / lh;
        real p[4] = {0, 0, 0, 0};
        if ((uint32_t)L->w - (uint32_t)cx > 1) {
            real sy = om_sin(OR_PI * ((real)cy + 0.5f) * inv_h);
            p[0] = texel_lum(L, cx, cy) * sy;
            p[1] = texel_lum(L, cx + 1, cy) * sy;
        }
        if ((uint32_t)L->h - (uint32_t)cy > 1) {
            real sy = om_sin(OR_PI * ((real)cy + 1.5f) * inv_h);
            p[2] = texel_lum(L, cx, cy + 1) * sy;
            p[3] = texel_lum(L, cx + 1, cy + 1) * sy;
        }
        real sum = p[0] + p[1] + p[2] + p[3];
        if (sum < 1e-6f) continue;
        for (int j = 0; j < 4; j++) p[j] = p[j] / sum;
        for (int j = 0; j < 4; j++) {
            if (j == 3 || rx < p[j]) {
                cx += j % 2; cy += j / 2;
                *pdf *= p[j];
                rx /= p[j];
                break;
            }
            rx -= p[j];
        }
        lw = L->w; lh = L->h;
    }
    *pdf *= (real)((uint32_t)lw * (uint32_t)lh);
    uv[0] = ((real)cx + rx) / (real)lw;
    uv[1] = ((real)cy + ry) / (real)lh;
}
/* SampleTexelPdf, lights.hlsli:98-139 */
static real sample_texel_pdf(const or_scene* s, real u, real v) {
    int levelCount = s->nlevels, lw = 1, lh = 1;
    real pdf = 1;
    int iend = levelCount - 1 < 11 ? levelCount - 1 : 11;
    for (int i = 1; i < iend; i++) {
        const level_t* L = &s->lv[levelCount - 1 - i];
        real inv_h = 1 / (real)L->h;
        int ux = (int)((real)L->w * u), uy = (int)((real)L->h * v);
        int cx = (ux / 2) * 2, cy = (uy / 2) * 2;
        int ox = ux - cx < 1 ? ux - cx : 1, oy = uy - cy < 1 ? uy - cy : 1;
        real p[4] = {0, 0, 0, 0};
        if ((uint32_t)L->w - (uint32_t)cx > 1) {
            real sy = om_sin(OR_PI * ((real)cy + 0.5f) * inv_h);
            p[0] = texel_lum(L, cx, cy) * sy;
            p[1] = texel_lum(L, cx + 1, cy) * sy;
        }
        if ((uint32_t)L->h - (uint32_t)cy > 1) {
            real sy = om_sin(OR_PI * ((real)cy + 1.5f) * inv_h);
            p[2] = texel_lum(L, cx, cy + 1) * sy;
            p[3] = texel_lum(L, cx + 1, cy + 1) * sy;
        }
        real sum = p[0] + p[1] + p[2] + p[3];
        if (sum < 1e-6f) continue;
        for (int j = 0; j < 4; j++) p[j] = p[j] / sum;
        pdf *= p[oy * 2 + ox];
        lw = L->w; lh = L->h;
    }
    return pdf * (real)((uint32_t)lw * (uint32_t)lh);
}
/* Alias table over the full-resolution map (the north-star environment
 * sampler, ST_FLAG_ENV_ALIAS; the reference warps the mip pyramid instead).
 * Walker's alias method with Vose's two-worklist construction (M. D. Vose,
 * "A linear algorithm for generating random numbers with a given
 * distribution", IEEE TSE 17(9), 1991): texel weights w_i = max(0, lum) *
 * sin(pi (y + 1/2) / H) in double; column i keeps probability min(1, N w_i /
 * sum w) of itself and gives the rest to one donor.  Worklists are stacks
 * (last in, first out), so the table is one fixed function of the image.
 * pdf[i] = N w_i / sum w is the texel's uv-space density. */
static int build_alias(or_scene* s) {
    const level_t* l = &s->lv[0];
    const size_t N = (size_t)l->w * l->h;
    double* wgt = (double*)malloc(sizeof(double) * N);
    double* sc = (double*)malloc(sizeof(double) * N);
    uint32_t* small = (uint32_t*)malloc(sizeof(uint32_t) * N);
    uint32_t* large = (uint32_t*)malloc(sizeof(uint32_t) * N);
    s->alias_prob = (float*)malloc(sizeof(float) * N);
    s->alias_idx = (uint32_t*)malloc(sizeof(uint32_t) * N);
    s->alias_pdf = (float*)malloc(sizeof(float) * N);
    if (!wgt || !sc || !small || !large || !s->alias_prob || !s->alias_idx || !s->alias_pdf) {
        free(wgt); free(sc); free(small); free(large);
        return 1;
    }
    double total = 0;
    for (int y = 0; y < l->h; y++) {
        const double st = sin(3.14159265358979323846 * (y + 0.5) / l->h);
        for (int x = 0; x < l->w; x++) {
            const float* t = texel(l, x, y);
            const double lum = 0.2126 * t[0] + 0.7152 * t[1] + 0.0722 * t[2];
            wgt[(size_t)y * l->w + x] = (lum > 0 ? lum : 0.0) * st;
            total += wgt[(size_t)y * l->w + x];
        }
    }
    if (total <= 0) {
        for (size_t i = 0; i < N; i++) wgt[i] = 1;
        total = (double)N;
    }
    size_t ns = 0, nl = 0;
    for (size_t i = 0; i < N; i++) {
        sc[i] = wgt[i] / total * N;
        s->alias_pdf[i] = (real)(wgt[i] / total * N);
        if (sc[i] < 1) small[ns++] = (uint32_t)i;
        else large[nl++] = (uint32_t)i;
    }
    while (ns > 0 && nl > 0) {
        const uint32_t a = small[--ns], g = large[nl - 1];
        s->alias_prob[a] = (real)sc[a];
        s->alias_idx[a] = g;
        sc[g] = (sc[g] + sc[a]) - 1;
        if (sc[g] < 1) {
            nl--;
            small[ns++] = g;
        }
    }
    /* leftovers hold probability 1 up to rounding */
    for (size_t j = 0; j < nl; j++) { s->alias_prob[large[j]] = 1; s->alias_idx[large[j]] = large[j]; }
    for (size_t j = 0; j < ns; j++) { s->alias_prob[small[j]] = 1; s->alias_idx[small[j]] = small[j]; }
    free(wgt); free(sc); free(small); free(large);
    /* a jitter this far below 1 keeps x + j and (int)((x + j) / W * W) inside
     * texel x for every x < 2^lg, so the pdf a sample reports is the pdf
     * evaluated at its point */
    int lg = 0;
    while ((1 << lg) < (l->w > l->h ? l->w : l->h)) lg++;
    s->alias_jmax = 1.0f - ldexpr(1.0f, lg - 21);
    return 0;
}
/* One texel by the alias table, a uniform point inside it.  rx picks the
 * column, ra is the accept test, ju / jv place the point in the texel; *pdf
 * is the uv-space density of the texel drawn. */
static void sample_alias(const or_scene* s, real rx, real ra, real ju, real jv, real uv[2], real* pdf) {
    const int W = s->lv[0].w, H = s->lv[0].h;
    const uint32_t N = (uint32_t)W * (uint32_t)H;
    uint32_t i = (uint32_t)(rx * (real)N);
    if (i >= N) i = N - 1;
    if (!(ra < s->alias_prob[i])) i = s->alias_idx[i];
    *pdf = s->alias_pdf[i];
    uv[0] = ((real)(i % (uint32_t)W) + minf(ju, s->alias_jmax)) / (real)W;
    uv[1] = ((real)(i / (uint32_t)W) + minf(jv, s->alias_jmax)) / (real)H;
}
/* uv-space density of the alias sampler at uv */
static real alias_pdf(const or_scene* s, real u, real v) {
    const int W = s->lv[0].w, H = s->lv[0].h;
    int x = (int)(u * (real)W), y = (int)(v * (real)H);
    x = x < 0 ? 0 : (x >= W ? W - 1 : x);
    y = y < 0 ? 0 : (y >= H ? H - 1 : y);
    return s->alias_pdf[(size_t)y * W + x];
}
int or_env_alias_sample(or_scene* s, uint32_t n, const float* rnd, float* uv, float* pdf) {
    if (!s->env_has_image) return 1;
    for (uint32_t i = 0; i < n; i++)
    {
        real u[2], pd;
        sample_alias(s, rnd[4 * i], rnd[4 * i + 1], rnd[4 * i + 2], rnd[4 * i + 3], u, &pd);
        uv[2 * i] = u[0];
        uv[2 * i + 1] = u[1];
        pdf[i] = pd;
    }
    return 0;
}
float or_env_alias_pdf(or_scene* s, const float uv[2]) { return s->env_has_image ? alias_pdf(s, uv[0], uv[1]) : 0; }
int or_env_alias_table(or_scene* s, float* prob, uint32_t* alias, float* pdf) {
    if (!s->env_has_image) return 1;
    const size_t N = (size_t)s->lv[0].w * s->lv[0].h;
    if (prob) memcpy(prob, s->alias_prob, sizeof(float) * N);
    if (alias) memcpy(alias, s->alias_idx, sizeof(uint32_t) * N);
    if (pdf) memcpy(pdf, s->alias_pdf, sizeof(float) * N);
    return 0;
}
int or_env_sample_texel(or_scene* s, const float rnd[2], float uv[2], float* pdf) {
    if (!s->env_has_image) return 1;
    real u[2], pd;
    sample_texel(s, rnd[0], rnd[1], u, &pd);
    uv[0] = u[0];
    uv[1] = u[1];
    *pdf = pd;
    return 0;
}
float or_env_sample_texel_pdf(or_scene* s, const float uv[2]) {
    return s->env_has_image ? sample_texel_pdf(s, uv[0], uv[1]) : 0;
}

typedef struct {
    uint32_t flags;
    real env_prob;
    uint32_t max_depth, min_len, cand, max_null, seed_px, seed_py, seed;
    uint32_t rr_start; /* Russian roulette from this path length on (0, 1: every vertex) */
    int lambert;      /* 0 Disney, 1 every material Lambert (gLambertian), 2 per material Lambert / GGX */
    real max_m;       /* mMaxM (ReSTIR reuse: M clamped at max_m x candidates) */
} cfg_t;

/* The continuation probability folded into the BSDF / phase pdfs at a vertex
 * of path length plen: Russian roulette runs at every vertex in the reference
 * (testrenderer.slang:890-892); the north-star mode starts it at rr_start. */
static real rr_cp(const cfg_t* c, uint32_t plen, real cp) { return plen >= c->rr_start ? cp : 1.0f; }

/* EvaluateEnvironment, lights.hlsli:143-163 */
static f3 evaluate_environment(const or_scene* s, const cfg_t* c, f3 d, real* pdfw) {
    if (!s->has_env) { *pdfw = 0; return v3(0, 0, 0); }
    f3 e = s->env_emission;
    if (s->env_has_image) {
        real uv[2];
        cart_to_uv(d, uv);
        e = mul(e, env_bilinear(&s->lv[0], uv[0], uv[1]));
        const real pu = (c->flags & ST_FLAG_ENV_ALIAS) ? alias_pdf(s, uv[0], uv[1]) : sample_texel_pdf(s, uv[0], uv[1]);
        *pdfw = pu / (2 * OR_PI * OR_PI * sqrtr(1 - d.y * d.y));
    } else {
        *pdfw = 1 / (4 * OR_PI);
    }
    if (s->nlights > 0) *pdfw *= c->env_prob;
    return e;
}

typedef struct {
    f3 radiance, dir, pos;
    real pdf, dist, cos_light;
    uint32_t normal;
    int finite;
} isr_t; /* IlluminationSampleRecord, lights.hlsli:32-45 */

/* SampleEmission + SampleIllumination, lights.hlsli:166-260.  The reference
 * leaves mDirectionToLight unset for environment samples (lights.hlsli:234-244);
 * here it is the sampled direction. */
static isr_t sample_illumination(const or_scene* s, const cfg_t* c, const real rnd_in[4], f3 ref) {
    isr_t r;
    memset(&r, 0, sizeof r);
    real rnd[4] = {rnd_in[0], rnd_in[1], rnd_in[2], rnd_in[3]};
    if (s->has_env) {
        if (s->nlights == 0 || rnd[3] < c->env_prob) {
            f3 d;
            real pdf;
            const int alias = (c->flags & ST_FLAG_ENV_ALIAS) && s->env_has_image;
            real uv[2] = {0, 0};
            if (s->env_has_image) {
                if (alias) {
                    /* rnd[2] is free on this branch; rnd[3], rescaled past the
                     * environment-or-emitter choice, is uniform again */
                    const real ju = s->nlights > 0 ? rnd[3] / c->env_prob : rnd[3];
                    sample_alias(s, rnd[0], rnd[2], ju, rnd[1], uv, &pdf);
                } else {
                    sample_texel(s, rnd[0], rnd[1], uv, &pdf);
                }
                d = uv_to_cart(uv[0], uv[1]);
                pdf /= (2 * OR_PI * OR_PI * sqrtr(1 - pow2(d.y)));
            } else {
                d = sample_uniform_sphere(rnd[0], rnd[1]);
                pdf = 1 / (4 * OR_PI);
            }
            if (s->nlights > 0) pdf *= c->env_prob;
            r.pdf = pdf;
            r.dist = INFINITY;
            r.pos = v3(INFINITY, INFINITY, INFINITY);
            r.cos_light = 1;
            r.finite = 0;
            r.dir = d;
            if (alias) {
                /* alias mode: the radiance at the sampled uv itself, without the
                 * direction -> uv round trip */
                r.radiance = mul(s->env_emission, env_bilinear(&s->lv[0], uv[0], uv[1]));
            } else {
                real tmp;
                r.radiance = evaluate_environment(s, c, d, &tmp);
            }
            return r;
        }
        if (s->nlights > 0) rnd[3] = (rnd[3] - c->env_prob) / (1 - c->env_prob);
    }
    if (s->nlights == 0) return r;
    uint32_t li = (uint32_t)(rnd[2] * (real)s->nlights) % s->nlights;
    uint32_t ii = s->light_inst[li];
    real pdf = 1 / (real)s->nlights;
    if (s->has_env) pdf *= 1 - c->env_prob;
    const st_instance* in = &s->inst[ii];
    const uint32_t type = in->type_material_address & 0xfu;
    shading_t sd;
    if (type == 1) { /* sphere: makeSphereShadingData at radius * sampleUniformSphere (lights.hlsli:215-220) */
        make_sphere_shading(s, ii, mulf(sample_uniform_sphere(rnd[0], rnd[1]), asfloat(in->data)), &sd);
    } else if (type == 0) {
        uint32_t pc = in->data >> 16;
        uint32_t prim = (uint32_t)(rnd[3] * (real)pc) % pc;
        pdf /= (real)pc;
        real b[2];
        sample_uniform_triangle(rnd[0], rnd[1], b);
        make_tri_shading(s, ii, prim, b[0], b[1], &sd);
    } else {
        memset(&r, 0, sizeof r);
        return r;
    }
    pdf /= sd.area;
    r.pdf = pdf;
    r.finite = 1;
    material_t m = load_material(s, sd.flags_material >> 4);
    r.radiance = m_emission(&m);
    r.pos = sd.pos;
    r.normal = sd.gn;
    r.dir = sub(r.pos, ref);
    r.dist = length3(r.dir);
    r.dir = divf(r.dir, r.dist);
    r.cos_light = -dot(unpack_normal(r.normal), r.dir);
    if (r.cos_light <= 0) { memset(&r, 0, sizeof r); }
    return r;
}

/* ------------------------------------------------------------ NanoVDB */
/* Float-grid read path of NanoVDB 32.3 (extern/nanovdb/PNanoVDB.h:1237-1560, 1988-2058):
 * grid (672 B) -> tree (64 B) -> root (64 B + 32 B tiles) -> upper (32^3) -> lower (16^3) -> leaf (8^3). */
static uint32_t rd32(const uint8_t* b, uint64_t o) { uint32_t v; memcpy(&v, b + o, 4); return v; }
static uint64_t rd64(const uint8_t* b, uint64_t o) { uint64_t v; memcpy(&v, b + o, 8); return v; }
static float rdf(const uint8_t* b, uint64_t o) { float v; memcpy(&v, b + o, 4); return v; }
static uint64_t nvdb_root(const uint8_t* b) { return 672 + rd64(b, 672 + 24); }
/* Byte offset of the value that the read at (i, j, k) returns (a leaf voxel,
 * a tile of an internal node or of the root, or the background). */
static uint64_t nvdb_addr(const uint8_t* b, int i, int j, int k) {
    const uint64_t root = nvdb_root(b);
    const uint32_t ntiles = rd32(b, root + 24);
    const uint64_t key = (uint64_t)((uint32_t)k >> 12) | ((uint64_t)((uint32_t)j >> 12) << 21) | ((uint64_t)((uint32_t)i >> 12) << 42);
    for (uint32_t t = 0; t < ntiles; t++) {
        const uint64_t tile = root + 64 + 32 * (uint64_t)t;
        if (rd64(b, tile) != key) continue;
        const int64_t child = (int64_t)rd64(b, tile + 8);
        if (child == 0) return tile + 20;
        const uint64_t up = root + (uint64_t)child;
        const uint32_t n2 = (uint32_t)(((i & 4095) >> 7) << 10) + (uint32_t)(((j & 4095) >> 7) << 5) + (uint32_t)((k & 4095) >> 7);
        if (!((rd32(b, up + 4128 + 4 * (n2 >> 5)) >> (n2 & 31)) & 1)) return up + 8256 + 8 * (uint64_t)n2;
        const uint64_t lo = up + rd64(b, up + 8256 + 8 * (uint64_t)n2);
        const uint32_t n1 = (uint32_t)(((i & 127) >> 3) << 8) + (uint32_t)(((j & 127) >> 3) << 4) + (uint32_t)((k & 127) >> 3);
        if (!((rd32(b, lo + 544 + 4 * (n1 >> 5)) >> (n1 & 31)) & 1)) return lo + 1088 + 8 * (uint64_t)n1;
        const uint64_t leaf = lo + rd64(b, lo + 1088 + 8 * (uint64_t)n1);
        const uint32_t n0 = (uint32_t)((i & 7) << 6) + (uint32_t)((j & 7) << 3) + (uint32_t)(k & 7);
        return leaf + 96 + 4 * (uint64_t)n0;
    }
    return root + 28; /* background */
}
static real nvdb_read(const uint8_t* b, int i, int j, int k) { return rdf(b, nvdb_addr(b, i, j, k)); }
static f3 nvdb_map(const uint8_t* b, uint64_t off, f3 v, int translate, int inverse) {
    const uint64_t m = 296 + off;
    if (inverse && translate) v = v3(v.x - rdf(b, 296 + 72), v.y - rdf(b, 296 + 76), v.z - rdf(b, 296 + 80));
    f3 r = v3(v.x * rdf(b, m) + v.y * rdf(b, m + 4) + v.z * rdf(b, m + 8),
              v.x * rdf(b, m + 12) + v.y * rdf(b, m + 16) + v.z * rdf(b, m + 20),
              v.x * rdf(b, m + 24) + v.y * rdf(b, m + 28) + v.z * rdf(b, m + 32));
    if (!inverse && translate) r = v3(r.x + rdf(b, 296 + 72), r.y + rdf(b, 296 + 76), r.z + rdf(b, 296 + 80));
    return r;
}
static f3 nvdb_world_to_index(const uint8_t* b, f3 v) { return nvdb_map(b, 36, v, 1, 1); }
static f3 nvdb_world_to_index_dir(const uint8_t* b, f3 v) { return nvdb_map(b, 36, v, 0, 1); }
static f3 nvdb_index_to_world(const uint8_t* b, f3 v) { return nvdb_map(b, 0, v, 1, 0); }
static f3 nvdb_index_to_world_dir(const uint8_t* b, f3 v) { return nvdb_map(b, 0, v, 0, 0); }
float or_nvdb_read(const void* grid, int i, int j, int k) { return nvdb_read((const uint8_t*)grid, i, j, k); }
void or_nvdb_world_to_index(const void* grid, const float p[3], int dir, float out[3]) {
    const f3 r = dir ? nvdb_world_to_index_dir((const uint8_t*)grid, v3(p[0], p[1], p[2])) : nvdb_world_to_index((const uint8_t*)grid, v3(p[0], p[1], p[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* Medium record (Material.hpp:43-59), medium.hlsli:17-40 */
typedef struct {
    f3 density;
    uint32_t albedo_packed, dens_vol, alb_vol;
    real g;
} medium_t;
static medium_t load_medium(const or_scene* s, uint32_t addr) {
    medium_t m;
    memset(&m, 0, sizeof m);
    if ((uint64_t)addr / 4 + 7 > s->nmatw) return m;
    const uint32_t* w = s->mat + addr / 4;
    m.density = v3(asfloat(w[0]), asfloat(w[1]), asfloat(w[2]));
    m.albedo_packed = w[3];
    m.dens_vol = w[4];
    m.alb_vol = w[5];
    m.g = asfloat(w[6]);
    return m;
}
static f3 med_albedo(const medium_t* m) { return v3(un8(m->albedo_packed, 0), un8(m->albedo_packed, 8), un8(m->albedo_packed, 16)); }
static int med_can_eval(const medium_t* m) { return any_gt0(m->density) && m->albedo_packed > 0; }
static int med_singular(const medium_t* m) { return fabsr(m->g) > 0.999f; }
static real med_cont_prob(const medium_t* m) { return saturate(luminance(mul(med_albedo(m), m->density))); }
/* Henyey-Greenstein evaluateReflectance / sampleDirection, medium.hlsli:44-78 */
static real hg_eval(real g, f3 din, f3 dout) {
    /* pow(d, 1.5) as d * sqrt(d), as the GPU library takes it */
    const real d = 1 + g * g + 2 * g * dot(din, dout);
    return 1 / (4 * OR_PI) * (1 - g * g) / (d * sqrtr(d));
}
static f3 hg_sample(real g, real r0, real r1, f3 din) {
    if (fabsr(g) < 1e-4f) return sample_uniform_sphere(r0, r1);
    const real g2 = g * g;
    const real tmp = (g2 - 1) / (2 * r0 * g - (g + 1));
    const real ce = (tmp * tmp - (1 + g2)) / (2 * g);
    const real se = sqrtr(maxf(1 - ce * ce, 0));
    const real az = 2 * OR_PI * r1;
    f3 fr[3];
    make_orthonormal(din, fr);
    return add(add(mulf(fr[0], se * om_cos(az)), mulf(fr[1], se * om_sin(az))), mulf(din, ce));
}
float or_hg_eval(float g, const float din[3], const float dout[3]) { return hg_eval(g, v3(din[0], din[1], din[2]), v3(dout[0], dout[1], dout[2])); }
void or_hg_sample(float g, float r0, float r1, const float din[3], float out[3]) {
    const f3 r = hg_sample(g, r0, r1, v3(din[0], din[1], din[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
static f3 expv(f3 a) { return v3(om_exp(a.x), om_exp(a.y), om_exp(a.z)); }
static real avg3(f3 a) { return (a.x + a.y + a.z) / 3; }
/* deltaTrack<Scattering>, medium.hlsli:104-181.  origin/direction in object space. */
/* Density-gradient sink of the adjoint: at each collision in grid `vol`,
 * grad[value offset / 4] += sum_c wgt_c * dlog(beta_c) / d(value), where wgt
 * = dLoss/dL (per channel) times the radiance the collision's factor
 * multiplies. */
typedef struct { double* grad; uint32_t vol; f3 wgt; } adjd_t;

static f3 delta_track(const or_scene* s, const medium_t* m, int scattering, rng_t* rng, f3 origin, f3 dir, real tmax, f3* beta,
                      f3* dirPdf, f3* neePdf, int* scattered, uint32_t max_null, const adjd_t* ad) {
    *scattered = 0;
    *dirPdf = v3(1, 1, 1);
    *neePdf = v3(1, 1, 1);
    uint32_t u4[4];
    rng_next(rng, u4);
    const int ch = (int)(u4[0] % 3);
    if (m->dens_vol == 0xFFFFFFFFu || m->dens_vol >= s->nvols) {
        const f3 st = m->density;
        const real sc = fget(st, ch);
        if (sc < 1e-6f) return v3(0, 0, 0);
        real r[4];
        rng_nextf(rng, r);
        const real t = -om_log(1 - r[0]) / sc;
        if (t < tmax && scattering) {
            *scattered = 1;
            *beta = mul(*beta, mul(mul(expv(mulf(neg(st), t)), st), med_albedo(m)));
            const real e = om_exp(-sc * t) * sc;
            *dirPdf = mulf(*dirPdf, e);
            return add(origin, mulf(dir, t));
        }
        *beta = mul(*beta, expv(mulf(neg(st), tmax)));
        *dirPdf = mulf(*dirPdf, om_exp(-sc * tmax));
        return v3(0, 0, 0);
    }
    const uint8_t* g = s->vol[m->dens_vol];
    /* detached sampling: decisions, pdfs and the majorant read the grid, the
     * path weights read its weight copy (the same grid unless one is set) */
    const uint8_t* gw = s->volw[m->dens_vol] ? s->volw[m->dens_vol] : g;
    const int sink = ad && ad->grad && ad->vol == m->dens_vol;
    const real rmax = rdf(g, nvdb_root(g) + 36);
    const f3 maj = mulf(m->density, rmax);
    if (fget(maj, ch) < 1e-6f) return v3(0, 0, 0);
    origin = nvdb_world_to_index(g, origin);
    dir = nvdb_world_to_index_dir(g, dir);
    const real invc = 1 / fget(maj, ch);
    const real invmax = 1 / max3f(maj);
    for (uint32_t it = 0; it < max_null && any_gt0(*beta); it++) {
        real r[4];
        rng_nextf(rng, r);
        const real t = -om_log(1 - r[0]) * invc;
        if (t >= tmax) {
            const f3 tr = expv(mulf(neg(maj), tmax));
            *beta = mul(*beta, tr);
            *neePdf = mul(*neePdf, tr);
            *dirPdf = mul(*dirPdf, tr);
            break;
        }
        origin = add(origin, mulf(dir, t));
        tmax -= t;
        const uint64_t va = nvdb_addr(g, (int)floorr(origin.x), (int)floorr(origin.y), (int)floorr(origin.z));
        const real dv = rdf(g, va), dvw = rdf(gw, va);
        const f3 sig_t = mulf(m->density, dv), sig_tw = mulf(m->density, dvw);
        const f3 sig_s = med_albedo(m); /* readAlbedo is 1 without an albedo volume */
        const f3 tr = mulf(expv(mulf(neg(maj), t)), invmax);
        int real_coll = 0;
        if (scattering) {
            real r2[4];
            rng_nextf(rng, r2);
            real_coll = r2[0] < fget(sig_t, ch) * invc;
        }
        if (real_coll) {
            /* beta_c gains sigma_t,c = density_c * value: dlog = 1 / value */
            if (sink && dvw > 0) ad->grad[va / 4] += ((double)ad->wgt.x + (double)ad->wgt.y + (double)ad->wgt.z) / (double)dvw;
            *beta = mul(*beta, mul(mul(tr, sig_tw), sig_s));
            *dirPdf = mul(*dirPdf, mul(tr, sig_t));
            *scattered = 1;
            return nvdb_index_to_world(g, origin);
        }
        const f3 d = sub(maj, sig_t), dw = sub(maj, sig_tw);
        /* beta_c gains maj_c - density_c * value: dlog = -density_c / (maj_c - sigma_t,c) */
        if (sink) {
            double gsum = 0;
            if (dw.x != 0) gsum -= (double)ad->wgt.x * (double)m->density.x / (double)dw.x;
            if (dw.y != 0) gsum -= (double)ad->wgt.y * (double)m->density.y / (double)dw.y;
            if (dw.z != 0) gsum -= (double)ad->wgt.z * (double)m->density.z / (double)dw.z;
            ad->grad[va / 4] += gsum;
        }
        *beta = mul(*beta, mul(tr, dw));
        *dirPdf = mul(*dirPdf, mul(tr, d));
        *neePdf = mul(*neePdf, mul(tr, maj));
    }
    return v3(0, 0, 0);
}

/* ------------------------------------------------------------- BVH */
static real tri_c(const or_scene* s, uint32_t t, int ax) {
    const float* v = s->tv + 9 * (size_t)t;
    return (v[ax] + v[3 + ax] + v[6 + ax]) * (1.0f / 3.0f);
}
static void select_nth(const or_scene* s, uint32_t* a, int64_t n, int64_t k, int ax) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        real piv = tri_c(s, a[(lo + hi) / 2], ax);
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (tri_c(s, a[i], ax) < piv) i++;
            while (tri_c(s, a[j], ax) > piv) j--;
            if (i <= j) { uint32_t t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else break;
    }
}
static void node_bounds(const or_scene* s, bnode* nd, const uint32_t* a, uint32_t n) {
    for (int k = 0; k < 3; k++) { nd->lo[k] = INFINITY; nd->hi[k] = -INFINITY; }
    for (uint32_t i = 0; i < n; i++) {
        const float* v = s->tv + 9 * (size_t)a[i];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                nd->lo[k] = minf(nd->lo[k], v[3 * j + k]);
                nd->hi[k] = maxf(nd->hi[k], v[3 * j + k]);
            }
    }
    for (int k = 0; k < 3; k++) { nd->lo[k] = nextafterr(nd->lo[k], -INFINITY); nd->hi[k] = nextafterr(nd->hi[k], INFINITY); }
}
static void build_rec(or_scene* s, uint32_t ni, uint32_t start, uint32_t n) {
    bnode* nd = &s->nodes[ni];
    node_bounds(s, nd, s->order + start, n);
    if (n <= 4) { nd->a = start; nd->n = n; return; }
    real c_lo[3] = {INFINITY, INFINITY, INFINITY}, c_hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            real c = tri_c(s, s->order[start + i], k);
            c_lo[k] = minf(c_lo[k], c); c_hi[k] = maxf(c_hi[k], c);
        }
    int ax = 0;
    for (int k = 1; k < 3; k++) if (c_hi[k] - c_lo[k] > c_hi[ax] - c_lo[ax]) ax = k;
    select_nth(s, s->order + start, n, n / 2, ax);
    uint32_t c0 = s->nnodes;
    s->nnodes += 2;
    nd = &s->nodes[ni];
    nd->a = c0; nd->n = 0;
    build_rec(s, c0, start, n / 2);
    build_rec(s, c0 + 1, start + n / 2, n - n / 2);
}

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
 * Returns 1 with t, b1 (weight of v1), b2 (weight of v2) on a hit with t > 0. */
typedef struct { f3 o, d; int kx, ky, kz; real Sx, Sy, Sz; } wray_t;
static wray_t wray_make(f3 o, f3 d) {
    wray_t r;
    r.o = o; r.d = d;
    real ax = fabsr(d.x), ay = fabsr(d.y), az = fabsr(d.z);
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = (r.kz + 1) % 3; r.ky = (r.kx + 1) % 3;
    if (fget(d, r.kz) < 0) { int t = r.kx; r.kx = r.ky; r.ky = t; }
    r.Sx = fget(d, r.kx) / fget(d, r.kz);
    r.Sy = fget(d, r.ky) / fget(d, r.kz);
    r.Sz = 1.0f / fget(d, r.kz);
    return r;
}
static int wt_intersect(const wray_t* r, const float* v, real* t, real* b1, real* b2) {
    f3 A = sub(v3(v[0], v[1], v[2]), r->o), B = sub(v3(v[3], v[4], v[5]), r->o), C = sub(v3(v[6], v[7], v[8]), r->o);
    real Ax = fget(A, r->kx) - r->Sx * fget(A, r->kz), Ay = fget(A, r->ky) - r->Sy * fget(A, r->kz);
    real Bx = fget(B, r->kx) - r->Sx * fget(B, r->kz), By = fget(B, r->ky) - r->Sy * fget(B, r->kz);
    real Cx = fget(C, r->kx) - r->Sx * fget(C, r->kz), Cy = fget(C, r->ky) - r->Sy * fget(C, r->kz);
    real U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (real)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (real)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (real)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return 0;
    real det = U + V + W;
    if (det == 0.0f) return 0;
    real Az = r->Sz * fget(A, r->kz), Bz = r->Sz * fget(B, r->kz), Cz = r->Sz * fget(C, r->kz);
    real T = U * Az + V * Bz + W * Cz;
    real rcp = 1.0f / det;
    real tt = T * rcp;
    if (!(tt > 0.0f)) return 0;
    *t = tt; *b1 = V * rcp; *b2 = W * rcp;
    return 1;
}

typedef struct { real t, b1, b2; uint32_t id; int hit; } hit_t;

static int box_hit(const bnode* n, f3 o, f3 inv, real tmax, real* tnear) {
    real t0x = (n->lo[0] - o.x) * inv.x, t1x = (n->hi[0] - o.x) * inv.x;
    real t0y = (n->lo[1] - o.y) * inv.y, t1y = (n->hi[1] - o.y) * inv.y;
    real t0z = (n->lo[2] - o.z) * inv.z, t1z = (n->hi[2] - o.z) * inv.z;
    real tn = fmaxr(fmaxr(fminr(t0x, t1x), fminr(t0y, t1y)), fmaxr(fminr(t0z, t1z), 0.0f));
    real tf = fminr(fminr(fmaxr(t0x, t1x), fmaxr(t0y, t1y)), fminr(fmaxr(t0z, t1z), tmax));
    *tnear = tn;
    return tn <= tf * 1.00000024f;
}

/* closest (any=0) or any hit (any=1) in (0, tmax) */
/* raySphere (compat/common.h:124-135) about the origin */
static void ray_sphere(f3 o, f3 d, real r, real* t0, real* t1) {
    const real a = dot(d, d), b = dot(o, d);
    const f3 l = sub(mulf(o, a), mulf(d, b));
    real det = pow2(a * r) - dot(l, l);
    if (det < 0) { *t0 = *t1 = 0; return; }
    const real inv_a = 1 / a;
    det = sqrtr(det * inv_a) * inv_a;
    const real m = -b * inv_a;
    *t0 = m - det;
    *t1 = m + det;
}
/* the procedural sphere hits (intersection.hlsli:82-90) after the triangles:
 * the entry distance, or the exit one from inside, when it is in (0, h.t);
 * ties go to the smaller id */
static void trace_spheres(const or_scene* s, f3 o, f3 d, int any, hit_t* h) {
    for (uint32_t k = 0; k < s->nsph; k++) {
        const uint32_t ii = s->sph_inst[k];
        const f3 oo = xf_point(&s->inv_xf[ii], o), od = xf_vector(&s->inv_xf[ii], d);
        real t0, t1;
        ray_sphere(oo, od, asfloat(s->inst[ii].data), &t0, &t1);
        if (!(t0 < t1)) continue;
        const real t = t0 > 0 ? t0 : t1;
        const uint32_t id = ii | 0xffff0000u;
        if (!(t > 0) || !(t < h->t || (t == h->t && h->hit && id < h->id))) continue;
        h->t = t; h->b1 = h->b2 = 0; h->id = id; h->hit = 1;
        if (any) return;
    }
}
static hit_t trace_tris(const or_scene* s, f3 o, f3 d, real tmax, int any, uint64_t* nodes, uint64_t* tris);
static hit_t trace(const or_scene* s, f3 o, f3 d, real tmax, int any, uint64_t* nodes, uint64_t* tris) {
    hit_t h = trace_tris(s, o, d, tmax, any, nodes, tris);
    if (s->nsph && !(any && h.hit)) trace_spheres(s, o, d, any, &h);
    return h;
}
/* the shading data of a hit of the ray (o, d): a triangle at its
 * barycentrics, a sphere at the object-space point o' + d' t
 * (intersection.hlsli:178-181) */
static void make_hit_shading(const or_scene* s, const hit_t* h, f3 o, f3 d, shading_t* r) {
    const uint32_t ii = h->id & 0xffffu, prim = h->id >> 16;
    if (prim == 0xFFFFu) {
        const f3 oo = xf_point(&s->inv_xf[ii], o), od = xf_vector(&s->inv_xf[ii], d);
        make_sphere_shading(s, ii, add(oo, mulf(od, h->t)), r);
        return;
    }
    make_tri_shading(s, ii, prim, h->b1, h->b2, r);
}
/* The alpha test of a candidate triangle hit (intersection.hlsli:125-147):
 * a triangle whose material names an alpha mask (material word 12, the
 * cutoff in word 13, material.hlsli:72-76) is committed only where channel 0
 * of the mask, sampled bilinearly at level 0 at the candidate's texture
 * coordinate, reaches the cutoff. */
static int alpha_pass(const or_scene* s, uint32_t id, real b1, real b2) {
    const st_instance* in = &s->inst[id & 0xffffu];
    const uint32_t a = (in->type_material_address >> 4) / 4, mask = s->mat[a + 12];
    if (mask >= ST_IMAGE_LIMIT) return 1;
    float cutoff;
    memcpy(&cutoff, &s->mat[a + 13], 4);
    const st_mesh_vertex_info* vi = &s->vi[in->data & 0xffffu];
    uint32_t tri[3] = {0, 0, 0};
    load_tri(s, vi, id >> 16, tri);
    real t0[2] = {0, 0}, t1[2] = {0, 0}, t2[2] = {0, 0};
    const uint32_t tb = vi_field(vi->packed_buffer_indices[1], 16, 16), ts = vi_field(vi->packed_strides, 24, 8);
    if (tb < ST_IMAGE_LIMIT) {
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[0], t0);
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[1], t1);
        load_f2(s, tb, vi->packed_offsets[3], ts, tri[2], t2);
    }
    const real u = t0[0] + (t1[0] - t0[0]) * b1 + (t2[0] - t0[0]) * b2;
    const real v = t0[1] + (t1[1] - t0[1]) * b1 + (t2[1] - t0[1]) * b2;
    real o[4];
    bilinear4(&s->tex[mask].lv[0], u, v, o);
    return o[0] >= (real)cutoff;
}

static hit_t trace_tris(const or_scene* s, f3 o, f3 d, real tmax, int any, uint64_t* nodes, uint64_t* tris) {
    hit_t h;
    h.hit = 0; h.t = tmax; h.b1 = h.b2 = 0; h.id = 0xffffffffu;
    if (s->nnodes == 0) return h;
    wray_t wr = wray_make(o, d);
    f3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const bnode* n = &s->nodes[stack[--sp]];
        real tn;
        (*nodes)++;
        if (!box_hit(n, o, inv, h.t, &tn)) continue;
        if (n->n > 0) {
            for (uint32_t i = 0; i < n->n; i++) {
                uint32_t ti = s->order[n->a + i];
                real t, b1, b2;
                (*tris)++;
                if (!wt_intersect(&wr, s->tv + 9 * (size_t)ti, &t, &b1, &b2)) continue;
                uint32_t id = s->tid[ti];
                if (t < h.t || (t == h.t && h.hit && id < h.id)) {
                    if (s->has_alpha && !alpha_pass(s, id, b1, b2)) continue;
                    h.t = t; h.b1 = b1; h.b2 = b2; h.id = id; h.hit = 1;
                    if (any) return h;
                }
            }
        } else {
            stack[sp++] = n->a + 1;
            stack[sp++] = n->a;
        }
    }
    return h;
}

/* traceRay over triangles and the procedural volume boxes, intersection.hlsli:68-193.
 * *vol_gn receives the packed box normal when the closest hit is a volume box. */
static hit_t trace_ray(const or_scene* s, f3 o, f3 d, real tmax, int any, int* is_vol, uint32_t* vol_gn, uint64_t* nodes,
                       uint64_t* tris) {
    hit_t h = trace(s, o, d, tmax, any, nodes, tris);
    *is_vol = 0;
    if (any && h.hit) return h;
    for (uint32_t v = 0; v < s->nvolinst; v++) {
        const uint32_t ii = s->volinst[v];
        const f3 oo = xf_point(&s->inv_xf[ii], o), od = xf_vector(&s->inv_xf[ii], d);
        const uint32_t vidx = s->inst[ii].data;
        f3 org, dir, bmin, bmax;
        const uint8_t* g = NULL;
        if (vidx == 0xFFFFFFFFu || vidx >= s->nvols) {
            org = oo; dir = od; bmin = v3(-1, -1, -1); bmax = v3(1, 1, 1);
        } else {
            g = s->vol[vidx];
            const uint64_t root = nvdb_root(g);
            org = nvdb_world_to_index(g, oo);
            dir = nvdb_world_to_index_dir(g, od);
            bmin = v3((real)(int32_t)rd32(g, root), (real)(int32_t)rd32(g, root + 4), (real)(int32_t)rd32(g, root + 8));
            bmax = add(v3((real)(int32_t)rd32(g, root + 12), (real)(int32_t)rd32(g, root + 16), (real)(int32_t)rd32(g, root + 20)),
                       v3(1, 1, 1));
        }
        const f3 t0 = divv(sub(bmin, org), dir), t1 = divv(sub(bmax, org), dir);
        const real sx = maxf(maxf(minf(t0.x, t1.x), minf(t0.y, t1.y)), minf(t0.z, t1.z));
        const real sy = minf(minf(maxf(t0.x, t1.x), maxf(t0.y, t1.y)), maxf(t0.z, t1.z));
        if (!(sx <= sy)) continue; /* the box itself is missed */
        const real t = sx > 0 ? sx : sy;
        if (!(t < h.t && t > 0)) continue;
        f3 ln = sub(v3(t1.x == t, t1.y == t, t1.z == t), v3(t0.x == t, t0.y == t, t0.z == t));
        if (g) ln = nvdb_index_to_world_dir(g, ln);
        *vol_gn = pack_normal(normalize3(xf_vector(&s->xf[ii], ln)));
        h.t = t;
        h.id = ii | (0xFFFFu << 16);
        h.hit = 1;
        h.b1 = h.b2 = 0;
        *is_vol = 1;
    }
    return h;
}

/* TransformData::inverse of a 3x4 affine transform, computed in double. */
static st_transform inverse_xf(const st_transform* T) {
    st_transform r;
    memset(&r, 0, sizeof r);
    double a[3][3], inv[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = T->m[i][j];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (det == 0) return r;
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det; inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det; inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det; inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det; inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int i = 0; i < 3; i++) {
        double tr = 0;
        for (int j = 0; j < 3; j++) { r.m[i][j] = (real)inv[i][j]; tr -= inv[i][j] * T->m[j][3]; }
        r.m[i][3] = (real)tr;
    }
    return r;
}

or_scene* or_scene_create(const st_scene_desc* d) {
    if (!d || (d->instance_count && (!d->instances || !d->transforms))) { fail("invalid scene description"); return NULL; }
    or_scene* s = (or_scene*)calloc(1, sizeof *s);
    s->ninst = d->instance_count;
    s->inst = (st_instance*)malloc(sizeof(st_instance) * (s->ninst + 1));
    s->xf = (st_transform*)malloc(sizeof(st_transform) * (s->ninst + 1));
    memcpy(s->inst, d->instances, sizeof(st_instance) * s->ninst);
    memcpy(s->xf, d->transforms, sizeof(st_transform) * s->ninst);
    s->nmatw = d->material_word_count;
    s->mat = (uint32_t*)malloc(4 * (s->nmatw + 16));
    memcpy(s->mat, d->material_data, 4 * (size_t)s->nmatw);
    s->nvi = d->mesh_vertex_info_count;
    s->vi = (st_mesh_vertex_info*)malloc(sizeof(st_mesh_vertex_info) * (s->nvi + 1));
    memcpy(s->vi, d->mesh_vertex_info, sizeof(st_mesh_vertex_info) * s->nvi);
    s->nvb = d->vertex_buffer_count;
    s->vb = (uint8_t**)calloc(s->nvb + 1, sizeof(uint8_t*));
    s->vbsz = (uint64_t*)calloc(s->nvb + 1, sizeof(uint64_t));
    for (uint32_t i = 0; i < s->nvb; i++) {
        s->vbsz[i] = d->vertex_buffers[i].size_bytes;
        s->vb[i] = (uint8_t*)malloc(s->vbsz[i] + 16);
        memcpy(s->vb[i], d->vertex_buffers[i].data, s->vbsz[i]);
    }
    /* environment record: EnvironmentMap::store, src/App/Scene.hpp:42-45 */
    s->has_env = d->environment_material_address != 0xFFFFFFFFu;
    if (s->has_env) {
        uint32_t a = d->environment_material_address / 4;
        if (a + 4 > s->nmatw) { fail("environment address out of range"); or_scene_destroy(s); return NULL; }
        s->env_addr = d->environment_material_address;
        s->env_emission = v3(asfloat(s->mat[a]), asfloat(s->mat[a + 1]), asfloat(s->mat[a + 2]));
        uint32_t img = s->mat[a + 3];
        if (img < ST_IMAGE_LIMIT) {
            if (img >= d->image_count) { fail("environment image index out of range"); or_scene_destroy(s); return NULL; }
            const st_image_desc* im = &d->images[img];
            s->env_has_image = 1;
            s->lv[0].w = (int)im->width; s->lv[0].h = (int)im->height;
            s->lv[0].rgba = (float*)malloc(sizeof(float) * 4 * (size_t)im->width * im->height);
            memcpy(s->lv[0].rgba, im->rgba, sizeof(float) * 4 * (size_t)im->width * im->height);
            build_mips(s);
            if (build_alias(s)) { or_scene_destroy(s); fail("out of memory"); return NULL; }
        }
    }
    /* media: volume instances, their grids and the inverse transforms (TransformData::inverse) */
    s->inv_xf = (st_transform*)calloc(s->ninst + 1, sizeof(st_transform));
    for (uint32_t i = 0; i < s->ninst; i++) s->inv_xf[i] = inverse_xf(&s->xf[i]);
    s->nvols = d->volume_count < 8 ? d->volume_count : 8;
    for (uint32_t v = 0; v < s->nvols; v++) {
        s->volsz[v] = d->volumes[v].size_bytes;
        s->vol[v] = (uint8_t*)malloc(s->volsz[v] + 16);
        memcpy(s->vol[v], d->volumes[v].data, s->volsz[v]);
    }
    for (uint32_t i = 0; i < s->ninst && s->nvolinst < 8; i++)
        if ((s->inst[i].type_material_address & 0xfu) == 2) s->volinst[s->nvolinst++] = i;
    s->nvolinfo = d->volume_info_count;
    s->volinfo = (st_volume_info*)calloc(s->nvolinfo + 1, sizeof(st_volume_info));
    if (s->nvolinfo) memcpy(s->volinfo, d->volume_info, sizeof(st_volume_info) * s->nvolinfo);
    /* lights, scene AABB and flattened triangles */
    s->light_inst = (uint32_t*)malloc(sizeof(uint32_t) * (s->ninst + 1));
    s->sph_inst = (uint32_t*)malloc(sizeof(uint32_t) * (s->ninst + 1));
    real amin[3] = {INFINITY, INFINITY, INFINITY}, amax[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint64_t ntri = 0;
    for (uint32_t i = 0; i < s->ninst; i++) {
        if ((s->inst[i].type_material_address & 0xfu) == 0) ntri += s->inst[i].data >> 16;
    }
    s->tv = (float*)malloc(sizeof(float) * 9 * (ntri + 1));
    s->tid = (uint32_t*)malloc(sizeof(uint32_t) * (ntri + 1));
    s->tri_base = (uint32_t*)calloc(s->ninst + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < s->ninst; i++) {
        const st_instance* in = &s->inst[i];
        s->tri_base[i] = s->ntris;
        uint32_t type = in->type_material_address & 0xfu;
        if (type == 2) { /* volume box corners, Scene.cpp:700-711 */
            f3 mn = v3(-1, -1, -1), mx = v3(1, 1, 1);
            if (in->data < s->nvols) {
                double b[6];
                memcpy(b, s->vol[in->data] + 560, 48);
                mn = v3((real)b[0], (real)b[1], (real)b[2]);
                mx = v3((real)b[3], (real)b[4], (real)b[5]);
            }
            /* as the reference does, the running min/max also feed the later
             * corners (Scene.cpp:701-709 updates mn/mx inside the corner loop) */
            for (int c = 0; c < 8; c++) {
                f3 corner = v3((c % 2) == 0 ? mn.x : mx.x, ((c % 4) / 2) == 0 ? mn.y : mx.y, (c / 4) == 0 ? mn.z : mx.z);
                corner = xf_point(&s->xf[i], corner);
                mn = v3(minf(mn.x, corner.x), minf(mn.y, corner.y), minf(mn.z, corner.z));
                mx = v3(maxf(mx.x, corner.x), maxf(mx.y, corner.y), maxf(mx.z, corner.z));
            }
            amin[0] = minf(amin[0], mn.x); amin[1] = minf(amin[1], mn.y); amin[2] = minf(amin[2], mn.z);
            amax[0] = maxf(amax[0], mx.x); amax[1] = maxf(amax[1], mx.y); amax[2] = maxf(amax[2], mx.z);
        }
        if (type == 1) { /* sphere: lights, the transformed corners of [-r, r]^3 */
            const real rad = asfloat(in->data);
            if (!(rad > 0) || !isfinite(rad)) { fail("sphere radius must be positive and finite"); or_scene_destroy(s); return NULL; }
            if ((uint64_t)(in->type_material_address >> 4) / 4 + 9 > s->nmatw) { fail("material address out of range"); or_scene_destroy(s); return NULL; }
            material_t m = load_material(s, in->type_material_address >> 4);
            if (any_gt0(v3(un8(m.p[1], 0), un8(m.p[1], 8), un8(m.p[1], 16)))) s->light_inst[s->nlights++] = i;
            s->sph_inst[s->nsph++] = i;
            for (int c = 0; c < 8; c++) {
                f3 corner = v3((c % 2) == 0 ? -rad : rad, ((c % 4) / 2) == 0 ? -rad : rad, (c / 4) == 0 ? -rad : rad);
                corner = xf_point(&s->xf[i], corner);
                amin[0] = minf(amin[0], corner.x); amin[1] = minf(amin[1], corner.y); amin[2] = minf(amin[2], corner.z);
                amax[0] = maxf(amax[0], corner.x); amax[1] = maxf(amax[1], corner.y); amax[2] = maxf(amax[2], corner.z);
            }
            continue;
        }
        if (type != 0) continue;
        material_t m = load_material(s, in->type_material_address >> 4);
        f3 e = v3(un8(m.p[1], 0), un8(m.p[1], 8), un8(m.p[1], 16));
        if (any_gt0(e)) s->light_inst[s->nlights++] = i;
        uint32_t vii = in->data & 0xffffu, pc = in->data >> 16;
        if (vii >= s->nvi) { fail("vertex info index out of range"); or_scene_destroy(s); return NULL; }
        const st_mesh_vertex_info* vi = &s->vi[vii];
        uint32_t pb = vi_field(vi->packed_buffer_indices[0], 16, 16), ps = vi_field(vi->packed_strides, 8, 8);
        real lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t p = 0; p < pc; p++) {
            uint32_t tri[3];
            if (!load_tri(s, vi, p, tri)) { fail("index buffer out of range"); or_scene_destroy(s); return NULL; }
            for (int j = 0; j < 3; j++) {
                f3 v;
                if (!load_f3(s, pb, vi->packed_offsets[1], ps, tri[j], &v)) { fail("position buffer out of range"); or_scene_destroy(s); return NULL; }
                lo[0] = minf(lo[0], v.x); lo[1] = minf(lo[1], v.y); lo[2] = minf(lo[2], v.z);
                hi[0] = maxf(hi[0], v.x); hi[1] = maxf(hi[1], v.y); hi[2] = maxf(hi[2], v.z);
                f3 w = xf_point(&s->xf[i], v);
                s->tv[9 * (size_t)s->ntris + 3 * j] = w.x;
                s->tv[9 * (size_t)s->ntris + 3 * j + 1] = w.y;
                s->tv[9 * (size_t)s->ntris + 3 * j + 2] = w.z;
            }
            s->tid[s->ntris] = i | (p << 16);
            s->ntris++;
        }
        if (pc == 0) continue;
        /* transformed corners of the mesh AABB, Scene.cpp:622-631 */
        for (int c = 0; c < 8; c++) {
            f3 corner = v3((c % 2) == 0 ? lo[0] : hi[0], ((c % 4) / 2) == 0 ? lo[1] : hi[1], (c / 4) == 0 ? lo[2] : hi[2]);
            corner = xf_point(&s->xf[i], corner);
            amin[0] = minf(amin[0], corner.x); amin[1] = minf(amin[1], corner.y); amin[2] = minf(amin[2], corner.z);
            amax[0] = maxf(amax[0], corner.x); amax[1] = maxf(amax[1], corner.y); amax[2] = maxf(amax[2], corner.z);
        }
    }
    if (s->nsph && s->nvolinst) { fail("sphere instances together with media are not supported"); or_scene_destroy(s); return NULL; }
    /* material images: base colour, packed words 2 and 3 (Material::store's
     * image indices) and the alpha masks of mesh materials (word 12); emission
     * images and normal maps are refused */
    s->tex = (struct tex_s*)calloc(d->image_count + 1, sizeof(struct tex_s));
    for (uint32_t i = 0; i < s->ninst; i++) {
        const uint32_t type = s->inst[i].type_material_address & 0xfu;
        if (type > 1) continue;
        const uint32_t a = (s->inst[i].type_material_address >> 4) / 4;
        if ((uint64_t)a + 14 > s->nmatw) continue;
        if (s->mat[a + 5] < ST_IMAGE_LIMIT || s->mat[a + 9] < ST_IMAGE_LIMIT) {
            fail("emission images and normal maps are not supported"); or_scene_destroy(s); return NULL;
        }
        if (s->mat[a + 12] < ST_IMAGE_LIMIT && type != 0) {
            fail("alpha masks apply to mesh materials only"); or_scene_destroy(s); return NULL;
        }
        if (s->mat[a + 12] < ST_IMAGE_LIMIT) s->has_alpha = 1;
        for (int k = 4; k < 13; k++) {
            if (k == 8) k = 12;
            const uint32_t img = s->mat[a + k];
            if (img >= ST_IMAGE_LIMIT) continue;
            if (img >= d->image_count) { fail("material image index out of range"); or_scene_destroy(s); return NULL; }
            struct tex_s* t = &s->tex[img];
            if (t->n) continue;
            const st_image_desc* im = &d->images[img];
            t->lv[0].w = (int)im->width; t->lv[0].h = (int)im->height;
            t->lv[0].rgba = (float*)malloc(sizeof(float) * 4 * (size_t)im->width * im->height);
            memcpy(t->lv[0].rgba, im->rgba, sizeof(float) * 4 * (size_t)im->width * im->height);
            t->n = build_level_chain(t->lv);
            s->ntex = d->image_count;
        }
    }
    /* scene sphere, TestRenderer.cpp:458-461 */
    f3 cen = divf(add(v3(amax[0], amax[1], amax[2]), v3(amin[0], amin[1], amin[2])), 2);
    s->sphere[0] = cen.x; s->sphere[1] = cen.y; s->sphere[2] = cen.z;
    s->sphere[3] = length3(sub(v3(amax[0], amax[1], amax[2]), cen));
    if (s->ntris > 0) {
        s->order = (uint32_t*)malloc(sizeof(uint32_t) * s->ntris);
        for (uint32_t i = 0; i < s->ntris; i++) s->order[i] = i;
        s->nodes = (bnode*)malloc(sizeof(bnode) * 2 * (size_t)s->ntris);
        s->nnodes = 1;
        build_rec(s, 0, 0, s->ntris);
    }
    return s;
}

void or_scene_destroy(or_scene* s) {
    if (!s) return;
    free(s->inst); free(s->xf); free(s->mat); free(s->vi);
    for (uint32_t i = 0; i < s->nvb; i++) free(s->vb[i]);
    free(s->vb); free(s->vbsz);
    for (int l = 0; l < 16; l++) free(s->lv[l].rgba);
    free(s->alias_prob); free(s->alias_idx); free(s->alias_pdf);
    free(s->light_inst); free(s->tv); free(s->tid); free(s->nodes); free(s->order);
    free(s->inv_xf); free(s->volinfo); free(s->sph_inst);
    if (s->tex)
        for (uint32_t i = 0; i < s->ntex; i++)
            for (int l = 0; l < 16; l++) free(s->tex[i].lv[l].rgba);
    free(s->tex);
    for (uint32_t v = 0; v < s->nvols; v++) { free(s->vol[v]); free(s->volw[v]); }
    free(s->tri_alb); free(s->tri_base);
    free(s);
}

int or_trace_rays(or_scene* s, const float* rays, uint32_t count, int closest, uint32_t* hits) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        const float* r = rays + 8 * i;
        uint64_t a = 0, b = 0;
        hit_t h = trace(s, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], !closest, &a, &b);
        if (closest) {
            hits[4 * i + 0] = h.hit ? asuint(h.t) : asuint(r[3]);
            hits[4 * i + 1] = h.hit ? h.id : 0xffffffffu;
            hits[4 * i + 2] = asuint(h.b1);
            hits[4 * i + 3] = asuint(h.b2);
        } else {
            hits[i] = (uint32_t)h.hit;
        }
    }
    return 0;
}

int or_trace_rays_bruteforce_f64(or_scene* s, const float* rays, uint32_t count, uint32_t* hits) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        const float* r = rays + 8 * i;
        double o[3] = {r[0], r[1], r[2]}, dd[3] = {r[4], r[5], r[6]};
        double best = r[3];
        uint32_t bid = 0xffffffffu;
        for (uint32_t t = 0; t < s->ntris; t++) {
            const float* v = s->tv + 9 * (size_t)t;
            double e1[3], e2[3], pv[3], tv[3], qv[3];
            for (int k = 0; k < 3; k++) { e1[k] = (double)v[3 + k] - v[k]; e2[k] = (double)v[6 + k] - v[k]; tv[k] = o[k] - v[k]; }
            pv[0] = dd[1] * e2[2] - dd[2] * e2[1]; pv[1] = dd[2] * e2[0] - dd[0] * e2[2]; pv[2] = dd[0] * e2[1] - dd[1] * e2[0];
            double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if (det == 0) continue;
            double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) / det;
            if (u < 0 || u > 1) continue;
            qv[0] = tv[1] * e1[2] - tv[2] * e1[1]; qv[1] = tv[2] * e1[0] - tv[0] * e1[2]; qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
            double vv = (dd[0] * qv[0] + dd[1] * qv[1] + dd[2] * qv[2]) / det;
            if (vv < 0 || u + vv > 1) continue;
            double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) / det;
            if (tt > 0 && tt < best) { best = tt; bid = s->tid[t]; }
        }
        hits[4 * i + 0] = asuint((real)best);
        hits[4 * i + 1] = bid;
        hits[4 * i + 2] = 0;
        hits[4 * i + 3] = 0;
    }
    return 0;
}

/* ------------------------------------------------------ path tracer */
typedef struct {
    f3 origin, thr;
    uint32_t dir;       /* packed oct-snorm16 (testrenderer.slang:119-122) */
    uint32_t path_len;  /* mPathLength */
    uint32_t inst_prim; /* mInstancePrimitiveIndex */
    real t16, pdfa16, dirpdf16; /* FP16-stored fields, kept as their f32 values */
    rng_t rng;
    uint32_t medium;    /* mCurrentMedium */
    real dvc, dvcm;    /* dVC, dVCM (FP16-stored, kept as their f32 values) */
    int vc_f32;         /* ST_FLAG_VC_FP32_MIS: dVC, dVCM kept in fp32 */
    int flag31;         /* isSpecular (camera paths) / isFiniteLight (light paths): one bit */
} pstate_t; /* PathState, testrenderer.slang:66-125 */

/* a dVC / dVCM store: fp16 as the reference's PathState (testrenderer.slang:75-82) */
static real qvc(const pstate_t* ps, real f) { return ps->vc_f32 ? f : q16(f); }

/* PackedLightVertex, testrenderer.slang:45-64 */
typedef struct {
    f3 lpos;            /* object-space position */
    uint32_t inst_prim;
    f3 thr;
    uint32_t dirin;     /* packed local incoming direction */
    uint32_t plen;
    real dvc, dvcm;
} lvert_t;

/* ShadowRay (testrenderer.slang:127-142) + the vertex medium and the key of
 * its own random stream for ratio tracking */
typedef struct { f3 o, d, c; real tmax, w; uint32_t medium, plen; } shadow_t;

typedef struct {
    const or_scene* s;
    const cfg_t* c;
    pstate_t ps;
    f3 out;           /* mOutput.rgb of this launch */
    uint32_t nee[3];  /* mOutputAtomic */
    shadow_t sh[64];
    int nsh;
    uint32_t vis[2];
    uint64_t rays[2], trav[4];
    uint32_t view_medium;
    /* auxiliary outputs of the first vertex (StoreAuxillaryData) */
    const st_view* view;
    st_transform inv_cam;
    uint32_t px, py, W, H;
    real alb[3], dep[2], uv[2];
    /* bidirectional mode: this sample's light vertex cache and subpath count */
    const lvert_t* lvc;
    uint32_t lvc_count, L;
    real ipd;
    const uint32_t* light_acc;  /* this pixel's light-pass connections (3 x u32) or NULL */
    /* screen-space ReSTIR DI reuse: the primary vertex keeps its RIS reservoir
     * (seed, target, W, M, hit distance) and shading record instead of making
     * a shadow ray */
    int reuse;
    int prim_hit;               /* the primary ray hit a surface */
    shading_t prim_sd;
    uint32_t prim_dir;          /* packed primary ray direction */
    real res[8];               /* seed[4], target, W, M, depth */
    struct adj_s* adj;          /* path replay of or_adjoint, or NULL */
    struct cache_s* cache;      /* or_render_cached: this path's cache-depth vertex, or NULL */
    /* hash-grid reservoir reuse (ST_FLAG_RESTIR_DI_HASHGRID): the previous
     * sample's grid (NULL for a call's first sample), this sample's, and this
     * path's pixel */
    const struct hgrid_s* hg_prev;
    struct hgrid_s* hg_cur;
    uint32_t hg_pix;
} pctx_t;

/* A sample's hash grid (HashGrid<HashGridData>, hashgrid.hlsli:14-120): the
 * entries of the first hg_dmax reservoir vertices of each path at slot
 * vertex * W * H + pixel (16 floats: {object-space position, instance |
 * primitive << 16}, light-sample seed, {target, W, M, cell checksum},
 * {local incoming direction, 0}; W = 0: empty), then the slots sorted by
 * (empty, checksum), stable in slot order, with their keys. */
typedef struct hgrid_s {
    real* rec;
    uint64_t* keys;
    uint32_t* slots;
    uint32_t nslots, npix, dmax;
    real cell;
} hgrid_t;

/* or_render_cached (st_render_cached without the network): a path reaching a
 * scattering vertex at path length `depth` either ends there in a cache query
 * or, as the training path of its run of `stride` paths, goes on unchanged
 * and records its training row.  g = sample * W * H + y * W + x. */
typedef struct cache_s {
    uint32_t depth, stride, frame, g;
    int kind;                /* 0 no record, 1 cache query, 2 training row */
    real q[12], thr[3];
    f3 out_before;           /* fp32 radiance (emission) up to and including the vertex */
    int nsh_before;          /* shadow rays made before the vertex */
    uint32_t nee_before[3];  /* their fixed-point NEE */
} cache_t;

/* Path replay (or_adjoint).  Both passes trace each shadow ray right after
 * the vertex that made it and add its fp32 contribution to acc, in the order
 * the GPU wavefront adds them (emission, then that vertex's NEE); pass 1 ends
 * with total = acc, pass 2 replays the path and differentiates: the radiance
 * still to come at any point is total - acc. */
typedef struct adj_s {
    int pass;            /* 1: totals, 2: gradients */
    f3 acc, total;
    f3 g;                /* dLoss/dL of this path: d_hdr / spp */
    double* grad_alb;    /* 3 per flattened triangle */
    double* grad_grid;   /* per 4-byte word of grid grad_vol */
    uint32_t grad_vol;
} adj_t;

/* ViewData::toRaster (compat/scene.h:158-165) with ProjectionData::projectPoint
 * (compat/transform.h:139-155); 0, with *ox, *oy unchanged, when p is not visible. */
static int to_raster(const st_view* V, f3 p, real* ox, real* oy) {
    const st_projection* pr = &V->projection;
    const real sz = p.z > 0 ? 1.0f : (p.z < 0 ? -1.0f : 0.0f);
    const real sn = pr->near_plane > 0 ? 1.0f : (pr->near_plane < 0 ? -1.0f : 0.0f);
    if (sz != sn) return 0;
    real r0, r1, r2, r3;
    if (pr->vertical_fov < 0) {
        r0 = p.x * pr->scale[0] + pr->offset[0];
        r1 = p.y * pr->scale[1] + pr->offset[1];
        r2 = (p.z - pr->far_plane) / (pr->near_plane - pr->far_plane);
        r3 = 1;
    } else {
        r0 = p.x * pr->scale[0] + p.z * pr->offset[0];
        r1 = p.y * pr->scale[1] + p.z * pr->offset[1];
        r2 = fabsr(pr->near_plane);
        r3 = p.z * sn;
    }
    const real aw = fabsr(r3);
    if (fabsr(r0) >= aw || fabsr(r1) >= aw || fabsr(r2) >= aw) return 0;
    const real ew = (real)(V->image_max[0] - V->image_min[0]), eh = (real)(V->image_max[1] - V->image_min[1]);
    *ox = (real)V->image_min[0] + ew * ((r0 / r3) * 0.5f + 0.5f);
    *oy = (real)V->image_min[1] + eh * ((-r1 / r3) * 0.5f + 0.5f);
    return 1;
}

/* StoreAuxillaryData (testrenderer.slang:195-218) for a static scene in steady
 * state: the instance motion transform is the identity and the previous view
 * is this view.  has_instance: the vertex is a surface or medium vertex. */
static void store_aux(pctx_t* x, int has_instance, f3 pos, f3 albedo) {
    real pcx = (real)x->px + 0.5f, pcy = (real)x->py + 0.5f;
    real prev_depth = x->ps.t16;
    if (has_instance) {
        const f3 pc = xf_point(&x->inv_cam, pos);
        prev_depth = length3(pc);
        to_raster(x->view, pc, &pcx, &pcy);
    }
    x->alb[0] = albedo.x; x->alb[1] = albedo.y; x->alb[2] = albedo.z;
    x->dep[0] = x->ps.t16;
    x->dep[1] = prev_depth;
    x->uv[0] = pcx / (real)x->W;
    x->uv[1] = pcy / (real)x->H;
}

/* AddRadiance, testrenderer.slang:220-235 */
static void add_radiance(pctx_t* x, f3 rad, real w) {
    if (w <= 0 || all_le0(rad)) return;
    x->out = add(x->out, mulf(rad, w));
    if (x->adj) x->adj->acc = add(x->adj->acc, mulf(rad, w));
}
/* AddEmission, testrenderer.slang:788-808 (+ --minPathLength gate, vcm.slang:960) */
static void add_emission(pctx_t* x, f3 e, real directPdfA, real emissionPdfW, real WtoA) {
    real w = 1;
    if (x->c->flags & ST_FLAG_BDPT_LVC) {
        /* partial eye sub-path MIS weight [tech. rep. (43)] */
        if (x->ps.path_len > 1 && !x->ps.flag31) {
            const real wc = mis1(directPdfA) * x->ps.dvcm + mis1(emissionPdfW) * x->ps.dvc;
            w = 1 / (1 + wc);
        }
    } else if (x->ps.path_len > 1) { /* isSpecular is never set on camera paths without VC */
        if ((x->c->flags & ST_FLAG_SAMPLE_DIRECT) && x->ps.dirpdf16 > 0) w = mis2(x->ps.dirpdf16 * WtoA, directPdfA);
    }
    if (x->ps.path_len < x->c->min_len) return;
    add_radiance(x, e, w);
}
/* ConnectToLight (plain NEE), testrenderer.slang:751-779 */
static void connect_to_light(pctx_t* x, const material_t* m, const shading_t* sd, f3 ldin) {
    real rnd[4];
    rng_nextf(&x->ps.rng, rnd);
    isr_t r = sample_illumination(x->s, x->c, rnd, sd->pos);
    if (all_le0(r.radiance) || r.pdf <= 0) return;
    real AtoW = 1;
    if (r.finite) AtoW = pow2(r.dist) / r.cos_light;
    real dpw = r.pdf * AtoW;
    f3 ldout = to_local(sd, r.dir);
    rer_t f = bsdf_eval(m, x->c->lambert, ldin, ldout);
    f.f = mulf(f.f, shading_normal_correction(sd, ldin, ldout));
    real cp = rr_cp(x->c, x->ps.path_len, m_cont_prob(m));
    f.fwd *= cp; f.rev *= cp;
    f3 contrib = divf(mul(mul(x->ps.thr, r.radiance), f.f), dpw);
    real w = (x->c->flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(r.pdf, f.fwd / AtoW); /* ComputeMIS_DI :530-566 */
    if (x->c->flags & ST_FLAG_BDPT_LVC) {
        /* ComputeMIS_DI with gUseVC [tech. rep. (44), (45)] */
        const real epw = r.pdf * cos_hemisphere_pdfw(r.cos_light);
        const real wl = mis1(f.fwd / dpw);
        const real wc = mis1(epw * fabsr(ldout.z) / (dpw * r.cos_light)) * (0 + x->ps.dvcm + x->ps.dvc * mis1(f.rev));
        w = 1 / (wl + 1 + wc);
    }
    if (x->ps.path_len + 1 < x->c->min_len) return;
    if (x->nsh >= 64) return;
    shadow_t* sr = &x->sh[x->nsh++];
    sr->o = ray_offset_dir(sd->pos, unpack_normal(sd->gn), r.dir);
    sr->d = r.dir;
    sr->tmax = maxf(r.dist * 0.99f, r.dist - 1e-2f); /* AdjustIntersectionDistance, intersection.hlsli:65 */
    sr->c = contrib;
    sr->w = w;
    sr->medium = x->ps.medium;
    sr->plen = x->ps.path_len;
}

/* SampleLightVertex + ConnectToLightSubpath, testrenderer.slang:355-437: one
 * uniformly chosen vertex of this sample's light vertex cache (picked by
 * next().x modulo the cache size), weighted by cache size / light subpaths. */
static void connect_to_light_subpath(pctx_t* x, const material_t* m, const shading_t* sd, const medium_t* cmed, f3 cpos,
                                     f3 ldin) {
    uint32_t rv[4];
    rng_next(&x->ps.rng, rv);
    if (x->lvc_count == 0) return;
    const lvert_t* v = &x->lvc[rv[0] % x->lvc_count];
    if (x->ps.path_len + 1 + v->plen > x->c->max_depth) return;
    const f3 lp = v->lpos;
    const uint32_t vi = v->inst_prim & 0xffffu;
    const int vmed = (x->s->inst[vi].type_material_address & 0xfu) == 2; /* a medium vertex (isMedium) */
    shading_t sv;
    memset(&sv, 0, sizeof sv);
    f3 vpos;
    if (vmed) vpos = xf_point(&x->s->xf[vi], lp); /* makeVolumeShadingData, shading_data.hlsli:186-194 */
    else {
        make_tri_shading_x(x->s, vi, v->inst_prim >> 16, 0, 0, &lp, &sv);
        vpos = sv.pos;
    }
    f3 dir = sub(vpos, cpos);
    const real dist = length3(dir);
    dir = divf(dir, dist);
    const f3 vin = unpack_normal(v->dirin);
    rer_t fv;
    real cos_light, cpv;
    if (vmed) {
        const medium_t mv = load_medium(x->s, x->s->inst[vi].type_material_address >> 4);
        const real ph = hg_eval(mv.g, vin, neg(dir));
        fv.f = v3(ph, ph, ph); fv.fwd = ph; fv.rev = ph;
        cos_light = 1;
        cpv = med_cont_prob(&mv);
    } else {
        const material_t mv = load_material_sd(x->s, &sv);
        const f3 vout = to_local(&sv, neg(dir));
        cos_light = fabsr(vout.z);
        fv = bsdf_eval_x(&mv, x->c->lambert, vin, vout, 1);
        fv.f = mulf(fv.f, shading_normal_correction_x(&sv, vin, vout, 1));
        cpv = m_cont_prob(&mv);
    }
    fv.fwd *= cpv; fv.rev *= cpv;
    if (all_le0(v->thr) || all_le0(fv.f)) return;
    rer_t f;
    real cos_cam;
    if (cmed) {
        const real ph = hg_eval(cmed->g, ldin, dir);
        f.f = v3(ph, ph, ph); f.fwd = ph; f.rev = ph;
        cos_cam = 1;
        const real cp = med_cont_prob(cmed);
        f.fwd *= cp; f.rev *= cp;
    } else {
        const f3 ldout = to_local(sd, dir);
        cos_cam = fabsr(ldout.z);
        f = bsdf_eval(m, x->c->lambert, ldin, ldout);
        f.f = mulf(f.f, shading_normal_correction(sd, ldin, ldout));
        const real cp = m_cont_prob(m);
        f.fwd *= cp; f.rev *= cp;
    }
    const real wn = (real)x->lvc_count / (real)x->L;
    const real inv_d2 = 1 / pow2(dist);
    const f3 contrib = mul(mul(x->ps.thr, v->thr), mulf(mulf(mul(f.f, fv.f), inv_d2), wn));
    const real cam_pdfa = f.fwd * cos_light * inv_d2;
    const real light_pdfa = fv.fwd * cos_cam * inv_d2;
    const real wl = mis1(cam_pdfa) * (0 + v->dvcm + v->dvc * mis1(fv.rev));   /* [tech. rep. (40)] */
    const real wc = mis1(light_pdfa) * (0 + x->ps.dvcm + x->ps.dvc * mis1(f.rev)); /* [tech. rep. (41)] */
    const real w = 1 / (wl + 1 + wc);
    if (x->nsh >= 64) return;
    shadow_t* sr = &x->sh[x->nsh++];
    sr->o = cmed ? cpos : ray_offset_dir(sd->pos, unpack_normal(sd->gn), dir); /* rayOffset of a medium vertex: itself */
    sr->d = dir;
    sr->tmax = maxf(dist * 0.99f, dist - 1e-2f); /* AdjustIntersectionDistance, intersection.hlsli:65 */
    sr->c = contrib;
    sr->w = w;
    sr->medium = x->ps.medium;
    sr->plen = x->ps.path_len;
}

/* ReservoirSampleDataDI, testrenderer.slang:568-606 */
typedef struct {
    real smp[4];
    real src_pdf, mis_w, wtoa;
    f3 contrib, vo, vd;
    real vtmax;
} rsd_t;
static real rsd_target(const rsd_t* r) { return luminance(r->contrib) * r->wtoa; }
static void rsd_init(rsd_t* r, pctx_t* x, const material_t* m, const shading_t* sd, f3 ldin, const real rnd[4]) {
    memset(r, 0, sizeof *r);
    memcpy(r->smp, rnd, sizeof r->smp);
    isr_t L = sample_illumination(x->s, x->c, rnd, sd->pos);
    r->src_pdf = L.pdf;
    if (luminance(L.radiance) > 0 && L.pdf > 0) {
        f3 ldout = to_local(sd, L.dir);
        rer_t f = bsdf_eval(m, x->c->lambert, ldin, ldout);
        f.f = mulf(f.f, shading_normal_correction(sd, ldin, ldout));
        real cp = rr_cp(x->c, x->ps.path_len, m_cont_prob(m));
        f.fwd *= cp; f.rev *= cp;
        r->contrib = mul(f.f, L.radiance);
        r->wtoa = L.finite ? L.cos_light / pow2(L.dist) : 1;
        real AtoW = L.finite ? pow2(L.dist) / L.cos_light : 1;
        r->mis_w = (x->c->flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(L.pdf, f.fwd / AtoW); /* ComputeMIS_DI, :530-566 */
        r->vo = ray_offset_dir(sd->pos, unpack_normal(sd->gn), L.dir);
        r->vd = L.dir;
        r->vtmax = maxf(L.dist * 0.99f, L.dist - 1e-2f);
    }
}
/* HashGrid::FindCellIndex's checksum of the cell holding p (hashgrid.hlsli:51-58) */
static uint32_t hg_checksum(f3 p, real cs) {
    const int ix = (int)floorr(p.x / cs), iy = (int)floorr(p.y / cs), iz = (int)floorr(p.z / cs);
    const uint32_t h = or_xxhash32(or_xxhash32(asuint(cs)) + or_xxhash32((uint32_t)iz + or_xxhash32((uint32_t)iy + or_xxhash32((uint32_t)ix))));
    return h > 1u ? h : 1u;
}
/* first index of keys[0, n) with key >= k (upper: > k) */
static uint32_t hg_bound(const uint64_t* keys, uint32_t n, uint64_t k, int upper) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (upper ? keys[mid] <= k : keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
static real mis_pair(real a, real b);
/* SampleIlluminationReSTIR's reuse and store (testrenderer.slang:686-723) with
 * ReservoirReuse (:608-663): a point jittered over one cell on the tangent
 * plane, one entry of the previous sample's grid in its cell, pairwise MIS;
 * then this vertex's reservoir into this sample's grid.  A current reservoir
 * of target 0 gets weight 0 (the reference's 0/0 makes the sum NaN). */
static void hg_reuse(pctx_t* x, const material_t* m, const shading_t* sd, f3 ldin, rsd_t* sample, real* W) {
    hgrid_t* g = x->hg_cur;
    real M = (real)x->c->cand;
    real target = rsd_target(sample);
    if (x->hg_prev) {
        const hgrid_t* pg = x->hg_prev;
        real dr[4], d[2];
        rng_nextf(&x->ps.rng, dr);
        sample_concentric_disc(dr[0], dr[1], d);
        const f3 qp = add(sd->pos, to_world(sd, v3(d[0] * g->cell, d[1] * g->cell, 0)));
        const uint64_t key = hg_checksum(qp, g->cell);
        const uint32_t lo = hg_bound(pg->keys, pg->nslots, key, 0), hi = hg_bound(pg->keys, pg->nslots, key, 1);
        if (hi > lo) {
            uint32_t u[4];
            rng_next(&x->ps.rng, u);
            const real* e = pg->rec + 16 * (size_t)pg->slots[lo + u[0] % (hi - lo)];
            rsd_t xn;
            rsd_init(&xn, x, m, sd, ldin, e + 4);
            const real tn = rsd_target(&xn);
            real phq = 0;
            {
                const uint32_t id = asuint(e[3]);
                const f3 lp = v3(e[0], e[1], e[2]);
                shading_t sq;
                make_tri_shading_x(x->s, id & 0xffffu, id >> 16, 0, 0, &lp, &sq);
                const material_t mq = load_material_sd(x->s, &sq);
                rsd_t xq;
                rsd_init(&xq, x, &mq, &sq, v3(e[12], e[13], e[14]), sample->smp);
                phq = rsd_target(&xq);
            }
            const real qM = e[10];
            const real m0 = mis_pair(M * target, qM * phq), m1 = mis_pair(qM * e[8], M * tn);
            const real w0 = m0 * target * *W, w1 = m1 * tn * e[9];
            M += qM;
            const real total = w0 + w1;
            real uu[4];
            rng_nextf(&x->ps.rng, uu);
            if (uu[0] * total <= w1) {
                *sample = xn;
                target = tn;
            }
            *W = target > 0 ? total / target : 0.0f;
        }
    }
    if (*W > 0 && x->ps.path_len - 1 < g->dmax) {
        real* e = g->rec + 16 * ((size_t)(x->ps.path_len - 1) * g->npix + x->hg_pix);
        const uint32_t id = x->ps.inst_prim;
        const f3 lp = xf_point(&x->s->inv_xf[id & 0xffffu], sd->pos);
        e[0] = lp.x; e[1] = lp.y; e[2] = lp.z; e[3] = asfloat(id);
        memcpy(e + 4, sample->smp, 4 * sizeof(real));
        const real mmax = x->c->max_m * (real)x->c->cand;
        e[8] = target; e[9] = *W; e[10] = M < mmax ? M : mmax; e[11] = asfloat(hg_checksum(sd->pos, g->cell));
        e[12] = ldin.x; e[13] = ldin.y; e[14] = ldin.z; e[15] = 0;
    }
}

/* SampleIlluminationReSTIR (initial RIS; with the hash grid, its reuse) +
 * ConnectToLight (ReSTIR), testrenderer.slang:665-726, 730-747 */
static void connect_to_light_restir(pctx_t* x, const material_t* m, const shading_t* sd, f3 ldin) {
    rsd_t sample, xi;
    memset(&sample, 0, sizeof sample);
    real wsum = 0;
    for (uint32_t i = 0; i < x->c->cand; i++) {
        real rnd[4];
        rng_nextf(&x->ps.rng, rnd);
        rsd_init(&xi, x, m, sd, ldin, rnd);
        if (rsd_target(&xi) <= 0) continue;
        wsum += rsd_target(&xi) / xi.src_pdf;
        real u[4];
        rng_nextf(&x->ps.rng, u);
        if (u[0] <= (rsd_target(&xi) / xi.src_pdf) / wsum) sample = xi;
    }
    real W = 0;
    if (rsd_target(&sample) > 0) W = wsum / ((real)x->c->cand * rsd_target(&sample));
    if (x->hg_cur) hg_reuse(x, m, sd, ldin, &sample, &W);
    if (W <= 0) return;
    W *= sample.wtoa;
    f3 contrib = mulf(mul(x->ps.thr, sample.contrib), W);
    if (x->ps.path_len + 1 < x->c->min_len) return;
    if (x->nsh >= 64) return;
    shadow_t* sr = &x->sh[x->nsh++];
    sr->o = sample.vo;
    sr->d = sample.vd;
    sr->tmax = sample.vtmax;
    sr->c = contrib;
    sr->w = sample.mis_w;
    sr->medium = x->ps.medium;
    sr->plen = x->ps.path_len;
}

/* RIS at a primary vertex of the screen-space reuse mode: the candidates of
 * SampleIlluminationReSTIR (testrenderer.slang:665-684), in the same draws,
 * with the winning seed kept for reuse instead of a shadow ray. */
static void ris_reservoir(pctx_t* x, const material_t* m, const shading_t* sd, f3 ldin) {
    rsd_t xi;
    real wsum = 0, target = 0;
    for (int j = 0; j < 4; j++) x->res[j] = 0;
    for (uint32_t i = 0; i < x->c->cand; i++) {
        real rnd[4];
        rng_nextf(&x->ps.rng, rnd);
        rsd_init(&xi, x, m, sd, ldin, rnd);
        const real t = rsd_target(&xi);
        if (t <= 0) continue;
        wsum += t / xi.src_pdf;
        real u[4];
        rng_nextf(&x->ps.rng, u);
        if (u[0] <= (t / xi.src_pdf) / wsum) {
            memcpy(x->res, rnd, sizeof rnd);
            target = t;
        }
    }
    x->res[4] = target;
    x->res[5] = target > 0 ? wsum / ((real)x->c->cand * target) : 0.0f;
    x->res[6] = target > 0 ? (real)x->c->cand : 0.0f;
}

#define OR_MAX_MEDIA_SEGMENTS 256
/* traceScatteringRay, intersection.hlsli:195-260: closest hits, delta tracking
 * through the current medium, medium boundaries crossed in place.
 * Returns 0 miss, 1 surface hit (h), 2 medium scattering at *pos. */
static int trace_scattering(pctx_t* x, f3 origin, f3 dir, hit_t* h, f3* pos, real* dirpdf, real* neepdf) {
    const or_scene* s = x->s;
    pstate_t* ps = &x->ps;
    *dirpdf = 1;
    *neepdf = 1;
    f3 o = origin;
    real tmax = INFINITY;
    /* at most OR_MAX_MEDIA_SEGMENTS traceRay calls; a ray still crossing
     * boundaries after that ends as a miss (the reference loop is unbounded) */
    for (int seg = 0; tmax > 1e-6f && seg < OR_MAX_MEDIA_SEGMENTS; seg++) {
        int is_vol = 0;
        uint32_t vgn = 0;
        *h = trace_ray(s, o, dir, tmax, 0, &is_vol, &vgn, &x->trav[0], &x->trav[1]);
        x->rays[0]++;
        if (ps->medium != 0xFFFFu) {
            const medium_t m = load_medium(s, s->inst[ps->medium].type_material_address >> 4);
            const st_transform* iT = &s->inv_xf[ps->medium];
            f3 dp3, np3;
            int sc;
            adjd_t ad, *adp = NULL;
            if (x->adj && x->adj->pass == 2) { /* every later contribution carries this segment's factors */
                ad.grad = x->adj->grad_grid;
                ad.vol = x->adj->grad_vol;
                ad.wgt = mul(x->adj->g, sub(x->adj->total, x->adj->acc));
                adp = &ad;
            }
            f3 p = delta_track(s, &m, 1, &ps->rng, xf_point(iT, o), xf_vector(iT, dir), h->hit ? h->t : tmax, &ps->thr, &dp3, &np3,
                               &sc, x->c->max_null, adp);
            *dirpdf *= avg3(dp3);
            *neepdf *= avg3(np3);
            if (sc) {
                *pos = xf_point(&s->xf[ps->medium], p);
                h->t = length3(sub(*pos, origin));
                h->id = ps->medium | (0xFFFFu << 16);
                return 2;
            }
        }
        if (!h->hit) {
            h->t = tmax;
            return 0;
        }
        if (!is_vol) return 1;
        const f3 bp = add(o, mulf(dir, h->t));
        const f3 ng = unpack_normal(vgn);
        if (dot(ng, dir) > 0) ps->medium = 0xFFFFu;
        else ps->medium = h->id & 0xFFFFu;
        o = ray_offset_dir(bp, ng, dir);
        tmax -= h->t;
    }
    h->hit = 0;
    h->t = tmax;
    return 0;
}

/* traceVisibilityRay with media, intersection.hlsli:262-338 (ratio tracking).
 * The deferred-shadow variant of the reference reads an uninitialised static
 * path state here (testrenderer.slang:1153-1158); this uses the vertex's
 * medium and a random stream of its own, keyed by pixel, sample and bounce. */
static void trace_visibility_media(pctx_t* x, const shadow_t* sr, uint32_t px, uint32_t py, uint32_t seed, f3* beta, real* neepdf,
                                   const adjd_t* ad) {
    const or_scene* s = x->s;
    rng_t rng;
    rng_init(&rng, seed, px, py, 0x40000000u | (sr->plen << 16));
    *neepdf = 1;
    uint32_t cur = sr->medium;
    f3 o = sr->o;
    real tmax = sr->tmax;
    for (int seg = 0; tmax > 1e-6f; seg++) {
        if (seg == OR_MAX_MEDIA_SEGMENTS) { /* bounded like the scattering ray: counts as blocked */
            *beta = v3(0, 0, 0);
            *neepdf = 0;
            return;
        }
        int is_vol = 0;
        uint32_t vgn = 0;
        hit_t h = trace_ray(s, o, sr->d, tmax, 0, &is_vol, &vgn, &x->trav[2], &x->trav[3]);
        x->rays[1]++;
        if (h.hit && !is_vol) {
            *beta = v3(0, 0, 0);
            *neepdf = 0;
            return;
        }
        if (cur != 0xFFFFu) {
            const medium_t m = load_medium(s, s->inst[cur].type_material_address >> 4);
            const st_transform* iT = &s->inv_xf[cur];
            f3 dp3, np3;
            int sc;
            delta_track(s, &m, 0, &rng, xf_point(iT, o), xf_vector(iT, sr->d), h.hit ? h.t : tmax, beta, &dp3, &np3, &sc,
                        x->c->max_null, ad);
            *neepdf *= avg3(np3);
            if (!(beta->x != 0 && beta->y != 0 && beta->z != 0)) return;
        }
        if (!h.hit) return;
        const f3 bp = add(o, mulf(sr->d, h.t));
        const f3 ng = unpack_normal(vgn);
        if (dot(ng, sr->d) < 0) {
            cur = h.id & 0xFFFFu;
            o = ray_offset(bp, neg(ng));
        } else {
            cur = 0xFFFFu;
            o = ray_offset(bp, ng);
        }
        tmax -= h.t;
    }
}

/* ProcessInteraction for a Medium vertex (testrenderer.slang:816-935 with isMedium). */
/* Sub-path MIS quantities after a BSDF sample, ProcessInteraction
 * (testrenderer.slang:909-922): [tech. rep. (53)-(55)] for a singular sample,
 * else (34)-(36); a camera path is no longer specular after a non-singular one. */
static void vc_after_sample(pstate_t* ps, const dsr_t* smp, real cos_out, int from_light) {
    if (smp->rough == 0) {
        ps->dvcm = qvc(ps, 0.0f);
        ps->dvc = qvc(ps, ps->dvc * mis1(cos_out));
    } else {
        if (!from_light) ps->flag31 = 0;
        ps->dvc = qvc(ps, mis1(cos_out / smp->fwd) * (ps->dvc * mis1(smp->rev) + ps->dvcm + 0));
        ps->dvcm = qvc(ps, mis1(1.0f / smp->fwd));
    }
}

/* The cache-depth vertex: the training path of a run records its row and
 * goes on, any other path records its query and ends (returns 1). */
static int cache_vertex(pctx_t* x, f3 pos, f3 wo, f3 nrm, f3 alb) {
    cache_t* c = x->cache;
    const uint32_t row = c->g / c->stride;
    const int train = c->g % c->stride == or_pcg(row ^ or_pcg(c->frame)) % c->stride;
    const real q[12] = {pos.x, pos.y, pos.z, wo.x, wo.y, wo.z, nrm.x, nrm.y, nrm.z, alb.x, alb.y, alb.z};
    memcpy(c->q, q, sizeof q);
    c->thr[0] = x->ps.thr.x; c->thr[1] = x->ps.thr.y; c->thr[2] = x->ps.thr.z;
    c->kind = train ? 2 : 1;
    c->out_before = x->out;
    c->nsh_before = x->nsh;
    return !train;
}

static int medium_interaction(pctx_t* x, f3 pos, f3 dir) {
    const or_scene* s = x->s;
    pstate_t* ps = &x->ps;
    const medium_t m = load_medium(s, s->inst[ps->medium].type_material_address >> 4);
    const f3 ldin = neg(dir);
    const int vc = (x->c->flags & ST_FLAG_BDPT_LVC) != 0;
    if (vc) { /* [tech. rep. (31)-(33)]; cosIn of a medium vertex is 1 (testrenderer.slang:826-834) */
        ps->dvcm = qvc(ps, ps->dvcm * mis1(pow2(ps->t16)));
        ps->dvcm = qvc(ps, ps->dvcm / mis1(1.0f));
        ps->dvc = qvc(ps, ps->dvc / mis1(1.0f));
    }
    if (ps->path_len == 1) {
        x->vis[0] = ps->inst_prim; x->vis[1] = 0;
        store_aux(x, 1, pos, med_albedo(&m));
    }
    if (ps->path_len >= x->c->max_depth) return 0;
    if (!med_can_eval(&m)) return 0;
    if (x->cache && ps->path_len == x->cache->depth && cache_vertex(x, pos, ldin, v3(0, 0, 0), med_albedo(&m))) return 0;
    if (!med_singular(&m) && vc) {
        real rnd[4];
        rng_nextf(&ps->rng, rnd);
        isr_t r = sample_illumination(s, x->c, rnd, pos);
        if (!(all_le0(r.radiance) || r.pdf <= 0)) {
            real AtoW = 1;
            if (r.finite) AtoW = pow2(r.dist) / r.cos_light;
            const real dpw = r.pdf * AtoW;
            const real v = hg_eval(m.g, ldin, r.dir);
            const real cp = med_cont_prob(&m);
            const real fwd = v * cp, rev = v * cp;
            const f3 contrib = divf(mulf(mul(ps->thr, r.radiance), v), dpw);
            /* ComputeMIS_DI with gUseVC; cosOut of a medium vertex is 1 */
            const real epw = r.pdf * cos_hemisphere_pdfw(r.cos_light);
            const real wl = mis1(fwd / dpw);
            const real wc = mis1(epw * 1.0f / (dpw * r.cos_light)) * (0 + ps->dvcm + ps->dvc * mis1(rev));
            if (x->nsh < 64) {
                shadow_t* sr = &x->sh[x->nsh++];
                sr->o = pos;
                sr->d = r.dir;
                sr->tmax = maxf(r.dist * 0.99f, r.dist - 1e-2f);
                sr->c = contrib;
                sr->w = 1 / (wl + 1 + wc);
                sr->medium = ps->medium;
                sr->plen = ps->path_len;
            }
        }
        connect_to_light_subpath(x, NULL, NULL, &m, pos, ldin);
    } else if (!med_singular(&m) && (x->c->flags & ST_FLAG_SAMPLE_DIRECT) && (x->c->flags & ST_FLAG_RESTIR_DI)) {
        /* SampleIlluminationReSTIR at a medium vertex (testrenderer.slang:568-606,
         * 665-684, 730-747): the phase function, cosOut 1, no ray offset */
        const real cp = rr_cp(x->c, ps->path_len, med_cont_prob(&m));
        real wsum = 0, sel_t = 0, sel_wtoa = 0, sel_mis = 0, sel_dist = 0;
        f3 sel_c = v3(0, 0, 0), sel_dir = v3(0, 0, 0);
        for (uint32_t i = 0; i < x->c->cand; i++) {
            real rnd[4];
            rng_nextf(&ps->rng, rnd);
            isr_t r = sample_illumination(s, x->c, rnd, pos);
            if (!(luminance(r.radiance) > 0 && r.pdf > 0)) continue;
            const real v = hg_eval(m.g, ldin, r.dir);
            const real AtoW = r.finite ? pow2(r.dist) / r.cos_light : 1;
            const f3 xc = mul(v3(v, v, v), r.radiance);
            const real xw = r.finite ? r.cos_light / pow2(r.dist) : 1;
            const real target = luminance(xc) * xw;
            if (target <= 0) continue;
            wsum += target / r.pdf;
            real u[4];
            rng_nextf(&ps->rng, u);
            if (u[0] <= (target / r.pdf) / wsum) {
                sel_t = target; sel_c = xc; sel_wtoa = xw; sel_dir = r.dir; sel_dist = r.dist;
                sel_mis = (x->c->flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(r.pdf, (v * cp) / AtoW);
            }
        }
        const real W = sel_t > 0 ? wsum / ((real)x->c->cand * sel_t) : 0;
        if (W > 0 && ps->path_len + 1 >= x->c->min_len && x->nsh < 64) {
            shadow_t* sr = &x->sh[x->nsh++];
            sr->o = pos;
            sr->d = sel_dir;
            sr->tmax = maxf(sel_dist * 0.99f, sel_dist - 1e-2f);
            sr->c = mulf(mul(ps->thr, sel_c), W * sel_wtoa);
            sr->w = sel_mis;
            sr->medium = ps->medium;
            sr->plen = ps->path_len;
        }
    } else if (!med_singular(&m) && (x->c->flags & ST_FLAG_SAMPLE_DIRECT)) {
        real rnd[4];
        rng_nextf(&ps->rng, rnd);
        isr_t r = sample_illumination(s, x->c, rnd, pos);
        if (!(all_le0(r.radiance) || r.pdf <= 0)) {
            real AtoW = 1;
            if (r.finite) AtoW = pow2(r.dist) / r.cos_light;
            const real dpw = r.pdf * AtoW;
            const real v = hg_eval(m.g, ldin, r.dir);
            const real cp = rr_cp(x->c, ps->path_len, med_cont_prob(&m));
            const real fwd = v * cp;
            const f3 contrib = divf(mulf(mul(ps->thr, r.radiance), v), dpw);
            const real w = (x->c->flags & ST_FLAG_DIRECT_ONLY) ? 1.0f : mis2(r.pdf, fwd / AtoW);
            if (ps->path_len + 1 >= x->c->min_len && x->nsh < 64) {
                shadow_t* sr = &x->sh[x->nsh++];
                sr->o = pos; /* rayOffset of a non-surface point is the point itself */
                sr->d = r.dir;
                sr->tmax = maxf(r.dist * 0.99f, r.dist - 1e-2f);
                sr->c = contrib;
                sr->w = w;
                sr->medium = ps->medium;
                sr->plen = ps->path_len;
            }
        }
    }
    if (x->c->flags & ST_FLAG_DIRECT_ONLY) return 0;
    const real cp = minf(1, med_cont_prob(&m));
    const int rr_on = ps->path_len >= x->c->rr_start;
    real rr[4];
    rng_nextf(&ps->rng, rr);
    if (rr_on && rr[0] > cp) return 0;
    real rb[4];
    rng_nextf(&ps->rng, rb);
    const f3 sdir = hg_sample(m.g, rb[0], rb[1], ldin);
    const real v = hg_eval(m.g, ldin, sdir);
    const real fwd = v * (rr_on ? cp : 1.0f);
    ps->thr = mulf(ps->thr, v / fwd);
    ps->dirpdf16 = (1 - fabsr(m.g)) == 0 ? 0 : q16(fwd);
    if (vc) {
        dsr_t smp;
        memset(&smp, 0, sizeof smp);
        smp.fwd = fwd; smp.rev = fwd; smp.rough = (1 - fabsr(m.g)) == 0 ? 0.0f : 1.0f;
        vc_after_sample(ps, &smp, 1.0f, 0);
    }
    ps->dir = pack_normal(sdir);
    ps->origin = pos;
    if (ps->thr.x != ps->thr.x || ps->thr.y != ps->thr.y || ps->thr.z != ps->thr.z) return 0;
    return 1;
}

/* TraceRay + ExtendPath + ProcessInteraction, testrenderer.slang:816-935, 1026-1071.
 * Returns 0 when the path ends. */
static int extend_path(pctx_t* x, const st_render_params* p, const st_view* view) {
    (void)p; (void)view;
    const or_scene* s = x->s;
    pstate_t* ps = &x->ps;
    f3 dir = unpack_normal(ps->dir);
    hit_t h;
    int kind;
    f3 mpos = v3(0, 0, 0);
    if (s->nvolinst > 0) {
        real dirpdf, neepdf;
        kind = trace_scattering(x, ps->origin, dir, &h, &mpos, &dirpdf, &neepdf);
        ps->thr = divf(ps->thr, dirpdf); /* TraceRay, testrenderer.slang:1030 */
    } else {
        h = trace(s, ps->origin, dir, INFINITY, 0, &x->trav[0], &x->trav[1]);
        x->rays[0]++;
        kind = h.hit ? 1 : 0;
    }
    ps->path_len++;
    shading_t sd;
    memset(&sd, 0, sizeof sd);
    real pick = 0;
    if (kind == 2) {
        /* medium scattering: pickPdf 0 and shape area 0 (intersection.hlsli:232-236) */
        ps->inst_prim = h.id;
        ps->t16 = q16(h.t);
        real pdfa = 0;
        if (s->nlights > 0) {
            pdfa = 0.0f / (0.0f * (real)s->nlights);
            if (s->has_env) pdfa *= 1 - x->c->env_prob;
        }
        ps->pdfa16 = q16(pdfa);
        if (all_le0(ps->thr)) return 0;
        return medium_interaction(x, mpos, dir);
    }
    if (h.hit) {
        uint32_t ii = h.id & 0xffffu, prim = h.id >> 16;
        pick = prim == 0xFFFFu ? 1.0f : 1.0f / (real)(s->inst[ii].data >> 16);
        make_hit_shading(s, &h, ps->origin, dir, &sd);
        ps->inst_prim = h.id;
        if (s->nvolinst > 0) h.t = length3(sub(sd.pos, ps->origin)); /* intersection.hlsli:244 */
    } else {
        ps->inst_prim = 0xffffffffu;
        sd.area = -1;
        sd.sn = pack_normal(dir);
    }
    ps->t16 = q16(h.hit ? h.t : INFINITY);
    /* LightSamplePdfA, lights.hlsli:264-271 */
    real pdfa = 0;
    if (s->nlights > 0) {
        pdfa = pick / (sd.area * (real)s->nlights);
        if (s->has_env) pdfa *= 1 - x->c->env_prob;
    }
    ps->pdfa16 = q16(pdfa);
    if (!h.hit) {
        if (ps->path_len == 1) {
            x->vis[0] = ps->inst_prim; x->vis[1] = sd.sn;
            store_aux(x, 0, v3(0, 0, 0), v3(1, 1, 1));
        }
        real dpw;
        f3 e = mul(ps->thr, evaluate_environment(s, x->c, dir, &dpw));
        if (any_gt0(e)) {
            real dpa = concentric_disc_pdfa() / pow2(s->sphere[3]);
            if (x->c->flags & ST_FLAG_ENV_MIS_BALANCED) dpa = dpw;
            add_emission(x, e, dpa, dpw, 1);
        }
        ps->thr = v3(0, 0, 0);
        return 0;
    }
    if (all_le0(ps->thr)) return 0; /* ExtendPath, testrenderer.slang:1062-1063 */
    material_t m = load_material_sd(s, &sd);
    uint32_t alb_tri = 0xffffffffu; /* the vertex's entry of the per-triangle albedo */
    if (s->tri_alb) {
        alb_tri = s->tri_base[h.id & 0xffffu] + (h.id >> 16);
        m.has_bo = 1;
        m.bo = v3(s->tri_alb[3 * (size_t)alb_tri], s->tri_alb[3 * (size_t)alb_tri + 1], s->tri_alb[3 * (size_t)alb_tri + 2]);
    }
    f3 ldin = to_local(&sd, neg(dir));
    real cos_in = fabsr(ldin.z);
    const int vc = (x->c->flags & ST_FLAG_BDPT_LVC) != 0;
    if (vc) { /* [tech. rep. (31)-(33)], completed at the hit (testrenderer.slang:826-834) */
        ps->dvcm = qvc(ps, ps->dvcm * mis1(pow2(ps->t16)));
        ps->dvcm = qvc(ps, ps->dvcm / mis1(cos_in));
        ps->dvc = qvc(ps, ps->dvc / mis1(cos_in));
    }
    if (ps->path_len == 1) {
        x->vis[0] = ps->inst_prim; x->vis[1] = sd.sn;
        store_aux(x, 1, sd.pos, m_base(&m));
        if (x->reuse) {
            x->prim_hit = 1;
            x->prim_sd = sd;
            x->prim_dir = ps->dir;
            x->res[7] = h.t; /* the raw hit distance */
        }
    }
    if (ldin.z > 0) {
        f3 e = mul(ps->thr, m_emission(&m));
        if (any_gt0(e)) {
            real pa = ps->pdfa16;
            add_emission(x, e, pa, pa * cos_hemisphere_pdfw(ldin.z), cos_in / pow2(ps->t16));
        }
    }
    if (x->adj && x->adj->pass == 2 && alb_tri != 0xffffffffu) {
        /* Lambertian f = albedo / pi: every later contribution (this vertex's
         * NEE, the continuation) carries one factor albedo_c of this vertex */
        const f3 wgt = mul(x->adj->g, sub(x->adj->total, x->adj->acc));
        const f3 a = m.bo;
        double* ga = x->adj->grad_alb + 3 * (size_t)alb_tri;
        if (a.x > 0) ga[0] += (double)(wgt.x / a.x);
        if (a.y > 0) ga[1] += (double)(wgt.y / a.y);
        if (a.z > 0) ga[2] += (double)(wgt.z / a.z);
    }
    if (ps->path_len >= x->c->max_depth) return 0;
    if (!m_can_eval(&m)) return 0;
    if (x->cache && ps->path_len == x->cache->depth && cache_vertex(x, sd.pos, neg(dir), unpack_normal(sd.sn), m_base(&m))) return 0;
    if (!m_singular(&m, x->c->lambert) && vc) {
        connect_to_light(x, &m, &sd, ldin);
        connect_to_light_subpath(x, &m, &sd, NULL, sd.pos, ldin);
    } else if (!m_singular(&m, x->c->lambert) && (x->c->flags & ST_FLAG_SAMPLE_DIRECT)) {
        /* with screen-space reuse, the primary vertex keeps its RIS reservoir
         * (its shadow ray comes after the reuse passes); deeper vertices take
         * RIS with ST_FLAG_RESTIR_DI_DEEP, else plain NEE */
        if (x->reuse && ps->path_len == 1) ris_reservoir(x, &m, &sd, ldin);
        else if ((x->c->flags & ST_FLAG_RESTIR_DI) && (!x->reuse || (x->c->flags & ST_FLAG_RESTIR_DI_DEEP)))
            connect_to_light_restir(x, &m, &sd, ldin);
        else connect_to_light(x, &m, &sd, ldin);
    }
    if (x->c->flags & ST_FLAG_DIRECT_ONLY) return 0; /* gSampleDirectIlluminationOnly, testrenderer.slang:886-888 */
    /* Russian roulette (testrenderer.slang:890-892); the draw is made at every
     * vertex, the test only from rr_start on */
    real cp = minf(1, m_cont_prob(&m));
    const int rr_on = ps->path_len >= x->c->rr_start;
    real rr[4];
    rng_nextf(&ps->rng, rr);
    if (rr_on && rr[0] > cp) return 0;
    if (!rr_on) cp = 1;
    real rb[4];
    rng_nextf(&ps->rng, rb);
    dsr_t smp = bsdf_sample(&m, x->c->lambert, v3(rb[0], rb[1], rb[2]), ldin);
    smp.fwd *= cp; smp.rev *= cp;
    ps->thr = mul(ps->thr, divf(smp.f, smp.fwd));
    if (smp.rough == 0) ps->dirpdf16 = 0;
    else ps->dirpdf16 = q16(smp.fwd);
    if (vc) vc_after_sample(ps, &smp, fabsr(smp.dir.z), 0);
    ps->thr = mulf(ps->thr, shading_normal_correction(&sd, ldin, smp.dir));
    f3 wd = to_world(&sd, smp.dir);
    ps->dir = pack_normal(wd);
    ps->origin = ray_offset_dir(sd.pos, unpack_normal(sd.gn), wd);
    if (ps->thr.x != ps->thr.x || ps->thr.y != ps->thr.y || ps->thr.z != ps->thr.z) return 0;
    return 1;
}

/* The path replay's shadow rays, traced as the GPU wavefront does after each
 * bounce: the fp32 contribution (ST_FLAG_FLOAT_NEE's) joins acc, and in pass
 * 2 ratio tracking is replayed with the contribution as the sink weight. */
static void adj_flush_shadows(pctx_t* x, uint32_t px, uint32_t py, uint32_t seed) {
    adj_t* A = x->adj;
    for (int i = 0; i < x->nsh; i++) {
        shadow_t sr = x->sh[i];
        f3 c = sr.c;
        if (x->s->nvolinst > 0) {
            real neepdf;
            trace_visibility_media(x, &sr, px, py, seed, &c, &neepdf, NULL);
            if (all_le0(c) || neepdf <= 0) continue;
            c = divf(c, neepdf);
        } else {
            hit_t h = trace(x->s, sr.o, sr.d, sr.tmax, 1, &x->trav[2], &x->trav[3]);
            x->rays[1]++;
            if (h.hit) continue;
        }
        if (sr.w <= 0 || all_le0(c)) continue;
        const f3 rw = mulf(c, sr.w);
        const f3 add3 = v3(maxf(0, rw.x), maxf(0, rw.y), maxf(0, rw.z));
        A->acc = add(A->acc, add3);
        if (A->pass == 2 && x->s->nvolinst > 0) {
            adjd_t ad = {A->grad_grid, A->grad_vol, mul(A->g, add3)};
            f3 cb = sr.c;
            real np;
            trace_visibility_media(x, &sr, px, py, seed, &cb, &np, &ad);
        }
    }
    x->nsh = 0;
}

/* One launch (1 spp) for one pixel: Render + RenderIteration x (maxDepth-1)
 * + ProcessShadowRays + ProcessAtomicOutput (testrenderer.slang:1087-1183). */
static void render_pixel(pctx_t* x, const st_render_params* p, const st_view* view, uint32_t px, uint32_t py,
                         uint32_t seed, f3* result, uint32_t* plen) {
    pstate_t* ps = &x->ps;
    memset(ps, 0, sizeof *ps);
    x->out = v3(0, 0, 0);
    x->nee[0] = x->nee[1] = x->nee[2] = 0;
    x->nsh = 0;
    if (x->cache) {
        x->cache->kind = 0;
        x->cache->nee_before[0] = x->cache->nee_before[1] = x->cache->nee_before[2] = 0;
    }
    rng_init(&ps->rng, seed, px, py, 0);
    real ox = 0.5f, oy = 0.5f;
    if (x->c->flags & ST_FLAG_PIXEL_JITTER) {
        real j[4];
        rng_nextf(&ps->rng, j);
        ox = j[0]; oy = j[1];
    }
    /* InitializeCameraPath, testrenderer.slang:938-982; ViewData::toWorld scene.h:152-157;
     * ProjectionData::backProject transform.h:159-170 */
    real ew = (real)(view->image_max[0] - view->image_min[0]), eh = (real)(view->image_max[1] - view->image_min[1]);
    real u = (((real)px + ox) - (real)view->image_min[0]) / ew, v = (((real)py + oy) - (real)view->image_min[1]) / eh;
    real cx = 2 * u - 1, cy = -(2 * v - 1);
    const st_projection* pr = &view->projection;
    real sg = pr->near_plane > 0 ? 1.0f : (pr->near_plane < 0 ? -1.0f : 0.0f);
    f3 bp = v3(pr->near_plane * (cx * sg - pr->offset[0]) / pr->scale[0], pr->near_plane * (cy * sg - pr->offset[1]) / pr->scale[1],
               pr->near_plane);
    if (pr->vertical_fov < 0) bp = v3((cx - pr->offset[0]) / pr->scale[0], (cy - pr->offset[1]) / pr->scale[1], pr->near_plane);
    f3 ldir = normalize3(bp);
    const st_transform* T = &view->camera_to_world;
    ps->origin = v3(T->m[0][3], T->m[1][3], T->m[2][3]);
    ps->medium = x->view_medium;
    ps->path_len = 0;
    ps->thr = v3(1, 1, 1);
    ps->dir = pack_normal(normalize3(xf_vector(T, ldir)));
    if (x->c->flags & ST_FLAG_BDPT_LVC) { /* testrenderer.slang:958-977 */
        const real cos_cam = fabsr(ldir.z);
        const real ipcd = x->ipd / cos_cam;
        const real cam_pdfw = pow2(ipcd) / cos_cam;
        ps->dirpdf16 = q16(cam_pdfw);
        ps->flag31 = 1;
        ps->vc_f32 = (x->c->flags & ST_FLAG_VC_FP32_MIS) != 0;
        ps->dvcm = qvc(ps, mis1((real)x->L / cam_pdfw));
        ps->dvc = qvc(ps, 0.0f);
    }
    /* Render always extends once; RenderIteration returns early when the
     * throughput is zero or the depth is reached (testrenderer.slang:1105, 1137-1141) */
    for (int first = 1;; first = 0) {
        if (!first && (all_le0(ps->thr) || ps->path_len + 1 > x->c->max_depth)) break;
        const int alive = extend_path(x, p, view);
        if (x->adj) adj_flush_shadows(x, px, py, seed);
        if (!alive) { ps->thr = v3(0, 0, 0); break; }
    }
    /* ProcessShadowRays + AddRadianceInterlocked, testrenderer.slang:237-261, 1147-1166 */
    for (int i = 0; i < x->nsh; i++) {
        shadow_t* sr = &x->sh[i];
        if (x->s->nvolinst > 0) {
            real neepdf;
            trace_visibility_media(x, sr, px, py, seed, &sr->c, &neepdf, NULL);
            if (all_le0(sr->c) || neepdf <= 0) continue;
            sr->c = divf(sr->c, neepdf);
        } else {
            hit_t h = trace(x->s, sr->o, sr->d, sr->tmax, 1, &x->trav[2], &x->trav[3]);
            x->rays[1]++;
            if (h.hit) continue;
        }
        if (sr->w <= 0 || all_le0(sr->c)) continue;
        f3 rw = mulf(sr->c, sr->w);
        real q[3] = {clamp0(rw.x) * (real)OR_Q, clamp0(rw.y) * (real)OR_Q, clamp0(rw.z) * (real)OR_Q};
        for (int k = 0; k < 3; k++) {
            if (x->c->flags & ST_FLAG_FLOAT_NEE) continue;
            int32_t iv = q[k] >= 2147483647.0f ? INT_MAX : (int32_t)q[k];
            x->nee[k] += (uint32_t)iv;
            if (x->cache && x->cache->kind && i < x->cache->nsh_before) x->cache->nee_before[k] += (uint32_t)iv;
        }
        if (x->c->flags & ST_FLAG_FLOAT_NEE) x->out = add(x->out, v3(maxf(0, rw.x), maxf(0, rw.y), maxf(0, rw.z)));
    }
    /* ProcessAtomicOutput, testrenderer.slang:1173-1183: one atomic sum per pixel
     * holds every connection (with VC, also the light pass's camera connections) */
    if (x->light_acc)
        for (int k = 0; k < 3; k++) x->nee[k] += x->light_acc[k];
    result->x = x->out.x + (real)x->nee[0] / (real)OR_Q;
    result->y = x->out.y + (real)x->nee[1] / (real)OR_Q;
    result->z = x->out.z + (real)x->nee[2] / (real)OR_Q;
    *plen = ps->path_len;
}

/* ------------------------------------------------------- light tracing */
/* TestRenderer with mLightTrace (gTraceFromLight, TestRenderer.cpp:106-125,
 * 534-582): the light dispatch of Render / RenderIteration
 * (testrenderer.slang:1087-1144) over mLightSubpathCount paths, then
 * ProcessShadowRays and ProcessAtomicOutput with gClearImage (:1147-1183). */
typedef struct {
    pctx_t b;       /* path state, scene, counters, view */
    uint32_t L;     /* light paths per sample */
    real ipd;      /* ViewData::imagePlaneDist */
    uint32_t seed;  /* this sample's seed */
    uint32_t* acc;  /* this thread's mOutputAtomic, 3 per pixel */
    lvert_t* verts; /* bidirectional mode: max_depth - 1 slots per light subpath */
    uint8_t* nverts;
} lctx_t;

/* SceneParameters::GetMediumIndex (scene.hlsli:32-40): the first volume whose
 * object-space bounds hold the point. */
static uint32_t medium_at(const or_scene* s, f3 pos) {
    for (uint32_t i = 0; i < s->nvolinfo; i++) {
        const st_volume_info* vi = &s->volinfo[i];
        const f3 lp = xf_point(&s->inv_xf[vi->instance_index], pos);
        if (lp.x >= vi->min[0] && lp.y >= vi->min[1] && lp.z >= vi->min[2] && lp.x <= vi->max[0] && lp.y <= vi->max[1] &&
            lp.z <= vi->max[2])
            return vi->instance_index;
    }
    return 0xFFFFu;
}

/* ConnectToCamera, testrenderer.slang:441-524 (no VC: weight 1).  med != NULL:
 * a medium vertex (phase function toward the camera, cosOut 1, the ray leaves
 * from the vertex itself); else a surface vertex with the adjoint BSDF and
 * shading-normal factor.  The shadow ray runs to the camera position,
 * unshortened, through the vertex's medium (ratio tracking keyed like the
 * camera paths' NEE, by the light path's dispatch index and length); its
 * contribution goes to pixel int(y) * W + int(x) (dropped beyond the image). */
static void connect_to_camera(lctx_t* lx, const material_t* m, const shading_t* sd, const medium_t* med, f3 pos, f3 ldin,
                              uint32_t q) {
    pctx_t* x = &lx->b;
    const f3 pc = xf_point(&x->inv_cam, pos);
    real pcx = 0, pcy = 0;
    if (!to_raster(x->view, pc, &pcx, &pcy)) return;
    const real cos_cam = fabsr(normalize3(pc).z);
    const st_transform* T = &x->view->camera_to_world;
    f3 tc = sub(v3(T->m[0][3], T->m[1][3], T->m[2][3]), pos);
    const real dist = length3(tc);
    tc = divf(tc, dist);
    f3 f, o;
    real cos_out, frev = 0;
    if (med) {
        const real v = hg_eval(med->g, ldin, tc);
        f = v3(v, v, v);
        cos_out = 1;
        o = pos;
        frev = v * med_cont_prob(med);
    } else {
        const f3 ldout = to_local(sd, tc);
        cos_out = fabsr(ldout.z);
        rer_t r = bsdf_eval_x(m, x->c->lambert, ldin, ldout, 1);
        f = mulf(r.f, shading_normal_correction_x(sd, ldin, ldout, 1));
        frev = r.rev * m_cont_prob(m);
        o = ray_offset_dir(pos, unpack_normal(sd->gn), tc);
    }
    const real image_point_to_camera = lx->ipd / cos_cam;
    const real image_to_solid_angle = pow2(image_point_to_camera) / cos_cam;
    const real image_to_surface = image_to_solid_angle * cos_out / pow2(dist);
    const real surface_to_image = 1.0f / image_to_surface;
    f3 contrib = divf(mul(x->ps.thr, f), (real)lx->L * surface_to_image * cos_out);
    real wmis = 1;
    if (x->c->flags & ST_FLAG_BDPT_LVC) { /* [tech. rep. (46)]; the eye sub-path weight is 0 (47) */
        const real wl = mis1(image_to_surface / (real)lx->L) * (0 + x->ps.dvcm + x->ps.dvc * mis1(frev));
        wmis = 1.f / (wl + 1.f);
    }
    const uint32_t pix = (uint32_t)(int32_t)pcy * x->W + (uint32_t)(int32_t)pcx;
    if (pix >= x->W * x->H) return;
    if (x->s->nvolinst > 0) {
        shadow_t sr;
        sr.o = o; sr.d = tc; sr.tmax = dist; sr.c = contrib; sr.w = 1;
        sr.medium = x->ps.medium; sr.plen = x->ps.path_len;
        real neepdf;
        trace_visibility_media(x, &sr, q % x->W, q / x->W, lx->seed, &contrib, &neepdf, NULL);
        if (all_le0(contrib) || neepdf <= 0) return;
        contrib = divf(contrib, neepdf);
    } else {
        hit_t h = trace(x->s, o, tc, dist, 1, &x->trav[2], &x->trav[3]);
        x->rays[1]++;
        if (h.hit) return;
    }
    if (all_le0(contrib) || wmis <= 0) return;
    /* AddRadianceInterlocked, testrenderer.slang:237-261 */
    contrib = mulf(contrib, wmis);
    const real qv[3] = {clamp0(contrib.x) * (real)OR_Q, clamp0(contrib.y) * (real)OR_Q, clamp0(contrib.z) * (real)OR_Q};
    for (int k = 0; k < 3; k++) lx->acc[3 * (size_t)pix + k] += (uint32_t)(qv[k] >= 2147483647.0f ? INT_MAX : (int32_t)qv[k]);
}

/* One light subpath: InitializeLightPath (testrenderer.slang:983-1024), then
 * TraceRay<true> + ExtendPath + ProcessInteraction<true> (:810-935, 1026-1074)
 * per segment.  A sample of zero pdf (no light, or a one-sided light facing
 * away from the origin) or with a NaN throughput ends at once: the reference
 * carries it on with a NaN throughput whose contributions clamp to zero. */
static void trace_light_path(lctx_t* lx, uint32_t q) {
    pctx_t* x = &lx->b;
    const or_scene* s = x->s;
    const cfg_t* c = x->c;
    pstate_t* ps = &x->ps;
    memset(ps, 0, sizeof *ps);
    rng_init(&ps->rng, lx->seed, q % x->W, q / x->W, 0xFFFFu);
    real rnd[4], dr[4];
    rng_nextf(&ps->rng, rnd);
    const isr_t r = sample_illumination(s, c, rnd, v3(0, 0, 0));
    rng_nextf(&ps->rng, dr);
    f3 thr = divf(r.radiance, r.pdf), o, d;
    real direct_pdfa = 1, emission_pdfw = 1, cos_light = 1;
    if (r.finite) {
        const f3 ld = sample_cos_hemisphere(dr[0], dr[1]);
        f3 onb[3];
        make_orthonormal(unpack_normal(r.normal), onb);
        o = ray_offset(r.pos, onb[2]);
        d = normalize3(add(add(mulf(onb[0], ld.x), mulf(onb[1], ld.y)), mulf(onb[2], ld.z)));
        const real dirpdf = q16(cos_hemisphere_pdfw(ld.z)); /* mDirPdfW is stored in fp16 */
        thr = mulf(thr, ld.z / dirpdf);
        direct_pdfa = r.pdf;
        emission_pdfw = r.pdf * dirpdf;
        cos_light = ld.z;
    } else {
        real xy[2];
        sample_concentric_disc(dr[0], dr[1], xy);
        f3 onb[3];
        make_orthonormal(neg(r.dir), onb);
        const f3 ctr = v3(s->sphere[0], s->sphere[1], s->sphere[2]);
        o = add(ctr, mulf(add(add(mulf(onb[0], xy[0]), mulf(onb[1], xy[1])), r.dir), s->sphere[3]));
        d = neg(r.dir);
    }
    if (!(r.pdf > 0) || thr.x != thr.x || thr.y != thr.y || thr.z != thr.z || !any_gt0(thr)) return;
    ps->thr = thr;
    ps->origin = o;
    ps->dir = pack_normal(d);
    ps->medium = (s->nvolinst > 0 && r.finite) ? medium_at(s, r.pos) : 0xFFFFu;
    const int vc = (c->flags & ST_FLAG_BDPT_LVC) != 0;
    if (vc) { /* testrenderer.slang:1016-1019; light samples are never singular (lights.hlsli:168) */
        ps->flag31 = r.finite;
        ps->vc_f32 = (c->flags & ST_FLAG_VC_FP32_MIS) != 0;
        ps->dvcm = qvc(ps, mis1(direct_pdfa / emission_pdfw));
        ps->dvc = qvc(ps, mis1(cos_light / emission_pdfw));
    }
    while (ps->path_len + 1 <= c->max_depth) {
        const f3 wd = unpack_normal(ps->dir);
        hit_t h;
        int kind;
        f3 mpos = v3(0, 0, 0);
        if (s->nvolinst > 0) {
            real dirpdf, neepdf;
            kind = trace_scattering(x, ps->origin, wd, &h, &mpos, &dirpdf, &neepdf);
            ps->thr = divf(ps->thr, dirpdf);
        } else {
            h = trace(s, ps->origin, wd, INFINITY, 0, &x->trav[0], &x->trav[1]);
            x->rays[0]++;
            kind = h.hit ? 1 : 0;
        }
        ps->path_len++;
        if (kind == 0) return;          /* no emission is gathered on light paths */
        if (all_le0(ps->thr)) return;   /* ExtendPath, testrenderer.slang:1062-1063 */
        if (kind == 2) {
            const medium_t m = load_medium(s, s->inst[ps->medium].type_material_address >> 4);
            const f3 ldin = neg(wd);
            if (vc) { /* testrenderer.slang:812-824 with cosIn 1 */
                ps->t16 = q16(h.t);
                if (ps->path_len > 1 || ps->flag31) ps->dvcm = qvc(ps, ps->dvcm * mis1(pow2(ps->t16)));
                ps->dvcm = qvc(ps, ps->dvcm / mis1(1.0f));
                ps->dvc = qvc(ps, ps->dvc / mis1(1.0f));
            }
            if (ps->path_len >= c->max_depth || !med_can_eval(&m)) return;
            if (vc && !med_singular(&m) && ps->path_len + 2 <= c->max_depth) {
                lvert_t* v = &lx->verts[(size_t)q * (c->max_depth - 1) + lx->nverts[q]++];
                v->lpos = xf_point(&s->inv_xf[h.id & 0xffffu], mpos);
                v->inst_prim = h.id;
                v->thr = ps->thr;
                v->dirin = pack_normal(ldin);
                v->plen = ps->path_len;
                v->dvc = ps->dvc;
                v->dvcm = ps->dvcm;
            }
            if (!med_singular(&m) && ps->path_len + 1 >= c->min_len) connect_to_camera(lx, NULL, NULL, &m, mpos, ldin, q);
            const real cp = minf(1, med_cont_prob(&m));
            const int rr_on = ps->path_len >= c->rr_start;
            real rr[4], rb[4];
            rng_nextf(&ps->rng, rr);
            if (rr_on && rr[0] > cp) return;
            rng_nextf(&ps->rng, rb);
            const f3 sdir = hg_sample(m.g, rb[0], rb[1], ldin);
            const real v = hg_eval(m.g, ldin, sdir);
            const real fwd = v * (rr_on ? cp : 1.0f);
            ps->thr = mulf(ps->thr, v / fwd);
            if (vc) {
                dsr_t smp;
                memset(&smp, 0, sizeof smp);
                smp.fwd = fwd; smp.rev = fwd; smp.rough = (1 - fabsr(m.g)) == 0 ? 0.0f : 1.0f;
                vc_after_sample(ps, &smp, 1.0f, 1);
            }
            ps->dir = pack_normal(sdir);
            ps->origin = mpos;
        } else {
            shading_t sd;
            make_hit_shading(s, &h, ps->origin, wd, &sd);
            const material_t m = load_material_sd(s, &sd);
            const f3 ldin = to_local(&sd, neg(wd));
            if (vc) { /* testrenderer.slang:812-824 */
                const real cos_in = fabsr(ldin.z);
                ps->t16 = q16(h.t);
                if (ps->path_len > 1 || ps->flag31) ps->dvcm = qvc(ps, ps->dvcm * mis1(pow2(ps->t16)));
                ps->dvcm = qvc(ps, ps->dvcm / mis1(cos_in));
                ps->dvc = qvc(ps, ps->dvc / mis1(cos_in));
            }
            if (ps->path_len >= c->max_depth || !m_can_eval(&m)) return;
            if (vc && !m_singular(&m, c->lambert) && ps->path_len + 2 <= c->max_depth) {
                /* StoreLightVertex, testrenderer.slang:329-341 */
                lvert_t* v = &lx->verts[(size_t)q * (c->max_depth - 1) + lx->nverts[q]++];
                v->lpos = xf_point(&s->inv_xf[h.id & 0xffffu], sd.pos);
                v->inst_prim = h.id;
                v->thr = ps->thr;
                v->dirin = pack_normal(ldin);
                v->plen = ps->path_len;
                v->dvc = ps->dvc;
                v->dvcm = ps->dvcm;
            }
            if (!m_singular(&m, c->lambert) && ps->path_len + 1 >= c->min_len) connect_to_camera(lx, &m, &sd, NULL, sd.pos, ldin, q);
            const real cp0 = minf(1, m_cont_prob(&m));
            const int rr_on = ps->path_len >= c->rr_start;
            real rr[4], rb[4];
            rng_nextf(&ps->rng, rr);
            if (rr_on && rr[0] > cp0) return;
            const real cp = rr_on ? cp0 : 1.0f;
            rng_nextf(&ps->rng, rb);
            dsr_t smp = bsdf_sample_x(&m, c->lambert, v3(rb[0], rb[1], rb[2]), ldin, 1);
            smp.fwd *= cp; smp.rev *= cp;
            ps->thr = mul(ps->thr, divf(smp.f, smp.fwd));
            if (vc) vc_after_sample(ps, &smp, fabsr(smp.dir.z), 1);
            ps->thr = mulf(ps->thr, shading_normal_correction_x(&sd, ldin, smp.dir, 1));
            const f3 w = to_world(&sd, smp.dir);
            ps->dir = pack_normal(w);
            ps->origin = ray_offset_dir(sd.pos, unpack_normal(sd.gn), w);
        }
        if (ps->thr.x != ps->thr.x || ps->thr.y != ps->thr.y || ps->thr.z != ps->thr.z) return;
        if (all_le0(ps->thr)) return; /* RenderIteration skips zero-throughput paths */
    }
}

static int render_light(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads) {
    if (!(p->light_subpath_factor >= 0) || !(p->light_subpath_factor <= 64)) return fail("light_subpath_factor out of range");
    cfg_t c;
    memset(&c, 0, sizeof c);
    c.flags = p->flags & ~(uint32_t)(ST_FLAG_SAMPLE_DIRECT | ST_FLAG_DIRECT_ONLY | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE |
                                      ST_FLAG_FLOAT_NEE | ST_FLAG_BDPT_LVC);
    c.env_prob = p->environment_sample_probability;
    c.max_depth = p->max_depth;
    c.min_len = p->min_path_length;
    c.lambert = (p->flags & ST_FLAG_LAMBERTIAN) ? 1 : ((p->flags & ST_FLAG_LAMBERT_GGX) ? 2 : 0);
    c.max_null = p->max_null_collisions;
    c.rr_start = p->rr_start_depth;
    const uint32_t W = p->width, H = p->height, npix = W * H;
    /* mLightSubpathCount = max(1, uint(W * H * factor)), TestRenderer.cpp:452 */
    const real Lf = (real)npix * p->light_subpath_factor; /* bounded before the conversion */
    if (!(Lf < 2147483648.0f)) return fail("too many light subpaths");
    uint32_t L = (uint32_t)Lf;
    if (L < 1) L = 1;
    /* ViewData::imagePlaneDist, compat/scene.h:147 */
    const real ipd = (real)abs(view->image_max[1] - view->image_min[1]) / (2 * tanr(view->projection.vertical_fov / 2));
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    uint32_t* total = calloc((size_t)npix * 3, 4);
    f3* sum = calloc(npix, sizeof(f3));
    if (!total || !sum) { free(total); free(sum); return fail("out of memory"); }
    uint64_t rays[2] = {0, 0}, trav[4] = {0, 0, 0, 0};
    int oom = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (uint32_t k = 0; k < p->spp; k++) {
        memset(total, 0, (size_t)npix * 3 * 4);
#pragma omp parallel
        {
            lctx_t x;
            memset(&x, 0, sizeof x);
            x.b.s = s; x.b.c = &c; x.b.view = view; x.b.inv_cam = inv_cam;
            x.b.W = W; x.b.H = H;
            x.L = L; x.ipd = ipd; x.seed = p->seed + k;
            x.acc = calloc((size_t)npix * 3, 4);
#pragma omp for schedule(dynamic, 256)
            for (int64_t q = 0; q < (int64_t)L; q++)
                if (x.acc) trace_light_path(&x, (uint32_t)q);
#pragma omp critical
            {
                /* integer sums: the same in any order, as the reference's atomics */
                if (x.acc) for (size_t i = 0; i < (size_t)npix * 3; i++) total[i] += x.acc[i];
                else oom = 1;
                rays[0] += x.b.rays[0]; rays[1] += x.b.rays[1];
                for (int j = 0; j < 4; j++) trav[j] += x.b.trav[j];
            }
            free(x.acc);
        }
        if (oom) { free(total); free(sum); return fail("out of memory"); }
        /* ProcessAtomicOutput with gClearImage, then the mean over samples */
        for (uint32_t i = 0; i < npix; i++) {
            const f3 r = v3(0.0f + (real)total[3 * i] / (real)OR_Q, 0.0f + (real)total[3 * i + 1] / (real)OR_Q,
                            0.0f + (real)total[3 * i + 2] / (real)OR_Q);
            sum[i] = add(sum[i], r);
        }
    }
    if (out->hdr)
        for (uint32_t i = 0; i < npix; i++) {
            out->hdr[4 * i + 0] = sum[i].x / (real)p->spp;
            out->hdr[4 * i + 1] = sum[i].y / (real)p->spp;
            out->hdr[4 * i + 2] = sum[i].z / (real)p->spp;
            out->hdr[4 * i + 3] = 1;
        }
    if (out->ray_counts) { out->ray_counts[0] = rays[0]; out->ray_counts[1] = rays[1]; }
    if (out->traversal) for (int j = 0; j < 4; j++) out->traversal[j] = trav[j];
    free(total);
    free(sum);
    return 0;
}

/* The camera's medium: the last volume whose object-space bounds hold the
 * camera position (TestRenderer.cpp:431-446). */
static uint32_t camera_medium(const or_scene* s, const st_view* view) {
    uint32_t view_medium = 0xFFFFu;
    const f3 cam = v3(view->camera_to_world.m[0][3], view->camera_to_world.m[1][3], view->camera_to_world.m[2][3]);
    for (uint32_t i = 0; i < s->nvolinfo; i++) {
        const st_volume_info* vi = &s->volinfo[i];
        if (vi->instance_index >= s->ninst) continue;
        const f3 lp = xf_point(&s->inv_xf[vi->instance_index], cam);
        if (lp.x >= vi->min[0] && lp.y >= vi->min[1] && lp.z >= vi->min[2] && lp.x <= vi->max[0] && lp.y <= vi->max[1] &&
            lp.z <= vi->max[2])
            view_medium = vi->instance_index;
    }
    return view_medium;
}

/* TestRenderer with gUseVC (TestRenderer.cpp:478-482, 559-629): per sample,
 * the light dispatch (light subpaths store their vertices and connect to the
 * camera), then the view dispatch (emission, NEE and one light-vertex
 * connection per vertex), then ProcessShadowRays and ProcessAtomicOutput into
 * the view paths' image.  The light vertex cache is ordered by light subpath
 * and depth. */
static int render_vc(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads) {
    /* Participating media stay rejected: the medium-vertex branches below
     * follow TestRenderer's volume code, but with them the four strategies'
     * weights sum to about 1.18 on a Cornell box with a scattering block (each
     * strategy alone matches the path tracer), so that combination is left
     * open. */
    if (s->nvolinst > 0 || (p->flags & (ST_FLAG_RESTIR_DI | ST_FLAG_FLOAT_NEE | ST_FLAG_DIRECT_ONLY)) ||
        p->min_path_length)
        return fail("the bidirectional mode supports surfaces with plain NEE only");
    if (p->rr_start_depth > 1) return fail("the bidirectional mode runs Russian roulette at every vertex");
    if (!(p->light_subpath_factor >= 0) || !(p->light_subpath_factor <= 64)) return fail("light_subpath_factor out of range");
    if (p->max_depth < 1) return fail("max_depth must be at least 1");
    cfg_t c;
    memset(&c, 0, sizeof c);
    c.flags = p->flags;
    c.env_prob = p->environment_sample_probability;
    c.max_depth = p->max_depth;
    c.lambert = (p->flags & ST_FLAG_LAMBERTIAN) ? 1 : ((p->flags & ST_FLAG_LAMBERT_GGX) ? 2 : 0);
    c.max_null = p->max_null_collisions;
    c.rr_start = p->rr_start_depth;
    const uint32_t W = p->width, H = p->height, npix = W * H;
    const real Lf = (real)npix * p->light_subpath_factor; /* bounded before the conversion */
    if (!(Lf < 2147483648.0f)) return fail("too many light subpaths");
    uint32_t L = (uint32_t)Lf;
    if (L < 1) L = 1;
    const real ipd = (real)abs(view->image_max[1] - view->image_min[1]) / (2 * tanr(view->projection.vertical_fov / 2));
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    const size_t slots = (size_t)L * (p->max_depth - 1 ? p->max_depth - 1 : 1);
    const uint32_t view_medium = camera_medium(s, view);
    uint32_t* total = calloc((size_t)npix * 3, 4);
    f3* sum = calloc(npix, sizeof(f3));
    lvert_t* verts = malloc(sizeof(lvert_t) * slots);
    lvert_t* lvc = malloc(sizeof(lvert_t) * slots);
    uint8_t* nverts = calloc(L, 1);
    if (!total || !sum || !verts || !lvc || !nverts) {
        free(total); free(sum); free(verts); free(lvc); free(nverts);
        return fail("out of memory");
    }
    uint64_t rays[2] = {0, 0}, trav[4] = {0, 0, 0, 0};
    int oom = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (uint32_t k = 0; k < p->spp && !oom; k++) {
        memset(total, 0, (size_t)npix * 3 * 4);
        memset(nverts, 0, L);
#pragma omp parallel
        {
            lctx_t x;
            memset(&x, 0, sizeof x);
            x.b.s = s; x.b.c = &c; x.b.view = view; x.b.inv_cam = inv_cam;
            x.b.W = W; x.b.H = H;
            x.L = L; x.ipd = ipd; x.seed = p->seed + k;
            x.verts = verts; x.nverts = nverts;
            x.acc = calloc((size_t)npix * 3, 4);
#pragma omp for schedule(dynamic, 256)
            for (int64_t q = 0; q < (int64_t)L; q++)
                if (x.acc) trace_light_path(&x, (uint32_t)q);
#pragma omp critical
            {
                if (x.acc) for (size_t i = 0; i < (size_t)npix * 3; i++) total[i] += x.acc[i];
                else oom = 1;
                rays[0] += x.b.rays[0]; rays[1] += x.b.rays[1];
                for (int j = 0; j < 4; j++) trav[j] += x.b.trav[j];
            }
            free(x.acc);
        }
        uint32_t count = 0;
        for (uint32_t q = 0; q < L; q++)
            for (uint32_t j = 0; j < nverts[q]; j++) lvc[count++] = verts[(size_t)q * (p->max_depth - 1) + j];
#pragma omp parallel
        {
            pctx_t x;
            memset(&x, 0, sizeof x);
            x.s = s; x.c = &c; x.view_medium = view_medium; x.view = view; x.inv_cam = inv_cam;
            x.W = W; x.H = H;
            x.lvc = lvc; x.lvc_count = count; x.L = L; x.ipd = ipd;
#pragma omp for schedule(dynamic, 64)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
                x.px = px; x.py = py;
                x.light_acc = total + 3 * (size_t)pix;
                f3 r;
                uint32_t pl;
                render_pixel(&x, p, view, px, py, p->seed + k, &r, &pl);
                if (k == 0 && out->visibility) { out->visibility[2 * pix] = x.vis[0]; out->visibility[2 * pix + 1] = x.vis[1]; }
                if (k == 0 && out->albedo) {
                    for (int j = 0; j < 3; j++) out->albedo[4 * pix + j] = x.alb[j];
                    out->albedo[4 * pix + 3] = 1;
                }
                if (k == 0 && out->depth) {
                    out->depth[4 * pix] = x.dep[0]; out->depth[4 * pix + 1] = x.dep[1];
                    out->depth[4 * pix + 2] = 1; out->depth[4 * pix + 3] = 1;
                }
                if (k == 0 && out->prev_uv) { out->prev_uv[2 * pix] = x.uv[0]; out->prev_uv[2 * pix + 1] = x.uv[1]; }
                if (out->path_length) out->path_length[(size_t)k * npix + pix] = pl;
                sum[pix] = add(sum[pix], r);
            }
#pragma omp critical
            {
                rays[0] += x.rays[0]; rays[1] += x.rays[1];
                for (int j = 0; j < 4; j++) trav[j] += x.trav[j];
            }
        }
    }
    if (!oom && out->hdr)
        for (uint32_t i = 0; i < npix; i++) {
            out->hdr[4 * i + 0] = sum[i].x / (real)p->spp;
            out->hdr[4 * i + 1] = sum[i].y / (real)p->spp;
            out->hdr[4 * i + 2] = sum[i].z / (real)p->spp;
            out->hdr[4 * i + 3] = 1;
        }
    if (out->ray_counts) { out->ray_counts[0] = rays[0]; out->ray_counts[1] = rays[1]; }
    if (out->traversal) for (int j = 0; j < 4; j++) out->traversal[j] = trav[j];
    free(total); free(sum); free(verts); free(lvc); free(nverts);
    return oom ? fail("out of memory") : 0;
}

/* ------------------------------------------- screen-space ReSTIR DI reuse */
/* The north-star reuse mode (ST_FLAG_RESTIR_DI_REUSE; the reference reuses
 * one entry of a world-space hash grid instead, testrenderer.slang:697-709).
 * Per sample, in sample order:
 *  1. every pixel's path, with RIS at the primary vertex keeping its
 *     reservoir (ris_reservoir) and plain NEE deeper;
 *  2. temporal reuse: the reservoir against the same pixel's final reservoir
 *     of the previous sample (ReservoirReuse, :608-663: pairwise MIS with
 *     mis(a, b) = a^2 / (a^2 + b^2), M clamped at maxM * candidates);
 *  3. spatial reuse over `spatial_neighbours` pixels drawn in a disc of
 *     `spatial_radius` pixels (depth within 10%, shading normals within
 *     cos = 0.9), then the shadow ray of the final sample (ConnectToLight
 *     with ReSTIR, :730-747) added to the pixel's fixed-point sum. */
typedef struct {
    int hit;
    shading_t sd;
    material_t m;
    f3 ldin;
} prim_t;
static void prim_make(const or_scene* s, const shading_t* sd, uint32_t dir, prim_t* P) {
    P->hit = 1;
    P->sd = *sd;
    P->m = load_material_sd(s, sd);
    P->ldin = to_local(sd, neg(unpack_normal(dir)));
}
/* target function of a light-sample seed at a primary vertex (0 when unusable) */
static real seed_target(pctx_t* x, const prim_t* P, const real seed[4], rsd_t* r) {
    x->ps.path_len = 1;
    rsd_init(r, x, &P->m, &P->sd, P->ldin, seed);
    return r->src_pdf > 0 ? rsd_target(r) : 0.0f;
}
static real mis_pair(real a, real b) { /* mis(a, b), testrenderer.slang:287-290 */
    const real a2 = a * a, d = a2 + b * b;
    return d > 0 ? a2 / d : 0.0f;
}

/* or_render_cached's record of path g: kind, query, throughput, the radiance
 * gathered before (training rows), the path's fp32 and fixed-point parts.
 * nee: the path's final fixed-point sums; extra: fixed point added to the
 * path before its cache-depth vertex outside its own shadow rays (the
 * screen-space reuse's primary NEE). */
#define OR_CACHE_REC 25
static void cache_store(const cache_t* c, f3 out_final, const uint32_t nee[3], const uint32_t extra[3], float* r) {
    memset(r, 0, OR_CACHE_REC * sizeof(float));
    r[0] = (float)c->kind;
    if (c->kind) {
        for (int k = 0; k < 12; k++) r[1 + k] = c->q[k];
        for (int k = 0; k < 3; k++) r[13 + k] = c->thr[k];
    }
    if (c->kind == 2) {
        const real ob[3] = {c->out_before.x, c->out_before.y, c->out_before.z};
        for (int k = 0; k < 3; k++) r[16 + k] = ob[k] + (real)(c->nee_before[k] + extra[k]) / (real)OR_Q;
    }
    r[19] = out_final.x; r[20] = out_final.y; r[21] = out_final.z;
    for (int k = 0; k < 3; k++) r[22 + k] = (real)nee[k] / (real)OR_Q;
}

static int render_reuse(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads,
                        const cfg_t* c, const uint32_t* cc, float* recs) {
    const uint32_t W = p->width, H = p->height;
    const size_t npix = (size_t)W * H;
    prim_t* prim = (prim_t*)calloc(npix, sizeof(prim_t));
    real* rinit = (real*)calloc(npix * 8, sizeof(real));
    real* rtemp = (real*)calloc(npix * 8, sizeof(real));
    real* rprev = (real*)calloc(npix * 8, sizeof(real));
    f3* em = (f3*)calloc(npix, sizeof(f3));
    uint32_t* nee = (uint32_t*)calloc(npix * 3, sizeof(uint32_t));
    f3* sum = (f3*)calloc(npix, sizeof(f3));
    cache_t* crec = cc ? (cache_t*)calloc(npix, sizeof(cache_t)) : NULL;
    uint32_t* nee_path = cc ? (uint32_t*)calloc(npix * 3, sizeof(uint32_t)) : NULL;
    if (!prim || !rinit || !rtemp || !rprev || !em || !nee || !sum || (cc && (!crec || !nee_path))) {
        free(prim); free(rinit); free(rtemp); free(rprev); free(em); free(nee); free(sum); free(crec); free(nee_path);
        return fail("out of memory");
    }
    const uint32_t view_medium = camera_medium(s, view);
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    uint64_t rays[2] = {0, 0}, trav[4] = {0, 0, 0, 0};
    const real mmax = p->max_m * (real)c->cand;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (uint32_t k = 0; k < p->spp; k++) {
        const uint32_t seed = p->seed + k;
#pragma omp parallel
        {
            pctx_t x;
            memset(&x, 0, sizeof x);
            x.s = s;
            x.c = c;
            x.view_medium = view_medium;
            x.view = view;
            x.inv_cam = inv_cam;
            x.W = W;
            x.H = H;
            x.reuse = 1;
            cache_t xc;
            if (cc) {
                memset(&xc, 0, sizeof xc);
                xc.depth = cc[0]; xc.stride = cc[1]; xc.frame = cc[2];
                x.cache = &xc;
            }
            /* 1. paths, primary reservoirs */
#pragma omp for schedule(dynamic, 64)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
                x.px = px;
                x.py = py;
                x.prim_hit = 0;
                memset(x.res, 0, sizeof x.res);
                f3 r;
                uint32_t pl;
                if (cc) xc.g = k * (uint32_t)npix + (uint32_t)pix;
                render_pixel(&x, p, view, px, py, seed, &r, &pl);
                if (cc) {
                    crec[pix] = xc;
                    for (int j = 0; j < 3; j++) nee_path[3 * pix + j] = x.nee[j];
                }
                if (k == 0 && out->visibility) { out->visibility[2 * pix] = x.vis[0]; out->visibility[2 * pix + 1] = x.vis[1]; }
                if (k == 0 && out->albedo) {
                    for (int j = 0; j < 3; j++) out->albedo[4 * pix + j] = x.alb[j];
                    out->albedo[4 * pix + 3] = 1;
                }
                if (k == 0 && out->depth) {
                    out->depth[4 * pix] = x.dep[0]; out->depth[4 * pix + 1] = x.dep[1];
                    out->depth[4 * pix + 2] = 1; out->depth[4 * pix + 3] = 1;
                }
                if (k == 0 && out->prev_uv) { out->prev_uv[2 * pix] = x.uv[0]; out->prev_uv[2 * pix + 1] = x.uv[1]; }
                if (out->path_length) out->path_length[(size_t)k * npix + pix] = pl;
                em[pix] = x.out;
                for (int j = 0; j < 3; j++) nee[3 * pix + j] = x.nee[j];
                prim[pix].hit = 0;
                if (x.prim_hit) prim_make(s, &x.prim_sd, x.prim_dir, &prim[pix]);
                memcpy(&rinit[8 * pix], x.res, sizeof x.res);
                if (!x.prim_hit) memset(&rinit[8 * pix], 0, sizeof x.res); /* a miss keeps no reservoir */
            }
            /* 2. temporal reuse against the previous sample's final reservoir */
#pragma omp for schedule(dynamic, 64)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
                real* b = &rtemp[8 * pix];
                memcpy(b, &rinit[8 * pix], 8 * sizeof(real));
                const real* q = &rprev[8 * pix];
                if (b[6] > 0 && k > 0 && q[6] > 0) {
                    rsd_t xr;
                    const real tq = seed_target(&x, &prim[pix], q, &xr);
                    rng_t rng;
                    rng_init(&rng, seed, px, py, 1u << 20);
                    real u[4];
                    rng_nextf(&rng, u);
                    const real m0 = mis_pair(b[6] * b[4], q[6] * b[4]), m1 = mis_pair(q[6] * q[4], b[6] * tq);
                    const real w0 = m0 * b[4] * b[5], w1 = m1 * tq * q[5];
                    const real total = w0 + w1;
                    const real M = b[6] + q[6];
                    real target = b[4];
                    if (u[0] * total <= w1 && tq > 0) {
                        memcpy(b, q, 4 * sizeof(real));
                        target = tq;
                    }
                    b[4] = target;
                    b[5] = target > 0 ? total / target : 0.0f;
                    b[6] = minf(M, mmax);
                }
            }
            /* 3. spatial reuse, then the final sample's shadow ray */
#pragma omp for schedule(dynamic, 64)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
                real b[8];
                memcpy(b, &rtemp[8 * pix], sizeof b);
                if (!(b[6] > 0)) {
                    memset(&rprev[8 * pix], 0, 8 * sizeof(real));
                    memcpy(&rprev[8 * pix], b, 4 * sizeof(real));
                    continue;
                }
                const prim_t* P = &prim[pix];
                const f3 nc = unpack_normal(P->sd.sn);
                rsd_t cur;
                seed_target(&x, P, b, &cur);
                rng_t rng;
                rng_init(&rng, seed, px, py, (1u << 20) + 64);
                for (uint32_t n = 0; n < p->spatial_neighbours; n++) {
                    real u[4], d[2];
                    rng_nextf(&rng, u);
                    sample_concentric_disc(u[0], u[1], d);
                    const int qx = (int)px + (int)floorr(d[0] * p->spatial_radius + 0.5f);
                    const int qy = (int)py + (int)floorr(d[1] * p->spatial_radius + 0.5f);
                    if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H || (qx == (int)px && qy == (int)py)) continue;
                    const size_t qp = (size_t)qy * W + (size_t)qx;
                    const real* q = &rtemp[8 * qp];
                    if (!(q[6] > 0)) continue;
                    if (fabsr(q[7] - b[7]) > 0.1f * b[7]) continue;
                    const prim_t* Q = &prim[qp];
                    if (dot(unpack_normal(Q->sd.sn), nc) < 0.9f) continue;
                    rsd_t xn, xq;
                    const real tn = seed_target(&x, P, q, &xn);   /* the neighbour's seed here */
                    const real tcq = seed_target(&x, Q, b, &xq);  /* our seed there */
                    const real m0 = mis_pair(b[6] * b[4], q[6] * tcq), m1 = mis_pair(q[6] * q[4], b[6] * tn);
                    const real w0 = m0 * b[4] * b[5], w1 = m1 * tn * q[5];
                    const real total = w0 + w1;
                    if (u[2] * total <= w1 && tn > 0) {
                        memcpy(b, q, 4 * sizeof(real));
                        b[4] = tn;
                        cur = xn;
                    }
                    b[5] = b[4] > 0 ? total / b[4] : 0.0f;
                    b[6] = minf(b[6] + q[6], mmax);
                }
                memcpy(&rprev[8 * pix], b, sizeof b);
                if (b[5] > 0 && b[4] > 0 && 2 >= c->min_len) {
                    const real Wp = b[5] * cur.wtoa; /* 1/pdfA -> 1/pdfW */
                    const f3 contrib = mulf(cur.contrib, Wp); /* primary throughput is 1 */
                    hit_t h = trace(s, cur.vo, cur.vd, cur.vtmax, 1, &x.trav[2], &x.trav[3]);
                    x.rays[1]++;
                    if (!h.hit && cur.mis_w > 0 && !all_le0(contrib)) {
                        const f3 rw = mulf(contrib, cur.mis_w);
                        const real qv[3] = {clamp0(rw.x) * (real)OR_Q, clamp0(rw.y) * (real)OR_Q, clamp0(rw.z) * (real)OR_Q};
                        for (int j = 0; j < 3; j++) nee[3 * pix + j] += (uint32_t)(qv[j] >= 2147483647.0f ? INT_MAX : (int32_t)qv[j]);
                    }
                }
            }
            /* resolve: the running sum over samples (ProcessAtomicOutput) */
#pragma omp for schedule(static)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                if (cc && recs) {
                    uint32_t extra[3];
                    for (int j = 0; j < 3; j++) extra[j] = nee[3 * pix + j] - nee_path[3 * pix + j];
                    cache_store(&crec[pix], em[pix], &nee[3 * pix], extra, recs + OR_CACHE_REC * ((size_t)k * npix + (size_t)pix));
                }
                sum[pix].x += em[pix].x + (real)nee[3 * pix] / (real)OR_Q;
                sum[pix].y += em[pix].y + (real)nee[3 * pix + 1] / (real)OR_Q;
                sum[pix].z += em[pix].z + (real)nee[3 * pix + 2] / (real)OR_Q;
            }
#pragma omp critical
            {
                rays[0] += x.rays[0]; rays[1] += x.rays[1];
                for (int j = 0; j < 4; j++) trav[j] += x.trav[j];
            }
        }
    }
    if (out->hdr)
        for (size_t i = 0; i < npix; i++) {
            out->hdr[4 * i + 0] = sum[i].x / (real)p->spp;
            out->hdr[4 * i + 1] = sum[i].y / (real)p->spp;
            out->hdr[4 * i + 2] = sum[i].z / (real)p->spp;
            out->hdr[4 * i + 3] = 1;
        }
    if (out->ray_counts) { out->ray_counts[0] = rays[0]; out->ray_counts[1] = rays[1]; }
    if (out->traversal) for (int j = 0; j < 4; j++) out->traversal[j] = trav[j];
    free(prim); free(rinit); free(rtemp); free(rprev); free(em); free(nee); free(sum); free(crec); free(nee_path);
    return 0;
}

/* The per-triangle albedo serves Lambertian camera paths with plain NEE. */
static int check_tri_albedo(const or_scene* s, uint32_t flags) {
    if (!s->tri_alb) return 0;
    if (!(flags & ST_FLAG_LAMBERTIAN)) return fail("a per-triangle albedo needs ST_FLAG_LAMBERTIAN");
    if (flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE))
        return fail("a per-triangle albedo serves camera paths with plain NEE");
    return 0;
}

static int render_camera(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads,
                         const uint32_t* cc, float* recs);

/* TestRenderer with gReSTIR_DI_Reuse: samples one after another, each
 * reusing the previous sample's hash grid (TestRenderer.cpp:595-628: clear,
 * render, build).  The build sorts a sample's entries by (empty, checksum),
 * stable in slot order. */
static int hg_cmp_n;
static const uint64_t* hg_cmp_keys;
static int hg_cmp(const void* a, const void* b) {
    const uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
    if (hg_cmp_keys[i] != hg_cmp_keys[j]) return hg_cmp_keys[i] < hg_cmp_keys[j] ? -1 : 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}
static int render_hashgrid(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads,
                           const cfg_t* c) {
    const uint32_t W = p->width, H = p->height;
    const size_t npix = (size_t)W * H;
    hgrid_t g[2];
    const uint32_t dmax = p->max_depth - 1 < 2 ? p->max_depth - 1 : 2;
    const size_t ns = npix * dmax;
    f3* sum = (f3*)calloc(npix, sizeof(f3));
    uint64_t* kin = (uint64_t*)calloc(ns, sizeof(uint64_t));
    int ok = sum && kin;
    for (int i = 0; i < 2; i++) {
        g[i].rec = (real*)calloc(ns * 16, sizeof(real));
        g[i].keys = (uint64_t*)calloc(ns, sizeof(uint64_t));
        g[i].slots = (uint32_t*)calloc(ns, sizeof(uint32_t));
        g[i].nslots = (uint32_t)ns;
        g[i].npix = (uint32_t)npix;
        g[i].dmax = dmax;
        g[i].cell = p->hashgrid_cell_size > 0 ? p->hashgrid_cell_size : 0.5f;
        ok = ok && g[i].rec && g[i].keys && g[i].slots;
    }
    if (!ok) {
        free(sum); free(kin);
        for (int i = 0; i < 2; i++) { free(g[i].rec); free(g[i].keys); free(g[i].slots); }
        return fail("out of memory");
    }
    const uint32_t view_medium = camera_medium(s, view);
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    uint64_t rays[2] = {0, 0}, trav[4] = {0, 0, 0, 0};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (uint32_t k = 0; k < p->spp; k++) {
        hgrid_t* cur = &g[k & 1];
        const hgrid_t* prev = k > 0 ? &g[(k + 1) & 1] : NULL;
        memset(cur->rec, 0, ns * 16 * sizeof(real));
#pragma omp parallel
        {
            pctx_t x;
            memset(&x, 0, sizeof x);
            x.s = s;
            x.c = c;
            x.view_medium = view_medium;
            x.view = view;
            x.inv_cam = inv_cam;
            x.W = W;
            x.H = H;
            x.hg_prev = prev;
            x.hg_cur = cur;
#pragma omp for schedule(dynamic, 64)
            for (int64_t pix = 0; pix < (int64_t)npix; pix++) {
                const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
                x.px = px;
                x.py = py;
                x.hg_pix = (uint32_t)pix;
                f3 r;
                uint32_t pl;
                render_pixel(&x, p, view, px, py, p->seed + k, &r, &pl);
                if (k == 0 && out->visibility) { out->visibility[2 * pix] = x.vis[0]; out->visibility[2 * pix + 1] = x.vis[1]; }
                if (k == 0 && out->albedo) {
                    for (int j = 0; j < 3; j++) out->albedo[4 * pix + j] = x.alb[j];
                    out->albedo[4 * pix + 3] = 1;
                }
                if (k == 0 && out->depth) {
                    out->depth[4 * pix] = x.dep[0]; out->depth[4 * pix + 1] = x.dep[1];
                    out->depth[4 * pix + 2] = 1; out->depth[4 * pix + 3] = 1;
                }
                if (k == 0 && out->prev_uv) { out->prev_uv[2 * pix] = x.uv[0]; out->prev_uv[2 * pix + 1] = x.uv[1]; }
                if (out->path_length) out->path_length[(size_t)k * npix + pix] = pl;
                sum[pix] = add(sum[pix], r);
            }
#pragma omp critical
            {
                rays[0] += x.rays[0]; rays[1] += x.rays[1];
                for (int j = 0; j < 4; j++) trav[j] += x.trav[j];
            }
        }
        /* HashGrid build (GpuHashGrid::build), deterministic */
        for (size_t i = 0; i < ns; i++) {
            const real* e = cur->rec + 16 * i;
            kin[i] = e[9] > 0 ? (uint64_t)asuint(e[11]) : (1ull << 32);
            cur->slots[i] = (uint32_t)i;
        }
        hg_cmp_keys = kin;
        hg_cmp_n = (int)ns;
        qsort(cur->slots, ns, sizeof(uint32_t), hg_cmp);
        for (size_t i = 0; i < ns; i++) cur->keys[i] = kin[cur->slots[i]];
    }
    if (out->hdr)
        for (size_t i = 0; i < npix; i++) {
            out->hdr[4 * i + 0] = sum[i].x / (real)p->spp;
            out->hdr[4 * i + 1] = sum[i].y / (real)p->spp;
            out->hdr[4 * i + 2] = sum[i].z / (real)p->spp;
            out->hdr[4 * i + 3] = 1;
        }
    if (out->ray_counts) { out->ray_counts[0] = rays[0]; out->ray_counts[1] = rays[1]; }
    if (out->traversal) for (int j = 0; j < 4; j++) out->traversal[j] = trav[j];
    free(sum); free(kin);
    for (int i = 0; i < 2; i++) { free(g[i].rec); free(g[i].keys); free(g[i].slots); }
    return 0;
}

int or_render(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads) {
    return render_camera(s, view, p, out, threads, NULL, NULL);
}

int or_render_cached(or_scene* s, const st_view* view, const st_render_params* p, uint32_t depth, uint32_t stride,
                     uint32_t frame, const or_outputs* out, float* recs, int threads) {
    if (!s || !view || !p || !out) return fail("null argument");
    if (p->flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC)) return fail("the radiance cache ends camera paths");
    if ((p->flags & ST_FLAG_RESTIR_DI) && !(p->flags & ST_FLAG_RESTIR_DI_REUSE))
        return fail("with the radiance cache, ReSTIR DI runs at primary vertices (ST_FLAG_RESTIR_DI_REUSE)");
    if (p->flags & ST_FLAG_RESTIR_DI_DEEP) return fail("with the radiance cache, deeper vertices take one-sample NEE");
    if (p->flags & ST_FLAG_FLOAT_NEE) return fail("fp32 NEE with the radiance cache is not restated");
    if (depth < 1 || depth >= p->max_depth || stride < 1) return fail("cache depth in [1, max_depth), stride >= 1");
    if ((p->flags & ST_FLAG_RESTIR_DI_REUSE) && depth < 2) return fail("with screen-space reuse: cache depth >= 2");
    const uint32_t cc[3] = {depth, stride, frame};
    return render_camera(s, view, p, out, threads, cc, recs);
}

static int render_camera(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out, int threads,
                         const uint32_t* cc, float* recs) {
    if (!s || !view || !p || !out) return fail("null argument");
    if (p->width == 0 || p->height == 0 || p->spp == 0) return fail("empty render");
    if (check_tri_albedo(s, p->flags)) return -1;
    if (p->flags & ST_FLAG_TRACE_FROM_LIGHT) return render_light(s, view, p, out, threads);
    if (p->flags & ST_FLAG_BDPT_LVC) return render_vc(s, view, p, out, threads);
    if (s->nvolinst > 0 && (p->flags & (ST_FLAG_RESTIR_DI_REUSE | ST_FLAG_RESTIR_DI_HASHGRID)))
        return fail("ReSTIR DI reuse with media is not supported (RIS is)");
    if (s->ntex && (p->flags & ST_FLAG_RESTIR_DI_REUSE)) return fail("the screen-space reuse with material images");
    if (s->ntex && (p->flags & ST_FLAG_RESTIR_DI_REUSE)) return fail("the screen-space reuse with material images");
    const int reuse = (p->flags & ST_FLAG_RESTIR_DI_REUSE) != 0;
    if (reuse && !(p->flags & ST_FLAG_RESTIR_DI)) return fail("ST_FLAG_RESTIR_DI_REUSE needs ST_FLAG_RESTIR_DI");
    if (reuse && (p->flags & (ST_FLAG_FLOAT_NEE | ST_FLAG_PIXEL_JITTER)))
        return fail("screen-space reuse with fp32 NEE or pixel jitter is not restated");
    cfg_t c;
    memset(&c, 0, sizeof c);
    c.flags = p->flags;
    c.env_prob = p->environment_sample_probability;
    c.max_depth = p->max_depth;
    c.min_len = p->min_path_length;
    c.cand = p->candidate_samples;
    c.lambert = (p->flags & ST_FLAG_LAMBERTIAN) ? 1 : ((p->flags & ST_FLAG_LAMBERT_GGX) ? 2 : 0);
    c.max_null = p->max_null_collisions;
    c.rr_start = p->rr_start_depth;
    c.max_m = p->max_m;
    if (reuse) return render_reuse(s, view, p, out, threads, &c, cc, recs);
    if (p->flags & ST_FLAG_RESTIR_DI_HASHGRID) {
        if (!(p->flags & ST_FLAG_RESTIR_DI)) return fail("ST_FLAG_RESTIR_DI_HASHGRID needs ST_FLAG_RESTIR_DI");
        if (cc) return fail("hash-grid reuse with the radiance cache is not restated");
        if (p->max_depth < 2) return fail("hash-grid reuse needs max_depth >= 2");
        return render_hashgrid(s, view, p, out, threads, &c);
    }
    const uint32_t view_medium = camera_medium(s, view);
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    uint64_t rays[2] = {0, 0}, trav[4] = {0, 0, 0, 0};
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const uint32_t W = p->width, H = p->height;
#pragma omp parallel
    {
        pctx_t x;
        memset(&x, 0, sizeof x);
        x.s = s;
        x.c = &c;
        x.view_medium = view_medium;
        x.view = view;
        x.inv_cam = inv_cam;
        x.W = W;
        x.H = H;
        cache_t xc;
        if (cc) {
            memset(&xc, 0, sizeof xc);
            xc.depth = cc[0]; xc.stride = cc[1]; xc.frame = cc[2];
            x.cache = &xc;
        }
#pragma omp for schedule(dynamic, 64)
        for (int64_t pix = 0; pix < (int64_t)W * H; pix++) {
            uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
            x.px = px;
            x.py = py;
            f3 sum = v3(0, 0, 0);
            for (uint32_t k = 0; k < p->spp; k++) {
                f3 r;
                uint32_t pl;
                if (cc) xc.g = k * W * H + (uint32_t)pix;
                render_pixel(&x, p, view, px, py, p->seed + k, &r, &pl);
                if (cc && recs) {
                    const uint32_t none[3] = {0, 0, 0};
                    cache_store(&xc, x.out, x.nee, none, recs + OR_CACHE_REC * ((size_t)k * W * H + (size_t)pix));
                }
                if (k == 0 && out->visibility) { out->visibility[2 * pix] = x.vis[0]; out->visibility[2 * pix + 1] = x.vis[1]; }
                if (k == 0 && out->albedo) {
                    for (int j = 0; j < 3; j++) out->albedo[4 * pix + j] = x.alb[j];
                    out->albedo[4 * pix + 3] = 1;
                }
                if (k == 0 && out->depth) {
                    out->depth[4 * pix] = x.dep[0]; out->depth[4 * pix + 1] = x.dep[1];
                    out->depth[4 * pix + 2] = 1; out->depth[4 * pix + 3] = 1;
                }
                if (k == 0 && out->prev_uv) { out->prev_uv[2 * pix] = x.uv[0]; out->prev_uv[2 * pix + 1] = x.uv[1]; }
                if (out->path_length) out->path_length[(size_t)k * W * H + pix] = pl;
                sum = add(sum, r);
            }
            if (out->hdr) {
                out->hdr[4 * pix + 0] = sum.x / (real)p->spp;
                out->hdr[4 * pix + 1] = sum.y / (real)p->spp;
                out->hdr[4 * pix + 2] = sum.z / (real)p->spp;
                out->hdr[4 * pix + 3] = 1;
            }
        }
#pragma omp critical
        {
            rays[0] += x.rays[0]; rays[1] += x.rays[1];
            for (int k = 0; k < 4; k++) trav[k] += x.trav[k];
        }
    }
    if (out->ray_counts) { out->ray_counts[0] = rays[0]; out->ray_counts[1] = rays[1]; }
    if (out->traversal) for (int k = 0; k < 4; k++) out->traversal[k] = trav[k];
    return 0;
}

/* ------------------------------------------------------- path replay */
uint32_t or_scene_triangle_count(const or_scene* s) { return s ? s->ntris : 0; }

int or_scene_set_triangle_albedo(or_scene* s, const float* rgb, uint32_t count) {
    if (!s) return fail("null scene");
    if (!rgb) { free(s->tri_alb); s->tri_alb = NULL; return 0; }
    if (s->nsph) return fail("a per-triangle albedo in a scene with sphere instances");
    if (count != s->ntris) return fail("one albedo per flattened triangle");
    if (!s->tri_alb) s->tri_alb = (float*)malloc(sizeof(float) * 3 * ((size_t)count + 1));
    memcpy(s->tri_alb, rgb, sizeof(float) * 3 * (size_t)count);
    return 0;
}

int or_scene_set_weight_grid(or_scene* s, uint32_t vol, const void* grid, uint64_t size) {
    if (!s || vol >= s->nvols) return fail("no such volume");
    if (!grid) { free(s->volw[vol]); s->volw[vol] = NULL; return 0; }
    if (size != s->volsz[vol]) return fail("the weight grid must have the grid's size");
    if (!s->volw[vol]) s->volw[vol] = (uint8_t*)malloc(size + 16);
    memcpy(s->volw[vol], grid, size);
    return 0;
}

/* Path-replay adjoint of the camera-path estimator (no tape): for each path,
 * pass 1 renders it and keeps its fp32 radiance (shadow rays traced after
 * each bounce), pass 2 replays it from the same seed and, at every
 * differentiable event, adds dLoss/dL times the radiance still to come times
 * the event's log-derivative: 1/albedo_c at a Lambertian vertex (the
 * triangle's entry of the albedo table), and at each delta- or
 * ratio-tracking collision in grid grad_vol the derivative of log sigma_t
 * (real) or log(majorant - sigma_t) (null) with respect to the value read.
 * Sampling decisions (tracking events, roulette, MIS) are held fixed
 * ("detached"), which keeps the estimator unbiased; a fixed-seed central
 * difference over a weight grid (or_scene_set_weight_grid) or over the
 * albedo table with roulette off is exactly its derivative.  NEE is fp32
 * (ST_FLAG_FLOAT_NEE).  d_hdr: dLoss/dHDR, RGBA per pixel; gradients are
 * added to grad_alb (3 per flattened triangle) and grad_grid (one per 4-byte
 * word of grid grad_vol); either may be NULL. */
int or_adjoint(or_scene* s, const st_view* view, const st_render_params* p, const float* d_hdr, double* grad_alb,
               double* grad_grid, uint32_t grad_vol, int threads) {
    if (!s || !view || !p || !d_hdr) return fail("null argument");
    if (p->width == 0 || p->height == 0 || p->spp == 0) return fail("empty render");
    if (p->flags & (ST_FLAG_TRACE_FROM_LIGHT | ST_FLAG_BDPT_LVC | ST_FLAG_RESTIR_DI | ST_FLAG_RESTIR_DI_REUSE))
        return fail("the adjoint replays camera paths with plain NEE");
    if (grad_alb && (!s->tri_alb || !(p->flags & ST_FLAG_LAMBERTIAN)))
        return fail("albedo gradients need a per-triangle albedo and ST_FLAG_LAMBERTIAN");
    if (grad_grid && grad_vol >= s->nvols) return fail("no such volume");
    if (check_tri_albedo(s, p->flags)) return -1;
    cfg_t c;
    memset(&c, 0, sizeof c);
    c.flags = p->flags | ST_FLAG_FLOAT_NEE;
    c.env_prob = p->environment_sample_probability;
    c.max_depth = p->max_depth;
    c.min_len = p->min_path_length;
    c.cand = p->candidate_samples;
    c.lambert = (p->flags & ST_FLAG_LAMBERTIAN) ? 1 : ((p->flags & ST_FLAG_LAMBERT_GGX) ? 2 : 0);
    c.max_null = p->max_null_collisions;
    c.rr_start = p->rr_start_depth;
    const uint32_t view_medium = camera_medium(s, view);
    const st_transform inv_cam = inverse_xf(&view->camera_to_world);
    const uint32_t W = p->width, H = p->height;
    const size_t na = grad_alb ? 3 * (size_t)s->ntris : 0, ng = grad_grid ? (size_t)(s->volsz[grad_vol] / 4) : 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        pctx_t x;
        memset(&x, 0, sizeof x);
        x.s = s;
        x.c = &c;
        x.view_medium = view_medium;
        x.view = view;
        x.inv_cam = inv_cam;
        x.W = W;
        x.H = H;
        adj_t A;
        memset(&A, 0, sizeof A);
        A.grad_alb = na ? (double*)calloc(na, sizeof(double)) : NULL;
        A.grad_grid = ng ? (double*)calloc(ng, sizeof(double)) : NULL;
        A.grad_vol = grad_vol;
        x.adj = &A;
#pragma omp for schedule(dynamic, 64)
        for (int64_t pix = 0; pix < (int64_t)W * H; pix++) {
            const uint32_t px = (uint32_t)(pix % W), py = (uint32_t)(pix / W);
            x.px = px;
            x.py = py;
            A.g = v3(d_hdr[4 * pix] / (real)p->spp, d_hdr[4 * pix + 1] / (real)p->spp, d_hdr[4 * pix + 2] / (real)p->spp);
            for (uint32_t k = 0; k < p->spp; k++) {
                f3 r;
                uint32_t pl;
                A.pass = 1;
                A.acc = v3(0, 0, 0);
                render_pixel(&x, p, view, px, py, p->seed + k, &r, &pl);
                A.total = A.acc;
                A.pass = 2;
                A.acc = v3(0, 0, 0);
                render_pixel(&x, p, view, px, py, p->seed + k, &r, &pl);
            }
        }
#pragma omp critical
        {
            for (size_t i = 0; i < na; i++) grad_alb[i] += A.grad_alb[i];
            for (size_t i = 0; i < ng; i++) grad_grid[i] += A.grad_grid[i];
        }
        free(A.grad_alb);
        free(A.grad_grid);
    }
    return 0;
}

/* ------------------------------------------------------------ tonemapper */
/* Tonemapper::render (src/App/Tonemapper.cpp:49-96): maxReduce then main of
 * src/Shaders/kernels/tonemap.slang:124-192, operators :30-117, viridisQuintic
 * and rgbToSrgb of compat/common.h:88-104, in the reference's operation order. */
static f3 tm_div1p(f3 c) { return v3(c.x / (1 + c.x), c.y / (1 + c.y), c.z / (1 + c.z)); }
static real tm_u2(real x) {
    const real A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
static real tm_filmic1(real c) {
    c = maxf(0.0f, c - 0.004f);
    return (c * (6.2f * c + 0.5f)) / (c * (6.2f * c + 1.7f) + 0.06f);
}
static real tm_rtt1(real v) {
    const real a = v * (v + 0.0245786f) - 0.000090537f;
    const real b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}
static real tm_aces_approx1(real v) {
    const real a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    v *= 0.6f;
    return saturate((v * (a * v + b)) / (v * (c * v + d) + e));
}
static f3 tm_viridis(real x) {
    const real x2 = x * x, x3 = x2 * x, x4 = x * x3, x5 = x2 * x3;
    return v3((0.280268003f + -0.143510503f * x + 2.225793877f * x2 + -14.815088879f * x3) + (25.212752309f * x4 + -11.772589584f * x5),
              (-0.002117546f + 1.617109353f * x + -1.909305070f * x2 + 2.701152864f * x3) + (-1.685288385f * x4 + 0.178738871f * x5),
              (0.300805501f + 2.614650302f * x + -12.019139090f * x2 + 28.933559110f * x3) + (-33.491294770f * x4 + 13.762053843f * x5));
}
static real tm_srgb1(real c) { return c <= 0.0031308f ? c * 12.92f : powr(c * 1.055f, 1 / 2.4f) - 0.055f; }

static f3 tm_apply(int mode, f3 c, const real mx[4]) {
    switch (mode) {
        case ST_TONEMAP_REINHARD: {
            const real l = luminance(c);
            const f3 tc = tm_div1p(c), a = divf(c, 1 + l);
            return add(a, mul(sub(tc, a), tc));
        }
        case ST_TONEMAP_REINHARD_EXTENDED: {
            const f3 m = v3(mx[0] == 0 ? 1.0f : mx[0], mx[1] == 0 ? 1.0f : mx[1], mx[2] == 0 ? 1.0f : mx[2]);
            return mul(tm_div1p(c), add(v3(1, 1, 1), divv(c, mul(m, m))));
        }
        case ST_TONEMAP_REINHARD_LUMINANCE: {
            const real l = luminance(c), l1 = l / (1 + l);
            return mulf(c, l1 / l);
        }
        case ST_TONEMAP_REINHARD_LUMINANCE_EXTENDED: {
            const real l = luminance(c), m = mx[3] == 0 ? 1.0f : mx[3];
            const real l1 = (l / (1 + l)) * (1 + l / (m * m));
            return mulf(c, l1 / l);
        }
        case ST_TONEMAP_UNCHARTED2: {
            const real w = tm_u2(mx[3] == 0 ? 1.0f : mx[3]);
            return divf(v3(tm_u2(c.x), tm_u2(c.y), tm_u2(c.z)), w);
        }
        case ST_TONEMAP_FILMIC: return v3(tm_filmic1(c.x), tm_filmic1(c.y), tm_filmic1(c.z));
        case ST_TONEMAP_ACES: {
            const f3 i = v3(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
                            0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
            const f3 r = v3(tm_rtt1(i.x), tm_rtt1(i.y), tm_rtt1(i.z));
            return v3(saturate(1.60475f * r.x + -0.53108f * r.y + -0.07367f * r.z),
                      saturate(-0.10208f * r.x + 1.10813f * r.y + -0.00605f * r.z),
                      saturate(-0.00327f * r.x + -0.07276f * r.y + 1.07602f * r.z));
        }
        case ST_TONEMAP_ACES_APPROX: return v3(tm_aces_approx1(c.x), tm_aces_approx1(c.y), tm_aces_approx1(c.z));
        case ST_TONEMAP_VIRIDIS_R: return tm_viridis(saturate(luminance(c)));
        case ST_TONEMAP_VIRIDIS_LENGTH_RGB: return tm_viridis(saturate(luminance(c) / (mx[3] == 0 ? 1.0f : mx[3])));
        default: return c;
    }
}

int or_tonemap(const float* hdr, const float* albedo, uint32_t w, uint32_t h, float exposure, int mode, int gamma, float* out,
               float max_out[4]) {
    if (!hdr || !out) return fail("null argument");
    const size_t n = (size_t)w * h;
    uint32_t mq[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        f3 v = v3(hdr[4 * i], hdr[4 * i + 1], hdr[4 * i + 2]);
        if (albedo) v = mul(v, v3(albedo[4 * i], albedo[4 * i + 1], albedo[4 * i + 2]));
        const real l = luminance(v);
        if (v.x != v.x || v.y != v.y || v.z != v.z || l != l || l <= 0) continue;
        const real q[4] = {v.x * 16384.0f, v.y * 16384.0f, v.z * 16384.0f, l * 16384.0f};
        for (int k = 0; k < 4; k++) {
            const real cl = minf(maxf(q[k], 0.0f), 4294967296.0f);
            const uint32_t u = cl >= 4294967296.0f ? 0xffffffffu : (uint32_t)cl;
            if (u > mq[k]) mq[k] = u;
        }
    }
    real mx[4];
    for (int k = 0; k < 4; k++) mx[k] = (real)mq[k] / 16384.0f;
    if (max_out) for (int k = 0; k < 4; k++) max_out[k] = mx[k];
    const real scale = powr(2.0f, exposure);
    for (size_t i = 0; i < n; i++) {
        f3 r = v3(hdr[4 * i], hdr[4 * i + 1], hdr[4 * i + 2]);
        if (albedo) r = mul(r, add(v3(1e-2f, 1e-2f, 1e-2f), v3(albedo[4 * i], albedo[4 * i + 1], albedo[4 * i + 2])));
        r = mulf(r, scale);
        r = tm_apply(mode, r, mx);
        if (gamma) r = v3(tm_srgb1(r.x), tm_srgb1(r.y), tm_srgb1(r.z));
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = 1;
    }
    return 0;
}
