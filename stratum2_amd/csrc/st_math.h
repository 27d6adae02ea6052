// st_math.h — the transcendental functions of the path tracer, written out
// as float operations in a fixed order so that every result is one function
// of its inputs (IEEE-754 single precision, round to nearest, no contraction:
// the library is built with -ffp-contract=off).  The CPU oracle carries a
// second copy with the same operation sequence, so the GPU and the oracle
// sample the same directions and free-flight distances bit for bit.
//
// Method: Cody-Waite argument reduction and the minimax polynomials of the
// Cephes library's single-precision routines (S. L. Moshier, "Methods and
// Programs for Mathematical Functions", 1989; sinf.c, atanf.c, asinf.c,
// logf.c, expf.c), evaluated entirely in float.  Errors against the exact
// values, measured over the ranges the renderer uses: sin/cos <= 2 ulp,
// atan2/acos <= 3 ulp, log <= 2 ulp, exp <= 2 ulp (tests/test_oracle_math.py).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

// sin and cos for |x| < 8192: octant j of x * 4/pi, x - j * pi/4 in three
// exactly representable parts, then the octant's polynomial on [-pi/4, pi/4].
__device__ __forceinline__ float poly_sin(float x, float z) {
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
__device__ __forceinline__ float poly_cos(float z) {
    const float y = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z;
    return (y - 0.5f * z) + 1.0f;
}
__device__ __forceinline__ float reduce_quarter_pi(float a, int& j) {
    j = (int)(a * 1.27323954473516f);
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    return ((a - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}
__device__ __forceinline__ float sin_(float x) {
    float sign = 1.0f, a = x;
    if (a < 0) {
        a = -a;
        sign = -1.0f;
    }
    int j;
    const float r = reduce_quarter_pi(a, j);
    if (j > 3) {
        sign = -sign;
        j -= 4;
    }
    const float z = r * r;
    const float y = (j == 1 || j == 2) ? poly_cos(z) : poly_sin(r, z);
    return sign < 0 ? -y : y;
}
__device__ __forceinline__ float cos_(float x) {
    const float a = x < 0 ? -x : x;
    int j;
    const float r = reduce_quarter_pi(a, j);
    float sign = 1.0f;
    if (j > 3) {
        j -= 4;
        sign = -sign;
    }
    if (j > 1) sign = -sign;
    const float z = r * r;
    const float y = (j == 1 || j == 2) ? poly_sin(r, z) : poly_cos(z);
    return sign < 0 ? -y : y;
}

// atan on [0, inf): reduced to [0, tan(pi/8)] by pi/2 - atan(1/x) and
// pi/4 + atan((x-1)/(x+1)).
__device__ __forceinline__ float atan_pos(float x) {
    float y = 0.0f;
    if (x > 2.414213562373095f) {
        y = 1.5707963267948966f;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = 0.7853981633974483f;
        x = (x - 1.0f) / (x + 1.0f);
    }
    const float z = x * x;
    return y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
}
// atan2 for x != 0 (the callers handle x == 0, compat/common.h:137-140)
__device__ __forceinline__ float atan2_(float y, float x) {
    const float q = y / x;
    const float t = q < 0 ? -atan_pos(-q) : atan_pos(q);
    if (x > 0) return t;
    return y < 0 ? t - 3.14159265358979f : t + 3.14159265358979f;
}
// asin on [0, 1] (acos below)
__device__ __forceinline__ float asin_pos(float a, bool& big) {
    float x, z;
    big = a > 0.5f;
    if (big) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
    } else {
        x = a;
        z = x * x;
    }
    return ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z +
            1.6666752422e-1f) * z * x + x;
}
// acos on [-1, 1]: pi/2 - asin(x), or 2 asin(sqrt((1 - |x|) / 2)) near |x| = 1
__device__ __forceinline__ float acos_(float x) {
    const float a = x < 0 ? -x : x;
    bool big;
    const float s = asin_pos(a, big);
    if (big) {
        const float r = s + s;  // acos(|x|)
        return x < 0 ? 3.14159265358979f - r : r;
    }
    return x < 0 ? 1.5707963267948966f + s : 1.5707963267948966f - s;
}

// log for x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log(m) by a
// polynomial in m - 1, e ln 2 in two parts.  log(0) = -inf, log(x < 0) = NaN.
__device__ __forceinline__ float log_(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__int_as_float(0x7f800000) : __int_as_float(0x7fc00000);
    if (x == __int_as_float(0x7f800000)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range first
        x *= 16777216.0f;
        e = -24;
    }
    const uint32_t b = __float_as_uint(x);
    e += (int)((b >> 23) & 0xffu) - 126;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f000000u);  // [0.5, 1)
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = m + m - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float y = ((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m +
                   1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m +
               3.3333331174e-1f) * m * z;
    const float fe = (float)e;
    y += -2.12194440e-4f * fe;
    y += -0.5f * z;
    return (m + y) + 0.693359375f * fe;
}

// exp: n = round(x / ln 2), x - n ln 2 in two parts, a polynomial, then 2^n
// as two exact power-of-two factors.  Results below 1.6e-38 (x < -87) are
// returned as 0, so no result is subnormal; x > 88.72 gives inf.
__device__ __forceinline__ float exp_(float x) {
    if (x != x) return x;
    if (x > 88.7228391f) return __int_as_float(0x7f800000);
    if (x < -87.0f) return 0.0f;
    const float fz = floorf(1.44269504088896341f * x + 0.5f);
    float r = x - fz * 0.693359375f;
    r = r - fz * -2.12194440e-4f;
    const int n = (int)fz;
    const float z = r * r;
    const float p = ((((((1.9875691500e-4f * r + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r + 4.1665795894e-2f) * r +
                      1.6666665459e-1f) * r + 5.0000001201e-1f) * z + r) + 1.0f;
    const int n1 = n / 2, n2 = n - n1;
    return (p * __uint_as_float((uint32_t)(n1 + 127) << 23)) * __uint_as_float((uint32_t)(n2 + 127) << 23);
}

// x^y for x > 0 as exp(y log x)
__device__ __forceinline__ float pow_(float x, float y) { return exp_(y * log_(x)); }

}  // namespace st
