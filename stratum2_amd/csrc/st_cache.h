// st_cache.h — learned radiance cache: a multiresolution hash-grid encoding
// (fp16 table) and a fully-fused 64-wide ReLU MLP on fp16 MFMA with fp32
// accumulation, with hand-written inference, backward and Adam kernels.
//
// Inputs per query (12 floats): world position, outgoing direction, shading
// normal, albedo.  Network input (64 fp16): 2 features x L levels of the
// hash grid (trilinear over each level's cell), (direction + 1) / 2,
// (normal + 1) / 2, albedo, a constant 1 (the first layer's bias), zeros.
// Layers: 64 -> 64 -> 64 -> 64 (ReLU) -> 3 (linear; the output matrix is
// stored 64 x 64 with rows 3..63 zero).  No other biases.
//
// Matrix layout on v_mfma_f32_16x16x16f16 (one wavefront, 16 x 16 x 16):
// lane l holds A[l % 16][4 (l / 16) + j], B[4 (l / 16) + j][l % 16] and
// D[4 (l / 16) + j][l % 16], j = 0..3.  With A = a 16 x 16 tile of W[out][in]
// and B = a tile of X^T (k = input feature, column = sample), each lane's A
// and B are 4 consecutive halves of a row of W and of X, and D is the tile of
// Y^T whose lane holds 4 consecutive output features of one sample: the
// B operand of the next layer's matching k block, and 4 consecutive halves
// of a row of Y.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace st {

constexpr int kMlpWidth = 64;      // input, hidden and (padded) output width
constexpr int kMlpLayers = 4;      // 3 hidden layers + output
constexpr int kCacheMaxLevels = 16;
constexpr int kCacheQuery = 12;    // floats per query

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

struct CacheGrid {
    float lo[3], inv_ext[3];            // position normalisation: (p - lo) * inv_ext in [0, 1]
    float res[kCacheMaxLevels];         // cells per axis of each level
    uint32_t levels, table_mask;        // table entries per level = table_mask + 1 (a power of 2)
    uint32_t dense;                     // bit l: level l's (res + 2)^3 corners fit the table, indexed directly
};

// pcg (rng.hlsli:17-21)
__device__ __forceinline__ uint32_t rc_pcg(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

// The 8 corners of a query's cell at one level: table entries and trilinear
// weights.  A level whose (res + 2)^3 corner points fit its table indexes
// them directly (x + n (y + n z), n = res + 2); a finer level hashes them
// with the spatial hash of Mueller et al., "Instant Neural Graphics
// Primitives", 2022 (x xor 2654435761 y xor 805459861 z), which keeps the
// x and x + 1 corners in neighbouring entries, mostly one cache line.
// res: the levels' resolutions in LDS (rc_load_res; a lane-dependent index
// into the kernel argument would put the argument in scratch)
__device__ __forceinline__ void rc_corners(const CacheGrid& g, const float* res, const float* q, uint32_t l, uint32_t idx[8],
                                           float wt[8]) {
    float p[3];
    int c[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
        float u = (q[a] - g.lo[a]) * g.inv_ext[a];
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        p[a] = u * res[l];
        const float fl = floorf(p[a]);
        c[a] = (int)fl;
        f[a] = p[a] - fl;
    }
    const bool dense = (g.dense >> l) & 1u;
    const uint32_t n = (uint32_t)res[l] + 2;
    for (int k = 0; k < 8; k++) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const uint32_t x = (uint32_t)(c[0] + dx), y = (uint32_t)(c[1] + dy), z = (uint32_t)(c[2] + dz);
        const uint32_t h = dense ? x + n * (y + n * z) : (x ^ (y * 2654435761u) ^ (z * 805459861u)) & g.table_mask;
        idx[k] = l * (g.table_mask + 1) + h;
        wt[k] = (dx ? f[0] : 1 - f[0]) * (dy ? f[1] : 1 - f[1]) * (dz ? f[2] : 1 - f[2]);
    }
}

// The two hash features of level l: trilinear over the level's cell
__device__ __forceinline__ void rc_level(const CacheGrid& g, const float* res, const _Float16* table, const float* q, uint32_t l,
                                         _Float16& fa, _Float16& fb) {
    uint32_t idx[8];
    float wt[8];
    rc_corners(g, res, q, l, idx, wt);
    h2 t[8];
    for (int k = 0; k < 8; k++) t[k] = *reinterpret_cast<const h2*>(table + 2 * (size_t)idx[k]);  // one entry: both features
    float a = 0, b = 0;
    for (int k = 0; k < 8; k++) {
        a += wt[k] * (float)t[k][0];
        b += wt[k] * (float)t[k][1];
    }
    fa = (_Float16)a;
    fb = (_Float16)b;
}
// Raw input i after the hash features: (direction + 1) / 2, (normal + 1) / 2,
// albedo, a constant 1, then zeros
__device__ __forceinline__ _Float16 rc_raw(const float* q, uint32_t i) {
    // selects over constant indices: q stays in registers
    float v = i == 9 ? 1.0f : 0.0f;
    for (int a = 0; a < 6; a++) v = i == (uint32_t)a ? (q[3 + a] + 1) * 0.5f : v;
    for (int a = 6; a < 9; a++) v = i == (uint32_t)a ? q[3 + a] : v;
    return (_Float16)v;
}
// Network inputs f and f + 1 (f even) of query q
__device__ __forceinline__ void rc_feature_pair(const CacheGrid& g, const float* res, const _Float16* table, const float* q,
                                                uint32_t f, _Float16& a, _Float16& b) {
    if (f / 2 < g.levels) {
        rc_level(g, res, table, q, f / 2, a, b);
    } else {
        a = rc_raw(q, f - 2 * g.levels);
        b = rc_raw(q, f + 1 - 2 * g.levels);
    }
}

// The levels' resolutions into LDS (constant indices into the argument);
// every thread of the block must call it
__device__ __forceinline__ void rc_load_res(const CacheGrid& g, float* res) {
    if (threadIdx.x == 0)
        for (int l = 0; l < kCacheMaxLevels; l++) res[l] = g.res[l];
    __syncthreads();
}

// Rows a training kernel covers: n, or fewer when the count is on the device
__device__ __forceinline__ uint32_t rc_rows(uint32_t n, const uint32_t* n_dev) { return n_dev ? min(*n_dev, n) : n; }

// X[s] (64 fp16) from query s: hash features, then the raw inputs
__global__ void k_rc_encode(CacheGrid g, const _Float16* table, const float* queries, uint32_t n_max, const uint32_t* n_dev,
                            _Float16* X) {
    __shared__ float res[kCacheMaxLevels];
    rc_load_res(g, res);
    const uint32_t n = rc_rows(n_max, n_dev);
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const float* q = queries + (size_t)kCacheQuery * s;
        _Float16* x = X + (size_t)kMlpWidth * s;
        for (uint32_t f = 0; f < (uint32_t)kMlpWidth; f += 2) rc_feature_pair(g, res, table, q, f, x[f], x[f + 1]);
    }
}

// Inference for the paths of a wave that ended in the cache, fused into one
// kernel: each wavefront takes 16 queries at a time, encodes them straight
// into the first layer's B operands (lane l: features 16 kb + 4 (l / 16) + j
// of query l % 16, so the 4 lanes of a query read 4 levels each) and runs
// the four layers in registers, each layer's D tile being the next layer's B
// operand; the weights sit in LDS.  The arithmetic is that of k_rc_encode +
// k_mlp_layer + k_rc_output, so a query gets the radiance st_cache_infer
// gives it.  em[path] += throughput x radiance.  cq: 4 float4 per query
// ({pos, thr.x}, {wo, thr.y}, {normal, thr.z}, {albedo, path}); count:
// queries (device).
__global__ void __launch_bounds__(256) k_rc_infer_paths(CacheGrid g, const _Float16* table, const _Float16* W16, const float4* cq,
                                                         const uint32_t* count, float4* em) {
    constexpr int kH4 = kMlpLayers * kMlpWidth * kMlpWidth / 4;
    __shared__ h4 Wl[kH4];
    __shared__ float res[kCacheMaxLevels];
    const h4* Wg = reinterpret_cast<const h4*>(W16);
    for (int i = threadIdx.x; i < kH4; i += blockDim.x) Wl[i] = Wg[i];
    rc_load_res(g, res);
    const uint32_t n = *count;
    const uint32_t lane = __lane_id(), r = lane % 16, qd = lane / 16;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, waves = gridDim.x * blockDim.x / 64;
    for (uint32_t base = wave * 16; base < n; base += waves * 16) {  // wave-uniform bound: every lane runs every MFMA
        const uint32_t s = base + r;
        const bool in = s < n;
        float q[kCacheQuery];
        float4 c0 = make_float4(0, 0, 0, 0), c1 = c0, c2 = c0, c3 = c0;
        if (in) {
            c0 = cq[4 * (size_t)s];
            c1 = cq[4 * (size_t)s + 1];
            c2 = cq[4 * (size_t)s + 2];
            c3 = cq[4 * (size_t)s + 3];
        }
        q[0] = c0.x, q[1] = c0.y, q[2] = c0.z, q[3] = c1.x, q[4] = c1.y, q[5] = c1.z;
        q[6] = c2.x, q[7] = c2.y, q[8] = c2.z, q[9] = c3.x, q[10] = c3.y, q[11] = c3.z;
        h4 B[4];
        for (int kb = 0; kb < 4; kb++) {
            const uint32_t f = kb * 16 + 4 * qd;
            _Float16 a, b, c, d;
            rc_feature_pair(g, res, table, q, f, a, b);
            rc_feature_pair(g, res, table, q, f + 2, c, d);
            B[kb] = in ? h4{a, b, c, d} : h4{0, 0, 0, 0};
        }
        f4v acc[4];
        for (int l = 0; l < kMlpLayers; l++) {
            const h4* A = Wl + l * (kMlpWidth * kMlpWidth / 4);
            for (int ob = 0; ob < 4; ob++) {
                acc[ob] = f4v{0, 0, 0, 0};
                for (int kb = 0; kb < 4; kb++)
                    acc[ob] = __builtin_amdgcn_mfma_f32_16x16x16f16(A[(ob * 16 + r) * 16 + kb * 4 + qd], B[kb], acc[ob], 0, 0, 0);
            }
            if (l < kMlpLayers - 1)
                for (int ob = 0; ob < 4; ob++)
                    for (int j = 0; j < 4; j++) B[ob][j] = (_Float16)fmaxf(acc[ob][j], 0.0f);
        }
        if (in && qd == 0) {  // lanes 0..15 hold output features 0..3 of their query
            const uint32_t p = __float_as_uint(c3.w);
            float4 e = em[p];
            e.x += c0.w * fmaxf((float)(_Float16)acc[0][0], 0.0f);
            e.y += c1.w * fmaxf((float)(_Float16)acc[0][1], 0.0f);
            e.z += c2.w * fmaxf((float)(_Float16)acc[0][2], 0.0f);
            em[p] = e;
        }
    }
}

// dTable += sum over queries of weight * dX[s][2 l + f].  Device-scope float
// atomics: 16 per sample and level, about 12 G/s here whatever the order of
// the samples (screen order, scattered, or summed across a wavefront first
// where lanes share a corner all measured the same).
__global__ void k_rc_encode_bwd(CacheGrid g, const float* queries, uint32_t n_max, const uint32_t* n_dev, const _Float16* dX,
                                float* dtable) {
    __shared__ float res[kCacheMaxLevels];
    rc_load_res(g, res);
    const uint32_t n = rc_rows(n_max, n_dev);
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const float* q = queries + (size_t)kCacheQuery * s;
        const _Float16* d = dX + (size_t)kMlpWidth * s;
        for (uint32_t l = 0; l < g.levels; l++) {
            const float ga = (float)d[2 * l], gb = (float)d[2 * l + 1];
            if (ga == 0 && gb == 0) continue;
            uint32_t idx[8];
            float wt[8];
            rc_corners(g, res, q, l, idx, wt);
            for (int k = 0; k < 8; k++) {
                atomicAdd(dtable + 2 * (size_t)idx[k], wt[k] * ga);
                atomicAdd(dtable + 2 * (size_t)idx[k] + 1, wt[k] * gb);
            }
        }
    }
}

// Y[s][o] = act(sum_k X[s][k] W[o][k]) over n rows, one wavefront per 64
// rows at a time: act = ReLU (relu), or a mask (mask[s][o] > 0 keeps the
// value: the backward pass through a ReLU), or none.  W: 64 x 64 fp16,
// row-major [out][in].  n need not be a multiple of 64 (rows past n are
// neither read nor written).
__global__ void __launch_bounds__(256) k_mlp_layer(const _Float16* W, const _Float16* X, _Float16* Y, const _Float16* mask,
                                                    uint32_t n_max, const uint32_t* n_dev, int relu) {
    const uint32_t n = rc_rows(n_max, n_dev);
    const uint32_t lane = __lane_id(), r = lane % 16, qd = lane / 16;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, waves = gridDim.x * blockDim.x / 64;
    h4 A[4][4];
    for (int ob = 0; ob < 4; ob++)
        for (int kb = 0; kb < 4; kb++) A[ob][kb] = *reinterpret_cast<const h4*>(W + (ob * 16 + r) * kMlpWidth + kb * 16 + 4 * qd);
    for (uint32_t base = wave * 64; base < n; base += waves * 64) {
        for (int sb = 0; sb < 4; sb++) {
            const uint32_t s = base + sb * 16 + r;
            const bool in = s < n;
            h4 B[4];
            for (int kb = 0; kb < 4; kb++) {
                if (in) B[kb] = *reinterpret_cast<const h4*>(X + (size_t)s * kMlpWidth + kb * 16 + 4 * qd);
                else B[kb] = h4{0, 0, 0, 0};
            }
            f4v acc[4];
            for (int ob = 0; ob < 4; ob++) {  // every lane runs every MFMA
                acc[ob] = f4v{0, 0, 0, 0};
                for (int kb = 0; kb < 4; kb++) acc[ob] = __builtin_amdgcn_mfma_f32_16x16x16f16(A[ob][kb], B[kb], acc[ob], 0, 0, 0);
            }
            if (in) {
                for (int ob = 0; ob < 4; ob++) {
                    const size_t off = (size_t)s * kMlpWidth + ob * 16 + 4 * qd;
                    h4 m4 = h4{1, 1, 1, 1};
                    if (mask) m4 = *reinterpret_cast<const h4*>(mask + off);
                    h4 y;
                    for (int j = 0; j < 4; j++) {
                        float v = acc[ob][j];
                        if (relu) v = fmaxf(v, 0.0f);
                        if (mask && !((float)m4[j] > 0.0f)) v = 0.0f;
                        y[j] = (_Float16)v;
                    }
                    *reinterpret_cast<h4*>(Y + off) = y;
                }
            }
        }
    }
}

// partial[w][o][i] = sum over wavefront w's chunk of `chunk` rows (a
// multiple of 16) of dY[s][o] X[s][i], all 16 output tiles in registers;
// k_mlp_wgrad_reduce adds the partials up.
__global__ void __launch_bounds__(256) k_mlp_wgrad(const _Float16* dY, const _Float16* X, float* partial, uint32_t n_max,
                                                    const uint32_t* n_dev, uint32_t chunk) {
    const uint32_t lane = __lane_id(), r = lane % 16, qd = lane / 16;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t n = rc_rows(n_max, n_dev);
    const uint32_t s0 = wave * chunk;
    if (s0 >= n) {  // whole wavefronts leave together; past a device count their partials are 0
        if (s0 < n_max)
            for (int e = lane; e < kMlpWidth * kMlpWidth; e += 64) partial[(size_t)wave * kMlpWidth * kMlpWidth + e] = 0.0f;
        return;
    }
    const uint32_t s1 = min(n, s0 + chunk);
    f4v acc[4][4];
    for (int ob = 0; ob < 4; ob++)
        for (int ib = 0; ib < 4; ib++) acc[ob][ib] = f4v{0, 0, 0, 0};
    for (uint32_t k0 = s0; k0 < s1; k0 += 16) {
        // A[o][s] = dY[s][o]: lane holds rows s = k0 + 4 qd + j of column o = ob 16 + r
        // B[s][i] = X[s][i]: lane holds rows s = k0 + 4 qd + j of column i = ib 16 + r
        h4 A[4], B[4];
        for (int t = 0; t < 4; t++)
            for (int j = 0; j < 4; j++) {
                const uint32_t s = k0 + 4 * qd + j;
                A[t][j] = s < s1 ? dY[(size_t)s * kMlpWidth + t * 16 + r] : (_Float16)0.0f;
                B[t][j] = s < s1 ? X[(size_t)s * kMlpWidth + t * 16 + r] : (_Float16)0.0f;
            }
        for (int ob = 0; ob < 4; ob++)
            for (int ib = 0; ib < 4; ib++) acc[ob][ib] = __builtin_amdgcn_mfma_f32_16x16x16f16(A[ob], B[ib], acc[ob][ib], 0, 0, 0);
    }
    // D[o][i]: lane holds o = ob 16 + 4 qd + j, i = ib 16 + r
    float* out = partial + (size_t)wave * kMlpWidth * kMlpWidth;
    for (int ob = 0; ob < 4; ob++)
        for (int ib = 0; ib < 4; ib++)
            for (int j = 0; j < 4; j++) out[(ob * 16 + 4 * qd + j) * kMlpWidth + ib * 16 + r] = acc[ob][ib][j];
}

// dW[e] += scale * the sum of the nw partials of element e: block (x, y)
// sums partials y, y + gridDim.y, ... of 256 consecutive elements, and the
// gridDim.y sums of an element meet in one atomic add each
__global__ void k_mlp_wgrad_reduce(const float* partial, uint32_t nw, float* dW, float scale) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint32_t)(kMlpWidth * kMlpWidth)) return;
    float sum = 0;
    for (uint32_t w = blockIdx.y; w < nw; w += gridDim.y) sum += partial[(size_t)w * kMlpWidth * kMlpWidth + e];
    atomicAdd(dW + e, scale * sum);
}

// Relative L2 loss of the radiance cache (Mueller et al., "Real-time Neural
// Radiance Caching for Path Tracing", 2021): per sample and channel
// (pred - target)^2 / (pred^2 + 0.01) / 3, with pred held fixed in the
// denominator.  dY[s][c] is the per-sample gradient (the 1/n goes into the
// weights' gradient scale).  targets: tstride floats per sample, rgb and,
// with tstride 4, a weight (0: the sample is left out).  loss_sum[0]
// receives the weighted loss sum, loss_sum[1] the weight sum.
__global__ void k_rc_loss(const _Float16* Y, const float* targets, uint32_t tstride, uint32_t n_max, const uint32_t* n_dev,
                          _Float16* dY, float* loss_sum) {
    const uint32_t n = rc_rows(n_max, n_dev);
    float local = 0, wsum = 0;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        _Float16* d = dY + (size_t)kMlpWidth * s;
        for (int c = 0; c < kMlpWidth; c++) d[c] = (_Float16)0.0f;
        const float* t = targets + (size_t)tstride * s;
        const float wgt = tstride == 4 ? t[3] : 1.0f;
        if (!(wgt > 0)) continue;
        wsum += wgt;
        for (int c = 0; c < 3; c++) {
            const float p = (float)Y[(size_t)kMlpWidth * s + c];
            const float den = p * p + 0.01f;
            local += wgt * ((p - t[c]) * (p - t[c]) / den / 3);
            d[c] = (_Float16)fminf(fmaxf(wgt * (2 * (p - t[c]) / den / 3), -30000.0f), 30000.0f);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        local += __shfl_xor(local, off);
        wsum += __shfl_xor(wsum, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(loss_sum, local);
        atomicAdd(loss_sum + 1, wsum);
    }
}

// Adam (Kingma and Ba 2015) over n parameters with fp32 master copies; g is
// scaled by gscale; entries whose gradient is exactly zero keep their
// moments and value (sparse: table entries no sample touched).  p16: the
// fp16 copy the training kernels read; p16t (optional): its 64 x 64 blocks
// transposed, for the backward pass.  ema (optional): the exponential moving
// average of the parameters, ema = decay ema + (1 - decay) p, whose fp16 copy
// ema16 inference reads (Polyak averaging, as Mueller et al. 2021 infer with).
__global__ void k_rc_adam(float* p, float* m, float* v, const float* g, uint32_t n, float gscale, float lr, float b1, float b2,
                          float eps, float c1, float c2, _Float16* p16, _Float16* p16t, int sparse, float* ema, _Float16* ema16,
                          float decay) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i] * gscale;
        if (!(sparse && gi == 0.0f)) {
            const float mi = b1 * m[i] + (1 - b1) * gi;
            const float vi = b2 * v[i] + (1 - b2) * gi * gi;
            m[i] = mi;
            v[i] = vi;
            p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        }
        p16[i] = (_Float16)p[i];
        if (p16t) {
            const uint32_t blk = i / (kMlpWidth * kMlpWidth), e = i % (kMlpWidth * kMlpWidth);
            const uint32_t o = e / kMlpWidth, in = e % kMlpWidth;
            p16t[(size_t)blk * kMlpWidth * kMlpWidth + in * kMlpWidth + o] = (_Float16)p[i];
        }
        if (ema) {
            const float e = decay * ema[i] + (1 - decay) * p[i];
            ema[i] = e;
            ema16[i] = (_Float16)e;
        }
    }
}

// fp16 copies of fp32 parameters (and, optionally, the 64 x 64 blocks transposed)
__global__ void k_rc_copy16(const float* p, uint32_t n, _Float16* p16, _Float16* p16t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        p16[i] = (_Float16)p[i];
        if (p16t) {
            const uint32_t blk = i / (kMlpWidth * kMlpWidth), e = i % (kMlpWidth * kMlpWidth);
            p16t[(size_t)blk * kMlpWidth * kMlpWidth + (e % kMlpWidth) * kMlpWidth + e / kMlpWidth] = (_Float16)p[i];
        }
    }
}

// radiance = max(0, Y[s][0..2]) as fp32
__global__ void k_rc_output(const _Float16* Y, uint32_t n, float* out) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) out[3 * (size_t)s + c] = fmaxf((float)Y[(size_t)kMlpWidth * s + c], 0.0f);
}

}  // namespace st
