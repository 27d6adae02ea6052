/*
 * pt_oracle.h — CPU ORACLE (test infrastructure only).
 *
 * A plain-C restatement of Stratum2's TestRenderer path tracer
 * (src/Shaders/kernels/testrenderer.slang and the headers it includes) used
 * as the parity checker of the HIP implementation.  Only tests, the smoke
 * entry and the bench's CPU-baseline leg may load it; the product never does.
 *
 * It consumes the same scene buffers as the C ABI (include/stratum2.h), with
 * its own triangle flattening, its own median-split BVH and its own
 * watertight triangle test, so that it shares no code with the product.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include "../include/stratum2.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* Per-pixel/per-sample debug streams filled by or_render (NULL to skip). */
typedef struct or_outputs {
    float* hdr;              /* RGBA32F W*H, mean over spp */
    uint32_t* visibility;    /* RG32UI W*H (sample 0) */
    uint32_t* path_length;   /* W*H*spp */
    uint64_t* ray_counts;    /* [2]: extension, shadow */
    uint64_t* traversal;     /* [4]: nodes visited (closest), tris tested (closest), nodes (shadow), tris (shadow) */
    float* albedo;           /* RGBA32F W*H (sample 0), as st_outputs */
    float* depth;            /* RGBA32F W*H (sample 0), as st_outputs */
    float* prev_uv;          /* RG32F W*H (sample 0), as st_outputs */
} or_outputs;

or_scene* or_scene_create(const st_scene_desc* desc);
void or_scene_destroy(or_scene* s);
const char* or_last_error(void);

/* Render like st_render.  threads: OpenMP threads (0 = all). */
int or_render(or_scene* s, const st_view* view, const st_render_params* p, const or_outputs* out,
              int threads);

/* st_render_cached without the network: as or_render, while a path reaching a
 * scattering vertex at path length depth ends there in a cache query (adding
 * nothing here) unless it is the training path of its run of stride paths
 * (offset pcg(run ^ pcg(frame)) % stride).  recs (may be NULL): 25 floats per
 * path (index sample * W * H + y * W + x): kind (0 none, 1 cache query,
 * 2 training row), query[12], throughput[3], the radiance gathered up to and
 * including the vertex's emission (kind 2)[3], the path's fp32 part[3] and
 * its fixed-point part / 16384 [3]. */
int or_render_cached(or_scene* s, const st_view* view, const st_render_params* p, uint32_t depth, uint32_t stride,
                     uint32_t frame, const or_outputs* out, float* recs, int threads);

/* Closest-hit / any-hit of raw rays, same layout as st_trace_rays. */
int or_trace_rays(or_scene* s, const float* rays, uint32_t count, int closest, uint32_t* hits);

/* Brute force closest hit in FP64 over every triangle (independent check of
 * the watertight test + BVH on tiny scenes). hits: {t (float bits), id, 0, 0}. */
int or_trace_rays_bruteforce_f64(or_scene* s, const float* rays, uint32_t count, uint32_t* hits);

/* Differentiable rendering.  The per-triangle albedo (3 floats per flattened
 * triangle: mesh instances in order, primitives in order; NULL clears)
 * replaces the base colour of Lambertian camera-path vertices.  A weight grid
 * (same size as volume `vol`; NULL clears) feeds the path weights while the
 * tracking decisions keep reading the grid: fixed-seed differences over it
 * are the detached-sampling derivative.  or_adjoint: the path-replay
 * gradient of sum(d_hdr * hdr) (see pt_oracle.c), added to grad_alb (3 per
 * triangle) and grad_grid (one per 4-byte word of grid grad_vol). */
uint32_t or_scene_triangle_count(const or_scene* s);
int or_scene_set_triangle_albedo(or_scene* s, const float* rgb, uint32_t count);
int or_scene_set_weight_grid(or_scene* s, uint32_t vol, const void* grid, uint64_t size);
int or_adjoint(or_scene* s, const st_view* view, const st_render_params* p, const float* d_hdr, double* grad_alb,
               double* grad_grid, uint32_t grad_vol, int threads);

/* Bit-exact helper exports (for golden vectors). */
void or_pcg4d(const uint32_t in[4], uint32_t out[4]);
uint32_t or_xxhash32(uint32_t p);
uint32_t or_pcg(uint32_t v);
uint32_t or_pack_normal(const float v[3]);
void or_unpack_normal(uint32_t p, float out[3]);
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
void or_ray_offset(const float p[3], const float n[3], float out[3]);

/* Environment sampling restated (lights.hlsli:49-163): returns uv and pdf. */
int or_env_sample_texel(or_scene* s, const float rnd[2], float uv[2], float* pdf);
float or_env_sample_texel_pdf(or_scene* s, const float uv[2]);
/* Alias-table sampling of the environment (ST_FLAG_ENV_ALIAS), n samples:
 * rnd[4i..4i+3] = {column, accept test, u jitter, v jitter} -> uv[2i..], pdf[i];
 * the table itself (prob, alias, uv pdf per texel of level 0; any pointer may
 * be null). */
int or_env_alias_sample(or_scene* s, uint32_t n, const float* rnd, float* uv, float* pdf);
float or_env_alias_pdf(or_scene* s, const float uv[2]);
/* SampleImage4 (scene.hlsli:101-109) of one RGBA32F image with its mip chain,
 * at uv with texture-coordinate screen size uvs (0: level 0). */
int or_tex_sample(const float* rgba, uint32_t w, uint32_t h, float u, float v, float uvs, float out[4]);
int or_env_alias_table(or_scene* s, float* prob, uint32_t* alias, float* pdf);

/* BSDF restated: evaluateReflectance / sampleDirection (disney.hlsli:391-429,
 * lambertian.hlsli:13-39). mat = 4 packed words + emission scale bits.
 * out_eval: {f.rgb, fwdPdf, revPdf}; out_sample: {dir.xyz, f.rgb, fwdPdf, revPdf, eta, roughness}. */
void or_bsdf_eval(const uint32_t mat[5], int lambertian, const float din[3], const float dout[3],
                  float out_eval[5]);
/* evaluateReflectance<Adjoint> (disney.hlsli:391-403): adjoint = 1 for light paths */
void or_bsdf_eval_adjoint(const uint32_t mat[5], int lambertian, int adjoint, const float din[3], const float dout[3],
                          float o[5]);
void or_bsdf_sample(const uint32_t mat[5], int lambertian, const float rnd[3], const float din[3],
                    float out_sample[10]);

/* NanoVDB 32.3 float grid read (PNanoVDB.h read accessor path, as
 * medium.hlsli:85-88 uses it) and the grid's world->index map
 * (pnanovdb_grid_world_to_index[_dir]f). */
float or_nvdb_read(const void* grid, int i, int j, int k);
void or_nvdb_world_to_index(const void* grid, const float p[3], int dir, float out[3]);

/* The deterministic transcendentals (the same float operations as the GPU
 * library) over n arguments: fn 0 sin(x), 1 cos(x), 2 atan2(x, y) (x the
 * numerator, y != 0), 3 acos(x), 4 log(x), 5 exp(x), 6 pow(x, y) for x > 0;
 * y may be null for the one-argument functions. */
void or_math(int fn, uint32_t n, const float* x, const float* y, float* out);

/* Henyey-Greenstein phase function, medium.hlsli:44-78. */
float or_hg_eval(float g, const float din[3], const float dout[3]);
void or_hg_sample(float g, float r0, float r1, const float din[3], float out[3]);

/* Tonemapper restated (tonemap.slang:30-192): like st_tonemap on host
 * arrays; max_out (may be null) receives the reduced maxima {r, g, b, lum}. */
int or_tonemap(const float* hdr, const float* albedo, uint32_t w, uint32_t h, float exposure, int mode, int gamma, float* out,
               float max_out[4]);

/* SVGF denoiser kernels restated (svgf_oracle.c): one call per kernel over
 * host images, as st_denoise runs them (Denoiser::denoise restated in
 * oracle.py). */
float or_svgf_normal_threshold(void);
void or_svgf_accumulate(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* map,
                        uint32_t map_count, const uint32_t* vis, const uint32_t* prev_vis, const float* depth,
                        const float* prev_depth, const float* prev_uv, const float* input, const float* albedo,
                        const float* prev_color, const float* prev_moments, float* color, float* moments);
void or_svgf_variance(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* map,
                      uint32_t map_count, const uint32_t* vis, const float* depth, const float* color,
                      const float* moments, float* out);
void or_svgf_atrous(const st_view* view, const st_denoise_params* p, uint32_t W, uint32_t H, const uint32_t* vis,
                    const float* depth, uint32_t iteration, const float* in, float* out);
void or_svgf_copy_rgb(uint32_t W, uint32_t H, const float* filtered, float* color);

#ifdef __cplusplus
}
#endif
#endif
