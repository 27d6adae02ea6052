/*
 * stratum2.h — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * This is the drop-in boundary for Stratum2's TestRenderer: the scene buffer
 * formats, the view, the push-constant block and the HDR output buffer are
 * the reference's own; the Vulkan objects around them are replaced by plain
 * pointers and sizes.  No C++ or torch types appear in any signature.
 *
 * Conventions
 *   - every call returns an int status (ST_OK == 0); st_last_error() returns a
 *     thread-local message for the last failing call on this thread;
 *   - one HIP stream per context; calls on one context must be serialised by
 *     the caller (the reference renderer is not re-entrant either,
 *     src/App/TestRenderer.cpp:284-649);
 *   - host arrays in st_scene_desc are copied, the caller keeps ownership;
 *   - output pointers in st_outputs are DEVICE pointers (hipMalloc'd or torch
 *     tensors) owned by the caller.
 */
#ifndef STRATUM2_H
#define STRATUM2_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_OK 0
#define ST_ERROR_INVALID_ARGUMENT 1
#define ST_ERROR_UNSUPPORTED 2
#define ST_ERROR_HIP 3
#define ST_ERROR_OUT_OF_MEMORY 4
#define ST_ERROR_LIMIT 5

#define ST_INVALID_INSTANCE 0xFFFFu  /* src/Shaders/compat/scene.h:17 */
#define ST_INVALID_PRIMITIVE 0xFFFFu /* src/Shaders/compat/scene.h:18 */
#define ST_NO_IMAGE 0xFFFFFFFFu      /* any index >= 2048 means "no image" (scene.h:14) */
#define ST_IMAGE_LIMIT 2048u         /* gImageCount, src/Shaders/compat/scene.h:14 */

/* InstanceData, src/Shaders/compat/scene.h:30-80.
 * type in bits 0..3 (0 mesh, 1 sphere, 2 volume), material byte address in 4..31.
 * mesh:   data = vertexInfoIndex | primitiveCount << 16
 * volume: data = volume index (0xFFFFFFFF = unit-cube homogeneous medium). */
typedef struct st_instance {
    uint32_t type_material_address;
    uint32_t data;
} st_instance;

/* TransformData: 3x4 row-major object-to-world, src/Shaders/compat/transform.h:7-24. */
typedef struct st_transform {
    float m[3][4];
} st_transform;

/* MeshVertexInfo, src/Shaders/compat/scene.h:89-132.
 * buffer indices: [0] = index | position << 16, [1] = normal | texcoord << 16
 * strides (bytes, 8 bits each): index | position << 8 | normal << 16 | texcoord << 24
 * offsets (bytes): index, position, normal, texcoord.  A buffer index >= 2048 means absent. */
typedef struct st_mesh_vertex_info {
    uint32_t packed_buffer_indices[2];
    uint32_t packed_strides;
    uint32_t pad;
    uint32_t packed_offsets[4];
} st_mesh_vertex_info;

/* VolumeInfo, src/Shaders/compat/scene.h:82-87 (object-space bounds of a medium instance). */
typedef struct st_volume_info {
    float min[3];
    uint32_t instance_index;
    float max[3];
    uint32_t pad;
} st_volume_info;

typedef struct st_buffer_view {
    const void* data;
    uint64_t size_bytes;
} st_buffer_view;

/* RGBA32F image, row 0 at the top (v = 0), the layout of an env map loaded by
 * src/Core/Image.cpp:57-73. */
typedef struct st_image_desc {
    const float* rgba;
    uint32_t width;
    uint32_t height;
} st_image_desc;

/* The scene buffers of Scene::FrameData (src/App/Scene.hpp:70-116), as
 * produced by Scene::updateFrameData (src/App/Scene.cpp:428-791). */
typedef struct st_scene_desc {
    uint32_t instance_count;
    const st_instance* instances;
    const st_transform* transforms;          /* object-to-world, one per instance */
    uint32_t material_word_count;
    const uint32_t* material_data;           /* Material::store (64 B), Medium::store (28 B),
                                                EnvironmentMap::store (16 B) records */
    uint32_t mesh_vertex_info_count;
    const st_mesh_vertex_info* mesh_vertex_info;
    uint32_t vertex_buffer_count;
    const st_buffer_view* vertex_buffers;    /* byte buffers (indices u16/u32, positions f32x3) */
    uint32_t image_count;
    const st_image_desc* images;             /* only the environment image may be referenced */
    uint32_t environment_material_address;   /* 0xFFFFFFFF = no environment */
    uint32_t volume_count;
    const st_buffer_view* volumes;           /* NanoVDB 32.3 float grids (bytes) */
    uint32_t volume_info_count;
    const st_volume_info* volume_info;
} st_scene_desc;

/* ProjectionData, src/Shaders/compat/transform.h:128-191. */
typedef struct st_projection {
    float scale[2];
    float offset[2];
    float near_plane;
    float far_plane;
    float sensor_area;
    float vertical_fov;
} st_projection;

/* ViewData (scene.h:134-167) + the camera node's TransformData (camera-to-world). */
typedef struct st_view {
    st_projection projection;
    int32_t image_min[2];
    int32_t image_max[2];
    st_transform camera_to_world;
} st_view;

/* Compile-time defines of TestRenderer (src/App/TestRenderer.cpp:19-36) become flags. */
#define ST_FLAG_LAMBERTIAN              (1u << 0)  /* gLambertian: all materials Lambert */
#define ST_FLAG_SAMPLE_DIRECT           (1u << 1)  /* gSampleDirectIllumination (default on) */
#define ST_FLAG_RESTIR_DI               (1u << 2)  /* gReSTIR_DI: RIS over candidate_samples */
#define ST_FLAG_RESTIR_DI_REUSE         (1u << 3)  /* screen-space temporal + spatial reuse of the
                                                       primary-vertex reservoirs (deeper vertices: plain NEE) */
#define ST_FLAG_ENV_ALIAS               (1u << 4)  /* alias-table env sampling (else mip warp) */
#define ST_FLAG_ENV_MIS_BALANCED        (1u << 5)  /* env hits weighted against env pdfW */
#define ST_FLAG_FLOAT_NEE               (1u << 6)  /* NEE added in fp32, not 1/16384 fixed point */
#define ST_FLAG_PIXEL_JITTER            (1u << 7)  /* gPixelJitter */
#define ST_FLAG_COUNT_TRAVERSAL         (1u << 8)  /* count nodes/triangles visited (gCountRays-like) */
#define ST_FLAG_DIRECT_ONLY             (1u << 9)  /* gSampleDirectIlluminationOnly: NEE, then stop */
/* TestRenderer light tracing ("Light tracing", mLightTrace, gTraceFromLight;
 * TestRenderer.cpp:106-125, 534-582): paths start on the lights and every
 * non-singular vertex connects to the camera; hdr is the atomic 1/16384
 * fixed-point image alone (ProcessAtomicOutput with gClearImage).  As the
 * reference forces, NEE, ReSTIR DI and the direct-only mode are off in this
 * mode (their flags are ignored), and visibility, albedo, depth, prev_uv and
 * path_length are not written. */
#define ST_FLAG_TRACE_FROM_LIGHT        (1u << 10)
/* TestRenderer's bidirectional mode with a light vertex cache (gUseVC;
 * TestRenderer.cpp:126-129, 478-482, 566-582; testrenderer.slang:325-440,
 * 483-497, 531-563, 788-806, 812-936, 958-1019).  Per sample, a light pass
 * traces max(1, W*H*light_subpath_factor) light subpaths that store their
 * non-singular vertices and connect them to the camera; then the camera paths
 * add emission, NEE and one connection to a uniformly chosen cached light
 * vertex at every non-singular vertex, all weighted with the VCM recursive MIS
 * quantities (no merging).  The cache is ordered by light subpath and depth
 * (the reference appends in atomic, run-dependent order; both pick uniformly).
 * Supported for surfaces with plain NEE: media, ReSTIR DI, --minPathLength,
 * fp32 NEE and the direct-only mode return ST_ERROR_UNSUPPORTED.  Waves hold
 * one sample (samples_per_wave is ignored); each of the up to four wave slots
 * keeps a cache of 96 B per light subpath and path segment (max_depth - 1).
 * ST_FLAG_TRACE_FROM_LIGHT takes precedence, as mLightTrace clears gUseVC. */
#define ST_FLAG_BDPT_LVC                (1u << 11)
/* With ST_FLAG_BDPT_LVC: keep dVC and dVCM in fp32.  The reference stores them
 * in fp16 (testrenderer.slang:75-82), where they overflow past 65504 and zero
 * the MIS weights of NEE and the connections: in a 64x64 Cornell box with one
 * light subpath per pixel, paths of three segments come out 8% too dark. */
#define ST_FLAG_VC_FP32_MIS             (1u << 12)
/* Run the waves one after another on the context stream instead of rotating
 * them through four wave slots on four streams.  The image is the same bit
 * for bit; no two kernels of the call overlap, so the per-kernel HIP-event
 * times of st_last_kernel_times are each kernel's own (profiling). */
#define ST_FLAG_SERIAL_WAVES            (1u << 13)
/* With ST_FLAG_RESTIR_DI: the reference's own reservoir reuse
 * (gReSTIR_DI_Reuse; SampleIlluminationReSTIR + ReservoirReuse,
 * testrenderer.slang:608-726; HashGrid, hashgrid.hlsli:14-120).  Every NEE
 * vertex runs RIS, then looks up the previous sample's world-space hash grid
 * at a point jittered in a concentric disc of one cell on its tangent plane,
 * picks one entry of that cell uniformly, combines it with pairwise MIS, and
 * appends its own reservoir (M clamped at max_m x candidates) to this
 * sample's grid.  Cells are hashgrid_cell_size wide (0: the reference's 0.5).
 * The grid of a sample keeps the first two reservoir vertices of each path
 * (the reference's capacity, W x H x min(2, maxDepth - 1)), and is built
 * deterministically: entries sorted by cell checksum, stable in pixel and
 * vertex order, looked up by binary search (the reference's atomic append
 * order, prefix sum and linear probing are run-dependent).  Samples run one
 * after another (one sample per wave, one stream).  As TestRenderer keeps
 * mPrevHashGrid from frame to frame (TestRenderer.cpp:472-493), a call's
 * first sample reuses the grid of the previous call's last sample when that
 * call, on the same context, was the last st_render, used this flag, and had
 * the same scene and W x H x min(2, maxDepth - 1); otherwise it has none.
 * Not with ST_FLAG_RESTIR_DI_REUSE or media. */
#define ST_FLAG_RESTIR_DI_HASHGRID      (1u << 14)
/* Drop the history kept from earlier calls on the context (the hash grid of
 * ST_FLAG_RESTIR_DI_HASHGRID), as at a camera cut. */
#define ST_FLAG_HISTORY_RESET           (1u << 15)
/* Per-material Lambert / GGX in place of the Disney BSDF: a material whose
 * metallic is at least 0.5 is a GGX conductor (the Disney BSDF's metallic lobe
 * alone: Schlick's Fresnel about the base colour, anisotropic GGX with
 * alpha = roughness^2, separable Smith masking, visible-normal sampling in
 * Heitz's view-aligned disk frame), any other is Lambert with its base colour.  The reference has no such
 * switch (gLambertian makes every material Lambert, testrenderer.slang:33-37);
 * ST_FLAG_LAMBERTIAN takes precedence. */
#define ST_FLAG_LAMBERT_GGX             (1u << 16)
/* Camera-path st_render only: return once the work is enqueued on the
 * context's streams, joined into the context stream, which waits for the
 * caller's null-stream work at entry.  The outputs (ray_counts included,
 * written on the device) are complete after st_sync(ctx); calls on one
 * context run in order.  No timings or statistics are collected for such a
 * call (st_last_timings and st_last_kernel_times keep the last synchronous
 * call's). */
#define ST_FLAG_ASYNC                   (1u << 17)
/* With ST_FLAG_RESTIR_DI_REUSE: RIS (no reuse) at the NEE vertices past the
 * primary one too, as the reference runs SampleIlluminationReSTIR at every
 * NEE vertex (testrenderer.slang:869-884); without it they take one-sample
 * NEE.  Not with the radiance cache. */
#define ST_FLAG_RESTIR_DI_DEEP          (1u << 18)
#define ST_FLAG_DEFAULT (ST_FLAG_SAMPLE_DIRECT)

/* PushConstants of testrenderer.slang:1-21 plus the CLI surface
 * (--maxPathLength, --minPathLength, --exposure) and the spp loop. */
typedef struct st_render_params {
    uint32_t width, height;            /* mOutputExtent */
    uint32_t max_depth;                /* mMaxDepth == --maxPathLength (segments) */
    uint32_t min_path_length;          /* --minPathLength: drop contributions of shorter paths */
    uint32_t seed;                     /* mRandomSeed of the first sample */
    uint32_t spp;                      /* samples: one launch each, seeds seed..seed+spp-1 */
    uint32_t max_null_collisions;      /* mMaxNullCollisions */
    float environment_sample_probability; /* mEnvironmentSampleProbability */
    uint32_t candidate_samples;        /* mCandidateSamples (ReSTIR DI) */
    float max_m;                       /* mMaxM */
    uint32_t flags;                    /* ST_FLAG_* */
    uint32_t rr_start_depth;           /* Russian roulette from this path length on (0: every bounce) */
    float exposure;                    /* --exposure: ldr = hdr * 2^exposure (only for st_tonemap) */
    uint32_t spatial_neighbours;       /* ReSTIR spatial reuse neighbours */
    float spatial_radius;              /* ReSTIR spatial reuse radius in pixels */
    uint32_t samples_per_wave;         /* 0 = automatic: samples traced concurrently */
    float light_subpath_factor;        /* mLightSubpathCount (TestRenderer.hpp:29, default 1): light paths
                                          per sample = max(1, uint(W*H * factor)) (TestRenderer.cpp:452) */
    float hashgrid_cell_size;          /* ST_FLAG_RESTIR_DI_HASHGRID: cell size in world units (0: 0.5,
                                          GpuHashGrid(..., 100000, 0.5f), TestRenderer.cpp:52) */
} st_render_params;

/* Device outputs.  hdr is required; the others may be NULL. */
typedef struct st_outputs {
    float* hdr;                 /* RGBA32F W*H: mean radiance over spp (mOutput) */
    uint32_t* visibility;       /* RG32UI W*H: {instance|primitive<<16, packed normal}, sample 0 */
    uint32_t* path_length;      /* u32 W*H*spp: segments traced by each sample's path */
    uint64_t* ray_counts;       /* u64[2]: {extension rays, shadow rays} added by this call */
    /* StoreAuxillaryData (testrenderer.slang:195-218) of sample 0's first vertex;
     * each may be null.  prev_uv and the previous depth assume a static scene
     * in steady state (identity motion transform, previous view = this view). */
    float* albedo;              /* RGBA32F W*H: {albedo, 1} (1 on a miss) (mAlbedo) */
    float* depth;               /* RGBA32F W*H: {fp16-stored distance, previous depth, 1, 1} (mDepth) */
    float* prev_uv;             /* RG32F W*H: previous-frame pixel coordinate / extent (mPrevUVs) */
} st_outputs;

typedef struct st_context st_context;
typedef struct st_scene st_scene;

const char* st_last_error(void);
int st_version(void);

int st_context_create(int device, st_context** out);
int st_context_destroy(st_context* ctx);
int st_sync(st_context* ctx);

/* Replaces Scene::updateFrameData + buildAccelerationStructure
 * (src/App/Scene.cpp:17-55, 428-791): copies the buffers, flattens mesh
 * instances and builds the SAH BVH on the host, uploads everything. */
int st_scene_create(st_context* ctx, const st_scene_desc* desc, st_scene** out);
int st_scene_destroy(st_scene* scene);

/* Moves the instances (instance_count host 3x4 transforms, the same
 * TransformData layout as st_scene_desc) and rebuilds the BVH on the GPU:
 * triangles re-flattened on the device, linear BVH (Morton order + radix
 * tree) collapsed to the 4-wide node format.  Replaces the reference's
 * per-update acceleration-structure rebuild (src/App/Scene.cpp:726-754).
 * build_ms (may be null) receives the device build time.  Hits do not depend
 * on the tree: ties resolve to the smaller instance|primitive id.  A scene
 * whose mesh materials carry alpha masks returns ST_ERROR_UNSUPPORTED. */
int st_scene_update(st_scene* scene, const st_transform* transforms, double* build_ms);
/* BVH statistics: node count, triangle count, max depth, bytes on device. */
int st_scene_stats(const st_scene* scene, uint64_t out[4]);

/* Replaces TestRenderer::render (src/App/TestRenderer.cpp:284-649):
 * Render, RenderIteration x (maxDepth-1), ProcessShadowRays, ProcessAtomicOutput
 * (testrenderer.slang:1087-1183) for every sample, averaged into outputs->hdr. */
int st_render(st_scene* scene, const st_view* view, const st_render_params* params,
              const st_outputs* outputs);

/* ---- Differentiable rendering (no reference counterpart: the reference has
 * no derivatives; this is the path-replay adjoint of st_render's camera-path
 * estimator, after Vicini et al., "Path Replay Backpropagation", 2021).
 *
 * st_scene_set_triangle_albedo: rgb = 3 floats per flattened triangle (mesh
 * instances in instance order, primitives in index order; count = the
 * triangle count of st_scene_stats), host or device memory, copied; the
 * values replace the material base colour at Lambertian camera-path vertices.
 * NULL removes the table.  While a table is set, st_render and batches need
 * ST_FLAG_LAMBERTIAN and plain NEE (no ReSTIR DI, light tracing or the
 * bidirectional mode): others return ST_ERROR_UNSUPPORTED.
 *
 * st_scene_set_volume_grid: replaces the values of grid `volume` (host or
 * device memory): the same byte size and tree layout as the grid the scene
 * was created with (only values may differ; the root's maximum value, byte
 * 36 of the root node, is the tracking majorant and must bound them).
 *
 * st_adjoint: the gradient of sum(d_hdr * hdr) over the image that st_render
 * with the same arguments and ST_FLAG_FLOAT_NEE renders.  d_hdr: RGBA float
 * per pixel (device).  grad_albedo (3 floats per flattened triangle; needs
 * the albedo table and ST_FLAG_LAMBERTIAN) and grad_grid (one float per
 * 4-byte word of grid grad_volume: the gradient of the value stored at byte
 * 4i is added to grad_grid[i]) are device buffers that the call adds to;
 * either may be NULL.  Each path is rendered, then replayed from its seed
 * (no tape): at a Lambertian vertex the radiance still to come is divided by
 * the albedo, at each delta-tracking or ratio-tracking collision it is
 * multiplied by the derivative of log(sigma_t) (real) or log(majorant -
 * sigma_t) (null) with respect to the grid value read.  Sampling decisions
 * (tracking events, roulette, MIS) are not differentiated, which keeps the
 * gradient unbiased.  Camera paths with plain NEE only (ReSTIR DI, light
 * tracing and the bidirectional mode return ST_ERROR_UNSUPPORTED). */
int st_scene_set_triangle_albedo(st_scene* scene, const float* rgb, uint32_t count);
int st_scene_set_volume_grid(st_scene* scene, uint32_t volume, const void* grid, uint64_t bytes);
int st_adjoint(st_scene* scene, const st_view* view, const st_render_params* params, const float* d_hdr,
               float* grad_albedo, float* grad_grid, uint32_t grad_volume);

/* ---- Learned radiance cache (no reference counterpart; after Mueller et
 * al., "Real-time Neural Radiance Caching for Path Tracing", 2021, with the
 * multiresolution hash encoding of "Instant Neural Graphics Primitives",
 * 2022).  An fp16 hash-grid table (levels x 2^log2_table entries x 2
 * features; a level whose (res + 2)^3 cell corners fit indexes them
 * directly, finer levels by that paper's spatial hash) and a 64-64-64-64-3
 * ReLU MLP on fp16 MFMA with fp32 accumulation, trained with Adam on a
 * relative L2 loss.
 * A query is 12 floats: world position, outgoing direction, shading normal,
 * albedo; positions are normalised by the box given at creation. */
typedef struct st_cache st_cache;
typedef struct st_cache_config {
    uint32_t levels;          /* 1-16 (16) */
    uint32_t log2_table;      /* entries per level, log2: 8-24 (19) */
    uint32_t base_resolution; /* cells per axis of the coarsest level (16) */
    uint32_t max_resolution;  /* of the finest level (2048) */
    float learning_rate;      /* Adam (1e-2) */
    float beta1, beta2, epsilon; /* (0.9, 0.99, 1e-15) */
    uint32_t seed;            /* initial parameters */
    float ema_decay;          /* inference reads the parameters' exponential moving average
                                 ema = ema_decay ema + (1 - ema_decay) p, updated at each step
                                 (0: inference reads the parameters themselves) */
} st_cache_config;
int st_cache_create(st_context* ctx, const st_cache_config* config, const float bbox_min[3], const float bbox_max[3],
                    st_cache** cache);
int st_cache_destroy(st_cache* cache);
/* queries: n x 12 floats, radiance: n x 3 floats (device); with the
 * parameters inference reads (the moving average when ema_decay > 0) */
int st_cache_infer(st_cache* cache, const float* queries, uint32_t n, float* radiance);
/* One Adam step on n samples (queries, n x 3 target radiances; device);
 * loss (host, may be null) receives the mean relative L2 loss before the step. */
int st_cache_train(st_cache* cache, const float* queries, const float* targets, uint32_t n, float* loss);
/* Parameter sizes in floats: the table (levels x entries x 2) and the four
 * 64 x 64 weight matrices ([layer][out][in]; the output layer's rows 3..63 are 0). */
int st_cache_param_counts(const st_cache* cache, uint32_t* table_floats, uint32_t* weight_floats);
/* Host copies of the fp32 parameters and of the last step's gradients (sums
 * over its samples of the per-sample loss gradient); any may be null. */
int st_cache_get_params(st_cache* cache, float* table, float* weights, float* table_grad, float* weight_grad);
int st_cache_set_params(st_cache* cache, const float* table, const float* weights);

/* Camera paths that end in the cache (no reference counterpart: st_render's
 * estimator with its paths cut short, Mueller et al. 2021).  A path reaching
 * a scattering vertex at path length cache_depth (2: bounce k is vertex k), with
 * cache_depth < max_depth, adds that vertex's emission as st_render does and
 * then, instead of NEE and a continuation, throughput x the cache's radiance
 * at the vertex (query: position, direction towards the previous vertex,
 * shading normal (0 at a medium vertex), albedo) and ends.  Training paths
 * are exempt: of each run of train_stride consecutive paths (path index =
 * sample * W * H + y * W + x, samples of this call counted from 0) the one at
 * offset pcg(run ^ pcg(frame)) % train_stride continues exactly as in
 * st_render, and its cache-depth vertex becomes training row `run`: the query
 * and target = (the path's final radiance - the radiance it had gathered up
 * to and including that vertex's emission) / its throughput there.  Rows of
 * training paths that end before the cache depth, or at a black vertex, have
 * weight 0.  After the render, train_steps Adam steps on the rows with
 * weight (all of them each step, moved to the front in row order;
 * train_steps 0 leaves the cache unchanged).  Inference in this call uses the
 * cache as it was on entry.
 * train_queries (rows x 12 floats) and train_targets (rows x 4: target rgb,
 * weight), rows = ceil(spp * W * H / train_stride), are optional device
 * outputs; loss (host, may be null) receives the mean loss of the first step.
 * ReSTIR DI is allowed only with ST_FLAG_RESTIR_DI_REUSE (RIS at primary
 * vertices); light tracing and the bidirectional mode are not. */
typedef struct st_cache_render_params {
    uint32_t cache_depth;     /* path length of the vertex that queries the cache (2) */
    uint32_t train_stride;    /* one training path per this many paths (>= 1) */
    uint32_t train_steps;     /* Adam steps after the render (1) */
    uint32_t frame;           /* selects this call's training paths */
    float* train_queries;     /* optional device output */
    float* train_targets;     /* optional device output */
    float* loss;              /* optional host output */
} st_cache_render_params;
int st_render_cached(st_scene* scene, st_cache* cache, const st_view* view, const st_render_params* params,
                     const st_cache_render_params* cache_params, const st_outputs* outputs);
/* The last st_render_cached's cache work: out[0] ms of inference, target and
 * row-compaction kernels (HIP events), out[1] their launches, out[2] ms of
 * training steps,
 * out[3] steps, out[4] paths that ended in a cache query, out[5] training
 * rows with weight 1 (0 unless the call asked for its loss).  The render's
 * own kernel classes are in st_last_kernel_times. */
int st_cache_last_times(st_context* ctx, double out[6]);

/* Primary-ray closest hit only (the minimum slice): visibility for each pixel
 * and, when t_out is not null, the hit distance as the path state stores it
 * (fp16-rounded, inf on a miss).  Device pointers. */
int st_trace_primary(st_scene* scene, const st_view* view, uint32_t width, uint32_t height,
                     uint32_t* visibility, float* t_out);

/* Raw ray queries (device pointers): rays are float4 {o.xyz, tmax}, float4 {d.xyz, 0};
 * closest: hits are uint4 {t bits, instance|primitive<<16, bary.x bits, bary.y bits};
 * any-hit: occluded[i] = 1 when something lies in (0, tmax). */
int st_trace_rays(st_scene* scene, const float* rays, uint32_t count, int closest,
                  uint32_t* hits);

/* st_trace_rays without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as st_trace_rays.  Its traversal scratch (ray-fetch counter and the
 * HBM part of the per-lane stacks) is one fixed 16 MiB + 1 KiB device array
 * per GPU, held until the process ends; the traversal grid shrinks to what
 * fits it (fewer lanes for deeper trees), and a tree too deep for even one
 * block returns ST_ERROR_LIMIT, as do counts of 2^31 or more.  Calls on one
 * GPU share the array: calls that overlap on one GPU (two threads, two
 * contexts) race and must be serialised by the caller. */
int st_trace_rays_noalloc(st_scene* scene, const float* rays, uint32_t count, int closest,
                          uint32_t* hits);

/* ---- Batched rendering (no reference counterpart: the reference renders one
 * view per frame from a host-driven dispatch loop, TestRenderer.cpp:613-621).
 * A batch renders many requests, each exactly as one st_render call with the
 * same arguments would (same images, bit for bit), in one hipGraph launch:
 * the requests are spread over the context's streams, every launch of every
 * request is captured once into a graph at the first st_batch_render, and
 * later calls replay it.  Camera paths only (light tracing and the
 * bidirectional mode return ST_ERROR_UNSUPPORTED).  Requests' scenes must
 * belong to the context; outputs are device pointers, written on every
 * render; ray_counts, when given, receive the request's rays.  Like
 * st_render, a call waits for work on the null stream first and returns when
 * the images are complete.  Calls on one context must be serialised.  A
 * scene moved by st_scene_update is captured again at the next render;
 * destroy a batch before the scenes it renders. */
typedef struct st_render_request {
    st_scene* scene;
    st_view view;
    st_render_params params;
    st_outputs outputs;
} st_render_request;

/* Launch shapes of a batch (st_batch_set_config; defaults: 16 trace and 8
 * shade blocks per CU, 2^23 paths per wave, 4 streams).  The package's
 * autotuner (stratum2_amd/autotune.py) times candidates on fresh batches. */
typedef struct st_batch_config {
    uint32_t trace_blocks_per_cu;  /* traversal grid: blocks of 128 lanes per CU (1-16) */
    uint32_t shade_blocks_per_cu;  /* shading grid: blocks of 256 lanes per CU (1-32) */
    uint32_t wave_paths;           /* paths per wave when samples_per_wave is 0 (1 to 2^26) */
    uint32_t streams;              /* streams the requests are spread over (1-4) */
    double ms;                     /* reserved (0) */
} st_batch_config;

typedef struct st_batch st_batch;
int st_batch_create(st_context* ctx, const st_render_request* requests, uint32_t count, st_batch** batch);
int st_batch_render(st_batch* batch);
/* Switched off: returns ST_ERROR_UNSUPPORTED (its timed search over
 * configurations aborted the process on the host with no cause found). */
int st_batch_autotune(st_batch* batch, uint32_t repeats, st_batch_config* chosen);
int st_batch_get_config(const st_batch* batch, st_batch_config* out);
/* Changes the launch shapes; the next render captures again.  Images do
 * not depend on the configuration. */
int st_batch_set_config(st_batch* batch, const st_batch_config* config);
int st_batch_destroy(st_batch* batch);
/* create + render + destroy */
int st_render_batch(st_context* ctx, const st_render_request* requests, uint32_t count);

/* Device time of the last st_render per kernel class (HIP events on the
 * context stream) and traversal statistics for the algorithmic-bytes roofline:
 * out[0..2]  ms in closest-hit trace, shade (k_shade alone: k_ris and the reuse
 *            passes are timed apart, st_last_kernel_times), shadow trace
 * out[3..5]  launches of each class
 * out[6..7]  nodes visited / triangles tested by closest-hit rays (COUNT_TRAVERSAL)
 * out[8..9]  nodes visited / triangles tested by shadow rays (COUNT_TRAVERSAL)
 * out[10..11] closest-hit rays, shadow rays traced (traceRay calls: with
 *             media, one per segment between medium boundaries)
 * In scenes with media, out[6] and out[8] count density-grid reads (free-flight
 * collisions) of scattering and shadow rays instead of BVH nodes. */
int st_last_timings(st_context* ctx, double out[12]);

/* Device time of the last st_render per kernel (HIP events around each
 * launch, on the stream it runs on):
 * out[0..4]  ms in k_trace_closest (or k_trace_media), shading (k_shade,
 *            k_shade_light, k_shade_vc), k_trace_shadow (or k_shadow_media),
 *            k_ris (primary-vertex RIS of the reuse mode), the reuse passes
 *            (k_restir_temporal + k_restir_spatial)
 * out[5..9]  launches of each (one per bounce and wave; reuse: per sample)
 * Waves in flight on several streams overlap, so a launch's interval can hold
 * other kernels too: render with ST_FLAG_SERIAL_WAVES for exclusive times.
 * No counterpart in the reference (its profiler scopes, Profiler.hpp, time
 * command-buffer regions on the host); replaces the GUI readout of
 * ReSTIRPT.cpp:89-95 for the roofline. */
int st_last_kernel_times(st_context* ctx, double out[10]);

/* Tonemapper modes (TonemapMode, src/Shaders/compat/tonemap.h:10-23). */
#define ST_TONEMAP_RAW 0
#define ST_TONEMAP_REINHARD 1
#define ST_TONEMAP_REINHARD_EXTENDED 2
#define ST_TONEMAP_REINHARD_LUMINANCE 3
#define ST_TONEMAP_REINHARD_LUMINANCE_EXTENDED 4
#define ST_TONEMAP_UNCHARTED2 5
#define ST_TONEMAP_FILMIC 6
#define ST_TONEMAP_ACES 7
#define ST_TONEMAP_ACES_APPROX 8
#define ST_TONEMAP_VIRIDIS_R 9
#define ST_TONEMAP_VIRIDIS_LENGTH_RGB 10

/* Replaces Tonemapper::render (src/App/Tonemapper.cpp:49-96; kernels
 * maxReduce and main of src/Shaders/kernels/tonemap.slang:124-192): the
 * maximum of the (albedo-modulated) input in 1/16384 fixed point, then
 * radiance * 2^exposure, the mode's operator and, when gamma_correct, the
 * reference's sRGB encode.  hdr and out are device RGBA32F W*H images (out may
 * equal hdr); albedo (device RGBA32F) may be null.  --exposure maps here. */
int st_tonemap(st_context* ctx, const float* hdr, const float* albedo, uint32_t width, uint32_t height, float exposure,
               int mode, int gamma_correct, float* out);

/* ---- SVGF denoiser: Denoiser::denoise (src/App/Denoiser.cpp:113-255) and
 * the kernels of src/Shaders/kernels/svgf/ (temporal_accumulation.slang,
 * estimate_variance.slang, atrous.slang).  It runs between st_render and
 * st_tonemap, as TestRenderer::render does (TestRenderer.cpp:656-670). */

/* Compile-time defines of the svgf kernels (Denoiser.cpp:160-168) become flags. */
#define ST_DENOISE_REPROJECTION      (1u << 0)  /* gReprojection (default on; --noReprojection) */
#define ST_DENOISE_DEMODULATE_ALBEDO (1u << 1)  /* gDemodulateAlbedo (default on) */
#define ST_DENOISE_CHECK_NORMAL      (1u << 2)  /* gCheckNormal (default on; --noNormalCheck) */
#define ST_DENOISE_CHECK_DEPTH       (1u << 3)  /* gCheckDepth (default on; --noDepthCheck) */
#define ST_DENOISE_SHADOW_HYSTERESIS (1u << 4)  /* gShadowHysteresis (default off) */
#define ST_DENOISE_DEFAULT (ST_DENOISE_REPROJECTION | ST_DENOISE_DEMODULATE_ALBEDO | ST_DENOISE_CHECK_NORMAL | ST_DENOISE_CHECK_DEPTH)

/* FilterKernelType, src/Shaders/compat/filter_type.h:5-13. */
#define ST_FILTER_ATROUS 0
#define ST_FILTER_BOX3 1
#define ST_FILTER_BOX5 2
#define ST_FILTER_SUBSAMPLED 3
#define ST_FILTER_BOX3_SUBSAMPLED 4
#define ST_FILTER_BOX5_SUBSAMPLED 5

/* Denoiser settings (Denoiser.hpp:41-58; defaults in brackets). */
typedef struct st_denoise_params {
    uint32_t flags;                 /* ST_DENOISE_* [ST_DENOISE_DEFAULT] */
    uint32_t history_limit;         /* mHistoryLimit: cap on the accumulated sample count, 0 = none [0] */
    uint32_t atrous_iterations;     /* mAtrousIterations: filter passes after accumulation, at most 15 [0] */
    uint32_t history_tap;           /* mHistoryTap: after pass i == history_tap - 1 the first filter image's
                                       colour replaces the history colour; 0 = never [0] */
    uint32_t filter_type;           /* ST_FILTER_* [ST_FILTER_BOX3] */
    float variance_boost_length;    /* estimate_variance's mVarianceBoostLength.  The reference pushes its
                                       uint32 default 4 into this float constant, which reads as 5.6e-45
                                       and leaves the variance unboosted: 0 gives the same result [0] */
    float sigma_luminance_boost;    /* mSigmaLuminanceBoost [3] */
    float shadow_preserve_offset;   /* mShadowPreserveOffset [0.1] */
    float shadow_preserve_scale;    /* mShadowPreserveScale [1] */
} st_denoise_params;

/* One frame's inputs, all device W*H images as st_render writes them
 * (DenoiserParameters, src/Shaders/compat/denoiser.h:36-52). */
typedef struct st_denoise_inputs {
    const float* radiance;               /* RGBA32F; alpha is the frame's sample weight (1) */
    const float* albedo;                 /* RGBA32F; read only with ST_DENOISE_DEMODULATE_ALBEDO */
    const float* prev_uv;                /* RG32F previous-frame pixel coordinate / extent */
    const uint32_t* visibility;          /* 2 x u32 {instance|primitive, oct-snorm16 normal} */
    const float* depth;                  /* RGBA32F {depth, previous depth, d/dx, d/dy} */
    const uint32_t* instance_index_map;  /* device: current instance -> previous frame's index
                                            (mInstanceIndexMap); null = identity */
    uint32_t instance_count;             /* entries of the map; indices beyond it map to 0xFFFF */
} st_denoise_inputs;

typedef struct st_denoiser st_denoiser;

/* History images for one extent (Denoiser's mPrevAccumColor, mPrevAccumMoments,
 * mPrevVisibility, mPrevDepth and the two filter images). */
int st_denoiser_create(st_context* ctx, uint32_t width, uint32_t height, st_denoiser** out);
int st_denoiser_destroy(st_denoiser* d);
/* Denoiser::resetAccumulation (Denoiser.hpp:30-33): the next call starts over. */
int st_denoiser_reset(st_denoiser* d);
/* Denoiser::accumulatedFrames: TestRenderer seeds its frames with it while
 * denoising (TestRenderer.cpp:468). */
uint32_t st_denoiser_accumulated_frames(const st_denoiser* d);
/* Denoiser::denoise: without history (first call, after a reset) the output is
 * the radiance and the history is cleared; otherwise temporal accumulation,
 * then, with atrous_iterations > 0, variance estimation and the filter
 * passes.  The frame's visibility and depth are kept as the next call's
 * history.  out (device RGBA32F W*H, may equal in->radiance) receives the
 * output image; with ST_DENOISE_DEMODULATE_ALBEDO it is demodulated, and
 * st_tonemap with the same albedo restores it. */
int st_denoise(st_denoiser* d, const st_view* view, const st_denoise_inputs* in, const st_denoise_params* p,
               float* out);
/* GPU time of the last st_denoise call per pass, as the reference's
 * ProfilerScopes record it (Denoiser.cpp:186,204,219): out[0..3] = ms of
 * accumulation (or the history restart), variance estimate, filter passes and
 * history copy; out[4..7] = their launch counts.  HIP events on the context's
 * stream. */
int st_denoiser_last_timings(const st_denoiser* d, double out[8]);

#ifdef __cplusplus
}
#endif

#endif
