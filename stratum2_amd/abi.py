"""ctypes mirror of include/stratum2.h and the loader of the HIP library.

The structures mirror the reference's scene buffer formats
(src/Shaders/compat/scene.h:30-198, transform.h:7-191) and the TestRenderer
push constants (src/Shaders/kernels/testrenderer.slang:1-21).  The loader
raises when the built library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

u32, i32, f32, u64 = C.c_uint32, C.c_int32, C.c_float, C.c_uint64

ST_OK = 0
ST_ERROR_INVALID_ARGUMENT = 1
ST_ERROR_UNSUPPORTED = 2
ST_ERROR_HIP = 3
ST_ERROR_OUT_OF_MEMORY = 4
ST_ERROR_LIMIT = 5
ST_INVALID_INSTANCE = 0xFFFF
ST_NO_IMAGE = 0xFFFFFFFF

ST_FLAG_LAMBERTIAN = 1 << 0
ST_FLAG_SAMPLE_DIRECT = 1 << 1
ST_FLAG_RESTIR_DI = 1 << 2
ST_FLAG_RESTIR_DI_REUSE = 1 << 3
ST_FLAG_ENV_ALIAS = 1 << 4
ST_FLAG_ENV_MIS_BALANCED = 1 << 5
ST_FLAG_FLOAT_NEE = 1 << 6
ST_FLAG_PIXEL_JITTER = 1 << 7
ST_FLAG_COUNT_TRAVERSAL = 1 << 8
ST_FLAG_DIRECT_ONLY = 1 << 9
ST_FLAG_TRACE_FROM_LIGHT = 1 << 10  # TestRenderer light tracing (gTraceFromLight)
ST_FLAG_BDPT_LVC = 1 << 11  # TestRenderer bidirectional mode with a light vertex cache (gUseVC)
ST_FLAG_VC_FP32_MIS = 1 << 12  # dVC / dVCM in fp32 instead of the reference's fp16
ST_FLAG_SERIAL_WAVES = 1 << 13  # waves one after another on one stream (exclusive per-kernel times)
ST_FLAG_RESTIR_DI_HASHGRID = 1 << 14  # the reference's hash-grid reservoir reuse (gReSTIR_DI_Reuse)
ST_FLAG_HISTORY_RESET = 1 << 15  # drop the hash grid kept from the previous call (a camera cut)
ST_FLAG_LAMBERT_GGX = 1 << 16  # per material: GGX conductor (metallic >= 0.5) or Lambert
ST_FLAG_ASYNC = 1 << 17  # st_render returns once enqueued; outputs complete after st_sync
ST_FLAG_RESTIR_DI_DEEP = 1 << 18  # with the screen-space reuse: RIS at the deeper NEE vertices too
ST_FLAG_DEFAULT = ST_FLAG_SAMPLE_DIRECT

# TonemapMode (src/Shaders/compat/tonemap.h:10-23)
(ST_TONEMAP_RAW, ST_TONEMAP_REINHARD, ST_TONEMAP_REINHARD_EXTENDED, ST_TONEMAP_REINHARD_LUMINANCE,
 ST_TONEMAP_REINHARD_LUMINANCE_EXTENDED, ST_TONEMAP_UNCHARTED2, ST_TONEMAP_FILMIC, ST_TONEMAP_ACES, ST_TONEMAP_ACES_APPROX,
 ST_TONEMAP_VIRIDIS_R, ST_TONEMAP_VIRIDIS_LENGTH_RGB) = range(11)


class st_instance(C.Structure):
    _fields_ = [("type_material_address", u32), ("data", u32)]


class st_transform(C.Structure):
    _fields_ = [("m", (f32 * 4) * 3)]


class st_mesh_vertex_info(C.Structure):
    _fields_ = [("packed_buffer_indices", u32 * 2), ("packed_strides", u32), ("pad", u32),
                ("packed_offsets", u32 * 4)]


class st_volume_info(C.Structure):
    _fields_ = [("min", f32 * 3), ("instance_index", u32), ("max", f32 * 3), ("pad", u32)]


class st_buffer_view(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size_bytes", u64)]


class st_image_desc(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("width", u32), ("height", u32)]


class st_scene_desc(C.Structure):
    _fields_ = [
        ("instance_count", u32), ("instances", C.c_void_p), ("transforms", C.c_void_p),
        ("material_word_count", u32), ("material_data", C.c_void_p),
        ("mesh_vertex_info_count", u32), ("mesh_vertex_info", C.c_void_p),
        ("vertex_buffer_count", u32), ("vertex_buffers", C.c_void_p),
        ("image_count", u32), ("images", C.c_void_p),
        ("environment_material_address", u32),
        ("volume_count", u32), ("volumes", C.c_void_p),
        ("volume_info_count", u32), ("volume_info", C.c_void_p),
    ]


class st_projection(C.Structure):
    _fields_ = [("scale", f32 * 2), ("offset", f32 * 2), ("near_plane", f32), ("far_plane", f32),
                ("sensor_area", f32), ("vertical_fov", f32)]


class st_view(C.Structure):
    _fields_ = [("projection", st_projection), ("image_min", i32 * 2), ("image_max", i32 * 2),
                ("camera_to_world", st_transform)]


class st_render_params(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("max_depth", u32), ("min_path_length", u32),
        ("seed", u32), ("spp", u32), ("max_null_collisions", u32),
        ("environment_sample_probability", f32), ("candidate_samples", u32), ("max_m", f32),
        ("flags", u32), ("rr_start_depth", u32), ("exposure", f32),
        ("spatial_neighbours", u32), ("spatial_radius", f32), ("samples_per_wave", u32),
        ("light_subpath_factor", f32), ("hashgrid_cell_size", f32),
    ]


class st_outputs(C.Structure):
    _fields_ = [("hdr", C.c_void_p), ("visibility", C.c_void_p), ("path_length", C.c_void_p),
                ("ray_counts", C.c_void_p), ("albedo", C.c_void_p), ("depth", C.c_void_p), ("prev_uv", C.c_void_p)]


class st_render_request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("view", st_view), ("params", st_render_params), ("outputs", st_outputs)]


class st_cache_config(C.Structure):
    _fields_ = [("levels", u32), ("log2_table", u32), ("base_resolution", u32), ("max_resolution", u32),
                ("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("seed", u32), ("ema_decay", C.c_float)]


def default_cache_config(**kw):
    c = st_cache_config(levels=16, log2_table=19, base_resolution=16, max_resolution=2048, learning_rate=1e-2,
                        beta1=0.9, beta2=0.99, epsilon=1e-15, seed=0, ema_decay=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class st_cache_render_params(C.Structure):
    _fields_ = [("cache_depth", u32), ("train_stride", u32), ("train_steps", u32), ("frame", u32),
                ("train_queries", C.c_void_p), ("train_targets", C.c_void_p), ("loss", C.POINTER(C.c_float))]


class st_batch_config(C.Structure):
    _fields_ = [("trace_blocks_per_cu", u32), ("shade_blocks_per_cu", u32), ("wave_paths", u32), ("streams", u32),
                ("ms", C.c_double)]


def default_params(width, height, **kw):
    """TestRenderer defaults (src/App/TestRenderer.cpp:38-43)."""
    p = st_render_params()
    p.width, p.height = width, height
    p.max_depth = 5
    p.min_path_length = 0
    p.seed = 0
    p.spp = 1
    p.max_null_collisions = 1000
    p.environment_sample_probability = 0.9
    p.candidate_samples = 32
    p.max_m = 3.0
    p.flags = ST_FLAG_DEFAULT
    p.rr_start_depth = 0
    p.exposure = 0.0
    p.spatial_neighbours = 5
    p.spatial_radius = 30.0
    p.samples_per_wave = 0
    p.light_subpath_factor = 1.0  # mLightSubpathCount, TestRenderer.hpp:29
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


# SVGF denoiser (Denoiser.cpp / svgf kernels)
ST_DENOISE_REPROJECTION = 1 << 0
ST_DENOISE_DEMODULATE_ALBEDO = 1 << 1
ST_DENOISE_CHECK_NORMAL = 1 << 2
ST_DENOISE_CHECK_DEPTH = 1 << 3
ST_DENOISE_SHADOW_HYSTERESIS = 1 << 4
ST_DENOISE_DEFAULT = (ST_DENOISE_REPROJECTION | ST_DENOISE_DEMODULATE_ALBEDO | ST_DENOISE_CHECK_NORMAL
                      | ST_DENOISE_CHECK_DEPTH)
ST_FILTER_ATROUS, ST_FILTER_BOX3, ST_FILTER_BOX5, ST_FILTER_SUBSAMPLED, ST_FILTER_BOX3_SUBSAMPLED, \
    ST_FILTER_BOX5_SUBSAMPLED = range(6)


class st_denoise_params(C.Structure):
    _fields_ = [("flags", u32), ("history_limit", u32), ("atrous_iterations", u32), ("history_tap", u32),
                ("filter_type", u32), ("variance_boost_length", f32), ("sigma_luminance_boost", f32),
                ("shadow_preserve_offset", f32), ("shadow_preserve_scale", f32)]


class st_denoise_inputs(C.Structure):
    _fields_ = [("radiance", C.c_void_p), ("albedo", C.c_void_p), ("prev_uv", C.c_void_p),
                ("visibility", C.c_void_p), ("depth", C.c_void_p), ("instance_index_map", C.c_void_p),
                ("instance_count", u32)]


def default_denoise_params(**kw):
    """Denoiser defaults (src/App/Denoiser.hpp:41-58 and the constructor,
    Denoiser.cpp:24-31).  The reference's variance boost length (uint32 4
    pushed into a float constant) reads as 5.6e-45 and boosts nothing; 0 is
    the same."""
    p = st_denoise_params()
    p.flags = ST_DENOISE_DEFAULT
    p.history_limit = 0
    p.atrous_iterations = 0
    p.history_tap = 0
    p.filter_type = ST_FILTER_BOX3
    p.variance_boost_length = 0.0
    p.sigma_luminance_boost = 3.0
    p.shadow_preserve_offset = 0.1
    p.shadow_preserve_scale = 1.0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libstratum2.so")

EXPORTS = [
    "st_last_error", "st_version", "st_context_create", "st_context_destroy", "st_sync",
    "st_scene_create", "st_scene_destroy", "st_scene_stats", "st_render", "st_trace_primary",
    "st_trace_rays", "st_trace_rays_noalloc", "st_last_timings", "st_last_kernel_times", "st_tonemap", "st_scene_update",
    "st_denoiser_create", "st_denoiser_destroy", "st_denoiser_reset", "st_denoiser_accumulated_frames", "st_denoise",
    "st_denoiser_last_timings",
    "st_batch_create", "st_batch_render", "st_batch_autotune", "st_batch_get_config", "st_batch_set_config",
    "st_batch_destroy", "st_render_batch", "st_scene_set_triangle_albedo", "st_scene_set_volume_grid", "st_adjoint",
    "st_cache_create", "st_cache_destroy", "st_cache_infer", "st_cache_train", "st_cache_param_counts", "st_cache_get_params",
    "st_cache_set_params", "st_render_cached", "st_cache_last_times",
]

_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library; raises when it has not been built."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"stratum2 HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.st_last_error.restype = C.c_char_p
    lib.st_version.restype = C.c_int
    lib.st_context_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.st_context_destroy.argtypes = [vp]
    lib.st_sync.argtypes = [vp]
    lib.st_scene_create.argtypes = [vp, C.POINTER(st_scene_desc), C.POINTER(vp)]
    lib.st_scene_destroy.argtypes = [vp]
    lib.st_scene_stats.argtypes = [vp, C.POINTER(u64)]
    lib.st_render.argtypes = [vp, C.POINTER(st_view), C.POINTER(st_render_params), C.POINTER(st_outputs)]
    lib.st_trace_primary.argtypes = [vp, C.POINTER(st_view), u32, u32, vp, vp]
    lib.st_trace_rays.argtypes = [vp, vp, u32, C.c_int, vp]
    lib.st_trace_rays_noalloc.argtypes = [vp, vp, u32, C.c_int, vp]
    lib.st_tonemap.argtypes = [vp, vp, vp, u32, u32, C.c_float, C.c_int, C.c_int, vp]
    lib.st_scene_update.argtypes = [vp, vp, C.POINTER(C.c_double)]
    lib.st_last_timings.argtypes = [vp, C.POINTER(C.c_double)]
    lib.st_last_kernel_times.argtypes = [vp, C.POINTER(C.c_double)]
    lib.st_denoiser_create.argtypes = [vp, u32, u32, C.POINTER(vp)]
    lib.st_denoiser_destroy.argtypes = [vp]
    lib.st_denoiser_reset.argtypes = [vp]
    lib.st_denoiser_accumulated_frames.argtypes = [vp]
    lib.st_denoiser_accumulated_frames.restype = u32
    lib.st_denoise.argtypes = [vp, C.POINTER(st_view), C.POINTER(st_denoise_inputs), C.POINTER(st_denoise_params), vp]
    lib.st_denoiser_last_timings.argtypes = [vp, C.POINTER(C.c_double)]
    lib.st_batch_create.argtypes = [vp, C.POINTER(st_render_request), u32, C.POINTER(vp)]
    lib.st_batch_render.argtypes = [vp]
    lib.st_batch_autotune.argtypes = [vp, u32, C.POINTER(st_batch_config)]
    lib.st_batch_get_config.argtypes = [vp, C.POINTER(st_batch_config)]
    lib.st_batch_set_config.argtypes = [vp, C.POINTER(st_batch_config)]
    lib.st_batch_destroy.argtypes = [vp]
    lib.st_scene_set_triangle_albedo.argtypes = [vp, vp, u32]
    lib.st_scene_set_volume_grid.argtypes = [vp, u32, vp, C.c_uint64]
    lib.st_adjoint.argtypes = [vp, C.POINTER(st_view), C.POINTER(st_render_params), vp, vp, vp, u32]
    lib.st_render_batch.argtypes = [vp, C.POINTER(st_render_request), u32]
    f3 = C.c_float * 3
    lib.st_cache_create.argtypes = [vp, C.POINTER(st_cache_config), f3, f3, C.POINTER(vp)]
    lib.st_cache_destroy.argtypes = [vp]
    lib.st_cache_infer.argtypes = [vp, vp, u32, vp]
    lib.st_cache_train.argtypes = [vp, vp, vp, u32, C.POINTER(C.c_float)]
    lib.st_cache_param_counts.argtypes = [vp, C.POINTER(u32), C.POINTER(u32)]
    lib.st_cache_get_params.argtypes = [vp, vp, vp, vp, vp]
    lib.st_cache_set_params.argtypes = [vp, vp, vp]
    lib.st_render_cached.argtypes = [vp, vp, C.POINTER(st_view), C.POINTER(st_render_params),
                                     C.POINTER(st_cache_render_params), C.POINTER(st_outputs)]
    lib.st_cache_last_times.argtypes = [vp, C.POINTER(C.c_double)]
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(lib, status, what):
    if status != ST_OK:
        msg = lib.st_last_error()
        raise RuntimeError(f"{what} failed ({status}): {msg.decode() if msg else ''}")
