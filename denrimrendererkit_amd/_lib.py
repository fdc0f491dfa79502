"""Loader of the in-tree gfx950 library and ctypes bindings of its C ABI
(include/denrim_hip.h). There is no fallback: if the library is missing the
import of the renderer fails with an error that says how to build it."""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdenrim_hip.so")

DR_OK = 0
STATUS_NAMES = {
    1: "noDevice", 2: "missingShaderFunction", 3: "invalidScene", 4: "commandBufferFailed",
    5: "pngExportFailed", 6: "invalidArgument", 7: "outOfMemory",
}

# Exported symbols declared in include/denrim_hip.h.
EXPORTS = (
    "dr_last_error", "dr_hip_device_create", "dr_hip_device_destroy", "dr_hip_scene_upload",
    "dr_hip_scene_release", "dr_hip_scene_info", "dr_hip_session_create", "dr_hip_session_destroy",
    "dr_hip_session_render", "dr_hip_session_encode", "dr_hip_session_synchronize",
    "dr_hip_session_render_tile", "dr_hip_session_reset", "dr_hip_session_update_camera",
    "dr_hip_session_sample_count", "dr_hip_session_read_output", "dr_hip_session_stats",
    "dr_hip_session_reset_stats", "dr_hip_session_set_stage_timing", "dr_hip_session_timing",
    "dr_sdf_build_sparse", "dr_sdf_volume_info", "dr_sdf_volume_export", "dr_sdf_volume_free",
    "dr_sdf_build_dense", "dr_hip_session_estimate_bytes", "dr_hip_session_set_graphs", "dr_hip_session_query",
    "dr_hip_server_create", "dr_hip_server_destroy", "dr_hip_server_submit", "dr_hip_server_poll",
    "dr_hip_server_wait", "dr_hip_server_result", "dr_hip_server_timing", "dr_hip_server_release",
    "dr_hip_server_usage", "dr_hip_session_retarget", "dr_hip_session_denoise",
    "dr_hip_session_set_denoiser_network", "dr_hip_unet_create", "dr_hip_unet_destroy", "dr_hip_unet_param_count",
    "dr_hip_unet_get_params", "dr_hip_unet_set_params", "dr_hip_unet_set_batch", "dr_hip_unet_forward_backward",
    "dr_hip_unet_get_grads", "dr_hip_unet_get_output", "dr_hip_unet_train_step", "dr_hip_unet_dataset_reserve",
    "dr_hip_unet_dataset_add", "dr_hip_unet_train_dataset", "dr_hip_server_set_denoiser_network",
    "dr_hip_scene_intersect",
    "dr_hip_scene_intersect_device",
)


class SceneDesc(C.Structure):
    _fields_ = [
        ("triangles", C.c_void_p), ("triangle_count", C.c_uint32),
        ("materials", C.c_void_p), ("material_count", C.c_uint32),
        ("lights", C.c_void_p), ("light_count", C.c_uint32),
        ("environment_samples", C.c_void_p), ("environment_sample_count", C.c_uint32),
        ("texture_descriptors", C.c_void_p), ("texture_count", C.c_uint32),
        ("texture_pixels", C.c_void_p), ("texture_pixel_count", C.c_uint64),
        ("environment_texture_index_plus_one", C.c_uint32),
        ("environment_intensity", C.c_float), ("environment_rotation_y", C.c_float),
        ("environment_max_radiance", C.c_float), ("max_leaf_triangles", C.c_uint32),
        ("volumes", C.c_void_p), ("volume_count", C.c_uint32),
        ("volume_samples", C.c_void_p), ("volume_sample_count", C.c_uint32),
        ("volume_bricks", C.c_void_p), ("volume_brick_count", C.c_uint32),
        ("volume_brick_samples", C.c_void_p), ("volume_brick_sample_count", C.c_uint32),
        ("volume_brick_grids", C.c_void_p), ("volume_brick_grid_count", C.c_uint32),
        ("volume_brick_grid_indices", C.c_void_p), ("volume_brick_grid_index_count", C.c_uint32),
    ]


class SdfPrimitive(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("material", C.c_uint32), ("center", C.c_float * 3), ("radius", C.c_float),
                ("axis", C.c_float * 3), ("minor_radius", C.c_float), ("half_extents", C.c_float * 3),
                ("corner_radius", C.c_float), ("half_height", C.c_float), ("smooth_radius", C.c_float),
                ("operation", C.c_uint32), ("has_transform", C.c_uint32), ("world_to_primitive", C.c_float * 16),
                ("distance_scale", C.c_float), ("lipschitz", C.c_float)]


class SdfBuildSettings(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("brick_size", C.c_uint32), ("apron", C.c_uint32),
                ("narrow_band", C.c_float), ("smooth_radius", C.c_float), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3), ("quantize", C.c_uint32)]


class Settings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("max_bounces", C.c_uint32), ("quality", C.c_uint32),
        ("sample_radiance_clamp", C.c_float), ("transparent_background", C.c_uint32),
        ("shows_environment_background", C.c_uint32), ("background_color", C.c_float * 3),
        ("denoiser_enabled", C.c_uint32), ("collects_stats", C.c_uint32), ("frame_seed", C.c_uint32),
        ("samples_per_launch", C.c_uint32), ("denoiser", C.c_uint32), ("denoise_radius", C.c_uint32),
        ("denoise_iterations", C.c_uint32), ("denoise_normal_sigma", C.c_float), ("denoise_depth_sigma", C.c_float),
        ("denoise_albedo_sigma", C.c_float), ("denoise_color_sigma", C.c_float), ("intersection", C.c_uint32),
        ("environment_sampling", C.c_uint32),
    ]


class ServerSettings(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("max_active", C.c_uint32), ("samples_per_batch", C.c_uint32),
                ("use_graphs", C.c_uint32), ("admission", C.c_uint32)]


class GpuCamera(C.Structure):
    _fields_ = [("origin", C.c_float * 4), ("lower_left", C.c_float * 4), ("horizontal", C.c_float * 4),
                ("vertical", C.c_float * 4), ("lens", C.c_float * 4)]


_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Loads libdenrim_hip.so (built by __graft_entry__.build() or
    `python -m denrimrendererkit_amd._build`)."""
    global _lib
    if _lib is not None:
        return _lib
    # DENRIM_HIP_LIBRARY names another build of the library in this package
    # folder (libdenrim_hip_lds4.so: the small-LDS-stack build the tests use).
    path = os.path.join(PKG, os.path.basename(os.environ.get("DENRIM_HIP_LIBRARY", "libdenrim_hip.so")))
    if not os.path.exists(path):
        raise LibraryMissing(f"{path} is not built; run `python -m denrimrendererkit_amd._build`")
    L = C.CDLL(path)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    L.dr_last_error.restype = C.c_char_p
    L.dr_hip_device_create.argtypes = [i32, C.POINTER(vp)]
    L.dr_hip_device_destroy.argtypes = [vp]
    L.dr_hip_scene_upload.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(vp)]
    L.dr_hip_scene_release.argtypes = [vp]
    L.dr_hip_scene_info.argtypes = [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_double)]
    L.dr_hip_session_create.argtypes = [vp, C.POINTER(Settings), C.POINTER(GpuCamera), C.POINTER(GpuCamera),
                                        C.POINTER(vp)]
    L.dr_hip_session_destroy.argtypes = [vp]
    L.dr_hip_session_retarget.argtypes = [vp, vp, C.POINTER(Settings), C.POINTER(GpuCamera), C.POINTER(GpuCamera)]
    L.dr_hip_session_render.argtypes = [vp, u32]
    L.dr_hip_session_encode.argtypes = [vp, u32]
    L.dr_hip_session_synchronize.argtypes = [vp]
    L.dr_hip_session_render_tile.argtypes = [vp, u32, u32, u32, u32, i32]
    L.dr_hip_session_reset.argtypes = [vp]
    L.dr_hip_session_update_camera.argtypes = [vp, C.POINTER(GpuCamera), C.POINTER(GpuCamera)]
    L.dr_hip_session_sample_count.argtypes = [vp, C.POINTER(u32)]
    L.dr_hip_session_read_output.argtypes = [vp, i32, vp, C.c_size_t]
    L.dr_hip_session_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.c_size_t]
    L.dr_hip_session_reset_stats.argtypes = [vp]
    L.dr_hip_session_set_stage_timing.argtypes = [vp, i32]
    L.dr_hip_session_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_uint64), C.c_size_t]
    L.dr_sdf_build_sparse.argtypes = [C.POINTER(SdfPrimitive), u32, C.POINTER(SdfBuildSettings), C.POINTER(vp)]
    L.dr_sdf_volume_info.argtypes = [vp, C.POINTER(u32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.dr_sdf_volume_export.argtypes = [vp] + [vp] * 8
    L.dr_sdf_volume_free.argtypes = [vp]
    L.dr_sdf_build_dense.argtypes = [C.POINTER(SdfPrimitive), u32, C.POINTER(SdfBuildSettings)] + [vp] * 4
    u64, dbl = C.c_uint64, C.c_double
    L.dr_hip_session_estimate_bytes.argtypes = [vp, C.POINTER(Settings), C.POINTER(u64)]
    L.dr_hip_session_set_graphs.argtypes = [vp, i32]
    L.dr_hip_session_query.argtypes = [vp, C.POINTER(i32)]
    L.dr_hip_server_create.argtypes = [vp, C.POINTER(ServerSettings), C.POINTER(vp)]
    L.dr_hip_server_destroy.argtypes = [vp]
    L.dr_hip_server_submit.argtypes = [vp, vp, C.POINTER(Settings), C.POINTER(GpuCamera), u32, C.POINTER(u64)]
    L.dr_hip_server_poll.argtypes = [vp, u64, C.POINTER(i32)]
    L.dr_hip_server_wait.argtypes = [vp, u64, u32, C.POINTER(i32)]
    L.dr_hip_server_result.argtypes = [vp, u64, i32, vp, C.c_size_t]
    L.dr_hip_server_timing.argtypes = [vp, u64, C.POINTER(dbl), C.POINTER(dbl)]
    L.dr_hip_server_release.argtypes = [vp, u64]
    L.dr_hip_server_usage.argtypes = [vp, C.POINTER(u64), C.POINTER(u32), C.POINTER(u32)]
    f32, fp = C.c_float, C.POINTER(C.c_float)
    L.dr_hip_session_denoise.argtypes = [vp]
    L.dr_hip_session_set_denoiser_network.argtypes = [vp, vp]
    L.dr_hip_unet_create.argtypes = [vp, u32, C.POINTER(vp)]
    L.dr_hip_unet_destroy.argtypes = [vp]
    L.dr_hip_unet_param_count.argtypes = [vp, C.POINTER(u64)]
    L.dr_hip_unet_get_params.argtypes = [vp, vp, C.c_size_t]
    L.dr_hip_unet_set_params.argtypes = [vp, vp, C.c_size_t]
    L.dr_hip_unet_set_batch.argtypes = [vp, vp, vp, u32, u32, u32]
    L.dr_hip_unet_forward_backward.argtypes = [vp, fp]
    L.dr_hip_unet_get_grads.argtypes = [vp, vp, C.c_size_t]
    L.dr_hip_unet_get_output.argtypes = [vp, vp, C.c_size_t]
    L.dr_hip_unet_train_step.argtypes = [vp, f32, fp]
    L.dr_hip_unet_dataset_reserve.argtypes = [vp, u32, u32, u32]
    L.dr_hip_unet_dataset_add.argtypes = [vp, vp, vp, C.POINTER(u32)]
    L.dr_hip_unet_train_dataset.argtypes = [vp, u32, u32, f32, f32, u32, vp]
    L.dr_hip_server_set_denoiser_network.argtypes = [vp, vp]
    L.dr_hip_scene_intersect.argtypes = [vp, vp, u32, u32, vp, vp]
    L.dr_hip_scene_intersect_device.argtypes = [vp, vp, u32, u32, vp, vp, vp]
    for name in EXPORTS:
        fn = getattr(L, name)
        if name != "dr_last_error":
            fn.restype = C.c_int
    _lib = L
    return L


def check(status: int, what: str):
    if status != DR_OK:
        from .scene import DenrimRendererError
        msg = load().dr_last_error().decode(errors="replace")
        raise DenrimRendererError(STATUS_NAMES.get(status, f"status{status}"), f"{what}: {msg}")
