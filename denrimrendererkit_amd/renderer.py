"""DenrimRenderer / RenderSession over the HIP backend's C ABI.

Mirrors the reference's public session API (Renderer/DenrimRenderer.swift:52-292,
Renderer/RenderSession.swift:582-1058) with Python spelling:

=====================================  ==========================================
reference                              here
=====================================  ==========================================
DenrimRenderer(device:)                DenrimRenderer(device_ordinal=0)
makeSession(scene:settings:...)        renderer.make_session(scene, settings)
session.render(samples:)               session.render(samples)
session.renderNextSample()             session.render_next_sample()
session.encodeNextSample(into:)        session.encode_next_sample(); session.synchronize()
session.renderNextTile(_:completes..)  session.render_next_tile(x, y, w, h, completes_sample)
session.resetAccumulation()            session.reset_accumulation()
session.updateCamera(_:previous...)    session.update_camera(camera, previous_camera)
session.sampleCount                    session.sample_count
session.pixels(for:)                   session.pixels(output) -> (H, W, 4) float32
session.sdfTraversalStats()            session.sdf_traversal_stats()
=====================================  ==========================================
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .scene import AccelerationBuild, Camera, DenrimRendererError, RenderScene, compile_scene
from .settings import RenderAccelerationMode, RenderOutput, RenderQuality, RenderSettings

STAGES = ("raygen", "extend", "shade", "shadow", "resolve", "finalize", "sss", "shade_resume", "preview")
N_STATS = 37


@dataclass(frozen=True)
class SDFTraversalStats:
    """SDFTraversalStats (Renderer/RenderSession.swift:41-73)."""
    dense_volume_tests: int = 0
    dense_march_steps: int = 0
    sparse_grid_cells_visited: int = 0
    sparse_grid_empty_cells: int = 0
    sparse_grid_macro_skips: int = 0
    sparse_brick_tests: int = 0
    sparse_brick_invalid: int = 0
    sparse_brick_range_culls: int = 0
    sparse_brick_marches: int = 0
    sparse_brick_march_steps: int = 0
    sparse_brick_hits: int = 0
    primary_scene_queries: int = 0
    bounce_scene_queries: int = 0
    shadow_scene_queries: int = 0


def _ptr(a: Optional[np.ndarray]):
    return None if a is None or a.size == 0 else a.ctypes.data


def gpu_camera_struct(g: np.ndarray) -> _lib.GpuCamera:
    c = _lib.GpuCamera()
    for name, field in (("origin", "origin"), ("lowerLeft", "lower_left"), ("horizontal", "horizontal"),
                        ("vertical", "vertical"), ("lens", "lens")):
        arr = getattr(c, field)
        for k in range(4):
            arr[k] = float(g[name][k])
    return c


class DenrimRenderer:
    """DenrimRenderer.init(device:) (DenrimRenderer.swift:52-63)."""

    def __init__(self, device_ordinal: int = 0):
        L = _lib.load()
        h = C.c_void_p()
        _lib.check(L.dr_hip_device_create(device_ordinal, C.byref(h)), "device")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib.load().dr_hip_device_destroy(self._h)
            self._h = None

    def upload(self, build: AccelerationBuild, max_leaf: int = 0) -> "GpuScene":
        return GpuScene(self, build, max_leaf)

    def make_session(self, scene, settings: RenderSettings,
                     acceleration_mode: RenderAccelerationMode = RenderAccelerationMode.HIP,
                     camera: Optional[Camera] = None, samples_per_launch: int = 0) -> "RenderSession":
        """makeSession(scene:settings:accelerationMode:) (DenrimRenderer.swift:253-292).
        `scene` is a RenderScene (compiled here) or an already uploaded GpuScene."""
        if acceleration_mode not in (RenderAccelerationMode.HIP, RenderAccelerationMode.AUTOMATIC,
                                     RenderAccelerationMode.FLAT_BVH):
            raise DenrimRendererError("missingShaderFunction", f"{acceleration_mode.value} is not available")
        if settings.width <= 0 or settings.height <= 0:
            raise DenrimRendererError.invalid_scene("Render dimensions must be positive.")
        if isinstance(scene, RenderScene):
            gpu_scene = self.upload(compile_scene(scene))
            cam = camera or scene.camera
        else:
            gpu_scene = scene
            if camera is None:
                raise DenrimRendererError("invalidArgument", "a camera is required with a pre-uploaded scene")
            cam = camera
        return RenderSession(gpu_scene, settings, cam, samples_per_launch)


class GpuScene:
    """A scene resident in HBM (records + traversal BVH); shareable by sessions."""

    def __init__(self, renderer: DenrimRenderer, build: AccelerationBuild, max_leaf: int = 0):
        self.renderer = renderer
        self.build = build
        keep = [np.ascontiguousarray(x) for x in (build.triangles, build.materials, build.lights,
                                                   build.environment_samples, build.texture_descriptors,
                                                   build.texture_pixels, build.volumes, build.volume_bricks,
                                                   build.volume_brick_samples, build.volume_brick_grids,
                                                   build.volume_brick_grid_indices, build.volume_samples)]
        t, m, li, e, td, tp, v, vb, vs, vg, vi, ds = keep
        d = _lib.SceneDesc(_ptr(t), len(t), _ptr(m), len(m), _ptr(li), len(li), _ptr(e), len(e), _ptr(td), len(td),
                           _ptr(tp), len(tp), build.environment_texture_index_plus_one, build.environment_intensity,
                           build.environment_rotation_y, build.environment_max_radiance, max_leaf,
                           _ptr(v), len(v), _ptr(ds), len(ds), _ptr(vb), len(vb), _ptr(vs), len(vs), _ptr(vg), len(vg),
                           _ptr(vi), len(vi))
        h = C.c_void_p()
        _lib.check(_lib.load().dr_hip_scene_upload(renderer._h, C.byref(d), C.byref(h)), "scene upload")
        self._h = h

    def info(self) -> dict:
        n, l, d = C.c_uint32(), C.c_uint32(), C.c_uint32()
        s = C.c_double()
        _lib.check(_lib.load().dr_hip_scene_info(self._h, C.byref(n), C.byref(l), C.byref(d), C.byref(s)), "info")
        return {"nodes": n.value, "leaves": l.value, "max_depth": d.value, "build_seconds": s.value}

    def intersect(self, rays: np.ndarray, intersection: int = 0):
        """Closest triangle hits of rays (N, 8) = (origin, _, direction, _):
        (t, triangle index), +inf / 0xffffffff on a miss
        (dr_hip_scene_intersect; CPUIntersector in the reference)."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(len(r), np.float32)
        ids = np.zeros(len(r), np.uint32)
        if len(r):
            _lib.check(_lib.load().dr_hip_scene_intersect(self._h, _ptr(r), len(r), int(intersection), _ptr(t),
                                                          _ptr(ids)), "intersect")
        return t, ids

    def intersect_device(self, rays_ptr: int, count: int, t_ptr: int, ids_ptr: int, stream: int = 0,
                         intersection: int = 0) -> None:
        """The same queries on device buffers (addresses of count x 8 floats,
        count floats, count u32), enqueued on `stream` with no allocation
        (dr_hip_scene_intersect_device); the caller synchronizes."""
        _lib.check(_lib.load().dr_hip_scene_intersect_device(self._h, C.c_void_p(rays_ptr), count, int(intersection),
                                                             C.c_void_p(t_ptr), C.c_void_p(ids_ptr),
                                                             C.c_void_p(stream or None)), "intersect_device")

    def release(self):
        if getattr(self, "_h", None):
            _lib.load().dr_hip_scene_release(self._h)
            self._h = None

    def __del__(self):
        self.release()


def settings_struct(settings: RenderSettings, samples_per_launch: int = 0) -> _lib.Settings:
    """The C ABI's dr_settings for RenderSettings (RenderSettings.swift:29-115)."""
    s = _lib.Settings()
    s.width, s.height = settings.width, settings.height
    s.max_bounces = max(1, settings.max_bounces)
    s.quality = settings.quality.value
    s.sample_radiance_clamp = settings.resolved_sample_radiance_clamp
    s.transparent_background = int(settings.transparent_background)
    s.shows_environment_background = int(settings.shows_environment_background)
    for k in range(3):
        s.background_color[k] = float(settings.background_color[k])
    d = settings.denoise
    s.denoiser_enabled = int(d.denoiser != 0)
    s.denoiser = int(d.denoiser)
    s.denoise_radius, s.denoise_iterations = int(d.radius), int(d.iterations)
    s.denoise_normal_sigma, s.denoise_depth_sigma = float(d.normal_sigma), float(d.depth_sigma)
    s.denoise_albedo_sigma, s.denoise_color_sigma = float(d.albedo_sigma), float(d.color_sigma)
    s.intersection = int(settings.intersection)
    s.environment_sampling = int(settings.environment_sampling)
    s.collects_stats = int(settings.collects_sdf_traversal_stats)
    s.frame_seed = settings.frame_seed & 0xFFFFFFFF
    s.samples_per_launch = samples_per_launch
    return s


class RenderSession:
    def __init__(self, scene: GpuScene, settings: RenderSettings, camera: Camera, samples_per_launch: int = 0):
        self.scene = scene
        self.settings = settings
        self.camera = camera
        self.previous_camera = settings.previous_camera or camera
        s = settings_struct(settings, samples_per_launch)
        self._settings = s
        cam = gpu_camera_struct(camera.gpu_camera(settings.width, settings.height))
        prev = gpu_camera_struct(self.previous_camera.gpu_camera(settings.width, settings.height))
        h = C.c_void_p()
        _lib.check(_lib.load().dr_hip_session_create(scene._h, C.byref(s), C.byref(cam), C.byref(prev), C.byref(h)),
                   "session")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().dr_hip_session_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- rendering -----------------------------------------------------------
    def render(self, samples: int) -> None:
        if samples < 0:  # RenderSession.swift:1011-1013
            raise DenrimRendererError.invalid_scene("Sample count must not be negative.")
        _lib.check(_lib.load().dr_hip_session_render(self._h, samples), "render")

    def render_next_sample(self) -> None:
        self.render(1)

    def encode_next_sample(self, samples: int = 1) -> None:
        _lib.check(_lib.load().dr_hip_session_encode(self._h, samples), "encode")

    def synchronize(self) -> None:
        _lib.check(_lib.load().dr_hip_session_synchronize(self._h), "synchronize")

    def render_next_tile(self, x: int, y: int, width: int, height: int, completes_sample: bool) -> None:
        _lib.check(_lib.load().dr_hip_session_render_tile(self._h, x, y, width, height, int(completes_sample)), "tile")

    def reset_accumulation(self) -> None:
        _lib.check(_lib.load().dr_hip_session_reset(self._h), "reset")

    def update_camera(self, camera: Camera, previous_camera: Optional[Camera] = None) -> None:
        old = self.camera
        self.camera = camera
        self.previous_camera = previous_camera or old
        W, H = self.settings.width, self.settings.height
        c = gpu_camera_struct(camera.gpu_camera(W, H))
        p = gpu_camera_struct(self.previous_camera.gpu_camera(W, H))
        _lib.check(_lib.load().dr_hip_session_update_camera(self._h, C.byref(c), C.byref(p)), "camera")

    def retarget(self, scene: GpuScene, settings: RenderSettings, camera: Camera, samples_per_launch: int = 0) -> None:
        """Rebinds this session to another scene/settings/camera in its own
        device memory (dr_hip_session_retarget); raises outOfMemory when the
        new settings need larger buffers than the session was created with."""
        prev_cam = settings.previous_camera or camera
        s = settings_struct(settings, samples_per_launch)
        cam = gpu_camera_struct(camera.gpu_camera(settings.width, settings.height))
        prev = gpu_camera_struct(prev_cam.gpu_camera(settings.width, settings.height))
        _lib.check(_lib.load().dr_hip_session_retarget(self._h, scene._h, C.byref(s), C.byref(cam), C.byref(prev)),
                   "retarget")
        self.scene, self.settings, self.camera, self.previous_camera = scene, settings, camera, prev_cam
        self._settings = s

    def set_denoiser_network(self, network) -> None:
        """The UNetDenoiser used when settings.denoise selects LEARNED."""
        _lib.check(_lib.load().dr_hip_session_set_denoiser_network(self._h, network._h if network else None),
                   "denoiser network")
        self._network = network  # kept alive with the session

    def denoise(self) -> None:
        """Enqueues the configured denoiser on the current beauty (no readback);
        pixels(BEAUTY) runs it when needed (RenderSession.swift:1115-1168)."""
        _lib.check(_lib.load().dr_hip_session_denoise(self._h), "denoise")

    @property
    def sample_count(self) -> int:
        v = C.c_uint32()
        _lib.check(_lib.load().dr_hip_session_sample_count(self._h, C.byref(v)), "sample count")
        return v.value

    # -- outputs -------------------------------------------------------------
    def pixels(self, output: RenderOutput = RenderOutput.BEAUTY) -> np.ndarray:
        W, H = self.settings.width, self.settings.height
        out = np.zeros((H, W, 4), np.float32)
        _lib.check(_lib.load().dr_hip_session_read_output(self._h, int(output), _ptr(out), out.size), "read output")
        return out

    def write_png(self, path: str, output: RenderOutput = RenderOutput.BEAUTY) -> None:
        """writePNG(output:to:) (RenderSession.swift:1026-1034): the output's
        display encoding (PNGWriter.swift:38-56) as an RGBA8 PNG."""
        from .image import visualized_rgba8, write_png
        write_png(path, visualized_rgba8(self.pixels(output), output))

    def triangle_ids(self) -> np.ndarray:
        """Internal parity plane: global triangle index + 1 (0 = no hit)."""
        return self.pixels(RenderOutput.TRIANGLE_ID)[..., 0].view(np.uint32).copy()

    def stats(self) -> np.ndarray:
        """All counters (include/denrim_hip.h dr_stat order), uint64."""
        out = (C.c_uint64 * N_STATS)()
        _lib.check(_lib.load().dr_hip_session_stats(self._h, out, N_STATS), "stats")
        return np.array(list(out), dtype=np.uint64)

    def sdf_traversal_stats(self) -> SDFTraversalStats:
        """sdfTraversalStats() (RenderSession.swift:807-831)."""
        s = [int(x) for x in self.stats()]
        c = [s[8 + k] + s[18 + k] for k in range(10)]  # extension/subsurface/AOV + shadow queries
        cells, tests = c[2], c[4]
        return SDFTraversalStats(c[0], c[1], cells, cells - min(cells, tests), c[3], tests, c[5], c[6], c[7], c[8],
                                 c[9], s[0], s[1], s[2])

    def reset_sdf_traversal_stats(self) -> None:
        self.reset_stats()

    def reset_stats(self) -> None:
        _lib.check(_lib.load().dr_hip_session_reset_stats(self._h), "reset stats")

    def set_graphs(self, enabled: bool) -> None:
        """Replay full-frame sample batches as captured HIP graphs."""
        _lib.check(_lib.load().dr_hip_session_set_graphs(self._h, int(enabled)), "graphs")

    def query(self) -> bool:
        """True when all encoded work has finished (non-blocking)."""
        d = C.c_int()
        _lib.check(_lib.load().dr_hip_session_query(self._h, C.byref(d)), "query")
        return bool(d.value)

    def set_stage_timing(self, enabled: bool) -> None:
        _lib.check(_lib.load().dr_hip_session_set_stage_timing(self._h, int(enabled)), "stage timing")

    def timing(self) -> dict:
        total = C.c_double()
        ms = (C.c_double * len(STAGES))()
        n = (C.c_uint64 * len(STAGES))()
        _lib.check(_lib.load().dr_hip_session_timing(self._h, C.byref(total), ms, n, len(STAGES)), "timing")
        return {"total_ms": total.value, "stage_ms": dict(zip(STAGES, list(ms))),
                "stage_launches": dict(zip(STAGES, [int(x) for x in n]))}


# ---------------------------------------------------------------- serving --
REQUEST_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class RenderServer:
    """Device-resident render server (dr_hip_server_*): requests on scenes
    already uploaded to HBM are admitted by memory budget while others run
    (continuous batching), each replaying graph-captured sample batches on
    its own stream. A request's pixels equal those of a session rendering it
    alone."""

    ADMISSION = {"arrival": 0, "least_work": 1}

    def __init__(self, renderer: DenrimRenderer, memory_budget_bytes: int = 8 << 30, max_active: int = 16,
                 samples_per_batch: int = 0, use_graphs: bool = True, admission: str = "arrival"):
        """admission: "arrival" admits queued requests in submission order,
        "least_work" the one with the fewest pixel samples first."""
        if admission not in self.ADMISSION:
            raise ValueError(f"admission must be one of {sorted(self.ADMISSION)}")
        cfg = _lib.ServerSettings(memory_budget_bytes, max_active, samples_per_batch, int(use_graphs),
                                  self.ADMISSION[admission])
        h = C.c_void_p()
        _lib.check(_lib.load().dr_hip_server_create(renderer._h, C.byref(cfg), C.byref(h)), "server")
        self.renderer = renderer
        self._h = h
        self._sizes = {}
        self._keep = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().dr_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, scene: GpuScene, settings: RenderSettings, camera: Camera, samples: int,
               samples_per_launch: int = 0) -> int:
        s = settings_struct(settings, samples_per_launch)
        cam = gpu_camera_struct(camera.gpu_camera(settings.width, settings.height))
        t = C.c_uint64()
        _lib.check(_lib.load().dr_hip_server_submit(self._h, scene._h, C.byref(s), C.byref(cam), samples, C.byref(t)),
                   "submit")
        self._sizes[t.value] = (settings.width, settings.height)
        self._keep[t.value] = scene  # the scene must outlive the request
        return t.value

    def state(self, ticket: int) -> str:
        st = C.c_int()
        _lib.check(_lib.load().dr_hip_server_poll(self._h, ticket, C.byref(st)), "poll")
        return REQUEST_STATES[st.value]

    def wait(self, ticket: int, timeout_ms: int = 600000) -> str:
        st = C.c_int()
        _lib.check(_lib.load().dr_hip_server_wait(self._h, ticket, timeout_ms, C.byref(st)), "wait")
        return REQUEST_STATES[st.value]

    def result(self, ticket: int, output: RenderOutput = RenderOutput.BEAUTY) -> np.ndarray:
        W, H = self._sizes[ticket]
        out = np.zeros((H, W, 4), np.float32)
        _lib.check(_lib.load().dr_hip_server_result(self._h, ticket, int(output), _ptr(out), out.size), "result")
        return out

    def timing(self, ticket: int) -> dict:
        q, r = C.c_double(), C.c_double()
        _lib.check(_lib.load().dr_hip_server_timing(self._h, ticket, C.byref(q), C.byref(r)), "timing")
        return {"queued_ms": q.value, "run_ms": r.value, "latency_ms": q.value + r.value}

    def release(self, ticket: int) -> None:
        _lib.check(_lib.load().dr_hip_server_release(self._h, ticket), "release")
        self._sizes.pop(ticket, None)
        self._keep.pop(ticket, None)

    def set_denoiser_network(self, network) -> None:
        """Network for requests whose settings select RenderDenoiser.LEARNED."""
        _lib.check(_lib.load().dr_hip_server_set_denoiser_network(self._h, network._h if network else None),
                   "denoiser network")
        self._network = network

    def usage(self) -> dict:
        b, a, q = C.c_uint64(), C.c_uint32(), C.c_uint32()
        _lib.check(_lib.load().dr_hip_server_usage(self._h, C.byref(b), C.byref(a), C.byref(q)), "usage")
        return {"used_bytes": b.value, "active": a.value, "queued": q.value}
