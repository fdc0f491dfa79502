"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CNT_PRIMARY, CNT_BOUNCE, CNT_SHADOW, CNT_NODES, CNT_TRIS = 0, 1, 2, 3, 4
CNT_SDF = 8  # + the reference's SDF counter slot (0 dense tests, 1 dense steps, 2 grid cells ... 9 brick hits)
N_COUNTERS = 24


def build(force: bool = False) -> str:
    """Builds liboracle.so and the BSDF pin library liboracle_pins.so."""
    pins = os.path.join(HERE, "liboracle_pins.so")
    if force or not os.path.exists(LIB_PATH) or not os.path.exists(pins):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "all"], check=True)
    return LIB_PATH


class _SceneDesc(C.Structure):
    _fields_ = [
        ("triangles", C.c_void_p), ("triangleCount", C.c_uint32),
        ("materials", C.c_void_p), ("materialCount", C.c_uint32),
        ("lights", C.c_void_p), ("lightCount", C.c_uint32),
        ("envSamples", C.c_void_p), ("envSampleCount", C.c_uint32),
        ("textureDescs", C.c_void_p), ("textureCount", C.c_uint32),
        ("texturePixels", C.c_void_p), ("texturePixelCount", C.c_uint64),
        ("volumes", C.c_void_p), ("volumeCount", C.c_uint32),
        ("volumeSamples", C.c_void_p), ("volumeSampleCount", C.c_uint32),
        ("bricks", C.c_void_p), ("brickCount", C.c_uint32),
        ("brickSamples", C.c_void_p), ("brickSampleCount", C.c_uint32),
        ("brickGrids", C.c_void_p), ("brickGridCount", C.c_uint32),
        ("brickGridIndices", C.c_void_p), ("brickGridIndexCount", C.c_uint32),
    ]


class _Settings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("maxBounces", C.c_uint32),
        ("frameSeed", C.c_uint32), ("transparentBackground", C.c_uint32),
        ("showsEnvironmentBackground", C.c_uint32), ("environmentTextureIndexPlusOne", C.c_uint32),
        ("denoiserEnabled", C.c_uint32), ("environmentIntensity", C.c_float),
        ("environmentRotationY", C.c_float), ("environmentMaxRadiance", C.c_float),
        ("sampleRadianceClamp", C.c_float), ("backgroundColor", C.c_float * 4),
        ("tileX", C.c_uint32), ("tileY", C.c_uint32), ("tileWidth", C.c_uint32), ("tileHeight", C.c_uint32),
        ("renderQuality", C.c_uint32), ("intersection", C.c_uint32), ("environmentSampling", C.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.or_scene_create.argtypes = [C.POINTER(_SceneDesc), C.c_uint32, C.POINTER(C.c_void_p)]
        L.or_scene_destroy.argtypes = [C.c_void_p]
        L.or_scene_node_count.argtypes = [C.c_void_p]
        L.or_scene_node_count.restype = C.c_uint32
        L.or_scene_export_bvh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_render.argtypes = [C.c_void_p, C.POINTER(_Settings), C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.or_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        L.or_environment_alias.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        L.or_hash.argtypes = [C.c_uint32]
        L.or_hash.restype = C.c_uint32
        L.or_random_float.argtypes = [C.POINTER(C.c_uint32)]
        L.or_random_float.restype = C.c_float
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None or a.size == 0 else a.ctypes.data


class OracleScene:
    """The oracle's copy of a compiled scene plus its count-median BVH."""

    def __init__(self, build, max_leaf: int = 4):
        self.build = build
        L = lib()
        self._keep = [np.ascontiguousarray(x) for x in (build.triangles, build.materials, build.lights,
                                                         build.environment_samples, build.texture_descriptors,
                                                         build.texture_pixels, build.volumes, build.volume_bricks,
                                                         build.volume_brick_samples, build.volume_brick_grids,
                                                         build.volume_brick_grid_indices, build.volume_samples)]
        t, m, li, e, td, tp, v, vb, vs, vg, vi, ds = self._keep
        d = _SceneDesc(_ptr(t), len(t), _ptr(m), len(m), _ptr(li), len(li), _ptr(e), len(e),
                       _ptr(td), len(td), _ptr(tp), len(tp), _ptr(v), len(v), _ptr(ds), len(ds),
                       _ptr(vb), len(vb), _ptr(vs), len(vs), _ptr(vg), len(vg), _ptr(vi), len(vi))
        h = C.c_void_p()
        rc = L.or_scene_create(C.byref(d), max_leaf, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"or_scene_create failed ({rc})")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            lib().or_scene_destroy(self.handle)
            self.handle = None

    @property
    def node_count(self) -> int:
        return int(lib().or_scene_node_count(self.handle))

    def export_bvh(self):
        from denrimrendererkit_amd import gpu_types as gt  # layout only
        nodes = np.zeros(self.node_count, dtype=gt.NODE)
        prim = np.zeros(len(self.build.triangles), dtype=np.uint32)
        lib().or_scene_export_bvh(self.handle, _ptr(nodes), _ptr(prim))
        return nodes, prim

    def environment_alias(self):
        """(prob float32, alias uint32) of the environment alias table."""
        n = len(self.build.environment_samples)
        prob = np.zeros(n, np.float32)
        alias = np.zeros(n, np.uint32)
        assert lib().or_environment_alias(self.handle, _ptr(prob), _ptr(alias), n) == 0
        return prob, alias

    def intersect(self, rays: np.ndarray, brute_force: bool = False, with_hit: bool = False, intersection: int = 0):
        """Closest hits of rays (N, 8) = (origin, _, direction, _): t, primitive
        id (~0 = miss) and, with_hit, (N, 8) = normal xyz, materialID,
        materialID2, materialBlend, objectID, frontFacing."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        t = np.zeros(len(rays), np.float32)
        ids = np.zeros(len(rays), np.uint32)
        hit = np.zeros((len(rays), 8), np.float32) if with_hit else None
        lib().or_intersect(self.handle, _ptr(rays), len(rays), 1 if brute_force else 0, int(intersection), _ptr(t),
                           _ptr(ids), _ptr(hit))
        return (t, ids, hit) if with_hit else (t, ids)


@dataclass
class OracleResult:
    accumulation: np.ndarray   # (H, W, 4)
    aovs: np.ndarray           # (7, H, W, 4): depth, normal, albedo, materialID, objectID, motion, triangleID
    counters: np.ndarray       # u64[8]


def make_settings(build, width, height, max_bounces=4, frame_seed=0x1234ABCD, clamp=64.0,
                  transparent_background=False, shows_environment_background=True, background=(0, 0, 0),
                  denoiser=False, tile=None, quality=2, intersection=0, environment_sampling=0) -> _Settings:
    s = _Settings()
    s.width, s.height, s.maxBounces, s.frameSeed = width, height, max(1, max_bounces), frame_seed
    s.transparentBackground = int(transparent_background)
    s.showsEnvironmentBackground = int(shows_environment_background)
    s.environmentTextureIndexPlusOne = build.environment_texture_index_plus_one
    s.denoiserEnabled = int(denoiser)
    s.environmentIntensity = build.environment_intensity
    s.environmentRotationY = build.environment_rotation_y
    s.environmentMaxRadiance = build.environment_max_radiance
    s.sampleRadianceClamp = clamp
    for k in range(3):
        s.backgroundColor[k] = background[k]
    s.backgroundColor[3] = 1.0
    x, y, w, h = tile if tile is not None else (0, 0, width, height)
    s.tileX, s.tileY, s.tileWidth, s.tileHeight = x, y, w, h
    s.renderQuality = quality
    s.intersection = intersection
    s.environmentSampling = environment_sampling
    return s


def render(scene: OracleScene, settings: _Settings, camera: np.ndarray, previous_camera: Optional[np.ndarray] = None,
           sample_begin: int = 0, sample_count: int = 1, precision: int = 32, brute_force: bool = False,
           accumulation: Optional[np.ndarray] = None, threads: int = 0, with_aovs: bool = True) -> OracleResult:
    W, H = settings.width, settings.height
    dt = np.float64 if precision == 64 else np.float32
    acc = np.zeros((H, W, 4), dt) if accumulation is None else accumulation
    assert acc.dtype == dt and acc.flags.c_contiguous
    aov = np.zeros((7, H, W, 4), dt) if with_aovs else None
    cnt = np.zeros(N_COUNTERS, np.uint64)
    cam = np.ascontiguousarray(camera)
    prev = np.ascontiguousarray(previous_camera) if previous_camera is not None else cam
    rc = lib().or_render(scene.handle, C.byref(settings), _ptr(cam), _ptr(prev), sample_begin, sample_count,
                         precision, 1 if brute_force else 0, _ptr(acc), _ptr(aov), _ptr(cnt), threads)
    if rc != 0:
        raise RuntimeError(f"or_render failed ({rc})")
    return OracleResult(acc, aov, cnt)


def hash_u32(v: int) -> int:
    return int(lib().or_hash(v & 0xFFFFFFFF))


def random_floats(seed: int, n: int):
    st = C.c_uint32(seed & 0xFFFFFFFF)
    out = []
    for _ in range(n):
        out.append(float(lib().or_random_float(C.byref(st))))
    return out, int(st.value)
