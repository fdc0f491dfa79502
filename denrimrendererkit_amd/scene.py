"""Scene model and scene compile, mirroring the reference's Swift API.

This is the host side above the C ABI. The reference is Swift; no Swift
toolchain exists on this platform, so the public types it exposes to apps are
restated here with the same names, fields, defaults and error behaviour:

* ``Transform``     -- Scene/Transform.swift:5-86
* ``Mesh``          -- Geometry/Mesh.swift:5-186
* ``Material``      -- Scene/Material.swift:16-284
* ``Camera``        -- Scene/Camera.swift:44-218
* ``Environment``   -- Scene/Environment.swift:4-30
* ``Texture2D``     -- Scene/Texture.swift (width, height, RGBA pixels, mode)
* ``QuadLight``     -- Scene/QuadLight.swift:5-50
* ``RenderScene``   -- Scene/RenderScene.swift:73-321

``compile_scene`` restates ``LinearTriangleAccelerationBackend.build``
(Acceleration/AccelerationBackend.swift:7-67) for the triangle path: instance
flattening (Acceleration/InstanceAcceleration.swift:23-128), material and
texture packing (Acceleration/MaterialGPUEncoding.swift:5-50), emissive
light records (Acceleration/LightingGPUEncoding.swift:5-64) and the HDRI
sampling table (Acceleration/LightingGPUEncoding.swift:66-119). It produces
the flat arrays the HIP backend's C ABI consumes. All arithmetic is float32
with an explicit, documented operation order (no fused multiply-add), so the
arrays are deterministic.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import gpu_types as gt

f32 = np.float32


class DenrimRendererError(Exception):
    """Errors of the renderer (DenrimRenderer.swift:5-26)."""

    def __init__(self, kind: str, message: str = ""):
        super().__init__(f"{kind}: {message}" if message else kind)
        self.kind = kind
        self.message = message

    @staticmethod
    def invalid_scene(message: str) -> "DenrimRendererError":
        return DenrimRendererError("invalidScene", message)


# ------------------------------------------------------------ float32 math --
def _v(a) -> np.ndarray:
    return np.asarray(a, dtype=f32)


def _dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """dot(a, b) = (ax*bx + ay*by) + az*bz in float32, last axis 3."""
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.stack([
        a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
        a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0],
    ], axis=-1)


def _normalize(a: np.ndarray) -> np.ndarray:
    """normalize(v) = v * (1 / sqrt(dot(v, v))), float32."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = f32(1.0) / np.sqrt(_dot(a, a))
    return a * inv[..., None]


def _length(a: np.ndarray) -> np.ndarray:
    return np.sqrt(_dot(a, a))


def luminance(v: np.ndarray) -> np.ndarray:
    """dot(v, (0.2126, 0.7152, 0.0722)) (LightingGPUEncoding.swift:121-123)."""
    v = _v(v)
    return (v[..., 0] * f32(0.2126) + v[..., 1] * f32(0.7152)) + v[..., 2] * f32(0.0722)


# --------------------------------------------------------------- Transform --
class Transform:
    """Local-to-world 4x4 transform, column-major like simd_float4x4."""

    def __init__(self, matrix: Optional[np.ndarray] = None):
        m = np.eye(4, dtype=f32) if matrix is None else np.asarray(matrix, dtype=f32).reshape(4, 4)
        self.matrix = m.copy()  # matrix[row, col]

    @staticmethod
    def identity() -> "Transform":
        return Transform()

    @staticmethod
    def translation(v: Sequence[float]) -> "Transform":
        m = np.eye(4, dtype=f32)
        m[0:3, 3] = _v(v)
        return Transform(m)

    @staticmethod
    def scale(v: Sequence[float]) -> "Transform":
        m = np.eye(4, dtype=f32)
        m[0, 0], m[1, 1], m[2, 2] = _v(v)
        return Transform(m)

    @staticmethod
    def rotation_y(radians: float) -> "Transform":
        c, s = f32(math.cos(radians)), f32(math.sin(radians))
        m = np.eye(4, dtype=f32)
        # columns (c,0,-s,0), (0,1,0,0), (s,0,c,0) -- Transform.swift:36-45
        m[0, 0], m[2, 0] = c, -s
        m[0, 2], m[2, 2] = s, c
        return Transform(m)

    @staticmethod
    def rotation_x(radians: float) -> "Transform":
        c, s = f32(math.cos(radians)), f32(math.sin(radians))
        m = np.eye(4, dtype=f32)
        m[1, 1], m[2, 1] = c, s
        m[1, 2], m[2, 2] = -s, c
        return Transform(m)

    @staticmethod
    def rotation_z(radians: float) -> "Transform":
        c, s = f32(math.cos(radians)), f32(math.sin(radians))
        m = np.eye(4, dtype=f32)
        m[0, 0], m[1, 0] = c, s
        m[0, 1], m[1, 1] = -s, c
        return Transform(m)

    def __mul__(self, other: "Transform") -> "Transform":
        return Transform((self.matrix.astype(np.float64) @ other.matrix.astype(np.float64)).astype(f32))

    def is_identity(self) -> bool:
        return bool(np.array_equal(self.matrix, np.eye(4, dtype=f32)))

    def transform_point(self, p: np.ndarray) -> np.ndarray:
        """matrix * (p, 1): ((c0*x + c1*y) + c2*z) + c3 (Transform.swift:77-80)."""
        p = _v(p)
        m = self.matrix
        if self.is_identity():
            return p + f32(0.0)
        return ((m[:3, 0] * p[..., 0:1] + m[:3, 1] * p[..., 1:2]) + m[:3, 2] * p[..., 2:3]) + m[:3, 3]

    def _inverse_transpose(self) -> np.ndarray:
        inv = np.linalg.inv(self.matrix.astype(np.float64))
        return inv.T.astype(f32)

    def transform_normal(self, n: np.ndarray) -> np.ndarray:
        """normalize((M^-1)^T * (n, 0)) (Transform.swift:82-86)."""
        n = _v(n)
        if self.is_identity():
            return _normalize(n)
        d = self._inverse_transpose()
        r = (d[:3, 0] * n[..., 0:1] + d[:3, 1] * n[..., 1:2]) + d[:3, 2] * n[..., 2:3]
        return _normalize(r)


# -------------------------------------------------------------------- Mesh --
class Mesh:
    """Indexed triangle mesh (Mesh.swift:5-86)."""

    def __init__(self, vertices, indices, normals=None, texcoords=None):
        self.vertices = _v(vertices).reshape(-1, 3)
        self.indices = np.asarray(indices, dtype=np.uint32).reshape(-1)
        self.normals = _v(normals).reshape(-1, 3) if normals is not None and len(normals) else np.zeros((0, 3), f32)
        self.texcoords = _v(texcoords).reshape(-1, 2) if texcoords is not None and len(texcoords) else np.zeros((0, 2), f32)

    @property
    def triangle_count(self) -> int:
        return len(self.indices) // 3

    @staticmethod
    def quad(a, b, c, d, texcoords=((0, 0), (1, 0), (1, 1), (0, 1))) -> "Mesh":
        return Mesh([a, b, c, d], [0, 1, 2, 0, 2, 3], texcoords=texcoords)

    @staticmethod
    def box_sized(size) -> "Mesh":
        """Mesh.box(size:) (Mesh.swift:53-60): a box centred at the origin."""
        half = _v(size) * f32(0.5)
        return Mesh.box(-half, half)

    @staticmethod
    def box(minimum, maximum) -> "Mesh":
        """Mesh.box(minimum:maximum:) (Mesh.swift:62-86)."""
        mn, mx = _v(minimum), _v(maximum)
        v = [
            (mn[0], mn[1], mn[2]), (mx[0], mn[1], mn[2]), (mx[0], mx[1], mn[2]), (mn[0], mx[1], mn[2]),
            (mn[0], mn[1], mx[2]), (mx[0], mn[1], mx[2]), (mx[0], mx[1], mx[2]), (mn[0], mx[1], mx[2]),
        ]
        idx = [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 4, 7, 0, 7, 3,
               1, 2, 6, 1, 6, 5, 3, 7, 6, 3, 6, 2, 0, 1, 5, 0, 5, 4]
        return Mesh(v, idx)

    def key(self) -> bytes:
        """Exact bit key used to deduplicate meshes (InstanceAcceleration.swift:153-210)."""
        return b"".join([self.vertices.tobytes(), self.indices.tobytes(), self.normals.tobytes(), self.texcoords.tobytes()])

    def gpu_triangles(self, material: int = 0, transform: Optional[Transform] = None, object_id: int = 0) -> np.ndarray:
        """Mesh.gpuTriangles (Mesh.swift:88-140), vectorised in float32."""
        t = transform or Transform.identity()
        n = self.triangle_count
        out = np.zeros(n, dtype=gt.TRIANGLE)
        if n == 0:
            return out
        i = self.indices.reshape(-1, 3).astype(np.int64)
        p = [self.vertices[i[:, k]] for k in range(3)]
        for k, name in enumerate(("v0", "v1", "v2")):
            out[name][:, :3] = t.transform_point(p[k])
        uv = [self._texcoord(i[:, k]) for k in range(3)]
        local_normal = _normalize(_cross(p[1] - p[0], p[2] - p[0]))
        for k, name in enumerate(("n0", "n1", "n2")):
            out[name][:, :3] = t.transform_normal(self._normal(i[:, k], local_normal))
        for k, name in enumerate(("uv0", "uv1", "uv2")):
            out[name][:, :2] = uv[k]
        tan, bit = _tangent_frame(p, uv, local_normal)
        out["tangent"][:, :3] = t.transform_normal(tan)
        out["bitangent"][:, :3] = t.transform_normal(bit)
        out["materialID"] = material
        out["objectID"] = object_id
        out["primitiveID"] = np.arange(n, dtype=np.uint32)
        return out

    def _texcoord(self, idx: np.ndarray) -> np.ndarray:
        out = np.zeros((len(idx), 2), f32)
        ok = idx < len(self.texcoords)
        out[ok] = self.texcoords[idx[ok]]
        return out

    def _normal(self, idx: np.ndarray, fallback: np.ndarray) -> np.ndarray:
        out = fallback.copy()
        ok = idx < len(self.normals)
        if np.any(ok):
            cand = self.normals[idx[ok]]
            good = _dot(cand, cand) > f32(1e-8)
            sub = out[ok]
            sub[good] = _normalize(cand[good])
            out[ok] = sub
        return out


def _tangent_frame(p, uv, normal):
    """Mesh.tangentFrame (Mesh.swift:161-186)."""
    e1 = p[1] - p[0]
    e2 = p[2] - p[0]
    d1 = uv[1] - uv[0]
    d2 = uv[2] - uv[0]
    den = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    degenerate = np.abs(den) <= f32(1e-6)
    helper = np.where((np.abs(normal[:, 1]) < f32(0.999))[:, None], _v([0, 1, 0]), _v([1, 0, 0]))
    h_t = _normalize(_cross(helper, normal))
    h_b = _cross(normal, h_t)
    with np.errstate(divide="ignore", invalid="ignore"):
        factor = f32(1.0) / den
        tan = _normalize(factor[:, None] * (d2[:, 1:2] * e1 - d1[:, 1:2] * e2))
        bit = _normalize(factor[:, None] * ((-d2[:, 0:1]) * e1 + d1[:, 0:1] * e2))
    tan = np.where(degenerate[:, None], h_t, tan).astype(f32)
    bit = np.where(degenerate[:, None], h_b, bit).astype(f32)
    return tan, bit


# ---------------------------------------------------------------- Texture --
class TextureSamplingMode:
    NEAREST = 0
    LINEAR = 1


@dataclass
class Texture2D:
    width: int
    height: int
    pixels: np.ndarray  # (H*W, 4) float32, row-major, row 0 first
    sampling_mode: int = TextureSamplingMode.NEAREST

    def __post_init__(self):
        self.pixels = _v(self.pixels).reshape(-1, 4)


# --------------------------------------------------------------- Material --
@dataclass
class Material:
    """Material.swift:16-220 (defaults at :137-174)."""
    base_color: Sequence[float] = (0.8, 0.8, 0.8)
    emission: Sequence[float] = (0, 0, 0)
    emission_strength: float = 0.0
    roughness: float = 0.5
    metallic: float = 0.0
    specular: float = 1.0
    specular_color: Sequence[float] = (1, 1, 1)
    index_of_refraction: float = 1.5
    specular_anisotropy: float = 0.0
    clearcoat: float = 0.0
    clearcoat_color: Sequence[float] = (1, 1, 1)
    clearcoat_attenuation_color: Optional[Sequence[float]] = None
    clearcoat_thickness: float = 0.0
    clearcoat_roughness: float = 0.1
    clearcoat_index_of_refraction: float = 1.5
    thin_film: float = 0.0
    thin_film_thickness_nanometers: float = 450.0
    thin_film_index_of_refraction: float = 1.45
    sheen: float = 0.0
    sheen_color: Sequence[float] = (1, 1, 1)
    sheen_roughness: float = 0.5
    subsurface: float = 0.0
    subsurface_color: Optional[Sequence[float]] = None
    subsurface_radius: Sequence[float] = (1, 1, 1)
    subsurface_scale: float = 1.0
    subsurface_anisotropy: float = 0.0
    opacity: float = 1.0
    transmission: float = 0.0
    transmission_color: Optional[Sequence[float]] = None
    transmission_roughness: Optional[float] = None
    transmission_index_of_refraction: Optional[float] = None
    transmission_absorption_color: Sequence[float] = (1, 1, 1)
    transmission_absorption_distance: float = 0.0
    thin_walled: bool = False
    volume_scattering: float = 0.0
    volume_scattering_color: Optional[Sequence[float]] = None
    volume_scattering_distance: float = 1.0
    volume_anisotropy: float = 0.0
    base_color_texture: Optional[Texture2D] = None
    normal_map: Optional[Texture2D] = None

    def gpu_material(self, base_color_texture_index: Optional[int] = None,
                     normal_map_index: Optional[int] = None) -> np.ndarray:
        """Material.gpuMaterial (Material.swift:222-284)."""
        m = np.zeros((), dtype=gt.MATERIAL)
        bc = _v(self.base_color)
        tc = _v(self.transmission_color) if self.transmission_color is not None else bc
        sc = _v(self.subsurface_color) if self.subsurface_color is not None else bc
        vc = (_v(self.volume_scattering_color) if self.volume_scattering_color is not None
              else (_v(self.transmission_color) if self.transmission_color is not None else bc))
        tr = self.transmission_roughness if self.transmission_roughness is not None else self.roughness
        ti = (self.transmission_index_of_refraction if self.transmission_index_of_refraction is not None
              else self.index_of_refraction)
        cca = (_v(self.clearcoat_attenuation_color) if self.clearcoat_attenuation_color is not None
               else _v(self.clearcoat_color))
        m["baseColor"] = [*bc, self.opacity]
        m["emission"] = [*(_v(self.emission) * f32(self.emission_strength)), self.transmission]
        m["parameters"] = [self.roughness, self.metallic,
                           float((base_color_texture_index if base_color_texture_index is not None else -1) + 1),
                           float((normal_map_index if normal_map_index is not None else -1) + 1)]
        m["parameters2"] = [self.specular, self.index_of_refraction, self.clearcoat, self.clearcoat_roughness]
        m["specularColor"] = [*_v(self.specular_color), self.clearcoat_index_of_refraction]
        m["sheenColor"] = [*_v(self.sheen_color), self.sheen]
        m["transmissionColor"] = [*tc, 1.0 if self.thin_walled else 0.0]
        m["parameters3"] = [self.sheen_roughness, tr, ti, self.specular_anisotropy]
        m["clearcoatColor"] = [*_v(self.clearcoat_color), 0.0]
        m["clearcoatAttenuation"] = [*cca, self.clearcoat_thickness]
        m["transmissionAbsorption"] = [*_v(self.transmission_absorption_color), self.transmission_absorption_distance]
        m["thinFilm"] = [self.thin_film, self.thin_film_thickness_nanometers, self.thin_film_index_of_refraction, 0.0]
        m["subsurfaceColor"] = [*sc, self.subsurface]
        m["subsurfaceRadius"] = [*_v(self.subsurface_radius), self.subsurface_scale]
        m["subsurfaceParameters"] = [self.subsurface_anisotropy, 0.0, 0.0, 0.0]
        m["volumeScattering"] = [*vc, self.volume_scattering]
        m["volumeParameters"] = [self.volume_scattering_distance, self.volume_anisotropy, 0.0, 0.0]
        return m


# ----------------------------------------------------------------- Camera --
@dataclass
class CameraLens:
    focus_distance: float = 1.0
    aperture_radius: float = 0.0

    def __post_init__(self):
        self.focus_distance = max(self.focus_distance, 0.0001)
        self.aperture_radius = max(self.aperture_radius, 0.0)


@dataclass
class Camera:
    """Camera.swift:44-218. ``orthographic_scale`` None means perspective."""
    origin: Sequence[float]
    target: Sequence[float]
    up: Sequence[float] = (0, 1, 0)
    vertical_field_of_view_degrees: float = 45.0
    orthographic_scale: Optional[float] = None
    lens: CameraLens = field(default_factory=CameraLens)

    def gpu_camera(self, width: int, height: int) -> np.ndarray:
        """Camera.gpuCamera (Camera.swift:167-200), float32."""
        origin, target, up = _v(self.origin), _v(self.target), _v(self.up)
        aspect = f32(width) / f32(height)
        view = target - origin
        forward = _normalize(view) if _dot(view, view) > f32(1e-8) else _v([0, 0, -1])
        upv = _normalize(up) if _dot(up, up) > f32(1e-8) else _v([0, 1, 0])
        helper = upv if abs(_dot(forward, upv)) < f32(0.999) else _v([1, 0, 0])
        right = _normalize(_cross(forward, helper))
        true_up = _cross(right, forward)
        if self.orthographic_scale is None:
            theta = f32(self.vertical_field_of_view_degrees) * f32(math.pi) / f32(180)
            viewport_height = f32(2) * f32(math.tan(theta / f32(2)))
            flag = 0.0
            center = origin + forward
        else:
            viewport_height = f32(max(self.orthographic_scale, 0.0001))
            flag = 1.0
            center = origin
        viewport_width = aspect * viewport_height
        horizontal = viewport_width * right
        vertical = viewport_height * true_up
        lower_left = (center - horizontal * f32(0.5)) - vertical * f32(0.5)
        g = np.zeros((), dtype=gt.CAMERA)
        g["origin"] = [*origin, flag]
        g["lowerLeft"] = [*lower_left, 0]
        g["horizontal"] = [*horizontal, 0]
        g["vertical"] = [*vertical, 0]
        g["lens"] = [self.lens.aperture_radius, self.lens.focus_distance, 0, 0]
        return g


# ------------------------------------------------------------ Environment --
@dataclass
class Environment:
    """Environment.swift:4-30."""
    texture: Optional[Texture2D] = None
    intensity: float = 1.0
    rotation_y: float = 0.0
    max_radiance: float = 16.0


@dataclass
class QuadLight:
    """QuadLight.swift:5-50."""
    a: Sequence[float]
    b: Sequence[float]
    c: Sequence[float]
    d: Sequence[float]
    color: Sequence[float] = (1, 1, 1)
    intensity: float = 1.0

    @property
    def material(self) -> Material:
        return Material(base_color=self.color, emission=self.color, emission_strength=self.intensity,
                        roughness=0.4, metallic=0.0, specular=0.0)


@dataclass
class MeshInstance:
    mesh: Mesh
    material: int
    transform: Transform


class RenderScene:
    """RenderScene.swift:73-321 (triangle content)."""

    def __init__(self, camera: Optional[Camera] = None, environment: Optional[Environment] = None):
        self.camera = camera or Camera(origin=(0, 1, 4), target=(0, 1, 0))
        self.environment = environment or Environment()
        self.materials: List[Material] = []
        self.mesh_instances: List[MeshInstance] = []
        self.sparse_volume_instances: list = []
        self.volume_instances: list = []
        self.render_defaults = None  # scenescript.RenderDefaults of a parsed script

    def add_material(self, material: Material) -> int:
        self.materials.append(material)
        return len(self.materials) - 1

    def add(self, mesh: Mesh, material: int, transform: Optional[Transform] = None) -> None:
        self.mesh_instances.append(MeshInstance(mesh, material, transform or Transform.identity()))

    def add_volume(self, volume, material: int, transform: Optional[Transform] = None) -> None:
        """add(volume:material:transform:) (RenderScene.swift:140-152): a dense grid."""
        from .sdf import VolumeInstance
        self.volume_instances.append(VolumeInstance(volume, material, transform or Transform.identity()))

    def add_sparse_volume(self, volume, material: int, transform: Optional[Transform] = None) -> None:
        """add(sparseVolume:material:transform:) (RenderScene.swift:154-166)."""
        from .sdf import SparseVolumeInstance
        self.sparse_volume_instances.append(SparseVolumeInstance(volume, material, transform or Transform.identity()))

    def add_quad_light(self, light: QuadLight) -> int:
        mid = self.add_material(light.material)
        self.add(Mesh.quad(light.a, light.b, light.c, light.d), mid)
        return mid

    @staticmethod
    def cornell_box() -> "RenderScene":
        """RenderScene.cornellBox (RenderScene.swift:324-378)."""
        s = RenderScene(camera=Camera(origin=(0, 1, 3.4), target=(0, 1, 0), vertical_field_of_view_degrees=42))
        white = s.add_material(Material(base_color=(0.78, 0.78, 0.74)))
        red = s.add_material(Material(base_color=(0.65, 0.05, 0.04)))
        green = s.add_material(Material(base_color=(0.12, 0.45, 0.15)))
        light = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 0.86, 0.62), emission_strength=14))
        s.add(Mesh.quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)), white)
        s.add(Mesh.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), white)
        s.add(Mesh.quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)), white)
        s.add(Mesh.quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1)), red)
        s.add(Mesh.quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), green)
        s.add(Mesh.quad((-0.28, 1.99, -0.25), (0.28, 1.99, -0.25), (0.28, 1.99, 0.25), (-0.28, 1.99, 0.25)), light)
        return s

    @staticmethod
    def _distance_volume_reference_contents():
        """distanceVolumeReferenceContents (RenderScene.swift:644-768)."""
        from .sdf import SDFPrimitive
        s = RenderScene(camera=Camera(origin=(0, 1.35, 4.8), target=(0, 0.68, 0), vertical_field_of_view_degrees=38))
        floor = s.add_material(Material(base_color=(0.58, 0.60, 0.56), roughness=0.74))
        back = s.add_material(Material(base_color=(0.32, 0.37, 0.43), roughness=0.82))
        red = s.add_material(Material(base_color=(0.92, 0.22, 0.14), roughness=0.62, clearcoat=0.18,
                                      clearcoat_roughness=0.08))
        transparent = s.add_material(Material(base_color=(0.18, 0.56, 0.95), roughness=0.08, opacity=0.42,
                                              transmission=0.82, transmission_color=(0.62, 0.84, 1.0),
                                              transmission_roughness=0.03, transmission_index_of_refraction=1.38,
                                              transmission_absorption_color=(0.68, 0.84, 1.0),
                                              transmission_absorption_distance=0.9))
        metal = s.add_material(Material(base_color=(0.74, 0.86, 0.95), roughness=0.18, metallic=1,
                                        specular_anisotropy=0.45, clearcoat=0.12, clearcoat_roughness=0.04))
        yellow = s.add_material(Material(base_color=(0.95, 0.72, 0.18), roughness=0.38, sheen=0.25,
                                         sheen_color=(1.0, 0.86, 0.42)))
        rear = s.add_material(Material(base_color=(0.18, 0.84, 0.42), emission=(0.18, 0.84, 0.42),
                                       emission_strength=0.7, roughness=0.65))
        light = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 0.9, 0.72), emission_strength=10))
        s.add(Mesh.quad((-3.0, 0, 2.0), (3.0, 0, 2.0), (3.0, 0, -2.0), (-3.0, 0, -2.0)), floor)
        s.add(Mesh.quad((-3.0, 0, -2.0), (3.0, 0, -2.0), (3.0, 2.5, -2.0), (-3.0, 2.5, -2.0)), back)
        s.add(Mesh.quad((-0.8, 2.36, -0.75), (0.8, 2.36, -0.75), (0.8, 2.36, 0.25), (-0.8, 2.36, 0.25)), light)
        s.add(Mesh.quad((-0.66, 0.16, -0.38), (0.36, 0.16, -0.38), (0.36, 1.16, -0.38), (-0.66, 1.16, -0.38)), rear)
        T = Transform
        prims = [
            SDFPrimitive("sphere", red, radius=0.5,
                         transform=T.translation((-1.25, 0.55, 0.1)) * T.scale((1.0, 1.1, 0.9))),
            SDFPrimitive("sphere", transparent, radius=0.5,
                         transform=T.translation((-0.2, 0.48, 0.0)) * T.scale((0.88, 0.96, 0.88)),
                         smooth_union_radius=0.08),
            SDFPrimitive("sphere", metal, radius=0.5,
                         transform=T.translation((0.86, 0.42, 0.06)) * T.scale((0.84, 0.84, 0.84))),
            SDFPrimitive("sphere", yellow, radius=0.5,
                         transform=T.translation((1.55, 0.26, 0.38)) * T.scale((0.52, 0.52, 0.52))),
            SDFPrimitive("sphere", yellow, radius=0.5,
                         transform=T.translation((-1.85, 0.23, 0.62)) * T.scale((0.46, 0.46, 0.46))),
        ]
        settings = dict(dims=(48, 36, 32), bounds_min=(-2.35, -0.02, -0.72), bounds_max=(2.1, 1.2, 0.92))
        return s, prims, settings, red

    @staticmethod
    def distance_volume_reference() -> "RenderScene":
        """RenderScene.distanceVolumeReference (RenderScene.swift:770-779): the
        SDF model built as a dense volume."""
        from .sdf import build_volume
        s, prims, st, red = RenderScene._distance_volume_reference_contents()
        s.add_volume(build_volume(prims, **st), red)
        return s

    @staticmethod
    def sparse_distance_volume_reference() -> "RenderScene":
        """RenderScene.sparseDistanceVolumeReference (RenderScene.swift:782-797):
        the same model as 8^3 bricks with narrow band 0.28 (reference format)."""
        from .sdf import build_sparse_reference
        s, prims, st, red = RenderScene._distance_volume_reference_contents()
        s.add_sparse_volume(build_sparse_reference(prims, st["dims"], st["bounds_min"], st["bounds_max"],
                                                   brick_size=8, narrow_band=0.28), red)
        return s

    @staticmethod
    def material_reference() -> "RenderScene":
        """RenderScene.materialReference (RenderScene.swift:383-447)."""
        s = RenderScene(camera=Camera(origin=(0, 1.65, 5.2), target=(0, 0.75, 0), vertical_field_of_view_degrees=38))
        floor = s.add_material(Material(base_color=(0.62, 0.62, 0.58)))
        back = s.add_material(Material(base_color=(0.38, 0.42, 0.48)))
        warm = s.add_material(Material(base_color=(0.95, 0.42, 0.24), roughness=0.8))
        cool = s.add_material(Material(base_color=(0.18, 0.44, 0.9), roughness=0.18, metallic=1))
        green = s.add_material(Material(base_color=(0.18, 0.72, 0.35), roughness=0.58, sheen=0.65,
                                        sheen_color=(0.72, 1.0, 0.82), sheen_roughness=0.72))
        violet = s.add_material(Material(base_color=(0.62, 0.36, 0.95), roughness=0.35, metallic=0.55,
                                         specular_color=(0.85, 0.72, 1.0), index_of_refraction=1.65, clearcoat=0.35,
                                         clearcoat_roughness=0.06, clearcoat_index_of_refraction=1.55))
        bright = s.add_material(Material(base_color=(0.95, 0.86, 0.42), roughness=0.08, metallic=1))
        light = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 0.9, 0.72), emission_strength=10))
        s.add(Mesh.quad((-3, 0, 2.2), (3, 0, 2.2), (3, 0, -2.2), (-3, 0, -2.2)), floor)
        s.add(Mesh.quad((-3, 0, -2.2), (3, 0, -2.2), (3, 2.8, -2.2), (-3, 2.8, -2.2)), back)
        s.add(Mesh.quad((-0.8, 2.75, -0.6), (0.8, 2.75, -0.6), (0.8, 2.75, 0.4), (-0.8, 2.75, 0.4)), light)
        box = Mesh.box_sized((0.58, 0.58, 0.58))
        placements = [(warm, (-1.6, 0.29, 0.15)), (cool, (-0.8, 0.29, -0.1)), (green, (0, 0.29, 0.1)),
                      (violet, (0.8, 0.29, -0.05)), (bright, (1.6, 0.29, 0.12))]
        for index, (mat, pos) in enumerate(placements):
            t = Transform.translation(pos) * Transform.rotation_y(float(index) * 0.28 - 0.45)
            s.add(box, mat, t)
        return s


# ---------------------------------------------------------- scene compile --
@dataclass
class AccelerationBuild:
    """The flat arrays of AccelerationBuild (AccelerationBuild.swift:4-44), triangle path."""
    triangles: np.ndarray
    materials: np.ndarray
    texture_descriptors: np.ndarray
    texture_pixels: np.ndarray
    environment_texture_index_plus_one: int
    environment_samples: np.ndarray
    lights: np.ndarray
    environment_intensity: float = 1.0
    environment_rotation_y: float = 0.0
    environment_max_radiance: float = 16.0
    volumes: Optional[np.ndarray] = None
    volume_bricks: Optional[np.ndarray] = None
    volume_brick_samples: Optional[np.ndarray] = None
    volume_brick_grids: Optional[np.ndarray] = None
    volume_brick_grid_indices: Optional[np.ndarray] = None
    volume_samples: Optional[np.ndarray] = None  # dense GPUVolumeSample records

    def __post_init__(self):
        from .gpu_types import BRICK, BRICK_GRID, BRICK_SAMPLE, VOLUME, VOLUME_SAMPLE
        if self.volumes is None:
            self.volumes = np.zeros(0, VOLUME)
        if self.volume_bricks is None:
            self.volume_bricks = np.zeros(0, BRICK)
            self.volume_brick_samples = np.zeros(0, BRICK_SAMPLE)
            self.volume_brick_grids = np.zeros(0, BRICK_GRID)
            self.volume_brick_grid_indices = np.zeros(0, np.uint32)
        if self.volume_samples is None:
            self.volume_samples = np.zeros(0, VOLUME_SAMPLE)


def _materials_and_textures(scene: RenderScene):
    """gpuMaterialsAndTextures (MaterialGPUEncoding.swift:5-50)."""
    descs: List[tuple] = []
    pixels: List[np.ndarray] = []
    count = [0]

    def append(tex: Optional[Texture2D]) -> Optional[int]:
        if tex is None or tex.width <= 0 or tex.height <= 0:
            return None
        need = tex.width * tex.height
        if len(tex.pixels) < need:
            return None
        index = len(descs)
        descs.append((count[0], tex.width, tex.height, tex.sampling_mode))
        pixels.append(tex.pixels[:need])
        count[0] += need
        return index

    mats = np.zeros(len(scene.materials), dtype=gt.MATERIAL)
    for k, m in enumerate(scene.materials):
        bi = append(m.base_color_texture)
        ni = append(m.normal_map)
        mats[k] = m.gpu_material(bi, ni)
    env_index = append(scene.environment.texture)
    d = np.zeros(len(descs), dtype=gt.TEXTURE_DESC)
    for k, row in enumerate(descs):
        d[k] = row
    px = np.concatenate(pixels).astype(f32) if pixels else np.zeros((0, 4), f32)
    return mats, d, px, (env_index + 1 if env_index is not None else 0)


def materialize_triangles(scene: RenderScene) -> np.ndarray:
    """InstanceAccelerationBuilder.build + materializedTriangles
    (InstanceAcceleration.swift:23-34, 74-128): meshes are deduplicated by
    exact key, local triangles are built once, then transformed per
    instance with objectID = instance index."""
    local_by_key = {}
    parts = []
    for instance_index, inst in enumerate(scene.mesh_instances):
        if inst.material >= len(scene.materials) or inst.material < 0:
            raise DenrimRendererError.invalid_scene("Mesh references an unknown material.")
        key = inst.mesh.key()
        local = local_by_key.get(key)
        if local is None:
            local = inst.mesh.gpu_triangles(0, Transform.identity(), 0)
            local_by_key[key] = local
        parts.append(_transform_triangles(local, inst.transform, inst.material, instance_index))
    if not parts:
        return np.zeros(0, dtype=gt.TRIANGLE)
    return np.concatenate(parts)


def _transform_triangles(local: np.ndarray, t: Transform, material: int, object_id: int) -> np.ndarray:
    """InstanceAcceleration.transform (InstanceAcceleration.swift:36-71)."""
    out = np.zeros(len(local), dtype=gt.TRIANGLE)
    for name in ("v0", "v1", "v2"):
        out[name][:, :3] = t.transform_point(local[name][:, :3])
    for name in ("n0", "n1", "n2", "tangent", "bitangent"):
        out[name][:, :3] = t.transform_normal(local[name][:, :3])
    for name in ("uv0", "uv1", "uv2"):
        out[name] = local[name]
    out["materialID"] = material
    out["objectID"] = object_id
    out["primitiveID"] = local["primitiveID"]
    out["padding2"] = 0
    return out


def light_records(triangles: np.ndarray, materials: np.ndarray):
    """lightRecordsAndTaggedTriangles (LightingGPUEncoding.swift:5-64).
    Tags triangles in place (padding2 = light index + 1) and returns records."""
    if len(materials) == 0 or len(triangles) == 0:
        return np.zeros(0, dtype=gt.LIGHT)
    mid = np.minimum(triangles["materialID"].astype(np.int64), len(materials) - 1)
    em = materials["emission"][:, :3]
    emissive_mat = np.max(em, axis=1) > 0
    cand = np.nonzero(emissive_mat[mid])[0]
    if len(cand) == 0:
        return np.zeros(0, dtype=gt.LIGHT)
    tri = triangles[cand]
    a = tri["v1"][:, :3] - tri["v0"][:, :3]
    b = tri["v2"][:, :3] - tri["v0"][:, :3]
    c = _cross(a, b)
    area = _length(c) * f32(0.5)
    normal = _normalize(c)
    power = luminance(em[mid[cand]]) * area
    keep = (area > 0) & (power > 0)
    cand, area, normal, power = cand[keep], area[keep], normal[keep], power[keep]
    if len(cand) == 0:
        return np.zeros(0, dtype=gt.LIGHT)
    cum = np.cumsum(power.astype(f32), dtype=f32)  # sequential float32 sums
    total = cum[-1]
    if not total > 0:
        return np.zeros(0, dtype=gt.LIGHT)
    cdf = np.minimum(cum / total, f32(1)).astype(f32)
    cdf[-1] = f32(1)
    lights = np.zeros(len(cand), dtype=gt.LIGHT)
    lights["triangleIndex"] = cand.astype(np.uint32)
    lights["materialIndex"] = mid[cand].astype(np.uint32)
    lights["area"] = area
    lights["selectionCDF"] = cdf
    lights["normal"][:, :3] = normal
    triangles["padding2"][cand] = np.arange(1, len(cand) + 1, dtype=np.uint32)
    return lights


def environment_samples(env: Environment) -> np.ndarray:
    """environmentSamples (LightingGPUEncoding.swift:66-119), float32."""
    tex = env.texture
    if tex is None or tex.width <= 0 or tex.height <= 0 or len(tex.pixels) < tex.width * tex.height:
        return np.zeros(0, dtype=gt.ENV_SAMPLE)
    w, h = tex.width, tex.height
    intensity = f32(max(env.intensity, 0.0))
    max_radiance = f32(max(env.max_radiance, 0.0))
    d_theta = f32(math.pi) / f32(h)
    d_phi = f32(2) * f32(math.pi) / f32(w)
    theta = (np.arange(h, dtype=f32) + f32(0.5)) * d_theta
    solid = np.maximum(np.sin(theta).astype(f32) * d_theta * d_phi, f32(1e-8)).astype(f32)
    color = tex.pixels[: w * h, :3] * intensity
    if max_radiance > 0:
        color = np.minimum(color, max_radiance)
    weight = (np.maximum(luminance(color), f32(0)).reshape(h, w) * solid[:, None]).reshape(-1).astype(f32)
    cum = np.cumsum(weight, dtype=f32)
    total = cum[-1]
    if not total > 0:
        return np.zeros(0, dtype=gt.ENV_SAMPLE)
    out = np.zeros(w * h, dtype=gt.ENV_SAMPLE)
    out["cdf"] = np.minimum(cum / total, f32(1))
    prob = weight / total
    out["pdf"] = (prob.reshape(h, w) / solid[:, None]).reshape(-1)
    return out


def compile_scene(scene: RenderScene) -> AccelerationBuild:
    """LinearTriangleAccelerationBackend.build (AccelerationBackend.swift:7-67):
    triangles, then dense volumes, then sparse volumes (gpuVolumes,
    VolumeGPUEncoding.swift:5-65, with its objectID / primitiveID bases)."""
    tris = materialize_triangles(scene)
    from .gpu_types import VOLUME
    from .sdf import encode_dense_volumes, encode_sparse_volumes
    n_mesh, n_dense = len(scene.mesh_instances), len(scene.volume_instances)
    dense, dense_samples = encode_dense_volumes(scene.volume_instances, n_mesh, len(scene.materials))
    sparse = (None,) * 5
    if scene.sparse_volume_instances:
        sparse = encode_sparse_volumes(scene.sparse_volume_instances, n_dense, n_mesh + n_dense, n_dense,
                                       len(scene.materials))
    vol_list = [v for v in (dense, sparse[0]) if v is not None and len(v)]
    volumes = np.concatenate(vol_list) if vol_list else np.zeros(0, VOLUME)
    if len(tris) == 0 and len(volumes) == 0:  # RenderSession.swift:322-324
        raise DenrimRendererError.invalid_scene("Scene contains no renderable geometry.")
    if not scene.materials:  # RenderSession.swift:325-327
        raise DenrimRendererError.invalid_scene("Scene contains no materials.")
    mats, descs, pixels, env_index = _materials_and_textures(scene)
    lights = light_records(tris, mats)
    env = environment_samples(scene.environment)
    return AccelerationBuild(
        triangles=tris, materials=mats, texture_descriptors=descs, texture_pixels=pixels,
        environment_texture_index_plus_one=env_index, environment_samples=env, lights=lights,
        environment_intensity=float(scene.environment.intensity),
        environment_rotation_y=float(scene.environment.rotation_y),
        environment_max_radiance=float(scene.environment.max_radiance),
        volumes=volumes, volume_bricks=sparse[1], volume_brick_samples=sparse[2], volume_brick_grids=sparse[3],
        volume_brick_grid_indices=sparse[4], volume_samples=dense_samples,
    )
