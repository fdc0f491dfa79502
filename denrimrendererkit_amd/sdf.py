"""Signed-distance volumes: scene model, host build, GPU encoding.

* ``SDFPrimitive`` -- SDFPrimitive (Scene/SDFModel.swift:4-70): sphere, rounded
  box and capped cylinder with a transform, smooth-union radius and union /
  subtract operation, plus a torus.
* ``build_volume`` / ``build_sparse_volume`` -- DistanceVolumeBuilder.build and
  .buildSparse (Scene/SDFModel.swift:189-355) run by the C++ builder in
  libdenrim_hip.so. ``build_sparse_volume(quantize=True)`` is the FP16 brick
  format of the HIP backend (the sparse-SDF workload); ``quantize=False`` with
  ``apron=2`` is the reference's own format.
* ``DistanceVolume`` -- a dense grid (Scene/DistanceVolume.swift:182-262) and
  ``encode_dense_volumes`` -- appendVolume (Acceleration/VolumeGPUEncoding.swift:84-182).
* ``SparseDistanceVolume`` -- the brick set (Scene/DistanceVolume.swift:309-491).
* ``encode_sparse_volumes`` -- gpuVolumeBricks + appendSparseVolumeDescriptor
  (Acceleration/SparseVolumeGPUEncoding.swift:9-299,
  Acceleration/VolumeGPUEncoding.swift:184-243): brick descriptors with core
  and sample bounds, the brick grid with its 4^3 macro-occupancy grid, and the
  volume descriptor with world-to-local and normal transforms.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import gpu_types as gt

f32 = np.float32
MACRO_SIZE = 4  # SparseVolumeGPUEncoding.swift:7


_KINDS = {"sphere": 0, "torus": 1, "box": 2, "cylinder": 3}


@dataclass
class SDFPrimitive:
    """One primitive of an SDF model. The shape is evaluated at
    inverse(transform) * p - center and its distance scaled by the
    transform's smallest axis scale (CompiledPrimitive, SDFModel.swift:664-713)."""
    kind: str                            # "sphere", "box", "cylinder" or "torus"
    material: int
    center: Sequence[float] = (0.0, 0.0, 0.0)
    radius: float = 0.5                  # sphere / cylinder radius, torus major radius
    axis: Sequence[float] = (0.0, 1.0, 0.0)  # torus axis
    minor_radius: float = 0.0            # torus tube radius
    half_extents: Sequence[float] = (0.5, 0.5, 0.5)  # box
    corner_radius: float = 0.0           # box
    half_height: float = 0.5             # cylinder
    transform: Optional["object"] = None  # scene.Transform (primitive-local to model space)
    smooth_union_radius: Optional[float] = None  # None: the build's default radius
    operation: str = "union"             # "union" or "subtract"


def _primitive_array(primitives):
    from . import _lib
    arr = (_lib.SdfPrimitive * len(primitives))()
    for i, p in enumerate(primitives):
        a = arr[i]
        if p.kind not in _KINDS or p.operation not in ("union", "subtract"):
            from .scene import DenrimRendererError
            raise DenrimRendererError("invalidArgument", f"unknown SDF primitive {p.kind}/{p.operation}")
        a.kind = _KINDS[p.kind]
        a.material = int(p.material)
        for k in range(3):
            a.center[k] = float(p.center[k])
            a.axis[k] = float(p.axis[k])
            a.half_extents[k] = float(p.half_extents[k])
        a.radius, a.minor_radius = float(p.radius), float(p.minor_radius)
        a.corner_radius, a.half_height = float(p.corner_radius), float(p.half_height)
        a.smooth_radius = -1.0 if p.smooth_union_radius is None else float(p.smooth_union_radius)
        a.operation = 0 if p.operation == "union" else 1
        if p.transform is not None:
            m = np.asarray(p.transform.matrix, np.float64)
            inv = np.linalg.inv(m).astype(f32)
            cols = np.asarray(p.transform.matrix, f32)[:3, :3].T  # columns as rows
            scale = max(float(min(np.sqrt((cols.astype(f32) ** 2).sum(axis=1, dtype=f32)))), 1e-6)
            lin_inv = np.linalg.inv(m[:3, :3])
            a.has_transform = 1
            for c in range(4):
                for r in range(4):
                    a.world_to_primitive[c * 4 + r] = float(inv[r, c])
            a.distance_scale = scale
            a.lipschitz = max(1.0, scale * float(np.linalg.norm(lin_inv, 2)) * (1 + 1e-5))
        else:
            a.distance_scale, a.lipschitz = 1.0, 1.0
    return arr


@dataclass
class SparseDistanceVolume:
    dimensions: np.ndarray      # (3,) virtual sample grid
    brick_size: int
    bounds_min: np.ndarray
    bounds_max: np.ndarray
    stored_origin: np.ndarray   # (B, 3) int32
    stored_dims: np.ndarray     # (B, 3)
    core_origin: np.ndarray     # (B, 3)
    core_dims: np.ndarray       # (B, 3)
    sample_offset: np.ndarray   # (B,) into the sample arrays
    distance: np.ndarray        # (S,) float32 (FP16-exact)
    material_a: np.ndarray      # (S,) uint32
    material_b: np.ndarray
    blend: np.ndarray           # (S,) float32
    candidates: int = 0

    @property
    def brick_count(self) -> int:
        return len(self.stored_origin)

    @property
    def grid_dimensions(self) -> np.ndarray:
        return (self.dimensions + self.brick_size - 1) // self.brick_size

    @staticmethod
    def from_bricks(dimensions, brick_size: int, bounds_min, bounds_max, bricks) -> "SparseDistanceVolume":
        """SparseDistanceVolume(dimensions:brickSize:boundsMin:boundsMax:bricks:)
        (Scene/DistanceVolume.swift:309-491) from explicit bricks. Each brick
        is a dict with origin, dimensions, core_origin, core_dimensions,
        distances (x fastest) and material_a, plus optional material_b and
        blend (scalars or per-sample arrays)."""
        so, sd, co, cd, dist, ma, mb, bl = [], [], [], [], [], [], [], []
        for br in bricks:
            dims = np.asarray(br["dimensions"], np.int32)
            n = int(np.prod(dims))
            d = np.asarray(br["distances"], f32).reshape(-1)
            if d.size != n:
                from .scene import DenrimRendererError
                raise DenrimRendererError.invalid_scene("Sparse distance volume brick sample count does not match its dimensions.")
            so.append(br["origin"]); sd.append(dims); co.append(br["core_origin"]); cd.append(br["core_dimensions"])
            dist.append(d)
            ma.append(np.broadcast_to(np.asarray(br["material_a"], np.uint32), (n,)))
            mb.append(np.broadcast_to(np.asarray(br.get("material_b", br["material_a"]), np.uint32), (n,)))
            bl.append(np.broadcast_to(np.asarray(br.get("blend", 0.0), f32), (n,)))
        counts = np.array([int(np.prod(x)) for x in sd], np.int64)
        offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64) if len(counts) else np.zeros(0, np.int64)
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)  # noqa: E731
        arr = lambda xs: np.asarray(xs, np.int32).reshape(-1, 3)  # noqa: E731
        return SparseDistanceVolume(np.asarray(dimensions, np.int64), int(brick_size), np.asarray(bounds_min, f32),
                                    np.asarray(bounds_max, f32), arr(so), arr(sd), arr(co), arr(cd), offsets,
                                    cat(dist, f32), cat(ma, np.uint32), cat(mb, np.uint32), cat(bl, f32))


def _settings(dims, bounds_min, bounds_max, brick_size=8, apron=1, narrow_band=0.0, smooth_radius=0.0,
              quantize=True):
    from . import _lib
    st = _lib.SdfBuildSettings()
    for k in range(3):
        st.dims[k] = int(dims[k])
        st.bounds_min[k] = float(bounds_min[k])
        st.bounds_max[k] = float(bounds_max[k])
    st.brick_size, st.apron = int(brick_size), int(apron)
    st.narrow_band, st.smooth_radius = float(narrow_band), float(smooth_radius)
    st.quantize = 1 if quantize else 0
    return st


def build_sparse_volume(primitives: List[SDFPrimitive], dims=(768, 768, 768), brick_size: int = 8, apron: int = 1,
                        narrow_band: float = 2.0 / 767.0, smooth_radius: float = 0.05,
                        bounds_min=(0.0, 0.0, 0.0), bounds_max=(1.0, 1.0, 1.0),
                        quantize: bool = True) -> SparseDistanceVolume:
    """buildSparse over the virtual grid `dims`. The defaults are the FP16
    brick format (1-sample apron, quantized); the reference's own format is
    apron=2, quantize=False (see build_sparse_reference)."""
    from . import _lib
    L = _lib.load()
    arr = _primitive_array(primitives)
    st = _settings(dims, bounds_min, bounds_max, brick_size, apron, narrow_band, smooth_radius, quantize)
    h = C.c_void_p()
    _lib.check(L.dr_sdf_build_sparse(arr, len(primitives), C.byref(st), C.byref(h)), "sdf build")
    try:
        nb, ns, nc = C.c_uint32(), C.c_uint64(), C.c_uint64()
        L.dr_sdf_volume_info(h, C.byref(nb), C.byref(ns), C.byref(nc))
        B, S = nb.value, ns.value
        so, sd = np.zeros((B, 3), np.int32), np.zeros((B, 3), np.int32)
        co, cd = np.zeros((B, 3), np.int32), np.zeros((B, 3), np.int32)
        dist, bl = np.zeros(S, f32), np.zeros(S, f32)
        ma, mb = np.zeros(S, np.uint32), np.zeros(S, np.uint32)
        p = lambda a: a.ctypes.data if a.size else None  # noqa: E731
        _lib.check(L.dr_sdf_volume_export(h, p(so), p(sd), p(co), p(cd), p(dist), p(ma), p(mb), p(bl)), "sdf export")
    finally:
        L.dr_sdf_volume_free(h)
    counts = np.prod(sd.astype(np.int64), axis=1)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    return SparseDistanceVolume(np.array(dims, np.int64), brick_size, np.array(bounds_min, f32),
                                np.array(bounds_max, f32), so, sd, co, cd, offsets, dist, ma, mb, bl, nc.value)


def build_sparse_reference(primitives: List[SDFPrimitive], dims, bounds_min, bounds_max, brick_size: int = 8,
                           narrow_band: float = 0.25, sample_scale: int = 1) -> SparseDistanceVolume:
    """DistanceVolumeBuilder.buildSparse with the reference's settings
    (SDFModel.swift:256-355): 2-sample overlap, FP32 distances, sampleScale
    multiplies the sample density (dims (d - 1) * s + 1, bricks b * s)."""
    s = max(int(sample_scale), 1)
    d = [max(int(x), 2) for x in dims]
    if s > 1:
        d = [(x - 1) * s + 1 for x in d]
    return build_sparse_volume(primitives, dims=d, brick_size=max(brick_size * s, 1), apron=2,
                               narrow_band=narrow_band, smooth_radius=0.0, bounds_min=bounds_min,
                               bounds_max=bounds_max, quantize=False)


# Material field flags (volumeMaterialField*, PathTrace.metal:255-262).
FIELD_BASE_COLOR, FIELD_OPACITY, FIELD_EMISSION, FIELD_ROUGHNESS = 1, 2, 4, 8
FIELD_METALLIC, FIELD_TRANSMISSION, FIELD_SPECULAR, FIELD_EMISSION_STRENGTH = 16, 32, 64, 128


@dataclass
class DistanceVolume:
    """A dense signed-distance grid (DistanceVolume, Scene/DistanceVolume.swift:182-262):
    samples x fastest over [bounds_min, bounds_max]. Without material arrays
    the instance material covers the volume. ``fields`` (optional) is a
    VOLUME_SAMPLE-shaped array whose baseColorOpacity, emissionTransmission,
    surface and materialFieldFlags columns override material channels."""
    dimensions: np.ndarray
    distances: np.ndarray
    bounds_min: np.ndarray
    bounds_max: np.ndarray
    material_a: Optional[np.ndarray] = None
    material_b: Optional[np.ndarray] = None
    blend: Optional[np.ndarray] = None
    fields: Optional[np.ndarray] = None


@dataclass
class MaterialFields:
    """DistanceVolumeMaterialFields (Scene/DistanceVolume.swift:95-150): the
    channels that are set override the material at hit time."""
    base_color: Optional[Sequence[float]] = None
    opacity: Optional[float] = None
    emission: Optional[Sequence[float]] = None
    roughness: Optional[float] = None
    metallic: Optional[float] = None
    specular: Optional[float] = None
    transmission: Optional[float] = None
    emission_strength: Optional[float] = None

    @property
    def flags(self) -> int:
        f = 0
        for bit, v in ((FIELD_BASE_COLOR, self.base_color), (FIELD_OPACITY, self.opacity),
                       (FIELD_EMISSION, self.emission), (FIELD_ROUGHNESS, self.roughness),
                       (FIELD_METALLIC, self.metallic), (FIELD_TRANSMISSION, self.transmission),
                       (FIELD_SPECULAR, self.specular), (FIELD_EMISSION_STRENGTH, self.emission_strength)):
            if v is not None:
                f |= bit
        return f

    def samples(self, n: int) -> np.ndarray:
        """n VOLUME_SAMPLE records carrying these fields
        (gpuVolumeMaterialFieldSample, VolumeGPUEncoding.swift:341-366)."""
        r = np.zeros(n, gt.VOLUME_SAMPLE)
        r["baseColorOpacity"][:, :3] = self.base_color if self.base_color is not None else 0
        r["baseColorOpacity"][:, 3] = self.opacity or 0
        r["emissionTransmission"][:, :3] = self.emission if self.emission is not None else 0
        r["emissionTransmission"][:, 3] = self.transmission or 0
        r["surface"] = [self.roughness or 0, self.metallic or 0, self.specular or 0, self.emission_strength or 0]
        r["materialFieldFlags"][:, 0] = self.flags
        return r


def sphere_volume(resolution: int, radius: float, bounds_min=(-1.0, -1.0, -1.0),
                  bounds_max=(1.0, 1.0, 1.0)) -> DistanceVolume:
    """DistanceVolume.sphere(resolution:radius:boundsMin:boundsMax:)
    (Scene/DistanceVolume.swift:228-262), float32."""
    r = max(int(resolution), 2)
    bmin, bmax = np.asarray(bounds_min, f32), np.asarray(bounds_max, f32)
    u = np.arange(r, dtype=f32) / f32(r - 1)
    z, y, x = np.meshgrid(u, u, u, indexing="ij")
    ext = (bmax - bmin).astype(f32)
    p = [bmin[0] + ext[0] * x, bmin[1] + ext[1] * y, bmin[2] + ext[2] * z]
    d = (np.sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]) - f32(radius)).astype(f32).reshape(-1)
    return DistanceVolume(np.array([r, r, r], np.int64), d, bmin, bmax)


def build_volume(primitives: List[SDFPrimitive], dims, bounds_min, bounds_max) -> DistanceVolume:
    """DistanceVolumeBuilder.build (SDFModel.swift:189-254): every sample of
    the dense grid, FP32."""
    from . import _lib
    L = _lib.load()
    d = [max(int(x), 2) for x in dims]
    n = d[0] * d[1] * d[2]
    dist, bl = np.zeros(n, f32), np.zeros(n, f32)
    ma, mb = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    st = _settings(d, bounds_min, bounds_max, quantize=False)
    _lib.check(L.dr_sdf_build_dense(_primitive_array(primitives), len(primitives), C.byref(st), dist.ctypes.data,
                                    ma.ctypes.data, mb.ctypes.data, bl.ctypes.data), "sdf dense build")
    return DistanceVolume(np.array(d, np.int64), dist, np.asarray(bounds_min, f32), np.asarray(bounds_max, f32),
                          ma, mb, bl)


@dataclass
class VolumeInstance:
    volume: DistanceVolume
    material: int
    transform: "object"


def _volume_descriptor(dims, bmin, bmax, material, t, sample_offset, object_id, primitive_id, kind):
    d = np.zeros((), gt.VOLUME)
    wmn, wmx = _transformed_bounds(bmin, bmax, t)
    d["worldBoundsMin"][:3], d["worldBoundsMax"][:3] = wmn, wmx
    d["localBoundsMin"][:3], d["localBoundsMax"][:3] = bmin, bmax
    d["dimensions"] = [*dims, material]
    d["metadata"] = [sample_offset, object_id, primitive_id, kind]
    d["materialProgram"] = [0xFFFFFFFF, 0, 0, 0]
    m = t.matrix.astype(np.float64)
    d["worldToLocal"] = np.linalg.inv(m).astype(f32).T   # rows of the stored array are columns
    d["normalTransform"] = np.linalg.inv(m.T).astype(f32).T
    return d


def encode_dense_volumes(instances: List[VolumeInstance], object_id_base: int, material_count: int):
    """appendVolume (Acceleration/VolumeGPUEncoding.swift:84-182): descriptors
    (metadata = sample offset, objectID, primitiveID, 0) and 80-B samples."""
    from .scene import DenrimRendererError
    descs, samples = [], []
    base = 0
    for i, inst in enumerate(instances):
        v = inst.volume
        if inst.material >= material_count or inst.material < 0:
            raise DenrimRendererError.invalid_scene("Distance volume references an unknown material.")
        dims = np.asarray(v.dimensions, np.int64)
        if np.any(dims < 2):
            raise DenrimRendererError.invalid_scene("Distance volume dimensions must be at least 2x2x2.")
        n = int(np.prod(dims))
        if len(v.distances) < n:
            raise DenrimRendererError.invalid_scene("Distance volume sample count does not match its dimensions.")
        if v.material_a is not None and len(v.material_a) < n:
            raise DenrimRendererError.invalid_scene(
                "Distance volume material sample count does not match its dimensions.")
        bmin, bmax = np.asarray(v.bounds_min, f32), np.asarray(v.bounds_max, f32)
        if np.any(bmax <= bmin):
            raise DenrimRendererError.invalid_scene("Distance volume bounds must have positive extent.")
        s = np.zeros(n, gt.VOLUME_SAMPLE)
        s["distance"] = np.asarray(v.distances, f32)[:n]
        if v.material_a is None:
            s["materialA"] = s["materialB"] = inst.material
        else:
            s["materialA"] = v.material_a[:n]
            s["materialB"] = (v.material_b if v.material_b is not None else v.material_a)[:n]
            if v.blend is not None:
                s["materialBlend"] = np.clip(np.asarray(v.blend, f32)[:n], 0, 1)
        if v.fields is not None:
            for k in ("baseColorOpacity", "emissionTransmission", "surface", "materialFieldFlags"):
                s[k] = v.fields[k][:n]
        samples.append(s)
        descs.append(_volume_descriptor(dims, bmin, bmax, inst.material, inst.transform, base, object_id_base + i, i, 0))
        base += n
    vols = np.array(descs, gt.VOLUME) if descs else np.zeros(0, gt.VOLUME)
    return vols, (np.concatenate(samples) if samples else np.zeros(0, gt.VOLUME_SAMPLE))


@dataclass
class SparseVolumeInstance:
    volume: SparseDistanceVolume
    material: int
    transform: "object"  # scene.Transform


def _transformed_bounds(mn, mx, transform):
    corners = np.array([[x, y, z] for x in (mn[0], mx[0]) for y in (mn[1], mx[1]) for z in (mn[2], mx[2])], f32)
    w = transform.transform_point(corners)
    return w.min(axis=0), w.max(axis=0)


def encode_sparse_volumes(instances: List[SparseVolumeInstance], volume_index_base: int, object_id_base: int,
                          primitive_id_base: int, material_count: int):
    """Returns (volumes, bricks, samples, grids, grid_indices) in the reference layouts."""
    from .scene import DenrimRendererError
    vols, bricks, samples, grids, gindex = [], [], [], [], []
    sample_base = 0
    for vi, inst in enumerate(instances):
        v = inst.volume
        if inst.material >= material_count or inst.material < 0:
            raise DenrimRendererError.invalid_scene("Sparse distance volume references an unknown material.")
        if np.any(v.dimensions < 2):
            raise DenrimRendererError.invalid_scene("Sparse distance volume dimensions must be at least 2x2x2.")
        if np.any(v.bounds_max <= v.bounds_min):
            raise DenrimRendererError.invalid_scene("Sparse distance volume bounds must have positive extent.")
        vol_index = volume_index_base + vi
        t = inst.transform
        # Volume descriptor (VolumeGPUEncoding.swift:184-243).
        vols.append(_volume_descriptor(v.dimensions, v.bounds_min, v.bounds_max, inst.material, t, 0,
                                       object_id_base + vi, primitive_id_base + vi, 1))
        # Brick grid (SparseVolumeGPUEncoding.swift:69-107).
        gd = v.grid_dimensions.astype(np.int64)
        md = (gd + MACRO_SIZE - 1) // MACRO_SIZE
        gio = sum(len(x) for x in gindex)
        grid_count = int(np.prod(gd))
        macro_offset = gio + grid_count
        gi = np.full(grid_count, 0xFFFFFFFF, np.uint32)
        mi = np.zeros(int(np.prod(md)), np.uint32)
        g = np.zeros((), gt.BRICK_GRID)
        g["dimensionsAndIndexOffset"] = [*gd, gio]
        g["brickSizeAndVolume"] = [v.brick_size] * 3 + [vol_index]
        g["macroDimensionsAndIndexOffset"] = [*md, macro_offset]
        g["macroSizeAndReserved"] = [MACRO_SIZE] * 3 + [0]
        grids.append(g)
        # Brick descriptors (SparseVolumeGPUEncoding.swift:143-170, 578-655).
        nb = v.brick_count
        b = np.zeros(nb, gt.BRICK)
        ext = (v.bounds_max - v.bounds_min).astype(f32)
        den = np.maximum(v.dimensions - 1, 1).astype(f32)
        co, cd = v.core_origin.astype(f32), v.core_dims.astype(f32)
        core_min_g = np.maximum(co - f32(0.5), f32(0))
        core_max_g = np.minimum((co + cd - f32(1)) + f32(0.5), den)
        core_min = v.bounds_min + ext * (core_min_g / den)
        core_max = v.bounds_min + ext * (core_max_g / den)
        so, sd = v.stored_origin.astype(f32), v.stored_dims.astype(f32)
        smp_min = v.bounds_min + ext * (so / den)
        smp_max = v.bounds_min + ext * ((so + sd - f32(1)) / den)
        counts = np.prod(v.stored_dims.astype(np.int64), axis=1)
        ends = v.sample_offset + counts
        dmin = np.minimum.reduceat(v.distance, v.sample_offset) if nb else np.zeros(0, f32)
        dmax = np.maximum.reduceat(v.distance, v.sample_offset) if nb else np.zeros(0, f32)
        b["localBoundsMin"][:, :3], b["localBoundsMax"][:, :3] = core_min, core_max
        b["sampleBoundsMin"][:, :3], b["sampleBoundsMax"][:, :3] = smp_min, smp_max
        b["sampleBoundsMin"][:, 3], b["sampleBoundsMax"][:, 3] = dmin, dmax
        for k in range(nb):
            wn, wx = _transformed_bounds(core_min[k], core_max[k], t) if not t.is_identity() else (core_min[k], core_max[k])
            b["worldBoundsMin"][k, :3], b["worldBoundsMax"][k, :3] = wn, wx
        b["gridOriginAndVolume"][:, :3] = v.stored_origin
        b["gridOriginAndVolume"][:, 3] = vol_index
        b["dimensionsAndSampleOffset"][:, :3] = v.stored_dims
        b["dimensionsAndSampleOffset"][:, 3] = (sample_base + v.sample_offset).astype(np.uint32)
        brick_base = sum(len(x) for x in bricks)
        cell = v.core_origin // v.brick_size
        ok = np.all((cell >= 0) & (cell < gd), axis=1)
        slot = cell[:, 0] + cell[:, 1] * gd[0] + cell[:, 2] * gd[0] * gd[1]
        gi[slot[ok]] = (brick_base + np.nonzero(ok)[0]).astype(np.uint32)
        mc = cell // MACRO_SIZE
        mslot = mc[:, 0] + mc[:, 1] * md[0] + mc[:, 2] * md[0] * md[1]
        mi[mslot[ok]] = 1
        bricks.append(b)
        gindex.append(gi)
        gindex.append(mi)
        s = np.zeros(len(v.distance), gt.BRICK_SAMPLE)
        s["distance"], s["materialA"], s["materialB"], s["materialBlend"] = v.distance, v.material_a, v.material_b, v.blend
        samples.append(s)
        sample_base += len(v.distance)
        del ends
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)  # noqa: E731
    return (np.array(vols, gt.VOLUME) if vols else np.zeros(0, gt.VOLUME), cat(bricks, gt.BRICK),
            cat(samples, gt.BRICK_SAMPLE), np.array(grids, gt.BRICK_GRID) if grids else np.zeros(0, gt.BRICK_GRID),
            cat(gindex, np.uint32))
