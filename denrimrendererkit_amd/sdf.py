"""Sparse signed-distance volumes: scene model, host build, GPU encoding.

* ``SDFPrimitive`` / ``build_sparse_volume`` -- DistanceVolumeBuilder.buildSparse
  (Scene/SDFModel.swift:256-355) run by the C++ builder in libdenrim_hip.so
  (spheres as in the reference, plus tori).
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


@dataclass
class SDFPrimitive:
    kind: str            # "sphere" or "torus"
    material: int
    center: Sequence[float]
    radius: float        # sphere radius / torus major radius
    axis: Sequence[float] = (0.0, 1.0, 0.0)
    minor_radius: float = 0.0


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


def build_sparse_volume(primitives: List[SDFPrimitive], dims=(768, 768, 768), brick_size: int = 8, apron: int = 1,
                        narrow_band: float = 2.0 / 767.0, smooth_radius: float = 0.05,
                        bounds_min=(0.0, 0.0, 0.0), bounds_max=(1.0, 1.0, 1.0)) -> SparseDistanceVolume:
    from . import _lib
    L = _lib.load()
    arr = (_lib.SdfPrimitive * len(primitives))()
    for i, p in enumerate(primitives):
        a = arr[i]
        a.kind = 0 if p.kind == "sphere" else 1
        a.material = p.material
        for k in range(3):
            a.center[k] = float(p.center[k])
            a.axis[k] = float(p.axis[k])
        a.radius = float(p.radius)
        a.minor_radius = float(p.minor_radius)
    st = _lib.SdfBuildSettings()
    for k in range(3):
        st.dims[k] = int(dims[k])
        st.bounds_min[k] = float(bounds_min[k])
        st.bounds_max[k] = float(bounds_max[k])
    st.brick_size, st.apron = brick_size, apron
    st.narrow_band, st.smooth_radius = narrow_band, smooth_radius
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
        # Volume descriptor (VolumeGPUEncoding.swift:209-242).
        d = np.zeros((), gt.VOLUME)
        wmn, wmx = _transformed_bounds(v.bounds_min, v.bounds_max, t)
        d["worldBoundsMin"][:3], d["worldBoundsMax"][:3] = wmn, wmx
        d["localBoundsMin"][:3], d["localBoundsMax"][:3] = v.bounds_min, v.bounds_max
        d["dimensions"] = [*v.dimensions, inst.material]
        d["metadata"] = [0, object_id_base + vi, primitive_id_base + vi, 1]
        d["materialProgram"] = [0xFFFFFFFF, 0, 0, 0]
        m = t.matrix.astype(np.float64)
        w2l = np.linalg.inv(m).astype(f32)
        nrm = np.linalg.inv(m.T).astype(f32)
        d["worldToLocal"] = w2l.T   # rows of the stored array are columns
        d["normalTransform"] = nrm.T
        vols.append(d)
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
