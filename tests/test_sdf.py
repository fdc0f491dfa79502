"""CPU tests of the sparse signed-distance path: the host brick builder and
encoding, and the oracle's brick tracing (no GPU).

Pins:
  * builder: brute force over the whole virtual grid (every sample of a small
    grid evaluated densely, in the builder's own float32 arithmetic, decides
    which bricks the narrow band keeps and what each stored FP16 sample is);
  * oracle tracing (closestSparseVolumeBrickDDAHit and
    intersectVolumeBrickPrepared, PathTrace.metal:2608-3088): closed-form
    ray/sphere hits and normals for untransformed and transformed instances;
  * counters and culling: the reference's own assertions in
    RenderSessionTests.swift:806-860, :862-919 and :1069-1120.
"""
import numpy as np
import pytest

import oracle as O
from denrimrendererkit_amd import Camera, Material, RenderScene, Transform, compile_scene
from denrimrendererkit_amd.sdf import SDFPrimitive, SparseDistanceVolume, build_sparse_volume
from denrimrendererkit_amd.scene import DenrimRendererError

f32 = np.float32
SDF = O.CNT_SDF  # counter base: + reference slot index (2 grid cells ... 9 brick hits)


def sdf_stats(counters):
    c = [int(x) for x in counters]
    cells, tests = c[SDF + 2], c[SDF + 4]
    return dict(cells=cells, empty=cells - min(cells, tests), macro=c[SDF + 3], tests=tests, invalid=c[SDF + 5],
                culls=c[SDF + 6], marches=c[SDF + 7], steps=c[SDF + 8], hits=c[SDF + 9])


# ----------------------------------------------------------------- builder --
def _dense_sphere_field(dims, bmin, bmax, c, r):
    ext = (np.asarray(bmax, f32) - np.asarray(bmin, f32)).astype(f32)
    axes = [f32(bmin[a]) + ext[a] * (np.arange(dims[a], dtype=f32) / f32(dims[a] - 1)) for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    q = [x - f32(c[0]), y - f32(c[1]), z - f32(c[2])]
    return (np.sqrt((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) - f32(r)).astype(f32)  # [z, y, x]


@pytest.mark.parametrize("dims,brick,apron", [((24, 24, 24), 8, 1), ((29, 20, 23), 5, 2), ((16, 16, 16), 4, 0)])
def test_builder_matches_dense_brute_force(dims, brick, apron):
    c, r = (0.1, -0.05, 0.02), 0.55
    bmin, bmax = (-0.8, -0.8, -0.8), (0.8, 0.8, 0.8)
    band = 0.13
    vol = build_sparse_volume([SDFPrimitive("sphere", 0, c, r)], dims=dims, brick_size=brick, apron=apron,
                              narrow_band=band, smooth_radius=0.05, bounds_min=bmin, bounds_max=bmax)
    dense = _dense_sphere_field(dims, bmin, bmax, c, r).astype(np.float16).astype(f32)
    grid = [(d + brick - 1) // brick for d in dims]
    expect = []
    for bz in range(grid[2]):
        for by in range(grid[1]):
            for bx in range(grid[0]):
                o = np.array([bx, by, bz]) * brick
                lo = np.maximum(o - apron, 0)
                hi = np.minimum(o + brick + apron, dims)
                block = dense[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
                if block.min() <= band and block.max() >= -band:
                    expect.append((tuple(o), tuple(lo), tuple(hi - lo), block.reshape(-1)))
    assert vol.brick_count == len(expect)
    for i, (o, lo, n, samples) in enumerate(expect):
        assert tuple(vol.core_origin[i]) == o
        assert tuple(vol.stored_origin[i]) == lo
        assert tuple(vol.stored_dims[i]) == n
        off = int(vol.sample_offset[i])
        got = vol.distance[off:off + len(samples)]
        np.testing.assert_array_equal(got, samples)  # FP16 round-to-nearest-even of the float32 distance
    assert np.all(vol.material_a == 0) and np.all(vol.blend == 0)


def test_builder_smooth_union_blends_materials():
    k = 0.05
    prims = [SDFPrimitive("sphere", 3, (0.4, 0.5, 0.5), 0.12), SDFPrimitive("sphere", 7, (0.62, 0.5, 0.5), 0.12)]
    vol = build_sparse_volume(prims, dims=(40, 40, 40), brick_size=8, apron=1, narrow_band=2 / 39, smooth_radius=k)
    # Between the spheres the smooth-min lies below both distances by at most k/4
    # and the material pair is (3, 7) with a blend weight in (0, 1).
    ext = np.float64(1.0) / 39
    mid = []
    for i in range(vol.brick_count):
        so, sd, off = vol.stored_origin[i], vol.stored_dims[i], int(vol.sample_offset[i])
        zz, yy, xx = np.meshgrid(*[np.arange(so[a], so[a] + sd[a]) for a in (2, 1, 0)], indexing="ij")
        p = np.stack([xx, yy, zz], -1).reshape(-1, 3) * ext
        d1 = np.linalg.norm(p - prims[0].center, axis=1) - 0.12
        d2 = np.linalg.norm(p - prims[1].center, axis=1) - 0.12
        got = vol.distance[off:off + len(p)].astype(np.float64)
        m = np.minimum(d1, d2)
        assert np.all(got <= m + 2e-3) and np.all(got >= m - k / 4 - 2e-3)
        sel = (np.abs(d1 - d2) < 0.5 * k)
        mid.append((vol.material_a[off:off + len(p)][sel], vol.material_b[off:off + len(p)][sel],
                    vol.blend[off:off + len(p)][sel]))
    a = np.concatenate([x[0] for x in mid])
    b = np.concatenate([x[1] for x in mid])
    w = np.concatenate([x[2] for x in mid])
    assert len(a) > 0 and np.all(a == 3) and np.all(b == 7) and np.all((w > 0) & (w < 1))
    # Blend weights are multiples of 1/65535.
    q = np.round(w.astype(np.float64) * 65535)
    np.testing.assert_array_equal((q / 65535).astype(f32), w)


def test_builder_rejects_bad_settings():
    with pytest.raises(DenrimRendererError):
        build_sparse_volume([], dims=(8, 8, 8))
    with pytest.raises(DenrimRendererError):
        build_sparse_volume([SDFPrimitive("sphere", 0, (0, 0, 0), 0.2)], dims=(8, 8, 8), bounds_min=(0, 0, 0),
                            bounds_max=(1, 0, 1))


# ---------------------------------------------------------------- tracing --
def _sphere_scene(res, transform=None, c=(0.5, 0.5, 0.5), r=0.3):
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0)))
    m = s.add_material(Material(base_color=(0.2, 0.5, 0.8)))
    vol = build_sparse_volume([SDFPrimitive("sphere", m, c, r)], dims=(res,) * 3, narrow_band=2.0 / (res - 1))
    s.add_sparse_volume(vol, m, transform)
    return s


def _rays_at_sphere(rng, n, C, R):
    o = rng.normal(size=(n, 3))
    o = C + 3.0 * R * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = C + rng.uniform(-0.7, 0.7, size=(n, 3)) * R - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), f32)
    rays[:, :3], rays[:, 4:7] = o, d
    return rays


def _analytic_sphere(rays, C, R):
    o, d = rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    oc = o - C
    b = (oc * d).sum(1)
    disc = b * b - ((oc * oc).sum(1) - R * R)
    t = -b - np.sqrt(np.maximum(disc, 0))
    return t, (o + t[:, None] * d - C) / R, disc


@pytest.mark.parametrize("res", [48, 96])
@pytest.mark.parametrize("transformed", [False, True])
def test_oracle_sparse_hits_match_closed_form_sphere(res, transformed):
    xf = (Transform.translation((0.3, -0.2, 0.1)) * Transform.rotation_y(0.7) * Transform.scale((1.5, 1.5, 1.5))
          if transformed else None)
    b = compile_scene(_sphere_scene(res, xf))
    sc = O.OracleScene(b)
    C, R = np.array([0.5, 0.5, 0.5]), 0.3
    if xf is not None:
        C, R = xf.transform_point(C[None].astype(f32))[0].astype(np.float64), R * 1.5
    rays = _rays_at_sphere(np.random.default_rng(res), 3000, C, R)
    t, ids, hit = sc.intersect(rays, with_hit=True)
    ta, na, disc = _analytic_sphere(rays, C, R)
    h = ids != 0xFFFFFFFF
    clear = disc > (0.05 * R) ** 2
    assert np.all(h[clear])  # every clearly hitting ray hits
    voxel = (1.0 / (res - 1)) * (R / 0.3)
    dt = np.abs(t[h] - ta[h])
    assert np.quantile(dt, 0.99) < 0.05 * voxel and dt.max() < 0.5 * voxel
    dn = np.linalg.norm(hit[h, :3] - na[h], axis=1)
    rel = (1.0 / (res - 1)) / 0.3  # sample spacing over radius, in the volume's own frame
    # Cubic gradients inside a brick, trilinear ones next to its border.
    assert np.median(dn) < 0.05 * rel and np.quantile(dn, 0.99) < 0.8 * rel and dn.max() < 1.0 * rel
    assert np.all(hit[h, 7] == 1)  # front facing from outside
    assert np.all(hit[h, 6] == 0)  # objectID: first instance


def test_oracle_volume_only_scene_counters():
    # RenderSessionTests.swift:806-860: a sparse-volume-only scene renders
    # through the brick path; cells >= empty + brick tests; marches, steps,
    # hits > 0; albedo shows the volume's material.
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), orthographic_scale=2.4))
    m = s.add_material(Material(base_color=(0.2, 0.65, 0.9), roughness=0.55))
    vol = build_sparse_volume([SDFPrimitive("sphere", m, (0, 0, 0), 0.55)], dims=(20, 20, 20), brick_size=5,
                              narrow_band=0.3, bounds_min=(-0.8, -0.8, -0.8), bounds_max=(0.8, 0.8, 0.8))
    s.add_sparse_volume(vol, m)
    b = compile_scene(s)
    assert len(b.triangles) == 0
    r = O.render(O.OracleScene(b), O.make_settings(b, 20, 20, max_bounces=1), s.camera.gpu_camera(20, 20))
    st = sdf_stats(r.counters)
    assert st["cells"] > 0 and st["tests"] > 0 and st["cells"] >= st["empty"] + st["tests"]
    assert st["marches"] > 0 and st["steps"] > 0 and st["hits"] > 0
    alb = r.aovs[2]
    assert np.any((alb[..., 0] < 0.35) & (alb[..., 1] > 0.45) & (alb[..., 2] > 0.65) & (alb[..., 3] > 0.9))


def no_crossing_scene():
    # RenderSessionTests.swift:862-919: one 3x3x3 brick whose samples are all
    # 0.05 has no zero crossing and is culled before any march.
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), orthographic_scale=2.0))
    m = s.add_material(Material(base_color=(0.8, 0.2, 0.2)))
    vol = SparseDistanceVolume.from_bricks((3, 3, 3), 3, (-1, -1, -1), (1, 1, 1), [dict(
        origin=(0, 0, 0), dimensions=(3, 3, 3), core_origin=(0, 0, 0), core_dimensions=(3, 3, 3),
        distances=np.full(27, 0.05, f32), material_a=m)])
    s.add_sparse_volume(vol, m)
    return s


def test_oracle_brick_without_zero_crossing_is_culled_before_march():
    s = no_crossing_scene()
    b = compile_scene(s)
    r = O.render(O.OracleScene(b), O.make_settings(b, 8, 8, max_bounces=1), s.camera.gpu_camera(8, 8))
    st = sdf_stats(r.counters)
    assert st["cells"] > 0 and st["tests"] > 0 and st["cells"] >= st["empty"] + st["tests"]
    assert st["culls"] > 0 and st["marches"] == 0 and st["steps"] == 0 and st["hits"] == 0


def macro_skip_scene():
    # RenderSessionTests.swift:1069-1120: a small sphere off-centre in a large
    # grid; rays crossing empty 4^3 macro cells skip them.
    s = RenderScene(camera=Camera(origin=(0.6, 0, 3), target=(0.6, 0, 0), orthographic_scale=2.4))
    m = s.add_material(Material(base_color=(0.9, 0.45, 0.2), roughness=0.55))
    vol = build_sparse_volume([SDFPrimitive("sphere", m, (0.6, 0, 0), 0.35)], dims=(48, 48, 48), brick_size=4,
                              narrow_band=0.18, bounds_min=(-1.5, -1.5, -1.5), bounds_max=(1.5, 1.5, 1.5))
    s.add_sparse_volume(vol, m)
    return s


def test_oracle_macro_grid_skips():
    s = macro_skip_scene()
    b = compile_scene(s)
    r = O.render(O.OracleScene(b), O.make_settings(b, 32, 32, max_bounces=1), s.camera.gpu_camera(32, 32))
    st = sdf_stats(r.counters)
    assert st["macro"] > 0 and st["cells"] > 0 and st["tests"] > 0


def quality_probe(res, quality, render):
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), orthographic_scale=1.7))
    m = s.add_material(Material(base_color=(0.9, 0.56, 0.22), roughness=0.42))
    prims = [SDFPrimitive("sphere", m, (0, 0, 0), 0.62),
             SDFPrimitive("torus", m, (0.05, 0.1, 0.3), 0.45, axis=(0.2, 1.0, 0.3), minor_radius=0.1)]
    vol = build_sparse_volume(prims, dims=(res,) * 3, brick_size=7, narrow_band=0.32, smooth_radius=0.05,
                              bounds_min=(-0.9, -0.9, -0.9), bounds_max=(0.9, 0.9, 0.9))
    s.add_sparse_volume(vol, m)
    return render(s, 44, 44, quality)


def aov_metrics(a, b):
    """primaryAOVDifferenceMetrics (RenderSessionTests.swift:2151-2210)."""
    (na, da), (nb, db) = a, b
    ha, hb = na[..., 3] > 0.5, nb[..., 3] > 0.5
    both = ha & hb
    va, vb = na[..., :3] * 2 - 1, nb[..., :3] * 2 - 1
    n_err = ((va - vb) ** 2).sum(-1)[both]
    d_err = ((da[..., 0] - db[..., 0]) ** 2)[both]
    cnt = max(1, both.sum())
    return dict(coverage=both.sum() / ha.size, normal_rmse=float(np.sqrt(n_err.sum() / cnt)),
                depth_rmse=float(np.sqrt(d_err.sum() / cnt)),
                silhouette=float((ha != hb).sum() / max(1, (ha | hb).sum())))


def _oracle_probe(scene, w, h, quality):
    b = compile_scene(scene)
    r = O.render(O.OracleScene(b), O.make_settings(b, w, h, max_bounces=1, quality=quality),
                 scene.camera.gpu_camera(w, h))
    return r.aovs[1], r.aovs[0]


def test_oracle_sparse_quality_close_to_denser_reference():
    # RenderSessionTests.swift:1024-1067 (thresholds as there; the shapes are
    # a sphere and a torus, which the brick builder here supports).
    low = quality_probe(28, 2, _oracle_probe)
    ref = quality_probe(64, 2, _oracle_probe)
    m = aov_metrics(low, ref)
    assert m["coverage"] > 0.20 and m["normal_rmse"] < 0.30 and m["depth_rmse"] < 0.05 and m["silhouette"] < 0.02


# ------------------------------------------------------ model and dense --
def _np_primitive_distance(p, prim, xf):
    """CompiledPrimitive.distance (SDFModel.swift:685-713) in float32 numpy."""
    if xf is not None:
        inv = np.linalg.inv(xf.matrix.astype(np.float64)).astype(f32)
        l = [((inv[r, 0] * p[..., 0] + inv[r, 1] * p[..., 1]) + inv[r, 2] * p[..., 2]) + inv[r, 3] for r in range(3)]
        l = np.stack(l, -1).astype(f32)
        cols = xf.matrix[:3, :3].T.astype(f32)
        scale = f32(max(float(np.sqrt((cols * cols).sum(1)).min()), 1e-6))
    else:
        l, scale = p.astype(f32), f32(1)
    if prim.kind == "sphere":
        d = np.sqrt((l * l).sum(-1)) - f32(prim.radius)
    elif prim.kind == "box":
        q = np.abs(l) - np.asarray(prim.half_extents, f32)
        d = (np.sqrt((np.maximum(q, 0) ** 2).sum(-1)) + np.minimum(q.max(-1), 0) - f32(max(prim.corner_radius, 0)))
    else:  # cylinder
        dx = np.sqrt(l[..., 0] ** 2 + l[..., 2] ** 2) - f32(prim.radius)
        dy = np.abs(l[..., 1]) - f32(prim.half_height)
        d = np.minimum(np.maximum(dx, dy), 0) + np.sqrt(np.maximum(dx, 0) ** 2 + np.maximum(dy, 0) ** 2)
    return (d * scale).astype(f32)


def test_dense_builder_shapes_transforms_and_subtract():
    # DistanceVolumeBuilder.build over a box, a capped cylinder and a
    # non-uniformly scaled, rotated sphere, then a subtracted cylinder
    # (sampleField, SDFModel.swift:480-505), against numpy on every sample.
    from denrimrendererkit_amd.sdf import build_volume
    T = Transform
    box = SDFPrimitive("box", 1, half_extents=(0.3, 0.2, 0.25), corner_radius=0.03,
                       transform=T.translation((-0.3, 0.0, 0.1)) * T.rotation_z(0.3))
    cyl = SDFPrimitive("cylinder", 2, radius=0.18, half_height=0.3, transform=T.translation((0.35, 0.05, 0.0)))
    sph = SDFPrimitive("sphere", 3, radius=0.25,
                       transform=T.translation((0.0, -0.3, -0.2)) * T.rotation_y(0.5) * T.scale((1.2, 0.8, 1.0)))
    cut = SDFPrimitive("cylinder", 4, radius=0.08, half_height=0.9, operation="subtract",
                       transform=T.translation((-0.3, 0.0, 0.1)) * T.rotation_x(1.5707964))
    dims, bmin, bmax = (21, 18, 19), (-0.8, -0.7, -0.6), (0.8, 0.7, 0.6)
    vol = build_volume([box, cyl, sph, cut], dims, bmin, bmax)
    ext = np.asarray(bmax, f32) - np.asarray(bmin, f32)
    axes = [f32(bmin[a]) + ext[a] * (np.arange(dims[a], dtype=f32) / f32(dims[a] - 1)) for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    p = np.stack([x, y, z], -1).reshape(-1, 3)
    ds = [_np_primitive_distance(p, q, q.transform) for q in (box, cyl, sph, cut)]
    union = np.minimum(np.minimum(ds[0], ds[1]), ds[2])
    expect = np.maximum(union, -ds[3])
    np.testing.assert_allclose(vol.distances, expect, atol=3e-6, rtol=0)
    owner = np.argmin(np.stack(ds[:3]), 0) + 1
    assert np.mean(vol.material_a == owner) > 0.99 and np.all(vol.blend == 0)


@pytest.mark.parametrize("res", [32, 64])
def test_oracle_dense_hits_match_closed_form_sphere(res):
    from denrimrendererkit_amd.sdf import sphere_volume
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0)))
    m = s.add_material(Material(base_color=(0.85, 0.18, 0.12)))
    s.add_volume(sphere_volume(res, 0.55), m, Transform.translation((0.2, -0.1, 0.3)))
    sc = O.OracleScene(compile_scene(s))
    C, R = np.array([0.2, -0.1, 0.3]), 0.55
    rays = _rays_at_sphere(np.random.default_rng(7 + res), 3000, C, R)
    t, ids, hit = sc.intersect(rays, with_hit=True)
    ta, na, disc = _analytic_sphere(rays, C, R)
    h = ids != 0xFFFFFFFF
    assert np.all(h[disc > (0.05 * R) ** 2])
    voxel = 2.0 / (res - 1)
    dt = np.abs(t[h] - ta[h])
    assert np.quantile(dt, 0.99) < 0.1 * voxel and dt.max() < 0.5 * voxel
    dn = np.linalg.norm(hit[h, :3] - na[h], axis=1)
    assert np.quantile(dn, 0.99) < voxel / R and dn.max() < 2 * voxel / R


def dense_only_scene(fields=None):
    # RenderSessionTests.swift:158-200 (and :337-376 with material fields).
    from denrimrendererkit_amd.sdf import sphere_volume
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), orthographic_scale=2.4))
    m = s.add_material(Material(base_color=(0.85, 0.18, 0.12), roughness=0.7))
    v = sphere_volume(20, 0.55)
    if fields is not None:
        v.fields = fields.samples(len(v.distances))
    s.add_volume(v, m)
    return s


def test_oracle_volume_only_scene_renders():
    s = dense_only_scene()
    b = compile_scene(s)
    assert len(b.triangles) == 0 and len(b.volume_samples) == 20 ** 3
    r = O.render(O.OracleScene(b), O.make_settings(b, 20, 20, max_bounces=1), s.camera.gpu_camera(20, 20))
    beauty, alb, nrm = r.accumulation, r.aovs[2], r.aovs[1]
    assert np.any(np.isfinite(beauty[..., :3]).all(-1) & (beauty[..., :3] > 0).any(-1))
    assert np.any((alb[..., 0] > 0.5) & (alb[..., 1] < 0.3) & (alb[..., 2] < 0.25) & (alb[..., 3] > 0.9))
    assert np.any((nrm[..., 3] > 0.9) & (np.abs(nrm[..., :3] - 0.5) > 0.01).any(-1))
    assert r.counters[SDF + 0] > 0 and r.counters[SDF + 1] > 0


def test_oracle_volume_material_fields_override_albedo():
    from denrimrendererkit_amd.sdf import MaterialFields
    s = dense_only_scene(MaterialFields(base_color=(0.08, 0.82, 0.24), opacity=0.64))
    b = compile_scene(s)
    r = O.render(O.OracleScene(b), O.make_settings(b, 20, 20, max_bounces=1), s.camera.gpu_camera(20, 20))
    a = r.aovs[2]
    assert np.any((a[..., 0] < 0.2) & (a[..., 1] > 0.7) & (a[..., 2] > 0.15) & (a[..., 2] < 0.35) & (a[..., 3] > 0.55)
                  & (a[..., 3] < 0.75))


@pytest.mark.parametrize("name", ["distance_volume_reference", "sparse_distance_volume_reference"])
def test_reference_volume_scenes_compile_and_render(name):
    s = getattr(RenderScene, name)()
    b = compile_scene(s)
    assert len(b.volumes) == 1 and len(b.triangles) == 8
    assert int(b.volumes["metadata"][0, 1]) == 4  # objectID after the four quads
    r = O.render(O.OracleScene(b), O.make_settings(b, 48, 32, max_bounces=2), s.camera.gpu_camera(48, 32))
    assert set(np.unique(r.aovs[4][..., 0]).tolist()) >= {1.0, 2.0, 5.0}
    assert np.isfinite(r.accumulation).all()
