"""GPU parity of the signed-distance paths against the CPU oracle, through
the C ABI: sparse bricks (DDA, march, bisection, cubic polish, normals,
material blend) and dense grids (march, bisection, central-difference
normals, material fields).

Tolerances:
  * material / object / primitive IDs: bit-exact;
  * depth, normal, albedo: |diff| <= 1e-5 (the same IEEE operations over
    the same FP32 values: FP16-stored distances widen exactly);
  * SDF counters: exactly equal when every ray is a primary ray
    (no lights, no environment texture, one bounce); within 1% otherwise
    (shading transcendentals differ by ulps between libm and the GPU and can
    move a rare secondary ray);
  * beauty: as the triangle parity tests (at most 1% of channels outside
    1e-3 * (1 + |ref|), mean |diff| < 2e-3).
"""
import numpy as np
import pytest

import oracle as O
from denrimrendererkit_amd import RenderQuality, RenderSettings, compile_scene, synthetic
from denrimrendererkit_amd.settings import RenderOutput
from test_sdf import aov_metrics, macro_skip_scene, no_crossing_scene, quality_probe, sdf_stats

pytestmark = pytest.mark.gpu

SDF_NAMES = ("cells", "macro", "tests", "invalid", "culls", "marches", "steps", "hits")


def _render_pair(renderer, scene, W, H, spp=1, bounces=4, quality=RenderQuality.FINAL, seed=0):
    build = compile_scene(scene)
    settings = RenderSettings(width=W, height=H, max_bounces=bounces, quality=quality, frame_seed=seed,
                              collects_sdf_traversal_stats=True)
    sess = renderer.make_session(renderer.upload(build), settings, camera=scene.camera)
    sess.render(spp)
    ost = O.make_settings(build, W, H, max_bounces=bounces, frame_seed=seed, clamp=settings.resolved_sample_radiance_clamp,
                          quality=quality.value)
    ref = O.render(O.OracleScene(build), ost, scene.camera.gpu_camera(W, H), sample_count=spp)
    return sess, ref


def _gpu_sdf(sess):
    s = sess.sdf_traversal_stats()
    return dict(cells=s.sparse_grid_cells_visited, macro=s.sparse_grid_macro_skips, tests=s.sparse_brick_tests,
                invalid=s.sparse_brick_invalid, culls=s.sparse_brick_range_culls, marches=s.sparse_brick_marches,
                steps=s.sparse_brick_march_steps, hits=s.sparse_brick_hits, empty=s.sparse_grid_empty_cells)


def _check_ids_and_aovs(sess, ref):
    prim = sess.triangle_ids()
    ref_prim = ref.aovs[6][..., 0].view(np.uint32)
    assert np.array_equal(prim, ref_prim), f"primitive IDs differ at {int((prim != ref_prim).sum())} px"
    for out, k in ((RenderOutput.MATERIAL_ID, 3), (RenderOutput.OBJECT_ID, 4)):
        assert np.array_equal(sess.pixels(out), ref.aovs[k]), out
    for out, k in ((RenderOutput.DEPTH, 0), (RenderOutput.NORMAL, 1), (RenderOutput.ALBEDO, 2)):
        d = np.abs(sess.pixels(out) - ref.aovs[k]).max()
        assert d <= 1e-5, (out, d)


def _check_beauty(sess, ref, frac_tol=0.01, mean_tol=2e-3):
    g, r = sess.pixels(RenderOutput.BEAUTY), ref.accumulation
    assert np.isfinite(g).all()
    diff = np.abs(g - r)
    bad = (diff > 1e-3 * (1.0 + np.abs(r))).mean()
    assert bad <= frac_tol, f"{bad:.4%} channels outside tolerance"
    assert diff.mean() < mean_tol, diff.mean()


def test_sdf_bricks_scene_parity(renderer):
    # The sparse-SDF workload at a reduced virtual grid (96^3) and image.
    scene = synthetic.sdf_bricks_scene(resolution=96, primitives=60, hdri=(256, 128))
    sess, ref = _render_pair(renderer, scene, 96, 96, spp=1, bounces=4)
    _check_ids_and_aovs(sess, ref)
    _check_beauty(sess, ref)
    g, o = _gpu_sdf(sess), sdf_stats(ref.counters)
    assert g["hits"] > 0 and g["marches"] > 0
    for k in SDF_NAMES:
        assert abs(g[k] - o[k]) <= 0.01 * max(o[k], 1) + 2, (k, g[k], o[k])


@pytest.mark.parametrize("quality", [RenderQuality.PREVIEW, RenderQuality.INTERACTIVE, RenderQuality.FINAL])
def test_sdf_primary_counters_exact(renderer, quality):
    # No lights and no environment texture, one bounce: every query is a
    # primary ray, so the counters are the same sums of the same traversals.
    from denrimrendererkit_amd import Camera, Material, RenderScene
    from denrimrendererkit_amd.sdf import build_sparse_volume
    scene = RenderScene(camera=Camera(origin=(0.5, 0.9, 2.0), target=(0.5, 0.45, 0.5)))
    metal = scene.add_material(Material(base_color=(0.9, 0.75, 0.55), metallic=1.0, roughness=0.3))
    matte = scene.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0, roughness=0.5))
    prims = synthetic.sdf_primitives(synthetic.content_rng(3), 40, metal, matte)
    scene.add_sparse_volume(build_sparse_volume(prims, dims=(64, 64, 64), narrow_band=2 / 63), metal)
    sess, ref = _render_pair(renderer, scene, 80, 64, spp=1, bounces=1, quality=quality)
    _check_ids_and_aovs(sess, ref)
    g, o = _gpu_sdf(sess), sdf_stats(ref.counters)
    for k in SDF_NAMES:
        assert g[k] == o[k], (k, g[k], o[k])
    assert sess.stats()[0] == ref.counters[O.CNT_PRIMARY]


def test_opaque_volume_shadow_test_equals_full_query(renderer):
    # In a volume scene where every hit blocks light, the timed shadow kernel
    # stops at the first triangle or sparse crossing known to lie in front of
    # the light, without refining it; the counting kernels run the full
    # closest-hit query. Both must give the same image, bit for bit, and the
    # oracle's within the 1-spp bar.
    scene = synthetic.sdf_bricks_scene(resolution=96, primitives=60, hdri=(256, 128))
    build = compile_scene(scene)
    gs = renderer.upload(build)
    out = []
    for stats in (False, True):
        st = RenderSettings(width=96, height=96, max_bounces=4, quality=RenderQuality.FINAL, frame_seed=0,
                            collects_sdf_traversal_stats=stats)
        sess = renderer.make_session(gs, st, camera=scene.camera)
        sess.render(2)
        out.append((sess, sess.pixels(RenderOutput.BEAUTY)))
    assert np.array_equal(out[0][1], out[1][1])
    ost = O.make_settings(build, 96, 96, max_bounces=4, frame_seed=0, clamp=64.0, quality=RenderQuality.FINAL.value)
    ref = O.render(O.OracleScene(build), ost, scene.camera.gpu_camera(96, 96), sample_count=2)
    _check_ids_and_aovs(out[0][0], ref)
    _check_beauty(out[0][0], ref)


def test_overlapping_brick_sample_ranges(renderer):
    # Bricks whose sample ranges overlap (the layout allows it: each header
    # names its own offset) and a brick whose range lies past the samples.
    # The upload gives overlapping bricks private copies of their samples and
    # keeps the out-of-range brick invalid; hits, IDs and every counter must
    # still equal the oracle's, which reads the caller's arrays as they are.
    from denrimrendererkit_amd import Camera, Material, RenderScene
    from denrimrendererkit_amd.sdf import build_sparse_volume
    scene = RenderScene(camera=Camera(origin=(0.5, 0.9, 2.0), target=(0.5, 0.45, 0.5)))
    metal = scene.add_material(Material(base_color=(0.9, 0.75, 0.55), metallic=1.0, roughness=0.3))
    matte = scene.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0, roughness=0.5))
    prims = synthetic.sdf_primitives(synthetic.content_rng(3), 40, metal, matte)
    scene.add_sparse_volume(build_sparse_volume(prims, dims=(64, 64, 64), narrow_band=2 / 63), metal)
    build = compile_scene(scene)
    bricks = build.volume_bricks
    dso = bricks["dimensionsAndSampleOffset"]
    assert len(bricks) > 40
    for k in range(1, 40, 3):  # brick k reads from brick k - 1's first sample
        dso[k, 3] = dso[k - 1, 3]
    dso[2, 3] = dso[1, 3] + 7  # a range straddling two bricks
    dso[5, 3] = len(build.volume_brick_samples) + 5  # past the end: invalid
    W, H = 80, 64
    settings = RenderSettings(width=W, height=H, max_bounces=1, quality=RenderQuality.FINAL, frame_seed=0,
                              collects_sdf_traversal_stats=True)
    sess = renderer.make_session(renderer.upload(build), settings, camera=scene.camera)
    sess.render(1)
    ost = O.make_settings(build, W, H, max_bounces=1, frame_seed=0, clamp=settings.resolved_sample_radiance_clamp,
                          quality=RenderQuality.FINAL.value)
    ref = O.render(O.OracleScene(build), ost, scene.camera.gpu_camera(W, H), sample_count=1)
    _check_ids_and_aovs(sess, ref)
    g, o = _gpu_sdf(sess), sdf_stats(ref.counters)
    assert g["hits"] > 0 and g["invalid"] > 0
    for k in SDF_NAMES:
        assert g[k] == o[k], (k, g[k], o[k])


def test_sparse_brick_without_zero_crossing_is_culled_before_march(renderer):
    # RenderSessionTests.swift:862-919; the samples (0.05) are not FP16-exact,
    # so this also covers the FP32 distance storage.
    scene = no_crossing_scene()
    sess, ref = _render_pair(renderer, scene, 8, 8, bounces=1)
    g = _gpu_sdf(sess)
    assert g["cells"] > 0 and g["tests"] > 0 and g["cells"] >= g["empty"] + g["tests"]
    assert g["culls"] > 0 and g["marches"] == 0 and g["steps"] == 0 and g["hits"] == 0
    assert g == {**sdf_stats(ref.counters)}
    sess.reset_sdf_traversal_stats()
    assert _gpu_sdf(sess)["cells"] == 0


def test_sparse_volume_uses_macro_grid_skips(renderer):
    # RenderSessionTests.swift:1069-1120.
    scene = macro_skip_scene()
    sess, ref = _render_pair(renderer, scene, 32, 32, bounces=1)
    g = _gpu_sdf(sess)
    assert g["macro"] > 0 and g["cells"] > 0 and g["tests"] > 0
    assert g == sdf_stats(ref.counters)
    _check_ids_and_aovs(sess, ref)


def test_sparse_quality_probe_matches_oracle_and_reference_bounds(renderer):
    # RenderSessionTests.swift:1024-1067 on the GPU, and each probe equal to
    # the oracle's.
    def gpu_probe(scene, w, h, quality):
        sess, ref = _render_pair(renderer, scene, w, h, bounces=1, quality=RenderQuality(quality))
        _check_ids_and_aovs(sess, ref)
        return sess.pixels(RenderOutput.NORMAL), sess.pixels(RenderOutput.DEPTH)
    low = quality_probe(28, 2, gpu_probe)
    ref = quality_probe(64, 2, gpu_probe)
    quality_probe(28, 0, gpu_probe)
    m = aov_metrics(low, ref)
    assert m["coverage"] > 0.20 and m["normal_rmse"] < 0.30 and m["depth_rmse"] < 0.05 and m["silhouette"] < 0.02


def test_two_volumes_with_transform_and_triangles(renderer):
    # Mixed scene: a ground quad, an untransformed volume and a second,
    # transformed volume (non-zero volume index), as in
    # RenderSessionTests.swift:1398-1456.
    from denrimrendererkit_amd import Camera, Material, Mesh, RenderScene, Transform
    from denrimrendererkit_amd.sdf import SDFPrimitive, build_sparse_volume
    scene = RenderScene(camera=Camera(origin=(0.2, 1.0, 3.0), target=(0.2, 0.3, 0.0)))
    ground = scene.add_material(Material(base_color=(0.5, 0.5, 0.5), roughness=0.8))
    red = scene.add_material(Material(base_color=(0.8, 0.2, 0.1), roughness=0.4))
    blue = scene.add_material(Material(base_color=(0.1, 0.3, 0.9), metallic=1.0, roughness=0.25))
    scene.add(Mesh.quad((-3, 0, 3), (3, 0, 3), (3, 0, -3), (-3, 0, -3)), ground)
    v1 = build_sparse_volume([SDFPrimitive("sphere", red, (0.5, 0.5, 0.5), 0.3)], dims=(40, 40, 40),
                             narrow_band=2 / 39)
    v2 = build_sparse_volume([SDFPrimitive("torus", blue, (0.5, 0.5, 0.5), 0.3, axis=(0, 1, 0), minor_radius=0.1)],
                             dims=(48, 48, 48), narrow_band=2 / 47)
    scene.add_sparse_volume(v1, red, Transform.translation((-1.2, -0.2, 0.0)))
    scene.add_sparse_volume(v2, blue, Transform.translation((0.6, 0.1, -0.3)) * Transform.rotation_x(0.6)
                            * Transform.scale((1.4, 1.4, 1.4)))
    sess, ref = _render_pair(renderer, scene, 96, 64, spp=2, bounces=3)
    _check_ids_and_aovs(sess, ref)
    _check_beauty(sess, ref)
    objs = set(np.unique(sess.pixels(RenderOutput.OBJECT_ID)[..., 0]).tolist())
    assert {1.0, 2.0, 3.0} <= objs  # ground (instance 0) and both volumes


@pytest.mark.parametrize("name", ["distance_volume_reference", "sparse_distance_volume_reference"])
def test_reference_volume_scene_parity(renderer, name):
    # RenderScene.distanceVolumeReference / sparseDistanceVolumeReference
    # (RenderScene.swift:770-797): triangles, lights, a dielectric and the SDF
    # model as a dense grid or as reference-format FP32 bricks.
    from denrimrendererkit_amd import RenderScene
    scene = getattr(RenderScene, name)()
    sess, ref = _render_pair(renderer, scene, 96, 64, spp=2, bounces=4)
    _check_ids_and_aovs(sess, ref)
    _check_beauty(sess, ref)


def test_dense_volume_primary_counters_exact(renderer):
    # Dense march (intersectVolume): every query a primary ray, so the dense
    # counters are the same sums of the same marches.
    from denrimrendererkit_amd import Camera, Material, RenderScene, Transform
    from denrimrendererkit_amd.sdf import build_volume, SDFPrimitive
    scene = RenderScene(camera=Camera(origin=(0.2, 0.6, 2.6), target=(0, 0.1, 0)))
    a = scene.add_material(Material(base_color=(0.9, 0.3, 0.2)))
    b = scene.add_material(Material(base_color=(0.2, 0.4, 0.9), metallic=1, roughness=0.3))
    prims = [SDFPrimitive("box", a, half_extents=(0.4, 0.2, 0.3), corner_radius=0.05),
             SDFPrimitive("sphere", b, radius=0.3, transform=Transform.translation((0.3, 0.2, 0.1)),
                          smooth_union_radius=0.1)]
    scene.add_volume(build_volume(prims, (40, 30, 34), (-0.7, -0.5, -0.6), (0.7, 0.6, 0.6)), a,
                     Transform.rotation_y(0.4))
    sess, ref = _render_pair(renderer, scene, 64, 48, spp=1, bounces=1)
    _check_ids_and_aovs(sess, ref)
    s = sess.sdf_traversal_stats()
    assert s.dense_volume_tests == int(ref.counters[O.CNT_SDF + 0]) > 0
    assert s.dense_march_steps == int(ref.counters[O.CNT_SDF + 1]) > 0


def test_volume_material_fields_override_albedo(renderer):
    # RenderSessionTests.swift:337-376 on the GPU, equal to the oracle.
    from denrimrendererkit_amd.sdf import MaterialFields
    from test_sdf import dense_only_scene
    scene = dense_only_scene(MaterialFields(base_color=(0.08, 0.82, 0.24), opacity=0.64, roughness=0.3,
                                            emission=(0.2, 0.1, 0.0), emission_strength=2.0))
    sess, ref = _render_pair(renderer, scene, 20, 20, bounces=2)
    _check_ids_and_aovs(sess, ref)
    _check_beauty(sess, ref)
    a = sess.pixels(RenderOutput.ALBEDO)
    assert np.any((a[..., 0] < 0.2) & (a[..., 1] > 0.7) & (a[..., 2] > 0.15) & (a[..., 2] < 0.35) & (a[..., 3] > 0.55)
                  & (a[..., 3] < 0.75))
