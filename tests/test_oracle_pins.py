"""CPU pins of the oracle (no GPU): the reference's own known answers and
golden data, closed forms, and brute force.

Each test names what pins it:
  * RNG: published inverse of Wellons' lowbias32 + committed golden vectors;
  * camera packing: APITests.swift:23-68 known answers;
  * BVH build / light records: BVHBuilderTests.swift:6-48,
    BVHFlattenerTests.swift:7-135 known answers;
  * triangle test: CPUIntersectorTests.swift:6-70 known answers;
  * traversal: the flat BVH equals brute force (intersectSceneLinear,
    PathTrace.metal:3265-3292) on random rays;
  * HDRI table: closed form for a constant map (pdf = 1 / sum of solid angles,
    sum p dOmega = 1);
  * integrator: the reference's stored render metrics
    (ReferenceMetrics/CornellBox.json, MaterialReference.json) through the
    reference PNG tonemap and ImageMetrics.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from denrimrendererkit_amd import Camera, CameraLens, Environment, Material, Mesh, RenderScene, Texture2D, Transform
from denrimrendererkit_amd import compile_scene
from denrimrendererkit_amd.image import tonemap
from denrimrendererkit_amd.scene import environment_samples

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M32 = 0xFFFFFFFF


# --------------------------------------------------------------------- RNG --
def _lowbias32_inverse(x):
    # Published inverse (hash-prospector): constants are the modular inverses
    # of the forward multipliers; a wrong forward constant fails this.
    x ^= x >> 16
    x = (x * 0x43021123) & M32
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M32
    x ^= x >> 16
    return x


def test_rng_hash_inverts_with_published_inverse():
    rng = np.random.default_rng(5)
    for v in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64).tolist() + [0, 1, M32]:
        assert _lowbias32_inverse(O.hash_u32(v)) == v


def test_rng_golden_vectors():
    g = json.load(open(os.path.join(GOLDEN, "rng_lowbias32.json")))
    for k, v in g["hash"].items():
        assert O.hash_u32(int(k)) == v
    for s in g["sequences"]:
        seed = s["frame_seed"] ^ O.hash_u32(s["px"] + s["py"] * s["width"]) ^ O.hash_u32(s["sample"] + 1)
        assert seed == s["seed"]
        vals, st = O.random_floats(seed, len(s["floats"]))
        assert vals == s["floats"]
        assert st == s["state_after"]
        assert all(0.0 <= x < 1.0 for x in vals)


# ------------------------------------------------------------------ camera --
def test_camera_packing_known_answers():
    g = Camera(origin=(0, 0, 3), target=(0, 0, 0), vertical_field_of_view_degrees=90).gpu_camera(200, 100)
    assert abs(g["origin"][3]) < 1e-4 and abs(g["origin"][2] - 3) < 1e-4
    assert abs(g["lowerLeft"][2] - 2) < 1e-4
    assert abs(np.linalg.norm(g["horizontal"][:3]) - 4) < 1e-4
    assert abs(np.linalg.norm(g["vertical"][:3]) - 2) < 1e-4
    o = Camera(origin=(0, 0, 3), target=(0, 0, 2), orthographic_scale=6).gpu_camera(200, 100)
    assert abs(o["origin"][3] - 1) < 1e-4
    assert np.allclose(o["lowerLeft"][:3], [-6, -3, 3], atol=1e-4)
    assert abs(np.linalg.norm(o["horizontal"][:3]) - 12) < 1e-4
    assert abs(np.linalg.norm(o["vertical"][:3]) - 6) < 1e-4
    l = Camera(origin=(0, 0, 3), target=(0, 0, 0), lens=CameraLens(4.5, 0.08)).gpu_camera(200, 100)
    assert abs(l["lens"][0] - 0.08) < 1e-4 and abs(l["lens"][1] - 4.5) < 1e-4


# --------------------------------------------------------------------- BVH --
def _quads_scene(quads, materials=None):
    s = RenderScene()
    m = s.add_material(Material(base_color=(0.8, 0.8, 0.8)))
    for q in quads:
        s.add(Mesh.quad(*q), m)
    return s


def _oracle_bvh(scene, max_leaf):
    b = compile_scene(scene)
    osc = O.OracleScene(b, max_leaf=max_leaf)
    nodes, prim = osc.export_bvh()
    return b, nodes, prim


def test_bvh_single_leaf():
    s = _quads_scene([((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))])
    _, nodes, prim = _oracle_bvh(s, 4)
    assert len(nodes) == 1
    assert sorted(prim.tolist()) == [0, 1]
    assert nodes[0]["metadata"].tolist() == [0, 0, 0, 2]
    assert abs(nodes[0]["boundsMin"][0] + 1) < 1e-4 and abs(nodes[0]["boundsMax"][0] - 1) < 1e-4


def test_bvh_splits_with_leaf_two():
    quads = [((x, 0, 0), (x + 0.5, 0, 0), (x + 0.5, 0.5, 0), (x, 0.5, 0)) for x in range(8)]
    _, nodes, prim = _oracle_bvh(_quads_scene(quads), 2)
    assert nodes[0]["metadata"][3] == 0 and len(nodes) > 1
    assert sorted(prim.tolist()) == list(range(16))
    leaves = nodes["metadata"][:, 3]
    assert leaves.max() <= 2
    assert abs(nodes[0]["boundsMin"][0]) < 1e-4 and abs(nodes[0]["boundsMax"][0] - 7.5) < 1e-4


def test_light_records_known_answers():
    s = RenderScene()
    matte = s.add_material(Material(base_color=(0.8, 0.8, 0.8)))
    light = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 0.8, 0.6), emission_strength=4))
    s.add(Mesh.quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)), matte)
    s.add(Mesh.quad((-0.5, 1, -0.5), (0.5, 1, -0.5), (0.5, 1, 0.5), (-0.5, 1, 0.5)), light)
    b = compile_scene(s)
    assert len(b.triangles) == 4
    assert b.lights["triangleIndex"].tolist() == [2, 3]
    assert np.all(np.abs(b.lights["area"] - 0.5) < 1e-4)
    assert np.all(b.lights["normal"][:, 1] < -0.99)
    assert b.triangles["padding2"].tolist() == [0, 0, 1, 2]
    assert abs(b.lights["selectionCDF"][-1] - 1) < 1e-4


def test_light_cdf_power_weighted():
    s = RenderScene()
    dim = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 1, 1), emission_strength=1))
    bright = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 1, 1), emission_strength=3))
    s.add(Mesh.quad((-1, 0, 1), (0, 0, 1), (0, 0, 0), (-1, 0, 0)), dim)
    s.add(Mesh.quad((0, 0, 1), (1, 0, 1), (1, 0, 0), (0, 0, 0)), bright)
    b = compile_scene(s)
    cdf = b.lights["selectionCDF"]
    assert len(cdf) == 4 and abs(cdf[-1] - 1) < 1e-4 and np.all(np.diff(cdf) > 0)
    assert abs(cdf[1] - 0.25) < 1e-4
    assert b.triangles["padding2"].tolist() == [1, 2, 3, 4]


def test_environment_table_constant_map_closed_form():
    w, h = 64, 32
    tex = Texture2D(w, h, np.ones((w * h, 4), np.float32) * 2.0)
    e = environment_samples(Environment(texture=tex, max_radiance=16))
    d_theta, d_phi = math.pi / h, 2 * math.pi / w
    solid = np.array([max(math.sin((y + 0.5) * d_theta) * d_theta * d_phi, 1e-8) for y in range(h)])
    pdf = e["pdf"].reshape(h, w)
    expected = 1.0 / (solid.sum() * w)
    assert np.allclose(pdf, expected, rtol=1e-4)
    assert abs((pdf * solid[:, None]).sum() - 1.0) < 1e-4
    assert abs(expected - 1 / (4 * math.pi)) < 2e-3 / (4 * math.pi) * 10
    assert abs(e["cdf"][-1] - 1.0) < 1e-6 and np.all(np.diff(e["cdf"]) >= 0)


# ------------------------------------------------------------- intersector --
def _rays(o, d):
    r = np.zeros((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    return r


def test_triangle_test_known_answers():
    s = _quads_scene([((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))])
    osc = O.OracleScene(compile_scene(s))
    t, ids = osc.intersect(_rays([[0, 0, 2], [3, 0, 2]], [[0, 0, -1], [0, 0, -1]]))
    assert abs(t[0] - 2) < 1e-5 and ids[0] != 0xFFFFFFFF
    assert ids[1] == 0xFFFFFFFF
    s2 = RenderScene()
    m = s2.add_material(Material(base_color=(1, 1, 1)))
    s2.add(Mesh.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)), m, Transform.translation((2, 0, 0)))
    osc2 = O.OracleScene(compile_scene(s2))
    t2, ids2 = osc2.intersect(_rays([[2, 0, 2], [0, 0, 2]], [[0, 0, -1], [0, 0, -1]]))
    assert abs(t2[0] - 2) < 1e-5 and ids2[1] == 0xFFFFFFFF


def _random_rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays(o, d)


@pytest.mark.parametrize("leaf", [1, 4])
def test_flat_bvh_equals_brute_force_room(leaf):
    from denrimrendererkit_amd import synthetic
    b = compile_scene(synthetic.cornell_room_scene())
    osc = O.OracleScene(b, max_leaf=leaf)
    rays = _random_rays(np.random.default_rng(11), 4000, [-1.9, 0.1, -1.9], [1.9, 2.9, 1.9])
    t0, i0 = osc.intersect(rays)
    t1, i1 = osc.intersect(rays, brute_force=True)
    assert np.array_equal(i0, i1) and np.array_equal(t0, t1)


def test_flat_bvh_equals_brute_force_soup_with_ties():
    # Duplicated coplanar triangles produce exact t ties: lowest index wins.
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    idx = np.arange(300, dtype=np.uint32)
    s = RenderScene()
    m = s.add_material(Material())
    mesh = Mesh(v, idx)
    s.add(mesh, m)
    s.add(mesh, m)  # same triangles again, higher indices
    b = compile_scene(s)
    osc = O.OracleScene(b)
    rays = _random_rays(rng, 3000, [-1.5] * 3, [1.5] * 3)
    t0, i0 = osc.intersect(rays)
    t1, i1 = osc.intersect(rays, brute_force=True)
    assert np.array_equal(i0, i1)
    hit = i0 != 0xFFFFFFFF
    assert hit.any() and (i0[hit] < 100).all()  # the first copy always wins


# ---------------------------------------------------- reference metrics --
def _image_metrics(rgba8):
    H, W = rgba8.shape[:2]
    px = rgba8.astype(np.int64)
    br = (px[..., 0] + px[..., 1] + px[..., 2]) / 3.0
    colors = set(((px[..., 0] // 32) << 16 | (px[..., 1] // 32) << 8 | (px[..., 2] // 32)).ravel().tolist())
    ys, xs = np.mgrid[0:H, 0:W]
    cx = (xs >= W * 3 // 8) & (xs <= W * 5 // 8)
    top = br[cx & (ys < H // 4)].max()
    bottom = br[cx & (ys > H * 3 // 4)].max()
    return {"averageBrightness": br.mean(), "maxBrightness": br.max(), "uniqueColorEstimate": len(colors),
            "topCenterMaxBrightness": top, "bottomCenterMaxBrightness": bottom}


def _assert_metrics(m, base):
    tol = base["tolerance"]
    ref = base["metrics"]
    assert abs(m["averageBrightness"] - ref["averageBrightness"]) <= tol["averageBrightness"], m
    assert abs(m["maxBrightness"] - ref["maxBrightness"]) <= tol["maxBrightness"], m
    assert abs(m["uniqueColorEstimate"] - ref["uniqueColorEstimate"]) <= tol["uniqueColorEstimate"], m
    assert abs(m["topCenterMaxBrightness"] - ref["topCenterMaxBrightness"]) <= tol["regionalBrightness"], m
    assert abs(m["bottomCenterMaxBrightness"] - ref["bottomCenterMaxBrightness"]) <= tol["regionalBrightness"], m


@pytest.mark.parametrize("name,factory", [("CornellBox", RenderScene.cornell_box),
                                          ("MaterialReference", RenderScene.material_reference)])
def test_oracle_matches_reference_render_metrics(name, factory):
    base = json.load(open(os.path.join(GOLDEN, "reference_metrics", f"{name}.json")))
    W, H = base["metrics"]["width"], base["metrics"]["height"]
    scene = factory()
    b = compile_scene(scene)
    st = O.make_settings(b, W, H, max_bounces=base["maxBounces"], frame_seed=0x1234ABCD, clamp=64.0)
    r = O.render(O.OracleScene(b), st, scene.camera.gpu_camera(W, H), sample_count=base["samples"])
    rgba8 = np.zeros((H, W, 4), np.uint8)
    rgba8[..., :3] = tonemap(r.accumulation[..., :3])
    _assert_metrics(_image_metrics(rgba8), base)


def test_fp32_and_fp64_oracles_agree_on_ids():
    from denrimrendererkit_amd import synthetic
    scene = synthetic.cornell_room_scene()
    b = compile_scene(scene)
    osc = O.OracleScene(b)
    st = O.make_settings(b, 64, 36, max_bounces=4, frame_seed=0)
    cam = scene.camera.gpu_camera(64, 36)
    r32 = O.render(osc, st, cam, precision=32)
    r64 = O.render(osc, st, cam, precision=64)
    # Material / object IDs come from the same first-hit triangle in both
    # precisions except where rays graze an edge.
    same = (r32.aovs[3][..., 0] == r64.aovs[3][..., 0]).mean()
    assert same > 0.99
    assert np.isfinite(r64.accumulation).all()


def test_oracle_matches_more_reference_fixtures():
    # ReferenceMetrics/ScriptedTexture{Albedo,Normal}.json (the albedo and
    # normal AOV PNGs of a checker-textured, normal-mapped quad, preview
    # quality) and DistanceVolumeReference.json (the dense distance-volume
    # scene, final quality, 4 spp), through the reference's output encodings.
    from denrimrendererkit_amd import scenescript
    from denrimrendererkit_amd.image import visualized_rgba8
    from denrimrendererkit_amd.settings import RenderOutput
    script = ("camera 0 0 3 0 0 0 34\ntexture checker checker 1 0 0 1 0 0 1 1\n"
              "texture tangentRight solid 1 0.5 0.5 1\n"
              "material textured 1 1 1 roughness 0.7 baseColorTexture checker normalMap tangentRight\n"
              "quad textured -1 -1 0 1 -1 0 1 1 0 -1 1 0\n")
    scene = scenescript.parse(script)
    b = compile_scene(scene)
    st = O.make_settings(b, 48, 48, max_bounces=1, frame_seed=0x1234ABCD, clamp=10.0, quality=0)
    r = O.render(O.OracleScene(b), st, scene.camera.gpu_camera(48, 48), sample_count=1)
    for k, out, name in ((2, RenderOutput.ALBEDO, "ScriptedTextureAlbedo"), (1, RenderOutput.NORMAL, "ScriptedTextureNormal")):
        base = json.load(open(os.path.join(GOLDEN, "reference_metrics", f"{name}.json")))
        _assert_metrics(_image_metrics(visualized_rgba8(r.aovs[k], out)), base)
    base = json.load(open(os.path.join(GOLDEN, "reference_metrics", "DistanceVolumeReference.json")))
    scene = RenderScene.distance_volume_reference()
    b = compile_scene(scene)
    W, H = base["metrics"]["width"], base["metrics"]["height"]
    st = O.make_settings(b, W, H, max_bounces=base["maxBounces"], frame_seed=0x1234ABCD, clamp=64.0)
    r = O.render(O.OracleScene(b), st, scene.camera.gpu_camera(W, H), sample_count=base["samples"])
    _assert_metrics(_image_metrics(visualized_rgba8(r.accumulation, RenderOutput.BEAUTY)), base)
