"""GPU parity of the path tracer's less common branches, each against the
oracle (IDs bit-exact, depth/normal/albedo/motion AOVs within 1e-5 of the
oracle, 1e-4 relative for motion, beauty within the 1-spp bound of
test_gpu_parity), plus the reference's own assertions and stored metrics
for the same scenes:

  * motion vectors from a shifted previous camera (AOVTests.swift:71-94;
    PathTrace.metal:1666-1701, :5526-5585);
  * transparent-background export (TransparentExportRenderTests.swift:9-120,
    ReferenceMetrics/TransparentExport.json);
  * opacity pass-through in primaries and shadow chains
    (OpacityTransportTests.swift:7-60; PathTrace.metal:5318-5321, :4221-4271)
    and the albedo AOV's opacity (AOVTests.swift:119-136);
  * base-colour textures, normal maps and linear sampling
    (TextureMaterialTests.swift:7-96; PathTrace.metal:3723-3763, :4005-4054)
    and the scripted-texture metrics (ScriptedTextureReferenceRenderTests.swift,
    ReferenceMetrics/ScriptedTexture*.json);
  * the dense distance-volume reference scene's stored metrics and AOVs
    (DistanceVolumeReferenceRenderTests.swift);
  * a thin-lens camera (PathTrace.metal:4056-4091, lens.x > 0);
  * the transmission-volume random walk of the liquid.milk preset
    (PathTrace.metal:3685-3721);
  * tile rendering (RenderSession.swift:876-889) and updateCamera
    (RenderSession.swift:595-608).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from denrimrendererkit_amd import (Camera, CameraLens, Material, Mesh, RenderQuality, RenderScene, RenderSettings,
                                   Texture2D, compile_scene, synthetic)
from denrimrendererkit_amd.scene import TextureSamplingMode
from denrimrendererkit_amd.settings import RenderOutput

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_metrics")
FINAL, PREVIEW = RenderQuality.FINAL, RenderQuality.PREVIEW


def _both(renderer, scene, W, H, spp=1, bounces=4, quality=FINAL, seed=0x1234ABCD, previous=None,
          transparent=False, camera=None):
    build = compile_scene(scene)
    cam = camera or scene.camera
    st = RenderSettings(width=W, height=H, max_bounces=bounces, quality=quality, frame_seed=seed,
                        previous_camera=previous, transparent_background=transparent)
    sess = renderer.make_session(renderer.upload(build), st, camera=cam)
    sess.render(spp)
    ost = O.make_settings(build, W, H, max_bounces=bounces, frame_seed=seed, clamp=st.resolved_sample_radiance_clamp,
                          transparent_background=transparent, quality=quality.value)
    ref = O.render(O.OracleScene(build), ost, cam.gpu_camera(W, H),
                   previous_camera=previous.gpu_camera(W, H) if previous is not None else None, sample_count=spp)
    return sess, ref


def _check(sess, ref, motion=True):
    tri = sess.triangle_ids()
    ref_tri = ref.aovs[6][..., 0].view(np.uint32)
    assert np.array_equal(tri, ref_tri), f"triangle IDs differ at {int((tri != ref_tri).sum())} px"
    for out, k in ((RenderOutput.MATERIAL_ID, 3), (RenderOutput.OBJECT_ID, 4)):
        assert np.array_equal(sess.pixels(out), ref.aovs[k]), out
    for out, k in ((RenderOutput.DEPTH, 0), (RenderOutput.NORMAL, 1), (RenderOutput.ALBEDO, 2)):
        d = np.abs(sess.pixels(out) - ref.aovs[k]).max()
        assert d <= 1e-5, (out, d)
    if motion:
        m, r = sess.pixels(RenderOutput.MOTION_VECTOR), ref.aovs[5]
        assert np.all(np.abs(m - r) <= 1e-4 * (1 + np.abs(r))), np.abs(m - r).max()
    g, r = sess.pixels(RenderOutput.BEAUTY), ref.accumulation
    assert np.isfinite(g).all()
    diff = np.abs(g - r)
    bad = (diff > 1e-3 * (1.0 + np.abs(r))).mean()
    assert bad <= 0.01, f"{bad:.4%} beauty channels outside tolerance"
    assert diff.mean() < 2e-3, diff.mean()


def _quad_scene(material, z=0.0, half=1.0):
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), vertical_field_of_view_degrees=34))
    m = s.add_material(material)
    s.add(Mesh.quad((-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z)), m)
    return s


def _png_rgba(sess, output, tmp_path):
    from PIL import Image
    path = os.path.join(tmp_path, f"out{int(output)}.png")
    sess.write_png(path, output)
    return np.asarray(Image.open(path).convert("RGBA"))


# ------------------------------------------------------------ motion --
@pytest.mark.parametrize("quality", [FINAL, PREVIEW])
def test_motion_vectors_with_shifted_previous_camera(renderer, quality):
    scene = RenderScene.material_reference()
    c = scene.camera
    prev = Camera(origin=tuple(np.add(c.origin, (0.15, 0, 0))), target=tuple(np.add(c.target, (0.15, 0, 0))),
                  up=c.up, vertical_field_of_view_degrees=c.vertical_field_of_view_degrees)
    for W, H in ((32, 32), (96, 80)):
        sess, ref = _both(renderer, scene, W, H, bounces=1, quality=quality, previous=prev)
        _check(sess, ref)
        motion = sess.pixels(RenderOutput.MOTION_VECTOR)
        assert ((np.abs(motion[..., 0]) > 0.01) | (np.abs(motion[..., 1]) > 0.01)).any()
        assert (motion[..., 3] > 0).any()


# ------------------------------------------------------- transparency --
def test_transparent_export_matches_reference_alpha_metrics(renderer, tmp_path):
    base = json.load(open(os.path.join(GOLDEN, "TransparentExport.json")))
    scene = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), vertical_field_of_view_degrees=34))
    m = scene.add_material(Material(base_color=(0.9, 0.2, 0.1), emission=(0.9, 0.2, 0.1), emission_strength=1))
    scene.add(Mesh.quad((-0.45, -0.45, 0), (0.45, -0.45, 0), (0.45, 0.45, 0), (-0.45, 0.45, 0)), m)
    W, H = base["metrics"]["width"], base["metrics"]["height"]
    for quality in (PREVIEW, FINAL):
        sess, ref = _both(renderer, scene, W, H, spp=base["samples"], bounces=base["maxBounces"], quality=quality,
                          transparent=True)
        _check(sess, ref)
    sess, _ = _both(renderer, scene, W, H, spp=base["samples"], bounces=base["maxBounces"], quality=PREVIEW,
                    transparent=True)
    alpha = _png_rgba(sess, RenderOutput.BEAUTY, tmp_path)[..., 3].astype(int)
    got = {"width": W, "height": H, "minAlpha": int(alpha.min()), "maxAlpha": int(alpha.max()),
           "transparentPixels": int((alpha < 8).sum()), "opaquePixels": int((alpha > 247).sum())}
    ref_m, tol = base["metrics"], base["tolerance"]
    assert got["transparentPixels"] > 0 and got["opaquePixels"] > 0
    assert got["transparentPixels"] + got["opaquePixels"] == W * H  # one centre ray per pixel: alpha 0 or 1
    assert (got["minAlpha"], got["maxAlpha"]) == (ref_m["minAlpha"], ref_m["maxAlpha"]), got
    # Exact coverage: the pixel centres whose camera ray meets |x|, |y| <= 0.45 at z = 0.
    cam = scene.camera.gpu_camera(W, H)
    ys, xs = np.mgrid[0:H, 0:W]
    u, v = (xs + 0.5) / W, 1.0 - (ys + 0.5) / H
    p = cam["lowerLeft"][:3] + u[..., None] * cam["horizontal"][:3] + v[..., None] * cam["vertical"][:3]
    d = p - cam["origin"][:3]
    t = -cam["origin"][2] / d[..., 2]
    hit = (np.abs(cam["origin"][0] + t * d[..., 0]) <= 0.45) & (np.abs(cam["origin"][1] + t * d[..., 1]) <= 0.45)
    assert got["opaquePixels"] == int(hit.sum()) == 256
    # The reference's stored counts (opaque 246 +- 4) are 10 lower. 16 pixel
    # centres lie on the quad's shared diagonal, the one place where a
    # non-watertight triangle test can miss both triangles under the
    # reference's Metal fast-math rounding; IEEE arithmetic (oracle and HIP
    # alike) hits one there. The allowance covers those 16.
    diag = 16
    assert abs(got["transparentPixels"] - ref_m["transparentPixels"]) <= tol["transparentPixels"] + diag, got
    assert abs(got["opaquePixels"] - ref_m["opaquePixels"]) <= tol["opaquePixels"] + diag, got


def _cutout_scene():
    s = RenderScene(camera=Camera(origin=(0, 0, 3), target=(0, 0, 0), vertical_field_of_view_degrees=34))
    cut = s.add_material(Material(base_color=(1, 0, 0), roughness=0.8, opacity=0))
    rear = s.add_material(Material(base_color=(0, 1, 0), emission=(0, 1, 0), emission_strength=1, roughness=0.8))
    s.add(Mesh.quad((-0.9, -0.9, 0), (0.9, -0.9, 0), (0.9, 0.9, 0), (-0.9, 0.9, 0)), cut)
    s.add(Mesh.quad((-0.9, -0.9, -0.25), (0.9, -0.9, -0.25), (0.9, 0.9, -0.25), (-0.9, 0.9, -0.25)), rear)
    return s


def test_fully_transparent_primary_reveals_rear_surface(renderer):
    sess, ref = _both(renderer, _cutout_scene(), 24, 24, bounces=2, quality=FINAL)
    _check(sess, ref)
    albedo = sess.pixels(RenderOutput.ALBEDO).reshape(-1, 4)
    albedo = albedo[albedo[:, 3] > 0]
    assert ((albedo[:, 0] < 0.1) & (albedo[:, 1] > 0.8) & (albedo[:, 2] < 0.1) & (albedo[:, 3] > 0.99)).any()
    assert not ((albedo[:, 0] > 0.8) & (albedo[:, 1] < 0.1) & (albedo[:, 2] < 0.1) & (albedo[:, 3] < 0.01)).any()
    beauty = sess.pixels(RenderOutput.BEAUTY).reshape(-1, 4)
    assert ((beauty[:, 1] > 0.8) & (beauty[:, 0] < 0.2) & (beauty[:, 2] < 0.2)).any()


def test_shadow_chains_pass_cutouts_and_dielectrics(renderer):
    # A floor lit by an area light through a fully transparent panel and a
    # glass panel: the NEE shadow rays step through both (the cutout
    # unattenuated, the glass with its tint), the reference's up-to-8-step
    # chain. Camera below the panels' height so primaries hit the floor.
    s = RenderScene(camera=Camera(origin=(0, 0.6, 3.2), target=(0, 0.2, 0), vertical_field_of_view_degrees=45))
    floor = s.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0))
    light = s.add_material(Material(base_color=(1, 1, 1), emission=(1, 1, 1), emission_strength=12))
    cut = s.add_material(Material(base_color=(1, 0, 0), opacity=0.0))
    glass = s.add_material(Material(base_color=(1, 1, 1), transmission=1.0, transmission_color=(0.6, 0.9, 1.0),
                                    roughness=0.05, index_of_refraction=1.5))
    s.add(Mesh.quad((-3, 0, -3), (-3, 0, 3), (3, 0, 3), (3, 0, -3)), floor)
    s.add(Mesh.quad((-0.6, 2.5, -0.6), (0.6, 2.5, -0.6), (0.6, 2.5, 0.6), (-0.6, 2.5, 0.6)), light)
    s.add(Mesh.quad((-2.0, 1.2, -2.0), (0.0, 1.2, -2.0), (0.0, 1.2, 2.0), (-2.0, 1.2, 2.0)), cut)
    s.add(Mesh.quad((0.0, 1.3, -2.0), (2.0, 1.3, -2.0), (2.0, 1.3, 2.0), (0.0, 1.3, 2.0)), glass)
    for spp in (1, 8):
        sess, ref = _both(renderer, s, 96, 64, spp=spp, bounces=3, quality=FINAL, seed=5)
        _check(sess, ref)
    beauty = sess.pixels(RenderOutput.BEAUTY)
    left, right = beauty[40:, 10:40, :3].mean(), beauty[40:, 56:86, :3].mean()
    assert left > 0.05 and right > 0.02, (left, right)  # light reaches the floor under both panels


def test_albedo_aov_keeps_material_opacity(renderer, tmp_path):
    scene = _quad_scene(Material(base_color=(0.7, 0.2, 0.1), roughness=0.8, opacity=0.35))
    for quality in (PREVIEW, FINAL):
        sess, ref = _both(renderer, scene, 16, 16, bounces=1, quality=quality)
        _check(sess, ref)
    sess, _ = _both(renderer, scene, 16, 16, bounces=1, quality=PREVIEW)
    albedo = sess.pixels(RenderOutput.ALBEDO).reshape(-1, 4)
    assert (np.abs(albedo[albedo[:, 3] > 0][:, 3] - 0.35) < 0.02).any()
    alpha = _png_rgba(sess, RenderOutput.ALBEDO, tmp_path)[..., 3].astype(int)
    assert (np.abs(alpha - 89) <= 3).any()


# ----------------------------------------------------------- textures --
_SCRIPT = """camera 0 0 3 0 0 0 34
texture checker checker 1 0 0 1 0 0 1 1
texture tangentRight solid 1 0.5 0.5 1
material textured 1 1 1 roughness 0.7 baseColorTexture checker normalMap tangentRight
quad textured -1 -1 0 1 -1 0 1 1 0 -1 1 0
"""


def test_textures_normal_maps_and_linear_sampling(renderer):
    from denrimrendererkit_amd import scenescript
    checker = scenescript.parse("camera 0 0 3 0 0 0 34\ntexture c checker 1 0 0 1 0 0 1 1\n"
                                "material m 1 1 1 roughness 0.8 baseColorTexture c\n"
                                "quad m -1 -1 0 1 -1 0 1 1 0 -1 1 0\n")
    normal = scenescript.parse("camera 0 0 3 0 0 0 34\ntexture n solid 1 0.5 0.5 1\n"
                               "material m 1 1 1 roughness 0.8 normalMap n\nquad m -1 -1 0 1 -1 0 1 1 0 -1 1 0\n")
    blended = Texture2D(2, 2, np.array([[1, 0, 0, 1], [0, 0, 1, 1], [1, 0, 0, 1], [0, 0, 1, 1]], np.float32),
                        sampling_mode=TextureSamplingMode.LINEAR)
    linear = _quad_scene(Material(base_color=(1, 1, 1), roughness=0.8, base_color_texture=blended))
    for scene, W in ((checker, 24), (normal, 16), (linear, 32)):
        for quality in (PREVIEW, FINAL):
            sess, ref = _both(renderer, scene, W, W, bounces=1 if quality == PREVIEW else 3, quality=quality)
            _check(sess, ref)
    a = _both(renderer, checker, 24, 24, bounces=1, quality=PREVIEW)[0].pixels(RenderOutput.ALBEDO).reshape(-1, 4)
    a = a[a[:, 3] > 0]
    assert ((a[:, 0] > 0.9) & (a[:, 1] < 0.1) & (a[:, 2] < 0.1)).any()
    assert ((a[:, 0] < 0.1) & (a[:, 1] < 0.1) & (a[:, 2] > 0.9)).any()
    n = _both(renderer, normal, 16, 16, bounces=1, quality=PREVIEW)[0].pixels(RenderOutput.NORMAL).reshape(-1, 4)
    n = n[n[:, 3] > 0]
    assert ((n[:, 0] > 0.9) & (n[:, 1] > 0.4) & (n[:, 1] < 0.6) & (n[:, 2] > 0.4) & (n[:, 2] < 0.6)).any()
    b = _both(renderer, linear, 32, 32, bounces=1, quality=PREVIEW)[0].pixels(RenderOutput.ALBEDO).reshape(-1, 4)
    b = b[b[:, 3] > 0]
    assert ((b[:, 0] > 0.25) & (b[:, 0] < 0.75) & (b[:, 1] < 0.05) & (b[:, 2] > 0.25) & (b[:, 2] < 0.75)).any()


def test_scripted_texture_matches_reference_metrics(renderer, tmp_path):
    from denrimrendererkit_amd import scenescript
    from test_oracle_pins import _assert_metrics, _image_metrics
    scene = scenescript.parse(_SCRIPT)
    for quality in (PREVIEW, FINAL):
        sess, ref = _both(renderer, scene, 48, 48, bounces=1, quality=quality)
        _check(sess, ref)
    sess, _ = _both(renderer, scene, 48, 48, bounces=1, quality=PREVIEW)
    for output, name in ((RenderOutput.ALBEDO, "ScriptedTextureAlbedo"), (RenderOutput.NORMAL, "ScriptedTextureNormal")):
        base = json.load(open(os.path.join(GOLDEN, f"{name}.json")))
        m = _image_metrics(_png_rgba(sess, output, tmp_path))
        _assert_metrics(m, base)
    m = _image_metrics(_png_rgba(sess, RenderOutput.ALBEDO, tmp_path))
    assert m["uniqueColorEstimate"] >= 2
    assert _image_metrics(_png_rgba(sess, RenderOutput.NORMAL, tmp_path))["averageBrightness"] > 100


# ------------------------------------------------------ dense volume --
def test_distance_volume_reference_metrics_and_aovs(renderer, tmp_path):
    from test_oracle_pins import _assert_metrics, _image_metrics
    base = json.load(open(os.path.join(GOLDEN, "DistanceVolumeReference.json")))
    scene = RenderScene.distance_volume_reference()
    W, H = base["metrics"]["width"], base["metrics"]["height"]
    sess, ref = _both(renderer, scene, W, H, spp=base["samples"], bounces=base["maxBounces"], quality=FINAL)
    _check(sess, ref)
    m = _image_metrics(_png_rgba(sess, RenderOutput.BEAUTY, tmp_path))
    assert m["averageBrightness"] > 15 and m["uniqueColorEstimate"] > 8 and m["maxBrightness"] > 90, m
    _assert_metrics(m, base)
    sess, ref = _both(renderer, scene, 48, 48, bounces=3, quality=FINAL)
    _check(sess, ref)
    albedo = sess.pixels(RenderOutput.ALBEDO).reshape(-1, 4)
    normal = sess.pixels(RenderOutput.NORMAL).reshape(-1, 4)
    beauty = sess.pixels(RenderOutput.BEAUTY).reshape(-1, 4)
    albedo, normal = albedo[albedo[:, 3] > 0], normal[normal[:, 3] > 0]
    assert ((albedo[:, 2] > 0.75) & (albedo[:, 0] < 0.35) & (albedo[:, 1] > 0.35) & (albedo[:, 3] > 0.35) &
            (albedo[:, 3] < 0.5)).any(), "semi-transparent blue volume"
    assert ((np.abs(normal[:, 0] - 0.5) > 0.05) & (np.abs(normal[:, 1] - 0.5) > 0.05) & (normal[:, 3] > 0.9)).any()
    assert ((beauty[:, 1] > 0.18) & (beauty[:, 0] < 0.25)).any(), "rear panel through the volume"


# ------------------------------------------------------------ camera --
def test_thin_lens_camera(renderer):
    scene = synthetic.cornell_room_scene()
    c = scene.camera
    lens_cam = Camera(origin=c.origin, target=c.target, up=c.up, vertical_field_of_view_degrees=c.vertical_field_of_view_degrees,
                      lens=CameraLens(focus_distance=2.5, aperture_radius=0.12))
    assert lens_cam.gpu_camera(16, 9)["lens"][0] > 0
    for spp in (1, 4):
        sess, ref = _both(renderer, scene, 160, 90, spp=spp, camera=lens_cam, seed=0)
        _check(sess, ref)
    pin = _both(renderer, scene, 160, 90, spp=4, seed=0)[0].pixels(RenderOutput.DEPTH)[..., 0]
    assert (np.abs(pin - sess.pixels(RenderOutput.DEPTH)[..., 0]) > 1e-3).mean() > 0.1, "the lens moves primary rays"


# ------------------------------------------------- transmission volume --
def test_transmission_volume_walk_liquid_milk(renderer):
    from denrimrendererkit_amd import materials
    milk = materials.material("liquid.milk")
    assert milk.volume_scattering > 0 and milk.transmission > 0 and not milk.thin_walled
    s = RenderScene(camera=Camera(origin=(0.0, 1.0, 2.6), target=(0.0, 0.7, 0.0), vertical_field_of_view_degrees=40))
    s.environment = synthetic.material_preview_scene(0, hdri=(64, 32), segments=8).environment
    plane = s.add_material(Material(base_color=(0.5, 0.5, 0.5), roughness=0.7))
    s.add(Mesh.quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4)), plane)
    from denrimrendererkit_amd import Transform
    s.add(synthetic.uv_sphere(32, 33, 0.7), s.add_material(milk), Transform.translation((0.0, 0.7, 0.0)))
    for spp in (1, 8):
        sess, ref = _both(renderer, s, 96, 80, spp=spp, bounces=8, seed=11)
        _check(sess, ref)
    # The walk changes the image: without volume scattering the oracle differs.
    import dataclasses
    plain = dataclasses.replace(milk, volume_scattering=0.0)
    s2 = RenderScene(camera=s.camera, environment=s.environment)
    s2.add(Mesh.quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4)), s2.add_material(Material(base_color=(0.5,) * 3,
                                                                                             roughness=0.7)))
    s2.add(synthetic.uv_sphere(32, 33, 0.7), s2.add_material(plain), Transform.translation((0.0, 0.7, 0.0)))
    _, ref2 = _both(renderer, s2, 96, 80, spp=8, bounces=8, seed=11)
    assert np.abs(ref2.accumulation - ref.accumulation).mean() > 1e-3


# ------------------------------------------------- tiles and cameras --
def test_render_tiles_equal_full_frame_sample(renderer):
    scene = synthetic.cornell_room_scene()
    W, H = 72, 50
    gs = renderer.upload(compile_scene(scene))
    st = RenderSettings(width=W, height=H, max_bounces=4, quality=FINAL, frame_seed=9)
    full = renderer.make_session(gs, st, camera=scene.camera)
    full.render(2)
    tiled = renderer.make_session(gs, st, camera=scene.camera)
    tiled.render(1)
    tiles = [(x, y, 32, 24) for y in range(0, H, 24) for x in range(0, W, 32)]
    for k, (x, y, w, h) in enumerate(tiles):
        tiled.render_next_tile(x, y, w, h, completes_sample=(k == len(tiles) - 1))
        assert tiled.sample_count == (2 if k == len(tiles) - 1 else 1)
    tiled.render_next_tile(W + 5, 0, 8, 8, completes_sample=False)  # clamped to nothing
    for out in (RenderOutput.BEAUTY, RenderOutput.DEPTH, RenderOutput.NORMAL, RenderOutput.ALBEDO,
                RenderOutput.MATERIAL_ID, RenderOutput.OBJECT_ID, RenderOutput.MOTION_VECTOR):
        assert np.array_equal(tiled.pixels(out), full.pixels(out)), out
    assert np.array_equal(tiled.triangle_ids(), full.triangle_ids())
    # the oracle renders tiles through the same settings fields
    ost = O.make_settings(compile_scene(scene), W, H, max_bounces=4, frame_seed=9, clamp=64.0, tile=(32, 24, 32, 24))
    acc = np.zeros((H, W, 4), np.float32)
    O.render(O.OracleScene(compile_scene(scene)), ost, scene.camera.gpu_camera(W, H), sample_count=1, accumulation=acc)
    fresh = renderer.make_session(gs, st, camera=scene.camera)
    fresh.render_next_tile(32, 24, 32, 24, completes_sample=True)
    g = fresh.pixels(RenderOutput.BEAUTY)
    assert np.array_equal(g[:24], np.zeros_like(g[:24])) and np.array_equal(g[:, :32], np.zeros_like(g[:, :32]))
    d = np.abs(g - acc)
    assert (d > 1e-3 * (1 + np.abs(acc))).mean() <= 0.01 and d.mean() < 2e-3


def test_update_camera_matches_oracle_and_new_session(renderer):
    # updateCamera (RenderSession.swift:595-608) resets the sample count but
    # not the accumulation, so the next sample is mix(old, new, 1) =
    # old + (new - old): equal to a fresh session's first sample up to the
    # rounding of that subtraction, and equal to the oracle carrying the same
    # accumulation. AOVs are rewritten, so they equal a fresh session exactly;
    # motion vectors use the previous camera (the one before the update).
    scene = synthetic.cornell_room_scene()
    W, H = 64, 48
    build = compile_scene(scene)
    gs = renderer.upload(build)
    c = scene.camera
    moved = Camera(origin=tuple(np.add(c.origin, (0.3, -0.1, -0.3))), target=c.target, up=c.up,
                   vertical_field_of_view_degrees=50)
    st = RenderSettings(width=W, height=H, max_bounces=4, quality=FINAL, frame_seed=4)
    sess = renderer.make_session(gs, st, camera=c, samples_per_launch=2)
    sess.set_graphs(True)
    sess.render(4)
    sess.update_camera(moved)  # the previous camera becomes the old one
    assert sess.sample_count == 0
    sess.render(3)
    st2 = RenderSettings(width=W, height=H, max_bounces=4, quality=FINAL, frame_seed=4, previous_camera=c)
    fresh = renderer.make_session(gs, st2, camera=moved, samples_per_launch=2)
    fresh.render(3)
    for out in (RenderOutput.DEPTH, RenderOutput.NORMAL, RenderOutput.MOTION_VECTOR, RenderOutput.OBJECT_ID):
        assert np.array_equal(sess.pixels(out), fresh.pixels(out)), out
    assert np.array_equal(sess.triangle_ids(), fresh.triangle_ids())
    assert np.allclose(sess.pixels(), fresh.pixels(), rtol=1e-5, atol=1e-6)
    assert np.abs(sess.pixels(RenderOutput.MOTION_VECTOR)[..., :2]).max() > 0.5
    osc = O.OracleScene(build)
    acc = np.zeros((H, W, 4), np.float32)
    ost = O.make_settings(build, W, H, max_bounces=4, frame_seed=4, clamp=64.0)
    O.render(osc, ost, c.gpu_camera(W, H), sample_count=4, accumulation=acc)
    ref = O.render(osc, ost, moved.gpu_camera(W, H), previous_camera=c.gpu_camera(W, H), sample_count=3,
                   accumulation=acc)
    _check(sess, ref)
