"""GPU parity on seeded random scenes: combinations the fixed scenes leave
out, each drawn from one seed. A scene mixes boxes, textured quads and UV
spheres (flat or with smooth normals and texcoords) under random transforms
and instancing. Materials are drawn from the six layered kinds with random
parameters, plus Lambertian, cut-out, emissive, anisotropic metal,
thin-walled and absorbing or scattering transmission, with random
base-colour textures (nearest or bilinear) and normal maps. Lights are
emissive quads and triangles, an HDRI or a plain background, and sometimes
a sparse SDF volume is added. The camera is pinhole, thin-lens or
orthographic. Settings are random: preview, interactive or final quality,
depth 1-6, Moller-Trumbore or watertight triangles, CDF or alias HDRI
sampling, transparent or background-colour misses, and 1 or 2 samples,
batched or not.

Bar, as for the fixed scenes (test_gpu_parity): primary triangle, material
and object IDs bit-exact; depth, normal and albedo within 1e-5; beauty at
most 1% of channels outside 1e-3 * (1 + |ref|) and mean |diff| < 2e-3,
against the oracle on the same sample streams.
"""
import math

import numpy as np
import pytest

import oracle as O
from denrimrendererkit_amd import (Camera, Environment, Material, Mesh, RenderQuality, RenderScene, RenderSettings,
                                   Transform, compile_scene, synthetic)
from denrimrendererkit_amd.scene import CameraLens, QuadLight, Texture2D, TextureSamplingMode
from denrimrendererkit_amd.sdf import SDFPrimitive, build_sparse_volume
from denrimrendererkit_amd.settings import EnvironmentSampling, RenderOutput, TriangleIntersection

pytestmark = pytest.mark.gpu

SEEDS = range(48)
W, H = 64, 48


def _u(rng, a, b):
    return float(rng.uniform(a, b))


def _rgb(rng, a, b):
    return tuple(float(x) for x in rng.uniform(a, b, 3))


def _texture(rng):
    w, h = int(rng.integers(2, 17)), int(rng.integers(2, 17))
    mode = TextureSamplingMode.LINEAR if rng.uniform() < 0.5 else TextureSamplingMode.NEAREST
    px = np.ones((w * h, 4), np.float32)
    px[:, :3] = rng.uniform(0, 1, (w * h, 3))
    return Texture2D(w, h, px, sampling_mode=mode)


def _normal_map(rng):
    w = h = int(rng.integers(2, 9))
    n = rng.standard_normal((w * h, 3))
    n[:, 2] = np.abs(n[:, 2]) + 1.0
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    px = np.ones((w * h, 4), np.float32)
    px[:, :3] = 0.5 * n + 0.5
    return Texture2D(w, h, px, sampling_mode=TextureSamplingMode.LINEAR if rng.uniform() < 0.5 else 0)


def _material(rng):
    k = int(rng.integers(13))
    if k < 6:
        m = synthetic._layered_material(rng, k)
    elif k == 6:  # Lambertian
        m = Material(base_color=_rgb(rng, 0.2, 0.9), specular=0.0, roughness=_u(rng, 0.3, 1.0))
    elif k == 7:  # cut-out: primaries and shadow rays pass it
        m = Material(base_color=_rgb(rng, 0.2, 0.9), opacity=float(rng.choice([0.0005, 0.3, 0.7])))
    elif k == 8:  # emissive triangles (NEE over the mesh)
        m = Material(base_color=_rgb(rng, 0.2, 0.9), emission=_rgb(rng, 0.5, 1.0), emission_strength=_u(rng, 1, 8))
    elif k == 9:  # anisotropic metal
        m = Material(base_color=_rgb(rng, 0.4, 0.95), metallic=1.0, roughness=_u(rng, 0.1, 0.6),
                     specular_anisotropy=_u(rng, -0.9, 0.9))
    elif k == 10:  # thin-walled transmission
        m = Material(base_color=_rgb(rng, 0.4, 0.95), transmission=_u(rng, 0.5, 1.0), thin_walled=True,
                     roughness=_u(rng, 0.02, 0.5), transmission_color=_rgb(rng, 0.6, 1.0))
    elif k == 11:  # absorbing and scattering transmission
        m = Material(base_color=_rgb(rng, 0.4, 0.95), transmission=1.0, roughness=_u(rng, 0.02, 0.4),
                     transmission_absorption_color=_rgb(rng, 0.3, 1.0), transmission_absorption_distance=_u(rng, 0.2, 2),
                     volume_scattering=_u(rng, 0, 1), volume_scattering_distance=_u(rng, 0.1, 1),
                     volume_anisotropy=_u(rng, -0.5, 0.8))
    else:
        m = Material(base_color=_rgb(rng, 0.2, 0.9), roughness=_u(rng, 0.1, 0.9), metallic=_u(rng, 0, 1),
                     specular=_u(rng, 0, 1), specular_color=_rgb(rng, 0.5, 1), index_of_refraction=_u(rng, 1.2, 2.0))
    if rng.uniform() < 0.25:
        m.base_color_texture = _texture(rng)
    if rng.uniform() < 0.15:
        m.normal_map = _normal_map(rng)
    return m


def _sphere(rng):
    seg, bands = int(rng.integers(6, 24)), int(rng.integers(5, 16))
    flat = synthetic.uv_sphere(seg, bands, 1.0)
    if rng.uniform() < 0.5:
        return flat
    v = flat.vertices.astype(np.float64)
    uv = np.stack([np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi) + 0.5, np.arccos(np.clip(v[:, 1], -1, 1)) / math.pi], 1)
    return Mesh(flat.vertices, flat.indices, normals=v, texcoords=uv)


def random_scene(seed):
    rng = np.random.default_rng(0xC0FFEE + 7919 * seed)
    look = tuple(float(x) for x in rng.uniform(-0.4, 0.4, 3))
    eye = (_u(rng, -2.5, 2.5), _u(rng, 0.5, 3.0), _u(rng, 3.0, 5.5))
    cam_kind = int(rng.integers(3))
    if cam_kind == 0:
        cam = Camera(origin=eye, target=look, vertical_field_of_view_degrees=_u(rng, 30, 70))
    elif cam_kind == 1:
        cam = Camera(origin=eye, target=look, vertical_field_of_view_degrees=_u(rng, 30, 70),
                     lens=CameraLens(focus_distance=_u(rng, 2.0, 6.0), aperture_radius=_u(rng, 0.02, 0.15)))
    else:
        cam = Camera(origin=eye, target=look, orthographic_scale=_u(rng, 3.0, 6.0))
    env = None
    if rng.uniform() < 0.6:
        env = Environment(texture=synthetic.lognormal_hdri(rng, 64, 32, sun_peak=_u(rng, 10, 500)),
                          intensity=_u(rng, 0.3, 1.5), rotation_y=_u(rng, 0, 2 * math.pi),
                          max_radiance=float(rng.choice([0.0, 0.0, 20.0])))
    s = RenderScene(camera=cam, environment=env) if env else RenderScene(camera=cam)
    fw = _u(rng, 1.5, 4.0)
    s.add(Mesh.quad((-fw, 0, -fw), (-fw, 0, fw), (fw, 0, fw), (fw, 0, -fw),
                    texcoords=((0, 0), (0, 2), (2, 2), (2, 0))), s.add_material(_material(rng)))
    box = Mesh.box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    spheres = [_sphere(rng) for _ in range(2)]
    for _ in range(int(rng.integers(2, 7))):
        m = s.add_material(_material(rng))
        mesh = spheres[int(rng.integers(2))] if rng.uniform() < 0.6 else box
        sc = _u(rng, 0.2, 0.8)
        xf = (Transform.translation((_u(rng, -2, 2), _u(rng, 0.2, 1.5), _u(rng, -2, 2))) * Transform.rotation_y(_u(rng, 0, 6))
              * Transform.rotation_x(_u(rng, 0, 6)) * Transform.scale((sc, sc * _u(rng, 0.5, 1.5), sc)))
        s.add(mesh, m, xf)
    for _ in range(int(rng.integers(0 if env else 1, 3))):
        y, x, z, w = _u(rng, 2.0, 3.5), _u(rng, -2, 2), _u(rng, -2, 2), _u(rng, 0.3, 1.2)
        s.add_quad_light(QuadLight((x - w, y, z - w), (x + w, y, z - w), (x + w, y, z + w), (x - w, y, z + w),
                                   color=_rgb(rng, 0.6, 1.0), intensity=_u(rng, 2, 15)))
    if rng.uniform() < 0.3:
        vm = s.add_material(_material(rng))
        prims = [SDFPrimitive("sphere", vm, tuple(rng.uniform(0.3, 0.7, 3)), _u(rng, 0.1, 0.3)),
                 SDFPrimitive("torus", vm, tuple(rng.uniform(0.3, 0.7, 3)), _u(rng, 0.15, 0.3), axis=(0, 1, 0),
                              minor_radius=_u(rng, 0.04, 0.1))]
        s.add_sparse_volume(build_sparse_volume(prims, dims=(32, 32, 32), narrow_band=2 / 31), vm,
                            Transform.translation((_u(rng, -1.5, 0.5), _u(rng, 0.0, 0.8), _u(rng, -1.5, 0.5))))
    q = float(rng.uniform())
    quality = RenderQuality.PREVIEW if q < 0.15 else RenderQuality.INTERACTIVE if q < 0.3 else RenderQuality.FINAL
    spp = int(rng.integers(1, 3)) if quality == RenderQuality.FINAL else 1
    settings = dict(quality=quality, spp=spp, spl=int(rng.choice([0, spp])), max_bounces=int(rng.integers(1, 7)),
                    intersection=int(rng.integers(2)), environment_sampling=int(rng.integers(2)) if env else 0,
                    transparent_background=bool(rng.uniform() < 0.25),
                    shows_environment_background=bool(rng.uniform() < 0.8),
                    background=_rgb(rng, 0, 0.3), frame_seed=int(rng.integers(1 << 31)))
    return s, settings


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_parity(renderer, seed):
    scene, st = random_scene(seed)
    build = compile_scene(scene)
    rs = RenderSettings(width=W, height=H, max_bounces=st["max_bounces"], quality=st["quality"],
                        frame_seed=st["frame_seed"], transparent_background=st["transparent_background"],
                        shows_environment_background=st["shows_environment_background"],
                        background_color=st["background"],
                        intersection=[TriangleIntersection.MOLLER_TRUMBORE, TriangleIntersection.WATERTIGHT][st["intersection"]],
                        environment_sampling=[EnvironmentSampling.CDF, EnvironmentSampling.ALIAS][st["environment_sampling"]])
    sess = renderer.make_session(renderer.upload(build), rs, camera=scene.camera, samples_per_launch=st["spl"])
    sess.render(st["spp"])
    ost = O.make_settings(build, W, H, max_bounces=st["max_bounces"], frame_seed=st["frame_seed"],
                          clamp=rs.resolved_sample_radiance_clamp, transparent_background=st["transparent_background"],
                          shows_environment_background=st["shows_environment_background"], background=st["background"],
                          quality=st["quality"].value, intersection=st["intersection"],
                          environment_sampling=st["environment_sampling"])
    ref = O.render(O.OracleScene(build), ost, scene.camera.gpu_camera(W, H), sample_count=st["spp"])
    print(f"seed {seed}: {len(build.triangles)} tris, {len(build.volumes)} volumes, {st}")
    tri, ref_tri = sess.triangle_ids(), ref.aovs[6][..., 0].view(np.uint32)
    assert np.array_equal(tri, ref_tri), f"triangle IDs differ at {int((tri != ref_tri).sum())} px"
    for out, k in ((RenderOutput.MATERIAL_ID, 3), (RenderOutput.OBJECT_ID, 4)):
        assert np.array_equal(sess.pixels(out), ref.aovs[k]), out
    for out, k in ((RenderOutput.DEPTH, 0), (RenderOutput.NORMAL, 1), (RenderOutput.ALBEDO, 2)):
        d = np.abs(sess.pixels(out) - ref.aovs[k]).max()
        assert d <= 1e-5, (out, d)
    g, r = sess.pixels(RenderOutput.BEAUTY), ref.accumulation
    assert np.isfinite(g).all()
    diff = np.abs(g - r)
    bad = (diff > 1e-3 * (1.0 + np.abs(r))).mean()
    print(f"  beauty: {bad:.4%} channels outside tol, mean |diff| {diff.mean():.2e}")
    assert bad <= 0.01 and diff.mean() < 2e-3, (seed, bad, diff.mean())
