"""SceneScript reader, mesh importers and image loaders on the CPU.

The cases follow the reference's SceneScriptTests
(Tests/DenrimRendererKitTests/SceneScriptTests.swift) with scripts and
assets written here: geometry commands and their named/grouped argument
forms, material keywords and presets, render defaults, textures (solid,
checker, image, normalFrom), environment images, OBJ/PLY meshes relative to
the script, includes and every error case. File formats are pinned by
known answers: Radiance RGBE texels are m * 2^(e - 136) (flat and RLE
scanlines), sRGB bytes decode by the sRGB curve, PLY ascii and binary of the
same mesh give the same arrays, OBJ polygons fan from their first vertex.
The reference's own example scripts parse when the reference checkout is
present (skipped otherwise).
"""
import math
import os
import struct

import numpy as np
import pytest

from denrimrendererkit_amd import compile_scene, materials, scenescript
from denrimrendererkit_amd.mesh_import import MeshLoadingError, load_mesh, parse_obj, parse_ply
from denrimrendererkit_amd.scene import TextureSamplingMode
from denrimrendererkit_amd.scenescript import SceneScriptError, parse, parse_file
from denrimrendererkit_amd.settings import RenderAccelerationMode, RenderDenoiser, RenderOutput, RenderQuality
from denrimrendererkit_amd.texture_io import load_radiance_hdr, load_texture

f32 = np.float32
REF_EXAMPLES = "/root/reference/Examples/SceneScripts"

BASIC = """
# comment line
camera origin(0, 1, 4) target(0, 0.5, 0) fov(40)   # trailing comment

material floor 0.8 0.8 0.8 roughness 0.9
material lamp 1 1 1 1 0.9 0.8 6
quad floor a(-2, 0, 2) b(2, 0, 2) c(2, 0, -2) d(-2, 0, -2)
quad lamp -0.5 2 -0.5 0.5 2 -0.5 0.5 2 0.5 -0.5 2 0.5
box floor 0 0.5 0 1 1 1
box material(floor) position(1, 0.25, 0) size(0.5, 0.5, 0.5) rotationY(0.4)
"""


def test_builds_renderable_scene_and_ignores_comments():
    s = parse(BASIC)
    assert len(s.materials) == 2 and len(s.mesh_instances) == 4
    assert list(s.camera.origin) == [0, 1, 4] and s.camera.vertical_field_of_view_degrees == f32(40)
    lamp = s.materials[1]
    assert tuple(lamp.emission) == (1, f32(0.9), f32(0.8)) and lamp.emission_strength == 6
    b = compile_scene(s)
    assert len(b.triangles) == 2 + 2 + 12 + 12
    assert len(b.lights) == 2  # the emissive quad's two triangles
    # The grouped box: translation * rotationY, centred size.
    t = s.mesh_instances[3].transform.matrix
    assert np.allclose(t[:3, 3], [1, 0.25, 0]) and abs(t[0, 0] - math.cos(0.4)) < 1e-6
    v = s.mesh_instances[3].mesh.vertices
    assert np.allclose(v.min(0), -0.25) and np.allclose(v.max(0), 0.25)


def test_named_grouped_geometry_arguments(tmp_path):
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    src = """
material m 0.5 0.5 0.5
mesh t tri.obj
quad material m a 0 0 0 b 1 0 0 c 1 1 0 d 0 1 0 uvA(0, 0) uvB(2, 0) uvC(2, 2) uvD(0, 2)
instance mesh(t) material(m) position(1, 2, 3) scale(2, 2, 2) rotationX(0.1) rotationY(0.2) rotationZ(0.3)
instance t m 0 0 0 1 1 1 0.5
"""
    s = parse(src, base_dir=str(tmp_path))
    q = s.mesh_instances[0].mesh
    assert np.array_equal(q.texcoords, np.array([[0, 0], [2, 0], [2, 2], [0, 2]], f32))
    from denrimrendererkit_amd.scene import Transform
    want = (Transform.translation([1, 2, 3]) * Transform.rotation_x(f32(0.1)) * Transform.rotation_y(f32(0.2))
            * Transform.rotation_z(f32(0.3)) * Transform.scale([2, 2, 2]))
    assert np.array_equal(s.mesh_instances[1].transform.matrix, want.matrix)
    want7 = Transform.translation([0, 0, 0]) * Transform.rotation_y(f32(0.5)) * Transform.scale([1, 1, 1])
    assert np.array_equal(s.mesh_instances[2].transform.matrix, want7.matrix)


def test_material_parameters():
    src = ("material m 0.2 0.3 0.4 roughness 0.25 metallic 0.5 specular 0.6 specularColor 0.9 0.8 0.7 ior 1.4 "
           "anisotropy 0.3 clearcoat 0.7 clearcoatColor 1 0.9 0.8 clearcoatAttenuationColor 0.5 0.6 0.7 "
           "clearcoatThickness 0.2 clearcoatRoughness 0.05 clearcoatIOR 1.6 thinFilm 0.4 thinFilmThickness 500 "
           "thinFilmIOR 1.3 sheen 0.5 sheenColor 0.1 0.2 0.3 sheenRoughness 0.6 subsurface 0.8 subsurfaceColor 0.9 0.5 "
           "0.4 subsurfaceRadius 1 0.5 0.25 subsurfaceScale 0.3 subsurfaceAnisotropy 0.1 opacity 0.9 transmission 0.5 "
           "transmissionColor 0.7 0.8 0.9 transmissionRoughness 0.2 transmissionIOR 1.33 absorptionColor 0.6 0.7 0.8 "
           "absorptionDistance 2 thinWalled 1 volumeScattering 0.4 volumeScatteringColor 0.3 0.3 0.3 "
           "volumeScatteringDistance 0.7 volumeAnisotropy 0.2 emission 1 0.5 0.25 3")
    m = parse(src + "\nquad m a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)").materials[0]
    expect = dict(roughness=0.25, metallic=0.5, specular=0.6, specular_color=(0.9, 0.8, 0.7), index_of_refraction=1.4,
                  specular_anisotropy=0.3, clearcoat=0.7, clearcoat_color=(1, 0.9, 0.8),
                  clearcoat_attenuation_color=(0.5, 0.6, 0.7), clearcoat_thickness=0.2, clearcoat_roughness=0.05,
                  clearcoat_index_of_refraction=1.6, thin_film=0.4, thin_film_thickness_nanometers=500,
                  thin_film_index_of_refraction=1.3, sheen=0.5, sheen_color=(0.1, 0.2, 0.3), sheen_roughness=0.6,
                  subsurface=0.8, subsurface_color=(0.9, 0.5, 0.4), subsurface_radius=(1, 0.5, 0.25),
                  subsurface_scale=0.3, subsurface_anisotropy=0.1, opacity=0.9, transmission=0.5,
                  transmission_color=(0.7, 0.8, 0.9), transmission_roughness=0.2,
                  transmission_index_of_refraction=1.33, transmission_absorption_color=(0.6, 0.7, 0.8),
                  transmission_absorption_distance=2, thin_walled=True, volume_scattering=0.4,
                  volume_scattering_color=(0.3, 0.3, 0.3), volume_scattering_distance=0.7, volume_anisotropy=0.2,
                  emission=(1, 0.5, 0.25), emission_strength=3, base_color=(0.2, 0.3, 0.4))
    for k, v in expect.items():
        got = getattr(m, k)
        if isinstance(v, tuple):
            assert tuple(got) == tuple(float(f32(x)) for x in v), k
        elif isinstance(v, bool):
            assert got is v, k
        else:
            assert got == float(f32(v)), k


def test_render_defaults(tmp_path):
    p = tmp_path / "s.denrim"
    p.write_text("render output out/img.png size 1080p spp 32 quality interactive maxBounces 3 backend flat-bvh "
                 "clamp 12 transparentBackground denoise simple sdfResolution 48\n"
                 "render defaults aov depth width 300\n")
    d = parse_file(str(p)).render_defaults
    assert d.output_path == os.path.join(str(tmp_path), "out", "img.png")
    assert (d.width, d.height, d.samples, d.max_bounces) == (300, 1080, 32, 3)
    assert d.quality is RenderQuality.INTERACTIVE and d.acceleration_mode is RenderAccelerationMode.FLAT_BVH
    assert d.sample_radiance_clamp == 12 and d.transparent_background is True and d.sdf_resolution == 48
    assert d.denoise.denoiser == RenderDenoiser.SIMPLE_SPATIAL and d.output == RenderOutput.DEPTH
    st = d.render_settings()
    assert (st.width, st.height, st.max_bounces, st.quality) == (300, 1080, 3, RenderQuality.INTERACTIVE)
    assert parse("render size 64").render_defaults.width == 64
    with pytest.raises(SceneScriptError) as e:
        parse("render quality ultra")
    assert e.value == SceneScriptError("invalidArgumentCount", "render", 1)


def test_builtin_presets_and_plain_materials():
    s = parse("material a preset metal.gold roughness 0.2\nmaterial b builtin Glass.Clear\nmaterial c 0.1 0.2 0.3")
    gold = materials.material("metal.gold")
    assert s.materials[0].metallic == 1 and tuple(s.materials[0].base_color) == tuple(gold.base_color)
    assert s.materials[0].roughness == float(f32(0.2)) and gold.roughness == 0.06  # the preset is copied
    assert s.materials[1].transmission == 1.0 and s.materials[1].transmission_index_of_refraction == 1.45
    assert tuple(s.materials[2].base_color) == tuple(float(f32(x)) for x in (0.1, 0.2, 0.3))
    assert len(materials.PRESETS) == 30 and len(set(materials.IDENTIFIERS)) == 30
    # Lookup is case-insensitive and reads '_' as '-' (BuiltInMaterialLibrary.swift:398-409).
    assert materials.material("METAL.BRUSHED_ALUMINUM").specular_anisotropy == 0.72
    with pytest.raises(SceneScriptError) as e:
        parse("material x preset metal.unobtanium")
    assert e.value == SceneScriptError("unknownMaterialPreset", "metal.unobtanium", 1)


def test_texture_bindings_checker_solid_and_normal_from():
    s = parse("texture c checker 1 0 0 1 0 0 1 1 linear\ntexture s solid 0.5 0.5 0.5 1\n"
              "texture n normalFrom c strength 2\n"
              "material m 1 1 1 baseColorTexture c normalMap n\nquad m a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)")
    m = s.materials[0]
    px = m.base_color_texture.pixels
    assert m.base_color_texture.sampling_mode == TextureSamplingMode.LINEAR
    assert np.array_equal(px, np.array([[1, 0, 0, 1], [0, 0, 1, 1], [0, 0, 1, 1], [1, 0, 0, 1]], f32))
    # normalFrom: central differences of luminance, x wrapped, y clamped.
    lr = f32(0.2126) - f32(0.0722)  # lum(red) - lum(blue)
    n = m.normal_map.pixels.reshape(2, 2, 4)
    # texel (0,0): left = right = texel (0,1) (wrap), so dx = 0; down = itself, up = (1,0).
    v = np.array([0, (f32(0.2126) - f32(0.0722)) * 2, 1], f32)
    v = v / np.sqrt(v @ v)
    assert np.allclose(n[0, 0, :3], v * 0.5 + 0.5, atol=1e-6) and lr > 0
    b = compile_scene(s)
    assert b.texture_descriptors.shape[0] >= 2


def _write_hdr(path, rows_flat, rows_rle):
    """Two-row 9-wide RGBE file: row 0 flat, row 1 new-style RLE."""
    w = 9
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 9\n")
    out += bytes(rows_flat.reshape(-1))
    out += bytes([2, 2, 0, w])
    for c in range(4):
        ch = rows_rle[:, c]
        # one run of the first 5 values (all equal), then 4 literals
        assert (ch[:5] == ch[0]).all()
        out += bytes([128 + 5, ch[0], 4]) + bytes(ch[5:])
    path.write_bytes(bytes(out))


def test_radiance_hdr_known_answers(tmp_path):
    rng = np.random.default_rng(3)
    flat = rng.integers(0, 256, (9, 4)).astype(np.uint8)
    flat[0] = [1, 2, 3, 0]  # exponent 0: black
    flat[1, :2] = [2, 2]    # only a scanline that starts 2 2 is RLE; this row starts with flat[0]
    rle = np.empty((9, 4), np.uint8)
    rle[:5] = [200, 100, 50, 130]
    rle[5:] = rng.integers(0, 256, (4, 4))
    rle[5:, 3] = [128, 129, 140, 0]
    _write_hdr(tmp_path / "e.hdr", flat, rle)
    w, h, px = load_radiance_hdr((tmp_path / "e.hdr").read_bytes())
    assert (w, h) == (9, 2)
    px = px.reshape(2, 9, 4)
    for y, rows in ((0, flat), (1, rle)):
        for x in range(9):
            r, g, b, e = (int(v) for v in rows[x])
            want = [0.0, 0.0, 0.0] if e == 0 else [math.ldexp(c, e - 136) for c in (r, g, b)]
            assert list(px[y, x, :3]) == want and px[y, x, 3] == 1.0
    # The scene command: environment image relative to the script.
    (tmp_path / "s.denrim").write_text("environment hdri e.hdr intensity 2 rotationY 0.5 maxRadiance 0\n"
                                       "material m 1 1 1\nquad m a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)\n")
    env = parse_file(str(tmp_path / "s.denrim")).environment
    assert env.texture.width == 9 and env.intensity == 2 and env.rotation_y == float(f32(0.5))
    assert env.max_radiance == 0 and env.texture.sampling_mode == TextureSamplingMode.LINEAR
    with pytest.raises(SceneScriptError) as e:
        parse_file_src(tmp_path, "environment image missing.hdr\n")
    assert e.value == SceneScriptError("environmentLoadFailed", "missing.hdr", 1)


def parse_file_src(tmp_path, src):
    p = tmp_path / "x.denrim"
    p.write_text(src)
    return parse_file(str(p))


def test_image_texture_srgb_and_linear(tmp_path):
    from PIL import Image
    a = np.array([[[0, 0, 0, 255], [255, 255, 255, 255]], [[188, 10, 128, 255], [64, 32, 16, 255]]], np.uint8)
    Image.fromarray(a, "RGBA").save(tmp_path / "t.png")
    t = load_texture(str(tmp_path / "t.png"), "srgb")
    lin = load_texture(str(tmp_path / "t.png"), "linear", TextureSamplingMode.NEAREST)
    assert (t.width, t.height) == (2, 2) and t.sampling_mode == TextureSamplingMode.LINEAR
    assert np.array_equal(lin.pixels[:, :3], (a.reshape(-1, 4)[:, :3] / f32(255)).astype(f32))

    def srgb(c):  # the sRGB transfer function, IEC 61966-2-1
        return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4

    for k, v in enumerate(a.reshape(-1, 4)):
        for c in range(3):
            assert abs(float(t.pixels[k, c]) - srgb(v[c] / 255)) < 2e-7
    s = parse_file_src(tmp_path, "texture img image t.png color linear sampler nearest\n")
    assert s is not None
    with pytest.raises(SceneScriptError) as e:
        parse_file_src(tmp_path, "texture img image nothere.png\n")
    assert e.value == SceneScriptError("textureLoadFailed", "nothere.png", 1)


def _ply_header(fmt, nv, nf, extra_vertex=""):
    return (f"ply\nformat {fmt} 1.0\ncomment test\nelement vertex {nv}\nproperty float x\nproperty float y\n"
            f"property float z\n{extra_vertex}property float nx\nproperty float ny\nproperty float nz\n"
            f"property float s\nproperty float t\nelement face {nf}\nproperty list uchar int vertex_indices\n"
            f"element extra 1\nproperty int junk\nend_header\n")


def test_ply_ascii_binary_and_obj_agree():
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1]], f32)
    N = np.array([[0, 0, 1]] * 5, f32)
    UV = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]], f32)
    faces = [[0, 1, 2, 3], [0, 1, 4]]
    txt = _ply_header("ascii", 5, 2, "property uchar red\n")
    for k in range(5):
        txt += " ".join(str(float(x)) for x in V[k]) + " 7 " + " ".join(str(float(x)) for x in (*N[k], *UV[k])) + "\n"
    txt += "4 0 1 2 3\n3 0 1 4\n9\n"
    a = parse_ply(txt.encode())
    hdr = _ply_header("binary_little_endian", 5, 2, "property uchar red\n").encode()
    body = b"".join(struct.pack("<3fB5f", *V[k], 7, *N[k], *UV[k]) for k in range(5))
    body += struct.pack("<B4i", 4, 0, 1, 2, 3) + struct.pack("<B3i", 3, 0, 1, 4) + struct.pack("<i", 9)
    b = parse_ply(hdr + body)
    for m in (a, b):
        assert np.array_equal(m.vertices, V) and np.array_equal(m.normals, N) and np.array_equal(m.texcoords, UV)
        assert list(m.indices) == [0, 1, 2, 0, 2, 3, 0, 1, 4]  # fanned from the first vertex
    obj = "# quad + tri\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1\nvt 0 0\nvn 0 0 2\n" \
          "f 1/1/1 2/1/1 3/1/1 4/1/1\nf -5/1/1 -4/1/1 -1\n"
    o = parse_obj(obj.encode())
    assert o.triangle_count == 3
    # (position, texcoord, normal) triples are vertices in first-use order:
    # -5/1/1 and -4/1/1 are the triples of 1/1/1 and 2/1/1 again, the bare -1
    # (position 5, no texcoord or normal) is a new vertex.
    assert list(o.indices) == [0, 1, 2, 0, 2, 3, 0, 1, 4] and len(o.vertices) == 5
    assert np.array_equal(o.normals[0], [0, 0, 1]) and np.array_equal(o.normals[4], [0, 0, 0])
    assert np.array_equal(o.vertices[4], [0.5, 0.5, 1]) and np.array_equal(o.texcoords[4], [0, 0])
    # OBJ numbers: sign, fraction and exponent forms.
    n = parse_obj(b"v -1.5 2.25e1 +3E-1\nv 0 0 0\nv 1 1 1\nf 1 2 3\n")
    assert list(n.vertices[0]) == [-1.5, 22.5, float(f32(0.3))]
    with pytest.raises(MeshLoadingError, match="Invalid mesh index '0'"):
        parse_obj(b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")
    with pytest.raises(MeshLoadingError, match="face requires at least three vertices"):
        parse_obj(b"v 0 0 0\nv 1 0 0\nf 1 2\n")


def test_mesh_relative_to_script_and_flipv(tmp_path):
    (tmp_path / "x.stl").write_bytes(b"solid x\n")
    with pytest.raises(MeshLoadingError, match="Unsupported mesh format 'stl'"):
        load_mesh(str(tmp_path / "x.stl"))
    (tmp_path / "m").mkdir()
    (tmp_path / "m" / "t.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0.25\nvt 1 0.25\nvt 0 1\n"
                                          "f 1/1 2/2 3/3\n")
    s = parse_file_src(tmp_path, "material m 1 1 1\nmesh a m/t.obj\nmesh b m/t.obj flipV\n"
                                 "instance a m position(0,0,0) scale(1,1,1)\ninstance b m position(0,0,1) scale(1,1,1)\n")
    ta, tb = s.mesh_instances[0].mesh.texcoords, s.mesh_instances[1].mesh.texcoords
    assert np.array_equal(tb[:, 1], 1 - ta[:, 1]) and np.array_equal(tb[:, 0], ta[:, 0])
    with pytest.raises(SceneScriptError) as e:
        parse_file_src(tmp_path, "mesh a m/none.ply\n")
    assert e.value == SceneScriptError("meshLoadFailed", "m/none.ply", 1)


def test_includes_and_cycles(tmp_path):
    (tmp_path / "mats.denrim").write_text("material m 0.5 0.5 0.5\n")
    (tmp_path / "loop.denrim").write_text("include loop.denrim\n")
    s = parse_file_src(tmp_path, "include mats.denrim\nquad m a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)\n")
    assert len(s.materials) == 1 and len(s.mesh_instances) == 1
    with pytest.raises(SceneScriptError) as e:
        parse_file_src(tmp_path, "include loop.denrim\n")
    assert e.value == SceneScriptError("includeCycle", "loop.denrim", 1)
    with pytest.raises(SceneScriptError) as e:
        parse("include x.denrim")
    assert e.value == SceneScriptError("includeResolverMissing", "x.denrim", 1)
    fragments = {"a": "material m 1 1 1\ninclude b", "b": "material n 0 0 0"}
    s = parse("include a\nquad n a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)", include_resolver=fragments.__getitem__)
    assert len(s.materials) == 2


@pytest.mark.parametrize("src,err", [
    ("quad nope a(0,0,0) b(1,0,0) c(1,1,0) d(0,1,0)", ("unknownMaterial", "nope", 1)),
    ("material m 1 1 1 baseColorTexture tex", ("unknownTexture", "tex", 1)),
    ("material m 1 1 1\ninstance ghost m position(0,0,0) scale(1,1,1)", ("unknownMesh", "ghost", 2)),
    ("material m 1 x 1", ("invalidNumber", "x", 1)),
    ("render spp 3.5", ("invalidNumber", "3.5", 1)),
    ("teleport now", ("unknownCommand", "teleport", 1)),
    ("camera origin(0,0,0) fov(40)", ("invalidArgumentCount", "camera", 1)),
    ("material m 1 1 1\nquad m a(0,0,0) b(1,0,0) c(1,1,0)", ("invalidArgumentCount", "quad", 2)),
    ("material m 1 1 1\nbox m position(0,0,0) rotationY(1)", ("invalidArgumentCount", "box", 2)),
    # Eight tokens select the positional box form (x y z sx sy sz), as in the reference.
    ("material m 1 1 1\nbox m position(0,0,0)", ("invalidNumber", "position", 2)),
])
def test_error_cases(src, err):
    with pytest.raises(SceneScriptError) as e:
        parse(src)
    assert e.value == SceneScriptError(*err)


def test_sdf_volume_command():
    s = parse("material a 0.8 0.2 0.2\nmaterial b 0.2 0.2 0.8\n"
              "sdf blob sparse material a resolution 24 brickSize 8 narrowBand 0.3 boundsMin -1 -1 -1 boundsMax 1 1 1 "
              "sphere a radius 0.6 box material b size(0.6, 0.6, 0.6) position(0.4, 0, 0) cornerRadius 0.05 "
              "smooth 0.1 subtract cylinder a radius 0.2 halfHeight 0.8\n"
              "volume solid dense material b resolution 16 sphere radius 0.5 position 0 0.1 0\n")
    assert len(s.sparse_volume_instances) == 1 and len(s.volume_instances) == 1
    sv = s.sparse_volume_instances[0].volume
    assert sv.brick_count > 0
    b = compile_scene(s)
    assert len(b.volume_bricks) == sv.brick_count and len(b.volumes) == 2
    with pytest.raises(SceneScriptError) as e:
        parse("material a 1 1 1\nsdf x sparse sphere a radius 0.5 opacity 0.5")
    assert e.value.kind == "unsupported"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference checkout not present")
def test_reference_example_scripts_parse():
    s = parse_file(os.path.join(REF_EXAMPLES, "MaterialTestBall", "material-testball.denrim"))
    b = compile_scene(s)
    # Three PLY meshes of 13,312 + 17,536 + 49,920 triangles, 5 quads.
    assert len(b.triangles) == 13312 + 17536 + 49920 + 10
    assert (s.environment.texture.width, s.environment.texture.height) == (1024, 512)
    assert s.environment.intensity == 3 and len(s.materials) == 7
    assert s.materials[0].metallic == 1 and f32(s.materials[0].specular_anisotropy) == f32(0.72)
    v = parse_file(os.path.join(REF_EXAMPLES, "MaterialVariants", "material-variants.denrim"))
    assert len(v.mesh_instances) == 3 + 5
    d = parse_file(os.path.join(REF_EXAMPLES, "SDF", "shadertoy-material-testball.denrim"), sdf_resolution_override=20)
    assert d.render_defaults.samples == 64 and d.render_defaults.sdf_resolution == 72
    assert len(d.sparse_volume_instances) == 2
