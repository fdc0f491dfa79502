"""A SceneScript scene through the HIP path against the CPU oracle: the
script, its OBJ/PLY meshes, checker texture and Radiance HDR environment are
written here, parsed by the reader and rendered through the C ABI; IDs and
AOVs must match bit-exactly / to 1e-5 and beauty to the 1-spp tolerance of
the parity tests (test_gpu_parity.py)."""
import struct

import numpy as np
import pytest

from denrimrendererkit_amd import compile_scene
from denrimrendererkit_amd.scenescript import parse_file
from test_gpu_parity import _check_aovs, _check_beauty, _render_both

pytestmark = pytest.mark.gpu

OCTA_OBJ = """v 0 1 0
v 1 0 0
v 0 0 1
v -1 0 0
v 0 0 -1
v 0 -1 0
f 1 3 2
f 1 4 3
f 1 5 4
f 1 2 5
f 6 2 3
f 6 3 4
f 6 4 5
f 6 5 2
"""


def _write_assets(d):
    (d / "octa.obj").write_text(OCTA_OBJ)
    # A binary PLY tetrahedron.
    hdr = (b"ply\nformat binary_little_endian 1.0\nelement vertex 4\nproperty float x\nproperty float y\n"
           b"property float z\nelement face 4\nproperty list uchar uint vertex_indices\nend_header\n")
    v = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    f = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)]
    body = b"".join(struct.pack("<3f", *p) for p in v) + b"".join(struct.pack("<B3I", 3, *t) for t in f)
    (d / "tet.ply").write_bytes(hdr + body)
    # 16 x 8 flat RGBE sky: brighter upper rows.
    w, h = 16, 8
    px = bytearray()
    for y in range(h):
        for x in range(w):
            px += bytes([120 + 8 * (h - y), 140, 200 - 4 * x, 129 if y < 3 else 127])
    (d / "sky.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 8 +X 16\n" + bytes(px))
    (d / "scene.denrim").write_text("""
camera origin(0.3, 1.6, 4.2) target(0, 0.4, 0) fov(38)
environment hdri sky.hdr intensity 1.5 rotationY 0.7
texture floorTex checker 0.8 0.8 0.78 1 0.2 0.22 0.25 1 nearest
material floor 1 1 1 roughness 0.7 baseColorTexture floorTex
material red preset coating.car-red
material gold preset metal.gold roughness 0.2
material jade preset subsurface.jade-green
material lamp 1 1 1 emission 1 0.9 0.8 8
mesh octa octa.obj
mesh tet tet.ply
quad floor a(-3, 0, 3) b(3, 0, 3) c(3, 0, -3) d(-3, 0, -3) uvA(0, 0) uvB(6, 0) uvC(6, 6) uvD(0, 6)
quad lamp a(-0.5, 2.5, -0.5) b(0.5, 2.5, -0.5) c(0.5, 2.5, 0.5) d(-0.5, 2.5, 0.5)
instance octa red position(-0.9, 0.6, 0) scale(0.6, 0.6, 0.6) rotationY(0.3)
instance tet gold position(0.5, 0, 0.3) scale(0.9, 0.9, 0.9) rotationY(-0.4)
box material(jade) position(0.9, 0.3, -0.9) size(0.5, 0.6, 0.5) rotationY(0.5)
""")


def test_scenescript_scene_parity(renderer, tmp_path):
    _write_assets(tmp_path)
    scene = parse_file(str(tmp_path / "scene.denrim"))
    assert len(compile_scene(scene).triangles) == 8 + 4 + 2 + 2 + 12
    sess, ref = _render_both(renderer, scene, 160, 96, spp=1, bounces=4, seed=7)
    _check_aovs(sess, ref)
    bad = _check_beauty(sess, ref)
    print(f"scenescript scene 160x96 1spp: channels outside tol {bad:.4%}")
