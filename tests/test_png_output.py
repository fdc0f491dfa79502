"""Output side (PNGWriter.swift restated in denrimrendererkit_amd/image.py):
known answers of each RenderOutput's display encoding and a PNG file that
decodes (zlib + PNG chunks, read back with Pillow) to the same RGBA8."""
import io
import os

import numpy as np
import pytest
from PIL import Image

from denrimrendererkit_amd.image import id_color, png_bytes, visualized_rgba8, write_png
from denrimrendererkit_amd.scene import DenrimRendererError
from denrimrendererkit_amd.settings import RenderOutput


def _px(vals):
    return np.asarray(vals, np.float32).reshape(1, -1, 4)


def test_beauty_and_display_linear_known_answers():
    b = visualized_rgba8(_px([[0, 0, 0, 1], [1e6, 1e6, 1e6, 0.5], [-1, 0.18, 2, 1]]), RenderOutput.BEAUTY)
    assert b[0, 0].tolist() == [0, 0, 0, 255] and b[0, 1].tolist() == [255, 255, 255, 127]
    x = np.float32(0.18) * np.float32(0.8)
    m = (x * (np.float32(2.51) * x + np.float32(0.03))) / (x * (np.float32(2.43) * x + np.float32(0.59)) + np.float32(0.14))
    assert b[0, 2, 0] == 0 and b[0, 2, 1] == int(np.float32(m ** np.float32(1 / 2.2)) * 255)
    n = visualized_rgba8(_px([[0.5, 1.0, 0.0, 0.3]]), RenderOutput.NORMAL)
    assert n[0, 0].tolist() == [int(np.float32(0.5) ** np.float32(1 / 2.2) * 255), 255, 0, 255]
    a = visualized_rgba8(_px([[0.5, 1.0, 0.0, 0.3]]), RenderOutput.ALBEDO)
    assert a[0, 0, 3] == int(np.float32(0.3) * 255)


def test_depth_id_motion_known_answers():
    d = visualized_rgba8(_px([[2, 0, 0, 1], [4, 0, 0, 1], [0, 0, 0, 0], [np.inf, 0, 0, 1]]), RenderOutput.DEPTH)
    assert d[0, 0].tolist() == [25, 25, 25, 255] and d[0, 1].tolist() == [255, 255, 255, 255]
    assert d[0, 2].tolist() == [0, 0, 0, 255] and d[0, 3].tolist() == [0, 0, 0, 255]
    assert visualized_rgba8(_px([[0, 0, 0, 0]]), RenderOutput.DEPTH)[0, 0].tolist() == [0, 0, 0, 0]
    ids = visualized_rgba8(_px([[3, 0, 0, 1], [0, 0, 0, 0]]), RenderOutput.MATERIAL_ID)
    # colorForID(3) by hand: the reference's PCG-style hash, channels 64 + (h & 0x7F).
    h = (3 * 747796405) & 0xFFFFFFFF
    h = (h + 2891336453) & 0xFFFFFFFF
    h = (((h >> ((h >> 28) + 4)) ^ h) * 277803737) & 0xFFFFFFFF
    h = (h >> 22) ^ h
    assert ids[0, 0, :3].tolist() == [64 + (h & 0x7F), 64 + ((h >> 8) & 0x7F), 64 + ((h >> 16) & 0x7F)]
    assert ids[0, 1].tolist() == [0, 0, 0, 255]
    assert id_color(np.array([3]))[0].tolist() == ids[0, 0, :3].tolist()
    mv = visualized_rgba8(_px([[4, -2, 0, 0], [0, 0, 0, 0]]), RenderOutput.MOTION_VECTOR)
    assert mv[0, 0].tolist() == [255, 63, 128, 255] and mv[0, 1].tolist() == [127, 127, 128, 255]


def test_png_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (17, 23, 4), dtype=np.uint8)
    back = np.asarray(Image.open(io.BytesIO(png_bytes(img))).convert("RGBA"))
    assert np.array_equal(back, img)
    p = os.path.join(tmp_path, "x.png")
    write_png(p, img)
    assert np.array_equal(np.asarray(Image.open(p).convert("RGBA")), img)
    with pytest.raises(DenrimRendererError):
        write_png(os.path.join(tmp_path, "missing-dir", "x.png"), img)
