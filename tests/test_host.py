"""Host-side logic and the C-ABI library (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

from denrimrendererkit_amd import (DenrimRendererError, Material, Mesh, RenderQuality, RenderScene, RenderSettings,
                                   Transform, compile_scene)
from denrimrendererkit_amd import gpu_types as gt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "denrim_hip.h")).read()
    return sorted(set(re.findall(r"\b(dr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from denrimrendererkit_amd import _lib
    lib = _lib.load()
    names = _declared_symbols()
    assert "dr_hip_session_render" in names and len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_device_error_is_typed_without_gpu():
    from denrimrendererkit_amd import DenrimRenderer
    try:
        DenrimRenderer(0)
    except DenrimRendererError as e:
        assert e.kind in ("noDevice", "missingShaderFunction")
    else:
        pytest.skip("a gfx950 device is present")


def test_record_layouts_match_reference_sizes():
    assert gt.TRIANGLE.itemsize == 192 and gt.MATERIAL.itemsize == 272 and gt.LIGHT.itemsize == 32
    assert gt.CAMERA.itemsize == 80 and gt.ENV_SAMPLE.itemsize == 8 and gt.NODE.itemsize == 48


def test_compile_errors_follow_reference():
    with pytest.raises(DenrimRendererError) as e:
        compile_scene(RenderScene())
    assert e.value.kind == "invalidScene"
    s = RenderScene()
    s.add_material(Material())
    s.add(Mesh.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), 5)
    with pytest.raises(DenrimRendererError):
        compile_scene(s)


def test_instances_share_mesh_and_set_object_ids():
    s = RenderScene()
    m = s.add_material(Material())
    box = Mesh.box((-1, -1, -1), (1, 1, 1))
    s.add(box, m)
    s.add(box, m, Transform.translation((3, 0, 0)))
    b = compile_scene(s)
    assert len(b.triangles) == 24
    assert b.triangles["objectID"].tolist() == [0] * 12 + [1] * 12
    assert b.triangles["primitiveID"].tolist() == list(range(12)) * 2
    assert np.allclose(b.triangles["v0"][12:, 0] - b.triangles["v0"][:12, 0], 3.0)


def test_transform_normal_is_inverse_transpose():
    t = Transform.scale((2, 1, 1))
    n = t.transform_normal(np.array([[1, 1, 0]], np.float32))[0]
    # (M^-1)^T n = (0.5, 1, 0) normalised.
    assert np.allclose(n, np.array([0.5, 1, 0]) / np.linalg.norm([0.5, 1, 0]), atol=1e-6)


def test_settings_defaults_and_clamps():
    s = RenderSettings()
    assert (s.width, s.height, s.max_bounces, s.quality) == (512, 512, 4, RenderQuality.PREVIEW)
    assert [RenderSettings(quality=q).resolved_sample_radiance_clamp for q in RenderQuality] == [10.0, 24.0, 64.0]
    assert RenderSettings(sample_radiance_clamp=0).resolved_sample_radiance_clamp == 0.0


def test_synthetic_workload_sizes():
    from denrimrendererkit_amd import synthetic
    b = compile_scene(synthetic.cornell_room_scene())
    assert len(b.triangles) == 8334 and len(b.lights) == 2
    small = compile_scene(synthetic.layered_terrain_scene(scale=1 / 64, hdri=(64, 32)))
    assert len(small.lights) == 64 and len(small.environment_samples) == 64 * 32
    kinds = {int(m) for m in small.triangles["materialID"]}
    assert len(kinds) > 4
