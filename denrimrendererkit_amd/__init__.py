"""MI355X (gfx950) backend of DenrimRendererKit's progressive path tracer.

The scene model and settings mirror the reference's Swift API; rendering runs
in hand-written HIP kernels in ``libdenrim_hip.so`` behind the C ABI declared
in ``include/denrim_hip.h``.
"""
from .scene import (AccelerationBuild, Camera, CameraLens, DenrimRendererError, Environment, Material, Mesh,
                    QuadLight, RenderScene, Texture2D, TextureSamplingMode, Transform, compile_scene)
from .settings import (DenoiseSettings, EnvironmentSampling, RenderAccelerationMode, RenderDenoiser, RenderOutput,
                       RenderQuality, RenderSettings,
                       TriangleIntersection)


def __getattr__(name):
    # The renderer needs the built library; import it lazily so the scene
    # model stays usable (and testable) without a GPU build.
    if name in ("DenrimRenderer", "RenderSession", "GpuScene", "RenderServer"):
        from . import renderer
        return getattr(renderer, name)
    raise AttributeError(name)


__all__ = [
    "AccelerationBuild", "Camera", "CameraLens", "DenrimRendererError", "Environment", "Material", "Mesh",
    "QuadLight", "RenderScene", "Texture2D", "TextureSamplingMode", "Transform", "compile_scene",
    "RenderAccelerationMode", "RenderOutput", "RenderQuality", "RenderSettings", "DenoiseSettings", "RenderDenoiser",
    "TriangleIntersection", "EnvironmentSampling",
    "DenrimRenderer", "RenderSession", "GpuScene", "RenderServer",
]
