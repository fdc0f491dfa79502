"""Render settings and output kinds (Renderer/RenderSettings.swift:5-115,
Renderer/RenderOutput.swift:4-25, RenderSession.swift:8-17)."""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import Optional, Sequence

from .scene import Camera


class RenderQuality(enum.Enum):
    """RenderSettings.swift:5-27."""
    PREVIEW = 0
    INTERACTIVE = 1
    FINAL = 2

    @property
    def default_sample_radiance_clamp(self) -> float:
        return {RenderQuality.PREVIEW: 10.0, RenderQuality.INTERACTIVE: 24.0, RenderQuality.FINAL: 64.0}[self]


class RenderOutput(enum.IntEnum):
    """RenderOutput.swift:4-25, in the reference's case order; TRIANGLE_ID
    is this backend's internal parity output (the global materialized
    triangle index + 1, stored as uint32 bits in every channel)."""
    BEAUTY = 0
    DEPTH = 1
    NORMAL = 2
    ALBEDO = 3
    MATERIAL_ID = 4
    OBJECT_ID = 5
    MOTION_VECTOR = 6
    TRIANGLE_ID = 7


class RenderAccelerationMode(enum.Enum):
    """RenderSession.swift:8-17 plus the new ``hip`` backend."""
    AUTOMATIC = "automatic"
    FLAT_BVH = "flat-bvh"
    METAL_RAY_TRACING = "metal-ray-tracing"
    HIP = "hip"


class TriangleIntersection(enum.IntEnum):
    """Ray/triangle test: the reference's Moller-Trumbore
    (PathTrace.metal:1801-1852) or the watertight test of Woop, Benthin and
    Wald (JCGT 2013), which no ray slips through at a shared edge."""
    MOLLER_TRUMBORE = 0
    WATERTIGHT = 1


class EnvironmentSampling(enum.IntEnum):
    """How NEE picks an HDRI texel: the reference's CDF search
    (PathTrace.metal:3929-3950) or an alias table of the same distribution
    (one more random number per sample; equal in the limit)."""
    CDF = 0
    ALIAS = 1


class RenderDenoiser(enum.IntEnum):
    """RenderDenoiser (Denoising/Denoiser.swift:8-18) plus this backend's
    learned U-Net; APPLE_SVGF is refused as on platforms without MPS."""
    NONE = 0
    SIMPLE_SPATIAL = 1
    APPLE_SVGF = 2
    LEARNED = 3


@dataclass
class DenoiseSettings:
    """DenoiseSettings (Denoising/Denoiser.swift:20-75), same defaults."""
    denoiser: RenderDenoiser = RenderDenoiser.NONE
    radius: int = 2
    normal_sigma: float = 0.18
    depth_sigma: float = 0.03
    albedo_sigma: float = 0.25
    color_sigma: float = 1.0
    iterations: int = 3

    @staticmethod
    def none() -> "DenoiseSettings":
        return DenoiseSettings()

    @staticmethod
    def simple_spatial() -> "DenoiseSettings":
        return DenoiseSettings(denoiser=RenderDenoiser.SIMPLE_SPATIAL)

    @staticmethod
    def apple_svgf() -> "DenoiseSettings":  # Denoiser.swift:68-76
        return DenoiseSettings(RenderDenoiser.APPLE_SVGF, 2, 0.18, 0.08, 0.25, 4.0, 5)

    @staticmethod
    def learned() -> "DenoiseSettings":
        return DenoiseSettings(denoiser=RenderDenoiser.LEARNED)


@dataclass
class RenderSettings:
    """RenderSettings.swift:29-101 (defaults at :75-87)."""
    width: int = 512
    height: int = 512
    max_bounces: int = 4
    quality: RenderQuality = RenderQuality.PREVIEW
    previous_camera: Optional[Camera] = None
    transparent_background: bool = False
    shows_environment_background: bool = True
    background_color: Sequence[float] = (0.0, 0.0, 0.0)
    denoise: DenoiseSettings = field(default_factory=DenoiseSettings)
    sample_radiance_clamp: Optional[float] = None
    collects_sdf_traversal_stats: bool = False
    frame_seed: int = 0x1234ABCD  # RenderSession.swift:917
    intersection: TriangleIntersection = TriangleIntersection.MOLLER_TRUMBORE
    environment_sampling: EnvironmentSampling = EnvironmentSampling.CDF

    @property
    def resolved_sample_radiance_clamp(self) -> float:
        if self.sample_radiance_clamp is not None:
            return float(self.sample_radiance_clamp)
        return self.quality.default_sample_radiance_clamp
