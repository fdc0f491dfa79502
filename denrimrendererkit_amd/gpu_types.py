"""Flat scene records in the reference's GPU layouts.

These numpy dtypes restate, byte for byte, the structs of
``Sources/DenrimRendererKit/Metal/GPUTypes.swift`` (GPUCamera :4,
GPUTriangle :12, GPUMaterial :30, GPUTextureDescriptor :51,
GPULightRecord :55, GPUEnvironmentSample :63, GPUAccelerationNode :145).
The HIP backend's C ABI (``include/denrim_hip.h``) takes arrays of exactly
these records, which is what the reference's ``AccelerationBuild``
(``Acceleration/AccelerationBuild.swift:4-44``) holds, so a Swift caller can
hand its compiled arrays over without repacking.
"""
import numpy as np

F4 = ("<f4", (4,))
U4 = ("<u4", (4,))

CAMERA = np.dtype([
    ("origin", *F4), ("lowerLeft", *F4), ("horizontal", *F4),
    ("vertical", *F4), ("lens", *F4),
], align=True)

TRIANGLE = np.dtype([
    ("v0", *F4), ("v1", *F4), ("v2", *F4),
    ("n0", *F4), ("n1", *F4), ("n2", *F4),
    ("uv0", *F4), ("uv1", *F4), ("uv2", *F4),
    ("tangent", *F4), ("bitangent", *F4),
    ("materialID", "<u4"), ("objectID", "<u4"),
    ("primitiveID", "<u4"), ("padding2", "<u4"),
], align=True)

MATERIAL_FIELDS = (
    "baseColor", "emission", "parameters", "parameters2", "specularColor",
    "sheenColor", "transmissionColor", "parameters3", "clearcoatColor",
    "clearcoatAttenuation", "transmissionAbsorption", "thinFilm",
    "subsurfaceColor", "subsurfaceRadius", "subsurfaceParameters",
    "volumeScattering", "volumeParameters",
)
MATERIAL = np.dtype([(name, *F4) for name in MATERIAL_FIELDS], align=True)

LIGHT = np.dtype([
    ("triangleIndex", "<u4"), ("materialIndex", "<u4"),
    ("area", "<f4"), ("selectionCDF", "<f4"), ("normal", *F4),
], align=True)

ENV_SAMPLE = np.dtype([("cdf", "<f4"), ("pdf", "<f4")])

TEXTURE_DESC = np.dtype([
    ("offset", "<u4"), ("width", "<u4"), ("height", "<u4"), ("mode", "<u4"),
])

NODE = np.dtype([
    ("boundsMin", *F4), ("boundsMax", *F4), ("metadata", *U4),
])

assert CAMERA.itemsize == 80
assert TRIANGLE.itemsize == 192
assert MATERIAL.itemsize == 272
assert LIGHT.itemsize == 32
assert ENV_SAMPLE.itemsize == 8
assert TEXTURE_DESC.itemsize == 16
assert NODE.itemsize == 48

# Sparse distance volumes (PathTrace.metal:74-148).
VOLUME = np.dtype([
    ("worldBoundsMin", *F4), ("worldBoundsMax", *F4), ("localBoundsMin", *F4), ("localBoundsMax", *F4),
    ("dimensions", *U4), ("metadata", *U4), ("materialProgram", *U4),
    ("worldToLocal", "<f4", (4, 4)),      # 4 columns
    ("normalTransform", "<f4", (4, 4)),   # 4 columns
])
BRICK = np.dtype([
    ("worldBoundsMin", *F4), ("worldBoundsMax", *F4), ("localBoundsMin", *F4), ("localBoundsMax", *F4),
    ("sampleBoundsMin", *F4), ("sampleBoundsMax", *F4),
    ("gridOriginAndVolume", *U4), ("dimensionsAndSampleOffset", *U4),
])
BRICK_SAMPLE = np.dtype([("distance", "<f4"), ("materialA", "<u4"), ("materialB", "<u4"), ("materialBlend", "<f4")])
VOLUME_SAMPLE = np.dtype([  # GPUVolumeSample (PathTrace.metal:92-101), dense volumes
    ("distance", "<f4"), ("materialA", "<u4"), ("materialB", "<u4"), ("materialBlend", "<f4"),
    ("baseColorOpacity", *F4), ("emissionTransmission", *F4), ("surface", *F4), ("materialFieldFlags", *U4),
])
BRICK_GRID = np.dtype([
    ("dimensionsAndIndexOffset", *U4), ("brickSizeAndVolume", *U4),
    ("macroDimensionsAndIndexOffset", *U4), ("macroSizeAndReserved", *U4),
])

assert VOLUME.itemsize == 240
assert BRICK.itemsize == 128
assert BRICK_SAMPLE.itemsize == 16
assert VOLUME_SAMPLE.itemsize == 80
assert BRICK_GRID.itemsize == 64
