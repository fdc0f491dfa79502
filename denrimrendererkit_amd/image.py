"""Output-side encoding of the reference's PNG writer (Image/PNGWriter.swift).

``visualized_rgba8`` restates PNGWriter.visualizedRGBA8Pixels
(PNGWriter.swift:38-56): per RenderOutput, beauty through ``tonemap``
(exposure 0.8, the ACES-fit curve (x(2.51x + 0.03)) / (x(2.43x + 0.59) +
0.14), gamma 1/2.2, alpha kept), depth normalized to [0.1, 1] over the
covered pixels, normal and albedo display-linear (gamma 1/2.2), material
and object IDs through the ID palette hash (``id_color``, :178-189), motion
vectors as signed values scaled by the largest component. 8-bit values
truncate as Swift's UInt8(Float) does. ``write_png`` writes RGBA8 as a PNG
file (zlib, no filter), replacing the reference's ImageIO writer
(:58-72); failures raise pngExportFailed.
"""
import struct
import zlib

import numpy as np


def tonemap(linear: np.ndarray) -> np.ndarray:
    x = np.maximum(0.0, np.asarray(linear, np.float32)) * np.float32(0.8)
    mapped = (x * (np.float32(2.51) * x + np.float32(0.03))) / (x * (np.float32(2.43) * x + np.float32(0.59)) + np.float32(0.14))
    g = np.power(np.clip(mapped, 0.0, 1.0), np.float32(1.0 / 2.2))
    return np.clip(g * np.float32(255.0), 0, 255).astype(np.uint8)


def beauty_to_rgba8(rgba: np.ndarray) -> np.ndarray:
    """(H, W, 4) linear beauty -> (H, W, 4) uint8 with alpha = alpha*255."""
    out = np.empty(rgba.shape[:2] + (4,), np.uint8)
    out[..., :3] = tonemap(rgba[..., :3])
    out[..., 3] = np.clip(rgba[..., 3] * 255.0, 0, 255).astype(np.uint8)
    return out


def id_color(ids: np.ndarray) -> np.ndarray:
    h = np.asarray(ids, np.uint64) & 0xFFFFFFFF
    h = (h * 747796405) & 0xFFFFFFFF
    h = (h + 2891336453) & 0xFFFFFFFF
    h = (((h >> ((h >> 28) + 4)) ^ h) * 277803737) & 0xFFFFFFFF
    h = (h >> 22) ^ h
    return np.stack([64 + (h & 0x7F), 64 + ((h >> 8) & 0x7F), 64 + ((h >> 16) & 0x7F)], -1).astype(np.uint8)


def _u8(x):
    """UInt8(max(0, min(255, x))) with Swift's truncation toward zero."""
    return np.clip(np.asarray(x, np.float32), 0.0, 255.0).astype(np.uint8)


def display_linear(linear: np.ndarray) -> np.ndarray:
    """encodeDisplayLinear (PNGWriter.swift:168-171)."""
    g = np.power(np.clip(np.asarray(linear, np.float32), 0.0, 1.0), np.float32(1.0 / 2.2))
    return _u8(g * np.float32(255.0))


def visualized_rgba8(pixels: np.ndarray, output) -> np.ndarray:
    """(H, W, 4) float pixels of a RenderOutput -> (H, W, 4) uint8 as
    PNGWriter.visualizedRGBA8Pixels encodes them."""
    from .settings import RenderOutput
    px = np.asarray(pixels, np.float32)
    H, W = px.shape[:2]
    out = np.zeros((H, W, 4), np.uint8)
    out[..., 3] = 255
    output = RenderOutput(output)
    if output == RenderOutput.BEAUTY:
        return beauty_to_rgba8(px)
    if output in (RenderOutput.NORMAL, RenderOutput.ALBEDO):
        out[..., :3] = display_linear(px[..., :3])
        if output == RenderOutput.ALBEDO:  # preservesAlpha
            out[..., 3] = _u8(px[..., 3] * np.float32(255.0))
        return out
    if output == RenderOutput.DEPTH:  # :78-107
        d = px[..., 0]
        ok = np.isfinite(d) & (d > 0)
        if not ok.any():
            out[..., 3] = 0
            return out
        lo, hi = d[ok].min(), d[ok].max()
        rng = np.float32(hi - lo)
        norm = np.ones_like(d) if rng <= np.finfo(np.float32).eps else (d - lo) / rng
        v = _u8((np.float32(0.1) + norm * np.float32(0.9)) * np.float32(255.0))
        for c in range(3):
            out[..., c] = np.where(ok, v, 0)
        return out
    if output in (RenderOutput.MATERIAL_ID, RenderOutput.OBJECT_ID):  # :109-129
        ids = np.rint(px[..., 0]).astype(np.int64)  # Float.rounded(): half away from zero
        ids = np.where(np.abs(px[..., 0] - np.trunc(px[..., 0])) == 0.5, np.trunc(px[..., 0]) + np.sign(px[..., 0]),
                       ids).astype(np.int64)
        pos = ids > 0
        col = id_color(np.where(pos, ids, 0))
        out[..., :3] = np.where(pos[..., None], col, 0)
        return out
    if output == RenderOutput.MOTION_VECTOR:  # :131-150
        scale = max(float(np.max(np.maximum(np.abs(px[..., 0]), np.abs(px[..., 1])))) if px.size else 0.0, 1.0)
        s = np.float32(scale)
        out[..., 0] = _u8((px[..., 0] / s * np.float32(0.5) + np.float32(0.5)) * np.float32(255.0))
        out[..., 1] = _u8((px[..., 1] / s * np.float32(0.5) + np.float32(0.5)) * np.float32(255.0))
        out[..., 2] = 128
        return out
    raise ValueError(f"no PNG encoding for {output}")


def png_bytes(rgba8: np.ndarray) -> bytes:
    """RGBA8 (H, W, 4) -> PNG file bytes (8-bit RGBA, filter 0, zlib)."""
    a = np.ascontiguousarray(rgba8, np.uint8)
    H, W = a.shape[:2]
    raw = b"".join(b"\x00" + a[y].tobytes() for y in range(H))

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def write_png(path: str, rgba8: np.ndarray) -> None:
    """PNGWriter.write (PNGWriter.swift:58-72): raises pngExportFailed(path)."""
    from .scene import DenrimRendererError
    try:
        with open(path, "wb") as f:
            f.write(png_bytes(rgba8))
    except OSError as e:
        raise DenrimRendererError("pngExportFailed", str(path)) from e
