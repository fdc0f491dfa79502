"""Output-side encoding of the reference's PNG writer (Image/PNGWriter.swift).

``tonemap`` restates the beauty encoding (PNGWriter.swift:160-166): exposure
0.8, the ACES-fit curve (x(2.51x + 0.03)) / (x(2.43x + 0.59) + 0.14), gamma
1/2.2, truncation to 8 bits. ``id_color`` restates the ID palette hash
(PNGWriter.swift:178-189).
"""
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
