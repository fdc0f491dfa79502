"""Image assets from disk as linear RGBA textures (Scene/Texture.swift).

* ``load_texture(path, color_encoding, sampling_mode)`` -- Texture2D(contentsOf:)
  (Texture.swift:70-143): ``.hdr`` files through the Radiance reader below;
  any other image decoded to 8-bit RGBA (PIL here, CoreGraphics there),
  un-premultiplied, clamped to [0, 1] and, for ``"srgb"``, converted with
  the sRGB transfer curve (Texture.swift:207-212). Rows are stored top row
  first, as decoded.
* ``load_radiance_hdr(data)`` -- loadRadianceHDR (Texture.swift:214-343):
  the resolution line is the first containing ``-Y`` or ``+Y`` (a leading Y
  token means height first); each scanline is new-style RLE (``2 2 hi lo``
  with four run-length coded channels) or flat RGBE; texel = RGB * 2^(E-136),
  E = 0 gives black; alpha 1.
* ``derived_normal_map(texture, strength)`` -- Texture2D.derivedNormalMap
  (Texture.swift:168-205): central differences of Rec. 709 luminance, x
  wrapped and y clamped.

What is not pinned: CoreGraphics draws the decoded image into an sRGB,
premultiplied 8-bit context, so embedded colour profiles are converted there
and translucent texels are premultiplied with its own rounding. Here the
8-bit values are taken as they are and premultiplied with round-to-nearest;
opaque sRGB or linear images decode to the same texels.
"""
from __future__ import annotations

import os
from typing import Tuple

import numpy as np

from .scene import Texture2D, TextureSamplingMode

f32 = np.float32


class TextureLoadingError(Exception):
    """TextureLoadingError (Texture.swift:25-40)."""

    def __init__(self, kind: str, path: str):
        self.kind, self.path = kind, path
        if kind == "invalidHDR":
            msg = f"Could not decode Radiance HDR texture at {path}."
        else:
            msg = f"Could not decode texture image at {path}."
        super().__init__(msg)


def srgb_to_linear(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, f32)
    lo = v / f32(12.92)
    hi = np.power((v + f32(0.055)) / f32(1.055), f32(2.4)).astype(f32)
    return np.where(v <= f32(0.04045), lo, hi).astype(f32)


def load_radiance_hdr(data: bytes, path: str = "<data>") -> Tuple[int, int, np.ndarray]:
    """Returns (width, height, (H*W, 4) float32 pixels)."""
    n = len(data)
    i = 0
    res = None
    while i < n:
        j = data.find(b"\n", i)
        end = n if j < 0 else j
        line = data[i:end]
        i = n if j < 0 else j + 1
        if line.endswith(b"\r"):
            line = line[:-1]
        t = line.decode("ascii", "replace").strip()
        if not t:
            continue
        if "-Y" in t or "+Y" in t:
            res = t
            break
    if res is None:
        raise TextureLoadingError("invalidHDR", path)
    tok = res.split()
    try:
        a, b = int(tok[1]), int(tok[3])
    except (IndexError, ValueError):
        raise TextureLoadingError("invalidHDR", path)
    if len(tok) != 4 or a <= 0 or b <= 0:
        raise TextureLoadingError("invalidHDR", path)
    if tok[0].endswith("Y"):
        height, width = a, b
    else:
        width, height = a, b
    buf = np.frombuffer(data, np.uint8)
    rgbe = np.zeros((height, width, 4), np.uint8)
    for y in range(height):
        if i + 4 > n:
            raise TextureLoadingError("invalidHDR", path)
        if 8 <= width <= 0x7FFF and data[i] == 2 and data[i + 1] == 2 and not (data[i + 2] & 0x80):
            if (data[i + 2] << 8 | data[i + 3]) != width:
                raise TextureLoadingError("invalidHDR", path)
            i += 4
            for c in range(4):
                x = 0
                row = rgbe[y, :, c]
                while x < width:
                    if i >= n:
                        raise TextureLoadingError("invalidHDR", path)
                    cnt = data[i]
                    i += 1
                    if cnt > 128:
                        run = cnt - 128
                        if x + run > width or i >= n:
                            raise TextureLoadingError("invalidHDR", path)
                        row[x:x + run] = data[i]
                        i += 1
                        x += run
                    else:
                        if cnt == 0 or x + cnt > width or i + cnt > n:
                            raise TextureLoadingError("invalidHDR", path)
                        row[x:x + cnt] = buf[i:i + cnt]
                        i += cnt
                        x += cnt
        else:
            if i + width * 4 > n:
                raise TextureLoadingError("invalidHDR", path)
            rgbe[y] = buf[i:i + width * 4].reshape(width, 4)
            i += width * 4
    e = rgbe[..., 3].astype(np.int32)
    scale = np.ldexp(f32(1), e - 136).astype(f32)
    px = np.ones((height, width, 4), f32)
    px[..., :3] = rgbe[..., :3].astype(f32) * scale[..., None]
    px[e == 0, :3] = 0
    return width, height, px.reshape(-1, 4)


def load_texture(path: str, color_encoding: str = "srgb",
                 sampling_mode: int = TextureSamplingMode.LINEAR) -> Texture2D:
    """Texture2D(contentsOf:colorEncoding:samplingMode:) (Texture.swift:70-143)."""
    if os.path.splitext(path)[1].lower() == ".hdr":
        try:
            with open(path, "rb") as f:
                data = f.read()
        except OSError:
            raise TextureLoadingError("unsupportedImage", path)
        w, h, px = load_radiance_hdr(data, path)
        return Texture2D(w, h, px, sampling_mode)
    try:
        from PIL import Image
        with Image.open(path) as im:
            rgba = np.asarray(im.convert("RGBA"), np.uint8)
    except Exception:
        raise TextureLoadingError("unsupportedImage", path)
    h, w = rgba.shape[:2]
    rgba = rgba.reshape(-1, 4)
    a8 = rgba[:, 3].astype(np.int32)
    premul = (rgba[:, :3].astype(np.int32) * a8[:, None] + 127) // 255
    alpha = a8.astype(f32) / f32(255)
    with np.errstate(divide="ignore"):
        unpre = np.where(alpha > 0, f32(1) / alpha, f32(0)).astype(f32)
    rgb = np.clip(premul.astype(f32) / f32(255) * unpre[:, None], f32(0), f32(1)).astype(f32)
    if color_encoding == "srgb":
        rgb = srgb_to_linear(rgb)
    px = np.concatenate([rgb, alpha[:, None]], 1).astype(f32)
    return Texture2D(w, h, px, sampling_mode)


def derived_normal_map(t: Texture2D, strength: float = 1.0) -> Texture2D:
    """Texture2D.derivedNormalMap(strength:) (Texture.swift:168-205)."""
    w, h = t.width, t.height
    if w <= 0 or h <= 0 or len(t.pixels) != w * h:
        return Texture2D(w, h, t.pixels.copy(), t.sampling_mode)
    s = f32(max(strength, 0.0))
    p = t.pixels.reshape(h, w, 4)
    lum = (p[..., 0] * f32(0.2126) + p[..., 1] * f32(0.7152)) + p[..., 2] * f32(0.0722)
    left = np.roll(lum, 1, axis=1)
    right = np.roll(lum, -1, axis=1)
    down = np.concatenate([lum[:1], lum[:-1]], 0)
    up = np.concatenate([lum[1:], lum[-1:]], 0)
    v = np.stack([(left - right) * s, (down - up) * s, np.ones_like(lum)], -1).astype(f32)
    from .scene import _normalize
    n = _normalize(v)
    out = np.ones((h, w, 4), f32)
    out[..., :3] = n * f32(0.5) + f32(0.5)
    return Texture2D(w, h, out.reshape(-1, 4), t.sampling_mode)
