"""The learned denoiser (dr_hip_unet_*) and render-crop cameras for its
training data.

The network is a 5-level U-Net (19 3x3 convolutions, channels 32-128,
634,576 parameters) whose convolutions run on the gfx950 matrix cores in
BF16 with FP32 accumulation, for inference and for training (forward,
backward and Adam on the device). It consumes a session's beauty, albedo,
normal and depth planes and predicts demodulated log radiance; a session
whose settings select ``RenderDenoiser.LEARNED`` returns the network's
output as its beauty (``RenderSession.set_denoiser_network``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _lib

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "unet_denoiser.npy")


def _ptr(a: np.ndarray):
    return a.ctypes.data


# (input channels, upsampled input channels, output channels, level) of the
# 19 convolutions, in order (the table of csrc/dr_unet.cpp).
LAYERS = ((16, 0, 32, 0), (32, 0, 32, 0), (32, 0, 32, 1), (32, 0, 32, 1), (32, 0, 48, 2), (48, 0, 48, 2),
          (48, 0, 64, 3), (64, 0, 64, 3), (64, 0, 128, 4), (128, 0, 128, 4), (64, 128, 64, 3), (64, 0, 64, 3),
          (48, 64, 48, 2), (48, 0, 48, 2), (32, 48, 32, 1), (32, 0, 32, 1), (32, 32, 32, 0), (32, 0, 32, 0),
          (32, 0, 16, 0))


def inference_flops(width: int, height: int) -> int:
    """Multiply-add FLOPs of one forward pass: sum over convolutions of
    2 * H_l * W_l * 9 * C_in * C_out on the padded (multiple-of-16) image."""
    W, H = (width + 15) // 16 * 16, (height + 15) // 16 * 16
    return sum(2 * (H >> lv) * (W >> lv) * 9 * (ca + cb) * co for ca, cb, co, lv in LAYERS)


class UNetDenoiser:
    def __init__(self, renderer, seed: int = 0, params: Optional[np.ndarray] = None):
        h = C.c_void_p()
        _lib.check(_lib.load().dr_hip_unet_create(renderer._h, seed, C.byref(h)), "unet")
        self._h = h
        self.renderer = renderer
        if params is not None:
            self.set_params(params)

    @classmethod
    def pretrained(cls, renderer, path: str = DEFAULT_WEIGHTS) -> "UNetDenoiser":
        """The network with the committed trained parameters (FP32 .npy)."""
        return cls(renderer, params=np.load(path, allow_pickle=False))

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().dr_hip_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def param_count(self) -> int:
        n = C.c_uint64()
        _lib.check(_lib.load().dr_hip_unet_param_count(self._h, C.byref(n)), "param count")
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        _lib.check(_lib.load().dr_hip_unet_get_params(self._h, _ptr(p), p.size), "get params")
        return p

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        _lib.check(_lib.load().dr_hip_unet_set_params(self._h, _ptr(p), p.size), "set params")

    def save(self, path: str) -> None:
        np.save(path, self.get_params())

    # -- host-fed batches (gradient checks, custom data) ------------------
    def set_batch(self, features: np.ndarray, targets: np.ndarray) -> None:
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        n, h, w, c = f.shape
        assert c == 16 and t.shape == (n, h, w, 4)
        _lib.check(_lib.load().dr_hip_unet_set_batch(self._h, _ptr(f), _ptr(t), n, h, w), "set batch")
        self._shape = (n, h, w)

    def forward_backward(self) -> float:
        loss = C.c_float()
        _lib.check(_lib.load().dr_hip_unet_forward_backward(self._h, C.byref(loss)), "forward/backward")
        return loss.value

    def grads(self) -> np.ndarray:
        g = np.zeros(self.param_count, np.float32)
        _lib.check(_lib.load().dr_hip_unet_get_grads(self._h, _ptr(g), g.size), "grads")
        return g

    def output(self) -> np.ndarray:
        n, h, w = self._shape
        y = np.zeros((n, h, w, 16), np.float32)
        _lib.check(_lib.load().dr_hip_unet_get_output(self._h, _ptr(y), y.size), "output")
        return y

    def train_step(self, learning_rate: float) -> float:
        loss = C.c_float()
        _lib.check(_lib.load().dr_hip_unet_train_step(self._h, learning_rate, C.byref(loss)), "train step")
        return loss.value

    # -- device-resident training set from sessions -------------------------
    def dataset_reserve(self, capacity: int, height: int, width: int) -> None:
        _lib.check(_lib.load().dr_hip_unet_dataset_reserve(self._h, capacity, height, width), "dataset reserve")

    def dataset_add(self, noisy_session, target_session) -> int:
        k = C.c_uint32()
        _lib.check(_lib.load().dr_hip_unet_dataset_add(self._h, noisy_session._h, target_session._h, C.byref(k)),
                   "dataset add")
        return k.value

    def train(self, steps: int, batch: int, learning_rate: float, final_learning_rate: float,
              seed: int = 0) -> np.ndarray:
        losses = np.zeros(steps, np.float32)
        _lib.check(_lib.load().dr_hip_unet_train_dataset(self._h, steps, batch, learning_rate, final_learning_rate,
                                                         seed, _ptr(losses)), "train")
        return losses


class CropCamera:
    """The camera of the (x0, y0, w, h) window of a full_w x full_h render:
    the same rays as that window (pixel centres and jitter span the window),
    so a small session renders a crop of a large frame."""

    def __init__(self, camera, full_w: int, full_h: int, x0: int, y0: int, w: int, h: int):
        self.camera, self.full = camera, (full_w, full_h)
        self.window = (x0, y0, w, h)

    def gpu_camera(self, width: int, height: int) -> np.ndarray:
        full_w, full_h = self.full
        x0, y0, w, h = self.window
        assert (width, height) == (w, h)
        g = self.camera.gpu_camera(full_w, full_h).copy()
        hor = np.asarray(g["horizontal"], np.float64)
        ver = np.asarray(g["vertical"], np.float64)
        ll = np.asarray(g["lowerLeft"], np.float64)
        # u = (x0 + u' w) / W; v = 1 - (y0 + (1 - v') h) / H  (row 0 on top).
        ll_new = ll + hor * (x0 / full_w) + ver * ((full_h - y0 - h) / full_h)
        for k in range(3):
            g["lowerLeft"][k] = ll_new[k]
            g["horizontal"][k] = hor[k] * (w / full_w)
            g["vertical"][k] = ver[k] * (h / full_h)
        return g


def relmse(x: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """Relative MSE over RGB: mean((x - ref)^2 / (ref^2 + eps))."""
    x = np.asarray(x, np.float64)[..., :3]
    ref = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))
