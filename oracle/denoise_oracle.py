"""CPU oracle of the denoisers (test infrastructure only: imported by tests/,
never by the product path).

* ``atrous`` restates simpleSpatialDenoiseKernel
  (Sources/DenrimRendererKit/Resources/Shaders/Denoise.metal:39-115) and the
  pass loop of SimpleSpatialDenoiser.encode (Denoising/Denoiser.swift:112-164)
  in float32 numpy, vectorised over pixels, with the kernel's operation order.
* ``UNet`` restates this backend's learned denoiser (no reference
  counterpart): the network of denrimrendererkit_amd/csrc/dr_unet.cpp, forward
  and backward, in float64 with the GPU path's BF16 roundings applied at the
  same points (features, weights, activations, pooled tensors, output
  gradients), so the two differ only by accumulation order. Its backward pass
  is pinned by finite differences of its own forward pass in exact float64
  (``UNet(bf16=False)``), which fixes the convolution/transposition/upsampling
  conventions independently of the GPU code.
"""
from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------- a-trous

def _tap(o):
    o = abs(o)
    return np.float32(0.375 if o == 0 else (0.25 if o == 1 else 0.0625))


def _decode_normal(e):
    n = e[..., :3] * np.float32(2.0) - np.float32(1.0)
    ln = np.sqrt(np.sum(n * n, axis=-1, keepdims=True, dtype=np.float32))
    with np.errstate(invalid="ignore", divide="ignore"):
        return n / ln


def atrous_pass(src, depth, normal, albedo, radius, step, normal_sigma, depth_sigma, albedo_sigma, color_sigma):
    """One pass of simpleSpatialDenoiseKernel over (H, W, 4) float32 planes."""
    f32 = np.float32
    H, W = src.shape[:2]
    src, depth, normal, albedo = (np.asarray(a, f32) for a in (src, depth, normal, albedo))
    cd = depth[..., 0]
    valid_c = ~((src[..., 3] <= f32(0.0001)) | ~np.isfinite(cd) | (cd <= 0))
    cn = _decode_normal(normal)
    ns2 = f32(normal_sigma) * f32(normal_sigma)
    ds = np.maximum(f32(depth_sigma) * np.maximum(cd, f32(1.0)), f32(0.0001))
    ds2 = ds * ds
    as2 = f32(albedo_sigma) * f32(albedo_sigma)
    cs2 = f32(color_sigma) * f32(color_sigma)
    acc = np.zeros((H, W, 3), f32)
    tot = np.zeros((H, W), f32)
    ys, xs = np.arange(H), np.arange(W)
    for oy in range(-radius, radius + 1):
        for ox in range(-radius, radius + 1):
            sy = np.clip(ys + oy * step, 0, H - 1)
            sx = np.clip(xs + ox * step, 0, W - 1)
            s = src[sy][:, sx]
            sd = depth[sy][:, sx][..., 0]
            ok = ~((s[..., 3] <= f32(0.0001)) | ~np.isfinite(sd) | (sd <= 0))
            spatial = _tap(ox) * _tap(oy)
            sn = cn[sy][:, sx]
            dot = (cn[..., 0] * sn[..., 0] + cn[..., 1] * sn[..., 1]) + cn[..., 2] * sn[..., 2]
            nd = f32(1.0) - np.minimum(np.maximum(dot, f32(0.0)), f32(1.0))
            with np.errstate(invalid="ignore", over="ignore"):
                wn = np.exp(-(nd * nd) / (f32(2.0) * ns2))
                dd = sd - cd
                wd = np.exp(-(dd * dd) / (f32(2.0) * ds2))
                da = albedo[..., :3] - albedo[sy][:, sx][..., :3]
                wa = np.exp(-((da[..., 0] * da[..., 0] + da[..., 1] * da[..., 1]) + da[..., 2] * da[..., 2])
                            / (f32(2.0) * as2))
                dc = src[..., :3] - s[..., :3]
                wc = np.exp(-((dc[..., 0] * dc[..., 0] + dc[..., 1] * dc[..., 1]) + dc[..., 2] * dc[..., 2])
                            / (f32(2.0) * cs2))
                w = (((spatial * wn) * wd) * wa) * wc
            w = np.where(ok, w, f32(0.0)).astype(f32)
            acc += s[..., :3] * w[..., None]
            tot += w
    with np.errstate(invalid="ignore", divide="ignore"):
        den = np.where(tot[..., None] > 0, acc / tot[..., None], src[..., :3])
    out = src.copy()
    mixed = src[..., :3] + (den - src[..., :3]) * f32(0.85)
    out[..., :3] = np.where(valid_c[..., None], mixed, src[..., :3])
    return out


def atrous(beauty, depth, normal, albedo, radius=2, normal_sigma=0.18, depth_sigma=0.03, albedo_sigma=0.25,
           color_sigma=1.0, iterations=3):
    """SimpleSpatialDenoiser.encode: 1-5 passes, step 1 << pass, parameters clamped."""
    n = max(1, min(int(iterations), 5))
    r = max(1, min(int(radius), 2))
    cur = np.asarray(beauty, np.float32)
    for i in range(n):
        cur = atrous_pass(cur, depth, normal, albedo, r, 1 << i, max(normal_sigma, 1e-4), max(depth_sigma, 1e-4),
                          max(albedo_sigma, 1e-4), max(color_sigma, 1e-4))
    return cur


# ------------------------------------------------------------------ U-Net

# (name, input a, input b (upsampled) or None, output, output channels); pools
# as (None, a, None, out). Mirrors the table of dr_unet.cpp.
TENSORS = {"x0": (16, 0), "t0a": (32, 0), "e0": (32, 0), "p0": (32, 1), "t1a": (32, 1), "e1": (32, 1),
           "p1": (32, 2), "t2a": (48, 2), "e2": (48, 2), "p2": (48, 3), "t3a": (64, 3), "e3": (64, 3),
           "p3": (64, 4), "t4a": (128, 4), "b4": (128, 4), "t5a": (64, 3), "d3": (64, 3), "t6a": (48, 2),
           "d2": (48, 2), "t7a": (32, 1), "d1": (32, 1), "t8a": (32, 0), "d0": (32, 0), "y": (16, 0)}
OPS = [("conv", "x0", None, "t0a"), ("conv", "t0a", None, "e0"), ("pool", "e0", None, "p0"),
       ("conv", "p0", None, "t1a"), ("conv", "t1a", None, "e1"), ("pool", "e1", None, "p1"),
       ("conv", "p1", None, "t2a"), ("conv", "t2a", None, "e2"), ("pool", "e2", None, "p2"),
       ("conv", "p2", None, "t3a"), ("conv", "t3a", None, "e3"), ("pool", "e3", None, "p3"),
       ("conv", "p3", None, "t4a"), ("conv", "t4a", None, "b4"), ("conv", "e3", "b4", "t5a"),
       ("conv", "t5a", None, "d3"), ("conv", "e2", "d3", "t6a"), ("conv", "t6a", None, "d2"),
       ("conv", "e1", "d2", "t7a"), ("conv", "t7a", None, "d1"), ("conv", "e0", "d1", "t8a"),
       ("conv", "t8a", None, "d0"), ("conv", "d0", None, "y")]


def layers():
    """[(op index, ca, cb, cout, w offset, b offset)] in parameter order."""
    out, off = [], 0
    for i, (kind, a, b, o) in enumerate(OPS):
        if kind != "conv":
            continue
        ca, cb, co = TENSORS[a][0], (TENSORS[b][0] if b else 0), TENSORS[o][0]
        w = off
        off += 9 * co * (ca + cb)
        out.append((i, ca, cb, co, w, off))
        off += co
    return out


PARAM_COUNT = layers()[-1][5] + TENSORS["y"][0]


def bf16(x):
    """Round-to-nearest-even to BF16, returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _conv(x, w, bias):
    """x (n, H, W, cin), w (9, cout, cin) -> (n, H, W, cout), zero padding."""
    n, H, W, _ = x.shape
    xp = np.zeros((n, H + 2, W + 2, x.shape[3]))
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((n, H, W, w.shape[1])) + (bias if bias is not None else 0.0)
    for tap in range(9):
        dy, dx = tap // 3, tap % 3
        out += xp[:, dy:dy + H, dx:dx + W] @ w[tap].T
    return out


def _up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def _pool(x):
    return ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2])) * 0.25


class UNet:
    """Forward/backward of the network for features (n, H, W, 16) and
    targets (n, H, W, 4); ``bf16`` applies the GPU path's roundings."""

    def __init__(self, params, bf16_rounding=True):
        self.p = np.asarray(params, np.float64)
        assert self.p.size == PARAM_COUNT
        self.r = bf16 if bf16_rounding else (lambda v: np.asarray(v, np.float64))

    def _w(self, L):
        _, ca, cb, co, w, b = L
        return self.r(self.p[w:w + 9 * co * (ca + cb)].reshape(9, co, ca + cb)), self.p[b:b + co]

    def forward(self, x):
        r = self.r
        act = {"x0": r(x)}
        li = 0
        ls = layers()
        for kind, a, b, o in OPS:
            if kind == "pool":
                act[o] = r(_pool(act[a]))
                continue
            w, bias = self._w(ls[li])
            li += 1
            inp = act[a] if b is None else np.concatenate([act[a], _up(act[b])], axis=-1)
            v = _conv(inp, w, bias)
            act[o] = v if o == "y" else r(np.maximum(v, 0.0))
        self.act = act
        return act["y"]

    def loss_and_grads(self, x, target):
        """Mean L1 over channels 0-2 and its gradient w.r.t. every parameter."""
        r = self.r
        y = self.forward(x)
        act = self.act
        n, H, W = y.shape[:3]
        scale = 1.0 / (n * H * W * 3)
        d = y[..., :3] - target[..., :3]
        loss = float(np.abs(d).sum() * scale)
        dy = np.zeros_like(y)
        dy[..., :3] = r(np.sign(d) * scale)
        grads = np.zeros_like(self.p)
        dact = {k: np.zeros_like(v) for k, v in act.items() if k not in ("x0", "y")}
        ls = layers()
        li = len(ls)
        for kind, a, b, o in reversed(OPS):
            if kind == "pool":
                dact[a] += _up(dact[o]) * 0.25
                continue
            li -= 1
            L = ls[li]
            _, ca, cb, co, wo, bo = L
            w, _ = self._w(L)
            dpre = dy if o == "y" else r(np.where(act[o] > 0, dact[o], 0.0))
            inp = act[a] if b is None else np.concatenate([act[a], _up(act[b])], axis=-1)
            ip = np.zeros((n, inp.shape[1] + 2, inp.shape[2] + 2, ca + cb))
            ip[:, 1:-1, 1:-1] = inp
            h, wd = inp.shape[1:3]
            gw = np.zeros((9, co, ca + cb))
            for tap in range(9):
                ty, tx = tap // 3, tap % 3
                sl = ip[:, ty:ty + h, tx:tx + wd].reshape(-1, ca + cb)
                gw[tap] = dpre.reshape(-1, co).T @ sl
            grads[wo:wo + gw.size] = gw.ravel()
            grads[bo:bo + co] = dpre.reshape(-1, co).sum(axis=0)
            if a == "x0":
                continue
            wt = w[::-1].transpose(0, 2, 1)  # [8 - tap][cin][cout]
            din = _conv(dpre, wt, None)
            dact[a] += din[..., :ca]
            if b is not None:
                db = din[..., ca:]
                dact[b] += (db[:, 0::2, 0::2] + db[:, 0::2, 1::2]) + (db[:, 1::2, 0::2] + db[:, 1::2, 1::2])
        return loss, grads


def init_params(seed=0):
    """He-uniform weights, zero biases (the same scheme as dr_hip_unet_create,
    not the same random stream)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAM_COUNT)
    for _, ca, cb, co, w, b in layers():
        bound = np.sqrt(6.0 / (9 * (ca + cb)))
        p[w:b] = rng.uniform(-bound, bound, b - w)
    return p
