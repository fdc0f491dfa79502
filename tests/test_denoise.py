"""CPU pins of the denoiser oracle (oracle/denoise_oracle.py).

* The U-Net oracle's backward pass equals central finite differences of its
  own forward pass (exact float64, no BF16 rounding) for sampled weights and
  biases of every layer: this fixes the tap order, the transposed weights of
  the data gradient, the nearest-upsampling/sum-pool pair and the pooling
  gradient independently of the GPU code.
* The a-trous oracle with edge-stopping weights switched off (huge sigmas)
  is 0.15 x input + 0.85 x a separable B3-spline blur with clamp-to-edge
  sampling, computed here with scipy.ndimage.correlate1d (mode "nearest"); a
  constant image is a fixed point; pixels with no coverage pass through.
"""
import numpy as np
import pytest
from scipy import ndimage

import denoise_oracle as D


def test_unet_layer_table():
    ls = D.layers()
    assert len(ls) == 19
    assert sum(9 * co * (ca + cb) for _, ca, cb, co, _, _ in ls) == 633600
    assert D.PARAM_COUNT == 633600 + sum(co for _, _, _, co, _, _ in ls) == 634576
    for _, ca, cb, co, _, _ in ls:
        assert ca % 16 == 0 and cb % 16 == 0 and co % 16 == 0 and (co // 16) * ((ca + cb) // 16) <= 64


def test_package_layer_table_matches_oracle():
    from denrimrendererkit_amd.denoise import LAYERS, inference_flops
    ls = D.layers()
    assert [(ca, cb, co) for _, ca, cb, co, _, _ in ls] == [(ca, cb, co) for ca, cb, co, _ in LAYERS]
    assert [D.TENSORS[D.OPS[i][3]][1] for i, *_ in ls] == [lv for *_, lv in LAYERS]
    assert inference_flops(32, 32) == sum(2 * (32 >> lv) ** 2 * 9 * (ca + cb) * co for ca, cb, co, lv in LAYERS)


def test_unet_oracle_backward_equals_finite_differences():
    rng = np.random.default_rng(1)
    p = D.init_params(3)
    for _, ca, cb, co, w, b in D.layers():
        p[b:b + co] = rng.uniform(-0.05, 0.05, co)
    x = rng.uniform(0, 1, (1, 16, 16, 16))
    t = rng.uniform(0, 2, (1, 16, 16, 4))
    _, g = D.UNet(p, bf16_rounding=False).loss_and_grads(x, t)
    idx = []
    for _, ca, cb, co, w, b in D.layers():
        idx += list(rng.integers(w, b, 2)) + [b + int(rng.integers(0, co))]
    worst = 0.0
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        num = (D.UNet(pp, False).loss_and_grads(x, t)[0] - D.UNet(pm, False).loss_and_grads(x, t)[0]) / (2 * e)
        worst = max(worst, abs(num - g[i]) / max(abs(g[i]), 1e-7))
    assert worst < 1e-3, worst


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -2.5], np.float32)
    assert list(D.bf16(v)) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5]


def _planes(H, W, rng):
    beauty = np.ones((H, W, 4), np.float32)
    beauty[..., :3] = rng.uniform(0, 2, (H, W, 3))
    depth = np.ones((H, W, 4), np.float32) * 2.0
    normal = np.zeros((H, W, 4), np.float32)
    normal[..., 2] = 1.0
    normal[..., :3] = normal[..., :3] * 0.5 + 0.5
    albedo = np.full((H, W, 4), 0.5, np.float32)
    return beauty, depth, normal, albedo


@pytest.mark.parametrize("step", [1, 2, 4])
def test_atrous_without_edge_stops_is_b3_spline_blur(step):
    rng = np.random.default_rng(2)
    beauty, depth, normal, albedo = _planes(23, 31, rng)
    out = D.atrous_pass(beauty, depth, normal, albedo, 2, step, 1e9, 1e9, 1e9, 1e9)
    k = np.zeros(4 * step + 1)
    k[::step] = [1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16]
    blur = ndimage.correlate1d(beauty[..., :3].astype(np.float64), k, axis=0, mode="nearest")
    blur = ndimage.correlate1d(blur, k, axis=1, mode="nearest")
    expect = 0.15 * beauty[..., :3] + 0.85 * blur
    assert np.abs(out[..., :3] - expect).max() < 1e-5
    assert np.array_equal(out[..., 3], beauty[..., 3])


def test_atrous_constant_fixed_point_and_uncovered_passthrough():
    rng = np.random.default_rng(3)
    beauty, depth, normal, albedo = _planes(12, 10, rng)
    beauty[..., :3] = 0.7
    out = D.atrous(beauty, depth, normal, albedo)
    assert np.abs(out[..., :3] - 0.7).max() < 1e-6
    beauty[..., :3] = rng.uniform(0, 2, (12, 10, 3))
    depth[3, 4] = 0.0  # no coverage: written unchanged (Denoise.metal:56-60)
    beauty[5, 5, 3] = 0.0
    out = D.atrous(beauty, depth, normal, albedo, iterations=2)
    assert np.array_equal(out[3, 4], beauty[3, 4]) and np.array_equal(out[5, 5], beauty[5, 5])
