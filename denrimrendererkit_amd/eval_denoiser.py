"""relMSE of plain accumulation, the learned denoiser and the a-trous filter
against a high-spp render, on held-out scenes the network was not trained on.

    python -m denrimrendererkit_amd.eval_denoiser preview      # material-preview scenes, 96^2 ... 1280x720
    python -m denrimrendererkit_amd.eval_denoiser room --size 1920x1080 --scenes 2

relMSE = mean over pixels and RGB of (x - ref)^2 / (ref^2 + 0.01); the
reference is a ``--ref-spp`` render with another frame seed. Prints one
JSON line per scene and the mean/median over scenes.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import DenrimRenderer, RenderQuality, RenderSettings, compile_scene, synthetic
from .denoise import DEFAULT_WEIGHTS, UNetDenoiser, relmse
from .settings import DenoiseSettings

PREVIEW_CASES = ((900001, (256, 256)), (900002, (512, 512)), (900003, (256, 256)), (900004, (1280, 720)),
                 (900005, (96, 96)), (900006, (512, 512)))
ROOM_SEEDS = (2001, 2002, 2003, 2004)


def evaluate(renderer, net, scene, W, H, ref_spp, plain_spp, learned_spp, atrous_spp=(), bounces=4):
    gs = renderer.upload(compile_scene(scene))

    def st(seed, d=None):
        return RenderSettings(width=W, height=H, max_bounces=bounces, quality=RenderQuality.FINAL, frame_seed=seed,
                              denoise=d or DenoiseSettings())

    ref = renderer.make_session(gs, st(99), camera=scene.camera)
    ref.render(ref_spp)
    target = ref.pixels()
    ref.close()
    res = {}
    for spp in plain_spp:
        s = renderer.make_session(gs, st(1), camera=scene.camera)
        s.render(spp)
        res[f"plain_{spp}"] = relmse(s.pixels(), target)
        s.close()
    for spp in learned_spp:
        s = renderer.make_session(gs, st(1, DenoiseSettings.learned()), camera=scene.camera)
        s.set_denoiser_network(net)
        s.render(spp)
        res[f"learned_{spp}"] = relmse(s.pixels(), target)
        s.close()
    for spp in atrous_spp:
        s = renderer.make_session(gs, st(1, DenoiseSettings.simple_spatial()), camera=scene.camera)
        s.render(spp)
        res[f"atrous_{spp}"] = relmse(s.pixels(), target)
        s.close()
    gs.release()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("family", choices=("preview", "room"))
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--size", default="320x180", help="room frame size")
    ap.add_argument("--scenes", type=int, default=0, help="number of held-out scenes (0: all)")
    args = ap.parse_args(argv)
    r = DenrimRenderer(0)
    net = UNetDenoiser(r, params=np.load(args.weights, allow_pickle=False))
    rows = []
    t0 = time.perf_counter()
    if args.family == "preview":
        for seed, (W, H) in PREVIEW_CASES[:args.scenes or None]:
            res = evaluate(r, net, synthetic.material_preview_scene(seed), W, H, args.ref_spp, (4, 16, 64, 256),
                           (1, 2, 4, 8, 16, 32))
            rows.append(res)
            print(json.dumps({"seed": seed, "size": f"{W}x{H}", **{k: round(v, 5) for k, v in res.items()}}),
                  flush=True)
    else:
        W, H = (int(v) for v in args.size.split("x"))
        for seed in ROOM_SEEDS[:args.scenes or None]:
            res = evaluate(r, net, synthetic.room_variant(seed), W, H, args.ref_spp, (1, 2, 4, 16, 64, 256), (1, 2, 4),
                           (4,))
            rows.append(res)
            print(json.dumps({"seed": seed, "size": f"{W}x{H}", **{k: round(v, 5) for k, v in res.items()}}),
                  flush=True)
    keys = rows[0].keys()
    print(json.dumps({"mean": {k: round(float(np.mean([r_[k] for r_ in rows])), 5) for k in keys},
                      "median": {k: round(float(np.median([r_[k] for r_ in rows])), 5) for k in keys},
                      "seconds": round(time.perf_counter() - t0, 1)}), flush=True)


if __name__ == "__main__":
    main()
