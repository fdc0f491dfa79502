"""Trains the learned denoiser on pairs the renderer generates itself.

Each example is a size x size window of a 1920x1080 frame of a config-1-style
room (synthetic.room_variant: seeded colours, materials, light, camera pose)
or, with --preview-scenes, of a material-preview frame (sphere on a plane
under an HDRI, synthetic.material_preview_scene, orbit poses):
the noisy input at 1, 2 or 4 spp (beauty + centre-ray albedo/normal/depth, as
a session with a denoiser renders them) and a target of the same window at
``target_spp`` spp with another frame seed. Examples stay on the device;
training (forward, backward, Adam) runs there too.

    python -m denrimrendererkit_amd.train_denoiser --out denrimrendererkit_amd/weights/unet_denoiser.npy
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import DenrimRenderer, RenderQuality, RenderSettings, compile_scene, synthetic
from .denoise import CropCamera, UNetDenoiser
from .scene import Camera
from .settings import DenoiseSettings

FULL_W, FULL_H = 1920, 1080


PREVIEW_FRAMES = ((256, 256), (512, 512), (1280, 720), (1920, 1080))


def _pose(rng):
    u = lambda a, b: float(rng.uniform(a, b))  # noqa: E731
    return Camera(origin=(u(-1.4, 1.4), u(0.6, 2.4), u(1.2, 1.9)), target=(u(-1.5, 1.5), u(0.3, 1.8), u(-1.9, -0.6)),
                  vertical_field_of_view_degrees=u(35, 65))


def _preview_pose(rng):
    """An orbit around the material-preview sphere."""
    u = lambda a, b: float(rng.uniform(a, b))  # noqa: E731
    d, th = u(2.4, 4.2), u(-math.pi, math.pi)
    return Camera(origin=(d * math.sin(th), u(0.4, 2.2), d * math.cos(th)), target=(u(-0.2, 0.2), u(0.5, 1.0), 0.0),
                  vertical_field_of_view_degrees=u(28, 45))


def generate(renderer, net: UNetDenoiser, scenes: int, crops: int, size: int, target_spp: int, seed: int = 0,
             max_bounces: int = 4, log=None, preview_scenes: int = 0) -> int:
    """Adds (scenes + preview_scenes) x crops examples to the network's
    device-resident set: config-1-style rooms (windows of 1920x1080 frames)
    and material-preview scenes (sphere on a plane under an HDRI, windows of
    256^2 to 1920x1080 frames from orbit poses)."""
    rng = np.random.default_rng(seed)
    total = scenes + preview_scenes
    net.dataset_reserve(total * crops, size, size)
    noisy_st = RenderSettings(width=size, height=size, max_bounces=max_bounces, quality=RenderQuality.FINAL,
                              frame_seed=1, denoise=DenoiseSettings.learned())
    target_st = RenderSettings(width=size, height=size, max_bounces=max_bounces, quality=RenderQuality.FINAL,
                               frame_seed=2)
    t0 = time.perf_counter()
    for k in range(total):
        preview = k >= scenes
        if preview:
            scene = synthetic.material_preview_scene(50000 + seed * 100003 + k)
        else:
            scene = synthetic.room_variant(seed * 100003 + k)
        gs = renderer.upload(compile_scene(scene))
        noisy = target = None
        for c in range(crops):
            fw, fh = PREVIEW_FRAMES[int(rng.integers(len(PREVIEW_FRAMES)))] if preview else (FULL_W, FULL_H)
            pose = _preview_pose(rng) if preview else _pose(rng)
            cam = CropCamera(pose, fw, fh, int(rng.integers(0, fw - size + 1)), int(rng.integers(0, fh - size + 1)),
                             size, size)
            if noisy is None:
                noisy = renderer.make_session(gs, noisy_st, camera=cam)
                target = renderer.make_session(gs, target_st, camera=cam)
            else:
                noisy.update_camera(cam, cam)
                target.update_camera(cam, cam)
            noisy.render(int(rng.choice([1, 2, 4])))
            target.render(target_spp)
            net.dataset_add(noisy, target)
        noisy.close()
        target.close()
        gs.release()
        if log:
            log(f"scene {k + 1}/{total}: {(k + 1) * crops} examples, {time.perf_counter() - t0:.1f} s")
    return total * crops


def train(net: UNetDenoiser, steps: int, batch: int, lr: float, lr_final: float, seed: int = 0, chunk: int = 250,
          log=None) -> np.ndarray:
    """Cosine-decayed Adam over the device-resident set, in chunks (a log line each)."""
    losses = []
    t0 = time.perf_counter()
    cos = lambda f: lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * f))  # noqa: E731
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        a, b = cos(done / steps), cos(min(1.0, (done + n) / steps))
        losses.append(net.train(n, batch, a, b, seed + done))
        done += n
        if log:
            log(f"step {done}/{steps}: loss {float(losses[-1][-min(n, 50):].mean()):.5f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return np.concatenate(losses)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=48)
    ap.add_argument("--preview-scenes", type=int, default=0)
    ap.add_argument("--crops", type=int, default=32)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=2e-5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    renderer = DenrimRenderer(0)
    net = UNetDenoiser(renderer, seed=args.seed)
    n = generate(renderer, net, args.scenes, args.crops, args.size, args.target_spp, args.seed, log=log,
                 preview_scenes=args.preview_scenes)
    log(f"dataset: {n} examples of {args.size}x{args.size}")
    losses = train(net, args.steps, args.batch, args.lr, args.lr_final, args.seed, log=log)
    net.save(args.out)
    np.save(args.out.replace(".npy", "_losses.npy"), losses)
    log(f"saved {args.out}")


if __name__ == "__main__":
    sys.exit(main())
