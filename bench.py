"""Throughput benchmark of the HIP progressive path tracer on one MI355X.

A step is one frame of the workload at its configured sample count (for the
default 1080p layered-terrain workload: 1920x1080 pixels x 64 spp, max depth
8), rendered one progressive sample at a time on the session stream. The
scene, its BVH and the path-state buffers are resident in HBM before the
timed region; the timed region is K steps enqueued back to back, bracketed by
a device synchronize on both sides and timed both by the host clock and by
HIP events on that stream.

Printed: one JSON line with Mrays/s as `value` (every scene query of the
timed samples -- extension, NEE shadow-chain steps, subsurface-walk segments
-- from the device counter every kernel variant keeps), pixel-samples/s, the
roofline of the dominant kernel, the sparse-SDF and room workloads as
sub-objects, the serving stream, the denoiser, and a CPU baseline (the
oracle on the host's cores over a bounded sample of the same frame).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mrays/s & spp/s at 1080p; preview p50/p99 latency at equal relMSE"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per unit of work, per ray query (the measurement plan):
# 32 B per node visit (a compact node), 36 B per triangle test (positions
# only), 48 B per query (32-B ray in, 16-B hit out). A visit of this build's
# four-wide node reads 128 B; that figure is reported beside the headline as
# `frac_node128`.
NODE_BYTES, NODE4_BYTES, TRI_BYTES, RAY_BYTES = 32, 128, 36, 48
# Sparse-SDF queries: 2 B per FP16 distance sample fetched (8 per march step;
# 8 x 8 bisection + 3 x 64 cubic polish/normal samples per hit), 4 B per
# brick-grid cell visited, 128 B per brick entered.
SAMPLE_BYTES, CELL_BYTES, BRICK_BYTES = 2, 4, 128
# k_shade: per shaded path the path state it reads (origin, direction,
# throughput, radiance, meta, hit: 96 B + its queue entry) and writes (80 B +
# the next queue entry and the NEE throughput); per shadow record 48 B + its
# queue entry.
SHADE_PATH_BYTES, SHADOW_RECORD_BYTES = 200, 52
SAMPLES_PER_STEP, SAMPLES_PER_HIT = 8, 8 * 8 + 3 * 64
STAT_RAY_QUERIES = 36


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(workload_name, kernel):
    """HBM bytes per dispatch of `kernel` on this workload from the committed
    rocprofv3 --pmc summary (profiles/traffic.json): FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, each scaled by the factor the same counter
    shows on a plain device copy of known size. None when not profiled."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["workloads"][workload_name]["kernels"][kernel]
        cal = t["copy_calibration"]
        return int((k["fetch_kib"] * cal["fetch_scale"] + k["write_kib"] * cal["write_scale"]) * 1024)
    except (OSError, KeyError, ValueError):
        return None


# VALU issue peak: a wave64 VALU instruction occupies a 16-lane SIMD for 4
# cycles; 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock.
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def valu_issue(insts, avg_ms):
    """The dominant kernel's VALU issue rate against the issue peak: the
    tracers run out of instruction issue slots, not bandwidth."""
    if not insts or avg_ms <= 0:
        return None
    rate = insts / (avg_ms / 1e3) / 1e9
    return {"insts_per_launch": insts, "achieved_ginst_per_s": round(rate, 1), "peak_ginst_per_s": VALU_PEAK_GINST,
            "frac": round(rate / VALU_PEAK_GINST, 4),
            "note": "peak = wave64 VALU instructions at 4 cycles each, 1024 SIMDs, 2.4 GHz; instructions with half "
                    "of EXEC clear can take fewer passes, so a divergent kernel can pass 1.0"}


def measured_valu(workload_name, kernel):
    """VALU instructions (wave64) per dispatch of `kernel` on this workload
    from the committed rocprofv3 --pmc summary (SQ_INSTS_VALU over the
    dispatches of the same batch shape), or None when not profiled."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return int(json.load(f)["workloads"][workload_name]["kernels"][kernel]["valu_insts"])
    except (OSError, KeyError, ValueError):
        return None


def workload(name):
    from denrimrendererkit_amd import synthetic
    if name == "terrain_1080p":
        return synthetic.layered_terrain_scene(), 1920, 1080, 8, {
            "workload": "layered-terrain 1920x1080, 6,168,640 tris (1M fBm terrain + 256 x 20k displaced spheres), "
                        "6 layered material kinds, 64 emissive tris, 2048x1024 HDRI, max depth 8",
            "width": 1920, "height": 1080, "max_bounces": 8, "triangles": 6168640, "spp": 64}
    if name == "sdf_1024":
        return synthetic.sdf_bricks_scene(resolution=768, primitives=200, hdri=(2048, 1024)), 1024, 1024, 4, {
            "workload": "sparse SDF bricks 1024x1024: virtual 768^3 grid of 8^3 bricks (1-voxel apron, FP16 "
                        "distances), smooth-min union (k=0.05) of 200 spheres and tori, metal/Lambertian, "
                        "2048x1024 HDRI, ground quad, max depth 4",
            "width": 1024, "height": 1024, "max_bounces": 4, "spp": 16}
    if name == "room_320x180":
        return synthetic.cornell_room_scene(), 320, 180, 4, {
            "workload": "room 320x180, 8,334 tris, Lambertian + GGX metal, 1 area light, max depth 4",
            "width": 320, "height": 180, "max_bounces": 4, "triangles": 8334, "spp": 1}
    raise SystemExit(f"unknown workload {name}")


SERVE_RES = ((96, 96), (256, 256), (512, 512), (1280, 720))
SERVE_RES_P = (0.5, 0.25, 0.15, 0.1)
SERVE_SPP = (16, 64, 256)
# Denoised stream: (resolution, requested spp) -> spp rendered before the
# learned denoiser, chosen so the denoised relMSE is at or below plain
# accumulation's at the requested spp on held-out preview scenes
# (profiles/denoiser_relmse_calibration.md); 0 = no denoiser (plain at the
# requested spp: at 96^2 the network does not reach the plain relMSE, nor at
# 256^2 for 256 spp).
SERVE_DENOISED_SPP = {
    (96, 96): {16: 0, 64: 0, 256: 0},
    (256, 256): {16: 1, 64: 2, 256: 0},
    (512, 512): {16: 1, 64: 1, 256: 4},
    (1280, 720): {16: 1, 64: 1, 256: 2},
}


def serving_bench(renderer, n_requests, n_scenes=12, seed=4, max_bounces=4, load=4.0, verify=8):
    """The material-preview request stream: sphere (20k tris) on a plane
    with seeded layered materials, 1024x512 HDRI; resolution from {96^2,
    256^2, 512^2, 1280x720} w.p. {.5, .25, .15, .1}; spp from {16, 64, 256};
    Poisson arrivals at `load` x the one-at-a-time plain service rate (from a
    pilot), so every stream runs saturated and images/s is its throughput.
    Scenes are resident in HBM before timing. The same arrival schedule is
    served (a) one request at a time on a plain session, (b) by the render
    server (continuous batching, graph replays) at the same spp (identical
    images), admitting in arrival order with 16 requests on the GPU and (b')
    least work first with 2, (c) by the server (least work first, 2) with the
    learned denoiser at the spp of SERVE_DENOISED_SPP. relMSE against 4096-spp renders is measured for up to
    `verify` requests of at most 256^2 in (a) and (c)."""
    import threading

    from denrimrendererkit_amd import RenderQuality, RenderSettings, compile_scene, synthetic
    from denrimrendererkit_amd.denoise import DEFAULT_WEIGHTS, UNetDenoiser, relmse
    from denrimrendererkit_amd.renderer import RenderServer
    from denrimrendererkit_amd.settings import DenoiseSettings
    rng = np.random.default_rng(0x5EED0004 + seed)
    scenes = [synthetic.material_preview_scene(k) for k in range(n_scenes)]
    gpu = [renderer.upload(compile_scene(sc)) for sc in scenes]
    net = UNetDenoiser.pretrained(renderer, DEFAULT_WEIGHTS)
    reqs = []
    for _ in range(n_requests):
        W, H = SERVE_RES[rng.choice(4, p=SERVE_RES_P)]
        reqs.append((int(rng.integers(n_scenes)), W, H, int(SERVE_SPP[rng.integers(3)])))
    check = [k for k, (_, W, H, _) in enumerate(reqs) if W * H <= 256 * 256][:verify]

    def settings(W, H, denoise=False):
        return RenderSettings(width=W, height=H, max_bounces=max_bounces, quality=RenderQuality.FINAL, frame_seed=0,
                              denoise=DenoiseSettings.learned() if denoise else DenoiseSettings())

    def plain(i, W, H, spp):
        sess = renderer.make_session(gpu[i], settings(W, H), camera=scenes[i].camera)
        sess.render(spp)
        px = sess.pixels()
        sess.close()
        return px

    pilot = [reqs[k] for k in range(min(6, n_requests))]
    t0 = time.perf_counter()
    for r in pilot:
        plain(*r)
    service = (time.perf_counter() - t0) / len(pilot)
    rate = load / max(service, 1e-6)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]

    def pct(x, q):
        return round(float(np.percentile(np.asarray(x) * 1e3, q)), 2)

    # (a) one at a time, in arrival order
    lat, images = [], {}
    t0 = time.perf_counter()
    for k, (r, a) in enumerate(zip(reqs, arrivals)):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        px = plain(*r)
        if k in check:
            images[k] = px
        lat.append(time.perf_counter() - t0 - a)
    span = time.perf_counter() - t0
    out = {"plain": {"images_per_s": round(n_requests / span, 3), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99)}}

    def served(denoise, admission="arrival", max_active=16):
        server = RenderServer(renderer, memory_budget_bytes=24 << 30, max_active=max_active, admission=admission)
        server.set_denoiser_network(net)
        outstanding, lat, errors, got = {}, [], [], {}
        lock = threading.Lock()
        submitted = threading.Event()

        def collect():  # reads each finished request's beauty plane and releases it
            while True:
                with lock:
                    items = list(outstanding.items())
                if not items and submitted.is_set():
                    return
                for t, (k, a) in items:
                    st = server.state(t)
                    if st in ("done", "failed"):
                        if st == "done":
                            px = server.result(t)
                            if k in check:
                                got[k] = px
                        else:
                            errors.append(t)
                        server.release(t)
                        lat.append(time.perf_counter() - t0 - a)
                        with lock:
                            del outstanding[t]
                time.sleep(2e-4)

        t0 = time.perf_counter()
        collector = threading.Thread(target=collect)
        collector.start()
        for k, ((i, W, H, spp), a) in enumerate(zip(reqs, arrivals)):
            now = time.perf_counter() - t0
            if now < a:
                time.sleep(a - now)
            d_spp = SERVE_DENOISED_SPP[(W, H)][spp] if denoise else 0
            st = settings(W, H, denoise=d_spp > 0)
            t = server.submit(gpu[i], st, scenes[i].camera, d_spp or spp)
            with lock:
                outstanding[t] = (k, a)
        submitted.set()
        collector.join()
        span = time.perf_counter() - t0
        server.close()
        if errors:
            raise RuntimeError("served requests failed: %s" % errors)
        return {"images_per_s": round(n_requests / span, 3), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99)}, got

    # 16 requests on the GPU at once in arrival order; 2 at once, least work
    # first (short previews are not held behind long renders, and with each
    # request batching up to 4 M paths two streams fill the GPU).
    out["served"], _ = served(False)
    out["served"]["admission"] = "arrival, 16 active"
    out["served_least_work"], _ = served(False, "least_work", 2)
    out["served_least_work"]["admission"] = "least work first, 2 active"
    out["served_denoised"], den_images = served(True, "least_work", 2)
    out["served_denoised"]["admission"] = "least work first, 2 active"
    # relMSE of the verified requests against 4096-spp renders (another frame seed).
    rel_plain, rel_den = [], []
    for k in check:
        i, W, H, spp = reqs[k]
        ref = renderer.make_session(gpu[i], RenderSettings(width=W, height=H, max_bounces=max_bounces,
                                                           quality=RenderQuality.FINAL, frame_seed=99),
                                    camera=scenes[i].camera)
        ref.render(4096)
        target = ref.pixels()
        ref.close()
        rel_plain.append(relmse(images[k], target))
        rel_den.append(relmse(den_images[k], target))
    out["relmse_verified"] = {"requests": len(check), "plain_mean": round(float(np.mean(rel_plain)), 5),
                              "served_denoised_mean": round(float(np.mean(rel_den)), 5)}
    out.update({"requests": n_requests, "arrival_rate_per_s": round(rate, 3), "load_vs_plain": load,
                "speedup_served": round(out["served"]["images_per_s"] / out["plain"]["images_per_s"], 3),
                "speedup_served_least_work": round(out["served_least_work"]["images_per_s"] /
                                                   out["plain"]["images_per_s"], 3),
                "speedup_served_denoised": round(out["served_denoised"]["images_per_s"] /
                                                 out["plain"]["images_per_s"], 3),
                "mix": "res {96^2:.5, 256^2:.25, 512^2:.15, 1280x720:.1}, spp {16,64,256}, max depth %d" %
                       max_bounces})
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0


def denoiser_bench(renderer, frames, spp=4):
    """Learned-denoiser inference on a 1920x1080 config-1-style room at
    `spp` spp: the U-Net forward pass (19 BF16 MFMA convolutions + pooling,
    feature packing and output) enqueued `frames` times on the session
    stream, timed by the host clock around a synchronize; FLOPs from the
    layer table (2 * H * W * 9 * Cin * Cout per convolution)."""
    import os as _os
    from denrimrendererkit_amd import RenderQuality, RenderSettings, compile_scene, synthetic
    from denrimrendererkit_amd.denoise import DEFAULT_WEIGHTS, UNetDenoiser, inference_flops
    from denrimrendererkit_amd.settings import DenoiseSettings
    W, H = 1920, 1080
    scene = synthetic.room_variant(2001)
    gs = renderer.upload(compile_scene(scene))
    net = UNetDenoiser.pretrained(renderer) if _os.path.exists(DEFAULT_WEIGHTS) else UNetDenoiser(renderer, seed=0)
    st = RenderSettings(width=W, height=H, max_bounces=4, quality=RenderQuality.FINAL, frame_seed=0,
                        denoise=DenoiseSettings.learned())
    sess = renderer.make_session(gs, st, camera=scene.camera)
    sess.render(spp)
    sess.set_denoiser_network(net)
    sess.denoise()  # warm-up: workspace allocation
    sess.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        sess.set_denoiser_network(net)  # marks the denoised plane stale
        sess.denoise()
    sess.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / frames
    flops = inference_flops(W, H)
    tflops = flops / (ms * 1e-3) / 1e12
    sess.close()
    return {"workload": f"U-Net inference {W}x{H} ({W}x1088 padded), {spp}-spp room input",
            "ms_per_frame": round(ms, 3), "gflop_per_frame": round(flops / 1e9, 2),
            "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)},
            "trained_weights": _os.path.exists(DEFAULT_WEIGHTS)}


def cpu_baseline(build, camera, W, H, bounces, target_seconds):
    """The oracle (C++ restatement, FP32, OpenMP over rows) on this host's
    cores over a band of rows of the same frame at 1 spp."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    threads = os.cpu_count() or 1
    osc = O.OracleScene(build)
    cam = camera.gpu_camera(W, H)
    rows = 2
    spent = 0.0
    while True:
        y0 = H // 2 - rows // 2
        st = O.make_settings(build, W, H, max_bounces=bounces, frame_seed=0, tile=(0, y0, W, rows))
        t0 = time.perf_counter()
        r = O.render(osc, st, cam, sample_count=1, threads=threads, with_aovs=False)
        dt = time.perf_counter() - t0
        spent += dt
        if dt >= target_seconds * 0.5 or rows >= H or spent > target_seconds * 2:
            break
        rows = min(H, max(rows * 2, int(rows * target_seconds / max(dt, 1e-3))))
    c = r.counters
    queries = int(c[O.CNT_PRIMARY] + c[O.CNT_BOUNCE] + c[O.CNT_SHADOW])
    return {"value": round(queries / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{rows} rows x {W} px x 1 spp of the same frame ({queries} ray queries, {dt:.2f} s); "
                      f"oracle FP32 restatement of the reference flat-BVH path tracer, OpenMP {threads} threads",
            "spp_per_s": round(rows * W / dt, 1)}


def measure(renderer, name, steps, warmup, spp_per_step=None, env_mode=None, isect_mode=None):
    """Renders `warmup` then `steps` frames of workload `name` (spp_per_step
    progressive samples each, default: the workload's configured spp) and
    returns the measurement dict: timed ray queries from the device counter,
    the dominant kernel's roofline from a counting sample's node, triangle
    and SDF counters, and the scene objects for the CPU baseline."""
    from denrimrendererkit_amd import RenderQuality, RenderSettings, compile_scene
    from denrimrendererkit_amd.settings import EnvironmentSampling, TriangleIntersection
    env_mode = env_mode or EnvironmentSampling.CDF
    isect_mode = isect_mode or TriangleIntersection.MOLLER_TRUMBORE
    t0 = time.perf_counter()
    scene, W, H, bounces, config = workload(name)
    spp = spp_per_step or config["spp"]
    build = compile_scene(scene)
    log(f"{name}: scene compiled: {len(build.triangles)} tris, {len(build.volume_bricks)} bricks in "
        f"{time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    gscene = renderer.upload(build)
    info = gscene.info()
    log(f"{name}: uploaded + BVH: {info} in {time.perf_counter() - t1:.1f} s")

    def settings(counting):
        return RenderSettings(width=W, height=H, max_bounces=bounces, quality=RenderQuality.FINAL, frame_seed=0,
                              collects_sdf_traversal_stats=counting, intersection=isect_mode,
                              environment_sampling=env_mode)

    # Counting sample: node fetches, triangle tests and SDF work per sample.
    cs = renderer.make_session(gscene, settings(True), camera=scene.camera, samples_per_launch=1)
    cs.render(1)
    c = cs.stats().astype(np.float64)
    cs.close()
    assert c[STAT_RAY_QUERIES] == c[0] + c[1] + c[2], "query counter disagrees with the counting kernels"
    sdf_ext, sdf_sh = c[8:18], c[18:28]
    counted_queries = c[0] + c[1] + c[2]

    # samples_per_launch left to the library (0): about 64 M path slots per
    # batch, so a 64-spp 1080p frame runs as 2 batches of 32 samples.
    sess = renderer.make_session(gscene, settings(False), camera=scene.camera, samples_per_launch=0)
    sess.encode_next_sample(warmup * spp)
    sess.synchronize()
    sess.reset_stats()
    sess.set_stage_timing(True)
    sess.synchronize()
    w0 = time.perf_counter()
    sess.encode_next_sample(steps * spp)
    sess.synchronize()
    wall = time.perf_counter() - w0
    tm = sess.timing()
    queries = float(sess.stats()[STAT_RAY_QUERIES])
    gpu_s = tm["total_ms"] / 1e3
    sec = max(wall, gpu_s)
    samples = steps * spp
    r0 = time.perf_counter()
    beauty = sess.pixels()  # one beauty plane over PCIe: what a caller reads after the K frames
    readback_s = time.perf_counter() - r0
    assert np.isfinite(beauty).all(), "non-finite beauty"
    sess.close()

    # Roofline of the dominant kernel (largest share of stage time). On
    # triangle scenes k_sss also traces the shadow rays queued by k_shade
    # (counters 30-32) and the next bounce's k_extend those of resumed paths
    # (33-35); k_shadow traces the last bounce's rest.
    st_ms, st_n = tm["stage_ms"], tm["stage_launches"]
    dom = max(("extend", "shadow", "shade", "sss"), key=lambda k: st_ms[k])
    if dom == "shadow":
        nodes, tris, rays, sdf = c[5] - c[31] - c[34], c[6] - c[32] - c[35], c[2] - c[30] - c[33], sdf_sh
    elif dom == "sss":
        nodes, tris, rays, sdf = c[28] + c[31], c[29] + c[32], c[7] + c[30], sdf_ext * 0.0
    elif dom == "extend":
        nodes, tris, rays, sdf = c[3] + c[34], c[4] + c[35], c[0] + c[1] - c[7] + c[33], sdf_ext
    else:  # k_shade: no traversal; path state and shadow records
        nodes, tris, rays, sdf = 0.0, 0.0, 0.0, sdf_ext * 0.0
    # Counters are per sample (the counting sample); the timed launches of
    # the dominant stage carry samples x those bytes between them.
    launch_share = samples / max(1, st_n[dom])
    sdf_bytes = (SAMPLE_BYTES * (SAMPLES_PER_STEP * sdf[8] + SAMPLES_PER_HIT * sdf[9]) + CELL_BYTES * sdf[2] +
                 BRICK_BYTES * sdf[4])
    tri_bytes = TRI_BYTES * tris + RAY_BYTES * rays
    shade_bytes = (SHADE_PATH_BYTES * (c[0] + c[1] - c[7]) + SHADOW_RECORD_BYTES * c[2]) if dom == "shade" else 0.0
    bytes_per_launch = (NODE_BYTES * nodes + tri_bytes + sdf_bytes + shade_bytes) * launch_share
    bytes4_per_launch = (NODE4_BYTES * nodes + tri_bytes + sdf_bytes + shade_bytes) * launch_share
    avg_ms = st_ms[dom] / max(1, st_n[dom])
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    kname = {"extend": "k_extend", "shadow": "k_shadow", "shade": "k_shade", "sss": "k_sss"}[dom]
    roofline = {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(name, kname),
                "valu_issue": valu_issue(measured_valu(name, kname), avg_ms),
                "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": int(bytes_per_launch),
                "frac_node128": round(bytes4_per_launch / (avg_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                "bytes_per_launch_node128": int(bytes4_per_launch),
                "bytes_note": "32 B per node visit + 36 B per triangle test + 48 B per query (+ SDF: 2 B per "
                              "distance sample, 4 B per grid cell, 128 B per brick); frac_node128 charges the "
                              "128 B a four-wide node visit reads; k_shade: 200 B per shaded path + 52 B per "
                              "shadow record",
                "stage_ms_per_sample": {k: round(v / samples, 3) for k, v in st_ms.items()}}
    batches = max(1, st_n["raygen"])
    config.update({"spp_per_step": spp, "frame_seed": 0, "bvh_nodes": info["nodes"],
                   "samples_per_launch": round(samples / batches, 2),
                   "ray_queries_per_sample": int(round(queries / samples)),
                   "counting_sample_queries": int(counted_queries),
                   "rays_per_path": round(queries / samples / (W * H), 3)})
    if len(build.volume_bricks):
        tot = sdf_ext + sdf_sh
        config.update({"bricks": int(len(build.volume_bricks)), "brick_samples": int(len(build.volume_brick_samples)),
                       "sdf_per_sample": {"grid_cells": int(tot[2]), "macro_skips": int(tot[3]),
                                          "brick_tests": int(tot[4]), "marches": int(tot[7]),
                                          "march_steps": int(tot[8]), "hits": int(tot[9])}})
    return {"value": round(queries / sec / 1e6, 2), "unit": "Mrays/s", "steps": steps, "warmup": warmup,
            "ms_per_step": round(sec * 1e3 / steps, 3), "spp_per_s": round(W * H * samples / sec, 1),
            "wall_s": round(wall, 4), "gpu_event_s": round(gpu_s, 4), "readback_ms": round(readback_s * 1e3, 3),
            "pcie_inclusive_mrays_per_s": round(queries / (sec + readback_s) / 1e6, 2),
            "config": config, "roofline": roofline,
            "_scene": (build, scene.camera, W, H, bounces)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="terrain_1080p")
    ap.add_argument("--spp-per-step", type=int, default=0,
                    help="progressive samples per step (0: the workload's configured spp, 64 for the terrain)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--environment-sampling", choices=("cdf", "alias"), default="cdf",
                    help="HDRI texel selection of the timed render (the reference's is a CDF search)")
    ap.add_argument("--intersection", choices=("moller-trumbore", "watertight"), default="moller-trumbore",
                    help="triangle test of the timed render (the reference's is Moller-Trumbore)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sub-workloads", action="store_true",
                    help="skip the sparse-SDF and room workloads measured beside the main one")
    ap.add_argument("--serving-requests", type=int, default=48,
                    help="requests in the material-preview serving run (0: skip)")
    ap.add_argument("--denoiser-frames", type=int, default=10,
                    help="1080p learned-denoiser inferences to time (0: skip)")
    ap.add_argument("--denoiser-only", action="store_true",
                    help="run only the 1080p learned-denoiser inference leg (--denoiser-frames times) and print "
                         "its JSON object: the command for kernel profiles of the U-Net")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 256 MiB (5 times): the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 << 20  # float32 elements: 256 MiB
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        assert bool((dst[-4:] == src[-4:]).all())
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every GPU box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from denrimrendererkit_amd import DenrimRenderer
    from denrimrendererkit_amd.settings import EnvironmentSampling, TriangleIntersection
    renderer = DenrimRenderer(0)
    if args.denoiser_only:
        print(json.dumps(denoiser_bench(renderer, max(1, args.denoiser_frames))), flush=True)
        return
    isect = TriangleIntersection.WATERTIGHT if args.intersection == "watertight" else TriangleIntersection.MOLLER_TRUMBORE
    env = EnvironmentSampling.ALIAS if args.environment_sampling == "alias" else EnvironmentSampling.CDF
    main_run = measure(renderer, args.workload, args.steps, args.warmup, args.spp_per_step or None, env, isect)
    main_run["config"].update({"intersection": args.intersection, "environment_sampling": args.environment_sampling})
    log(f"{args.workload}: {main_run['value']} Mrays/s, {main_run['ms_per_step']} ms per step")

    subs = {}
    if not args.no_sub_workloads:
        for name, steps in (("sdf_1024", 2), ("room_320x180", 64)):
            if name == args.workload:
                continue
            r = measure(renderer, name, steps, 1)
            r.pop("_scene")
            subs[name] = {k: r[k] for k in ("value", "unit", "steps", "ms_per_step", "spp_per_s", "config", "roofline")}
            log(f"{name}: {r['value']} Mrays/s, {r['ms_per_step']} ms per step")

    serving = None
    if args.serving_requests > 0:
        serving = serving_bench(renderer, args.serving_requests)
        log(f"serving: {serving}")

    denoiser = None
    if args.denoiser_frames > 0:
        denoiser = denoiser_bench(renderer, args.denoiser_frames)
        log(f"denoiser: {denoiser}")

    build, camera, W, H, bounces = main_run.pop("_scene")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(build, camera, W, H, bounces, args.cpu_seconds)

    out = {"metric": METRIC, "value": main_run["value"], "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": main_run["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    out.update({k: main_run[k] for k in ("spp_per_s", "wall_s", "gpu_event_s", "readback_ms",
                                         "pcie_inclusive_mrays_per_s", "config", "roofline")})
    out["cpu_baseline"] = cpu
    if subs:
        out["workloads"] = subs
    if denoiser is not None:
        out["denoiser"] = denoiser
    if serving is not None:
        out["preview_p50_ms"] = serving["served_denoised"]["p50_ms"]
        out["preview_p99_ms"] = serving["served_denoised"]["p99_ms"]
        out["serving"] = serving
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
