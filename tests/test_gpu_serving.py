"""GPU tests of batch graphs and the render server.

  * Graph replays of sample batches equal the launch-by-launch path
    bit for bit, over several calls with eager work and synchronizations in
    between (the device-side sample base advances inside the graph).
  * Served requests: each result equals the same request rendered alone on
    a plain session, bit for bit (tolerance: exact), whether it ran alone or
    in a mix of sizes and sample counts admitted while others were running;
    the memory budget bounds what runs at once.
  * A retargeted session (another scene, size, settings and camera in the
    same device memory) renders exactly what a new session renders, and
    refuses settings that need more memory than it holds.
  * Learned-denoiser requests' U-Net workspaces count against the budget
    (server count and device free memory), mixed with plain requests.
  * Injected encode failures: requests released the moment their wait
    returns, failed or done, while the scheduler keeps serving.
"""
import numpy as np
import pytest

from denrimrendererkit_amd import RenderQuality, RenderScene, RenderSettings, compile_scene, synthetic
from denrimrendererkit_amd.renderer import RenderServer
from denrimrendererkit_amd.settings import RenderOutput

pytestmark = pytest.mark.gpu


def _settings(W, H, bounces=4):
    return RenderSettings(width=W, height=H, max_bounces=bounces, quality=RenderQuality.FINAL, frame_seed=7)


def test_graph_replays_match_eager(renderer):
    scene = RenderScene.cornell_box()
    gs = renderer.upload(compile_scene(scene))
    st = _settings(40, 32)
    eager = renderer.make_session(gs, st, camera=scene.camera, samples_per_launch=3)
    graph = renderer.make_session(gs, st, camera=scene.camera, samples_per_launch=3)
    graph.set_graphs(True)
    for n in (3, 3, 5, 1, 3):  # full batches, a partial batch, a single sample
        eager.render(n)
        graph.render(n)
        eager.pixels()  # eager work and synchronization between replays
        for out in (RenderOutput.BEAUTY, RenderOutput.DEPTH, RenderOutput.NORMAL, RenderOutput.MATERIAL_ID):
            assert np.array_equal(graph.pixels(out), eager.pixels(out)), (n, out)
        assert graph.sample_count == eager.sample_count
    graph.reset_accumulation()
    eager.reset_accumulation()
    graph.render(4)
    eager.render(4)
    assert np.array_equal(graph.pixels(), eager.pixels())


def test_retarget_equals_new_session(renderer):
    scenes = [synthetic.material_preview_scene(seed, hdri=(64, 32), segments=20) for seed in (3, 4)]
    gs = [renderer.upload(compile_scene(s)) for s in scenes]
    sess = renderer.make_session(gs[0], _settings(64, 48), camera=scenes[0].camera, samples_per_launch=4)
    sess.render(6)
    for i, (W, H, spp, k) in ((1, (48, 40, 5, 2)), (0, (32, 32, 3, 3)), (1, (64, 48, 4, 4))):
        st = _settings(W, H, bounces=3)
        sess.retarget(gs[i], st, scenes[i].camera, samples_per_launch=k)
        sess.render(spp)
        fresh = renderer.make_session(gs[i], st, camera=scenes[i].camera, samples_per_launch=k)
        fresh.render(spp)
        assert sess.sample_count == fresh.sample_count == spp
        for out in (RenderOutput.BEAUTY, RenderOutput.DEPTH, RenderOutput.OBJECT_ID):
            assert np.array_equal(sess.pixels(out), fresh.pixels(out)), (i, W, H, out)
        assert np.array_equal(sess.stats(), fresh.stats())
    from denrimrendererkit_amd.scene import DenrimRendererError
    with pytest.raises(DenrimRendererError):
        sess.retarget(gs[0], _settings(128, 128), scenes[0].camera, samples_per_launch=4)


def _alone(renderer, gs, camera, st, spp):
    s = renderer.make_session(gs, st, camera=camera)
    s.render(spp)
    return s.pixels(RenderOutput.BEAUTY), s.pixels(RenderOutput.ALBEDO)


@pytest.mark.parametrize("admission", ["arrival", "least_work"])
def test_served_requests_equal_alone_in_mix(renderer, admission):
    scenes = [synthetic.material_preview_scene(seed, hdri=(64, 32), segments=24) for seed in range(3)]
    uploaded = [renderer.upload(compile_scene(s)) for s in scenes]
    reqs = [(0, 48, 48, 8), (1, 64, 32, 4), (2, 24, 40, 16), (0, 32, 32, 2), (1, 48, 48, 8), (2, 64, 64, 3)]
    server = RenderServer(renderer, memory_budget_bytes=4 << 30, max_active=3, admission=admission)
    try:
        tickets = [server.submit(uploaded[i], _settings(W, H), scenes[i].camera, spp) for i, W, H, spp in reqs]
        for t in tickets:
            assert server.wait(t) == "done"
        for t, (i, W, H, spp) in zip(tickets, reqs):
            beauty, albedo = _alone(renderer, uploaded[i], scenes[i].camera, _settings(W, H), spp)
            assert np.array_equal(server.result(t), beauty), t
            assert np.array_equal(server.result(t, RenderOutput.ALBEDO), albedo), t
            tm = server.timing(t)
            assert tm["run_ms"] > 0 and tm["queued_ms"] >= 0
            server.release(t)
        # Alone: one request on an idle server gives the same pixels again.
        i, W, H, spp = reqs[4]
        t = server.submit(uploaded[i], _settings(W, H), scenes[i].camera, spp)
        assert server.wait(t) == "done"
        assert np.array_equal(server.result(t), _alone(renderer, uploaded[i], scenes[i].camera, _settings(W, H), spp)[0])
        server.release(t)
        use = server.usage()
        assert use["active"] == 0 and use["queued"] == 0
        assert 0 < use["used_bytes"] <= 4 << 30  # pooled sessions stay within the budget
    finally:
        server.close()


@pytest.mark.parametrize("admission", ["arrival", "least_work"])
def test_admission_order(renderer, admission):
    # One request on the GPU at a time: a long request A runs, a long B and
    # a short C queue behind it. Arrival order runs B before C; least work
    # first runs C before B, so B is still unfinished when C is done.
    scene = synthetic.material_preview_scene(2, hdri=(64, 32), segments=24)
    gs = renderer.upload(compile_scene(scene))
    server = RenderServer(renderer, memory_budget_bytes=2 << 30, max_active=1, admission=admission)
    try:
        a = server.submit(gs, _settings(256, 256), scene.camera, 64)
        b = server.submit(gs, _settings(256, 256), scene.camera, 64)
        c = server.submit(gs, _settings(16, 16), scene.camera, 1)
        assert server.wait(c) == "done"
        state_b = server.state(b)
        assert state_b in (("done",) if admission == "arrival" else ("queued", "running")), state_b
        for t in (a, b):
            assert server.wait(t) == "done"
        beauty, _ = _alone(renderer, gs, scene.camera, _settings(16, 16), 1)
        assert np.array_equal(server.result(c), beauty)
        for t in (a, b, c):
            server.release(t)
    finally:
        server.close()
    with pytest.raises(ValueError):
        RenderServer(renderer, admission="shortest")


def test_server_budget_limits_concurrency(renderer):
    scene = synthetic.material_preview_scene(5, hdri=(64, 32), segments=16)
    gs = renderer.upload(compile_scene(scene))
    st = _settings(32, 32, bounces=2)
    from denrimrendererkit_amd import _lib
    import ctypes as C
    from denrimrendererkit_amd.renderer import settings_struct
    need = C.c_uint64()
    ss = settings_struct(st, 4)
    _lib.check(_lib.load().dr_hip_session_estimate_bytes(gs._h, C.byref(ss), C.byref(need)), "estimate")
    server = RenderServer(renderer, memory_budget_bytes=int(need.value * 1.5), max_active=8)
    try:
        tickets = [server.submit(gs, st, scene.camera, 4, samples_per_launch=4) for _ in range(4)]
        for t in tickets:
            assert server.wait(t) == "done"
        ref = server.result(tickets[0])
        for t in tickets[1:]:
            assert np.array_equal(server.result(t), ref)
        too_big = server.submit(gs, _settings(512, 512), scene.camera, 4, samples_per_launch=64)
        assert server.wait(too_big) == "failed"
    finally:
        server.close()


def test_request_holds_released_scene(renderer):
    # A queued request keeps its scene alive after the caller releases it.
    scene = synthetic.material_preview_scene(9, hdri=(64, 32), segments=16)
    st = _settings(40, 40)
    keep = renderer.upload(compile_scene(scene))
    alone = _alone(renderer, keep, scene.camera, st, 4)[0]
    gs = renderer.upload(compile_scene(scene))
    server = RenderServer(renderer, memory_budget_bytes=1 << 30, max_active=1)
    try:
        busy = server.submit(keep, _settings(256, 256), scene.camera, 32)
        t = server.submit(gs, st, scene.camera, 4)
        server._keep.pop(t)
        gs.release()
        assert server.wait(t) == "done" and server.wait(busy) == "done"
        assert np.array_equal(server.result(t), alone)
        server.release(t)
    finally:
        server.close()


def test_denoised_requests_equal_alone_in_mix(renderer):
    # Served requests with the learned denoiser: denoised on the device when
    # their samples are done; pixels equal a session with the same network.
    from denrimrendererkit_amd.denoise import UNetDenoiser
    from denrimrendererkit_amd.settings import DenoiseSettings
    net = UNetDenoiser(renderer, seed=11)
    scenes = [synthetic.material_preview_scene(seed, hdri=(64, 32), segments=20) for seed in (6, 7)]
    uploaded = [renderer.upload(compile_scene(s)) for s in scenes]

    def st(W, H, den):
        s = _settings(W, H)
        if den:
            s.denoise = DenoiseSettings.learned()
        return s

    reqs = [(0, 48, 32, 2, True), (1, 64, 64, 4, False), (1, 40, 24, 1, True), (0, 32, 48, 3, True)]
    server = RenderServer(renderer, memory_budget_bytes=2 << 30, max_active=3)
    server.set_denoiser_network(net)
    try:
        tickets = [server.submit(uploaded[i], st(W, H, d), scenes[i].camera, spp) for i, W, H, spp, d in reqs]
        for t, (i, W, H, spp, d) in zip(tickets, reqs):
            assert server.wait(t) == "done"
            alone = renderer.make_session(uploaded[i], st(W, H, d), camera=scenes[i].camera)
            if d:
                alone.set_denoiser_network(net)
            alone.render(spp)
            assert np.array_equal(server.result(t), alone.pixels()), (t, d)
            server.release(t)
    finally:
        server.close()


def _device_free_bytes():
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_learned_and_plain_requests_stay_within_budget(renderer):
    # The U-Net workspace of learned requests counts against the budget: a
    # pooled session is handed to a learned request only when it already
    # holds the workspace, a new one gets it at once. Alternating learned and
    # plain requests never hold more device memory than the budget (the
    # server's own count and the device's free memory agree), and the
    # results equal sessions rendering them alone.
    import ctypes as C
    from denrimrendererkit_amd import _lib
    from denrimrendererkit_amd.denoise import UNetDenoiser
    from denrimrendererkit_amd.renderer import settings_struct
    from denrimrendererkit_amd.settings import DenoiseSettings
    net = UNetDenoiser(renderer, seed=5)
    scene = synthetic.material_preview_scene(12, hdri=(64, 32), segments=16)
    gs = renderer.upload(compile_scene(scene))

    def st(W, H, den):
        s = _settings(W, H, bounces=3)
        if den:
            s.denoise = DenoiseSettings.learned()
        return s

    reqs = [(64, 64, False), (96, 96, True), (64, 64, False), (80, 48, True), (96, 96, False), (64, 64, True),
            (48, 48, False), (96, 96, True)]

    def need(W, H, den):
        b = C.c_uint64()
        ss = settings_struct(st(W, H, den), 2)
        _lib.check(_lib.load().dr_hip_session_estimate_bytes(gs._h, C.byref(ss), C.byref(b)), "estimate")
        return b.value

    budget = int(max(need(W, H, d) for W, H, d in reqs) * 1.7)
    free0 = _device_free_bytes()
    server = RenderServer(renderer, memory_budget_bytes=budget, max_active=4)
    server.set_denoiser_network(net)
    try:
        tickets = [server.submit(gs, st(W, H, d), scene.camera, 2, samples_per_launch=2) for W, H, d in reqs]
        peak_used, peak_dev = 0, 0
        for t in tickets:
            while server.state(t) not in ("done", "failed"):
                peak_used = max(peak_used, server.usage()["used_bytes"])
                peak_dev = max(peak_dev, free0 - _device_free_bytes())
        peak_used = max(peak_used, server.usage()["used_bytes"])
        peak_dev = max(peak_dev, free0 - _device_free_bytes())
        assert peak_used <= budget, (peak_used, budget)
        # stream, event and graph objects take a little device memory beside the sessions
        assert peak_dev <= budget + (64 << 20), (peak_dev, budget)
        for t, (W, H, d) in zip(tickets, reqs):
            assert server.wait(t) == "done"
            alone = renderer.make_session(gs, st(W, H, d), camera=scene.camera, samples_per_launch=2)
            if d:
                alone.set_denoiser_network(net)
            alone.render(2)
            assert np.array_equal(server.result(t), alone.pixels()), (t, W, H, d)
            server.release(t)
    finally:
        server.close()


def test_failed_encode_released_at_once(renderer, monkeypatch):
    # Every 7th encode call fails (fault injection read at server creation).
    # Each request is released as soon as its wait returns, failed or done;
    # the scheduler must not touch a failed request after publishing it.
    monkeypatch.setenv("DENRIM_HIP_SERVER_FAIL_ENCODE", "7")
    scene = synthetic.material_preview_scene(13, hdri=(64, 32), segments=16)
    gs = renderer.upload(compile_scene(scene))
    server = RenderServer(renderer, memory_budget_bytes=1 << 30, max_active=4, samples_per_batch=1)
    monkeypatch.delenv("DENRIM_HIP_SERVER_FAIL_ENCODE")
    alone = _alone(renderer, gs, scene.camera, _settings(32, 32), 3)[0]
    try:
        tickets = [server.submit(gs, _settings(32, 32), scene.camera, 3) for _ in range(14)]
        states = []
        for t in tickets:
            s = server.wait(t)
            if s == "done":
                assert np.array_equal(server.result(t), alone)
            server.release(t)
            states.append(s)
        assert states.count("failed") >= 1 and states.count("done") >= 1, states
        t = server.submit(gs, _settings(32, 32), scene.camera, 3)
        s = server.wait(t)
        server.release(t)
        assert s in ("done", "failed")
    finally:
        server.close()


def test_material_preview_stream_as_configured(renderer):
    # The material-preview serving workload at its stated sizes: a 20k-
    # triangle sphere on a 2-triangle plane with the 1024x512 synthetic HDRI
    # and seeded layered materials; one request each of 96^2, 256^2, 512^2
    # and 1280x720 at 16 spp submitted together. Each served result equals
    # the same request rendered alone, bit for bit; a 96^2 request at 1 spp
    # equals the oracle (IDs exact, beauty within the 1-spp bound).
    import oracle as O
    scenes = [synthetic.material_preview_scene(seed) for seed in range(4)]
    builds = [compile_scene(s) for s in scenes]
    assert 20_000 <= len(builds[0].triangles) <= 20_300 and builds[0].environment_samples.shape[0] == 1024 * 512
    uploaded = [renderer.upload(b) for b in builds]
    reqs = [(0, 96, 96), (1, 256, 256), (2, 512, 512), (3, 1280, 720)]
    server = RenderServer(renderer, memory_budget_bytes=24 << 30, max_active=4)
    try:
        tickets = [server.submit(uploaded[i], _settings(W, H), scenes[i].camera, 16) for i, W, H in reqs]
        for t, (i, W, H) in zip(tickets, reqs):
            assert server.wait(t) == "done"
            beauty, albedo = _alone(renderer, uploaded[i], scenes[i].camera, _settings(W, H), 16)
            assert np.array_equal(server.result(t), beauty), (W, H)
            assert np.array_equal(server.result(t, RenderOutput.ALBEDO), albedo), (W, H)
            server.release(t)
    finally:
        server.close()
    st = _settings(96, 96)
    sess = renderer.make_session(uploaded[0], st, camera=scenes[0].camera)
    sess.render(1)
    ref = O.render(O.OracleScene(builds[0]), O.make_settings(builds[0], 96, 96, max_bounces=4, frame_seed=7, clamp=64.0),
                   scenes[0].camera.gpu_camera(96, 96), sample_count=1)
    assert np.array_equal(sess.triangle_ids(), ref.aovs[6][..., 0].view(np.uint32))
    assert np.array_equal(sess.pixels(RenderOutput.MATERIAL_ID), ref.aovs[3])
    g, r = sess.pixels(RenderOutput.BEAUTY), ref.accumulation
    d = np.abs(g - r)
    assert (d > 1e-3 * (1 + np.abs(r))).mean() <= 0.01 and d.mean() < 2e-3
