"""GPU test of dr_hip_scene_intersect_device, the ray-query entry point that
allocates nothing (its traversal-stack spill is a static device array): its
results equal dr_hip_scene_intersect's bit for bit, on two calls in a row on
the same device (the second smaller), on a scene deep enough that the stack
spills and with enough rays that the queries run in several launches. Device
buffers for the test come from hipMalloc through ctypes (the test may
allocate; the entry point may not)."""
import ctypes as C

import numpy as np
import pytest

from denrimrendererkit_amd import compile_scene, synthetic

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Dev:
    """hipMalloc'ed buffer of the HIP runtime the library uses."""
    hip = None

    def __init__(self, nbytes):
        if _Dev.hip is None:
            _Dev.hip = C.CDLL("libamdhip64.so.7")
        self.ptr = C.c_void_p()
        assert _Dev.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert _Dev.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, a):
        assert _Dev.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        _Dev.hip.hipFree(self.ptr)


def _rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    r[:, 4:7] = rng.standard_normal((n, 3))
    return r


@pytest.mark.parametrize("mode", [0, 1])
def test_intersect_device_equals_host_entry_point(renderer, mode):
    cases = [(synthetic.cornell_room_scene(), [-1.9, 0.1, -1.9], [1.9, 2.9, 1.9]),
             (synthetic.layered_terrain_scene(scale=1 / 16, hdri=(64, 32)), [-6.0, 0.5, -3.0], [6.0, 3.0, 3.0])]
    for scene, lo, hi in cases:
        gs = renderer.upload(compile_scene(scene))
        depth = gs.info()["max_depth"]
        assert 7 * depth + 2 > 24, "the scene must be deep enough that the traversal stack spills"
        n_big = 400_000
        rays_d, t_d, ids_d = _Dev(n_big * 32), _Dev(n_big * 4), _Dev(n_big * 4)
        try:
            for call, n in enumerate((n_big, 5000)):  # two calls in a row, the second smaller
                rays = _rays(n, lo, hi, 100 + call)
                t_ref, ids_ref = gs.intersect(rays, intersection=mode)
                rays_d.put(rays)
                gs.intersect_device(rays_d.ptr.value, n, t_d.ptr.value, ids_d.ptr.value, intersection=mode)
                assert _Dev.hip.hipDeviceSynchronize() == 0
                t = t_d.get(np.empty(n, np.float32))
                ids = ids_d.get(np.empty(n, np.uint32))
                assert np.array_equal(ids, ids_ref), (call, int((ids != ids_ref).sum()))
                assert np.array_equal(t.view(np.uint32), t_ref.view(np.uint32)), call
                assert (ids != 0xFFFFFFFF).mean() > 0.3
        finally:
            for b in (rays_d, t_d, ids_d):
                b.free()


def test_intersect_device_refuses_host_buffers(renderer):
    from denrimrendererkit_amd.scene import DenrimRendererError
    gs = renderer.upload(compile_scene(synthetic.cornell_room_scene()))
    rays = _rays(16, [-1, 0.5, -1], [1, 2, 1], 3)
    t = np.zeros(16, np.float32)
    ids = np.zeros(16, np.uint32)
    with pytest.raises(DenrimRendererError):
        gs.intersect_device(rays.ctypes.data, 16, t.ctypes.data, ids.ctypes.data)
