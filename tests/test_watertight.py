"""CPU pins of the watertight triangle test of the oracle (this build's
DR_INTERSECTION_WATERTIGHT mode; the reference only has Moller-Trumbore).

* Closed form: a ray through a triangle's interior gives the same t, u, v as
  the reference's Moller-Trumbore test to float rounding, with u, v the
  weights of v1, v2.
* Watertightness (the property the test exists for): rays from the centre of
  a closed icosphere through every vertex and every edge midpoint hit the
  surface; the Moller-Trumbore test is allowed to miss some (the count is
  printed), the watertight test misses none.
* Edge inclusion: a ray through the exact midpoint of an edge or through a
  vertex counts as a hit.
"""
import numpy as np

import oracle as O
from denrimrendererkit_amd import Material, Mesh, RenderScene, compile_scene, synthetic

WT = 1


def _rays(o, d):
    r = np.zeros((len(o), 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    return r


def _tri_scene(v):
    s = RenderScene()
    m = s.add_material(Material())
    s.add(Mesh(np.asarray(v, np.float32), np.arange(len(v), dtype=np.uint32)), m)
    return O.OracleScene(compile_scene(s))


def test_watertight_known_answer_and_moller_trumbore_agreement():
    osc = _tri_scene([(0, 0, 1), (1, 0, 1), (0, 1, 1)])
    rays = _rays([[0.25, 0.25, 0], [0.6, 0.1, 0], [0.7, 0.7, 0]], [[0, 0, 1], [0, 0, 1], [0, 0, 1]])
    t, ids, hit = osc.intersect(rays, brute_force=True, with_hit=True, intersection=WT)
    assert ids[0] == 0 and abs(t[0] - 1.0) < 1e-6 and ids[1] == 0 and ids[2] == 0xFFFFFFFF
    tm, im = osc.intersect(rays, brute_force=True, intersection=0)
    assert np.array_equal(im, ids) and np.allclose(tm[:2], t[:2], rtol=1e-6)
    # Random interior rays on a random soup: same hits, same t to rounding.
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    soup = _tri_scene(v)
    o = rng.uniform(-1.5, 1.5, (4000, 3)).astype(np.float32)
    d = rng.standard_normal((4000, 3)).astype(np.float32)
    t0, i0 = soup.intersect(_rays(o, d), brute_force=True, intersection=0)
    t1, i1 = soup.intersect(_rays(o, d), brute_force=True, intersection=WT)
    same = i0 == i1
    assert same.mean() > 0.999
    h = same & (i0 != 0xFFFFFFFF)
    assert np.allclose(t0[h], t1[h], rtol=1e-5, atol=1e-6)


def test_edge_and_vertex_rays_are_hits():
    osc = _tri_scene([(0, 0, 1), (1, 0, 1), (0, 1, 1)])
    # Exact hypotenuse midpoint, a vertex and an edge point on the x axis.
    rays = _rays([[0.5, 0.5, 0], [1, 0, 0], [0.25, 0, 0]], [[0, 0, 1]] * 3)
    _, ids = osc.intersect(rays, brute_force=True, intersection=WT)
    assert (ids == 0).all()


def test_closed_mesh_has_no_cracks():
    s = RenderScene()
    m = s.add_material(Material())
    sphere = synthetic.icosphere(3, 1.0)
    s.add(sphere, m)
    b = compile_scene(s)
    osc = O.OracleScene(b)
    v = np.asarray(sphere.vertices, np.float32)
    tri = np.asarray(sphere.indices, np.int64).reshape(-1, 3)
    edges = {tuple(sorted((int(a), int(c)))) for t in tri for a, c in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    mids = np.array([(v[a] + v[c]) * np.float32(0.5) for a, c in sorted(edges)], np.float32)
    dirs = np.concatenate([v, mids])
    # Slightly off-centre origins too: the rays then cross edges at general angles.
    origins = [np.zeros(3, np.float32), np.array([0.013, -0.021, 0.007], np.float32)]
    for o in origins:
        rays = _rays(np.broadcast_to(o, dirs.shape), dirs - o)
        _, wt = osc.intersect(rays, brute_force=True, intersection=WT)
        _, mt = osc.intersect(rays, brute_force=True, intersection=0)
        print(f"icosphere edge/vertex rays from {o}: Moller-Trumbore misses {(mt == 0xFFFFFFFF).sum()} of {len(rays)}, "
              f"watertight misses {(wt == 0xFFFFFFFF).sum()}")
        assert (wt != 0xFFFFFFFF).all()
