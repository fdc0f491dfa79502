"""CPU tests of the host-side scene compile that feeds both the HIP path and
the oracle (they consume the same arrays, so GPU parity cannot catch a
mistake here):

  * Mesh.gpuTriangles known answers (MeshTests.swift:7-68): box counts and
    bounds, material / object / primitive IDs, imported normals kept;
  * its fallbacks (Mesh.swift:149-186) against closed forms: a missing or
    zero normal becomes the face normal, a non-unit one is normalised,
    degenerate UVs give the helper-axis frame (t = normalize(helper x n),
    b = n x t, helper x when |n.y| >= 0.999), proper UVs give
    t = dP/du and b = dP/dv; normals and frames transform by the inverse
    transpose (Transform.swift:77-86);
  * GPUMaterial packing known answers (APITests.swift:2140-2233,
    2328-2342): every slot of the reference's layout.
"""
import numpy as np

from denrimrendererkit_amd import Material, Mesh, Transform


def test_mesh_known_answers():
    box = Mesh.box_sized((2, 4, 6))
    assert len(box.vertices) == 8 and len(box.indices) == 36 and len(box.gpu_triangles(0)) == 12
    b = Mesh.box((-1, -2, -3), (4, 5, 6))
    assert b.vertices[:, 0].min() == -1 and b.vertices[:, 0].max() == 4
    assert b.vertices[:, 1].min() == -2 and b.vertices[:, 1].max() == 5
    assert b.vertices[:, 2].min() == -3 and b.vertices[:, 2].max() == 6
    q = Mesh.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)).gpu_triangles(3, object_id=9)
    assert q[0]["materialID"] == 3 and q[0]["objectID"] == 9
    assert q[0]["primitiveID"] == 0 and q[1]["primitiveID"] == 1
    m = Mesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 2], normals=[(1, 0, 0), (0, 1, 0), (0, 0, 1)])
    t = m.gpu_triangles(0)[0]
    assert list(t["n0"][:3]) == [1, 0, 0] and list(t["n1"][:3]) == [0, 1, 0] and list(t["n2"][:3]) == [0, 0, 1]


def _face_normal(p0, p1, p2):
    n = np.cross(np.subtract(p1, p0, dtype=np.float64), np.subtract(p2, p0, dtype=np.float64))
    return n / np.linalg.norm(n)


def test_missing_zero_and_unnormalised_normals():
    p = [(0.0, 0.0, 0.0), (2.0, 0.0, 1.0), (0.5, 3.0, -1.0)]
    face = _face_normal(*p)
    t = Mesh(p, [0, 1, 2]).gpu_triangles(0)[0]  # no normals at all
    for k in ("n0", "n1", "n2"):
        assert np.allclose(t[k][:3], face, atol=1e-6)
    t = Mesh(p, [0, 1, 2], normals=[(0, 0, 0), (0, 0, 2.0), (3.0, 4.0, 0)]).gpu_triangles(0)[0]
    assert np.allclose(t["n0"][:3], face, atol=1e-6)  # zero length: the face normal
    assert np.allclose(t["n1"][:3], (0, 0, 1), atol=1e-7) and np.allclose(t["n2"][:3], (0.6, 0.8, 0), atol=1e-7)
    # a mesh with normals for only some vertices falls back for the rest
    t = Mesh(p + [(1, 1, 1)], [3, 1, 2], normals=[(0, 1, 0)]).gpu_triangles(0)[0]
    assert np.allclose(t["n0"][:3], _face_normal((1, 1, 1), p[1], p[2]), atol=1e-6)


def test_tangent_frames():
    # proper UVs: tangent = dP/du, bitangent = dP/dv (here the quad's x and y axes)
    q = Mesh.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)).gpu_triangles(0)
    for t in q:
        assert np.allclose(t["tangent"][:3], (1, 0, 0), atol=1e-7) and np.allclose(t["bitangent"][:3], (0, 1, 0), atol=1e-7)
    # stretched UVs scale but do not turn the frame
    m = Mesh([(0, 0, 0), (4, 0, 0), (0, 2, 0)], [0, 1, 2], texcoords=[(0, 0), (1, 0), (0, 1)]).gpu_triangles(0)[0]
    assert np.allclose(m["tangent"][:3], (1, 0, 0), atol=1e-7) and np.allclose(m["bitangent"][:3], (0, 1, 0), atol=1e-7)
    # degenerate UVs (none given): the helper-axis frame around the face normal
    p = [(0.0, 0.0, 0.0), (2.0, 0.0, 1.0), (0.5, 3.0, -1.0)]
    n = _face_normal(*p)
    t = Mesh(p, [0, 1, 2]).gpu_triangles(0)[0]
    tan = np.cross((0, 1, 0), n)
    tan /= np.linalg.norm(tan)
    assert np.allclose(t["tangent"][:3], tan, atol=1e-6) and np.allclose(t["bitangent"][:3], np.cross(n, tan), atol=1e-6)
    assert abs(np.dot(t["tangent"][:3], n)) < 1e-6 and abs(np.dot(t["bitangent"][:3], n)) < 1e-6
    # a face normal along y: the helper is the x axis
    t = Mesh([(0, 0, 0), (0, 0, 1), (1, 0, 0)], [0, 1, 2]).gpu_triangles(0)[0]
    assert np.allclose(t["n0"][:3], (0, 1, 0), atol=1e-7)
    assert np.allclose(t["tangent"][:3], np.cross((1, 0, 0), (0, 1, 0)), atol=1e-7)
    assert np.allclose(t["bitangent"][:3], np.cross((0, 1, 0), np.cross((1, 0, 0), (0, 1, 0))), atol=1e-7)


def test_transformed_normals_use_the_inverse_transpose():
    m = Mesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 2], normals=[(1, 1, 1)] * 3)
    xf = Transform.translation((5, -2, 3))
    t = m.gpu_triangles(0, xf)[0]
    assert np.allclose(t["v1"][:3], (6, -2, 3)) and np.allclose(t["n0"][:3], np.ones(3) / np.sqrt(3), atol=1e-6)
    xf = Transform.scale((2.0, 1.0, 0.5))
    t = m.gpu_triangles(0, xf)[0]
    want = np.array([1 / 2.0, 1.0, 1 / 0.5])
    want /= np.linalg.norm(want)
    assert np.allclose(t["n0"][:3], want, atol=1e-6)
    assert np.allclose(t["v1"][:3], (2, 0, 0))


def test_material_packing_known_answers():
    m = Material(base_color=(0.8, 0.7, 0.6), roughness=0.35, metallic=0.2, specular=0.6, specular_color=(0.9, 0.7, 0.5),
                 index_of_refraction=1.8, specular_anisotropy=0.65, clearcoat=0.4, clearcoat_color=(0.25, 0.8, 0.9),
                 clearcoat_attenuation_color=(0.5, 0.65, 0.8), clearcoat_thickness=0.6, clearcoat_roughness=0.08,
                 clearcoat_index_of_refraction=1.6, thin_film=0.8, thin_film_thickness_nanometers=520,
                 thin_film_index_of_refraction=1.42, sheen=0.55, sheen_color=(0.3, 0.4, 0.5), sheen_roughness=0.7,
                 subsurface=0.8, subsurface_color=(0.95, 0.42, 0.28), subsurface_radius=(1.2, 0.45, 0.25),
                 subsurface_scale=0.32, subsurface_anisotropy=0.18, transmission=0.75, transmission_color=(0.2, 0.6, 0.9),
                 transmission_roughness=0.18, transmission_index_of_refraction=1.33,
                 transmission_absorption_color=(0.5, 0.75, 0.95), transmission_absorption_distance=1.25,
                 thin_walled=True, volume_scattering=0.7, volume_scattering_color=(0.8, 0.85, 0.9),
                 volume_scattering_distance=0.6, volume_anisotropy=0.22)
    g = m.gpu_material()
    want = {
        ("parameters", 0): 0.35, ("parameters", 1): 0.2,
        ("parameters2", 0): 0.6, ("parameters2", 1): 1.8, ("parameters2", 2): 0.4, ("parameters2", 3): 0.08,
        ("specularColor", 0): 0.9, ("specularColor", 1): 0.7, ("specularColor", 2): 0.5, ("specularColor", 3): 1.6,
        ("sheenColor", 0): 0.3, ("sheenColor", 1): 0.4, ("sheenColor", 2): 0.5, ("sheenColor", 3): 0.55,
        ("transmissionColor", 0): 0.2, ("transmissionColor", 1): 0.6, ("transmissionColor", 2): 0.9,
        ("transmissionColor", 3): 1.0,
        ("parameters3", 0): 0.7, ("parameters3", 1): 0.18, ("parameters3", 2): 1.33, ("parameters3", 3): 0.65,
        ("clearcoatColor", 0): 0.25, ("clearcoatColor", 1): 0.8, ("clearcoatColor", 2): 0.9, ("clearcoatColor", 3): 0.0,
        ("clearcoatAttenuation", 0): 0.5, ("clearcoatAttenuation", 1): 0.65, ("clearcoatAttenuation", 2): 0.8,
        ("clearcoatAttenuation", 3): 0.6,
        ("transmissionAbsorption", 0): 0.5, ("transmissionAbsorption", 1): 0.75,
        ("transmissionAbsorption", 2): 0.95, ("transmissionAbsorption", 3): 1.25,
        ("thinFilm", 0): 0.8, ("thinFilm", 1): 520.0, ("thinFilm", 2): 1.42,
        ("subsurfaceColor", 0): 0.95, ("subsurfaceColor", 1): 0.42, ("subsurfaceColor", 2): 0.28,
        ("subsurfaceColor", 3): 0.8,
        ("subsurfaceRadius", 0): 1.2, ("subsurfaceRadius", 1): 0.45, ("subsurfaceRadius", 2): 0.25,
        ("subsurfaceRadius", 3): 0.32,
        ("subsurfaceParameters", 0): 0.18,
        ("volumeScattering", 0): 0.8, ("volumeScattering", 1): 0.85, ("volumeScattering", 2): 0.9,
        ("volumeScattering", 3): 0.7,
        ("volumeParameters", 0): 0.6, ("volumeParameters", 1): 0.22,
        ("emission", 3): 0.75,
    }
    for (field, k), v in want.items():
        assert abs(float(g[field][k]) - v) <= 1e-4, (field, k, float(g[field][k]), v)
    # transmission colour, roughness and IOR default to the base surface's
    t = Material(base_color=(0.35, 0.6, 0.9), roughness=0.42, index_of_refraction=1.47, transmission=1).gpu_material()
    assert np.allclose(t["transmissionColor"][:3], (0.35, 0.6, 0.9)) and t["transmissionColor"][3] == 0.0
    assert abs(t["parameters3"][1] - 0.42) < 1e-6 and abs(t["parameters3"][2] - 1.47) < 1e-6
    # clearcoat attenuation defaults to the clearcoat colour
    c = Material(base_color=(0.4, 0.5, 0.6), clearcoat=0.7, clearcoat_color=(0.72, 0.86, 1),
                 clearcoat_thickness=0.4).gpu_material()
    assert np.allclose(c["clearcoatAttenuation"], (0.72, 0.86, 1.0, 0.4), atol=1e-6)
