"""Seeded synthetic scenes for the benchmark and parity workloads.

Scene *content* uses numpy's PCG64 generator seeded with 0x5EED0000 + the
workload number, so every run builds bit-identical triangles, materials and
textures. Render randomness is the reference's per-pixel stream
(PathTrace.metal:264-276, 5212-5214) with ``frame_seed``.

* ``cornell_room_scene``   -- closed 4x3x4 room (12 tris), 1x1 emissive ceiling
  quad (radiance 15), level-4 icosphere (5,120 tris, r = 0.8, Lambertian 0.7),
  40x40-quad height-perturbed floor panel (3,200 tris, sigma 0.02, GGX metal
  roughness 0.3); pinhole camera, 45 deg fov; 320x180, 1 spp, depth 4.
* ``layered_terrain_scene`` -- 1024x512-quad fBm heightfield (6 octaves,
  1,048,576 tris) plus 256 instances of a 20,000-tri displaced sphere with
  random transforms, materials drawn uniformly from {dielectric, metal,
  clearcoat, thin-film, sheen, random-walk subsurface}, 64 emissive
  triangles and a 2048x1024 log-normal HDRI with a 1e4 Gaussian sun;
  1920x1080, 64 spp, depth 8. ``scale`` < 1 shrinks every count for tests.
* ``sdf_bricks_scene``      -- sparse SDF volume: virtual 768^3 grid of 8^3
  bricks with a 1-sample apron, narrow band 2 voxels, FP16-exact distances;
  the field is the smooth union (k = 0.05) of 200 seeded spheres and tori
  (radii U[0.02, 0.15]) in the unit cube, each a metal or Lambertian
  material, on a 2-triangle ground under the 1024x512 HDRI; 1024x1024, 16 spp.
* ``material_preview_scene`` -- a 20,000-tri sphere on a 2-tri plane under a
  1024x512 synthetic HDRI, with seeded layered-material parameters.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from .scene import (Camera, Environment, Material, Mesh, QuadLight, RenderScene, Texture2D, TextureSamplingMode,
                    Transform)

f32 = np.float32


def content_rng(workload: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(0x5EED0000 + workload))


# ------------------------------------------------------------------ meshes --
def icosphere(level: int, radius: float = 1.0) -> Mesh:
    """Subdivided icosahedron: 20 * 4**level triangles, smooth normals."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, np.float64)
    return Mesh((v * radius).astype(f32), np.array(faces, np.uint32).reshape(-1), normals=v.astype(f32))


def uv_sphere(segments: int, bands: int, radius: float = 1.0, displacement=None) -> Mesh:
    """Latitude-longitude sphere
    (poles are fans, bands - 1 rings): 2 * segments * (bands - 1) triangles;
    optional radial displacement fn(dirs) -> scale."""
    verts = [(0.0, 1.0, 0.0)]
    for i in range(1, bands):
        th = math.pi * i / bands
        for j in range(segments):
            ph = 2 * math.pi * j / segments
            verts.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
    verts.append((0.0, -1.0, 0.0))
    v = np.array(verts, np.float64)
    idx = []
    for j in range(segments):  # top fan
        idx += [0, 1 + (j + 1) % segments, 1 + j]
    for i in range(bands - 2):
        r0, r1 = 1 + i * segments, 1 + (i + 1) * segments
        for j in range(segments):
            a, b = r0 + j, r0 + (j + 1) % segments
            c, d = r1 + j, r1 + (j + 1) % segments
            idx += [a, b, c, b, d, c]
    last = len(verts) - 1
    r0 = 1 + (bands - 2) * segments
    for j in range(segments):  # bottom fan
        idx += [r0 + j, r0 + (j + 1) % segments, last]
    dirs = v.copy()
    scale = np.ones(len(v)) if displacement is None else displacement(dirs)
    pos = dirs * (radius * scale)[:, None]
    mesh = Mesh(pos.astype(f32), np.array(idx, np.uint32))
    return mesh


def value_noise_2d(rng: np.random.Generator, shape: Tuple[int, int], cells: int) -> np.ndarray:
    """Bilinear value noise over a (cells+1)^2 lattice sampled at `shape`."""
    lattice = rng.standard_normal((cells + 1, cells + 1))
    h, w = shape
    y = np.linspace(0, cells, h, endpoint=False)
    x = np.linspace(0, cells, w, endpoint=False)
    y0, x0 = np.floor(y).astype(int), np.floor(x).astype(int)
    fy, fx = (y - y0)[:, None], (x - x0)[None, :]
    fy, fx = fy * fy * (3 - 2 * fy), fx * fx * (3 - 2 * fx)
    a = lattice[y0][:, x0]
    b = lattice[y0][:, x0 + 1]
    c = lattice[y0 + 1][:, x0]
    d = lattice[y0 + 1][:, x0 + 1]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def fbm_heightfield(rng: np.random.Generator, nx: int, nz: int, octaves: int = 6) -> np.ndarray:
    """(nz+1, nx+1) heights: sum of `octaves` value-noise layers, gain 0.5."""
    out = np.zeros((nz + 1, nx + 1))
    amp, cells = 1.0, 4
    for _ in range(octaves):
        out += amp * value_noise_2d(rng, (nz + 1, nx + 1), cells)
        amp *= 0.5
        cells *= 2
    return out


def grid_mesh(heights: np.ndarray, x0: float, x1: float, z0: float, z1: float) -> Mesh:
    nz, nx = heights.shape[0] - 1, heights.shape[1] - 1
    xs = np.linspace(x0, x1, nx + 1)
    zs = np.linspace(z0, z1, nz + 1)
    X, Z = np.meshgrid(xs, zs)
    verts = np.stack([X, heights, Z], axis=-1).reshape(-1, 3)
    i = np.arange(nz)[:, None] * (nx + 1) + np.arange(nx)[None, :]
    a, b, c, d = i, i + 1, i + nx + 1, i + nx + 2
    # Winding gives +y face normals.
    idx = np.stack([a, c, b, b, c, d], axis=-1).reshape(-1)
    uv = np.stack([(X - x0) / (x1 - x0), (Z - z0) / (z1 - z0)], axis=-1).reshape(-1, 2)
    return Mesh(verts.astype(f32), idx.astype(np.uint32), texcoords=uv.astype(f32))


# ------------------------------------------------------------------- HDRI --
def lognormal_hdri(rng: np.random.Generator, width: int, height: int, sun_peak: float = 1e4,
                   sun_dir=(0.3, 0.6, -0.74), sun_sigma_deg: float = 1.5) -> Texture2D:
    """Per-texel log-normal radiance (mu 0, sigma 1) plus a Gaussian sun."""
    base = np.exp(rng.standard_normal((height, width)))
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * math.pi
    phi = (u - 0.5) * 2 * math.pi
    # Texel direction as the reference maps it (PathTrace.metal:3952-3993).
    dx = np.cos(phi)[None, :] * np.sin(theta)[:, None]
    dy = np.cos(theta)[:, None] * np.ones((1, width))
    dz = np.sin(phi)[None, :] * np.sin(theta)[:, None]
    s = np.array(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    cosang = np.clip(dx * s[0] + dy * s[1] + dz * s[2], -1, 1)
    ang = np.degrees(np.arccos(cosang))
    sun = sun_peak * np.exp(-0.5 * (ang / sun_sigma_deg) ** 2)
    tint = np.array([0.9, 0.95, 1.0])
    rgb = base[..., None] * tint[None, None, :] + sun[..., None]
    px = np.concatenate([rgb, np.ones((height, width, 1))], axis=-1).astype(f32).reshape(-1, 4)
    return Texture2D(width, height, px, TextureSamplingMode.LINEAR)


# ------------------------------------------------------------- workloads --
def _facing(poly, toward):
    """Orders a planar polygon so its winding normal cross(p1-p0, p2-p0)
    points toward `toward`."""
    p = np.array(poly, np.float64)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    if np.dot(n, np.array(toward, np.float64) - p.mean(axis=0)) < 0:
        p = p[::-1]
    return [tuple(float(x) for x in v) for v in p]


def cornell_room_scene() -> RenderScene:
    """320x180 room workload (see module docstring)."""
    rng = content_rng(1)
    cam = Camera(origin=(0.0, 1.55, 1.85), target=(0.1, 0.9, -1.0), vertical_field_of_view_degrees=45)
    s = RenderScene(camera=cam, environment=Environment())
    wall = s.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0, roughness=0.5))
    # Inward-facing room quads, [-2,2] x [0,3] x [-2,2].
    quads = [
        ((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2)),      # floor, +y
        ((-2, 3, -2), (-2, 3, 2), (2, 3, 2), (2, 3, -2)),      # ceiling, -y
        ((-2, 0, -2), (-2, 3, -2), (2, 3, -2), (2, 0, -2)),    # back, +z
        ((-2, 0, 2), (2, 0, 2), (2, 3, 2), (-2, 3, 2)),        # front, -z
        ((-2, 0, -2), (-2, 0, 2), (-2, 3, 2), (-2, 3, -2)),    # left, +x
        ((2, 0, -2), (2, 3, -2), (2, 3, 2), (2, 0, 2)),        # right, -x
    ]
    for q in quads:
        s.add(Mesh.quad(*_facing(q, (0.0, 1.5, 0.0))), wall)
    s.add_quad_light(QuadLight(*_facing(((-0.5, 2.999, -0.5), (0.5, 2.999, -0.5), (0.5, 2.999, 0.5),
                                         (-0.5, 2.999, 0.5)), (0.0, 0.0, 0.0)), color=(1, 1, 1), intensity=15.0))
    ball = s.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0, roughness=0.5))
    s.add(icosphere(4, 0.8), ball, Transform.translation((-0.75, 0.8, -0.6)))
    metal = s.add_material(Material(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.3))
    heights = 0.02 + rng.normal(0.0, 0.02, size=(41, 41))
    s.add(grid_mesh(heights, -0.1, 1.9, -1.9, 0.1), metal)
    return s


def room_variant(seed: int) -> RenderScene:
    """Config-1-style scene family for the learned denoiser: the closed 4x3x4
    room with a level-4 icosphere and a height-perturbed floor panel, with
    seeded wall colours, sphere material (Lambertian, metal or one of the
    six layered kinds), sphere size and place, ceiling-light size and
    strength, floor-panel roughness and camera pose inside the room."""
    rng = np.random.Generator(np.random.PCG64(0x5EED0005 * 7919 + seed))
    u = lambda a, b: float(rng.uniform(a, b))  # noqa: E731
    origin = (u(-1.4, 1.4), u(0.6, 2.4), u(1.2, 1.9))
    target = (u(-1.5, 1.5), u(0.3, 1.8), u(-1.9, -0.6))
    cam = Camera(origin=origin, target=target, vertical_field_of_view_degrees=u(35, 65))
    s = RenderScene(camera=cam, environment=Environment())
    grey = u(0.45, 0.8)
    walls = [s.add_material(Material(base_color=(grey, grey, grey), specular=0.0, roughness=0.5))]
    for _ in range(2):
        walls.append(s.add_material(Material(base_color=tuple(float(x) for x in rng.uniform(0.1, 0.85, 3)),
                                             specular=0.0, roughness=0.5)))
    quads = [
        ((-2, 0, -2), (2, 0, -2), (2, 0, 2), (-2, 0, 2)),
        ((-2, 3, -2), (-2, 3, 2), (2, 3, 2), (2, 3, -2)),
        ((-2, 0, -2), (-2, 3, -2), (2, 3, -2), (2, 0, -2)),
        ((-2, 0, 2), (2, 0, 2), (2, 3, 2), (-2, 3, 2)),
        ((-2, 0, -2), (-2, 0, 2), (-2, 3, 2), (-2, 3, -2)),
        ((2, 0, -2), (2, 3, -2), (2, 3, 2), (2, 0, 2)),
    ]
    for k, q in enumerate(quads):
        s.add(Mesh.quad(*_facing(q, (0.0, 1.5, 0.0))), walls[1] if k == 4 else (walls[2] if k == 5 else walls[0]))
    h, lx, lz = u(0.25, 0.7), u(-1.0, 1.0), u(-1.0, 1.0)
    s.add_quad_light(QuadLight(*_facing(((lx - h, 2.999, lz - h), (lx + h, 2.999, lz - h), (lx + h, 2.999, lz + h),
                                         (lx - h, 2.999, lz + h)), (lx, 0.0, lz)), color=(1, 1, 1),
                               intensity=u(8.0, 30.0)))
    kind = int(rng.integers(0, 8))
    base = tuple(float(x) for x in rng.uniform(0.15, 0.9, 3))
    if kind == 6:
        ball_mat = Material(base_color=base, specular=0.0, roughness=0.5)
    elif kind == 7:
        ball_mat = Material(base_color=base, metallic=1.0, roughness=u(0.05, 0.6))
    else:
        ball_mat = _layered_material(rng, kind)
    r = u(0.4, 0.8)
    s.add(icosphere(4, r), s.add_material(ball_mat), Transform.translation((u(-1.1, 1.1), r, u(-1.4, 0.0))))
    metal = s.add_material(Material(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=u(0.1, 0.6)))
    heights = 0.02 + rng.normal(0.0, 0.02, size=(41, 41))
    x0, z0 = u(-1.9, -0.1), u(-1.9, -0.1)
    s.add(grid_mesh(heights, x0, x0 + 2.0, z0, z0 + 2.0), metal)
    return s


def _layered_material(rng: np.random.Generator, kind: int) -> Material:
    base = tuple(float(x) for x in rng.uniform(0.2, 0.9, 3))
    rough = float(rng.uniform(0.05, 0.7))
    if kind == 0:  # rough dielectric
        return Material(base_color=base, transmission=1.0, transmission_index_of_refraction=float(rng.uniform(1.3, 1.7)),
                        transmission_roughness=float(rng.uniform(0.02, 0.8)), roughness=rough,
                        transmission_color=tuple(float(x) for x in rng.uniform(0.6, 1.0, 3)))
    if kind == 1:  # metal
        return Material(base_color=base, metallic=1.0, roughness=rough)
    if kind == 2:  # clearcoat
        return Material(base_color=base, roughness=rough, clearcoat=1.0,
                        clearcoat_roughness=float(rng.uniform(0.02, 0.3)), clearcoat_thickness=float(rng.uniform(0, 0.5)),
                        clearcoat_attenuation_color=tuple(float(x) for x in rng.uniform(0.6, 1.0, 3)))
    if kind == 3:  # thin film
        return Material(base_color=base, metallic=float(rng.uniform(0, 1)), roughness=rough, thin_film=1.0,
                        thin_film_thickness_nanometers=float(rng.uniform(200, 800)))
    if kind == 4:  # sheen
        return Material(base_color=base, roughness=rough, sheen=1.0,
                        sheen_color=tuple(float(x) for x in rng.uniform(0.5, 1.0, 3)),
                        sheen_roughness=float(rng.uniform(0.2, 0.9)))
    # random-walk subsurface
    mfp = float(rng.uniform(0.01, 0.3))
    return Material(base_color=base, roughness=rough, subsurface=1.0, subsurface_radius=(mfp, mfp * 0.6, mfp * 0.35),
                    subsurface_color=base)


def layered_terrain_scene(scale: float = 1.0, hdri: Tuple[int, int] = (2048, 1024)) -> RenderScene:
    """The 1080p layered-material workload (see module docstring).
    `scale` shrinks grid, instance count and sphere resolution for tests."""
    rng = content_rng(2)
    nx, nz = max(8, int(1024 * scale)), max(4, int(512 * scale))
    n_inst = max(4, int(256 * scale))
    seg = max(8, int(100 * math.sqrt(scale)))
    width_x, width_z = 128.0, 64.0
    env = Environment(texture=lognormal_hdri(rng, *hdri), intensity=1.0, max_radiance=0.0)
    cam = Camera(origin=(0.0, 14.0, 36.0), target=(0.0, 2.0, 0.0), vertical_field_of_view_degrees=45)
    s = RenderScene(camera=cam, environment=env)
    h = fbm_heightfield(rng, nx, nz) * 3.0
    terrain = s.add_material(Material(base_color=(0.45, 0.42, 0.38), roughness=0.8))
    s.add(grid_mesh(h, -width_x / 2, width_x / 2, -width_z / 2, width_z / 2), terrain)

    lat = rng.standard_normal(64)

    def displace(d):
        k = np.floor((np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi) + 0.5) * 8).astype(int) % 8
        j = np.clip(np.floor(np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi * 8).astype(int), 0, 7)
        return 1.0 + 0.08 * lat[j * 8 + k] + 0.03 * np.sin(7 * d[:, 0]) * np.cos(5 * d[:, 2])

    sphere = uv_sphere(seg, seg + 1, 1.0, displace)  # 2*seg*seg triangles

    def terrain_height(x, z):
        ix = int(np.clip((x + width_x / 2) / width_x * nx, 0, nx))
        iz = int(np.clip((z + width_z / 2) / width_z * nz, 0, nz))
        return float(h[iz, ix])

    for _ in range(n_inst):
        kind = int(rng.integers(0, 6))
        mat = s.add_material(_layered_material(rng, kind))
        sc = float(rng.uniform(0.8, 2.5))
        x = float(rng.uniform(-0.45, 0.45) * width_x)
        z = float(rng.uniform(-0.45, 0.45) * width_z)
        y = terrain_height(x, z) + sc * 0.9
        t = Transform.translation((x, y, z)) * Transform.rotation_y(float(rng.uniform(0, 2 * math.pi))) * \
            Transform.rotation_x(float(rng.uniform(-0.5, 0.5))) * Transform.scale((sc, sc, sc))
        s.add(sphere, mat, t)

    # 64 emissive triangles floating above the terrain.
    lights = s.add_material(Material(base_color=(1, 1, 1), emission=(1.0, 0.9, 0.75), emission_strength=40.0,
                                     specular=0.0))
    tris = []
    for _ in range(64):
        cx, cz = rng.uniform(-0.45, 0.45) * width_x, rng.uniform(-0.45, 0.45) * width_z
        cy = terrain_height(cx, cz) + float(rng.uniform(4, 8))
        r = float(rng.uniform(0.4, 1.0))
        a = float(rng.uniform(0, 2 * math.pi))
        p = [(cx + r * math.cos(a + k * 2.0944), cy, cz + r * math.sin(a + k * 2.0944)) for k in range(3)]
        tris.extend(_facing(p, (cx, cy - 1.0, cz)))  # emits downwards
    s.add(Mesh(np.array(tris, f32), np.arange(len(tris), dtype=np.uint32)), lights)
    return s


def layered_material_gallery(hdri: Tuple[int, int] = (64, 32), segments: int = 16) -> RenderScene:
    """Parity scene for the layered BSDF: two spheres of each of the six
    material kinds of the terrain workload (rough dielectric, metal,
    clearcoat, thin film, sheen, random-walk subsurface) on a ground plane,
    lit by 8 emissive triangles and a log-normal HDRI with a sun."""
    rng = content_rng(12)
    env = Environment(texture=lognormal_hdri(rng, *hdri, sun_peak=500.0), intensity=1.0, max_radiance=0.0)
    cam = Camera(origin=(0.0, 2.2, 7.5), target=(0.0, 0.6, 0.0), vertical_field_of_view_degrees=45)
    s = RenderScene(camera=cam, environment=env)
    ground = s.add_material(Material(base_color=(0.45, 0.42, 0.38), roughness=0.8))
    s.add(Mesh.quad((-8, 0, -8), (-8, 0, 8), (8, 0, 8), (8, 0, -8)), ground)
    sphere = uv_sphere(segments, segments + 1, 0.45)
    for kind in range(6):
        for row in range(2):
            mat = s.add_material(_layered_material(rng, kind))
            x, z = -3.0 + 1.2 * kind, -0.7 + 1.4 * row
            s.add(sphere, mat, Transform.translation((x, 0.45, z)))
    lights = s.add_material(Material(base_color=(1, 1, 1), emission=(1.0, 0.9, 0.75), emission_strength=25.0,
                                     specular=0.0))
    tris = []
    for k in range(8):
        cx, cz, cy, r = -3.5 + k, float(rng.uniform(-1.5, 1.5)), float(rng.uniform(2.0, 3.0)), 0.4
        p = [(cx + r * math.cos(k + j * 2.0944), cy, cz + r * math.sin(k + j * 2.0944)) for j in range(3)]
        tris.extend(_facing(p, (cx, cy - 1.0, cz)))
    s.add(Mesh(np.array(tris, f32), np.arange(len(tris), dtype=np.uint32)), lights)
    return s


def sdf_primitives(rng: np.random.Generator, count: int = 200, metal: int = 1, matte: int = 2):
    from .sdf import SDFPrimitive
    prims = []
    for _ in range(count):
        r = float(rng.uniform(0.02, 0.15))
        c = tuple(float(x) for x in rng.uniform(0.15, 0.85, 3))
        mat = metal if rng.uniform() < 0.5 else matte
        if rng.uniform() < 0.5:
            prims.append(SDFPrimitive("sphere", mat, c, r))
        else:
            axis = rng.standard_normal(3)
            axis /= np.linalg.norm(axis)
            prims.append(SDFPrimitive("torus", mat, c, r, tuple(float(x) for x in axis),
                                      float(r * rng.uniform(0.2, 0.45))))
    return prims


def sdf_bricks_scene(resolution: int = 768, primitives: int = 200, hdri: Tuple[int, int] = (1024, 512)) -> RenderScene:
    """The sparse-SDF workload (see module docstring); `resolution` shrinks
    the virtual grid for tests (the narrow band stays 2 voxels)."""
    from .sdf import build_sparse_volume
    rng = content_rng(3)
    env = Environment(texture=lognormal_hdri(content_rng(2), *hdri), max_radiance=0.0)
    cam = Camera(origin=(0.5, 0.95, 2.2), target=(0.5, 0.45, 0.5), vertical_field_of_view_degrees=45)
    s = RenderScene(camera=cam, environment=env)
    ground = s.add_material(Material(base_color=(0.5, 0.5, 0.5), roughness=0.8))
    metal = s.add_material(Material(base_color=(0.9, 0.75, 0.55), metallic=1.0, roughness=0.3))
    matte = s.add_material(Material(base_color=(0.7, 0.7, 0.7), specular=0.0, roughness=0.5))
    s.add(Mesh.quad(*_facing(((-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)), (0, 1, 0))), ground)
    prims = sdf_primitives(rng, primitives, metal, matte)
    vol = build_sparse_volume(prims, dims=(resolution,) * 3, brick_size=8, apron=1,
                              narrow_band=2.0 / (resolution - 1), smooth_radius=0.05)
    s.add_sparse_volume(vol, metal)
    return s


def material_preview_scene(seed: int, hdri: Tuple[int, int] = (1024, 512), segments: int = 100) -> RenderScene:
    """Sphere-on-plane material preview with seeded layered parameters."""
    rng = np.random.Generator(np.random.PCG64(0x5EED0004 * 7919 + seed))
    env_rng = content_rng(4)
    env = Environment(texture=lognormal_hdri(env_rng, *hdri, sun_peak=200.0), max_radiance=0.0)
    cam = Camera(origin=(0.0, 1.2, 3.2), target=(0.0, 0.8, 0.0), vertical_field_of_view_degrees=38)
    s = RenderScene(camera=cam, environment=env)
    plane = s.add_material(Material(base_color=(0.5, 0.5, 0.5), roughness=0.7))
    s.add(Mesh.quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4)), plane)
    mat = s.add_material(_layered_material(rng, int(rng.integers(0, 6))))
    s.add(uv_sphere(segments, segments + 1, 0.8), mat, Transform.translation((0.0, 0.8, 0.0)))
    return s
