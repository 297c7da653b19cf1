"""Seeded synthetic scenes, built through the Ray scene API (capi.Scene), so the
same construction runs on any backend that exports the rhx_* ABI.

Workloads (seed = 0x5EED0000 + k for workload k):
  cornell_box      — Cornell box: 5 wall quads + 2 boxes rotated about Y by seeded
                     angles (34 triangles) + a 2-triangle emissive ceiling quad,
                     Lambertian albedos U[0.2, 0.8]; 256x256, 16 spp, depth 4.
  triangle_soup    — 512 jittered icospheres (radii log-uniform 0.05..2, centres
                     U[-20, 20]^3) over a ground plane, ~200k triangles, two sphere
                     lights plus a constant environment, 50/50 diffuse / glossy
                     (roughness U[0.05, 1]); 1920x1080, 64 spp, depth 8.
  mesh_lights      — the reference's own mesh-lights test scene (test_scene1.h).
  batch_workload   — 64 icosphere soups (seed 0x5EED0003): triangle counts
                     log-uniform 1k..500k, frames drawn from {64, 128, 256, 512,
                     1024}^2, 4..32 spp, depth 4..8, each rendered alone or all
                     together by the batch renderer.
  cornell_refined  — the Cornell box with every wall and box face split into a
                     k x k grid (k = 17: 9,826 triangles + the 2-triangle light),
                     one diffuse material per triangle with albedo U[0.1, 0.9]
                     (seed 0x5EED0004; the hidden target uses 0x5EED0005), one
                     emissive quad; 512x512, 32 spp, depth 6, Russian roulette off
                     (the estimator is then smooth in the albedos).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import capi as A

SEED_BASE = 0x5EED0000


@dataclass
class SceneSpec:
    name: str
    width: int
    height: int
    spp: int
    depth: int
    triangles: int = 0
    info: dict = field(default_factory=dict)


def _white(scene):
    return scene.add_texture(np.array([[[255, 255, 255, 255]]], np.uint8), is_srgb=False, generate_mipmaps=True)


def _quad(p0, p1, p2, p3, n):
    """Two triangles with PxyzNxyzTuv attributes, wound so that the geometric
    normal (e0 x e1, the front face of PreprocessTri) agrees with n."""
    uv = [(0, 0), (1, 0), (1, 1), (0, 1)]
    attrs = [list(p) + list(n) + list(t) for p, t in zip((p0, p1, p2, p3), uv)]
    geo = np.cross(np.subtract(p1, p0), np.subtract(p2, p0))
    idx = [0, 1, 2, 0, 2, 3] if np.dot(geo, n) > 0 else [0, 2, 1, 0, 3, 2]
    return np.array(attrs, np.float32), np.array(idx, np.uint32)


def _box(center, size, angle):
    """Axis box rotated about Y: 12 triangles, outward normals."""
    cx, cy, cz = center
    sx, sy, sz = size
    c, s = math.cos(angle), math.sin(angle)

    def tr(x, y, z):
        return (cx + c * x + s * z, cy + y, cz - s * x + c * z)

    def rn(x, y, z):
        return (c * x + s * z, y, -s * x + c * z)

    hx, hy, hz = sx / 2, sy / 2, sz / 2
    faces = [
        ((-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz), (0, 0, 1)),
        ((hx, -hy, -hz), (-hx, -hy, -hz), (-hx, hy, -hz), (hx, hy, -hz), (0, 0, -1)),
        ((hx, -hy, hz), (hx, -hy, -hz), (hx, hy, -hz), (hx, hy, hz), (1, 0, 0)),
        ((-hx, -hy, -hz), (-hx, -hy, hz), (-hx, hy, hz), (-hx, hy, -hz), (-1, 0, 0)),
        ((-hx, hy, hz), (hx, hy, hz), (hx, hy, -hz), (-hx, hy, -hz), (0, 1, 0)),
        ((-hx, -hy, -hz), (hx, -hy, -hz), (hx, -hy, hz), (-hx, -hy, hz), (0, -1, 0)),
    ]
    attrs, idx = [], []
    for f in faces:
        a, i = _quad(tr(*f[0]), tr(*f[1]), tr(*f[2]), tr(*f[3]), rn(*f[4]))
        idx.append(i + len(attrs) * 4)
        attrs.append(a)
    return np.concatenate(attrs), np.concatenate(idx)


def cornell_geometry(seed: int = SEED_BASE + 1):
    """(attrs, indices, shape list of (material slot, vtx_start, vtx_count), albedos)."""
    rng = np.random.default_rng(seed)
    L = 5.0
    parts = []
    # walls: floor, ceiling, back, left, right (inward normals)
    parts.append(_quad((-L, 0, -L), (L, 0, -L), (L, 0, L), (-L, 0, L), (0, 1, 0)))
    parts.append(_quad((-L, 2 * L, L), (L, 2 * L, L), (L, 2 * L, -L), (-L, 2 * L, -L), (0, -1, 0)))
    parts.append(_quad((-L, 0, L), (L, 0, L), (L, 2 * L, L), (-L, 2 * L, L), (0, 0, -1)))
    parts.append(_quad((-L, 0, -L), (-L, 0, L), (-L, 2 * L, L), (-L, 2 * L, -L), (1, 0, 0)))
    parts.append(_quad((L, 0, L), (L, 0, -L), (L, 2 * L, -L), (L, 2 * L, L), (-1, 0, 0)))
    a1, a2 = rng.uniform(0.0, math.pi / 2, size=2)
    parts.append(_box((-1.8, 1.5, 1.0), (2.6, 3.0, 2.6), a1))
    parts.append(_box((1.8, 3.0, -0.8), (2.6, 6.0, 2.6), a2))
    # emissive ceiling quad, slightly below the ceiling
    parts.append(_quad((-1.2, 2 * L - 0.01, 1.2), (1.2, 2 * L - 0.01, 1.2), (1.2, 2 * L - 0.01, -1.2),
                       (-1.2, 2 * L - 0.01, -1.2), (0, -1, 0)))
    attrs, idx, shapes = [], [], []
    nv = 0
    ni = 0
    for k, (a, i) in enumerate(parts):
        attrs.append(a)
        idx.append(i + nv)
        shapes.append((k, ni, len(i)))
        nv += len(a)
        ni += len(i)
    albedos = rng.uniform(0.2, 0.8, size=(7, 3)).astype(np.float32)
    return np.concatenate(attrs).astype(np.float32), np.concatenate(idx).astype(np.uint32), shapes, albedos


def cornell_box(scene: "A.Scene", seed: int = SEED_BASE + 1, width=256, height=256, spp=16, depth=4) -> SceneSpec:
    attrs, idx, shapes, albedos = cornell_geometry(seed)
    tex = _white(scene)
    mats = [scene.add_material(A.DiffuseMaterial, tex, main_color=tuple(float(v) for v in albedos[k]))
            for k in range(7)]
    mats.append(scene.add_material(A.EmissiveMaterial, tex, main_color=(1.0, 1.0, 1.0), strength=10.0))
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(mats[k], s, c) for k, s, c in shapes])
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.0, 0.0, 0.0))
    cam = scene.add_camera(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth)
    scene.set_current_cam(cam)
    return SceneSpec("cornell_box", width, height, spp, depth, triangles=len(idx) // 3)


def _quad_grid(p0, p1, p2, p3, n, k):
    """A planar quad split into k x k cells, two triangles each, wound like _quad."""
    p0, p1, p3 = (np.asarray(p, np.float64) for p in (p0, p1, p3))
    ex, ey = p1 - p0, p3 - p0
    geo = np.cross(np.subtract(p1, p0), np.subtract(p2, p0))
    flip = np.dot(geo, n) <= 0
    verts = []
    for j in range(k + 1):
        for i in range(k + 1):
            p = p0 + ex * (i / k) + ey * (j / k)
            verts.append(list(p) + list(n) + [i / k, j / k])
    idx = []
    for j in range(k):
        for i in range(k):
            a, b, c, d = j * (k + 1) + i, j * (k + 1) + i + 1, (j + 1) * (k + 1) + i + 1, (j + 1) * (k + 1) + i
            idx += [a, c, b, a, d, c] if flip else [a, b, c, a, c, d]
    return np.array(verts, np.float32), np.array(idx, np.uint32)


def cornell_refined_geometry(k: int = 17, seed: int = SEED_BASE + 1):
    """Cornell box faces as k x k grids: (attrs, indices, n_emissive_tris) with the
    two emissive triangles last."""
    rng = np.random.default_rng(seed)
    L = 5.0
    quads = [
        ((-L, 0, -L), (L, 0, -L), (L, 0, L), (-L, 0, L), (0, 1, 0)),
        ((-L, 2 * L, L), (L, 2 * L, L), (L, 2 * L, -L), (-L, 2 * L, -L), (0, -1, 0)),
        ((-L, 0, L), (L, 0, L), (L, 2 * L, L), (-L, 2 * L, L), (0, 0, -1)),
        ((-L, 0, -L), (-L, 0, L), (-L, 2 * L, L), (-L, 2 * L, -L), (1, 0, 0)),
        ((L, 0, L), (L, 0, -L), (L, 2 * L, -L), (L, 2 * L, L), (-1, 0, 0)),
    ]
    a1, a2 = rng.uniform(0.0, math.pi / 2, size=2)
    for center, size, angle in (((-1.8, 1.5, 1.0), (2.6, 3.0, 2.6), a1), ((1.8, 3.0, -0.8), (2.6, 6.0, 2.6), a2)):
        cx, cy, cz = center
        sx, sy, sz = size
        c, s = math.cos(angle), math.sin(angle)
        tr = lambda x, y, z: (cx + c * x + s * z, cy + y, cz - s * x + c * z)
        rn = lambda x, y, z: (c * x + s * z, y, -s * x + c * z)
        hx, hy, hz = sx / 2, sy / 2, sz / 2
        for f in (((-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz), (0, 0, 1)),
                  ((hx, -hy, -hz), (-hx, -hy, -hz), (-hx, hy, -hz), (hx, hy, -hz), (0, 0, -1)),
                  ((hx, -hy, hz), (hx, -hy, -hz), (hx, hy, -hz), (hx, hy, hz), (1, 0, 0)),
                  ((-hx, -hy, -hz), (-hx, -hy, hz), (-hx, hy, hz), (-hx, hy, -hz), (-1, 0, 0)),
                  ((-hx, hy, hz), (hx, hy, hz), (hx, hy, -hz), (-hx, hy, -hz), (0, 1, 0)),
                  ((-hx, -hy, -hz), (hx, -hy, -hz), (hx, -hy, hz), (-hx, -hy, hz), (0, -1, 0))):
            quads.append((tr(*f[0]), tr(*f[1]), tr(*f[2]), tr(*f[3]), rn(*f[4])))
    attrs, idx, nv = [], [], 0
    for q in quads:
        a, i = _quad_grid(*q, k)
        attrs.append(a)
        idx.append(i + nv)
        nv += len(a)
    a, i = _quad((-1.2, 2 * L - 0.01, 1.2), (1.2, 2 * L - 0.01, 1.2), (1.2, 2 * L - 0.01, -1.2),
                 (-1.2, 2 * L - 0.01, -1.2), (0, -1, 0))
    attrs.append(a)
    idx.append(i + nv)
    return np.concatenate(attrs).astype(np.float32), np.concatenate(idx).astype(np.uint32), 2


def refined_albedos(n_tris: int, seed: int = SEED_BASE + 4) -> np.ndarray:
    """Per-triangle RGB albedo U[0.1, 0.9] (float32, n x 3)."""
    return np.random.default_rng(seed).uniform(0.1, 0.9, size=(n_tris, 3)).astype(np.float32)


def cornell_refined(scene: "A.Scene", k: int = 17, albedo_seed: int = SEED_BASE + 4, width=512, height=512, spp=32,
                    depth=6, albedos=None) -> SceneSpec:
    """Materials 0..n-1 are the diffuse triangles in index order, material n the
    emissive light (strength 10, main_color (1, 1, 1))."""
    attrs, idx, n_light = cornell_refined_geometry(k)
    n_tris = len(idx) // 3 - n_light
    alb = refined_albedos(n_tris, albedo_seed) if albedos is None else np.asarray(albedos, np.float32)
    tex = _white(scene)
    mats = [scene.add_material(A.DiffuseMaterial, tex, main_color=(float(a[0]), float(a[1]), float(a[2])))
            for a in alb]
    light = scene.add_material(A.EmissiveMaterial, tex, main_color=(1.0, 1.0, 1.0), strength=10.0)
    shapes = [(mats[t], 3 * t, 3) for t in range(n_tris)] + [(light, 3 * n_tris, 3 * n_light)]
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.0, 0.0, 0.0))
    cam = scene.add_camera(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth, termination_start_depth=16)
    scene.set_current_cam(cam)
    return SceneSpec("cornell_refined", width, height, spp, depth, triangles=len(idx) // 3,
                     info={"materials": n_tris + 1, "diffuse": n_tris, "light_material": n_tris})


def icosphere(subdiv: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = np.array(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    verts = [tuple(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = (np.array(verts[a]) + np.array(verts[b])) / 2
                p /= np.linalg.norm(p)
                cache[key] = len(verts)
                verts.append(tuple(p))
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def soup_geometry(seed: int = SEED_BASE + 2, n_spheres: int = 512, n_fine: int = 38):
    """Icosphere soup: subdivision 2 (320 tris), n_fine of them subdivision 3 (1280 tris)."""
    rng = np.random.default_rng(seed)
    base = {2: icosphere(2), 3: icosphere(3)}
    fine = set(rng.choice(n_spheres, size=n_fine, replace=False).tolist())
    radii = np.exp(rng.uniform(math.log(0.05), math.log(2.0), size=n_spheres))
    centres = rng.uniform(-20.0, 20.0, size=(n_spheres, 3))
    attrs, idx, spans = [], [], []
    nv = ni = 0
    for s in range(n_spheres):
        v, f = base[3 if s in fine else 2]
        jit = 1.0 + 0.05 * rng.standard_normal(len(v))
        p = centres[s] + v * (radii[s] * jit)[:, None]
        n = v
        uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi), 0.5 - np.arcsin(np.clip(v[:, 1], -1, 1)) / math.pi], 1)
        attrs.append(np.concatenate([p, n, uv], 1))
        idx.append(f.reshape(-1) + nv)
        spans.append((ni, f.size))
        nv += len(v)
        ni += f.size
    # ground plane
    g = 200.0
    ga, gi = _quad((-g, -21.0, -g), (-g, -21.0, g), (g, -21.0, g), (g, -21.0, -g), (0, 1, 0))
    attrs.append(ga.astype(np.float64))
    idx.append(gi + nv)
    spans.append((ni, 6))
    mat_kind = rng.integers(0, 2, size=n_spheres)
    albedo = rng.uniform(0.2, 0.9, size=(n_spheres, 3))
    rough = rng.uniform(0.05, 1.0, size=n_spheres)
    return (np.concatenate(attrs).astype(np.float32), np.concatenate(idx).astype(np.uint32), spans, mat_kind, albedo,
            rough)


def triangle_soup(scene: "A.Scene", seed: int = SEED_BASE + 2, width=1920, height=1080, spp=64, depth=8,
                  n_spheres=512, n_fine=38) -> SceneSpec:
    attrs, idx, spans, kind, albedo, rough = soup_geometry(seed, n_spheres, n_fine)
    tex = _white(scene)
    mats = []
    for s in range(len(kind)):
        col = tuple(float(c) for c in albedo[s])
        if kind[s] == 0:
            mats.append(scene.add_material(A.DiffuseMaterial, tex, main_color=col, roughness=0.0))
        else:
            mats.append(scene.add_material(A.GlossyMaterial, tex, main_color=col, roughness=float(rough[s])))
    ground = scene.add_material(A.DiffuseMaterial, tex, main_color=(0.6, 0.6, 0.6))
    shapes = [(mats[s], spans[s][0], spans[s][1]) for s in range(len(kind))] + [(ground, spans[-1][0], spans[-1][1])]
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.5, 0.5, 0.5))
    scene.add_light(A.PointLight, position=(-10.0, 25.0, -10.0), radius=0.5, color=(50.0, 50.0, 50.0))
    scene.add_light(A.PointLight, position=(12.0, 22.0, 8.0), radius=0.5, color=(50.0, 50.0, 50.0))
    cam = scene.add_camera(origin=(0.0, 8.0, -48.0), fwd=(0.0, -0.15, 1.0), fov=50.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth)
    scene.set_current_cam(cam)
    return SceneSpec("triangle_soup", width, height, spp, depth, triangles=len(idx) // 3)


def mesh_lights(scene: "A.Scene", data) -> SceneSpec:
    """The reference's mesh-lights test scene (tests/test_mesh_lights.cpp:51-170)."""
    tex = scene.add_texture(np.array([[[255, 255, 255, 255]]], np.uint8), generate_mipmaps=True)
    m1 = scene.add_material(A.EmissiveMaterial, tex, main_color=(1, 0, 0), strength=10.0)
    m2 = scene.add_material(A.EmissiveMaterial, tex, main_color=(0, 1, 0), strength=10.0)
    m3 = scene.add_material(A.EmissiveMaterial, tex, main_color=(0, 0, 1), strength=10.0)
    m4 = scene.add_material(A.DiffuseMaterial, tex, main_color=(1, 1, 1))
    g = [int(x) for x in data["groups"]]
    cam = scene.add_camera(origin=(2.0, 2.0, 0.0), fwd=(-1.0, 0.0, 0.0), fov=45.0, filter=A.Box, dtype=A.DeviceNone)
    scene.set_current_cam(cam)
    scene.set_environment((0.0, 0.0, 0.0))
    mesh = scene.add_mesh(A.PxyzNxyzTuv, data["attrs"], data["indices"],
                          [(m1, g[0], g[1]), (m2, g[2], g[3]), (m3, g[4], g[5]), (m4, g[6], g[7])])
    scene.add_mesh_instance(mesh)
    return SceneSpec("mesh_lights", 64, 64, 2048, 4)


@dataclass
class BatchItem:
    seed: int
    triangles: int   # target
    n_spheres: int
    n_fine: int
    width: int
    height: int
    spp: int
    depth: int


def batch_workload(n: int = 64, seed: int = SEED_BASE + 3) -> list:
    """Scene list of the many-scene workload (see the module docstring)."""
    rng = np.random.default_rng(seed)
    items = []
    for k in range(n):
        tris = int(round(math.exp(rng.uniform(math.log(1e3), math.log(5e5)))))
        n_spheres = max(2, int(round((tris - 2) / 391.0)))
        n_fine = int(round(n_spheres * 38 / 512))
        res = int(rng.choice([64, 128, 256, 512, 1024]))
        spp = int(rng.integers(4, 33))
        depth = int(rng.integers(4, 9))
        items.append(BatchItem(seed + 1000 + k, tris, n_spheres, n_fine, res, res, spp, depth))
    return items


def batch_scene(scene: "A.Scene", item: BatchItem) -> SceneSpec:
    return triangle_soup(scene, seed=item.seed, width=item.width, height=item.height, spp=item.spp, depth=item.depth,
                         n_spheres=item.n_spheres, n_fine=item.n_fine)


def textured_cornell(scene: "A.Scene", k: int = 8, tex_size: int = 64, seed: int = SEED_BASE + 6, width=128,
                     height=128, spp=8, depth=4, mipmaps: bool = True, textures=None) -> SceneSpec:
    """Refined Cornell box whose floor, back, left and right walls carry their
    own sRGB texture (tex_size^2, seeded noise around a base colour, or the
    given (4, n, n, 3) uint8 textures); ceiling and boxes diffuse, one
    emissive quad, Russian roulette off.  info["textures"] are the texture ids."""
    attrs, idx, n_light = cornell_refined_geometry(k)
    rng = np.random.default_rng(seed)
    per_quad = 2 * k * k
    white = _white(scene)
    tex_ids, mats = [], []
    for q in range(4):
        if textures is None:
            base = rng.uniform(0.25, 0.8, size=3)
            img = np.clip(base + 0.15 * rng.standard_normal((tex_size, tex_size, 3)), 0.0, 1.0)
            rgb = np.round(img * 255).astype(np.uint8)
        else:
            rgb = np.asarray(textures[q], np.uint8)
        rgba = np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], axis=2)
        t = scene.add_texture(rgba, is_srgb=True, generate_mipmaps=mipmaps)
        tex_ids.append(t)
        mats.append(scene.add_material(A.DiffuseMaterial, t, main_color=(1.0, 1.0, 1.0)))
    plain = scene.add_material(A.DiffuseMaterial, white, main_color=(0.7, 0.7, 0.7))
    light = scene.add_material(A.EmissiveMaterial, white, main_color=(1.0, 1.0, 1.0), strength=10.0)
    # quads: floor, ceiling, back, left, right, then 12 box faces
    quad_mat = {0: mats[0], 2: mats[1], 3: mats[2], 4: mats[3]}
    shapes = []
    n_quads = (len(idx) // 3 - n_light) // per_quad
    for qi in range(n_quads):
        shapes.append((quad_mat.get(qi, plain), 3 * qi * per_quad, 3 * per_quad))
    shapes.append((light, 3 * n_quads * per_quad, 3 * n_light))
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.0, 0.0, 0.0))
    cam = scene.add_camera(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth, termination_start_depth=16)
    scene.set_current_cam(cam)
    return SceneSpec("textured_cornell", width, height, spp, depth, triangles=len(idx) // 3,
                     info={"textures": tex_ids, "tex_size": tex_size})
