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
  large_instanced  — two-level scene (seed 0x5EED0005): 16 base meshes of
                     jittered icosphere clusters, 1.2k..80k triangles each
                     (log-uniform), 4096 instances with seeded affine transforms
                     (uniform rotation, log-uniform scale, positions in a 120 x 20
                     x 120 field; instances spread over the meshes with weight
                     size^-2.5, ~7.5 M instanced triangles) over a textured ground
                     quad; 32 RGBA8 sRGB textures of 2048^2 with mips (seeded
                     checker + noise), each mesh diffuse-textured or a mix of
                     textured diffuse and glossy (roughness U[0.1, 0.6]); four
                     sphere lights and a constant environment; 3840x2160, 16 spp,
                     depth 12 (beyond the reference's cap of 10: the renderer's
                     "max_bounces" extension).
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


def cornell_box(scene: "A.Scene", seed: int = SEED_BASE + 1, width=256, height=256, spp=16, depth=4,
                fast_bvh=False, coherent=False, spatial_splits=False) -> SceneSpec:
    attrs, idx, shapes, albedos = cornell_geometry(seed)
    tex = _white(scene)
    mats = [scene.add_material(A.DiffuseMaterial, tex, main_color=tuple(float(v) for v in albedos[k]))
            for k in range(7)]
    mats.append(scene.add_material(A.EmissiveMaterial, tex, main_color=(1.0, 1.0, 1.0), strength=10.0))
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(mats[k], s, c) for k, s, c in shapes],
                          use_fast_bvh_build=fast_bvh, allow_spatial_splits=spatial_splits)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.0, 0.0, 0.0))
    cam = scene.add_camera(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth, use_coherent_sampling=coherent)
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


def lightmap_box(scene: "A.Scene", k: int = 2, width=64, height=64, spp=4, depth=4, output_sh=True,
                 perspective=False) -> SceneSpec:
    """The Cornell box (faces as k x k grids, one material per face kind) with
    a second UV layer that packs every triangle into its own cell of a square
    grid (a lightmap atlas without overlaps), and a geometry camera on mesh
    instance 0, UV layer 1 (camera_desc_t type Geo): each pass bakes one sample
    per lightmap texel; with output_sh the SH projection of the pass's
    secondary-ray directions too (OutputSH).  perspective: a perspective
    camera instead (with the same SH flag)."""
    attrs, idx, n_light = cornell_refined_geometry(k)
    n_tris = len(idx) // 3
    verts = attrs[idx]                           # one vertex per triangle corner
    g = int(math.ceil(math.sqrt(n_tris)))
    cell = 1.0 / g
    corner = np.array([(0.15, 0.15), (0.85, 0.15), (0.15, 0.85)], np.float32)
    uv1 = np.zeros((len(verts), 2), np.float32)
    for t in range(n_tris):
        cx, cy = (t % g) * cell, (t // g) * cell
        uv1[3 * t:3 * t + 3] = np.array([cx, cy], np.float32) + corner * cell
    a = np.zeros((len(verts), 10), np.float32)
    a[:, :8] = verts
    a[:, 8:10] = uv1
    tex = _white(scene)
    kinds = [scene.add_material(A.DiffuseMaterial, tex, main_color=c)
             for c in ((0.75, 0.75, 0.75), (0.8, 0.2, 0.2), (0.2, 0.75, 0.25), (0.6, 0.6, 0.9))]
    light = scene.add_material(A.EmissiveMaterial, tex, main_color=(1.0, 1.0, 1.0), strength=10.0)
    n_face = 2 * k * k
    shapes = []
    for t in range(n_tris - n_light):
        face = t // n_face
        shapes.append((kinds[0 if face < 3 else (1 if face == 3 else (2 if face == 4 else 3))], 3 * t, 3))
    shapes.append((light, 3 * (n_tris - n_light), 3 * n_light))
    mesh = scene.add_mesh(A.PxyzNxyzTuvTuv, a, np.arange(len(verts), dtype=np.uint32), shapes)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.05, 0.05, 0.05))
    scene.add_light(A.PointLight, position=(0.0, 8.5, 0.0), radius=0.3, color=(20.0, 20.0, 20.0))
    common = dict(max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth, max_transp_depth=depth,
                  max_total_depth=depth, termination_start_depth=16, output_sh=output_sh, dtype=A.DeviceNone)
    if perspective:
        cam = scene.add_camera(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box, **common)
    else:
        cam = scene.add_camera(origin=(0.0, 0.0, 0.0), fwd=(0.0, 0.0, 1.0), type=A.Geo, mi_index=0, uv_index=1,
                               **common)
    scene.set_current_cam(cam)
    return SceneSpec("lightmap_box", width, height, spp, depth, triangles=n_tris)


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
                  n_spheres=512, n_fine=38, fast_bvh=False, spatial_splits=False) -> SceneSpec:
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
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes, use_fast_bvh_build=fast_bvh,
                          allow_spatial_splits=spatial_splits)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.5, 0.5, 0.5))
    scene.add_light(A.PointLight, position=(-10.0, 25.0, -10.0), radius=0.5, color=(50.0, 50.0, 50.0))
    scene.add_light(A.PointLight, position=(12.0, 22.0, 8.0), radius=0.5, color=(50.0, 50.0, 50.0))
    cam = scene.add_camera(**soup_camera(depth))
    scene.set_current_cam(cam)
    return SceneSpec("triangle_soup", width, height, spp, depth, triangles=len(idx) // 3)


def sliver_geometry(seed: int = SEED_BASE + 7, n: int = 400):
    """Long thin triangles crossing a 20^3 box in random directions over a
    ground quad: every split of their boxes leaves the halves overlapping, the
    case spatial splits are for.  (attrs, indices) in the PxyzNxyzTuv layout."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-8.0, 8.0, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    half = rng.uniform(4.0, 12.0, size=n)[:, None]
    w = rng.standard_normal((n, 3))
    w -= (w * d).sum(1, keepdims=True) * d
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    width = rng.uniform(0.05, 0.4, size=n)[:, None]
    p0, p1, p2 = c - d * half, c + d * half, c + w * width
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    verts = np.stack([p0, p1, p2], 1).reshape(-1, 3)
    nv = np.repeat(nrm, 3, axis=0)
    uv = np.tile(np.float64([[0, 0], [1, 0], [0, 1]]), (n, 1))
    attrs = np.concatenate([verts, nv, uv], 1)
    ga, gi = _quad((-30.0, -10.0, -30.0), (-30.0, -10.0, 30.0), (30.0, -10.0, 30.0), (30.0, -10.0, -30.0), (0, 1, 0))
    idx = np.concatenate([np.arange(3 * n), gi + 3 * n])
    return np.concatenate([attrs, ga]).astype(np.float32), idx.astype(np.uint32)


def sliver_scene(scene: "A.Scene", seed: int = SEED_BASE + 7, n: int = 400, width=96, height=64, spp=4, depth=4,
                 spatial_splits=True) -> SceneSpec:
    """sliver_geometry lit by two sphere lights and a grey environment, one
    diffuse material for the slivers and one for the ground."""
    attrs, idx = sliver_geometry(seed, n)
    tex = _white(scene)
    m0 = scene.add_material(A.DiffuseMaterial, tex, main_color=(0.7, 0.5, 0.3))
    m1 = scene.add_material(A.DiffuseMaterial, tex, main_color=(0.5, 0.5, 0.5))
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(m0, 0, 3 * n), (m1, 3 * n, 6)],
                          allow_spatial_splits=spatial_splits)
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.3, 0.3, 0.3))
    scene.add_light(A.PointLight, position=(-6.0, 14.0, -8.0), radius=0.5, color=(40.0, 40.0, 40.0))
    scene.add_light(A.PointLight, position=(9.0, 12.0, 4.0), radius=0.5, color=(40.0, 40.0, 40.0))
    cam = scene.add_camera(origin=(0.0, 2.0, -30.0), fwd=(0.0, -0.05, 1.0), fov=50.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth)
    scene.set_current_cam(cam)
    return SceneSpec("sliver_scene", width, height, spp, depth, triangles=len(idx) // 3)


def furnace_box(scene: "A.Scene", p: float = 0.9, rho: float = 0.95, emission: float = 1.0, width=64, height=64,
                spp=32, depth=12) -> SceneSpec:
    """A closed cube seen from its centre, every wall the same mix material:
    with probability 1 - p an emissive surface of strength `emission` (the
    path ends), with probability p a Lambertian surface of albedo rho (the
    path bounces).  The mix's interior IOR is tiny, so its Fresnel factor is 1
    (total internal reflection at every angle but normal incidence) and the
    choice probability is p at every hit.  Russian roulette is off
    (termination_start_depth past the depth), so the expected pixel value is
    (1 - p) emission sum_{k=0..depth} (p rho)^k (furnace_expected)."""
    L = 5.0
    faces = [((-L, -L, -L), (L, -L, -L), (L, -L, L), (-L, -L, L), (0, 1, 0)),
             ((-L, L, L), (L, L, L), (L, L, -L), (-L, L, -L), (0, -1, 0)),
             ((-L, -L, L), (L, -L, L), (L, L, L), (-L, L, L), (0, 0, -1)),
             ((L, -L, -L), (-L, -L, -L), (-L, L, -L), (L, L, -L), (0, 0, 1)),
             ((-L, -L, -L), (-L, -L, L), (-L, L, L), (-L, L, -L), (1, 0, 0)),
             ((L, -L, L), (L, -L, -L), (L, L, -L), (L, L, L), (-1, 0, 0))]
    attrs, idx = [], []
    for f in faces:
        a, i = _quad(*f)
        idx.append(i + 4 * len(attrs))
        attrs.append(a)
    attrs, idx = np.concatenate(attrs), np.concatenate(idx)
    tex = _white(scene)
    emissive = scene.add_material(A.EmissiveMaterial, tex, main_color=(1.0, 1.0, 1.0), strength=emission)
    diffuse = scene.add_material(A.DiffuseMaterial, tex, main_color=(rho, rho, rho), roughness=0.0)
    mix = scene.add_material(A.MixMaterial, tex, mix_materials=(emissive, diffuse), strength=p, int_ior=1e-4,
                             ext_ior=1e-4)
    mesh = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(mix, 0, len(idx))])
    scene.add_mesh_instance(mesh)
    scene.set_environment((0.0, 0.0, 0.0))
    cam = scene.add_camera(origin=(0.3, -0.2, 0.1), fwd=(0.2, 0.1, 1.0), fov=70.0, filter=A.Box, dtype=A.DeviceNone,
                           max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth, termination_start_depth=min(255, depth + 1))
    scene.set_current_cam(cam)
    return SceneSpec("furnace_box", width, height, spp, depth, triangles=len(idx) // 3,
                     info={"p": p, "rho": rho, "emission": emission})


def furnace_expected(p: float, rho: float, emission: float, depth: int) -> float:
    """The closed form of furnace_box: sum over the hit k = 0..depth at which
    the path ends in emission (probability (1 - p) p^k, throughput rho^k)."""
    return (1.0 - p) * emission * sum((p * rho) ** k for k in range(depth + 1))


def soup_camera(depth=8, coherent=False) -> dict:
    """The soup's camera (camera_desc_t fields); coherent: UseCoherentSampling."""
    return dict(origin=(0.0, 8.0, -48.0), fwd=(0.0, -0.15, 1.0), fov=50.0, filter=A.Box, dtype=A.DeviceNone,
                max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth, max_transp_depth=depth,
                max_total_depth=depth, use_coherent_sampling=coherent)


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


def _pattern_texture(rng, n: int) -> np.ndarray:
    """n x n RGBA8: a two-colour checker of seeded period plus a tiled noise."""
    base, alt = rng.uniform(0.15, 0.85, 3).astype(np.float32), rng.uniform(0.15, 0.85, 3).astype(np.float32)
    period = int(rng.choice([8, 16, 32, 64, 128]))
    k = np.arange(n) // period
    chk = ((k[:, None] + k[None, :]) & 1).astype(bool)
    img = np.where(chk[..., None], base, alt)
    t = min(64, n)
    tile = rng.normal(0.0, 0.06, (t, t, 3)).astype(np.float32)
    img = img + np.tile(tile, (n // t, n // t, 1))
    rgb = np.clip(np.round(img * 255.0), 0, 255).astype(np.uint8)
    return np.concatenate([rgb, np.full((n, n, 1), 255, np.uint8)], axis=2)


def _cluster_geometry(rng, target: int):
    """Jittered subdivision-2 icospheres (320 triangles each) in a unit cube,
    about `target` triangles: (attrs (n, 8), indices)."""
    v, f = icosphere(2)
    m = max(1, int(round(target / 320.0)))
    centres = rng.uniform(-1.0, 1.0, (m, 3))
    radii = np.exp(rng.uniform(math.log(0.05), math.log(0.35), m))
    uv = np.stack([0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2 * math.pi),
                   0.5 - np.arcsin(np.clip(v[:, 1], -1, 1)) / math.pi], 1)
    jit = 1.0 + 0.04 * rng.standard_normal((m, len(v)))
    p = centres[:, None, :] + v[None, :, :] * (radii[:, None] * jit)[..., None]
    n = np.broadcast_to(v, p.shape)
    attrs = np.concatenate([p, n, np.broadcast_to(uv, p.shape[:2] + (2,))], axis=2).reshape(-1, 8)
    idx = (f.reshape(1, -1) + (np.arange(m) * len(v))[:, None]).reshape(-1)
    return attrs.astype(np.float32), idx.astype(np.uint32)


def _random_rotation(rng) -> np.ndarray:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def large_instanced(scene: "A.Scene", seed: int = SEED_BASE + 5, width=3840, height=2160, spp=16, depth=12,
                    n_meshes=16, n_instances=4096, tri_range=(1200, 80000), n_textures=32, tex_size=2048,
                    field_size=120.0, fast_bvh=False) -> SceneSpec:
    """The large two-level workload (see the module docstring); the keyword
    arguments scale it down for tests.  info["textures"] are the texture ids,
    info["adjoint_textures"] the 4 differentiated ones (the ground's and those
    of the three most instanced meshes), info["adjoint_texels"] their level-0
    RGB texels.  fast_bvh builds the meshes' trees with the fast (HLBVH) build."""
    rng = np.random.default_rng(seed)
    white = _white(scene)
    tex, texels = [], []
    for _ in range(n_textures):
        rgba = _pattern_texture(rng, tex_size)
        tex.append(scene.add_texture(rgba, is_srgb=True, generate_mipmaps=True))
        texels.append(rgba)
    sizes = np.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]), n_meshes))
    meshes, mesh_tex = [], []
    tris_per_mesh = []
    for k in range(n_meshes):
        attrs, idx = _cluster_geometry(rng, float(sizes[k]))
        t0 = tex[(2 * k + 1) % n_textures]
        t1 = tex[(2 * k + 2) % n_textures]
        diff0 = scene.add_material(A.DiffuseMaterial, t0, main_color=(1.0, 1.0, 1.0))
        if rng.random() < 0.5:
            second = scene.add_material(A.DiffuseMaterial, t1, main_color=(1.0, 1.0, 1.0))
        else:
            d1 = scene.add_material(A.DiffuseMaterial, t1, main_color=(1.0, 1.0, 1.0))
            gl = scene.add_material(A.GlossyMaterial, white, main_color=(0.9, 0.9, 0.9),
                                    roughness=float(rng.uniform(0.1, 0.6)))
            second = scene.add_material(A.MixMaterial, white, strength=0.5, mix_materials=(d1, gl), int_ior=1.5)
        half = (len(idx) // 6) * 3
        shapes = [(diff0, 0, half), (second, half, len(idx) - half)]
        meshes.append(scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes, use_fast_bvh_build=fast_bvh))
        mesh_tex.append((t0, t1))
        tris_per_mesh.append(len(idx) // 3)
    tris_per_mesh = np.array(tris_per_mesh, np.float64)
    # instances per mesh: weight size^-2.5, at least one each
    w = tris_per_mesh ** -2.5
    w /= w.sum()
    counts = np.maximum(1, np.floor(w * n_instances)).astype(int)
    counts[int(np.argmin(tris_per_mesh))] += n_instances - counts.sum()
    order = np.repeat(np.arange(n_meshes), counts)
    rng.shuffle(order)
    half_f = field_size / 2
    for k in order:
        x = np.eye(4, dtype=np.float64)
        s_ = math.exp(rng.uniform(math.log(0.5), math.log(4.0)))
        x[:3, :3] = (_random_rotation(rng) * s_).T  # row-vector convention: rows are the basis images
        x[3, :3] = (rng.uniform(-half_f, half_f), rng.uniform(s_, s_ + 20.0), rng.uniform(-half_f, half_f))
        scene.add_mesh_instance(meshes[int(k)], x.astype(np.float32))
    # ground
    g = field_size * 1.5
    ga, gi = _quad((-g, 0.0, -g), (-g, 0.0, g), (g, 0.0, g), (g, 0.0, -g), (0, 1, 0))
    ga = ga.astype(np.float32)
    ga[:, 6:8] *= 24.0  # tile the ground texture
    ground_mat = scene.add_material(A.DiffuseMaterial, tex[0], main_color=(1.0, 1.0, 1.0))
    ground = scene.add_mesh(A.PxyzNxyzTuv, ga, gi, [(ground_mat, 0, 6)], use_fast_bvh_build=fast_bvh)
    scene.add_mesh_instance(ground)
    scene.set_environment((0.3, 0.35, 0.4))
    for lx, lz in ((-0.3, -0.3), (0.3, -0.2), (-0.2, 0.3), (0.25, 0.3)):
        scene.add_light(A.PointLight, position=(lx * field_size, 60.0, lz * field_size), radius=3.0,
                        color=(900.0, 880.0, 850.0))
    cam = scene.add_camera(origin=(0.0, 30.0, -0.85 * field_size), fwd=(0.0, -0.25, 1.0), fov=60.0, filter=A.Box,
                           dtype=A.DeviceNone, max_diff_depth=depth, max_glossy_depth=depth, max_refr_depth=depth,
                           max_transp_depth=depth, max_total_depth=depth, termination_start_depth=16)
    scene.set_current_cam(cam)
    most = [int(m) for m in np.argsort(-counts)[:3]]
    adj_tex = [tex[0]] + [mesh_tex[m][0] for m in most]
    instanced = int((counts * tris_per_mesh).sum())
    adj_rgb = [texels[tex.index(t)][..., :3] for t in adj_tex]
    return SceneSpec("large_instanced", width, height, spp, depth, triangles=instanced + 2,
                     info={"textures": tex, "adjoint_textures": adj_tex, "adjoint_texels": adj_rgb,
                           "instances": int(counts.sum()) + 1, "base_triangles": int(tris_per_mesh.sum()),
                           "instanced_triangles": instanced})
