"""The demo's scene formats, read into a scene of this library.

load_scene follows the demo's LoadScene (src/DemoLib/load/Load.cpp:24-443):
a JSON object with "environment" (env_col, env_clamp, env_map, and a sun:
sun_dir, sun_col, sun_softness, added as a directional light), "camera"
(view_origin, view_dir or view_target, view_up, fov, filter "box"/"tent",
srgb), "materials" (type diffuse/glossy/refractive/emissive/mix/transparent,
main_texture, main_color, normal_map, roughness or strength, fresnel,
int_ior, ext_ior; a mix names its two materials), "meshes" (vertex_data in
the .bin format, per-group materials, allow_spatial_splits,
use_fast_bvh_build) and "mesh_instances" (mesh, pos, rot in degrees applied
z, then x, then y).  Objects are read in file order, so textures, materials,
meshes and instances get the handles the demo's loader gives them.

load_bin reads the .bin mesh format (Load.cpp:616-638): three uint32 counts
(attribute floats, indices, group words), then the attributes (8 floats a
vertex: position, normal, uv), the indices and the groups (start index,
index count per material group).

Textures: JPEG/PNG/TGA are decoded with Pillow (the demo uses stb_image,
whose JPEG decoder may differ from Pillow's in the last bit of some texels),
Radiance .hdr (RGBE, new-style run-length encoding) by read_hdr, kept as RGBE
texels as the demo's LoadHDR does.

load_obj follows the demo's OBJ reader (Load.cpp:445-614): triangles only,
each corner v/vt/vn; a corner reuses one of the last 1000 vertices whose 8
attributes all lie within 1e-7 of its own (searched from the newest), else it
appends a vertex; "g" lines start material groups.
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np

from . import capi as A

_MAT_TYPES = {"diffuse": A.DiffuseMaterial, "glossy": A.GlossyMaterial, "refractive": A.RefractiveMaterial,
              "emissive": A.EmissiveMaterial, "mix": A.MixMaterial, "transparent": A.TransparentMaterial}


def load_bin(path: str):
    """(attrs float32 (n, 8), indices uint32, groups uint32 (k, 2))."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 12:
        raise ValueError(f"{path}: truncated header")
    n_attrs, n_idx, n_groups = struct.unpack_from("<3I", data, 0)
    need = 12 + 4 * (n_attrs + n_idx + n_groups)
    if len(data) < need or n_attrs % 8 or n_idx % 3 or n_groups % 2:
        raise ValueError(f"{path}: inconsistent sizes ({n_attrs} attrs, {n_idx} indices, {n_groups} group words)")
    off = 12
    attrs = np.frombuffer(data, "<f4", n_attrs, off).reshape(-1, 8).copy()
    off += 4 * n_attrs
    idx = np.frombuffer(data, "<u4", n_idx, off).copy()
    off += 4 * n_idx
    groups = np.frombuffer(data, "<u4", n_groups, off).reshape(-1, 2).copy()
    if len(idx) and int(idx.max()) >= len(attrs):
        raise ValueError(f"{path}: index out of range")
    return attrs, idx, groups


def load_obj(path: str):
    """(attrs float32 (n, 8), indices uint32, groups uint32 (k, 2)), as load_bin."""
    f32 = np.float32
    v, vn, vt = [], [], []
    attrs = np.zeros((0, 8), f32)
    rows, idx, groups = [], [], []
    eps = f32(1e-7)

    def flush():
        nonlocal attrs, rows
        if rows:
            attrs = np.concatenate([attrs, np.array(rows, f32)])
            rows = []

    with open(path, "r") as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                v.append([f32(float(x)) for x in tok[1:4]])
            elif tok[0] == "vn":
                vn.append([f32(float(x)) for x in tok[1:4]])
            elif tok[0] == "vt":
                vt.append([f32(float(x)) for x in tok[1:3]])
            elif tok[0] == "f":
                for corner in tok[1:4]:
                    parts = (corner.split("/") + ["", ""])[:3]
                    i1, i2, i3 = (int(p) - 1 if p else -1 for p in parts)
                    a = np.array(v[i1] + vn[i3] + vt[i2], f32)
                    flush()
                    n = len(attrs)
                    lo = max(0, n - 1000)
                    hit = -1
                    if n:
                        win = attrs[lo:n]
                        ok = np.all(np.abs(win - a) < eps, axis=1)
                        where = np.nonzero(ok)[0]
                        if len(where):
                            hit = lo + int(where[-1])
                    if hit >= 0:
                        idx.append(hit)
                    else:
                        idx.append(n + len(rows))
                        rows.append(a)
                        flush()
            elif tok[0] == "g":
                if groups:
                    groups.append(len(idx) - groups[-1])
                groups.append(len(idx))
    flush()
    if not groups:
        groups.append(0)
    groups.append(len(idx) - groups[-1])
    return attrs, np.array(idx, np.uint32), np.array(groups, np.uint32).reshape(-1, 2)


def save_bin(path: str, attrs, indices, groups):
    """Writes the .bin mesh format (the inverse of load_bin)."""
    attrs = np.ascontiguousarray(attrs, "<f4").ravel()
    indices = np.ascontiguousarray(indices, "<u4").ravel()
    groups = np.ascontiguousarray(groups, "<u4").ravel()
    with open(path, "wb") as f:
        f.write(struct.pack("<3I", attrs.size, indices.size, groups.size))
        f.write(attrs.tobytes())
        f.write(indices.tobytes())
        f.write(groups.tobytes())


def read_hdr(path: str) -> np.ndarray:
    """A Radiance RGBE image (new-style RLE scanlines) as (h, w, 4) uint8 RGBE."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode("latin-1")
        pos = e + 1
        return s

    if line() != "#?RADIANCE":
        raise ValueError(f"{path}: not a Radiance file")
    fmt = None
    while True:
        s = line()
        if not s:
            break
        if s.startswith("FORMAT="):
            fmt = s[7:]
    if fmt != "32-bit_rle_rgbe":
        raise ValueError(f"{path}: unsupported format {fmt}")
    parts = line().split()
    if len(parts) != 4 or parts[0] != "-Y" or parts[2] != "+X":
        raise ValueError(f"{path}: unsupported orientation")
    h, w = int(parts[1]), int(parts[3])
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if data[pos] != 2 or data[pos + 1] != 2 or (data[pos + 2] << 8 | data[pos + 3]) != w:
            raise ValueError(f"{path}: only run-length encoded scanlines are read")
        pos += 4
        for c in range(4):
            x = 0
            while x < w:
                n = data[pos]
                pos += 1
                if n > 128:
                    n -= 128
                    out[y, x:x + n, c] = data[pos]
                    pos += 1
                else:
                    out[y, x:x + n, c] = np.frombuffer(data, np.uint8, n, pos)
                    pos += n
                x += n
    return out


def _read_image(path: str) -> np.ndarray:
    if path.lower().endswith(".hdr"):
        return read_hdr(path)
    from PIL import Image
    with Image.open(path) as im:
        return np.asarray(im.convert("RGBA"), np.uint8).copy()


def _translate(m, v):
    r = m.copy()
    r[3] = r[3] + (m[0] * v[0] + m[1] * v[1] + m[2] * v[2])
    return r


def _rotate(m, angle, axis):
    """Ren::Rotate (src/Ren/MMat.h:409-437) in float32; m[i] is column i."""
    f = np.float32
    c = f(math.cos(float(angle)))
    s = f(math.sin(float(angle)))
    axis = np.asarray(axis, f)
    dot = axis[0] * axis[0]
    dot = dot + axis[1] * axis[1]
    dot = dot + axis[2] * axis[2]
    axis = axis / np.sqrt(dot)
    temp = (f(1) - c) * axis
    rot = np.zeros((3, 3), f)
    rot[0] = (c + temp[0] * axis[0], f(0) + temp[0] * axis[1] + s * axis[2], f(0) + temp[0] * axis[2] - s * axis[1])
    rot[1] = (f(0) + temp[1] * axis[0] - s * axis[2], c + temp[1] * axis[1], f(0) + temp[1] * axis[2] + s * axis[0])
    rot[2] = (f(0) + temp[2] * axis[0] + s * axis[1], f(0) + temp[2] * axis[1] - s * axis[0], c + temp[2] * axis[2])
    r = np.empty_like(m)
    for i in range(3):
        r[i] = m[0] * rot[i][0] + m[1] * rot[i][1] + m[2] * rot[i][2]
    r[3] = m[3]
    return r


def instance_transform(pos=None, rot=None) -> np.ndarray:
    """The demo's instance matrix (Load.cpp:322-351) as 16 floats, columns
    first: translate by pos, then rotate by rot (degrees) about z, x and y."""
    f = np.float32
    m = np.eye(4, dtype=f)
    if pos is not None:
        m = _translate(m, [f(v) for v in pos])
    if rot is not None:
        k = f(math.pi) / f(180.0)
        rx, ry, rz = (f(v) * k for v in rot)
        m = _rotate(m, rz, (0.0, 0.0, 1.0))
        m = _rotate(m, rx, (1.0, 0.0, 0.0))
        m = _rotate(m, ry, (0.0, 1.0, 0.0))
    return m.reshape(16)


def load_scene(scene, js, base_dir: str = ".", missing_texture: str = "raise") -> dict:
    """Fills `scene` (a capi.Scene) from the demo's JSON (a dict, or the path
    of a .json file); relative file names are taken from base_dir (the demo
    runs from the repository root, where they start with assets/).
    missing_texture: "raise" (the demo's behaviour: the scene fails to load)
    or "white" (a 1x1 white texture stands in).  Returns the handles:
    {"textures", "materials", "meshes", "instances", "camera", "sun"}."""
    if isinstance(js, (str, os.PathLike)):
        with open(js, "r") as f:
            js = json.load(f)
    textures, materials, meshes, instances = {}, {}, {}, []

    def texture(name, srgb, mips):
        key = name
        if key in textures:
            return textures[key]
        path = os.path.join(base_dir, name)
        if not os.path.exists(path):
            if missing_texture != "white":
                raise FileNotFoundError(f"texture {name} not found under {base_dir}")
            rgba = np.full((1, 1, 4), 255, np.uint8)
        else:
            rgba = _read_image(path)
        textures[key] = scene.add_texture(rgba, is_srgb=srgb, generate_mipmaps=mips)
        return textures[key]

    sun = None
    env = js.get("environment")
    if env is not None:
        col = env.get("env_col", (0.0, 0.0, 0.0))
        env_map = texture(env["env_map"], False, False) if "env_map" in env else A.INVALID
        scene.set_environment(tuple(float(c) for c in col), float(env.get("env_clamp", 0.0)), env_map)
        sun = scene.add_light(A.DirectionalLight, direction=tuple(float(v) for v in env["sun_dir"]),
                              color=tuple(float(v) for v in env["sun_col"]),
                              angle=float(env.get("sun_softness", 0.0)))
    for name, m in js["materials"].items():
        kind = m["type"]
        if kind not in _MAT_TYPES:
            raise ValueError(f"unknown material type {kind}")
        main_tex = texture(m["main_texture"], kind != "mix", kind != "mix")
        kw = {}
        if "main_color" in m:
            kw["main_color"] = tuple(float(v) for v in m["main_color"])
        if "normal_map" in m:
            kw["normal_map"] = texture(m["normal_map"], False, False)
        if "roughness" in m:
            kw["roughness"] = float(m["roughness"])
        elif "strength" in m:
            kw["strength"] = float(m["strength"])
        for k in ("fresnel", "int_ior", "ext_ior"):
            if k in m:
                kw[k] = float(m[k])
        if kind == "mix":
            mix = [materials[x] for x in m["materials"] if x in materials]
            mix = (mix + [A.INVALID, A.INVALID])[:2]
            kw["mix_materials"] = tuple(mix)
        materials[name] = scene.add_material(_MAT_TYPES[kind], main_tex, **kw)
    for name, m in js["meshes"].items():
        vd = m["vertex_data"]
        if ".obj" in vd:
            attrs, idx, groups = load_obj(os.path.join(base_dir, vd))
        elif ".bin" in vd:
            attrs, idx, groups = load_bin(os.path.join(base_dir, vd))
        else:
            raise ValueError(f"{vd}: unknown mesh format")
        shapes = []
        for g, (start, count) in enumerate(groups):
            mat = materials[m["materials"][g]]
            shapes.append((mat, mat, int(start), int(count)))
        meshes[name] = scene.add_mesh(A.PxyzNxyzTuv, attrs, idx, shapes,
                                      allow_spatial_splits=bool(m.get("allow_spatial_splits", False)),
                                      use_fast_bvh_build=bool(m.get("use_fast_bvh_build", False)))
    for inst in js["mesh_instances"]:
        x = instance_transform(inst.get("pos"), inst.get("rot"))
        instances.append(scene.add_mesh_instance(meshes[inst["mesh"]], x.reshape(4, 4)))
    cam = js.get("camera", {})
    f = np.float32
    origin = np.array(cam.get("view_origin", (0.0, 0.0, 0.0)), f)
    fwd = np.array(cam.get("view_dir", (0.0, 0.0, -1.0)), f)
    up = np.array(cam.get("view_up", (0.0, 0.0, 0.0)), f)
    if "view_target" in cam:
        # Normalize(-(origin - target)), Ren::Normalize (src/Ren/MVec.h:200-203)
        d = -(origin - np.array(cam["view_target"], f))
        dot = d[0] * d[0]
        dot = dot + d[1] * d[1]
        dot = dot + d[2] * d[2]
        fwd = d / np.sqrt(dot)
    filt = {"box": A.Box, "tent": A.Tent}.get(cam.get("filter", "tent"))
    if filt is None:
        raise ValueError("unknown camera filter")
    c = scene.add_camera(origin=tuple(float(v) for v in origin), fwd=tuple(float(v) for v in fwd),
                         up=tuple(float(v) for v in up), fov=float(cam.get("fov", 45.0)), filter=filt,
                         dtype=A.DeviceSRGB if cam.get("srgb", True) else A.DeviceNone, clamp=True)
    scene.set_current_cam(c)
    return {"textures": textures, "materials": materials, "meshes": meshes, "instances": instances, "camera": c,
            "sun": sun}
