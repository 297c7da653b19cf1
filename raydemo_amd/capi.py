"""ctypes bindings for the flat C ABI of the Ray renderer/scene API (include/rh_scene.h).

The same binding drives any shared library that exports the ``rhx_*`` entry
points: the HIP backend's host layer (``libray_hip_host.so``) and, in tests,
the reference library built behind the same ABI.  Structures mirror
include/rh_types.h and include/rh_scene.h field for field.

Reference interface mirrored: Ray::RendererBase (src/Ray/RendererBase.h:55-103),
Ray::SceneBase (src/Ray/SceneBase.h:157-285), Ray::CreateRenderer
(src/Ray/RendererFactory.h:19).
"""
from __future__ import annotations

import ctypes as C
from ctypes import POINTER, Structure, c_float, c_int, c_int32, c_size_t, c_uint8, c_uint16, c_uint32, c_uint64, c_void_p

import numpy as np

# ---- enums (src/Ray/SceneBase.h, src/Ray/Types.h) ----
DiffuseMaterial, GlossyMaterial, RefractiveMaterial, EmissiveMaterial, MixMaterial, TransparentMaterial = range(6)
PointLight, SpotLight, DirectionalLight = range(3)
Persp, Ortho, Geo = range(3)
Box, Tent = range(2)
DeviceNone, DeviceSRGB = range(2)
PxyzNxyzTuv, PxyzNxyzTuvTuv, PxyzNxyzBxyzTuv, PxyzNxyzBxyzTuvTuv = range(4)
ATTR_STRIDES = (8, 10, 11, 13)

RENDERER_REF = 1 << 0
RENDERER_SSE2 = 1 << 1
RENDERER_AVX = 1 << 2
RENDERER_AVX2 = 1 << 3
RENDERER_HIP = 1 << 6

INVALID = 0xFFFFFFFF
HALTON_COUNT = 24
HALTON_SEQ_LEN = 256


class MatDesc(Structure):
    _fields_ = [("type", c_int32), ("main_color", c_float * 3), ("main_texture", c_uint32), ("normal_map", c_uint32),
                ("mix_materials", c_uint32 * 2), ("roughness", c_float), ("strength", c_float), ("fresnel", c_float),
                ("int_ior", c_float), ("ext_ior", c_float)]


class ShapeDesc(Structure):
    _fields_ = [("mat_index", c_uint32), ("back_mat_index", c_uint32), ("vtx_start", c_uint64), ("vtx_count", c_uint64)]


class LightDesc(Structure):
    _fields_ = [("type", c_int32), ("position", c_float * 3), ("radius", c_float), ("color", c_float * 3),
                ("direction", c_float * 3), ("angle", c_float)]


class CamDesc(Structure):
    _fields_ = [("type", c_int32), ("filter", c_int32), ("dtype", c_int32), ("origin", c_float * 3), ("fwd", c_float * 3),
                ("up", c_float * 3), ("fov", c_float), ("gamma", c_float), ("focus_distance", c_float),
                ("focus_factor", c_float), ("mi_index", c_uint32), ("uv_index", c_uint32),
                ("lighting_only", c_uint8), ("skip_direct_lighting", c_uint8), ("skip_indirect_lighting", c_uint8),
                ("no_background", c_uint8), ("clamp", c_uint8), ("output_sh", c_uint8),
                ("use_coherent_sampling", c_uint8), ("pad0", c_uint8), ("max_diff_depth", c_uint8),
                ("max_glossy_depth", c_uint8), ("max_refr_depth", c_uint8), ("max_transp_depth", c_uint8),
                ("max_total_depth", c_uint8), ("termination_start_depth", c_uint8), ("pad1", c_uint8 * 2)]


class Settings(Structure):
    _fields_ = [("w", c_int32), ("h", c_int32), ("device_index", c_int32), ("use_wide_bvh", c_int32),
                ("sort_rays", c_int32), ("lds_stack_depth", c_int32)]


class Stats(Structure):
    _fields_ = [("time_primary_ray_gen_us", c_uint64), ("time_primary_trace_us", c_uint64),
                ("time_primary_shade_us", c_uint64), ("time_secondary_sort_us", c_uint64),
                ("time_secondary_trace_us", c_uint64), ("time_secondary_shade_us", c_uint64)]


class RayCounts(Structure):
    _fields_ = [("primary", c_uint64), ("secondary", c_uint64), ("shadow", c_uint64), ("sorted", c_uint64),
                ("adjoint", c_uint64), ("adjoint_shadow", c_uint64), ("overflow", c_uint64)]


TRACE_STAGES = 17  # RH_TRACE_STAGES: the primary rays, then bounces 0..15


class KernelProfile(Structure):
    _fields_ = [("ms", C.c_double * 5), ("launches", c_uint64 * 5), ("node_visits", c_uint64),
                ("tri_tests", c_uint64), ("instance_tests", c_uint64), ("counted_rays", c_uint64),
                ("shadow_visits", c_uint64 * 5), ("tail_visits", c_uint64 * 5),
                ("trace_stage_ms", C.c_double * TRACE_STAGES), ("trace_stage_launches", c_uint64 * TRACE_STAGES),
                ("trace_stage_visits", (c_uint64 * 4) * TRACE_STAGES)]


class PassSettings(Structure):
    _fields_ = [("max_diff_depth", c_uint8), ("max_glossy_depth", c_uint8), ("max_refr_depth", c_uint8),
                ("max_transp_depth", c_uint8), ("max_total_depth", c_uint8), ("termination_start_depth", c_uint8),
                ("pad", c_uint8 * 2), ("flags", c_uint32)]


class Camera(Structure):
    _fields_ = [("type", c_int32), ("filter", c_int32), ("dtype", c_int32), ("fov", c_float), ("gamma", c_float),
                ("focus_distance", c_float), ("focus_factor", c_float), ("origin", c_float * 3), ("fwd", c_float * 3),
                ("side", c_float * 3), ("up", c_float * 3), ("mi_index", c_uint32), ("uv_index", c_uint32),
                ("pass_settings", PassSettings)]


class PassInfo(Structure):
    _fields_ = [("index", c_int32), ("rand_index", c_int32), ("iteration", c_int32), ("bounce", c_int32),
                ("settings", PassSettings)]


class SceneView(Structure):
    _fields_ = [
        ("env", c_void_p),
        ("mesh_instances", c_void_p), ("mesh_instances_count", c_uint32),
        ("mi_indices", c_void_p), ("mi_indices_count", c_uint32),
        ("meshes", c_void_p), ("meshes_count", c_uint32),
        ("transforms", c_void_p), ("transforms_count", c_uint32),
        ("vtx_indices", c_void_p), ("vtx_indices_count", c_uint32),
        ("vertices", c_void_p), ("vertices_count", c_uint32),
        ("nodes", c_void_p), ("nodes_count", c_uint32),
        ("tris", c_void_p), ("tris_count", c_uint32),
        ("tri_indices", c_void_p), ("tri_indices_count", c_uint32),
        ("materials", c_void_p), ("materials_count", c_uint32),
        ("textures", c_void_p), ("textures_count", c_uint32),
        ("tex_mips", c_void_p),
        ("texels", c_void_p), ("texels_count", c_uint32),
        ("lights", c_void_p), ("lights_count", c_uint32),
        ("li_indices", c_void_p), ("li_indices_count", c_uint32),
        ("macro_nodes_root", c_uint32),
        ("light_nodes_root", c_uint32),
        ("mnodes", c_void_p), ("mnodes_count", c_uint32),
    ]


# numpy dtypes for the record arrays (include/rh_types.h)
TRI_DTYPE = np.dtype([("nu", "<f4"), ("nv", "<f4"), ("np", "<f4"), ("pu", "<f4"), ("pv", "<f4"), ("ci", "<i4"),
                      ("e0u", "<f4"), ("e0v", "<f4"), ("e1u", "<f4"), ("e1v", "<f4"), ("mi", "<u4"), ("back_mi", "<u4")])
NODE_DTYPE = np.dtype([("bbox_min", "<f4", 3), ("prim_index", "<u4"), ("bbox_max", "<f4", 3), ("prim_count", "<u4")])
MBVH_DTYPE = np.dtype([("bbox_min", "<f4", (3, 8)), ("bbox_max", "<f4", (3, 8)), ("child", "<u4", 8)])
MBVH_EMPTY = 0x7FFFFFFF
VERTEX_DTYPE = np.dtype([("p", "<f4", 3), ("n", "<f4", 3), ("b", "<f4", 3), ("t", "<f4", (2, 2))])
MESH_DTYPE = np.dtype([("node_index", "<u4"), ("node_count", "<u4"), ("tris_index", "<u4"), ("tris_count", "<u4")])
TRANSFORM_DTYPE = np.dtype([("xform", "<f4", 16), ("inv_xform", "<f4", 16)])
MI_DTYPE = np.dtype([("bbox_min", "<f4", 3), ("tr_index", "<u4"), ("bbox_max", "<f4", 3), ("mesh_index", "<u4")])
MATERIAL_DTYPE = np.dtype([("textures", "<u4", 4), ("main_color", "<f4", 3), ("pad", "<f4"), ("type", "<u4"),
                           ("roughness", "<f4"), ("int_ior", "<f4"), ("ext_ior", "<f4")])
TEXTURE_DTYPE = np.dtype([("width", "<u2"), ("height", "<u2"), ("page", "u1", 14), ("pos", "<u2", (14, 2))])
TEX_MIP_DTYPE = np.dtype([("offset", "<u4"), ("bw", "<u2"), ("bh", "<u2")])
LIGHT_DTYPE = np.dtype([("pos", "<f4", 3), ("radius", "<f4"), ("col", "<f4", 3), ("brightness", "<f4"),
                        ("dir", "<f4", 3), ("spot", "<f4")])
RAY_PACKET_DTYPE = np.dtype([("o", "<f4", 3), ("d", "<f4", 3), ("c", "<f4", 3), ("ior", "<f4"), ("do_dx", "<f4", 3),
                             ("dd_dx", "<f4", 3), ("do_dy", "<f4", 3), ("dd_dy", "<f4", 3), ("xy", "<i4"),
                             ("ray_depth", "<i4")])
HIT_DTYPE = np.dtype([("mask", "<i4"), ("obj", "<i4"), ("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("xy", "<i4"), ("pad", "<i4")])
PASS_INFO_DTYPE = np.dtype([("index", "<i4"), ("rand_index", "<i4"), ("iteration", "<i4"), ("bounce", "<i4"),
                            ("max_diff_depth", "u1"), ("max_glossy_depth", "u1"), ("max_refr_depth", "u1"),
                            ("max_transp_depth", "u1"), ("max_total_depth", "u1"), ("termination_start_depth", "u1"),
                            ("pad", "u1", 2), ("flags", "<u4")])
assert TRI_DTYPE.itemsize == 48 and NODE_DTYPE.itemsize == 32 and VERTEX_DTYPE.itemsize == 52
assert MATERIAL_DTYPE.itemsize == 48 and TEXTURE_DTYPE.itemsize == 74 and LIGHT_DTYPE.itemsize == 48
assert RAY_PACKET_DTYPE.itemsize == 96 and HIT_DTYPE.itemsize == 32 and PASS_INFO_DTYPE.itemsize == 28


def _arr(ptr, count, dtype):
    if not ptr or count == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


class RayLib:
    """A loaded library exporting the rhx_* ABI."""

    def __init__(self, path: str):
        self.path = path
        self.lib = C.CDLL(path, mode=C.RTLD_LOCAL)
        L = self.lib
        P = c_void_p
        sig = {
            "rhx_last_error": (C.c_char_p, []),
            "rhx_create_renderer": (c_int, [POINTER(Settings), c_uint32, POINTER(P)]),
            "rhx_destroy_renderer": (None, [P]),
            "rhx_renderer_type": (c_uint32, [P]),
            "rhx_resize": (c_int, [P, c_int, c_int]),
            "rhx_clear": (c_int, [P, POINTER(c_float)]),
            "rhx_get_pixels": (c_int, [P, P, c_size_t]),
            "rhx_get_stats": (c_int, [P, POINTER(Stats)]),
            "rhx_reset_stats": (c_int, [P]),
            "rhx_get_ray_counts": (c_int, [P, POINTER(RayCounts)]),
            "rhx_set_option": (c_int, [P, C.c_char_p, c_int]),
            "rhx_synchronize": (c_int, [P]),
            "rhx_get_kernel_profile": (c_int, [P, POINTER(KernelProfile)]),
            "rhx_reset_kernel_profile": (c_int, [P]),
            "rhx_create_scene": (c_int, [P, POINTER(P)]),
            "rhx_destroy_scene": (None, [P]),
            "rhx_region_create": (c_int, [c_int, c_int, c_int, c_int, POINTER(P)]),
            "rhx_region_destroy": (None, [P]),
            "rhx_region_iteration": (c_int, [P]),
            "rhx_region_clear": (c_int, [P]),
            "rhx_region_halton": (c_int, [P, P, c_size_t]),
            "rhx_render_scene": (c_int, [P, P, P]),
            "rhx_set_environment": (c_int, [P, POINTER(c_float), c_float, c_uint32]),
            "rhx_add_texture": (c_uint32, [P, P, c_int, c_int, c_int, c_int]),
            "rhx_add_material": (c_uint32, [P, POINTER(MatDesc)]),
            "rhx_add_mesh": (c_uint32, [P, c_int, P, c_size_t, P, c_size_t, c_int, P, c_size_t, c_int]),
            "rhx_remove_mesh": (None, [P, c_uint32]),
            "rhx_add_light": (c_uint32, [P, POINTER(LightDesc)]),
            "rhx_add_mesh_instance": (c_uint32, [P, c_uint32, POINTER(c_float)]),
            "rhx_set_mesh_instance_transform": (None, [P, c_uint32, POINTER(c_float)]),
            "rhx_remove_mesh_instance": (None, [P, c_uint32]),
            "rhx_add_camera": (c_uint32, [P, POINTER(CamDesc)]),
            "rhx_set_camera": (None, [P, c_uint32, POINTER(CamDesc)]),
            "rhx_set_current_cam": (None, [P, c_uint32]),
            "rhx_get_camera_record": (c_int, [P, c_uint32, POINTER(Camera)]),
            "rhx_triangle_count": (c_uint32, [P]),
            "rhx_node_count": (c_uint32, [P]),
            "rhx_scene_view": (c_int, [P, POINTER(SceneView)]),
            "rhx_create_scene_standalone": (c_int, [c_int, POINTER(P)]),
            "rhx_get_sh_data": (c_int, [P, P, c_size_t]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # differentiation entry points (HIP backend only)
        optional = {
            "rhx_adjoint_scene": (c_int, [P, P, P, P, P, P, c_float]),
            "rhx_set_texture_data": (c_int, [P, P, c_uint32, P]),
            "rhx_set_material_colors": (c_int, [P, P, c_uint32, c_uint32, P]),
            "rhx_read_accum": (c_int, [P, P, c_size_t]),
            "rhx_set_gradient_textures": (c_int, [P, P, c_uint32]),
            "rhx_build_fast_bvh": (c_int, [P, c_size_t, c_int, P, c_size_t, POINTER(c_uint32), P]),
            "rhx_build_sah": (c_int, [P, c_size_t, c_int, c_float, c_float, P, c_size_t, POINTER(c_uint32), P,
                                      POINTER(c_int)]),
            "rhx_get_option": (c_int, [P, C.c_char_p, POINTER(c_int)]),
        }
        optional.update({
            "rhx_batch_create": (c_int, [c_int, POINTER(P)]),
            "rhx_batch_destroy": (None, [P]),
            "rhx_batch_add": (c_int, [P, P, c_int, c_int, c_int]),
            "rhx_batch_passes": (c_int, [P]),
            "rhx_batch_render_pass": (c_int, [P]),
            "rhx_batch_get_pixels": (c_int, [P, c_int, P, c_size_t]),
            "rhx_batch_set_option": (c_int, [P, C.c_char_p, c_int]),
            "rhx_batch_read_queue": (c_int, [P, c_int, c_int, c_int, P, c_size_t, POINTER(c_int)]),
            "rhx_batch_get_ray_counts": (c_int, [P, POINTER(RayCounts)]),
            "rhx_batch_reset_stats": (c_int, [P]),
            "rhx_batch_synchronize": (c_int, [P]),
            "rhx_batch_restart": (c_int, [P]),
        })
        self.has_adjoint = all(hasattr(L, n) for n in ("rhx_adjoint_scene", "rhx_set_material_colors", "rhx_read_accum",
                                                       "rhx_set_texture_data"))
        self.has_batch = hasattr(L, "rhx_batch_create")
        for name, (res, args) in optional.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args

    def error(self) -> str:
        e = self.lib.rhx_last_error()
        return e.decode() if e else ""

    def check(self, rc: int, what: str):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.error()}")


class Renderer:
    def __init__(self, lib: RayLib, w: int, h: int, flags: int, use_wide_bvh: bool = False, sort_rays: bool = True,
                 device_index: int = -1, lds_stack_depth: int = 0):
        self.lib = lib
        s = Settings(w, h, device_index, int(use_wide_bvh), int(sort_rays), lds_stack_depth)
        h_ = c_void_p()
        lib.check(lib.lib.rhx_create_renderer(C.byref(s), flags, C.byref(h_)), "rhx_create_renderer")
        self.h = h_
        self.w, self.hgt = w, h

    def close(self):
        if self.h:
            self.lib.lib.rhx_destroy_renderer(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def type(self) -> int:
        return self.lib.lib.rhx_renderer_type(self.h)

    def clear(self, rgba=(0.0, 0.0, 0.0, 0.0)):
        c = (c_float * 4)(*rgba)
        self.lib.check(self.lib.lib.rhx_clear(self.h, c), "rhx_clear")

    def resize(self, w, h):
        self.lib.check(self.lib.lib.rhx_resize(self.h, w, h), "rhx_resize")
        self.w, self.hgt = w, h

    def pixels(self) -> np.ndarray:
        out = np.zeros((self.hgt, self.w, 4), dtype=np.float32)
        self.lib.check(self.lib.lib.rhx_get_pixels(self.h, _ptr(out), out.size), "rhx_get_pixels")
        return out

    def sh_data(self):
        """The accumulated SH data (h, w, 12: coeff_r[4], coeff_g[4],
        coeff_b[4]), or None before any pass with OutputSH."""
        out = np.zeros((self.hgt, self.w, 12), dtype=np.float32)
        rc = self.lib.lib.rhx_get_sh_data(self.h, _ptr(out), out.size)
        if rc == 1:
            return None
        self.lib.check(rc, "rhx_get_sh_data")
        return out

    def stats(self) -> dict:
        st = Stats()
        self.lib.check(self.lib.lib.rhx_get_stats(self.h, C.byref(st)), "rhx_get_stats")
        return {f: getattr(st, f) for f, _ in Stats._fields_}

    def reset_stats(self):
        self.lib.lib.rhx_reset_stats(self.h)

    def ray_counts(self) -> dict:
        rc = RayCounts()
        ok = self.lib.lib.rhx_get_ray_counts(self.h, C.byref(rc))
        if ok != 0:
            return {}
        return {f: getattr(rc, f) for f, _ in RayCounts._fields_}

    def set_option(self, name: str, value: int):
        self.lib.check(self.lib.lib.rhx_set_option(self.h, name.encode(), int(value)), f"rhx_set_option({name})")

    def get_option(self, name: str) -> int:
        v = c_int(0)
        self.lib.check(self.lib.lib.rhx_get_option(self.h, name.encode(), C.byref(v)), f"rhx_get_option({name})")
        return v.value

    def synchronize(self):
        self.lib.check(self.lib.lib.rhx_synchronize(self.h), "rhx_synchronize")

    def kernel_profile(self) -> dict:
        p = KernelProfile()
        self.lib.check(self.lib.lib.rhx_get_kernel_profile(self.h, C.byref(p)), "rhx_get_kernel_profile")
        kinds = ("trace", "shade", "sort", "shadow", "tail")
        d = {f"{k}_ms": p.ms[i] for i, k in enumerate(kinds)}
        d.update({f"{k}_launches": p.launches[i] for i, k in enumerate(kinds)})
        d.update(node_visits=p.node_visits, tri_tests=p.tri_tests, instance_tests=p.instance_tests,
                 counted_rays=p.counted_rays)
        names = ("nodes", "tris", "instances")
        d["shadow"] = dict(zip(names + ("entries", "segments"), list(p.shadow_visits)))
        d["tail"] = dict(zip(names + ("segments", "shadow_segments"), list(p.tail_visits)))
        # closest-hit trace launches of the passes by stage (0: primary rays, b + 1: bounce b)
        d["trace_stages"] = [
            {"stage": k, "ms": p.trace_stage_ms[k], "launches": p.trace_stage_launches[k],
             **dict(zip(names + ("rays",), list(p.trace_stage_visits[k])))}
            for k in range(TRACE_STAGES) if p.trace_stage_launches[k] or p.trace_stage_visits[k][3]]
        return d

    def reset_kernel_profile(self):
        self.lib.check(self.lib.lib.rhx_reset_kernel_profile(self.h), "rhx_reset_kernel_profile")

    def create_scene(self) -> "Scene":
        return Scene(self)

    def render(self, scene: "Scene", region: "Region"):
        self.lib.check(self.lib.lib.rhx_render_scene(self.h, scene.h, region.h), "rhx_render_scene")

    # ---- differentiation (HIP backend)
    def adjoint(self, scene: "Scene", region: "Region", d_adj_rgba: int, d_grad: int, scale: float,
                d_texel_grad: int | None = None):
        """Replays the pass rhx_render_scene would render next for region and adds
        scale * d loss / d main_color to the device array d_grad (4 floats per
        material), and d loss / d stored texel to d_texel_grad when given (4
        floats per texel), for the device adjoint image d_adj_rgba (w*h RGBA)."""
        self.lib.check(self.lib.lib.rhx_adjoint_scene(self.h, scene.h, region.h, c_void_p(d_adj_rgba),
                                                      c_void_p(d_grad), c_void_p(d_texel_grad or 0), float(scale)),
                       "rhx_adjoint_scene")

    def set_texture_data(self, scene: "Scene", tex: int, rgba: np.ndarray):
        rgba = np.ascontiguousarray(rgba, np.uint8)
        self.lib.check(self.lib.lib.rhx_set_texture_data(self.h, scene.h, tex, _ptr(rgba)), "rhx_set_texture_data")

    def set_gradient_textures(self, ids=None):
        """Texel gradients only for these textures (None: every texture)."""
        arr = np.ascontiguousarray(list(ids) if ids else [0], np.uint32)
        self.lib.check(self.lib.lib.rhx_set_gradient_textures(self.h, _ptr(arr), len(ids) if ids else 0),
                       "rhx_set_gradient_textures")

    def set_material_colors(self, scene: "Scene", first: int, rgb: np.ndarray):
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        self.lib.check(self.lib.lib.rhx_set_material_colors(self.h, scene.h, first, rgb.shape[0], _ptr(rgb)),
                       "rhx_set_material_colors")

    def read_accum(self, dst: int | None = None) -> np.ndarray | None:
        """The linear accumulated image (h, w, 4); into the device or host
        address dst when given."""
        w, h = self.w, self.hgt
        if dst is not None:
            self.lib.check(self.lib.lib.rhx_read_accum(self.h, c_void_p(dst), w * h * 4), "rhx_read_accum")
            return None
        out = np.zeros((h, w, 4), np.float32)
        self.lib.check(self.lib.lib.rhx_read_accum(self.h, _ptr(out), out.size), "rhx_read_accum")
        return out


class _NoRenderer:
    def __init__(self, lib, wide=False):
        self.lib = lib
        self.h = None
        self.wide = wide


def build_fast_bvh(lib: RayLib, boxes: np.ndarray, device: int = -1):
    """The fast (HLBVH) tree over boxes (n x 6: bbox_min, bbox_max) on the
    host (device < 0) or with its treelets built on GPU `device`: (nodes,
    indices), nodes in NODE_DTYPE with the root first."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
    idx = np.zeros(max(1, n), np.uint32)
    count = c_uint32(0)
    lib.check(lib.lib.rhx_build_fast_bvh(_ptr(boxes), n, int(device), _ptr(nodes), len(nodes), C.byref(count),
                                         _ptr(idx)), "rhx_build_fast_bvh")
    return nodes[:count.value], idx[:n]


def build_sah(lib: RayLib, boxes: np.ndarray, device: int = -1, oversplit: float = 0.95, cost: float = 0.025):
    """The SAH tree (object splits) over boxes (n x 6), as the instance tree is
    built, on the host (device < 0) or on GPU `device`: (nodes, indices,
    on_device), nodes in NODE_DTYPE with the root first; on_device is False
    where the device declined (tied keys) and the host built it."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    n = len(boxes)
    nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
    idx = np.zeros(max(1, n), np.uint32)
    count = c_uint32(0)
    on_device = c_int(0)
    lib.check(lib.lib.rhx_build_sah(_ptr(boxes), n, int(device), float(oversplit), float(cost), _ptr(nodes), len(nodes),
                                    C.byref(count), _ptr(idx), C.byref(on_device)), "rhx_build_sah")
    return nodes[:count.value], idx[:n], bool(on_device.value)


def standalone_scene(lib: RayLib, wide: bool = False) -> "Scene":
    """A scene built on the host without a renderer, with binary or 8-wide
    trees (use_wide_bvh); any renderer (or batch) of the same library can
    render it."""
    return Scene(_NoRenderer(lib, wide))


class Batch:
    """Many scenes per pass (HIP backend): Ray::Hip::BatchRenderer."""

    def __init__(self, lib: RayLib, device_index: int = -1):
        if not lib.has_batch:
            raise RuntimeError(f"{lib.path} has no batch entry points (HIP backend required)")
        self.lib = lib
        self.h = c_void_p()
        lib.check(lib.lib.rhx_batch_create(device_index, C.byref(self.h)), "rhx_batch_create")
        self.sizes = []

    def close(self):
        if self.h:
            self.lib.lib.rhx_batch_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scene: "Scene", w: int, h: int, spp: int) -> int:
        i = self.lib.lib.rhx_batch_add(self.h, scene.h, w, h, spp)
        if i < 0:
            raise RuntimeError(f"rhx_batch_add: {self.lib.error()}")
        self.sizes.append((w, h))
        return i

    @property
    def passes(self) -> int:
        return self.lib.lib.rhx_batch_passes(self.h)

    def render_pass(self):
        self.lib.check(self.lib.lib.rhx_batch_render_pass(self.h), "rhx_batch_render_pass")

    def render_all(self):
        for _ in range(self.passes):
            self.render_pass()

    def pixels(self, i: int) -> np.ndarray:
        w, h = self.sizes[i]
        out = np.zeros((h, w, 4), np.float32)
        self.lib.check(self.lib.lib.rhx_batch_get_pixels(self.h, i, _ptr(out), out.size), "rhx_batch_get_pixels")
        return out

    def set_option(self, name: str, value: int):
        self.lib.check(self.lib.lib.rhx_batch_set_option(self.h, name.encode(), int(value)),
                       f"rhx_batch_set_option({name})")

    def ray_counts(self) -> dict:
        rc = RayCounts()
        self.lib.check(self.lib.lib.rhx_batch_get_ray_counts(self.h, C.byref(rc)), "rhx_batch_get_ray_counts")
        return {f: getattr(rc, f) for f, _ in RayCounts._fields_}

    def reset_stats(self):
        self.lib.check(self.lib.lib.rhx_batch_reset_stats(self.h), "rhx_batch_reset_stats")

    def read_queue(self, scene: int, stage: int, buffer: int) -> np.ndarray:
        """The traversal halves of the rays entering `stage` of `scene` in ray
        buffer `buffer` after the last pass, as (n, 8) float32 rows
        (o xyz, xy bits, d xyz, depth bits)."""
        w, h = self.sizes[scene]
        out = np.zeros((w * h, 8), np.float32)
        n = c_int(0)
        self.lib.check(self.lib.lib.rhx_batch_read_queue(self.h, scene, stage, buffer, _ptr(out), w * h, C.byref(n)),
                       "rhx_batch_read_queue")
        return out[:n.value]

    def synchronize(self):
        self.lib.check(self.lib.lib.rhx_batch_synchronize(self.h), "rhx_batch_synchronize")

    def restart(self):
        self.lib.check(self.lib.lib.rhx_batch_restart(self.h), "rhx_batch_restart")


class Region:
    def __init__(self, lib: RayLib, x=0, y=0, w=0, h=0):
        self.lib = lib
        h_ = c_void_p()
        lib.check(lib.lib.rhx_region_create(x, y, w, h, C.byref(h_)), "rhx_region_create")
        self.h = h_

    @property
    def iteration(self) -> int:
        return self.lib.lib.rhx_region_iteration(self.h)

    def halton(self) -> np.ndarray:
        out = np.zeros(HALTON_COUNT * HALTON_SEQ_LEN, dtype=np.float32)
        rc = self.lib.lib.rhx_region_halton(self.h, _ptr(out), out.size)
        if rc != 0:
            raise RuntimeError("region has no Halton table yet")
        return out

    def clear(self):
        self.lib.lib.rhx_region_clear(self.h)

    def __del__(self):
        try:
            if self.h:
                self.lib.lib.rhx_region_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Scene:
    def __init__(self, renderer: Renderer):
        self.r = renderer
        self.lib = renderer.lib
        h_ = c_void_p()
        if renderer.h is None:
            self.lib.check(self.lib.lib.rhx_create_scene_standalone(int(getattr(renderer, "wide", False)),
                                                                    C.byref(h_)), "rhx_create_scene_standalone")
        else:
            self.lib.check(self.lib.lib.rhx_create_scene(renderer.h, C.byref(h_)), "rhx_create_scene")
        self.h = h_
        self._keep = []

    def __del__(self):
        try:
            if self.h:
                self.lib.lib.rhx_destroy_scene(self.h)
                self.h = None
        except Exception:
            pass

    def _added(self, index: int, what: str) -> int:
        """The index an adder returned; RuntimeError with the library's reason
        when it returned 0xffffffff (nothing was added)."""
        if index == INVALID:
            raise RuntimeError(f"{what} failed: {self.lib.error()}")
        return index

    def set_environment(self, env_col=(0.0, 0.0, 0.0), env_clamp=0.0, env_map=INVALID):
        self.lib.check(self.lib.lib.rhx_set_environment(self.h, (c_float * 3)(*env_col), env_clamp, env_map),
                       "rhx_set_environment")

    def add_texture(self, rgba: np.ndarray, is_srgb=True, generate_mipmaps=False) -> int:
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        h, w = rgba.shape[:2]
        return self._added(self.lib.lib.rhx_add_texture(self.h, _ptr(rgba), w, h, int(is_srgb), int(generate_mipmaps)),
                           "rhx_add_texture")

    def add_material(self, type, main_texture, main_color=(1.0, 1.0, 1.0), normal_map=INVALID,
                     mix_materials=(INVALID, INVALID), roughness=0.0, strength=1.0, fresnel=1.0, int_ior=1.0,
                     ext_ior=1.0) -> int:
        m = MatDesc(type, (c_float * 3)(*main_color), main_texture, normal_map, (c_uint32 * 2)(*mix_materials),
                    roughness, strength, fresnel, int_ior, ext_ior)
        return self._added(self.lib.lib.rhx_add_material(self.h, C.byref(m)), "rhx_add_material")

    def add_mesh(self, layout, attrs: np.ndarray, indices: np.ndarray, shapes, base_vertex=0,
                 allow_spatial_splits=False, use_fast_bvh_build=False) -> int:
        attrs = np.ascontiguousarray(attrs, dtype=np.float32).ravel()
        indices = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
        sh = (ShapeDesc * len(shapes))()
        for i, s in enumerate(shapes):
            if len(s) == 3:
                sh[i] = ShapeDesc(s[0], INVALID, s[1], s[2])
            else:
                sh[i] = ShapeDesc(s[0], s[1], s[2], s[3])
        n_vtx = attrs.size // ATTR_STRIDES[layout]
        return self._added(self.lib.lib.rhx_add_mesh(self.h, layout, _ptr(attrs), n_vtx, _ptr(indices), indices.size,
                                                     base_vertex, C.cast(sh, c_void_p), len(shapes),
                                                     int(bool(allow_spatial_splits)) | (2 if use_fast_bvh_build else 0)),
                           "rhx_add_mesh")

    def remove_mesh(self, i):
        self.lib.lib.rhx_remove_mesh(self.h, i)

    def add_light(self, type, position=(0, 0, 0), radius=1.0, color=(1, 1, 1), direction=(0, -1, 0), angle=0.0) -> int:
        l = LightDesc(type, (c_float * 3)(*position), radius, (c_float * 3)(*color), (c_float * 3)(*direction), angle)
        return self.lib.lib.rhx_add_light(self.h, C.byref(l))

    def add_mesh_instance(self, mesh, xform=None) -> int:
        x = np.eye(4, dtype=np.float32).ravel() if xform is None else np.asarray(xform, dtype=np.float32).ravel()
        return self.lib.lib.rhx_add_mesh_instance(self.h, mesh, (c_float * 16)(*x))

    def set_mesh_instance_transform(self, mi, xform):
        x = np.asarray(xform, dtype=np.float32).ravel()
        self.lib.lib.rhx_set_mesh_instance_transform(self.h, mi, (c_float * 16)(*x))

    def remove_mesh_instance(self, mi):
        self.lib.lib.rhx_remove_mesh_instance(self.h, mi)

    @staticmethod
    def cam_desc(origin, fwd, up=(0, 0, 0), fov=45.0, type=Persp, filter=Tent, dtype=DeviceSRGB, gamma=1.0,
                 focus_distance=1.0, focus_factor=0.0, max_diff_depth=4, max_glossy_depth=4, max_refr_depth=4,
                 max_transp_depth=4, max_total_depth=4, termination_start_depth=3, output_sh=False, clamp=False,
                 no_background=False, skip_direct_lighting=False, skip_indirect_lighting=False, lighting_only=False,
                 use_coherent_sampling=False, mi_index=0, uv_index=0) -> CamDesc:
        c = CamDesc()
        c.type, c.filter, c.dtype = type, filter, dtype
        c.origin = (c_float * 3)(*origin)
        c.fwd = (c_float * 3)(*fwd)
        c.up = (c_float * 3)(*up)
        c.fov, c.gamma, c.focus_distance, c.focus_factor = fov, gamma, focus_distance, focus_factor
        c.mi_index, c.uv_index = mi_index, uv_index
        c.lighting_only, c.skip_direct_lighting, c.skip_indirect_lighting = lighting_only, skip_direct_lighting, skip_indirect_lighting
        c.no_background, c.clamp, c.output_sh, c.use_coherent_sampling = no_background, clamp, output_sh, use_coherent_sampling
        c.max_diff_depth, c.max_glossy_depth, c.max_refr_depth = max_diff_depth, max_glossy_depth, max_refr_depth
        c.max_transp_depth, c.max_total_depth, c.termination_start_depth = max_transp_depth, max_total_depth, termination_start_depth
        return c

    def add_camera(self, **kw) -> int:
        c = self.cam_desc(**kw)
        return self.lib.lib.rhx_add_camera(self.h, C.byref(c))

    def set_camera(self, i, **kw):
        c = self.cam_desc(**kw)
        self.lib.lib.rhx_set_camera(self.h, i, C.byref(c))

    def set_current_cam(self, i):
        self.lib.lib.rhx_set_current_cam(self.h, i)

    def camera_record(self, i) -> Camera:
        c = Camera()
        self.lib.check(self.lib.lib.rhx_get_camera_record(self.h, i, C.byref(c)), "rhx_get_camera_record")
        return c

    def triangle_count(self) -> int:
        return self.lib.lib.rhx_triangle_count(self.h)

    def node_count(self) -> int:
        return self.lib.lib.rhx_node_count(self.h)

    def raw_view(self) -> SceneView:
        v = SceneView()
        self.lib.check(self.lib.lib.rhx_scene_view(self.h, C.byref(v)), "rhx_scene_view")
        return v

    def view(self) -> dict:
        """Copies of the preprocessed scene arrays (for comparison and for the oracle)."""
        v = self.raw_view()
        d = {
            "mesh_instances": _arr(v.mesh_instances, v.mesh_instances_count, MI_DTYPE),
            "mi_indices": _arr(v.mi_indices, v.mi_indices_count, np.uint32),
            "meshes": _arr(v.meshes, v.meshes_count, MESH_DTYPE),
            "transforms": _arr(v.transforms, v.transforms_count, TRANSFORM_DTYPE),
            "vtx_indices": _arr(v.vtx_indices, v.vtx_indices_count, np.uint32),
            "vertices": _arr(v.vertices, v.vertices_count, VERTEX_DTYPE),
            "nodes": _arr(v.nodes, v.nodes_count, NODE_DTYPE),
            "mnodes": _arr(v.mnodes, v.mnodes_count, MBVH_DTYPE),
            "tris": _arr(v.tris, v.tris_count, TRI_DTYPE),
            "tri_indices": _arr(v.tri_indices, v.tri_indices_count, np.uint32),
            "materials": _arr(v.materials, v.materials_count, MATERIAL_DTYPE),
            "textures": _arr(v.textures, v.textures_count, TEXTURE_DTYPE),
            "tex_mips": _arr(v.tex_mips, v.textures_count * 14, TEX_MIP_DTYPE).reshape(-1, 14),
            "texels_count": v.texels_count,
            "lights": _arr(v.lights, v.lights_count, LIGHT_DTYPE),
            "li_indices": _arr(v.li_indices, v.li_indices_count, np.uint32),
            "macro_nodes_root": v.macro_nodes_root,
            "light_nodes_root": v.light_nodes_root,
        }
        env = (C.c_float * 5).from_address(v.env)
        d["env"] = np.array(env[:4], dtype=np.float32), C.cast(v.env + 16, POINTER(c_uint32))[0]
        return d
