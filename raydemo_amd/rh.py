"""ctypes bindings for the kernel library's C ABI (include/rh_hip.h): the
per-stage parity entry points that run one pipeline stage on caller-supplied
rays.  Scene data comes from a scene view (capi.Scene.raw_view())."""
from __future__ import annotations

import ctypes as C
from ctypes import POINTER, Structure, c_float, c_int, c_int32, c_size_t, c_uint32, c_void_p

import numpy as np

from .capi import (HALTON_COUNT, HALTON_SEQ_LEN, HIT_DTYPE, PASS_INFO_DTYPE, RAY_PACKET_DTYPE, Camera, KernelProfile,
                   SceneView)


class PassDesc(Structure):
    _fields_ = [("cam", Camera), ("rect", c_int32 * 4), ("iteration", c_int32), ("sort_rays", c_int32),
                ("root_min", c_float * 3), ("cell_size", c_float * 3)]


def _p(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None


class DeviceLib:
    def __init__(self, path: str):
        self.lib = C.CDLL(path, mode=C.RTLD_LOCAL)
        L = self.lib
        P = c_void_p
        sig = {
            "rh_last_error": (C.c_char_p, []),
            "rh_create": (c_int, [c_int, POINTER(P)]),
            "rh_destroy": (None, [P]),
            "rh_resize": (c_int, [P, c_int, c_int]),
            "rh_clear": (c_int, [P, P]),
            "rh_upload_scene": (c_int, [P, POINTER(SceneView)]),
            "rh_set_halton": (c_int, [P, P, c_size_t]),
            "rh_render_pass": (c_int, [P, POINTER(PassDesc)]),
            "rh_read_pixels": (c_int, [P, P, c_size_t]),
            "rh_synchronize": (c_int, [P]),
            "rh_set_option": (c_int, [P, C.c_char_p, c_int]),
            "rh_get_option": (c_int, [P, C.c_char_p, POINTER(c_int)]),
            "rh_gen_primary_rays": (c_int, [P, POINTER(Camera), P, c_int, c_int, c_int, P, P]),
            "rh_trace_rays": (c_int, [P, P, c_size_t, c_int, P, P]),
            "rh_ray_hashes": (c_int, [P, P, c_size_t, P, P, P]),
            "rh_sort_rays": (c_int, [P, P, c_size_t, P, P]),
            "rh_sort_rays_noalloc": (c_int, [P, P, c_size_t, P, P, P, P, P]),
            "rh_shade_rays": (c_int, [P, P, P, P, c_size_t, P, P, P, POINTER(c_int), P]),
            "rh_get_kernel_profile": (c_int, [P, POINTER(KernelProfile)]),
            "rh_reset_kernel_profile": (c_int, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    def error(self) -> str:
        e = self.lib.rh_last_error()
        return e.decode() if e else ""

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.error()}")


class Context:
    """One device context (rh_context) with a scene uploaded from a scene view."""

    def __init__(self, dev: DeviceLib, device: int = -1, w: int = 64, h: int = 64):
        self.dev = dev
        h_ = c_void_p()
        dev.check(dev.lib.rh_create(device, C.byref(h_)), "rh_create")
        self.h = h_
        dev.check(dev.lib.rh_resize(self.h, w, h), "rh_resize")

    def close(self):
        if self.h:
            self.dev.lib.rh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, view: SceneView):
        self.dev.check(self.dev.lib.rh_upload_scene(self.h, C.byref(view)), "rh_upload_scene")

    def set_option(self, name, value):
        self.dev.check(self.dev.lib.rh_set_option(self.h, name.encode(), int(value)), f"rh_set_option({name})")

    def get_option(self, name) -> int:
        v = c_int()
        self.dev.check(self.dev.lib.rh_get_option(self.h, name.encode(), C.byref(v)), f"rh_get_option({name})")
        return v.value

    def visit_counts(self) -> tuple:
        """(node visits, triangle tests, instance tests, rays) counted by the
        trace launches since the last reset_visit_counts, with "count_visits" on."""
        p = KernelProfile()
        self.dev.check(self.dev.lib.rh_get_kernel_profile(self.h, C.byref(p)), "rh_get_kernel_profile")
        return int(p.node_visits), int(p.tri_tests), int(p.instance_tests), int(p.counted_rays)

    def reset_visit_counts(self):
        self.dev.check(self.dev.lib.rh_reset_kernel_profile(self.h), "rh_reset_kernel_profile")

    def gen_primary(self, cam: Camera, rect, w, h, iteration, halton) -> np.ndarray:
        n = rect[2] * rect[3]
        out = np.zeros(n, RAY_PACKET_DTYPE)
        r = np.array(rect, np.int32)
        halton = np.ascontiguousarray(halton, np.float32)
        assert halton.size >= HALTON_COUNT * HALTON_SEQ_LEN
        self.dev.check(self.dev.lib.rh_gen_primary_rays(self.h, C.byref(cam), _p(r), w, h, iteration, _p(halton),
                                                        _p(out)), "rh_gen_primary_rays")
        return out

    def trace(self, rays: np.ndarray, any_hit=False, tmax=None) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_PACKET_DTYPE)
        out = np.zeros(len(rays), HIT_DTYPE)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        self.dev.check(self.dev.lib.rh_trace_rays(self.h, _p(rays), len(rays), int(any_hit), _p(tm), _p(out)),
                       "rh_trace_rays")
        return out

    def ray_hashes(self, rays, root_min, cell_size) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_PACKET_DTYPE)
        out = np.zeros(len(rays), np.uint32)
        rm = np.array(root_min, np.float32)
        cs = np.array(cell_size, np.float32)
        self.dev.check(self.dev.lib.rh_ray_hashes(self.h, _p(rays), len(rays), _p(rm), _p(cs), _p(out)),
                       "rh_ray_hashes")
        return out

    def sort_rays(self, rays, root_min, cell_size) -> np.ndarray:
        rays = np.array(rays, RAY_PACKET_DTYPE, copy=True)
        rm = np.array(root_min, np.float32)
        cs = np.array(cell_size, np.float32)
        self.dev.check(self.dev.lib.rh_sort_rays(self.h, _p(rays), len(rays), _p(rm), _p(cs)), "rh_sort_rays")
        return rays

    def shade(self, pass_infos, rays, hits, halton):
        rays = np.ascontiguousarray(rays, RAY_PACKET_DTYPE)
        hits = np.ascontiguousarray(hits, HIT_DTYPE)
        pis = np.ascontiguousarray(pass_infos, PASS_INFO_DTYPE)
        halton = np.ascontiguousarray(halton, np.float32)
        n = len(rays)
        rgba = np.zeros((n, 4), np.float32)
        sec = np.zeros(max(n, 1), RAY_PACKET_DTYPE)
        src = np.zeros(max(n, 1), np.int32)
        cnt = c_int(0)
        self.dev.check(self.dev.lib.rh_shade_rays(self.h, _p(pis), _p(rays), _p(hits), n, _p(halton), _p(rgba),
                                                  _p(sec), C.byref(cnt), _p(src)), "rh_shade_rays")
        return rgba, sec[:cnt.value], src[:cnt.value]
