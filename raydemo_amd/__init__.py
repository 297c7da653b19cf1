"""MI355X (gfx950) HIP rendering backend for the Ray path tracer.

The package holds the native libraries (built in-tree by ``raydemo_amd.build``)
and thin Python bindings:

* ``libray_hip.so`` — hand-written HIP kernels behind the C ABI of include/rh_hip.h
* ``libray_hip_host.so`` — the C++ host layer (Ray::Hip::Renderer / Ray::Hip::Scene)
  behind the flat ABI of include/rh_scene.h

Nothing here falls back to a CPU path: if a library is missing, loading fails.
"""
from __future__ import annotations

from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
HIP_LIB_PATH = PKG_DIR / "libray_hip.so"
HOST_LIB_PATH = PKG_DIR / "libray_hip_host.so"

_host = None
_dev = None


def host_lib():
    """The HIP backend's renderer/scene library (RayLib over rhx_*)."""
    global _host
    if _host is None:
        from .capi import RayLib
        if not HOST_LIB_PATH.exists():
            raise RuntimeError(f"{HOST_LIB_PATH} is not built; run `python -m raydemo_amd.build`")
        _host = RayLib(str(HOST_LIB_PATH))
    return _host


def device_lib():
    """The kernel library's C ABI (rh_*), for the per-stage parity entry points."""
    global _dev
    if _dev is None:
        from .rh import DeviceLib
        if not HIP_LIB_PATH.exists():
            raise RuntimeError(f"{HIP_LIB_PATH} is not built; run `python -m raydemo_amd.build`")
        _dev = DeviceLib(str(HIP_LIB_PATH))
    return _dev
