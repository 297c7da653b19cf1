"""The C-ABI libraries load on a machine without a GPU and export every
function their headers declare (no compute calls are made here)."""
import ctypes as C
import re

import pytest

from conftest import ROOT


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rhx?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("lib,header", [("raydemo_amd/libray_hip.so", "rh_hip.h"),
                                        ("raydemo_amd/libray_hip_host.so", "rh_scene.h")])
def test_library_exports_declared_symbols(lib, header):
    path = ROOT / lib
    assert path.exists(), f"{lib} not built (python -m raydemo_amd.build)"
    so = C.CDLL(str(path), mode=C.RTLD_LOCAL)
    names = declared(header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_no_device_fails_loudly(host_lib):
    """Without a usable device the HIP renderer refuses to be created (the
    reference factory's OCL branch throws and falls back the same way,
    RendererFactory.cpp:26-35); it never silently renders on the CPU."""
    import os
    from raydemo_amd import capi as A
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        A.Renderer(host_lib, 16, 16, A.RENDERER_HIP)
    with pytest.raises(RuntimeError):
        A.Renderer(host_lib, 16, 16, A.RENDERER_REF)
