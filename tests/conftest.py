"""Shared fixtures.  GPU tests are marked ``gpu``; CPU tests cover the host
scene build, the CPU oracle against the golden vectors and the reference
library, and that the C-ABI libraries load and export every declared symbol."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
REF_LIB = ROOT / "oracle" / "_ref" / "libray_ref.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ref_lib():
    """The reference Ray library (CPU Ref backend) built from its own sources."""
    if not REF_LIB.exists():
        pytest.skip("reference library oracle/_ref/libray_ref.so not built")
    from raydemo_amd.capi import RayLib
    lib = RayLib(str(REF_LIB))
    import ctypes as C
    L = lib.lib
    P = C.c_void_p
    for name, res, args in [
        ("ref_halton_table", C.c_int, [C.c_int, P]),
        ("ref_halton_perms", C.c_int, [P, C.c_size_t]),
        ("ref_gen_primary", C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P]),
        ("ref_trace", C.c_int, [P, P, C.c_size_t, C.c_int, P, P]),
        ("ref_ray_hashes", C.c_int, [P, P, C.c_size_t, P]),
        ("ref_sort_rays", C.c_int, [P, P, C.c_size_t]),
        ("ref_shade", C.c_int, [P, P, P, P, C.c_size_t, P, P, P, P, P]),
        ("ref_sample_bilinear", C.c_int, [P, C.c_uint32, P, C.c_int, C.c_size_t, P]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return lib


@pytest.fixture(scope="session")
def host_lib():
    from raydemo_amd import host_lib as hl
    return hl()


@pytest.fixture(scope="session")
def dev_lib():
    from raydemo_amd import device_lib
    return device_lib()


@pytest.fixture(scope="session")
def golden_scene1():
    d = np.load(GOLDEN / "scene1.npz")
    return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def golden_scene2():
    d = np.load(GOLDEN / "scene2.npz")
    return {k: d[k] for k in d.files}


def standalone_scene(lib, wide=False):
    from raydemo_amd.capi import standalone_scene as make
    return make(lib, wide)


@pytest.fixture(scope="session")
def oracle():
    if not ORACLE_LIB.exists():
        pytest.skip("CPU oracle oracle/liboracle.so not built")
    from helpers import Oracle
    return Oracle(ORACLE_LIB)


def no_overflow(renderer_or_batch) -> bool:
    """True when no traversal stack push was dropped (a dropped push is a
    possibly missed hit): every GPU image comparison checks it.  The reference
    library's renderers have no device counters (an empty dict) and pass."""
    rc = renderer_or_batch.ray_counts()
    return not rc or rc["overflow"] == 0
