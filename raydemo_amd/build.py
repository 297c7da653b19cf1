"""In-tree build of the HIP rendering backend.

  raydemo_amd/libray_hip.so       gfx950 kernels + the C ABI of include/rh_hip.h   (hipcc)
  raydemo_amd/libray_hip_host.so  C++ host layer: Ray::Hip::Renderer/Scene and the
                                  flat ABI of include/rh_scene.h                  (g++)

Usage: python -m raydemo_amd.build   (or raydemo_amd.build.build_all()).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INC = ROOT / "include"
CSRC = PKG / "csrc"
HOST = PKG / "host"

HIP_LIB = PKG / "libray_hip.so"
HOST_LIB = PKG / "libray_hip_host.so"

# Every float operation rounds separately (no contraction to FMA): the kernels
# reproduce the CPU reference's arithmetic.  Packed-FP32 instructions are kept
# out of the generated code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    return "hipcc"


def build_hip():
    """Each source compiled to its own object (in parallel), then linked."""
    srcs = [CSRC / "rh_kernels.hip", CSRC / "rh_batch.hip", CSRC / "rh_bvh_build.hip"]
    obj_dir = ROOT / "_build"
    obj_dir.mkdir(exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs, objs = [], []
    for src in srcs:
        obj = obj_dir / (src.stem + ".o")
        cmd = [hipcc(), *flags, "-c", f"-I{INC}", f"-I{CSRC}", src, "-o", obj]
        print("+", " ".join(str(c) for c in cmd), flush=True)
        procs.append(subprocess.Popen([str(c) for c in cmd]))
        objs.append(obj)
    failed = [p.wait() for p in procs]
    if any(failed):
        raise subprocess.CalledProcessError(max(failed), "hipcc -c")
    tmp = HIP_LIB.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "--offload-arch=gfx950", *objs, "-o", tmp])
    os.replace(tmp, HIP_LIB)


def build_host():
    srcs = sorted(HOST.glob("*.cpp"))
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run(["g++", *GXX_FLAGS, f"-I{INC}", f"-I{HOST}", *srcs, "-o", tmp, f"-L{PKG}", "-lray_hip",
          "-Wl,-rpath,$ORIGIN", "-lpthread"])
    os.replace(tmp, HOST_LIB)


def build_oracle(with_reference: bool | None = None):
    odir = ROOT / "oracle"
    _run(["make", "-C", odir, "liboracle.so"])
    ref = Path(os.environ.get("RAY_REFERENCE", "/root/reference"))
    if with_reference is None:
        with_reference = (ref / "src" / "Ray" / "_Ray.cpp").exists()
    if with_reference:
        _run(["make", "-C", odir, "-j4", "ref", f"REF={ref}"])


def build_all(oracle: bool = True):
    build_hip()
    build_host()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "oracle": build_oracle}[w]()
