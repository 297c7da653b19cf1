"""AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU code: the C++
host layer (scene preprocessing, every BVH build, texture atlas and mips,
light tree, scene edits) and the CPU oracle (generate, trace, hash and sort,
render, SH render, adjoint), built with g++ -fsanitize=address,undefined by
tests/sanitize/Makefile and driven by tests/sanitize/host_oracle_driver.cpp
over standalone scenes (binary and 8-wide trees).  Any report, leak included,
fails the run."""
import shutil
import subprocess

import pytest

from conftest import ROOT

SAN = ROOT / "tests" / "sanitize"


def _sanitizer_runtime() -> bool:
    if not shutil.which("g++"):
        return False
    res = subprocess.run(["g++", "-fsanitize=address,undefined", "-x", "c++", "-", "-o", "/dev/null"],
                         input="int main() { return 0; }\n", capture_output=True, text=True)
    return res.returncode == 0


@pytest.mark.timeout(900)
def test_host_layer_and_oracle_are_clean_under_asan_ubsan():
    if not (ROOT / "raydemo_amd" / "libray_hip.so").exists():
        pytest.skip("raydemo_amd/libray_hip.so not built")
    if not _sanitizer_runtime():
        pytest.skip("g++ sanitizer runtime not available")
    res = subprocess.run(["make", "-s", "-C", str(SAN), "run"], capture_output=True, text=True)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-6000:]
    assert "ERROR: AddressSanitizer" not in out and "ERROR: LeakSanitizer" not in out, out[-6000:]
    assert "runtime error:" not in out, out[-6000:]
    assert "sanitized host layer and oracle: OK" in res.stdout
