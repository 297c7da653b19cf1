"""A scene read from the demo's formats (raydemo_amd/load.py: JSON, .bin
meshes, a PNG texture, an HDR environment map, every material type, a sun)
rendered on the GPU against the oracle and against the reference Ref backend
fed by the same loader."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from raydemo_amd import capi as A, load
from test_load import write_scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("wide", [False, True])
def test_loaded_scene_renders_as_oracle_and_reference(tmp_path, ref_lib, host_lib, oracle, wide):
    path = write_scene(str(tmp_path))
    w, h, spp = 64, 48, 4
    r = A.Renderer(host_lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
    s = r.create_scene()
    load.load_scene(s, path, base_dir=str(tmp_path))
    reg = A.Region(host_lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    assert no_overflow(r)
    img = r.pixels()
    hs = standalone_scene(host_lib, wide=wide)
    load.load_scene(hs, path, base_dir=str(tmp_path))
    ref_img, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
    rc = r.ray_counts()
    assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(c) for c in counts)
    e = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - ref_img[..., :3]) ** 2)))
    rr = A.Renderer(ref_lib, w, h, A.RENDERER_REF, use_wide_bvh=wide)
    rs = rr.create_scene()
    load.load_scene(rs, path, base_dir=str(tmp_path))
    rreg = A.Region(ref_lib, 0, 0, w, h)
    for _ in range(spp):
        rr.render(rs, rreg)
    e2 = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - rr.pixels()[..., :3]) ** 2)))
    print(f"wide={wide}: RMSE vs oracle {e:.3e}, vs reference {e2:.3e}, mean {img[..., :3].mean():.3f}")
    assert e <= 1e-3 and e2 <= 1e-3
    assert img[..., :3].mean() > 0.01
