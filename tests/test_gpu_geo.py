"""The geometry camera (lightmap baking) and the SH output on the GPU:
k_geo_init / k_geo_raster / k_geo_finish (first covering triangle per texel
by atomic minimum) and k_sh_* against the oracle's sequential restatement of
SampleMeshInTextureSpace (CoreRef.cpp:507-613) and of the SH framebuffer
(FramebufferRef.cpp:48-94), and against the reference Ref backend."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from raydemo_amd import capi as A, scenes
from test_geo_oracle import _ref_render

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gpu(lib, build, w, h, spp, wide=False):
    r = A.Renderer(lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
    s = r.create_scene()
    build(s)
    assert r.sh_data() is None
    reg = A.Region(lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    acc = np.zeros((h, w, 4), np.float32)
    r.read_accum(acc.ctypes.data)
    assert no_overflow(r)
    return r, acc, r.pixels(), r.sh_data()


@pytest.mark.parametrize("perspective,wide", [(False, False), (False, True), (True, True)])
def test_lightmap_and_sh_match_oracle(host_lib, oracle, perspective, wide):
    w, h, spp = 48, 40, 4
    build = lambda s: scenes.lightmap_box(s, k=2, width=w, height=h, spp=spp, depth=3, perspective=perspective)
    r, acc, img, sh = _gpu(host_lib, build, w, h, spp, wide)
    hs = standalone_scene(host_lib, wide=wide)
    build(hs)
    clean, img_o, counts, sh_o = oracle.render_sh(hs.raw_view(), hs.camera_record(0), w, h, spp)
    rc = r.ray_counts()
    assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(c) for c in counts)
    same = np.mean(np.all(acc == clean, axis=-1))
    e = float(np.sqrt(np.mean((acc[..., :3].astype(np.float64) - clean[..., :3]) ** 2)))
    se = float(np.abs(sh - sh_o).max())
    print(f"perspective={perspective} wide={wide}: identical texels {same:.4f}, RMSE {e:.3e}, "
          f"SH max diff {se:.3e} of {np.abs(sh_o).max():.3e}")
    assert e <= 1e-3
    assert se <= 1e-5 * np.abs(sh_o).max()
    covered = np.any(acc[..., :3] != 0, axis=-1).mean()
    assert covered > 0.3


def test_lightmap_and_sh_match_reference(ref_lib, host_lib):
    w, h, spp = 40, 40, 3
    build = lambda s: scenes.lightmap_box(s, k=2, width=w, height=h, spp=spp, depth=3)
    _, _, img, sh = _gpu(host_lib, build, w, h, spp)
    img_r, sh_r = _ref_render(ref_lib, build, w, h, spp)
    assert float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - img_r[..., :3]) ** 2))) <= 1e-3
    assert np.abs(sh - sh_r).max() <= 1e-5 * np.abs(sh_r).max()


def test_sh_cleared_and_batch_refuses_geometry_cameras(host_lib):
    w, h = 24, 24
    build = lambda s: scenes.lightmap_box(s, k=1, width=w, height=h, spp=2, depth=2)
    r, _, _, sh = _gpu(host_lib, build, w, h, 2)
    assert np.abs(sh).max() > 0
    r.clear()
    assert np.abs(r.sh_data()).max() == 0
    s = standalone_scene(host_lib)
    build(s)
    b = A.Batch(host_lib)
    with pytest.raises(RuntimeError):
        b.add(s, w, h, 2)


def test_lightmap_regions_match_oracle(host_lib, oracle):
    """The lightmap baked in two regions (each its own RegionContext): triangle
    boxes clipped to each region, as the reference clips them."""
    w, h, spp = 40, 36, 3
    rects = [(0, 0, 40, 15), (5, 15, 30, 21)]
    build = lambda s: scenes.lightmap_box(s, k=2, width=w, height=h, spp=spp, depth=3)
    r = A.Renderer(host_lib, w, h, A.RENDERER_HIP)
    s = r.create_scene()
    build(s)
    for rect in rects:
        reg = A.Region(host_lib, *rect)
        for _ in range(spp):
            r.render(s, reg)
    acc = np.zeros((h, w, 4), np.float32)
    r.read_accum(acc.ctypes.data)
    hs = standalone_scene(host_lib)
    build(hs)
    clean, _, counts, sh_o = oracle.render_sh_regions(hs.raw_view(), hs.camera_record(0), w, h, spp, rects)
    e = float(np.sqrt(np.mean((acc[..., :3].astype(np.float64) - clean[..., :3]) ** 2)))
    se = float(np.abs(r.sh_data() - sh_o).max())
    print(f"regions: RMSE {e:.3e}, SH max diff {se:.3e}")
    assert e <= 1e-3 and se <= 1e-5 * np.abs(sh_o).max()
    assert np.all(acc[:15, :, :3] >= 0) and np.abs(acc[15:, :5]).max() == 0  # outside both regions


def test_coherent_sampling_matches_oracle_and_reference(ref_lib, host_lib, oracle):
    w, h, spp = 48, 40, 3
    build = lambda s: scenes.cornell_box(s, width=w, height=h, spp=spp, depth=3, coherent=True)
    r = A.Renderer(host_lib, w, h, A.RENDERER_HIP, use_wide_bvh=True)
    s = r.create_scene()
    build(s)
    reg = A.Region(host_lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    assert no_overflow(r)
    img = r.pixels()
    hs = standalone_scene(host_lib, wide=True)
    build(hs)
    _, img_o, _, _ = oracle.render_sh(hs.raw_view(), hs.camera_record(0), w, h, spp)
    img_r, _ = _ref_render(ref_lib, build, w, h, spp, sh=False)
    for ref in (img_o, img_r):
        assert float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - ref[..., :3]) ** 2))) <= 1e-3
