"""The fast BVH build on the device (rh_build_lbvh, raydemo_amd/csrc/
rh_bvh_build.hip): a scene created by the HIP renderer builds its
use_fast_bvh_build meshes' treelets on the GPU; its arrays equal the host
build's (a standalone scene), which equals the reference's PreprocessPrims_HLBVH
(tests/test_hlbvh.py), byte for byte; renders of fast-built scenes equal the
oracle's and the reference Ref backend's."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from raydemo_amd import capi as A, scenes
from test_hlbvh import CASES, _mesh_scene, _random_tris

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _device_scene(lib, wide=False):
    r = A.Renderer(lib, 16, 16, A.RENDERER_HIP, use_wide_bvh=wide)
    return r, r.create_scene()


def _same_arrays(dv, hv):
    for k in ("tris", "tri_indices", "meshes", "nodes"):
        assert dv[k].tobytes() == hv[k].tobytes(), k


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_fast_build_equals_host(host_lib, name):
    attrs, idx = CASES[name]
    r, ds = _device_scene(host_lib)
    hs = standalone_scene(host_lib)
    assert _mesh_scene(ds, attrs, idx) != A.INVALID
    _mesh_scene(hs, attrs, idx)
    print(f"{name}: {len(idx) // 3} triangles, {len(hs.view()['nodes'])} nodes")
    _same_arrays(ds.view(), hs.view())


def test_device_fast_build_many_meshes_and_big_mesh(host_lib):
    """Several meshes in one scene (leaf indices offset by the earlier meshes'
    triangles; the builder's buffers grow and shrink back) and a 600k-triangle
    mesh with thousands of treelets."""
    rng = np.random.default_rng(77)
    meshes = [_random_tris(rng, n) for n in (3000, 600000, 50, 20000)]
    meshes.append(_random_tris(rng, 4000, flat_axis=2))
    r, ds = _device_scene(host_lib)
    hs = standalone_scene(host_lib)
    for attrs, idx in meshes:
        _mesh_scene(ds, attrs, idx)
        _mesh_scene(hs, attrs, idx)
    _same_arrays(ds.view(), hs.view())


@pytest.mark.parametrize("wide", [False, True])
def test_fast_built_soup_renders_as_reference(ref_lib, host_lib, oracle, wide):
    """The 100k-triangle soup with fast-built trees: GPU image and ray counts
    equal the oracle's render of the same (host-built) scene; the reference
    Ref backend's image within RMSE 1e-3."""
    w, h, spp = 64, 48, 3
    build = lambda s: scenes.triangle_soup(s, n_spheres=256, n_fine=20, fast_bvh=True)
    r = A.Renderer(host_lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
    s = r.create_scene()
    build(s)
    reg = A.Region(host_lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    assert no_overflow(r)
    img = r.pixels()
    hs = standalone_scene(host_lib, wide=wide)
    build(hs)
    _same_arrays(s.view(), hs.view())
    ref_img, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
    rc = r.ray_counts()
    assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(c) for c in counts)
    assert float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - ref_img[..., :3]) ** 2))) <= 1e-3
    rr = A.Renderer(ref_lib, w, h, A.RENDERER_REF, use_wide_bvh=wide)
    rsc = rr.create_scene()
    build(rsc)
    rreg = A.Region(ref_lib, 0, 0, w, h)
    for _ in range(spp):
        rr.render(rsc, rreg)
    e = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - rr.pixels()[..., :3]) ** 2)))
    print(f"wide={wide}: RMSE vs reference {e:.3e}")
    assert e <= 1e-3


def test_device_fast_bvh_over_boxes(host_lib):
    """rhx_build_fast_bvh on the device equals the host build on 1M random
    boxes (some degenerate: points, repeated boxes, a flat slab)."""
    rng = np.random.default_rng(123)
    n = 1_000_000
    lo = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    ext = rng.exponential(0.5, (n, 3)).astype(np.float32)
    ext[: n // 100] = 0.0                      # points
    lo[n // 100: n // 50] = lo[n // 100]       # repeated boxes
    ext[n // 100: n // 50] = ext[n // 100]
    lo[n // 50: n // 20, 1] = 3.0              # a slab
    ext[n // 50: n // 20, 1] = 0.0
    boxes = np.concatenate([lo, lo + ext], axis=1)
    hn, hi = A.build_fast_bvh(host_lib, boxes)
    dn, di = A.build_fast_bvh(host_lib, boxes, device=0)
    print(f"{n} boxes: {len(hn)} nodes")
    assert hn.tobytes() == dn.tobytes() and hi.tobytes() == di.tobytes()


# ---- the SAH build on the device (rh_build_sah): instance tree and fast-build top tree

def _distinct_boxes(rng, n, spread=100.0):
    """n random boxes whose bbox_max coordinates differ along every axis (the
    device build's precondition; with equal keys the host builds)."""
    while True:
        lo = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
        ext = rng.exponential(spread / max(1.0, np.cbrt(n)) / 2, (n, 3)).astype(np.float32)
        hi = lo + ext
        if all(len(np.unique(hi[:, a])) == n for a in range(3)):
            return np.concatenate([lo, hi], axis=1)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 4096, 4097, 9000])
def test_device_sah_equals_host(host_lib, n):
    """Node for node and index for index the host's PreprocessPrims_SAH (the
    LDS sort up to 4096 boxes a node, global scratch above)."""
    boxes = _distinct_boxes(np.random.default_rng(1000 + n), n)
    hn, hi, h_dev = A.build_sah(host_lib, boxes)
    dn, di, d_dev = A.build_sah(host_lib, boxes, device=0)
    assert not h_dev and d_dev
    print(f"{n} boxes: {len(hn)} nodes")
    assert hn.tobytes() == dn.tobytes() and hi.tobytes() == di.tobytes()


def test_device_sah_top_tree_settings(host_lib):
    """The fast build's top tree settings (oversplit FLT_MAX, traversal cost
    0: every box its own leaf) and a clustered set of boxes."""
    rng = np.random.default_rng(5)
    boxes = _distinct_boxes(rng, 700)
    boxes[:350, :3] *= 0.01
    boxes[:350, 3:] = boxes[:350, :3] + np.abs(boxes[:350, 3:] - boxes[:350, :3]) * 0.01
    assert all(len(np.unique(boxes[:, 3 + a])) == 700 for a in range(3))
    kw = dict(oversplit=3.4028234e38, cost=0.0)
    hn, hi, _ = A.build_sah(host_lib, boxes, **kw)
    dn, di, d_dev = A.build_sah(host_lib, boxes, device=0, **kw)
    assert d_dev and hn.tobytes() == dn.tobytes() and hi.tobytes() == di.tobytes()
    leaves = hn[(hn["prim_index"] & 0x80000000) != 0]
    assert len(leaves) == 700 and (leaves["prim_count"] == 1).all()


def test_device_sah_ties_build_on_host(host_lib):
    """Equal bbox_max keys: the device declines and the host's tree is used."""
    boxes = _distinct_boxes(np.random.default_rng(9), 300)
    boxes[10:20, 4] = boxes[10, 4]
    hn, hi, _ = A.build_sah(host_lib, boxes)
    dn, di, d_dev = A.build_sah(host_lib, boxes, device=0)
    assert not d_dev
    assert hn.tobytes() == dn.tobytes() and hi.tobytes() == di.tobytes()


def _instanced_scene(s, rng, n_inst):
    attrs, idx = _random_tris(rng, 400)
    mesh = _mesh_scene(s, attrs, idx)
    for _ in range(n_inst):
        x = np.eye(4, dtype=np.float32)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x[:3, :3] = (q * rng.uniform(0.3, 2.0)).astype(np.float32)
        x[3, :3] = rng.uniform(-200, 200, 3)
        s.add_mesh_instance(mesh, x)


@pytest.mark.parametrize("wide", [False, True])
def test_instance_tree_built_on_device_equals_host(host_lib, wide):
    """A renderer's scene builds its instance tree (RebuildMacroBVH) on the
    device: 3000 instances give the standalone (host-built) scene's arrays
    byte for byte, and the device does build this set (no tied keys)."""
    r, ds = _device_scene(host_lib, wide)
    hs = standalone_scene(host_lib, wide)
    for s in (ds, hs):
        _instanced_scene(s, np.random.default_rng(31), 3000)
    dv, hv = ds.view(), hs.view()
    for k in ("nodes", "mnodes", "mi_indices", "mesh_instances", "meshes"):
        assert dv[k].tobytes() == hv[k].tobytes(), k
    mi = hv["mesh_instances"]
    boxes = np.concatenate([mi["bbox_min"], mi["bbox_max"]], axis=1)
    _, _, on_device = A.build_sah(host_lib, boxes, device=0, oversplit=0.95, cost=0.025)
    assert on_device
