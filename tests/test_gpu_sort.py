"""Ray sorting on the GPU (hash keys + stable LSB radix passes) against the
reference's get_ray_hash / SortRays_CPU (src/Ray/internal/CoreRef.cpp:316-333,
615-657): keys bit-exact, permutation identical (a stable sort by key), and
the sort is invisible in the image."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from helpers import ref_ray_hashes, ref_sort_rays
from raydemo_amd import capi as A, scenes
from raydemo_amd.rh import Context

pytestmark = pytest.mark.gpu


def _rays(n, seed, lo=-6.0, hi=12.0, degenerate=False):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, A.RAY_PACKET_DTYPE)
    r["o"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    if degenerate:  # axis-aligned directions and repeated keys
        d[::3] = [0, 0, 1]
        r["o"][::5] = r["o"][0]
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["xy"] = np.arange(n)
    return r


@pytest.mark.parametrize("n,seed,deg", [(1, 0, False), (1000, 1, False), (4096, 5, False), (70001, 2, True),
                                        (300000, 3, False), (1048577, 4, True)])
def test_sort_matches_reference(ref_lib, host_lib, dev_lib, n, seed, deg):
    rs, hs = standalone_scene(ref_lib), standalone_scene(host_lib)
    scenes.cornell_box(rs)
    scenes.cornell_box(hs)
    v = hs.view()
    root = v["nodes"][v["macro_nodes_root"]]
    rmin = root["bbox_min"]
    cell = ((root["bbox_max"] - root["bbox_min"]) / np.float32(255)).astype(np.float32)
    rays = _rays(n, seed, degenerate=deg)
    ctx = Context(dev_lib, w=64, h=64)
    ctx.upload(hs.raw_view())
    k_gpu = ctx.ray_hashes(rays, rmin, cell)
    assert (k_gpu == ref_ray_hashes(ref_lib, rs, rays)).all()
    s_gpu = ctx.sort_rays(rays, rmin, cell)
    s_ref = ref_sort_rays(ref_lib, rs, rays)
    assert (s_gpu["xy"] == s_ref["xy"]).all()
    assert s_gpu.tobytes() == s_ref.tobytes()
    # invariants: permutation, keys non-decreasing, stable
    ks = k_gpu[s_gpu["xy"]]
    assert sorted(s_gpu["xy"].tolist()) == list(range(n))
    assert (np.diff(ks.astype(np.int64)) >= 0).all()
    eq = np.diff(ks.astype(np.int64)) == 0
    assert (np.diff(s_gpu["xy"])[eq] > 0).all()
    assert ctx.get_option("sort_stalls") == 0  # no tile look-back gave up


@pytest.mark.parametrize("lds_depth", [24, 8])
def test_sort_is_invisible_in_the_image(host_lib, lds_depth):
    """Sorting (by the full key or only its top bits), the LDS share of the
    traversal stack, deferred or inline shadow rays, the path-tail launch and
    the primary queue's order (8x8 tiles by default, row-major; 90 rows leave
    a remainder below the tiles) change only the order of the work: the image
    is identical."""
    imgs = []
    for opts in ({"sort_rays": 0, "path_tail": -1}, {"sort_rays": 1, "path_tail": -1},
                 {"sort_rays": 1, "sort_bits": 8}, {"sort_rays": 1, "sort_bits": 16},
                 {"sort_rays": 1, "sort_bits": 24, "path_tail": -1}, {"sort_rays": 1, "sort_bits": 32, "path_tail": 2},
                 {"sort_rays": 0, "defer_nee": 0, "path_tail": -1}, {"path_tail": 0}, {"path_tail": 1, "sort_rays": 1},
                 {"path_tail": 3, "defer_nee": 0}, {"path_tail": 0, "tail_occupancy": 3},
                 {"path_tail": 2, "tail_occupancy": 3, "defer_nee": 0}, {"refill": 1},
                 {"refill": 1, "path_tail": -1, "sort_rays": 1, "lds_stack_depth": 4},
                 {"path_tail": 0, "tail_refill": 1}, {"path_tail": 2, "tail_refill": 1, "tail_occupancy": 3},
                 {"path_tail": 1, "tail_refill": 1, "defer_nee": 0}, {"tiled_primary": 0},
                 {"tiled_primary": 0, "sort_rays": 1, "path_tail": -1}, {"tiled_primary": 0, "refill": 1}):
        r = A.Renderer(host_lib, 160, 90, A.RENDERER_HIP)
        r.set_option("lds_stack_depth", lds_depth)
        for k, v in opts.items():
            r.set_option(k, v)
        s = r.create_scene()
        scenes.triangle_soup(s, n_spheres=64, n_fine=4)
        reg = A.Region(host_lib, 0, 0, 160, 90)
        for _ in range(4):
            r.render(s, reg)
        assert no_overflow(r)
        assert r.get_option("sort_stalls") == 0
        imgs.append(r.pixels())
    for im in imgs[1:]:
        assert im.tobytes() == imgs[0].tobytes()
