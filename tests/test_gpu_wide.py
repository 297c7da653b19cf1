"""The 8-wide trees on the GPU (settings use_wide_bvh, the reference's
default): the wide traversal kernels against the reference's wide traversals
(CoreRef.cpp:1133-1368) hit for hit, and against the oracle's visit counts;
whole renders against the oracle and the reference Ref backend built with
use_wide_bvh; batches mixing binary and wide scenes; the path-replay adjoint
on a wide scene against the oracle's FP64 gradient."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from helpers import ref_gen_primary, ref_halton, ref_trace, secondary_rays_from_hits
from raydemo_amd import capi as A, scenes
from raydemo_amd.rh import Context
from test_gpu_trace import _check_hits, _instanced

pytestmark = pytest.mark.gpu

# torch joins the process before the backend initialises the device
torch = pytest.importorskip("torch")

BUILDS = {
    "cornell_box": lambda s, g: scenes.cornell_box(s),
    "soup_small": lambda s, g: scenes.triangle_soup(s, n_spheres=48, n_fine=4),
    "mesh_lights": lambda s, g: scenes.mesh_lights(s, g),
    "instanced": lambda s, g: _instanced(s),
}


@pytest.mark.parametrize("knobs", [{}, {"refill": 1}, {"refill": 1, "trace_occupancy": 5}, {"trace_group": 2},
                                   {"trace_group": 4}, {"trace_group": 8, "lds_stack_depth": 4}],
                         ids=["plain", "refill", "refill_occ5", "group2", "group4", "group8_lds4"])
@pytest.mark.parametrize("name", sorted(BUILDS))
def test_wide_trace_bit_exact_and_visits(ref_lib, host_lib, dev_lib, oracle, golden_scene1, name, knobs):
    """Closest-hit and any-hit traces equal the reference library's, and the
    visit counters the oracle's; with refill the closest-hit walk runs as the
    persistent lane-refill kernel (the instrumented and the plain one); with
    trace_group each ray is walked by a group of 2, 4 or 8 lanes."""
    rs, hs = standalone_scene(ref_lib, wide=True), standalone_scene(host_lib, wide=True)
    BUILDS[name](rs, golden_scene1)
    BUILDS[name](hs, golden_scene1)
    w, h = 128, 96
    rays = ref_gen_primary(ref_lib, rs.camera_record(0), (0, 0, w, h), w, h, 1, ref_halton(ref_lib, 1))
    view = hs.raw_view()
    ctx = Context(dev_lib, w=w, h=h)
    for k, v in knobs.items():
        ctx.set_option(k, v)
    ctx.set_option("count_visits", 1)
    ctx.upload(view)
    hr = ref_trace(ref_lib, rs, rays)
    sec = secondary_rays_from_hits(rays, hr, np.random.default_rng(7), n_dirs=2)
    for what, rs_ in (("primary", rays), ("secondary", sec)):
        ctx.reset_visit_counts()
        hg = ctx.trace(rs_)
        nodes, tris, inst, counted = ctx.visit_counts()
        _check_hits(hg, ref_trace(ref_lib, rs, rs_), what)
        _, cnt = oracle.trace(view, rs_, counters=True)
        assert counted == len(rs_)
        assert (nodes, tris, inst) == tuple(int(c) for c in cnt), f"{what}: {(nodes, tris, inst)} vs {cnt}"
        ctx.set_option("count_visits", 0)
        _check_hits(ctx.trace(rs_), ref_trace(ref_lib, rs, rs_), what + " (uninstrumented)")
        ctx.set_option("count_visits", 1)
    tmax = np.full(len(sec), 6.0, np.float32)
    _check_hits(ctx.trace(sec, any_hit=True, tmax=tmax), ref_trace(ref_lib, rs, sec, True, tmax), "any-hit")


def _render(lib, flags, build, w, h, spp, wide=True, **opts):
    r = A.Renderer(lib, w, h, flags, use_wide_bvh=wide)
    for k, v in opts.items():
        r.set_option(k, v)
    s = r.create_scene()
    build(s)
    r.clear()
    reg = A.Region(lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    assert no_overflow(r)
    return r.pixels(), r


def _rmse(a, b):
    return float(np.sqrt(np.mean((a[..., :3].astype(np.float64) - b[..., :3]) ** 2)))


@pytest.mark.parametrize("opts", [{}, {"sort_rays": 1, "lds_stack_depth": 8, "path_tail": -1}, {"defer_nee": 0},
                                  {"path_tail": 0, "tail_occupancy": 3}, {"refill": 1},
                                  {"trace_occupancy": 5, "lds_stack_depth": 8, "path_tail": 1},
                                  {"refill": 1, "trace_occupancy": 6, "lds_stack_depth": 4, "path_tail": -1,
                                   "sort_rays": 1},
                                  {"trace_group": 4, "path_tail": -1},
                                  {"trace_group": 8, "path_tail": 4, "lds_stack_depth": 4, "sort_rays": 1}])
def test_wide_render_matches_oracle(oracle, host_lib, golden_scene1, opts):
    """Image and ray counts of the wide-tree renderer equal the oracle's wide
    render of the same passes."""
    for name, w, h, spp in (("soup_small", 48, 32, 3), ("mesh_lights", 32, 32, 4)):
        build = lambda s: BUILDS[name](s, golden_scene1)
        img, r = _render(host_lib, A.RENDERER_HIP, build, w, h, spp, **opts)
        hs = standalone_scene(host_lib, wide=True)
        build(hs)
        ref_img, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
        rc = r.ray_counts()
        assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(c) for c in counts), name
        e = _rmse(img, ref_img)
        same = np.mean(np.all(img[..., :3] == ref_img[..., :3], axis=-1))
        print(f"{name} {opts}: RMSE vs oracle {e:.3e}, identical pixels {same:.4f}")
        assert e <= 1e-3


def test_wide_cornell_matches_reference(ref_lib, host_lib):
    """The small plain-method workload (Cornell box, 256x256, 16 spp, depth 4)
    against the reference Ref backend, both with their default 8-wide trees."""
    build = lambda s: scenes.cornell_box(s)
    img_g, _ = _render(host_lib, A.RENDERER_HIP, build, 256, 256, 16)
    img_r, _ = _render(ref_lib, A.RENDERER_REF, build, 256, 256, 16)
    e = _rmse(img_g, img_r)
    print(f"cornell wide 16 spp: RMSE {e:.3e}")
    assert e <= 1e-3


def test_batch_of_binary_and_wide_scenes(host_lib, golden_scene1):
    """A batch may mix tree kinds; every scene's image equals its single render."""
    cases = [("soup", True, lambda s: scenes.triangle_soup(s, n_spheres=24, n_fine=2), 96, 64, 3),
             ("cornell", False, lambda s: scenes.cornell_box(s), 64, 48, 4),
             ("mesh_lights", True, lambda s: scenes.mesh_lights(s, golden_scene1), 40, 40, 2)]
    for subset in (cases, [c for c in cases if c[1]]):
        singles = [_render(host_lib, A.RENDERER_HIP, b, w, h, spp, wide=wd)[0] for _, wd, b, w, h, spp in subset]
        batch = A.Batch(host_lib)
        keep = []
        for _, wd, build, w, h, spp in subset:
            s = standalone_scene(host_lib, wide=wd)
            build(s)
            keep.append(s)
            batch.add(s, w, h, spp)
        batch.render_all()
        for i, c in enumerate(subset):
            assert batch.pixels(i).tobytes() == singles[i].tobytes(), c[0]


def test_wide_adjoint_matches_oracle(host_lib, oracle):
    from raydemo_amd import adjoint
    W, H, SPP, DEPTH, K = 32, 24, 3, 3, 2
    r = A.Renderer(host_lib, W, H, A.RENDERER_HIP, use_wide_bvh=True)
    dr, colors, spec = adjoint.refined_cornell(host_lib, k=K, width=W, height=H, spp=SPP, depth=DEPTH, renderer=r)
    wts = np.random.default_rng(11).uniform(-1, 1, size=(H, W, 4)).astype(np.float32)
    wts[..., 3] = 0
    img = dr.render(torch.tensor(colors))
    g = dr.adjoint(torch.tensor(wts, device="cuda")).cpu().numpy().astype(np.float64)
    hs = standalone_scene(host_lib, wide=True)
    scenes.cornell_refined(hs, k=K, width=W, height=H, spp=SPP, depth=DEPTH)
    assert img.cpu().numpy().tobytes() == oracle.render_linear(hs.raw_view(), hs.camera_record(0), W, H, SPP).tobytes()
    ref = oracle.adjoint(hs.raw_view(), hs.camera_record(0), W, H, SPP, wts)
    err = np.abs(g - ref)
    assert err.max() <= 1e-5 * np.abs(ref).max()
