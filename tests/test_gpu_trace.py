"""Traversal parity on identical ray sets: the HIP trace kernel against the
reference's Traverse_MacroTree_WithStack_{ClosestHit,AnyHit}
(src/Ray/internal/CoreRef.cpp:1091-1131, 1224-1268) on the same scene, BVH and
rays.  Hit instance/primitive IDs must be bit-exact; t/u/v within 1e-6
relative; an FP64 Moller-Trumbore of every hit states the FP32 error."""
import numpy as np
import pytest

from conftest import standalone_scene
from helpers import (camera_record, moller_trumbore_f64, ref_gen_primary, ref_halton, ref_trace,
                     secondary_rays_from_hits)
from raydemo_amd import capi as A, scenes
from raydemo_amd.rh import Context

pytestmark = pytest.mark.gpu


def _pair(ref_lib, host_lib, build):
    rs, hs = standalone_scene(ref_lib), standalone_scene(host_lib)
    spec = build(rs)
    build(hs)
    return rs, hs, spec


def _check_hits(h_gpu, h_ref, what):
    assert len(h_gpu) == len(h_ref)
    for f in ("mask", "obj", "prim"):
        bad = np.nonzero(h_gpu[f] != h_ref[f])[0]
        assert bad.size == 0, f"{what}: {f} differs on {bad.size} rays, first {bad[:5]}"
    m = h_ref["mask"] != 0
    for f in ("t", "u", "v"):
        a, b = h_gpu[f][m].astype(np.float64), h_ref[f][m].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        small = np.abs(b) < 1e-6
        rel[small] = np.abs(a - b)[small]
        assert rel.max(initial=0) <= 1e-6, f"{what}: {f} max rel err {rel.max()}"
    return m.sum()


@pytest.mark.parametrize("refill", [0, 1])
@pytest.mark.parametrize("scene_name", ["mesh_lights", "cornell_box", "soup_small"])
def test_trace_primary_and_secondary_bit_exact(ref_lib, host_lib, dev_lib, golden_scene1, scene_name, refill):
    build = {
        "mesh_lights": lambda s: scenes.mesh_lights(s, golden_scene1),
        "cornell_box": lambda s: scenes.cornell_box(s),
        "soup_small": lambda s: scenes.triangle_soup(s, n_spheres=48, n_fine=4),
    }[scene_name]
    rs, hs, spec = _pair(ref_lib, host_lib, build)
    w, h = (128, 96) if scene_name == "soup_small" else (spec.width, spec.height)
    cam = rs.camera_record(0)
    halton = ref_halton(ref_lib, 1)
    rays = ref_gen_primary(ref_lib, cam, (0, 0, w, h), w, h, 1, halton)

    ctx = Context(dev_lib, w=w, h=h)
    ctx.set_option("refill", refill)
    ctx.upload(hs.raw_view())
    hg = ctx.trace(rays)
    hr = ref_trace(ref_lib, rs, rays)
    n_hit = _check_hits(hg, hr, "primary")
    assert n_hit > 0

    rng = np.random.default_rng(7)
    sec = secondary_rays_from_hits(rays, hr, rng, n_dirs=2)
    hg2 = ctx.trace(sec)
    hr2 = ref_trace(ref_lib, rs, sec)
    _check_hits(hg2, hr2, "secondary")

    # FP64 oracle: the FP32 t/u/v error of the reported hits
    view = hs.view()
    m = hr2["mask"] != 0
    t64, u64, v64 = moller_trumbore_f64(view, sec[m], hr2[m])
    err_t = np.abs(hg2["t"][m] - t64) / np.maximum(np.abs(t64), 1e-12)
    print(f"{scene_name}: {m.sum()} secondary hits, FP32 vs FP64 t rel err max {err_t.max():.3e} "
          f"median {np.median(err_t):.3e}; u abs err max {np.abs(hg2['u'][m] - u64).max():.3e}")
    assert np.median(err_t) < 1e-5


def test_trace_any_hit_bit_exact(ref_lib, host_lib, dev_lib):
    rs, hs, spec = _pair(ref_lib, host_lib, lambda s: scenes.cornell_box(s))
    rng = np.random.default_rng(11)
    n = 20000
    o = rng.uniform(-4.5, 4.5, size=(n, 3)) + np.array([0, 5, 0])
    p2 = rng.uniform(-4.5, 4.5, size=(n, 3)) + np.array([0, 5, 0])
    d = p2 - o
    dist = np.linalg.norm(d, axis=1)
    rays = np.zeros(n, A.RAY_PACKET_DTYPE)
    rays["o"] = o.astype(np.float32)
    rays["d"] = (d / dist[:, None]).astype(np.float32)
    tmax = dist.astype(np.float32)
    ctx = Context(dev_lib, w=64, h=64)
    ctx.upload(hs.raw_view())
    hg = ctx.trace(rays, any_hit=True, tmax=tmax)
    hr = ref_trace(ref_lib, rs, rays, any_hit=True, tmax=tmax)
    _check_hits(hg, hr, "any-hit")
    assert 0 < (hr["mask"] != 0).mean() < 1


def _instanced(s):
    """Many instances of the Cornell geometry with seeded affine transforms."""
    attrs, idx, shapes, alb = scenes.cornell_geometry()
    tex = s.add_texture(np.full((1, 1, 4), 255, np.uint8))
    m = s.add_material(A.DiffuseMaterial, tex)
    mesh = s.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(m, 0, len(idx))])
    rng = np.random.default_rng(3)
    for k in range(40):
        a = rng.uniform(0, 2 * np.pi)
        sc = rng.uniform(0.2, 0.6)
        x = np.eye(4, dtype=np.float32)
        x[0, 0], x[0, 2], x[2, 0], x[2, 2] = sc * np.cos(a), -sc * np.sin(a), sc * np.sin(a), sc * np.cos(a)
        x[1, 1] = sc
        x[3, :3] = rng.uniform(-15, 15, 3)
        s.add_mesh_instance(mesh, x)
    c = s.add_camera(origin=(0, 5, -40), fwd=(0, -0.1, 1), fov=60.0, filter=A.Box, dtype=A.DeviceNone)
    s.set_current_cam(c)


def test_trace_instanced_scene(ref_lib, host_lib, dev_lib):
    """Two-level traversal: many instances with seeded affine transforms."""
    rs, hs, _ = _pair(ref_lib, host_lib, _instanced)
    cam = rs.camera_record(0)
    rays = ref_gen_primary(ref_lib, cam, (0, 0, 160, 120), 160, 120, 1, ref_halton(ref_lib, 1))
    ctx = Context(dev_lib, w=160, h=120)
    ctx.upload(hs.raw_view())
    hg, hr = ctx.trace(rays), ref_trace(ref_lib, rs, rays)
    _check_hits(hg, hr, "instanced")
    assert len(np.unique(hr["obj"][hr["mask"] != 0])) > 5


@pytest.mark.parametrize("refill", [0, 1])
@pytest.mark.parametrize("scene_name", ["soup_small", "instanced"])
def test_visit_counts_match_oracle(ref_lib, host_lib, dev_lib, oracle, scene_name, refill):
    """The instrumented trace kernel ("count_visits") counts the node visits,
    leaf triangle tests and instance tests of the reference traversal, which
    the roofline's algorithmic bytes are priced from: equal to the oracle's
    instrumented Traverse_MacroTree_WithStack_ClosestHit on the same rays.
    Any-hit: triangle and instance tests equal; nodes at least the oracle's
    (far children tested before an early exit are counted, the reference
    leaves them on its stack)."""
    build = {"soup_small": lambda s: scenes.triangle_soup(s, n_spheres=48, n_fine=4), "instanced": _instanced}
    rs, hs, spec = _pair(ref_lib, host_lib, build[scene_name])
    w, h = 128, 96
    rays = ref_gen_primary(ref_lib, rs.camera_record(0), (0, 0, w, h), w, h, 1, ref_halton(ref_lib, 1))
    hr = ref_trace(ref_lib, rs, rays)
    sec = secondary_rays_from_hits(rays, hr, np.random.default_rng(9), n_dirs=2)
    view = hs.raw_view()
    ctx = Context(dev_lib, w=w, h=h)
    ctx.set_option("refill", refill)
    ctx.set_option("count_visits", 1)
    ctx.upload(view)
    for what, rs_ in (("primary", rays), ("secondary", sec)):
        ctx.reset_visit_counts()
        hg = ctx.trace(rs_)
        nodes, tris, inst, counted = ctx.visit_counts()
        ho, cnt = oracle.trace(view, rs_, counters=True)
        _check_hits(hg, ho, what)
        assert counted == len(rs_)
        assert (nodes, tris, inst) == tuple(int(c) for c in cnt), f"{what}: {(nodes, tris, inst)} vs {cnt}"
    tmax = np.full(len(sec), 6.0, np.float32)
    ctx.reset_visit_counts()
    hg = ctx.trace(sec, any_hit=True, tmax=tmax)
    nodes, tris, inst, _ = ctx.visit_counts()
    ho, cnt = oracle.trace(view, sec, True, tmax, counters=True)
    _check_hits(hg, ho, "any-hit")
    assert (tris, inst) == (int(cnt[1]), int(cnt[2]))
    assert nodes >= int(cnt[0])
