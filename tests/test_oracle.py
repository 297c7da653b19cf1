"""Pinning the CPU oracle (oracle/ray_oracle.cpp).

Each oracle function is checked against (a) the reference's own golden data
(tests/golden/, from the reference's tests), (b) the reference library built
from its own sources (oracle/_ref), and (c) closed forms / invariants that a
plausible restatement mistake would break."""
import json
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN, standalone_scene
from helpers import (camera_record, pass_infos_for, ref_gen_primary, ref_halton, ref_perms, ref_ray_hashes,
                     ref_shade, ref_sort_rays, ref_trace, secondary_rays_from_hits)
from raydemo_amd import capi as A, scenes

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89]


def test_halton_permutations_are_permutations(oracle):
    p = oracle.perms()
    assert len(p) == sum(PRIMES) == 963
    off = 0
    for b in PRIMES:
        assert sorted(p[off:off + b].tolist()) == list(range(b))
        off += b


def test_halton_matches_reference(oracle, ref_lib):
    assert (oracle.perms() == ref_perms(ref_lib)).all()
    for it in (1, 2, 255, 256, 1000):
        assert oracle.halton(it).tobytes() == ref_halton(ref_lib, it).tobytes()


def test_halton_closed_form(oracle):
    """Scrambled radical inverse of a with permutation p in base b equals
    sum_k p[d_k] b^-(k+1) + p[0] b^-(n+1) / (1 - 1/b) (exact rational), to float rounding."""
    p = oracle.perms()
    for it in (1, 37, 500):
        tab = oracle.halton(it).reshape(256, 24)
        for i in (0, 5, 255):
            a = it + i
            off = 0
            for d, b in enumerate(PRIMES):
                perm = p[off:off + b]
                off += b
                x, k, acc = a, 0, Fraction(0)
                while x:
                    acc += Fraction(int(perm[x % b]), b ** (k + 1))
                    x //= b
                    k += 1
                acc += Fraction(int(perm[0]), b ** (k + 1)) / (1 - Fraction(1, b))
                exact = min(float(acc), 1 - 2 ** -23)
                assert abs(tab[i, d] - exact) <= 4e-7 * max(1.0, exact), (it, i, d)


def test_primary_rays_golden_vectors(oracle, host_lib):
    g = json.loads((GOLDEN / "primary_ray_gen.json").read_text())
    cam = camera_record(host_lib, origin=(0, 0, 4), fwd=(0, 0, -1), up=(0, 0, 0), fov=53.13, filter=A.Box,
                        dtype=A.DeviceNone, gamma=2.2)
    rays = oracle.gen_primary(cam, (0, 0, 4, 4), 4, 4, 0, np.zeros(24 * 256, np.float32))
    exp = np.array(g["rays"], np.float64)
    for i in range(16):
        r = rays[i]
        got = np.array([*r["o"], *r["d"], *r["c"], r["ior"], *r["do_dx"], *r["dd_dx"], *r["do_dy"], *r["dd_dy"]])
        assert int(r["xy"]) == int(exp[i, 0])
        assert np.abs(got - exp[i, 1:]).max() < g["tolerance_abs"]


def test_primary_rays_match_reference(oracle, ref_lib, host_lib):
    for filt, ff in ((A.Box, 0.0), (A.Tent, 0.25)):
        cam = camera_record(host_lib, origin=(1, 2, -7), fwd=(0.1, -0.2, 1), fov=55.0, filter=filt, focus_factor=ff)
        hal = ref_halton(ref_lib, 9)
        a = oracle.gen_primary(cam, (3, 2, 40, 30), 64, 48, 9, hal)
        b = ref_gen_primary(ref_lib, cam, (3, 2, 40, 30), 64, 48, 9, hal)
        assert a.tobytes() == b.tobytes()


def same_hits(a, b):
    """Bit-equal hits; u/v of misses are unset in the reference and not compared."""
    if len(a) != len(b):
        return False
    for f in ("mask", "obj", "prim", "t", "xy"):
        if a[f].tobytes() != b[f].tobytes():
            return False
    m = b["mask"] != 0
    return a["u"][m].tobytes() == b["u"][m].tobytes() and a["v"][m].tobytes() == b["v"][m].tobytes()


def _pair(ref_lib, host_lib, build):
    rs, hs = standalone_scene(ref_lib), standalone_scene(host_lib)
    build(rs)
    build(hs)
    return rs, hs


def test_trace_matches_reference(oracle, ref_lib, host_lib):
    rs, hs = _pair(ref_lib, host_lib, lambda s: scenes.triangle_soup(s, n_spheres=24, n_fine=2))
    cam = rs.camera_record(0)
    rays = ref_gen_primary(ref_lib, cam, (0, 0, 64, 40), 64, 40, 1, ref_halton(ref_lib, 1))
    view = hs.raw_view()
    h_o, cnt = oracle.trace(view, rays, counters=True)
    h_r = ref_trace(ref_lib, rs, rays)
    assert same_hits(h_o, h_r)
    assert cnt[0] > len(rays)
    sec = secondary_rays_from_hits(rays, h_r, np.random.default_rng(5), 3)
    assert same_hits(oracle.trace(view, sec), ref_trace(ref_lib, rs, sec))
    tmax = np.full(len(sec), 7.5, np.float32)
    assert same_hits(oracle.trace(view, sec, True, tmax), ref_trace(ref_lib, rs, sec, True, tmax))


def test_trace_brute_force_closest_hit(oracle, host_lib):
    """Closest hit through the BVH equals a brute-force loop over all triangles
    (any visit order gives the same closest t unless two triangles tie)."""
    hs = standalone_scene(host_lib)
    scenes.cornell_box(hs)
    view = hs.raw_view()
    v = hs.view()
    rng = np.random.default_rng(1)
    n = 3000
    rays = np.zeros(n, A.RAY_PACKET_DTYPE)
    rays["o"] = rng.uniform(-4, 4, (n, 3)).astype(np.float32) + np.float32([0, 5, 0])
    d = rng.standard_normal((n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    hits = oracle.trace(view, rays)
    # brute force in float64 with a plain Moller-Trumbore over every triangle
    P = v["vertices"]["p"].astype(np.float64)[v["vtx_indices"]].reshape(-1, 3, 3)
    o, dd = rays["o"].astype(np.float64), rays["d"].astype(np.float64)
    e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    best_t = np.full(n, np.inf)
    best = np.full(n, -1)
    for k in range(len(P)):
        pv = np.cross(dd, e2[k])
        det = pv @ e1[k]
        tv = o - P[k, 0]
        u = np.einsum("ij,ij->i", tv, pv) / det
        qv = np.cross(tv, e1[k])
        vv = (dd * qv).sum(1) / det
        t = qv @ e2[k] / det
        ok = (u >= 0) & (vv >= 0) & (u + vv <= 1) & (t > 0) & (t < best_t)
        best_t[ok] = t[ok]
        best[ok] = k
    m = best >= 0
    assert (hits["mask"] != 0)[m].all()
    rel = np.abs(hits["t"][m] - best_t[m]) / best_t[m]
    # FP32 projected-plane test vs FP64: worst cases are grazing rays
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4
    # a different triangle is only acceptable at a tie (coplanar faces, e.g. box bottoms on the floor)
    diff = m & (hits["prim"] != best)
    assert np.allclose(hits["t"][diff], best_t[diff], rtol=1e-6)
    assert diff.mean() < 0.01


def test_ray_hash_and_sort_match_reference(oracle, ref_lib, host_lib):
    rs, hs = _pair(ref_lib, host_lib, lambda s: scenes.cornell_box(s))
    v = hs.view()
    root = v["nodes"][v["macro_nodes_root"]]
    rmin = root["bbox_min"]
    cell = (root["bbox_max"] - root["bbox_min"]) / np.float32(255)
    rng = np.random.default_rng(2)
    n = 5000
    rays = np.zeros(n, A.RAY_PACKET_DTYPE)
    rays["o"] = rng.uniform(-6, 12, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["xy"] = np.arange(n)
    k_o = oracle.ray_hashes(rays, rmin, cell)
    assert (k_o == ref_ray_hashes(ref_lib, rs, rays)).all()
    s_o = oracle.sort_rays(rays, rmin, cell)
    s_r = ref_sort_rays(ref_lib, rs, rays)
    assert s_o.tobytes() == s_r.tobytes()
    # invariants: a permutation, keys non-decreasing, stable among equal keys
    assert sorted(s_o["xy"].tolist()) == list(range(n))
    ks = k_o[s_o["xy"]]
    assert (np.diff(ks.astype(np.int64)) >= 0).all()
    eq = np.diff(ks.astype(np.int64)) == 0
    assert (np.diff(s_o["xy"])[eq] > 0).all()


def test_shade_matches_reference(oracle, ref_lib, host_lib):
    rs, hs = _pair(ref_lib, host_lib, lambda s: scenes.triangle_soup(s, n_spheres=24, n_fine=2))
    cam = rs.camera_record(0)
    w, h = 64, 40
    hal = ref_halton(ref_lib, 3)
    rays = ref_gen_primary(ref_lib, cam, (0, 0, w, h), w, h, 3, hal)
    hits = ref_trace(ref_lib, rs, rays)
    st = dict(max_diff_depth=8, max_glossy_depth=8, max_refr_depth=8, max_transp_depth=8, max_total_depth=8,
              termination_start_depth=3)
    pis = pass_infos_for(rays, w, 3, 2, st)
    a = oracle.shade(hs.raw_view(), pis, rays, hits, hal)
    b = ref_shade(ref_lib, rs, pis, rays, hits, hal)
    assert a[0].tobytes() == b[0].tobytes()
    assert a[1].tobytes() == b[1].tobytes() and (a[2] == b[2]).all()


def test_render_matches_reference(oracle, ref_lib, host_lib, golden_scene1):
    for build, w, h, spp in ((lambda s: scenes.mesh_lights(s, golden_scene1), 32, 32, 8),
                             (lambda s: scenes.triangle_soup(s, n_spheres=24, n_fine=2), 48, 32, 3)):
        r = A.Renderer(ref_lib, w, h, A.RENDERER_REF, use_wide_bvh=False)
        s = r.create_scene()
        build(s)
        reg = A.Region(ref_lib, 0, 0, w, h)
        for _ in range(spp):
            r.render(s, reg)
        hs = standalone_scene(host_lib)
        build(hs)
        img, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
        assert img.tobytes() == r.pixels().tobytes()
        assert counts[0] == w * h * spp
