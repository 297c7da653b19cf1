"""Passes in flight on several streams (option pass_overlap: passes in flight
besides the newest, 3 by default): plain RenderScene passes rotate over that
many pass slots plus one, so that a pass's first stages run while the earlier
passes' late bounces drain.  The
accumulated image must be the one passes run one after another give, bit for
bit, whatever entry points are called between the passes (clears, colour and
texel uploads, reads, region passes, resizes, option changes)."""
import numpy as np
import pytest

from raydemo_amd import capi as A, scenes

from conftest import no_overflow, standalone_scene

pytestmark = pytest.mark.gpu

W, H = 160, 96
# the tuned knobs of the 1080p soup: staged bounces 0-1, then the refilling path tail
TAIL = dict(path_tail=2, tail_refill=1, tail_occupancy=2, trace_occupancy=5, lds_stack_depth=8)
STAGED_SORT = dict(path_tail=-1, sort_rays=1, sort_bits=32, lds_stack_depth=16)


def soup(s):
    scenes.triangle_soup(s, width=W, height=H, spp=8, depth=8, n_spheres=64, n_fine=4)


def run(lib, overlap, wide, knobs, passes=8, between=None, build=soup, w=W, h=H):
    r = A.Renderer(lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
    r.set_option("pass_overlap", overlap)
    for k, v in knobs.items():
        r.set_option(k, v)
    s = r.create_scene()
    build(s)
    reg = A.Region(lib, 0, 0, w, h)
    slots = []
    for i in range(passes):
        r.render(s, reg)
        slots.append(r.get_option("bound_slot"))
        if between:
            between(i, r, s, reg)
    img = r.pixels()
    rc = r.ray_counts()
    assert no_overflow(r)
    return img, rc, slots, r


@pytest.mark.parametrize("wide", [False, True], ids=["binary", "wide"])
@pytest.mark.parametrize("knobs", [TAIL, STAGED_SORT], ids=["tail", "staged_sort"])
@pytest.mark.parametrize("overlap", [1, 2, 3])
def test_overlapped_passes_equal_serial_passes(host_lib, wide, knobs, overlap):
    img0, rc0, slots0, _ = run(host_lib, 0, wide, knobs)
    img1, rc1, slots1, _ = run(host_lib, overlap, wide, knobs)
    print(f"slots serial {slots0}, overlapped {slots1}; rays {rc1}")
    assert len(set(slots0)) == 1, "serial passes stay on one slot"
    assert all(a != b for a, b in zip(slots1, slots1[1:])), "overlapped passes alternate slots"
    assert len(set(slots1)) == overlap + 1
    assert np.array_equal(img0.view(np.uint32), img1.view(np.uint32))
    assert rc0 == rc1


@pytest.mark.parametrize("group", [2, 4, 8])
def test_lane_group_trace_equals_lane_trace(host_lib, group):
    """The staged traces of bounce 2 on walked by groups of 2, 4 or 8 lanes per
    ray (trace_group): the image and ray counts of one lane per ray, bit for
    bit, with passes in flight and with a shallow LDS stack whose entries
    spill."""
    for knobs in (STAGED_SORT, dict(TAIL, path_tail=5, lds_stack_depth=4)):
        img0, rc0, _, _ = run(host_lib, 3, True, knobs)
        img1, rc1, _, _ = run(host_lib, 3, True, dict(knobs, trace_group=group))
        assert np.array_equal(img0.view(np.uint32), img1.view(np.uint32))
        assert rc0 == rc1


def test_calls_between_overlapped_passes(host_lib):
    """Every entry point a caller may use between passes: the pass after it
    waits for it, and the image is the serial one."""
    rng = np.random.default_rng(11)
    colors = rng.uniform(0.1, 0.9, size=(5, 3)).astype(np.float32)
    reads = []

    def between(i, r, s, reg):
        if i == 1:
            reads.append(r.pixels())  # a read of the accumulated image
        elif i == 2:
            r.clear((0.25, 0.0, 0.5, 0.0))
        elif i == 3:
            r.set_material_colors(s, 3, colors)
        elif i == 4:
            r.reset_stats()
        elif i == 5:
            r.set_option("lds_stack_depth", 24)  # buffer shapes change at the next pass
        elif i == 6:
            reads.append(r.read_accum())

    img0, rc0, _, _ = run(host_lib, 0, True, TAIL, passes=10, between=between)
    serial_reads = list(reads)
    reads.clear()
    img1, rc1, slots1, _ = run(host_lib, 2, True, TAIL, passes=10, between=between)
    assert all(a != b for a, b in zip(slots1, slots1[1:]))
    for a, b in zip(serial_reads, reads):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(img0.view(np.uint32), img1.view(np.uint32))
    assert rc0 == rc1


def test_region_passes_and_resize(host_lib):
    """Two regions with their own sample counters rendered in turn (the
    reference demo's bucket scheme), then a resize and full-frame passes."""

    def go(overlap):
        r = A.Renderer(host_lib, W, H, A.RENDERER_HIP, use_wide_bvh=True)
        r.set_option("pass_overlap", overlap)
        for k, v in TAIL.items():
            r.set_option(k, v)
        s = r.create_scene()
        soup(s)
        left, right = A.Region(host_lib, 0, 0, W // 2, H), A.Region(host_lib, W // 2, 0, W - W // 2, H)
        for _ in range(4):
            r.render(s, left)
            r.render(s, right)
        first = r.pixels()
        r.resize(W // 2, H // 2)
        full = A.Region(host_lib, 0, 0, W // 2, H // 2)
        for _ in range(5):
            r.render(s, full)
        out = first, r.pixels()
        assert no_overflow(r)
        r.close()
        return out

    a0, b0 = go(0)
    a1, b1 = go(3)
    assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
    assert np.array_equal(b0.view(np.uint32), b1.view(np.uint32))


def test_overlapped_passes_match_oracle(host_lib, oracle):
    """The overlapped image and ray counts against the CPU oracle's on the
    same scene, passes, seeds and 8-wide trees."""
    w, h, spp = 64, 48, 4
    build = lambda s: scenes.triangle_soup(s, width=w, height=h, spp=spp, depth=8, n_spheres=64, n_fine=4)
    img, rc, slots, _ = run(host_lib, 3, True, TAIL, passes=spp, w=w, h=h, build=build)
    assert all(a != b for a, b in zip(slots, slots[1:]))
    hs = standalone_scene(host_lib, wide=True)
    build(hs)
    ref, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
    e = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - ref[..., :3]) ** 2)))
    print(f"overlapped soup {w}x{h} {spp} spp vs oracle: RMSE {e:.2e}")
    assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(c) for c in counts)
    assert e <= 1e-3
