"""The 8-wide trees (settings use_wide_bvh, the reference's default).

Host side: FlattenBVH (Core.cpp:683-819) of every mesh, macro and light tree
must give the reference's wide trees (node for node, bounds bit for bit).
Oracle side: the wide traversals (CoreRef.cpp:1133-1578) and the wide light
walk (CoreRef.cpp:1949-1970) against the reference library built from its
sources, hit for hit and image for image, plus what the mathematics fixes: the
closest hit's t does not depend on which tree found it."""
import numpy as np
import pytest

from conftest import standalone_scene
from helpers import ref_gen_primary, ref_halton, ref_trace, secondary_rays_from_hits
from raydemo_amd import capi as A, scenes

EMPTY = A.MBVH_EMPTY
LEAF = 0x80000000


def _instanced(s):
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


BUILDS = {
    "cornell_box": lambda s, g: scenes.cornell_box(s),
    "soup_small": lambda s, g: scenes.triangle_soup(s, n_spheres=48, n_fine=4),
    "mesh_lights": lambda s, g: scenes.mesh_lights(s, g),
    "instanced": lambda s, g: _instanced(s),
}


def _pair(ref_lib, host_lib, name, golden):
    rs, hs = standalone_scene(ref_lib, wide=True), standalone_scene(host_lib, wide=True)
    BUILDS[name](rs, golden)
    spec = BUILDS[name](hs, golden)
    return rs, hs, spec


def _same_tree(a, ra, b, rb, seen):
    """Wide trees rooted at a[ra] and b[rb] are equal node for node (child
    indices compared through the trees' own structure)."""
    na, nb = a[ra], b[rb]
    assert na["bbox_min"].tobytes() == nb["bbox_min"].tobytes()
    assert na["bbox_max"].tobytes() == nb["bbox_max"].tobytes()
    ca, cb = na["child"], nb["child"]
    seen[0] += 1
    if ca[0] & LEAF:
        assert ca[0] == cb[0] and ca[1] == cb[1]
        return
    for i in range(8):
        assert (ca[i] == EMPTY) == (cb[i] == EMPTY), f"slot {i}"
        if ca[i] != EMPTY:
            _same_tree(a, int(ca[i]), b, int(cb[i]), seen)


@pytest.mark.parametrize("name", sorted(BUILDS))
def test_wide_trees_match_reference(ref_lib, host_lib, golden_scene1, name):
    rs, hs, _ = _pair(ref_lib, host_lib, name, golden_scene1)
    vr, vh = rs.view(), hs.view()
    assert len(vr["nodes"]) == 0 and len(vh["nodes"]) == 0
    a, b = vr["mnodes"], vh["mnodes"]
    assert len(a) == len(b) > 0
    seen = [0]
    for k in range(len(vr["meshes"])):
        _same_tree(a, int(vr["meshes"][k]["node_index"]), b, int(vh["meshes"][k]["node_index"]), seen)
        assert vr["meshes"][k]["node_count"] == vh["meshes"][k]["node_count"]
    _same_tree(a, vr["macro_nodes_root"], b, vh["macro_nodes_root"], seen)
    if vr["light_nodes_root"] != 0xFFFFFFFF:
        _same_tree(a, vr["light_nodes_root"], b, vh["light_nodes_root"], seen)
    assert seen[0] == len(a)  # every node belongs to one of the trees
    # instance bounds come from the wide mesh roots
    assert vr["mesh_instances"].tobytes() == vh["mesh_instances"].tobytes()
    assert vr["mi_indices"].tobytes() == vh["mi_indices"].tobytes()


def test_wide_node_slots_are_morton_ordered(host_lib):
    """A wide node's used slots hold boxes whose centres sit in the octant the
    slot names (or were displaced to the next free slot); unused slots have NaN
    bounds; every level-2 descendant of the binary node appears once."""
    hs = standalone_scene(host_lib, wide=True)
    scenes.triangle_soup(hs, n_spheres=48, n_fine=4)
    mn = hs.view()["mnodes"]
    interior = mn[(mn["child"][:, 0] & LEAF) == 0]
    assert len(interior) > 10
    for n in interior[:200]:
        used = n["child"] != EMPTY
        assert used.sum() >= 2
        assert np.isnan(n["bbox_min"][:, ~used]).all() and np.isnan(n["bbox_max"][:, ~used]).all()
        assert (n["bbox_min"][:, used] <= n["bbox_max"][:, used]).all()


@pytest.mark.parametrize("name", sorted(BUILDS))
def test_wide_trace_matches_reference(oracle, ref_lib, host_lib, golden_scene1, name):
    rs, hs, spec = _pair(ref_lib, host_lib, name, golden_scene1)
    w, h = 64, 48
    rays = ref_gen_primary(ref_lib, rs.camera_record(0), (0, 0, w, h), w, h, 1, ref_halton(ref_lib, 1))
    view = hs.raw_view()
    from test_oracle import same_hits
    hr = ref_trace(ref_lib, rs, rays)
    ho, cnt = oracle.trace(view, rays, counters=True)
    assert same_hits(ho, hr)
    assert cnt[0] > 0
    sec = secondary_rays_from_hits(rays, hr, np.random.default_rng(5), 3)
    assert same_hits(oracle.trace(view, sec), ref_trace(ref_lib, rs, sec))
    tmax = np.full(len(sec), 7.5, np.float32)
    assert same_hits(oracle.trace(view, sec, True, tmax), ref_trace(ref_lib, rs, sec, True, tmax))


def test_wide_and_binary_find_the_same_closest_hit(oracle, host_lib):
    """The closest hit along a ray is a property of the triangles, not of the
    tree: the binary and wide traversals agree on t (and on the primitive,
    except where two triangles tie exactly)."""
    hb, hw = standalone_scene(host_lib), standalone_scene(host_lib, wide=True)
    scenes.triangle_soup(hb, n_spheres=48, n_fine=4)
    scenes.triangle_soup(hw, n_spheres=48, n_fine=4)
    cam = hb.camera_record(0)
    rays = oracle.gen_primary(cam, (0, 0, 96, 64), 96, 64, 1, oracle.halton(1))
    a, b = oracle.trace(hb.raw_view(), rays), oracle.trace(hw.raw_view(), rays)
    assert (a["mask"] == b["mask"]).all()
    assert a["t"].tobytes() == b["t"].tobytes()
    assert (a["prim"] == b["prim"]).mean() > 0.999
    # a wide node stands for up to 7 binary interior nodes: far fewer visits
    _, cb = oracle.trace(hb.raw_view(), rays, counters=True)
    _, cw = oracle.trace(hw.raw_view(), rays, counters=True)
    assert cw[0] * 3 < cb[0]


@pytest.mark.parametrize("name,w,h,spp", [("mesh_lights", 32, 32, 8), ("soup_small", 48, 32, 3),
                                          ("cornell_box", 32, 32, 4)])
def test_wide_render_matches_reference(oracle, ref_lib, host_lib, golden_scene1, name, w, h, spp):
    r = A.Renderer(ref_lib, w, h, A.RENDERER_REF, use_wide_bvh=True)
    s = r.create_scene()
    BUILDS[name](s, golden_scene1)
    reg = A.Region(ref_lib, 0, 0, w, h)
    for _ in range(spp):
        r.render(s, reg)
    hs = standalone_scene(host_lib, wide=True)
    BUILDS[name](hs, golden_scene1)
    img, counts = oracle.render(hs.raw_view(), hs.camera_record(0), w, h, spp)
    assert img.tobytes() == r.pixels().tobytes()
    assert counts[0] == w * h * spp
