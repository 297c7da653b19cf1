"""The traversal visit counters that price the roofline's algorithmic bytes,
checked on a tree whose shape is known: one instance of a two-triangle mesh
(the triangles 20 apart along x), so the mesh's binary tree is a root over two
one-triangle leaves and the macro tree a single leaf.  For each ray the visits
are counted by hand from the walk's rules and compared with the oracle's
counters (orc_trace), which the device's instrumented kernels equal on the
same rays (tests/test_gpu_trace.py, tests/test_gpu_wide.py).

Rules (binary trees): the macro root's box is tested (1 visit); an interior
node tests both children's boxes (2 visits) and a popped far child is not
counted again; each instance of a reached macro leaf is one instance test, and
a hit instance's mesh root is one visit.  8-wide trees: each interior wide
node entered is one visit (its 8 slot tests included); a leaf root is entered
without a test."""
import numpy as np
import pytest

from conftest import standalone_scene
from raydemo_amd import capi as A


def two_triangle_scene(lib, wide):
    s = standalone_scene(lib, wide)
    tex = s.add_texture(np.full((1, 1, 4), 255, np.uint8), False, False)
    m = s.add_material(A.DiffuseMaterial, tex)
    P = np.array([[-11, 0, -1], [-9, 0, -1], [-10, 1, 1], [9, 0, -1], [11, 0, -1], [10, 1, 1]], np.float32)
    attrs = np.concatenate([P, np.tile([0, 1, 0], (6, 1)), np.zeros((6, 2))], 1).astype(np.float32)
    mesh = s.add_mesh(A.PxyzNxyzTuv, attrs, np.arange(6, dtype=np.uint32), [(m, 0, 6)])
    s.add_mesh_instance(mesh)
    return s


RAYS = [  # origin, direction, hit triangle, (binary nodes, tris, instances), (wide nodes, tris, instances)
    ((-10.0, 5.0, 0.5), (0.0, -1.0, 0.0), 0, (4, 1, 1), (1, 1, 1)),    # down onto the first triangle
    ((10.0, 5.0, 0.5), (0.0, -1.0, 0.0), 1, (4, 1, 1), (1, 1, 1)),     # down onto the second
    ((0.0, 5.0, 0.0), (0.0, 1.0, 0.0), -1, (1, 0, 0), (0, 0, 1)),      # away from everything
    ((0.0, 0.5, 5.0), (0.0, 0.0, -1.0), -1, (4, 0, 1), (1, 0, 1)),     # between the leaves' boxes
    ((-20.0, 0.5, 0.5), (1.0, 0.0, 0.0), -1, (4, 2, 1), (1, 2, 1)),    # through both boxes, parallel to both
]


@pytest.mark.parametrize("wide", [False, True])
def test_visit_counts_by_hand(oracle, host_lib, wide):
    s = two_triangle_scene(host_lib, wide)
    v = s.view()
    if wide:
        assert len(v["mnodes"]) == 4  # the mesh's root, its two leaves, the macro tree's leaf root
    else:
        root = v["nodes"][v["meshes"][0]["node_index"]]
        assert not root["prim_index"] & 0x80000000
        for c in (root["prim_index"], root["prim_count"] & 0x3FFFFFFF):
            assert v["nodes"][c]["prim_index"] & 0x80000000 and v["nodes"][c]["prim_count"] == 1
        assert v["nodes"][v["macro_nodes_root"]]["prim_index"] & 0x80000000
    view = s.raw_view()
    for o, d, prim, binary, wide_counts in RAYS:
        ray = np.zeros(1, A.RAY_PACKET_DTYPE)
        ray["o"], ray["d"] = o, d
        hit, cnt = oracle.trace(view, ray, counters=True)
        assert int(hit["prim"][0]) == prim, (o, d)
        assert tuple(int(c) for c in cnt) == (wide_counts if wide else binary), (o, d)
