"""Indices a caller passes the scene are checked when they come in: the
reference stores them as they are (SceneRef.cpp:118-148 AddMaterial,
SceneRef.cpp:156-263 AddMesh) and its renderers read through them, so an index
that names nothing reads outside the scene's arrays.  Here the adder fails
with a message and leaves the scene as it was."""
import numpy as np
import pytest

from conftest import standalone_scene
from raydemo_amd import capi as A


def _tri():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    return np.concatenate([P, np.tile([0, 0, 1], (3, 1)), np.zeros((3, 2))], 1).astype(np.float32)


@pytest.fixture
def scene(host_lib):
    s = standalone_scene(host_lib)
    tex = s.add_texture(np.full((2, 2, 4), 200, np.uint8), False, False)
    mat = s.add_material(A.DiffuseMaterial, tex)
    return s, tex, mat


def test_material_indices_are_checked(scene):
    s, tex, mat = scene
    n = s.view()["materials"].shape[0]
    with pytest.raises(RuntimeError, match="texture"):
        s.add_material(A.DiffuseMaterial, A.INVALID)
    with pytest.raises(RuntimeError, match="texture"):
        s.add_material(A.DiffuseMaterial, tex + 5)
    with pytest.raises(RuntimeError, match="normal_map"):
        s.add_material(A.DiffuseMaterial, tex, normal_map=tex + 1)
    with pytest.raises(RuntimeError, match="mix_materials"):
        s.add_material(A.MixMaterial, tex, mix_materials=(mat, A.INVALID))
    with pytest.raises(RuntimeError, match="material type"):
        s.add_material(17, tex)
    assert s.view()["materials"].shape[0] == n
    # the sentinel stays allowed where it means something: no normal map
    assert s.add_material(A.GlossyMaterial, tex, normal_map=A.INVALID) == n


def test_mesh_indices_are_checked(scene):
    s, _, mat = scene
    attrs = _tri()
    tris = s.triangle_count()
    with pytest.raises(RuntimeError, match="vertex"):
        s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 3], np.uint32), [(mat, 0, 3)])
    with pytest.raises(RuntimeError, match="vertex"):
        s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 2], np.uint32), [(mat, 0, 3)], base_vertex=1)
    with pytest.raises(RuntimeError, match="mat_index"):
        s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 2], np.uint32), [(mat + 3, 0, 3)])
    with pytest.raises(RuntimeError, match="back_mat_index"):
        s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 2], np.uint32), [(mat, mat + 3, 0, 3)])
    with pytest.raises(RuntimeError, match="shape range"):
        s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 2], np.uint32), [(mat, 0, 6)])
    assert s.triangle_count() == tris
    assert s.add_mesh(A.PxyzNxyzTuv, attrs, np.array([0, 1, 2], np.uint32), [(mat, 0, 3)]) == 0
    assert s.triangle_count() == tris + 1


def test_environment_map_is_checked(scene):
    s, tex, _ = scene
    with pytest.raises(RuntimeError, match="env_map"):
        s.set_environment((1, 1, 1), 0.0, tex + 9)
    s.set_environment((1, 1, 1), 0.0, tex)
    s.set_environment((1, 1, 1), 0.0, A.INVALID)


def test_deep_mix_chain(scene):
    """A chain of mixes deeper than the reference's 32-entry walk stack
    (SceneRef.cpp:187) with a transparent material at the bottom: the
    triangles are marked non-solid."""
    s, tex, mat = scene
    top = s.add_material(A.TransparentMaterial, tex)
    for _ in range(40):
        top = s.add_material(A.MixMaterial, tex, mix_materials=(mat, top))
    s.add_mesh(A.PxyzNxyzTuv, _tri(), np.array([0, 1, 2], np.uint32), [(top, 0, 3)])
    assert not int(s.view()["tris"]["ci"][-1]) & 0x10
