"""Instances moved between passes (SceneBase::SetMeshInstanceTransform,
SceneBase.h:265, the animated-scene case): the renderer rebuilds the instance
tree (on the device) and replaces only the arrays that follow the instances
(rh_update_instances), keeping triangles, textures and materials resident.
The passes after the move equal, bit for bit, those of a renderer given the
moved scene from the start (which uploads everything), and the oracle's
render of that scene."""
import numpy as np
import pytest

from conftest import no_overflow
from raydemo_amd import capi as A

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _xforms(rng, n, t):
    out = []
    for k in range(n):
        x = np.eye(4, dtype=np.float32)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x[:3, :3] = (q * rng.uniform(0.2, 0.6)).astype(np.float32)
        x[3, :3] = (rng.uniform(-4, 4), rng.uniform(0.3, 3.0) + 0.5 * np.sin(t + k), rng.uniform(2, 10))
        out.append(x)
    return out


def _scene(s, xforms):
    rng = np.random.default_rng(3)
    tex = s.add_texture(np.full((4, 4, 4), 220, np.uint8), False, True)
    mats = [s.add_material(A.DiffuseMaterial, tex, main_color=tuple(rng.uniform(0.2, 0.9, 3))) for _ in range(3)]
    glossy = s.add_material(A.GlossyMaterial, tex, main_color=(0.9, 0.9, 0.9), roughness=0.2)
    light = s.add_material(A.EmissiveMaterial, tex, main_color=(1, 1, 1), strength=8.0)
    # a blob of 60 triangles, instanced; a ground quad; an emissive quad
    c = rng.uniform(-1, 1, (60, 1, 3))
    P = (c + rng.normal(0, 0.35, (60, 3, 3))).reshape(-1, 3).astype(np.float32)
    attrs = np.zeros((len(P), 8), np.float32)
    attrs[:, :3] = P
    attrs[:, 4] = 1.0
    blob = s.add_mesh(A.PxyzNxyzTuv, attrs, np.arange(len(P), dtype=np.uint32),
                      [(mats[0], 0, 60), (glossy, 60, 60), (mats[1], 120, 60)])
    g = np.array([[-20, 0, -5], [20, 0, -5], [20, 0, 30], [-20, 0, 30]], np.float32)
    ga = np.zeros((4, 8), np.float32)
    ga[:, :3] = g
    ga[:, 4] = 1.0
    ground = s.add_mesh(A.PxyzNxyzTuv, ga, np.array([0, 2, 1, 0, 3, 2], np.uint32), [(mats[2], 0, 6)])
    q = np.array([[-2, 8, 4], [2, 8, 4], [2, 8, 8], [-2, 8, 8]], np.float32)
    qa = np.zeros((4, 8), np.float32)
    qa[:, :3] = q
    qa[:, 4] = -1.0
    lamp = s.add_mesh(A.PxyzNxyzTuv, qa, np.array([0, 1, 2, 0, 2, 3], np.uint32), [(light, 0, 6)])
    s.add_mesh_instance(ground)
    s.add_mesh_instance(lamp)
    ids = [s.add_mesh_instance(blob, x) for x in xforms]
    s.add_light(A.PointLight, position=(0, 6, 3), radius=0.3, color=(30, 30, 30))
    cam = s.add_camera(origin=(0, 3, -6), fwd=(0, -0.15, 1), fov=50.0, max_total_depth=4, max_diff_depth=4,
                       max_glossy_depth=4)
    s.set_current_cam(cam)
    return ids


def _passes(r, s, spp):
    r.clear()
    reg = A.Region(r.lib, 0, 0, W, H)
    for _ in range(spp):
        r.render(s, reg)
    assert no_overflow(r)
    return r.pixels(), r.ray_counts()


@pytest.mark.parametrize("wide", [False, True])
def test_moved_instances_render_as_fresh_upload(host_lib, oracle, wide):
    n = 300
    before, after = _xforms(np.random.default_rng(7), n, 0.0), _xforms(np.random.default_rng(7), n, 1.3)
    r = A.Renderer(host_lib, W, H, A.RENDERER_HIP, use_wide_bvh=wide)
    s = r.create_scene()
    ids = _scene(s, before)
    _, counts0 = _passes(r, s, 2)
    uploads = r.get_option("scene_uploads")
    for mi, x in zip(ids, after):
        s.set_mesh_instance_transform(mi, x)
    img, counts1 = _passes(r, s, 3)
    counts = {k: counts1[k] - counts0[k] for k in counts1}  # the counts run from the renderer's start
    assert r.get_option("scene_uploads") == uploads
    assert r.get_option("instance_updates") >= 1

    r2 = A.Renderer(host_lib, W, H, A.RENDERER_HIP, use_wide_bvh=wide)
    s2 = r2.create_scene()
    _scene(s2, after)
    img2, counts2 = _passes(r2, s2, 3)
    assert img.tobytes() == img2.tobytes()
    assert counts == counts2

    ref = oracle.render_linear(s2.raw_view(), s2.camera_record(0), W, H, 3)
    lin = r.read_accum()
    e = float(np.sqrt(np.mean((lin[..., :3].astype(np.float64) - ref[..., :3]) ** 2)))
    assert e <= 1e-3, e


def test_material_change_after_move_is_kept(host_lib):
    """A colour change made between a move and the next pass reaches the
    device (the moved-only update keeps the materials it already has, so the
    change is patched in place)."""
    before, after = _xforms(np.random.default_rng(8), 50, 0.0), _xforms(np.random.default_rng(8), 50, 0.7)
    r = A.Renderer(host_lib, W, H, A.RENDERER_HIP)
    s = r.create_scene()
    ids = _scene(s, before)
    _passes(r, s, 1)
    for mi, x in zip(ids, after):
        s.set_mesh_instance_transform(mi, x)
    r.set_material_colors(s, 0, np.array([[0.1, 0.8, 0.3]], np.float32))
    img, _ = _passes(r, s, 2)

    r2 = A.Renderer(host_lib, W, H, A.RENDERER_HIP)
    s2 = r2.create_scene()
    _scene(s2, after)
    r2.set_material_colors(s2, 0, np.array([[0.1, 0.8, 0.3]], np.float32))
    img2, _ = _passes(r2, s2, 2)
    assert img.tobytes() == img2.tobytes()
