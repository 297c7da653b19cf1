"""Features off the benchmark path against the reference Ref backend and the
oracle: a spot light, a normal-mapped quad, and the camera settings (Tent
filter, depth of field, sRGB with gamma, clamp, skip direct / indirect light,
lighting only, no background), each rendered on the GPU and by the reference
from the same scene description."""
import numpy as np
import pytest

from conftest import no_overflow, standalone_scene
from raydemo_amd import capi as A, scenes
from test_geo_oracle import _ref_render

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

W, H, SPP = 48, 40, 3

FLAGS = {
    "plain": {},
    "tent_dof": {"filter": A.Tent, "focus_factor": 0.5, "focus_distance": 12.0},
    "srgb_gamma_clamp": {"dtype": A.DeviceSRGB, "gamma": 2.2, "clamp": True},
    "skip_direct": {"skip_direct_lighting": True},
    "skip_indirect": {"skip_indirect_lighting": True},
    "lighting_only": {"lighting_only": True},
    "no_background": {"no_background": True},
}


def _build(flags):
    def build(s):
        scenes.cornell_box(s, width=W, height=H, spp=SPP, depth=3)
        s.add_light(A.SpotLight, position=(0.0, 9.0, 0.0), direction=(0.0, -1.0, 0.0), angle=35.0, radius=0.3,
                    color=(40.0, 38.0, 36.0))
        rng = np.random.default_rng(9)
        nm = np.zeros((8, 8, 4), np.uint8)
        nm[..., :2] = rng.integers(96, 160, (8, 8, 2))
        nm[..., 2] = 230
        nm[..., 3] = 255
        ntex = s.add_texture(nm, is_srgb=False, generate_mipmaps=False)
        white = s.add_texture(np.full((1, 1, 4), 255, np.uint8), is_srgb=False, generate_mipmaps=False)
        mat = s.add_material(A.DiffuseMaterial, white, main_color=(0.7, 0.7, 0.8), normal_map=ntex)
        qa, qi = scenes._quad((-3.0, 0.01, -3.0), (-3.0, 0.01, 1.0), (1.0, 0.01, 1.0), (1.0, 0.01, -3.0), (0, 1, 0))
        m = s.add_mesh(A.PxyzNxyzTuv, qa.astype(np.float32), qi, [(mat, 0, 6)])
        s.add_mesh_instance(m)
        kw = dict(origin=(0.0, 5.0, -14.0), fwd=(0.0, 0.0, 1.0), fov=45.0, filter=A.Box, dtype=A.DeviceNone,
                  max_diff_depth=3, max_glossy_depth=3, max_refr_depth=3, max_transp_depth=3, max_total_depth=3)
        kw.update(flags)
        s.set_camera(0, **kw)
    return build


@pytest.mark.parametrize("name", sorted(FLAGS))
def test_feature_renders_as_reference(ref_lib, host_lib, oracle, name):
    build = _build(FLAGS[name])
    r = A.Renderer(host_lib, W, H, A.RENDERER_HIP, use_wide_bvh=True)
    s = r.create_scene()
    build(s)
    reg = A.Region(host_lib, 0, 0, W, H)
    for _ in range(SPP):
        r.render(s, reg)
    assert no_overflow(r)
    img = r.pixels()
    img_r, _ = _ref_render(ref_lib, build, W, H, SPP, sh=False)
    e = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - img_r[..., :3]) ** 2)))
    hs = standalone_scene(host_lib, wide=True)
    build(hs)
    img_o, _ = oracle.render(hs.raw_view(), hs.camera_record(0), W, H, SPP)
    eo = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - img_o[..., :3]) ** 2)))
    print(f"{name}: RMSE vs reference {e:.3e}, vs oracle {eo:.3e}, mean {img[..., :3].mean():.4f}")
    assert e <= 1e-3 and eo <= 1e-3
