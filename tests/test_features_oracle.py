"""The oracle against the reference Ref backend on the features off the
benchmark path (spot light, normal map, Tent filter, depth of field, sRGB
with gamma, clamp, skip direct / indirect light, lighting only, no
background): images byte for byte."""
import pytest

from conftest import standalone_scene
from test_geo_oracle import _ref_render


def _cases():
    import test_gpu_features as F
    return F


@pytest.mark.parametrize("name", ["lighting_only", "no_background", "plain", "skip_direct", "skip_indirect",
                                  "srgb_gamma_clamp", "tent_dof"])
def test_oracle_features_match_reference(ref_lib, host_lib, oracle, name):
    F = _cases()
    build = F._build(F.FLAGS[name])
    img_r, _ = _ref_render(ref_lib, build, F.W, F.H, F.SPP, sh=False)
    hs = standalone_scene(host_lib)
    build(hs)
    img_o, _ = oracle.render(hs.raw_view(), hs.camera_record(0), F.W, F.H, F.SPP)
    assert img_o.tobytes() == img_r.tobytes()
