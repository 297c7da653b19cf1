"""The CPU oracle's FP64 path-replay gradient (orc_adjoint) against central
finite differences of the oracle renderer itself, on a small refined Cornell
box with one material per triangle.  The loss is linear in the image,
sum_p w_p . I_p, so d loss / d main_color is exactly what the adjoint returns;
with Russian roulette off, the sampled paths do not depend on the colours and
the rendered mean image is a polynomial in them, so the differences converge
to the derivative (h = 1e-3, FP32 rendering)."""
import ctypes as C

import numpy as np

from conftest import standalone_scene
from raydemo_amd import capi as A, scenes


def _scene(host_lib, k=2):
    s = standalone_scene(host_lib)
    spec = scenes.cornell_refined(s, k=k, width=24, height=24, spp=3, depth=3)
    return s, spec


def _with_materials(view, mats):
    v = A.SceneView()
    C.memmove(C.byref(v), C.byref(view), C.sizeof(A.SceneView))
    v.materials = mats.ctypes.data
    return v


def test_oracle_gradient_matches_finite_differences(oracle, host_lib):
    s, spec = _scene(host_lib)
    w, h, spp = spec.width, spec.height, spec.spp
    view = s.raw_view()
    cam = s.camera_record(0)
    mats = A._arr(view.materials, view.materials_count, A.MATERIAL_DTYPE)
    rng = np.random.default_rng(7)
    wts = rng.uniform(-1.0, 1.0, size=(h, w, 4)).astype(np.float32)
    wts[..., 3] = 0.0
    grad = oracle.adjoint(_with_materials(view, mats), cam, w, h, spp, wts)
    assert np.abs(grad).max() > 0

    def loss(m):
        img = oracle.render_linear(_with_materials(view, m), cam, w, h, spp)
        return float(np.sum(img[..., :3].astype(np.float64) * wts[..., :3]))

    hstep = 1e-3
    # materials seen by the camera the most, plus the emissive light
    order = np.argsort(-np.abs(grad).sum(axis=1))
    picks = list(order[:6]) + [spec.info["light_material"]]
    for m in picks:
        for ch in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp["main_color"][m, ch] += hstep
            mm["main_color"][m, ch] -= hstep
            fd = (loss(mp) - loss(mm)) / (2 * hstep)
            g = grad[m, ch]
            assert abs(fd - g) <= 5e-4 * abs(g) + 1e-5, (m, ch, fd, g)


def test_oracle_gradient_of_emission_is_linear(oracle, host_lib):
    """Closed form: with the light's main_color scaled by a, the image scales by
    a (all radiance comes from the one emissive material), so d loss / d
    main_color_light . main_color_light = loss."""
    s, spec = _scene(host_lib)
    w, h, spp = spec.width, spec.height, spec.spp
    view = s.raw_view()
    cam = s.camera_record(0)
    wts = np.ones((h, w, 4), np.float32)
    wts[..., 3] = 0.0
    grad = oracle.adjoint(view, cam, w, h, spp, wts)
    img = oracle.render_linear(view, cam, w, h, spp)
    lm = spec.info["light_material"]
    for ch in range(3):
        total = float(np.sum(img[..., ch].astype(np.float64)))
        assert abs(grad[lm, ch] * 1.0 - total) <= 1e-5 * total
