"""The many-scene workload through the batch pipeline against the CPU oracle
(the restatement of RendererRef.cpp:19-325, itself equal to the reference
library byte for byte): the first 8 scenes of the 64-scene workload with their
seeded triangle counts (1.9k-267k) and depths (5-8), resolution capped at
128^2 and spp at 4, binary and 8-wide trees alternating.  Every scene's image
must match the oracle's render of the same scene, and the batch's ray counts
the sum of the oracle's."""
import numpy as np
import pytest

from conftest import standalone_scene
from raydemo_amd import capi as A, scenes

pytestmark = pytest.mark.gpu


def reduced_items(n=8):
    return [scenes.BatchItem(it.seed, it.triangles, it.n_spheres, it.n_fine, min(it.width, 128),
                             min(it.height, 128), min(it.spp, 4), it.depth) for it in scenes.batch_workload(64)[:n]]


@pytest.fixture(scope="module")
def batch_cases(host_lib, oracle):
    cases = []
    for k, it in enumerate(reduced_items()):
        s = standalone_scene(host_lib, wide=(k % 2 == 1))
        scenes.batch_scene(s, it)
        img, counts = oracle.render(s.raw_view(), s.camera_record(0), it.width, it.height, it.spp)
        cases.append((it, s, img, counts))
    return cases


@pytest.mark.parametrize("opts", [{}, {"path_tail": -1}, {"path_tail": 1, "lds_stack_depth": 8},
                                  {"sort_bits": 32, "path_tail": -1}, {"sort_bits": 16, "path_tail": 2},
                                  {"sort_bits": 8, "path_tail": -1, "use_graph": 0},
                                  {"queue_rays": 20000, "sort_bits": 32, "path_tail": -1}, {"queue_rays": 40000}])
def test_batch_matches_oracle(host_lib, batch_cases, opts):
    b = A.Batch(host_lib)
    for k, v in opts.items():
        b.set_option(k, v)
    for it, s, _, _ in batch_cases:
        b.add(s, it.width, it.height, it.spp)
    b.render_all()
    rc = b.ray_counts()
    assert rc["overflow"] == 0
    total = np.sum([c[3] for c in batch_cases], axis=0)
    assert (rc["primary"], rc["secondary"], rc["shadow"]) == tuple(int(x) for x in total)
    for i, (it, _, ref, _) in enumerate(batch_cases):
        img = b.pixels(i)
        e = float(np.sqrt(np.mean((img[..., :3].astype(np.float64) - ref[..., :3]) ** 2)))
        same = float(np.mean(img.view(np.uint32) == ref.view(np.uint32)))
        print(f"scene {i} ({it.triangles} tris, {it.width}^2, {it.spp} spp, depth {it.depth}): RMSE {e:.2e}, "
              f"identical values {same:.5f}")
        assert e <= 1e-3, i
        assert same >= 0.999, i
    b.close()
