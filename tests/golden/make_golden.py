"""Writes the golden fixtures under tests/golden/ from the data arrays the
reference's own tests hold.  Only numbers are extracted (C array literals),
no code:

  primary_ray_gen.json  <- src/Ray/tests/test_data.cpp:2-18 (16 rays x 23 fields)
                           camera of src/Ray/tests/test_primary_ray_gen.cpp:24-30
  scene1.npz            <- src/Ray/tests/test_scene1.h (attrs, indices, groups)
                           + golden image src/Ray/tests/test_img1.h (64x64 RGBA8)
  scene2.npz            <- src/Ray/tests/test_scene2.h (texture-test plane)

Run in a container that has the reference checkout:
    python tests/golden/make_golden.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _array(text: str, name: str) -> list:
    m = re.search(r"\b" + re.escape(name) + r"\s*\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise ValueError(f"array {name} not found")
    body = m.group(1)
    toks = [t.strip() for t in body.replace("\n", " ").split(",")]
    out = []
    for t in toks:
        if not t:
            continue
        t = t.rstrip("fF") if not t.lower().startswith("0x") else t
        out.append(int(t, 16) if t.lower().startswith("0x") else float(t))
    return out


def main(ref_root: str = "/root/reference"):
    tests = Path(ref_root) / "src" / "Ray" / "tests"

    data = _array((tests / "test_data.cpp").read_text(), "g_primary_ray_gen_test_data")
    assert len(data) == 16 * 23
    (HERE / "primary_ray_gen.json").write_text(json.dumps({
        "source": "src/Ray/tests/test_data.cpp:2-18; camera src/Ray/tests/test_primary_ray_gen.cpp:24-30",
        "camera": {"origin": [0, 0, 4], "fwd": [0, 0, -1], "up": [0, 0, 0], "fov": 53.13, "filter": "Box",
                   "dtype": "None", "gamma": 2.2, "focus_distance": 1.0, "focus_factor": 0.0},
        "image": [4, 4], "iteration": 0, "halton": "all zeros", "tolerance_abs": 1e-3,
        "fields": ["xy", "o0", "o1", "o2", "d0", "d1", "d2", "c0", "c1", "c2", "ior", "do_dx0", "do_dx1", "do_dx2",
                   "dd_dx0", "dd_dx1", "dd_dx2", "do_dy0", "do_dy1", "do_dy2", "dd_dy0", "dd_dy1", "dd_dy2"],
        "rays": [data[i * 23:(i + 1) * 23] for i in range(16)],
    }, indent=1))

    s1 = (tests / "test_scene1.h").read_text()
    img = (tests / "test_img1.h").read_text()
    attrs = np.array(_array(s1, "attrs"), dtype=np.float32)
    indices = np.array(_array(s1, "indices"), dtype=np.uint32)
    groups = np.array(_array(s1, "groups"), dtype=np.uint32)
    img_data = np.array(_array(img, "img_data"), dtype=np.uint8).reshape(64, 64, 4)
    np.savez_compressed(HERE / "scene1.npz", attrs=attrs, indices=indices, groups=groups, img=img_data)

    s2 = (tests / "test_scene2.h").read_text()
    np.savez_compressed(HERE / "scene2.npz", attrs=np.array(_array(s2, "attrs"), dtype=np.float32),
                        indices=np.array(_array(s2, "indices"), dtype=np.uint32),
                        groups=np.array(_array(s2, "groups"), dtype=np.uint32))
    print("wrote", sorted(p.name for p in HERE.glob("*.json")) + sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main(*sys.argv[1:])
