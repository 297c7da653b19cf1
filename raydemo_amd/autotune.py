"""Autotuner for the HIP backend's launch knobs.

Every knob here changes only how the work is laid out on the device, never
what is computed: the traversal stack's LDS share (the rest spills to a
global slice), whether bounce rays are sorted by their origin/direction hash
before tracing, how many top bits of the 32-bit key the sort orders by, the
bounce from which the remaining bounces run as one path-per-lane launch
instead of one trace/shade/shadow stage each, the register budget (waves per
SIMD) of the path-tail and 8-wide trace kernels, the lane-refill trace, and
the path tail as a persistent kernel whose lanes take a new path when theirs
ends, and how many passes are in flight at once (each on its own stream:
a pass's first stages run while the earlier passes' late bounces drain).
The image is bit-identical for every setting (tests/test_gpu_sort.py), so the
tuner may pick whichever is fastest on the scene at hand.

The sweep renders a few progressive passes per candidate through the normal
renderer API (device-side timing with hipEvents around each pass loop) and
restores the renderer's accumulation afterwards by clearing it.  Results are
cached per (device, scene signature, frame size) in a JSON file.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from dataclasses import dataclass, field
from pathlib import Path

LDS_DEPTHS = (8, 16, 24, 32)
# (sort_rays, sort_bits): off, then the 8-bit radix passes over the top 8..32 key bits
SORT_MODES = ((0, 32), (1, 8), (1, 32))
# first bounce of the path-tail launch (-1: every bounce staged)
PATH_TAILS = (-1, 0, 1, 2, 3, 4)
TAIL_OCCUPANCY = (2, 3)
REFILL = (0, 1)
TRACE_OCCUPANCY = (0, 5, 6)
SHADOW_OCCUPANCY = (-1, 0, 5)
PASS_OVERLAP = (3, 2, 1, 0)  # passes in flight besides the newest
TRACE_GROUP = (0, 2, 4, 8)   # lanes per ray of the late staged bounces' traces (0: one)
KNOBS = ("lds_stack_depth", "sort_rays", "sort_bits", "path_tail", "tail_occupancy", "refill", "trace_occupancy",
         "shadow_occupancy", "tail_refill", "pass_overlap", "trace_group")


@dataclass
class TuneResult:
    best: dict
    table: list = field(default_factory=list)  # [{<knobs>, ms_per_pass, mrays}]
    seconds: float = 0.0

    def to_json(self) -> dict:
        return {"best": self.best, "table": self.table, "seconds": round(self.seconds, 3)}


def candidates(depths=LDS_DEPTHS, sort_modes=SORT_MODES, tails=PATH_TAILS, occupancies=TAIL_OCCUPANCY):
    for d, (so, sb), pt, occ in itertools.product(depths, sort_modes, tails, occupancies):
        if pt < 0 and occ != occupancies[0]:
            continue  # no tail launch: occupancy is moot
        yield {"lds_stack_depth": d, "sort_rays": so, "sort_bits": sb, "path_tail": pt, "tail_occupancy": occ,
               "refill": 0, "trace_occupancy": 0, "shadow_occupancy": -1, "tail_refill": 0, "pass_overlap": PASS_OVERLAP[0],
               "trace_group": 0}


def apply(renderer, cfg: dict):
    for k, v in cfg.items():
        renderer.set_option(k, int(v))


def measure(renderer, scene, region_factory, cfg: dict, passes: int, warmup: int = 2) -> dict:
    """Mrays/s of `passes` progressive passes under cfg (after `warmup` passes).
    With several passes in flight the first and last few of a timed run
    overlap less, so a candidate needs enough passes to show its rate."""
    apply(renderer, cfg)
    renderer.clear()
    reg = region_factory()
    for _ in range(warmup):
        renderer.render(scene, reg)
    renderer.synchronize()
    renderer.reset_stats()
    t0 = time.perf_counter()
    for _ in range(passes):
        renderer.render(scene, reg)
    renderer.synchronize()
    dt = time.perf_counter() - t0
    rc = renderer.ray_counts()
    rays = rc["primary"] + rc["secondary"] + rc["shadow"]
    return dict(cfg, passes=passes, ms_per_pass=round(dt / passes * 1e3, 4), mrays=round(rays / dt / 1e6, 2))


def tune(renderer, scene, region_factory, passes: int = 6, depths=LDS_DEPTHS, sort_modes=SORT_MODES,
         tails=PATH_TAILS, cache: str | os.PathLike | None = None, key: str | None = None) -> TuneResult:
    """Sweep the candidates and leave the renderer set to the fastest one."""
    if cache and key:
        p = Path(cache)
        if p.exists():
            try:
                hit = json.loads(p.read_text()).get(key)
            except (OSError, ValueError):
                hit = None
            if hit:
                apply(renderer, hit["best"])
                renderer.clear()
                return TuneResult(best=hit["best"], table=hit.get("table", []), seconds=0.0)
    t0 = time.perf_counter()
    mr = lambda r: r["mrays"]
    knobs_of = lambda r: {k: r[k] for k in KNOBS}
    table = [measure(renderer, scene, region_factory, cfg, passes) for cfg in candidates(depths, sort_modes, tails)]
    # The leaders are often within a few per cent of each other: the best five
    # are timed again over four times the passes, and every later stage times
    # its candidates that long and compares them only with the best of those.
    long_run = 4 * passes

    def trial(cfg):
        row = measure(renderer, scene, region_factory, cfg, long_run)
        table.append(row)
        return row

    best = max((trial(knobs_of(r)) for r in sorted(table, key=mr, reverse=True)[:5]), key=mr)
    # the refilled path tail, the refill trace, the capped-register trace and
    # shadow kernel, lane groups for the late staged bounces, then how many
    # passes are in flight, each on the best so far (all leave the image
    # unchanged)
    stages = [
        lambda b: [dict(knobs_of(b), path_tail=pt, tail_occupancy=occ, tail_refill=1)
                   for pt in sorted({p for p in (b["path_tail"], 0, 1, 2) if p >= 0}) for occ in TAIL_OCCUPANCY],
        lambda b: [dict(knobs_of(b), refill=rf) for rf in REFILL[1:]],
        lambda b: [dict(knobs_of(b), trace_occupancy=occ, lds_stack_depth=d)
                   for occ in TRACE_OCCUPANCY[1:] for d in sorted({b["lds_stack_depth"], LDS_DEPTHS[0]})],
        lambda b: [dict(knobs_of(b), shadow_occupancy=so) for so in SHADOW_OCCUPANCY[1:] if so != b["trace_occupancy"]],
        lambda b: [dict(knobs_of(b), trace_group=g) for g in TRACE_GROUP[1:] if b["path_tail"] < 0 or b["path_tail"] > 2],
        lambda b: [dict(knobs_of(b), pass_overlap=ov) for ov in PASS_OVERLAP if ov != b["pass_overlap"]],
    ]
    for stage in stages:
        best = max([best] + [trial(cfg) for cfg in stage(best)], key=mr)
    best = knobs_of(best)
    apply(renderer, best)
    renderer.clear()
    res = TuneResult(best=best, table=table, seconds=time.perf_counter() - t0)
    if cache and key:
        p = Path(cache)
        data = {}
        if p.exists():
            try:
                data = json.loads(p.read_text())
            except (OSError, ValueError):
                data = {}
        data[key] = res.to_json()
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(data, indent=1))
    return res
