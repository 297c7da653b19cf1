"""Benchmark: Mrays/s of the HIP backend on the 1080p depth-8 triangle-soup workload.

One step = one RenderScene pass (1 spp) over the full 1920x1080 frame of the
~200k-triangle icosphere soup (sphere lights + environment, diffuse/glossy,
max depth 8), through the Ray renderer API of the HIP backend.  The timed
region is exactly K passes bracketed by device synchronisation; rays counted
on the device (primary + secondary + shadow) divided by the wall time give
Mrays/s.  Scene build and upload happen before the timed region.

Also reported on the same JSON line:
  roofline      the trace kernel's achieved algorithmic bandwidth (bytes per
                traced ray from the node/triangle/instance visit counts of an
                instrumented replay of the same passes) over its HIP-event time
                inside the timed region, against 8 TB/s
  cpu_baseline  the reference Ray library's Ref backend (or, where that is not
                built, the CPU oracle) on the host cores, 128x128 buckets on a
                thread pool, over a bounded sample of the same workload
  modes         Mrays/s with secondary-ray sorting on and off

Usage: python bench.py [--gpus 1] [--steps 64] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import queue
import sys
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes (layouts of include/rh_types.h): per traced ray 32 (ray) + 32 (hit);
# per node visit 32; per instance test 112 (instance 32 + inverse transform 64 + mesh 16);
# per triangle test 52 (index 4 + triangle record 48)
B_RAY, B_HIT, B_NODE, B_INST, B_TRI = 32, 32, 32, 112, 52


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(args, lib, A, scenes):
    W, H = args.width, args.height
    r = A.Renderer(lib, W, H, A.RENDERER_HIP, use_wide_bvh=False)
    s = r.create_scene()
    t0 = time.time()
    spec = scenes.triangle_soup(s, width=W, height=H, spp=args.steps, depth=args.depth)
    log(f"scene: {spec.triangles} triangles, host build {time.time() - t0:.1f}s")

    results = {}
    for sort_on in (1, 0):
        r.set_option("sort_rays", sort_on)
        r.set_option("kernel_timing", 1)
        r.clear()
        reg = A.Region(lib, 0, 0, W, H)
        for _ in range(args.warmup):
            r.render(s, reg)
        r.synchronize()
        r.reset_stats()
        r.reset_kernel_profile()
        r.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r.render(s, reg)
        r.synchronize()
        dt = time.perf_counter() - t1
        rc = r.ray_counts()
        rays = rc["primary"] + rc["secondary"] + rc["shadow"]
        prof = r.kernel_profile()
        results[sort_on] = dict(seconds=dt, rays=rays, counts=rc, prof=prof, mrays=rays / dt / 1e6)
        log(f"sort={sort_on}: {args.steps} passes in {dt:.3f}s, {rays / 1e6:.1f} Mrays -> {rays / dt / 1e6:.1f} Mrays/s")

    # instrumented replay of the same passes (sorting on) for the visit counts
    r.set_option("sort_rays", 1)
    r.set_option("kernel_timing", 0)
    r.set_option("count_visits", 1)
    r.clear()
    reg = A.Region(lib, 0, 0, W, H)
    for _ in range(args.warmup):
        r.render(s, reg)
    r.synchronize()
    r.reset_kernel_profile()
    r.reset_stats()
    for _ in range(args.steps):
        r.render(s, reg)
    r.synchronize()
    vis = r.kernel_profile()
    rc2 = r.ray_counts()
    r.set_option("count_visits", 0)
    traced = rc2["primary"] + rc2["secondary"]
    return spec, results, vis, traced


def run_cpu_baseline(args, A, scenes, spp_sample):
    """Reference Ref backend (or the CPU oracle) over spp_sample full-frame passes."""
    ref_path = ROOT / "oracle" / "_ref" / "libray_ref.so"
    kind = "reference"
    if not ref_path.exists():
        return None
    lib = A.RayLib(str(ref_path))
    W, H = args.width, args.height
    r = A.Renderer(lib, W, H, A.RENDERER_REF, use_wide_bvh=False)
    s = r.create_scene()
    scenes.triangle_soup(s, width=W, height=H, depth=args.depth)
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    regions = [A.Region(lib, x, y, min(128, W - x), min(128, H - y)) for y in range(0, H, 128) for x in range(0, W, 128)]

    def one_pass():
        q = queue.Queue()
        for g in regions:
            q.put(g)

        def work():
            while True:
                try:
                    g = q.get_nowait()
                except queue.Empty:
                    return
                r.render(s, g)

        th = [threading.Thread(target=work) for _ in range(cores)]
        for t in th:
            t.start()
        for t in th:
            t.join()

    t0 = time.perf_counter()
    for _ in range(spp_sample):
        one_pass()
    dt = time.perf_counter() - t0
    return dict(kind=kind, seconds=dt, cores=cores, spp=spp_sample)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-spp", type=int, default=8, help="full-frame passes of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from raydemo_amd import capi as A, host_lib, scenes

    lib = host_lib()
    spec, res, vis, traced = run_gpu(args, lib, A, scenes)
    on, off = res[1], res[0]
    best = on
    prof = on["prof"]

    # roofline of the trace kernel (closest hit), sorting on
    trace_ms = prof["trace_ms"]
    bytes_alg = (traced * (B_RAY + B_HIT) + vis["node_visits"] * B_NODE + vis["instance_tests"] * B_INST +
                 vis["tri_tests"] * B_TRI)
    achieved = bytes_alg / (trace_ms / 1e3) / 1e9 if trace_ms > 0 else 0.0
    traffic_file = ROOT / "profiles" / "trace_traffic.json"
    traffic = None
    if traffic_file.exists():
        try:
            traffic = json.loads(traffic_file.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "k_trace<false,false>", "launches": prof["trace_launches"],
        "avg_launch_ms": round(trace_ms / max(prof["trace_launches"], 1), 4),
        "alg_bytes_per_launch": int(bytes_alg / max(prof["trace_launches"], 1)),
    }

    cpu = None
    if not args.no_cpu:
        cpu_run = run_cpu_baseline(args, A, scenes, args.cpu_spp)
        if cpu_run is not None:
            # rays of the identical sample, counted on the device (same paths, same seeds)
            r = A.Renderer(lib, args.width, args.height, A.RENDERER_HIP, use_wide_bvh=False)
            s = r.create_scene()
            scenes.triangle_soup(s, width=args.width, height=args.height, depth=args.depth)
            reg = A.Region(lib, 0, 0, args.width, args.height)
            r.reset_stats()
            for _ in range(cpu_run["spp"]):
                r.render(s, reg)
            r.synchronize()
            rc = r.ray_counts()
            rays = rc["primary"] + rc["secondary"] + rc["shadow"]
            cpu = {"value": round(rays / cpu_run["seconds"] / 1e6, 3), "unit": "Mrays/s", "cores": cpu_run["cores"],
                   "kind": cpu_run["kind"],
                   "sample": f"{cpu_run['spp']} full-frame 1920x1080 passes (spp 1..{cpu_run['spp']}) of the same scene, "
                             f"128x128 buckets on {cpu_run['cores']} threads, {cpu_run['seconds']:.1f}s"}
            log(f"cpu baseline: {cpu}")

    line = {
        "metric": "Mrays/s at 1080p depth-8 (triangle soup, primary+secondary+shadow rays)",
        "value": round(best["mrays"], 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["seconds"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"triangle_soup {spec.triangles} tris, {args.width}x{args.height}, "
                               f"{args.steps} spp, depth {args.depth}, sorted secondary rays",
                   "triangles": spec.triangles, "width": args.width, "height": args.height, "spp": args.steps,
                   "depth": args.depth, "parallelism": "single GPU"},
        "modes": {"sort_on_mrays": round(on["mrays"], 2), "sort_off_mrays": round(off["mrays"], 2),
                  "sort_off_kernel_ms": {k: round(v, 3) for k, v in off["prof"].items() if k.endswith("_ms")},
                  "rays_per_pass": int(on["rays"] / args.steps), "ray_counts": on["counts"]},
        "kernel_ms": {k: round(v, 3) for k, v in prof.items() if k.endswith("_ms")},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
