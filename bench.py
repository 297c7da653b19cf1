"""Benchmark: Mrays/s of the HIP backend on the 1080p depth-8 triangle-soup workload.

One step = one RenderScene pass (1 spp) over the full 1920x1080 frame of the
~200k-triangle icosphere soup (sphere lights + environment, diffuse/glossy,
max depth 8), through the Ray renderer API of the HIP backend.  The timed
region is exactly K passes bracketed by device synchronisation; rays counted
on the device (primary + secondary + shadow) divided by the wall time give
Mrays/s.  Scene build and upload happen before the timed region.

The launch knobs (LDS share of the traversal stack, ray sort on/off and the
number of key bits it orders by) are chosen first by raydemo_amd.autotune on
a few passes per candidate; every candidate renders the identical image.  The
headline value is the tuned setting (by default four passes in flight, each
on its own stream: a pass's first stages run while the earlier passes' late
bounces drain); the same knobs with one pass in flight (serial), and the staged
sort-off (coherent order) and full 32-bit sort modes are timed the same way
and reported beside it.

Also reported on the same JSON line:
  roofline      the trace kernel's achieved algorithmic bandwidth (bytes per
                traced ray from the node/triangle/instance visit counts of an
                instrumented replay of the same passes, which count the
                reference traversal's visits: equal to the oracle's counters on
                the same rays, tests/test_gpu_trace.py) over its HIP-event time
                inside the timed region, against 8 TB/s; traffic = HBM bytes
                per launch from the copy-calibrated FETCH_SIZE/WRITE_SIZE
                counters committed in profiles/trace_traffic.json;
                per_kernel adds the shadow kernel (per queue entry) and the
                path tail (per shaded segment and light sample) from the same
                replay's counters
  cpu_baseline  the reference Ray library's Ref backend (and its AVX2 backend
                beside it) on the host cores, 128x128 buckets on a thread pool,
                over a bounded sample of the same workload; the CPU model and
                core counts are recorded
  modes         Mrays/s and per-kernel ms of each timed mode

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
# per interior node visit of an 8-wide tree 224 (mbvh_node_t: 8 boxes + 8 children)
B_RAY, B_HIT, B_NODE, B_INST, B_TRI = 32, 32, 32, 112, 52
B_WNODE = 224
# per shadow-queue entry: queue index 4 + shadow ray record 24 (nee_ray 16, nee_aux 8) + visibility byte 1
B_SHADOW_ENTRY = 29
# per shaded path segment, the reference's shading bytes (ray 112 + hit 32 + triangle 48 + vertex indices 12 +
# 3 vertices 156 + material 48 + instance and transform 96 + 2 textures 148 + 2 bilinear footprints 32 +
# secondary ray 112 + framebuffer read-modify-write 32); per deferred light sample its record 40
B_SHADE, B_NEE_RECORD = 828, 40


def traversal_bytes(v: dict, wide: bool) -> float:
    """Bytes of the reference structures a traversal reads, from its visit counts."""
    return v["nodes"] * (B_WNODE if wide else B_NODE) + v["instances"] * B_INST + v["tris"] * B_TRI


def kernel_roofline(name: str, alg_bytes: float, ms: float, launches: int, traffic=None) -> dict:
    achieved = alg_bytes / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "launches": launches,
            "avg_launch_ms": round(ms / max(launches, 1), 4), "alg_bytes_per_launch": int(alg_bytes / max(launches, 1))}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_passes(r, s, A, lib, W, H, steps, warmup, opts):
    """K progressive passes under opts, bracketed by device synchronisation."""
    for k, v in opts.items():
        r.set_option(k, v)
    r.set_option("kernel_timing", 1)
    r.clear()
    reg = A.Region(lib, 0, 0, W, H)
    for _ in range(warmup):
        r.render(s, reg)
    r.synchronize()
    r.reset_stats()
    r.reset_kernel_profile()
    r.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        r.render(s, reg)
    r.synchronize()
    dt = time.perf_counter() - t1
    rc = r.ray_counts()
    rays = rc["primary"] + rc["secondary"] + rc["shadow"]
    prof = r.kernel_profile()
    r.set_option("kernel_timing", 0)
    return dict(opts=dict(opts), seconds=dt, rays=rays, counts=rc, prof=prof, mrays=rays / dt / 1e6)


def run_gpu(args, lib, A, scenes, wide, staged_modes=True, tuned=None):
    """Autotune and time the workload on one tree kind (binary or the
    reference's default 8-wide trees); then the instrumented replay."""
    from raydemo_amd import autotune

    W, H = args.width, args.height
    r = A.Renderer(lib, W, H, A.RENDERER_HIP, use_wide_bvh=wide)
    s = r.create_scene()
    t0 = time.time()
    spec = scenes.triangle_soup(s, width=W, height=H, spp=args.steps, depth=args.depth)
    log(f"scene ({'8-wide' if wide else 'binary'} trees): {spec.triangles} triangles, "
        f"host build {time.time() - t0:.1f}s")

    # launch knobs (LDS stack share, sort on/off and key bits, path tail): all give the same image
    if tuned is not None:
        pass  # knobs chosen by an earlier sweep of the same scene
    elif args.knobs:
        fixed = dict(zip(autotune.KNOBS, (int(v) for v in args.knobs.split(","))))
        tuned = autotune.TuneResult(best=fixed)
        log(f"knobs given: {fixed}")
    else:
        tuned = autotune.tune(r, s, lambda: A.Region(lib, 0, 0, W, H), passes=args.tune_passes)
        log(f"autotune ({tuned.seconds:.1f}s): best {tuned.best}")
    depth = tuned.best["lds_stack_depth"]

    results = {}
    modes = {"tuned": dict(tuned.best),
             # the tuned knobs with one pass in flight (each pass waits for the last one's path tail)
             "serial": dict(tuned.best, pass_overlap=0),
             "sort_off": {"lds_stack_depth": depth, "sort_rays": 0, "sort_bits": 32, "path_tail": -1,
                          "pass_overlap": tuned.best.get("pass_overlap", 3)},
             "sort_on": {"lds_stack_depth": depth, "sort_rays": 1, "sort_bits": 32, "path_tail": -1,
                         "pass_overlap": tuned.best.get("pass_overlap", 3)}}
    if args.only_tuned or not staged_modes:
        modes = {"tuned": modes["tuned"], "serial": modes["serial"]}
    if not tuned.best.get("pass_overlap", 1):
        del modes["serial"]  # the tuned setting is already one pass in flight
    for name, opts in modes.items():
        res = timed_passes(r, s, A, lib, W, H, args.steps, args.warmup, opts)
        results[name] = res
        log(f"{name} {opts}: {args.steps} passes in {res['seconds']:.3f}s, {res['rays'] / 1e6:.1f} Mrays -> "
            f"{res['mrays']:.1f} Mrays/s")
    if staged_modes and not args.only_tuned:
        # the tuned knobs with UseCoherentSampling (the 8x8 pattern as the random index)
        s.set_camera(0, **scenes.soup_camera(args.depth, coherent=True))
        res = timed_passes(r, s, A, lib, W, H, args.steps, args.warmup, dict(tuned.best))
        results["coherent"] = res
        log(f"coherent sampling, tuned knobs: {res['mrays']:.1f} Mrays/s")
        s.set_camera(0, **scenes.soup_camera(args.depth, coherent=False))

    # instrumented replay of the tuned passes for the traversal visit counts
    autotune.apply(r, tuned.best)
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
    traced = vis["counted_rays"]  # rays of the counting k_trace launches (the tail and shadow kernels count apart)
    r.close()
    return spec, results, vis, traced, tuned


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> dict:
    """Threads the CPU baseline uses: every core this process may run on, up
    to 16 (the host share of one GPU on the machines this runs on)."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = nproc
    return {"nproc": nproc, "affinity": allowed, "threads": max(1, min(allowed, 16))}


def run_cpu_baseline(args, A, scenes, spp_sample, flags):
    """A reference CPU backend (flags: RENDERER_REF or RENDERER_AVX2) over
    spp_sample full-frame passes."""
    ref_path = ROOT / "oracle" / "_ref" / "libray_ref.so"
    kind = "reference"
    if not ref_path.exists():
        return None
    lib = A.RayLib(str(ref_path))
    W, H = args.width, args.height
    # the reference's default settings (settings_t::use_wide_bvh = true, RendererBase.h:29):
    # its 8-wide BVH is the faster of its two trees on the host
    r = A.Renderer(lib, W, H, flags, use_wide_bvh=True)
    backend = {A.RENDERER_REF: "Ref", A.RENDERER_SSE2: "SSE2", A.RENDERER_AVX: "AVX",
               A.RENDERER_AVX2: "AVX2"}.get(r.type(), str(r.type()))
    s = r.create_scene()
    scenes.triangle_soup(s, width=W, height=H, depth=args.depth)
    cores = host_threads()["threads"]
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
    return dict(kind=kind, seconds=dt, cores=cores, spp=spp_sample, backend=backend)


def device_used_bytes() -> int:
    import torch
    free, total = torch.cuda.mem_get_info()
    return int(total - free)


def run_adjoint(args, lib, A, scenes, knobs, wide):
    """Refined Cornell box (9,828 triangles, one material per triangle), 512x512,
    32 spp, depth 6: Adam on the 9,826 albedos against a target rendered from a
    hidden albedo set (seed 0x5EED0005).  One iteration = forward (32 passes) +
    adjoint by path replay (32 passes) + Adam update: the headline, with memory
    O(pixels).  The same with the tape (the forward keeps every pass's vertex
    records, memory O(pixels x passes x depth), no replay) is timed beside it."""
    import torch
    from raydemo_amd import adjoint

    r = A.Renderer(lib, args.adj_size, args.adj_size, A.RENDERER_HIP, use_wide_bvh=wide)
    dr, init, spec = adjoint.refined_cornell(lib, width=args.adj_size, height=args.adj_size, spp=args.adj_spp,
                                            depth=6, renderer=r)
    for k, v in knobs.items():
        dr.r.set_option(k, v)
    n = spec.info["diffuse"]
    hidden = init.copy()
    hidden[:n] = scenes.refined_albedos(n, scenes.SEED_BASE + 5)
    target = dr.render(torch.tensor(hidden))[..., :3].clone()
    # stage timing of one forward and one backward (after a warm-up)
    x = torch.tensor(init, device="cuda")
    for _ in range(2):
        dr.r.reset_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        img = dr.render(x)[..., :3]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = dr.adjoint(2 * (img - target) / img.numel())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    rc = dr.r.ray_counts()
    fwd_rays = rc["primary"] + rc["secondary"] + rc["shadow"]
    adj_rays = rc["adjoint"] + rc["adjoint_shadow"]
    res = adjoint.descend(dr, init, target, hidden, steps=args.adj_steps, lr=0.02, optimize=slice(0, n))
    used_replay = device_used_bytes()
    # the tape beside it: same workload, the forward keeps its vertex records
    dr.tape = True
    dr.r.set_option("adjoint_tape", dr.spp)
    tape_steps = max(1, args.adj_steps // 4)
    res_tape = adjoint.descend(dr, init, target, hidden, steps=tape_steps, lr=0.02, optimize=slice(0, n))
    used_tape = device_used_bytes()
    tape_bytes = dr.tape_bytes()
    dr.tape = False
    dr.r.set_option("adjoint_tape", 0)
    out = {
        "metric": "adjoint s/iter (forward + path-replay adjoint + Adam)",
        "s_per_iter": round(res.s_per_iter, 5), "steps": res.steps, "seconds": round(res.seconds, 3),
        "forward_ms": round((t1 - t0) * 1e3, 3), "adjoint_ms": round((t2 - t1) * 1e3, 3),
        "forward_mrays": round(fwd_rays / (t1 - t0) / 1e6, 2), "adjoint_mrays": round(adj_rays / (t2 - t1) / 1e6, 2),
        "loss_first": res.loss_first, "loss_last": res.loss_last,
        "albedo_rmse_first": round(res.albedo_rmse_first, 5), "albedo_rmse_last": round(res.albedo_rmse_last, 5),
        "knobs": knobs,
        "tape": False,
        "replay_record_bytes": dr.replay_bytes(), "device_used_bytes_replay": used_replay,
        "tape_s_per_iter": round(res_tape.s_per_iter, 5), "tape_steps": tape_steps, "tape_record_bytes": tape_bytes,
        "device_used_bytes_tape": used_tape,
        "config": {"workload": f"cornell_refined {spec.triangles} tris, {n} albedo parameters x 3, "
                               f"{args.adj_size}x{args.adj_size}, {args.adj_spp} spp, depth 6, roulette off",
                   "trees": "8-wide" if wide else "binary",
                   "gradient": "d L2(image, target) / d main_color, all materials"},
    }
    log(f"adjoint: {out}")
    return out


def run_batch(args, lib, A, scenes, knobs, wide):
    """The 64-scene workload: rendered scene after scene by single-scene
    renderers (the plain wavefront path, each scene's passes back to back),
    then all together by the batch renderer.  Same passes, same rays (the
    images are identical); Mrays/s of each."""
    from concurrent.futures import ThreadPoolExecutor

    items = scenes.batch_workload(args.batch_scenes)
    t0 = time.time()

    def build(it):
        s = A.standalone_scene(lib, wide)
        scenes.batch_scene(s, it)
        s.triangle_count()
        return s

    with ThreadPoolExecutor(max_workers=16) as ex:
        ss = list(ex.map(build, items))
    log(f"batch: {len(items)} scenes, {sum(s.triangle_count() for s in ss)} triangles, host build {time.time() - t0:.1f}s")

    def total(rc):
        return rc["primary"] + rc["secondary"] + rc["shadow"]

    # plain: one renderer per scene (scene uploaded by a warm-up pass), scenes rendered in turn
    rs = []
    for it, s in zip(items, ss):
        r = A.Renderer(lib, it.width, it.height, A.RENDERER_HIP, use_wide_bvh=wide)
        for k, v in knobs.items():
            r.set_option(k, v)
        r.render(s, A.Region(lib, 0, 0, it.width, it.height))
        r.synchronize()
        r.clear()
        r.reset_stats()
        rs.append(r)
    # plain, scene after scene: each renderer's passes, then a wait before
    # the next scene (every renderer has its own stream; without the wait the
    # scenes' passes overlap on the GPU, reported apart as "overlapped")
    t1 = time.perf_counter()
    for r, s, it in zip(rs, ss, items):
        reg = A.Region(lib, 0, 0, it.width, it.height)
        for _ in range(it.spp):
            r.render(s, reg)
        r.synchronize()
    t_plain = time.perf_counter() - t1
    rays_plain = sum(total(r.ray_counts()) for r in rs)
    for r in rs:
        r.clear()
        r.reset_stats()
    t1 = time.perf_counter()
    for r, s, it in zip(rs, ss, items):
        reg = A.Region(lib, 0, 0, it.width, it.height)
        for _ in range(it.spp):
            r.render(s, reg)
    for r in rs:
        r.synchronize()
    t_overlap = time.perf_counter() - t1
    rays_overlap = sum(total(r.ray_counts()) for r in rs)
    for r in rs:
        r.close()
    del rs

    b = A.Batch(lib)
    for it, s in zip(items, ss):
        b.add(s, it.width, it.height, it.spp)
    # the batch's knobs by coordinate descent from the defaults, 3 passes per
    # setting: path tail, LDS share of the traversal stack, persistent grid
    # size, sort-key bits of the per-scene queue sort, and the queue size (rays
    # in flight: scenes grouped to fit, 0 = every scene at once)
    space = {"path_tail": (-1, 1, 2, 3), "lds_stack_depth": (8, 16, 24, 32), "blocks_per_cu": (4, 5, 8, 16),
             "sort_bits": (0, 8, 16, 24, 32), "queue_rays": (65536, 262144, 1048576, 4194304, 0),
             "pass_overlap": (3, 2, 1, 0)}
    cur = {"path_tail": 2, "lds_stack_depth": 16, "blocks_per_cu": 8, "sort_bits": 0, "queue_rays": 0, "pass_overlap": 3}
    tune_passes = 6
    table, seen = [], {}

    def measure(cfg):
        key = tuple(cfg[k] for k in space)
        if key in seen:
            return seen[key]
        for k, v in cfg.items():
            b.set_option(k, v)
        b.restart()
        b.render_pass()  # (re)build, graph capture and warm-up
        b.synchronize()
        b.restart()
        b.reset_stats()
        t2 = time.perf_counter()
        for _ in range(tune_passes):
            b.render_pass()
        b.synchronize()
        mr = round(total(b.ray_counts()) / (time.perf_counter() - t2) / 1e6, 2)
        seen[key] = mr
        table.append(list(key) + [mr])
        return mr

    best_mr = measure(cur)
    for _ in range(2):
        for knob, values in space.items():
            for v in values:
                cfg = dict(cur, **{knob: v})
                mr = measure(cfg)
                if mr > best_mr:
                    best_mr, cur = mr, cfg
    for k, v in cur.items():
        b.set_option(k, v)
    b.restart()
    b.render_pass()
    b.synchronize()
    b.restart()
    b.reset_stats()
    t3 = time.perf_counter()
    b.render_all()
    b.synchronize()
    t_batch = time.perf_counter() - t3
    rays_batch = total(b.ray_counts())
    out = {
        "metric": "Mrays/s, 64 scenes: batched vs rendered one after another",
        "batched_mrays": round(rays_batch / t_batch / 1e6, 2), "plain_mrays": round(rays_plain / t_plain / 1e6, 2),
        "speedup": round(t_plain / t_batch, 3), "batched_s": round(t_batch, 4), "plain_s": round(t_plain, 4),
        "rays": rays_batch, "rays_plain": rays_plain, "passes": b.passes,
        "plain_overlapped_mrays": round(rays_overlap / t_overlap / 1e6, 2),
        "speedup_vs_overlapped": round(rays_batch / t_batch / (rays_overlap / t_overlap), 3),
        "knobs": dict(cur, use_graph=1),
        "plain_knobs": knobs,
        "autotune": {"method": f"coordinate descent from the defaults, two rounds, {tune_passes} passes per setting",
                     "columns": list(space) + [f"mrays_first_{tune_passes}_passes"], "table": table},
        "config": {"workload": f"{len(items)} icosphere soups, {sum(s.triangle_count() for s in ss)} triangles, "
                               f"{sum(i.width * i.height for i in items)} pixels, 4..32 spp, depth 4..8",
                   "scenes": len(items), "trees": "8-wide" if wide else "binary"},
    }
    log(f"batch: {out}")
    b.close()
    return out


def run_large(args, lib, A, scenes, knobs, wide):
    """The large two-level workload (scenes.large_instanced: 4097 instances of
    16 textured meshes, 32 sRGB textures of 2048^2, 3840x2160, 16 spp, depth
    12 through the max_bounces extension): Mrays/s of its passes; texel-gradient
    descent on 4 of its textures against a target rendered from them perturbed
    with Gaussian noise (sigma 0.1), s per iteration; and the scene rendered in
    one batch with 16 small soups against each rendered alone."""
    import numpy as np
    import torch
    from raydemo_amd import adjoint

    t0 = time.time()
    r = A.Renderer(lib, args.large_width, args.large_height, A.RENDERER_HIP, use_wide_bvh=wide)
    for k, v in knobs.items():
        r.set_option(k, v)
    r.set_option("max_bounces", 12)
    s = r.create_scene()
    # the library's own build time (scene calls and the lazy tree builds),
    # apart from the generation of the synthetic data in Python
    lib_time = [0.0]
    timed = {}
    for name in ("add_texture", "add_material", "add_mesh", "add_mesh_instance", "add_light", "add_camera"):
        f = getattr(s, name)
        def wrap(*a, _f=f, **k):
            t = time.perf_counter()
            try:
                return _f(*a, **k)
            finally:
                lib_time[0] += time.perf_counter() - t
        timed[name] = wrap
        setattr(s, name, wrap)
    spec = scenes.large_instanced(s, width=args.large_width, height=args.large_height, spp=args.large_spp)
    for name in timed:
        delattr(s, name)
    t_fin = time.perf_counter()
    s.triangle_count()
    lib_time[0] += time.perf_counter() - t_fin
    s.view()
    t_build = time.time() - t0
    log(f"large: {spec.info['instances']} instances, {spec.info['instanced_triangles']} instanced triangles "
        f"({spec.info['base_triangles']} in the meshes), host build {t_build:.1f}s")
    W, H, spp = args.large_width, args.large_height, args.large_spp

    def total(rc):
        return rc["primary"] + rc["secondary"] + rc["shadow"]

    reg = A.Region(lib, 0, 0, W, H)
    r.render(s, reg)  # upload + warm-up
    r.synchronize()
    # a short sweep of the knobs that matter at this depth and size (2 passes each)
    sweep = []
    for cfg in ({"path_tail": -1, "lds_stack_depth": 16, "trace_occupancy": 0, "trace_group": 0},
                {"path_tail": -1, "lds_stack_depth": 8, "trace_occupancy": 5, "trace_group": 0},
                {"path_tail": -1, "lds_stack_depth": 16, "trace_occupancy": 5, "trace_group": 0},
                {"path_tail": -1, "lds_stack_depth": 24, "trace_occupancy": 0, "trace_group": 0},
                {"path_tail": -1, "lds_stack_depth": 16, "trace_occupancy": 0, "trace_group": 4},
                {"path_tail": -1, "lds_stack_depth": 16, "trace_occupancy": 0, "trace_group": 8},
                {"path_tail": 8, "lds_stack_depth": 16, "trace_occupancy": 0, "tail_occupancy": 3, "trace_group": 0},
                {"path_tail": knobs.get("path_tail", 2), "lds_stack_depth": knobs.get("lds_stack_depth", 8),
                 "trace_occupancy": knobs.get("trace_occupancy", 0), "tail_occupancy": knobs.get("tail_occupancy", 2),
                 "trace_group": knobs.get("trace_group", 0)}):
        for k, v in cfg.items():
            r.set_option(k, v)
        r.reset_stats()
        t = time.perf_counter()
        for _ in range(2):
            r.render(s, reg)
        r.synchronize()
        sweep.append((total(r.ray_counts()) / (time.perf_counter() - t) / 1e6, cfg))
    best_cfg = max(sweep, key=lambda x: x[0])[1]
    for k, v in best_cfg.items():
        r.set_option(k, v)
    log(f"large knobs: {[(round(m, 1), c) for m, c in sweep]} -> {best_cfg}")
    r.clear()
    r.reset_stats()
    reg = A.Region(lib, 0, 0, W, H)
    t1 = time.perf_counter()
    for _ in range(spp):
        r.render(s, reg)
    r.synchronize()
    dt = time.perf_counter() - t1
    rays = total(r.ray_counts())
    out = {"mrays": round(rays / dt / 1e6, 2), "seconds": round(dt, 4), "rays": rays,
           "config": {"workload": f"large_instanced: {spec.info['instances']} instances, "
                                  f"{spec.info['instanced_triangles']} instanced triangles, 32 sRGB textures "
                                  f"2048^2 with mips, {W}x{H}, {spp} spp, depth 12",
                      "trees": "8-wide" if wide else "binary", "host_build_s": round(t_build, 1),
                      "library_build_s": round(lib_time[0], 2),
                      "synthetic_data_s": round(t_build - lib_time[0], 1), "knobs": best_cfg}}
    log(f"large render: {out['mrays']} Mrays/s ({spp} passes in {dt:.3f}s)")

    # an animation step: every instance moved (SetMeshInstanceTransform), then
    # one pass, which rebuilds the instance tree on the device and replaces
    # only the instance arrays there; against one pass with nothing changed
    # and one after an edit that re-uploads the whole scene
    def one_pass():
        r.synchronize()
        t = time.perf_counter()
        r.render(s, reg)
        r.synchronize()
        return time.perf_counter() - t

    v = s.view()
    mis, trs = v["mesh_instances"], v["transforms"]
    orig_x = [trs[mis[i]["tr_index"]]["xform"].copy() for i in range(len(mis))]
    def move(dy):
        for i, x in enumerate(orig_x):
            y = x.copy()
            y[13] += dy
            s.set_mesh_instance_transform(i, y)
    base_s = min(one_pass() for _ in range(2))
    updates0, uploads0 = r.get_option("instance_updates"), r.get_option("scene_uploads")
    moved = []
    for dy in (0.25, 0.0):
        move(dy)
        moved.append(one_pass())
    counted = {"instance_updates": r.get_option("instance_updates") - updates0,
               "full_uploads": r.get_option("scene_uploads") - uploads0}
    s.set_environment((0.3, 0.35, 0.4))  # an edit outside the instances: the next pass uploads everything
    full_s = one_pass()
    mi_boxes = np.concatenate([mis["bbox_min"], mis["bbox_max"]], axis=1)
    out["instance_update"] = {
        "metric": "one pass after moving every instance, ms (instance tree rebuilt on the device, instance arrays "
                  "replaced) vs after a full scene upload",
        "instances": len(mis), "pass_ms_unchanged": round(base_s * 1e3, 2),
        "pass_ms_after_move": round(min(moved) * 1e3, 2), "pass_ms_after_full_upload": round(full_s * 1e3, 2),
        "update_ms": round((min(moved) - base_s) * 1e3, 2), "full_upload_ms": round((full_s - base_s) * 1e3, 2),
        "moves_counted": counted, "instance_tree_on_device": A.build_sah(lib, mi_boxes, device=0)[2]}
    log(f"large instance update: {out['instance_update']}")

    # texel-gradient descent on 4 textures
    if args.large_adj_steps > 0:
        dr = adjoint.DiffRenderer(lib, r, s, W, H, spp, len(s.view()["materials"]), tape_texels=True)
        ids = tuple(spec.info["adjoint_textures"])
        orig = np.stack(spec.info["adjoint_texels"]).astype(np.float32) / 255.0
        noise = np.random.default_rng(scenes.SEED_BASE + 55).normal(0.0, 0.1, orig.shape).astype(np.float32)
        target_tex = torch.tensor(np.clip(orig + noise, 0.0, 1.0), device="cuda")
        target = adjoint.render_textures(target_tex, dr, ids).detach().clone()
        def descent(steps):
            params = torch.tensor(orig, device="cuda").requires_grad_(True)
            opt = torch.optim.Adam([params], lr=0.01)
            losses = []
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for _ in range(steps):
                opt.zero_grad(set_to_none=True)
                img = adjoint.render_textures(params, dr, ids)
                loss = torch.mean((img - target) ** 2)
                loss.backward()
                opt.step()
                with torch.no_grad():
                    params.clamp_(0.0, 1.0)
                losses.append(float(loss.detach()))
            torch.cuda.synchronize()
            return (time.perf_counter() - t2) / steps, losses, int(params.numel())

        s_iter, losses, n_params = descent(args.large_adj_steps)
        used_replay = device_used_bytes()
        # the tape beside it (the forward keeps every pass's vertex and texel records)
        dr.tape = True
        dr.r.set_option("adjoint_tape", dr.spp)
        dr.r.set_option("adjoint_tape_texels", 1)
        tape_steps = max(1, args.large_adj_steps // 2)
        s_tape, _, _ = descent(tape_steps)
        used_tape = device_used_bytes()
        tape_bytes = dr.tape_bytes()
        dr.tape = False
        dr.r.set_option("adjoint_tape", 0)
        out["texel_adjoint"] = {
            "metric": "texel-gradient descent s/iter (forward + path-replay adjoint + fold + Adam)",
            "s_per_iter": round(s_iter, 4), "steps": args.large_adj_steps,
            "parameters": n_params, "loss_first": losses[0], "loss_last": losses[-1],
            "textures": list(ids), "tape": False, "replay_record_bytes": dr.replay_bytes(),
            "device_used_bytes_replay": used_replay, "tape_s_per_iter": round(s_tape, 4), "tape_steps": tape_steps,
            "tape_record_bytes": tape_bytes, "device_used_bytes_tape": used_tape}
        log(f"large texel adjoint: {out['texel_adjoint']}")
        # back to the generator's textures for the batch leg
        adjoint.render_textures(torch.tensor(orig, device="cuda"), dr, ids)
    r.close()

    # the scene in one batch with 16 small soups, against each alone
    if args.large_batch_scenes > 0:
        items = scenes.batch_workload(args.large_batch_scenes, seed=scenes.SEED_BASE + 3)
        builds = [(None, W, H, spp)] + [
            (lambda sc, it=it: scenes.batch_scene(sc, it), it.width, it.height, it.spp) for it in items]
        ss = [s]  # the large scene itself (its textures are back to the generator's)
        for build, w, h, n in builds[1:]:
            sc = A.standalone_scene(lib, wide)
            build(sc)
            sc.triangle_count()
            ss.append(sc)
        rs = []
        for sc, (build, w, h, n) in zip(ss, builds):
            rr = A.Renderer(lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
            for k, v in (best_cfg if sc is s else knobs).items():
                rr.set_option(k, v)
            rr.set_option("max_bounces", 12)
            rr.render(sc, A.Region(lib, 0, 0, w, h))
            rr.synchronize()
            rr.clear()
            rr.reset_stats()
            rs.append(rr)
        t3 = time.perf_counter()
        for rr, sc, (build, w, h, n) in zip(rs, ss, builds):  # scene after scene
            rg = A.Region(lib, 0, 0, w, h)
            for _ in range(n):
                rr.render(sc, rg)
            rr.synchronize()
        t_plain = time.perf_counter() - t3
        rays_plain = sum(total(rr.ray_counts()) for rr in rs)
        for rr in rs:
            rr.close()
        b = A.Batch(lib)
        b.set_option("max_bounces", 12)
        for sc, (build, w, h, n) in zip(ss, builds):
            b.add(sc, w, h, n)
        # batch knobs: the 64-scene workload's, or every bounce staged (deep paths)
        bsweep = []
        base_b = dict(args.large_batch_knobs or {})
        for cfg in (base_b, dict(base_b, path_tail=-1), dict(base_b, path_tail=-1, blocks_per_cu=5, lds_stack_depth=16),
                    dict(base_b, path_tail=-1, blocks_per_cu=8), dict(base_b, path_tail=8, blocks_per_cu=5),
                    dict(base_b, path_tail=-1, lds_stack_depth=8),
                    # per-scene queue sort (the 4K scene's bounces gain from coherent rays)
                    dict(base_b, path_tail=-1, lds_stack_depth=16, sort_bits=24),
                    dict(base_b, path_tail=-1, lds_stack_depth=16, sort_bits=24, blocks_per_cu=5),
                    dict(base_b, path_tail=-1, lds_stack_depth=16, sort_bits=16),
                    dict(base_b, path_tail=-1, lds_stack_depth=16, sort_bits=24, blocks_per_cu=8, pass_overlap=3)):
            for k, v in cfg.items():
                b.set_option(k, v)
            b.restart()
            b.render_pass()  # graph capture and warm-up
            b.synchronize()
            b.restart()
            b.reset_stats()
            t = time.perf_counter()
            for _ in range(2):
                b.render_pass()
            b.synchronize()
            bsweep.append((total(b.ray_counts()) / (time.perf_counter() - t), cfg))
        bcfg = max(bsweep, key=lambda x: x[0])[1]
        for k, v in bcfg.items():
            b.set_option(k, v)
        b.restart()
        b.render_pass()
        b.synchronize()
        b.restart()
        b.reset_stats()
        t4 = time.perf_counter()
        b.render_all()
        b.synchronize()
        t_batch = time.perf_counter() - t4
        rays_batch = total(b.ray_counts())
        b.close()
        out["batch"] = {"scenes": len(builds), "knobs": bcfg, "batched_mrays": round(rays_batch / t_batch / 1e6, 2),
                        "plain_mrays": round(rays_plain / t_plain / 1e6, 2), "speedup": round(t_plain / t_batch, 3),
                        "rays": rays_batch, "rays_plain": rays_plain}
        log(f"large batch: {out['batch']}")
    return out


def run_build(args, lib, A, scenes):
    """The fast (HLBVH) build, mesh_desc_t::use_fast_bvh_build, of the tree
    over the triangle boxes of a soup of about 1M triangles: on the host, and
    with the treelets built on the GPU (the SAH top tree over at most 4096
    treelet roots stays on the host); best of 3 each, identical output
    required.  Also the whole AddMesh of that soup (triangle records, vertices
    and tangent frames included) for a standalone scene and a renderer's."""
    import numpy as np
    attrs, idx, spans, kind, albedo, rough = scenes.soup_geometry(scenes.SEED_BASE + 9, args.build_spheres, 38)
    n_tris = len(idx) // 3
    P = attrs[:, :3][idx].reshape(-1, 3, 3)
    boxes = np.ascontiguousarray(np.concatenate([P.min(1), P.max(1)], axis=1), np.float32)
    out = {"triangles": n_tris}
    res = {}
    for where, dev in (("host", -1), ("gpu", 0)):
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            res[where] = A.build_fast_bvh(lib, boxes, device=dev)
            ts.append(time.perf_counter() - t)
        out[f"{where}_ms"] = round(min(ts) * 1e3, 2)
    out["identical"] = all(a.tobytes() == b.tobytes() for a, b in zip(res["host"], res["gpu"]))
    out["nodes"] = len(res["host"][0])
    out["speedup"] = round(out["host_ms"] / out["gpu_ms"], 2)
    out["mprims_per_s_gpu"] = round(n_tris / out["gpu_ms"] / 1e3, 1)
    r = A.Renderer(lib, 16, 16, A.RENDERER_HIP, use_wide_bvh=False)
    for where in ("host", "gpu"):
        sc = A.standalone_scene(lib, False) if where == "host" else r.create_scene()
        tex = sc.add_texture(np.full((1, 1, 4), 255, np.uint8), is_srgb=False, generate_mipmaps=False)
        mat = sc.add_material(A.DiffuseMaterial, tex, main_color=(0.5, 0.5, 0.5))
        t = time.perf_counter()
        if sc.add_mesh(A.PxyzNxyzTuv, attrs, idx, [(mat, 0, len(idx))], use_fast_bvh_build=True) == A.INVALID:
            raise RuntimeError("fast build failed")
        out[f"add_mesh_{where}_ms"] = round((time.perf_counter() - t) * 1e3, 1)
    r.close()
    log(f"fast build: {out}")
    return out


def run_lightmap(args, lib, A, scenes, knobs, wide):
    """Lightmap baking with a geometry camera and SH output: the Cornell box
    (faces as 16 x 16 grids, 8,706 triangles, each in its own lightmap cell)
    baked into a 1024^2 lightmap, 16 passes, depth 4; Mrays/s of the passes
    (secondary + shadow rays: the geometry camera traces no primary rays)."""
    w = h = args.lightmap_size
    r = A.Renderer(lib, w, h, A.RENDERER_HIP, use_wide_bvh=wide)
    for k, v in knobs.items():
        r.set_option(k, v)
    s = r.create_scene()
    spec = scenes.lightmap_box(s, k=16, width=w, height=h, spp=16, depth=4)
    reg = A.Region(lib, 0, 0, w, h)
    r.render(s, reg)  # upload + warm-up
    r.synchronize()
    r.clear()
    r.reset_stats()
    reg = A.Region(lib, 0, 0, w, h)
    t = time.perf_counter()
    for _ in range(spec.spp):
        r.render(s, reg)
    r.synchronize()
    dt = time.perf_counter() - t
    rc = r.ray_counts()
    rays = rc["primary"] + rc["secondary"] + rc["shadow"]
    sh = r.sh_data()
    out = {"metric": "lightmap bake Mrays/s (geometry camera + SH output)", "mrays": round(rays / dt / 1e6, 2),
           "ms_per_pass": round(dt / spec.spp * 1e3, 3), "texels": w * h, "triangles": spec.triangles,
           "spp": spec.spp, "sh_nonzero": bool(sh is not None and float(abs(sh).max()) > 0)}
    r.close()
    log(f"lightmap: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-spp", type=int, default=64, help="full-frame passes of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune-passes", type=int, default=24,
                    help="passes per autotuner candidate (enough that the passes in flight fill and drain "
                         "only at its ends)")
    ap.add_argument("--knobs", default="",
                    help="skip the autotuner: lds_stack_depth,sort_rays,sort_bits,path_tail"
                         "[,tail_occupancy[,refill[,trace_occupancy[,shadow_occupancy[,tail_refill"
                         "[,pass_overlap[,trace_group]]]]]]]")
    ap.add_argument("--only-tuned", action="store_true", help="time only the tuned (or given) knobs")
    ap.add_argument("--trees", default="both", choices=("both", "binary", "wide"),
                    help="BVH kind(s) to time: binary, the reference's default 8-wide, or both (the faster is the "
                         "headline)")
    ap.add_argument("--adj-steps", type=int, default=200, help="Adam steps of the differentiation workload (0: skip)")
    ap.add_argument("--adj-size", type=int, default=512)
    ap.add_argument("--adj-spp", type=int, default=32)
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batch workload (0: skip)")
    ap.add_argument("--large", type=int, default=1, help="the large two-level workload (0: skip)")
    ap.add_argument("--large-width", type=int, default=3840)
    ap.add_argument("--large-height", type=int, default=2160)
    ap.add_argument("--large-spp", type=int, default=16)
    ap.add_argument("--large-adj-steps", type=int, default=5, help="texel-gradient descent steps (0: skip)")
    ap.add_argument("--large-batch-scenes", type=int, default=16, help="small scenes batched with it (0: skip)")
    ap.add_argument("--lightmap-size", type=int, default=1024, help="lightmap bake leg (0: skip)")
    ap.add_argument("--build-spheres", type=int, default=3000,
                    help="icospheres of the fast-build mesh (about 390 triangles each; 0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    args.large_batch_knobs = None
    if args.adj_steps > 0 or args.large:
        import torch  # noqa: F401  (loaded before the HIP libraries: one HIP runtime in the process)
    from raydemo_amd import autotune, capi as A, host_lib, scenes

    lib = host_lib()
    kinds = {"both": (False, True), "binary": (False,), "wide": (True,)}[args.trees]
    runs = {}
    for wide in kinds:
        runs[wide] = run_gpu(args, lib, A, scenes, wide, staged_modes=False)
    wide = max(runs, key=lambda k: runs[k][1]["tuned"]["mrays"])
    per_tree = {("8-wide" if k else "binary"): round(v[1]["tuned"]["mrays"], 2) for k, v in runs.items()}
    log(f"trees: {per_tree}; headline on the {'8-wide' if wide else 'binary'} trees")
    if not args.only_tuned:
        # the staged modes (every bounce its own launches; sort off / 32-bit sort) on the faster trees
        first = runs[wide][4]
        runs[wide] = run_gpu(args, lib, A, scenes, wide, staged_modes=True, tuned=first)
    spec, res, vis, traced, tuned = runs[wide]
    best = res["tuned"]
    # Kernel durations for the roofline come from the tuned knobs with one pass
    # in flight: with two, a kernel shares the chip with the other pass's
    # kernels and its event-timed duration is not its own.  The same passes
    # render the same rays, so the replay's visit counts fit both.
    ser = res.get("serial", best)
    prof = ser["prof"]

    # roofline of the trace kernel (closest hit): bytes of the reference
    # structures the reference traversal reads, per visit
    trace_ms = prof["trace_ms"]
    b_node = B_WNODE if wide else B_NODE
    bytes_alg = (traced * (B_RAY + B_HIT) + vis["node_visits"] * b_node + vis["instance_tests"] * B_INST +
                 vis["tri_tests"] * B_TRI)
    achieved = bytes_alg / (trace_ms / 1e3) / 1e9 if trace_ms > 0 else 0.0
    min_blocks = 5 if (wide and tuned.best.get("trace_occupancy", 0) == 5) else 1
    kernel = f"k_trace<false,false,{1 if wide else 0},{min_blocks}>"
    traffic_file = ROOT / "profiles" / "trace_traffic.json"
    traffic = None
    if traffic_file.exists():
        try:
            rec = json.loads(traffic_file.read_text()).get(kernel) or {}
            traffic = rec.get("bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": kernel, "trees": "8-wide" if wide else "binary", "launches": prof["trace_launches"],
        "avg_launch_ms": round(trace_ms / max(prof["trace_launches"], 1), 4),
        "alg_bytes_per_launch": int(bytes_alg / max(prof["trace_launches"], 1)),
        "measured_in": "serial" if ser is not best else "tuned",
        "overlapped_avg_launch_ms": round(best["prof"]["trace_ms"] / max(best["prof"]["trace_launches"], 1), 4),
    }
    # the trace launches by stage (the primary rays, then each staged bounce):
    # the same byte model over each stage's counted visits and launch time
    vis_st = {st["stage"]: st for st in vis.get("trace_stages", [])}
    per_stage = []
    for st in prof.get("trace_stages", []):
        cv = vis_st.get(st["stage"])
        if not cv or not st["launches"] or st["ms"] <= 0:
            continue
        b = (cv["rays"] * (B_RAY + B_HIT) + cv["nodes"] * b_node + cv["instances"] * B_INST + cv["tris"] * B_TRI)
        per_stage.append({"stage": "primary" if st["stage"] == 0 else f"bounce {st['stage'] - 1}",
                          "launches": st["launches"], "avg_launch_ms": round(st["ms"] / st["launches"], 4),
                          "rays_per_launch": int(cv["rays"] / st["launches"]),
                          "alg_bytes_per_launch": int(b / st["launches"]),
                          "achieved": round(b / (st["ms"] / 1e3) / 1e9, 1),
                          "frac": round(b / (st["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)})
    roofline["per_stage"] = per_stage
    # the other traversal kernels from the same replay's counters: the shadow
    # kernel per queue entry, the path tail per shaded segment and light sample
    sh, tl = vis["shadow"], vis["tail"]
    per_kernel = [{k: v for k, v in roofline.items() if k != "per_stage"}]
    if prof["shadow_launches"]:
        b_sh = sh["entries"] * B_SHADOW_ENTRY + traversal_bytes(sh, wide)
        per_kernel.append(kernel_roofline("k_shadow", b_sh, prof["shadow_ms"], prof["shadow_launches"]))
    if prof["tail_launches"]:
        b_tl = tl["segments"] * B_SHADE + tl["shadow_segments"] * B_NEE_RECORD + traversal_bytes(tl, wide)
        per_kernel.append(kernel_roofline("k_path_tail", b_tl, prof["tail_ms"], prof["tail_launches"]))
    roofline["per_kernel"] = per_kernel

    cpu = None
    if not args.no_cpu:
        runs = {}
        for flags in (A.RENDERER_REF, A.RENDERER_AVX2):
            run = run_cpu_baseline(args, A, scenes, args.cpu_spp, flags)
            if run is not None:
                runs[run["backend"]] = run
        if runs:
            # rays of the identical sample, counted on the device (the Ref backend's paths and seeds;
            # the AVX2 backend traces the same rays up to its packets' arithmetic)
            r = A.Renderer(lib, args.width, args.height, A.RENDERER_HIP, use_wide_bvh=True)
            s = r.create_scene()
            scenes.triangle_soup(s, width=args.width, height=args.height, depth=args.depth)
            reg = A.Region(lib, 0, 0, args.width, args.height)
            r.reset_stats()
            for _ in range(args.cpu_spp):
                r.render(s, reg)
            r.synchronize()
            rc = r.ray_counts()
            rays = rc["primary"] + rc["secondary"] + rc["shadow"]
            r.close()
            per = {name: {"value": round(rays / run["seconds"] / 1e6, 3), "seconds": round(run["seconds"], 2)}
                   for name, run in runs.items()}
            head = runs.get("Ref") or next(iter(runs.values()))
            th = host_threads()
            cpu = {"value": per[head["backend"]]["value"], "unit": "Mrays/s", "cores": head["cores"],
                   "kind": head["kind"], "backend": head["backend"], "backends": per,
                   "cpu_model": cpu_model(), "nproc": th["nproc"], "affinity": th["affinity"],
                   "sample": f"{args.cpu_spp} full-frame 1920x1080 passes (spp 1..{args.cpu_spp}) of the same scene, "
                             f"128x128 buckets on {head['cores']} threads (every allowed core, at most 16: one GPU's "
                             f"host share), 8-wide BVH (the reference default); the reference's Ref (scalar, the "
                             f"parity target) and AVX2 (4x2 ray packets) backends"}
            log(f"cpu baseline: {cpu}")

    plain_knobs = {k: v for k, v in tuned.best.items() if k in autotune.KNOBS}
    adj = run_adjoint(args, lib, A, scenes, plain_knobs, wide) if args.adj_steps > 0 else None
    batch = run_batch(args, lib, A, scenes, plain_knobs, wide) if args.batch_scenes > 0 else None
    if batch is not None:
        args.large_batch_knobs = {k: v for k, v in batch["knobs"].items() if k != "use_graph"}
    large = run_large(args, lib, A, scenes, plain_knobs, wide) if args.large else None
    build = run_build(args, lib, A, scenes) if args.build_spheres > 0 else None
    lightmap = run_lightmap(args, lib, A, scenes, plain_knobs, wide) if args.lightmap_size > 0 else None

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
                               f"{args.steps} spp, depth {args.depth}, autotuned launch knobs",
                   "triangles": spec.triangles, "width": args.width, "height": args.height, "spp": args.steps,
                   "depth": args.depth, "trees": "8-wide" if wide else "binary", "knobs": tuned.best,
                   "parallelism": "single GPU"},
        "trees": per_tree,
        "modes": {name: {"mrays": round(m["mrays"], 2), "opts": m["opts"],
                         "kernel_ms": {k: round(v, 3) for k, v in m["prof"].items() if k.endswith("_ms")}}
                  for name, m in res.items()},
        "rays_per_pass": int(best["rays"] / args.steps), "ray_counts": best["counts"],
        "overflow": best["counts"]["overflow"],
        "autotune": {"seconds": round(tuned.seconds, 2), "passes_per_candidate": args.tune_passes,
                     "columns": list(autotune.KNOBS) + ["passes", "mrays"],
                     "table": [[t[k] for k in autotune.KNOBS] + [t.get("passes"), t["mrays"]] for t in tuned.table]},
        "kernel_ms": {k: round(v, 3) for k, v in prof.items() if k.endswith("_ms")},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adj,
        "batch": batch,
        "large": large,
        "fast_build": build,
        "lightmap": lightmap,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
