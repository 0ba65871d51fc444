#!/usr/bin/env python3
"""Benchmark: path-traced frames of the 1024x1024x64spp displaced-sphere
Cornell box (69,620-triangle GGX mesh + box, depth 16) on one MI355X.

A step is one full frame (PathTracingIntegrator.Render) with the scene
already resident in HBM.  Prints one JSON line:
  value          Mrays/s = (extension + shadow rays of one frame) * steps / time
  roofline       dominant kernel (trace_closest): algorithmic bytes per launch
                 (48 per ray + 32 per node + 36 per triangle + 48 per instance
                 visited) / HIP-event launch duration, against 8 TB/s HBM;
                 traffic = HBM bytes per launch from the committed rocprofv3
                 FETCH_SIZE / WRITE_SIZE summary
  batched        32 seeded scenes rendered one after the other (plain) and as
                 one batch; speedup = plain time / batched time
  adjoint        inverse rendering of the 4096-patch scene: Adam iterations
                 (forward + path-replay backward) per second, and their time
                 over the forward render's
  pssmlt         the caustic scene: mutations per second and acceptance
  rooflines      the same for trace_shadow and shade_c (192 B per path
                 vertex), each with its L2 hit rate (committed TCC_HIT / TCC_MISS
                 summary) and measured-traffic fraction
  cpu_baseline   the CPU oracle (C++ restatement of the F# renderer) on a
                 bounded pixel window of the same frame, on one GPU's share of
                 the host (16 threads), with the CPU model; plus bounded CPU
                 legs of PSSMLT (config 4) and the adjoint (config 3)
  parity         GPU vs oracle on a 128x128x2spp render of the same scene

--copy-calibration additionally copies 1 GiB device to device (hipMemcpy) so
that a rocprofv3 --pmc run of this script sees FETCH_SIZE / WRITE_SIZE of a
copy of known size next to the trace kernels.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from barnacle_amd import configs, native, scene  # noqa: E402

HBM_PEAK_GBS = 8000.0


def closest_bytes(st):
    return 48 * st["rays_closest"] + 32 * st["nodes_closest"] + 36 * st["tris_closest"] + 48 * st["insts_closest"]


def shadow_bytes(st):
    return 36 * st["rays_shadow"] + 32 * st["nodes_shadow"] + 36 * st["tris_shadow"] + 48 * st["insts_shadow"]


def shade_bytes(st):
    return 192 * st["vertices"]


def frame_bytes(st):
    return closest_bytes(st) + shadow_bytes(st) + shade_bytes(st) + 12 * st["paths"]


def host_cpu():
    """Host threads this process may use and the CPU model (/proc/cpuinfo)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    lscpu = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)"):
                lscpu[k.strip()] = v.strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    # One GPU's share of the GPU box's host is 16 threads; use all of them.
    return {"threads": min(16, avail), "available_threads": avail, "cpu_count": os.cpu_count(), "model": model,
            "lscpu": lscpu}


def cpu_baseline(sc, params, gpu_film=None, budget_s=12.0):
    """Oracle rays/s on 64x64 pixel windows of the same frame, centre outwards.
    With the GPU's film of that frame, the windows also give a full-size
    parity check: rel-RMSE and the bit-equal fraction over their pixels."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cpu = host_cpu()
    threads = cpu["threads"]
    osc = oracle.Oracle().scene(sc)
    W, H = params.width, params.height
    tiles = sorted(((x, y) for y in range(0, H, 64) for x in range(0, W, 64)),
                   key=lambda t: (t[0] + 32 - W / 2) ** 2 + (t[1] + 32 - H / 2) ** 2)
    rays, paths, elapsed, n = 0, 0, 0.0, 0
    got, ref = [], []
    for (x, y) in tiles:
        x1, y1 = min(x + 64, W), min(y + 64, H)
        t0 = time.perf_counter()
        film, st = osc.render(params, threads=threads, window=(x, y, x1, y1))
        elapsed += time.perf_counter() - t0
        if gpu_film is not None:
            # Film.SetPixel stores row y at H - y - 1 (Film.fs:43)
            rows = np.arange(H - y1, H - y)[:, None] * W + np.arange(x, x1)[None, :]
            got.append(gpu_film[rows.reshape(-1)])
            ref.append(film[rows.reshape(-1)])
        rays += st["rays_closest"] + st["rays_shadow"]
        paths += st["paths"]
        n += 1
        if elapsed >= budget_s:
            break
    out = {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port", "host": cpu,
           "sample": f"{n} 64x64-pixel windows at the film centre of the same 1024x1024x64spp frame "
                     f"({paths} paths, {rays} rays, {elapsed:.1f} s); oracle = C++ restatement of the F# "
                     f"renderer (no .NET runtime available)",
           "mpaths_per_s": paths / elapsed / 1e6}
    if got:
        g, o = np.concatenate(got), np.concatenate(ref)
        ok = np.isfinite(o).all(axis=1)
        out["frame_parity"] = {"pixels": int(len(o)),
                               "rel_rmse": float(np.sqrt(np.sum((g[ok] - o[ok]) ** 2) / np.sum(o[ok] ** 2))),
                               "bit_equal_fraction": float(np.mean(g.view(np.int32) == o.view(np.int32))),
                               "nonfinite_masks_equal": bool(np.array_equal(np.isfinite(g), np.isfinite(o)))}
    return out


def batched(ctx, n=32, pause_s=0.0, small=True):
    """Plain per-scene renders against one batched render of the same items.
    pause_s > 0 idles the GPU that long between the warm-up and timed phases,
    so that a kernel trace of the run splits into them at its longest gaps."""
    t0 = time.perf_counter()
    scs = [scene.load_scene(j) for j in configs.config2(n)]
    load_s = time.perf_counter() - t0
    ctx.set_option("bvh_gpu", 0)                          # the same scene builds with every BVH on the host
    t0 = time.perf_counter()
    for s in scs:
        ctx.scene(s).release()
    host_build_s = time.perf_counter() - t0
    ctx.set_option("bvh_gpu", 1)
    t0 = time.perf_counter()
    ds = [ctx.scene(s) for s in scs]                      # BLAS built on the GPU
    build_s = time.perf_counter() - t0
    setup = load_s + build_s
    jobs = [(d, s.pt_params()) for d, s in zip(ds, scs)]
    for d, p in jobs:
        d.render_pt(p)
    time.sleep(pause_s)
    t0 = time.perf_counter()
    plain = [d.render_pt(p) for d, p in jobs]
    plain_s = time.perf_counter() - t0
    time.sleep(pause_s)
    ctx.render_batch(jobs)
    time.sleep(pause_s)
    t0 = time.perf_counter()
    films, st = ctx.render_batch(jobs)
    batch_s = time.perf_counter() - t0
    time.sleep(pause_s)
    equal = all(np.array_equal(a.view(np.int32), b.view(np.int32)) for a, b in zip(films, plain))
    paths = sum(p.width * p.height * p.spp for _, p in jobs)
    return {"config": f"config2: {n} seeded scenes, 1-256 instances of 1k-500k-triangle displaced icospheres, "
                      "256^2-1024^2, 4-64 spp, depth 4-16",
            "paths": paths, "plain_s": plain_s, "batched_s": batch_s, "speedup": plain_s / batch_s,
            "batched_mpaths_per_s": paths / batch_s / 1e6, "rounds": st["rounds"], "graph": st["graph"],
            "films_bit_equal": equal, "scene_setup_s": setup, "scene_load_s": load_s, "scene_build_s": build_s,
            "scene_build_host_bvh_s": host_build_s,
            "small_items": small_items(ctx) if small else None, "_scenes": scs}


def bvh_builds(ctx, scs):
    """BLAS construction (BVHNode.Build over triangle bounds, Mesh.fs:159-169)
    of every mesh of the given scenes: the host build and the GPU build, one
    call per mesh each (the GPU time includes copying the bounds in and the
    nodes and permutation out); nodes and permutations must be equal."""
    sets = []
    for sc in scs:
        v = sc.vertices.reshape(-1, 3)
        for pr in sc.primitives:
            if not pr.get("n_indices"):
                continue
            ix = sc.indices[pr["index_offset"]:pr["index_offset"] + pr["n_indices"]].reshape(-1, 3)
            tri = v[pr["vert_offset"]:pr["vert_offset"] + pr["n_verts"]][ix]
            sets.append(np.ascontiguousarray(np.concatenate([tri.min(1), tri.max(1)], 1), np.float32))
    ctx.bvh_build(sets[0])                                # allocation outside the timing
    host_t, gpu_t, equal = [], [], True
    for b in sets:
        t0 = time.perf_counter()
        hn, hp = native.bvh_build(b)
        host_t.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        gn, gp = ctx.bvh_build(b)
        gpu_t.append(time.perf_counter() - t0)
        equal = equal and gn.tobytes() == hn.tobytes() and np.array_equal(gp, hp)
    tris = [len(b) for b in sets]
    k = int(np.argmax(tris))
    return {"meshes": len(sets), "triangles": int(sum(tris)), "host_s": float(sum(host_t)),
            "gpu_s": float(sum(gpu_t)), "host_mtri_per_s": sum(tris) / sum(host_t) / 1e6,
            "gpu_mtri_per_s": sum(tris) / sum(gpu_t) / 1e6, "nodes_and_perm_equal": bool(equal),
            "largest": {"triangles": tris[k], "host_ms": host_t[k] * 1e3, "gpu_ms": gpu_t[k] * 1e3},
            "host": "one thread (the reference's build is sequential too)"}


def counter_summary(out_path, dbs, commands):
    """Per-kernel averages per dispatch of the counters in rocprofv3 result
    databases (one counter pass each; counters_collection view) of the config-1
    frame, written as the JSON that roofline() reads: FETCH_SIZE / WRITE_SIZE
    in KiB with the calibration of the run's largest device copy (a 1 GiB
    hipMemcpy reads and writes exactly 2^30 bytes), TCC_HIT_sum / TCC_MISS_sum,
    and any SQ counters.  Kernel names drop their parameter lists."""
    import re
    import sqlite3
    kern = {}
    copy = {}
    for db in dbs:
        con = sqlite3.connect(db)
        per = {}
        for kname, cname, val, did, dur in con.execute(
                "select kernel_name, counter_name, value, dispatch_id, duration from counters_collection"):
            short = re.sub(r"\(.*", "", kname).replace("void ", "").replace("bnd::", "")
            d = per.setdefault((short, cname), {})
            d[did] = d.get(did, 0.0) + val                    # summed over dimensions (XCDs, channels)
            if cname in ("FETCH_SIZE", "WRITE_SIZE") and "copyBuffer" in short:
                copy[cname] = max(copy.get(cname, 0.0), d[did])
            kern.setdefault(short, {}).setdefault("_dur", {})[did] = dur
        for (short, cname), d in per.items():
            k = kern[short]
            k[cname] = sum(d.values()) / len(d)
            k["dispatches"] = max(k.get("dispatches", 0), len(d))
    out = {"workload": "config1 1024x1024x64spp: one counting frame (<true, ...> traversal variants) and one plain "
                       "frame", "commands": commands, "kernels": {}, "copy_calibration_bytes": 1 << 30,
           "copy_kib": copy}
    for short, k in sorted(kern.items()):
        durs = k.pop("_dur")
        k["avg_ms_under_counters"] = sum(durs.values()) / len(durs) / 1e6
        if "TCC_HIT_sum" in k and "TCC_MISS_sum" in k:
            k["l2_hit_rate"] = k["TCC_HIT_sum"] / max(k["TCC_HIT_sum"] + k["TCC_MISS_sum"], 1.0)
        if "FETCH_SIZE" in k and "WRITE_SIZE" in k and copy.get("FETCH_SIZE") and copy.get("WRITE_SIZE"):
            ff, wf = copy["FETCH_SIZE"] * 1024 / (1 << 30), copy["WRITE_SIZE"] * 1024 / (1 << 30)
            k["hbm_bytes_per_launch"] = k["FETCH_SIZE"] * 1024 / ff + k["WRITE_SIZE"] * 1024 / wf
            out["fetch_factor"], out["write_factor"] = ff, wf
        out["kernels"][short] = k
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    return out


def timeline_summary(csv_path, min_phase_gap_s=0.2):
    """Idle time of the GPU in a rocprofv3 kernel trace (kernel_trace.csv) of
    `bench.py --batch-timeline`: the trace splits into phases at gaps of at
    least min_phase_gap_s (warm-up plain, timed plain, warm-up batch, timed
    batch).  Per phase: span, time with no kernel running, the gaps between
    consecutive kernels by size, and the time per kernel."""
    import csv
    with open(csv_path) as fh:
        rows = list(csv.DictReader(fh))
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in rows)
    phases, cur = [], [ks[0]]
    for k in ks[1:]:
        if k[0] - max(e for _, e, _ in cur) >= min_phase_gap_s * 1e9:
            phases.append(cur)
            cur = []
        cur.append(k)
    phases.append(cur)
    names = ["plain (warm-up)", "plain (timed)", "batch (warm-up)", "batch (timed)"]
    names = ["setup"] * max(0, len(phases) - 4) + names[-len(phases):]   # uploads come first
    out = []
    for i, ph in enumerate(phases):
        span = ph[-1][1] - ph[0][0]
        busy, end, gaps = 0, ph[0][0], []
        for s0, e0, _ in ph:                                  # union of kernel intervals
            if s0 > end:
                gaps.append(s0 - end)
            busy += max(0, e0 - max(s0, end))
            end = max(end, e0)
        bins = [(0, 10e3), (10e3, 100e3), (100e3, 1e6), (1e6, float("inf"))]
        hist = {f"{int(lo / 1e3)}-{'inf' if hi == float('inf') else int(hi / 1e3)}us":
                {"count": sum(lo <= g < hi for g in gaps), "ms": sum(g for g in gaps if lo <= g < hi) / 1e6}
                for lo, hi in bins}
        per = {}
        for s0, e0, nm in ph:
            short = nm.split("(")[0].split("<")[0].replace("void ", "").strip()
            c, t = per.get(short, (0, 0.0))
            per[short] = (c + 1, t + (e0 - s0) / 1e6)
        out.append({"phase": names[i], "kernels": len(ph),
                    "span_ms": span / 1e6, "busy_ms": busy / 1e6, "idle_ms": (span - busy) / 1e6,
                    "idle_frac": (span - busy) / span, "gaps": hist,
                    "per_kernel_ms": {k: round(v[1], 3) for k, v in sorted(per.items(), key=lambda kv: -kv[1][1])},
                    "per_kernel_launches": {k: v[0] for k, v in per.items()}})
    return out


def small_items(ctx, n=256):
    """Many small render items (the serving case batching is for): n frames
    of the Cornell box at 128x128, 4 spp, depth 8, frame ids 0..n-1, one at a
    time against one batched render."""
    sc = scene.load_scene(configs.cornell_box(128, 128, 4, 8))
    d = ctx.scene(sc)
    jobs = [(d, sc.pt_params(frame_id=f)) for f in range(n)]
    for j in jobs[:4]:
        j[0].render_pt(j[1])
    t0 = time.perf_counter()
    plain = [dd.render_pt(p) for dd, p in jobs]
    plain_s = time.perf_counter() - t0
    ctx.render_batch(jobs)
    t0 = time.perf_counter()
    films, st = ctx.render_batch(jobs)
    batch_s = time.perf_counter() - t0
    equal = all(np.array_equal(a.view(np.int32), b.view(np.int32)) for a, b in zip(films, plain))
    return {"config": f"{n} frames of the Cornell box, 128x128, 4 spp, depth 8", "plain_s": plain_s,
            "batched_s": batch_s, "speedup": plain_s / batch_s, "films_bit_equal": equal}


def adjoint(ctx, steps=200):
    """config3 inverse rendering: Adam iterations of path-replay gradients,
    the full 200-step optimisation, with the parameter errors along it."""
    import torch  # noqa: F401  (the autograd wrapper runs on host tensors)
    from barnacle_amd import diff
    info = configs.config3()
    sc = scene.load_scene(info)
    ds = ctx.scene(sc)
    p = sc.pt_params()
    (tm, tl), _, _ = configs.config3_params(sc, info)
    ds.set_params(tm, tl)
    target = torch.from_numpy(ds.render_pt(p))
    ctx.set_option("skip_dark", 2)
    _, st_fwd = ds.render_pt(p, stats=True)                 # rays the timed forward render traces
    ctx.set_option("skip_dark", 1)
    fwd = []
    for _ in range(3):
        t0 = time.perf_counter()
        ds.render_pt(p)
        fwd.append(time.perf_counter() - t0)
    stamps = []
    _, hist = diff.optimise(ds, info, sc, p, steps=12, target=target,
                            callback=lambda i, l, m, li: stamps.append(time.perf_counter()))
    # The optimisation itself: 200 Adam steps from the start point, fresh
    # random streams each step with that step's target rendered from the
    # truth on the same streams, learning rate 1e-2 decaying to 1e-3.
    seen = diff.seen_patches(ds, info, sc, p)
    _, (m0, l0), _ = configs.config3_params(sc, info)
    errors = [dict(step=0, **diff.parameter_errors(info, sc, m0, l0, seen))]

    def on_step(i, loss, mat, light):
        if (i + 1) % 50 == 0 or i + 1 == steps:
            errors.append(dict(step=i + 1, loss=loss, **diff.parameter_errors(info, sc, mat, light, seen)))
    t_opt = time.perf_counter()
    diff.optimise(ds, info, sc, p, steps=steps, callback=on_step, fresh_samples=True, lr_decay=0.1 ** (1 / steps))
    t_opt = time.perf_counter() - t_opt
    iters = np.diff(np.asarray(stamps))[1:]                 # steady state: the first step allocates
    it_s = float(np.median(iters))
    f_s = float(np.median(fwd))
    # An iteration renders forward (no shadow ray whose NEE term is exactly
    # zero) and then replays the same paths for the gradients (every shadow
    # ray the reference casts: a zero NEE term can still have a gradient), at
    # the last step's parameters.
    ctx.set_option("skip_dark", 2)
    _, st_it_fwd = ds.render_pt(p, stats=True)
    ctx.set_option("skip_dark", 1)
    _, st_it = ds.render_pt(p, stats=True)
    paths = p.width * p.height * p.spp
    peak = HBM_PEAK_GBS * 1e9
    f_bytes, it_bytes = frame_bytes(st_fwd), frame_bytes(st_it_fwd) + frame_bytes(st_it)
    return {"config": "config3: config1 geometry, 4096 GGX sphere patches (albedo rgb, roughness) + 4096 "
                      "light-only back-wall tiles (emission rgb), 512x512, 32 spp, depth 8, 200 Adam steps, lr 1e-2",
            "optimisation": {"steps": steps, "seconds": t_opt, "seen_patches": int(seen.sum()),
                             "parameter_errors": errors},
            "iters_per_s": 1.0 / it_s, "iter_ms": it_s * 1e3, "forward_ms": f_s * 1e3,
            "iter_over_forward": it_s / f_s, "timed_steps": 12, "loss_first": hist[0], "loss_last": hist[-1],
            "forward_mpaths_per_s": paths / f_s / 1e6, "forward_algorithmic_bytes": f_bytes,
            "forward_hbm_roofline_frac": f_bytes / f_s / peak, "iter_mpaths_per_s": paths / it_s / 1e6,
            "iter_algorithmic_bytes": it_bytes, "iter_hbm_roofline_frac": it_bytes / it_s / peak,
            "nonfinite_pixel_fraction": float((~torch.isfinite(target).all(dim=1)).float().mean())}


def pssmlt(ctx):
    """config4: PSSMLT on the caustic scene, 2^20 bootstrap paths, 2^20 chains x 256 mutations."""
    sc = scene.load_scene(configs.caustic_scene())
    ds = ctx.scene(sc)
    p = sc.mlt_params()
    ds.render_pssmlt(sc.mlt_params(spp=4))
    t0 = time.perf_counter()
    film, st = ds.render_pssmlt(p)
    el = time.perf_counter() - t0
    return {"config": "config4: Cornell box + dielectric sphere (IOR 1.5) + GGX metal sphere (roughness 0.1), "
                      "1024x1024, depth 16, 2^20 bootstrap, 2^20 chains x 256 mutations, large step 0.3, "
                      "sigma 1/64",
            "seconds": el, "mmutations_per_s": st["proposed"] / el / 1e6,
            "acceptance": st["accepted"] / max(1, st["proposed"]), "b": float(st["b"]),
            "nan_pixel_fraction": float((~np.isfinite(film).all(axis=1)).mean())}


def cpu_pssmlt_and_adjoint(budget_s=8.0):
    """Bounded CPU legs of the other two workloads, on the oracle: PSSMLT
    mutations/s on a chain subsample of config 4 (its whole bootstrap, then
    chains [0, n)), and path-replay adjoint paths/s on a pixel window of
    config 3 (the FP64 oracle adjoint of the same definition)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = host_cpu()["threads"]
    orc = oracle.Oracle()
    out = {"cores": threads}
    sc = scene.load_scene(configs.caustic_scene())
    osc = orc.scene(sc)
    p = sc.mlt_params(n_bootstrap=1 << 16)
    t0 = time.perf_counter()
    osc.render_pssmlt(p, threads=threads, chains=(0, 1))         # bootstrap + one chain
    boot_s = time.perf_counter() - t0
    n_ch = 4096
    t0 = time.perf_counter()
    _, st, _, _ = osc.render_pssmlt(p, threads=threads, chains=(0, n_ch))
    el = time.perf_counter() - t0 - boot_s
    muts = n_ch * st["mutations_per_chain"]
    out["pssmlt"] = {"value": muts / el / 1e6, "unit": "Mmutations/s",
                     "sample": f"config4 with 2^16 bootstrap paths, chains 0..{n_ch} of 2^20 x "
                               f"{st['mutations_per_chain']} mutations ({muts} mutations in {el:.1f} s, "
                               f"bootstrap time {boot_s:.2f} s subtracted)"}
    info = configs.config3()
    sc3 = scene.load_scene(info)
    osc3 = orc.scene(sc3)
    p3 = sc3.pt_params()
    d = np.random.default_rng(0).normal(size=(p3.width * p3.height, 3)).astype(np.float32)
    side = 16
    while True:
        x0 = y0 = 256 - side // 2
        t0 = time.perf_counter()
        osc3.adjoint_pt(p3, d, threads=threads, window=(x0, y0, x0 + side, y0 + side))
        el = time.perf_counter() - t0
        if el > budget_s / 4 or side >= 128:
            break
        side *= 2
    paths = side * side * p3.spp
    out["adjoint"] = {"value": paths / el / 1e6, "unit": "Mpaths/s (forward replay + FP64 adjoint)",
                      "sample": f"config3, a {side}x{side} pixel window at the film centre, {p3.spp} spp "
                                f"({paths} paths, {el:.1f} s)"}
    return out


def copy_calibration(nbytes=1 << 30):
    """One device-to-device hipMemcpy (the runtime's copy kernel) of nbytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:      # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(f"# copy calibration: {nbytes} bytes device to device", file=sys.stderr, flush=True)


def parity(ctx):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sc = scene.load_scene(configs.config1(width=128, height=128, spp=2))
    p = sc.pt_params()
    g = ctx.scene(sc).render_pt(p)
    o, _ = oracle.Oracle().scene(sc).render(p, threads=host_cpu()["threads"])
    rel = float(np.sqrt(np.sum((g - o) ** 2) / np.sum(o ** 2)))
    out = {"config": "same scene at 128x128x2spp", "rel_rmse": rel, "bit_equal_fraction": float(np.mean(g == o)),
           "tolerance": 1e-3}
    # config 0, the reference's CPU-runnable case, whole frame on both sides
    sc0 = scene.load_scene(configs.config0())
    p0 = sc0.pt_params()
    ds0 = ctx.scene(sc0)
    ds0.render_pt(p0)
    t0 = time.perf_counter()
    g0 = ds0.render_pt(p0)
    gpu_s = time.perf_counter() - t0
    threads = host_cpu()["threads"]
    t0 = time.perf_counter()
    o0, st0 = oracle.Oracle().scene(sc0).render(p0, threads=threads)
    cpu_s = time.perf_counter() - t0
    out["config0"] = {"config": "config0: Cornell box, 256x256, 16 spp, depth 8, whole frame",
                      "gpu_ms": gpu_s * 1e3, "oracle_s": cpu_s, "oracle_threads": threads,
                      "speedup": cpu_s / gpu_s,
                      "gpu_mrays_per_s": (st0["rays_closest"] + st0["rays_shadow"]) / gpu_s / 1e6,
                      "rel_rmse": float(np.sqrt(np.sum((g0 - o0) ** 2) / np.sum(o0 ** 2))),
                      "bit_equal_fraction": float(np.mean(g0.view(np.int32) == o0.view(np.int32)))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-pssmlt", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-overlap", action="store_true",
                    help="trace shadow rays after shading on the one stream in every frame")
    ap.add_argument("--batch-timeline", action="store_true",
                    help="run only the config-2 plain/batched leg, with pauses between its phases")
    ap.add_argument("--timeline-summary", metavar="KERNEL_TRACE_CSV",
                    help="summarise a rocprofv3 kernel trace of --batch-timeline (no GPU use)")
    ap.add_argument("--counter-summary", nargs="+", metavar=("OUT_JSON", "RESULTS_DB"),
                    help="merge rocprofv3 counter databases of the config-1 frame into OUT_JSON (no GPU use)")
    ap.add_argument("--command", action="append", default=[],
                    help="with --counter-summary: a command that made the databases (recorded)")
    args = ap.parse_args()
    if args.timeline_summary:
        print(json.dumps(timeline_summary(args.timeline_summary), indent=1))
        return
    if args.counter_summary:
        r = counter_summary(args.counter_summary[0], args.counter_summary[1:], args.command)
        print(json.dumps({k: {c: v for c, v in d.items() if not isinstance(v, dict)} for k, d in r["kernels"].items()},
                         indent=1))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the renderer has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.batch_timeline:
        ctx = native.Context(0)
        r = batched(ctx, pause_s=0.5, small=False)
        print(json.dumps({k: v for k, v in r.items() if k not in ("small_items", "_scenes")}), flush=True)
        return

    sc = scene.load_scene(configs.config1(spp=args.spp))
    params = sc.pt_params()
    ctx = native.Context(0)
    ds = ctx.scene(sc)
    if args.no_overlap:
        ctx.set_option("overlap_shadow", 0)
    print(f"# scene: {sc.n_triangles} triangles, {params.width}x{params.height}x{params.spp}spp depth "
          f"{params.max_depth}", file=sys.stderr, flush=True)

    # Counting passes (outside the timed region): exact per-frame work counts.
    # st: the rays the reference casts for this frame.  st_tr: the rays the
    # GPU traces (plain frames cast no shadow ray whose NEE term is exactly
    # zero: it could add nothing to the film).  `value` counts the former per
    # second of frame time; the rooflines use the work actually traced.
    _, st = ds.render_pt(params, stats=True)
    ctx.set_option("skip_dark", 2)
    _, st_tr = ds.render_pt(params, stats=True)
    ctx.set_option("skip_dark", 1)
    rays = st["rays_closest"] + st["rays_shadow"]
    rays_tr = st_tr["rays_closest"] + st_tr["rays_shadow"]
    print(f"# counts: {json.dumps(st)}", file=sys.stderr, flush=True)
    print(f"# traced: {json.dumps(st_tr)}", file=sys.stderr, flush=True)

    # Timed frames: the scene resident in HBM, the film left there (the frame
    # is complete on the device; nothing is copied to the host).  The same
    # frames with the film copied into a host array are timed afterwards.
    for _ in range(args.warmup):
        ds.render_pt(params, host=False)
    t0 = time.perf_counter()
    for i in range(args.steps):
        ds.render_pt(params, host=False)
        print(f"# step {i} done", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    film = np.zeros((params.width * params.height, 3), np.float32)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ds.render_pt(params, out=film)
    ms_with_film_copy = (time.perf_counter() - t1) / args.steps * 1e3
    # Per-kernel launch times for the rooflines: one more frame with HIP events
    # around every launch and the kernels one after another (the timed frames
    # trace each iteration's shadow rays beside the next closest-hit traversal,
    # so a kernel's own duration is not defined there).
    ctx.set_option("overlap_shadow", 0)
    ctx.set_option("timing", 1)
    ctx.reset_kernel_times()
    ds.render_pt(params)
    kt = ctx.kernel_times()
    ctx.set_option("timing", 0)
    ctx.set_option("overlap_shadow", 0 if args.no_overlap else 1)

    ms_per_step = elapsed / args.steps * 1e3

    # Committed rocprofv3 summaries of the same workload: HBM bytes per
    # launch (FETCH_SIZE / WRITE_SIZE) and L2 hit rates (TCC_HIT / TCC_MISS).
    def load_profile(name):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            return {}
        with open(path) as fh:
            return json.load(fh)
    prof = load_profile("counters_config1.json").get("kernels", {})
    names = {"trace_closest": "k_trace_closest<false, false, 16>", "trace_shadow": "k_trace_shadow<false, false, 16>",
             "shade_c": "k_shade_c<false, true>"}
    traffic_of = {k: prof.get(v, {}).get("hbm_bytes_per_launch") for k, v in names.items()}
    l2 = {k: {"l2_hit_rate": prof.get(v, {}).get("l2_hit_rate")} for k, v in names.items()}

    def roofline(kernel, timer_name, frame_algorithmic):
        launches, total_ms = kt[timer_name]
        per_launch_ms = total_ms / launches
        bytes_per_launch = frame_algorithmic / launches        # one frame
        achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
        traffic = traffic_of.get(kernel)
        r = {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
             "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_ms": per_launch_ms,
             "launches": launches, "l2_hit_rate": l2.get(kernel, {}).get("l2_hit_rate")}
        if traffic:
            # what the hardware moved: measured HBM bytes over the same launch time
            r["traffic_gbs"] = traffic / (per_launch_ms * 1e-3) / 1e9
            r["traffic_frac"] = r["traffic_gbs"] / HBM_PEAK_GBS
        return r
    rl_closest = roofline("trace_closest", "trace_closest", closest_bytes(st_tr))
    # The closest-hit traversal keeps 87% of its VALU issue slots busy
    # (profiles/trace_config1_sq_counters.md): the byte roofline is the
    # contract's yardstick, instruction issue is what limits it.
    rl_closest["limiter"] = "VALU instruction issue (87% of issue slots, profiles/trace_config1_sq_counters.md)"

    out = {
        "metric": "Mrays/s (batched vs plain, %HBM peak) + adjoint iters/s at RMSE<=1e-3 vs CPU",
        "value": rays * args.steps / elapsed / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "config1: Cornell box + 69,620-tri displaced icosphere (GGX), 1024x1024, "
                               f"{params.spp} spp, max depth {params.max_depth}, rr depth {params.rr_depth}",
                   "parallelism": "single GPU"},
        "value_note": "value and frame_hbm_roofline_frac count the frame's work as the reference does it: its "
                      "rays (rays_per_frame) and their algorithmic bytes per second of frame time.  The GPU "
                      "traces rays_traced_per_frame (no shadow ray whose NEE term is exactly zero; film "
                      "bit-identical): mrays_traced_per_s and frame_hbm_roofline_frac_traced count those, and "
                      "the per-kernel rooflines count what each kernel traces",
        "mrays_traced_per_s": rays_tr * args.steps / elapsed / 1e6,
        "mpaths_per_s": params.width * params.height * params.spp * args.steps / elapsed / 1e6,
        "rays_per_frame": {"extension": st["rays_closest"], "shadow": st["rays_shadow"]},
        "rays_traced_per_frame": {"extension": st_tr["rays_closest"], "shadow": st_tr["rays_shadow"]},
        "frame_algorithmic_bytes": frame_bytes(st),
        "frame_hbm_roofline_frac": frame_bytes(st) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "frame_algorithmic_bytes_traced": frame_bytes(st_tr),
        "frame_hbm_roofline_frac_traced": frame_bytes(st_tr) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "roofline": rl_closest,
        "rooflines": {"trace_shadow": roofline("trace_shadow", "trace_shadow", shadow_bytes(st_tr)),
                      "shade_c": roofline("shade_c", "shade", shade_bytes(st_tr))},
        "ms_per_step_with_film_to_host": ms_with_film_copy,
        "kernel_ms_per_step": {k: v[1] for k, v in kt.items()},
        "kernel_ms_note": "one frame with the kernels one after another and HIP events around each launch",
        "overlap_shadow": not args.no_overlap,
    }
    if not args.no_batch:
        out["batched"] = batched(ctx)
        out["bvh_build"] = bvh_builds(ctx, out["batched"].pop("_scenes"))
    if not args.no_adjoint:
        out["adjoint"] = adjoint(ctx)
    if not args.no_pssmlt:
        out["pssmlt"] = pssmlt(ctx)
    if not args.no_parity:
        out["parity"] = parity(ctx)
    if args.copy_calibration:
        copy_calibration()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, params, film)
        out["cpu_baseline"]["other_workloads"] = cpu_pssmlt_and_adjoint()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
