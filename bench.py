#!/usr/bin/env python3
"""Benchmark of the volumetric light-transport path on one MI355X.

One step = one full-frame forward render of the 256^3 sphere-union cloud
(1024x576, 64 paths per pixel, up to 32 scattering events, sigma_t 16), inputs
resident in HBM, outputs left in HBM. Prints ONE JSON line with the throughput
in Msamples/s (one sample = one camera path), the HBM roofline of the
path-tracing kernel (algorithmic bytes from a counting pass with the same RNG
keys, duration from HIP events recorded on the launch stream) and the CPU
oracle's rate on a bounded sample of the same workload, timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz peak engine
# clock = 39.3 T lane-operations per second (one plain FP32 operation per
# lane per cycle; wave64 instructions take 4 cycles on a 16-lane SIMD).
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
PMC_PASSES = [
    # FETCH_SIZE uses 3 TCC counters; 4 SQ and 1 GRBM counter share the pass
    ["FETCH_SIZE", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVES", "GRBM_GUI_ACTIVE"],
    # WRITE_SIZE uses 2 TCC counters, the L2 hit/miss pair the other 2
    ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"],
]
METRIC = "delta-tracking Msamples/s fwd & fwd+adjoint vs plain CPU, 256³ cloud, 1 GPU"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="target CPU-baseline wall time per build")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device copies (a kernel of known traffic, for scaling the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--pmc-child", action="store_true",
                    help="internal: the program rocprofv3 runs for the counter passes (copies + one forward)")
    ap.add_argument("--sample-base", type=int, default=0)
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--pmc-report", default=None, help="write the counter summary to this file")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(requested --gpus %d, WORLD_SIZE %d)" % (args.gpus, world), file=sys.stderr, flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    from __graft_entry__ import load_package

    if args.copy_calibration or args.pmc_child:
        n = 1 << 28   # floats: 1 GiB per copy
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 3}), flush=True)
        del a, b
        if args.copy_calibration:
            return

    pkg = load_package()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    if args.pmc_child:
        # One forward of the benchmark's workload, as a timed step renders it.
        wl = getattr(pkg.configs, args.workload)()
        k = wl.kernel()
        npix = wl.w * wl.h
        out_l = torch.empty(npix * 4, dtype=torch.float32, device="cuda")
        out_e = torch.empty(npix * 4, dtype=torch.float32, device="cuda")
        k.render_device(0, 0, wl.w, wl.h, wl.spp, args.seed, args.sample_base, out_l, out_e)
        torch.cuda.synchronize()
        k.Destroy()
        return

    if args.workload == "c2":
        bench_batch(pkg, torch, args)
        return
    if args.workload == "c3":
        bench_fit(pkg, torch, args)
        return
    if args.workload == "c4":
        bench_large(pkg, torch, args)
        return

    t_build = time.perf_counter()
    wl = getattr(pkg.configs, args.workload)()
    t_build = time.perf_counter() - t_build
    k = wl.kernel()
    k.set_timing(True)
    npix = wl.w * wl.h
    out_l = torch.empty(npix * 4, dtype=torch.float32, device="cuda")
    out_e = torch.empty(npix * 4, dtype=torch.float32, device="cuda")
    stream = torch.cuda.Stream()

    def step(i):
        k.render_device(0, 0, wl.w, wl.h, wl.spp, args.seed, i * wl.spp, out_l, out_e, stream=stream)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    k.timing()   # discard warm-up launches

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern_ms, launches = k.timing()
    wall = t1 - t0
    value = wl.samples * args.steps / wall / 1e6
    kern_avg_ms = kern_ms / max(launches, 1)

    # Algorithmic bytes: B = 64*N_node + 32*N_lookup per launch + 32 B output per pixel.
    n_node = n_look = 0
    for i in range(args.steps):
        a, b = k.count(0, 0, wl.w, wl.h, wl.spp, args.seed, (args.warmup + i) * wl.spp)
        n_node += a
        n_look += b
    n_node /= args.steps
    n_look /= args.steps
    alg_bytes = 64.0 * n_node + 32.0 * n_look + 32.0 * npix
    pmc = None if args.no_pmc else measure_pmc(args, args.warmup * wl.spp)

    # Forward + adjoint: one step = the forward render (keeping the per-sample
    # path record) plus the path-replay adjoint of the same samples for a
    # dense image gradient (every pixel weighted), accumulated into FP32
    # gradient bricks. Also timed: the adjoint without the record (it replays
    # each path twice and keeps no per-sample state).
    fwd_adj = None
    if not args.no_adjoint:
        g = torch.Generator(device="cpu").manual_seed(5)
        adj_l = (torch.rand(npix * 4, generator=g) * 1000.0).cuda()
        adj_e = torch.rand(npix * 4, generator=g).cuda()
        grad = torch.zeros(wl.scene.leafx[1] * wl.scene.lvoxc3, dtype=torch.float32, device="cuda")
        rec_bytes = k.record_bytes(0, 0, wl.w, wl.h, wl.spp)
        rec = torch.empty(rec_bytes // 4, dtype=torch.float32, device="cuda")

        def step_adj(i):
            k.render_record_device(0, 0, wl.w, wl.h, wl.spp, args.seed, i * wl.spp, out_l, out_e, rec, stream=stream)
            k.adjoint_record_device(0, 0, wl.w, wl.h, wl.spp, args.seed, i * wl.spp, adj_l, adj_e, grad, rec,
                                    stream=stream)

        def step_replay(i):
            k.render_device(0, 0, wl.w, wl.h, wl.spp, args.seed, i * wl.spp, out_l, out_e, stream=stream)
            k.adjoint_device(0, 0, wl.w, wl.h, wl.spp, args.seed, i * wl.spp, adj_l, adj_e, grad, stream=stream)

        def timed(fn, n):
            fn(0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for i in range(n):
                fn(args.warmup + i)
            torch.cuda.synchronize()
            return time.perf_counter() - t

        k.set_timing(False)
        torch.cuda.reset_peak_memory_stats()
        ta = timed(step_adj, args.adj_steps)
        tr = timed(step_replay, max(1, args.adj_steps // 2)) / max(1, args.adj_steps // 2) * args.adj_steps
        k.set_timing(True)
        fwd_ms = wall / args.steps
        fwd_adj = {"value": round(wl.samples * args.adj_steps / ta / 1e6, 3), "unit": "Msamples/s",
                   "steps": args.adj_steps, "ms_per_step": round(ta / args.adj_steps * 1e3, 3),
                   "ratio_to_forward": round((ta / args.adj_steps) / fwd_ms, 3),
                   "estimator": "path replay, differential delta tracking, trilinear splat of every term; "
                                "forward keeps the per-sample path record (40 B/sample)",
                   "no_record_ms_per_step": round(tr / args.adj_steps * 1e3, 3),
                   "no_record_ratio_to_forward": round((tr / args.adj_steps) / fwd_ms, 3),
                   "record_bytes": rec_bytes,
                   "kernel_device_bytes": k.footprint(),
                   "device_bytes": k.footprint() + wl.scene.footprint() + rec_bytes + 4 * grad.numel() +
                   4 * adj_l.numel() + 4 * adj_e.numel(),
                   "device_bytes_note": "kernel buffers + scene (nodes, bricks, the adjoint's event buffer) + path "
                                        "record + gradient + image gradients",
                   "gradient_finite": bool(torch.isfinite(grad).all().item())}
        del rec

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import pyoracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        threads = max(1, min(threads, os.cpu_count() or 1))
        # Bounded sample: the full frame at a reduced number of paths per
        # pixel (sample indices 0..spp_cpu-1 of every pixel), sized from a
        # 1-spp probe to about --cpu-seconds of work, for each build of the
        # restatement: one lane per path, and the reference's structure (2x2
        # pixel packets traced in four lanes under masks, camera-ray leaves
        # kept for all samples, secondary rays on a step traversal), each
        # compiled -O2 and -O3 -march=x86-64-v3. All four give the same sums
        # bit for bit.
        variants = []
        ref = None
        for packets in (False, True):
            for fast in (False, True):
                tp = time.perf_counter()
                O.render(wl, seed=args.seed, samples=1, threads=threads, fast=fast, packets=packets)
                tp = time.perf_counter() - tp
                spp_cpu = int(max(1, min(wl.spp, args.cpu_seconds / max(tp, 1e-3))))
                tc = time.perf_counter()
                r = O.render(wl, seed=args.seed, samples=spp_cpu, threads=threads, fast=fast, packets=packets)
                tc = time.perf_counter() - tc
                variants.append({"structure": "2x2 packets, 4 lanes" if packets else "one lane per path",
                                 "build": "-O3 -march=x86-64-v3" if fast else "-O2", "spp": spp_cpu,
                                 "seconds": round(tc, 2), "value": round(npix * spp_cpu / tc / 1e6, 4)})
                if ref is None or spp_cpu < ref[1]:
                    ref = (r, spp_cpu)
        best = max(variants, key=lambda v: v["value"])
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            pass
        cpu = {"value": best["value"], "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": "CPU oracle (oracle/oracle.c, %s on %d threads, %s), full %dx%d frame at %d of "
                         "%d paths per pixel (%.1f s), same scene/camera/seed; fastest of the builds listed"
                         % (best["structure"], threads, best["build"], wl.w, wl.h, best["spp"], wl.spp,
                            best["seconds"]),
               "builds": variants, "cpu_model": model, "nproc": os.cpu_count()}
        ref, spp_cpu = ref
        # Same sample on the GPU (same RNG keys) for a path-level parity check.
        k.set_seed(args.seed, 0)
        k.Render(0, 0, wl.w, wl.h, spp_cpu)
        g = k.phb
        a_eq = float(np.mean(g[0][..., 3] == ref[0][..., 3]))
        close = []
        for b in range(2):
            rel = np.abs(g[b][..., :3] - ref[b][..., :3]) / np.maximum(np.abs(ref[b][..., :3]), 1e-6)
            close.append(float(np.mean(rel.max(axis=2) < 1e-4)))
        parity = {"alpha_equal_frac": round(a_eq, 6), "rgb_rel1e-4_frac": [round(c, 6) for c in close],
                  "gpu_cpu_speedup": round(value / cpu["value"], 1)}

    name = (pkg._lib.C.create_string_buffer(256))
    cu = pkg._lib.C.c_uint32()
    pkg._lib.check(pkg._lib.lib().drk_device_info(name, 256, pkg._lib.C.byref(cu)))
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s: 256^3 fBm-displaced union of 64 spheres, sigma_t 16, albedo 0.99, Mie phase, "
                        "Hosek sky T3/a0.1 + sun at 30 deg, %dx%d, %d spp, %d scattering events, forward"
                        % (wl.name, wl.w, wl.h, wl.spp, wl.scattevs),
            "image": [wl.w, wl.h], "spp": wl.spp, "scattevs": wl.scattevs,
            "octree_nodes": wl.scene.node_count, "bricks": wl.scene.leafx[1],
            "brick_occupancy": round(wl.scene.occupancy(), 4), "scene_build_s": round(t_build, 3),
            "device": name.value.decode(), "cus": cu.value,
        },
        "roofline": roofline(pmc, kern_avg_ms, alg_bytes, n_node, n_look),
        "fwd_adjoint": fwd_adj,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    k.Destroy()
    print(json.dumps(line), flush=True)


def _pmc_run(args, counters, sample_base, outdir):
    """One rocprofv3 counter pass over `bench.py --pmc-child` (three 1 GiB
    copies, then one forward of the workload); returns {kernel: {counter:
    mean value per dispatch}} from its CSV, or raises."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    exe = shutil.which("rocprofv3")
    if not exe:
        raise RuntimeError("rocprofv3 not found")
    cmd = [exe, "--pmc"] + counters + ["-d", outdir, "-o", "run", "--output-format", "csv", "--",
                                       sys.executable, os.path.abspath(__file__), "--pmc-child",
                                       "--workload", args.workload, "--seed", str(args.seed),
                                       "--sample-base", str(sample_base)]
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         start_new_session=True)
    try:
        out, _ = p.communicate(timeout=150)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError("rocprofv3 pass timed out: %s" % " ".join(counters))
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("no counter file (rc %d): %s" % (p.returncode, out.decode(errors="replace")[-400:]))
    per = {}
    for fn in files:
        with open(fn) as f:
            for row in csv.DictReader(f):
                key = (row.get("Kernel_Name", ""), row.get("Dispatch_Id", ""))
                per.setdefault(key, {})
                c = row["Counter_Name"]
                per[key][c] = per[key].get(c, 0.0) + float(row["Counter_Value"])
    return [(name, vals) for (name, _), vals in per.items()]


def _pick(dispatches, pattern, largest=None):
    """Mean counters per dispatch of the kernels whose name matches; with
    `largest`, only the dispatches within 10% of the largest value of that
    counter (the 1 GiB copies, not the small copies of the scene upload)."""
    import re
    hits = [v for n, v in dispatches if re.search(pattern, n)]
    if largest:
        top = max(v.get(largest, 0.0) for v in hits) if hits else 0.0
        hits = [v for v in hits if v.get(largest, 0.0) >= 0.9 * top]
    if not hits:
        return None
    return {c: sum(v.get(c, 0.0) for v in hits) / len(hits) for c in hits[0]}


def measure_pmc(args, sample_base):
    """HBM traffic, L2 hit rate and VALU counters of the path-tracing kernel
    for the benchmark's workload, measured now by two rocprofv3 --pmc passes
    of a child process (FETCH_SIZE and WRITE_SIZE scaled by what they read
    for a copy of known size in the same process). Returns a dict, with
    'error' set if the passes could not run."""
    import tempfile
    res = {"passes": [" ".join(c) for c in PMC_PASSES]}
    try:
        vals = {}
        copy = {}
        for i, counters in enumerate(PMC_PASSES):
            with tempfile.TemporaryDirectory(prefix="drk_pmc_", dir="/tmp") as d:
                agg = _pmc_run(args, counters, sample_base, d)
            k = _pick(agg, r"k_render<0>|k_renderILi0E")
            c = _pick(agg, r"(?i)copy", largest=counters[0])
            if k is None or c is None:
                raise RuntimeError("kernels not found in pass %d: %s" % (i, sorted({n for n, _ in agg})[:6]))
            vals.update(k)
            copy.update(c)
        gib_kib = float(1 << 20)
        ff = gib_kib / copy["FETCH_SIZE"]
        wf = gib_kib / copy["WRITE_SIZE"]
        fetch = vals["FETCH_SIZE"] * 1024.0 * ff
        write = vals["WRITE_SIZE"] * 1024.0 * wf
        res.update({
            "fetch_bytes": fetch, "write_bytes": write, "fetch_factor": ff, "write_factor": wf,
            "l2_hit_rate": vals["TCC_HIT_sum"] / max(vals["TCC_HIT_sum"] + vals["TCC_MISS_sum"], 1.0),
            "valu_lane_ops": vals["SQ_THREAD_CYCLES_VALU"], "valu_insts": vals["SQ_INSTS_VALU"],
            "valu_active_insts": vals["SQ_ACTIVE_INST_VALU"], "waves": vals["SQ_WAVES"],
            "gui_active_cycles": vals["GRBM_GUI_ACTIVE"], "raw": vals, "copy_raw": copy,
        })
    except Exception as e:   # the benchmark line still prints; the roofline says why it lacks counters
        res["error"] = str(e)[:300]
    if args.pmc_report:
        with open(args.pmc_report, "w") as f:
            json.dump(res, f, indent=1)
    return res


def roofline(pmc, kern_ms, alg_bytes, n_node, n_look):
    """The path-tracing kernel's roofline: VALU lane-operations (the bound:
    divergent per-lane state machines saturate VALU issue with part of the
    lanes active; the working set is served from L2/MALL), with the HBM
    traffic and the algorithmic bytes beside it."""
    sec = kern_ms * 1e-3
    alg = {"bytes_per_launch": int(alg_bytes), "node_reads_per_launch": int(n_node),
           "brick_lookups_per_launch": int(n_look), "gbs": round(alg_bytes / sec / 1e9, 1),
           "note": "cache-served, not an HBM fraction: 64 B per octree node record the walk visits, 32 B per "
                   "trilinear lookup, 32 B per pixel; the octree and bricks stay in L2/MALL"}
    line = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_LANE_OPS / 1e12, 2), "unit": "T lane-op/s",
            "frac": None, "traffic": None, "kernel": "k_render<0>", "kernel_ms": round(kern_ms, 4),
            "peak_note": "256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (plain FP32 lane-ops)",
            "algorithmic_bytes": alg}
    if pmc and "error" not in pmc:
        ops = pmc["valu_lane_ops"]
        traffic = pmc["fetch_bytes"] + pmc["write_bytes"]
        line.update({
            "achieved": round(ops / sec / 1e12, 3), "frac": round(ops / sec / VALU_PEAK_LANE_OPS, 4),
            "traffic": int(traffic),
            "lane_ops_per_launch": int(ops),
            "lane_utilisation": round(ops / (pmc["valu_active_insts"] * 64.0), 4),
            "valu_issue_per_simd_cycle": round(pmc["valu_insts"] / (1024.0 * pmc["gui_active_cycles"] / 8.0), 4),
            "hbm": {"bytes_per_launch": int(traffic), "fetch": int(pmc["fetch_bytes"]),
                    "write": int(pmc["write_bytes"]), "gbs": round(traffic / sec / 1e9, 1),
                    "frac_of_8TBs": round(traffic / sec / 1e9 / HBM_PEAK_GBS, 4),
                    "scale": "FETCH_SIZE x %.3f, WRITE_SIZE x %.3f (copy of 1 GiB)" % (pmc["fetch_factor"],
                                                                                      pmc["write_factor"])},
            "l2_hit_rate": round(pmc["l2_hit_rate"], 4),
            "counters": "rocprofv3 --pmc, measured in this run: " + "; ".join(pmc["passes"]),
        })
    elif pmc:
        line["counters_error"] = pmc["error"]
    return line


def bench_batch(pkg, torch, args):
    """Batched forward workload: 32 independent clouds of uneven size, each
    with its own image and sample count, in one persistent launch replayed as
    a hipGraph; the autotuner picks the grid and work-unit granularity first.
    Also timed: the same jobs as separate launches, back to back."""
    t_build = time.perf_counter()
    wls = pkg.configs.c2()
    t_build = time.perf_counter() - t_build
    kernels = [wl.kernel() for wl in wls]
    jobs = []
    for i, (wl, k) in enumerate(zip(wls, kernels)):
        n = wl.w * wl.h * 4
        jobs.append((k, 0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, torch.empty(n, device="cuda"),
                     torch.empty(n, device="cuda")))
    samples = sum(wl.samples for wl in wls)
    b = pkg.batch.Batch(jobs)
    report = b.autotune(reps=1)
    chosen = report.strip().splitlines()[-1].split()
    for _ in range(args.warmup):
        b.launch_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch_graph()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # The same jobs as separate launches (each with the default chunking).
    for j in jobs:
        j[0].render_device(*j[1:8], j[8], j[9])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        for j in jobs:
            j[0].render_device(*j[1:8], j[8], j[9])
    torch.cuda.synchronize()
    sep = (time.perf_counter() - t1) / max(1, args.steps // 2)
    value = samples * args.steps / wall / 1e6
    line = {
        "metric": METRIC + " (batched variant)",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: %d clouds (fBm / sphere union), grids %d^3..%d^3, voxel occupancy 2-30%%, "
                               "images 128^2..1024^2, 16-64 spp, 16 scattering events, one batched launch "
                               "replayed as a hipGraph" % (len(wls), min(w.grid.shape[0] for w in wls),
                                                            max(w.grid.shape[0] for w in wls)),
                   "scenes": len(wls), "samples_per_step": samples, "scene_build_s": round(t_build, 1),
                   "autotune_choice": {"grid_blocks": int(chosen[1]), "unit_scale": float(chosen[2])}},
        "separate_launches_ms_per_step": round(sep * 1e3, 3),
        "batch_speedup_vs_separate": round(sep / (wall / args.steps), 3),
        "autotune_table": report.strip().splitlines()[1:-1],
    }
    b.destroy()
    for k in kernels:
        k.Destroy()
    print(json.dumps(line), flush=True)


def bench_fit(pkg, torch, args):
    """Differentiable density fitting (128^3 fBm ground truth, 16 ring views,
    256^2, 32 spp, 8 scattering events): Adam iterations, each a forward with
    path record and an adjoint per view; also the forward alone."""
    from droplet_render_amd.fit import DensityFit
    t_build = time.perf_counter()
    fw = pkg.configs.c3()
    fit = DensityFit(fw)
    t_build = time.perf_counter() - t_build
    samples = len(fw.views) * fw.w * fw.h * fw.spp
    rho = torch.tensor(fw.init.pvol[1], device="cuda")
    rho, losses = fit.run(args.warmup, seed=1000, rho=rho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rho, more = fit.run(args.steps, seed=2000, rho=rho)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    losses += more
    with torch.no_grad():
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(args.steps):
            fit.loss(rho, seed=3000 + i)
        torch.cuda.synchronize()
        fwd = (time.perf_counter() - t1) / args.steps
    it = wall / args.steps
    line = {
        "metric": METRIC + " (density-fitting variant)",
        "value": round(samples / it / 1e6, 3), "unit": "Msamples/s (forward+adjoint)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(it * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c3: fit a uniform 0.5 start to a 128^3 fBm cloud (seed 7) from 16 views on a "
                               "20-degree ring, 256^2, 32 spp, 8 scattering events, L2 image loss, Adam lr 0.05",
                   "samples_per_step": samples, "setup_s": round(t_build, 2)},
        "forward_only_ms_per_step": round(fwd * 1e3, 3),
        "ratio_to_forward": round(it / fwd, 3),
        "loss_first_last": [round(losses[0], 4), round(losses[-1], 4)],
    }
    fit.destroy()
    print(json.dumps(line), flush=True)


def bench_large(pkg, torch, args):
    """Large forward+adjoint: 512^3 sparse cloud, 8 views at 1920x1080, 16 spp,
    64 scattering events. Forward: the 8 views in one batched launch.
    Forward+adjoint: the 8 views' record forward in one views launch, then
    their adjoints from the records in one views launch into one FP32
    gradient (and, for comparison, as 8 separate calls each). Reports the
    device footprint and the algorithmic GB/s of the forward."""
    t_build = time.perf_counter()
    wl, views = pkg.configs.c4()
    t_build = time.perf_counter() - t_build
    from droplet_render_amd.fit import _kernel
    ks = [_kernel(wl.scene, vi, pi, wl.w, wl.h, wl.scattevs, wl.sigma_t) for vi, pi in views]
    n = wl.w * wl.h * 4
    outs = [(torch.empty(n, device="cuda"), torch.empty(n, device="cuda")) for _ in ks]
    jobs = [(k, 0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, o[0], o[1]) for i, (k, o) in enumerate(zip(ks, outs))]
    b = pkg.batch.Batch(jobs)
    samples = len(ks) * wl.samples
    b.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / args.steps
    # the same forward as separate launches (the fwd+adjoint leg is per view)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for i, k in enumerate(ks):
        k.render_device(0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, outs[i][0], outs[i][1])
    torch.cuda.synchronize()
    fwd_sep = time.perf_counter() - t2
    # forward with path records + adjoint (dense random image gradient): the
    # 8 views in one views launch each, one event reduction; and as 8
    # separate calls each
    g = torch.Generator(device="cpu").manual_seed(5)
    adj_l = (torch.rand(n, generator=g) * 1000.0).cuda()
    adj_e = torch.rand(n, generator=g).cuda()
    grad = torch.zeros(wl.scene.leafx[1] * wl.scene.lvoxc3, device="cuda")
    rb = ks[0].record_bytes(0, 0, wl.w, wl.h, wl.spp)
    recs = [torch.empty(rb // 4, device="cuda") for _ in ks]
    jf = [(k, 0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, outs[i][0], outs[i][1], recs[i]) for i, k in enumerate(ks)]
    ja = [(k, 0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, adj_l, adj_e, recs[i]) for i, k in enumerate(ks)]

    def views_step():
        pkg.batch.render_record_views(jf)
        pkg.batch.adjoint_record_views(ja, grad)

    def separate_step():
        for i, k in enumerate(ks):
            k.render_record_device(0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, outs[i][0], outs[i][1], recs[i])
        for i, k in enumerate(ks):
            k.adjoint_record_device(0, 0, wl.w, wl.h, wl.spp, args.seed + i, 0, adj_l, adj_e, grad, recs[i])

    views_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        views_step()
    torch.cuda.synchronize()
    fa = (time.perf_counter() - t1) / args.steps
    separate_step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    separate_step()
    torch.cuda.synchronize()
    fa_sep = time.perf_counter() - t3
    # algorithmic bytes of one view's forward (counting pass)
    nn, nl = ks[0].count(0, 0, wl.w, wl.h, wl.spp, args.seed, 0)
    alg = 64.0 * nn + 32.0 * nl + 32.0 * wl.w * wl.h
    footprint = wl.scene.footprint() + sum(k.footprint() for k in ks) + rb * len(recs) + 4 * grad.numel() + \
        2 * 4 * n * len(outs)
    line = {
        "metric": METRIC + " (large fwd+adjoint variant)",
        "value": round(samples / fwd / 1e6, 3), "unit": "Msamples/s (forward)", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(fwd * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c4: 512^3 sparse cloud (3 fBm-displaced spheres, %.1f%% of the leaf lattice "
                               "occupied), sigma_t 32, 8 views 1920x1080 on a ring, 16 spp, 64 scattering events"
                               % (100 * wl.scene.occupancy()),
                   "samples_per_step": samples, "bricks": wl.scene.leafx[1], "scene_build_s": round(t_build, 1)},
        "fwd_adjoint": {"value": round(samples / fa / 1e6, 3), "unit": "Msamples/s", "ms_per_step": round(fa * 1e3, 3),
                        "ratio_to_forward": round(fa / fwd, 3),
                        "launches": "record forward and adjoint of the 8 views, one views launch each",
                        "separate_calls_ms": round(fa_sep * 1e3, 3),
                        "forward_separate_launches_ms": round(fwd_sep * 1e3, 3),
                        "gradient_finite": bool(torch.isfinite(grad).all().item())},
        "forward_algorithmic_GBps": round(alg * len(ks) / fwd / 1e9, 1),
        "device_footprint_bytes": int(footprint),
    }
    b.destroy()
    for k in ks:
        k.Destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
