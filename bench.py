"""Benchmark of the MI355X APSP + Brandes path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload spatial200k|spatial50k|fixed5k|cohort64|adjoint20k]

One step is one batched call of the hot path (ng_brandes_accumulate with per-source SSSP summaries, so
every source yields both its APSP summary and its betweenness contribution) over one batch of sources of
a synthetic graph resident in HBM. The timed region is exactly K steps, bracketed by device
synchronisation; value = sources processed / wall time.

Workloads (default spatial200k, the headline configuration):
  fixed5k      5000 nodes, exactly 6 distinct out-arcs per node, integer weights U{1..10}, seed 0, all sources
  spatial50k   50k nodes uniform in a 1000^3 um box, p ~ exp(-d/150 um), mean out-degree 20, FP64 lognormal(0,0.5),
               all 50k sources
  spatial200k  200k nodes, mean out-degree 100 (~20M arcs), FP32 lognormal(0,0.5), 8192 sampled sources, batches of 1024
  cohort64     64 spatial graphs, nodes U[1k, 20k], mean out-degree U[10, 60], FP32 lognormal(0,0.5), all sources of all
               graphs; one step = one replay of the hipGraph-captured persistent cohort launch, reported beside the
               same work as 64 plain launches (one ng_brandes_accumulate per graph)
  adjoint20k   20k nodes, mean out-degree 30, FP64 weights U[0.5, 1.5]; one step = the gradient of CPL and GE with
               respect to every arc weight over all 20k sources (ng_apsp_adjoint), reported beside the plain
               forward (ng_sssp_summaries) and Brandes (ng_brandes_accumulate) passes over the same sources

Roofline: algorithmic bytes per source B_bc = 2*(A*(s_i+s_w) + (V+1)*s_o) + V*(2*s_d + 32) (one pass over the
CSR for the forward sweep, one for the backward sweep, per-node distance/sigma/delta/betweenness state) divided
by the batched kernel's launch duration (HIP events on the library's stream), against 8 TB/s.
Traffic: the hot kernel's profiled bytes per launch (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each scaled by
what the counter reads for a plain copy of known size), read from the committed profile of the workload; these
are L2-to-fabric bytes (they include hits in the MALL, the last-level cache in front of HBM).
CPU baseline: the oracle's restatement of the reference (PriorityQueueVariant APSP + BrandesVariant), one worker
thread per host core available to this process (its CPU affinity, capped by its cgroup CPU quota and by
OMP_NUM_THREADS where those are set), over contiguous source blocks, timed on a bounded sample of the same
sources; the median of 3 timed runs.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sources/sec + GTEPS for APSP & Brandes BC, GPU vs host CPU; HBM GB/s fraction"
HBM_PEAK_GBS = 8000.0


def make_workload(name: str):
    import neurograph_amd as ng

    if name == "fixed5k":
        g = ng.generate_fixed_degree(5000, 6, 1, 10, 0)
        return g, np.arange(g.n, dtype=np.uint32), 5000, {"nodes": 5000, "out_degree": 6, "weights": "int U{1..10}"}
    if name == "spatial50k":
        g = ng.generate_spatial(50000, 20, box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5, weights=np.float64, seed=0)
        return g, np.arange(g.n, dtype=np.uint32), 10000, {"nodes": 50000, "mean_out_degree": 20, "weights": "f64 lognormal(0,0.5)"}
    if name == "spatial200k":
        g = ng.generate_spatial(200000, 100, box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5, weights=np.float32, seed=0)
        rng = np.random.default_rng(0)
        src = np.sort(rng.choice(g.n, size=8192, replace=False)).astype(np.uint32)
        return g, src, 1024, {"nodes": 200000, "mean_out_degree": 100, "weights": "f32 lognormal(0,0.5)", "sampled_sources": 8192}
    raise SystemExit(f"unknown workload {name}")


def bytes_per_source(g) -> float:
    """B_bc with field sizes s_i = 4 (target), s_w = weight bytes, s_o = 4 (row offset), s_d = distance bytes."""
    s_w = g.weights.dtype.itemsize
    s_d = 8 if g.weights.dtype != np.float32 else 4
    return 2.0 * (g.m * (4 + s_w) + (g.n + 1) * 4) + g.n * (2 * s_d + 32)


def read_traffic(workload: str):
    """(bytes per launch, source) of the hot kernel from the committed rocprofv3 counter summary of this workload
    (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a copy of known size), or (None, None).
    The counters count L2-to-fabric bytes, MALL hits included."""
    rel = os.path.join("profiles", f"traffic_{workload}.json")
    prof = os.path.join(ROOT, rel)
    if not os.path.exists(prof):
        return None, None
    with open(prof) as f:
        return json.load(f).get("fabric_bytes_per_launch"), rel


def roofline(achieved_gbs: float, workload: str) -> dict:
    traffic, src = read_traffic(workload)
    return {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_kind": "profiled L2-to-fabric bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated; MALL hits "
                            "included)" if traffic else None,
            "traffic_source": src}


def host_cpu() -> dict:
    """Host cores this process may use: CPU affinity, cgroup CPU quota (cgroup v2 cpu.max or v1 cfs quota), and
    OMP_NUM_THREADS where set; threads = the smallest of them. Plus the lscpu model name."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["nproc"]
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        try:
            with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
                q = int(f.read())
            with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                per = int(f.read())
            if q > 0:
                quota = q / per
        except (OSError, ValueError):
            pass
    info["cgroup_cpus"] = None if quota is None else round(quota, 2)
    omp = os.environ.get("OMP_NUM_THREADS")
    info["omp_num_threads"] = int(omp) if omp and omp.isdigit() else None
    caps = [info["affinity"]]
    if quota is not None:
        caps.append(max(1, math.floor(quota)))
    if info["omp_num_threads"]:
        caps.append(info["omp_num_threads"])
    info["threads"] = max(1, min(caps))
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    info["model"] = model
    return info


def copy_calibration(ctx, mib: int):
    """One device copy of a known size (outside the timed region): the reference kernel for scaling profiler
    byte counters."""
    if mib <= 0:
        return
    nb = mib << 20
    cms = ctx.copy_calibration(nb)
    print(f"[bench] copy_kernel {nb} bytes read + {nb} written in {cms:.3f} ms ({2 * nb / cms / 1e6:.1f} GB/s)",
          file=sys.stderr, flush=True)


def make_cohort(seed: int = 0):
    """64 member graphs: n ~ U[1000, 20000], mean out-degree ~ U[10, 60] (seeded), spatial generator, FP32."""
    import neurograph_amd as ng

    rng = np.random.default_rng(seed)
    ns = rng.integers(1000, 20001, size=64)
    ds = rng.uniform(10.0, 60.0, size=64)
    return [ng.generate_spatial(int(n), float(d), box=1000.0, lam=150.0, wdist="lognormal", a=0.0, b=0.5,
                                weights=np.float32, seed=1000 + i) for i, (n, d) in enumerate(zip(ns, ds))]


def run_cohort(args):
    import neurograph_amd as ng

    t0 = time.perf_counter()
    hgs = make_cohort()
    gen_s = time.perf_counter() - t0
    ctx = ng.Context(0)
    gs = [ng.Graph(ctx, h) for h in hgs]
    n_total = sum(h.n for h in hgs)
    m_total = sum(h.m for h in hgs)
    print(f"[bench] cohort64: {n_total} sources, {m_total} arcs in 64 graphs, generated in {gen_s:.1f}s",
          file=sys.stderr, flush=True)
    co = ng.Cohort(ctx, gs)
    t0 = time.perf_counter()
    co.autotune(256)
    print(f"[bench] cohort autotune {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    steps = args.steps or 3
    for _ in range(args.warmup):
        co.brandes()
    kernel_ms, arcs = [], 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, _, ms = co.brandes()
        kernel_ms.append(ms)
        arcs += co.last_stats()["arcs_reached"]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    copy_calibration(ctx, args.copy_calibration_mib)
    # the same work as 64 plain launches (each graph's own tuned width, all its sources, one call per graph)
    for g in gs:
        g.brandes(delta=0.0)
    ctx.synchronize()
    t1 = time.perf_counter()
    plain_kernel = 0.0
    for _ in range(steps):
        for g in gs:
            g.brandes(delta=0.0)
            plain_kernel += g.last_stats()["kernel_ms"]
    ctx.synchronize()
    plain_wall = time.perf_counter() - t1
    value = n_total * steps / wall
    plain_value = n_total * steps / plain_wall
    algo = sum(bytes_per_source(h) * h.n for h in hgs)
    achieved = algo * steps / (sum(kernel_ms) / 1e3) / 1e9
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "sources/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cohort64", "graphs": 64, "nodes": "U[1000,20000]", "mean_out_degree": "U[10,60]",
                   "weights": "f32 lognormal(0,0.5)", "sources": n_total, "arcs": m_total,
                   "lanes_per_source": co.last_stats()["block"], "slots": co.last_stats()["slots"],
                   "outputs": "per-source APSP summaries + betweenness of every graph in one replay"},
        "gteps": round(arcs / wall / 1e9, 3),
        "kernel_ms_per_step": round(sum(kernel_ms) / steps, 3),
        "plain_64_launches": {"value": round(plain_value, 3), "unit": "sources/s",
                              "ms_per_step": round(plain_wall * 1e3 / steps, 3),
                              "kernel_ms_per_step": round(plain_kernel / steps, 3)},
        "cohort_vs_plain": round(value / plain_value, 3),
        "roofline": roofline(achieved, "cohort64"),
    }
    if not args.no_cpu:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline_cohort(hgs, args.cpu_seconds)
        res["speedup_vs_cpu"] = round(value / res["cpu_baseline"]["value"], 2)
    print(json.dumps(res), flush=True)


def run_adjoint(args):
    import neurograph_amd as ng

    hg = ng.generate_spatial(20000, 30, box=1000.0, lam=150.0, wdist="uniform", a=0.5, b=1.5, weights=np.float64, seed=5)
    ctx = ng.Context(0)
    g = ng.Graph(ctx, hg)
    src = np.arange(hg.n, dtype=np.uint32)
    sample = np.arange(0, hg.n, 6, dtype=np.uint32)
    d = args.delta  # > 0: a fixed bucket width (and --block) for every pass instead of the tuned shapes
    if args.block:
        g.set_options(block=args.block)
    shape = {0: (d, args.block), 1: (d, args.block)}  # (bucket width, lanes) per mode: SSSP 0, Brandes/adjoint 1
    if d <= 0.0:
        shape[1] = g.autotune_shape(sample, mode=1)
        shape[0] = g.autotune_shape(sample, mode=0)
    steps = args.steps or 3
    for _ in range(args.warmup):
        g.adjoint(src, delta=d)
    times = {}
    kms = {}
    for name, fn in (("adjoint", lambda: g.adjoint(src, delta=d)), ("forward", lambda: g.sssp(src, delta=d)),
                     ("brandes", lambda: g.brandes(src, delta=d))):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        k = 0.0
        for _ in range(steps):
            fn()
            k += g.last_stats()["kernel_ms"]
        ctx.synchronize()
        times[name] = (time.perf_counter() - t0) / steps
        kms[name] = k / steps
    copy_calibration(ctx, args.copy_calibration_mib)
    # one training step of the torch binding's shape (new weights -> forward summaries -> adjoint), with the device
    # graph rebuilt per step (ng_graph_create) or updated in place (ng_graph_set_weights)
    rng = np.random.default_rng(2)
    w_steps = [hg.weights * np.exp(rng.normal(0.0, 1e-3, hg.m)) for _ in range(steps + 1)]
    def train_rebuild(w):  # the same widths and lanes as the tuned graph
        gg = ng.Graph(ctx, ng.HostGraph(hg.n, hg.offsets, hg.targets, w))
        gg.set_options(block=shape[0][1])
        gg.sssp(src, delta=shape[0][0])
        gg.set_options(block=shape[1][1])
        gg.adjoint(src, delta=shape[1][0])
        gg.close()
    def train_in_place(w):
        g.set_weights(w)
        g.sssp(src, delta=d)
        g.adjoint(src, delta=d)
    train = {}
    for name, fn in (("rebuild", train_rebuild), ("set_weights", train_in_place)):
        fn(w_steps[-1])
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            fn(w_steps[k])
        ctx.synchronize()
        train[name] = (time.perf_counter() - t0) / steps * 1e3
    g.set_weights(hg.weights)
    value = hg.n / times["adjoint"]
    # adjoint traffic model: B_bc plus the gradient read-modify-write of the two per-arc vectors
    algo = bytes_per_source(hg) * hg.n + 2 * 16.0 * hg.m
    achieved = algo / (kms["adjoint"] / 1e3) / 1e9
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "sources/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(times["adjoint"] * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint20k", "nodes": hg.n, "arcs": hg.m, "mean_out_degree": 30, "weights": "f64 U[0.5,1.5]",
                   "lanes_per_source": g.last_stats()["block"],
                   "outputs": "dCPL/dw and dGE/dw for every arc, all sources"},
        "kernel_ms_per_step": round(kms["adjoint"], 3),
        "plain_forward": {"value": round(hg.n / times["forward"], 3), "unit": "sources/s", "kernel_ms": round(kms["forward"], 3)},
        "plain_brandes": {"value": round(hg.n / times["brandes"], 3), "unit": "sources/s", "kernel_ms": round(kms["brandes"], 3)},
        "adjoint_vs_forward": round(times["forward"] / times["adjoint"], 3),
        "train_step_ms": {"graph_rebuilt_per_step": round(train["rebuild"], 3),
                          "graph_updated_in_place": round(train["set_weights"], 3)},
        "roofline": roofline(achieved, "adjoint20k"),
    }
    print(json.dumps(res), flush=True)


def cpu_baseline_cohort(hgs, seconds: float):
    """The oracle on every member graph in turn, each on an equal share of a bounded time budget (sources of a
    graph are independent); the rate is the cohort's total sources over the extrapolated total time."""
    from oracle import oracle as orc

    host = host_cpu()
    threads = host["threads"]
    total_t, sampled = 0.0, 0
    per_graph = seconds / len(hgs)
    for h in hgs:
        c = orc.Csr(h.n, h.offsets, h.targets, h.weights)
        k = threads
        while True:
            src = np.arange(0, h.n, max(1, h.n // k), dtype=np.uint32)[:k]
            t0 = time.perf_counter()
            orc.apsp(c, sources=src, threads=threads)
            orc.brandes(c, sources=src, threads=threads)
            dt = time.perf_counter() - t0
            if dt >= per_graph / 4 or k >= h.n:
                break
            k = min(h.n, k * 4)
        total_t += dt * h.n / len(src)
        sampled += len(src)
    n_total = sum(h.n for h in hgs)
    return {"value": round(n_total / total_t, 3), "unit": "sources/s", "cores": threads, "kind": "port",
            "host": host,
            "sample": f"{sampled} sources spread over all 64 graphs, each graph's time extrapolated to all its sources; "
                      f"avg-apsp + avg-centrality (oracle/oracle.cpp, {threads} threads)"}


def cpu_baseline(g, sources, seconds: float, runs: int = 3):
    """Oracle restatement of the reference's CPU path (APSP via PriorityQueueVariant + BrandesVariant: the
    reference's avg-apsp and avg-centrality runs), one worker thread per available host core, each a contiguous
    block of sources like one MPI rank per core, on a bounded sample of the same sources (sources are independent,
    so the rate scales linearly). `runs` timed runs of about seconds / runs each; the median is reported."""
    from oracle import oracle as orc

    host = host_cpu()
    threads = host["threads"]
    c = orc.Csr(g.n, g.offsets, g.targets, g.weights)
    probe = sources[: 2 * threads]
    t0 = time.perf_counter()
    orc.apsp(c, sources=probe, threads=threads)
    orc.brandes(c, sources=probe, threads=threads)
    per_source_wall = (time.perf_counter() - t0) / len(probe)
    k = int(max(len(probe), min(len(sources), seconds / runs / max(per_source_wall, 1e-9))))
    k = max(threads, (k // threads) * threads)
    sample = sources[:k]
    rates, parts = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        orc.apsp(c, sources=sample, threads=threads)
        t1 = time.perf_counter()
        orc.brandes(c, sources=sample, threads=threads)
        t2 = time.perf_counter()
        rates.append(len(sample) / (t2 - t0))
        parts.append((t1 - t0, t2 - t1))
        print(f"[bench] cpu run: {len(sample)} sources in {t2 - t0:.2f} s", file=sys.stderr, flush=True)
    med = sorted(range(runs), key=lambda i: rates[i])[runs // 2]
    return {"value": round(rates[med], 3), "unit": "sources/s", "cores": threads, "kind": "port",
            "runs": [round(r, 3) for r in rates], "statistic": f"median of {runs}", "host": host,
            "sample": f"{len(sample)} of the workload's sources per run; median run avg-apsp {parts[med][0]:.2f} s + "
                      f"avg-centrality {parts[med][1]:.2f} s (oracle/oracle.cpp, {threads} threads, contiguous source blocks)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: one pass over the workload's sources)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spatial200k",
                    choices=["fixed5k", "spatial50k", "spatial200k", "cohort64", "adjoint20k"])
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=24.0, help="CPU baseline budget, split over 3 timed runs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--delta", type=float, default=0.0, help="bucket width (0: autotune width and lanes per source)")
    ap.add_argument("--block", type=int, default=0, choices=[0, 64, 256, 512, 1024],
                    help="lanes per source workgroup when --delta is given (0: the library default)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (a known-size kernel for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path (sources would be split across GPUs with "
              "one final all-reduce of the n-entry betweenness vector and of the scalar sums)", file=sys.stderr)
        sys.exit(2)

    import neurograph_amd as ng

    if args.workload == "cohort64":
        run_cohort(args)
        return
    if args.workload == "adjoint20k":
        run_adjoint(args)
        return
    t0 = time.perf_counter()
    hg, sources, batch, cfg = make_workload(args.workload)
    gen_s = time.perf_counter() - t0
    batch = args.batch or batch
    batches = [sources[i:i + batch] for i in range(0, len(sources), batch)]
    steps = args.steps or len(batches)
    ctx = ng.Context(0)
    g = ng.Graph(ctx, hg)
    if args.block:
        g.set_options(block=args.block)  # lanes per workgroup for every call (with --delta: no autotuning)
    print(f"[bench] {args.workload}: n={hg.n} m={hg.m} generated in {gen_s:.1f}s; {len(sources)} sources, batch {batch}, "
          f"{steps} steps", file=sys.stderr, flush=True)

    delta = args.delta
    block = None
    if delta <= 0:
        # bucket width and lanes per workgroup, tuned together on a sample large enough to fill the GPU
        rng = np.random.default_rng(1)
        sample = np.sort(rng.choice(sources, size=min(len(sources), 3072), replace=False))
        t0 = time.perf_counter()
        delta, block = g.autotune_shape(sample, mode=1)
        print(f"[bench] autotune {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    print(f"[bench] delta {delta:.6g} block {block}", file=sys.stderr, flush=True)

    sync = ctx.synchronize  # every ng call synchronizes already; this brackets the timed region explicitly

    bc = np.zeros(hg.n)
    for w in range(args.warmup):
        g.brandes(batches[w % len(batches)], delta=delta, bc=bc, summaries=True)
    bc[:] = 0.0
    kernel_ms, arcs, srcs = [], 0, 0
    summaries = []
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        b = batches[k % len(batches)]
        _, summ = g.brandes(b, delta=delta, bc=bc, summaries=True)
        st = g.last_stats()
        kernel_ms.append(st["kernel_ms"])
        arcs += st["arcs_reached"]
        srcs += len(b)
        summaries.append(summ)
    sync()
    wall = time.perf_counter() - t0

    value = srcs / wall
    copy_calibration(ctx, args.copy_calibration_mib)
    launches_bytes = [bytes_per_source(hg) * len(batches[k % len(batches)]) for k in range(steps)]
    achieved = sum(launches_bytes) / (sum(kernel_ms) / 1e3) / 1e9
    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "sources/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # distances are computed in this type; sigma, delta and betweenness always in f64
        "dtype": {np.dtype(np.int32): "u32", np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[hg.weights.dtype],
        "data": "synthetic",
        "config": {"workload": args.workload, **cfg, "arcs": hg.m, "sources_per_step": batch, "delta": delta,
                   "lanes_per_source": g.last_stats()["block"],
                   "outputs": "per-source APSP summaries + betweenness in one pass"},
        "gteps": round(arcs / wall / 1e9, 3),
        "kernel_ms_per_step": round(sum(kernel_ms) / steps, 3),
        "roofline": roofline(achieved, args.workload),
    }
    if not args.no_cpu:
        print("[bench] timing the CPU baseline", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(hg, sources, args.cpu_seconds)
        res["speedup_vs_cpu"] = round(value / res["cpu_baseline"]["value"], 2)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
