"""Self-join benchmark: result pairs/s and join time, FP32+FP64-recheck vs
plain FP64, with the roofline of the dominant kernel and a CPU baseline.

Workloads (--workload, default exp2d):
  exp2d       2,000,000 points, 2-D Exp(40) clipped to [0,1], seed 2, eps for
              ~64 neighbours/pt; workload-sorted work queue, B-Unicomp.
  uniform2d   100,000 points, U[0,100]^2, seed 1, eps 0.5; Unicomp.
  uniform6d   2,000,000 points, U[0,1]^6, seed 3, eps for ~32 neighbours/pt;
              linear-ID Unicomp, result batches forced to >= 4 rounds.
  clusters64d 500,000 points, 64-D, 100 Gaussian clusters, seed 4, 6 indexed
              dims, eps for ~50 neighbours/pt; FP32 distances on the matrix
              cores (roofline: FP32 MFMA).
  many64      64 independent joins (N log-uniform 1e4..5e5, d in
              {2,3,4,6,8}, uniform or Exp(40) data, 8..128 neighbours/pt,
              seed 5) in one graph-recorded persistent launch per pass,
              parameters picked by the autotuner.

One step = grid construction + count pass + offsets + fill pass over points
already resident in HBM, producing the exact neighbour table in HBM (for
many64: one replay of the recorded graph over the planned grids).  Prints
one JSON line.  roofline.achieved / frac: the join's algorithmic bytes
N*d*8 + |S|*4 + (N+1)*8 over the step time (clusters64d: FP32 matrix-core
flops 2*d*C; uniform6d and many64: the exact predicate's (3d-1) flops per
claimed candidate against the vector FP64 peak, bytes under roofline.hbm);
kernel_achieved / kernel_frac: the same work over the dominant kernel's
time, from HIP events on the stream it runs on.
cpu_baseline is the CPU oracle (grid join, OpenMP) on the host cores; the
oracle's table is also the reference of config.exact_vs_oracle (an
order-free fingerprint of every GPU table), and a mismatch exits 1.
end_to_end times the reference's region: host points to host neighbour
table.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "self-join-hpbdc2019_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
# vector FP64 peak (FMA = 2 flops); unpacked FP32 VALU issues at the same rate
VALU_FP64_PEAK_TFLOPS = 78.6


def predicate_flops(candidates, d):
    """Flops of the exact predicate over `candidates` pairs: d differences,
    one square and d - 1 fused multiply-adds per pair."""
    return float(candidates) * (3 * d - 1)
METRIC = "result pairs/sec & join time at 2M pts, FP32+recheck vs FP64, % HBM roofline"

SINGLE = {
    "exp2d": dict(form="bunicomp", schedule="workqueue", rounds=0,
                  desc="2,000,000 points x 2-D, Exp(40) clipped [0,1], seed 2, ~64 neighbours/pt"),
    "uniform2d": dict(form="unicomp", schedule="strided", rounds=0,
                      desc="100,000 points x 2-D, U[0,100]^2, seed 1"),
    "uniform6d": dict(form="lid", schedule="strided", rounds=4,
                      desc="2,000,000 points x 6-D, U[0,1]^6, seed 3, ~32 neighbours/pt, >= 4 result rounds"),
    "clusters64d": dict(form="lid", schedule="strided", rounds=0,
                        desc="500,000 points x 64-D, 100 Gaussian clusters (sigma 0.5), seed 4, "
                             "6 indexed dims, ~50 neighbours/pt"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="exp2d", choices=list(SINGLE) + ["many64"])
    ap.add_argument("--form", default=None)
    ap.add_argument("--schedule", default=None)
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="timed steps of the plain FP64 form (default: --steps; 1 for clusters64d)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-record", action="store_true",
                    help="fill evaluates every distance again instead of replaying the count's hit record")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--probe", action="store_true",
                    help="counter probe: a 512 MiB device copy, then one FP32+recheck and one FP64 join "
                         "of the workload (run under rocprofv3 --pmc)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two --probe counter runs (FETCH_SIZE, WRITE_SIZE) into "
                         "profiles/traffic_<workload>_<kernel>.json")
    return ap.parse_args()


def cores_used():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        cores = min(cores, int(env))
    return cores


def run_gpu(ctx, x, k, eps, kw, steps, warmup, torch):
    stats = []
    for _ in range(warmup):
        r = ctx.self_join(x, k, eps, **kw)
        r.free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    total = None
    for _ in range(steps):
        r = ctx.self_join(x, k, eps, **kw)
        stats.append(r.stats())
        total = r.total
        r.free()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(steps, 1)
    avg = {key: sum(s[key] for s in stats) / len(stats) for key in stats[0]}
    return dt, total, avg


def fingerprint(off, ids, rows, torch):
    """Order-free fingerprint of a neighbour table on the device: |S|, sum
    of (row ID + 1) * neighbour ID and sum of squared IDs."""
    rr = torch.repeat_interleave(rows.to(torch.int64), (off[1:] - off[:-1]).to(torch.int64))
    v = ids.to(torch.int64)
    return (int(off[-1].item() - off[0].item()), int((v * (rr + 1)).sum().item()), int((v * v).sum().item()))


def checksum(ctx, x, k, eps, kw, torch):
    import sjamd

    r = ctx.self_join(x, k, eps, **kw)
    if r.stats()["num_batches"] > 1:
        off, ids = r.csr()  # batches live on the host
        rows = torch.arange(x.shape[0], device=x.device)
        fp = fingerprint(torch.from_numpy(off.astype("int64")).to(x.device),
                         torch.from_numpy(ids).to(x.device), rows, torch)
    else:
        o, i, rid = r.device_table()
        n = x.shape[0]
        fp = fingerprint(sjamd.device_copy(o, n + 1, torch.int64), sjamd.device_copy(i, r.total, torch.int32),
                         sjamd.device_copy(rid, n, torch.int32), torch)
    r.free()
    return fp


def canonical_pairs_C(g):
    """Unordered candidate pairs of the canonical grid (cell width eps):
    sum over cells n_c(n_c+1)/2 plus sum over adjacent unordered cell pairs
    n_a n_b, from the device grid's occupancy."""
    import itertools

    import numpy as np

    a = g.arrays()
    lin = a["cellLinear"].astype(np.int64)
    size = (a["indexMax"] - a["indexMin"] + 1).astype(np.int64)
    coord = a["cellCoord"].astype(np.int64)
    _, _, nc, _ = g.dims()
    k = coord.shape[1]
    mult = np.cumprod(np.concatenate([[1], nc[:-1].astype(np.int64)]))
    pos = {int(l): i for i, l in enumerate(lin)}
    C = int((size * (size + 1) // 2).sum())
    offs = [o for o in itertools.product([-1, 0, 1], repeat=k) if o > tuple([0] * k)]
    for o in offs:
        o = np.array(o, dtype=np.int64)
        c2 = coord + o
        ok = np.all((c2 >= 0) & (c2 < nc.astype(np.int64)), axis=1)
        l2 = (c2 * mult).sum(axis=1)
        for i in np.nonzero(ok)[0]:
            j = pos.get(int(l2[i]))
            if j is not None:
                C += int(size[i] * size[j])
    return C


def fingerprint_host(off, ids):
    """fingerprint() of a CSR in original-ID order (rows are point IDs)."""
    import numpy as np

    rows = np.repeat(np.arange(off.size - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    v = ids.astype(np.int64)
    return (int(off[-1]), int((v * (rows + 1)).sum()), int((v * v).sum()))


def cpu_baseline_join(points, k, eps, reps):
    """The oracle's grid join (grid build + LID join, OpenMP on the host
    cores), best of reps; also returns the fingerprint of its table, the
    reference the GPU tables are checked against."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = cores_used()
    best, total, fp = None, 0, None
    for _ in range(reps):
        t0 = time.perf_counter()
        g = O.OracleGrid(points, k, eps)
        off, ids = g.join(O.LID, nthreads=cores, sort_rows=False)
        dt = time.perf_counter() - t0
        total = int(off[-1])
        if fp is None:
            fp = fingerprint_host(off, ids)
        del g, off, ids
        best = dt if best is None else min(best, dt)
    return {"value": total / best, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"full workload ({points.shape[0]} points, grid build + LID join), best of {reps}",
            "seconds": best}, fp


def cpu_baseline_sampled(points, k, eps, stride):
    """Grid join of every stride-th query point against the whole grid (full
    stencil): pairs found per second of that work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = cores_used()
    g = O.OracleGrid(points, k, eps)
    t0 = time.perf_counter()
    cnt = int(g.sample_count(stride))
    dt = time.perf_counter() - t0
    return {"value": cnt / dt, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"queries = every {stride}th of {points.shape[0]} points against the whole grid "
                      f"(full stencil, {cnt} pairs, grid build excluded)", "seconds": dt}


def bench_single(a, torch):
    import sjamd
    from sjamd import datasets as D

    spec = SINGLE[a.workload]
    w = D.WORKLOADS[a.workload]
    pts = D.make_points(a.workload)
    eps = D.workload_eps(a.workload)
    k = w["k"]
    n, d = pts.shape
    form = sjamd.FORMS[a.form or spec["form"]]
    sched_name = a.schedule or spec["schedule"]
    sched = sjamd.WORKQUEUE if sched_name == "workqueue" else sjamd.STRIDED
    fp64_steps = a.fp64_steps if a.fp64_steps is not None else (1 if a.workload == "clusters64d" else a.steps)

    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = sjamd.Context(0, stream=stream)
        x = ctx.to_device(pts)
        torch.cuda.synchronize()
        kw = dict(form=form, schedule=sched, hit_record=not a.no_record)
        if spec["rounds"]:
            r = ctx.self_join(x, k, eps, **kw)
            kw["capacity"] = r.total // spec["rounds"] + 1
            r.free()
        fp32 = run_gpu(ctx, x, k, eps, dict(kw, precision=sjamd.FP32_RECHECK), a.steps, a.warmup, torch)
        fp64 = run_gpu(ctx, x, k, eps, dict(kw, precision=sjamd.FP64), fp64_steps,
                       min(a.warmup, fp64_steps), torch)
        fp_a = checksum(ctx, x, k, eps, dict(kw, precision=sjamd.FP64), torch)
        fp_b = checksum(ctx, x, k, eps, dict(kw, precision=sjamd.FP32_RECHECK), torch)
        C = None
        if a.workload == "clusters64d":
            g = ctx.grid_build(x, k, eps)
            C = canonical_pairs_C(g)
            g.free()
        e2e = end_to_end(ctx, pts, k, eps, dict(kw, precision=sjamd.FP32_RECHECK), torch)
    exact = fp_a == fp_b

    dt, total, st = fp32
    ms = dt * 1e3
    if st["ms_fill"] >= st["ms_count"]:
        kname, kms = "fill", st["ms_fill"]
    else:
        kname, kms = "count", st["ms_count"]
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{a.workload}_{kname}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    # join level: the algorithmic work of the whole join over the step time;
    # kernel level: the same work over the dominant kernel's HIP-event time
    if a.workload == "clusters64d":
        flops = 2.0 * d * C
        achieved = flops / dt / 1e12
        kach = flops / (kms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / MFMA_FP32_PEAK_TFLOPS, "traffic": traffic, "alg_flops": flops,
                "canonical_pairs_C": C, "step_ms": ms,
                "kernel": kname, "kernel_ms": kms, "kernel_achieved": kach,
                "kernel_frac": kach / MFMA_FP32_PEAK_TFLOPS}
    else:
        balg = n * d * 8 + total * 4 + (n + 1) * 8
        achieved = balg / dt / 1e9
        kach = balg / (kms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes": balg, "step_ms": ms,
                "kernel": kname, "kernel_ms": kms, "kernel_achieved": kach, "kernel_frac": kach / HBM_PEAK_GBS}
        if a.workload == "uniform6d":
            # the 6-D walk is bound by distance arithmetic, not bytes: the
            # predicate's flops over the claimed candidates, against the
            # vector FP64 peak, for the step and for the count pass
            flops = predicate_flops(st["candidates"], d)
            hbm = roof
            roof = {"bound": "valu", "achieved": flops / dt / 1e12, "peak": VALU_FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": flops / dt / 1e12 / VALU_FP64_PEAK_TFLOPS, "traffic": traffic,
                    "alg_flops": flops, "candidates": st["candidates"], "step_ms": ms,
                    "kernel": "count", "kernel_ms": st["ms_count"],
                    "kernel_achieved": flops / (st["ms_count"] * 1e-3) / 1e12,
                    "kernel_frac": flops / (st["ms_count"] * 1e-3) / 1e12 / VALU_FP64_PEAK_TFLOPS,
                    "candidates_per_s": st["candidates"] / dt, "hbm": hbm}
    out = {
        "metric": METRIC,
        "value": total / dt,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {spec['desc']}, eps={eps.hex()}",
                   "form": a.form or spec["form"], "schedule": sched_name,
                   "precision": "fp32+fp64-recheck", "pairs": total, "exact_vs_fp64": exact,
                   "result_rounds": int(st["num_batches"])},
        "fp64": {"ms_per_step": fp64[0] * 1e3, "pairs_per_s": fp64[1] / fp64[0], "steps": fp64_steps,
                 "ms_count": fp64[2]["ms_count"], "ms_fill": fp64[2]["ms_fill"]},
        "fp32_recheck": {"ms_per_step": ms, "ms_grid": st["ms_grid"], "ms_workload": st["ms_workload"],
                         "ms_count": st["ms_count"], "ms_scan": st["ms_scan"], "ms_fill": st["ms_fill"],
                         "band_pairs": st["fp32_band_pairs"],
                         "misclassified_before_recheck": st["fp32_misclassified"],
                         "candidates": st["candidates"], "fill_from_record": st["fill_from_record"]},
        "fp32_speedup_vs_fp64": fp64[0] / dt,
        "end_to_end": e2e,
        "roofline": roof,
    }
    out["config"]["exact_vs_oracle"] = None
    if not a.no_cpu:
        if a.workload == "clusters64d":
            out["cpu_baseline"] = cpu_baseline_sampled(pts, k, eps, 1000)
        else:
            out["cpu_baseline"], fp_o = cpu_baseline_join(pts, k, eps, a.cpu_reps)
            out["config"]["exact_vs_oracle"] = fp_a == fp_o and fp_b == fp_o
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def end_to_end(ctx, pts, k, eps, kw, torch, reps=3):
    """The reference's timed region (main.cu:190-197: H2D, join, D2H and the
    host neighbour table): host points -> device -> join -> the reference
    neighbour table on the host (sj_result_neighbor_table: the table
    downloaded into pinned memory and indexed per point).  Median of reps."""
    times = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = ctx.to_device(pts)
        r = ctx.self_join(x, k, eps, **kw)
        tab = r.neighbor_table()
        dt = time.perf_counter() - t0
        r.free()
        del x, tab
        times.append(dt)
    times = sorted(times[1:])
    return {"ms": times[len(times) // 2] * 1e3,
            "region": "host points -> H2D -> grid + join -> D2H of the table -> host neighbour table"}


def bench_many(a, torch):
    import sjamd
    from sjamd import datasets as D

    insts = D.multi_join_instances()
    batch = [(D.instance_points(i), i["d"], D.multi_eps(j)) for j, i in enumerate(insts)]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = sjamd.Context(0, stream=stream)
        res = {}
        for name, prec in [("fp32", sjamd.FP32_RECHECK), ("fp64", sjamd.FP64)]:
            m = ctx.many(batch, form="bunicomp", precision=prec, schedule=sjamd.WORKQUEUE)
            tuned = m.autotune(tune_form=True, reps=2)
            for _ in range(a.warmup):
                m.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dev = 0.0
            for _ in range(a.steps):
                m.run()
                dev += m.times()["ms_run"]
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / max(a.steps, 1)
            st = m.stats()
            flops = sum(predicate_flops(m.stats(i)["candidates"], batch[i][1]) for i in range(len(batch)))
            fps = []
            for i in range(len(batch)):
                o, ids, rid, tot = m.device_table(i)
                n = batch[i][0].shape[0]
                fps.append(fingerprint(sjamd.device_copy(o, n + 1, torch.int64),
                                       sjamd.device_copy(ids, tot, torch.int32),
                                       sjamd.device_copy(rid, n, torch.int32), torch))
            res[name] = dict(dt=dt, dev_ms=dev / max(a.steps, 1), total=st["total_pairs"], stats=st,
                             times=m.times(), tuning=tuned, fps=fps, flops=flops)
            m.free()
    f32 = res["fp32"]
    total = f32["total"]
    exact = res["fp32"]["fps"] == res["fp64"]["fps"]
    kbytes = sum(p.shape[0] * p.shape[1] * 8 + (p.shape[0] + 1) * 8 for p, _, _ in batch) + total * 4
    kms = f32["dev_ms"]
    achieved = kbytes / (kms * 1e-3) / 1e9
    names = {v: k for k, v in sjamd.FORMS.items()}
    tune = dict(f32["tuning"])
    tune["form"] = names[tune["form"]]
    tune["schedule"] = "workqueue" if tune["schedule"] == sjamd.WORKQUEUE else "strided"
    out = {
        "metric": METRIC,
        "value": total / f32["dt"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": f32["dt"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "many64: 64 joins, N log-uniform 1e4..5e5, d in {2,3,4,6,8}, uniform or "
                               "Exp(40) data, 8..128 neighbours/pt, seed 5; one step = one graph replay "
                               "(count + offsets + fill of all 64 joins over the planned grids)",
                   "instances": len(batch), "points": int(sum(p.shape[0] for p, _, _ in batch)),
                   "pairs": total, "exact_vs_fp64": exact, "exact_vs_oracle": None, "autotuned": tune,
                   "precision": "fp32+fp64-recheck"},
        "plan_ms": f32["times"]["ms_plan"], "plan_grid_ms": f32["times"]["ms_grid"],
        "fp64": {"ms_per_step": res["fp64"]["dt"] * 1e3, "pairs_per_s": res["fp64"]["total"] / res["fp64"]["dt"],
                 "autotuned": res["fp64"]["tuning"]},
        "fp32_recheck": {"ms_per_step": f32["dt"] * 1e3, "graph_device_ms": kms,
                         "band_pairs": f32["stats"]["fp32_band_pairs"],
                         "misclassified_before_recheck": f32["stats"]["fp32_misclassified"],
                         "candidates": f32["stats"]["candidates"]},
        "fp32_speedup_vs_fp64": res["fp64"]["dt"] / f32["dt"],
        # bound by distance arithmetic: the predicate's flops over every
        # instance's claimed candidates against the vector FP64 peak
        "roofline": {"bound": "valu", "achieved": f32["flops"] / f32["dt"] / 1e12, "peak": VALU_FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": f32["flops"] / f32["dt"] / 1e12 / VALU_FP64_PEAK_TFLOPS,
                     "traffic": None, "alg_flops": f32["flops"], "step_ms": f32["dt"] * 1e3,
                     "kernel": "graph replay (count + offsets + fill, all instances)", "kernel_ms": kms,
                     "kernel_achieved": f32["flops"] / (kms * 1e-3) / 1e12,
                     "kernel_frac": f32["flops"] / (kms * 1e-3) / 1e12 / VALU_FP64_PEAK_TFLOPS,
                     "candidates_per_s": f32["stats"]["candidates"] / f32["dt"],
                     "hbm": {"bound": "hbm", "achieved": kbytes / f32["dt"] / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": kbytes / f32["dt"] / 1e9 / HBM_PEAK_GBS, "alg_bytes": kbytes,
                             "kernel_achieved": achieved, "kernel_frac": achieved / HBM_PEAK_GBS}},
    }
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        cores = cores_used()
        tot = 0
        sec = 0.0
        fps_o = []
        for p, k, e in batch:
            t0 = time.perf_counter()
            off, ids = O.OracleGrid(p, k, e).join(O.LID, nthreads=cores, sort_rows=False)
            sec += time.perf_counter() - t0
            tot += int(off[-1])
            fps_o.append(fingerprint_host(off, ids))
            del off, ids
        out["config"]["exact_vs_oracle"] = fps_o == res["fp32"]["fps"] and fps_o == res["fp64"]["fps"]
        out["cpu_baseline"] = {"value": tot / sec, "unit": "pairs/s", "cores": cores, "kind": "port",
                               "sample": "all 64 joins (grid build + LID join each), one pass", "seconds": sec}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


COPY_BYTES = 512 * 2**20


def probe(a, torch):
    import sjamd
    from sjamd import datasets as D

    src = torch.ones(COPY_BYTES // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)  # the calibration dispatch: COPY_BYTES read, COPY_BYTES written
    torch.cuda.synchronize()
    del src, dst
    spec = SINGLE[a.workload]
    pts = D.make_points(a.workload)
    eps = D.workload_eps(a.workload)
    k = D.WORKLOADS[a.workload]["k"]
    ctx = sjamd.Context(0)
    x = ctx.to_device(pts)
    form = sjamd.FORMS[a.form or spec["form"]]
    sched = sjamd.WORKQUEUE if (a.schedule or spec["schedule"]) == "workqueue" else sjamd.STRIDED
    # the first join sizes the context's hit-record space; the measured
    # joins then replay their records as in the timed steps
    for prec in (sjamd.FP32_RECHECK, sjamd.FP32_RECHECK, sjamd.FP64):
        r = ctx.self_join(x, k, eps, form=form, schedule=sched, precision=prec)
        st = r.stats()
        print(json.dumps({"precision": prec, "pairs": r.total, "ms_fill": st["ms_fill"],
                          "fill_from_record": st["fill_from_record"]}), flush=True)
        r.free()


def traffic_from(a):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of the traversal kernels,
    scaled by what the same counter reads for the calibration copy."""
    import csv
    import glob

    def read(d, counter):
        rows = []
        for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if r["Counter_Name"] == counter:
                        rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        rows.sort()
        return rows

    out = {}
    for counter, d in (("FETCH_SIZE", a.traffic_from[0]), ("WRITE_SIZE", a.traffic_from[1])):
        rows = read(d, counter)
        first_sj = next(i for i, r in enumerate(rows) if "sj::" in r[1])
        copy = max(rows[:first_sj], key=lambda r: r[2])
        factor = (COPY_BYTES / 1024.0) / copy[2]
        joins = [r for r in rows if "k_join" in r[1]]
        # probe order: FP32+recheck count and fill twice, then FP64; the fill
        # measured is the first FP32 one that replays the count's record
        rp = [r for r in joins if "k_join_rp" in r[1] and r[1].rstrip(")").split("(")[0].endswith("1>")]
        kern = rp[0] if rp else joins[3]
        out[counter] = {"copy_kernel": copy[1], "copy_kib": copy[2], "scale": factor,
                        "kernel": kern[1], "kib": kern[2], "bytes_scaled": kern[2] * 1024.0 * factor}
    res = {"workload": a.workload, "kernel": out["FETCH_SIZE"]["kernel"],
           "fetch_bytes": out["FETCH_SIZE"]["bytes_scaled"], "write_bytes": out["WRITE_SIZE"]["bytes_scaled"],
           "hbm_bytes_per_launch": out["FETCH_SIZE"]["bytes_scaled"] + out["WRITE_SIZE"]["bytes_scaled"],
           "calibration": {c: {k: v for k, v in out[c].items() if k != "kernel"} for c in out},
           "command": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -- python3 bench.py --probe "
                      f"--workload {a.workload}; python3 bench.py --workload {a.workload} --traffic-from F W"}
    path = os.path.join(ROOT, "profiles", f"traffic_{a.workload}_fill.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


def main():
    a = parse()
    if a.traffic_from:
        traffic_from(a)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the join would shard by query cells over a replicated grid)", file=sys.stderr)
        sys.exit(2)
    import torch

    if a.probe:
        probe(a, torch)
        return
    out = bench_many(a, torch) if a.workload == "many64" else bench_single(a, torch)
    print(json.dumps(out))
    if out["config"].get("exact_vs_oracle") is False:
        print("bench.py: the GPU neighbour table differs from the oracle's", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
