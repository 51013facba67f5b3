"""Benchmark: stationary solves per second at N=20, batched FP32 + FP64
refinement against the plain per-solve FP64 path, at <= 1e-10 L1.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sets B] [--seed S] [--no-cpu]

Workload: a seeded synthetic random Boolean network with N = 20 nodes (2^20
states), each rule reading 2 or 3 other nodes (probability 1/2 each), truth
bits ~ Bernoulli(1/2), 5 seeded fixed nodes in the initial state (uniform over
the 2^15 matching states) and up/down rates log-uniform in [0.1, 10].  The
table and graph stages are built once; a step is one batched solve of B rate
sets (all sharing the topology and its zero pattern) on inputs resident in
HBM.  `value` = rate sets solved per second.

Also measured (in the same JSON line):
  plain    the per-solve FP64 path, one rate set per call, graph reused;
  plain_full  table + graph + FP64 solve per rate set (the reference's flow);
  accuracy L1 and max-relative error of the batched results against the
           plain FP64 results (and the oracle for one set);
  roofline the dominant kernel's achieved bandwidth (algorithmic bytes over
           its measured launch time, HIP events on the library stream);
  cpu_baseline  the FP64 CPU oracle (numpy/scipy, 1 thread) on one rate set.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sets", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-sets", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--graph", type=int, default=1, help="replay the numeric sequence as a HIP graph")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="residual reduction an FP32 cycle aims for (library default when omitted)")
    ap.add_argument("--calibrate", action="store_true",
                    help="after the run, copy 1 GiB device to device three times (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def traffic_record(kernel, workload):
    """HBM bytes per launch of `kernel` measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE passes, scaled by the same counters on a 1 GiB copy), when a
    record for this kernel and workload exists under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for r in recs:
        if r.get("kernel") == kernel and r.get("workload") == workload:
            return r
    return None


def copy_calibration(L, ctx):
    """Three 1 GiB device-to-device copies by the library's copy kernel (each
    reads and writes 1 GiB once)."""
    from paraexastolog_amd._lib import check

    nbytes = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    check(L.esl_device_alloc(ctx.handle, nbytes, C.byref(src)), ctx.handle)
    check(L.esl_device_alloc(ctx.handle, nbytes, C.byref(dst)), ctx.handle)
    for _ in range(3):
        check(L.esl_memcpy_d2d(ctx.handle, dst, src, nbytes), ctx.handle)
    L.esl_device_free(ctx.handle, src)
    L.esl_device_free(ctx.handle, dst)
    print(json.dumps({"calibration_copy_bytes": nbytes, "launches": 3}), file=sys.stderr)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: only the one-GPU form exists (asked for --gpus {a.gpus}); rate sets are independent, "
              "so more GPUs would split the sets -- that multi-GPU path is not built.", file=sys.stderr)
        sys.exit(2)

    from paraexastolog_amd import cu_context, model_builder, solve_opts, synth, transition_table
    from paraexastolog_amd._lib import SolveReport, check, lib

    L = lib()
    net, rng = synth.config("c1", a.seed)
    N = net.N
    n = 1 << N
    B = a.sets
    rates = synth.rate_sets(rng, N, B)
    fixed, vals = net.fixed_spec()

    ctx = cu_context(0)
    model = model_builder().parse_model(net.bnet())
    x0 = np.zeros(n)
    from paraexastolog_amd.api import initial_state

    x0 = initial_state(model.nodes, fixed, vals).state

    # shared symbolic part: table + graph, built once
    t0 = time.perf_counter()
    table = transition_table(ctx, model, rates[0])
    table.construct_table()
    check(L.esl_graph_reorganize(ctx.handle, table.handle), ctx.handle)
    t_graph = time.perf_counter() - t0
    sizes = table.sizes()
    stage_ms = table.stage_ms()

    def dalloc(nbytes):
        p = C.c_void_p()
        check(L.esl_device_alloc(ctx.handle, nbytes, C.byref(p)), ctx.handle)
        return p

    def h2d(dst, arr):
        arr = np.ascontiguousarray(arr)
        check(L.esl_memcpy_h2d(ctx.handle, dst, arr.ctypes.data_as(C.c_void_p), arr.nbytes), ctx.handle)

    def d2h(arr, src):
        check(L.esl_memcpy_d2h(ctx.handle, arr.ctypes.data_as(C.c_void_p), src, arr.nbytes), ctx.handle)

    d_rates = dalloc(rates.nbytes)
    h2d(d_rates, rates)
    d_x0 = dalloc(x0.nbytes)
    h2d(d_x0, x0)
    n_check = min(a.plain_sets, B)
    d_final = dalloc(8 * n * B)
    d_plain = dalloc(8 * n * n_check)

    # the rate-independent part (reachability, terminal masks, unknowns, coupling
    # structure) is built once per topology and initial state
    t = time.perf_counter()
    pl = C.c_void_p()
    check(L.esl_plan_create(ctx.handle, table.handle, d_x0, C.byref(pl)), ctx.handle)
    t_plan = time.perf_counter() - t
    info = np.zeros(8, dtype=np.int64)
    check(L.esl_plan_query(pl, info.ctypes.data_as(C.POINTER(C.c_int64))))

    def solve_batch(count, d_r, d_out, precision, tol, graph=0):
        o = solve_opts(precision=precision, tol=tol, inner_rtol=a.inner_rtol)
        o.graph = graph
        rep = SolveReport()
        check(L.esl_plan_solve(ctx.handle, pl, count, d_r, d_out, C.byref(o), C.byref(rep)), ctx.handle)
        return rep

    # ---- plain FP64, one rate set per call (plan reused)
    plain_times = []
    plain_reports = []
    for k in range(n_check):
        r_k = C.c_void_p(d_rates.value + 8 * 2 * N * k)
        out_k = C.c_void_p(d_plain.value + 8 * n * k)
        t = time.perf_counter()
        rep = solve_batch(1, r_k, out_k, 0, 1e-13)
        plain_times.append(time.perf_counter() - t)
        plain_reports.append(rep.as_dict())
    plain_s = float(np.median(plain_times[1:] if len(plain_times) > 1 else plain_times))

    # ---- plain, full reference flow per rate set: table + graph + FP64 solve
    full_times = []
    for k in range(min(3, B)):
        t = time.perf_counter()
        tb = transition_table(ctx, model, rates[k])
        tb.construct_table()
        check(L.esl_graph_reorganize(ctx.handle, tb.handle), ctx.handle)
        o = solve_opts(precision=0, tol=1e-13)
        rep = SolveReport()
        out_k = C.c_void_p(d_plain.value)
        r_k = C.c_void_p(d_rates.value + 8 * 2 * N * k)
        # the one-shot entry point: plan built and dropped inside the call
        check(L.esl_solve_batch_device(ctx.handle, tb.handle, 1, r_k, d_x0, out_k, C.byref(o), C.byref(rep)),
              ctx.handle)
        full_times.append(time.perf_counter() - t)
        del tb
    # restore set 0's plain result (buffer 0 was reused above)
    solve_batch(1, d_rates, d_plain, 0, 1e-13)
    plain_full_s = float(np.median(full_times))

    # ---- batched + refined: warmup, then K timed steps
    tol = 1e-12
    # kernel timing events on (they are captured into the HIP graph as event nodes)
    check(L.esl_profile(ctx.handle, 1), ctx.handle)
    for _ in range(max(a.warmup, 1)):  # the first call captures the HIP graph
        solve_batch(B, d_rates, d_final, a.precision, tol, graph=a.graph)
    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731
    sync()
    check(L.esl_profile(ctx.handle, 1), ctx.handle)  # clears the warm-up figures
    reports = []
    t = time.perf_counter()
    for _ in range(a.steps):
        reports.append(solve_batch(B, d_rates, d_final, a.precision, tol, graph=a.graph).as_dict())
    sync()
    elapsed = time.perf_counter() - t
    stats = []
    i = 0
    name = C.create_string_buffer(128)
    while True:
        launches, ms, by = C.c_int64(), C.c_double(), C.c_double()
        st = L.esl_profile_stat(ctx.handle, i, name, 128, C.byref(launches), C.byref(ms), C.byref(by))
        if st != 0:
            break
        stats.append((name.value.decode(), launches.value, ms.value, by.value))
        i += 1
    check(L.esl_profile(ctx.handle, 0), ctx.handle)
    ms_per_step = 1000.0 * elapsed / a.steps
    value = B * a.steps / elapsed

    # ---- accuracy: batched vs plain FP64 on the checked sets
    xb = np.zeros((n_check, n))
    d2h(xb, d_final)  # the first n_check sets (set-major layout)
    xp = np.zeros((n_check, n))
    d2h(xp, d_plain)
    l1 = [float(np.abs(xb[k] - xp[k]).sum()) for k in range(n_check)]
    nzp = xp != 0
    maxrel = [float((np.abs(xb[k] - xp[k])[nzp[k]] / np.abs(xp[k][nzp[k]])).max()) if nzp[k].any() else 0.0
              for k in range(n_check)]
    sums = np.abs(xb.sum(axis=1) - 1.0).max()

    # ---- dominant kernel roofline
    dom = max(stats, key=lambda s: s[2]) if stats else ("none", 0, 0.0, 0.0)
    avg_ms = dom[2] / max(dom[1], 1)
    bytes_per_launch = dom[3] / max(dom[1], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    # ---- CPU baseline: the FP64 oracle on one rate set (bounded sample)
    cpu = None
    oracle_l1 = None
    if not a.no_cpu:
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        from oracle import pipeline as P
        from oracle.bnet import read_bnet_text

        t = time.perf_counter()
        om = read_bnet_text(net.bnet())
        ip, rows, _ = P.transition_table(om, rates[0])
        gr = P.transition_graph(n, ip, rows)
        xo = P.solve_final(n, ip, rows, rates[0], P.initial_state(om, fixed, vals), gr)
        cpu_s = time.perf_counter() - t
        oracle_l1 = float(np.abs(xp[0] - xo).sum())
        cpu = {"value": 1.0 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"1 rate set of this workload through the FP64 oracle (table, SCCs, order, "
                         f"reachability-pruned solve), numpy/scipy single thread, {cpu_s:.1f} s"}

    rep0 = reports[-1]
    out = {
        "metric": "stationary solves/sec at N=20 (batched+refined vs plain FP64) at <=1e-10 L1",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64" if a.precision == 1 else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"seeded random Boolean network N=20 (K in {{2,3}}, 5 fixed nodes), generator v1 seed "
                        f"{a.seed}; {B} log-uniform [0.1,10] rate sets per step, batched FP32 GMRES + FP64 "
                        f"residual refinement, tol {tol:g}",
            "states": int(sizes[0]), "transitions": int(sizes[1]), "sccs": int(sizes[2]),
            "terminal_sccs": int(sizes[3]), "rate_sets": B,
            "unknowns_per_set": int(rep0["unknowns"]), "terminals_reached": int(rep0["terminals_reached"]),
        },
        "plain_fp64": {"solves_per_s": 1.0 / plain_s, "ms_per_solve": 1000 * plain_s,
                       "stage_ms": plain_reports[-1]["ms_stage"]},
        "plain_full_pipeline": {"solves_per_s": 1.0 / plain_full_s, "ms_per_solve": 1000 * plain_full_s},
        "speedup_vs_plain_fp64": value * plain_s,
        "shared_stages_ms": {"table": float(stage_ms[0]), "graph": float(stage_ms[1]),
                             "wall_table_graph": 1000 * t_graph, "plan": 1000 * t_plan},
        "plan": {"unknowns": int(info[0]), "terminal_unknowns": int(info[1]), "ambiguous": int(info[2]),
                 "terminals_reached": int(info[3]), "coupling_nnz": int(info[5])},
        "accuracy": {"max_l1_vs_plain_fp64": max(l1), "max_rel_vs_plain_fp64": max(maxrel),
                     "plain_fp64_l1_vs_oracle": oracle_l1, "max_abs_sum_minus_1": float(sums),
                     "max_residual": max(r["residual"] for r in reports),
                     "gmres_iterations": rep0["iterations"], "refinement_cycles": rep0["refinements"]},
        "kernels": [{"name": s[0], "launches": s[1], "ms": s[2], "algorithmic_GB": s[3] / 1e9} for s in stats],
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch, "traffic": None},
        "cpu_baseline": cpu,
    }
    tr = traffic_record(dom[0], f"N=20 seed {a.seed} sets {B}")
    if tr is not None:
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        out["roofline"]["traffic_source"] = tr["source"]
    print(json.dumps(out))
    sys.stdout.flush()
    if a.calibrate:
        copy_calibration(L, ctx)
    if max(l1) > 1e-10:
        print(f"bench.py: accuracy bar missed: L1 {max(l1):.3e} > 1e-10", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
