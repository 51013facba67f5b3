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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--sets", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=None,
                    help="network seed (default: 0 for n20 and uneven, 11 for n16 and warm, 8 for n22)")
    ap.add_argument("--plain-sets", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cap", type=float, default=20.0,
                    help="wall-clock cap (s) of each run of the reference CPU path restatement")
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--graph", type=int, default=1, help="replay the numeric sequence as a HIP graph")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="residual reduction an FP32 cycle aims for (library default when omitted)")
    ap.add_argument("--autotune", type=int, default=1, help="pick the schedule with esl_plan_autotune first")
    ap.add_argument("--schedule", type=int, default=None,
                    help="pin a schedule instead of autotuning (0 per-set, 1 per-set global, 2 device-wide, "
                         "3 per-set LDS with the basis in global memory, 4/5 per-set at 3/2 workgroups per CU)")
    ap.add_argument("--restart", type=int, default=32, help="GMRES restart length of a pinned --schedule")
    ap.add_argument("--workload", default="n20", choices=["n20", "n16", "uneven", "n22", "warm", "seeds"],
                    help="n20 (default) is the headline metric; the others print their own JSON line")
    ap.add_argument("--epochs", type=int, default=100, help="training epochs of the warm workload")
    ap.add_argument("--calibrate", action="store_true",
                    help="after the run, copy 1 GiB device to device three times (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def profile_stats(L, ctx):
    """(name, launches, ms, algorithmic bytes) of every kernel timed since profiling started."""
    stats = []
    name = C.create_string_buffer(128)
    i = 0
    while True:
        launches, ms, by = C.c_int64(), C.c_double(), C.c_double()
        if L.esl_profile_stat(ctx.handle, i, name, 128, C.byref(launches), C.byref(ms), C.byref(by)) != 0:
            break
        stats.append((name.value.decode(), launches.value, ms.value, by.value))
        i += 1
    return stats


TABLE_KERNELS = ("k_flip_mask", "k_table_fill")
MASK_PASS_KERNELS = ("k_trim", "k_color", "k_close", "k_height_check")


def stage_bandwidth(stats, n, m, counts=None):
    """Per-stage device time, algorithmic bytes and GB/s of the table and graph
    stages.  Graph stage: passes over the per-state flip masks, counted by the
    library (`counts`, esl_graph_stage_counts; passes that found their
    predecessor converged return at once and are not counted), with the
    compulsory-bytes form 4 n * passes + 3 * 4 n (mask per pass; labels,
    heights, order written once) as the algorithmic bytes."""
    tab = [s for s in stats if s[0] in TABLE_KERNELS]
    gra = [s for s in stats if s[0] not in TABLE_KERNELS and s[0].startswith("k_")]
    out = {}
    for key, ks in (("table", tab), ("graph", gra)):
        ms = sum(k[2] for k in ks)
        by = sum(k[3] for k in ks)
        out[key] = {"kernel_ms": ms, "algorithmic_GB": by / 1e9, "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                    "kernels": {k[0]: {"launches": k[1], "ms": k[2]} for k in ks}}
    out["table"]["compulsory_GB"] = (4.0 * (n + 1) + 8.0 * m + 4.0 * n) / 1e9  # indptr, rows, cols, masks written
    passes = counts["mask_passes"] if counts else sum(k[1] for k in gra if k[0] in MASK_PASS_KERNELS)
    g = out["graph"]
    g["mask_passes"] = passes
    g["compulsory_GB"] = (4.0 * n * passes + 12.0 * n) / 1e9
    g["algorithmic_GB"] = g["compulsory_GB"]
    g["GBps"] = g["compulsory_GB"] / (g["kernel_ms"] * 1e-3) if g["kernel_ms"] > 0 else None
    if counts:
        g["scc_rounds"] = counts["scc_rounds"]
        g["host_round_trips"] = counts["host_round_trips"]
    return out


def traffic_record(kernel, workload, restart=None, name="traffic.json"):
    """HBM bytes per launch of `kernel` measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE passes, scaled by the same counters on a 1 GiB copy), when a
    record for this kernel, workload and restart length exists under profiles/
    (name="issue.json": its VALU instruction count and issue fraction, from the
    SQ counters)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    for r in recs:
        if r.get("kernel") == kernel and r.get("workload") == workload and r.get("restart") in (None, restart):
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


SCHEDULE_NAMES = ["per-set LDS-resident", "per-set global", "device-wide", "per-set LDS, basis in memory",
                  "per-set LDS-resident, 3 workgroups per CU", "per-set LDS-resident, 2 workgroups per CU"]


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_reference_path(seed, kernel, threads, cap_s, mem_gib=32):
    """The restatement of the reference's CPU path (python/exastolog.py, in
    oracle/exastolog_cpu.py) on rate set 0 of the N=20 workload, in a child
    process: OMP/OPENBLAS threads = `threads`, wall-clock cap `cap_s` (the
    child stops itself when its determinant loop would pass the cap; the parent
    kills it 5 s after), address space `mem_gib`.  Must run before this process
    touches the GPU (the child is started by fork + exec)."""
    import resource
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads),
               MKL_NUM_THREADS=str(threads))
    lim = mem_gib << 30

    def limit():
        resource.setrlimit(resource.RLIMIT_AS, (lim, lim))

    cmd = [sys.executable, "-m", "oracle.exastolog_cpu", "--config", "c1", "--seed", str(seed), "--kernel", kernel,
           "--cap", str(cap_s)]
    t = time.perf_counter()
    killed = False
    try:
        p = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=cap_s + 5,
                           preexec_fn=limit)
        out = p.stdout
    except subprocess.TimeoutExpired as e:
        out = e.stdout.decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
        killed = True
    wall = time.perf_counter() - t
    stages, result = [], None
    for line in out.splitlines():
        try:
            rec = json.loads(line)
        except ValueError:
            continue
        if "stage" in rec:
            stages.append([rec["stage"], round(rec["s"], 3)])
        elif "result" in rec:
            result = rec["result"]
    if result is None:
        done = {st for st, _ in stages}
        if "column nullspace (terminal kernels)" in done:
            at = "row nullspace (SuperLU on N^T)"
        elif "graph (weak components, SCCs, toposort, sorted sub-tables)" in done:
            at = "column nullspace (terminal kernels)"
        else:
            at = "transition table / graph"
        result = {"status": "DNF", "stage": at,
                  "detail": f"killed at the {cap_s:.0f} s cap" if killed else "child ended without a result"}
    return {"kernel": "determinants (exastolog.py:333-354)" if kernel == "det" else
            "row replacement (labelled variant of exastolog.py:333-354)", "threads": threads, "cap_s": cap_s,
            "wall_s": round(wall, 2), "stages_s": stages, "result": result}


def cpu_oracle_line(net, rates0, threads):
    """The project's FP64 oracle (reachability-pruned solve) on one rate set, in
    this process with BLAS threads limited by threadpoolctl when available."""
    from oracle import pipeline as P
    from oracle.bnet import read_bnet_text

    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    n = 1 << net.N
    fixed, vals = net.fixed_spec()
    ctxm = threadpool_limits(limits=threads) if threadpool_limits else None
    try:
        t = time.perf_counter()
        om = read_bnet_text(net.bnet())
        ip, rows, _ = P.transition_table(om, rates0)
        gr = P.transition_graph(n, ip, rows)
        xo = P.solve_final(n, ip, rows, rates0, P.initial_state(om, fixed, vals), gr)
        sec = time.perf_counter() - t
    finally:
        if ctxm is not None:
            ctxm.__exit__(None, None, None)
    return sec, xo


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: only the one-GPU form exists (asked for --gpus {a.gpus}); rate sets are independent, "
              "so more GPUs would split the sets -- that multi-GPU path is not built.", file=sys.stderr)
        sys.exit(2)

    from paraexastolog_amd import cu_context, model_builder, solve_opts, synth, transition_table
    from paraexastolog_amd._lib import SolveReport, check, lib

    if a.workload != "n20":
        WORKLOADS[a.workload](a, cu_context(0))
        return

    if a.seed is None:
        a.seed = 0
    net, rng = synth.config("c1", a.seed)
    N = net.N
    n = 1 << N
    B = a.sets
    rates = synth.rate_sets(rng, N, B)
    fixed, vals = net.fixed_spec()

    # ---- CPU baselines first: the child processes of the reference-path leg are
    # started before this process touches the GPU
    cpu = None
    xo = None
    if not a.no_cpu:
        host = host_cpu()
        nthr = max(1, min(host["usable"], 16))
        ref_lines = [cpu_reference_path(a.seed, "det", nthr, a.cpu_cap),
                     cpu_reference_path(a.seed, "det", 1, a.cpu_cap),
                     cpu_reference_path(a.seed, "rowrep", nthr, a.cpu_cap)]
        o1, xo = cpu_oracle_line(net, rates[0], 1)
        on, _ = cpu_oracle_line(net, rates[0], nthr)
        ok_ref = [r for r in ref_lines if r["result"].get("status") == "ok"]
        if ok_ref:
            best = min(ok_ref, key=lambda r: r["result"]["s"])
            value, kind, cores = 1.0 / best["result"]["s"], "port", best["threads"]
            sample = (f"1 rate set of this workload through the restatement of the reference's CPU path "
                      f"(python/exastolog.py, {best['kernel']}), {best['threads']} BLAS threads")
        else:
            value, kind, cores = 1.0 / o1, "port", 1
            sample = (f"1 rate set of this workload through the project's FP64 oracle (table, SCCs, order, "
                      f"reachability-pruned solve), numpy/scipy, 1 thread, {o1:.1f} s; the reference CPU path "
                      f"(restated) did not finish on this workload, see reference_path")
        cpu = {"value": value, "unit": "solves/s", "cores": cores, "kind": kind, "sample": sample,
               "host": dict(host, threads_used=nthr),
               "reference_path": ref_lines,
               "oracle_pruned": {"s_1_thread": round(o1, 3), f"s_{nthr}_threads": round(on, 3)}}

    L = lib()

    ctx = cu_context(0)
    model = model_builder().parse_model(net.bnet())
    x0 = np.zeros(n)
    from paraexastolog_amd.api import initial_state

    x0 = initial_state(model.nodes, fixed, vals).state

    # shared symbolic part: table + graph, built once (kernel timing on, to report
    # each stage's algorithmic bandwidth), after one untimed build (module load,
    # first-touch allocations)
    warm = transition_table(ctx, model, rates[0])
    warm.construct_table()
    check(L.esl_graph_reorganize(ctx.handle, warm.handle), ctx.handle)
    del warm
    check(L.esl_profile(ctx.handle, 1), ctx.handle)
    t0 = time.perf_counter()
    table = transition_table(ctx, model, rates[0])
    table.construct_table()
    check(L.esl_graph_reorganize(ctx.handle, table.handle), ctx.handle)
    t_graph = time.perf_counter() - t0
    sizes = table.sizes()
    stage_ms = table.stage_ms()
    stage_stats = profile_stats(L, ctx)
    check(L.esl_profile(ctx.handle, 0), ctx.handle)
    cnt = np.zeros(3, dtype=np.int64)
    check(L.esl_graph_stage_counts(table.handle, cnt.ctypes.data_as(C.POINTER(C.c_int64))))
    stages = stage_bandwidth(stage_stats, n, int(sizes[1]),
                             {"mask_passes": int(cnt[0]), "scc_rounds": int(cnt[1]), "host_round_trips": int(cnt[2])})

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

    def solve_batch(count, d_r, d_out, precision, tol, graph=0, restart=32):
        o = solve_opts(precision=precision, tol=tol, inner_rtol=a.inner_rtol, restart=restart)
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

    # ---- batched + refined: the autotuner picks the schedule on this batch, then
    # warmup and K timed steps with the tuned schedule (restart 0 = tuned)
    tol = 1e-12
    schedule = None
    restart = 32
    if a.schedule is not None:
        check(L.esl_plan_set_schedule(pl, a.schedule, a.restart))
        schedule = {"path": SCHEDULE_NAMES[a.schedule],
                    "restart": a.restart, "pinned": True}
        restart = 0
    elif a.autotune:
        o = solve_opts(precision=a.precision, tol=tol, inner_rtol=a.inner_rtol)
        o.graph = a.graph
        ch = (C.c_int32 * 2)()
        best = C.c_double()
        t = time.perf_counter()
        check(L.esl_plan_autotune(ctx.handle, pl, B, d_rates, d_final, C.byref(o), ch, C.byref(best)), ctx.handle)
        schedule = {"path": SCHEDULE_NAMES[ch[0]], "restart": int(ch[1]),
                    "best_ms": best.value, "tuning_s": time.perf_counter() - t}
        restart = 0
    # kernel timing events on (they are captured into the HIP graph as event nodes)
    check(L.esl_profile(ctx.handle, 1), ctx.handle)
    for _ in range(max(a.warmup, 1)):  # the first call captures the HIP graph
        solve_batch(B, d_rates, d_final, a.precision, tol, graph=a.graph, restart=restart)
    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731
    sync()
    check(L.esl_profile(ctx.handle, 1), ctx.handle)  # clears the warm-up figures
    # the options and the report structures are built before the timed region;
    # each step is one esl_plan_solve call
    o_step = solve_opts(precision=a.precision, tol=tol, inner_rtol=a.inner_rtol, restart=restart)
    o_step.graph = a.graph
    raw = [SolveReport() for _ in range(a.steps)]
    t = time.perf_counter()
    for k in range(a.steps):
        check(L.esl_plan_solve(ctx.handle, pl, B, d_rates, d_final, C.byref(o_step), C.byref(raw[k])), ctx.handle)
    sync()
    elapsed = time.perf_counter() - t
    reports = [r.as_dict() for r in raw]
    stats = profile_stats(L, ctx)
    check(L.esl_profile(ctx.handle, 0), ctx.handle)
    ms_per_step = 1000.0 * elapsed / a.steps
    # per-set start / end of the last step's solve launch (device wall clock): the
    # launch is bound by its slowest set's dependent chain, not by bandwidth
    t_start, t_end = np.zeros(B), np.zeros(B)
    latency = None
    if L.esl_plan_set_times(ctx.handle, pl, B, t_start.ctypes.data_as(C.POINTER(C.c_double)),
                            t_end.ctypes.data_as(C.POINTER(C.c_double))) == 0:
        dur = t_end - t_start
        span = float(t_end.max() - t_start.min())
        latency = {"bound": "slowest set (dependent chain per set)", "launch_span_ms": span,
                   "set_ms": {"mean": float(dur.mean()), "median": float(np.median(dur)),
                              "p90": float(np.percentile(dur, 90)), "max": float(dur.max()),
                              "min": float(dur.min())},
                   "slot_busy_frac": float(dur.sum() / (B * span)) if span > 0 else None,
                   "note": "slot_busy_frac = sum of set times / (sets x launch span): every set has its own "
                           "workgroup slot (1024 slots at 4 workgroups per CU), so 1 - frac is the time "
                           "slots idle waiting for the slowest set"}
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

    # ---- accuracy of the whole batch: every set against one batched FP64 run
    # (precision 0, tol 1e-13), compared in chunks of 64 sets on the host
    d_fp64 = dalloc(8 * n * B)
    rep64 = solve_batch(B, d_rates, d_fp64, 0, 1e-13).as_dict()
    full_l1, full_rel, full_sum = np.zeros(B), np.zeros(B), np.zeros(B)
    chunk = 64
    for c0 in range(0, B, chunk):
        cnt = min(chunk, B - c0)
        got = np.empty((cnt, n))
        ref = np.empty((cnt, n))
        d2h(got, C.c_void_p(d_final.value + 8 * n * c0))
        d2h(ref, C.c_void_p(d_fp64.value + 8 * n * c0))
        diff = np.abs(got - ref)
        full_l1[c0:c0 + cnt] = diff.sum(axis=1)
        nz = ref != 0
        rel = np.where(nz, diff / np.where(nz, np.abs(ref), 1.0), 0.0)
        full_rel[c0:c0 + cnt] = rel.max(axis=1)
        full_sum[c0:c0 + cnt] = np.abs(got.sum(axis=1) - 1.0)
        if c0 == 0:  # the first two sets also go to the oracle below
            xo_sets = got[:2].copy()
    L.esl_device_free(ctx.handle, d_fp64)
    sums = float(full_sum.max())

    # ---- two sets against the FP64 oracle (numpy/scipy, reachability-pruned solve)
    from oracle import pipeline as P
    from oracle.bnet import read_bnet_text

    om = read_bnet_text(net.bnet())
    oracle_sets = []
    for k in range(min(2, B)):
        if k == 0 and xo is not None:
            xk = xo
        else:
            ip, rows_, _ = P.transition_table(om, rates[k])
            gr = P.transition_graph(n, ip, rows_)
            xk = P.solve_final(n, ip, rows_, rates[k], P.initial_state(om, fixed, vals), gr)
            if k == 0:
                xo = xk
        oracle_sets.append({"set": k, "l1_batched_vs_oracle": float(np.abs(xo_sets[k] - xk).sum())})

    # ---- dominant kernel roofline
    dom = max(stats, key=lambda s: s[2]) if stats else ("none", 0, 0.0, 0.0)
    avg_ms = dom[2] / max(dom[1], 1)
    bytes_per_launch = dom[3] / max(dom[1], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    oracle_l1 = float(np.abs(xp[0] - xo).sum()) if xo is not None else None

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
        "stage_bandwidth": stages,
        "plan": {"unknowns": int(info[0]), "terminal_unknowns": int(info[1]), "ambiguous": int(info[2]),
                 "terminals_reached": int(info[3]), "coupling_nnz": int(info[5])},
        "accuracy": {"full_batch_vs_batched_fp64": {
                         "sets": B, "max_l1": float(full_l1.max()), "median_l1": float(np.median(full_l1)),
                         "worst_set": int(full_l1.argmax()), "max_rel": float(full_rel.max()),
                         "median_rel": float(np.median(full_rel)), "fp64_residual": rep64["residual"],
                         "fp64_tol": 1e-13},
                     "vs_oracle": oracle_sets,
                     "max_l1_vs_plain_fp64": max(l1), "max_rel_vs_plain_fp64": max(maxrel),
                     "plain_sets_checked": n_check,
                     "plain_fp64_l1_vs_oracle": oracle_l1, "max_abs_sum_minus_1": float(sums),
                     "max_residual": max(r["residual"] for r in reports),
                     "gmres_iterations": rep0["iterations"], "refinement_cycles": rep0["refinements"],
                     "mean_iterations_per_set": rep0["iterations_sum"] / B,
                     "fp64_fallback_sets": rep0["fp64_fallback"]},
        "schedule": schedule,
        "kernels": [{"name": s[0], "launches": s[1], "ms": s[2], "algorithmic_GB": s[3] / 1e9} for s in stats],
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "avg_launch_ms": avg_ms,
                     "bytes_per_launch": bytes_per_launch, "traffic": None},
        "cpu_baseline": cpu,
        "latency_roof": latency,
    }
    tr = traffic_record(dom[0], f"N=20 seed {a.seed} sets {B}", (schedule or {}).get("restart"))
    if tr is not None:
        out["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        out["roofline"]["traffic_source"] = tr["source"]
    # the issue roof beside the HBM one: VALU instructions per launch against one
    # wave instruction per CU per cycle (rocprofv3 SQ counters, kernel alone)
    ir = traffic_record(dom[0], f"N=20 seed {a.seed} sets {B}", (schedule or {}).get("restart"), "issue.json")
    if ir is not None:
        out["issue_roof"] = {"bound": "valu issue", "valu_insts_per_launch": ir["valu_insts_per_launch"],
                             "frac_alone": ir["valu_issue_frac"], "launch_ms_counted": ir["launch_ms_counted"],
                             "frac_in_step": (ir["valu_insts_per_launch"] / (256 * 2.4e9 * avg_ms * 1e-3)
                                              if avg_ms > 0 else None),
                             "lds_bank_conflict_frac": ir["lds_bank_conflict_frac"], "source": ir["source"]}
    print(json.dumps(out))
    sys.stdout.flush()
    if a.calibrate:
        copy_calibration(L, ctx)
    worst = max(max(l1), float(full_l1.max()), max(o["l1_batched_vs_oracle"] for o in oracle_sets))
    if worst > 1e-10:
        print(f"bench.py: accuracy bar missed: L1 {worst:.3e} > 1e-10", file=sys.stderr)
        sys.exit(1)


# ---------------------------------------------------------------------------
# further workloads (each prints its own JSON line; the default run is the one above)


class _Topo:
    """One network on the device: model, table + graph, initial state, plan."""

    def __init__(self, ctx, net, rates0):
        from paraexastolog_amd import initial_state, model_builder, solve_plan, transition_graph, transition_table
        from paraexastolog_amd.surrogate import DeviceArray

        self.net = net
        self.n = 1 << net.N
        t0 = time.perf_counter()
        self.model = model_builder().parse_model(net.bnet())
        self.table = transition_table(ctx, self.model, rates0)
        self.table.construct_table()
        self.graph = transition_graph(ctx, self.table)
        self.graph.reorganize_vertices()
        fixed, vals = net.fixed_spec()
        self.x0 = initial_state(self.model.nodes, fixed, vals).state
        self.d_x0 = DeviceArray.from_host(ctx, self.x0)
        self.plan = solve_plan(ctx, self.table, self.d_x0.ptr)
        self.setup_s = time.perf_counter() - t0
        self.n_t = int(self.graph.sccs_offsets[self.graph.terminals_count])
        self.info = self.plan.info()


def _timed(fn, steps):
    fn()  # warm-up (graph capture, buffer growth)
    t = time.perf_counter()
    rep = None
    for _ in range(steps):
        rep = fn()
    return (time.perf_counter() - t) / steps, rep


def _errors(x, ref):
    """max L1 and max relative error (over nonzero reference entries) per row."""
    l1 = np.abs(x - ref).sum(axis=1)
    nz = ref != 0
    rel = np.where(nz, np.abs(x - ref) / np.where(nz, np.abs(ref), 1.0), 0.0).max(axis=1)
    return float(l1.max()), float(np.median(l1)), float(rel.max())


def workload_n16(a, ctx):
    """One N=16 topology (K=2, generator v1 seed 11: one terminal SCC of 2048
    states), 1024 rate sets sharing the plan: FP64 and FP32+refinement, HIP graph
    off/on, autotuned schedule; against the plain per-solve FP64 path."""
    from paraexastolog_amd import synth
    from paraexastolog_amd.surrogate import DeviceArray

    seed = 11 if a.seed is None else a.seed
    net, rng = synth.config("c2", seed)
    B = a.sets
    rates = synth.rate_sets(rng, net.N, B)
    tp = _Topo(ctx, net, rates[0])
    dr = DeviceArray.from_host(ctx, rates)
    df = DeviceArray(ctx, (B, tp.n), np.float64)
    dp = DeviceArray(ctx, (8, tp.n), np.float64)
    plain = []
    for k in range(8):
        t = time.perf_counter()
        tp.plan.solve(1, dr.offset(2 * net.N * k), dp.offset(tp.n * k), precision=0, tol=1e-13)
        plain.append(time.perf_counter() - t)
    plain_s = float(np.median(plain[1:]))
    ref = dp.download()
    out = {}
    for prec in (0, 1):
        for graph in (0, 1):
            sec, rep = _timed(lambda: tp.plan.solve(B, dr.ptr, df.ptr, precision=prec, tol=1e-12, graph=graph),
                              a.steps)
            l1, l1_med, rel = _errors(df.download()[:8], ref)
            out[f"{'fp32+fp64' if prec else 'fp64'}_graph{graph}"] = {
                "sets_per_s": B / sec, "ms_per_batch": 1000 * sec, "max_l1_vs_plain": l1, "max_rel_vs_plain": rel,
                "mean_iterations": rep["iterations_sum"] / B}
    path, restart, best_ms = tp.plan.autotune(B, dr.ptr, df.ptr, precision=1, tol=1e-12, graph=1)
    sec, rep = _timed(lambda: tp.plan.solve(B, dr.ptr, df.ptr, precision=1, tol=1e-12, graph=1, restart=0), a.steps)
    out["autotuned"] = {"sets_per_s": B / sec, "ms_per_batch": 1000 * sec,
                        "schedule": SCHEDULE_NAMES[path], "restart": restart}
    best = max(v["sets_per_s"] for v in out.values())
    print(json.dumps({
        "metric": "rate sets/s, N=16 topology, batch of shared-plan rate sets", "value": best, "unit": "sets/s",
        "workload": f"N=16 K=2 generator v1 seed {seed}; {B} log-uniform [0.1,10] rate sets; tol 1e-12",
        "plan": tp.info, "setup_ms_table_graph_plan": 1000 * tp.setup_s,
        "plain_fp64": {"sets_per_s": 1.0 / plain_s, "ms_per_solve": 1000 * plain_s},
        "speedup_vs_plain_fp64": best * plain_s, "variants": out}))


def workload_uneven(a, ctx):
    """Eight networks of N = 12..18, 18 (K=2), 128 rate sets each, in one call
    (esl_plan_solve_many, concurrent streams, one HIP graph) against the same
    plans solved one after another."""
    from paraexastolog_amd import solve_plan, synth
    from paraexastolog_amd.surrogate import DeviceArray

    per = 128
    a.seed = 0 if a.seed is None else a.seed
    nets = synth.uneven(a.seed)
    tops, drs, dfs, rates_all = [], [], [], []
    for net, rng in nets:
        rates = synth.rate_sets(rng, net.N, per)
        tp = _Topo(ctx, net, rates[0])
        tops.append(tp)
        rates_all.append(rates)
        drs.append(DeviceArray.from_host(ctx, rates))
        dfs.append(DeviceArray(ctx, (per, tp.n), np.float64))
    plans = [t.plan for t in tops]
    res = {}
    for prec in (0, 1):
        seq_s, _ = _timed(lambda: [t.plan.solve(per, d.ptr, f.ptr, precision=prec, tol=1e-12, graph=1)
                                   for t, d, f in zip(tops, drs, dfs)], a.steps)
        seq = [f.download() for f in dfs]
        for graph in (0, 1):
            sec, reps = _timed(lambda: solve_plan.solve_many(ctx, plans, [per] * 8, [d.ptr for d in drs],
                                                             [f.ptr for f in dfs], precision=prec, tol=1e-12,
                                                             graph=graph), a.steps)
            diff = max(np.abs(f.download() - s_).sum(axis=1).max() for f, s_ in zip(dfs, seq))
            res[f"{'fp32+fp64' if prec else 'fp64'}_graph{graph}"] = {
                "sets_per_s": 8 * per / sec, "ms": 1000 * sec, "sequential_sets_per_s": 8 * per / seq_s,
                "max_l1_vs_sequential": float(diff), "max_residual": max(r["residual"] for r in reps)}
    best = max(v["sets_per_s"] for v in res.values())
    print(json.dumps({
        "metric": "rate sets/s, 8 uneven networks x 128 sets in one call", "value": best, "unit": "sets/s",
        "workload": f"N=12..18,18 K=2 generator v1 seeds 1000*{a.seed}+k; 128 rate sets each; tol 1e-12",
        "networks": [{"N": t.net.N, **{k: t.info[k] for k in ("unknowns", "ambiguous", "terminals_reached")}}
                     for t in tops], "variants": res}))


def workload_seeds(a, ctx):
    """Representative instances: N=20 seeds 0..7 (the metric's network family),
    N=22 seeds 0..3 (the mixed-precision family) and N=16 seeds 0..3 (the batch
    family).  Per seed: the plan's unknowns, the schedule, batched FP32 + FP64
    refinement sets/s (1024 sets when the system fits the per-set kernels,
    autotuned; device-wide otherwise, 256 sets at N=16 and 64 above), the plain
    per-solve FP64 rate on the same plan, the speedup, and the L1 distance of
    batched set 0 to its plain FP64 solve."""
    from paraexastolog_amd import synth
    from paraexastolog_amd._lib import check, lib
    from paraexastolog_amd.surrogate import DeviceArray

    L = lib()
    recs = []
    for cfg, seeds in (("c1", range(8)), ("c3", range(4)), ("c2", range(4))):
        for seed in seeds:
            net, rng = synth.config(cfg, seed)
            tp = _Topo(ctx, net, synth.rate_sets(rng, net.N, 1)[0])
            nU = tp.info["unknowns"]
            B = 1024 if nU <= 4096 else (256 if net.N <= 16 else 64)
            net2, rng2 = synth.config(cfg, seed)
            rates = synth.rate_sets(rng2, net2.N, B)
            dr = DeviceArray.from_host(ctx, rates)
            df = DeviceArray(ctx, (B, tp.n), np.float64)
            d1 = DeviceArray(ctx, (1, tp.n), np.float64)
            if nU <= 4096:
                path, restart, _ = tp.plan.autotune(B, dr.ptr, df.ptr, precision=1, tol=1e-12, graph=1)
                sched = f"{SCHEDULE_NAMES[path]}, restart {restart}"
                run = lambda: tp.plan.solve(B, dr.ptr, df.ptr, precision=1, tol=1e-12, graph=1, restart=0)  # noqa: E731
                t_batch, rep = _timed(run, 1)
            else:
                # device-wide: the faster of restart 32 and 64 (neither wins everywhere)
                best = None
                for m in (32, 64):
                    t_m, rep_m = _timed(lambda: tp.plan.solve(B, dr.ptr, df.ptr, precision=1, tol=1e-12,  # noqa: B023
                                                              restart=m), 1)
                    if best is None or t_m < best[0]:
                        best = (t_m, rep_m, m)
                t_batch, rep, m = best
                sched = f"device-wide, restart {m} (best of 32, 64)"
            t_plain, _ = _timed(lambda: tp.plan.solve(1, dr.ptr, d1.ptr, precision=0, tol=1e-13), 1)
            xb0 = np.empty(tp.n)  # batched set 0 only
            check(L.esl_memcpy_d2h(ctx.handle, xb0.ctypes.data_as(C.c_void_p), df.ptr, 8 * tp.n), ctx.handle)
            l1 = float(np.abs(d1.download()[0] - xb0).sum())
            rec = {"N": net.N, "seed": seed, "unknowns": nU, "terminals_reached": tp.info["terminals_reached"],
                   "ambiguous": tp.info["ambiguous"], "sets": B, "schedule": sched,
                   "sets_per_s": B / t_batch, "plain_fp64_per_s": 1.0 / t_plain,
                   "speedup_vs_plain_fp64": (B / t_batch) * t_plain, "l1_set0_vs_plain_fp64": l1,
                   "mean_iterations": rep["iterations_sum"] / B, "exact_solves": rep["exact_solves"]}
            recs.append(rec)
            print(json.dumps(rec), file=sys.stderr, flush=True)
            del tp, dr, df, d1
    out = {"metric": "rate sets/s per seed (batched FP32 + FP64 refinement) and speedup vs plain per-solve FP64",
           "workload": "N=20 (K in {2,3}, 5 fixed nodes) seeds 0..7, N=22 (K=3 + 3 spliced 3-cycles) seeds 0..3 "
                       "and N=16 (K=2) seeds 0..3, generator v1; tol 1e-12 batched, 1e-13 plain",
           "seeds": recs}
    for N in (20, 22, 16):
        r = [x for x in recs if x["N"] == N]
        out[f"median_N{N}"] = {"sets_per_s": float(np.median([x["sets_per_s"] for x in r])),
                               "speedup_vs_plain_fp64": float(np.median([x["speedup_vs_plain_fp64"] for x in r])),
                               "unknowns": float(np.median([x["unknowns"] for x in r]))}
    print(json.dumps(out))


def workload_n22(a, ctx):
    """N=22, K=3 with three spliced negative-feedback 3-cycles (generator v1
    seed 8: one terminal SCC of 221,184 states), 64 rate sets: FP32 Krylov +
    FP64 refinement against direct FP64, errors of every result."""
    from paraexastolog_amd import synth
    from paraexastolog_amd.surrogate import DeviceArray

    seed = 8 if a.seed is None else a.seed
    net, rng = synth.config("c3", seed)
    B = min(a.sets, 64)
    rates = synth.rate_sets(rng, net.N, B)
    tp = _Topo(ctx, net, rates[0])
    dr = DeviceArray.from_host(ctx, rates)
    d64 = DeviceArray(ctx, (B, tp.n), np.float64)
    d32 = DeviceArray(ctx, (B, tp.n), np.float64)
    # 64 sets of 2e5-2e6 unknowns: the device-wide schedule (every vector operation over all sets at
    # once), restart 64 (measured faster than 32 on these systems: N=22 seed 8 377 vs 203 sets/s)
    tp.plan.set_schedule(2, 64)
    steps = max(1, a.steps // 3)
    t64, r64 = _timed(lambda: tp.plan.solve(B, dr.ptr, d64.ptr, precision=0, tol=1e-12, restart=0), steps)
    print(f"n22: fp64 {t64:.3f} s per batch", file=sys.stderr, flush=True)
    t32, r32 = _timed(lambda: tp.plan.solve(B, dr.ptr, d32.ptr, precision=1, tol=1e-12, restart=0), steps)
    x64 = d64.download()
    x32 = d32.download()
    l1 = np.abs(x32 - x64).sum(axis=1)
    nz = x64 != 0
    rel = np.where(nz, np.abs(x32 - x64) / np.where(nz, np.abs(x64), 1.0), 0.0).max(axis=1)
    print(json.dumps({
        "metric": "rate sets/s, N=22 mixed precision (FP32 Krylov + FP64 refinement vs FP64)",
        "value": B / t32, "unit": "sets/s",
        "workload": f"N=22 K=3 + 3 spliced negative-feedback 3-cycles, generator v1 seed {seed}, 5 fixed nodes; "
                    f"{B} rate sets; tol 1e-12",
        "plan": tp.info, "terminal_scc_states": tp.n_t,
        "fp64": {"sets_per_s": B / t64, "s_per_batch": t64, "mean_iterations": r64["iterations_sum"] / B},
        "fp32_refined": {"sets_per_s": B / t32, "s_per_batch": t32, "mean_iterations": r32["iterations_sum"] / B,
                         "refinement_cycles": r32["refinements"], "fp64_fallback_sets": r32["fp64_fallback"]},
        "speedup": t64 / t32,
        "errors_fp32_vs_fp64": {"l1_max": float(l1.max()), "l1_median": float(np.median(l1)),
                                "max_rel_max": float(rel.max()), "max_rel_median": float(np.median(rel))},
        "sums": [float(np.abs(x64.sum(axis=1) - 1).max()), float(np.abs(x32.sum(axis=1) - 1).max())]}))


def surrogate_gemm_rate(L, ctx, sur, X, T, steps=40, batch=512):
    """Device time of the surrogate's BF16 GEMM kernels over `steps` training
    steps (one esl_mlp_train_steps call at learning rate 0, so the trained
    network is left as it is) on minibatches of 512 rows, timed per launch with
    HIP events: the algorithmic 2 M N K flops of every product over the summed
    kernel time, against the 2.5 PFLOP/s dense BF16 roof.  The products are
    512 x 256 x 32 up to 512 x 2048 x 256 (and their transposes for the
    gradients); event timing adds a few microseconds to each launch, so the
    kernel-trace profile gives the tighter figure."""
    from paraexastolog_amd._lib import check
    from paraexastolog_amd.surrogate import DeviceArray

    X = np.ascontiguousarray(X[:steps * batch], dtype=np.float32)
    T = np.ascontiguousarray(T[:steps * batch], dtype=np.float32)
    dX, dT = DeviceArray.from_host(ctx, X), DeviceArray.from_host(ctx, T)
    check(L.esl_profile(ctx.handle, 1), ctx.handle)
    t = time.perf_counter()
    sur.train_steps_device(steps, batch, dX.ptr, dT.ptr, 0.0, 0.0)
    wall = time.perf_counter() - t
    stats = profile_stats(L, ctx)
    check(L.esl_profile(ctx.handle, 0), ctx.handle)
    g = [x for x in stats if x[0] == "k_gemm_bf16"]
    r = [x for x in stats if x[0] == "k_gemm_reduce"]
    launches = sum(x[1] for x in g)
    ms = sum(x[2] for x in g)
    flops = sum(x[3] for x in g)
    tf = flops / (ms * 1e-3) / 1e12 if ms > 0 else None
    return {"launches": launches, "ms": ms, "tflops": tf, "peak_tflops": 2500.0,
            "frac": tf / 2500.0 if tf else None, "reduce_ms": sum(x[2] for x in r),
            "step_ms_profiled": 1e3 * wall / steps,
            "all_kernels_ms_per_step": sum(x[2] for x in stats) / steps}


def workload_warm(a, ctx):
    """The learned warm start on the N=16 topology (seed 11, 2048 terminal
    states): 50k pipeline-generated training solves, the MLP trained on the GPU,
    4096 held-out rate sets solved cold and warm (prediction included in the
    warm timing)."""
    from paraexastolog_amd import synth
    from paraexastolog_amd._lib import lib
    from paraexastolog_amd.surrogate import DeviceArray, Surrogate, log_rates, terminal_targets, train

    seed = 11 if a.seed is None else a.seed
    net, rng = synth.config("c2", seed)
    n_train, n_test = 50000, 4096
    rates = synth.rate_sets(rng, net.N, n_train + n_test)
    tp = _Topo(ctx, net, rates[0])
    t = time.perf_counter()
    T = terminal_targets(ctx, tp.plan, rates, tp.n, tp.n_t, batch=2048, precision=1, tol=1e-12, graph=1)
    gen_s = time.perf_counter() - t
    X = log_rates(rates)
    sur = Surrogate(ctx, 2 * net.N, tp.n_t, seed=3)
    t = time.perf_counter()
    hist = train(sur, X[:n_train], T[:n_train], epochs=a.epochs, batch=512, lr=1e-2)
    train_s = time.perf_counter() - t
    gemm = surrogate_gemm_rate(lib(), ctx, sur, X[:n_train], T[:n_train])
    Tt = T[n_train:]
    ent = float(-(np.where(Tt > 0, Tt * np.log(np.maximum(Tt, 1e-300)), 0.0)).sum(axis=1).mean())
    dX = DeviceArray.from_host(ctx, X[n_train:])
    dW = DeviceArray(ctx, (n_test, tp.n_t), np.float64)
    dr = DeviceArray.from_host(ctx, rates[n_train:])
    dc = DeviceArray(ctx, (n_test, tp.n), np.float64)
    dw = DeviceArray(ctx, (n_test, tp.n), np.float64)
    cold_s, cold = _timed(lambda: tp.plan.solve(n_test, dr.ptr, dc.ptr, precision=1, tol=1e-12, graph=1), a.steps)

    def warm_step():
        sur.predict_warm_device(n_test, dX.ptr, dW.ptr)
        return tp.plan.solve(n_test, dr.ptr, dw.ptr, precision=1, tol=1e-12, graph=1, warm=dW.ptr)

    warm_s, warm = _timed(warm_step, a.steps)
    W = dW.download()
    err = np.abs(W - Tt).sum(axis=1)
    base = np.abs(T[:n_train].mean(axis=0) - Tt).sum(axis=1)
    diff = float(np.abs(dc.download() - dw.download()).sum(axis=1).max())
    n_par = sur.n_params
    print(json.dumps({
        "metric": "rate sets/s with vs without the learned warm start (N=16, held-out sets)",
        "value": n_test / warm_s, "unit": "sets/s",
        "workload": f"N=16 K=2 generator v1 seed {seed}; MLP 32-256-256-256-{tp.n_t} (GELU, softmax), BF16 MFMA / "
                    f"FP32 accumulate, Adam; {n_train} training solves, {n_test} held out; tol 1e-12",
        "data_generation": {"solves": n_train + n_test, "s": gen_s, "solves_per_s": (n_train + n_test) / gen_s},
        "training": {"epochs": a.epochs, "steps": a.epochs * (n_train // 512), "s": train_s,
                     "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1], "held_out_target_entropy": ent,
                     "tflops": 6.0 * 512 * n_par * a.epochs * (n_train // 512) / train_s / 1e12,
                     "gemm_kernel": gemm},
        "surrogate_error_l1": {"median": float(np.median(err)), "p90": float(np.percentile(err, 90)),
                               "max": float(err.max()), "constant_predictor_median": float(np.median(base))},
        "cold": {"sets_per_s": n_test / cold_s, "mean_iterations": cold["iterations_sum"] / n_test},
        "warm": {"sets_per_s": n_test / warm_s, "mean_iterations": warm["iterations_sum"] / n_test,
                 "sets_refined": warm["sets_iterated"], "fp64_fallback_sets": warm["fp64_fallback"]},
        "gain": cold_s / warm_s, "max_l1_warm_vs_cold": diff}))


WORKLOADS = {"n16": workload_n16, "uneven": workload_uneven, "n22": workload_n22, "warm": workload_warm,
             "seeds": workload_seeds}


if __name__ == "__main__":
    main()
