"""Plans, batches, HIP-graph replays and warm starts through the C ABI."""
import numpy as np
import pytest

from oracle import pipeline as P
from oracle.bnet import read_bnet_text
from paraexastolog_amd import model_builder, solve_plan, synth, transition_table, transition_graph, initial_state
from paraexastolog_amd import solver
from paraexastolog_amd.surrogate import DeviceArray

pytestmark = pytest.mark.gpu


def setup_net(ctx, cfg, seed, sets):
    net, rng = synth.config(cfg, seed)
    rates = synth.rate_sets(rng, net.N, sets)
    m = model_builder().parse_model(net.bnet())
    t = transition_table(ctx, m, rates[0])
    t.construct_table()
    g = transition_graph(ctx, t)
    g.reorganize_vertices()
    fixed, vals = net.fixed_spec()
    x0 = initial_state(m.nodes, fixed, vals).state
    return net, rates, m, t, g, x0


@pytest.mark.parametrize("cfg,seed", [("c0", 1), ("c2", 37), ("c2", 0)])
@pytest.mark.parametrize("precision", [0, 1])
def test_plan_batch_matches_single_solves(ctx, cfg, seed, precision):
    B = 12
    net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=precision, tol=1e-12)
    xb = dout.download()
    for k in (0, 5, B - 1):
        s = solver(ctx, t, g, rates[k], x0)
        s.solve()
        assert np.abs(xb[k] - s.final_state).sum() <= 1e-10
        # mass is conserved up to the residual target (1e-12 relative) times the conditioning
        assert abs(xb[k].sum() - 1) < 1e-11
    assert rep["residual"] <= 1e-12


def test_large_batches(ctx):
    # 2,500 sets on the per-set kernels (more workgroups than slots) and 1,100 on
    # the device-wide solver (two chunks of at most 1,024 sets): the chunk and
    # set offsets agree with single solves and with each other
    B = 2500
    net, rates, m, t, g, x0 = setup_net(ctx, "c0", 1, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    assert plan.info()["unknowns"] > 0
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12, graph=1)
    xb = dout.download()
    assert rep["residual"] <= 1e-12
    assert np.abs(xb.sum(axis=1) - 1).max() < 1e-11
    for k in (0, 1023, 1024, 2047, B - 1):
        s = solver(ctx, t, g, rates[k], x0)
        s.solve()
        assert np.abs(xb[k] - s.final_state).sum() <= 1e-10
    # more sets than the context's pinned staging buffer holds (8,192): the
    # per-set outputs come back array by array
    Bl = 9000
    rl = np.concatenate([rates, rates, rates, rates])[:Bl]
    drl = DeviceArray.from_host(ctx, rl)
    dl = DeviceArray(ctx, (Bl, n), np.float64)
    rep = plan.solve(Bl, drl.ptr, dl.ptr, precision=1, tol=1e-12)
    assert rep["residual"] <= 1e-12
    res, its, meth = plan.set_results(Bl)
    assert len(res) == Bl and res.max() <= 1e-12 and (its > 0).all() and (meth >= 0).all()
    xl = dl.download()
    for k in (0, 2499, 2500, 8191, 8192, Bl - 1):
        assert np.abs(xl[k] - xb[k % B]).sum() <= 1e-10
    del dl, drl, xl
    Bw = 1100
    dw = DeviceArray(ctx, (Bw, n), np.float64)
    plan.set_schedule(2, 32)
    rep = plan.solve(Bw, dr.ptr, dw.ptr, precision=1, tol=1e-12, restart=0)
    assert rep["residual"] <= 1e-12
    assert np.abs(dw.download() - xb[:Bw]).sum(axis=1).max() <= 1e-10


def test_plan_matches_oracle_c0(ctx):
    B = 4
    net, rates, m, t, g, x0 = setup_net(ctx, "c0", 3, B)
    n = 1 << net.N
    om = read_bnet_text(net.bnet())
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    plan.solve(B, dr.ptr, dout.ptr, precision=0, tol=1e-13)
    xb = dout.download()
    ip, rows, _ = P.transition_table(om, rates[0])
    gr = P.transition_graph(n, ip, rows)
    for k in range(B):
        ref = P.solve(n, ip, rows, rates[k], x0, gr, full=True).final_state
        assert np.abs(xb[k] - ref).sum() <= 1e-12


def test_graph_replays_match_eager(ctx):
    # a captured numeric sequence replayed several times, with eager work and
    # synchronisations in between; every replay must equal the eager result
    B = 16
    net, rates, m, t, g, x0 = setup_net(ctx, "c2", 37, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    d_eager = DeviceArray(ctx, (B, n), np.float64)
    d_graph = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    plan.solve(B, dr.ptr, d_eager.ptr, precision=1, tol=1e-12)
    eager = d_eager.download()
    for rep in range(4):
        plan.solve(B, dr.ptr, d_graph.ptr, precision=1, tol=1e-12, graph=1)
        got = d_graph.download()
        assert np.abs(got - eager).max() <= 1e-13, rep
        # eager work between replays
        plan.solve(2, dr.ptr, d_eager.ptr, precision=0, tol=1e-13)
    # new rates in the same buffer: the replay must pick them up
    rates2 = rates[::-1].copy()
    dr.upload(rates2)
    plan.solve(B, dr.ptr, d_graph.ptr, precision=1, tol=1e-12, graph=1)
    got = d_graph.download()
    plan.solve(B, dr.ptr, d_eager.ptr, precision=1, tol=1e-12)
    assert np.abs(got - d_eager.download()).max() <= 1e-13


def test_warm_start_from_exact_and_perturbed(ctx):
    B = 6
    net, rates, m, t, g, x0 = setup_net(ctx, "c2", 37, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    cold = plan.solve(B, dr.ptr, dout.ptr, precision=0, tol=1e-13)
    exact = dout.download()
    n_t = int(g.sccs_offsets[g.terminals_count])
    term = g.reordered_vertices[:n_t]
    warm = np.ascontiguousarray(exact[:, term])
    dwarm = DeviceArray.from_host(ctx, warm)
    hot = plan.solve(B, dr.ptr, dout.ptr, precision=0, tol=1e-13, warm=dwarm.ptr)
    assert np.abs(dout.download() - exact).sum(axis=1).max() <= 1e-12
    # an exact prediction of the terminal part cannot need more Krylov iterations
    # than a cold start (the absorption part starts cold in both)
    assert hot["iterations"] <= cold["iterations"] + 2
    noisy = warm * np.exp(np.random.default_rng(0).normal(0, 0.3, warm.shape))
    dwarm.upload(noisy)
    plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12, warm=dwarm.ptr)
    assert np.abs(dout.download() - exact).sum(axis=1).max() <= 1e-10


@pytest.mark.parametrize("pre", [1, 2])
def test_device_wide_solver_matches_per_set(ctx, pre):
    # one rate set with 33k unknowns runs device-wide (kernels per ILU level);
    # 16 sets run one workgroup per set: same answers
    B = 16
    net, rates, m, t, g, x0 = setup_net(ctx, "c2", 37, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    plan = solve_plan(ctx, t, dx0.ptr)
    assert plan.info()["unknowns"] > 16384
    d16 = DeviceArray(ctx, (B, n), np.float64)
    plan.set_schedule(1, 32)  # one workgroup per set, vectors in global memory
    plan.solve(B, dr.ptr, d16.ptr, precision=0, tol=1e-13, preconditioner=pre, restart=0)
    d1 = DeviceArray(ctx, (1, n), np.float64)
    plan.set_schedule(2, 32)  # device-wide
    rep = plan.solve(1, dr.offset(2 * net.N * 3), d1.ptr, precision=0, tol=1e-13, preconditioner=pre, restart=0)
    assert np.abs(d1.download()[0] - d16.download()[3]).sum() <= 1e-11
    if pre == 2:
        assert rep["iterations"] < 500


@pytest.mark.parametrize("cfg,seed", [("c2", 11), ("c0", 2)])
@pytest.mark.parametrize("precision", [0, 1])
def test_lds_resident_solver_matches_oracle(ctx, cfg, seed, precision):
    # c2 seed 11: one terminal SCC of 2048 states (8 unknowns per thread);
    # c0 seed 2: small systems (4 per thread); both run the LDS-resident kernel
    B = 3
    net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
    n = 1 << net.N
    om = read_bnet_text(net.bnet())
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=precision, tol=1e-13)
    xb = dout.download()
    for k in range(B):
        ip, rows, _ = P.transition_table(om, rates[k])
        gr = P.transition_graph(n, ip, rows)
        ref = P.solve_final(n, ip, rows, rates[k], x0, gr)
        assert np.abs(xb[k] - ref).sum() <= 1e-10, (k, np.abs(xb[k] - ref).sum())
    assert rep["residual"] <= 1e-13


def test_device_copy(ctx):
    from paraexastolog_amd._lib import check, lib
    import ctypes as C

    rng = np.random.default_rng(5)
    for nbytes, off in ((1 << 20, 0), (1000003, 0), (4096 + 7, 3)):
        src = rng.integers(0, 256, nbytes + off, dtype=np.uint8)
        ds = DeviceArray.from_host(ctx, src)
        dd = DeviceArray(ctx, (nbytes + off,), np.uint8)
        check(lib().esl_memcpy_d2d(ctx.handle, C.c_void_p(dd.ptr.value + off), C.c_void_p(ds.ptr.value + off),
                                   nbytes), ctx.handle)
        got = dd.download()
        np.testing.assert_array_equal(got[off:], src[off:])


@pytest.mark.parametrize("graph", [0, 1])
def test_solve_many_matches_single_plans(ctx, graph):
    # networks of uneven size solved concurrently == each plan on its own
    specs = [("c0", 1, 5), ("c2", 11, 3), ("c0", 4, 7), ("c2", 37, 2)]
    plans, nsets, drs, dfs, singles, keep = [], [], [], [], [], []
    for cfg, seed, B in specs:
        net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
        n = 1 << net.N
        dx0 = DeviceArray.from_host(ctx, x0)
        dr = DeviceArray.from_host(ctx, rates)
        df = DeviceArray(ctx, (B, n), np.float64)
        ds = DeviceArray(ctx, (B, n), np.float64)
        p = solve_plan(ctx, t, dx0.ptr)
        p.solve(B, dr.ptr, ds.ptr, precision=1, tol=1e-12)
        plans.append(p)
        nsets.append(B)
        drs.append(dr.ptr)
        dfs.append(df.ptr)
        singles.append(ds)
        keep += [t, g, dx0, dr, df]
    for rep_i in range(3 if graph else 1):
        reps = solve_plan.solve_many(ctx, plans, nsets, drs, dfs, precision=1, tol=1e-12, graph=graph)
        for k in range(len(specs)):
            got = keep[5 * k + 4].download()
            ref = singles[k].download()
            assert np.abs(got - ref).sum(axis=1).max() <= 1e-10, (rep_i, k)
            assert reps[k]["residual"] <= 1e-12


def test_solve_many_error_on_a_device_wide_plan(ctx):
    # c2 seed 37 (33,280 unknowns) runs device-wide on a host thread of its own;
    # one Krylov iteration without the exact fallback misses the tolerance there,
    # the error reaches the caller after every thread has joined, and the context
    # then solves the same plans normally
    from paraexastolog_amd._lib import EslError

    specs = [("c0", 1, 5), ("c2", 37, 2)]
    plans, nsets, drs, dfs, singles, keep = [], [], [], [], [], []
    for cfg, seed, B in specs:
        net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
        n = 1 << net.N
        dx0 = DeviceArray.from_host(ctx, x0)
        dr = DeviceArray.from_host(ctx, rates)
        df = DeviceArray(ctx, (B, n), np.float64)
        ds = DeviceArray(ctx, (B, n), np.float64)
        p = solve_plan(ctx, t, dx0.ptr)
        p.solve(B, dr.ptr, ds.ptr, precision=1, tol=1e-12)
        plans.append(p)
        nsets.append(B)
        drs.append(dr.ptr)
        dfs.append(df.ptr)
        singles.append(ds)
        keep += [t, g, dx0, dr, df]
    assert plans[1].info()["unknowns"] > 4096
    with pytest.raises(EslError) as e:
        solve_plan.solve_many(ctx, plans, nsets, drs, dfs, precision=1, tol=1e-12, max_iters=1, exact_fallback=0)
    assert "ESL_E_NUMERIC" in str(e.value)
    reps = solve_plan.solve_many(ctx, plans, nsets, drs, dfs, precision=1, tol=1e-12)
    for k in range(len(specs)):
        assert np.abs(keep[5 * k + 4].download() - singles[k].download()).sum(axis=1).max() <= 1e-10
        assert reps[k]["residual"] <= 1e-12


@pytest.mark.parametrize("cfg,seed,B", [("c0", 1, 8), ("c2", 11, 4)])
def test_autotune_choice_keeps_results(ctx, cfg, seed, B):
    net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    d_ref = DeviceArray(ctx, (B, n), np.float64)
    d_out = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    plan.solve(B, dr.ptr, d_ref.ptr, precision=0, tol=1e-13)
    assert plan.info()["tuned"] == -1
    path, restart, ms = plan.autotune(B, dr.ptr, d_out.ptr, precision=1, tol=1e-12, graph=1)
    assert path in (0, 1, 2, 3, 4, 5) and restart in (16, 24, 32, 64) and ms > 0
    assert plan.info()["tuned"] == 1000 * path + restart
    for graph in (0, 1):
        rep = plan.solve(B, dr.ptr, d_out.ptr, precision=1, tol=1e-12, restart=0, graph=graph)
        assert rep["residual"] <= 1e-12
        assert np.abs(d_out.download() - d_ref.download()).sum(axis=1).max() <= 1e-10


_N20_ORACLE = {}


def _n20_oracle(net, rates, k, x0):
    """FP64 oracle result of rate set k of the N=20 seed 0 network (cached)."""
    if k not in _N20_ORACLE:
        om = read_bnet_text(net.bnet())
        n = 1 << net.N
        ip, rows, _ = P.transition_table(om, rates[k])
        gr = P.transition_graph(n, ip, rows)
        _N20_ORACLE[k] = P.solve_final(n, ip, rows, rates[k], x0, gr)
    return _N20_ORACLE[k]


@pytest.mark.parametrize("restart", [16, 24, 32])
def test_headline_kernel_each_restart(ctx, restart):
    # the bench's schedule: per-set LDS-resident kernel with the FP32 basis in
    # registers (k_gmres_priv<restart>), N=20 seed 0 (1022 unknowns), FP32 +
    # FP64 refinement at tol 1e-12, replayed as a HIP graph; every set against
    # one batched FP64 run, two sets against the oracle
    from paraexastolog_amd._lib import check, lib

    B = 64
    net, rates, m, t, g, x0 = setup_net(ctx, "c1", 0, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    d64 = DeviceArray(ctx, (B, n), np.float64)
    d32 = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    assert plan.info()["unknowns"] == 1022
    plan.solve(B, dr.ptr, d64.ptr, precision=0, tol=1e-13, restart=32)
    plan.set_schedule(0, restart)
    check(lib().esl_profile(ctx.handle, 1), ctx.handle)
    rep = plan.solve(B, dr.ptr, d32.ptr, precision=1, tol=1e-12, restart=0, graph=1)
    import ctypes as C
    name = C.create_string_buffer(64)
    names = []
    i = 0
    while lib().esl_profile_stat(ctx.handle, i, name, 64, None, None, None) == 0:
        names.append(name.value.decode())
        i += 1
    check(lib().esl_profile(ctx.handle, 0), ctx.handle)
    assert "k_gmres_priv<float>" in names, names
    assert rep["residual"] <= 1e-12 and rep["fp64_fallback"] == 0
    x32, x64 = d32.download(), d64.download()
    l1 = np.abs(x32 - x64).sum(axis=1)
    assert l1.max() <= 1e-10, (l1.argmax(), l1.max())
    assert np.abs(x32.sum(axis=1) - 1).max() <= 1e-11
    for k in (0, 1):
        ref = _n20_oracle(net, rates, k, x0)
        assert np.abs(x32[k] - ref).sum() <= 1e-10
        assert np.abs(x64[k] - ref).sum() <= 1e-12


def test_uneven_batch_matches_oracle(ctx):
    # the bench's uneven batch: 8 networks N = 12..18, 18 (K = 2), 128 rate sets
    # each, in one esl_plan_solve_many call (FP32 + FP64 refinement, one HIP
    # graph); three sets of every network against the FP64 oracle, every set
    # against an FP64 call, mass, non-negativity and global balance K x = 0
    per = 128
    specs = []
    for net, rng in synth.uneven(0):
        rates = synth.rate_sets(rng, net.N, per)
        m = model_builder().parse_model(net.bnet())
        t = transition_table(ctx, m, rates[0])
        t.construct_table()
        g = transition_graph(ctx, t)
        g.reorganize_vertices()
        fixed, vals = net.fixed_spec()
        x0 = initial_state(m.nodes, fixed, vals).state
        dx0 = DeviceArray.from_host(ctx, x0)
        plan = solve_plan(ctx, t, dx0.ptr)
        n = 1 << net.N
        specs.append(dict(net=net, rates=rates, t=t, g=g, x0=x0, dx0=dx0, plan=plan,
                          dr=DeviceArray.from_host(ctx, rates), d32=DeviceArray(ctx, (per, n), np.float64),
                          d64=DeviceArray(ctx, (per, n), np.float64)))
    print("uneven networks:", [(s["net"].N, s["plan"].info()["unknowns"]) for s in specs], flush=True)
    plans = [s["plan"] for s in specs]
    reps32 = solve_plan.solve_many(ctx, plans, [per] * 8, [s["dr"].ptr for s in specs],
                                   [s["d32"].ptr for s in specs], precision=1, tol=1e-12, graph=1)
    reps64 = solve_plan.solve_many(ctx, plans, [per] * 8, [s["dr"].ptr for s in specs],
                                   [s["d64"].ptr for s in specs], precision=0, tol=1e-13)
    for s, r32, r64 in zip(specs, reps32, reps64):
        net = s["net"]
        n = 1 << net.N
        x32, x64 = s["d32"].download(), s["d64"].download()
        assert r32["residual"] <= 1e-12 and r64["residual"] <= 1e-13, (net.N, r32, r64)
        assert np.abs(x32 - x64).sum(axis=1).max() <= 1e-10, net.N
        assert np.abs(x32.sum(axis=1) - 1).max() <= 1e-11, net.N
        assert x32.min() >= -1e-14, net.N
        om = read_bnet_text(net.bnet())
        for k in (0, 1, per - 1):
            ip, rows, _ = P.transition_table(om, s["rates"][k])
            gr = P.transition_graph(n, ip, rows)
            ref = P.solve_final(n, ip, rows, s["rates"][k], s["x0"], gr)
            assert np.abs(x32[k] - ref).sum() <= 1e-10, (net.N, k, np.abs(x32[k] - ref).sum())
            assert np.abs(x64[k] - ref).sum() <= 1e-12, (net.N, k, np.abs(x64[k] - ref).sum())
            K = P.kinetic(n, ip, rows, s["rates"][k])
            assert np.abs(K @ x32[k]).sum() <= 1e-7, (net.N, k)  # ||K||_1 <= 2 * 10 * N times the L1 bar


def _oracle_final(net, rates, k, x0):
    om = read_bnet_text(net.bnet())
    n = 1 << net.N
    ip, rows, _ = P.transition_table(om, rates[k])
    gr = P.transition_graph(n, ip, rows)
    return P.solve_final(n, ip, rows, rates[k], x0, gr)


@pytest.mark.parametrize("cfg,seed,B", [("c0", 10, 5), ("c1", 0, 3), ("c2", 11, 2)])
def test_exact_lu_matches_oracle(ctx, cfg, seed, B):
    # exact_fallback 2: every set solved by the dense LU (+ refinement) of each
    # diagonal block -- c0 seed 10 has a terminal segment and ambiguous states,
    # c1 seed 0 one terminal SCC of 1022 states, c2 seed 11 one of 2048
    net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12, exact_fallback=2)
    assert rep["exact_solves"] == B
    res, its, meth = plan.set_results(B)
    assert (meth == 2).all() and res.max() <= 1e-14, (meth, res)
    xb = dout.download()
    for k in range(B):
        ref = _oracle_final(net, rates, k, x0)
        assert np.abs(xb[k] - ref).sum() <= 1e-12, (k, np.abs(xb[k] - ref).sum())


@pytest.mark.parametrize("schedule", [None, 2])
def test_missed_tolerance_falls_back_to_exact_lu(ctx, schedule):
    # one Krylov iteration cannot reach 1e-12: every set goes to the exact LU
    # (per-set kernels, and the device-wide solver when pinned); without the
    # fallback the call fails with ESL_E_NUMERIC and every set is marked
    from paraexastolog_amd import EslError

    B = 4
    net, rates, m, t, g, x0 = setup_net(ctx, "c0", 10, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    restart = 32
    if schedule is not None:
        plan.set_schedule(schedule, 32)
        restart = 0
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12, max_iters=1, restart=restart)
    assert rep["exact_solves"] == B and rep["residual"] <= 1e-12
    res, its, meth = plan.set_results(B)
    assert (meth == 2).all()
    xb = dout.download()
    for k in range(B):
        assert np.abs(xb[k] - _oracle_final(net, rates, k, x0)).sum() <= 1e-12
    with pytest.raises(EslError) as e:
        plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12, max_iters=1, restart=restart, exact_fallback=0)
    assert "NUMERIC" in str(e.value) and "exact fallback off" in str(e.value)
    res, its, meth = plan.set_results(B)
    assert (meth == -1).all() and (res > 1e-12).all()


def test_exact_fallback_refuses_blocks_beyond_its_limit(ctx):
    # c2 seed 37: 32,768 ambiguous states in one block, beyond the exact
    # solver's 16384: a forced Krylov failure is reported, not rescued
    from paraexastolog_amd import EslError

    B = 2
    net, rates, m, t, g, x0 = setup_net(ctx, "c2", 37, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    assert plan.info()["ambiguous"] > 16384
    with pytest.raises(EslError) as e:
        plan.solve(B, dr.ptr, dout.ptr, precision=0, tol=1e-13, max_iters=1)
    assert "exceeds" in str(e.value)
    res, its, meth = plan.set_results(B)
    assert (meth == -1).all()


def latch_bnet(n_latch):
    """n_latch memory nodes m_i = m_i | t and a trigger t = false: while t is on,
    any memory node may switch on, and t switches off for good.  The 2^n_latch
    states with t off are fixed points (terminal SCCs); a state with t on and
    memory pattern M reaches every fixed point whose pattern contains M."""
    lines = ["targets, factors"]
    for i in range(n_latch):
        lines.append(f"m{i}, m{i} | t")
    lines.append("t, t & !t")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("n_latch", [5, 6, 7, 9])
@pytest.mark.parametrize("precision", [0, 1])
def test_many_terminal_sccs_match_oracle(ctx, n_latch, precision):
    # 32, 64 (every bit of the terminal masks), 128 and 512 reachable terminal
    # SCCs; beyond 64 every reachable transient state enters the absorption system
    text = latch_bnet(n_latch)
    m = model_builder().parse_model(text)
    N = n_latch + 1
    n = 1 << N
    B = 3
    rng = np.random.default_rng(n_latch)
    rates = np.exp(rng.uniform(np.log(0.1), np.log(10), (B, 2 * N)))
    t = transition_table(ctx, m, rates[0])
    t.construct_table()
    g = transition_graph(ctx, t)
    g.reorganize_vertices()
    assert g.terminals_count == 1 << n_latch
    x0 = initial_state(m.nodes).state
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    assert plan.info()["terminals_reached"] == 1 << n_latch
    rep = plan.solve(B, dr.ptr, dout.ptr, precision=precision, tol=1e-13 if precision == 0 else 1e-12)
    assert rep["exact_solves"] == 0
    xb = dout.download()
    om = read_bnet_text(text)
    for k in range(B):
        ip, rows, _ = P.transition_table(om, rates[k])
        gr = P.transition_graph(n, ip, rows)
        ref = P.solve(n, ip, rows, rates[k], x0, gr, full=False).final_state
        bar = 1e-12 if precision == 0 else 1e-10
        assert np.abs(xb[k] - ref).sum() <= bar, (k, np.abs(xb[k] - ref).sum())


@pytest.mark.parametrize("case", ["c0-10", "c2-37", "latch-7", "uneven-15"])
def test_results_are_bitwise_reproducible(ctx, case):
    # two plans built from scratch, each solved twice (once replayed as a HIP
    # graph): every final vector bit for bit the same -- no floating-point
    # accumulation depends on the order in which threads or blocks run
    B = 6
    if case.startswith("latch"):
        text = latch_bnet(int(case.split("-")[1]))
        m = model_builder().parse_model(text)
        N = len(m.nodes)
        rates = np.exp(np.random.default_rng(1).uniform(np.log(0.1), np.log(10), (B, 2 * N)))
        x0 = initial_state(m.nodes).state
        t = transition_table(ctx, m, rates[0])
        t.construct_table()
        g = transition_graph(ctx, t)
        g.reorganize_vertices()
    elif case.startswith("uneven"):
        net, rng = synth.uneven(0)[3]
        rates = synth.rate_sets(rng, net.N, B)
        m = model_builder().parse_model(net.bnet())
        t = transition_table(ctx, m, rates[0])
        t.construct_table()
        g = transition_graph(ctx, t)
        g.reorganize_vertices()
        x0 = initial_state(m.nodes).state
    else:
        cfg, seed = case.split("-")
        net, rates, m, t, g, x0 = setup_net(ctx, cfg, int(seed), B)
    n = len(x0)
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    outs = []
    for build in range(2):
        plan = solve_plan(ctx, t, dx0.ptr)
        for graph in (0, 1):
            d = DeviceArray(ctx, (B, n), np.float64)
            plan.solve(B, dr.ptr, d.ptr, precision=1, tol=1e-12, graph=graph)
            outs.append(d.download())
        plan.close()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_per_set_times_of_a_launch(ctx):
    # esl_plan_set_times: every set of the last per-set launch has a start and
    # an end, end >= start, all within one launch span
    import ctypes as C

    from paraexastolog_amd._lib import EslError, check, lib

    B = 16
    net, rates, m, t, g, x0 = setup_net(ctx, "c1", 0, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (B, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    plan.solve(B, dr.ptr, dout.ptr, precision=1, tol=1e-12)
    s0, s1 = np.zeros(B), np.zeros(B)
    check(lib().esl_plan_set_times(ctx.handle, plan.handle, B, s0.ctypes.data_as(C.POINTER(C.c_double)),
                                   s1.ctypes.data_as(C.POINTER(C.c_double))), ctx.handle)
    assert s0.min() == 0.0 and (s1 >= s0).all() and 0 < s1.max() < 1000.0
    assert lib().esl_plan_set_times(ctx.handle, plan.handle, B + 1, None, None) != 0  # another batch size
