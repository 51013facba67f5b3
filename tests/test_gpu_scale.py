"""The path at the sizes of the benchmark workloads: the N=20 network of the
headline metric against the oracle (bit-exact table and graph stages, final
vector within 1e-12 L1), and the N=22 network with spliced negative-feedback
cycles through size-independent checks (terminal SCC size, global balance of
the stationary vector, FP32+refinement against FP64)."""
import numpy as np
import pytest

from oracle import pipeline as P
from oracle.bnet import read_bnet_text
from paraexastolog_amd import initial_state, model_builder, solve_plan, solver, synth, transition_graph, transition_table
from paraexastolog_amd.surrogate import DeviceArray

pytestmark = pytest.mark.gpu


def test_n20_metric_network_matches_oracle(ctx):
    net, rng = synth.config("c1", 0)
    rates = synth.rate_sets(rng, net.N, 1)[0]
    n = 1 << net.N
    m = model_builder().parse_model(net.bnet())
    t = transition_table(ctx, m, rates)
    t.construct_table()
    g = transition_graph(ctx, t)
    g.reorganize_vertices()
    om = read_bnet_text(net.bnet())
    ip, rows, _ = P.transition_table(om, rates)
    assert np.array_equal(t.indptr, ip) and np.array_equal(t.rows, rows)
    # no rule reads its own node and every rate is positive: m = N 2^(N-1) exactly
    assert len(rows) == net.N * (1 << (net.N - 1))
    reordered, offsets, T, _ = gr = P.transition_graph(n, ip, rows)
    assert np.array_equal(g.reordered_vertices, reordered)
    assert np.array_equal(g.sccs_offsets, offsets) and g.terminals_count == T
    fixed, vals = net.fixed_spec()
    x0 = initial_state(m.nodes, fixed, vals).state
    s = solver(ctx, t, g, rates, x0)
    x = s.solve()
    ref = P.solve_final(n, ip, rows, rates, P.initial_state(om, fixed, vals), gr)
    assert np.abs(x - ref).sum() <= 1e-12


def _edge_rates(rates, src, dst):
    i = np.log2((src ^ dst).astype(np.float64)).astype(np.int64)
    up = dst > src  # the flipped bit is set in the target
    return rates[2 * i + np.where(up, 0, 1)]


def test_n22_spliced_cycles_give_large_terminal_scc_and_balanced_vectors(ctx):
    net, rng = synth.config("c3", 8)
    # the splice rule: f_b = a, f_c = b, f_a = !c for every drawn triple
    assert len(net.cycles) == 3 and len({v for tr in net.cycles for v in tr}) == 9
    for a, b, c in net.cycles:
        assert net.inputs[b] == [a] and net.tables[b] == [0, 1]
        assert net.inputs[c] == [b] and net.tables[c] == [0, 1]
        assert net.inputs[a] == [c] and net.tables[a] == [1, 0]
    B = 2
    rates = synth.rate_sets(rng, net.N, B)
    n = 1 << net.N
    m = model_builder().parse_model(net.bnet())
    t = transition_table(ctx, m, rates[0])
    t.construct_table()
    g = transition_graph(ctx, t)
    g.reorganize_vertices()
    sizes = np.diff(np.asarray(g.sccs_offsets)[: g.terminals_count + 1])
    assert sizes.max() >= 10 ** 4  # what the splice is for
    fixed, vals = net.fixed_spec()
    x0 = initial_state(m.nodes, fixed, vals).state
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    plan = solve_plan(ctx, t, dx0.ptr)
    d64 = DeviceArray(ctx, (B, n), np.float64)
    d32 = DeviceArray(ctx, (B, n), np.float64)
    plan.solve(B, dr.ptr, d64.ptr, precision=0, tol=1e-12)
    plan.solve(B, dr.ptr, d32.ptr, precision=1, tol=1e-12)
    x64, x32 = d64.download(), d32.download()
    assert np.abs(x32 - x64).sum(axis=1).max() <= 1e-10
    cols = np.asarray(t.cols, dtype=np.int64)
    rws = np.asarray(t.rows, dtype=np.int64)
    for k in range(B):
        x = x64[k]
        assert abs(x.sum() - 1) < 1e-12 and x.min() >= -1e-15
        # global balance on the support: what flows into each state equals what leaves it
        sel = x[cols] > 0
        src, dst = cols[sel], rws[sel]
        flow = _edge_rates(rates[k], src, dst) * x[src]
        inflow = np.bincount(dst, weights=flow, minlength=n)
        outflow = np.bincount(src, weights=flow, minlength=n)
        assert np.abs(inflow - outflow).sum() <= 1e-10 * outflow.sum()
