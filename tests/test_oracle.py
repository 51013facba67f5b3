"""The oracle pinned against the reference's own goldens (cuda/test/test.cu),
closed forms and brute force -- before it is trusted as the checker."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pipeline as P
from oracle.bnet import BnetError, read_bnet, read_bnet_text

from .helpers import GOLDEN, golden, model_path, nonzero, oracle_rates

G = golden()

SMALL_CASES = [k for k, v in G.items() if v["model"] in ("toy.bnet", "toy2.bnet", "toy3.bnet")]


def run_oracle(case, full=True, rates_key="rates"):
    m = read_bnet(model_path(case["model"]))
    rates = oracle_rates(m, case, rates_key)
    x0 = P.initial_state(m, case.get("fixed", []), case.get("vals", []), case.get("p", 1.0))
    return m, P.Pipeline(m, rates, x0, full=full)


@pytest.mark.parametrize("name", ["trans_table.toy"])
def test_table_golden(name):
    c = G[name]
    m = read_bnet(model_path(c["model"]))
    ip, rows, cols = P.transition_table(m, oracle_rates(m, c))
    assert ip.tolist() == c["indptr"]
    assert cols.tolist() == c["cols"]


def test_initial_state_golden():
    c = G["initial_value.toy"]
    m = read_bnet(model_path(c["model"]))
    x0 = P.initial_state(m, c["fixed"], c["vals"], c["p"])
    assert x0.tolist() == c["state"]


@pytest.mark.parametrize("name", sorted(SMALL_CASES))
def test_small_goldens(name):
    c = G[name]
    if "nonzero_indices" not in c and "reordered_vertices" not in c:
        pytest.skip("table/x0 only")
    _, pl = run_oracle(c)
    if "reordered_vertices" in c:
        assert pl.reordered_vertices.tolist() == c["reordered_vertices"]
        assert pl.sccs_offsets[: pl.terminals_count + 1].tolist() == c["terminal_offsets"]
    if "nonzero_indices" in c:
        nz, v = nonzero(pl.final_state)
        assert nz.tolist() == c["nonzero_indices"]
        np.testing.assert_allclose(v, c["nonzero_data"], rtol=0, atol=max(c["tol"], 1e-15))
    for k in ("term_indptr", "term_indices", "nonterm_indptr", "nonterm_indices"):
        if k in c:
            assert getattr(pl.sol, k).tolist() == c[k]
    for k in ("term_data", "nonterm_data"):
        if k in c:
            np.testing.assert_allclose(getattr(pl.sol, k), c[k], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", ["solver.kras", "solver.mammal"])
def test_real_model_goldens(name):
    c = G[name]
    _, pl = run_oracle(c, full=False)
    nz, v = nonzero(pl.final_state)
    assert nz.tolist() == c["nonzero_indices"]
    np.testing.assert_allclose(v, c["nonzero_data"], rtol=0, atol=c["tol"])
    assert abs(pl.final_state.sum() - 1.0) < 1e-12


def test_symbolic_with_refactor_toy():
    # the reference reuses the symbolic factorisation of uniform rates for d_C = 2
    c = G["symbolic_with_refactor.toy"]
    _, pl = run_oracle(c)
    nz, v = nonzero(pl.final_state)
    assert nz.tolist() == c["nonzero_indices"]
    np.testing.assert_allclose(v, c["nonzero_data"], rtol=0, atol=c["tol"])


def test_closed_forms_rates_toy():
    # rates.toy (d_C = 2): x*(1) = 1/2, x*(2) = 11/48, x*(4) = 13/48
    _, pl = run_oracle(G["rates.toy"])
    x = pl.final_state
    assert abs(x[1] - 0.5) < 1e-15 and abs(x[2] - 11 / 48) < 1e-15 and abs(x[4] - 13 / 48) < 1e-15
    # zero_rate.toy (d_C = 0): state 3 becomes terminal
    _, pl = run_oracle(G["zero_rate.toy"])
    np.testing.assert_allclose(pl.final_state[[1, 2, 3, 4]], [1 / 4, 5 / 16, 1 / 4, 3 / 16], atol=1e-15)


def test_two_cycle_kernel_closed_form():
    # A 2-state irreducible chain a <-> b: pi ∝ (rate b->a, rate a->b)
    for ra, rb in [(1.0, 1.0), (0.3, 7.0), (10.0, 0.1)]:
        K = sp.csc_matrix(np.array([[-ra, rb], [ra, -rb]]))
        np.testing.assert_allclose(P.terminal_kernel(K), [rb / (ra + rb), ra / (ra + rb)], rtol=1e-14)


def test_row_replacement_equals_tree_theorem():
    # the reference's determinant form (Markov chain tree theorem) vs the row-replaced solve
    rng = np.random.default_rng(3)
    for s in range(2, 8):
        A = rng.uniform(0.1, 10.0, (s, s)) * (rng.random((s, s)) < 0.7)
        np.fill_diagonal(A, 0)
        A[np.arange(1, s), np.arange(s - 1)] += 0.5  # ensure irreducible (a cycle)
        A[0, s - 1] += 0.5
        K = A - np.diag(A.sum(axis=0))
        a = P.terminal_kernel(sp.csc_matrix(K))
        b = P.terminal_kernel_det(sp.csc_matrix(K))
        np.testing.assert_allclose(a, b, rtol=1e-10)
        assert np.abs(K @ a).max() < 1e-12


def _closure_scc(n, edges):
    R = np.eye(n, dtype=bool)
    for u, v in edges:
        R[u, v] = True
    for k in range(n):
        R |= R[:, [k]] & R[[k], :]
    same = R & R.T
    ids = np.array([np.nonzero(same[v])[0].min() for v in range(n)])
    term = np.array([not any(R[v, w] and not same[v, w] for w in range(n)) for v in range(n)])
    return ids, term, R


@pytest.mark.parametrize("seed", range(12))
def test_graph_brute_force(seed):
    # random orientation of a hypercube with random deletions: SCC ids (min vertex),
    # terminal flags, heights and the canonical order against transitive closure
    rng = np.random.default_rng(seed)
    N = 5 + seed % 3
    n = 1 << N
    mask = np.zeros(n, dtype=np.int64)
    for s in range(n):
        for i in range(N):
            t = s ^ (1 << i)
            if s < t and rng.random() < 0.9:
                if rng.random() < 0.5:
                    mask[s] |= 1 << i
                else:
                    mask[t] |= 1 << i
            if s < t and rng.random() < 0.08:  # some bidirectional pairs
                mask[s] |= 1 << i
                mask[t] |= 1 << i
    deg = np.array([bin(int(x)).count("1") for x in mask])
    indptr = np.r_[0, np.cumsum(deg)]
    rows = np.array([s ^ (1 << i) for s in range(n) for i in range(N) if (mask[s] >> i) & 1], dtype=np.int64)
    edges = [(s, s ^ (1 << i)) for s in range(n) for i in range(N) if (mask[s] >> i) & 1]
    ids_bf, term_bf, R = _closure_scc(n, edges)
    ids = P.scc_labels(n, indptr, rows)
    assert ids.tolist() == ids_bf.tolist()
    roots, h = P.scc_heights(n, indptr, rows, ids)
    # height by brute force: longest path in the condensation DAG
    hb = {}

    def height(r):
        if r in hb:
            return hb[r]
        succ = {ids_bf[w] for v in range(n) if ids_bf[v] == r for (a, w) in edges if a == v and ids_bf[w] != r}
        hb[r] = 1 + max((height(x) for x in succ), default=0)
        return hb[r]

    assert [height(r) for r in roots] == h.tolist()
    reord, off, T, _ = P.transition_graph(n, indptr, rows)
    hmap = dict(zip(roots.tolist(), h.tolist()))
    keys = [((0, r) if hmap[r] == 1 else (1, -hmap[r], -r)) for r in roots.tolist()]
    order = [r for _, r in sorted(zip(keys, roots.tolist()))]
    expect = [v for r in order for v in range(n) if ids_bf[v] == r]
    assert reord.tolist() == expect
    assert T == sum(1 for r in roots if hmap[r] == 1)


def test_single_rhs_equals_full_L():
    # A.8: R (L x0) through the full L (T right-hand sides) equals the single
    # right-hand side N w = x0_nt, on every database-small model small enough
    import os

    for f in sorted(os.listdir(model_path("database-small")))[:8]:
        m = read_bnet(model_path("database-small/" + f))
        if m.n_nodes > 11:
            continue
        rates = np.exp(np.random.default_rng(len(f)).uniform(np.log(0.1), np.log(10), 2 * m.n_nodes))
        x0 = P.initial_state(m)
        a = P.Pipeline(m, rates, x0, full=True).final_state
        b = P.Pipeline(m, rates, x0, full=False).final_state
        assert np.abs(a - b).sum() < 1e-12
        assert abs(a.sum() - 1) < 1e-12 and a.min() >= -1e-15


def test_reachability_pruned_solve_equals_full():
    # the pruned final-only oracle (used at N=20) against the full solve
    import os

    for f in sorted(os.listdir(model_path("database-small"))):
        m = read_bnet(model_path("database-small/" + f))
        if m.n_nodes > 12:
            continue
        rates = np.exp(np.random.default_rng(len(f) + 7).uniform(np.log(0.1), np.log(10), 2 * m.n_nodes))
        n = 1 << m.n_nodes
        x0 = P.initial_state(m, [m.nodes[0]], [1])
        ip, rows, _ = P.transition_table(m, rates)
        g = P.transition_graph(n, ip, rows)
        a = P.solve(n, ip, rows, rates, x0, g, full=True).final_state
        b = P.solve_final(n, ip, rows, rates, x0, g)
        assert np.abs(a - b).sum() < 1e-12


def test_parser_errors_and_grammar():
    with pytest.raises(BnetError):
        read_bnet_text("targets, factors\nA A\n")
    with pytest.raises(BnetError):
        read_bnet_text("targets, factors\nA, (A\n")
    with pytest.raises(BnetError):
        read_bnet_text("targets, factors\nA, A B\n")
    m = read_bnet_text("targets, factors\nA, !!A | B & !C\nB, B\nC, (A|B)&C")  # no trailing newline kept
    assert m.nodes == ["A", "B", "C"]
    # right associative, & binds tighter than |
    for a, b, c in itertools.product([0, 1], repeat=3):
        s = np.array([a | (b << 1) | (c << 2)])
        assert bool(m.evaluate(0, s)[0]) == bool(a or (b and not c))
        assert bool(m.evaluate(2, s)[0]) == bool((a or b) and c)
