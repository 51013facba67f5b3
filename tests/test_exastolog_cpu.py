"""The restatement of the reference's CPU path (oracle/exastolog_cpu.py, the
bench's reference-path CPU baseline) pinned against the reference's goldens,
against the FP64 oracle, and on the defects it fixes."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import exastolog_cpu as E
from oracle import pipeline as P
from oracle.bnet import read_bnet, read_bnet_text
from paraexastolog_amd import synth

from .helpers import golden, model_path, oracle_rates

G = golden()


@pytest.mark.parametrize("name", ["solver.toy", "solver.toy2", "solver.toy3", "rates.toy", "zero_rate.toy"])
def test_reference_goldens(name):
    c = G[name]
    m = read_bnet(model_path(c["model"]))
    rates = oracle_rates(m, c)
    x0 = E.initial_state(m, c.get("fixed", []), c.get("vals", []), c.get("p", 1.0))
    for kernel in ("det", "rowrep"):
        x, _ = E.run(m, rates, x0, kernel=kernel)
        nz = np.nonzero(x)[0]
        assert nz.tolist() == c["nonzero_indices"], (name, kernel)
        np.testing.assert_allclose(x[nz], c["nonzero_data"], atol=max(c["tol"], 1e-12))


@pytest.mark.parametrize("seed", range(8))
def test_matches_oracle_on_seeded_networks(seed):
    net, rng = synth.config("c0", seed)
    rates = synth.rate_sets(rng, net.N, 1)[0]
    m = read_bnet_text(net.bnet())
    fixed, vals = net.fixed_spec()
    x0 = E.initial_state(m, fixed, vals)
    np.testing.assert_array_equal(x0, P.initial_state(m, fixed, vals))
    ref = P.Pipeline(m, rates, P.initial_state(m, fixed, vals), full=False).final_state
    for kernel in ("det", "rowrep"):
        x, _ = E.run(m, rates, x0, kernel=kernel)
        assert np.abs(x - ref).sum() <= 1e-12, (seed, kernel)


def test_table_equals_oracle_table():
    # the clause enumeration (exastolog.py:81-118) gives the oracle's edge set
    net, rng = synth.config("c0", 3)
    rates = synth.rate_sets(rng, net.N, 1)[0]
    m = read_bnet_text(net.bnet())
    T = E.transition_table(m, rates).tocoo()
    ip, rows, cols = P.transition_table(m, rates)
    got = sorted(zip(T.col.tolist(), T.row.tolist()))
    want = sorted(zip(cols.tolist(), rows.tolist()))
    assert got == want
    K = P.kinetic(1 << net.N, ip, rows, rates)
    assert abs(E.create_kinetic(T.tocsr()) - K).max() <= 1e-12  # diagonal sums in another order


def _buggy_column_nullspace(table, offsets, kernel="det", deadline=None):
    """compute_column_nullspace as written (exastolog.py:371-375): the data
    index is the terminal index i, not the SCC's offset."""
    T = len(offsets) - 1
    ts = offsets[0]
    n = table.shape[0]
    data = np.empty(n - ts)
    indices, indptr = [], [0] * (ts + 1)
    for i in range(T):
        a, b = offsets[i], offsets[i + 1]
        size = b - a
        for _ in range(size):
            indptr.append(indptr[-1] + 1)
        indices += [i] * size
        if size == 1:
            data[i] = 1.0
        else:
            data[i:i + size] = E.scc_kernel_rowrep(E.create_kinetic(table[a:b, a:b]).toarray())
    return sp.csr_matrix((data, indices, indptr), shape=(n, T))


def test_column_nullspace_offset_defect_is_fixed(monkeypatch):
    # c0 seed 10: two reachable terminal SCCs of 10 and 1 states; the written
    # index lets the singleton's 1 overwrite the second entry of the first kernel
    net, rng = synth.config("c0", 10)
    rates = synth.rate_sets(rng, net.N, 1)[0]
    m = read_bnet_text(net.bnet())
    fixed, vals = net.fixed_spec()
    x0 = E.initial_state(m, fixed, vals)
    ref = P.Pipeline(m, rates, P.initial_state(m, fixed, vals), full=False).final_state
    x, _ = E.run(m, rates, x0, kernel="rowrep")
    assert np.abs(x - ref).sum() <= 1e-12
    monkeypatch.setattr(E, "column_nullspace", _buggy_column_nullspace)
    xb, _ = E.run(m, rates, x0, kernel="rowrep")
    assert np.abs(xb - ref).sum() > 1e-3


def test_determinant_kernel_leaves_fp64_range_on_large_sccs():
    # a directed cycle of 80 states at rate 1e5: every minor's determinant is
    # 1e395 > 1.8e308, so the written method yields non-finite kernels, while
    # row replacement gives the uniform kernel
    s = 80
    idx = np.arange(s)
    T = sp.csr_matrix((np.full(s, 1e5), ((idx + 1) % s, idx)), shape=(s, s))
    K = E.create_kinetic(T).toarray()
    det = E.scc_kernel_det(K)
    assert not np.isfinite(det).all()
    np.testing.assert_allclose(E.scc_kernel_rowrep(K), np.full(s, 1.0 / s), rtol=1e-12)


def test_cap_stops_the_determinant_loop():
    s = 300
    idx = np.arange(s)
    T = sp.csr_matrix((np.full(s, 1.0), ((idx + 1) % s, idx)), shape=(s, s))
    K = E.create_kinetic(T).toarray()
    import time

    with pytest.raises(E.CapExceeded):
        E.scc_kernel_det(K, deadline=time.perf_counter() + 1e-4)
