"""Test infrastructure and CPU baseline only: a stage-for-stage restatement of
the reference's CPU path, python/exastolog.py (numpy + scipy, FP64).  The
product (paraexastolog_amd + libesl.so) never imports it; tests use it as a
checker and bench.py times it as the reference CPU path.

python/exastolog.py itself cannot run here (pyeda and sortednp are not
installed, exastolog.py:2-3, and scikits.umfpack is absent, :395-396), so each
stage is restated from its source with these substitutions, each marked where
it happens:

  S1  DNF: the clauses of a rule (and of its negation) are the minterms of the
      rule's truth table over the nodes it reads (oracle.bnet) instead of
      pyeda's minimised DNF (:34-79).  Same set of states per clause union.
  S2  sortednp.kway_merge(..., duplicates=DROP) (:116-117) is np.unique of the
      concatenated sorted clause enumerations (the same sorted union).
  S3  rates: the table holds the rate of each flip (rates[2i] up, [2i+1] down,
      the CUDA reference's layout and semantics, kernels/solver.cu:33-36;
      a zero rate removes that direction, transition_table.cu:79-83) instead of
      ones scaled by 1/(2N) in create_kinetic (:328-331).
  S4  scipy SuperLU (splu) instead of UMFPACK in the row-nullspace solve
      (:395-398).
Reference defects fixed (and flagged by the tests):
  B1  compute_column_nullspace writes data[i : i + scc_size] and data[i] with
      the terminal index i instead of the SCC's offset (:371-375); here the
      offset is used.
  B2  an isolated state is stored as (1, ...) and compute_final_states then
      calls len() on the int (:292-294, :407); here isolated states keep their
      initial mass, which is what the intent (":406 just a single disconnected
      state") says.
  B3  the determinant sign `-1 ** (n - 1)` parses as -(1 ** (n - 1)) = -1
      (:349); the sign cancels in the normalisation, so it is kept.

The determinant kernel (:333-354) costs one (s-1)x(s-1) determinant per state
of a terminal SCC, Theta(s^4) in all.  It runs under a wall-clock deadline:
after the first determinants the remaining time is extrapolated and, past the
deadline, CapExceeded is raised (the caller reports DNF).  kernel="rowrep" is
the labelled variant: the same column nullspace from one row-replaced dense
solve per SCC (K_j with row 0 replaced by ones, right-hand side e_0).

Command line (used by bench.py in a child process with a time and memory
limit): python -m oracle.exastolog_cpu --config c1 --seed 0 [--kernel det]
[--cap 30]; it prints one JSON line per finished stage and a last line with
the result summary.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as sla

from .bnet import Model, read_bnet_text


class CapExceeded(RuntimeError):
    """The wall-clock deadline of the run would be (or was) passed in `stage`."""

    def __init__(self, stage, detail):
        super().__init__(f"{stage}: {detail}")
        self.stage = stage
        self.detail = detail


# ----------------------------------------------------------------------------
# TransitionTable (exastolog.py:29-172)


def _minterm_clauses(model: Model, i, value):
    """S1: the clauses of f_i (value True) or of !f_i (value False) as
    {node index: literal} dicts, one per satisfying assignment of the inputs."""
    ins = model.inputs[i]
    tab = model.tables[i]
    out = []
    for idx in np.nonzero(tab == value)[0]:
        out.append({j: bool((int(idx) >> k) & 1) for k, j in enumerate(ins)})
    return out


def _clause_transitions(N, node_idx, clauses, up_down):
    """get_transitions (exastolog.py:81-118): for every clause, the source states
    that satisfy it with the node at `not up_down`, enumerated by doubling over
    the free nodes; the sorted union over the clauses (S2)."""
    node_values = 2 ** np.arange(N, dtype=np.int64)
    parts = [np.empty(0, dtype=np.int64)]
    for clause in clauses:
        if node_idx in clause and clause[node_idx] == up_down:
            continue  # no change (:86-88)
        idxs = list(clause.keys())
        vals = [clause[j] for j in idxs]
        if node_idx not in clause:
            idxs.append(node_idx)
            vals.append(not up_down)
        static_val = int(np.sum(node_values[idxs] * np.asarray(vals, dtype=np.int64)))
        free = [j for j in range(N) if j not in idxs]
        tr = np.empty(1 << len(free), dtype=np.int64)
        tr[0] = static_val
        ln = 1
        for j in free:  # (:104-111)
            tr[ln:2 * ln] = tr[:ln] + (1 << j)
            ln *= 2
        parts.append(tr)
    return np.unique(np.concatenate(parts))


def transition_table(model: Model, rates):
    """build_transition_table (exastolog.py:137-172) with S3: T[dst, src] = rate
    of the flip src -> dst (CSR, n x n)."""
    N = model.n_nodes
    n = 1 << N
    src_all, dst_all, val_all = [], [], []
    for i in range(N):
        up_src = _clause_transitions(N, i, _minterm_clauses(model, i, True), True)
        dn_src = _clause_transitions(N, i, _minterm_clauses(model, i, False), False)
        for srcs, sign, r in ((up_src, 1, rates[2 * i]), (dn_src, -1, rates[2 * i + 1])):
            if r <= 0 or len(srcs) == 0:
                continue  # a zero rate removes the direction (S3)
            src_all.append(srcs)
            dst_all.append(srcs + sign * (1 << i))
            val_all.append(np.full(len(srcs), float(r)))
    if src_all:
        src = np.concatenate(src_all)
        dst = np.concatenate(dst_all)
        val = np.concatenate(val_all)
    else:
        src = dst = np.empty(0, dtype=np.int64)
        val = np.empty(0)
    return sp.csr_matrix((val, (dst, src)), shape=(n, n))


# ----------------------------------------------------------------------------
# InitialState (exastolog.py:175-203)


def initial_state(model: Model, fixed_nodes=(), fixed_values=(), p=1.0):
    N = model.n_nodes
    n = 1 << N
    states = np.arange(n, dtype=np.int64)
    mask = np.ones(n, dtype=bool)
    for name, v in zip(fixed_nodes, fixed_values):
        j = model.nodes.index(name)
        mask &= ((states >> j) & 1).astype(bool) == bool(v)
    x0 = np.empty(n)
    nf = int(mask.sum())
    x0[mask] = p / nf
    x0[~mask] = (1 - p) / (n - nf) if n > nf else 0.0
    return x0


# ----------------------------------------------------------------------------
# TransitionGraph (exastolog.py:206-316)


def toposort(table):
    """Peeling toposort of the metagraph (exastolog.py:213-232): metavertices
    whose every out-edge leads to already peeled ones go to the back."""
    orig = table
    ordering = np.empty(table.shape[0], dtype=np.int64)
    mask = np.zeros(table.shape[0], dtype=bool)
    terminals_idx = None
    fill = table.shape[0]
    while table.shape[0] != 0:
        new_mask = np.asarray(table.sum(axis=0) == 0).ravel()
        idx = np.nonzero(new_mask ^ mask)[0]
        fill -= len(idx)
        ordering[fill:fill + len(idx)] = idx
        mask = new_mask
        table = orig[~mask, :]
        if terminals_idx is None:
            terminals_idx = fill
    return ordering, terminals_idx


def create_metagraph(table):
    """Strong components and the condensation (exastolog.py:234-268)."""
    count, labels = csg.connected_components(table, directed=True, connection="strong")
    if count == table.shape[0]:
        return [np.array([v]) for v in range(count)], table
    sort_idx = np.argsort(labels, kind="stable")
    ls = labels[sort_idx]
    first = np.concatenate(([True], ls[1:] != ls[:-1]))
    cnt = np.diff(np.nonzero(first)[0])
    sccs = np.split(sort_idx, np.cumsum(cnt))
    rows, cols = table.nonzero()
    cross = labels[rows] != labels[cols]
    meta = sp.csr_matrix((np.ones(int(cross.sum())), (labels[rows[cross]], labels[cols[cross]])),
                         shape=(count, count))
    meta.data = np.ones(meta.data.shape)
    return sccs, meta


def sort_graph(table, x0):
    """TransitionGraph.sort (exastolog.py:277-316): per weakly connected
    component with initial mass, its sorted sub-table, terminal offsets and
    vertex order; components are skipped when x0 vanishes on them."""
    count, comp = csg.connected_components(table, directed=True, connection="weak")
    order = np.argsort(comp, kind="stable")
    bounds = np.concatenate(([0], np.cumsum(np.bincount(comp, minlength=count))))
    out = []
    for c in range(count):
        sub = order[bounds[c]:bounds[c + 1]]
        if np.all(x0[sub] == 0):
            continue
        if len(sub) == 1:
            out.append((sub, None, [0, 1], np.array([0])))  # B2: kept as a 1-state component
            continue
        subtable = table[sub, :][:, sub]
        sccs, meta = create_metagraph(subtable)
        if len(sccs) == 1:
            out.append((sub, subtable, [0, len(sub)], np.arange(len(sub))))
            continue
        ordering, term_start = toposort(meta)
        sorted_vertices = np.concatenate([sccs[k] for k in ordering])
        sorted_sub = subtable[sorted_vertices, :][:, sorted_vertices]
        offs = [0]
        for k in range(term_start, len(ordering)):
            offs.append(offs[-1] + len(sccs[ordering[k]]))
        offs = [x + (len(sub) - offs[-1]) for x in offs]
        out.append((sub, sorted_sub, offs, sorted_vertices))
    return out


# ----------------------------------------------------------------------------
# Solution (exastolog.py:319-416)


def create_kinetic(trans_table):
    """K = T - diag(column sums) (exastolog.py:328-331; S3: T holds the rates)."""
    return trans_table - sp.diags(np.asarray(trans_table.sum(axis=0)).ravel())


def scc_kernel_det(K, deadline=None):
    """compute_scc_kernel (exastolog.py:333-354): delete row 0 of the SCC's
    dense generator, kernel[i] = sign * det(minor without column i), then
    normalise (B3: the constant sign cancels).  Raises CapExceeded when the
    extrapolated time of the remaining determinants passes `deadline`."""
    s = K.shape[0]
    Kd = np.delete(K, 0, axis=0)
    kernel = np.empty(s)
    sign = -1.0  # -1 ** (s - 1) as written parses as -1 (B3)
    t0 = time.perf_counter()
    for i in range(s):
        kernel[i] = sign * scipy.linalg.det(np.delete(Kd, i, axis=1))
        sign *= -1.0
        if deadline is not None and (i in (0, 3, 15, 63) or i % 256 == 255):
            est = (time.perf_counter() - t0) / (i + 1) * s
            if t0 + est > deadline:
                raise CapExceeded("terminal kernel (determinants)",
                                  f"SCC of {s} states: {s} determinants of {s - 1}x{s - 1}, "
                                  f"{est:.0f} s extrapolated from {i + 1}")
    with np.errstate(invalid="ignore", over="ignore"):
        # determinants of large minors leave the FP64 range (inf / inf = nan),
        # as they do in the reference; the caller reports a non-finite result
        return kernel / kernel.sum()


def scc_kernel_rowrep(K):
    """Labelled variant of compute_scc_kernel: the normalised null vector from
    one dense solve with row 0 of K replaced by ones (right-hand side e_0)."""
    s = K.shape[0]
    A = np.array(K, dtype=float)
    A[0, :] = 1.0
    b = np.zeros(s)
    b[0] = 1.0
    v = scipy.linalg.solve(A, b)
    return v / v.sum()


def column_nullspace(table, offsets, kernel="det", deadline=None):
    """compute_column_nullspace (exastolog.py:356-377), B1 fixed: R (n x T)
    with the kernel of terminal SCC i in column i on that SCC's rows."""
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
            data[a - ts] = 1.0
        else:
            K = create_kinetic(table[a:b, a:b]).toarray()
            data[a - ts:b - ts] = scc_kernel_det(K, deadline) if kernel == "det" else scc_kernel_rowrep(K)
    return sp.csr_matrix((data, indices, indptr), shape=(n, T))


def row_nullspace(table, offsets, R):
    """compute_row_nullspace (exastolog.py:379-400), S4: L = [X | U] with
    U = R[ts:]^T (ones) and X = -U B N^{-1} from a SuperLU factorisation of N^T."""
    ts = offsets[0]
    U = R[ts:, :].transpose().tocsr()
    U.data = np.ones(len(U.data))
    if ts == 0:
        return U
    K = create_kinetic(table).tocsr()
    Nm = K[:ts, :][:, :ts]
    Bm = K[ts:, :][:, :ts]
    X = -(U @ Bm)  # T x ts
    lu = sla.splu(Nm.transpose().tocsc())
    Xs = lu.solve(np.asarray(X.toarray().T))  # ts x T
    return sp.hstack((sp.csr_matrix(Xs.T), U)).tocsr()


def final_states(table, x0, kernel="det", deadline=None, progress=None):
    """compute_final_states (exastolog.py:402-416): x <- R L x on each sorted
    component.  `progress(stage, seconds)` is called after each stage."""
    t = time.perf_counter()
    comps = sort_graph(table, x0)
    if progress:
        progress("graph (weak components, SCCs, toposort, sorted sub-tables)", time.perf_counter() - t)
    x = x0.copy()
    t_R = t_L = 0.0
    for sub, subtable, offs, sorted_v in comps:
        if len(sub) == 1:
            continue  # B2
        t = time.perf_counter()
        R = column_nullspace(subtable, offs, kernel, deadline)
        t_R += time.perf_counter() - t
        if progress:
            progress("column nullspace (terminal kernels)", t_R)
            if not np.isfinite(R.data).all():
                progress("column nullspace: non-finite kernel entries (determinant out of FP64 range)",
                         float(np.count_nonzero(~np.isfinite(R.data))))
        if deadline is not None and time.perf_counter() > deadline:
            raise CapExceeded("row nullspace", "deadline passed before the transient solve")
        t = time.perf_counter()
        L = row_nullspace(subtable, offs, R)
        t_L += time.perf_counter() - t
        if progress:
            progress("row nullspace (SuperLU on N^T)", t_L)
        v = sub[sorted_v]
        x[v] = R @ (L @ x[v])
    return x


def run(model: Model, rates, x0, kernel="det", cap_s=None, progress=None):
    """The whole CPU path; returns (final vector, stage seconds)."""
    deadline = time.perf_counter() + cap_s if cap_s else None
    stages = {}

    def prog(stage, sec):
        stages[stage] = sec
        if progress:
            progress(stage, sec)

    t = time.perf_counter()
    T = transition_table(model, rates)
    prog("transition table", time.perf_counter() - t)
    x = final_states(T, x0, kernel, deadline, prog)
    return x, stages


def main(argv=None):
    import argparse
    import os

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c1")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rate-set", type=int, default=0)
    ap.add_argument("--kernel", default="det", choices=["det", "rowrep"])
    ap.add_argument("--cap", type=float, default=30.0)
    a = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from paraexastolog_amd import synth

    net, rng = synth.config(a.config, a.seed)
    rates = synth.rate_sets(rng, net.N, a.rate_set + 1)[a.rate_set]
    model = read_bnet_text(net.bnet())
    fixed, vals = net.fixed_spec()
    x0 = initial_state(model, fixed, vals)

    def progress(stage, sec):
        print(json.dumps({"stage": stage, "s": sec}), flush=True)

    t = time.perf_counter()
    try:
        x, stages = run(model, rates, x0, a.kernel, a.cap, progress)
        res = {"status": "ok" if np.isfinite(x).all() else "nan", "s": time.perf_counter() - t, "stages": stages,
               "sum": float(x.sum()), "nonzero": int(np.count_nonzero(x))}
        if os.environ.get("ESL_CPU_OUT"):
            np.save(os.environ["ESL_CPU_OUT"], x)
    except CapExceeded as e:
        res = {"status": "DNF", "stage": e.stage, "detail": e.detail, "s": time.perf_counter() - t}
    except MemoryError:
        res = {"status": "DNF", "stage": "memory", "detail": "memory limit of the run", "s": time.perf_counter() - t}
    print(json.dumps({"result": res}), flush=True)


if __name__ == "__main__":
    main()
