"""Oracle (test infrastructure only): FP64 CPU restatement of the stationary
distribution pipeline of asynchronous Boolean networks.

Used by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py`` as the CHECKER.  The product (paraexastolog_amd + libesl.so) never
imports it.

Stages and the reference code each one restates:

=====================  =====================================================
rates_uniform          transition_rates::generate / generate_uniform
                       (cuda/src/transition_rates.cpp:7-73)
initial_state          initial_state::initial_state (cuda/src/initial_state.cu:7-60),
                       python/exastolog.py:175-203
transition_table       transition_table::generate_transitions / compute_rows_and_cols /
                       construct_table (cuda/src/transition_table.cu:76-174);
                       python/exastolog.py:81-172
transition_graph       transition_graph::reorganize_vertices, toposort,
                       create_metagraph (cuda/src/transition_graph.cu:76-202);
                       topological_labelling / reorganize
                       (cuda/src/kernels/transition_graph.cu:3-52)
kinetic                scatter_rows_data (cuda/src/kernels/solver.cu:4-45);
                       python/exastolog.py:328-331
terminal_kernels       solver::solve_terminal_part (cuda/src/solver.cu:102-159);
                       python/exastolog.py:333-377
absorption_full        solver::solve_nonterminal_part (cuda/src/solver.cu:205-300);
                       python/exastolog.py:379-400
final_state            solver::compute_final_states (cuda/src/solver.cu:302-309)
=====================  =====================================================

Documented substitutions (the reference's own methods are not reproduced
where they are numerically or asymptotically unusable):

* the terminal kernel is the normalised null vector of the SCC's generator,
  obtained by replacing one row with ones (direct FP64 solve) instead of one
  determinant per vertex (solver.cu:137-150).  ``terminal_kernel_det`` keeps
  the determinant (Markov-chain tree theorem) form for small SCCs so tests can
  pin the two against each other;
* SuperLU (scipy) replaces cuSOLVER/cusparse and UMFPACK;
* FP64 throughout (the CUDA reference is FP32, cuda/src/types.h:7).

Canonical SCC identity: an SCC is named by its minimum vertex.  The
reference's SCC root ids come from an atomicCAS race (cuda/src/sga/bitset.cu:415-423);
all of its order goldens agree with the min-vertex naming.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as sla

from .bnet import Model

# ----------------------------------------------------------------------------
# rates and initial state


def rates_uniform(model: Model, up=None, down=None, default=1.0):
    """Layout [2i] = up rate of node i, [2i+1] = down rate (transition_rates.cpp:40-53)."""
    up = dict(up or {})
    down = dict(down or {})
    r = np.full(2 * model.n_nodes, float(default))
    for d, off in ((up, 0), (down, 1)):
        for name, v in d.items():
            if name not in model.nodes:
                raise KeyError(f"node {name} not found")
            if v < 0:
                raise ValueError("transition rate must be non-negative")
            r[2 * model.nodes.index(name) + off] = float(v)
    return r


def initial_state(model: Model, fixed=(), vals=(), p=1.0):
    """x0[s] = p/|F| on states matching every fixed (node, value), (1-p)/(n-|F|)
    elsewhere; uniform 1/n without fixed nodes (initial_state.cu:7-60)."""
    N = model.n_nodes
    n = 1 << N
    if len(fixed) != len(vals):
        raise ValueError("number of fixed node names not equal to number of fixed node values")
    if len(fixed) == 0:
        return np.full(n, 1.0 / n)
    s = np.arange(n, dtype=np.int64)
    m = np.ones(n, dtype=bool)
    for name, v in zip(fixed, vals):
        if name not in model.nodes:
            raise KeyError(f"invalid initial node name: {name}")
        i = model.nodes.index(name)
        m &= ((s >> i) & 1) == int(bool(v))
    nf = int(m.sum())
    x0 = np.empty(n)
    x0[m] = p / nf
    x0[~m] = (1.0 - p) / (n - nf) if n > nf else 0.0
    return x0


# ----------------------------------------------------------------------------
# transition table


def flip_masks(model: Model, rates):
    """mask[s] bit i set iff s -> s^(1<<i) is a transition: f_i(s) != s_i and the
    rate of that direction is > 0 (transition_table.cu:44-55,79-83,102-116)."""
    N = model.n_nodes
    n = 1 << N
    s = np.arange(n, dtype=np.int64)
    mask = np.zeros(n, dtype=np.int64)
    for i in range(N):
        f = model.evaluate(i, s)
        bit = ((s >> i) & 1).astype(bool)
        flip = f != bit
        if rates[2 * i] == 0.0:
            flip &= bit  # no up transitions
        if rates[2 * i + 1] == 0.0:
            flip &= ~bit  # no down transitions
        mask |= flip.astype(np.int64) << i
    return mask


def transition_table(model: Model, rates):
    """CSC over source states: indptr[n+1], rows = destinations, cols = sources
    (expanded).  Inside a column the destinations follow ascending node index."""
    N = model.n_nodes
    mask = flip_masks(model, rates)
    n = mask.shape[0]
    deg = np.zeros(n, dtype=np.int64)
    for i in range(N):
        deg += (mask >> i) & 1
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=indptr[1:])
    m = int(indptr[-1])
    rows = np.empty(m, dtype=np.int64)
    cols = np.repeat(np.arange(n, dtype=np.int64), deg)
    fill = indptr[:-1].copy()
    s = np.arange(n, dtype=np.int64)
    for i in range(N):
        has = ((mask >> i) & 1).astype(bool)
        rows[fill[has]] = s[has] ^ (1 << i)
        fill[has] += 1
    return indptr, rows, cols


# ----------------------------------------------------------------------------
# graph: SCCs, heights, canonical order


def scc_labels(n, indptr, rows):
    """Canonical SCC id of every vertex = minimum vertex of its SCC (scipy's
    strong components, renamed)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    A = sp.csr_matrix((np.ones(len(rows), dtype=np.int8), (cols, rows)), shape=(n, n))
    _, lab = csg.connected_components(A, directed=True, connection="strong")
    root = np.full(lab.max() + 1, n, dtype=np.int64)
    np.minimum.at(root, lab, np.arange(n, dtype=np.int64))
    return root[lab]


def scc_heights(n, indptr, rows, ids):
    """Height per SCC root: 1 for terminal SCCs, else 1 + max over successor
    SCCs (the fixpoint of topological_labelling, kernels/transition_graph.cu:3-27).
    Returns (roots ascending, heights aligned with roots)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    cs, cd = ids[cols], ids[rows]
    cross = cs != cd
    roots = np.unique(ids)
    S = len(roots)
    pos = np.full(n, -1, dtype=np.int64)
    pos[roots] = np.arange(S)
    u = pos[cs[cross]]
    v = pos[cd[cross]]
    if len(u):
        key = np.unique(u * S + v)
        u, v = key // S, key % S
    outdeg = np.bincount(u, minlength=S)
    # predecessors CSR keyed by v
    order = np.argsort(v, kind="stable")
    pu = u[order]
    pptr = np.zeros(S + 1, dtype=np.int64)
    np.cumsum(np.bincount(v, minlength=S), out=pptr[1:])
    h = np.zeros(S, dtype=np.int64)
    frontier = np.nonzero(outdeg == 0)[0]
    label = 1
    remaining = outdeg.copy()
    while len(frontier):
        h[frontier] = label
        starts, ends = pptr[frontier], pptr[frontier + 1]
        cnt = ends - starts
        if cnt.sum() == 0:
            break
        idx = np.repeat(starts - np.cumsum(np.r_[0, cnt[:-1]]), cnt) + np.arange(cnt.sum())
        preds = pu[idx]
        dec = np.bincount(preds, minlength=S)
        before = remaining > 0
        remaining -= dec
        frontier = np.nonzero(before & (remaining == 0))[0]
        label += 1
    assert (h > 0).all(), "condensation graph is not a DAG"
    return roots, h


def transition_graph(n, indptr, rows):
    """(reordered_vertices, sccs_offsets, terminals_count) in the canonical
    order: terminal SCCs by ascending id, then transient SCCs by (height
    descending, id descending), vertices ascending inside each SCC
    (transition_graph.cu:128-198)."""
    ids = scc_labels(n, indptr, rows)
    roots, h = scc_heights(n, indptr, rows, ids)
    S = len(roots)
    if S == 1:
        return np.arange(n, dtype=np.int64), np.array([0, n], dtype=np.int64), 1, ids
    term = h == 1
    # SCC order keys
    t_roots = roots[term]
    nt = np.nonzero(~term)[0]
    nt_order = nt[np.lexsort((-roots[nt], -h[nt]))]
    scc_order = np.concatenate([np.nonzero(term)[0], nt_order])
    rank = np.empty(S, dtype=np.int64)
    rank[scc_order] = np.arange(S)
    pos = np.full(n, -1, dtype=np.int64)
    pos[roots] = np.arange(S)
    vrank = rank[pos[ids]]
    reordered = np.argsort(vrank, kind="stable").astype(np.int64)
    sizes = np.bincount(vrank, minlength=S)
    offsets = np.zeros(S + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    return reordered, offsets, int(len(t_roots)), ids


# ----------------------------------------------------------------------------
# numerics


def kinetic(n, indptr, rows, rates):
    """Column generator K: K[d, s] = rate(s -> d), K[s, s] = -sum of out-rates.
    rate = rates[2*i + (0 if up else 1)] for the flipped node i
    (kernels/solver.cu:33-44)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    diff = rows ^ cols
    node = np.log2(diff).astype(np.int64) if len(diff) else diff
    # exact bit index (log2 of a power of two is exact in double up to 2^52)
    up = rows > cols
    val = rates[2 * node + np.where(up, 0, 1)]
    out = np.bincount(cols, weights=val, minlength=n)
    K = sp.csc_matrix((np.r_[val, -out], (np.r_[rows, np.arange(n)], np.r_[cols, np.arange(n)])),
                      shape=(n, n))
    K.sum_duplicates()
    return K


def terminal_kernel(Kj):
    """Normalised null vector of an irreducible generator block (row 0 replaced
    by ones, right-hand side e_0)."""
    s = Kj.shape[0]
    if s == 1:
        return np.ones(1)
    if s <= 2000:
        A = Kj.toarray()
        A[0, :] = 1.0
        b = np.zeros(s)
        b[0] = 1.0
        pi = np.linalg.solve(A, b)
    else:
        A = sp.lil_matrix(Kj)
        A[0, :] = np.ones(s)
        b = np.zeros(s)
        b[0] = 1.0
        pi = sla.spsolve(sp.csc_matrix(A), b)
    return pi / pi.sum()


def terminal_kernel_det(Kj):
    """Markov-chain tree theorem form used by the reference: pi_i proportional to
    |det(K_j without row i and column i)| (solver.cu:137-157).  Small SCCs only."""
    A = Kj.toarray()
    s = A.shape[0]
    if s == 1:
        return np.ones(1)
    m = np.array([abs(np.linalg.det(np.delete(np.delete(A, i, 0), i, 1))) for i in range(s)])
    return m / m.sum()


def vertex_heights(n, indptr, rows, ids):
    """Height of every vertex's SCC (1 for terminal SCCs)."""
    roots, h = scc_heights(n, indptr, rows, ids)
    pos = np.full(n, -1, dtype=np.int64)
    pos[roots] = np.arange(len(roots))
    return h[pos[ids]]


def solve_by_levels(A, b, hv):
    """Solves A w = b for a generator block A over transient vertices listed in
    canonical order (heights `hv` non-increasing): an entry A[d, s] (s -> d)
    joins two SCCs only when height(s) > height(d), so A is block lower
    triangular by height level and, inside a level, block diagonal by SCC.
    Forward substitution level by level, one SuperLU factorisation per level
    (its fill stays inside each SCC).  Exact up to rounding, like one sparse LU
    of A, without the fill a global ordering leaves across SCCs."""
    A = sp.csr_matrix(A)
    m = len(b)
    w = np.zeros(m)
    if m == 0:
        return w
    cuts = np.flatnonzero(np.diff(hv)) + 1
    starts = np.r_[0, cuts]
    ends = np.r_[cuts, m]
    for a, e in zip(starts, ends):
        r = np.array(b[a:e], dtype=float)
        if a > 0:
            r -= A[a:e, :a] @ w[:a]
        D = A[a:e, a:e]
        if e - a == 1:
            w[a] = r[0] / D[0, 0]
        else:
            w[a:e] = sla.splu(D.tocsc()).solve(r)
    return w


class Solution:
    def __init__(self):
        self.final_state = None
        self.term_indptr = self.term_indices = self.term_data = None
        self.nonterm_indptr = self.nonterm_indices = self.nonterm_data = None


def solve(n, indptr, rows, rates, x0, graph, full=True):
    """Terminal kernels (R), absorption (L = [U | X], X = -U B N^-1) and the
    final vector x* = R (L x0).  With full=False only the final vector is made,
    through the single right-hand side N w = x0_nt (a = -B w)."""
    reordered, offsets, T, _ = graph
    K = kinetic(n, indptr, rows, rates)
    n_t = int(offsets[T])
    tv = reordered[:n_t]
    ntv = reordered[n_t:]
    sol = Solution()
    # terminal part
    pi = np.empty(n_t)
    for j in range(T):
        a, b = offsets[j], offsets[j + 1]
        V = reordered[a:b]
        pi[a:b] = terminal_kernel(K[V][:, V]) if b - a > 1 else 1.0
    sol.term_indptr = offsets[: T + 1].copy()
    sol.term_indices = tv.copy()
    sol.term_data = pi
    scc_of_t = np.repeat(np.arange(T), np.diff(offsets[: T + 1]))
    if len(ntv) == 0:
        y = np.bincount(scc_of_t, weights=x0[tv], minlength=T)
        sol.nonterm_indptr = sol.term_indptr.copy()
        sol.nonterm_indices = tv.copy()
        sol.nonterm_data = np.ones(n_t)
    else:
        Nm = K[ntv][:, ntv].tocsc()
        B = K[tv][:, ntv].tocsc()
        if full:
            U = sp.csr_matrix((np.ones(n_t), (scc_of_t, np.arange(n_t))), shape=(T, n_t))
            A = -(U @ B)  # T x n_nt
            X = sla.spsolve(Nm.T.tocsc(), A.T.toarray() if T > 1 else A.T.toarray().ravel())
            X = np.asarray(X).reshape(len(ntv), T).T if T > 1 else np.asarray(X).reshape(1, -1)
            # X[j, c] is zero exactly when c cannot reach terminal SCC j; the sparse LU
            # can leave rounding noise there, the reference's triangular solves do not
            cols_ = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
            At = sp.csr_matrix((np.ones(len(rows), dtype=np.int8), (rows, cols_)), shape=(n, n))
            for j in range(T):
                seen = np.zeros(n, dtype=bool)
                seen[reordered[offsets[j]:offsets[j + 1]]] = True
                frontier = seen.copy()
                while frontier.any():
                    nxt = (At.T @ frontier.astype(np.int8)) > 0
                    new = nxt & ~seen
                    seen |= new
                    frontier = new
                X[j, ~seen[ntv]] = 0.0
            ip = [0]
            ind, dat = [], []
            for j in range(T):
                a, b = offsets[j], offsets[j + 1]
                nz = np.nonzero(X[j] != 0.0)[0]
                ind.append(reordered[a:b])
                dat.append(np.ones(b - a))
                ind.append(ntv[nz])
                dat.append(X[j, nz])
                ip.append(ip[-1] + (b - a) + len(nz))
            sol.nonterm_indptr = np.array(ip, dtype=np.int64)
            sol.nonterm_indices = np.concatenate(ind)
            sol.nonterm_data = np.concatenate(dat)
            y = X @ x0[ntv] + np.bincount(scc_of_t, weights=x0[tv], minlength=T)
        else:
            hv = vertex_heights(n, indptr, rows, graph[3])
            w = solve_by_levels(Nm, x0[ntv], hv[ntv])
            a_ = -(B @ w)
            y = np.bincount(scc_of_t, weights=x0[tv] + a_, minlength=T)
    xs = np.zeros(n)
    xs[tv] = pi * y[scc_of_t]
    sol.final_state = xs
    sol.y = y
    return sol


def reachable(n, indptr, rows, x0):
    """States reachable from supp(x0) (breadth-first over the CSC table)."""
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    A = sp.csr_matrix((np.ones(len(rows), dtype=np.float32), (rows, cols)), shape=(n, n))
    reach = x0 > 0
    frontier = reach.copy()
    while frontier.any():
        nxt = (A @ frontier.astype(np.float32)) > 0
        new = nxt & ~reach
        reach |= new
        frontier = new
    return reach


def solve_final(n, indptr, rows, rates, x0, graph):
    """Final vector only, on the part reachable from supp(x0): unreachable states
    end at 0 and unreachable terminal SCCs receive nothing, so the transient
    system shrinks to the reachable transient states; with a single reachable
    terminal SCC all the mass ends there and no transient solve is needed.
    Same result as solve(..., full=False) (pinned by the tests on small models)."""
    reordered, offsets, T, _ = graph
    reach = reachable(n, indptr, rows, x0)
    n_t = int(offsets[T])
    K = kinetic(n, indptr, rows, rates)
    xs = np.zeros(n)
    term_j = [j for j in range(T) if reach[reordered[offsets[j]:offsets[j + 1]]].any()]
    pis = {}
    for j in term_j:
        V = reordered[offsets[j]:offsets[j + 1]]
        pis[j] = terminal_kernel(K[V][:, V]) if len(V) > 1 else np.ones(1)
    if len(term_j) == 1:
        y = {term_j[0]: float(x0[reach].sum())}
    else:
        ntv = reordered[n_t:]
        ntv = ntv[reach[ntv]]
        tv = reordered[:n_t]
        scc_of_t = np.repeat(np.arange(T), np.diff(offsets[: T + 1]))
        if len(ntv):
            hv = vertex_heights(n, indptr, rows, graph[3])
            w = solve_by_levels(K[ntv][:, ntv], x0[ntv], hv[ntv])
            a_ = -(K[tv][:, ntv] @ w)
        else:
            a_ = np.zeros(n_t)
        yy = np.bincount(scc_of_t, weights=x0[tv] + a_, minlength=T)
        y = {j: yy[j] for j in term_j}
    for j in term_j:
        V = reordered[offsets[j]:offsets[j + 1]]
        xs[V] = pis[j] * y[j]
    return xs


# ----------------------------------------------------------------------------
# end to end


class Pipeline:
    """model -> rates -> x0 -> table -> graph -> solution, as the reference's
    test.cu drives it."""

    def __init__(self, model: Model, rates, x0, full=True):
        self.model = model
        self.rates = np.asarray(rates, dtype=np.float64)
        self.x0 = np.asarray(x0, dtype=np.float64)
        self.n = 1 << model.n_nodes
        self.indptr, self.rows, self.cols = transition_table(model, self.rates)
        self.graph = transition_graph(self.n, self.indptr, self.rows)
        self.sol = solve(self.n, self.indptr, self.rows, self.rates, self.x0, self.graph, full=full)

    @property
    def reordered_vertices(self):
        return self.graph[0]

    @property
    def sccs_offsets(self):
        return self.graph[1]

    @property
    def terminals_count(self):
        return self.graph[2]

    @property
    def final_state(self):
        return self.sol.final_state


def nonzero_final(x):
    nz = np.nonzero(x)[0]
    return nz, x[nz]
