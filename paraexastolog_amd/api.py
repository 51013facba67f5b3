"""Host-side mirror of the reference's C++ class API, over the C ABI of libesl.so.

The names, arguments and error behaviour follow cuda/src/*.h of the reference
so that code written against it reads the same:

    builder = model_builder()
    model = builder.construct_model("toy.bnet")         # model.h:41
    ctx = cu_context()                                    # cu_context.h:15
    r = transition_rates(model); r.generate_uniform()     # transition_rates.h:24-27
    table = transition_table(ctx, model, r.rates)         # transition_table.h:21
    table.construct_table()                               # transition_table.h:25
    g = transition_graph(ctx, table.rows, table.cols, table.indptr)   # transition_graph.h:24
    g.reorganize_vertices()                               # transition_graph.h:26
    st = initial_state(model.nodes, ["A"], [False], 1.0)  # initial_state.h:9
    s = solver(ctx, table, g, r, st)                      # solver.h:47
    s.solve()                                             # solver.h:62
    s.final_state, s.solution_term, s.solution_nonterm

Errors: input errors raise EslError (the reference throws std::runtime_error);
device errors raise EslError too (the reference exits the process).
All compute runs in libesl.so on the GPU; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import EslError, SolveOpts, SolveReport, check, lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _cstrs(names):
    arr = (C.c_char_p * max(len(names), 1))()
    for i, s in enumerate(names):
        arr[i] = s.encode()
    return arr


class cu_context:
    """One device and one HIP stream (replaces the reference's cuSPARSE/cuSOLVER handles)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().esl_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().esl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NodeList(list):
    """model.nodes: the node names, remembering their model."""

    def __init__(self, names, model):
        super().__init__(names)
        self.model = model


class model_t:
    def __init__(self, handle):
        self.handle = handle
        L = lib()
        n = L.esl_model_num_nodes(handle)
        self.nodes = NodeList([L.esl_model_node_name(handle, i).decode() for i in range(n)], self)

    def rule(self, i):
        """(inputs, truth table) of node i."""
        k = C.c_int()
        check(lib().esl_model_rule(self.handle, i, C.byref(k), None, None))
        ins = np.zeros(max(k.value, 1), dtype=np.int32)
        tab = np.zeros(1 << k.value, dtype=np.uint8)
        check(lib().esl_model_rule(self.handle, i, C.byref(k), _ip(ins),
                                   tab.ctypes.data_as(C.POINTER(C.c_uint8))))
        return ins[: k.value].tolist(), tab.astype(bool)

    def __del__(self):
        try:
            if self.handle:
                lib().esl_model_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class model_builder:
    def construct_model(self, file: str) -> model_t:
        h = C.c_void_p()
        check(lib().esl_model_load(file.encode(), C.byref(h)))
        return model_t(h)

    def parse_model(self, text: str) -> model_t:
        h = C.c_void_p()
        check(lib().esl_model_parse(text.encode(), C.byref(h)))
        return model_t(h)


class transition_rates:
    """rates[2i] = up rate of node i, rates[2i+1] = down rate (transition_rates.cpp:40-53)."""

    def __init__(self, model: model_t):
        self.model = model
        self.rates = None

    def _fill(self, kind, mean, std, up, down):
        up = list(up or [])
        down = list(down or [])
        un = _cstrs([u[0] for u in up])
        uv = np.array([float(u[1]) for u in up] or [0.0])
        dn = _cstrs([d[0] for d in down])
        dv = np.array([float(d[1]) for d in down] or [0.0])
        out = np.zeros(2 * len(self.model.nodes))
        check(lib().esl_rates_fill(self.model.handle, kind, mean, std, un, _dp(uv), len(up), dn, _dp(dv),
                                   len(down), _dp(out)))
        self.rates = out

    def generate_uniform(self, up_transition_rates=(), down_transition_rates=()):
        self._fill(0, 0.0, 0.0, up_transition_rates, down_transition_rates)

    def generate_normal(self, mean, std, up_transition_rates=(), down_transition_rates=()):
        self._fill(1, mean, std, up_transition_rates, down_transition_rates)


class initial_state:
    """x0 over the 2^N states (initial_state.cu:7-60)."""

    def __init__(self, node_names, fixed_node_names=(), fixed_node_values=(), fixed_probability=1.0):
        model = node_names.model if isinstance(node_names, NodeList) else node_names
        if not isinstance(model, model_t):
            raise EslError(3, "initial_state needs model.nodes (or the model)")
        fixed = list(fixed_node_names)
        vals = list(fixed_node_values)
        if len(fixed) != len(vals):
            raise EslError(3, "number of fixed node names not equal to number of fixed node values")
        v = np.array([1 if x else 0 for x in vals] or [0], dtype=np.uint8)
        self.state = np.zeros(1 << len(model.nodes))
        check(lib().esl_initial_state(model.handle, _cstrs(fixed), v.ctypes.data_as(C.POINTER(C.c_uint8)),
                                      len(fixed), float(fixed_probability), _dp(self.state)))


class transition_table:
    """CSC table over source states: indptr (n+1), rows = targets, cols = sources."""

    def __init__(self, context: cu_context, model: model_t, rates):
        self.context = context
        self.model = model
        self.rates_ = np.ascontiguousarray(rates, dtype=np.float64)
        self.handle = None
        self._cache = {}

    def construct_table(self):
        h = C.c_void_p()
        check(lib().esl_table_build(self.context.handle, self.model.handle, _dp(self.rates_), C.byref(h)),
              self.context.handle)
        self.handle = h
        self._cache = {}

    def sizes(self):
        s = np.zeros(7, dtype=np.int64)
        check(lib().esl_graph_query(self.handle, s.ctypes.data_as(C.POINTER(C.c_int64))))
        return s

    def _export(self):
        if "indptr" not in self._cache:
            n, nnz = (int(x) for x in self.sizes()[:2])
            ip = np.zeros(n + 1, dtype=np.int32)
            rows = np.zeros(max(nnz, 1), dtype=np.int32)
            cols = np.zeros(max(nnz, 1), dtype=np.int32)
            check(lib().esl_table_export(self.context.handle, self.handle, _ip(ip), _ip(rows), _ip(cols)),
                  self.context.handle)
            self._cache = {"indptr": _tag(ip, self), "rows": _tag(rows[:nnz], self), "cols": _tag(cols[:nnz], self)}
        return self._cache

    @property
    def indptr(self):
        return self._export()["indptr"]

    @property
    def rows(self):
        return self._export()["rows"]

    @property
    def cols(self):
        return self._export()["cols"]

    def stage_ms(self):
        ms = np.zeros(2)
        check(lib().esl_graph_stage_ms(self.handle, _dp(ms)))
        return ms

    def __del__(self):
        try:
            if self.handle:
                lib().esl_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class _TableArray(np.ndarray):
    """Table member that remembers its table (lets transition_graph take the
    reference's (rows, cols, indptr) arguments)."""


def _tag(arr, table):
    a = arr.view(_TableArray)
    a.table = table
    return a


class transition_graph:
    """SCCs, heights and the canonical order (transition_graph.h:21-27)."""

    def __init__(self, context: cu_context, rows_or_table, cols=None, indptr=None):
        self.context = context
        if isinstance(rows_or_table, transition_table):
            self.table = rows_or_table
        else:
            t = getattr(rows_or_table, "table", None) or getattr(indptr, "table", None)
            if t is None:
                raise EslError(3, "transition_graph needs the table's arrays (table.rows, table.cols, table.indptr)")
            self.table = t
        self.terminals_count = None
        self.sccs_offsets = None
        self.reordered_vertices = None
        self.scc_ids = None

    def reorganize_vertices(self):
        t = self.table
        check(lib().esl_graph_reorganize(self.context.handle, t.handle), self.context.handle)
        s = t.sizes()
        n, S = int(s[0]), int(s[2])
        self.reordered_vertices = np.zeros(n, dtype=np.int32)
        self.sccs_offsets = np.zeros(S + 1, dtype=np.int32)
        self.scc_ids = np.zeros(n, dtype=np.int32)
        tc = C.c_int32()
        check(lib().esl_graph_export(self.context.handle, t.handle, _ip(self.reordered_vertices),
                                     _ip(self.sccs_offsets), C.byref(tc), _ip(self.scc_ids)), self.context.handle)
        self.terminals_count = int(tc.value)
        self.max_height = int(s[6])

    def stage_counts(self):
        """Counters of the graph stage (include/esl.h esl_graph_stage_counts)."""
        a = np.zeros(3, dtype=np.int64)
        check(lib().esl_graph_stage_counts(self.table.handle, a.ctypes.data_as(C.POINTER(C.c_int64))))
        return {"mask_passes": int(a[0]), "scc_rounds": int(a[1]), "host_round_trips": int(a[2])}


@dataclass
class sparse_cs_matrix:
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray

    def nnz(self):
        return len(self.indices)

    def n(self):
        return len(self.indptr) - 1


def solve_opts(precision=0, tol=1e-13, max_iters=20000, restart=0, full=0, preconditioner=0, inner_rtol=None,
               exact_fallback=None):
    o = SolveOpts()
    lib().esl_solve_opts_default(C.byref(o))
    o.precision, o.tol, o.max_iters, o.restart, o.full = precision, tol, max_iters, restart, full
    o.preconditioner = preconditioner
    if inner_rtol is not None:
        o.inner_rtol = inner_rtol
    if exact_fallback is not None:
        o.exact_fallback = exact_fallback
    return o


class solver:
    """Stationary vector and terminal kernels (solver.h:10-67)."""

    def __init__(self, context: cu_context, table: transition_table, graph: transition_graph,
                 rates: transition_rates, state: initial_state, **opts):
        self.context = context
        self.table = table
        self.graph = graph
        self.rates = np.ascontiguousarray(rates.rates if isinstance(rates, transition_rates) else rates,
                                          dtype=np.float64)
        self.x0 = np.ascontiguousarray(state.state if isinstance(state, initial_state) else state,
                                       dtype=np.float64)
        self.opts = solve_opts(**opts)
        self.final_state = None
        self.report = None

    def solve(self):
        n = int(self.table.sizes()[0])
        if self.x0.shape[0] != n:
            raise EslError(3, f"initial state has {self.x0.shape[0]} entries, the model has {n} states")
        out = np.zeros(n)
        rep = SolveReport()
        check(lib().esl_solve(self.context.handle, self.table.handle, _dp(self.rates), _dp(self.x0), _dp(out),
                              C.byref(self.opts), C.byref(rep)), self.context.handle)
        self.final_state = out
        self.report = rep.as_dict()
        self._solution = None
        return out

    def _export_solution(self):
        if getattr(self, "_solution", None) is None:
            if not self.opts.full:
                self.opts.full = 1
                self.solve()
            T = self.graph.terminals_count
            n_t = int(self.graph.sccs_offsets[T])
            ti = np.zeros(T + 1, dtype=np.int32)
            tx = np.zeros(max(n_t, 1), dtype=np.int32)
            td = np.zeros(max(n_t, 1))
            nnz = C.c_int64()
            check(lib().esl_solution_export(self.context.handle, self.table.handle, None, None, None, C.byref(nnz),
                                            None, None, None), self.context.handle)
            ni = np.zeros(T + 1, dtype=np.int32)
            nx = np.zeros(max(nnz.value, 1), dtype=np.int32)
            nd = np.zeros(max(nnz.value, 1))
            check(lib().esl_solution_export(self.context.handle, self.table.handle, _ip(ti), _ip(tx), _dp(td), None,
                                            _ip(ni), _ip(nx), _dp(nd)), self.context.handle)
            self._solution = (sparse_cs_matrix(ti, tx[:n_t], td[:n_t]),
                              sparse_cs_matrix(ni, nx[: nnz.value], nd[: nnz.value]))
        return self._solution

    @property
    def solution_term(self):
        """R: CSC n x T -- indptr = terminal offsets, indices = terminal vertices, data = pi."""
        return self._export_solution()[0]

    @property
    def solution_nonterm(self):
        """L: CSR T x n -- per terminal SCC its vertices (1) then the transient vertices
        with a nonzero absorption probability, in canonical order."""
        return self._export_solution()[1]

    def print_final_state(self, model_nodes, file=None):
        """Non-zero states as node names joined by ' --- ', '<nil>' for state 0,
        then the value with 6 significant digits (solver.cu:336-372)."""
        lines = []
        for s in np.nonzero(self.final_state)[0]:
            names = [model_nodes[i] for i in range(len(model_nodes)) if (int(s) >> i) & 1]
            lines.append(f"{' --- '.join(names) if names else '<nil>'} {self.final_state[s]:g}")
        text = "\n".join(lines) + ("\n" if lines else "")
        if file is not None:
            file.write(text)
        else:
            print(text, end="")
        return text


class solve_plan:
    """Rate-independent part of a solve (one graph, one initial state), reused
    for any number of rate sets of the same zero pattern (include/esl.h
    esl_plan_*).  Device buffers come from paraexastolog_amd.surrogate.DeviceArray."""

    def __init__(self, context: cu_context, table: transition_table, d_x0):
        self.context = context
        self.table = table
        h = C.c_void_p()
        check(lib().esl_plan_create(context.handle, table.handle, d_x0, C.byref(h)), context.handle)
        self.handle = h

    def info(self):
        a = np.zeros(8, dtype=np.int64)
        check(lib().esl_plan_query(self.handle, a.ctypes.data_as(C.POINTER(C.c_int64))))
        keys = ["unknowns", "terminal_unknowns", "ambiguous", "terminals_reached", "terminal_segments", "nnz", "n_t",
                "tuned"]
        return dict(zip(keys, a.tolist()))

    def solve(self, n_sets, d_rates, d_final, **opts):
        o = solve_opts(**{k: v for k, v in opts.items() if k not in ("graph", "warm")})
        o.graph = int(opts.get("graph", 0))
        warm = opts.get("warm")
        if warm is not None:
            o.warm_pi = C.cast(warm, C.POINTER(C.c_double))
        rep = SolveReport()
        check(lib().esl_plan_solve(self.context.handle, self.handle, n_sets, d_rates, d_final, C.byref(o),
                                   C.byref(rep)), self.context.handle)
        return rep.as_dict()

    def solve_noalloc(self, n_sets, d_rates, d_final, **opts):
        """As solve(), but every per-run buffer is a slice of the library's static
        device scratch (esl_plan_solve_noalloc: no device allocation; 64 MiB)."""
        o = solve_opts(**{k: v for k, v in opts.items() if k not in ("graph", "warm")})
        warm = opts.get("warm")
        if warm is not None:
            o.warm_pi = C.cast(warm, C.POINTER(C.c_double))
        rep = SolveReport()
        check(lib().esl_plan_solve_noalloc(self.context.handle, self.handle, n_sets, d_rates, d_final, C.byref(o),
                                           C.byref(rep)), self.context.handle)
        return rep.as_dict()

    def gather_terminal(self, n_sets, d_final, d_out):
        """d_out[n_sets][n_t] = final vectors on the terminal states (warm-start layout)."""
        check(lib().esl_plan_gather_terminal(self.context.handle, self.handle, n_sets, d_final, d_out),
              self.context.handle)

    def set_results(self, n_sets):
        """Per-set outcome of the last solve: (residual, iterations, method) arrays,
        method 0 Krylov, 1 FP64 Krylov finish, 2 exact dense LU, -1 missed."""
        res = np.zeros(n_sets)
        its = np.zeros(n_sets, dtype=np.int32)
        meth = np.zeros(n_sets, dtype=np.int32)
        check(lib().esl_plan_set_results(self.handle, n_sets, res.ctypes.data_as(C.POINTER(C.c_double)),
                                         its.ctypes.data_as(C.POINTER(C.c_int32)),
                                         meth.ctypes.data_as(C.POINTER(C.c_int32))))
        return res, its, meth

    def set_schedule(self, schedule, restart=32):
        """Pins a schedule (0 per-set LDS, 1 per-set global, 2 device-wide, 3 per-set LDS with
        the basis in global memory; -1 forgets)."""
        check(lib().esl_plan_set_schedule(self.handle, schedule, restart))

    def autotune(self, n_sets, d_rates, d_final, **opts):
        """Picks the fastest schedule for this batch (include/esl.h esl_plan_autotune);
        later solves with restart=0 use it.  Returns (schedule, restart, ms)."""
        o = solve_opts(**{k: v for k, v in opts.items() if k != "graph"})
        o.graph = int(opts.get("graph", 0))
        ch = (C.c_int32 * 2)()
        ms = C.c_double()
        check(lib().esl_plan_autotune(self.context.handle, self.handle, n_sets, d_rates, d_final, C.byref(o), ch,
                                      C.byref(ms)), self.context.handle)
        return int(ch[0]), int(ch[1]), float(ms.value)

    @staticmethod
    def solve_many(context, plans, n_sets, d_rates, d_finals, **opts):
        """Several plans in one call (include/esl.h esl_plan_solve_many): the
        per-set solves of all plans run concurrently.  Returns one report per plan."""
        k = len(plans)
        o = solve_opts(**{key: v for key, v in opts.items() if key != "graph"})
        o.graph = int(opts.get("graph", 0))
        reps = (SolveReport * k)()
        arr = lambda vals: (C.c_void_p * k)(*[C.c_void_p(getattr(v, "value", v)) for v in vals])  # noqa: E731
        check(lib().esl_plan_solve_many(context.handle, k, (C.c_void_p * k)(*[p.handle for p in plans]),
                                        (C.c_int * k)(*n_sets), arr(d_rates), arr(d_finals), C.byref(o), reps),
              context.handle)
        return [r.as_dict() for r in reps]

    def close(self):
        if self.handle:
            lib().esl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
