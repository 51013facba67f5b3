"""Call-order and argument errors of the C ABI on a device: each is reported as
a status (no exception or exit crosses the boundary, unlike the reference's
cu_context.cpp:7-32), nothing is launched for it, and the context keeps
working afterwards."""
import ctypes as C

import numpy as np
import pytest

from paraexastolog_amd import initial_state, model_builder, solve_plan, solver, synth, transition_graph, \
    transition_table
from paraexastolog_amd._lib import ESL_E_ARG, ESL_E_STATE, lib
from paraexastolog_amd.api import solve_opts
from paraexastolog_amd.surrogate import DeviceArray

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_call_order_and_argument_errors(ctx):
    L = lib()
    net, rng = synth.config("c0", 1)
    rates = synth.rate_sets(rng, net.N, 2)
    m = model_builder().parse_model(net.bnet())
    n = 1 << net.N
    t = transition_table(ctx, m, rates[0])
    t.construct_table()
    x0 = initial_state(m.nodes).state
    out = np.zeros(n)
    o = solve_opts()
    h = ctx.handle

    # a solve before the graph stage
    assert L.esl_solve(h, t.handle, _dp(rates[0]), _dp(x0), _dp(out), C.byref(o), None) == ESL_E_STATE
    assert b"reorganize" in L.esl_last_error(h)
    g = transition_graph(ctx, t)
    g.reorganize_vertices()
    # null arguments, a negative rate, another zero pattern, a bad precision
    assert L.esl_solve(h, t.handle, _dp(rates[0]), None, _dp(out), C.byref(o), None) == ESL_E_ARG
    assert L.esl_solve(h, None, _dp(rates[0]), _dp(x0), _dp(out), C.byref(o), None) == ESL_E_ARG
    bad = rates[0].copy()
    bad[1] = -1.0
    assert L.esl_solve(h, t.handle, _dp(bad), _dp(x0), _dp(out), C.byref(o), None) == ESL_E_ARG
    bad = rates[0].copy()
    bad[1] = 0.0
    assert L.esl_solve(h, t.handle, _dp(bad), _dp(x0), _dp(out), C.byref(o), None) == ESL_E_ARG
    o2 = solve_opts()
    o2.precision = 2
    assert L.esl_solve(h, t.handle, _dp(rates[0]), _dp(x0), _dp(out), C.byref(o2), None) == ESL_E_ARG
    # the solution export before a full solve
    nnz = C.c_int64()
    assert L.esl_solution_export(h, t.handle, None, None, None, C.byref(nnz), None, None, None) == ESL_E_STATE
    # a graph import that is not a permutation
    t2 = transition_table(ctx, m, rates[0])
    t2.construct_table()
    order = np.arange(n, dtype=np.int32)
    order[1] = 0
    offs = np.array([0, n], dtype=np.int32)
    assert L.esl_graph_import(h, t2.handle, order.ctypes.data_as(C.POINTER(C.c_int32)),
                              offs.ctypes.data_as(C.POINTER(C.c_int32)), 1, 1) == ESL_E_ARG

    # plans: no sets, full export, results of a batch of another size
    dx0 = DeviceArray.from_host(ctx, x0)
    dr = DeviceArray.from_host(ctx, rates)
    dout = DeviceArray(ctx, (2, n), np.float64)
    plan = solve_plan(ctx, t, dx0.ptr)
    res = np.zeros(2)
    assert L.esl_plan_set_results(plan.handle, 2, _dp(res), None, None) == ESL_E_STATE
    assert L.esl_plan_solve(h, plan.handle, 0, dr.ptr, dout.ptr, C.byref(o), None) == ESL_E_ARG
    of = solve_opts(full=1)
    assert L.esl_plan_solve(h, plan.handle, 2, dr.ptr, dout.ptr, C.byref(of), None) == ESL_E_ARG
    assert L.esl_solve_batch_device(h, t.handle, 2, dr.ptr, dx0.ptr, dout.ptr, C.byref(of), None) == ESL_E_ARG
    plan.solve(2, dr.ptr, dout.ptr, precision=0, tol=1e-13)
    assert L.esl_plan_set_results(plan.handle, 3, _dp(np.zeros(3)), None, None) == ESL_E_ARG

    # the context still solves: the plan's results and a fresh solve agree
    s = solver(ctx, t, g, rates[1], x0)
    s.solve()
    assert np.abs(dout.download()[1] - s.final_state).sum() <= 1e-12
