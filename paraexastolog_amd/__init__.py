"""paraexastolog_amd: MI355X-native (gfx950) stationary-distribution pipeline
for asynchronous Boolean networks -- .bnet model -> transition table ->
SCCs / toposort -> terminal kernels and absorption -> stationary vector.

All compute runs in libesl.so (hand-written HIP kernels); this package is the
host-side mirror of the reference's C++ class API over the library's C ABI.
"""
from ._lib import EslError, LIB_PATH, lib  # noqa: F401
from .api import (  # noqa: F401
    cu_context,
    initial_state,
    model_builder,
    model_t,
    solve_opts,
    solve_plan,
    solver,
    sparse_cs_matrix,
    transition_graph,
    transition_rates,
    transition_table,
)
