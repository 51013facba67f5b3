"""esl_plan_solve_noalloc (per-run buffers in the library's static device
scratch) against esl_plan_solve on the same plan, through the C ABI: two calls
in a row, the second with fewer rate sets and new rates, so that the second
call runs over what the first left in the scratch."""
import numpy as np
import pytest

from paraexastolog_amd import EslError, solve_plan
from paraexastolog_amd.surrogate import DeviceArray

from .test_gpu_batch import setup_net

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,seed,B1,B2,precision", [
    ("c1", 0, 96, 24, 1),   # N=20 bench network: per-set kernel with the FP32 basis in registers
    ("c2", 11, 32, 8, 1),   # N=16, 2048 unknowns: 8 rows per thread
    ("c0", 1, 40, 7, 0),    # N=10, FP64 per-set kernel
    ("c0", 10, 33, 5, 1),   # N=10, two terminal SCCs (10 + 1 states): absorption system and exit flows
])
def test_noalloc_matches_plan_solve(ctx, cfg, seed, B1, B2, precision):
    net, rates, m, t, g, x0 = setup_net(ctx, cfg, seed, B1 + B2)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    plan = solve_plan(ctx, t, dx0.ptr)
    print(f"noalloc case {cfg} seed {seed} B {B1}/{B2} precision {precision} info {plan.info()}", flush=True)
    for B, rs in ((B1, rates[:B1]), (B2, rates[B1:])):
        dr = DeviceArray.from_host(ctx, np.ascontiguousarray(rs))
        d_ref = DeviceArray(ctx, (B, n), np.float64)
        d_out = DeviceArray(ctx, (B, n), np.float64)
        # poison the output: every entry must be written by the call
        d_out.upload(np.full((B, n), np.nan))
        rep_ref = plan.solve(B, dr.ptr, d_ref.ptr, precision=precision, tol=1e-12)
        rep = plan.solve_noalloc(B, dr.ptr, d_out.ptr, precision=precision, tol=1e-12)
        ref, got = d_ref.download(), d_out.download()
        assert np.isfinite(got).all()
        l1 = np.abs(got - ref).sum(axis=1)
        assert l1.max() <= 1e-12, (B, l1.max())
        assert np.abs(got.sum(axis=1) - 1.0).max() <= 1e-11
        assert rep["residual"] <= 1e-12 and rep_ref["residual"] <= 1e-12
        assert rep["unknowns"] == rep_ref["unknowns"]


def test_noalloc_rejects_batches_beyond_the_scratch(ctx):
    net, rates, m, t, g, x0 = setup_net(ctx, "c1", 0, 1)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    plan = solve_plan(ctx, t, dx0.ptr)
    B = 1024  # about 350 MB of per-run buffers at restart 32: beyond the 64 MiB scratch
    dr = DeviceArray(ctx, (B, 2 * net.N), np.float64)
    d_out = DeviceArray(ctx, (1, n), np.float64)  # never written: the size check comes first
    with pytest.raises(EslError) as e:
        plan.solve_noalloc(B, dr.ptr, d_out.ptr, precision=1, tol=1e-12)
    assert "scratch" in str(e.value) and "ARG" in str(e.value)


def test_noalloc_reports_missed_sets_without_exact_fallback(ctx):
    # the no-allocation solve has no exact fallback (its dense matrix would need
    # an allocation): a forced Krylov failure is ESL_E_NUMERIC with every set marked
    B = 3
    net, rates, m, t, g, x0 = setup_net(ctx, "c0", 10, B)
    n = 1 << net.N
    dx0 = DeviceArray.from_host(ctx, x0)
    plan = solve_plan(ctx, t, dx0.ptr)
    dr = DeviceArray.from_host(ctx, rates)
    d_out = DeviceArray(ctx, (B, n), np.float64)
    with pytest.raises(EslError) as e:
        plan.solve_noalloc(B, dr.ptr, d_out.ptr, precision=1, tol=1e-12, max_iters=1)
    assert "NUMERIC" in str(e.value) and "no exact fallback" in str(e.value)
    res, its, meth = plan.set_results(B)
    assert (meth == -1).all()
    rep = plan.solve_noalloc(B, dr.ptr, d_out.ptr, precision=1, tol=1e-12)
    res, its, meth = plan.set_results(B)
    assert (meth >= 0).all() and res.max() <= 1e-12 and rep["exact_solves"] == 0
