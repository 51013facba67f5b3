"""Surrogate MLP: the FP64 CPU reference is pinned by finite differences; the
GPU forward/backward (BF16 MFMA, FP32 sums) is checked against it."""
import numpy as np
import pytest

from oracle import mlp as R

DIMS = [32, 64, 64, 64, 40]


def _params(dims, seed):
    rng = np.random.default_rng(seed)
    ps = []
    for a, b in zip(dims[:-1], dims[1:]):
        ps.append(rng.normal(0, np.sqrt(2.0 / a), a * b))
        ps.append(rng.normal(0, 0.1, b))
    return np.concatenate(ps)


def _data(rows, dims, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(np.log(0.1), np.log(10), (rows, dims[0]))
    T = rng.dirichlet(np.ones(dims[-1]) * 0.3, rows)
    return X, T


def test_reference_gradient_finite_differences():
    dims = [6, 5, 5, 4]
    p = _params(dims, 1)
    X, T = _data(7, dims, 2)
    _, loss, g = R.forward_backward(p, dims, X, T, round_bf16=False)
    eps = 1e-6
    rng = np.random.default_rng(3)
    for i in rng.choice(len(p), 25, replace=False):
        q = p.copy()
        q[i] += eps
        lp = R.forward_backward(q, dims, X, T, round_bf16=False)[1]
        q[i] -= 2 * eps
        lm = R.forward_backward(q, dims, X, T, round_bf16=False)[1]
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-7 * max(1.0, abs(g[i]))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265], dtype=np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-12


@pytest.mark.gpu
def test_gpu_forward_matches_reference(ctx):
    from paraexastolog_amd.surrogate import Surrogate

    s = Surrogate(ctx, DIMS[0], DIMS[-1], hidden=DIMS[1], hidden_layers=3, seed=5)
    p = _params(DIMS, 4).astype(np.float32)
    s.set_params(p)
    X, _ = _data(70, DIMS, 6)  # 70 rows: exercises partial 64-row tiles
    got = s.forward(X.astype(np.float32))
    ref, _, _ = R.forward_backward(p.astype(np.float64), DIMS, X.astype(np.float32), None)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=1e-6)
    np.testing.assert_allclose(got.sum(axis=1), 1.0, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("n_out", [3000, 4500])
def test_gpu_wide_softmax_matches_reference(ctx, n_out):
    # 3000 outputs: 16 logits per thread held in registers; 4500: read per pass
    from paraexastolog_amd.surrogate import Surrogate

    dims = [32, 64, 64, 64, n_out]
    s = Surrogate(ctx, 32, n_out, hidden=64, hidden_layers=3, seed=5)
    p = _params(dims, 14).astype(np.float32)
    s.set_params(p)
    X, T = _data(20, dims, 15)
    got = s.forward(X.astype(np.float32))
    ref, _, _ = R.forward_backward(p.astype(np.float64), dims, X.astype(np.float32), None)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=1e-7)
    g, loss = s.gradients(X.astype(np.float32), T.astype(np.float32))
    _, rloss, _ = R.forward_backward(p.astype(np.float64), dims, X.astype(np.float32), T.astype(np.float32))
    assert abs(loss - rloss) < 1e-4 * max(1.0, abs(rloss))


@pytest.mark.gpu
def test_gpu_gradients_match_reference(ctx):
    from paraexastolog_amd.surrogate import Surrogate

    s = Surrogate(ctx, DIMS[0], DIMS[-1], hidden=DIMS[1], hidden_layers=3, seed=5)
    p = _params(DIMS, 8).astype(np.float32)
    s.set_params(p)
    X, T = _data(96, DIMS, 9)
    g, loss = s.gradients(X.astype(np.float32), T.astype(np.float32))
    _, rloss, rg = R.forward_backward(p.astype(np.float64), DIMS, X.astype(np.float32), T.astype(np.float32))
    assert abs(loss - rloss) < 1e-4 * max(1.0, abs(rloss))
    # relative to the size of each layer's gradient
    off = 0
    for a, b in zip(DIMS[:-1], DIMS[1:]):
        for size in (a * b, b):
            gg, rr = g[off:off + size], rg[off:off + size]
            scale = np.abs(rr).max() + 1e-12
            assert np.abs(gg - rr).max() / scale < 2e-2, (off, np.abs(gg - rr).max() / scale)
            off += size


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [[27, 70, 70, 70, 1100], [32, 68, 68, 68, 1100]])
def test_gpu_gradients_split_k_and_ragged_tiles(ctx, dims):
    # 601 rows: the weight gradients reduce over K = 601 in two splits whose
    # second ends on an odd k; n_out = 1100: dX of the last layer reduces over
    # K = 1100 in four splits, the last one ragged; hidden 70 / 68 give partial
    # 64-row and 64-column tiles.  Widths 27 and 70 take the single-float loads
    # (odd K for the first layer), 32 and 68 the 16-byte loads.
    from paraexastolog_amd.surrogate import Surrogate

    s = Surrogate(ctx, dims[0], dims[-1], hidden=dims[1], hidden_layers=3, seed=5)
    p = _params(dims, 12).astype(np.float32)
    s.set_params(p)
    X, T = _data(601, dims, 13)
    g, loss = s.gradients(X.astype(np.float32), T.astype(np.float32))
    g2, loss2 = s.gradients(X.astype(np.float32), T.astype(np.float32))
    assert np.array_equal(g, g2) and loss == loss2  # splits are summed in a fixed order
    _, rloss, rg = R.forward_backward(p.astype(np.float64), dims, X.astype(np.float32), T.astype(np.float32))
    assert abs(loss - rloss) < 1e-4 * max(1.0, abs(rloss))
    off = 0
    for a, b in zip(dims[:-1], dims[1:]):
        for size in (a * b, b):
            gg, rr = g[off:off + size], rg[off:off + size]
            scale = np.abs(rr).max() + 1e-12
            assert np.abs(gg - rr).max() / scale < 1e-2, (off, np.abs(gg - rr).max() / scale)
            off += size


@pytest.mark.gpu
def test_gpu_training_reduces_loss(ctx):
    from paraexastolog_amd.surrogate import DeviceArray, Surrogate

    dims = [32, 256, 256, 256, 16]
    s = Surrogate(ctx, dims[0], dims[-1], seed=1)
    rng = np.random.default_rng(0)
    X = rng.uniform(np.log(0.1), np.log(10), (512, 32)).astype(np.float32)
    W = rng.normal(0, 0.3, (32, 16))
    logits = X @ W
    T = np.exp(logits - logits.max(axis=1, keepdims=True))
    T = (T / T.sum(axis=1, keepdims=True)).astype(np.float32)
    dx, dt = DeviceArray.from_host(ctx, X), DeviceArray.from_host(ctx, T)
    losses = [s.train_step_device(512, dx.ptr, dt.ptr, 1e-3) for _ in range(300)]
    ent = float(-(T * np.log(T)).sum(axis=1).mean())
    assert losses[-1] < losses[0]
    assert losses[-1] - ent < 0.25 * (losses[0] - ent)


@pytest.mark.gpu
def test_gpu_train_steps_equals_single_steps(ctx):
    # esl_mlp_train_steps enqueues the same kernels as repeated
    # esl_mlp_train_step calls; at a constant rate the parameters match bitwise
    from paraexastolog_amd.surrogate import DeviceArray, Surrogate

    dims = [32, 64, 64, 64, 40]
    rng = np.random.default_rng(4)
    X = rng.uniform(np.log(0.1), np.log(10), (6 * 128, 32)).astype(np.float32)
    T = rng.dirichlet(np.ones(40) * 0.3, 6 * 128).astype(np.float32)
    dx, dt = DeviceArray.from_host(ctx, X), DeviceArray.from_host(ctx, T)
    a = Surrogate(ctx, 32, 40, hidden=64, hidden_layers=3, seed=7)
    b = Surrogate(ctx, 32, 40, hidden=64, hidden_layers=3, seed=7)
    la = [a.train_step_device(128, dx.offset(j * 128 * 32), dt.offset(j * 128 * 40), 1e-3) for j in range(6)]
    lb = b.train_steps_device(6, 128, dx.ptr, dt.ptr, 1e-3, 1e-3)
    assert np.array_equal(a.params(), b.params())
    assert abs(lb - np.mean(la)) < 1e-5 * abs(lb)

