"""Oracle (test infrastructure only): FP64 forward/backward of the surrogate MLP.

Same network as paraexastolog_amd/csrc/mlp.hip: n_in -> H (GELU, tanh form)
-> ... -> n_out (softmax), mean cross-entropy.  Products optionally round their
operands to BF16 (round to nearest even) so the GPU's BF16-operand / FP32-sum
products can be compared tightly; the sums themselves are FP64 here.
"""
import numpy as np

C0 = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = C0 * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C0 * (1.0 + 3 * 0.044715 * z * z)


def unpack(params, dims):
    Ws, bs, off = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(params[off:off + a * b].reshape(a, b))
        off += a * b
        bs.append(params[off:off + b])
        off += b
    return Ws, bs


def forward_backward(params, dims, X, T=None, round_bf16=True):
    """Returns (probs, loss, grads) with grads in the packed parameter layout."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unpack(np.asarray(params, dtype=np.float64), dims)
    L = len(Ws)
    acts = [np.asarray(X, dtype=np.float64)]
    pres = []
    for l in range(L):
        z = r(acts[-1]) @ r(Ws[l]) + bs[l]
        pres.append(z)
        acts.append(gelu(z) if l < L - 1 else z)
    z = pres[-1]
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    if T is None:
        return p, None, None
    rows = X.shape[0]
    T = np.asarray(T, dtype=np.float64)
    logp = np.log(np.maximum(p, 1e-300))
    loss = -(T * logp).sum() / rows
    grads = []
    dz = (p - T) / rows
    gW = [None] * L
    gb = [None] * L
    for l in range(L - 1, -1, -1):
        gW[l] = r(acts[l]).T @ r(dz)
        gb[l] = dz.sum(axis=0)
        if l > 0:
            da = r(dz) @ r(Ws[l]).T
            dz = da * gelu_grad(pres[l - 1])
    for l in range(L):
        grads.append(gW[l].ravel())
        grads.append(gb[l])
    return p, loss, np.concatenate(grads)
