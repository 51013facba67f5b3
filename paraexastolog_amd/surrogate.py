"""Surrogate MLP (log-rates -> terminal-state distribution) over the C ABI of
include/esl_mlp.h.  Training and inference run in libesl.so on the matrix
cores; this module only moves arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    """A float32/float64 device buffer owned through esl_device_alloc."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().esl_device_alloc(ctx.handle, max(self.nbytes, 1), C.byref(p)), ctx.handle)
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(ctx, arr.shape, arr.dtype)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        check(lib().esl_memcpy_h2d(self.ctx.handle, self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes),
              self.ctx.handle)

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().esl_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes),
              self.ctx.handle)
        return out

    def offset(self, elems):
        return C.c_void_p(self.ptr.value + elems * self.dtype.itemsize)

    def free(self):
        if self.ptr:
            lib().esl_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Surrogate:
    def __init__(self, ctx, n_in, n_out, hidden=256, hidden_layers=3, seed=0):
        self.ctx = ctx
        self.dims = [n_in] + [hidden] * hidden_layers + [n_out]
        h = C.c_void_p()
        check(lib().esl_mlp_create(ctx.handle, n_in, hidden, hidden_layers, n_out, seed, C.byref(h)), ctx.handle)
        self.handle = h
        self.n_params = int(lib().esl_mlp_num_params(h))

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        check(lib().esl_mlp_get_params(self.ctx.handle, self.handle, p.ctypes.data_as(C.POINTER(C.c_float))),
              self.ctx.handle)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().esl_mlp_set_params(self.ctx.handle, self.handle, p.ctypes.data_as(C.POINTER(C.c_float))),
              self.ctx.handle)

    def forward_device(self, rows, d_x, d_probs):
        check(lib().esl_mlp_forward(self.ctx.handle, self.handle, rows, d_x, d_probs), self.ctx.handle)

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        dx = DeviceArray.from_host(self.ctx, x)
        dp = DeviceArray(self.ctx, (x.shape[0], self.dims[-1]), np.float32)
        self.forward_device(x.shape[0], dx.ptr, dp.ptr)
        return dp.download()

    def predict_warm_device(self, rows, d_x, d_warm):
        """FP64 warm start [rows][n_out] on the device (esl_mlp_predict_warm)."""
        check(lib().esl_mlp_predict_warm(self.ctx.handle, self.handle, rows, d_x, d_warm), self.ctx.handle)

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = DeviceArray.from_host(self.ctx, x), DeviceArray.from_host(self.ctx, t)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        check(lib().esl_mlp_gradients(self.ctx.handle, self.handle, x.shape[0], dx.ptr, dt.ptr,
                                      g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)), self.ctx.handle)
        return g, float(loss.value)

    def train_step_device(self, rows, d_x, d_t, lr):
        loss = C.c_float()
        check(lib().esl_mlp_train_step(self.ctx.handle, self.handle, rows, d_x, d_t, C.c_float(lr), C.byref(loss)),
              self.ctx.handle)
        return float(loss.value)

    def train_steps_device(self, steps, rows, d_x, d_t, lr_first, lr_last):
        """`steps` Adam steps on consecutive minibatches of `rows` rows, one host
        synchronisation (esl_mlp_train_steps); returns the mean loss."""
        loss = C.c_float()
        check(lib().esl_mlp_train_steps(self.ctx.handle, self.handle, steps, rows, d_x, d_t, C.c_float(lr_first),
                                        C.c_float(lr_last), C.byref(loss)), self.ctx.handle)
        return float(loss.value)

    def __del__(self):
        try:
            if self.handle:
                lib().esl_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def log_rates(rates):
    """Network inputs: the 2N natural-log rates of each set, FP32."""
    return np.log(np.asarray(rates, dtype=np.float64)).astype(np.float32)


def terminal_targets(ctx, plan, rates, n, n_t, batch=1024, **opts):
    """Solves every rate set with `plan` and returns the final vectors on the
    terminal states in canonical order, [sets][n_t] FP64 -- the training targets
    (they sum to 1 per set: all mass ends on terminal states)."""
    rates = np.ascontiguousarray(rates, dtype=np.float64)
    S = rates.shape[0]
    batch = min(batch, S)
    out = np.empty((S, n_t))
    dr = DeviceArray(ctx, (batch, rates.shape[1]), np.float64)
    df = DeviceArray(ctx, (batch, n), np.float64)
    dt = DeviceArray(ctx, (batch, n_t), np.float64)
    for a in range(0, S, batch):
        k = min(S, a + batch) - a
        dr.upload(rates[a:a + k])
        plan.solve(k, dr.ptr, df.ptr, **opts)
        plan.gather_terminal(k, df.ptr, dt.ptr)
        out[a:a + k] = dt.download()[:k]
    return out


def train(sur, X, T, epochs, batch=512, lr=2e-3, lr_final=1e-4):
    """Adam on the mean cross-entropy over minibatches of consecutive rows (the
    rate sets are i.i.d., so no shuffling is needed); the learning rate decays
    geometrically from lr to lr_final.  Returns the mean loss of each epoch."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    rows, n_in = X.shape
    n_out = T.shape[1]
    batch = min(batch, rows)
    steps = rows // batch
    dX = DeviceArray.from_host(sur.ctx, X)
    dT = DeviceArray.from_host(sur.ctx, T)
    total = max(epochs * steps - 1, 1)
    hist = []
    k = 0
    for _ in range(epochs):
        # one call per epoch: the steps are enqueued without host round trips
        first = lr * (lr_final / lr) ** (k / total)
        last = lr * (lr_final / lr) ** ((k + steps - 1) / total)
        hist.append(sur.train_steps_device(steps, batch, dX.ptr, dT.ptr, first, last))
        k += steps
    return hist
