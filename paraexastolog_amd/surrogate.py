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

    def __del__(self):
        try:
            if self.handle:
                lib().esl_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
