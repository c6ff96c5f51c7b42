"""The surrogate MLP (ctypes binding of rsg_mlp_*): GELU hidden layers, linear output,
forward / backward / Adam kernels on the device."""
import ctypes as C

import numpy as np

from . import _check, _err, lib


def _declare():
    L = lib()
    if getattr(L, "_mlp_declared", False):
        return L
    vp, i64, i32, d, cp = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_char_p
    L.rsg_mlp_create.restype, L.rsg_mlp_create.argtypes = i32, [vp, i32, vp, C.c_uint64, C.POINTER(vp), cp]
    L.rsg_mlp_destroy.restype, L.rsg_mlp_destroy.argtypes = None, [vp, vp]
    L.rsg_mlp_set_precision.restype, L.rsg_mlp_set_precision.argtypes = i32, [vp, i32]
    L.rsg_mlp_set_scaling.restype, L.rsg_mlp_set_scaling.argtypes = i32, [vp, vp, vp, vp, vp, vp, cp]
    L.rsg_mlp_set_range.restype, L.rsg_mlp_set_range.argtypes = i32, [vp, vp, vp]
    L.rsg_mlp_in_range.restype, L.rsg_mlp_in_range.argtypes = i32, [vp, vp]
    L.rsg_mlp_set_params.restype, L.rsg_mlp_set_params.argtypes = i32, [vp, vp, vp, vp, cp]
    L.rsg_mlp_get_params.restype, L.rsg_mlp_get_params.argtypes = i32, [vp, vp, vp, vp, cp]
    L.rsg_mlp_forward.restype, L.rsg_mlp_forward.argtypes = i32, [vp, vp, vp, i64, vp, cp]
    L.rsg_mlp_grad.restype, L.rsg_mlp_grad.argtypes = i32, [vp, vp, vp, vp, i64, vp, vp, C.POINTER(d), cp]
    L.rsg_mlp_train.restype = i32
    L.rsg_mlp_train.argtypes = [vp, vp, vp, vp, i64, i32, i32, d, C.c_uint64, C.POINTER(d), cp]
    L._mlp_declared = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class MLP:
    """dims = (d_in, h1, ..., d_out).  bf16=True: BF16 operands, FP32 accumulation."""

    def __init__(self, ctx, dims, seed=1, bf16=True):
        self.L = _declare()
        self.ctx, self.dims = ctx, tuple(int(d) for d in dims)
        self.nl = len(self.dims) - 1
        h = C.c_void_p()
        err = _err()
        da = (C.c_int * len(self.dims))(*self.dims)
        _check(self.L.rsg_mlp_create(ctx.h, self.nl, da, seed, C.byref(h), err), err, "rsg_mlp_create")
        self.h = h.value
        self.set_precision(bf16)

    def close(self):
        if getattr(self, "h", None):
            self.L.rsg_mlp_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, bf16):
        self.bf16 = bool(bf16)
        self.L.rsg_mlp_set_precision(self.h, 1 if bf16 else 0)

    def set_scaling(self, xmean=None, xstd=None, ymean=None, ystd=None):
        arrs = [None if a is None else _f32(a) for a in (xmean, xstd, ymean, ystd)]
        err = _err()
        _check(self.L.rsg_mlp_set_scaling(self.ctx.h, self.h, *[None if a is None else a.ctypes.data for a in arrs], err),
               err, "rsg_mlp_set_scaling")
        self._scaling = arrs

    def set_range(self, xmin=None, xmax=None):
        if xmin is None:
            self.L.rsg_mlp_set_range(self.h, None, None)
            return
        self._range = (_f32(xmin), _f32(xmax))
        self.L.rsg_mlp_set_range(self.h, self._range[0].ctypes.data, self._range[1].ctypes.data)

    def in_range(self, x):
        x = _f32(x)
        return bool(self.L.rsg_mlp_in_range(self.h, x.ctypes.data))

    def _shapes(self):
        return [(self.dims[l + 1], self.dims[l]) for l in range(self.nl)]

    def get_params(self):
        W = [np.zeros(s, np.float32) for s in self._shapes()]
        b = [np.zeros(s[0], np.float32) for s in self._shapes()]
        err = _err()
        _check(self.L.rsg_mlp_get_params(self.ctx.h, self.h, _ptrs(W), _ptrs(b), err), err, "rsg_mlp_get_params")
        return W, b

    def set_params(self, W, b):
        W = [_f32(w) for w in W]
        b = [_f32(x) for x in b]
        for w, x, s in zip(W, b, self._shapes()):
            assert w.shape == s and x.shape == (s[0],)
        err = _err()
        _check(self.L.rsg_mlp_set_params(self.ctx.h, self.h, _ptrs(W), _ptrs(b), err), err, "rsg_mlp_set_params")

    def forward(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        err = _err()
        _check(self.L.rsg_mlp_forward(self.ctx.h, self.h, x.ctypes.data, x.shape[0], y.ctypes.data, err),
               err, "rsg_mlp_forward")
        return y

    def grad(self, x, t):
        """Gradients of mean((y - t)^2) and the loss."""
        x, t = _f32(x), _f32(t)
        dW = [np.zeros(s, np.float32) for s in self._shapes()]
        db = [np.zeros(s[0], np.float32) for s in self._shapes()]
        loss = C.c_double()
        err = _err()
        _check(self.L.rsg_mlp_grad(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                   _ptrs(dW), _ptrs(db), C.byref(loss), err), err, "rsg_mlp_grad")
        return dW, db, loss.value

    def train(self, x, t, epochs=10, batch=512, lr=1e-3, seed=7):
        x, t = _f32(x), _f32(t)
        loss = C.c_double()
        err = _err()
        _check(self.L.rsg_mlp_train(self.ctx.h, self.h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                    epochs, batch, lr, seed, C.byref(loss), err), err, "rsg_mlp_train")
        return loss.value
