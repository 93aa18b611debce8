"""Python handle of the device MLP (nb_mlp_*): BF16 MFMA forward/backward, Adam on the GPU."""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import lib


def _fp(a):
    return a.ctypes.data_as(L._fp)


class MLP:
    def __init__(self, d_in, hidden=128, layers=3, d_out=3, seed=0, device=0):
        self._m = C.c_void_p()
        rc = lib.nb_mlp_create(device, d_in, hidden, layers, d_out, seed, C.byref(self._m))
        if rc != L.NB_OK:
            raise L.NBodyError(rc, "nb_mlp_create failed: " + lib.nb_last_error(None).decode())
        self.d_in, self.hidden, self.layers, self.d_out = d_in, hidden, layers, d_out
        n = C.c_int64()
        self._check(lib.nb_mlp_param_count(self._m, C.byref(n)))
        self.n_params = n.value

    def _check(self, rc):
        if rc != L.NB_OK:
            raise L.NBodyError(rc, lib.nb_mlp_last_error(self._m).decode())

    def close(self):
        if self._m:
            lib.nb_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(lib.nb_mlp_get_params(self._m, _fp(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(lib.nb_mlp_set_params(self._m, _fp(p), p.size))

    def unpack(self, p):
        """Flat parameters -> [(W [in x out], b [out]), ...]."""
        dims = [self.d_in] + [self.hidden] * self.layers + [self.d_out]
        out, off = [], 0
        for i, o in zip(dims[:-1], dims[1:]):
            W = p[off:off + i * o].reshape(i, o)
            off += i * o
            b = p[off:off + o]
            off += o
            out.append((W, b))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.d_out), np.float32)
        self._check(lib.nb_mlp_forward(self._m, _fp(x), x.shape[0], _fp(y)))
        return y

    def infer(self, x):
        """Fused one-kernel inference (the surrogate's path); equal to forward() bit for bit."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.d_out), np.float32)
        self._check(lib.nb_mlp_infer(self._m, _fp(x), x.shape[0], _fp(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        self._check(lib.nb_mlp_grad(self._m, _fp(x), _fp(t), x.shape[0], _fp(g),
                                    C.byref(loss)))
        return g, loss.value

    def train(self, x, t, steps, batch=4096, lr=1e-3):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        self._check(lib.nb_mlp_train(self._m, _fp(x), _fp(t), x.shape[0], steps, batch, lr,
                                     C.byref(loss)))
        return loss.value
