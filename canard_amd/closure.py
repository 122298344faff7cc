"""Python handle of the learned closure (include/cnd.h, cnd_net_*): a
three-layer 3x3x3 convolution network evaluated and trained with the BF16
matrix-core kernels of libcnd."""
import ctypes as C

import numpy as np

from . import _lib


class ClosureError(RuntimeError):
    pass


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Net:
    """Network for an n^3 periodic coarse grid (n a multiple of 8) with room
    for `capacity` training pairs on the device."""

    def __init__(self, n, capacity=0, seed=0):
        self.L = _lib.load()
        self.n, self.V = n, n ** 3
        h = C.c_void_p()
        rc = self.L.cnd_net_create(n, capacity, seed, C.byref(h))
        if rc != 0:
            raise ClosureError(f"cnd_net_create: {self.L.cnd_net_error(None).decode()} ({rc})")
        self.h = h

    def _check(self, rc, what):
        if rc < 0:
            raise ClosureError(f"{what}: {self.L.cnd_net_error(self.h).decode()} ({rc})")
        return rc

    def close(self):
        if self.h:
            self.L.cnd_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparam(self):
        return int(self.L.cnd_net_param_count())

    def params(self):
        w = np.empty(self.nparam, dtype=np.float32)
        self._check(self.L.cnd_net_get_params(self.h, _fp(w)), "cnd_net_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.L.cnd_net_set_params(self.h, _fp(w)), "cnd_net_set_params")

    def set_norm(self, in_mean, in_scale, out_scale):
        a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (in_mean, in_scale, out_scale))
        self._check(self.L.cnd_net_set_norm(self.h, _dp(a), _dp(b), _dp(c)), "cnd_net_set_norm")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((self.V, 5), dtype=np.float32)
        self._check(self.L.cnd_net_forward_host(self.h, _fp(x), _fp(y)), "cnd_net_forward_host")
        return y

    def grad(self, x, gy):
        x = np.ascontiguousarray(x, dtype=np.float32)
        gy = np.ascontiguousarray(gy, dtype=np.float32)
        gx = np.empty((self.V, 5), dtype=np.float32)
        gp = np.empty(self.nparam, dtype=np.float32)
        self._check(self.L.cnd_net_grad_host(self.h, _fp(x), _fp(gy), _fp(gx), _fp(gp)), "cnd_net_grad_host")
        return gx, gp

    def add_pair(self, coarse, bc, target, bt):
        return self._check(self.L.cnd_net_add_pair(self.h, coarse.h, bc, target.h, bt), "cnd_net_add_pair")

    def add_pair_host(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        return self._check(self.L.cnd_net_add_pair_host(self.h, _fp(x), _fp(t)), "cnd_net_add_pair_host")

    @property
    def pairs(self):
        return int(self.L.cnd_net_pairs(self.h))

    def train(self, iters, lr=1e-3, seed=0):
        loss = C.c_double()
        self._check(self.L.cnd_net_train(self.h, iters, lr, seed, C.byref(loss)), "cnd_net_train")
        return loss.value

    def eval(self, first=0, count=None):
        loss = C.c_double()
        count = self.pairs - first if count is None else count
        self._check(self.L.cnd_net_eval(self.h, first, count, C.byref(loss)), "cnd_net_eval")
        return loss.value

    def apply(self, case, b=0):
        self._check(self.L.cnd_net_apply(self.h, case.h, b), "cnd_net_apply")

    def apply_noalloc(self, case, b=0):
        """cnd_net_apply with the library's fixed device scratch (n <= 64)."""
        self._check(self.L.cnd_net_apply_noalloc(self.h, case.h, b), "cnd_net_apply_noalloc")


def box_filter(fine, bf, coarse, bc):
    """State of coarse block bc = top-hat filter of fine block bf (include/cnd.h)."""
    L = _lib.load()
    rc = L.cnd_box_filter(fine.h, bf, coarse.h, bc)
    if rc != 0:
        raise ClosureError(f"cnd_box_filter failed ({rc})")
