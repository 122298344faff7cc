"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker.  The product
package never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")


class Params(C.Structure):
    _fields_ = [("fltk", C.c_double), ("fltrbc", C.c_double), ("reoo", C.c_double),
                ("tempoo", C.c_double), ("amach1", C.c_double), ("amach2", C.c_double),
                ("amach3", C.c_double), ("timf", C.c_double), ("nsmf", C.c_int),
                ("cfl", C.c_double), ("viscous", C.c_int), ("sponge", C.c_int),
                ("szco", C.c_double), ("szth0", C.c_double), ("szth1", C.c_double),
                ("doml0", C.c_double), ("doml1", C.c_double), ("quirk_send2", C.c_int)]


class BlockDesc(C.Structure):
    _fields_ = [("lxi", C.c_int), ("let", C.c_int), ("lze", C.c_int),
                ("nbc", (C.c_int * 2) * 3), ("mcd", (C.c_int * 2) * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(_LIB)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(Params), P(BlockDesc), C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_set_grid.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_get_metrics.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp]
        L.orc_set_state.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_get_state.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_set_clock.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 4
        L.orc_get_clock.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)] + [dp] * 4
        L.orc_run.argtypes = [C.c_void_p, C.c_int]
        L.orc_filter_phase.argtypes = [C.c_void_p]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int]
        L.orc_deriv_all.argtypes = [C.c_void_p, C.c_int, P(dp), P(dp)]
        L.orc_filter_all.argtypes = [C.c_void_p, C.c_int, P(dp)]
        L.orc_coefs.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_coefs.restype = C.c_int
        L.orc_penta.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_junctions.argtypes = [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_gcbc_mats.argtypes = [dp] * 5
        L.orc_stage_tail.argtypes = [C.c_void_p, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(**kw):
    p = Params()
    defaults = dict(fltk=0.49, fltrbc=0.9, reoo=16000.0, tempoo=300.0, amach1=0.0, amach2=0.0,
                    amach3=0.0, timf=1.0, nsmf=1, cfl=0.8, viscous=1, sponge=0, szco=0.0,
                    szth0=1.0, szth1=1.0, doml0=0.0, doml1=1.0, quirk_send2=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def make_blocks(blocks):
    """blocks: list of dicts with lxi, let, lze, nbc [[a,b]]*3, mcd [[a,b]]*3"""
    arr = (BlockDesc * len(blocks))()
    for i, b in enumerate(blocks):
        arr[i].lxi, arr[i].let, arr[i].lze = b["lxi"], b["let"], b["lze"]
        for d in range(3):
            for s in range(2):
                arr[i].nbc[d][s] = b["nbc"][d][s]
                arr[i].mcd[d][s] = b["mcd"][d][s]
    return arr


class Oracle:
    """One oracle case (all blocks in this process)."""

    def __init__(self, blocks, threads=1, **params):
        L = lib()
        L.orc_set_threads(int(threads))
        self.blocks = blocks
        self.params = make_params(**params)
        self._barr = make_blocks(blocks)
        h = L.orc_create(C.byref(self.params), self._barr, len(blocks))
        if not h:
            raise ValueError(L.orc_error().decode())
        self.h = h
        self.npt = [(b["lxi"] + 1) * (b["let"] + 1) * (b["lze"] + 1) for b in blocks]

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, xyz_list):
        for b, xyz in enumerate(xyz_list):
            a = np.ascontiguousarray(xyz, dtype=np.float64)
            assert a.size == 3 * self.npt[b]
            lib().orc_set_grid(self.h, b, _dp(a))

    def metrics(self, b):
        n = self.npt[b]
        xim, etm, zem, yaco = (np.empty(3 * n), np.empty(3 * n), np.empty(3 * n), np.empty(n))
        lib().orc_get_metrics(self.h, b, _dp(xim), _dp(etm), _dp(zem), _dp(yaco))
        return xim.reshape(3, n), etm.reshape(3, n), zem.reshape(3, n), yaco

    def set_state(self, q_list):
        for b, q in enumerate(q_list):
            a = np.ascontiguousarray(q, dtype=np.float64)
            assert a.size == 5 * self.npt[b]
            lib().orc_set_state(self.h, b, _dp(a))

    def state(self, b):
        q = np.empty(5 * self.npt[b])
        lib().orc_get_state(self.h, b, _dp(q))
        return q.reshape(5, self.npt[b])

    def set_clock(self, n=0, ndt=0, dt=0.0, dts=0.0, dte=0.0, timo=0.0):
        lib().orc_set_clock(self.h, n, ndt, dt, dts, dte, timo)

    def clock(self):
        n, ndt = C.c_int(), C.c_int()
        v = [C.c_double() for _ in range(4)]
        lib().orc_get_clock(self.h, C.byref(n), C.byref(ndt), *[C.byref(x) for x in v])
        return dict(n=n.value, ndt=ndt.value, dt=v[0].value, dts=v[1].value, dte=v[2].value,
                    timo=v[3].value)

    def run(self, nsteps):
        lib().orc_run(self.h, int(nsteps))

    def filter_phase(self):
        lib().orc_filter_phase(self.h)

    def stage(self, nk):
        lib().orc_stage(self.h, int(nk))

    def deriv(self, nn, fields):
        ins = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        outs = [np.empty_like(f) for f in ins]
        pin = (C.POINTER(C.c_double) * len(ins))(*[_dp(a) for a in ins])
        pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(a) for a in outs])
        lib().orc_deriv_all(self.h, nn, pin, pout)
        return outs

    def filter(self, nn, fields):
        fs = [np.array(f, dtype=np.float64, copy=True) for f in fields]
        pf = (C.POINTER(C.c_double) * len(fs))(*[_dp(a) for a in fs])
        lib().orc_filter_all(self.h, nn, pf)
        return fs

    def coefs(self):
        out = np.zeros(200)
        n = lib().orc_coefs(self.h, _dp(out), 200)
        o = out[:n]
        d = {}
        k = 0

        def take(name, cnt, shape=None):
            nonlocal k
            v = o[k:k + cnt].copy()
            k += cnt
            d[name] = v.reshape(shape) if shape else (v[0] if cnt == 1 else v)

        for nm in ("alphf", "betf", "fa", "fb", "fc"):
            take(nm, 1)
        take("albef0", 15, (3, 5))
        take("fbc", 15, (3, 5))
        take("pbci", 48, (2, 2, 12))
        take("pbco", 48, (2, 2, 12))
        take("pbcot", 4, (2, 2))
        take("sbci", 4)
        return d

    def penta(self, b, nn, nt):
        blk = self.blocks[b]
        n = [blk["lxi"], blk["let"], blk["lze"]][nn] + 1
        xl, xu = np.empty(2 * n), np.empty(3 * n)
        lib().orc_penta(self.h, b, nn, nt, _dp(xl), _dp(xu))
        return xl.reshape(n, 2), xu.reshape(n, 3)

    def stage_tail(self, umf=(0.0, 0.0, 0.0)):
        """Wall update, junction and surface averaging of the current state."""
        u = np.ascontiguousarray(umf, dtype=np.float64)
        lib().orc_stage_tail(self.h, _dp(u))

    def junctions(self, b):
        p8, l12 = (C.c_int * 8)(), (C.c_int * 12)()
        lib().orc_junctions(self.h, b, p8, l12)
        return list(p8), list(l12)


def gcbc_mats(cm, q, umf=(0.0, 0.0, 0.0)):
    """(xtq2r, xtr2q) 5x5 of the oracle at state q, unit normal cm."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (cm, q, umf)]
    q2r, r2q = np.empty(25), np.empty(25)
    lib().orc_gcbc_mats(_dp(a[0]), _dp(a[1]), _dp(a[2]), _dp(q2r), _dp(r2q))
    return q2r.reshape(5, 5), r2q.reshape(5, 5)
