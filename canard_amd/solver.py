"""Python host mirror of the reference's driver surface over libcnd.

`Case` owns one cnd_ctx: the blocks of a (multi-block or ensemble) case with
their namelist parameters, grids and states, and exposes the time loop of
src/mo_canard_driver.f90:226-431 (step, filter, rk_stage) plus the numerics
sweeps of src/mo_numerics.f90 (deriv, filter_sweep).  Errors from the
library raise CanardError with the library's message.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CND_FP64, CND_MIXED, Block, Params  # noqa: F401


class CanardError(RuntimeError):
    pass


DEFAULT_PARAMS = dict(fltk=0.49, fltrbc=0.9, reoo=16000.0, tempoo=300.0, amach1=0.0, amach2=0.0, amach3=0.0,
                      timf=1.0, nsmf=1, cfl=0.8, viscous=1, sponge=0, szco=0.0, szth0=1.0, szth1=1.0,
                      doml0=0.0, doml1=1.0)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def shapes():
    """Compiled sweep shapes: list of dict(seg_len, pencils, run, xi)."""
    L = _lib.load()
    out = []
    while True:
        v = [C.c_int() for _ in range(4)]
        if L.cnd_shape_info(len(out), *[C.byref(x) for x in v]) != 0:
            return out
        out.append(dict(seg_len=v[0].value, pencils=v[1].value, run=v[2].value, xi=bool(v[3].value)))


class Case:
    """blocks: list of dicts (lxi, let, lze, nbc, mcd[, inst]); params: one
    dict per instance (missing keys take DEFAULT_PARAMS)."""

    def __init__(self, blocks, params, precision=CND_FP64):
        self.L = _lib.load()
        if isinstance(params, dict):
            params = [params]
        self.blocks = [dict(b) for b in blocks]
        for b in self.blocks:
            b.setdefault("inst", 0)
        self.params = [dict(DEFAULT_PARAMS, **p) for p in params]
        P = (Params * len(self.params))()
        for i, p in enumerate(self.params):
            for k, v in p.items():
                setattr(P[i], k, v)
        B = (Block * len(self.blocks))()
        for i, b in enumerate(self.blocks):
            B[i].lxi, B[i].let, B[i].lze, B[i].inst = b["lxi"], b["let"], b["lze"], b["inst"]
            for d in range(3):
                for s in range(2):
                    B[i].nbc[d][s] = b["nbc"][d][s]
                    B[i].mcd[d][s] = b["mcd"][d][s]
        h = C.c_void_p()
        rc = self.L.cnd_create(P, len(self.params), B, len(self.blocks), precision, C.byref(h))
        if rc != 0:
            raise CanardError(f"cnd_create: {self.L.cnd_last_error(None).decode()} ({rc})")
        self.h = h
        self.npt = [(b["lxi"] + 1) * (b["let"] + 1) * (b["lze"] + 1) for b in self.blocks]
        self.shape = [(b["lxi"] + 1, b["let"] + 1, b["lze"] + 1) for b in self.blocks]

    # -- plumbing -------------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            raise CanardError(f"{what}: {self.L.cnd_last_error(self.h).decode()} ({rc})")

    def close(self):
        if getattr(self, "h", None):
            self.L.cnd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- set-up ---------------------------------------------------------------
    def set_grids(self, grids):
        for b, xyz in enumerate(grids):
            a = np.ascontiguousarray(xyz, dtype=np.float64)
            assert a.size == 3 * self.npt[b], "grid size"
            self._check(self.L.cnd_set_grid(self.h, b, _dp(a)), "cnd_set_grid")
        self._check(self.L.cnd_init_metrics(self.h), "cnd_init_metrics")

    def make_grid(self, b, kind, *par):
        """Generate block b's grid on the device: kind 'toy' (lxi0, let0,
        lze0, doml0, doml1, domh, span), 'toy_multi' (lxi0, lxi1, lxi2, let0,
        lze0, doml0, doml1, domh, span, mb) or 'sheared' (x0, h, amp, lx, ly);
        call init_metrics once every block has its grid."""
        a = np.ascontiguousarray(par, dtype=np.float64)
        self._check(self.L.cnd_make_grid(self.h, b, _lib.GRIDS[kind], _dp(a), len(a)), "cnd_make_grid")

    def init_metrics(self):
        self._check(self.L.cnd_init_metrics(self.h), "cnd_init_metrics")

    def grid(self, b):
        xyz = np.empty(3 * self.npt[b])
        self._check(self.L.cnd_get_grid(self.h, b, _dp(xyz)), "cnd_get_grid")
        return xyz.reshape(3, self.npt[b])

    def metrics(self, b):
        n = self.npt[b]
        xim, etm, zem, yaco = np.empty(3 * n), np.empty(3 * n), np.empty(3 * n), np.empty(n)
        self._check(self.L.cnd_get_metrics(self.h, b, _dp(xim), _dp(etm), _dp(zem), _dp(yaco)), "cnd_get_metrics")
        return xim.reshape(3, n), etm.reshape(3, n), zem.reshape(3, n), yaco

    def set_states(self, states):
        for b, q in enumerate(states):
            a = np.ascontiguousarray(q, dtype=np.float64)
            assert a.size == 5 * self.npt[b], "state size"
            self._check(self.L.cnd_set_state(self.h, b, _dp(a)), "cnd_set_state")

    def state(self, b):
        q = np.empty(5 * self.npt[b])
        self._check(self.L.cnd_get_state(self.h, b, _dp(q)), "cnd_get_state")
        return q.reshape(5, self.npt[b])

    def set_clock(self, inst=0, n=0, ndt=0, dt=0.0, dts=0.0, dte=0.0, timo=0.0):
        self._check(self.L.cnd_set_clock(self.h, inst, n, ndt, dt, dts, dte, timo), "cnd_set_clock")

    def clock(self, inst=0):
        n, ndt = C.c_int(), C.c_int()
        v = [C.c_double() for _ in range(4)]
        self._check(self.L.cnd_get_clock(self.h, inst, C.byref(n), C.byref(ndt), *[C.byref(x) for x in v]),
                    "cnd_get_clock")
        return dict(n=n.value, ndt=ndt.value, dt=v[0].value, dts=v[1].value, dte=v[2].value, timo=v[3].value)

    # -- time stepping ----------------------------------------------------------
    def step(self, nsteps=1):
        self._check(self.L.cnd_step(self.h, int(nsteps)), "cnd_step")

    def filter(self):
        self._check(self.L.cnd_filter(self.h), "cnd_filter")

    def rk_stage(self, nk):
        self._check(self.L.cnd_rk_stage(self.h, int(nk)), "cnd_rk_stage")

    def end_step(self):
        self._check(self.L.cnd_end_step(self.h), "cnd_end_step")

    def record(self, on=True):
        """Accumulate time averages at the end of every step (cnd_record)."""
        self._check(self.L.cnd_record(self.h, int(bool(on))), "cnd_record")

    def record_read(self, b):
        """Output variables of block b from the recorded averages: (5, points)
        float32 -- density, three velocity components, pressure."""
        out = np.empty(5 * self.npt[b], dtype=np.float32)
        self._check(self.L.cnd_record_read(self.h, b, out.ctypes.data_as(C.POINTER(C.c_float))), "cnd_record_read")
        return out.reshape(5, self.npt[b])

    def record_reset(self):
        self._check(self.L.cnd_record_reset(self.h), "cnd_record_reset")

    # -- output (Tecplot files) ------------------------------------------------
    def output_schedule(self, tmax, ndata, inst=0):
        """Output k is due on the step whose interval holds k * tmax / ndata."""
        self._check(self.L.cnd_output_schedule(self.h, inst, float(tmax), int(ndata)), "cnd_output_schedule")

    def output_due(self, inst=0):
        """(due, ndati, time) of the step just taken."""
        due, nd, t = C.c_int(), C.c_int(), C.c_double()
        self._check(self.L.cnd_output_due(self.h, inst, C.byref(due), C.byref(nd), C.byref(t)), "cnd_output_due")
        return bool(due.value), nd.value, t.value

    def output_write(self, path, inst=0, wait=True):
        """Tecplot file of the recorded output; wait=False writes it from a
        host thread while the device goes on (join with output_wait)."""
        self._check(self.L.cnd_output_write(self.h, inst, str(path).encode(), int(not wait)), "cnd_output_write")

    def output_grid(self, path, inst=0):
        self._check(self.L.cnd_output_grid(self.h, inst, str(path).encode()), "cnd_output_grid")

    def output_wait(self):
        self._check(self.L.cnd_output_wait(self.h), "cnd_output_wait")

    def use_graph(self, on=True):
        self._check(self.L.cnd_use_graph(self.h, int(bool(on))), "cnd_use_graph")

    def synchronize(self):
        self._check(self.L.cnd_synchronize(self.h), "cnd_synchronize")

    def profile(self, on=True):
        self._check(self.L.cnd_profile(self.h, int(bool(on))), "cnd_profile")

    def profile_read(self):
        """{section: (milliseconds, kernel launches)} since profiling was enabled."""
        out = {}
        for i, name in enumerate(_lib.SECTIONS):
            ms, n = C.c_double(), C.c_longlong()
            self._check(self.L.cnd_profile_read(self.h, i, C.byref(ms), C.byref(n)), "cnd_profile_read")
            out[name] = (ms.value, n.value)
        return out

    def tune(self):
        """Re-run the sweep-shape autotuner now (it runs once by itself before
        the first step unless CANARD_TUNE=0)."""
        self._check(self.L.cnd_tune(self.h), "cnd_tune")

    def set_shape(self, kind, d, shape):
        """Pin the sweep shape of kind ('filter', 'grad', 'flux', 'field') along d
        (-1: tiled sweep); turns automatic tuning off."""
        self._check(self.L.cnd_set_shape(self.h, _lib.KINDS.index(kind), d, shape), "cnd_set_shape")

    def tuning(self):
        """{kind: [per direction dict(shape, seg_len, pencils, segments, ms)]}."""
        out = {}
        for k, name in enumerate(_lib.KINDS):
            rows = []
            for d in range(3):
                v = [C.c_int() for _ in range(4)]
                ms = C.c_double()
                self._check(self.L.cnd_tuning(self.h, k, d, *[C.byref(x) for x in v], C.byref(ms)), "cnd_tuning")
                rows.append(dict(shape=v[0].value, seg_len=v[1].value, pencils=v[2].value, segments=v[3].value,
                                 ms=ms.value))
            out[name] = rows
        return out

    def stream(self):
        return self.L.cnd_stream(self.h)

    @property
    def points(self):
        return int(self.L.cnd_points(self.h))

    @property
    def cells(self):
        return int(self.L.cnd_cells(self.h))

    # -- single sweeps -----------------------------------------------------------
    def deriv(self, direction, fields):
        ins = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        outs = [np.empty_like(f) for f in ins]
        pin = (C.POINTER(C.c_double) * len(ins))(*[_dp(a) for a in ins])
        pout = (C.POINTER(C.c_double) * len(outs))(*[_dp(a) for a in outs])
        self._check(self.L.cnd_deriv_sweep(self.h, int(direction), pin, pout), "cnd_deriv_sweep")
        return outs

    def filter_sweep(self, direction, fields):
        fs = [np.array(f, dtype=np.float64, copy=True) for f in fields]
        pf = (C.POINTER(C.c_double) * len(fs))(*[_dp(a) for a in fs])
        self._check(self.L.cnd_filter_sweep(self.h, int(direction), pf), "cnd_filter_sweep")
        return fs


def write_plt(path, title, names, zones, time=0.0, file_type=2):
    """Tecplot binary file (cnd_write_plt): zones is a list of (dims, data)
    with dims = (imax, jmax, kmax) and data a (len(names), points) array of
    FP32 values, i fastest."""
    L = _lib.load()
    arrs = [np.ascontiguousarray(d, dtype=np.float32) for _, d in zones]
    dims = (C.c_int * (3 * len(zones)))(*[int(v) for dz, _ in zones for v in dz])
    nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
    dp = (C.POINTER(C.c_float) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    rc = L.cnd_write_plt(str(path).encode(), title.encode(), int(file_type), len(names), nm, len(zones), dims,
                         float(time), dp)
    if rc != 0:
        raise CanardError(f"cnd_write_plt: {L.cnd_last_error(None).decode()} ({rc})")
