"""Shared helpers for the tests: matching oracle and product cases."""
import numpy as np

from canard_amd import cases


def tgv_case(n, mach=0.1, re=1600.0, pert=0.0, seed=20261202, nxyz=None):
    """Single triply periodic TGV block on gridtoy's [0,2pi]x[-pi,pi]x[pi,-pi]."""
    lxi, let, lze = nxyz or (n, n, n)
    blk = cases.periodic_block(n, lxi=lxi, let=let, lze=lze)
    xyz = cases.gridtoy(lxi, let, lze, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi)
    q = cases.tgv_state(xyz, (lxi + 1, let + 1, lze + 1), mach=mach, pert=pert, seed=seed)
    params = dict(fltk=0.49, fltrbc=0.9, reoo=re / mach, tempoo=300.0, cfl=0.8, viscous=1)
    return [blk], [xyz], [q], params


def rel_err(a, b):
    """Per-variable max-norm relative error max|a-b| / scale.  The scale of
    rho and rho E is their own max; the three momentum components share the
    max of the momentum vector's components (a component that is ~0 in the
    flow, e.g. rho w of a Taylor-Green vortex, is measured against the
    momentum it belongs to)."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    scale = np.max(np.abs(b), axis=1)
    if b.shape[0] == 5:
        scale[1:4] = np.max(scale[1:4])
    return np.max(np.abs(a - b), axis=1) / np.maximum(scale, 1e-300)


def oracle_for(blocks, grids, states, params, threads=4, **extra):
    import orc
    ob = [dict(lxi=b["lxi"], let=b["let"], lze=b["lze"], nbc=b["nbc"], mcd=b["mcd"]) for b in blocks]
    o = orc.Oracle(ob, threads=threads, **dict(params, **extra))
    o.set_grid(grids)
    o.set_state(states)
    o.set_clock()
    return o
