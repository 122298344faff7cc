"""Whole steps of every boundary and frame option the library accepts,
through the C ABI against the oracle: inviscid (BC 20) and viscous
isothermal (BC 25) walls, characteristic interfaces between curvilinear
blocks (BC 30, with the face exchange of gcbc_comm), sub-domain interfaces
(BC 40) and the accelerating frame (nsmf = 0).

Tolerances as for the other step tests: 1e-12 per conserved variable
(max-norm relative), and, where the flow changes by ~1e-3 of the state, the
change from the initial state within 1e-10 of the largest change (the
segmented sweeps add their homogeneous corrections in another order than
the sequential recurrence, which shows up magnified in the change)."""
import numpy as np
import pytest

from canard_amd import cases
from helpers import oracle_for, rel_err

pytestmark = pytest.mark.gpu

GAM = 1.4


def _case(blocks, grids, states, params):
    from canard_amd.solver import Case
    c = Case(blocks, params)
    c.set_grids(grids)
    c.set_states(states)
    c.set_clock()
    return c


def _compare(name, blocks, grids, states, params, steps, tol_change=1e-10):
    print(f"[bc] {name}: {len(blocks)} block(s), {steps} steps", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params, threads=16)
    c.step(steps)
    o.run(steps)
    qc = [c.state(b) for b in range(len(blocks))]
    qo = [o.state(b) for b in range(len(blocks))]
    big = np.max([np.max(np.abs(a - s), axis=1) for a, s in zip(qo, states)], axis=0)
    big[1:4] = np.max(big[1:4])
    for b in range(len(blocks)):
        e = rel_err(qc[b], qo[b])
        ec = np.max(np.abs((qc[b] - states[b]) - (qo[b] - states[b])), axis=1) / np.maximum(big, 1e-300)
        print(f"  block {b}: rel err {e}, change {ec}", flush=True)
        assert np.all(np.isfinite(qc[b]))
        assert np.all(e < 1e-12), (name, b, e)
        assert np.all(ec < tol_change), (name, b, ec)
    cg, co = c.clock(), o.clock()
    assert cg["n"] == co["n"] == steps
    assert abs(cg["timo"] - co["timo"]) <= 1e-13 * max(co["timo"], 1.0)
    c.close()
    o.close()
    return big


def _sheared_grid(lxi, let, lze, amp=0.05):
    """lxi x let x lze cells of unit spacing, x sheared by amp * lxi *
    sin(2 pi y / let) (periodic in y and z)."""
    g = cases.gridtoy(lxi, let, lze, doml0=0.0, doml1=float(lxi), domh=let / 2.0, span=float(lze))
    g[0] += amp * lxi * np.sin(2 * np.pi * (g[1] + let / 2.0) / let)
    return g


def _pulse(xyz, centre, eps=1e-3, width=3.0, u=(0.0, 0.0, 0.0)):
    x, y, z = xyz
    r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
    g = eps * np.exp(-r2 / (2 * width ** 2))
    p = 1.0 / GAM + g / GAM
    rho = (GAM * p) ** (1.0 / GAM)
    e = p / (GAM - 1.0) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    return np.stack([rho, rho * u[0], rho * u[1], rho * u[2], e])


@pytest.mark.parametrize("wall,viscous", [(20, 0), (25, 1)])
def test_walls_match_oracle(wall, viscous):
    """Walls on the high xi face and the low zeta face of a sheared block
    (the GCBC wall branch, mo_gcbc.f90:285-308, and wall_condition_update,
    :424-459), non-reflective faces opposite them, periodic eta; a pressure
    pulse next to the corner the two walls share reaches both within the
    run, so the reflected waves, the corner's shared weights and the wall
    state update all act; the viscous wall also holds the stream at rest.
    40 steps: dt reaches its computed value dte."""
    lxi, let, lze = 28, 16, 20
    blk = dict(lxi=lxi, let=let, lze=lze, nbc=[[10, wall], [45, 45], [wall, 10]], mcd=[[-1, -1], [0, 0], [-1, -1]])
    g = _sheared_grid(lxi, let, lze)
    # a Mach 0.2 stream along eta, tangential to both walls, so that every
    # momentum component is of the size of the flow (the relative error of a
    # fluid at rest would measure rounding against the pulse's 1e-4)
    q = _pulse(g, centre=(lxi - 6.0, 0.0, -lze / 2.0 + 6.0), u=(0.0, 0.2, 0.0))
    prm = dict(reoo=500.0, viscous=viscous, cfl=0.8)
    _compare(f"wall {wall} viscous={viscous}", [blk], [g], [q], prm, 40)


def test_viscous_walls_in_an_accelerating_frame():
    """nsmf = 0 (mo_physics.f90:150-168): the frame accelerates along eta
    from rest to Mach 0.2 over timf, so the fluxes carry the ramping frame
    velocity and the viscous walls (BC 25 on both xi faces) hold the fluid at
    minus that velocity; 40 steps cover most of the ramp."""
    lxi, let, lze = 24, 16, 12
    blk = dict(lxi=lxi, let=let, lze=lze, nbc=[[25, 25], [45, 45], [45, 45]], mcd=[[-1, -1], [0, 0], [0, 0]])
    g = cases.gridtoy(lxi, let, lze, doml0=0.0, doml1=float(lxi), domh=let / 2.0, span=float(lze))
    q = _pulse(g, centre=(lxi / 2.0, 0.0, 0.0))
    prm = dict(reoo=400.0, viscous=1, amach2=0.2, nsmf=0, timf=30.0)
    _compare("accelerating frame, viscous walls", [blk], [g], [q], prm, 40)


def test_accelerating_frame_periodic_tgv():
    """nsmf = 0 on a periodic Taylor-Green block (frame ramping along xi to
    Mach 0.1 over timf): only the fluxes and time step see the frame."""
    n = 24
    blk = cases.periodic_block(n)
    g = cases.gridtoy(n, n, n)
    q = cases.tgv_state(g, (n + 1,) * 3, mach=0.1, pert=1e-3)
    prm = dict(reoo=1600.0 / 0.1, viscous=1, amach1=0.1, nsmf=0, timf=0.5)
    _compare("accelerating frame, periodic TGV", [blk], [g], [q], prm, 40)


@pytest.mark.parametrize("viscous", [0, 1])
def test_curvilinear_interface_matches_oracle(viscous):
    """Two sheared blocks joined on their xi faces by a BC 30 interface:
    characteristic conditions on both sides (gcbc_setup, gcbc_comm's
    exchange of the face values, the upwind choice of each wave in
    gcbc_update, mo_gcbc.f90:133-338) and the BC 30 surface averaging;
    non-reflective outer faces, periodic zeta; Mach 0.3 flow carries a
    pulse across the interface within the 40 steps."""
    sizes = [(20, 16, 12), (16, 16, 12)]
    blocks = cases.sheared_pair_blocks(sizes)
    for b in blocks:
        b["nbc"][0] = [30 if c == cases.INTER_STRAIGHT else c for c in b["nbc"][0]]
    grids = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
    states = [_pulse(g, centre=(14.0, 8.0, 6.0), u=(0.3, 0.0, 0.0)) for g in grids]
    prm = dict(reoo=300.0, viscous=viscous, amach1=0.3, nsmf=1)
    _compare(f"BC 30 interface viscous={viscous}", blocks, grids, states, prm, 40)


def test_subdomain_interface_matches_oracle():
    """Two blocks joined by a BC 40 sub-domain interface (quasi-disjoint
    closure rows with the partner's points, no averaging; the reference's
    split of one block over ranks), periodic through each other's outer xi
    faces (BC 45) and in eta and zeta: a Taylor-Green flow across the
    split."""
    n0, n1, let, lze = 20, 16, 16, 14
    b0 = dict(lxi=n0, let=let, lze=lze, nbc=[[45, 40], [45, 45], [45, 45]], mcd=[[1, 1], [0, 0], [0, 0]])
    b1 = dict(lxi=n1, let=let, lze=lze, nbc=[[40, 45], [45, 45], [45, 45]], mcd=[[0, 0], [1, 1], [1, 1]])
    h = 2 * np.pi / (n0 + n1)
    g0 = cases.gridtoy(n0, let, lze, doml0=0.0, doml1=n0 * h, domh=np.pi, span=2 * np.pi)
    g1 = cases.gridtoy(n1, let, lze, doml0=-n0 * h, doml1=(n0 + n1) * h, domh=np.pi, span=2 * np.pi)
    states = [cases.tgv_state(g, (b["lxi"] + 1, let + 1, lze + 1), mach=0.1) for g, b in zip((g0, g1), (b0, b1))]
    prm = dict(reoo=1600.0 / 0.1, viscous=1)
    _compare("BC 40 sub-domain interfaces", [b0, b1], [g0, g1], states, prm, 30)
