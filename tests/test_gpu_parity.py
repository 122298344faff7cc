"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on
the same inputs.  Tolerances: single sweeps and metrics 1e-13 relative to the
field's max (the device follows the oracle's evaluation order; only math
library ulps differ); whole steps 1e-12 per conserved variable (max-norm
relative)."""
import numpy as np
import pytest

from canard_amd import cases
from helpers import oracle_for, rel_err, tgv_case

pytestmark = pytest.mark.gpu


def _case(blocks, grids, states, params, **kw):
    from canard_amd.solver import Case
    c = Case(blocks, params, **kw)
    c.set_grids(grids)
    c.set_states(states)
    c.set_clock()
    return c


def two_block_x(n=(20, 16), let=16, lze=12, code=35):
    """Blocks 0 and 1 joined on xi faces (BC code), periodic in x through each
    other, periodic in y and z."""
    b0 = dict(lxi=n[0], let=let, lze=lze, nbc=[[code, code], [45, 45], [45, 45]], mcd=[[1, 1], [0, 0], [0, 0]])
    b1 = dict(lxi=n[1], let=let, lze=lze, nbc=[[code, code], [45, 45], [45, 45]], mcd=[[0, 0], [1, 1], [1, 1]])
    h = 2 * np.pi / (n[0] + n[1])
    g0 = cases.gridtoy(n[0], let, lze, doml0=0.0, doml1=n[0] * h, domh=np.pi, span=2 * np.pi)
    g1 = cases.gridtoy(n[1], let, lze, doml0=-n[0] * h, doml1=(n[0] + n[1]) * h, domh=np.pi, span=2 * np.pi)
    return [b0, b1], [g0, g1]


SWEEP_CASES = {
    "periodic16": lambda: ([cases.periodic_block(16)], [cases.gridtoy(16, 16, 16)]),
    # long pencils: several segments per pencil in every direction
    "periodic_100x60x40": lambda: ([dict(lxi=100, let=60, lze=40, nbc=[[45, 45]] * 3, mcd=[[0, 0]] * 3)],
                                   [cases.gridtoy(100, 60, 40)]),
    "mixed_bc_80x70x50": lambda: ([dict(lxi=80, let=70, lze=50, nbc=[[10, 10], [45, 45], [25, 20]],
                                        mcd=[[-1, -1], [0, 0], [-1, -1]])], [cases.gridtoy(80, 70, 50)]),
    "two_block_60_50": lambda: two_block_x(n=(60, 50), let=40, lze=36),
    "mixed_bc": lambda: ([dict(lxi=24, let=20, lze=16, nbc=[[10, 10], [45, 45], [25, 20]],
                               mcd=[[-1, -1], [0, 0], [-1, -1]])], [cases.gridtoy(24, 20, 16)]),
    "two_block": lambda: two_block_x(),
}


@pytest.fixture(params=["segmented", "tiled"])
def sweep_mode(request, monkeypatch):
    """Both device sweep algorithms: register-resident segments (default) and
    the tiled sweep with forward values in scratch (CANARD_SWEEP=tiled)."""
    if request.param == "tiled":
        monkeypatch.setenv("CANARD_SWEEP", "tiled")
    else:
        monkeypatch.delenv("CANARD_SWEEP", raising=False)
    return request.param


@pytest.mark.parametrize("name", list(SWEEP_CASES))
def test_single_sweeps_match_oracle(name, sweep_mode):
    blocks, grids = SWEEP_CASES[name]()
    rng = np.random.default_rng(11)
    npt = [(b["lxi"] + 1) * (b["let"] + 1) * (b["lze"] + 1) for b in blocks]
    fields = [np.sin(g[0]) * np.cos(2 * g[1]) + 0.3 * rng.standard_normal(n) for g, n in zip(grids, npt)]
    states = [cases.uniform_state(n) for n in npt]
    params = dict(reoo=1000.0)
    print(f"[sweep] case={name} mode={sweep_mode}", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params, quirk_send2=0)
    for d in range(3):
        got = c.deriv(d, fields)
        want = o.deriv(d, fields)
        for g, w in zip(got, want):
            assert np.max(np.abs(g - w)) <= 1e-13 * np.max(np.abs(w)), (name, "deriv", d)
        got = c.filter_sweep(d, fields)
        want = o.filter(d, fields)
        for g, w in zip(got, want):
            assert np.max(np.abs(g - w)) <= 1e-13 * np.max(np.abs(w)), (name, "filter", d)
    c.close()
    o.close()


def test_metrics_match_oracle():
    # periodic (45) faces between the two blocks carry the coordinate jump
    blocks, grids = two_block_x(code=45)
    # shear the grids so every metric term is non-trivial
    sheared = []
    for g in grids:
        x, y, z = g.copy()
        sheared.append(np.stack([x + 0.05 * np.sin(y), y + 0.03 * np.cos(x), z + 0.02 * np.sin(x + y)]))
    states = [cases.uniform_state(g.shape[1]) for g in sheared]
    print("[metrics] two_block sheared", flush=True)
    c = _case(blocks, sheared, states, dict(reoo=1000.0))
    o = oracle_for(blocks, sheared, states, dict(reoo=1000.0))
    for b in range(2):
        for got, want in zip(c.metrics(b), o.metrics(b)):
            assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    c.close()
    o.close()


@pytest.mark.parametrize("n,steps,pert", [(16, 12, 0.0), (24, 12, 1e-3), (48, 6, 1e-3)])
def test_tgv_steps_match_oracle(n, steps, pert):
    blocks, grids, states, params = tgv_case(n, pert=pert)
    print(f"[steps] tgv n={n} steps={steps}", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(steps)
    o.run(steps)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    cg, co = c.clock(), o.clock()
    assert cg["n"] == co["n"] == steps
    assert abs(cg["timo"] - co["timo"]) <= 1e-14 * max(co["timo"], 1.0)
    assert abs(cg["dte"] - co["dte"]) <= 1e-14 * co["dte"]
    c.close()
    o.close()


def test_eight_block_periodic_tgv_matches_oracle():
    """A 2 x 2 x 2 arrangement of 16^3-cell blocks tiling the periodic TGV box
    (z falls with zeta as in gridtoy), every interface joining two different
    blocks (BC 35 inside the box, BC 45 across its periodic wrap): corners are
    shared by eight blocks and edges by four, so the corner and edge
    averaging of every stage acts on every block edge (and corners lie on
    several averaged edges of their block).  8 steps against the oracle,
    1e-12 per variable; a second context gives bit-identical states."""
    n, h = 16, np.pi / 16
    idx = lambda i, j, k: i + 2 * j + 4 * k
    blocks, grids, states = [], [], []
    kk, jj, ii = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    for k in range(2):
        for j in range(2):
            for i in range(2):
                # faces across the box's periodic wrap are BC 45 (coordinate jump)
                nbc = [[45 if c == 0 else 35, 45 if c == 1 else 35] for c in (i, j, k)]
                blocks.append(dict(lxi=n, let=n, lze=n, nbc=nbc,
                                   mcd=[[idx(1 - i, j, k)] * 2, [idx(i, 1 - j, k)] * 2, [idx(i, j, 1 - k)] * 2]))
                xyz = np.stack([(ii.ravel() + i * n) * h, (jj.ravel() + j * n) * h, -(kk.ravel() + k * n) * h])
                grids.append(np.ascontiguousarray(xyz, dtype=np.float64))
                states.append(cases.tgv_state(grids[-1], (n + 1,) * 3, mach=0.1))
    params = dict(fltk=0.49, fltrbc=0.9, reoo=1600.0 / 0.1, tempoo=300.0, cfl=0.8, viscous=1)
    print("[steps] eight-block periodic tgv", flush=True)
    c = _case(blocks, grids, states, params)
    c2 = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(8)
    c2.step(8)
    o.run(8)
    for b in range(8):
        err = rel_err(c.state(b), o.state(b))
        print("  block", b, "rel err per variable", err)
        assert np.all(err < 1e-12), (b, err)
        assert np.array_equal(c.state(b), c2.state(b)), b
    c.close()
    c2.close()
    o.close()


def test_noncubic_periodic_matches_oracle():
    blocks, grids, states, params = tgv_case(0, nxyz=(20, 16, 24), pert=1e-3)
    print("[steps] noncubic 20x16x24", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(8)
    o.run(8)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    c.close()
    o.close()


def test_multiblock_gcbc_sponge_matches_oracle(sweep_mode):
    """Two sheared blocks joined by a BC 35 interface, non-reflective outer
    faces, periodic z, sponge, uniform Mach 0.5 flow with a pressure pulse.
    The change from the initial state is ~1e-3 of the state, so its relative
    error shows rounding differences magnified 1e3 times: the tiled sweep
    follows the oracle's operation order (1e-11), the segmented sweep adds
    homogeneous corrections in another order (1e-10)."""
    tol_change = 1e-11 if sweep_mode == "tiled" else 1e-10
    sizes = [(24, 16, 14), (16, 16, 14)]
    blocks = cases.sheared_pair_blocks(sizes)
    grids = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
    states = [cases.pulse_state(g, mach=0.5, centres=[(20.0, 8.0, 7.0)], amp=1e-3, width=3.0) for g in grids]
    params = dict(reoo=200.0, amach1=0.5, nsmf=1, cfl=0.8, sponge=1, szco=0.5, szth0=6.0, szth1=6.0,
                  doml0=0.0, doml1=40.0)
    print("[steps] multiblock gcbc sponge, 30 steps", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(30)   # dt reaches its computed value dte at step 11 and is recomputed at 21
    o.run(30)
    for b in range(2):
        err = rel_err(c.state(b) - states[b], o.state(b) - states[b])
        print("  block", b, "rel err of the change", err)
        assert np.all(err < tol_change), err
        err = rel_err(c.state(b), o.state(b))
        assert np.all(err < 1e-12), err
    c.close()
    o.close()


@pytest.mark.parametrize("case", ["tgv32", "gcbc_sponge_pair"])
def test_200_steps_match_oracle(case):
    """Long horizon: 200 steps (dt recomputed 20 times) of a 32^3 TGV and of
    the sheared two-block GCBC + sponge case, state within 1e-12 of the
    oracle per conserved variable; the error of the change is printed."""
    if case == "tgv32":
        blocks, grids, states, params = tgv_case(32, pert=1e-3)
    else:
        sizes = [(24, 16, 14), (16, 16, 14)]
        blocks = cases.sheared_pair_blocks(sizes)
        grids = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
        states = [cases.pulse_state(g, mach=0.5, centres=[(20.0, 8.0, 7.0)], amp=1e-3, width=3.0) for g in grids]
        params = dict(reoo=200.0, amach1=0.5, nsmf=1, cfl=0.8, sponge=1, szco=0.5, szth0=6.0, szth1=6.0,
                      doml0=0.0, doml1=40.0)
    print(f"[long] {case}, 200 steps", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params, threads=8)
    c.step(200)
    o.run(200)
    for b in range(len(blocks)):
        err = rel_err(c.state(b), o.state(b))
        chg = rel_err(c.state(b) - states[b], o.state(b) - states[b])
        print(f"  block {b}: rel err {err}, of the change {chg}", flush=True)
        assert np.all(err < 1e-12), (b, err)
    cg, co = c.clock(), o.clock()
    assert cg["n"] == co["n"] == 200
    assert abs(cg["timo"] - co["timo"]) <= 1e-13 * co["timo"]
    c.close()
    o.close()


def test_inviscid_steps_match_oracle(sweep_mode):
    """Euler steps (viscous = 0): the fluxes are evaluated inside the flux
    sweeps instead of being assembled with the stresses.  A TGV block and
    the two-block GCBC + sponge case; tolerances as for the viscous cases."""
    blocks, grids, states, params = tgv_case(24, pert=1e-3)
    params = dict(params, viscous=0)
    print("[steps] inviscid tgv 24", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(6)
    o.run(6)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    c.close()
    o.close()
    sizes = [(24, 16, 14), (16, 16, 14)]
    blocks = cases.sheared_pair_blocks(sizes)
    grids = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
    states = [cases.pulse_state(g, mach=0.5, centres=[(20.0, 8.0, 7.0)], amp=1e-3, width=3.0) for g in grids]
    params = dict(reoo=200.0, amach1=0.5, nsmf=1, cfl=0.8, sponge=1, szco=0.5, szth0=6.0, szth1=6.0,
                  doml0=0.0, doml1=40.0, viscous=0)
    print("[steps] inviscid multiblock gcbc sponge", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params)
    c.step(6)
    o.run(6)
    tol_change = 1e-11 if sweep_mode == "tiled" else 1e-10
    for b in range(2):
        err = rel_err(c.state(b) - states[b], o.state(b) - states[b])
        print("  block", b, "rel err of the change", err)
        assert np.all(err < tol_change), err
        assert np.all(rel_err(c.state(b), o.state(b)) < 1e-12)
    c.close()
    o.close()


def test_ensemble_instances_match_separate_oracles():
    """Three independent instances of different size, Mach and Re in one
    context; each must equal its own oracle run."""
    specs = [(16, 0.1, 1600.0), (20, 0.2, 800.0), (14, 0.05, 3000.0)]
    blocks, grids, states, params = [], [], [], []
    for i, (n, m, re) in enumerate(specs):
        b, g, s, p = tgv_case(n, mach=m, re=re, pert=1e-3, seed=i)
        b[0]["inst"] = i
        b[0]["mcd"] = [[i, i]] * 3
        blocks += b; grids += g; states += s; params.append(p)
    print("[steps] ensemble of 3", flush=True)
    c = _case(blocks, grids, states, params)
    c.step(6)
    for i, (n, m, re) in enumerate(specs):
        b, g, s, p = tgv_case(n, mach=m, re=re, pert=1e-3, seed=i)
        o = oracle_for(b, g, s, p)
        o.run(6)
        err = rel_err(c.state(i), o.state(0))
        print("  instance", i, err)
        assert np.all(err < 1e-12), err
        assert c.clock(i)["n"] == 6
        o.close()
    c.close()


def test_graph_replay_equals_eager():
    blocks, grids, states, params = tgv_case(16, pert=1e-3)
    print("[graph] 16^3", flush=True)
    a = _case(blocks, grids, states, params)
    b = _case(blocks, grids, states, params)
    b.use_graph(True)
    for k in range(3):
        a.step(4)
        b.step(4)
        qa_, qb_ = a.state(0), b.state(0)     # eager work between replays
        assert np.array_equal(qa_, qb_), k
    assert a.clock() == b.clock()
    a.close()
    b.close()


def test_stage_by_stage_api_equals_step():
    blocks, grids, states, params = tgv_case(16, pert=1e-3)
    a = _case(blocks, grids, states, params)
    b = _case(blocks, grids, states, params)
    a.step(3)
    for _ in range(3):
        b.filter()
        for nk in range(1, 5):
            b.rk_stage(nk)
        b.end_step()
    assert np.array_equal(a.state(0), b.state(0))
    a.close()
    b.close()


@pytest.mark.parametrize("fuse_dir", ["xi", "zeta"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_update_bitwise_equals_separate(monkeypatch, precision, fuse_dir):
    """Viscous stages without characteristic faces or sponge fold the RK
    update into the last flux sweep, the xi one (default) or the zeta one
    (CANARD_FUSE_DIR=2); on a periodic TGV (several segments per pencil) the
    fused step gives bit for bit the state of the separate update kernel
    (CANARD_FUSE_UPDATE=0) in both precisions, and the FP64 one matches the
    oracle."""
    from canard_amd.solver import CND_FP64, CND_MIXED
    blocks, grids, states, params = tgv_case(0, nxyz=(24, 20, 40), pert=1e-3)
    prec = CND_FP64 if precision == "fp64" else CND_MIXED
    monkeypatch.setenv("CANARD_TUNE", "0")
    monkeypatch.setenv("CANARD_FUSE_DIR", "0" if fuse_dir == "xi" else "2")
    print(f"[fused update] {precision} {fuse_dir}", flush=True)
    monkeypatch.setenv("CANARD_FUSE_UPDATE", "0")
    sep = _case(blocks, grids, states, params, precision=prec)
    monkeypatch.setenv("CANARD_FUSE_UPDATE", "1")
    fus = _case(blocks, grids, states, params, precision=prec)
    sep.step(6)
    fus.step(6)
    assert np.array_equal(sep.state(0), fus.state(0))
    assert sep.clock() == fus.clock()
    if precision == "fp64":
        o = oracle_for(blocks, grids, states, params)
        o.run(6)
        err = rel_err(fus.state(0), o.state(0))
        print("  rel err per variable", err)
        assert np.all(err < 1e-12), err
        o.close()
    sep.close()
    fus.close()


def _side_by_side_case(case):
    if case == "tgv":
        return tgv_case(0, nxyz=(32, 28, 24), pert=1e-3)
    sizes = [(24, 16, 14), (16, 16, 14)]
    blocks = cases.sheared_pair_blocks(sizes)
    grids = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
    states = [cases.pulse_state(g, mach=0.5, centres=[(20.0, 8.0, 7.0)], amp=1e-3, width=3.0) for g in grids]
    params = dict(reoo=200.0, amach1=0.5, nsmf=1, cfl=0.8, sponge=1, szco=0.5, szth0=6.0, szth1=6.0,
                  doml0=0.0, doml1=40.0, viscous=0 if case == "gcbc_inviscid" else 1)
    return blocks, grids, states, params


@pytest.mark.parametrize("precision,case", [("fp64", "tgv"), ("mixed", "tgv"), ("fp64", "gcbc_sponge"),
                                            ("fp64", "gcbc_inviscid")])
def test_side_by_side_sweeps_bitwise_equal_serial(monkeypatch, precision, case):
    """Independent sweeps on a second stream (the time step beside the xi
    gradient sweep, eta beside zeta gradients, eta beside the other flux
    sweeps; CANARD_OVERLAP=1, the default from 4 M points) give bit for bit
    the states of the serial order, eagerly and replayed as a graph: on a
    periodic TGV (update fused into the xi flux sweep), on the sheared
    two-block case with characteristic faces and sponge (all three flux
    sweeps, then GCBC and the separate update), and on that case inviscid
    (fluxes evaluated in the sweeps)."""
    from canard_amd.solver import CND_FP64, CND_MIXED
    blocks, grids, states, params = _side_by_side_case(case)
    prec = CND_FP64 if precision == "fp64" else CND_MIXED
    print(f"[side by side] {precision} {case}", flush=True)
    monkeypatch.setenv("CANARD_TUNE", "0")      # the same (default) shapes in all three
    monkeypatch.setenv("CANARD_OVERLAP", "0")
    ser = _case(blocks, grids, states, params, precision=prec)
    monkeypatch.setenv("CANARD_OVERLAP", "1")
    par = _case(blocks, grids, states, params, precision=prec)
    gra = _case(blocks, grids, states, params, precision=prec)
    gra.use_graph(True)
    for k in range(3):
        for c in (ser, par, gra):
            c.step(4)
        for b in range(len(blocks)):
            assert np.array_equal(ser.state(b), par.state(b)), (k, b)
            assert np.array_equal(ser.state(b), gra.state(b)), (k, b)
    assert ser.clock() == par.clock() == gra.clock()
    for c in (ser, par, gra):
        c.close()
