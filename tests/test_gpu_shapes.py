"""Every compiled sweep shape (segment length = lanes per pencil, pencils per
workgroup, staging run length) and the tiled sweep give the oracle's
results, and the autotuner's choice is one of them.  Tolerances as in
test_gpu_parity: sweeps 1e-13 of the field maximum, whole steps 1e-12 per
conserved variable."""
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


def _shapes_for(d):
    from canard_amd.solver import shapes
    return [-1] + [i for i, s in enumerate(shapes()) if s["xi"] == (d == 0)]


@pytest.mark.parametrize("nxyz", [(100, 60, 40), (99, 59, 41)])
def test_every_shape_sweeps_match_oracle(nxyz):
    """Every compiled shape against the oracle, on pencils of odd (101, 61,
    41) and even (100, 60, 42) point counts: the xi tile in LDS keeps its
    pencils n | 1 points apart."""
    lx, ly, lz = nxyz
    blocks = [dict(lxi=lx, let=ly, lze=lz, nbc=[[45, 45], [10, 10], [25, 20]], mcd=[[0, 0], [-1, -1], [-1, -1]])]
    grids = [cases.gridtoy(lx, ly, lz)]
    npt = (lx + 1) * (ly + 1) * (lz + 1)
    rng = np.random.default_rng(3)
    fields = [np.sin(grids[0][0]) * np.cos(2 * grids[0][1]) + 0.3 * rng.standard_normal(npt)]
    states = [cases.uniform_state(npt)]
    params = dict(reoo=1000.0)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params, quirk_send2=0)
    for d in range(3):
        want_d, want_f = o.deriv(d, fields)[0], o.filter(d, fields)[0]
        for sh in _shapes_for(d):
            print(f"[shape] d={d} shape={sh}", flush=True)
            c.set_shape("field", d, sh)
            c.set_shape("filter", d, sh)
            got = c.deriv(d, fields)[0]
            assert np.max(np.abs(got - want_d)) <= 1e-13 * np.max(np.abs(want_d)), (d, sh, "deriv")
            got = c.filter_sweep(d, fields)[0]
            assert np.max(np.abs(got - want_f)) <= 1e-13 * np.max(np.abs(want_f)), (d, sh, "filter")
    c.close()
    o.close()


@pytest.mark.parametrize("pick,nxyz", [("first", (40, 36, 34)), ("last", (40, 36, 34)), ("first", (41, 35, 33)),
                                       ("last", (41, 35, 33))])
def test_every_shape_steps_match_oracle(pick, nxyz):
    """A few TGV steps with every kind pinned to the first (or last) compiled
    shape that fits each direction, on odd and on even point counts."""
    blocks, grids, states, params = tgv_case(0, nxyz=nxyz, pert=1e-3)
    from canard_amd.solver import CanardError
    c = _case(blocks, grids, states, params)
    for d in range(3):
        cands = _shapes_for(d)[1:]
        for kind in ("filter", "grad", "flux"):
            for sh in (cands if pick == "first" else cands[::-1]):
                try:
                    c.set_shape(kind, d, sh)
                    break
                except CanardError:     # segments too short for this pencil length
                    continue
    print(f"[shape steps] {pick}: {c.tuning()}", flush=True)
    o = oracle_for(blocks, grids, states, params)
    c.step(6)
    o.run(6)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    c.close()
    o.close()


def test_tiled_xi_gradients_with_segmented_eta_zeta():
    """The xi gradient sweep on the tiled algorithm, which stores no
    primitives: the eta and zeta sweeps then read those of the pointwise
    primitives kernel."""
    blocks, grids, states, params = tgv_case(0, nxyz=(40, 36, 34), pert=1e-3)
    c = _case(blocks, grids, states, params)
    c.set_shape("grad", 0, -1)
    t = c.tuning()
    print("[tiled xi grad]", t["grad"], flush=True)
    assert t["grad"][0]["shape"] == -1 and t["grad"][1]["shape"] >= 0 and t["grad"][2]["shape"] >= 0
    o = oracle_for(blocks, grids, states, params)
    c.step(6)
    o.run(6)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    c.close()
    o.close()


@pytest.mark.parametrize("nxyz", [(600, 16, 14), (2111, 12, 12)])
def test_long_xi_pencils_fall_back_to_tiled(nxyz):
    """Pencils of 601 and 2112 points (the largest a block may have) along
    xi fit no segmented xi shape (528 at most): every xi sweep runs tiled,
    the primitives for the segmented eta and zeta gradient sweeps come from
    the pointwise kernel; the second case also has the fewest points a
    direction may have (13) along eta and zeta."""
    blocks, grids, states, params = tgv_case(0, nxyz=nxyz, pert=1e-3)
    c = _case(blocks, grids, states, params)
    t = c.tuning()
    print("[long xi]", {k: [r["shape"] for r in v] for k, v in t.items()}, flush=True)
    assert all(t[k][0]["shape"] == -1 for k in ("filter", "grad", "flux"))
    assert t["grad"][1]["shape"] >= 0 and t["grad"][2]["shape"] >= 0
    o = oracle_for(blocks, grids, states, params)
    c.step(5)
    o.run(5)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable", err)
    assert np.all(err < 1e-12), err
    c.close()
    o.close()


def test_autotuned_step_matches_oracle():
    blocks, grids, states, params = tgv_case(40, pert=1e-3)
    c = _case(blocks, grids, states, params)
    c.tune()
    t = c.tuning()
    print("[tuned]", t, flush=True)
    for kind in ("filter", "grad", "flux"):
        for d in range(3):
            assert t[kind][d]["ms"] > 0.0
    o = oracle_for(blocks, grids, states, params)
    c.step(6)
    o.run(6)
    err = rel_err(c.state(0), o.state(0))
    assert np.all(err < 1e-12), err
    c.close()
    o.close()


def test_ensemble_launch_bitwise_equals_separate_contexts(monkeypatch):
    """Instances of different sizes in one context (one launch per kernel over
    all of them) give bit-for-bit the states of one context per instance with
    the same shapes."""
    sizes, machs, res = [24, 40, 32], [0.1, 0.2, 0.15], [400.0, 1600.0, 800.0]
    blocks, grids, states, params = [], [], [], []
    for i, (n, m, re) in enumerate(zip(sizes, machs, res)):
        b, g, s, p = tgv_case(n, mach=m, re=re, pert=1e-3, seed=7 + i)
        b[0]["inst"] = i
        b[0]["mcd"] = [[i, i]] * 3       # its own periodic partner
        blocks += b; grids += g; states += s; params.append(p)
    from canard_amd.solver import Case

    monkeypatch.setenv("CANARD_TUNE", "0")
    ens = Case(blocks, params)
    ens.set_grids(grids); ens.set_states(states); ens.set_clock(inst=0); ens.set_clock(inst=1); ens.set_clock(inst=2)
    chosen = ens.tuning()

    def pin(c):
        for d in range(3):
            for kind in ("filter", "grad", "flux"):
                c.set_shape(kind, d, chosen[kind][d]["shape"])

    ens.step(5)
    for i in range(3):
        one = _case([dict(blocks[i], inst=0, mcd=[[0, 0]] * 3)], [grids[i]], [states[i]], params[i])
        pin(one)
        one.step(5)
        assert np.array_equal(one.state(0), ens.state(i)), i
        assert one.clock() == ens.clock(i)
        one.close()
    ens.close()
