"""Mixed precision (precision=CND_MIXED): FP32 line solves with one FP64
residual refinement in every sweep, FP32 storage of the velocity/temperature
gradients; stresses, fluxes, state, metrics, RK accumulation and GCBC stay
FP64.  Pinned by the FP64 device path and the oracle.

Tolerances: a refined sweep is within 1e-12 of the field maximum of the FP64
result (one refinement of an FP32 solve of a well-conditioned banded system
leaves ~1e-14); whole steps within 1e-6 per conserved variable (max-norm
relative), the mixed-precision target."""
import numpy as np
import pytest

from canard_amd import cases
from helpers import oracle_for, rel_err, tgv_case

pytestmark = pytest.mark.gpu


def _case(blocks, grids, states, params, precision):
    from canard_amd.solver import Case
    c = Case(blocks, params, precision=precision)
    c.set_grids(grids)
    c.set_states(states)
    c.set_clock()
    return c


@pytest.mark.parametrize("name", ["periodic_100x60x40", "mixed_bc_80x70x50", "two_block_60_50"])
def test_mixed_sweeps_match_oracle(name):
    from canard_amd.solver import CND_MIXED
    from test_gpu_parity import SWEEP_CASES
    blocks, grids = SWEEP_CASES[name]()
    rng = np.random.default_rng(5)
    npt = [(b["lxi"] + 1) * (b["let"] + 1) * (b["lze"] + 1) for b in blocks]
    fields = [np.sin(g[0]) * np.cos(2 * g[1]) + 0.3 * rng.standard_normal(n) for g, n in zip(grids, npt)]
    states = [cases.uniform_state(n) for n in npt]
    params = dict(reoo=1000.0)
    print(f"[mixed sweep] case={name}", flush=True)
    c = _case(blocks, grids, states, params, CND_MIXED)
    o = oracle_for(blocks, grids, states, params, quirk_send2=0)
    for d in range(3):
        for kind, got, want in (("deriv", c.deriv(d, fields), o.deriv(d, fields)),
                                ("filter", c.filter_sweep(d, fields), o.filter(d, fields))):
            for g, w in zip(got, want):
                e = np.max(np.abs(g - w)) / np.max(np.abs(w))
                print(f"  {kind} d={d} rel err {e:.2e}")
                assert e <= 1e-12, (name, kind, d, e)
    c.close()
    o.close()


@pytest.mark.parametrize("viscous", [1, 0])
def test_mixed_steps_within_target_of_fp64(viscous):
    from canard_amd.solver import CND_FP64, CND_MIXED
    blocks, grids, states, params = tgv_case(48, pert=1e-3)
    params = dict(params, viscous=viscous)
    print(f"[mixed steps] tgv 48^3, 20 steps, viscous={viscous}", flush=True)
    a = _case(blocks, grids, states, params, CND_FP64)
    b = _case(blocks, grids, states, params, CND_MIXED)
    a.step(20)
    b.step(20)
    err = rel_err(b.state(0), a.state(0))
    print("  mixed vs fp64 rel err per variable", err)
    assert np.all(err < 1e-6), err
    assert a.clock()["n"] == b.clock()["n"] == 20
    a.close()
    b.close()


def test_mixed_rejects_long_pencils():
    """Mixed precision runs on the segmented sweeps only."""
    from canard_amd.solver import CND_MIXED, Case
    with pytest.raises(RuntimeError, match="segmented"):
        Case([dict(lxi=600, let=16, lze=16, nbc=[[45, 45]] * 3, mcd=[[0, 0]] * 3)], dict(reoo=100.0),
             precision=CND_MIXED)
