"""Recording of intermediate results (mo_canard_driver.f90:373-425) against
a numpy restatement on the states the stage-by-stage API exposes: the
filtered state qo after cnd_filter and the new state qa after the step.
The device evaluates the same FP64 operations in the same order and rounds
to FP32 once, so the outputs are compared bit for bit."""
import numpy as np
import pytest

from helpers import tgv_case

pytestmark = pytest.mark.gpu

GM1 = 1.4 - 1.0


def _outputs(q, umf=(0.0, 0.0, 0.0)):
    """Density, velocity plus frame velocity, pressure of the conserved
    variables q (5, points), as the reference writes them (FP32)."""
    rr = 1.0 / q[0]
    out = np.empty(q.shape, dtype=np.float32)
    out[0] = q[0]
    for m in range(1, 4):
        out[m] = rr * q[m] + umf[m - 1]
    out[4] = GM1 * (q[4] - (0.5 * rr) * ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]))
    return out


def _case(blocks, grids, states, params):
    from canard_amd.solver import Case
    c = Case(blocks, params)
    c.set_grids(grids)
    c.set_states(states)
    c.set_clock()
    return c


def test_record_matches_restatement_and_graph_replay():
    blocks, grids, states, params = tgv_case(16, pert=1e-3)
    print("[record] 16^3 TGV, 6 steps", flush=True)
    c = _case(blocks, grids, states, params)
    c.record(True)
    qb, dtsum = None, 0.0
    for s in range(6):
        c.filter()
        f = c.state(0)
        for nk in range(1, 5):
            c.rk_stage(nk)
        c.end_step()
        q = c.state(0)
        dt = c.clock()["dt"]
        qb = (0.5 * dt) * (f + q) if qb is None else qb + (0.5 * dt) * (f + q)
        dtsum += dt
        if s == 0:   # first step: the filtered state itself
            assert np.array_equal(c.record_read(0), _outputs(f))
    assert dtsum > 0.0
    got = c.record_read(0)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, _outputs((1.0 / dtsum) * qb))

    # the same steps replayed as graphs record the same averages
    g = _case(blocks, grids, states, params)
    g.use_graph(True)
    g.record(True)
    g.step(6)
    assert np.array_equal(g.record_read(0), got)

    # a reset starts a new average
    c.record_reset()
    qb, dtsum = None, 0.0
    for s in range(2):
        c.filter()
        f = c.state(0)
        for nk in range(1, 5):
            c.rk_stage(nk)
        c.end_step()
        q = c.state(0)
        dt = c.clock()["dt"]
        qb = (0.5 * dt) * (f + q) if qb is None else qb + (0.5 * dt) * (f + q)
        dtsum += dt
    assert np.array_equal(c.record_read(0), _outputs((1.0 / dtsum) * qb))
    c.close()
    g.close()


def test_record_read_refuses_before_enable():
    from canard_amd.solver import CanardError
    blocks, grids, states, params = tgv_case(16)
    c = _case(blocks, grids, states, params)
    with pytest.raises(CanardError, match="never enabled"):
        c.record_read(0)
    c.close()
