"""Invalid cases and out-of-order calls are refused with a status and a
message, as the reference stops on them (its STOP 1 becomes a non-zero
return and cnd_last_error).  Creation checks run on the host before any
device allocation; they need a device only because cnd_create looks for one
first."""
import numpy as np
import pytest

from helpers import tgv_case

pytestmark = pytest.mark.gpu

PER = [[45, 45]] * 3


def _blk(lxi=16, let=16, lze=16, nbc=None, mcd=None, inst=0):
    return dict(lxi=lxi, let=let, lze=lze, nbc=nbc or PER, mcd=mcd or [[0, 0]] * 3, inst=inst)


def _pair(n0, n1, inst=(0, 0), code=35, reciprocal=True):
    """Blocks 0 and 1 joined on their xi faces, each periodic in eta and zeta."""
    b0 = _blk(lxi=n0[0], let=n0[1], lze=n0[2], nbc=[[code, code], [45, 45], [45, 45]],
              mcd=[[1, 1], [0, 0], [0, 0]], inst=inst[0])
    part1 = [[0, 0], [1, 1], [1, 1]] if reciprocal else [[1, 1], [1, 1], [1, 1]]
    b1 = _blk(lxi=n1[0], let=n1[1], lze=n1[2], nbc=[[code, code], [45, 45], [45, 45]], mcd=part1, inst=inst[1])
    return [b0, b1]


CREATE_CASES = {
    "too_few_points": ([_blk(lxi=11)], [{}], ">= 13 points"),
    "too_many_points": ([_blk(lxi=2112)], [{}], "at most 2112"),
    "unknown_bc": ([_blk(nbc=[[99, 45], [45, 45], [45, 45]])], [{}], "unknown boundary code 99"),
    "interface_without_partner": ([_blk(mcd=[[-1, 0], [0, 0], [0, 0]])], [{}], "without partner"),
    "partner_out_of_range": ([_blk(mcd=[[3, 0], [0, 0], [0, 0]])], [{}], "without partner"),
    "not_reciprocal": (_pair((16, 16, 16), (16, 16, 16), reciprocal=False), [{}], "not reciprocal"),
    "non_conforming": (_pair((16, 16, 16), (16, 20, 16)), [{}], "non-conforming"),
    "across_instances": (_pair((16, 16, 16), (16, 16, 16), inst=(0, 1)), [{}, {}], "two instances"),
    "instance_out_of_range": ([_blk(inst=1)], [{}], "instance out of range"),
    "fltrbc_one": ([_blk()], [dict(fltrbc=1.0)], "bad parameters"),
    "nonpositive_reoo": ([_blk()], [dict(reoo=0.0)], "bad parameters"),
    "sponge_without_width": ([_blk()], [dict(sponge=1, szth0=0.0)], "sponge needs"),
    "fixed_dt_without_timf": ([_blk()], [dict(nsmf=0, timf=0.0)], "needs timf"),
    "viscous_and_euler": ([_blk(), _blk(inst=1, mcd=[[1, 1]] * 3)], [dict(viscous=1), dict(viscous=0)],
                          "mixed viscous/Euler"),
}


@pytest.mark.parametrize("name", list(CREATE_CASES))
def test_create_rejects(name):
    from canard_amd.solver import Case, CanardError
    blocks, params, msg = CREATE_CASES[name]
    print(f"[errors] create {name}", flush=True)
    with pytest.raises(CanardError, match=msg):
        Case(blocks, params)


def test_create_rejects_unknown_precision():
    from canard_amd.solver import Case, CanardError
    with pytest.raises(CanardError, match="unknown precision"):
        Case([_blk()], [{}], precision=7)


def test_calls_out_of_order_and_bad_arguments():
    """Steps before the metrics, stages out of order (the filter twice, a
    stage other than the next one, the end of a step before its four stages,
    a stage after a finished step without the filter), and bad indices are
    refused without touching the state; the context stays usable."""
    from canard_amd.solver import Case, CanardError
    blocks, grids, states, params = tgv_case(16, pert=1e-3)
    c = Case(blocks, params)
    with pytest.raises(CanardError, match="not set|not initialised"):
        c.step(1)
    c.set_grids(grids)
    c.set_states(states)
    c.set_clock()
    with pytest.raises(CanardError, match="nk must be 1..4"):
        c.rk_stage(5)
    with pytest.raises(CanardError, match="stage 1 must follow cnd_filter"):
        c.rk_stage(1)
    with pytest.raises(CanardError, match="negative step count"):
        c.step(-1)
    with pytest.raises(CanardError, match="bad instance"):
        c.set_clock(inst=3)
    with pytest.raises(CanardError, match="bad argument"):
        c.set_shape("grad", 3, 0)
    with pytest.raises(CanardError, match="four stages of the step have not run"):
        c.end_step()
    with pytest.raises(CanardError, match="stage 2 out of order"):
        c.rk_stage(2)
    c.filter()
    with pytest.raises(CanardError, match="already ran"):
        c.filter()
    with pytest.raises(CanardError, match="half done"):
        c.step(1)
    with pytest.raises(CanardError, match="four stages of the step have not run"):
        c.end_step()
    c.rk_stage(1)
    with pytest.raises(CanardError, match="expected cnd_rk_stage\\(2\\)"):
        c.end_step()
    with pytest.raises(CanardError, match="stage 3 out of order"):
        c.rk_stage(3)
    with pytest.raises(CanardError, match="half done"):
        c.filter()
    with pytest.raises(CanardError, match="half done"):
        c.record(True)
    for nk in range(2, 5):
        c.rk_stage(nk)
    with pytest.raises(CanardError, match="stage 1 out of order|half done"):
        c.rk_stage(1)
    c.end_step()
    with pytest.raises(CanardError, match="stage 2 out of order"):
        c.rk_stage(2)
    c.step(1)
    q = c.state(0)
    assert np.all(np.isfinite(q))
    assert c.clock()["n"] == 2
    c.close()
