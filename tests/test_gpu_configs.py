"""The multi-block and ensemble configurations at their full sizes.

Config 3: four uneven sheared blocks (384x256x64, 192x256x64, 512x128x64,
128x128x64; 15.0 M points) in two BC-35 pairs with non-reflective faces,
periodic z, Mach 0.5 flow with pressure pulses and a sponge; FP64 against
the oracle (1e-12 per conserved variable; 1e-9 on the change from the
initial state measured against the largest change over the four blocks, as
the change is ~1e-3 of the state and rounding differences of the reordered
segmented solves show up that much larger), mixed precision against FP64
(1e-6).
Config 4: 32 periodic Taylor-Green instances of sizes drawn from
{32,48,64,96,128}^3 with drawn Mach and Re, in one context (one launch per
kernel over all instances, hipGraph replay); a sample of instances against
their own oracles (1e-12), mixed precision against FP64 (1e-6)."""
import numpy as np
import pytest

from canard_amd import cases
from helpers import oracle_for, rel_err

pytestmark = pytest.mark.gpu


def _case(blocks, grids, states, params, precision=0):
    from canard_amd.solver import Case
    c = Case(blocks, params, precision=precision)
    c.set_grids(grids)
    c.set_states(states)
    for i in range(len(params) if isinstance(params, list) else 1):
        c.set_clock(inst=i)
    return c


def test_config3_multiblock_fp64_and_mixed():
    from canard_amd.solver import CND_MIXED
    blocks, grids, states, params = cases.config3()
    print("[config3] 4 blocks,", sum(g.shape[1] for g in grids), "points", flush=True)
    c = _case(blocks, grids, states, params)
    m = _case(blocks, grids, states, params, precision=CND_MIXED)
    o = oracle_for(blocks, grids, states, params, threads=16)
    steps = 4
    c.step(steps)
    m.step(steps)
    o.run(steps)
    qcs, qos = [c.state(b) for b in range(4)], [o.state(b) for b in range(4)]
    big = np.max([np.max(np.abs(qo - s0), axis=1) for qo, s0 in zip(qos, states)], axis=0)
    big[1:4] = np.max(big[1:4])
    for b in range(4):
        qc, qo, qm = qcs[b], qos[b], m.state(b)
        e = rel_err(qc, qo)
        ec = np.max(np.abs((qc - states[b]) - (qo - states[b])), axis=1) / big
        em = rel_err(qm, qc)
        print(f"  block {b}: fp64 vs oracle {e}, change {ec}, mixed vs fp64 {em}", flush=True)
        assert np.all(e < 1e-12), (b, e)
        assert np.all(ec < 1e-9), (b, ec)
        assert np.all(em < 1e-6), (b, em)
    for x in (c, m, o):
        x.close()


def test_config4_ensemble_fp64_graph_and_mixed():
    from canard_amd.solver import CND_MIXED
    blocks, grids, states, params, draws = cases.config4()
    print("[config4] sizes", [d["n"] for d in draws], flush=True)
    c = _case(blocks, grids, states, params)
    c.use_graph(True)
    m = _case(blocks, grids, states, params, precision=CND_MIXED)
    m.use_graph(True)
    steps = 4
    c.step(steps)
    m.step(steps)
    sample = sorted({0, 5, 11, 17, 23, 31})
    for i in sample:
        o = oracle_for([dict(blocks[i], inst=0, mcd=[[0, 0]] * 3)], [grids[i]], [states[i]], params[i], threads=16)
        o.run(steps)
        e = rel_err(c.state(i), o.state(0))
        print(f"  instance {i} n={draws[i]['n']}: fp64 vs oracle {e}", flush=True)
        assert np.all(e < 1e-12), (i, e)
        assert c.clock(i)["n"] == o.clock()["n"] == steps
        o.close()
    worst = np.zeros(5)
    for i in range(len(blocks)):
        worst = np.maximum(worst, rel_err(m.state(i), c.state(i)))
    print("  mixed vs fp64, worst over instances", worst, flush=True)
    assert np.all(worst < 1e-6), worst
    c.close()
    m.close()
