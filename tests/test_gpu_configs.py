"""The benchmark configurations at their full sizes.

Config 1: 64^3 periodic Taylor-Green vortex, Mach 0.1, Re 1600, fltk 0.49,
cfl 0.8, 100 steps, against the oracle (1e-12 per conserved variable).
Config 2: the 256^3 Taylor-Green vortex with the seeded 1e-3 perturbation,
compared with the oracle at 128^3 (20 steps) and 256^3 (12 steps; steps 0
and 1 have dt = 0, dt then ramps to the computed dte at step 11), and run
for its full 200 steps on the device alone with checks that hold at any
size: finite, kinetic energy non-increasing, and (without the perturbation)
the vortex's mirror symmetry z -> -z.

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
from helpers import oracle_for, rel_err, tgv_case

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
    steps = 30   # dt reaches its computed value dte at step 11 and is recomputed at 21
    c.step(steps)
    m.step(steps)
    o.run(steps)
    qcs, qos = [c.state(b) for b in range(4)], [o.state(b) for b in range(4)]
    big = np.max([np.max(np.abs(qo - s0), axis=1) for qo, s0 in zip(qos, states)], axis=0)
    big[1:4] = np.max(big[1:4])
    for b in range(4):
        qc, qo = qcs[b], qos[b]
        e = rel_err(qc, qo)
        ec = np.max(np.abs((qc - states[b]) - (qo - states[b])), axis=1) / big
        em = rel_err(m.state(b), qc)
        print(f"  block {b}: fp64 vs oracle {e}, change {ec}, mixed vs fp64 {em}", flush=True)
        assert np.all(e < 1e-12), (b, e)
        assert np.all(ec < 1e-9), (b, ec)
        assert np.all(em < 1e-6), (b, em)
    for x in (c, m, o):
        x.close()


def test_config1_tgv64_100_steps():
    blocks, grids, states, params = tgv_case(64)
    assert params["fltk"] == 0.49 and params["cfl"] == 0.8 and params["reoo"] == 1600.0 / 0.1
    print("[config1] 64^3 TGV, 100 steps", flush=True)
    c = _case(blocks, grids, states, params)
    o = oracle_for(blocks, grids, states, params, threads=16)
    c.step(100)
    o.run(100)
    err = rel_err(c.state(0), o.state(0))
    print("  rel err per variable after 100 steps", err, "clock", c.clock(), flush=True)
    assert np.all(err < 1e-12), err
    cg, co = c.clock(), o.clock()
    assert cg["n"] == co["n"] == 100
    assert abs(cg["timo"] - co["timo"]) <= 1e-13 * co["timo"]
    c.close()
    o.close()


@pytest.mark.parametrize("n,steps", [(128, 20), (256, 12)])
def test_config2_parity_sizes(n, steps):
    blocks, grids, states, params = tgv_case(n, pert=1e-3)
    print(f"[config2] {n}^3 TGV with perturbation, {steps} steps", flush=True)
    c = _case(blocks, grids, states, params)
    c.step(steps)
    got = c.state(0)
    clk = c.clock()
    c.close()
    o = oracle_for(blocks, grids, states, params, threads=16)
    o.run(steps)
    err = rel_err(got, o.state(0))
    print("  rel err per variable", err, "clock", clk, flush=True)
    assert np.all(err < 1e-12), err
    assert clk["n"] == o.clock()["n"] == steps
    o.close()


def _kinetic_energy(q, n):
    s = q.reshape(5, n + 1, n + 1, n + 1)[:, :n, :n, :n]
    return 0.5 * float(np.sum((s[1] ** 2 + s[2] ** 2 + s[3] ** 2) / s[0])) / n ** 3


@pytest.mark.parametrize("pert", [1e-3, 0.0])
def test_config2_tgv256_200_steps(pert):
    n = 256
    blocks, grids, states, params = tgv_case(n, pert=pert)
    print(f"[config2] 256^3 TGV, 200 steps on the device, perturbation {pert}", flush=True)
    c = _case(blocks, grids, states, params)
    c.use_graph(True)
    ke = [_kinetic_energy(states[0], n)]
    for _ in range(20):
        c.step(10)
        q = c.state(0)
        assert np.all(np.isfinite(q))
        ke.append(_kinetic_energy(q, n))
    clk = c.clock()
    c.close()
    print("  kinetic energy every 10 steps", ke, "clock", clk, flush=True)
    assert clk["n"] == 200
    assert all(b <= a * (1.0 + 1e-14) for a, b in zip(ke, ke[1:])), ke
    assert ke[-1] < ke[0]
    if pert == 0.0:
        s = q.reshape(5, n + 1, n + 1, n + 1)
        m = s[:, ::-1]                    # k -> lze - k is z -> -z
        sign = np.array([1, 1, 1, -1, 1]).reshape(5, 1, 1, 1)
        err = np.max(np.abs(s - sign * m).reshape(5, -1), axis=1) / np.max(np.abs(s).reshape(5, -1), axis=1)
        err[1:4] = np.max(np.abs(s - sign * m).reshape(5, -1)[1:4], axis=1) / np.max(np.abs(s[1:4]))
        print("  z-mirror symmetry error per variable", err, flush=True)
        assert np.all(err < 1e-10), err


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
    for x in (c, m):
        x.close()
