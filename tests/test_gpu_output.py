"""Output of results on the device: the output schedule of the time loop
(mo_canard_driver.f90:286-291) against a host restatement of its test on
the clock, the Tecplot files of the recorded averages and of the grid
(read back with an independent reader), and the asynchronous write (host
thread, pinned copy on a second stream) against the synchronous one while
steps go on."""
import numpy as np
import pytest

from canard_amd import cases
from helpers import tgv_case
from plt_reader import read_plt

pytestmark = pytest.mark.gpu


def _case(n=16, blocks=None):
    from canard_amd.solver import Case
    b, g, s, p = tgv_case(n, pert=1e-3)
    c = Case(b, p)
    c.set_grids(g)
    c.set_states(s)
    c.set_clock()
    return c, g


def test_output_schedule_and_files(tmp_path):
    c, grids = _case()
    tmax, ndata = 1.5, 6
    c.record(True)
    c.output_schedule(tmax, ndata)
    c.output_grid(tmp_path / "grid.plt")
    gf = read_plt(tmp_path / "grid.plt")
    assert gf["file_type"] == 1 and gf["names"] == ["x", "y", "z"]
    assert np.array_equal(gf["zones"][0]["data"], grids[0].astype(np.float32))
    ndati, dtsum, written = -1, 0.0, []
    timo = c.clock()["timo"]
    while timo < tmax:
        c.step(1)
        clk = c.clock()
        dtsum += clk["dt"]
        # the reference's test, on the step's start time and dt
        res = (ndati + 1) * tmax / ndata
        want = (timo - res) * (timo + clk["dt"] - res) <= 0.0
        due, nd, t = c.output_due()
        assert due == want, (clk["n"], timo, clk["dt"], res)
        if want:
            ndati += 1
            assert nd == ndati
            assert t == clk["timo"] - 0.5 * dtsum
            rec = c.record_read(0)
            path = tmp_path / f"n{nd:03d}.plt"
            c.output_write(path)
            f = read_plt(path)
            assert f["title"] == f"n{nd:03d}" and f["file_type"] == 2
            assert f["names"] == ["rho", "u", "v", "w", "p"]
            z = f["zones"][0]
            assert z["time"] == t and z["dims"] == (17, 17, 17)
            assert np.array_equal(z["data"], rec)
            written.append(nd)
            dtsum = 0.0
        timo = clk["timo"]
    print(f"[output] {len(written)} outputs over {c.clock()['n']} steps", flush=True)
    assert written == list(range(len(written))) and len(written) >= ndata
    c.close()


def test_output_schedule_under_graph_replay():
    """The schedule is decided on the device inside the step, so steps
    replayed as a graph report the same outputs as eager steps."""
    a, _ = _case()
    b, _ = _case()
    b.use_graph(True)
    for c in (a, b):
        c.output_schedule(1.0, 8)
    seq = {0: [], 1: []}
    for _ in range(40):
        for k, c in enumerate((a, b)):
            c.step(1)
            seq[k].append(c.output_due())
    print("[output graph]", sum(d for d, _, _ in seq[0]), "outputs in 40 steps", flush=True)
    assert seq[0] == seq[1]
    assert sum(d for d, _, _ in seq[0]) >= 2
    for c in (a, b):
        c.close()


def test_async_output_equals_sync(tmp_path):
    a, _ = _case(24)
    b, _ = _case(24)
    for c in (a, b):
        c.record(True)
        c.step(7)
    a.output_write(tmp_path / "sync.plt")
    b.output_write(tmp_path / "async.plt", wait=False)
    b.step(5)             # steps overlap the write
    a.step(5)
    b.output_wait()
    assert open(tmp_path / "sync.plt", "rb").read() == open(tmp_path / "async.plt", "rb").read()
    # the sums were cleared after forming the output, before the new steps
    assert np.array_equal(a.record_read(0), b.record_read(0))
    assert np.array_equal(a.state(0), b.state(0))
    for c in (a, b):
        c.close()


def test_output_of_one_ensemble_instance(tmp_path):
    """Each block of the instance is a zone; blocks of other instances are
    left out and keep their sums."""
    from canard_amd.solver import Case
    specs = [(12, 0.1), (16, 0.2)]
    blocks, grids, states, params = [], [], [], []
    for i, (n, m) in enumerate(specs):
        b, g, s, p = tgv_case(n, mach=m, pert=1e-3, seed=i)
        b[0]["inst"] = i
        b[0]["mcd"] = [[i, i]] * 3
        blocks += b; grids += g; states += s; params.append(p)
    c = Case(blocks, params)
    c.set_grids(grids)
    c.set_states(states)
    for i in range(2):
        c.set_clock(inst=i)
    c.record(True)
    c.step(4)
    other = c.record_read(0)
    want = c.record_read(1)
    c.output_write(tmp_path / "i1.plt", inst=1)
    f = read_plt(tmp_path / "i1.plt")
    assert len(f["zones"]) == 1 and f["zones"][0]["dims"] == (17, 17, 17)
    assert np.array_equal(f["zones"][0]["data"], want)
    assert np.array_equal(c.record_read(0), other)
    c.close()
