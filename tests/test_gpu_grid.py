"""Grids generated on the device (cnd_make_grid) against the host
generators: gridtoy and gridtoy_multi bit for bit (the same operations in
the same order, src/mo_gridgen.f90:510-654), the sheared block to rounding
of the sine, and metrics computed from a device grid equal to those of the
uploaded one."""
import numpy as np
import pytest

from canard_amd import cases

pytestmark = pytest.mark.gpu


def test_gridtoy_and_metrics_match_the_uploaded_grid():
    from canard_amd.solver import Case
    n = (20, 16, 12)
    blk = dict(lxi=n[0], let=n[1], lze=n[2], nbc=[[45, 45]] * 3, mcd=[[0, 0]] * 3)
    par = (n[0], n[1], n[2], 0.0, 2 * np.pi, np.pi, 2 * np.pi)
    a = Case([blk], {})
    a.make_grid(0, "toy", *par)
    a.init_metrics()
    want = cases.gridtoy(*n, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi)
    assert np.array_equal(a.grid(0), want)
    b = Case([blk], {})
    b.set_grids([want])
    for x, y in zip(a.metrics(0), b.metrics(0)):
        assert np.array_equal(x, y)
    a.close()
    b.close()


def test_gridtoy_multi_six_blocks():
    from canard_amd.solver import Case
    lxi0, lxi1, lxi2, let0, lze0 = 14, 18, 16, 13, 12
    blocks = []
    for mb in range(6):
        lxi = (lxi0, lxi1, lxi2)[mb % 3]
        blocks.append(dict(lxi=lxi, let=let0, lze=lze0, nbc=[[10, 10], [10, 10], [45, 45]],
                           mcd=[[-1, -1], [-1, -1], [mb, mb]]))
    c = Case(blocks, {})
    for mb in range(6):
        c.make_grid(mb, "toy_multi", lxi0, lxi1, lxi2, let0, lze0, 1.0, 3.0, 2.0, 4.0, mb)
    for mb in range(6):
        want = cases.gridtoy_multi(mb, lxi0, lxi1, lxi2, let0, lze0, doml0=1.0, doml1=3.0, domh=2.0, span=4.0)
        assert np.array_equal(c.grid(mb), want), mb
    c.close()


def test_sheared_pair_matches_host_generator():
    from canard_amd.solver import Case
    sizes = [(24, 16, 14), (16, 16, 14)]
    blocks = cases.sheared_pair_blocks(sizes)
    want = cases.sheared_pair_grids(sizes, h=1.0, amp=0.05)
    c = Case(blocks, {})
    lx, ly = float(sizes[0][0] + sizes[1][0]), float(sizes[0][1])
    c.make_grid(0, "sheared", 0.0, 1.0, 0.05, lx, ly)
    c.make_grid(1, "sheared", float(sizes[0][0]), 1.0, 0.05, lx, ly)
    for b in range(2):
        got = c.grid(b)
        assert np.max(np.abs(got - want[b])) <= 4e-15 * np.max(np.abs(want[b])), b
    c.init_metrics()
    c.close()


def test_bad_generator_arguments_are_refused():
    from canard_amd.solver import Case, CanardError
    c = Case([cases.periodic_block(16)], {})
    with pytest.raises(CanardError, match="gridtoy takes"):
        c.make_grid(0, "toy", 16, 16)
    with pytest.raises(CanardError, match="not set"):
        c.grid(0)
    c.close()
