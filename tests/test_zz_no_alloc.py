"""cnd_net_apply_noalloc (the allocation-free apply, scratch in a fixed
device array) against cnd_net_apply, bit for bit, on two calls in a row,
the second on a smaller grid; and its refusal of grids above the limit."""
import numpy as np
import pytest

from canard_amd import cases

pytestmark = pytest.mark.gpu


def _pair(n, seed):
    from canard_amd.closure import Net
    from canard_amd.solver import Case
    g = cases.gridtoy(n, n, n)
    q = cases.tgv_state(g, (n + 1,) * 3, mach=0.2, pert=1e-2, seed=seed)
    ctxs = []
    for _ in range(2):
        c = Case([cases.periodic_block(n)], dict(reoo=1000.0))
        c.set_grids([g])
        c.set_states([q])
        ctxs.append(c)
    net = Net(n, seed=seed)
    net.set_norm(np.array([1.0, 0, 0, 0, 1.8]), np.array([0.1, 0.2, 0.2, 0.2, 0.3]), np.full(5, 1e-2))
    return net, ctxs, q


def test_noalloc_apply_equals_apply_twice_smaller_second():
    for n, seed in ((64, 3), (32, 4)):
        print(f"[no-alloc] n={n}", flush=True)
        net, (a, b), q0 = _pair(n, seed)
        net.apply(a, 0)
        net.apply_noalloc(b, 0)
        qa, qb = a.state(0), b.state(0)
        assert np.array_equal(qa, qb), n
        assert not np.array_equal(qa, q0)
        for x in (a, b, net):
            x.close()


def test_noalloc_refuses_grids_above_the_limit():
    from canard_amd.closure import ClosureError
    net, (a, b), _ = _pair(72, 5)
    with pytest.raises(ClosureError, match="exceeds the scratch limit"):
        net.apply_noalloc(a, 0)
    for x in (a, b, net):
        x.close()
