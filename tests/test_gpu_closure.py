"""Learned closure on the GPU: the BF16 matrix-core network against the CPU
FP64 reference (tests/closure_ref.py) with the same BF16 roundings, Adam
training on a known target, the box filter and the state update.

Tolerances: forward 2e-3 and gradients 1e-2 of the field's max norm (the
reference rounds to BF16 where the device stores BF16, so what is left is
FP32 accumulation and rounding flips of single BF16 values); box filter
1e-13 (FP64 arithmetic, different summation order)."""
import numpy as np
import pytest

import closure_ref as R
from canard_amd import cases

pytestmark = pytest.mark.gpu


def _net(n, cap=0, seed=0):
    from canard_amd.closure import Net
    return Net(n, cap, seed)


def test_forward_and_gradients_match_fp64_reference():
    n = 8
    rng = np.random.default_rng(4)
    net = _net(n)
    p = R.init_params(rng, gain3=1.0).astype(np.float32)
    net.set_params(p)
    x = rng.standard_normal((n ** 3, 5)).astype(np.float32)
    y = net.forward(x)
    yr = R.forward(p.astype(np.float64), x.astype(np.float64), n, bf=True)[0]
    e = np.max(np.abs(y - yr)) / np.max(np.abs(yr))
    print(f"[closure] forward max rel err {e:.2e}", flush=True)
    assert e < 2e-3
    gy = rng.standard_normal((n ** 3, 5)).astype(np.float32)
    gx, gp = net.grad(x, gy)
    gxr, gpr = R.grad(p.astype(np.float64), x.astype(np.float64), gy.astype(np.float64), n, bf=True)
    ex = np.max(np.abs(gx - gxr)) / np.max(np.abs(gxr))
    print(f"[closure] input gradient max rel err {ex:.2e}", flush=True)
    assert ex < 1e-2
    for i in range(6):
        a, b = R.OFFS[i], R.OFFS[i + 1]
        ep = np.max(np.abs(gp[a:b] - gpr[a:b])) / np.max(np.abs(gpr[a:b]))
        print(f"[closure] parameter block {i} gradient max rel err {ep:.2e}", flush=True)
        assert ep < 1e-2, i
    net.close()


def test_training_reduces_loss_on_a_known_target():
    """Targets made by another network of the same form: Adam on the device
    must bring the loss well down."""
    n = 16
    rng = np.random.default_rng(5)
    teacher = R.init_params(rng, gain3=1.0)
    net = _net(n, cap=8, seed=3)
    for _ in range(8):
        x = rng.standard_normal((n ** 3, 5))
        net.add_pair_host(x, R.forward(teacher, x, n)[0])
    before = net.eval()
    net.train(1000, lr=2e-3, seed=1)
    after = net.eval()
    print(f"[closure] loss {before:.3e} -> {after:.3e}", flush=True)
    assert after < 0.3 * before


def test_box_filter_and_apply():
    from canard_amd.closure import box_filter
    from canard_amd.solver import Case
    nf, n = 64, 16
    fine = Case([cases.periodic_block(nf)], dict(reoo=1000.0))
    coarse = Case([cases.periodic_block(n)], dict(reoo=1000.0))
    gf = cases.gridtoy(nf, nf, nf)
    gc = cases.gridtoy(n, n, n)
    fine.set_grids([gf]); coarse.set_grids([gc])
    rng = np.random.default_rng(6)
    qf = cases.tgv_state(gf, (nf + 1,) * 3, mach=0.2, pert=1e-2, seed=9)
    fine.set_states([qf]); coarse.set_states([cases.tgv_state(gc, (n + 1,) * 3, mach=0.2)])
    box_filter(fine, 0, coarse, 0)
    got = coarse.state(0).reshape(5, n + 1, n + 1, n + 1)
    f = qf.reshape(5, nf + 1, nf + 1, nf + 1)[:, :nf, :nf, :nf]   # unique points [m][z][y][x]
    w = np.array([0.5, 1, 1, 1, 0.5]) / 4.0
    want = np.zeros((5, n, n, n))
    for c, wc in enumerate(w):
        for b, wb in enumerate(w):
            for a, wa in enumerate(w):
                sh = np.roll(f, (-(c - 2), -(b - 2), -(a - 2)), axis=(1, 2, 3))
                want += wc * wb * wa * sh[:, ::4, ::4, ::4]
    full = np.pad(want, ((0, 0), (0, 1), (0, 1), (0, 1)), mode="wrap")
    assert np.max(np.abs(got - full)) <= 1e-13 * np.max(np.abs(full))

    # q += net(q): compare with the host forward of the same normalised input
    net = _net(n, seed=7)
    mean, scale, oscale = np.array([1.0, 0, 0, 0, 1.8]), np.array([0.1, 0.2, 0.2, 0.2, 0.3]), np.full(5, 1e-3)
    net.set_norm(mean, scale, oscale)
    q0 = coarse.state(0)
    net.apply(coarse, 0)
    q1 = coarse.state(0)
    u = q0.reshape(5, n + 1, n + 1, n + 1)[:, :n, :n, :n].reshape(5, -1).T   # [v][m], v = x + n (y + n z)
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)   # the device keeps the norms in FP32
    x = ((u - f32(mean)) / f32(scale)).astype(np.float32)
    dy = net.forward(x).astype(np.float64) * f32(oscale)
    d = (q1 - q0).reshape(5, n + 1, n + 1, n + 1)
    dfull = np.pad(dy.T.reshape(5, n, n, n), ((0, 0), (0, 1), (0, 1), (0, 1)), mode="wrap")
    assert np.max(np.abs(d - dfull)) <= 1e-6 * np.max(np.abs(dfull)) + 1e-15
    for c in (fine, coarse, net):
        c.close()


def test_apply_to_two_contexts_in_a_row_equals_serial_applies():
    """One network applied to two contexts back to back (each apply runs on
    its context's stream, sharing the network's activation buffers) gives
    bit for bit what two applies separated by a synchronisation give."""
    from canard_amd.solver import Case
    n = 32
    g = cases.gridtoy(n, n, n)
    states = [cases.tgv_state(g, (n + 1,) * 3, mach=0.2, pert=1e-2, seed=s) for s in (1, 2)]
    net = _net(n, seed=5)
    net.set_norm(np.array([1.0, 0, 0, 0, 1.8]), np.array([0.1, 0.2, 0.2, 0.2, 0.3]), np.full(5, 1e-2))

    def make(q):
        c = Case([cases.periodic_block(n)], dict(reoo=1000.0))
        c.set_grids([g])
        c.set_states([q])
        return c

    a = [make(q) for q in states]
    for c in a:
        net.apply(c, 0)           # no synchronisation between the two
    got = [c.state(0) for c in a]
    b = [make(q) for q in states]
    want = []
    for c in b:
        net.apply(c, 0)
        want.append(c.state(0))   # synchronises
    for x, y in zip(got, want):
        assert np.array_equal(x, y)
    assert not np.array_equal(got[0], states[0])
    for c in a + b + [net]:
        c.close()
