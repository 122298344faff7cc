"""Pins of the CPU reference of the learned-closure network: finite
differences of its gradients and periodic shift equivariance."""
import numpy as np

import closure_ref as R


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    n = 4
    p = R.init_params(rng, gain3=1.0)
    x = rng.standard_normal((n ** 3, R.C0))
    gy = rng.standard_normal((n ** 3, R.C3))
    gx, gp = R.grad(p, x, gy, n)
    L = lambda pp, xx: float(np.sum(gy * R.forward(pp, xx, n)[0]))
    h = 1e-6
    for k in rng.choice(R.NPARAM, 40, replace=False):
        e = np.zeros(R.NPARAM); e[k] = h
        fd = (L(p + e, x) - L(p - e, x)) / (2 * h)
        assert abs(fd - gp[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, gp[k])
    for _ in range(20):
        v, c = rng.integers(n ** 3), rng.integers(R.C0)
        e = np.zeros_like(x); e[v, c] = h
        fd = (L(p, x + e) - L(p, x - e)) / (2 * h)
        assert abs(fd - gx[v, c]) <= 1e-6 * max(1.0, abs(fd))


def test_periodic_shift_equivariance():
    rng = np.random.default_rng(2)
    n = 4
    p = R.init_params(rng)
    x = rng.standard_normal((n ** 3, R.C0))
    y = R.forward(p, x, n)[0]
    s = lambda a: np.roll(a.reshape(n, n, n, -1), (1, 2, 3), axis=(0, 1, 2)).reshape(n ** 3, -1)
    assert np.allclose(R.forward(p, s(x), n)[0], s(y), rtol=0, atol=1e-12)


def test_bf16_rounding():
    assert R.bf16(1.0) == 1.0
    assert R.bf16(1.0 + 2 ** -9) == 1.0                    # below half an ulp (2^-7)
    assert R.bf16(1.0 + 2 ** -8) == 1.0                    # tie: to even
    assert R.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6      # tie: to even
    assert abs(R.bf16(np.pi) - np.pi) < 2 ** -8 * np.pi


def test_closure_norm_stats_share_momentum_scale():
    from canard_amd.closure_train import norm_stats
    rng = np.random.default_rng(3)
    ins = rng.normal(size=(4, 5, 1000))
    ins[:, 0] += 1.0
    ins[:, 3] *= 1e-6          # a momentum component that starts near zero
    outs = rng.normal(size=(4, 5, 1000)) * np.array([1, 2, 2, 2e-6, 3])[None, :, None]
    m, s, o = norm_stats(ins, outs)
    assert abs(m[0] - 1.0) < 0.1
    assert s[1] == s[2] == s[3] and 0.5 < s[1] < 1.0
    assert o[1] == o[2] == o[3] and 1.0 < o[1] < 2.0
    assert abs(o[4] - 3.0) < 0.2
