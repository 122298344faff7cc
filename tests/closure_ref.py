"""CPU FP64 reference of the learned-closure network (the checker of the
BF16 device kernels; test infrastructure only).

Three periodic 3x3x3 convolutions, 5 -> 32 -> 32 -> 5 channels, ReLU after
the first two; parameters in the layout of include/cnd.h: W1[32][27][5],
b1[32], W2[32][27][32], b2[32], W3[5][27][32], b3[5], tap t = (dx+1) +
3 (dy+1) + 9 (dz+1), out[v][o] = sum_{t,c} a[v + d_t][c] W[o][t][c] + b[o].
Fields are channel-last [n^3][C] with voxel v = x + n (y + n z).

With bf16=True every value the device stores in BF16 (inputs, weights,
activations, back-propagated gradients) is rounded the same way, so only
the FP32-vs-FP64 accumulation differs from the device."""
import numpy as np

C0, C1, C2, C3, T = 5, 32, 32, 5, 27
SIZES = [C1 * T * C0, C1, C2 * T * C1, C2, C3 * T * C2, C3]
NPARAM = sum(SIZES)
OFFS = np.cumsum([0] + SIZES)


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(a, dtype=np.float32).copy()
    u = f.view(np.uint32)
    u += np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))
    u &= np.uint32(0xFFFF0000)
    return f.astype(np.float64)


def split(p):
    p = np.asarray(p, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = (p[OFFS[i]:OFFS[i + 1]] for i in range(6))
    return W1.reshape(C1, T, C0), b1, W2.reshape(C2, T, C1), b2, W3.reshape(C3, T, C2), b3


def _d(t):
    return t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def _shift(a, t, n, sign=1):
    """a[v + sign d_t] for a of shape [V][C]."""
    dx, dy, dz = _d(t)
    g = a.reshape(n, n, n, -1)   # [z][y][x][c]
    g = np.roll(g, (-sign * dz, -sign * dy, -sign * dx), axis=(0, 1, 2))
    return g.reshape(n ** 3, -1)


def _conv(a, W, b, n):
    out = np.zeros((a.shape[0], W.shape[0]))
    for t in range(T):
        out += _shift(a, t, n) @ W[:, t, :].T
    return out + (b if b is not None else 0.0)


def forward(p, x, n, bf=False):
    W1, b1, W2, b2, W3, b3 = split(p)
    r = bf16 if bf else (lambda v: v)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    x = r(x)
    z1 = _conv(x, W1, b1, n)
    a1 = r(np.maximum(z1, 0.0))
    z2 = _conv(a1, W2, b2, n)
    a2 = r(np.maximum(z2, 0.0))
    y = _conv(a2, W3, b3, n)
    return y, (x, a1, a2, W1, W2, W3)


def _wgrad(g, a, n, co):
    gw = np.zeros((co, T, a.shape[1]))
    for t in range(T):
        gw[:, t, :] = g.T @ _shift(a, t, n)
    return gw


def _dgrad(g, W, n):
    out = np.zeros((g.shape[0], W.shape[2]))
    for t in range(T):
        out += _shift(g, t, n, sign=-1) @ W[:, t, :]
    return out


def grad(p, x, gy, n, bf=False):
    """Gradients of L = sum(gy * y) w.r.t. x and the flat parameters."""
    r = bf16 if bf else (lambda v: v)
    _, (x, a1, a2, W1, W2, W3) = forward(p, x, n, bf)
    g3 = r(gy)
    gW3, gb3 = _wgrad(g3, a2, n, C3), g3.sum(0)
    g2 = r(_dgrad(g3, W3, n) * (a2 > 0))
    gW2, gb2 = _wgrad(g2, a1, n, C2), g2.sum(0)
    g1 = r(_dgrad(g2, W2, n) * (a1 > 0))
    gW1, gb1 = _wgrad(g1, x, n, C1), g1.sum(0)
    gx = _dgrad(g1, W1, n)
    gp = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return gx, gp


def init_params(rng, gain3=0.1):
    p = np.zeros(NPARAM)
    for i, (fan, g) in enumerate([(T * C0, 1.0), (T * C1, 1.0), (T * C2, gain3)]):
        a = g * np.sqrt(6.0 / fan)
        p[OFFS[2 * i]:OFFS[2 * i + 1]] = rng.uniform(-a, a, SIZES[2 * i])
        p[OFFS[2 * i + 1]:OFFS[2 * i + 2]] = rng.uniform(-0.1, 0.1, SIZES[2 * i + 1])
    return p
