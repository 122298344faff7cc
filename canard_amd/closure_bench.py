"""Timing of the learned closure's kernels on an n^3 coarse grid (default
64^3): Adam training iterations (forward, loss, backward, update) on random
pairs, and applications q <- q + N(q) to a periodic Taylor-Green block.

    python -m canard_amd.closure_bench [--n 64] [--iters 50] [--applies 50]

Prints one JSON line with the wall time per iteration and per apply and
the algorithmic FLOP of each (multiply-adds of the three 3x3x3
convolutions at their real channel counts, 5 -> 32 -> 32 -> 5, times two);
run under rocprofv3 --kernel-trace --stats for the per-kernel durations."""
import argparse
import json
import sys
import time

import numpy as np

from . import cases
from .closure import Net
from .solver import Case

TAPS = 27
CHANNELS = (5, 32, 32, 5)


def layer_flop(n):
    """Algorithmic FLOP of each convolution layer's forward pass on n^3 voxels."""
    v = n ** 3
    return [2 * v * TAPS * ci * co for ci, co in zip(CHANNELS[:-1], CHANNELS[1:])]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--applies", type=int, default=50)
    ap.add_argument("--pairs", type=int, default=8)
    a = ap.parse_args(argv)
    n = a.n
    rng = np.random.default_rng(7)
    net = Net(n, capacity=a.pairs, seed=1)
    net.set_norm(np.array([1.0, 0, 0, 0, 1.8]), np.array([0.1, 0.2, 0.2, 0.2, 0.3]), np.full(5, 1e-2))
    for _ in range(a.pairs):
        net.add_pair_host(rng.standard_normal((n ** 3, 5)).astype(np.float32),
                          0.1 * rng.standard_normal((n ** 3, 5)).astype(np.float32))
    print(f"[closure_bench] {n}^3, {a.pairs} pairs: training", file=sys.stderr, flush=True)
    net.train(2, lr=1e-4, seed=1)
    t0 = time.perf_counter()
    net.train(a.iters, lr=1e-4, seed=2)
    train_s = time.perf_counter() - t0

    print(f"[closure_bench] {n}^3: applies", file=sys.stderr, flush=True)
    g = cases.gridtoy(n, n, n)
    c = Case([cases.periodic_block(n)], dict(reoo=1000.0))
    c.set_grids([g])
    c.set_states([cases.tgv_state(g, (n + 1,) * 3, mach=0.1)])
    net.apply(c, 0)
    c.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.applies):
        net.apply(c, 0)
    c.synchronize()
    apply_s = time.perf_counter() - t0
    finite = bool(np.all(np.isfinite(c.state(0))))
    c.close()
    net.close()

    fwd = layer_flop(n)
    # an iteration: the forward, the data gradients of layers 3 and 2 (the
    # forward flops of those layers) and the three weight gradients
    # (the forward flops of all layers)
    it_flop = sum(fwd) + fwd[2] + fwd[1] + sum(fwd)
    out = {"n": n, "voxels": n ** 3, "iters": a.iters, "applies": a.applies,
           "ms_per_iter": train_s / a.iters * 1e3, "ms_per_apply": apply_s / a.applies * 1e3,
           "gflop_forward_per_layer": [f / 1e9 for f in fwd], "gflop_forward": sum(fwd) / 1e9,
           "gflop_per_iter": it_flop / 1e9,
           "tflops_iter": it_flop / (train_s / a.iters) / 1e12,
           "tflops_apply": sum(fwd) / (apply_s / a.applies) / 1e12,
           "finite": finite}
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    main()
