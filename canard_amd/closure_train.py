"""Train and judge the learned closure on Taylor-Green runs.

    python -m canard_amd.closure_train [--fine 256] [--coarse 64] [--runs 4]
        [--pairs 500] [--iters 4000] [--eval-steps 300] [--alphas 0.1,0.25,0.5,1]
        [--out DIR]

Training pairs: seeded fine-grid TGV runs (FP64, Mach 0.1, Re 1600, 1e-3
grid-scale velocity perturbation plus resolved-scale Fourier modes of 5% of
the Mach number drawn from the seed, so every seed is a different flow)
box-filtered to the coarse grid.  For pair k the
coarse solver takes one step of dt_c = r dt_f from the filtered state F_k;
the input is that coarse state C_k and the target F_{k+1} - C_k, where
F_{k+1} is the fine solution after r fine steps, filtered.  The time step
is held fixed (dt_f, dt_c) through the clock so both grids cover the same
interval.  Training: Adam, one pair per step, on the device.

Judge: from the filtered state of a held-out seed (a flow of its own:
its resolved-scale modes differ from every training and calibration
seed's), the coarse solver runs
eval-steps steps with and without the network (the plain compact filter is
the fallback); both are compared with the filtered fine run over the same
time in kinetic energy (every step) and in the kinetic-energy spectrum at
the end.  The network is kept only if it lowers both errors.  With several
--alphas, the network's correction is applied scaled by the alpha whose
coarse run stays closest to the filtered fine run in kinetic energy on a
separate calibration seed, before the judge sees the held-out seed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import cases
from .closure import Net, box_filter
from .solver import Case


def log(*a):
    print("[closure]", *a, file=sys.stderr, flush=True)


MODES = 0.05   # relative amplitude of the resolved-scale modes that make each seed's flow its own


def tgv(n, seed, mach=0.1, re=1600.0, modes=None):
    """Taylor-Green run of seed: the 1e-3 grid-scale perturbation and
    resolved-scale Fourier modes (relative amplitude modes) drawn from it."""
    modes = MODES if modes is None else modes
    xyz = cases.gridtoy(n, n, n, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi)
    q = cases.tgv_state(xyz, (n + 1,) * 3, mach=mach, pert=1e-3, seed=seed, modes=modes, mode_seed=seed)
    c = Case([cases.periodic_block(n)], dict(fltk=0.49, fltrbc=0.9, reoo=re / mach, tempoo=300.0, cfl=0.8,
                                             viscous=1))
    c.set_grids([xyz])
    c.set_states([q])
    c.set_clock()
    return c


def fixed_step(c, dt, steps=1):
    """Steps of exactly dt: the clock is set so the dt schedule keeps dt
    (n = 2: no new estimate, dts = dte = dt)."""
    for _ in range(steps):
        t = c.clock()["timo"]
        c.set_clock(n=2, ndt=0, dt=dt, dts=dt, dte=dt, timo=t)
        c.step(1)


def kinetic_energy(q):
    """Mean kinetic energy of the unique points of a periodic block."""
    n = round(q.shape[1] ** (1 / 3)) - 1
    g = q.reshape(5, n + 1, n + 1, n + 1)[:, :n, :n, :n]
    return float(np.mean(0.5 * (g[1] ** 2 + g[2] ** 2 + g[3] ** 2) / g[0]))


def spectrum(q):
    """Shell-summed kinetic-energy spectrum of the velocity, k = 1..n/2."""
    n = round(q.shape[1] ** (1 / 3)) - 1
    g = q.reshape(5, n + 1, n + 1, n + 1)[:, :n, :n, :n]
    e = np.zeros(n // 2 + 1)
    k = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    shell = np.rint(kk).astype(int)
    for m in (1, 2, 3):
        f = np.fft.fftn(g[m] / g[0]) / n ** 3
        p = 0.5 * np.abs(f) ** 2
        np.add.at(e, np.minimum(shell, n // 2), p)
    return e[1:]


def norm_stats(ins, outs):
    """Per-channel mean and scale of inputs and targets ([samples, 5, points]).
    The three momentum channels share one scale (the flow has no preferred
    direction), so a component that starts at zero is not blown up."""
    in_mean = ins.mean(axis=(0, 2))
    in_var = ins.var(axis=(0, 2))
    out_var = np.mean(outs ** 2, axis=(0, 2))
    in_var[1:4] = in_var[1:4].mean()
    out_var[1:4] = out_var[1:4].mean()
    in_scale = np.maximum(np.sqrt(in_var), 1e-6 + 1e-3 * np.abs(in_mean))
    out_scale = np.maximum(np.sqrt(out_var), 1e-12)
    return in_mean, in_scale, out_scale


def main(argv=None):
    global MODES
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=256)
    ap.add_argument("--coarse", type=int, default=64)
    ap.add_argument("--runs", type=int, default=4)
    ap.add_argument("--pairs", type=int, default=500, help="pairs per run")
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-steps", type=int, default=300)
    ap.add_argument("--dt", type=float, default=0.01, help="fine time step")
    ap.add_argument("--out", default="closure_out", help="folder for the weights, losses and summary")
    ap.add_argument("--modes", type=float, default=MODES,
                    help="relative amplitude of the resolved-scale modes of each seed's initial flow")
    ap.add_argument("--alphas", default="0.1,0.25,0.5,1",
                    help="comma-separated scales of the network's correction tried on a calibration run (seed "
                         "998, its own fine truth); the judge uses the one with the lowest kinetic-energy error")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    MODES = a.modes
    nf, n = a.fine, a.coarse
    r = nf // n
    dtf, dtc = a.dt, a.dt * r
    t0 = time.perf_counter()

    # normalisation: pooled statistics of pairs sampled over a whole run of
    # a separate seed (the flow's scales grow by orders of magnitude as the
    # vortex breaks down, so the first state alone is not representative)
    fine = tgv(nf, seed=1000)
    coarse, target = tgv(n, seed=0), tgv(n, seed=0)
    every = max(1, a.pairs // 10)
    ins, outs = [], []
    for k in range(a.pairs):
        sample = k % every == 0
        if sample:
            box_filter(fine, 0, coarse, 0)
            fixed_step(coarse, dtc)
        fixed_step(fine, dtf, r)
        if sample:
            box_filter(fine, 0, target, 0)
            qc = coarse.state(0)
            ins.append(qc)
            outs.append(target.state(0) - qc)
    fine.close()
    in_mean, in_scale, out_scale = norm_stats(np.stack(ins), np.stack(outs))
    log(f"norm ({len(ins)} samples): in_mean {in_mean}, in_scale {in_scale}, out_scale {out_scale}")

    net = Net(n, capacity=a.runs * a.pairs, seed=11)
    net.set_norm(in_mean, in_scale, out_scale)
    for run in range(a.runs):
        fine = tgv(nf, seed=1 + run)
        box_filter(fine, 0, coarse, 0)
        for k in range(a.pairs):
            box_filter(fine, 0, coarse, 0)
            fixed_step(coarse, dtc)
            fixed_step(fine, dtf, r)
            box_filter(fine, 0, target, 0)
            net.add_pair(coarse, 0, target, 0)
            if k % 100 == 0:
                log(f"run {run} pair {k} t={fine.clock()['timo']:.3f} ({time.perf_counter() - t0:.0f} s)")
        fine.close()
    npairs = net.pairs
    held = max(1, a.pairs // 5)            # the last pairs of the last run: validation
    log(f"{npairs} pairs; training {a.iters} iterations")
    hist = []
    before = net.eval(npairs - held, held)
    done = 0
    while done < a.iters:
        step = min(500, a.iters - done)
        tl = net.train(step, lr=a.lr, seed=done)
        done += step
        vl = net.eval(npairs - held, held)
        hist.append(dict(iters=done, train=tl, valid=vl))
        log(f"iter {done}: train {tl:.4e} valid {vl:.4e}")

    def ke_err(series, truth):
        kt = np.array(truth)
        return float(np.sqrt(np.mean((np.array(series) - kt) ** 2)) / float(np.mean(np.abs(kt))))

    def rollout(seed, alphas, tag):
        """Coarse runs from the filtered state of `seed`: plain and with the
        network's correction scaled by each alpha, beside the filtered fine
        run.  Returns the kinetic-energy series and the final coarse cases."""
        fine = tgv(nf, seed=seed)
        plain = tgv(n, seed=0)
        nets = [tgv(n, seed=0) for _ in alphas]
        for c in [plain] + nets:
            box_filter(fine, 0, c, 0)
        ke = {"truth": [], "plain": [], "net": [[] for _ in alphas]}
        for s in range(a.eval_steps):
            fixed_step(plain, dtc)
            for al, c in zip(alphas, nets):
                fixed_step(c, dtc)
                net.set_norm(in_mean, in_scale, al * out_scale)
                net.apply(c, 0)
            fixed_step(fine, dtf, r)
            box_filter(fine, 0, target, 0)
            ke["truth"].append(kinetic_energy(target.state(0)))
            ke["plain"].append(kinetic_energy(plain.state(0)))
            for i, c in enumerate(nets):
                ke["net"][i].append(kinetic_energy(c.state(0)))
            if s % 50 == 0:
                log(f"{tag} step {s}: KE truth {ke['truth'][-1]:.6e} plain {ke['plain'][-1]:.6e} "
                    f"net {' '.join(f'{v[-1]:.6e}' for v in ke['net'])}")
        net.set_norm(in_mean, in_scale, out_scale)
        return fine, plain, nets, ke

    # the correction's scale, chosen on a calibration seed
    alphas = [float(x) for x in a.alphas.split(",")]
    calib = None
    if len(alphas) > 1:
        cf, cp, cn, cke = rollout(998, alphas, "calibration")
        calib = {"alphas": alphas, "ke_rms_rel_err_plain": ke_err(cke["plain"], cke["truth"]),
                 "ke_rms_rel_err_net": [ke_err(v, cke["truth"]) for v in cke["net"]]}
        for c in [cf, cp] + cn:
            c.close()
        alpha = alphas[int(np.argmin(calib["ke_rms_rel_err_net"]))]
        log(f"calibration: KE err plain {calib['ke_rms_rel_err_plain']:.3e}, per alpha "
            f"{dict(zip(alphas, calib['ke_rms_rel_err_net']))}; alpha {alpha}")
    else:
        alpha = alphas[0]

    # judge on a held-out seed
    fine, plain, (withnet,), ke = rollout(999, [alpha], "judge")
    ke["net"] = ke["net"][0]
    err = {m: ke_err(ke[m], ke["truth"]) for m in ("plain", "net")}
    st, sp, sn = spectrum(target.state(0)), spectrum(plain.state(0)), spectrum(withnet.state(0))
    serr = {"plain": float(np.linalg.norm(sp - st) / np.linalg.norm(st)),
            "net": float(np.linalg.norm(sn - st) / np.linalg.norm(st))}
    keep = err["net"] < err["plain"] and serr["net"] < serr["plain"]
    summary = dict(fine=nf, coarse=n, runs=a.runs, pairs=npairs, dt_fine=dtf, dt_coarse=dtc, iters=a.iters,
                   lr=a.lr, modes=a.modes, valid_before=before, history=hist, eval_steps=a.eval_steps, alpha=alpha,
                   calibration=calib,
                   t_end=float(fine.clock()["timo"]), ke_rms_rel_err=err, spectrum_rel_err=serr,
                   verdict="network kept" if keep else "fallback: plain compact filter",
                   ke_truth=ke["truth"][::10], ke_plain=ke["plain"][::10], ke_net=ke["net"][::10],
                   spectrum_truth=st.tolist(), spectrum_plain=sp.tolist(), spectrum_net=sn.tolist(),
                   seconds=time.perf_counter() - t0)
    # everything needed to rebuild the judged network:
    #   z = np.load("closure.npz"); net = Net(n); net.set_params(z["params"])
    #   net.set_norm(z["in_mean"], z["in_scale"], z["alpha"] * z["out_scale"])
    np.savez(os.path.join(a.out, "closure.npz"), params=net.params(), in_mean=in_mean, in_scale=in_scale,
             out_scale=out_scale, alpha=np.float64(alpha), coarse=np.int64(n))
    with open(os.path.join(a.out, "summary.json"), "w") as f:
        json.dump(summary, f)
    log(f"KE rms rel err plain {err['plain']:.3e} net {err['net']:.3e}; spectrum rel err plain "
        f"{serr['plain']:.3e} net {serr['net']:.3e}; {summary['verdict']}")
    print(json.dumps({k: summary[k] for k in ("pairs", "iters", "alpha", "calibration", "ke_rms_rel_err",
                                             "spectrum_rel_err", "verdict", "valid_before", "seconds")}))
    for c in (fine, plain, withnet, coarse, target, net):
        c.close()
    return summary


if __name__ == "__main__":
    main()
