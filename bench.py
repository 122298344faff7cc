"""Benchmark of the plain FP64 ILES time step on one MI355X.

Workload (default): single triply periodic block of 256^3 cells (257^3
stored points), Taylor-Green vortex at Mach 0.1, Re = 1600, seeded uniform
velocity perturbation of amplitude 1e-3, viscous, RK4 with cfl = 0.8, FP64.
Synthetic inputs are generated on the host and uploaded before timing; the
timed region is K whole time steps (filter phase + 4 RK stages each) with the
state resident in HBM.

    python bench.py [--steps K] [--warmup W] [--n 256] [--precision fp64|mixed] [--compare-mixed 0|1]

Prints one JSON line.  value = nominal cells x 4 RK stages x K / seconds.
With --compare-mixed 1 (the default for an FP64 run) the same workload is
then run in mixed precision (FP32 line solves with FP64 refinement, FP32
gradients) for the same W + K steps; the line's "mixed" object
gives its rate, its speed-up over FP64 and the max-norm relative error of
each conserved variable against the FP64 state after W + K steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s per RK step + % HBM peak; mixed vs plain FP64 at stated error"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s (AMD public figure)
STEP_BYTES_MODEL = 5048.0  # compulsory FP64 bytes per stored point per viscous step

# Algorithmic HBM bytes per stored point of one launch of each kernel group
# (reads + writes of FP64 fields the kernel must touch):
#   filter sweep (5 variables, one direction): read 5 + write 5 fields
#   gradient sweeps (u,v,w,T, three directions): the xi sweep reads q (5)
#                   and writes the 4 gradients and the 4 primitives, the
#                   eta and zeta sweeps read the primitives (4) and write
#                   4 gradients: 232 B over three launches, 77.33 per launch
#   stress + flux assembly: read q 5, 12 gradients, 9 metrics + Jacobian;
#                           write the fluxes of all three directions (15)
#   flux sweep (one direction): read the 5 fluxes of d, write its own
#                               right-hand-side buffer (5)
#   update: read qo 5, the three right-hand-side buffers (15), Jacobian;
#           write qa 5.  Viscous stages without characteristic faces run it
#           in the xi flux sweep (the "flux" group then has the eta and zeta
#           launches): read the 5 xi fluxes, the eta and zeta buffers (10),
#           qo 5, Jacobian; write qa 5 -- the same 208 B
SECTION_BYTES = {"filter": 80.0, "grad": 232.0 / 3.0, "stress": 336.0, "flux": 80.0, "update": 208.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay whole steps as a hipGraph (section timers then stay empty)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--cpu-n", type=int, default=0, help="CPU baseline grid (default: same as --n)")
    ap.add_argument("--compare-mixed", type=int, default=1,
                    help="FP64 runs: also time mixed precision and report its error against FP64")
    ap.add_argument("--compare-ensemble", type=int, default=1,
                    help="FP64 runs: also time Taylor-Green ensembles (32 instances of 32..128^3 cells, 64 of "
                         "24..32^3) batched in one context against one context per instance")
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--shapes", default="", metavar="JSON",
                    help='pin sweep shapes instead of autotuning, e.g. \'{"grad": [7, 6, 6]}\' '
                         '(shape index per direction, -1: tiled; see canard_amd.solver.shapes())')
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run 3 device-to-device copies of MIB MiB (a known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run)")
    return ap.parse_args()


def traffic_from_profiles(section, n):
    """HBM bytes per launch of `section` from the committed rocprofv3 PMC
    summary (profiles/traffic_tgv<n>.md), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_tgv{n}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| " + section + " |"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def host_description():
    """The host the CPU baseline runs on: logical CPUs of the machine and of
    this process's affinity mask, the CPU model (lscpu), and the OpenMP
    thread settings of the environment."""
    import platform
    import subprocess
    model = ""
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "model": model or platform.processor(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")}


def cpu_baseline(n, steps):
    """Times the CPU oracle (C restatement of the reference, OpenMP over
    pencils) on the same TGV workload, on this host's cores: as many OpenMP
    threads as OMP_NUM_THREADS names (the CPU share a job is given), else one
    per CPU of the process's affinity mask.  Steps 0 and 1 of a run have
    dt = 0 but do the full work, so every timed step costs the same."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=False)
    from helpers import oracle_for, tgv_case
    host = host_description()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host["affinity_cpus"] or os.cpu_count() or 1
    blocks, grids, states, params = tgv_case(n, pert=1e-3)
    o = oracle_for(blocks, grids, states, params, threads=threads)
    o.run(1)   # first touch of the oracle's arrays
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    o.close()
    cells = n ** 3
    return {"value": cells * 4 * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} time steps of the {n}^3 TGV (4 RK stages each) after one untimed step, on "
                      f"{threads} OpenMP threads, {dt:.1f} s",
            "host": host}


def ensemble_compare(steps, warmup, count=32, sizes=(32, 48, 64, 96, 128)):
    """A Taylor-Green ensemble (cases.config4: count instances of n^3 cells,
    n drawn from sizes, Mach and Re drawn per instance), FP64, each step
    replayed as a hipGraph.  Batched: every instance in one context, one
    launch per kernel over all of them.  Separate: one context (and stream)
    per instance, all queued before any is waited for.  Batched mixed: the
    batched context in mixed precision (FP32 line solves with FP64
    refinement).  All are autotuned and warmed up before the timed steps."""
    from canard_amd import cases
    from canard_amd.solver import CND_FP64, CND_MIXED, Case
    blocks, grids, states, params, draws = cases.config4(count=count, sizes=sizes)
    cells = sum(d["n"] ** 3 for d in draws)

    def prepare(blks, grds, sts, prms, precision=CND_FP64):
        c = Case(blks, prms, precision=precision)
        c.set_grids(grds)
        c.set_states(sts)
        for i in range(len(prms)):
            c.set_clock(inst=i)
        c.use_graph(True)
        c.step(warmup)
        return c

    def timed(ctxs):
        for c in ctxs:
            c.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for c in ctxs:
                c.step(1)
        for c in ctxs:
            c.synchronize()
        return time.perf_counter() - t0

    print("[bench] ensemble: batched", file=sys.stderr, flush=True)
    bat = prepare(blocks, grids, states, params)
    dtb = timed([bat])
    print("[bench] ensemble: one context per instance", file=sys.stderr, flush=True)
    sep = [prepare([dict(blocks[i], inst=0, mcd=[[0, 0]] * 3)], [grids[i]], [states[i]], [params[i]])
           for i in range(len(blocks))]
    dts = timed(sep)
    diff = 0.0
    for i, c in enumerate(sep):
        a, b = bat.state(i), c.state(0)
        diff = max(diff, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
        c.close()
    print("[bench] ensemble: batched, mixed precision", file=sys.stderr, flush=True)
    mix = prepare(blocks, grids, states, params, precision=CND_MIXED)
    dtm = timed([mix])
    merr = 0.0
    for i in range(len(blocks)):
        a, b = mix.state(i), bat.state(i)
        merr = max(merr, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
    mix.close()
    bat.close()
    return {"workload": f"taylor_green_ensemble_{count}x({min(sizes)}..{max(sizes)})^3_periodic_viscous_rk4",
            "instances": len(blocks),
            "cells": cells, "steps": steps, "graph": True,
            "batched": {"value": cells * 4 * steps / dtb, "ms_per_step": dtb / steps * 1e3},
            "separate": {"value": cells * 4 * steps / dts, "ms_per_step": dts / steps * 1e3},
            "speedup_batched_vs_separate": dts / dtb,
            "max_rel_diff_batched_vs_separate": diff,
            "batched_mixed": {"value": cells * 4 * steps / dtm, "ms_per_step": dtm / steps * 1e3,
                              "speedup_vs_batched_fp64": dtb / dtm, "max_rel_err_vs_batched_fp64": merr},
            "note": "value in cell-updates/s per RK stage over all instances; the two runs may pick different "
                    "sweep shapes, so their states agree to rounding, not bitwise"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the library's step runs on one GPU per "
              "process and does not shard across GPUs.", file=sys.stderr)
        sys.exit(2)
    from canard_amd.solver import CND_FP64, CND_MIXED, Case

    if a.copy_calibration > 0:
        import torch
        nbytes = a.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: 3 copies of {nbytes} bytes", file=sys.stderr, flush=True)
        del src, dst

    n = a.n
    from canard_amd import cases
    blocks = [cases.periodic_block(n)]
    xyz = cases.gridtoy(n, n, n, doml0=0.0, doml1=2 * np.pi, domh=np.pi, span=2 * np.pi)
    q0 = cases.tgv_state(xyz, (n + 1, n + 1, n + 1), mach=0.1, pert=1e-3, seed=20261202)
    params = dict(fltk=0.49, fltrbc=0.9, reoo=1600.0 / 0.1, tempoo=300.0, cfl=0.8, viscous=1)
    prec = CND_FP64 if a.precision == "fp64" else CND_MIXED

    tuning = []

    def run(precision):
        """W warm-up steps, then K timed steps; returns (seconds, section
        timers, clock, final state, stored points, cells)."""
        c = Case(blocks, params, precision=precision)
        c.set_grids([xyz])
        c.set_states([q0])
        c.set_clock()
        c.use_graph(bool(a.graph))
        for kind, per_dir in (json.loads(a.shapes).items() if a.shapes else []):
            for d, sh in enumerate(per_dir):
                c.set_shape(kind, d, int(sh))
        c.step(a.warmup)
        c.synchronize()
        tuning.append(c.tuning())
        print(f"[bench] warm-up done, timing {a.steps} steps", file=sys.stderr, flush=True)
        c.profile(True)
        t0 = time.perf_counter()
        c.step(a.steps)
        c.synchronize()
        dt = time.perf_counter() - t0
        sections = c.profile_read()
        c.profile(False)
        clock = c.clock()
        q = c.state(0)
        points, cells = c.points, c.cells
        c.close()
        return dt, sections, clock, q, points, cells

    print(f"[bench] {n}^3 TGV, {a.precision}: set-up", file=sys.stderr, flush=True)
    dt, sections, clock, q, points, cells = run(prec)
    finite = bool(np.all(np.isfinite(q)))
    mixed = None
    if prec == CND_FP64 and a.compare_mixed:
        print(f"[bench] {n}^3 TGV, mixed: set-up", file=sys.stderr, flush=True)
        dtm, msec, _, qm, _, _ = run(CND_MIXED)
        err = []
        for m in range(5):   # q is (5, points)
            ref = np.max(np.abs(q[1:4])) if 1 <= m <= 3 else np.max(np.abs(q[m]))
            err.append(float(np.max(np.abs(qm[m] - q[m])) / ref))
        mixed = {"value": cells * 4 * a.steps / dtm, "ms_per_step": dtm / a.steps * 1e3,
                 "speedup_vs_fp64": dt / dtm, "steps_compared": a.warmup + a.steps,
                 "max_rel_err_per_variable": err, "error_target": 1e-6,
                 "sections_ms_per_step": {k: v[0] / a.steps for k, v in msec.items()},
                 "note": "FP32 line solves with one FP64 residual refinement per sweep and FP32 velocity and "
                         "temperature gradients; state, metrics, fluxes, right-hand sides, RK update and GCBC in "
                         "FP64. Error: max-norm over the field after warmup + steps steps against the FP64 run, "
                         "momentum components against the largest momentum component."}
        del qm
    value = cells * 4 * a.steps / dt
    # dominant kernel group by measured time
    timed = {k: v for k, v in sections.items() if k in SECTION_BYTES and v[1] > 0}
    roof = None
    if timed:
        dom = max(timed, key=lambda k: timed[k][0])
        ms, launches = timed[dom]
        avg_s = ms / 1e3 / launches
        bytes_launch = SECTION_BYTES[dom] * points
        ach = bytes_launch / avg_s
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": traffic_from_profiles(dom, n),
                "bytes_per_launch": bytes_launch, "avg_launch_ms": avg_s * 1e3}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == CND_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"taylor_green_{n}^3_periodic_viscous_rk4", "cells": cells,
                   "stored_points": points, "mach": 0.1, "re": 1600.0, "perturbation": 1e-3,
                   "precision": a.precision, "graph": bool(a.graph),
                   "tuning": {k: [r["shape"] for r in v] for k, v in tuning[0].items()}},
        "hbm_pct_step_model": 100.0 * STEP_BYTES_MODEL * points * a.steps / dt / HBM_PEAK,
        "roofline": roof,
        "sections_ms_per_step": {k: v[0] / a.steps for k, v in sections.items()},
        "final_clock": clock,
        "finite": finite,
    }
    if mixed is not None:
        out["mixed"] = mixed
    if prec == CND_FP64 and a.compare_ensemble:
        out["ensemble"] = ensemble_compare(a.ensemble_steps, 3)
        out["ensemble_small"] = ensemble_compare(a.ensemble_steps, 3, count=64, sizes=(24, 32))
    print(f"[bench] GPU {dt:.2f} s; CPU baseline", file=sys.stderr, flush=True)
    if a.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(a.cpu_n or n, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
