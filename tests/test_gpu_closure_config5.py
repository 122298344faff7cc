"""The learned closure on its benchmark configuration, reduced: 256^3 FP64
Taylor-Green runs box-filtered to 64^3 give the training pairs; the network
is trained on the device by the library's own forward and backward
kernels; its correction's scale is calibrated on one seed and the judge
compares coarse runs with and without it against the filtered fine run of
another seed (kinetic energy every step, spectrum at the end).  Sweep
shapes are pinned (CANARD_TUNE=0) and every seed is fixed.

What the test holds to: training lowers the held-out one-step error; the
calibration and judge flows differ (their plain coarse runs' kinetic-energy
errors differ); the judge's verdict is the rule it states (the network is
kept only if it lowers both errors); and every error is printed."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_closure_config5_reduced(tmp_path, monkeypatch):
    from canard_amd import closure_train
    monkeypatch.setenv("CANARD_TUNE", "0")
    argv = ["--fine", "256", "--coarse", "64", "--runs", "1", "--pairs", "120", "--iters", "600",
            "--eval-steps", "60", "--alphas", "0.5,1", "--out", str(tmp_path)]
    print("[config5] closure_train", " ".join(argv), flush=True)
    s = closure_train.main(argv)
    err, serr = s["ke_rms_rel_err"], s["spectrum_rel_err"]
    print(f"  validation MSE {s['valid_before']:.3e} -> {s['history'][-1]['valid']:.3e}", flush=True)
    print(f"  calibration: {s['calibration']}", flush=True)
    print(f"  judge: KE error plain {err['plain']:.3e} net {err['net']:.3e}; "
          f"spectrum error plain {serr['plain']:.3e} net {serr['net']:.3e}; {s['verdict']}", flush=True)
    assert s["history"][-1]["valid"] < s["valid_before"]
    assert all(math.isfinite(v) for v in list(err.values()) + list(serr.values()))
    assert s["calibration"]["ke_rms_rel_err_plain"] != err["plain"]
    keep = err["net"] < err["plain"] and serr["net"] < serr["plain"]
    assert s["verdict"] == ("network kept" if keep else "fallback: plain compact filter")
    assert (tmp_path / "closure.npz").exists()
