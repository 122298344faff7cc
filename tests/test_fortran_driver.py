"""The Fortran driver (canard_amd/fortran/canard_gpu): input.canard namelists,
grid generation, restart files and the time loop over libcnd."""
import os
import subprocess

import numpy as np
import pytest

from canard_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "canard_amd", "fortran", "canard_gpu")


def write_input(path, n=16, maxsteps=5, nts=0, tgv=True, extra="", ndata=-1, tmax=100.0):
    txt = f"""&nml_driver
nio = 0
mbk = 0
ndata = {ndata}
/
&nml_canard
nts = {nts}
nrestart = 1
cfl = 0.8
tmax = {tmax}
ltimer = .true.
/
&nml_domdcomp
nbpc(0,:) = 1,1,1
lximb = {n}
letmb = {n}
lzemb = {n}
nbbc(0,1,:) = 45,45
nbbc(0,2,:) = 45,45
nbbc(0,3,:) = 45,45
mbcd(0,1,:) = 0,0
mbcd(0,2,:) = 0,0
mbcd(0,3,:) = 0,0
/
&nml_numerics
fltk = 0.49
fltrbc = 0.9
/
&nml_physics
reoo = 16000.0
tempoo = 300.0
amach1 = 0.0
amach2 = 0.0
amach3 = 0.0
timf = 1.0
nsmf = 1
/
&nml_gridgen
gridtype = 0
lxi0 = {n}
let0 = {n}
lze0 = {n}
doml0 = 0.0
doml1 = 6.283185307179586
domh = 3.141592653589793
span = 6.283185307179586
/
&nml_gpu
viscous = 1
maxsteps = {maxsteps}
icase = '{'tgv' if tgv else 'vortex'}'
tgvmach = 0.1
{extra}/
"""
    open(os.path.join(path, "input.canard"), "w").write(txt)
    os.makedirs(os.path.join(path, "misc"), exist_ok=True)


def run(path, timeout=120):
    return subprocess.run([EXE], cwd=path, capture_output=True, text=True, timeout=timeout)


def test_driver_built():
    assert os.path.exists(EXE), "build the driver with python -m canard_amd.build"


def test_driver_reports_missing_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    write_input(str(tmp_path))
    r = run(str(tmp_path))
    assert r.returncode != 0
    assert "no HIP device" in r.stdout


def test_driver_rejects_bad_namelist(tmp_path):
    write_input(str(tmp_path), extra="bogus_key = 3\n")
    r = run(str(tmp_path))
    assert r.returncode != 0
    assert "error reading nml_gpu" in r.stdout


@pytest.mark.gpu
def test_driver_matches_library_and_restarts_bitwise(tmp_path):
    """5 steps through the Fortran driver equal 5 steps through the Python
    binding, and 3 steps + restart + 2 steps equal 5 straight steps."""
    from canard_amd.solver import Case
    n = 16
    a = tmp_path / "straight"
    a.mkdir()
    write_input(str(a), n=n, maxsteps=5)
    r = run(str(a))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    npt = (n + 1) ** 3
    q_f, clock_f = cases.read_restart(str(a / "misc" / "restartz000.dat"), npt)
    assert clock_f["n"] == 5

    xyz = cases.gridtoy(n, n, n)
    q0 = cases.tgv_state(xyz, (n + 1,) * 3, mach=0.1)
    c = Case([cases.periodic_block(n)], dict(reoo=16000.0))
    c.set_grids([xyz])
    c.set_states([q0])
    c.set_clock()
    c.step(5)
    q_p = c.state(0)
    cp = c.clock()
    c.close()
    # the driver evaluates the TGV formula in Fortran: same state to rounding
    assert np.max(np.abs(q_f - q_p)) < 1e-13
    assert abs(cp["timo"] - clock_f["timo"]) < 1e-13

    b = tmp_path / "restart"
    b.mkdir()
    write_input(str(b), n=n, maxsteps=3)
    assert run(str(b)).returncode == 0
    write_input(str(b), n=n, maxsteps=5, nts=1)
    r = run(str(b))
    assert r.returncode == 0, r.stdout
    q_r, clock_r = cases.read_restart(str(b / "misc" / "restartz000.dat"), npt)
    assert clock_r == clock_f
    assert np.array_equal(q_r, q_f)


@pytest.mark.gpu
def test_driver_writes_tecplot_output(tmp_path):
    """With ndata >= 0 the driver writes data/grid.plt at set-up and
    data/nNNN.plt whenever a step crosses k tmax / ndata, from the recorded
    averages (asynchronously); the files hold what the Python binding's
    recording of the same run gives."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from plt_reader import read_plt
    from canard_amd.solver import Case
    n = 16
    write_input(str(tmp_path), n=n, maxsteps=12, ndata=40, tmax=4.0)
    r = run(str(tmp_path))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    xyz = cases.gridtoy(n, n, n)
    g = read_plt(str(tmp_path / "data" / "grid.plt"))
    assert np.array_equal(g["zones"][0]["data"], xyz.astype(np.float32))
    files = sorted(f for f in os.listdir(tmp_path / "data") if f.startswith("n"))
    assert len(files) >= 2 and files[0] == "n000.plt"
    c = Case([cases.periodic_block(n)], dict(reoo=16000.0))
    c.set_grids([xyz])
    c.set_states([cases.tgv_state(xyz, (n + 1,) * 3, mach=0.1)])
    c.set_clock()
    c.record(True)
    c.output_schedule(4.0, 40)
    got = {}
    for _ in range(12):
        c.step(1)
        due, nd, t = c.output_due()
        if due:
            got[f"n{nd:03d}.plt"] = (t, c.record_read(0))
            c.output_write(tmp_path / "py.plt")
    c.close()
    assert sorted(got) == files
    for f in files:
        z = read_plt(str(tmp_path / "data" / f))["zones"][0]
        t, rec = got[f]
        assert abs(z["time"] - t) < 1e-13
        assert np.max(np.abs(z["data"] - rec)) <= 1e-6 * np.max(np.abs(rec))
