"""Junction averaging as one gather and one scatter (the plan the library
uploads, built by canard_amd/csrc/cnd_scheme.cpp) gives bit for bit the
corner averaging followed by the edge averaging, with a point on several
averaged edges keeping the last edge's value, as average_line writes them
(reference src/mo_domdcomp.f90:439-559).  Host logic only: compiled with g++
and run on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "canard_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_junction_program_equals_sequential_plans(tmp_path):
    exe = tmp_path / "junction_program_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-I", SRC,
                    os.path.join(HERE, "junction_program_check.cpp"), os.path.join(SRC, "cnd_scheme.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if "mismatches" in ln]
    assert len(lines) == 4
    # the multi-block periodic cases hold points on several averaged edges
    assert any("kept" in ln and int(ln.split("edges ")[1].split()[0]) > int(ln.split("kept ")[1].split()[0])
               for ln in lines)
