"""Build libcnd.so (HIP, gfx950) and the Fortran driver in-tree.

    python -m canard_amd.build

The shared library is written next to this file (canard_amd/libcnd.so) so it
travels with the source tree; nothing is installed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcnd.so")
FDIR = os.path.join(HERE, "fortran")
FEXE = os.path.join(FDIR, "canard_gpu")

ARCH = os.environ.get("CND_ARCH", "gfx950")
# Every product and sum is rounded separately (-ffp-contract=off), as the
# reference evaluates them; packed FP32 instructions are kept out of the code.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value"]
SOURCES = ["cnd_scheme.cpp", "cnd_gpu.hip", "cnd_run_field.hip", "cnd_run_grad.hip", "cnd_run_flux.hip",
           "cnd_run_update.hip", "cnd_output.hip", "cnd_grid.hip",
           "cnd_closure.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_lib(verbose=False, lib=LIB, defines=()):
    """Compile the HIP sources (in parallel) into lib; defines (e.g.
    "CND_CHUNK_ETA=16") override compile-time tile constants for tuning
    builds."""
    objdir = os.path.join(HERE, "_obj")
    if lib != LIB:
        objdir = os.path.join(objdir, os.path.basename(lib).replace(".", "_"))
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(objdir, s + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmd = [hipcc()] + HIPFLAGS + ["-D" + d for d in defines] + lang + ["-c", src, "-o", obj]
        jobs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    objs = []
    failed = None
    for cmd, obj, proc in jobs:
        out, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out)
            failed = failed or cmd
        elif verbose and out.strip():
            print(out)
        objs.append(obj)
    if failed:
        raise RuntimeError("build step failed: " + " ".join(failed))
    tmp = lib + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_fortran(verbose=False):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        raise RuntimeError("amdflang not found")
    srcs = [os.path.join(FDIR, f) for f in ("mo_cnd.f90", "canard_gpu.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    _run([fc, "-O2", "-o", FEXE] + srcs + [f"-L{HERE}", "-lcnd", f"-Wl,-rpath,{HERE}"], cwd=FDIR)
    return FEXE


def build(verbose=False):
    build_lib(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
