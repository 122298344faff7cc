"""canard_amd: MI355X-native implementation of Canard's compressible ILES
time step (compact FD derivatives, compact filters, fluxes, metrics, RK
stages, GCBC/sponge and block interfaces) behind the C ABI of include/cnd.h.

The solver needs the in-tree libcnd.so (built by `python -m canard_amd.build`)
and a HIP device; there is no CPU fallback.
"""
from . import cases  # noqa: F401

__all__ = ["cases", "solver"]


def __getattr__(name):
    if name == "solver":
        from . import solver
        return solver
    raise AttributeError(name)
