"""ctypes binding of libcnd.so (include/cnd.h).

The library is loaded from this package directory (or from the file that
CANARD_LIB names, e.g. a build with other tile constants); if it is missing
the import of the solver fails loudly (there is no CPU fallback in the
product).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("CANARD_LIB") or os.path.join(HERE, "libcnd.so")

CND_FP64, CND_MIXED = 0, 1
# grid generators (include/cnd.h CND_GRID_*)
GRIDS = {"toy": 0, "toy_multi": 1, "sheared": 2}
# section timers (include/cnd.h CND_SEC_*)
SECTIONS = ["filter", "tstep", "grad", "stress", "flux", "gcbc", "update", "avg"]
# sweep kinds of the autotuner (include/cnd.h CND_KIND_*)
KINDS = ["filter", "grad", "flux", "field"]


class Params(C.Structure):
    """cnd_params: one instance's namelist parameters."""
    _fields_ = [("fltk", C.c_double), ("fltrbc", C.c_double), ("reoo", C.c_double), ("tempoo", C.c_double),
                ("amach1", C.c_double), ("amach2", C.c_double), ("amach3", C.c_double), ("timf", C.c_double),
                ("nsmf", C.c_int), ("cfl", C.c_double), ("viscous", C.c_int), ("sponge", C.c_int),
                ("szco", C.c_double), ("szth0", C.c_double), ("szth1", C.c_double), ("doml0", C.c_double),
                ("doml1", C.c_double)]


class Block(C.Structure):
    """cnd_block: one block of the multi-block topology."""
    _fields_ = [("lxi", C.c_int), ("let", C.c_int), ("lze", C.c_int), ("nbc", (C.c_int * 2) * 3),
                ("mcd", (C.c_int * 2) * 3), ("inst", C.c_int)]


# (name, restype, argtypes) of every symbol include/cnd.h declares
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
SIGNATURES = [
    ("cnd_create", C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(Block), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("cnd_destroy", None, [C.c_void_p]),
    ("cnd_last_error", C.c_char_p, [C.c_void_p]),
    ("cnd_set_grid", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("cnd_init_metrics", C.c_int, [C.c_void_p]),
    ("cnd_make_grid", C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int]),
    ("cnd_get_grid", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("cnd_get_metrics", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]),
    ("cnd_set_state", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("cnd_get_state", C.c_int, [C.c_void_p, C.c_int, _dp]),
    ("cnd_set_clock", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_double] * 4),
    ("cnd_get_clock", C.c_int, [C.c_void_p, C.c_int, _ip, _ip] + [_dp] * 4),
    ("cnd_step", C.c_int, [C.c_void_p, C.c_int]),
    ("cnd_filter", C.c_int, [C.c_void_p]),
    ("cnd_rk_stage", C.c_int, [C.c_void_p, C.c_int]),
    ("cnd_end_step", C.c_int, [C.c_void_p]),
    ("cnd_record", C.c_int, [C.c_void_p, C.c_int]),
    ("cnd_record_read", C.c_int, [C.c_void_p, C.c_int, _fp]),
    ("cnd_record_reset", C.c_int, [C.c_void_p]),
    ("cnd_output_schedule", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
    ("cnd_output_due", C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp]),
    ("cnd_output_write", C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int]),
    ("cnd_output_grid", C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
    ("cnd_output_wait", C.c_int, [C.c_void_p]),
    ("cnd_write_plt", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_int, _ip,
                                C.c_double, C.POINTER(_fp)]),
    ("cnd_use_graph", C.c_int, [C.c_void_p, C.c_int]),
    ("cnd_synchronize", C.c_int, [C.c_void_p]),
    ("cnd_deriv_sweep", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_dp), C.POINTER(_dp)]),
    ("cnd_filter_sweep", C.c_int, [C.c_void_p, C.c_int, C.POINTER(_dp)]),
    ("cnd_profile", C.c_int, [C.c_void_p, C.c_int]),
    ("cnd_profile_read", C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_longlong)]),
    ("cnd_tune", C.c_int, [C.c_void_p]),
    ("cnd_set_shape", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    ("cnd_shape_info", C.c_int, [C.c_int, _ip, _ip, _ip, _ip]),
    ("cnd_tuning", C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _dp]),
    ("cnd_net_param_count", C.c_int, []),
    ("cnd_net_create", C.c_int, [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]),
    ("cnd_net_destroy", None, [C.c_void_p]),
    ("cnd_net_error", C.c_char_p, [C.c_void_p]),
    ("cnd_net_get_params", C.c_int, [C.c_void_p, _fp]),
    ("cnd_net_set_params", C.c_int, [C.c_void_p, _fp]),
    ("cnd_net_set_norm", C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    ("cnd_net_add_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    ("cnd_net_add_pair_host", C.c_int, [C.c_void_p, _fp, _fp]),
    ("cnd_net_pairs", C.c_int, [C.c_void_p]),
    ("cnd_net_train", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_ulonglong, _dp]),
    ("cnd_net_eval", C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    ("cnd_net_apply", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("cnd_net_apply_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("cnd_net_forward_host", C.c_int, [C.c_void_p, _fp, _fp]),
    ("cnd_net_grad_host", C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp]),
    ("cnd_box_filter", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    ("cnd_stream", C.c_void_p, [C.c_void_p]),
    ("cnd_points", C.c_longlong, [C.c_void_p]),
    ("cnd_cells", C.c_longlong, [C.c_void_p]),
]

_lib = None


def load():
    """Load libcnd.so and bind every entry point; raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m canard_amd.build`")
        L = C.CDLL(LIBPATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
