"""ctypes binding of libifgf_amd.so (include/ifgf_amd.h).

This is the stub a maintainer of the reference's Python side would add: every entry point
of the C ABI with its argument types.  The library has no CPU path for mult; loading fails
loudly when the shared object is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libifgf_amd.so")

IFGF_OK = 0
IFGF_EINVAL = -1
IFGF_ENODEV = -2
IFGF_ESTATE = -3
IFGF_EINTERNAL = -4

HELMHOLTZ, MODHELMHOLTZ, LAPLACE, DOUBLELAYER, COMBINED, GRAD = range(6)
PREC_FP64, PREC_MIXED = 0, 1
INIT_PLAN_ONLY = 1
INIT_HOST_PLAN = 2
INIT_LEVEL_RING = 4
DEVICE_POINTERS = 1
MULT_GRAPH = 2

CUTOFF_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_double, ctypes.c_void_p)

# name -> (restype, argtypes): the exported symbols of include/ifgf_amd.h
_P, _D, _I, _I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64
SIGNATURES = {
    "ifgf_create": (_I, [_I, _D, _D, _I64, _I, _I, _D, _I, ctypes.POINTER(_P)]),
    "ifgf_set_dx": (_I, [_P, _I]),
    "ifgf_init": (_I, [_P, _P, _I64, _P, _I64, _P, CUTOFF_FN, _P, _I]),
    "ifgf_plan_save": (_I, [_P, ctypes.c_char_p]),
    "ifgf_init_from_plan": (_I, [_P, ctypes.c_char_p, _I]),
    "ifgf_mult": (_I, [_P, _P, _I64, _P, _I, _P]),
    "ifgf_mult_noalloc": (_I, [_P, _P, _P, _P]),
    "ifgf_mult_adjoint": (_I, [_P, _P, _I64, _P, _I, _P]),
    "ifgf_merge": (_I, [_P, _I, _I, ctypes.POINTER(_P)]),
    "ifgf_sizes": (_I, [_P, _P, _P]),
    "ifgf_destroy": (None, [_P]),
    "ifgf_last_error": (ctypes.c_char_p, [_P]),
    "ifgf_version": (ctypes.c_char_p, []),
    "ifgf_device_count": (_I, []),
    "ifgf_plan_counts": (_I, [_P, _P]),
    "ifgf_plan_orders": (_I, [_P, _P]),
    "ifgf_plan_tree": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ifgf_plan_pairs": (_I, [_P, _I, _P, _P]),
    "ifgf_plan_cones": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ifgf_plan_work": (_I, [_P, _P]),
    "ifgf_last_timings": (_I, [_P, _P]),
    "ifgf_autotune": (_I, [_P, _I64, _I, _P]),
    "ifgf_get_launch_shape": (_I, [_P, _P]),
    "ifgf_set_launch_shape": (_I, [_P, _P]),
}

_lib = None


def _load_torch_first():
    """torch ships its own HIP runtime; it must be loaded before this library's runtime is
    initialised, otherwise torch cannot initialise HIP in the same process."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the in-tree shared object (raises if it has not been built)."""
    global _lib
    if _lib is None:
        _load_torch_first()
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m ifgf_amd.build` "
                "(ifgf_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class IfgfError(RuntimeError):
    pass


def check(rc, handle=None, what="call"):
    if rc != IFGF_OK:
        msg = lib().ifgf_last_error(handle).decode() if handle else ""
        raise IfgfError(f"{what} failed ({rc}): {msg}")
