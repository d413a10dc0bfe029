"""The C ABI: the shared object loads, exports every function include/ifgf_amd.h declares,
and refuses to compute without a GPU (there is no CPU path)."""
import ctypes
import os
import re

import numpy as np
import pytest

import ifgf_amd
from ifgf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ifgf_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ifgf_\w+)\s*\(", src)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_version_and_errors_without_device():
    L = _lib.lib()
    assert L.ifgf_version().startswith(b"ifgf_amd")
    h = ctypes.c_void_p()
    assert L.ifgf_create(99, 1.0, 0.0, 64, 6, 1, -1.0, 1, ctypes.byref(h)) == _lib.IFGF_EINVAL
    assert L.ifgf_create(0, 1.0, 0.0, 64, 6, 1, -1.0, 7, ctypes.byref(h)) == _lib.IFGF_EINVAL
    assert L.ifgf_create(0, 1.0, 0.0, 64, 6, 1, -1.0, 1, ctypes.byref(h)) == _lib.IFGF_OK
    w = np.zeros(8, complex)
    assert L.ifgf_mult(h, w.ctypes.data, 1, w.ctypes.data, 0, None) == _lib.IFGF_ESTATE
    assert L.ifgf_mult_adjoint(h, w.ctypes.data, 1, w.ctypes.data, 0, None) == _lib.IFGF_ESTATE
    assert L.ifgf_mult_noalloc(h, w.ctypes.data, w.ctypes.data, None) == _lib.IFGF_ESTATE
    assert L.ifgf_set_dx(h, 1) == _lib.IFGF_EINVAL  # not a gradient operator
    shape = np.zeros(6, np.int32)
    assert L.ifgf_autotune(h, 1, 1, shape.ctypes.data) == _lib.IFGF_ESTATE
    assert L.ifgf_set_launch_shape(h, shape.ctypes.data) == _lib.IFGF_ESTATE
    assert L.ifgf_get_launch_shape(h, shape.ctypes.data) == _lib.IFGF_OK
    assert shape.tolist() == [4, 0, 16, 0, 1, 1]  # the defaults
    if L.ifgf_device_count() == 0:
        x = np.random.default_rng(0).standard_normal((50, 3))
        rc = L.ifgf_init(h, x.ctypes.data, 50, x.ctypes.data, 50, None, _lib.CUTOFF_FN(0), None, 0)
        assert rc == _lib.IFGF_ENODEV
        assert b"no HIP device" in L.ifgf_last_error(h)
    L.ifgf_destroy(h)


def test_python_classes_mirror_reference_api():
    # constructors take the reference's positional arguments (py_ifgf.cpp:19, 46)
    ifgf_amd.MofifiedHelmholtzIfgfOperator(1 + 2j, 64, 6, 1, -1.0)
    ifgf_amd.GradHelmholtzIfgfOperator(1 + 2j, 64, 6, 1, -1.0).setDx(2)
    ifgf_amd.LaplaceIfgfOperator(64, 6, 1, -1.0)
    with pytest.raises(_lib.IfgfError):
        ifgf_amd.GradHelmholtzIfgfOperator(1 + 2j, 64, 6, 1, -1.0).setDx(3)


def test_import_pyifgf_unchanged():
    """`import pyifgf` (py_ifgf.cpp:25) works and gives the reference's classes, taking the
    reference's positional arguments; a plan-only init and the plan's orders work on the CPU."""
    import pyifgf
    for name in ("MofifiedHelmholtzIfgfOperator", "GradHelmholtzIfgfOperator",
                 "LaplaceIfgfOperator"):
        assert getattr(pyifgf, name) is getattr(ifgf_amd, name)
    op = pyifgf.MofifiedHelmholtzIfgfOperator(0.5 - 3.0j, 64, 8, 1, -1.0)
    x = np.random.default_rng(1).standard_normal((3, 500))
    op.init(x, x, plan_only=True)
    assert op.plan_orders()["base"].tolist() == [5, 8, 8]
