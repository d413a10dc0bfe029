"""`pyifgf`: the module name TimHRO/ifgf's pybind11 bindings export (py_ifgf.cpp:25), so code
written against the reference (bempp-cl's IFGF assembler) runs unchanged with

    import pyifgf
    op = pyifgf.MofifiedHelmholtzIfgfOperator(k, leafSize, order, n_elem, tol)
    op.init(srcs, targets)
    u = op.mult(weights)

The classes are those of ifgf_amd.pyifgf (the HIP library behind the C ABI): the three the
reference exports (MofifiedHelmholtzIfgfOperator [sic], GradHelmholtzIfgfOperator with setDx,
LaplaceIfgfOperator, py_ifgf.cpp:16-48) and the Helmholtz, double-layer and combined-field
operators of its C++ API.
"""
from ifgf_amd.pyifgf import (  # noqa: F401
    CombinedFieldHelmholtzIfgfOperator,
    DoubleLayerHelmholtzIfgfOperator,
    GradHelmholtzIfgfOperator,
    HelmholtzIfgfOperator,
    LaplaceIfgfOperator,
    MofifiedHelmholtzIfgfOperator,
    ModifiedHelmholtzIfgfOperator,
)

__all__ = [
    "MofifiedHelmholtzIfgfOperator",
    "GradHelmholtzIfgfOperator",
    "LaplaceIfgfOperator",
    "HelmholtzIfgfOperator",
    "ModifiedHelmholtzIfgfOperator",
    "DoubleLayerHelmholtzIfgfOperator",
    "CombinedFieldHelmholtzIfgfOperator",
]
