"""ifgf_amd: IFGF operator application on AMD MI355X (gfx950).

A from-scratch HIP implementation of TimHRO/ifgf's IfgfOperator init/mult path behind the
reference's `pyifgf` API (see ifgf_amd.pyifgf) and a C ABI (include/ifgf_amd.h).
"""
from ._lib import IfgfError  # noqa: F401
from .pyifgf import (  # noqa: F401
    CombinedFieldHelmholtzIfgfOperator,
    DoubleLayerHelmholtzIfgfOperator,
    GradHelmholtzIfgfOperator,
    HelmholtzIfgfOperator,
    LaplaceIfgfOperator,
    MergedIfgfOperator,
    MofifiedHelmholtzIfgfOperator,
    ModifiedHelmholtzIfgfOperator,
    merge,
)

__all__ = [
    "IfgfError",
    "HelmholtzIfgfOperator",
    "MofifiedHelmholtzIfgfOperator",
    "ModifiedHelmholtzIfgfOperator",
    "GradHelmholtzIfgfOperator",
    "LaplaceIfgfOperator",
    "DoubleLayerHelmholtzIfgfOperator",
    "CombinedFieldHelmholtzIfgfOperator",
    "MergedIfgfOperator",
    "merge",
]
