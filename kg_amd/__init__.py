"""kg_amd -- MI355X-native engine for the Klein-Gordon method-of-lines step of FlavioRosales/KG.

The path (right-hand side on a Kerr-Schild background, 2nd-order stencils,
ghost fills, PIRK2/PIRK3/RK3/RK4 stage combines) runs as hand-written HIP
kernels for gfx950 in ``kg_amd/libkg_hip.so`` behind the C ABI of
``include/kg.h``.  ``LinesMethod`` mirrors the reference's ``type(lines_method)``
interface on top of that ABI.
"""
from ._capi import KGError, load  # noqa: F401
from .lines_method import LinesMethod  # noqa: F401

__all__ = ["LinesMethod", "KGError", "load"]
