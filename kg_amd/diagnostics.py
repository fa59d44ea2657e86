"""On-device diagnostics of a grid's state (MoL/diagnostic_lib.f90:185-274).

The reference calls ``diagnostic(l)`` after every step (main/evolve.f90:26):
every step it integrates the energy density (``Mass_scalar``, :207); every
``save`` steps it forms the ADM currents and, at detector radii, their fluxes
through the sphere and the spherical-harmonic coefficients of phi (:197-199).
Here those run on the GPU where the state lives:

- ``integrate(lm, quantity)``: ``integrate`` (tools/integral.f90:646-669) of a
  state field (0..4) or of ``density() * sqrtg`` ("density");
- ``detectors(projector, index_det)``: ``currents_detectors`` and
  ``proyector_detectors`` at the given spectral radii;
- ``detector_radii`` / ``detector_indices``: the reference's detector radii
  (:124-134) and the spectral radius index of each (the reference allocates
  ``index_det`` but never assigns it; here it is the last spectral radius at
  or below the detector).

Note: in the reference the ``call this%density()`` line is commented out, so
its ``Mass_scalar`` integrates an array that is never set; ``integrate(lm,
"density")`` integrates what ``density()`` defines.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _capi as capi

QUANTITIES = {"phi": 0, "psi_r": 1, "psi_theta": 2, "psi_phi": 3, "Pi": 4, "density": capi.KG_DIAG_DENSITY}


def integrate(lm, quantity="density"):
    """Volume integral over r in [x_0, x_Nx], theta rows 0..Ny, all phi planes.
    Returns a float, or a complex for a field of a complex state."""
    q = QUANTITIES[quantity] if isinstance(quantity, str) else int(quantity)
    out = (C.c_double * 2)()
    lib = capi.load()
    st = lib.kg_diag_integrate(lm._h, q, out)
    if st != capi.KG_OK:
        raise capi.KGError(st, lib.kg_diag_last_error().decode())
    if lm.is_complex and q != capi.KG_DIAG_DENSITY:
        return complex(out[0], out[1])
    return float(out[0])


def detector_radii(n_det, a=0.0, xmax=60.0, metric="Kerr_Schild"):
    """r_det of scalar_field_constructor (diagnostic_lib.f90:124-134)."""
    r = np.zeros(n_det)
    r[0] = 1.0 + math.sqrt(1.0 - a * a)
    r[n_det - 1] = math.exp(xmax) if metric == "MKS" else xmax * 0.99
    for n in range(2, n_det):
        r[n - 1] = r[0] + (r[n_det - 1] - r[0]) / float(n_det) * float(n - 1)
    return r


def detector_indices(r_det, r_spec):
    """For each detector inside the spectral grid (r_1 <= r_det < r_Nx, the
    reference's test at :243-244) the last spectral radius at or below it;
    detectors outside are dropped.  Returns (indices, kept detector numbers)."""
    r_spec = np.asarray(r_spec)
    idx, keep = [], []
    for q, rd in enumerate(r_det):
        if r_spec[0] <= rd < r_spec[-1]:
            idx.append(int(np.searchsorted(r_spec, rd, side="right") - 1))
            keep.append(q)
    return np.asarray(idx, dtype=np.int32), keep


def detectors(projector, index_det):
    """(momentum complex [n_det][3], alm complex [n_det][nlm]) at the spectral
    radius indices index_det."""
    idx = np.ascontiguousarray(index_det, dtype=np.int32)
    n = len(idx)
    mom = np.zeros((n, 3, 2))
    alm = np.zeros((n, projector.nlm, 2))
    lib = capi.load()
    st = lib.kg_spec_detectors(projector._p, n, idx.ctypes.data, mom.ctypes.data, alm.ctypes.data)
    projector._check(st, projector._p)
    return mom[..., 0] + 1j * mom[..., 1], alm[..., 0] + 1j * alm[..., 1]
