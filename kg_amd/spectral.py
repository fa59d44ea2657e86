"""Spherical-Bessel spectral projection of a grid's state on the GPU.

Mirrors the set-up and projection routines of the reference's
tools/FFT_sph.f90 as ``scalar_field_constructor`` uses them
(MoL/diagnostic_lib.f90:144-149,239-247):

- ``sph_harmonics`` (FFT_sph.f90:282-336) and ``generate_full_bessel_basis``
  (:13-95) run once, when a ``SpectralProjector`` is created;
- ``domain_to_sph`` (:525-651) interpolates a state field onto the spectral
  grid r_i = xmin + i dx, theta_j = j dtheta, phi_k = k dphi;
- ``project_sh`` (:247-279) gives f_lm(r_i) at every radius;
- ``radial`` is the contraction ``SphFT_discrete_radial`` (:213-244) defines,
  ft(n, l, m) = trapezoid_i(f_lm(r_i) j_l(k_n r_i) r_i^2, dr), in FP64 or on
  BF16 matrix cores with split operands (``precision="split_bf16"``).

Arrays: lm = l*l + l + m.  ``project_sh`` returns complex [lm][i];
``radial`` returns complex [lm][n]; ``domain_to_sph`` returns real
[part][k][j][i] (part = real[, imaginary]).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

_BASES = {"zeros": capi.KG_BASIS_ZEROS, "uniform": capi.KG_BASIS_UNIFORM}
_PRECISIONS = {"fp64": capi.KG_SPEC_FP64, "split_bf16": capi.KG_SPEC_SPLIT_BF16}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class SpectralProjector:
    """Y_lm and Bessel tables of one ``LinesMethod`` grid, resident on its GPU.

    basis="zeros": k_{l,n} = z_n(J_{l+1/2}) / R_max, j_l normalised as in
    generate_full_bessel_basis; basis="uniform": k_n = kmax n / nk.
    """

    def __init__(self, lm, lmax=8, nk=32, basis="zeros", R_max=None, kmax=4.0):
        if basis not in _BASES:
            raise capi.KGError(capi.KG_ERR_INVALID, "unknown basis %r" % basis)
        self._lib = capi.load()
        self.lm = lm
        d = capi.KGSpecDesc()
        d.lmax, d.nk, d.basis = lmax, nk, _BASES[basis]
        d.R_max = float(R_max if R_max is not None else lm.desc.xmax)
        d.kmax = float(kmax)
        p = C.c_void_p()
        st = self._lib.kg_spec_create(lm._h, C.byref(d), C.byref(p))
        self._check(st, None)
        self._p = p
        self.lmax, self.nk = lmax, nk
        self.nlm = (lmax + 1) ** 2
        self.Nr, self.Nt, self.Np = lm.Nx, lm.Ny, lm.Nz

    def _check(self, st, p):
        if st != capi.KG_OK:
            msg = self._lib.kg_spec_last_error(p)
            raise capi.KGError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_p", None):
            self._lib.kg_spec_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- tables
    def tables(self):
        """(Y_lm complex [lm][theta][phi], k_values [l][n], j_basis [l][n][i])."""
        ylm = np.zeros((self.nlm, self.Nt, self.Np, 2))
        kv = np.zeros((self.lmax + 1, self.nk))
        jb = np.zeros((self.lmax + 1, self.nk, self.Nr))
        self._check(self._lib.kg_spec_get_tables(self._p, _ptr(ylm), _ptr(kv), _ptr(jb)), self._p)
        return ylm[..., 0] + 1j * ylm[..., 1], kv, jb

    # ------------------------------------------------------------- projections
    def domain_to_sph(self, field=0):
        out = np.zeros((self.lm.nparts, self.Np, self.Nt, self.Nr))
        self._check(self._lib.kg_spec_domain_to_sph(self._p, field, _ptr(out), 0), self._p)
        return out

    def project_sh(self, field=0):
        out = np.zeros((self.nlm, self.Nr, 2))
        self._check(self._lib.kg_spec_project_sh(self._p, field, _ptr(out), 0), self._p)
        return out[..., 0] + 1j * out[..., 1]

    def radial(self, field=0, precision="fp64", out_device_ptr=None):
        """Spherical-Bessel spectrum complex [lm][n].  With out_device_ptr (a
        device address of nlm*nk*2 doubles) the result stays on the GPU and the
        call returns None without waiting."""
        prec = _PRECISIONS[precision]
        if out_device_ptr is not None:
            self._check(self._lib.kg_spec_radial(self._p, field, prec, C.c_void_p(out_device_ptr), 1), self._p)
            return None
        out = np.zeros((self.nlm, self.nk, 2))
        self._check(self._lib.kg_spec_radial(self._p, field, prec, _ptr(out), 0), self._p)
        return out[..., 0] + 1j * out[..., 1]

    def radial_from_flm(self, flm, precision="fp64"):
        """The radial contraction alone on host f_lm (complex [lm][i])."""
        flm = np.asarray(flm, dtype=np.complex128)
        if flm.shape != (self.nlm, self.Nr):
            raise capi.KGError(capi.KG_ERR_INVALID, "flm must be [%d][%d]" % (self.nlm, self.Nr))
        src = np.ascontiguousarray(np.stack([flm.real, flm.imag], axis=-1))
        out = np.zeros((self.nlm, self.nk, 2))
        self._check(self._lib.kg_spec_radial_flm(self._p, _ptr(src), _PRECISIONS[precision], _ptr(out)), self._p)
        return out[..., 0] + 1j * out[..., 1]


def radial_batch(projectors, field=0, precision="split_bf16", out_device_ptrs=None):
    """``radial`` of many projectors (grids of any size, all lmax <= 8 or all
    > 8) with one launch per pass over all of them (``kg_spec_radial_batch``):
    a list of complex [lm][n] arrays, equal bit for bit to each projector's
    own ``radial``.  With out_device_ptrs (one device address of nlm*nk*2
    doubles per projector) the results stay on the GPU and the call returns
    None without waiting."""
    projectors = list(projectors)
    n = len(projectors)
    if n == 0:
        return []
    lib = projectors[0]._lib
    ps = (C.c_void_p * n)(*[p._p for p in projectors])
    prec = _PRECISIONS[precision]
    if out_device_ptrs is not None:
        outs = (C.c_void_p * n)(*[C.c_void_p(int(a)) for a in out_device_ptrs])
        projectors[0]._check(lib.kg_spec_radial_batch(ps, n, field, prec, outs, 1), projectors[0]._p)
        return None
    arrays = [np.zeros((p.nlm, p.nk, 2)) for p in projectors]
    outs = (C.c_void_p * n)(*[a.ctypes.data for a in arrays])
    projectors[0]._check(lib.kg_spec_radial_batch(ps, n, field, prec, outs, 0), projectors[0]._p)
    return [a[..., 0] + 1j * a[..., 1] for a in arrays]
