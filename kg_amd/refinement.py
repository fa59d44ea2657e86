"""Two fixed refinement levels (base grid + a 2:1 radially refined patch)
advanced Berger-Oliger style on the GPU (``kg_refined_*`` in include/kg.h).

The reference has no refinement levels (its mesh_refinement_lib.f90:16-147
is the MPI slab decomposition of one uniform grid); this is new.  ``base``
and ``fine`` are ``LinesMethod`` views of the two levels, through which the
initial data are set and results read; ``step`` advances both.
"""
from __future__ import annotations

import ctypes as C

from . import _capi as capi
from .lines_method import LinesMethod, make_desc


def _check(lib, st, r):
    if st != capi.KG_OK:
        msg = lib.kg_refined_last_error(r)
        raise capi.KGError(st, msg.decode() if msg else "")


class RefinedLinesMethod:
    """A base grid (the arguments of LinesMethod) with a refined patch on
    [xmin, r_patch]: dr/2, dt/2, two fine steps per base step, a buffer of
    `buffer` fine columns beyond r_patch refilled from the base level."""

    def __init__(self, r_patch, buffer=16, **kw):
        self._lib = lib = capi.load()
        args = dict(nvars=5, g_pts=2, xmin=1.8, xmax=60.0, Nx=256, Ny=33, Nz=65, tmin=0.0, CFL=0.25,
                    boundary_type_rmax="outflow", metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.0, scheme="pirk2",
                    dt=0.0, is_complex=False, operators="ref2")
        args.update(kw)
        d = make_desc(**args)
        rd = capi.KGRefineDesc()
        rd.r_patch, rd.buffer = r_patch, buffer
        r = C.c_void_p()
        _check(lib, lib.kg_refined_create(C.byref(d), C.byref(rd), C.byref(r)), None)
        self._r = r
        hb, hf = C.c_void_p(), C.c_void_p()
        _check(lib, lib.kg_refined_levels(r, C.byref(hb), C.byref(hf)), r)
        self.base = LinesMethod(_handle=hb, **args)
        fine_args = dict(args)
        fine_args.update(Nx=self._fine_dims(hf)[0], boundary_type_rmax="inflow")
        self.fine = LinesMethod(_handle=hf, **fine_args)

    def _fine_dims(self, h):
        dims = (C.c_int64 * 8)()
        capi.check(self._lib.kg_get_dims(h, dims), h)
        return [int(v) for v in dims]

    @property
    def t(self) -> float:
        return self.base.t

    def step(self, nsteps: int = 1) -> None:
        """nsteps base steps (each with two fine steps)."""
        _check(self._lib, self._lib.kg_refined_step(self._r, int(nsteps)), self._r)

    def close(self):
        if getattr(self, "_r", None):
            self.base.close()
            self.fine.close()
            self._lib.kg_refined_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
