"""Host-side mirror of the reference's ``type(lines_method)`` interface.

The reference drives the path through Fortran type-bound procedures of
``type(lines_method)`` (MoL/lines_method_lib.f90:19-44): the constructor
``lines_method_constructor(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, tmax,
CFL, boundary_type_rmax, metric, a, M_bh, m, ...)`` (:49-129), the integrators
``pirk2``/``pirk3``/``rk3``/``rk4`` (:211-474), the ghost fills
``boundary_*`` (:507-597) and the state ``u(nvars, i, j, k)``.  This class
keeps those names, argument meanings and the stop-on-error behaviour (a
``KGError`` is raised where the reference prints and stops), and runs every
call on the GPU through the C ABI of include/kg.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

_SCHEMES = {"pirk2": capi.KG_PIRK2, "pirk3": capi.KG_PIRK3, "rk3": capi.KG_RK3, "rk4": capi.KG_RK4}
_METRICS = {"Minkowski": capi.KG_METRIC_MINKOWSKI, "Kerr_Schild": capi.KG_METRIC_KERR_SCHILD,
            "MKS": capi.KG_METRIC_MKS, "Schw_EF": capi.KG_METRIC_SCHW_EF}
# 'sommerfeld': a radiative outer boundary the reference does not have (the
# outer column is evolved with d_t Pi = -v (d_r Pi + Pi/r)); see include/kg.h
_BCS = {"outflow": capi.KG_BC_OUTFLOW, "inflow": capi.KG_BC_INFLOW, "sommerfeld": capi.KG_BC_SOMMERFELD}
# operator set of the RHS: the reference's live 2nd-order stencils, or the
# 4th-order set its finite_differences.f90 defines (RK3 and RK4 only)
_OPS = {"ref2": capi.KG_OPS_REF2, "fd4": capi.KG_OPS_FD4}


def make_desc(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, CFL, boundary_type_rmax, metric, a, M_bh, m, scheme, dt,
              is_complex, operators):
    """struct kg_desc from the constructor's arguments (lines_method_lib.f90:49-51)."""
    if boundary_type_rmax not in _BCS:
        raise capi.KGError(capi.KG_ERR_UNSUPPORTED, "This boundary condition type for r_max is not implemented yet")
    if metric not in _METRICS:
        raise capi.KGError(capi.KG_ERR_UNSUPPORTED, "That metric is not implmented yet")
    if operators not in _OPS:
        raise capi.KGError(capi.KG_ERR_INVALID, "operators must be 'ref2' or 'fd4'")
    d = capi.KGDesc()
    d.nvars, d.g_pts, d.Nx, d.Ny, d.Nz = nvars, g_pts, Nx, Ny, Nz
    d.scheme = _SCHEMES[scheme]
    d.metric = _METRICS[metric]
    d.bc_rmax = _BCS[boundary_type_rmax]
    d.is_complex = 1 if is_complex else 0
    d.op_set = _OPS[operators]
    d.xmin, d.xmax, d.tmin, d.CFL, d.dt = xmin, xmax, tmin, CFL, dt
    d.a, d.M_bh, d.m = a, M_bh, m
    return d


class LinesMethod:
    """One grid of the Klein-Gordon method-of-lines solver, resident on the GPU.

    Arrays exchanged with the host are the padded state in SoA order
    ``[part][field][k][j][i]`` (part = real[, imaginary]) -- ``u_soa`` -- or
    the reference's own ``u(nvars, i, j, k)`` Fortran layout -- ``u_fortran``.
    """

    def __init__(self, nvars=5, g_pts=2, xmin=1.8, xmax=60.0, Nx=256, Ny=33, Nz=65, tmin=0.0, tmax=None,
                 CFL=0.25, boundary_type_rmax="outflow", metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.0,
                 scheme="pirk2", dt=0.0, is_complex=False, operators="ref2", _handle=None):
        self._lib = capi.load()
        d = make_desc(nvars, g_pts, xmin, xmax, Nx, Ny, Nz, tmin, CFL, boundary_type_rmax, metric, a, M_bh, m,
                      scheme, dt, is_complex, operators)
        if _handle is None:
            h = C.c_void_p()
            capi.check(self._lib.kg_create(C.byref(d), C.byref(h)))
            self._owned = True
        else:  # a level of a refined pair: the pair owns the handle
            h = _handle
            self._owned = False
        self._h = h
        self.desc = d
        dims = (C.c_int64 * 8)()
        capi.check(self._lib.kg_get_dims(h, dims), h)
        self.Nx, self.Ny, self.Nz, self.g_pts, self.ni, self.nj, self.nk, self.npts = [int(v) for v in dims]
        self.nparts = 2 if is_complex else 1
        self.is_complex = bool(is_complex)
        self.tmin = tmin
        self.tmax = tmax
        self.CFL = CFL
        self.metric = metric
        self.boundary_type_rmax = boundary_type_rmax
        self.m = m
        self.a = a
        self.M_bh = M_bh
        self.operators = operators

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._lib.kg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- scalars (this%t, this%dt, this%Nt: lines_method_lib.f90:24,112-116) --
    @property
    def t(self) -> float:
        t = C.c_double()
        capi.check(self._lib.kg_get_time(self._h, C.byref(t), None), self._h)
        return t.value

    @t.setter
    def t(self, value: float) -> None:
        capi.check(self._lib.kg_set_time(self._h, float(value)), self._h)

    @property
    def dt(self) -> float:
        dt = C.c_double()
        capi.check(self._lib.kg_get_time(self._h, None, C.byref(dt)), self._h)
        return dt.value

    @property
    def Nt(self) -> int:
        """int((tmax - tmin)/dt), rounded up to even (lines_method_lib.f90:115-116)."""
        if self.tmax is None:
            raise ValueError("tmax not set")
        nt = int((self.tmax - self.tmin) / self.dt)
        return nt + 1 if nt % 2 else nt

    @property
    def soa_shape(self):
        return (self.nparts, 5, self.nk, self.nj, self.ni)

    @property
    def fortran_shape(self):
        """numpy shape of u(nvars, ni, nj, nk) in Fortran order, complex or real."""
        return (5, self.ni, self.nj, self.nk)

    # -- state transfer ---------------------------------------------------------
    def set_u_soa(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.soa_shape)
        capi.check(self._lib.kg_upload_state(self._h, u.ctypes.data, capi.KG_LAYOUT_SOA), self._h)

    def u_soa(self) -> np.ndarray:
        out = np.empty(self.soa_shape, dtype=np.float64)
        capi.check(self._lib.kg_download_state(self._h, out.ctypes.data, capi.KG_LAYOUT_SOA), self._h)
        return out

    def set_u_fortran(self, u: np.ndarray) -> None:
        """u: numpy array of shape (5, ni, nj, nk) in Fortran order (complex128 or float64)."""
        dt = np.complex128 if self.is_complex else np.float64
        u = np.asfortranarray(u, dtype=dt)
        if u.shape != self.fortran_shape:
            raise ValueError("u must have shape %s" % (self.fortran_shape,))
        capi.check(self._lib.kg_upload_state(self._h, u.ctypes.data, capi.KG_LAYOUT_FORTRAN), self._h)

    def u_fortran(self) -> np.ndarray:
        dt = np.complex128 if self.is_complex else np.float64
        out = np.empty(self.fortran_shape, dtype=dt, order="F")
        capi.check(self._lib.kg_download_state(self._h, out.ctypes.data, capi.KG_LAYOUT_FORTRAN), self._h)
        return out

    def set_u_stage_soa(self, which: int, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.soa_shape)
        capi.check(self._lib.kg_upload_aux(self._h, which, u.ctypes.data, capi.KG_LAYOUT_SOA), self._h)

    def u_stage_soa(self, which: int = 0) -> np.ndarray:
        out = np.empty(self.soa_shape, dtype=np.float64)
        capi.check(self._lib.kg_download_aux(self._h, which, out.ctypes.data, capi.KG_LAYOUT_SOA), self._h)
        return out

    def copy_state_device(self, dev_ptr: int, to_handle: bool) -> None:
        """Exchange the state with a device buffer in SoA order (e.g. a torch CUDA tensor's data_ptr())."""
        capi.check(self._lib.kg_copy_state_device(self._h, C.c_void_p(dev_ptr), 0 if to_handle else 1), self._h)

    # -- checkpoint / restart (lines_method_lib.f90:599-667) ---------------------
    def check_point(self, path: str) -> None:
        """Write the state, time and (PIRK) stage arrays to `path`."""
        capi.check(self._lib.kg_checkpoint_save(self._h, str(path).encode()), self._h)

    def read_from_check_point(self, path: str) -> None:
        """Restore what check_point wrote (same grid, scheme and physics)."""
        capi.check(self._lib.kg_checkpoint_load(self._h, str(path).encode()), self._h)

    # -- the path ---------------------------------------------------------------
    def boundary(self) -> None:
        """xL, xR, yL, yR, zL, zR fills of u (lines_method_lib.f90:313-330)."""
        capi.check(self._lib.kg_fill_ghosts(self._h), self._h)

    def pirk2(self) -> None:
        capi.check(self._lib.kg_pirk2(self._h), self._h)

    def pirk3(self) -> None:
        capi.check(self._lib.kg_pirk3(self._h), self._h)

    def rk3(self) -> None:
        capi.check(self._lib.kg_rk3(self._h), self._h)

    def rk4(self) -> None:
        capi.check(self._lib.kg_rk4(self._h), self._h)

    def step(self, nsteps: int = 1) -> None:
        """nsteps steps of the scheme given at construction (graph replay)."""
        capi.check(self._lib.kg_step(self._h, int(nsteps)), self._h)

    def sync(self) -> None:
        capi.check(self._lib.kg_sync(self._h), self._h)

    # -- launch shape and timing -------------------------------------------------
    def set_tuning(self, kc: int, tj: int) -> None:
        """phi planes per workgroup and theta rows per workgroup of the stage kernel."""
        capi.check(self._lib.kg_set_tuning(self._h, int(kc), int(tj)), self._h)

    def set_batch_persistent(self, wg_per_cu: int) -> None:
        """Batched launches this grid leads (first member): 0 = one workgroup
        per tile, n > 0 = one persistent launch of n workgroups per CU per stage."""
        capi.check(self._lib.kg_set_batch_persistent(self._h, int(wg_per_cu)), self._h)

    def profile(self, on: bool) -> None:
        capi.check(self._lib.kg_profile_enable(self._h, 1 if on else 0), self._h)

    def profile_read(self, kind: int):
        ms = C.c_double()
        n = C.c_int64()
        capi.check(self._lib.kg_profile_read(self._h, kind, C.byref(ms), C.byref(n)), self._h)
        return ms.value, n.value
