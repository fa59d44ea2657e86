"""CPU FP64 oracle for the Klein-Gordon method-of-lines path (test infrastructure only).

Loads ``oracle/libkg_oracle.so`` (the C restatement in ``kg_oracle.c``) and, when
it has been built, ``oracle/_ref/libkgref_fd.so`` (the reference's own
``tools/finite_differences.f90`` compiled from source with bind(C) wrappers).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, and only as the checker.  The product (``kg_amd``)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libkg_oracle.so")
_REF = os.path.join(_HERE, "_ref", "libkgref_fd.so")

PIRK2, PIRK3, RK3, RK4 = 0, 1, 2, 3
SCHEMES = {"pirk2": PIRK2, "pirk3": PIRK3, "rk3": RK3, "rk4": RK4}
MINKOWSKI, KERR_SCHILD, MKS, SCHW_EF = 0, 1, 2, 3
METRICS = {"Minkowski": MINKOWSKI, "Kerr_Schild": KERR_SCHILD, "MKS": MKS, "Schw_EF": SCHW_EF}
OUTFLOW, INFLOW, SOMMERFELD = 0, 1, 2
BCS = {"outflow": OUTFLOW, "inflow": INFLOW, "sommerfeld": SOMMERFELD}
# deliberate mistakes the oracle can plant (kg_oracle.h KGO_PLANT_*), for the
# tests' mutation checks
PLANTS = {"drop_pi_dxbeta": 1, "swap_phi_flux": 2, "psi2_alpha_pi": 3, "drop_psi1_dxbeta": 4,
          "gdn13_no_sin2": 5, "advect_pi": 6, "mass_sign": 7}

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and, when /root/reference exists, the reference stencils)."""
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference")
    target = "ref" if with_ref else "all"
    subprocess.run(["make", "-s", "-C", _HERE, target], check=True)


def _load():
    if not os.path.exists(_LIB):
        raise RuntimeError("oracle/libkg_oracle.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(_LIB)
    lib.kgo_create.restype = C.c_void_p
    lib.kgo_create.argtypes = [C.c_int] * 4 + [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                               C.c_double, C.c_int, C.c_double, C.c_double]
    lib.kgo_destroy.argtypes = [C.c_void_p]
    lib.kgo_field.restype = _dp
    lib.kgo_field.argtypes = [C.c_void_p, C.c_int]
    lib.kgo_table.restype = _dp
    lib.kgo_table.argtypes = [C.c_void_p, C.c_int]
    lib.kgo_info.argtypes = [C.c_void_p, _i64p, _dp]
    lib.kgo_reference_dt.restype = C.c_double
    lib.kgo_reference_dt.argtypes = [C.c_int] * 4 + [C.c_double] * 3
    for nm in ("kgo_dx1_raw", "kgo_dx2_raw", "kgo_dx3_raw", "kgo_dx1_4_raw", "kgo_dx2_4_raw", "kgo_dx3_4_raw"):
        getattr(lib, nm).argtypes = [C.c_int] * 3 + [C.c_double, _dp, _dp]
    lib.kgo_advec_x_raw.argtypes = [C.c_int] * 3 + [C.c_double, _dp, _dp, _dp]
    lib.kgo_advec_x_4_raw.argtypes = [C.c_int] * 3 + [C.c_double, _dp, _dp, _dp]
    lib.kgo_set_order.argtypes = [C.c_void_p, C.c_int]
    lib.kgo_set_order.restype = C.c_int
    lib.kgo_plant.argtypes = [C.c_void_p, C.c_int]
    lib.kgo_plant.restype = C.c_int
    lib.kgo_threads.restype = C.c_int
    lib.kgo_threads.argtypes = []
    for nm in ("kgo_rhs_phi", "kgo_rhs_psi1", "kgo_rhs_psi2", "kgo_rhs_psi3", "kgo_rhs_pi", "kgo_rhs_all"):
        getattr(lib, nm).argtypes = [C.c_void_p, _dp, _dp]
    lib.kgo_fill.argtypes = [C.c_void_p, _dp, _dp]
    lib.kgo_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.kgo_fill_source_map.argtypes = [C.c_void_p, _i64p]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def threads() -> int:
    """Host threads the oracle's point loops use (OMP_NUM_THREADS)."""
    return lib().kgo_threads()


def reference_dt(Nx, Ny, Nz, g, xmin, xmax, CFL):
    return lib().kgo_reference_dt(Nx, Ny, Nz, g, xmin, xmax, CFL)


class Sim:
    """One simulation of the oracle: padded real FP64 arrays [5][nk][nj][ni]."""

    def __init__(self, Nx, Ny, Nz, g=2, xmin=1.8, xmax=60.0, metric="Kerr_Schild", a=0.0, M_bh=1.0,
                 m=0.0, bc="outflow", CFL=0.25, dt=0.0):
        L = lib()
        self._h = L.kgo_create(Nx, Ny, Nz, g, xmin, xmax, METRICS[metric], a, M_bh, m, BCS[bc], CFL, dt)
        if not self._h:
            raise ValueError("invalid oracle configuration")
        o = (C.c_int64 * 8)()
        d = (C.c_double * 8)()
        L.kgo_info(self._h, o, d)
        self.Nx, self.Ny, self.Nz, self.g, self.ni, self.nj, self.nk, self.npts = [int(v) for v in o]
        self.dx, self.dy, self.dz, self.dt = d[0], d[1], d[2], d[3]
        self.shape = (5, self.nk, self.nj, self.ni)

    def close(self):
        if getattr(self, "_h", None):
            lib().kgo_destroy(self._h)
            self._h = None

    __del__ = close

    def _view(self, which):
        p = lib().kgo_field(self._h, which)
        return np.ctypeslib.as_array(p, shape=self.shape)

    @property
    def u(self):
        return self._view(0)

    @property
    def u_stage(self):
        return self._view(2)

    @property
    def u_stage_2(self):
        return self._view(3)

    def table(self, which):
        names = {"alpha": 0, "beta_r": 1, "dx_beta": 2, "sqrtg": 3, "g_rr": 4, "g_rp": 5, "g_tt": 6,
                 "g_pp": 7}
        if which in names:
            p = lib().kgo_table(self._h, names[which])
            return np.ctypeslib.as_array(p, shape=(self.nj, self.ni)).copy()
        if which in ("g_dn", "g_up"):
            p = lib().kgo_table(self._h, 8 if which == "g_dn" else 9)
            return np.ctypeslib.as_array(p, shape=(4, 4, self.nj, self.ni)).copy()
        n = {"x": self.ni, "y": self.nj, "z": self.nk}[which]
        p = lib().kgo_table(self._h, {"x": 10, "y": 11, "z": 12}[which])
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def set_order(self, order):
        """Operator set of the RHS: 2 (the reference's live stencils) or 4 (its fd4 set)."""
        if lib().kgo_set_order(self._h, order) != 0:
            raise ValueError("order must be 2 or 4")

    def plant(self, code):
        """Plant deliberate mistake `code` (a PLANTS value; 0 removes it) for a mutation check."""
        if lib().kgo_plant(self._h, int(code)) != 0:
            raise ValueError("unknown planted mistake %r" % (code,))

    def step(self, scheme, nsteps=1):
        lib().kgo_step(self._h, SCHEMES[scheme] if isinstance(scheme, str) else scheme, nsteps)

    def rhs_all(self, vars5):
        out = np.zeros(self.shape)
        v = np.ascontiguousarray(vars5, dtype=np.float64)
        lib().kgo_rhs_all(self._h, _ptr(out), _ptr(v))
        return out

    def fill(self, u, u_p=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        up = np.ascontiguousarray(u if u_p is None else u_p, dtype=np.float64).copy()
        lib().kgo_fill(self._h, _ptr(u), _ptr(up))
        return u

    def fill_source_map(self):
        out = np.zeros(self.npts, dtype=np.int64)
        lib().kgo_fill_source_map(self._h, out.ctypes.data_as(_i64p))
        return out.reshape(self.nk, self.nj, self.ni)


def stencil(name, f, h, beta=None):
    """Oracle stencil on a real padded array f[nk][nj][ni]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    nk, nj, ni = f.shape
    out = np.zeros_like(f)
    L = lib()
    if name in ("advec_x", "advec_x_4"):
        b = np.ascontiguousarray(beta, dtype=np.float64)
        getattr(L, "kgo_%s_raw" % name)(ni, nj, nk, h, _ptr(f), _ptr(b), _ptr(out))
    else:
        getattr(L, "kgo_%s_raw" % name)(ni, nj, nk, h, _ptr(f), _ptr(out))
    return out


def have_reference() -> bool:
    return os.path.exists(_REF)


_ref = None


def ref_stencil(name, f, h, g, beta=None):
    """The reference's own stencil (complex(8) in and out; real(8) for the fd4 set, names *_4)."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(_REF)
        for nm in ("kgref_dx1", "kgref_dx2", "kgref_dx3"):
            getattr(_ref, nm).argtypes = [C.c_int] * 4 + [C.c_double, C.c_void_p, C.c_void_p]
        _ref.kgref_advec_x.argtypes = [C.c_int] * 4 + [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        for nm in ("kgref_dx1_4", "kgref_dx2_4", "kgref_dx3_4"):
            getattr(_ref, nm).argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p, C.c_void_p]
        _ref.kgref_advec_x_4.argtypes = [C.c_int] * 3 + [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    if name.endswith("_4"):  # the fd4 set: real arrays
        f = np.ascontiguousarray(f, dtype=np.float64)
        nk, nj, ni = f.shape
        out = np.zeros_like(f)
        if name == "advec_x_4":
            b = np.ascontiguousarray(beta, dtype=np.float64)
            _ref.kgref_advec_x_4(ni, nj, nk, h, f.ctypes.data, b.ctypes.data, out.ctypes.data)
        else:
            getattr(_ref, "kgref_" + name)(ni, nj, nk, h, f.ctypes.data, out.ctypes.data)
        return out
    f = np.ascontiguousarray(f, dtype=np.complex128)
    nk, nj, ni = f.shape
    out = np.zeros_like(f)
    if name == "advec_x":
        b = np.ascontiguousarray(beta, dtype=np.float64)
        _ref.kgref_advec_x(ni, nj, nk, g, h, f.ctypes.data, b.ctypes.data, out.ctypes.data)
    else:
        getattr(_ref, "kgref_" + name)(ni, nj, nk, g, h, f.ctypes.data, out.ctypes.data)
    return out


# --------------------------------------------------------------------------
# The integrators' stage arithmetic with a caller-supplied right-hand side
# (kgo_*_generic), for pinning the combine formulas on linear systems.
# --------------------------------------------------------------------------
_RHS_CB = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)
_FILL_CB = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)


class _Ops(C.Structure):
    _fields_ = [("rhs14", _RHS_CB), ("rhs5", _RHS_CB), ("fill", C.c_void_p), ("ctx", C.c_void_p),
                ("n", C.c_int64)]


class _Bufs(C.Structure):
    _fields_ = [(nm, _dp) for nm in ("u", "u_p", "u_stage", "u_stage_2", "rhs_u", "k1", "k2", "k3", "k4")]


class GenericIntegrator:
    """Runs one of the oracle's integrators on n values per field with Python
    right-hand sides: rhs14(vars[5,n]) -> out[4,n], rhs5(vars[5,n]) -> out[n]."""

    def __init__(self, n, rhs14, rhs5):
        self.n = n
        names = [f[0] for f in _Bufs._fields_]
        self.bufs = {nm: np.zeros(5 * n) for nm in names}

        def c14(ctx, vars_p, out_p):
            v = np.ctypeslib.as_array(vars_p, shape=(5, n))
            o = np.ctypeslib.as_array(out_p, shape=(5, n))
            o[:4] = rhs14(v.copy())

        def c5(ctx, vars_p, out_p):
            v = np.ctypeslib.as_array(vars_p, shape=(5, n))
            o = np.ctypeslib.as_array(out_p, shape=(n,))
            o[:] = rhs5(v.copy())

        self._c14 = _RHS_CB(c14)
        self._c5 = _RHS_CB(c5)
        self.ops = _Ops(self._c14, self._c5, None, None, n)
        self.b = _Bufs(*[self.bufs[nm].ctypes.data_as(_dp) for nm in names])
        L = lib()
        for nm in ("kgo_pirk2_generic", "kgo_pirk3_generic", "kgo_rk3_generic", "kgo_rk4_generic"):
            getattr(L, nm).argtypes = [C.POINTER(_Ops), C.POINTER(_Bufs), C.c_double]

    @property
    def u(self):
        return self.bufs["u"].reshape(5, self.n)

    def stage_array(self, which=0):
        return self.bufs["u_stage" if which == 0 else "u_stage_2"].reshape(5, self.n)

    def step(self, scheme, dt, nsteps=1):
        fn = getattr(lib(), "kgo_%s_generic" % scheme)
        for _ in range(nsteps):
            fn(C.byref(self.ops), C.byref(self.b), dt)


# --------------------------------------------------------------------------
# Spectral projection (kg_spectral_oracle.c)
# --------------------------------------------------------------------------
def _spec():
    L = lib()
    if not getattr(L, "_spec_ready", False):
        L.kgs_plm.restype = C.c_double
        L.kgs_plm.argtypes = [C.c_int, C.c_int, C.c_double]
        L.kgs_sph_jl.restype = C.c_double
        L.kgs_sph_jl.argtypes = [C.c_int, C.c_double]
        L.kgs_jl_zero.restype = C.c_double
        L.kgs_jl_zero.argtypes = [C.c_int, C.c_int]
        L.kgs_sph_harmonics.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.kgs_project_sh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                     _dp, _dp]
        L.kgs_bessel_basis.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_int, C.c_double, _dp, _dp]
        L.kgs_radial_project.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.kgs_domain_to_sph.argtypes = [C.c_int] * 4 + [C.c_double] * 4 + [_dp, _dp]
        L._spec_ready = True
    return L


def plm(l, m, x):
    return _spec().kgs_plm(l, m, x)


def sph_jl(l, x):
    return _spec().kgs_sph_jl(l, x)


def jl_zero(l, n):
    return _spec().kgs_jl_zero(l, n)


def sph_harmonics(lmax, theta, phi):
    """Y_lm on the spectral grid: complex [lm][len(theta)][len(phi)], lm = l*l + m + l."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = (lmax + 1) ** 2
    yr = np.zeros((n, len(theta), len(phi)))
    yi = np.zeros_like(yr)
    _spec().kgs_sph_harmonics(lmax, len(theta), len(phi), _ptr(theta), _ptr(phi), _ptr(yr), _ptr(yi))
    return yr + 1j * yi


def project_sh(f, ylm, lmax, theta, dth, dph):
    """project_sh of one complex slice f[theta][phi] -> flm[lm] complex."""
    f = np.asarray(f, dtype=np.complex128)
    fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
    yr, yi = np.ascontiguousarray(ylm.real), np.ascontiguousarray(ylm.imag)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    n = (lmax + 1) ** 2
    orr, oi = np.zeros(n), np.zeros(n)
    _spec().kgs_project_sh(f.shape[0], f.shape[1], dth, dph, _ptr(theta), _ptr(fr), _ptr(fi), lmax, _ptr(yr),
                           _ptr(yi), _ptr(orr), _ptr(oi))
    return orr + 1j * oi


def bessel_basis(lmax, nk, r, R_max, kind=0, kmax=0.0):
    """(k_values [l][n], j_basis [l][n][i])."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    kv = np.zeros((lmax + 1, nk))
    jb = np.zeros((lmax + 1, nk, len(r)))
    _spec().kgs_bessel_basis(lmax, nk, len(r), _ptr(r), R_max, kind, kmax, _ptr(kv), _ptr(jb))
    return kv, jb


def radial_project(lmax, nk, r, dr, jb, flm):
    """flm [i][lm] complex -> ft [n][lm] complex."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    fr, fi = np.ascontiguousarray(flm.real), np.ascontiguousarray(flm.imag)
    jb = np.ascontiguousarray(jb, dtype=np.float64)
    n = (lmax + 1) ** 2
    orr, oi = np.zeros((nk, n)), np.zeros((nk, n))
    _spec().kgs_radial_project(lmax, nk, len(r), _ptr(r), dr, _ptr(jb), _ptr(fr), _ptr(fi), _ptr(orr), _ptr(oi))
    return orr + 1j * oi


def domain_to_sph(fd, Nx, Ny, Nz, g, xmin, dx, dy, dz):
    """Padded real field fd[nk][nj][ni] -> spectral grid f[Nz][Ny][Nx]."""
    fd = np.ascontiguousarray(fd, dtype=np.float64)
    out = np.zeros((Nz, Ny, Nx))
    _spec().kgs_domain_to_sph(Nx, Ny, Nz, g, xmin, dx, dy, dz, _ptr(fd), _ptr(out))
    return out


# --------------------------------------------------------------------------
# Diagnostics (kg_diag_oracle.c)
# --------------------------------------------------------------------------
def _diag():
    L = lib()
    if not getattr(L, "_diag_ready", False):
        L.kgd_density.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.kgd_currents.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.kgd_integrate.restype = C.c_double
        L.kgd_integrate.argtypes = [C.c_void_p, _dp]
        L.kgd_trapezium_2d_complex.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L._diag_ready = True
    return L


def _opt(a):
    return _ptr(a) if a is not None else None


def density(sim, u_re, u_im=None):
    """density() (diagnostic_lib.f90:258-271) of a padded state [5][nk][nj][ni]."""
    u_re = np.ascontiguousarray(u_re, dtype=np.float64)
    u_im = None if u_im is None else np.ascontiguousarray(u_im, dtype=np.float64)
    rho = np.zeros(u_re.shape[1:])
    _diag().kgd_density(sim._h, _ptr(u_re), _opt(u_im), _ptr(rho))
    return rho


def currents(sim, u_re, u_im=None):
    """currents_ADM (diagnostic_lib.f90:214-224): complex j[3][nk][nj][ni]."""
    u_re = np.ascontiguousarray(u_re, dtype=np.float64)
    u_im = None if u_im is None else np.ascontiguousarray(u_im, dtype=np.float64)
    jr = np.zeros((3,) + u_re.shape[1:])
    ji = np.zeros_like(jr)
    _diag().kgd_currents(sim._h, _ptr(u_re), _opt(u_im), _ptr(jr), _ptr(ji))
    return jr + 1j * ji


def integrate(sim, f):
    """integrate_real (integral.f90:646-669) of a padded real field [nk][nj][ni]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    return _diag().kgd_integrate(sim._h, _ptr(f))


def trapezium_2d_complex(f, dx, dy):
    """trapezium_2D_complex (integral.f90:330-362) of f[i][j]."""
    f = np.asarray(f, dtype=np.complex128)
    fr, fi = np.ascontiguousarray(f.real), np.ascontiguousarray(f.imag)
    o_r, o_i = C.c_double(), C.c_double()
    _diag().kgd_trapezium_2d_complex(f.shape[0], f.shape[1], dx, dy, _ptr(fr), _ptr(fi), C.byref(o_r), C.byref(o_i))
    return complex(o_r.value, o_i.value)


# --------------------------------------------------------------------------
# Initial data (kg_ic_oracle.c)
# --------------------------------------------------------------------------
IC_TYPES = {"spherical_shell": 0, "random": 1, "random_shell": 2}


def _ic():
    L = lib()
    if not getattr(L, "_ic_ready", False):
        L.kgi_uniforms.argtypes = [C.c_uint64, C.c_int64, _dp]
        L.kgi_random_field.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.kgi_sample.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.kgi_initial.restype = C.c_double
        L.kgi_initial.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, _dp, _dp,
                                  _dp, _dp]
        L._ic_ready = True
    return L


def uniforms(seed, n):
    """The n**3 phase numbers U [k][j][i] drawn from `seed`."""
    out = np.zeros((n, n, n))
    _ic().kgi_uniforms(seed, n ** 3, _ptr(out))
    return out


def random_field(n, r_max, p0, U):
    """psi_x = ifft_3D(psi_p) (random.f90:17-59): complex [k][j][i]."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    assert U.shape == (n, n, n)
    re, im = np.zeros((n, n, n)), np.zeros((n, n, n))
    _ic().kgi_random_field(n, r_max, p0, _ptr(U), _ptr(re), _ptr(im))
    return re + 1j * im


def sample(sim, field, r_max):
    """convert_to_spherical (random.f90:128-156) of a complex [k][j][i] field: complex [nk][nj][ni]."""
    n = field.shape[0]
    fr, fi = np.ascontiguousarray(field.real), np.ascontiguousarray(field.imag)
    pr, pi = np.zeros((sim.nk, sim.nj, sim.ni)), np.zeros((sim.nk, sim.nj, sim.ni))
    _ic().kgi_sample(sim._h, n, r_max, _ptr(fr), _ptr(fi), _ptr(pr), _ptr(pi))
    return pr + 1j * pi


def initial_conditions(sim, kind, mu=0.0, sigma=1.0, field=None, r_max=0.0):
    """initial_conditions (initial_conditions.f90:23-188): (state complex [5][nk][nj][ni], Amp).

    For the random kinds `field` is psi_x of random_field and r_max the value
    it was built with."""
    t = IC_TYPES[kind]
    ure = np.zeros(sim.shape)
    uim = np.zeros(sim.shape)
    if t == 0:
        amp = _ic().kgi_initial(sim._h, 0, mu, sigma, 0, 0.0, None, None, _ptr(ure), _ptr(uim))
    else:
        n = field.shape[0]
        fr, fi = np.ascontiguousarray(field.real), np.ascontiguousarray(field.imag)
        amp = _ic().kgi_initial(sim._h, t, mu, sigma, n, r_max, _ptr(fr), _ptr(fi), _ptr(ure), _ptr(uim))
    return ure + 1j * uim, amp
