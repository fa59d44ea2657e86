"""Initial data.

- ``initial_conditions``: the reference's ``subroutine initial_conditions``
  (main/initial_conditions.f90:1-220) for 'spherical_shell', 'random' and
  'random_shell', built on the GPU (``kg_initial_data``), and
  ``random_field``, the Cartesian random field of tools/random.f90:17-59
  (``kg_ic_random_field``).
- ``shell_state`` and its relatives: synthetic workloads, Gaussian shells
  times real Y_lm evaluated at every padded point (r = x_i, theta = y_j,
  phi = z_k of MoL/mesh_refinement_lib.f90:130-132) with exact partial
  derivatives and Pi = 0; host-side input generation for tests and benches.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
from scipy.special import lpmv

from . import _capi as capi

IC_TYPES = {"spherical_shell": capi.KG_IC_SPHERICAL_SHELL, "random": capi.KG_IC_RANDOM,
            "random_shell": capi.KG_IC_RANDOM_SHELL}


def _uniform_ptr(uniforms, n):
    if uniforms is None:
        return None, None
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    if u.size != n ** 3:
        raise ValueError("uniforms must hold ncart**3 values")
    return u, u.ctypes.data_as(C.POINTER(C.c_double))


def initial_conditions(lm, initial_conditions_type, mu=0.0, sigma=1.0, p0=1.0, ncart=256, seed=0, uniforms=None):
    """Fill lm's state with the reference's initial data; returns Amp.

    mu, sigma: the shell (global_numbers); p0: momentum width of the random
    field (the reference passes 1); ncart: Cartesian FFT points per axis (the
    reference fixes 256); uniforms: optional random phases U [k][j][i] in
    [0, 1), otherwise drawn from `seed`.  The random types need a complex lm.
    """
    if initial_conditions_type not in IC_TYPES:
        raise capi.KGError(capi.KG_ERR_UNSUPPORTED, "Intial condition is not implemented yet")
    d = capi.KGICDesc()
    d.type = IC_TYPES[initial_conditions_type]
    d.ncart = ncart
    d.mu, d.sigma, d.p0 = mu, sigma, p0
    d.seed = seed & 0xFFFFFFFFFFFFFFFF
    keep, ptr = _uniform_ptr(uniforms, ncart)
    d.uniforms = ptr
    amp = C.c_double()
    lib = capi.load()
    st = lib.kg_initial_data(lm._h, C.byref(d), C.byref(amp))
    if st != capi.KG_OK:
        raise capi.KGError(st, lib.kg_ic_last_error().decode())
    del keep
    return amp.value


def random_field(ncart, r_max, p0=1.0, seed=0, uniforms=None, device=0):
    """psi_x = ifft_3D(psi_p) of random.f90:17-59 on the GPU: complex [k][j][i]."""
    keep, ptr = _uniform_ptr(uniforms, ncart)
    out = np.empty((ncart, ncart, ncart), dtype=np.complex128)
    lib = capi.load()
    st = lib.kg_ic_random_field(device, ncart, r_max, p0, seed & 0xFFFFFFFFFFFFFFFF, ptr,
                                out.ctypes.data_as(C.POINTER(C.c_double)))
    if st != capi.KG_OK:
        raise capi.KGError(st, lib.kg_ic_last_error().decode())
    del keep
    return out


def grid_coords(Nx, Ny, Nz, g, xmin, xmax):
    """x_i, y_j, z_k of the padded grid (mesh_refinement_lib.f90:109-132, nproc = 1)."""
    dx = (xmax - xmin) / Nx
    dy = math.acos(-1.0) / (Ny - 1)
    dz = 2.0 * math.acos(-1.0) / (Nz - 1)
    ii = np.arange(Nx + 2 * g + 1)
    jj = np.arange(Ny + 2 * g + 1)
    x = xmin + ii.astype(np.float64) * dx
    y = -dy / 2.0 + (jj - g).astype(np.float64) * dy
    z = np.arange(Nz + 1, dtype=np.float64) * dz
    return x, y, z


def real_ylm(l, m, theta, phi):
    """Real spherical harmonic and its theta, phi derivatives on a (theta, phi) mesh.

    Y_l0 = N P_l(cos t); Y_lm = sqrt(2) N P_l^m(cos t) cos(m p) (m > 0);
    Y_l,-m = sqrt(2) N P_l^m(cos t) sin(m p), N = sqrt((2l+1)/4pi (l-m)!/(l+m)!).
    Uses scipy's lpmv (with the Condon-Shortley phase).
    """
    am = abs(m)
    N = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    ct = np.cos(theta)
    st = np.sin(theta)
    P = lpmv(am, l, ct)
    # dP/dtheta = -sin(t) dP/dx via the recurrence (1-x^2) dP_l^m/dx = (l+m) P_{l-1}^m - l x P_l^m
    Plm1 = lpmv(am, l - 1, ct) if l >= 1 else np.zeros_like(ct)
    with np.errstate(divide="ignore", invalid="ignore"):
        dPdt = np.where(np.abs(st) > 1e-300, -((l + am) * Plm1 - l * ct * P) / st, 0.0)
    if m == 0:
        Y = N * P[:, None] * np.ones_like(phi)[None, :]
        Yt = N * dPdt[:, None] * np.ones_like(phi)[None, :]
        Yp = np.zeros_like(Y)
    elif m > 0:
        c = math.sqrt(2.0) * N
        Y = c * P[:, None] * np.cos(am * phi)[None, :]
        Yt = c * dPdt[:, None] * np.cos(am * phi)[None, :]
        Yp = -c * am * P[:, None] * np.sin(am * phi)[None, :]
    else:
        c = math.sqrt(2.0) * N
        Y = c * P[:, None] * np.sin(am * phi)[None, :]
        Yt = c * dPdt[:, None] * np.sin(am * phi)[None, :]
        Yp = c * am * P[:, None] * np.cos(am * phi)[None, :]
    return Y, Yt, Yp


def shell_state(Nx, Ny, Nz, g, xmin, xmax, r0, width, amp, coeffs):
    """Padded SoA state [5][nk][nj][ni] for phi = amp exp(-(r-r0)^2/w^2) sum c_lm Y_lm.

    coeffs: dict {(l, m): c}.
    """
    x, y, z = grid_coords(Nx, Ny, Nz, g, xmin, xmax)
    R = amp * np.exp(-((x - r0) ** 2) / width ** 2)
    dR = R * (-2.0 * (x - r0) / width ** 2)
    A = np.zeros((len(y), len(z)))
    At = np.zeros_like(A)
    Ap = np.zeros_like(A)
    for (l, m), c in coeffs.items():
        Y, Yt, Yp = real_ylm(l, m, y, z)
        A += c * Y
        At += c * Yt
        Ap += c * Yp
    u = np.empty((5, len(z), len(y), len(x)))
    u[0] = R[None, None, :] * A.T[:, :, None]
    u[1] = dR[None, None, :] * A.T[:, :, None]
    u[2] = R[None, None, :] * At.T[:, :, None]
    u[3] = R[None, None, :] * Ap.T[:, :, None]
    u[4] = 0.0
    return u


def small_workload_state(Nx=256, Ny=33, Nz=65, g=2, xmin=1.8, xmax=60.0):
    """phi = 1e-3 exp(-(r-15)^2/1.5^2) Y_2^0, Pi = 0 (the small CPU-runnable workload)."""
    return shell_state(Nx, Ny, Nz, g, xmin, xmax, 15.0, 1.5, 1e-3, {(2, 0): 1.0})


def random_lm_coeffs(seed, lmax=4, std=1e-3):
    rng = np.random.default_rng(seed)
    return {(l, m): float(rng.normal(0.0, std)) for l in range(lmax + 1) for m in range(-l, l + 1)}


def full_workload_state(seed=20261203, Nx=2048, Ny=129, Nz=257, g=2, xmin=1.8, xmax=200.0):
    """phi = exp(-(r-30)^2/3^2) sum_{l<=4} c_lm Y_lm, c_lm ~ N(0, 1e-3) (the full single-GPU workload)."""
    return shell_state(Nx, Ny, Nz, g, xmin, xmax, 30.0, 3.0, 1.0, random_lm_coeffs(seed))


def shell_state_device(Nx, Ny, Nz, g, xmin, xmax, r0, width, amp, coeffs, device="cuda"):
    """shell_state built on the GPU with torch (same values, SoA [5][nk][nj][ni]).

    The radial profile and the angular sum are formed on the host in FP64 (small
    1-D/2-D arrays) and their outer products on the device, so large inputs do
    not pass through host memory.
    """
    import torch

    x, y, z = grid_coords(Nx, Ny, Nz, g, xmin, xmax)
    R = amp * np.exp(-((x - r0) ** 2) / width ** 2)
    dR = R * (-2.0 * (x - r0) / width ** 2)
    A = np.zeros((len(y), len(z)))
    At = np.zeros_like(A)
    Ap = np.zeros_like(A)
    for (l, m), c in coeffs.items():
        Y, Yt, Yp = real_ylm(l, m, y, z)
        A += c * Y
        At += c * Yt
        Ap += c * Yp
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=torch.float64)
    Rt, dRt = t(R), t(dR)
    At_, Att, Apt = t(A.T), t(At.T), t(Ap.T)  # [nk][nj]
    u = torch.empty((5, len(z), len(y), len(x)), dtype=torch.float64, device=device)
    u[0] = At_[:, :, None] * Rt[None, None, :]
    u[1] = At_[:, :, None] * dRt[None, None, :]
    u[2] = Att[:, :, None] * Rt[None, None, :]
    u[3] = Apt[:, :, None] * Rt[None, None, :]
    u[4] = 0.0
    return u
