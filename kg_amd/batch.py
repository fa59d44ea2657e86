"""Batched evolutions: many independent Klein-Gordon grids advanced together.

Each member is a ``LinesMethod`` (its own grid sizes, field mass, black-hole
mass, initial data and buffers in HBM); ``batch_step`` advances all of them
with one kernel launch per RK4 stage over the tiles of every member
(``kg_batch_step`` in include/kg.h).  A member's result is the same as
stepping it alone.

``uneven_specs`` draws the uneven population of the batched workload: 256
evolutions with Nr uniform in [256, 2048], (Ntheta, Nphi) from {(16,32),
(32,64), (64,128)}, field mass m ~ U[0, 0.5], black-hole mass M ~ U[0.5, 2],
excision radius 0.9*2M, the radial spacing of the Nr=256, r in [1.8, 60]
grid, and a seeded l <= 4 Gaussian-shell pulse.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from .initial_data import shell_state, random_lm_coeffs
from .lines_method import LinesMethod

ANGULAR = [(16, 32), (32, 64), (64, 128)]
DR = 58.2 / 256  # radial spacing of Nr = 256 on r in [1.8, 60]


def batch_step(members, nsteps: int = 1, noalloc: bool = False) -> None:
    """Advance every LinesMethod in `members` by nsteps steps of their common
    scheme (rk4, rk3, pirk2 or pirk3) in batched launches: one per stage
    kernel and one per ghost-fill pass over all members.  noalloc: through
    kg_batch_step_noalloc, which keeps its launch descriptors in the
    library's own fixed device array (no allocation in the call)."""
    if not members:
        return
    lib = capi.load()
    arr = (C.c_void_p * len(members))(*[m._h for m in members])
    fn = lib.kg_batch_step_noalloc if noalloc else lib.kg_batch_step
    st = fn(arr, len(members), int(nsteps))
    capi.check(st, members[0]._h)


@dataclass
class MemberSpec:
    Nr: int
    Ntheta: int
    Nphi: int
    m: float
    M_bh: float
    xmin: float
    xmax: float
    r0: float
    width: float
    coeffs: dict = field(default_factory=dict)
    seed: int = 0

    @property
    def points(self) -> int:
        return self.Nr * self.Ntheta * self.Nphi


def uneven_specs(n: int = 256, seed: int = 20261203):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        Nr = int(rng.integers(256, 2049))
        Nt, Np = ANGULAR[int(rng.integers(0, 3))]
        m = float(rng.uniform(0.0, 0.5))
        M = float(rng.uniform(0.5, 2.0))
        xmin = 0.9 * 2.0 * M
        xmax = xmin + Nr * DR
        r0 = float(rng.uniform(xmin + 10.0, xmin + 0.5 * (xmax - xmin)))
        s = int(rng.integers(0, 2**31))
        out.append(MemberSpec(Nr, Nt, Np, m, M, xmin, xmax, r0, 1.5, random_lm_coeffs(s), s))
    return out


def make_member(spec: MemberSpec, CFL: float = 0.25, on_device: bool = False, operators: str = "ref2",
                scheme: str = "rk4", is_complex: bool = False) -> LinesMethod:
    """Create the grid on the GPU and upload its ghost-filled initial data
    (built on the host, or with torch on the GPU when on_device; a complex
    member gets the shell as its real part and half of it as its imaginary
    part)."""
    lm = LinesMethod(Nx=spec.Nr, Ny=spec.Ntheta + 1, Nz=spec.Nphi + 1, xmin=spec.xmin, xmax=spec.xmax, CFL=CFL,
                     m=spec.m, M_bh=spec.M_bh, scheme=scheme, operators=operators, is_complex=is_complex)
    args = (spec.Nr, spec.Ntheta + 1, spec.Nphi + 1, 2, spec.xmin, spec.xmax, spec.r0, spec.width, 1.0, spec.coeffs)
    if on_device:
        import torch
        from .initial_data import shell_state_device
        u0 = shell_state_device(*args)
        if is_complex:
            u0 = torch.stack([u0, 0.5 * u0])
        torch.cuda.synchronize()
        lm.copy_state_device(u0.data_ptr(), to_handle=True)
        del u0
    else:
        u0 = shell_state(*args)
        lm.set_u_soa(np.stack([u0, 0.5 * u0]) if is_complex else u0[None])
    lm.boundary()
    return lm
