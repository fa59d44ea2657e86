"""Two-level refinement on the oracle (test infrastructure only).

CPU restatement of kg_amd/csrc/kg_refine.hip: a base-level oracle.Sim and a
fine-level oracle.Sim on [xmin, r_patch + buffer] with dr/2 and dt/2,
advanced Berger-Oliger style -- base step; buffer refill (r: copies at
coincident points, (3, -25, 150, 150, -25, 3)/256 at midpoints; t: the last
three base times, -1/8, 3/4, 3/8 at the half step, 1/2, 1/2 while two are
known) and fine step, twice; injection of the fine values into base columns
0 .. I-3.  Every expression keeps the kernels' association, so the GPU
reproduces these arrays bit for bit.  The reference has no refinement levels
(mesh_refinement_lib.f90:16-147 is a slab decomposition), so there is no
reference statement to follow; the algorithm is pinned by exactness and
convergence checks in tests/test_oracle_refinement.py.
"""
from __future__ import annotations

import math

import numpy as np

from . import Sim


class RefinedSim:
    def __init__(self, Nx, Ny, Nz, g, xmin, xmax, metric, a, M_bh, m, bc, CFL, r_patch, buffer=16, dt=0.0,
                 order=2):
        self.base = Sim(Nx, Ny, Nz, g, xmin, xmax, metric, a, M_bh, m, bc, CFL, dt)
        self.base.set_order(order)
        dx0 = (xmax - xmin) / 1.0
        dx0 = ((xmin + 1.0 * dx0) - (xmin + 0.0 * dx0)) / Nx  # the grid's own dx (mesh_refinement_lib.f90:109-115)
        I = int(math.floor((r_patch - xmin) / dx0 + 0.5))  # lround of a positive value
        B = buffer
        self.I, self.B = I, B
        Nxf = 2 * I + B - 2 * g
        xmaxf = xmin + float(Nxf) * (dx0 / 2.0)
        self.fine = Sim(Nxf, Ny, Nz, g, xmin, xmaxf, metric, a, M_bh, m, "inflow", CFL, self.base.dt / 2.0)
        self.fine.set_order(order)
        self.c_lo, self.ncol = I - 2, B // 2 + 6
        self.hist = {}  # base time index -> base buffer columns
        self.n = 0      # base steps taken

    def _save(self, key):
        self.hist[key] = self.base.u[..., self.c_lo:self.c_lo + self.ncol].copy()

    def _prolong(self, slots, weights):
        """Refill the fine buffer columns from the base history (time
        combination first, then the spatial rule)."""
        if len(slots) == 1:
            V = slots[0]
        elif len(slots) == 2:
            V = weights[0] * slots[0] + weights[1] * slots[1]
        else:
            V = weights[0] * slots[0] + weights[1] * slots[1] + weights[2] * slots[2]
        u = self.fine.u
        ni = self.fine.ni
        for fi in range(ni - self.B, ni):
            if fi % 2 == 0:
                u[..., fi] = V[..., fi // 2 - self.c_lo]
            else:
                c = (fi - 1) // 2 - self.c_lo
                u[..., fi] = (3.0 * V[..., c - 2] - 25.0 * V[..., c - 1] + 150.0 * V[..., c] + 150.0 * V[..., c + 1]
                              - 25.0 * V[..., c + 2] + 3.0 * V[..., c + 3]) * 0.00390625

    def step(self, scheme, nsteps=1):
        for _ in range(nsteps):
            t = self.n
            if t not in self.hist:
                self._save(t)
            self.base.step(scheme, 1)
            self._save(t + 1)
            self._prolong([self.hist[t]], [1.0])
            self.fine.step(scheme, 1)
            if t - 1 in self.hist:
                self._prolong([self.hist[t - 1], self.hist[t], self.hist[t + 1]], [-0.125, 0.75, 0.375])
            else:
                self._prolong([self.hist[t], self.hist[t + 1]], [0.5, 0.5])
            self.fine.step(scheme, 1)
            nc = self.I - 2
            self.base.u[..., 0:nc] = self.fine.u[..., 0:2 * nc:2]
            self.hist.pop(t - 1, None)
            self.n += 1
