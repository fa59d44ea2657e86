"""The Sommerfeld outer boundary of the oracle, pinned by an exact solution.

The reference has 'outflow' and 'inflow' outer boundaries only
(lines_method_lib.f90:522-535); the Sommerfeld boundary is new physics here:
the outer column evolves with R_Pi = -(v (D_r Pi + Pi/r)), v = alpha
sqrt(gamma^rr) - beta^r, and keeps the other fields' own right-hand sides.

In flat space (Kerr-Schild with M = 0: alpha = 1, beta = 0, gamma^rr = 1) the
outgoing spherical wave phi = f(t - r)/r has Pi = f'/r, which satisfies
d_t Pi = -(d_r Pi + Pi/r) exactly, so the condition adds no error of its own:
after the pulse has left the grid, what remains (reflection plus truncation)
falls at the operator set's order, 4x per halving of dr with the 2nd-order set
and 16x with fd4.  The reference's outflow copy reflects the pulse at a fixed
amplitude that does not converge.
"""
import numpy as np
import pytest

import oracle


def _outgoing_pulse(Nr, bc, order, T=16.0, xmin=5.0, xmax=25.0, r0=15.0, w=2.0):
    sim = oracle.Sim(Nr, 5, 9, 2, xmin, xmax, "Kerr_Schild", 0.0, 0.0, 0.0, bc, 0.25)
    sim.set_order(order)
    x = sim.table("x")

    def exact(t):
        u = t - x + r0
        f = np.exp(-(u ** 2) / w ** 2)
        fp = -2.0 * u / w ** 2 * f
        out = np.zeros(sim.shape)
        out[0] = f / x
        out[1] = -fp / x - f / x ** 2
        out[4] = fp / x
        return out

    sim.u[...] = exact(0.0)
    n = int(round(T / sim.dt))
    sim.step("rk4", n)
    want = exact(n * sim.dt)
    sel = x >= xmin + 2.0  # away from the inner extrapolation
    err = [float(np.abs(sim.u[c][..., sel] - want[c][..., sel]).max()) for c in (0, 1, 4)]
    sim.close()
    return err


@pytest.mark.parametrize("order,lo,hi", [(2, 3.5, 4.5), (4, 13.0, 19.0)])
def test_outgoing_pulse_leaves_at_the_schemes_order(order, lo, hi):
    som = [_outgoing_pulse(N, "sommerfeld", order) for N in (80, 160, 320)]
    for i in range(2):
        # phi: the operator set's order; psi_r and Pi at least that fast
        assert lo < som[i][0] / som[i + 1][0] < hi, (order, som)
        for c in (1, 2):
            assert som[i][c] / som[i + 1][c] > lo, (order, c, som)
    # the reference's outflow copy reflects the pulse: no convergence, and at
    # the finest grid far above the Sommerfeld run's error
    out = [_outgoing_pulse(N, "outflow", order) for N in (160, 320)]
    assert out[0][0] / out[1][0] < 1.5
    assert out[1][0] > 100.0 * som[2][0]


def test_sommerfeld_column_is_evolved_not_filled():
    """With Sommerfeld the ghost fill leaves the outer column alone (xL, y and z
    fills as before); with outflow it copies the column inside."""
    for bc in ("sommerfeld", "outflow"):
        sim = oracle.Sim(12, 7, 9, 2, 2.0, 8.0, "Kerr_Schild", 0.0, 1.0, 0.0, bc, 0.25)
        rng = np.random.default_rng(3)
        u = rng.standard_normal(sim.shape)
        f = sim.fill(u.copy())
        j = slice(sim.g + 1, sim.g + sim.Ny)  # rows the theta fill does not write
        k = slice(1, sim.Nz)                  # planes the phi fill does not write
        if bc == "sommerfeld":
            assert np.array_equal(f[:, k, j, -1], u[:, k, j, -1])
        else:
            assert np.array_equal(f[:, k, j, -1], f[:, k, j, -2])
