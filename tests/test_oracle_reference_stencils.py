"""Pin the oracle's stencils against the reference's own code.

oracle/_ref/libkgref_fd.so is tools/finite_differences.f90 of the reference,
compiled from its source with amdflang (oracle/Makefile).  The oracle's C
restatement of D_r, D_theta, D_phi and the upwind A_r must agree with it bit
for bit on complex data (real and imaginary parts separately), including the
edge rows, the pole row j = -g that reads phi+pi (finite_differences.f90:218-225)
and the sign-reversed phi row k = Nz (:335).
"""
import zlib

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.have_reference(),
                                reason="oracle/_ref not built (needs the reference source)")

SHAPES = [(13, 9, 10), (21, 13, 18), (20, 12, 17), (9, 7, 5), (70, 38, 66)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("name", ["dx1", "dx2", "dx3", "advec_x"])
def test_stencil_bit_exact(shape, name):
    ni, nj, nk = shape
    rng = np.random.default_rng(zlib.crc32(repr((shape, name)).encode()))
    fr = rng.standard_normal((nk, nj, ni))
    fi = rng.standard_normal((nk, nj, ni))
    beta = rng.choice([-0.7, 0.0, 1.3], size=(nk, nj, ni))
    h = float(rng.uniform(0.05, 0.5))
    ref = oracle.ref_stencil(name, fr + 1j * fi, h, 2, beta=beta)
    got_r = oracle.stencil(name, fr, h, beta=beta)
    got_i = oracle.stencil(name, fi, h, beta=beta)
    assert np.array_equal(ref.real, got_r)
    assert np.array_equal(ref.imag, got_i)


def test_reference_quirk_rows_are_what_the_oracle_states():
    """The two reference quirks seen through the reference itself: D_theta at
    j = -g pairs f(j+1, k) with f(j+2, k_target) and D_phi at k = Nz uses
    (f(Nz-1) - f(2)) -- checked on index-tag data so each output names its
    inputs."""
    ni, nj, nk = 9, 8, 11
    Nz = nk - 1
    kk, jj, ii = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    tag = (kk * 1000 + jj * 10).astype(np.float64)  # depends on (j, k) only
    h = 0.5  # res = 1
    d2 = oracle.ref_stencil("dx2", tag.astype(complex), h, 2).real
    for k in range(nk):
        kt = (k + Nz // 2) % (Nz - 1)
        assert np.all(d2[k, 0, :] == -(k * 1000 + 10) + (kt * 1000 + 20))
    d3 = oracle.ref_stencil("dx3", tag.astype(complex), h, 2).real
    assert np.all(d3[Nz] == tag[Nz - 1] - tag[2])
    assert np.all(d3[0] == tag[1] - tag[Nz - 2])


@pytest.mark.parametrize("shape", [(13, 9, 10), (21, 13, 18), (9, 7, 5), (70, 38, 66)])
@pytest.mark.parametrize("name", ["dx1_4", "dx2_4", "dx3_4", "advec_x_4"])
def test_fd4_stencil_bit_exact(shape, name):
    """The 4th-order set (finite_differences.f90:124-151, 241-268, 340-367,
    922-957) against the reference's compiled functions, every row."""
    ni, nj, nk = shape
    rng = np.random.default_rng(zlib.crc32(repr((shape, name, 4)).encode()))
    f = rng.standard_normal((nk, nj, ni))
    beta = rng.choice([-0.7, 0.0, 1.3], size=(nk, nj, ni))
    h = float(rng.uniform(0.05, 0.5))
    ref = oracle.ref_stencil(name, f, h, 2, beta=beta)
    got = oracle.stencil(name, f, h, beta=beta)
    assert np.array_equal(ref, got)
