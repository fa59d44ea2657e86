"""The oracle's stencils against committed outputs of the reference's own
functions (tests/golden/reference_stencils.npz, written by
tests/golden/make_reference_stencils.py from tools/finite_differences.f90
compiled from its source): the live 2nd-order set on complex data in the
reference's padded layout (edge rows, the pole row that reads phi+pi, the
sign-reversed phi row) and the 4th-order set, every row, beta of all signs.
Bit equality; this pin holds where the reference itself is not available."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_stencils.npz")


def _cases():
    with np.load(GOLDEN) as z:
        keys = sorted({k.rsplit("_", 1)[0] for k in z.files})
    return keys


@pytest.mark.parametrize("case", _cases())
def test_oracle_matches_reference_output(case):
    with np.load(GOLDEN) as z:
        f, beta, h, ref = z[case + "_f"], z[case + "_beta"], float(z[case + "_h"]), z[case + "_out"]
    name = case.rsplit("_", 1)[0]  # the stencil, without the shape
    if np.iscomplexobj(f):
        got_r = oracle.stencil(name, np.ascontiguousarray(f.real), h, beta=beta)
        got_i = oracle.stencil(name, np.ascontiguousarray(f.imag), h, beta=beta)
        assert np.array_equal(got_r, ref.real), case
        assert np.array_equal(got_i, ref.imag), case
    else:
        assert np.array_equal(oracle.stencil(name, f, h, beta=beta), ref), case


def test_golden_cases_cover_every_stencil():
    names = {c.rsplit("_", 1)[0] for c in _cases()}
    assert names == {"dx1", "dx2", "dx3", "advec_x", "dx1_4", "dx2_4", "dx3_4", "advec_x_4"}, names
