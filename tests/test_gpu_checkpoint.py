"""Checkpoint / restart (lines_method_lib.f90:599-667): a run stopped,
written, read back into a fresh handle and continued ends bit-identical to
the run that never stopped -- for PIRK this includes the stage arrays whose
Pi the next step reads (lines_method_lib.f90:288-289,366-367)."""
import numpy as np
import pytest

import kg_amd
from kg_amd.initial_data import random_lm_coeffs, shell_state

pytestmark = pytest.mark.gpu


def _make(scheme, cplx=False):
    lm = kg_amd.LinesMethod(Nx=20, Ny=12, Nz=14, xmin=1.9, xmax=12.0, a=0.6, m=0.3, boundary_type_rmax="inflow",
                            scheme=scheme, is_complex=cplx)
    return lm


@pytest.mark.parametrize("scheme", ["pirk2", "pirk3", "rk3", "rk4"])
@pytest.mark.parametrize("cplx", [False, True])
def test_restart_is_bit_identical(tmp_path, scheme, cplx):
    u0 = shell_state(20, 12, 14, 2, 1.9, 12.0, 6.0, 1.5, 1.0, random_lm_coeffs(5, 3, 0.3))
    u0 = np.stack([u0, 0.5 * u0] if cplx else [u0])
    ref = _make(scheme, cplx)
    ref.set_u_soa(u0)
    ref.boundary()
    ref.step(4)
    path = tmp_path / "ck.bin"
    ref.check_point(path)
    ref.step(5)
    want, t_want = ref.u_soa(), ref.t

    new = _make(scheme, cplx)
    new.read_from_check_point(path)
    new.step(5)
    assert np.array_equal(new.u_soa(), want)
    assert new.t == t_want


def test_checkpoint_of_another_grid_is_refused(tmp_path):
    a = _make("rk4")
    a.check_point(tmp_path / "a.bin")
    b = kg_amd.LinesMethod(Nx=24, Ny=12, Nz=14, xmin=1.9, xmax=12.0, scheme="rk4")
    with pytest.raises(kg_amd.KGError):
        b.read_from_check_point(tmp_path / "a.bin")
    with pytest.raises(kg_amd.KGError):
        b.read_from_check_point(tmp_path / "missing.bin")


def test_checkpoint_of_another_operator_set_is_refused(tmp_path):
    # same grid and scheme, the other stencil set: restarting would continue
    # with a different right-hand side
    kw = dict(Nx=20, Ny=12, Nz=14, xmin=1.9, xmax=12.0, a=0.6, m=0.3, boundary_type_rmax="inflow", scheme="rk4")
    a = kg_amd.LinesMethod(operators="fd4", **kw)
    a.check_point(tmp_path / "fd4.bin")
    b = kg_amd.LinesMethod(operators="ref2", **kw)
    with pytest.raises(kg_amd.KGError):
        b.read_from_check_point(tmp_path / "fd4.bin")
    b.check_point(tmp_path / "ref2.bin")
    with pytest.raises(kg_amd.KGError):
        a.read_from_check_point(tmp_path / "ref2.bin")
    a.read_from_check_point(tmp_path / "fd4.bin")
