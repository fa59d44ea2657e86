"""Edge cases of the C ABI on the GPU: sizes at and beyond the limits, empty
step counts, repeated or mismatched handles, refinement patches that do not
fit; each is either a no-op or refused with a status and a message, and
leaves the handles usable."""
import numpy as np
import pytest

from kg_amd import KGError, LinesMethod
from kg_amd import _capi as capi
from kg_amd.batch import batch_step
from kg_amd.initial_data import shell_state
from kg_amd.refinement import RefinedLinesMethod

pytestmark = pytest.mark.gpu


def _small(**kw):
    args = dict(Nx=24, Ny=9, Nz=9, xmin=1.8, xmax=10.0, m=0.2, scheme="rk4")
    args.update(kw)
    lm = LinesMethod(**args)
    u0 = shell_state(args["Nx"], args["Ny"], args["Nz"], 2, args["xmin"], args["xmax"], 5.0, 1.0, 1.0, {(1, 0): 1.0})
    lm.set_u_soa(u0[None])
    return lm, u0


def test_grid_beyond_the_field_size_limit_is_refused():
    """One field of one system must stay below 2 GiB (32-bit byte offsets in
    the kernels): 8192 x 2048 x 16 is refused before any allocation, and a
    grid just below the limit is accepted."""
    print("too-large grid", flush=True)
    with pytest.raises(KGError) as ei:
        LinesMethod(Nx=8192, Ny=2049, Nz=17, xmin=1.8, xmax=200.0)
    assert ei.value.status == capi.KG_ERR_INVALID and "2 GiB" in str(ei.value)
    lm = LinesMethod(Nx=4096, Ny=1025, Nz=33, xmin=1.8, xmax=200.0)  # 4112 x 1030 x 34 doubles: 1.15 GB a field
    lm.close()


def test_zero_steps_change_nothing():
    lm, u0 = _small()
    t0 = lm.t
    lm.step(0)
    batch_step([lm], 0)
    assert lm.t == t0
    assert np.array_equal(lm.u_soa()[0], u0)
    lm.step(1)
    assert lm.t > t0
    lm.close()


def test_bad_tuning_and_batches_are_refused_and_handles_stay_usable():
    a, _ = _small()
    b, _ = _small(scheme="pirk2")
    for kc, tj in ((0, 8), (5000, 8), (8, 5)):
        with pytest.raises(KGError):
            a.set_tuning(kc, tj)
    with pytest.raises(KGError):
        batch_step([a, a], 1)  # repeated handle
    with pytest.raises(KGError):
        batch_step([a, b], 1)  # different schemes
    a.step(2)
    batch_step([a], 2)
    assert np.isfinite(a.u_soa()).all()
    a.close()
    b.close()


def test_refinement_patch_must_fit_the_base_grid():
    print("refinement patch bounds", flush=True)
    for r_patch in (1.9, 9.9):  # too close to the inner edge; no room for the buffer before the outer edge
        with pytest.raises(KGError):
            RefinedLinesMethod(r_patch, Nx=48, Ny=9, Nz=9, xmin=1.8, xmax=10.0, scheme="pirk2")
    ok = RefinedLinesMethod(5.0, Nx=48, Ny=9, Nz=9, xmin=1.8, xmax=10.0, scheme="pirk2")
    ok.step(1)
    ok.close()
