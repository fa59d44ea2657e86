"""Host-side helpers of the diagnostics (no GPU): the reference's detector
radii (MoL/diagnostic_lib.f90:124-134) and their spectral radius indices."""
import numpy as np
import pytest

from kg_amd.diagnostics import detector_indices, detector_radii


def test_detector_radii_follow_the_reference():
    r = detector_radii(5, a=0.0, xmax=60.0)
    assert r[0] == 2.0 and r[-1] == pytest.approx(59.4)
    assert np.allclose(np.diff(r[:-1]), (59.4 - 2.0) / 5)
    idx, keep = detector_indices(r, 1.8 + np.arange(256) * 0.22734375)
    assert keep == [0, 1, 2, 3, 4] and idx[0] == 0
