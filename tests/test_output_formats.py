"""Output formats on the host: the HDF5 layout of the reference's snapshot
files (hdf5/hdf5_lib.f90:60-111, main/global_numbers.f90:280-396) and its
ASCII time series / harmonic tables (IO/save_data.f90:11-110)."""
import numpy as np
import pytest

from kg_amd.output import HDF5, save_a, save_t


def _h5():
    try:
        return HDF5()
    except RuntimeError as e:  # no HDF5 C library on this machine
        pytest.skip(str(e))


def test_hdf5_group_and_datasets_round_trip(tmp_path):
    h5 = _h5()
    fn = str(tmp_path / "phi.h5")
    h5.create_file_with_group(fn, "/refinement_1")
    a = np.random.default_rng(0).normal(size=(5, 4, 7))
    h5.write(fn, "/refinement_1", "Xcoord", a)
    h5.write(fn, "/refinement_1", "phi_3", 2 * a)
    assert h5.exists(fn, "/refinement_1/Xcoord") and h5.exists(fn, "/refinement_1/phi_3")
    assert not h5.exists(fn, "/refinement_1/phi_4")
    assert np.array_equal(h5.read(fn, "/refinement_1/Xcoord"), a)
    assert np.array_equal(h5.read(fn, "/refinement_1/phi_3"), 2 * a)


def test_ascii_time_series_and_lm_table(tmp_path):
    name = str(tmp_path / "Mass_scalar")
    save_t(name, 0.0, 1.5)
    save_t(name, 0.25, -2.0e-3)
    d = np.loadtxt(name + ".t.asc")
    assert d.shape == (2, 2) and d[1, 0] == 0.25 and d[1, 1] == -2.0e-3
    lmax = 2
    alm = np.arange(9) + 1j * np.arange(9)[::-1]
    save_a(str(tmp_path / "Harmonic_coeficients01"), alm, lmax)
    text = open(str(tmp_path / "Harmonic_coeficients01") + ".lm.asc").read()
    assert text.endswith("\n\n\n")
    rows = np.loadtxt(str(tmp_path / "Harmonic_coeficients01") + ".lm.asc")
    assert rows.shape == (9, 5)
    # rows in the reference's (l, m) order with Re, Im, |a|
    assert list(rows[:, 0]) == [0, 1, 1, 1, 2, 2, 2, 2, 2] and list(rows[:, 1]) == [0, -1, 0, 1, -2, -1, 0, 1, 2]
    np.testing.assert_allclose(rows[:, 4], np.abs(alm))
