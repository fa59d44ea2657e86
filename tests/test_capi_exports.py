"""The C-ABI library loads and exports every entry point include/kg.h declares.

CPU-only: the library is loaded and its symbols resolved; nothing that needs a
GPU is called except kg_create, which must refuse cleanly without one.
"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ("kg_create", "kg_destroy", "kg_pirk2", "kg_pirk3", "kg_rk3", "kg_rk4", "kg_step",
                 "kg_upload_state", "kg_download_state", "kg_fill_ghosts", "kg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from kg_amd import _capi
    lib = _capi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _capi.SIGNATURES, "ctypes binding missing for " + name


def test_desc_struct_matches_header_layout():
    from kg_amd import _capi
    # 10 int32 + 8 doubles, no padding
    assert C.sizeof(_capi.KGDesc) == 10 * 4 + 8 * 8
    assert _capi.KGDesc.xmin.offset == 40


def test_spec_desc_struct_matches_header_layout():
    from kg_amd import _capi
    # 4 int32 + 2 doubles
    assert C.sizeof(_capi.KGSpecDesc) == 4 * 4 + 2 * 8
    assert _capi.KGSpecDesc.R_max.offset == 16


def test_ic_desc_struct_matches_header_layout():
    from kg_amd import _capi
    # 2 int32, 3 doubles, uint64, pointer
    assert C.sizeof(_capi.KGICDesc) == 8 + 3 * 8 + 8 + 8
    assert _capi.KGICDesc.mu.offset == 8 and _capi.KGICDesc.uniforms.offset == 40


def test_random_field_rejects_bad_sizes_before_touching_a_device():
    from kg_amd import _capi
    lib = _capi.load()
    out = (C.c_double * 2)()
    for n in (0, 3, 6, 2048):
        assert lib.kg_ic_random_field(0, n, 5.0, 1.0, 0, None, out) == _capi.KG_ERR_INVALID
    assert b"power of two" in lib.kg_ic_last_error()
    assert lib.kg_ic_random_field(0, 16, -1.0, 1.0, 0, None, out) == _capi.KG_ERR_INVALID
    assert lib.kg_ic_random_field(0, 16, 5.0, 1.0, 0, None, None) == _capi.KG_ERR_INVALID


def test_create_without_gpu_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    import kg_amd
    with pytest.raises(kg_amd.KGError) as ei:
        kg_amd.LinesMethod(Nx=16, Ny=9, Nz=9)
    assert ei.value.status in (5, 3)


def test_unsupported_metric_is_rejected_like_the_reference():
    """rhs_pi stops for metrics other than Kerr_Schild and Schw_EF
    (lines_method_lib.f90:744-748); kg_create returns KG_ERR_UNSUPPORTED."""
    from kg_amd import _capi
    lib = _capi.load()
    d = _capi.KGDesc(nvars=5, g_pts=2, Nx=16, Ny=9, Nz=9, scheme=0, metric=_capi.KG_METRIC_MINKOWSKI,
                     bc_rmax=0, is_complex=0, xmin=1.8, xmax=9.0, CFL=0.25)
    h = C.c_void_p()
    st = lib.kg_create(C.byref(d), C.byref(h))
    assert st == _capi.KG_ERR_UNSUPPORTED
    assert b"not implemented" in lib.kg_last_error(None)
    d.metric = _capi.KG_METRIC_KERR_SCHILD
    d.nvars = 4
    assert lib.kg_create(C.byref(d), C.byref(h)) == _capi.KG_ERR_INVALID
