"""Argument checks of the C ABI that run before any GPU work, on the CPU.

Every entry point validates its arguments first and returns a status with a
message (the reference prints and stops on the same conditions,
lines_method_lib.f90:529-533,744-748); these tests call the library without a
GPU and check that each bad descriptor or argument is refused with the right
status, and that a good descriptor gets as far as the device check.
"""
import ctypes as C

import pytest

from kg_amd import _capi as capi
from kg_amd.lines_method import make_desc


def _lib():
    return capi.load()


def _desc(**kw):
    args = dict(nvars=5, g_pts=2, xmin=1.8, xmax=20.0, Nx=32, Ny=9, Nz=9, tmin=0.0, CFL=0.25,
                boundary_type_rmax="outflow", metric="Kerr_Schild", a=0.0, M_bh=1.0, m=0.1, scheme="rk4", dt=0.0,
                is_complex=False, operators="ref2")
    args.update(kw)
    return make_desc(**args)


def _create(d):
    lib = _lib()
    h = C.c_void_p()
    st = lib.kg_create(C.byref(d), C.byref(h))
    msg = lib.kg_last_error(None)
    return st, (msg.decode() if msg else ""), h


@pytest.mark.parametrize("field,value,status,text", [
    ("nvars", 4, capi.KG_ERR_INVALID, "nvars"),
    ("g_pts", 0, capi.KG_ERR_INVALID, "g_pts"),
    ("Nx", 3, capi.KG_ERR_INVALID, "Nx"),
    ("Ny", 3, capi.KG_ERR_INVALID, "Ny"),
    ("Nz", 3, capi.KG_ERR_INVALID, "Nz"),
    ("scheme", 7, capi.KG_ERR_INVALID, "scheme"),
    ("op_set", 5, capi.KG_ERR_INVALID, "op_set"),
])
def test_create_refuses_bad_descriptors(field, value, status, text):
    d = _desc()
    setattr(d, field, value)
    st, msg, _ = _create(d)
    assert st == status and text in msg, (st, msg)


def test_create_refuses_bad_geometry_and_combinations():
    st, msg, _ = _create(_desc(xmin=5.0, xmax=5.0))
    assert st == capi.KG_ERR_INVALID and "xmax" in msg
    st, msg, _ = _create(_desc(CFL=0.0, dt=0.0))
    assert st == capi.KG_ERR_INVALID and "dt" in msg
    st, msg, _ = _create(_desc(operators="fd4", scheme="pirk2"))
    assert st == capi.KG_ERR_UNSUPPORTED and "fd4" in msg
    d = _desc()
    d.metric = 2  # MKS: the reference's rhs_pi stops on it (lines_method_lib.f90:744-748)
    st, msg, _ = _create(d)
    assert st == capi.KG_ERR_UNSUPPORTED and "Metric" in msg
    d = _desc()
    d.bc_rmax = 7
    st, msg, _ = _create(d)
    assert st == capi.KG_ERR_UNSUPPORTED and "boundary" in msg


def test_good_descriptor_reaches_the_device_check():
    """Without a GPU a valid descriptor is refused only for want of a device."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: kg_create would succeed")
    for ops, sch, bc in (("ref2", "pirk2", "outflow"), ("fd4", "rk4", "sommerfeld"), ("ref2", "rk3", "inflow")):
        st, msg, _ = _create(_desc(operators=ops, scheme=sch, boundary_type_rmax=bc))
        assert st == capi.KG_ERR_NO_DEVICE, (ops, sch, bc, st, msg)


def test_batch_and_refinement_and_probe_argument_checks():
    lib = _lib()
    # a batch needs at least one handle; nsteps >= 0
    assert lib.kg_batch_step(None, 0, 1) == capi.KG_ERR_INVALID
    assert lib.kg_batch_step_noalloc(None, 0, 1) == capi.KG_ERR_INVALID
    # the refinement buffer is even and in [8, 64]
    d = _desc(scheme="pirk2")
    r = C.c_void_p()
    for buf in (7, 6, 66):
        rd = capi.KGRefineDesc()
        rd.r_patch, rd.buffer = 6.0, buf
        assert lib.kg_refined_create(C.byref(d), C.byref(rd), C.byref(r)) == capi.KG_ERR_INVALID, buf
        assert b"buffer" in lib.kg_refined_last_error(None)
    # the HBM probe's kinds and sizes
    ms = C.c_double()
    for kind, n, blocks, reps in ((6, 1024, 8, 1), (0, 1023, 8, 1), (0, 1024, 0, 1), (0, 1024, 8, 0)):
        assert lib.kg_hbm_probe(0, kind, n, blocks, reps, C.byref(ms)) == capi.KG_ERR_INVALID
    # the spectral batch needs projectors and outputs
    assert lib.kg_spec_radial_batch(None, 0, 0, 0, None, 0) == capi.KG_ERR_INVALID
