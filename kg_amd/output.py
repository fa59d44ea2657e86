"""Output in the reference's formats, so its notebooks read this engine's runs.

- HDF5 snapshots (main/save_gnuplot.f90:121-144, main/global_numbers.f90:
  280-396, hdf5/hdf5_lib.f90:60-111): one file per quantity (phi.h5, rho.h5,
  psix.h5, psiy.h5, psiz.h5, pi.h5, rePhi.h5, imPhi.h5), group
  ``/refinement_1``, the coordinate datasets ``Xcoord``/``Ycoord``/``Zcoord``
  written when the file is created and one dataset ``<name>_<n>`` per save.
  Datasets are the Fortran arrays (i, j, k) over i = -g..Nx, j = 0..Ny,
  k = 0..Nz-1, so they read back as [k][j][i].  The arrays are formed on the
  GPU (``kg_diag_output``); files are written through the HDF5 C library
  (serial: one GPU holds the whole grid).
- ASCII time series ``<name>.t.asc`` (``t value`` per line, IO/save_data.f90:
  11-51) and harmonic coefficients ``<name>.lm.asc`` (``l m re im |a|`` per
  (l, m), then two blank lines, :53-110).

The reference writes ``rho`` from an array that its diagnostics never fill
(the density call is commented out); here rho.h5 holds density() as defined.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

from . import _capi as capi

# quantity code of kg_diag_output -> (file stem, dataset stem)
SNAPSHOT = [(0, "phi"), (1, "rho"), (2, "psix"), (3, "psiy"), (4, "psiz"), (5, "pi"), (6, "rePhi"), (7, "imPhi")]


def output_array(lm, quantity):
    """One of the reference's snapshot arrays, formed on the GPU: [Nz][Ny+1][Nx+g+1]."""
    out = np.empty((lm.Nz, lm.Ny + 1, lm.Nx + lm.g_pts + 1))
    lib = capi.load()
    st = lib.kg_diag_output(lm._h, int(quantity), out.ctypes.data_as(C.POINTER(C.c_double)))
    if st != capi.KG_OK:
        raise capi.KGError(st, lib.kg_diag_last_error().decode())
    return out


def snapshot_coordinates(lm):
    """x, y, z (mesh_refinement_lib.f90:126-132) over the snapshot slab, each [Nz][Ny+1][Nx+g+1]."""
    g = lm.g_pts
    dx = ((lm.desc.xmin + 1.0 * (lm.desc.xmax - lm.desc.xmin)) - lm.desc.xmin) / lm.Nx
    dy = np.arccos(-1.0) / (lm.Ny - 1)
    dz = 2.0 * np.arccos(-1.0) / (lm.Nz - 1)
    x = lm.desc.xmin + np.arange(0, lm.Nx + g + 1, dtype=np.float64) * dx
    y = -dy / 2.0 + np.arange(0, lm.Ny + 1, dtype=np.float64) * dy
    z = np.arange(0, lm.Nz, dtype=np.float64) * dz
    shape = (lm.Nz, lm.Ny + 1, lm.Nx + g + 1)
    return (np.broadcast_to(x[None, None, :], shape), np.broadcast_to(y[None, :, None], shape),
            np.broadcast_to(z[:, None, None], shape))


# ----------------------------------------------------------------------------- HDF5
class HDF5:
    """The few HDF5 C calls the snapshot files need (HDF5 >= 1.10: hid_t is 64-bit)."""

    ACC_RDONLY, ACC_RDWR, ACC_TRUNC = 0x0000, 0x0001, 0x0002
    P_DEFAULT = 0
    S_ALL = 0

    def __init__(self, path=None):
        cands = [path, os.environ.get("KG_HDF5_LIB"), ctypes.util.find_library("hdf5"), "libhdf5.so",
                 "/opt/conda/lib/libhdf5.so"]
        self.lib = None
        for c in cands:
            if not c:
                continue
            try:
                self.lib = C.CDLL(c)
                break
            except OSError:
                continue
        if self.lib is None:
            raise RuntimeError("the HDF5 C library (libhdf5.so) was not found; set KG_HDF5_LIB")
        L = self.lib
        maj, mi, rel = C.c_uint(), C.c_uint(), C.c_uint()
        L.H5get_libversion(C.byref(maj), C.byref(mi), C.byref(rel))
        if (maj.value, mi.value) < (1, 10):
            raise RuntimeError("HDF5 >= 1.10 needed, found %d.%d" % (maj.value, mi.value))
        hid = C.c_int64
        hs = C.POINTER(C.c_ulonglong)
        sig = {
            "H5open": (C.c_int, []),
            "H5Fcreate": (hid, [C.c_char_p, C.c_uint, hid, hid]),
            "H5Fopen": (hid, [C.c_char_p, C.c_uint, hid]),
            "H5Fclose": (C.c_int, [hid]),
            "H5Gcreate2": (hid, [hid, C.c_char_p, hid, hid, hid]),
            "H5Gopen2": (hid, [hid, C.c_char_p, hid]),
            "H5Gclose": (C.c_int, [hid]),
            "H5Screate_simple": (hid, [C.c_int, hs, hs]),
            "H5Sclose": (C.c_int, [hid]),
            "H5Sget_simple_extent_ndims": (C.c_int, [hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid, hs, hs]),
            "H5Dcreate2": (hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
            "H5Dopen2": (hid, [hid, C.c_char_p, hid]),
            "H5Dwrite": (C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dread": (C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dget_space": (hid, [hid]),
            "H5Dclose": (C.c_int, [hid]),
            "H5Lexists": (C.c_int, [hid, C.c_char_p, hid]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        L.H5open()
        self.DOUBLE = C.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value

    def _ok(self, v, what):
        if v < 0:
            raise RuntimeError("HDF5 call failed: " + what)
        return v

    def create_file_with_group(self, filename, group):
        f = self._ok(self.lib.H5Fcreate(filename.encode(), self.ACC_TRUNC, self.P_DEFAULT, self.P_DEFAULT), filename)
        g = self._ok(self.lib.H5Gcreate2(f, group.encode(), 0, 0, 0), group)
        self.lib.H5Gclose(g)
        self.lib.H5Fclose(f)

    def write(self, filename, group, name, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        f = self._ok(self.lib.H5Fopen(filename.encode(), self.ACC_RDWR, self.P_DEFAULT), filename)
        try:
            g = self._ok(self.lib.H5Gopen2(f, group.encode(), 0), group)
            dims = (C.c_ulonglong * data.ndim)(*data.shape)
            sp = self._ok(self.lib.H5Screate_simple(data.ndim, dims, None), "dataspace")
            d = self._ok(self.lib.H5Dcreate2(g, name.encode(), self.DOUBLE, sp, 0, 0, 0), name)
            self._ok(self.lib.H5Dwrite(d, self.DOUBLE, self.S_ALL, self.S_ALL, self.P_DEFAULT, data.ctypes.data),
                     "write " + name)
            self.lib.H5Dclose(d)
            self.lib.H5Sclose(sp)
            self.lib.H5Gclose(g)
        finally:
            self.lib.H5Fclose(f)

    def read(self, filename, path):
        f = self._ok(self.lib.H5Fopen(filename.encode(), self.ACC_RDONLY, self.P_DEFAULT), filename)
        try:
            d = self._ok(self.lib.H5Dopen2(f, path.encode(), 0), path)
            sp = self.lib.H5Dget_space(d)
            nd = self.lib.H5Sget_simple_extent_ndims(sp)
            dims = (C.c_ulonglong * nd)()
            self.lib.H5Sget_simple_extent_dims(sp, dims, None)
            out = np.empty(tuple(dims), dtype=np.float64)
            self._ok(self.lib.H5Dread(d, self.DOUBLE, self.S_ALL, self.S_ALL, self.P_DEFAULT, out.ctypes.data),
                     "read " + path)
            self.lib.H5Sclose(sp)
            self.lib.H5Dclose(d)
            return out
        finally:
            self.lib.H5Fclose(f)

    def exists(self, filename, path):
        f = self._ok(self.lib.H5Fopen(filename.encode(), self.ACC_RDONLY, self.P_DEFAULT), filename)
        try:
            # check every prefix, as H5Lexists requires
            parts = [p for p in path.split("/") if p]
            cur = ""
            for p in parts:
                cur += "/" + p
                if self.lib.H5Lexists(f, cur.encode(), 0) <= 0:
                    return False
            return True
        finally:
            self.lib.H5Fclose(f)


def write_snapshot(lm, index, directory=".", h5=None):
    """save_gnuplot's HDF5 block for save index `index` (dataset <name>_<index>);
    files and their coordinate datasets are created on first use."""
    h5 = h5 or HDF5()
    X = None
    for q, name in SNAPSHOT:
        fn = os.path.join(directory, name + ".h5")
        if not os.path.exists(fn):
            if X is None:
                X = snapshot_coordinates(lm)
            h5.create_file_with_group(fn, "/refinement_1")
            for cname, arr in zip(("Xcoord", "Ycoord", "Zcoord"), X):
                h5.write(fn, "/refinement_1", cname, arr)
        h5.write(fn, "/refinement_1", "%s_%d" % (name, index), output_array(lm, q))


# ----------------------------------------------------------------------------- ASCII
def save_t(name, t, f):
    """Append `t f` to name.t.asc (save_t / save0d)."""
    with open(name + ".t.asc", "a") as fh:
        fh.write("%26.17E %26.17E\n" % (t, f))


def save_a(name, alm, lmax):
    """Append the (l, m) table of a complex alm [lm] to name.lm.asc (save_a / savealm)."""
    with open(name + ".lm.asc", "a") as fh:
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                a = complex(alm[l * l + l + m])
                fh.write("%6d %6d %26.17E %26.17E %26.17E\n" % (l, m, a.real, a.imag, abs(a)))
        fh.write("\n\n")
