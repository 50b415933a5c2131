"""MatrixMarket files (psb_mmio_mod: util/psb_mmio_mod.F90, util/psb_d_mmio_impl.f90).

The parsing and formatting run in the library's host code
(psb_hip_mm_read_coo / _read_vec / _write_coo / _write_vec); failures raise
PsbError carrying the reference's info value (901 cannot open, 902 short
file, 904 unsupported type or symmetry, 905 bad entry, 909 unsupported
object or format).
"""
import ctypes as C

import numpy as np

from . import _native as N
from .mat import psb_d_coo_sparse_mat


def _unsupported_info(path, fmt):
    """909 when the banner's object/format is not "matrix <fmt>", else 904
    (the data type or symmetry)."""
    try:
        with open(path) as f:
            words = f.readline().lower().split()
    except OSError:
        return 909
    return 904 if words[1:3] == ["matrix", fmt] else 909


def _raise(rc, name, path, fmt="coordinate"):
    info = {N.RUNTIME: 901, N.INVALID: 905}.get(rc, rc)
    if rc == N.UNSUPPORTED:
        info = _unsupported_info(path, fmt)
    raise N.PsbError(info, name, f"{path}")


def mm_mat_read(filename):
    """mm_mat_read (util/psb_d_mmio_impl.f90:315-440): a coordinate file into a
    fixed psb_d_coo_sparse_mat (rows ascending, columns ascending, duplicates
    added); symmetric files are de-symmetrized."""
    path = str(filename).encode()
    nr, nc, nz = C.c_int(0), C.c_int(0), C.c_int64(0)
    rc = N.fn("psb_hip_mm_read_coo")(path, 0, C.byref(nr), C.byref(nc), C.byref(nz), None, None,
                                     None)
    if rc:
        _raise(rc, "mm_mat_read", filename)
    cap = nz.value
    ia = np.empty(max(cap, 1), np.int32)
    ja = np.empty(max(cap, 1), np.int32)
    val = np.empty(max(cap, 1), np.float64)
    rc = N.fn("psb_hip_mm_read_coo")(path, cap, C.byref(nr), C.byref(nc), C.byref(nz),
                                     ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    if rc:
        _raise(rc, "mm_mat_read", filename)
    k = nz.value
    coo = psb_d_coo_sparse_mat(nr.value, nc.value, ia[:k].copy(), ja[:k].copy(), val[:k].copy())
    return coo.fix()


def mm_array_read(filename):
    """mm_array_read (mm_dvet_read, util/psb_d_mmio_impl.f90:32-102): the first
    column of a "matrix array real general" file."""
    path = str(filename).encode()
    n = C.c_int64(0)
    rc = N.fn("psb_hip_mm_read_vec")(path, 0, C.byref(n), None)
    if rc:
        _raise(rc, "mm_array_read", filename, "array")
    b = np.empty(n.value, np.float64)
    rc = N.fn("psb_hip_mm_read_vec")(path, n.value, C.byref(n), b.ctypes.data if n.value else None)
    if rc:
        _raise(rc, "mm_array_read", filename, "array")
    return b


def mm_mat_write(a, title, filename):
    """mm_mat_write (util/psb_d_mmio_impl.f90:443-486, through the COO print):
    a is a psb_d_coo_sparse_mat (written in its entry order)."""
    ia = np.ascontiguousarray(a.ia, np.int32)
    ja = np.ascontiguousarray(a.ja, np.int32)
    val = np.ascontiguousarray(a.val, np.float64)
    rc = N.fn("psb_hip_mm_write_coo")(str(filename).encode(), title.encode(), a.nr, a.nc,
                                      len(val), ia.ctypes.data, ja.ctypes.data, val.ctypes.data)
    if rc:
        _raise(rc, "mm_mat_write", filename)


def mm_array_write(b, title, filename):
    """mm_array_write (mm_dvet1_write, util/psb_d_mmio_impl.f90:226-280)."""
    b = np.ascontiguousarray(b, np.float64)
    rc = N.fn("psb_hip_mm_write_vec")(str(filename).encode(), title.encode(), len(b),
                                      b.ctypes.data)
    if rc:
        _raise(rc, "mm_array_write", filename)
