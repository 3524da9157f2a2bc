"""The reference's on-disk formats through the C ABI (include/inmsom_bt.h):
DIOGIN direct-access records (Service/input_output_data.f90: rdstd,
wdstd, rdstd8, wdstd8) and GrADS descriptors (Service/rw_ctl_file.f90:
ctl_file_write).  Arrays are (nz, ny, nx) or (ny, nx) C-order numpy
arrays, the same memory as the reference's column-major (nx, ny, nz);
grid indices (nxb .. nze) are the reference's 1-based ones."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import check, lib

UNDEF = np.float32(-1.0e32)  # Inc/reclen.fi


def _field(field, dtype):
    a = np.asarray(field)
    if a.dtype != dtype or not a.flags.c_contiguous:
        raise ValueError(f"field must be a C-contiguous {np.dtype(dtype).name} array")
    return a


def _dims(a):
    return (1,) + a.shape if a.ndim == 2 else a.shape


def _box(a, lu, box):
    nz, ny, nx = _dims(a)
    if lu.shape != (ny, nx):
        raise ValueError("lu must be (ny, nx)")
    if box is None:
        box = (1, nx, 1, ny, 1, nz)
    return nx, ny, nz, box


def wdstd(path: str, fname: str, nfield: int, field: np.ndarray, lu: np.ndarray, box=None):
    """input_output_data.f90:206-327: field (float32) records of the box
    (nxb, nxe, nyb, nye, nzb, nze), land as undef in the file; land points
    of ``field`` are left 0 as the reference leaves them."""
    a = _field(field, np.float32)
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    nx, ny, nz, box = _box(a, lu, box)
    check(lib().inmsom_bt_wdstd(path.encode(), fname.encode(), int(nfield), a.ctypes.data_as(C.POINTER(C.c_float)),
                                lu.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz, *map(int, box)),
          "inmsom_bt_wdstd")


def rdstd(path: str, fname: str, nfield: int, field: np.ndarray, lu: np.ndarray, box=None):
    """input_output_data.f90:117-204: fills the box of ``field`` (float32)
    from the file's records, land points 0."""
    a = _field(field, np.float32)
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    nx, ny, nz, box = _box(a, lu, box)
    check(lib().inmsom_bt_rdstd(path.encode(), fname.encode(), int(nfield), a.ctypes.data_as(C.POINTER(C.c_float)),
                                lu.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz, *map(int, box)),
          "inmsom_bt_rdstd")


def wdstd8(path: str, fname: str, nfield: int, field: np.ndarray, lu: np.ndarray, box=None):
    """input_output_data.f90:419-533, float64 records."""
    a = _field(field, np.float64)
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    nx, ny, nz, box = _box(a, lu, box)
    check(lib().inmsom_bt_wdstd8(path.encode(), fname.encode(), int(nfield), a.ctypes.data_as(C.POINTER(C.c_double)),
                                 lu.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz, *map(int, box)),
          "inmsom_bt_wdstd8")


def rdstd8(path: str, fname: str, nfield: int, field: np.ndarray, lu: np.ndarray, box=None):
    """input_output_data.f90:329-417, float64 records."""
    a = _field(field, np.float64)
    lu = np.ascontiguousarray(lu, dtype=np.float32)
    nx, ny, nz, box = _box(a, lu, box)
    check(lib().inmsom_bt_rdstd8(path.encode(), fname.encode(), int(nfield), a.ctypes.data_as(C.POINTER(C.c_double)),
                                 lu.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz, *map(int, box)),
          "inmsom_bt_rdstd8")


def ctl_file_write(fname: str, undef: float, nx: int, ny: int, nz: int, nt: int, xtype: int, x0, hx: float,
                   ytype: int, y0, hy: float, ztype: int, z0, hz: float, yr_type: int, year0: int, month0: int,
                   day0: int, hour0: int, minute0: int, ht: float, title: str, varname: str):
    """rw_ctl_file.f90:1-138 with the reference's argument list: x0/y0/z0
    hold the first value (type 0, with the step) or every value (type 1)."""
    def arr(v, n, t):
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))
        if a.size < (n if t else 1):
            raise ValueError("axis values shorter than the axis")
        return a
    ax, ay, az = arr(x0, nx, xtype), arr(y0, ny, ytype), arr(z0, nz, ztype)
    dp = C.POINTER(C.c_double)
    check(lib().inmsom_bt_ctl_write(fname.encode(), float(undef), nx, ny, nz, nt, xtype, ax.ctypes.data_as(dp),
                                    float(hx), ytype, ay.ctypes.data_as(dp), float(hy), ztype, az.ctypes.data_as(dp),
                                    float(hz), yr_type, year0, month0, day0, hour0, minute0, float(ht),
                                    title.encode(), varname.encode()), "inmsom_bt_ctl_write")
