"""pandas-HDF5 tables (HDF5DataSet<D>, src/HDF5DataSet.hpp:18-167), through
libmmc_h5.so (include/minimc_amd_h5.h).

read(path, D) returns (axes, values) with values shaped by the level sizes
(level 0 slowest).  write(path, axes, values) writes the same layout.  The
reference's checks and messages are kept: a missing file raises
"File not found: <path>" and a level count other than D raises
"<path>: Expected D dimensions, but got L" (ReadPandasAxis, :89-101).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MMC_H5_LIB: another build of the same library (the sanitizer test loads one)
LIB_PATH = os.environ.get("MMC_H5_LIB") or os.path.join(_HERE, "_build", "libmmc_h5.so")
MAX_LEVELS = 8
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmmc_h5.so not built (needs the HDF5 C library): run "
                               "`python -m minimc_amd.build`")
        L = C.CDLL(LIB_PATH)
        dpp = C.POINTER(C.POINTER(C.c_double))
        L.mmc_h5_info.argtypes = [C.c_char_p, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint64), C.c_int32, C.POINTER(C.c_uint64)]
        L.mmc_h5_read.argtypes = [C.c_char_p, C.c_int32, dpp, C.POINTER(C.c_double)]
        L.mmc_h5_write.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_uint64), dpp,
                                   C.POINTER(C.c_double)]
        L.mmc_h5_last_error.restype = C.c_char_p
        for f in (L.mmc_h5_info, L.mmc_h5_read, L.mmc_h5_write):
            f.restype = C.c_int
        _lib = L
    return _lib


def _raise():
    raise RuntimeError(lib().mmc_h5_last_error().decode())


def read(path, D=None):
    """(axes, values) of a pandas-HDF5 file; D, if given, is checked."""
    L = lib()
    p = str(path).encode()
    levels = C.c_int32()
    sizes = (C.c_uint64 * MAX_LEVELS)()
    n = C.c_uint64()
    if L.mmc_h5_info(p, C.byref(levels), sizes, MAX_LEVELS, C.byref(n)):
        _raise()
    d = levels.value
    if D is not None and d != D:
        raise RuntimeError("%s: Expected %d dimensions, but got %d" % (path, D, d))
    if not 1 <= d <= MAX_LEVELS:
        raise RuntimeError("%s: %d levels" % (path, d))
    shape = tuple(int(sizes[i]) for i in range(d))
    if int(np.prod(shape)) != n.value:
        raise RuntimeError("%s: %d values for levels of sizes %s" % (path, n.value, shape))
    axes = [np.zeros(s) for s in shape]
    values = np.zeros(n.value)
    ptrs = (C.POINTER(C.c_double) * d)(*[a.ctypes.data_as(C.POINTER(C.c_double))
                                         for a in axes])
    if L.mmc_h5_read(p, d, ptrs, values.ctypes.data_as(C.POINTER(C.c_double))):
        _raise()
    return axes, values.reshape(shape)


def write(path, axes, values):
    """Writes values (shape = the axes' lengths, or flat) over the axes."""
    axes = [np.ascontiguousarray(a, dtype=np.float64) for a in axes]
    shape = tuple(len(a) for a in axes)
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    if v.size != int(np.prod(shape)):
        raise ValueError("%d values for levels of sizes %s" % (v.size, shape))
    sizes = (C.c_uint64 * len(axes))(*shape)
    ptrs = (C.POINTER(C.c_double) * len(axes))(*[a.ctypes.data_as(C.POINTER(C.c_double))
                                                 for a in axes])
    if lib().mmc_h5_write(str(path).encode(), len(axes), sizes, ptrs,
                          v.ctypes.data_as(C.POINTER(C.c_double))):
        _raise()


def read_table(path):
    """A one-level file as a ContinuousMap table (HDF5DataSet<1>::
    ToContinuousMap, :49-63): (keys, values)."""
    axes, values = read(path, 1)
    return axes[0], values
