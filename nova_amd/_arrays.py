"""Bulk copies between numpy arrays and the library-owned C arrays of the ABI structures."""
import ctypes as ct

import numpy as np


def fill(ptr, values, dtype):
    """Copies values (any array-like) into the C array ptr, which must hold values.size items."""
    a = np.ascontiguousarray(values, dtype=dtype).ravel()
    if a.size:
        ct.memmove(ptr, a.ctypes.data, a.nbytes)


def view(ptr, count, dtype):
    """A numpy copy of count items of the C array ptr."""
    if count == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    ctype = {np.float32: ct.c_float, np.int32: ct.c_int, np.uint32: ct.c_uint}[np.dtype(dtype).type]
    buf = ct.cast(ptr, ct.POINTER(ctype * count)).contents
    return np.frombuffer(buf, dtype=dtype, count=count).copy()


def float_array(values):
    """A ctypes float array holding values (kept alive by the caller)."""
    a = np.ascontiguousarray(values, dtype=np.float32).ravel()
    arr = (ct.c_float * a.size)()
    ct.memmove(arr, a.ctypes.data, a.nbytes)
    return arr
