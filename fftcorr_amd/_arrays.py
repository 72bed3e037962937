"""numpy argument checks of the reference's as_RowMajorArrayPtr
(fftcorr/array/numpy_adaptor.h:45-72): same checks, order and messages."""

import numpy as np


def check_row_major(arr, ndim, dtype=np.float64):
    if not isinstance(arr, np.ndarray):
        raise TypeError("expected a numpy.ndarray, got %s" % type(arr).__name__)
    if arr.ndim != ndim:
        raise RuntimeError("Number of dimensions of numpy array is not equal to N.")
    if arr.dtype != dtype:
        raise RuntimeError("Numpy array does not have the expected type.")
    if not arr.flags["C_CONTIGUOUS"]:
        raise RuntimeError("Numpy array is not C-contiguous.")
    if not arr.flags["WRITEABLE"]:
        raise RuntimeError("Numpy array is not writeable.")
    return arr


def readonly_copy(a):
    a = np.array(a, copy=True)
    a.setflags(write=False)
    return a
