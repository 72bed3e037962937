"""Legacy wrappers of fftcorr/array/row_major_array.pyx:7-35: hold a 3D
C-contiguous float64 / complex128 array and expose its shape."""

import numpy as np

from fftcorr_amd._arrays import check_row_major


class _RowMajorArrayPtr3D:
    _dtype = None

    def __init__(self, arr):
        check_row_major(arr, 3, self._dtype)
        self._arr = arr

    @property
    def shape(self):
        return self._arr.shape

    @property
    def data(self):
        return self._arr


class RowMajorArrayPtr3D_Float(_RowMajorArrayPtr3D):
    _dtype = np.float64


class RowMajorArrayPtr3D_Complex(_RowMajorArrayPtr3D):
    _dtype = np.complex128
