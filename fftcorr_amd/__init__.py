"""MI355X implementation of fftcorr's FFT-based anisotropic 2-point
correlation and power-spectrum multipole path.

The sub-packages mirror the reference's Python API (fftcorr.grid,
fftcorr.particle_mesh, fftcorr.correlate, fftcorr.histogram, fftcorr.fftw,
fftcorr.array, fftcorr.utils); the work runs in libfftcorr_hip.so on the GPU.
"""

import enum

__version__ = "0.1.0"


class DistributionScheme(enum.IntEnum):
    """cc/particle_mesh/distribution_functions.h:11-15 (+ CIC, new)."""
    NEAREST_CELL = 0
    TRIANGULAR_SHAPED_CLOUD = 1
    WAVELET = 2
    CLOUD_IN_CELL = 3


class WindowCorrection(enum.IntEnum):
    """cc/correlate/window_correct.h:8-12 (+ CIC codes, new)."""
    NO_CORRECTION = 0
    TSC_CORRECTION = 1
    TSC_ALIASED_CORRECTION = 2
    CIC_CORRECTION = 3
    CIC_ALIASED_CORRECTION = 4


NEAREST_CELL = DistributionScheme.NEAREST_CELL
TRIANGULAR_SHAPED_CLOUD = DistributionScheme.TRIANGULAR_SHAPED_CLOUD
WAVELET = DistributionScheme.WAVELET
CLOUD_IN_CELL = DistributionScheme.CLOUD_IN_CELL
NO_CORRECTION = WindowCorrection.NO_CORRECTION
TSC_CORRECTION = WindowCorrection.TSC_CORRECTION
TSC_ALIASED_CORRECTION = WindowCorrection.TSC_ALIASED_CORRECTION
CIC_CORRECTION = WindowCorrection.CIC_CORRECTION
CIC_ALIASED_CORRECTION = WindowCorrection.CIC_ALIASED_CORRECTION
