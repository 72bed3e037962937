"""Plan-cache ("wisdom") functions under the reference's names
(fftcorr/fftw/fftw.pyx:1-14) and its FFTW flag constants
(fftcorr/fftw/fftw.pxd:8-17).

The cache holds the FFT tile shapes chosen for each transform length and
precision (see fftcorr_amd.autotune); exporting writes them to a text file,
importing makes later plans use them, forgetting drops them.  The flag
constants are accepted by the correlators and have no effect on the GPU
plans.
"""

from fftcorr_amd import _lib

FFTW_MEASURE = 0
FFTW_DESTROY_INPUT = 1
FFTW_UNALIGNED = 2
FFTW_CONSERVE_MEMORY = 4
FFTW_EXHAUSTIVE = 8
FFTW_PRESERVE_INPUT = 16
FFTW_PATIENT = 32
FFTW_ESTIMATE = 64
FFTW_WISDOM_ONLY = 2097152


def export_wisdom_to_file(filename):
    if _lib.lib().fc_plan_cache_export(str(filename).encode("utf-8")) != _lib.FC_OK:
        # Message text as the reference writes it (fftw.pyx:5).
        raise IOError("Failed to import wisdom file")


def import_wisdom_from_file(filename):
    if _lib.lib().fc_plan_cache_import(str(filename).encode("utf-8")) != _lib.FC_OK:
        raise IOError("Failed to export wisdom file")


def forget_wisdom():
    _lib.check(_lib.lib().fc_plan_cache_forget())
