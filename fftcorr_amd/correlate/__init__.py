from fftcorr_amd.correlate.correlator import Correlator, PeriodicCorrelator
from fftcorr_amd.correlate.utils import compute_xi
from fftcorr_amd.correlate.batch import PeriodicBatch
