from fftcorr_amd.utils.timer import Timer
