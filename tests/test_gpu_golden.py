"""The GPU Correlator (FP64) reproduces the reference's own golden outputs
test_data/{north,south}/nonperiodic/corr{DD,RR}.dat.out at rtol 1e-7 (as
test_fftcorr_aniso.py:52-67 compares them).  The D12-wavelet density grids
are built by the reference's own code (oracle/_ref); the correlation
(256^3, l<=2, 13 FFTs) runs on the GPU.  FP32-mixed is checked at 2e-5
(floor metric)."""
import os

import numpy as np
import pytest

from fftcorr_amd.correlate import Correlator
from gpu_helpers import floor_rel_err
from test_oracle_golden import CASES, GOLDEN, golden_inputs, golden_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hemi,kind", CASES)
def test_gpu_reproduces_golden(ref, hemi, kind):
    grid, posmin, cell, st = golden_inputs(ref, hemi, kind)
    gold = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "corr%s.dat.out" % kind))
    c = Correlator([256] * 3, cell, posmin, 0, 200.0, 10.0, 0.03, 0.01, 2)
    c.autocorrelate(grid)
    rows = golden_rows(c.results())
    np.testing.assert_allclose(rows, gold, rtol=1e-7)
    c32 = Correlator([256] * 3, cell, posmin, 0, 200.0, 10.0, 0.03, 0.01, 2, precision="fp32")
    c32.autocorrelate(grid)
    r32 = c32.results()
    r64 = c.results()
    assert floor_rel_err(r32["r_hist"], r64["r_hist"]).max() < 2e-5
    assert floor_rel_err(r32["k_hist"], r64["k_hist"]).max() < 2e-5
