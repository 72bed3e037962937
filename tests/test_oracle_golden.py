"""Pins the oracle's Correlator path (cc/correlate/correlator.h:360-505) on the
reference's own golden outputs: test_data/{north,south}/nonperiodic/corr{DD,RR}.dat.out
(256^3, l<=2, r<=200 by 10, k<=0.03 by 0.01, D12 wavelet gridding by the
reference's own code in oracle/_ref), compared as test_fftcorr_aniso.py does
(rtol 1e-7, the printed precision)."""
import os

import numpy as np
import pytest

from oracle import fftcorr_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("north", "DD"), ("north", "RR"), ("south", "DD"), ("south", "RR")]


def golden_inputs(ref, hemi, kind):
    f1 = os.path.join(GOLDEN, hemi, "corr%s.dat" % kind)
    f2 = os.path.join(GOLDEN, hemi, "corrRR.dat") if kind == "DD" else None
    return ref.cli_grid(f1, f2, 256, 200.0, 0, O.WAVELET)


def golden_rows(res):
    ps = O.print_hist_rows(res["k"], res["k_counts"], res["k_hist"], 1, True)
    xi = O.print_hist_rows(res["r"], res["r_counts"], res["r_hist"], 0, False)
    return np.vstack([ps, xi])


@pytest.mark.slow
@pytest.mark.parametrize("hemi,kind", CASES)
def test_oracle_reproduces_golden(ref, hemi, kind):
    grid, posmin, cell, st = golden_inputs(ref, hemi, kind)
    c = O.Correlator([256] * 3, cell, posmin, O.NO_CORRECTION, 200.0, 10.0, 0.03, 0.01, 2)
    c.fast_hist = True
    res = c.autocorrelate(grid)
    gold = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "corr%s.dat.out" % kind))
    np.testing.assert_allclose(golden_rows(res), gold, rtol=1e-7)
