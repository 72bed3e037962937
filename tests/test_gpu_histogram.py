"""GPU HistogramList (cc/histogram/histogram_list.h) against the oracle:
counts bit-exact, FP64 sums to 1e-12 relative."""
import numpy as np
import pytest

from fftcorr_amd.histogram import HistogramList
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu


def test_known_answers():
    h = HistogramList(2, -10, 10, 2)
    h.accumulate(0, np.array([[1, 7], [3, -5]], float), np.array([[11, 77], [33, -55]], float))
    h.accumulate(1, np.array([1, -7], float), np.array([100, -700], float))
    np.testing.assert_array_equal(h.counts[0], [0, 0, 1, 0, 0, 1, 1, 0, 1, 0])
    np.testing.assert_array_equal(h.hist_values[0], [0, 0, -55, 0, 0, 11, 33, 0, 77, 0])
    np.testing.assert_array_equal(h.counts[1], [0, 1, 0, 0, 0, 1, 0, 0, 0, 0])
    np.testing.assert_array_equal(h.hist_values[1], [0, -700, 0, 0, 0, 100, 0, 0, 0, 0])
    with pytest.raises(ValueError):
        h.hist_values[0][0] = 5
    h.reset()
    assert h.counts.sum() == 0


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 300_000])
def test_random_against_oracle(n):
    rng = np.random.default_rng(n)
    x = np.concatenate([rng.uniform(-5, 210, n), np.arange(0, 200, 5.0)])
    y = rng.normal(size=x.size)
    h = HistogramList(3, 0.0, 200.0, 5.0)
    o = O.HistogramList(3, 0.0, 200.0, 5.0)
    h.accumulate(2, x, y)
    o.accumulate_fast(2, x, y)
    np.testing.assert_array_equal(h.counts, o.counts)
    np.testing.assert_allclose(h.hist_values, o.hist_values, rtol=1e-12, atol=1e-12)


def test_errors():
    h = HistogramList(1, 0, 1, 0.1)
    with pytest.raises(ValueError):
        h.accumulate(0, np.zeros(3), np.zeros(4))
    with pytest.raises(IndexError):
        h.accumulate(1, np.zeros(3), np.zeros(3))
    with pytest.raises(RuntimeError):
        h.accumulate(0, np.zeros(3, np.float32), np.zeros(3, np.float32))
