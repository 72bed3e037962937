"""Pins the oracle's HistogramList (cc/histogram/histogram_list.h:10-67)."""
import numpy as np
import pytest

from oracle import fftcorr_oracle as O


def test_known_answers_from_reference_smoke_script():
    # Inputs of test_histogram.py:5-36; expected values follow from
    # histogram_list.h:15 (nbins), :24 (midpoints), :57 (floor binning).
    h = O.HistogramList(2, -10, 10, 2)
    np.testing.assert_array_equal(h.bins, np.arange(-9, 10, 2.0))
    h.accumulate(0, np.array([[1, 7], [3, -5]], float), np.array([[11, 77], [33, -55]], float))
    h.accumulate(1, np.array([1, -7], float), np.array([100, -700], float))
    np.testing.assert_array_equal(h.counts[0], [0, 0, 1, 0, 0, 1, 1, 0, 1, 0])
    np.testing.assert_array_equal(h.hist_values[0], [0, 0, -55, 0, 0, 11, 33, 0, 77, 0])
    np.testing.assert_array_equal(h.counts[1], [0, 1, 0, 0, 0, 1, 0, 0, 0, 0])
    np.testing.assert_array_equal(h.hist_values[1], [0, -700, 0, 0, 0, 100, 0, 0, 0, 0])
    h.reset()
    assert h.counts.sum() == 0 and not h.hist_values.any()


def test_edges_dropped_unlike_numpy_histogram():
    h = O.HistogramList(1, 0.0, 10.0, 2.5)
    h.accumulate(0, [10.0, -1e-300, 0.0, 9.999999, 2.5], [1, 1, 1, 1, 1])
    np.testing.assert_array_equal(h.counts[0], [1, 1, 0, 1])  # x=max excluded


def test_bad_arguments():
    for args in [(0, 0, 1, 1), (1, 1, 1, 1), (1, 0, 1, 0), (1, 0, 1, -1)]:
        with pytest.raises(ValueError):
            O.HistogramList(*args)
    with pytest.raises(IndexError):
        O.HistogramList(1, 0, 1, 0.1).accumulate(1, [0.5], [1.0])


def test_bit_exact_against_reference(ref):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-1, 205, 5000), np.arange(0, 205, 5.0),
                        np.sqrt(np.arange(0, 40000, 25.0))])
    y = rng.normal(size=x.size)
    bins, counts, hist = ref.histogram(3, 0.0, 200.0, 5.0, 1, x, y)
    h = O.HistogramList(3, 0.0, 200.0, 5.0)
    h.accumulate(1, x, y)
    np.testing.assert_array_equal(h.bins, bins)
    np.testing.assert_array_equal(h.counts, counts)
    np.testing.assert_array_equal(h.hist_values, hist)
