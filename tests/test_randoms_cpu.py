"""Counter-based random stream (numpy restatement) against the Philox4x32-10
known answers, and the displacement-field oracle (fftcorr/grid/utils.pyx)
against the definition."""
import numpy as np
import pytest

from oracle import fftcorr_oracle as O
from philox_ref import counters, keys, philox4x32_10, uniform_positions

# Known-answer vectors of Philox4x32-10 published with the algorithm
# (Random123 kat_vectors): counter, key -> output.
KAT = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
       ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
        (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
       ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = philox4x32_10(np.array([ctr], np.uint64), np.array([key], np.uint64))[0]
    assert [int(v) for v in got] == list(want)


def test_uniform_positions_range_and_moments():
    p = uniform_positions(200000, [1, -2, 3], [5, 2, 11], seed=42)
    assert p.min(axis=0).tolist() >= [1, -2, 3] and np.all(p < [5, 2, 11])
    np.testing.assert_allclose(p.mean(axis=0), [3, 0, 7], atol=0.02)
    np.testing.assert_allclose(p.var(axis=0), np.array([16, 16, 64]) / 12, rtol=0.02)
    # batching does not change the stream
    np.testing.assert_array_equal(uniform_positions(100, [0] * 3, [1] * 3, 7, first=50),
                                  uniform_positions(150, [0] * 3, [1] * 3, 7)[50:])


def test_counter_layout():
    c = counters(2 ** 32 + 5, 2, 1, 9)
    assert c[0].tolist() == [5, 1, 1, 9] and c[1].tolist() == [6, 1, 1, 9]
    k = keys(2 ** 33 + 3, 1)
    assert k[0].tolist() == [3, 2]


def test_displacement_oracle():
    shape = (4, 5, 6)
    posmin = np.array([2.0, -5.0, 77.0])
    cell = 2.0
    rng = np.random.default_rng(0)
    disp = rng.uniform(size=shape + (3,))
    pos = np.array([[2.0, -5.0, 77.0], [9.9, 4.9, 88.9], [5.5, -1.2, 80.1]])
    out = O.apply_displacement_field(pos, disp, posmin, cell, shape, False)
    for p, o in zip(pos, out):
        ic = np.floor((p - posmin) / cell).astype(int)
        np.testing.assert_array_equal(o, p + disp[tuple(ic)])
    with pytest.raises(ValueError):
        O.apply_displacement_field(np.array([[1.0, 0.0, 77.0]]), disp, posmin, cell, shape, False)
    w = O.apply_displacement_field(np.array([[1.0, 0.0, 77.0]]), disp, posmin, cell, shape, True)
    np.testing.assert_array_equal(w[0], [1.0, 0.0, 77.0] + disp[3, 2, 0])
