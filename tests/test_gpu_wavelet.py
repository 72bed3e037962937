"""D12 wavelet mass assignment on the GPU (distribution_functions.h:162-217,
sample table d12.h) with the table registered by the caller.

The grid is checked against the oracle (the table-driven stencil, itself
pinned against the reference's own code) at 1e-13 relative; the full CLI at
the configuration of the reference's golden outputs (256^3, l <= 2, r <= 200
by 10, k <= 0.03 by 0.01, D12; DD - RR and RR on the north and south
catalogues) reproduces test_data/{north,south}/nonperiodic/corr{DD,RR}.dat.out
at the printed precision."""
import ctypes
import os

import numpy as np
import pytest

from fftcorr_amd import _lib, cli
from fftcorr_amd.grid import ConfigSpaceGrid
from fftcorr_amd.particle_mesh import MassAssignor, set_wavelet_table
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture
def table(ref):
    tab = ref.wavelet_table()
    set_wavelet_table(tab)
    yield tab
    set_wavelet_table(None)


def _grid(shape, posmin, cell, precision="fp64"):
    posmin = np.asarray(posmin, float)
    return ConfigSpaceGrid(shape, posmin, posmin + cell * np.array(shape), cell_size=cell,
                           precision=precision)


def _rows(text):
    return np.array([[float(v) for v in l.split()] for l in text.splitlines()
                     if l and not l.startswith("#")])


@pytest.mark.parametrize("shape", [(8, 8, 8), (9, 12, 10), (16, 16, 16), (40, 24, 70)])
@pytest.mark.parametrize("wrap", [False, True])
def test_d12_grid_against_oracle(table, shape, wrap):
    rng = np.random.default_rng(sum(shape) * 5 + wrap)
    n = 5000
    posmin = np.array([-10.0, 3.0, 100.0])
    cell = 7.0
    margin = 2 * cell if wrap else 0.0
    pos = rng.uniform(posmin - margin, posmin + cell * np.array(shape) + margin, size=(n, 3))
    pos[: n // 4] = pos[0] + rng.normal(0, 0.3 * cell, size=(n // 4, 3))
    w = rng.uniform(-1.0, 1.5, n)
    g = _grid(shape, posmin, cell)
    ma = MassAssignor(g, 2, periodic_wrap=wrap, buffer_size=1500)
    ma.add_particles(pos, w)
    x, ok = O.grid_coords(pos, posmin, cell, shape, wrap)
    want = O.assign_mass(np.zeros(shape), x[ok], w[ok], O.WAVELET, table)
    scale = np.abs(want).max()
    np.testing.assert_allclose(g.data, want, rtol=1e-13, atol=1e-13 * scale)
    assert ma.num_added == ok.sum() and ma.num_skipped == (~ok).sum()
    g32 = _grid(shape, posmin, cell, "fp32")
    MassAssignor(g32, 2, periodic_wrap=wrap).add_particles(pos, w)
    np.testing.assert_allclose(g32.data, want, rtol=1e-6, atol=1e-6 * scale)


def test_d12_needs_table_and_eight_cells(table):
    with pytest.raises(ValueError, match="8 cells"):
        MassAssignor(_grid((8, 7, 8), [0, 0, 0], 1.0), 2).add_particles(
            np.array([[1.0, 1.0, 1.0]]))
    bad = np.zeros(100)
    rc = _lib.lib().fc_set_wavelet_table(bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          100)
    assert rc != 0 and "512" in _lib.lib().fc_last_error().decode()
    set_wavelet_table(None)
    with pytest.raises(ValueError, match="sample table"):
        MassAssignor(_grid((8, 8, 8), [0, 0, 0], 1.0), 2)
    with pytest.raises(SystemExit, match="not available on the GPU"):
        cli.run(["-r", "50", "-d", "2", "-in", os.path.join(GOLDEN, "north", "corrDD.dat")])


@pytest.mark.parametrize("hemi,kind", [("north", "DD"), ("north", "RR"), ("south", "DD"),
                                       ("south", "RR")])
def test_cli_d12_reproduces_reference_golden(table, tmp_path, hemi, kind):
    # DD runs take the randoms as the second file (D - R), RR runs the
    # randoms alone, as the golden outputs were made.
    f1 = os.path.join(GOLDEN, hemi, "corr%s.dat" % kind)
    extra = ["-in2", os.path.join(GOLDEN, hemi, "corrRR.dat")] if kind == "DD" else []
    text = cli.run(["-n", "256", "-ell", "2", "-r", "200", "-dr", "10", "-in", f1] + extra +
                   ["-wisdom", str(tmp_path / "w.txt")])
    assert "# Using D12 wavelet" in text
    golden = os.path.join(GOLDEN, hemi, "nonperiodic", "corr%s.dat.out" % kind)
    gold = np.loadtxt(golden)
    got = _rows(text)
    assert got.shape == gold.shape
    # Rows are printed with 10 significant digits; the density grids differ
    # from the reference's by summation order only.
    np.testing.assert_allclose(got, gold, rtol=1e-7, atol=1e-9 * np.abs(gold).max())
    gold_text = open(golden).read()
    for key in ("# Sum of squares of density", "# Sum of squares of weights"):
        line = [l for l in text.splitlines() if l.startswith(key)]
        gline = [l for l in gold_text.splitlines() if l.startswith(key)]
        assert line == gline
