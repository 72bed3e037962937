"""The CLI (fftcorr_amd.cli) on the reference's north catalogues with TSC:
its printed rows equal the oracle's -- grid built by the reference's own CLI
gridding code (oracle/_ref), FP64 Correlator restated in numpy -- and a
periodic CIC run equals the Python API path."""
import os

import numpy as np
import pytest

from fftcorr_amd import cli
from oracle import fftcorr_oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rows(text):
    return np.array([[float(v) for v in l.split()] for l in text.splitlines()
                     if l and not l.startswith("#")])


def test_cli_matches_oracle_on_reference_catalogues(ref, tmp_path):
    f1 = os.path.join(GOLDEN, "north", "corrDD.dat")
    f2 = os.path.join(GOLDEN, "north", "corrRR.dat")
    out = tmp_path / "out.txt"
    cli.run(["-n", "64", "-ell", "2", "-r", "100", "-dr", "10", "-d", "1", "-in", f1,
             "-in2", f2, "-out", str(out), "-wisdom", str(tmp_path / "w.txt")])
    got = _rows(out.read_text())
    grid, posmin, cell, st = ref.cli_grid(f1, f2, 64, 100.0, False, 1)
    c = O.Correlator([64] * 3, cell, posmin, 2, 100.0, 10.0, 0.03, 0.01, 2)
    c.fast_hist = True
    res = c.autocorrelate(grid)
    want = np.vstack([O.print_hist_rows(res["k"], res["k_counts"], res["k_hist"], 1, True),
                      O.print_hist_rows(res["r"], res["r_counts"], res["r_hist"], 0, False)])
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())
    text = out.read_text()
    assert "# Found %d particles" % st["num_added"] in text
    assert os.path.exists(tmp_path / "w.txt")


def test_cli_periodic_cic(tmp_path):
    from fftcorr_amd.catalog import uniform_catalog, write_catalog
    from fftcorr_amd.correlate import PeriodicCorrelator
    from fftcorr_amd.grid import ConfigSpaceGrid
    from fftcorr_amd.particle_mesh import MassAssignor
    pos = uniform_catalog(20000, 0.0, 200.0, 3)
    rec = np.column_stack([pos, np.ones(len(pos))])
    p = tmp_path / "cat.dat"
    write_catalog(p, rec, [0, 0, 0], [200, 200, 200])
    text = cli.run(["-n", "32", "-ell", "4", "-r", "40", "-dr", "5", "-k", "0.2", "-dk", "0.02",
                    "-p", "-d", "3", "-in", str(p), "-wisdom", str(tmp_path / "w.txt")])
    got = _rows(text)
    g = ConfigSpaceGrid([32] * 3, [0, 0, 0], [200] * 3)
    ma = MassAssignor(g, 3, True)
    ma.add_particles(rec)
    g.normalize(ma.totw / 32 ** 3)
    c = PeriodicCorrelator([32] * 3, g.cell_size, 4, 40.0, 5.0, 0.2, 0.02, 4)
    c.autocorrelate(g)
    r = c.results()
    want = np.vstack([O.print_hist_rows(r["k"], r["k_counts"], r["k_hist"], 1, True),
                      O.print_hist_rows(r["r"], r["r_counts"], r["r_hist"], 0, True)])
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-12)
