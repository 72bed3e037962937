"""CPU-only checks: the C-ABI library loads and exports every entry point of
include/fftcorr_hip.h; host-side API logic that needs no GPU."""
import os
import re

import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.correlate import compute_xi
from fftcorr_amd.grid import ConfigSpaceGrid
from fftcorr_amd.histogram import HistogramList
from fftcorr_amd.utils import Timer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fftcorr_hip.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fc_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 30
    L = _lib.lib()
    for s in syms:
        assert hasattr(L, s), s
    # ctypes signatures cover exactly the declared entry points.
    assert set(_lib.SIGNATURES) == set(syms)


def test_plan_cache_roundtrip_without_gpu(tmp_path):
    from fftcorr_amd.fftw import export_wisdom_to_file, forget_wisdom, import_wisdom_from_file
    p = tmp_path / "w.txt"
    p.write_text("fftcorr_hip_tiles v1\n512 1 0 16\n256 1 1 8\n")
    import_wisdom_from_file(str(p))
    out = tmp_path / "out.txt"
    export_wisdom_to_file(str(out))
    lines = out.read_text().splitlines()
    assert lines[0] == "fftcorr_hip_tiles v1"
    assert "512 1 0 16" in lines and "256 1 1 8" in lines
    forget_wisdom()
    export_wisdom_to_file(str(out))
    assert out.read_text().splitlines() == ["fftcorr_hip_tiles v1"]
    with pytest.raises(IOError):
        import_wisdom_from_file(str(tmp_path / "missing"))
    bad = tmp_path / "bad.txt"
    bad.write_text("not wisdom\n")
    with pytest.raises(IOError):
        import_wisdom_from_file(str(bad))


def test_fft_radix_plans_without_gpu(tmp_path):
    from fftcorr_amd.autotune import fft_radices
    from fftcorr_amd.fftw import export_wisdom_to_file, forget_wisdom, import_wisdom_from_file
    assert fft_radices(512, 0) == [8, 8, 8]
    assert fft_radices(256, 0) == [8, 8, 4]
    assert fft_radices(512, 1) == [16, 16, 2]
    assert fft_radices(256, 1) == [16, 16]
    assert fft_radices(128, 1) == [16, 8]
    assert fft_radices(64, 1) is None and fft_radices(384, 0) is None
    assert fft_radices(512, 2) is None
    for n in (8, 16, 32, 64, 128, 256, 512):
        for v in (0, 1):
            r = fft_radices(n, v)
            if r is not None:
                assert int(np.prod(r)) == n and r[0] == (8 if v == 0 else 16)
    # Variant choices travel with the plan cache; a variant that does not
    # exist is refused.
    p = tmp_path / "v.txt"
    p.write_text("fftcorr_hip_tiles v1\n512 1 3 1\n256 1 4 1\n")
    import_wisdom_from_file(str(p))
    out = tmp_path / "out.txt"
    export_wisdom_to_file(str(out))
    lines = out.read_text().splitlines()
    assert "512 1 3 1" in lines and "256 1 4 1" in lines
    forget_wisdom()
    bad = tmp_path / "bad.txt"
    bad.write_text("fftcorr_hip_tiles v1\n512 1 3 2\n")
    with pytest.raises(IOError):
        import_wisdom_from_file(str(bad))
    bad.write_text("fftcorr_hip_tiles v1\n512 1 -1 4\n")
    with pytest.raises(IOError):
        import_wisdom_from_file(str(bad))
    forget_wisdom()


def test_grid_construction_and_cell_adoption(capsys):
    # test_grid.py:7-106.
    g = ConfigSpaceGrid([2, 2, 2], [2, -5, 77], [50, 45, 120], cell_size=25,
                        metadata={"hello": 12345})
    assert g.cell_size == 25
    np.testing.assert_array_equal(g.posmax, [52, 45, 127])
    assert g.metadata == {"hello": 12345}
    with pytest.raises(ValueError):
        g.posmin[0] = 3
    g2 = ConfigSpaceGrid([2, 2, 2], [2, -5, 77], [4, 5, 79])
    assert g2.cell_size == 5.0
    for kwargs in [dict(shape=[2, 2], posmin=[2, -5, 77], posmax=[50, 45, 120], cell_size=25),
                   dict(shape=[2, 2, 2], posmin=[2, -5, 77], posmax=[4, 5, 79], cell_size=4),
                   dict(shape=[2, 2, 2], posmin=[2, -5, 77], posmax=[4, 5, 72]),
                   dict(shape=[2, 0, 2], posmin=[0, 0, 0], posmax=[1, 1, 1])]:
        with pytest.raises(ValueError):
            ConfigSpaceGrid(**kwargs)
    # to_grid_coords (test_grid.py:112-119)
    np.testing.assert_allclose(g.to_grid_coords([2, -4, 79]), [0, 0.04, 0.08], rtol=1e-15)
    with pytest.raises(ValueError):
        g.to_grid_coords([1, 0, 77])
    np.testing.assert_allclose(g.to_grid_coords([1, 0, 77], periodic_wrap=True),
                               [1.96, 0.2, 0.0], rtol=1e-15)
    # Host-side views behave like the reference's zero-copy numpy view.
    d = g.data
    d += 7
    assert np.all(g.data == 7)
    g.clear()
    assert np.all(g.data == 0) and g.metadata == {}


def test_extra_pad_matches_reference_arithmetic(capsys):
    g = ConfigSpaceGrid([256] * 3, [500.0] * 3, [1500.0] * 3, extra_pad=200 / 1.5)
    assert g.cell_size == pytest.approx((1000 + 2 * 200 / 1.5) / 256)
    assert g.padding == pytest.approx(200 / 1.5)


def test_histogram_list_host_checks():
    h = HistogramList(2, -10, 10, 2)
    np.testing.assert_array_equal(h.bins, np.arange(-9, 10, 2.0))
    assert h.counts.dtype == np.intc and h.hist_values.dtype == np.float64
    with pytest.raises(ValueError):
        h.hist_values[0][0] = 5
    for args in [(0, 0, 1, 1), (1, 1, 1, 1), (1, 0, 1, 0)]:
        with pytest.raises(ValueError):
            HistogramList(*args)


@pytest.mark.parametrize("hemi", ["north", "south"])
def test_compute_xi_against_reference_goldens(hemi):
    # test_fftcorr_aniso.py:115-134 at rtol 1e-5.
    d = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "corrDD.dat.out"))
    r = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "corrRR.dat.out"))
    ref = np.loadtxt(os.path.join(GOLDEN, hemi, "nonperiodic", "xi.txt")).T
    d = d[d[:, 0] == 0]
    r = r[r[:, 0] == 0]
    np.testing.assert_allclose(d[:, 1], ref[0])
    np.testing.assert_allclose(compute_xi(d[:, 3:], r[:, 3:]), ref[1:].T, rtol=1e-5)


def test_compute_xi_identity_without_boundary():
    # With RR isotropic (f_j = 0 for j > 0), xi = NN / RR_0.
    nn = np.array([[2.0, 0.5, 0.1], [1.0, -0.2, 0.05]])
    rr = np.array([[4.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    np.testing.assert_allclose(compute_xi(nn, rr), nn / rr[:, [0]], rtol=1e-15)


def test_timer():
    t = Timer()
    with t:
        pass
    assert t.elapsed >= 0
    with pytest.raises(ValueError):
        t.stop()
