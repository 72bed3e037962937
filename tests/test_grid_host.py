"""Host logic of ConfigSpaceGrid and of the CLI's grid set-up, on the CPU.

The device copy is replaced by a numpy stand-in that counts transfers, so the
coherence rule of the live ``.data`` view (fftcorr/grid/config_space_grid.pyx:
the view is the grid's own storage) is checked without a GPU."""
import numpy as np
import pytest

from fftcorr_amd import _lib
from fftcorr_amd.cli import grid_geometry
from fftcorr_amd.grid import ConfigSpaceGrid


class _FakeDeviceGrid:
    def __init__(self, shape, posmin=None, cell_size=None, precision=None, ctx=None):
        self.a = np.zeros(shape)
        self.uploads = 0
        self.downloads = 0

    def upload(self, host):
        self.a[...] = host
        self.uploads += 1

    def download(self, out):
        out[...] = self.a
        self.downloads += 1
        return out

    def stats(self):
        return float(self.a.sum()), float((self.a ** 2).sum())

    def normalize(self, mean):
        self.a = (self.a - mean) / mean

    def add_scalar(self, v):
        self.a += v

    def multiply_by(self, v):
        self.a *= v

    def clear(self):
        self.a[...] = 0


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(_lib, "DeviceGrid", _FakeDeviceGrid)
    monkeypatch.setattr(_lib, "context", lambda device=None: None)


def _grid():
    return ConfigSpaceGrid([4, 4, 4], [0, 0, 0], [4, 4, 4])


def test_kept_view_sees_writes_and_device_results(fake_device):
    g = _grid()
    d = g.data
    assert g.sum() == 0.0
    d[...] = 2.0  # write through the kept view after a device use
    assert g.sum() == 2.0 * 64  # the device copy picked it up
    g.normalize(4.0)  # device work while the view is held ...
    np.testing.assert_array_equal(d, -0.5)  # ... shows in the view at once
    d[0, 0, 0] = 10.0
    assert g.sum() == pytest.approx(-0.5 * 63 + 10.0)
    sub = g.data[1]  # a view of a view keeps the host copy primary too
    del d
    g += 1.0
    np.testing.assert_array_equal(sub, 0.5)


def test_no_kept_view_stays_device_resident(fake_device):
    g = _grid()
    np.copyto(g.data, np.full((4, 4, 4), 3.0))
    g.normalize(3.0)
    dev = g.device_grid()
    assert dev.uploads == 1 and dev.downloads == 0
    g *= 2.0  # scalar ops run on the device while it holds the current copy
    assert dev.downloads == 0
    np.testing.assert_array_equal(g.data, 0.0)
    assert dev.downloads == 1
    g.sum()
    assert dev.uploads == 2  # .data was handed out: the next device use uploads


def test_construction_contract():
    g = ConfigSpaceGrid([2, 3, 4], [1.0, 2.0, 3.0], [3.0, 5.0, 5.0])
    assert g.cell_size == 1.0
    np.testing.assert_array_equal(g.posmax, [3.0, 5.0, 7.0])
    g = ConfigSpaceGrid([2, 2, 2], [0, 0, 0], [1, 1, 1], extra_pad=1.0, padding=0.5)
    np.testing.assert_array_equal(g.posmin, [-1, -1, -1])
    assert g.cell_size == 1.5 and g.padding == 1.5
    with pytest.raises(ValueError):
        g.posmax[0] = 0.0
    for bad in [dict(shape=[2, 2, 2, 2]), dict(shape=[2, -1, 2]), dict(posmin=[0, 0]),
                dict(posmax=[1, 0, 1]), dict(cell_size=0.0), dict(cell_size=0.4),
                dict(padding=-1), dict(extra_pad=-0.5)]:
        kw = dict(shape=[2, 2, 2], posmin=[0, 0, 0], posmax=[1, 1, 1])
        kw.update(bad)
        with pytest.raises(ValueError):
            ConfigSpaceGrid(**kw)


def test_cli_grid_setup_over_random_headers():
    """Every catalogue header the CLI can read gives a grid: the adopted cell
    is used as is (cc/fftcorr.cc:335-346), never re-derived and re-checked."""
    rng = np.random.default_rng(0)
    for _ in range(20000):
        lo = rng.uniform(-3000, 3000, 3)
        hi = lo + rng.uniform(1.0, 4000.0, 3)
        n = [int(v) for v in rng.integers(64, 513, 3)]
        periodic = bool(rng.integers(2))
        sep = float(rng.uniform(10, 300))
        posmin, cell = grid_geometry(lo, hi, n, -123.0, sep, periodic)
        g = ConfigSpaceGrid.from_cell(n, posmin, cell)
        assert g.cell_size == cell
        np.testing.assert_array_equal(g.posmin, posmin)
        assert np.all(g.posmax >= hi + (0.0 if periodic else sep / 1.5) - 1e-9 * np.abs(hi))
        # The covering constructor accepts the same cell despite rounding.
        ConfigSpaceGrid(n, posmin, posmin + np.asarray(n) * cell, cell_size=cell)


def test_covering_constructor_tolerates_one_ulp():
    # A header whose re-derived cell rounds one ulp above the adopted one.
    posmin = np.array([1717.93, -1948.31, -2606.94])
    n = [128, 128, 128]
    cell = 28.992218238440167
    g = ConfigSpaceGrid(n, posmin, posmin + np.asarray(n) * cell, cell_size=cell)
    assert g.cell_size == cell
