"""ConfigSpaceGrid: the density grid of the correlation path.

Interface of fftcorr/grid/config_space_grid.pyx:11-207 (Python) over
cc/grid/config_space_grid.h:10-56 (C++): a regular [n0, n1, n2] grid of cubic
cells of side ``cell_size`` whose corner sits at ``posmin``.

Geometry contract (what the reference's constructor guarantees):
  * shape: 3 positive ints; posmin < posmax elementwise; cell_size > 0 when
    given; padding and extra_pad >= 0.  A violation raises ValueError.
  * extra_pad widens the box by that much on every side before adoption.
  * Adoption: the cell is the smallest that covers the box on every axis,
    max_i (posmax_i - posmin_i) / shape_i, or the requested cell_size if it is
    at least that; posmax becomes posmin + shape * cell.
  * posmin and posmax are exposed read-only; padding reports
    padding + extra_pad.

Storage: ``.data`` is a writable float64 host array, as in the reference,
and the grid also keeps a device copy in HBM that mass assignment and the
correlators work on.  Coherence rule, so that ``.data`` behaves as a live
view of the grid:
  * device work marks the host copy stale; reading ``.data`` downloads it;
  * handing out ``.data`` marks the device copy stale, and the next device use
    uploads it;
  * while any reference to the host array (or a view of it) is held outside
    this object, the host copy is primary: every device use uploads first, and
    every device modification is downloaded at once, so a kept view both sees
    device results and has its writes picked up.  Whether such a reference
    exists is read from the array's reference count.
"""

import logging
import math
import sys

import numpy as np

from fftcorr_amd import _lib

_log = logging.getLogger("fftcorr_amd.grid")

# Slack of the coverage check, in units of the box coordinates' ulp: a cell
# computed elsewhere as max((posmax - posmin) / shape) (the CLI's adoption)
# may fall short of this module's by the rounding of that difference.
_COVER_ULPS = 8 * np.finfo(np.float64).eps


def _triple(name, value, dtype):
    arr = np.array(value, dtype=dtype)
    if arr.shape != (3,):
        raise ValueError("%s must hold 3 values, got shape %s" % (name, arr.shape))
    return arr


class ConfigSpaceGrid:
    def __init__(self, shape, posmin, posmax, cell_size=None, padding=0, extra_pad=0,
                 metadata=None, precision="fp64", device=None):
        n = _triple("shape", shape, np.intc)
        if (n <= 0).any():
            raise ValueError("shape entries must be positive, got %s" % (n.tolist(),))
        lo = _triple("posmin", posmin, np.float64)
        hi = _triple("posmax", posmax, np.float64)
        if not (lo < hi).all():
            raise ValueError("need posmin < posmax on every axis, got posmin=%s posmax=%s"
                             % (lo.tolist(), hi.tolist()))
        if padding < 0 or extra_pad < 0:
            raise ValueError("padding and extra_pad must be non-negative, got %r and %r"
                             % (padding, extra_pad))
        if cell_size is not None and not cell_size > 0:
            raise ValueError("cell_size must be positive, got %r" % (cell_size,))
        lo = lo - extra_pad
        hi = hi + extra_pad
        cover = float(np.max((hi - lo) / n))
        if cell_size is None:
            cell = cover
        else:
            cell = float(cell_size)
            reach = lo + n * cell
            slack = _COVER_ULPS * np.maximum(np.maximum(np.abs(lo), np.abs(hi)), n * cell)
            if (reach < hi - slack).any():
                raise ValueError("cell_size %.6g does not cover the box: it needs at least %.6g"
                                 % (cell, cover))
        self._setup(n, lo, cell, padding + extra_pad, metadata, precision, device)

    @classmethod
    def from_cell(cls, shape, posmin, cell_size, padding=0, metadata=None, precision="fp64",
                  device=None):
        """A grid of ``shape`` cells of side ``cell_size`` from ``posmin``,
        taken as given (no coverage check): the construction of the
        reference's CLI, which builds its grid from (ngrid, posmin, cell)
        (cc/fftcorr.cc:335-346, cc/grid/config_space_grid.h:12-16)."""
        n = _triple("shape", shape, np.intc)
        if (n <= 0).any():
            raise ValueError("shape entries must be positive, got %s" % (n.tolist(),))
        if not cell_size > 0:
            raise ValueError("cell_size must be positive, got %r" % (cell_size,))
        self = cls.__new__(cls)
        self._setup(n, _triple("posmin", posmin, np.float64), float(cell_size), padding,
                    metadata, precision, device)
        return self

    def _setup(self, n, lo, cell, padding, metadata, precision, device):
        hi = lo + n * cell
        lo.setflags(write=False)
        hi.setflags(write=False)
        self._shape = tuple(int(s) for s in n)
        self._posmin = lo
        self._posmax = hi
        self._cell_size = cell
        self._padding = padding
        self._metadata = dict(metadata) if metadata else {}
        self._precision = _lib.grid_precision_code(precision)
        self._device = device
        self._host = np.zeros(self._shape, dtype=np.float64)
        self._host_refs = sys.getrefcount(self._host)
        self._dev = None
        self._host_valid = True
        self._dev_valid = False
        _log.info("grid %s, cell %.6g, posmin %s, posmax %s", self._shape, cell, lo.tolist(),
                  hi.tolist())

    # ---- host / device coherence -------------------------------------------
    def _view_held(self):
        """True while something outside this object references the host
        array (directly or through a numpy view, whose base it is)."""
        return sys.getrefcount(self._host) > self._host_refs

    def device_grid(self):
        """The device copy (an _lib.DeviceGrid), uploaded if stale."""
        if self._dev is None:
            self._dev = _lib.DeviceGrid(self._shape, self._posmin, self._cell_size,
                                        self._precision, _lib.context(self._device))
            self._dev_valid = False
        if self._view_held():
            self._sync_host()
            self._dev_valid = False
        if not self._dev_valid:
            self._dev.upload(self._host)
            self._dev_valid = True
        return self._dev

    def _device_modified(self):
        """Called after device work changed the grid."""
        self._dev_valid = True
        self._host_valid = False
        if self._view_held():
            self._sync_host()

    def _sync_host(self):
        if not self._host_valid:
            self._dev.download(self._host)
            self._host_valid = True

    # ---- reference API -----------------------------------------------------
    def to_grid_coords(self, survey_coords, periodic_wrap=False):
        """Survey coordinates of one point to grid units, (p - posmin) / cell
        per axis (cc/grid/config_space_grid.h:31-46).  Out of range: wrapped
        into [0, n) when periodic_wrap, else ValueError."""
        p = np.asarray(survey_coords, dtype=np.float64)
        out = (p - self._posmin) / self._cell_size
        for i, ni in enumerate(self._shape):
            if 0.0 <= out[i] < ni:
                continue
            if not periodic_wrap:
                raise ValueError("point %s lies outside the grid" % (p.tolist(),))
            w = math.fmod(out[i], ni)
            if w < 0:
                w += ni
            out[i] = 0.0 if w >= ni else w  # -tiny + n rounds to n: that is cell 0
        return out

    @property
    def shape(self):
        return self._shape

    @property
    def posmin(self):
        return self._posmin

    @property
    def posmax(self):
        return self._posmax

    @property
    def cell_size(self):
        return self._cell_size

    @property
    def padding(self):
        return self._padding

    @property
    def data(self):
        self._sync_host()
        self._dev_valid = False  # the caller may write through the view
        return self._host

    @property
    def metadata(self):
        return self._metadata

    @property
    def precision(self):
        return "fp64" if self._precision == _lib.FC_FP64 else "fp32"

    def clear(self):
        self._host.fill(0.0)
        self._host_valid = True
        if self._dev is not None:
            self._dev.clear()
            self._dev_valid = True
        self._metadata.clear()

    def sum(self):
        """FP64 sum of the grid on the device (array_ops::sum)."""
        return self.device_grid().stats()[0]

    def sumsq(self):
        """FP64 sum of squares on the device (array_ops::sumsq)."""
        return self.device_grid().stats()[1]

    def normalize(self, mean=None):
        """delta = rho / mean - 1 in FP64 on the device (cc/fftcorr.cc:369-375).
        mean defaults to the grid's mean."""
        dev = self.device_grid()
        if mean is None:
            mean = dev.stats()[0] / float(np.prod(self._shape))
        dev.normalize(mean)
        self._device_modified()
        return mean

    # ---- asdf files (the reference's on-disk format) --------------------------
    def _header(self):
        return {"shape": list(self._shape), "posmin": self._posmin.tolist(),
                "posmax": self._posmax.tolist(), "cell_size": self._cell_size,
                "padding": self._padding}

    def write(self, filename, dtype=np.float32):
        """Writes header, data (as dtype) and metadata to an asdf file; needs
        the asdf package."""
        import asdf
        tree = {"header": self._header(), "data": np.asarray(self.data, dtype=dtype),
                "metadata": dict(self._metadata)}
        asdf.AsdfFile(tree).write_to(filename)

    @classmethod
    def read(cls, filename, precision="fp64", device=None):
        """Reads a grid written by write() (or by the reference)."""
        import asdf
        with asdf.open(filename) as f:
            h = f.tree["header"]
            grid = cls.from_cell(h["shape"], h["posmin"], float(h["cell_size"]),
                                 padding=h.get("padding", 0), metadata=f.tree.get("metadata"),
                                 precision=precision, device=device)
            if not np.allclose(grid.posmax, h["posmax"]):
                raise ValueError("inconsistent header in %s: posmax %s, expected %s"
                                 % (filename, h["posmax"], grid.posmax.tolist()))
            np.copyto(grid.data, np.asarray(f.tree["data"]))
        return grid

    # ---- in-place arithmetic ---------------------------------------------------
    def _inplace(self, other, op, dev_op):
        device_current = self._dev is not None and self._dev_valid and not self._host_valid
        if np.isscalar(other) and dev_op is not None and device_current \
                and not self._view_held():
            dev_op(float(other))
            self._device_modified()
        else:
            op(self.data, other, out=self.data)
        return self

    def __iadd__(self, other):
        return self._inplace(other, np.add, lambda v: self._dev.add_scalar(v))

    def __isub__(self, other):
        return self._inplace(other, np.subtract, lambda v: self._dev.add_scalar(-v))

    def __imul__(self, other):
        return self._inplace(other, np.multiply, lambda v: self._dev.multiply_by(v))

    def __itruediv__(self, other):
        return self._inplace(other, np.divide, None)

    def __array__(self, dtype=None, copy=None):
        d = self.data
        return d if dtype is None else d.astype(dtype)
