"""ConfigSpaceGrid (fftcorr/grid/config_space_grid.pyx:11-207,
cc/grid/config_space_grid.h:10-56).

The grid keeps the reference's host view: ``.data`` is a writable float64
[n0, n1, n2] numpy array.  Alongside it lives a device copy in HBM that mass
assignment and the correlators use directly.  Which copy is current is
tracked: device work (mass assignment, in-place scalar ops) marks the host
copy stale and ``.data`` downloads it; handing out ``.data`` marks the device
copy stale (the caller may write through the view) and the next device use
uploads it.
"""

import math

import numpy as np

from fftcorr_amd import _lib


class ConfigSpaceGrid:
    def __init__(self, shape, posmin, posmax, cell_size=None, padding=0, extra_pad=0,
                 metadata=None, precision="fp64", device=None):
        shape = np.ascontiguousarray(shape, dtype=np.intc)
        posmin = np.asarray(posmin, dtype=np.double).copy(order="C")
        posmax = np.asarray(posmax, dtype=np.double).copy(order="C")
        # Checks and messages of config_space_grid.pyx:27-62.
        if shape.shape != (3,):
            raise ValueError("Expected shape to have shape (3,), got: {}".format(shape.shape))
        if np.any(shape <= 0):
            raise ValueError("Expected shape to be positive, got: {}".format(shape))
        if posmin.shape != (3,):
            raise ValueError("Expected posmin to have shape (3,), got: {}".format(posmin.shape))
        if posmax.shape != (3,):
            raise ValueError("Expected posmax to have shape (3,), got: {}".format(posmax.shape))
        if np.any(posmax <= posmin):
            raise ValueError("Expected posmin < posmax, got posmin={}, posmax={}".format(
                posmin, posmax))
        if cell_size is not None and cell_size <= 0:
            raise ValueError("Expected cell_size to be positive, got: {}".format(cell_size))
        if padding < 0:
            raise ValueError("Expected padding > 0, got {}".format(padding))
        if extra_pad < 0:
            raise ValueError("Expected extra_pad > 0, got {}".format(extra_pad))
        if extra_pad:
            print("Padding boundaries by {:.6g}".format(extra_pad))
            posmin -= extra_pad
            posmax += extra_pad
        # Cell-size adoption, config_space_grid.pyx:68-78.
        minimum_cell_size = np.amax((posmax - posmin) / shape)
        if cell_size is not None:
            print("Requested cell size {:.6g}".format(cell_size))
            if cell_size < minimum_cell_size:
                raise ValueError("Requested cell size too small to cover entire region. "
                                 "Need at least {:.6g}".format(minimum_cell_size))
        else:
            cell_size = minimum_cell_size
        posmax = posmin + shape * cell_size
        print("Adopting:")
        print("  shape = [" + ", ".join(["{}".format(x) for x in shape]) + "]")
        print("  cell_size = {:.6g}".format(cell_size))
        print("  posmin = [" + ", ".join(["{:.6g}".format(x) for x in posmin]) + "]")
        print("  posmax = [" + ", ".join(["{:.6g}".format(x) for x in posmax]) + "]")
        for arr in (posmin, posmax):
            arr.setflags(write=False)
        self._shape = tuple(int(s) for s in shape)
        self._posmin = posmin
        self._posmax = posmax
        self._cell_size = float(cell_size)
        self._padding = padding + extra_pad
        self._metadata = {}
        if metadata:
            self._metadata.update(metadata)
        self._precision = _lib.grid_precision_code(precision)
        self._device = device
        self._host = np.zeros(self._shape, dtype=np.float64)
        self._dev = None
        self._host_valid = True
        self._dev_valid = False

    # ---- host / device coherence -------------------------------------------
    def device_grid(self):
        """The device copy (an _lib.DeviceGrid), uploaded if stale."""
        if self._dev is None:
            self._dev = _lib.DeviceGrid(self._shape, self._posmin, self._cell_size,
                                        self._precision, _lib.context(self._device))
            self._dev_valid = False
        if not self._dev_valid:
            self._dev.upload(self._host)
            self._dev_valid = True
        return self._dev

    def _device_modified(self):
        self._dev_valid = True
        self._host_valid = False

    def _sync_host(self):
        if not self._host_valid:
            self._dev.download(self._host)
            self._host_valid = True

    # ---- reference API -----------------------------------------------------
    def to_grid_coords(self, survey_coords, periodic_wrap=False):
        """cc/grid/config_space_grid.h:31-46 for one point."""
        survey_coords = np.asarray(survey_coords, dtype=np.float64)
        out = np.empty_like(survey_coords)
        for i in range(3):
            x = (survey_coords[i] - self._posmin[i]) / self._cell_size
            xmax = float(self._shape[i])
            if x < 0 or x >= xmax:
                if not periodic_wrap:
                    raise ValueError(f"Coordinates {survey_coords} out of bounds")
                x = math.fmod(x, xmax)
                if x < 0:
                    x += xmax
                if x >= xmax:
                    x = 0.0
            out[i] = x
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
        self._host[...] = 0.0
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

    def write(self, filename, dtype=np.float32):
        import asdf  # the reference's format; not bundled here
        tree = {"header": {"shape": self.shape, "posmin": self.posmin, "posmax": self.posmax,
                           "cell_size": self.cell_size, "padding": self.padding},
                "data": self.data.astype(dtype), "metadata": self.metadata}
        with asdf.AsdfFile(tree) as af:
            af.write_to(filename)

    @classmethod
    def read(cls, filename):
        import asdf
        with asdf.open(filename) as af:
            header = af.tree["header"]
            grid = cls(shape=header["shape"], posmin=header["posmin"], posmax=header["posmax"],
                       padding=header["padding"], metadata=af.tree.get("metadata"))
            assert np.allclose(grid.cell_size, header["cell_size"])
            np.copyto(grid.data, af.tree["data"])
        return grid

    def _inplace(self, other, op, dev_op):
        if np.isscalar(other) and self._dev_valid and not self._host_valid and dev_op:
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
