"""Output tables.  The reference returns astropy Tables
(fftcorr/correlate/correlator.pyx:42-70); astropy is used when importable,
otherwise this column-compatible stand-in: t["r"], t.colnames, len(t),
iteration over column names, t.as_array()."""

import numpy as np

try:  # pragma: no cover - astropy is not in this image
    import astropy.table as _astropy_table
except ImportError:  # pragma: no cover
    _astropy_table = None


class Table:
    def __init__(self, data, names, copy=True):
        if len(data) != len(names):
            raise ValueError("data and names differ in length")
        self._cols = {}
        n = None
        for name, col in zip(names, data):
            col = np.array(col, copy=True) if copy else np.asarray(col)
            if n is None:
                n = len(col)
            elif len(col) != n:
                raise ValueError("Inconsistent data column lengths")
            self._cols[name] = col
        self._len = n or 0

    @property
    def colnames(self):
        return list(self._cols)

    def __getitem__(self, key):
        return self._cols[key]

    def __len__(self):
        return self._len

    def __iter__(self):
        return iter(self._cols)

    def keys(self):
        return self.colnames

    def as_array(self):
        dt = [(k, v.dtype, v.shape[1:]) for k, v in self._cols.items()]
        out = np.empty(self._len, dtype=dt)
        for k, v in self._cols.items():
            out[k] = v
        return out

    def __repr__(self):
        return "<Table length=%d cols=%s>" % (self._len, self.colnames)


def make_table(data, names):
    if _astropy_table is not None:  # pragma: no cover
        return _astropy_table.Table(data=data, names=names, copy=True)
    return Table(data, names, copy=True)
