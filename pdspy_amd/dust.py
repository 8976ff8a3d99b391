"""Dust opacities per model draw, as pdspy's fits make them.

``DustGenerator`` restates pdspy's ``dust.DustGenerator`` call
(dust/DustGenerator.py:71-97, new-format tables with a ``p`` axis): opacity
tables ``kabs``, ``ksca`` [n_p][n_amax][n_lam] over a grid of the size
distribution's slope ``p`` and maximum grain size ``amax`` (cm) are interpolated
linearly in (p, amax) -- in amax itself, not its logarithm -- in log10 of the
opacity at the tables' own wavelengths (scipy's RegularGridInterpolator over
(p, amax, lam) queried at each tabulated lam), and returned as a ``Dust`` with
``set_properties`` (dust/Dust.py:181-188).  run_disk_model.py:79-85 calls it for
every walker; ``DeviceModel.set_dust`` / ``set_opacity`` take the result to the
device.  pdspy keeps the tables in HDF5 files read with h5py
(DustGenerator.py:99-125); here they come as arrays (``from_npz`` reads an .npz
with the same dataset names).  The old format without a ``p`` axis
(interp2d, DustGenerator.py:72-82) is not supported.
"""
import numpy as np

from .inputs import Dust  # the fields pdspy's Dust.set_properties fills (dust/Dust.py:181-188)


def _bracket(grid, x, name):
    if not (grid[0] <= x <= grid[-1]):
        raise ValueError("DustGenerator: %s = %g outside the table [%g, %g]" % (name, x, grid[0], grid[-1]))
    i = int(np.clip(np.searchsorted(grid, x, side="right") - 1, 0, len(grid) - 2))
    return i, (x - grid[i]) / (grid[i + 1] - grid[i])


def _bracket_many(grid, x, name):
    bad = ~((grid[0] <= x) & (x <= grid[-1]))
    if bad.any():
        v = x[bad][0]
        raise ValueError("DustGenerator: %s = %g outside the table [%g, %g]" % (name, v, grid[0], grid[-1]))
    i = np.clip(np.searchsorted(grid, x, side="right") - 1, 0, len(grid) - 2)
    return i, (x - grid[i]) / (grid[i + 1] - grid[i])


class DustGenerator:
    def __init__(self, lam, amax, p, kabs, ksca):
        self.lam = np.asarray(lam, dtype=float)
        self.amax = np.asarray(amax, dtype=float)
        self.p = np.asarray(p, dtype=float)
        shape = (self.p.size, self.amax.size, self.lam.size)
        self.kabs = np.asarray(kabs, dtype=float).reshape(shape)
        self.ksca = np.asarray(ksca, dtype=float).reshape(shape)
        for name, ax in (("p", self.p), ("amax", self.amax)):
            if ax.size < 2 or np.any(np.diff(ax) <= 0):
                raise ValueError("DustGenerator: the %s axis must hold >= 2 increasing values" % name)
        if np.any(self.kabs <= 0) or np.any(self.ksca <= 0):
            raise ValueError("DustGenerator: opacities must be positive (they are interpolated in log10)")
        self.kext = self.kabs + self.ksca
        self.albedo = self.ksca / self.kext
        self._lk = (np.log10(self.kabs), np.log10(self.ksca))

    @classmethod
    def from_npz(cls, path):
        f = np.load(path)
        return cls(f["lam"], f["amax"], f["p"], f["kabs"], f["ksca"])

    def __call__(self, amax, p=None):
        if p is None:
            raise ValueError("DustGenerator: tables with a p axis need p")
        i, tp = _bracket(self.p, float(p), "p")
        j, ta = _bracket(self.amax, float(amax), "amax")
        out = []
        for L in self._lk:
            v = ((1 - tp) * (1 - ta) * L[i, j] + (1 - tp) * ta * L[i, j + 1] +
                 tp * (1 - ta) * L[i + 1, j] + tp * ta * L[i + 1, j + 1])
            out.append(10.0 ** v)
        d = Dust()
        d.set_properties(self.lam, out[0], out[1])
        return d

    def draw_many(self, amax, p):
        """kabs, ksca [n][n_lam] of the draws (amax[k], p[k]) at the tables'
        wavelengths: the values ``self(amax[k], p[k])`` holds (the same
        elementwise operations), for many walkers in one pass."""
        amax = np.atleast_1d(np.asarray(amax, dtype=float))
        p = np.atleast_1d(np.asarray(p, dtype=float))
        if amax.shape != p.shape or amax.ndim != 1:
            raise ValueError("DustGenerator.draw_many: amax and p must be 1-d of one length")
        i, tp = _bracket_many(self.p, p, "p")
        j, ta = _bracket_many(self.amax, amax, "amax")
        tp, ta = tp[:, None], ta[:, None]
        out = []
        for L in self._lk:
            v = ((1 - tp) * (1 - ta) * L[i, j] + (1 - tp) * ta * L[i, j + 1] +
                 tp * (1 - ta) * L[i + 1, j] + tp * ta * L[i + 1, j + 1])
            out.append(10.0 ** v)
        return out[0], out[1]
