"""RADMC-3D file interop: the text files pdspy writes and reads around its
RADMC-3D runs, so a temperature computed on the GPU can feed an unchanged
RADMC-3D imaging run, and a RADMC-3D model directory (or RADMC-3D's own
``dust_temperature.dat``) can be loaded into the GPU path.

Formats follow pdspy's writer and reader:

* ``radmc3d.inp``          -- ``key = value`` lines (pdspy/radmc3d/write.py:1-146)
* ``stars.inp``            -- iformat 2, blackbody stars as -T (write.py:148-171)
* ``wavelength_micron.inp``-- write.py:173-184
* ``amr_grid.inp``         -- regular spherical grid, walls in cm / rad / rad
  (write.py:198-238; Model.py:747-756)
* ``dust_density.inp``     -- iformat 1, r fastest (write.py:240-264)
* ``dust_temperature.dat`` -- iformat 1, ``%f`` values, r fastest
  (write.py:266-289); read as in read.pyx:433-478, reshaped to
  ``[nr, ntheta, nphi]`` as Model.py:249-254 does
* ``dustopac.inp`` / ``dustkappa_<name>.inp`` -- write.py:291-341

``write_model`` is Model.write_radmc3d (Model.py:723-778) for the continuum
part: densities, the rho-weighted species-mean temperature (Model.py:759-769)
when temperatures exist, opacities (without g, as pdspy writes them), stars,
wavelengths, grid and control keywords.  ``read_model_arrays`` goes the other
way and returns the arrays ``rt.DeviceModel`` takes.
"""
import os

import numpy as np

from .inputs import AU, M_sun, R_sun
from .rt import _loglog_interp, imaging_temperature


# ---------------------------------------------------------------- writing
def _fmt_value(v):
    if isinstance(v, (bool, np.bool_)):
        return "%d" % int(v)
    if isinstance(v, (int, np.integer)):
        return "%d" % int(v)
    if isinstance(v, float) and float(v).is_integer() and abs(v) < 2 ** 62:
        return "%d" % int(v)
    return "%f" % float(v)


def write_control(path, setthreads=1, **keywords):
    """radmc3d.inp: one ``key = value`` line per keyword that is not None
    (write.py:1-146; ``setthreads`` is written only when it is not 1)."""
    with open(path, "w") as f:
        for k, v in keywords.items():
            if v is not None:
                f.write("%s = %s\n" % (k, _fmt_value(v)))
        if setthreads != 1:
            f.write("setthreads = %d\n" % int(setthreads))


def _write_column(f, values, fmt):
    np.savetxt(f, np.asarray(values, dtype=float).reshape(-1, 1), fmt=fmt)


def write_stars(path, rstar_cm, mstar_g, lam_um, xyz_cm, tstar):
    """stars.inp (write.py:148-171): iformat 2; blackbody stars as -T."""
    n = len(rstar_cm)
    with open(path, "w") as f:
        f.write("2\n%d  %d\n" % (n, len(lam_um)))
        for i in range(n):
            f.write("%e   %e   %e   %e   %e\n" % (rstar_cm[i], mstar_g[i], xyz_cm[i][0],
                                                  xyz_cm[i][1], xyz_cm[i][2]))
        _write_column(f, lam_um, "%e")
        for i in range(n):
            f.write("%f\n" % -float(tstar[i]))


def write_wavelength_micron(path, lam_um):
    with open(path, "w") as f:
        f.write("%d\n" % len(lam_um))
        _write_column(f, lam_um, "%f")


def write_amr_grid(path, rw_cm, tw, pw):
    """amr_grid.inp for a regular spherical grid (write.py:198-238)."""
    rw_cm, tw, pw = (np.asarray(x, dtype=float) for x in (rw_cm, tw, pw))
    n = (rw_cm.size - 1, tw.size - 1, pw.size - 1)
    with open(path, "w") as f:
        f.write("1\n0\n100\n0\n")
        f.write("%d  %d  %d\n" % tuple(int(x > 1) for x in n))
        f.write("%d  %d  %d\n" % n)
        for w in (rw_cm, tw, pw):
            _write_column(f, w, "%12.9e")


def _write_cells(path, fields, fmt):
    """iformat 1, ncells, nspecies, then each [nr, ntheta, nphi] field with r
    fastest (write.py:240-289)."""
    fields = [np.asarray(x, dtype=float) for x in fields]
    ncell = fields[0].size
    with open(path, "w") as f:
        f.write("1\n%d\n%d\n" % (ncell, len(fields)))
        for x in fields:
            _write_column(f, x.transpose(2, 1, 0).ravel(), fmt)


def write_dust_density(path, density):
    _write_cells(path, density, "%e")


def write_dust_temperature(path, temperature):
    """dust_temperature.dat, ``%f`` values as pdspy writes them."""
    _write_cells(path, temperature, "%f")


def write_dustkappa(path, lam_um, kabs, ksca=None, g=None):
    """dustkappa_<name>.inp (write.py:316-341): iformat 1, 2 or 3."""
    cols = [lam_um, kabs]
    if ksca is not None:
        cols.append(ksca)
        if g is not None:
            cols.append(g)
    with open(path, "w") as f:
        f.write("%d\n%d\n" % (len(cols) - 1, len(lam_um)))
        np.savetxt(f, np.column_stack([np.asarray(c, dtype=float) for c in cols]),
                   fmt="   ".join(["%e"] * len(cols)))


def write_dustopac(path, names):
    """dustopac.inp (write.py:291-314) for dustkappa_<name>.inp files."""
    with open(path, "w") as f:
        f.write("2\n%d\n" % len(names))
        f.write("=" * 62 + "\n")
        for name in names:
            f.write("1\n0\n%s\n" % name)
            f.write("-" * 58 + "\n")


def write_model(model, directory=".", **control):
    """Writes the RADMC-3D inputs of a pdspy Model's continuum problem into
    ``directory`` (Model.write_radmc3d, Model.py:723-778).  ``control`` goes to
    radmc3d.inp (e.g. nphot_therm, iseed, modified_random_walk)."""
    g = model.grid
    if getattr(g, "coordsystem", "spherical") != "spherical":
        raise ValueError("only spherical grids are supported")
    os.makedirs(directory, exist_ok=True)
    p = lambda name: os.path.join(directory, name)
    write_control(p("radmc3d.inp"), **control)
    write_stars(p("stars.inp"), [s.radius * R_sun for s in g.stars],
                [s.mass * M_sun for s in g.stars], g.lam,
                [(s.x * AU, s.y * AU, s.z * AU) for s in g.stars],
                [s.temperature for s in g.stars])
    write_wavelength_micron(p("wavelength_micron.inp"), g.lam)
    write_amr_grid(p("amr_grid.inp"), np.asarray(g.w1) * AU, g.w2, g.w3)
    write_dust_density(p("dust_density.inp"), g.density)
    if len(g.temperature) > 0:
        nr, nt, np_ = len(g.w1) - 1, len(g.w2) - 1, len(g.w3) - 1
        T = to_grid_layout(imaging_temperature(g), nr, nt, np_)
        write_dust_temperature(p("dust_temperature.dat"), [T] * len(g.density))
    names = []
    for i, d in enumerate(g.dust):
        names.append("%d" % i)
        write_dustkappa(p("dustkappa_%d.inp" % i), np.asarray(d.lam) * 1e4, d.kabs, ksca=d.ksca)
    write_dustopac(p("dustopac.inp"), names)


# ---------------------------------------------------------------- reading
def _tokens(path):
    with open(path) as f:
        return f.read().split()


def read_dust_temperature(path="dust_temperature.dat", binary=False):
    """List (one per species) of flat temperature arrays, r fastest
    (read.pyx:433-478).  Binary files: four 8-byte header words (format,
    precision, ncells, nspecies), then doubles -- the data start at word 4
    (pdspy's reader starts at word 3, one word early; it never uses that path)."""
    if binary:
        hdr = np.fromfile(path, dtype="<i8", count=4)
        ncell, nspec = int(hdr[2]), int(hdr[3])
        data = np.fromfile(path, dtype="<f8", offset=32, count=ncell * nspec)
    else:
        t = _tokens(path)
        ncell, nspec = int(t[1]), int(t[2])
        data = np.array(t[3:3 + ncell * nspec], dtype=float)
    if data.size != ncell * nspec:
        raise ValueError("%s: expected %d values, found %d" % (path, ncell * nspec, data.size))
    return [data[i * ncell:(i + 1) * ncell].copy() for i in range(nspec)]


def read_amr_grid(path="amr_grid.inp"):
    """Walls (r in cm, theta, phi in rad) of a regular spherical amr_grid.inp."""
    t = _tokens(path)
    if int(t[1]) != 0 or int(t[2]) != 100:
        raise ValueError("%s: only regular spherical grids are supported" % path)
    nx, ny, nz = int(t[7]), int(t[8]), int(t[9])
    w = np.array(t[10:10 + nx + ny + nz + 3], dtype=float)
    return w[:nx + 1], w[nx + 1:nx + ny + 2], w[nx + ny + 2:]


def to_grid_layout(flat, nr, nt, np_):
    """r-fastest flat cell values -> [nr, ntheta, nphi] (Model.py:249-254)."""
    return np.asarray(flat).reshape((np_, nt, nr)).transpose(2, 1, 0)


def read_temperature_into(model, path="dust_temperature.dat", binary=False):
    """Sets ``model.grid.temperature`` from a RADMC-3D temperature file, as
    Model.run_thermal_radmc3d does after mctherm (Model.py:249-254)."""
    g = model.grid
    nr, nt, np_ = len(g.w1) - 1, len(g.w2) - 1, len(g.w3) - 1
    model.grid.temperature = [to_grid_layout(x, nr, nt, np_)
                              for x in read_dust_temperature(path, binary=binary)]
    return model.grid.temperature


def _read_stars(path):
    t = _tokens(path)
    if int(t[0]) != 2:
        raise ValueError("%s: only iformat 2 is supported" % path)
    n, nlam = int(t[1]), int(t[2])
    k = 3
    rows = []
    for _ in range(n):
        rows.append([float(x) for x in t[k:k + 5]])
        k += 5
    k += nlam
    stars = []
    for i in range(n):
        first = float(t[k])
        if first >= 0:
            raise ValueError("%s: star %d has a tabulated spectrum; only blackbody stars "
                             "(-T) are supported" % (path, i))
        stars.append([rows[i][0], -first, rows[i][2], rows[i][3], rows[i][4]])
        k += 1
    return np.array(stars, dtype=float).reshape(-1, 5)


def _read_dustkappa(path):
    t = _tokens(path)
    iformat, nlam = int(t[0]), int(t[1])
    ncol = {1: 2, 2: 3, 3: 4}[iformat]
    a = np.array(t[2:2 + nlam * ncol], dtype=float).reshape(nlam, ncol)
    ksca = a[:, 2] if ncol > 2 else np.zeros(nlam)
    g = a[:, 3] if ncol > 3 else None
    return a[:, 0], a[:, 1], ksca, g


def _read_dustopac(path):
    lines = [ln.strip() for ln in open(path) if ln.strip()]
    nspec = int(lines[1])
    names, k = [], 3
    for _ in range(nspec):
        style = int(lines[k])
        if style != 1:
            raise ValueError("%s: only dustkappa (input style 1) files are supported" % path)
        names.append(lines[k + 2])
        k += 4
    return names


def read_model_arrays(directory="."):
    """A RADMC-3D model directory -> the arrays ``rt.DeviceModel`` takes (the
    same layout ``rt.model_arrays`` builds from a pdspy Grid)."""
    p = lambda name: os.path.join(directory, name)
    rw, tw, pw = read_amr_grid(p("amr_grid.inp"))
    t = _tokens(p("dust_density.inp"))
    ncell, nspec = int(t[1]), int(t[2])
    if ncell != (len(rw) - 1) * (len(tw) - 1) * (len(pw) - 1):
        raise ValueError("dust_density.inp has %d cells, amr_grid.inp %d" %
                         (ncell, (len(rw) - 1) * (len(tw) - 1) * (len(pw) - 1)))
    rho = np.array(t[3:3 + ncell * nspec], dtype=float).reshape(
        nspec, len(pw) - 1, len(tw) - 1, len(rw) - 1)
    lam = np.array(_tokens(p("wavelength_micron.inp"))[1:], dtype=float)
    names = _read_dustopac(p("dustopac.inp"))
    if len(names) != nspec:
        raise ValueError("dustopac.inp lists %d species, dust_density.inp %d" %
                         (len(names), nspec))
    kabs, ksca, gg, has_g = [], [], [], False
    for name in names:
        lam_d, ka, ks, g = _read_dustkappa(p("dustkappa_%s.inp" % name))
        kabs.append(_loglog_interp(lam, lam_d, ka))
        ksca.append(_loglog_interp(lam, lam_d, ks) if np.any(ks > 0) else np.zeros_like(lam))
        has_g |= g is not None
        gg.append(np.interp(np.log(lam), np.log(lam_d), g) if g is not None
                  else np.zeros_like(lam))
    return dict(rw=rw, tw=tw, pw=pw, rho=np.ascontiguousarray(rho), lam_um=lam,
                kabs=np.ascontiguousarray(kabs), ksca=np.ascontiguousarray(ksca),
                gsca=np.ascontiguousarray(gg) if has_g else None,
                stars=_read_stars(p("stars.inp")))


__all__ = ["write_control", "write_stars", "write_wavelength_micron", "write_amr_grid",
           "write_dust_density", "write_dust_temperature", "write_dustkappa", "write_dustopac",
           "write_model", "read_dust_temperature", "read_amr_grid", "to_grid_layout",
           "read_temperature_into", "read_model_arrays"]
