"""pdspy-compatible model inputs: Grid, Star, Dust, density models, containers.

These mirror the attribute names and units of pdspy's own classes, so that the
GPU path reads a real pdspy ``Model`` and one of these interchangeably (duck
typing; nothing here is needed when pdspy itself is importable):

* ``Grid``  -- pdspy/modeling/Grid.py:7-87 (``w1/w2/w3`` walls in au / rad / rad,
  cell centres as arithmetic midpoints, ``density``/``dust``/``temperature``
  lists of ``[nr, ntheta, nphi]`` arrays, ``stars``, ``lam`` in micron).
* ``Star``  -- pdspy/modeling/Star.py:6-16 (radius from L and T in R_sun).
* ``Dust``  -- pdspy/dust/Dust.py:181-188 (``lam`` in cm, ``kabs``/``ksca``).
* ``Disk.density`` -- pdspy/modeling/Disk.py:48-70, 86-135.
* ``UlrichEnvelopeExtended.density`` -- pdspy/modeling/UlrichEnvelopeExtended.py:37-110.
* ``Model`` -- the subset of pdspy/modeling/YSOModel.py:29-115, 595-651 used to
  build the workloads (``add_star``, ``set_spherical_grid``, ``add_disk``,
  ``add_ulrichextended_envelope``).
* ``Image`` / ``Visibilities`` -- pdspy/imaging/libimaging.pyx:7-46 and
  pdspy/interferometry/libinterferometry.pyx:10-44 (same field names).
"""
import numpy as np
from scipy.optimize import brenth

# pdspy CGS constants (pdspy/constants/astronomy.py, physics.py, math.py)
AU = 1.495978707e13
pc = 3.085677581e18
M_sun = 1.99e33
R_sun = 6.96e10
L_sun = 3.9e33
Jy = 1.0e-23
arcsec = 4.84813681e-6
c = 2.99792458e10
h = 6.6260755e-27
k = 1.380658e-16
sigma = 5.67051e-5
pi_pdspy = 3.14159265  # pdspy/constants/math.py:1 (truncated on purpose)


def _trapz(y, x, axis=-1):
    return np.trapezoid(y, x, axis=axis)


class Star:
    def __init__(self, mass=0.5, luminosity=1.0, temperature=4000., x=0.0, y=0.0, z=0.0):
        self.mass = mass
        self.luminosity = luminosity
        self.temperature = temperature
        self.radius = (luminosity * L_sun / (4 * np.pi * sigma * temperature ** 4)) ** 0.5 / R_sun
        self.x, self.y, self.z = x, y, z


class Dust:
    def set_properties(self, lam, kabs, ksca, g=None):
        self.lam = np.asarray(lam, dtype=float)
        self.nu = c / self.lam
        self.kabs = np.asarray(kabs, dtype=float)
        self.ksca = np.asarray(ksca, dtype=float)
        self.kext = self.kabs + self.ksca
        self.albedo = self.ksca / self.kext
        if g is not None:
            self.g = np.asarray(g, dtype=float)


class Grid:
    def __init__(self):
        self.density = []
        self.dust = []
        self.temperature = []
        self.stars = []

    def add_density(self, density, dust):
        self.density.append(density)
        self.dust.append(dust)

    def add_star(self, star):
        self.stars.append(star)

    def add_temperature(self, temperature):
        self.temperature.append(temperature)

    def set_spherical_grid(self, w1, w2, w3):
        self.coordsystem = "spherical"
        self.r = 0.5 * (w1[:-1] + w1[1:])
        self.theta = 0.5 * (w2[:-1] + w2[1:])
        self.phi = 0.5 * (w3[:-1] + w3[1:])
        self.w1, self.w2, self.w3 = w1, w2, w3

    def set_wavelength_grid(self, lmin, lmax, nlam, log=False):
        if log:
            self.lam = np.logspace(np.log10(lmin), np.log10(lmax), nlam)
        else:
            self.lam = np.linspace(lmin, lmax, nlam)


class Disk:
    """Flared power-law disk density (pdspy/modeling/Disk.py:48-135)."""

    def __init__(self, mass=1.0e-3, rmin=0.1, rmax=300, plrho=2.37, h0=0.1, plh=58. / 45.):
        self.mass, self.rmin, self.rmax = mass, rmin, rmax
        self.plrho, self.h0, self.plh = plrho, h0, plh

    def surface_density(self, r, normalize=True):
        rin, rout = self.rmin * AU, self.rmax * AU
        mass = self.mass * M_sun
        gamma = self.plrho - self.plh
        Sigma0 = (2 - gamma) * mass / (2 * pi_pdspy * AU ** gamma) / \
            (rout ** (-gamma + 2) - rin ** (-gamma + 2))
        r = np.asarray(r, dtype=float)
        with np.errstate(divide="ignore"):
            Sigma = Sigma0 * r ** (-gamma)
        Sigma = np.where((r >= rout / AU) ^ (r <= rin / AU), 0.0, Sigma)
        if np.any(r == 0):
            dr = r[r > 0].min()
            Sigma = np.where(r == 0, Sigma0 * (0.7 * dr) ** (-gamma), Sigma)
        if normalize:
            r_high = np.logspace(np.log10(self.rmin), np.log10(self.rmax), 1000)
            Sigma_high = self.surface_density(r_high, normalize=False)
            scale = mass / (2 * np.pi * _trapz(r_high * AU * Sigma_high, r_high * AU))
            Sigma = Sigma * scale
        return Sigma

    def scale_height(self, r):
        return self.h0 * AU * r ** self.plh

    def density(self, x1, x2, x3):
        rt, tt, pp = np.meshgrid(x1 * AU, x2, x3, indexing="ij")
        rr = rt * np.sin(tt)
        zz = rt * np.cos(tt)
        Sigma = self.surface_density(rr / AU)
        hh = self.scale_height(rr / AU)
        return Sigma / (np.sqrt(2 * np.pi) * hh) * np.exp(-0.5 * (zz / hh) ** 2)


class UlrichEnvelopeExtended:
    """Rotating-collapse envelope with a conical cavity
    (pdspy/modeling/UlrichEnvelopeExtended.py:37-110)."""

    def __init__(self, mass=1.0e-3, rmin=0.1, rmax=1000, rcent=30, cavpl=1.0,
                 cavrfact=0.2, theta_open=1., zoffset=1.):
        self.mass, self.rmin, self.rmax, self.rcent = mass, rmin, rmax, rcent
        self.cavpl, self.cavrfact = cavpl, cavrfact
        self.theta_open, self.zoffset = theta_open, zoffset

    def density(self, r, theta, phi):
        rin, rout = self.rmin * AU, self.rmax * AU
        mass = self.mass * M_sun
        rcent = self.rcent * AU
        cavz0 = self.zoffset * AU
        cavpl, cavrfact = self.cavpl, self.cavrfact
        theta_open = self.theta_open * np.pi / 180.
        cava = (rout * np.sin(theta_open / 2) / AU / np.tan(theta_open / 2) - cavz0 / AU) * \
            (rout * np.sin(theta_open / 2) / AU) ** -cavpl
        rr, tt, pp = np.meshgrid(r * AU, theta, phi, indexing="ij")
        mu = np.cos(tt)
        RR = rr * np.sin(tt)
        zz = rr * np.cos(tt)

        def func(mu0, r_, mu_, R_c):
            return mu0 ** 3 - mu0 * (1 - r_ / R_c) - mu_ * (r_ / R_c)

        mu0 = mu * 0.
        for ir in range(rr.shape[0]):
            for it in range(rr.shape[1]):
                mu0[ir, it, 0] = brenth(func, -1.0, 1.0, args=(rr[ir, it, 0], mu[ir, it, 0], rcent))
        if mu0.shape[2] > 1:
            mu0[:, :, 1:] = mu0[:, :, :1]
        rho0 = 1.0
        with np.errstate(divide="ignore", invalid="ignore"):
            rho = rho0 * (rr / rcent) ** (-1.5) * (1 + mu / mu0) ** (-0.5) * \
                (mu / mu0 + 2 * mu0 ** 2 * rcent / rr) ** (-1)
            mid1 = (np.abs(mu) < 1.0e-10) & (rr < rcent)
            rho[mid1] = rho0 * (rr[mid1] / rcent) ** (-0.5) * (1. - rr[mid1] / rcent) ** (-1) / 2.
            mid2 = (np.abs(mu) < 1.0e-10) & (rr > rcent)
            rho[mid2] = rho0 * (2. * rr[mid2] / rcent - 1) ** (-0.5) * (rr[mid2] / rcent - 1.) ** (-1)
        rho[(rr >= rout) ^ (rr <= rin)] = 0e0
        rho[np.abs(zz) / AU - cavz0 / AU - cava * (RR / AU) ** cavpl > 0.0] *= cavrfact
        if theta.max() > pi_pdspy / 2:
            mdot = mass / (2 * pi_pdspy * _trapz(_trapz(rho * rr ** 2 * np.sin(tt), tt, axis=1),
                                                 rr[:, 0, :], axis=0))[0]
        else:
            mdot = mass / (4 * pi_pdspy * _trapz(_trapz(rho * rr ** 2 * np.sin(tt), tt, axis=1),
                                                 rr[:, 0, :], axis=0))[0]
        return rho * mdot


class Image:
    """Image container with pdspy's layout: image[ny, nx, nfreq, npol]."""

    def __init__(self, image=None, x=None, y=None, wave=None, freq=None, unc=None):
        self.image, self.x, self.y, self.unc = image, x, y, unc
        if wave is not None and freq is None:
            self.wave = wave
            self.freq = c / wave
        elif freq is not None and wave is None:
            self.freq = freq
            self.wave = c / freq
        else:
            self.wave, self.freq = wave, freq


class Visibilities:
    """Visibility container with pdspy's fields (u, v in wavelengths)."""

    def __init__(self, u=None, v=None, freq=None, real=None, imag=None, weights=None):
        self.u, self.v, self.freq = u, v, freq
        if u is not None and v is not None:
            self.uvdist = np.sqrt(u ** 2 + v ** 2)
        self.real, self.imag = real, imag
        if real is not None and imag is not None:
            self.amp = np.sqrt(real ** 2 + imag ** 2)
            self.phase = np.arctan2(imag, real)
            self.weights = weights if weights is not None else np.ones(real.shape)


class Model:
    """The subset of pdspy's YSOModel used to describe the workloads."""

    def __init__(self):
        self.grid = Grid()
        self.images = {}
        self.visibilities = {}

    def add_star(self, mass=0.5, luminosity=1, temperature=4000.):
        self.grid.add_star(Star(mass=mass, luminosity=luminosity, temperature=temperature))

    def set_spherical_grid(self, rmin, rmax, nr, ntheta, nphi, log=True):
        """Takes WALL counts, as pdspy does (YSOModel.py:47-63, code="radmc3d")."""
        r = np.logspace(np.log10(rmin), np.log10(rmax), nr) if log else np.linspace(rmin, rmax, nr)
        theta = np.linspace(0.0, np.pi / 2, ntheta)
        phi = np.linspace(0.0, 2 * np.pi, nphi)
        self.grid.set_spherical_grid(r, theta, phi)

    def add_disk(self, mass=1.0e-3, rmin=0.1, rmax=300, plrho=2.37, h0=0.1, plh=58. / 45.,
                 dust=None):
        self.disk = Disk(mass=mass, rmin=rmin, rmax=rmax, plrho=plrho, h0=h0, plh=plh)
        if dust is not None:
            self.grid.add_density(self.disk.density(self.grid.r, self.grid.theta, self.grid.phi),
                                  dust)

    def add_ulrichextended_envelope(self, mass=1.0e-3, rmin=0.1, rmax=1000, rcent=300,
                                    cavpl=1.0, cavrfact=0.2, theta_open=45., zoffset=1.,
                                    dust=None, rcent_ne_rdisk=False):
        if not rcent_ne_rdisk and hasattr(self, "disk"):
            rcent = self.disk.rmax
        self.envelope = UlrichEnvelopeExtended(mass=mass, rmin=rmin, rmax=rmax, rcent=rcent,
                                               cavpl=cavpl, cavrfact=cavrfact,
                                               theta_open=theta_open, zoffset=zoffset)
        if dust is not None:
            self.grid.add_density(self.envelope.density(self.grid.r, self.grid.theta,
                                                        self.grid.phi), dust)
