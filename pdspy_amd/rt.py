"""Host side of the MI355X continuum RT path, mirroring pdspy's Model API.

The functions take a pdspy ``Model`` (or anything with the same ``grid``
attributes: ``w1/w2/w3``, ``density``, ``dust``, ``stars``, ``lam``,
``temperature``) and fill the same outputs pdspy's RADMC-3D path fills:

* ``run_thermal``      -- Model.run_thermal / run_thermal_radmc3d
  (pdspy/modeling/Model.py:54-72, 221-256): ``grid.temperature`` becomes a list
  (one per species) of float64 ``[nr, ntheta, nphi]`` arrays, exactly the
  reshape/transpose of Model.py:249-254.
* ``run_image``        -- Model.run_image_radmc3d (Model.py:515-577): an
  ``Image`` in Jy/pixel with x, y in arcsec (Model.py:566-572).
* ``run_visibilities`` -- Model.run_visibilities_radmc3d (Model.py:682-721) and
  imtovis (pdspy/imaging/imtovis.py:6-27).
* ``interpolate_model`` -- pdspy/interferometry/interpolate_model.py:11-57
  (exact DFT at arbitrary (u, v)).
* ``GPUModelMixin``    -- adds ``code="mi355x"`` to a Model subclass and defers
  every other ``code`` to pdspy's own implementation.

All heavy work runs in the HIP library (``pdspy_amd/lib/libpdspy_amd.so``);
there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib
from .inputs import AU, R_sun, Jy, arcsec, c as c_light, pc, Image, Visibilities

# radmc3d.inp keys pdspy's write.control accepts (pdspy/radmc3d/write.py:1-21);
# those that change the thermal MC are used, the rest are accepted and ignored
_CONTROL_KEYS = {
    "incl_dust", "incl_lines", "incl_freefree", "nphot_therm", "nphot_scat", "nphot_spec",
    "iseed", "ifast", "enthres", "itempdecoup", "istar_sphere", "ntemp", "temp0", "temp1",
    "scattering_mode_max", "rto_style", "camera_tracemode", "camera_nrrefine",
    "camera_refine_criterion", "camera_incl_stars", "camera_starsphere_nrpix",
    "camera_spher_cavity_relres", "camera_localobs_projection", "camera_min_dangle",
    "camera_max_dangle", "camera_min_dr", "camera_diagnostics_subpix", "camera_secondorder",
    "camera_interpol_jnu", "mc_weighted_photons", "optimized_motion", "lines_mode",
    "lines_maxdoppler", "lines_partition_ntempint", "lines_partition_temp0",
    "lines_partition_temp1", "lines_show_pictograms", "tgas_eq_tdust", "subbox_nx",
    "subbox_ny", "subbox_nz", "subbox_x0", "subbox_x1", "subbox_y0", "subbox_y1", "subbox_z0",
    "subbox_z1", "modified_random_walk", "mrw_gamma", "mrw_tauthres", "mrw_count_trigger",
    "setthreads", "camera_scatsrc_allfreq", "mc_scat_maxtauabs", "writeimage_unformatted",
    "camera_circ_nrphiinf", "camera_circ_dbdr",
}
# RADMC-3D's defaults for camera_nrrefine / camera_refine_criterion, which
# pdspy leaves unset unless the caller passes them
RADMC3D_NRREFINE = 100
RADMC3D_REFINE_CRITERION = 1.0

# keys of the run_* wrappers themselves (Model.py:221-222, 515-519)
_RUN_KEYS = {"verbose", "timelimit", "nice"}


# run_image_radmc3d's line-transfer and unstructured-camera arguments
# (Model.py:515-519, 682-686) with their defaults: accepted at the defaults,
# refused otherwise (continuum only on this path)
_LINE_DEFAULTS = {"imolspec": None, "iline": None, "widthkms": None, "vkms": None,
                  "linenlam": None, "doppcatch": False, "unstructured": False}


def _pop_line_options(keywords):
    for k, d in _LINE_DEFAULTS.items():
        if k in keywords:
            v = keywords.pop(k)
            if v is not d and v != d:
                raise NotImplementedError("%s=%r: line transfer and unstructured cameras are "
                                          "not on this path (continuum images only)" % (k, v))


def _check_keywords(keywords):
    bad = set(keywords) - _CONTROL_KEYS - _RUN_KEYS
    if bad:
        raise TypeError("unexpected keyword(s) for the RADMC-3D option set: %s" % sorted(bad))


def _loglog_interp(x_new, x, y):
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    if x.shape == np.shape(x_new) and np.allclose(x, x_new, rtol=1e-10, atol=0):
        return y.copy()
    if x[0] > x[-1]:
        x, y = x[::-1], y[::-1]
    pos = y > 0
    if np.all(pos):
        lx, ly = np.log(x), np.log(y)
        out = np.interp(np.log(x_new), lx, ly)
        # power-law extrapolation beyond the table ends
        lo = np.log(x_new) < lx[0]
        hi = np.log(x_new) > lx[-1]
        if lo.any():
            s = (ly[1] - ly[0]) / (lx[1] - lx[0])
            out[lo] = ly[0] + s * (np.log(x_new)[lo] - lx[0])
        if hi.any():
            s = (ly[-1] - ly[-2]) / (lx[-1] - lx[-2])
            out[hi] = ly[-1] + s * (np.log(x_new)[hi] - lx[-1])
        return np.exp(out)
    return np.interp(x_new, x, y)


def loglog_resample(lam_um, lam_src_um, k):
    """One opacity row k(lam_src_um) on the wavelengths lam_um, as
    ``dust_arrays`` resamples a Dust's kabs or ksca."""
    return _loglog_interp(np.asarray(lam_um, dtype=float), np.asarray(lam_src_um, dtype=float), k)


def dust_arrays(dust, lam_um):
    """pdspy Dust objects (lam in cm, kabs, ksca, optional g) -> kabs, ksca, g
    [nspec][nlam] on the wavelengths lam_um (micron), log-log interpolated as
    write_radmc3d's dustkappa tables are read (Model.py:770-775), and whether
    any species has g."""
    lam = np.asarray(lam_um, dtype=float)
    kabs, ksca, gg = [], [], []
    has_g = any(getattr(d, "g", None) is not None for d in dust)
    for d in dust:
        lam_d = np.asarray(d.lam, dtype=float) * 1e4
        kabs.append(_loglog_interp(lam, lam_d, d.kabs))
        ksca.append(_loglog_interp(lam, lam_d, d.ksca))
        g = getattr(d, "g", None)
        gg.append(np.interp(np.log(lam), np.log(lam_d), g) if g is not None else np.zeros_like(lam))
    return kabs, ksca, gg, has_g


def model_arrays(grid, hg=False):
    """pdspy Grid -> the flat FP64 arrays of rt_model_desc (units of
    write_radmc3d, Model.py:733-778: walls in cm/rad/rad, rho r-fastest,
    opacities on the grid's wavelengths in micron).  pdspy writes the dust
    opacity files with kabs and ksca only (Model.py:771-775), so RADMC-3D
    scatters isotropically even for Dust objects that carry an asymmetry
    parameter ``g``; the same holds here unless ``hg=True`` asks for
    Henyey-Greenstein scattering with the Dust's g."""
    if getattr(grid, "coordsystem", "spherical") != "spherical":
        raise ValueError("only spherical grids are supported")
    lam = np.asarray(grid.lam, dtype=float)
    rho = np.ascontiguousarray(np.array([np.asarray(d, dtype=float).transpose(2, 1, 0)
                                         for d in grid.density]))
    kabs, ksca, gg, has_g = dust_arrays(grid.dust, lam)
    stars = np.array([[s.radius * R_sun, s.temperature, s.x * AU, s.y * AU, s.z * AU]
                      for s in grid.stars], dtype=float).reshape(-1, 5)
    return dict(rw=np.asarray(grid.w1, dtype=float) * AU, tw=np.asarray(grid.w2, dtype=float),
                pw=np.asarray(grid.w3, dtype=float), rho=rho, lam_um=lam,
                kabs=np.ascontiguousarray(kabs), ksca=np.ascontiguousarray(ksca),
                gsca=np.ascontiguousarray(gg) if (has_g and hg) else None, stars=stars)


class DeviceModel:
    """A model resident on the GPU (``rt_model``)."""

    def __init__(self, arrays, device=0, ntemp=1000, temp0=0.01, temp1=1e5, ctx=None):
        # ctx: a Context of its own (its own stream, e.g. for a host thread that
        # runs models concurrently with others); default: the device's shared one
        self.ctx = ctx if ctx is not None else _lib.context(device)
        a = {k: (None if v is None else np.ascontiguousarray(v, dtype=float))
             for k, v in arrays.items()}
        self._keep = a
        rho = a["rho"]
        self.nspec = rho.shape[0]
        self.nr, self.nt, self.np = len(a["rw"]) - 1, len(a["tw"]) - 1, len(a["pw"]) - 1
        self.ncell = self.nr * self.nt * self.np
        P = lambda x: None if x is None else x.ctypes.data
        d = _lib.ModelDesc(self.nr, self.nt, self.np, P(a["rw"]), P(a["tw"]), P(a["pw"]),
                           self.nspec, P(rho), len(a["lam_um"]), P(a["lam_um"]), P(a["kabs"]),
                           P(a["ksca"]), P(a.get("gsca")), a["stars"].shape[0], P(a["stars"]),
                           int(ntemp), float(temp0), float(temp1))
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.rt_model_create(self.ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self._ntemp = int(ntemp) if ntemp else 1000
        self.rho_host = rho.reshape(self.nspec, self.ncell)
        self._rho_dev = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.L.rt_model_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_opacity(self, kabs, ksca, g=None, on_device=False):
        """New dust opacities [nspec][nlam] (cm^2/g) on the model's wavelengths
        (rt_model_set_opacity); the species grouping must stay as it was.
        ``on_device``: the per-species arrays and thermal tables are made by
        kernels on torch's current stream (rt_model_set_opacity_device) instead
        of on the host, and the call returns without waiting for them."""
        nl = len(self._keep["lam_um"])
        kabs = np.ascontiguousarray(kabs, dtype=float).reshape(self.nspec, nl)
        ksca = np.ascontiguousarray(ksca, dtype=float).reshape(self.nspec, nl)
        gg = None if g is None else np.ascontiguousarray(g, dtype=float).reshape(self.nspec, nl)
        # the device variant is ordered on torch's current stream; the host
        # variant copies synchronously
        fn = self.ctx.L.rt_model_set_opacity_device if on_device else self.ctx.L.rt_model_set_opacity
        self.ctx.sync_stream()
        self.ctx.check(fn(self.h, kabs.ctypes.data, ksca.ctypes.data,
                          None if gg is None else gg.ctypes.data))
        self._keep.update(kabs=kabs, ksca=ksca)
        if gg is not None:
            self._keep["gsca"] = gg

    def set_dust(self, dust, on_device=False):
        """New opacities from pdspy Dust objects, one per species (lam in cm,
        kabs, ksca and optionally g), resampled onto the model's wavelengths as
        the model was built (model_arrays); ``on_device`` as in set_opacity."""
        if len(dust) != self.nspec:
            raise ValueError("set_dust: %d Dust objects for %d species" % (len(dust), self.nspec))
        kabs, ksca, gg, _ = dust_arrays(dust, self._keep["lam_um"])
        self.set_opacity(np.array(kabs), np.array(ksca),
                         np.array(gg) if self._keep.get("gsca") is not None else None,
                         on_device=on_device)

    def tables(self):
        """The device's thermal tables: lnq [G][ntemp], cdf [G][ntemp][nlam],
        guide [G][ntemp][128], kplanck [G][ntemp], kross [nspec][ntemp]."""
        self.ctx.sync_stream()
        G = int(self.ctx.L.rt_model_ngroups(self.h))
        nT = int(self._ntemp)
        nl = len(self._keep["lam_um"])
        out = dict(lnq=np.zeros((G, nT)), cdf=np.zeros((G, nT, nl), dtype=np.float32),
                   guide=np.zeros((G, nT, 128), dtype=np.uint8),
                   kplanck=np.zeros((G, nT), dtype=np.float32),
                   kross=np.zeros((self.nspec, nT), dtype=np.float32))
        self.ctx.check(self.ctx.L.rt_model_get_tables(
            self.h, out["lnq"].ctypes.data, out["cdf"].ctypes.data, out["guide"].ctypes.data,
            out["kplanck"].ctypes.data, out["kross"].ctypes.data))
        return out

    def opacity_arrays(self):
        """The device's per-species opacities as the photon kernels read them:
        ka, ks, g [nspec][nlam] (float32), gka [G][nlam], ginvk [G] (float32),
        econv [cells][G] (float64)."""
        G = int(self.ctx.L.rt_model_ngroups(self.h))
        nl = len(self._keep["lam_um"])
        out = dict(ka=np.zeros((self.nspec, nl), dtype=np.float32),
                   ks=np.zeros((self.nspec, nl), dtype=np.float32),
                   g=np.zeros((self.nspec, nl), dtype=np.float32),
                   gka=np.zeros((G, nl), dtype=np.float32),
                   ginvk=np.zeros(G, dtype=np.float32),
                   econv=np.zeros((self.ncell, G)))
        self.ctx.check(self.ctx.L.rt_model_get_opacity(
            self.h, *(out[k].ctypes.data for k in ("ka", "ks", "g", "gka", "ginvk", "econv"))))
        return out

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho, dtype=float).reshape(self.nspec, self.ncell)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_model_set_density(self.h, rho.ctypes.data))
        self.rho_host = rho
        self._rho_dev = None

    def thermal(self, nphot, seed=0, **opts):
        o = _thermal_opts(opts)
        st = _lib.ThermalStats()
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_thermal(self.h, int(nphot), int(seed), C.byref(o), C.byref(st)))
        return st.as_dict()

    def temperature(self, with_visits=False):
        T = np.zeros((self.nspec, self.ncell))
        vis = np.zeros(self.ncell, dtype=np.uint64) if with_visits else None
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_get_temperature(
            self.h, T.ctypes.data, None if vis is None else vis.ctypes.data))
        return (T, vis) if with_visits else T

    def timeline(self):
        """ms from the start of the last thermal launch to each epoch's snapshot,
        and (last entry) to the launch's end."""
        n = self.ctx.L.rt_thermal_timeline(self.h, None, 0)
        out = np.zeros(max(n, 0))
        if n > 0:
            self.ctx.L.rt_thermal_timeline(self.h, out.ctypes.data, n)
        return out

    def temperature_device(self):
        import torch
        out = torch.empty((self.nspec, self.ncell), dtype=torch.float32, device="cuda")
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_get_temperature_device(self.h, C.c_void_p(out.data_ptr())))
        return out

    def rho_device(self):
        """The model's densities on the device, float [nspec, cells]."""
        import torch
        if self._rho_dev is None:
            out = torch.empty((self.nspec, self.ncell), dtype=torch.float32, device="cuda")
            self.ctx.sync_stream()
            self.ctx.check(self.ctx.L.rt_get_density_device(self.h, C.c_void_p(out.data_ptr())))
            self.ctx.sync_stream()
            self._rho_dev = out
        return self._rho_dev

    def generate_density(self, components):
        """Replaces the densities with components computed on the device:
        dicts with ``kind`` ("disk" or "ulrich_ext"), ``species`` and the
        parameters of pdspy's add_disk / add_ulrichextended_envelope (Msun,
        au, degrees).  See ``density_components``.  The kernels run on torch's
        current stream; the call returns without waiting for them."""
        arr = (_lib.DensityComp * len(components))()
        for k, c in enumerate(components):
            c = dict(c)
            kind = {"disk": _lib.RT_COMP_DISK, "ulrich_ext": _lib.RT_COMP_ULRICH_EXT}[c.pop("kind")]
            arr[k].kind = kind
            for key, v in c.items():
                if not hasattr(arr[k], key):
                    raise TypeError("unknown density component parameter %r" % key)
                setattr(arr[k], key, v)
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_model_generate_density(self.h, arr, len(components)))
        self.rho_host = None
        self._rho_dev = None

    @staticmethod
    def camera(npix, pix_cm, lam_um, incl=0.0, pa=0.0, phi=0.0, mode=0, incl_star=True,
               nsub=1, tau_stop=40.0, ny=None, tile=(0, 0), jscat=None, nrefine=0,
               refine_crit=1.0):
        """Camera; ``jscat``: the mean intensity J[cells] (float32 device tensor)
        of ``scattering_source`` at ``lam_um`` for images with the scattering
        source (the camera keeps a reference to it).  ``nrefine`` /
        ``refine_crit``: RADMC-3D's recursive sub-pixeling (camera_nrrefine /
        camera_refine_criterion), off at nrefine = 0."""
        if jscat is not None:
            jscat = jscat.contiguous()
        cam = _lib.Camera(int(npix), int(ny if ny is not None else npix), int(nsub),
                          float(pix_cm), float(incl), float(pa), float(phi), float(lam_um),
                          int(mode), 1 if incl_star else 0, float(tau_stop), int(tile[0]),
                          int(tile[1]), None if jscat is None else jscat.data_ptr(),
                          int(nrefine), float(refine_crit))
        cam._jscat = jscat
        return cam

    def image_crossings(self, reset=True):
        """(cells crossed by this context's image rays, by its adjoint sweeps)
        since the last reset"""
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.rt_image_crossings(self.ctx.h, C.byref(a), C.byref(b),
                                                     1 if reset else 0))
        return a.value, b.value

    def refine_stats(self):
        """(rays traced by the last sub-pixelled image of this model's context,
        sub-pixels left unsplit for want of pool space since the last call)"""
        n, u = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.L.rt_refine_stats(self.ctx.h, C.byref(n), C.byref(u)))
        return n.value, u.value

    def scattering_source(self, lam_um, nphot, T_dev, seed=0):
        """Mean intensity J[cells] (float32 device tensor, erg/s/cm^2/Hz/sr) at
        lam_um from the monochromatic scattering Monte Carlo with nphot packets
        (stars and the dust's thermal emission at the imaging temperature
        T_dev[cells]); returns (J, stats dict)."""
        import torch
        J = torch.empty(self.ncell, dtype=torch.float32, device="cuda")
        T_dev = T_dev.contiguous()
        st = _lib.ScatStats()
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_scattering_source(
            self.h, float(lam_um), int(nphot), int(seed), C.c_void_p(T_dev.data_ptr()),
            C.c_void_p(J.data_ptr()), C.byref(st)))
        return J, {k: getattr(st, k) for k, _ in _lib.ScatStats._fields_}

    def image(self, cam, T_dev, rho_dev=None):
        """Intensity image (erg/s/cm^2/Hz/sr, or tau) as a float32 torch tensor [ny, nx]."""
        import torch
        img = torch.empty((cam.ny, cam.nx), dtype=torch.float32, device="cuda")
        T_dev = T_dev.contiguous() if T_dev is not None else None
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_image(
            self.h, C.byref(cam), None if T_dev is None else C.c_void_p(T_dev.data_ptr()),
            None if rho_dev is None else C.c_void_p(rho_dev.contiguous().data_ptr()),
            C.c_void_p(img.data_ptr())))
        return img

    def image_adjoint(self, cam, T_dev, rho_dev, dimg):
        import torch
        g_lr = torch.empty((self.nspec, self.ncell), dtype=torch.float32, device="cuda")
        g_T = torch.empty((self.ncell,), dtype=torch.float32, device="cuda")
        dimg = dimg.contiguous().float()
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_image_adjoint(
            self.h, C.byref(cam), C.c_void_p(T_dev.contiguous().data_ptr()),
            None if rho_dev is None else C.c_void_p(rho_dev.contiguous().data_ptr()),
            C.c_void_p(dimg.data_ptr()), C.c_void_p(g_lr.data_ptr()), C.c_void_p(g_T.data_ptr())))
        return g_lr, g_T


def _thermal_opts(opts):
    o = _lib.ThermalOpts()
    for k, v in opts.items():
        if v is None:
            continue
        if not hasattr(o, k):
            raise TypeError("unknown thermal option %r" % k)
        setattr(o, k, v)
    return o


class Batch:
    """Several ``DeviceModel``s run as one sampler step (``rt_batch``): the
    thermal Monte Carlo of all of them in one persistent launch, then per model
    image -> visibilities -> chi^2 -> adjoint, optionally replayed from a
    captured hipGraph.  This is the per-step work of pdspy's fitting loop
    (run_disk_model.py:31-511) for a whole batch of walkers."""

    def __init__(self, models, device=0):
        self.ctx = _lib.context(device)
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m.h.value for m in self.models])
        h = C.c_void_p()
        self.ctx.check(self.ctx.L.rt_batch_create(self.ctx.h, arr, len(self.models), C.byref(h)))
        self.h = h
        self._keep = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.ctx.L.rt_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _arrays(self, nphot, seeds):
        n = len(self.models)
        nphot = np.ascontiguousarray(np.broadcast_to(np.asarray(nphot, dtype=np.uint64), (n,)))
        seeds = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.uint64), (n,)))
        return nphot, seeds

    def thermal(self, nphot, seeds, **opts):
        """One launch for all models; returns one stats dict per model."""
        nphot, seeds = self._arrays(nphot, seeds)
        o = _thermal_opts(opts)
        st = (_lib.ThermalStats * len(self.models))()
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_batch_thermal(self.h, nphot.ctypes.data, seeds.ctypes.data,
                                                   C.byref(o), st))
        return [x.as_dict() for x in st]

    @staticmethod
    def item(cam, dxy, u, v, re, im, w, img, chi2, g_logrho=None, g_T=None, scale=1.0,
             dra=0.0, ddec=0.0, nphot_scat=0):
        """A step item from device tensors (u, v float64; re, im, w, img,
        g_logrho, g_T float32; chi2 float64 [1]); nphot_scat > 0 adds the
        scattering source (that many packets) before the image."""
        P = lambda t: None if t is None else t.data_ptr()
        return _lib.StepItem(cam, float(dxy), float(scale), P(u), P(v), int(u.numel()),
                             float(dra), float(ddec), P(re), P(im), P(w), P(img), P(chi2),
                             P(g_logrho), P(g_T), int(nphot_scat))

    def step(self, items, nphot, seeds, use_graph=True, **opts):
        """One sampler step; returns (stats per model, device ms of the step)."""
        nphot, seeds = self._arrays(nphot, seeds)
        o = _thermal_opts(opts)
        arr = (_lib.StepItem * len(self.models))(*items)
        self._keep = arr
        st = (_lib.ThermalStats * len(self.models))()
        ms = C.c_double()
        self.ctx.sync_stream()
        self.ctx.check(self.ctx.L.rt_batch_step(self.h, nphot.ctypes.data, seeds.ctypes.data,
                                                C.byref(o), arr, 1 if use_graph else 0, st,
                                                C.byref(ms)))
        return [x.as_dict() for x in st], ms.value


def vis_dft(img, dxy, u, v, dra=0.0, ddec=0.0, device=0):
    """V at (u, v) of a float32 torch image [ny, nx]; returns (re, im) float32 tensors."""
    import torch
    ctx = _lib.context(device)
    img = img.contiguous().float()
    u = torch.as_tensor(u, dtype=torch.float64, device="cuda").contiguous()
    v = torch.as_tensor(v, dtype=torch.float64, device="cuda").contiguous()
    K = u.numel()
    re = torch.empty(K, dtype=torch.float32, device="cuda")
    im = torch.empty(K, dtype=torch.float32, device="cuda")
    ny, nx = img.shape
    ctx.check(ctx.L.rt_vis_dft(ctx.h, C.c_void_p(img.data_ptr()), nx, ny, float(dxy),
                               C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), K, float(dra),
                               float(ddec), C.c_void_p(re.data_ptr()), C.c_void_p(im.data_ptr())))
    return re, im


def vis_grid(img, dxy, u, v, dra=0.0, ddec=0.0, device=0, work=None, conv="imtovis"):
    """V at (u, v) of a float32 torch image [ny, nx] by galario's method (the
    image's DFT on its pixel-conjugate grid, bilinear interpolation, phase-centre
    shift; galario sampleImage as called at interpolate_model.py:20-27).
    ``conv="imtovis"``: the imtovis sign convention (rt_vis_grid);
    ``conv="galario"``: what pdspy's interpolate_model makes of galario's
    output -- sampleImage(image[::-1]) with the imaginary part negated
    (rt_interpolate_galario).  A point outside the grid's interpolation range
    raises ValueError.  ``work``: an optional contiguous device tensor of at
    least ``rt_vis_grid_workspace(nx, ny)`` bytes, else the library allocates
    one stream-ordered per call."""
    import torch
    if conv not in ("imtovis", "galario"):
        raise ValueError("vis_grid: conv must be 'imtovis' or 'galario'")
    ctx = _lib.context(device)
    img = img.contiguous().float()
    ny, nx = img.shape
    u = torch.as_tensor(u, dtype=torch.float64, device="cuda").contiguous()
    v = torch.as_tensor(v, dtype=torch.float64, device="cuda").contiguous()
    K = u.numel()
    if K:
        # the four nodes around each point must lie on the grid (galario's
        # sampleImage checks the uv coverage against the image the same way);
        # the galario convention interpolates at (-u, -v)
        sg = -1.0 if conv == "galario" else 1.0
        fa = sg * u * (nx * dxy) + nx // 2
        fb = sg * v * (ny * dxy) + ny // 2
        ok = ((fa >= 0) & (fa < nx - 1) & (fb >= 0) & (fb < ny - 1)).all()
        if not bool(ok):
            raise ValueError("vis_grid: (u,v) beyond the image's uv grid (|u| < (nx/2 - 1) / "
                             "(nx dxy)); use a finer pixel size or vis_dft")
    re = torch.empty(K, dtype=torch.float32, device="cuda")
    im = torch.empty(K, dtype=torch.float32, device="cuda")
    need = int(ctx.L.rt_vis_grid_workspace(nx, ny))
    if work is not None:
        have = work.numel() * work.element_size()
        if not (work.is_cuda and work.is_contiguous()) or have < need:
            raise ValueError("vis_grid: work must be a contiguous device tensor of >= %d bytes "
                             "(got %d)" % (need, have))
    fn = ctx.L.rt_interpolate_galario if conv == "galario" else ctx.L.rt_vis_grid
    ctx.check(fn(ctx.h, C.c_void_p(img.data_ptr()), nx, ny, float(dxy),
                 C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()), K, float(dra),
                 float(ddec), C.c_void_p(re.data_ptr()), C.c_void_p(im.data_ptr()),
                 C.c_void_p(work.data_ptr() if work is not None else None)))
    return re, im


def vis_dft_adjoint(gre, gim, nx, ny, dxy, u, v, dra=0.0, ddec=0.0, device=0):
    import torch
    ctx = _lib.context(device)
    gre = gre.contiguous().float()
    gim = gim.contiguous().float()
    u = torch.as_tensor(u, dtype=torch.float64, device="cuda").contiguous()
    v = torch.as_tensor(v, dtype=torch.float64, device="cuda").contiguous()
    g = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ctx.check(ctx.L.rt_vis_dft_adjoint(ctx.h, C.c_void_p(gre.data_ptr()), C.c_void_p(gim.data_ptr()),
                                       nx, ny, float(dxy), C.c_void_p(u.data_ptr()),
                                       C.c_void_p(v.data_ptr()), u.numel(), float(dra), float(ddec),
                                       C.c_void_p(g.data_ptr())))
    return g


def chisq(dre, dim, w, mre, mim, with_grad=False, device=0):
    """chi^2 (float64 tensor, shape [1]) and optionally d chi^2 / d(mre, mim)."""
    import torch
    ctx = _lib.context(device)
    ts = [t.contiguous().float() for t in (dre, dim, w, mre, mim)]
    K = ts[0].numel()
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    gre = torch.empty(K, dtype=torch.float32, device="cuda") if with_grad else None
    gim = torch.empty(K, dtype=torch.float32, device="cuda") if with_grad else None
    ctx.check(ctx.L.rt_chisq(ctx.h, *[C.c_void_p(t.data_ptr()) for t in ts], K,
                             C.c_void_p(out.data_ptr()),
                             None if gre is None else C.c_void_p(gre.data_ptr()),
                             None if gim is None else C.c_void_p(gim.data_ptr())))
    return (out, gre, gim) if with_grad else out


def lnlike_vis(data, model, device=0):
    """pdspy's log-likelihood of one visibility data set (utils/emcee.py:31-43):
    -1/2 sum w dRe^2 - sum_{w>0} log(w / 2 pi) - 1/2 sum w dIm^2 - sum_{w>0}
    log(w / 2 pi), the weighted sums on the device (rt_chisq), the weight term
    -- which does not depend on the model -- in FP64 numpy as pdspy has it.
    ``data``: Visibilities (real, imag, weights [K, 1]); ``model``: Visibilities
    or a (re, im) pair of float32 device tensors."""
    import torch
    w = np.asarray(data.weights, dtype=float).ravel()
    lw = np.sum(np.log(w[w > 0] / (2 * np.pi)))

    def dev(x):
        return torch.as_tensor(np.asarray(x, dtype=float).ravel(), dtype=torch.float32,
                               device="cuda")

    if isinstance(model, tuple):
        mre, mim = model
    else:
        mre, mim = dev(model.real), dev(model.imag)
    c2 = chisq(dev(data.real), dev(data.imag), dev(w), mre, mim, device=device)
    return -0.5 * float(c2.item()) - lw - lw


# ---------------------------------------------------------------- pdspy mirror
def _arrays_digest(arrays):
    """digest of a model's input arrays (walls, densities, opacities, stars)"""
    import hashlib
    h = hashlib.blake2b(digest_size=16)
    for k in sorted(arrays):
        v = arrays[k]
        h.update(k.encode())
        if v is None:
            h.update(b"-")
        else:
            v = np.ascontiguousarray(v, dtype=float)
            h.update(str(v.shape).encode())
            h.update(v.tobytes())
    return h.hexdigest()


def _device_model(model, device=0, ntemp=1000, temp0=0.01, temp1=1e5, hg=False):
    """The model's resident device copy.  pdspy rewrites every input file on
    each run, so any change of walls, densities (also in place), dust or stars
    is honoured: the inputs are gathered again and compared by digest, and the
    device model is rebuilt when they differ."""
    arrays = model_arrays(model.grid, hg=hg)
    key = (_arrays_digest(arrays), ntemp, temp0, temp1, device)
    dm = getattr(model, "_pdspy_amd_dev", None)
    if dm is None or dm[0] != key:
        dev = DeviceModel(arrays, device=device, ntemp=ntemp, temp0=temp0, temp1=temp1)
        model._pdspy_amd_dev = (key, dev)
        return dev
    return dm[1]


def run_thermal(model, nphot=1e6, seed=None, device=0, max_events=None, epoch_first=None,
                epoch_growth=None, count_visits=False, return_stats=False, timelimit=7200,
                hg=False, **keywords):
    """Model.run_thermal(code="mi355x"): fills ``model.grid.temperature``.

    Accepts the RADMC-3D keyword set pdspy forwards (``iseed`` is the seed,
    ``modified_random_walk``/``mrw_*`` select MRW, ``ntemp``/``temp0``/``temp1``
    set the emission tables); ``timelimit`` (seconds, pdspy's default 7200,
    Model.py:221) bounds the kernel time; thread, niceness and verbosity
    keywords are accepted and ignored.  ``hg=True`` (no pdspy counterpart)
    scatters with the Dust objects' Henyey-Greenstein g.
    """
    _check_keywords(keywords)
    if seed is None:
        seed = keywords.get("iseed", 0) or 0
    dm = _device_model(model, device, keywords.get("ntemp") or 1000,
                       keywords.get("temp0") or 0.01, keywords.get("temp1") or 1e5, hg=hg)
    opts = dict(max_events=max_events, epoch_first=epoch_first, epoch_growth=epoch_growth,
                count_visits=1 if count_visits else 0,
                mrw=1 if keywords.get("modified_random_walk") else 0,
                mrw_tauthres=keywords.get("mrw_tauthres"),
                mrw_count_trigger=keywords.get("mrw_count_trigger"),
                mrw_gamma=keywords.get("mrw_gamma"),
                max_seconds=float(timelimit) if timelimit else None)
    stats = dm.thermal(int(nphot), int(seed), **opts)
    T, vis = dm.temperature(with_visits=True)
    nr, nt, np_ = dm.nr, dm.nt, dm.np
    model.grid.temperature = [np.transpose(T[s].reshape((np_, nt, nr)), axes=(2, 1, 0))
                              for s in range(dm.nspec)]
    model._pdspy_amd_visits = np.transpose(vis.reshape((np_, nt, nr)), axes=(2, 1, 0))
    return stats if return_stats else None


def imaging_temperature(grid):
    """Density-weighted species-mean temperature, as write_radmc3d hands
    RADMC-3D for imaging (Model.py:759-769); returns [cells] r-fastest."""
    density = np.array(grid.density, dtype=float)
    temperature = np.array(grid.temperature, dtype=float)
    density[density == 0] = 1.0e-30
    T = (density * temperature).sum(axis=0) / density.sum(axis=0)
    return np.ascontiguousarray(T.transpose(2, 1, 0)).reshape(-1)


def _camera_lams(model, lam, loadlambda):
    if loadlambda:
        return np.atleast_1d(np.asarray(model.camera_wavelength, dtype=float))
    return np.atleast_1d(np.asarray([float(x) for x in np.atleast_1d(lam)], dtype=float))


def _scatters(model, keywords):
    """RADMC-3D adds the scattering source when the opacities scatter, unless
    scattering_mode_max = 0 (pdspy writes the option only when given)."""
    if keywords.get("scattering_mode_max", None) == 0:
        return False
    return any(np.any(np.asarray(d.ksca) > 0) for d in model.grid.dust)


def _render(model, npix, pixelsize, lams, incl, pa, phi, dpc, nostar, nsub, device,
            nphot=0, keywords=None):
    import torch
    keywords = keywords or {}
    dm = _device_model(model, device)
    if len(model.grid.temperature) == 0:
        raise RuntimeError("run_thermal (or grid.add_temperature) must come before imaging")
    T = torch.tensor(imaging_temperature(model.grid), dtype=torch.float32, device="cuda")
    pix_cm = pixelsize * dpc * AU
    scat = int(nphot) > 0 and _scatters(model, keywords)
    # RADMC-3D's recursive sub-pixeling at its defaults unless pdspy passes the
    # options (write.py:57-61 writes them only when given)
    nrefine = keywords.get("camera_nrrefine")
    nrefine = RADMC3D_NRREFINE if nrefine is None else int(nrefine)
    crit = keywords.get("camera_refine_criterion")
    crit = RADMC3D_REFINE_CRITERION if crit is None else float(crit)
    seed = int(keywords.get("iseed") or 0) & 0xFFFFFFFFFFFFFFFF
    chans = []
    for lam_um in lams:
        J = dm.scattering_source(lam_um, int(nphot), T, seed=seed)[0] if scat else None
        cam = dm.camera(npix, pix_cm, lam_um, incl=incl, pa=pa, phi=phi, incl_star=not nostar,
                        nsub=nsub, jscat=J, nrefine=nrefine, refine_crit=crit)
        chans.append(dm.image(cam, T).double().cpu().numpy())
    img = np.stack(chans, axis=-1)  # [ny, nx, nlam], erg/s/cm^2/Hz/sr
    # Jy/pixel (Model.py:566-567)
    img = img / Jy * (pix_cm / (dpc * pc)) ** 2
    return img[:, :, :, None]


def run_image(model, name=None, nphot=1e6, npix=256, pixelsize=1.0, lam="1300",
              loadlambda=False, incl=0, pa=0, phi=0, dpc=1, nostar=False, nsub=1, device=0,
              **keywords):
    """Model.run_image(code="mi355x") (Model.py:515-577): ``model.images[name]``."""
    _pop_line_options(keywords)
    _check_keywords(keywords)
    lams = _camera_lams(model, lam, loadlambda)
    image = _render(model, npix, pixelsize, lams, incl, pa, phi, dpc, nostar, nsub, device,
                    nphot=nphot, keywords=keywords)
    x = (np.arange(npix) - npix // 2) * float(pixelsize)
    model.images[name] = Image(image, x=x.astype(float), y=x.astype(float).copy(),
                               wave=lams * 1.0e-4)
    return model.images[name]


def fft_grid_uv(npix, pixelsize):
    """u, v of imtovis (imtovis.py:18-23): fftshift(fftfreq(n, dx[rad])) meshgrid, raveled."""
    uu = np.fft.fftshift(np.fft.fftfreq(npix, pixelsize * arcsec))
    u, v = np.meshgrid(uu, uu)
    return u.reshape(-1), v.reshape(-1)


def run_visibilities(model, name=None, nphot=1e6, npix=256, pixelsize=1.0, lam="1300",
                     loadlambda=False, incl=0, pa=0, phi=0, dpc=1, nostar=False, nsub=1,
                     device=0, **keywords):
    """Model.run_visibilities(code="mi355x") (Model.py:682-721): the image's
    visibilities on its pixel-conjugate grid (imtovis.py:6-27)."""
    import torch
    _pop_line_options(keywords)
    _check_keywords(keywords)
    lams = _camera_lams(model, lam, loadlambda)
    image = _render(model, npix, pixelsize, lams, incl, pa, phi, dpc, nostar, nsub, device,
                    nphot=nphot, keywords=keywords)
    u, v = fft_grid_uv(npix, pixelsize)
    real = np.empty((npix * npix, len(lams)))
    imag = np.empty((npix * npix, len(lams)))
    for i in range(len(lams)):
        img = torch.tensor(image[:, :, i, 0], dtype=torch.float32, device="cuda")
        re, im = vis_dft(img, pixelsize * arcsec, u, v, device=device)
        real[:, i] = re.double().cpu().numpy()
        imag[:, i] = im.double().cpu().numpy()
    freq = c_light / (lams * 1.0e-4)
    model.visibilities[name] = Visibilities(u, v, freq, real, imag, np.ones(real.shape))
    return model.visibilities[name]


def interpolate_model(u, v, freq, model, nthreads=1, dRA=0.0, dDec=0.0, code="galario",
                      nxy=1024, dxy=0.01, device=0):
    """Model visibilities of an Image at arbitrary (u, v) (interpolate_model.py:11-57,
    same arguments in the same order); dRA/dDec in arcsec.  ``code="galario"``
    (pdspy's default) gives what pdspy computes there: galario's
    sampleImage(image[::-1], dxy, u, v, dRA, dDec) with the imaginary part
    negated (interpolate_model.py:23-27), i.e. bilinear interpolation of the
    flipped image's transform on its pixel-conjugate grid
    (``vis_grid(conv="galario")``).  ``code="exact"`` evaluates the same
    visibility function exactly at each point instead of interpolating: on the
    grid nodes pdspy's galario call equals
    exp(-2 pi i (u dRA + v (dDec + dxy))) DFT(-u, v) with DFT the imtovis
    transform of the unflipped image (the flip moves the phase origin by one
    pixel), so ``vis_dft`` is called at (-u, v) with the shift (-dRA, dDec + dxy).
    pdspy's other codes (galario-unstructured, trift) are not on this path;
    ``nthreads`` (galario's host threads) and ``nxy``/``dxy`` (used only by
    galario-unstructured) are accepted and unused."""
    import torch
    if code not in ("galario", "exact"):
        raise ValueError("interpolate_model: code %r is not supported (galario, exact)" % (code,))
    pix = (model.x[1] - model.x[0]) * arcsec  # rad (interpolate_model.py:20)
    u_h = np.asarray(u, dtype=float).reshape(-1)
    v_h = np.asarray(v, dtype=float).reshape(-1)
    real, imag = [], []
    for i in range(model.image.shape[2]):
        img = torch.tensor(model.image[:, :, i, 0], dtype=torch.float32, device="cuda")
        if code == "galario":
            re, im = vis_grid(img, pix, u_h, v_h, dRA * arcsec, dDec * arcsec, device=device,
                              conv="galario")
        else:
            re, im = vis_dft(img, pix, -u_h, v_h, -dRA * arcsec, dDec * arcsec + pix,
                             device=device)
        real.append(re.double().cpu().numpy().reshape(-1, 1))
        imag.append(im.double().cpu().numpy().reshape(-1, 1))
    real = np.concatenate(real, axis=1)
    imag = np.concatenate(imag, axis=1)
    return Visibilities(np.asarray(u, float), np.asarray(v, float), np.asarray(freq, float),
                        real, imag, np.ones(real.shape))


class GPUModelMixin:
    """Mix into a pdspy Model subclass: ``code="mi355x"`` runs here, every other
    ``code`` goes to pdspy's own method (RADMC-3D or Hyperion)."""

    def run_thermal(self, nphot=1e6, code="radmc3d", **keywords):
        if code == "mi355x":
            return run_thermal(self, nphot=nphot, **keywords)
        return super().run_thermal(nphot=nphot, code=code, **keywords)

    def run_image(self, name=None, nphot=1e6, code="radmc3d", **keywords):
        if code == "mi355x":
            return run_image(self, name=name, nphot=nphot, **keywords)
        return super().run_image(name=name, nphot=nphot, code=code, **keywords)

    def run_visibilities(self, name=None, nphot=1e6, code="radmc3d", **keywords):
        if code == "mi355x":
            return run_visibilities(self, name=name, nphot=nphot, **keywords)
        return super().run_visibilities(name=name, nphot=nphot, code=code, **keywords)
