"""One batched sampler step: the work pdspy's fitting loop does for every walker
of an MCMC step (pdspy/modeling/run_disk_model.py:31-511 with
utils/emcee.py:183-203: thermal run, 1.3 mm image, visibilities at the data's
(u,v), chi^2), here for a whole batch of models at once, plus the adjoint
gradient of chi^2 with respect to per-cell log density and temperature.

``SamplerStep`` builds the models of a parameter draw (``configs.sampler_batch``:
uneven grids and photon counts), synthetic data for each (its own visibilities
at a reference seed plus Gaussian noise of 1 % of the peak), and runs (images
with the scattering source of ``nphot_scat`` packets, as pdspy's fits call
run_image with nphot=1e5, run_disk_model.py:301-315)

* ``step()``       -- the batched step (one thermal launch for all models,
  optionally replayed from a captured hipGraph), and
* ``sequential()`` -- the same work model by model with the single-model API.
"""
import time

import numpy as np

from . import configs, rt
from .inputs import AU, arcsec, c as c_light, pc

LAM_UM = 1300.0
DPC = 140.0
JY = 1e-23
UPDATE_STREAMS = 4   # streams a new parameter draw's per-model updates run on


class SamplerStep:
    SCAT_KEY = 0x9E3779B97F4A7C15  # scattering key = model seed ^ this (as the batch step)

    def __init__(self, params, npix=256, field_arcsec=2.0, K=5000, seed=0, max_events=1000000,
                 mrw=1, device=0, data_seed=12345, nphot_scat=100000, dust_gen=None,
                 env_dust=None):
        """``dust_gen``: optional ``pdspy_amd.dust.DustGenerator``; then each
        parameter set's ``a_max`` (micron) and ``p`` draw the disk's dust, as
        run_disk_model.py:79-85 does per walker, and ``env_dust`` (default: the
        generator at 1 micron, p = 3.5) is the envelope's."""
        import torch
        self.params = list(params)
        self.dust_gen = dust_gen
        self.env_dust = env_dust if env_dust is not None or dust_gen is None else dust_gen(1e-4, 3.5)
        self.nphot_scat = int(nphot_scat)
        self.n = len(self.params)
        self.npix = npix
        self.K = K
        self.opts = dict(max_events=int(max_events), mrw=int(mrw))
        self.nphot = np.array([int(p["nphot"]) for p in self.params], dtype=np.uint64)
        self.models = [rt.DeviceModel(self._arrays(p), device=device) for p in self.params]
        self.batch = rt.Batch(self.models, device=device)
        pixel = field_arcsec / npix                      # arcsec
        self.pix_cm = pixel * DPC * AU
        self.dxy = pixel * arcsec                        # rad
        # intensity (erg/s/cm^2/Hz/sr) -> Jy per pixel
        self.scale = (pixel * arcsec) ** 2 / JY
        u, v = configs.uv_points(K, lam_m=LAM_UM * 1e-6, seed=seed)
        dev = "cuda"
        self.u = torch.tensor(u, dtype=torch.float64, device=dev)
        self.v = torch.tensor(v, dtype=torch.float64, device=dev)
        # model visibilities as pdspy's fits compute them (interpolate_model
        # with galario, interpolate_model.py:23-27), evaluated exactly: on the
        # grid nodes that is exp(-2 pi i v dxy) DFT(-u, v) of the image, so the
        # DFT runs at (-u, v) with a phase-centre shift of one pixel in Dec
        self.u_vis = (-self.u).contiguous()
        self.ddec = self.dxy
        self.cams = [rt.DeviceModel.camera(npix, self.pix_cm, LAM_UM, incl=p["incl"])
                     for p in self.params]
        f32 = dict(dtype=torch.float32, device=dev)
        self.img = [torch.zeros((npix, npix), **f32) for _ in range(self.n)]
        self.chi2 = [torch.zeros(1, dtype=torch.float64, device=dev) for _ in range(self.n)]
        self.g_logrho = [torch.zeros((m.nspec, m.ncell), **f32) for m in self.models]
        self.g_T = [torch.zeros(m.ncell, **f32) for m in self.models]
        self.w = [torch.ones(K, **f32) for _ in range(self.n)]
        self.re = [torch.zeros(K, **f32) for _ in range(self.n)]
        self.im = [torch.zeros(K, **f32) for _ in range(self.n)]
        # synthetic data: each model's own visibilities at a reference seed + noise
        self._items = None
        self.step(seed=data_seed, use_graph=False)
        rng = np.random.default_rng(data_seed)
        for i in range(self.n):
            re, im = rt.vis_dft(self.img[i] * self.scale, self.dxy, self.u_vis, self.v, 0.0,
                                self.ddec)
            sig = 0.01 * float(torch.sqrt(re * re + im * im).max())
            sig = sig if sig > 0 else 1.0
            self.re[i] = (re + torch.tensor(rng.normal(0, sig, K), **f32)).contiguous()
            self.im[i] = (im + torch.tensor(rng.normal(0, sig, K), **f32)).contiguous()
            self.w[i] = torch.full((K,), 1.0 / sig ** 2, **f32)
        self._items = None

    def autotune(self, frac=0.1):
        """Picks the thermal launch shape (timed on `frac` of each model's packets)
        and the ray tile (timed on the largest model's image + adjoint); returns
        the choices.  The results of a step do not depend on them."""
        import torch
        from . import tune
        nph = np.maximum(1024, (self.nphot * frac).astype(np.uint64))
        seeds = np.arange(self.n, dtype=np.uint64)
        key = ("batch", tuple(int(x) for x in self.nphot),
               tuple((m.nr, m.nt, m.np) for m in self.models))
        best = tune.tune_thermal(key, lambda **c: self.batch.thermal(nph, seeds, **self.opts, **c))
        self.opts.update(best)
        big = max(range(self.n), key=lambda i: self.models[i].ncell)
        m = self.models[big]
        T = torch.ones(m.ncell, dtype=torch.float32, device="cuda") * 20.0
        dimg = torch.ones((self.npix, self.npix), dtype=torch.float32, device="cuda")

        def render(tile):
            cam = rt.DeviceModel.camera(self.npix, self.pix_cm, LAM_UM,
                                        incl=self.params[big]["incl"], tile=tile)
            m.image(cam, T)
            m.image_adjoint(cam, T, None, dimg)

        tile = tune.tune_tiles(("tiles", self.npix, m.nr, m.nt, m.np), render)
        self.cams = [rt.DeviceModel.camera(self.npix, self.pix_cm, LAM_UM, incl=p["incl"],
                                           tile=tile) for p in self.params]
        self._items = None
        return dict(thermal=best, tile=tile)

    def _dust(self, p):
        return [self.dust_gen(p["a_max"] / 1e4, p["p"]), self.env_dust]

    def _set_disk_dust(self, i, kabs_t, ksca_t):
        """Model i's opacities from a disk dust drawn at the generator's
        wavelengths (``DustGenerator.draw_many``) and the envelope dust: the
        arrays ``DeviceModel.set_dust`` would make from the two Dust objects."""
        m = self.models[i]
        if m._keep.get("gsca") is not None:
            raise ValueError("sampler models scatter isotropically (no g)")
        lam = m._keep["lam_um"]
        if not hasattr(self, "_env_rows"):
            self._env_rows = {}
        env = self._env_rows.get(i)
        if env is None:
            ek, es, _, _ = rt.dust_arrays([self.env_dust], lam)
            env = self._env_rows[i] = (np.asarray(ek[0]), np.asarray(es[0]))
        lam_t = self.dust_gen.lam * 1e4
        kabs = np.stack([rt.loglog_resample(lam, lam_t, kabs_t), env[0]])
        ksca = np.stack([rt.loglog_resample(lam, lam_t, ksca_t), env[1]])
        m.set_opacity(kabs, ksca, on_device=True)

    def _arrays(self, p):
        a = rt.model_arrays(configs.batch_model(p).grid)
        if self.dust_gen is not None:
            kabs, ksca, _, _ = rt.dust_arrays(self._dust(p), a["lam_um"])
            a["kabs"], a["ksca"] = np.array(kabs), np.array(ksca)
        return a

    def set_params(self, params):
        """A new parameter draw for the same models (grids and photon counts
        stay): densities are rebuilt on the device (disk + Ulrich envelope
        generator), the disk's dust is drawn again when the step has a dust
        generator, and the cameras take the new inclinations.  Returns the wall
        seconds of the density (and opacity) update."""
        import torch
        params = list(params)
        if len(params) != self.n:
            raise ValueError("expected %d parameter sets" % self.n)
        for p, q in zip(params, self.params):
            if (p["nr"], p["nt"], int(p["nphot"])) != (q["nr"], q["nt"], int(q["nphot"])):
                raise ValueError("set_params keeps each model's grid and photon count")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # the models' updates are small kernels (tens of waves each): they run
        # on UPDATE_STREAMS streams, forked from and joined to the caller's
        main = torch.cuda.current_stream()
        if not hasattr(self, "_upd_streams"):
            self._upd_streams = [torch.cuda.Stream() for _ in range(UPDATE_STREAMS)]
        for st in self._upd_streams:
            st.wait_stream(main)
        if self.dust_gen is not None:
            # the new disk dust of every model in one pass on the host (2 x nlam
            # values per model, which the host keeps: the group check and the
            # camera-wavelength opacities read them); the per-species arrays and
            # thermal tables are then made by kernels on the device
            ka, ks = self.dust_gen.draw_many([p["a_max"] / 1e4 for p in params],
                                             [p["p"] for p in params])
        for i, (m, p) in enumerate(zip(self.models, params)):
            with torch.cuda.stream(self._upd_streams[i % UPDATE_STREAMS]):
                m.generate_density(configs.batch_components(p))
                if self.dust_gen is not None:
                    self._set_disk_dust(i, ka[i], ks[i])
        for st in self._upd_streams:
            main.wait_stream(st)
        self.models[0].ctx.sync_stream()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.params = params
        tile = (self.cams[0].tile_w, self.cams[0].tile_h)
        self.cams = [rt.DeviceModel.camera(self.npix, self.pix_cm, LAM_UM, incl=p["incl"],
                                           tile=tile) for p in self.params]
        self._items = None
        return dt

    def items(self):
        if self._items is None:
            self._items = [rt.Batch.item(self.cams[i], self.dxy, self.u_vis, self.v, self.re[i],
                                         self.im[i], self.w[i], self.img[i], self.chi2[i],
                                         self.g_logrho[i], self.g_T[i], scale=self.scale,
                                         ddec=self.ddec, nphot_scat=self.nphot_scat)
                           for i in range(self.n)]
        return self._items

    def step(self, seed=0, use_graph=True):
        """The batched step; returns (stats per model, device ms)."""
        seeds = np.arange(self.n, dtype=np.uint64) + np.uint64(seed) * np.uint64(1000)
        return self.batch.step(self.items(), self.nphot, seeds, use_graph=use_graph, **self.opts)

    def single_model(self, i, seed):
        """Model i's share of the step through the single-model calls:
        thermal run, imaging temperature, scattering source, image,
        visibilities, chi^2 and the adjoint; returns (image, chi^2,
        d chi^2 / d log rho, d chi^2 / d T)."""
        import torch
        m = self.models[i]
        m.thermal(int(self.nphot[i]), int(seed), **self.opts)
        T = m.temperature_device()
        rho = m.rho_device()
        dens = torch.where(rho == 0, torch.full_like(rho, 1e-30), rho)
        Timg = ((dens * T).sum(0) / dens.sum(0)).contiguous()
        cam = self.cams[i]
        if self.nphot_scat:
            J, _ = m.scattering_source(LAM_UM, self.nphot_scat, Timg,
                                       seed=int(seed) ^ self.SCAT_KEY)
            cam = rt.DeviceModel.camera(self.npix, self.pix_cm, LAM_UM,
                                        incl=self.params[i]["incl"],
                                        tile=(cam.tile_w, cam.tile_h), jscat=J)
        img = m.image(cam, Timg)
        re, im = rt.vis_dft(img * self.scale, self.dxy, self.u_vis, self.v, 0.0, self.ddec)
        c2, gre, gim = rt.chisq(self.re[i], self.im[i], self.w[i], re, im, with_grad=True)
        gimg = rt.vis_dft_adjoint(gre, gim, self.npix, self.npix, self.dxy, self.u_vis,
                                  self.v, 0.0, self.ddec) * self.scale
        glr, gT = m.image_adjoint(cam, Timg, None, gimg)
        return img, c2, glr, gT

    def sequential(self, seed=0, progress=None):
        """The same work, model by model through the single-model calls; returns
        wall seconds (device synchronised) and per-model chi^2.  progress(i, s):
        called after model i with the seconds so far."""
        import torch
        seeds = np.arange(self.n, dtype=np.uint64) + np.uint64(seed) * np.uint64(1000)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chis = []
        for i in range(self.n):
            _, c2, _, _ = self.single_model(i, int(seeds[i]))
            chis.append(c2)
            if progress is not None:
                torch.cuda.synchronize()
                progress(i, time.perf_counter() - t0)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, chis
