"""FP64 CPU oracle for the continuum radiative-transfer path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, as the checker or
the timed CPU baseline.  The product package ``pdspy_amd`` never imports it.

The C source (``rt_oracle.c``) restates what RADMC-3D computes for pdspy's
``Model.run_thermal`` / ``run_image`` (pdspy/modeling/Model.py:221-256,
:515-577) plus the visibility DFT and chi^2 (imaging/imtovis.py:6-27,
interferometry/interpolate_model.py:11-57, interferometry/libinterferometry.pyx:610-633).
``vis_grid`` (numpy, in this file) restates galario's sampling method that
interpolate_model calls: the grid transform, bilinear interpolation and the
phase-centre shift.

Model inputs are plain numpy arrays (a dict or any object with the fields
``rw, tw, pw, rho, lam_um, kabs, ksca, gsca, stars``), never product objects.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PDSPY_ORACLE_LIB selects another build of the same source (the sanitizer
# build of tests/sanitize)
_LIB_PATH = os.environ.get("PDSPY_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


class _Model(C.Structure):
    _fields_ = [("nr", C.c_int), ("nt", C.c_int), ("np", C.c_int),
                ("rw", C.c_void_p), ("tw", C.c_void_p), ("pw", C.c_void_p),
                ("nspec", C.c_int), ("rho", C.c_void_p), ("nlam", C.c_int),
                ("lam_um", C.c_void_p), ("kabs", C.c_void_p), ("ksca", C.c_void_p),
                ("gsca", C.c_void_p), ("nstar", C.c_int), ("star", C.c_void_p)]


class _ThermOpts(C.Structure):
    _fields_ = [("nphot", C.c_uint64), ("seed", C.c_uint64), ("ntemp", C.c_int),
                ("temp0", C.c_double), ("temp1", C.c_double),
                ("max_events", C.c_uint64), ("lagged", C.c_int),
                ("epoch_first", C.c_uint64), ("epoch_growth", C.c_double),
                ("nthreads", C.c_int), ("mrw", C.c_int), ("mrw_tauthres", C.c_double),
                ("mrw_count", C.c_int), ("mrw_gamma", C.c_double), ("lag", C.c_int),
                ("slice_events", C.c_uint64)]


class _ThermStats(C.Structure):
    _fields_ = [("n_cross", C.c_uint64), ("n_abs", C.c_uint64), ("n_scat", C.c_uint64),
                ("n_escape", C.c_uint64), ("n_killed", C.c_uint64),
                ("n_events", C.c_uint64), ("n_mrw", C.c_uint64), ("eps", C.c_double),
                ("lum", C.c_double),
                ("tally_tau_sum", C.c_double), ("n_epochs", C.c_int),
                ("n_nudged", C.c_uint64)]


class _ScatStats(C.Structure):
    _fields_ = [("n_escape", C.c_uint64), ("n_abs", C.c_uint64), ("n_scat", C.c_uint64),
                ("n_cross", C.c_uint64), ("n_killed", C.c_uint64), ("lum_star", C.c_double),
                ("lum_dust", C.c_double), ("eps", C.c_double)]


class _Camera(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nsub", C.c_int),
                ("pix_cm", C.c_double), ("incl", C.c_double), ("posang", C.c_double),
                ("phi", C.c_double), ("lam_um", C.c_double), ("mode", C.c_int),
                ("incl_star", C.c_int), ("tau_stop", C.c_double), ("nrefine", C.c_int),
                ("refine_crit", C.c_double)]


def _setup(L):
    P = C.c_void_p
    L.or_thermal.argtypes = [C.POINTER(_Model), C.POINTER(_ThermOpts), P, P,
                             C.POINTER(_ThermStats)]
    L.or_image.argtypes = [C.POINTER(_Model), P, P, C.POINTER(_Camera), P]
    L.or_image_last_nrays.argtypes = []
    L.or_image_last_nrays.restype = C.c_long
    L.or_image_adjoint.argtypes = [C.POINTER(_Model), P, P, C.POINTER(_Camera), P, P, P]
    L.or_trace_ray.argtypes = [C.POINTER(_Model), P, P, C.c_int, P, P, P]
    L.or_trace_ray_n.argtypes = [C.POINTER(_Model), P, P, C.c_int, P, P, P, P]
    L.or_trace_ray_from.argtypes = [C.POINTER(_Model), P, P, P, C.c_int, P, P, P]
    L.or_scat_source.argtypes = [C.POINTER(_Model), P, P, P, C.c_double, C.c_uint64,
                                 C.c_uint64, C.c_uint64, C.c_int, P, C.POINTER(_ScatStats)]
    L.or_image_scat.argtypes = [C.POINTER(_Model), P, P, P, P, C.POINTER(_Camera), P]
    L.or_image_adjoint_scat.argtypes = [C.POINTER(_Model), P, P, P, P, C.POINTER(_Camera), P,
                                        P, P]
    L.or_vis_dft.argtypes = [P, C.c_int, C.c_int, C.c_double, P, P, C.c_int,
                             C.c_double, C.c_double, P, P]
    L.or_vis_dft_adjoint.argtypes = [P, P, C.c_int, C.c_int, C.c_double, P, P,
                                     C.c_int, C.c_double, C.c_double, P]
    L.or_chisq.argtypes = [P, P, P, P, P, C.c_int]
    L.or_chisq.restype = C.c_double
    L.or_philox4x32_10.argtypes = [P, P, P]
    L.or_planck.argtypes = [C.c_double, C.c_double]
    L.or_planck.restype = C.c_double
    L.or_dplanck_dT.argtypes = [C.c_double, C.c_double]
    L.or_dplanck_dT.restype = C.c_double
    L.or_emis_per_gram.argtypes = [P, P, C.c_int, C.c_double]
    L.or_emis_per_gram.restype = C.c_double
    L.or_freq_grid.argtypes = [P, C.c_int, P, P]
    L.or_tables_create.argtypes = [C.POINTER(_Model), C.c_int, C.c_double, C.c_double]
    L.or_tables_create.restype = C.c_void_p
    L.or_tables_free.argtypes = [P]
    L.or_tables_shape.argtypes = [P, P, P, P]
    L.or_tables_get.argtypes = [P, P, P, P, P, P]
    L.or_reemit_bin.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double]
    L.or_reemit_bin.restype = C.c_int
    L.or_reemit_hist.argtypes = [P, C.c_int, C.c_double, C.c_uint64, C.c_uint64, P]
    L.or_set_mutation.argtypes = [C.c_int]
    L.or_mrw_survival.argtypes = [C.c_double]
    L.or_mrw_survival.restype = C.c_double
    L.or_mrw_invert.argtypes = [C.c_double]
    L.or_mrw_invert.restype = C.c_double


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _get(m, k):
    return m[k] if isinstance(m, dict) else getattr(m, k)


class _Pinned:
    """Holds numpy arrays alive while the C struct points at them."""

    def __init__(self, m):
        self.rw = _f64(_get(m, "rw"))
        self.tw = _f64(_get(m, "tw"))
        self.pw = _f64(_get(m, "pw"))
        rho = _f64(_get(m, "rho"))
        self.nspec = rho.shape[0]
        self.rho = rho
        self.lam = _f64(_get(m, "lam_um"))
        self.kabs = _f64(_get(m, "kabs")).reshape(self.nspec, -1)
        self.ksca = _f64(_get(m, "ksca")).reshape(self.nspec, -1)
        g = _get(m, "gsca")
        self.gsca = None if g is None else _f64(g).reshape(self.nspec, -1)
        self.stars = _f64(_get(m, "stars")).reshape(-1, 5)
        nr, nt, np_ = len(self.rw) - 1, len(self.tw) - 1, len(self.pw) - 1
        assert rho.size == self.nspec * nr * nt * np_, "rho must be [nspec][np][nt][nr]"
        self.s = _Model(nr, nt, np_, self.rw.ctypes.data, self.tw.ctypes.data,
                        self.pw.ctypes.data, self.nspec, self.rho.ctypes.data,
                        len(self.lam), self.lam.ctypes.data, self.kabs.ctypes.data,
                        self.ksca.ctypes.data,
                        None if self.gsca is None else self.gsca.ctypes.data,
                        self.stars.shape[0], self.stars.ctypes.data)
        self.ncell = nr * nt * np_


def thermal(m, nphot, seed=0, ntemp=1000, temp0=0.01, temp1=1e5, max_events=0,
            lagged=True, epoch_first=None, epoch_growth=1.25, nthreads=0, mrw=False,
            mrw_tauthres=10.0, mrw_count=100, mrw_gamma=2.0, lag=2, slice_events=10000):
    """Returns (T[nspec][cells] float64, visits[cells] uint64, stats dict)."""
    pm = _Pinned(m)
    if epoch_first is None:
        epoch_first = max(1024, int(nphot) // 4096)
    o = _ThermOpts(int(nphot), int(seed), ntemp, temp0, temp1, int(max_events),
                   1 if lagged else 0, int(epoch_first), float(epoch_growth), int(nthreads),
                   1 if mrw else 0, float(mrw_tauthres), int(mrw_count), float(mrw_gamma),
                   int(lag), int(slice_events))
    T = np.zeros((pm.nspec, pm.ncell))
    visits = np.zeros(pm.ncell, dtype=np.uint64)
    st = _ThermStats()
    rc = lib().or_thermal(C.byref(pm.s), C.byref(o), T.ctypes.data, visits.ctypes.data,
                          C.byref(st))
    if rc != 0:
        raise RuntimeError("or_thermal failed: %d" % rc)
    stats = {k: getattr(st, k) for k, _ in _ThermStats._fields_}
    return T, visits, stats


def _camera(npix, pix_cm, incl, posang, phi, lam_um, mode, incl_star, nsub, tau_stop,
            ny=None, nrefine=0, refine_crit=1.0):
    return _Camera(int(npix), int(ny if ny is not None else npix), int(nsub), float(pix_cm),
                   float(incl), float(posang), float(phi), float(lam_um), int(mode),
                   int(incl_star), float(tau_stop), int(nrefine), float(refine_crit))


def scat_source(m, T, kap_abs, kap_sca, lam_um, nphot, seed=0, max_events=1000000,
                nthreads=0):
    """Mean intensity J[cells] at lam_um from the monochromatic scattering
    Monte Carlo; returns (J, stats dict)."""
    pm = _Pinned(m)
    T = _f64(T).reshape(-1)
    ka = _f64(kap_abs).reshape(-1)
    ks = _f64(kap_sca).reshape(-1)
    J = np.zeros(pm.ncell)
    st = _ScatStats()
    rc = lib().or_scat_source(C.byref(pm.s), T.ctypes.data, ka.ctypes.data, ks.ctypes.data,
                              float(lam_um), int(nphot), int(seed), int(max_events),
                              int(nthreads), J.ctypes.data, C.byref(st))
    if rc != 0:
        raise RuntimeError("or_scat_source failed: %d" % rc)
    return J, {k: getattr(st, k) for k, _ in _ScatStats._fields_}


def image(m, T, kap, npix, pix_cm, incl=0.0, posang=0.0, phi=0.0, lam_um=1300.0,
          mode=0, incl_star=False, nsub=1, tau_stop=40.0, ksca=None, J=None, nrefine=0,
          refine_crit=1.0, with_nrays=False):
    """Returns img[ny][nx] (erg/s/cm^2/Hz/sr, or tau for mode=1).  With ksca
    and J (from scat_source): isotropic scattering source.  nrefine > 0:
    recursive sub-pixeling (RADMC-3D camera_nrrefine / camera_refine_criterion);
    with_nrays: also the number of rays traced."""
    pm = _Pinned(m)
    T = _f64(T).reshape(-1)
    kap = _f64(kap).reshape(-1)
    cam = _camera(npix, pix_cm, incl, posang, phi, lam_um, mode, incl_star, nsub, tau_stop,
                  nrefine=nrefine, refine_crit=refine_crit)
    img = np.zeros((cam.ny, cam.nx))
    ks = None if ksca is None else _f64(ksca).reshape(-1)
    Jv = None if J is None else _f64(J).reshape(-1)
    rc = lib().or_image_scat(C.byref(pm.s), T.ctypes.data, kap.ctypes.data,
                             None if ks is None else ks.ctypes.data,
                             None if Jv is None else Jv.ctypes.data, C.byref(cam),
                             img.ctypes.data)
    if rc != 0:
        raise RuntimeError("or_image failed: %d" % rc)
    if with_nrays:
        return img, int(lib().or_image_last_nrays())
    return img


def image_adjoint(m, T, kap, dimg, npix, pix_cm, incl=0.0, posang=0.0, phi=0.0,
                  lam_um=1300.0, incl_star=False, nsub=1, tau_stop=40.0, ksca=None, J=None,
                  nrefine=0, refine_crit=1.0):
    pm = _Pinned(m)
    T = _f64(T).reshape(-1)
    kap = _f64(kap).reshape(-1)
    dimg = _f64(dimg)
    cam = _camera(npix, pix_cm, incl, posang, phi, lam_um, 0, incl_star, nsub, tau_stop,
                  nrefine=nrefine, refine_crit=refine_crit)
    g_lr = np.zeros((pm.nspec, pm.ncell))
    g_T = np.zeros(pm.ncell)
    ks = None if ksca is None else _f64(ksca).reshape(-1)
    Jv = None if J is None else _f64(J).reshape(-1)
    rc = lib().or_image_adjoint_scat(C.byref(pm.s), T.ctypes.data, kap.ctypes.data,
                                     None if ks is None else ks.ctypes.data,
                                     None if Jv is None else Jv.ctypes.data,
                                     C.byref(cam), dimg.ctypes.data, g_lr.ctypes.data,
                                     g_T.ctypes.data)
    if rc != 0:
        raise RuntimeError("or_image_adjoint failed: %d" % rc)
    return g_lr, g_T


def trace_ray(m, o, d, maxseg=100000, with_nudges=False):
    """cells, segment lengths and hemispheres along a ray from o along d
    (and, with_nudges, how often the ping-pong guard moved it and the number
    of segments, which the walk caps at 1e7)"""
    pm = _Pinned(m)
    o = _f64(o)
    d = _f64(d)
    cell = np.zeros(maxseg, dtype=np.int32)
    ln = np.zeros(maxseg)
    hemi = np.zeros(maxseg, dtype=np.int32)
    nud = C.c_int(0)
    n = lib().or_trace_ray_n(C.byref(pm.s), o.ctypes.data, d.ctypes.data, maxseg,
                             cell.ctypes.data, ln.ctypes.data, hemi.ctypes.data, C.byref(nud))
    if n < 0:
        raise RuntimeError("or_trace_ray failed: %d" % n)
    m_ = min(n, maxseg)
    if with_nudges:
        return cell[:m_], ln[:m_], hemi[:m_], nud.value, n
    return cell[:m_], ln[:m_], hemi[:m_]


def trace_ray_from(m, o, d, idx, maxseg=100000):
    """cells, segment lengths and hemispheres along a ray from o (inside the
    grid) along d, starting in the cell idx = (ir, it, ip, hemi) as given
    rather than the one o lies in"""
    pm = _Pinned(m)
    o = _f64(o)
    d = _f64(d)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    cell = np.zeros(maxseg, dtype=np.int32)
    ln = np.zeros(maxseg)
    hemi = np.zeros(maxseg, dtype=np.int32)
    n = lib().or_trace_ray_from(C.byref(pm.s), o.ctypes.data, d.ctypes.data, ix.ctypes.data,
                                maxseg, cell.ctypes.data, ln.ctypes.data, hemi.ctypes.data)
    if n < 0:
        raise RuntimeError("or_trace_ray_from failed: %d" % n)
    m_ = min(n, maxseg)
    return cell[:m_], ln[:m_], hemi[:m_]


def vis_dft(img, dxy, u, v, dra=0.0, ddec=0.0):
    img = _f64(img)
    u = _f64(u)
    v = _f64(v)
    re = np.zeros(len(u))
    im = np.zeros(len(u))
    ny, nx = img.shape
    lib().or_vis_dft(img.ctypes.data, nx, ny, float(dxy), u.ctypes.data, v.ctypes.data,
                     len(u), float(dra), float(ddec), re.ctypes.data, im.ctypes.data)
    return re, im


def vis_dft_adjoint(gre, gim, nx, ny, dxy, u, v, dra=0.0, ddec=0.0):
    gre = _f64(gre)
    gim = _f64(gim)
    u = _f64(u)
    v = _f64(v)
    g = np.zeros((ny, nx))
    lib().or_vis_dft_adjoint(gre.ctypes.data, gim.ctypes.data, nx, ny, float(dxy),
                             u.ctypes.data, v.ctypes.data, len(u), float(dra),
                             float(ddec), g.ctypes.data)
    return g


def vis_grid(img, dxy, u, v, dra=0.0, ddec=0.0):
    """galario's sampling method in FP64 numpy: the image's transform on its
    pixel-conjugate grid (imtovis's fftshift(fft2(ifftshift(I))), imaging/
    imtovis.py:13-24), bilinear interpolation at (u, v), phase-centre shift
    (interferometry/interpolate_model.py:20-27 calls galario's sampleImage).
    Pinned in the tests by the exact DFT at grid nodes and by scipy's
    RegularGridInterpolator between them."""
    img = _f64(img)
    u = _f64(u)
    v = _f64(v)
    ny, nx = img.shape
    F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img)))
    fa = u * nx * dxy + nx // 2
    fb = v * ny * dxy + ny // 2
    a0 = np.floor(fa)
    b0 = np.floor(fb)
    ok = (a0 >= 0) & (a0 <= nx - 2) & (b0 >= 0) & (b0 <= ny - 2)
    a = np.where(ok, a0, 0).astype(np.int64)
    b = np.where(ok, b0, 0).astype(np.int64)
    ta = fa - a0
    tb = fb - b0
    V = ((1 - ta) * (1 - tb) * F[b, a] + ta * (1 - tb) * F[b, a + 1] +
         (1 - ta) * tb * F[b + 1, a] + ta * tb * F[b + 1, a + 1])
    V = V * np.exp(-2j * np.pi * (u * dra + v * ddec))
    return np.where(ok, V.real, np.nan), np.where(ok, V.imag, np.nan)


class Tables:
    """The thermal Monte Carlo's tables for a model (or_tables_create): ``lnq``
    [ngroup][ntemp] (ln of the emission per gram), ``kplanck`` [ngroup][ntemp],
    ``kross`` [nspec][ntemp], ``cdf`` [ngroup][ntemp][nlam] (re-emission),
    ``grp`` [nspec]; ``temps`` the temperature grid.  ``reemit(q, T, u_row,
    u_freq)`` is the thermal MC's re-emission draw, ``hist(q, T, n, seed)``
    the histogram of 2n such draws."""

    def __init__(self, m, ntemp=1000, temp0=0.01, temp1=1e5):
        pm = _Pinned(m)
        h = lib().or_tables_create(C.byref(pm.s), int(ntemp), float(temp0), float(temp1))
        if not h:
            raise RuntimeError("or_tables_create failed")
        self.h = h
        ng, nt, nl = C.c_int(), C.c_int(), C.c_int()
        lib().or_tables_shape(h, C.byref(ng), C.byref(nt), C.byref(nl))
        self.ngroup, self.ntemp, self.nlam = ng.value, nt.value, nl.value
        self.lnq = np.zeros((self.ngroup, self.ntemp))
        self.kplanck = np.zeros((self.ngroup, self.ntemp))
        self.kross = np.zeros((pm.nspec, self.ntemp))
        self.cdf = np.zeros((self.ngroup, self.ntemp, self.nlam))
        self.grp = np.zeros(pm.nspec, dtype=np.int32)
        lib().or_tables_get(h, self.lnq.ctypes.data, self.kplanck.ctypes.data,
                            self.kross.ctypes.data, self.cdf.ctypes.data, self.grp.ctypes.data)
        self.temps = np.exp(np.linspace(np.log(temp0), np.log(temp1), ntemp))
        self.lnt0 = np.log(temp0)
        self.dlnt = (np.log(temp1) - np.log(temp0)) / (ntemp - 1)

    def reemit(self, q, T, u_row, u_freq):
        return lib().or_reemit_bin(self.h, int(q), float(T), float(u_row), float(u_freq))

    def hist(self, q, T, n, seed=0):
        out = np.zeros(self.nlam, dtype=np.uint64)
        lib().or_reemit_hist(self.h, int(q), float(T), int(n), int(seed), out.ctypes.data)
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().or_tables_free(self.h)
                self.h = None
        except Exception:
            pass


def set_mutation(code):
    """Test only: 0 = none, 1 = re-emission spectrum kappa B instead of
    kappa dB/dT, 2 = temperature row shifted by one (applies to tables built
    and draws made afterwards), 3 = ping-pong guard off, 4 = a wall crossing
    found just behind the position is dropped instead of taken at once."""
    lib().or_set_mutation(int(code))


def galario_sample_image(image, dxy, u, v, dRA=0.0, dDec=0.0, origin="upper"):
    """galario's sampleImage (Tazzari, Beaujean & Testi 2018, MNRAS 476, 4527),
    restated in FP64 numpy from its published interface and method: ``image``
    [N, N] (Jy/pixel), ``dxy`` the pixel size (rad), (u, v) in wavelengths,
    dRA / dDec (rad) translate the image towards the East (left, decreasing
    column index) / North (up), ``origin`` says whether row 0 is the top
    ('upper', galario's default) or the bottom ('lower') of the image.
    Method: real-to-complex transform of fftshift(image) (phase origin at pixel
    (N/2, N/2)), fftshift along the row axis, bilinear interpolation at
    (|u|/du, N/2 + s v_o v/du) with du = 1/(N dxy), s = sign(u), v_o = +1
    ('upper') or -1 ('lower'), complex-conjugated where u < 0 (the stored half
    plane), then the factor exp(+2 pi i (u dRA + v dDec)).  pdspy calls it at
    interferometry/interpolate_model.py:23-24.  Returns complex V."""
    img = _f64(image)
    u = _f64(u).reshape(-1)
    v = _f64(v).reshape(-1)
    n = img.shape[0]
    if img.shape != (n, n):
        raise ValueError("galario_sample_image: square image expected")
    vo = {"upper": 1.0, "lower": -1.0}[origin]
    F = np.fft.fftshift(np.fft.rfft2(np.fft.fftshift(img)), axes=0)   # [n, n/2 + 1]
    du = 1.0 / (n * dxy)
    neg = u < 0
    fu = np.abs(u) / du
    fv = n // 2 + np.where(neg, -1.0, 1.0) * vo * v / du
    a0 = np.floor(fu)
    b0 = np.floor(fv)
    ok = (a0 >= 0) & (a0 <= F.shape[1] - 2) & (b0 >= 0) & (b0 <= n - 2)
    a = np.where(ok, a0, 0).astype(np.int64)
    b = np.where(ok, b0, 0).astype(np.int64)
    ta = fu - a0
    tb = fv - b0
    V = ((1 - ta) * (1 - tb) * F[b, a] + ta * (1 - tb) * F[b, a + 1] +
         (1 - ta) * tb * F[b + 1, a] + ta * tb * F[b + 1, a + 1])
    V = np.where(neg, np.conj(V), V)
    V = V * np.exp(2j * np.pi * (u * dRA + v * dDec))
    return np.where(ok, V, np.nan + 1j * np.nan)


def pdspy_interpolate_galario(image, dxy, u, v, dRA=0.0, dDec=0.0):
    """What pdspy's interpolate_model(code="galario") computes for one channel
    (interferometry/interpolate_model.py:20-27): galario's sampleImage of the
    row-flipped image (pdspy's row 0 is the bottom of the image, galario's
    default origin is the top) and the imaginary part negated.  ``image``
    [ny, nx] in pdspy's orientation; dRA, dDec in rad.  Returns (re, im)."""
    V = galario_sample_image(_f64(image)[::-1], dxy, u, v, dRA, dDec)
    return V.real, -V.imag


def chisq(dre, dim, w, mre, mim):
    a = [_f64(x) for x in (dre, dim, w, mre, mim)]
    return lib().or_chisq(*[x.ctypes.data for x in a], len(a[0]))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def planck(nu, T):
    return lib().or_planck(float(nu), float(T))


def dplanck_dT(nu, T):
    return lib().or_dplanck_dT(float(nu), float(T))


def emis_per_gram(kabs, lam_um, T):
    k = _f64(kabs)
    lam = _f64(lam_um)
    return lib().or_emis_per_gram(k.ctypes.data, lam.ctypes.data, len(lam), float(T))


def freq_grid(lam_um):
    lam = _f64(lam_um)
    nu = np.zeros(len(lam))
    dnu = np.zeros(len(lam))
    lib().or_freq_grid(lam.ctypes.data, len(lam), nu.ctypes.data, dnu.ctypes.data)
    return nu, dnu


def mrw_survival(y):
    return lib().or_mrw_survival(float(y))


def mrw_invert(xi):
    return lib().or_mrw_invert(float(xi))
