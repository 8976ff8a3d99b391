"""CPU tests of the host side: input builders (pdspy restatements), layout
conversions, keyword handling, and the C ABI surface of the built library."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.optimize import brenth

from pdspy_amd import configs, inputs, rt
from pdspy_amd._lib import EXPORTS, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_star_radius_pdspy_formula():
    # Star.py:12-13; the 4000 K, L = 0.906329 L_sun star has R = 2 R_sun
    s = inputs.Star(luminosity=0.906329, temperature=4000.0)
    assert s.radius == pytest.approx(2.0, rel=1e-5)


def test_wall_count_convention():
    m = configs.disk_c1()
    assert len(m.grid.w1) == 65 and len(m.grid.w2) == 65 and len(m.grid.w3) == 2
    assert m.grid.density[0].shape == (64, 64, 1)
    assert m.grid.w2[-1] == pytest.approx(np.pi / 2)


def _grid_mass(m, i):
    g = m.grid
    r, t, p = g.w1 * inputs.AU, g.w2, g.w3
    vr = np.diff(r ** 3) / 3
    vt = -np.diff(np.cos(t))
    vp = np.diff(p)
    V = vr[:, None, None] * vt[None, :, None] * vp[None, None, :] * 2  # mirror
    return np.sum(g.density[i] * V)


def test_disk_mass_normalisation():
    m = configs.disk_envelope(nr=256, nt=256, envelope=False)
    assert _grid_mass(m, 0) / inputs.M_sun == pytest.approx(1e-3, rel=0.03)


def test_ulrich_envelope_mass_and_roots():
    m = configs.disk_envelope(nr=128, nt=64)
    # UlrichEnvelopeExtended normalises with a trapezoid rule over cell centres
    # and 4 pi over the half theta grid (:100-106); the cell-volume sum agrees to
    # the discretisation of that rule
    g = m.grid
    rr, tt = np.meshgrid(g.r * inputs.AU, g.theta, indexing="ij")
    rho = g.density[1][:, :, 0]
    trap = 4 * inputs.pi_pdspy * np.trapezoid(np.trapezoid(rho * rr ** 2 * np.sin(tt), tt, axis=1),
                                              rr[:, 0])
    assert trap / inputs.M_sun == pytest.approx(0.1, rel=1e-10)
    assert _grid_mass(m, 1) / inputs.M_sun == pytest.approx(0.1, rel=0.12)
    # outside the centrifugal radius the density follows Ulrich's closed form
    # rho ~ (r/Rc)^-1.5 (1 + mu/mu0)^-0.5 (mu/mu0 + 2 mu0^2 Rc/r)^-1 with the
    # single real (Cardano) root mu0; compare shapes (normalisation cancels)
    def ulrich(r, mu, rc=50.0):
        p, q = -(1 - r / rc), -mu * r / rc
        D = (q / 2) ** 2 + (p / 3) ** 3
        mu0 = np.cbrt(-q / 2 + np.sqrt(D)) + np.cbrt(-q / 2 - np.sqrt(D))
        return (r / rc) ** -1.5 * (1 + mu / mu0) ** -0.5 * (mu / mu0 + 2 * mu0 ** 2 * rc / r) ** -1
    it = 20
    mu = np.cos(g.theta[it])
    ref = np.searchsorted(g.r, 200.0)
    for r_au in (400.0, 1000.0, 3000.0):
        i = np.searchsorted(g.r, r_au)
        got = g.density[1][i, it, 0] / g.density[1][ref, it, 0]
        want = ulrich(g.r[i], mu) / ulrich(g.r[ref], mu)
        assert got == pytest.approx(want, rel=1e-9)
    # mu0 solves mu0^3 - mu0 (1 - r/Rc) - mu r/Rc = 0 (UlrichEnvelopeExtended.py:68-75);
    # for r > Rc the cubic has one real root (Cardano)
    rc = 50.0
    for r, mu in [(200.0, 0.3), (1000.0, 0.9), (75.0, 0.01)]:
        root = brenth(lambda x: x ** 3 - x * (1 - r / rc) - mu * r / rc, -1, 1)
        p, q = -(1 - r / rc), -mu * r / rc
        D = (q / 2) ** 2 + (p / 3) ** 3
        card = np.cbrt(-q / 2 + np.sqrt(D)) + np.cbrt(-q / 2 - np.sqrt(D))
        assert root == pytest.approx(card, abs=1e-10)


def test_model_arrays_layout():
    m = configs.disk_c1(nr=8, nt=6)
    a = rt.model_arrays(m.grid)
    assert a["rho"].shape == (1, 1, 6, 8)
    # r fastest: rho[s][ip][it][ir] == density[s][ir, it, ip] (write.py:261-264)
    assert a["rho"][0, 0, 2, 5] == m.grid.density[0][5, 2, 0]
    assert np.allclose(a["rw"], m.grid.w1 * inputs.AU)
    assert a["stars"][0, 0] == pytest.approx(2.0 * inputs.R_sun, rel=1e-5)
    assert np.allclose(a["kabs"][0], m.grid.dust[0].kabs)


def test_imaging_temperature_is_density_weighted_mean():
    # Model.py:759-769
    m = configs.disk_envelope(nr=16, nt=8)
    rng = np.random.default_rng(0)
    m.grid.temperature = [rng.uniform(10, 100, d.shape) for d in m.grid.density]
    T = rt.imaging_temperature(m.grid)
    d = [np.where(x == 0, 1e-30, x) for x in m.grid.density]
    want = (d[0] * m.grid.temperature[0] + d[1] * m.grid.temperature[1]) / (d[0] + d[1])
    assert np.allclose(T, want.transpose(2, 1, 0).ravel())


def test_keywords_follow_radmc3d_option_set():
    rt._check_keywords({"modified_random_walk": True, "setthreads": 4, "verbose": False,
                        "mrw_gamma": 2, "timelimit": 10})
    with pytest.raises(TypeError):
        rt._check_keywords({"not_a_radmc3d_key": 1})


def test_fft_grid_uv_matches_imtovis():
    u, v = rt.fft_grid_uv(8, 0.01)
    uu = np.fft.fftshift(np.fft.fftfreq(8, 0.01 * inputs.arcsec))
    U, V = np.meshgrid(uu, uu)
    assert np.array_equal(u, U.ravel()) and np.array_equal(v, V.ravel())


def test_loglog_opacity_interpolation():
    lam = np.logspace(-1, 4, 50)
    k = 1e3 * lam ** -1.0
    lam_new = np.array([0.05, 3.3, 1300.0, 2e4])
    out = rt._loglog_interp(lam_new, lam, k)
    assert np.allclose(out, 1e3 * lam_new ** -1.0, rtol=1e-10)


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "pdspy_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char \*)\s*(rt_\w+)\(", txt, re.M)))


def test_header_and_binding_agree():
    assert _header_symbols() == sorted(EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name in _header_symbols():
        assert hasattr(lib, name), name


def test_product_never_imports_oracle():
    pkg = os.path.join(ROOT, "pdspy_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "from oracle" not in src, f
                assert "liboracle" not in src, f


def test_interpolate_model_rejects_other_codes_before_any_device_work():
    from types import SimpleNamespace
    im = SimpleNamespace(x=np.arange(4.0), y=np.arange(4.0), image=np.zeros((4, 4, 1, 1)))
    for code in ("trift", "galario-unstructured"):
        with pytest.raises(ValueError, match="not supported"):
            rt.interpolate_model(np.zeros(3), np.zeros(3), np.array([1e11]), im, code=code)


def test_interpolate_model_signature_is_pdspys():
    """Same parameters in the same order as interpolate_model.py:11-12, so
    positional and keyword calls written for pdspy carry over."""
    import inspect
    names = list(inspect.signature(rt.interpolate_model).parameters)
    assert names[:10] == ["u", "v", "freq", "model", "nthreads", "dRA", "dDec", "code", "nxy",
                          "dxy"]


def test_line_options_accepted_at_defaults_only():
    """run_image's line-transfer arguments (Model.py:515-519) pass at pdspy's
    defaults and are refused when a line is asked for."""
    kw = dict(imolspec=None, iline=None, widthkms=None, vkms=None, linenlam=None,
              doppcatch=False, unstructured=False, verbose=True)
    rt._pop_line_options(kw)
    assert kw == dict(verbose=True)
    for k, v in (("imolspec", 1), ("doppcatch", True), ("unstructured", True)):
        with pytest.raises(NotImplementedError):
            rt._pop_line_options({k: v})


def test_dust_arrays_match_model_arrays_resampling():
    """DeviceModel.set_dust resamples a Dust as model_arrays does at creation."""
    m = configs.disk_envelope(nr=8, nt=6)
    a = rt.model_arrays(m.grid)
    kabs, ksca, gg, has_g = rt.dust_arrays(m.grid.dust, a["lam_um"])
    assert np.array_equal(np.array(kabs), a["kabs"]) and np.array_equal(np.array(ksca), a["ksca"])
    assert not has_g and a["gsca"] is None
    # a Dust on its own coarser grid is log-log interpolated onto the model's
    gen = configs.synthetic_dust_generator(np.logspace(-1, 4, 40))
    d = gen(1e-2, 3.2)
    k = rt.dust_arrays([d], a["lam_um"])[0][0]
    sel = np.isin(np.round(np.log10(a["lam_um"]), 9), np.round(np.log10(d.lam * 1e4), 9))
    assert sel.sum() >= 2
    assert np.allclose(k[sel], d.kabs[np.isin(np.round(np.log10(d.lam * 1e4), 9),
                                              np.round(np.log10(a["lam_um"]), 9))], rtol=1e-10)


def test_dust_with_g_scatters_isotropically_unless_asked():
    """pdspy writes dustkappa files with kabs and ksca only (Model.py:771-775),
    so a Dust carrying g (pdspy's PAH class sets one) scatters isotropically in
    RADMC-3D; model_arrays does the same unless hg=True."""
    m = configs.spiral_3d(nr=6, nt=4, nphi=4)
    assert getattr(m.grid.dust[0], "g", None) is not None
    assert rt.model_arrays(m.grid)["gsca"] is None
    g = rt.model_arrays(m.grid, hg=True)["gsca"]
    assert g is not None and np.allclose(g, 0.6)


def test_device_model_cache_key_follows_every_input():
    """The resident device model is reused only while walls, densities (also
    edited in place), dust and stars are unchanged."""
    m = configs.disk_envelope(nr=8, nt=6)
    d0 = rt._arrays_digest(rt.model_arrays(m.grid))
    assert rt._arrays_digest(rt.model_arrays(m.grid)) == d0
    m.grid.density[0][3, 2, 0] *= 1.5                 # in place
    d1 = rt._arrays_digest(rt.model_arrays(m.grid))
    assert d1 != d0
    m.grid.dust[1].kabs = m.grid.dust[1].kabs * 1.01  # another dust
    d2 = rt._arrays_digest(rt.model_arrays(m.grid))
    assert d2 != d1
    m.grid.stars[0].temperature = 4100.0
    assert rt._arrays_digest(rt.model_arrays(m.grid)) != d2


def test_run_thermal_time_limit_default_is_pdspys():
    import inspect
    assert inspect.signature(rt.run_thermal).parameters["timelimit"].default == 7200
