"""Pins of the FP64 CPU oracle against known answers and closed forms.

RADMC-3D (the reference's engine for this path) is absent, and the reference's
tests hold no golden vectors for it, so every oracle function is checked here
against something the mathematics fixes: published known-answer vectors,
closed-form radiative-transfer solutions, brute-force geometry, library FFTs,
and internal consistency of the Monte Carlo estimators.
"""
import numpy as np
import pytest

import oracle
from pdspy_amd.inputs import AU
from tests.helpers import dT_stats, oracle_arrays
from pdspy_amd import configs

# pdspy CGS constants (pdspy/constants/physics.py)
H, C_LIGHT, K_B = 6.6260755e-27, 2.99792458e10, 1.380658e-16


def np_planck(nu, T):
    return 2 * H * nu ** 3 / C_LIGHT ** 2 / np.expm1(H * nu / (K_B * T))


# ------------------------------------------------------------------ Philox
def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10 (Salmon et al. 2011)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle.philox(ctr, key)) == want


# ------------------------------------------------------------------ Planck
def test_planck_matches_pdspy_formula_and_derivative():
    for lam_um in (0.5, 10.0, 1300.0):
        nu = C_LIGHT / (lam_um * 1e-4)
        for T in (20.0, 300.0, 4000.0):
            assert oracle.planck(nu, T) == pytest.approx(np_planck(nu, T), rel=1e-12)
            h = 1e-4 * T
            fd = (oracle.planck(nu, T + h) - oracle.planck(nu, T - h)) / (2 * h)
            assert oracle.dplanck_dT(nu, T) == pytest.approx(fd, rel=1e-6)


def test_grey_emission_per_gram_is_4_kappa_sigma_T4():
    lam = np.logspace(-2, 5, 3000)
    sigma = 2 * np.pi ** 5 * K_B ** 4 / (15 * C_LIGHT ** 2 * H ** 3)
    for T in (10.0, 100.0, 1500.0):
        q = oracle.emis_per_gram(np.full(lam.size, 2.0), lam, T)
        assert q == pytest.approx(4 * 2.0 * sigma * T ** 4, rel=2e-4)


def test_frequency_grid_trapezoid_weights():
    lam = np.array([1.0, 2.0, 4.0, 8.0])
    nu, dnu = oracle.freq_grid(lam)
    assert np.allclose(nu, C_LIGHT / (lam * 1e-4))
    assert dnu[0] == pytest.approx(0.5 * (nu[0] - nu[1]))
    assert dnu[1] == pytest.approx(0.5 * (nu[0] - nu[2]))
    assert dnu.sum() == pytest.approx(nu[0] - nu[-1])


# ------------------------------------------------------------------ geometry
def _small_model(nr=5, nt=4, nphi=1, mirror=True, rin=1.0, rout=10.0):
    rw = np.logspace(np.log10(rin), np.log10(rout), nr + 1) * AU
    tw = np.linspace(0, np.pi / 2 if mirror else np.pi, nt + 1)
    pw = np.linspace(0, 2 * np.pi, nphi + 1)
    lam = np.array([1.0, 10.0])
    return dict(rw=rw, tw=tw, pw=pw, rho=np.ones((1, nphi, nt, nr)) * 1e-20, lam_um=lam,
                kabs=np.ones((1, 2)), ksca=np.zeros((1, 2)), gsca=None,
                stars=np.array([[7e10, 4000.0, 0, 0, 0]]))


def _locate(m, p):
    r = np.linalg.norm(p)
    rw, tw, pw = m["rw"], m["tw"], m["pw"]
    nr, nt, np_ = len(rw) - 1, len(tw) - 1, len(pw) - 1
    if r < rw[0]:
        return -1
    if r >= rw[-1]:
        return None
    ir = np.searchsorted(rw, r, side="right") - 1
    z = abs(p[2]) if abs(tw[-1] - np.pi / 2) < 1e-9 else p[2]
    th = np.arctan2(np.hypot(p[0], p[1]), z)
    it = min(np.searchsorted(tw, th, side="right") - 1, nt - 1)
    ph = np.arctan2(p[1], p[0]) % (2 * np.pi)
    ip = min(np.searchsorted(pw, ph, side="right") - 1, np_ - 1)
    return ir + nr * (it + nt * ip)


@pytest.mark.parametrize("nphi,mirror", [(1, True), (3, True), (1, False), (4, False)])
def test_trace_ray_matches_brute_force(nphi, mirror):
    m = _small_model(nphi=nphi, mirror=mirror)
    rng = np.random.default_rng(1)
    R = m["rw"][-1]
    for _ in range(40):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o = rng.uniform(-0.9, 0.9, 3) * R / np.sqrt(3)
        o = o - d * 3 * R  # start outside, aim near the centre
        cell, ln, _ = oracle.trace_ray(m, o, d)
        if len(cell) == 0:
            continue
        # total length equals the chord through the outer sphere
        b = o @ d
        chord = 2 * np.sqrt(b * b - (o @ o - R * R))
        assert ln.sum() == pytest.approx(chord, rel=1e-9)
        # every segment's interior points lie in the reported cell
        t0 = -b - np.sqrt(b * b - (o @ o - R * R))
        t = t0
        for c, l in zip(cell, ln):
            for f in (0.05, 0.5, 0.95):
                if l > 1e-6 * R:
                    assert _locate(m, o + d * (t + f * l)) == c
            t += l


def test_trace_ray_through_origin_along_a_wall():
    # the central pixel's ray is radial; on a theta wall every point of it lies
    # in the cone: the traversal must neither cross it nor stall on it
    m = _small_model(nphi=1, mirror=True)
    R = m["rw"][-1]
    for th in m["tw"][1:-1]:
        d = np.array([0.0, -np.sin(th), -np.cos(th)])
        o = -d * 3 * R
        cell, ln, _ = oracle.trace_ray(m, o, d)
        assert ln.sum() == pytest.approx(2 * R, rel=1e-9)
        assert len(cell) <= 4 * len(m["rw"])


def _past_wall_cases():
    """(model, position a hair past a wall, direction through it, cell index
    one cell behind): a theta cone (mirror grid), a phi half-plane (3D grid)
    and the equatorial plane (grid over [0, pi])"""
    eps = 1e-11
    m = _small_model(nt=8)
    r = np.sqrt(m["rw"][2] * m["rw"][3])
    th = m["tw"][4] + eps
    e_r = np.array([np.sin(th), 0.0, np.cos(th)])
    e_th = np.array([np.cos(th), 0.0, -np.sin(th)])
    cone = (m, r * e_r, 0.8 * e_th + 0.6 * e_r, (2, 3, 0, 1))
    m = _small_model(nphi=6)
    th, ph = 0.3, m["pw"][2] + eps
    e_r = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    e_ph = np.array([-np.sin(ph), np.cos(ph), 0.0])
    phi = (m, r * e_r, 0.8 * e_ph + 0.6 * e_r, (2, 0, 1, 1))
    m = _small_model(mirror=False)
    plane = (m, np.array([r, 0.0, -eps * r]), np.array([0.6, 0.0, -0.8]), (2, 1, 0, 1))
    return [cone, phi, plane]


@pytest.mark.parametrize("case", [0, 1, 2], ids=["cone", "phi", "plane"])
def test_wall_just_behind_the_position_is_crossed_at_once(case):
    """Two crossings closer than the arithmetic resolves can leave a ray a
    hair past a wall while its cell index still names the cell before it.
    The walk then crosses that wall at once (a zero-length segment) and
    continues exactly as a walk started from the cell the position lies in;
    with the crossing planted to be dropped instead (the rule before), the
    index stays behind and the cells walked differ."""
    m, o, d, idx = _past_wall_cases()[case]
    d = d / np.linalg.norm(d)
    nr, nt = len(m["rw"]) - 1, len(m["tw"]) - 1
    c_ref, l_ref, h_ref = oracle.trace_ray(m, o, d)
    assert c_ref[0] == _locate(m, o) != idx[0] + nr * (idx[1] + nt * idx[2])
    c, ln, h = oracle.trace_ray_from(m, o, d, idx)
    assert c[0] == idx[0] + nr * (idx[1] + nt * idx[2]) and ln[0] == 0.0
    assert np.array_equal(c[1:], c_ref) and np.array_equal(h[1:], h_ref)
    assert np.allclose(ln[1:], l_ref, rtol=1e-12, atol=1e-9 * m["rw"][-1])
    try:
        oracle.set_mutation(4)
        c4, _, _ = oracle.trace_ray_from(m, o, d, idx)
    finally:
        oracle.set_mutation(0)
    assert len(c4) != len(c_ref) + 1 or not np.array_equal(c4[1:], c_ref)


# ------------------------------------------------------------------ thermal MC
def _thin_shell_model(nr=4, nt=4, rho=1e-24, kappa=10.0):
    lam = np.logspace(-1, 4, 100)
    rw = np.logspace(np.log10(1.0), np.log10(50.0), nr + 1) * AU
    return dict(rw=rw, tw=np.linspace(0, np.pi / 2, nt + 1), pw=np.array([0, 2 * np.pi]),
                rho=np.full((1, 1, nt, nr), rho), lam_um=lam, kabs=np.full((1, 100), kappa),
                ksca=np.zeros((1, 100)), gsca=None,
                stars=np.array([[1.392e11, 4000.0, 0.0, 0.0, 0.0]]))


def _star_lum(lam_um, R, T):
    nu = C_LIGHT / (lam_um * 1e-4)
    dnu = np.empty_like(nu)
    dnu[1:-1] = 0.5 * np.abs(nu[2:] - nu[:-2])
    dnu[0] = 0.5 * abs(nu[1] - nu[0])
    dnu[-1] = 0.5 * abs(nu[-1] - nu[-2])
    return 4 * np.pi * R ** 2 * np.pi * np.sum(np_planck(nu, T) * dnu)


def test_thermal_optically_thin_closed_form():
    """tau << 1 grey absorber around a point star: every packet crosses each shell
    radially, so the absorbed power per gram in shell [r1, r2] is
    L kappa (r2 - r1) / (4/3 pi (r2^3 - r1^3)); T solves q(T) = that."""
    m = _thin_shell_model()
    T, visits, st = oracle.thermal(m, 200000, seed=3, lagged=True, nthreads=4)
    L = _star_lum(m["lam_um"], 1.392e11, 4000.0)
    assert st["lum"] == pytest.approx(L, rel=1e-12)
    rw = m["rw"]
    for i in range(4):
        e = L * 10.0 * (rw[i + 1] - rw[i]) / (4.0 / 3.0 * np.pi * (rw[i + 1] ** 3 - rw[i] ** 3))
        # grey: q(T) = 4 pi kappa sum B dnu -> find T by bisection with the oracle's q
        lo, hi = 1.0, 3000.0
        for _ in range(80):
            mid = np.sqrt(lo * hi)
            if oracle.emis_per_gram(np.full(100, 10.0), m["lam_um"], mid) < e:
                lo = mid
            else:
                hi = mid
        T_exp = lo
        Tc = T[0].reshape(4, 4)[:, i]  # [theta, r]
        assert np.all(np.abs(Tc / T_exp - 1) < 6e-3), (i, Tc, T_exp)


def test_path_length_tally_matches_absorption_count():
    m = oracle_arrays(configs.disk_c1(nr=24, nt=24))
    T, vis, st = oracle.thermal(m, 20000, seed=1, max_events=100000, lagged=True, nthreads=8)
    n = st["n_abs"]
    assert abs(st["tally_tau_sum"] - n) < 5 * np.sqrt(n), (st["tally_tau_sum"], n)
    assert st["n_escape"] + st["n_killed"] == 20000


def test_lagged_oracle_is_thread_count_independent():
    m = oracle_arrays(configs.disk_c1(nr=16, nt=16))
    T1, v1, s1 = oracle.thermal(m, 20000, seed=5, max_events=20000, lagged=True, nthreads=1)
    T8, v8, s8 = oracle.thermal(m, 20000, seed=5, max_events=20000, lagged=True, nthreads=8)
    assert np.array_equal(T1, T8)
    assert np.array_equal(v1, v8)
    assert s1["n_events"] == s8["n_events"]


def test_epoch_lag_bias_below_mc_noise():
    """Epoch-lagged temperatures (what the GPU computes) against true immediate
    Bjorkman & Wood re-emission: the difference must not exceed the seed-to-seed
    Monte Carlo scatter of the immediate method itself."""
    m = oracle_arrays(configs.disk_c1(nr=24, nt=24))
    n = 60000
    Ti0, vi0, _ = oracle.thermal(m, n, seed=11, max_events=20000, lagged=False)
    Ti1, vi1, _ = oracle.thermal(m, n, seed=12, max_events=20000, lagged=False)
    Tl, vl, _ = oracle.thermal(m, n, seed=11, max_events=20000, lagged=True, nthreads=8)
    vis = np.minimum(vi0, vi1)
    noise = dT_stats(Ti0, Ti1, vis)
    lag = dT_stats(Ti0, Tl, vis)
    assert lag[0] <= 1.5 * noise[0] + 1e-3, (lag, noise)
    assert lag[1] <= 1.5 * noise[1] + 5e-3, (lag, noise)


# ------------------------------------------------------------------ camera
def _uniform_sphere(rho=1e-18, kappa=1.0, T=50.0, nr=6, nt=8, nphi=1, rin=2.0, rout=40.0):
    rw = np.logspace(np.log10(rin), np.log10(rout), nr + 1) * AU
    tw = np.linspace(0, np.pi / 2, nt + 1)
    pw = np.linspace(0, 2 * np.pi, nphi + 1)
    m = dict(rw=rw, tw=tw, pw=pw, rho=np.full((1, nphi, nt, nr), rho), lam_um=np.array([100., 2000.]),
             kabs=np.full((1, 2), kappa), ksca=np.zeros((1, 2)), gsca=None,
             stars=np.array([[1.392e11, 4000.0, 0, 0, 0]]))
    return m, np.full(nr * nt * nphi, T)


@pytest.mark.parametrize("incl", [0.0, 37.0, 90.0])
def test_image_uniform_shell_closed_form(incl):
    m, T = _uniform_sphere()
    rin, rout = m["rw"][0], m["rw"][-1]
    npix, pix = 33, 2.6 * AU
    lam = 1300.0
    img = oracle.image(m, T, [1.0], npix, pix, incl=incl, lam_um=lam, tau_stop=1e30)
    tau_img = oracle.image(m, T, [1.0], npix, pix, incl=incl, lam_um=lam, mode=1)
    k = np.arange(npix) - npix // 2
    X, Y = np.meshgrid(k * pix, k * pix)
    b2 = X ** 2 + Y ** 2
    path = 2 * np.sqrt(np.clip(rout ** 2 - b2, 0, None)) - 2 * np.sqrt(np.clip(rin ** 2 - b2, 0, None))
    tau = 1e-18 * 1.0 * path
    B = np_planck(C_LIGHT / (lam * 1e-4), 50.0)
    assert np.allclose(tau_img, tau, rtol=1e-9, atol=1e-14)
    assert np.allclose(img, B * (-np.expm1(-tau)), rtol=1e-9, atol=1e-30)


def test_image_thin_disk_matches_pdspy_simple_dust_image():
    """Vertically isothermal, geometrically thin disk: I = B(T)(1 - exp(-kappa
    Sigma / cos i)), the closed form of YSOModel.run_simple_dust_image
    (pdspy/modeling/YSOModel.py:794-806), at radii where the disk is thin."""
    mdl = configs.disk_c1(nr=128, nt=256, rmax=100.0)
    arr = oracle_arrays(mdl)
    kap = 0.01
    T = np.full(arr["rho"][0].size, 30.0)
    incl = 30.0
    npix, pix = 101, 1.0 * AU
    img = oracle.image(arr, T, [kap], npix, pix, incl=incl, lam_um=1300.0, tau_stop=1e30)
    B = np_planck(C_LIGHT / 1300e-4, 30.0)
    k = np.arange(npix) - npix // 2
    X, Y = np.meshgrid(k * pix, k * pix)
    R = np.sqrt(X ** 2 + (Y / np.cos(np.radians(incl))) ** 2) / AU
    Sigma = mdl.disk.surface_density(R)
    want = B * (1 - np.exp(-kap * Sigma / np.cos(np.radians(incl))))
    sel = (R > 5) & (R < 40)
    assert np.allclose(img[sel], want[sel], rtol=0.03)


def test_image_adjoint_matches_finite_differences():
    m, T = _uniform_sphere(rho=3e-16, nr=4, nt=4)
    rng = np.random.default_rng(2)
    m["rho"] = m["rho"] * rng.uniform(0.5, 2.0, m["rho"].shape)
    T = T * rng.uniform(0.7, 1.3, T.shape)
    npix, pix = 17, 5.0 * AU
    kw = dict(incl=50.0, lam_um=1300.0, tau_stop=1e30)
    dimg = rng.normal(size=(npix, npix))
    g_lr, g_T = oracle.image_adjoint(m, T, [1.0], dimg, npix, pix, **kw)

    def L(rho, TT):
        mm = dict(m)
        mm["rho"] = rho
        return np.sum(dimg * oracle.image(mm, TT, [1.0], npix, pix, **kw))

    ncell = T.size
    for c in rng.choice(ncell, 6, replace=False):
        h = 1e-4
        rp = m["rho"].copy().reshape(1, -1)
        rm = rp.copy()
        rp[0, c] *= np.exp(h)
        rm[0, c] *= np.exp(-h)
        fd = (L(rp.reshape(m["rho"].shape), T) - L(rm.reshape(m["rho"].shape), T)) / (2 * h)
        assert g_lr[0, c] == pytest.approx(fd, rel=1e-5, abs=1e-12 * abs(fd) + 1e-30)
        Tp, Tm = T.copy(), T.copy()
        dT = 1e-4 * T[c]
        Tp[c] += dT
        Tm[c] -= dT
        fdT = (L(m["rho"], Tp) - L(m["rho"], Tm)) / (2 * dT)
        assert g_T[c] == pytest.approx(fdT, rel=1e-5, abs=1e-30)


def test_image_star_point_source():
    m, T = _uniform_sphere(rho=1e-30)
    npix, pix = 9, 4.0 * AU
    img0 = oracle.image(m, np.zeros_like(T), [1.0], npix, pix, lam_um=1300.0, incl_star=False)
    img1 = oracle.image(m, np.zeros_like(T), [1.0], npix, pix, lam_um=1300.0, incl_star=True)
    R = 1.392e11
    F = np.pi * R ** 2 * np_planck(C_LIGHT / 1300e-4, 4000.0) / pix ** 2
    diff = img1 - img0
    assert diff[npix // 2, npix // 2] == pytest.approx(F, rel=1e-9)
    diff[npix // 2, npix // 2] = 0
    assert np.all(diff == 0)


# ------------------------------------------------------------------ visibilities
def test_dft_equals_imtovis_fft_convention():
    """On the pixel-conjugate grid the DFT is fftshift(fft2(ifftshift(I))) with
    u, v = fftshift(fftfreq(n, dx)) (pdspy/imaging/imtovis.py:13-24)."""
    rng = np.random.default_rng(0)
    for n in (16, 15):
        img = rng.normal(size=(n, n))
        dx = 1e-7
        V = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img)))
        uu = np.fft.fftshift(np.fft.fftfreq(n, dx))
        u, v = np.meshgrid(uu, uu)
        re, im = oracle.vis_dft(img, dx, u.ravel(), v.ravel())
        assert np.allclose(re, V.real.ravel(), atol=1e-10)
        assert np.allclose(im, V.imag.ravel(), atol=1e-10)


def test_dft_shift_theorem_and_hermitian():
    rng = np.random.default_rng(3)
    n, dx = 32, 1e-7
    img = np.zeros((n, n))
    img[10:20, 8:18] = rng.uniform(size=(10, 10))
    sh = np.roll(np.roll(img, 3, axis=1), -2, axis=0)  # +3 px in x, -2 px in y
    u = rng.uniform(-2e6, 2e6, 50)
    v = rng.uniform(-2e6, 2e6, 50)
    a = oracle.vis_dft(sh, dx, u, v)
    b = oracle.vis_dft(img, dx, u, v, dra=3 * dx, ddec=-2 * dx)
    assert np.allclose(a[0], b[0], atol=1e-9) and np.allclose(a[1], b[1], atol=1e-9)
    c = oracle.vis_dft(img, dx, -u, -v)
    d = oracle.vis_dft(img, dx, u, v)
    assert np.allclose(c[0], d[0], atol=1e-9) and np.allclose(c[1], -d[1], atol=1e-9)


def test_grid_sampler_equals_dft_at_nodes_and_scipy_between():
    """galario's method (oracle.vis_grid): at the nodes of the pixel-conjugate
    grid it is the exact DFT (also with a phase-centre shift); between nodes it
    is scipy's linear RegularGridInterpolator over numpy's FFT on the grid
    u = fftshift(fftfreq(nx, dxy)); points without four nodes are NaN."""
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(11)
    for ny, nx in ((16, 16), (15, 20)):
        dx = 1e-7
        img = rng.normal(size=(ny, nx))
        ug = np.fft.fftshift(np.fft.fftfreq(nx, dx))
        vg = np.fft.fftshift(np.fft.fftfreq(ny, dx))
        U, Vv = np.meshgrid(ug[:-1], vg[:-1])  # nodes with a right/upper neighbour
        for dra, ddec in ((0.0, 0.0), (3e-7, -1e-7)):
            g = oracle.vis_grid(img, dx, U.ravel(), Vv.ravel(), dra, ddec)
            d = oracle.vis_dft(img, dx, U.ravel(), Vv.ravel(), dra, ddec)
            assert np.allclose(g[0], d[0], atol=1e-9) and np.allclose(g[1], d[1], atol=1e-9)
        F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img)))
        u = rng.uniform(ug[0], ug[-1], 300)
        v = rng.uniform(vg[0], vg[-1], 300)
        want = RegularGridInterpolator((vg, ug), F, method="linear")(np.stack([v, u], axis=1))
        re, im = oracle.vis_grid(img, dx, u, v)
        assert np.allclose(re, want.real, atol=1e-10) and np.allclose(im, want.imag, atol=1e-10)
        # half a cell beyond the last node in u, just below the first in v
        out = oracle.vis_grid(img, dx, np.array([ug[-1] + 0.5 * (ug[1] - ug[0]), 0.0]),
                              np.array([0.0, vg[0] - 1e-3 * (vg[1] - vg[0])]))
        assert np.all(np.isnan(out[0])) and np.all(np.isnan(out[1]))


def test_grid_sampler_tracks_the_dft_on_a_resolved_image():
    """For a source much wider than a pixel the visibility function is smooth on
    the grid, and bilinear interpolation stays within h^2/8 max|F''| of the DFT
    (here: a Gaussian of sigma s, F'' <= (2 pi s)^2 |F|, h = 1 / (n dx))."""
    n, dx, s = 128, 1e-7, 8e-7
    x = (np.arange(n) - n // 2) * dx
    X, Y = np.meshgrid(x - 2e-7, x + 1e-7)
    img = np.exp(-(X ** 2 + Y ** 2) / (2 * s * s))
    rng = np.random.default_rng(12)
    lim = 0.2 / s
    u = rng.uniform(-lim, lim, 400)
    v = rng.uniform(-lim, lim, 400)
    g = oracle.vis_grid(img, dx, u, v)
    d = oracle.vis_dft(img, dx, u, v)
    h = 1.0 / (n * dx)
    # F'' along each axis bounded by ((2 pi s)^2 + (2 pi x0)^2 + 2 (2 pi)^2 s |x0|) |F|max
    x0 = 2e-7
    bound = 2 * h * h / 8 * ((2 * np.pi * s) ** 2 + (2 * np.pi * x0) ** 2 +
                             2 * (2 * np.pi) ** 2 * s * x0) * img.sum()
    err = np.abs((g[0] - d[0]) + 1j * (g[1] - d[1])).max()
    assert err <= bound
    # and the error is not trivially zero: the sampler interpolates
    assert err > 1e-6 * bound


def test_dft_adjoint_is_transpose():
    rng = np.random.default_rng(4)
    ny, nx, K, dx = 12, 10, 37, 2e-7
    u = rng.uniform(-1e6, 1e6, K)
    v = rng.uniform(-1e6, 1e6, K)
    x = rng.normal(size=(ny, nx))
    gr, gi = rng.normal(size=K), rng.normal(size=K)
    re, im = oracle.vis_dft(x, dx, u, v, 1e-7, -3e-7)
    g = oracle.vis_dft_adjoint(gr, gi, nx, ny, dx, u, v, 1e-7, -3e-7)
    assert np.dot(re, gr) + np.dot(im, gi) == pytest.approx(np.sum(g * x), rel=1e-10)


def test_chisq_closed_form():
    rng = np.random.default_rng(5)
    a = [rng.normal(size=100) for _ in range(4)]
    w = rng.uniform(0.5, 2, 100)
    want = np.sum(((a[0] - a[2]) ** 2 + (a[1] - a[3]) ** 2) * w)
    assert oracle.chisq(a[0], a[1], w, a[2], a[3]) == pytest.approx(want, rel=1e-12)


# ------------------------------------------------------------------ modified random walk
def test_mrw_escape_law_mean_first_passage():
    """Diffusion from the centre of a sphere: mean first-passage time R0^2/(6D),
    i.e. int_0^inf P(y) dy = 1/6; the inverse samples the same law."""
    from scipy.integrate import quad
    assert quad(oracle.mrw_survival, 0, 60, limit=200)[0] == pytest.approx(1 / 6, rel=1e-10)
    assert oracle.mrw_survival(0.0) == 1.0
    # the two series representations join continuously at y = 0.1
    assert oracle.mrw_survival(0.1 - 1e-12) == pytest.approx(oracle.mrw_survival(0.1), abs=1e-10)
    xi = (np.arange(4000) + 0.5) / 4000
    ys = np.array([oracle.mrw_invert(x) for x in xi])
    assert ys.mean() == pytest.approx(1 / 6, rel=2e-3)
    for x, y in zip(xi[::500], ys[::500]):
        assert oracle.mrw_survival(y) == pytest.approx(x, abs=1e-9)


def test_mrw_temperatures_agree_with_plain_walk():
    """The MRW approximation leaves the bulk of the disk temperatures within the
    seed-to-seed Monte Carlo noise of the plain random walk (median), with far
    fewer interactions per packet; in the few deepest, optically thickest cells
    the diffusion approximation differs more (tail bound 2.5x the noise)."""
    m = oracle_arrays(configs.disk_c1(nr=32, nt=32))
    n = 40000
    T0, v0, s0 = oracle.thermal(m, n, seed=1, max_events=1000000, lagged=True, nthreads=8)
    T1, v1, _ = oracle.thermal(m, n, seed=2, max_events=1000000, lagged=True, nthreads=8)
    Tm, vm, sm = oracle.thermal(m, n, seed=3, max_events=1000000, lagged=True, nthreads=8,
                                mrw=True)
    noise = dT_stats(T0, T1, np.minimum(v0, v1))
    mrw = dT_stats(T0, Tm, np.minimum(v0, vm))
    assert sm["n_mrw"] > 0 and sm["n_killed"] == 0
    assert sm["n_events"] < 0.5 * s0["n_events"]
    assert mrw[0] <= 1.5 * noise[0] + 1e-3, (mrw, noise)
    assert mrw[1] <= 2.5 * noise[1], (mrw, noise)


# ---------------------------------------------------------------- scattering source
def _shell_avg_inv_r2(r1, r2):
    # volume average of 1 / r^2 over the shell [r1, r2]
    return 3.0 * (r2 - r1) / (r2 ** 3 - r1 ** 3)


def test_scat_source_optically_thin_star_dilution():
    """Point star in an optically thin medium, dust emission off (T = 0): the
    mean intensity is the diluted stellar one, J = L_nu / (16 pi^2 r^2),
    averaged over each shell."""
    m, _ = _uniform_sphere(rho=1e-30, nr=8, nt=3)
    T = np.zeros(8 * 3)
    lam = 1.0
    J, st = oracle.scat_source(m, T, [1.0], [1.0], lam, 200000, seed=4)
    R, Ts = m["stars"][0, 0], m["stars"][0, 1]
    Lnu = 4 * np.pi ** 2 * R ** 2 * np_planck(C_LIGHT / (lam * 1e-4), Ts)
    assert st["lum_star"] == pytest.approx(Lnu, rel=1e-12) and st["lum_dust"] == 0.0
    assert st["n_escape"] == 200000
    rw = m["rw"]
    want = Lnu / (16 * np.pi ** 2) * _shell_avg_inv_r2(rw[:-1], rw[1:])
    got = J.reshape(3, 8)
    for it in range(3):
        assert np.allclose(got[it], want, rtol=0.03), (it, got[it] / want)


def test_scat_source_absorbing_sphere_attenuation():
    """Pure absorber (albedo 0) around a central star: J = L e^{-alpha r} /
    (16 pi^2 r^2) (shell averages in closed form), and the absorbed fraction
    is 1 - e^{-alpha (r_out - r_in)}."""
    rho, kap = 2e-15, 1.0
    m, _ = _uniform_sphere(rho=rho, kappa=kap, nr=6, nt=2, rin=0.5, rout=40.0)
    T = np.zeros(6 * 2)
    n = 200000
    J, st = oracle.scat_source(m, T, [kap], [0.0], 1.0, n, seed=5)
    rw = m["rw"]
    a = rho * kap
    L = st["lum_star"]
    r1, r2 = rw[:-1], rw[1:]
    want = L / (4 * np.pi) * (np.exp(-a * r1) - np.exp(-a * r2)) / a / \
        (4 * np.pi / 3 * (r2 ** 3 - r1 ** 3))
    got = J.reshape(2, 6)
    for it in range(2):
        assert np.allclose(got[it], want, rtol=0.04), (it, got[it] / want)
    p_abs = 1 - np.exp(-a * (rw[-1] - rw[0]))
    assert st["n_abs"] / n == pytest.approx(p_abs, abs=4 * np.sqrt(p_abs * (1 - p_abs) / n))
    assert st["n_abs"] + st["n_escape"] == n and st["n_scat"] == 0


def test_scat_source_thin_isothermal_dust_emission():
    """No star, optically thin isothermal sphere: volume-averaged J = j 3R/4
    (mean distance from an interior point to the surface), j = rho kappa B."""
    rho, kap, Td = 1e-19, 1.0, 40.0
    m, _ = _uniform_sphere(rho=rho, kappa=kap, nr=10, nt=4, rin=0.01, rout=40.0)
    m["stars"] = np.array([[1.0, 1.0, 0, 0, 0]])   # negligible
    T = np.full(40, Td)
    lam = 300.0
    J, st = oracle.scat_source(m, T, [kap], [kap], lam, 400000, seed=6)
    assert st["lum_star"] < 1e-30 * st["lum_dust"]
    rw, tw = m["rw"], m["tw"]
    r1, r2 = rw[:-1], rw[1:]
    vol = (np.outer(np.cos(tw[:-1]) - np.cos(tw[1:]), (r2 ** 3 - r1 ** 3) / 3) * 2 * np.pi * 2).ravel()
    j = rho * kap * np_planck(C_LIGHT / (lam * 1e-4), Td)
    R = rw[-1]
    assert (J * vol).sum() / vol.sum() == pytest.approx(j * 0.75 * R, rel=0.02)
    # each shell: J(r) = j/2 (R^2 - r^2)/(2r) ... use the exact thin-sphere law
    # J(r) = (j / 2) [R + (R^2 - r^2) / (2 r) ln((R + r) / (R - r))]
    rc = 0.5 * (r1 + r2)
    Jr = 0.5 * j * (R + (R ** 2 - rc ** 2) / (2 * rc) * np.log((R + rc) / (R - rc)))
    shell = (J.reshape(4, 10) * vol.reshape(4, 10)).sum(0) / vol.reshape(4, 10).sum(0)
    assert np.allclose(shell[2:-1], Jr[2:-1], rtol=0.05)


def test_image_with_scattering_source_identities():
    """With J = B(T) the source function is B whatever the albedo, so the
    image equals the emission-only image with extinction kabs + ksca; with
    ksca = 0 it equals the plain image whatever J."""
    m, T = _uniform_sphere(rho=3e-16, nr=5, nt=4)
    lam, npix, pix = 1300.0, 17, 5.0 * AU
    B = np_planck(C_LIGHT / (lam * 1e-4), 50.0)
    J = np.full(T.size, B)
    a = oracle.image(m, T, [1.0], npix, pix, incl=40.0, lam_um=lam, ksca=[2.5], J=J,
                     tau_stop=1e30)
    b = oracle.image(m, T, [3.5], npix, pix, incl=40.0, lam_um=lam, tau_stop=1e30)
    assert np.allclose(a, b, rtol=1e-12, atol=0)
    c = oracle.image(m, T, [1.0], npix, pix, incl=40.0, lam_um=lam, ksca=[0.0], J=J * 7.0,
                     tau_stop=1e30)
    d = oracle.image(m, T, [1.0], npix, pix, incl=40.0, lam_um=lam, tau_stop=1e30)
    assert np.allclose(c, d, rtol=1e-12, atol=0)


def test_image_adjoint_with_scattering_matches_finite_differences():
    m, T = _uniform_sphere(rho=3e-16, nr=4, nt=4)
    rng = np.random.default_rng(3)
    m["rho"] = m["rho"] * rng.uniform(0.5, 2.0, m["rho"].shape)
    T = T * rng.uniform(0.7, 1.3, T.size)
    lam, npix, pix = 1300.0, 9, 9.0 * AU
    B = np_planck(C_LIGHT / (lam * 1e-4), 50.0)
    J = B * rng.uniform(0.1, 2.0, T.size)
    kw = dict(incl=30.0, lam_um=lam, ksca=[1.7], J=J, tau_stop=1e30)
    dimg = rng.normal(size=(npix, npix))
    g_lr, g_T = oracle.image_adjoint(m, T, [1.0], dimg, npix, pix, **kw)
    L = lambda mm, TT: float((oracle.image(mm, TT, [1.0], npix, pix, **kw) * dimg).sum())
    for c in rng.choice(T.size, 8, replace=False):
        h = 1e-4
        mp, mm_ = dict(m), dict(m)
        rp, rm = m["rho"].copy(), m["rho"].copy()
        rp.reshape(-1)[c] *= np.exp(h)
        rm.reshape(-1)[c] *= np.exp(-h)
        mp["rho"], mm_["rho"] = rp, rm
        fd = (L(mp, T) - L(mm_, T)) / (2 * h)
        assert g_lr[0, c] == pytest.approx(fd, rel=1e-4, abs=1e-12 * abs(fd) + 1e-40)
        Tp, Tm = T.copy(), T.copy()
        dT = 1e-4 * T[c]
        Tp[c] += dT
        Tm[c] -= dT
        fdT = (L(m, Tp) - L(m, Tm)) / (2 * dT)
        assert g_T[c] == pytest.approx(fdT, rel=1e-4, abs=1e-40)


# ---- galario's sampleImage and pdspy's call of it (interpolate_model.py:20-27)

def _grid_nodes(n, dxy, rng, k):
    """k random nodes of the pixel-conjugate grid, away from its edges"""
    du = 1.0 / (n * dxy)
    a = rng.integers(-n // 2 + 2, n // 2 - 2, size=k)
    b = rng.integers(-n // 2 + 2, n // 2 - 2, size=k)
    return a * du, b * du


def _bordered_image(n, rng, border=6):
    img = np.zeros((n, n))
    img[border:-border, border:-border] = rng.random((n - 2 * border, n - 2 * border))
    return img


def test_galario_restatement_documented_shifts():
    """galario's documented meaning of dRA / dDec: dRA > 0 translates the image
    to the left (East), dDec > 0 translates it up (North); 'upper' puts row 0
    at the top, 'lower' at the bottom.  Translating the array by k pixels must
    equal the shift factor with dRA (dDec) = k dxy, at every grid node; a sign
    error in the shift factor or in the transform's kernel fails this."""
    rng = np.random.default_rng(3)
    n, dxy = 64, 1e-7
    img = _bordered_image(n, rng)
    u, v = _grid_nodes(n, dxy, rng, 200)
    for k in (1, 3):
        left = np.roll(img, -k, axis=1)             # k columns towards column 0
        up = np.roll(img, -k, axis=0)               # k rows towards row 0
        down = np.roll(img, k, axis=0)
        V0 = oracle.galario_sample_image(img, dxy, u, v, dRA=k * dxy)
        np.testing.assert_allclose(V0, oracle.galario_sample_image(left, dxy, u, v),
                                   rtol=1e-10, atol=1e-10 * np.abs(V0).max())
        V1 = oracle.galario_sample_image(img, dxy, u, v, dDec=k * dxy)
        np.testing.assert_allclose(V1, oracle.galario_sample_image(up, dxy, u, v),
                                   rtol=1e-10, atol=1e-10 * np.abs(V1).max())
        V2 = oracle.galario_sample_image(img, dxy, u, v, dDec=k * dxy, origin="lower")
        np.testing.assert_allclose(V2, oracle.galario_sample_image(down, dxy, u, v,
                                                                   origin="lower"),
                                   rtol=1e-10, atol=1e-10 * np.abs(V2).max())
        # the wrong direction does not match
        assert np.abs(V0 - oracle.galario_sample_image(np.roll(img, k, axis=1), dxy, u,
                                                       v)).max() > 1e-3 * np.abs(V0).max()


def test_galario_restatement_point_source_and_half_plane():
    """A point at galario's phase origin (pixel (N/2, N/2)) has V = its flux at
    every (u, v), on and off the grid; for a real image the stored half plane
    and its conjugate give V(-u, -v) = conj V(u, v); at grid nodes the
    restatement equals the full complex FFT of fftshift(image)."""
    rng = np.random.default_rng(4)
    n, dxy = 64, 1e-7
    du = 1.0 / (n * dxy)
    img = np.zeros((n, n))
    img[n // 2, n // 2] = 2.5
    u = rng.uniform(-0.45, 0.45, 500) * n * du
    v = rng.uniform(-0.45, 0.45, 500) * n * du
    V = oracle.galario_sample_image(img, dxy, u, v)
    np.testing.assert_allclose(V, 2.5 + 0j, atol=1e-12)
    img = _bordered_image(n, rng)
    Vp = oracle.galario_sample_image(img, dxy, u, v)
    Vm = oracle.galario_sample_image(img, dxy, -u, -v)
    np.testing.assert_allclose(Vm, np.conj(Vp), rtol=1e-10, atol=1e-10 * np.abs(Vp).max())
    a, b = _grid_nodes(n, dxy, rng, 100)
    F = np.fft.fftshift(np.fft.fft2(np.fft.fftshift(img)))
    ia = np.rint(a / du).astype(int) + n // 2
    ib = np.rint(b / du).astype(int) + n // 2
    np.testing.assert_allclose(oracle.galario_sample_image(img, dxy, a, b), F[ib, ia],
                               rtol=1e-10, atol=1e-10 * np.abs(F).max())


def test_pdspy_galario_call_equals_shifted_mirrored_dft():
    """pdspy's interpolate_model negates galario's imaginary part and flips the
    image rows.  On grid nodes this is exactly
    exp(-2 pi i (u dRA + v (dDec + dxy))) * DFT(-u, v) of the unflipped image
    in the imtovis convention (the flip moves the phase origin by one pixel);
    between nodes it is the imtovis-convention bilinear sampler of the flipped
    image at (-u, -v) with the shift factor exp(-2 pi i (u dRA + v dDec)).
    These are the identities the GPU's interpolate_model uses; dropping the
    flip, the negation or the one-pixel shift fails them."""
    rng = np.random.default_rng(5)
    n, dxy = 64, 4e-8
    img = _bordered_image(n, rng)
    img[20:24, 40:47] += 5.0                        # asymmetric, off centre
    dra, ddec = 0.37 * dxy, -1.9 * dxy
    u, v = _grid_nodes(n, dxy, rng, 300)
    re, im = oracle.pdspy_interpolate_galario(img, dxy, u, v, dra, ddec)
    er, ei = oracle.vis_dft(img, dxy, -u, v, -dra, ddec + dxy)
    scale = np.abs(er + 1j * ei).max()
    np.testing.assert_allclose(re, er, atol=1e-10 * scale)
    np.testing.assert_allclose(im, ei, atol=1e-10 * scale)
    # without the one-pixel shift, or without the mirror, the identity fails
    xr, xi = oracle.vis_dft(img, dxy, -u, v, -dra, ddec)
    assert np.abs((re - xr) + 1j * (im - xi)).max() > 1e-3 * scale
    xr, xi = oracle.vis_dft(img, dxy, u, v, dra, ddec + dxy)
    assert np.abs((re - xr) + 1j * (im - xi)).max() > 1e-3 * scale
    # off the grid: the imtovis-convention sampler of image[::-1] at (-u, -v)
    du = 1.0 / (n * dxy)
    u = rng.uniform(-0.4, 0.4, 400) * n * du
    v = rng.uniform(-0.4, 0.4, 400) * n * du
    re, im = oracle.pdspy_interpolate_galario(img, dxy, u, v, dra, ddec)
    gr, gi = oracle.vis_grid(img[::-1], dxy, -u, -v)
    G = (gr + 1j * gi) * np.exp(-2j * np.pi * (u * dra + v * ddec))
    np.testing.assert_allclose(re, G.real, atol=1e-10 * scale)
    np.testing.assert_allclose(im, G.imag, atol=1e-10 * scale)


# ---- Bjorkman & Wood re-emission spectrum and the mean-opacity tables
#      (temperature grid: RADMC-3D's ntemp / temp0 / temp1, radmc3d/write.py:47-52)

def _np_dB_nu(nu, T):
    """pdspy's dB_nu (misc/dB_nu.py:4-7); it carries a minus sign, the
    spectrum uses its magnitude"""
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        x = H * nu / (K_B * T)
        d = (-2 * H ** 2 * nu ** 4 / (C_LIGHT ** 2 * K_B * T ** 2)) / np.expm1(x) / (-np.expm1(-x))
    return np.nan_to_num(np.abs(d))


def _np_B_nu(nu, T):
    with np.errstate(over="ignore"):
        return (2 * H * nu ** 3 / C_LIGHT ** 2) / np.expm1(H * nu / (K_B * T))


def _np_dnu(lam_um):
    """trapezoid weights of the frequency grid nu = c / lam"""
    nu = C_LIGHT / (lam_um * 1e-4)
    d = np.zeros_like(nu)
    d[1:-1] = 0.5 * np.abs(nu[:-2] - nu[2:])
    d[0] = 0.5 * abs(nu[0] - nu[1])
    d[-1] = 0.5 * abs(nu[-2] - nu[-1])
    return nu, d


def _two_dust_model():
    lam = np.logspace(-1, 4, 100)
    m = _small_model()
    k1 = np.where(lam > 1.0, 1e3 / lam, 1e3)
    k2 = np.where(lam > 1.0, 1e3 * lam ** -0.6, 1e3)
    m.update(lam_um=lam, rho=np.ones((2, 1, 4, 5)) * 1e-20, kabs=np.array([k1, k2]),
             ksca=np.array([k1, 2 * k2]))
    return m


def _reemit_expected(m, q, T, tabs):
    """the expected distribution of the draw at T: kappa_nu |dB_nu/dT| dnu on
    the two temperature rows around T, mixed linearly in the row coordinate"""
    nu, dnu = _np_dnu(m["lam_um"])
    ka = m["kabs"][q]
    x = (np.log(T) - tabs.lnt0) / tabs.dlnt
    lo = int(np.floor(x))
    w = x - lo
    out = np.zeros(len(nu))
    for row, wt in ((lo, 1 - w), (lo + 1, w)):
        if wt > 0:
            p = ka * _np_dB_nu(nu, np.exp(tabs.lnt0 + row * tabs.dlnt)) * dnu
            out += wt * p / p.sum()
    return out


def _chi2(counts, expected_p):
    n = counts.sum()
    e = expected_p * n
    keep = e >= 20
    c = counts[keep].astype(float)
    # everything in the sparse tail as one bin
    c = np.append(c, counts[~keep].sum())
    e = np.append(e[keep], e[~keep].sum())
    sel = e > 0
    return float(((c[sel] - e[sel]) ** 2 / e[sel]).sum()), int(sel.sum() - 1)


# temperatures on a table row (row 400 = 6.5 K... row 860 = 2.6e3 K) and between rows
_REEMIT_CASES = [(0, 400, 0.0), (0, 612, 0.37), (1, 700, 0.0), (1, 860, 0.81), (0, 520, 0.5)]


def _reemit_pin(m, seed=11, n=1000000):
    """largest chi^2 excess over the cases, in standard deviations"""
    tabs = oracle.Tables(m)
    worst = -np.inf
    for q, row, frac in _REEMIT_CASES:
        T = np.exp(tabs.lnt0 + (row + frac) * tabs.dlnt)
        counts = tabs.hist(q, T, n, seed=seed + row)
        c2, dof = _chi2(counts, _reemit_expected(m, q, T, tabs))
        worst = max(worst, (c2 - dof) / np.sqrt(2 * dof))
    return worst


def test_reemission_draw_follows_kappa_dBdT():
    """2e6 draws of the thermal MC's re-emission frequency (the same function
    run_packet and the MRW step call) at temperatures on and between table
    rows, against kappa_nu dB_nu/dT dnu from pdspy's dB_nu formula in numpy,
    mixed linearly between the two rows around T: chi^2 within 5 sigma of its
    degrees of freedom."""
    assert _reemit_pin(_two_dust_model()) < 5.0


def test_reemission_pin_catches_planted_mistakes():
    """The same pin fails when the oracle's re-emission spectrum is kappa B
    instead of kappa dB/dT, and when the temperature row is off by one."""
    m = _two_dust_model()
    try:
        for code in (1, 2):
            oracle.set_mutation(code)
            assert _reemit_pin(m) > 50.0, code
    finally:
        oracle.set_mutation(0)
    assert _reemit_pin(m) < 5.0


def test_mean_opacity_tables_match_numpy_quadrature():
    """Planck-mean absorption, Rosseland-mean extinction and the emission per
    gram q(T) = 4 pi sum kappa B dnu on the temperature grid, against numpy
    quadrature of pdspy's B_nu / dB_nu on the same frequency weights; species
    with different kabs are different absorption groups."""
    m = _two_dust_model()
    tabs = oracle.Tables(m)
    assert list(tabs.grp) == [0, 1]
    nu, dnu = _np_dnu(m["lam_um"])
    for row in (150, 400, 620, 800, 990):
        T = np.exp(tabs.lnt0 + row * tabs.dlnt)
        B = _np_B_nu(nu, T)
        dB = _np_dB_nu(nu, T)
        for q in (0, 1):
            ka = m["kabs"][q]
            if B.sum() > 0:
                assert tabs.kplanck[q, row] == pytest.approx((ka * B * dnu).sum() / (B * dnu).sum(),
                                                             rel=1e-10)
                assert tabs.lnq[q, row] == pytest.approx(np.log(4 * np.pi * (ka * B * dnu).sum()),
                                                         rel=1e-12, abs=1e-9)
        for s in (0, 1):
            ke = m["kabs"][s] + m["ksca"][s]
            if dB.sum() > 0:
                want = (dB * dnu).sum() / (dB * dnu / ke).sum()
                assert tabs.kross[s, row] == pytest.approx(want, rel=1e-10)
    # identical absorption opacity: one group
    m2 = _two_dust_model()
    m2["kabs"] = np.array([m2["kabs"][0], m2["kabs"][0]])
    assert list(oracle.Tables(m2).grp) == [0, 0]


# ---- ping-pong guard: rays that graze a sphere at FP64 resolution
#      (found by a random search over tangent rays on the 64x32 disk+envelope
#      grid; each makes the traversal bounce between two cells without moving)
_GRAZING_RAYS = [
    ([-325621223417906.8, 49386887234287.49, 58561204497278.32],
     [-0.1493119313775693, -0.9887901431286116, -3.648638127115638e-15]),
    ([142879425523101.4, 254365375019529.78, 163650054660660.66],
     [-0.8718814457963383, 0.48971700447920624, 7.225896500954678e-15]),
    ([-4725679007135499.0, -1204404995739213.8, 1112849858861335.4],
     [0.24716839945556102, -0.9689725395028366, 3.8854976619789003e-16]),
]


def test_ping_pong_guard_moves_grazing_rays_on():
    """Each grazing ray leaves the grid after a few dozen segments with the
    guard's move counted; with the guard planted off the same ray bounces
    until the walk's 1e7-segment cap.  The thermal MC and the camera use the
    same traversal and guard."""
    arr = oracle_arrays(configs.disk_envelope(nr=64, nt=32))
    for o, d in _GRAZING_RAYS:
        c, ln, h, nud, n = oracle.trace_ray(arr, o, d, maxseg=1000, with_nudges=True)
        assert nud >= 1 and n < 1000
        # it walks from its start to the outer sphere (it passes outside the
        # inner hole): the chord's closed form
        o_, d_ = np.array(o), np.array(d) / np.linalg.norm(d)
        b = o_ @ d_
        t_exit = -b + np.sqrt(b * b - (o_ @ o_ - arr["rw"][-1] ** 2))
        assert ln.sum() == pytest.approx(t_exit, rel=1e-9)
    try:
        oracle.set_mutation(3)
        c, ln, h, nud, n = oracle.trace_ray(arr, *_GRAZING_RAYS[0], maxseg=10, with_nudges=True)
        assert nud == 0 and n > 9000000
    finally:
        oracle.set_mutation(0)


# ---- recursive sub-pixeling (RADMC-3D camera_nrrefine / camera_refine_criterion)

def _shell_with_cavity(nr=60, nt=16, r0=0.01, R=10.0, alpha_au=0.1):
    """uniform isothermal sphere of radius R (au) with a tiny empty cavity of
    radius r0: the log-spaced radial grid has tiny cells at the centre"""
    rw = np.logspace(np.log10(r0), np.log10(R), nr + 1) * AU
    tw = np.linspace(0, np.pi / 2, nt + 1)
    pw = np.array([0.0, 2 * np.pi])
    rho = alpha_au / AU  # kappa = 1
    m = dict(rw=rw, tw=tw, pw=pw, rho=np.full((1, 1, nt, nr), rho), lam_um=np.array([100., 2000.]),
             kabs=np.ones((1, 2)), ksca=np.zeros((1, 2)), gsca=None,
             stars=np.array([[1.392e11, 4000.0, 0, 0, 0]]))
    return m, np.full(nr * nt, 40.0)


def _shell_pixel_means(npix, pix_au, r0, R, alpha_au, B, n=600):
    """pixel averages of I(b) = B (1 - exp(-alpha L(b))), L the chord through
    the shell at impact parameter b, by the midpoint rule on n x n points per
    pixel (the central 4 x 4 pixels, where I is smooth apart from the cavity's
    kink, whose area share is ~1e-5)"""
    out = {}
    t = (np.arange(n) + 0.5) / n - 0.5
    for j in range(npix // 2 - 2, npix // 2 + 2):
        for k in range(npix // 2 - 2, npix // 2 + 2):
            X, Y = np.meshgrid((k - npix // 2 + t) * pix_au, (j - npix // 2 + t) * pix_au)
            b2 = X ** 2 + Y ** 2
            L = 2 * np.sqrt(np.clip(R * R - b2, 0, None)) - 2 * np.sqrt(np.clip(r0 * r0 - b2, 0, None))
            out[(j, k)] = B * float(np.mean(-np.expm1(-alpha_au * L)))
    return out


def test_subpixeling_converges_to_pixel_integrated_shell():
    """A face-on uniform isothermal shell with a tiny inner cavity: the centre
    pixels' exact means (numerical integration of the closed form over the
    pixel) are approached as the refinement criterion tightens; without
    sub-pixeling the centre value misses them by ~1e-3."""
    r0, R, alpha = 0.01, 10.0, 0.1
    m, T = _shell_with_cavity(r0=r0, R=R, alpha_au=alpha)
    npix, pix = 8, 2.5
    lam = 1300.0
    B = np_planck(C_LIGHT / (lam * 1e-4), 40.0)
    want = _shell_pixel_means(npix, pix, r0, R, alpha, B)

    def err(**kw):
        img = oracle.image(m, T, [1.0], npix, pix * AU, lam_um=lam, tau_stop=1e30, **kw)
        return max(abs(img[j, k] - v) / B for (j, k), v in want.items())

    e0 = err()
    e_coarse = err(nrefine=100, refine_crit=4.0)
    e_def = err(nrefine=100, refine_crit=1.0)
    e_fine = err(nrefine=100, refine_crit=0.25)
    print("sub-pixeling errors: off %.2e, crit 4 %.2e, 1 %.2e, 0.25 %.2e"
          % (e0, e_coarse, e_def, e_fine))
    assert e0 > 3e-3
    assert e_def < e0 / 10
    assert e_fine < e_def / 5 and e_def < e_coarse <= e0
    assert e_fine < 5e-5
    # the refinement depth cap: one level is what nrefine=1 allows
    img1, n1 = oracle.image(m, T, [1.0], npix, pix * AU, lam_um=lam, tau_stop=1e30, nrefine=1,
                            with_nrays=True)
    assert n1 <= npix * npix * 5
    _, n_def = oracle.image(m, T, [1.0], npix, pix * AU, lam_um=lam, tau_stop=1e30,
                            nrefine=100, with_nrays=True)
    assert n_def > 4 * n1


def test_subpixeling_adjoint_matches_finite_differences():
    m, T = _shell_with_cavity(nr=12, nt=6, r0=0.05, R=10.0, alpha_au=0.3)
    rng = np.random.default_rng(9)
    m["rho"] = m["rho"] * rng.uniform(0.5, 2.0, m["rho"].shape)
    T = T * rng.uniform(0.7, 1.3, T.shape)
    npix, pix = 6, 3.0 * AU
    kw = dict(incl=35.0, lam_um=1300.0, tau_stop=1e30, nrefine=6, refine_crit=1.0)
    dimg = rng.normal(size=(npix, npix))
    g_lr, g_T = oracle.image_adjoint(m, T, [1.0], dimg, npix, pix, **kw)

    def L(rho, TT):
        mm = dict(m)
        mm["rho"] = rho
        return np.sum(dimg * oracle.image(mm, TT, [1.0], npix, pix, **kw))

    for c in rng.choice(T.size, 6, replace=False):
        h = 1e-4
        rp = m["rho"].copy().reshape(1, -1)
        rm = rp.copy()
        rp[0, c] *= np.exp(h)
        rm[0, c] *= np.exp(-h)
        fd = (L(rp.reshape(m["rho"].shape), T) - L(rm.reshape(m["rho"].shape), T)) / (2 * h)
        assert g_lr[0, c] == pytest.approx(fd, rel=1e-5, abs=1e-12 * abs(fd) + 1e-30)
        Tp, Tm = T.copy(), T.copy()
        dT = 1e-4 * T[c]
        Tp[c] += dT
        Tm[c] -= dT
        fdT = (L(m["rho"], Tp) - L(m["rho"], Tm)) / (2 * dT)
        assert g_T[c] == pytest.approx(fdT, rel=1e-5, abs=1e-30)
