"""GPU images, adjoint, visibilities and chi^2 against closed forms and the
FP64 oracle.  Tolerances are written per test (FP32 kernels vs FP64 oracle)."""
import numpy as np
import pytest
import torch

import oracle
from pdspy_amd import autograd as rtag
from pdspy_amd import configs, inputs, rt
from pdspy_amd._lib import RTError
from tests.helpers import oracle_arrays
from tests.test_oracle import C_LIGHT, _uniform_sphere, np_planck

pytestmark = pytest.mark.gpu
AU = inputs.AU


def _dev(arr):
    return rt.DeviceModel(arr)


@pytest.mark.parametrize("incl", [0.0, 37.0, 90.0])
def test_image_uniform_shell_closed_form_gpu(incl):
    m, T = _uniform_sphere()
    dm = _dev(m)
    npix, pix = 33, 2.6 * AU
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    cam = dm.camera(npix, pix, 1300.0, incl=incl, incl_star=False, tau_stop=1e30)
    img = dm.image(cam, Td).double().cpu().numpy()
    camt = dm.camera(npix, pix, 1300.0, incl=incl, mode=1, incl_star=False)
    tau_img = dm.image(camt, Td).double().cpu().numpy()
    rin, rout = m["rw"][0], m["rw"][-1]
    k = np.arange(npix) - npix // 2
    X, Y = np.meshgrid(k * pix, k * pix)
    b2 = X ** 2 + Y ** 2
    path = 2 * np.sqrt(np.clip(rout ** 2 - b2, 0, None)) - 2 * np.sqrt(np.clip(rin ** 2 - b2, 0, None))
    tau = 1e-18 * path
    B = np_planck(C_LIGHT / 1300e-4, 50.0)
    # FP32 traversal: 1e-4 of the peak
    assert np.allclose(tau_img, tau, rtol=1e-4, atol=1e-4 * tau.max())
    assert np.allclose(img, B * (-np.expm1(-tau)), rtol=1e-4, atol=1e-4 * B * tau.max())


def _disk_setup(nr=48, nt=32, nphi=1, seed=0):
    arr = oracle_arrays(configs.disk_envelope(nr=nr, nt=nt, nphi=nphi))
    rng = np.random.default_rng(seed)
    ncell = nr * nt * nphi
    T = 20.0 + 200.0 * rng.uniform(size=ncell)
    return arr, T


@pytest.mark.parametrize("incl,pa", [(60.0, 0.0), (25.0, 33.0)])
def test_image_matches_oracle(incl, pa):
    arr, T = _disk_setup()
    dm = _dev(arr)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 64, 4.0 * AU
    cam = dm.camera(npix, pix, 1300.0, incl=incl, pa=pa, incl_star=True)
    img = dm.image(cam, torch.tensor(T, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    ref = oracle.image(arr, T, [kap, kap], npix, pix, incl=incl, posang=pa, lam_um=1300.0,
                       incl_star=True)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    print("image max rel err %.2e" % err)
    assert err < 2e-5
    # FP32 vs FP64 traversal: total flux to 1e-6
    assert img.sum() == pytest.approx(ref.sum(), rel=1e-6)


def test_image_matches_oracle_distinct_dust():
    """Disk and envelope with different opacities (each species its own kappa at
    the camera wavelength), as in pdspy's fits."""
    print("case: image, two species with different dust", flush=True)
    arr, T = _disk_setup(seed=3)
    lam = arr["lam_um"]
    arr["kabs"] = np.array(arr["kabs"], dtype=float)
    arr["kabs"][0] = np.where(lam > 1.0, 1e3 * lam ** -0.6, 1e3)
    dm = _dev(arr)
    kaps = [rt._loglog_interp(np.array([1300.0]), lam, arr["kabs"][s])[0] for s in (0, 1)]
    assert kaps[0] > 3 * kaps[1]
    npix, pix = 64, 4.0 * AU
    cam = dm.camera(npix, pix, 1300.0, incl=55.0, pa=20.0, incl_star=True)
    img = dm.image(cam, torch.tensor(T, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    ref = oracle.image(arr, T, kaps, npix, pix, incl=55.0, posang=20.0, lam_um=1300.0,
                       incl_star=True)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    print("image (distinct dust) max rel err %.2e" % err)
    assert err < 5e-5
    assert img.sum() == pytest.approx(ref.sum(), rel=1e-6)


def test_ray_tiles_do_not_change_results():
    # the launch geometry (pixels per block) must not change any ray
    arr, T = _disk_setup()
    dm = _dev(arr)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    npix, pix = 72, 4.0 * AU  # not a multiple of the tile sizes: edge tiles
    ref = None
    dimg = torch.rand((npix, npix), dtype=torch.float32, device="cuda")
    for tile in [(0, 0), (64, 1), (8, 8), (16, 4), (16, 16), (32, 8)]:
        cam = dm.camera(npix, pix, 1300.0, incl=50.0, pa=10.0, tile=tile)
        img = dm.image(cam, Td).cpu().numpy()
        glr, gT = dm.image_adjoint(cam, Td, None, dimg)
        if ref is None:
            ref = (img, glr.cpu().numpy(), gT.cpu().numpy())
            continue
        assert np.array_equal(img, ref[0]), tile
        # the adjoint's per-cell sums are fixed-point integers: the gradient
        # does not depend on the order in which rays and blocks add to a cell
        assert np.array_equal(gT.cpu().numpy(), ref[2]), tile
        assert np.array_equal(glr.cpu().numpy(), ref[1]), tile
    with pytest.raises(RTError):
        dm.image(dm.camera(npix, pix, 1300.0, tile=(10, 3)), Td)


def test_image_3d_and_tau_mode_match_oracle():
    arr, T = _disk_setup(nr=24, nt=16, nphi=12)
    dm = _dev(arr)
    kap = rt._loglog_interp(np.array([870.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 40, 6.0 * AU
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    for mode in (0, 1):
        cam = dm.camera(npix, pix, 870.0, incl=44.0, pa=10.0, phi=31.0, mode=mode, incl_star=False)
        img = dm.image(cam, Td).double().cpu().numpy()
        ref = oracle.image(arr, T, [kap, kap], npix, pix, incl=44.0, posang=10.0, phi=31.0,
                           lam_um=870.0, mode=mode)
        err = np.abs(img - ref).max() / np.abs(ref).max()
        print("mode %d max rel err %.2e" % (mode, err))
        assert err < 2e-5


def test_image_3d_central_ray_on_a_theta_and_a_phi_wall():
    """At incl 45 deg and phi 30 deg the central pixel's radial ray lies in a
    theta cone (walls every 5.625 deg) and in a phi half-plane (walls every
    30 deg).  Which side of each wall it is charged to is a matter of rounding,
    FP32 and FP64 alike, so that pixel is held to the range the oracle spans
    with the camera turned 1e-5 deg to either side of both walls; every other
    pixel to the oracle itself."""
    arr, T = _disk_setup(nr=24, nt=16, nphi=12)
    dm = _dev(arr)
    kap = rt._loglog_interp(np.array([870.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 40, 6.0 * AU
    cam = dm.camera(npix, pix, 870.0, incl=45.0, pa=10.0, phi=30.0, incl_star=False)
    img = dm.image(cam, torch.tensor(T, dtype=torch.float32, device="cuda")).double().cpu().numpy()

    def ref(di=0.0, dp=0.0):
        return oracle.image(arr, T, [kap, kap], npix, pix, incl=45.0 + di, posang=10.0,
                            phi=30.0 + dp, lam_um=870.0)

    r0 = ref()
    c = npix // 2
    d = np.abs(img - r0) / r0.max()
    d[c, c] = 0.0
    print("off-centre max rel err %.2e" % d.max())
    assert d.max() < 2e-5
    side = [ref(di, dp)[c, c] for di in (-1e-5, 1e-5) for dp in (-1e-5, 1e-5)]
    lo, hi = min(side), max(side)
    print("centre %.5e, oracle beside the walls [%.5e, %.5e]" % (img[c, c], lo, hi))
    assert lo - 0.02 * hi <= img[c, c] <= 1.02 * hi


def test_image_adjoint_matches_oracle_fp64_gradient():
    arr, T = _disk_setup(nr=32, nt=16)
    dm = _dev(arr)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 48, 5.0 * AU
    rng = np.random.default_rng(1)
    dimg = rng.normal(size=(npix, npix))
    cam = dm.camera(npix, pix, 1300.0, incl=60.0, incl_star=True)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    g_lr, g_T = dm.image_adjoint(cam, Td, None, torch.tensor(dimg, dtype=torch.float32, device="cuda"))
    o_lr, o_T = oracle.image_adjoint(arr, T, [kap, kap], dimg, npix, pix, incl=60.0, lam_um=1300.0,
                                     incl_star=True)
    e_lr = np.linalg.norm(g_lr.double().cpu().numpy() - o_lr) / np.linalg.norm(o_lr)
    e_T = np.linalg.norm(g_T.double().cpu().numpy() - o_T) / np.linalg.norm(o_T)
    print("relative gradient error: logrho %.2e  T %.2e" % (e_lr, e_T))
    assert e_lr < 5e-5 and e_T < 5e-5


def test_image_adjoint_central_differences_gpu():
    """Relative error of the hand-written adjoint against central finite
    differences of the GPU forward pass (50 random cells, FP32 forward)."""
    arr, T = _disk_setup(nr=32, nt=16, seed=3)
    dm = _dev(arr)
    npix, pix = 40, 5.0 * AU
    cam = dm.camera(npix, pix, 1300.0, incl=60.0, incl_star=False)
    logrho = torch.log(torch.tensor(arr["rho"].reshape(2, -1), dtype=torch.float64,
                                    device="cuda").clamp_min(1e-40)).float()
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(4)
    w = torch.tensor(rng.normal(size=(npix, npix)), dtype=torch.float32, device="cuda")
    lr = logrho.clone().requires_grad_(True)
    Tt = Td.clone().requires_grad_(True)
    loss = (rtag.image(lr, Tt, dm, cam) * w).sum()
    loss.backward()
    g = lr.grad[0].double().cpu().numpy()
    gT = Tt.grad.double().cpu().numpy()
    # cells that matter: largest gradients
    cells = np.argsort(-np.abs(g))[:200]
    cells = rng.choice(cells, 50, replace=False)
    errs, errsT = [], []
    for c in cells:
        h = 2e-2
        lp = logrho.clone()
        lp[0, c] += h
        lm = logrho.clone()
        lm[0, c] -= h
        fp = (dm.image(cam, Td, torch.exp(lp)).double() * w.double()).sum().item()
        fm = (dm.image(cam, Td, torch.exp(lm)).double() * w.double()).sum().item()
        fd = (fp - fm) / (2 * h)
        errs.append(abs(fd - g[c]) / (abs(g[c]) + 1e-30))
        dT = 1e-2 * T[c]
        Tp = Td.clone()
        Tp[c] += dT
        Tm = Td.clone()
        Tm[c] -= dT
        rho = torch.exp(logrho)
        fpT = (dm.image(cam, Tp, rho).double() * w.double()).sum().item()
        fmT = (dm.image(cam, Tm, rho).double() * w.double()).sum().item()
        if abs(gT[c]) > 0:
            errsT.append(abs((fpT - fmT) / (2 * dT) - gT[c]) / abs(gT[c]))
    print("FD relative error (50 cells): logrho median %.2e max %.2e; T median %.2e"
          % (np.median(errs), np.max(errs), np.median(errsT)))
    assert np.median(errs) < 2e-2
    assert np.median(errsT) < 2e-2


def test_vis_dft_matches_numpy_fft_and_oracle():
    rng = np.random.default_rng(5)
    n, dx = 64, 1e-7
    img = rng.uniform(size=(n, n))
    u, v = rt.fft_grid_uv(n, dx / inputs.arcsec)
    re, im = rt.vis_dft(torch.tensor(img, dtype=torch.float32, device="cuda"), dx, u, v)
    V = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img))).ravel()
    scale = np.abs(V).max()
    assert np.abs(re.double().cpu().numpy() - V.real).max() < 1e-5 * scale
    assert np.abs(im.double().cpu().numpy() - V.imag).max() < 1e-5 * scale
    uu, vv = configs.uv_points(K=3000, lam_m=1.3e-3, seed=1)
    dxy = 0.0078125 * inputs.arcsec
    img2 = rng.uniform(size=(256, 256))
    re2, im2 = rt.vis_dft(torch.tensor(img2, dtype=torch.float32, device="cuda"), dxy, uu, vv,
                          0.1 * inputs.arcsec, -0.05 * inputs.arcsec)
    ore, oim = oracle.vis_dft(img2, dxy, uu, vv, 0.1 * inputs.arcsec, -0.05 * inputs.arcsec)
    scale = np.abs(ore + 1j * oim).max()
    err = max(np.abs(re2.double().cpu().numpy() - ore).max(), np.abs(im2.double().cpu().numpy() - oim).max())
    print("vis max abs err / max|V| = %.2e" % (err / scale))
    assert err < 2e-5 * scale
    # non-square, with a row count the kernel's four row groups split unevenly
    img3 = rng.uniform(size=(37, 45))
    re3, im3 = rt.vis_dft(torch.tensor(img3, dtype=torch.float32, device="cuda"), dxy, uu[:700],
                          vv[:700], 0.02 * inputs.arcsec, 0.03 * inputs.arcsec)
    ore3, oim3 = oracle.vis_dft(img3, dxy, uu[:700], vv[:700], 0.02 * inputs.arcsec,
                                0.03 * inputs.arcsec)
    scale3 = np.abs(ore3 + 1j * oim3).max()
    assert np.abs(re3.double().cpu().numpy() - ore3).max() < 2e-5 * scale3
    assert np.abs(im3.double().cpu().numpy() - oim3).max() < 2e-5 * scale3
    # enough points to fill the GPU without row groups (one group per workgroup)
    u4, v4 = configs.uv_points(K=140000, lam_m=1.3e-3, seed=4)
    img4 = rng.uniform(size=(24, 32))
    re4, im4 = rt.vis_dft(torch.tensor(img4, dtype=torch.float32, device="cuda"), dxy, u4, v4)
    ore4, oim4 = oracle.vis_dft(img4, dxy, u4, v4)
    scale4 = np.abs(ore4 + 1j * oim4).max()
    assert np.abs(re4.double().cpu().numpy() - ore4).max() < 2e-5 * scale4
    assert np.abs(im4.double().cpu().numpy() - oim4).max() < 2e-5 * scale4


def test_vis_grid_matches_oracle_and_dft_at_nodes():
    """galario's method on the GPU (rt_vis_grid) against the FP64 oracle at the
    workload's (u,v) sampling, with a phase-centre shift, on a square and an odd
    non-square image; at grid nodes it equals the exact DFT kernel."""
    rng = np.random.default_rng(21)
    uu, vv = configs.uv_points(K=20000, lam_m=1.3e-3, seed=3)
    dxy = 0.0078125 * inputs.arcsec
    for ny, nx in ((256, 256), (75, 64)):
        print("case: vis_grid %dx%d" % (ny, nx), flush=True)
        img = rng.uniform(size=(ny, nx))
        t = torch.tensor(img, dtype=torch.float32, device="cuda")
        d = dxy * (256 / nx)
        for dra, ddec in ((0.0, 0.0), (0.1 * inputs.arcsec, -0.05 * inputs.arcsec)):
            re, im = rt.vis_grid(t, d, uu, vv, dra, ddec)
            ore, oim = oracle.vis_grid(img, d, uu, vv, dra, ddec)
            scale = np.abs(ore + 1j * oim).max()
            err = max(np.abs(re.double().cpu().numpy() - ore).max(),
                      np.abs(im.double().cpu().numpy() - oim).max())
            print("vis_grid max abs err / max|V| = %.2e" % (err / scale))
            assert err < 2e-5 * scale
        ug = (np.arange(nx - 1) - nx // 2) / (nx * d)
        vg = (np.arange(ny - 1) - ny // 2) / (ny * d)
        U, V = np.meshgrid(ug, vg)
        g = rt.vis_grid(t, d, U.ravel(), V.ravel(), 3e-7, 1e-7)
        e = rt.vis_dft(t, d, U.ravel(), V.ravel(), 3e-7, 1e-7)
        scale = float(torch.hypot(e[0], e[1]).max())
        assert float((g[0] - e[0]).abs().max()) < 2e-5 * scale
        assert float((g[1] - e[1]).abs().max()) < 2e-5 * scale
    # a caller's workspace gives the same result as the stream-ordered one
    need = int(rt._lib.context(0).L.rt_vis_grid_workspace(nx, ny))
    work = torch.empty(need, dtype=torch.uint8, device="cuda")
    g2 = rt.vis_grid(t, d, U.ravel(), V.ravel(), 3e-7, 1e-7, work=work)
    assert torch.equal(g2[0], g[0]) and torch.equal(g2[1], g[1])
    with pytest.raises(ValueError):
        rt.vis_grid(t, d, np.array([0.5 / d]), np.array([0.0]))


def test_interpolate_model_codes():
    """interpolate_model: code="galario" samples by galario's method, "exact" by
    the DFT; on a Gaussian of width s the two agree to the bilinear bound
    h^2/8 |F''| summed over both axes, |F''| <= (2 pi s)^2 F(0), h = 1/(n dxy)."""
    from types import SimpleNamespace
    n, pix = 256, 0.01
    x = (np.arange(n) - n // 2) * pix
    X, Y = np.meshgrid(x, x)
    img = np.exp(-(X ** 2 + Y ** 2) / (2 * 0.15 ** 2))
    im = SimpleNamespace(x=x, y=x, image=img[:, :, None, None])
    u, v = configs.uv_points(K=5000, lam_m=1.3e-3, seed=4)
    a = rt.interpolate_model(u, v, np.array([2.3e11]), im, dRA=0.05, dDec=-0.02)
    b = rt.interpolate_model(u, v, np.array([2.3e11]), im, dRA=0.05, dDec=-0.02, code="exact")
    d = np.abs((a.real - b.real) + 1j * (a.imag - b.imag)).max() / img.sum()
    print("galario-method vs exact DFT: max |dV| / V(0) = %.2e" % d)
    h = 1.0 / (n * pix * inputs.arcsec)
    bound = 2 * h * h / 8 * (2 * np.pi * 0.15 * inputs.arcsec) ** 2
    assert 0.1 * bound < d <= bound
    with pytest.raises(ValueError):
        rt.interpolate_model(u, v, np.array([2.3e11]), im, code="trift")


def test_interpolate_model_follows_pdspys_galario_call():
    """interpolate_model(code="galario") against the oracle's restatement of
    what pdspy computes (interpolate_model.py:23-27: galario's sampleImage of
    image[::-1], imaginary part negated) on an asymmetric, off-centre image
    with nonzero dRA and dDec; code="exact" against the exact form of the same
    visibility function (exp(-2 pi i (u dRA + v (dDec + dxy))) DFT(-u, v))."""
    from types import SimpleNamespace
    rng = np.random.default_rng(31)
    n, pix = 256, 0.0078125
    x = (np.arange(n) - n // 2) * pix
    X, Y = np.meshgrid(x, x)
    img = (np.exp(-((X - 0.21) ** 2 + (Y + 0.08) ** 2) / (2 * 0.05 ** 2)) +
           0.4 * np.exp(-((X + 0.3) ** 2 / (2 * 0.12 ** 2) + (Y - 0.25) ** 2 / (2 * 0.03 ** 2))) +
           0.05 * rng.uniform(size=(n, n)))
    im = SimpleNamespace(x=x, y=x, image=img[:, :, None, None])
    u, v = configs.uv_points(K=20000, lam_m=1.3e-3, seed=8)
    dra, ddec = 0.037, -0.052                      # arcsec
    dxy = pix * inputs.arcsec
    a = rt.interpolate_model(u, v, np.array([2.3e11]), im, dRA=dra, dDec=ddec)
    ore, oim = oracle.pdspy_interpolate_galario(img, dxy, u, v, dra * inputs.arcsec,
                                                ddec * inputs.arcsec)
    scale = np.abs(ore + 1j * oim).max()
    err = max(np.abs(a.real[:, 0] - ore).max(), np.abs(a.imag[:, 0] - oim).max()) / scale
    print("galario convention: max |dV| / max|V| = %.2e" % err)
    assert err < 2e-5
    b = rt.interpolate_model(u, v, np.array([2.3e11]), im, dRA=dra, dDec=ddec, code="exact")
    ere, eim = oracle.vis_dft(img, dxy, -u, v, -dra * inputs.arcsec, ddec * inputs.arcsec + dxy)
    err = max(np.abs(b.real[:, 0] - ere).max(), np.abs(b.imag[:, 0] - eim).max()) / scale
    print("exact form: max |dV| / max|V| = %.2e" % err)
    assert err < 2e-5
    # the imtovis convention is a different function for this image
    c, d = oracle.vis_grid(img, dxy, u, v, dra * inputs.arcsec, ddec * inputs.arcsec)
    assert np.abs((c - ore) + 1j * (d - oim)).max() > 0.05 * scale


@pytest.mark.parametrize("nx,ny,K", [(48, 48, 1500), (40, 56, 777)])
def test_vis_adjoint_matches_oracle(nx, ny, K):
    # square and non-square images, partial 16x16 tiles and a partial last
    # chunk of points (the adjoint's tile update runs on the matrix cores)
    print("case: vis adjoint vs oracle nx=%d ny=%d K=%d" % (nx, ny, K), flush=True)
    rng = np.random.default_rng(6)
    dxy = 0.03 * inputs.arcsec
    u, v = configs.uv_points(K=K, lam_m=1.3e-3, seed=2)
    gr, gi = rng.normal(size=K), rng.normal(size=K)
    g = rt.vis_dft_adjoint(torch.tensor(gr, dtype=torch.float32, device="cuda"),
                           torch.tensor(gi, dtype=torch.float32, device="cuda"), nx, ny, dxy, u, v,
                           0.02 * inputs.arcsec, 0.01 * inputs.arcsec).double().cpu().numpy()
    o = oracle.vis_dft_adjoint(gr, gi, nx, ny, dxy, u, v, 0.02 * inputs.arcsec, 0.01 * inputs.arcsec)
    assert g.shape == o.shape == (ny, nx)
    assert np.linalg.norm(g - o) / np.linalg.norm(o) < 1e-5
    assert np.abs(g - o).max() <= 1e-5 * np.abs(o).max()


def test_chisq_and_gradient():
    rng = np.random.default_rng(7)
    K = 10007
    a = [rng.normal(size=K).astype(np.float32) for _ in range(4)]
    w = rng.uniform(0.5, 2, K).astype(np.float32)
    t = [torch.tensor(x, device="cuda") for x in (a[0], a[1], w, a[2], a[3])]
    out, gre, gim = rt.chisq(*t, with_grad=True)
    want = oracle.chisq(a[0], a[1], w, a[2], a[3])
    assert out.item() == pytest.approx(want, rel=1e-6)
    assert np.allclose(gre.cpu().numpy(), -2 * w * (a[0] - a[2]), rtol=1e-5)
    # deterministic reduction
    out2 = rt.chisq(*t)
    assert out2.item() == out.item()


def test_lnlike_vis_follows_pdspy():
    """rt.lnlike_vis against pdspy's formula (utils/emcee.py:31-43) in FP64 on
    the same FP32-rounded inputs, with zero-weight points; a model equal to the
    data leaves only the weight term."""
    from types import SimpleNamespace
    rng = np.random.default_rng(8)
    K = 20011
    f32 = lambda x: x.astype(np.float32).astype(float)
    d = SimpleNamespace(real=f32(rng.normal(size=(K, 1))), imag=f32(rng.normal(size=(K, 1))),
                        weights=f32(rng.uniform(0.1, 5.0, (K, 1))))
    d.weights[::97] = 0.0
    m = SimpleNamespace(real=f32(rng.normal(size=(K, 1))), imag=f32(rng.normal(size=(K, 1))))
    good = d.weights > 0
    want = (-0.5 * np.sum((d.real - m.real) ** 2 * d.weights) -
            np.sum(np.log(d.weights[good] / (2 * np.pi))) +
            -0.5 * np.sum((d.imag - m.imag) ** 2 * d.weights) -
            np.sum(np.log(d.weights[good] / (2 * np.pi))))
    got = rt.lnlike_vis(d, m)
    assert got == pytest.approx(want, rel=1e-9, abs=1e-6 * abs(want))
    same = rt.lnlike_vis(d, d)
    assert same == -2 * np.sum(np.log(d.weights[good] / (2 * np.pi)))


def test_chi2_of_visibilities_gradient_matches_oracle_chain():
    """d chi^2 / d log rho through image -> DFT -> chi^2 (config-3 chain, small)."""
    arr, T = _disk_setup(nr=32, nt=16, seed=8)
    dm = _dev(arr)
    npix = 48
    pixsize = 0.05  # arcsec
    dpc = 140.0
    pix_cm = pixsize * dpc * AU
    dxy = pixsize * inputs.arcsec
    cam = dm.camera(npix, pix_cm, 1300.0, incl=60.0, incl_star=True)
    u, v = configs.uv_points(K=2000, bmax_m=300.0, lam_m=1.3e-3, seed=3)
    logrho = torch.log(torch.tensor(arr["rho"].reshape(2, -1), dtype=torch.float32,
                                    device="cuda").clamp_min(1e-38))
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    with torch.no_grad():
        img = rtag.image(logrho, Td, dm, cam)
        re, im = rtag.visibilities(img, torch.tensor(u, device="cuda"), torch.tensor(v, device="cuda"), dxy)
    rng = np.random.default_rng(9)
    sig = 0.01 * torch.sqrt(re ** 2 + im ** 2).max().item()
    dre = re + torch.tensor(rng.normal(size=len(u)) * sig * 3, dtype=torch.float32, device="cuda")
    dim = im + torch.tensor(rng.normal(size=len(u)) * sig * 3, dtype=torch.float32, device="cuda")
    w = torch.full_like(dre, 1.0 / sig ** 2)
    lr = logrho.clone().requires_grad_(True)
    Tt = Td.clone().requires_grad_(True)
    mre, mim = rtag.visibilities(rtag.image(lr, Tt, dm, cam), torch.tensor(u, device="cuda"),
                                 torch.tensor(v, device="cuda"), dxy)
    loss = rtag.chi2(mre, mim, dre, dim, w)
    loss.backward()
    # FP64 oracle chain
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    oimg = oracle.image(arr, T, [kap, kap], npix, pix_cm, incl=60.0, lam_um=1300.0, incl_star=True)
    ore, oim = oracle.vis_dft(oimg, dxy, u, v)
    d_re, d_im, ww = dre.double().cpu().numpy(), dim.double().cpu().numpy(), w.double().cpu().numpy()
    g_lr = lr.grad.double().cpu().numpy()
    g_T = Tt.grad.double().cpu().numpy()

    def oracle_chain(mre_, mim_):
        gre, gim = -2 * ww * (d_re - mre_), -2 * ww * (d_im - mim_)
        gimg = oracle.vis_dft_adjoint(gre, gim, npix, npix, dxy, u, v)
        o_lr, o_T = oracle.image_adjoint(arr, T, [kap, kap], gimg, npix, pix_cm, incl=60.0,
                                         lam_um=1300.0, incl_star=True)
        return (np.linalg.norm(g_lr - o_lr) / np.linalg.norm(o_lr),
                np.linalg.norm(g_T - o_T) / np.linalg.norm(o_T))

    # the chain itself: the oracle's adjoints applied to the GPU model's
    # residual (data - model)
    e_lr, e_T = oracle_chain(re.double().cpu().numpy(), im.double().cpu().numpy())
    print("chi2 chain relative gradient error, same residual: logrho %.2e T %.2e" % (e_lr, e_T))
    assert e_lr < 5e-5 and e_T < 5e-5
    # end to end, the oracle with its own FP64 model: the data here are the
    # model plus noise, so the residual is pure noise and the ~1e-6 difference
    # of the FP32 model (exp of a float32 log-density included) enters it
    # coherently; the DFT adjoint sums that smooth difference over the
    # (u,v) points in phase and the noise out of phase, which raises it ~40x
    e_lr, e_T = oracle_chain(ore, oim)
    print("chi2 chain relative gradient error, FP64 model: logrho %.2e T %.2e" % (e_lr, e_T))
    assert e_lr < 1e-2 and e_T < 1e-2


def test_run_image_and_visibilities_units():
    m = configs.disk_envelope(nr=32, nt=16)
    rt.run_thermal(m, nphot=20000, iseed=1, max_events=20000)
    im = rt.run_image(m, name="1300", npix=32, pixelsize=0.1, lam="1300", incl=60, dpc=140,
                      scattering_mode_max=0)
    assert im.image.shape == (32, 32, 1, 1)
    assert im.x[16] == 0.0 and im.x[17] == pytest.approx(0.1)
    assert im.wave[0] == pytest.approx(0.13)
    # Jy/pixel conversion (Model.py:566-567) against an explicit rendering in cgs
    dm = rt._device_model(m)
    T = torch.tensor(rt.imaging_temperature(m.grid), dtype=torch.float32, device="cuda")
    pix_cm = 0.1 * 140 * AU
    # RADMC-3D's sub-pixeling at its defaults (camera_nrrefine 100, criterion 1)
    cgs = dm.image(dm.camera(32, pix_cm, 1300.0, incl=60, nrefine=100), T).double().cpu().numpy()
    assert np.allclose(im.image[:, :, 0, 0], cgs / inputs.Jy * (pix_cm / (140 * inputs.pc)) ** 2,
                       rtol=1e-6)
    im0 = rt.run_image(m, name="p", npix=32, pixelsize=0.1, lam="1300", incl=60, dpc=140,
                       scattering_mode_max=0, camera_nrrefine=0)
    plain = dm.image(dm.camera(32, pix_cm, 1300.0, incl=60), T).double().cpu().numpy()
    assert np.allclose(im0.image[:, :, 0, 0],
                       plain / inputs.Jy * (pix_cm / (140 * inputs.pc)) ** 2, rtol=1e-6)
    assert not np.allclose(im0.image, im.image, rtol=1e-3)
    # by default the image carries the scattering source (nphot packets, iseed)
    ims = rt.run_image(m, name="s", nphot=100000, npix=32, pixelsize=0.1, lam="1300", incl=60,
                       dpc=140, iseed=3)
    J, _ = dm.scattering_source(1300.0, 100000, T, seed=3)
    cgs_s = dm.image(dm.camera(32, pix_cm, 1300.0, incl=60, jscat=J, nrefine=100),
                     T).double().cpu().numpy()
    assert np.allclose(ims.image[:, :, 0, 0],
                       cgs_s / inputs.Jy * (pix_cm / (140 * inputs.pc)) ** 2, rtol=1e-6)
    vis = rt.run_visibilities(m, name="1300", npix=32, pixelsize=0.1, lam="1300", incl=60, dpc=140,
                              scattering_mode_max=0)
    V = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(im.image[:, :, 0, 0])))
    scale = np.abs(V).max()
    assert np.abs(vis.real[:, 0] - V.real.ravel()).max() < 1e-5 * scale
    assert np.abs(vis.imag[:, 0] - V.imag.ravel()).max() < 1e-5 * scale


# ---------------------------------------------------------------- scattering source
def _kap_at(arr, key, lam):
    return [rt._loglog_interp(np.array([lam]), arr["lam_um"], arr[key][s])[0]
            for s in range(arr[key].shape[0])]


def test_scattering_source_thin_shell_closed_form_gpu():
    # star only (T = 0), optically thin: J = L_nu / (16 pi^2 r^2), shell averages
    print("case: scattering source, thin shell around a star", flush=True)
    m, _ = _uniform_sphere(rho=1e-30, nr=8, nt=3)
    dm = _dev(m)
    T0 = torch.zeros(24, dtype=torch.float32, device="cuda")
    J, st = dm.scattering_source(1.0, 400000, T0, seed=4)
    R, Ts = m["stars"][0, 0], m["stars"][0, 1]
    Lnu = 4 * np.pi ** 2 * R ** 2 * np_planck(C_LIGHT / 1e-4, Ts)
    assert st["lum_star"] == pytest.approx(Lnu, rel=1e-12) and st["lum_dust"] == 0.0
    assert st["n_escape"] == 400000
    rw = m["rw"]
    want = Lnu / (16 * np.pi ** 2) * 3.0 * (rw[1:] - rw[:-1]) / (rw[1:] ** 3 - rw[:-1] ** 3)
    got = J.double().cpu().numpy().reshape(3, 8)
    for it in range(3):
        assert np.allclose(got[it], want, rtol=0.03), (it, got[it] / want)


def test_scattering_source_matches_oracle_and_is_reproducible():
    # the same Philox streams as the FP64 oracle: per-cell J agrees up to the
    # packets whose FP32 and FP64 paths part
    print("case: scattering source vs oracle, disk+envelope 32x24 at 1.3 mm", flush=True)
    arr, T = _disk_setup(nr=32, nt=24)
    dm = _dev(arr)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    lam, n = 1300.0, 400000
    J1, st = dm.scattering_source(lam, n, Td, seed=11)
    J2, _ = dm.scattering_source(lam, n, Td, seed=11)
    assert torch.equal(J1, J2)
    ka, ks = _kap_at(arr, "kabs", lam), _kap_at(arr, "ksca", lam)
    Jo, so = oracle.scat_source(arr, T, ka, ks, lam, n, seed=11)
    Jg = J1.double().cpu().numpy()
    assert st["lum_dust"] == pytest.approx(so["lum_dust"], rel=1e-5)
    for k in ("n_escape", "n_abs", "n_scat"):
        assert st[k] == pytest.approx(so[k], rel=2e-3), k
    vol = np.asarray(oracle_vol(arr))
    assert (Jg * vol).sum() == pytest.approx((Jo * vol).sum(), rel=2e-3)
    sel = Jo > 1e-3 * Jo.max()
    rel = np.abs(Jg[sel] - Jo[sel]) / Jo[sel]
    print("J vs oracle: median %.2e, p90 %.2e" % (np.median(rel), np.percentile(rel, 90)))
    assert np.median(rel) < 5e-3 and np.percentile(rel, 90) < 5e-2


def test_scattering_source_3d_hg_grid_matches_oracle():
    # phi walls and Henyey-Greenstein dust (the image's scattering source is
    # isotropic, as RADMC-3D's for pdspy's g-less opacity files)
    print("case: scattering source vs oracle, 3D spiral 24x16x12 at 0.8 mm", flush=True)
    m = configs.spiral_3d(nr=24, nt=16, nphi=12)
    arr = oracle_arrays(m)
    dm = _dev(arr)
    rng = np.random.default_rng(5)
    T = 15.0 + 150.0 * rng.uniform(size=dm.ncell)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    lam, n = 800.0, 300000
    J, st = dm.scattering_source(lam, n, Td, seed=21)
    Jo, so = oracle.scat_source(arr, T, _kap_at(arr, "kabs", lam), _kap_at(arr, "ksca", lam),
                                lam, n, seed=21)
    Jg = J.double().cpu().numpy()
    vol = np.asarray(oracle_vol(arr))
    assert (Jg * vol).sum() == pytest.approx((Jo * vol).sum(), rel=3e-3)
    sel = Jo > 1e-3 * Jo.max()
    rel = np.abs(Jg[sel] - Jo[sel]) / Jo[sel]
    print("3D J vs oracle: median %.2e, p90 %.2e" % (np.median(rel), np.percentile(rel, 90)))
    assert np.median(rel) < 5e-3 and np.percentile(rel, 90) < 5e-2


def oracle_vol(arr):
    rw, tw, pw = arr["rw"], arr["tw"], arr["pw"]
    vr = (rw[1:] ** 3 - rw[:-1] ** 3) / 3
    vt = np.cos(tw[:-1]) - np.cos(tw[1:])
    vp = pw[1:] - pw[:-1]
    mirror = abs(tw[-1] - np.pi / 2) < 1e-9
    return (vp[:, None, None] * vt[None, :, None] * vr[None, None, :]).ravel() * (2 if mirror else 1)


def _scat_image_setup(seed=0):
    arr, T = _disk_setup(nr=32, nt=16, seed=seed)
    rng = np.random.default_rng(seed + 10)
    B = np_planck(C_LIGHT / 1300e-4, T)
    J = (B * rng.uniform(0.05, 1.5, T.size)).astype(np.float32).astype(np.float64)
    return arr, T, J


def test_image_with_scattering_source_matches_oracle():
    print("case: image with scattering source vs oracle", flush=True)
    arr, T, J = _scat_image_setup()
    dm = _dev(arr)
    lam, npix, pix = 1300.0, 64, 4.0 * AU
    Jd = torch.tensor(J, dtype=torch.float32, device="cuda")
    cam = dm.camera(npix, pix, lam, incl=60.0, incl_star=True, jscat=Jd)
    img = dm.image(cam, torch.tensor(T, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    ref = oracle.image(arr, T, _kap_at(arr, "kabs", lam), npix, pix, incl=60.0, lam_um=lam,
                       incl_star=True, ksca=_kap_at(arr, "ksca", lam), J=J)
    plain = oracle.image(arr, T, _kap_at(arr, "kabs", lam), npix, pix, incl=60.0, lam_um=lam,
                         incl_star=True)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    print("image (scattering) max rel err %.2e; scattering changes the image by %.2e" % (
        err, np.abs(ref - plain).max() / np.abs(plain).max()))
    assert err < 2e-5
    assert img.sum() == pytest.approx(ref.sum(), rel=1e-6)
    assert np.abs(ref - plain).max() > 0.05 * np.abs(plain).max()


def test_image_adjoint_with_scattering_source_matches_oracle():
    print("case: image adjoint with scattering source vs oracle", flush=True)
    arr, T, J = _scat_image_setup(seed=2)
    dm = _dev(arr)
    lam, npix, pix = 1300.0, 48, 5.0 * AU
    rng = np.random.default_rng(1)
    dimg = rng.normal(size=(npix, npix))
    Jd = torch.tensor(J, dtype=torch.float32, device="cuda")
    cam = dm.camera(npix, pix, lam, incl=60.0, incl_star=True, jscat=Jd)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    g_lr, g_T = dm.image_adjoint(cam, Td, None, torch.tensor(dimg, dtype=torch.float32,
                                                              device="cuda"))
    o_lr, o_T = oracle.image_adjoint(arr, T, _kap_at(arr, "kabs", lam), dimg, npix, pix,
                                     incl=60.0, lam_um=lam, incl_star=True,
                                     ksca=_kap_at(arr, "ksca", lam), J=J)
    e_lr = np.linalg.norm(g_lr.double().cpu().numpy() - o_lr) / np.linalg.norm(o_lr)
    e_T = np.linalg.norm(g_T.double().cpu().numpy() - o_T) / np.linalg.norm(o_T)
    print("relative gradient error (scattering): logrho %.2e  T %.2e" % (e_lr, e_T))
    assert e_lr < 5e-5 and e_T < 5e-5


def test_run_image_adds_scattering_source_unless_disabled():
    print("case: run_image with and without the scattering source", flush=True)
    m = configs.disk_c1(nr=24, nt=16)
    rt.run_thermal(m, nphot=20000, seed=1, max_events=100000)
    a = rt.run_image(m, name="s", nphot=100000, npix=32, pixelsize=0.5, lam="1300", incl=45,
                     dpc=140).image
    b = rt.run_image(m, name="n", nphot=100000, npix=32, pixelsize=0.5, lam="1300", incl=45,
                     dpc=140, scattering_mode_max=0).image
    c = rt.run_image(m, name="z", nphot=0, npix=32, pixelsize=0.5, lam="1300", incl=45,
                     dpc=140).image
    assert np.array_equal(b, c)
    assert np.all(np.isfinite(a)) and a.sum() > 0
    assert np.abs(a - b).max() > 1e-3 * np.abs(b).max()


def test_image_crossing_counts_match_oracle_segments():
    """rt_image_crossings: cells crossed by an image's rays (and by its
    adjoint's reverse sweeps) equal the number of in-grid segments the oracle
    finds along the same camera rays."""
    print("case: image crossing counters vs oracle segments, 24x24 px", flush=True)
    arr, T = _disk_setup(nr=32, nt=16)
    dm = _dev(arr)
    npix, pix, incl = 24, 5.0 * AU, 55.0
    cam = dm.camera(npix, pix, 1300.0, incl=incl, incl_star=False)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    dm.image_crossings(reset=True)
    dm.image(cam, Td)
    n_img, n_adj0 = dm.image_crossings(reset=True)
    dm.image_adjoint(cam, Td, None, torch.ones((npix, npix), dtype=torch.float32,
                                               device="cuda"))
    n_img1, n_adj = dm.image_crossings(reset=True)
    assert n_adj0 == 0 and n_img1 == 0
    i = np.radians(incl)
    n = np.array([0.0, -np.sin(i), np.cos(i)])
    ex = np.array([1.0, 0.0, 0.0])
    ey = np.cross(n, ex)
    R = arr["rw"][-1]
    ref = 0
    for j in range(npix):
        for k in range(npix):
            o = (k - npix // 2) * pix * ex + (j - npix // 2) * pix * ey + 2 * R * n
            c, _, _ = oracle.trace_ray(arr, o, -n)
            ref += int(np.sum(c >= 0))
    print("crossings: image %d, adjoint %d, oracle %d" % (n_img, n_adj, ref), flush=True)
    assert n_img == n_adj
    assert abs(n_img - ref) <= 1e-3 * ref
