"""GPU recursive sub-pixeling (RADMC-3D camera_nrrefine /
camera_refine_criterion) against the FP64 oracle and the closed form of a
uniform shell with a tiny cavity."""
import numpy as np
import pytest
import torch

import oracle
from pdspy_amd import inputs, rt
from tests.test_oracle import C_LIGHT, _shell_pixel_means, _shell_with_cavity, np_planck
from tests.test_gpu_image import _disk_setup

pytestmark = pytest.mark.gpu
AU = inputs.AU


def test_subpixel_shell_closed_form_and_oracle():
    """Face-on shell with a 0.01 au cavity on a log grid: the refined centre
    pixels reach the pixel-integrated closed form (to FP32 traversal
    accuracy), the unrefined ones miss it by ~1e-2, and the GPU traces the
    same number of sub-pixel rays as the oracle."""
    print("case: sub-pixeling, shell with cavity, 8x8 px", flush=True)
    r0, R, alpha = 0.01, 10.0, 0.1
    m, T = _shell_with_cavity(r0=r0, R=R, alpha_au=alpha)
    dm = rt.DeviceModel(m)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    npix, pix = 8, 2.5
    B = np_planck(C_LIGHT / 1300e-4, 40.0)
    want = _shell_pixel_means(npix, pix, r0, R, alpha, B)
    errs = {}
    for crit in (None, 1.0, 0.25):
        kw = dict(nrefine=100, refine_crit=crit) if crit else {}
        cam = dm.camera(npix, pix * AU, 1300.0, incl_star=False, tau_stop=1e30, **kw)
        img = dm.image(cam, Td).double().cpu().numpy()
        errs[crit] = max(abs(img[j, k] - v) / B for (j, k), v in want.items())
        if crit:
            nr_gpu, uns = dm.refine_stats()
            ref, nr_or = oracle.image(m, T, [1.0], npix, pix * AU, lam_um=1300.0, tau_stop=1e30,
                                      nrefine=100, refine_crit=crit, with_nrays=True)
            err_o = np.abs(img - ref).max() / B
            print("crit %g: rays gpu %d oracle %d, max |gpu - oracle| / B %.2e"
                  % (crit, nr_gpu, nr_or, err_o))
            assert uns == 0
            assert abs(nr_gpu - nr_or) <= 0.002 * nr_or
            assert err_o < 2e-5
    print("errors vs closed form:", errs)
    assert errs[None] > 3e-3
    assert errs[1.0] < errs[None] / 10
    assert errs[0.25] < 6e-5


@pytest.mark.parametrize("incl,pa", [(60.0, 0.0), (35.0, 20.0)])
def test_subpixel_image_matches_oracle_disk_envelope(incl, pa):
    """Disk + envelope (tiny cells at the 0.1 au inner rim), pdspy's defaults
    camera_nrrefine = 100, camera_refine_criterion = 1."""
    print("case: sub-pixeling disk+envelope 48x32, incl %g pa %g" % (incl, pa), flush=True)
    arr, T = _disk_setup()
    dm = rt.DeviceModel(arr)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 48, 4.0 * AU
    cam = dm.camera(npix, pix, 1300.0, incl=incl, pa=pa, incl_star=True, nrefine=100)
    img = dm.image(cam, torch.tensor(T, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    nr_gpu, uns = dm.refine_stats()
    ref, nr_or = oracle.image(arr, T, [kap, kap], npix, pix, incl=incl, posang=pa, lam_um=1300.0,
                              incl_star=True, nrefine=100, with_nrays=True)
    plain = oracle.image(arr, T, [kap, kap], npix, pix, incl=incl, posang=pa, lam_um=1300.0,
                         incl_star=True)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    print("rays gpu %d oracle %d (plain %d); max rel err %.2e; refinement changes the image "
          "by %.2e" % (nr_gpu, nr_or, npix * npix, err, np.abs(plain - ref).max() / ref.max()))
    assert uns == 0
    assert abs(nr_gpu - nr_or) <= 0.002 * nr_or and nr_or > 2 * npix * npix
    assert err < 2e-5
    assert img.sum() == pytest.approx(ref.sum(), rel=1e-6)


def test_subpixel_adjoint_matches_oracle_and_repeats():
    print("case: sub-pixeled adjoint vs oracle", flush=True)
    arr, T = _disk_setup(nr=32, nt=16)
    dm = rt.DeviceModel(arr)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    npix, pix = 32, 6.0 * AU
    rng = np.random.default_rng(4)
    dimg = rng.normal(size=(npix, npix))
    cam = dm.camera(npix, pix, 1300.0, incl=50.0, pa=15.0, incl_star=True, nrefine=100)
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    dd = torch.tensor(dimg, dtype=torch.float32, device="cuda")
    g_lr, g_T = dm.image_adjoint(cam, Td, None, dd)
    o_lr, o_T = oracle.image_adjoint(arr, T, [kap, kap], dimg, npix, pix, incl=50.0, posang=15.0,
                                     lam_um=1300.0, incl_star=True, nrefine=100)
    for g, o, name in ((g_lr.double().cpu().numpy(), o_lr, "logrho"),
                       (g_T.double().cpu().numpy(), o_T, "T")):
        rel = np.linalg.norm(g - o) / np.linalg.norm(o)
        print("adjoint %s: relative L2 error %.2e" % (name, rel))
        assert rel < 5e-5
    # the image and its adjoint are bit-reproducible (sub-pixel slots are
    # handed out in arrival order; sums do not depend on it)
    a = dm.image(cam, Td)
    b = dm.image(cam, Td)
    assert torch.equal(a, b)
    g_lr2, g_T2 = dm.image_adjoint(cam, Td, None, dd)
    assert torch.equal(g_lr, g_lr2) and torch.equal(g_T, g_T2)


def test_subpixel_pool_grows_and_cap_of_levels():
    """A pool too small for the refinement is grown and the image traced again
    (same result as with a large pool); nrefine=1 allows one level."""
    print("case: sub-pixel pool growth", flush=True)
    arr, T = _disk_setup()
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    from pdspy_amd import _lib
    ctx = _lib.Context(0)
    dm = rt.DeviceModel(arr, ctx=ctx)
    cam = dm.camera(48, 4.0 * AU, 1300.0, incl=60.0, nrefine=100)
    ctx.sync_stream()
    ctx.check(ctx.L.rt_refine_reserve(ctx.h, 48 * 48 * 2))
    small = dm.image(cam, Td)
    n_small, uns = dm.refine_stats()
    ctx.check(ctx.L.rt_refine_reserve(ctx.h, 1 << 24))
    big = dm.image(cam, Td)
    n_big, _ = dm.refine_stats()
    assert uns == 0 and n_small == n_big and torch.equal(small, big)
    one = dm.camera(48, 4.0 * AU, 1300.0, incl=60.0, nrefine=1)
    dm.image(one, Td)
    n1, _ = dm.refine_stats()
    assert 48 * 48 < n1 <= 5 * 48 * 48
