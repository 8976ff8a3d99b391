"""Every benchmark configuration at its stated size on the GPU, against the
FP64 oracle (same Philox streams and epoch schedule) or closed properties:

* C1 -- flared disk, 64x64 cells, 1e6 packets, seed 0;
* C2 -- disk + envelope, 256x128 cells: 1e6 packets against the oracle, and
  the full 1e8-packet run twice (bitwise repeat, packet accounting, the
  ping-pong guard having moved some packets);
* C3 -- the 256x256 1.3 mm image of the C2 model (2" field at 140 pc, 60 deg)
  with and without RADMC-3D's sub-pixeling, its visibilities at the 5e4 (u,v)
  points, and the chi^2 gradient against the FP64 oracle's and against FP64
  central differences on 50 cells;
* C4 -- the 64-model batched sampler step: batch == single runs bitwise for a
  sample of models, and one model's chi^2 chain against the oracle;
* C5 -- the 256x128x128 spiral disk (HG g = 0.6) at reduced nphot against the
  oracle, and a 512x512 image and its adjoint against the oracle.
Tolerances: temperatures median |dT|/T <= 1 %, p99 <= 5 % on cells with >= 100
visits; images and gradients as stated per test."""
import os

import numpy as np
import pytest
import torch

import oracle
from pdspy_amd import configs, inputs, rt
from tests.helpers import dT_stats, oracle_arrays

pytestmark = pytest.mark.gpu
AU = inputs.AU
NTHR = min(16, os.cpu_count() or 1)


def _gpu_T(arr, nphot, seed, **opts):
    dm = rt.DeviceModel(arr)
    st = dm.thermal(nphot, seed, count_visits=1, **opts)
    T, vis = dm.temperature(with_visits=True)
    return dm, T, vis, st


def _imaging_T(arr, T):
    rho = np.array(arr["rho"], dtype=float).reshape(T.shape[0], -1)
    rho = np.where(rho == 0, 1e-30, rho)
    return (rho * T).sum(0) / rho.sum(0)


def test_c1_disk_64x64_1e6_packets_vs_oracle():
    print("case: C1 disk 64x64, 1e6 packets, seed 0", flush=True)
    arr = oracle_arrays(configs.disk_c1())
    n = 1000000
    _, Tg, vg, sg = _gpu_T(arr, n, 0, max_events=1000000)
    To, vo, so = oracle.thermal(arr, n, seed=0, max_events=1000000, lagged=True, nthreads=NTHR)
    med, p99, ncell = dT_stats(To, Tg, vo)
    print("C1: median %.2e p99 %.2e over %d cells" % (med, p99, ncell), flush=True)
    assert ncell > 2000
    assert med <= 0.01 and p99 <= 0.05
    assert sg["n_escape"] + sg["n_killed"] == n and sg["n_epochs"] == so["n_epochs"]
    for k in ("n_abs", "n_cross"):
        assert sg[k] == pytest.approx(so[k], rel=0.02), k


def test_c2_disk_envelope_1e6_packets_vs_oracle():
    print("case: C2 disk+envelope 256x128, MRW, 1e6 packets, seed 0", flush=True)
    arr = oracle_arrays(configs.disk_envelope())
    n = 1000000
    _, Tg, vg, sg = _gpu_T(arr, n, 0, max_events=1000000, mrw=1)
    To, vo, so = oracle.thermal(arr, n, seed=0, max_events=1000000, lagged=True, mrw=True,
                                nthreads=NTHR)
    med, p99, ncell = dT_stats(To, Tg, vo)
    print("C2: median %.2e p99 %.2e over %d cells; MRW steps gpu %d oracle %d"
          % (med, p99, ncell, sg["n_mrw"], so["n_mrw"]), flush=True)
    assert ncell > 10000
    assert med <= 0.01 and p99 <= 0.05
    assert sg["n_escape"] + sg["n_killed"] == n
    assert sg["n_mrw"] == pytest.approx(so["n_mrw"], rel=0.05)


def test_c2_full_1e8_run_repeats_bitwise():
    print("case: C2 disk+envelope 256x128, MRW, 1e8 packets, seed 0, twice", flush=True)
    dm = rt.DeviceModel(rt.model_arrays(configs.disk_envelope().grid))
    n = 100000000
    st = []
    Ts = []
    for rep in range(2):
        st.append(dm.thermal(n, 0, max_events=1000000, mrw=1))
        Ts.append(dm.temperature())
        print("C2 1e8 run %d: %.0f ms, %d killed, %d nudged" % (rep, st[-1]["kernel_ms"],
                                                              st[-1]["n_killed"],
                                                              st[-1]["n_nudged"]), flush=True)
    assert np.array_equal(Ts[0], Ts[1])
    for s in st:
        assert s["n_escape"] + s["n_killed"] == n
        assert s["n_killed"] <= 1e-5 * n
    # near-tangent crossings occur at this size: the guard moved packets on and
    # every run still ended normally (a stalled packet would end in the watchdog)
    assert st[0]["n_nudged"] > 0 and st[0]["n_nudged"] == st[1]["n_nudged"]
    assert st[0]["watchdog_fired"] == 0


@pytest.fixture(scope="module")
def c3():
    """C2 model with GPU temperatures from 1e6 packets, its imaging temperature
    and the C3 camera"""
    arr = oracle_arrays(configs.disk_envelope())
    dm, T, vis, st = _gpu_T(arr, 1000000, 3, max_events=1000000, mrw=1)
    Timg = _imaging_T(arr, T)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    pix = 2.0 / 256 * 140 * AU
    return dict(arr=arr, dm=dm, Timg=Timg, kap=kap, pix=pix,
                Td=torch.tensor(Timg, dtype=torch.float32, device="cuda"))


@pytest.mark.parametrize("nrefine", [0, 100])
def test_c3_image_256_vs_oracle(c3, nrefine):
    print("case: C3 image 256x256 at 1.3 mm, 60 deg, nrefine %d" % nrefine, flush=True)
    dm = c3["dm"]
    cam = dm.camera(256, c3["pix"], 1300.0, incl=60.0, incl_star=True, nrefine=nrefine)
    img = dm.image(cam, c3["Td"]).double().cpu().numpy()
    ref, nr = oracle.image(c3["arr"], c3["Timg"], [c3["kap"], c3["kap"]], 256, c3["pix"],
                           incl=60.0, lam_um=1300.0, incl_star=True, nrefine=nrefine,
                           with_nrays=True)
    d = np.abs(img - ref) / np.abs(ref).max()
    err = d.max()
    j, k = np.unravel_index(np.argmax(d), d.shape)
    print("C3 image: %d rays, max rel err %.2e at pixel (%d, %d), 99.9th pct %.2e, flux rel err "
          "%.2e" % (nr, err, j, k, np.percentile(d, 99.9), abs(img.sum() / ref.sum() - 1)),
          flush=True)
    # FP32 rays re-anchored on the exact line at radial crossings: < 1e-6
    # of the peak observed
    assert err < 2e-5
    assert img.sum() == pytest.approx(ref.sum(), rel=1e-6)
    if nrefine:
        ng, uns = dm.refine_stats()
        assert uns == 0 and abs(ng - nr) <= 0.002 * nr


def test_c3_visibilities_and_chi2_gradient_vs_oracle_and_fd(c3):
    print("case: C3 visibilities K=5e4, chi^2 gradient, FD on 50 cells", flush=True)
    dm, arr = c3["dm"], c3["arr"]
    kaps = [c3["kap"], c3["kap"]]
    cam = dm.camera(256, c3["pix"], 1300.0, incl=60.0, incl_star=True)
    img_d = dm.image(cam, c3["Td"])
    img = img_d.double().cpu().numpy()
    dxy = 2.0 / 256 * inputs.arcsec
    u, v = configs.uv_points(K=50000, lam_m=1.3e-3, seed=0)
    re, im = rt.vis_dft(img_d, dxy, u, v)
    ore, oim = oracle.vis_dft(img, dxy, u, v)
    scale = np.abs(ore + 1j * oim).max()
    verr = max(np.abs(re.double().cpu().numpy() - ore).max(),
               np.abs(im.double().cpu().numpy() - oim).max()) / scale
    print("C3 vis max |dV| / max|V| = %.2e" % verr, flush=True)
    assert verr < 2e-5
    # noisy target: model + N(0, 1 % of the peak); chi^2 gradient through the
    # DFT adjoint to the image, then the image adjoint
    rng = np.random.default_rng(11)
    sig = 0.01 * scale
    dre = ore + rng.normal(0, sig, ore.size)
    dim = oim + rng.normal(0, sig, oim.size)
    w = np.full(ore.size, 1.0 / sig ** 2)
    gre, gim = 2 * w * (ore - dre), 2 * w * (oim - dim)
    dimg = oracle.vis_dft_adjoint(gre, gim, 256, 256, dxy, u, v)
    g_lr, g_T = dm.image_adjoint(cam, c3["Td"], None,
                                 torch.tensor(dimg, dtype=torch.float32, device="cuda"))
    g_lr = g_lr.double().cpu().numpy()
    g_T = g_T.double().cpu().numpy()
    o_lr, o_T = oracle.image_adjoint(arr, c3["Timg"], kaps, dimg, 256, c3["pix"], incl=60.0,
                                     lam_um=1300.0, incl_star=True)
    e_lr = np.linalg.norm(g_lr - o_lr) / np.linalg.norm(o_lr)
    e_T = np.linalg.norm(g_T - o_T) / np.linalg.norm(o_T)
    print("C3 gradient vs FP64 oracle: logrho %.2e T %.2e" % (e_lr, e_T), flush=True)
    assert e_lr < 5e-5 and e_T < 5e-5
    # FP64 central differences of L = sum(dimg * image) (first order in chi^2)
    tot = o_lr.sum(0)
    cand = np.argsort(-np.abs(tot))[:500]
    cells = rng.choice(cand, 50, replace=False)

    def L(rho=None, T=None):
        a = dict(arr)
        if rho is not None:
            a["rho"] = rho
        return np.sum(dimg * oracle.image(a, c3["Timg"] if T is None else T, kaps, 256,
                                          c3["pix"], incl=60.0, lam_um=1300.0, incl_star=True))

    h = 1e-3
    rho0 = np.array(arr["rho"], dtype=float)
    worst = 0.0
    for c in cells:
        rp, rm = rho0.copy().reshape(2, -1), rho0.copy().reshape(2, -1)
        rp[:, c] *= np.exp(h)
        rm[:, c] *= np.exp(-h)
        fd = (L(rho=rp.reshape(rho0.shape)) - L(rho=rm.reshape(rho0.shape))) / (2 * h)
        worst = max(worst, abs(g_lr[:, c].sum() - fd) / abs(fd))
        assert o_lr[:, c].sum() == pytest.approx(fd, rel=1e-4)
    print("C3 FD on 50 cells: worst GPU relative error %.2e" % worst, flush=True)
    assert worst < 5e-5
    for c in cells[:10]:
        Tp, Tm = c3["Timg"].copy(), c3["Timg"].copy()
        dT = 1e-3 * Tp[c]
        Tp[c] += dT
        Tm[c] -= dT
        fdT = (L(T=Tp) - L(T=Tm)) / (2 * dT)
        assert g_T[c] == pytest.approx(fdT, rel=1e-4, abs=1e-6 * np.abs(g_T).max())


def test_c4_batched_step_64_models():
    from pdspy_amd.sampler import LAM_UM, SamplerStep
    print("case: C4 batched sampler step, 64 models, 256x256 images, K=5e4", flush=True)
    params = configs.sampler_batch(64, seed=0)
    s = SamplerStep(params, npix=256, K=50000, max_events=1000000, mrw=1, nphot_scat=0)
    stats, ms = s.step(seed=1, use_graph=True)
    print("C4 step: %.0f ms on the device, %d packets" % (ms, int(s.nphot.sum())), flush=True)
    T_b = [m.temperature() for m in s.models]
    chi_b = [float(c.item()) for c in s.chi2]
    img_b = [x.double().cpu().numpy() for x in s.img]
    assert all(np.isfinite(chi_b))
    # one model's chain against the oracle: image of its batch temperatures,
    # DFT at the data's (u,v), chi^2 against the data
    i = int(np.argmin([m.ncell for m in s.models]))
    arr = rt.model_arrays(configs.batch_model(s.params[i]).grid)
    Timg = _imaging_T(arr, T_b[i])
    kaps = [rt._loglog_interp(np.array([LAM_UM]), arr["lam_um"], arr["kabs"][k])[0]
            for k in range(arr["rho"].shape[0])]
    oimg = oracle.image(arr, Timg, kaps, 256, s.pix_cm, incl=s.params[i]["incl"], lam_um=LAM_UM,
                        incl_star=True)
    e_img = np.abs(img_b[i] - oimg).max() / np.abs(oimg).max()
    u = s.u_vis.cpu().numpy()
    v = s.v.cpu().numpy()
    ore, oim = oracle.vis_dft(oimg * s.scale, s.dxy, u, v, 0.0, s.ddec)
    dre = s.re[i].double().cpu().numpy()
    dim = s.im[i].double().cpu().numpy()
    w = s.w[i].double().cpu().numpy()
    chi_o = float(np.sum(w * ((dre - ore) ** 2 + (dim - oim) ** 2)))
    print("C4 model %d: image rel err %.2e, chi^2 gpu %.6e oracle %.6e"
          % (i, e_img, chi_b[i], chi_o), flush=True)
    assert e_img < 2e-5
    assert chi_b[i] == pytest.approx(chi_o, rel=1e-5)
    # batch == single runs, bit for bit, for three models of different sizes
    order = np.argsort([int(n) for n in s.nphot])
    seeds = np.arange(s.n, dtype=np.uint64) + np.uint64(1000)
    for j in (int(order[0]), int(order[s.n // 2]), int(order[-1])):
        m = s.models[j]
        m.thermal(int(s.nphot[j]), int(seeds[j]), **s.opts)
        assert np.array_equal(m.temperature(), T_b[j]), j
    assert all(st["n_escape"] + st["n_killed"] == int(n) for st, n in zip(stats, s.nphot))


def test_c5_spiral_3d_full_grid_vs_oracle():
    print("case: C5 spiral 256x128x128, HG g=0.6, MRW, 1e6 packets, seed 1", flush=True)
    arr = oracle_arrays(configs.spiral_3d())
    assert arr["gsca"] is not None
    n = 1000000
    dm, Tg, vg, sg = _gpu_T(arr, n, 1, max_events=1000000, mrw=1)
    To, vo, so = oracle.thermal(arr, n, seed=1, max_events=1000000, lagged=True, mrw=True,
                                nthreads=NTHR)
    med, p99, ncell = dT_stats(To, Tg, vo)
    print("C5: median %.2e p99 %.2e over %d cells" % (med, p99, ncell), flush=True)
    assert ncell > 500
    assert med <= 0.01 and p99 <= 0.05
    assert sg["n_escape"] + sg["n_killed"] == n
    print("case: C5 512x512 image at 1.3 mm and its adjoint", flush=True)
    T = np.maximum(To[0], 1.0)
    kap = rt._loglog_interp(np.array([1300.0]), arr["lam_um"], arr["kabs"][0])[0]
    pix = 2.0 / 512 * 140 * AU
    Td = torch.tensor(T, dtype=torch.float32, device="cuda")
    cam = dm.camera(512, pix, 1300.0, incl=60.0, pa=20.0, incl_star=True)
    img = dm.image(cam, Td).double().cpu().numpy()
    ref = oracle.image(arr, T, [kap], 512, pix, incl=60.0, posang=20.0, lam_um=1300.0,
                       incl_star=True)
    err = np.abs(img - ref).max() / np.abs(ref).max()
    print("C5 image max rel err %.2e" % err, flush=True)
    assert err < 2e-5 and img.sum() == pytest.approx(ref.sum(), rel=1e-6)
    rng = np.random.default_rng(2)
    dimg = rng.normal(size=(512, 512))
    g_lr, g_T = dm.image_adjoint(cam, Td, None, torch.tensor(dimg, dtype=torch.float32,
                                                            device="cuda"))
    o_lr, o_T = oracle.image_adjoint(arr, T, [kap], dimg, 512, pix, incl=60.0, posang=20.0,
                                     lam_um=1300.0, incl_star=True)
    e_lr = np.linalg.norm(g_lr.double().cpu().numpy() - o_lr) / np.linalg.norm(o_lr)
    e_T = np.linalg.norm(g_T.double().cpu().numpy() - o_T) / np.linalg.norm(o_T)
    print("C5 adjoint vs oracle: logrho %.2e T %.2e" % (e_lr, e_T), flush=True)
    assert e_lr < 5e-5 and e_T < 5e-5
