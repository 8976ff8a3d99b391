"""New dust opacities on a resident model (rt_model_set_opacity): a model whose
opacities were replaced runs exactly as a model created with them, also inside a
batch whose step graph was captured before the change; a change of the species
grouping is refused and leaves the model as it was."""
import numpy as np
import pytest
import torch

from pdspy_amd import configs, rt
from pdspy_amd._lib import RTError
from pdspy_amd.dust import DustGenerator
from pdspy_amd.sampler import SamplerStep
from tests.test_gpu_batch import _close, _small_params, _snapshot

pytestmark = pytest.mark.gpu


def _two_species(nr=8, nt=6):
    m = rt.model_arrays(configs.batch_model(dict(mdisk=1e-3, gamma=1.0, plh=1.1, menv=0.05,
                                                 incl=30.0, nr=nr, nt=nt, nphot=1)).grid)
    lam = m["lam_um"]
    m["kabs"] = np.stack([1e3 * (lam / 1.0) ** -1.0, 5e2 * (lam / 1.0) ** -0.7]).clip(max=3e3)
    m["ksca"] = 0.5 * m["kabs"]
    return m


def _generator(lam_um):
    lam = lam_um * 1e-4
    amax = np.logspace(-4, 0, 9)
    p = np.linspace(2.5, 4.5, 5)
    P, A, L = np.meshgrid(p, amax, lam, indexing="ij")
    kabs = 1e3 * (L / 1e-4) ** (-1.0 + 0.1 * (P - 3.5)) * (1 + A)
    return DustGenerator(lam, amax, p, kabs, 0.4 * kabs)


def test_set_opacity_runs_as_a_fresh_model():
    print("case: set_opacity vs fresh model, thermal + image", flush=True)
    base = _two_species()
    dm = rt.DeviceModel(base)
    dm.thermal(20000, 4, max_events=100000, mrw=1)
    gen = _generator(base["lam_um"])
    new = dict(base)
    d0, d1 = gen(3e-2, 3.1), gen(1e-3, 3.9)
    dm.set_dust([d0, d1])
    new["kabs"] = dm._keep["kabs"].copy()
    new["ksca"] = dm._keep["ksca"].copy()
    assert not np.allclose(new["kabs"], base["kabs"])
    fresh = rt.DeviceModel(new)
    for seed in (4, 9):
        s1 = dm.thermal(20000, seed, max_events=100000, mrw=1)
        s2 = fresh.thermal(20000, seed, max_events=100000, mrw=1)
        assert np.array_equal(dm.temperature(), fresh.temperature()), seed
        assert s1["n_abs"] == s2["n_abs"] and s1["n_cross"] == s2["n_cross"]
    T = torch.tensor(dm.temperature()[0], dtype=torch.float32, device="cuda")
    pix = 2.0 * 140 * 1.496e13 / 32
    for lam in (870.0, 1300.0):
        a = dm.image(dm.camera(32, pix, lam, incl=40.0), T)
        b = fresh.image(fresh.camera(32, pix, lam, incl=40.0), T)
        assert torch.equal(a, b), lam


def test_set_opacity_refuses_a_new_grouping():
    print("case: set_opacity that would merge two species", flush=True)
    base = _two_species()
    dm = rt.DeviceModel(base)
    dm.thermal(10000, 2, max_events=100000)
    T0 = dm.temperature()
    same = np.stack([base["kabs"][0], base["kabs"][0]])
    with pytest.raises(RTError):
        dm.set_opacity(same, 0.5 * same)
    with pytest.raises(RTError):
        dm.set_opacity(same, 0.5 * same, on_device=True)
    dm.thermal(10000, 2, max_events=100000)
    assert np.array_equal(dm.temperature(), T0)
    with pytest.raises(RTError):
        dm.set_opacity(-base["kabs"], base["ksca"])


def test_batch_graph_recaptured_after_set_opacity():
    print("case: sampler step graph, set_opacity on one model, replay", flush=True)
    s = SamplerStep(_small_params(), npix=32, K=300, max_events=100000, nphot_scat=20000)
    s.step(seed=3, use_graph=True)
    m0 = s.models[0]
    k = m0._keep["kabs"] * 1.7
    m0.set_opacity(k, m0._keep["ksca"] * 0.6)
    s.step(seed=3, use_graph=False)
    eager = _snapshot(s)
    s.step(seed=3, use_graph=True)
    got = _snapshot(s)
    for i in range(s.n):
        assert np.array_equal(got[0][i], eager[0][i]), i
        assert got[1][i] == pytest.approx(eager[1][i], rel=1e-5), i
        assert _close(got[2][i], eager[2][i], 1e-6), i
    # and equals model 0's single-model calls with its new opacities
    img, c2, _, _ = s.single_model(0, 3 * 1000)
    assert np.array_equal(m0.temperature(), got[0][0])
    assert _close(img.cpu().numpy(), got[2][0], 1e-6)
    assert float(c2.item()) == pytest.approx(got[1][0], rel=1e-5)


def test_sampler_step_draws_dust_per_model():
    """A sampler step with a dust generator: each model's disk dust comes from
    its (a_max, p) draw (run_disk_model.py:79-85); after set_params the graph
    step equals the single-model calls with the new opacities."""
    import time
    print("case: sampler step with DustGenerator, set_params, graph replay", flush=True)
    ps = _small_params()
    lam_um = rt.model_arrays(configs.batch_model(ps[0]).grid)["lam_um"]
    gen = _generator(lam_um)
    rng = np.random.default_rng(3)
    for p in ps:
        p.update(a_max=float(10 ** rng.uniform(0, 3)), p=float(rng.uniform(2.6, 4.4)))
    s = SamplerStep(ps, npix=32, K=300, max_events=100000, nphot_scat=20000, dust_gen=gen)
    want = gen(ps[1]["a_max"] / 1e4, ps[1]["p"])
    assert np.allclose(s.models[1]._keep["kabs"][0], rt.dust_arrays([want], lam_um)[0][0])
    s.step(seed=2, use_graph=True)
    new = [dict(p) for p in ps]
    for p in new:
        p.update(a_max=float(10 ** rng.uniform(0, 3)), p=float(rng.uniform(2.6, 4.4)),
                 mdisk=p["mdisk"] * 1.5)
    t0 = time.perf_counter()
    s.set_params(new)
    print("set_params with new dust for %d models: %.1f ms" % (s.n, (time.perf_counter() - t0) * 1e3))
    s.step(seed=2, use_graph=True)
    got = _snapshot(s)
    for i, m in enumerate(s.models):
        d = gen(new[i]["a_max"] / 1e4, new[i]["p"])
        # the vectorized draw is the single draw's values bit for bit
        assert np.array_equal(m._keep["kabs"][0], rt.dust_arrays([d], lam_um)[0][0])
        assert np.array_equal(m._keep["ksca"][0], rt.dust_arrays([d], lam_um)[1][0])
        img, c2, _, _ = s.single_model(i, 2 * 1000 + i)
        assert np.array_equal(m.temperature(), got[0][i]), i
        assert _close(img.cpu().numpy(), got[2][i], 1e-6), i
        assert float(c2.item()) == pytest.approx(got[1][i], rel=1e-5), i


def test_device_built_tables_match_host_build():
    """rt_model_set_opacity_device builds the emission, re-emission and mean
    tables with kernels; they equal the host build of a model created with the
    same opacities (the re-emission CDFs and guides exactly, the rest to FP32
    rounding), and the thermal MC run on them is that model's to within the
    Monte Carlo tolerance."""
    print("case: device-built opacity tables vs host build", flush=True)
    base = _two_species(nr=24, nt=16)
    gen = _generator(base["lam_um"])
    new = rt.dust_arrays([gen(3e-3, 3.1), gen(1e-4, 3.5)], base["lam_um"])
    fresh = dict(base, kabs=np.array(new[0]), ksca=np.array(new[1]))
    host = rt.DeviceModel(fresh)
    dev = rt.DeviceModel(base)
    # two stream-ordered updates in a row (the second reuses the staging
    # buffer while the first may still be in flight); the last one counts
    other = rt.dust_arrays([gen(1e-2, 2.7), gen(1e-4, 3.5)], base["lam_um"])
    dev.set_opacity(np.array(other[0]), np.array(other[1]), on_device=True)
    dev.set_opacity(np.array(new[0]), np.array(new[1]), on_device=True)
    # the per-species arrays are the host build's bit for bit
    oa, ob = host.opacity_arrays(), dev.opacity_arrays()
    for k in ("ka", "ks", "g", "gka", "ginvk", "econv"):
        assert np.array_equal(oa[k], ob[k]), k
    assert oa["econv"].min() > 0 and oa["ginvk"].min() > 0
    a, b = host.tables(), dev.tables()
    assert np.allclose(a["lnq"], b["lnq"], rtol=1e-12, atol=0)
    d = np.abs(a["cdf"].astype(float) - b["cdf"].astype(float))
    print("cdf max |diff| %.2e, entries differing %d of %d" % (d.max(), (d > 0).sum(), d.size))
    # each row is summed in the host's order from the same FP64 terms: the
    # CDFs and guides come out equal (exp and log of the device and the host
    # may still differ in a last bit elsewhere)
    assert d.max() == 0.0
    assert np.array_equal(a["guide"], b["guide"])
    for k in ("kplanck", "kross"):
        assert np.allclose(a[k], b[k], rtol=3e-7, atol=0), k
    arr = fresh
    sh = host.thermal(50000, 3, max_events=100000, mrw=1, count_visits=1)
    Th, vis = host.temperature(with_visits=True)
    dev.thermal(50000, 3, max_events=100000, mrw=1)
    Td = dev.temperature()
    sel = (vis >= 100)[None, :] & (Th > 0)
    rel = np.abs(Td[sel] - Th[sel]) / Th[sel]
    print("thermal on device tables vs host tables: median %.2e p99 %.2e"
          % (np.median(rel), np.percentile(rel, 99)))
    assert np.median(rel) <= 0.01 and np.percentile(rel, 99) <= 0.05
