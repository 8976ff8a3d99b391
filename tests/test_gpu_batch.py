"""Batched sampler step on the GPU: one thermal launch for several models, the
per-model image/visibility/chi^2/adjoint chain, and its hipGraph replay."""
import numpy as np
import pytest

from pdspy_amd import configs, rt
from pdspy_amd.sampler import SamplerStep

pytestmark = pytest.mark.gpu


def _small_params():
    # uneven grids, photon counts, masses and inclinations
    return [
        dict(mdisk=1e-3, gamma=1.0, plh=1.1, menv=0.05, incl=30.0, nr=24, nt=16, nphot=30000),
        dict(mdisk=1e-4, gamma=0.5, plh=1.25, menv=0.01, incl=60.0, nr=16, nt=24, nphot=50000),
        dict(mdisk=3e-3, gamma=1.4, plh=1.0, menv=0.2, incl=80.0, nr=32, nt=12, nphot=20000),
    ]


def test_batch_thermal_equals_single_runs_bitwise():
    print("case: batch of 3 models vs 3 single runs", flush=True)
    ps = _small_params()
    models = [rt.DeviceModel(rt.model_arrays(configs.batch_model(p).grid)) for p in ps]
    opts = dict(max_events=100000, mrw=1)
    seeds = [11, 22, 33]
    single = []
    for m, p, s in zip(models, ps, seeds):
        st = m.thermal(p["nphot"], s, **opts)
        single.append((m.temperature(), st))
    b = rt.Batch(models)
    stats = b.thermal([p["nphot"] for p in ps], seeds, **opts)
    for i, m in enumerate(models):
        T = m.temperature()
        assert np.array_equal(T, single[i][0]), "model %d differs" % i
        for k in ("n_cross", "n_abs", "n_scat", "n_escape", "n_killed", "n_events", "n_mrw"):
            assert stats[i][k] == single[i][1][k], (i, k)
        assert stats[i]["n_epochs"] == single[i][1]["n_epochs"]
    # a second batched launch reproduces itself
    b.thermal([p["nphot"] for p in ps], seeds, **opts)
    for i, m in enumerate(models):
        assert np.array_equal(m.temperature(), single[i][0])


def test_batch_parking_keeps_results_bitwise():
    # many small models with uneven packet counts: waves run out of work in one
    # model while others still hand out epochs, so packets get parked
    print("case: batch of 12 models, parking on vs off", flush=True)
    ps = configs.sampler_batch(12, seed=3)
    for k, p in enumerate(ps):
        p.update(nr=16 + 8 * (k % 3), nt=16, nphot=(20000, 60000, 150000)[k % 3])
    models = [rt.DeviceModel(rt.model_arrays(configs.batch_model(p).grid)) for p in ps]
    nph = [p["nphot"] for p in ps]
    seeds = list(range(100, 112))
    b = rt.Batch(models)
    opts = dict(max_events=100000, mrw=1)
    st0 = b.thermal(nph, seeds, no_park=1, **opts)
    T0 = [m.temperature() for m in models]
    assert sum(x["n_parked"] for x in st0) == 0
    parked = 0
    for rep in range(2):
        st1 = b.thermal(nph, seeds, **opts)
        parked += sum(x["n_parked"] for x in st1)
        for i, m in enumerate(models):
            assert np.array_equal(m.temperature(), T0[i]), (rep, i)
            for k in ("n_cross", "n_abs", "n_scat", "n_escape", "n_killed", "n_events", "n_mrw"):
                assert st1[i][k] == st0[i][k], (rep, i, k)
    print("parked packets over 2 launches: %d" % parked, flush=True)


def _snapshot(s):
    return ([m.temperature() for m in s.models], [float(c.item()) for c in s.chi2],
            [x.cpu().numpy().copy() for x in s.img], [g.cpu().numpy().copy() for g in s.g_T],
            [g.cpu().numpy().copy() for g in s.g_logrho])


def _close(a, b, rtol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


def test_batch_step_matches_single_model_calls():
    import torch
    print("case: sampler step of 3 models, 32x32 images, K=300", flush=True)
    s = SamplerStep(_small_params(), npix=32, K=300, max_events=100000, nphot_scat=20000)
    stats, ms = s.step(seed=5, use_graph=False)
    T_b, chi_b, img_b, gT_b, glr_b = _snapshot(s)
    assert all(st["n_killed"] <= 0.01 * st["n_escape"] for st in stats)
    # the same work through the single-model calls (thermal, scattering
    # source, image, visibilities, chi^2, adjoint)
    seeds = np.arange(s.n) + 5000
    for i, m in enumerate(s.models):
        img, c2, glr, gT = s.single_model(i, int(seeds[i]))
        assert np.array_equal(m.temperature(), T_b[i])
        assert _close(img.cpu().numpy(), img_b[i], 1e-6)
        assert float(c2.item()) == pytest.approx(chi_b[i], rel=1e-5)
        assert _close(gT.cpu().numpy(), gT_b[i], 1e-4)
        assert _close(glr.cpu().numpy(), glr_b[i], 1e-4)
        assert np.abs(gT_b[i]).max() > 0


def test_batch_step_graph_replays_match_eager():
    import torch
    print("case: graph capture + 3 replays vs eager steps", flush=True)
    s = SamplerStep(_small_params(), npix=32, K=300, max_events=100000, nphot_scat=20000)
    ref = {}
    for seed in (7, 8):
        s.step(seed=seed, use_graph=False)
        ref[seed] = _snapshot(s)
    for k, seed in enumerate((7, 8, 7)):
        # eager work between replays
        x = torch.randn(1 << 16, device="cuda")
        (x * 2).sum().item()
        s.models[0].thermal(20000, 99, max_events=100000)
        s.step(seed=seed, use_graph=True)
        got = _snapshot(s)
        exp = ref[seed]
        for i in range(s.n):
            assert np.array_equal(got[0][i], exp[0][i]), (k, i)
            assert got[1][i] == pytest.approx(exp[1][i], rel=1e-5), (k, i)
            assert _close(got[2][i], exp[2][i], 1e-6), (k, i)
            assert _close(got[3][i], exp[3][i], 1e-4), (k, i)
            assert _close(got[4][i], exp[4][i], 1e-4), (k, i)
