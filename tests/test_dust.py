"""Dust opacities per model draw: the DustGenerator restatement against scipy's
RegularGridInterpolator (the routine pdspy's DustGenerator.__call__ calls,
dust/DustGenerator.py:84-95) and closed forms."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from pdspy_amd.dust import DustGenerator


def _tables(seed=0):
    rng = np.random.default_rng(seed)
    lam = np.logspace(-5, 0, 40)          # cm
    amax = np.logspace(-4.0, 1.0, 12)     # cm, log-spaced like pdspy's tables
    p = np.linspace(2.5, 4.5, 5)
    kabs = 10 ** rng.uniform(-1, 4, (p.size, amax.size, lam.size))
    ksca = 10 ** rng.uniform(-2, 3, (p.size, amax.size, lam.size))
    return lam, amax, p, kabs, ksca


def test_generator_equals_scipy_regular_grid_interpolator():
    lam, amax, p, kabs, ksca = _tables()
    gen = DustGenerator(lam, amax, p, kabs, ksca)
    fa = RegularGridInterpolator((p, amax, lam), np.log10(kabs))
    fs = RegularGridInterpolator((p, amax, lam), np.log10(ksca))
    rng = np.random.default_rng(1)
    draws = [(a, q) for a, q in zip(10 ** rng.uniform(-4, 1, 30), rng.uniform(2.5, 4.5, 30))]
    draws += [(amax[0], p[0]), (amax[-1], p[-1]), (amax[5], p[2]), (amax[5], 3.3)]
    for a, q in draws:
        d = gen(a, q)
        pts = np.array([[q, a, l] for l in lam])
        assert np.allclose(d.kabs, 10 ** fa(pts), rtol=1e-12, atol=0)
        assert np.allclose(d.ksca, 10 ** fs(pts), rtol=1e-12, atol=0)
        assert np.allclose(d.kext, d.kabs + d.ksca) and np.allclose(d.albedo, d.ksca / d.kext)
        assert np.array_equal(d.lam, lam)


def test_generator_closed_forms():
    lam, amax, p, _, _ = _tables()
    # log10 k linear in p and in amax (not log amax): bilinear interpolation is exact
    P, A, L = np.meshgrid(p, amax, lam, indexing="ij")
    lk = 0.3 * P - 0.05 * A + np.log10(L) * -1.0 + 2.0
    gen = DustGenerator(lam, amax, p, 10 ** lk, 10 ** (lk - 1))
    a, q = 3.7e-2, 3.1
    want = 10 ** (0.3 * q - 0.05 * a - np.log10(lam) + 2.0)
    d = gen(a, q)
    assert np.allclose(d.kabs, want, rtol=1e-12) and np.allclose(d.ksca, want / 10, rtol=1e-12)
    # at grid nodes the tables come back unchanged
    d = gen(amax[3], p[1])
    assert np.allclose(d.kabs, 10 ** lk[1, 3], rtol=1e-13)
    # outside the tables, as RegularGridInterpolator's default bounds_error
    with pytest.raises(ValueError):
        gen(amax[-1] * 1.01, 3.0)
    with pytest.raises(ValueError):
        gen(1e-3, 2.4)


def test_draw_many_equals_single_draws():
    """The vectorized draw of many walkers (a sampler step's new parameters)
    holds the single draws' values bit for bit, and resampled onto a model's
    wavelengths equals dust_arrays of the Dust objects."""
    from pdspy_amd import rt
    lam, amax, p, kabs, ksca = _tables(2)
    gen = DustGenerator(lam, amax, p, kabs, ksca)
    rng = np.random.default_rng(3)
    a = np.concatenate([10 ** rng.uniform(-4, 1, 40), [amax[0], amax[-1], amax[4]]])
    q = np.concatenate([rng.uniform(2.5, 4.5, 40), [p[0], p[-1], 3.3]])
    ka, ks = gen.draw_many(a, q)
    assert ka.shape == (a.size, lam.size)
    for k in range(a.size):
        d = gen(a[k], q[k])
        assert np.array_equal(d.kabs, ka[k]) and np.array_equal(d.ksca, ks[k]), k
    # onto the tables' own wavelengths and onto a different grid (log-log
    # interpolation with power-law ends)
    for lam_um in (lam * 1e4, np.logspace(-1.5, 4.5, 57)):
        for k in (0, 17, a.size - 1):
            want = rt.dust_arrays([gen(a[k], q[k])], lam_um)
            assert np.array_equal(rt.loglog_resample(lam_um, lam * 1e4, ka[k]), want[0][0])
            assert np.array_equal(rt.loglog_resample(lam_um, lam * 1e4, ks[k]), want[1][0])
    with pytest.raises(ValueError):
        gen.draw_many([amax[-1] * 1.01], [3.0])
    with pytest.raises(ValueError):
        gen.draw_many([1e-3, 1e-3], [3.0])
