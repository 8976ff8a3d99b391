"""Test helpers: independent conversion of pdspy-style models to the oracle's
plain arrays (walls cm/rad/rad, rho [nspec][nphi][ntheta][nr], opacities on
the grid wavelengths), and the comparison statistic of the thermal MC."""
import numpy as np

from pdspy_amd.inputs import AU, R_sun


def oracle_arrays(m):
    g = m.grid
    rho = np.array([np.asarray(d, float).transpose(2, 1, 0) for d in g.density])
    lam = np.asarray(g.lam, float)
    kabs = np.array([np.exp(np.interp(np.log(lam), np.log(d.lam * 1e4), np.log(d.kabs)))
                     for d in g.dust])
    ksca = np.array([np.exp(np.interp(np.log(lam), np.log(d.lam * 1e4), np.log(d.ksca)))
                     for d in g.dust])
    gs = None
    if any(getattr(d, "g", None) is not None for d in g.dust):
        gs = np.array([np.asarray(getattr(d, "g", np.zeros_like(lam)), float) for d in g.dust])
    stars = np.array([[s.radius * R_sun, s.temperature, s.x * AU, s.y * AU, s.z * AU]
                      for s in g.stars])
    return dict(rw=np.asarray(g.w1, float) * AU, tw=np.asarray(g.w2, float),
                pw=np.asarray(g.w3, float), rho=rho, lam_um=lam, kabs=kabs, ksca=ksca,
                gsca=gs, stars=stars)


def dT_stats(T_ref, T, visits, min_visits=100):
    """median and 99th percentile of |dT|/T over cells with >= min_visits entries
    and a positive reference temperature (all species)."""
    T_ref = np.asarray(T_ref, float)
    T = np.asarray(T, float)
    sel = (np.asarray(visits) >= min_visits)[None, :] & (T_ref > 0)
    d = np.abs(T[sel] - T_ref[sel]) / T_ref[sel]
    return float(np.median(d)), float(np.percentile(d, 99)), int(sel.sum())
