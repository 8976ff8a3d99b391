"""Writes the small fixtures under tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).  Each one is data: inputs and the
outputs a library routine or closed form gives for them, so that a change of
numpy's generator, scipy's root finder or of this package's restatements shows
up as a test failure.

* uv_points_c3.npz  -- the first 256 (u, v) points of the C3 visibility set
  (configs.uv_points(K=5e4, seed=0): baselines log-uniform in 15-2000 m,
  position angle uniform, lambda = 1.3 mm) and the sums over all 5e4.
* ulrich_mu0.npz    -- roots mu0 of mu0^3 - mu0 (1 - r/Rc) - mu r/Rc = 0 from
  scipy.optimize.brenth on [-1, 1] (the call UlrichEnvelope.py:64-68 makes)
  for r/Rc in [0.01, 20] and mu in [-0.98, 0.98], including the three-real-root
  region r < Rc.
* disk_sigma.npz    -- the normalised surface density of the C2 disk
  (Disk.py:48-70: 1e-3 Msun, Sigma ~ r^-0.5 out to 100 au) at 64 radii.
"""
import os
import sys

import numpy as np
from scipy.optimize import brenth

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pdspy_amd import configs  # noqa: E402
from pdspy_amd.inputs import Disk  # noqa: E402


def main():
    u, v = configs.uv_points(50000, seed=0)
    np.savez(os.path.join(HERE, "uv_points_c3.npz"), u=u[:256], v=v[:256],
             sum_u=u.sum(), sum_v=v.sum(), sum_b=np.hypot(u, v).sum())
    rr = np.concatenate([np.linspace(0.01, 0.99, 25), np.linspace(1.01, 20.0, 15)])
    mu = np.linspace(-0.98, 0.98, 41)  # open interval: at mu = +-1 a root sits on the bracket
    roots = np.array([[brenth(lambda x: x ** 3 - x * (1 - r) - m * r, -1.0, 1.0) for m in mu]
                      for r in rr])
    np.savez(os.path.join(HERE, "ulrich_mu0.npz"), r_over_rc=rr, mu=mu, mu0=roots)
    d = Disk(mass=1e-3, rmin=0.1, rmax=100.0, plrho=1.75, h0=0.1, plh=1.25)
    r = np.logspace(-1.2, 2.2, 64)
    np.savez(os.path.join(HERE, "disk_sigma.npz"), r_au=r, sigma=d.surface_density(r))


if __name__ == "__main__":
    main()
