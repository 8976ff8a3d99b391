"""The allocation-free gridded sampler (rt_vis_grid_noalloc, scratch in a
static device array) against rt_vis_grid, called through the C ABI on two
calls in a row, the second with a smaller image, and its size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

from pdspy_amd import _lib, configs, inputs

pytestmark = pytest.mark.gpu


def _call(L, ctx, fn, img, dxy, u, v, dra, ddec, extra=()):
    ny, nx = img.shape
    K = u.numel()
    re = torch.empty(K, dtype=torch.float32, device="cuda")
    im = torch.empty(K, dtype=torch.float32, device="cuda")
    rc = fn(ctx.h, C.c_void_p(img.data_ptr()), nx, ny, float(dxy), C.c_void_p(u.data_ptr()),
            C.c_void_p(v.data_ptr()), K, float(dra), float(ddec), C.c_void_p(re.data_ptr()),
            C.c_void_p(im.data_ptr()), *extra)
    return rc, re, im


def test_noalloc_sampler_equals_allocating_one_twice():
    ctx = _lib.context(0)
    L = ctx.L
    rng = np.random.default_rng(12)
    for (ny, nx), K, seed in (((256, 256), 20000, 1), ((48, 64), 3000, 2)):
        print("case: noalloc vis_grid %dx%d, K=%d" % (ny, nx, K), flush=True)
        img = torch.tensor(rng.uniform(size=(ny, nx)), dtype=torch.float32, device="cuda")
        dxy = 0.0078125 * inputs.arcsec * 256 / nx
        u, v = configs.uv_points(K=K, lam_m=1.3e-3, seed=seed)
        # keep the points inside the grid's interpolation range
        umax = (min(nx, ny) // 2 - 2) / (max(nx, ny) * dxy)
        r = np.hypot(u, v)
        u, v = np.where(r < umax, u, u * umax / r * 0.9), np.where(r < umax, v, v * umax / r * 0.9)
        ud = torch.tensor(u, dtype=torch.float64, device="cuda")
        vd = torch.tensor(v, dtype=torch.float64, device="cuda")
        dra, ddec = 0.03 * inputs.arcsec, -0.02 * inputs.arcsec
        rc0, re0, im0 = _call(L, ctx, L.rt_vis_grid, img, dxy, ud, vd, dra, ddec, (None,))
        rc1, re1, im1 = _call(L, ctx, L.rt_vis_grid_noalloc, img, dxy, ud, vd, dra, ddec)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0
        assert torch.isfinite(re1).all() and torch.isfinite(im1).all()
        assert torch.equal(re0, re1) and torch.equal(im0, im1)
    # beyond its scratch: a clear error, nothing launched
    big = torch.zeros((8, 1100), dtype=torch.float32, device="cuda")
    ud = torch.zeros(4, dtype=torch.float64, device="cuda")
    rc, _, _ = _call(L, ctx, L.rt_vis_grid_noalloc, big, 1e-8, ud, ud, 0.0, 0.0)
    assert rc == -6
    assert b"scratch" in L.rt_last_error(ctx.h)
