"""torch.autograd wrappers around the HIP image, visibility and chi^2 kernels.

Gradients of an image (and of model visibilities / chi^2 built on it) with
respect to per-cell log-density (per species) and temperature, with the
temperature held fixed from the thermal MC.  pdspy has no counterpart; the
chi^2 form is pdspy's (pdspy/utils/emcee.py:31-43,
interferometry/libinterferometry.pyx:619-633).
"""
import torch

from . import rt


class RTImage(torch.autograd.Function):
    """img[ny, nx] (erg/s/cm^2/Hz/sr) = formal solution through the model.
    Inputs: logrho [nspec, ncell] (float32, ln g/cm^3), T [ncell] (K)."""

    @staticmethod
    def forward(ctx, logrho, T, dm, cam):
        rho = torch.exp(logrho.detach().float()).contiguous()
        Tc = T.detach().float().contiguous()
        img = dm.image(cam, Tc, rho)
        ctx.save_for_backward(rho, Tc)
        ctx.dm = dm
        ctx.cam = cam
        return img

    @staticmethod
    def backward(ctx, gimg):
        rho, T = ctx.saved_tensors
        g_lr, g_T = ctx.dm.image_adjoint(ctx.cam, T, rho, gimg)
        return g_lr, g_T, None, None


class RTVis(torch.autograd.Function):
    """(re, im)[K] = DFT of img at (u, v); dxy, dra, ddec in radians."""

    @staticmethod
    def forward(ctx, img, u, v, dxy, dra=0.0, ddec=0.0):
        re, im = rt.vis_dft(img.detach(), dxy, u, v, dra, ddec)
        ctx.save_for_backward(u, v)
        ctx.shape = img.shape
        ctx.par = (dxy, dra, ddec)
        return re, im

    @staticmethod
    def backward(ctx, gre, gim):
        u, v = ctx.saved_tensors
        dxy, dra, ddec = ctx.par
        ny, nx = ctx.shape
        if gre is None:
            gre = torch.zeros(u.numel(), dtype=torch.float32, device="cuda")
        if gim is None:
            gim = torch.zeros(u.numel(), dtype=torch.float32, device="cuda")
        g = rt.vis_dft_adjoint(gre, gim, nx, ny, dxy, u, v, dra, ddec)
        return g, None, None, None, None, None


class RTChi2(torch.autograd.Function):
    """chi^2 = sum w ((dre - mre)^2 + (dim - mim)^2), float64 scalar."""

    @staticmethod
    def forward(ctx, mre, mim, dre, dim, w):
        out, gre, gim = rt.chisq(dre, dim, w, mre.detach(), mim.detach(), with_grad=True)
        ctx.save_for_backward(gre, gim)
        return out[0]

    @staticmethod
    def backward(ctx, g):
        gre, gim = ctx.saved_tensors
        s = g.float()
        return gre * s, gim * s, None, None, None


def image(logrho, T, dm, cam):
    return RTImage.apply(logrho, T, dm, cam)


def visibilities(img, u, v, dxy, dra=0.0, ddec=0.0):
    return RTVis.apply(img, u, v, dxy, dra, ddec)


def chi2(mre, mim, dre, dim, w):
    return RTChi2.apply(mre, mim, dre, dim, w)
