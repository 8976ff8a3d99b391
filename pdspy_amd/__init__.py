"""pdspy_amd -- MI355X (gfx950) continuum radiative transfer for pdspy models.

A drop-in beside pdspy's RADMC-3D path: ``run_thermal``, ``run_image``,
``run_visibilities`` and ``interpolate_model`` take pdspy ``Model`` objects and
fill the same outputs; the work runs in hand-written HIP kernels reached
through a C ABI (``include/pdspy_amd.h``).  ``GPUModelMixin`` adds
``code="mi355x"`` to a pdspy Model subclass.  ``lnlike_vis`` is pdspy's
visibility log-likelihood.  ``pdspy_amd.sampler`` runs the batched sampler
step, ``pdspy_amd.dust`` draws dust opacities per model as pdspy's
DustGenerator does, ``pdspy_amd.radmc3d`` reads and writes RADMC-3D's files.
"""
from .rt import (Batch, DeviceModel, GPUModelMixin, chisq, dust_arrays, fft_grid_uv,
                 imaging_temperature, interpolate_model, lnlike_vis, model_arrays, run_image,
                 run_thermal, run_visibilities, vis_dft, vis_dft_adjoint, vis_grid)

__all__ = ["Batch", "DeviceModel", "GPUModelMixin", "chisq", "dust_arrays", "fft_grid_uv",
           "imaging_temperature", "interpolate_model", "lnlike_vis", "model_arrays", "run_image",
           "run_thermal", "run_visibilities", "vis_dft", "vis_dft_adjoint", "vis_grid"]
