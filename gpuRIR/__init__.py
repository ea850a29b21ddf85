"""``import gpuRIR`` on AMD MI355X: the API of gpuRIR/__init__.py (DavidDiazGuerra/gpuRIR v1.2.0)
served by gpurir_amd, whose hand-written gfx950 kernels sit behind the C ABI of libgpurir.so.

Scripts written for the reference run unchanged (``import gpuRIR`` then ``gpuRIR.simulateRIR``,
``gpuRIR.simulateTrajectory``, ``gpuRIR.activateMixedPrecision`` ...): the names, arguments,
defaults, assertion messages and output layouts are the reference's.  ``gpuRIR.torch_ops`` adds
the device-resident, batched and differentiable forms.  There is no CPU fallback: importing
fails if libgpurir.so has not been built.
"""
from gpurir_amd import (activate_lut, activate_mixed_precision, activateLUT, activateMixedPrecision,  # noqa: F401
                        att2t_SabineEstimator, beta_SabineEstimation, beta_SabineEstimation_batch, polar_patterns,
                        set_seed, simulateRIR, simulateTrajectory, t2n)
from gpurir_amd import _lib as _native
from gpurir_amd import torch_ops  # noqa: F401

# the reference's list (gpuRIR/__init__.py:11), whose two snake_case names exist here
__all__ = ["polar_patterns", "beta_SabineEstimation", "att2t_SabineEstimator", "t2n", "simulateRIR",
           "simulateTrajectory", "activate_mixed_precision", "activate_lut"]

_native.load()  # fail at import, as the reference does when its compiled module is missing
