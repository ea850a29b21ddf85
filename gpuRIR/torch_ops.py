"""``gpuRIR.torch_ops``: the device-resident, batched and differentiable API (gpurir_amd.torch_ops)."""
from gpurir_amd.torch_ops import (BatchPlan, DiffPlan, image_list, n_samples, options, simulate_batch,  # noqa: F401
                                  simulate_batch_diff, simulate_rir, simulate_rir_diff, trajectory,
                                  TrajectoryPlan)
