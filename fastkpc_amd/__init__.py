"""MI355X (gfx950) implementation of the kernel-PC conditional-independence
hot path: regrXonS GAM residualisation with GCV smoothing selection and the
kpcalg dcov.gamma, hsic.gamma, hsic.perm and dcc.perm tests, behind the
reference's skeleton API."""
from .api import (ci_batch, dcov_components, fastkpc_compatible_cuda_skeleton, gcv_grad_hess,
                  kernelCItest, last_fit_timing, permutation_table, regrXonS, regrxons_batch, regrxons_mixed,
                  resolve_max_conditioning_size, task_seed, tuning_report, empty_cache,
                  cache_control, cache_state, fastkpc_orient_wanpdag,
                  fastkpc_compatible_cuda_wanpdag)

__all__ = ["fastkpc_compatible_cuda_skeleton", "kernelCItest", "regrXonS", "regrxons_batch",
           "regrxons_mixed", "last_fit_timing", "tuning_report", "empty_cache",
           "ci_batch", "gcv_grad_hess", "dcov_components", "resolve_max_conditioning_size",
           "task_seed", "permutation_table", "cache_control", "cache_state",
           "fastkpc_orient_wanpdag", "fastkpc_compatible_cuda_wanpdag"]
