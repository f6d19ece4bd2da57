"""yolohtli_amd: MI355X-native cardiac tissue time-stepper (HIP kernels behind a C ABI).

The compute path is libyolohtli_amd.so (hand-written gfx950 kernels).  This
package is the host-side mirror of the reference solver surface; it loads the
library on first use and raises if it is missing.
"""
from . import abi_types  # noqa: F401
from .abi_types import (  # noqa: F401
    YH_KERNEL_AUTO, YH_KERNEL_DIRECT, YH_KERNEL_TILED, YH_MODEL_2V_REF, YH_MODEL_3V_FK,
    default_params, grid_params, copy_params, nvars)
from .solver import (  # noqa: F401
    BASE_SEED, Plan, countour_wrapper, derive, initGates, init_random, make_holes, params,
    reactionDiffusion_wrapper, seed, tip_wrapper)

__all__ = [
    "Plan", "params", "derive", "seed", "initGates", "init_random", "make_holes",
    "reactionDiffusion_wrapper", "tip_wrapper", "countour_wrapper", "default_params", "grid_params",
    "copy_params", "nvars", "BASE_SEED",
]
