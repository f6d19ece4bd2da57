"""Python face of the native autotuner (yh_autotune in csrc/autotune.cpp)."""
import ctypes as C

from . import abi_types as T
from ._lib import check, lib
from .solver import _stream_handle


def tune_plan(plan, fields=None, steps=400, stream=None, log=None):
    """Times the candidate kernels on the plan's state and installs the fastest.

    steps: the plan's call size (each candidate runs whole calls of it).
    Returns {"cfg": (variant, K, rows, waves, cols), "ms_per_step": t}."""
    cfg = T.YhKernelCfg()
    ms = C.c_float()
    check(lib().yh_autotune(plan._h, int(steps), _stream_handle(stream), C.byref(cfg),
                            C.byref(ms)), "yh_autotune")
    out = {"cfg": (cfg.variant, cfg.substeps, cfg.rows, cfg.waves, cfg.cols),
           "ms_per_step": ms.value}
    if log:
        log(f"autotune: variant={cfg.variant} K={cfg.substeps} rows={cfg.rows} "
            f"waves={cfg.waves} cols={cfg.cols}: {ms.value * 1e3:.3f} us/step")
    return out
