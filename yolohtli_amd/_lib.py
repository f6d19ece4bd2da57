"""Loads libyolohtli_amd.so and declares the C-ABI prototypes.

The library must have been built (``python -m yolohtli_amd.build`` or
``__graft_entry__.build()``).  There is no fallback: every product call goes
through the HIP library, and a missing library raises immediately.
"""
import ctypes as C
import os

from .abi_types import YhDerived, YhKernelCfg, YhParams, YhTip, YhFkParams, YhSrInfo  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyolohtli_amd.so")

_lib = None

P = C.c_void_p
I = C.c_int
D = C.c_double
F_P = C.POINTER(C.c_float)
PARAMS_P = C.POINTER(YhParams)

# name -> (restype, argtypes)
PROTOTYPES = {
    "yh_strerror": (C.c_char_p, [I]),
    "yh_params_default": (None, [PARAMS_P]),
    "yh_fk_params_default": (None, [C.POINTER(YhFkParams)]),
    "yh_params_derive": (I, [PARAMS_P, C.POINTER(YhDerived)]),
    "yh_seed": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "yh_init_cross_field": (I, [I, I, I, C.c_float, C.c_float, P, P, P]),
    "yh_init_random": (I, [I, I, C.c_uint64, C.c_float, P, P, P]),
    "yh_make_holes": (I, [I, I, I, D, D, C.c_uint64, P]),
    "yh_rd_step_dev": (I, [PARAMS_P, P, P, P, P, P]),
    "yh_rd_step_dev_ex": (I, [PARAMS_P, P, P, P, P, I, I, I, P]),
    "yh_step_noalloc": (I, [PARAMS_P, P, P, P, I, P]),
    "yh_tip_detect_dev": (I, [PARAMS_P, P, P, D, P, P, P, I, P]),
    "yh_contour_dev": (I, [I, I, I, D, D, D, P, P, P, C.c_float, P, P, P, I, P]),
    "yh_contour": (I, [P, I, D, D, D, P, P, I, C.POINTER(I)]),
    "yh_plan_create": (I, [PARAMS_P, C.POINTER(P)]),
    "yh_plan_destroy": (None, [P]),
    "yh_plan_params": (I, [P, PARAMS_P, C.POINTER(YhDerived)]),
    "yh_set_mask": (I, [P, P]),
    "yh_set_state": (I, [P, P, P, P]),
    "yh_get_state": (I, [P, P, P, P]),
    "yh_state_dev": (I, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "yh_set_kernel": (I, [P, C.POINTER(YhKernelCfg)]),
    "yh_get_kernel": (I, [P, C.POINTER(YhKernelCfg)]),
    "yh_step": (I, [P, I, P]),
    "yh_step_count": (C.c_int64, [P]),
    "yh_track_tips": (I, [P, P]),
    "yh_read_tips": (I, [P, P, I, C.POINTER(I)]),
    "yh_clear_tips": (I, [P]),
    "yh_check_finite": (I, [P, P]),
    "yh_set_option": (I, [P, I, C.c_int64]),
    "yh_set_stimulus": (I, [P, I, I, I]),
    "yh_apd_enable": (I, [P, I, D]),
    "yh_apd_get": (I, [P, P, P, P, P]),
    "yh_autotune": (I, [P, I, P, C.POINTER(YhKernelCfg), C.POINTER(C.c_float)]),
    "yh_sr_begin": (I, [P, C.c_float, C.c_float, P]),
    "yh_sr_step": (I, [P, I, P, P]),
    "yh_sr_set_graph": (I, [P, I]),
    "yh_sr_info_get": (I, [P, C.POINTER(YhSrInfo)]),
    "yh_solve_matrix": (I, [P, P, P]),
    "yh_batch_create": (I, [PARAMS_P, I, C.POINTER(P)]),
    "yh_batch_destroy": (None, [P]),
    "yh_batch_size": (I, [P, C.POINTER(I), C.POINTER(C.c_int64)]),
    "yh_batch_set_mask": (I, [P, I, P]),
    "yh_batch_set_state": (I, [P, I, P, P, P]),
    "yh_batch_get_state": (I, [P, I, P, P, P]),
    "yh_batch_set_kernel": (I, [P, C.POINTER(YhKernelCfg), I]),
    "yh_batch_get_kernel": (I, [P, C.POINTER(YhKernelCfg), C.POINTER(I)]),
    "yh_batch_set_graph": (I, [P, I]),
    "yh_batch_step": (I, [P, I, P]),
    "yh_batch_step_each": (I, [P, I, P]),
    "yh_batch_autotune": (I, [P, I, P, C.POINTER(YhKernelCfg), C.POINTER(C.c_float)]),
    "yh_adjoint_create": (I, [PARAMS_P, I, I, I, P, I, C.POINTER(P)]),
    "yh_adjoint_destroy": (None, [P]),
    "yh_adjoint_set_substeps": (I, [P, I]),
    "yh_adjoint_set_mask": (I, [P, P]),
    "yh_adjoint_set_initial": (I, [P, P, P, P]),
    "yh_adjoint_set_targets": (I, [P, P]),
    "yh_adjoint_forward": (I, [P, P, P, P, C.POINTER(D), P]),
    "yh_adjoint_backward": (I, [P, P, P, P]),
    "yh_params_write_csv": (I, [C.c_char_p, PARAMS_P, C.c_char_p, C.c_char_p, C.c_float,
                                C.c_float]),
    "yh_params_read_csv": (I, [C.c_char_p, PARAMS_P]),
    "yh_state_write_dat": (I, [C.c_char_p, I, I, P, P, P]),
    "yh_state_read_dat": (I, [C.c_char_p, I, I, P, P, P]),
    "yh_tips_write_dat": (I, [C.c_char_p, C.c_char_p, P, I, I]),
    "yh_sym_write_dat": (I, [C.c_char_p, P, I]),
    "yh_mask_read_dat": (I, [C.c_char_p, I, I, P]),
    "yh_mask_write_dat": (I, [C.c_char_p, I, I, P]),
}


class YhError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().yh_strerror(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m yolohtli_amd.build); there is no CPU fallback")
        try:  # share torch's HIP runtime when torch is present (one runtime per process)
            import torch  # noqa: F401
        except ImportError:
            pass
        h = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(code, what=""):
    if code != 0:
        raise YhError(code, what)
    return code
