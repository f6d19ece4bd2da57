"""Batched tissues over the C ABI (yh_batch_*), and the batched workload."""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import check, lib
from .solver import _ptr, _stream_handle, init_random, make_holes, seed

SIDES = (128, 256, 384, 512, 768)


class Batch:
    """Many tissues of one model advanced by one work-list launch per k steps."""

    def __init__(self, params_list):
        arr = (T.YhParams * len(params_list))(*params_list)
        self.params = [T.copy_params(p) for p in params_list]
        self.nv = T.nvars(params_list[0])
        h = C.c_void_p()
        check(lib().yh_batch_create(arr, len(params_list), C.byref(h)), "yh_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().yh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n, tot = C.c_int(), C.c_int64()
        check(lib().yh_batch_size(self._h, C.byref(n), C.byref(tot)))
        return n.value, tot.value

    def set_mask(self, t, solid):
        m = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        check(lib().yh_batch_set_mask(self._h, t, _ptr(m)), "yh_batch_set_mask")

    def set_state(self, t, *fields):
        fs = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1) for f in fields]
        w = fs[2] if self.nv == 3 else None
        check(lib().yh_batch_set_state(self._h, t, _ptr(fs[0]), _ptr(fs[1]), _ptr(w)),
              "yh_batch_set_state")

    def get_state(self, t):
        p = self.params[t]
        out = [np.empty(p.nx * p.ny, np.float32) for _ in range(self.nv)]
        w = out[2] if self.nv == 3 else None
        check(lib().yh_batch_get_state(self._h, t, _ptr(out[0]), _ptr(out[1]), _ptr(w)),
              "yh_batch_get_state")
        return out

    def set_kernel(self, substeps=8, rows=4, waves=16, cols=1, persistent=1):
        cfg = T.YhKernelCfg(T.YH_KERNEL_TILED, substeps, rows, waves, cols)
        check(lib().yh_batch_set_kernel(self._h, C.byref(cfg), int(persistent)),
              "yh_batch_set_kernel")

    def get_kernel(self):
        cfg, pers = T.YhKernelCfg(), C.c_int()
        check(lib().yh_batch_get_kernel(self._h, C.byref(cfg), C.byref(pers)))
        return (cfg.variant, cfg.substeps, cfg.rows, cfg.waves, cfg.cols), pers.value

    def set_graph(self, on):
        check(lib().yh_batch_set_graph(self._h, int(bool(on))))

    def step(self, nsteps, stream=None):
        check(lib().yh_batch_step(self._h, int(nsteps), _stream_handle(stream)), "yh_batch_step")

    def step_each(self, nsteps, stream=None):
        check(lib().yh_batch_step_each(self._h, int(nsteps), _stream_handle(stream)),
              "yh_batch_step_each")

    def autotune(self, steps=64, stream=None):
        cfg, ms = T.YhKernelCfg(), C.c_float()
        check(lib().yh_batch_autotune(self._h, int(steps), _stream_handle(stream),
                                      C.byref(cfg), C.byref(ms)), "yh_batch_autotune")
        return (cfg.variant, cfg.substeps, cfg.rows, cfg.waves, cfg.cols), ms.value


def batched_workload(ntissues=256, config=3, model=T.YH_MODEL_3V_FK, sides=SIDES,
                     max_holes=40, dx=0.025):
    """The batched configuration: sides uniform over `sides`, every FK time
    constant and D scaled by U[0.8, 1.2], U{0..max_holes} holes with radii
    U[3,12] cells, u ~ U[0,1) and v = w = 1.  Returns (params, masks, fields)."""
    rng = np.random.default_rng(seed(config, 0, 0) & 0xFFFFFFFF)
    params, masks, fields = [], [], []
    taus = ("tau_d", "tau_r", "tau_si", "tau_0", "tau_vp", "tau_v1m", "tau_v2m", "tau_wp",
            "tau_wm")
    for t in range(ntissues):
        n = int(sides[rng.integers(len(sides))])
        D = 0.001 * rng.uniform(0.8, 1.2)
        fk = {name: getattr(T.default_fk(), name) * rng.uniform(0.8, 1.2) for name in taus}
        nh = int(rng.integers(0, max_holes + 1))
        p = T.grid_params(n, n, dx=dx, model=model, solidSwitch=1, diff_par=D, diff_per=D,
                          fk=fk)
        params.append(p)
        masks.append(make_holes(n, n, nh, 3.0, 12.0, seed(config, t, 1)))
        fields.append(init_random(n, n, seed(config, t, 2), gate=1.0, nvars=T.nvars(p)))
    return params, masks, fields
