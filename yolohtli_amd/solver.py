"""Host-side mirror of the reference solver surface over the C ABI.

Names follow the reference where a reference function exists:
``initGates`` (main.cu:577-621), ``domainObjects`` hole masks (main.cu:623-684),
``reactionDiffusion_wrapper`` (reactionDiffusion.cu:568-577) and
``tip_wrapper`` (tipTracker.cu:569-611).  ``Plan`` owns device state the way
main.cu's globals do, and steps it as display() does (main.cu:876-885).
Errors raise ``YhError`` carrying the C error code; nothing here computes on
the CPU.
"""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import check, lib

BASE_SEED = 20261202


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def params(nx=512, ny=512, dx=None, **over):
    """yh_params with the reference defaults; dx (cm) sets Lx = dx*(nx-1)."""
    if dx is not None:
        return T.grid_params(nx, ny, dx=dx, **over)
    return T.default_params(nx=nx, ny=ny, **over)


def derive(p):
    d = T.YhDerived()
    check(lib().yh_params_derive(C.byref(p), C.byref(d)), "yh_params_derive")
    return d


def seed(config, tissue=0, purpose=0, base=BASE_SEED):
    return int(lib().yh_seed(base, config, tissue, purpose))


def initGates(nx, ny, u_cols=None, rest=0.0, refr=1.0, nvars=2):
    """Cross-field IC of main.cu:577-621 (u_cols defaults to nx/8)."""
    u = np.empty(nx * ny, np.float32)
    v = np.empty(nx * ny, np.float32)
    w = np.empty(nx * ny, np.float32) if nvars == 3 else None
    cols = nx // 8 if u_cols is None else u_cols
    check(lib().yh_init_cross_field(nx, ny, cols, rest, refr, _ptr(u), _ptr(v), _ptr(w)))
    return (u, v, w) if nvars == 3 else (u, v)


def init_random(nx, ny, seed_value, gate=1.0, nvars=3):
    u = np.empty(nx * ny, np.float32)
    v = np.empty(nx * ny, np.float32)
    w = np.empty(nx * ny, np.float32) if nvars == 3 else None
    check(lib().yh_init_random(nx, ny, seed_value, gate, _ptr(u), _ptr(v), _ptr(w)))
    return (u, v, w) if nvars == 3 else (u, v)


def make_holes(nx, ny, nholes, rmin=3.0, rmax=12.0, seed_value=0):
    m = np.empty(nx * ny, np.uint8)
    check(lib().yh_make_holes(nx, ny, nholes, rmin, rmax, seed_value, _ptr(m)))
    return m


def _dev_ptr(t):
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("device arrays must be contiguous GPU tensors")
        return t.data_ptr()
    return int(t)


def _stream_handle(stream):
    if stream is None:
        try:
            import torch
            if torch.cuda.is_available():
                return torch.cuda.current_stream().cuda_stream
        except Exception:  # pragma: no cover
            pass
        return None
    return getattr(stream, "cuda_stream", stream)


def reactionDiffusion_wrapper(p, gOut, gIn, velTan=None, solid=None, stream=None,
                              stimLockMouse=False, point=(0, 0)):
    """One time step on caller-owned device fields (lists of tensors u,v[,w]);
    stimLockMouse/point: the mouse stimulus of reactionDiffusion.cu:54-61."""
    nv = T.nvars(p)
    arr = C.c_void_p * 3
    out = arr(*([_dev_ptr(x) for x in gOut[:nv]] + [None] * (3 - nv)))
    inp = arr(*([_dev_ptr(x) for x in gIn[:nv]] + [None] * (3 - nv)))
    vt = None
    if velTan is not None:
        vt = arr(*([_dev_ptr(x) for x in velTan[:nv]] + [None] * (3 - nv)))
    check(lib().yh_rd_step_dev_ex(C.byref(p), out, inp, vt, _dev_ptr(solid),
                                  int(bool(stimLockMouse)), int(point[0]), int(point[1]),
                                  _stream_handle(stream)), "yh_rd_step_dev_ex")


def tip_wrapper(p, g_past, g_present, time, tips, keys, count, cap, stream=None):
    """Append spiral tips between two u fields (device tensors) to tips/keys/count."""
    check(lib().yh_tip_detect_dev(C.byref(p), _dev_ptr(g_past), _dev_ptr(g_present),
                                  float(time), _dev_ptr(tips), _dev_ptr(keys),
                                  _dev_ptr(count), int(cap), _stream_handle(stream)),
          "yh_tip_detect_dev")


def countour_wrapper(nx, ny, field1, field2, area, pts, keys, count, cap, physicalTime, mode,
                     thresholds=(0.8, 0.85, 0.7), stream=None):
    """Contour points between two device fields (spaceAPD.cu:257-275, the
    reference's spelling): pts (cap, 3) float32, keys (cap,) int32 and count
    (1,) int32 device tensors; the count is zeroed first, as there."""
    check(lib().yh_contour_dev(int(nx), int(ny), int(mode), *(float(x) for x in thresholds),
                               _dev_ptr(field1), _dev_ptr(field2), _dev_ptr(area),
                               float(physicalTime), _dev_ptr(pts), _dev_ptr(keys),
                               _dev_ptr(count), int(cap), _stream_handle(stream)),
          "yh_contour_dev")


def tips_to_array(tips):
    """(n,5) float32 array x, y, vx, vy, t from a ctypes YhTip array or device bytes."""
    return np.ctypeslib.as_array(tips).view(np.float32).reshape(-1, 5) if len(tips) else \
        np.zeros((0, 5), np.float32)


class Plan:
    """A tissue on one GPU: ping-pong state, mask, tip log."""

    def __init__(self, p):
        self.p = T.copy_params(p)
        self.nv = T.nvars(p)
        self.n = p.nx * p.ny
        h = C.c_void_p()
        check(lib().yh_plan_create(C.byref(self.p), C.byref(h)), "yh_plan_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().yh_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def derived(self):
        d = T.YhDerived()
        check(lib().yh_plan_params(self._h, None, C.byref(d)))
        return d

    def set_mask(self, solid):
        m = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        assert m.size == self.n
        check(lib().yh_set_mask(self._h, _ptr(m)), "yh_set_mask")

    def set_state(self, *fields):
        fs = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1) for f in fields]
        assert len(fs) >= self.nv and all(f.size == self.n for f in fs[:self.nv])
        w = fs[2] if self.nv == 3 else None
        check(lib().yh_set_state(self._h, _ptr(fs[0]), _ptr(fs[1]), _ptr(w)), "yh_set_state")

    def get_state(self):
        out = [np.empty(self.n, np.float32) for _ in range(self.nv)]
        w = out[2] if self.nv == 3 else None
        check(lib().yh_get_state(self._h, _ptr(out[0]), _ptr(out[1]), _ptr(w)), "yh_get_state")
        return out

    def state_dev(self):
        u, v, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().yh_state_dev(self._h, C.byref(u), C.byref(v), C.byref(w)))
        return u.value, v.value, w.value

    def set_kernel(self, variant=T.YH_KERNEL_AUTO, substeps=8, rows=8, waves=16, cols=1):
        cfg = T.YhKernelCfg(variant, substeps, rows, waves, cols)
        check(lib().yh_set_kernel(self._h, C.byref(cfg)), "yh_set_kernel")

    def get_kernel(self):
        cfg = T.YhKernelCfg()
        check(lib().yh_get_kernel(self._h, C.byref(cfg)))
        return cfg

    def step_count(self):
        """Steps taken since the last set_state (param.count)."""
        return int(lib().yh_step_count(self._h))

    def autotune(self, steps=400, stream=None):
        """Times the tiled shapes and depths on the current state (the state
        and step count are restored) and installs the fastest (yh_autotune).
        steps is the call size the plan will be driven with: each candidate
        runs whole step(steps) calls, so a call's remainder launch counts.
        Returns ((variant, K, rows, waves, cols), ms per step)."""
        cfg, ms = T.YhKernelCfg(), C.c_float()
        check(lib().yh_autotune(self._h, int(steps), _stream_handle(stream), C.byref(cfg),
                                C.byref(ms)), "yh_autotune")
        return (cfg.variant, cfg.substeps, cfg.rows, cfg.waves, cfg.cols), ms.value

    def step(self, nsteps, stream=None):
        check(lib().yh_step(self._h, int(nsteps), _stream_handle(stream)), "yh_step")

    @property
    def count(self):
        return int(lib().yh_step_count(self._h))

    def track_tips(self, stream=None):
        check(lib().yh_track_tips(self._h, _stream_handle(stream)), "yh_track_tips")

    def read_tips(self, cap=500000):
        buf = (T.YhTip * cap)()
        n = C.c_int()
        rc = lib().yh_read_tips(self._h, buf, cap, C.byref(n))
        if rc not in (0, T.YH_EOVERFLOW):
            check(rc, "yh_read_tips")
        m = min(n.value, cap)
        return np.ctypeslib.as_array(buf).view(np.float32).reshape(-1, 5)[:m].copy()

    def clear_tips(self):
        check(lib().yh_clear_tips(self._h))

    # ---- symmetry reduction (main.cu:894-953)
    def sr_begin(self, tip0=(0.0, 0.0), stream=None):
        check(lib().yh_sr_begin(self._h, float(tip0[0]), float(tip0[1]),
                                _stream_handle(stream)), "yh_sr_begin")

    def sr_step(self, nsteps, stream=None, history=False):
        """nsteps SR steps; with history, returns (nsteps, 6) c and phi per step."""
        out = np.empty((nsteps, 6)) if history else None
        check(lib().yh_sr_step(self._h, int(nsteps), _stream_handle(stream),
                               _ptr(out)), "yh_sr_step")
        return out

    def sr_set_graph(self, on):
        check(lib().yh_sr_set_graph(self._h, int(bool(on))))

    def sr_info(self):
        info = T.YhSrInfo()
        rc = lib().yh_sr_info_get(self._h, C.byref(info))
        if rc not in (0, T.YH_ESINGULAR):
            check(rc, "yh_sr_info_get")
        return {"c": np.array(info.c[:]), "phi": np.array(info.phi[:]),
                "I": np.array(info.I[:]), "cx": info.cx, "cy": info.cy,
                "count": info.count, "status": info.status, "ntips": info.ntips}

    # ---- action-potential durations (spaceAPD.cu:278-374)
    def apd_enable(self, threshold=0.15, on=True):
        check(lib().yh_apd_enable(self._h, int(bool(on)), float(threshold)), "yh_apd_enable")

    def apd_get(self):
        """dict of per-cell apd1, apd2, front, back (float32)."""
        out = {k: np.empty(self.n, np.float32) for k in ("apd1", "apd2", "front", "back")}
        check(lib().yh_apd_get(self._h, _ptr(out["apd1"]), _ptr(out["apd2"]), _ptr(out["front"]),
                               _ptr(out["back"])), "yh_apd_get")
        return out

    def set_stimulus(self, on, point=(0, 0)):
        """Mouse stimulus (stimLockMouse, point) for the following steps."""
        check(lib().yh_set_stimulus(self._h, int(bool(on)), int(point[0]), int(point[1])),
              "yh_set_stimulus")

    def contour(self, mode, thresholds=(0.8, 0.85, 0.7), area=None, cap=None):
        """(n, 3) contour points (x, y, t) of the current state in cell order
        (countour_kernel, spaceAPD.cu:18-153); thresholds = conTh1..3."""
        cap = cap or 2 * self.n
        out = np.empty((cap, 3), np.float32)
        n = C.c_int()
        a = None if area is None else np.ascontiguousarray(area, np.uint8).reshape(-1)
        rc = lib().yh_contour(self._h, int(mode), *(float(x) for x in thresholds), _ptr(a),
                              _ptr(out), cap, C.byref(n))
        if rc not in (0, T.YH_EOVERFLOW):
            check(rc, "yh_contour")
        return out[:min(n.value, cap)].copy()

    def set_option(self, option, value):
        check(lib().yh_set_option(self._h, int(option), int(value)), "yh_set_option")

    def check_finite(self, stream=None):
        return lib().yh_check_finite(self._h, _stream_handle(stream))
