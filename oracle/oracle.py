"""ctypes wrapper of liboracle.so, the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from yolohtli_amd.abi_types import YhParams, YhTip, nvars

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


class OrcCoef(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "hx", "hy", "dt", "Dxx", "Dyy", "rx", "ry", "invdx", "invdy")]


class OrcSrState(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("phi", C.c_double * 3), ("ac", C.c_double * 3),
                ("aphi", C.c_double), ("cx", C.c_int), ("cy", C.c_int),
                ("tip0x", C.c_float), ("tip0y", C.c_float), ("count", C.c_long),
                ("ntips", C.c_int), ("I", C.c_double * 12)]


def build():
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        P = C.c_void_p
        PP = C.POINTER(YhParams)
        h.orc_derive.argtypes = [PP, C.POINTER(OrcCoef)]
        h.orc_coord.argtypes = [C.c_int, C.c_int]
        h.orc_coord.restype = C.c_int
        for nm in ("orc_rd_step_d", "orc_rd_step_f"):
            getattr(h, nm).argtypes = [PP, P, P, P, P]
        for nm in ("orc_run_d", "orc_run_f"):
            getattr(h, nm).argtypes = [PP, P, P, C.c_int, C.c_int]
        h.orc_tips_d.argtypes = [PP, P, P, C.c_double, P, P, C.c_int]
        h.orc_tips_d.restype = C.c_int
        h.orc_slice_field_d.argtypes = [PP, P, C.c_int, C.c_int, P, C.c_double, P, P, P, P, P, P]
        h.orc_sr_integrals_d.argtypes = [PP, P, P, C.c_int, C.c_int, P, C.c_double, P]
        h.orc_solve_matrix.argtypes = [P, P, P]
        h.orc_bfecc_d.argtypes = [PP, P, P, P, P, C.c_double]
        h.orc_sr_init.argtypes = [C.POINTER(OrcSrState), C.c_float, C.c_float]
        h.orc_sr_step_d.argtypes = [PP, P, P, C.POINTER(OrcSrState)]
        h.orc_sr_step_d.restype = C.c_int
        h.orc_apd_update_d.argtypes = [C.c_int, C.c_double, C.c_long, P, P, P, P, P, P, P]
        h.orc_rd_step_forced_d.argtypes = [PP, P, P, P, P, P, C.c_int]
        h.orc_run_apd_d.argtypes = [PP, P, P, C.c_int, C.c_long, C.c_double, P, P, P, P, P]
        h.orc_rd_step_general_d.argtypes = [PP, P, C.c_int, C.c_int, C.c_int, P, P, P]
        h.orc_rd_step_general_d.restype = C.c_int
        h.orc_contour_d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, P, P, P, C.c_float, P, P, C.c_int]
        h.orc_contour_d.restype = C.c_int
        h.orc_fk_dfield_step_d.argtypes = [PP, P, P, P, P]
        h.orc_fk_loss_d.argtypes = [PP, P, P, P, C.c_int, C.c_int, C.c_int, P, P, P]
        h.orc_fk_loss_d.restype = C.c_double
        h.orc_fk_adjoint_d.argtypes = [PP, P, P, P, C.c_int, C.c_int, C.c_int, P, P, P, P]
        h.orc_fk_adjoint_d.restype = C.c_double
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _arr3(fields, dtype):
    fs = [np.ascontiguousarray(f, dtype=dtype).reshape(-1) for f in fields]
    return fs, (C.c_void_p * 3)(*([_p(f) for f in fs] + [None] * (3 - len(fs))))


def derive(p):
    c = OrcCoef()
    lib().orc_derive(C.byref(p), C.byref(c))
    return c


def coord(k, n):
    return lib().orc_coord(k, n)


def rd_step(p, fields, solid=None, dtype=np.float64, veltan=False):
    """One Euler step; returns (out_fields, veltan_fields or None)."""
    nv = nvars(p)
    n = p.nx * p.ny
    fin, cin = _arr3(fields[:nv], dtype)
    out = [np.empty(n, dtype) for _ in range(nv)]
    cout = (C.c_void_p * 3)(*([_p(f) for f in out] + [None] * (3 - nv)))
    vt = [np.empty(n, dtype) for _ in range(nv)] if veltan else None
    cvt = (C.c_void_p * 3)(*([_p(f) for f in vt] + [None] * (3 - nv))) if veltan else None
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    fn = lib().orc_rd_step_d if dtype == np.float64 else lib().orc_rd_step_f
    fn(C.byref(p), _p(m), cin, cout, cvt)
    return out, vt


def run(p, fields, nsteps, solid=None, dtype=np.float64, nthreads=None):
    """nsteps of the plain loop; returns new fields (inputs untouched)."""
    nv = nvars(p)
    fs = [np.array(f, dtype=dtype).reshape(-1) for f in fields[:nv]]
    cst = (C.c_void_p * 3)(*([_p(f) for f in fs] + [None] * (3 - nv)))
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    fn = lib().orc_run_d if dtype == np.float64 else lib().orc_run_f
    fn(C.byref(p), _p(m), cst, int(nsteps), int(nthreads or os.cpu_count() or 1))
    return fs


def rd_step_forced(p, fields, force_p, force_q, solid=None, nthreads=None):
    """One FP64 Euler step of the FK model with imposed Heaviside decisions:
    force_p / force_q (int8 per cell) = 0 or 1 replace STEP(u - u_c) /
    STEP(u - u_v), -1 keeps the cell's own decision."""
    nv = nvars(p)
    n = p.nx * p.ny
    fin, cin = _arr3(fields[:nv], np.float64)
    out = [np.empty(n) for _ in range(nv)]
    cout = (C.c_void_p * 3)(*([_p(f) for f in out] + [None] * (3 - nv)))
    fp = np.ascontiguousarray(force_p, dtype=np.int8).reshape(-1)
    fq = np.ascontiguousarray(force_q, dtype=np.int8).reshape(-1)
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    lib().orc_rd_step_forced_d(C.byref(p), _p(m), cin, cout, _p(fp), _p(fq),
                               int(nthreads or os.cpu_count() or 1))
    return out


def rd_step_general(p, fields, solid=None, stim=None, veltan=False):
    """One step of the whole reference kernel (every boundary, stencil and
    time-integration branch), FP64.  stim: (x, y) of the mouse stimulus."""
    nv = nvars(p)
    fs = [np.ascontiguousarray(f, dtype=np.float64).reshape(-1) for f in fields[:nv]]
    n = fs[0].size
    outs = [np.empty(n) for _ in range(nv)]
    vts = [np.empty(n) for _ in range(nv)] if veltan else None
    cin = (C.c_void_p * 3)(*([_p(f) for f in fs] + [None] * (3 - nv)))
    cout = (C.c_void_p * 3)(*([_p(f) for f in outs] + [None] * (3 - nv)))
    cvt = (C.c_void_p * 3)(*([_p(f) for f in vts] + [None] * (3 - nv))) if veltan else None
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    sx, sy = (int(stim[0]), int(stim[1])) if stim is not None else (0, 0)
    rc = lib().orc_rd_step_general_d(C.byref(p), _p(m), int(stim is not None), sx, sy, cin, cout,
                                     cvt)
    if rc != 0:
        raise ValueError("unsupported timeIntOrder")
    return (outs, vts) if veltan else outs


def contour(nx, ny, mode, th, f1, f2, area=None, t=0.0, cap=200000):
    """(n, 3) contour points (x, y, t) in (cell, sub) order and their keys."""
    a = np.ascontiguousarray(f1, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(f2, dtype=np.float64).reshape(-1)
    m = None if area is None else np.ascontiguousarray(area, dtype=np.uint8).reshape(-1)
    out = np.zeros((cap, 3), np.float32)
    keys = np.zeros(cap, np.int32)
    n = lib().orc_contour_d(nx, ny, mode, float(th[0]), float(th[1]), float(th[2]), _p(a), _p(b),
                            _p(m), float(t), _p(out), _p(keys), cap)
    if n < 0:
        raise ValueError("unknown contour mode")
    return out[:min(n, cap)].copy(), keys[:min(n, cap)].copy()


def run_general(p, fields, nsteps, solid=None, stim=None):
    X = [np.asarray(f, np.float64) for f in fields]
    for _ in range(nsteps):
        X = rd_step_general(p, X, solid=solid, stim=stim)
    return X


def tips(p, g_past, g_present, time, cap=100000):
    """(n,5) array of tips in (cell, root) order, and their keys."""
    a = np.ascontiguousarray(g_past, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(g_present, dtype=np.float64).reshape(-1)
    out = (YhTip * cap)()
    keys = np.zeros(cap, np.int32)
    n = lib().orc_tips_d(C.byref(p), _p(a), _p(b), float(time), out, _p(keys), cap)
    m = min(n, cap)
    arr = np.ctypeslib.as_array(out).view(np.float32).reshape(-1, 5)[:m].copy()
    return arr, keys[:m].copy()


def slice_field(p, f, cx, cy, ac=(0.0, 0.0, 0.0), aphi=0.0):
    n = p.nx * p.ny
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    outs = [np.zeros(n) for _ in range(6)]
    acv = np.array(ac, np.float64)
    lib().orc_slice_field_d(C.byref(p), _p(f), int(cx), int(cy), _p(acv), float(aphi),
                            *[_p(o) for o in outs])
    return outs


def sr_integrals(p, X, vt, cx, cy, ac=(0.0, 0.0, 0.0), aphi=0.0):
    nv = nvars(p)
    _, cX = _arr3(X[:nv], np.float64)
    fX = _
    _, cV = _arr3(vt[:nv], np.float64)
    I = np.zeros(12)
    acv = np.array(ac, np.float64)
    lib().orc_sr_integrals_d(C.byref(p), cX, cV, int(cx), int(cy), _p(acv), float(aphi), _p(I))
    del fX
    return I


def solve_matrix(phi, I):
    phi = np.array(phi, np.float64)
    I = np.array(I, np.float64)
    c = np.zeros(3)
    lib().orc_solve_matrix(_p(phi), _p(I), _p(c))
    return c


def bfecc(p, fields, c, phit, solid=None):
    nv = nvars(p)
    n = p.nx * p.ny
    fin, cin = _arr3(fields[:nv], np.float64)
    out = [np.empty(n) for _ in range(nv)]
    cout = (C.c_void_p * 3)(*([_p(f) for f in out] + [None] * (3 - nv)))
    cc = np.array(c, np.float64)
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    lib().orc_bfecc_d(C.byref(p), _p(m), cin, cout, _p(cc), float(phit))
    return out


class SR:
    """Symmetry-reduction session of the oracle (main.cu:894-953)."""

    def __init__(self, p, fields, tip0=(0.0, 0.0), solid=None):
        self.p = p
        self.nv = nvars(p)
        self.X = [np.array(f, dtype=np.float64).reshape(-1) for f in fields[:self.nv]]
        self.solid = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        self.s = OrcSrState()
        lib().orc_sr_init(C.byref(self.s), float(tip0[0]), float(tip0[1]))

    def step(self, n=1):
        rc = 0
        for _ in range(n):
            cX = (C.c_void_p * 3)(*([_p(f) for f in self.X] + [None] * (3 - self.nv)))
            rc = lib().orc_sr_step_d(C.byref(self.p), _p(self.solid), cX, C.byref(self.s))
        return rc

    @property
    def c(self):
        return np.array(self.s.c[:])

    @property
    def phi(self):
        return np.array(self.s.phi[:])


def fk_dfield_step(p, D, fields, solid=None):
    n = p.nx * p.ny
    fin, cin = _arr3(fields[:3], np.float64)
    out = [np.empty(n) for _ in range(3)]
    cout = (C.c_void_p * 3)(*[_p(f) for f in out])
    Dv = np.ascontiguousarray(D, np.float64).reshape(-1)
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    lib().orc_fk_dfield_step_d(C.byref(p), _p(m), _p(Dv), cin, cout)
    return out


def fk_loss(p, D, X0, nsteps, rec, sensors, targets=None, solid=None, traces=False):
    fin, cin = _arr3(X0[:3], np.float64)
    Dv = np.ascontiguousarray(D, np.float64).reshape(-1)
    s = np.ascontiguousarray(sensors, np.int32)
    t = None if targets is None else np.ascontiguousarray(targets, np.float64).reshape(-1)
    nrec = nsteps // rec
    tr = np.zeros(nrec * len(s)) if traces else None
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    L = lib().orc_fk_loss_d(C.byref(p), _p(m), _p(Dv), cin, int(nsteps), int(rec), len(s),
                            _p(s), _p(t), _p(tr))
    return (L, tr.reshape(nrec, len(s))) if traces else L


def fk_adjoint(p, D, X0, nsteps, rec, sensors, targets, solid=None):
    fin, cin = _arr3(X0[:3], np.float64)
    Dv = np.ascontiguousarray(D, np.float64).reshape(-1)
    s = np.ascontiguousarray(sensors, np.int32)
    t = np.ascontiguousarray(targets, np.float64).reshape(-1)
    gD = np.zeros(p.nx * p.ny)
    gt = np.zeros(6)
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    L = lib().orc_fk_adjoint_d(C.byref(p), _p(m), _p(Dv), cin, int(nsteps), int(rec), len(s),
                               _p(s), _p(t), _p(gD), _p(gt))
    return L, gD, gt


def apd_update(dt, count, uold, unew, st):
    """st: dict of apd1, apd2, front, back (float64) and first (uint8) arrays, updated in place."""
    n = uold.size
    lib().orc_apd_update_d(n, float(dt), int(count), _p(np.ascontiguousarray(uold, np.float64)),
                           _p(np.ascontiguousarray(unew, np.float64)), _p(st["apd1"]),
                           _p(st["apd2"]), _p(st["front"]), _p(st["back"]), _p(st["first"]))


def run_apd(p, fields, nsteps, solid=None, th=0.15, count0=0, state=None):
    """nsteps with the APD update; returns (fields, apd state dict)."""
    nv = nvars(p)
    n = p.nx * p.ny
    fs = [np.array(f, dtype=np.float64).reshape(-1) for f in fields[:nv]]
    cst = (C.c_void_p * 3)(*([_p(f) for f in fs] + [None] * (3 - nv)))
    st = state or {k: np.zeros(n) for k in ("apd1", "apd2", "front", "back")}
    if "first" not in st:
        st["first"] = np.zeros(n, np.uint8)
    m = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
    lib().orc_run_apd_d(C.byref(p), _p(m), cst, int(nsteps), int(count0), float(th),
                        _p(st["apd1"]), _p(st["apd2"]), _p(st["front"]), _p(st["back"]),
                        _p(st["first"]))
    return fs, st
