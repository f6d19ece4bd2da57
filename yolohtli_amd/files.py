"""The reference's text files (parameter CSV, raw state, tip log, symmetry
history, obstacle mask), written and read by the C library with the
reference's printf formats.  Names follow the reference's functions:
printParameters / loadParamValues (printFunctions.cu:266-403,
saveFiles.cu:540-711), print2D2column / loadData (printFunctions.cu:58-78,
saveFiles.cu:512-536), printTip (printFunctions.cu:149-195), printSym
(printFunctions.cu:249-264) and the mask read of main.cu:663-681.
"""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import check, lib


def _b(path):
    return str(path).encode()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def printParameters(path, p, ic_path=None, results_path=None, tip=(0.0, 0.0)):
    """dataparamcsv.csv for parameters p."""
    check(lib().yh_params_write_csv(_b(path), C.byref(p),
                                    None if ic_path is None else _b(ic_path),
                                    None if results_path is None else _b(results_path),
                                    float(tip[0]), float(tip[1])), "yh_params_write_csv")


def loadParamValues(path, p=None):
    """Solver parameters from dataparamcsv.csv (fields the file does not hold
    come from p, or the defaults)."""
    q = T.copy_params(p) if p is not None else T.default_params()
    check(lib().yh_params_read_csv(_b(path), C.byref(q)), "yh_params_read_csv")
    return q


def print2D2column(path, u, v, w=None):
    """raw_data.dat: 'u v' (or 'u v w') per cell, i fastest."""
    u = np.ascontiguousarray(u, np.float32).reshape(-1)
    v = np.ascontiguousarray(v, np.float32).reshape(-1)
    w = None if w is None else np.ascontiguousarray(w, np.float32).reshape(-1)
    n = u.size
    if v.size != n or (w is not None and w.size != n):
        raise ValueError("fields differ in size")
    check(lib().yh_state_write_dat(_b(path), n, 1, _ptr(u), _ptr(v), _ptr(w)),
          "yh_state_write_dat")


def loadData(path, nx, ny, nvars=2):
    """Fields written by print2D2column, as float32 arrays of nx*ny."""
    out = [np.empty(nx * ny, np.float32) for _ in range(nvars)]
    w = out[2] if nvars > 2 else None
    check(lib().yh_state_read_dat(_b(path), nx, ny, _ptr(out[0]), _ptr(out[1]), _ptr(w)),
          "yh_state_read_dat")
    return out


def printTip(tip_path, size_path, tips, append=False):
    """Append (or start) dataTip.dat / dataTipSize.dat with tips (n, 5)."""
    t = np.ascontiguousarray(tips, np.float32).reshape(-1, 5)
    check(lib().yh_tips_write_dat(_b(tip_path), _b(size_path), _ptr(t) if len(t) else None,
                                  len(t), int(bool(append))), "yh_tips_write_dat")


def printSym(path, c_phi):
    """c_phi_list_sym.dat from the (nsteps, 6) history of Plan.sr_step."""
    a = np.ascontiguousarray(c_phi, np.float64).reshape(-1, 6)
    check(lib().yh_sym_write_dat(_b(path), _ptr(a) if len(a) else None, len(a)),
          "yh_sym_write_dat")


def read_mask(path, nx, ny):
    m = np.empty(nx * ny, np.uint8)
    check(lib().yh_mask_read_dat(_b(path), nx, ny, _ptr(m)), "yh_mask_read_dat")
    return m


def write_mask(path, mask, nx, ny):
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
    if m.size != nx * ny:
        raise ValueError("mask size")
    check(lib().yh_mask_write_dat(_b(path), nx, ny, _ptr(m)), "yh_mask_write_dat")
