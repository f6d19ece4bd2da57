"""yh_step_noalloc (no device allocation) through the C ABI against yh_step.

yh_step_noalloc runs the kernels of a plan left on YH_KERNEL_AUTO with its
stencil codes and intermediate state in the library's fixed scratch array;
tiled results do not depend on where launches split, so its output must
equal the plan's state bit for bit.  Calls follow each other with shrinking
grids (the scratch keeps the previous call's contents), and the caller's
input fields must be left untouched.
"""
import ctypes as C

import numpy as np
import pytest

import yolohtli_amd as Y
from yolohtli_amd import abi_types as T
from yolohtli_amd._lib import lib

pytestmark = pytest.mark.gpu


def _ptrs(ts, nv):
    return (C.c_void_p * 3)(*([t.data_ptr() for t in ts[:nv]] + [None] * (3 - nv)))


def _noalloc(torch, p, f, solid, nsteps):
    nv = T.nvars(p)
    gin = [torch.from_numpy(x.copy()).cuda() for x in f]
    gout = [torch.full_like(x, float("nan")) for x in gin]
    sd = torch.from_numpy(solid).cuda() if solid is not None else None
    torch.cuda.synchronize()
    rc = lib().yh_step_noalloc(C.byref(p), _ptrs(gout, nv), _ptrs(gin, nv),
                               sd.data_ptr() if sd is not None else None, nsteps, None)
    assert rc == 0, lib().yh_strerror(rc).decode()
    torch.cuda.synchronize()
    for a in range(nv):
        assert np.array_equal(gin[a].cpu().numpy(), f[a]), "input field written"
    return [x.cpu().numpy() for x in gout]


def _plan(p, f, solid, nsteps):
    pl = Y.Plan(p)
    if solid is not None:
        pl.set_mask(solid)
    pl.set_state(*f)
    pl.step(nsteps)
    out = pl.get_state()
    pl.close()
    return out


CASES = [
    # (nx, ny, model, holes, steps): the headline grid first, then smaller ones
    (1024, 1024, T.YH_MODEL_3V_FK, 200, 50),
    (300, 200, T.YH_MODEL_3V_FK, 12, 37),
    (130, 90, T.YH_MODEL_2V_REF, 0, 23),
]


def test_noalloc_equals_plan_step_on_successive_calls(torch_cuda):
    torch = torch_cuda
    for nx, ny, model, holes, steps in CASES:
        print(f"yh_step_noalloc {nx}x{ny} model {model} holes {holes} steps {steps}", flush=True)
        p = T.grid_params(nx, ny, model=model, solidSwitch=int(holes > 0))
        rng = np.random.default_rng(nx + ny)
        f = [rng.random(nx * ny).astype(np.float32) for _ in range(T.nvars(p))]
        solid = Y.make_holes(nx, ny, holes, 3.0, 12.0, Y.seed(1, 0, nx)) if holes else None
        got = _noalloc(torch, p, f, solid, steps)
        want = _plan(p, f, solid, steps)
        for a in range(len(f)):
            assert np.array_equal(got[a], want[a]), (nx, ny, a)


def test_noalloc_edge_cases(torch_cuda):
    torch = torch_cuda
    p = T.grid_params(64, 48, model=T.YH_MODEL_3V_FK)
    rng = np.random.default_rng(5)
    f = [rng.random(64 * 48).astype(np.float32) for _ in range(3)]
    got = _noalloc(torch, p, f, None, 0)  # no steps: a copy
    for a in range(3):
        assert np.array_equal(got[a], f[a])
    # grids beyond the scratch are refused before anything runs
    big = T.grid_params(2048, 1025, model=T.YH_MODEL_3V_FK)
    arr = (C.c_void_p * 3)(1, 2, 3)
    arr2 = (C.c_void_p * 3)(4, 5, 6)
    assert lib().yh_step_noalloc(C.byref(big), arr, arr2, None, 1, None) == T.YH_ESCRATCH
    # another branch than Euler + 5-point + Neumann is refused
    rk = T.grid_params(64, 48, model=T.YH_MODEL_3V_FK, timeIntOrder=4)
    assert lib().yh_step_noalloc(C.byref(rk), arr, arr2, None, 1, None) == T.YH_EUNSUPPORTED
    # a mask is required with solidSwitch; in and out must not alias
    ms = T.grid_params(64, 48, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    assert lib().yh_step_noalloc(C.byref(ms), arr, arr2, None, 1, None) == T.YH_EINVAL
    assert lib().yh_step_noalloc(C.byref(p), arr, arr, None, 1, None) == T.YH_EINVAL
