"""The C ABI called from compiled C: tests/c/shim_driver links the
reference-side shim (yolohtli_amd/shim/yh_shim.c, the binding a maintainer
adds to the reference's display() in place of the wrappers of
hostPrototypes.h:22-54) against include/yolohtli.h and libyolohtli_amd.so.

  * host-only calls (defaults, derive, solve_matrix, the parameter CSV round
    trip) against the Python binding and numpy, on the CPU;
  * yh_shim_init + yh_shim_plain and yh_shim_sr on 128^2 from the reference's
    cross-field start (SR after 1000 plain steps, from the last tip), bit for
    bit against the Python Plan, on the GPU.
"""
import os
import subprocess

import numpy as np
import pytest

import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "c", "shim_driver")
N = 128


def _driver():
    if not os.path.exists(DRIVER):
        import __graft_entry__ as G
        G.build_shim_driver()
    return DRIVER


def _run(mode, tmp_path):
    r = subprocess.run([_driver(), mode, str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}


def test_host_calls_from_c(tmp_path):
    out = _run("host", tmp_path)
    p = T.default_params()
    assert [int(out["default"][0]), int(out["default"][1])] == [p.nx, p.ny]
    assert [float(x) for x in out["default"][2:7]] == [p.Lx, p.dt, p.Uth, p.fk.tau_v1m,
                                                       p.fk.tau_v2m]
    assert [int(out["default"][7]), int(out["default"][8])] == [p.tipOffsetX, p.timeIntOrder]
    d = Y.derive(T.grid_params(512, 512))
    assert [float(x) for x in out["derive"]] == [d.hx, d.rx, d.dt]
    phi = np.array([0.1, -0.2, 0.3])
    I = np.array([3, 0.1, 0.2, 0.1, 3, 0.3, 0.2, 0.3, 3, 0.5, -0.4, 0.7])
    cp, sp = np.cos(phi[2]), np.sin(phi[2])
    A = np.array([[I[3 * r] * cp + I[3 * r + 1] * sp, I[3 * r + 1] * cp - I[3 * r] * sp,
                   I[3 * r + 2]] for r in range(3)])
    np.testing.assert_allclose([float(x) for x in out["solve"]], np.linalg.solve(A, I[9:12]),
                               rtol=1e-12)
    # the CSV carries the values at float precision (strtof, saveFiles.cu:540-711)
    assert out["csv"][:2] == ["512", "512"] and float(out["csv"][2]) == pytest.approx(0.015)
    assert out["csv"][3] == "77"
    assert os.path.getsize(tmp_path / "dataparamcsv.csv") > 0


# the paramVar fields the shim reads as REAL (float in the shim's build)
REAL_FIELDS = ("Lx", "Ly", "dt", "diff_par", "diff_per", "rscale", "degrad", "boundaryVal",
               "Uth", "tc", "alpha", "beta", "gamma", "delta", "eps", "mu", "theta")


def _shim_params(reduce_sym):
    """The parameters yh_shim_init builds from the driver's paramVar: the
    REAL fields pass through float (the SR step uses hx, hy and dt in FP64,
    so the rounding shows there)."""
    p = T.grid_params(N, N, reduceSym=reduce_sym)
    for name in REAL_FIELDS:
        setattr(p, name, float(np.float32(getattr(p, name))))
    return p


def _plan_start(reduce_sym):
    p = _shim_params(reduce_sym)
    pl = Y.Plan(p)
    pl.set_state(*Y.initGates(N, N))
    return pl


@pytest.mark.gpu
def test_shim_plain_and_sr_equal_the_python_plan(torch_cuda, tmp_path):
    _run("plain", tmp_path)
    st = np.fromfile(tmp_path / "plain_state.bin", np.float32).reshape(2, N * N)
    tips = np.fromfile(tmp_path / "plain_tips.bin", np.float32).reshape(-1, 5)
    pl = _plan_start(0)
    for _ in range(4):
        pl.step(50)
        pl.track_tips()
    want = pl.get_state()
    want_tips = pl.read_tips()
    pl.close()
    for a in range(2):
        assert np.array_equal(st[a], want[a])
    assert np.array_equal(tips, want_tips)

    _run("sr", tmp_path)
    st = np.fromfile(tmp_path / "sr_state.bin", np.float32).reshape(2, N * N)
    cphi = np.fromfile(tmp_path / "sr_cphi.bin", np.float64).reshape(200, 6)
    pl = _plan_start(0)
    pl.step(1000)
    pl.track_tips()
    tips = pl.read_tips()
    warm = pl.get_state()
    pl.close()
    tip0 = (float(tips[-1, 0]), float(tips[-1, 1])) if len(tips) else (N // 2, N // 2)
    pl = Y.Plan(_shim_params(1))
    pl.set_state(*warm)
    pl.sr_begin(tip0)
    hist = pl.sr_step(200, history=True)
    want = pl.get_state()
    pl.close()
    assert np.isfinite(hist).all() and np.isfinite(want[0]).all()
    for a in range(2):
        assert np.array_equal(st[a], want[a])
    assert np.array_equal(cphi, hist)
