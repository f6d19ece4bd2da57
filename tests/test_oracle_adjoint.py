"""Pins the oracle's gradient path (oracle.c orc_fk_dfield_step_d,
orc_fk_loss_d, orc_fk_adjoint_d) without the GPU:
  * with a uniform D the flux-form operator equals the reference operator
    (c_W (u_W-u) + c_E (u_E-u)) rx + ... of the plain step, masked or not;
  * the smoothed Heaviside tends to the sharp one as the width shrinks;
  * the discrete adjoint matches FP64 central finite differences of the loss
    in D (single cells, including mirror-boundary and hole-adjacent cells)
    and in each of the six time constants.
"""
import numpy as np
import pytest

import oracle.oracle as O
from yolohtli_amd import abi_types as T

TAUS = ("tau_d", "tau_r", "tau_si", "tau_0", "tau_vp", "tau_wp")


def _setup(n=16, masked=False, width=0.05, seed=0):
    rng = np.random.default_rng(seed)
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=int(masked),
                      heaviside_width=width)
    u = np.zeros((n, n))
    u[:, :3] = 1.0
    v = np.ones((n, n))
    v[n // 2:, :] = 0.0
    w = v.copy()
    X0 = [u.reshape(-1), v.reshape(-1), w.reshape(-1)]
    solid = None
    if masked:
        solid = np.ones(n * n, np.uint8)
        solid[[5 * n + 7, 5 * n + 8, 9 * n + 2]] = 0
    D = 0.001 * (1.0 + 0.3 * rng.random(n * n))
    return p, X0, D, solid


@pytest.mark.parametrize("masked", [False, True])
def test_uniform_D_operator_equals_reference_step(masked):
    p, X0, _, solid = _setup(masked=masked)
    rng = np.random.default_rng(1)
    X = [rng.random(p.nx * p.ny) for _ in range(3)]
    D = np.full(p.nx * p.ny, 0.001)
    a = O.fk_dfield_step(p, D, X, solid)
    b, _ = O.rd_step(p, X, solid)
    for q in range(3):
        np.testing.assert_allclose(a[q], b[q], rtol=0, atol=1e-14)


def test_smoothed_heaviside_tends_to_sharp():
    p, X0, _, _ = _setup(width=0.0)
    rng = np.random.default_rng(2)
    X = [rng.random(p.nx * p.ny) for _ in range(3)]
    sharp, _ = O.rd_step(p, X)
    errs = []
    for wdt in (1e-2, 1e-3, 1e-4):
        q = T.copy_params(p, heaviside_width=wdt)
        sm, _ = O.rd_step(q, X)
        errs.append(np.mean(np.abs(sm[1] - sharp[1])))
    assert errs[0] > errs[1] > errs[2]


def _loss(p, D, X0, nsteps, rec, sensors, targets, solid):
    return O.fk_loss(p, D, X0, nsteps, rec, sensors, targets, solid)


@pytest.mark.parametrize("masked", [False, True])
def test_adjoint_matches_central_fd(masked):
    p, X0, D, solid = _setup(masked=masked)
    n = p.nx
    nsteps, rec = 60, 3
    sensors = np.array([2 * n + 5, 7 * n + 7, 12 * n + 3, 14 * n + 13, 6 * n + 9], np.int32)
    D_true = D * (1.0 + 0.2 * np.sin(np.arange(n * n)))
    _, targets = O.fk_loss(p, D_true, X0, nsteps, rec, sensors, None, solid, traces=True)
    L, gD, gt = O.fk_adjoint(p, D, X0, nsteps, rec, sensors, targets, solid)
    assert L == pytest.approx(_loss(p, D, X0, nsteps, rec, sensors, targets, solid), rel=1e-13)
    cells = [0, 1, n - 1, n, 5 * n + 6, 6 * n + 7, 4 * n + 8, 7 * n + 7, n * n - 1, n * n - n + 2]
    for c in cells:
        h = 1e-6 * D[c]
        Dp, Dm = D.copy(), D.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (_loss(p, Dp, X0, nsteps, rec, sensors, targets, solid)
              - _loss(p, Dm, X0, nsteps, rec, sensors, targets, solid)) / (2 * h)
        assert gD[c] == pytest.approx(fd, rel=2e-5, abs=1e-9 * np.max(np.abs(gD))), c
    for k, name in enumerate(TAUS):
        t0 = getattr(p.fk, name)
        h = 1e-6 * t0
        pp = T.copy_params(p, fk={name: t0 + h})
        pm = T.copy_params(p, fk={name: t0 - h})
        fd = (_loss(pp, D, X0, nsteps, rec, sensors, targets, solid)
              - _loss(pm, D, X0, nsteps, rec, sensors, targets, solid)) / (2 * h)
        assert gt[k] == pytest.approx(fd, rel=2e-5, abs=1e-12), name
    assert np.max(np.abs(gD)) > 0 and np.max(np.abs(gt)) > 0
