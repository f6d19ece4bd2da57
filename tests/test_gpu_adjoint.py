"""GPU gradient path against the FP64 oracle.

  * forward traces and loss of the GPU (FP32, smoothed Heaviside, D field)
    against the oracle's FP64 forward;
  * the GPU adjoint (FP32 state, FP64 accumulation) against the oracle's
    discrete adjoint: relative error of the whole D-gradient and of each
    time-constant gradient <= 1e-3;
  * against FP64 central finite differences of the oracle loss at single
    cells and for the six time constants: relative error <= 1e-3;
  * the torch.autograd.Function returns the same gradients, and re-runs its
    own forward when another forward ran in between.
"""
import numpy as np
import pytest

import oracle.oracle as O
from yolohtli_amd import abi_types as T
from yolohtli_amd.adjoint import SensorProblem, TAU_NAMES, adjoint_workload, sensor_loss

pytestmark = pytest.mark.gpu


def _case(masked, n=64, nsteps=600, rec=5):
    p, x0, sensors, D_hidden, D0 = adjoint_workload(n=n, nsteps=nsteps, nsensors=24, config=41)
    # sensors in the band the wave crosses within nsteps (well-scaled loss)
    rng = np.random.default_rng(5)
    k = min(24, n // 2)
    cols = rng.integers(4, min(24, n), size=k)
    rows = rng.choice(n, size=k, replace=False)
    sensors = (rows * n + cols).astype(np.int32)
    solid = None
    if masked:
        p = T.copy_params(p, solidSwitch=1)
        solid = np.ones(n * n, np.uint8)
        solid[(np.arange(n * n) % 97) == 13] = 0
        sensors = np.array([s for s in sensors if solid[s]], np.int32)
    D = np.full(n * n, D0, np.float32)
    return p, x0, sensors, D_hidden.reshape(-1), D, solid, nsteps, rec


@pytest.mark.parametrize("masked", [False, True])
def test_gpu_adjoint_matches_oracle(torch_cuda, masked):
    torch = torch_cuda
    p, x0, sensors, Dh, D, solid, nsteps, rec = _case(masked)
    pr = SensorProblem(p, nsteps, sensors, x0, rec_every=rec, ckpt_every=70, mask=solid)
    # GPU forward traces in the hidden field against the oracle's
    _, tr = pr.forward(torch.from_numpy(Dh).cuda())
    _, want_tr = O.fk_loss(p, Dh.astype(np.float64), [f.astype(np.float64) for f in x0], nsteps,
                           rec, sensors, None, solid, traces=True)
    print("trace error", np.max(np.abs(tr - want_tr)), "trace range", tr.min(), tr.max())
    assert np.max(np.abs(tr - want_tr)) < 1e-3
    # gradient checks on a well-scaled loss: constant targets 0.3
    targets = np.full_like(tr, 0.3)
    pr.set_targets(targets)
    Dt = torch.from_numpy(D).cuda()
    L, _ = pr.forward(Dt)
    gD, gt = pr.backward()
    gD = gD.cpu().numpy()
    X0 = [f.astype(np.float64) for f in x0]
    tg = targets.astype(np.float64)
    Lo, gDo, gto = O.fk_adjoint(p, D.astype(np.float64), X0, nsteps, rec, sensors, tg, solid)
    print("loss gpu", L, "oracle", Lo)
    assert L == pytest.approx(Lo, rel=1e-3)
    rel = np.linalg.norm(gD - gDo) / np.linalg.norm(gDo)
    print("relative D-gradient error", rel)
    assert rel < 1e-3
    for k in range(6):
        assert gt[k] == pytest.approx(gto[k], rel=1e-3), (TAU_NAMES[k], gt[k], gto[k])
    # central finite differences of the FP64 loss
    Dd = D.astype(np.float64)
    big = np.argsort(-np.abs(gDo))[:4]
    for c in big:
        h = 1e-5 * Dd[c]
        Dp, Dm = Dd.copy(), Dd.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (O.fk_loss(p, Dp, X0, nsteps, rec, sensors, tg, solid)
              - O.fk_loss(p, Dm, X0, nsteps, rec, sensors, tg, solid)) / (2 * h)
        assert gD[c] == pytest.approx(fd, rel=1e-3), (c, gD[c], fd)
    for k, name in enumerate(TAU_NAMES):
        t0 = getattr(p.fk, name)
        h = 1e-5 * t0
        fd = (O.fk_loss(T.copy_params(p, fk={name: t0 + h}), Dd, X0, nsteps, rec, sensors, tg, solid)
              - O.fk_loss(T.copy_params(p, fk={name: t0 - h}), Dd, X0, nsteps, rec, sensors, tg,
                          solid)) / (2 * h)
        assert gt[k] == pytest.approx(fd, rel=1e-3), (name, gt[k], fd)


@pytest.mark.parametrize("masked,n,k", [(False, 64, 8), (True, 64, 5), (True, 45, 8),
                                         (False, 45, 3), (True, 12, 8)])
def test_tiled_steps_equal_one_step(torch_cuda, masked, n, k):
    """The LDS-tiled forward and reverse (k steps per launch; forward launches
    end on every checkpoint, 70 steps here, reverse launches on segment
    starts) against the one-step kernels, on ragged tiles (45 = 2 x 16 + 13),
    a grid smaller than one tile (12) and with a mask: traces and loss bit
    for bit; the gradients to 1e-12, since
    only the grouping of their FP64 sums over steps and workgroups differs
    (lambda itself is computed by the same per-cell code)."""
    torch = torch_cuda
    p, x0, sensors, Dh, D, solid, nsteps, rec = _case(masked, n=n, nsteps=300, rec=3)
    out = []
    for kk in (1, k):
        pr = SensorProblem(p, nsteps, sensors, x0, rec_every=rec, ckpt_every=70, mask=solid)
        pr.set_substeps(kk)
        _, tr = pr.forward(torch.from_numpy(Dh).cuda())
        pr.set_targets(np.full_like(tr, 0.3))
        L, _ = pr.forward(torch.from_numpy(D).cuda())
        gD, gt = pr.backward()
        out.append((tr, L, gD.cpu().numpy(), np.array(gt)))
        pr.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    scale = np.max(np.abs(out[0][2]))
    np.testing.assert_allclose(out[1][2], out[0][2], rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(out[1][3], out[0][3], rtol=1e-12)


def test_autograd_function(torch_cuda):
    torch = torch_cuda
    p, x0, sensors, Dh, D, solid, nsteps, rec = _case(False, n=48, nsteps=240, rec=4)
    pr = SensorProblem(p, nsteps, sensors, x0, rec_every=rec, ckpt_every=50)
    _, tr = pr.forward(torch.from_numpy(Dh).cuda())
    pr.set_targets(np.full_like(tr, 0.3))
    Dv = torch.from_numpy(D.reshape(48, 48)).cuda().requires_grad_(True)
    taus = torch.tensor(pr.default_taus(), dtype=torch.float64, requires_grad=True)
    loss = sensor_loss(Dv, taus, pr)
    pr.forward(torch.from_numpy(Dh).cuda())  # an unrelated forward in between
    (2.0 * loss).backward()
    pr.forward(torch.from_numpy(D).cuda())
    gD, gt = pr.backward()
    np.testing.assert_allclose(Dv.grad.cpu().numpy().reshape(-1), 2.0 * gD.cpu().numpy(),
                               rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(taus.grad.numpy(), 2.0 * gt, rtol=1e-12)
    # one gradient-descent step on D lowers the loss
    with torch.no_grad():
        lr = 0.01 * 0.001 / Dv.grad.abs().max()  # largest change: 1% of D0
        Dn = (Dv - lr * Dv.grad).contiguous()
    l1 = sensor_loss(Dn, taus, pr).item()
    assert l1 < loss.item()


def test_adjoint_argument_errors(torch_cuda):
    torch = torch_cuda
    p, x0, sensors, Dh, D, solid, nsteps, rec = _case(False, n=32, nsteps=20, rec=2)
    with pytest.raises(RuntimeError):
        SensorProblem(T.copy_params(p, heaviside_width=0.0), nsteps, sensors, x0)
    with pytest.raises(RuntimeError):
        SensorProblem(p, nsteps, np.array([3, 3], np.int32), x0)
    pr = SensorProblem(p, nsteps, sensors, x0, rec_every=rec)
    pr.forward(torch.from_numpy(D).cuda())
    with pytest.raises(RuntimeError):
        pr.backward()  # no targets yet
