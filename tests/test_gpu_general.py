"""GPU parity of the other stepper branches (general.hip: Dirichlet, lap4,
anisotropy, RK2/RK4, mouse stimulus) against the oracle's race-free FP64
restatement of reactionDiffusion.cu:36-561.

Tolerances (FP32 GPU vs FP64 oracle on the same FP32 inputs): one step 3e-6;
40 steps 1e-4 for the 2V model, and for Fenton-Karma (whose sharp Heaviside
can flip a cell sitting on a threshold) at most 0.2% of the cells above 1e-4.
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

pytestmark = pytest.mark.gpu

ANISO = dict(anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0)
CASES = [
    (T.YH_MODEL_2V_REF, 0, dict(lap4=1)),
    (T.YH_MODEL_2V_REF, 0, ANISO),
    (T.YH_MODEL_2V_REF, 0, dict(lap4=1, **ANISO)),
    (T.YH_MODEL_2V_REF, 0, dict(neumannBC=0, boundaryVal=0.1)),
    (T.YH_MODEL_2V_REF, 0, dict(neumannBC=0, boundaryVal=0.1, **ANISO)),
    (T.YH_MODEL_2V_REF, 0, dict(timeIntOrder=2)),
    (T.YH_MODEL_2V_REF, 0, dict(timeIntOrder=4)),
    (T.YH_MODEL_2V_REF, 0, dict(timeIntOrder=4, lap4=1)),
    (T.YH_MODEL_2V_REF, 1, dict(neumannBC=0, boundaryVal=0.1)),
    (T.YH_MODEL_2V_REF, 1, dict(neumannBC=0, boundaryVal=0.1, **ANISO)),
    (T.YH_MODEL_2V_REF, 1, dict(timeIntOrder=4)),
    (T.YH_MODEL_3V_FK, 0, dict(lap4=1)),
    (T.YH_MODEL_3V_FK, 0, dict(neumannBC=0, boundaryVal=0.0)),
    (T.YH_MODEL_3V_FK, 0, dict(timeIntOrder=4)),
    (T.YH_MODEL_3V_FK, 0, ANISO),
    (T.YH_MODEL_3V_FK, 1, dict(neumannBC=0, boundaryVal=0.0)),
    (T.YH_MODEL_3V_FK, 1, dict(timeIntOrder=2)),
]
NX, NY = 70, 52


def _setup(model, masked, over, seed):
    rng = np.random.default_rng(seed)
    p = T.grid_params(NX, NY, model=model, solidSwitch=masked, **over)
    nv = T.nvars(p)
    f = list(Y.initGates(NX, NY, u_cols=12, rest=1.0 if model else 0.0,
                         refr=0.0 if model else 1.0, nvars=nv))
    f[0] = np.clip(f[0] + 0.05 * rng.random(NX * NY).astype(np.float32), 0, 1)
    m = Y.make_holes(NX, NY, 4, 2.0, 5.0, seed) if masked else None
    return p, f, m


def _gate(model, got, want):
    for a in range(len(want)):
        d = np.abs(got[a].astype(np.float64) - want[a])
        if model == T.YH_MODEL_2V_REF:
            assert d.max() < 1e-4, (a, d.max())
        else:
            assert np.mean(d > 1e-4) < 2e-3, (a, np.mean(d > 1e-4), d.max())


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plan_steps_match_oracle(torch_cuda, case):
    model, masked, over = CASES[case]
    p, f, m = _setup(model, masked, over, 300 + case)
    pl = Y.Plan(p)
    if masked:
        pl.set_mask(m)
    pl.set_state(*f)
    pl.step(40)
    got = pl.get_state()
    want = O.run_general(p, [x.astype(np.float64) for x in f], 40, solid=m)
    print(CASES[case], [float(np.abs(g - w).max()) for g, w in zip(got, want)])
    _gate(model, got, want)
    pl.close()


@pytest.mark.parametrize("model,masked,over", [
    (T.YH_MODEL_2V_REF, 0, dict(timeIntOrder=4, lap4=1, **ANISO)),
    (T.YH_MODEL_3V_FK, 1, dict(neumannBC=0, boundaryVal=0.0, **ANISO)),
    (T.YH_MODEL_3V_FK, 0, dict(timeIntOrder=2)),
])
def test_larger_grid_100_steps_match_oracle(torch_cuda, model, masked, over):
    """The heaviest branch combinations on 512 x 384 (ragged against every
    block) for 100 steps, with the gates of the 70 x 52 cases."""
    nx, ny = 512, 384
    rng = np.random.default_rng(77 + model)
    p = T.grid_params(nx, ny, model=model, solidSwitch=masked, **over)
    nv = T.nvars(p)
    f = list(Y.initGates(nx, ny, u_cols=48, rest=1.0 if model else 0.0,
                         refr=0.0 if model else 1.0, nvars=nv))
    f[0] = np.clip(f[0] + 0.05 * rng.random(nx * ny).astype(np.float32), 0, 1)
    m = Y.make_holes(nx, ny, 20, 3.0, 9.0, 77) if masked else None
    pl = Y.Plan(p)
    if masked:
        pl.set_mask(m)
    pl.set_state(*f)
    pl.step(100)
    got = pl.get_state()
    pl.close()
    want = O.run_general(p, [x.astype(np.float64) for x in f], 100, solid=m)
    print(over, [float(np.abs(g - w).max()) for g, w in zip(got, want)])
    _gate(model, got, want)


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
@pytest.mark.parametrize("over", [dict(), dict(neumannBC=0, boundaryVal=0.2, **ANISO)])
def test_wrapper_with_mouse_stimulus(torch_cuda, model, over):
    torch = torch_cuda
    p, f, _ = _setup(model, 0, over, 17)
    nv = T.nvars(p)
    gin = [torch.from_numpy(x).cuda() for x in f]
    gout = [torch.empty_like(x) for x in gin]
    vt = [torch.empty_like(x) for x in gin]
    Y.reactionDiffusion_wrapper(p, gout, gin, velTan=vt, stimLockMouse=True, point=(30, 21))
    torch.cuda.synchronize()
    want, wvt = O.rd_step_general(p, [x.astype(np.float64) for x in f], stim=(30, 21),
                                  veltan=True)
    dt = O.derive(p).dt
    for a in range(nv):
        got = gout[a].cpu().numpy().astype(np.float64)
        assert np.max(np.abs(got - want[a])) < 3e-6, (a, np.max(np.abs(got - want[a])))
        assert np.max(np.abs(vt[a].cpu().numpy() - wvt[a])) * dt < 3e-6


def test_plan_stimulus_switches_on_and_off(torch_cuda):
    p, f, _ = _setup(T.YH_MODEL_2V_REF, 0, {}, 23)
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.set_stimulus(True, (45, 30))
    pl.step(10)
    pl.set_stimulus(False)
    pl.step(10)
    got = pl.get_state()
    X = [x.astype(np.float64) for x in f]
    X = O.run_general(p, X, 10, stim=(45, 30))
    X = O.run_general(p, X, 10)
    _gate(T.YH_MODEL_2V_REF, got, X)


def test_general_branches_rejected_where_unsupported(torch_cuda):
    torch = torch_cuda
    p = T.grid_params(64, 64, timeIntOrder=4)
    f = Y.initGates(64, 64)
    gin = [torch.from_numpy(x).cuda() for x in f]
    gout = [torch.empty_like(x) for x in gin]
    with pytest.raises(RuntimeError):  # RK stages need a plan's scratch fields
        Y.reactionDiffusion_wrapper(p, gout, gin)
    q = T.grid_params(64, 64, lap4=1, reduceSym=1, tipOffsetX=20, tipOffsetY=20)
    pl = Y.Plan(q)
    pl.set_state(*f)
    with pytest.raises(RuntimeError):  # the SR step runs the standard branch
        pl.sr_begin((32.0, 32.0))
    from yolohtli_amd.batch import Batch
    with pytest.raises(RuntimeError):
        Batch([T.grid_params(64, 64, model=T.YH_MODEL_3V_FK, solidSwitch=1, neumannBC=0)])


@pytest.mark.parametrize("mode,th", [(1, (0.0, 0.0, 0.0)), (2, (0.5, 0.25, 0.0)),
                                     (3, (0.5, 0.25, 0.375))])
def test_contours_match_oracle(torch_cuda, mode, th):
    """Contour points of two smooth random fields: the same cells (keys) and,
    for mode 1, the same interpolated positions to FP32 rounding."""
    torch = torch_cuda
    nx, ny = 96, 80
    rng = np.random.default_rng(40 + mode)
    j, i = np.mgrid[0:ny, 0:nx]
    f1 = (0.5 + 0.5 * np.sin(0.11 * i + 0.7) * np.cos(0.07 * j)).astype(np.float32)
    f2 = (np.sin(0.09 * i - 0.3) * np.cos(0.13 * j + 0.2) * 0.6 + 0.3).astype(np.float32)
    if mode == 1:
        f2 = np.where(f2 > 0.3, 1.0, -1.0).astype(np.float32)
    area = (rng.random((ny, nx)) > 0.1).astype(np.uint8)
    cap = 2 * nx * ny
    d1, d2 = torch.from_numpy(f1.reshape(-1)).cuda(), torch.from_numpy(f2.reshape(-1)).cuda()
    da = torch.from_numpy(area.reshape(-1)).cuda()
    pts = torch.zeros((cap, 3), dtype=torch.float32, device="cuda")
    keys = torch.zeros(cap, dtype=torch.int32, device="cuda")
    cnt = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    Y.solver.countour_wrapper(nx, ny, d1, d2, da, pts, keys, cnt, cap, 3.5, mode, th)
    torch.cuda.synchronize()
    n = int(cnt.item())
    k = keys[:n].cpu().numpy()
    order = np.argsort(k)
    got, gk = pts[:n].cpu().numpy()[order], k[order]
    want, wk = O.contour(nx, ny, mode, th, f1.astype(np.float64), f2.astype(np.float64),
                         area=area, t=3.5)
    assert n > 20
    np.testing.assert_array_equal(gk, wk)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)


def test_contour_dev_graph_replays_match_eager(torch_cuda):
    """yh_contour_dev captured into a graph: every replay zeroes the count on
    the stream (with a kernel, not a memset node) and finds the same points as
    the eager call, with eager work and synchronisations between replays."""
    torch = torch_cuda
    nx, ny, cap, th = 96, 80, 2 * 96 * 80, (0.5, 0.25, 0.0)
    j, i = np.mgrid[0:ny, 0:nx]
    f1 = (0.5 + 0.5 * np.sin(0.11 * i + 0.7) * np.cos(0.07 * j)).astype(np.float32)
    f2 = (np.sin(0.09 * i - 0.3) * np.cos(0.13 * j + 0.2) * 0.6 + 0.3).astype(np.float32)
    d1, d2 = torch.from_numpy(f1.reshape(-1)).cuda(), torch.from_numpy(f2.reshape(-1)).cuda()
    pts = torch.zeros((cap, 3), dtype=torch.float32, device="cuda")
    keys = torch.zeros(cap, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")

    def result():
        torch.cuda.synchronize()
        n = int(cnt.item())
        k = keys[:n].cpu().numpy()
        order = np.argsort(k)
        return k[order], pts[:n].cpu().numpy()[order]

    Y.solver.countour_wrapper(nx, ny, d1, d2, None, pts, keys, cnt, cap, 3.5, 2, th)
    wk, wp = result()
    assert len(wk) > 20
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        Y.solver.countour_wrapper(nx, ny, d1, d2, None, pts, keys, cnt, cap, 3.5, 2, th)
    for rep in range(3):
        cnt.fill_(777 + rep)
        keys.fill_(-1)
        pts.zero_()
        torch.cuda.synchronize()
        g.replay()
        gk, gp = result()
        np.testing.assert_array_equal(gk, wk)
        np.testing.assert_array_equal(gp, wp)


def test_plan_contour_modes(torch_cuda):
    """Plan contours of (u, v) and, with APD on, of sign(APD1 - APD2)."""
    p = T.grid_params(64, 64)
    f = Y.initGates(64, 64)
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.step(300)
    u, v = pl.get_state()
    got = pl.contour(2, (0.5, 0.25, 0.0))
    want, _ = O.contour(64, 64, 2, (0.5, 0.25, 0.0), u.astype(np.float64), v.astype(np.float64),
                        t=O.derive(p).dt * 300)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-6)
    with pytest.raises(RuntimeError):
        pl.contour(1)  # APD not enabled
    pl.apd_enable()
    pl.step(3000)
    a = pl.apd_get()
    sapd = np.where(a["apd1"] - a["apd2"] > 0, 1.0, -1.0)
    got = pl.contour(1)
    want, _ = O.contour(64, 64, 1, (0, 0, 0), np.zeros(64 * 64), sapd, t=O.derive(p).dt * 3300)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
