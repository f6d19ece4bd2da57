"""GPU symmetry-reduction step against the oracle's SR step (main.cu:894-953).

The GPU keeps the state in FP32 and does the slices, inner products and the
3x3 solve in FP64; the oracle is FP64 throughout.  Tolerances: c and phi
relative 1e-3 (the translation/rotation velocities are small differences of
large inner products), state 1e-4 absolute after 40 steps.
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

pytestmark = pytest.mark.gpu

N = 128
R = 40


def _spiral_state(model):
    """Cross-field start, plain steps until a broken wave exists (FP32 on GPU)."""
    if model == T.YH_MODEL_2V_REF:
        p = T.grid_params(N, N)
        f = Y.initGates(N, N)
        warm = 2500
    else:
        p = T.grid_params(N, N, model=model)
        f = Y.initGates(N, N, u_cols=16, rest=1.0, refr=0.0, nvars=3)
        warm = 1500
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.step(warm)
    pl.track_tips()
    tips = pl.read_tips()
    st = pl.get_state()
    pl.close()
    tip0 = (float(tips[-1, 0]), float(tips[-1, 1])) if len(tips) else (N / 2, N / 2)
    return st, tip0


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_sr_steps_match_oracle(torch_cuda, model):
    st, tip0 = _spiral_state(model)
    p = T.grid_params(N, N, model=model, reduceSym=1, tipOffsetX=R, tipOffsetY=R)
    pl = Y.Plan(p)
    pl.set_state(*st)
    pl.sr_begin(tip0)
    hist = pl.sr_step(40, history=True)
    info = pl.sr_info()
    got = pl.get_state()
    ref = O.SR(p, [x.astype(np.float64) for x in st], tip0=tip0)
    want_hist = []
    for _ in range(40):
        want_hist.append(np.concatenate([ref.c, ref.phi]))
        ref.step(1)
    want_hist = np.array(want_hist)
    print("cx,cy gpu", info["cx"], info["cy"], "oracle", ref.s.cx, ref.s.cy)
    print("c gpu", info["c"], "oracle", ref.c)
    assert (info["cx"], info["cy"]) == (ref.s.cx, ref.s.cy)
    scale = np.max(np.abs(want_hist), axis=0) + 1e-12
    rel = np.max(np.abs(hist - want_hist) / scale, axis=0)
    print("relative history error per component", rel)
    assert np.all(rel < 1e-3), rel
    np.testing.assert_allclose(info["I"], np.array(ref.s.I[:]), rtol=1e-3,
                               atol=1e-6 * np.max(np.abs(ref.s.I[:])))
    for a in range(len(got)):
        err = np.max(np.abs(got[a] - ref.X[a]))
        print("state", a, err)
        assert err < 1e-4, (a, err)


def test_sr_graph_replay_equals_eager(torch_cuda):
    st, tip0 = _spiral_state(T.YH_MODEL_2V_REF)
    p = T.grid_params(N, N, reduceSym=1, tipOffsetX=R, tipOffsetY=R)
    outs = []
    for graph in (True, False):
        pl = Y.Plan(p)
        pl.set_state(*st)
        pl.sr_begin(tip0)
        pl.sr_set_graph(graph)
        h1 = pl.sr_step(7, history=True)
        h2 = pl.sr_step(9, history=True)
        outs.append((np.concatenate([h1, h2]), pl.get_state(), pl.read_tips()))
        pl.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    for a in range(2):
        np.testing.assert_array_equal(outs[0][1][a], outs[1][1][a])
    np.testing.assert_array_equal(outs[0][2], outs[1][2])


def test_sr_singular_and_state_errors(torch_cuda):
    p = T.grid_params(64, 64, reduceSym=1, tipOffsetX=20, tipOffsetY=20)
    pl = Y.Plan(p)
    pl.set_state(np.zeros(64 * 64, np.float32), np.zeros(64 * 64, np.float32))
    with pytest.raises(RuntimeError):
        pl.sr_step(1)  # before sr_begin
    pl.sr_begin((32.0, 32.0))
    pl.sr_step(2)
    info = pl.sr_info()
    assert info["status"] & 1  # all-zero integrals: singular solve flagged
    q = Y.Plan(T.grid_params(64, 64))
    q.set_state(np.zeros(64 * 64, np.float32), np.zeros(64 * 64, np.float32))
    with pytest.raises(RuntimeError):
        q.sr_begin((32.0, 32.0))  # plan without reduceSym


@pytest.mark.parametrize("edge_holes", [False, True])
def test_sr_masked_matches_oracle(torch_cuda, edge_holes):
    """The masked branch: holes in the tissue (zero state, mirrored no-flux
    stencil, BFECC neighbour rules of advFDBFECC.cu:38-103) and the tip
    search restricted to the centred disk (tipTracker.cu:47-51).  With
    edge_holes, single-cell holes next to all four domain edges, where the
    BFECC rule's substitute neighbour lies outside the domain and reads as a
    hole."""
    model = T.YH_MODEL_3V_FK
    st, tip0 = _spiral_state(model)
    mask = Y.make_holes(N, N, 6, 3.0, 5.0, Y.seed(9, 1, 1))
    if edge_holes:
        m = mask.reshape(N, N).copy()
        m[::5, 1] = 0
        m[2::5, N - 2] = 0
        m[1, ::7] = 0
        m[N - 2, 3::7] = 0
        mask = m.reshape(-1)
    st = [np.where(mask != 0, x, 0.0).astype(np.float32) for x in st]
    p = T.grid_params(N, N, model=model, reduceSym=1, tipOffsetX=R, tipOffsetY=R, solidSwitch=1)
    pl = Y.Plan(p)
    pl.set_mask(mask)
    pl.set_state(*st)
    pl.sr_begin(tip0)
    hist = pl.sr_step(40, history=True)
    info = pl.sr_info()
    got = pl.get_state()
    ref = O.SR(p, [x.astype(np.float64) for x in st], tip0=tip0, solid=mask)
    want_hist = []
    for _ in range(40):
        want_hist.append(np.concatenate([ref.c, ref.phi]))
        ref.step(1)
    want_hist = np.array(want_hist)
    print("cx,cy gpu", info["cx"], info["cy"], "oracle", ref.s.cx, ref.s.cy)
    assert (info["cx"], info["cy"]) == (ref.s.cx, ref.s.cy)
    scale = np.max(np.abs(want_hist), axis=0) + 1e-12
    rel = np.max(np.abs(hist - want_hist) / scale, axis=0)
    print("relative history error per component", rel)
    assert np.all(rel < 1e-3), rel
    for a in range(len(got)):
        assert np.all(got[a][mask == 0] == 0)
        err = np.max(np.abs(got[a] - ref.X[a]))
        print("state", a, err)
        assert err < 1e-4, (a, err)


def _sr_run(p, st, tip0, launches, graph, mask=None, interrupt=False):
    """100 SR steps in calls of 37 + 63 (with `interrupt`, 5 plain steps
    between the calls: the state changed, so with two launches the next call
    recomputes the RD output and tips of its first step); state, history and
    tip log."""
    pl = Y.Plan(p)
    if mask is not None:
        pl.set_mask(mask)
    pl.set_state(*st)
    pl.set_option(T.YH_OPT_SR_LAUNCHES, launches)
    pl.sr_begin(tip0)
    pl.sr_set_graph(graph)
    h1 = pl.sr_step(37, history=True)
    if interrupt:
        pl.step(5)
    h2 = pl.sr_step(63, history=True)
    out = (np.concatenate([h1, h2]), pl.get_state(), pl.read_tips())
    pl.close()
    return out


@pytest.mark.parametrize("model,masked", [(T.YH_MODEL_3V_FK, False), (T.YH_MODEL_3V_FK, True),
                                          (T.YH_MODEL_2V_REF, False)])
def test_two_launch_sr_step_equals_three_launches(torch_cuda, model, masked):
    """The default SR step runs in two launches (the BFECC launch also takes
    the RD step and the tip test of the next step, k_bf_rd); the three-launch
    step (YH_OPT_SR_LAUNCHES 3) runs them in k_sr_rdtips.  Over 100 steps the
    two agree bit for bit in state, (c, phi) history and tip log, eagerly and
    from the graphs, also across a change of the state between calls."""
    st, tip0 = _spiral_state(model)
    mask = None
    if masked:
        rng = np.random.default_rng(4)
        mask = (rng.random(N * N) > 0.02).astype(np.uint8)
        st = [np.where(mask > 0, x, 0).astype(np.float32) for x in st]
    p = T.grid_params(N, N, model=model, reduceSym=1, tipOffsetX=R, tipOffsetY=R,
                      solidSwitch=int(masked))
    for graph in (True, False):
        for interrupt in (False, True):
            a = _sr_run(p, st, tip0, 3, graph, mask, interrupt)
            b = _sr_run(p, st, tip0, 2, graph, mask, interrupt)
            np.testing.assert_array_equal(a[0], b[0])
            for q in range(len(a[1])):
                np.testing.assert_array_equal(a[1][q], b[1][q])
            np.testing.assert_array_equal(a[2], b[2])
            assert len(a[2]) > 0
