"""GPU parity of the plain stepper (libyolohtli_amd.so) against the oracle.

Tolerances (FP32 GPU vs FP64 oracle on the same FP32 inputs):
  * one step: 2e-6 absolute (FP32 rounding of O(1) values, a few ops);
  * tiled vs direct kernel after 23 steps: 2e-5 absolute;
  * 1000 steps at 512^2 from the cross-field IC: max|du|, max|dv| <= 1e-4
    (the accuracy gate of the plain stepper); for the FK model with its sharp
    Heaviside switches, the gate holds once the threshold decisions the GPU
    took differently from FP64 are imposed on the oracle (_fk_gate).
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

pytestmark = pytest.mark.gpu

MODELS = [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK]


def _fields(rng, p):
    nv = T.nvars(p)
    n = p.nx * p.ny
    f = [rng.random(n).astype(np.float32) for _ in range(nv)]
    return f


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("masked", [False, True])
def test_single_step_wrapper_matches_oracle(torch_cuda, model, masked):
    torch = torch_cuda
    rng = np.random.default_rng(100 + model * 2 + masked)
    p = T.grid_params(67, 45, model=model, solidSwitch=int(masked))
    nv = T.nvars(p)
    f = _fields(rng, p)
    solid = (rng.random(p.nx * p.ny) > 0.3).astype(np.uint8) if masked else None
    gin = [torch.from_numpy(x).cuda() for x in f]
    gout = [torch.empty_like(x) for x in gin]
    vt = [torch.empty_like(x) for x in gin]
    sd = torch.from_numpy(solid).cuda() if masked else None
    Y.reactionDiffusion_wrapper(p, gout, gin, velTan=vt, solid=sd)
    torch.cuda.synchronize()
    want, wvt = O.rd_step(p, [x.astype(np.float64) for x in f], solid, veltan=True)
    for a in range(nv):
        got = gout[a].cpu().numpy().astype(np.float64)
        assert np.max(np.abs(got - want[a])) < 2e-6, (a, np.max(np.abs(got - want[a])))
        gv = vt[a].cpu().numpy().astype(np.float64)
        c = O.derive(p)
        # velTan = rhs/dt: compare rhs (scaled back by dt) at FP32 resolution
        assert np.max(np.abs(gv - wvt[a])) * c.dt < 2e-6


CFGS = [(K, r, w, c) for K in (1, 2, 3, 4, 8, 12)
        for (r, w) in ((4, 4), (8, 4), (16, 4), (4, 8), (8, 8), (16, 8), (4, 16), (8, 16))
        for c in (1, 2) if 2 * K < r * w and (c == 1 or (r * w <= 64 and r <= 8))]
CFGS += [(K, 7, 16, 1) for K in (1, 3, 8, 9, 12)]  # 7 rows per wave: not a power of two


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("masked", [False, True])
def test_tiled_configs_match_direct(torch_cuda, model, masked):
    rng = np.random.default_rng(7 + model + 2 * masked)
    p = T.grid_params(131, 77, model=model, solidSwitch=int(masked))
    f = _fields(rng, p)
    solid = (rng.random(p.nx * p.ny) > 0.25).astype(np.uint8) if masked else None
    ref = Y.Plan(p)
    if masked:
        ref.set_mask(solid)
    ref.set_kernel(T.YH_KERNEL_DIRECT)
    ref.set_state(*f)
    ref.step(23)
    want = ref.get_state()
    for K, r, w, c in CFGS:
        pl = Y.Plan(p)
        if masked:
            pl.set_mask(solid)
        pl.set_kernel(T.YH_KERNEL_TILED, K, r, w, c)
        pl.set_state(*f)
        pl.step(23)
        got = pl.get_state()
        for a in range(len(got)):
            err = np.max(np.abs(got[a].astype(np.float64) - want[a]))
            assert err < 2e-5, (K, r, w, c, a, err)
        pl.close()


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("masked", [False, True])
def test_unequal_spacing_tiled_matches_oracle_and_direct(torch_cuda, model, masked):
    """hy = 1.3 hx, so the x and y stencil weights differ (trx != try): the
    shapes that also exist with isotropic weights fixed at compile time
    (7 x 16, 8 x 16, 8 x 8) must take the general weights here.  One step
    against the FP64 oracle (2e-6, as the single-step test), 23 steps against
    the direct kernel (2e-5, as the tiled configurations)."""
    rng = np.random.default_rng(31 + model + 2 * masked)
    p = T.grid_params(131, 77, model=model, solidSwitch=int(masked))
    p.Ly = 1.3 * p.Ly
    c = O.derive(p)
    assert abs(c.rx / c.ry - 1.69) < 1e-9
    f = _fields(rng, p)
    solid = (rng.random(p.nx * p.ny) > 0.25).astype(np.uint8) if masked else None
    shapes = [(9, 7, 16, 1), (3, 8, 16, 1), (8, 8, 8, 1), (4, 4, 16, 1), (4, 8, 8, 2)]

    def run(cfg, steps):
        pl = Y.Plan(p)
        if masked:
            pl.set_mask(solid)
        pl.set_kernel(*cfg)
        pl.set_state(*f)
        pl.step(steps)
        out = pl.get_state()
        pl.close()
        return out

    one, _ = O.rd_step(p, [x.astype(np.float64) for x in f], solid)
    want = run((T.YH_KERNEL_DIRECT,), 23)
    for K, r, w, cc in shapes:
        got1 = run((T.YH_KERNEL_TILED, 1, r, w, cc), 1)
        got = run((T.YH_KERNEL_TILED, K, r, w, cc), 23)
        for a in range(len(got)):
            e1 = np.max(np.abs(got1[a].astype(np.float64) - one[a]))
            assert e1 < 2e-6, ("one step", r, w, cc, a, e1)
            err = np.max(np.abs(got[a].astype(np.float64) - want[a]))
            assert err < 2e-5, (K, r, w, cc, a, err)


def _fk_shadow(p, f, nsteps, solid=None):
    """Runs the GPU plan one launch per step (tiled results do not depend on
    where launches split, which the caller checks) beside an FP64 oracle run
    that takes the GPU's Heaviside decisions wherever they differ from its
    own.  The GPU decides STEP(u - u_c) and STEP(u - u_v) on its FP32 u
    against the FP32 thresholds; a cell whose FP64 u lies within FP32
    rounding of a threshold can decide the other way one step earlier or
    later, and one such switch moves v by up to dt/tau_vp ~ 6e-3.  Returns
    the GPU state, the shadowed FP64 state, the switches (step, cell,
    threshold, |u_gpu - u_fp64| at that step) and the largest |u_gpu - u_fp64|
    over all steps and cells."""
    pl = Y.Plan(p)
    if solid is not None:
        pl.set_mask(solid)
    pl.set_state(*f)
    x = [a.astype(np.float64) for a in f]
    ug = f[0]
    uc, uv = np.float32(p.fk.u_c), np.float32(p.fk.u_v)
    flips, drift = [], 0.0
    for s in range(nsteps):
        pg, qg = ug >= uc, ug >= uv
        fp = np.where(pg != (x[0] >= p.fk.u_c), pg.astype(np.int8), np.int8(-1))
        fq = np.where(qg != (x[0] >= p.fk.u_v), qg.astype(np.int8), np.int8(-1))
        for name, ff in (("u_c", fp), ("u_v", fq)):
            for c in np.nonzero(ff >= 0)[0]:
                flips.append((s, int(c), name, abs(float(ug[c]) - x[0][c])))
        x = O.rd_step_forced(p, x, fp, fq, solid=solid)
        pl.step(1)
        g = pl.get_state()
        ug = g[0]
        drift = max(drift, float(np.max(np.abs(ug - x[0]))))
    pl.close()
    return g, x, flips, drift


def _near(cells, centres, nx, r):
    """Whether each cell lies within Chebyshev distance r of a centre."""
    if len(cells) == 0:
        return np.ones(0, bool)
    if len(centres) == 0:
        return np.zeros(len(cells), bool)
    ci, cj = np.asarray(cells) % nx, np.asarray(cells) // nx
    ki, kj = np.asarray(centres) % nx, np.asarray(centres) // nx
    d = np.maximum(np.abs(ci[:, None] - ki[None, :]), np.abs(cj[:, None] - kj[None, :]))
    return d.min(axis=1) <= r


def _fk_gate(p, f, nsteps, solid=None, radius=8):
    """The 1e-4 gate for the FK model (sharp Heaviside switches):
      * the GPU's result does not depend on one launch per step;
      * with the GPU's threshold decisions imposed on the FP64 oracle, every
        variable agrees to 1e-4 everywhere, and u agrees to 1e-4 at every step
        (so each differing decision was taken where FP64's u lay within the
        gate of the threshold);
      * against the FP64 oracle alone, every cell off by more than 1e-4 lies
        within `radius` cells of a cell whose decision differed.
    Prints the maxima of both comparisons."""
    nv = T.nvars(p)
    g, x, flips, drift = _fk_shadow(p, f, nsteps, solid)
    pl = Y.Plan(p)
    if solid is not None:
        pl.set_mask(solid)
    pl.set_state(*f)
    pl.step(nsteps)
    whole = pl.get_state()
    pl.close()
    for a in range(nv):
        assert np.array_equal(whole[a], g[a]), "one launch per step changed the result"
    want = O.run(p, [a.astype(np.float64) for a in f], nsteps, solid=solid)
    shadow = [float(np.max(np.abs(g[a] - x[a]))) for a in range(nv)]
    plain = [float(np.max(np.abs(g[a] - want[a]))) for a in range(nv)]
    bad = np.nonzero(np.max([np.abs(g[a] - want[a]) for a in range(nv)], axis=0) > 1e-4)[0]
    centres = sorted({c for _, c, _, _ in flips})
    print(f"FK {p.nx}x{p.ny} {nsteps} steps: max|d| vs FP64 {plain}; with the GPU's "
          f"decisions imposed {shadow}; u drift over the run {drift:.2e}; "
          f"{len(flips)} differing decisions {flips[:8]}; {len(bad)} cells above 1e-4")
    assert max(shadow) <= 1e-4, shadow
    assert drift <= 1e-4, drift
    near = _near(bad, centres, p.nx, radius)
    assert np.all(near), [(int(c) % p.nx, int(c) // p.nx) for c in np.asarray(bad)[~near][:10]]
    return g, want, flips


@pytest.mark.parametrize("model", MODELS)
def test_512_cross_field_1000_steps_within_1e4(torch_cuda, model):
    """The accuracy gate of the plain stepper at 512^2 from the cross-field
    start: max|d| <= 1e-4 after 1000 steps against the FP64 oracle; for FK
    with the threshold decisions separated out (_fk_gate)."""
    nx = ny = 512
    p = T.grid_params(nx, ny, model=model)
    nv = T.nvars(p)
    if model == T.YH_MODEL_2V_REF:
        f = Y.initGates(nx, ny)  # reference IC: u=1 for i<nx/8, v=1 for j>=ny/2
        pl = Y.Plan(p)
        pl.set_state(*f)
        pl.step(1000)
        got = pl.get_state()
        want = O.run(p, [x.astype(np.float64) for x in f], 1000)
        errs = [float(np.max(np.abs(got[a] - want[a]))) for a in range(nv)]
        print("max|d| per variable:", errs)
        assert max(errs) <= 1e-4, errs
    else:
        f = Y.initGates(nx, ny, u_cols=32, rest=1.0, refr=0.0, nvars=3)
        _fk_gate(p, f, 1000)


def test_1024_masked_fk_against_oracle(torch_cuda):
    """The headline workload's first 200 steps (1024^2, 200 holes, random
    start) under the same gate; holes stay 0."""
    nx = ny = 1024
    p = T.grid_params(nx, ny, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    mask = Y.make_holes(nx, ny, 200, 3.0, 12.0, Y.seed(1, 0, 1))
    f = Y.init_random(nx, ny, Y.seed(1, 0, 2), gate=1.0, nvars=3)
    got, want, _ = _fk_gate(p, f, 200, solid=mask)
    hole = mask == 0
    for a in range(3):
        assert np.all(got[a][hole] == 0)
        assert np.median(np.abs(got[a] - want[a])) < 1e-6


def test_tips_match_oracle_and_plan_log(torch_cuda):
    torch = torch_cuda
    nx = ny = 256
    p = T.grid_params(nx, ny)
    f = Y.initGates(nx, ny)
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.step(3000)
    s1 = pl.get_state()
    pl.step(1)
    s2 = pl.get_state()
    pl.track_tips()
    log = pl.read_tips()
    want, keys = O.tips(p, s1[0].astype(np.float64), s2[0].astype(np.float64), O.derive(p).dt * 3001)
    assert len(want) > 0
    assert len(log) == len(want)
    np.testing.assert_allclose(log, want, rtol=0, atol=1e-5)
    # the device-pointer wrapper gives the same set
    cap = 1024
    d_tips = torch.zeros(cap * 5, dtype=torch.float32, device="cuda")
    d_keys = torch.zeros(cap, dtype=torch.int32, device="cuda")
    d_cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    Y.tip_wrapper(p, torch.from_numpy(s1[0]).cuda(), torch.from_numpy(s2[0]).cuda(),
                  O.derive(p).dt * 3001, d_tips, d_keys, d_cnt, cap)
    torch.cuda.synchronize()
    n = int(d_cnt.item())
    assert n == len(want)
    got = d_tips.cpu().numpy().reshape(-1, 5)[:n]
    order = np.argsort(d_keys.cpu().numpy()[:n], kind="stable")
    np.testing.assert_array_equal(np.sort(d_keys.cpu().numpy()[:n]), keys)
    np.testing.assert_allclose(got[order], want, rtol=0, atol=1e-5)


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("cfg", [(T.YH_KERNEL_TILED, 8, 8, 8, 1), (T.YH_KERNEL_TILED, 8, 8, 4, 1),
                                 (T.YH_KERNEL_TILED, 9, 7, 16, 1)])
def test_tips_after_multistep_call_match_oracle(torch_cuda, cfg, masked):
    """Tips after a call of 37 steps with K = 8 or 9 per launch.  The
    previous u comes from the last launch (8 rows x 8 waves and 7 x 16: it
    writes u one substep before its end; with K = 9 that launch is the
    one-substep 37th step) or from a single-step tail (8 rows x 4 waves).  A second
    plan takes 36 steps without keeping the previous state, then 1, and
    exposes steps 3036 and 3037: the first plan's final state equals the
    second's bit for bit, and its tips are the oracle's tips between those
    two states.  With a mask the search covers the centred disk only
    (tipTracker.cu:47-51)."""
    nx = ny = 256
    # with a mask a disk of radius 100 (bounding box [29, 228)^2)
    p = T.grid_params(nx, ny, solidSwitch=int(masked), tipOffsetX=100, tipOffsetY=100)
    mask = Y.make_holes(nx, ny, 6, 3.0, 6.0, Y.seed(9, 0, 1)) if masked else None
    plans = []
    for _ in range(2):
        pl = Y.Plan(p)
        if masked:
            pl.set_mask(mask)
        pl.set_state(*Y.initGates(nx, ny))
        pl.set_kernel(*cfg)
        pl.step(3000)
        plans.append(pl)
    a, b = plans
    a.step(37)
    a.track_tips()
    log = a.read_tips()
    fa = a.get_state()
    # 36 steps in K-substep launches, then the last step as the first plan
    # takes it: inside a tiled launch (fused) or as the single-step tail
    # (a tiled cell update does not depend on where launches split: reflected
    # halo cells stay bitwise mirrors of the cells they reflect)
    b.set_option(T.YH_OPT_KEEP_PREV, 0)
    b.step(36)
    s36 = b.get_state()
    b.set_option(T.YH_OPT_KEEP_PREV, 1)
    b.step(1)
    s37 = b.get_state()
    for pl in plans:
        pl.close()
    want, _ = O.tips(p, s36[0].astype(np.float64), s37[0].astype(np.float64),
                     O.derive(p).dt * 3037)
    assert len(want) > 0
    for x, y in zip(fa, s37):
        np.testing.assert_array_equal(x, y)
    assert len(log) == len(want)
    np.testing.assert_allclose(log, want, rtol=0, atol=1e-5)


@pytest.mark.parametrize("model", MODELS)
def test_tiled_results_do_not_depend_on_launch_split(torch_cuda, model):
    """With one column per lane, reflected halo cells add their stencil terms
    in the order of the cells they mirror, so they stay bitwise mirrors and a
    cell's value after n steps does not depend on the fused depth, the tile
    shape or where calls split: masked runs of 100 steps with K = 3, 5, 8, 9
    on five shapes, and in calls of 37 + 63 (where remainders merge into
    the shape one row taller: 7 x 16 -> 8 x 16, 4 x 16 -> 5 x 16), agree
    bit for bit.  The previous
    state is not kept, so every step runs in the tiled kernel (a shape built
    without the fused previous-u output would end each call with the
    one-step kernel, whose stencil sums in another order)."""
    n = 256
    p = T.grid_params(n, n, model=model, solidSwitch=1)
    mask = Y.make_holes(n, n, 12, 3.0, 9.0, Y.seed(5, 0, 1))
    f = Y.init_random(n, n, Y.seed(5, 0, 2), gate=1.0, nvars=T.nvars(p))
    runs = []
    for cfg, calls in (((3, 8, 4, 1), [100]), ((5, 4, 16, 1), [100]), ((8, 8, 8, 1), [100]),
                       ((9, 7, 16, 1), [100]), ((9, 7, 16, 1), [37, 63]),
                       ((8, 4, 16, 1), [100]), ((8, 4, 16, 1), [37, 63]),
                       ((12, 5, 16, 1), [100])):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(T.YH_KERNEL_TILED, *cfg)
        pl.set_option(T.YH_OPT_KEEP_PREV, 0)
        for c in calls:
            pl.step(c)
        runs.append(pl.get_state())
        pl.close()
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("model", MODELS)
def test_write_through_launches_equal_plain_ones(torch_cuda, model):
    """At 1024^2, 4 x 16 launches of K = 8 (418 tiles: two workgroups on most
    CUs, all resident) and the 5 x 16 launches their remainders merge into
    write the state through to memory; at K = 16 (1024 tiles, several
    rounds) the same shape stores plainly.  A store's cache policy changes no
    value: 96 masked steps in the three forms, with and without the kept
    previous state, agree bit for bit."""
    n = 1024
    p = T.grid_params(n, n, model=model, solidSwitch=1)
    mask = Y.make_holes(n, n, 60, 3.0, 12.0, Y.seed(6, 0, 1))
    f = Y.init_random(n, n, Y.seed(6, 0, 2), gate=1.0, nvars=T.nvars(p))
    runs = []
    for K, calls, keep in ((16, [96], 0), (8, [96], 0), (8, [37, 59], 0), (8, [37, 59], 1),
                           (16, [37, 59], 1)):
        print(f"write-through case K={K} calls={calls} keep_prev={keep}", flush=True)
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(T.YH_KERNEL_TILED, K, 4, 16, 1)
        pl.set_option(T.YH_OPT_KEEP_PREV, keep)
        for c in calls:
            pl.step(c)
        runs.append(pl.get_state())
        pl.close()
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            np.testing.assert_array_equal(x, y)


def test_random_grids_shapes_and_splits_agree(torch_cuda):
    """24 seeded cases: grid sides 5..400 (ragged against every tile), model,
    mask and hole count, a one-column shape, K and a split of the steps into
    one to three calls all drawn at random.  Each run equals, bit for bit,
    the same problem stepped one substep per launch on 4 x 4 regions, and
    stays within 2e-5 of the one-step kernel (which sums the stencil in
    another order) except where FK's sharp Heaviside flips on a threshold
    that u meets to within FP32 rounding (at most 1 cell in 1000)."""
    rng = np.random.default_rng(20261202)
    shapes = [(4, 4), (8, 4), (16, 4), (4, 8), (8, 8), (16, 8), (4, 16), (5, 16), (8, 16),
              (7, 16)]
    for case in range(24):
        nx, ny = (int(x) for x in rng.integers(5, 401, size=2))
        model = MODELS[int(rng.integers(2))]
        masked = bool(rng.integers(2))
        r, w = shapes[int(rng.integers(len(shapes)))]
        kmax = min(12, (r * w - 1) // 2, min(nx, ny) - 1)
        K = int(rng.integers(1, kmax + 1))
        steps = int(rng.integers(1, 41))
        cuts = sorted(int(x) for x in rng.integers(0, steps + 1, size=int(rng.integers(0, 3))))
        calls = [b - a for a, b in zip([0] + cuts, cuts + [steps]) if b > a]
        print(f"case {case}: {nx}x{ny} model {model} masked {masked} shape {r}x{w} K {K} "
              f"calls {calls}", flush=True)
        p = T.grid_params(nx, ny, model=model, solidSwitch=int(masked))
        mask = (Y.make_holes(nx, ny, int(rng.integers(0, 6)), 1.0, 4.0, Y.seed(7, case, 1))
                if masked else None)
        f = Y.init_random(nx, ny, Y.seed(7, case, 2), gate=1.0, nvars=T.nvars(p))
        out = []
        for cfg, parts in (((T.YH_KERNEL_TILED, K, r, w, 1), calls),
                           ((T.YH_KERNEL_TILED, 1, 4, 4, 1), [steps]),
                           ((T.YH_KERNEL_DIRECT, 1, 0, 0, 1), [steps])):
            pl = Y.Plan(p)
            if masked:
                pl.set_mask(mask)
            pl.set_state(*f)
            pl.set_kernel(*cfg)
            pl.set_option(T.YH_OPT_KEEP_PREV, 0)
            for c in parts:
                pl.step(c)
            out.append(pl.get_state())
            pl.close()
        for a in range(len(out[0])):
            np.testing.assert_array_equal(out[0][a], out[1][a])
            d = np.abs(out[0][a] - out[2][a])
            assert np.isfinite(out[0][a]).all() and np.mean(d > 2e-5) < 1e-3, (a, d.max())


def test_remainder_merge_is_bitwise_equal(torch_cuda):
    """A 100-step call with K = 9 on 7 rows x 16 waves ends with one launch of
    10 substeps on 8 rows x 16 waves (the remainder merged into the last full
    launch: at 512^2 that shape needs 60 tiles against 72).  State and tips
    equal those of eleven 9-step calls and one 1-step call, which merge
    nothing."""
    n = 512
    # the masked tip search covers the centred disk (tipTracker.cu:47-51):
    # radius 363 takes in the whole grid, where the end of the broken front
    # of the FK cross-field start is a tip from the first samples on
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1, tipOffsetX=363,
                      tipOffsetY=363)
    mask = Y.make_holes(n, n, 50, 3.0, 12.0, Y.seed(1, 0, 1))
    f = Y.initGates(n, n, rest=1.0, refr=0.0, nvars=3)
    out = []
    for calls in ([100], [9] * 11 + [1]):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(T.YH_KERNEL_TILED, 9, 7, 16, 1)
        for rep in range(30):
            for c in calls:
                pl.step(c)
            pl.track_tips()
        out.append((pl.get_state(), pl.read_tips()))
        pl.close()
    for x, y in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(x, y)
    assert len(out[0][1]) > 0
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_autotune_keeps_the_state_and_installs_its_choice(torch_cuda):
    """yh_autotune times candidates on the plan's state: the state and the
    step count come back bit for bit, the returned configuration is the one
    installed, and stepping on from there equals a plan that was set to that
    configuration directly.  The tip test needs a new step first (the
    buffer of the previous state held candidates' outputs)."""
    n = 256
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    mask = Y.make_holes(n, n, 10, 3.0, 9.0, Y.seed(3, 0, 1))
    f = Y.init_random(n, n, Y.seed(3, 0, 2), gate=1.0, nvars=3)
    out = []
    for tune in (True, False):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.step(37)
        if tune:
            before = pl.get_state()
            cfg, ms = pl.autotune(steps=64)
            assert ms > 0
            for x, y in zip(before, pl.get_state()):
                np.testing.assert_array_equal(x, y)
            assert pl.step_count() == 37
            got = pl.get_kernel()
            assert (got.variant, got.substeps, got.rows, got.waves, got.cols) == cfg
            with pytest.raises(RuntimeError):
                pl.track_tips()
        else:
            pl.set_kernel(*cfg)
        pl.step(10)
        pl.track_tips()
        out.append(pl.get_state())
        pl.close()
    for x, y in zip(*out):
        np.testing.assert_array_equal(x, y)


def test_autotune_times_whole_calls_of_the_given_size(torch_cuda):
    """yh_autotune's steps is the call size: each call size is tuned and
    cached on its own, and a 20-step call on a chosen tiled configuration
    gives the bits of the same 20 steps on the default one (tiled results do
    not depend on the shape, the depth or where launches split; the one-step
    kernel, if chosen, agrees to rounding)."""
    n = 384
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    mask = Y.make_holes(n, n, 20, 3.0, 9.0, Y.seed(4, 0, 1))
    f = Y.init_random(n, n, Y.seed(4, 0, 2), gate=1.0, nvars=3)
    got = []
    for steps in (20, 100, 20):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        cfg, ms = pl.autotune(steps=steps)
        assert ms > 0 and cfg[0] in (T.YH_KERNEL_DIRECT, T.YH_KERNEL_TILED)
        pl.step(20)
        got.append((cfg, pl.get_state()))
        pl.close()
    assert got[0][0] == got[2][0]  # the 20-step choice comes from the cache
    ref = Y.Plan(p)
    ref.set_mask(mask)
    ref.set_state(*f)
    ref.step(20)
    want = ref.get_state()
    ref.close()
    for cfg, st in got:
        for x, y in zip(st, want):
            if cfg[0] == T.YH_KERNEL_TILED:
                np.testing.assert_array_equal(x, y)
            else:
                np.testing.assert_allclose(x, y, atol=1e-5)


@pytest.mark.parametrize("n,want", [(256, (16, 4, 16)), (512, (16, 4, 16)), (540, (15, 4, 16)),
                                    (1024, (8, 4, 16)), (4096, (8, 8, 8))])
def test_auto_kernel_follows_the_grid(torch_cuda, n, want):
    """Without timing, AUTO takes 4 x 16 at K=8 where 64 x 112 regions at
    K=9 would fill between half and all of the 256 CUs and two 4 x 16
    workgroups fit on a CU (1024^2: 484 tiles on 512 places), 8 x 8 waves
    at K=8 when there are more tiles than CUs (4096^2), else 4 x 16 at the
    largest K <= 16 whose tiles number at most 256 (512^2: 16 x 16 tiles of
    32 x 32 outputs; 540^2: K=15, 16 x 16 tiles of 34 x 34)."""
    pl = Y.Plan(T.grid_params(n, n, model=T.YH_MODEL_3V_FK))
    c = pl.get_kernel()
    pl.close()
    assert (c.variant, c.substeps, c.rows, c.waves, c.cols) == (T.YH_KERNEL_TILED, *want, 1)


def test_check_finite_and_argument_errors(torch_cuda):
    p = T.grid_params(64, 64)
    pl = Y.Plan(p)
    u, v = Y.initGates(64, 64)
    pl.set_state(u, v)
    pl.step(5)
    assert pl.check_finite() == 0
    u2 = u.copy()
    u2[100] = np.nan
    pl.set_state(u2, v)
    assert pl.check_finite() == T.YH_ENONFINITE
    with pytest.raises(RuntimeError):
        pl.step(-1)
    with pytest.raises(RuntimeError):
        pl.track_tips()  # no previous state after set_state
    pm = Y.Plan(T.grid_params(64, 64, solidSwitch=1))
    pm.set_state(u, v)
    with pytest.raises(RuntimeError):
        pm.step(1)  # mask required first


@pytest.mark.parametrize("shape", [(4, 4), (5, 7), (1000, 5), (6, 700), (129, 65), (257, 3 + 64 * 2)])
@pytest.mark.parametrize("model", MODELS)
def test_small_and_ragged_grids_match_oracle(torch_cuda, shape, model):
    """Grids smaller than a tile, thinner than the halo, and ragged edges: the
    tiled kernel (K clamped to the grid) against the FP64 oracle."""
    nx, ny = shape
    p = T.grid_params(nx, ny, model=model, solidSwitch=1)
    rng = np.random.default_rng(nx * 1000 + ny)
    mask = (rng.random(nx * ny) > 0.15).astype(np.uint8)
    f = _fields(rng, p)
    for cfg in ((T.YH_KERNEL_TILED, 8, 8, 8, 1), (T.YH_KERNEL_TILED, 3, 4, 16, 2),
                (T.YH_KERNEL_DIRECT, 1, 0, 0, 1)):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(*cfg)
        pl.step(30)
        got = pl.get_state()
        pl.close()
        want = O.run(p, [x.astype(np.float64) for x in f], 30, solid=mask)
        for a in range(len(f)):
            d = np.abs(got[a] - want[a])
            assert np.all(got[a][mask == 0] == 0)
            if model == T.YH_MODEL_2V_REF:
                assert d.max() < 1e-4, (cfg, a, d.max())
            else:
                assert np.mean(d > 1e-4) < 2e-3, (cfg, a, d.max())


def test_4096_tiled_equals_one_step_kernel(torch_cuda):
    """The largest bench grid (4096^2, 3200 holes): the headline tile shape and
    the one-step kernel stay together over 48 steps; holes stay zero."""
    n = 4096
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    mask = Y.make_holes(n, n, 3200, 3.0, 12.0, Y.seed(1, 1, 1))
    f = Y.init_random(n, n, Y.seed(1, 1, 2), gate=1.0, nvars=3)
    out = []
    for cfg in ((T.YH_KERNEL_TILED, 8, 8, 8, 1), (T.YH_KERNEL_DIRECT, 1, 0, 0, 1)):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(*cfg)
        pl.step(48)
        assert pl.check_finite() == 0
        out.append(pl.get_state())
        pl.close()
    hole = mask == 0
    for a in range(3):
        assert np.all(out[0][a][hole] == 0) and np.all(out[1][a][hole] == 0)
        d = np.abs(out[0][a] - out[1][a])
        print(a, d.max(), float(np.mean(d > 1e-5)))
        assert np.mean(d > 1e-5) < 1e-4 and np.median(d) < 1e-7


def test_max_grid_tiled_equals_one_step_kernel(torch_cuda):
    """The largest grid the ABI accepts, 2^30 cells (32768^2, 4 GiB per
    variable, so byte offsets pass 2^32), at the 1024^2 config's hole density
    (204800 holes): the headline tile shape against the one-step kernel over
    16 steps; holes stay zero, and the last rows (the largest offsets) agree
    as the first rows do."""
    n = 32768
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    with pytest.raises(RuntimeError):
        Y.Plan(T.grid_params(n + 1, n, model=T.YH_MODEL_3V_FK, solidSwitch=1))
    mask = Y.make_holes(n, n, 200 * 1024, 3.0, 12.0, Y.seed(1, 1, 3))
    f = Y.init_random(n, n, Y.seed(1, 1, 4), gate=1.0, nvars=3)
    out = []
    for cfg in ((T.YH_KERNEL_TILED, 8, 8, 8, 1), (T.YH_KERNEL_DIRECT, 1, 0, 0, 1)):
        pl = Y.Plan(p)
        pl.set_mask(mask)
        pl.set_state(*f)
        pl.set_kernel(*cfg)
        pl.step(16)
        assert pl.check_finite() == 0
        out.append(pl.get_state())
        pl.close()
    del f
    holes = np.flatnonzero(mask == 0)
    del mask
    edge = 64 * n
    for a in range(3):
        assert not np.any(out[0][a][holes]) and not np.any(out[1][a][holes])
        d = np.abs(out[0][a] - out[1][a])
        # an offset bug would spoil whole rows; a cell whose u sits within
        # rounding of a gate threshold may switch in one kernel only
        far = float(np.count_nonzero(d[-edge:] > 1e-5)) / edge
        near = float(np.count_nonzero(d[:edge] > 1e-5)) / edge
        frac = float(np.count_nonzero(d > 1e-5)) / d.size
        med = float(np.median(d[::1021]))
        print(a, float(d.max()), near, far, frac, med, flush=True)
        assert frac < 1e-4 and med < 1e-7
        assert far < 1e-3 and near < 1e-3
        del d
