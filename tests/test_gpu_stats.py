"""Long-run statistics on the GPU against the FP64 oracle.

FP32 and FP64 trajectories of an excitable medium drift apart over thousands
of steps, so these gates compare statistics, as the accuracy target for long
runs states: action-potential durations (the fused sAPD update) and the
spiral-tip trajectory, each within 1%.
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", [(T.YH_KERNEL_TILED, 8, 4, 16, 1), (T.YH_KERNEL_TILED, 6, 8, 8, 1),
                                 (T.YH_KERNEL_DIRECT, 1, 0, 0, 1)])
def test_fused_apd_matches_oracle(torch_cuda, cfg):
    n, nsteps = 64, 20000
    p = T.grid_params(n, n)
    f = Y.initGates(n, n)
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.set_kernel(*cfg)
    pl.apd_enable(0.15)
    for _ in range(nsteps // 1000):  # several calls: the step count carries over
        pl.step(1000)
    g = pl.apd_get()
    _, st = O.run_apd(p, [x.astype(np.float64) for x in f], nsteps)
    a_g, a_o = g["apd1"].astype(np.float64), st["apd1"]
    both = (a_g != 0) & (a_o != 0)
    assert both.sum() > 1000
    assert np.mean((a_g != 0) != (a_o != 0)) < 0.01
    rel_mean = abs(a_g[both].mean() - a_o[both].mean()) / a_o[both].mean()
    print("cells", both.sum(), "mean APD gpu", a_g[both].mean(), "oracle", a_o[both].mean())
    assert rel_mean < 0.01
    assert np.quantile(np.abs(a_g[both] - a_o[both]), 0.99) < 0.01 * a_o[both].mean()


def test_apd_with_mask_skips_holes(torch_cuda):
    n = 64
    p = T.grid_params(n, n, solidSwitch=1)
    mask = Y.make_holes(n, n, 5, 3.0, 6.0, 7)
    f = Y.initGates(n, n)
    pl = Y.Plan(p)
    pl.set_mask(mask)
    pl.set_state(*f)
    pl.apd_enable(0.15)
    pl.step(12000)
    g = pl.apd_get()
    _, st = O.run_apd(p, [x.astype(np.float64) for x in f], 12000, solid=mask)
    assert np.all(g["apd1"][mask == 0] == 0)
    both = (g["apd1"] != 0) & (st["apd1"] != 0)
    assert both.sum() > 100
    assert abs(g["apd1"][both].mean() - st["apd1"][both].mean()) < 0.01 * st["apd1"][both].mean()


def _physical(tips, past, present, n, uth=0.7):
    """Tips whose cell has Uth between the corner extremes of both fields.

    The closed-form root of tipTracker.cu:86-97 can land inside (0,1)^2 by
    cancellation where both fields are nearly flat far below Uth; FP64 keeps
    exponentially small tails ahead of a wave that FP32 underflows, so such
    roots differ between the two and are no spiral tips."""
    keep = []
    P, Q = past.reshape(n, n), present.reshape(n, n)

    for t in tips:
        i, j = int(t[0]), int(t[1])
        cp = P[j:j + 2, i:i + 2]
        cq = Q[j:j + 2, i:i + 2]
        keep.append(cp.min() < uth < cp.max() and cq.min() < uth < cq.max())
    return tips[np.array(keep, bool)] if len(tips) else tips


def _persistent(samples, radius=2.0):
    """Tips that have a tip within `radius` cells in the previous or the next
    sample.  u^n and u^{n+1} isolines are one step apart, so along a
    translating front they nearly coincide and sporadic roots there are
    decided by rounding; a spiral tip persists from sample to sample."""
    out = []
    for q, s in enumerate(samples):
        nb = [samples[r] for r in (q - 1, q + 1) if 0 <= r < len(samples) and len(samples[r])]
        for t in s:
            if any(np.min(np.hypot(o[:, 0] - t[0], o[:, 1] - t[1])) <= radius for o in nb):
                out.append(t)
    return np.array(out).reshape(-1, 5)


def test_tip_trajectory_statistics(torch_cuda):
    """Broken-wave tip from the reference's cross-field start, every 100 steps."""
    n, warm, nsteps, every = 256, 6000, 12000, 100
    p = T.grid_params(n, n)
    f = Y.initGates(n, n)
    c = O.derive(p)
    pl = Y.Plan(p)
    pl.set_state(*f)
    pl.step(warm)
    gt = []
    for q in range(nsteps // every):
        pl.step(every - 1)
        prev = pl.get_state()[0]
        pl.step(1)
        cur = pl.get_state()[0]
        pl.clear_tips()
        pl.track_tips()
        gt.append(_physical(pl.read_tips(), prev, cur, n))
    gt = _persistent(gt)
    X = [x.astype(np.float64) for x in f]
    X = O.run(p, X, warm)
    ot = []
    for q in range(nsteps // every):
        X = O.run(p, X, every - 1)
        prev = X[0].copy()
        X = O.run(p, X, 1)
        t, _ = O.tips(p, prev, X[0], c.dt * (warm + (q + 1) * every))
        ot.append(_physical(t, prev, X[0], n))
    ot = _persistent(ot)
    print("tips gpu", len(gt), "oracle", len(ot))
    assert len(ot) > 50
    assert abs(len(gt) - len(ot)) <= 0.01 * len(ot) + 2
    for k in (0, 1):
        mg, mo = gt[:, k].mean(), ot[:, k].mean()
        sg, so = gt[:, k].std(), ot[:, k].std()
        print("axis", k, "mean", mg, mo, "std", sg, so)
        assert abs(mg - mo) <= 0.01 * mo
        assert abs(sg - so) <= 0.01 * so + 0.05


def _long_run_stats():
    import json
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_long_run_stats as S
    with open(os.path.join(here, "long_run_stats.json")) as fh:
        return S, json.load(fh)


@pytest.mark.parametrize("name", ["headline", "cross_field_512"])
def test_long_run_statistics_match_fp64(torch_cuda, name):
    """The long-run statistics of tests/golden/long_run_stats.json (FP64
    oracle, tests/golden/make_long_run_stats.py) from the library: the
    headline workload (FK, 1024^2, 200 holes, random start, 50000 steps, tips
    every 100 in the centred disk, APD at 0.15) and the 512^2 cross-field
    spiral (10000 steps, tips over the interior).  Counts and means within
    1%; the tip positions' mean and spread within 1% of the grid size.  Tips
    count where they are genuine (the closed form's roots where the two
    isolines nearly coincide depend on rounding: make_long_run_stats.genuine);
    the raw counts are printed.  Without tips in the FP64 run the GPU must log
    none either."""
    S, ref = _long_run_stats()
    want = ref[name]
    p, mask, f, steps = S.workload(name)
    pl = Y.Plan(p)
    if mask is not None:
        pl.set_mask(mask)
    pl.set_state(*f)
    pl.apd_enable(0.15)
    samples = steps // S.TIP_EVERY
    rows = []
    fields = want["tips_raw"] > 0  # the filter reads both states at each sample
    for k in range(samples):
        pl.step(S.TIP_EVERY - 1)
        past = pl.get_state()[0] if fields else None
        pl.step(1)
        pl.track_tips()
        tips = pl.read_tips()
        if len(tips):
            on = (S.genuine(tips[:, :2], past, pl.get_state()[0], p.nx, p.ny, p.Uth) if fields
                  else np.ones(len(tips), bool))
            rows.extend((k, float(t[0]), float(t[1]), float(o)) for t, o in zip(tips, on))
            pl.clear_tips()
    a = pl.apd_get()
    pl.close()
    got = S.summarize(rows, a["apd1"], a["apd2"], mask, samples)
    print(name, "GPU", got, "FP64", want)
    if want["tips_raw"] == 0:
        assert got["tips_raw"] == 0
    for key in ("tips", "samples_with_tips", "apd1_cells", "apd2_cells", "apd1_mean_ms",
                "apd2_mean_ms"):
        assert got[key] == pytest.approx(want[key], rel=0.01, abs=1e-9), key
    for key in ("tip_x_mean", "tip_y_mean", "tip_x_std", "tip_y_std"):
        if key in want:
            assert abs(got[key] - want[key]) <= 0.01 * p.nx, key
