"""Writes tests/golden/long_run_stats.json: long-run statistics of two plain
stepper workloads computed by the FP64 oracle.

  headline        FK, 1024 x 1024, 200 seeded circular holes (radii 3-12
                  cells), random u in [0, 1) with v = w = 1, 50000 steps;
                  tips searched in the centred disk of radius^2
                  tipOffsetX*tipOffsetY, as the masked branch of
                  spiralTip_kernel does (tipTracker.cu:47-51).  The random
                  start excites the tissue once and nothing re-enters, so
                  there are no tips and one action potential per cell.
  cross_field_512 FK, 512 x 512, no holes, the cross-field start (u = 1 for
                  i < 32, v = w = 0 for j >= ny/2), 10000 steps: a spiral
                  whose tip is tracked over the interior (tipTracker.cu:126).

Both use dx = 0.025 cm, dt = 0.02 ms; every 100 steps the tips between the
last two states are recorded (the statistics count the genuine ones, see
genuine()), and every step updates the per-cell
action-potential durations of sAPD_kernel (spaceAPD.cu:278-374, threshold
0.15).  tests/test_gpu_stats.py computes the same statistics from the library
and compares them within 1%.

Run from the repository root (about 20 minutes on 8 cores):
    python tests/golden/make_long_run_stats.py [workload ...]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.oracle as O  # noqa: E402
import yolohtli_amd as Y  # noqa: E402
from yolohtli_amd import abi_types as T  # noqa: E402

TIP_EVERY = 100
WORKLOADS = {"headline": 50000, "cross_field_512": 10000}


def workload(name):
    """(params, mask or None, initial fields, steps) of a named workload."""
    if name == "headline":
        n = 1024
        p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
        mask = Y.make_holes(n, n, 200, 3.0, 12.0, Y.seed(1, 0, 1))
        f = Y.init_random(n, n, Y.seed(1, 0, 2), gate=1.0, nvars=3)
        return p, mask, f, WORKLOADS[name]
    n = 512
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK)
    f = Y.initGates(n, n, u_cols=32, rest=1.0, refr=0.0, nvars=3)
    return p, None, f, WORKLOADS[name]


def genuine(tips, u_past, u_present, nx, ny, uth):
    """Whether each tip's cell holds a spiral tip: the isolines u = uth of
    both fields and the zero line of u_present - u_past cross it (corners
    (i, j), (i+1, j), (i+1, j+1), (i, j+1), the cell itself past the last
    column or row, as the tip test reads them).  The closed form of
    spiralTip_kernel (tipTracker.cu:86-97) also returns roots in (0, 1)^2
    where the two isolines nearly coincide (along a travelling front, and on
    plateaus the isoline does not reach); which of those appear depends on
    rounding (FP64 finds thousands where FP32 finds none), so the statistics
    count the tips where u rises on one side and falls on the other."""
    tips = np.asarray(tips, np.float64).reshape(-1, 2)
    if not len(tips):
        return np.zeros(0, bool)
    i = np.clip(np.floor(tips[:, 0]).astype(int), 0, nx - 1)
    j = np.clip(np.floor(tips[:, 1]).astype(int), 0, ny - 1)
    i1, j1 = np.minimum(i + 1, nx - 1), np.minimum(j + 1, ny - 1)

    def corners(u):
        U = np.asarray(u, np.float64).reshape(ny, nx)
        return np.stack([U[j, i], U[j, i1], U[j1, i1], U[j1, i]], 1)

    a, b = corners(u_past), corners(u_present)
    d = b - a
    return ((a.min(1) < uth) & (a.max(1) > uth) & (b.min(1) < uth) & (b.max(1) > uth)
            & (d.min(1) < 0) & (d.max(1) > 0))


def summarize(tip_rows, apd1, apd2, mask, samples):
    """Statistics of a run: tips as (sample, x, y, on the isoline) rows;
    per-cell APDs.  Tip position statistics use the tips on the isoline."""
    raw = np.asarray(tip_rows, np.float64).reshape(-1, 4)
    tips = raw[raw[:, 3] > 0, :3]
    tissue = np.ones(np.asarray(apd1).size, bool) if mask is None else \
        np.asarray(mask).reshape(-1) > 0
    a1 = np.asarray(apd1, np.float64).reshape(-1)[tissue]
    a2 = np.asarray(apd2, np.float64).reshape(-1)[tissue]
    per_sample = np.bincount(tips[:, 0].astype(int), minlength=samples) if len(tips) else \
        np.zeros(samples, int)
    out = {"samples": int(samples), "tips": int(len(tips)), "tips_raw": int(len(raw)),
           "samples_with_tips": int(np.count_nonzero(per_sample)),
           "first_sample_with_tips": int(np.argmax(per_sample > 0)) if len(tips) else -1,
           "last_sample_with_tips": int(samples - 1 - np.argmax(per_sample[::-1] > 0))
           if len(tips) else -1,
           "apd1_cells": int(np.count_nonzero(a1 > 0)), "apd2_cells": int(np.count_nonzero(a2 > 0)),
           "apd1_mean_ms": float(a1[a1 > 0].mean()) if np.any(a1 > 0) else 0.0,
           "apd2_mean_ms": float(a2[a2 > 0].mean()) if np.any(a2 > 0) else 0.0}
    if len(tips):
        out.update({"tip_x_mean": float(tips[:, 1].mean()), "tip_y_mean": float(tips[:, 2].mean()),
                    "tip_x_std": float(tips[:, 1].std()), "tip_y_std": float(tips[:, 2].std())})
    return out


def run(name):
    p, mask, f, steps = workload(name)
    x = [a.astype(np.float64) for a in f]
    st = None
    rows = []
    samples = steps // TIP_EVERY
    t0 = time.time()
    for k in range(samples):
        x, st = O.run_apd(p, x, TIP_EVERY - 1, solid=mask, count0=k * TIP_EVERY, state=st)
        prev = x[0].copy()
        x, st = O.run_apd(p, x, 1, solid=mask, count0=(k + 1) * TIP_EVERY - 1, state=st)
        tt, _ = O.tips(p, prev, x[0], 0.0)
        on = genuine(tt[:, :2], prev, x[0], p.nx, p.ny, p.Uth)
        rows.extend((k, float(t[0]), float(t[1]), float(o)) for t, o in zip(tt, on))
        if k % 50 == 49:
            print(f"{name}: {(k + 1) * TIP_EVERY} steps, {len(rows)} tips, "
                  f"{time.time() - t0:.0f} s", flush=True)
    stats = summarize(rows, st["apd1"], st["apd2"], mask, samples)
    stats.update({"steps": steps, "n": p.nx, "tip_every": TIP_EVERY})
    return stats


def main(names):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "long_run_stats.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for name in names:
        out[name] = run(name)
        print(name, json.dumps(out[name], sort_keys=True), flush=True)
    out["source"] = "FP64 oracle (oracle/oracle.c), tests/golden/make_long_run_stats.py"
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
