"""Benchmark of the plain tissue stepper on one MI355X.

Workload (the 1024x1024 headline configuration): Fenton-Karma 3V, FP32,
1024x1024 cells, dx = 0.025 cm, D = 0.001 cm^2/ms, dt = 0.02 ms, no-flux
boundaries, 200 seeded circular holes (radii U[3,12] cells), u ~ U[0,1) with
v = w = 1, tip tracking every 100 steps.  A "step" is one explicit-Euler time
step of the whole tissue; the run is --steps such steps after --warmup.

Printed: one JSON line with Gcell-steps/s, the roofline of the dominant kernel
(timed live with HIP events on the stream it runs on) and the CPU baseline
(the oracle's OpenMP loop of the same update on this host).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E peak


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--holes", type=int, default=200)
    ap.add_argument("--model", choices=["fk", "2v"], default="fk")
    ap.add_argument("--tip-every", type=int, default=100)
    ap.add_argument("--cfg", default="auto", help="auto or K,rows,waves,cols")
    ap.add_argument("--kernel-launches", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="timed CPU-baseline seconds, split between FP32 and FP64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sr-n", type=int, default=512)
    ap.add_argument("--sr-steps", type=int, default=20000)
    ap.add_argument("--no-sr", action="store_true")
    ap.add_argument("--control-n", type=int, default=4096,
                    help="side of the non-cache-resident control grid")
    ap.add_argument("--no-control", action="store_true")
    ap.add_argument("--batch-tissues", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--batch-each-steps", type=int, default=1000,
                    help="steps of the per-tissue-launch comparison")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adj-n", type=int, default=256)
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--adj-fd", type=int, default=1, help="FP64 FD check of tau gradients")
    ap.add_argument("--adj-fd-taus", type=lambda s: [int(x) for x in s.split(",")],
                    default=[0, 1, 2, 3, 4, 5])
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes behind roofline.traffic and "
                         "roofline.valu_issue")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the dominant-kernel timing (for profiling); needs --cfg")
    ap.add_argument("--batch-kernel-only", action="store_true",
                    help="only --kernel-launches launches of the batch kernel (for profiling)")
    ap.add_argument("--batch-cfg", default="6,8,8,1,0",
                    help="K,rows,waves,cols,persistent of --batch-kernel-only")
    ap.add_argument("--c0-steps", type=int, default=10000,
                    help="steps of the 512^2 cross-field leg (0 skips it)")
    return ap.parse_args()


def cpu_baseline(p, fields, mask, seconds):
    """The oracle's restatement of the update as an OpenMP host loop, in a
    child process with OMP_PROC_BIND=close and OMP_PLACES=cores: 100 warm-up
    steps, then the median of five timed windows, FP32 and FP64."""
    import subprocess
    import tempfile
    nv = len(fields)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "workload.npz")
        np.savez(path, nx=p.nx, ny=p.ny, model=p.model,
                 mask=mask if mask is not None else np.zeros(0, np.uint8),
                 u=fields[0], v=fields[1], w=fields[2] if nv > 2 else np.zeros(0, np.float32))
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", path,
                            "--seconds", str(seconds)], cwd=ROOT, env=env,
                           capture_output=True, text=True, timeout=30 * seconds + 300)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline failed:\n" + r.stderr[-2000:])
    c = json.loads(r.stdout.strip().splitlines()[-1])
    f32, f64 = c["fp32"], c["fp64"]
    return {"value": round(f32["gcell_steps_per_s"], 4), "unit": "Gcell-steps/s",
            "cores": c["threads"], "kind": "port",
            "sample": f"{p.nx}x{p.ny} same model/mask/IC; after 100 warm-up steps, median of "
                      f"5 windows of {f32['steps_per_window']} steps (FP32) and "
                      f"{f64['steps_per_window']} steps (FP64); {c['threads']} OpenMP threads, "
                      f"OMP_PROC_BIND=close, OMP_PLACES=cores; {c['cpu']}",
            "fp32_GBps": round(f32["GBps"], 2),
            "fp64_gcell_steps_per_s": round(f64["gcell_steps_per_s"], 4),
            "fp64_GBps": round(f64["GBps"], 2)}


def copy_bandwidth(stream, mib=256, reps=20):
    """Device-to-device copy of `mib` MiB (read + write counted), HIP events on
    the stream: the bandwidth a plain streaming kernel reaches on this GPU."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * (mib << 20) / (ms / 1e3) / 1e9


def _kernel_only_args(a, cfg, nx=None, holes=None, launches=50):
    """Arguments of the --kernel-only run of the single-tissue stepper."""
    return ["--kernel-only", "--cfg", ",".join(str(x) for x in cfg[1:]),
            "--kernel-launches", str(launches), "--nx", str(nx or a.nx),
            "--holes", str(a.holes if holes is None else holes), "--model", a.model]


def _counter_pass(a, args, counter):
    """One rocprofv3 pass with a single counter over a --kernel-only run of
    bench.py with `args` (the program itself follows "--"; KILL after 120 s).
    Returns {kernel name: [values per dispatch]} or None."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    d = tempfile.mkdtemp(prefix="yh_traffic_", dir="/tmp")
    cmd = ["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--kernel-trace",
           "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
           os.path.abspath(__file__)] + list(args)
    try:
        subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=150)
    except (subprocess.TimeoutExpired, OSError):
        pass
    out = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") == counter:
                    out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    shutil.rmtree(d, ignore_errors=True)
    return out or None


def hbm_traffic(a, args, kernel="k_tiled"):
    """HBM bytes per launch of a kernel: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch, one rocprofv3 pass each) over the --kernel-only run given by
    `args`, each scaled by what the same counter reads for the 256 MiB device
    copy that run makes first (256 MiB read and 256 MiB written).  None if a
    pass fails."""
    total = 0.0
    parts = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = _counter_pass(a, args, counter)
        if not rows:
            return None
        kern = [v for k, vs in rows.items() if kernel in k for v in vs]
        copy = sorted(v for k, vs in rows.items() if "copyBuffer" in k for v in vs)
        if not kern or not copy:
            return None
        copy_kib = float(np.median(copy[-3:]))  # the three 256 MiB copies
        if copy_kib <= 0:
            return None
        factor = (256 << 10) / copy_kib
        kib = float(np.mean(kern[len(kern) // 5:]))  # after the first launches
        parts[counter] = {"kib_per_dispatch": round(kib, 1), "copy_kib": round(copy_kib, 1),
                          "scale": round(factor, 4)}
        total += kib * factor * 1024.0
    return total, parts


VALU_CLOCK_GHZ = 2.4  # MI355X peak engine clock


def valu_issue(a, cfg, t_launch_ms):
    """VALU wave instructions per launch of the dominant kernel (SQ_INSTS_VALU,
    one rocprofv3 pass) and the time the chip's SIMDs need to issue them at
    the peak engine clock: a wave instruction occupies a 16-lane SIMD for 4
    cycles, 4 SIMDs per CU.  Transcendentals take longer, so `frac` (issue
    time / launch time) is a lower bound on how busy the VALUs are.  None if
    the pass fails."""
    import torch
    rows = _counter_pass(a, _kernel_only_args(a, cfg), "SQ_INSTS_VALU")
    if not rows:
        return None
    kern = [v for k, vs in rows.items() if "k_tiled" in k for v in vs]
    if not kern:
        return None
    insts = float(np.mean(kern[len(kern) // 5:]))
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    issue_us = insts * 4.0 / (4 * cus) / (VALU_CLOCK_GHZ * 1e3)
    return {"counter": "SQ_INSTS_VALU", "insts_per_launch": round(insts), "cus": cus,
            "clock_GHz": VALU_CLOCK_GHZ, "issue_us": round(issue_us, 3),
            "frac": round(issue_us / (t_launch_ms * 1e3), 4)}


def kernel_timing(a, pl, cfg, stream, n, bytes_per_cell, launches=None):
    """Average duration of the dominant kernel: `launches` (default
    --kernel-launches) back-to-back launches of K fused steps each (no
    single-step tail, no tip kernel), bracketed by two HIP events on the
    stream the kernel runs on."""
    import torch
    from yolohtli_amd import abi_types as T
    pl.set_option(T.YH_OPT_KEEP_PREV, 0)
    K = cfg[1] if cfg[0] == T.YH_KERNEL_TILED else 1
    for _ in range(20):
        pl.step(K, stream)
    nl = launches or a.kernel_launches
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    pl.step(K * nl, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t_launch_ms = e0.elapsed_time(e1) / nl
    pl.set_option(T.YH_OPT_KEEP_PREV, 1)
    alg_bytes = n * bytes_per_cell  # per launch, independent of K
    achieved = alg_bytes / (t_launch_ms / 1e3) / 1e9
    return t_launch_ms, K, alg_bytes, achieved


def symmetry_reduction(a):
    """Symmetry-reduction configuration: 512^2 FK, cross-field start (u=1 for
    i<32, v=w=1 for j<ny/2, 0 above), plain steps until a tip exists (the
    count is recorded), then --sr-steps co-moving-frame steps at dt=0.01 ms,
    each two launches (disk centre + slices with the inner products; 3x3
    solve + the three BFECC phases + the next step's RD and tip test);
    steady steps replayed from captured hipGraphs of 8 steps."""
    import torch
    import yolohtli_amd as Y
    from yolohtli_amd import abi_types as T
    n = a.sr_n
    p0 = T.grid_params(n, n, model=T.YH_MODEL_3V_FK)
    pl = Y.Plan(p0)
    pl.set_state(*Y.initGates(n, n, u_cols=32, rest=1.0, refr=0.0, nvars=3))
    warm, tips = 0, []
    while warm < 20000 and not len(tips):
        pl.step(100)
        warm += 100
        pl.track_tips()
        tips = pl.read_tips()
    st = pl.get_state()
    pl.close()
    if not len(tips):
        raise RuntimeError("no spiral tip after 20000 plain steps")
    tip0 = (float(tips[-1, 0]), float(tips[-1, 1]))
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, reduceSym=1)
    pl = Y.Plan(p)
    pl.set_state(*st)
    stream = torch.cuda.current_stream()
    pl.sr_begin(tip0, stream)
    pl.sr_step(100, stream)  # eager first steps, graph capture, warm-up
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    pl.sr_step(a.sr_steps, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    info = pl.sr_info()
    finite = pl.check_finite() == 0
    pl.close()
    rate = n * n * a.sr_steps / (ms / 1e3) / 1e9
    log(f"SR: {a.sr_steps} steps at {n}^2 in {ms:.1f} ms = {rate:.2f} Gcell-steps/s "
        f"({1e3 * ms / a.sr_steps:.2f} us/step); c={info['c']}")
    if not finite:
        raise RuntimeError("SR state is not finite")
    # two state passes per step (RD, BFECC), 2*nvars*4 B each
    bpc = 2 * (2 * 3 * 4)
    return {"n": n, "model": "FK 3V", "plain_warmup_steps": warm, "steps": a.sr_steps,
            "ms": round(ms, 3), "us_per_step": round(1e3 * ms / a.sr_steps, 3),
            "gcell_steps_per_s": round(rate, 3), "alg_bytes_per_cell_step": bpc,
            "GBps_alg": round(rate * bpc, 1), "graph": True,
            "c": [float(x) for x in info["c"]], "phi": [float(x) for x in info["phi"]],
            "singular": int(info["status"] & 1)}


def control_grid(a, cfg):
    """Non-cache-resident control: the headline kernel, the 8 x 8 shape at
    K=8 and the one-step kernel on a --control-n^2 grid (FK, holes at the
    headline density), whose state exceeds the 256 MB Infinity Cache; launch
    time by HIP events."""
    import torch
    import yolohtli_amd as Y
    from yolohtli_amd import abi_types as T
    n = a.control_n
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK, solidSwitch=1)
    holes = int(round(a.holes * (n / 1024.0) ** 2))
    pl = Y.Plan(p)
    pl.set_mask(Y.make_holes(n, n, holes, 3.0, 12.0, Y.seed(1, 1, 1)))
    pl.set_state(*Y.init_random(n, n, Y.seed(1, 1, 2), gate=1.0, nvars=3))
    stream = torch.cuda.current_stream()
    out = {"n": n, "holes": holes, "state_bytes": n * n * 25}
    # the headline configuration, 8 rows x 8 waves (two workgroups per CU,
    # which pipeline better once every CU has many tiles) and the one-step kernel
    for name, c in (("tiled", cfg), ("tiled_8x8", (T.YH_KERNEL_TILED, 8, 8, 8, 1)),
                    ("one_step", (T.YH_KERNEL_DIRECT, 1, 0, 0, 1))):
        pl.set_kernel(*c)
        t_ms, K, alg, ach = kernel_timing(a, pl, c, stream, n * n, 25, launches=50)
        out[name] = {"K": K, "rows": c[2], "waves": c[3], "cols": c[4],
                     "launch_us": round(t_ms * 1e3, 2), "achieved_GBps": round(ach, 1),
                     "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "gcell_steps_per_s": round(n * n * K / (t_ms / 1e3) / 1e9, 2)}
        log(f"control {n}^2 {name}: {t_ms * 1e3:.1f} us/launch K={K}, {ach:.0f} GB/s")
    pl.close()
    if not a.no_traffic:
        # the state (n^2 x 25 B) exceeds the Infinity Cache: these counters
        # see DRAM traffic, not cache hits
        for name, c in (("tiled", cfg), ("tiled_8x8", (T.YH_KERNEL_TILED, 8, 8, 8, 1))):
            tr = hbm_traffic(a, _kernel_only_args(a, c, nx=n, holes=holes, launches=20))
            if tr:
                alg = n * n * 25
                out[name]["traffic"] = round(tr[0])
                out[name]["traffic_over_alg"] = round(tr[0] / alg, 4)
                out[name]["traffic_counters"] = tr[1]
                log(f"control {n}^2 {name}: HBM traffic {tr[0] / 1e6:.1f} MB per launch "
                    f"({tr[0] / alg:.3f}x the algorithmic {alg / 1e6:.1f} MB)")
    return out


def batch_kernel_only(a):
    """--batch-kernel-only: the batched workload on a fixed kernel, a 256 MiB
    device copy (the byte counters' calibration), then --kernel-launches
    direct launches of K steps each (no graph), for a profiler to count."""
    import torch
    from yolohtli_amd.batch import Batch, batched_workload
    params, masks, fields = batched_workload(ntissues=a.batch_tissues)
    b = Batch(params)
    for t in range(len(params)):
        b.set_mask(t, masks[t])
        b.set_state(t, *fields[t])
    K, r, w, c, pers = (int(x) for x in a.batch_cfg.split(","))
    b.set_kernel(K, r, w, c, pers)
    b.set_graph(False)
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    stream = torch.cuda.current_stream()
    b.step(K * a.kernel_launches, stream)
    torch.cuda.synchronize()
    b.close()


def batched(a):
    """Batched configuration: --batch-tissues tissues of uneven size and
    parameters, --batch-steps steps in one captured graph of work-list
    launches, against per-tissue launches of the same kernel and k."""
    import torch
    from yolohtli_amd.batch import Batch, batched_workload
    params, masks, fields = batched_workload(ntissues=a.batch_tissues)
    b = Batch(params)
    for t in range(len(params)):
        b.set_mask(t, masks[t])
        b.set_state(t, *fields[t])
    del masks, fields
    _, cells = b.size()
    stream = torch.cuda.current_stream()
    cfg, ms = b.autotune(steps=32, stream=stream)
    cfg, pers = b.get_kernel()
    log(f"batch autotune: K={cfg[1]} rows={cfg[2]} waves={cfg[3]} cols={cfg[4]} "
        f"persistent={pers}: {ms:.3f} ms/step")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    b.set_graph(True)
    b.step(a.batch_steps, stream)  # captures the graph, warms up
    torch.cuda.synchronize()
    e0.record(stream)
    b.step(a.batch_steps, stream)  # one graph launch
    e1.record(stream)
    torch.cuda.synchronize()
    ms_b = e0.elapsed_time(e1)
    rate_b = cells * a.batch_steps / (ms_b / 1e3) / 1e9
    log(f"batched: {a.batch_steps} steps of {cells} cells in {ms_b:.1f} ms = {rate_b:.1f} "
        f"Gcell-steps/s")
    n_each = a.batch_each_steps
    b.step_each(cfg[1], stream)
    torch.cuda.synchronize()
    e0.record(stream)
    b.step_each(n_each, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_e = e0.elapsed_time(e1)
    rate_e = cells * n_each / (ms_e / 1e3) / 1e9
    log(f"per-tissue launches: {n_each} steps in {ms_e:.1f} ms = {rate_e:.1f} Gcell-steps/s")
    b.close()
    out = {"tissues": len(params), "cells": int(cells), "steps": a.batch_steps,
           "gcell_steps_per_s": round(rate_b, 3), "ms": round(ms_b, 3),
           "per_tissue_launches_gcell_steps_per_s": round(rate_e, 3),
           "per_tissue_steps": n_each, "speedup_vs_per_tissue": round(rate_b / rate_e, 3),
           "kernel": {"K": cfg[1], "rows": cfg[2], "waves": cfg[3], "cols": cfg[4],
                      "persistent": pers, "graph": True}}
    # per launch: every cell's three state variables read and written once
    # plus its mask code (25 B), K steps
    alg = int(cells) * 25
    out["launch_us"] = round(1e3 * ms_b / (a.batch_steps / cfg[1]), 2)
    out["alg_bytes_per_launch"] = alg
    out["achieved_GBps"] = round(alg / (out["launch_us"] / 1e6) / 1e9, 1)
    out["frac"] = round(out["achieved_GBps"] / HBM_PEAK_GBPS, 4)
    if not a.no_traffic:
        args = ["--batch-kernel-only", "--batch-tissues", str(a.batch_tissues),
                "--batch-cfg", ",".join(str(x) for x in (cfg[1], cfg[2], cfg[3], cfg[4], pers)),
                "--kernel-launches", "10"]
        tr = hbm_traffic(a, args, kernel="k_batch")
        if tr:
            out["traffic"] = round(tr[0])
            out["traffic_over_alg"] = round(tr[0] / alg, 4)
            out["traffic_counters"] = tr[1]
            log(f"batched: HBM traffic {tr[0] / 1e6:.1f} MB per launch ({tr[0] / alg:.3f}x "
                f"the algorithmic {alg / 1e6:.1f} MB)")
    return out


def cross_field_512(a):
    """The 512^2 plain configuration: FK, no holes, the cross-field start
    (u = 1 for i < 32, v = w = 0 for j >= ny/2), --c0-steps steps on the
    autotuned kernel, the dominant kernel's launch time by HIP events, and
    the CPU baseline of the same workload."""
    import torch
    import yolohtli_amd as Y
    from yolohtli_amd import abi_types as T
    from yolohtli_amd import autotune
    n = 512
    p = T.grid_params(n, n, model=T.YH_MODEL_3V_FK)
    f = Y.initGates(n, n, u_cols=32, rest=1.0, refr=0.0, nvars=3)
    pl = Y.Plan(p)
    pl.set_state(*f)
    best = autotune.tune_plan(pl, f, steps=400, log=log)
    cfg = best["cfg"]
    pl.set_kernel(*cfg)
    pl.set_state(*f)
    stream = torch.cuda.current_stream()
    pl.step(100, stream)
    pl.set_state(*f)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    pl.step(a.c0_steps, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    rate = n * n * a.c0_steps / (ms / 1e3) / 1e9
    t_ms, K, alg, ach = kernel_timing(a, pl, cfg, stream, n * n, 24)
    pl.close()
    log(f"512^2 cross-field: {a.c0_steps} steps in {ms:.2f} ms = {rate:.1f} Gcell-steps/s; "
        f"kernel {t_ms * 1e3:.2f} us per launch of K={K}")
    out = {"workload": "plain stepper 512x512, FK 3V, cross-field start, no holes",
           "steps": a.c0_steps, "ms": round(ms, 3), "gcell_steps_per_s": round(rate, 3),
           "kernel": {"K": cfg[1], "rows": cfg[2], "waves": cfg[3], "cols": cfg[4]},
           "launch_us": round(t_ms * 1e3, 3), "alg_bytes_per_launch": alg,
           "achieved_GBps": round(ach, 1), "frac": round(ach / HBM_PEAK_GBPS, 4)}
    if not a.no_cpu:
        cpu = cpu_baseline(p, f, None, a.cpu_seconds / 2)
        out["cpu_baseline"] = cpu
        log(f"512^2 cross-field CPU baseline {cpu['value']:.3f} Gcell-steps/s")
    return out


def adjoint(a):
    """Gradient configuration: 256^2 FK with tanh-smoothed Heaviside (width
    0.01), 2000 steps, 64 sensors recorded every step, checkpoint every 100;
    targets from the hidden D field; gradient w.r.t. D (256^2) and 6 taus at
    the uniform D0.  Reports the time per gradient and the relative error of
    the tau-gradient against FP64 central differences of the oracle's loss."""
    import torch
    import oracle.oracle as O
    from yolohtli_amd.adjoint import SensorProblem, TAU_NAMES, adjoint_workload
    from yolohtli_amd import abi_types as T
    n, nsteps = a.adj_n, a.adj_steps
    p, x0, sensors, Dh, D0 = adjoint_workload(n=n, nsteps=nsteps)
    pr = SensorProblem(p, nsteps, sensors, x0, rec_every=1, ckpt_every=100)
    _, targets = pr.forward(torch.from_numpy(Dh.reshape(-1)).cuda())
    pr.set_targets(targets)
    D = torch.full((n * n,), D0, dtype=torch.float32, device="cuda")
    L, _ = pr.forward(D)
    pr.backward()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, _ = pr.forward(D)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gD, gt = pr.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"adjoint: loss {L:.6e}; forward {1e3 * (t1 - t0):.1f} ms, backward "
        f"{1e3 * (t2 - t1):.1f} ms")
    out = {"n": n, "steps": nsteps, "sensors": int(len(sensors)), "checkpoint_every": 100,
           "loss": L, "forward_ms": round(1e3 * (t1 - t0), 2),
           "backward_ms": round(1e3 * (t2 - t1), 2),
           "ms_per_gradient": round(1e3 * (t2 - t0), 2),
           "grad_tau": [float(x) for x in gt], "grad_D_norm": float(gD.norm().item())}
    if a.adj_fd:
        # the FP64 problem takes its targets from its own run with the hidden
        # D (mixing the GPU's FP32 targets into an FP64 loss shifts the small
        # tau gradients, whose terms cancel, by ~1e-3 relative)
        X0 = [f.astype(np.float64) for f in x0]
        Dd = np.full(n * n, D0, np.float64)
        _, tg = O.fk_loss(p, Dh.reshape(-1).astype(np.float64), X0, nsteps, 1, sensors,
                          traces=True)
        out["target_max_abs_diff"] = float(np.max(np.abs(targets.astype(np.float64) - tg)))
        errs = []
        for k in a.adj_fd_taus:
            name = TAU_NAMES[k]
            tk = getattr(p.fk, name)
            h = 1e-4 * tk
            lp = O.fk_loss(T.copy_params(p, fk={name: tk + h}), Dd, X0, nsteps, 1, sensors, tg)
            lm = O.fk_loss(T.copy_params(p, fk={name: tk - h}), Dd, X0, nsteps, 1, sensors, tg)
            fd = (lp - lm) / (2 * h)
            errs.append(abs(gt[k] - fd) / max(abs(fd), 1e-300))
            log(f"adjoint FD check {name}: adjoint {gt[k]:.6e} FP64 central FD {fd:.6e}")
        # D along a smooth bump in the band the wave crosses
        jj, ii = np.mgrid[0:n, 0:n].astype(np.float64)
        bump = np.exp(-((ii - 0.15 * n) ** 2 + (jj - 0.5 * n) ** 2)
                      / (2 * (0.05 * n) ** 2)).reshape(-1)
        h = 1e-3 * D0
        fd = (O.fk_loss(p, Dd + h * bump, X0, nsteps, 1, sensors, tg)
              - O.fk_loss(p, Dd - h * bump, X0, nsteps, 1, sensors, tg)) / (2 * h)
        gb = float(gD.double().cpu().numpy() @ bump)
        log(f"adjoint FD check D.bump: adjoint {gb:.6e} FP64 central FD {fd:.6e}")
        out["fd_check"] = {"taus": [TAU_NAMES[k] for k in a.adj_fd_taus],
                           "rel_err": [float(e) for e in errs],
                           "max_rel_err": float(max(errs)), "h_rel": 1e-4,
                           "D_bump_rel_err": abs(gb - fd) / abs(fd),
                           "reference": "FP64 oracle loss, central differences"}
    pr.close()
    return out


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multi-GPU (domain-decomposed) stepper is not "
              "implemented; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import yolohtli_amd as Y
    from yolohtli_amd import abi_types as T
    from yolohtli_amd import autotune

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.batch_kernel_only:
        batch_kernel_only(a)
        return
    nx = ny = a.nx
    model = T.YH_MODEL_3V_FK if a.model == "fk" else T.YH_MODEL_2V_REF
    masked = a.holes > 0
    p = T.grid_params(nx, ny, model=model, solidSwitch=int(masked))
    nv = T.nvars(p)
    mask = Y.make_holes(nx, ny, a.holes, 3.0, 12.0, Y.seed(1, 0, 1)) if masked else None
    fields = Y.init_random(nx, ny, Y.seed(1, 0, 2), gate=1.0, nvars=nv)
    n = nx * ny
    bytes_per_cell = 2 * nv * 4 + (1 if masked else 0)

    stream = torch.cuda.current_stream()
    pl = Y.Plan(p)
    if masked:
        pl.set_mask(mask)
    pl.set_state(*fields)

    # ---- kernel choice (untimed): the project's autotuner, or a fixed config
    if a.cfg == "auto":
        # tuned for the call size of the timed loop below (tip_every-step
        # blocks, or one call when fewer steps are timed)
        call = min(a.steps, a.tip_every) if a.tip_every else a.steps
        best = autotune.tune_plan(pl, fields, steps=max(1, call), log=log)
        cfg = best["cfg"]
    else:
        K, r, w, c = (int(x) for x in a.cfg.split(","))
        cfg = (T.YH_KERNEL_TILED, K, r, w, c)
    pl.set_kernel(*cfg)
    pl.set_state(*fields)
    log(f"kernel cfg variant={cfg[0]} K={cfg[1]} rows={cfg[2]} waves={cfg[3]} cols={cfg[4]}")
    if a.kernel_only:
        # calibration for the HBM byte counters: a plain copy of known size
        # (64 Mi floats = 256 MiB read + 256 MiB written per launch)
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        t_ms, K, alg_bytes, achieved = kernel_timing(a, pl, cfg, stream, n, bytes_per_cell)
        print(json.dumps({"kernel_launch_us": t_ms * 1e3, "K": K,
                          "alg_bytes_per_launch": alg_bytes, "achieved_GBps": achieved,
                          "gcell_steps_per_s_in_launch": n * K / (t_ms / 1e3) / 1e9}),
              flush=True)
        return

    def run(nsteps):
        done = 0
        while done < nsteps:
            m = min(a.tip_every, nsteps - done)
            pl.step(m, stream)
            done += m
            if a.tip_every and pl.count % a.tip_every == 0:
                pl.track_tips(stream)

    run(a.warmup)
    torch.cuda.synchronize()
    pl.clear_tips()
    log(f"warmup {a.warmup} steps done")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    run(a.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    log(f"timed {a.steps} steps: {ms:.2f} ms (wall {wall * 1e3:.2f} ms)")
    if pl.check_finite() != 0:
        log("state is not finite")
        sys.exit(1)
    tips = pl.read_tips()
    value = n * a.steps / (ms / 1e3) / 1e9

    t_launch_ms, K, alg_bytes, achieved = kernel_timing(a, pl, cfg, stream, n, bytes_per_cell)
    log(f"kernel: {t_launch_ms * 1e3:.2f} us per launch of K={K} "
        f"({n * K / (t_launch_ms / 1e3) / 1e9:.1f} Gcell-steps/s inside one launch)")

    pl.close()
    traffic = valu = None
    if not a.no_traffic and cfg[0] == T.YH_KERNEL_TILED:
        log("HBM traffic: FETCH_SIZE and WRITE_SIZE passes under rocprofv3 ...")
        traffic = hbm_traffic(a, _kernel_only_args(a, cfg))
        if traffic:
            log(f"HBM traffic {traffic[0] / 1e6:.2f} MB per launch "
                f"({traffic[0] / alg_bytes:.3f}x the algorithmic {alg_bytes / 1e6:.2f} MB)")
        else:
            log("HBM traffic: counter passes unavailable")
        valu = valu_issue(a, cfg, t_launch_ms)
        if valu:
            log(f"VALU: {valu['insts_per_launch'] / 1e6:.3f} M wave instructions per launch = "
                f"{valu['issue_us']:.2f} us of issue at {VALU_CLOCK_GHZ} GHz "
                f"({valu['frac']:.3f} of the launch)")
    copy_gbps = copy_bandwidth(stream)
    log(f"device copy of 256 MiB: {copy_gbps:.0f} GB/s ({copy_gbps / HBM_PEAK_GBPS:.3f} of peak)")
    control = None
    if not a.no_control and cfg[0] == T.YH_KERNEL_TILED and a.model == "fk":
        log("non-cache-resident control grid ...")
        control = control_grid(a, cfg)
    c0 = None
    if a.c0_steps > 0:
        log("512^2 cross-field configuration ...")
        c0 = cross_field_512(a)
    sr = None
    if not a.no_sr:
        log("symmetry-reduction configuration ...")
        sr = symmetry_reduction(a)
    batch = None
    if not a.no_batch:
        log("batched configuration ...")
        batch = batched(a)
    adj = None
    if not a.no_adjoint:
        log("adjoint configuration ...")
        adj = adjoint(a)

    cpu = None
    if not a.no_cpu:
        log("CPU baseline ...")
        cpu = cpu_baseline(p, fields, mask, a.cpu_seconds)
        log(f"CPU baseline {cpu['value']:.3f} Gcell-steps/s on {cpu['cores']} threads")

    kname = "k_tiled" if cfg[0] == T.YH_KERNEL_TILED else "k_direct_codes"
    out = {
        "metric": "Gcell-steps/s + % HBM roofline at 1024^2; batched and adjoint vs plain",
        "value": round(value, 3),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded splitmix64 holes and random IC)",
        "config": {
            "workload": f"plain stepper {nx}x{ny}, {'FK 3V' if nv == 3 else '2V reference'} "
                        f"model, {a.holes} holes, {a.steps} steps, tips every {a.tip_every}",
            "nx": nx, "ny": ny, "holes": a.holes, "dt_ms": 0.02, "dx_cm": 0.025,
            "D_cm2_per_ms": 0.001, "kernel": {"name": kname, "K": cfg[1], "rows": cfg[2],
                                              "waves": cfg[3], "cols": cfg[4]},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": round(traffic[0]) if traffic else None,
            "traffic_counters": traffic[1] if traffic else None,
            "kernel": kname,
            "launch_us": round(t_launch_ms * 1e3, 3),
            "alg_bytes_per_launch": alg_bytes,
            "steps_per_launch": K,
            "equiv_GBps_per_step": round(bytes_per_cell * n * K / (t_launch_ms / 1e3) / 1e9, 1),
            "copy_GBps": round(copy_gbps, 1),
            "valu_issue": valu,
        },
        "cpu_baseline": cpu,
        "batched": batch,
        "control_grid": control,
        "cross_field_512": c0,
        "symmetry_reduction": sr,
        "adjoint": adj,
        "tips_logged": int(len(tips)),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
