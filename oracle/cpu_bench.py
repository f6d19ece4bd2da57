"""CPU baseline: the oracle's restatement of the update as an OpenMP host loop.

TEST INFRASTRUCTURE ONLY: run by bench.py's CPU-baseline leg as a child
process, so that its OpenMP thread binding (OMP_PROC_BIND=close,
OMP_PLACES=cores) never touches the threads of the process that drives the
GPU.  It reads the workload (state and mask) from an .npz file, runs 100
warm-up steps, then times five windows of the same update with a monotonic
clock in FP32 and in FP64 (the reference's REAL = double) and prints one JSON
line with the medians.

    python -m oracle.cpu_bench workload.npz --seconds 20
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from yolohtli_amd import abi_types as T  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure(p, fields, mask, dtype, seconds, threads, windows=5, warm=100):
    n = p.nx * p.ny
    fs = O.run(p, fields, warm // 2, solid=mask, dtype=dtype, nthreads=threads)
    t0 = time.perf_counter()
    fs = O.run(p, fs, warm - warm // 2, solid=mask, dtype=dtype, nthreads=threads)
    per_step = (time.perf_counter() - t0) / (warm - warm // 2)
    steps = max(4, int(seconds / windows / max(per_step, 1e-9)))
    rates = []
    for _ in range(windows):
        t0 = time.perf_counter()
        fs = O.run(p, fs, steps, solid=mask, dtype=dtype, nthreads=threads)
        rates.append(n * steps / (time.perf_counter() - t0) / 1e9)
    return float(np.median(rates)), steps, rates


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("--seconds", type=float, default=20.0,
                    help="total timed seconds, split between FP32 and FP64")
    a = ap.parse_args()
    z = np.load(a.workload)  # written by bench.py in this run (no pickles)
    nx, ny, model = int(z["nx"]), int(z["ny"]), int(z["model"])
    mask = z["mask"] if z["mask"].size else None
    p = T.grid_params(nx, ny, model=model, solidSwitch=int(mask is not None))
    nv = T.nvars(p)
    fields = [z[k] for k in ("u", "v", "w")[:nv]]
    threads = len(os.sched_getaffinity(0))
    threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS", threads))))
    out = {"threads": threads, "cpu": cpu_model(),
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "omp_places": os.environ.get("OMP_PLACES")}
    masked = mask is not None
    for name, dt, word in (("fp32", np.float32, 4), ("fp64", np.float64, 8)):
        rate, steps, rates = measure(p, fields, mask, dt, a.seconds / 2, threads)
        bpc = 2 * nv * word + (1 if masked else 0)
        out[name] = {"gcell_steps_per_s": rate, "GBps": rate * bpc,
                     "bytes_per_cell_step": bpc, "steps_per_window": steps,
                     "windows": [round(r, 4) for r in rates]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
