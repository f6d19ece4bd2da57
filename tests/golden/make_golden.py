"""Writes tests/golden/oracle_small.npz: seeded inputs and the FP64 oracle's
outputs for small cases of the stepper (both models, with and without a
mask), two other branches of the reference kernel and the tip test.

The reference holds no tests or fixtures, and nothing of it is run here:
these vectors come from the oracle (oracle/, itself pinned by the checks in
tests/test_oracle_*.py), and freeze its results so that a later change to
the oracle or the kernels shows up as a difference.  Run from the repository
root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.oracle as O  # noqa: E402
from yolohtli_amd import abi_types as T  # noqa: E402

NX, NY = 24, 17
ANISO = dict(anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0)
# name, grid_params overrides, masked, steps, path ("plain" or "general")
CASES = [
    ("2v", dict(model=T.YH_MODEL_2V_REF), False, 7, "plain"),
    ("2v_masked", dict(model=T.YH_MODEL_2V_REF, solidSwitch=1), True, 7, "plain"),
    ("fk", dict(model=T.YH_MODEL_3V_FK), False, 7, "plain"),
    ("fk_masked", dict(model=T.YH_MODEL_3V_FK, solidSwitch=1), True, 7, "plain"),
    ("2v_rk4_lap4_aniso", dict(model=T.YH_MODEL_2V_REF, timeIntOrder=4, lap4=1, **ANISO), False,
     4, "general"),
    ("fk_dirichlet_masked", dict(model=T.YH_MODEL_3V_FK, solidSwitch=1, neumannBC=0,
                                 boundaryVal=0.1), True, 4, "general"),
]


def case_inputs(k, over, masked):
    rng = np.random.default_rng(1000 + k)
    nv = 3 if over["model"] == T.YH_MODEL_3V_FK else 2
    u = rng.random(NX * NY).astype(np.float32)
    gates = [rng.uniform(0.3, 1.0, NX * NY).astype(np.float32) for _ in range(nv - 1)]
    mask = (rng.random(NX * NY) > 0.15).astype(np.uint8) if masked else None
    return [u] + gates, mask


def outputs(over, fields, mask, steps, path):
    p = T.grid_params(NX, NY, **over)
    X = [f.astype(np.float64) for f in fields]
    if path == "plain":
        return O.run(p, X, steps, solid=mask)
    return O.run_general(p, X, steps, solid=mask)


def tip_case():
    # two smooth fields whose u = Uth isolines cross in several cells
    j, i = np.mgrid[0:NY, 0:NX].astype(np.float64)
    # (stored in FP32, the precision the library takes them in)
    past = (0.7 + 0.4 * np.sin(0.45 * i + 0.2) * np.cos(0.5 * j - 0.3)).astype(np.float32)
    present = (0.7 + 0.4 * np.sin(0.4 * i - 0.1 + 0.3 * j) * np.cos(0.45 * j + 0.2)).astype(
        np.float32)
    p = T.grid_params(NX, NY)
    tips, keys = O.tips(p, past.reshape(-1), present.reshape(-1), 1.25)
    return past.reshape(-1), present.reshape(-1), tips, keys


def main():
    arrays = {"meta": np.array(json.dumps({"nx": NX, "ny": NY, "cases": [
        {"name": n, "over": o, "masked": m, "steps": s, "path": pth}
        for n, o, m, s, pth in CASES]}))}
    for k, (name, over, masked, steps, path) in enumerate(CASES):
        fields, mask = case_inputs(k, over, masked)
        out = outputs(over, fields, mask, steps, path)
        for a, f in enumerate(fields):
            arrays[f"{name}/in{a}"] = f
            arrays[f"{name}/out{a}"] = np.asarray(out[a], np.float64)
        if mask is not None:
            arrays[f"{name}/mask"] = mask
    past, present, tips, keys = tip_case()
    arrays["tips/past"] = past
    arrays["tips/present"] = present
    arrays["tips/tips"] = tips
    arrays["tips/keys"] = keys
    path = os.path.join(ROOT, "tests", "golden", "oracle_small.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path}: {len(arrays)} arrays, {len(tips)} tips")


if __name__ == "__main__":
    main()
