"""Batched tissues: one work-list launch per k steps must give, tissue by
tissue, exactly what the single-tissue kernel gives with the same tile shape
and k (same arithmetic, so bit-for-bit), whether it runs persistent or one
workgroup per item, eagerly or from a captured graph; and it must agree with
the oracle like the plain stepper does."""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T
from yolohtli_amd.batch import Batch, batched_workload

pytestmark = pytest.mark.gpu


def _small_workload(model=T.YH_MODEL_3V_FK):
    return batched_workload(ntissues=7, config=99, model=model, sides=(64, 96, 130, 160),
                            max_holes=6)


def _batch(params, masks, fields, **cfg):
    b = Batch(params)
    for t, (m, f) in enumerate(zip(masks, fields)):
        b.set_mask(t, m)
        b.set_state(t, *f)
    if cfg:
        b.set_kernel(**cfg)
    return b


@pytest.mark.parametrize("aniso", [False, True])
@pytest.mark.parametrize("shape", [(6, 4, 16, 1), (8, 7, 16, 1), (6, 8, 8, 1)])
@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_batch_equals_single_tissue_kernel(torch_cuda, model, shape, aniso):
    """aniso: one tissue with hy = 1.3 hx, so the batch (8 x 8 has an
    all-isotropic variant) takes each tissue's weights at run time, while the
    single-tissue launches of the other tissues keep the isotropic variant."""
    params, masks, fields = _small_workload(model)
    if aniso:
        params[1].Ly = 1.3 * params[1].Ly
    K, r, w, c = shape
    cfg = dict(substeps=K, rows=r, waves=w, cols=c)
    b = _batch(params, masks, fields, persistent=1, **cfg)
    b.set_graph(False)
    b.step(27)
    for t, p in enumerate(params):
        pl = Y.Plan(p)
        pl.set_mask(masks[t])
        pl.set_state(*fields[t])
        pl.set_option(T.YH_OPT_KEEP_PREV, 0)
        pl.set_kernel(T.YH_KERNEL_TILED, K, r, w, c)
        pl.step(27)
        want = pl.get_state()
        got = b.get_state(t)
        for a in range(len(got)):
            np.testing.assert_array_equal(got[a], want[a])
        pl.close()


def test_batch_modes_agree_bitwise(torch_cuda):
    params, masks, fields = _small_workload()
    results = []
    for persistent, graph, each in ((1, False, False), (0, False, False), (1, True, False),
                                    (1, False, True)):
        b = _batch(params, masks, fields, substeps=8, rows=4, waves=8, cols=2,
                   persistent=persistent)
        b.set_graph(graph)
        for n in (16, 16, 5):  # a graph replayed twice, and a shorter tail
            if each:
                b.step_each(n)
            else:
                b.step(n)
        results.append([b.get_state(t) for t in range(len(params))])
        b.close()
    for r in results[1:]:
        for t in range(len(params)):
            for a in range(3):
                np.testing.assert_array_equal(r[t][a], results[0][t][a])


def test_batch_against_oracle(torch_cuda):
    params, masks, fields = _small_workload()
    b = _batch(params, masks, fields, substeps=8, rows=4, waves=16, cols=1, persistent=1)
    b.step(200)
    for t in (0, 3, 6):
        got = b.get_state(t)
        want = O.run(params[t], [x.astype(np.float64) for x in fields[t]], 200, solid=masks[t])
        for a in range(3):
            d = np.abs(got[a] - want[a])
            assert np.mean(d > 1e-4) < 1e-3, (t, a, float(np.mean(d > 1e-4)))
            assert np.median(d) < 1e-6


def test_batch_autotune_and_errors(torch_cuda):
    params, masks, fields = _small_workload()
    b = _batch(params, masks, fields)
    before = [b.get_state(t) for t in range(len(params))]
    cfg, ms = b.autotune(steps=16)
    assert cfg[0] == T.YH_KERNEL_TILED and ms > 0
    after = [b.get_state(t) for t in range(len(params))]  # autotuning restores the state
    for t in range(len(params)):
        np.testing.assert_array_equal(after[t][0], before[t][0])
    with pytest.raises(RuntimeError):
        b.set_kernel(substeps=40, rows=4, waves=16, cols=1)
    b2 = Batch(params)  # masks not set
    with pytest.raises(RuntimeError):
        b2.step(1)
    with pytest.raises(RuntimeError):
        Batch([params[0], T.copy_params(params[1], model=T.YH_MODEL_2V_REF)])


def test_default_depth_is_capped_by_the_smallest_tissue(torch_cuda):
    """With tissues of side 5 and 6 in the batch the default K (6) is capped
    at 4, the most substeps a reflected halo is exact for, and every tissue
    matches the one-step kernel."""
    sides = (5, 6, 64)
    params = [T.grid_params(n, n, model=T.YH_MODEL_3V_FK) for n in sides]
    b = Batch(params)
    fields = [Y.init_random(n, n, Y.seed(9, t, 2), gate=1.0, nvars=3) for t, n in enumerate(sides)]
    for t in range(len(sides)):
        b.set_state(t, *fields[t])
    cfg, _ = b.get_kernel()
    assert cfg[1] == 4
    b.set_graph(False)
    b.step(20)
    for t, n in enumerate(sides):
        pl = Y.Plan(params[t])
        pl.set_state(*fields[t])
        pl.set_kernel(T.YH_KERNEL_DIRECT, 1, 0, 0, 1)
        pl.step(20)
        for x, y in zip(b.get_state(t), pl.get_state()):
            assert np.max(np.abs(x - y)) < 2e-5
        pl.close()
    b.close()
