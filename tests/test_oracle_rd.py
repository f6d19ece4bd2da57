"""Pins the oracle's reaction-diffusion step (oracle/oracle_rd.inc) without the GPU.

Checks are chosen so that a plausible mistake fails one of them:
  * mirror index against its definition (a wrong reflection or off-by-one);
  * an independent vectorised formulation (np.pad 'reflect' + masked weights
    written as neighbour substitution, not as the reference's coefficient
    table) on random masks covering every coefficient case;
  * closed forms: zero state is a fixed point, a uniform field follows the
    single-cell ODE, masked pure diffusion keeps a constant field exactly,
    unmasked pure diffusion conserves the trapezoid-weighted sum, mirror
    symmetry is preserved;
  * FP32 twin within FP32 rounding of FP64.
"""
import numpy as np
import pytest

import oracle.oracle as O
from yolohtli_amd import abi_types as T


def _p(nx, ny, **kw):
    return T.grid_params(nx, ny, **kw)


def test_coord_mirror_definition():
    for n in (4, 5, 17):
        for k in range(-n + 1, 2 * n - 1):
            want = -k if k < 0 else (2 * (n - 1) - k if k >= n else k)
            assert O.coord(k, n) == want


def test_derive_matches_baseline_numbers():
    p = _p(512, 512)
    c = O.derive(p)
    assert c.hx == pytest.approx(0.025, rel=1e-15)
    assert c.rx == pytest.approx(0.02 * 0.001 / 0.025 ** 2, rel=1e-14)  # 0.032
    q = T.copy_params(p, reduceSym=1)
    assert O.derive(q).dt == 0.01  # main.cu:148
    q = T.copy_params(p, degrad=90.0, diff_par=0.002)
    c = O.derive(q)
    assert c.Dyy == pytest.approx(0.002, rel=1e-9) and c.Dxx == pytest.approx(0.001, rel=1e-9)


def _indep_step(p, fields, solid):
    """Independent numpy formulation of one Euler step (2V or FK)."""
    nx, ny = p.nx, p.ny
    c = O.derive(p)
    u = fields[0].reshape(ny, nx)
    v = fields[1].reshape(ny, nx)
    s = np.ones((ny, nx), bool) if solid is None else solid.reshape(ny, nx).astype(bool)

    def lap(f):
        fp = np.pad(f, 1, mode="reflect")
        sp = np.pad(s, 1, mode="reflect")
        C = fp[1:-1, 1:-1]
        out = np.zeros_like(f)
        for axis, r in ((1, c.rx), (0, c.ry)):
            if axis == 1:
                m1, p1 = fp[1:-1, :-2], fp[1:-1, 2:]
                s1, s2 = sp[1:-1, :-2], sp[1:-1, 2:]
            else:
                m1, p1 = fp[:-2, 1:-1], fp[2:, 1:-1]
                s1, s2 = sp[:-2, 1:-1], sp[2:, 1:-1]
            # a hole neighbour is replaced by the opposite neighbour (mirror at
            # the hole wall); with both neighbours holes the axis term vanishes
            a = np.where(s1, m1, np.where(s2, p1, C))
            b = np.where(s2, p1, np.where(s1, m1, C))
            out = out + r * ((a - C) + (b - C))
        return np.where(s, out, 0.0)

    lu = lap(u)
    if p.model == T.YH_MODEL_2V_REF:
        lv = lap(v) * p.rscale if p.gateDiff else 0.0
        Is = -(p.mu * u * (1 - u) * (u - p.alpha) - u * v)
        Iv = -(p.eps * (p.delta * (u - p.gamma) * (p.beta - u) - v - p.theta))
        un = u + p.tc * (lu - c.dt * Is)
        vn = v + p.tc * (lv - c.dt * Iv)
        res = [un, vn]
    else:
        w = fields[2].reshape(ny, nx)
        f = p.fk
        pp = (u >= f.u_c).astype(float)
        qq = (u >= f.u_v).astype(float)
        I = (-v * pp * (1 - u) * (u - f.u_c) / f.tau_d + u * (1 - pp) / f.tau_0 + pp / f.tau_r
             - w * (1 + np.tanh(f.k * (u - f.u_csi))) / (2 * f.tau_si))
        tvm = (1 - qq) * f.tau_v1m + qq * f.tau_v2m
        Gv = -((1 - pp) * (1 - v) / tvm - pp * v / f.tau_vp)
        Gw = -((1 - pp) * (1 - w) / f.tau_wm - pp * w / f.tau_wp)
        res = [u + p.tc * (lu - c.dt * I), v - p.tc * c.dt * Gv, w - p.tc * c.dt * Gw]
    return [np.where(s, r, 0.0).reshape(-1) for r in res]


def _random_mask(rng, nx, ny, frac=0.3):
    m = (rng.random(nx * ny) > frac).astype(np.uint8)
    return m


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("shape", [(8, 8), (13, 9), (64, 48)])
def test_step_matches_independent_formulation(model, masked, shape):
    rng = np.random.default_rng(11 + shape[0] + 7 * masked + 3 * model)
    nx, ny = shape
    p = _p(nx, ny, model=model, solidSwitch=int(masked))
    nv = T.nvars(p)
    fields = [rng.random(nx * ny) for _ in range(nv)]
    solid = _random_mask(rng, nx, ny) if masked else None
    got, _ = O.rd_step(p, fields, solid)
    want = _indep_step(p, fields, solid)
    for a in range(nv):
        np.testing.assert_allclose(got[a], want[a], rtol=0, atol=2e-15)


def test_random_masks_cover_every_coefficient_case():
    rng = np.random.default_rng(5)
    nx = ny = 32
    m = _random_mask(rng, nx, ny).reshape(ny, nx).astype(bool)
    mp = np.pad(m, 1, mode="reflect")
    sw, se = mp[1:-1, :-2], mp[1:-1, 2:]
    cases = set()
    for a, b, c in zip(sw[m], se[m], m[m]):
        cases.add((bool(a), bool(b)))
    assert cases == {(True, True), (True, False), (False, True), (False, False)}


def test_zero_state_is_fixed_point():
    for model in (T.YH_MODEL_2V_REF,):
        p = _p(16, 16, model=model)
        f = [np.zeros(256), np.zeros(256)]
        out, _ = O.rd_step(p, f)
        assert np.all(out[0] == 0) and np.all(out[1] == 0)


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_uniform_field_follows_single_cell_ode(model):
    p = _p(8, 8, model=model)
    c = O.derive(p)
    if model == T.YH_MODEL_2V_REF:
        x = [0.3, 0.05]
    else:
        x = [0.2, 0.9, 0.95]
    nv = len(x)
    fields = [np.full(64, val) for val in x]
    out = O.run(p, fields, 200, nthreads=1)
    st = list(x)
    for _ in range(200):
        u, v = st[0], st[1]
        if model == T.YH_MODEL_2V_REF:
            Is = -(p.mu * u * (1.0 - u) * (u - p.alpha) - u * v)
            Iv = -(p.eps * (p.delta * (u - p.gamma) * (p.beta - u) - v - p.theta))
            st = [u + p.tc * (0.0 - c.dt * Is), v + p.tc * (0.0 - c.dt * Iv)]
        else:
            w = st[2]
            f = p.fk
            pp = 1.0 if u >= f.u_c else 0.0
            qq = 1.0 if u >= f.u_v else 0.0
            I = (-v * pp * (1 - u) * (u - f.u_c) / f.tau_d + u * (1 - pp) / f.tau_0
                 + pp / f.tau_r - w * (1 + np.tanh(f.k * (u - f.u_csi))) / (2 * f.tau_si))
            tvm = (1 - qq) * f.tau_v1m + qq * f.tau_v2m
            Gv = -((1 - pp) * (1 - v) / tvm - pp * v / f.tau_vp)
            Gw = -((1 - pp) * (1 - w) / f.tau_wm - pp * w / f.tau_wp)
            st = [u + p.tc * (0.0 - c.dt * I), v - p.tc * c.dt * Gv, w - p.tc * c.dt * Gw]
    for a in range(nv):
        np.testing.assert_allclose(out[a], st[a], rtol=1e-13, atol=1e-15)


def test_masked_pure_diffusion_keeps_constant_field():
    rng = np.random.default_rng(3)
    p = _p(40, 30, solidSwitch=1, mu=0.0, eps=0.0)
    solid = _random_mask(rng, 40, 30, 0.4)
    u = np.where(solid > 0, 0.625, 0.0)
    v = np.zeros(1200)
    out = O.run(p, [u, v], 50, solid=solid, nthreads=2)
    assert np.array_equal(out[0], u)


def test_unmasked_pure_diffusion_conserves_trapezoid_sum():
    rng = np.random.default_rng(4)
    nx, ny = 33, 21
    p = _p(nx, ny, mu=0.0, eps=0.0, gateDiff=0)
    u = rng.random(nx * ny)
    v = np.zeros(nx * ny)
    wx = np.ones(nx); wx[0] = wx[-1] = 0.5
    wy = np.ones(ny); wy[0] = wy[-1] = 0.5
    W = np.outer(wy, wx).reshape(-1)
    s0 = np.sum(W * u)
    out = O.run(p, [u, v], 300, nthreads=2)
    assert np.sum(W * out[0]) == pytest.approx(s0, rel=1e-12)
    # and diffusion actually acted
    assert np.max(np.abs(out[0] - u)) > 1e-2


def test_mirror_symmetry_preserved():
    rng = np.random.default_rng(9)
    nx, ny = 24, 24
    p = _p(nx, ny, solidSwitch=1)
    half = rng.random((ny, nx // 2))
    u = np.concatenate([half, half[:, ::-1]], axis=1)
    mh = (rng.random((ny, nx // 2)) > 0.2)
    m = np.concatenate([mh, mh[:, ::-1]], axis=1).astype(np.uint8)
    v = np.zeros_like(u)
    out = O.run(p, [u.reshape(-1), v.reshape(-1)], 40, solid=m.reshape(-1), nthreads=1)
    U = out[0].reshape(ny, nx)
    np.testing.assert_allclose(U, U[:, ::-1], rtol=0, atol=1e-14)


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_fp32_twin_close_to_fp64(model):
    nx = ny = 48
    p = _p(nx, ny, model=model)
    nv = T.nvars(p)
    rng = np.random.default_rng(2)
    fields = [rng.random(nx * ny) for _ in range(nv)]
    a = O.run(p, fields, 100, dtype=np.float64, nthreads=2)
    b = O.run(p, fields, 100, dtype=np.float32, nthreads=2)
    for q in range(nv):
        assert np.max(np.abs(a[q] - b[q])) < 1e-4


def test_velTan_is_rhs_over_dt():
    rng = np.random.default_rng(1)
    p = _p(16, 12, solidSwitch=1)
    solid = _random_mask(rng, 16, 12)
    f = [rng.random(192), rng.random(192)]
    out, vt = O.rd_step(p, f, solid, veltan=True)
    c = O.derive(p)
    s = solid > 0
    np.testing.assert_allclose(vt[0][s], (out[0][s] - f[0][s]) / c.dt, rtol=1e-9, atol=1e-9)
    assert np.all(vt[0][~s] == 0) and np.all(out[0][~s] == 0)


def test_openmp_run_equals_serial():
    rng = np.random.default_rng(8)
    p = _p(50, 40, solidSwitch=1, model=T.YH_MODEL_3V_FK)
    solid = _random_mask(rng, 50, 40)
    f = [rng.random(2000) for _ in range(3)]
    a = O.run(p, f, 30, solid=solid, nthreads=1)
    b = O.run(p, f, 30, solid=solid, nthreads=4)
    for q in range(3):
        assert np.array_equal(a[q], b[q])
