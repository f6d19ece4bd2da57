"""The oracle's restatement of the reference kernel's other branches
(reactionDiffusion.cu:36-561), pinned by checks that do not restate it:
a constant-padded 5-point Laplacian for Dirichlet boundaries, closed forms of
separable polynomials for lap4 and the anisotropic term, rotation and
reflection symmetries for their boundary corrections, a brute-force disk for
the mouse stimulus, and convergence orders against a tight ODE solve for
RK2/RK4.  On the standard branch it equals the Euler oracle bit for bit."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

NX, NY = 14, 11


def _quiet(**over):
    """2V parameters with a zero ionic current when v = 0 (mu = eps = theta = 0)."""
    base = dict(mu=0.0, eps=0.0, theta=0.0)
    base.update(over)
    return T.grid_params(NX, NY, **base)


def _grid():
    j, i = np.mgrid[0:NY, 0:NX]
    return i.astype(np.float64), j.astype(np.float64)


def _step(p, u, v, w=None, solid=None, stim=None):
    f = [u.reshape(-1), v.reshape(-1)] + ([w.reshape(-1)] if w is not None else [])
    return [x.reshape(NY, NX) for x in O.rd_step_general(p, f, solid=solid, stim=stim)]


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
@pytest.mark.parametrize("masked", [0, 1])
def test_standard_branch_equals_euler_oracle(model, masked):
    p = T.grid_params(NX, NY, model=model, solidSwitch=masked)
    rng = np.random.default_rng(7)
    f = [rng.random(NX * NY) for _ in range(T.nvars(p))]
    m = Y.make_holes(NX, NY, 2, 1.5, 3.0, 5) if masked else None
    a, va = O.rd_step(p, f, solid=m, veltan=True)
    b, vb = O.rd_step_general(p, f, solid=m, veltan=True)
    for x, y in zip(a + va, b + vb):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("masked", [0, 1])
def test_dirichlet_is_the_constant_padded_laplacian(masked):
    bval = 0.37
    p = _quiet(neumannBC=0, boundaryVal=bval, solidSwitch=masked)
    d = O.derive(p)
    rng = np.random.default_rng(3)
    u = rng.random((NY, NX))
    v = np.zeros((NY, NX))
    m = np.ones((NY, NX), np.uint8)
    if masked:
        m[4:6, 5:8] = 0
    un, _ = _step(p, u, v, solid=m.reshape(-1) if masked else None)
    # ghost cells at the boundary value: outside the domain and in holes
    g = np.pad(np.where(m > 0, u, bval), 1, constant_values=bval)
    lap = ((g[1:-1, :-2] - 2 * u + g[1:-1, 2:]) * d.rx + (g[2:, 1:-1] - 2 * u + g[:-2, 1:-1]) * d.ry)
    want = np.where(m > 0, u + lap, 0.0)
    np.testing.assert_allclose(un, want, rtol=0, atol=1e-14)


def test_lap4_closed_forms_interior():
    p = _quiet(lap4=1)
    d = O.derive(p)
    dt = d.dt
    q = dt * d.Dyy / (d.hy ** 2 * 12) + dt * d.Dxx / (d.hx ** 2 * 12)
    i, j = _grid()
    inner = (slice(2, NY - 2), slice(2, NX - 2))
    # u = i^2 j^2: second differences 2 j^2 and 2 i^2, mixed fourth difference 4
    u = i ** 2 * j ** 2
    un, _ = _step(p, u, np.zeros_like(u))
    want = 2 * d.rx * j ** 2 + 2 * d.ry * i ** 2 + 4 * q
    np.testing.assert_allclose((un - u)[inner], want[inner], rtol=1e-12)
    # I_sum = u (mu = 0, v = 1): the compact correction subtracts dt/12 * second
    # difference of the current; u = i^2 has delta_x^2 = 2, delta_y^2 = 0
    u = i ** 2
    un, vn = _step(p, u, np.ones_like(u))
    want = 2 * d.rx - 2 * dt / 12.0 - dt * u
    np.testing.assert_allclose((un - u)[inner], want[inner], rtol=1e-12, atol=1e-14)


def test_anisotropic_cross_term_interior():
    par, per, deg = 0.0016, 0.0006, 30.0
    p = _quiet(anisotropy=1, diff_par=par, diff_per=per, degrad=deg)
    d = O.derive(p)
    th = deg * 3.14159265359 / 180.0
    Dxy = (par - per) * math.sin(th) * math.cos(th)
    i, j = _grid()
    u = i * j  # u_xy = 1 in index units; the 5-point part vanishes
    un, _ = _step(p, u, np.zeros_like(u))
    inner = (slice(1, NY - 1), slice(1, NX - 1))
    # rxy * (SW + NE - SE - NW) = rxy * 4 = 2 Dxy dt / (hx hy)
    want = 2 * Dxy * d.dt / (d.hx * d.hy)
    np.testing.assert_allclose((un - u)[inner], want, rtol=1e-12)


BRANCHES = [
    dict(lap4=1),
    dict(anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0),
    dict(lap4=1, anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0),
    dict(neumannBC=0, boundaryVal=0.2, anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0),
    dict(neumannBC=0, boundaryVal=0.2, anisotropy=1, diff_par=0.0016, diff_per=0.0006, degrad=30.0,
         solidSwitch=1),
]


@pytest.mark.parametrize("over", BRANCHES)
@pytest.mark.parametrize("order", [1, 4])
def test_half_turn_symmetry(over, order):
    """The scheme commutes with the rotation by 180 degrees (the cross term is
    even under (x, y) -> (-x, -y)): catches swapped edge or corner corrections."""
    p = T.grid_params(NX, NY, timeIntOrder=order, **over)
    rng = np.random.default_rng(11)
    u, v = rng.random((NY, NX)), 0.3 * rng.random((NY, NX))
    m = None
    if p.solidSwitch:
        mm = np.ones((NY, NX), np.uint8)
        mm[3:5, 4:6] = 0
        m = mm
    R = lambda a: a[::-1, ::-1].copy()
    a = _step(p, u, v, solid=None if m is None else m.reshape(-1))
    b = _step(p, R(u), R(v), solid=None if m is None else R(m).reshape(-1))
    for x, y in zip(a, b):
        np.testing.assert_allclose(R(x), y, rtol=0, atol=1e-15)


@pytest.mark.parametrize("over", BRANCHES)
def test_mirror_symmetry_flips_the_fibre_angle(over):
    """Reflecting x maps D_xy to -D_xy (fibre angle -> -angle)."""
    p = T.grid_params(NX, NY, **over)
    q = T.copy_params(p, degrad=-p.degrad)
    rng = np.random.default_rng(12)
    u, v = rng.random((NY, NX)), 0.3 * rng.random((NY, NX))
    m = None
    if p.solidSwitch:
        mm = np.ones((NY, NX), np.uint8)
        mm[3:5, 4:7] = 0
        m = mm
    M = lambda a: a[:, ::-1].copy()
    a = _step(p, u, v, solid=None if m is None else m.reshape(-1))
    b = _step(q, M(u), M(v), solid=None if m is None else M(m).reshape(-1))
    for x, y in zip(a, b):
        np.testing.assert_allclose(M(x), y, rtol=0, atol=1e-15)


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_mouse_stimulus_disk(model):
    p = T.grid_params(48, 40, model=model)
    d = O.derive(p)
    rng = np.random.default_rng(5)
    f = [rng.random(48 * 40) for _ in range(T.nvars(p))]
    x0, y0 = 17, 22
    a = O.rd_step_general(p, f)
    b = O.rd_step_general(p, f, stim=(x0, y0))
    du = (b[0] - a[0]).reshape(40, 48)
    j, i = np.mgrid[0:40, 0:48]
    disk = (i - x0) ** 2 + (j - y0) ** 2 < 400
    assert disk.sum() == sum(1 for a_ in range(-20, 21) for b_ in range(-20, 21)
                             if a_ * a_ + b_ * b_ < 400 and 0 <= x0 + a_ < 48 and 0 <= y0 + b_ < 40)
    np.testing.assert_allclose(du[disk], p.tc * d.dt * 24.7, rtol=1e-9)
    assert np.all(du[~disk] == 0)
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)  # gates see no stimulus


def _ode_2v(p):
    def f(t, y):
        u, v = y
        Iu = -(p.mu * u * (1 - u) * (u - p.alpha) - u * v)
        Iv = -(p.eps * (p.delta * (u - p.gamma) * (p.beta - u) - v - p.theta))
        return [-p.tc * Iu, -p.tc * Iv]
    return f


def _ode_fk(p):
    fk = p.fk

    def f(t, y):
        u, v, w = y
        pp = 1.0 if u >= fk.u_c else 0.0
        q = 1.0 if u >= fk.u_v else 0.0
        J_fi = -v * pp * (1 - u) * (u - fk.u_c) / fk.tau_d
        J_so = u * (1 - pp) / fk.tau_0 + pp / fk.tau_r
        J_si = -w * (1 + math.tanh(fk.k * (u - fk.u_csi))) / (2 * fk.tau_si)
        tvm = (1 - q) * fk.tau_v1m + q * fk.tau_v2m
        Gv = -((1 - pp) * (1 - v) / tvm - pp * v / fk.tau_vp)
        Gw = -((1 - pp) * (1 - w) / fk.tau_wm - pp * w / fk.tau_wp)
        return [-(J_fi + J_so + J_si), -Gv, -Gw]
    return f


@pytest.mark.parametrize("model,y0,T_end", [(T.YH_MODEL_2V_REF, (0.55, 0.05), 2.0),
                                            (T.YH_MODEL_3V_FK, (0.9, 0.6, 0.7), 0.4)])
def test_time_integration_orders(model, y0, T_end):
    """A uniform field sees no diffusion: each scheme integrates the cell ODE
    with global order 1 (Euler), 2 (RK2) and 4 (RK4)."""
    nx = ny = 6
    base = T.grid_params(nx, ny, dx=0.1, model=model)
    ode = _ode_2v(base) if model == T.YH_MODEL_2V_REF else _ode_fk(base)
    ref = solve_ivp(ode, (0.0, T_end), list(y0), method="DOP853", rtol=1e-13, atol=1e-15).y[:, -1]
    for order, lo, hi in ((1, 1.6, 2.4), (2, 3.2, 4.8), (4, 11.0, 21.0)):
        errs = []
        for N in (8, 16, 32):
            p = T.copy_params(base, dt=T_end / N, timeIntOrder=order)
            X = [np.full(nx * ny, c) for c in y0]
            X = O.run_general(p, X, N)
            errs.append(max(abs(X[a][0] - ref[a]) for a in range(len(y0))))
        ratios = [errs[k] / errs[k + 1] for k in range(2)]
        print(model, order, errs, ratios)
        assert all(lo < r < hi for r in ratios), (order, errs, ratios)


# ----------------------------------------------------------------- contours
CX, CY = 40, 45


def test_contour_mode1_interpolates_a_linear_crossing():
    j, i = np.mgrid[0:CY, 0:CX].astype(np.float64)
    f2 = i - 10.3
    pts, keys = O.contour(CX, CY, 1, (0, 0, 0), np.zeros_like(f2), f2, t=2.5)
    assert len(pts) == CY  # one crossing per row, between i = 10 and 11
    np.testing.assert_allclose(pts[:, 0], 10.3, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(pts[:, 1], np.arange(CY))
    assert np.all(pts[:, 2] == 2.5)
    np.testing.assert_array_equal(keys, 2 * (10 + CX * np.arange(CY)))


def test_contour_mode1_keeps_the_reference_precedence():
    """zpmx < 0 || (zpmy < 0 && sc): stimArea masks only the y crossings."""
    j, i = np.mgrid[0:CY, 0:CX].astype(np.float64)
    off = np.zeros((CY, CX), np.uint8)
    px, _ = O.contour(CX, CY, 1, (0, 0, 0), np.zeros_like(i), i - 10.3, area=off)
    py, _ = O.contour(CX, CY, 1, (0, 0, 0), np.zeros_like(i), j - 10.3, area=off)
    assert len(px) == CY and len(py) == 0


def test_contour_mode2_single_threshold():
    rng = np.random.default_rng(8)
    j, i = np.mgrid[0:CY, 0:CX].astype(np.float64)
    f1 = rng.random((CY, CX))
    area = (rng.random((CY, CX)) > 0.2).astype(np.uint8)
    pts, keys = O.contour(CX, CY, 2, (0.5, 7.0, 0.0), f1, j + 0.5, area=area)
    want = [(x, 6) for x in range(CX) if f1[6, x] < 0.5 and area[6, x]]
    assert [tuple(p[:2]) for p in pts] == want


def test_contour_mode3_double_threshold():
    rng = np.random.default_rng(9)
    j, i = np.mgrid[0:CY, 0:CX].astype(np.float64)
    f1 = rng.random((CY, CX))
    f2 = 0.02 * j - 0.15
    # th2 = 0.3, th3 = 0.4: band |f2 - th2| < 0.5; crossings of f2 - 2*th2
    # (rows 37|38, sub 0) and of f2 - th2 + th3 (rows 2|3, sub 1)
    pts, keys = O.contour(CX, CY, 3, (0.5, 0.3, 0.4), f1, f2)
    want = sorted([2 * (x + CX * 37) for x in range(CX) if f1[37, x] < 0.5]
                  + [2 * (x + CX * 2) + 1 for x in range(CX) if f1[2, x] < 0.5])
    assert sorted(keys.tolist()) == want
