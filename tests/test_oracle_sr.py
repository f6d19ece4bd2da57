"""Pins the oracle's symmetry-reduction pieces, tips and APD by closed forms.

  * solve_matrix against numpy.linalg.solve on well-conditioned systems;
  * one-sided 2nd-order slices exact on quadratics; the template's 4th-order
    central difference exact (with the reference's sign) on quartics;
  * the 12 disk inner products of linear fields against lattice-point sums;
  * BFECC: identity at c = 0, constants preserved, second-order convergence of
    a translating Gaussian;
  * tips: bilinear fields with a known isoline intersection;
  * APD: linear up/down strokes with known threshold crossing times;
  * the SR step's composition (main.cu:894-953): each step's c is the 3x3
    solve of that step's own integrals with the phi from before the step,
    and phi then advances by dt times that c (main.cu:935-941).
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T


def _A(phi, I):
    cp, sp = np.cos(phi[2]), np.sin(phi[2])
    return np.array([[I[3 * r] * cp + I[3 * r + 1] * sp, I[3 * r + 1] * cp - I[3 * r] * sp,
                      I[3 * r + 2]] for r in range(3)])


@pytest.mark.parametrize("seed", range(6))
def test_solve_matrix_matches_linalg(seed):
    rng = np.random.default_rng(seed)
    I = rng.normal(size=12) + np.array([3, 0, 0, 0, 3, 0, 0, 0, 3, 0, 0, 0])
    phi = [0.0, 0.0, 0.0] if seed == 0 else rng.normal(size=3)
    c = O.solve_matrix(phi, I)
    want = np.linalg.solve(_A(phi, I), I[9:12])
    np.testing.assert_allclose(c, want, rtol=1e-10, atol=1e-12)


def test_solve_matrix_zero_integrals_is_not_finite():
    c = O.solve_matrix([0, 0, 0], np.zeros(12))
    assert not np.all(np.isfinite(c))


def _sr_params(n=64, r=20):
    return T.grid_params(n, n, reduceSym=1, tipOffsetX=r, tipOffsetY=r)


def test_slices_exact_on_polynomials():
    p = _sr_params()
    c = O.derive(p)
    n = p.nx
    jj, ii = np.mgrid[0:n, 0:n].astype(float)
    a, b, q = 0.3, -0.2, 0.15
    f = (a * ii ** 2 + b * ii * jj + q * jj ** 2).reshape(-1)
    for ac in ([0, 0, 0], [1.0, -0.5, 0.0]):  # backward and forward one-sided branches
        fx, fy, ft, fx0, fy0, ft0 = O.slice_field(p, f, 0, 0, ac, 0.0)
        inner = ((ii - n // 2) ** 2 + (jj - n // 2) ** 2 < 20 * 20).reshape(-1)
        dfx = ((2 * a * ii + b * jj) / c.hx).reshape(-1)
        dfy = ((b * ii + 2 * q * jj) / c.hy).reshape(-1)
        np.testing.assert_allclose(fx[inner], dfx[inner], rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(fy[inner], dfy[inner], rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(ft[inner], (c.hx * ii.reshape(-1) * dfy - c.hy * jj.reshape(-1) * dfx)[inner],
                                   rtol=1e-10, atol=1e-8)
        assert np.all(fx[~inner] == 0) and np.all(ft0[~inner] == 0)
    # 4th-order central (negated, symmetryReduction.cu:229) exact on quartics
    g = (0.01 * ii ** 4 - 0.02 * ii ** 3 * jj + 0.005 * jj ** 4).reshape(-1)
    _, _, _, gx0, gy0, _ = O.slice_field(p, g, 0, 0)
    dgx = ((0.04 * ii ** 3 - 0.06 * ii ** 2 * jj) / c.hx).reshape(-1)
    dgy = ((-0.02 * ii ** 3 + 0.02 * jj ** 3) / c.hy).reshape(-1)
    np.testing.assert_allclose(gx0[inner], -dgx[inner], rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(gy0[inner], -dgy[inner], rtol=1e-9, atol=1e-6)


def test_integrals_of_linear_field_are_lattice_sums():
    p = _sr_params(64, 18)
    c = O.derive(p)
    n = p.nx
    jj, ii = np.mgrid[0:n, 0:n].astype(float)
    u = (ii * c.hx).reshape(-1)  # fx = 1 exactly, template fx0 = -1, fy = fy0 = 0
    v = np.zeros(n * n)
    vt = [np.full(n * n, 2.0), np.zeros(n * n)]
    cx, cy = 3, -2
    I = O.sr_integrals(p, [u, v], vt, cx, cy)
    disk = ((ii - n // 2 - cx) ** 2 + (jj - n // 2 - cy) ** 2 < 18 * 18)
    N = disk.sum()
    hh = c.hx * c.hy
    assert I[0] == pytest.approx(-N * hh, rel=1e-12)
    assert I[1] == pytest.approx(0.0, abs=1e-12)
    # ft = hx*i*fy - hy*j*fx = -hy*j ; I2 = sum fx0*ft = sum hy*j
    assert I[2] == pytest.approx(c.hy * jj[disk].sum() * hh, rel=1e-12)
    # template ft0 = -hy*j*fx0 = +hy*j, so I8 = sum ft0*ft = -sum (hy*j)^2
    assert I[8] == pytest.approx(-(c.hy ** 2) * (jj[disk] ** 2).sum() * hh, rel=1e-12)
    assert I[9] == pytest.approx(-2.0 * N * hh, rel=1e-12)
    assert I[10] == pytest.approx(0.0, abs=1e-12)


def test_bfecc_identity_and_constants():
    p = _sr_params(40, 10)
    rng = np.random.default_rng(0)
    f = [rng.random(1600), rng.random(1600)]
    out = O.bfecc(p, f, [0.0, 0.0, 0.0], 0.0)
    assert np.array_equal(out[0], f[0]) and np.array_equal(out[1], f[1])
    const = [np.full(1600, 0.37), np.full(1600, 0.2)]
    out = O.bfecc(p, const, [2.0, -1.0, 0.3], 0.4)
    np.testing.assert_allclose(out[0], 0.37, rtol=0, atol=1e-15)


def _gauss_error(n):
    L = 4.0
    p = T.default_params(nx=n, ny=n, Lx=L, Ly=L, reduceSym=1, dt=0.02 * (64 / n))
    c = O.derive(p)
    jj, ii = np.mgrid[0:n, 0:n].astype(float)
    x, y = ii * c.hx, jj * c.hy
    sig = 0.4
    g0 = np.exp(-((x - 1.6) ** 2 + (y - 2.0) ** 2) / (2 * sig ** 2))
    vel = 1.5  # cm/ms; c = (vel, 0, 0) advects +x by vel*dt per step
    steps = int(round(0.5 / c.dt))
    f = [g0.reshape(-1), np.zeros(n * n)]
    for _ in range(steps):
        f = O.bfecc(p, f, [vel, 0.0, 0.0], 0.0)
    t = steps * c.dt
    exact = np.exp(-((x - 1.6 - vel * t) ** 2 + (y - 2.0) ** 2) / (2 * sig ** 2))
    return np.max(np.abs(f[0].reshape(n, n) - exact))


def test_bfecc_translating_gaussian_second_order():
    e1, e2 = _gauss_error(64), _gauss_error(128)
    assert e2 < 0.05
    assert e1 / e2 > 3.0, (e1, e2)


def test_tips_of_crossing_linear_fields():
    p = T.grid_params(32, 32)
    jj, ii = np.mgrid[0:32, 0:32].astype(float)
    x0, y0 = 13.3, 17.6
    # bilinear fields (exactly representable cell by cell) whose Uth isolines
    # contain the lines x = x0 and y = y0; the cross terms keep the closed
    # form's quadratic non-degenerate (purely linear fields make den1 = 0)
    present = (0.7 + 0.05 * (ii - x0) + 0.02 * (ii - x0) * (jj - y0)).reshape(-1)
    past = (0.7 + 0.05 * (jj - y0) - 0.03 * (ii - x0) * (jj - y0)).reshape(-1)
    tips, keys = O.tips(p, past, present, 1.25)
    assert len(tips) >= 1
    d = np.hypot(tips[:, 0] - x0, tips[:, 1] - y0)
    assert d.min() < 1e-5
    assert np.all(tips[:, 4] == np.float32(1.25))
    assert np.all(np.diff(keys) > 0)
    # the cell containing the crossing produced it
    best = tips[np.argmin(d)]
    assert int(best[0]) == 13 and int(best[1]) == 17


def test_apd_of_linear_strokes():
    n = 4
    dt = 0.02
    st = {k: np.zeros(n) for k in ("apd1", "apd2", "front", "back")}
    st["first"] = np.zeros(n, np.uint8)
    up_at, down_at = 10.3, 31.7  # step indices of the 0.15 crossings (fractional)
    prev = np.zeros(n)
    for count in range(1, 60):
        t = np.full(n, float(count))
        # u rises through 0.15 at up_at and falls through it at down_at
        u = np.where(t < 20, 0.15 + 0.01 * (t - up_at), 0.15 - 0.01 * (t - down_at))
        O.apd_update(dt, count, prev, u, st)
        prev = u
    np.testing.assert_allclose(st["apd1"], dt * (down_at - up_at), rtol=1e-9)


@pytest.mark.parametrize("model", [T.YH_MODEL_2V_REF, T.YH_MODEL_3V_FK])
def test_sr_step_solves_with_old_phi_then_advances_phi(model):
    """Over 8 oracle SR steps: c equals solve_matrix(phi before the step, the
    integrals of the step) bit for bit, and phi after the step equals
    phi before + dt * c bit for bit.  A composition that solved with the
    advanced phi, advanced phi with the previous c, or solved with the
    previous step's integrals fails one of the two."""
    n = 40
    p0 = T.grid_params(n, n, model=model)
    nv = T.nvars(p0)
    f = [np.asarray(x, np.float64) for x in
         Y.initGates(n, n, u_cols=5, rest=1.0 if model else 0.0, refr=0.0 if model else 1.0,
                     nvars=nv)]
    X = O.run(p0, f, 300)
    p = T.grid_params(n, n, model=model, reduceSym=1, tipOffsetX=12, tipOffsetY=12)
    ref = O.SR(p, X, tip0=(n / 2, n / 2))
    dt = O.derive(p).dt
    for _ in range(8):
        phi0 = ref.phi.copy()
        ref.step(1)
        I = np.array(ref.s.I[:])
        c = ref.c
        assert np.isfinite(c).all() and np.abs(I).max() > 0
        np.testing.assert_array_equal(c, O.solve_matrix(phi0, I))
        np.testing.assert_array_equal(ref.phi, phi0 + dt * c)


def test_bfecc_outside_substitute_reads_as_a_hole():
    """Masked BFECC: a cell with a hole on one side and the domain edge on the
    other takes its substitute neighbour outside the domain, which reads as a
    hole.  So the domain equals the same domain grown by a ring of holes
    (where that neighbour is an actual hole), bit for bit on the original
    cells, to rounding.  Translation only (c_theta = 0): the advection field
    is uniform, so the grown grid's shifted positions do not change it."""
    rng = np.random.default_rng(11)
    nx, ny = 23, 19
    m = (rng.random((ny, nx)) > 0.15).astype(np.uint8)
    for j in range(0, ny, 3):  # holes next to the east and west edges
        m[j, 1] = 0
        m[j, nx - 2] = 0
    for i in range(0, nx, 4):  # and next to the south and north edges
        m[1, i] = 0
        m[ny - 2, i] = 0
    f = [np.where(m > 0, rng.random((ny, nx)), 0.0) for _ in range(2)]
    for c in ([0.7, -0.4, 0.0], [-0.5, 0.6, 0.0]):  # both upwind directions
        p = T.grid_params(nx, ny, reduceSym=1, solidSwitch=1)
        out = O.bfecc(p, [x.reshape(-1) for x in f], c, 0.3, solid=m.reshape(-1))
        M = np.zeros((ny + 2, nx + 2), np.uint8)
        M[1:-1, 1:-1] = m
        F = [np.zeros((ny + 2, nx + 2)) for _ in range(2)]
        for a in range(2):
            F[a][1:-1, 1:-1] = f[a]
        P = T.grid_params(nx + 2, ny + 2, reduceSym=1, solidSwitch=1)
        assert O.derive(P).hx == pytest.approx(O.derive(p).hx, rel=1e-15)
        big = O.bfecc(P, [x.reshape(-1) for x in F], c, 0.3, solid=M.reshape(-1))
        for a in range(2):
            got = out[a].reshape(ny, nx)
            want = big[a].reshape(ny + 2, nx + 2)[1:-1, 1:-1]
            # hx = Lx/(nx-1) of the grown grid differs in its last bit; a
            # substitute read as anything but 0 differs by ~1e-2 here
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
