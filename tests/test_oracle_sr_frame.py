"""Pins the oracle's symmetry-reduction step composition (main.cu:894-953) by
the reference's own check of its co-moving frame.

The reference validates a symmetry-reduced run by rebuilding the lab-frame
tip path from the co-moving tips and phi (DATA/processSymmetry.m:88-89),
  x_lab = -phi_x + x cos(phi_t) + y sin(phi_t)
  y_lab = -phi_y - x sin(phi_t) + y cos(phi_t),
and comparing it with a plain run (the reference's README.md:10-12).  With x = hx*i,
y = hy*j measured from cell 0 this is the exact inverse of the frame the
advection field of Cxy_field_kernel (symmetryReduction.cu:20-62) moves: the
co-moving state is u_lab(R(-phi_t) x - (phi_x, phi_y)).  Both runs start
from the same formed spiral with phi = 0, so the two paths coincide up to
discretisation error without the script's circle-fit centring (which only
removes an offset when the start frames differ).

The step is restated here from the oracle's individually pinned pieces
(RD step, tips, disk integrals, solve, BFECC); that restatement equals the
oracle's orc_sr_step_d bit for bit, and a wrong composition departs from the
plain path by more than a cell within 600 steps where the right one stays
within 0.25 of a cell:
  * BFECC applied to the state before RD (the RD output discarded);
  * tips taken between consecutive co-moving states instead of between the
    state and its RD output (the co-moving pattern barely moves);
  * the rotation advected with the wrong sign.
Reversing the order of RD and advection inside a step is another consistent
splitting and is not told apart by this check (nor meant to be).
"""
import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import abi_types as T

N, R = 192, 60          # grid and disk radius (cells)
WARM, STEPS = 8000, 600  # plain steps to a formed spiral; compared steps


def _sr_run(p, X0, tip0, nsteps, variant=None):
    """nsteps SR steps composed as main.cu:894-953 from the oracle's pieces;
    returns the final state, phi and each step's tips mapped to the lab frame
    (in cells) with the phi of that step."""
    c = O.derive(p)
    nx, ny = p.nx, p.ny
    half_x, half_y = np.float32(nx // 2), np.float32(ny // 2)
    X, Xprev = [a.copy() for a in X0], None
    phi, ac, aphi = np.zeros(3), np.zeros(3), 0.0
    cx = int(np.rint(np.float32(tip0[0]) - half_x))
    cy = int(np.rint(np.float32(tip0[1]) - half_y))
    lab = []
    for s in range(nsteps):
        Xo, vt = O.rd_step(p, X, veltan=True)
        if variant == "tips_comoving" and Xprev is not None:
            tt, _ = O.tips(p, X[0], Xprev[0], c.dt * s)
        else:
            tt, _ = O.tips(p, Xo[0], X[0], c.dt * s)  # g_past = RD output (main.cu:898-899)
        if s > 0 and len(tt):  # the disk follows the last tip (symmetryReduction.cu:98-107)
            cx = int(np.rint(tt[-1, 0] - half_x))
            cy = int(np.rint(tt[-1, 1] - half_y))
        ct, st = np.cos(phi[2]), np.sin(phi[2])
        x, y = c.hx * tt[:, 0].astype(np.float64), c.hy * tt[:, 1].astype(np.float64)
        lab.append(np.stack([(x * ct + y * st - phi[0]) / c.hx,
                             (-x * st + y * ct - phi[1]) / c.hy], 1))
        if s == 0:  # count 0: solve, rebuild the field, slice again (main.cu:907-920)
            I = O.sr_integrals(p, X, vt, cx, cy, ac, aphi)
            ac, aphi = O.solve_matrix(phi, I), phi[2]
        I = O.sr_integrals(p, X, vt, cx, cy, ac, aphi)  # slices of gateIn (main.cu:906)
        cc = O.solve_matrix(phi, I)
        ac, aphi = cc.copy(), phi[2]
        Xprev = X
        if variant == "bfecc_before_rd":
            X = O.bfecc(p, X, cc, phi[2])
        elif variant == "rotation_sign":
            X = O.bfecc(p, Xo, [cc[0], cc[1], -cc[2]], phi[2])
        else:
            X = O.bfecc(p, Xo, cc, phi[2])  # advFDBFECC of the RD output (main.cu:930)
        phi = phi + cc * c.dt  # main.cu:935-941
    return X, phi, lab


@pytest.fixture(scope="module")
def spiral():
    """A formed FK spiral on 192^2 (cross-field start, dt = 0.01 ms), its
    last tip, and the plain run's tips over the next STEPS steps."""
    p0 = T.grid_params(N, N, model=T.YH_MODEL_3V_FK, dt=0.01)
    f = Y.initGates(N, N, u_cols=N // 8, rest=1.0, refr=0.0, nvars=3)
    x = O.run(p0, [a.astype(np.float64) for a in f], WARM)
    y = O.run(p0, x, 1)
    tp, _ = O.tips(p0, y[0], x[0], 0.0)
    assert len(tp) == 1
    plain, xp = [], x
    for _ in range(STEPS):
        yp = O.run(p0, xp, 1)
        tt, _ = O.tips(p0, yp[0], xp[0], 0.0)
        plain.append(tt[:, :2].astype(np.float64))
        xp = yp
    # the tip moves: the comparison is not of a fixed point
    assert np.hypot(*(plain[-1][-1] - plain[0][-1])) > 5.0
    p = T.grid_params(N, N, model=T.YH_MODEL_3V_FK, reduceSym=1, tipOffsetX=R, tipOffsetY=R)
    assert O.derive(p).dt == 0.01  # main.cu:148
    return p, x, (float(tp[-1, 0]), float(tp[-1, 1])), plain


def _path_error(plain, lab):
    """Per step, the distance (cells) from the plain run's tips to the nearest
    rebuilt tip; inf where either run has none."""
    out = []
    for a, b in zip(plain, lab):
        if len(a) and len(b):
            out.append(float(np.sqrt(((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)).min()))
        else:
            out.append(np.inf)
    return np.array(out)


def test_composition_equals_the_oracle_step(spiral):
    p, x, tip0, _ = spiral
    sr = O.SR(p, x, tip0=tip0)
    sr.step(20)
    X, phi, _ = _sr_run(p, x, tip0, 20)
    for a in range(3):
        assert np.array_equal(X[a], sr.X[a])
    assert np.array_equal(phi, sr.phi)


def test_rebuilt_lab_frame_tips_follow_the_plain_run(spiral):
    p, x, tip0, plain = spiral
    _, phi, lab = _sr_run(p, x, tip0, STEPS)
    err = _path_error(plain, lab)
    print("rebuilt vs plain tip path (cells), max over the first 100/300/600 steps:",
          [float(err[:k].max()) for k in (100, 300, 600)], "phi", phi)
    assert np.all(err <= 0.25), float(err.max())
    assert abs(phi[2]) > 0.05  # the frame did rotate


@pytest.mark.parametrize("variant", ["bfecc_before_rd", "tips_comoving", "rotation_sign"])
def test_wrong_compositions_leave_the_plain_path(spiral, variant):
    p, x, tip0, plain = spiral
    _, _, lab = _sr_run(p, x, tip0, STEPS, variant)
    err = _path_error(plain, lab)
    print(variant, "max path error (cells):", float(err.max()))
    assert err.max() > 1.0
