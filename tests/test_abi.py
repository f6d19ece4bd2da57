"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
keeps the reference defaults, rejects unbuilt branches, and its workload
generators agree with an independent Python restatement of splitmix64."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle.oracle as O
import yolohtli_amd as Y
from yolohtli_amd import _lib
from yolohtli_amd import abi_types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(yh_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    h = C.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    # and the Python prototypes cover the declared surface
    assert not [n for n in names if n not in _lib.PROTOTYPES], \
        [n for n in names if n not in _lib.PROTOTYPES]


def test_defaults_are_reference_parameterSetup():
    p = T.YhParams()
    _lib.lib().yh_params_default(C.byref(p))
    q = T.default_params()
    for name, _ in T.YhParams._fields_:
        if name == "fk":
            continue
        assert getattr(p, name) == getattr(q, name), name
    for name, _ in T.YhFkParams._fields_:
        assert getattr(p.fk, name) == getattr(q.fk, name), name
    # saveFiles.cu:140-148,164,194-201,218-227
    assert (p.nx, p.ny, p.Lx, p.dt, p.diff_par, p.rscale) == (512, 512, 12.0, 0.02, 0.001, 0.01)
    assert (p.tc, p.alpha, p.beta, p.gamma, p.delta, p.eps, p.mu, p.theta) == \
        (1.0, 0.2, 1.1, 0.0, 1.0, 0.005, 1.0, 0.0)
    assert (p.Uth, p.tipOffsetX, p.tipOffsetY, p.neumannBC, p.gateDiff) == (0.7, 160, 160, 1, 1)


@pytest.mark.parametrize("over", [dict(), dict(reduceSym=1), dict(degrad=30.0, diff_per=0.0005),
                                  dict(nx=100, ny=37, Lx=3.0, Ly=1.0)])
def test_derive_matches_oracle(over):
    p = T.default_params(**over)
    d = Y.derive(p)
    c = O.derive(p)
    for name in ("hx", "hy", "dt", "Dxx", "Dyy", "rx", "ry", "invdx", "invdy"):
        assert getattr(d, name) == pytest.approx(getattr(c, name), rel=1e-15), name


@pytest.mark.parametrize("over,code", [
    (dict(timeIntOrder=3), T.YH_EINVAL),
    (dict(timeIntOrder=0), T.YH_EINVAL),
    (dict(dt=2.0), T.YH_EUNSTABLE),
    (dict(nx=2), T.YH_EINVAL),
    (dict(nx=32769, ny=32768), T.YH_EINVAL),
    (dict(model=7), T.YH_EINVAL),
    (dict(model=1, fk=dict(tau_d=0.0)), T.YH_EINVAL),
])
def test_derive_rejects(over, code):
    p = T.default_params(**over)
    d = T.YhDerived()
    assert _lib.lib().yh_params_derive(C.byref(p), C.byref(d)) == code
    with pytest.raises(RuntimeError):
        Y.derive(p)


M64 = (1 << 64) - 1


class _SM:
    def __init__(self, s):
        self.s = s & M64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53


def test_splitmix_known_answer():
    # splitmix64 published first outputs for seed 0
    m = _SM(0)
    assert m.next() == 0xE220A8397B1DCDAF
    assert m.next() == 0x6E789E6AA1B965F4


def test_seed_and_random_ic_follow_splitmix():
    s = Y.seed(3, 5, 7)
    m = _SM(Y.BASE_SEED)
    h = m.next()
    for v in (3, 5, 7):
        h = _SM(h ^ ((v * 0xD1B54A32D192ED03) & M64)).next()
    assert s == h
    assert len({Y.seed(1, t, 0) for t in range(64)}) == 64
    u, v, w = Y.init_random(16, 8, 12345, gate=1.0, nvars=3)
    m = _SM(12345)
    want = np.array([m.uniform() for _ in range(128)], np.float32)
    assert np.array_equal(u, want)
    assert np.all(v == 1.0) and np.all(w == 1.0)


def test_holes_match_bruteforce():
    nx, ny, nh = 96, 80, 12
    m = Y.make_holes(nx, ny, nh, 3.0, 12.0, 777)
    g = _SM(777)
    want = np.ones((ny, nx), np.uint8)
    jj, ii = np.mgrid[0:ny, 0:nx]
    for _ in range(nh):
        cx, cy = g.uniform() * nx, g.uniform() * ny
        r = 3.0 + g.uniform() * 9.0
        want[(ii - cx) ** 2 + (jj - cy) ** 2 <= r * r] = 0
    assert np.array_equal(m.reshape(ny, nx), want)


def test_cross_field_layout():
    u, v = Y.initGates(64, 32)
    U, V = u.reshape(32, 64), v.reshape(32, 64)
    assert np.all(U[:, :8] == 1) and np.all(U[:, 8:] == 0)  # main.cu:606-611
    assert np.all(V[16:, :] == 1) and np.all(V[:16, :] == 0)  # main.cu:613-618


def test_solve_matrix_host_matches_oracle():
    import ctypes as C
    from yolohtli_amd._lib import lib
    rng = np.random.default_rng(3)
    for _ in range(5):
        I = rng.normal(size=12) + np.array([3, 0, 0, 0, 3, 0, 0, 0, 3, 0, 0, 0])
        phi = rng.normal(size=3)
        c = np.zeros(3)
        rc = lib().yh_solve_matrix(phi.ctypes.data_as(C.c_void_p), I.ctypes.data_as(C.c_void_p),
                                   c.ctypes.data_as(C.c_void_p))
        assert rc == 0
        np.testing.assert_allclose(c, O.solve_matrix(phi, I), rtol=1e-13)
