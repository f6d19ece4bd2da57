"""The reference's text files, written and read through the C library.

The layout contract is the reference's: printParameters writes one
"label,value" line per field in a fixed order (printFunctions.cu:290-399) and
loadParamValues reads the value column back by position (saveFiles.cu:596-708).
The positions and labels below are the ones those two functions share; the
values are checked against printf formatting done here in Python.
"""
import ctypes as C

import numpy as np
import pytest

from yolohtli_amd import abi_types as T
from yolohtli_amd import files as FF
from yolohtli_amd._lib import YhError, lib

# (line, label) of the fields loadParamValues assigns to solver parameters
SOLVER_LINES = [
    (10, "Reduce symmetry:"), (11, "Solid boundary:"), (12, "Neumann BCs:"),
    (13, "Gate diffusion:"), (15, "Anisotropic tisue:"), (18, "Laplacian order:"),
    (19, "Scheme order for time integration (Euler or RK4):"), (22, "# grid points X ="),
    (23, "# grid points Y ="), (24, "Physical Lx length"), (25, "Physical Ly length"),
    (26, "Physical dx"), (27, "Physical dy"), (28, "Time step:"),
    (29, "Diffusion parallel component:"), (30, "Diffusion perpendicular component:"),
    (31, "Initial fiber angle:"), (38, "rx (Dxx*dt/(dx*dx)):"), (39, "ry (Dyy*dt/(dy*dy)):"),
    (40, "Gate r-scale:"), (63, "Tip offset x"), (64, "Tip offset y"),
    (66, "Dirichlet BC value:"), (77, "Last tip point X"), (78, "Last tip point Y"),
    (82, "Filament voltage threshold:"), (83, "time scale (tc):"), (84, "alpha:"),
    (85, "beta:"), (86, "gamma:"), (87, "delta:"), (88, "epsilon:"), (89, "mu:"),
    (90, "theta:"),
]


def _f(x):
    """The reference's '%f' of (float)x."""
    return "%f" % float(np.float32(x))


def _lines(path):
    with open(path) as fh:
        return [ln.rstrip("\n").split(",", 1) for ln in fh]


@pytest.fixture
def odd_params():
    return T.grid_params(300, 200, dx=0.0275, diff_par=0.0011, diff_per=0.0011, rscale=0.02,
                         solidSwitch=1, reduceSym=1, tipOffsetX=40, tipOffsetY=30, Uth=0.65,
                         tc=1.5, alpha=0.21, beta=1.05, gamma=0.01, delta=0.9, eps=0.006,
                         mu=1.1, theta=0.002, boundaryVal=0.25)


def test_params_csv_layout(tmp_path, odd_params):
    p = odd_params
    path = tmp_path / "dataparamcsv.csv"
    FF.printParameters(path, p, tip=(12.5, 7.25))
    rows = _lines(path)
    assert len(rows) == 91
    assert rows[0] == ["Initial condition path:", "NA"] and rows[1][1] == "NA"
    for line, label in SOLVER_LINES:
        assert rows[line][0] == label, (line, rows[line][0])
    hx = p.Lx / (p.nx - 1)
    dt_half = 0.5 * p.dt
    want = {10: "1", 11: "1", 12: "1", 13: "1", 15: "0", 18: "0", 19: "1", 22: "300",
            23: "200", 24: _f(p.Lx), 25: _f(p.Ly), 26: _f(hx),
            # reduceSym: the halved step is printed doubled (printFunctions.cu:320)
            28: "%f" % (2.0 * float(np.float32(dt_half))),
            29: _f(0.0011), 38: _f(dt_half * 0.0011 / hx ** 2), 40: _f(0.02), 63: "40",
            64: "30", 66: _f(0.25), 77: _f(12.5), 78: _f(7.25), 82: _f(0.65), 83: _f(1.5),
            84: _f(0.21), 90: _f(0.002)}
    for line, val in want.items():
        assert rows[line][1] == val, (line, rows[line][1], val)
    assert rows[14][1] == " 1"  # "Tip trajectory algorithm:, %d" keeps its space


def test_params_csv_round_trip(tmp_path, odd_params):
    p = odd_params
    path = tmp_path / "dataparamcsv.csv"
    FF.printParameters(path, p)
    q = FF.loadParamValues(path, T.default_params(model=T.YH_MODEL_3V_FK))
    for name in ("nx", "ny", "reduceSym", "solidSwitch", "neumannBC", "gateDiff", "anisotropy",
                 "lap4", "timeIntOrder", "tipOffsetX", "tipOffsetY"):
        assert getattr(q, name) == getattr(p, name), name
    # values go through '%f' and strtof (saveFiles.cu:576)
    for name in ("Lx", "Ly", "diff_par", "diff_per", "degrad", "rscale", "boundaryVal", "Uth",
                 "tc", "alpha", "beta", "gamma", "delta", "eps", "mu", "theta"):
        want = float(np.float32(float("%f" % float(np.float32(getattr(p, name))))))
        assert getattr(q, name) == want, name
    assert q.dt == float(np.float32(0.02))
    assert q.model == T.YH_MODEL_3V_FK  # not in the file: kept


def test_params_csv_short_file_is_an_error(tmp_path):
    path = tmp_path / "short.csv"
    path.write_text("a,1\nb,2\n")
    with pytest.raises(YhError):
        FF.loadParamValues(path)


@pytest.mark.parametrize("value", ["nan", "1e20", "-inf"])
def test_params_csv_bad_integer_field_is_an_error(tmp_path, value):
    """An integer field that is not a finite value in int range is rejected,
    and the caller's parameters stay as they were."""
    path = tmp_path / "dataparamcsv.csv"
    FF.printParameters(path, T.default_params())
    lines = path.read_text().splitlines()
    k = next(i for i, l in enumerate(lines) if l.startswith("# grid points X ="))
    lines[k] = "# grid points X =," + value
    path.write_text("\n".join(lines) + "\n")
    q = T.default_params(model=T.YH_MODEL_3V_FK)
    q.nx = 77
    with pytest.raises(YhError):
        FF.loadParamValues(path, q)
    assert lib().yh_params_read_csv(str(path).encode(), C.byref(q)) == T.YH_EINVAL
    assert q.nx == 77 and q.model == T.YH_MODEL_3V_FK


def test_state_dat_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 37 * 23
    u, v, w = (rng.uniform(-0.2, 1.2, n).astype(np.float32) for _ in range(3))
    path = tmp_path / "raw_data.dat"
    FF.print2D2column(path, u, v)
    first = path.read_text().splitlines()[0]
    assert first == "%f %f" % (u[0], v[0])
    uu, vv = FF.loadData(path, 37, 23)
    assert np.max(np.abs(uu - u)) <= 5e-7 + 1e-7 and np.max(np.abs(vv - v)) <= 6e-7
    FF.print2D2column(path, u, v, w)
    uu, vv, ww = FF.loadData(path, 37, 23, nvars=3)
    assert np.max(np.abs(ww - w)) <= 6e-7
    with pytest.raises(YhError):
        FF.loadData(path, 37, 24, nvars=3)  # file too short


def test_tip_files_append(tmp_path):
    t1 = np.arange(15, dtype=np.float32).reshape(3, 5) * 0.5
    t2 = -np.arange(10, dtype=np.float32).reshape(2, 5)
    p1, p2 = tmp_path / "dataTip.dat", tmp_path / "dataTipSize.dat"
    FF.printTip(p1, p2, t1, append=False)
    FF.printTip(p1, p2, t2, append=True)
    FF.printTip(p1, p2, np.zeros((0, 5)), append=True)  # no tips: nothing written
    lines = p1.read_text().splitlines()
    assert len(lines) == 5
    assert lines[0] == "%f %f %f %f %f" % tuple(t1[0])
    assert p2.read_text() == "3\n2\n"
    np.testing.assert_array_equal(np.loadtxt(p1), np.concatenate([t1, t2]))


def test_sym_file(tmp_path):
    rng = np.random.default_rng(4)
    h = rng.normal(size=(11, 6))
    path = tmp_path / "c_phi_list_sym.dat"
    FF.printSym(path, h)
    got = np.loadtxt(path)
    assert got.shape == (11, 6)
    assert np.max(np.abs(got - h)) <= 5e-7


def test_mask_file(tmp_path):
    rng = np.random.default_rng(5)
    m = (rng.random(20 * 13) > 0.2).astype(np.uint8)
    path = tmp_path / "holes.dat"
    FF.write_mask(path, m, 20, 13)
    assert path.read_text().splitlines()[0] in ("   1.0000000e+00", "   0.0000000e+00")
    np.testing.assert_array_equal(FF.read_mask(path, 20, 13), m)
    # any white space, any number format; > 0.5 is tissue (main.cu:679)
    path.write_text("0.5 0.51\n1\t0\n   1.0000000e+00\n2 -1 0.4999\n")
    np.testing.assert_array_equal(FF.read_mask(path, 4, 2), [0, 1, 1, 0, 1, 1, 0, 0])
    with pytest.raises(YhError):
        FF.read_mask(path, 3, 3)
