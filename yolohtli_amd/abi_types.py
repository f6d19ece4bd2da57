"""ctypes mirrors of the structs in include/yolohtli.h (no library is loaded here)."""
import ctypes as C

YH_OK = 0
YH_EINVAL = 1
YH_EUNSUPPORTED = 2
YH_EHIP = 3
YH_ENOMEM = 4
YH_EUNSTABLE = 5
YH_ESINGULAR = 6
YH_ENOTIP = 7
YH_EOVERFLOW = 8
YH_ENONFINITE = 9
YH_ESTATE = 10
YH_ESCRATCH = 11

YH_MODEL_2V_REF = 0
YH_MODEL_3V_FK = 1

YH_KERNEL_AUTO = 0
YH_KERNEL_DIRECT = 1
YH_KERNEL_TILED = 2

YH_OPT_KEEP_PREV = 1
YH_OPT_SR_LAUNCHES = 2


class YhFkParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "tau_d", "tau_r", "tau_si", "tau_0", "tau_vp", "tau_v1m", "tau_v2m",
        "tau_wp", "tau_wm", "u_c", "u_v", "u_csi", "k")]


class YhParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("dt", C.c_double),
        ("diff_par", C.c_double), ("diff_per", C.c_double),
        ("degrad", C.c_double),
        ("rscale", C.c_double),
        ("tc", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
        ("gamma", C.c_double), ("delta", C.c_double), ("eps", C.c_double),
        ("mu", C.c_double), ("theta", C.c_double),
        ("model", C.c_int),
        ("neumannBC", C.c_int), ("solidSwitch", C.c_int), ("gateDiff", C.c_int),
        ("reduceSym", C.c_int),
        ("timeIntOrder", C.c_int), ("lap4", C.c_int), ("anisotropy", C.c_int),
        ("boundaryVal", C.c_double), ("Uth", C.c_double),
        ("tipOffsetX", C.c_int), ("tipOffsetY", C.c_int),
        ("fk", YhFkParams),
        ("heaviside_width", C.c_double),
    ]


class YhDerived(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "hx", "hy", "dt", "Dxx", "Dyy", "rx", "ry", "invdx", "invdy")]


class YhKernelCfg(C.Structure):
    _fields_ = [("variant", C.c_int), ("substeps", C.c_int), ("rows", C.c_int),
                ("waves", C.c_int), ("cols", C.c_int)]


class YhSrInfo(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("phi", C.c_double * 3), ("I", C.c_double * 12),
                ("cx", C.c_int), ("cy", C.c_int), ("count", C.c_int64), ("status", C.c_int),
                ("ntips", C.c_int)]


class YhTip(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("vx", C.c_float),
                ("vy", C.c_float), ("t", C.c_float)]


def default_fk():
    """Fenton-Karma 1998 Beeler-Reuter parameter set (same values as yh_fk_params_default).

    tau_v1m is the v recovery constant below u_v (rest), tau_v2m the one for
    u_v <= u < u_c: the paper's tau_v2- = 19.6 and tau_v1- = 1250, whose
    tau_v-(u) = H(u - u_v) tau_v1- + H(u_v - u) tau_v2- names them the other
    way round (Fenton et al. 2002 use this library's naming).
    """
    return YhFkParams(tau_d=0.25, tau_r=33.33, tau_si=29.0, tau_0=12.5, tau_vp=3.33,
                      tau_v1m=19.6, tau_v2m=1250.0, tau_wp=870.0, tau_wm=41.0,
                      u_c=0.13, u_v=0.04, u_csi=0.85, k=10.0)


def default_params(**over):
    """parameterSetup defaults (saveFiles.cu:105-231) with Euler + 5-point selected."""
    p = YhParams(nx=512, ny=512, Lx=12.0, Ly=12.0, dt=0.02, diff_par=0.001,
                 diff_per=0.001, degrad=0.0, rscale=0.01, tc=1.0, alpha=0.2, beta=1.1,
                 gamma=0.0, delta=1.0, eps=0.005, mu=1.0, theta=0.0,
                 model=YH_MODEL_2V_REF, neumannBC=1, solidSwitch=0, gateDiff=1,
                 reduceSym=0, timeIntOrder=1, lap4=0, anisotropy=0, boundaryVal=0.0,
                 Uth=0.7, tipOffsetX=160, tipOffsetY=160, fk=default_fk(),
                 heaviside_width=0.0)
    for k, v in over.items():
        if k == "fk" and isinstance(v, dict):
            for fk_k, fk_v in v.items():
                setattr(p.fk, fk_k, fk_v)
        else:
            setattr(p, k, v)
    return p


def copy_params(p, **over):
    q = YhParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(YhParams))
    for k, v in over.items():
        if k == "fk" and isinstance(v, dict):
            for fk_k, fk_v in v.items():
                setattr(q.fk, fk_k, fk_v)
        else:
            setattr(q, k, v)
    return q


def grid_params(nx, ny, dx=0.025, **over):
    """Parameters for an nx x ny grid of spacing dx cm: Lx = dx*(nx-1), since
    hx = Lx/(nx-1) (saveFiles.cu:144)."""
    return default_params(nx=nx, ny=ny, Lx=dx * (nx - 1), Ly=dx * (ny - 1), **over)


def nvars(p):
    return 3 if p.model == YH_MODEL_3V_FK else 2
