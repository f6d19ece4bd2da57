"""Gradient path: a sensor-mismatch loss through the FK stepper, exposed to
PyTorch as a torch.autograd.Function.

The forward and the hand-written adjoint both run in libyolohtli_amd.so
(yh_adjoint_*); torch only provides the device tensors and the stream.
Gradients are with respect to the spatial diffusion field D (cm^2/ms per
cell) and six ionic time constants (tau_d, tau_r, tau_si, tau_0, tau_v+,
tau_w+).  The forward uses the tanh-smoothed Heaviside of width
``params.heaviside_width``.
"""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import check, lib
from .solver import _ptr, _stream_handle, seed

TAU_NAMES = ("tau_d", "tau_r", "tau_si", "tau_0", "tau_vp", "tau_wp")

def _protos():
    return lib()


class SensorProblem:
    """Forward + adjoint of one FK tissue with sensors (owns device buffers)."""

    def __init__(self, p, nsteps, sensors, x0, rec_every=1, ckpt_every=100, mask=None):
        L = _protos()
        self.p = T.copy_params(p)
        self.nsteps = int(nsteps)
        self.rec = int(rec_every)
        self.sensors = np.ascontiguousarray(sensors, dtype=np.int32)
        self.nrec = self.nsteps // self.rec
        h = C.c_void_p()
        check(L.yh_adjoint_create(C.byref(self.p), self.nsteps, self.rec, int(ckpt_every),
                                  _ptr(self.sensors), len(self.sensors), C.byref(h)),
              "yh_adjoint_create")
        self._h = h
        if mask is not None:
            m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
            check(L.yh_adjoint_set_mask(self._h, _ptr(m)), "yh_adjoint_set_mask")
        xs = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1) for f in x0]
        check(L.yh_adjoint_set_initial(self._h, _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2])),
              "yh_adjoint_set_initial")
        self._last = None
        self._gen = 0

    def set_substeps(self, k):
        """Steps per launch of the forward and reverse sweeps (1: one-step
        kernels; 2..8: LDS-tiled)."""
        check(_protos().yh_adjoint_set_substeps(self._h, int(k)), "yh_adjoint_set_substeps")

    def close(self):
        if getattr(self, "_h", None):
            _protos().yh_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def default_taus(self):
        return np.array([getattr(self.p.fk, n) for n in TAU_NAMES], np.float64)

    def set_targets(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        assert t.size == self.nrec * len(self.sensors)
        check(_protos().yh_adjoint_set_targets(self._h, _ptr(t)), "yh_adjoint_set_targets")

    def forward(self, D, taus=None, stream=None):
        """D: contiguous float32 CUDA tensor of nx*ny cells.  Returns (loss, traces)."""
        taus = self.default_taus() if taus is None else np.asarray(taus, np.float64)
        if D.dtype.itemsize != 4 or not D.is_cuda or not D.is_contiguous():
            raise ValueError("D must be a contiguous float32 CUDA tensor")
        tr = np.empty((self.nrec, len(self.sensors)), np.float32)
        loss = C.c_double()
        check(_protos().yh_adjoint_forward(self._h, D.data_ptr(), _ptr(np.ascontiguousarray(taus)),
                                           _stream_handle(stream), C.byref(loss), _ptr(tr)),
              "yh_adjoint_forward")
        self._gen += 1
        self._last = (D, taus.copy(), self._gen)
        return loss.value, tr

    def backward(self, stream=None):
        """Gradient of the last forward's loss: (gD float64 CUDA tensor, gtau[6])."""
        import torch
        D = self._last[0]
        gD = torch.empty(D.numel(), dtype=torch.float64, device=D.device)
        gt = np.zeros(6)
        check(_protos().yh_adjoint_backward(self._h, _stream_handle(stream), gD.data_ptr(),
                                            _ptr(gt)), "yh_adjoint_backward")
        return gD, gt


def _make_function():
    import torch

    class FKSensorLoss(torch.autograd.Function):
        """loss = FKSensorLoss.apply(D, taus, problem); differentiable in D and taus."""

        @staticmethod
        def forward(ctx, D, taus, problem):
            Dc = D.detach().contiguous().to(torch.float32)
            tv = taus.detach().to("cpu", torch.float64).numpy()
            loss, _ = problem.forward(Dc, tv)
            ctx.problem = problem
            ctx.gen = problem._gen
            ctx.Dc = Dc
            ctx.tv = tv
            ctx.tdev, ctx.tdtype, ctx.Dshape, ctx.Ddtype = taus.device, taus.dtype, D.shape, D.dtype
            return torch.tensor(loss, dtype=torch.float64, device=D.device)

        @staticmethod
        def backward(ctx, gL):
            pr = ctx.problem
            if pr._gen != ctx.gen:  # another forward ran since: recompute this one
                pr.forward(ctx.Dc, ctx.tv)
            gD, gt = pr.backward()
            g = gL.to(torch.float64)
            gD = (g * gD).to(ctx.Ddtype).view(ctx.Dshape)
            gtau = (g.cpu() * torch.from_numpy(gt)).to(ctx.tdtype).to(ctx.tdev)
            return gD, gtau, None

    return FKSensorLoss


_FN = None


def sensor_loss(D, taus, problem):
    """Differentiable loss for torch tensors D (ny, nx) and taus (6,)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(D, taus, problem)


def adjoint_workload(n=256, nsteps=2000, nsensors=64, config=4, width=0.01, dx=0.025):
    """The gradient configuration: FK with a tanh-smoothed Heaviside, cross-field
    start (a plane wave leaves the u = 1 strip at the west edge), hidden
    D = D0*(1 + sum of 4 Gaussian bumps) and seeded sensor sites, both in the
    west band the wave crosses during the run (x < 0.3 n), so the sensors see
    the bumps.  Returns (params, x0, sensors, D_hidden, D0)."""
    rng = np.random.default_rng(seed(config, 0, 0) & 0xFFFFFFFF)
    p = T.grid_params(n, n, dx=dx, model=T.YH_MODEL_3V_FK, heaviside_width=width)
    from .solver import initGates
    x0 = initGates(n, n, u_cols=max(4, n // 16), rest=1.0, refr=0.0, nvars=3)
    cells = set()
    while len(cells) < nsensors:
        i = int(rng.integers(int(0.05 * n), int(0.3 * n)))
        j = int(rng.integers(0, n))
        cells.add(i + n * j)
    sensors = np.array(sorted(cells), np.int32)
    D0 = 0.001
    jj, ii = np.mgrid[0:n, 0:n].astype(np.float64)
    bump = np.zeros((n, n))
    for _ in range(4):
        xm = rng.uniform(0.08, 0.25) * n
        ym = rng.uniform(0.1, 0.9) * n
        sig = rng.uniform(0.04, 0.08) * n
        amp = rng.uniform(-0.3, 0.3)
        bump += amp * np.exp(-((ii - xm) ** 2 + (jj - ym) ** 2) / (2 * sig ** 2))
    D_hidden = (D0 * (1.0 + bump)).astype(np.float32)
    return p, x0, sensors, D_hidden, D0
