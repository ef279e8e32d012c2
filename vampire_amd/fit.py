"""Gradient-based fitting of exchange, anisotropy, damping and site fields to a magnetisation
trajectory.

The trajectory is the deterministic (T = 0) Heun step of sim::LLG_Heun
(src/simulate/LLGHeun.cpp:160-282) run on the device by libvmx; its gradient comes from the
library's hand-written discrete adjoint (vmx_fit_backward), with segments between checkpoints
recomputed.  PyTorch only wraps the two calls as a torch.autograd.Function so a loss written in
torch can be minimised with torch optimisers; no spin dynamics runs in torch.

Parameters are in SI: J (J), Ku (J), alpha, and per-site fields h (T, [3, N]).  The library works
with the reduced J/mu_s and Ku/mu_s (initialise_variables.cpp:427-429, anisotropy/initialize.cpp:
59-62); the chain factor is 1/mu_s for both.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .context import Context
from .system import MU_B, System


class Trajectory:
    """A device context set up for fitting (one material, uniform isotropic exchange)."""

    def __init__(self, system: System, spins0: np.ndarray, nsteps: int, record_every: int = 1,
                 checkpoint_every: int = 100, precision: str = "fp64", H: Sequence[float] = (0.0, 0.0, 0.0),
                 topology: str = "auto"):
        self.ctx = Context(system, precision=precision, topology=topology)
        self.n = system.n
        self.mu_s = system.materials[0]["mu_s_SI"]
        self.spins0 = np.ascontiguousarray(spins0, dtype=np.float64).reshape(3, self.n)
        self.nsteps, self.record_every, self.checkpoint_every = nsteps, record_every, checkpoint_every
        self.H = tuple(float(v) for v in H)

    def forward(self, J: float, Ku: float, alpha: float, h: Optional[np.ndarray] = None) -> np.ndarray:
        p = L.vmx_fit_params(J=J / self.mu_s, ku2=Ku / self.mu_s, alpha=alpha, H=(C.c_double * 3)(*self.H))
        mz = np.empty(self.nsteps // self.record_every)
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).reshape(-1)
        s0 = self.spins0.reshape(-1)
        L.check(self.ctx.lib.vmx_fit_forward(self.ctx.h, C.byref(p), None if hh is None else L.dptr(hh), L.dptr(s0),
                                             self.nsteps, self.record_every, self.checkpoint_every, L.dptr(mz)))
        return mz

    def backward(self, dL_dmz: np.ndarray, want_h: bool = True):
        """(dL/dJ, dL/dKu, dL/dalpha) in SI and dL/dh [3, N] for the last forward."""
        g = np.empty(3)
        dh = np.empty(3 * self.n) if want_h else None
        d = np.ascontiguousarray(dL_dmz, dtype=np.float64)
        L.check(self.ctx.lib.vmx_fit_backward(self.ctx.h, L.dptr(d), L.dptr(g), None if dh is None else L.dptr(dh)))
        return g[0] / self.mu_s, g[1] / self.mu_s, g[2], (None if dh is None else dh.reshape(3, self.n))

    def close(self):
        self.ctx.close()


def torch_function():
    """The torch.autograd.Function over a Trajectory (imported lazily: torch is optional)."""
    import torch

    class HeunTrajectory(torch.autograd.Function):
        @staticmethod
        def forward(ctx, J, Ku, alpha, h, traj: Trajectory):
            mz = traj.forward(float(J), float(Ku), float(alpha),
                              None if h is None else h.detach().cpu().double().numpy())
            ctx.traj = traj
            ctx.has_h = h is not None
            ctx.dev = J.device
            return torch.as_tensor(mz, dtype=torch.float64, device=J.device)

        @staticmethod
        def backward(ctx, grad_mz):
            dJ, dKu, da, dh = ctx.traj.backward(grad_mz.detach().cpu().double().numpy(), want_h=ctx.has_h)
            t = lambda v: torch.tensor(v, dtype=torch.float64, device=ctx.dev)
            gh = torch.as_tensor(dh, dtype=torch.float64, device=ctx.dev) if ctx.has_h else None
            return t(dJ), t(dKu), t(da), gh, None

    return HeunTrajectory
