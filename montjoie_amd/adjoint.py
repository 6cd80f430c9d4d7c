"""Receiver-misfit gradients with respect to per-element mu (= c^2, rho = 1)
through the discrete adjoint leapfrog of the C ABI, and a
``torch.autograd.Function`` around it.

The adjoint has no counterpart in the reference (Montjoie has no time-domain
adjoint); it differentiates the forward path that reproduces Montjoie's
leapfrog (Instationary/ModifiedEquation.cxx:295-345).  torch is imported
before the HIP library is loaded so that both use one HIP runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import dptr
from .problem import WaveProblem


class AdjointSolver:
    """Forward / backward pair on one acoustic WaveProblem (single shot,
    zero initial state, sources / receivers / pulse already set)."""

    def __init__(self, problem: WaveProblem, dt: float, nsteps: int, nsnap: int = 20,
                 t_begin: float = 0.0):
        if problem.ncomp != 1 or problem.nshots != 1:
            raise ValueError("adjoint: single-shot acoustic problems only")
        self.P = problem
        self.dt = dt
        self.nsteps = nsteps
        self.nsnap = nsnap
        self.t_begin = t_begin

    def set_mu(self, mu):
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, np.float64), (self.P.mesh.E,)))
        self.P._chk(_lib.lib().mjw_set_mu(self.P.handle, dptr(mu)), "set_mu")

    def forward(self, mu=None) -> np.ndarray:
        """Traces [nsteps+1, nrec] of the zero-initial-state solve."""
        if mu is not None:
            self.set_mu(mu)
        self.P.leapfrog_init(self.dt, self.t_begin)
        tr = np.zeros((self.nsteps + 1, self.P.nrec))
        self.P._chk(_lib.lib().mjw_adjoint_forward(self.P.handle, self.nsteps, self.nsnap,
                                                   dptr(tr)), "adjoint_forward")
        return tr

    def backward(self, dJ_dtrace) -> np.ndarray:
        """dJ/dmu_e for the traces of the last forward (FP64)."""
        g = np.ascontiguousarray(dJ_dtrace, np.float64).reshape(self.nsteps + 1, self.P.nrec)
        out = np.zeros(self.P.mesh.E)
        self.P._chk(_lib.lib().mjw_adjoint_backward(self.P.handle, dptr(g), dptr(out)),
                    "adjoint_backward")
        return out

    def timing(self):
        a = (C.c_double * 2)()
        self.P._chk(_lib.lib().mjw_adjoint_timing(self.P.handle, a), "adjoint_timing")
        return a[0], a[1]


class WaveTraces(torch.autograd.Function):
    """traces = WaveTraces.apply(mu, solver): receiver traces as a function of
    per-element mu; backward runs the adjoint (float64 tensors)."""

    @staticmethod
    def forward(ctx, mu: torch.Tensor, solver: AdjointSolver):
        tr = solver.forward(mu.detach().cpu().numpy().astype(np.float64))
        ctx.solver = solver
        ctx.mu_device = mu.device
        return torch.from_numpy(tr).to(device=mu.device, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_traces: torch.Tensor):
        g = ctx.solver.backward(grad_traces.detach().cpu().numpy())
        return torch.from_numpy(g).to(device=ctx.mu_device, dtype=torch.float64), None
