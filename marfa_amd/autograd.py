"""Differentiable PT tables: torch.autograd.Function over the marfa operators.

Forward: the records of every level x subinterval on the GPU (torch.ops.marfa.pt_tables) for
per-line S, gamma_air, n_T, delta and per-level T, P; every other input (centres,
gamma_self, E'', isotopes, densities, TIPS) is held fixed.  The line parameters reach the
library as device tensors (device-to-device copies); the levels, a few values each, through
the host.  Backward: torch.ops.marfa.pt_tables_adjoint with the incoming dL/d(records),
which recomputes the grid sums and returns dL/d(line parameters) and dL/d(T, P) as device
tensors -- no grids are stored between the two calls.  PyTorch only supplies device buffers
and the stream; all arithmetic is in the HIP library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lbl, ops


class PTTableModel:
    """A line list and atmosphere whose S, gamma_air, n_T, delta and T, P can vary.  The
    model holds one parameter set at a time; `version` changes with every set_params, so a
    backward pass can tell whether the parameters of its forward pass are still loaded."""

    def __init__(self, job, tips, precision=lbl.FP32K, device: int = 0):
        self.job = job
        self.device = torch.device("cuda", device)
        self.ctx = lbl.LBLContext(device)
        self.ctx.keep_adjoint_state(True)  # forwards keep what the adjoint reuses
        self.ctx.set_tips(tips)
        self.spec = lbl.JobSpec.from_job(job, precision)
        self.ctx.set_lines(job.lines)
        self.ctx.set_levels(job.atm)
        lo, _, _, _ = self.ctx.windows(self.spec)
        self.shape = (len(job.atm), lo.shape[0], lbl.NT)
        f = lambda x: torch.tensor(np.asarray(x, np.float32), device=self.device)
        L, A = job.lines, job.atm
        self._params = (f(L.S), f(L.gair), f(L.nT), f(L.delta), f(A.T), f(A.P))
        self._rho = f(A.rho)
        self.version = 0
        ops.load()

    def params(self):
        """The current parameters as float32 device tensors (copies)."""
        return tuple(p.clone() for p in self._params)

    def set_params(self, S, gair, nT, delta, T, P):
        f = lambda t: t.detach().to(self.device, torch.float32).contiguous()
        new = tuple(f(t) for t in (S, gair, nT, delta, T, P))
        ops.set_line_params(self.ctx, *new[:4])
        # one host synchronisation for both level comparisons
        if torch.cat([new[4] != self._params[4], new[5] != self._params[5]]).any().item():
            ops.set_levels(self.ctx, new[5], new[4], self._rho)
        self._params = tuple(t.clone() for t in new)
        self.version += 1

    def records(self):
        """Records [nlev, nrec, 20481] (float32, device) of the current parameters."""
        return ops.pt_tables(self.ctx, self.spec)

    def records_f64(self):
        return self.ctx.pt_tables_f64_torch(self.spec)

    def adjoint(self, grad_out):
        """(dL/d(S, gamma_air, n_T, delta) [4, n], dL/d(T, P) [2, nlev]): float64 device tensors."""
        return ops.pt_tables_adjoint(self.ctx, self.spec, grad_out.detach())

    def close(self):
        self.ctx.close()


class PTTables(torch.autograd.Function):
    """records = PTTables.apply(S, gamma_air, n_T, delta, T, P, model)"""

    @staticmethod
    def forward(ctx, S, gair, nT, delta, T, P, model):
        model.set_params(S, gair, nT, delta, T, P)
        ctx.model = model
        ctx.version = model.version
        ctx.save_for_backward(S, gair, nT, delta, T, P)
        return model.records()

    @staticmethod
    def backward(ctx, grad_out):
        saved = ctx.saved_tensors
        if ctx.model.version != ctx.version:
            # the model was applied again since this node's forward: reload its parameters
            ctx.model.set_params(*saved)
        gl, gv = ctx.model.adjoint(grad_out)
        outs = (gl[0], gl[1], gl[2], gl[3], gv[0], gv[1])
        return tuple(g.to(device=t.device, dtype=t.dtype) for g, t in zip(outs, saved)) + (None,)


class SpectralLoss(torch.autograd.Function):
    """sum w (ln k - ln k_hat)^2 by torch.ops.marfa.log_loss (one pass over the points, FP64
    logarithms and a fixed-order sum); backward: dL/dk by torch.ops.marfa.log_loss_grad."""

    @staticmethod
    def forward(ctx, records, target, weight, wscale):
        ctx.save_for_backward(records, target, weight)
        ctx.wscale = wscale
        return ops.load().log_loss(records, target, weight, wscale)

    @staticmethod
    def backward(ctx, g):
        records, target, weight = ctx.saved_tensors
        grad = ops.load().log_loss_grad(records, target, weight, ctx.wscale, g)
        return grad, None, None, None


def spectral_loss(records, target, weight=1.0):
    """sum w (log k - log k_hat)^2 over the points where both spectra are positive (a float64
    scalar).  records: float32 or float64 CUDA tensor; target: the same shape (cast to the
    records' dtype); weight: a number or a tensor of the records' shape."""
    target = target.to(device=records.device, dtype=records.dtype).contiguous()
    if isinstance(weight, torch.Tensor):
        w = weight.to(device=records.device, dtype=torch.float32).expand_as(records).contiguous()
        return SpectralLoss.apply(records.contiguous(), target, w, 1.0)
    return SpectralLoss.apply(records.contiguous(), target, None, float(weight))
