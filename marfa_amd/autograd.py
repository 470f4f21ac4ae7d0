"""Differentiable PT tables: torch.autograd.Function over marfa_pt_tables and its adjoint.

Forward: the records of every level x subinterval on the GPU (marfa_pt_tables) for per-line
S, gamma_air, n_T, delta and per-level T, P; every other input (centres, gamma_self, E'',
isotopes, densities, TIPS) is held fixed.  Backward: marfa_pt_tables_adjoint with the
incoming dL/d(records) (a CUDA tensor), which recomputes the grid sums and returns
dL/d(line parameters) and dL/d(T, P) -- no grids are stored between the two calls.
PyTorch only supplies device buffers and the stream; all arithmetic is in the HIP library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lbl, synth


class PTTableModel:
    """A line list and atmosphere whose S, gamma_air, n_T, delta and T, P can vary."""

    def __init__(self, job, tips, precision=lbl.FP32K, device: int = 0):
        self.job = job
        self.ctx = lbl.LBLContext(device)
        self.ctx.set_tips(tips)
        self.spec = lbl.JobSpec.from_job(job, precision)
        self.lines = job.lines
        self.atm = job.atm
        self.ctx.set_lines(job.lines)
        self.ctx.set_levels(job.atm)
        lo, _, _, _ = self.ctx.windows(self.spec)
        self.shape = (len(job.atm), lo.shape[0], lbl.NT)
        self.last_grads = None

    def params(self, device="cuda"):
        """The current parameters as float32 tensors (the reference's FP32 inputs)."""
        L, A = self.lines, self.atm
        f = lambda x: torch.tensor(np.asarray(x, np.float32), device=device)
        return f(L.S), f(L.gair), f(L.nT), f(L.delta), f(A.T), f(A.P)

    def set_params(self, S, gair, nT, delta, T, P):
        host = lambda t: t.detach().to("cpu", torch.float32).numpy()
        L = self.lines
        self.lines = synth.Lines(nu=L.nu, S=host(S), gair=host(gair), gself=L.gself, elow=L.elow,
                                 nT=host(nT), delta=host(delta), molIso=L.molIso)
        self.atm = synth.Atmosphere(P=host(P), T=host(T), rho=self.atm.rho, h=self.atm.h)
        self.ctx.set_lines(self.lines)
        self.ctx.set_levels(self.atm)

    def records(self):
        out = torch.empty(self.shape, dtype=torch.float32, device="cuda")
        self.ctx.pt_tables_device(self.spec, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        return out

    def records_f64(self):
        return self.ctx.pt_tables_f64_torch(self.spec)

    def adjoint(self, grad_out):
        g = grad_out.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream().synchronize()
        gl, gv = self.ctx.pt_tables_adjoint(self.spec, g.data_ptr(), torch.cuda.current_stream().cuda_stream)
        self.last_grads = (gl, gv)
        return gl, gv

    def close(self):
        self.ctx.close()


class PTTables(torch.autograd.Function):
    """records = PTTables.apply(S, gamma_air, n_T, delta, T, P, model)"""

    @staticmethod
    def forward(ctx, S, gair, nT, delta, T, P, model):
        model.set_params(S, gair, nT, delta, T, P)
        ctx.model = model
        ctx.meta = [(t.dtype, t.device) for t in (S, gair, nT, delta, T, P)]
        return model.records()

    @staticmethod
    def backward(ctx, grad_out):
        gl, gv = ctx.model.adjoint(grad_out)
        outs = [gl[0], gl[1], gl[2], gl[3], gv[0], gv[1]]
        grads = tuple(torch.as_tensor(g, dtype=dt, device=dev) for g, (dt, dev) in zip(outs, ctx.meta))
        return grads + (None,)


def spectral_loss(records, target, weight=1.0):
    """sum w (log k - log k_hat)^2 over the points where both spectra are positive."""
    m = (records > 0) & (target > 0)
    d = torch.where(m, torch.log(torch.where(m, records, 1.0)) - torch.log(torch.where(m, target, 1.0)), 0.0)
    return (weight * d.double() ** 2).sum()
