"""torch.ops.marfa: PyTorch-ROCm operators over the C ABI (csrc/torch_ops.cpp).

The records of a job are produced as a CUDA (HIP) tensor on the caller's current stream,
and the adjoint takes dL/d(records) as a tensor and returns its gradients as device
tensors, so a differentiable caller never moves line data through the host.  The context
(line list, TIPS, levels) is created and loaded through marfa_amd.lbl.LBLContext.  The
operator library libmarfa_torch.so is built in place by marfa_amd.build; without it (or
without a GPU) these functions raise -- there is no other implementation.
"""
from __future__ import annotations

import os

import torch

from . import lbl

LIB_PATH = os.path.join(lbl.HERE, "libmarfa_torch.so")
_ops = None


def load():
    """Load libmarfa_torch.so (and libmarfa.so, which it links) once; returns torch.ops.marfa."""
    global _ops
    if _ops is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        lbl.library()
        torch.ops.load_library(LIB_PATH)
        _ops = torch.ops.marfa
    return _ops


def _job(spec: lbl.JobSpec):
    return (float(spec.start_wv), float(spec.end_wv), float(spec.cutoff), int(spec.shape_id), int(spec.chi_id),
            int(spec.mol_code), int(spec.target), int(spec.precision))


def handle(ctx: lbl.LBLContext) -> int:
    return int(ctx.h.value)


def pt_tables(ctx: lbl.LBLContext, spec: lbl.JobSpec) -> torch.Tensor:
    """Records [nlev, nrec, 20481] float32 on the context's GPU (marfa_pt_tables)."""
    return load().pt_tables(handle(ctx), *_job(spec))


def pt_tables_adjoint(ctx: lbl.LBLContext, spec: lbl.JobSpec, grad: torch.Tensor):
    """(dL/d(S, gamma_air, n_T, delta) [4, n], dL/d(T, P) [2, nlev]), float64 device tensors
    (marfa_pt_tables_adjoint_dev)."""
    return load().pt_tables_adjoint(handle(ctx), *_job(spec), grad)


def set_line_params(ctx: lbl.LBLContext, S, gair, nT, delta) -> None:
    """Per-line S, gamma_air, n_T, delta from float32 device tensors (marfa_update_line_params)."""
    f = lambda t: t.detach().to(torch.float32).contiguous()
    load().set_line_params(handle(ctx), f(S), f(gair), f(nT), f(delta))


def set_levels(ctx: lbl.LBLContext, P, T, rho) -> None:
    """Level pressure, temperature and density (marfa_set_levels)."""
    load().set_levels(handle(ctx), P.detach(), T.detach(), rho.detach())
    ctx.nlev = int(P.numel())
