"""The hand-written adjoint (marfa_pt_tables_adjoint) and its torch.autograd.Function.

References:
  * tests/fp64_model.py -- the gridded model in FP64 PyTorch, geometry and values pinned to
    the oracle here (CPU tests), differentiated by autograd;
  * central finite differences (h = 1e-4 relative) of the GPU forward with FP64 records.
The model is piecewise smooth: region boundaries of the reference's Voigt (e.g. the switch
to the chi-corrected Lorentz at VX = 15 jumps by ~0.5/VX^2 relative), the cascade limiter
and the grid selection are discontinuities.  The adjoint is the derivative between them
(what autograd of the FP64 model gives); a finite difference across a jump is not, so the
finite-difference check is held to 1e-3 where the parameter moves no boundary (S, which
scales a line) and to its median elsewhere.
"""
import numpy as np
import pytest
import torch

import fp64_model as fm
import pyoracle as po
from marfa_amd import synth

TIPS = synth.load_tips()
SEED = synth.SEED_BASE + 4


def adjoint_job(n=2000, nlev=3, chi=0, target=0):
    """Configuration-4 laws (1000-1100 cm-1, P 10 -> 1e-3 atm, T 300 -> 180 K, cutoff 25)."""
    return synth.Job(synth.uniform_lines(n, 1000.0, 1100.0, SEED), synth.log_levels(nlev, 10.0, 1e-3, 300.0, 180.0, 0.965),
                     1000.0, 1100.0, 25.0, chi_id=chi, target=target)


def perturbed(params, rng):
    """k-hat's parameters: S and gamma_air scaled by 1 + U(-0.05, 0.05)."""
    out = [p.detach().clone() for p in params]
    for k in (0, 1):
        f = torch.as_tensor(1 + rng.uniform(-0.05, 0.05, tuple(out[k].shape)), dtype=out[k].dtype, device=out[k].device)
        out[k] = out[k] * f
    return out


@pytest.mark.parametrize("chi,target", [(0, 0), (2, 1)])
def test_fp64_model_matches_oracle(chi, target):
    job = synth.Job(synth.uniform_lines(300, 985.0, 1035.0, 5), synth.log_levels(2, 10.0, 1e-3, 300.0, 180.0, 0.965),
                    1000.0, 1020.0, 25.0, chi_id=chi, target=target)
    m = fm.Model(job, TIPS)
    p = m.params()
    for lev in range(2):
        f = m.level_samples(p, lev).detach().numpy()
        v = m.geo.levels[lev]["val"]
        assert np.max(np.abs(f - v) / np.abs(v)) < 1e-12
    rec = m.records(p).detach().numpy()
    ref, _ = po.OracleJob(job, TIPS).pt_tables(po.F64)
    assert np.max(np.abs(rec - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-12


def test_fp64_model_cascade_matches_oracle():
    rng = np.random.default_rng(1)
    g = rng.uniform(0.5, 2.0, (2, fm.NNODE))
    got = fm.cascade(torch.as_tensor(g)).numpy()
    for r in range(2):
        grids = np.ascontiguousarray(g[r, :fm.NGRID])
        rk = np.ascontiguousarray(g[r, fm.NGRID:])  # direct deposits in, record out
        po.lib().mo_cascade_f64(grids.ctypes.data_as(po.C.POINTER(po.C.c_double)),
                                rk.ctypes.data_as(po.C.POINTER(po.C.c_double)))
        assert np.max(np.abs(got[r] - rk) / np.abs(rk)) < 1e-14


@pytest.fixture(scope="module")
def reference_grads():
    """FP64 autograd gradients of L = sum (log k - log k_hat)^2 at 2,000 lines."""
    rng = np.random.default_rng(SEED)
    job = adjoint_job()
    ref = fm.Model(job, TIPS)
    p = ref.params()
    target = ref.records(perturbed(p, rng)).detach()
    rec = ref.records(p)
    rec.retain_grad()
    fm.loss(rec, target).backward()
    return job, target, rec.grad.detach(), [x.grad.numpy() for x in p]


NAMES = ("S", "gamma_air", "n_T", "delta", "T", "P")


def compare(got, ref, tol_norm, tol_median):
    for name, a, b in zip(NAMES, got, ref):
        scale = np.abs(b).max()
        err = np.abs(np.asarray(a) - b)
        rel = err / np.maximum(np.abs(b), 1e-3 * scale)
        print(f"{name}: max err / max |g| {err.max() / scale:.3e}, median rel {np.median(rel):.3e}", flush=True)
        assert err.max() <= tol_norm * scale, name
        assert np.median(rel) <= tol_median, name


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [1, 0])
def test_adjoint_operator_vs_fp64_autograd(reference_grads, prec):
    """Same dL/dk in: the adjoint of the deposit + cascade is autograd's to FP32/FP64 accuracy."""
    from marfa_amd.autograd import PTTableModel
    job, _, gk, gref = reference_grads
    model = PTTableModel(job, TIPS, precision=prec)
    try:
        gl, gv = (t.cpu().numpy() for t in model.adjoint(gk.to("cuda", torch.float32)))
    finally:
        model.close()
    compare([gl[0], gl[1], gl[2], gl[3], gv[0], gv[1]], gref, 1e-5 if prec == 1 else 1e-6, 1e-6)


@pytest.mark.gpu
def test_autograd_function_end_to_end(reference_grads):
    """torch.autograd through PTTables (FP32 records, FP32K form) against FP64 autograd:
    the records' FP32 rounding enters the log residuals, so the bound is norm-wise."""
    from marfa_amd.autograd import PTTableModel, PTTables, spectral_loss
    job, target, _, gref = reference_grads
    model = PTTableModel(job, TIPS)
    try:
        tp = [t.requires_grad_(True) for t in model.params()]
        loss = spectral_loss(PTTables.apply(*tp, model), target.to("cuda", torch.float32))
        loss.backward()
        got = [t.grad.double().cpu().numpy() for t in tp]
    finally:
        model.close()
    compare(got, gref, 1e-3, 1e-4)


@pytest.mark.gpu
def test_adjoint_deterministic():
    from marfa_amd.autograd import PTTableModel
    job = adjoint_job(n=1500, nlev=2)
    model = PTTableModel(job, TIPS)
    try:
        rec = model.records()
        g = torch.randn(rec.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
        a = model.adjoint(g)
        b = model.adjoint(g)
    finally:
        model.close()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


@pytest.mark.gpu
def test_adjoint_rejects_perrin():
    from marfa_amd import lbl
    from marfa_amd.autograd import PTTableModel
    model = PTTableModel(adjoint_job(n=500, nlev=1, chi=3), TIPS)
    try:
        rec = model.records()
        with pytest.raises(RuntimeError, match="Perrin"):
            model.adjoint(torch.ones_like(rec))
    finally:
        model.close()


@pytest.mark.gpu
def test_finite_differences_config4():
    """Configuration 4 (200,000 lines, 30 levels): central differences (h = 1e-4 relative,
    FP64 form with FP64 records) on 64 random parameters -- 10 each of S, gamma_air, n_T,
    delta and 12 each of T, P, drawn among those whose loss change 2 h |g| is above the
    FP64 loss's resolution -- against the FP32K adjoint."""
    from marfa_amd import lbl
    from marfa_amd.autograd import PTTableModel, PTTables, spectral_loss
    rng = np.random.default_rng(SEED)
    job = synth.config4()
    m64 = PTTableModel(job, TIPS, precision=lbl.FP64)
    m32 = PTTableModel(job, TIPS, precision=lbl.FP32K)
    try:
        base = list(m32.params())
        m64.set_params(*perturbed(base, rng))
        tgt = m64.records_f64()

        def loss_at(ps):
            m64.set_params(*ps)
            return float(spectral_loss(m64.records_f64(), tgt))

        tp = [b.clone().requires_grad_(True) for b in base]
        L0 = spectral_loss(PTTables.apply(*tp, m32).double(), tgt)
        L0.backward()
        g = [t.grad.double().cpu().numpy() for t in tp]
        floor = 1e-12 * float(L0.detach())
        rels = {}
        for w, cnt in ((0, 10), (1, 10), (2, 10), (3, 10), (4, 12), (5, 12)):
            vals = base[w].double().cpu().numpy()
            ok = np.nonzero(2e-4 * np.abs(vals * g[w]) > 1e3 * floor)[0]
            for i in rng.choice(ok, size=min(cnt, ok.size), replace=False):
                h = 1e-4 * abs(float(vals[i]))
                pp = [b.clone() for b in base]
                pm = [b.clone() for b in base]
                pp[w][i] += h
                pm[w][i] -= h
                step = float(pp[w][i] - pm[w][i])  # the FP32 step actually taken
                fd = (loss_at(pp) - loss_at(pm)) / step
                rels.setdefault(NAMES[w], []).append(abs(g[w][i] - fd) / abs(fd))
                print(f"{NAMES[w]}[{i}]: fd {fd:.6e} adjoint {g[w][i]:.6e} rel {rels[NAMES[w]][-1]:.2e}", flush=True)
    finally:
        m64.close()
        m32.close()
    allr = np.concatenate([np.array(v) for v in rels.values()])
    assert allr.size >= 60
    assert max(rels["S"]) <= 1e-3
    assert np.median(allr) <= 1e-3
