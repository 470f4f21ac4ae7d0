"""The hand-written adjoint (marfa_pt_tables_adjoint) and its torch.autograd.Function.

References:
  * tests/fp64_model.py -- the gridded model in FP64 PyTorch, geometry and values pinned to
    the oracle here (CPU tests), differentiated by autograd;
  * central finite differences (h = 1e-4 relative) of the GPU forward with FP64 records.
The model is piecewise smooth: region boundaries of the reference's Voigt (e.g. the switch
to the chi-corrected Lorentz at VX = 15 jumps by ~0.5/VX^2 relative), the cascade limiter
and the grid selection are discontinuities.  The adjoint is the derivative between them
(what autograd of the FP64 model gives); a finite difference across a jump is not, so the
finite-difference check detects such draws with the oracle's decision trace, replaces
them, and holds every other draw of every kind to 1e-3.
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


@pytest.mark.parametrize("chi,target", [(0, 0), (2, 1), (3, 0)])
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


@pytest.fixture(scope="module")
def perrin_grads():
    """The same loss with the Perrin chi factor (temperature-dependent band exponents) at
    1,000 lines: its dL/dT carries the chi factor's temperature derivative."""
    rng = np.random.default_rng(SEED + 3)
    job = adjoint_job(n=1000, chi=3)
    ref = fm.Model(job, TIPS)
    p = ref.params()
    target = ref.records(perturbed(p, rng)).detach()
    rec = ref.records(p)
    rec.retain_grad()
    fm.loss(rec, target).backward()
    return job, rec.grad.detach(), [x.grad.numpy() for x in p]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [1, 0])
def test_adjoint_perrin_vs_fp64_autograd(perrin_grads, prec):
    from marfa_amd.autograd import PTTableModel
    job, gk, gref = perrin_grads
    model = PTTableModel(job, TIPS, precision=prec)
    try:
        gl, gv = (t.cpu().numpy() for t in model.adjoint(gk.to("cuda", torch.float32)))
    finally:
        model.close()
    compare([gl[0], gl[1], gl[2], gl[3], gv[0], gv[1]], gref, 1e-5 if prec == 1 else 1e-6, 1e-6)


def draw_nonstraddling(rng, candidates, want, straddles, limit):
    """Up to `want` of the candidates (in a seeded random order) for which straddles(c) is
    false, trying at most `limit`; returns (kept, number straddling)."""
    kept, bad = [], 0
    for c in rng.permutation(np.asarray(candidates))[:limit]:
        if len(kept) == want:
            break
        if straddles(c):
            bad += 1
        else:
            kept.append(c)
    return kept, bad


def fd_check(name, pairs, tol=1e-3):
    """pairs: (finite difference, FP64-form adjoint, FP32K end-to-end adjoint) per draw."""
    a = np.array(pairs)
    r64 = np.abs(a[:, 1] - a[:, 0]) / np.abs(a[:, 0])
    r32 = np.abs(a[:, 2] - a[:, 0]) / np.abs(a[:, 0])
    print(f"{name}: {len(a)} draws, FP64 adjoint max rel {r64.max():.2e}, FP32K adjoint max rel {r32.max():.2e}",
          flush=True)
    assert r64.max() <= tol, (name, r64.max())
    assert r32.max() <= tol, (name, r32.max())


@pytest.mark.gpu
def test_finite_differences_config4():
    """Central differences of the GPU forward (FP64 form, FP64 records) against the adjoint,
    per parameter, 64 draws: 10 each of S, gamma_air, n_T, delta on configuration 4 (200,000
    lines, 30 levels, h = 1e-4 relative) and 12 each of T, P on configuration 4's laws at 300
    lines (h = 1e-5 relative: a level parameter moves every sample of the level, so at full
    size nearly every step straddles a branch; at 300 lines most do not).  The model is
    piecewise smooth (Voigt regions, chi bands, grid selection, cascade limiter); a draw
    whose +h and -h evaluations take different branches -- the oracle's decision trace of
    the line's samples (or the level's) and the limiter choices of the touched records
    differ -- is a straddle: its difference quotient is not a derivative, so it is counted,
    reported and replaced by the next candidate.  Every kept draw must agree within 1e-3
    relative, for the FP64-form adjoint (dL/dk from FP64 records) and for the FP32K form
    end to end through torch.autograd (FP32 records)."""
    from marfa_amd import lbl
    from marfa_amd.autograd import PTTableModel, PTTables, spectral_loss
    rng = np.random.default_rng(SEED)
    fields = ("S", "gair", "nT", "delta")
    straddled = {}

    def gradients(m64, m32, base, tgt):
        tp = [b.clone().requires_grad_(True) for b in base]
        spectral_loss(PTTables.apply(*tp, m32).double(), tgt).backward()
        g32 = [t.grad.double().cpu().numpy() for t in tp]
        m64.set_params(*base)
        rec = m64.records_f64().requires_grad_(True)
        L0 = spectral_loss(rec, tgt)
        L0.backward()
        gl, gv = (t.cpu().numpy() for t in m64.adjoint(rec.grad.float()))
        return g32, [gl[0], gl[1], gl[2], gl[3], gv[0], gv[1]], float(L0.detach())

    def fd(m64, base, tgt, w, i, h):
        pp = [b.clone() for b in base]
        pm = [b.clone() for b in base]
        pp[w][i] += h
        pm[w][i] -= h
        step = float(pp[w][i] - pm[w][i])  # the FP32 step actually taken
        vals = []
        for ps in (pp, pm):
            m64.set_params(*ps)
            vals.append(float(spectral_loss(m64.records_f64(), tgt)))
        return (vals[0] - vals[1]) / step, pp, pm

    # ---- per-line parameters, configuration 4
    job = synth.config4()
    m64 = PTTableModel(job, TIPS, precision=lbl.FP64)
    m32 = PTTableModel(job, TIPS, precision=lbl.FP32K)
    oj = po.OracleJob(job, TIPS)
    spec64 = lbl.JobSpec.from_job(job, lbl.FP64)
    try:
        base = list(m32.params())
        m64.set_params(*perturbed(base, rng))
        tgt = m64.records_f64()
        g32, g64, L0 = gradients(m64, m32, base, tgt)
        floor = 1e-12 * L0
        for w, name in enumerate(fields):
            vals = base[w].double().cpu().numpy()
            ok = np.nonzero(2e-4 * np.abs(vals * g64[w]) > 1e3 * floor)[0]
            pairs = []

            def straddles(i, w=w, name=name, vals=vals):
                h = 1e-4 * abs(float(vals[i]))
                f, pp, pm = fd(m64, base, tgt, w, int(i), h)
                ojp = oj.with_param(name, int(i), np.float32(pp[w][i].item()))
                ojm = oj.with_param(name, int(i), np.float32(pm[w][i].item()))
                touched = []
                for lev in range(len(job.atm)):
                    dp, recs = ojp.decisions_line(lev, int(i))
                    dm, _ = ojm.decisions_line(lev, int(i))
                    if dp != dm:
                        return True
                    touched += [(lev, r) for r in recs]
                lim = []
                for ps in (pp, pm):
                    m64.set_params(*ps)
                    torch.cuda.synchronize()
                    lim.append(po.limiter_hash([np.concatenate(m64.ctx.grid_sums(spec64, lev, r)[:2])
                                                for lev, r in touched]))
                if lim[0] != lim[1]:
                    return True
                pairs.append((f, g64[w][i], g32[w][i]))
                print(f"{name}[{i}]: fd {f:.6e} adjoint fp64 {g64[w][i]:.6e} fp32k {g32[w][i]:.6e}", flush=True)
                return False

            kept, straddled[name] = draw_nonstraddling(rng, ok, 10, straddles, 40)
            assert len(kept) == 10, (name, len(kept), straddled[name])
            fd_check(name, pairs)
    finally:
        m64.close()
        m32.close()

    # ---- level parameters T and P: configuration 4's laws at 300 lines
    job = synth.Job(synth.uniform_lines(300, 1000.0, 1100.0, SEED), synth.log_levels(30, 10.0, 1e-3, 300.0, 180.0, 0.965),
                    1000.0, 1100.0, 25.0)
    m64 = PTTableModel(job, TIPS, precision=lbl.FP64)
    m32 = PTTableModel(job, TIPS, precision=lbl.FP32K)
    oj = po.OracleJob(job, TIPS)
    try:
        base = list(m32.params())
        m64.set_params(*perturbed(base, rng))
        tgt = m64.records_f64()
        g32, g64, _ = gradients(m64, m32, base, tgt)
        for w, name in ((4, "T"), (5, "P")):
            pairs = []

            def straddles(lev, w=w, name=name):
                v = float(base[w][lev])
                f, pp, pm = fd(m64, base, tgt, w, int(lev), 1e-5 * abs(v))
                kw = lambda ps: {name: np.float32(ps[w][lev].item())}
                if oj.with_level(int(lev), **kw(pp)).decisions_level(int(lev)) != \
                        oj.with_level(int(lev), **kw(pm)).decisions_level(int(lev)):
                    return True
                pairs.append((f, g64[w][lev], g32[w][lev]))
                print(f"{name}[{lev}]: fd {f:.6e} adjoint fp64 {g64[w][lev]:.6e} fp32k {g32[w][lev]:.6e}", flush=True)
                return False

            kept, straddled[name] = draw_nonstraddling(rng, np.arange(len(job.atm)), 12, straddles, 30)
            assert len(kept) == 12, (name, len(kept), straddled[name])
            fd_check(name, pairs)
    finally:
        m64.close()
        m32.close()
    print(f"straddling draws replaced: {straddled}", flush=True)


@pytest.mark.gpu
def test_adjoint_checkpoint_reuse_and_invalidation():
    """The adjoint reuses the forward's level-0 sums and cascade limiter choices (the model's
    context keeps them: marfa_keep_adjoint_state) only for the inputs they were computed
    from: recomputed without a checkpoint, reused (no grid sums recomputed), recomputed after
    the same parameters are set again, and after a forward at other parameters followed by a
    return to these (no forward), the gradients are the same, bit for bit."""
    from marfa_amd.autograd import PTTableModel
    job = adjoint_job(n=1500, nlev=2)
    model = PTTableModel(job, TIPS)
    try:
        p = model.params()
        g = torch.randn(model.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
        a0 = model.adjoint(g)  # no checkpoint: everything recomputed
        model.records()        # forward: the checkpoint holds these parameters' sums and choices
        a1 = model.adjoint(g)  # reuses both
        model.set_params(*p)   # same values, new generation: recomputed
        a2 = model.adjoint(g)
        model.set_params(*perturbed(p, np.random.default_rng(2)))
        model.records()        # checkpoint of other parameters
        model.set_params(*p)   # back, without a forward
        a3 = model.adjoint(g)
    finally:
        model.close()
    for a in (a0, a2, a3):
        assert torch.equal(a1[0], a[0]) and torch.equal(a1[1], a[1])


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_loss_against_numpy(dtype):
    """marfa_log_loss / marfa_log_loss_grad against the same sums in numpy (FP64), with
    zeros and negative values (excluded points), a weight tensor and a scalar weight;
    repeated calls give the same bits."""
    from marfa_amd import ops
    rng = np.random.default_rng(3)
    n = 300_001
    k = 10.0 ** rng.uniform(-30, -20, n)
    kh = k * (1 + rng.uniform(-0.2, 0.2, n))
    k[rng.integers(0, n, 500)] = 0.0
    kh[rng.integers(0, n, 500)] = -1.0
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    kt = torch.tensor(k, dtype=dtype, device="cuda")
    kht = torch.tensor(kh, dtype=dtype, device="cuda")
    wt = torch.tensor(w, device="cuda")
    kk, hh = kt.double().cpu().numpy(), kht.double().cpu().numpy()
    m = (kk > 0) & (hh > 0)
    d = np.where(m, np.log(np.where(m, kk, 1.0)) - np.log(np.where(m, hh, 1.0)), 0.0)
    want = float(np.sum(0.5 * w.astype(np.float64) * d * d))
    got = ops.load().log_loss(kt, kht, wt, 0.5)
    assert abs(float(got) - want) <= 1e-12 * want
    assert torch.equal(got, ops.load().log_loss(kt, kht, wt, 0.5))
    gs = torch.tensor([3.0], dtype=torch.float64, device="cuda")
    grad = ops.load().log_loss_grad(kt, kht, wt, 0.5, gs).double().cpu().numpy()
    want_g = np.where(m, 3.0 * 2.0 * 0.5 * w.astype(np.float64) * d / np.where(m, kk, 1.0), 0.0)
    # the residual d = ln k - ln k_hat cancels: an ulp of either logarithm (|ln k| ~ 60) is an
    # absolute error of d, hence of dL/dk relative to 3 w / k
    tol = 1e-14 if dtype == torch.float64 else 1e-6
    slack = np.where(m, 3.0 * w.astype(np.float64) * 4.0 * np.finfo(np.float64).eps * np.abs(np.log(np.where(m, kk, 1.0)))
                     / np.where(m, kk, 1.0), 0.0)
    assert np.all(np.abs(grad - want_g) <= tol * np.abs(want_g) + slack)
    assert np.all(grad[~m] == 0.0)
    got1 = ops.load().log_loss(kt, kht, None, 1.0)
    assert abs(float(got1) - float(np.sum(d * d))) <= 1e-12 * float(np.sum(d * d))


@pytest.mark.gpu
def test_keep_adjoint_state_switch():
    """marfa_keep_adjoint_state only decides what the forward keeps: the adjoint after a
    forward that kept the limiter choices equals, bit for bit, the one after a forward that
    did not (which recomputes the finer grids)."""
    from marfa_amd import lbl
    job = adjoint_job(n=1500, nlev=2, chi=2)
    out = []
    for keep in (False, True):
        ctx = lbl.LBLContext(0)
        try:
            ctx.keep_adjoint_state(keep)
            ctx.load_job(job, TIPS)
            spec = lbl.JobSpec.from_job(job, lbl.FP32K)
            lo, _, _, _ = ctx.windows(spec)
            rec = torch.empty((len(job.atm), lo.shape[0], lbl.NT), dtype=torch.float32, device="cuda")
            ctx.pt_tables_device(spec, rec.data_ptr(), 0)  # the context's stream, as the adjoint's
            g = torch.randn(rec.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
            out.append(ctx.pt_tables_adjoint(spec, g.data_ptr()))
        finally:
            ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
