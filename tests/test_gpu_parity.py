"""GPU parity: the HIP path through the C ABI against the CPU oracle.

FP64 form: grid sums within 1e-11 relative of the FP64 oracle (same selection, different
summation order), records within 1e-6 (their FP32 rounding).  FP32K form: grid sums and
records within 1e-4 relative of the FP64 oracle.  Selection is checked node by node in
both forms: the set of grid nodes that receive any sample is identical.
"""
import numpy as np
import pytest

import json
import os
import sys

import pyoracle as po
from marfa_amd import synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import gen_golden  # noqa: E402  (case definitions only; nothing here runs the reference)

GOLDEN = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "records.npz"))
GOLDEN_META = json.loads(str(GOLDEN["meta"]))

pytestmark = pytest.mark.gpu

NGRID = 30690
LEVEL_OFF = [30 * ((1 << k) - 1) for k in range(11)]


@pytest.fixture(scope="module")
def tips():
    return synth.load_tips()


@pytest.fixture(scope="module")
def ctx():
    from marfa_amd import lbl
    c = lbl.LBLContext(0)
    yield c
    c.close()


def small_job(seed=7, n=2000, lo=3985.0, hi=4045.0, a=4000.0, b=4030.0, cut=25.0, nlev=3,
              chi=0, target=0, shape=0, p0=90.0, p1=1e-3):
    return synth.Job(synth.uniform_lines(n, lo, hi, seed), synth.log_levels(nlev, p0, p1, 735.0, 170.0, 0.965),
                     a, b, cut, chi_id=chi, target=target, shape_id=shape)


def rel_err(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), 1e-300)
    return np.abs(x - ref) / den


def describe_grid_diff(g, ref, tol):
    bad = np.nonzero(rel_err(g, ref) > tol)[0]
    out = []
    for e in bad[:12]:
        k = max(i for i in range(10) if LEVEL_OFF[i] <= e)
        o = e - LEVEL_OFF[k]
        ntk = 10 << k
        out.append(f"level {k} role {'PML'[o // ntk]} cell {o % ntk + 1}: gpu {g[e]:.6e} ref {ref[e]:.6e}")
    return f"{bad.size} nodes differ; " + "; ".join(out)


GRID_CASES = {
    "plain": dict(),
    "wide": dict(seed=9, n=6000, lo=3960.0, hi=4080.0, cut=25.0),
    "chi2": dict(seed=10, n=3000, lo=3900.0, hi=4150.0, cut=125.0, chi=2),
    "lowp": dict(seed=11, n=2500, p0=1e-2, p1=1e-4),
    "dense": dict(seed=12, n=45000, lo=3995.0, hi=4025.0, a=4000.0, b=4020.0, cut=5.0, nlev=2),
    # cutoff below the largest pressure shift (|delta P| up to 0.9 cm-1 at 90 atm) and more
    # centre lines per tile than the LDS stage holds: lines outside a record's window lie
    # between window lines in shifted order in the global-memory lane passes
    "shifted": dict(seed=14, n=45000, lo=3995.0, hi=4025.0, a=4000.0, b=4020.0, cut=0.5, nlev=2),
    # far wings in every Tonkov band (3-150, 150-300, beyond 300 cm-1) and across its band
    # edges; Perrin's temperature-dependent bands (edges 30 and 120 cm-1) per node
    "chi1far": dict(seed=15, n=5000, lo=3670.0, hi=4360.0, a=4000.0, b=4020.0, cut=330.0, nlev=2, chi=1),
    "chi3far": dict(seed=16, n=3000, lo=3840.0, hi=4190.0, a=4000.0, b=4020.0, cut=160.0, nlev=2, chi=3),
}


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cfg", sorted(GRID_CASES))
def test_grid_sums(ctx, tips, cfg, prec):
    """Every node of every grid receives samples exactly where the FP64 oracle's does (the
    selection is bit-exact in both forms), with values within 1e-11 (FP64 form: same
    selection, another summation order) or 1e-4 (FP32K form: FP32 samples).  'dense' and
    'shifted' (1,500 lines per cm-1) put more centre lines in a tile's reach than the LDS
    stage holds, so their tiles take the global-memory gathers on every grid."""
    from marfa_amd import lbl
    job = small_job(**GRID_CASES[cfg])
    ctx.load_job(job, tips)
    oj = po.OracleJob(job, tips)
    spec = lbl.JobSpec.from_job(job, prec)
    tol, rtol = (1e-11, 1e-6) if prec == lbl.FP64 else (1e-4, 1e-4)
    for lev in range(len(job.atm)):
        for r in range(oj.nrec):
            print(f"case {cfg} precision {prec} level {lev} record {r}", flush=True)
            rk_ref, g_ref, _ = oj.record(po.F64, lev, r, grids=True)
            g, rkd, rec = ctx.grid_sums(spec, lev, r)
            # selection: the same nodes receive samples
            assert np.array_equal(g != 0, g_ref[:NGRID] != 0), describe_grid_diff(g, g_ref[:NGRID], 0)
            assert np.array_equal(rkd != 0, g_ref[NGRID:] != 0), "fine-point selection differs"
            assert rel_err(g, g_ref[:NGRID]).max() < tol, describe_grid_diff(g, g_ref[:NGRID], tol)
            assert rel_err(rkd, g_ref[NGRID:]).max() < tol
            assert rel_err(rec, rk_ref).max() < rtol  # record is FP32


def check_records(out, ref, tol, what):
    m = np.abs(ref) > np.finfo(np.float32).tiny
    err = rel_err(out[m], ref[m])
    print(f"{what}: max rel {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} over {m.sum()} points", flush=True)
    assert err.max() < tol, (what, err.max())


@pytest.mark.parametrize("case", sorted(gen_golden.record_cases()))
def test_golden_cases(ctx, tips, case):
    """The reference's own record cases (tests/golden/records.npz: tiny, carry, the Tonkov,
    Pollack (VAC) and Perrin chi factors, lorentzShape, the branch boundaries, low pressure,
    H2O) through both GPU forms: against the FP64 oracle (FP64 form 1e-6 = the FP32
    rounding of the records, FP32K 1e-4), and against the reference's FP32 records (every
    16th point) within 2e-4, the bound the FP64 oracle itself meets there."""
    from marfa_amd import lbl
    meta = GOLDEN_META[case]
    job = gen_golden.make_job(meta["case"])
    assert gen_golden.lines_sha(job.lines) == meta["lines_sha"], "synthetic generator drifted"
    ctx.load_job(job, tips)
    ref, _ = po.OracleJob(job, tips).pt_tables(po.F64)
    refd = GOLDEN[f"{case}__strided"].astype(np.float64)
    for prec, tol in ((lbl.FP64, 1e-6), (lbl.FP32K, 1e-4)):
        out = ctx.pt_tables(lbl.JobSpec.from_job(job, prec)).astype(np.float64)
        check_records(out, ref, tol, f"{case} precision {prec} vs FP64 oracle")
        check_records(out[:, :, ::gen_golden.STRIDE], refd, 2e-4, f"{case} precision {prec} vs reference")


@pytest.mark.parametrize("cfg", ["config2_window", "config5_window"])
def test_reduced_configs(ctx, tips, cfg):
    """Configuration 2 (banded lines around the densest 100 cm-1, 400,000 lines, 6 levels to
    1e-4 atm, cutoff 125, Pollack) and configuration 5 (H2O laws: S down to 1e-32, E'' to
    8,000, P to 1e-5 atm, cutoff 250, on 10000-10100 cm-1) at sizes the oracle runs in
    seconds: every record of every level in both forms against the FP64 oracle.  The FP32K
    form's power-of-two sample scale is what keeps configuration 5's weak lines in range."""
    from marfa_amd import lbl
    job = getattr(synth, cfg)()
    ctx.load_job(job, tips)
    oj = po.OracleJob(job, tips)
    ref, cnt = oj.pt_tables(po.F64)
    for prec, tol in ((lbl.FP64, 1e-6), (lbl.FP32K, 1e-4)):
        spec = lbl.JobSpec.from_job(job, prec)
        out = ctx.pt_tables(spec).astype(np.float64)
        check_records(out, ref, tol, f"{cfg} precision {prec}")
        assert ctx.count_samples(spec) == cnt.samples


def test_empty_windows(ctx, tips):
    """A range above every line: every window is empty, so the records are zero, no sample
    is counted and the gradients are zero -- and no line beyond the list is read."""
    import torch
    from marfa_amd import lbl
    job = small_job(a=5000.0, b=5030.0)
    ctx.load_job(job, tips)
    lo, hi, _, _ = ctx.windows(lbl.JobSpec.from_job(job))
    assert np.all(lo == len(job.lines)) and np.all(hi == lo)
    for prec in (lbl.FP64, lbl.FP32K):
        spec = lbl.JobSpec.from_job(job, prec)
        out = ctx.pt_tables(spec)
        assert out.shape == (len(job.atm), 3, lbl.NT) and not out.any()
        assert ctx.count_samples(spec) == 0
        g = torch.ones(out.shape, dtype=torch.float32, device="cuda")
        gl, gv = ctx.pt_tables_adjoint(spec, g.data_ptr())
        assert not gl.any() and not gv.any()


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-4)])
def test_pt_tables_config1(ctx, tips, prec, tol):
    from marfa_amd import lbl
    job = synth.config1()
    ctx.load_job(job, tips)
    oj = po.OracleJob(job, tips)
    ref, cnt = oj.pt_tables(po.F64)
    out = ctx.pt_tables(lbl.JobSpec.from_job(job, prec))
    err = rel_err(out, ref)
    print(f"config1 precision {prec}: max rel {err.max():.3e} rms {np.sqrt((err**2).mean()):.3e}", flush=True)
    assert err.max() < tol
    ns = ctx.count_samples(lbl.JobSpec.from_job(job, prec))
    assert ns == cnt.samples, (ns, cnt.samples)


def test_determinism(ctx, tips):
    from marfa_amd import lbl
    job = small_job(seed=21)
    ctx.load_job(job, tips)
    for prec in (lbl.FP64, lbl.FP32K):
        a = ctx.pt_tables(lbl.JobSpec.from_job(job, prec))
        b = ctx.pt_tables(lbl.JobSpec.from_job(job, prec))
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_level_and_window_changes_between_runs(ctx, tips):
    """A context keeps the level and window arrays of its last run on the device and copies
    them again only when they change: runs whose levels, windows or both differ from the
    previous run's (with the same array sizes) equal the same runs in fresh contexts."""
    from marfa_amd import lbl
    base = small_job(seed=23)
    other_lev = small_job(seed=23, p0=50.0, p1=2e-3)
    other_win = small_job(seed=23, a=4001.0, b=4031.0)
    seq = [base, other_lev, other_win, base, base]
    got = []
    ctx.load_job(base, tips)
    for job in seq:
        ctx.set_levels(job.atm)
        got.append(ctx.pt_tables(lbl.JobSpec.from_job(job, lbl.FP32K)))
    for i, job in enumerate(seq[:3]):
        fresh = lbl.LBLContext(0)
        try:
            fresh.load_job(job, tips)
            ref = fresh.pt_tables(lbl.JobSpec.from_job(job, lbl.FP32K))
        finally:
            fresh.close()
        assert np.array_equal(got[i].view(np.uint32), ref.view(np.uint32)), f"run {i}"
    assert np.array_equal(got[3].view(np.uint32), got[0].view(np.uint32))
    assert np.array_equal(got[4].view(np.uint32), got[0].view(np.uint32))


@pytest.mark.parametrize("case", ["aligned", "offset", "start_below_cutoff", "wide_cut", "past_lines"])
def test_windows_match_oracle(ctx, tips, case):
    """Subinterval windows and record numbers (app/main.f90:116-144, :196-248)."""
    from marfa_amd import lbl
    job = {"aligned": small_job(),
           "offset": small_job(a=4003.5, b=4041.25),
           "start_below_cutoff": small_job(lo=5.0, hi=80.0, a=10.0, b=60.0, cut=25.0),
           "wide_cut": small_job(lo=3700.0, hi=4400.0, a=4000.0, b=4100.0, cut=250.0),
           "past_lines": small_job(lo=3985.0, hi=4020.0, a=4000.0, b=4060.0)}[case]
    ctx.load_job(job, tips)
    oj = po.OracleJob(job, tips)
    lo, hi, rec, org = ctx.windows(lbl.JobSpec.from_job(job, lbl.FP64))
    olo, ohi, orec, oorg = oj.windows()
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    assert np.array_equal(rec, orec) and np.array_equal(org, oorg)


def test_lbl_scheme_drop_in(ctx, tips):
    """marfa_lbl_scheme called the way app/main.f90:116-160 calls lblScheme: the carried
    line index and each subinterval's record match the windows and the FP64 oracle."""
    from marfa_amd import lbl
    job = small_job(seed=13, n=2500, lo=3960.0, hi=4070.0, a=4000.0, b=4040.0)
    ctx.load_job(job, tips)
    oj = po.OracleJob(job, tips)
    ref, _ = oj.pt_tables(po.F64)
    lo, hi, _, _ = ctx.windows(lbl.JobSpec.from_job(job, lbl.FP64))
    nu = np.asarray(job.lines.nu, np.float64)
    cut = np.float32(job.cutoff)
    start = 1
    if job.start_wv > cut:
        start = int(np.searchsorted(nu, job.start_wv - cut, side="right")) + 1
    start_line_wv = nu[start - 1]
    pself = synth.p_self(job.atm)
    for lev in range(len(job.atm)):
        idx, a, r = start, float(job.start_wv), 0
        while a < job.end_wv:
            print(f"lbl_scheme level {lev} subinterval {r}", flush=True)
            assert idx - 1 == lo[r]
            ext_start = a - float(cut)
            cap = max(ext_start + 10.0, start_line_wv)
            idx, rk = ctx.lbl_scheme(idx, cap, a, a + 10.0 + float(cut), float(cut),
                                     float(job.atm.P[lev]), float(job.atm.T[lev]), float(pself[lev]))
            assert rel_err(rk, ref[lev, r]).max() < 1e-6
            a, r = a + 10.0, r + 1
