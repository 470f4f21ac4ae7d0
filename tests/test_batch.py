"""Batched jobs in one hipGraph (marfa_batch_*): every replay reproduces, bit for bit, what
each job gives on its own through marfa_pt_tables, whatever variant the autotuner picked,
and agrees with the FP64 oracle on sampled records."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from marfa_amd import synth

pytestmark = pytest.mark.gpu
TIPS = synth.load_tips()


def uneven_jobs(seed=synth.SEED_BASE + 3):
    """Six small jobs with the configuration-3 laws: line counts, level counts, window widths,
    molar masses and cutoffs all different."""
    rng = np.random.Generator(np.random.PCG64(seed))
    specs = [(800, 2, 20.0, 44, 25.0), (6000, 3, 60.0, 18, 125.0), (3000, 4, 30.0, 16, 250.0),
             (20000, 2, 40.0, 44, 125.0), (1500, 3, 50.0, 28, 25.0), (9000, 2, 20.0, 48, 250.0)]
    jobs = []
    for j, (n, nlev, w, mass, cut) in enumerate(specs):
        a = 10.0 * np.floor(rng.uniform(500, 9000) / 10.0)
        molIso, mol = synth.MOLIISO_FOR_MASS[mass]
        sub = np.random.Generator(np.random.PCG64(seed * 1000 + j))
        lines = synth.make_lines(sub.uniform(a - cut, a + w + cut, n), sub, molIso=molIso)
        atm = synth.log_levels(nlev, 90.0, 1e-4, 735.0, 170.0, 0.965)
        jobs.append(synth.Job(lines, atm, a, a + w, cut, mol_code=mol, chi_id=2 if mol == 2 else 0))
    return jobs


def hip_memset(ptr, nbytes):
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    assert hip.hipMemset(C.c_void_p(ptr), 0, nbytes) == 0
    hip.hipDeviceSynchronize()


def test_batch_graph_matches_single_jobs():
    from marfa_amd import lbl
    jobs = uneven_jobs()
    ctx = lbl.LBLContext(0)
    single = []
    try:
        for job in jobs:
            ctx.load_job(job, TIPS)
            single.append(ctx.pt_tables(lbl.JobSpec.from_job(job, lbl.FP32K)))
    finally:
        ctx.close()
    b = lbl.Batch(0, lanes=3)
    try:
        b.set_tips(TIPS)
        ids = [b.add_job(job) for job in jobs]
        variants, tms = b.tune()
        print("variants", variants.tolist(), flush=True)
        assert set(variants.tolist()) <= set(range(lbl.NVARIANTS))
        for i in ids:
            assert np.array_equal(b.records(i), single[i]), f"eager run of job {i}"
        b.capture()
        for rep in range(3):
            # wipe the records, then one replay must rebuild all of them
            for i in ids:
                nlev, nrec = C.c_int32(), C.c_int32()
                ptr = C.c_void_p()
                b._check(b.lib.marfa_batch_info(b.h, i, C.byref(nlev), C.byref(nrec), C.byref(ptr), None))
                hip_memset(ptr.value, nlev.value * nrec.value * lbl.NT * 4)
            ms = b.launch()
            print(f"replay {rep}: {ms:.3f} ms", flush=True)
            for i in ids:
                assert np.array_equal(b.records(i), single[i]), f"replay {rep}, job {i}"
    finally:
        b.close()


def test_batch_jobs_against_oracle():
    from marfa_amd import lbl
    jobs = uneven_jobs()
    b = lbl.Batch(0, lanes=2)
    try:
        b.set_tips(TIPS)
        ids = [b.add_job(job) for job in jobs]
        b.tune()
        b.capture()
        b.launch()
        rng = np.random.default_rng(0)
        for i, job in zip(ids, jobs):
            rec = b.records(i)
            lev, r = int(rng.integers(len(job.atm))), int(rng.integers(rec.shape[1]))
            ref, _ = po.OracleJob(job, TIPS).record(po.F64, lev, r)
            m = np.abs(ref) > 1e-30 * np.abs(ref).max()
            err = np.max(np.abs(rec[lev, r][m] - ref[m]) / np.abs(ref[m]))
            print(f"job {i}: level {lev} record {r} max rel err {err:.2e}", flush=True)
            assert err < 1e-4
    finally:
        b.close()


def test_batch_waves_and_pieces_match_single_jobs():
    """A lane budget of 1/4 MiB cuts the jobs into one-level pieces and the arena (12 lane
    budgets) into several waves, so the graph prepares waves beside the previous wave's
    deposits and runs several work lists per kernel variant: the records still equal the
    jobs run one by one."""
    from marfa_amd import lbl
    jobs = uneven_jobs()[:4]
    ctx = lbl.LBLContext(0)
    single = []
    try:
        for job in jobs:
            ctx.load_job(job, TIPS)
            single.append(ctx.pt_tables(lbl.JobSpec.from_job(job, lbl.FP32K)))
    finally:
        ctx.close()
    b = lbl.Batch(0, lanes=3, lane_budget_gb=1.0 / 4096)
    try:
        b.set_tips(TIPS)
        ids = [b.add_job(job) for job in jobs]
        b.tune()
        b.capture()
        for rep in range(2):
            for i in ids:
                nlev, nrec = C.c_int32(), C.c_int32()
                ptr = C.c_void_p()
                b._check(b.lib.marfa_batch_info(b.h, i, C.byref(nlev), C.byref(nrec), C.byref(ptr), None))
                hip_memset(ptr.value, nlev.value * nrec.value * lbl.NT * 4)
            ms = b.launch()
            print(f"replay {rep}: {ms:.3f} ms", flush=True)
            for i in ids:
                assert np.array_equal(b.records(i), single[i]), f"replay {rep}, job {i}"
    finally:
        b.close()
