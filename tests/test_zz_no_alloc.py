"""marfa_pt_tables_noalloc: marfa_pt_tables with its scratch in the library's fixed array
(no device allocation), called through the C ABI and checked bit for bit against
marfa_pt_tables on the same context, two calls in a row (the second smaller, so stale
scratch from the first would show), in both precision forms; and a job whose one level
needs more than the fixed scratch is refused with MARFA_E_ARG before any device work."""
import numpy as np
import pytest
import torch

from marfa_amd import lbl, synth

pytestmark = pytest.mark.gpu
TIPS = synth.load_tips()


def job_of(n, lo, hi, a, b, nlev, seed, chi=0):
    return synth.Job(synth.uniform_lines(n, lo, hi, seed), synth.log_levels(nlev, 90.0, 1e-3, 735.0, 170.0, 0.965),
                     a, b, 25.0, chi_id=chi)


@pytest.mark.parametrize("prec", [1, 0])
def test_noalloc_matches_pt_tables(prec):
    # first call: 600,000 lines x 10 levels -- 96 MB (FP64) / 77 MB (FP32K) of scratch per
    # level, so the fixed scratch runs the levels in several batches; second call: smaller
    jobs = [job_of(600_000, 3970.0, 4130.0, 4000.0, 4100.0, 10, 41, chi=2),
            job_of(3_000, 3985.0, 4045.0, 4000.0, 4030.0, 3, 42)]
    ctx = lbl.LBLContext(0)
    try:
        for k, job in enumerate(jobs):
            print(f"noalloc call {k}: {len(job.lines)} lines, {len(job.atm)} levels, precision {prec}", flush=True)
            ctx.load_job(job, TIPS)
            spec = lbl.JobSpec.from_job(job, prec)
            want = ctx.pt_tables(spec)
            got = torch.full(want.shape, float("nan"), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            ctx.pt_tables_noalloc(spec, got.data_ptr())
            torch.cuda.synchronize()  # the call is asynchronous on the context's stream
            out = got.cpu().numpy()
            assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), (k, prec)
    finally:
        ctx.close()


def test_noalloc_refuses_a_job_beyond_the_fixed_scratch():
    # 2.2 million lines in one window: 2.2e6 x 128 B > 256 MiB for a single level
    job = job_of(2_200_000, 3990.0, 4030.0, 4000.0, 4010.0, 1, 43)
    ctx = lbl.LBLContext(0)
    try:
        ctx.load_job(job, TIPS)
        spec = lbl.JobSpec.from_job(job, lbl.FP32K)
        out = torch.zeros((1, 1, lbl.NT), dtype=torch.float32, device="cuda")
        with pytest.raises(lbl.MarfaError) as e:
            ctx.pt_tables_noalloc(spec, out.data_ptr())
        assert e.value.code == 20 and "bytes of scratch" in str(e.value)
        torch.cuda.synchronize()
        assert not out.any()
    finally:
        ctx.close()
