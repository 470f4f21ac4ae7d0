"""torch.ops.marfa (marfa_amd/csrc/torch_ops.cpp): the PyTorch operators over the C ABI.

The operators return device tensors on the caller's current stream; they must give the
same bits as the C ABI called through ctypes, and the autograd Function built on them must
give every applied node the gradient of its own parameters."""
import numpy as np
import pytest
import torch

from marfa_amd import lbl, ops, synth

TIPS = synth.load_tips()


def test_operators_registered():
    """The operator library loads (no GPU needed) and registers its four operators."""
    m = ops.load()
    for name in ("pt_tables", "pt_tables_adjoint", "set_line_params", "set_levels"):
        assert hasattr(m, name), name
    s = str(torch.ops.marfa.pt_tables.default._schema)
    assert "int ctx" in s and "Tensor" in s


def small_job(n=1500, nlev=2, chi=0):
    return synth.Job(synth.uniform_lines(n, 3985.0, 4045.0, 31), synth.log_levels(nlev, 10.0, 1e-3, 300.0, 200.0, 0.965),
                     4000.0, 4020.0, 25.0, chi_id=chi)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_pt_tables_op_matches_c_abi(prec):
    job = small_job(chi=2)
    ctx = lbl.LBLContext(0)
    try:
        ctx.load_job(job, TIPS)
        spec = lbl.JobSpec.from_job(job, prec)
        want = ctx.pt_tables(spec)
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):   # the operator runs on the current stream
            got = ops.pt_tables(ctx, spec)
        side.synchronize()
        assert got.is_cuda and got.dtype == torch.float32 and tuple(got.shape) == want.shape
        assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))
    finally:
        ctx.close()


@pytest.mark.gpu
def test_adjoint_op_matches_c_abi():
    job = small_job()
    ctx = lbl.LBLContext(0)
    try:
        ctx.load_job(job, TIPS)
        spec = lbl.JobSpec.from_job(job, lbl.FP32K)
        rec = ops.pt_tables(ctx, spec)
        g = torch.randn(rec.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
        gl, gv = ops.pt_tables_adjoint(ctx, spec, g)
        torch.cuda.synchronize()
        hl, hv = ctx.pt_tables_adjoint(spec, g.data_ptr())
        assert gl.dtype == torch.float64 and tuple(gl.shape) == (4, len(job.lines)) and tuple(gv.shape) == (2, 2)
        assert np.array_equal(gl.cpu().numpy(), hl) and np.array_equal(gv.cpu().numpy(), hv)
        assert np.abs(hl).max() > 0 and np.abs(hv).max() > 0
    finally:
        ctx.close()


@pytest.mark.gpu
def test_device_line_update_matches_reload():
    """set_line_params (device-to-device) gives the records of a full set_lines reload."""
    job = small_job()
    rng = np.random.default_rng(5)
    ctx = lbl.LBLContext(0)
    try:
        ctx.load_job(job, TIPS)
        spec = lbl.JobSpec.from_job(job, lbl.FP32K)
        L = job.lines
        new = [(np.asarray(getattr(L, f)) * rng.uniform(0.9, 1.1, len(L))).astype(np.float32)
               for f in ("S", "gair", "nT", "delta")]
        ops.set_line_params(ctx, *[torch.tensor(x, device="cuda") for x in new])
        got = ops.pt_tables(ctx, spec).cpu().numpy()
        ctx.set_lines(synth.Lines(nu=L.nu, S=new[0], gair=new[1], gself=L.gself, elow=L.elow, nT=new[2],
                                  delta=new[3], molIso=L.molIso))
        want = ctx.pt_tables(spec)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    finally:
        ctx.close()


@pytest.mark.gpu
def test_two_applies_of_one_model():
    """Two PTTables nodes of one model in one loss: each gets the gradient of its own
    parameters (the model is reloaded for the node whose parameters were overwritten)."""
    from marfa_amd.autograd import PTTableModel, PTTables, spectral_loss
    model = PTTableModel(small_job(), TIPS)
    try:
        base = model.params()
        other = [b.clone() for b in base]
        other[0] = other[0] * 1.3
        other[1] = other[1] * 0.9
        other[4] = other[4] + 7.0
        target = model.records() * 1.05

        def grads(*sets):
            leaves = [[t.clone().requires_grad_(True) for t in ps] for ps in sets]
            loss = sum(spectral_loss(PTTables.apply(*lv, model), target) for lv in leaves)
            loss.backward()
            return [[t.grad.clone() for t in lv] for lv in leaves]

        g_base, = grads(base)
        g_other, = grads(other)
        g_both = grads(base, other)
        for got, want in ((g_both[0], g_base), (g_both[1], g_other)):
            for a, b in zip(got, want):
                assert torch.equal(a, b)
        assert not torch.equal(g_base[0], g_other[0])
    finally:
        model.close()
