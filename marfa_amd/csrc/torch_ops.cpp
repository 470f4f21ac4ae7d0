// torch_ops.cpp -- PyTorch operators (torch.ops.marfa.*) over the C ABI of include/marfa.h.
//
// The Python side of the reference reads PT-table records from disk into numpy arrays
// (handlers/parsers.py:17-65, handlers/convert.py:12-53); here the records are produced
// and stay on the GPU as torch tensors, on the caller's current HIP stream, and the adjoint
// returns its gradients as device tensors.  A context is created and loaded through the C
// ABI (marfa_amd/lbl.py) and passed to the operators as its handle (an int).  These
// functions only allocate tensors, pick the stream and forward to the library; every
// computation is in libmarfa.so.
#include <ATen/ATen.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <string>
#include <tuple>

#include "../../include/marfa.h"

namespace {

marfa_ctx *ctx_of(int64_t h) {
    TORCH_CHECK(h != 0, "marfa: null context");
    return reinterpret_cast<marfa_ctx *>(static_cast<intptr_t>(h));
}

void check(marfa_ctx *c, int rc, const char *what) {
    TORCH_CHECK(rc == 0, "marfa: ", what, " failed with code ", rc, ": ", marfa_last_error(c));
}

marfa_job make_job(double start_wv, double end_wv, double cutoff, int64_t shape_id, int64_t chi_id,
                   int64_t mol_code, int64_t target, int64_t precision) {
    marfa_job j;
    j.start_wv = start_wv;
    j.end_wv = end_wv;
    j.cutoff = (float)cutoff;
    j.shape_id = (int32_t)shape_id;
    j.chi_id = (int32_t)chi_id;
    j.mol_code = (int32_t)mol_code;
    j.target = (int32_t)target;
    j.precision = (int32_t)precision;
    return j;
}

struct Info {
    int64_t n;
    int32_t nlev, device;
};

Info info(marfa_ctx *c) {
    Info i{};
    check(c, marfa_ctx_info(c, &i.n, &i.nlev, &i.device), "marfa_ctx_info");
    return i;
}

void check_lines(const at::Tensor &t, const Info &i, const char *name) {
    TORCH_CHECK(t.is_cuda() && t.device().index() == i.device, "marfa: ", name, " must be on the context's GPU");
    TORCH_CHECK(t.scalar_type() == at::kFloat && t.dim() == 1 && t.numel() == i.n && t.is_contiguous(),
                "marfa: ", name, " must be a contiguous float32 vector of the context's ", i.n, " lines");
}

// records [nlev][nrec][20481] (float32) of the loaded job, enqueued on the current stream
at::Tensor pt_tables(int64_t h, double start_wv, double end_wv, double cutoff, int64_t shape_id, int64_t chi_id,
                     int64_t mol_code, int64_t target, int64_t precision) {
    marfa_ctx *c = ctx_of(h);
    const Info i = info(c);
    const marfa_job j = make_job(start_wv, end_wv, cutoff, shape_id, chi_id, mol_code, target, precision);
    int32_t nrec = 0;
    check(c, marfa_windows(c, &j, 0, &nrec, nullptr, nullptr, nullptr, nullptr), "marfa_windows");
    at::Tensor out = at::empty({i.nlev, nrec, MARFA_NT}, at::TensorOptions().dtype(at::kFloat).device(at::kCUDA, i.device));
    const hipStream_t st = c10::hip::getCurrentHIPStream(i.device).stream();
    check(c, marfa_pt_tables(c, &j, out.data_ptr<float>(), nullptr, st, 1), "marfa_pt_tables");
    return out;
}

// dL/d(S, gamma_air, n_T, delta) [4][n] and dL/d(T, P) [2][nlev] (float64, device) from
// dL/d(records), enqueued on the current stream
std::tuple<at::Tensor, at::Tensor> pt_tables_adjoint(int64_t h, double start_wv, double end_wv, double cutoff,
                                                     int64_t shape_id, int64_t chi_id, int64_t mol_code,
                                                     int64_t target, int64_t precision, const at::Tensor &grad) {
    marfa_ctx *c = ctx_of(h);
    const Info i = info(c);
    const marfa_job j = make_job(start_wv, end_wv, cutoff, shape_id, chi_id, mol_code, target, precision);
    int32_t nrec = 0;
    check(c, marfa_windows(c, &j, 0, &nrec, nullptr, nullptr, nullptr, nullptr), "marfa_windows");
    TORCH_CHECK(grad.is_cuda() && grad.device().index() == i.device, "marfa: grad must be on the context's GPU");
    TORCH_CHECK(grad.dim() == 3 && grad.size(0) == i.nlev && grad.size(1) == nrec && grad.size(2) == MARFA_NT,
                "marfa: grad must have the records' shape [nlev, nrec, 20481]");
    const at::Tensor g = grad.to(at::kFloat).contiguous();
    const auto opt = at::TensorOptions().dtype(at::kDouble).device(at::kCUDA, i.device);
    at::Tensor gl = at::empty({4, i.n}, opt), gv = at::empty({2, i.nlev}, opt);
    const hipStream_t st = c10::hip::getCurrentHIPStream(i.device).stream();
    check(c, marfa_pt_tables_adjoint_dev(c, &j, g.data_ptr<float>(), st, gl.data_ptr<double>(), gv.data_ptr<double>()),
          "marfa_pt_tables_adjoint_dev");
    return {gl, gv};
}

// per-line S, gamma_air, n_T, delta from device tensors (device-to-device, current stream)
void set_line_params(int64_t h, const at::Tensor &S, const at::Tensor &gair, const at::Tensor &nT,
                     const at::Tensor &delta) {
    marfa_ctx *c = ctx_of(h);
    const Info i = info(c);
    check_lines(S, i, "S");
    check_lines(gair, i, "gamma_air");
    check_lines(nT, i, "n_T");
    check_lines(delta, i, "delta");
    const hipStream_t st = c10::hip::getCurrentHIPStream(i.device).stream();
    check(c, marfa_update_line_params(c, S.data_ptr<float>(), gair.data_ptr<float>(), nT.data_ptr<float>(),
                                      delta.data_ptr<float>(), st),
          "marfa_update_line_params");
}

// level state P, T, density (a few hundred values at most: copied through the host, where
// the library keeps the levels)
void set_levels(int64_t h, const at::Tensor &P, const at::Tensor &T, const at::Tensor &rho) {
    marfa_ctx *c = ctx_of(h);
    TORCH_CHECK(P.numel() == T.numel() && P.numel() == rho.numel(), "marfa: P, T and rho must have one value per level");
    const at::Tensor p = P.to(at::kCPU, at::kFloat).contiguous(), t = T.to(at::kCPU, at::kFloat).contiguous(),
                     r = rho.to(at::kCPU, at::kFloat).contiguous();
    check(c, marfa_set_levels(c, (int32_t)p.numel(), p.data_ptr<float>(), t.data_ptr<float>(), r.data_ptr<float>()),
          "marfa_set_levels");
}

// the spectral loss sum w (ln k - ln khat)^2 (a 0-dim float64 tensor) and its gradient
// dL/dk times gscale (float32, k's shape), on the current stream (marfa_log_loss[_grad])
void check_loss_args(const at::Tensor &k, const at::Tensor &kh, const c10::optional<at::Tensor> &w) {
    TORCH_CHECK(k.is_cuda() && (k.scalar_type() == at::kFloat || k.scalar_type() == at::kDouble) && k.is_contiguous(),
                "marfa: records must be a contiguous float32 or float64 CUDA tensor");
    TORCH_CHECK(kh.device() == k.device() && kh.scalar_type() == k.scalar_type() && kh.is_contiguous() &&
                    kh.numel() == k.numel(),
                "marfa: target must be a contiguous tensor of the records' size, dtype and GPU");
    if (w) {
        TORCH_CHECK(w->device() == k.device() && w->scalar_type() == at::kFloat && w->is_contiguous() &&
                        w->numel() == k.numel(),
                    "marfa: weights must be a contiguous float32 tensor of the records' size on their GPU");
    }
}

at::Tensor log_loss(const at::Tensor &k, const at::Tensor &kh, const c10::optional<at::Tensor> &w, double wscale) {
    check_loss_args(k, kh, w);
    const auto opt = at::TensorOptions().dtype(at::kDouble).device(k.device());
    at::Tensor part = at::empty({MARFA_LOSS_PARTIALS}, opt), loss = at::empty({}, opt);
    const hipStream_t st = c10::hip::getCurrentHIPStream(k.device().index()).stream();
    const float *wp = w ? w->data_ptr<float>() : nullptr;
    const int rc = k.scalar_type() == at::kFloat
                       ? marfa_log_loss(k.numel(), k.data_ptr<float>(), kh.data_ptr<float>(), wp, wscale,
                                        part.data_ptr<double>(), loss.data_ptr<double>(), st)
                       : marfa_log_loss_f64(k.numel(), k.data_ptr<double>(), kh.data_ptr<double>(), wp, wscale,
                                            part.data_ptr<double>(), loss.data_ptr<double>(), st);
    TORCH_CHECK(rc == 0, "marfa: marfa_log_loss failed with code ", rc);
    return loss;
}

at::Tensor log_loss_grad(const at::Tensor &k, const at::Tensor &kh, const c10::optional<at::Tensor> &w, double wscale,
                         const at::Tensor &gscale) {
    check_loss_args(k, kh, w);
    const at::Tensor g = gscale.to(k.device(), at::kDouble).reshape({1}).contiguous();
    at::Tensor grad = at::empty_like(k);
    const hipStream_t st = c10::hip::getCurrentHIPStream(k.device().index()).stream();
    const float *wp = w ? w->data_ptr<float>() : nullptr;
    const int rc = k.scalar_type() == at::kFloat
                       ? marfa_log_loss_grad(k.numel(), k.data_ptr<float>(), kh.data_ptr<float>(), wp, wscale,
                                             g.data_ptr<double>(), grad.data_ptr<float>(), st)
                       : marfa_log_loss_grad_f64(k.numel(), k.data_ptr<double>(), kh.data_ptr<double>(), wp, wscale,
                                                 g.data_ptr<double>(), grad.data_ptr<double>(), st);
    TORCH_CHECK(rc == 0, "marfa: marfa_log_loss_grad failed with code ", rc);
    return grad;
}

}  // namespace

TORCH_LIBRARY(marfa, m) {
    m.def("pt_tables(int ctx, float start_wv, float end_wv, float cutoff, int shape_id, int chi_id, int mol_code, "
          "int target, int precision) -> Tensor",
          &pt_tables);
    m.def("pt_tables_adjoint(int ctx, float start_wv, float end_wv, float cutoff, int shape_id, int chi_id, "
          "int mol_code, int target, int precision, Tensor grad) -> (Tensor, Tensor)",
          &pt_tables_adjoint);
    m.def("set_line_params(int ctx, Tensor S, Tensor gair, Tensor nT, Tensor delta) -> ()", &set_line_params);
    m.def("set_levels(int ctx, Tensor P, Tensor T, Tensor rho) -> ()", &set_levels);
    m.def("log_loss(Tensor k, Tensor khat, Tensor? w, float wscale) -> Tensor", &log_loss);
    m.def("log_loss_grad(Tensor k, Tensor khat, Tensor? w, float wscale, Tensor gscale) -> Tensor", &log_loss_grad);
}
