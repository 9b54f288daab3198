// torch::autograd::Function over the differentiable step of libhns (LibTorch-ROCm, host code only:
// the forward and backward passes are the library's own HIP kernels, called through the C ABI of
// include/hns.h on the caller's stream). No reference counterpart; the reference is forward only.
#include <torch/extension.h>

#include <string>
#include <vector>

#include "../../include/hns.h"

namespace {

void check(hns_status st) {
  if (st != HNS_OK) throw std::runtime_error(std::string("libhns: ") + hns_last_error());
}

void check_tensor(const torch::Tensor& t, const char* what) {
  TORCH_CHECK(t.defined() && !t.is_cpu() && t.is_contiguous() && t.scalar_type() == torch::kFloat32, what,
              " must be a contiguous float32 device tensor");
}

// Owns a tape (checkpoints in HBM) for the lifetime of an autograd node: freed by backward, or by
// the destructor when the node goes without one (outputs dropped before backward).
struct TapeHolder : torch::CustomClassHolder {
  hns_diff_tape* tape = nullptr;
  explicit TapeHolder(hns_diff_tape* t) : tape(t) {}
  ~TapeHolder() override {
    if (tape) (void)hns_diff_tape_destroy(&tape);
  }
};
const auto kTapeHolderClass = torch::class_<TapeHolder>("hns_torch", "TapeHolder");

struct DiffStepsFn : public torch::autograd::Function<DiffStepsFn> {
  static torch::autograd::variable_list forward(torch::autograd::AutogradContext* ctx, torch::Tensor vel0,
                                                torch::Tensor rho0, int64_t topo, int64_t steps, double dt, double dx,
                                                int64_t iterations, double omega, int64_t every, int64_t stream) {
    check_tensor(vel0, "vel0");
    check_tensor(rho0, "rho0");
    TORCH_CHECK(vel0.numel() == 3 * rho0.numel(), "vel0 must hold 3 N values (SoA) for N densities");
    auto v = torch::empty_like(vel0);
    auto r = torch::empty_like(rho0);
    hns_diff_tape* tape = nullptr;
    check(hnsk_diff_forward(reinterpret_cast<const hns_topology*>(topo), vel0.data_ptr<float>(), rho0.data_ptr<float>(),
                            (int32_t)steps, (float)dt, (float)dx, (int32_t)iterations, (float)omega, (int32_t)every,
                            v.data_ptr<float>(), r.data_ptr<float>(), &tape, reinterpret_cast<void*>(stream)));
    auto holder = c10::make_intrusive<TapeHolder>(tape);
    if (ctx) ctx->saved_data["tape"] = holder;  // else: no gradient wanted, the holder frees it now
    if (ctx) ctx->saved_data["stream"] = stream;
    return {v, r};
  }

  static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::variable_list g) {
    auto holder = ctx->saved_data["tape"].toCustomClass<TapeHolder>();
    hns_diff_tape* tape = holder->tape;
    void* stream = reinterpret_cast<void*>(ctx->saved_data["stream"].toInt());
    TORCH_CHECK(tape != nullptr, "backward called twice on a differentiable step");
    torch::Tensor gv = g[0].defined() ? g[0].contiguous() : torch::Tensor();
    torch::Tensor gr = g[1].defined() ? g[1].contiguous() : torch::Tensor();
    const torch::Tensor& like_v = gv.defined() ? gv : gr;
    int64_t n = 0;
    int32_t nck = 0;
    (void)n;
    (void)nck;
    torch::Tensor gv0, gr0;
    if (gv.defined()) {
      gv0 = torch::empty_like(gv);
      gr0 = gr.defined() ? torch::empty_like(gr) : torch::empty({gv.numel() / 3}, gv.options());
    } else {
      gr0 = torch::empty_like(gr);
      gv0 = torch::empty({3 * gr.numel()}, like_v.options());
    }
    check(hnsk_diff_backward(tape, gv.defined() ? gv.data_ptr<float>() : nullptr,
                             gr.defined() ? gr.data_ptr<float>() : nullptr, gv0.data_ptr<float>(), gr0.data_ptr<float>(),
                             stream));
    check(hns_diff_tape_destroy(&holder->tape));
    return {gv0, gr0, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
            torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

std::vector<torch::Tensor> diff_steps(torch::Tensor vel0, torch::Tensor rho0, int64_t topo, int64_t steps, double dt,
                                      double dx, int64_t iterations, double omega, int64_t every, int64_t stream) {
  if (!torch::GradMode::is_enabled() || !(vel0.requires_grad() || rho0.requires_grad())) {
    // no gradient can be asked for: run the forward and free its tape at once
    auto out = DiffStepsFn::forward(nullptr, vel0, rho0, topo, steps, dt, dx, iterations, omega, every, stream);
    return {out[0], out[1]};
  }
  auto out = DiffStepsFn::apply(vel0, rho0, topo, steps, dt, dx, iterations, omega, every, stream);
  return {out[0], out[1]};
}

}  // namespace

PYBIND11_MODULE(hns_torch, m) {
  m.doc() = "torch autograd function over the libhns differentiable step";
  m.def("diff_steps", &diff_steps, "(u_S, rho_S) after `steps` differentiable steps", pybind11::arg("vel0"),
        pybind11::arg("rho0"), pybind11::arg("topo"), pybind11::arg("steps"), pybind11::arg("dt"),
        pybind11::arg("dx"), pybind11::arg("iterations"), pybind11::arg("omega"), pybind11::arg("every"),
        pybind11::arg("stream"));
}
