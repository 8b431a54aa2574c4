// pn_torch.cpp -- the adjoint form's loss as a libtorch autograd function.
//
// WeibelLossFn::forward runs the PIC steps of the adjoint form on the GPU
// (pn_adj_forward: checkpoints every K steps) and returns L = 1/2 sum B^2;
// backward runs the hand-written discrete adjoint kernels (pn_adj_backward:
// recomputation from the checkpoints, deposit adjoint = gather, gather adjoint =
// scatter) and returns dL/dtheta scaled by the incoming gradient, on theta's
// device and in theta's dtype.  The C ABI (include/picnix_adjoint.h) keeps one
// tape per handle -- the last forward's checkpoints -- so backward checks that
// no other forward ran on the same handle since its own forward.
#include <torch/extension.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/picnix_adjoint.h"

namespace {

// forwards run on each handle so far: a backward needs the tape of its own forward
std::mutex                          tape_mu;
std::unordered_map<int64_t, int64_t> tape_count;

pn_adj* handle_of(int64_t h)
{
  TORCH_CHECK(h != 0, "picnix adjoint: null adjoint-form handle");
  return reinterpret_cast<pn_adj*>(static_cast<intptr_t>(h));
}

std::string last_error(pn_adj* a)
{
  const char* m = pn_adj_last_error(a);
  return m ? std::string(m) : std::string();
}

struct WeibelLossFn : public torch::autograd::Function<WeibelLossFn> {
  static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor theta,
                               int64_t handle, int64_t nsteps, int64_t every)
  {
    pn_adj*       a = handle_of(handle);
    torch::Tensor t = theta.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    TORCH_CHECK(t.dim() == 1, "picnix adjoint: theta must be a vector (ratio, eps_0 .. eps_M-1)");
    double loss = 0.0;
    const int rc = pn_adj_forward(a, t.data_ptr<double>(), static_cast<int>(nsteps),
                                  static_cast<int>(every), &loss);
    TORCH_CHECK(rc == 0, "pn_adj_forward failed (", rc, "): ", last_error(a));
    int64_t tape;
    {
      std::lock_guard<std::mutex> lock(tape_mu);
      tape = ++tape_count[handle];
    }
    ctx->saved_data["handle"] = handle;
    ctx->saved_data["tape"]   = tape;
    ctx->saved_data["n"]      = t.size(0);
    ctx->saved_data["device"] = theta.device();
    ctx->saved_data["dtype"]  = static_cast<int64_t>(theta.scalar_type());
    return torch::tensor(loss, torch::dtype(torch::kFloat64)).to(theta.options());
  }

  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                               torch::autograd::tensor_list grad_outputs)
  {
    const int64_t handle = ctx->saved_data["handle"].toInt();
    const int64_t tape   = ctx->saved_data["tape"].toInt();
    {
      std::lock_guard<std::mutex> lock(tape_mu);
      TORCH_CHECK(tape_count[handle] == tape,
                  "picnix adjoint: another forward ran on this adjoint form after the one being "
                  "differentiated; backward needs the checkpoints of its own forward");
    }
    pn_adj*       a = handle_of(handle);
    torch::Tensor g = torch::zeros({ctx->saved_data["n"].toInt()}, torch::kFloat64);
    const int     rc = pn_adj_backward(a, g.data_ptr<double>());
    TORCH_CHECK(rc == 0, "pn_adj_backward failed (", rc, "): ", last_error(a));
    const auto dev   = ctx->saved_data["device"].toDevice();
    const auto dtype = static_cast<c10::ScalarType>(ctx->saved_data["dtype"].toInt());
    g                = g.to(dev, dtype) * grad_outputs[0].to(dev, dtype);
    return {g, torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor weibel_loss(torch::Tensor theta, int64_t handle, int64_t nsteps, int64_t every)
{
  return WeibelLossFn::apply(theta, handle, nsteps, every);
}

} // namespace

PYBIND11_MODULE(picnix_torch, m)
{
  m.doc() = "libtorch autograd function over the picnix adjoint form (include/picnix_adjoint.h)";
  m.def("weibel_loss", &weibel_loss,
        "L = 1/2 sum B^2 after nsteps steps of the adjoint form (checkpoint every `every`); "
        "differentiable in theta through the hand-written adjoint kernels",
        pybind11::arg("theta"), pybind11::arg("handle"), pybind11::arg("nsteps"),
        pybind11::arg("every"));
}
