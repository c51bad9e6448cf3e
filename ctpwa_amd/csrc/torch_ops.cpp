// PyTorch extension over the C ABI: the likelihood as a C++ autograd function.
//
// analysis.getNLL(params) returns a 0-dim float64 tensor whose backward gives the analytic
// gradient, as the reference's NLLFunction (src/Analysis.cu:625-1261: forward evaluates the
// NLL and keeps the gradient, backward scales it by the incoming gradient).  Here one
// ctpwa_nll_grad call computes both; the function keeps the gradient part of its output
// and backward returns gradient * grad_out.  Written in C++ so that a fit step (forward,
// autograd.grad, backward) does not cross into the Python interpreter on the autograd
// engine's device thread.
//
// No HIP types appear here: the caller passes the handle and the stream (the current torch
// stream of the parameters' device) as integers, and every device launch happens inside
// libctpwa_amd.so.
#include <torch/extension.h>

#include <stdexcept>
#include <string>

#include "../../include/ctpwa_amd.h"

namespace {

[[noreturn]] void raise(const char* what) {
    const char* m = ctpwa_last_error();
    throw std::runtime_error(std::string(what) + ": " + (m ? m : "failed"));
}

struct NLLFunction : public torch::autograd::Function<NLLFunction> {
    static at::Tensor forward(torch::autograd::AutogradContext* ctx, const at::Tensor& params, int64_t handle,
                              int64_t stream, int64_t n_params) {
        const at::Tensor p = params.contiguous();
        at::Tensor out = at::empty({1 + n_params}, p.options());
        if (ctpwa_nll_grad(reinterpret_cast<ctpwa_handle*>(handle), p.data_ptr<double>(), out.data_ptr<double>(),
                           reinterpret_cast<void*>(stream)) != 0)
            raise("ctpwa_nll_grad");
        ctx->save_for_backward({out.narrow(0, 1, n_params)});
        return out.select(0, 0);
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_out) {
        const at::Tensor g = ctx->get_saved_variables()[0];
        return {g * grad_out[0], at::Tensor(), at::Tensor(), at::Tensor()};
    }
};

// params: float64, 1-D, n_params entries, on the handle's device (checked by the caller)
at::Tensor nll(const at::Tensor& params, int64_t handle, int64_t stream, int64_t n_params) {
    return NLLFunction::apply(params, handle, stream, n_params);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "ctpwa likelihood as a C++ autograd function over libctpwa_amd.so";
    // the GIL is released while the function records its graph node and launches the
    // evaluation (other Python threads, such as a service's admission thread, run meanwhile)
    m.def("nll", &nll, pybind11::call_guard<pybind11::gil_scoped_release>(),
          "NLL with the analytic gradient as its backward");
}
