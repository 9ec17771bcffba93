// autograd.cpp -- torch::autograd::Function over the C ABI adjoint rollout.
//
// J(v0, alpha, q0, eta) = sum over steps of dt * sum_{i in obj} f_i
// (f = p^2 or u^2) for nsteps of time_integration started from a fixed
// initial state with initial velocities v0 (n x 2), artificial-viscosity
// alpha, heat source amplitude q0 and dynamic viscosity eta.  forward() runs
// arm_rollout_forward (checkpoints kept in the device context); backward()
// runs arm_rollout_backward with the incoming gradient as the adjoint seed
// and returns dJ/dv0, dJ/dalpha, dJ/dq0 and dJ/deta.  The context holds the
// checkpoints of one forward at a time: every forward gets a generation
// number, and a backward whose forward is no longer the live one first runs
// that forward again from its saved inputs.  Host code only; the device work
// is in libarmando.so.
#include <torch/extension.h>

#include <cstdint>
#include <string>
#include <vector>

#include "armando.h"

namespace {

struct Rollout {
  arm_ctx* ctx = nullptr;
  arm_params prm{};
  int np = 0, nv = 0;
  std::vector<double> x, rho, u;  // x is (2, n): x and y of a particle adjacent
  std::vector<int32_t> obj;
  int nsteps = 0, ckpt = 0;
  int64_t next_gen = 0;  // generation of the next forward
  int64_t live_gen = -1;  // generation whose checkpoints the context holds
};

void check(const Rollout* r, int rc, const char* what) {
  TORCH_CHECK(rc == ARM_OK, what, " failed (", rc, "): ",
              r->ctx ? arm_last_error(r->ctx) : "no context");
}

Rollout* as_rollout(int64_t h) {
  TORCH_CHECK(h != 0, "null rollout handle");
  return reinterpret_cast<Rollout*>(h);
}

std::vector<double> to_vec(const torch::Tensor& t) {
  auto c = t.detach().to(torch::kCPU, torch::kFloat64).contiguous();
  return std::vector<double>(c.data_ptr<double>(), c.data_ptr<double>() + c.numel());
}

double scalar(const torch::Tensor& t) {
  return t.detach().to(torch::kCPU, torch::kFloat64).item<double>();
}

// arm_rollout_forward from the fixed initial state with initial velocities v
// ((n, 2), CPU, FP64) and the three scalars; returns J.
double run_forward(Rollout* r, const torch::Tensor& v, double alpha, double q0, double eta) {
  arm_params P = r->prm;
  P.av_alpha = alpha;
  P.hs_q0 = q0;
  P.visc_eta = eta;
  check(r, arm_update_params(r->ctx, &P), "arm_update_params");
  check(r, arm_upload(r->ctx, r->np, r->nv, r->x.data(), v.data_ptr<double>(), r->rho.data(),
                      r->u.data()),
        "arm_upload");
  double J = 0.0;
  check(r, arm_rollout_forward(r->ctx, r->nsteps, r->ckpt, r->obj.data(), (int)r->obj.size(),
                               &J, nullptr),
        "arm_rollout_forward");
  return J;
}

class RolloutFn : public torch::autograd::Function<RolloutFn> {
 public:
  static torch::Tensor forward(torch::autograd::AutogradContext* actx, torch::Tensor v0,
                               torch::Tensor alpha, torch::Tensor q0, torch::Tensor eta,
                               int64_t handle) {
    Rollout* r = as_rollout(handle);
    const int n = r->np + r->nv;
    TORCH_CHECK(v0.dim() == 2 && v0.size(0) == n && v0.size(1) == 2, "v0 must be (n, 2)");
    auto v = v0.detach().to(torch::kCPU, torch::kFloat64).contiguous().clone();
    const double a = scalar(alpha), q = scalar(q0), e = scalar(eta);
    r->live_gen = -1;  // a failed forward leaves no usable checkpoints
    const double J = run_forward(r, v, a, q, e);
    const int64_t gen = r->next_gen++;
    r->live_gen = gen;
    actx->saved_data["handle"] = handle;
    actx->saved_data["gen"] = gen;
    actx->saved_data["v"] = v;
    actx->saved_data["alpha"] = a;
    actx->saved_data["q0"] = q;
    actx->saved_data["eta"] = e;
    actx->saved_data["v_device"] = v0.device().str();
    return torch::tensor(J, torch::dtype(torch::kFloat64).device(alpha.device()));
  }

  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* actx,
                                               torch::autograd::tensor_list grads) {
    Rollout* r = as_rollout(actx->saved_data["handle"].toInt());
    const int64_t gen = actx->saved_data["gen"].toInt();
    if (r->live_gen != gen) {
      // a later forward replaced this one's checkpoints: run it again
      r->live_gen = -1;
      run_forward(r, actx->saved_data["v"].toTensor(), actx->saved_data["alpha"].toDouble(),
                  actx->saved_data["q0"].toDouble(), actx->saved_data["eta"].toDouble());
      r->live_gen = gen;
    }
    const std::string dev = actx->saved_data["v_device"].toStringRef();
    const double seed = grads[0].detach().to(torch::kCPU, torch::kFloat64).item<double>();
    const int n = r->np + r->nv;
    auto dv = torch::zeros({n, 2}, torch::dtype(torch::kFloat64));
    check(r, arm_rollout_backward(r->ctx, seed, nullptr, dv.data_ptr<double>(), nullptr),
          "arm_rollout_backward");
    double g[ARM_NGRAD] = {0.0, 0.0, 0.0};
    check(r, arm_rollout_param_grads(r->ctx, g, ARM_NGRAD), "arm_rollout_param_grads");
    auto opts = torch::dtype(torch::kFloat64).device(grads[0].device());
    return {dv.to(torch::Device(dev)), torch::tensor(g[ARM_GRAD_ALPHA], opts),
            torch::tensor(g[ARM_GRAD_Q0], opts), torch::tensor(g[ARM_GRAD_VISC], opts),
            torch::Tensor()};
  }
};

int64_t make_rollout(py::dict p, torch::Tensor x, torch::Tensor rho, torch::Tensor u, int64_t np,
                     int64_t nv, torch::Tensor obj, int64_t nsteps, int64_t ckpt,
                     int64_t objective) {
  auto* r = new Rollout();
  auto d = [&](const char* k) { return p[k].cast<double>(); };
  auto i = [&](const char* k) { return p[k].cast<int>(); };
  arm_params& P = r->prm;
  P.precision = i("precision");
  P.flags = i("flags");
  P.pa_sph = i("pa_sph");
  P.density_method = i("density_method");
  P.h = d("h");
  P.mass = d("mass");
  P.rho0 = d("rho0");
  P.c0 = d("c0");
  P.pmin = d("pmin");
  P.av_alpha = d("av_alpha");
  P.av_beta = d("av_beta");
  P.av_eta = d("av_eta");
  P.ah_g1 = d("ah_g1");
  P.ah_g2 = d("ah_g2");
  P.ah_eta = d("ah_eta");
  P.xsph_eps = d("xsph_eps");
  P.gx = d("gx");
  P.gy = d("gy");
  P.dt = d("dt");
  P.ox = d("ox");
  P.oy = d("oy");
  P.cell = d("cell");
  P.nx = i("nx");
  P.ny = i("ny");
  P.visc_eta = p.contains("visc_eta") ? d("visc_eta") : 0.0;
  P.eos = p.contains("eos") ? i("eos") : ARM_EOS_TAIT;
  P.eos_p0 = p.contains("eos_p0") ? d("eos_p0") : 0.0;
  P.hs_q0 = p.contains("hs_q0") ? d("hs_q0") : 0.0;
  P.hs_x = p.contains("hs_x") ? d("hs_x") : 0.0;
  P.hs_y = p.contains("hs_y") ? d("hs_y") : 0.0;
  P.hs_sigma = p.contains("hs_sigma") ? d("hs_sigma") : 0.0;
  P.dim = 2;  // the adjoint rollout is two-dimensional
  r->np = (int)np;
  r->nv = (int)nv;
  r->x = to_vec(x);
  r->rho = to_vec(rho);
  r->u = to_vec(u);
  auto o = obj.detach().to(torch::kCPU, torch::kInt32).contiguous();
  r->obj.assign(o.data_ptr<int32_t>(), o.data_ptr<int32_t>() + o.numel());
  r->nsteps = (int)nsteps;
  r->ckpt = (int)ckpt;
  const int n = (int)(np + nv);
  TORCH_CHECK((int)r->x.size() == 2 * n && (int)r->rho.size() == n && (int)r->u.size() == n,
              "x must be (n, 2), rho and u (n)");
  int rc = arm_create(&P, n, &r->ctx);
  if (rc == ARM_OK) rc = arm_set_objective(r->ctx, (int)objective);
  if (rc != ARM_OK) {
    const std::string msg = r->ctx ? arm_last_error(r->ctx) : "";
    arm_destroy(r->ctx);
    delete r;
    TORCH_CHECK(false, "rollout context failed (", rc, "): ", msg);
  }
  return reinterpret_cast<int64_t>(r);
}

void free_rollout(int64_t h) {
  if (!h) return;
  Rollout* r = as_rollout(h);
  arm_destroy(r->ctx);
  delete r;
}

torch::Tensor rollout_loss(int64_t h, torch::Tensor v0, torch::Tensor alpha, torch::Tensor q0,
                           torch::Tensor eta) {
  return RolloutFn::apply(v0, alpha, q0, eta, h);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "torch.autograd over the MI355X WCSPH adjoint rollout";
  m.def("make_rollout", &make_rollout, "create a rollout context (returns a handle)");
  m.def("free_rollout", &free_rollout, "destroy a rollout context");
  m.def("rollout_loss", &rollout_loss,
        "J(v0, alpha, q0, eta) with a hand-written adjoint backward");
}
