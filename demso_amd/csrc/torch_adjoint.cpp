// torch_adjoint.cpp -- the DEM trajectory as a torch::autograd::Function.
//
//   final_pos_by_id = DemTrajectory(pos0_by_id, params)
//
// pos0_by_id: (n, 2) float32 GPU tensor of initial positions indexed by
// particle id; params: (num_types, 4) float64 CPU tensor of (kn, kt, cn, mu)
// per material.  forward() writes the parameters and positions into an
// existing dem_sim and runs `steps` steps through dem_adj_forward (keeping
// checkpoints); backward() calls dem_adj_backward (hand-written adjoint of
// collide + integrate, FP64 gradient accumulation) and returns dL/dpos0 and
// dL/dparams.  Everything runs on the library's context stream, which is
// the null stream, as torch's default stream.
#include <torch/extension.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "demso.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::Function;
using torch::autograd::tensor_list;

void check(int rc, dem_context* ctx, const char* what) {
  if (rc != DEM_OK) throw std::runtime_error(std::string(what) + ": " + dem_last_error(ctx));
}

struct DemTrajectory : public Function<DemTrajectory> {
  static torch::Tensor forward(AutogradContext* actx, torch::Tensor pos0, torch::Tensor params, int64_t ctx_h,
                               int64_t sim_h, int64_t adj_h, int64_t steps) {
    auto* ctx = reinterpret_cast<dem_context*>(ctx_h);
    auto* sim = reinterpret_cast<dem_sim*>(sim_h);
    auto* adj = reinterpret_cast<dem_adjoint*>(adj_h);
    TORCH_CHECK(ctx && sim && adj, "null context, sim or adjoint handle");
    TORCH_CHECK(steps >= 0 && steps <= UINT32_MAX, "steps out of range");
    uint32_t n = 0;
    int device = -1;
    check(dem_sim_info(sim, &n, &device), ctx, "dem_sim_info");
    // every check happens before any launch: the kernels read pos0[id] for
    // each of the sim's n particle ids
    TORCH_CHECK(pos0.is_cuda() && pos0.scalar_type() == torch::kFloat32 && pos0.dim() == 2 && pos0.size(1) == 2,
                "pos0 must be an (n, 2) float32 GPU tensor");
    TORCH_CHECK(pos0.size(0) == (int64_t)n, "pos0 has ", pos0.size(0), " rows; the simulation has ", n,
                " particles");
    TORCH_CHECK(pos0.get_device() == device, "pos0 is on cuda:", pos0.get_device(),
                "; the simulation runs on cuda:", device);
    auto p = params.detach().to(torch::kCPU, torch::kFloat64).contiguous();
    TORCH_CHECK(p.dim() == 2 && p.size(1) == 4, "params must be (num_types, 4)");
    auto x0 = pos0.detach().contiguous();
    check(dem_set_contact_params(ctx, p.data_ptr<double>(), (uint32_t)p.size(0)), ctx, "dem_set_contact_params");
    check(dem_sim_set_by_id(sim, x0.data_ptr<float>(), nullptr), ctx, "dem_sim_set_by_id");
    check(dem_adj_forward(adj, (uint32_t)steps), ctx, "dem_adj_forward");
    uint64_t gen = 0;
    check(dem_adj_generation(adj, &gen), ctx, "dem_adj_generation");
    auto out = torch::empty_like(x0);
    check(dem_sim_get_by_id(sim, out.data_ptr<float>(), nullptr), ctx, "dem_sim_get_by_id");
    check(dem_synchronize(ctx), ctx, "dem_synchronize");
    actx->saved_data["ctx"] = ctx_h;
    actx->saved_data["adj"] = adj_h;
    actx->saved_data["ntypes"] = (int64_t)p.size(0);
    actx->saved_data["n"] = (int64_t)n;
    actx->saved_data["device"] = (int64_t)device;
    actx->saved_data["generation"] = (int64_t)gen;
    return out;
  }

  static tensor_list backward(AutogradContext* actx, tensor_list grads) {
    auto* ctx = reinterpret_cast<dem_context*>(actx->saved_data["ctx"].toInt());
    auto* adj = reinterpret_cast<dem_adjoint*>(actx->saved_data["adj"].toInt());
    const int64_t nt = actx->saved_data["ntypes"].toInt();
    // the adjoint holds one trajectory (checkpoints, final state): the
    // latest forward's.  A graph built by an earlier call cannot be
    // differentiated through it any more.
    uint64_t gen = 0;
    check(dem_adj_generation(adj, &gen), ctx, "dem_adj_generation");
    TORCH_CHECK((int64_t)gen == actx->saved_data["generation"].toInt(),
                "this trajectory was replaced by a later call on the same Trajectory; call backward before "
                "running it again, or use one Trajectory per differentiated run");
    auto g = grads[0].to(torch::kFloat32).contiguous();
    TORCH_CHECK(g.is_cuda() && g.get_device() == actx->saved_data["device"].toInt() && g.dim() == 2 &&
                    g.size(0) == actx->saved_data["n"].toInt() && g.size(1) == 2,
                "the gradient of the final positions must be an (n, 2) tensor on the simulation's device");
    auto gp0 = torch::empty_like(g);
    auto gpar = torch::empty({nt, 4}, g.options().dtype(torch::kFloat64));
    check(dem_adj_backward(adj, g.data_ptr<float>(), nullptr, gp0.data_ptr<float>(), nullptr, gpar.data_ptr<double>()),
          ctx, "dem_adj_backward");
    return {gp0, gpar.cpu(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
  }
};

torch::Tensor trajectory(torch::Tensor pos0, torch::Tensor params, int64_t ctx, int64_t sim, int64_t adj,
                         int64_t steps) {
  return DemTrajectory::apply(pos0, params, ctx, sim, adj, steps);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("trajectory", &trajectory,
        "Differentiable DEM trajectory: final positions (by id) from initial positions and (kn, kt, cn, mu)");
}
