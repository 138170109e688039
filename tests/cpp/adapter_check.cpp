// Exercises the C++ adapters (zisafvm_amd/csrc/adapters.hpp) the way the
// reference's TimeLoop drives RateOfChange / TimeIntegration
// (src/zisa/fvm_loops/time_loop.cpp:36-75).  Built by tests/test_cpp_adapters.py;
// run on a GPU box.  Prints "adapter_check: OK" on success.
#include <cmath>
#include <cstdio>
#include <memory>

#include "../../zisafvm_amd/csrc/adapters.hpp"

int main() {
  using namespace zfv;
  zfv_mesh_buf *buf = nullptr;
  check(zfv_mesh_lattice2d(16, 0.2, 7, 1, 1, 0, &buf));
  zfv_mesh mesh;
  check(zfv_mesh_view(buf, &mesh));

  zfv_params p{};
  p.n_dims = 2;
  p.n_stencils = 4;
  const int orders[4] = {3, 2, 2, 2}, biases[4] = {0, 1, 1, 1};
  const double overfit[4] = {2.0, 1.5, 1.5, 1.5}, lw[4] = {100.0, 1.0, 1.0, 1.0};
  for (int k = 0; k < 4; ++k) {
    p.orders[k] = orders[k];
    p.biases[k] = biases[k];
    p.overfit[k] = overfit[k];
    p.linear_weights[k] = lw[k];
  }
  p.epsilon = 1e-6;
  p.exponent = 4.0;
  p.gamma = 1.4;
  p.gas_constant = 1.0;
  p.edge_deg = 3;
  p.volume_deg = 4;
  p.moments_deg = -1;
  p.steps_per_recompute = 1;
  p.recompute_threshold = 1e300;

  const int n = mesh.n_cells;
  auto inst = std::make_shared<Instance>(mesh, p);
  auto u0 = std::make_shared<AllVariables>(n);
  for (int i = 0; i < n; ++i) {
    const double *c = mesh.corners + 9 * i;
    double x = (c[0] + c[3] + c[6]) / 3.0, y = (c[1] + c[4] + c[7]) / 3.0;
    double rho = 1.0 + 0.2 * std::sin(2 * M_PI * x) * std::cos(2 * M_PI * y);
    double *u = u0->data() + 5 * i;
    u[0] = rho;
    u[1] = rho * 0.5;
    u[2] = rho * -0.25;
    u[3] = 0.0;
    u[4] = 2.5 + 0.5 * rho * (0.25 + 0.0625);
  }

  // RateOfChange::compute adds into the caller-owned tendency
  HipRateOfChange roc(inst);
  AllVariables k1(n), k2(n);
  for (auto &v : k2.cvars) v = 1.0;
  roc.compute(k1, *u0, 0.0);
  roc.compute(k2, *u0, 0.0);
  for (size_t q = 0; q < k1.cvars.size(); ++q)
    if (k2.cvars[q] != 1.0 + k1.cvars[q]) {
      std::printf("adapter_check: accumulation mismatch at %zu\n", q);
      return 1;
    }

  // TimeIntegration::compute_step keeps u0 and swaps buffers
  HipCFL cfl(inst, 0.45);
  HipRungeKutta rk(inst, n);
  const AllVariables before = *u0;
  double t = 0.0, dt = cfl(*u0);
  auto u1 = rk.compute_step(u0, t, dt);
  if (u1.get() == u0.get() || u0->cvars != before.cvars) {
    std::printf("adapter_check: buffer contract violated\n");
    return 1;
  }
  auto u2 = rk.compute_step(u1, t + dt, dt);
  if (u2.get() != u0.get()) {
    std::printf("adapter_check: expected the old u0 to be reused as the next output buffer\n");
    return 1;
  }

  // TimeLoop may change the returned state in place before the next step
  // (instantaneous physics, src/zisa/fvm_loops/time_loop.cpp:60): the next
  // step must start from the changed values, not from the device copy
  auto inst3 = std::make_shared<Instance>(mesh, p);
  HipRungeKutta rk3(inst3, n);
  auto v0 = std::make_shared<AllVariables>(before);
  auto v1 = rk3.compute_step(v0, t, dt);
  for (int i = 0; i < n; ++i) v1->cvars[5 * size_t(i) + 4] *= 1.001;  // heat every cell
  const AllVariables changed = *v1;
  auto v2 = rk3.compute_step(v1, t + dt, dt);
  auto inst4 = std::make_shared<Instance>(mesh, p);
  check(zfv_set_state(inst4->get(), changed.data()));
  check(zfv_step(inst4->get(), t + dt, dt));
  AllVariables ref4(n);
  check(zfv_get_state(inst4->get(), ref4.data()));
  if (ref4.cvars != v2->cvars) {
    std::printf("adapter_check: a step after an in-place change of the state did not start from it\n");
    return 1;
  }

  // the same two steps through the C-ABI on a fresh instance
  auto inst2 = std::make_shared<Instance>(mesh, p);
  check(zfv_set_state(inst2->get(), before.data()));
  check(zfv_steps(inst2->get(), 2, 0.0, dt));
  AllVariables ref(n);
  check(zfv_get_state(inst2->get(), ref.data()));
  for (size_t q = 0; q < ref.cvars.size(); ++q)
    if (ref.cvars[q] != u2->cvars[q]) {
      std::printf("adapter_check: step mismatch at %zu\n", q);
      return 1;
    }
  zfv_mesh_free(buf);
  std::printf("adapter_check: OK\n");
  return 0;
}
