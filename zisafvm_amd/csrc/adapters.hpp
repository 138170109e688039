// C++ adapters that put libzfv_hip.so behind the reference's class interfaces:
//
//   RateOfChange::compute(AllVariables &tendency, const AllVariables &state, double t)
//       include/zisa/ode/rate_of_change.hpp:37-39
//   TimeIntegration::compute_step(const shared_ptr<AllVariables> &u0, double t, double dt)
//       include/zisa/ode/time_integration.hpp:42-44
//   CFLCondition::operator()(const AllVariables &)
//       include/zisa/model/cfl_condition.hpp:17
//
// `AllVariables` here is a minimal stand-in with the reference's layout
// (cvars n_cells x 5 row-major, src/zisa/model/all_variables.cpp:25-30); a
// maintainer replaces it by zisa::AllVariables and the base classes by the
// reference's own (include/zisa/ode/rate_of_change.hpp,
// include/zisa/ode/time_integration.hpp).  Header-only; link against
// libzfv_hip.so.
#pragma once

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/zfv.h"

namespace zfv {

struct AllVariables {
  int n_cells = 0;
  std::vector<double> cvars;  // n_cells x 5, row-major
  explicit AllVariables(int n = 0) : n_cells(n), cvars(size_t(n) * 5, 0.0) {}
  double *data() { return cvars.data(); }
  const double *data() const { return cvars.data(); }
};

inline void check(int rc) {
  if (rc != ZFV_OK) throw std::runtime_error(std::string("zfv: ") + zfv_last_error());
}

// Owns one device instance (mesh, stencils, pseudo-inverses, state).
class Instance {
 public:
  Instance(const zfv_mesh &mesh, const zfv_params &params) { check(zfv_create(&mesh, &params, &h_)); }
  ~Instance() { zfv_destroy(h_); }
  Instance(const Instance &) = delete;
  Instance &operator=(const Instance &) = delete;
  zfv_handle *get() const { return h_; }

 private:
  zfv_handle *h_ = nullptr;
};

class RateOfChange {
 public:
  virtual ~RateOfChange() = default;
  virtual void compute(AllVariables &tendency, const AllVariables &current_state, double t) const = 0;
};

// The whole physical rate of change (FluxLoop + GravitySourceLoop + FluxBC)
// evaluated on the GPU; like every reference term it *adds* into `tendency`.
class HipRateOfChange : public RateOfChange {
 public:
  explicit HipRateOfChange(std::shared_ptr<Instance> inst) : inst_(std::move(inst)) {}
  void compute(AllVariables &tendency, const AllVariables &current_state, double t) const override {
    check(zfv_rate(inst_->get(), current_state.data(), tendency.data(), t));
  }

 private:
  std::shared_ptr<Instance> inst_;
};

class TimeIntegration {
 public:
  virtual ~TimeIntegration() = default;
  virtual std::shared_ptr<AllVariables> compute_step(const std::shared_ptr<AllVariables> &u0, double t,
                                                     double dt) = 0;
};

// SSP3 in Butcher form with the frozen BC after every stage sum, on the
// device.  Keeps the reference's buffer contract: u0 is not modified, the
// returned buffer is the previous scratch buffer and u0 becomes the next
// scratch (src/zisa/ode/runge_kutta.cpp:87-112).  The device holds the
// state between calls.  The caller may change the returned buffer in place
// before passing it back (TimeLoop applies its instantaneous physics to it,
// src/zisa/fvm_loops/time_loop.cpp:60), so the upload is skipped only when
// u0 is the buffer returned last and still holds, bit for bit, what was
// returned; anything else is uploaded.
class HipRungeKutta : public TimeIntegration {
 public:
  explicit HipRungeKutta(std::shared_ptr<Instance> inst, int n_cells)
      : inst_(std::move(inst)), ux_(std::make_shared<AllVariables>(n_cells)) {}

  std::shared_ptr<AllVariables> compute_step(const std::shared_ptr<AllVariables> &u0, double t,
                                             double dt) override {
    const bool resident = u0.get() == last_.get() && u0->cvars.size() == returned_.size() &&
                          std::memcmp(u0->data(), returned_.data(), returned_.size() * sizeof(double)) == 0;
    if (!resident) check(zfv_set_state(inst_->get(), u0->data()));
    check(zfv_step(inst_->get(), t, dt));
    check(zfv_get_state(inst_->get(), ux_->data()));
    returned_ = ux_->cvars;
    auto out = ux_;
    ux_ = u0;
    last_ = out;
    return out;
  }

 private:
  std::shared_ptr<Instance> inst_;
  std::shared_ptr<AllVariables> ux_, last_;
  std::vector<double> returned_;  // the values handed out last (the device state)
};

// LocalCFL on the device-resident state.
class HipCFL {
 public:
  HipCFL(std::shared_ptr<Instance> inst, double cfl) : inst_(std::move(inst)), cfl_(cfl) {}
  double operator()(const AllVariables &u) const {
    check(zfv_set_state(inst_->get(), u.data()));
    double dt = 0.0;
    check(zfv_cfl_dt(inst_->get(), cfl_, &dt));
    return dt;
  }

 private:
  std::shared_ptr<Instance> inst_;
  double cfl_;
};

}  // namespace zfv
