// zfv_ref reconstruction (CWENO-AO / WENO-AO with LDLT least squares),
// physics (ideal gas, HLLC, Rusanov, gravity, local equilibria), rate of
// change (FluxLoop + GravitySourceLoop + FluxBC), SSP3 and CFL.
// TEST INFRASTRUCTURE ONLY (see zfv_ref.hpp).
#include "zfv_ref.hpp"
#include "../../include/zfv.h"

#include <algorithm>
#include <type_traits>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace zfv_ref {

// ---------------------------------------------------------------------------
// Ideal gas (include/zisa/model/ideal_gas_eos.hpp, euler_variables.hpp:50-56)
double kinetic_energy(const Vars &u) { return 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0]; }
double internal_energy(const Vars &u) { return u[4] - kinetic_energy(u); }
double pressure(const EOS &eos, const Vars &u) { return (u[4] - kinetic_energy(u)) * (eos.gamma - 1.0); }
double sound_speed(const EOS &eos, const Vars &u) { return std::sqrt(eos.gamma * pressure(eos, u) / u[0]); }

// include/zisa/model/euler_impl.hpp:23-36
Vars euler_flux(const Vars &u, double p) {
  double v = u[1] / u[0];
  return {u[1], v * u[1] + p, v * u[2], v * u[3], v * (u[4] + p)};
}

// src/zisa/model/euler_variables.cpp:29-66
void coord_transform(Vars &u, const Face &f) {
  const XYZ &n = f.normal, &t1 = f.t1, &t2 = f.t2;
  double un = u[1] * n[0] + u[2] * n[1] + u[3] * n[2];
  double ut1 = u[1] * t1[0] + u[2] * t1[1] + u[3] * t1[2];
  double ut2 = u[1] * t2[0] + u[2] * t2[1] + u[3] * t2[2];
  u[1] = un;
  u[2] = ut1;
  u[3] = ut2;
}
void inv_coord_transform(Vars &u, const Face &f) {
  const XYZ &n = f.normal, &t1 = f.t1, &t2 = f.t2;
  double ux = u[1] * n[0] + u[2] * t1[0] + u[3] * t2[0];
  double uy = u[1] * n[1] + u[2] * t1[1] + u[3] * t2[1];
  double uz = u[1] * n[2] + u[2] * t1[2] + u[3] * t2[2];
  u[1] = ux;
  u[2] = uy;
  u[3] = uz;
}

// include/zisa/flux/hllc.hpp:17-176 (HLLC with Einfeldt-Batten speeds)
Vars hllc_flux(const EOS &eos, const Vars &uL, const Vars &uR) {
  double pL = pressure(eos, uL), aL = std::sqrt(eos.gamma * pL / uL[0]);
  double pR = pressure(eos, uR), aR = std::sqrt(eos.gamma * pR / uR[0]);
  double roe_ratio = std::sqrt(uR[0] / uL[0]);
  auto roe = [roe_ratio](double qL, double qR) { return (qL + qR * roe_ratio) / (1.0 + roe_ratio); };
  double vL = uL[1] / uL[0];
  double vR = uR[1] / uR[0];
  double v_tilda = roe(vL, vR);
  double HL = (uL[4] + pL) / uL[0];
  double HR = (uR[4] + pR) / uR[0];
  double H_tilda = roe(HL, HR);
  double r1 = roe(uL[1] / uL[0], uR[1] / uR[0]);
  double r2 = roe(uL[2] / uL[0], uR[2] / uR[0]);
  double r3 = roe(uL[3] / uL[0], uR[3] / uR[0]);
  double vroe_square = r1 * r1 + r2 * r2 + r3 * r3;
  double a_tilda = std::sqrt((eos.gamma - 1.0) * (H_tilda - 0.5 * vroe_square));
  double sL = std::min(vL - aL, v_tilda - a_tilda);
  double sR = std::max(vR + aR, v_tilda + a_tilda);
  double s_star = (uR[1] * (sR - vR) - uL[1] * (sL - vL) + pL - pR) / (uR[0] * (sR - vR) - uL[0] * (sL - vL));

  const Vars &uK = (0.0 <= s_star ? uL : uR);
  double pK = (0.0 <= s_star ? pL : pR);
  Vars nf = euler_flux(uK, pK);
  if (sL < 0.0 && 0.0 <= sR) {
    double sK = (0.0 <= s_star ? sL : sR);
    double vK = (0.0 <= s_star ? uL[1] / uL[0] : uR[1] / uR[0]);
    double cK = (sK - vK) / (sK - s_star);
    nf[0] += sK * (cK * uK[0] - uK[0]);
    nf[1] += sK * (cK * uK[0] * s_star - uK[1]);
    nf[2] += sK * (cK * uK[2] - uK[2]);
    nf[3] += sK * (cK * uK[3] - uK[3]);
    nf[4] += sK * (cK * (uK[4] + (s_star - vK) * (uK[0] * s_star + pK / (sK - vK))) - uK[4]);
  }
  return nf;
}

// Rusanov (local Lax-Friedrichs).  Not in the reference: a build extension,
// F = (F_L + F_R)/2 - s_max/2 (u_R - u_L), s_max = max(|v_n| + a) over sides.
Vars rusanov_flux(const EOS &eos, const Vars &uL, const Vars &uR) {
  double pL = pressure(eos, uL), aL = std::sqrt(eos.gamma * pL / uL[0]);
  double pR = pressure(eos, uR), aR = std::sqrt(eos.gamma * pR / uR[0]);
  Vars fL = euler_flux(uL, pL), fR = euler_flux(uR, pR);
  double s = std::max(std::abs(uL[1] / uL[0]) + aL, std::abs(uR[1] / uR[0]) + aR);
  Vars f;
  for (int v = 0; v < 5; ++v) f[v] = 0.5 * (fL[v] + fR[v]) - 0.5 * s * (uR[v] - uL[v]);
  return f;
}

// ---------------------------------------------------------------------------
// Gravity (include/zisa/model/gravity_decl.hpp:17-140, gravity_impl.hpp:13-59)
double Gravity::phi(const XYZ &x) const {
  if (kind == 1) return g * dot(x, axis);
  if (kind == 2) return g * norm(x);
  if (kind == 3) {
    double alpha = std::sqrt(2.0 * M_PI * G / K);
    double ce = alpha * (norm(x) + 2.2250738585072014e-308);
    return -2.0 * K * rhoC * std::sin(ce) / ce;
  }
  return 0.0;
}
XYZ Gravity::grad_phi(const XYZ &x) const {
  if (kind == 1) return {g * axis[0], g * axis[1], g * axis[2]};
  double r = norm(x);
  double d = 0.0;
  if (kind == 2) {
    d = g;
  } else if (kind == 3) {
    double alpha = std::sqrt(2.0 * M_PI * G / K);
    double ce = alpha * (r + 2.2250738585072014e-308);
    double dphi = (std::cos(ce) - std::sin(ce) / ce) / ce;
    d = -2.0 * K * rhoC * dphi * alpha;
  } else {
    return {0.0, 0.0, 0.0};
  }
  return {d * (x[0] / (r + 1e-50)), d * (x[1] / (r + 1e-50)), d * (x[2] / (r + 1e-50))};
}

// ---------------------------------------------------------------------------
// Local equilibria (include/zisa/model/local_equilibrium_impl.hpp:41-186,
// isentropic_equilibrium.hpp:16-69).  kind 2 (isothermal) is an extension.
struct RhoE {
  double rho = 0.0, E = 0.0;
};
struct EqPoint {
  double rho = 0.0, E = 0.0, p = 0.0, a = 0.0;
};

template <class F>
static auto cell_average(const DenormalizedRule &qr, const F &f) {
  auto ret = f(qr.points[0]);
  auto mul = [](double w, decltype(ret) v) {
    for (auto &c : v) c = w * c;
    return v;
  };
  ret = mul(qr.weights[0], ret);
  for (size_t i = 1; i < qr.weights.size(); ++i) {
    auto t = mul(qr.weights[i], f(qr.points[i]));
    for (size_t c = 0; c < ret.size(); ++c) ret[c] = ret[c] + t[c];
  }
  ret = mul(1.0, ret);
  for (auto &c : ret) c = c / qr.volume;
  return ret;
}

// quasi_newton (include/zisa/math/quasi_newton.hpp:12-50) with the
// RollingConvergenceRate of its steps (include/zisa/math/rolling_convergence_rate.hpp):
// x <- x - inv_df(x, f(x)) until every |dx_k| <= atol_k or max_iter steps.
// From the fifth step on, unless the last step is within atol (signed, as
// the reference compares it), every component's rate
// log(|c|/|b|) / log(|b|/|a|) over the last three steps must be >= 0, else
// the solve is "not converging" (the reference aborts with LOG_ERR).  After
// max_iter steps without every |dx_k| <= 1000 atol_k: not converged, and x0
// is returned, as the reference does.
enum QNStatus { QN_CONVERGED = 0, QN_NOT_CONVERGED = 1, QN_NOT_CONVERGING = 2 };

template <int N, class F, class IDF>
QNStatus quasi_newton(const F &f, const IDF &inv_df, const std::array<double, N> &x0,
                      const std::array<double, N> &atol, int max_iter, std::array<double, N> &x,
                      std::vector<std::array<double, N>> *steps = nullptr) {
  x = x0;
  auto fx = f(x);
  std::array<double, N> dx;
  for (int k = 0; k < N; ++k) dx[k] = 2.0 * atol[k] + 1.0;
  auto converged = [&](double factor) {
    for (int k = 0; k < N; ++k)
      if (!(std::abs(dx[k]) <= factor * atol[k])) return false;
    return true;
  };
  std::vector<std::array<double, N>> hist;
  int iter = 0;
  while (!converged(1.0) && iter < max_iter) {
    dx = inv_df(x, fx);
    for (int k = 0; k < N; ++k) x[k] -= dx[k];
    fx = f(x);
    hist.push_back(dx);
    if (steps) steps->push_back(dx);
    if (iter >= 4) {
      const auto &c = hist[hist.size() - 1];
      bool within = true;
      for (int k = 0; k < N; ++k) within = within && c[k] <= atol[k];
      if (!within) {
        const auto &a = hist[hist.size() - 3], &b = hist[hist.size() - 2];
        for (int k = 0; k < N; ++k) {
          double rate = std::log(std::abs(c[k]) / std::abs(b[k])) / std::log(std::abs(b[k]) / std::abs(a[k]));
          if (!(rate >= 0.0)) {
            x = x0;
            return QN_NOT_CONVERGING;
          }
        }
      }
    }
    ++iter;
  }
  if (iter == max_iter && !converged(1000.0)) {
    x = x0;
    return QN_NOT_CONVERGED;
  }
  return QN_CONVERGED;
}

struct LocalEquilibrium {
  int kind = 0;  // 0 none, 1 isentropic, 2 isothermal
  const EOS *eos = nullptr;
  const Gravity *gravity = nullptr;
  double theta[2] = {0.0, 0.0};
  XYZ x_ref{};
  bool found = false;

  // theta -> (rho, E) at x (IdealGasEOS::rhoE(EnthalpyEntropy), ideal_gas_eos.hpp:162-168)
  std::array<double, 2> extrapolate_with(const double *th, const XYZ &x) const {
    double gamma = eos->gamma;
    if (kind == 1) {
      double h = th[0] + gravity->phi(x_ref) - gravity->phi(x);
      double K = th[1];
      double base = 1.0 / K * (gamma - 1.0) / gamma * h;
      double rho = std::pow(base, 1.0 / (gamma - 1.0));
      double p = K * std::pow(rho, gamma);
      return {rho, p / (gamma - 1.0)};
    }
    double RT = th[1];
    double rho = th[0] * std::exp(-(gravity->phi(x) - gravity->phi(x_ref)) / RT);
    return {rho, rho * RT / (gamma - 1.0)};
  }

  void solve(const RhoE &bar, const DenormalizedRule &cell) {
    if (kind == 0) return;
    x_ref = cell.points[0];
    double gamma = eos->gamma;
    if (kind == 2) {
      double RT = (gamma - 1.0) * bar.E / bar.rho;
      double th1[2] = {1.0, RT};
      auto avg = cell_average(cell, [&](const XYZ &x) { return extrapolate_with(th1, x); });
      theta[0] = bar.rho / avg[0];
      theta[1] = RT;
      found = true;
      return;
    }
    // isentropic: quasi-Newton on (h, K) with a central finite-difference
    // Jacobian (include/zisa/model/local_equilibrium_impl.hpp:41-93)
    double p = bar.E * (gamma - 1.0);
    std::array<double, 2> guess{gamma / (gamma - 1.0) * p / bar.rho, p / std::pow(bar.rho, gamma)};
    auto f = [&](const std::array<double, 2> &th) {
      auto avg = cell_average(cell, [&](const XYZ &x) { return extrapolate_with(th.data(), x); });
      return std::array<double, 2>{bar.rho - avg[0], bar.E - avg[1]};
    };
    auto df = [&](const std::array<double, 2> &x, int dir) {
      double eps = 1e-6 * std::abs(x[dir]);
      std::array<double, 2> xp = x, xm = x;
      xp[dir] = x[dir] + 0.5 * eps;
      xm[dir] = x[dir] - 0.5 * eps;
      auto fp = f(xp), fm = f(xm);
      return std::array<double, 2>{(fp[0] - fm[0]) / eps, (fp[1] - fm[1]) / eps};
    };
    auto inv_df = [&](const std::array<double, 2> &x, const std::array<double, 2> &fx) {
      auto d0 = df(x, 0), d1 = df(x, 1);
      double inv_det = 1.0 / (d0[0] * d1[1] - d0[1] * d1[0]);
      return std::array<double, 2>{inv_det * (d1[1] * fx[0] - d1[0] * fx[1]),
                                   inv_det * (-d0[1] * fx[0] + d0[0] * fx[1])};
    };
    std::array<double, 2> atol{1e-13 * guess[0], 1e-13 * guess[1]};
    std::array<double, 2> x;
    std::vector<std::array<double, 2>> hist;
    const QNStatus st = quasi_newton<2>(f, inv_df, guess, atol, 20, x, &hist);
    if (st == QN_NOT_CONVERGING) {
      std::string msg = "quasi_newton: Is not converging. guess=(" + std::to_string(guess[0]) + "," +
                        std::to_string(guess[1]) + ") dx:";
      char buf[64];
      for (const auto &h : hist) {
        snprintf(buf, sizeof(buf), " (%.3e,%.3e)", h[0], h[1]);
        msg += buf;
      }
      throw Error(msg);
    }
    theta[0] = x[0];
    theta[1] = x[1];
    found = st == QN_CONVERGED;
  }

  RhoE extrapolate(const XYZ &x) const {
    if (kind == 0 || !found) return {0.0, 0.0};
    auto r = extrapolate_with(theta, x);
    return {r[0], r[1]};
  }
  EqPoint extrapolate_full(const XYZ &x) const {
    if (kind == 0 || !found) return {};
    auto r = extrapolate_with(theta, x);
    EqPoint e;
    e.rho = r[0];
    e.E = r[1];
    e.p = r[1] * (eos->gamma - 1.0);
    e.a = std::sqrt(eos->gamma * e.p / e.rho);
    return e;
  }
  RhoE extrapolate_cell(const DenormalizedRule &cell) const {
    if (kind == 0 || !found) return {0.0, 0.0};
    auto avg = cell_average(cell, [&](const XYZ &x) { return extrapolate_with(theta, x); });
    return {avg[0], avg[1]};
  }
};

// ---------------------------------------------------------------------------
// PolyND<35, 5> (include/zisa/math/poly2d_impl.hpp, polynomial_expr.hpp)
struct Poly {
  int degree = 0;
  int n_dims = 2;
  double a[MAX_COEFFS * NVARS];
  double c[MAX_COEFFS];
  XYZ center{};
  double length = 1.0;
  Poly() {
    std::fill(a, a + MAX_COEFFS * NVARS, 0.0);
    std::fill(c, c + MAX_COEFFS, 0.0);
  }
  Vars operator()(const XYZ &xyz) const {
    XYZ X = (xyz - center) / length;
    Vars px{0, 0, 0, 0, 0};
    double pxk = 1.0;
    int d = degree;
    for (int kx = 0; kx <= d; ++kx) {
      double pyk = 1.0;
      for (int ky = 0; ky <= d - kx; ++ky) {
        if (n_dims == 2) {
          int i = poly_index2(kx, ky);
          for (int k = 0; k < NVARS; ++k) px[k] += a[i * NVARS + k] * (pxk * pyk - c[i]);
        } else {
          double pzk = 1.0;
          for (int kz = 0; kz <= d - kx - ky; ++kz) {
            int i = poly_index3(kx, ky, kz);
            for (int k = 0; k < NVARS; ++k) px[k] += a[i * NVARS + k] * (pxk * pyk * pzk - c[i]);
            pzk *= X[2];
          }
        }
        pyk *= X[1];
      }
      pxk *= X[0];
    }
    return px;
  }
};

// result metadata of a pointwise binary op: larger degree wins, ties -> e2
static void binary_meta(Poly &out, const Poly &e1, const Poly &e2) {
  const Poly &m = (e1.degree > e2.degree ? e1 : e2);
  std::copy(m.c, m.c + MAX_COEFFS, out.c);
  out.center = m.center;
  out.length = m.length;
  out.degree = std::max(e1.degree, e2.degree);
  out.n_dims = std::max(e1.n_dims, e2.n_dims);
}

// include/zisa/math/poly2d_impl.hpp:325-337
static Vars smoothness_indicator(const Poly &p) {
  Vars beta{0, 0, 0, 0, 0};
  int n = poly_dof(p.degree, p.n_dims);
  for (int i = 1; i < n; ++i)
    for (int k = 0; k < NVARS; ++k) beta[k] += p.a[i * NVARS + k] * p.a[i * NVARS + k];
  return beta;
}

// ---------------------------------------------------------------------------
// LSQSolver (src/zisa/reconstruction/lsq_solver.cpp:40-85)
struct LSQ {
  int order = 1, rows = 1, cols = 1;
  std::vector<double> AT;  // cols × rows
  LDLT ldlt;
};

// ---------------------------------------------------------------------------
// LocalReconstruction (include/zisa/reconstruction/local_reconstruction.hpp)
struct LocalRC {
  int i = 0;
  const StencilFamily *fam = nullptr;
  std::vector<LSQ> lsq;
  std::vector<double> linear_weights;
  double epsilon = 1e-6, exponent = 4.0;
  int mode = 0;  // 0 CWENO-AO, 1 WENO-AO
  int scaling = 0;  // 0 EulerScaling, 1 UnityScaling
  LocalEquilibrium eq;
  Vars scale{0, 0, 0, 0, 0};
  std::vector<RhoE> rhoEbar;
  Poly poly;
  int steps_since_recompute = 0;
  int steps_per_recompute = 1;
  double recompute_threshold = 1e300;
};

struct Solver::Impl {
  EOS eos;
  Gravity gravity;
  std::vector<LocalRC> rc;
  std::vector<int> interior_cells, exterior_cells, interior_faces, exterior_faces;
  std::vector<double> steady;  // frozen BC values (n_cells × 5) or empty
  int n_threads = 1;
};

Solver::~Solver() = default;

static void build_lsq(const Grid &grid, const Stencil &st, LSQ &out) {
  out.order = st.order;
  auto A = assemble_weno_ao_matrix(grid, st.global, st.size, st.order, out.rows, out.cols);
  out.AT.assign(size_t(out.cols) * out.rows, 0.0);
  for (int r = 0; r < out.rows; ++r)
    for (int c = 0; c < out.cols; ++c) out.AT[size_t(c) * out.rows + r] = A[size_t(r) * out.cols + c];
  if (st.order == 1) return;
  std::vector<double> AtA(size_t(out.cols) * out.cols, 0.0);
  for (int p = 0; p < out.cols; ++p)
    for (int q = 0; q < out.cols; ++q) {
      double s = 0.0;
      for (int r = 0; r < out.rows; ++r) s += A[size_t(r) * out.cols + p] * A[size_t(r) * out.cols + q];
      AtA[size_t(p) * out.cols + q] = s;
    }
  out.ldlt.compute(AtA, out.cols);
}

Solver::Solver(std::shared_ptr<Grid> grid, const Params &params,
               const std::vector<StencilFamily> *stencils_override)
    : grid_(std::move(grid)), params_(params), impl_(new Impl) {
  const Grid &g = *grid_;
  impl_->eos.gamma = params.gamma;
  impl_->eos.R = params.gas_constant;
  impl_->gravity.kind = params.gravity;
  impl_->gravity.g = params.gravity_g;
  impl_->gravity.axis = {params.gravity_axis[0], params.gravity_axis[1], params.gravity_axis[2]};
  impl_->gravity.rhoC = params.polytrope_rhoC;
  impl_->gravity.K = params.polytrope_K;
  impl_->gravity.G = params.polytrope_G;
  families_ = stencils_override ? *stencils_override : compute_stencil_families(g, params);

  // make_reconstruction_array (include/zisa/reconstruction/global_reconstruction_decl.hpp:105-147)
  impl_->rc.resize(g.n_cells);
  // cells are independent: built on the host's threads, identical to a serial loop
  std::string setup_err;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < g.n_cells; ++i) {
   try {
    LocalRC &r = impl_->rc[i];
    r.i = i;
    r.fam = &families_[i];
    r.epsilon = params.epsilon;
    r.exponent = params.exponent;
    r.mode = params.reconstruction;
    r.scaling = params.scaling;
    r.steps_per_recompute = params.steps_per_recompute;
    r.recompute_threshold = params.recompute_threshold;
    const bool first_order = (r.fam->stencils.size() == 1 && r.fam->order == 1);
    std::vector<double> lw;
    if (first_order) lw = {1.0};
    else lw.assign(params.linear_weights, params.linear_weights + params.n_stencils);
    // src/zisa/reconstruction/hybrid_weno.cpp:15-31
    double tot = 0.0;
    for (double w : lw) tot += w;
    for (double w : lw) r.linear_weights.push_back(w / tot);
    r.lsq.resize(r.fam->stencils.size());
    for (size_t k = 0; k < r.fam->stencils.size(); ++k) build_lsq(g, r.fam->stencils[k], r.lsq[k]);
    r.eq.kind = params.equilibrium;
    r.eq.eos = &impl_->eos;
    r.eq.gravity = &impl_->gravity;
    r.rhoEbar.assign(r.fam->l2g.size(), RhoE{});
   } catch (const std::exception &e) {
#pragma omp critical
    if (setup_err.empty()) setup_err = "cell " + std::to_string(i) + ": " + e.what();
   }
  }
  if (!setup_err.empty()) throw Error(setup_err);

  // FluxLoop constructor: interior/exterior split (include/zisa/fvm_loops/flux_loop.hpp:29-90)
  for (int i = 0; i < g.n_cells; ++i) {
    bool ext = false;
    for (const auto &m : families_[i].l2g)
      if (g.cell_flags[m.cell].ghost_cell) ext = true;
    (ext ? impl_->exterior_cells : impl_->interior_cells).push_back(i);
  }
  auto contains = [](const std::vector<int> &v, int i) {
    return std::lower_bound(v.begin(), v.end(), i) != v.end();
  };
  for (int e = 0; e < g.n_interior_edges; ++e) {
    auto [iL, iR] = g.left_right[e];
    bool is_interior = !contains(impl_->exterior_cells, iL) && !contains(impl_->exterior_cells, iR);
    if (is_interior) {
      impl_->interior_faces.push_back(e);
    } else if (!(g.cell_flags[iL].ghost_cell && g.cell_flags[iR].ghost_cell)) {
      impl_->exterior_faces.push_back(e);
    }
  }
}

void Solver::set_steady_state(const double *u) {
  impl_->steady.assign(u, u + size_t(grid_->n_cells) * 5);
}

// LocalReconstruction::compute_equilibrium / recompute_equilibrium / compute
// (include/zisa/reconstruction/local_reconstruction.hpp:69-120) followed by
// CWENO_AO::reconstruct_impl (src/zisa/reconstruction/cweno_ao.cpp:36-53) or
// WENO_AO::reconstruct (src/zisa/reconstruction/weno_ao.cpp:14-35).
static void compute_local(const Grid &g, const EOS &eos, LocalRC &r, std::vector<Vars> &u_local) {
  const auto &l2g = r.fam->l2g;
  // recompute_equilibrium
  auto compute_equilibrium = [&]() {
    const Vars &u0 = u_local[0];
    RhoE self{u0[0], internal_energy(u0)};
    if (r.scaling == 1) {
      // UnityScaling (include/zisa/model/characteristic_scale.hpp:39-46)
      r.scale = {1.0, 1.0, 1.0, 1.0, 1.0};
    } else {
      // EulerScaling (include/zisa/model/characteristic_scale.hpp:23-32)
      if (self.rho <= 0.0) throw Error("Invalid density.");
      if (self.E <= 0.0) throw Error("Invalid energy.");
      double p = self.E * (eos.gamma - 1.0);
      double cs = std::sqrt(eos.gamma * p / self.rho);
      r.scale = {self.rho, cs, cs, cs, self.E};
    }
    r.eq.solve(self, g.cells[r.i]);
    for (size_t il = 0; il < l2g.size(); ++il) {
      // the equilibrium of cell i averaged over stencil member il (in i's frame)
      DenormalizedRule cq = g.cells[l2g[il].cell];
      for (auto &x : cq.points) x = x + l2g[il].shift;
      r.rhoEbar[il] = r.eq.extrapolate_cell(cq);
    }
    r.steps_since_recompute = 0;
  };
  if (r.steps_since_recompute % r.steps_per_recompute == 0) {
    compute_equilibrium();
  } else {
    const Vars &u0 = u_local[0];
    double rho = u0[0], E = internal_energy(u0);
    double d0 = (rho - r.rhoEbar[0].rho) / r.scale[0], d1 = (E - r.rhoEbar[0].E) / r.scale[4];
    if (std::sqrt(d0 * d0 + d1 * d1) >= r.recompute_threshold) compute_equilibrium();
  }

  for (size_t il = 0; il < l2g.size(); ++il) {
    u_local[il][0] -= r.rhoEbar[il].rho;
    u_local[il][4] -= r.rhoEbar[il].E;
    for (int v = 0; v < NVARS; ++v) u_local[il][v] = u_local[il][v] / r.scale[v];
  }

  // HybridWENO::compute_polys_impl (src/zisa/reconstruction/hybrid_weno.cpp:72-92)
  const auto &fam = *r.fam;
  const int ns = int(fam.stencils.size());
  std::vector<Poly> polys(ns);
  const XYZ xc = g.cell_centers[r.i];
  const double len = g.characteristic_length[r.i];
  for (int k = 0; k < ns; ++k) {
    const Stencil &st = fam.stencils[k];
    const LSQ &ls = r.lsq[k];
    Poly &p = polys[k];
    p.n_dims = g.n_dims;
    p.center = xc;
    p.length = len;
    if (st.order == 1) {
      p.degree = 0;
    } else {
      p.degree = st.order - 1;
      const auto &mom = g.normalized_moments[r.i];
      for (size_t q = 3; q < mom.size() && q < size_t(MAX_COEFFS); ++q) p.c[q] = mom[q];
      // coeffs = LDLT.solve(A^T rhs), rhs(ig, v) = q(local(ig+1), v) - q(0, v)
      std::vector<double> rhs(size_t(ls.rows) * NVARS);
      for (int ig = 0; ig < ls.rows; ++ig) {
        int il = st.local[ig + 1];
        for (int v = 0; v < NVARS; ++v) rhs[size_t(ig) * NVARS + v] = u_local[il][v] - u_local[0][v];
      }
      std::vector<double> b(size_t(ls.cols) * NVARS, 0.0);
      for (int c = 0; c < ls.cols; ++c)
        for (int v = 0; v < NVARS; ++v) {
          double s = 0.0;
          for (int rr = 0; rr < ls.rows; ++rr) s += ls.AT[size_t(c) * ls.rows + rr] * rhs[size_t(rr) * NVARS + v];
          b[size_t(c) * NVARS + v] = s;
        }
      ls.ldlt.solve(b, NVARS);
      for (int c = 0; c < ls.cols; ++c)
        for (int v = 0; v < NVARS; ++v) p.a[(c + 1) * NVARS + v] = b[size_t(c) * NVARS + v];
    }
    for (int v = 0; v < NVARS; ++v) p.a[v] = u_local[0][v];
  }

  if (r.mode == 0) {
    // CWENO: p_hi <- (p_hi - sum_{k != hi} c_k p_k) / c_hi
    int kh = fam.k_high;
    for (int k = 0; k < ns; ++k) {
      if (k == kh) continue;
      Poly t;
      binary_meta(t, polys[kh], polys[k]);
      for (int q = 0; q < MAX_COEFFS * NVARS; ++q) t.a[q] = polys[kh].a[q] - r.linear_weights[k] * polys[k].a[q];
      polys[kh] = t;
    }
    double inv = 1.0 / r.linear_weights[kh];
    for (int q = 0; q < MAX_COEFFS * NVARS; ++q) polys[kh].a[q] = inv * polys[kh].a[q];
  }

  // eno_hybridize (src/zisa/reconstruction/hybrid_weno.cpp:110-127)
  std::vector<double> al(ns);
  double al_tot = 0.0;
  for (int k = 0; k < ns; ++k) {
    Vars beta = smoothness_indicator(polys[k]);
    double IS = beta[0];
    for (int v = 1; v < NVARS; ++v) IS = std::max(IS, beta[v]);
    al[k] = r.linear_weights[k] / (r.epsilon + std::pow(IS, r.exponent));
    al_tot += al[k];
  }
  Poly p;
  p.degree = 0;
  p.n_dims = 2;
  p.center = {0, 0, 0};
  p.length = 1.0;
  for (int k = 0; k < ns; ++k) {
    double w = al[k] / al_tot;
    Poly t;
    binary_meta(t, p, polys[k]);
    for (int q = 0; q < MAX_COEFFS * NVARS; ++q) t.a[q] = p.a[q] + w * polys[k].a[q];
    p = t;
  }
  r.poly = p;
  ++r.steps_since_recompute;
}

// LocalReconstruction::operator() (local_reconstruction.hpp:141-156)
static Vars rc_eval(const LocalRC &r, const XYZ &x) {
  Vars d = r.poly(x);
  auto bg = r.eq.extrapolate_full(x);
  Vars u{bg.rho, 0.0, 0.0, 0.0, bg.E};
  for (int v = 0; v < NVARS; ++v) u[v] = u[v] + r.scale[v] * d[v];
  return u;
}
static Vars rc_delta(const LocalRC &r, const XYZ &x) {
  Vars d = r.poly(x);
  for (int v = 0; v < NVARS; ++v) d[v] = r.scale[v] * d[v];
  return d;
}

Vars Solver::eval_reconstruction(int i, const XYZ &x) const { return rc_eval(impl_->rc[i], x); }

static Vars load(const double *u, int i) {
  return {u[5 * size_t(i)], u[5 * size_t(i) + 1], u[5 * size_t(i) + 2], u[5 * size_t(i) + 3], u[5 * size_t(i) + 4]};
}

void Solver::rate(const double *u, double *dudt, double /*t*/) {
  const Grid &g = *grid_;
  Impl &im = *impl_;
  const int nb = g.max_neighbours;

  // ZeroRateOfChange (src/zisa/ode/rate_of_change.cpp:49-54)
  std::fill(dudt, dudt + size_t(g.n_cells) * 5, 0.0);

  // FluxLoop::compute -> compute_patch(interior), compute_patch(exterior)
  // (include/zisa/fvm_loops/flux_loop.hpp:96-195)
  auto patch = [&](const std::vector<int> &cells, const std::vector<int> &edges) {
    // EulerGlobalReconstruction::compute (global_reconstruction_impl.hpp:133-163)
    std::string err;
#pragma omp parallel for schedule(static, 8) if (im.n_threads > 1)
    for (size_t ii = 0; ii < cells.size(); ++ii) {
      int i = cells[ii];
      LocalRC &r = im.rc[i];
      std::vector<Vars> u_local(r.fam->l2g.size());
      for (size_t q = 0; q < u_local.size(); ++q) u_local[q] = load(u, r.fam->l2g[q].cell);
      try {
        compute_local(g, im.eos, r, u_local);
      } catch (const std::exception &ex) {
#pragma omp critical
        if (err.empty()) err = "cell " + std::to_string(i) + ": " + ex.what();
      }
    }
    if (!err.empty()) throw Error(err);
#pragma omp parallel for schedule(static, 8) if (im.n_threads > 1)
    for (size_t ie = 0; ie < edges.size(); ++ie) {
      int e = edges[ie];
      const Face &face = g.faces[e];
      auto [iL, iR] = g.left_right[e];
      const XYZ sh = g.edge_shift[e];
      Vars nf{0, 0, 0, 0, 0};
      for (size_t k = 0; k < face.qr.weights.size(); ++k) {
        double w = face.qr.weights[k];
        XYZ x = face.qr.points[k];
        Vars uL = rc_eval(im.rc[iL], x);
        Vars uR = rc_eval(im.rc[iR], x - sh);
        coord_transform(uL, face);
        coord_transform(uR, face);
        Vars f = params_.flux == 0 ? hllc_flux(im.eos, uL, uR) : rusanov_flux(im.eos, uL, uR);
        for (int v = 0; v < 5; ++v) nf[v] = nf[v] + w * f[v];
      }
      inv_coord_transform(nf, face);
      for (int v = 0; v < 5; ++v) {
        double nfL = nf[v] / g.volumes[iL];
#pragma omp atomic
        dudt[5 * size_t(iL) + v] -= nfL;
        double nfR = nf[v] / g.volumes[iR];
#pragma omp atomic
        dudt[5 * size_t(iR) + v] += nfR;
      }
    }
  };
  patch(im.interior_cells, im.interior_faces);
  patch(im.exterior_cells, im.exterior_faces);

  // GravitySourceLoop (include/zisa/fvm_loops/gravity_source_loop.hpp:32-148)
  if (im.gravity.kind != 0) {
#pragma omp parallel for schedule(static, 8) if (im.n_threads > 1)
    for (int i = 0; i < g.n_cells; ++i) {
      const LocalRC &r = im.rc[i];
      const DenormalizedRule &cell = g.cells[i];
      Vars s{0, 0, 0, 0, 0};
      if (params_.equilibrium != 0) {
        const XYZ x_cell = g.cell_centers[i];
        for (int k = 0; k < nb; ++k) {
          const Face &face = g.face(i, k);
          XYZ sh = g.face_to_cell_shift(i, k);
          XYZ n = sign(dot(face.normal, face.qr.points[0] + sh - x_cell)) * face.normal;
          auto s_eq = [&](const XYZ &x) {
            double p_eq = r.eq.extrapolate_full(x).p;
            return Vars{0.0, p_eq * n[0], p_eq * n[1], p_eq * n[2], 0.0};
          };
          Vars ret = s_eq(face.qr.points[0] + sh);
          for (int v = 0; v < 5; ++v) ret[v] = face.qr.weights[0] * ret[v];
          for (size_t q = 1; q < face.qr.weights.size(); ++q) {
            Vars t = s_eq(face.qr.points[q] + sh);
            for (int v = 0; v < 5; ++v) ret[v] = ret[v] + face.qr.weights[q] * t[v];
          }
          for (int v = 0; v < 5; ++v) s[v] += 1.0 * ret[v];
        }
        auto s_delta = [&](const XYZ &x) {
          auto bg = r.eq.extrapolate_full(x);
          Vars du = rc_delta(r, x);
          Vars uu{bg.rho + du[0], 0.0 + du[1], 0.0 + du[2], 0.0 + du[3], bg.E + du[4]};
          double drho = du[0];
          XYZ gp = im.gravity.grad_phi(x);
          return Vars{0.0, -drho * gp[0], -drho * gp[1], -drho * gp[2],
                      -(uu[1] * gp[0] + uu[2] * gp[1] + uu[3] * gp[2])};
        };
        Vars ret = s_delta(cell.points[0]);
        for (int v = 0; v < 5; ++v) ret[v] = cell.weights[0] * ret[v];
        for (size_t q = 1; q < cell.weights.size(); ++q) {
          Vars t = s_delta(cell.points[q]);
          for (int v = 0; v < 5; ++v) ret[v] = ret[v] + cell.weights[q] * t[v];
        }
        for (int v = 0; v < 5; ++v) s[v] += 1.0 * ret[v];
        for (int v = 0; v < 5; ++v) dudt[5 * size_t(i) + v] += s[v] / cell.volume;
      } else {
        auto src = [&](const XYZ &x) {
          Vars uu = rc_eval(r, x);
          XYZ gp = im.gravity.grad_phi(x);
          return Vars{0.0, -uu[0] * gp[0], -uu[0] * gp[1], -uu[0] * gp[2],
                      -(uu[1] * gp[0] + uu[2] * gp[1] + uu[3] * gp[2])};
        };
        Vars ret = src(cell.points[0]);
        for (int v = 0; v < 5; ++v) ret[v] = cell.weights[0] * ret[v];
        for (size_t q = 1; q < cell.weights.size(); ++q) {
          Vars t = src(cell.points[q]);
          for (int v = 0; v < 5; ++v) ret[v] = ret[v] + cell.weights[q] * t[v];
        }
        for (int v = 0; v < 5; ++v) dudt[5 * size_t(i) + v] += (1.0 * ret[v]) / cell.volume;
      }
    }
  }

  // FluxBC (include/zisa/boundary/flux_bc.hpp:25-42)
  if (params_.flux_bc == 1) {
    for (int e = g.n_interior_edges; e < g.n_edges; ++e) {
      int i = g.left_right[e].first;
      const Face &face = g.faces[e];
      Vars uu = load(u, i);
      coord_transform(uu, face);
      Vars f = euler_flux(uu, pressure(im.eos, uu));
      inv_coord_transform(f, face);
      double fac = face.qr.volume / g.volumes[i];
      for (int v = 0; v < 5; ++v) dudt[5 * size_t(i) + v] -= fac * f[v];
    }
  }

  // EquilibriumFluxBC<IsentropicEquilibrium> (include/zisa/boundary/equilibrium_flux_bc.hpp:36-62):
  // per exterior face, the isentropic equilibrium of the owner cell solved
  // from its average (LocalEquilibrium::solve -> solve_exact), evaluated at
  // the face points as a state at rest, whose Euler flux in the face frame is
  // (0, p, 0, 0, 0); rotated back, integrated with the face rule and
  // subtracted with 1/vol.  An equilibrium that is not found extrapolates to
  // zeros, whose pressure is NaN (0/0), as in the reference.
  if (params_.flux_bc == 2) {
    for (int e = g.n_interior_edges; e < g.n_edges; ++e) {
      int i = g.left_right[e].first;
      const Face &face = g.faces[e];
      const Vars ui = load(u, i);
      LocalEquilibrium eq;
      eq.kind = 1;
      eq.eos = &im.eos;
      eq.gravity = &im.gravity;
      eq.solve(RhoE{ui[0], internal_energy(ui)}, g.cells[i]);
      auto flux = [&](const XYZ &x) {
        RhoE r = eq.extrapolate(x);
        Vars uu{r.rho, 0.0, 0.0, 0.0, r.E};
        Vars f = euler_flux(uu, pressure(im.eos, uu));
        inv_coord_transform(f, face);
        return f;
      };
      Vars f = flux(face.qr.points[0]);
      for (int v = 0; v < 5; ++v) f[v] = face.qr.weights[0] * f[v];
      for (size_t q = 1; q < face.qr.weights.size(); ++q) {
        Vars t = flux(face.qr.points[q]);
        for (int v = 0; v < 5; ++v) f[v] = f[v] + face.qr.weights[q] * t[v];
      }
      for (int v = 0; v < 5; ++v) dudt[5 * size_t(i) + v] -= f[v] / g.volumes[i];
    }
  }
}

// RungeKutta::compute_step + runge_kutta_sum + FrozenBC::apply
// (src/zisa/ode/runge_kutta.cpp:87-143, 154-165;
//  src/zisa/boundary/frozen_boundary_condition.cpp:39-55)
// Butcher tableau "ssp3" (src/zisa/ode/runge_kutta.cpp:154-165)
const Tableau &ssp3_tableau() {
  static const Tableau t = {3,
                            {{0.0, 0.0, 0.0, 0.0}, {1.0, 0.0, 0.0, 0.0}, {0.25, 0.25, 0.0, 0.0}, {0, 0, 0, 0}},
                            {1.0 / 6, 1.0 / 6, 2.0 / 3, 0.0},
                            {0.0, 1.0, 0.5, 0.0}};
  return t;
}

// RungeKutta::compute_step + runge_kutta_sum
// (src/zisa/ode/runge_kutta.cpp:87-143): one accumulator per entry, zero
// coefficients skipped, u_s = u0 + dt * dudt, the BC after every sum.
void rk_step(const Tableau &tab, size_t n, double *u, double t, double dt,
             const std::function<void(const double *, double *, double)> &rate_fn,
             const std::function<void(double *)> &bc_fn) {
  std::vector<double> u0(u, u + n), ux(n);
  std::vector<std::vector<double>> k(tab.n_stages, std::vector<double>(n));
  auto rk_sum = [&](const double *coeffs) {
    for (size_t q = 0; q < n; ++q) {
      double dudt = 0.0;
      for (int s = 0; s < tab.n_stages; ++s)
        if (coeffs[s] != 0.0) dudt += coeffs[s] * k[s][q];
      ux[q] = u0[q] + dt * dudt;
    }
  };
  std::fill(k[0].begin(), k[0].end(), 0.0);
  rate_fn(u0.data(), k[0].data(), t);
  for (int s = 1; s < tab.n_stages; ++s) {
    rk_sum(tab.a[s]);
    bc_fn(ux.data());
    std::fill(k[s].begin(), k[s].end(), 0.0);
    rate_fn(ux.data(), k[s].data(), t + tab.c[s] * dt);
  }
  rk_sum(tab.b);
  bc_fn(ux.data());
  std::copy(ux.begin(), ux.end(), u);
}

void Solver::step(double *u, double t, double dt) {
  const size_t n = size_t(grid_->n_cells) * 5;
  auto bc = [&](double *w) {
    // FrozenBC::apply (src/zisa/boundary/frozen_boundary_condition.cpp:39-55)
    if (impl_->steady.empty()) return;
    for (int i = 0; i < grid_->n_cells; ++i)
      if (grid_->cell_flags[i].ghost_cell)
        for (int v = 0; v < 5; ++v) w[5 * size_t(i) + v] = impl_->steady[5 * size_t(i) + v];
  };
  rk_step(ssp3_tableau(), n, u, t, dt, [&](const double *x, double *k, double tt) { rate(x, k, tt); }, bc);
}

// LocalCFL (include/zisa/model/local_cfl_condition_impl.hpp:25-40)
double Solver::cfl_dt(const double *u, double cfl) const {
  double m = 1e300;
  for (int i = 0; i < grid_->n_cells; ++i) {
    Vars uu = load(u, i);
    double v2 = (uu[1] * uu[1] + uu[2] * uu[2] + uu[3] * uu[3]) / (uu[0] * uu[0]);
    double ev = std::sqrt(v2) + sound_speed(impl_->eos, uu);
    m = std::min(m, grid_->inradii[i] / ev);
  }
  return cfl * m;
}

// TimeLoop::operator() (src/zisa/fvm_loops/time_loop.cpp:36-75): a step of
// the picked dt, the step-rejection check, on rejection reject_step
// (:182-196: retry from u0 with the factor times the rejected dt), else the
// clock advances and the next dt is picked from the new state
// (pick_time_step, :77-81: factor * LocalCFL).  RejectLargeDensityChange
// (src/zisa/ode/step_rejection.cpp:17-57): a step is bad when any cell's
// density changes by more than drho_crit_rel times its old value; the factor
// halves after a bad step, doubles after two good steps in a row, and stays
// within [2^-max_exponent, 1].  The simulation clock (its library is not in
// the reference's tree) clamps every step to the end time; a step that no
// longer advances the time, or is not positive, ends the loop.
double Solver::advance(double *u, double cfl, double t0, double t_end, int max_steps, double drho_crit_rel,
                       int max_exponent, int *steps, int *rejected) {
  const size_t n = size_t(grid_->n_cells) * 5;
  const bool reject = drho_crit_rel > 0.0;
  const double min_factor = reject ? std::pow(2.0, -double(max_exponent)) : 1.0;
  double factor = 1.0;
  bool was_good = true;
  double t = t0;
  int k = 0, r = 0;
  if (max_steps <= 0) max_steps = 0x7fffffff;
  auto clamp = [&](double dt) { return dt > t_end - t ? t_end - t : dt; };
  std::vector<double> u0(n);
  double dt = clamp(factor * cfl_dt(u, cfl));
  while (t < t_end && k < max_steps) {
    if (!(dt > 0.0) || t + dt == t) break;
    std::copy(u, u + n, u0.begin());
    step(u, t, dt);
    if (reject) {
      bool good = true;
      for (int i = 0; i < grid_->n_cells; ++i)
        if (std::abs(u0[5 * size_t(i)] - u[5 * size_t(i)]) > drho_crit_rel * u0[5 * size_t(i)]) good = false;
      if (!good) {
        was_good = false;
        factor /= 2.0;
      } else {
        if (was_good) factor *= 2.0;
        was_good = true;
      }
      factor = std::max(min_factor, std::min(factor, 1.0));
      if (!good) {
        ++r;
        dt = clamp(factor * dt);
        std::copy(u0.begin(), u0.end(), u);
        continue;
      }
    }
    t += dt;
    ++k;
    dt = clamp(factor * cfl_dt(u, cfl));
  }
  if (steps) *steps = k;
  if (rejected) *rejected = r;
  return t;
}

}  // namespace zfv_ref

// n_threads control used by the CPU-baseline timing (OpenMP like the reference)
namespace zfv_ref {
void Solver::set_threads(int n) { impl_->n_threads = n < 1 ? 1 : n; }
}  // namespace zfv_ref

// ---------------------------------------------------------------------------
// Entry points of the oracle's pieces for the known-answer tests of the
// reference's unit tests (test/zisa/unit_test/math/{quasi_newton,poly2d}.cpp,
// test/zisa/unit_test/model/local_equilibrium.cpp).  Test infrastructure.
namespace {
zfv_ref::EOS kat_eos(const zfv_params *p) {
  zfv_ref::EOS e;
  e.gamma = p->gamma;
  e.R = p->gas_constant;
  return e;
}
zfv_ref::Gravity kat_gravity(const zfv_params *p) {
  zfv_ref::Gravity g;
  g.kind = p->gravity;
  g.g = p->gravity_g;
  g.axis = {p->gravity_axis[0], p->gravity_axis[1], p->gravity_axis[2]};
  g.rhoC = p->polytrope_rhoC;
  g.K = p->polytrope_K;
  g.G = p->polytrope_G;
  return g;
}
zfv_ref::DenormalizedRule kat_cell(int n_dims, const double *corners, int deg) {
  zfv_ref::XYZ v[4];
  for (int k = 0; k <= n_dims; ++k) v[k] = {corners[3 * k], corners[3 * k + 1], corners[3 * k + 2]};
  return zfv_ref::simplex_cell_rule(n_dims, v, deg);
}
thread_local std::string kat_err;
}  // namespace

extern "C" {

const char *zfv_ref_kat_error(void) { return kat_err.c_str(); }

// quasi_newton on an n-dimensional system given by callbacks:
// f(x, fx) and inv_df(x, fx, dx).  Returns the status (0 converged,
// 1 not converged, 2 not converging) and the result in x.
typedef void (*zfv_ref_qn_f)(const double *x, double *fx);
typedef void (*zfv_ref_qn_idf)(const double *x, const double *fx, double *dx);
int zfv_ref_quasi_newton(int n, zfv_ref_qn_f f, zfv_ref_qn_idf inv_df, const double *x0, const double *atol,
                         int max_iter, double *x, int *n_steps) {
  auto run = [&](auto tag) {
    constexpr int N = decltype(tag)::value;
    std::array<double, N> a0, at, xr;
    for (int k = 0; k < N; ++k) {
      a0[k] = x0[k];
      at[k] = atol[k];
    }
    auto ff = [&](const std::array<double, N> &y) {
      std::array<double, N> r;
      f(y.data(), r.data());
      return r;
    };
    auto idf = [&](const std::array<double, N> &y, const std::array<double, N> &fy) {
      std::array<double, N> d;
      inv_df(y.data(), fy.data(), d.data());
      return d;
    };
    std::vector<std::array<double, N>> steps;
    int st = zfv_ref::quasi_newton<N>(ff, idf, a0, at, max_iter, xr, &steps);
    for (int k = 0; k < N; ++k) x[k] = xr[k];
    if (n_steps) *n_steps = int(steps.size());
    return st;
  };
  if (n == 1) return run(std::integral_constant<int, 1>{});
  if (n == 2) return run(std::integral_constant<int, 2>{});
  kat_err = "quasi_newton: n must be 1 or 2";
  return -1;
}

// LocalEquilibrium::solve (solve_exact) of p->equilibrium on one simplex
// (n_dims + 1 corners, xyz each) with its cell rule of degree deg.
int zfv_ref_eq_solve(const zfv_params *p, int n_dims, const double *corners, int deg, double rho_bar,
                     double E_bar, double *theta, double *x_ref, int *found) {
  try {
    zfv_ref::EOS eos = kat_eos(p);
    zfv_ref::Gravity grav = kat_gravity(p);
    zfv_ref::LocalEquilibrium eq;
    eq.kind = p->equilibrium;
    eq.eos = &eos;
    eq.gravity = &grav;
    eq.solve(zfv_ref::RhoE{rho_bar, E_bar}, kat_cell(n_dims, corners, deg));
    theta[0] = eq.theta[0];
    theta[1] = eq.theta[1];
    for (int d = 0; d < 3; ++d) x_ref[d] = eq.x_ref[d];
    *found = eq.found ? 1 : 0;
    return 0;
  } catch (const std::exception &e) {
    kat_err = e.what();
    return -1;
  }
}

// Equilibrium.extrapolate(theta, x_ref, x) at n points (rho, E per point)
int zfv_ref_eq_extrapolate(const zfv_params *p, const double *theta, const double *x_ref, int n, const double *x,
                           double *rhoE) {
  zfv_ref::EOS eos = kat_eos(p);
  zfv_ref::Gravity grav = kat_gravity(p);
  zfv_ref::LocalEquilibrium eq;
  eq.kind = p->equilibrium;
  eq.eos = &eos;
  eq.gravity = &grav;
  eq.theta[0] = theta[0];
  eq.theta[1] = theta[1];
  eq.x_ref = {x_ref[0], x_ref[1], x_ref[2]};
  eq.found = true;
  for (int q = 0; q < n; ++q) {
    auto r = eq.extrapolate({x[3 * q], x[3 * q + 1], x[3 * q + 2]});
    rhoE[2 * q] = r.rho;
    rhoE[2 * q + 1] = r.E;
  }
  return 0;
}

// average(cell, extrapolate) over one simplex with its cell rule of degree deg
int zfv_ref_eq_average(const zfv_params *p, const double *theta, const double *x_ref, int n_dims,
                       const double *corners, int deg, double *rhoE) {
  zfv_ref::EOS eos = kat_eos(p);
  zfv_ref::Gravity grav = kat_gravity(p);
  zfv_ref::LocalEquilibrium eq;
  eq.kind = p->equilibrium;
  eq.eos = &eos;
  eq.gravity = &grav;
  eq.theta[0] = theta[0];
  eq.theta[1] = theta[1];
  eq.x_ref = {x_ref[0], x_ref[1], x_ref[2]};
  eq.found = true;
  auto r = eq.extrapolate_cell(kat_cell(n_dims, corners, deg));
  rhoE[0] = r.rho;
  rhoE[1] = r.E;
  return 0;
}

// PolyND evaluation: coefficients a[alpha][var] (dof x 5), moments c[alpha]
int zfv_ref_poly_eval(int n_dims, int degree, const double *coeffs, const double *moments, const double *center,
                      double length, const double *x, double *out) {
  zfv_ref::Poly p;
  p.n_dims = n_dims;
  p.degree = degree;
  const int dof = zfv_ref::poly_dof(degree, n_dims);
  for (int i = 0; i < dof * zfv_ref::NVARS; ++i) p.a[i] = coeffs[i];
  for (int i = 0; i < dof; ++i) p.c[i] = moments[i];
  p.center = {center[0], center[1], center[2]};
  p.length = length;
  auto v = p({x[0], x[1], x[2]});
  for (int k = 0; k < zfv_ref::NVARS; ++k) out[k] = v[k];
  return 0;
}

// sum_k w_k p_k formed as the reconstruction forms its WENO sum (starting
// from the zero polynomial of degree 0 in 2D), each partial sum taking the
// metadata of the larger-degree operand, ties to the right one; evaluated at x
int zfv_ref_poly_lincomb(int n_polys, const int *n_dims, const int *degrees, const double *coeffs,
                         const double *moments, const double *weights, const double *x, double *out) {
  zfv_ref::Poly acc;
  acc.degree = 0;
  acc.n_dims = 2;
  for (int j = 0; j < n_polys; ++j) {
    zfv_ref::Poly p;
    p.n_dims = n_dims[j];
    p.degree = degrees[j];
    const int dof = zfv_ref::poly_dof(degrees[j], n_dims[j]);
    for (int i = 0; i < dof * zfv_ref::NVARS; ++i) p.a[i] = coeffs[j * zfv_ref::MAX_COEFFS * zfv_ref::NVARS + i];
    for (int i = 0; i < dof; ++i) p.c[i] = moments[j * zfv_ref::MAX_COEFFS + i];
    zfv_ref::Poly t;
    zfv_ref::binary_meta(t, acc, p);
    for (int q = 0; q < zfv_ref::MAX_COEFFS * zfv_ref::NVARS; ++q) t.a[q] = acc.a[q] + weights[j] * p.a[q];
    acc = t;
  }
  auto v = acc({x[0], x[1], x[2]});
  for (int k = 0; k < zfv_ref::NVARS; ++k) out[k] = v[k];
  return 0;
}

}  // extern "C"
