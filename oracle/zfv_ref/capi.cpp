// C entry points of the oracle (loaded by tests and the bench's CPU-baseline
// leg through ctypes).  TEST INFRASTRUCTURE ONLY (see zfv_ref.hpp).
//
// The parameter and mesh structs reuse the layout of include/zfv.h so that a
// test can hand the same inputs to the oracle and to the HIP library.
#include "../../include/zfv.h"
#include "zfv_ref.hpp"

#include <cmath>
#include <cstring>
#include <string>

using namespace zfv_ref;

static thread_local std::string g_err;

static Params to_params(const zfv_params *p) {
  Params q;
  q.n_dims = p->n_dims;
  q.n_stencils = p->n_stencils;
  for (int k = 0; k < 5; ++k) {
    q.orders[k] = p->orders[k];
    q.biases[k] = p->biases[k];
    q.overfit[k] = p->overfit[k];
    q.linear_weights[k] = p->linear_weights[k];
  }
  q.epsilon = p->epsilon;
  q.exponent = p->exponent;
  q.reconstruction = p->reconstruction;
  q.flux = p->flux;
  q.equilibrium = p->equilibrium;
  q.gravity = p->gravity;
  q.gravity_g = p->gravity_g;
  for (int k = 0; k < 3; ++k) q.gravity_axis[k] = p->gravity_axis[k];
  q.polytrope_rhoC = p->polytrope_rhoC;
  q.polytrope_K = p->polytrope_K;
  q.polytrope_G = p->polytrope_G;
  q.gamma = p->gamma;
  q.gas_constant = p->gas_constant;
  q.edge_deg = p->edge_deg;
  q.volume_deg = p->volume_deg;
  q.moments_deg = p->moments_deg;
  q.flux_bc = p->flux_bc;
  q.steps_per_recompute = p->steps_per_recompute;
  q.recompute_threshold = p->recompute_threshold;
  q.scaling = p->scaling;
  return q;
}

#define GUARD(body)                   \
  try {                               \
    body;                             \
  } catch (const std::exception &e) { \
    g_err = e.what();                 \
    return -1;                        \
  }

extern "C" {

const char *zfv_ref_last_error(void) { return g_err.c_str(); }

int zfv_ref_create(const zfv_mesh *mesh, const zfv_params *p, void **out) {
  GUARD({
    Params q = to_params(p);
    auto grid = make_grid(mesh->n_dims, mesh->n_cells, mesh->vertex_ids, mesh->corners, mesh->ghost_mask,
                          q.edge_deg, q.volume_deg, q.moments_deg);
    *out = new Solver(grid, q);
    return 0;
  })
}

int zfv_ref_destroy(void *h) {
  delete static_cast<Solver *>(h);
  return 0;
}

int zfv_ref_rate(void *h, const double *u, double *dudt, double t) {
  GUARD({
    static_cast<Solver *>(h)->rate(u, dudt, t);
    return 0;
  })
}

int zfv_ref_step(void *h, double *u, double t, double dt) {
  GUARD({
    static_cast<Solver *>(h)->step(u, t, dt);
    return 0;
  })
}

int zfv_ref_cfl_dt(void *h, const double *u, double cfl, double *dt) {
  GUARD({
    *dt = static_cast<Solver *>(h)->cfl_dt(u, cfl);
    return 0;
  })
}

int zfv_ref_advance(void *h, double *u, double cfl, double t0, double t_end, int max_steps, double drho_crit_rel,
                    int max_exponent, double *t_final, int *steps, int *rejected) {
  GUARD({
    *t_final = static_cast<Solver *>(h)->advance(u, cfl, t0, t_end, max_steps, drho_crit_rel, max_exponent, steps,
                                                 rejected);
    return 0;
  })
}

int zfv_ref_set_steady_state(void *h, const double *u) {
  static_cast<Solver *>(h)->set_steady_state(u);
  return 0;
}

int zfv_ref_set_threads(void *h, int n) {
  static_cast<Solver *>(h)->set_threads(n);
  return 0;
}

// reconstruction of cell i evaluated at x (caller supplies 3 doubles)
int zfv_ref_eval(void *h, int i, const double *x, double *out5) {
  GUARD({
    Vars v = static_cast<Solver *>(h)->eval_reconstruction(i, {x[0], x[1], x[2]});
    for (int k = 0; k < 5; ++k) out5[k] = v[k];
    return 0;
  })
}

// geometry export: volumes, centers (3), inradii, char length, ghost flag
int zfv_ref_geometry(void *h, double *vol, double *centers, double *inradii, double *length,
                     uint8_t *ghost) {
  const Grid &g = static_cast<Solver *>(h)->grid();
  for (int i = 0; i < g.n_cells; ++i) {
    if (vol) vol[i] = g.volumes[i];
    if (centers)
      for (int k = 0; k < 3; ++k) centers[3 * i + k] = g.cell_centers[i][k];
    if (inradii) inradii[i] = g.inradii[i];
    if (length) length[i] = g.characteristic_length[i];
    if (ghost) ghost[i] = g.cell_flags[i].ghost_cell ? 1 : 0;
  }
  return 0;
}

// topology export: neighbours (-1 invalid) and edge indices per (cell, face),
// face normals per edge (3 doubles)
int zfv_ref_topology(void *h, int *nbr, int *edge_idx, double *normals) {
  const Grid &g = static_cast<Solver *>(h)->grid();
  for (int i = 0; i < g.n_cells; ++i)
    for (int k = 0; k < g.max_neighbours; ++k) {
      nbr[i * g.max_neighbours + k] = g.neighbour(i, k);
      edge_idx[i * g.max_neighbours + k] = g.edge_indices[size_t(i) * g.max_neighbours + k];
    }
  for (int e = 0; e < g.n_edges; ++e)
    for (int d = 0; d < 3; ++d) normals[3 * e + d] = g.faces[e].normal[d];
  return 0;
}

// counts: n_interior_edges, n_exterior_edges
int zfv_ref_counts(void *h, int *n_int, int *n_ext) {
  const Grid &g = static_cast<Solver *>(h)->grid();
  *n_int = g.n_interior_edges;
  *n_ext = g.n_exterior_edges;
  return 0;
}

// same layout as zfv_get_stencils (include/zfv.h)
int zfv_ref_get_stencils(void *h, int i, int *l2g, int cap, int *l2g_size, int *members, int members_cap) {
  const auto &fam = static_cast<Solver *>(h)->stencils()[i];
  *l2g_size = int(fam.l2g.size());
  for (int q = 0; q < int(fam.l2g.size()) && q < cap; ++q) l2g[q] = fam.l2g[q].cell;
  int w = 0;
  for (const auto &st : fam.stencils) {
    if (w + 2 + int(st.global.size()) > members_cap) return -1;
    members[w++] = st.order;
    members[w++] = int(st.global.size());
    for (const auto &m : st.global) members[w++] = m.cell;
  }
  return w;
}

// LSQ building blocks for known-answer tests
int zfv_ref_ldlt_solve(int n, const double *M, int nrhs, double *b) {
  GUARD({
    LDLT l;
    l.compute(std::vector<double>(M, M + size_t(n) * n), n);
    std::vector<double> bb(b, b + size_t(n) * nrhs);
    l.solve(bb, nrhs);
    std::memcpy(b, bb.data(), bb.size() * sizeof(double));
    return 0;
  })
}

int zfv_ref_rank(int rows, int cols, const double *A) {
  return matrix_rank(std::vector<double>(A, A + size_t(rows) * cols), rows, cols);
}

int zfv_ref_hllc(double gamma, const double *uL, const double *uR, double *f) {
  EOS eos;
  eos.gamma = gamma;
  Vars a{uL[0], uL[1], uL[2], uL[3], uL[4]}, b{uR[0], uR[1], uR[2], uR[3], uR[4]};
  Vars r = hllc_flux(eos, a, b);
  for (int k = 0; k < 5; ++k) f[k] = r[k];
  return 0;
}

int zfv_ref_rusanov(double gamma, const double *uL, const double *uR, double *f) {
  EOS eos;
  eos.gamma = gamma;
  Vars a{uL[0], uL[1], uL[2], uL[3], uL[4]}, b{uR[0], uR[1], uR[2], uR[3], uR[4]};
  Vars r = rusanov_flux(eos, a, b);
  for (int k = 0; k < 5; ++k) f[k] = r[k];
  return 0;
}

// SSP3 on the scalar ODEs of test/zisa/unit_test/ode/runge_kutta.cpp:
// kind 0: u' = u (u(0) = 1), kind 1: u' = t^2 (u(0) = 0), kind 2: u' = 1.
// Integrates to t_end with fixed dt, returns u(t_end).
int zfv_ref_rk_ode(int kind, double dt, double t_end, double *u_out) {
  double u = (kind == 0) ? 1.0 : 0.0;
  double t = 0.0;
  int steps = int(std::lround(t_end / dt));
  auto f = [kind](const double *x, double *k, double tt) {
    k[0] += (kind == 0) ? x[0] : (kind == 1 ? tt * tt : 1.0);
  };
  for (int s = 0; s < steps; ++s) {
    rk_step(ssp3_tableau(), 1, &u, t, dt, f, [](double *) {});
    t += dt;
  }
  *u_out = u;
  return 0;
}

// quadrature rule export: kind 0 edge, 1 triangle, 2 tetrahedron
int zfv_ref_rule(int kind, int deg, double *w, double *bary, int cap) {
  const RefRule &r = kind == 0 ? edge_rule(deg) : (kind == 1 ? triangular_rule(deg) : tetrahedral_rule(deg));
  int n = int(r.weights.size());
  for (int q = 0; q < n && q < cap; ++q) {
    w[q] = r.weights[q];
    for (int k = 0; k < 4; ++k) bary[4 * q + k] = r.points[q][k];
  }
  return n;
}

}  // extern "C"
