// zfv_ref — CPU oracle for the rate-of-change hot path of ZisaFVM.
//
// TEST INFRASTRUCTURE ONLY.  This is a from-spec C++ restatement of the
// reference's CPU rate of change (reconstruction + face fluxes + gravity
// source + flux BC) and its SSP3 Runge-Kutta step.  It exists to check the
// HIP implementation in `zisafvm_amd/`; only `tests/`, `__graft_entry__.smoke()`
// and the CPU-baseline leg of `bench.py` may load it.  The product never links
// it.
//
// Structure deliberately mirrors the reference: a `Grid` object, one
// `LocalReconstruction` per cell holding an `HybridWENO` with per-stencil LDLT
// factorisations that are *re-solved on every call* (no precomputed
// pseudo-inverses), a face loop that processes interior edges in edge order and
// scatters into the tendency, and a per-cell gravity source loop.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root).
//
// Parity status per part (how each is pinned) is listed in tests/test_oracle_*.py.
// Parts without a reference known-answer test are pinned by closed forms,
// invariants or brute force; Eigen's LDLT rounding is restated from Eigen's
// published algorithm (diagonal pivoting), not bit-matched.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace zfv_ref {

using XYZ = std::array<double, 3>;
using Vars = std::array<double, 5>;  // euler_var_t: rho, mx, my, mz, E

constexpr int MAX_COEFFS = 35;  // PolyND<35, 5> (include/zisa/reconstruction/weno_poly.hpp:12)
constexpr int NVARS = 5;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// small vector helpers (restating Cartesian<3> arithmetic)
inline XYZ operator+(const XYZ &a, const XYZ &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline XYZ operator-(const XYZ &a, const XYZ &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline XYZ operator*(double s, const XYZ &a) { return {s * a[0], s * a[1], s * a[2]}; }
inline XYZ operator*(const XYZ &a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline XYZ operator/(const XYZ &a, double s) { return {a[0] / s, a[1] / s, a[2] / s}; }
inline double dot(const XYZ &a, const XYZ &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(const XYZ &a) { return std::sqrt(dot(a, a)); }
inline XYZ normalize(const XYZ &a) { return a / norm(a); }
inline XYZ cross(const XYZ &a, const XYZ &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double det3(const XYZ &a, const XYZ &b, const XYZ &c) { return dot(a, cross(b, c)); }
inline double sign(double x) { return double((0.0 < x) - (x < 0.0)); }

// ---------------------------------------------------------------------------
// Parameters (the JSON keys of the reference's config, flattened).
struct Params {
  int n_dims = 2;
  int n_stencils = 4;
  int orders[5] = {3, 2, 2, 2, 2};
  int biases[5] = {0, 1, 1, 1, 1};  // 0 = "c", 1 = "b"
  double overfit[5] = {2.0, 1.5, 1.5, 1.5, 1.5};
  double linear_weights[5] = {100.0, 1.0, 1.0, 1.0, 1.0};
  double epsilon = 1e-6;
  double exponent = 4.0;
  int reconstruction = 0;  // 0 CWENO-AO, 1 WENO-AO
  int flux = 0;            // 0 HLLC (Batten), 1 Rusanov (extension)
  int equilibrium = 0;     // 0 none, 1 isentropic, 2 isothermal (extension)
  int gravity = 0;         // 0 none, 1 axial constant, 2 radial constant, 3 polytrope
  double gravity_g = 0.0;
  double gravity_axis[3] = {0.0, 1.0, 0.0};
  double polytrope_rhoC = 1.0, polytrope_K = 1.0, polytrope_G = 1.0;
  double gamma = 1.4;
  double gas_constant = 1.0;
  int edge_deg = 3, volume_deg = 4, moments_deg = 5;
  int flux_bc = 0;  // 0 none, 1 FluxBC (constant extrapolation)
  int steps_per_recompute = 1;
  double recompute_threshold = 1e300;
  int scaling = 0;  // 0 EulerScaling, 1 UnityScaling
};

// ---------------------------------------------------------------------------
// Quadrature (src/zisa/math/{triangular_rule,tetrahedral_rule,edge_rule}.cpp)
struct RefRule {
  std::vector<double> weights;                // normalized
  std::vector<std::array<double, 4>> points;  // barycentric (edge: [xi, 0, 0, 0])
};
const RefRule &triangular_rule(int deg);
const RefRule &tetrahedral_rule(int deg);
const RefRule &edge_rule(int deg);

struct DenormalizedRule {
  std::vector<double> weights;
  std::vector<XYZ> points;
  double volume = 0.0;
};
// the cell rule of one simplex with n_dims + 1 corners
DenormalizedRule simplex_cell_rule(int n_dims, const XYZ *v, int deg);


struct Face {
  DenormalizedRule qr;
  XYZ normal{};
  XYZ t1{}, t2{};
};

struct CellFlags {
  bool interior = true;
  bool ghost_cell = false;
  bool ghost_cell_l1 = false;
};

// ---------------------------------------------------------------------------
// Grid (src/zisa/grid/grid.cpp:651-723).  Periodic meshes are an extension:
// each cell carries its own (unwrapped) corner coordinates, neighbours are
// found through shared vertex ids and `shift(i,k)` is the translation that maps
// neighbour(i,k)'s frame into cell i's frame.
struct Grid {
  int n_dims = 2;
  int n_cells = 0;
  int max_neighbours = 3;
  std::vector<int> vertex_ids;     // n_cells × (n_dims+1)
  std::vector<XYZ> corners;        // n_cells × (n_dims+1)
  std::vector<int> neighbours;     // n_cells × max_neighbours, -1 = invalid
  std::vector<XYZ> shifts;         // n_cells × max_neighbours
  int n_interior_edges = 0, n_exterior_edges = 0, n_edges = 0;
  std::vector<int> edge_indices;   // n_cells × max_neighbours
  std::vector<std::pair<int, int>> left_right;  // -1 = magic_index_value
  std::vector<XYZ> edge_shift;     // translation R-frame -> L-frame per edge
  std::vector<double> volumes, inradii, circum_radii, characteristic_length;
  std::vector<DenormalizedRule> cells;
  std::vector<XYZ> cell_centers;
  std::vector<Face> faces;
  std::vector<XYZ> face_centers;
  std::vector<std::vector<double>> normalized_moments;
  std::vector<CellFlags> cell_flags;

  XYZ corner(int i, int k) const { return corners[size_t(i) * (n_dims + 1) + k]; }
  bool is_valid(int i, int k) const { return neighbours[size_t(i) * max_neighbours + k] >= 0; }
  int neighbour(int i, int k) const { return neighbours[size_t(i) * max_neighbours + k]; }
  const Face &face(int i, int k) const { return faces[edge_indices[size_t(i) * max_neighbours + k]]; }
  // translation that maps a point of face(i,k) (stored in the frame of the
  // face's left cell) into cell i's frame.
  XYZ face_to_cell_shift(int i, int k) const;
};

std::shared_ptr<Grid> make_grid(int n_dims, int n_cells, const int *vertex_ids,
                                const double *corners, const uint8_t *ghost_mask,
                                int edge_deg, int volume_deg, int moments_deg);

// ---------------------------------------------------------------------------
// Stencils (src/zisa/reconstruction/stencil.cpp, stencil_family.cpp)
struct Member {
  int cell;
  XYZ shift;  // maps the member's frame into the centre cell's frame
};

struct Stencil {
  int max_order = 1, order = 1;
  int size = 1, max_size = 1;
  int bias = 0;
  double overfit = 1.0;
  std::vector<int> local;    // index into l2g
  std::vector<Member> global;
};

struct StencilFamily {
  std::vector<Stencil> stencils;
  std::vector<Member> l2g;
  int order = 1;
  int k_high = 0;
};

std::vector<StencilFamily> compute_stencil_families(const Grid &grid, const Params &params);
StencilFamily make_first_order_family(int i);
int poly_dof(int deg, int n_dims);
int poly_index2(int a, int b);
int poly_index3(int a, int b, int c);
int required_stencil_size(int deg, double factor, int n_dims);

// dense LSQ matrix A (rows × cols, row-major) of the WENO-AO fit
// (src/zisa/reconstruction/lsq_solver.cpp:168-403)
std::vector<double> assemble_weno_ao_matrix(const Grid &grid, const std::vector<Member> &stencil,
                                            int n_points, int order, int &rows, int &cols);
int matrix_rank(const std::vector<double> &A, int rows, int cols);

// Eigen::LDLT<MatrixXd> restated: diagonal pivoting, Lower storage.
struct LDLT {
  int n = 0;
  std::vector<double> mat;  // n×n, row-major; lower triangle holds L, diagonal holds D
  std::vector<int> transpositions;
  void compute(const std::vector<double> &M, int n);
  // solves in place for nrhs right-hand sides stored row-major n × nrhs
  void solve(std::vector<double> &b, int nrhs) const;
};

// ---------------------------------------------------------------------------
// Physics (include/zisa/model/*, include/zisa/flux/hllc.hpp)
struct EOS {
  double gamma = 1.4, R = 1.0;
};
struct Gravity {
  int kind = 0;
  double g = 0.0;
  XYZ axis{0.0, 1.0, 0.0};
  double rhoC = 1.0, K = 1.0, G = 1.0;
  double phi(const XYZ &x) const;
  XYZ grad_phi(const XYZ &x) const;
};

double kinetic_energy(const Vars &u);
double internal_energy(const Vars &u);
double pressure(const EOS &eos, const Vars &u);
double sound_speed(const EOS &eos, const Vars &u);
Vars euler_flux(const Vars &u, double p);
void coord_transform(Vars &u, const Face &face);
void inv_coord_transform(Vars &u, const Face &face);
Vars hllc_flux(const EOS &eos, const Vars &uL, const Vars &uR);
Vars rusanov_flux(const EOS &eos, const Vars &uL, const Vars &uR);

// ---------------------------------------------------------------------------
// Explicit Runge-Kutta in Butcher form (src/zisa/ode/runge_kutta.cpp)
struct Tableau {
  int n_stages;
  double a[4][4];
  double b[4];
  double c[4];
};
const Tableau &ssp3_tableau();
void rk_step(const Tableau &tab, size_t n, double *u, double t, double dt,
             const std::function<void(const double *, double *, double)> &rate_fn,
             const std::function<void(double *)> &bc_fn);

// ---------------------------------------------------------------------------
// The rate of change object (FluxLoop + GravitySourceLoop + FluxBC) and the
// SSP3 time integration with frozen ghost cells.
class Solver {
 public:
  Solver(std::shared_ptr<Grid> grid, const Params &params,
         const std::vector<StencilFamily> *stencils_override = nullptr);
  ~Solver();

  // RateOfChange::compute semantics on an n_cells × 5 row-major state:
  // tendency is *accumulated into* (SumRatesOfChange starts from zero).
  void rate(const double *u, double *dudt, double t);
  // one SSP3 step in Butcher form, frozen BC after every stage sum.
  void step(double *u, double t, double dt);
  double cfl_dt(const double *u, double cfl) const;
  // TimeLoop::operator() with LocalCFL and RejectNothing (drho_crit_rel <= 0)
  // or RejectLargeDensityChange(drho_crit_rel, max_exponent); every step
  // clamped to t_end; advances u in place, returns the final time
  double advance(double *u, double cfl, double t0, double t_end, int max_steps, double drho_crit_rel,
                 int max_exponent, int *steps, int *rejected);
  void set_steady_state(const double *u);
  void set_threads(int n);

  const std::vector<StencilFamily> &stencils() const { return families_; }
  const Grid &grid() const { return *grid_; }

  // reconstruction values for tests: the per-cell reconstruction evaluated at x
  Vars eval_reconstruction(int i, const XYZ &x) const;

  struct Impl;

 private:
  std::shared_ptr<Grid> grid_;
  Params params_;
  std::vector<StencilFamily> families_;
  std::unique_ptr<Impl> impl_;
};

}  // namespace zfv_ref
