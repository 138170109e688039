// zfv_ref stencils, WENO-AO least-squares matrices and the LDLT solver.
// TEST INFRASTRUCTURE ONLY (see zfv_ref.hpp).
#include "zfv_ref.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>

namespace zfv_ref {

static int rel_vertex(int n_dims, int k, int rel) {
  if (n_dims == 2) return (k + rel) % 3;
  static const int tab[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  return tab[k][rel];
}
static int rel_off_vertex(int n_dims, int k) {
  if (n_dims == 2) return (k + 2) % 3;
  static const int tab[4] = {2, 3, 1, 0};
  return tab[k];
}

// src/zisa/reconstruction/stencil.cpp:168-175
int required_stencil_size(int deg, double factor, int n_dims) {
  if (deg == 0) return 1;
  return int(double(poly_dof(deg, n_dims) - 1) * factor + 1);
}

// src/zisa/reconstruction/stencil.cpp:159-166
static int deduce_max_order(int stencil_size, double factor, int n_dims) {
  int deg = 0;
  while (required_stencil_size(deg + 1, factor, n_dims) <= stencil_size) deg += 1;
  return deg + 1;
}

// ---------------------------------------------------------------------------
// Regions (src/zisa/math/cone.cpp)
struct Region {
  int kind = 0;  // 0 full sphere, 1 triangular cone, 2 tetrahedral cone
  XYZ A{}, dB{}, dC{}, dD{};
  bool is_inside(const XYZ &x) const {
    if (kind == 0) return true;
    XYZ dx = x - A;
    if (kind == 1) return cross(dB, dx)[2] >= 0.0 && cross(dx, dC)[2] >= 0.0;
    return det3(dB, dC, dx) >= 0.0 && det3(dC, dD, dx) >= 0.0 && det3(dD, dB, dx) >= 0.0;
  }
};

// src/zisa/reconstruction/stencil.cpp:178-190: query points are the maximal
// degree rule of the candidate cell.
static std::vector<XYZ> selection_points(const Grid &grid, int i) {
  const RefRule &r = grid.n_dims == 2 ? triangular_rule(5) : tetrahedral_rule(3);
  std::vector<XYZ> pts;
  const int nv = grid.n_dims + 1;
  for (size_t q = 0; q < r.weights.size(); ++q) {
    XYZ x = grid.corner(i, 0) * r.points[q][0];
    for (int k = 1; k < nv; ++k) x = x + grid.corner(i, k) * r.points[q][k];
    pts.push_back(x);
  }
  return pts;
}

static bool not_found(const std::vector<Member> &c, int cand) {
  for (const auto &m : c)
    if (m.cell == cand) return false;
  return true;
}

// src/zisa/reconstruction/stencil.cpp:192-236
static std::vector<Member> region_based_candidates(const Grid &grid, int i_center, int n_points,
                                                   const Region &region) {
  int max_points = 5 * n_points;
  std::vector<Member> candidates;
  candidates.push_back({i_center, {0.0, 0.0, 0.0}});

  auto is_inside = [&](int cand, const XYZ &shift) {
    for (const auto &x : selection_points(grid, cand))
      if (region.is_inside(x + shift)) return true;
    return region.is_inside(grid.cell_centers[cand] + shift);
  };

  for (int p = 0; p < max_points; ++p) {
    if (p >= int(candidates.size())) break;
    Member mj = candidates[p];
    for (int k = 0; k < grid.max_neighbours; ++k) {
      if (!grid.is_valid(mj.cell, k)) continue;
      int cand = grid.neighbour(mj.cell, k);
      XYZ shift = mj.shift + grid.shifts[size_t(mj.cell) * grid.max_neighbours + k];
      if (not_found(candidates, cand) && is_inside(cand, shift)) candidates.push_back({cand, shift});
    }
  }
  return candidates;
}

// src/zisa/reconstruction/stencil.cpp:238-254.  The reference sorts with
// std::sort; ties are broken here by BFS order (stable sort) so that the
// selection is deterministic.
static std::vector<Member> region_based_stencil(const Grid &grid, int i_center, int n_points,
                                                const Region &region) {
  auto candidates = region_based_candidates(grid, i_center, n_points, region);
  const XYZ xc = grid.cell_centers[i_center];
  std::vector<double> dist(candidates.size());
  for (size_t q = 0; q < candidates.size(); ++q)
    dist[q] = norm(grid.cell_centers[candidates[q].cell] + candidates[q].shift - xc);
  std::vector<size_t> order(candidates.size());
  std::iota(order.begin(), order.end(), size_t(0));
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return dist[a] < dist[b]; });
  std::vector<Member> out;
  for (size_t q = 0; q < order.size() && int(out.size()) < n_points; ++q) out.push_back(candidates[order[q]]);
  return out;
}

// src/zisa/reconstruction/stencil.cpp:256-275
static Region make_cone(const Grid &grid, int i_center, const XYZ &apex, int k) {
  Region r;
  r.A = apex;
  XYZ B = grid.corner(i_center, rel_vertex(grid.n_dims, k, 0));
  XYZ C = grid.corner(i_center, rel_vertex(grid.n_dims, k, 1));
  r.dB = B - apex;
  r.dC = C - apex;
  if (grid.n_dims == 2) {
    r.kind = 1;
  } else {
    r.kind = 2;
    r.dD = grid.corner(i_center, rel_vertex(grid.n_dims, k, 2)) - apex;
  }
  return r;
}

// ---------------------------------------------------------------------------
// WENO-AO matrix (src/zisa/reconstruction/lsq_solver.cpp:168-403)

static void assemble_2d(std::vector<double> &A, int rows, int cols, const Grid &grid,
                        const std::vector<Member> &st, int order) {
  if (order == 1) return;
  const int nd = grid.n_dims;
  auto idx = [nd](int i, int j) { return nd == 2 ? poly_index2(i, j) : poly_index3(i, j, 0); };
  int i0 = st[0].cell;
  XYZ x0 = grid.cell_centers[i0];
  double l0 = grid.characteristic_length[i0];
  const auto &C0 = grid.normalized_moments[i0];
  auto at = [&](int r, int c) -> double & { return A[size_t(r) * cols + c]; };
  for (int ii = 0; ii < rows; ++ii) {
    const Member &mj = st[ii + 1];
    int j = mj.cell;
    XYZ X = (grid.cell_centers[j] + mj.shift - x0) / l0;
    double x_10 = X[0], x_01 = X[1];
    double lj = grid.characteristic_length[j] / l0;
    const auto &Cj = grid.normalized_moments[j];
    if (order >= 2) {
      at(ii, idx(1, 0) - 1) = x_10;
      at(ii, idx(0, 1) - 1) = x_01;
    }
    if (order >= 3) {
      int i_20 = idx(2, 0), i_11 = idx(1, 1), i_02 = idx(0, 2);
      double x_20 = x_10 * x_10, x_11 = x_10 * x_01, x_02 = x_01 * x_01;
      double lj_2 = lj * lj;
      at(ii, i_20 - 1) = x_20 - C0[i_20] + lj_2 * Cj[i_20];
      at(ii, i_11 - 1) = x_11 - C0[i_11] + lj_2 * Cj[i_11];
      at(ii, i_02 - 1) = x_02 - C0[i_02] + lj_2 * Cj[i_02];
      if (order >= 4) {
        int i_30 = idx(3, 0), i_21 = idx(2, 1), i_12 = idx(1, 2), i_03 = idx(0, 3);
        double x_30 = x_20 * x_10, x_21 = x_20 * x_01, x_12 = x_11 * x_01, x_03 = x_02 * x_01;
        double lj_3 = lj_2 * lj;
        at(ii, i_30 - 1) = x_30 - C0[i_30] + 3.0 * x_10 * lj_2 * Cj[i_20] + lj_3 * Cj[i_30];
        at(ii, i_21 - 1) = x_21 - C0[i_21] + x_01 * lj_2 * Cj[i_20] + 2.0 * x_10 * lj_2 * Cj[i_11] +
                           lj_3 * Cj[i_21];
        at(ii, i_12 - 1) = x_12 - C0[i_12] + x_10 * lj_2 * Cj[i_02] + 2.0 * x_01 * lj_2 * Cj[i_11] +
                           lj_3 * Cj[i_12];
        at(ii, i_03 - 1) = x_03 - C0[i_03] + 3.0 * x_01 * lj_2 * Cj[i_02] + lj_3 * Cj[i_03];
        if (order >= 5) {
          int i_40 = idx(4, 0), i_31 = idx(3, 1), i_22 = idx(2, 2), i_13 = idx(1, 3), i_04 = idx(0, 4);
          double x_40 = x_30 * x_10, x_31 = x_30 * x_01, x_22 = x_21 * x_01, x_13 = x_12 * x_01,
                 x_04 = x_03 * x_01;
          double lj_4 = lj_3 * lj;
          at(ii, i_40 - 1) = x_40 - C0[i_40] + 6.0 * x_20 * lj_2 * Cj[i_20] +
                             4.0 * x_10 * lj_3 * Cj[i_30] + lj_4 * Cj[i_40];
          at(ii, i_31 - 1) = x_31 - C0[i_31] + 3 * x_11 * lj_2 * Cj[i_20] + 3.0 * x_20 * lj_2 * Cj[i_11] +
                             x_01 * lj_3 * Cj[i_30] + 3.0 * x_10 * lj_3 * Cj[i_21] + lj_4 * Cj[i_31];
          at(ii, i_22 - 1) = x_22 - C0[i_22] + x_02 * lj_2 * Cj[i_20] + x_20 * lj_2 * Cj[i_02] +
                             4 * x_11 * lj_2 * Cj[i_11] + 2.0 * x_01 * lj_3 * Cj[i_21] +
                             2 * x_10 * lj_3 * Cj[i_12] + lj_4 * Cj[i_22];
          at(ii, i_13 - 1) = x_13 - C0[i_13] + 3 * x_11 * lj_2 * Cj[i_02] + 3.0 * x_02 * lj_2 * Cj[i_11] +
                             x_10 * lj_3 * Cj[i_03] + 3.0 * x_01 * lj_3 * Cj[i_12] + lj_4 * Cj[i_13];
          at(ii, i_04 - 1) = x_04 - C0[i_04] + 6.0 * x_02 * lj_2 * Cj[i_02] +
                             4.0 * x_01 * lj_3 * Cj[i_03] + lj_4 * Cj[i_04];
        }
      }
    }
    if (order >= 6) throw Error("2D WENO-AO matrix: order >= 6 not implemented");
  }
}

static void assemble_3d(std::vector<double> &A, int rows, int cols, const Grid &grid,
                        const std::vector<Member> &st, int order) {
  if (order == 1) return;
  int i0 = st[0].cell;
  XYZ x0 = grid.cell_centers[i0];
  double l0 = grid.characteristic_length[i0];
  const auto &C0 = grid.normalized_moments[i0];
  auto at = [&](int r, int c) -> double & { return A[size_t(r) * cols + c]; };
  for (int ii = 0; ii < rows; ++ii) {
    const Member &mj = st[ii + 1];
    int j = mj.cell;
    XYZ X = (grid.cell_centers[j] + mj.shift - x0) / l0;
    double x = X[0], y = X[1], z = X[2];
    double lj = grid.characteristic_length[j] / l0;
    const auto &Cj = grid.normalized_moments[j];
    if (order >= 2) at(ii, poly_index3(0, 0, 1) - 1) = z;
    if (order >= 3) {
      int i_002 = poly_index3(0, 0, 2), i_101 = poly_index3(1, 0, 1), i_011 = poly_index3(0, 1, 1);
      double lj_2 = lj * lj;
      at(ii, i_002 - 1) = z * z - C0[i_002] + lj_2 * Cj[i_002];
      at(ii, i_101 - 1) = x * z - C0[i_101] + lj_2 * Cj[i_101];
      at(ii, i_011 - 1) = y * z - C0[i_011] + lj_2 * Cj[i_011];
      if (order >= 4) {
        int i_003 = poly_index3(0, 0, 3), i_102 = poly_index3(1, 0, 2), i_012 = poly_index3(0, 1, 2);
        int i_201 = poly_index3(2, 0, 1), i_111 = poly_index3(1, 1, 1), i_021 = poly_index3(0, 2, 1);
        int i_200 = poly_index3(2, 0, 0), i_020 = poly_index3(0, 2, 0), i_110 = poly_index3(1, 1, 0);
        double x_003 = z * z * z, x_102 = x * z * z, x_012 = y * z * z;
        double x_201 = x * x * z, x_111 = x * y * z, x_021 = y * y * z;
        double lj_3 = lj * lj * lj;
        at(ii, i_003 - 1) = x_003 - C0[i_003] + 3.0 * z * lj_2 * Cj[i_002] + lj_3 * Cj[i_003];
        at(ii, i_102 - 1) =
            x_102 - C0[i_102] + x * lj_2 * Cj[i_002] + 2.0 * z * lj_2 * Cj[i_101] + lj_3 * Cj[i_102];
        at(ii, i_012 - 1) =
            x_012 - C0[i_012] + y * lj_2 * Cj[i_002] + 2.0 * z * lj_2 * Cj[i_011] + lj_3 * Cj[i_012];
        at(ii, i_201 - 1) =
            x_201 - C0[i_201] + z * lj_2 * Cj[i_200] + 2.0 * x * lj_2 * Cj[i_101] + lj_3 * Cj[i_201];
        at(ii, i_021 - 1) =
            x_021 - C0[i_021] + z * lj_2 * Cj[i_020] + 2.0 * y * lj_2 * Cj[i_011] + lj_3 * Cj[i_021];
        at(ii, i_111 - 1) = x_111 - C0[i_111] + z * lj_2 * Cj[i_110] + y * lj_2 * Cj[i_101] +
                            x * lj_2 * Cj[i_011] + lj_3 * Cj[i_111];
      }
      if (order >= 5) throw Error("3D WENO-AO matrix: order >= 5 not implemented");
    }
  }
}

std::vector<double> assemble_weno_ao_matrix(const Grid &grid, const std::vector<Member> &stencil,
                                            int n_points, int order, int &rows, int &cols) {
  if (order <= 0) throw Error("non-positive order");
  if (order == 1) {
    rows = cols = 1;
    return {1.0};
  }
  rows = n_points - 1;
  cols = poly_dof(order - 1, grid.n_dims) - 1;
  std::vector<double> A(size_t(rows) * cols, 0.0);
  assemble_2d(A, rows, cols, grid, stencil, order);
  if (grid.n_dims == 3) assemble_3d(A, rows, cols, grid, stencil, order);
  return A;
}

// Singular values by one-sided Jacobi; rank as Eigen's JacobiSVD::rank():
// count sigma_i >= max(sigma_0 * min(m,n) * eps, DBL_MIN).
int matrix_rank(const std::vector<double> &A0, int rows, int cols) {
  std::vector<double> A = A0;
  auto col = [&](int r, int c) -> double & { return A[size_t(r) * cols + c]; };
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < cols; ++p)
      for (int q = p + 1; q < cols; ++q) {
        double a = 0, b = 0, c = 0;
        for (int r = 0; r < rows; ++r) {
          a += col(r, p) * col(r, p);
          b += col(r, q) * col(r, q);
          c += col(r, p) * col(r, q);
        }
        if (std::abs(c) <= 1e-300 || std::abs(c) <= 1e-17 * std::sqrt(a * b)) continue;
        off = std::max(off, std::abs(c) / std::sqrt(a * b));
        double zeta = (b - a) / (2.0 * c);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < rows; ++r) {
          double xp = col(r, p), xq = col(r, q);
          col(r, p) = cs * xp - sn * xq;
          col(r, q) = sn * xp + cs * xq;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> sv(cols);
  for (int c = 0; c < cols; ++c) {
    double s = 0;
    for (int r = 0; r < rows; ++r) s += col(r, c) * col(r, c);
    sv[c] = std::sqrt(s);
  }
  std::sort(sv.begin(), sv.end(), std::greater<double>());
  int diag = std::min(rows, cols);
  if (diag == 0) return 0;
  double thr = std::max(sv[0] * diag * 2.220446049250313e-16, 2.2250738585072014e-308);
  int rank = 0;
  for (int k = 0; k < diag; ++k)
    if (sv[k] >= thr) ++rank;
  return rank;
}

// ---------------------------------------------------------------------------
// Eigen::LDLT (Lower, diagonal pivoting) restated from Eigen 3.3's
// ldlt_inplace<Lower>::unblocked and LDLT::_solve_impl.
void LDLT::compute(const std::vector<double> &M, int n_) {
  n = n_;
  mat = M;
  transpositions.assign(n, 0);
  auto m = [&](int r, int c) -> double & { return mat[size_t(r) * n + c]; };
  std::vector<double> temp(n);
  for (int k = 0; k < n; ++k) {
    int big = k;
    double bigv = std::abs(m(k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::abs(m(i, i)) > bigv) {
        bigv = std::abs(m(i, i));
        big = i;
      }
    transpositions[k] = big;
    if (k != big) {
      for (int j = 0; j < k; ++j) std::swap(m(k, j), m(big, j));
      for (int i = big + 1; i < n; ++i) std::swap(m(i, k), m(i, big));
      std::swap(m(k, k), m(big, big));
      for (int i = k + 1; i < big; ++i) {
        double tmp = m(i, k);
        m(i, k) = m(big, i);
        m(big, i) = tmp;
      }
    }
    int rs = n - k - 1;
    if (k > 0) {
      for (int j = 0; j < k; ++j) temp[j] = m(j, j) * m(k, j);
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += m(k, j) * temp[j];
      m(k, k) -= s;
      for (int i = k + 1; i < k + 1 + rs; ++i) {
        double t = 0.0;
        for (int j = 0; j < k; ++j) t += m(i, j) * temp[j];
        m(i, k) -= t;
      }
    }
    double akk = m(k, k);
    bool valid = std::abs(akk) > 0.0;
    if (k == 0 && !valid) {
      for (int j = 0; j < n; ++j) transpositions[j] = j;
      break;
    }
    if (rs > 0 && valid)
      for (int i = k + 1; i < n; ++i) m(i, k) /= akk;
  }
}

void LDLT::solve(std::vector<double> &b, int nrhs) const {
  auto m = [&](int r, int c) { return mat[size_t(r) * n + c]; };
  auto row_swap = [&](int a, int c) {
    if (a == c) return;
    for (int v = 0; v < nrhs; ++v) std::swap(b[size_t(a) * nrhs + v], b[size_t(c) * nrhs + v]);
  };
  for (int k = 0; k < n; ++k) row_swap(k, transpositions[k]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      for (int v = 0; v < nrhs; ++v) b[size_t(i) * nrhs + v] -= m(i, j) * b[size_t(j) * nrhs + v];
  const double tol = 2.2250738585072014e-308;
  for (int i = 0; i < n; ++i) {
    double d = m(i, i);
    for (int v = 0; v < nrhs; ++v) {
      if (std::abs(d) > tol) b[size_t(i) * nrhs + v] /= d;
      else b[size_t(i) * nrhs + v] = 0.0;
    }
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j)
      for (int v = 0; v < nrhs; ++v) b[size_t(i) * nrhs + v] -= m(j, i) * b[size_t(j) * nrhs + v];
  for (int k = n - 1; k >= 0; --k) row_swap(k, transpositions[k]);
}

// ---------------------------------------------------------------------------
// Biased stencils (src/zisa/reconstruction/stencil.cpp:277-399)

static bool is_good(const Grid &grid, const std::vector<Member> &s, int n_points, int order) {
  int rows, cols;
  auto A = assemble_weno_ao_matrix(grid, s, n_points, order, rows, cols);
  return matrix_rank(A, rows, cols) == cols;
}

// deterministic stand-in for the reference's std::random_device shuffle
static uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static std::vector<Member> tryhard_stencil(const Grid &grid, int i_center, int k, int n_points,
                                           int order) {
  int e = grid.edge_indices[size_t(i_center) * grid.max_neighbours + k];
  XYZ apex = grid.face_centers[e] + grid.face_to_cell_shift(i_center, k);
  Region region = make_cone(grid, i_center, apex, k);
  auto candidates = region_based_candidates(grid, i_center, n_points, region);
  if (int(candidates.size()) < n_points) return {candidates[0]};
  if (!is_good(grid, candidates, int(candidates.size()), order))
    throw Error("It is impossible to find a suitable stencil.");
  uint64_t seed = uint64_t(i_center) * 8u + uint64_t(k);
  for (int iter = 0; iter < 100; ++iter) {
    int n = int(candidates.size()) - 1;
    for (int m = n - 1; m >= 1; --m) {
      int r = int(splitmix64(seed) % uint64_t(m + 1));
      std::swap(candidates[1 + m], candidates[1 + r]);
    }
    if (is_good(grid, candidates, n_points, order)) {
      candidates.resize(n_points);
      return candidates;
    }
  }
  throw Error("Did not find a suitable stencil given the candidates.");
}

static std::vector<Member> biased_stencil(const Grid &grid, int i_center, int k, int n_points, int order) {
  const XYZ &off = grid.corner(i_center, rel_off_vertex(grid.n_dims, k));
  auto s = region_based_stencil(grid, i_center, n_points, make_cone(grid, i_center, off, k));
  if (int(s.size()) == n_points && is_good(grid, s, n_points, order)) return s;
  s = region_based_stencil(grid, i_center, n_points,
                           make_cone(grid, i_center, grid.cell_centers[i_center], k));
  if (int(s.size()) == n_points && is_good(grid, s, n_points, order)) return s;
  return tryhard_stencil(grid, i_center, k, n_points, order);
}

// src/zisa/reconstruction/stencil.cpp:77-95 (assign_local_indices)
static void assign_local_indices(Stencil &st, const std::vector<Member> &members,
                                 std::vector<Member> &l2g) {
  st.global = members;
  st.local.resize(members.size());
  for (size_t q = 0; q < members.size(); ++q) {
    int pos = -1;
    for (size_t r = 0; r < l2g.size(); ++r)
      if (l2g[r].cell == members[q].cell) pos = int(r);
    if (pos < 0) {
      pos = int(l2g.size());
      l2g.push_back(members[q]);
    }
    st.local[q] = pos;
  }
}

// src/zisa/reconstruction/stencil_family.cpp:15-52, 121-136
static StencilFamily make_family(const Grid &grid, int i, const int *orders, const int *biases,
                                 const double *overfit, int n_stencils) {
  StencilFamily f;
  int k_biased = 0;
  const int nd = grid.n_dims;
  for (int s = 0; s < n_stencils; ++s) {
    Stencil st;
    st.max_order = orders[s];
    st.bias = biases[s];
    st.overfit = overfit[s];
    st.max_size = required_stencil_size(st.max_order - 1, st.overfit, nd);
    std::vector<Member> members;
    if (biases[s] == 1) {
      members = biased_stencil(grid, i, k_biased, st.max_size, st.max_order);
      ++k_biased;
    } else {
      Region full;
      members = region_based_stencil(grid, i, st.max_size, full);
    }
    assign_local_indices(st, members, f.l2g);
    st.order = deduce_max_order(int(st.local.size()), st.overfit, nd);
    st.size = required_stencil_size(st.order - 1, st.overfit, nd);
    f.stencils.push_back(st);
  }
  f.order = 1;
  for (auto &st : f.stencils) f.order = std::max(f.order, st.order);
  int k_high = 0;
  for (int s = 1; s < int(f.stencils.size()); ++s) {
    int ko = f.stencils[k_high].order;
    if (f.stencils[s].order > ko) k_high = s;
    else if (f.stencils[s].order == ko && f.stencils[s].bias == 0) k_high = s;
  }
  f.k_high = k_high;
  return f;
}

StencilFamily make_first_order_family(int i) {
  int o[1] = {1}, b[1] = {0};
  double f[1] = {1.0};
  (void)o;
  StencilFamily fam;
  Stencil st;
  st.max_order = 1;
  st.order = 1;
  st.size = 1;
  st.max_size = 1;
  st.bias = 0;
  st.overfit = f[0];
  st.global = {{i, {0.0, 0.0, 0.0}}};
  st.local = {0};
  (void)b;
  fam.stencils.push_back(st);
  fam.l2g = {{i, {0.0, 0.0, 0.0}}};
  fam.order = 1;
  fam.k_high = 0;
  return fam;
}

// src/zisa/reconstruction/stencil_family.cpp:96-119
// (cells are independent: the loop runs on the host's threads; every cell's
// family is the same as in a serial loop)
std::vector<StencilFamily> compute_stencil_families(const Grid &grid, const Params &params) {
  std::vector<StencilFamily> out(grid.n_cells);
  std::string err;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < grid.n_cells; ++i) {
    try {
      const auto &fl = grid.cell_flags[i];
      if (fl.interior || fl.ghost_cell_l1) {
        out[i] = make_family(grid, i, params.orders, params.biases, params.overfit, params.n_stencils);
      } else {
        // StencilFamilyParams{{1}, {"c"}, {1.0}}: a one-cell central stencil
        int o[1] = {1}, b[1] = {0};
        double f[1] = {1.0};
        out[i] = make_family(grid, i, o, b, f, 1);
      }
    } catch (const std::exception &e) {
#pragma omp critical
      if (err.empty()) err = "cell " + std::to_string(i) + ": " + e.what();
    }
  }
  if (!err.empty()) throw Error(err);
  return out;
}

}  // namespace zfv_ref
