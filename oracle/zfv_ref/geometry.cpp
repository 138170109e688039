// zfv_ref geometry: quadrature rules, simplex metrics and the Grid.
// TEST INFRASTRUCTURE ONLY (see zfv_ref.hpp).
#include "zfv_ref.hpp"

#include <algorithm>
#include <climits>
#include <map>

namespace zfv_ref {

// ---------------------------------------------------------------------------
// Quadrature tables.

// Gauss-Legendre on [-1, 1], ascending nodes.  The reference computes these at
// run time by Fourier-Newton (include/zisa/math/gauss_legendre.hpp:45-170); the
// values below are the closed forms / published tables it converges to.
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  if (n == 1) {
    x = {0.0};
    w = {2.0};
  } else if (n == 2) {
    double a = 1.0 / std::sqrt(3.0);
    x = {-a, a};
    w = {1.0, 1.0};
  } else if (n == 3) {
    double a = std::sqrt(3.0 / 5.0);
    x = {-a, 0.0, a};
    w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  } else if (n == 4) {
    double a = 0.33998104358485626480, b = 0.86113631159405257522;
    double wa = 0.65214515486254614263, wb = 0.34785484513745385737;
    x = {-b, -a, a, b};
    w = {wb, wa, wa, wb};
  } else {
    throw Error("gauss_legendre: unsupported number of points");
  }
}

// src/zisa/math/edge_rule.cpp:10-31: deg -> deg/2 + 1 points, weights 0.5*GL.
const RefRule &edge_rule(int deg) {
  static std::map<int, RefRule> cache;
  if (deg == -1) deg = 7;
  auto it = cache.find(deg);
  if (it != cache.end()) return it->second;
  RefRule r;
  std::vector<double> x, w;
  gauss_legendre(deg / 2 + 1, x, w);
  for (size_t k = 0; k < x.size(); ++k) {
    r.weights.push_back(0.5 * w[k]);
    r.points.push_back({x[k], 0.0, 0.0, 0.0});
  }
  return cache[deg] = r;
}

// src/zisa/math/triangular_rule.cpp:115-136 (permutate).  n == 3 is unused for
// deg <= 5.
static void permutate(RefRule &r, double w, const std::vector<double> &l) {
  if (l.size() == 1) {
    r.weights.push_back(w);
    r.points.push_back({l[0], l[0], l[0], 0.0});
  } else if (l.size() == 2) {
    for (int k = 0; k < 3; ++k) r.weights.push_back(w);
    r.points.push_back({l[0], l[1], l[1], 0.0});
    r.points.push_back({l[1], l[0], l[1], 0.0});
    r.points.push_back({l[1], l[1], l[0], 0.0});
  } else {
    throw Error("permutate: unsupported");
  }
}

// src/zisa/math/triangular_rule.cpp:33-91 (Dunavant, degrees 1-5).
const RefRule &triangular_rule(int deg) {
  static std::map<int, RefRule> cache;
  if (deg == -1) deg = 5;
  auto it = cache.find(deg);
  if (it != cache.end()) return it->second;
  RefRule r;
  if (deg == 1) {
    permutate(r, 1.0, {1.0 / 3.0});
  } else if (deg == 2) {
    permutate(r, 1.0 / 3.0, {2.0 / 3.0, 1.0 / 6.0});
  } else if (deg == 3) {
    permutate(r, -0.5625, {1.0 / 3.0});
    permutate(r, 1.5625 / 3.0, {0.6, 0.2});
  } else if (deg == 4) {
    permutate(r, 0.109951743655322, {0.816847572980459, 0.091576213509771});
    permutate(r, 0.223381589678011, {0.108103018168070, 0.445948490915965});
  } else if (deg == 5) {
    permutate(r, 0.225, {1.0 / 3.0});
    permutate(r, 0.132394152788506, {0.059715871789770, 0.470142064105115});
    permutate(r, 0.125939180544827, {0.797426985353087, 0.101286507323456});
  } else {
    throw Error("triangular_rule: unsupported degree");
  }
  return cache[deg] = r;
}

// src/zisa/math/tetrahedral_rule.cpp:10-122 (Shunn-Ham, degrees 1-3).
const RefRule &tetrahedral_rule(int deg) {
  static std::map<int, RefRule> cache;
  if (deg == -1) deg = 3;
  auto it = cache.find(deg);
  if (it != cache.end()) return it->second;
  RefRule r;
  if (deg <= 1) {
    r.weights = {1.0};
    r.points = {{0.25, 0.25, 0.25, 0.25}};
  } else if (deg == 2) {
    double a = 0.5854101966249680, b = 0.1381966011250110;
    r.weights = {0.25, 0.25, 0.25, 0.25};
    r.points = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}, {b, b, b, a}};
  } else if (deg == 3) {
    double a = 0.7784952948213300, b = 0.0738349017262234;
    double c = 0.4062443438840510, d = 0.0937556561159491;
    double w1 = 0.0476331348432089, w2 = 0.1349112434378610;
    r.weights = {w1, w1, w1, w1, w2, w2, w2, w2, w2, w2};
    r.points = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}, {b, b, b, a}, {c, c, d, d},
                {c, d, c, d}, {c, d, d, c}, {d, c, c, d}, {d, c, d, c}, {d, d, c, c}};
  } else {
    throw Error("tetrahedral_rule: unsupported degree");
  }
  return cache[deg] = r;
}

// ---------------------------------------------------------------------------
// Simplex metrics.

// src/zisa/math/triangle.cpp:11-18 (Heron's formula on the side lengths).
static double triangle_volume(const XYZ &A, const XYZ &B, const XYZ &C) {
  double a = norm(B - C), b = norm(A - C), c = norm(A - B);
  double s = 0.5 * (a + b + c);
  return std::sqrt(s * (s - a) * (s - b) * (s - c));
}

// src/zisa/math/tetrahedron.cpp:18-31
static double tet_volume(const XYZ &v0, const XYZ &v1, const XYZ &v2, const XYZ &v3) {
  XYZ d1 = v1 - v0, d2 = v2 - v0, d3 = v3 - v0;
  double d11 = d1[0], d12 = d1[1], d13 = d1[2];
  double d21 = d2[0], d22 = d2[1], d23 = d2[2];
  double d31 = d3[0], d32 = d3[1], d33 = d3[2];
  double vol = 1.0 / 6.0 *
               (d11 * d22 * d33 + d21 * d32 * d13 + d31 * d12 * d23 - d11 * d32 * d23 -
                d21 * d12 * d33 - d31 * d22 * d13);
  return std::abs(vol);
}

// src/zisa/grid/gmsh_reader.cpp:23-56
static int rel_vertex(int n_dims, int k, int rel) {
  if (n_dims == 2) return (k + rel) % 3;
  static const int tab[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  return tab[k][rel];
}
// src/zisa/grid/gmsh_reader.cpp:58-77
static int rel_off_vertex(int n_dims, int k) {
  if (n_dims == 2) return (k + 2) % 3;
  static const int tab[4] = {2, 3, 1, 0};
  return tab[k];
}

// barycentric -> physical (src/zisa/math/barycentric.cpp:22-24, 59-62)
static XYZ bary_coord(const std::vector<XYZ> &v, const std::array<double, 4> &l) {
  XYZ x = v[0] * l[0];
  for (size_t k = 1; k < v.size(); ++k) x = x + v[k] * l[k];
  return x;
}

// include/zisa/math/denormalized_rule.hpp:35-51
static DenormalizedRule denormalize_simplex(const RefRule &rule, const std::vector<XYZ> &v,
                                            double vol) {
  DenormalizedRule qr;
  for (size_t i = 0; i < rule.weights.size(); ++i) {
    qr.weights.push_back(vol * rule.weights[i]);
    qr.points.push_back(bary_coord(v, rule.points[i]));
  }
  qr.volume = vol;
  return qr;
}

// edge: include/zisa/math/edge.hpp:26-30 and edge_rule weights.
static DenormalizedRule denormalize_edge(const RefRule &rule, const XYZ &a, const XYZ &b) {
  DenormalizedRule qr;
  double vol = norm(b - a);
  for (size_t i = 0; i < rule.weights.size(); ++i) {
    double xi = rule.points[i][0];
    qr.weights.push_back(vol * rule.weights[i]);
    qr.points.push_back((0.5 - 0.5 * xi) * a + (0.5 + 0.5 * xi) * b);
  }
  qr.volume = vol;
  return qr;
}

// include/zisa/math/quadrature.hpp:23-56: seeded with the first point.
static XYZ average_x(const DenormalizedRule &qr) {
  XYZ ret = qr.weights[0] * qr.points[0];
  for (size_t i = 1; i < qr.weights.size(); ++i) ret = ret + qr.weights[i] * qr.points[i];
  ret = 1.0 * ret;
  return ret / qr.volume;
}

static DenormalizedRule make_cell(int n_dims, const std::vector<XYZ> &v, int deg) {
  if (n_dims == 2) return denormalize_simplex(triangular_rule(deg), v, triangle_volume(v[0], v[1], v[2]));
  return denormalize_simplex(tetrahedral_rule(deg), v, tet_volume(v[0], v[1], v[2], v[3]));
}

// make_cell (src/zisa/math/cell_factory.cpp) of one free-standing simplex
DenormalizedRule simplex_cell_rule(int n_dims, const XYZ *v, int deg) {
  return make_cell(n_dims, std::vector<XYZ>(v, v + n_dims + 1), deg);
}

static double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

// src/zisa/math/cell.cpp:15-27
static double avg_moment(const DenormalizedRule &cell, int kx, int ky, int kz) {
  XYZ c = average_x(cell);
  auto f = [&](const XYZ &x) {
    XYZ d = x - c;
    return ipow(d[0], kx) * ipow(d[1], ky) * ipow(d[2], kz);
  };
  double ret = cell.weights[0] * f(cell.points[0]);
  for (size_t i = 1; i < cell.weights.size(); ++i) ret = ret + cell.weights[i] * f(cell.points[i]);
  ret = 1.0 * ret;
  return ret / cell.volume;
}

int poly_dof(int deg, int n_dims) {
  if (deg < 0) return 0;
  if (n_dims == 2) return ((deg + 1) * (deg + 2)) / 2;
  return ((deg + 1) * (deg + 2) * (deg + 3)) / 6;
}
// include/zisa/math/poly2d_impl.hpp:34-41
int poly_index2(int a, int b) {
  int n = a + b;
  return ((n + 1) * n) / 2 + b;
}
int poly_index3(int a, int b, int c) { return poly_dof(a + b + c - 1, 3) + poly_index2(b, c); }

// src/zisa/grid/grid.cpp:1030-1098 (normalized_moments for triangles/tets)
static std::vector<double> normalized_moments(int n_dims, const std::vector<XYZ> &v,
                                              double length, int quad_deg) {
  int degree = quad_deg;
  if (quad_deg == -1) {
    quad_deg = (n_dims == 2 ? 5 : 3);
    degree = quad_deg;
  }
  auto cell = make_cell(n_dims, v, quad_deg);
  std::vector<double> m(poly_dof(degree, n_dims), 0.0);
  double length_d = 1.0;
  for (int d = 0; d <= degree; ++d) {
    for (int k = 0; k <= d; ++k) {
      if (n_dims == 2) {
        int l = d - k;
        m[poly_index2(k, l)] = avg_moment(cell, k, l, 0) / length_d;
      } else {
        for (int l = 0; l <= d - k; ++l) {
          int mm = d - k - l;
          m[poly_index3(k, l, mm)] = avg_moment(cell, k, l, mm) / length_d;
        }
      }
    }
    length_d *= length;
  }
  return m;
}

XYZ Grid::face_to_cell_shift(int i, int k) const {
  int j = neighbour(i, k);
  if (j < 0 || i < j) return {0.0, 0.0, 0.0};
  return shifts[size_t(i) * max_neighbours + k];
}

// src/zisa/grid/grid.cpp:651-723 (Grid constructor)
std::shared_ptr<Grid> make_grid(int n_dims, int n_cells, const int *vertex_ids,
                                const double *corners, const uint8_t *ghost_mask,
                                int edge_deg, int volume_deg, int moments_deg) {
  auto g = std::make_shared<Grid>();
  const int nv = n_dims + 1;
  g->n_dims = n_dims;
  g->n_cells = n_cells;
  g->max_neighbours = nv;
  g->vertex_ids.assign(vertex_ids, vertex_ids + size_t(n_cells) * nv);
  g->corners.resize(size_t(n_cells) * nv);
  for (size_t q = 0; q < g->corners.size(); ++q)
    g->corners[q] = {corners[3 * q], corners[3 * q + 1], corners[3 * q + 2]};

  auto swap_corner = [&](int i, int a, int b) {
    std::swap(g->vertex_ids[size_t(i) * nv + a], g->vertex_ids[size_t(i) * nv + b]);
    std::swap(g->corners[size_t(i) * nv + a], g->corners[size_t(i) * nv + b]);
  };

  // enforce_standard_vertex_order (src/zisa/grid/grid.cpp:152-203)
  for (int i = 0; i < n_cells; ++i) {
    if (n_dims == 2) {
      XYZ v0 = g->corner(i, 0), v1 = g->corner(i, 1), v2 = g->corner(i, 2);
      if (cross(v1 - v0, v2 - v0)[2] < 0.0) swap_corner(i, 1, 2);
    } else {
      XYZ c = 0.25 * (((g->corner(i, 0) + g->corner(i, 1)) + g->corner(i, 2)) + g->corner(i, 3));
      for (int k = 0; k < 4; ++k) {
        XYZ a = g->corner(i, rel_vertex(3, k, 0)), b = g->corner(i, rel_vertex(3, k, 1)),
            cc = g->corner(i, rel_vertex(3, k, 2));
        XYZ n = cross(b - a, cc - a);
        XYZ fc = 1.0 / 3.0 * ((a + b) + cc);
        if (dot(n, fc - c) < 0.0) {
          swap_corner(i, 2, 3);
          break;
        }
      }
    }
  }

  // compute_neighbours (src/zisa/grid/grid.cpp:486-540): two cells are
  // neighbours across local face k when they share that face's vertices.
  const int nb = nv;
  const int nfv = n_dims;  // vertices per face
  g->neighbours.assign(size_t(n_cells) * nb, -1);
  g->shifts.assign(size_t(n_cells) * nb, XYZ{0.0, 0.0, 0.0});
  std::map<std::array<int, 3>, std::pair<int, int>> open;
  for (int i = 0; i < n_cells; ++i) {
    for (int k = 0; k < nb; ++k) {
      std::array<int, 3> key{-1, -1, -1};
      for (int r = 0; r < nfv; ++r) key[r] = g->vertex_ids[size_t(i) * nv + rel_vertex(n_dims, k, r)];
      std::sort(key.begin(), key.begin() + nfv);
      auto it = open.find(key);
      if (it == open.end()) {
        open[key] = {i, k};
      } else {
        auto [j, kj] = it->second;
        g->neighbours[size_t(i) * nb + k] = j;
        g->neighbours[size_t(j) * nb + kj] = i;
        open.erase(it);
      }
    }
  }
  // periodic extension: translation mapping neighbour frame -> own frame,
  // measured on a shared vertex.
  for (int i = 0; i < n_cells; ++i) {
    for (int k = 0; k < nb; ++k) {
      int j = g->neighbour(i, k);
      if (j < 0) continue;
      int lv = rel_vertex(n_dims, k, 0);
      int vid = g->vertex_ids[size_t(i) * nv + lv];
      for (int m = 0; m < nv; ++m) {
        if (g->vertex_ids[size_t(j) * nv + m] == vid) {
          g->shifts[size_t(i) * nb + k] = g->corner(i, lv) - g->corner(j, m);
        }
      }
    }
  }

  // count_interior_edges / compute_edge_indices / compute_left_right
  // (src/zisa/grid/grid.cpp:58-96, 385-442)
  int n_int = 0, n_ext = 0;
  for (int i = 0; i < n_cells; ++i)
    for (int k = 0; k < nb; ++k) {
      int j = g->neighbour(i, k);
      if (j < 0) ++n_ext;
      else if (i < j) ++n_int;
    }
  g->n_interior_edges = n_int;
  g->n_exterior_edges = n_ext;
  g->n_edges = n_int + n_ext;
  g->edge_indices.assign(size_t(n_cells) * nb, -1);
  g->left_right.assign(g->n_edges, {-1, -1});
  g->edge_shift.assign(g->n_edges, XYZ{0.0, 0.0, 0.0});
  int ei_int = 0, ei_ext = n_int;
  for (int i = 0; i < n_cells; ++i) {
    for (int k = 0; k < nb; ++k) {
      int j = g->neighbour(i, k);
      if (j < 0) {
        int e = ei_ext++;
        g->edge_indices[size_t(i) * nb + k] = e;
        g->left_right[e] = {i, -1};
      } else if (i < j) {
        int e = ei_int++;
        g->edge_indices[size_t(i) * nb + k] = e;
        g->left_right[e] = {i, j};
        g->edge_shift[e] = g->shifts[size_t(i) * nb + k];
      } else {
        int kj = -1;
        for (int m = 0; m < nb; ++m)
          if (g->neighbour(j, m) == i) kj = m;
        if (kj < 0) throw Error("Failed to find myself.");
        g->edge_indices[size_t(i) * nb + k] = g->edge_indices[size_t(j) * nb + kj];
      }
    }
  }

  g->volumes.resize(n_cells);
  g->inradii.resize(n_cells);
  g->circum_radii.resize(n_cells);
  g->characteristic_length.resize(n_cells);
  g->cells.resize(n_cells);
  g->cell_centers.resize(n_cells);
  g->normalized_moments.resize(n_cells);
  for (int i = 0; i < n_cells; ++i) {
    std::vector<XYZ> v(nv);
    for (int k = 0; k < nv; ++k) v[k] = g->corner(i, k);
    if (n_dims == 2) {
      // src/zisa/math/triangle.cpp:20-42
      double a = norm(v[1] - v[2]), b = norm(v[0] - v[2]), c = norm(v[0] - v[1]);
      g->volumes[i] = triangle_volume(v[0], v[1], v[2]);
      g->inradii[i] = 0.5 * std::sqrt((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c));
      XYZ bc = ((v[0] + v[1]) + v[2]) / 3.0;
      double cr = std::max(std::max(norm(v[0] - bc), norm(v[1] - bc)), norm(v[2] - bc));
      g->circum_radii[i] = cr;
      g->characteristic_length[i] = cr;
    } else {
      // src/zisa/math/tetrahedron.cpp:18-73
      g->volumes[i] = tet_volume(v[0], v[1], v[2], v[3]);
      XYZ c = 0.25 * (((v[0] + v[1]) + v[2]) + v[3]);
      double l = 0.0, rmin = 1e300, rmax = 0.0;
      for (int k = 0; k < 4; ++k) l += 2.0 * norm(c - v[k]);
      for (int k = 0; k < 4; ++k) {
        XYZ fc = ((v[rel_vertex(3, k, 0)] + v[rel_vertex(3, k, 1)]) + v[rel_vertex(3, k, 2)]) / 3.0;
        rmin = std::min(rmin, norm(fc - c));
        rmax = std::max(rmax, norm(v[k] - c));
      }
      g->characteristic_length[i] = 0.25 * l;
      g->inradii[i] = rmin;
      g->circum_radii[i] = rmax;
    }
    // compute_cells + compute_barycenters (grid.cpp:561-581, 318-328)
    g->cells[i] = make_cell(n_dims, v, volume_deg);
    g->cell_centers[i] = average_x(g->cells[i]);
    g->normalized_moments[i] = normalized_moments(n_dims, v, g->characteristic_length[i], moments_deg);
  }

  // compute_faces (grid.cpp:583-611) + make_face (src/zisa/math/face_factory.cpp:11-34)
  g->faces.resize(g->n_edges);
  g->face_centers.resize(g->n_edges);
  for (int i = 0; i < n_cells; ++i) {
    for (int k = 0; k < nb; ++k) {
      int j = g->neighbour(i, k);
      if (!(j < 0 || i < j)) continue;
      int e = g->edge_indices[size_t(i) * nb + k];
      Face f;
      XYZ v0 = g->corner(i, rel_vertex(n_dims, k, 0)), v1 = g->corner(i, rel_vertex(n_dims, k, 1));
      if (n_dims == 2) {
        f.qr = denormalize_edge(edge_rule(edge_deg), v0, v1);
        XYZ t = normalize(v1 - v0);
        f.normal = {t[1], -t[0], t[2]};  // rotate_right
        f.t1 = t;
        f.t2 = cross(f.normal, f.t1);
      } else {
        XYZ v2 = g->corner(i, rel_vertex(n_dims, k, 2));
        f.qr = denormalize_simplex(triangular_rule(edge_deg), {v0, v1, v2}, triangle_volume(v0, v1, v2));
        f.normal = normalize(cross(v1 - v0, v2 - v0));
        f.t1 = normalize(v1 - v0);
        f.t2 = cross(f.normal, f.t1);
      }
      g->faces[e] = f;
      g->face_centers[e] = average_x(f.qr);
    }
  }

  // mask_ghost_cells (grid.cpp:1122-1136)
  g->cell_flags.assign(n_cells, CellFlags());
  if (ghost_mask) {
    for (int i = 0; i < n_cells; ++i) {
      if (!ghost_mask[i]) continue;
      g->cell_flags[i].interior = false;
      g->cell_flags[i].ghost_cell = true;
      for (int k = 0; k < nb; ++k) {
        int j = g->neighbour(i, k);
        if (j >= 0 && !ghost_mask[j]) g->cell_flags[i].ghost_cell_l1 = true;
      }
    }
  }
  return g;
}

}  // namespace zfv_ref
