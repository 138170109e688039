// Host setup of one instance.  All cell-local steps run in parallel
// (OpenMP) and are deterministic: the result does not depend on the thread
// count.
//
// The geometry and stencil rules restate the reference's so that the
// reconstruction is the same operator:
//   geometry     src/zisa/grid/grid.cpp:152-203 (vertex order), 486-540
//                (neighbours), 385-442 (edges), src/zisa/math/{triangle,
//                tetrahedron}.cpp (metrics), grid.cpp:1030-1098 (moments)
//   stencils     src/zisa/reconstruction/stencil.cpp:159-399,
//                stencil_family.cpp:15-136
//   LSQ matrices src/zisa/reconstruction/lsq_solver.cpp:168-403
// The per-call LDLT solve of the reference (lsq_solver.cpp:52-85) becomes a
// pseudo-inverse P = (A^T A)^{-1} A^T formed once here with the same pivoted
// LDLT, so a kernel applies one small mat-vec per stencil.
#include "setup.hpp"
#include "stage_math.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

namespace zfv {

namespace {

using V3 = std::array<double, 3>;
inline V3 add(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 sub(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 scl(double s, const V3 &a) { return {s * a[0], s * a[1], s * a[2]}; }
inline double dot3(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double nrm(const V3 &a) { return std::sqrt(dot3(a, a)); }
inline V3 crs(const V3 &a, const V3 &b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

int dof(int deg, int dim) {
  if (deg < 0) return 0;
  return dim == 2 ? (deg + 1) * (deg + 2) / 2 : (deg + 1) * (deg + 2) * (deg + 3) / 6;
}
int pidx2(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }
int pidx3(int a, int b, int c) { return dof(a + b + c - 1, 3) + pidx2(b, c); }

// GMSH local face -> vertex table (src/zisa/grid/gmsh_reader.cpp:23-77)
int face_vertex(int dim, int k, int r) {
  if (dim == 2) return (k + r) % 3;
  static const int t[4][3] = {{0, 1, 3}, {0, 2, 1}, {0, 3, 2}, {1, 2, 3}};
  return t[k][r];
}
int off_vertex(int dim, int k) {
  if (dim == 2) return (k + 2) % 3;
  static const int t[4] = {2, 3, 1, 0};
  return t[k];
}

// ---- normalized quadrature rules ------------------------------------------
struct Rule {
  std::vector<double> w;
  std::vector<std::array<double, 4>> x;
};

Rule gl_edge(int deg) {
  if (deg == -1) deg = 7;
  int n = deg / 2 + 1;
  Rule r;
  std::vector<double> x, w;
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
    throw ZfvError(ZFV_ERR_INVALID, "edge rule: unsupported degree");
  }
  for (int k = 0; k < n; ++k) {
    r.w.push_back(0.5 * w[k]);
    r.x.push_back({x[k], 0, 0, 0});
  }
  return r;
}

void perm3(Rule &r, double w, std::initializer_list<double> l) {
  std::vector<double> v(l);
  if (v.size() == 1) {
    r.w.push_back(w);
    r.x.push_back({v[0], v[0], v[0], 0});
  } else {
    for (int k = 0; k < 3; ++k) r.w.push_back(w);
    r.x.push_back({v[0], v[1], v[1], 0});
    r.x.push_back({v[1], v[0], v[1], 0});
    r.x.push_back({v[1], v[1], v[0], 0});
  }
}

Rule dunavant(int deg) {
  if (deg == -1) deg = 5;
  Rule r;
  switch (deg) {
    case 1: perm3(r, 1.0, {1.0 / 3.0}); break;
    case 2: perm3(r, 1.0 / 3.0, {2.0 / 3.0, 1.0 / 6.0}); break;
    case 3:
      perm3(r, -0.5625, {1.0 / 3.0});
      perm3(r, 1.5625 / 3.0, {0.6, 0.2});
      break;
    case 4:
      perm3(r, 0.109951743655322, {0.816847572980459, 0.091576213509771});
      perm3(r, 0.223381589678011, {0.108103018168070, 0.445948490915965});
      break;
    case 5:
      perm3(r, 0.225, {1.0 / 3.0});
      perm3(r, 0.132394152788506, {0.059715871789770, 0.470142064105115});
      perm3(r, 0.125939180544827, {0.797426985353087, 0.101286507323456});
      break;
    default: throw ZfvError(ZFV_ERR_INVALID, "triangle rule: unsupported degree");
  }
  return r;
}

Rule shunn_ham(int deg) {
  if (deg == -1) deg = 3;
  Rule r;
  if (deg <= 1) {
    r.w = {1.0};
    r.x = {{0.25, 0.25, 0.25, 0.25}};
  } else if (deg == 2) {
    double a = 0.5854101966249680, b = 0.1381966011250110;
    r.w = {0.25, 0.25, 0.25, 0.25};
    r.x = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}, {b, b, b, a}};
  } else if (deg == 3) {
    double a = 0.7784952948213300, b = 0.0738349017262234, c = 0.4062443438840510, d = 0.0937556561159491;
    double w1 = 0.0476331348432089, w2 = 0.1349112434378610;
    r.w = {w1, w1, w1, w1, w2, w2, w2, w2, w2, w2};
    r.x = {{a, b, b, b}, {b, a, b, b}, {b, b, a, b}, {b, b, b, a}, {c, c, d, d},
           {c, d, c, d}, {c, d, d, c}, {d, c, c, d}, {d, c, d, c}, {d, d, c, c}};
  } else {
    throw ZfvError(ZFV_ERR_INVALID, "tetrahedron rule: unsupported degree");
  }
  return r;
}

Rule volume_rule(int dim, int deg) { return dim == 2 ? dunavant(deg) : shunn_ham(deg); }

// ---- geometry --------------------------------------------------------------
struct Geo {
  int dim = 2, nv = 3, n = 0;
  std::vector<int32_t> vid;
  std::vector<V3> crn;  // n * nv
  std::vector<int32_t> nbr;
  std::vector<V3> shf;  // n * nv, neighbour frame -> own frame
  std::vector<double> vol, inr, len;
  std::vector<V3> ctr;
  std::vector<std::vector<double>> mom;
  std::vector<uint8_t> ghost, ghost_l1;
  const V3 &c(int i, int k) const { return crn[size_t(i) * nv + k]; }
  int nb(int i, int k) const { return nbr[size_t(i) * nv + k]; }
  const V3 &sh(int i, int k) const { return shf[size_t(i) * nv + k]; }
};

double heron(const V3 &A, const V3 &B, const V3 &C) {
  double a = nrm(sub(B, C)), b = nrm(sub(A, C)), c = nrm(sub(A, B));
  double s = 0.5 * (a + b + c);
  return std::sqrt(s * (s - a) * (s - b) * (s - c));
}

double tet_vol(const V3 &v0, const V3 &v1, const V3 &v2, const V3 &v3) {
  V3 d1 = sub(v1, v0), d2 = sub(v2, v0), d3 = sub(v3, v0);
  double v = 1.0 / 6.0 *
             (d1[0] * d2[1] * d3[2] + d2[0] * d3[1] * d1[2] + d3[0] * d1[1] * d2[2] - d1[0] * d3[1] * d2[2] -
              d2[0] * d1[1] * d3[2] - d3[0] * d2[1] * d1[2]);
  return std::abs(v);
}

V3 bary_point(const V3 *v, int nv, const std::array<double, 4> &l) {
  V3 x = scl(l[0], v[0]);
  for (int k = 1; k < nv; ++k) x = add(x, scl(l[k], v[k]));
  return x;
}

double cell_volume(int dim, const V3 *v) { return dim == 2 ? heron(v[0], v[1], v[2]) : tet_vol(v[0], v[1], v[2], v[3]); }

// quadrature-average of x (the reference's cell barycenter, include/zisa/math/quadrature.hpp:23-56)
V3 quad_center(int dim, const V3 *v, const Rule &r) {
  double vol = cell_volume(dim, v);
  V3 ret = scl(vol * r.w[0], bary_point(v, dim + 1, r.x[0]));
  for (size_t q = 1; q < r.w.size(); ++q) ret = add(ret, scl(vol * r.w[q], bary_point(v, dim + 1, r.x[q])));
  return scl(1.0 / vol, scl(1.0, ret));
}

double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

std::vector<double> moments(int dim, const V3 *v, double length, int mdeg) {
  int degree = mdeg;
  if (mdeg == -1) degree = mdeg = (dim == 2 ? 5 : 3);
  Rule r = volume_rule(dim, mdeg);
  double vol = cell_volume(dim, v);
  V3 c = quad_center(dim, v, r);
  std::vector<double> m(dof(degree, dim), 0.0);
  auto avg = [&](int kx, int ky, int kz) {
    double ret = 0.0;
    for (size_t q = 0; q < r.w.size(); ++q) {
      V3 d = sub(bary_point(v, dim + 1, r.x[q]), c);
      double f = ipow(d[0], kx) * ipow(d[1], ky) * ipow(d[2], kz);
      ret = (q == 0) ? (vol * r.w[q]) * f : ret + (vol * r.w[q]) * f;
    }
    return (1.0 * ret) / vol;
  };
  double ld = 1.0;
  for (int d = 0; d <= degree; ++d) {
    for (int k = 0; k <= d; ++k) {
      if (dim == 2) {
        m[pidx2(k, d - k)] = avg(k, d - k, 0) / ld;
      } else {
        for (int l = 0; l <= d - k; ++l) m[pidx3(k, l, d - k - l)] = avg(k, l, d - k - l) / ld;
      }
    }
    ld *= length;
  }
  return m;
}

void build_geometry(const zfv_mesh &mesh, const zfv_params &p, Geo &g) {
  g.dim = mesh.n_dims;
  g.nv = g.dim + 1;
  g.n = mesh.n_cells;
  const int nv = g.nv, n = g.n;
  g.vid.assign(mesh.vertex_ids, mesh.vertex_ids + size_t(n) * nv);
  g.crn.resize(size_t(n) * nv);
  for (size_t q = 0; q < g.crn.size(); ++q)
    g.crn[q] = {mesh.corners[3 * q], mesh.corners[3 * q + 1], mesh.corners[3 * q + 2]};

  // standard (positive) vertex order
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3 *v = &g.crn[size_t(i) * nv];
    int32_t *id = &g.vid[size_t(i) * nv];
    if (g.dim == 2) {
      if (crs(sub(v[1], v[0]), sub(v[2], v[0]))[2] < 0.0) {
        std::swap(v[1], v[2]);
        std::swap(id[1], id[2]);
      }
    } else {
      V3 c = scl(0.25, add(add(add(v[0], v[1]), v[2]), v[3]));
      for (int k = 0; k < 4; ++k) {
        V3 a = v[face_vertex(3, k, 0)], b = v[face_vertex(3, k, 1)], cc = v[face_vertex(3, k, 2)];
        V3 nn = crs(sub(b, a), sub(cc, a));
        V3 fc = scl(1.0 / 3.0, add(add(a, b), cc));
        if (dot3(nn, sub(fc, c)) < 0.0) {
          std::swap(v[2], v[3]);
          std::swap(id[2], id[3]);
          break;
        }
      }
    }
  }

  // neighbours: sort (face key, cell, local face) and pair equal keys
  struct FK {
    int32_t a, b, c;
    int32_t cell, k;
  };
  std::vector<FK> fk(size_t(n) * nv);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < nv; ++k) {
      int32_t t[3] = {-1, -1, -1};
      for (int r = 0; r < g.dim; ++r) t[r] = g.vid[size_t(i) * nv + face_vertex(g.dim, k, r)];
      std::sort(t, t + g.dim);
      fk[size_t(i) * nv + k] = {t[0], t[1], t[2], i, k};
    }
  std::sort(fk.begin(), fk.end(), [](const FK &x, const FK &y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    if (x.c != y.c) return x.c < y.c;
    return x.cell < y.cell;
  });
  g.nbr.assign(size_t(n) * nv, -1);
  for (size_t q = 0; q + 1 < fk.size(); ++q) {
    const FK &x = fk[q], &y = fk[q + 1];
    if (x.a == y.a && x.b == y.b && x.c == y.c) {
      if (q + 2 < fk.size() && fk[q + 2].a == x.a && fk[q + 2].b == x.b && fk[q + 2].c == x.c)
        throw ZfvError(ZFV_ERR_INVALID, "mesh: a face is shared by more than two cells");
      g.nbr[size_t(x.cell) * nv + x.k] = y.cell;
      g.nbr[size_t(y.cell) * nv + y.k] = x.cell;
      ++q;
    }
  }
  g.shf.assign(size_t(n) * nv, V3{0, 0, 0});
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < nv; ++k) {
      int j = g.nb(i, k);
      if (j < 0) continue;
      int lv = face_vertex(g.dim, k, 0);
      int id = g.vid[size_t(i) * nv + lv];
      for (int m = 0; m < nv; ++m)
        if (g.vid[size_t(j) * nv + m] == id) g.shf[size_t(i) * nv + k] = sub(g.c(i, lv), g.c(j, m));
    }

  Rule vr = volume_rule(g.dim, p.volume_deg);
  g.vol.resize(n);
  g.inr.resize(n);
  g.len.resize(n);
  g.ctr.resize(n);
  g.mom.resize(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const V3 *v = &g.crn[size_t(i) * nv];
    if (g.dim == 2) {
      double a = nrm(sub(v[1], v[2])), b = nrm(sub(v[0], v[2])), c = nrm(sub(v[0], v[1]));
      g.vol[i] = heron(v[0], v[1], v[2]);
      g.inr[i] = 0.5 * std::sqrt((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c));
      V3 bc = scl(1.0 / 3.0, add(add(v[0], v[1]), v[2]));
      bc = {(v[0][0] + v[1][0] + v[2][0]) / 3.0, (v[0][1] + v[1][1] + v[2][1]) / 3.0,
            (v[0][2] + v[1][2] + v[2][2]) / 3.0};
      g.len[i] = std::max(std::max(nrm(sub(v[0], bc)), nrm(sub(v[1], bc))), nrm(sub(v[2], bc)));
    } else {
      g.vol[i] = tet_vol(v[0], v[1], v[2], v[3]);
      V3 c = scl(0.25, add(add(add(v[0], v[1]), v[2]), v[3]));
      double l = 0.0, rmin = 1e300;
      for (int k = 0; k < 4; ++k) l += 2.0 * nrm(sub(c, v[k]));
      for (int k = 0; k < 4; ++k) {
        V3 fc = add(add(v[face_vertex(3, k, 0)], v[face_vertex(3, k, 1)]), v[face_vertex(3, k, 2)]);
        fc = {fc[0] / 3.0, fc[1] / 3.0, fc[2] / 3.0};
        rmin = std::min(rmin, nrm(sub(fc, c)));
      }
      g.len[i] = 0.25 * l;
      g.inr[i] = rmin;
    }
    g.ctr[i] = quad_center(g.dim, v, vr);
    g.mom[i] = moments(g.dim, v, g.len[i], p.moments_deg);
  }

  g.ghost.assign(n, 0);
  g.ghost_l1.assign(n, 0);
  if (mesh.ghost_mask) {
    for (int i = 0; i < n; ++i) g.ghost[i] = mesh.ghost_mask[i] ? 1 : 0;
    for (int i = 0; i < n; ++i) {
      if (!g.ghost[i]) continue;
      for (int k = 0; k < nv; ++k) {
        int j = g.nb(i, k);
        if (j >= 0 && !g.ghost[j]) g.ghost_l1[i] = 1;
      }
    }
  }
}

// ---- stencils ----------------------------------------------------------------
struct Mem {
  int cell;
  V3 shift;
};

struct Cone {
  int kind = 0;  // 0 everything, 1 wedge (2D), 2 trihedral cone (3D)
  V3 A{}, dB{}, dC{}, dD{};
  bool inside(const V3 &x) const {
    if (kind == 0) return true;
    V3 dx = sub(x, A);
    if (kind == 1) return crs(dB, dx)[2] >= 0.0 && crs(dx, dC)[2] >= 0.0;
    return dot3(dB, crs(dC, dx)) >= 0.0 && dot3(dC, crs(dD, dx)) >= 0.0 && dot3(dD, crs(dB, dx)) >= 0.0;
  }
};

struct StencilBuilder {
  const Geo &g;
  Rule query;  // maximal-degree cell rule used to test region membership
  Rule face;   // face rule (for face barycenters)
  StencilBuilder(const Geo &geo, int edge_deg)
      : g(geo), query(volume_rule(geo.dim, -1)), face(geo.dim == 2 ? gl_edge(edge_deg) : dunavant(edge_deg)) {}

  bool in_region(int cand, const V3 &shift, const Cone &r) const {
    const V3 *v = &g.crn[size_t(cand) * g.nv];
    for (size_t q = 0; q < query.w.size(); ++q) {
      V3 x = scl(query.x[q][0], v[0]);
      for (int k = 1; k < g.nv; ++k) x = add(x, scl(query.x[q][k], v[k]));
      if (r.inside(add(x, shift))) return true;
    }
    return r.inside(add(g.ctr[cand], shift));
  }

  // breadth-first candidates (stencil.cpp:192-236)
  std::vector<Mem> candidates(int ic, int npts, const Cone &r) const {
    std::vector<Mem> c{{ic, {0, 0, 0}}};
    const int maxp = 5 * npts;
    for (int p = 0; p < maxp && p < int(c.size()); ++p) {
      Mem mj = c[p];
      for (int k = 0; k < g.nv; ++k) {
        int cand = g.nb(mj.cell, k);
        if (cand < 0) continue;
        bool seen = false;
        for (const auto &m : c)
          if (m.cell == cand) {
            seen = true;
            break;
          }
        if (seen) continue;
        V3 s = add(mj.shift, g.sh(mj.cell, k));
        if (in_region(cand, s, r)) c.push_back({cand, s});
      }
    }
    return c;
  }

  // nearest npts candidates; ties keep breadth-first order
  std::vector<Mem> nearest(int ic, int npts, const Cone &r) const {
    auto c = candidates(ic, npts, r);
    std::vector<double> d(c.size());
    for (size_t q = 0; q < c.size(); ++q) d[q] = nrm(sub(add(g.ctr[c[q].cell], c[q].shift), g.ctr[ic]));
    std::vector<int> o(c.size());
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return d[a] < d[b]; });
    std::vector<Mem> out;
    for (size_t q = 0; q < o.size() && int(out.size()) < npts; ++q) out.push_back(c[o[q]]);
    return out;
  }

  Cone cone(int ic, const V3 &apex, int k) const {
    Cone r;
    r.kind = g.dim == 2 ? 1 : 2;
    r.A = apex;
    r.dB = sub(g.c(ic, face_vertex(g.dim, k, 0)), apex);
    r.dC = sub(g.c(ic, face_vertex(g.dim, k, 1)), apex);
    if (g.dim == 3) r.dD = sub(g.c(ic, face_vertex(g.dim, k, 2)), apex);
    return r;
  }
};

// LSQ matrix rows for members st[1..rows] relative to st[0] (lsq_solver.cpp:168-403)
void lsq_matrix(const Geo &g, const std::vector<Mem> &st, int rows, int order, std::vector<double> &A, int &cols) {
  const int dim = g.dim;
  cols = dof(order - 1, dim) - 1;
  A.assign(size_t(rows) * cols, 0.0);
  if (order <= 1) return;
  auto ix = [dim](int i, int j, int k) { return dim == 2 ? pidx2(i, j) : pidx3(i, j, k); };
  int i0 = st[0].cell;
  const V3 x0 = g.ctr[i0];
  const double l0 = g.len[i0];
  const auto &C0 = g.mom[i0];
  for (int r = 0; r < rows; ++r) {
    const Mem &mj = st[r + 1];
    const auto &Cj = g.mom[mj.cell];
    V3 X = sub(add(g.ctr[mj.cell], mj.shift), x0);
    X = {X[0] / l0, X[1] / l0, X[2] / l0};
    const double x = X[0], y = X[1], z = X[2];
    const double lj = g.len[mj.cell] / l0;
    double *a = &A[size_t(r) * cols];
    auto set = [&](int idx, double v) { a[idx - 1] = v; };
    // two-dimensional monomials (also used in 3D with a zero z-exponent)
    set(ix(1, 0, 0), x);
    set(ix(0, 1, 0), y);
    if (order >= 3) {
      double l2 = lj * lj;
      int i20 = ix(2, 0, 0), i11 = ix(1, 1, 0), i02 = ix(0, 2, 0);
      double x20 = x * x, x11 = x * y, x02 = y * y;
      set(i20, x20 - C0[i20] + l2 * Cj[i20]);
      set(i11, x11 - C0[i11] + l2 * Cj[i11]);
      set(i02, x02 - C0[i02] + l2 * Cj[i02]);
      if (order >= 4) {
        double l3 = l2 * lj;
        int i30 = ix(3, 0, 0), i21 = ix(2, 1, 0), i12 = ix(1, 2, 0), i03 = ix(0, 3, 0);
        double x30 = x20 * x, x21 = x20 * y, x12 = x11 * y, x03 = x02 * y;
        set(i30, x30 - C0[i30] + 3.0 * x * l2 * Cj[i20] + l3 * Cj[i30]);
        set(i21, x21 - C0[i21] + y * l2 * Cj[i20] + 2.0 * x * l2 * Cj[i11] + l3 * Cj[i21]);
        set(i12, x12 - C0[i12] + x * l2 * Cj[i02] + 2.0 * y * l2 * Cj[i11] + l3 * Cj[i12]);
        set(i03, x03 - C0[i03] + 3.0 * y * l2 * Cj[i02] + l3 * Cj[i03]);
        if (order >= 5) {
          if (dim == 3) throw ZfvError(ZFV_ERR_INVALID, "3D reconstruction supports order <= 4");
          double l4 = l3 * lj;
          int i40 = ix(4, 0, 0), i31 = ix(3, 1, 0), i22 = ix(2, 2, 0), i13 = ix(1, 3, 0), i04 = ix(0, 4, 0);
          double x40 = x30 * x, x31 = x30 * y, x22 = x21 * y, x13 = x12 * y, x04 = x03 * y;
          set(i40, x40 - C0[i40] + 6.0 * x20 * l2 * Cj[i20] + 4.0 * x * l3 * Cj[i30] + l4 * Cj[i40]);
          set(i31, x31 - C0[i31] + 3 * x11 * l2 * Cj[i20] + 3.0 * x20 * l2 * Cj[i11] + y * l3 * Cj[i30] +
                       3.0 * x * l3 * Cj[i21] + l4 * Cj[i31]);
          set(i22, x22 - C0[i22] + x02 * l2 * Cj[i20] + x20 * l2 * Cj[i02] + 4 * x11 * l2 * Cj[i11] +
                       2.0 * y * l3 * Cj[i21] + 2 * x * l3 * Cj[i12] + l4 * Cj[i22]);
          set(i13, x13 - C0[i13] + 3 * x11 * l2 * Cj[i02] + 3.0 * x02 * l2 * Cj[i11] + x * l3 * Cj[i03] +
                       3.0 * y * l3 * Cj[i12] + l4 * Cj[i13]);
          set(i04, x04 - C0[i04] + 6.0 * x02 * l2 * Cj[i02] + 4.0 * y * l3 * Cj[i03] + l4 * Cj[i04]);
          if (order >= 6) throw ZfvError(ZFV_ERR_INVALID, "2D reconstruction supports order <= 5");
        }
      }
    }
    if (dim == 3) {
      set(pidx3(0, 0, 1), z);
      if (order >= 3) {
        double l2 = lj * lj;
        int i002 = pidx3(0, 0, 2), i101 = pidx3(1, 0, 1), i011 = pidx3(0, 1, 1);
        set(i002, z * z - C0[i002] + l2 * Cj[i002]);
        set(i101, x * z - C0[i101] + l2 * Cj[i101]);
        set(i011, y * z - C0[i011] + l2 * Cj[i011]);
        if (order >= 4) {
          double l3 = lj * lj * lj;
          int i003 = pidx3(0, 0, 3), i102 = pidx3(1, 0, 2), i012 = pidx3(0, 1, 2);
          int i201 = pidx3(2, 0, 1), i111 = pidx3(1, 1, 1), i021 = pidx3(0, 2, 1);
          int i200 = pidx3(2, 0, 0), i020 = pidx3(0, 2, 0), i110 = pidx3(1, 1, 0);
          set(i003, z * z * z - C0[i003] + 3.0 * z * l2 * Cj[i002] + l3 * Cj[i003]);
          set(i102, x * z * z - C0[i102] + x * l2 * Cj[i002] + 2.0 * z * l2 * Cj[i101] + l3 * Cj[i102]);
          set(i012, y * z * z - C0[i012] + y * l2 * Cj[i002] + 2.0 * z * l2 * Cj[i011] + l3 * Cj[i012]);
          set(i201, x * x * z - C0[i201] + z * l2 * Cj[i200] + 2.0 * x * l2 * Cj[i101] + l3 * Cj[i201]);
          set(i021, y * y * z - C0[i021] + z * l2 * Cj[i020] + 2.0 * y * l2 * Cj[i011] + l3 * Cj[i021]);
          set(i111, x * y * z - C0[i111] + z * l2 * Cj[i110] + y * l2 * Cj[i101] + x * l2 * Cj[i011] +
                        l3 * Cj[i111]);
        }
      }
    }
  }
}

// numerical rank with Eigen JacobiSVD's default threshold (one-sided Jacobi)
int rank_of(std::vector<double> A, int rows, int cols) {
  auto at = [&](int r, int c) -> double & { return A[size_t(r) * cols + c]; };
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < cols; ++p)
      for (int q = p + 1; q < cols; ++q) {
        double a = 0, b = 0, c = 0;
        for (int r = 0; r < rows; ++r) {
          a += at(r, p) * at(r, p);
          b += at(r, q) * at(r, q);
          c += at(r, p) * at(r, q);
        }
        if (std::abs(c) <= 1e-300 || std::abs(c) <= 1e-17 * std::sqrt(a * b)) continue;
        off = std::max(off, std::abs(c) / std::sqrt(a * b));
        double zeta = (b - a) / (2.0 * c);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int r = 0; r < rows; ++r) {
          double xp = at(r, p), xq = at(r, q);
          at(r, p) = cs * xp - sn * xq;
          at(r, q) = sn * xp + cs * xq;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> sv(cols);
  for (int c = 0; c < cols; ++c) {
    double s = 0;
    for (int r = 0; r < rows; ++r) s += at(r, c) * at(r, c);
    sv[c] = std::sqrt(s);
  }
  std::sort(sv.begin(), sv.end(), std::greater<double>());
  int dg = std::min(rows, cols);
  if (dg == 0) return 0;
  double thr = std::max(sv[0] * dg * 2.220446049250313e-16, 2.2250738585072014e-308);
  int rk = 0;
  for (int k = 0; k < dg; ++k)
    if (sv[k] >= thr) ++rk;
  return rk;
}

int required_size(int deg, double factor, int dim) {
  if (deg == 0) return 1;
  return int(double(dof(deg, dim) - 1) * factor + 1);
}
int deduce_order(int size, double factor, int dim) {
  int deg = 0;
  while (required_size(deg + 1, factor, dim) <= size) ++deg;
  return deg + 1;
}

uint64_t mix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct CellStencils {
  int n = 0;
  int order[ZFV_MAX_STENCILS] = {0};
  std::vector<Mem> st[ZFV_MAX_STENCILS];  // members (first `size` used)
  int size[ZFV_MAX_STENCILS] = {0};
  int khigh = 0;
  bool first_order = false;
  std::vector<Mem> l2g;
};

bool good(const Geo &g, const std::vector<Mem> &s, int npts, int order) {
  std::vector<double> A;
  int cols;
  lsq_matrix(g, s, npts - 1, order, A, cols);
  return rank_of(A, npts - 1, cols) == cols;
}

std::vector<Mem> biased(const StencilBuilder &b, int ic, int k, int npts, int order) {
  const Geo &g = b.g;
  auto s = b.nearest(ic, npts, b.cone(ic, g.c(ic, off_vertex(g.dim, k)), k));
  if (int(s.size()) == npts && good(g, s, npts, order)) return s;
  s = b.nearest(ic, npts, b.cone(ic, g.ctr[ic], k));
  if (int(s.size()) == npts && good(g, s, npts, order)) return s;
  // apex at the face barycenter (src/zisa/math/face.cpp:9-11): quadrature
  // average of x over the face rule, with the face's vertices in its left
  // cell's order, mapped into this cell's frame.
  const int j = g.nb(ic, k);
  V3 v[3];
  if (j >= 0 && j < ic) {
    int kj = -1;
    for (int m = 0; m < g.nv; ++m)
      if (g.nb(j, m) == ic) kj = m;
    for (int r = 0; r < g.dim; ++r) v[r] = add(g.c(j, face_vertex(g.dim, kj, r)), g.sh(ic, k));
  } else {
    for (int r = 0; r < g.dim; ++r) v[r] = g.c(ic, face_vertex(g.dim, k, r));
  }
  V3 apex{0, 0, 0};
  if (g.dim == 2) {
    double area = nrm(sub(v[1], v[0]));
    for (size_t q = 0; q < b.face.w.size(); ++q) {
      double xi = b.face.x[q][0];
      V3 x = add(scl(0.5 - 0.5 * xi, v[0]), scl(0.5 + 0.5 * xi, v[1]));
      apex = q == 0 ? scl(area * b.face.w[q], x) : add(apex, scl(area * b.face.w[q], x));
    }
    apex = scl(1.0 / area, scl(1.0, apex));
  } else {
    double area = heron(v[0], v[1], v[2]);
    for (size_t q = 0; q < b.face.w.size(); ++q) {
      V3 x = bary_point(v, 3, b.face.x[q]);
      apex = q == 0 ? scl(area * b.face.w[q], x) : add(apex, scl(area * b.face.w[q], x));
    }
    apex = scl(1.0 / area, scl(1.0, apex));
  }
  Cone r = b.cone(ic, apex, k);
  auto c = b.candidates(ic, npts, r);
  if (int(c.size()) < npts) return {c[0]};
  if (!good(g, c, int(c.size()), order))
    throw ZfvError(ZFV_ERR_INVALID, "stencil: it is impossible to find a suitable one-sided stencil");
  uint64_t seed = uint64_t(ic) * 8u + uint64_t(k);
  for (int it = 0; it < 100; ++it) {
    int nn = int(c.size()) - 1;
    for (int m = nn - 1; m >= 1; --m) {
      int rr = int(mix64(seed) % uint64_t(m + 1));
      std::swap(c[1 + m], c[1 + rr]);
    }
    if (good(g, c, npts, order)) {
      c.resize(npts);
      return c;
    }
  }
  throw ZfvError(ZFV_ERR_INVALID, "stencil: no suitable one-sided stencil among the candidates");
}

void cell_stencils(const StencilBuilder &b, const zfv_params &p, int ic, CellStencils &cs) {
  const Geo &g = b.g;
  const bool full = !g.ghost[ic] || g.ghost_l1[ic];
  int ns = full ? p.n_stencils : 1;
  cs.n = ns;
  cs.first_order = !full;
  int kb = 0;
  for (int s = 0; s < ns; ++s) {
    int mo = full ? p.orders[s] : 1;
    double of = full ? p.overfit[s] : 1.0;
    int bias = full ? p.biases[s] : 0;
    int maxsize = required_size(mo - 1, of, g.dim);
    std::vector<Mem> m;
    if (bias == 1) {
      m = biased(b, ic, kb, maxsize, mo);
      ++kb;
    } else {
      m = b.nearest(ic, maxsize, Cone{});
    }
    for (const auto &x : m) {
      bool seen = false;
      for (const auto &y : cs.l2g) seen = seen || (y.cell == x.cell);
      if (!seen) cs.l2g.push_back(x);
    }
    cs.order[s] = deduce_order(int(m.size()), of, g.dim);
    cs.size[s] = required_size(cs.order[s] - 1, of, g.dim);
    cs.st[s] = m;
  }
  int kh = 0;
  for (int s = 1; s < ns; ++s) {
    int ko = cs.order[kh];
    int bias = full ? p.biases[s] : 0;
    if (cs.order[s] > ko) kh = s;
    else if (cs.order[s] == ko && bias == 0) kh = s;
  }
  cs.khigh = kh;
}

// ---- pivoted LDLT and pseudo-inverse -----------------------------------------
// Diagonal-pivoting LDL^T of a symmetric matrix (lower part used), the
// factorisation the reference takes through Eigen::LDLT.
void pseudo_inverse(const std::vector<double> &A, int rows, int cols, std::vector<double> &P, double &cond) {
  std::vector<double> M(size_t(cols) * cols, 0.0);
  for (int a = 0; a < cols; ++a)
    for (int b = 0; b < cols; ++b) {
      double s = 0.0;
      for (int r = 0; r < rows; ++r) s += A[size_t(r) * cols + a] * A[size_t(r) * cols + b];
      M[size_t(a) * cols + b] = s;
    }
  const int n = cols;
  auto m = [&](int r, int c) -> double & { return M[size_t(r) * n + c]; };
  std::vector<int> tr(n);
  std::vector<double> tmp(n);
  for (int k = 0; k < n; ++k) {
    int big = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(m(i, i)) > std::abs(m(big, big))) big = i;
    tr[k] = big;
    if (big != k) {
      for (int j = 0; j < k; ++j) std::swap(m(k, j), m(big, j));
      for (int i = big + 1; i < n; ++i) std::swap(m(i, k), m(i, big));
      std::swap(m(k, k), m(big, big));
      for (int i = k + 1; i < big; ++i) {
        double t = m(i, k);
        m(i, k) = m(big, i);
        m(big, i) = t;
      }
    }
    if (k > 0) {
      for (int j = 0; j < k; ++j) tmp[j] = m(j, j) * m(k, j);
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += m(k, j) * tmp[j];
      m(k, k) -= s;
      for (int i = k + 1; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < k; ++j) t += m(i, j) * tmp[j];
        m(i, k) -= t;
      }
    }
    double akk = m(k, k);
    if (k == 0 && !(std::abs(akk) > 0.0)) {
      for (int j = 0; j < n; ++j) tr[j] = j;
      break;
    }
    if (std::abs(akk) > 0.0)
      for (int i = k + 1; i < n; ++i) m(i, k) /= akk;
  }
  double dmax = 0.0, dmin = 1e300;
  for (int k = 0; k < n; ++k) {
    dmax = std::max(dmax, std::abs(m(k, k)));
    dmin = std::min(dmin, std::abs(m(k, k)));
  }
  cond = dmin > 0 ? dmax / dmin : 1e300;
  // P = (A^T A)^{-1} A^T: solve with the columns of A^T as right-hand sides
  P.assign(size_t(cols) * rows, 0.0);
  std::vector<double> b(size_t(n) * rows);
  for (int a = 0; a < n; ++a)
    for (int r = 0; r < rows; ++r) b[size_t(a) * rows + r] = A[size_t(r) * cols + a];
  auto row_swap = [&](int x, int y) {
    if (x != y)
      for (int r = 0; r < rows; ++r) std::swap(b[size_t(x) * rows + r], b[size_t(y) * rows + r]);
  };
  for (int k = 0; k < n; ++k) row_swap(k, tr[k]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      for (int r = 0; r < rows; ++r) b[size_t(i) * rows + r] -= m(i, j) * b[size_t(j) * rows + r];
  for (int i = 0; i < n; ++i) {
    double d = m(i, i);
    for (int r = 0; r < rows; ++r) b[size_t(i) * rows + r] = std::abs(d) > 2.2250738585072014e-308 ? b[size_t(i) * rows + r] / d : 0.0;
  }
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j)
      for (int r = 0; r < rows; ++r) b[size_t(i) * rows + r] -= m(j, i) * b[size_t(j) * rows + r];
  for (int k = n - 1; k >= 0; --k) row_swap(k, tr[k]);
  P = b;
}

}  // namespace

// ---- Hilbert order -------------------------------------------------------------
uint64_t hilbert2(uint32_t x, uint32_t y, int bits) {
  uint64_t d = 0;
  for (uint32_t s = 1u << (bits - 1); s > 0; s >>= 1) {
    uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
    d += uint64_t(s) * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

// Skilling's transpose form for 3D
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t X[3] = {x, y, z};
  uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) X[0] ^= P;
      else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t d = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) d = (d << 1) | ((X[i] >> b) & 1u);
  return d;
}


bool volume_mean_is_constant_term(const zfv_params &p, const Shape &s) {
  int degc = 0;
  while (dof(degc + 1, s.dim) <= s.dofc) ++degc;
  const int mdeg = p.moments_deg == -1 ? (s.dim == 2 ? 5 : 3) : p.moments_deg;
  return p.volume_deg >= degc && mdeg >= degc;
}

Shape shape_for(const zfv_params &p) {
  Shape s;
  s.dim = p.n_dims;
  if (s.dim != 2 && s.dim != 3) throw ZfvError(ZFV_ERR_INVALID, "n_dims must be 2 or 3");
  s.nvars = s.dim + 2;
  s.nfaces = s.dim + 1;
  if (p.n_stencils < 1 || p.n_stencils > ZFV_MAX_STENCILS) throw ZfvError(ZFV_ERR_INVALID, "n_stencils out of range");
  if (p.biases[0] != 0) throw ZfvError(ZFV_ERR_INVALID, "the first stencil must be central");
  for (int k = 1; k < p.n_stencils; ++k)
    if (p.biases[k] != 1 || p.orders[k] != 2)
      throw ZfvError(ZFV_ERR_INVALID, "kernels support one central stencil followed by order-2 one-sided stencils");
  s.nb = p.n_stencils - 1;
  if (s.nb > s.nfaces) throw ZfvError(ZFV_ERR_INVALID, "more one-sided stencils than faces");
  int degc = p.orders[0] - 1;
  s.dofc = dof(degc, s.dim);
  s.rc = required_size(degc, p.overfit[0], s.dim) - 1;
  s.dofb = dof(1, s.dim);
  s.rb = s.nb > 0 ? required_size(1, p.overfit[1], s.dim) - 1 : 0;
  for (int k = 2; k < p.n_stencils; ++k)
    if (p.overfit[k] != p.overfit[1]) throw ZfvError(ZFV_ERR_INVALID, "one-sided stencils must share an overfit factor");
  s.m = (s.dofc - 1) * s.rc + s.nb * (s.dofb - 1) * s.rb;
  s.nrows = s.rc + s.nb * s.rb;
  s.nmom = std::max(0, s.dofc - 3);
  s.q = int((s.dim == 2 ? gl_edge(p.edge_deg) : dunavant(p.edge_deg)).w.size());
  s.qv = int(volume_rule(s.dim, p.volume_deg).w.size());
  // the reference rejects unknown modes (LOG_ERR in its factories); so does
  // every enumerated field here, rather than running another physics
  if (p.reconstruction != ZFV_CWENO_AO && p.reconstruction != ZFV_WENO_AO)
    throw ZfvError(ZFV_ERR_INVALID, "reconstruction must be ZFV_CWENO_AO (0) or ZFV_WENO_AO (1)");
  if (p.flux != ZFV_FLUX_HLLC && p.flux != ZFV_FLUX_RUSANOV)
    throw ZfvError(ZFV_ERR_INVALID, "flux must be ZFV_FLUX_HLLC (0) or ZFV_FLUX_RUSANOV (1)");
  if (p.equilibrium < ZFV_EQ_NONE || p.equilibrium > ZFV_EQ_ISOTHERMAL)
    throw ZfvError(ZFV_ERR_INVALID, "equilibrium must be ZFV_EQ_NONE, ZFV_EQ_ISENTROPIC or ZFV_EQ_ISOTHERMAL");
  if (p.gravity < ZFV_GRAVITY_NONE || p.gravity > ZFV_GRAVITY_POLYTROPE)
    throw ZfvError(ZFV_ERR_INVALID, "gravity must be one of ZFV_GRAVITY_NONE/AXIAL/RADIAL/POLYTROPE");
  if (p.flux_bc < ZFV_FLUX_BC_NONE || p.flux_bc > ZFV_FLUX_BC_ISENTROPIC)
    throw ZfvError(ZFV_ERR_INVALID, "flux_bc must be ZFV_FLUX_BC_NONE (0), _CONSTANT (1) or _ISENTROPIC (2)");
  if (p.flux_bc == ZFV_FLUX_BC_ISENTROPIC && p.equilibrium != ZFV_EQ_ISENTROPIC)
    throw ZfvError(ZFV_ERR_INVALID,
                   "flux_bc isentropic (EquilibriumFluxBC) is implemented with isentropic well-balancing only");
  if (!(p.gamma > 1.0)) throw ZfvError(ZFV_ERR_INVALID, "gamma must be > 1");
  if (p.scaling != 0 && p.scaling != 1)
    throw ZfvError(ZFV_ERR_INVALID, "scaling must be 0 (EulerScaling) or 1 (UnityScaling)");
  if (p.steps_per_recompute < 1) throw ZfvError(ZFV_ERR_INVALID, "steps_per_recompute must be >= 1");
  if (p.steps_per_recompute > 1 && p.flux_bc == ZFV_FLUX_BC_ISENTROPIC)
    throw ZfvError(ZFV_ERR_INVALID,
                   "steps_per_recompute > 1 with flux_bc isentropic is not implemented: EquilibriumFluxBC solves the "
                   "current equilibrium on every call while the reconstruction keeps its own");
  if (p.equilibrium != ZFV_EQ_NONE && p.gravity == ZFV_GRAVITY_NONE)
    throw ZfvError(ZFV_ERR_INVALID, "a well-balanced equilibrium needs gravity");
  return s;
}

void build_image(const zfv_mesh &mesh, const zfv_params &p, HostImage &img) {
  if (mesh.n_dims != p.n_dims) throw ZfvError(ZFV_ERR_INVALID, "mesh and params disagree on n_dims");
  if (mesh.n_cells <= 0) throw ZfvError(ZFV_ERR_INVALID, "empty mesh");
  const Shape s = shape_for(p);
  img.s = s;
  Geo g;
  build_geometry(mesh, p, g);
  const int n = g.n, nv = g.nv, dim = g.dim;
  img.n = n;

  // stencils + pseudo-inverses, caller numbering
  StencilBuilder sb(g, p.edge_deg);
  std::vector<CellStencils> cs(n);
  std::vector<double> Pc(size_t(n) * s.m, 0.0);
  std::vector<int32_t> rows_c(size_t(n) * s.nrows, 0);
  std::vector<double> cond(n, 0.0);
  std::vector<V3> row_shift(p.equilibrium == ZFV_EQ_NONE ? 0 : size_t(n) * s.nrows, V3{0.0, 0.0, 0.0});
  int err = 0;
  std::string err_msg;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    try {
      cell_stencils(sb, p, i, cs[i]);
      const CellStencils &c = cs[i];
      double *P = &Pc[size_t(i) * s.m];
      int32_t *R = &rows_c[size_t(i) * s.nrows];
      for (int r = 0; r < s.nrows; ++r) R[r] = i;  // padding rows read the cell itself (zero weight)
      if (c.first_order) continue;
      for (int k = 0; k < c.n; ++k) {
        int order = c.order[k];
        int rows = c.size[k] - 1;
        int dofk = (k == 0 ? s.dofc : s.dofb);
        int rmax = (k == 0 ? s.rc : s.rb);
        int off = (k == 0 ? 0 : (s.dofc - 1) * s.rc + (k - 1) * (s.dofb - 1) * s.rb);
        int roff = (k == 0 ? 0 : s.rc + (k - 1) * s.rb);
        if (order <= 1 || rows <= 0) continue;
        if (rows > rmax || dof(order - 1, dim) > dofk) throw ZfvError(ZFV_ERR_INVALID, "stencil larger than the kernel shape");
        std::vector<double> A, Pk;
        int cols;
        lsq_matrix(g, c.st[k], rows, order, A, cols);
        double cn;
        pseudo_inverse(A, rows, cols, Pk, cn);
        cond[i] = std::max(cond[i], cn);
        // P block of stencil k: [alpha (dofk-1)][row (rmax)], zero padded
        for (int a = 0; a < cols; ++a)
          for (int r = 0; r < rows; ++r) P[off + a * rmax + r] = Pk[size_t(a) * rows + r];
        for (int r = 0; r < rows; ++r) R[roff + r] = c.st[k][r + 1].cell;
      }
      // periodic translations of the members (read only by the equilibrium
      // averages of well-balanced runs); coded after the loop
      if (p.equilibrium == ZFV_EQ_NONE) continue;
      for (int k = 0; k < c.n; ++k) {
        int rows = c.size[k] - 1;
        int roff = (k == 0 ? 0 : s.rc + (k - 1) * s.rb);
        if (c.order[k] <= 1 || rows <= 0) continue;
        for (int r = 0; r < rows; ++r) row_shift[size_t(i) * s.nrows + roff + r] = c.st[k][r + 1].shift;
      }
    } catch (const std::exception &e) {
#pragma omp critical
      {
        err = 1;
        err_msg = e.what();
      }
    }
  }
  if (err) throw ZfvError(ZFV_ERR_INVALID, err_msg);
  img.max_cond = *std::max_element(cond.begin(), cond.end());
  if (n > (1 << 27)) throw ZfvError(ZFV_ERR_INVALID, "more than 2^27 cells");
  // Translation table and row codes, in caller order so that the same input
  // always gives the same table.  Translations computed from different
  // corner pairs agree to round-off: they are matched with a tolerance.
  {
    std::vector<V3> table(1, V3{0.0, 0.0, 0.0});
    for (size_t q = 0; q < row_shift.size(); ++q) {
      const V3 &t = row_shift[q];
      if (t[0] == 0.0 && t[1] == 0.0 && t[2] == 0.0) continue;
      const double tol = 1e-9 * std::max(1.0, std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
      int code = -1;
      for (int k = 1; k < int(table.size()) && code < 0; ++k)
        if (std::fabs(table[k][0] - t[0]) <= tol && std::fabs(table[k][1] - t[1]) <= tol &&
            std::fabs(table[k][2] - t[2]) <= tol)
          code = k;
      if (code < 0) {
        if (table.size() == 16) throw ZfvError(ZFV_ERR_INVALID, "more than 15 distinct periodic translations");
        table.push_back(t);
        code = int(table.size()) - 1;
      }
      rows_c[q] |= int32_t(code) << 27;
    }
    for (size_t q = 0; q < table.size(); ++q)
      for (int d = 0; d < 3; ++d) img.shift[3 * q + d] = table[q][d];
  }

  // reference-layout stencil export (caller numbering)
  img.l2g_off.assign(n + 1, 0);
  img.mem_off.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    img.l2g_off[i + 1] = img.l2g_off[i] + int(cs[i].l2g.size());
    int w = 0;
    for (int k = 0; k < cs[i].n; ++k) w += 2 + int(cs[i].st[k].size());
    img.mem_off[i + 1] = img.mem_off[i] + w;
  }
  img.l2g.resize(img.l2g_off[n]);
  img.members.resize(img.mem_off[n]);
  for (int i = 0; i < n; ++i) {
    int o = img.l2g_off[i];
    for (const auto &m : cs[i].l2g) img.l2g[o++] = m.cell;
    o = img.mem_off[i];
    for (int k = 0; k < cs[i].n; ++k) {
      img.members[o++] = cs[i].order[k];
      img.members[o++] = int(cs[i].st[k].size());
      for (const auto &m : cs[i].st[k]) img.members[o++] = m.cell;
    }
  }

  // Hilbert renumbering of cell centroids
  std::vector<uint64_t> key(n);
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < dim; ++d) {
        lo[d] = std::min(lo[d], g.ctr[i][d]);
        hi[d] = std::max(hi[d], g.ctr[i][d]);
      }
    const int bits = dim == 2 ? 20 : 16;
    const double scale = double((1u << bits) - 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      uint32_t c[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        double t = hi[d] > lo[d] ? (g.ctr[i][d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
        c[d] = uint32_t(t * scale);
      }
      key[i] = dim == 2 ? hilbert2(c[0], c[1], bits) : hilbert3(c[0], c[1], c[2], bits);
    }
  }
  img.perm.resize(n);
  std::iota(img.perm.begin(), img.perm.end(), 0);
  std::stable_sort(img.perm.begin(), img.perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  img.iperm.resize(n);
  for (int i = 0; i < n; ++i) img.iperm[img.perm[i]] = i;

  // faces: interior faces numbered in product order of first encounter; the
  // left cell is the one with the lower caller index (src/zisa/grid/grid.cpp:415-442).
  // Faces between two ghost cells get no number: the reference's flux loop
  // drops them (include/zisa/fvm_loops/flux_loop.hpp:71-83; every ghost cell
  // is an "exterior" cell there, so no such face is ever "interior").
  constexpr int32_t FID_GHOST_PAIR = -2;
  std::vector<int32_t> fid(size_t(n) * nv, -1);  // caller numbering [i][k]
  int64_t ne = 0;
  for (int ip = 0; ip < n; ++ip) {
    int i = img.perm[ip];
    for (int k = 0; k < nv; ++k) {
      int j = g.nb(i, k);
      if (j < 0 || fid[size_t(i) * nv + k] != -1) continue;
      if (g.ghost[i] && g.ghost[j]) {
        fid[size_t(i) * nv + k] = FID_GHOST_PAIR;
        continue;
      }
      int32_t e = int32_t(ne++);
      fid[size_t(i) * nv + k] = e;
      for (int m = 0; m < nv; ++m)
        if (g.nb(j, m) == i) fid[size_t(j) * nv + m] = e;
    }
  }
  img.ne = ne;

  // pack SoA in product order
  const size_t N = size_t(n);
  img.P.assign(size_t(s.m) * N, 0.0);
  img.rows.assign(size_t(s.nrows) * N, 0);
  img.twin.assign(size_t(std::max(s.nrows - s.rc, 1)) * N, int8_t(-1));
  img.crn.assign(size_t(dim == 2 ? 6 : 16) * N, 0.0);  // per cell, each vertex on 16 bytes
  const int xw = dim == 2 ? 2 : 4, mw = (s.nmom + 1) & ~1;  // per-cell centre and moment words
  img.xc.assign(size_t(xw) * N, 0.0);
  img.len.assign(N, 0.0);
  img.vol.assign(N, 0.0);
  img.inr.assign(N, 0.0);
  img.mom.assign(size_t(std::max(mw, 1)) * N, 0.0);
  img.meta.assign(N, 0);
  img.cface.assign(size_t(4) * N, 0);  // per cell 4 words (2D: the fourth unused)
  img.fperm.assign(N, 0u);
  img.ghost.assign(N, 0);
#pragma omp parallel for schedule(static)
  for (int ip = 0; ip < n; ++ip) {
    const int i = img.perm[ip];
    // central coefficient a of row r (flat a rc + r) at p_central, one-sided
    // coefficient a of row rr of stencil kb (flat kb nbc rb + a rb + rr after
    // the central ones) at p_onesided
    const int nc = s.dofc - 1, nbc = s.dofb - 1;
    for (int m = 0; m < s.m; ++m) {
      size_t at;
      if (m < nc * s.rc) {
        at = size_t(p_central(nc, s.rc, n, m / s.rc, m % s.rc, ip));
      } else {
        const int o = m - nc * s.rc, kb = o / (nbc * s.rb), a = (o % (nbc * s.rb)) / s.rb, rr = o % s.rb;
        at = size_t(p_onesided(nc, s.rc, nbc, s.nb, s.rb, n, kb, a, rr, ip));
      }
      img.P[at] = Pc[size_t(i) * s.m + m];
    }
    for (int r = 0; r < s.nrows; ++r) {
      const int32_t rc = rows_c[size_t(i) * s.nrows + r];
      img.rows[size_t(r) * N + ip] = img.iperm[rc & ((1 << 27) - 1)] | (rc & ~((1 << 27) - 1));
    }
    for (int r = s.rc; r < s.nrows; ++r)
      for (int rp = 0; rp < s.rc && rp < 127; ++rp)
        if (img.rows[size_t(rp) * N + ip] == img.rows[size_t(r) * N + ip]) {
          img.twin[size_t(r - s.rc) * N + ip] = int8_t(rp);
          break;
        }
    for (int k = 0; k < nv; ++k)
      for (int d = 0; d < dim; ++d) img.crn[size_t(ip) * (dim == 2 ? 6 : 16) + size_t(k) * (dim == 2 ? 2 : 4) + d] = g.c(i, k)[d];
    for (int d = 0; d < dim; ++d) img.xc[size_t(ip) * xw + d] = g.ctr[i][d];
    img.len[ip] = g.len[i];
    img.vol[ip] = g.vol[i];
    img.inr[ip] = g.inr[i];
    for (int a = 3; a < s.dofc; ++a)
      img.mom[size_t(ip) * mw + (a - 3)] = a < int(g.mom[i].size()) ? g.mom[i][a] : 0.0;
    int32_t meta = cs[i].khigh & META_KHIGH_MASK;
    if (cs[i].first_order) meta |= META_FIRST_ORDER;
    if (g.ghost[i]) meta |= META_GHOST;
    img.meta[ip] = meta;
    img.ghost[ip] = g.ghost[i];
    uint32_t fp = 0;
    for (int k = 0; k < nv; ++k) {
      int j = g.nb(i, k);
      int32_t cf;
      int lv[3] = {face_vertex(dim, k, 0), face_vertex(dim, k, 1), dim == 3 ? face_vertex(dim, k, 2) : 0};
      if (j < 0) {
        cf = CFACE_EXTERIOR;
      } else {
        bool right = j < i;  // left = lower caller index
        const int32_t e = fid[size_t(i) * nv + k];
        cf = e == FID_GHOST_PAIR ? CFACE_GHOST_PAIR : (e << 2) | (right ? CFACE_SIDE_R : 0);
        if (right) {
          // express the left cell's vertex order of this face in own corners
          int kj = -1;
          for (int m = 0; m < nv; ++m)
            if (g.nb(j, m) == i) kj = m;
          for (int r = 0; r < dim; ++r) {
            int idj = g.vid[size_t(j) * nv + face_vertex(dim, kj, r)];
            for (int m = 0; m < nv; ++m)
              if (g.vid[size_t(i) * nv + m] == idj) lv[r] = m;
          }
        }
      }
      img.cface[size_t(ip) * 4 + k] = cf;
      for (int r = 0; r < dim; ++r) fp |= uint32_t(lv[r]) << (2 * (k * 3 + r));
    }
    img.fperm[ip] = fp;
  }
  if (int64_t(n) * nv > INT32_MAX) throw ZfvError(ZFV_ERR_INVALID, "mesh too large for 32-bit cell-face slots");
  img.fslot.assign(size_t(2) * std::max<int64_t>(img.ne, 1), 0);
#pragma omp parallel for schedule(static)
  for (int ip = 0; ip < n; ++ip)
    for (int k = 0; k < nv; ++k) {
      const int32_t cf = img.cface[size_t(ip) * 4 + k];
      if (cf & CFACE_EXTERIOR) continue;
      img.fslot[size_t(cf >> 2) * 2 + size_t(cf & CFACE_SIDE_R)] = ip * nv + k;
    }

  // normalized rules for the kernels
  Rule fr = dim == 2 ? gl_edge(p.edge_deg) : dunavant(p.edge_deg);
  for (size_t q = 0; q < fr.w.size(); ++q) {
    img.face_w.push_back(fr.w[q]);
    for (int d = 0; d < (dim == 2 ? 1 : 3); ++d) img.face_x.push_back(fr.x[q][d]);
  }
  Rule vr = volume_rule(dim, p.volume_deg);
  for (size_t q = 0; q < vr.w.size(); ++q) {
    img.vol_w.push_back(vr.w[q]);
    for (int d = 0; d < dim + 1; ++d) img.vol_x.push_back(vr.x[q][d]);
  }
  double tot = 0.0;
  for (int k = 0; k < p.n_stencils; ++k) tot += p.linear_weights[k];
  for (int k = 0; k < p.n_stencils; ++k) img.linear_weights[k] = p.linear_weights[k] / tot;
}

void to_soa(const HostImage &img, const double *aos, double *soa) {
  const int64_t n = img.n;
  const int nv = img.s.nvars, dim = img.s.dim;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < n; ++ip) {
    const double *u = aos + 5 * int64_t(img.perm[ip]);
    soa[ip] = u[0];
    for (int d = 0; d < dim; ++d) soa[(1 + d) * n + ip] = u[1 + d];
    soa[(nv - 1) * n + ip] = u[4];
  }
}

void from_soa(const HostImage &img, const double *soa, double *aos, bool accumulate) {
  const int64_t n = img.n;
  const int nv = img.s.nvars, dim = img.s.dim;
  // perm is a permutation: every iteration writes its own row
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < n; ++ip) {
    double *u = aos + 5 * int64_t(img.perm[ip]);
    double w[5] = {soa[ip], 0.0, 0.0, 0.0, soa[(nv - 1) * n + ip]};
    for (int d = 0; d < dim; ++d) w[1 + d] = soa[(1 + d) * n + ip];
    for (int v = 0; v < 5; ++v) u[v] = accumulate ? u[v] + w[v] : w[v];
  }
}

}  // namespace zfv
