// Seeded synthetic meshes for the benchmark workloads.  The reference loads
// meshes from .msh.h5 files (src/zisa/grid/grid.cpp:889-922) and has no
// generator; these produce the jittered-lattice meshes the workloads name.
#include "common.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace zfv {

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [-1, 1)
  double sym() { return 2.0 * (double(next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0; }
};

double angle_at(const double *p, const double *a, const double *b) {
  double ux = a[0] - p[0], uy = a[1] - p[1], vx = b[0] - p[0], vy = b[1] - p[1];
  double c = (ux * vx + uy * vy) / std::sqrt((ux * ux + uy * uy) * (vx * vx + vy * vy));
  c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
  return std::acos(c);
}

}  // namespace

MeshBuf *mesh_lattice2d(int n, double jitter, uint64_t seed, bool px, bool py, int ghost_layers) {
  if (n < 4) throw ZfvError(ZFV_ERR_INVALID, "lattice2d: n must be >= 4");
  if (!(jitter >= 0.0 && jitter < 0.3)) throw ZfvError(ZFV_ERR_INVALID, "lattice2d: jitter must be in [0, 0.3)");
  const int gx = px ? 0 : ghost_layers, gy = py ? 0 : ghost_layers;
  // vertex index range per direction: periodic -> [0, n), else [-g, n + g]
  const int nvx = px ? n : n + 1 + 2 * gx, nvy = py ? n : n + 1 + 2 * gy;
  const double h = 1.0 / n;
  Rng rng{seed};
  std::vector<double> vx(size_t(nvx) * nvy), vy(size_t(nvx) * nvy);
  for (int j = 0; j < nvy; ++j)
    for (int i = 0; i < nvx; ++i) {
      int ii = px ? i : i - gx, jj = py ? j : j - gy;
      double dx = jitter * rng.sym(), dy = jitter * rng.sym();
      // keep the physical boundary x in {0,1} / y in {0,1} straight
      if (!px && (ii == 0 || ii == n)) dx = 0.0;
      if (!py && (jj == 0 || jj == n)) dy = 0.0;
      vx[size_t(j) * nvx + i] = (ii + dx) * h;
      vy[size_t(j) * nvx + i] = (jj + dy) * h;
    }
  const int nsx = px ? n : n + 2 * gx, nsy = py ? n : n + 2 * gy;
  auto *m = new MeshBuf;
  m->n_dims = 2;
  m->n_cells = 2 * nsx * nsy;
  m->vertex_ids.resize(size_t(m->n_cells) * 3);
  m->corners.resize(size_t(m->n_cells) * 9, 0.0);
  m->ghost.assign(m->n_cells, 0);
  int c = 0;
  for (int sj = 0; sj < nsy; ++sj)
    for (int si = 0; si < nsx; ++si) {
      // corner vertex (di, dj) of square (si, sj): id and unwrapped position
      int id[2][2];
      double pos[2][2][2];
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          int i = si + di, j = sj + dj;
          double ox = 0.0, oy = 0.0;
          if (px && i == n) {
            i = 0;
            ox = 1.0;
          }
          if (py && j == n) {
            j = 0;
            oy = 1.0;
          }
          id[dj][di] = j * nvx + i;
          pos[dj][di][0] = vx[size_t(j) * nvx + i] + ox;
          pos[dj][di][1] = vy[size_t(j) * nvx + i] + oy;
        }
      // Delaunay diagonal: split 00-11 unless the opposite angles exceed pi
      const double *p00 = pos[0][0], *p10 = pos[0][1], *p11 = pos[1][1], *p01 = pos[1][0];
      bool diag_0011 = angle_at(p10, p00, p11) + angle_at(p01, p11, p00) <= M_PI;
      int tri[2][3][2];  // [triangle][corner] -> (dj, di), CCW
      if (diag_0011) {
        int t0[3][2] = {{0, 0}, {0, 1}, {1, 1}}, t1[3][2] = {{0, 0}, {1, 1}, {1, 0}};
        std::memcpy(tri[0], t0, sizeof(t0));
        std::memcpy(tri[1], t1, sizeof(t1));
      } else {
        int t0[3][2] = {{0, 0}, {0, 1}, {1, 0}}, t1[3][2] = {{0, 1}, {1, 1}, {1, 0}};
        std::memcpy(tri[0], t0, sizeof(t0));
        std::memcpy(tri[1], t1, sizeof(t1));
      }
      for (int t = 0; t < 2; ++t, ++c) {
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < 3; ++k) {
          int dj = tri[t][k][0], di = tri[t][k][1];
          m->vertex_ids[size_t(c) * 3 + k] = id[dj][di];
          m->corners[size_t(c) * 9 + 3 * k + 0] = pos[dj][di][0];
          m->corners[size_t(c) * 9 + 3 * k + 1] = pos[dj][di][1];
          cx += pos[dj][di][0];
          cy += pos[dj][di][1];
        }
        cx /= 3.0;
        cy /= 3.0;
        bool ghost = (!px && (cx < 0.0 || cx > 1.0)) || (!py && (cy < 0.0 || cy > 1.0));
        m->ghost[c] = ghost ? 1 : 0;
      }
    }
  return m;
}

MeshBuf *mesh_ball3d(int n, double radius, double ghost_radius, double jitter, uint64_t seed) {
  if (n < 4) throw ZfvError(ZFV_ERR_INVALID, "ball3d: n must be >= 4");
  if (!(jitter >= 0.0 && jitter < 0.25)) throw ZfvError(ZFV_ERR_INVALID, "ball3d: jitter must be in [0, 0.25)");
  // lattice over [-radius, radius]^3 with n cubes per axis
  const int nv = n + 1;
  const double h = 2.0 * radius / n;
  Rng rng{seed};
  std::vector<double> P(size_t(nv) * nv * nv * 3);
  for (int k = 0; k < nv; ++k)
    for (int j = 0; j < nv; ++j)
      for (int i = 0; i < nv; ++i) {
        size_t v = (size_t(k) * nv + j) * nv + i;
        P[3 * v + 0] = -radius + (i + jitter * rng.sym()) * h;
        P[3 * v + 1] = -radius + (j + jitter * rng.sym()) * h;
        P[3 * v + 2] = -radius + (k + jitter * rng.sym()) * h;
      }
  // Kuhn (Freudenthal) subdivision: one tet per axis permutation, all sharing
  // the cube diagonal 000-111; conforming across cube faces.
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  auto *m = new MeshBuf;
  m->n_dims = 3;
  for (int ck = 0; ck < n; ++ck)
    for (int cj = 0; cj < n; ++cj)
      for (int ci = 0; ci < n; ++ci)
        for (int t = 0; t < 6; ++t) {
          int o[3] = {ci, cj, ck};
          int ids[4];
          double pts[4][3];
          for (int q = 0; q < 4; ++q) {
            if (q > 0) o[perm[t][q - 1]] += 1;
            size_t v = (size_t(o[2]) * nv + o[1]) * nv + o[0];
            ids[q] = int(v);
            for (int d = 0; d < 3; ++d) pts[q][d] = P[3 * v + d];
          }
          double c[3] = {0, 0, 0};
          for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 3; ++d) c[d] += 0.25 * pts[q][d];
          double r = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
          if (r > radius) continue;
          for (int q = 0; q < 4; ++q) {
            m->vertex_ids.push_back(ids[q]);
            for (int d = 0; d < 3; ++d) m->corners.push_back(pts[q][d]);
          }
          m->ghost.push_back(r > ghost_radius ? 1 : 0);
        }
  m->n_cells = int(m->ghost.size());
  return m;
}

}  // namespace zfv
