// Per-cell and per-face work of one rate-of-change evaluation, written once
// for the gfx950 kernels (kernels.hip).  Everything is FP64.
//
// Reference semantics restated here:
//   reconstruction   include/zisa/reconstruction/local_reconstruction.hpp:69-156,
//                    src/zisa/reconstruction/{cweno_ao,weno_ao,hybrid_weno}.cpp,
//                    include/zisa/math/poly2d_impl.hpp:271-337
//   face flux        include/zisa/fvm_loops/flux_loop.hpp:126-193,
//                    include/zisa/flux/hllc.hpp:17-176
//   gravity source   include/zisa/fvm_loops/gravity_source_loop.hpp:32-148
//   equilibria       include/zisa/model/local_equilibrium_impl.hpp:41-93,
//                    include/zisa/model/isentropic_equilibrium.hpp:28-63
//   flux BC          include/zisa/boundary/flux_bc.hpp:25-42
//   RK stage         src/zisa/ode/runge_kutta.cpp:87-143
#pragma once

// Rounding: everything here is compiled with -ffp-contract=off, so no
// compiler decides where to fuse a multiply-add; the least-squares
// accumulation and the polynomial evaluation use explicit fma().  The same
// source therefore rounds identically in every kernel and in the host
// emulation, whatever the surrounding code.

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ZFV_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define ZFV_HD inline
#endif

namespace zfv {

// cell meta word: bits 0-2 index of the CWENO high-order stencil, 8 = first
// order (a ghost cell away from the interior), 16 = ghost cell
constexpr int32_t META_FAR_GHOST = 8 | 16;

constexpr int ROW_CELL_BITS = 27;
constexpr int32_t ROW_CELL_MASK = (1 << ROW_CELL_BITS) - 1;

// What the kernels read.  Pointers are device pointers (or host pointers in
// the CPU emulation used by the tests).
struct Img {
  int64_t n;   // cells
  int64_t ne;  // interior faces
  const double *P;
  const int32_t *rows;  // cell | shift_code << 27
  const int8_t *twin;   // (nrows - rc) * n: central row with the same member, or -1
  const double *crn;
  const double *xc;
  const double *len;
  const double *vol;
  const double *inr;
  const double *mom;
  const int32_t *meta;
  const int32_t *cface;  // per cell its face words, 4 int32 (one 16-byte load; 2D's fourth unused)
  const int32_t *fslot;  // ne pairs: cell-face slots of each interior face's left and right cell
  const uint32_t *fperm;
  const double *steady;  // nvars * n (ghost cells only are read)
  const double *phim;    // n pairs: cell mean of phi and of (phi - mean)^2 (one 16-byte load), or null
  const double *pot;     // potential tables of the cell's own points (pot_index), or null
  const double *frm;     // face frames of the cells' local faces (load_face_frame)
  const double *ffr;     // per interior face its frame and area, t1, area (2D) or n, t1, area (3D), x ne
  // gravity: per cell vol sum_q vw_q b_alpha(x_q) d_d phi(x_q) over its volume
  // points, n x (alpha DIM + d), each cell's DOFC x DIM values contiguous
  // (store_source_moments), or null
  const double *smom;
  // the potential at every cell's corners ((d + 1) x n; the light set with an
  // equilibrium, whose axial potential is affine), or null
  const double *cphi;
  // stateful equilibria (steps_per_recompute > 1), or null: per cell theta
  // (2), found flag, EulerScaling (rho, a, E_int) and the equilibrium's own
  // cell average (rho, E) of the last recompute (8 x n), and calls since it
  // (n); launches that must leave them alone (autotuning, kernel timing) run
  // with eqc null and recompute
  double *eqc;
  int32_t *eqcnt;
  int spr;               // steps_per_recompute
  double eq_threshold;   // recompute_threshold
  // physics
  double gamma;
  double eps, exponent;
  double lw[5];  // normalized linear weights
  int mode;      // 0 CWENO-AO, 1 WENO-AO
  int scaling;   // 0 EulerScaling, 1 UnityScaling
  int vol_mean;  // 1: a reconstruction's volume-rule mean over its cell is its constant coefficient
  int flux;      // 0 HLLC, 1 Rusanov
  int eq;        // 0 none, 1 isentropic, 2 isothermal
  int gravity;   // 0 none, 1 axial, 2 radial, 3 polytrope
  double g, axis[3], rhoC, K, G;
  int flux_bc;
  double shift[16][3];  // periodic translations addressed by the row shift code
  double shift_phi[16];  // axial potential of each translation (added to cphi)
  // normalized quadrature rules
  int qv;                   // points of the volume rule
  double fw[8], fx[8][3];   // face rule: weights, GL abscissa (2D) or barycentric (3D)
  double vw[16], vx[16][4]; // volume rule
};

template <int DIM>
struct Dims {
  static constexpr int NV = DIM + 2;
  static constexpr int NF = DIM + 1;
  static constexpr int DOFB = DIM + 1;
};

ZFV_HD double sqr(double x) { return x * x; }

ZFV_HD double dmin(double a, double b) { return a < b ? a : b; }
ZFV_HD double dmax(double a, double b) { return a > b ? a : b; }

// two adjacent doubles, one 16-byte load
struct alignas(16) PPair {
  double lo, hi;
};

// an interior face's two cell-face slots (left, right), one 8-byte load
struct alignas(8) SlotPair {
  int32_t l, r;
};

// Face fluxes G: 2D per face its 4 values (two 16-byte pieces), 3D one
// plane of ne values per variable (an odd count would need a pad word per
// face; measured slower)

// a cell's four face words (face id << 2 | exterior << 1 | side), one load
struct alignas(16) FaceWords {
  int32_t w[4];
};
ZFV_HD FaceWords load_face_words(const Img &m, int64_t i) { return reinterpret_cast<const FaceWords *>(m.cface)[i]; }

// ---- gravity -----------------------------------------------------------------
// PH selects the physics a kernel is compiled for: PH_LIGHT covers no or axial
// gravity with no or isothermal equilibrium (exp only); PH_FULL adds radial
// and polytrope gravity and the isentropic equilibrium.  PH_LIN is PH_FULL
// restricted to the isentropic equilibrium of gamma = 2 with potential
// moments and no periodic images, where a member's equilibrium average needs
// only the member's two potential moments (linear_isentropic_avg), not its
// corners.  The transcendental functions of the full set and the member
// corners need many registers, so runs that do not need them use kernels
// without them.
constexpr int PH_LIGHT = 0;
constexpr int PH_FULL = 1;
constexpr int PH_LIN = 2;

template <int DIM, int PH = PH_FULL>
ZFV_HD double phi(const Img &m, const double *x) {
  if (m.gravity == 1) {
    double d = x[0] * m.axis[0] + x[1] * m.axis[1];
    if (DIM == 3) d += x[2] * m.axis[2];
    return m.g * d;
  }
  if (PH == PH_LIGHT) return 0.0;
  double r2 = x[0] * x[0] + x[1] * x[1];
  if (DIM == 3) r2 += x[2] * x[2];
  double r = sqrt(r2);
  if (m.gravity == 2) return m.g * r;
  if (m.gravity == 3) {
    double alpha = sqrt(2.0 * 3.14159265358979323846 * m.G / m.K);
    double ce = alpha * (r + 2.2250738585072014e-308);
    return -2.0 * m.K * m.rhoC * sin(ce) / ce;
  }
  return 0.0;
}

template <int DIM, int PH = PH_FULL>
ZFV_HD void grad_phi(const Img &m, const double *x, double *gp) {
  if (m.gravity == 1 || PH == PH_LIGHT) {
    for (int d = 0; d < DIM; ++d) gp[d] = m.gravity == 1 ? m.g * m.axis[d] : 0.0;
    return;
  }
  double r2 = x[0] * x[0] + x[1] * x[1];
  if (DIM == 3) r2 += x[2] * x[2];
  double r = sqrt(r2);
  double dd = 0.0;
  if (m.gravity == 2) {
    dd = m.g;
  } else if (m.gravity == 3) {
    double alpha = sqrt(2.0 * 3.14159265358979323846 * m.G / m.K);
    double ce = alpha * (r + 2.2250738585072014e-308);
    double dphi = (cos(ce) - sin(ce) / ce) / ce;
    dd = -2.0 * m.K * m.rhoC * dphi * alpha;
  }
  for (int d = 0; d < DIM; ++d) gp[d] = dd * (x[d] / (r + 1e-50));
}

// ---- local equilibrium: theta -> (rho, E) at x ------------------------------
template <int DIM, int PH = PH_FULL>
ZFV_HD void eq_eval(const Img &m, const double *th, double phi_ref, const double *x, double &rho, double &E) {
  const double gamma = m.gamma;
  if (PH != PH_LIGHT && m.eq == 1) {
    double h = th[0] + phi_ref - phi<DIM, PH>(m, x);
    double K = th[1];
    double base = 1.0 / K * (gamma - 1.0) / gamma * h;
    rho = pow(base, 1.0 / (gamma - 1.0));
    double p = K * pow(rho, gamma);
    E = p / (gamma - 1.0);
  } else {
    double RT = th[1];
    rho = th[0] * exp(-(phi<DIM, PH>(m, x) - phi_ref) / RT);
    E = rho * RT / (gamma - 1.0);
  }
}

// The same equilibrium with its per-cell constants hoisted out of the point
// loops: isentropic (the reference's): a = (1/K)(gamma-1)/gamma and
// b = 1/(gamma-1), the very subexpressions eq_eval forms at every point, so
// the values are identical; isothermal (an extension): a = -1/RT and
// b = RT/(gamma-1), which replace two divisions per point by products.
struct EqC {
  double th0, th1, phi_ref, a, b;
};

template <int PH>
ZFV_HD EqC eq_consts(const Img &m, const double *th, double phi_ref) {
  EqC c;
  c.th0 = th[0];
  c.th1 = th[1];
  c.phi_ref = phi_ref;
  if (PH != PH_LIGHT && m.eq == 1) {
    c.a = 1.0 / th[1] * (m.gamma - 1.0) / m.gamma;
    c.b = 1.0 / (m.gamma - 1.0);
  } else {
    c.a = -1.0 / th[1];
    c.b = th[1] / (m.gamma - 1.0);
  }
  return c;
}

// x^y with the exponents an ideal gas of gamma = 2 (the n = 1 polytrope)
// produces, 1 and 2, done exactly (x^1 = x; x^2 is one correctly rounded
// product); any other exponent goes to pow
ZFV_HD double ipow(double x, double y) {
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  return pow(x, y);
}

// the equilibrium at a point of potential ph
template <int PH>
ZFV_HD void eq_at_phi(const Img &m, const EqC &c, double ph, double &rho, double &E) {
  if (PH != PH_LIGHT && m.eq == 1) {
    double h = c.th0 + c.phi_ref - ph;
    // PH_LIN has gamma = 2: the exponents are 1 and 2, as ipow takes them
    rho = PH == PH_LIN ? c.a * h : ipow(c.a * h, c.b);
    double p = c.th1 * (PH == PH_LIN ? rho * rho : ipow(rho, m.gamma));
    E = p / (m.gamma - 1.0);
  } else {
    rho = c.th0 * exp((ph - c.phi_ref) * c.a);
    E = rho * c.b;
  }
}

template <int DIM, int PH>
ZFV_HD void eq_point(const Img &m, const EqC &c, const double *x, double &rho, double &E) {
  eq_at_phi<PH>(m, c, phi<DIM, PH>(m, x), rho, E);
}

// The isentropic equilibrium of an ideal gas with gamma = 2 is linear in the
// potential (rho = (h0 - phi) / (2K), E = K rho^2), so its average over a cell
// needs only the cell's quadrature mean of phi and of its squared deviation:
//   rho_avg = a (A - phim),  E_avg = th1 a^2 ((A - phim)^2 + phiv)
// with a = 1/(2K), A = th0 + phi_ref.  Exact in real arithmetic for the same
// quadrature rule; moments are per cell and time independent.
ZFV_HD bool linear_isentropic(const Img &m) { return m.eq == 1 && m.gamma == 2.0 && m.phim != nullptr; }

// the narrowest physics set whose kernel covers this instance
inline int physics_set(const Img &m) {
  if ((m.eq == 0 || m.eq == 2) && (m.gravity == 0 || m.gravity == 1)) return PH_LIGHT;
  if (linear_isentropic(m)) {
    bool shifted = false;  // a periodic image among the stencil members
    for (int c = 1; c < 16; ++c)
      for (int d = 0; d < 3; ++d) shifted = shifted || m.shift[c][d] != 0.0;
    if (!shifted) return PH_LIN;
  }
  return PH_FULL;
}

// the physics set of a kernel that runs instances of sets a and b
inline int physics_union(int a, int b) { return a == b ? a : PH_FULL; }

ZFV_HD void linear_isentropic_avg(const EqC &c, double phim, double phiv, double &ar, double &aE) {
  const double h = c.th0 + c.phi_ref - phim;
  ar = c.a * h;
  aE = c.th1 * (c.a * c.a) * (h * h + phiv);
}

// ---- geometry helpers ---------------------------------------------------------
// Corners are stored per cell, vertex by vertex, each vertex on 16 bytes:
// 2D (x, y), 3D (x, y, z, 0); corner_words doubles per cell
template <int DIM>
constexpr int corner_words() { return DIM == 2 ? 6 : 16; }

// vertex c of cell i
template <int DIM>
ZFV_HD void load_corner(const Img &m, int64_t i, int c, double *w) {
  const PPair *p = reinterpret_cast<const PPair *>(m.crn + i * corner_words<DIM>());
  if (DIM == 2) {
    const PPair a = p[c];
    w[0] = a.lo;
    w[1] = a.hi;
    w[2] = 0.0;
  } else {
    const PPair a = p[2 * c], b = p[2 * c + 1];
    w[0] = a.lo;
    w[1] = a.hi;
    w[2] = b.lo;
  }
}

// The centre (2D x y, 3D x y z 0) and the moments (indices 3 .. DOFC-1,
// padded to an even count) are stored per cell as well, in 16-byte pieces
template <int DIM>
constexpr int centre_words() { return DIM == 2 ? 2 : 4; }
template <int DOFC>
constexpr int moment_words() { return ((DOFC > 3 ? DOFC - 3 : 0) + 1) & ~1; }

template <int DIM>
ZFV_HD void load_centre(const Img &m, int64_t i, double *xc) {
  const PPair *p = reinterpret_cast<const PPair *>(m.xc + i * centre_words<DIM>());
  const PPair a = p[0];
  xc[0] = a.lo;
  xc[1] = a.hi;
  xc[2] = DIM == 3 ? p[1].lo : 0.0;
}

// cm[alpha] for alpha < DOFC: zero below 3, the cell's moments from 3 on
template <int DOFC>
ZFV_HD void load_moments(const Img &m, int64_t i, double *cm) {
  constexpr int MW = moment_words<DOFC>();
  for (int al = 0; al < DOFC && al < 3; ++al) cm[al] = 0.0;
  if (MW > 0) {
    const PPair *p = reinterpret_cast<const PPair *>(m.mom + i * MW);
    for (int k = 0; k < MW / 2; ++k) {
      const PPair v = p[k];
      if (3 + 2 * k < DOFC) cm[3 + 2 * k] = v.lo;
      if (4 + 2 * k < DOFC) cm[4 + 2 * k] = v.hi;
    }
  }
}

template <int DIM>
ZFV_HD void load_corners(const Img &m, int64_t i, double (*v)[3]) {
  for (int k = 0; k < DIM + 1; ++k) load_corner<DIM>(m, i, k, v[k]);
}

// vertices of local face f in the order of the face's left cell
template <int DIM>
ZFV_HD void face_vertices(uint32_t fperm, int f, const double (*v)[3], double (*w)[3]) {
  for (int r = 0; r < DIM; ++r) {
    int c = (fperm >> (2 * (f * 3 + r))) & 3u;
    w[r][0] = v[c][0];
    w[r][1] = v[c][1];
    w[r][2] = v[c][2];
  }
}

// the same vertices read from the cell's corner arrays (cache hits) instead
// of selected from registers
template <int DIM>
ZFV_HD void load_face_vertices(const Img &m, int64_t i, uint32_t fperm, int f, double (*w)[3]) {
  for (int r = 0; r < DIM; ++r) load_corner<DIM>(m, i, int((fperm >> (2 * (f * 3 + r))) & 3u), w[r]);
}

// Face frame (src/zisa/math/face_factory.cpp:11-34): n, t1, t2 and the face
// measure (edge length / Heron area).
template <int DIM>
ZFV_HD void face_frame(const double (*w)[3], double *n, double *t1, double *t2, double &area) {
  if (DIM == 2) {
    double dx = w[1][0] - w[0][0], dy = w[1][1] - w[0][1];
    double l = sqrt(dx * dx + dy * dy);
    area = l;
    t1[0] = dx / l;
    t1[1] = dy / l;
    t1[2] = 0.0 / l;
    n[0] = t1[1];
    n[1] = -t1[0];
    n[2] = t1[2];
  } else {
    double a[3] = {w[1][0] - w[0][0], w[1][1] - w[0][1], w[1][2] - w[0][2]};
    double b[3] = {w[2][0] - w[0][0], w[2][1] - w[0][1], w[2][2] - w[0][2]};
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    n[0] = c[0] / lc;
    n[1] = c[1] / lc;
    n[2] = c[2] / lc;
    double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    t1[0] = a[0] / la;
    t1[1] = a[1] / la;
    t1[2] = a[2] / la;
    // Heron on the side lengths (src/zisa/math/triangle.cpp:11-18)
    double e0[3] = {w[1][0] - w[2][0], w[1][1] - w[2][1], w[1][2] - w[2][2]};
    double e1[3] = {w[0][0] - w[2][0], w[0][1] - w[2][1], w[0][2] - w[2][2]};
    double sa = sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
    double sb = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    double sc = la;
    double s = 0.5 * (sa + sb + sc);
    area = sqrt(s * (s - sa) * (s - sb) * (s - sc));
  }
  t2[0] = n[1] * t1[2] - n[2] * t1[1];
  t2[1] = n[2] * t1[0] - n[0] * t1[2];
  t2[2] = n[0] * t1[1] - n[1] * t1[0];
}

// Face frames are time independent: each cell's local faces keep theirs in a
// table written once at setup (store_face_frames, with face_frame itself),
// 3 (2D: t1, area) or 7 (3D: n, t1, area) doubles per face, and the interior
// faces' frames (the left cell's view) are copied into a table in face order
// for the face kernel (store_face_table); the flux BC of the cell update
// reads a cell's own.  n (2D) and t2 follow from the stored values with
// face_frame's own formulas, so every value equals face_frame's.  (The
// reconstruction, which loads the face vertices for its quadrature points
// anyway, rebuilds them: measured faster.)
template <int DIM>
constexpr int frame_words() { return DIM == 2 ? 3 : 7; }
// doubles of one cell-face slot of face-point values: Q x NV rounded up to
// even, so that every slot starts on 16 bytes
template <int DIM, int Q>
constexpr int slot_doubles() { return (Q * (DIM + 2) + 1) & ~1; }

// writes the light set's potential at cell i's corners (tab: (d + 1) n)
template <int DIM>
ZFV_HD void store_corner_phi(const Img &m, int64_t i, double *tab) {
  double v[DIM + 1][3];
  load_corners<DIM>(m, i, v);
  for (int k = 0; k < DIM + 1; ++k) tab[k * m.n + i] = phi<DIM, PH_LIGHT>(m, v[k]);
}

// writes the frames of cell i's local faces (tab has frame_words (d+1) n entries)
template <int DIM>
ZFV_HD void store_face_frames(const Img &m, int64_t i, double *tab) {
  constexpr int W = frame_words<DIM>();
  double vtx[DIM + 1][3];
  load_corners<DIM>(m, i, vtx);
  const uint32_t fperm = m.fperm[i];
  for (int f = 0; f < DIM + 1; ++f) {
    double w3[3][3], n[3], t1[3], t2[3], area;
    face_vertices<DIM>(fperm, f, vtx, w3);
    face_frame<DIM>(w3, n, t1, t2, area);
    double *p = tab + int64_t(f * W) * m.n + i;
    if (DIM == 2) {
      p[0] = t1[0];
      p[m.n] = t1[1];
      p[2 * m.n] = area;
    } else {
      for (int d = 0; d < 3; ++d) p[int64_t(d) * m.n] = n[d];
      for (int d = 0; d < 3; ++d) p[int64_t(3 + d) * m.n] = t1[d];
      p[int64_t(6) * m.n] = area;
    }
  }
}

template <int DIM>
ZFV_HD void load_face_frame(const Img &m, int64_t i, int f, double *n, double *t1, double *t2, double &area) {
  constexpr int W = frame_words<DIM>();
  const double *p = m.frm + int64_t(f * W) * m.n + i;
  if (DIM == 2) {
    t1[0] = p[0];
    t1[1] = p[m.n];
    t1[2] = 0.0;
    area = p[2 * m.n];
    n[0] = t1[1];
    n[1] = -t1[0];
    n[2] = t1[2];
  } else {
    for (int d = 0; d < 3; ++d) n[d] = p[int64_t(d) * m.n];
    for (int d = 0; d < 3; ++d) t1[d] = p[int64_t(3 + d) * m.n];
    area = p[int64_t(6) * m.n];
  }
  t2[0] = n[1] * t1[2] - n[2] * t1[1];
  t2[1] = n[2] * t1[0] - n[0] * t1[2];
  t2[2] = n[0] * t1[1] - n[1] * t1[0];
}

// Per interior face its frame in face order (the face kernel reads it
// coalesced): written once at setup from the left cell's frame table entry
template <int DIM>
ZFV_HD void store_face_table(const Img &m, int64_t e, double *tab) {
  constexpr int NF = DIM + 1;
  const int64_t s = m.fslot[2 * e];
  double n[3], t1[3], t2[3], area;
  load_face_frame<DIM>(m, s / NF, int(s % NF), n, t1, t2, area);
  if (DIM == 2) {
    tab[e] = t1[0];
    tab[m.ne + e] = t1[1];
    tab[2 * m.ne + e] = area;
  } else {
    for (int d = 0; d < 3; ++d) tab[int64_t(d) * m.ne + e] = n[d];
    for (int d = 0; d < 3; ++d) tab[int64_t(3 + d) * m.ne + e] = t1[d];
    tab[int64_t(6) * m.ne + e] = area;
  }
}

// the frame of interior face e from the face table (the same values, and the
// same derived n (2D) and t2, as load_face_frame)
template <int DIM>
ZFV_HD void face_table_frame(const Img &m, int64_t e, double *n, double *t1, double *t2, double &area) {
  if (DIM == 2) {
    t1[0] = m.ffr[e];
    t1[1] = m.ffr[m.ne + e];
    t1[2] = 0.0;
    area = m.ffr[2 * m.ne + e];
    n[0] = t1[1];
    n[1] = -t1[0];
    n[2] = t1[2];
  } else {
    for (int d = 0; d < 3; ++d) n[d] = m.ffr[int64_t(d) * m.ne + e];
    for (int d = 0; d < 3; ++d) t1[d] = m.ffr[int64_t(3 + d) * m.ne + e];
    area = m.ffr[int64_t(6) * m.ne + e];
  }
  t2[0] = n[1] * t1[2] - n[2] * t1[1];
  t2[1] = n[2] * t1[0] - n[0] * t1[2];
  t2[2] = n[0] * t1[1] - n[1] * t1[0];
}

// quadrature point q of a face given its vertices in left-cell order
template <int DIM>
ZFV_HD void face_point(const Img &m, int q, const double (*w)[3], double *x) {
  if (DIM == 2) {
    double xi = m.fx[q][0];
    double a = 0.5 - 0.5 * xi, b = 0.5 + 0.5 * xi;
    x[0] = fma(b, w[1][0], a * w[0][0]);
    x[1] = fma(b, w[1][1], a * w[0][1]);
    x[2] = 0.0;
  } else {
    for (int d = 0; d < 3; ++d) x[d] = (m.fx[q][0] * w[0][d] + m.fx[q][1] * w[1][d]) + m.fx[q][2] * w[2][d];
  }
}

template <int DIM>
ZFV_HD void volume_point(const Img &m, int q, const double (*v)[3], double *x) {
  for (int d = 0; d < DIM; ++d) {
    double s = m.vx[q][0] * v[0][d];
    for (int k = 1; k < DIM + 1; ++k) s = fma(m.vx[q][k], v[k][d], s);
    x[d] = s;
  }
  if (DIM == 2) x[2] = 0.0;
}

// Quadrature mean of the potential over cell i and mean squared deviation
// from it (for linear_isentropic_avg)
template <int DIM>
ZFV_HD void phi_moments(const Img &m, int64_t i, double &mean, double &var) {
  double vtx[DIM + 1][3];
  load_corners<DIM>(m, i, vtx);
  double s = 0.0;
  for (int q = 0; q < m.qv; ++q) {
    double x[3];
    volume_point<DIM>(m, q, vtx, x);
    const double p = phi<DIM, PH_FULL>(m, x);
    s = (q == 0) ? m.vw[q] * p : s + m.vw[q] * p;
  }
  double v = 0.0;
  for (int q = 0; q < m.qv; ++q) {
    double x[3];
    volume_point<DIM>(m, q, vtx, x);
    const double d = phi<DIM, PH_FULL>(m, x) - s;
    v = (q == 0) ? m.vw[q] * (d * d) : v + m.vw[q] * (d * d);
  }
  mean = s;
  var = v;
}

// Potential tables.  Radial and polytrope gravity have no affine closed form,
// and the potential and its gradient at a cell's own quadrature points do not
// change in time, so instances of the full physics sets with gravity keep
// them per cell (SoA, n cells), computed once with the functions below:
//   pot[q n + i]                       phi at volume point q      (q < qv)
//   pot[(qv + q DIM + d) n + i]        d-th component of grad phi there
//   pot[(qv (1 + DIM) + f Q + q) n + i] phi at point q of local face f
// The reconstruction then evaluates no transcendental function at its own
// points; the values are those the functions give at the same points.
template <int DIM>
ZFV_HD int64_t pot_index_grad(const Img &m, int q, int d) { return int64_t(m.qv + q * DIM + d) * m.n; }
template <int DIM>
ZFV_HD int64_t pot_index_face(const Img &m, int nqf, int f, int q) { return int64_t(m.qv * (1 + DIM) + f * nqf + q) * m.n; }
// rows of the tables for nqf face points
template <int DIM>
ZFV_HD int pot_rows(int qv, int nqf) { return qv * (1 + DIM) + (DIM + 1) * nqf; }

// fills the potential tables of cell i (tab has pot_rows * n entries)
template <int DIM>
ZFV_HD void potential_points(const Img &m, int nqf, int64_t i, double *tab) {
  double vtx[DIM + 1][3];
  load_corners<DIM>(m, i, vtx);
  for (int q = 0; q < m.qv; ++q) {
    double x[3], gp[3];
    volume_point<DIM>(m, q, vtx, x);
    tab[int64_t(q) * m.n + i] = phi<DIM, PH_FULL>(m, x);
    grad_phi<DIM, PH_FULL>(m, x, gp);
    for (int d = 0; d < DIM; ++d) tab[pot_index_grad<DIM>(m, q, d) + i] = gp[d];
  }
  const uint32_t fperm = m.fperm[i];
  for (int f = 0; f < DIM + 1; ++f) {
    double w3[3][3];
    face_vertices<DIM>(fperm, f, vtx, w3);
    for (int q = 0; q < nqf; ++q) {
      double x[3];
      face_point<DIM>(m, q, w3, x);
      tab[pot_index_face<DIM>(m, nqf, f, q) + i] = phi<DIM, PH_FULL>(m, x);
    }
  }
}

// ---- polynomial evaluation (include/zisa/math/poly2d_impl.hpp:271-323) -------
// a[alpha][v] for alpha < DOF, c = moments (0 for alpha < 3)
template <int DIM, int DEG, int NV>
ZFV_HD void poly_eval(const double (*a)[NV], const double *c, const double *X, double *px) {
  for (int v = 0; v < NV; ++v) px[v] = 0.0;
  double pxk = 1.0;
#pragma unroll
  for (int kx = 0; kx <= DEG; ++kx) {
    double pyk = 1.0;
#pragma unroll
    for (int ky = 0; ky <= DEG - kx; ++ky) {
      if (DIM == 2) {
        const int n = kx + ky;
        const int i = (n + 1) * n / 2 + ky;
        const double b = fma(pxk, pyk, -c[i]);
#pragma unroll
        for (int v = 0; v < NV; ++v) px[v] = fma(a[i][v], b, px[v]);
      } else {
        double pzk = 1.0;
#pragma unroll
        for (int kz = 0; kz <= DEG - kx - ky; ++kz) {
          const int n3 = kx + ky + kz - 1;
          const int d3 = n3 < 0 ? 0 : (n3 + 1) * (n3 + 2) * (n3 + 3) / 6;
          const int n2 = ky + kz;
          const int i = d3 + (n2 + 1) * n2 / 2 + kz;
          const double b = fma(pxk * pyk, pzk, -c[i]);
#pragma unroll
          for (int v = 0; v < NV; ++v) px[v] = fma(a[i][v], b, px[v]);
          pzk *= X[2];
        }
      }
      pyk *= X[1];
    }
    pxk *= X[0];
  }
}

// the basis of poly_eval at X, in its order: b[i] = X^alpha_i - c[i]
template <int DIM, int DEG>
ZFV_HD void poly_basis(const double *c, const double *X, double *b) {
  double pxk = 1.0;
#pragma unroll
  for (int kx = 0; kx <= DEG; ++kx) {
    double pyk = 1.0;
#pragma unroll
    for (int ky = 0; ky <= DEG - kx; ++ky) {
      if (DIM == 2) {
        const int n = kx + ky;
        b[(n + 1) * n / 2 + ky] = fma(pxk, pyk, -c[(n + 1) * n / 2 + ky]);
      } else {
        double pzk = 1.0;
#pragma unroll
        for (int kz = 0; kz <= DEG - kx - ky; ++kz) {
          const int n3 = kx + ky + kz - 1;
          const int d3 = n3 < 0 ? 0 : (n3 + 1) * (n3 + 2) * (n3 + 3) / 6;
          const int n2 = ky + kz;
          const int i = d3 + (n2 + 1) * n2 / 2 + kz;
          b[i] = fma(pxk * pyk, pzk, -c[i]);
          pzk *= X[2];
        }
      }
      pyk *= X[1];
    }
    pxk *= X[0];
  }
}

template <int DIM>
constexpr int dof_of(int deg) {
  return DIM == 2 ? (deg + 1) * (deg + 2) / 2 : (deg + 1) * (deg + 2) * (deg + 3) / 6;
}

template <int DIM>
constexpr int deg_of(int dof) {
  int d = 0;
  while (dof_of<DIM>(d + 1) <= dof) ++d;
  return d;
}

// The gravity source moments of cell i: M[alpha][d] = vol sum_q vw_q
// b_alpha(x_q) d_d phi(x_q) with the volume rule, the polynomial basis of the
// cell's reconstruction (poly_basis) and the potential's gradient at the
// points.  The volume terms of the source, sum_q w_q delta(x_q) d_d phi(x_q)
// for delta = scale sum_alpha a_alpha b_alpha, are then scale sum_alpha
// a_alpha M[alpha][d]: the same quadrature, summed over the points once at
// setup instead of per call.
template <int DIM, int DOFC>
ZFV_HD void store_source_moments(const Img &m, int64_t i, double *tab) {
  constexpr int DEG = deg_of<DIM>(DOFC);
  const int64_t n = m.n;
  double vtx[DIM + 1][3];
  load_corners<DIM>(m, i, vtx);
  double cm[DOFC];
  load_moments<DOFC>(m, i, cm);
  double xc[3];
  load_centre<DIM>(m, i, xc);
  const double inv_len = 1.0 / m.len[i];
  const double vol_i = m.vol[i];
  double acc[DOFC][DIM];
  for (int q = 0; q < m.qv; ++q) {
    double x[3], gp[3], b[DOFC];
    volume_point<DIM>(m, q, vtx, x);
    grad_phi<DIM, PH_FULL>(m, x, gp);
    const double Xs[3] = {(x[0] - xc[0]) * inv_len, (x[1] - xc[1]) * inv_len,
                          DIM == 3 ? (x[2] - xc[2]) * inv_len : 0.0};
    poly_basis<DIM, DEG>(cm, Xs, b);
    const double w = vol_i * m.vw[q];
    for (int al = 0; al < DOFC; ++al)
      for (int d = 0; d < DIM; ++d) acc[al][d] = (q == 0) ? w * (b[al] * gp[d]) : fma(w, b[al] * gp[d], acc[al][d]);
  }
  for (int al = 0; al < DOFC; ++al)
    for (int d = 0; d < DIM; ++d) tab[int64_t(i) * (DOFC * DIM) + al * DIM + d] = acc[al][d];
}

// ---- ideal gas / fluxes -------------------------------------------------------
// Rotated 2D states carry 4 variables (rho, m_n, m_t, E): the t2 component of
// a 2D state is identically zero and stays zero through HLLC.
template <int NV>
ZFV_HD double pressure_of(const double *u, double gamma) {
  double m2 = u[1] * u[1] + u[2] * u[2];
  if (NV == 5) m2 += u[3] * u[3];
  double kin = 0.5 * m2 / u[0];
  return (u[NV - 1] - kin) * (gamma - 1.0);
}

template <int NV>
ZFV_HD void euler_flux(const double *u, double p, double *f) {
  double v = u[1] / u[0];
  f[0] = u[1];
  f[1] = v * u[1] + p;
  f[2] = v * u[2];
  if (NV == 5) f[3] = v * u[3];
  f[NV - 1] = v * (u[NV - 1] + p);
}

template <int NV>
ZFV_HD void hllc(const double *uL, const double *uR, double gamma, double *nf) {
  double pL = pressure_of<NV>(uL, gamma), aL = sqrt(gamma * pL / uL[0]);
  double pR = pressure_of<NV>(uR, gamma), aR = sqrt(gamma * pR / uR[0]);
  double roe_ratio = sqrt(uR[0] / uL[0]);
  double inv1 = 1.0 + roe_ratio;
  double vL = uL[1] / uL[0];
  double vR = uR[1] / uR[0];
  double v_tilda = (vL + vR * roe_ratio) / inv1;
  double HL = (uL[NV - 1] + pL) / uL[0];
  double HR = (uR[NV - 1] + pR) / uR[0];
  double H_tilda = (HL + HR * roe_ratio) / inv1;
  double r1 = (uL[1] / uL[0] + (uR[1] / uR[0]) * roe_ratio) / inv1;
  double r2 = (uL[2] / uL[0] + (uR[2] / uR[0]) * roe_ratio) / inv1;
  double vroe = r1 * r1 + r2 * r2;
  if (NV == 5) {
    double r3 = (uL[3] / uL[0] + (uR[3] / uR[0]) * roe_ratio) / inv1;
    vroe += r3 * r3;
  }
  double a_tilda = sqrt((gamma - 1.0) * (H_tilda - 0.5 * vroe));
  double sL = dmin(vL - aL, v_tilda - a_tilda);
  double sR = dmax(vR + aR, v_tilda + a_tilda);
  double s_star = (uR[1] * (sR - vR) - uL[1] * (sL - vL) + pL - pR) / (uR[0] * (sR - vR) - uL[0] * (sL - vL));
  const bool left = 0.0 <= s_star;
  const double *uK = left ? uL : uR;
  double pK = left ? pL : pR;
  euler_flux<NV>(uK, pK, nf);
  if (sL < 0.0 && 0.0 <= sR) {
    double sK = left ? sL : sR;
    double vK = left ? vL : vR;
    double cK = (sK - vK) / (sK - s_star);
    nf[0] += sK * (cK * uK[0] - uK[0]);
    nf[1] += sK * (cK * uK[0] * s_star - uK[1]);
    nf[2] += sK * (cK * uK[2] - uK[2]);
    if (NV == 5) nf[3] += sK * (cK * uK[3] - uK[3]);
    nf[NV - 1] += sK * (cK * (uK[NV - 1] + (s_star - vK) * (uK[0] * s_star + pK / (sK - vK))) - uK[NV - 1]);
  }
}

template <int NV>
ZFV_HD void rusanov(const double *uL, const double *uR, double gamma, double *f) {
  double pL = pressure_of<NV>(uL, gamma), aL = sqrt(gamma * pL / uL[0]);
  double pR = pressure_of<NV>(uR, gamma), aR = sqrt(gamma * pR / uR[0]);
  double fL[NV], fR[NV];
  euler_flux<NV>(uL, pL, fL);
  euler_flux<NV>(uR, pR, fR);
  double s = dmax(fabs(uL[1] / uL[0]) + aL, fabs(uR[1] / uR[0]) + aR);
  for (int v = 0; v < NV; ++v) f[v] = 0.5 * (fL[v] + fR[v]) - 0.5 * s * (uR[v] - uL[v]);
}

// ---- the per-cell reconstruction, one lane per conserved variable --------------
//
// Reads the cell's stencil, forms every stencil polynomial with the
// precomputed pseudo-inverses, applies the CWENO correction and the WENO
// weights, then evaluates the result at the quadrature points of the cell's
// interior faces (rotated into each face's frame) and, with gravity, computes
// the cell's source term.
//
// A group of NV lanes shares one cell: lane g owns conserved variable g (rho,
// momentum components, energy) and keeps only that variable's polynomial
// coefficients, so a lane needs a fraction of the registers of a whole cell.
// The lanes exchange values through X::bcast(v, k) / X::bcast_i(v, k) (value
// v of lane k of the same group) and a per-group LDS slice X::lds() of
// NV x (DOFC - 1) doubles ordered by X::sync():
//   - the equilibrium averages of the stencil members are computed by lane
//     (row mod NV) and broadcast to the density and energy lanes,
//   - the smoothness indicator of a stencil is the maximum over the variables,
//   - the face-frame rotation and the gravity source mix momentum components.
// Every call of bcast sits in control flow that is uniform over the group.
//
// U: state, nvars x n SoA.  fpv: per cell-face slot (i * NF + f) its Q x NV
// point values, written for interior faces (one group writes its cell's whole
// block, so cache lines fill within one XCD).  src: nvars x n or null.
// LDS of one wavefront: first a row region of wave_lds_doubles holding each
// group's NV pseudo-inverse rows of NC entries (group c at c NV NC: a stride
// of NV NC doubles keeps the groups' reads on different banks), then per
// cell group its central rows' member equilibrium averages (er, eE per row,
// for the one-sided rows that share their member; group_lds_doubles).
template <int DIM, int DOFC>
constexpr int wave_lds_doubles() { return (64 / (DIM + 2)) * (DIM + 2) * (DOFC - 1); }
template <int RC>
constexpr int group_lds_doubles() { return 2 * RC; }
template <int DIM, int DOFC, int RC>
constexpr int wave_block_doubles() { return wave_lds_doubles<DIM, DOFC>() + (64 / (DIM + 2)) * group_lds_doubles<RC>(); }
// the index of central coefficient a of row r of cell i in the pseudo-inverse
// array (pairs a, a + 1 of one row and cell adjacent, one 16-byte load; an
// odd last coefficient in its own block); the one-sided rows follow at nc rc n
// the index of coefficient a of row rr of one-sided stencil kb (nbc
// coefficients, rb rows each): after the central block, pairs a, a + 1 of a
// row adjacent as for the central rows, an odd last coefficient in its own
// block
ZFV_HD int64_t p_onesided(int nc, int rc, int nbc, int nb, int rb, int64_t n, int kb, int a, int rr, int64_t i) {
  const int64_t base = int64_t(nc) * rc * n;
  const int npair = nbc / 2;
  const int64_t row = int64_t(kb) * rb + rr;
  if (a < 2 * npair) return base + ((row * npair + (a >> 1)) * n + i) * 2 + (a & 1);
  return base + int64_t(2 * npair) * nb * rb * n + row * n + i;
}

ZFV_HD int64_t p_central(int nc, int rc, int64_t n, int a, int r, int64_t i) {
  const int ncp = nc / 2;
  if (a < 2 * ncp) return ((int64_t(a >> 1) * rc + r) * n + i) * 2 + (a & 1);
  return int64_t(2 * ncp) * rc * n + int64_t(r) * n + i;
}

template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH, class X>
ZFV_HD void reconstruct_lane(const Img &m, int64_t i, const int g, X &xch, const double *__restrict__ U,
                             double *__restrict__ fpv, double *__restrict__ src) {
  constexpr int NV = DIM + 2;
  constexpr int NF = DIM + 1;
  constexpr int NC = DOFC - 1;
  constexpr int NBC = DIM;
  constexpr int DEGC = deg_of<DIM>(DOFC);
  const int64_t n = m.n;
  const double gamma = m.gamma;

  // own state and scale (include/zisa/model/characteristic_scale.hpp:23-32)
  double u0[NV];
  for (int v = 0; v < NV; ++v) u0[v] = U[v * n + i];
  double msq = u0[1] * u0[1] + u0[2] * u0[2];
  if (DIM == 3) msq += u0[3] * u0[3];
  const double rho0 = u0[0];
  const double Eint0 = u0[NV - 1] - 0.5 * msq / rho0;
  // this call's scale (UnityScaling: all ones,
  // include/zisa/model/characteristic_scale.hpp:39-46); a stateful instance
  // may keep the last recompute's
  double sc_rho = rho0, sc_E = Eint0, sc_a = sqrt(gamma * (Eint0 * (gamma - 1.0)) / rho0);
  if (m.scaling == 1) sc_rho = sc_E = sc_a = 1.0;
  const bool is_rho = g == 0, is_E = g == NV - 1;
  // The first two chunks' stencil row entries (and in 3D the first chunk's
  // member states) are loaded before the equilibrium solve, whose arithmetic
  // then covers their latency (every row entry names a cell, first-order
  // cells included)
  constexpr int R_ALL = RC + NB * RB;
  const int32_t rw_pre0 = g < R_ALL ? m.rows[g * n + i] : int32_t(0);
  const int32_t rw_pre1 = NV + g < R_ALL ? m.rows[(NV + g) * n + i] : int32_t(0);
  double uk_pre[NV];
  if (DIM == 3)
    for (int k = 0; k < NV; ++k) uk_pre[k] = U[g * n + (xch.bcast_i(rw_pre0, k) & ROW_CELL_MASK)];

  double vtx[DIM + 1][3];
  load_corners<DIM>(m, i, vtx);

  // The potential at the cell's own quadrature points: from the instance's
  // tables in the full sets (always present in PH_LIN), else evaluated.
  const bool tab = PH == PH_LIN || (PH == PH_FULL && m.pot != nullptr);
  auto vol_phi = [&](int q, const double *x) { return tab ? m.pot[int64_t(q) * n + i] : phi<DIM, PH>(m, x); };

  // Average of an equilibrium over this cell with the quadrature points split
  // over the group's lanes: lane k sums the points q = k (mod NV) in order,
  // then every lane adds the NV partial sums in lane order, so all lanes hold
  // the same value.
  auto group_avg = [&](const EqC &tc, double &ar, double &aE) {
    if (PH == PH_LIN || (PH == PH_FULL && linear_isentropic(m))) {
      linear_isentropic_avg(tc, m.phim[2 * i], m.phim[2 * i + 1], ar, aE);
      return;
    }
    const double vol_i = m.vol[i];
    double pr = 0.0, pE = 0.0;
    for (int q = g; q < m.qv; q += NV) {
      double x[3];
      volume_point<DIM>(m, q, vtx, x);
      double r, E;
      eq_at_phi<PH>(m, tc, vol_phi(q, x), r, E);
      const double w = vol_i * m.vw[q];
      pr = pr + w * r;
      pE = pE + w * E;
    }
    double sr = xch.bcast(pr, 0), sE = xch.bcast(pE, 0);
    for (int k = 1; k < NV; ++k) {
      sr = sr + xch.bcast(pr, k);
      sE = sE + xch.bcast(pE, k);
    }
    ar = (1.0 * sr) / vol_i;
    aE = (1.0 * sE) / vol_i;
  };

  // local equilibrium of this cell: every lane solves it (identical results)
  double th[2] = {0.0, 0.0};
  double phi_ref = 0.0;
  double ar1 = -1.0;  // isothermal: the average of exp(a (phi - phi_ref)) over the cell, once solved
  const bool wb = m.eq != 0;
  bool eq_found = false;
  if (wb) {
    double xref[3];
    volume_point<DIM>(m, 0, vtx, xref);
    phi_ref = vol_phi(0, xref);
  }
  // Stateful equilibria (LocalReconstruction::recompute_equilibrium,
  // include/zisa/reconstruction/local_reconstruction.hpp:87-100): between
  // recomputes (every steps_per_recompute calls) the cell keeps the theta and
  // the scale of its last recompute, unless its (rho, E_int) has moved from
  // the kept equilibrium's own cell average by the threshold or more,
  // measured in the kept scale; a kept equilibrium that was not found
  // averages to zero there.
  int32_t calls = 0;
  bool reuse = false;
  if (m.eqc != nullptr) {
    calls = m.eqcnt[i];
    if (calls % m.spr != 0) {
      const bool kept_found = m.eqc[2 * n + i] != 0.0;
      const double kth[2] = {m.eqc[i], m.eqc[n + i]};
      const double c_rho = m.eqc[3 * n + i], c_a = m.eqc[4 * n + i], c_E = m.eqc[5 * n + i];
      const double er = m.eqc[6 * n + i], eE = m.eqc[7 * n + i];
      const double d0 = (rho0 - er) / c_rho, d1 = (Eint0 - eE) / c_E;
      reuse = !(sqrt(d0 * d0 + d1 * d1) >= m.eq_threshold);
      if (reuse) {
        th[0] = kth[0];
        th[1] = kth[1];
        eq_found = wb && kept_found;
        sc_rho = c_rho;
        sc_a = c_a;
        sc_E = c_E;
      }
    }
  }
  if (wb && !reuse) {
    double vol_i = m.vol[i];
    // cell average of the equilibrium with parameters t over this cell
    auto cell_avg = [&](const double *t, double &ar, double &aE) {
      group_avg(eq_consts<PH>(m, t, phi_ref), ar, aE);
    };
    if (PH == PH_LIGHT || m.eq == 2) {
      double RT = (gamma - 1.0) * Eint0 / rho0;
      double t1[2] = {1.0, RT};
      double ar, aE;
      cell_avg(t1, ar, aE);
      ar1 = ar;
      th[0] = rho0 / ar;
      th[1] = RT;
      eq_found = true;
    } else {
      // quasi-Newton with a central finite-difference Jacobian
      // (include/zisa/math/quasi_newton.hpp:12-50)
      double p = Eint0 * (gamma - 1.0);
      double guess[2] = {gamma / (gamma - 1.0) * p / rho0, p / (PH == PH_LIN ? rho0 * rho0 : ipow(rho0, gamma))};
      double atol[2] = {1e-13 * guess[0], 1e-13 * guess[1]};
      double x0 = guess[0], x1 = guess[1];
      auto f = [&](double a, double b, double &f0, double &f1) {
        double t[2] = {a, b};
        double ar, aE;
        cell_avg(t, ar, aE);
        f0 = rho0 - ar;
        f1 = Eint0 - aE;
      };
      double fx0, fx1;
      f(x0, x1, fx0, fx1);
      double dx0 = 2.0 * atol[0] + 1.0, dx1 = 2.0 * atol[1] + 1.0;
      double h00 = 0, h01 = 0, h10 = 0, h11 = 0;  // previous two steps
      bool ok = true;
      int iter = 0;
      while (!(fabs(dx0) <= atol[0] && fabs(dx1) <= atol[1]) && iter < 20) {
        double J00, J01, J10, J11;
        {
          double e = 1e-6 * fabs(x0);
          double fp0, fp1, fm0, fm1;
          f(x0 + 0.5 * e, x1, fp0, fp1);
          f(x0 - 0.5 * e, x1, fm0, fm1);
          J00 = (fp0 - fm0) / e;
          J01 = (fp1 - fm1) / e;
        }
        {
          double e = 1e-6 * fabs(x1);
          double fp0, fp1, fm0, fm1;
          f(x0, x1 + 0.5 * e, fp0, fp1);
          f(x0, x1 - 0.5 * e, fm0, fm1);
          J10 = (fp0 - fm0) / e;
          J11 = (fp1 - fm1) / e;
        }
        double inv_det = 1.0 / (J00 * J11 - J01 * J10);
        dx0 = inv_det * (J11 * fx0 - J10 * fx1);
        dx1 = inv_det * (-J01 * fx0 + J00 * fx1);
        x0 -= dx0;
        x1 -= dx1;
        f(x0, x1, fx0, fx1);
        if (iter >= 4 && !(dx0 <= atol[0] && dx1 <= atol[1])) {
          double rate0 = log(fabs(dx0) / fabs(h10)) / log(fabs(h10) / fabs(h00));
          double rate1 = log(fabs(dx1) / fabs(h11)) / log(fabs(h11) / fabs(h01));
          if (!(rate0 >= 0.0) || !(rate1 >= 0.0)) ok = false;
        }
        h00 = h10;
        h01 = h11;
        h10 = dx0;
        h11 = dx1;
        if (!ok) break;
        ++iter;
      }
      if (ok && !(iter == 20 && !(fabs(dx0) <= 1000.0 * atol[0] && fabs(dx1) <= 1000.0 * atol[1]))) {
        th[0] = x0;
        th[1] = x1;
        eq_found = true;
      }
    }
  }
  if (m.eqc != nullptr) {
    xch.sync();  // every lane of the group has read the kept state
  }
  if (m.eqc != nullptr && g == 0) {
    if (!reuse) {
      m.eqc[i] = th[0];
      m.eqc[n + i] = th[1];
      m.eqc[2 * n + i] = eq_found ? 1.0 : 0.0;
      m.eqc[3 * n + i] = sc_rho;
      m.eqc[4 * n + i] = sc_a;
      m.eqc[5 * n + i] = sc_E;
    }
    m.eqcnt[i] = reuse ? calls + 1 : 1;
  }
  const double scale_g = g == 0 ? sc_rho : (g == NV - 1 ? sc_E : sc_a);
  const double inv_g = 1.0 / scale_g;

  const EqC thc = eq_consts<PH>(m, th, phi_ref);

  // equilibrium average of this cell's equilibrium over a member cell given
  // its (translated) corners and volume
  auto member_avg = [&](const double (*w3)[3], double vol_j, double &er, double &eE) {
    if (PH == PH_LIGHT) {
      // axial gravity: phi is affine, so its point values follow from the
      // corner potentials load_member left in w3[k][0]; rho_eq = th0
      // exp(a (phi - phi_ref)) and E_eq = b rho_eq (b constant), so both
      // averages come from one weighted sum
      double pv[DIM + 1];
      for (int k = 0; k < DIM + 1; ++k) pv[k] = w3[k][0];
      // two points per iteration: independent exponentials interleave; the
      // sum still runs in point order.  An odd last point pairs with itself
      // at weight zero (adding +0 leaves the sum unchanged), so the loop body
      // is the only copy of the exponential.
      auto point = [&](int q) {
        double ph = m.vx[q][0] * pv[0];
        for (int k = 1; k < DIM + 1; ++k) ph = fma(m.vx[q][k], pv[k], ph);
        return exp((ph - thc.phi_ref) * thc.a);
      };
      double se = 0.0;
      for (int q = 0; q < m.qv; q += 2) {
        const bool pair = q + 1 < m.qv;
        const double e0 = point(q), e1 = point(pair ? q + 1 : q);
        const double w1 = pair ? m.vw[q + 1] : 0.0;
        se = (q == 0) ? m.vw[q] * e0 : fma(m.vw[q], e0, se);
        se = fma(w1, e1, se);
      }
      er = thc.th0 * se;
      eE = er * thc.b;
      return;
    }
    double sr = 0.0, sE = 0.0;
    for (int q = 0; q < m.qv; ++q) {
      double x[3];
      volume_point<DIM>(m, q, w3, x);
      double r, E;
      eq_point<DIM, PH>(m, thc, x, r, E);
      double w = vol_j * m.vw[q];
      if (q == 0) {
        sr = w * r;
        sE = w * E;
      } else {
        sr = sr + w * r;
        sE = sE + w * E;
      }
    }
    er = (1.0 * sr) / vol_j;
    eE = (1.0 * sE) / vol_j;
  };
  const bool eq_sub = wb && eq_found;
  // this cell's equilibrium averaged over the cell itself, the points split
  // over the group's lanes
  auto own_eq = [&](double &er, double &eE) { group_avg(thc, er, eE); };
  double q0;
  {
    double er = 0.0, eE = 0.0;
    if (eq_sub) {
      // isothermal: the equilibrium is th0 times the function the solve just
      // averaged over this cell (same points, weights and order), so its own
      // average is th0 ar1 (to rounding), with E = b rho pointwise
      if (ar1 > 0.0) {
        er = th[0] * ar1;
        eE = er * thc.b;
      } else {
        own_eq(er, eE);
      }
    }
    if (m.eqc != nullptr && !reuse && g == 0) {
      // the recomputed equilibrium's own average (zero if none), which the
      // next calls measure the threshold against
      m.eqc[6 * n + i] = er;
      m.eqc[7 * n + i] = eE;
    }
    q0 = u0[g];
    if (is_rho) q0 -= er;
    if (is_E) q0 -= eE;
    q0 = q0 * inv_g;
  }

  const int32_t meta = m.meta[i];
  const bool first_order = (meta & 8) != 0;
  const int kh = meta & 7;

  // this variable's coefficients: central c[alpha] (alpha = 1..DOFC-1),
  // one-sided b[k][alpha] (alpha = 1..DIM)
  double c[NC];
  // (b is zeroed after the central rows, where it starts to be used, so it
  // holds no registers during them)
  double b[NB > 0 ? NB : 1][NBC];
  for (int a = 0; a < NC; ++a) c[a] = 0.0;

  if (!first_order) {
    // Rows in chunks of NV: lane k of the group owns row r0 + k of a chunk.
    // It computes that row's member equilibrium average (member data loaded
    // one chunk ahead, row entries two chunks ahead, so memory latency
    // overlaps arithmetic), then loads the row of the pseudo-inverse, which
    // reaches every lane through the group's LDS slice; the averages reach
    // the lanes by wavefront shuffles.
    double *pl = xch.wave_lds() + xch.lane0() * NC;  // this group's NV rows x NC coefficients
    double *ceq = xch.lds();                         // central rows' member averages (er, eE)
    constexpr int NCP = NC / 2;                      // coefficient pairs of a central row
    // scaled deviation of row k of the chunk in this lane's variable
    // (the equilibrium parts enter with factors 1 or 0: u - 1 er = u - er
    // and x - 0 eE = x exactly, so this equals the branchy form bit for bit)
    const double f_rho = (wb && is_rho) ? 1.0 : 0.0, f_E = (wb && is_E) ? 1.0 : 0.0;
    auto deviation = [&](double u, double er, double eE) {
      const double d = fma(-f_E, eE, fma(-f_rho, er, u));
      return fma(d, inv_g, -q0);
    };
    // a member's equilibrium averages as lane k computed them; in the light
    // set the energy average is the density average times the cell constant
    // b (member_avg), formed here from the broadcast density average
    auto member_eq = [&](double my_er, double my_eE, int k, double &er, double &eE) {
      er = xch.bcast(my_er, k);
      if (PH == PH_LIGHT) eE = eq_sub ? er * thc.b : 0.0;
      else eE = xch.bcast(my_eE, k);
    };

    // what the equilibrium average over a member (row entry rw) needs: its potential
    // moments (linear isentropic equilibrium, untranslated member) or its
    // translated corners and volume
    struct Member {
      double w3[DIM + 1][3];
      double vol;
      double pm, pv;
      double sphi;  // light set: the translation's potential, added at use
      int lin;
    };
    auto load_member = [&](int32_t rw, Member &mb) {
      const int64_t j = rw & ROW_CELL_MASK;
      const int code = rw >> ROW_CELL_BITS;
      if (PH == PH_LIGHT) {
        // the member's corner potentials in w3[k][0] and its translation's
        // potential, added where they are used (so that the loads stay in
        // flight until then)
        for (int k = 0; k < DIM + 1; ++k) mb.w3[k][0] = m.cphi[k * n + j];
        mb.sphi = m.shift_phi[code];
        return;
      }
      mb.lin = PH == PH_LIN || (PH == PH_FULL && linear_isentropic(m) && code == 0) ? 1 : 0;
      if (PH == PH_LIN || mb.lin) {
        const PPair pv2 = reinterpret_cast<const PPair *>(m.phim)[j];
        mb.pm = pv2.lo;
        mb.pv = pv2.hi;
        return;
      }
      load_corners<DIM>(m, j, mb.w3);
      for (int k = 0; k < DIM + 1; ++k)
        for (int d = 0; d < DIM; ++d) mb.w3[k][d] += m.shift[code][d];
      mb.vol = m.vol[j];
    };
    auto member_avg_of = [&](const Member &mb, double &er, double &eE) {
      if (PH == PH_LIN || (PH == PH_FULL && mb.lin)) {
        linear_isentropic_avg(thc, mb.pm, mb.pv, er, eE);
      } else if (PH == PH_LIGHT) {
        double w3[DIM + 1][3];
        for (int k = 0; k < DIM + 1; ++k) w3[k][0] = mb.w3[k][0] + mb.sphi;
        member_avg(w3, mb.vol, er, eE);
      } else {
        member_avg(mb.w3, mb.vol, er, eE);
      }
    };
    // All rows R = RC + NB RB (central rows, then the one-sided stencils'
    // rows kb RB + rr) in one stream of chunks: the row entries run two chunks
    // ahead and lane g's member one chunk ahead of the chunk being
    // accumulated.  Chunks of central rows only run in a first rolled loop,
    // the rest (the last central rows and the one-sided ones) in a second, so
    // the one-sided coefficients b hold registers only there.
    constexpr int R = RC + NB * RB;
    int32_t rw_cur = rw_pre0;
    int32_t rw_nxt = rw_pre1;
    Member mb_cur{};
    if (eq_sub && g < R) load_member(rw_cur, mb_cur);
    // stages chunk r0: loads the chunk's member states, advances the
    // member/row pipeline, computes this lane's member average, then loads
    // this lane's row of the pseudo-inverse (central: NC entries; one-sided:
    // NBC, the rest 0; loaded after the average so that they hold no
    // registers during it) and shares the rows through the LDS slice
    auto chunk = [&](int r0, double *uk, double &my_er, double &my_eE) {
      const int my_r = r0 + g;
      const bool mine = my_r < R;
#pragma unroll
      for (int k = 0; k < NV; ++k)
        uk[k] = (DIM == 3 && r0 == 0) ? uk_pre[k] : U[g * n + (xch.bcast_i(rw_cur, k) & ROW_CELL_MASK)];
      // a one-sided row whose member is a central row of an earlier chunk
      // takes that row's averages from LDS (the same values: same member,
      // same function) instead of loading the member and averaging again;
      // not in PH_LIN, whose member averages cost less than the lookup
      auto twin_before = [&](int r, int chunk0) {
        if (PH == PH_LIN || !eq_sub || r < RC || r >= R) return -1;
        const int t = m.twin[(r - RC) * n + i];
        return t < chunk0 ? t : -1;
      };
      const int tw = twin_before(my_r, r0);
      Member mb_nxt{};
      if (eq_sub && my_r + NV < R && twin_before(my_r + NV, r0 + NV) < 0) load_member(rw_nxt, mb_nxt);
      const int32_t rw_far = my_r + 2 * NV < R ? m.rows[(my_r + 2 * NV) * n + i] : int32_t(0);
      my_er = 0.0;
      my_eE = 0.0;
      if (eq_sub && mine) {
        if (tw >= 0) {
          my_er = ceq[2 * tw];
          my_eE = ceq[2 * tw + 1];
        } else {
          member_avg_of(mb_cur, my_er, my_eE);
        }
        if (PH != PH_LIN && my_r < RC) {
          ceq[2 * my_r] = my_er;
          ceq[2 * my_r + 1] = my_eE;
        }
      }
      double pr[NC];
      if (my_r < RC) {
        // coefficient pairs in one 16-byte load each
        const PPair *pp = reinterpret_cast<const PPair *>(m.P);
#pragma unroll
        for (int p = 0; p < NCP; ++p) {
          const PPair v = pp[p_central(NC, RC, n, 2 * p, my_r, i) >> 1];
          pr[2 * p] = v.lo;
          pr[2 * p + 1] = v.hi;
        }
        if (NC & 1) pr[NC - 1] = m.P[p_central(NC, RC, n, NC - 1, my_r, i)];
      } else {
        const int kb = (my_r - RC) / RB, rr = (my_r - RC) - kb * RB;
#pragma unroll
        for (int a = 0; a < NC; ++a) pr[a] = 0.0;
        if (mine) {
          const PPair *pp = reinterpret_cast<const PPair *>(m.P);
#pragma unroll
          for (int p = 0; p < NBC / 2; ++p) {
            const PPair v = pp[p_onesided(NC, RC, NBC, NB, RB, n, kb, 2 * p, rr, i) >> 1];
            pr[2 * p] = v.lo;
            pr[2 * p + 1] = v.hi;
          }
          if (NBC & 1) pr[NBC - 1] = m.P[p_onesided(NC, RC, NBC, NB, RB, n, kb, NBC - 1, rr, i)];
        }
      }
      xch.sync();
#pragma unroll
      for (int a = 0; a < NC; ++a) pl[g * NC + a] = pr[a];
      xch.sync();
      rw_cur = rw_nxt;
      rw_nxt = rw_far;
      mb_cur = mb_nxt;
    };
    // Rows are accumulated in a rolled loop over the chunk (one row's
    // entries in registers at a time).  Central rows past the last of a
    // chunk would add exact zeros; rows past R are skipped.
    constexpr int R0 = (RC / NV) * NV;  // central rows in whole chunks
#pragma unroll 1
    for (int r0 = 0; r0 < R0; r0 += NV) {
      double uk[NV], my_er, my_eE;
      chunk(r0, uk, my_er, my_eE);
#pragma unroll 1
      for (int k = 0; k < NV; ++k) {
        double er = 0.0, eE = 0.0;
        if (wb) member_eq(my_er, my_eE, k, er, eE);
        const double d = deviation(uk[k], er, eE);
#pragma unroll
        for (int a = 0; a < NC; ++a) c[a] = fma(pl[k * NC + a], d, c[a]);
      }
    }
    for (int k = 0; k < NB; ++k)
      for (int a = 0; a < NBC; ++a) b[k][a] = 0.0;
#pragma unroll 1
    for (int r0 = R0; r0 < R; r0 += NV) {
      double uk[NV], my_er, my_eE;
      chunk(r0, uk, my_er, my_eE);
#pragma unroll 1
      for (int k = 0; k < NV; ++k) {
        const int r = r0 + k;  // uniform over the group
        if (r >= R) break;
        double er = 0.0, eE = 0.0;
        if (wb) member_eq(my_er, my_eE, k, er, eE);
        const double d = deviation(uk[k], er, eE);
        if (r < RC) {
#pragma unroll
          for (int a = 0; a < NC; ++a) c[a] = fma(pl[k * NC + a], d, c[a]);
        } else {
          const int kb = (r - RC) / RB;
#pragma unroll
          for (int a = 0; a < NBC; ++a) {
            const double p = pl[k * NC + a];
#pragma unroll
            for (int kk = 0; kk < NB; ++kk)
              if (kk == kb) b[kk][a] = fma(p, d, b[kk][a]);
          }
        }
      }
    }
  }

  // final polynomial of this variable
  double a[DOFC];
  if (first_order) {
    a[0] = 0.0 + 1.0 * q0;
    for (int al = 1; al < DOFC; ++al) a[al] = 0.0;
  } else {
    double a0[NB + 1];
    for (int k = 0; k <= NB; ++k) a0[k] = q0;
    if (m.mode == 0) {
      // CWENO: p_hi <- (p_hi - sum_{k != hi} c_k p_k) * (1 / c_hi)
      if (kh == 0) {
#pragma unroll
        for (int k = 1; k <= NB; ++k) {
          const double w = m.lw[k];
          a0[0] = a0[0] - w * a0[k];
          for (int al = 0; al < NBC; ++al) c[al] = c[al] - w * b[k - 1][al];
        }
        const double iw = 1.0 / m.lw[0];
        a0[0] = iw * a0[0];
        for (int al = 0; al < NC; ++al) c[al] = iw * c[al];
      } else if (NB > 0) {
        // the corrected polynomial is a one-sided one (central stencil reduced
        // to order 1); compile-time indices keep b in registers
#pragma unroll
        for (int hb = 0; hb < NB; ++hb) {
          if (hb != kh - 1) continue;
#pragma unroll
          for (int k = 0; k <= NB; ++k) {
            if (k == hb + 1) continue;
            const double w = m.lw[k];
            a0[hb + 1] = a0[hb + 1] - w * a0[k];
            for (int al = 0; al < NBC; ++al) {
              double o = (k == 0) ? c[al] : b[k > 0 ? k - 1 : 0][al];
              b[hb][al] = b[hb][al] - w * o;
            }
          }
          const double iw = 1.0 / m.lw[hb + 1];
          a0[hb + 1] = iw * a0[hb + 1];
          for (int al = 0; al < NBC; ++al) b[hb][al] = iw * b[hb][al];
        }
      }
    }
    // smoothness indicators (maximum over the variables, in variable order)
    // and nonlinear weights (src/zisa/reconstruction/hybrid_weno.cpp:110-127)
    double alpha[NB + 1];
    double al_tot = 0.0;
#pragma unroll
    for (int k = 0; k <= NB; ++k) {
      double s = 0.0;
      if (k == 0) {
        for (int al = 0; al < NC; ++al) s += c[al] * c[al];
      } else {
        for (int al = 0; al < NBC; ++al) s += b[k - 1][al] * b[k - 1][al];
      }
      double IS = xch.bcast(s, 0);
#pragma unroll
      for (int v = 1; v < NV; ++v) IS = dmax(IS, xch.bcast(s, v));
      double isp = m.exponent == 4.0 ? sqr(sqr(IS)) : pow(IS, m.exponent);
      alpha[k] = m.lw[k] / (m.eps + isp);
      al_tot += alpha[k];
    }
    const double w0 = alpha[0] / al_tot;
    a[0] = 0.0 + w0 * a0[0];
    for (int al = 1; al < DOFC; ++al) a[al] = 0.0 + w0 * c[al - 1];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const double w = alpha[k + 1] / al_tot;
      a[0] = a[0] + w * a0[k + 1];
      for (int al = 1; al <= NBC; ++al) a[al] = a[al] + w * b[k][al - 1];
    }
  }

  // moments: zero for alpha < 3
  double cm[DOFC];
  load_moments<DOFC>(m, i, cm);
  double xc[3];
  load_centre<DIM>(m, i, xc);
  const double len = m.len[i];
  const uint32_t fperm = m.fperm[i];

  // this variable of the scaled reconstruction at x, given the equilibrium
  // point values (r_eq, E_eq) there; delta is the scaled deviation
  const double inv_len = 1.0 / len;
  auto eval = [&](const double *x, double r_eq, double E_eq, double &u, double &delta) {
    double Xs[3] = {(x[0] - xc[0]) * inv_len, (x[1] - xc[1]) * inv_len, DIM == 3 ? (x[2] - xc[2]) * inv_len : 0.0};
    double px;
    poly_eval<DIM, DEGC, 1>(reinterpret_cast<const double(*)[1]>(a), cm, Xs, &px);
    delta = scale_g * px;
    u = delta;
    if (eq_sub) u = is_rho ? r_eq + delta : (is_E ? E_eq + delta : 0.0 + delta);
  };

  double s_acc = 0.0;
  const bool gravity_src = (m.gravity != 0) && (src != nullptr);
  // EquilibriumFluxBC (flux_bc 2, include/zisa/boundary/equilibrium_flux_bc.hpp:36-62):
  // at the cell's boundary faces, the flux of its isentropic equilibrium
  // (this cell's own, solved above: the same solve_exact from the same
  // average) as a state at rest, (0, p_eq n, 0), integrated with the face
  // rule and subtracted.  It has the form of the well-balanced source's face
  // term, w p_eq (sgn n), so it enters that sum as w p_eq ((sgn - 1) n): the
  // two cancel at a face whose frame normal points out of the cell.  An
  // equilibrium that was not found gives NaN, as the reference's zero
  // extrapolation does.  (Isentropic only: never in the light physics set.)
  const bool eq_flux_bc = PH != PH_LIGHT && m.flux_bc == 2 && gravity_src;

  const FaceWords cfw = load_face_words(m, i);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int32_t cf = cfw.w[f];
    double w3[3][3];
    load_face_vertices<DIM>(m, i, fperm, f, w3);
    double nn[3], t1[3], t2[3], area;
    face_frame<DIM>(w3, nn, t1, t2, area);  // (the vertices are needed here anyway)
    const bool exterior = (cf & 2) != 0;
    // outward normal sign for the source term (src/zisa/math/face.cpp:24-27)
    double x0f[3];
    face_point<DIM>(m, 0, w3, x0f);
    double sgn = 0.0;
    if (gravity_src && wb) {
      double dd = nn[0] * (x0f[0] - xc[0]) + nn[1] * (x0f[1] - xc[1]) + (DIM == 3 ? nn[2] * (x0f[2] - xc[2]) : 0.0);
      sgn = double((0.0 < dd) - (dd < 0.0));
    }
    // equilibrium point values of this face: lane k evaluates point q = k
    // (mod NV), every lane receives all of them
    double eqr[Q], eqE[Q];
    for (int q = 0; q < Q; ++q) eqr[q] = eqE[q] = 0.0;
    const bool need_eq = eq_sub && (!exterior || gravity_src);
    if (need_eq) {
#pragma unroll
      for (int q0 = 0; q0 < Q; q0 += NV) {
        double my_r = 0.0, my_E = 0.0;
        if (q0 + g < Q) {
          double x[3];
          face_point<DIM>(m, q0 + g, w3, x);
          const double ph = tab ? m.pot[pot_index_face<DIM>(m, Q, f, q0 + g) + i] : phi<DIM, PH>(m, x);
          eq_at_phi<PH>(m, thc, ph, my_r, my_E);
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const double r = xch.bcast(my_r, k), E = xch.bcast(my_E, k);
          if (q0 + k < Q) {
            eqr[q0 + k] = r;
            eqE[q0 + k] = E;
          }
        }
      }
    }
    double fs = 0.0;
    const double bc_sub = (eq_flux_bc && cf == 2) ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (!exterior) {
        double x[3];
        face_point<DIM>(m, q, w3, x);
        double u, dl;
        eval(x, eqr[q], eqE[q], u, dl);
        // in the global frame: the face kernel rotates the momentum into the
        // face frame, with every component in one thread
        fpv[(int64_t(i) * NF + f) * slot_doubles<DIM, Q>() + q * NV + g] = u;
      }
      if (gravity_src && eq_sub && g >= 1 && g <= DIM) {
        double p_eq = eqE[q] * (gamma - 1.0);
        double w = area * m.fw[q];
        double t = p_eq * ((sgn - bc_sub) * nn[g - 1]);
        fs = (q == 0) ? w * t : fma(w, t, fs);
      }
    }
    if (gravity_src && wb) s_acc += 1.0 * fs;
    if (eq_flux_bc && cf == 2 && !eq_sub) s_acc = __builtin_nan("");
  }

  if (gravity_src) {
    const double vol_i = m.vol[i];
    // The volume terms, -(rho - rho_eq) grad phi for the momentum and
    // -m . grad phi for the energy, need only the lanes' own deviations (the
    // equilibrium is at rest and differs from the state only in rho and E;
    // without one the deviation is the state): every lane forms -sum_q w_q
    // delta(x_q) d_d phi(x_q) for every d; the density lane's sums feed the
    // momentum lanes, momentum lane d's d-th sum the energy lane.
    double acc[DIM];
    if (PH == PH_LIGHT && m.vol_mean) {
      // constant gradient (axial gravity): sum_q w_q delta(x_q) grad phi is
      // vol grad phi times the rule's mean of delta over the cell, which is
      // its constant coefficient (the rule and the moments integrate every
      // other basis function, X^alpha - c_alpha, to zero)
      double gp[3];
      grad_phi<DIM, PH>(m, xc, gp);
      const double s0 = vol_i * (scale_g * a[0]);
#pragma unroll
      for (int d = 0; d < DIM; ++d) acc[d] = -s0 * gp[d];
    } else if (PH == PH_LIGHT) {
      // constant gradient, rules that do not integrate the basis exactly:
      // the quadrature over the points
      load_corners<DIM>(m, i, vtx);
#pragma unroll 1
      for (int q = 0; q < m.qv; ++q) {
        double x[3];
        volume_point<DIM>(m, q, vtx, x);
        double u, dl;
        eval(x, 0.0, 0.0, u, dl);
        double gp[3];
        grad_phi<DIM, PH>(m, x, gp);
        const double w = vol_i * m.vw[q];
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = (q == 0) ? w * (-dl * gp[d]) : fma(w, -dl * gp[d], acc[d]);
      }
    } else {
      // from the polynomial and the cell's source moments (store_source_moments)
#pragma unroll
      for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
      // (the cell's DOFC x DIM moments are contiguous: 16-byte loads)
      double sm[DOFC * DIM];
      {
        const PPair *sp = reinterpret_cast<const PPair *>(m.smom + int64_t(i) * (DOFC * DIM));
#pragma unroll
        for (int k = 0; k < DOFC * DIM / 2; ++k) {
          const PPair v = sp[k];
          sm[2 * k] = v.lo;
          sm[2 * k + 1] = v.hi;
        }
      }
#pragma unroll
      for (int al = 0; al < DOFC; ++al)
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = fma(a[al], sm[al * DIM + d], acc[d]);
#pragma unroll
      for (int d = 0; d < DIM; ++d) acc[d] = -scale_g * acc[d];
    }
    double cs_ = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double rs = xch.bcast(acc[d], 0);      // the density lane's sum for direction d
      const double bm = xch.bcast(acc[d], 1 + d);  // momentum lane d's sum for its own direction
      if (g == 1 + d) cs_ = rs;
      if (is_E) cs_ = (d == 0) ? bm : cs_ + bm;
    }
    double sum = wb ? (s_acc + 1.0 * cs_) : 1.0 * cs_;
    src[g * n + i] = sum / vol_i;
  }
}

// ---- the per-face flux ----------------------------------------------------------
// G[v][e] = sum_q w_q F(uL_q, uR_q) in the face frame, w_q the normalized
// face weights (the face measure is applied by the cell update).
// a flux g given in the face frame (normal, t1, t2 components of the
// momentum) in the global frame
template <int DIM>
ZFV_HD void world_flux(const double *g, const double *nn, const double *t1, const double *t2, double *wv) {
  constexpr int NV = DIM + 2;
  wv[0] = g[0];
  wv[NV - 1] = g[NV - 1];
  if (DIM == 2) {
    wv[1] = g[1] * nn[0] + g[2] * t1[0];
    wv[2] = g[1] * nn[1] + g[2] * t1[1];
  } else {
    wv[1] = g[1] * nn[0] + g[2] * t1[0] + g[3] * t2[0];
    wv[2] = g[1] * nn[1] + g[2] * t1[1] + g[3] * t2[1];
    wv[3] = g[1] * nn[2] + g[2] * t1[2] + g[3] * t2[2];
  }
}

template <int DIM, int Q>
ZFV_HD void face_flux(const Img &m, int64_t e, const double *__restrict__ fpv, double *__restrict__ G) {
  constexpr int NV = DIM + 2;
  constexpr int NF = DIM + 1;
  const int64_t ne = m.ne;
  // the two cells' point values of this face (Q x NV doubles at the start of
  // each slot, momentum in the global frame), in 16-byte loads
  constexpr int SQ = slot_doubles<DIM, Q>();
  double bL[SQ], bR[SQ];
  const SlotPair sl = reinterpret_cast<const SlotPair *>(m.fslot)[e];
  const PPair *pL = reinterpret_cast<const PPair *>(fpv + int64_t(sl.l) * SQ);
  const PPair *pR = reinterpret_cast<const PPair *>(fpv + int64_t(sl.r) * SQ);
#pragma unroll
  for (int k = 0; k < SQ / 2; ++k) {
    const PPair l = pL[k], r = pR[k];
    bL[2 * k] = l.lo;
    bL[2 * k + 1] = l.hi;
    bR[2 * k] = r.lo;
    bR[2 * k + 1] = r.hi;
  }
  // the face frame (the left cell's vertex order, as both cells see it)
  double nn[3], t1[3], t2[3], area;
  face_table_frame<DIM>(m, e, nn, t1, t2, area);
  auto rotate = [&](double *u) {
    const double mc[3] = {u[1], u[2], DIM == 3 ? u[3] : 0.0};
    u[1] = DIM == 2 ? mc[0] * nn[0] + mc[1] * nn[1] : mc[0] * nn[0] + mc[1] * nn[1] + mc[2] * nn[2];
    u[2] = DIM == 2 ? mc[0] * t1[0] + mc[1] * t1[1] : mc[0] * t1[0] + mc[1] * t1[1] + mc[2] * t1[2];
    if (DIM == 3) u[3] = mc[0] * t2[0] + mc[1] * t2[1] + mc[2] * t2[2];
  };
  double acc[NV];
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double uL[NV], uR[NV], f[NV];
    for (int v = 0; v < NV; ++v) {
      uL[v] = bL[q * NV + v];
      uR[v] = bR[q * NV + v];
    }
    rotate(uL);
    rotate(uR);
    if (m.flux == 0) hllc<NV>(uL, uR, m.gamma, f);
    else rusanov<NV>(uL, uR, m.gamma, f);
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + m.fw[q] * f[v];
  }
  // back to the global frame, times the face's measure: the flux through the
  // face, which the left cell loses and the right cell gains (cell_rate)
  double wv[NV];
  world_flux<DIM>(acc, nn, t1, t2, wv);
  if (DIM == 2) {
    PPair *gp = reinterpret_cast<PPair *>(G + e * NV);
    gp[0] = PPair{area * wv[0], area * wv[1]};
    gp[1] = PPair{area * wv[2], area * wv[NV - 1]};
  } else {
    for (int v = 0; v < NV; ++v) G[v * ne + e] = area * wv[v];
  }
}

// ---- the per-cell rate of change -------------------------------------------------
// L = -(1/vol) sum_faces (+-) |face| R^T G  (+ source) (+ flux BC)
template <int DIM>
ZFV_HD void cell_rate(const Img &m, int64_t i, const double *__restrict__ U, const double *__restrict__ G,
                      const double *__restrict__ src, double *L) {
  constexpr int NV = DIM + 2;
  constexpr int NF = DIM + 1;
  const int64_t n = m.n;
  double acc[NV];
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  const FaceWords cfw = load_face_words(m, i);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int32_t cf = cfw.w[f];
    const bool exterior = (cf & 2) != 0;
    // ghost-pair faces (cf == 3) carry neither a flux nor the flux BC; the
    // equilibrium flux BC (flux_bc 2) is part of the source term
    if (exterior && (m.flux_bc != 1 || (cf & 1))) continue;
    if (!exterior) {
      // the face kernel's flux through the face (global frame, times area)
      const int64_t e = cf >> 2;
      const double sgn = (cf & 1) ? 1.0 : -1.0;
      double g[NV];
      if (DIM == 2) {
        const PPair *gp = reinterpret_cast<const PPair *>(G + e * NV);
        const PPair g0 = gp[0], g1 = gp[1];
        g[0] = g0.lo;
        g[1] = g0.hi;
        g[2] = g1.lo;
        g[NV - 1] = g1.hi;
      } else {
        for (int v = 0; v < NV; ++v) g[v] = G[v * m.ne + e];
      }
      for (int v = 0; v < NV; ++v) acc[v] += sgn * g[v];
      continue;
    }
    // FluxBC: physical flux of the cell average in the face frame
    double nn[3], t1[3], t2[3], area;
    load_face_frame<DIM>(m, i, f, nn, t1, t2, area);
    double g[NV], u[NV], r[NV];
    for (int v = 0; v < NV; ++v) u[v] = U[v * n + i];
    r[0] = u[0];
    r[NV - 1] = u[NV - 1];
    if (DIM == 2) {
      r[1] = u[1] * nn[0] + u[2] * nn[1];
      r[2] = u[1] * t1[0] + u[2] * t1[1];
    } else {
      r[1] = u[1] * nn[0] + u[2] * nn[1] + u[3] * nn[2];
      r[2] = u[1] * t1[0] + u[2] * t1[1] + u[3] * t1[2];
      r[3] = u[1] * t2[0] + u[2] * t2[1] + u[3] * t2[2];
    }
    euler_flux<NV>(r, pressure_of<NV>(r, m.gamma), g);
    // the flux BC uses the face measure directly (weights sum to one)
    double wv[NV];
    world_flux<DIM>(g, nn, t1, t2, wv);
    for (int v = 0; v < NV; ++v) acc[v] += -1.0 * (area * wv[v]);
  }
  const double vol = m.vol[i];
  for (int v = 0; v < NV; ++v) L[v] = acc[v] / vol;
  if (m.gravity != 0 && src != nullptr)
    for (int v = 0; v < NV; ++v) L[v] += src[v * n + i];
}

}  // namespace zfv
