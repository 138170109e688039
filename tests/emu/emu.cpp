// TEST-ONLY CPU emulation of the HIP kernels.
//
// Runs the product's own setup (zisafvm_amd/csrc/setup.cpp) and the same
// per-cell / per-face functions the gfx950 kernels call
// (zisafvm_amd/csrc/stage_math.hpp) in host loops, so the CPU test suite can
// check the product's numerics against the oracle without a GPU.  The
// product library never loads this; GPU tests call libzfv_hip.so.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <memory>
#include <string>
#include <vector>

#include "../../zisafvm_amd/csrc/common.hpp"
#include "../../zisafvm_amd/csrc/setup.hpp"
#include "../../zisafvm_amd/csrc/stage_math.hpp"

using namespace zfv;

namespace {
thread_local std::string g_err;

struct Emu {
  HostImage img;
  Img m{};
  std::vector<double> U, fpv, G, src, phim, pot, frm;
};

// The lanes of one cell's group run as NV host threads in lockstep: every
// bcast is a barrier-separated exchange through shared slots, as the
// wavefront's cross-lane moves are on the GPU.  A lane that calls bcast a
// different number of times than its group deadlocks the group; the
// barrier then times out and the emulation reports the error.
struct LaneBarrier {
  int n;
  std::atomic<int> count{0};
  std::atomic<long> gen{0};
  std::atomic<bool> broken{false};
  double slot[8] = {0};
  double lds[128] = {0};  // the group's LDS slice (slice_doubles <= 128)
  explicit LaneBarrier(int n_) : n(n_) {}
  void wait() {
    if (broken.load()) throw std::runtime_error("lane group broken");
    const long my = gen.load(std::memory_order_acquire);
    if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
      count.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
      return;
    }
    const auto t0 = std::chrono::steady_clock::now();
    long spins = 0;
    while (gen.load(std::memory_order_acquire) == my) {
      if (broken.load()) throw std::runtime_error("lane group broken");
      if (++spins % 1024 == 0) {
        std::this_thread::yield();
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) {
          broken = true;
          throw std::runtime_error("lane group: bcast calls differ between lanes");
        }
      }
    }
  }
};

struct EmuXch {
  LaneBarrier *b;
  int g;
  double bcast(double v, int k) {
    b->slot[g] = v;
    b->wait();
    const double r = b->slot[k];
    b->wait();
    return r;
  }
  int32_t bcast_i(int32_t v, int k) { return int32_t(bcast(double(v), k)); }
  double *lds() { return b->lds; }
  void sync() { b->wait(); }
};

template <int DIM, int DOFC, int RC, int NB, int RB, int Q>
void run(Emu &e, const double *U, double *L) {
  constexpr int NV = DIM + 2;
  const int64_t n = e.m.n;
  {
    LaneBarrier bar(NV);
    std::vector<std::thread> lanes;
    std::vector<std::string> errs(NV);
    for (int g = 0; g < NV; ++g)
      lanes.emplace_back([&, g] {
        try {
          EmuXch x{&bar, g};
          // the physics set the GPU picks for these parameters
          const int ph = physics_set(e.m);
          double *src = e.src.empty() ? nullptr : e.src.data();
          for (int64_t i = 0; i < n; ++i) {
            if (ph == PH_LIGHT)
              reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH_LIGHT>(e.m, i, g, x, U, e.fpv.data(), src);
            else if (ph == PH_LIN)
              reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH_LIN>(e.m, i, g, x, U, e.fpv.data(), src);
            else
              reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH_FULL>(e.m, i, g, x, U, e.fpv.data(), src);
          }
        } catch (const std::exception &ex) {
          errs[g] = ex.what();
        }
      });
    for (auto &t : lanes) t.join();
    for (auto &m : errs)
      if (!m.empty()) throw ZfvError(ZFV_ERR_INVALID, "emulation: " + m);
  }
  for (int64_t f = 0; f < e.m.ne; ++f) face_flux<DIM, Q>(e.m, f, e.fpv.data(), e.G.data());
  for (int64_t i = 0; i < n; ++i) {
    double l[NV];
    cell_rate<DIM>(e.m, i, U, e.G.data(), e.src.empty() ? nullptr : e.src.data(), l);
    for (int v = 0; v < NV; ++v) L[v * n + i] = l[v];
  }
}

void dispatch(Emu &e, const double *U, double *L) {
  const Shape &s = e.img.s;
#define EMU_SHAPE(D, DOF, RC_, NB_, RB_, Q_) \
  if (s.dim == D && s.dofc == DOF && s.rc == RC_ && s.nb == NB_ && s.rb == RB_ && s.q == Q_) \
    return run<D, DOF, RC_, NB_, RB_, Q_>(e, U, L);
  EMU_SHAPE(2, 6, 10, 3, 3, 2)
  EMU_SHAPE(2, 6, 10, 3, 3, 3)
  EMU_SHAPE(2, 10, 18, 3, 3, 3)
  EMU_SHAPE(2, 10, 18, 3, 3, 2)
  EMU_SHAPE(2, 3, 4, 3, 3, 1)
  EMU_SHAPE(2, 3, 4, 3, 3, 2)
  EMU_SHAPE(3, 10, 18, 4, 4, 3)
  EMU_SHAPE(3, 10, 18, 4, 4, 4)
  EMU_SHAPE(3, 10, 18, 3, 4, 3)
#undef EMU_SHAPE
  throw ZfvError(ZFV_ERR_INVALID, "emulation: shape not instantiated");
}
}  // namespace

extern "C" {

const char *zfv_emu_last_error(void) { return g_err.c_str(); }

int zfv_emu_create(const zfv_mesh *mesh, const zfv_params *p, void **out) {
  try {
    auto e = std::make_unique<Emu>();
    build_image(*mesh, *p, e->img);
    const HostImage &im = e->img;
    Img &m = e->m;
    m.n = im.n;
    m.ne = im.ne;
    m.P = im.P.data();
    m.rows = im.rows.data();
    m.twin = im.twin.data();
    m.crn = im.crn.data();
    m.xc = im.xc.data();
    m.len = im.len.data();
    m.vol = im.vol.data();
    m.inr = im.inr.data();
    m.mom = im.mom.data();
    m.meta = im.meta.data();
    m.cface = im.cface.data();
    m.fperm = im.fperm.data();
    m.steady = nullptr;
    m.phim = nullptr;
    m.pot = nullptr;
    m.frm = nullptr;
    m.gamma = p->gamma;
    m.eps = p->epsilon;
    m.exponent = p->exponent;
    for (int k = 0; k < 5; ++k) m.lw[k] = im.linear_weights[k];
    m.mode = p->reconstruction;
    m.flux = p->flux;
    m.eq = p->equilibrium;
    m.gravity = p->gravity;
    m.g = p->gravity_g;
    for (int d = 0; d < 3; ++d) m.axis[d] = p->gravity_axis[d];
    m.rhoC = p->polytrope_rhoC;
    m.K = p->polytrope_K;
    m.G = p->polytrope_G;
    m.flux_bc = p->flux_bc;
    for (int q = 0; q < 16; ++q)
      for (int d = 0; d < 3; ++d) m.shift[q][d] = im.shift[3 * q + d];
    const int fdim = im.s.dim == 2 ? 1 : 3;
    for (size_t q = 0; q < im.face_w.size(); ++q) {
      m.fw[q] = im.face_w[q];
      for (int d = 0; d < fdim; ++d) m.fx[q][d] = im.face_x[q * fdim + d];
    }
    m.qv = int(im.vol_w.size());
    for (size_t q = 0; q < im.vol_w.size(); ++q) {
      m.vw[q] = im.vol_w[q];
      for (int d = 0; d < im.s.dim + 1; ++d) m.vx[q][d] = im.vol_x[q * (im.s.dim + 1) + d];
    }
    if (p->equilibrium == ZFV_EQ_ISENTROPIC && p->gravity != ZFV_GRAVITY_NONE && p->gamma == 2.0) {
      e->phim.assign(size_t(2) * im.n, 0.0);
      for (int64_t i = 0; i < im.n; ++i) {
        if (im.s.dim == 2)
          phi_moments<2>(m, i, e->phim[i], e->phim[im.n + i]);
        else
          phi_moments<3>(m, i, e->phim[i], e->phim[im.n + i]);
      }
      m.phim = e->phim.data();
    }
    // the face frame tables the GPU path keeps
    e->frm.assign(size_t(im.s.dim == 2 ? 3 * frame_words<2>() : 4 * frame_words<3>()) * im.n, 0.0);
    for (int64_t i = 0; i < im.n; ++i) {
      if (im.s.dim == 2)
        store_face_frames<2>(m, i, e->frm.data());
      else
        store_face_frames<3>(m, i, e->frm.data());
    }
    m.frm = e->frm.data();
    if (physics_set(m) != PH_LIGHT && p->gravity != ZFV_GRAVITY_NONE) {
      // the potential tables the GPU path keeps for these physics sets
      const int rows = im.s.dim == 2 ? pot_rows<2>(m.qv, im.s.q) : pot_rows<3>(m.qv, im.s.q);
      e->pot.assign(size_t(rows) * im.n, 0.0);
      for (int64_t i = 0; i < im.n; ++i) {
        if (im.s.dim == 2)
          potential_points<2>(m, im.s.q, i, e->pot.data());
        else
          potential_points<3>(m, im.s.q, i, e->pot.data());
      }
      m.pot = e->pot.data();
    }
    const int nv = im.s.nvars;
    e->fpv.assign(size_t(2) * im.s.q * nv * std::max<int64_t>(im.ne, 1), 0.0);
    e->G.assign(size_t(nv) * std::max<int64_t>(im.ne, 1), 0.0);
    if (p->gravity != 0) e->src.assign(size_t(nv) * im.n, 0.0);
    *out = e.release();
    return 0;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

int zfv_emu_destroy(void *h) {
  delete static_cast<Emu *>(h);
  return 0;
}

// dudt += L(u), both n x 5 row-major in caller order
int zfv_emu_rate(void *h, const double *u, double *dudt) {
  try {
    Emu &e = *static_cast<Emu *>(h);
    const HostImage &im = e.img;
    const int64_t n = im.n;
    const int nv = im.s.nvars, dim = im.s.dim;
    std::vector<double> U(size_t(nv) * n), L(size_t(nv) * n);
    for (int64_t ip = 0; ip < n; ++ip) {
      const double *a = u + 5 * int64_t(im.perm[ip]);
      U[ip] = a[0];
      for (int d = 0; d < dim; ++d) U[(1 + d) * n + ip] = a[1 + d];
      U[(nv - 1) * n + ip] = a[4];
    }
    dispatch(e, U.data(), L.data());
    for (int64_t ip = 0; ip < n; ++ip) {
      double *a = dudt + 5 * int64_t(im.perm[ip]);
      a[0] += L[ip];
      for (int d = 0; d < dim; ++d) a[1 + d] += L[(1 + d) * n + ip];
      a[4] += L[(nv - 1) * n + ip];
    }
    return 0;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

int zfv_emu_stencils(void *h, int i, int *l2g, int cap, int *l2g_size, int *members, int members_cap) {
  const HostImage &im = static_cast<Emu *>(h)->img;
  const int a = im.l2g_off[i], b = im.l2g_off[i + 1];
  *l2g_size = b - a;
  for (int q = a; q < b && q - a < cap; ++q) l2g[q - a] = im.l2g[q];
  const int ma = im.mem_off[i], mb = im.mem_off[i + 1];
  if (mb - ma > members_cap) return -1;
  for (int q = ma; q < mb; ++q) members[q - ma] = im.members[q];
  return mb - ma;
}

// Every index the kernels derive from the instance's setup data, checked
// against the array it addresses: stencil row entries (cell < n, translation
// code inside the table), twin rows, face words (face id < ne, ghost pairs),
// face-vertex permutations, the CWENO index.  Returns 0 when all hold, else
// writes the first violation into msg.
int zfv_emu_validate(void *h, char *msg, int cap) {
  const Emu &e = *static_cast<Emu *>(h);
  const HostImage &im = e.img;
  const Shape &s = im.s;
  const int64_t n = im.n, ne = im.ne;
  std::string err;
  auto bad = [&](const std::string &what, int64_t i) {
    if (err.empty()) err = what + " at cell " + std::to_string(i);
  };
  int n_codes = 1;
  for (int q = 1; q < 16; ++q)
    if (im.shift[3 * q] != 0.0 || im.shift[3 * q + 1] != 0.0 || im.shift[3 * q + 2] != 0.0) n_codes = q + 1;
  if (int64_t(im.rows.size()) != int64_t(s.nrows) * n) bad("rows size", -1);
  if (int64_t(im.cface.size()) != int64_t(s.nfaces) * n) bad("cface size", -1);
  std::vector<int> face_uses(size_t(ne > 0 ? ne : 1), 0);
  for (int64_t i = 0; i < n && err.empty(); ++i) {
    for (int r = 0; r < s.nrows; ++r) {
      const int32_t rw = im.rows[size_t(r) * n + i];
      if (rw < 0) bad("negative row entry", i);
      const int64_t j = rw & ROW_CELL_MASK;
      const int code = int((uint32_t(rw) >> ROW_CELL_BITS) & 15u);
      if (j >= n) bad("row cell out of range", i);
      if (code >= n_codes) bad("row translation code outside the table", i);
    }
    for (int r = s.rc; r < s.nrows; ++r) {
      const int t = im.twin[size_t(r - s.rc) * n + i];
      if (t < -1 || t >= s.rc) bad("twin row out of range", i);
      if (t >= 0 && im.rows[size_t(t) * n + i] != im.rows[size_t(r) * n + i]) bad("twin row names another member", i);
    }
    const int32_t meta = im.meta[i];
    if ((meta & 7) > s.nb) bad("CWENO stencil index out of range", i);
    for (int f = 0; f < s.nfaces; ++f) {
      const int32_t cf = im.cface[size_t(f) * n + i];
      if ((cf & 2) == 0) {
        const int64_t fe = cf >> 2;
        if (fe < 0 || fe >= ne) bad("face id out of range", i);
        else face_uses[size_t(fe)] += 1;
      } else if (cf != 2 && cf != 3) {
        bad("exterior face word with bits above the flags", i);
      }
      for (int r = 0; r < s.dim; ++r)
        if (int((im.fperm[i] >> (2 * (f * 3 + r))) & 3u) > s.dim) bad("face vertex index out of range", i);
    }
  }
  for (int64_t fe = 0; fe < ne && err.empty(); ++fe)
    if (face_uses[size_t(fe)] != 2) bad("interior face " + std::to_string(fe) + " is not shared by exactly two cells", -1);
  if (err.empty()) return 0;
  if (msg && cap > 0) {
    std::strncpy(msg, err.c_str(), size_t(cap) - 1);
    msg[cap - 1] = 0;
  }
  return 1;
}

// The kernels' polynomial evaluation (poly_eval of one variable) at the
// scaled point X = (x - center) / length: coefficients a[alpha], moments c[alpha].
int zfv_emu_poly_eval(int n_dims, int degree, const double *a, const double *c, const double *X, double *out) {
  const double(*aa)[1] = reinterpret_cast<const double(*)[1]>(a);
  double px = 0.0;
  if (n_dims == 2 && degree == 1) poly_eval<2, 1, 1>(aa, c, X, &px);
  else if (n_dims == 2 && degree == 2) poly_eval<2, 2, 1>(aa, c, X, &px);
  else if (n_dims == 2 && degree == 3) poly_eval<2, 3, 1>(aa, c, X, &px);
  else if (n_dims == 2 && degree == 4) poly_eval<2, 4, 1>(aa, c, X, &px);
  else if (n_dims == 3 && degree == 1) poly_eval<3, 1, 1>(aa, c, X, &px);
  else if (n_dims == 3 && degree == 2) poly_eval<3, 2, 1>(aa, c, X, &px);
  else if (n_dims == 3 && degree == 3) poly_eval<3, 3, 1>(aa, c, X, &px);
  else return -1;
  *out = px;
  return 0;
}

int zfv_emu_info(void *h, int64_t *n, int64_t *ne, double *max_cond) {
  const HostImage &im = static_cast<Emu *>(h)->img;
  *n = im.n;
  *ne = im.ne;
  *max_cond = im.max_cond;
  return 0;
}

}  // extern "C"
