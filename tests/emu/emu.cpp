// TEST-ONLY CPU emulation of the HIP kernels.
//
// Runs the product's own setup (zisafvm_amd/csrc/setup.cpp) and the same
// per-cell / per-face functions the gfx950 kernels call
// (zisafvm_amd/csrc/stage_math.hpp) in host loops, so the CPU test suite can
// check the product's numerics against the oracle without a GPU.  The
// product library never loads this; GPU tests call libzfv_hip.so.
#include <algorithm>
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
#include "../../zisafvm_amd/csrc/batch.hpp"

using namespace zfv;

namespace {
thread_local std::string g_err;

struct Emu {
  HostImage img;
  Img m{};
  std::vector<double> U, fpv, G, src, phim, pot, frm, ffr, eqc, cphi, smom;
  std::vector<int32_t> eqcnt;
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
  double lds[128] = {0};   // the group's LDS slice (group_lds_doubles <= 128)
  double wave[640] = {0};  // the row region of its wavefront (wave_lds_doubles <= 640), group 0
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
  double *wave_lds() { return b->wave; }
  int lane0() const { return 0; }  // the emulated group is lane group 0 of its wavefront
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
    m.fslot = im.fslot.data();
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
    m.scaling = p->scaling;
    m.vol_mean = volume_mean_is_constant_term(*p, im.s) ? 1 : 0;
    m.flux = p->flux;
    m.eq = p->equilibrium;
    m.gravity = p->gravity;
    m.g = p->gravity_g;
    for (int d = 0; d < 3; ++d) m.axis[d] = p->gravity_axis[d];
    m.rhoC = p->polytrope_rhoC;
    m.K = p->polytrope_K;
    m.G = p->polytrope_G;
    m.flux_bc = p->flux_bc;
    m.spr = p->steps_per_recompute;
    m.eq_threshold = p->recompute_threshold;
    if (p->steps_per_recompute > 1) {
      // kept equilibria, as the handle keeps them on the device
      e->eqc.assign(size_t(8) * im.n, 0.0);
      e->eqcnt.assign(size_t(im.n), 0);
      m.eqc = e->eqc.data();
      m.eqcnt = e->eqcnt.data();
    }
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
          phi_moments<2>(m, i, e->phim[2 * i], e->phim[2 * i + 1]);
        else
          phi_moments<3>(m, i, e->phim[2 * i], e->phim[2 * i + 1]);
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
    e->ffr.assign(size_t(im.s.dim == 2 ? frame_words<2>() : frame_words<3>()) * std::max<int64_t>(im.ne, 1), 0.0);
    for (int64_t f = 0; f < im.ne; ++f) {
      if (im.s.dim == 2)
        store_face_table<2>(m, f, e->ffr.data());
      else
        store_face_table<3>(m, f, e->ffr.data());
    }
    m.ffr = e->ffr.data();
    for (int c = 0; c < 16; ++c)
      m.shift_phi[c] = im.s.dim == 2 ? phi<2, PH_LIGHT>(m, m.shift[c]) : phi<3, PH_LIGHT>(m, m.shift[c]);
    if (physics_set(m) == PH_LIGHT && m.eq != 0) {
      // the corner potentials the GPU path keeps for the light set
      e->cphi.assign(size_t(im.s.dim + 1) * im.n, 0.0);
      for (int64_t i = 0; i < im.n; ++i) {
        if (im.s.dim == 2)
          store_corner_phi<2>(m, i, e->cphi.data());
        else
          store_corner_phi<3>(m, i, e->cphi.data());
      }
      m.cphi = e->cphi.data();
    }
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
    if (p->gravity != ZFV_GRAVITY_NONE) {
      // the source moments the GPU path keeps
      e->smom.assign(size_t(im.s.dofc) * im.s.dim * im.n, 0.0);
      for (int64_t i = 0; i < im.n; ++i) {
        if (im.s.dim == 3)
          store_source_moments<3, 10>(m, i, e->smom.data());
        else if (im.s.dofc == 10)
          store_source_moments<2, 10>(m, i, e->smom.data());
        else if (im.s.dofc == 6)
          store_source_moments<2, 6>(m, i, e->smom.data());
        else
          store_source_moments<2, 3>(m, i, e->smom.data());
      }
      m.smom = e->smom.data();
    }
    const int nv = im.s.nvars;
    e->fpv.assign(size_t(im.s.nfaces) * ((im.s.q * nv + 1) & ~1) * im.n, 0.0);
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
  if (int64_t(im.cface.size()) != int64_t(4) * n) bad("cface size", -1);
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
      const int32_t cf = im.cface[size_t(i) * 4 + f];
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

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch emulation: the batch kernels' per-block functions (batch.hpp) run in
// host loops over the block tables the library builds, with each member's
// arrays on the host (U, X, K, face values, fluxes, source).  Every launch of
// a step is emulated in the order of the captured step: begin, CFL, dt,
// then per stage reconstruction, face fluxes, update; then the end of step.
namespace {

struct EmuMember {
  Emu *e = nullptr;
  std::vector<double> U, X, K, fpv, G, src, steady;
};

struct EmuBatch {
  std::vector<EmuMember> mem;
  std::vector<DevInst> inst;
  Shape shape;
  int ph = PH_LIGHT;
  // counters of what the last run touched, for the tests
  int64_t rec_groups = 0, face_threads = 0, cell_threads = 0;
};

template <int DIM, int DOFC, int RC, int NB, int RB, int Q>
void batch_step(EmuBatch &B, int waves, bool xcd, int rec_sample) {
  constexpr int NV = DIM + 2;
  constexpr int GPW = groups_per_wave<NV>();
  const int n_inst = int(B.inst.size());
  std::vector<int64_t> cells(n_inst, 0), faces(n_inst, 0);
  for (int s = 0; s < n_inst; ++s)
    if (B.inst[s].active) {
      cells[s] = B.inst[s].m.n;
      faces[s] = B.inst[s].m.ne;
    }
  const int64_t big = int64_t(1) << 40;
  const auto tc = block_table(cells, 256, big, xcd);
  const auto tf = block_table(faces, 256, big, xcd);
  const auto tr = block_table(cells, int64_t(waves) * GPW, big, xcd);
  DevInst *insts = B.inst.data();
  // CFL (the minimum starts at 1e300: set when a member joins and at every
  // step's end): per block the minimum of its threads' candidates, then the minimum
  // over blocks per instance (the kernel's atomic min on the bits of
  // positive doubles)
  for (const BlockRef &b : tc) {
    if (b.inst < 0 || !insts[b.inst].active) continue;
    double r = 1.0e300;
    for (int t = 0; t < 256; ++t) r = std::min(r, kb_cfl_thread<DIM>(insts[b.inst], b, t));
    unsigned long long bits;
    std::memcpy(&bits, &r, sizeof(bits));
    insts[b.inst].dt_bits = std::min(insts[b.inst].dt_bits, bits);
  }
  for (int stage = 0; stage < 3; ++stage) {
    // reconstruction: the NV lanes of a group run as host threads in lockstep
    LaneBarrier bar(NV);
    std::vector<std::thread> lanes;
    std::vector<std::string> errs(NV);
    std::atomic<int64_t> groups{0};
    for (int g = 0; g < NV; ++g)
      lanes.emplace_back([&, g] {
        try {
          EmuXch x{&bar, g};
          for (size_t bi = 0; bi < tr.size(); ++bi) {
            if (rec_sample > 1 && bi % size_t(rec_sample) != 0) continue;
            for (int wave = 0; wave < waves; ++wave)
              for (int grp = 0; grp < GPW; ++grp) {
                if (B.ph == PH_LIGHT)
                  kb_reconstruct_group<DIM, DOFC, RC, NB, RB, Q, PH_LIGHT>(insts, tr[bi], waves, wave, grp, g, stage, x);
                else if (B.ph == PH_LIN)
                  kb_reconstruct_group<DIM, DOFC, RC, NB, RB, Q, PH_LIN>(insts, tr[bi], waves, wave, grp, g, stage, x);
                else
                  kb_reconstruct_group<DIM, DOFC, RC, NB, RB, Q, PH_FULL>(insts, tr[bi], waves, wave, grp, g, stage, x);
                if (g == 0) groups++;
              }
          }
        } catch (const std::exception &ex) {
          errs[g] = ex.what();
        }
      });
    for (auto &t : lanes) t.join();
    for (auto &m : errs)
      if (!m.empty()) throw ZfvError(ZFV_ERR_INVALID, "batch emulation: " + m);
    B.rec_groups += groups.load();
    for (const BlockRef &b : tf)
      for (int t = 0; t < 256; ++t) {
        kb_face_flux_thread<DIM, Q>(insts, b, t);
        ++B.face_threads;
      }
    for (const BlockRef &b : tc)
      for (int t = 0; t < 256; ++t) {
        kb_update_thread<DIM>(insts, b, t, stage);
        ++B.cell_threads;
      }
  }
  for (int s = 0; s < n_inst; ++s) kb_end_instance(insts[s]);
}

void batch_dispatch(EmuBatch &B, int waves, bool xcd, int rec_sample) {
  const Shape &s = B.shape;
#define EMU_BSHAPE(D, DOF, RC_, NB_, RB_, Q_) \
  if (s.dim == D && s.dofc == DOF && s.rc == RC_ && s.nb == NB_ && s.rb == RB_ && s.q == Q_) \
    return batch_step<D, DOF, RC_, NB_, RB_, Q_>(B, waves, xcd, rec_sample);
  EMU_BSHAPE(2, 6, 10, 3, 3, 2)
  EMU_BSHAPE(2, 6, 10, 3, 3, 3)
  EMU_BSHAPE(2, 10, 18, 3, 3, 3)
  EMU_BSHAPE(2, 10, 18, 3, 3, 2)
  EMU_BSHAPE(2, 3, 4, 3, 3, 1)
  EMU_BSHAPE(2, 3, 4, 3, 3, 2)
  EMU_BSHAPE(3, 10, 18, 4, 4, 3)
  EMU_BSHAPE(3, 10, 18, 4, 4, 4)
  EMU_BSHAPE(3, 10, 18, 3, 4, 3)
#undef EMU_BSHAPE
  throw ZfvError(ZFV_ERR_INVALID, "batch emulation: shape not instantiated");
}

}  // namespace

extern "C" {

// A batch of emulated instances (all of one kernel shape), each with its
// initial state (n x 5, caller order), optional steady state (frozen BC),
// CFL number, end time and step limit.
int zfv_emu_batch_create(void **members, int n, const double *const *u0, const double *const *steady,
                         const double *cfl, const double *t_end, const int *max_steps, void **out) {
  try {
    auto B = std::make_unique<EmuBatch>();
    B->mem.resize(n);
    B->inst.resize(n);
    int ph = -1;
    for (int k = 0; k < n; ++k) {
      Emu *e = static_cast<Emu *>(members[k]);
      const HostImage &im = e->img;
      const Shape &s = im.s;
      if (k == 0) B->shape = s;
      const Shape &b = B->shape;
      if (s.dim != b.dim || s.dofc != b.dofc || s.rc != b.rc || s.nb != b.nb || s.rb != b.rb || s.q != b.q)
        throw ZfvError(ZFV_ERR_INVALID, "members of different kernel shapes");
      EmuMember &M = B->mem[k];
      M.e = e;
      const int64_t nn = im.n, ne = std::max<int64_t>(im.ne, 1);
      const int nv = s.nvars;
      M.U.assign(size_t(nv) * nn, 0.0);
      M.X.assign(size_t(nv) * nn, 0.0);
      M.K.assign(size_t(3) * nv * nn, 0.0);
      M.fpv.assign(size_t(s.nfaces) * ((s.q * nv + 1) & ~1) * nn, 0.0);
      M.G.assign(size_t(nv) * ne, 0.0);
      if (e->m.gravity != 0) M.src.assign(size_t(nv) * nn, 0.0);
      to_soa(im, u0[k], M.U.data());
      DevInst d{};
      d.m = e->m;
      if (steady && steady[k]) {
        M.steady.assign(size_t(nv) * nn, 0.0);
        to_soa(im, steady[k], M.steady.data());
        d.m.steady = M.steady.data();
      }
      d.U = M.U.data();
      d.X = M.X.data();
      d.K = M.K.data();
      d.fpv = M.fpv.data();
      d.G = M.G.data();
      d.src = M.src.empty() ? nullptr : M.src.data();
      d.cfl = cfl[k];
      d.t = 0.0;
      d.t_end = t_end[k];
      d.dt_bits = kb_no_bound();
      d.active = 1;
      d.max_steps = max_steps[k] > 0 ? max_steps[k] : 0x7fffffff;
      B->inst[k] = d;
      const int p = physics_set(d.m);
      ph = ph < 0 ? p : physics_union(ph, p);
    }
    B->ph = ph < 0 ? PH_LIGHT : ph;
    *out = B.release();
    return 0;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

int zfv_emu_batch_destroy(void *b) {
  delete static_cast<EmuBatch *>(b);
  return 0;
}

// n_steps steps of the batch with reconstruction blocks of `waves`
// wavefronts; rec_sample > 1 reconstructs only every rec_sample-th block of
// the table (a bounds run at large sizes: the skipped blocks' face values
// are stale, the results not meaningful)
int zfv_emu_batch_run(void *bp, int n_steps, int waves, int xcd_chunks, int rec_sample) {
  try {
    EmuBatch &B = *static_cast<EmuBatch *>(bp);
    if (waves != 1 && waves != 2 && waves != 4) throw ZfvError(ZFV_ERR_INVALID, "waves must be 1, 2 or 4");
    for (int k = 0; k < n_steps; ++k) batch_dispatch(B, waves, xcd_chunks != 0, rec_sample);
    return 0;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

int zfv_emu_batch_member(void *bp, int k, double *u_out, double *t, int *steps, int *active) {
  EmuBatch &B = *static_cast<EmuBatch *>(bp);
  if (k < 0 || k >= int(B.mem.size())) return -1;
  from_soa(B.mem[k].e->img, B.mem[k].U.data(), u_out, false);
  *t = B.inst[k].t;
  *steps = B.inst[k].steps_done;
  *active = B.inst[k].active;
  return 0;
}

int zfv_emu_batch_counts(void *bp, int64_t *rec_groups, int64_t *face_threads, int64_t *cell_threads) {
  EmuBatch &B = *static_cast<EmuBatch *>(bp);
  *rec_groups = B.rec_groups;
  *face_threads = B.face_threads;
  *cell_threads = B.cell_threads;
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
