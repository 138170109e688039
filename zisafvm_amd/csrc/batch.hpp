// Batched execution: many instances of one kernel shape advance together.
// Each instance keeps its own arrays; a device table of DevInst describes
// them, and block tables map every launched block to (instance, local block).
//
// The per-block work of the batch kernels is written here once, for the
// gfx950 kernels (kernels.hip) and for the test-only host emulation, which
// runs these same functions over the same block tables with bounds-checked
// host arrays.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "stage_math.hpp"

namespace zfv {

struct DevInst {
  Img m;
  double *U, *X, *K, *fpv, *G, *src;
  double cfl, t, t_end, dt;
  unsigned long long dt_bits;
  int active;
  int steps_done;
  int max_steps;
  int pad;
};

struct BlockRef {
  int inst;   // -1: no work
  int local;  // block index inside the instance
};

// the step's dt of a member: cfl times its CFL minimum (the bits of a
// positive double in dt_bits), clamped to its end time
ZFV_HD double kb_step_dt(const DevInst &d) {
  double m;
  const unsigned long long bits = d.dt_bits;
  memcpy(&m, &bits, sizeof(m));
  double dt = d.cfl * m;
  const double rest = d.t_end - d.t;
  if (dt > rest) dt = rest;
  return dt;
}

// groups of NV lanes per 64-wide wavefront (2D: 16 cells, 3D: 12 cells)
template <int NV>
constexpr int groups_per_wave() {
  return 64 / NV;
}

// Block table: logical blocks of every counted instance, in slot order
// (counts[s] items of instance s, per_block items per block), laid out so
// that hardware block b (dispatched to XCD b mod 8) serves the (b mod 8)-th
// contiguous eighth of the logical blocks when xcd_chunks is set:
// neighbouring cells then share one XCD's L2.  The table has as many entries
// (a multiple of 8, at least 8) as the instances need; the kernels are
// launched with exactly that many blocks.  Returns an empty table when it
// would exceed cap entries.
inline std::vector<BlockRef> block_table(const std::vector<int64_t> &counts, int64_t per_block, int64_t cap,
                                         bool xcd_chunks) {
  std::vector<BlockRef> logical;
  for (size_t s = 0; s < counts.size(); ++s)
    for (int64_t k = 0; k < (counts[s] + per_block - 1) / per_block; ++k) logical.push_back({int(s), int(k)});
  const int64_t size = std::max<int64_t>(8, ((int64_t(logical.size()) + 7) / 8) * 8);
  if (size > cap) return {};
  std::vector<BlockRef> t(size, BlockRef{-1, 0});
  if (!xcd_chunks) {
    std::copy(logical.begin(), logical.end(), t.begin());
    return t;
  }
  const int64_t g8 = size / 8;
  for (int64_t j = 0; j < int64_t(logical.size()); ++j) {
    const int64_t x = j / g8, k = j % g8;
    t[8 * k + x] = logical[j];
  }
  return t;
}

// Reconstruction of the cell that lane group grp of wave `wave` in logical
// block b serves (blocks of `waves` wavefronts); X is the group's lane exchange
template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH, class X>
ZFV_HD void kb_reconstruct_group(const DevInst *insts, BlockRef b, int waves, int wave, int grp, int g, int stage,
                                 X &x) {
  constexpr int GPW = groups_per_wave<DIM + 2>();
  if (b.inst < 0) return;
  const DevInst &d = insts[b.inst];
  if (!d.active) return;
  if (stage == 0 && !(kb_step_dt(d) > 0.0)) return;  // stops this step (kb_update_thread)
  const int64_t i = (int64_t(b.local) * waves + wave) * GPW + grp;
  if (grp >= GPW || i >= d.m.n) return;
  if (d.m.steady != nullptr && (d.m.meta[i] & META_FAR_GHOST) == META_FAR_GHOST) return;
  reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH>(d.m, i, g, x, stage == 0 ? d.U : d.X, d.fpv, d.src);
}

// the face flux of thread t of logical block b (256 faces per block)
template <int DIM, int Q>
ZFV_HD void kb_face_flux_thread(const DevInst *insts, BlockRef b, int t) {
  if (b.inst < 0) return;
  const DevInst &d = insts[b.inst];
  if (!d.active) return;
  const int64_t e = int64_t(b.local) * 256 + t;
  if (e >= d.m.ne) return;
  face_flux<DIM, Q>(d.m, e, d.fpv, d.G);
}

// the SSP3 stage update of thread t's cell (256 cells per block):
// K[stage] = L(U_stage), U_next = U0 + dt sum_j c_j K[j]
// (src/zisa/ode/runge_kutta.cpp:122-165), then the frozen BC
template <int DIM>
ZFV_HD void kb_update_thread(DevInst *insts, BlockRef b, int t, int stage) {
  constexpr int NV = DIM + 2;
  if (b.inst < 0) return;
  DevInst &d = insts[b.inst];
  if (!d.active) return;
  double dt = d.dt;
  if (stage == 0) {
    // the step's dt, from the CFL minimum of this step: every thread of the
    // member forms the same value; the member's first thread records it for
    // the later stages and the step's end (and stops a member that cannot
    // advance, whose threads all leave here)
    dt = kb_step_dt(d);
    if (b.local == 0 && t == 0) {
      d.dt = dt;
      if (!(dt > 0.0)) d.active = 0;
    }
    if (!(dt > 0.0)) return;
  }
  const int64_t i = int64_t(b.local) * 256 + t;
  const Img &m = d.m;
  const int64_t n = m.n;
  if (i >= n) return;
  const double *Us = stage == 0 ? d.U : d.X;
  double *Un = stage == 2 ? d.U : d.X;
  if (m.steady != nullptr && (m.meta[i] & META_FAR_GHOST) == META_FAR_GHOST) {
    for (int v = 0; v < NV; ++v) {
      d.K[(int64_t(stage) * NV + v) * n + i] = 0.0;
      Un[v * n + i] = m.steady[v * n + i];
    }
    return;
  }
  double L[NV];
  cell_rate<DIM>(m, i, Us, d.G, d.src, L);
  // SSP3 next-state coefficients (src/zisa/ode/runge_kutta.cpp:154-165)
  const double c0 = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 1.0 / 6);
  const double c1 = stage == 1 ? 0.25 : 1.0 / 6;
  const double c2 = 2.0 / 3;
  const bool ghost = (m.meta[i] & 16) != 0 && m.steady != nullptr;
  for (int v = 0; v < NV; ++v) {
    if (stage < 2) d.K[(int64_t(stage) * NV + v) * n + i] = L[v];  // the last stage's is read by no later one
    double dudt = 0.0;
    if (stage == 0) {
      dudt += c0 * L[v];
    } else if (stage == 1) {
      dudt += c0 * d.K[(int64_t(0) * NV + v) * n + i];
      dudt += c1 * L[v];
    } else {
      dudt += c0 * d.K[(int64_t(0) * NV + v) * n + i];
      dudt += c1 * d.K[(int64_t(1) * NV + v) * n + i];
      dudt += c2 * L[v];
    }
    double un = d.U[v * n + i] + dt * dudt;
    if (ghost) un = m.steady[v * n + i];
    Un[v * n + i] = un;
  }
}

// LocalCFL candidate of thread t's cell, inradius / (|v| + a), or 1e300
// past the instance's cells; 0 for a NaN or non-positive state
// (include/zisa/model/local_cfl_condition_impl.hpp:25-40)
template <int DIM>
ZFV_HD double kb_cfl_thread(const DevInst &d, BlockRef b, int t) {
  constexpr int NV = DIM + 2;
  const int64_t i = int64_t(b.local) * 256 + t;
  const Img &m = d.m;
  double r = 1.0e300;
  if (i < m.n) {
    const int64_t n = m.n;
    double u[NV];
    for (int v = 0; v < NV; ++v) u[v] = d.U[v * n + i];
    double m2 = u[1] * u[1] + u[2] * u[2];
    if (DIM == 3) m2 += u[3] * u[3];
    double v2 = m2 / (u[0] * u[0]);
    double p = (u[NV - 1] - 0.5 * m2 / u[0]) * (m.gamma - 1.0);
    double a = sqrt(m.gamma * p / u[0]);
    r = m.inr[i] / (sqrt(v2) + a);
    if (!(r > 0.0)) r = 0.0;
  }
  return r;
}

// per-instance step bookkeeping: the CFL minimum of a step starts at 1e300
// (set when the member joins and at the end of every step); the step's dt is
// cfl times the minimum, clamped to t_end (an instance that cannot advance
// stops: kb_update_thread of stage 0); the end of a step advances the time
ZFV_HD unsigned long long kb_no_bound() {
  const double big = 1.0e300;
  unsigned long long bits;
  memcpy(&bits, &big, sizeof(bits));
  return bits;
}

ZFV_HD void kb_end_instance(DevInst &d) {
  if (!d.active) return;
  d.t += d.dt;
  d.steps_done += 1;
  if (d.t >= d.t_end || d.steps_done >= d.max_steps) d.active = 0;
  d.dt_bits = kb_no_bound();
}

}  // namespace zfv
