// gfx950 kernels of the rate-of-change hot path and the C-ABI of
// libzfv_hip.so (declared in include/zfv.h).
//
// One SSP3 stage = three launches over SoA arrays in HBM:
//   k_reconstruct  one group of nvars lanes per cell (lane g: variable g):
//                  stencil gather, LSQ pseudo-inverse mat-vecs, CWENO/WENO-AO,
//                  equilibrium solve (well-balanced), point values on the
//                  cell's faces (face frame), gravity source
//   k_face_flux    one thread per interior face: HLLC/Rusanov at the face
//                  points, integrated once per face, in the global frame
//                  times the face's area
//   k_update       one thread per cell: atomic-free gather of the cell's face
//                  fluxes, source, flux BC, Butcher-form stage sum, frozen BC
// FP64 throughout; no MFMA (the per-cell mat-vecs are memory-bound).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "setup.hpp"
#include "batch.hpp"
#include "stage_math.hpp"

namespace zfv {

static thread_local std::string g_last_error;

#define HIP_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw ::zfv::ZfvError(ZFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));            \
  } while (0)

// ---------------------------------------------------------------------------
// kernels

// Blocks are dispatched to the 8 XCDs round robin (block b to XCD b mod 8).
// The logical block of hardware block b makes every XCD serve one contiguous
// eighth of the cells, so that neighbouring cells share an L2; grids are
// launched as a multiple of 8 blocks and surplus logical blocks find no cells.
__device__ __forceinline__ int64_t xcd_block(unsigned b, unsigned grid) {
  const unsigned per = grid / 8;
  return int64_t(b % 8) * per + b / 8;
}


// Cross-lane exchange inside a group of NV consecutive lanes of a wavefront
// (one cell per group; lane g of the group owns conserved variable g).
struct WaveGroup {
  int base;       // first lane of the group
  double *slice;  // the group's LDS slice
  double *wave;   // the wavefront's row region
  __device__ __forceinline__ double bcast(double v, int k) const { return __shfl(v, base + k, 64); }
  __device__ __forceinline__ int32_t bcast_i(int32_t v, int k) const { return __shfl(v, base + k, 64); }
  __device__ __forceinline__ double *lds() const { return slice; }
  __device__ __forceinline__ double *wave_lds() const { return wave; }
  __device__ __forceinline__ int lane0() const { return base; }
  // orders the group's LDS writes before its reads (and reads before the
  // next writes): the lanes of a group share a wavefront, whose LDS
  // operations execute in order
  __device__ __forceinline__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// A group of 4 lanes (2D) is a DPP quad, lanes 4j .. 4j+3.  A broadcast whose
// lane index the compiler knows (unrolled loops) is a quad_perm move in the
// VALU, with no LDS round trip; a run-time index keeps the LDS permute (a
// select over the four moves measured slower).  The values are the same.
struct QuadGroup : WaveGroup {
  template <int K>
  __device__ __forceinline__ static int q32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, K * 0x55, 0xF, 0xF, false);
  }
  __device__ __forceinline__ static int sel32(int v, int k) {
    switch (k) {
      case 0: return q32<0>(v);
      case 1: return q32<1>(v);
      case 2: return q32<2>(v);
      default: return q32<3>(v);
    }
  }
  __device__ __forceinline__ double bcast(double v, int k) const {
    if (!__builtin_constant_p(k)) return __shfl(v, base + k, 64);
    const long long b = __double_as_longlong(v);
    const unsigned lo = unsigned(sel32(int(b), k)), hi = unsigned(sel32(int(b >> 32), k));
    return __longlong_as_double((long long)((unsigned long long)hi << 32 | lo));
  }
  __device__ __forceinline__ int32_t bcast_i(int32_t v, int k) const {
    if (!__builtin_constant_p(k)) return __shfl(v, base + k, 64);
    return sel32(v, k);
  }
};

// k_reconstruct: blocks of `blockDim.x / 64` waves, 64 / NV cells per wave
template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH>
__global__ __launch_bounds__(256) void k_reconstruct(
    Img m, const double *__restrict__ U, double *__restrict__ fpv, double *__restrict__ src, int skip_far_ghosts) {
  constexpr int NV = DIM + 2;
  constexpr int GPW = groups_per_wave<NV>();
  // one slice of slice_doubles doubles per cell group, sized at launch
  // (rec_lds_bytes) for the block's waves
  extern __shared__ double slices[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int grp = lane / NV;
  const int g = lane - grp * NV;
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t i = (blk * (blockDim.x >> 6) + wave) * GPW + grp;
  if (grp >= GPW || i >= m.n) return;  // whole groups leave together
  // under the frozen BC a ghost cell away from the interior feeds no flux and
  // its own tendency is overwritten: nothing to reconstruct
  if (skip_far_ghosts && (m.meta[i] & META_FAR_GHOST) == META_FAR_GHOST) return;
  double *wv = &slices[wave * wave_block_doubles<DIM, DOFC, RC>()];
  double *sl = wv + wave_lds_doubles<DIM, DOFC>() + grp * group_lds_doubles<RC>();
  if constexpr (NV == 4) {
    QuadGroup x{{lane - g, sl, wv}};
    reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH>(m, i, g, x, U, fpv, src);
  } else {
    WaveGroup x{lane - g, sl, wv};
    reconstruct_lane<DIM, DOFC, RC, NB, RB, Q, PH>(m, i, g, x, U, fpv, src);
  }
}

template <int DIM, int Q>
__global__ __launch_bounds__(256) void k_face_flux(Img m, const double *__restrict__ fpv, double *__restrict__ G) {
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  face_flux<DIM, Q>(m, e, fpv, G);
}

struct StageCoeffs {
  double c[4];  // coefficients of k_0 .. k_stage in the next state's sum
  int nk;       // number of k's in the sum (stage + 1)
  int write_rate_only;  // 1: store L in K[stage] and stop (zfv_rate)
};

// k_update: K[stage] = L(Ustage); Unext = U0 + dt * sum_j c_j K[j]
// (src/zisa/ode/runge_kutta.cpp:122-143), then FrozenBC for ghost cells.
// Ustage, U0 and Unext may name the same buffer (stages 1 and 2 read and
// write the stage state X; a thread reads only its own cell's entries before
// it writes them), so none of them is declared __restrict__.
template <int DIM>
__global__ __launch_bounds__(256) void k_update(Img m, StageCoeffs sc, int stage, double dt, const double *dtp,
                                                const double *Ustage, const double *__restrict__ G,
                                                const double *__restrict__ src, const double *U0,
                                                double *__restrict__ K, double *Unext) {
  constexpr int NV = DIM + 2;
  if (dtp) dt = *dtp;  // the step's dt chosen on the device (zfv_advance)
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t n = m.n;
  if (i >= n) return;
  if (!sc.write_rate_only && m.steady != nullptr && (m.meta[i] & META_FAR_GHOST) == META_FAR_GHOST) {
    // frozen far ghost: its tendency is never used
    for (int v = 0; v < NV; ++v) {
      K[(int64_t(stage) * NV + v) * n + i] = 0.0;
      Unext[v * n + i] = m.steady[v * n + i];
    }
    return;
  }
  double L[NV];
  cell_rate<DIM>(m, i, Ustage, G, src, L);
  // the last stage's tendency enters only its own sum (no later stage reads it)
  if (sc.write_rate_only || stage + 1 < 3)
    for (int v = 0; v < NV; ++v) K[(int64_t(stage) * NV + v) * n + i] = L[v];
  if (sc.write_rate_only) return;
  const bool ghost = (m.meta[i] & 16) != 0 && m.steady != nullptr;
  for (int v = 0; v < NV; ++v) {
    double dudt = 0.0;
    for (int j = 0; j < sc.nk; ++j) {
      if (sc.c[j] != 0.0) {
        const double kj = (j == stage) ? L[v] : K[(int64_t(j) * NV + v) * n + i];
        dudt += sc.c[j] * kj;
      }
    }
    double un = U0[v * n + i] + dt * dudt;
    if (ghost) un = m.steady[v * n + i];
    Unext[v * n + i] = un;
  }
}

// LocalCFL: min_i inradius_i / (|v_i| + a_i); positive doubles order like
// their bit patterns, so an unsigned 64-bit atomic min reduces them.
template <int DIM>
__global__ __launch_bounds__(256) void k_cfl(Img m, const double *__restrict__ U, unsigned long long *out) {
  constexpr int NV = DIM + 2;
  const int64_t n = m.n;
  double r = 1.0e300;
  // grid-stride over the cells: few blocks, so few atomics on one word
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    double u[NV];
    for (int v = 0; v < NV; ++v) u[v] = U[v * n + i];
    double m2 = u[1] * u[1] + u[2] * u[2];
    if (DIM == 3) m2 += u[3] * u[3];
    double v2 = m2 / (u[0] * u[0]);
    double p = (u[NV - 1] - 0.5 * m2 / u[0]) * (m.gamma - 1.0);
    double a = sqrt(m.gamma * p / u[0]);
    double ri = m.inr[i] / (sqrt(v2) + a);
    if (!(ri > 0.0)) ri = 0.0;  // NaN / non-positive states surface as dt = 0
    r = fmin(r, ri);
  }
  // wavefront reduction
  for (int off = 32; off > 0; off >>= 1) r = fmin(r, __shfl_down(r, off, 64));
  __shared__ double wmin[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wmin[wid] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wmin[0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) b = fmin(b, wmin[w]);
    unsigned long long bits;
    memcpy(&bits, &b, sizeof(bits));
    atomicMin(out, bits);
  }
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// per-cell face frame tables
template <int DIM>
__global__ __launch_bounds__(256) void k_face_frames(Img m, double *tab) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  store_face_frames<DIM>(m, i, tab);
}

// per interior face its frame, in face order
template <int DIM>
__global__ __launch_bounds__(256) void k_face_table(Img m, double *tab) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  store_face_table<DIM>(m, e, tab);
}

// the gravity source moments of every cell
template <int DIM, int DOFC>
__global__ __launch_bounds__(256) void k_source_moments(Img m, double *tab) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  store_source_moments<DIM, DOFC>(m, i, tab);
}

// the light set's corner potentials
template <int DIM>
__global__ __launch_bounds__(256) void k_corner_phi(Img m, double *tab) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  store_corner_phi<DIM>(m, i, tab);
}

// per-cell potential tables of the cells' own quadrature points
template <int DIM>
__global__ __launch_bounds__(256) void k_potential_points(Img m, int nqf, double *tab) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  potential_points<DIM>(m, nqf, i, tab);
}

// per-cell potential moments for the linear isentropic equilibrium averages
template <int DIM>
__global__ __launch_bounds__(256) void k_phi_moments(Img m, double *out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m.n) return;
  double mean, var;
  phi_moments<DIM>(m, i, mean, var);
  out[2 * i] = mean;
  out[2 * i + 1] = var;
}

// ---------------------------------------------------------------------------
// Batch sizes: the largest batches run on the GPU with results checked are
// 65,536 cells and 8 members.  A batched run of 64 members and 4.0M cells
// ended in a GPU memory fault whose cause has not been found (the block
// tables and the per-block indexing at that size are in bounds on the host
// emulation), so zfv_batch_create refuses larger batches.
constexpr int64_t kBatchMaxCells = 65536;
constexpr int32_t kBatchMaxMembers = 8;

// ---------------------------------------------------------------------------
// Batched execution (the per-block work is in batch.hpp, shared with the host
// emulation of the tests).  The graph of one step is captured once for the
// batch's capacity; adding or removing instances only rewrites the device
// tables between replays.  Block tables already place neighbouring blocks on
// one XCD, so the kernels index them by blockIdx directly.

template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH>
__global__ __launch_bounds__(256) void kb_reconstruct(const DevInst *insts, const BlockRef *blk, int stage) {
  constexpr int NV = DIM + 2;
  constexpr int GPW = groups_per_wave<NV>();
  extern __shared__ double slices[];  // as in k_reconstruct
  const int lane = threadIdx.x & 63;
  const int grp = lane / NV;
  const int g = lane - grp * NV;
  const int wave = threadIdx.x >> 6;
  double *wv = &slices[wave * wave_block_doubles<DIM, DOFC, RC>()];
  WaveGroup x{lane - g, wv + wave_lds_doubles<DIM, DOFC>() + grp * group_lds_doubles<RC>(), wv};
  kb_reconstruct_group<DIM, DOFC, RC, NB, RB, Q, PH>(insts, blk[blockIdx.x], blockDim.x >> 6, wave, grp, g, stage,
                                                      x);
}

template <int DIM, int Q>
__global__ __launch_bounds__(256) void kb_face_flux(const DevInst *insts, const BlockRef *blk) {
  kb_face_flux_thread<DIM, Q>(insts, blk[blockIdx.x], threadIdx.x);
}

template <int DIM>
__global__ __launch_bounds__(256) void kb_update(DevInst *insts, const BlockRef *blk, int stage) {
  kb_update_thread<DIM>(insts, blk[blockIdx.x], threadIdx.x, stage);
}

template <int DIM>
__global__ __launch_bounds__(256) void kb_cfl(DevInst *insts, const BlockRef *blk) {
  const BlockRef b = blk[blockIdx.x];
  if (b.inst < 0) return;  // the whole block shares b: it leaves together
  DevInst &d = insts[b.inst];
  if (!d.active) return;
  double r = kb_cfl_thread<DIM>(d, b, threadIdx.x);
  for (int off = 32; off > 0; off >>= 1) r = fmin(r, __shfl_down(r, off, 64));
  __shared__ double wmin[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wmin[wid] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bmin = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
    unsigned long long bits;
    memcpy(&bits, &bmin, sizeof(bits));
    atomicMin(&d.dt_bits, bits);
  }
}


// A single instance's time loop on the device (zfv_advance): TimeLoop's
// step with LocalCFL (src/zisa/fvm_loops/time_loop.cpp:36-81,
// include/zisa/model/local_cfl_condition_impl.hpp:25-40), dt clamped to the
// end time as the simulation clock does (its library is not vendored in the
// reference), with no host round trip per step.
struct AdvState {
  double t, t_end, cfl, dt;
  unsigned long long dt_bits;
  int steps, max_steps, active, pad;
  // RejectLargeDensityChange (src/zisa/ode/step_rejection.cpp:17-57; off
  // when drho_crit <= 0): factor scales the CFL step, halves on a rejected
  // step (which is retried with factor * its dt, TimeLoop::reject_step,
  // src/zisa/fvm_loops/time_loop.cpp:182-196) and doubles after two good
  // steps in a row, within [min_factor, 1]
  double drho_crit, factor, min_factor, retry_dt;
  int bad, prev_good, retry, rejections;
};

// start value of the CFL reduction (bits of 1e300)
__host__ __device__ inline unsigned long long adv_no_bound() {
  const double big = 1.0e300;
  unsigned long long bits;
  memcpy(&bits, &big, sizeof(bits));
  return bits;
}

// keeps u0 for a step that may be rejected (only with step rejection on)
__global__ __launch_bounds__(256) void k_adv_backup(const AdvState *a, int64_t count, const double *U,
                                                    double *Ubk) {
  if (!(a->drho_crit > 0.0) || !a->active) return;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) Ubk[k] = U[k];
}

// RejectLargeDensityChange::check: |rho1 - rho0| > drho_crit rho0 in any cell
__global__ __launch_bounds__(256) void k_adv_check(AdvState *a, int64_t n, const double *U, const double *Ubk) {
  if (!(a->drho_crit > 0.0) || !a->active) return;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fabs(Ubk[i] - U[i]) > a->drho_crit * Ubk[i]) atomicOr(&a->bad, 1);
}

// a rejected step leaves the state as it was
__global__ __launch_bounds__(256) void k_adv_restore(const AdvState *a, int64_t count, double *U,
                                                     const double *Ubk) {
  if (!a->bad || !a->active) return;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) U[k] = Ubk[k];
}

__global__ void k_adv_choose(AdvState *a) {
  double m;
  unsigned long long bits = a->dt_bits;
  memcpy(&m, &bits, sizeof(m));
  // StepRejection::pick_time_step: factor * the CFL step, or the retry step
  double dt = a->retry ? a->retry_dt : a->factor * (a->cfl * m);
  const double rest = a->t_end - a->t;
  if (dt > rest) dt = rest;
  // stop when finished, when the CFL gives no positive step (a non-physical
  // state) or when the step no longer advances the time
  if (!a->active || !(dt > 0.0) || a->t + dt == a->t) {
    dt = 0.0;  // a finished run's stages leave U as it is (U + 0 * dudt)
    a->active = 0;
  }
  a->dt = dt;
}

__global__ void k_adv_end(AdvState *a) {
  if (!a->active) return;
  const int bad = a->bad;
  // the next step's CFL reduction and density check start here
  a->dt_bits = adv_no_bound();
  a->bad = 0;
  a->retry = 0;
  if (a->drho_crit > 0.0) {
    const double f = a->factor;
    if (bad) {
      a->prev_good = 0;
      a->factor = fmax(a->min_factor, fmin(f / 2.0, 1.0));
      a->retry = 1;
      a->retry_dt = a->factor * a->dt;
      a->rejections += 1;
      return;  // the time does not advance; the state was restored
    }
    a->factor = fmax(a->min_factor, fmin(a->prev_good ? f * 2.0 : f, 1.0));
    a->prev_good = 1;
  }
  a->t += a->dt;
  a->steps += 1;
  if (a->t >= a->t_end || a->steps >= a->max_steps) a->active = 0;
}

__global__ void kb_end_step(DevInst *insts, int n_inst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_inst) kb_end_instance(insts[k]);
}

// plain streaming copy (16-byte vectors, grid-stride) for bandwidth calibration
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ a, double2 *__restrict__ b,
                                                    int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

// ---------------------------------------------------------------------------
// launch plan for one shape

struct Launcher {
  // rec[ph]: physics set ph (PH_LIGHT / PH_FULL / PH_LIN); cfg: reconstruction
  // launch configuration (rec_cfg_waves)
  void (*rec[3])(const Img &, const double *, double *, double *, int cfg, int skip_far_ghosts, hipStream_t);
  void (*face)(const Img &, const double *, double *, hipStream_t);
  void (*update)(const Img &, StageCoeffs, int, double, const double *, const double *, const double *,
                 const double *, const double *, double *, double *, hipStream_t);
  void (*cfl)(const Img &, const double *, unsigned long long *, hipStream_t);
  void (*source_moments)(const Img &, double *, hipStream_t);  // setup: Img::smom
  // batched variants: grid = number of block-table entries
  void (*brec[3])(DevInst *, const BlockRef *, unsigned, int, int cfg, hipStream_t);
  void (*bface)(DevInst *, const BlockRef *, unsigned, hipStream_t);
  void (*bupdate)(DevInst *, const BlockRef *, unsigned, int, hipStream_t);
  void (*bcfl)(DevInst *, const BlockRef *, unsigned, hipStream_t);
  int dim = 2;
  int groups_per_wave = 16;  // cells per reconstruction wave
};

inline unsigned blocks_for(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }
// blocks for n items at bs per block, rounded up to a multiple of 8 (xcd_block)
inline unsigned xcd_grid(int64_t n, int bs) { return (blocks_for(n, bs) + 7u) / 8u * 8u; }

// Reconstruction launch configurations: cfg = 0, 1, 2 runs 1, 2, 4 waves per
// block.  (Variants compiled for 4 waves per SIMD, i.e. at most 128 VGPRs,
// spilled to scratch and were never faster; they are not built.)
constexpr int kRecConfigs = 3;
inline int rec_cfg_waves(int cfg) { return 1 << cfg; }
// LDS of a reconstruction block of `waves` wavefronts: one slice per cell group
template <int DIM, int DOFC, int RC>
inline unsigned rec_lds_bytes(int waves) {
  return unsigned(waves * wave_block_doubles<DIM, DOFC, RC>() * sizeof(double));
}

template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH>
void launch_rec(const Img &m, const double *U, double *fpv, double *src, int cfg, int skip, hipStream_t s) {
  const int waves = rec_cfg_waves(cfg);
  const int cpb = waves * groups_per_wave<DIM + 2>();
  const unsigned lds = rec_lds_bytes<DIM, DOFC, RC>(waves);
  hipLaunchKernelGGL((k_reconstruct<DIM, DOFC, RC, NB, RB, Q, PH>), dim3(xcd_grid(m.n, cpb)), dim3(64 * waves), lds, s,
                     m, U, fpv, src, skip);
}

template <int DIM, int DOFC, int RC, int NB, int RB, int Q, int PH>
void launch_brec(DevInst *d, const BlockRef *b, unsigned nb, int stage, int cfg, hipStream_t s) {
  const int waves = rec_cfg_waves(cfg);
  const unsigned lds = rec_lds_bytes<DIM, DOFC, RC>(waves);
  hipLaunchKernelGGL((kb_reconstruct<DIM, DOFC, RC, NB, RB, Q, PH>), dim3(nb), dim3(64 * waves), lds, s, d, b, stage);
}

template <int DIM, int DOFC, int RC, int NB, int RB, int Q>
Launcher make_launcher() {
  Launcher l;
  l.rec[PH_LIGHT] = [](const Img &m, const double *U, double *fpv, double *src, int cfg, int skip, hipStream_t s) {
    launch_rec<DIM, DOFC, RC, NB, RB, Q, PH_LIGHT>(m, U, fpv, src, cfg, skip, s);
  };
  l.rec[PH_FULL] = [](const Img &m, const double *U, double *fpv, double *src, int cfg, int skip, hipStream_t s) {
    launch_rec<DIM, DOFC, RC, NB, RB, Q, PH_FULL>(m, U, fpv, src, cfg, skip, s);
  };
  l.rec[PH_LIN] = [](const Img &m, const double *U, double *fpv, double *src, int cfg, int skip, hipStream_t s) {
    launch_rec<DIM, DOFC, RC, NB, RB, Q, PH_LIN>(m, U, fpv, src, cfg, skip, s);
  };
  l.face = [](const Img &m, const double *fpv, double *G, hipStream_t s) {
    if (m.ne > 0)
      hipLaunchKernelGGL((k_face_flux<DIM, Q>), dim3(xcd_grid(m.ne, 256)), dim3(256), 0, s, m, fpv, G);
  };
  l.update = [](const Img &m, StageCoeffs sc, int stage, double dt, const double *dtp, const double *Us,
                const double *G, const double *src, const double *U0, double *K, double *Un, hipStream_t s) {
    hipLaunchKernelGGL((k_update<DIM>), dim3(xcd_grid(m.n, 256)), dim3(256), 0, s, m, sc, stage, dt, dtp, Us, G,
                       src, U0, K, Un);
  };
  l.cfl = [](const Img &m, const double *U, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL((k_cfl<DIM>), dim3(std::min(blocks_for(m.n, 256), 512u)), dim3(256), 0, s, m, U, out);
  };
  l.brec[PH_LIGHT] = [](DevInst *d, const BlockRef *b, unsigned nb, int stage, int cfg, hipStream_t s) {
    launch_brec<DIM, DOFC, RC, NB, RB, Q, PH_LIGHT>(d, b, nb, stage, cfg, s);
  };
  l.brec[PH_FULL] = [](DevInst *d, const BlockRef *b, unsigned nb, int stage, int cfg, hipStream_t s) {
    launch_brec<DIM, DOFC, RC, NB, RB, Q, PH_FULL>(d, b, nb, stage, cfg, s);
  };
  l.brec[PH_LIN] = [](DevInst *d, const BlockRef *b, unsigned nb, int stage, int cfg, hipStream_t s) {
    launch_brec<DIM, DOFC, RC, NB, RB, Q, PH_LIN>(d, b, nb, stage, cfg, s);
  };
  l.bface = [](DevInst *d, const BlockRef *b, unsigned nb, hipStream_t s) {
    hipLaunchKernelGGL((kb_face_flux<DIM, Q>), dim3(nb), dim3(256), 0, s, d, b);
  };
  l.bupdate = [](DevInst *d, const BlockRef *b, unsigned nb, int stage, hipStream_t s) {
    hipLaunchKernelGGL((kb_update<DIM>), dim3(nb), dim3(256), 0, s, d, b, stage);
  };
  l.bcfl = [](DevInst *d, const BlockRef *b, unsigned nb, hipStream_t s) {
    hipLaunchKernelGGL((kb_cfl<DIM>), dim3(nb), dim3(256), 0, s, d, b);
  };
  l.source_moments = [](const Img &m, double *tab, hipStream_t s) {
    hipLaunchKernelGGL((k_source_moments<DIM, DOFC>), dim3(blocks_for(m.n, 256)), dim3(256), 0, s, m, tab);
  };
  l.dim = DIM;
  l.groups_per_wave = groups_per_wave<DIM + 2>();
  return l;
}

Launcher pick_launcher(const Shape &s) {
  // instantiated shapes: (dim, dof_central, central rows, one-sided stencils, rows each, face points)
#define ZFV_SHAPE(D, DOF, RC_, NB_, RB_, Q_) \
  if (s.dim == D && s.dofc == DOF && s.rc == RC_ && s.nb == NB_ && s.rb == RB_ && s.q == Q_) \
    return make_launcher<D, DOF, RC_, NB_, RB_, Q_>();
  ZFV_SHAPE(2, 6, 10, 3, 3, 2)   // 2D CWENO-AO order 3 {3,2,2,2}, overfit {2,1.5,...}, edge deg 2-3
  ZFV_SHAPE(2, 6, 10, 3, 3, 3)   // ... edge deg 4-5
  ZFV_SHAPE(2, 10, 18, 3, 3, 3)  // 2D order 4, edge deg 4-5
  ZFV_SHAPE(2, 10, 18, 3, 3, 2)  // 2D order 4, edge deg 2-3
  ZFV_SHAPE(2, 3, 4, 3, 3, 1)    // 2D order 2 (overfit 2.0 central), midpoint rule
  ZFV_SHAPE(2, 3, 4, 3, 3, 2)
  ZFV_SHAPE(3, 10, 18, 4, 4, 3)  // 3D order 3 {3,2,2,2,2}, face deg 2
  ZFV_SHAPE(3, 10, 18, 4, 4, 4)  // ... face deg 3
  ZFV_SHAPE(3, 10, 18, 3, 4, 3)  // 3D order 3 with the reference default {3,2,2,2}
#undef ZFV_SHAPE
  throw ZfvError(ZFV_ERR_INVALID, "no kernel instantiation for this (dimension, order, stencil, quadrature) shape");
}

}  // namespace zfv

// ---------------------------------------------------------------------------
// the handle

struct zfv_handle {
  zfv::HostImage img;
  zfv::Img dimg{};
  zfv::Launcher L{};
  zfv_params params{};
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> allocations;
  int64_t device_bytes = 0;
  // state buffers (device)
  double *U = nullptr, *X = nullptr, *K = nullptr, *fpv = nullptr, *G = nullptr, *src = nullptr;
  double *steady = nullptr;
  unsigned long long *dt_bits = nullptr;
  double *host_stage = nullptr;  // pinned staging n * nvars
  bool has_steady = false;
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = -1.0;
  int graph_steps = 0;
  zfv::AdvState *adv = nullptr;           // device time-loop state (zfv_advance)
  double *Ubk = nullptr;                  // u0 of a step that may be rejected
  hipGraphExec_t adv_graph[2] = {nullptr, nullptr};  // one adaptive step, without / with step rejection
  int rec_cfg = -1;   // reconstruction launch configuration; -1 = not tuned yet
  double rec_ms[zfv::kRecConfigs] = {0, 0, 0};
  int ph = 1;         // physics set of the reconstruction kernel (PH_LIGHT / PH_FULL / PH_LIN)
  // asynchronous snapshot of the state (zfv_snapshot_begin / zfv_snapshot_end)
  double *snap_dev = nullptr, *snap_host = nullptr;
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ready = nullptr;
  bool snap_pending = false;
  // the batch this instance is a member of (zfv_batch_add), or null
  zfv_batch *batch = nullptr;
  int batch_slot = -1;

  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_CHECK(hipMalloc(&p, bytes));
    allocations.push_back(p);
    device_bytes += int64_t(bytes);
    return static_cast<T *>(p);
  }
  template <class T>
  const T *upload(const std::vector<T> &v) {
    T *p = alloc<T>(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  ~zfv_handle() {
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (hipGraphExec_t g : adv_graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (void *p : allocations) (void)hipFree(p);
    if (host_stage) (void)hipHostFree(host_stage);
    if (snap_host) (void)hipHostFree(snap_host);
    if (snap_ready) (void)hipEventDestroy(snap_ready);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace zfv {

// batch membership hooks of the instance calls (defined with the batch)
static void batch_forget(zfv_batch *b, int slot);
static void batch_member_steady(zfv_batch *b, int slot, const double *steady);
static void batch_after_snapshot(zfv_batch *b, hipEvent_t copied);
static void member_settle(zfv_handle *h);

static int fail(const std::exception &e) {
  g_last_error = e.what();
  if (auto *z = dynamic_cast<const ZfvError *>(&e)) return z->code;
  return ZFV_ERR_INVALID;
}

int api_fail(const std::exception &e) { return fail(e); }

#define ZFV_API_BEGIN try {
#define ZFV_API_END                     \
  }                                     \
  catch (const std::exception &e) {     \
    return fail(e);                     \
  }

static void check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw ZfvError(ZFV_ERR_NO_DEVICE, "no HIP device visible");
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    throw ZfvError(ZFV_ERR_NO_DEVICE, std::string("libzfv_hip is built for gfx950; device is ") + prop.gcnArchName);
}

static StageCoeffs ssp3_coeffs(int stage) {
  // next-state coefficients after stage s: a[s+1][:] and b for the last stage
  // (src/zisa/ode/runge_kutta.cpp:154-165)
  StageCoeffs c{};
  if (stage == 0) {
    c.c[0] = 1.0;
    c.nk = 1;
  } else if (stage == 1) {
    c.c[0] = 0.25;
    c.c[1] = 0.25;
    c.nk = 2;
  } else {
    c.c[0] = 1.0 / 6;
    c.c[1] = 1.0 / 6;
    c.c[2] = 2.0 / 3;
    c.nk = 3;
  }
  c.write_rate_only = 0;
  return c;
}

// the image with stateful equilibria switched off: its launches recompute
// every cell's equilibrium and leave the kept ones and their counters alone
static Img stateless(const Img &m) {
  Img r = m;
  r.eqc = nullptr;
  r.eqcnt = nullptr;
  return r;
}

// steps run under the frozen BC once a steady state is set
static int frozen(const zfv_handle *h) { return h->dimg.steady != nullptr ? 1 : 0; }

// Autotuner of the reconstruction launch configuration: every candidate (1,
// 2, 4 waves per block) runs the instance's own state through stages as the
// steps do (reconstruction, face fluxes, a rate-only update into K[0]), one
// warm-up and four timed stages, the reconstruction alone between HIP events
// on the handle's stream; the fastest average wins.  (Launched back to back
// the reconstruction finds the caches as the previous launch left them and
// runs up to 6% faster than inside a step, enough to reorder close
// candidates.)  The launches write only scratch buffers: face-point values,
// source, face fluxes and K[0].
static void tune_rec(zfv_handle *h, const double *Ustate) {
  const Img m = stateless(h->dimg);
  constexpr int kTimed = 4;
  hipEvent_t ev[2 * kTimed];
  for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
  StageCoeffs rate_only{};
  rate_only.write_rate_only = 1;
  // two rounds over the configurations, the faster round counting for each:
  // the first launches after an idle spell run at lower clocks
  for (int round = 0; round < 2; ++round)
    for (int c = 0; c < kRecConfigs; ++c) {
      for (int k = -1; k < kTimed; ++k) {
        if (k >= 0) HIP_CHECK(hipEventRecord(ev[2 * k], h->stream));
        h->L.rec[h->ph](m, Ustate, h->fpv, h->src, c, frozen(h), h->stream);
        if (k >= 0) HIP_CHECK(hipEventRecord(ev[2 * k + 1], h->stream));
        h->L.face(m, h->fpv, h->G, h->stream);
        h->L.update(m, rate_only, 0, 0.0, nullptr, Ustate, h->G, h->src, Ustate, h->K, nullptr, h->stream);
      }
      HIP_CHECK(hipEventSynchronize(ev[2 * kTimed - 1]));
      HIP_CHECK(hipGetLastError());
      double tot = 0.0;
      for (int k = 0; k < kTimed; ++k) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        tot += ms;
      }
      h->rec_ms[c] = round == 0 ? tot / kTimed : std::min(h->rec_ms[c], tot / kTimed);
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
  // the smallest block unless a larger one is clearly (2%) faster: near-ties
  // are clock noise, and a fixed choice keeps runs comparable
  int best = 0;
  for (int c = 1; c < kRecConfigs; ++c)
    if (h->rec_ms[c] < 0.98 * h->rec_ms[best]) best = c;
  h->rec_cfg = best;
}

// reconstruction launch configuration, tuning on first use (never inside a capture)
static int rec_config(zfv_handle *h, const double *Ustate) {
  if (h->rec_cfg < 0) tune_rec(h, Ustate);
  return h->rec_cfg;
}

static void enqueue_step(zfv_handle *h, double dt) {
  const Img &m = h->dimg;
  const double *Us = h->U;
  for (int s = 0; s < 3; ++s) {
    h->L.rec[h->ph](m, Us, h->fpv, h->src, h->rec_cfg, frozen(h), h->stream);
    h->L.face(m, h->fpv, h->G, h->stream);
    double *Un = (s == 2) ? h->U : h->X;
    h->L.update(m, ssp3_coeffs(s), s, dt, nullptr, Us, h->G, h->src, h->U, h->K, Un, h->stream);
    Us = h->X;
  }
}

// captured step graphs, rebuilt on next use
static void drop_graphs(zfv_handle *h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  for (hipGraphExec_t &g : h->adv_graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  h->graph_exec = nullptr;
}

// one step with the dt the device chooses (LocalCFL, clamped to t_end);
// with reject, the step is undone when the density changes too much.  adv is
// the loop state and Ubk the backup of u0 (used only with reject).
static void enqueue_adaptive_step(zfv_handle *h, bool reject, int cfg, AdvState *adv, double *Ubk) {
  const Img &m = h->dimg;
  const int64_t count = int64_t(h->img.s.nvars) * m.n;
  if (reject)
    hipLaunchKernelGGL(k_adv_backup, dim3(blocks_for(count, 256)), dim3(256), 0, h->stream, adv, count, h->U, Ubk);
  h->L.cfl(m, h->U, &adv->dt_bits, h->stream);
  hipLaunchKernelGGL(k_adv_choose, dim3(1), dim3(1), 0, h->stream, adv);
  const double *Us = h->U;
  for (int s = 0; s < 3; ++s) {
    h->L.rec[h->ph](m, Us, h->fpv, h->src, cfg, frozen(h), h->stream);
    h->L.face(m, h->fpv, h->G, h->stream);
    double *Un = (s == 2) ? h->U : h->X;
    h->L.update(m, ssp3_coeffs(s), s, 0.0, &adv->dt, Us, h->G, h->src, h->U, h->K, Un, h->stream);
    Us = h->X;
  }
  if (reject) {
    // density is the first SoA plane of U
    hipLaunchKernelGGL(k_adv_check, dim3(blocks_for(m.n, 256)), dim3(256), 0, h->stream, adv, m.n, h->U, Ubk);
    hipLaunchKernelGGL(k_adv_restore, dim3(blocks_for(count, 256)), dim3(256), 0, h->stream, adv, count, h->U, Ubk);
  }
  hipLaunchKernelGGL(k_adv_end, dim3(1), dim3(1), 0, h->stream, adv);
}

// The loop state of a device time loop starting at t0 (zfv_advance and its
// allocation-free variant fill it the same way)
static AdvState adv_start(double cfl_number, double t0, double t_end, int32_t max_steps, double drho_crit_rel,
                          int32_t max_exponent) {
  AdvState a{};
  a.t = t0;
  a.t_end = t_end;
  a.cfl = cfl_number;
  a.steps = 0;
  a.max_steps = max_steps > 0 ? max_steps : 0x7fffffff;
  a.active = t0 < t_end ? 1 : 0;
  a.drho_crit = drho_crit_rel > 0.0 ? drho_crit_rel : 0.0;
  a.factor = 1.0;
  a.min_factor = drho_crit_rel > 0.0 ? std::pow(2.0, -double(max_exponent)) : 1.0;
  a.prev_good = 1;
  a.dt_bits = adv_no_bound();
  a.bad = 0;
  return a;
}

static void check_advance_args(double cfl_number, double drho_crit_rel, int32_t max_exponent) {
  if (!(cfl_number > 0.0)) throw ZfvError(ZFV_ERR_INVALID, "cfl_number must be > 0");
  if (drho_crit_rel > 0.0 && max_exponent < 0) throw ZfvError(ZFV_ERR_INVALID, "max_exponent must be >= 0");
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free device time loop (zfv_advance_noalloc): the
// loop state (one AdvState) in the first kNoAllocStateBytes, then the backup
// of u0 that step rejection restores from.  A file-scope device array: each
// GPU the library runs on holds its own copy from module load until the
// process ends, and every call on that GPU shares it.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr size_t kNoAllocStateBytes = 256;
static_assert(sizeof(AdvState) <= kNoAllocStateBytes, "loop state larger than its scratch slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on device `dev`, which must be current
// (looked up once per device; the array's name is only a host-side stand-in)
static unsigned char *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw ZfvError(ZFV_ERR_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ZfvError(ZFV_ERR_HIP, "no device address for the time-loop scratch");
    g_scratch_dev[dev] = p;
  }
  return static_cast<unsigned char *>(g_scratch_dev[dev]);
}

static void to_soa(const zfv_handle *h, const double *aos, double *soa) { zfv::to_soa(h->img, aos, soa); }

static void from_soa(const zfv_handle *h, const double *soa, double *aos, bool accumulate) {
  zfv::from_soa(h->img, soa, aos, accumulate);
}

// launch configuration named by its threads per block
static int rec_cfg_of(int32_t rec_threads) {
  if (rec_threads != 64 && rec_threads != 128 && rec_threads != 256)
    throw ZfvError(ZFV_ERR_INVALID, "rec_threads must be 0, -1, 64, 128 or 256");
  return rec_threads == 64 ? 0 : (rec_threads == 128 ? 1 : 2);
}

static void report_tuning(int cfg, int groups_per_wave, const double *ms, zfv_tuning *out) {
  if (!out) return;
  const int c = cfg < 0 ? kRecConfigs - 1 : cfg;
  out->rec_threads = 64 * rec_cfg_waves(c);
  out->cells_per_block = rec_cfg_waves(c) * groups_per_wave;
  out->tuned = cfg >= 0;
  for (int k = 0; k < kRecConfigs; ++k) out->rec_ms[k] = ms[k];
}

}  // namespace zfv

using namespace zfv;

extern "C" {

const char *zfv_last_error(void) { return g_last_error.c_str(); }

int zfv_create(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out) {
  if (!out) {
    g_last_error = "null argument";
    return ZFV_ERR_INVALID;
  }
  zfv_handle *h = nullptr;
  int rc = zfv::create_host(mesh, params, &h);
  if (rc == ZFV_OK) rc = zfv::create_device(h);
  if (rc != ZFV_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return ZFV_OK;
}

}  // extern "C"

namespace zfv {

// the host half of zfv_create: the setup image (stencils, pseudo-inverses,
// tables), no device call
int create_host(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out) {
  ZFV_API_BEGIN
  if (!mesh || !params || !out) throw ZfvError(ZFV_ERR_INVALID, "null argument");
  auto h = std::make_unique<zfv_handle>();
  h->params = *params;
  build_image(*mesh, *params, h->img);
  *out = h.release();
  return ZFV_OK;
  ZFV_API_END
}

// the device half: stream, allocations, uploads and the setup kernels
int create_device(zfv_handle *h) {
  ZFV_API_BEGIN
  check_device();
  const zfv_params *params = &h->params;
  const HostImage &im = h->img;
  const Shape &s = im.s;
  h->L = pick_launcher(s);
  HIP_CHECK(hipGetDevice(&h->dev));
  HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));

  Img &m = h->dimg;
  m.n = im.n;
  m.ne = im.ne;
  m.P = h->upload(im.P);
  m.rows = h->upload(im.rows);
  m.twin = h->upload(im.twin);
  m.crn = h->upload(im.crn);
  m.xc = h->upload(im.xc);
  m.len = h->upload(im.len);
  m.vol = h->upload(im.vol);
  m.inr = h->upload(im.inr);
  m.mom = h->upload(im.mom);
  m.meta = h->upload(im.meta);
  m.cface = h->upload(im.cface);
  m.fslot = h->upload(im.fslot);
  m.fperm = h->upload(im.fperm);
  m.steady = nullptr;
  m.phim = nullptr;
  m.pot = nullptr;
  m.frm = nullptr;
  m.gamma = params->gamma;
  m.eps = params->epsilon;
  m.exponent = params->exponent;
  for (int k = 0; k < 5; ++k) m.lw[k] = im.linear_weights[k];
  m.mode = params->reconstruction;
  m.scaling = params->scaling;
  m.vol_mean = volume_mean_is_constant_term(*params, s) ? 1 : 0;
  m.flux = params->flux;
  m.eq = params->equilibrium;
  m.gravity = params->gravity;
  m.g = params->gravity_g;
  for (int d = 0; d < 3; ++d) m.axis[d] = params->gravity_axis[d];
  m.rhoC = params->polytrope_rhoC;
  m.K = params->polytrope_K;
  m.G = params->polytrope_G;
  m.flux_bc = params->flux_bc;
  m.spr = params->steps_per_recompute;
  m.eq_threshold = params->recompute_threshold;
  for (int q = 0; q < 16; ++q)
    for (int d = 0; d < 3; ++d) m.shift[q][d] = h->img.shift[3 * q + d];
  if (im.face_w.size() > 8 || im.vol_w.size() > 16) throw ZfvError(ZFV_ERR_INVALID, "quadrature rule too large");
  const int fdim = s.dim == 2 ? 1 : 3;
  for (size_t q = 0; q < im.face_w.size(); ++q) {
    m.fw[q] = im.face_w[q];
    for (int d = 0; d < fdim; ++d) m.fx[q][d] = im.face_x[q * fdim + d];
  }
  m.qv = int(im.vol_w.size());
  for (size_t q = 0; q < im.vol_w.size(); ++q) {
    m.vw[q] = im.vol_w[q];
    for (int d = 0; d < s.dim + 1; ++d) m.vx[q][d] = im.vol_x[q * (s.dim + 1) + d];
  }

  const int64_t n = im.n;
  const int nv = s.nvars;
  h->U = h->alloc<double>(size_t(nv) * n);
  h->X = h->alloc<double>(size_t(nv) * n);
  h->K = h->alloc<double>(size_t(3) * nv * n);
  h->fpv = h->alloc<double>(size_t(s.nfaces) * ((s.q * nv + 1) & ~1) * n);  // Q x NV per cell-face slot, even
  h->G = h->alloc<double>(size_t(nv) * std::max<int64_t>(im.ne, 1));
  if (params->gravity != ZFV_GRAVITY_NONE) h->src = h->alloc<double>(size_t(nv) * n);
  h->dt_bits = h->alloc<unsigned long long>(1);
  if (params->equilibrium == ZFV_EQ_ISENTROPIC && params->gravity != ZFV_GRAVITY_NONE && params->gamma == 2.0) {
    double *pm = h->alloc<double>(size_t(2) * n);
    if (s.dim == 2)
      hipLaunchKernelGGL(k_phi_moments<2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, pm);
    else
      hipLaunchKernelGGL(k_phi_moments<3>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, pm);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(h->stream));
    m.phim = pm;
  }
  {
    // face frames of every cell's local faces
    double *tab = h->alloc<double>(size_t(s.dim == 2 ? 3 * frame_words<2>() : 4 * frame_words<3>()) * n);
    if (s.dim == 2)
      hipLaunchKernelGGL(k_face_frames<2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, tab);
    else
      hipLaunchKernelGGL(k_face_frames<3>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, tab);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(h->stream));
    m.frm = tab;
  }
  {
    // the interior faces' frames in face order, for the face kernel
    double *tab = h->alloc<double>(size_t(s.dim == 2 ? frame_words<2>() : frame_words<3>()) * std::max<int64_t>(im.ne, 1));
    if (im.ne > 0) {
      if (s.dim == 2)
        hipLaunchKernelGGL(k_face_table<2>, dim3(blocks_for(im.ne, 256)), dim3(256), 0, h->stream, m, tab);
      else
        hipLaunchKernelGGL(k_face_table<3>, dim3(blocks_for(im.ne, 256)), dim3(256), 0, h->stream, m, tab);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    m.ffr = tab;
  }
  h->ph = physics_set(m);
  if (params->gravity != ZFV_GRAVITY_NONE) {
    double *tab = h->alloc<double>(size_t(s.dofc) * s.dim * n);
    h->L.source_moments(m, tab, h->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(h->stream));
    m.smom = tab;
  }
  for (int c = 0; c < 16; ++c) m.shift_phi[c] = s.dim == 2 ? phi<2, PH_LIGHT>(m, m.shift[c]) : phi<3, PH_LIGHT>(m, m.shift[c]);
  if (h->ph == PH_LIGHT && m.eq != 0) {
    // member equilibrium averages read the members' corner potentials
    double *tab = h->alloc<double>(size_t(s.dim + 1) * n);
    if (s.dim == 2)
      hipLaunchKernelGGL(k_corner_phi<2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, tab);
    else
      hipLaunchKernelGGL(k_corner_phi<3>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, tab);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(h->stream));
    m.cphi = tab;
  }
  if (h->ph != PH_LIGHT && params->gravity != ZFV_GRAVITY_NONE) {
    // potential tables of the cells' own quadrature points
    double *tab = h->alloc<double>(size_t(s.dim == 2 ? pot_rows<2>(m.qv, s.q) : pot_rows<3>(m.qv, s.q)) * n);
    if (s.dim == 2)
      hipLaunchKernelGGL(k_potential_points<2>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, s.q, tab);
    else
      hipLaunchKernelGGL(k_potential_points<3>, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, m, s.q, tab);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(h->stream));
    m.pot = tab;
  }
  if (params->steps_per_recompute > 1) {
    // kept equilibria and their call counters (zero: the first call recomputes)
    m.eqc = h->alloc<double>(size_t(8) * n);
    m.eqcnt = h->alloc<int32_t>(size_t(n));
    HIP_CHECK(hipMemset(m.eqcnt, 0, sizeof(int32_t) * size_t(n)));
  }
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&h->host_stage), sizeof(double) * size_t(nv) * n));
  return ZFV_OK;
  ZFV_API_END
}

}  // namespace zfv

extern "C" {

int zfv_destroy(zfv_handle *h) {
  if (!h) return ZFV_OK;
  int rc = ZFV_OK;
  if (h->batch) {
    // a member leaves its batch first: the batch must not run on freed arrays
    try {
      batch_forget(h->batch, h->batch_slot);
    } catch (const std::exception &e) {
      rc = fail(e);
    }
  }
  delete h;
  return rc;
}

int zfv_set_state(zfv_handle *h, const double *cvars) {
  ZFV_API_BEGIN
  member_settle(h);
  to_soa(h, cvars, h->host_stage);
  const size_t bytes = sizeof(double) * h->img.s.nvars * h->img.n;
  HIP_CHECK(hipMemcpyAsync(h->U, h->host_stage, bytes, hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_get_state(zfv_handle *h, double *cvars) {
  ZFV_API_BEGIN
  member_settle(h);
  const size_t bytes = sizeof(double) * h->img.s.nvars * h->img.n;
  HIP_CHECK(hipMemcpyAsync(h->host_stage, h->U, bytes, hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  from_soa(h, h->host_stage, cvars, false);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_snapshot_begin(zfv_handle *h) {
  ZFV_API_BEGIN
  member_settle(h);
  if (h->snap_pending) throw ZfvError(ZFV_ERR_INVALID, "a snapshot is pending: call zfv_snapshot_end first");
  const size_t count = size_t(h->img.s.nvars) * h->img.n;
  if (!h->snap_dev) {
    h->snap_dev = h->alloc<double>(count);
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&h->snap_host), sizeof(double) * count));
    HIP_CHECK(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&h->snap_ready, hipEventDisableTiming));
  }
  // the state as of the work enqueued so far, copied on the device in stream
  // order; the copy to the host runs on its own stream behind later steps
  HIP_CHECK(hipMemcpyAsync(h->snap_dev, h->U, sizeof(double) * count, hipMemcpyDeviceToDevice, h->stream));
  HIP_CHECK(hipEventRecord(h->snap_ready, h->stream));
  HIP_CHECK(hipStreamWaitEvent(h->copy_stream, h->snap_ready, 0));
  // a batch member's later steps run on the batch's stream: after the copy
  if (h->batch) batch_after_snapshot(h->batch, h->snap_ready);
  HIP_CHECK(hipMemcpyAsync(h->snap_host, h->snap_dev, sizeof(double) * count, hipMemcpyDeviceToHost,
                           h->copy_stream));
  h->snap_pending = true;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_snapshot_end(zfv_handle *h, double *cvars) {
  ZFV_API_BEGIN
  if (!h->snap_pending) throw ZfvError(ZFV_ERR_INVALID, "no snapshot pending");
  h->snap_pending = false;
  HIP_CHECK(hipStreamSynchronize(h->copy_stream));
  from_soa(h, h->snap_host, cvars, false);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_set_steady_state(zfv_handle *h, const double *cvars) {
  ZFV_API_BEGIN
  member_settle(h);
  const int64_t n = h->img.n;
  const int nv = h->img.s.nvars;
  if (!h->steady) h->steady = h->alloc<double>(size_t(nv) * n);
  to_soa(h, cvars, h->host_stage);
  HIP_CHECK(hipMemcpyAsync(h->steady, h->host_stage, sizeof(double) * nv * n, hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  h->dimg.steady = h->steady;
  drop_graphs(h);  // the captured steps hold the BC and the skip flag
  if (h->batch) batch_member_steady(h->batch, h->batch_slot, h->steady);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_rate(zfv_handle *h, const double *u, double *dudt, double t) {
  (void)t;
  ZFV_API_BEGIN
  member_settle(h);
  const int64_t n = h->img.n;
  const int nv = h->img.s.nvars;
  to_soa(h, u, h->host_stage);
  HIP_CHECK(hipMemcpyAsync(h->X, h->host_stage, sizeof(double) * nv * n, hipMemcpyHostToDevice, h->stream));
  const Img &m = h->dimg;
  h->L.rec[h->ph](m, h->X, h->fpv, h->src, rec_config(h, h->X), 0, h->stream);
  h->L.face(m, h->fpv, h->G, h->stream);
  StageCoeffs sc{};
  sc.write_rate_only = 1;
  h->L.update(m, sc, 0, 0.0, nullptr, h->X, h->G, h->src, h->X, h->K, h->X, h->stream);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h->host_stage, h->K, sizeof(double) * nv * n, hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  from_soa(h, h->host_stage, dudt, true);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_step(zfv_handle *h, double t, double dt) {
  (void)t;
  ZFV_API_BEGIN
  member_settle(h);
  rec_config(h, h->U);
  enqueue_step(h, dt);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_steps(zfv_handle *h, int32_t n_steps, double t, double dt) {
  (void)t;
  ZFV_API_BEGIN
  member_settle(h);
  if (n_steps < 0) throw ZfvError(ZFV_ERR_INVALID, "n_steps < 0");
  rec_config(h, h->U);
  if (!h->graph_exec || h->graph_dt != dt) {
    if (h->graph_exec) {
      HIP_CHECK(hipGraphExecDestroy(h->graph_exec));
      h->graph_exec = nullptr;
    }
    hipGraph_t graph;
    HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(h, dt);
    HIP_CHECK(hipStreamEndCapture(h->stream, &graph));
    HIP_CHECK(hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
    h->graph_dt = dt;
  }
  for (int k = 0; k < n_steps; ++k) HIP_CHECK(hipGraphLaunch(h->graph_exec, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_advance(zfv_handle *h, double cfl_number, double t0, double t_end, int32_t max_steps, int32_t check_every,
                double drho_crit_rel, int32_t max_exponent, double *t_final, int32_t *steps_done,
                int32_t *rejected) {
  ZFV_API_BEGIN
  member_settle(h);
  check_advance_args(cfl_number, drho_crit_rel, max_exponent);
  if (check_every < 1) check_every = 16;
  rec_config(h, h->U);
  if (!h->adv) h->adv = h->alloc<AdvState>(1);
  const bool reject = drho_crit_rel > 0.0;
  if (reject && !h->Ubk) h->Ubk = h->alloc<double>(size_t(h->img.s.nvars) * h->img.n);
  hipGraphExec_t &step = h->adv_graph[reject ? 1 : 0];
  if (!step) {
    hipGraph_t graph;
    HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_adaptive_step(h, reject, h->rec_cfg, h->adv, h->Ubk);
    HIP_CHECK(hipStreamEndCapture(h->stream, &graph));
    HIP_CHECK(hipGraphInstantiate(&step, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
  }
  AdvState a = adv_start(cfl_number, t0, t_end, max_steps, drho_crit_rel, max_exponent);
  HIP_CHECK(hipMemcpyAsync(h->adv, &a, sizeof(a), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  while (a.active) {
    for (int k = 0; k < check_every; ++k) HIP_CHECK(hipGraphLaunch(step, h->stream));
    HIP_CHECK(hipMemcpyAsync(&a, h->adv, sizeof(a), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(hipStreamSynchronize(h->stream));
  }
  if (t_final) *t_final = a.t;
  if (steps_done) *steps_done = a.steps;
  if (rejected) *rejected = a.rejections;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_advance_noalloc(zfv_handle *h, double cfl_number, double t0, double t_end, int32_t max_steps,
                        int32_t check_every, double drho_crit_rel, int32_t max_exponent, double *t_final,
                        int32_t *steps_done, int32_t *rejected) {
  ZFV_API_BEGIN
  member_settle(h);
  check_advance_args(cfl_number, drho_crit_rel, max_exponent);
  if (check_every < 1) check_every = 16;
  int dev = -1;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev != h->dev) throw ZfvError(ZFV_ERR_INVALID, "zfv_advance_noalloc: make the handle's device current first");
  const bool reject = drho_crit_rel > 0.0;
  const size_t backup = sizeof(double) * size_t(h->img.s.nvars) * size_t(h->img.n);
  if (reject && backup > kNoAllocScratchBytes - kNoAllocStateBytes)
    throw ZfvError(ZFV_ERR_MEMORY, "zfv_advance_noalloc: step rejection backs up the state in " +
                                       std::to_string((kNoAllocScratchBytes - kNoAllocStateBytes) >> 20) +
                                       " MiB of scratch; this instance needs " + std::to_string(backup) +
                                       " bytes (use zfv_advance)");
  unsigned char *sp = scratch_on(dev);
  AdvState *adv = reinterpret_cast<AdvState *>(sp);
  double *Ubk = reinterpret_cast<double *>(sp + kNoAllocStateBytes);
  // no tuning here (it creates events): the tuned launch configuration when
  // there is one, else one wavefront per block; results do not depend on it
  const int cfg = h->rec_cfg >= 0 ? h->rec_cfg : 0;
  // the whole loop state is written before the first step reads any of it
  AdvState a = adv_start(cfl_number, t0, t_end, max_steps, drho_crit_rel, max_exponent);
  HIP_CHECK(hipMemcpyAsync(adv, &a, sizeof(a), hipMemcpyHostToDevice, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  while (a.active) {
    for (int k = 0; k < check_every; ++k) enqueue_adaptive_step(h, reject, cfg, adv, Ubk);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(&a, adv, sizeof(a), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(hipStreamSynchronize(h->stream));
  }
  if (t_final) *t_final = a.t;
  if (steps_done) *steps_done = a.steps;
  if (rejected) *rejected = a.rejections;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_cfl_dt(zfv_handle *h, double cfl_number, double *dt) {
  ZFV_API_BEGIN
  const double big = 1.0e300;
  unsigned long long bits;
  std::memcpy(&bits, &big, sizeof(bits));
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, h->stream, h->dt_bits, bits);
  h->L.cfl(h->dimg, h->U, h->dt_bits, h->stream);
  HIP_CHECK(hipGetLastError());
  unsigned long long r = 0;
  HIP_CHECK(hipMemcpyAsync(&r, h->dt_bits, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_CHECK(hipStreamSynchronize(h->stream));
  double m;
  std::memcpy(&m, &r, sizeof(m));
  *dt = cfl_number * m;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_synchronize(zfv_handle *h) {
  ZFV_API_BEGIN
  HIP_CHECK(hipStreamSynchronize(h->stream));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_get_info(zfv_handle *h, zfv_info *info) {
  ZFV_API_BEGIN
  const Shape &s = h->img.s;
  const double n = double(h->img.n), ne = double(h->img.ne);
  info->n_cells = h->img.n;
  info->n_faces = h->img.ne;
  info->device_bytes = h->device_bytes;
  // algorithmic bytes of the reconstruction kernel per cell: pseudo-inverse,
  // stencil indices, own state, geometry (corners, centre, length, volume,
  // moments), meta/face words, the potential tables and moments where the
  // instance has them, the face-point values written once per interior face
  // side and the source term written (with gravity)
  const double nv = s.nvars;
  const double frame = s.dim == 2 ? frame_words<2>() : frame_words<3>();
  const Img &dm = h->dimg;
  // tables: the potential at the cell's own volume and face points (its
  // gradient rows are not read: the source moments hold the volume terms),
  // the potential moments, the source moments (read outside the light set)
  const double tables = (dm.pot ? 8.0 * (dm.qv + s.nfaces * s.q) : 0.0) + (dm.phim ? 16.0 : 0.0) +
                        (dm.smom && h->ph != PH_LIGHT ? 8.0 * s.dofc * s.dim : 0.0);
  const double rec = 8.0 * s.m + 4.0 * s.nrows + 8.0 * nv + 8.0 * (s.nfaces * s.dim + s.dim + 2 + s.nmom) +
                     4.0 * (1 + s.nfaces + 1) + tables +
                     8.0 * (2.0 * ne / n) * s.q * nv + (h->src ? 8.0 * nv : 0.0);
  const double face = (ne / n) * (8.0 * 2 * s.q * nv + 8.0 * nv + 4.0 * 2 + 8.0 * frame);
  const double upd = 4.0 * s.nfaces + 8.0 * s.nfaces * nv + 8.0 + 4.0 + 8.0 * nv * 5;
  info->recon_bytes_per_cell = rec;
  info->bytes_per_cell_stage = rec + face + upd;
  int maxs = 0;
  for (size_t i = 0; i + 1 < h->img.l2g_off.size(); ++i)
    maxs = std::max(maxs, h->img.l2g_off[i + 1] - h->img.l2g_off[i]);
  info->max_stencil_size = maxs;
  info->n_points_face = s.q;
  int64_t far = 0;
  for (int32_t v : h->img.meta) far += (v & zfv::META_FAR_GHOST) == zfv::META_FAR_GHOST;
  info->n_far_ghost_cells = far;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_tune(zfv_handle *h, int32_t rec_threads, zfv_tuning *out) {
  ZFV_API_BEGIN
  if (rec_threads == 0) {
    tune_rec(h, h->U);
  } else if (rec_threads > 0) {
    h->rec_cfg = rec_cfg_of(rec_threads);
  }
  if (rec_threads >= 0) drop_graphs(h);  // the captured steps hold the old configuration
  report_tuning(h->rec_cfg, h->L.groups_per_wave, h->rec_ms, out);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_get_stencils(zfv_handle *h, int32_t i, int32_t *l2g, int32_t cap, int32_t *l2g_size, int32_t *members,
                     int32_t members_cap) {
  ZFV_API_BEGIN
  const HostImage &im = h->img;
  if (i < 0 || i >= im.n) throw ZfvError(ZFV_ERR_INVALID, "cell index out of range");
  const int a = im.l2g_off[i], b = im.l2g_off[i + 1];
  *l2g_size = b - a;
  for (int q = a; q < b && q - a < cap; ++q) l2g[q - a] = im.l2g[q];
  const int ma = im.mem_off[i], mb = im.mem_off[i + 1];
  if (mb - ma > members_cap) throw ZfvError(ZFV_ERR_INVALID, "members buffer too small");
  for (int q = ma; q < mb; ++q) members[q - ma] = im.members[q];
  return mb - ma;
  ZFV_API_END
}

int zfv_time_kernels(zfv_handle *h, int32_t n_steps, double dt, zfv_kernel_times *out) {
  ZFV_API_BEGIN
  const Img m = stateless(h->dimg);
  const int cfg = rec_config(h, h->U);
  std::vector<hipEvent_t> ev(2 * 9 * std::max(n_steps, 1));
  for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
  int k = 0;
  for (int st = 0; st < n_steps; ++st) {
    const double *Us = h->U;
    for (int s = 0; s < 3; ++s) {
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      h->L.rec[h->ph](m, Us, h->fpv, h->src, cfg, frozen(h), h->stream);
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      h->L.face(m, h->fpv, h->G, h->stream);
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      double *Un = (s == 2) ? h->U : h->X;
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      h->L.update(m, ssp3_coeffs(s), s, dt, nullptr, Us, h->G, h->src, h->U, h->K, Un, h->stream);
      HIP_CHECK(hipEventRecord(ev[k++], h->stream));
      Us = h->X;
    }
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(h->stream));
  zfv_kernel_times t{};
  for (int q = 0; q < k; q += 6) {
    float a = 0, b = 0, c = 0;
    HIP_CHECK(hipEventElapsedTime(&a, ev[q], ev[q + 1]));
    HIP_CHECK(hipEventElapsedTime(&b, ev[q + 2], ev[q + 3]));
    HIP_CHECK(hipEventElapsedTime(&c, ev[q + 4], ev[q + 5]));
    t.reconstruct_ms += a;
    t.face_flux_ms += b;
    t.update_ms += c;
    t.n_launches += 1;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  *out = t;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_estimate_device_bytes(const zfv_mesh *mesh, const zfv_params *params, int64_t *bytes) {
  ZFV_API_BEGIN
  const Shape s = shape_for(*params);
  const double n = mesh->n_cells;
  const double ne = n * (s.nfaces / 2.0);  // interior faces <= cells * faces / 2
  const double nv = s.nvars;
  const double frame = s.dim == 2 ? frame_words<2>() : frame_words<3>();
  double per_cell = 8.0 * s.m + 4.0 * s.nrows + 1.0 * (s.nrows - s.rc) +
                    8.0 * ((s.dim == 2 ? 6 : 16) + s.dim + 3 + s.nmom) + 4.0 * (2 + 4) +
                    8.0 * s.nfaces * frame +    // face frames
                    8.0 * nv * (1 + 1 + 3);     // U, X, K[3]
  if (params->gravity != ZFV_GRAVITY_NONE) {
    per_cell += 8.0 * nv;                       // source
    per_cell += 8.0 * 2;                        // potential moments
    per_cell += 8.0 * (16 * (1 + s.dim) + s.nfaces * s.q);  // potential tables (at most)
  }
  if (mesh->ghost_mask) per_cell += 8.0 * nv;                          // steady state
  if (params->steps_per_recompute > 1) per_cell += 8.0 * 8 + 4.0;       // kept equilibria
  per_cell += 8.0 * s.nfaces * ((s.q * s.nvars + 1) & ~1);            // face-point values per cell-face slot
  if (params->equilibrium != ZFV_EQ_NONE) per_cell += 8.0 * s.nfaces;  // corner potentials (light set)
  if (params->gravity != ZFV_GRAVITY_NONE) per_cell += 8.0 * s.dofc * s.dim;  // source moments
  double per_face = 8.0 * nv + 4.0 * 2 + 8.0 * frame;  // fluxes, the two slots, the frame and area
  *bytes = int64_t(n * per_cell + ne * per_face) + (1 << 20);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_bandwidth_probe(int64_t bytes, int32_t iters, double *gbytes_per_s) {
  ZFV_API_BEGIN
  check_device();
  if (bytes < 16 || iters < 1) throw ZfvError(ZFV_ERR_INVALID, "bandwidth probe: bad size");
  const int64_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  HIP_CHECK(hipMalloc(&a, n * 16));
  HIP_CHECK(hipMalloc(&b, n * 16));
  hipStream_t s;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipMemsetAsync(a, 0, n * 16, s));
  const unsigned grid = 256 * 8 * 4;  // a few waves per CU on every XCD
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
  HIP_CHECK(hipEventRecord(e0, s));
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
  HIP_CHECK(hipEventRecord(e1, s));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *gbytes_per_s = 2.0 * double(n) * 16.0 * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_mesh_lattice2d(int32_t n, double jitter, uint64_t seed, int32_t periodic_x, int32_t periodic_y,
                       int32_t ghost_layers, zfv_mesh_buf **out) {
  ZFV_API_BEGIN
  *out = reinterpret_cast<zfv_mesh_buf *>(mesh_lattice2d(n, jitter, seed, periodic_x != 0, periodic_y != 0, ghost_layers));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_mesh_ball3d(int32_t n, double radius, double ghost_radius, double jitter, uint64_t seed, zfv_mesh_buf **out) {
  ZFV_API_BEGIN
  *out = reinterpret_cast<zfv_mesh_buf *>(mesh_ball3d(n, radius, ghost_radius, jitter, seed));
  return ZFV_OK;
  ZFV_API_END
}

int zfv_mesh_view(const zfv_mesh_buf *buf, zfv_mesh *view) {
  const MeshBuf *m = reinterpret_cast<const MeshBuf *>(buf);
  view->n_dims = m->n_dims;
  view->n_cells = m->n_cells;
  view->vertex_ids = m->vertex_ids.data();
  view->corners = m->corners.data();
  view->ghost_mask = m->ghost.data();
  return ZFV_OK;
}

int zfv_mesh_free(zfv_mesh_buf *buf) {
  delete reinterpret_cast<MeshBuf *>(buf);
  return ZFV_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches

struct zfv_batch {
  zfv::Launcher L{};
  zfv::Shape shape{};
  hipStream_t stream = nullptr;
  int max_inst = 0;
  int64_t cap_cell_blocks = 0, cap_face_blocks = 0, cap_rec_blocks = 0;
  int64_t n_cell_blocks = 0, n_face_blocks = 0, n_rec_blocks = 0;  // current launch sizes
  bool xcd_chunks = true;
  int rec_cfg = -1;   // reconstruction launch configuration; -1 = not tuned yet
  double rec_ms[zfv::kRecConfigs] = {0, 0, 0};
  int ph = 0;         // physics set: the widest any member needs
  zfv::DevInst *d_inst = nullptr;
  zfv::BlockRef *d_bcell = nullptr, *d_bface = nullptr, *d_brec = nullptr;
  std::vector<zfv::DevInst> h_inst;
  std::vector<zfv_handle *> members;
  hipGraphExec_t graph = nullptr;
  bool dirty = true;
  bool in_flight = false;  // steps launched (zfv_batch_launch) and not yet waited for
  ~zfv_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (zfv_handle *h : members)
      if (h) h->batch = nullptr;
    if (graph) (void)hipGraphExecDestroy(graph);
    if (d_inst) (void)hipFree(d_inst);
    if (d_bcell) (void)hipFree(d_bcell);
    if (d_bface) (void)hipFree(d_bface);
    if (d_brec) (void)hipFree(d_brec);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace zfv {

static bool same_shape(const Shape &a, const Shape &b) {
  return a.dim == b.dim && a.dofc == b.dofc && a.rc == b.rc && a.nb == b.nb && a.rb == b.rb && a.q == b.q;
}

// Block table: logical blocks of every active instance, in slot order, laid
// out so that hardware block b (dispatched to XCD b mod 8) serves the
// (b mod 8)-th contiguous eighth of the logical blocks when xcd_chunks is set:
// neighbouring cells then share one XCD's L2.
// The table has as many entries (a multiple of 8) as the members need; the
// step graph is launched with exactly that many blocks.
// a block table of at most cap entries, or ZFV_ERR_MEMORY
static std::vector<BlockRef> checked_table(const std::vector<int64_t> &counts, int64_t per_block, int64_t cap,
                                           bool xcd_chunks) {
  auto t = block_table(counts, per_block, cap, xcd_chunks);
  if (t.empty()) throw ZfvError(ZFV_ERR_MEMORY, "batch capacity exceeded");
  return t;
}

// waits for the batch's stream (and so for steps in flight) and refreshes
// the host copy of the members' loop state
static void batch_sync_down(zfv_batch *b) {
  HIP_CHECK(hipMemcpyAsync(b->h_inst.data(), b->d_inst, sizeof(DevInst) * b->max_inst, hipMemcpyDeviceToHost,
                           b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  b->in_flight = false;
}

// a member's own calls wait for its batch's steps in flight
static void member_settle(zfv_handle *h) {
  if (h->batch && h->batch->in_flight) batch_sync_down(h->batch);
}

static void batch_upload(zfv_batch *b) {
  std::vector<int64_t> cells(b->max_inst, 0), faces(b->max_inst, 0);
  for (int s = 0; s < b->max_inst; ++s)
    if (b->members[s] && b->h_inst[s].active) {
      cells[s] = b->h_inst[s].m.n;
      faces[s] = b->h_inst[s].m.ne;
    }
  int ph = -1;
  for (int s = 0; s < b->max_inst; ++s)
    if (b->members[s] && b->h_inst[s].active) {
      const int p = physics_set(b->h_inst[s].m);
      ph = ph < 0 ? p : physics_union(ph, p);
    }
  if (ph < 0) ph = PH_LIGHT;  // no active member: nothing is launched
  if (ph != b->ph && b->graph) {
    HIP_CHECK(hipGraphExecDestroy(b->graph));
    b->graph = nullptr;
  }
  b->ph = ph;
  auto tc = checked_table(cells, 256, b->cap_cell_blocks, b->xcd_chunks);
  auto tf = checked_table(faces, 256, b->cap_face_blocks, b->xcd_chunks);
  const int waves = rec_cfg_waves(b->rec_cfg >= 0 ? b->rec_cfg : kRecConfigs - 1);
  auto tr = checked_table(cells, int64_t(waves) * b->L.groups_per_wave, b->cap_rec_blocks, b->xcd_chunks);
  if (int64_t(tc.size()) != b->n_cell_blocks || int64_t(tf.size()) != b->n_face_blocks ||
      int64_t(tr.size()) != b->n_rec_blocks) {
    // launch sizes change: the step graph is captured again on the next run
    b->n_cell_blocks = int64_t(tc.size());
    b->n_face_blocks = int64_t(tf.size());
    b->n_rec_blocks = int64_t(tr.size());
    if (b->graph) {
      HIP_CHECK(hipGraphExecDestroy(b->graph));
      b->graph = nullptr;
    }
  }
  HIP_CHECK(hipMemcpyAsync(b->d_inst, b->h_inst.data(), sizeof(DevInst) * b->max_inst, hipMemcpyHostToDevice,
                           b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_bcell, tc.data(), sizeof(BlockRef) * tc.size(), hipMemcpyHostToDevice, b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_bface, tf.data(), sizeof(BlockRef) * tf.size(), hipMemcpyHostToDevice, b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_brec, tr.data(), sizeof(BlockRef) * tr.size(), hipMemcpyHostToDevice, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  b->dirty = false;
}

// The batch's reconstruction geometry, tuned like a single instance's: each
// candidate's block table is uploaded and stage-0 reconstructions of all
// active members are timed (one warm-up, three timed launches).
static void batch_tune(zfv_batch *b) {
  // the timed launches leave the members' kept equilibria alone
  std::vector<Img> kept(b->max_inst);
  for (int s = 0; s < b->max_inst; ++s) {
    kept[s] = b->h_inst[s].m;
    b->h_inst[s].m = stateless(kept[s]);
  }
  auto restore = [&] {
    for (int s = 0; s < b->max_inst; ++s) b->h_inst[s].m = kept[s];
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  try {
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    int best = kRecConfigs - 1;
    double best_ms = 1e300;
    for (int c = 0; c < kRecConfigs; ++c) {
      b->rec_cfg = c;
      batch_upload(b);
      const unsigned nr = unsigned(b->n_rec_blocks);
      b->L.brec[b->ph](b->d_inst, b->d_brec, nr, 0, c, b->stream);
      HIP_CHECK(hipEventRecord(e0, b->stream));
      for (int k = 0; k < 3; ++k) b->L.brec[b->ph](b->d_inst, b->d_brec, nr, 0, c, b->stream);
      HIP_CHECK(hipEventRecord(e1, b->stream));
      HIP_CHECK(hipEventSynchronize(e1));
      HIP_CHECK(hipGetLastError());
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      b->rec_ms[c] = ms / 3.0;
      if (b->rec_ms[c] < best_ms) {
        best_ms = b->rec_ms[c];
        best = c;
      }
    }
    b->rec_cfg = best;
  } catch (...) {
    restore();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  restore();
  batch_upload(b);
}

static bool batch_has_active(const zfv_batch *b) {
  for (int s = 0; s < b->max_inst; ++s)
    if (b->members[s] && b->h_inst[s].active) return true;
  return false;
}

static void batch_enqueue_step(zfv_batch *b) {
  const unsigned nc = unsigned(b->n_cell_blocks), nf = unsigned(b->n_face_blocks), nr = unsigned(b->n_rec_blocks);
  const unsigned ib = unsigned((b->max_inst + 63) / 64);
  // the step's CFL minimum (reset at the end of the previous step or when
  // the member joined); stage 0's update turns it into the step's dt
  b->L.bcfl(b->d_inst, b->d_bcell, nc, b->stream);
  for (int s = 0; s < 3; ++s) {
    b->L.brec[b->ph](b->d_inst, b->d_brec, nr, s, b->rec_cfg, b->stream);
    b->L.bface(b->d_inst, b->d_bface, nf, b->stream);
    b->L.bupdate(b->d_inst, b->d_bcell, nc, s, b->stream);
  }
  hipLaunchKernelGGL(kb_end_step, dim3(ib), dim3(64), 0, b->stream, b->d_inst, b->max_inst);
}

static void batch_forget(zfv_batch *b, int slot) {
  batch_sync_down(b);
  if (b->members[slot]) {
    b->members[slot]->batch = nullptr;
    b->members[slot]->batch_slot = -1;
  }
  b->members[slot] = nullptr;
  b->h_inst[slot].active = 0;
  batch_upload(b);
}

static void batch_member_steady(zfv_batch *b, int slot, const double *steady) {
  batch_sync_down(b);
  b->h_inst[slot].m.steady = steady;
  batch_upload(b);
}

static void batch_after_snapshot(zfv_batch *b, hipEvent_t copied) {
  HIP_CHECK(hipStreamWaitEvent(b->stream, copied, 0));
}

}  // namespace zfv

extern "C" {

int zfv_batch_create(const zfv_params *params, int64_t cell_capacity, int64_t face_capacity, int32_t max_instances,
                     zfv_batch **out) {
  ZFV_API_BEGIN
  if (!params || !out || max_instances < 1 || cell_capacity < 1) throw ZfvError(ZFV_ERR_INVALID, "bad batch arguments");
  if (cell_capacity > kBatchMaxCells || max_instances > kBatchMaxMembers)
    throw ZfvError(ZFV_ERR_INVALID, "batch capacity above " + std::to_string(kBatchMaxCells) + " cells or " +
                                        std::to_string(kBatchMaxMembers) +
                                        " members: larger batches are refused (an unexplained GPU memory fault)");
  check_device();
  auto b = std::make_unique<zfv_batch>();
  b->shape = shape_for(*params);
  b->L = pick_launcher(b->shape);
  b->max_inst = max_instances;
  // capacity in blocks, rounded to a multiple of 8 (one chunk per XCD)
  auto round8 = [](int64_t x) { return ((x + 7) / 8) * 8; };
  b->cap_cell_blocks = round8((cell_capacity + 255) / 256 + max_instances);
  b->cap_face_blocks = round8((std::max<int64_t>(face_capacity, 1) + 255) / 256 + max_instances);
  b->cap_rec_blocks = round8((cell_capacity + b->L.groups_per_wave - 1) / b->L.groups_per_wave + max_instances);
  HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_CHECK(hipMalloc(&b->d_inst, sizeof(DevInst) * max_instances));
  HIP_CHECK(hipMalloc(&b->d_bcell, sizeof(BlockRef) * b->cap_cell_blocks));
  HIP_CHECK(hipMalloc(&b->d_bface, sizeof(BlockRef) * b->cap_face_blocks));
  HIP_CHECK(hipMalloc(&b->d_brec, sizeof(BlockRef) * b->cap_rec_blocks));
  b->h_inst.assign(max_instances, DevInst{});
  b->members.assign(max_instances, nullptr);
  batch_upload(b.get());
  *out = b.release();
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_destroy(zfv_batch *b) {
  delete b;
  return ZFV_OK;
}

int zfv_batch_set_xcd_chunks(zfv_batch *b, int32_t on) {
  ZFV_API_BEGIN
  batch_sync_down(b);
  b->xcd_chunks = on != 0;
  batch_upload(b);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_add(zfv_batch *b, zfv_handle *h, double cfl, double t0, double t_end, int32_t max_steps,
                  int32_t *slot) {
  ZFV_API_BEGIN
  if (!same_shape(b->shape, h->img.s)) throw ZfvError(ZFV_ERR_INVALID, "instance shape differs from the batch's");
  if (h->batch) throw ZfvError(ZFV_ERR_INVALID, "the instance is already a member of a batch");
  batch_sync_down(b);
  int s = -1;
  for (int k = 0; k < b->max_inst; ++k)
    if (!b->members[k]) {
      s = k;
      break;
    }
  if (s < 0) throw ZfvError(ZFV_ERR_MEMORY, "batch is full");
  DevInst d{};
  d.m = h->dimg;
  d.U = h->U;
  d.X = h->X;
  d.K = h->K;
  d.fpv = h->fpv;
  d.G = h->G;
  d.src = h->src;
  d.cfl = cfl;
  d.t = t0;
  d.t_end = t_end;
  d.dt = 0.0;
  d.dt_bits = kb_no_bound();
  d.active = 1;
  d.steps_done = 0;
  d.max_steps = max_steps > 0 ? max_steps : 0x7fffffff;
  b->h_inst[s] = d;
  b->members[s] = h;
  try {
    batch_upload(b);
  } catch (...) {
    b->members[s] = nullptr;
    b->h_inst[s].active = 0;
    batch_upload(b);
    throw;
  }
  h->batch = b;
  h->batch_slot = s;
  *slot = s;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_remove(zfv_batch *b, int32_t slot) {
  ZFV_API_BEGIN
  if (slot < 0 || slot >= b->max_inst || !b->members[slot]) throw ZfvError(ZFV_ERR_INVALID, "bad slot");
  batch_forget(b, slot);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_launch(zfv_batch *b, int32_t n_steps) {
  ZFV_API_BEGIN
  if (n_steps < 0) throw ZfvError(ZFV_ERR_INVALID, "n_steps must be >= 0");
  if (b->dirty) batch_upload(b);
  if (b->rec_cfg < 0 && batch_has_active(b)) batch_tune(b);
  if (b->rec_cfg < 0) {  // nothing to run on: any configuration will do
    b->rec_cfg = kRecConfigs - 1;
    batch_upload(b);
  }
  if (!b->graph) {
    hipGraph_t g;
    HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    batch_enqueue_step(b);
    HIP_CHECK(hipStreamEndCapture(b->stream, &g));
    HIP_CHECK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(g));
  }
  for (int k = 0; k < n_steps; ++k) HIP_CHECK(hipGraphLaunch(b->graph, b->stream));
  b->in_flight = true;
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_wait(zfv_batch *b) {
  ZFV_API_BEGIN
  if (b->in_flight) batch_sync_down(b);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_run(zfv_batch *b, int32_t n_steps) {
  const int rc = zfv_batch_launch(b, n_steps);
  if (rc != ZFV_OK) return rc;
  return zfv_batch_wait(b);
}

int zfv_batch_tune(zfv_batch *b, int32_t rec_threads, zfv_tuning *out) {
  ZFV_API_BEGIN
  if (b->in_flight) batch_sync_down(b);
  if (b->dirty) batch_upload(b);
  if (rec_threads == 0) {
    if (batch_has_active(b)) batch_tune(b);
  } else if (rec_threads > 0) {
    b->rec_cfg = rec_cfg_of(rec_threads);
    batch_upload(b);
  }
  if (b->graph && rec_threads >= 0) {
    HIP_CHECK(hipGraphExecDestroy(b->graph));
    b->graph = nullptr;
  }
  report_tuning(b->rec_cfg, b->L.groups_per_wave, b->rec_ms, out);
  return ZFV_OK;
  ZFV_API_END
}

int zfv_batch_query(zfv_batch *b, int32_t slot, double *t, double *dt, int32_t *steps_done, int32_t *active) {
  ZFV_API_BEGIN
  if (slot < 0 || slot >= b->max_inst || !b->members[slot]) throw ZfvError(ZFV_ERR_INVALID, "bad slot");
  if (b->in_flight) batch_sync_down(b);
  const DevInst &d = b->h_inst[slot];
  if (t) *t = d.t;
  if (dt) *dt = d.dt;
  if (steps_done) *steps_done = d.steps_done;
  if (active) *active = d.active;
  return ZFV_OK;
  ZFV_API_END
}

}  // extern "C"
