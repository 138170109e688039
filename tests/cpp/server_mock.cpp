// TEST-ONLY: the serving runtime (zisafvm_amd/csrc/server.cpp) linked against
// a host mock of the instance and batch C-ABI, so its scheduling and memory
// lifetimes can run under AddressSanitizer / ThreadSanitizer on the CPU.
//
// The mock keeps every handle's state on the heap and, in zfv_batch_run,
// reads and writes the state of every active member, as the batch kernels
// do.  A member used after zfv_destroy, a slot that still names a destroyed
// handle, or a race on server data is reported by the sanitizers; the
// invariants below are checked explicitly as well.  Requests follow the
// serving benchmark's mix: sizes, end times and Poisson arrivals.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zfv.h"
#include "../../zisafvm_amd/csrc/common.hpp"

namespace {
std::mutex g_mu;                    // guards the registries below
std::set<const void *> g_live;      // handles alive
std::map<const void *, const void *> g_member_of;  // handle -> batch
thread_local std::string g_err;
std::atomic<long> g_steps{0};

void fail(const char *what) {
  std::fprintf(stderr, "server_mock: invariant violated: %s\n", what);
  std::abort();
}
}  // namespace

struct zfv_handle {
  int64_t n = 0;
  int nv = 5;
  std::vector<double> U;
  uint64_t magic = 0x5a4656ULL;
};

struct MockMember {
  zfv_handle *h = nullptr;
  double cfl = 0, t = 0, t_end = 0, dt = 0;
  int steps = 0, max_steps = 0, active = 0;
};

struct zfv_batch {
  int max_inst = 0;
  std::vector<MockMember> m;
};

namespace zfv {
int api_fail(const std::exception &e) {
  g_err = e.what();
  if (auto *z = dynamic_cast<const ZfvError *>(&e)) return z->code;
  return ZFV_ERR_INVALID;
}
}  // namespace zfv

extern "C" {

const char *zfv_last_error(void) { return g_err.c_str(); }
int zfv_create(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out);
}

// zfv_create's two halves: the mock builds the whole handle in the host half
namespace zfv {
int create_host(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out) {
  return zfv_create(mesh, params, out);
}
int create_device(zfv_handle *) { return ZFV_OK; }
}  // namespace zfv

extern "C" {

int zfv_create(const zfv_mesh *mesh, const zfv_params *params, zfv_handle **out) {
  (void)params;
  auto *h = new zfv_handle;
  h->n = mesh->n_cells;
  h->U.assign(size_t(h->n) * 5, 0.0);
  // setup takes time in the real library; let other threads interleave
  std::this_thread::sleep_for(std::chrono::microseconds(200 + (h->n % 500)));
  std::lock_guard<std::mutex> lk(g_mu);
  g_live.insert(h);
  *out = h;
  return ZFV_OK;
}

int zfv_destroy(zfv_handle *h) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_live.count(h)) fail("destroy of an unknown handle");
    if (g_member_of.count(h)) fail("destroy of a handle that is still a batch member");
    g_live.erase(h);
  }
  delete h;
  return ZFV_OK;
}

int zfv_set_state(zfv_handle *h, const double *u) {
  std::memcpy(h->U.data(), u, sizeof(double) * h->U.size());
  return ZFV_OK;
}

int zfv_get_state(zfv_handle *h, double *u) {
  std::memcpy(u, h->U.data(), sizeof(double) * h->U.size());
  return ZFV_OK;
}

int zfv_set_steady_state(zfv_handle *h, const double *u) {
  (void)h;
  (void)u;
  return ZFV_OK;
}

int zfv_estimate_device_bytes(const zfv_mesh *mesh, const zfv_params *params, int64_t *bytes) {
  (void)params;
  *bytes = int64_t(mesh->n_cells) * 1500 + (1 << 20);
  return ZFV_OK;
}

int zfv_batch_create(const zfv_params *p, int64_t cells, int64_t faces, int32_t max_inst, zfv_batch **out) {
  (void)p;
  (void)cells;
  (void)faces;
  auto *b = new zfv_batch;
  b->max_inst = max_inst;
  b->m.assign(max_inst, MockMember{});
  *out = b;
  return ZFV_OK;
}

int zfv_batch_destroy(zfv_batch *b) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &mm : b->m)
    if (mm.h) g_member_of.erase(mm.h);
  delete b;
  return ZFV_OK;
}

int zfv_batch_add(zfv_batch *b, zfv_handle *h, double cfl, double t0, double t_end, int32_t max_steps, int32_t *slot) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_live.count(h)) fail("add of a destroyed handle");
    if (g_member_of.count(h)) fail("handle added twice");
  }
  for (int s = 0; s < b->max_inst; ++s)
    if (!b->m[s].h) {
      b->m[s] = MockMember{h, cfl, t0, t_end, 0.0, 0, max_steps > 0 ? max_steps : 0x7fffffff, 1};
      std::lock_guard<std::mutex> lk(g_mu);
      g_member_of[h] = b;
      *slot = s;
      return ZFV_OK;
    }
  g_err = "batch is full";
  return ZFV_ERR_MEMORY;
}

int zfv_batch_remove(zfv_batch *b, int32_t slot) {
  if (slot < 0 || slot >= b->max_inst || !b->m[slot].h) fail("remove of an empty slot");
  std::lock_guard<std::mutex> lk(g_mu);
  g_member_of.erase(b->m[slot].h);
  b->m[slot] = MockMember{};
  return ZFV_OK;
}

int zfv_batch_run(zfv_batch *b, int32_t n_steps) {
  for (int k = 0; k < n_steps; ++k)
    for (auto &mm : b->m) {
      if (!mm.h || !mm.active) continue;
      if (mm.h->magic != 0x5a4656ULL) fail("batch member state is not a live handle");
      // the "kernels": read and write the member's state (first, middle and
      // last value: enough for the sanitizers to see a freed member)
      mm.dt = std::min(mm.cfl * 1e-3 * (1.0 + 1e-6 * double(mm.h->n)), mm.t_end - mm.t);
      const size_t last = mm.h->U.size() - 1;
      for (size_t q : {size_t(0), last / 2, last}) mm.h->U[q] = mm.h->U[q] + mm.dt;
      mm.t += mm.dt;
      mm.steps += 1;
      g_steps++;
      if (mm.t >= mm.t_end || mm.steps >= mm.max_steps || !(mm.dt > 0.0)) mm.active = 0;
    }
  return ZFV_OK;
}

// the mock's steps run at launch; wait has nothing left to wait for
int zfv_batch_launch(zfv_batch *b, int32_t n_steps) { return zfv_batch_run(b, n_steps); }
int zfv_batch_wait(zfv_batch *) { return ZFV_OK; }

int zfv_batch_query(zfv_batch *b, int32_t slot, double *t, double *dt, int32_t *steps, int32_t *active) {
  if (slot < 0 || slot >= b->max_inst || !b->m[slot].h) fail("query of an empty slot");
  const MockMember &mm = b->m[slot];
  if (t) *t = mm.t;
  if (dt) *dt = mm.dt;
  if (steps) *steps = mm.steps;
  if (active) *active = mm.active;
  return ZFV_OK;
}

}  // extern "C"

int main(int argc, char **argv) {
  const int n_req = argc > 1 ? std::atoi(argv[1]) : 48;
  const double rate = argc > 2 ? std::atof(argv[2]) : 400.0;  // requests per second
  const int max_members = argc > 3 ? std::atoi(argv[3]) : 64;  // slots per batch
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U01(0.0, 1.0);
  zfv_server *S = nullptr;
  if (zfv_server_create(int64_t(1) << 30, int64_t(1) << 23, max_members, 2, &S) != ZFV_OK) fail("create");
  struct Req {
    zfv_mesh mesh;
    std::vector<int32_t> vid;
    std::vector<double> crn, u0, steady;
    std::vector<uint8_t> ghost;
    zfv_params p;
    double t_end;
    int64_t id = -1;
  };
  std::vector<Req> reqs(n_req);
  for (int k = 0; k < n_req; ++k) {
    Req &r = reqs[k];
    const bool d3 = k % 4 == 3;
    const int nv = d3 ? 4 : 3;
    const int64_t n = d3 ? int64_t(10000 + 40000 * U01(rng)) : int64_t(std::exp(std::log(1e4) + U01(rng) * std::log(20.0)));
    r.vid.assign(size_t(n) * nv, 0);
    r.crn.assign(size_t(n) * nv * 3, 0.0);
    r.ghost.assign(n, 0);
    r.u0.assign(size_t(n) * 5, 1.0);
    if (k % 2) r.steady.assign(size_t(n) * 5, 1.0);
    r.mesh = zfv_mesh{d3 ? 3 : 2, int32_t(n), r.vid.data(), r.crn.data(), r.ghost.data()};
    std::memset(&r.p, 0, sizeof(r.p));
    r.p.n_dims = d3 ? 3 : 2;
    r.p.n_stencils = d3 ? 5 : 4;
    r.p.orders[0] = 3;
    for (int s = 1; s < 5; ++s) r.p.orders[s] = 2;
    for (int s = 1; s < 5; ++s) r.p.biases[s] = 1;
    r.p.overfit[0] = 2.0;
    for (int s = 1; s < 5; ++s) r.p.overfit[s] = 1.5;
    for (int s = 0; s < 5; ++s) r.p.linear_weights[s] = s == 0 ? 100.0 : 1.0;
    r.p.epsilon = 1e-6;
    r.p.exponent = 4.0;
    r.p.gamma = d3 ? 2.0 : 1.4;
    r.p.edge_deg = d3 ? 2 : 3;
    r.p.volume_deg = d3 ? 3 : 4;
    r.p.moments_deg = -1;
    r.p.steps_per_recompute = 1;
    if (d3) {
      r.p.flux = ZFV_FLUX_RUSANOV;
      r.p.equilibrium = ZFV_EQ_ISENTROPIC;
      r.p.gravity = ZFV_GRAVITY_POLYTROPE;
    }
    r.t_end = 0.05 + 0.45 * U01(rng);
  }
  std::exponential_distribution<double> gap(rate);
  for (auto &r : reqs) {
    std::this_thread::sleep_for(std::chrono::duration<double>(gap(rng)));
    int rc = zfv_server_submit(S, &r.mesh, &r.p, r.u0.data(), r.steady.empty() ? nullptr : r.steady.data(), r.t_end,
                               0.45, 0, &r.id);
    if (rc != ZFV_OK) {
      std::fprintf(stderr, "submit failed: %s\n", zfv_last_error());
      return 1;
    }
  }
  // poll until done, reading results and releasing as they finish
  int done = 0;
  std::vector<double> out;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<bool> seen(n_req, false);
  while (done < n_req) {
    for (int k = 0; k < n_req; ++k) {
      if (seen[k]) continue;
      zfv_request_status st{};
      if (zfv_server_poll(S, reqs[k].id, &st) != ZFV_OK) fail("poll");
      if (st.state == ZFV_REQ_FAILED) {
        std::fprintf(stderr, "request %d failed\n", k);
        return 1;
      }
      if (st.state == ZFV_REQ_DONE) {
        out.assign(size_t(reqs[k].mesh.n_cells) * 5, 0.0);
        if (zfv_server_result(S, reqs[k].id, out.data()) != ZFV_OK) fail("result");
        const double expect = 1.0 + st.t_final;  // the mock adds dt to every value each step
        if (std::fabs(out[0] - expect) > 1e-9 * st.steps + 1e-12) fail("result of a request does not match its steps");
        zfv_server_release(S, reqs[k].id);
        seen[k] = true;
        ++done;
      }
    }
    if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) fail("requests did not finish");
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  zfv_server_stats_t stt{};
  zfv_server_stats(S, &stt);
  if (stt.reserved_bytes != 0 || stt.running != 0 || stt.pending != 0) fail("server not drained");
  zfv_server_destroy(S);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_live.empty()) fail("handles leaked");
  }
  std::printf("server_mock: OK %d requests, %ld member steps\n", n_req, long(g_steps.load()));
  return 0;
}
