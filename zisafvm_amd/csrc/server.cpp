// Serving runtime: continuous batching of solve requests on one GPU.
//
// Not in the reference (it runs one experiment per process).  Built on the
// library's public C-ABI (include/zfv.h) and the two halves of zfv_create
// (common.hpp):
//   submit   -> pending queue (request copied)
//   admit    -> when the request's estimated HBM fits the budget, a setup
//               worker builds its instance (zfv_create's two halves: the
//               host setup image, then allocation and upload) while the GPU
//               keeps stepping others
//   join     -> at the next step boundary the instance joins the batch of its
//               kernel shape (zfv_batch_add); every batch advances one SSP3
//               step per loop iteration with each member's own CFL dt
//   retire   -> members that reach their end time (or step limit) are copied
//               out, removed, destroyed, and their HBM returns to the budget
// A member's result equals the same request run alone (same kernels, same
// per-member dt), which the tests check bit for bit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "setup.hpp"

namespace zfv {

int api_fail(const std::exception &e);  // kernels.hip

namespace {

using clock_t_ = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  MeshBuf mesh;
  zfv_params params{};
  std::vector<double> u0, steady;
  double t_end = 0.0, cfl = 0.45;
  int max_steps = 0;
  int64_t est_bytes = 0;
  int state = ZFV_REQ_QUEUED;
  std::string error;
  zfv_handle *h = nullptr;
  zfv_batch *batch = nullptr;
  int slot = -1;
  std::vector<double> result;
  double t_final = 0.0;
  int steps = 0;
  clock_t_::time_point t_submit, t_admit, t_start, t_done;
};

struct ShapeKey {
  int dim, dofc, rc, nb, rb, q;
  bool operator<(const ShapeKey &o) const {
    return std::tie(dim, dofc, rc, nb, rb, q) < std::tie(o.dim, o.dofc, o.rc, o.nb, o.rb, o.q);
  }
};

}  // namespace
}  // namespace zfv

struct zfv_server {
  int64_t budget = 0;
  int64_t reserved = 0;
  int64_t batch_cells = 0;
  int32_t max_members = 0;
  std::mutex mu;
  // held by a setup worker around its instance calls (allocation, uploads,
  // setup kernels) and by the GPU loop around its batch launches, which may
  // capture a step graph: no other thread's device calls run while a graph
  // is being captured
  std::mutex hip_mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<zfv::Request>> pending;  // waiting for HBM
  std::deque<std::shared_ptr<zfv::Request>> setup_q;  // admitted, waiting for a worker
  std::deque<std::shared_ptr<zfv::Request>> ready;    // built, waiting for a step boundary
  std::map<int64_t, std::shared_ptr<zfv::Request>> all;
  std::map<zfv::ShapeKey, zfv_batch *> batches;
  std::vector<std::shared_ptr<zfv::Request>> running;
  std::vector<std::thread> workers;
  std::thread gpu;
  std::atomic<bool> stop{false};
  int64_t next_id = 1;
  int32_t building = 0;  // requests a setup worker is building now
  int64_t cell_stages = 0;
  int64_t steps_run = 0;
};

namespace zfv {
namespace {

ShapeKey key_of(const Shape &s) { return {s.dim, s.dofc, s.rc, s.nb, s.rb, s.q}; }

void admit_locked(zfv_server *S) {
  // FIFO with skipping: any pending request whose estimate fits is admitted
  for (auto it = S->pending.begin(); it != S->pending.end();) {
    auto &r = *it;
    if (S->reserved + r->est_bytes <= S->budget) {
      S->reserved += r->est_bytes;
      r->state = ZFV_REQ_SETUP;
      r->t_admit = clock_t_::now();
      S->setup_q.push_back(r);
      it = S->pending.erase(it);
    } else {
      ++it;
    }
  }
}

void setup_worker(zfv_server *S) {
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->setup_q.empty(); });
      if (S->stop) return;
      r = S->setup_q.front();
      S->setup_q.pop_front();
      S->building += 1;
    }
    zfv_mesh m{r->mesh.n_dims, r->mesh.n_cells, r->mesh.vertex_ids.data(), r->mesh.corners.data(),
               r->mesh.ghost.data()};
    zfv_handle *h = nullptr;
    std::string err;
    int rc = zfv::create_host(&m, &r->params, &h);
    {
      std::lock_guard<std::mutex> hl(S->hip_mu);
      if (rc == ZFV_OK) rc = zfv::create_device(h);
      if (rc == ZFV_OK) rc = zfv_set_state(h, r->u0.data());
      if (rc == ZFV_OK && !r->steady.empty()) rc = zfv_set_steady_state(h, r->steady.data());
      if (rc != ZFV_OK) {
        err = zfv_last_error();
        if (h) zfv_destroy(h);
        h = nullptr;
      }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->building -= 1;
    if (rc != ZFV_OK) {
      r->state = ZFV_REQ_FAILED;
      r->error = err;
      S->reserved -= r->est_bytes;
      r->t_done = clock_t_::now();
      admit_locked(S);
      S->cv.notify_all();
      continue;
    }
    r->h = h;
    S->ready.push_back(r);
    S->cv.notify_all();
  }
}

void gpu_loop(zfv_server *S) {
  for (;;) {
    std::vector<std::shared_ptr<Request>> joining;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->ready.empty() || !S->running.empty(); });
      if (S->stop) return;
      joining.assign(S->ready.begin(), S->ready.end());
      S->ready.clear();
    }
    // join at the step boundary; a request whose batch has no free slot or
    // cells right now waits for the next boundary while other members of
    // that batch are still running (they free slots as they retire), and
    // fails only when it could not join even an empty batch
    std::vector<std::shared_ptr<Request>> deferred;
    for (auto &r : joining) {
      Shape s = shape_for(r->params);
      ShapeKey k = key_of(s);
      zfv_batch *b = nullptr;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->batches.find(k);
        if (it != S->batches.end()) b = it->second;
      }
      int rc = ZFV_OK;
      if (!b) {
        rc = zfv_batch_create(&r->params, S->batch_cells, 4 * S->batch_cells, S->max_members, &b);
        if (rc == ZFV_OK) {
          std::lock_guard<std::mutex> lk(S->mu);
          S->batches[k] = b;
        }
      }
      int32_t slot = -1;
      if (rc == ZFV_OK) rc = zfv_batch_add(b, r->h, r->cfl, 0.0, r->t_end, r->max_steps, &slot);
      if (rc == ZFV_ERR_MEMORY && b) {
        bool others = false;
        for (const auto &q : S->running) others = others || q->batch == b;
        if (others) {
          deferred.push_back(r);
          continue;
        }
      }
      std::lock_guard<std::mutex> lk(S->mu);
      if (rc != ZFV_OK) {
        r->state = ZFV_REQ_FAILED;
        r->error = zfv_last_error();
        zfv_destroy(r->h);
        r->h = nullptr;
        S->reserved -= r->est_bytes;
        r->t_done = clock_t_::now();
        admit_locked(S);
        S->cv.notify_all();
        continue;
      }
      r->batch = b;
      r->slot = slot;
      r->state = ZFV_REQ_RUNNING;
      r->t_start = clock_t_::now();
      S->running.push_back(r);
    }
    if (!deferred.empty()) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->ready.insert(S->ready.begin(), deferred.begin(), deferred.end());
    }
    // one step of every batch with active members
    std::vector<zfv_batch *> bs;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      for (auto &kv : S->batches) bs.push_back(kv.second);
    }
    // every batch's step launched before any is waited for: batches of
    // different shapes (each on its own stream) overlap on the device
    int run_rc = ZFV_OK;
    {
      std::lock_guard<std::mutex> hl(S->hip_mu);
      for (zfv_batch *b : bs) {
        int rc = zfv_batch_launch(b, 1);
        if (rc != ZFV_OK && run_rc == ZFV_OK) run_rc = rc;
      }
    }
    for (zfv_batch *b : bs) {
      int rc = zfv_batch_wait(b);
      if (rc != ZFV_OK && run_rc == ZFV_OK) run_rc = rc;
    }
    // retire finished members; request fields that poll() reads are
    // written under the lock below
    struct Progress {
      std::shared_ptr<Request> r;
      int32_t steps;
      double t;
      bool finished;
      int rc;
    };
    std::vector<Progress> prog;
    std::string run_err = run_rc == ZFV_OK ? std::string() : std::string(zfv_last_error());
    for (auto &r : S->running) {
      double t = 0, dt = 0;
      int32_t steps = 0, active = 0;
      int rc = run_rc;
      if (rc == ZFV_OK) rc = zfv_batch_query(r->batch, r->slot, &t, &dt, &steps, &active);
      Progress pg{r, steps, t, rc != ZFV_OK || !active, rc};
      if (pg.finished) {
        std::vector<double> out;
        if (rc == ZFV_OK) {
          out.assign(size_t(r->mesh.n_cells) * 5, 0.0);
          pg.rc = zfv_get_state(r->h, out.data());
        }
        zfv_batch_remove(r->batch, r->slot);
        zfv_destroy(r->h);
        std::lock_guard<std::mutex> lk(S->mu);
        r->result.swap(out);
        r->h = nullptr;
      }
      prog.push_back(std::move(pg));
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->steps_run += 1;
    std::vector<std::shared_ptr<Request>> still;
    bool any_done = false;
    for (auto &pg : prog) {
      Request &r = *pg.r;
      if (pg.rc == ZFV_OK) {
        S->cell_stages += int64_t(r.mesh.n_cells) * 3 * (pg.steps - r.steps);
        r.steps = pg.steps;
        r.t_final = pg.t;
      }
      if (!pg.finished) {
        still.push_back(pg.r);
        continue;
      }
      any_done = true;
      r.state = pg.rc == ZFV_OK ? ZFV_REQ_DONE : ZFV_REQ_FAILED;
      if (pg.rc != ZFV_OK) r.error = run_err.empty() ? std::string(zfv_last_error()) : run_err;
      r.t_done = clock_t_::now();
      S->reserved -= r.est_bytes;
    }
    S->running = still;
    if (any_done) admit_locked(S);
    S->cv.notify_all();
  }
}

}  // namespace
}  // namespace zfv

using namespace zfv;

extern "C" {

int zfv_server_create(int64_t hbm_budget_bytes, int64_t batch_cell_capacity, int32_t max_members,
                      int32_t setup_workers, zfv_server **out) {
  try {
    if (hbm_budget_bytes <= 0 || batch_cell_capacity <= 0 || max_members <= 0 || setup_workers <= 0)
      throw ZfvError(ZFV_ERR_INVALID, "bad server arguments");
    auto S = std::make_unique<zfv_server>();
    S->budget = hbm_budget_bytes;
    S->batch_cells = batch_cell_capacity;
    S->max_members = max_members;
    zfv_server *p = S.get();
    for (int k = 0; k < setup_workers; ++k) S->workers.emplace_back(setup_worker, p);
    S->gpu = std::thread(gpu_loop, p);
    *out = S.release();
    return ZFV_OK;
  } catch (const std::exception &e) {
    return api_fail(e);
  }
}

int zfv_server_submit(zfv_server *S, const zfv_mesh *mesh, const zfv_params *params, const double *u0,
                      const double *steady, double t_end, double cfl, int32_t max_steps, int64_t *id) {
  try {
    auto r = std::make_shared<Request>();
    const int nv = mesh->n_dims + 1;
    r->mesh.n_dims = mesh->n_dims;
    r->mesh.n_cells = mesh->n_cells;
    r->mesh.vertex_ids.assign(mesh->vertex_ids, mesh->vertex_ids + size_t(mesh->n_cells) * nv);
    r->mesh.corners.assign(mesh->corners, mesh->corners + size_t(mesh->n_cells) * nv * 3);
    if (mesh->ghost_mask) r->mesh.ghost.assign(mesh->ghost_mask, mesh->ghost_mask + mesh->n_cells);
    else r->mesh.ghost.assign(mesh->n_cells, 0);
    r->params = *params;
    r->u0.assign(u0, u0 + size_t(mesh->n_cells) * 5);
    if (steady) r->steady.assign(steady, steady + size_t(mesh->n_cells) * 5);
    r->t_end = t_end;
    r->cfl = cfl;
    r->max_steps = max_steps;
    int64_t est = 0;
    int rc = zfv_estimate_device_bytes(mesh, params, &est);
    if (rc != ZFV_OK) return rc;
    r->est_bytes = est;
    if (est > S->budget) throw ZfvError(ZFV_ERR_MEMORY, "request larger than the HBM budget");
    r->t_submit = clock_t_::now();
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    S->all[r->id] = r;
    S->pending.push_back(r);
    admit_locked(S);
    S->cv.notify_all();
    *id = r->id;
    return ZFV_OK;
  } catch (const std::exception &e) {
    return api_fail(e);
  }
}

int zfv_server_poll(zfv_server *S, int64_t id, zfv_request_status *st) {
  try {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) throw ZfvError(ZFV_ERR_INVALID, "unknown request id");
    const Request &r = *it->second;
    auto secs = [&](clock_t_::time_point a) {
      return a.time_since_epoch().count() == 0 ? -1.0 : std::chrono::duration<double>(a - r.t_submit).count();
    };
    st->state = r.state;
    st->t_final = r.t_final;
    st->steps = r.steps;
    st->n_cells = r.mesh.n_cells;
    st->admit_s = secs(r.t_admit);
    st->start_s = secs(r.t_start);
    st->latency_s = r.state == ZFV_REQ_DONE || r.state == ZFV_REQ_FAILED ? secs(r.t_done) : -1.0;
    return ZFV_OK;
  } catch (const std::exception &e) {
    return api_fail(e);
  }
}

int zfv_server_result(zfv_server *S, int64_t id, double *u_out) {
  try {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) throw ZfvError(ZFV_ERR_INVALID, "unknown request id");
    const Request &r = *it->second;
    if (r.state == ZFV_REQ_FAILED) throw ZfvError(ZFV_ERR_INVALID, "request failed: " + r.error);
    if (r.state != ZFV_REQ_DONE) throw ZfvError(ZFV_ERR_INVALID, "request not finished");
    std::memcpy(u_out, r.result.data(), r.result.size() * sizeof(double));
    return ZFV_OK;
  } catch (const std::exception &e) {
    return api_fail(e);
  }
}

int zfv_server_release(zfv_server *S, int64_t id) {
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->all.find(id);
  if (it != S->all.end() && (it->second->state == ZFV_REQ_DONE || it->second->state == ZFV_REQ_FAILED))
    S->all.erase(it);
  return ZFV_OK;
}

int zfv_server_stats(zfv_server *S, zfv_server_stats_t *out) {
  std::lock_guard<std::mutex> lk(S->mu);
  out->reserved_bytes = S->reserved;
  out->pending = int32_t(S->pending.size());
  out->in_setup = int32_t(S->setup_q.size()) + S->building;
  out->running = int32_t(S->running.size());
  out->ready = int32_t(S->ready.size());
  out->cell_stages = S->cell_stages;
  out->steps_run = S->steps_run;
  return ZFV_OK;
}

int zfv_server_destroy(zfv_server *S) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv.notify_all();
  }
  for (auto &w : S->workers) w.join();
  if (S->gpu.joinable()) S->gpu.join();
  for (auto &kv : S->all)
    if (kv.second->h) zfv_destroy(kv.second->h);
  for (auto &kv : S->batches) zfv_batch_destroy(kv.second);
  delete S;
  return ZFV_OK;
}

}  // extern "C"
