// Continuous-batching solve server. Requests (a mesh, an element kind with its parameters, Dirichlet flags and
// loads) are admitted in arrival order while a slot is free and their estimated device memory fits the budget;
// worker threads, each on its own stream, do a request's host symbolic work (pattern, block-multicolour
// ordering, sweep slabs), assembly and ILU(0) factorisation while the batch of admitted systems keeps
// iterating on the main stream; a prepared system joins the running batch at the next step and leaves it when
// converged. Every system keeps its own Krylov scalars, so its result does not depend on what else is served.
// No counterpart in the reference (single-threaded, one system at a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dfm_gpu.h"
#include "dfm_internal.hpp"

namespace dfm {

namespace {

enum { Q_QUEUED = 0, Q_SETUP = 1, Q_SOLVING = 2, Q_CONVERGED = 3, Q_MAXITER = 4, Q_FAILED = 5 };

struct Request {
  long long id = 0;
  int ndim = 0, len = 0, kind = 0, nnode = 0, nelem = 0, npar = 0, pstride = 0, nface = 0;
  std::vector<double> coords, params, trac, glob;
  std::vector<int> lnods, bc, faces;
  double tol = 1e-8;
  unsigned max_iter = 1000;
  int method = DFM_METHOD_PCG;
  long long est_bytes = 0;
  int state = Q_QUEUED;
  dfm_sys sys = nullptr;
  int worker = -1;  // the worker that set it up: its stream is the system's own
  double t_submit = 0, t_start = 0, t_done = 0;
  unsigned iters = 0;
  double conv = 0;
  double true_rel = -1;  // ||b - A x|| / ||b|| (FP64 matrix) of the finished solve
  int fell_back = 0;
  std::string err;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device bytes of one request: FP64 matrix and factor, FP32 matrix, FP64 and FP32 sweep slabs, their indices,
// and the vectors of a mixed-precision solve (P1: ~6 off-diagonal blocks per row in 2D, ~14 in 3D).
long long estimate_bytes(int ndim, int len, long long nnode) {
  const long long blocks = nnode * (ndim == 2 ? 7 : 15);
  const long long per_block = 32LL * len * len + 12;
  const long long per_dof = 96;
  return (blocks * per_block + nnode * len * per_dof) * 5 / 4 + (1LL << 20);
}

}  // namespace

struct Server {
  dfm_ctx ctx = nullptr;
  int precision = DFM_MIXED;
  long long budget = 0;
  int max_active = 1;
  dfm_batch batch = nullptr;
  std::vector<dfm_ctx> wctx;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_ready;
  bool stop = false;
  long long next_id = 1;
  std::deque<Request*> pending, setup_q, ready_q;
  // released systems, destroyed by the worker that created them (destruction waits for the system's own stream,
  // which is that worker's; the client thread would otherwise wait for whatever setup the worker has queued)
  std::vector<std::vector<dfm_sys>> grave;
  std::map<long long, Request*> all;
  std::vector<Request*> solving;
  long long bytes_admitted = 0;
  int active = 0;  // admitted and not finished
  // host-time counters (seconds): batch steps, waits with nothing to iterate, per-result true residuals; the
  // workers' setup time (under mu)
  long long n_step = 0, n_setup = 0;
  double t_step = 0, t_idle = 0, t_check = 0, t_setup = 0, t_join = 0, t_progress = 0;

  void worker(int w) {
    dfm_ctx c = wctx[w];
    (void)hipSetDevice(c->c.device);
    std::vector<dfm_sys> dead;
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !setup_q.empty() || !grave[w].empty(); });
        if (stop) return;
        dead.swap(grave[w]);
        if (!setup_q.empty()) {
          r = setup_q.front();
          setup_q.pop_front();
          r->worker = w;
        }
      }
      for (dfm_sys d : dead) dfm_sys_destroy(d);
      dead.clear();
      if (!r) continue;
      const double t0 = now_s();
      int rc = setup(c, *r);
      const double t1 = now_s();
      // PBiCGSTAB requests are solved here, on this worker's stream, beside the batch (which runs PCG only)
      int bst = 0;
      if (rc == DFM_OK && r->method == DFM_METHOD_PBICGSTAB) {
        {
          std::lock_guard<std::mutex> lk(mu);
          r->state = Q_SOLVING;
        }
        double conv = r->tol;
        unsigned it = r->max_iter;
        rc = dfm_solve_pbicgstab(r->sys, precision, &conv, &it);
        if (rc >= 0) {
          bst = (rc == 1 && it < r->max_iter) ? Q_CONVERGED : Q_MAXITER;
          double tr = 0.0;
          rc = sys_true_residual(r->sys->s, r->sys->s.ctx->stream, &tr);
          std::lock_guard<std::mutex> lk(mu);
          r->conv = conv;
          r->iters = it;
          r->true_rel = tr;
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        t_setup += t1 - t0;
        n_setup++;
        if (rc != DFM_OK) {
          r->state = Q_FAILED;
          r->err = dfm_last_error();
          r->t_done = now_s();
          active--;
        } else if (bst) {
          r->state = bst;
          r->t_done = now_s();
          active--;
          admit_locked();
        } else {
          ready_q.push_back(r);
        }
      }
      cv_ready.notify_all();
    }
  }

  int setup(dfm_ctx c, Request& r) {
    dfm_sys s = nullptr;
    int rc = dfm_sys_create_mesh(c, r.ndim, r.nnode, r.coords.data(), r.nelem, r.lnods.data(), r.len, &s);
    if (rc != DFM_OK) return rc;
    std::vector<int> order(r.nnode);
    int ncol = 0;
    rc = dfm_sys_make_ordering(s, DFM_ORDER_BLOCKCOLOR, 0, order.data(), &ncol);
    if (rc == DFM_OK) rc = dfm_sys_set_ordering(s, order.data());
    if (rc == DFM_OK) rc = dfm_sys_set_bc(s, r.bc.data());
    if (rc == DFM_OK)
      rc = dfm_sys_assemble(s, r.kind, r.params.data(), r.npar, r.pstride, r.glob.empty() ? nullptr : r.glob.data(),
                            r.nface, r.faces.empty() ? nullptr : r.faces.data(),
                            r.trac.empty() ? nullptr : r.trac.data(), nullptr);
    if (rc == DFM_OK) rc = dfm_prec_factor(s);
    if (rc == DFM_OK) rc = sys_solve_setup(s->s, precision);
    if (rc == DFM_OK && hipStreamSynchronize(s->s.ctx->stream) != hipSuccess) rc = DFM_ERR_HIP;
    if (rc != DFM_OK) {
      const std::string e = dfm_last_error();
      dfm_sys_destroy(s);
      set_error(e);
      return rc;
    }
    r.sys = s;
    return DFM_OK;
  }

  // Admission in arrival order: a free slot and the estimated bytes within the budget (a request larger than
  // the whole budget is admitted alone).
  void admit_locked() {
    bool any = false;
    while (!pending.empty() && active < max_active) {
      Request* r = pending.front();
      if (active > 0 && bytes_admitted + r->est_bytes > budget) break;
      pending.pop_front();
      r->state = Q_SETUP;
      r->t_start = now_s();
      bytes_admitted += r->est_bytes;
      active++;
      setup_q.push_back(r);
      any = true;
    }
    if (any) cv_work.notify_all();
  }

  int progress(int iterations, int wait_ms) {
    std::vector<Request*> joining;
    {
      std::unique_lock<std::mutex> lk(mu);
      admit_locked();
      if (solving.empty() && ready_q.empty() && active > 0 && wait_ms > 0) {
        const double t0 = now_s();
        cv_ready.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !ready_q.empty(); });
        t_idle += now_s() - t0;
      }
      joining.assign(ready_q.begin(), ready_q.end());
      ready_q.clear();
    }
    // a request that cannot join (e.g. no device memory for its solve buffers) fails alone; the others join
    int failed = 0;
    const double tj = now_s();
    for (Request* r : joining) {
      const int rc = dfm_batch_add(batch, r->sys, r->tol, r->max_iter);
      if (rc != DFM_OK) {
        std::lock_guard<std::mutex> lk(mu);
        r->state = Q_FAILED;
        r->err = dfm_last_error();
        r->t_done = now_s();
        active--;
        failed++;
        continue;
      }
      r->state = Q_SOLVING;
      solving.push_back(r);
    }
    t_join += now_s() - tj;
    if (failed) {
      std::lock_guard<std::mutex> lk(mu);
      admit_locked();
    }
    if (solving.empty()) return DFM_OK;
    int nrun = 0;
    const double t0 = now_s();
    int rc = dfm_batch_step(batch, iterations, &nrun);
    t_step += now_s() - t0;
    n_step++;
    if (rc != DFM_OK) return rc;
    std::vector<Request*> still;
    int finished = 0;
    for (Request* r : solving) {
      int st = 0;
      rc = dfm_batch_status(batch, r->sys, &st, &r->conv, &r->iters, nullptr);
      if (rc != DFM_OK) return rc;
      if (st == 0) {
        still.push_back(r);
        continue;
      }
      r->t_done = now_s();
      rc = dfm_batch_info(batch, r->sys, nullptr, &r->fell_back);
      if (rc != DFM_OK) return rc;
      // the true residual of every result (the FP64 loop's conv_ratio is its recursive one)
      const double tc = now_s();
      rc = sys_true_residual(r->sys->s, ctx->c.stream, &r->true_rel);
      t_check += now_s() - tc;
      if (rc != DFM_OK) return rc;
      r->state = st == 1 ? Q_CONVERGED : Q_MAXITER;
      rc = dfm_batch_remove(batch, r->sys);
      if (rc != DFM_OK) return rc;
      finished++;
    }
    solving.swap(still);
    if (finished) {
      std::lock_guard<std::mutex> lk(mu);
      active -= finished;
      admit_locked();
    }
    return DFM_OK;
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto& t : workers) t.join();
    for (auto& g : grave)
      for (dfm_sys d : g) dfm_sys_destroy(d);
    if (batch) dfm_batch_destroy(batch);
    for (auto& kv : all) {
      if (kv.second->sys) dfm_sys_destroy(kv.second->sys);
      delete kv.second;
    }
    for (dfm_ctx c : wctx) dfm_ctx_destroy(c);
  }
};

}  // namespace dfm

using namespace dfm;

struct dfm_server_s {
  Server s;
};

extern "C" {

int dfm_server_create(dfm_ctx ctx, int precision, long long budget_bytes, int max_active, int nworkers,
                      dfm_server* out) {
  DFM_CHECK(ctx && out, DFM_ERR_ARG, "dfm_server_create: null argument");
  DFM_CHECK(budget_bytes > 0 && max_active > 0 && nworkers > 0 && nworkers <= 64, DFM_ERR_ARG,
            "dfm_server_create: bad budget, slots or workers");
  auto* h = new dfm_server_s;
  Server& sv = h->s;
  sv.ctx = ctx;
  sv.precision = precision;
  sv.budget = budget_bytes;
  sv.max_active = max_active;
  int rc = dfm_batch_create(ctx, precision, max_active, &sv.batch);
  for (int w = 0; rc == DFM_OK && w < nworkers; w++) {
    dfm_ctx c = nullptr;
    rc = dfm_ctx_create(ctx->c.device, &c);
    if (rc == DFM_OK) sv.wctx.push_back(c);
  }
  if (rc != DFM_OK) {
    delete h;
    return rc;
  }
  sv.grave.resize((size_t)nworkers);
  for (int w = 0; w < nworkers; w++) sv.workers.emplace_back([&sv, w] { sv.worker(w); });
  *out = h;
  return DFM_OK;
}

int dfm_server_destroy(dfm_server s) {
  delete s;
  return DFM_OK;
}

int dfm_server_submit(dfm_server s, const dfm_request* q, long long* ticket) {
  DFM_CHECK(s && q && ticket, DFM_ERR_ARG, "dfm_server_submit: null argument");
  DFM_CHECK((q->ndim == 2 || q->ndim == 3) && q->len >= 1 && q->len <= 3 && q->nnode > 0 && q->nelem > 0 &&
                q->coords && q->lnods && q->params && q->bc && q->npar > 0,
            DFM_ERR_ARG, "dfm_server_submit: incomplete request");
  auto* r = new Request;
  r->ndim = q->ndim; r->len = q->len; r->kind = q->kind; r->nnode = q->nnode; r->nelem = q->nelem;
  r->npar = q->npar; r->pstride = q->pstride; r->nface = q->nface; r->tol = q->tol; r->max_iter = q->max_iter;
  DFM_CHECK(q->method == DFM_METHOD_PCG || q->method == DFM_METHOD_PBICGSTAB, DFM_ERR_ARG,
            "dfm_server_submit: unknown method");
  r->method = q->method;
  r->coords.assign(q->coords, q->coords + (size_t)q->nnode * q->ndim);
  r->lnods.assign(q->lnods, q->lnods + (size_t)q->nelem * (q->ndim + 1));
  const size_t np = q->pstride ? (size_t)q->pstride * q->nelem : (size_t)q->npar;
  r->params.assign(q->params, q->params + np);
  r->bc.assign(q->bc, q->bc + (size_t)q->nnode * q->len);
  // exactly the kind's own scalars, as assembly reads them
  if (q->glob) r->glob.assign(q->glob, q->glob + kind_glob_count(q->kind));
  if (q->nface > 0) {
    DFM_CHECK(q->faces && q->trac, DFM_ERR_ARG, "dfm_server_submit: faces without tractions");
    r->faces.assign(q->faces, q->faces + (size_t)q->nface * 3);
    r->trac.assign(q->trac, q->trac + (size_t)q->nface * 3);
  }
  r->est_bytes = estimate_bytes(q->ndim, q->len, q->nnode);
  r->t_submit = now_s();
  Server& sv = s->s;
  std::lock_guard<std::mutex> lk(sv.mu);
  r->id = sv.next_id++;
  sv.all[r->id] = r;
  sv.pending.push_back(r);
  *ticket = r->id;
  return DFM_OK;
}

int dfm_server_progress(dfm_server s, int iterations, int wait_ms, int* n_active, int* n_queued) {
  DFM_CHECK(s && iterations > 0, DFM_ERR_ARG, "dfm_server_progress: bad argument");
  Server& sv = s->s;
  const double t0 = now_s();
  const int rc = sv.progress(iterations, wait_ms);
  sv.t_progress += now_s() - t0;
  std::lock_guard<std::mutex> lk(sv.mu);
  if (n_active) *n_active = sv.active;
  if (n_queued) *n_queued = (int)sv.pending.size();
  return rc;
}

int dfm_server_poll(dfm_server s, long long ticket, int* state, double* latency_s, double* solve_s,
                    unsigned* iteration, double* conv_ratio, int* fell_back) {
  DFM_CHECK(s, DFM_ERR_ARG, "dfm_server_poll: null server");
  Server& sv = s->s;
  std::lock_guard<std::mutex> lk(sv.mu);
  auto it = sv.all.find(ticket);
  DFM_CHECK(it != sv.all.end(), DFM_ERR_ARG, "dfm_server_poll: unknown ticket");
  const Request& r = *it->second;
  if (state) *state = r.state;
  const bool done = r.state >= Q_CONVERGED;
  if (latency_s) *latency_s = done ? r.t_done - r.t_submit : 0.0;
  if (solve_s) *solve_s = done ? r.t_done - r.t_start : 0.0;
  if (iteration) *iteration = r.iters;
  if (conv_ratio) *conv_ratio = r.conv;
  if (fell_back) *fell_back = r.fell_back;
  return DFM_OK;
}

int dfm_server_stats(dfm_server s, double* out, int n) {
  DFM_CHECK(s && out && n > 0, DFM_ERR_ARG, "dfm_server_stats: bad argument");
  Server& sv = s->s;
  double b[4] = {0, 0, 0, 0};
  batch_stats(batch_of(sv.batch), b);
  std::lock_guard<std::mutex> lk(sv.mu);
  const double v[12] = {(double)sv.n_step, sv.t_step, b[0], b[1], b[2], b[3], sv.t_idle, sv.t_setup,
                        (double)sv.n_setup, sv.t_check, sv.t_join, sv.t_progress};
  for (int k = 0; k < n && k < 12; k++) out[k] = v[k];
  return DFM_OK;
}

int dfm_server_true_residual(dfm_server s, long long ticket, double* rel) {
  DFM_CHECK(s && rel, DFM_ERR_ARG, "dfm_server_true_residual: null argument");
  Server& sv = s->s;
  std::lock_guard<std::mutex> lk(sv.mu);
  auto it = sv.all.find(ticket);
  DFM_CHECK(it != sv.all.end(), DFM_ERR_ARG, "dfm_server_true_residual: unknown ticket");
  DFM_CHECK(it->second->state == Q_CONVERGED || it->second->state == Q_MAXITER, DFM_ERR_STATE,
            "dfm_server_true_residual: not finished");
  *rel = it->second->true_rel;
  return DFM_OK;
}

int dfm_server_result(dfm_server s, long long ticket, double* u) {
  DFM_CHECK(s && u, DFM_ERR_ARG, "dfm_server_result: null argument");
  Server& sv = s->s;
  Request* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    auto it = sv.all.find(ticket);
    DFM_CHECK(it != sv.all.end(), DFM_ERR_ARG, "dfm_server_result: unknown ticket");
    r = it->second;
  }
  DFM_CHECK(r->state == Q_CONVERGED || r->state == Q_MAXITER, DFM_ERR_STATE, "dfm_server_result: not finished");
  // read on the server's stream, where the solve ran (the system's own stream belongs to the worker that set it
  // up and may hold another request's setup)
  return sys_download(r->sys->s, -2, u, sv.ctx->c.stream);
}

int dfm_server_release(dfm_server s, long long ticket) {
  DFM_CHECK(s, DFM_ERR_ARG, "dfm_server_release: null server");
  Server& sv = s->s;
  Request* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    auto it = sv.all.find(ticket);
    DFM_CHECK(it != sv.all.end(), DFM_ERR_ARG, "dfm_server_release: unknown ticket");
    r = it->second;
    DFM_CHECK(r->state >= Q_CONVERGED, DFM_ERR_STATE, "dfm_server_release: request not finished");
    sv.all.erase(it);
    sv.bytes_admitted -= r->est_bytes;
    if (r->sys) sv.grave[(size_t)r->worker].push_back(r->sys);
    sv.admit_locked();
  }
  sv.cv_work.notify_all();
  delete r;
  return DFM_OK;
}

}  // extern "C"
