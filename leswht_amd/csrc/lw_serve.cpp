// A request loop over independent cases sharing one GPU.
//
// The reference runs one case per process (run_solver.sh:41: solver_exec with
// a case directory as cwd). Here a server takes such case directories as
// requests while others run: each request is admitted when a slot is free and
// its estimated device footprint fits the memory budget, then a worker thread
// creates its handle (fields resident in HBM for the whole run), steps it on
// the handle's own streams, writes the final fld file and frees the memory.
// Kernels of admitted cases overlap on the device because every handle
// launches on its own non-blocking streams from its own host thread; each case
// runs exactly the kernels and the arithmetic it would run alone, so its
// results do not depend on the mix.
//
// Admission is first-come first-served: the oldest queued request is admitted
// next, and it waits (blocking those behind it) until enough of the budget
// is free, so a large case is never starved by a stream of small ones.
//
// Continuous batching (lw_server_create_batched): a small case, once its
// worker has created it, joins the server's batch instead of being stepped
// by the worker. One batch thread steps every joined case together
// (lw_batch_step: each kernel launched once for all of them), one step at a
// time; between those steps it takes in the cases that joined and lets go
// of those that reached their step count (their fields are written, their
// memory freed). The worker is free again as soon as its case has joined.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/leswht.h"

void lw_set_last_error(const std::string &msg);

namespace {

struct Request {
  std::string dir, fld_out;
  lw_body body{};
  int64_t nsteps = 0;
  lw_case_result res{};
  std::string err;
  lw_handle *h = nullptr;  // a batched case's handle while it is in the batch
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct lw_server {
  int64_t budget = 0, nslots = 1;
  int64_t batch_max_cells = 0;
  std::vector<Request *> joining;  // batched cases created and waiting to join
  std::thread batcher;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::vector<std::unique_ptr<Request>> all;  // index = ticket
  int64_t reserved = 0, running = 0, peak_reserved = 0, peak_running = 0;
  bool stopping = false;
  std::vector<std::thread> workers;

  // Runs one admitted case without holding mu. Everything it learns goes
  // into `out` and `err`, which the worker publishes under mu together with
  // the completion time, so a poll or wait never sees a half-written result.
  struct Outcome {
    int rc = LW_OK;
    int64_t steps_done = 0, mg_cycles_sum = 0, device_bytes = 0;
    double time = 0.0, dt = 0.0;
    std::string err;
  };
  static Outcome run_case(const Request &r) {
    Outcome o;
    lw_handle *h = nullptr;
    lw_body body = r.body;
    o.rc = lw_create_from_dir(r.dir.c_str(), &body, &h);
    if (o.rc == LW_OK) {
      o.device_bytes = lw_device_bytes(h);
      lw_step_stats st{};
      for (int64_t n = 0; n < r.nsteps && o.rc == LW_OK; ++n) {
        o.rc = lw_step(h, 1, &st);
        if (o.rc == LW_OK) {
          o.steps_done = n + 1;
          o.mg_cycles_sum += st.mg_cycles_sum;
          o.time = st.time;
          o.dt = st.dt;
        }
      }
      if (o.rc == LW_OK && !r.fld_out.empty()) o.rc = lw_write_field(h, r.fld_out.c_str());
    }
    if (o.rc != LW_OK) o.err = lw_last_error();
    if (h) lw_destroy(h);
    return o;
  }

  // with mu held: a case's end (done or failed), its budget released
  void finish_locked(Request *r, int rc, const std::string &err) {
    r->err = err;
    r->res.t_done = seconds_since(t0);
    r->res.status = rc == LW_OK ? LW_CASE_DONE : rc;
    reserved -= r->res.reserved_bytes;
    running -= 1;
    cv_done.notify_all();
    cv_work.notify_all();
  }

  std::condition_variable cv_batch;

  // The batch thread: steps the joined cases together, takes in new ones
  // and lets go of finished ones between calls of lw_batch_step.
  void batch_loop() {
    std::vector<Request *> act;
    lw_batch *b = nullptr;
    bool changed = false;
    std::vector<lw_step_stats> st;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_batch.wait(lk, [&] { return !joining.empty() || !act.empty() || (stopping && queue.empty() && running == 0); });
      if (act.empty() && joining.empty()) break;
      if (!joining.empty()) {
        act.insert(act.end(), joining.begin(), joining.end());
        joining.clear();
        changed = true;
      }
      lk.unlock();
      // one step per call: lw_batch_step's statistics are those of the
      // call's last step, and cases join and leave between calls
      int64_t k = 1;
      for (Request *r : act) k = std::min(k, r->nsteps - r->res.steps_done);
      int rc = LW_OK;
      std::string err;
      if (changed) {
        if (b) lw_batch_destroy(b);
        b = nullptr;
        std::vector<lw_handle *> hs;
        for (Request *r : act) hs.push_back(r->h);
        rc = lw_batch_create(hs.data(), (int64_t)hs.size(), &b);
        changed = false;
      }
      if (rc == LW_OK && k > 0) {
        st.assign(act.size(), lw_step_stats{});
        rc = lw_batch_step(b, k, st.data());
      }
      if (rc != LW_OK) err = lw_last_error();
      // results of the step (published under mu below; this thread is the
      // only writer of a batched case's result); finished or failed cases
      // leave the batch
      std::vector<Request *> keep, out;
      for (size_t i = 0; i < act.size(); ++i) {
        Request *r = act[i];
        const int64_t done = r->res.steps_done + (rc == LW_OK ? k : 0);
        (rc != LW_OK || done >= r->nsteps ? out : keep).push_back(r);
      }
      if (!out.empty()) {
        if (b) lw_batch_destroy(b);
        b = nullptr;
        changed = true;
      }
      std::vector<std::pair<int, std::string>> ends;
      for (Request *r : out) {
        int rr = rc;
        std::string e = err;
        if (rr == LW_OK && !r->fld_out.empty()) {
          rr = lw_write_field(r->h, r->fld_out.c_str());
          if (rr != LW_OK) e = lw_last_error();
        }
        lw_destroy(r->h);
        r->h = nullptr;
        ends.emplace_back(rr, e);
      }
      lk.lock();
      if (rc == LW_OK && k > 0)
        for (size_t i = 0; i < act.size(); ++i) {
          Request *r = act[i];
          r->res.steps_done += k;
          r->res.mg_cycles_sum += st[i].mg_cycles_sum;
          r->res.time = st[i].time;
          r->res.dt = st[i].dt;
        }
      for (size_t q = 0; q < out.size(); ++q) finish_locked(out[q], ends[q].first, ends[q].second);
      act.swap(keep);
    }
    if (b) lw_batch_destroy(b);
  }

  void worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_work.wait(lk, [&] {
        return (!queue.empty() && reserved + queue.front()->res.reserved_bytes <= budget) ||
               (stopping && queue.empty());
      });
      if (queue.empty()) return;
      Request *r = queue.front();
      queue.pop_front();
      reserved += r->res.reserved_bytes;
      running += 1;
      if (reserved > peak_reserved) peak_reserved = reserved;
      if (running > peak_running) peak_running = running;
      r->res.status = LW_CASE_RUNNING;
      r->res.t_admit = seconds_since(t0);
      lk.unlock();
      if (batch_max_cells > 0 && r->res.cells <= batch_max_cells) {
        // a batched case: create it and hand it to the batch thread
        lw_body body = r->body;
        lw_handle *h = nullptr;
        const int rc = lw_create_from_dir(r->dir.c_str(), &body, &h);
        const std::string err = rc == LW_OK ? std::string() : std::string(lw_last_error());
        lk.lock();
        if (rc == LW_OK) {
          r->h = h;
          r->res.device_bytes = lw_device_bytes(h);
          joining.push_back(r);
          cv_batch.notify_all();
        } else {
          finish_locked(r, rc, err);
        }
        continue;
      }
      Outcome o = run_case(*r);
      lk.lock();
      r->res.steps_done = o.steps_done;
      r->res.mg_cycles_sum = o.mg_cycles_sum;
      r->res.device_bytes = o.device_bytes;
      r->res.time = o.time;
      r->res.dt = o.dt;
      finish_locked(r, o.rc, o.err);
    }
  }
};

extern "C" {

int lw_server_create_batched(int64_t mem_budget_bytes, int64_t slots, int64_t batch_max_cells,
                             lw_server **out) {
  if (!out || mem_budget_bytes <= 0 || slots < 1 || slots > 64 || batch_max_cells < 0) {
    lw_set_last_error("lw_server_create: budget must be > 0, slots in 1..64 and batch_max_cells >= 0");
    return LW_ERR_ARG;
  }
  auto *s = new lw_server();
  s->budget = mem_budget_bytes;
  s->nslots = slots;
  s->batch_max_cells = batch_max_cells;
  s->t0 = std::chrono::steady_clock::now();
  for (int64_t q = 0; q < slots; ++q) s->workers.emplace_back([s] { s->worker(); });
  if (batch_max_cells > 0) s->batcher = std::thread([s] { s->batch_loop(); });
  *out = s;
  return LW_OK;
}

int lw_server_create(int64_t mem_budget_bytes, int64_t slots, lw_server **out) {
  return lw_server_create_batched(mem_budget_bytes, slots, 0, out);
}

int lw_server_submit(lw_server *s, const char *solver_dir, const lw_body *body, int64_t nsteps,
                     const char *fld_out, int64_t *ticket) {
  if (!s || !solver_dir || !ticket || nsteps < 0) {
    lw_set_last_error("lw_server_submit: null argument or negative step count");
    return LW_ERR_ARG;
  }
  auto r = std::make_unique<Request>();
  r->res.t_submit = seconds_since(s->t0);
  r->dir = solver_dir;
  r->fld_out = fld_out ? fld_out : "";
  if (body) r->body = *body; else r->body.p[0] = 0.5;
  r->nsteps = nsteps;
  lw_settings st{};
  int64_t nm[3];
  int rc = lw_probe_dir(solver_dir, &st, nm);
  if (rc != LW_OK) return rc;
  const int64_t est = lw_estimate_device_bytes(&st, nm[0], nm[1], nm[2]);
  if (est < 0) return (int)est;
  if (est > s->budget) {
    lw_set_last_error("lw_server_submit: the case needs about " + std::to_string(est >> 20) +
                      " MiB, more than the whole budget");
    return LW_ERR_ARG;
  }
  r->res.cells = nm[0] * nm[1] * nm[2];
  r->res.reserved_bytes = est;
  r->res.status = LW_CASE_QUEUED;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stopping) {
    lw_set_last_error("lw_server_submit: server is shutting down");
    return LW_ERR_ARG;
  }
  r->res.ticket = (int64_t)s->all.size();
  *ticket = r->res.ticket;
  s->queue.push_back(r.get());
  s->all.push_back(std::move(r));
  s->cv_work.notify_all();
  return LW_OK;
}

int lw_server_poll(lw_server *s, int64_t ticket, lw_case_result *res) {
  if (!s || !res) return LW_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (ticket < 0 || ticket >= (int64_t)s->all.size()) {
    lw_set_last_error("lw_server_poll: unknown ticket");
    return LW_ERR_ARG;
  }
  *res = s->all[ticket]->res;
  if (res->status < 0) lw_set_last_error(s->all[ticket]->err);
  return LW_OK;
}

int lw_server_wait(lw_server *s, int64_t ticket, lw_case_result *res) {
  if (!s || !res) return LW_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (ticket < 0 || ticket >= (int64_t)s->all.size()) {
    lw_set_last_error("lw_server_wait: unknown ticket");
    return LW_ERR_ARG;
  }
  Request *r = s->all[ticket].get();
  s->cv_done.wait(lk, [&] { return r->res.status != LW_CASE_QUEUED && r->res.status != LW_CASE_RUNNING; });
  *res = r->res;
  if (res->status < 0) lw_set_last_error(r->err);
  return LW_OK;
}

int lw_server_stats(lw_server *s, lw_server_info *info) {
  if (!s || !info) return LW_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  info->submitted = (int64_t)s->all.size();
  info->queued = (int64_t)s->queue.size();
  info->running = s->running;
  info->reserved_bytes = s->reserved;
  info->peak_reserved_bytes = s->peak_reserved;
  info->peak_running = s->peak_running;
  info->budget_bytes = s->budget;
  info->slots = s->nslots;
  info->elapsed = seconds_since(s->t0);
  return LW_OK;
}

int lw_server_destroy(lw_server *s) {
  if (!s) return LW_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->cv_work.notify_all();
  }
  for (auto &t : s->workers) t.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cv_batch.notify_all();
  }
  if (s->batcher.joinable()) s->batcher.join();
  delete s;
  return LW_OK;
}

}  // extern "C"
