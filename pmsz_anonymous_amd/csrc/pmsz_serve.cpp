// pmsz_serve.cpp -- the correction server of libpmsz.so (include/pmsz.h,
// pmsz_server_*, pmsz_submit, pmsz_poll): continuous batching of correction jobs
// over lanes, admitted by an HBM budget.
//
// The reference corrects one field per process run (main, pMSz.cu:2596-3188).
// A server keeps `lanes` correction contexts (pmsz_ctx: own HIP stream, scratch
// state and lock) and one host thread per lane.  A lane that is free takes the
// oldest queued job it can admit and, when that job is small (fewer than
// alone_voxels voxels), every further queued small job of the same value type
// and loop, up to max_batch jobs, while the library-owned device memory of all
// lanes -- each running lane's workspace (pmsz_workspace_bytes) or what its
// context already holds if that is more, idle lanes' holdings included -- stays
// within hbm_budget.  The group is corrected as one batched run
// (pmsz_ctx_correct_batch); a job of alone_voxels or more runs by itself and
// takes the one-field schedule.  Runs on different lanes overlap on the device:
// the full-volume passes of one run fill the GPU while another is in its
// latency-bound tail.  An idle lane whose holdings block an admission frees them
// (pmsz_ctx_trim).
//
// Each result equals the job corrected alone: a batched run is exact per field,
// and lanes share no state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pmsz.h"

namespace {

double now_s() {  // CLOCK_MONOTONIC seconds, the clock of Python's time.perf_counter
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct ServeJob {
  uint64_t id = 0;
  pmsz_params p{};
  const void *f = nullptr;
  void *g = nullptr;
  uint8_t *dc = nullptr;
  hipEvent_t ready = nullptr;  // the submitter's stream reached the submission
  int64_t voxels = 0;
  bool done = false;
  int rc = PMSZ_OK;
  pmsz_stats st{};
  pmsz_job_info info{};
};

bool same_key(const ServeJob &a, const ServeJob &b) {
  return a.p.dtype == b.p.dtype && a.p.mode == b.p.mode && a.p.max_sweeps == b.p.max_sweeps;
}

}  // namespace

struct pmsz_server {
  pmsz_server_config cfg{};
  int device = -1;
  std::vector<pmsz_ctx *> ctxs;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServeJob *> queue;
  std::unordered_map<uint64_t, ServeJob *> jobs;
  std::vector<int64_t> reserved, held;
  uint64_t next_id = 1;
  bool stopping = false;
  int64_t runs = 0, jobs_done = 0;

  // The jobs lane `lane` admits now, oldest first (caller holds mu).  Nothing
  // is admitted only when the queue is empty or the budget blocks its head.
  std::vector<ServeJob *> take(int lane) {
    std::vector<ServeJob *> group;
    if (queue.empty()) return group;
    int64_t others = 0;
    for (size_t k = 0; k < ctxs.size(); k++)
      if ((int)k != lane) others += std::max(reserved[k], held[k]);
    const int64_t have = held[lane];
    ServeJob *head = queue.front();
    const bool alone = head->voxels >= cfg.alone_voxels;
    std::vector<pmsz_params> shapes;
    for (ServeJob *j : queue) {
      if (j != head && (alone || !same_key(*j, *head) || j->voxels >= cfg.alone_voxels))
        continue;
      shapes.push_back(j->p);
      const int64_t ws = pmsz_workspace_bytes((int)shapes.size(), shapes.data());
      if (ws < 0 || others + std::max(have, ws) > cfg.hbm_budget) {
        shapes.pop_back();
        break;
      }
      group.push_back(j);
      if (alone || (int)group.size() >= cfg.max_batch) break;
    }
    if (!group.empty()) {
      for (ServeJob *j : group) queue.erase(std::find(queue.begin(), queue.end(), j));
      reserved[lane] = std::max(have, pmsz_workspace_bytes((int)shapes.size(), shapes.data()));
    }
    return group;
  }

  void lane_loop(int lane) {
    (void)hipSetDevice(device);
    pmsz_ctx *ctx = ctxs[lane];
    for (;;) {
      std::vector<ServeJob *> group;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          group = take(lane);
          if (!group.empty()) break;
          if (stopping && queue.empty()) return;
          // the budget blocks the queue's head: memory this idle lane holds may be
          // what blocks it
          if (!queue.empty() && held[lane]) break;
          cv_work.wait_for(lk, std::chrono::milliseconds(20));
        }
        if (group.empty()) {  // free this idle lane's buffers, then look again
          lk.unlock();
          (void)pmsz_ctx_trim(ctx);
          lk.lock();
          held[lane] = 0;
          cv_work.notify_all();
          continue;
        }
      }
      const double t0 = now_s();
      const int n = (int)group.size();
      std::vector<pmsz_params> ps(n);
      std::vector<const void *> fs(n);
      std::vector<void *> gs(n);
      std::vector<uint8_t *> dcs(n);
      std::vector<pmsz_stats> sts(n);
      bool any_dc = false;
      for (int k = 0; k < n; k++) {
        ServeJob *j = group[k];
        if (j->ready) (void)hipEventSynchronize(j->ready);  // the inputs are written
        ps[k] = j->p;
        fs[k] = j->f;
        gs[k] = j->g;
        dcs[k] = j->dc;
        any_dc = any_dc || j->dc;
      }
      const int rc = pmsz_ctx_correct_batch(ctx, n, ps.data(), fs.data(), gs.data(),
                                            any_dc ? dcs.data() : nullptr, sts.data());
      const double t1 = now_s();
      const int64_t h = pmsz_ctx_bytes(ctx);
      std::lock_guard<std::mutex> lk(mu);
      for (int k = 0; k < n; k++) {
        ServeJob *j = group[k];
        j->rc = rc == PMSZ_OK ? sts[k].status : rc;
        j->st = sts[k];
        j->info.status = j->rc;
        j->info.batch_size = n;
        j->info.lane = lane;
        j->info.t_start = t0;
        j->info.t_done = t1;
        j->done = true;
      }
      reserved[lane] = 0;
      held[lane] = h > 0 ? h : 0;
      runs++;
      jobs_done += n;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" {

void pmsz_default_server_config(pmsz_server_config *c) {
  if (!c) return;
  c->lanes = 4;
  c->max_batch = 32;
  c->hbm_budget = (int64_t)64 << 30;
  c->alone_voxels = (int64_t)1 << 20;
}

int pmsz_server_create(const pmsz_server_config *cfg, pmsz_server **out) {
  if (!cfg || !out || cfg->lanes < 1 || cfg->lanes > 7 || cfg->max_batch < 1 ||
      cfg->hbm_budget <= 0 || cfg->alone_voxels < 0)
    return PMSZ_EINVAL;
  pmsz_server *s = new (std::nothrow) pmsz_server();
  if (!s) return PMSZ_ENOMEM;
  s->cfg = *cfg;
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return PMSZ_EDEVICE;
  }
  for (int k = 0; k < cfg->lanes; k++) {
    pmsz_ctx *c = nullptr;
    const int rc = pmsz_ctx_create(&c);
    if (rc) {
      for (pmsz_ctx *d : s->ctxs) pmsz_ctx_destroy(d);
      delete s;
      return rc;
    }
    s->ctxs.push_back(c);
  }
  s->reserved.assign(cfg->lanes, 0);
  s->held.assign(cfg->lanes, 0);
  for (int k = 0; k < cfg->lanes; k++) s->threads.emplace_back([s, k] { s->lane_loop(k); });
  *out = s;
  return PMSZ_OK;
}

int pmsz_server_destroy(pmsz_server *s) {
  if (!s) return PMSZ_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (std::thread &t : s->threads) t.join();
  for (pmsz_ctx *c : s->ctxs) pmsz_ctx_destroy(c);
  for (auto &kv : s->jobs) {
    if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
    delete kv.second;
  }
  delete s;
  return PMSZ_OK;
}

int pmsz_submit(pmsz_server *s, const pmsz_params *p, const void *d_f, void *d_g,
                uint8_t *d_dc, void *stream, uint64_t *job) {
  if (!s || !p || !d_f || !d_g || !job) return PMSZ_EINVAL;
  if (p->sub[0] * p->sub[1] * p->sub[2] != 1 || p->mode == PMSZ_NAIVE) return PMSZ_EINVAL;
  const int64_t ws = pmsz_workspace_bytes(1, p);
  if (ws < 0) return (int)ws;
  if (ws > s->cfg.hbm_budget) return PMSZ_EBUSY;  // could never be admitted
  ServeJob *j = new (std::nothrow) ServeJob();
  if (!j) return PMSZ_ENOMEM;
  j->p = *p;
  j->f = d_f;
  j->g = d_g;
  j->dc = d_dc;
  j->voxels = p->nx * p->ny * p->nz;
  j->info.t_submit = now_s();
  if (hipEventCreateWithFlags(&j->ready, hipEventDisableTiming) != hipSuccess ||
      hipEventRecord(j->ready, (hipStream_t)stream) != hipSuccess) {
    if (j->ready) (void)hipEventDestroy(j->ready);
    delete j;
    return PMSZ_EDEVICE;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) {
      (void)hipEventDestroy(j->ready);
      delete j;
      return PMSZ_EBUSY;
    }
    j->id = s->next_id++;
    s->jobs[j->id] = j;
    s->queue.push_back(j);
    *job = j->id;
  }
  s->cv_work.notify_all();
  return PMSZ_OK;
}

int pmsz_poll(pmsz_server *s, uint64_t job, int timeout_ms, pmsz_stats *stats,
              pmsz_job_info *info) {
  if (!s) return PMSZ_EINVAL;
  ServeJob *j = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return PMSZ_EINVAL;
    j = it->second;
    if (!j->done) {
      if (timeout_ms < 0)
        s->cv_done.wait(lk, [&] { return j->done; });
      else if (timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return j->done; });
    }
    if (!j->done) return PMSZ_EBUSY;
    s->jobs.erase(it);
  }
  if (stats) *stats = j->st;
  if (info) *info = j->info;
  if (j->ready) (void)hipEventDestroy(j->ready);
  delete j;
  return PMSZ_OK;
}

int pmsz_server_stats(pmsz_server *s, int64_t *runs, int64_t *jobs_done, int64_t *queued) {
  if (!s) return PMSZ_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (runs) *runs = s->runs;
  if (jobs_done) *jobs_done = s->jobs_done;
  if (queued) *queued = (int64_t)s->queue.size();
  return PMSZ_OK;
}

}  // extern "C"
