// Serving loop over the batch API: scene-bake jobs are admitted against a device-memory budget,
// keep their particle and grid state resident as tanks of one continuously batched table, advance
// one frame per tick together with every other running job, and retire when their frame count is
// reached. No counterpart in the reference (its page and bake.py run one scene at a time); each
// job follows the single-simulation semantics of one substep of its own dt per frame, so its
// result does not depend on the jobs it shares the device with.
//
// Host code only: it calls the library through its C ABI.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "suiri.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
  uint64_t id = 0;
  sl_job spec{};
  std::vector<float> objects;
  std::vector<sl_emit> fill;
  std::vector<sl_item> items;
  std::string bake_dir;
  uint32_t baked = 0;   // bake frames written
  int32_t status = SL_JOB_QUEUED;
  uint32_t slot = UINT32_MAX;
  uint32_t frames = 0;  // frames enqueued
  size_t bytes = 0;
  double t_submit = 0, t_admit = -1, t_done = -1;
  double particle_steps = 0;
  uint32_t count = 0, ww_count = 0;
  bool cancel = false;
};

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  double pos = q / 100.0 * (double)(v.size() - 1);
  size_t lo = (size_t)pos;
  size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct sl_server {
  sl_server_config cfg{};
  sl_batch *batch = nullptr;
  std::string err;
  std::map<uint64_t, Job> jobs;
  std::deque<uint64_t> queue;          // arrival order
  std::vector<uint64_t> slot_job;      // per batch slot: job id (0 = free)
  size_t used = 0;                     // device bytes of resident jobs
  uint64_t next_id = 1;
  std::vector<sl_job_info> finished;   // retired since the last poll
  Clock::time_point t0 = Clock::now();
  bool pending = false;                // a frame is enqueued and not yet waited for
  uint64_t ticks = 0, frames = 0;
  double particle_steps = 0;
  std::vector<double> step_ms;         // device time of each batch frame
  std::vector<double> latency;         // submit -> done of retired jobs [s]
  std::vector<sl_emit> emit_buf;
  std::vector<uint32_t> emit_counts;
  uint32_t stable = 0;                 // ticks since the membership last changed
  bool tuned = false;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
};

static void fill_info(const sl_server *s, const Job &j, sl_job_info *o) {
  memset(o, 0, sizeof *o);
  o->id = j.id;
  o->status = j.status;
  o->frames = j.frames;
  o->steps = j.spec.steps;
  o->t_submit = j.t_submit;
  o->t_admit = j.t_admit;
  o->t_done = j.t_done;
  o->particle_steps = j.particle_steps;
  o->bytes = j.bytes;
  o->count = j.count;
  o->ww_count = j.ww_count;
  o->slot = j.slot;
  (void)s;
}

extern "C" int sl_server_config_default(sl_server_config *c) {
  if (!c) return SL_EARG;
  memset(c, 0, sizeof *c);
  c->hbm_budget = 200ull << 30;  // 200 GiB of the device's 288 GB
  c->max_jobs = 128;
  c->device = 0;
  // no variant tuning by default: the one tuning runs on the jobs resident at that moment, and on a
  // stream of short jobs those are freshly filled tanks whose particles are still lattice-ordered, on
  // which chunked P2G variants win and which they then lose to the binned default as the particles
  // scatter (a stream of small scenes ran 45% slower per frame after such a choice)
  c->autotune_after = 0;
  c->graph_after = 0;
  c->frames_per_tick = 4;
  return SL_OK;
}

extern "C" int sl_server_create(const sl_server_config *cfg, sl_server **out) {
  if (!cfg || !out || !cfg->max_jobs || cfg->frames_per_tick > SL_SERVER_MAX_FRAMES_PER_TICK) return SL_EARG;
  *out = nullptr;
  sl_server *s = new sl_server();
  s->cfg = *cfg;
  int r = sl_batch_open(cfg->max_jobs, cfg->device, &s->batch);
  if (r) {
    delete s;
    return r;
  }
  sl_batch_timing_enable(s->batch, 1);
  s->slot_job.assign(cfg->max_jobs, 0);
  s->emit_counts.assign(cfg->max_jobs, 0);
  *out = s;
  return SL_OK;
}

extern "C" void sl_server_destroy(sl_server *s) {
  if (!s) return;
  if (s->batch) sl_batch_destroy(s->batch);
  delete s;
}

extern "C" const char *sl_server_last_error(const sl_server *s) { return s ? s->err.c_str() : ""; }

extern "C" int sl_server_submit(sl_server *s, const sl_job *job, uint64_t *id) {
  if (!s || !job) return SL_EARG;
  if (!(job->dt > 0) || !job->steps) return s->fail(SL_EARG, "sl_server_submit: steps and dt must be positive");
  if ((job->n_objects && !job->objects) || (job->n_fill && !job->fill) || (job->n_items && !job->items))
    return s->fail(SL_EARG, "sl_server_submit: null object, fill or item list");
  if (job->n_objects && job->n_items)
    return s->fail(SL_EARG, "sl_server_submit: give the obstacles as packed objects or as items, not both");
  if (job->cfg.device != s->cfg.device) return s->fail(SL_EARG, "sl_server_submit: job for another device");
  Job j;
  j.id = s->next_id++;
  j.spec = *job;
  j.objects.assign(job->objects, job->objects + (size_t)job->n_objects * 28);
  j.fill.assign(job->fill, job->fill + job->n_fill);
  if (job->n_items) j.items.assign(job->items, job->items + job->n_items);
  j.spec.objects = nullptr;
  j.spec.fill = nullptr;
  j.spec.items = nullptr;
  if (job->bake_dir) j.bake_dir = job->bake_dir;
  j.spec.bake_dir = nullptr;
  j.bytes = sl_job_bytes(&job->cfg);
  j.t_submit = s->now();
  if (id) *id = j.id;
  if (!j.bytes) {
    j.status = SL_JOB_REFUSED;
    j.t_done = j.t_submit;
    s->jobs[j.id] = j;
    return s->fail(SL_EARG, "sl_server_submit: invalid tank configuration");
  }
  if (j.bytes > s->cfg.hbm_budget) {  // can never fit: refused at once
    j.status = SL_JOB_REFUSED;
    j.t_done = j.t_submit;
    sl_job_info info;
    fill_info(s, j, &info);
    s->finished.push_back(info);
    s->jobs[j.id] = j;
    return s->fail(SL_EBUDGET, "sl_server_submit: the job alone exceeds the device-memory budget");
  }
  s->jobs[j.id] = j;
  s->queue.push_back(j.id);
  return SL_OK;
}

extern "C" int sl_server_cancel(sl_server *s, uint64_t id) {
  if (!s) return SL_EARG;
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return s->fail(SL_EARG, "sl_server_cancel: unknown job");
  Job &j = it->second;
  if (j.status == SL_JOB_QUEUED || j.status == SL_JOB_RUNNING) j.cancel = true;  // takes effect at the next tick
  return SL_OK;
}

// bake output of a job's tank at frame f (after the frame completed)
static int bake(sl_server *s, Job &j, uint32_t f) {
  int r = sl_bake_frame(sl_batch_sim(s->batch, j.slot), j.bake_dir.c_str(), f);
  if (r) return s->fail(r, "bake of job " + std::to_string(j.id) + " frame " + std::to_string(f) + " into " + j.bake_dir);
  j.baked++;
  return SL_OK;
}

// frees a resident job's tank
static int evict(sl_server *s, Job &j) {
  if (j.slot == UINT32_MAX) return SL_OK;
  int r = sl_batch_remove(s->batch, j.slot);
  s->slot_job[j.slot] = 0;
  j.slot = UINT32_MAX;
  s->used -= j.bytes;
  s->stable = 0;
  return r;
}

static void retire(sl_server *s, Job &j, int32_t status, double t) {
  j.status = status;
  j.t_done = t;
  sl_job_info info;
  fill_info(s, j, &info);
  s->finished.push_back(info);
  if (status == SL_JOB_DONE) s->latency.push_back(j.t_done - j.t_submit);
}

extern "C" int sl_server_tick(sl_server *s, uint32_t *running) {
  if (!s) return SL_EARG;
  int r;
  // 1. the frame enqueued by the previous tick has completed
  if (s->pending) {
    if ((r = sl_batch_sync(s->batch))) return s->fail(r, sl_batch_last_error(s->batch));
    // the device time of every frame the previous tick enqueued (at most frames_per_tick)
    double ms[SL_SERVER_MAX_FRAMES_PER_TICK];
    uint32_t n = 0;
    sl_batch_timing_read(s->batch, ms, SL_SERVER_MAX_FRAMES_PER_TICK, &n);
    for (uint32_t i = 0; i < n; i++) s->step_ms.push_back(ms[i]);
    s->pending = false;
  }
  const double t = s->now();
  // 2. retire finished and cancelled jobs
  for (uint32_t slot = 0; slot < s->cfg.max_jobs; slot++) {
    if (!s->slot_job[slot]) continue;
    Job &j = s->jobs[s->slot_job[slot]];
    if (j.status != SL_JOB_RUNNING) continue;
    sl_sim *sim = sl_batch_sim(s->batch, slot);
    double v = 0;
    if (sl_get_scalar(sim, "count", &v) == SL_OK) j.count = (uint32_t)v;
    if (sl_get_scalar(sim, "ww_count", &v) == SL_OK) j.ww_count = (uint32_t)v;
    if (!j.cancel && !j.bake_dir.empty() && j.frames &&
        (j.frames >= j.spec.steps || (j.spec.bake_every && j.frames % j.spec.bake_every == 0)) &&
        (r = bake(s, j, j.frames)))
      return r;
    if (j.cancel) {
      evict(s, j);
      retire(s, j, SL_JOB_CANCELLED, t);
    } else if (j.frames >= j.spec.steps) {
      if (j.spec.keep) {
        sl_batch_set_stepped(s->batch, slot, 0);  // stays resident for readback until released
        s->stable = 0;
      } else {
        evict(s, j);
      }
      retire(s, j, SL_JOB_DONE, t);
    }
  }
  // 3. admit queued jobs that fit, first fit in arrival order
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Job &j = s->jobs[*it];
    if (j.cancel) {
      retire(s, j, SL_JOB_CANCELLED, t);
      it = s->queue.erase(it);
      continue;
    }
    if (s->used + j.bytes > s->cfg.hbm_budget) { ++it; continue; }
    uint32_t slot = 0;
    r = sl_batch_add(s->batch, &j.spec.cfg, &slot);
    if (r == SL_ECAP) break;  // every slot is taken
    if (r) {                  // the device refused the allocation: the job cannot run
      retire(s, j, SL_JOB_REFUSED, t);
      it = s->queue.erase(it);
      continue;
    }
    sl_sim *sim = sl_batch_sim(s->batch, slot);
    const uint32_t ni = (uint32_t)j.items.size();
    std::vector<sl_emit> blocks(ni);
    uint32_t nb = 0;
    if ((r = ni ? sl_set_items(sim, j.items.data(), ni, 0, nullptr)
                : sl_set_objects(sim, j.objects.data(), (uint32_t)(j.objects.size() / 28), nullptr)) ||
        (j.fill.size() && (r = sl_emit_now(sim, j.fill.data(), (uint32_t)j.fill.size(), 0, nullptr))) ||
        (ni && (r = sl_scene_block_emits(sl_sim_scene(sim), j.items.data(), ni, 0, blocks.data(), ni, &nb))) ||
        (nb && (r = sl_emit_now(sim, blocks.data(), nb, 0, nullptr))) ||
        (r = sl_batch_set_dt(s->batch, slot, j.spec.dt))) {
      std::string m = sl_last_error(sim);
      sl_batch_remove(s->batch, slot);
      retire(s, j, SL_JOB_REFUSED, t);
      s->queue.erase(it);
      return s->fail(r, "admission: " + m);
    }
    j.slot = slot;
    j.status = SL_JOB_RUNNING;
    j.t_admit = t;
    s->slot_job[slot] = j.id;
    s->used += j.bytes;
    s->stable = 0;
    it = s->queue.erase(it);
    if (!j.bake_dir.empty() && (r = bake(s, j, 0))) return r;  // the start frame, like bake.py:208-214
  }
  // 4. the frames of this tick: one, or up to frames_per_tick back to back while no running job
  // reaches its last frame or a bake frame inside them (the host work of a tick then overlaps the
  // frames already enqueued instead of leaving the device idle after every frame)
  uint32_t burst = std::max<uint32_t>(1u, s->cfg.frames_per_tick), nrun = 0;
  for (uint32_t slot = 0; slot < s->cfg.max_jobs; slot++) {
    if (!s->slot_job[slot]) continue;
    const Job &j = s->jobs[s->slot_job[slot]];
    if (j.status != SL_JOB_RUNNING) continue;
    nrun++;
    burst = std::min(burst, j.spec.steps > j.frames ? j.spec.steps - j.frames : 1u);
    if (!j.bake_dir.empty() && j.spec.bake_every) burst = std::min(burst, j.spec.bake_every - j.frames % j.spec.bake_every);
  }
  for (uint32_t f = 0; nrun && f < burst; f++) {
    // per-frame emission of every running job with a source, in one batched emission
    s->emit_buf.clear();
    bool any_emit = false;
    for (uint32_t slot = 0; slot < s->cfg.max_jobs; slot++) {
      s->emit_counts[slot] = 0;
      if (!s->slot_job[slot]) continue;
      Job &j = s->jobs[s->slot_job[slot]];
      if (j.status != SL_JOB_RUNNING || (!j.spec.source && j.items.empty())) continue;
      sl_sim *sim = sl_batch_sim(s->batch, slot);
      sl_emit list[SL_MAX_EMIT];
      uint32_t n = 0;
      if (j.spec.source) {
        double time = 0;
        sl_get_scalar(sim, "time", &time);
        n = std::min<uint32_t>(j.spec.source(j.spec.user, sim, j.spec.dt, time, list, SL_MAX_EMIT), SL_MAX_EMIT);
      }
      if (!j.items.empty()) {
        uint32_t m = 0;
        if ((r = sl_scene_source_emits(sl_sim_scene(sim), j.items.data(), (uint32_t)j.items.size(), j.spec.dt,
                                       list + n, SL_MAX_EMIT - n, &m)))
          return s->fail(r, "sources of job " + std::to_string(j.id));
        n += m;
      }
      s->emit_buf.insert(s->emit_buf.end(), list, list + n);
      s->emit_counts[slot] = n;
      any_emit |= n > 0;
    }
    if (any_emit && (r = sl_batch_emit(s->batch, s->emit_buf.data(), s->emit_counts.data())))
      return s->fail(r, sl_batch_last_error(s->batch));
    // one frame of every running job
    if (f == 0) {
      if (s->cfg.autotune_after && !s->tuned && s->stable >= s->cfg.autotune_after) {
        uint32_t ch[6];
        if ((r = sl_batch_autotune(s->batch, 1, ch))) return s->fail(r, sl_batch_last_error(s->batch));
        s->tuned = true;
      }
      if (s->cfg.graph_after && s->stable == s->cfg.graph_after && (r = sl_batch_capture(s->batch, 0)))
        return s->fail(r, sl_batch_last_error(s->batch));
    }
    if ((r = sl_batch_step(s->batch, 0))) return s->fail(r, sl_batch_last_error(s->batch));
    s->pending = true;
    s->frames++;
    for (uint32_t slot = 0; slot < s->cfg.max_jobs; slot++) {
      if (!s->slot_job[slot]) continue;
      Job &j = s->jobs[s->slot_job[slot]];
      if (j.status != SL_JOB_RUNNING) continue;
      j.frames++;
      double v = 0;
      sl_get_scalar(sl_batch_sim(s->batch, slot), "count", &v);
      j.particle_steps += v;
      s->particle_steps += v;
    }
  }
  s->stable++;
  s->ticks++;
  if (running) *running = nrun;
  return SL_OK;
}

extern "C" int sl_server_poll(sl_server *s, sl_job_info *out, uint32_t max, uint32_t *n) {
  if (!s || (max && !out)) return SL_EARG;
  uint32_t k = std::min<uint32_t>(max, (uint32_t)s->finished.size());
  for (uint32_t i = 0; i < k; i++) out[i] = s->finished[i];
  s->finished.erase(s->finished.begin(), s->finished.begin() + k);
  if (n) *n = k;
  return SL_OK;
}

extern "C" int sl_server_info(const sl_server *s, uint64_t id, sl_job_info *out) {
  if (!s || !out) return SL_EARG;
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return SL_EARG;
  fill_info(s, it->second, out);
  return SL_OK;
}

extern "C" sl_sim *sl_server_job_sim(sl_server *s, uint64_t id) {
  if (!s) return nullptr;
  auto it = s->jobs.find(id);
  if (it == s->jobs.end() || it->second.slot == UINT32_MAX) return nullptr;
  if (s->pending) {  // readers see the completed frame
    if (sl_batch_sync(s->batch)) return nullptr;
  }
  return sl_batch_sim(s->batch, it->second.slot);
}

extern "C" int sl_server_release(sl_server *s, uint64_t id) {
  if (!s) return SL_EARG;
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return s->fail(SL_EARG, "sl_server_release: unknown job");
  Job &j = it->second;
  if (j.status == SL_JOB_RUNNING || j.status == SL_JOB_QUEUED)
    return s->fail(SL_EARG, "sl_server_release: the job has not finished");
  int r = evict(s, j);
  return r ? s->fail(r, sl_batch_last_error(s->batch)) : SL_OK;
}

extern "C" int sl_server_stats(const sl_server *s, sl_server_stats_t *o) {
  if (!s || !o) return SL_EARG;
  memset(o, 0, sizeof *o);
  o->budget = s->cfg.hbm_budget;
  o->used = s->used;
  o->queued = (uint32_t)s->queue.size();
  for (uint32_t slot = 0; slot < s->cfg.max_jobs; slot++) {
    if (!s->slot_job[slot]) continue;
    o->resident++;
    if (s->jobs.at(s->slot_job[slot]).status == SL_JOB_RUNNING) o->running++;
  }
  o->ticks = s->ticks;
  o->frames = s->frames;
  o->particle_steps = s->particle_steps;
  o->uptime = s->now();
  o->step_ms_p50 = pct(s->step_ms, 50);
  o->step_ms_p99 = pct(s->step_ms, 99);
  o->step_ms_sum = 0;
  for (double v : s->step_ms) o->step_ms_sum += v;
  o->latency_p50 = pct(s->latency, 50);
  o->latency_p99 = pct(s->latency, 99);
  o->done = (uint32_t)s->latency.size();
  return SL_OK;
}
