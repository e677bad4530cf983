// Host driver of the suiri library: the C ABI of include/suiri.h.
//
// Restates the orchestration of the reference's FluidSim (src/js/50-sim.js) and RippleLayer
// (src/js/64-ripple.js:149-267) on one HIP stream per simulation. The grid sizing and the parameter
// values follow the reference's double-precision JavaScript arithmetic, rounded to f32 where the
// reference writes them into its uniform buffer.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include "../../include/suiri.h"
#include "sl_internal.h"
#include "sl_scene.h"

static thread_local std::string g_err;
// the message of sl_last_error(NULL) for failures of entry points without a simulation
void sl_set_error(const char *msg) { g_err = msg ? msg : ""; }

// dynamic per-substep fields of a tank, written on the device by k_update (kernel arguments are
// captured at launch, so the host can change them freely while earlier work is in flight)
struct SlTankDyn {
  double time_d, dt_frame;
  float dt, gdt, time_f, rLx, rLz, rdt, rmd;
  uint32_t frame, obj_count, emit_k, emit_total, iters;
  int32_t mask_bound;
};
__global__ void k_update(SlTank *T, SlTankDyn d) {
  if (threadIdx.x) return;
  T->dt = d.dt; T->gdt = d.gdt; T->time_f = d.time_f;
  T->rLx = d.rLx; T->rLz = d.rLz; T->rdt = d.rdt; T->rmd = d.rmd;
  T->frame = d.frame; T->obj_count = d.obj_count; T->emit_k = d.emit_k; T->emit_total = d.emit_total;
  T->iters = d.iters; T->mask_bound = d.mask_bound;
  T->time_d = d.time_d;
  T->dt_frame = d.dt_frame;
}

// pinned staging for host->device uploads issued on the stream: two halves, each reused only after
// the event recorded when the stream last left it has completed
struct Staging {
  uint8_t *buf = nullptr;
  size_t half = 0, off = 0;
  int cur = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ev_valid[2] = {false, false};
};

struct sl_sim {
  sl_config cfg;
  // set when the simulation is a tank of a batch: the batch owns the stream, the table slot of the
  // tank record and the counters slot
  hipStream_t ext_stream = nullptr;
  SlTank *ext_tank = nullptr;
  SlCounters *ext_counters = nullptr;
  SlEmitDev *ext_emits = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  uint32_t gpu_errors = 0;
  // host-side FluidSim state (50-sim.js:29-58)
  int nx = 0, ny = 0, nz = 0;
  double h = 0, origin[3] = {0, 0, 0}, rim = 0, vp = 0, radius = 0, time = 0, vmax = 1.0;
  uint32_t nodes = 0, cells = 0, hw = 0, hcells = 0, cap = 0, ww_cap = 0, R = 0, M = 0, iters = 0;
  uint32_t frame = 0, obj_count = 0, count = 0, ww_count = 0, impacts = 0, frames = 0;
  bool full = false, geom_dirty = true, mask_dirty = true;  // mask_dirty: objMask out of date
  int cur = 0, ww_cur = 0;
  uint32_t forced_substeps = 0, substeps_last = 0;
  uint32_t pending_add = 0;  // particles emitted in this substep, committed by its cell-count clear
  // spatial re-ordering
  uint32_t *sort_keys = nullptr, *sort_vals = nullptr;
  void *sort_temp = nullptr;
  size_t sort_temp_bytes = 0;
  int sort_end_bit = 1;
  uint32_t reorder_clock = 0;
  double batch_dt = 0;    // frame length of a batched tank stepped with dt = 0
  size_t bytes = 0;       // device memory held (sum of the allocation sizes)
  int p2g_variant = SL_P2G_BINNED;  // kernel variant (binned, or chunk size / LDS word / lane mapping); see k_grid.hip
  int g2p_variant = 1;  // occupancy target of G2P's register allocation; see k_particles.hip
  double cur_dt = 1.0 / 60.0;
  int mask_bound = 1;
  float objs[SL_MAX_OBJ * 28];
  // particle count upper bound for launch sizes
  uint64_t emitted_total = 0, emitted_at_copy = 0;
  uint32_t p_ub = 0;
  // device
  SlTank ht;             // host mirror of the device record
  SlTank *d_tank = nullptr;
  std::vector<void *> allocs;
  SlCounters *h_counters = nullptr;  // pinned readback slot
  hipEvent_t stats_ev = nullptr;
  bool stats_pending = false;
  Staging stg;
  // host-side scene state: source progress, previous object poses, seed stream (scene.cpp)
  sl_scene scene;
  // saved water (50-sim.js:111-141): particles, whitewater, counters, ripple state and the sources'
  // progress
  struct Saved {
    bool valid = false;
    uint32_t id = 0;
    void *parts = nullptr, *ww = nullptr, *counters = nullptr, *ripple = nullptr;
    double time = 0;
    uint32_t count = 0, ww_count = 0, impacts = 0;
    bool full = false;
    int cur = 0, ww_cur = 0;
    std::map<uint32_t, SlSourceState> sources;
    uint32_t seed_state = 0;
  } saved;
  // timing
  bool timing = false;
  struct TEv { int stage; hipEvent_t a, b; };
  std::vector<TEv> tev;
  std::vector<hipEvent_t> ev_pool;
  double t_ms[SL_T_COUNT_] = {0};
  uint32_t t_launch[SL_T_COUNT_] = {0};
  // pressure solver (SL_SOLVER_SOR: the reference's sweeps; SL_SOLVER_MGPCG: converged solve)
  int solver = SL_SOLVER_SOR;
  SlMg mg;              // hierarchy and vectors, allocated when the converged solve is first chosen
  bool mg_seed = false; // the next solve starts from the SOR pressure
};

static uint32_t g_copies = 0;

#define FAIL(sim, code, ...)                                               \
  do {                                                                     \
    char _b[512];                                                          \
    snprintf(_b, sizeof _b, __VA_ARGS__);                                  \
    if (sim) (sim)->err = _b; else g_err = _b;                             \
    return code;                                                           \
  } while (0)

#define HIPCHK(sim, call)                                                                  \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) {                                                                \
      if (sim) (sim)->gpu_errors++;                                                        \
      FAIL(sim, SL_EHIP, "%s failed: %s", #call, hipGetErrorString(_e));                   \
    }                                                                                      \
  } while (0)

static int check_launch(sl_sim *s, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    s->gpu_errors++;
    FAIL(s, SL_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  }
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// timing
// ------------------------------------------------------------------------------------------------
static hipEvent_t ev_get(sl_sim *s) {
  if (!s->ev_pool.empty()) { hipEvent_t e = s->ev_pool.back(); s->ev_pool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
struct StageScope {
  sl_sim *s; int stage; hipEvent_t a = nullptr;
  StageScope(sl_sim *s_, int st, uint32_t launches) : s(s_), stage(st) {
    if (!s->timing) return;
    a = ev_get(s);
    (void)hipEventRecord(a, s->stream);
    s->t_launch[st] += launches;
  }
  ~StageScope() {
    if (!s->timing || !a) return;
    hipEvent_t b = ev_get(s);
    (void)hipEventRecord(b, s->stream);
    s->tev.push_back({stage, a, b});
  }
};
static void timing_collect(sl_sim *s) {
  if (s->tev.empty()) return;
  (void)hipStreamSynchronize(s->stream);
  for (auto &t : s->tev) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) s->t_ms[t.stage] += ms;
    s->ev_pool.push_back(t.a);
    s->ev_pool.push_back(t.b);
  }
  s->tev.clear();
}

// ------------------------------------------------------------------------------------------------
// staging
// ------------------------------------------------------------------------------------------------
static int stage_upload(sl_sim *s, void *dst, const void *src, size_t bytes) {
  Staging &g = s->stg;
  if (bytes > g.half) {  // large: synchronous
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SL_OK;
  }
  if (g.off + bytes > g.half) {
    HIPCHK(s, hipEventRecord(g.ev[g.cur], s->stream));
    g.ev_valid[g.cur] = true;
    g.cur ^= 1;
    g.off = 0;
    if (g.ev_valid[g.cur]) HIPCHK(s, hipEventSynchronize(g.ev[g.cur]));
  }
  uint8_t *p = g.buf + (size_t)g.cur * g.half + g.off;
  memcpy(p, src, bytes);
  g.off += (bytes + 255) & ~(size_t)255;
  HIPCHK(s, hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s->stream));
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// work maps of a single tank
// ------------------------------------------------------------------------------------------------
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
static SlLaunch launch_of(sl_sim *s, uint64_t items, uint32_t per_wg = SL_WG) {
  SlLaunch L;
  L.w.tanks = s->d_tank;
  L.w.start = nullptr;
  L.w.ntanks = 1;
  L.w.total = cdiv(items, per_wg);
  L.w.list = nullptr;
  L.stream = s->stream;
  // a simulation of its own: the host mirror equals the device record (push_dyn writes both, and
  // the host flips the buffer indices where it enqueues their flips); a tank of a batch is advanced
  // by the batch's kernels on the device, so its launches read the table
  if (!s->ext_tank) {
    s->ht.cur = s->cur;
    s->ht.ww_cur = s->ww_cur;
    L.one = &s->ht;
  }
  return L;
}

// multiplier and shifts for udiv_mag by d (1 <= d < 2^31): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, s1 = min(l, 1), s2 = max(l - 1, 0)
static void sl_magic_u32(uint32_t d, uint32_t g[3]) {
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  g[0] = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  g[1] = l < 1 ? l : 1;
  g[2] = l > 0 ? l - 1 : 0;
}

static int push_dyn(sl_sim *s, uint32_t emit_k = 0, uint32_t emit_total = 0) {
  SlTankDyn d;
  d.dt = s->ht.dt; d.gdt = s->ht.gdt; d.time_f = s->ht.time_f;
  d.rLx = s->ht.rLx; d.rLz = s->ht.rLz; d.rdt = s->ht.rdt; d.rmd = s->ht.rmd;
  // the object count of the parameter block (50-sim.js:155), as of the last writeParams: setObjects
  // alone does not change it (50-sim.js:206), so emitNow before a step sees the count of the last step
  d.frame = s->frame; d.obj_count = s->ht.obj_count; d.emit_k = emit_k; d.emit_total = emit_total;
  d.iters = s->iters; d.mask_bound = s->mask_bound;
  d.time_d = s->time;
  d.dt_frame = s->batch_dt;
  // the host mirror keeps the same values: single-simulation launches pass it by value
  s->ht.frame = d.frame; s->ht.emit_k = emit_k; s->ht.emit_total = emit_total;
  s->ht.iters = d.iters; s->ht.mask_bound = d.mask_bound; s->ht.time_d = d.time_d; s->ht.dt_frame = d.dt_frame;
  // A simulation of its own reads these fields only from its by-value record: the kernels that take
  // its device record (COUNT with an explicit emission total, the converged solve) read none of them,
  // so only a tank of a batch, whose kernels read the table, needs the device copy updated.
  if (!s->ext_tank) return SL_OK;
  hipLaunchKernelGGL(k_update, dim3(1), dim3(64), 0, s->stream, s->d_tank, d);
  return check_launch(s, "update");
}

// The particle buffer index flips on the device when a re-ordering's gather makes the other buffer
// current; the host flips its copy where it enqueues that flip. Code that copies particle buffers
// reads the device copy at a synchronisation point.
static int dev_cur(sl_sim *s) {
  int32_t c = 0;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(&c, (const uint8_t *)s->d_tank + offsetof(SlTank, cur), 4, hipMemcpyDeviceToHost));
  s->cur = c & 1;
  return SL_OK;
}

// writeParams, 50-sim.js:151-162 (the per-substep part; the static part is set at creation)
static void write_params(sl_sim *s, double dt) {
  s->cur_dt = dt;
  s->ht.dt = (float)dt;
  s->ht.gdt = s->ht.g * s->ht.dt;
  s->ht.time_f = (float)s->time;
  s->ht.frame = s->frame;
  s->ht.obj_count = s->obj_count;
}

// ------------------------------------------------------------------------------------------------
// config + creation
// ------------------------------------------------------------------------------------------------
extern "C" int sl_config_default(sl_config *c, const char *quality) {
  if (!c) FAIL((sl_sim *)nullptr, SL_EARG, "sl_config_default: null config");
  memset(c, 0, sizeof *c);
  c->tank_w = 0.6; c->tank_h = 0.32; c->tank_d = 0.4;  // scene.tank default, 10-scene.js:246
  // QUALITY, 50-sim.js:4-8
  c->h = 0.0085; c->cap = 320000; c->ww_cap = 140000; c->iters = 40; c->ripple_R = 256;
  if (quality && !strcmp(quality, "low")) { c->h = 0.012; c->cap = 140000; c->ww_cap = 60000; c->iters = 28; c->ripple_R = 128; }
  else if (quality && !strcmp(quality, "high")) { c->h = 0.0065; c->cap = 700000; c->ww_cap = 260000; c->iters = 56; c->ripple_R = 512; }
  else if (quality && strcmp(quality, "standard")) FAIL((sl_sim *)nullptr, SL_EARG, "unknown quality '%s'", quality);
  // writeParams constants, 50-sim.js:11, :157-160
  c->rc_factor = 0.32;
  c->gravity = 9.81f; c->flip = 0.96f; c->ppc = 8.0f; c->drift = 6.0f; c->omega = 1.6f;
  c->friction = 0.03f; c->ww_rate = 7.0f;
  // ripple constants, 64-ripple.js:225, :268, :88
  c->r_sigma_rho = (float)(0.0728 / 998.0); c->r_nu = 1.0e-6f; c->r_film = 0.3f;
  c->r_hp_k = (float)(2.0 * M_PI / 0.12); c->r_rho_kick = 998.0f;
  c->wmap = 256;
  c->whitewater = 1;
  c->device = 0;
  c->reorder_every = 0;
  c->seed = 1;
  // splashCrown, 31-wgsl-sim2.js:121-126
  c->splash_rho = 998.0; c->splash_sigma = 0.0728; c->splash_mu = 0.001;
  c->splash_k_jet = 700.0f; c->splash_k_crown = 2100.0f;
  c->ripple_depth = 0.0;
  // droplet layer off; when switched on: median 1.5 mm, sigma_ln 0.4, 1-4 m/s
  c->droplets = 0; c->drop_seed = 1;
  c->drop_d50 = 0.0015f; c->drop_sigma_ln = 0.4f; c->drop_vmin = 1.0f; c->drop_vmax = 4.0f;
  return SL_OK;
}

static int alloc(sl_sim *s, void **p, size_t bytes) {
  if (!bytes) bytes = 256;
  s->bytes += bytes;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) FAIL(s, SL_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  s->allocs.push_back(*p);
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) FAIL(s, SL_EHIP, "hipMemset failed: %s", hipGetErrorString(e));
  return SL_OK;
}
#define ALLOC(ptr, bytes)                                           \
  do {                                                              \
    void *_p = nullptr;                                             \
    int _r = alloc(s, &_p, (bytes));                                \
    if (_r) return _r;                                              \
    (ptr) = (decltype(ptr))_p;                                      \
  } while (0)

// grid sizing, 50-sim.js:33-45: at least 8 cells per axis, the cell size that fits all three, the
// origin at the min corner with x and z centred and the floor at y = 0
int sl_grid_of(const sl_config &c, int n[3], double &h, double origin[3]) {
  if (!(c.h > 0) || !(c.tank_w > 0) || !(c.tank_h > 0) || !(c.tank_d > 0)) return SL_EARG;
  n[0] = (int)fmax(8.0, round(c.tank_w / c.h));
  n[1] = (int)fmax(8.0, round(c.tank_h / c.h));
  n[2] = (int)fmax(8.0, round(c.tank_d / c.h));
  h = fmin(fmin(c.tank_w / n[0], c.tank_h / n[1]), c.tank_d / n[2]);
  origin[0] = -n[0] * h / 2; origin[1] = 0; origin[2] = -n[2] * h / 2;
  return SL_OK;
}

// configuration checks and grid sizing (50-sim.js:33-45); host only
static int size_grid(sl_sim *s) {
  const sl_config &c = s->cfg;
  if (!(c.h > 0) || !(c.tank_w > 0) || !(c.tank_h > 0) || !(c.tank_d > 0))
    FAIL(s, SL_EARG, "tank size and cell size must be positive");
  if (c.ripple_R && (c.ripple_R & (c.ripple_R - 1))) FAIL(s, SL_EARG, "ripple_R must be a power of two");
  if (c.ripple_R > 2048) FAIL(s, SL_ECAP, "ripple_R above 2048");
  if (c.cap == 0) FAIL(s, SL_EARG, "capacity must be positive");
  if (!(c.splash_rho > 0) || !(c.splash_sigma > 0) || !(c.splash_mu > 0))
    FAIL(s, SL_EARG, "splash density, surface tension and viscosity must be positive");
  if (!(c.ripple_depth >= 0)) FAIL(s, SL_EARG, "ripple_depth must be >= 0 (0: mean depth of the count)");
  if (c.droplets) {
    if (c.droplets > (1u << 22)) FAIL(s, SL_ECAP, "droplets above 2^22 per frame");
    if (!c.ripple_R) FAIL(s, SL_EARG, "droplets need a ripple layer (ripple_R > 0)");
    if (!(c.drop_d50 > 0) || !(c.drop_sigma_ln >= 0) || !(c.drop_vmin >= 0) || !(c.drop_vmax >= c.drop_vmin))
      FAIL(s, SL_EARG, "droplets: need drop_d50 > 0, drop_sigma_ln >= 0 and 0 <= drop_vmin <= drop_vmax");
  }
  if (c.cap >= (1u << 30) || c.ww_cap >= (1u << 28))  // 32-bit byte offsets into the particle arrays
    FAIL(s, SL_ECAP, "capacity too large (2^30 particles, 2^28 whitewater at most)");
  {
    int n[3];
    sl_grid_of(c, n, s->h, s->origin);
    s->nx = n[0]; s->ny = n[1]; s->nz = n[2];
  }
  s->rim = s->ny * s->h;
  uint64_t nodes = (uint64_t)(s->nx + 1) * (s->ny + 1) * (s->nz + 1);
  if (nodes >= (1ull << 30)) FAIL(s, SL_ECAP, "grid too large (2^30 nodes at most)");  // 32-bit byte offsets
  s->nodes = (uint32_t)nodes;
  s->cells = (uint32_t)s->nx * s->ny * s->nz;
  s->hw = (uint32_t)(s->nx + 1) / 2;
  s->hcells = s->hw * (uint32_t)s->ny * (uint32_t)s->nz;
  s->cap = c.cap; s->ww_cap = c.ww_cap; s->iters = c.iters;
  s->R = c.ripple_R; s->M = 2 * c.ripple_R;
  s->vp = pow(s->h / 2, 3);
  s->radius = 0.5 * s->h * 1.05;
  return SL_OK;
}

// Every device buffer of a tank, in allocation order (50-sim.js:59-80, 64-ripple.js:152-165,
// structure-of-arrays). `take(&ptr, bytes)` allocates, or only counts for sl_job_bytes; the sizes
// come from the grid of size_grid.
#define TAKE(ptr, bytes)                                            \
  do {                                                              \
    void *_p = nullptr;                                             \
    int _r = take(&_p, (size_t)(bytes));                            \
    if (_r) return _r;                                              \
    (ptr) = (decltype(ptr))_p;                                      \
  } while (0)
template <class Take>
static int tank_buffers(sl_sim *s, SlTank &T, Take take) {
  const size_t N = s->nodes, C = s->cells, H = s->hcells;
  for (int b = 0; b < 2; b++) {
    TAKE(T.px[b], (size_t)s->cap * 4); TAKE(T.py[b], (size_t)s->cap * 4); TAKE(T.pz[b], (size_t)s->cap * 4);
    TAKE(T.vx[b], (size_t)s->cap * 4); TAKE(T.vy[b], (size_t)s->cap * 4); TAKE(T.vz[b], (size_t)s->cap * 4);
    TAKE(T.pf[b], (size_t)s->cap);
    TAKE(T.wp[b], (size_t)(s->ww_cap ? s->ww_cap : 1) * 16);
    TAKE(T.wv[b], (size_t)(s->ww_cap ? s->ww_cap : 1) * 16);
  }
  if (s->ext_counters) T.C = s->ext_counters;
  else TAKE(T.C, sizeof(SlCounters));
  TAKE(T.args, 64);
  TAKE(T.vslot, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE * 4);
  TAKE(T.acc, N * 24);
  TAKE(T.cellCount, C * 4);
  TAKE(T.vel, N * 12); TAKE(T.velOld, N * 12);
  TAKE(T.frac, N * 12); TAKE(T.svel, N * 12);
  TAKE(T.divg, C * 4);
  {
    // both colour arrays with hw*ny + 1 guard slots before and after: every SOR neighbour slot of
    // an in-range slot lies in the allocation, so the kernel needs no clamps
    const size_t G = (size_t)s->hw * (size_t)s->ny + 1;
    TAKE(T.pres, (H * 2 + 2 * G) * 4);
    if (T.pres) T.pres += G;
  }
  TAKE(T.rhs, H * 8); TAKE(T.sflag, H * 4); TAKE(T.scut, H * 64);
  TAKE(T.objs, sizeof(SlObj) * SL_MAX_OBJ);
  TAKE(T.objMask, (size_t)cdiv(C, 64) * 8 + 64);
  if (s->ext_emits) T.emits = s->ext_emits;
  else TAKE(T.emits, sizeof(SlEmitDev) * SL_MAX_EMIT);
  TAKE(T.scanBlocks, (size_t)cdiv(s->cap, SL_COMPACT_ITEMS) * 4 + 64);
  TAKE(T.binCell, (size_t)s->cap * 4); TAKE(T.binRank, (size_t)s->cap * 4); TAKE(T.perm, (size_t)s->cap * 4);
  TAKE(T.cellStart, C * 4 + 64);
  TAKE(T.binTop, (size_t)sl_bin_blocks(s->cells) * 4 + 64);
  if (s->R) {
    const size_t M2 = (size_t)s->M * s->M, R1 = (size_t)(s->R + 1) * (s->R + 1);
    TAKE(T.rsrc, R1 * 4);
    TAKE(T.rstate, M2 * 8); TAKE(T.ra, M2 * 8); TAKE(T.rb, M2 * 8); TAKE(T.rhat, M2 * 8);
    TAKE(T.tw, (size_t)s->M * 8);
    TAKE(T.rmask, R1 * 4);
    TAKE(T.rtex, R1 * 16);
    const size_t wm = s->cfg.wmap ? s->cfg.wmap : 256;
    TAKE(T.topH, wm * wm * 4);
    if (s->cfg.top_height) TAKE(T.topB, wm * wm * 4);
  }
  // re-ordering keys: cell indices 0 .. C-1 and C for particles outside the grid (the tanks of a
  // batch sort in the batch's buffers)
  {
    int bits = 0;
    while ((1ull << bits) <= (uint64_t)s->cells) bits++;
    s->sort_end_bit = bits;
  }
  if (!s->ext_tank) {
    TAKE(s->sort_keys, (size_t)s->cap * 8);
    TAKE(s->sort_vals, (size_t)s->cap * 8);
    s->sort_temp_bytes = sl_sort_temp_bytes(s->cap, s->sort_end_bit);
    TAKE(s->sort_temp, s->sort_temp_bytes);
  }
  if (s->ext_tank) s->d_tank = s->ext_tank;
  else TAKE(s->d_tank, sizeof(SlTank));
  return SL_OK;
}
#undef TAKE

static int create_impl(sl_sim *s) {
  const sl_config &c = s->cfg;
  {
    int r = size_grid(s);
    if (r) return r;
  }
  HIPCHK(s, hipSetDevice(c.device));
  s->device = c.device;
  if (s->ext_stream) s->stream = s->ext_stream;
  else HIPCHK(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  s->time = 0; s->vmax = 1.0; s->frame = 0;
  s->reorder_clock = c.reorder_every ? c.reorder_phase % c.reorder_every : 0;
  if (sl_scene_init(s->scene, c)) FAIL(s, SL_EARG, "scene state of the configuration");
  memset(s->objs, 0, sizeof s->objs);

  SlTank &T = s->ht;
  memset(&T, 0, sizeof T);
  T.nx = s->nx; T.ny = s->ny; T.nz = s->nz; T.hw = (int32_t)s->hw;
  sl_magic_u32(s->hw, T.hw_mag);
  sl_magic_u32(s->hw * (uint32_t)s->ny, T.rowl_mag);
  sl_magic_u32((uint32_t)s->nx, T.nx_mag);
  sl_magic_u32((uint32_t)(s->nx * s->ny), T.nxy_mag);
  sl_magic_u32((uint32_t)(s->nx + 1), T.nx1_mag);
  sl_magic_u32((uint32_t)((s->nx + 1) * (s->ny + 1)), T.nxy1_mag);
  T.nodes = s->nodes; T.cells = s->cells; T.hcells = s->hcells;
  T.cap = s->cap; T.ww_cap = s->ww_cap; T.R = s->R; T.M = s->M; T.iters = s->iters;
  uint32_t lm = 0;
  while (s->M && (1u << lm) < s->M) lm++;
  T.logM = lm;
  T.wmap = c.wmap ? c.wmap : 256;
  T.mask_bound = s->R ? 1 : 0;
  s->mask_bound = T.mask_bound;
  T.ox = (float)s->origin[0]; T.oy = (float)s->origin[1]; T.oz = (float)s->origin[2];
  T.h = (float)s->h; T.invh = (float)(1.0 / s->h); T.hh = T.h * T.h;
  T.g = c.gravity; T.flip = c.flip; T.ppc = c.ppc; T.drift = c.drift; T.omega = c.omega;
  T.omega1 = 1.0f - T.omega;
  T.rc = (float)(c.rc_factor * s->h);
  T.rim = (float)s->rim;
  T.ww_rate = c.whitewater ? c.ww_rate : 0.0f;
  T.whitewater_on = c.whitewater ? 1 : 0;
  T.ww_cap_f = (float)s->ww_cap;
  T.friction = c.friction;
  T.radius = (float)s->radius;
  T.R_f = (float)s->R;
  T.h_d = s->h; T.vp_d = s->vp; T.rim_d = s->rim; T.time_d = 0;
  T.top_height = c.top_height && s->R ? 1 : 0;
  T.tankH = (float)c.tank_h;
  T.rsr = c.r_sigma_rho; T.rnu = c.r_nu; T.rfilm = c.r_film; T.rhpk = c.r_hp_k; T.rrho = c.r_rho_kick;
  T.rdepth = (float)c.ripple_depth;
  // splashCrown's constant expressions, folded in double and rounded once as the WGSL folds its
  // abstract-float constants (31-wgsl-sim2.js:121-123: 3.14159265 * 998.0, 998.0 * 0.0728)
  T.sp_pirho = (float)(3.14159265 * c.splash_rho);
  T.sp_rs = (float)(c.splash_rho * c.splash_sigma);
  T.sp_rho = (float)c.splash_rho; T.sp_sigma = (float)c.splash_sigma; T.sp_mu = (float)c.splash_mu;
  T.sp_kjet = c.splash_k_jet; T.sp_kcrown = c.splash_k_crown;
  T.drops = c.droplets; T.drop_seed = c.drop_seed;
  T.drop_d50 = c.drop_d50; T.drop_sig = c.drop_sigma_ln; T.drop_vmin = c.drop_vmin; T.drop_vmax = c.drop_vmax;
  T.drop_mc = (float)(M_PI / 6.0 * c.splash_rho);
  {
    int r = tank_buffers(s, T, [s](void **p, size_t bytes) { return alloc(s, p, bytes); });
    if (r) return r;
  }
  if (s->R) {
    // twiddles e^(2 pi i t/M) in double, rounded to f32 (64-ripple.js:175-178)
    std::vector<float> tw(2 * (size_t)s->M);
    for (uint32_t t = 0; t < s->M; t++) {
      tw[2 * t] = (float)cos((2 * M_PI * t) / s->M);
      tw[2 * t + 1] = (float)sin((2 * M_PI * t) / s->M);
    }
    HIPCHK(s, hipMemcpy(T.tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice));
  }
  // pinned host memory
  HIPCHK(s, hipHostMalloc((void **)&s->h_counters, sizeof(SlCounters), hipHostMallocDefault));
  memset(s->h_counters, 0, sizeof(SlCounters));
  HIPCHK(s, hipEventCreateWithFlags(&s->stats_ev, hipEventDisableTiming));
  s->stg.half = 1 << 20;
  HIPCHK(s, hipHostMalloc((void **)&s->stg.buf, 2 * s->stg.half, hipHostMallocDefault));
  HIPCHK(s, hipEventCreateWithFlags(&s->stg.ev[0], hipEventDisableTiming));
  HIPCHK(s, hipEventCreateWithFlags(&s->stg.ev[1], hipEventDisableTiming));
  // writeParams(1/60) + COUNT(4), 50-sim.js:85-88
  write_params(s, 1.0 / 60.0);
  HIPCHK(s, hipMemcpy(s->d_tank, &s->ht, sizeof(SlTank), hipMemcpyHostToDevice));
  HIPCHK(s, hipDeviceSynchronize());
  return SL_OK;
}

// device bytes a simulation of this configuration holds (the sum of its allocation sizes); host only
extern "C" size_t sl_job_bytes(const sl_config *cfg) {
  if (!cfg) return 0;
  sl_sim tmp;
  tmp.cfg = *cfg;
  if (size_grid(&tmp)) return 0;
  SlTank T;
  memset(&T, 0, sizeof T);
  size_t sum = 0;
  tank_buffers(&tmp, T, [&sum](void **p, size_t bytes) {
    *p = nullptr;
    sum += bytes ? bytes : 256;
    return SL_OK;
  });
  return sum;
}

extern "C" int sl_create(const sl_config *cfg, sl_sim **out) {
  if (!cfg || !out) FAIL((sl_sim *)nullptr, SL_EARG, "sl_create: null argument");
  *out = nullptr;
  sl_sim *s = new (std::nothrow) sl_sim();
  if (!s) FAIL((sl_sim *)nullptr, SL_ENOMEM, "out of host memory");
  s->cfg = *cfg;
  int r = create_impl(s);
  if (r) {
    g_err = s->err;
    sl_destroy(s);
    return r;
  }
  *out = s;
  return SL_OK;
}

extern "C" void sl_destroy(sl_sim *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void *p : s->allocs) (void)hipFree(p);
  if (s->saved.parts) { (void)hipFree(s->saved.parts); (void)hipFree(s->saved.ww); (void)hipFree(s->saved.counters); }
  if (s->saved.ripple) (void)hipFree(s->saved.ripple);
  if (s->h_counters) (void)hipHostFree(s->h_counters);
  if (s->stg.buf) (void)hipHostFree(s->stg.buf);
  for (int i = 0; i < 2; i++) if (s->stg.ev[i]) (void)hipEventDestroy(s->stg.ev[i]);
  if (s->stats_ev) (void)hipEventDestroy(s->stats_ev);
  for (auto &t : s->tev) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->stream && !s->ext_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

extern "C" const char *sl_last_error(const sl_sim *s) { return s ? s->err.c_str() : g_err.c_str(); }

extern "C" int sl_grid_info(const sl_sim *s, sl_grid_t *o) {
  if (!s || !o) FAIL((sl_sim *)nullptr, SL_EARG, "sl_grid_info: null argument");
  o->nx = s->nx; o->ny = s->ny; o->nz = s->nz; o->nodes = s->nodes; o->cells = s->cells;
  o->cap = s->cap; o->ww_cap = s->ww_cap; o->R = s->R; o->M = s->M; o->iters = s->iters;
  o->h = s->h; o->origin[0] = s->origin[0]; o->origin[1] = s->origin[1]; o->origin[2] = s->origin[2];
  o->rim = s->rim; o->vp = s->vp; o->radius = s->radius;
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// objects, 50-sim.js:194-210
// ------------------------------------------------------------------------------------------------
static bool same_f(float a, float b) { return a == b || (a != a && b != b); }
static uint32_t f32_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float bits_f32(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
// Radius of a sphere about an object's position that contains its solid, from the shape table of
// shape_sdf (sl_internal.h), stored in the unused row pad r0.w of the device copy. For the shapes
// below the signed distance is at least the distance to that sphere (the exact distance of a solid
// inside it; for the cup and container the max() of the outer solid's exact distance and more), so
// G2P may skip an obstacle whose sphere is farther than the contact radius, and a drain whose
// cylinder sphere does not hold the point. +inf (never skipped) where that is not guaranteed: the
// gutter (an intersection of three distances), negative or degenerate dimensions, and rotation rows
// that are not orthonormal; -inf for types whose distance is the constant 1e9.
static float cull_radius(const float *row) {
  const float *a = row, *b = row + 4, *c = row + 8, *d = row + 16;
  auto dot = [](const float *x, const float *y) { return (double)x[0] * y[0] + (double)x[1] * y[1] + (double)x[2] * y[2]; };
  const double err = fabs(dot(a, a) - 1) + fabs(dot(b, b) - 1) + fabs(dot(c, c) - 1) + fabs(dot(a, b)) +
                     fabs(dot(a, c)) + fabs(dot(b, c));
  if (!(err < 1e-4)) return INFINITY;
  const float flag = row[27];
  const double d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
  if (flag == 2.0f) return d0 >= 0 && d1 >= 0 ? (float)hypot(d0, d1) : INFINITY;  // in_drain: sd_cyl < 0
  if (flag != 1.0f) return INFINITY;
  const double t = row[15] + 0.5;
  if (!(t >= 0 && t < 4294967296.0)) return -INFINITY;
  switch ((uint32_t)t) {
    case 0:
      return d0 >= 0 && d1 >= 0 && d2 >= 0 && d3 >= 0 && d3 <= std::min(std::min(d0, d1), d2)
                 ? (float)sqrt(d0 * d0 + d1 * d1 + d2 * d2) : INFINITY;
    case 1: return (float)d0;  // |q| - r exactly
    case 2: case 3: case 5: return d0 >= 0 && d1 >= 0 ? (float)hypot(d0, d1) : INFINITY;
    case 4: return d1 >= 0 ? (float)(fabs(d0) + d1) : INFINITY;
    case 6:
      return d0 >= 0.002 && d1 >= 0.002 && d2 >= 0.002 ? (float)sqrt(d0 * d0 + d1 * d1 + d2 * d2) : INFINITY;
    case 7: return INFINITY;
    default: return -INFINITY;
  }
}

extern "C" int sl_set_objects(sl_sim *s, const float *rows, uint32_t n, int *changed) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (n && !rows) FAIL(s, SL_EARG, "sl_set_objects: null rows");
  if (n > SL_MAX_OBJ) n = SL_MAX_OBJ;
  bool ch = n != s->obj_count;
  for (uint32_t i = 0; !ch && i < n * 28; i++) ch = !same_f(s->objs[i], rows[i]);
  if (changed) *changed = ch ? 1 : 0;
  if (!ch) return SL_OK;
  memset(s->objs, 0, sizeof s->objs);
  if (n) memcpy(s->objs, rows, (size_t)n * 28 * 4);
  s->obj_count = n;
  s->geom_dirty = true;
  s->mask_dirty = true;
  float dev[SL_MAX_OBJ * 28];
  memcpy(dev, s->objs, sizeof dev);
  for (uint32_t i = 0; i < n; i++) dev[i * 28 + 3] = cull_radius(dev + i * 28);
  return stage_upload(s, s->ht.objs, dev, sizeof(SlObj) * (n ? n : 1));
}

extern "C" sl_scene *sl_sim_scene(sl_sim *s) { return s ? &s->scene : nullptr; }

extern "C" int sl_set_items(sl_sim *s, const sl_item *items, uint32_t n, double dt_move, int *moving) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (n && !items) FAIL(s, SL_EARG, "sl_set_items: null items");
  float rows[SL_MAX_OBJ * 28];
  uint32_t k = 0;
  int r = sl_scene_pack_objects(&s->scene, items, n, dt_move, rows, &k, moving);
  if (r) FAIL(s, r, "sl_set_items: packing failed");
  return sl_set_objects(s, rows, k, nullptr);
}

// ------------------------------------------------------------------------------------------------
// emission: writeEmits (50-sim.js:222-245) + EMIT + COUNT(2) (50-sim.js:267-274)
// ------------------------------------------------------------------------------------------------
static int write_emits(sl_sim *s, const sl_emit *list, uint32_t n, uint32_t *total_out, uint32_t *k_out) {
  SlEmitDev em[SL_MAX_EMIT];
  uint32_t off = 0, k = 0;
  int64_t room = (int64_t)s->cap - (int64_t)s->count - 64;
  int64_t budget = room > 0 ? room : 0;
  bool any = false;
  for (uint32_t i = 0; i < n; i++) if (list[i].count > 0) any = true;
  for (uint32_t i = 0; i < n; i++) {
    if (k >= SL_MAX_EMIT || budget <= 0) break;
    int64_t cnt = (int64_t)list[i].count < budget ? (int64_t)list[i].count : budget;
    if (cnt <= 0) continue;
    memcpy(em[k].a, list[i].a, 16); memcpy(em[k].b, list[i].b, 16);
    memcpy(em[k].c, list[i].c, 16); memcpy(em[k].d, list[i].d, 16);
    em[k].info[0] = off; em[k].info[1] = (uint32_t)cnt; em[k].info[2] = list[i].seed; em[k].info[3] = list[i].shape;
    off += (uint32_t)cnt; budget -= cnt; k++;
  }
  if (any && budget <= 0) s->full = true;
  *total_out = off; *k_out = k;
  if (off > 0) {
    int r = stage_upload(s, s->ht.emits, em, sizeof(SlEmitDev) * k);
    if (r) return r;
    s->count += off;
  }
  return SL_OK;
}
// commit = false (inside a substep) leaves count += total to the substep's cell-count clear, the
// next kernel, which reads nothing that needs it (the launch is saved)
static int encode_emit(sl_sim *s, uint32_t k, uint32_t total, bool commit = true) {
  int r = push_dyn(s, k, total);
  if (r) return r;
  {
    StageScope sc(s, SL_T_EMIT, commit ? 2 : 1);
    sl_k_emit(launch_of(s, total));
    if (commit) sl_k_count(launch_of(s, 1), 2, total);
    else s->pending_add += total;
  }
  s->emitted_total += total;
  uint64_t ub = (uint64_t)s->p_ub + total;
  s->p_ub = (uint32_t)(ub < s->cap ? ub : s->cap);
  return check_launch(s, "emit");
}

extern "C" int sl_emit_now(sl_sim *s, const sl_emit *list, uint32_t n, int fill, uint32_t *emitted) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (n && !list) FAIL(s, SL_EARG, "sl_emit_now: null list");
  uint32_t sum = 0;
  for (uint32_t i = 0; i < n; i += SL_MAX_EMIT) {
    uint32_t m = n - i < SL_MAX_EMIT ? n - i : SL_MAX_EMIT;
    uint32_t total = 0, k = 0;
    int r = write_emits(s, list + i, m, &total, &k);
    if (r) return r;
    if (!total) continue;
    if (fill) {
      // fresh per-cell counts so that the pour only fills cells without water (50-sim.js:253-261)
      sl_k_fill_u32(s->stream, s->ht.cellCount, 0, s->cells);
      sl_k_p2g(launch_of(s, s->p_ub, 512), 1, 0);
    }
    r = encode_emit(s, k, total);
    if (r) return r;
    sum += total;
  }
  if (emitted) *emitted = sum;
  return check_launch(s, "emit_now");
}

// ------------------------------------------------------------------------------------------------
// step, 50-sim.js:300-336
// ------------------------------------------------------------------------------------------------
extern "C" int sl_set_substeps(sl_sim *s, uint32_t n) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  s->forced_substeps = n;
  return SL_OK;
}

// levels of the converged solve: the cell grid, then every dimension halved (rounding up) until the
// level has at most SL_MG_COARSE cells
static int mg_alloc(sl_sim *s) {
  SlMg &m = s->mg;
  if (m.nlev) return SL_OK;
  int32_t X = s->nx, Y = s->ny, Z = s->nz;
  int l = 0;
  for (;;) {
    if (l == SL_MG_MAXLEV) FAIL(s, SL_EARG, "pressure solver: grid %dx%dx%d needs more than %d levels", s->nx, s->ny, s->nz, SL_MG_MAXLEV);
    SlMgLevel &L = m.lev[l];
    L.X = X; L.Y = Y; L.Z = Z;
    L.cells = (uint32_t)X * (uint32_t)Y * (uint32_t)Z;
    sl_magic_u32((uint32_t)X, L.magX);
    sl_magic_u32((uint32_t)X * (uint32_t)Y, L.magXY);
    ALLOC(L.A, (size_t)L.cells * sizeof(float4));
    ALLOC(L.F, (size_t)L.cells * 2);
    ALLOC(L.b, (size_t)L.cells * 4);
    ALLOC(L.x, (size_t)L.cells * 4);
    l++;
    if (L.cells <= SL_MG_COARSE) break;
    X = (X + 1) / 2; Y = (Y + 1) / 2; Z = (Z + 1) / 2;
  }
  const size_t C = s->cells;
  ALLOC(m.xs, C * 8); ALLOC(m.bs, C * 8); ALLOC(m.r, C * 8); ALLOC(m.p[0], C * 8); ALLOC(m.p[1], C * 8);
  ALLOC(m.q, C * 8);
  ALLOC(m.dlo, C * 4);
  ALLOC(m.part, (size_t)sl_mg_parts(s->cells) * 8);
  ALLOC(m.scal, 8 * 8);
  ALLOC(m.ctl, 8 * 4);
  m.nlev = l;
  return SL_OK;
}

extern "C" int sl_set_pressure_solver(sl_sim *s, int solver, double tol_per_s, uint32_t max_iters) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (solver == SL_SOLVER_SOR) { s->solver = solver; return SL_OK; }
  if (solver != SL_SOLVER_MGPCG) FAIL(s, SL_EARG, "sl_set_pressure_solver: unknown solver %d", solver);
  if (s->ext_stream) FAIL(s, SL_EARG, "sl_set_pressure_solver: the tanks of a batch use the SOR sweeps");
  if (!(tol_per_s >= 0) || max_iters < 1 || max_iters > 1000)
    FAIL(s, SL_EARG, "sl_set_pressure_solver: tolerance %g must be >= 0 and max_iters %u in 1..1000", tol_per_s, max_iters);
  int r = mg_alloc(s);
  if (r) return r;
  if (s->solver != SL_SOLVER_MGPCG) s->mg_seed = true;
  s->solver = solver;
  s->mg.rtol = tol_per_s * s->h * s->h;
  s->mg.max_iters = max_iters;
  return SL_OK;
}

extern "C" int sl_pressure_solver_stats(sl_sim *s, uint32_t *iters, int *breakdown, double *resid_per_s,
                                        double *resid0_per_s) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (!s->mg.nlev) FAIL(s, SL_EARG, "sl_pressure_solver_stats: the converged solve has not been chosen");
  uint32_t ctl[3];
  double scal[5];
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(ctl, s->mg.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  HIPCHK(s, hipMemcpy(scal, s->mg.scal, sizeof scal, hipMemcpyDeviceToHost));
  const double h2 = s->h * s->h;
  if (iters) *iters = ctl[1];
  if (breakdown) *breakdown = (int)ctl[2];
  if (resid_per_s) *resid_per_s = scal[3] / h2;
  if (resid0_per_s) *resid0_per_s = scal[4] / h2;
  return SL_OK;
}

// the pressure solve and PROJECT of a substep, by the chosen solver
static void encode_pressure(sl_sim *s) {
  if (s->solver == SL_SOLVER_MGPCG) {
    {
      StageScope sc(s, SL_T_SOR, 1);
      sl_mg_solve(s->stream, s->d_tank, s->mg, s->mg_seed ? 1 : 0);
      s->mg_seed = false;
    }
    StageScope sc(s, SL_T_PROJECT, 1);
    sl_k_project_mg(s->stream, s->d_tank, s->nodes, s->mg.xs);
    return;
  }
  {
    StageScope sc(s, SL_T_SOR, 2 * s->iters);
    const SlSorRec R = sl_sor_rec(s->ht);
    for (uint32_t i = 0; i < s->iters; i++) { sl_k_sor_rec(s->stream, R, 0); sl_k_sor_rec(s->stream, R, 1); }
  }
  StageScope sc(s, SL_T_PROJECT, 1);
  sl_k_project(launch_of(s, s->nodes));
}

extern "C" int sl_step_begin(sl_sim *s, double frame_dt, uint32_t *nsub, double *dt) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (!(frame_dt > 0)) { if (nsub) *nsub = 0; if (dt) *dt = 0; return SL_OK; }
  uint32_t n;
  if (s->forced_substeps) n = s->forced_substeps;
  else {
    double vmax = s->vmax > 0.4 ? s->vmax : 0.4;
    double q = ceil(frame_dt / (1.4 * s->h / vmax));
    n = (uint32_t)(q < 1 ? 1 : (q > 6 ? 6 : q));
  }
  s->substeps_last = n;
  if (nsub) *nsub = n;
  if (dt) *dt = frame_dt / n;
  sl_k_count(launch_of(s, 1), 5, 0);  // counters.vmax = 0
  return check_launch(s, "step_begin");
}

// P2G of one tank (cell counts cleared by the caller): a chunked LDS variant or the binned one
static void encode_p2g(sl_sim *s, int variant) {
  const uint32_t pb = s->p_ub;
  if (variant == SL_P2G_BINNED)
    sl_k_p2g_binned(launch_of(s, pb, SL_BIN_PER_WG), launch_of(s, sl_bin_blocks(s->cells), 1), launch_of(s, 1, 1),
                    launch_of(s, sl_bin_tiles(s->nx, s->ny, s->nz), 1));
  else
    sl_k_p2g(launch_of(s, pb, sl_p2g_chunk(variant)), 0, variant);
}

// G2P's object mask, after objects changed (the objects and the count the substep's parameters hold)
static void encode_objmask(sl_sim *s) {
  if (!s->mask_dirty) return;
  sl_k_objmask(launch_of(s, s->cells));
  s->mask_dirty = false;
}

static int encode_substep(sl_sim *s) {
  const uint32_t pb = s->p_ub;
  if (s->geom_dirty) {
    StageScope sc(s, SL_T_GEOM, 1);
    sl_k_geom(launch_of(s, s->nodes));
    s->geom_dirty = false;
  }
  {
    StageScope sc(s, SL_T_P2G, 2);
    sl_k_clear_cells_commit(s->stream, s->ht.cellCount, s->cells, s->ht.C, s->pending_add, s->cap);
    s->pending_add = 0;
    encode_p2g(s, s->p2g_variant);
  }
  { StageScope sc(s, SL_T_NORMFORCE, 1); sl_k_normforce(launch_of(s, s->nodes)); }
  { StageScope sc(s, SL_T_DIVERGENCE, 1); sl_k_divsetup(launch_of(s, s->cells)); }
  encode_pressure(s);
  encode_objmask(s);
  { StageScope sc(s, SL_T_G2P, 1); sl_k_g2p(launch_of(s, pb), s->g2p_variant); }
  return check_launch(s, "substep");
}

extern "C" int sl_substep(sl_sim *s, double dt, const sl_emit *list, uint32_t n, uint32_t *emitted) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (n && !list) FAIL(s, SL_EARG, "sl_substep: null list");
  s->frame++;
  write_params(s, dt);
  uint32_t total = 0, k = 0;
  int r = write_emits(s, list, n, &total, &k);
  if (r) return r;
  if (total > 0) { r = encode_emit(s, k, total, false); if (r) return r; }
  r = push_dyn(s);
  if (r) return r;
  r = encode_substep(s);
  if (r) return r;
  s->time += dt;
  if (emitted) *emitted = total;
  return SL_OK;
}

// the ripple depth of a frame: the mean depth of the host's count clamped to [0.003, rim]
// (50-sim.js:332), or the configured fixed depth
static double frame_mean_depth(const sl_sim *s) {
  if (s->cfg.ripple_depth > 0) return s->cfg.ripple_depth;
  double liters = s->count * s->vp * 1000;
  double md = liters / 1000 / (s->nx * s->nz * s->h * s->h);
  return md < 0.003 ? 0.003 : (md > s->rim ? s->rim : md);
}
// the frame's ripple uniform (64-ripple.js:223-226), read by the droplets and the ripple frame
static int set_ripple_uniform(sl_sim *s, double frame_dt, double mean_depth) {
  if (!s->R) return SL_OK;
  s->ht.rLx = (float)(s->nx * s->h); s->ht.rLz = (float)(s->nz * s->h);
  s->ht.rdt = (float)frame_dt; s->ht.rmd = (float)mean_depth;
  return push_dyn(s);
}
static int encode_droplets(sl_sim *s) {
  if (!s->ht.drops) return SL_OK;
  StageScope sc(s, SL_T_DROPLETS, 1);
  sl_k_droplets(launch_of(s, s->ht.drops, 64));
  return check_launch(s, "droplets");
}
static int encode_ripple(sl_sim *s) {
  if (!s->R) return SL_OK;
  StageScope sc(s, SL_T_RIPPLE, s->mask_bound ? 10 : 6);
  const uint32_t M = s->M, R1 = s->R + 1;
  uint32_t CW = 4096 / M; if (CW > M) CW = M; if (CW < 1) CW = 1;
  const SlLaunch rows = launch_of(s, M / sl_fft_rows_per_wg(M), 1);
  sl_k_ripple_frame_ex(rows.w, launch_of(s, M, CW).w, launch_of(s, (uint64_t)M * M).w,
                       launch_of(s, (uint64_t)R1 * R1).w, s->stream, s->mask_bound, M, rows.one);
  return check_launch(s, "ripple");
}

// commit = false leaves the count commit (count = newCount) to the whitewater commit of the same
// frame (COUNT mode 8): the kernels between them read only the whitewater counters
static int encode_compact(sl_sim *s, bool commit = true) {
  StageScope sc(s, SL_T_COMPACT, commit ? 4 : 3);
  SlLaunch L = launch_of(s, s->p_ub, SL_COMPACT_ITEMS);
  sl_k_compact_count(L);
  sl_k_compact_scan(L, 1);
  sl_k_compact_move(L);
  if (commit) sl_k_count(launch_of(s, 1), 1, 0);
  return check_launch(s, "compact");
}

static int encode_stats_copy(sl_sim *s) {
  if (s->stats_pending) return SL_OK;  // one readback in flight, like the staging ring being busy
  HIPCHK(s, hipMemcpyAsync(s->h_counters, s->ht.C, sizeof(SlCounters), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipEventRecord(s->stats_ev, s->stream));
  s->stats_pending = true;
  s->emitted_at_copy = s->emitted_total;
  return SL_OK;
}

static int encode_reorder(sl_sim *s) {
  if (!s->sort_keys) FAIL(s, SL_EARG, "re-ordering: the tanks of a batch re-order with sl_batch_reorder");
  StageScope sc(s, SL_T_COMPACT, 5);
  int e = sl_k_reorder(launch_of(s, 1), s->p_ub, s->sort_keys, s->sort_vals, s->sort_temp, s->sort_temp_bytes,
                       s->sort_end_bit);
  if (e) { s->gpu_errors++; FAIL(s, SL_EHIP, "particle re-ordering failed: %s", hipGetErrorString((hipError_t)e)); }
  sl_k_count(launch_of(s, 1), 6, 0);
  s->cur ^= 1;  // the gathered buffer becomes current
  return check_launch(s, "reorder");
}

extern "C" int sl_reorder(sl_sim *s) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  return encode_reorder(s);
}

extern "C" int sl_step_end(sl_sim *s, double frame_dt) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (!(frame_dt > 0)) return SL_OK;
  // a re-ordering reads the committed count; without one the commit joins the whitewater commit
  const bool reorder_due = s->cfg.reorder_every && s->reorder_clock + 1 >= s->cfg.reorder_every;
  int r = encode_compact(s, reorder_due);
  if (r) return r;
  if (s->cfg.reorder_every && ++s->reorder_clock >= s->cfg.reorder_every) {
    // the cadence restarts here only: an explicit sl_reorder leaves it (as the oracle's step does)
    s->reorder_clock = 0;
    r = encode_reorder(s);
    if (r) return r;
  }
  // the ripple uniform first: the droplets splash at the frame's water level
  if ((r = set_ripple_uniform(s, frame_dt, frame_mean_depth(s)))) return r;
  if ((r = encode_droplets(s))) return r;
  {
    StageScope sc(s, SL_T_WW, 2);
    sl_k_ww(launch_of(s, s->ww_cap), (float)frame_dt);
    sl_k_count(launch_of(s, 1), reorder_due ? 3 : 8, 0);
    s->ww_cur ^= 1;
  }
  r = encode_ripple(s);
  if (r) return r;
  if (s->ht.top_height) {  // drawn after the step, read by the next frame's ripple mask
    StageScope sc(s, SL_T_RIPPLE, 6);
    const uint32_t wm = s->ht.wmap;
    const SlLaunch parts = launch_of(s, s->p_ub);
    sl_k_toph(parts.w, launch_of(s, (uint64_t)wm * wm).w, s->stream, parts.one);
    if ((r = check_launch(s, "top height"))) return r;
  }
  s->frames++;
  return encode_stats_copy(s);
}

extern "C" int sl_step(sl_sim *s, double frame_dt, sl_emit_fn fn, void *user) {
  uint32_t n = 0;
  double dt = 0;
  int r = sl_step_begin(s, frame_dt, &n, &dt);
  if (r || n == 0) return r;
  sl_emit list[SL_MAX_EMIT];
  for (uint32_t i = 0; i < n; i++) {
    uint32_t m = fn ? fn(user, s, dt, s->time, list, SL_MAX_EMIT) : 0;
    if (m > SL_MAX_EMIT) m = SL_MAX_EMIT;
    r = sl_substep(s, dt, list, m, nullptr);
    if (r) return r;
  }
  return sl_step_end(s, frame_dt);
}

// FluidSim.step with the sources of `items` as the emitter (55-emit.js:13-69 per substep)
extern "C" int sl_step_items(sl_sim *s, double frame_dt, const sl_item *items, uint32_t n) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (n && !items) FAIL(s, SL_EARG, "sl_step_items: null items");
  uint32_t ns = 0;
  double dt = 0;
  int r = sl_step_begin(s, frame_dt, &ns, &dt);
  if (r || ns == 0) return r;
  sl_emit list[SL_MAX_EMIT];
  for (uint32_t i = 0; i < ns; i++) {
    uint32_t m = 0;
    if ((r = sl_scene_source_emits(&s->scene, items, n, dt, list, SL_MAX_EMIT, &m))) FAIL(s, r, "sl_step_items: emitters");
    if ((r = sl_substep(s, dt, list, m, nullptr))) return r;
  }
  return sl_step_end(s, frame_dt);
}

// ------------------------------------------------------------------------------------------------
// stats, 50-sim.js:338-363
// ------------------------------------------------------------------------------------------------
static void apply_stats(sl_sim *s) {
  const SlCounters &u = *s->h_counters;
  s->count = u.count;
  s->impacts = u.impacts;
  s->ww_count = u.wwCount < s->ww_cap ? u.wwCount : s->ww_cap;
  float vm;
  memcpy(&vm, &u.vmax, 4);
  if (isfinite(vm)) s->vmax = vm;
  if (s->count < s->cap * 0.9) s->full = false;
  uint64_t ub = (uint64_t)u.count + (s->emitted_total - s->emitted_at_copy);
  s->p_ub = (uint32_t)(ub < s->cap ? ub : s->cap);
}
extern "C" int sl_poll_stats(sl_sim *s, int *updated) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (updated) *updated = 0;
  if (!s->stats_pending) return SL_OK;
  hipError_t e = hipEventQuery(s->stats_ev);
  if (e == hipErrorNotReady) return SL_OK;
  if (e != hipSuccess) { s->gpu_errors++; FAIL(s, SL_EHIP, "stats readback: %s", hipGetErrorString(e)); }
  s->stats_pending = false;
  apply_stats(s);
  if (updated) *updated = 1;
  return SL_OK;
}
extern "C" int sl_wait_idle(const sl_sim *s) {
  if (!s || !s->stream) return SL_EARG;
  return hipStreamSynchronize(s->stream) == hipSuccess ? SL_OK : SL_EHIP;
}
extern "C" int sl_sync(sl_sim *s) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return SL_OK;
}
extern "C" int sl_stats(sl_sim *s, sl_stats_t *o) {
  if (!s || !o) FAIL((sl_sim *)nullptr, SL_EARG, "sl_stats: null argument");
  sl_k_count(launch_of(s, 1), 0, 0);  // G2P's speed slots into vmax
  HIPCHK(s, hipStreamSynchronize(s->stream));
  int r = sl_poll_stats(s, nullptr);
  if (r) return r;
  // a fresh readback of the counters of the last finished work
  HIPCHK(s, hipMemcpy(s->h_counters, s->ht.C, sizeof(SlCounters), hipMemcpyDeviceToHost));
  s->emitted_at_copy = s->emitted_total;
  apply_stats(s);
  memset(o, 0, sizeof *o);
  o->count = s->count; o->ww_count = s->ww_count; o->impacts = s->impacts; o->frames = s->frames;
  o->full = s->full; o->gpu_errors = s->gpu_errors; o->substeps_last = s->substeps_last;
  o->vmax = s->vmax; o->time = s->time; o->liters = s->count * s->vp * 1000;
  o->grid[0] = s->nx; o->grid[1] = s->ny; o->grid[2] = s->nz;
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// clear / save / restore, 50-sim.js:111-141, :365-373
// ------------------------------------------------------------------------------------------------
// A counters readback enqueued before a clear or restore carries the old count: wait for it and drop
// it, so that the next sl_poll_stats cannot replace the host count (and the launch bound p_ub) with it.
static int drop_pending_stats(sl_sim *s) {
  if (!s->stats_pending) return SL_OK;
  HIPCHK(s, hipEventSynchronize(s->stats_ev));
  s->stats_pending = false;
  s->emitted_at_copy = s->emitted_total;
  return SL_OK;
}

extern "C" int sl_clear_water(sl_sim *s) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  {
    int r = drop_pending_stats(s);
    if (r) return r;
  }
  sl_k_count(launch_of(s, 1), 4, 0);
  sl_k_fill_u32(s->stream, (uint32_t *)s->ht.pres, 0, (size_t)s->hcells * 2);
  if (s->R) {
    sl_k_fill_u32(s->stream, (uint32_t *)s->ht.rstate, 0, (size_t)s->M * s->M * 2);
    sl_k_fill_u32(s->stream, (uint32_t *)s->ht.rsrc, 0, (size_t)(s->R + 1) * (s->R + 1));
  }
  s->count = 0; s->ww_count = 0; s->impacts = 0; s->full = false; s->time = 0; s->vmax = 1;
  s->p_ub = 0;
  s->scene.sources.clear();
  return check_launch(s, "clear_water");
}

extern "C" int sl_save_water(sl_sim *s, uint32_t *id) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  auto &v = s->saved;
  const size_t pbytes = (size_t)s->cap * 25, wbytes = (size_t)s->ww_cap * 32;
  const size_t rbytes = (size_t)s->M * s->M * 8;
  if (!v.parts) {
    HIPCHK(s, hipMalloc(&v.parts, pbytes));
    HIPCHK(s, hipMalloc(&v.ww, wbytes ? wbytes : 32));
    HIPCHK(s, hipMalloc(&v.counters, 32));
    if (rbytes) HIPCHK(s, hipMalloc(&v.ripple, rbytes));
  }
  const SlTank &T = s->ht;
  int r = dev_cur(s);
  if (r) return r;
  const int c = s->cur, w = s->ww_cur;
  uint8_t *p = (uint8_t *)v.parts;
  const size_t cf = (size_t)s->cap * 4;
  const void *src[6] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c]};
  for (int i = 0; i < 6; i++) HIPCHK(s, hipMemcpyAsync(p + i * cf, src[i], cf, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(p + 6 * cf, T.pf[c], s->cap, hipMemcpyDeviceToDevice, s->stream));
  if (wbytes) {
    HIPCHK(s, hipMemcpyAsync(v.ww, T.wp[w], wbytes / 2, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync((uint8_t *)v.ww + wbytes / 2, T.wv[w], wbytes / 2, hipMemcpyDeviceToDevice, s->stream));
  }
  sl_k_count(launch_of(s, 1), 0, 0);  // G2P's speed slots into vmax before the counters are copied
  HIPCHK(s, hipMemcpyAsync(v.counters, T.C, 32, hipMemcpyDeviceToDevice, s->stream));
  if (rbytes) HIPCHK(s, hipMemcpyAsync(v.ripple, T.rstate, rbytes, hipMemcpyDeviceToDevice, s->stream));
  v.time = s->time; v.count = s->count; v.ww_count = s->ww_count; v.impacts = s->impacts; v.full = s->full;
  v.cur = c; v.ww_cur = w;
  v.sources = s->scene.sources;
  v.seed_state = s->scene.seed_state;
  v.valid = true;
  v.id = ++g_copies;
  if (id) *id = v.id;
  return SL_OK;
}
extern "C" int sl_has_water_copy(const sl_sim *s, uint32_t id) { return s && s->saved.valid && s->saved.id == id; }
extern "C" int sl_restore_water(sl_sim *s, uint32_t id) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (!sl_has_water_copy(s, id)) return 0;
  auto &v = s->saved;
  const SlTank &T = s->ht;
  if (dev_cur(s)) return 0;
  const int c = s->cur, w = s->ww_cur;
  const size_t cf = (size_t)s->cap * 4, wbytes = (size_t)s->ww_cap * 32, rbytes = (size_t)s->M * s->M * 8;
  uint8_t *p = (uint8_t *)v.parts;
  void *dst[6] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c]};
  for (int i = 0; i < 6; i++) HIPCHK(s, hipMemcpyAsync(dst[i], p + i * cf, cf, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(T.pf[c], p + 6 * cf, s->cap, hipMemcpyDeviceToDevice, s->stream));
  if (wbytes) {
    HIPCHK(s, hipMemcpyAsync(T.wp[w], v.ww, wbytes / 2, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(T.wv[w], (uint8_t *)v.ww + wbytes / 2, wbytes / 2, hipMemcpyDeviceToDevice, s->stream));
  }
  HIPCHK(s, hipMemcpyAsync(T.C, v.counters, 32, hipMemcpyDeviceToDevice, s->stream));
  sl_k_fill_u32(s->stream, T.vslot, 0, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE);
  if (rbytes) HIPCHK(s, hipMemcpyAsync(T.rstate, v.ripple, rbytes, hipMemcpyDeviceToDevice, s->stream));
  s->time = v.time; s->count = v.count; s->ww_count = v.ww_count; s->impacts = v.impacts; s->full = v.full;
  s->scene.sources = v.sources;
  s->scene.seed_state = v.seed_state;
  // dev_cur waited for the stream, so a readback enqueued before the restore has landed: drop it, and
  // take the launch bound from the restored device count
  s->stats_pending = false;
  {
    SlCounters c;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&c, T.C, sizeof c, hipMemcpyDeviceToHost));
    s->emitted_at_copy = s->emitted_total;
    s->p_ub = c.count < s->cap ? c.count : s->cap;
  }
  v.valid = false;
  return 1;
}

extern "C" int sl_set_ripple_mask(sl_sim *s, int bound) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  s->mask_bound = (bound && s->R) ? 1 : 0;
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// state transfer in the reference layouts
// ------------------------------------------------------------------------------------------------
extern "C" size_t sl_field_bytes(const sl_sim *s, int f) {
  if (!s) return 0;
  const size_t N = s->nodes, C = s->cells, M2 = (size_t)s->M * s->M, R1 = (size_t)(s->R + 1) * (s->R + 1);
  switch (f) {
    case SL_F_PARTICLES: return (size_t)s->cap * 32;
    case SL_F_WW: return (size_t)s->ww_cap * 32;
    case SL_F_COUNTERS: return 32;
    case SL_F_ACC: return N * 24;
    case SL_F_CELLCOUNT: case SL_F_CELLTYPE: case SL_F_DIVG: case SL_F_PRES: return C * 4;
    case SL_F_VEL: case SL_F_VELOLD: case SL_F_FRAC: case SL_F_SVEL: return N * 16;
    case SL_F_OBJS: return SL_MAX_OBJ * 112;
    case SL_F_RSRC: case SL_F_RMASK: return s->R ? R1 * 4 : 0;
    case SL_F_RSTATE: case SL_F_RHAT: case SL_F_RFIELD: return s->R ? M2 * 8 : 0;
    case SL_F_RTEX: return s->R ? R1 * 16 : 0;
    case SL_F_RTEX16: return s->R ? R1 * 8 : 0;
    case SL_F_ARGS: return 64;
    case SL_F_TOPH: return s->R ? (size_t)s->ht.wmap * s->ht.wmap * 4 : 0;
  }
  return 0;
}

static int d2h(sl_sim *s, void *dst, const void *src, size_t bytes) {
  HIPCHK(s, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SL_OK;
}
static int h2d(sl_sim *s, void *dst, const void *src, size_t bytes) {
  HIPCHK(s, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SL_OK;
}
// G2P's speed slots folded into the device vmax, waited for
static int fold_vmax(sl_sim *s) {
  sl_k_count(launch_of(s, 1), 0, 0);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return check_launch(s, "vmax fold");
}
static inline uint32_t packed_slot(const sl_sim *s, uint32_t c) {
  uint32_t i = c % s->nx, j = (c / s->nx) % s->ny, k = c / (s->nx * s->ny);
  return (((i + j + k) & 1) ? s->hcells : 0) + (i >> 1) + s->hw * (j + s->ny * k);
}
static uint16_t f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, e = (x >> 23) & 0xffu, m = x & 0x7fffffu;
  if (e == 0xff) return (uint16_t)(sign | 0x7c00u | (m ? 0x200u : 0));
  int32_t ee = (int32_t)e - 127 + 15;
  if (ee >= 31) return (uint16_t)(sign | 0x7c00u);
  if (ee <= 0) {
    if (ee < -10) return (uint16_t)sign;
    m |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - ee);
    uint32_t hm = m >> shift, rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t hm = m >> 13, rem = m & 0x1fffu;
  uint32_t out = sign | ((uint32_t)ee << 10) | hm;
  if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) out++;
  return (uint16_t)out;
}

extern "C" int sl_readback(sl_sim *s, int f, void *dst, size_t bytes) {
  if (!s || !dst) FAIL((sl_sim *)nullptr, SL_EARG, "sl_readback: null argument");
  size_t need = sl_field_bytes(s, f);
  if (!need || bytes < need) FAIL(s, SL_EARG, "sl_readback: field %d needs %zu bytes, got %zu", f, need, bytes);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const SlTank &T = s->ht;
  const size_t N = s->nodes, C = s->cells;
  int r = SL_OK;
  float *o = (float *)dst;
  switch (f) {
    case SL_F_PARTICLES: {
      const size_t P = s->cap;
      std::vector<float> a(P * 6);
      std::vector<uint8_t> fl(P);
      if ((r = dev_cur(s))) return r;
      const int c = s->cur;
      const float *srcs[6] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c]};
      for (int q = 0; q < 6; q++) if ((r = d2h(s, a.data() + q * P, srcs[q], P * 4))) return r;
      if ((r = d2h(s, fl.data(), T.pf[c], P))) return r;
      for (size_t i = 0; i < P; i++) {
        float *p = o + i * 8;
        p[0] = a[i]; p[1] = a[P + i]; p[2] = a[2 * P + i]; p[3] = (fl[i] & 1) ? 1.0f : 0.0f;
        p[4] = a[3 * P + i]; p[5] = a[4 * P + i]; p[6] = a[5 * P + i]; p[7] = (fl[i] & 2) ? 1.0f : 0.0f;
      }
      return SL_OK;
    }
    case SL_F_WW: {
      const size_t W = s->ww_cap;
      std::vector<float> a(W * 4), b(W * 4);
      if ((r = d2h(s, a.data(), T.wp[s->ww_cur], W * 16))) return r;
      if ((r = d2h(s, b.data(), T.wv[s->ww_cur], W * 16))) return r;
      for (size_t i = 0; i < W; i++) { memcpy(o + i * 8, &a[i * 4], 16); memcpy(o + i * 8 + 4, &b[i * 4], 16); }
      return SL_OK;
    }
    case SL_F_COUNTERS:
      if ((r = fold_vmax(s))) return r;
      return d2h(s, dst, T.C, 32);
    case SL_F_ACC: {  // device: 3 u64 planes (mom << 32 | wt); reference: 6 i32 per node
      std::vector<uint64_t> a(N * 3);
      if ((r = d2h(s, a.data(), T.acc, N * 24))) return r;
      int32_t *oi = (int32_t *)dst;
      for (size_t n = 0; n < N; n++)
        for (int c = 0; c < 3; c++) {
          uint64_t v = a[c * N + n];
          oi[n * 6 + 2 * c] = (int32_t)(uint32_t)(v >> 32);
          oi[n * 6 + 2 * c + 1] = (int32_t)(uint32_t)v;
        }
      return SL_OK;
    }
    case SL_F_CELLCOUNT: return d2h(s, dst, T.cellCount, C * 4);
    case SL_F_CELLTYPE: {
      std::vector<uint32_t> a(C);
      if ((r = d2h(s, a.data(), T.cellCount, C * 4))) return r;
      uint32_t *ou = (uint32_t *)dst;
      for (size_t c = 0; c < C; c++) ou[c] = a[c] > 0 ? 1u : 0u;
      return SL_OK;
    }
    case SL_F_VEL: case SL_F_VELOLD: case SL_F_FRAC: case SL_F_SVEL: {
      const float *src = f == SL_F_VEL ? T.vel : f == SL_F_VELOLD ? T.velOld : f == SL_F_FRAC ? T.frac : T.svel;
      std::vector<float> a(N * 3);
      if ((r = d2h(s, a.data(), src, N * 12))) return r;
      for (size_t n = 0; n < N; n++) {
        uint32_t m = 0;
        for (int c = 0; c < 3; c++) {
          float v = a[c * N + n];
          if (f == SL_F_VELOLD) {  // the mask bits out of velOld (which reads 0 where the mask is off)
            const bool on = f32_bits(v) != SL_MASKED_OUT;
            m |= (on ? 1u : 0u) << c;
            if (!on) v = 0.0f;
          }
          o[n * 4 + c] = v;
        }
        o[n * 4 + 3] = (float)m;
      }
      return SL_OK;
    }
    case SL_F_DIVG: return d2h(s, dst, T.divg, C * 4);
    case SL_F_PRES: {
      std::vector<float> a((size_t)s->hcells * 2);
      if ((r = d2h(s, a.data(), T.pres, a.size() * 4))) return r;
      for (uint32_t c = 0; c < C; c++) o[c] = a[packed_slot(s, c)];
      return SL_OK;
    }
    case SL_F_OBJS: memcpy(dst, s->objs, sizeof s->objs); return SL_OK;
    case SL_F_RSRC: return d2h(s, dst, T.rsrc, need);
    case SL_F_RSTATE: return d2h(s, dst, T.rstate, need);
    case SL_F_RHAT: return d2h(s, dst, T.rhat, need);
    case SL_F_RFIELD: return d2h(s, dst, T.rb, need);
    case SL_F_RMASK: return d2h(s, dst, T.rmask, need);
    case SL_F_RTEX: return d2h(s, dst, T.rtex, need);
    case SL_F_TOPH: return d2h(s, dst, T.topH, need);
    case SL_F_ARGS: return d2h(s, dst, T.args, need);
    case SL_F_RTEX16: {
      size_t R1 = (size_t)(s->R + 1) * (s->R + 1);
      std::vector<float> a(R1 * 4);
      if ((r = d2h(s, a.data(), T.rtex, R1 * 16))) return r;
      uint16_t *oh = (uint16_t *)dst;
      for (size_t i = 0; i < R1 * 4; i++) oh[i] = f32_to_f16(a[i]);
      return SL_OK;
    }
  }
  FAIL(s, SL_EARG, "sl_readback: unknown field %d", f);
}

extern "C" int sl_upload(sl_sim *s, int f, const void *src, size_t bytes) {
  if (!s || !src) FAIL((sl_sim *)nullptr, SL_EARG, "sl_upload: null argument");
  size_t need = sl_field_bytes(s, f);
  if (!need || bytes < need) FAIL(s, SL_EARG, "sl_upload: field %d needs %zu bytes, got %zu", f, need, bytes);
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const SlTank &T = s->ht;
  const size_t N = s->nodes, C = s->cells;
  const float *in = (const float *)src;
  int r = SL_OK;
  switch (f) {
    case SL_F_PARTICLES: {
      const size_t P = s->cap;
      std::vector<float> a(P * 6);
      std::vector<uint8_t> fl(P);
      for (size_t i = 0; i < P; i++) {
        const float *p = in + i * 8;
        a[i] = p[0]; a[P + i] = p[1]; a[2 * P + i] = p[2];
        a[3 * P + i] = p[4]; a[4 * P + i] = p[5]; a[5 * P + i] = p[6];
        fl[i] = (uint8_t)((p[3] < 0.5f ? 0u : 1u) | (p[7] > 0.5f ? 2u : 0u));
      }
      if ((r = dev_cur(s))) return r;
      const int c = s->cur;
      float *dsts[6] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c]};
      for (int q = 0; q < 6; q++) if ((r = h2d(s, dsts[q], a.data() + q * P, P * 4))) return r;
      return h2d(s, T.pf[c], fl.data(), P);
    }
    case SL_F_WW: {
      const size_t W = s->ww_cap;
      std::vector<float> a(W * 4), b(W * 4);
      for (size_t i = 0; i < W; i++) { memcpy(&a[i * 4], in + i * 8, 16); memcpy(&b[i * 4], in + i * 8 + 4, 16); }
      if ((r = h2d(s, T.wp[s->ww_cur], a.data(), W * 16))) return r;
      return h2d(s, T.wv[s->ww_cur], b.data(), W * 16);
    }
    case SL_F_COUNTERS: {
      r = h2d(s, T.C, src, 32);
      if (r) return r;
      sl_k_fill_u32(s->stream, T.vslot, 0, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE);
      HIPCHK(s, hipStreamSynchronize(s->stream));
      const SlCounters *cc = (const SlCounters *)src;
      s->p_ub = cc->count < s->cap ? cc->count : s->cap;
      return SL_OK;
    }
    case SL_F_ACC: {
      std::vector<uint64_t> a(N * 3);
      const int32_t *ii = (const int32_t *)src;
      for (size_t n = 0; n < N; n++)
        for (int c = 0; c < 3; c++)
          a[c * N + n] = ((uint64_t)(uint32_t)ii[n * 6 + 2 * c] << 32) + (uint64_t)(uint32_t)ii[n * 6 + 2 * c + 1];
      return h2d(s, T.acc, a.data(), N * 24);
    }
    case SL_F_CELLCOUNT: return h2d(s, T.cellCount, src, C * 4);
    case SL_F_VEL: case SL_F_VELOLD: case SL_F_FRAC: case SL_F_SVEL: {
      float *d = f == SL_F_VEL ? T.vel : f == SL_F_VELOLD ? T.velOld : f == SL_F_FRAC ? T.frac : T.svel;
      std::vector<float> a(N * 3);
      for (size_t n = 0; n < N; n++) {
        a[n] = in[n * 4]; a[N + n] = in[n * 4 + 1]; a[2 * N + n] = in[n * 4 + 2];
        if (f != SL_F_VELOLD) continue;
        // velOld: the mask bits (w: bit c for component c) go in as the masked-out pattern
        const float w = in[n * 4 + 3];
        const uint32_t m = w > 0 ? (uint32_t)w & 7u : 0u;
        for (int c = 0; c < 3; c++) {
          float &v = a[c * N + n];
          if (!((m >> c) & 1u)) v = bits_f32(SL_MASKED_OUT);
          else if (f32_bits(v) == SL_MASKED_OUT) v = bits_f32(0x7fc00000u);  // a NaN, not a mask
        }
      }
      if ((r = h2d(s, d, a.data(), N * 12))) return r;
      if (f == SL_F_FRAC || f == SL_F_SVEL) s->geom_dirty = false;  // explicit geometry replaces GEOM
      return SL_OK;
    }
    case SL_F_DIVG: return h2d(s, T.divg, src, C * 4);
    case SL_F_PRES: {
      std::vector<float> a((size_t)s->hcells * 2, 0.0f);
      for (uint32_t c = 0; c < C; c++) a[packed_slot(s, c)] = in[c];
      return h2d(s, T.pres, a.data(), a.size() * 4);
    }
    case SL_F_RSRC: return h2d(s, T.rsrc, src, need);
    case SL_F_RSTATE: return h2d(s, T.rstate, src, need);
    case SL_F_TOPH: return h2d(s, T.topH, src, need);
  }
  FAIL(s, SL_EARG, "sl_upload: field %d is read-only", f);
}

extern "C" int sl_set_scalar(sl_sim *s, const char *name, double v) {
  if (!s || !name) FAIL((sl_sim *)nullptr, SL_EARG, "sl_set_scalar: null argument");
  if (!strcmp(name, "count")) s->count = (uint32_t)v;
  else if (!strcmp(name, "time")) s->time = v;
  else if (!strcmp(name, "frame")) s->frame = (uint32_t)v;
  else if (!strcmp(name, "vmax")) s->vmax = v;
  else if (!strcmp(name, "full")) s->full = v != 0;
  else if (!strcmp(name, "p_ub")) s->p_ub = (uint32_t)v;
  else if (!strcmp(name, "geom_dirty")) s->geom_dirty = v != 0;
  else if (!strcmp(name, "p2g_variant")) {
    if (!(v >= 0 && v <= SL_P2G_BINNED)) FAIL(s, SL_EARG, "p2g_variant must be 0..%d", SL_P2G_BINNED);
    s->p2g_variant = (int)v;
  }
  else if (!strcmp(name, "g2p_variant")) {
    if (!(v >= 0 && v <= 3)) FAIL(s, SL_EARG, "g2p_variant must be 0..3");
    s->g2p_variant = (int)v;
  }
  else if (!strcmp(name, "reorder_every")) s->cfg.reorder_every = (uint32_t)v;
  else FAIL(s, SL_EARG, "unknown scalar '%s'", name);
  return SL_OK;
}
extern "C" int sl_get_scalar(const sl_sim *s, const char *name, double *v) {
  if (!s || !name || !v) FAIL((sl_sim *)nullptr, SL_EARG, "sl_get_scalar: null argument");
  if (!strcmp(name, "count")) *v = s->count;
  else if (!strcmp(name, "time")) *v = s->time;
  else if (!strcmp(name, "frame")) *v = s->frame;
  else if (!strcmp(name, "vmax")) *v = s->vmax;
  else if (!strcmp(name, "full")) *v = s->full;
  else if (!strcmp(name, "p_ub")) *v = s->p_ub;
  else if (!strcmp(name, "cur")) {
    int r = dev_cur(const_cast<sl_sim *>(s));
    if (r) return r;
    *v = s->cur;
  }
  else if (!strcmp(name, "ww_cur")) *v = s->ww_cur;
  else if (!strcmp(name, "impacts")) *v = s->impacts;
  else if (!strcmp(name, "bytes")) *v = (double)s->bytes;
  else if (!strcmp(name, "ww_count")) *v = s->ww_count;
  else if (!strcmp(name, "frames")) *v = s->frames;
  else if (!strcmp(name, "p2g_variant")) *v = s->p2g_variant;
  else if (!strcmp(name, "g2p_variant")) *v = s->g2p_variant;
  else { ((sl_sim *)s)->err = std::string("unknown scalar '") + name + "'"; return SL_EARG; }
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// single stages (stage-level parity)
// ------------------------------------------------------------------------------------------------
extern "C" int sl_run_stage(sl_sim *s, int stage, uint32_t arg, double dt) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  write_params(s, dt);
  int r = push_dyn(s);
  if (r) return r;
  switch (stage) {
    case SL_S_GEOM: sl_k_geom(launch_of(s, s->nodes)); s->geom_dirty = false; break;
    case SL_S_P2G:
      sl_k_fill_u32(s->stream, (uint32_t *)s->ht.acc, 0, (size_t)s->nodes * 6);
      sl_k_fill_u32(s->stream, s->ht.cellCount, 0, s->cells);
      encode_p2g(s, s->p2g_variant);
      break;
    case SL_S_NORMFORCE: sl_k_normforce(launch_of(s, s->nodes)); break;
    case SL_S_DIVERGENCE: sl_k_divsetup(launch_of(s, s->cells)); break;
    case SL_S_PRESSURE: sl_k_sor_rec(s->stream, sl_sor_rec(s->ht), arg & 1u); break;
    case SL_S_SOR:
      if (s->solver == SL_SOLVER_MGPCG) {
        sl_mg_solve(s->stream, s->d_tank, s->mg, s->mg_seed ? 1 : 0);
        s->mg_seed = false;
        break;
      }
      {
        const SlSorRec R = sl_sor_rec(s->ht);
        for (uint32_t i = 0; i < s->iters; i++) { sl_k_sor_rec(s->stream, R, 0); sl_k_sor_rec(s->stream, R, 1); }
      }
      break;
    case SL_S_PROJECT:
      if (s->solver == SL_SOLVER_MGPCG) sl_k_project_mg(s->stream, s->d_tank, s->nodes, s->mg.xs);
      else sl_k_project(launch_of(s, s->nodes));
      break;
    case SL_S_G2P:
      encode_objmask(s);
      sl_k_g2p(launch_of(s, s->p_ub), s->g2p_variant);
      sl_k_count(launch_of(s, 1), 0, 0);  // the speed slots into vmax
      break;
    case SL_S_COMPACT: return encode_compact(s);
    case SL_S_WW:
      sl_k_ww(launch_of(s, s->ww_cap), (float)dt);
      sl_k_count(launch_of(s, 1), 3, 0);
      s->ww_cur ^= 1;
      break;
    case SL_S_RIPPLE:
      if ((r = set_ripple_uniform(s, dt, frame_mean_depth(s)))) return r;
      return encode_ripple(s);
    case SL_S_DROPLETS:
      if ((r = set_ripple_uniform(s, dt, frame_mean_depth(s)))) return r;
      return encode_droplets(s);
    default: FAIL(s, SL_EARG, "unknown stage %d", stage);
  }
  return check_launch(s, "stage");
}

// ------------------------------------------------------------------------------------------------
// timing
// ------------------------------------------------------------------------------------------------
extern "C" int sl_timing_enable(sl_sim *s, int enable) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  timing_collect(s);
  s->timing = enable != 0;
  return SL_OK;
}
extern "C" int sl_timing_read(sl_sim *s, double *ms, uint32_t *launches, int n, int reset) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  timing_collect(s);
  for (int i = 0; i < n && i < SL_T_COUNT_; i++) {
    if (ms) ms[i] = s->t_ms[i];
    if (launches) launches[i] = s->t_launch[i];
  }
  if (reset) { memset(s->t_ms, 0, sizeof s->t_ms); memset(s->t_launch, 0, sizeof s->t_launch); }
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// autotuning of kernel variants on the live state
// ------------------------------------------------------------------------------------------------
// Every variant of a kernel computes the same result bit for bit (integer sums, identical
// arithmetic); only the work decomposition and the register budget differ. The tuner times each
// variant with HIP events on the simulation's own state between two frames, restores everything the
// trials touched, and keeps the fastest. It needs a state produced by at least one completed step.
static float time_region(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}
// First P2G variant the tuners time: above 256 Ki cells only the binned variant, whose cost does not
// depend on the particle order; the chunked variants measured 10-30x slower there (a tuning trial of
// one of them on a 4 M-cell tank costs ~10-30 ms of device time).
static int p2g_first_trial(uint32_t cells) { return cells > 262144u ? SL_P2G_BINNED : 0; }

extern "C" int sl_autotune(sl_sim *s, uint32_t reps, uint32_t *choice) {
  if (!s) FAIL((sl_sim *)nullptr, SL_EARG, "null sim");
  if (!reps) reps = 3;
  int rc = dev_cur(s);
  if (rc) return rc;
  const SlTank &T = s->ht;
  const int c = s->cur;
  const size_t P = s->p_ub, C = s->cells;
  const size_t R1 = s->R ? (size_t)(s->R + 1) * (s->R + 1) : 0;
  // snapshot of what P2G and G2P write: particles, counters, speed slots, cell counts, ripple impulses
  std::vector<void *> snap;
  bool ok = true;
  auto dup = [&](const void *src, size_t bytes) -> void * {
    void *p = nullptr;
    if (!bytes) return nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) { ok = false; return nullptr; }
    (void)hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToDevice, s->stream);
    snap.push_back(p);
    return p;
  };
  void *psrc[7] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c], T.pf[c]};
  void *pcpy[7];
  for (int i = 0; i < 7; i++) pcpy[i] = dup(psrc[i], i < 6 ? P * 4 : P);
  void *ccpy = dup(T.C, sizeof(SlCounters));
  void *vcpy = dup(T.vslot, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE * 4);
  void *cccpy = dup(T.cellCount, C * 4);
  void *rcpy = R1 ? dup(T.rsrc, R1 * 4) : nullptr;
  if (!ok) {
    (void)hipStreamSynchronize(s->stream);
    for (void *p : snap) (void)hipFree(p);
    FAIL(s, SL_ENOMEM, "sl_autotune: no memory for the state snapshot");
  }
  auto restore = [&]() {
    for (int i = 0; i < 7; i++)
      if (pcpy[i]) (void)hipMemcpyAsync(psrc[i], pcpy[i], i < 6 ? P * 4 : P, hipMemcpyDeviceToDevice, s->stream);
    if (ccpy) (void)hipMemcpyAsync(T.C, ccpy, sizeof(SlCounters), hipMemcpyDeviceToDevice, s->stream);
    if (vcpy) (void)hipMemcpyAsync(T.vslot, vcpy, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE * 4, hipMemcpyDeviceToDevice, s->stream);
    if (cccpy) (void)hipMemcpyAsync(T.cellCount, cccpy, C * 4, hipMemcpyDeviceToDevice, s->stream);
    if (rcpy) (void)hipMemcpyAsync(T.rsrc, rcpy, R1 * 4, hipMemcpyDeviceToDevice, s->stream);
  };
  hipEvent_t a = ev_get(s), b = ev_get(s);
  // P2G variants: clear, accumulate, time; the accumulators are cleared again afterwards
  float best = 1e30f;
  int bestv = s->p2g_variant;
  for (int v = p2g_first_trial(s->cells); v <= SL_P2G_BINNED; v++) {
    float tot = 0;
    for (uint32_t r = 0; r <= reps; r++) {
      sl_k_fill_u32(s->stream, (uint32_t *)T.acc, 0, (size_t)s->nodes * 6);
      sl_k_fill_u32(s->stream, T.cellCount, 0, s->cells);
      (void)hipEventRecord(a, s->stream);
      encode_p2g(s, v);
      (void)hipEventRecord(b, s->stream);
      float ms = time_region(a, b);
      if (r) tot += ms;  // the first run warms the caches
    }
    if (tot < best) { best = tot; bestv = v; }
  }
  s->p2g_variant = bestv;
  sl_k_fill_u32(s->stream, (uint32_t *)T.acc, 0, (size_t)s->nodes * 6);
  // G2P variants: restore the particles before every run
  best = 1e30f;
  int bestg = s->g2p_variant;
  for (int v = 0; v < 4; v++) {
    float tot = 0;
    for (uint32_t r = 0; r <= reps; r++) {
      restore();
      (void)hipEventRecord(a, s->stream);
      sl_k_g2p(launch_of(s, P), v);
      (void)hipEventRecord(b, s->stream);
      float ms = time_region(a, b);
      if (r) tot += ms;
    }
    if (tot < best) { best = tot; bestg = v; }
  }
  s->g2p_variant = bestg;
  restore();
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (void *p : snap) (void)hipFree(p);
  s->ev_pool.push_back(a);
  s->ev_pool.push_back(b);
  if (choice) { choice[0] = (uint32_t)s->p2g_variant; choice[1] = (uint32_t)s->g2p_variant; }
  return check_launch(s, "autotune");
}


// ================================================================================================
// Batched tanks: many independent simulations in one tank table, stepped with one launch per stage
// for all of them and optionally replayed from a captured hipGraph. No counterpart in the reference
// (it runs one tank per page). A tank of a batch keeps the single-simulation semantics of one
// substep of dt per frame. Kernels find their tank through work maps (prefix tables of workgroup
// counts per table slot) and read the tank's buffer index, counts and parameters from the device
// table, so a captured frame stays valid across frames, emissions, object moves and re-orderings.
// Slots are added and removed while the batch runs (continuous batching); per-tank kernels walk
// the list of stepped slots.
// ================================================================================================
#define SL_BUCKETS 3
#define SL_SOR_BANDS 16  // SOR launch maps: band j covers the tanks with more than band_start[j] iterations
enum {
  BM_NODES, BM_CELLS, BM_HCELLS, BM_COMPACT, BM_WW, BM_ELEMS, BM_TEX, BM_ROWS, BM_COLS,
  BM_P2G0, BM_G2P0 = BM_P2G0 + SL_BUCKETS, BM_SCRATCH = BM_G2P0 + SL_BUCKETS,
  BM_BINP, BM_BINS, BM_BINT, BM_BINTILE,  // binned P2G: particles, scan blocks, tanks, tiles
  BM_TOPP, BM_TOPT,                       // top-height map: particles, texels
  BM_SOR0, BM_COUNT = BM_SOR0 + SL_SOR_BANDS
};

// frame++ and the per-substep uniform (50-sim.js:151-162, :308-317), and the frame's ripple uniform
// (64-ripple.js:223-226) with the mean depth of 50-sim.js:330-333 from the count at frame start
__global__ void k_batch_advance(SlTank *tanks, const uint32_t *list, int n, double dtd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SlTank &T = tanks[list[i]];
  if (!(dtd > 0.0)) dtd = T.dt_frame;  // each tank's own frame length
  const float dt = (float)dtd;
  T.frame += 1;
  T.dt = dt;
  T.gdt = T.g * T.dt;
  T.time_f = (float)T.time_d;
  T.time_d += dtd;
  T.emit_k = 0;
  T.emit_total = 0;
  if (T.R) {
    double liters = (double)T.C->count * T.vp_d * 1000;
    double md = liters / 1000 / ((double)(T.nx * T.nz) * T.h_d * T.h_d);
    md = md < 0.003 ? 0.003 : (md > T.rim_d ? T.rim_d : md);
    if (T.rdepth > 0.0f) md = T.rdepth;
    T.rLx = (float)(T.nx * T.h_d);
    T.rLz = (float)(T.nz * T.h_d);
    T.rdt = dt;
    T.rmd = (float)md;
  }
}
__global__ __launch_bounds__(SL_WG) void k_batch_clear_cells(SlWork W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = W.tanks[t];
  uint32_t c = lb * SL_WG + threadIdx.x;
  if (c < T.cells) T.cellCount[c] = 0u;
}
// kt[2 slot] = emission records, kt[2 slot + 1] = particles requested (0: the slot does not emit)
__global__ void k_batch_emit_setup(SlTank *tanks, const uint32_t *kt, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !kt[2 * t + 1]) return;
  tanks[t].emit_k = kt[2 * t];
  tanks[t].emit_total = kt[2 * t + 1];
}

struct sl_batch {
  uint32_t cap = 0;  // table slots
  int device = 0;
  std::vector<sl_sim *> sims;   // per slot (nullptr: empty)
  std::vector<char> stepped;    // per slot: advanced by sl_batch_step
  std::vector<char> copy_valid; // per slot: the last counters copy holds this tank
  std::vector<int> bucket;      // per slot: size bucket for the variant choice
  hipStream_t stream = nullptr;
  SlTank *d_tanks = nullptr;
  SlCounters *d_counters = nullptr, *h_counters = nullptr;
  // a device block with two pinned host copies, written by stream-ordered uploads (kernels of
  // earlier frames keep reading the previous contents)
  struct Ring {
    uint8_t *d = nullptr, *h[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool valid[2] = {false, false};
    int cur = 0;
    size_t bytes = 0;
  } maps, emit;
  // maps block: BM_COUNT prefix tables of cap+1 entries, then the stepped-slot list (cap entries)
  std::vector<uint32_t> maps_host;
  // the maps resolved per workgroup (k_wtab_fill after every maps upload): map w's entries start at
  // wtab + wtab_off[w]
  uint2 *wtab = nullptr;
  size_t wtab_cap = 0;
  uint64_t wtab_off[BM_COUNT] = {};
  uint32_t total[BM_COUNT] = {};
  int map_variant[BM_COUNT];
  uint32_t n_stepped = 0, max_iters = 0;
  bool any_toph = false;
  uint32_t sor_bands = 0, band_start[SL_SOR_BANDS + 1] = {};
  bool any_ripple = false, maps_dirty = true;
  uint32_t max_M = 0;  // largest ripple M of the stepped tanks
  int p2g_variant[SL_BUCKETS] = {SL_P2G_BINNED, SL_P2G_BINNED, SL_P2G_BINNED}, g2p_variant[SL_BUCKETS] = {1, 1, 1};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  uint64_t steps = 0;
  bool timing = false;  // per-step events while enabled
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  std::vector<void *> allocs;
  // the re-orderings due in one frame run as one sort: keys and values (2 x ro_cap each) and the
  // sort's temporary storage, grown on demand
  uint32_t *ro_keys = nullptr, *ro_vals = nullptr;
  void *ro_temp = nullptr;
  size_t ro_temp_bytes = 0;
  uint64_t ro_cap = 0;
  std::string err;
};

static int batch_fail(sl_batch *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  g_err = msg;
  return code;
}

// size buckets for the variant choice: <= 64Ki cells, <= 256Ki cells, larger
static int bucket_of(const sl_sim *s) { return s->cells <= 65536 ? 0 : (s->cells <= 262144 ? 1 : 2); }

static int ring_init(sl_batch *b, sl_batch::Ring &r, size_t bytes) {
  r.bytes = bytes;
  if (hipMalloc((void **)&r.d, bytes) != hipSuccess) return batch_fail(b, SL_ENOMEM, "batch: device block");
  b->allocs.push_back(r.d);
  if (hipMemset(r.d, 0, bytes) != hipSuccess) return batch_fail(b, SL_EHIP, "batch: clear");
  for (int i = 0; i < 2; i++) {
    if (hipHostMalloc((void **)&r.h[i], bytes, hipHostMallocDefault) != hipSuccess)
      return batch_fail(b, SL_ENOMEM, "batch: pinned block");
    memset(r.h[i], 0, bytes);
    if (hipEventCreateWithFlags(&r.ev[i], hipEventDisableTiming) != hipSuccess) return batch_fail(b, SL_EHIP, "event");
  }
  return SL_OK;
}
static void ring_free(sl_batch::Ring &r) {
  for (int i = 0; i < 2; i++) {
    if (r.h[i]) (void)hipHostFree(r.h[i]);
    if (r.ev[i]) (void)hipEventDestroy(r.ev[i]);
  }
}
// the host copy that may be written now (its previous upload has completed)
static uint8_t *ring_host(sl_batch::Ring &r) {
  if (r.valid[r.cur]) { (void)hipEventSynchronize(r.ev[r.cur]); r.valid[r.cur] = false; }
  return r.h[r.cur];
}
static int ring_push(sl_batch *b, sl_batch::Ring &r, size_t bytes) {
  hipError_t e = hipMemcpyAsync(r.d, r.h[r.cur], bytes, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipEventRecord(r.ev[r.cur], b->stream);
  if (e != hipSuccess) return batch_fail(b, SL_EHIP, std::string("batch upload: ") + hipGetErrorString(e));
  r.valid[r.cur] = true;
  r.cur ^= 1;
  return SL_OK;
}

static uint32_t *map_dev(const sl_batch *b, int which) { return (uint32_t *)b->maps.d + (size_t)which * (b->cap + 1); }
static const uint32_t *list_dev(const sl_batch *b) { return map_dev(b, BM_COUNT); }
static SlLaunch batch_launch(const sl_batch *b, int which) {
  SlLaunch L;
  L.w.tanks = b->d_tanks;
  L.w.start = map_dev(b, which);
  L.w.ntanks = (int32_t)b->cap;
  L.w.total = b->total[which];
  L.w.list = nullptr;
  L.w.wtab = b->total[which] ? b->wtab + b->wtab_off[which] : nullptr;
  L.stream = b->stream;
  return L;
}
// per-tank kernels: one thread / workgroup per stepped slot
static SlLaunch batch_tanks(const sl_batch *b) {
  SlLaunch L;
  L.w.tanks = b->d_tanks;
  L.w.start = nullptr;
  L.w.ntanks = (int32_t)b->n_stepped;
  L.w.total = 0;
  L.w.list = list_dev(b);
  L.stream = b->stream;
  return L;
}
static bool live(const sl_batch *b, uint32_t t) { return b->sims[t] && b->stepped[t]; }

// one prefix table in the host copy of the maps block; `inc` selects the slots that get workgroups
template <class Items, class Inc>
static int put_map(sl_batch *b, int which, Items items, uint32_t per_wg, Inc inc) {
  uint32_t *m = b->maps_host.data() + (size_t)which * (b->cap + 1);
  uint64_t acc = 0;
  for (uint32_t t = 0; t < b->cap; t++) {
    m[t] = (uint32_t)acc;
    if (live(b, t) && inc(t)) acc += ((uint64_t)items(b->sims[t]) + per_wg - 1) / per_wg;
    if (acc > 0x7fffffffull) return batch_fail(b, SL_ECAP, "batch: launch too large");
  }
  m[b->cap] = (uint32_t)acc;
  b->total[which] = (uint32_t)acc;
  return SL_OK;
}
// workgroup b of a map: its table slot and its index among the slot's workgroups
__global__ __launch_bounds__(SL_WG) void k_wtab_fill(const uint32_t *start, int ntanks, uint2 *out, uint32_t total) {
  const uint32_t b = blockIdx.x * SL_WG + threadIdx.x;
  if (b >= total) return;
  const int t = tank_of(start, ntanks, b);
  out[b] = make_uint2((uint32_t)t, b - start[t]);
}
static void drop_graph(sl_batch *b);
// the maps block, then every map resolved per workgroup (both stream-ordered after the kernels of
// earlier frames, which keep reading the previous contents)
static int maps_upload(sl_batch *b) {
  const size_t bytes = b->maps_host.size() * 4;
  memcpy(ring_host(b->maps), b->maps_host.data(), bytes);
  int r = ring_push(b, b->maps, bytes);
  if (r) return r;
  uint64_t need = 0;
  for (int w = 0; w < BM_COUNT; w++) { b->wtab_off[w] = need; need += b->total[w]; }
  if (need > b->wtab_cap) {
    // a larger table: the kernels still in flight read the old one, and a captured frame holds it
    if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "batch: work tables");
    drop_graph(b);
    if (b->wtab) (void)hipFree(b->wtab);
    b->wtab = nullptr;
    b->wtab_cap = 0;
    const size_t cap = (size_t)(need + need / 4 + 4096);
    if (hipMalloc((void **)&b->wtab, cap * sizeof(uint2)) != hipSuccess)
      return batch_fail(b, SL_ENOMEM, "batch: no device memory for the work tables");
    b->wtab_cap = cap;
  }
  for (int w = 0; w < BM_COUNT; w++)
    if (b->total[w])
      hipLaunchKernelGGL(k_wtab_fill, dim3(cdiv(b->total[w], SL_WG)), dim3(SL_WG), 0, b->stream, map_dev(b, w),
                         (int)b->cap, b->wtab + b->wtab_off[w], b->total[w]);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SL_OK : batch_fail(b, SL_EHIP, std::string("batch: work tables: ") + hipGetErrorString(e));
}
static void drop_graph(sl_batch *b) {
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}
static uint32_t cols_of(const sl_sim *s) {
  const uint32_t M = s->M;
  if (!M) return 0;
  uint32_t CW = 4096 / M;
  if (CW > M) CW = M;
  if (CW < 1) CW = 1;
  return (M + CW - 1) / CW;
}

// the four maps of the binned P2G over the slots `inc` selects
template <class Inc>
static int put_bin_maps(sl_batch *b, Inc inc) {
  int r;
  if ((r = put_map(b, BM_BINP, [](const sl_sim *s) { return (uint64_t)s->cap; }, SL_BIN_PER_WG, inc)) ||
      (r = put_map(b, BM_BINS, [](const sl_sim *s) { return (uint64_t)sl_bin_blocks(s->cells); }, 1, inc)) ||
      (r = put_map(b, BM_BINT, [](const sl_sim *) { return (uint64_t)1; }, 1, inc)) ||
      (r = put_map(b, BM_BINTILE, [](const sl_sim *s) { return (uint64_t)sl_bin_tiles(s->nx, s->ny, s->nz); }, 1, inc)))
    return r;
  return SL_OK;
}
static void batch_p2g_binned(const sl_batch *b);

// every work map and the stepped-slot list from the current membership and variant choice
static int batch_rebuild(sl_batch *b) {
  std::fill(b->maps_host.begin(), b->maps_host.end(), 0u);
  uint32_t *list = b->maps_host.data() + (size_t)BM_COUNT * (b->cap + 1);
  b->n_stepped = 0;
  b->max_iters = 0;
  b->any_ripple = false;
  b->max_M = 0;
  for (uint32_t t = 0; t < b->cap; t++) {
    if (!live(b, t)) continue;
    list[b->n_stepped++] = t;
    b->max_iters = std::max(b->max_iters, b->sims[t]->iters);
    if (b->sims[t]->R) b->any_ripple = true;
    b->max_M = std::max(b->max_M, b->sims[t]->M);
  }
  auto all = [](uint32_t) { return true; };
  int r;
  if ((r = put_map(b, BM_NODES, [](const sl_sim *s) { return (uint64_t)s->nodes; }, SL_WG, all)) ||
      (r = put_map(b, BM_CELLS, [](const sl_sim *s) { return (uint64_t)s->cells; }, SL_WG, all)) ||
      (r = put_map(b, BM_HCELLS, [](const sl_sim *s) { return (uint64_t)s->hcells; }, SL_WG, all)) ||
      (r = put_map(b, BM_COMPACT, [](const sl_sim *s) { return (uint64_t)s->cap; }, SL_COMPACT_ITEMS, all)) ||
      (r = put_map(b, BM_WW, [](const sl_sim *s) { return (uint64_t)s->ww_cap; }, SL_WG, all)) ||
      (r = put_map(b, BM_ELEMS, [](const sl_sim *s) { return (uint64_t)s->M * s->M; }, SL_WG, all)) ||
      (r = put_map(b, BM_TEX, [](const sl_sim *s) { return s->R ? (uint64_t)(s->R + 1) * (s->R + 1) : 0; }, SL_WG, all)) ||
      (r = put_map(b, BM_ROWS, [](const sl_sim *s) { return (uint64_t)(s->M / sl_fft_rows_per_wg(s->M)); }, 1, all)) ||
      (r = put_map(b, BM_COLS, [](const sl_sim *s) { return (uint64_t)cols_of(s); }, 1, all)))
    return r;
  auto toph = [&](uint32_t t) { return b->sims[t]->ht.top_height != 0; };
  if ((r = put_map(b, BM_TOPP, [](const sl_sim *s) { return (uint64_t)s->cap; }, SL_WG, toph)) ||
      (r = put_map(b, BM_TOPT, [](const sl_sim *s) { return (uint64_t)s->ht.wmap * s->ht.wmap; }, SL_WG, toph)))
    return r;
  b->any_toph = b->total[BM_TOPT] > 0;
  // P2G / G2P: one map per distinct variant among the buckets that hold stepped tanks
  for (int pass = 0; pass < 2; pass++) {
    const int *var = pass ? b->g2p_variant : b->p2g_variant;
    const int base = pass ? BM_G2P0 : BM_P2G0;
    int used = 0;
    for (int k = 0; k < SL_BUCKETS; k++) { b->map_variant[base + k] = -1; b->total[base + k] = 0; }
    for (int k = 0; k < SL_BUCKETS; k++) {
      bool has = false, seen = false;
      for (uint32_t t = 0; t < b->cap; t++) has |= live(b, t) && b->bucket[t] == k;
      for (int u = 0; u < used; u++) seen |= b->map_variant[base + u] == var[k];
      if (!has || seen) continue;
      const int v = var[k];
      auto inc = [&](uint32_t t) { return var[b->bucket[t]] == v; };
      if (!pass && v == SL_P2G_BINNED) {
        r = put_bin_maps(b, inc);
        b->total[base + used] = 1;  // launched through the BM_BIN* maps
      } else {
        r = put_map(b, base + used, [](const sl_sim *s) { return (uint64_t)s->cap; },
                    pass ? SL_WG : sl_p2g_chunk(v), inc);
      }
      if (r) return r;
      b->map_variant[base + used++] = v;
    }
  }
  b->total[BM_SCRATCH] = 0;
  // SOR bands: iteration i runs over the tanks with iters > i. Band j (iterations band_start[j] ..
  // band_start[j+1]-1) covers the tanks with iters > band_start[j], a superset (k_sor returns for
  // passes beyond a tank's own count); with more distinct counts than bands, bands are merged.
  {
    std::vector<uint32_t> its;
    for (uint32_t t = 0; t < b->cap; t++)
      if (live(b, t)) its.push_back(b->sims[t]->iters);
    std::sort(its.begin(), its.end());
    its.erase(std::unique(its.begin(), its.end()), its.end());
    std::vector<uint32_t> starts{0};
    for (size_t q = 0; q + 1 < its.size(); q++) starts.push_back(its[q]);  // a tank set shrinks after each count
    if (starts.size() > SL_SOR_BANDS) {
      std::vector<uint32_t> keep;
      for (uint32_t j = 0; j < SL_SOR_BANDS; j++) keep.push_back(starts[(size_t)j * starts.size() / SL_SOR_BANDS]);
      starts = keep;
    }
    b->sor_bands = (uint32_t)starts.size();
    for (uint32_t j = 0; j < b->sor_bands; j++) {
      const uint32_t a = starts[j];
      b->band_start[j] = a;
      r = put_map(b, BM_SOR0 + (int)j, [](const sl_sim *s) { return (uint64_t)s->hcells; }, SL_SOR_WG,
                  [&](uint32_t t) { return b->sims[t]->iters > a; });
      if (r) return r;
    }
    b->band_start[b->sor_bands] = b->max_iters;
  }
  drop_graph(b);  // a captured frame holds the old launch shapes
  b->maps_dirty = false;
  return maps_upload(b);
}

extern "C" void sl_batch_destroy(sl_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (sl_sim *s : b->sims) if (s) sl_destroy(s);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->ro_keys) (void)hipFree(b->ro_keys);
  if (b->ro_vals) (void)hipFree(b->ro_vals);
  if (b->ro_temp) (void)hipFree(b->ro_temp);
  if (b->wtab) (void)hipFree(b->wtab);
  ring_free(b->maps);
  ring_free(b->emit);
  if (b->h_counters) (void)hipHostFree(b->h_counters);
  for (hipEvent_t e : b->tev) (void)hipEventDestroy(e);
  if (b->ev_a) (void)hipEventDestroy(b->ev_a);
  if (b->ev_b) (void)hipEventDestroy(b->ev_b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int sl_batch_open(uint32_t capacity, int device, sl_batch **out) {
  if (out) *out = nullptr;
  if (!out || !capacity || capacity > 65536) return batch_fail(nullptr, SL_EARG, "sl_batch_open: bad arguments");
  sl_batch *b = new (std::nothrow) sl_batch();
  if (!b) return batch_fail(nullptr, SL_ENOMEM, "out of host memory");
  auto fail = [&](int code, const std::string &msg) {
    sl_batch_destroy(b);
    return batch_fail(nullptr, code, msg);
  };
  const uint32_t n = capacity;
  b->cap = n;
  b->device = device;
  b->sims.assign(n, nullptr);
  b->stepped.assign(n, 0);
  b->copy_valid.assign(n, 0);
  b->bucket.assign(n, 0);
  for (int i = 0; i < BM_COUNT; i++) b->map_variant[i] = -1;
  if (hipSetDevice(device) != hipSuccess) return fail(SL_EHIP, "sl_batch_open: no such device");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(SL_EHIP, "stream");
  if (hipMalloc((void **)&b->d_tanks, sizeof(SlTank) * n) != hipSuccess) return fail(SL_ENOMEM, "tank table");
  b->allocs.push_back(b->d_tanks);
  if (hipMemset(b->d_tanks, 0, sizeof(SlTank) * n) != hipSuccess) return fail(SL_EHIP, "tank table");
  if (hipMalloc((void **)&b->d_counters, sizeof(SlCounters) * n) != hipSuccess) return fail(SL_ENOMEM, "counters");
  b->allocs.push_back(b->d_counters);
  if (hipMemset(b->d_counters, 0, sizeof(SlCounters) * n) != hipSuccess) return fail(SL_EHIP, "counters");
  if (hipHostMalloc((void **)&b->h_counters, sizeof(SlCounters) * n, hipHostMallocDefault) != hipSuccess)
    return fail(SL_ENOMEM, "pinned counters");
  memset(b->h_counters, 0, sizeof(SlCounters) * n);
  b->maps_host.assign((size_t)BM_COUNT * (n + 1) + n, 0u);
  int r;
  if ((r = ring_init(b, b->maps, b->maps_host.size() * 4))) return fail(r, b->err);
  // emission block: SL_MAX_EMIT records per slot, then kt[2n], the map [n+1] and the emitting slots [n]
  if ((r = ring_init(b, b->emit, sizeof(SlEmitDev) * SL_MAX_EMIT * n + 4 * (4 * (size_t)n + 1)))) return fail(r, b->err);
  if (hipEventCreate(&b->ev_a) != hipSuccess || hipEventCreate(&b->ev_b) != hipSuccess) return fail(SL_EHIP, "events");
  *out = b;
  return SL_OK;
}

extern "C" int sl_batch_add(sl_batch *b, const sl_config *cfg, uint32_t *slot) {
  if (!b || !cfg) return batch_fail(b, SL_EARG, "sl_batch_add: bad arguments");
  if (cfg->device != b->device) return batch_fail(b, SL_EARG, "sl_batch_add: tank on another device");
  if (cfg->droplets) return batch_fail(b, SL_EARG, "sl_batch_add: the droplet layer runs on single simulations only");
  uint32_t t = 0;
  while (t < b->cap && b->sims[t]) t++;
  if (t == b->cap) return batch_fail(b, SL_ECAP, "sl_batch_add: every slot is taken");
  // work in flight reads the table: the new record is written with the stream idle
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "sl_batch_add: sync");
  if (hipMemset(b->d_counters + t, 0, sizeof(SlCounters)) != hipSuccess) return batch_fail(b, SL_EHIP, "counters");
  sl_sim *s = new (std::nothrow) sl_sim();
  if (!s) return batch_fail(b, SL_ENOMEM, "out of host memory");
  s->cfg = *cfg;
  s->ext_stream = b->stream;
  s->ext_tank = b->d_tanks + t;
  s->ext_counters = b->d_counters + t;
  s->ext_emits = (SlEmitDev *)b->emit.d + (size_t)t * SL_MAX_EMIT;
  int r = create_impl(s);
  if (r) {
    std::string m = s->err;
    sl_destroy(s);
    return batch_fail(b, r, m);
  }
  b->sims[t] = s;
  b->stepped[t] = 1;
  b->copy_valid[t] = 0;
  b->bucket[t] = bucket_of(s);
  b->maps_dirty = true;
  if (slot) *slot = t;
  return SL_OK;
}

extern "C" int sl_batch_remove(sl_batch *b, uint32_t slot) {
  if (!b || slot >= b->cap || !b->sims[slot]) return batch_fail(b, SL_EARG, "sl_batch_remove: no tank in that slot");
  sl_destroy(b->sims[slot]);  // waits for the stream first
  b->sims[slot] = nullptr;
  b->stepped[slot] = 0;
  b->copy_valid[slot] = 0;
  b->maps_dirty = true;
  return SL_OK;
}

extern "C" int sl_batch_set_stepped(sl_batch *b, uint32_t slot, int on) {
  if (!b || slot >= b->cap || !b->sims[slot]) return batch_fail(b, SL_EARG, "sl_batch_set_stepped: no tank in that slot");
  if ((b->stepped[slot] != 0) != (on != 0)) {
    b->stepped[slot] = on ? 1 : 0;
    b->maps_dirty = true;
  }
  return SL_OK;
}

extern "C" int sl_batch_create(const sl_config *cfgs, uint32_t n, sl_batch **out) {
  if (out) *out = nullptr;
  if (!cfgs || !out || !n) return batch_fail(nullptr, SL_EARG, "sl_batch_create: bad arguments");
  sl_batch *b = nullptr;
  int r = sl_batch_open(n, cfgs[0].device, &b);
  if (r) return r;
  for (uint32_t t = 0; t < n; t++) {
    if ((r = sl_batch_add(b, cfgs + t, nullptr))) {
      std::string m = "tank " + std::to_string(t) + ": " + b->err;
      sl_batch_destroy(b);
      return batch_fail(nullptr, r, m);
    }
  }
  *out = b;
  return SL_OK;
}

extern "C" const char *sl_batch_last_error(const sl_batch *b) { return b ? b->err.c_str() : g_err.c_str(); }
extern "C" uint32_t sl_batch_size(const sl_batch *b) { return b ? b->cap : 0; }
extern "C" sl_sim *sl_batch_sim(sl_batch *b, uint32_t slot) { return b && slot < b->cap ? b->sims[slot] : nullptr; }

// Emission into many tanks with one upload and three launches: counts[slot] records per slot
// (recs holds them concatenated in slot order). The budget of writeEmits (cap - count - 64,
// 50-sim.js:222-245) is applied on the device count, so the result does not depend on when the
// host last read the counters.
extern "C" int sl_batch_emit(sl_batch *b, const sl_emit *recs, const uint32_t *counts) {
  if (!b || !counts) return batch_fail(b, SL_EARG, "sl_batch_emit: bad arguments");
  const uint32_t n = b->cap;
  size_t nrec = 0;
  for (uint32_t t = 0; t < n; t++) {
    if (counts[t] && !b->sims[t]) return batch_fail(b, SL_EARG, "sl_batch_emit: records for an empty slot");
    nrec += counts[t];
  }
  if (nrec && !recs) return batch_fail(b, SL_EARG, "sl_batch_emit: null records");
  if (!nrec) return SL_OK;
  uint8_t *h = ring_host(b->emit);
  const size_t eb = sizeof(SlEmitDev) * SL_MAX_EMIT * n;
  SlEmitDev *he = (SlEmitDev *)h;
  uint32_t *kt = (uint32_t *)(h + eb), *emap = kt + 2 * (size_t)n, *elist = emap + n + 1;
  uint32_t nlist = 0;
  uint64_t wg = 0;
  size_t off = 0;
  for (uint32_t t = 0; t < n; t++) {
    emap[t] = (uint32_t)wg;
    kt[2 * t] = kt[2 * t + 1] = 0;
    const sl_emit *L = recs + off;
    const uint32_t m = counts[t];
    off += m;
    if (!m) continue;
    sl_sim *s = b->sims[t];
    if (s->ht.obj_count != s->obj_count) {  // the single path's emission follows writeParams
      s->ht.obj_count = s->obj_count;
      const int r = push_dyn(s);
      if (r) return batch_fail(b, r, s->err);
    }
    uint32_t k = 0, tot = 0;
    for (uint32_t i = 0; i < m && k < SL_MAX_EMIT && tot < s->cap; i++) {
      if (!L[i].count) continue;
      uint32_t cnt = std::min(L[i].count, s->cap - tot);
      SlEmitDev &e = he[(size_t)t * SL_MAX_EMIT + k];
      memcpy(e.a, L[i].a, 16); memcpy(e.b, L[i].b, 16); memcpy(e.c, L[i].c, 16); memcpy(e.d, L[i].d, 16);
      e.info[0] = tot; e.info[1] = cnt; e.info[2] = L[i].seed; e.info[3] = L[i].shape;
      tot += cnt;
      k++;
    }
    if (!tot) continue;
    kt[2 * t] = k;
    kt[2 * t + 1] = tot;
    elist[nlist++] = t;
    wg += (tot + SL_WG - 1) / SL_WG;
    // host estimates until the next counters copy
    if ((uint64_t)s->count + tot + 64 > s->cap) s->full = true;
    s->count = (uint32_t)std::min<uint64_t>((uint64_t)s->count + tot, s->cap);
    s->emitted_total += tot;
    s->p_ub = (uint32_t)std::min<uint64_t>((uint64_t)s->p_ub + tot, s->cap);
  }
  emap[n] = (uint32_t)wg;
  if (!nlist) return SL_OK;
  if (wg > 0x7fffffffull) return batch_fail(b, SL_ECAP, "sl_batch_emit: too many particles");
  int r = ring_push(b, b->emit, b->emit.bytes);
  if (r) return r;
  const uint32_t *d_kt = (const uint32_t *)(b->emit.d + eb);
  hipLaunchKernelGGL(k_batch_emit_setup, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_tanks, d_kt, (int)n);
  SlLaunch L;
  L.w.tanks = b->d_tanks;
  L.w.start = d_kt + 2 * (size_t)n;
  L.w.ntanks = (int32_t)n;
  L.w.total = (uint32_t)wg;
  L.w.list = nullptr;
  L.stream = b->stream;
  sl_k_emit(L, 64);
  SlLaunch P = L;
  P.w.start = nullptr;
  P.w.ntanks = (int32_t)nlist;
  P.w.list = d_kt + 3 * (size_t)n + 1;
  sl_k_count(P, 7, 0);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SL_OK : batch_fail(b, SL_EHIP, std::string("batch emit: ") + hipGetErrorString(e));
}

static void batch_p2g_binned(const sl_batch *b) {
  sl_k_p2g_binned(batch_launch(b, BM_BINP), batch_launch(b, BM_BINS), batch_launch(b, BM_BINT),
                  batch_launch(b, BM_BINTILE));
}

// the kernels of one frame of every stepped tank, in stream order (no host synchronisation and no
// memset, so the sequence can be captured)
static int batch_encode_frame(sl_batch *b, double dt) {
  const int n = (int)b->n_stepped;
  if (!n) return SL_OK;
  const SlLaunch P = batch_tanks(b);
  hipLaunchKernelGGL(k_batch_advance, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_tanks, list_dev(b), n, dt);
  sl_k_count(P, 5, 0);  // vmax = 0 (50-sim.js:306)
  hipLaunchKernelGGL(k_batch_clear_cells, dim3(b->total[BM_CELLS]), dim3(SL_WG), 0, b->stream,
                     batch_launch(b, BM_CELLS).w);
  for (int u = 0; u < SL_BUCKETS; u++) {
    const int v = b->map_variant[BM_P2G0 + u];
    if (v == SL_P2G_BINNED) batch_p2g_binned(b);
    else if (v >= 0) sl_k_p2g(batch_launch(b, BM_P2G0 + u), 0, v);
  }
  sl_k_normforce(batch_launch(b, BM_NODES));
  sl_k_divsetup(batch_launch(b, BM_CELLS));
  for (uint32_t j = 0; j < b->sor_bands; j++) {  // only the tanks that still iterate
    const SlLaunch Ls = batch_launch(b, BM_SOR0 + (int)j);
    for (uint32_t i = b->band_start[j]; i < b->band_start[j + 1]; i++) { sl_k_sor(Ls, 0, 2 * i); sl_k_sor(Ls, 1, 2 * i + 1); }
  }
  sl_k_project(batch_launch(b, BM_NODES));
  for (int u = 0; u < SL_BUCKETS; u++)
    if (b->map_variant[BM_G2P0 + u] >= 0) sl_k_g2p(batch_launch(b, BM_G2P0 + u), b->map_variant[BM_G2P0 + u]);
  const SlLaunch Lc = batch_launch(b, BM_COMPACT);
  sl_k_compact_count(Lc);
  SlLaunch Lscan = Lc;
  Lscan.w.list = list_dev(b);
  sl_k_compact_scan(Lscan, n);
  sl_k_compact_move(Lc);
  sl_k_ww(batch_launch(b, BM_WW), (float)dt);  // dt = 0: each tank's own (its substep dt)
  sl_k_count(P, 8, 0);  // count = newCount and the whitewater commit (WW reads neither count)
  if (b->any_ripple)  // the mask stages return for tanks whose mask is off
    sl_k_ripple_frame_ex(batch_launch(b, BM_ROWS).w, batch_launch(b, BM_COLS).w, batch_launch(b, BM_ELEMS).w,
                         batch_launch(b, BM_TEX).w, b->stream, 1, b->max_M);
  if (b->any_toph) sl_k_toph(batch_launch(b, BM_TOPP).w, batch_launch(b, BM_TOPT).w, b->stream);
  hipError_t e = hipMemcpyAsync(b->h_counters, b->d_counters, sizeof(SlCounters) * b->cap, hipMemcpyDeviceToHost,
                                b->stream);
  if (e == hipSuccess) e = hipGetLastError();
  return e == hipSuccess ? SL_OK : batch_fail(b, SL_EHIP, std::string("batch frame: ") + hipGetErrorString(e));
}

// GEOM for tanks whose objects changed (outside any captured graph), and the work maps after a
// membership change
static int batch_prepare(sl_batch *b) {
  for (uint32_t t = 0; t < b->cap; t++) {
    sl_sim *s = b->sims[t];
    if (!s || !b->stepped[t] || !(s->geom_dirty || s->mask_dirty)) continue;
    s->ht.obj_count = s->obj_count;  // this frame's parameters, as the single path writes before GEOM
    int r = push_dyn(s);
    if (r) return batch_fail(b, r, s->err);
    if (s->geom_dirty) sl_k_geom(launch_of(s, s->nodes));
    s->geom_dirty = false;
    encode_objmask(s);
  }
  if (b->maps_dirty) {
    int r = batch_rebuild(b);
    if (r) return r;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SL_OK : batch_fail(b, SL_EHIP, hipGetErrorString(e));
}

// Re-orderings of batch tanks as one stable radix sort per group: each tank's keys carry its rank
// in the group above the cell-index bits, so the tanks' particles come out in group order, and each
// tank's in exactly the order of its own sort (stable, same cell keys, the same sentinel outside the
// grid). One sort instead of one per tank: a radix sort is a dozen or so launches whatever its
// length. A group takes tanks while the key fits 32 bits and the group stays below 2^31 particles;
// a single tank always fits (fewer than 2^30 cells and particles). The tanks of a batch hold no sort
// buffers of their own: the batch's serve them all (sized for every stepped tank at once, so that
// no later frame allocates or waits until the membership grows). `restart` restarts the tanks'
// cadence (the periodic re-orderings; an explicit one leaves it).
static int batch_sort_group(sl_batch *b, const uint32_t *ts, size_t k, uint64_t ntot, int cbits, bool restart) {
  const int rbits = k > 1 ? 32 - __builtin_clz((uint32_t)k - 1u) : 0;
  const int end_bit = cbits + rbits;
  if (ntot > b->ro_cap) {
    uint64_t all = 0;
    for (uint32_t t = 0; t < b->cap; t++)
      if (b->sims[t] && b->stepped[t]) all += b->sims[t]->cap;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "re-ordering: sync");
    (void)hipFree(b->ro_keys); (void)hipFree(b->ro_vals); (void)hipFree(b->ro_temp);
    b->ro_keys = b->ro_vals = nullptr; b->ro_temp = nullptr; b->ro_cap = 0; b->ro_temp_bytes = 0;
    const uint64_t cap = std::min<uint64_t>(std::max(ntot, all), (1ull << 31) - 1);
    const size_t tb = sl_sort_temp_bytes((uint32_t)cap, 32);
    if (hipMalloc((void **)&b->ro_keys, cap * 8) != hipSuccess || hipMalloc((void **)&b->ro_vals, cap * 8) != hipSuccess ||
        hipMalloc(&b->ro_temp, tb ? tb : 256) != hipSuccess)
      return batch_fail(b, SL_ENOMEM, "re-ordering: no device memory for the batch sort");
    b->ro_cap = cap;
    b->ro_temp_bytes = tb ? tb : 256;
  }
  uint32_t off = 0;
  for (size_t r = 0; r < k; r++) {
    sl_sim *s = b->sims[ts[r]];
    sl_k_sort_keys(launch_of(s, 1), s->p_ub, b->ro_keys, b->ro_vals, (uint32_t)r << cbits, off);
    off += s->p_ub;
  }
  // (empty tanks sort nothing but still swap their buffers, as a tank alone does)
  int e = ntot ? sl_sort_pairs(b->ro_temp, b->ro_temp_bytes, b->ro_keys, b->ro_vals, (uint32_t)ntot, end_bit, b->stream) : 0;
  if (e) return batch_fail(b, SL_EHIP, std::string("re-ordering sort: ") + hipGetErrorString((hipError_t)e));
  off = 0;
  for (size_t r = 0; r < k; r++) {
    sl_sim *s = b->sims[ts[r]];
    sl_k_sort_gather(launch_of(s, 1), s->p_ub, b->ro_vals + ntot, off);
    sl_k_count(launch_of(s, 1), 6, 0);  // the gathered buffer becomes current
    s->cur ^= 1;
    if (restart) s->reorder_clock = 0;
    off += s->p_ub;
  }
  hipError_t le = hipGetLastError();
  return le == hipSuccess ? SL_OK : batch_fail(b, SL_EHIP, std::string("re-ordering: ") + hipGetErrorString(le));
}
static int batch_reorder(sl_batch *b, const std::vector<uint32_t> &due, bool restart) {
  size_t i = 0;
  while (i < due.size()) {
    size_t j = i;
    uint64_t n = 0;
    int cbits = 1;
    while (j < due.size()) {
      const sl_sim *s = b->sims[due[j]];
      const int cb = std::max(cbits, s->sort_end_bit);
      const size_t k = j - i + 1;
      const int rb = k > 1 ? 32 - __builtin_clz((uint32_t)k - 1u) : 0;
      if (j > i && (cb + rb > 32 || n + s->p_ub >= (1ull << 31))) break;
      cbits = cb;
      n += s->p_ub;
      j++;
    }
    int r = batch_sort_group(b, due.data() + i, j - i, n, cbits, restart);
    if (r) return r;
    i = j;
  }
  return SL_OK;
}

// host mirror of one frame of every stepped tank, plus each tank's re-ordering cadence
static int batch_host_frame(sl_batch *b, double dt) {
  std::vector<uint32_t> due;
  for (uint32_t t = 0; t < b->cap; t++) {
    sl_sim *s = b->sims[t];
    if (s) b->copy_valid[t] = 1;  // the frame's counters copy covers every tank in the table
    if (!s || !b->stepped[t]) continue;
    const double d = dt > 0 ? dt : s->batch_dt;
    s->frame++;
    write_params(s, d);
    s->time += d;
    s->ww_cur ^= 1;
    s->frames++;
    s->substeps_last = 1;
    s->emitted_at_copy = s->emitted_total;
    if (s->cfg.reorder_every && ++s->reorder_clock >= s->cfg.reorder_every) due.push_back(t);
  }
  return batch_reorder(b, due, true);
}

// dt > 0: every stepped tank advances by dt; dt = 0: each by its own frame length (sl_batch_set_dt)
static int check_dt(sl_batch *b, double dt, const char *who) {
  if (dt > 0) return SL_OK;
  if (dt != 0) return batch_fail(b, SL_EARG, std::string(who) + ": dt must be positive, or 0 for per-tank dt");
  for (uint32_t t = 0; t < b->cap; t++)
    if (live(b, t) && !(b->sims[t]->batch_dt > 0))
      return batch_fail(b, SL_EARG, std::string(who) + ": a stepped tank has no frame length (sl_batch_set_dt)");
  return SL_OK;
}

extern "C" int sl_batch_set_dt(sl_batch *b, uint32_t slot, double dt) {
  if (!b || slot >= b->cap || !b->sims[slot] || !(dt > 0)) return batch_fail(b, SL_EARG, "sl_batch_set_dt: bad arguments");
  b->sims[slot]->batch_dt = dt;
  int r = push_dyn(b->sims[slot]);
  return r ? batch_fail(b, r, b->sims[slot]->err) : SL_OK;
}

extern "C" int sl_batch_step(sl_batch *b, double dt) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  int rc = check_dt(b, dt, "sl_batch_step");
  if (rc) return rc;
  int r = batch_prepare(b);
  if (r) return r;
  if (!b->n_stepped) return SL_OK;
  hipEvent_t t1 = nullptr;
  if (b->timing) {
    while (b->tev.size() < b->tev_used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return batch_fail(b, SL_EHIP, "timing event");
      b->tev.push_back(e);
    }
    (void)hipEventRecord(b->tev[b->tev_used], b->stream);
    t1 = b->tev[b->tev_used + 1];
    b->tev_used += 2;
  }
  if (b->exec && b->graph_dt == dt) {
    hipError_t e = hipGraphLaunch(b->exec, b->stream);
    if (e != hipSuccess) return batch_fail(b, SL_EHIP, std::string("graph launch: ") + hipGetErrorString(e));
  } else if ((r = batch_encode_frame(b, dt))) {
    return r;
  }
  b->steps++;
  const int hr = batch_host_frame(b, dt);  // enqueues the frame's re-orderings
  if (t1) (void)hipEventRecord(t1, b->stream);
  return hr;
}

// Records one frame of every stepped tank into a hipGraph; sl_batch_step with the same dt replays it
// (one graph launch per frame instead of 2*iters + ~25 kernel launches) until the membership or the
// variant choice changes.
extern "C" int sl_batch_capture(sl_batch *b, double dt) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  int rc = check_dt(b, dt, "sl_batch_capture");
  if (rc) return rc;
  int r = batch_prepare(b);
  if (r) return r;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "sync before capture");
  drop_graph(b);
  if (!b->n_stepped) return SL_OK;
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return batch_fail(b, SL_EHIP, "begin capture failed");
  r = batch_encode_frame(b, dt);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (r || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return r ? r : batch_fail(b, SL_EHIP, "end capture failed");
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->exec = nullptr;
    drop_graph(b);
    return batch_fail(b, SL_EHIP, "graph instantiate failed");
  }
  b->graph_dt = dt;
  return SL_OK;
}
extern "C" int sl_batch_captured(const sl_batch *b) { return b && b->exec ? 1 : 0; }

extern "C" int sl_batch_sync(sl_batch *b) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) {
    for (sl_sim *s : b->sims) if (s) s->gpu_errors++;
    return batch_fail(b, SL_EHIP, std::string("batch sync: ") + hipGetErrorString(e));
  }
  // the counters of the last frame are in pinned memory: refresh the host state of the tanks
  for (uint32_t t = 0; t < b->cap; t++) {
    sl_sim *s = b->sims[t];
    if (!s || !b->copy_valid[t]) continue;
    *s->h_counters = b->h_counters[t];
    apply_stats(s);
    if (s->count + 64 >= s->cap) s->full = true;
  }
  return SL_OK;
}

extern "C" int sl_batch_reorder(sl_batch *b) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  std::vector<uint32_t> all;
  for (uint32_t t = 0; t < b->cap; t++)
    if (live(b, t)) all.push_back(t);
  return batch_reorder(b, all, false);
}

extern "C" int sl_batch_timing_enable(sl_batch *b, int enable) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "sync");
  b->timing = enable != 0;
  b->tev_used = 0;
  return SL_OK;
}

// device time of each step since timing was enabled or last read (oldest first), then starts over
extern "C" int sl_batch_timing_read(sl_batch *b, double *ms, uint32_t max, uint32_t *n) {
  if (!b || (max && !ms)) return batch_fail(b, SL_EARG, "sl_batch_timing_read: bad arguments");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, SL_EHIP, "sync");
  uint32_t k = 0;
  for (size_t i = 0; i + 1 < b->tev_used && k < max; i += 2) {
    float t = 0;
    (void)hipEventElapsedTime(&t, b->tev[i], b->tev[i + 1]);
    ms[k++] = t;
  }
  if (n) *n = k;
  b->tev_used = 0;
  return SL_OK;
}

// Per size bucket, the fastest P2G and G2P variant on the live state of that bucket's stepped tanks
// (all variants give bit-identical results). Particles, counters, cell counts and ripple impulses
// are snapshotted and restored, as sl_autotune does for one simulation.
extern "C" int sl_batch_autotune(sl_batch *b, uint32_t reps, uint32_t *choice) {
  if (!b) return batch_fail(nullptr, SL_EARG, "null batch");
  if (!reps) reps = 2;
  int r = batch_prepare(b);
  if (r || (r = sl_batch_sync(b))) return r;
  struct Saved { void *live, *copy; size_t bytes; };
  std::vector<Saved> items;
  auto release = [&]() { for (auto &it : items) (void)hipFree(it.copy); };
  for (uint32_t t = 0; t < b->cap; t++) {
    if (!live(b, t)) continue;
    sl_sim *s = b->sims[t];
    if ((r = dev_cur(s))) return batch_fail(b, r, s->err);
    const SlTank &T = s->ht;
    const int c = s->cur;
    const size_t P = s->cap;
    void *src[11] = {T.px[c], T.py[c], T.pz[c], T.vx[c], T.vy[c], T.vz[c], T.pf[c], T.C, T.cellCount, T.rsrc,
                     T.vslot};
    size_t by[11] = {P * 4, P * 4, P * 4, P * 4, P * 4, P * 4, P, sizeof(SlCounters), (size_t)s->cells * 4,
                     s->R ? (size_t)(s->R + 1) * (s->R + 1) * 4 : 0, (size_t)SL_VSLOTS * SL_VSLOT_STRIDE * 4};
    for (int i = 0; i < 11; i++) {
      if (!by[i]) continue;
      void *p = nullptr;
      if (hipMalloc(&p, by[i]) != hipSuccess) { release(); return batch_fail(b, SL_ENOMEM, "batch autotune: snapshot"); }
      items.push_back({src[i], p, by[i]});
      (void)hipMemcpyAsync(p, src[i], by[i], hipMemcpyDeviceToDevice, b->stream);
    }
  }
  auto restore = [&]() {
    for (auto &it : items) (void)hipMemcpyAsync(it.live, it.copy, it.bytes, hipMemcpyDeviceToDevice, b->stream);
  };
  auto clear_acc = [&]() {
    for (uint32_t t = 0; t < b->cap; t++)
      if (live(b, t)) sl_k_fill_u32(b->stream, (uint32_t *)b->sims[t]->ht.acc, 0, (size_t)b->sims[t]->nodes * 6);
  };
  auto timed = [&](auto fn) {
    (void)hipEventRecord(b->ev_a, b->stream);
    fn();
    (void)hipEventRecord(b->ev_b, b->stream);
    (void)hipEventSynchronize(b->ev_b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, b->ev_a, b->ev_b);
    return ms;
  };
  for (int k = 0; k < SL_BUCKETS && !r; k++) {
    bool has = false;
    for (uint32_t t = 0; t < b->cap; t++) has |= live(b, t) && b->bucket[t] == k;
    if (!has) continue;
    auto in_k = [&](uint32_t t) { return b->bucket[t] == k; };
    float best = 1e30f;
    // the largest bucket holds the tanks above 256 Ki cells: binned only (p2g_first_trial)
    for (int v = k == SL_BUCKETS - 1 ? SL_P2G_BINNED : 0; v <= SL_P2G_BINNED; v++) {
      if (v == SL_P2G_BINNED) r = put_bin_maps(b, in_k);
      else r = put_map(b, BM_SCRATCH, [](const sl_sim *s) { return (uint64_t)s->cap; }, sl_p2g_chunk(v), in_k);
      if (r || (r = maps_upload(b))) break;
      float tot = 0;
      for (uint32_t rep = 0; rep <= reps; rep++) {
        clear_acc();
        hipLaunchKernelGGL(k_batch_clear_cells, dim3(b->total[BM_CELLS]), dim3(SL_WG), 0, b->stream,
                           batch_launch(b, BM_CELLS).w);
        float ms = timed([&] {
          if (v == SL_P2G_BINNED) batch_p2g_binned(b);
          else sl_k_p2g(batch_launch(b, BM_SCRATCH), 0, v);
        });
        if (rep) tot += ms;
      }
      if (tot < best) { best = tot; b->p2g_variant[k] = v; }
    }
    if (r) break;
    if ((r = put_map(b, BM_SCRATCH, [](const sl_sim *s) { return (uint64_t)s->cap; }, SL_WG, in_k)) ||
        (r = maps_upload(b)))
      break;
    best = 1e30f;
    for (int v = 0; v < 4; v++) {
      float tot = 0;
      for (uint32_t rep = 0; rep <= reps; rep++) {
        restore();
        float ms = timed([&] { sl_k_g2p(batch_launch(b, BM_SCRATCH), v); });
        if (rep) tot += ms;
      }
      if (tot < best) { best = tot; b->g2p_variant[k] = v; }
    }
  }
  clear_acc();  // the next P2G adds into cleared accumulators
  restore();
  hipError_t e = hipStreamSynchronize(b->stream);
  release();
  if (r) return r;
  if (e != hipSuccess || (e = hipGetLastError()) != hipSuccess) return batch_fail(b, SL_EHIP, hipGetErrorString(e));
  if ((r = batch_rebuild(b))) return r;
  if (choice)
    for (int k = 0; k < SL_BUCKETS; k++) { choice[2 * k] = (uint32_t)b->p2g_variant[k]; choice[2 * k + 1] = (uint32_t)b->g2p_variant[k]; }
  return SL_OK;
}

// sets every bucket's variants (values as sl_batch_autotune reports them)
extern "C" int sl_batch_set_variants(sl_batch *b, const uint32_t *choice) {
  if (!b || !choice) return batch_fail(b, SL_EARG, "sl_batch_set_variants: bad arguments");
  for (int k = 0; k < SL_BUCKETS; k++)
    if (choice[2 * k] > SL_P2G_BINNED || choice[2 * k + 1] > 3) return batch_fail(b, SL_EARG, "variant out of range");
  for (int k = 0; k < SL_BUCKETS; k++) { b->p2g_variant[k] = (int)choice[2 * k]; b->g2p_variant[k] = (int)choice[2 * k + 1]; }
  b->maps_dirty = true;
  return SL_OK;
}

extern "C" int sl_version(void) { return SL_ABI_VERSION; }
extern "C" const char *sl_build_info(void) {
  return "suiri gfx950 HIP (" __DATE__ ") fp32, -ffp-contract=off";
}
