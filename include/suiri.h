/*
 * suiri — MI355X-native liquid step (FLIP on a MAC grid with cut-cell face fractions, red-black SOR,
 * projection, advection, capillary-gravity ripple layer, whitewater), C ABI.
 *
 * Every entry point below replaces one method of the reference's JavaScript simulation driver
 * (`class FluidSim`, src/js/50-sim.js, and `class RippleLayer`, src/js/64-ripple.js). The reference has
 * no FFI; its callers are the main loop (src/js/90-main.js:5-30, :72-104), the UI
 * (src/js/80-ui.js:27-93, :334-380) and the automation API (src/js/90-main.js:145-205). A Node N-API
 * addon (hahenn-_amd/csrc/napi_addon.cpp) binds these functions for the JavaScript host.
 *
 * Conventions
 *   - Status: every int-returning function returns SL_OK (0) or a negative SL_E* code; the message is
 *     available from sl_last_error(sim) (sl_last_error(NULL) for failures before a sim exists).
 *     The reference never throws from step() and counts GPU errors (src/js/40-gpu.js:22-26); here
 *     device errors are sticky per sim and reported by the next call and in sl_stats.gpu_errors.
 *   - Ownership: the sim owns all device memory. Host inputs are copied at call time (like
 *     GPUQueue.writeBuffer); readbacks copy into caller-owned memory.
 *   - Threading: one sim = one HIP stream. Calls on one sim must not overlap. Work is enqueued and
 *     the call returns (like queue.submit); sl_sync / sl_stats / sl_readback wait for the stream.
 *   - Layouts on upload/readback are the reference's buffer layouts (array-of-structs particles of
 *     2 x vec4f, vec4f per MAC node, x-fastest indices); the device layout is structure-of-arrays.
 */
#ifndef SUIRI_H
#define SUIRI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SL_ABI_VERSION 2

#define SL_OK 0
#define SL_EARG -1     /* bad argument */
#define SL_EHIP -2     /* HIP runtime / device error (sticky) */
#define SL_ENOMEM -3   /* device or host allocation failed */
#define SL_ESTATE -4   /* call not valid in the current state */
#define SL_EBUDGET -5  /* admission refused: HBM budget */
#define SL_ECAP -6     /* size above a fixed maximum */

#define SL_MAX_OBJ 64   /* obstacles and drains per tank (MAX_OBJ, 50-sim.js:9) */
#define SL_MAX_EMIT 48  /* emission records per substep (MAX_EMIT, 50-sim.js:10) */

typedef struct sl_sim sl_sim;

/* Quality preset + physics (QUALITY, src/js/50-sim.js:4-8; constants of writeParams, :151-162;
   ripple constants, src/js/64-ripple.js:225, :268, :88). sl_config_default fills the reference values. */
typedef struct sl_config {
  double tank_w, tank_h, tank_d; /* tank size [m] (scene.tank, src/js/10-scene.js:246) */
  double h;                      /* requested cell size [m] */
  uint32_t cap, ww_cap;          /* particle / whitewater capacity */
  uint32_t iters;                /* SOR iterations per substep */
  uint32_t ripple_R;             /* ripple physical cells, M = 2R (power of two); 0 = ripple off */
  double rc_factor;              /* collision radius in cells (0.32) */
  float gravity, flip, ppc, drift, omega, friction, ww_rate;
  float r_sigma_rho, r_nu, r_film, r_hp_k, r_rho_kick;
  uint32_t wmap;                 /* top-height map size for the ripple mask (256) */
  uint32_t whitewater;           /* settings.whitewater: 0 sets the whitewater rate to 0 (50-sim.js:159) */
  int32_t device;                /* HIP device ordinal */
  uint32_t reorder_every;        /* spatial particle re-ordering period in frames (0 = never) */
  uint32_t reorder_phase;        /* frames already counted towards the first re-ordering (staggers
                                    the re-orderings of many tanks) */
  uint32_t top_height;           /* 1: each frame's end draws the top-surface height map from the
                                    particles (63-wgsl-water.js:5-55, what the renderer draws for the
                                    next frame's ripple mask); 0: the map stays as uploaded (zeros, as
                                    in the reference's headless runs) */
  uint32_t seed;                 /* seed of the stream of emission-record seeds (sl_scene_source_emits,
                                    sl_scene_block_emits; the reference draws Math.random(), 50-sim.js:234) */
  /* splash of a drop (splashCrown, 31-wgsl-sim2.js:119-148): liquid density, surface tension and
     viscosity of K = We * Oh^-0.4 (998, 0.0728, 0.001 at :121-123); jets when K > splash_k_jet (700),
     a crown ring of round(4 (K / splash_k_crown - 1)) droplets (2100, :125-126) */
  double splash_rho, splash_sigma, splash_mu;
  float splash_k_jet, splash_k_crown;
  /* ripple depth D [m] of the dispersion relation: > 0 a fixed depth; 0 the mean depth of the
     particle count, as the reference (50-sim.js:332) */
  double ripple_depth;
  /* sub-grid droplets deposited on the ripple layer each frame (no counterpart in the reference):
     `droplets` per frame, diameter lognormal with median drop_d50 [m] and drop_sigma_ln (sigma of
     ln D), speed uniform in [drop_vmin, drop_vmax] [m/s], impact points uniform over the tank
     footprint at the water level; each deposits the impulse 2 m v (mass m from splash_rho) and
     splashes as splashCrown. 0 = none. Not for the tanks of a batch. */
  uint32_t droplets, drop_seed;
  float drop_d50, drop_sigma_ln, drop_vmin, drop_vmax;
  uint32_t reserved[2];
} sl_config;

/* One emission record, as produced by sourceEmits / burstEmit / blockEmits (src/js/55-emit.js:13-106)
   and packed by writeEmits (src/js/50-sim.js:222-245). `seed` replaces Math.random() (50-sim.js:234). */
typedef struct sl_emit {
  float a[4], b[4], c[4], d[4];
  uint32_t shape;   /* 0 disc outlet, 1 sphere, 2 jittered lattice, 3 rain box */
  uint32_t count;   /* requested particles */
  uint32_t seed;
  uint32_t pad;
} sl_emit;

/* FluidSim grid sizing (src/js/50-sim.js:33-45) */
typedef struct sl_grid_t {
  int32_t nx, ny, nz;
  uint32_t nodes, cells, cap, ww_cap, R, M, iters;
  double h, origin[3], rim, vp, radius;
} sl_grid_t;

/* suiriLab.stats() (src/js/90-main.js:196) + counters (31-wgsl-sim2.js:5) */
typedef struct sl_stats_t {
  uint32_t count, ww_count, impacts, frames;
  uint32_t full, gpu_errors, substeps_last, reserved;
  double vmax, time, liters;
  int32_t grid[3];
} sl_stats_t;

/* Buffers for sl_upload / sl_readback (reference layouts). */
enum sl_field {
  SL_F_PARTICLES = 0, /* cap x {pos.xyz, alive, vel.xyz, mark} f32 (parts[cur], 50-sim.js:65) */
  SL_F_WW = 1,        /* ww_cap x {pos.xyz, life, vel.xyz, kind} f32 (ww[wwCur], 50-sim.js:66) */
  SL_F_COUNTERS = 2,  /* 8 x u32 {count,newCount,wwCount,wwNew,vmax,impacts,0,0} */
  SL_F_ACC = 3,       /* N x 6 i32 (mom,wt per axis) */
  SL_F_CELLCOUNT = 4, /* C x u32 */
  SL_F_VEL = 5,       /* N x vec4f */
  SL_F_VELOLD = 6,    /* N x vec4f, .w = participation mask */
  SL_F_FRAC = 7,      /* N x vec4f face open fractions */
  SL_F_SVEL = 8,      /* N x vec4f solid velocity at faces */
  SL_F_CELLTYPE = 9,  /* C x u32 (readback only; = cellCount > 0) */
  SL_F_DIVG = 10,     /* C x f32 */
  SL_F_PRES = 11,     /* C x f32 */
  SL_F_OBJS = 12,     /* 64 x 28 f32 (readback only) */
  SL_F_RSRC = 13,     /* (R+1)^2 i32 ripple impulses (1e9 fixed point) */
  SL_F_RSTATE = 14,   /* M^2 x vec2f ripple spectral state */
  SL_F_RHAT = 15,     /* M^2 x vec2f forward transform of the source (readback only) */
  SL_F_RFIELD = 16,   /* M^2 x vec2f inverse transform of the state (height, velocity) (readback only) */
  SL_F_RMASK = 17,    /* (R+1)^2 f32 absorption mask (readback only) */
  SL_F_RTEX = 18,     /* (R+1)^2 x 4 f32: height, slope x, slope z, 0 (readback only) */
  SL_F_TOPH = 19,     /* wmap^2 f32 top-surface height map, -1 dry (input of the ripple mask) */
  SL_F_RTEX16 = 20,   /* (R+1)^2 x 4 f16 rgba16float texture of the reference (64-ripple.js:160) */
  SL_F_ARGS = 21,     /* 16 u32 dispatch / draw-indirect arguments written by every COUNT
                         (31-wgsl-sim2.js:343-348): [ceil(n/256),1,1, ceil(w/256),1,1, 6,n,0,0, 6,w,0,0]
                         for n live particles and w whitewater particles (readback only) */
  SL_F_COUNT_ = 22
};

/* Single stages, for stage-level parity checks. One call = one dispatch of the reference kernel(s). */
enum sl_stage {
  SL_S_GEOM = 0,       /* 30-wgsl-sim.js:118-157 */
  SL_S_P2G = 1,        /* clear acc + cellCount, then 30-wgsl-sim.js:159-200 */
  SL_S_NORMFORCE = 2,  /* 30-wgsl-sim.js:203-238 (NORMALIZE then FORCES) */
  SL_S_DIVERGENCE = 3, /* 30-wgsl-sim.js:240-264 */
  SL_S_PRESSURE = 4,   /* one colour pass (arg = parity), 30-wgsl-sim.js:267-297 */
  SL_S_PROJECT = 5,    /* 30-wgsl-sim.js:299-341 */
  SL_S_G2P = 6,        /* 31-wgsl-sim2.js:149-251 */
  SL_S_COMPACT = 7,    /* COMPACT + COUNT(1), 31-wgsl-sim2.js:317-339 */
  SL_S_WW = 8,         /* WW + COUNT(3), with the call's dt as the frame length */
  SL_S_RIPPLE = 9,     /* RippleLayer.encodeFrame, 64-ripple.js:222-253 */
  SL_S_SOR = 10,       /* the full SOR loop (iters x red, black) */
  SL_S_DROPLETS = 11,  /* the frame's sub-grid droplets (cfg.droplets) with the frame's ripple uniform */
  SL_S_COUNT_ = 12
};

int sl_config_default(sl_config *cfg, const char *quality); /* "low" | "standard" | "high" */

/* new FluidSim(tank, quality) — src/js/50-sim.js:29-89 (sizing, buffers, COUNT(4)) */
int sl_create(const sl_config *cfg, sl_sim **out);
/* FluidSim.destroy() — src/js/50-sim.js:143-149 */
void sl_destroy(sl_sim *sim);
const char *sl_last_error(const sl_sim *sim);
int sl_grid_info(const sl_sim *sim, sl_grid_t *out);

/* The object upload of FluidSim.setObjects(items, dtMove) — src/js/50-sim.js:194-210: `rows` are n
   packed 28-float Obj records (rows of R^T, pos+type, dims, lin+bound, ang+flag). The host computes the
   packing and the moving-object velocities (50-sim.js:171-201); an unchanged set is not re-uploaded
   and does not mark the geometry dirty. *changed (optional) receives 1 when it was. */
int sl_set_objects(sl_sim *sim, const float *rows, uint32_t n, int *changed);

/* ---- scene items (the scene model of src/js/10-scene.js:17-87 and its file schema) ----
   One implementation of the host steps that turn a scene into device records, shared by every host
   (Node addon, Python, the serving loop): object packing (FluidSim.setObjects, 50-sim.js:165-211) and
   the emitters (sourceEmits / burstEmit / blockEmits / blockLiters, 55-emit.js:13-113). */
enum { SL_ITEM_OBSTACLE = 0, SL_ITEM_DRAIN = 1, SL_ITEM_SOURCE = 2, SL_ITEM_BLOCK = 3 };
/* obstacle shapes (10-scene.js:5-15) */
enum { SL_SHAPE_BOX = 0, SL_SHAPE_PLATE, SL_SHAPE_SPHERE, SL_SHAPE_CYLINDER, SL_SHAPE_CONE, SL_SHAPE_TORUS,
       SL_SHAPE_CUP, SL_SHAPE_CONTAINER, SL_SHAPE_GUTTER, SL_SHAPE_COUNT_ };
/* source modes (10-scene.js:60-68) */
enum { SL_MODE_STREAM = 0, SL_MODE_DROPS, SL_MODE_SHOWER, SL_MODE_BURST, SL_MODE_RAIN, SL_MODE_COUNT_ };
typedef struct sl_item {
  uint32_t id;       /* item id (object velocities and source progress are kept per id) */
  uint32_t kind;     /* SL_ITEM_* */
  uint32_t shape;    /* obstacle: SL_SHAPE_*; source: SL_MODE_* */
  uint32_t enabled;  /* sources: 0 = off */
  double pos[3];     /* [m], y up */
  double rot[3];     /* Euler angles [deg], applied x, then y, then z (obstacles, drains) */
  /* obstacle dims by shape: box (w, h, d), plate (w, d, t), sphere (r), cylinder and cone (r, h),
     torus (R, r), cup (r, h, t), container (w, h, d, t), gutter (len, r, t); block (w, h, d);
     drain (r) */
  double dims[4];
  /* sources: pitch / yaw [deg], flow [L/min], speed [m/s], drop diameter [mm], drops per second,
     shower spread [deg], burst volume [L], burst period [s] (0: once), rain rate [mm/h], rain square
     side [m], total litres (0: unlimited) */
  double pitch, yaw, flow, speed, drop_d, drop_rate, spread, burst, burst_every, rain_rate, rain_size, total;
} sl_item;

/* Host-side scene state (source progress, previous object poses, seed stream) for a tank of this
   configuration; no device is touched. A simulation keeps its own (sl_sim_scene): sl_save_water /
   sl_restore_water save and restore it with the water, sl_clear_water resets the sources. */
typedef struct sl_scene sl_scene;
int sl_scene_create(const sl_config *cfg, sl_scene **out);
void sl_scene_destroy(sl_scene *sc);
const char *sl_scene_last_error(const sl_scene *sc);
sl_scene *sl_sim_scene(sl_sim *sim);
/* The Obj records of the obstacles and drains among `items` (at most SL_MAX_OBJ, in item order) into
   rows[SL_MAX_OBJ * 28]: R^T rows, position + shape type, shape dimensions, linear velocity + bound
   radius, angular velocity + flag. Velocities come from each id's pose at the previous call over
   dt_move (none when dt_move <= 0 or the id is new), clamped to 4 m/s and 40 rad/s; *moving = 1 when
   an object moved (50-sim.js:165-201). The poses are replaced by this call's. */
int sl_scene_pack_objects(sl_scene *sc, const sl_item *items, uint32_t n, double dt_move, float *rows,
                          uint32_t *n_rows, int *moving);
/* FluidSim.setObjects(items, dtMove): sl_scene_pack_objects with the simulation's scene, then
   sl_set_objects */
int sl_set_items(sl_sim *sim, const sl_item *items, uint32_t n, double dt_move, int *moving);
/* sourceEmits(sim, dt, t) — 55-emit.js:13-69: the emission records of the enabled sources among
   `items` for one substep of length dt (stream and shower discs, drops, bursts, rain), at most `max`;
   each record takes the next seed of the scene's stream */
int sl_scene_source_emits(sl_scene *sc, const sl_item *items, uint32_t n, double dt, sl_emit *out, uint32_t max,
                          uint32_t *n_out);
/* blockEmits(sim, items, fill) — 55-emit.js:83-106: one jittered half-cell lattice per block,
   clipped to the tank; `fill` pours skip cells that already hold water */
int sl_scene_block_emits(sl_scene *sc, const sl_item *items, uint32_t n, int fill, sl_emit *out, uint32_t max,
                         uint32_t *n_out);
/* blockLiters(b) — 55-emit.js:108-113: the litres of a block that lie inside a tank of w x h x d */
double sl_block_liters(const sl_item *block, double tank_w, double tank_h, double tank_d);
/* FluidSim.step(frameDt, sourceEmits): a frame whose substeps emit the sources among `items` */
int sl_step_items(sl_sim *sim, double frame_dt, const sl_item *items, uint32_t n);

/* FluidSim.emitNow(list, fill) — src/js/50-sim.js:248-265 (writeEmits budget cap-count-64, `full`) */
int sl_emit_now(sl_sim *sim, const sl_emit *list, uint32_t n, int fill, uint32_t *emitted);

/* FluidSim.step(frameDt, emitFn) — src/js/50-sim.js:300-336, split so that a host language can call
   its own emitter between substeps:
     sl_step_begin: substep count n = clamp(ceil(frameDt / (1.4 h / max(vmax, 0.4))), 1, 6) (or the
                    forced count of sl_set_substeps) and counters.vmax = 0;
     sl_substep:    frame++, params(dt), writeEmits(list) + EMIT + COUNT(2), COUNT(0), encodeSubstep;
     sl_step_end:   COMPACT, COUNT(1), WW, COUNT(3), ripple frame, stats readback. */
int sl_step_begin(sl_sim *sim, double frame_dt, uint32_t *nsub, double *dt);
int sl_substep(sl_sim *sim, double dt, const sl_emit *list, uint32_t n, uint32_t *emitted);
int sl_step_end(sl_sim *sim, double frame_dt);
/* whole step with a C emitter callback (returns the number of records written to `out`) */
typedef uint32_t (*sl_emit_fn)(void *user, sl_sim *sim, double dt, double t, sl_emit *out, uint32_t max);
int sl_step(sl_sim *sim, double frame_dt, sl_emit_fn fn, void *user);
/* 0 = adaptive substeps as the reference; n > 0 pins the substep count (fixed-dt configurations) */
int sl_set_substeps(sl_sim *sim, uint32_t n);

/* Spatial re-ordering of the particles (stable sort by the Morton code of their 4x4x4-cell brick).
   No counterpart in the reference: its compaction leaves the particle order arbitrary
   (src/js/31-wgsl-sim2.js:317-328), so any permutation is a valid state. Also run automatically every
   cfg.reorder_every frames at the end of sl_step_end. */
int sl_reorder(sl_sim *sim);

/* FluidSim.clearWater() — src/js/50-sim.js:365-373 */
int sl_clear_water(sl_sim *sim);
/* FluidSim.saveWater / hasWaterCopy / restoreWater — src/js/50-sim.js:111-141 */
int sl_save_water(sl_sim *sim, uint32_t *id);
int sl_has_water_copy(const sl_sim *sim, uint32_t id);
int sl_restore_water(sl_sim *sim, uint32_t id);

/* FluidSim.pollStats() — src/js/50-sim.js:346-363: non-blocking; *updated = 1 when a readback landed */
int sl_poll_stats(sl_sim *sim, int *updated);
/* suiriLab.stats() after onSubmittedWorkDone — src/js/90-main.js:159-165, :196 (waits, then polls) */
int sl_stats(sl_sim *sim, sl_stats_t *out);
/* device.queue.onSubmittedWorkDone() */
int sl_sync(sl_sim *sim);
/* waits for the work submitted to the simulation so far without touching its state, so another
   thread may wait while the owning thread keeps calling (the Node addon's statsAsync waits on a
   worker thread); returns SL_OK or SL_EHIP */
int sl_wait_idle(const sl_sim *sim);

/* Ripple mask on/off: suiriLab.noRippleMask() — src/js/90-main.js:195 */
int sl_set_ripple_mask(sl_sim *sim, int bound);

/* State transfer in the reference layouts (new; needed for parity). */
size_t sl_field_bytes(const sl_sim *sim, int field);
int sl_upload(sl_sim *sim, int field, const void *src, size_t bytes);
int sl_readback(sl_sim *sim, int field, void *dst, size_t bytes);
/* Host-side FluidSim scalars that the reference keeps in JS (count, time, frame, vmax, full):
   names "count", "time", "frame", "vmax", "full", "cur_dt" */
int sl_set_scalar(sl_sim *sim, const char *name, double value);
int sl_get_scalar(const sl_sim *sim, const char *name, double *value);

/* One stage with the current params (stage-level parity). `dt` sets params.dt for the call. */
int sl_run_stage(sl_sim *sim, int stage, uint32_t arg, double dt);

/* Pressure solver. SL_SOLVER_SOR (the default) is the reference's PRESSURE pass: cfg.iters red-black
   SOR sweeps (src/js/30-wgsl-sim.js:267-297). SL_SOLVER_MGPCG is a converged solve with no
   counterpart in the reference, for callers that need the projection's residual small rather than the
   reference's values: conjugate gradients preconditioned with a multigrid V-cycle on the same system,
   in double precision, iterated until the largest residual, as a divergence, is at most tol_per_s
   (1/s) or max_iters (1..1000) iterations ran; PROJECT then reads the double-precision solution. The
   SL_S_SOR and SL_S_PROJECT stages follow the choice. Not for the tanks of a batch. */
enum sl_solver { SL_SOLVER_SOR = 0, SL_SOLVER_MGPCG = 1 };
int sl_set_pressure_solver(sl_sim *sim, int solver, double tol_per_s, uint32_t max_iters);
/* The last converged solve (waits for the stream): iterations run, 1 if it stopped on a breakdown,
   the largest residual at the end and at the start, in 1/s. */
int sl_pressure_solver_stats(sl_sim *sim, uint32_t *iters, int *breakdown, double *resid_per_s,
                             double *resid0_per_s);

/* Per-stage device timing with HIP events on the sim's stream. */
enum sl_tstage {
  SL_T_EMIT = 0, SL_T_CLEAR, SL_T_GEOM, SL_T_P2G, SL_T_NORMFORCE, SL_T_DIVERGENCE, SL_T_SOR,
  SL_T_PROJECT, SL_T_G2P, SL_T_COMPACT, SL_T_WW, SL_T_RIPPLE, SL_T_DROPLETS, SL_T_COUNT_
};
int sl_timing_enable(sl_sim *sim, int enable);
/* ms[i] = accumulated milliseconds of stage i since the last reset, launches[i] = kernel launches */
int sl_timing_read(sl_sim *sim, double *ms, uint32_t *launches, int n, int reset);

/* Chooses the fastest work decomposition of P2G and G2P for this simulation by timing each variant
   on its live state (between two frames, after at least one step) and restoring everything the
   trials touched. All variants compute bit-identical results. choice (optional, 2 entries) receives
   the selected P2G and G2P variants. No counterpart in the reference. */
int sl_autotune(sl_sim *sim, uint32_t reps, uint32_t *choice);

/* Batched tanks (no counterpart in the reference, which runs one tank per page). A batch is a table
   of tank slots stepped with one kernel launch per stage for all stepped tanks; each tank keeps the
   single-simulation semantics (one substep of dt per frame). Tanks are added and removed while the
   batch runs (continuous batching). Per-tank set-up, single emissions, readback and stats go
   through the tank's sl_sim (sl_batch_sim), which shares the batch's stream. */
typedef struct sl_batch sl_batch;
/* an empty batch of `capacity` slots on `device` */
int sl_batch_open(uint32_t capacity, int device, sl_batch **out);
/* sl_batch_open(n) + sl_batch_add of each configuration, in slot order */
int sl_batch_create(const sl_config *cfgs, uint32_t n, sl_batch **out);
void sl_batch_destroy(sl_batch *batch);
const char *sl_batch_last_error(const sl_batch *batch);
uint32_t sl_batch_size(const sl_batch *batch);                /* slots */
sl_sim *sl_batch_sim(sl_batch *batch, uint32_t slot);         /* NULL for an empty slot */
/* creates a tank in a free slot (stepped from the next sl_batch_step on); SL_ECAP when full */
int sl_batch_add(sl_batch *batch, const sl_config *cfg, uint32_t *slot);
/* destroys the tank of a slot and frees its device memory */
int sl_batch_remove(sl_batch *batch, uint32_t slot);
/* pauses (0) or resumes (1) stepping a tank; a paused tank stays resident for readback */
int sl_batch_set_stepped(sl_batch *batch, uint32_t slot, int on);
/* emission into many tanks at once: counts[slot] records per slot (capacity entries), records
   concatenated in slot order. writeEmits' budget (cap - count - 64, 50-sim.js:222-245) is applied
   on the device count. No `fill` pours here (sl_emit_now on the tank does them). */
int sl_batch_emit(sl_batch *batch, const sl_emit *recs, const uint32_t *counts);
/* one frame of every stepped tank (frame++, P2G .. G2P, compaction, whitewater, ripple, counters
   copy); the ripple mean depth comes from each tank's count at the start of the frame */
int sl_batch_step(sl_batch *batch, double dt);
/* records one frame into a hipGraph; sl_batch_step with the same dt then replays it until the
   membership or the variant choice changes */
int sl_batch_capture(sl_batch *batch, double dt);
int sl_batch_captured(const sl_batch *batch);
/* waits for the batch and refreshes every tank's counters */
int sl_batch_sync(sl_batch *batch);
/* spatial re-ordering of every stepped tank */
int sl_batch_reorder(sl_batch *batch);
/* per size bucket (<= 64Ki cells, <= 256Ki, larger): fastest P2G / G2P variant on the live state;
   choice (optional, 6 entries) receives (p2g, g2p) per bucket */
int sl_batch_autotune(sl_batch *batch, uint32_t reps, uint32_t *choice);
/* sets the (p2g, g2p) variants of the three buckets (6 entries, as sl_batch_autotune reports) */
int sl_batch_set_variants(sl_batch *batch, const uint32_t *choice);
/* per-step device time: while enabled, each sl_batch_step is bracketed by events; read returns the
   durations [ms] of the steps since enable or the last read, oldest first */
int sl_batch_timing_enable(sl_batch *batch, int enable);
int sl_batch_timing_read(sl_batch *batch, double *ms, uint32_t max, uint32_t *n);
/* a tank's own frame length, used by sl_batch_step(batch, 0) (tanks of one batch may differ) */
int sl_batch_set_dt(sl_batch *batch, uint32_t slot, double dt);

/* Bake output of the current frame into `dir`, in the layout of the reference's Blender cache
   (blender_addon/suiri_lab/bake.py:226-249): frame_%05d.npz with pos, vel (live particles), spray,
   foam, bubble (whitewater positions by kind), liters, particles, time, in Blender axes
   (X, Y, Z) = (x, -z, y); and, for a tank with a ripple layer, surface_%05d.exr (exr.py:1-41:
   uncompressed float R = top height, G = ripple slope d/dX, B = d/dY at the map's texel centres).
   Synchronous (reads the tank back). */
int sl_bake_frame(sl_sim *sim, const char *dir, uint32_t frame);
/* the EXR layout of the bake maps (exr.py:11-41); r, g, b: H x W, row 0 at the top; host only */
int sl_write_exr_rgb(const char *path, uint32_t W, uint32_t H, const float *r, const float *g, const float *b);

/* device bytes a simulation of this configuration holds (sum of its allocation sizes; host only) */
size_t sl_job_bytes(const sl_config *cfg);

/* Serving loop (no counterpart in the reference, whose bake.py runs one scene at a time).
   Scene-bake jobs are submitted at any time; each tick admits queued jobs (first fit in arrival
   order) while a batch slot and the device-memory budget allow, advances every running job by one
   frame of its own dt in one batched step, and retires the jobs that reached their frame count.
   State stays resident on the device from admission to retirement (or to sl_server_release for
   jobs submitted with keep = 1). A job's result does not depend on the other jobs. */
typedef struct sl_server sl_server;
#define SL_SERVER_MAX_FRAMES_PER_TICK 64
typedef struct {
  uint64_t hbm_budget;      /* device bytes the resident jobs may hold (sl_job_bytes each) */
  uint32_t max_jobs;        /* batch slots */
  int32_t device;
  uint32_t autotune_after;  /* ticks of unchanged membership before the one variant tuning (0 = never,
                               the default) */
  uint32_t graph_after;     /* capture the frame after this many unchanged ticks (0 = eager launches) */
  uint32_t frames_per_tick; /* frames one tick enqueues back to back (0 = 1, default 4, at most
                               SL_SERVER_MAX_FRAMES_PER_TICK); a tick stops early at any running
                               job's last frame or bake frame, so retirement and bakes keep their
                               frames, and jobs submitted meanwhile wait at most this many */
  uint32_t reserved[3];
} sl_server_config;
typedef struct {
  sl_config cfg;            /* the tank */
  const float *objects;     /* obstacles and drains, 28 floats each (as sl_set_objects); copied */
  uint32_t n_objects;
  const sl_emit *fill;      /* emitted at admission (water blocks); copied */
  uint32_t n_fill;
  uint32_t steps;           /* frames to run */
  double dt;                /* frame length, one substep per frame */
  sl_emit_fn source;        /* per-frame emission (sources), or NULL */
  void *user;
  uint32_t keep;            /* keep the final state resident until sl_server_release */
  uint32_t bake_every;      /* with bake_dir: sl_bake_frame at admission, every bake_every frames and at
                               the last frame (0: only at the last) */
  const char *bake_dir;     /* NULL: no bake output; copied at submission */
  /* scene items (copied; instead of `objects`): at admission the obstacles and drains are packed
     (sl_set_items) and the blocks fill the tank (sl_scene_block_emits, after `fill`); every frame the
     sources emit (sl_scene_source_emits, after the records of `source`) */
  const sl_item *items;
  uint32_t n_items;
  uint32_t reserved;
} sl_job;
enum { SL_JOB_QUEUED = 0, SL_JOB_RUNNING = 1, SL_JOB_DONE = 2, SL_JOB_CANCELLED = 3, SL_JOB_REFUSED = 4 };
typedef struct {
  uint64_t id;
  int32_t status;
  uint32_t frames, steps, slot;
  double t_submit, t_admit, t_done;   /* seconds on the server clock; -1 when not reached */
  double particle_steps;
  uint64_t bytes;
  uint32_t count, ww_count;
} sl_job_info;
typedef struct {
  uint64_t budget, used;
  uint32_t queued, running, resident, done;
  uint64_t ticks, frames;
  double particle_steps, uptime;      /* uptime [s] */
  double step_ms_p50, step_ms_p99, step_ms_sum;   /* device time per batched frame */
  double latency_p50, latency_p99;    /* submit -> done of finished jobs [s] */
} sl_server_stats_t;
int sl_server_config_default(sl_server_config *cfg);
int sl_server_create(const sl_server_config *cfg, sl_server **out);
void sl_server_destroy(sl_server *srv);
const char *sl_server_last_error(const sl_server *srv);
/* queues a job; SL_EBUDGET (job refused, *id still set) when it alone exceeds the budget */
int sl_server_submit(sl_server *srv, const sl_job *job, uint64_t *id);
/* cancels a queued or running job at the next tick (its memory is freed then) */
int sl_server_cancel(sl_server *srv, uint64_t id);
/* one scheduling round and one frame (up to frames_per_tick frames) of every running job;
   *running = jobs stepped */
int sl_server_tick(sl_server *srv, uint32_t *running);
/* jobs retired (done, cancelled, refused) since the last poll */
int sl_server_poll(sl_server *srv, sl_job_info *out, uint32_t max, uint32_t *n);
int sl_server_info(const sl_server *srv, uint64_t id, sl_job_info *out);
/* the resident state of a running or kept job (readback through the sl_sim entry points) */
sl_sim *sl_server_job_sim(sl_server *srv, uint64_t id);
/* frees a finished job kept resident */
int sl_server_release(sl_server *srv, uint64_t id);
int sl_server_stats(const sl_server *srv, sl_server_stats_t *out);

/* One unmasked ripple frame (RippleLayer.encodeFrame, src/js/64-ripple.js:222-253, the mask unbound) on
   caller-owned device buffers, allocating no device memory. No counterpart in the reference: for
   callers that cannot allocate (code under HIP graph capture, tight serving loops). The same kernels
   as sl_run_stage(sim, SL_S_RIPPLE) with the mask off: `state` (M^2 float2, read and written), `src`
   ((R+1)^2 i32 impulses at 1e9 per N s, read and cleared), `tex` ((R+1)^2 float4 height, slope x,
   slope z, 0; written), on the surface Lx x Lz [m] with frame length dt [s] and depth [m] (at least
   0.002 is used); the physics constants are cfg's ripple fields (r_*, gravity). M = 2R, a power of
   two, at most SL_RIPPLE_NOALLOC_MAX_M. `stream` is a hipStream_t (NULL: the null stream), on the
   current device. The tank record, the twiddles and the three M x M work fields live in one static
   device array of 6.3 MB per GPU, kept until the process ends; calls that overlap on one GPU race. */
#define SL_RIPPLE_NOALLOC_MAX_M 512
int sl_ripple_frame_noalloc(const sl_config *cfg, uint32_t R, double Lx, double Lz, double dt, double depth,
                            void *state, int32_t *src, void *tex, void *stream);

/* ABI version + build description */
int sl_version(void);
const char *sl_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
