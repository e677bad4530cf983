/*
 * Oracle (test infrastructure only): a plain-C, fp32 restatement of the reference liquid step.
 *
 * This library is the CHECKER for the HIP implementation. It is never linked into, loaded by or
 * called from the product path; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg use it.
 *
 * It restates, kernel by kernel, the WGSL compute shaders of the reference and the driver order of
 * FluidSim / RippleLayer. Buffers keep the reference layouts exactly (array-of-structs particles,
 * vec4f per MAC node, x-fastest indices) so that a reader can put each function next to the WGSL
 * it follows. Citations are path:line in the reference repository.
 *
 * Numeric conventions (WGSL semantics): round() is half-to-even (rintf); i32()/u32() truncate and
 * saturate; mix(a,b,t) = a*(1-t) + b*t; u32 arithmetic wraps; constant sub-expressions of abstract
 * floats are folded in double precision and rounded once to f32; transcendentals (exp, pow, sin, cos,
 * sinh) are the correctly rounded f32 values (double evaluation rounded once), the most accurate
 * results the WGSL error bounds allow. Build with -ffp-contract=off.
 *
 * One deliberate, documented choice where the reference is nondeterministic: particle compaction
 * (src/js/31-wgsl-sim2.js:317-328 appends survivors with an atomic counter, so the order changes from
 * run to run) is restated as a deterministic in-place compaction: survivors below the new count n'
 * keep their slots, and the dead slots below n' take, in ascending order, the survivors at or above
 * n', in ascending order. That order is one of the orders the reference can produce, and it makes
 * every later step reproducible.
 */
#ifndef SUIRI_ORACLE_H
#define SUIRI_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_OBJ 64   /* src/js/50-sim.js:9 */
#define OR_MAX_EMIT 48  /* src/js/50-sim.js:10 */

/* Params uniform, src/js/30-wgsl-sim.js:8-16 (7 x vec4 = 112 B used of a 128 B buffer). */
typedef struct {
  uint32_t dims[4];   /* nx, ny, nz, node count */
  uint32_t ext[4];    /* cell count, particle capacity, object count, frame */
  float origin[4];    /* grid min corner, cell size h */
  float phys[4];      /* dt, gravity, flip ratio, 1/h */
  float phys2[4];     /* ppc, drift, omega, collision radius */
  float tank[4];      /* rim, ww rate, ww capacity, friction */
  float misc[4];      /* time, render radius, ripple R, 0 */
  uint32_t pad[4];
} or_params;

typedef struct { float pos[4]; float vel[4]; } or_particle;            /* 30-wgsl-sim.js:17 */
typedef struct { float r0[4], r1[4], r2[4], pos[4], dims[4], vel[4], angv[4]; } or_obj; /* :18 */
typedef struct { uint32_t count, newCount, wwCount, wwNew, vmax, impacts, pad0, pad1; } or_counters; /* 31-wgsl-sim2.js:5 */
typedef struct { float a[4], b[4], c[4], d[4]; uint32_t info[4]; } or_emit;  /* 31-wgsl-sim2.js:255 */

/* Quality / solver settings. Defaults are the reference constants (50-sim.js:4-8, :151-162). */
typedef struct {
  double h;          /* requested cell size [m] */
  uint32_t cap;      /* particle capacity */
  uint32_t ww_cap;   /* whitewater capacity */
  uint32_t iters;    /* SOR iterations */
  uint32_t ripple_R; /* ripple physical cells (M = 2R); 0 = no ripple layer */
  double rc_factor;  /* collision radius / h; the reference multiplies 0.32 * h in double (50-sim.js:158) */
  float gravity, flip, ppc, drift, omega, friction, ww_rate;
  /* ripple physics: sigma/rho, nu, film damping, high-pass k0, kick density */
  float r_sigma_rho, r_nu, r_film, r_hp_k, r_rho_kick;
  uint32_t wmap;     /* top-height map size used by the ripple mask (256 in the reference) */
  uint32_t reorder_every; /* spatial re-ordering period in frames (0 = never) */
  uint32_t reorder_phase; /* frames already counted towards the first re-ordering */
  uint32_t top_height;    /* 1: produce the top-height map from the particles at each frame's end */
  /* splashCrown liquid constants and K thresholds (31-wgsl-sim2.js:121-126: 998, 0.0728, 0.001, 700, 2100) */
  double splash_rho, splash_sigma, splash_mu;
  float splash_k_jet, splash_k_crown;
  double ripple_depth;    /* > 0: fixed ripple depth; 0: the mean depth of the count (50-sim.js:332) */
  /* sub-grid droplets per frame (0 = none): seed, lognormal median diameter and sigma of ln D, speeds */
  uint32_t droplets, drop_seed;
  float drop_d50, drop_sigma_ln, drop_vmin, drop_vmax;
} or_quality;

typedef struct or_sim {
  /* host-side state of FluidSim (50-sim.js:29-58) */
  int nx, ny, nz;
  double h, origin[3], rim, vp, radius, time, vmax;
  uint32_t nodes, cells, cap, ww_cap, frame, obj_count, count, ww_count, impacts, iters;
  int full, geom_dirty, cur, ww_cur;
  or_quality q;
  or_params P;
  /* device buffers (50-sim.js:59-80) in the reference layout */
  or_particle *parts[2];
  or_particle *ww[2];
  or_counters C;
  uint32_t args[16];
  int32_t *acc;        /* N*6 */
  uint32_t *cellCount; /* C */
  float *vel, *velOld, *frac, *svel; /* N*4 */
  uint32_t *cellType;  /* C */
  float *divg, *pres;  /* C */
  or_obj objs[OR_MAX_OBJ];
  or_emit emits[OR_MAX_EMIT];
  uint32_t emitU[4];
  /* ripple layer (64-ripple.js:149-197) */
  uint32_t R, M;
  int32_t *rsrc;       /* (R+1)^2 */
  float *rstate, *ra, *rb, *rhat; /* M*M*2 */
  float *tw;           /* M*2 */
  float *mask;         /* (R+1)^2 */
  float *tex;          /* (R+1)^2 * 4 : height, slope x, slope z, 0 (f32; the reference stores rgba16f) */
  float *topH;         /* wmap*wmap, -1 where dry; zero-initialised like a fresh WebGPU texture */
  float *topB;         /* blur ping-pong of topH */
  double tank_h;       /* scene tank height (the renderer's F.tank.y) */
  int mask_bound;
  float rU_l[4];       /* last ripple uniform l = (Lx, Lz, dt, meanDepth) */
  int nthreads;
  uint32_t reorder_clock;
} or_sim;

/* sizing + allocation, src/js/50-sim.js:29-89 */
void or_quality_default(or_quality *q, const char *name);
or_sim *or_create(double tank_w, double tank_h, double tank_d, const or_quality *q);
void or_destroy(or_sim *s);
void or_set_threads(or_sim *s, int n);
void or_write_params(or_sim *s, double dt);                   /* 50-sim.js:151-162 */
int or_set_objects(or_sim *s, const float *rows, uint32_t n); /* packed Obj rows; 50-sim.js:194-209 */

/* kernels (one call = one dispatch over its full range) */
void or_geom(or_sim *s);                   /* 30-wgsl-sim.js:118-157 */
void or_p2g(or_sim *s);                    /* 30-wgsl-sim.js:159-200 */
void or_normalize(or_sim *s);              /* 30-wgsl-sim.js:203-222 */
void or_forces(or_sim *s);                 /* 30-wgsl-sim.js:224-238 */
void or_divergence(or_sim *s);             /* 30-wgsl-sim.js:240-264 */
void or_pressure(or_sim *s, uint32_t parity); /* 30-wgsl-sim.js:267-297 */
void or_project(or_sim *s);                /* 30-wgsl-sim.js:299-341 */
void or_g2p(or_sim *s);                    /* 31-wgsl-sim2.js:149-251 */
void or_emit_kernel(or_sim *s, uint32_t k, uint32_t total); /* 31-wgsl-sim2.js:262-315 */
void or_compact(or_sim *s);                /* 31-wgsl-sim2.js:317-328 (in place, holes filled from the tail) */
/* stable sort of the live particles by the Morton code of their 4x4x4-cell brick (outside the grid
   last); no counterpart in the reference, whose particle order is arbitrary after compaction */
void or_reorder(or_sim *s);
void or_toph(or_sim *s);       /* or_toph_splat + or_toph_blur */
void or_toph_splat(or_sim *s);
void or_toph_blur(or_sim *s);
void or_count(or_sim *s, uint32_t mode, uint32_t y); /* 31-wgsl-sim2.js:332-349 */
void or_ww(or_sim *s, float dt);           /* 31-wgsl-sim2.js:361-409 */
void or_crown(or_sim *s, const float *x, float U, float m, uint32_t id); /* splashCrown, 31-wgsl-sim2.js:119-148 */
/* the frame's sub-grid droplets (q.droplets): impulses into rsrc, splashes into the whitewater; the
   water level is origin y + depth */
void or_droplets(or_sim *s, double depth);
void or_droplet_draws(const or_sim *s, float *out);  /* per droplet (x, z, D, v) of the current frame */

/* ripple kernels, src/js/64-ripple.js */
void or_rsrc(or_sim *s, float *out);                                   /* :18-31 */
void or_rfft_pass(or_sim *s, const float *in, float *out, uint32_t ns, uint32_t horiz, float sign); /* :34-56 */
float *or_rfft2d(or_sim *s, float *input, float sign, float *final_out);/* :205-216 */
void or_revolve(or_sim *s);                                            /* :58-90 */
void or_rmaskbuild(or_sim *s);                                         /* :106-126 */
void or_rmask(or_sim *s, const float *field, float *out);              /* :91-104 */
void or_rscale(or_sim *s);                                             /* :127-133 */
void or_rextract(or_sim *s, const float *field);                       /* :134-147 */
void or_ripple_frame(or_sim *s, double dt, double mean_depth);         /* :222-253 */

/* driver, src/js/50-sim.js:222-336 */
uint32_t or_write_emits(or_sim *s, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds);
void or_emit_now(or_sim *s, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds, int fill);
void or_substep(or_sim *s);                /* encodeSubstep, 50-sim.js:276-297 */
/* one substep of step(): params, emission, count, substep; 50-sim.js:307-318 */
void or_step_substep(or_sim *s, double dt, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds);
void or_step_begin(or_sim *s);             /* counters.vmax = 0, 50-sim.js:306 */
void or_frame_tail(or_sim *s, double frame_dt); /* 50-sim.js:319-335 */
void or_poll_stats(or_sim *s);             /* 50-sim.js:346-363 (synchronous here) */
void or_clear_water(or_sim *s);            /* 50-sim.js:365-373 */
int or_substep_count(const or_sim *s, double frame_dt); /* 50-sim.js:301-305 */

/* raw access for the Python checker: pointer + byte size of a named buffer
   ("parts", "ww", "counters", "acc", "cellCount", "vel", "velOld", "frac", "svel", "cellType", "divg",
   "pres", "objs", "emits", "rsrc", "rstate", "ra", "rb", "rhat", "tw", "mask", "tex", "topH", "params") */
void *or_field(or_sim *s, const char *name, size_t *bytes);
/* scalars: nx, ny, nz, h, ox, oy, oz, rim, vp, nodes, cells, cap, ww_cap, frame, obj_count, count,
   ww_count, impacts, time, vmax, full, cur, ww_cur, R, M, iters (26 doubles) */
void or_info(const or_sim *s, double *out);
void or_set_scalar(or_sim *s, const char *name, double v);

/* helpers exported for known-answer tests */
uint32_t or_hashU(uint32_t a);
float or_scene_sdf(const or_sim *s, const float p[3]);
void or_scene_normal(const or_sim *s, const float p[3], float out[3]);
float or_face_open(const or_sim *s, const float c[3], uint32_t axis);

#ifdef __cplusplus
}
#endif
#endif
