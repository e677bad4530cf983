// Node N-API binding of the suiri C ABI (include/suiri.h), so that a JavaScript host keeps the
// reference's FluidSim interface (src/js/50-sim.js) with the HIP library underneath.
//
// Every function runs on the JS thread and calls the C ABI synchronously; the C ABI itself only
// enqueues GPU work (like GPUQueue.submit), so step() returns without waiting for the device.
// Status codes other than SL_OK become JS exceptions carrying sl_last_error().
#include <node_api.h>
#include <stdint.h>
#include <string.h>
#include <map>
#include <string>
#include <vector>

#include "../../include/suiri.h"

#define NAPI_CALL(env, call)                                             \
  do {                                                                   \
    if ((call) != napi_ok) {                                             \
      napi_throw_error((env), nullptr, "N-API call failed: " #call);     \
      return nullptr;                                                    \
    }                                                                    \
  } while (0)

static napi_value throw_sl(napi_env env, const sl_sim *s, int code) {
  std::string m = "suiri error " + std::to_string(code) + ": " + (sl_last_error(s) ? sl_last_error(s) : "");
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

static napi_value num(napi_env env, double v) {
  napi_value r;
  napi_create_double(env, v, &r);
  return r;
}
static napi_value boolean(napi_env env, bool v) {
  napi_value r;
  napi_get_boolean(env, v, &r);
  return r;
}
static double get_num(napi_env env, napi_value obj, const char *key, double dflt, bool *found = nullptr) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (found) *found = has;
  if (!has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  double d;
  napi_get_value_double(env, v, &d);
  return d;
}
static void set_num(napi_env env, napi_value obj, const char *key, double v) {
  napi_set_named_property(env, obj, key, num(env, v));
}

struct Handle {
  sl_sim *sim = nullptr;
  bool owned = true;  // false: a tank of a served job, destroyed with its server
  int pending = 0;    // statsAsync waits in flight on worker threads
  bool doomed = false;  // destroy() came while a wait was in flight: destroyed when the last one ends
};
static void finalize(napi_env, void *data, void *) {
  Handle *h = (Handle *)data;
  if (h->sim && h->owned) sl_destroy(h->sim);
  delete h;
}
static sl_sim *get_sim(napi_env env, napi_value v) {
  Handle *h = nullptr;
  if (napi_get_value_external(env, v, (void **)&h) != napi_ok || !h || !h->sim || h->doomed) {
    napi_throw_error(env, nullptr, "suiri: invalid or destroyed simulation handle");
    return nullptr;
  }
  return h->sim;
}
static bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "suiri: missing arguments");
    return false;
  }
  return true;
}

// ---- config ----
#define CFG_FIELDS(X)                                                                           \
  X(tank_w) X(tank_h) X(tank_d) X(h) X(cap) X(ww_cap) X(iters) X(ripple_R) X(rc_factor) X(gravity) \
  X(flip) X(ppc) X(drift) X(omega) X(friction) X(ww_rate) X(r_sigma_rho) X(r_nu) X(r_film)         \
  X(r_hp_k) X(r_rho_kick) X(wmap) X(whitewater) X(device) X(reorder_every) X(reorder_phase) X(top_height)  \
  X(seed) X(splash_rho) X(splash_sigma) X(splash_mu) X(splash_k_jet) X(splash_k_crown) X(ripple_depth)     \
  X(droplets) X(drop_seed) X(drop_d50) X(drop_sigma_ln) X(drop_vmin) X(drop_vmax)

static napi_value config_to_js(napi_env env, const sl_config &c) {
  napi_value o;
  napi_create_object(env, &o);
#define X(f) set_num(env, o, #f, (double)c.f);
  CFG_FIELDS(X)
#undef X
  return o;
}
static void config_from_js(napi_env env, napi_value o, sl_config &c) {
#define X(f) c.f = (decltype(c.f))get_num(env, o, #f, (double)c.f);
  CFG_FIELDS(X)
#undef X
}

// configDefault(quality) -> object
static napi_value ConfigDefault(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  size_t argc = 1;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  char q[32] = "standard";
  if (argc >= 1) {
    size_t len = 0;
    napi_get_value_string_utf8(env, argv[0], q, sizeof q, &len);
  }
  sl_config c;
  int r = sl_config_default(&c, q);
  if (r) return throw_sl(env, nullptr, r);
  return config_to_js(env, c);
}

// create(config) -> handle
static napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_config c;
  sl_config_default(&c, "standard");
  config_from_js(env, argv[0], c);
  sl_sim *s = nullptr;
  int r = sl_create(&c, &s);
  if (r) return throw_sl(env, nullptr, r);
  Handle *h = new Handle();
  h->sim = s;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, finalize, nullptr, &ext));
  return ext;
}

static napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  Handle *h = nullptr;
  if (napi_get_value_external(env, argv[0], (void **)&h) == napi_ok && h && h->sim) {
    if (h->pending) { h->doomed = true; return nullptr; }  // a worker still waits on its stream
    sl_destroy(h->sim);
    h->sim = nullptr;
  }
  return nullptr;
}

static napi_value GridInfo(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  sl_grid_t g;
  int r = sl_grid_info(s, &g);
  if (r) return throw_sl(env, s, r);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "nx", g.nx); set_num(env, o, "ny", g.ny); set_num(env, o, "nz", g.nz);
  set_num(env, o, "nodes", g.nodes); set_num(env, o, "cells", g.cells); set_num(env, o, "cap", g.cap);
  set_num(env, o, "wwCap", g.ww_cap); set_num(env, o, "R", g.R); set_num(env, o, "M", g.M);
  set_num(env, o, "iters", g.iters); set_num(env, o, "h", g.h); set_num(env, o, "rim", g.rim);
  set_num(env, o, "vp", g.vp); set_num(env, o, "radius", g.radius);
  napi_value org;
  napi_create_array_with_length(env, 3, &org);
  for (uint32_t i = 0; i < 3; i++) napi_set_element(env, org, i, num(env, g.origin[i]));
  napi_set_named_property(env, o, "origin", org);
  return o;
}

// bytes of a typed array / array buffer / buffer
static bool get_bytes(napi_env env, napi_value v, void **data, size_t *len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (is) {
    napi_typedarray_type t;
    size_t n, off;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &t, &n, data, &ab, &off) != napi_ok) return false;
    size_t es = 1;
    switch (t) {
      case napi_int16_array: case napi_uint16_array: es = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: es = 4; break;
      case napi_float64_array: es = 8; break;
      default: es = 1;
    }
    *len = n * es;
    return true;
  }
  napi_is_arraybuffer(env, v, &is);
  if (is) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  napi_is_buffer(env, v, &is);
  if (is) return napi_get_buffer_info(env, v, data, len) == napi_ok;
  return false;
}

// setObjects(h, Float32Array rows) -> changed
static napi_value SetObjects(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  void *data = nullptr;
  size_t len = 0;
  if (!get_bytes(env, argv[1], &data, &len)) {
    napi_throw_type_error(env, nullptr, "setObjects: rows must be a Float32Array");
    return nullptr;
  }
  int changed = 0;
  int r = sl_set_objects(s, (const float *)data, (uint32_t)(len / (28 * 4)), &changed);
  if (r) return throw_sl(env, s, r);
  return boolean(env, changed != 0);
}

// emission records: [{shape, a:[4], b:[4], c:[4]?, d:[4]?, count, seed}]
static bool read_vec4(napi_env env, napi_value rec, const char *key, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  bool has = false;
  napi_has_named_property(env, rec, key, &has);
  if (!has) return true;
  napi_value arr;
  napi_get_named_property(env, rec, key, &arr);
  napi_valuetype t;
  napi_typeof(env, arr, &t);
  if (t == napi_undefined || t == napi_null) return true;
  bool isarr = false;
  napi_is_array(env, arr, &isarr);
  if (!isarr) return false;
  uint32_t n = 0;
  napi_get_array_length(env, arr, &n);
  for (uint32_t i = 0; i < n && i < 4; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    double d = 0;
    napi_get_value_double(env, e, &d);
    out[i] = (float)d;
  }
  return true;
}
static bool read_emits(napi_env env, napi_value list, std::vector<sl_emit> &out) {
  bool isarr = false;
  napi_is_array(env, list, &isarr);
  if (!isarr) return false;
  uint32_t n = 0;
  napi_get_array_length(env, list, &n);
  out.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value rec;
    napi_get_element(env, list, i, &rec);
    sl_emit &e = out[i];
    memset(&e, 0, sizeof e);
    if (!read_vec4(env, rec, "a", e.a) || !read_vec4(env, rec, "b", e.b) || !read_vec4(env, rec, "c", e.c) ||
        !read_vec4(env, rec, "d", e.d))
      return false;
    e.shape = (uint32_t)get_num(env, rec, "shape", 0);
    double cnt = get_num(env, rec, "count", 0);
    e.count = cnt > 0 ? (uint32_t)cnt : 0u;
    e.seed = (uint32_t)(int64_t)get_num(env, rec, "seed", 0);
  }
  return true;
}

// emitNow(h, list, fill) -> emitted
static napi_value EmitNow(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  size_t argc = 3;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  if (argc < 2) { napi_throw_type_error(env, nullptr, "emitNow(h, list, fill)"); return nullptr; }
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  std::vector<sl_emit> list;
  if (!read_emits(env, argv[1], list)) { napi_throw_type_error(env, nullptr, "emitNow: bad list"); return nullptr; }
  bool fill = false;
  if (argc >= 3) napi_get_value_bool(env, argv[2], &fill);
  uint32_t emitted = 0;
  int r = sl_emit_now(s, list.data(), (uint32_t)list.size(), fill ? 1 : 0, &emitted);
  if (r) return throw_sl(env, s, r);
  return num(env, emitted);
}

// stepBegin(h, frameDt) -> {n, dt}
static napi_value StepBegin(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  double fdt = 0;
  napi_get_value_double(env, argv[1], &fdt);
  uint32_t n = 0;
  double dt = 0;
  int r = sl_step_begin(s, fdt, &n, &dt);
  if (r) return throw_sl(env, s, r);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "n", n);
  set_num(env, o, "dt", dt);
  return o;
}

// substep(h, dt, list) -> emitted
static napi_value Substep(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  double dt = 0;
  napi_get_value_double(env, argv[1], &dt);
  std::vector<sl_emit> list;
  if (!read_emits(env, argv[2], list)) { napi_throw_type_error(env, nullptr, "substep: bad list"); return nullptr; }
  uint32_t emitted = 0;
  int r = sl_substep(s, dt, list.data(), (uint32_t)list.size(), &emitted);
  if (r) return throw_sl(env, s, r);
  return num(env, emitted);
}

static napi_value StepEnd(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  double fdt = 0;
  napi_get_value_double(env, argv[1], &fdt);
  int r = sl_step_end(s, fdt);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}

// simple (h) -> status calls
#define SIMPLE(NAME, CALL)                                        \
  static napi_value NAME(napi_env env, napi_callback_info info) { \
    napi_value argv[1];                                           \
    if (!get_args(env, info, 1, argv)) return nullptr;            \
    sl_sim *s = get_sim(env, argv[0]);                            \
    if (!s) return nullptr;                                       \
    int r = CALL(s);                                              \
    if (r) return throw_sl(env, s, r);                            \
    return nullptr;                                               \
  }
SIMPLE(ClearWater, sl_clear_water)
SIMPLE(Sync, sl_sync)
SIMPLE(Reorder, sl_reorder)

static napi_value SetSubsteps(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[1], &n);
  int r = sl_set_substeps(s, n);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}
// setPressureSolver(sim, solver, tolPerS, maxIters): 0 = the reference's SOR sweeps, 1 = converged solve
static napi_value SetPressureSolver(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  int32_t solver = 0;
  double tol = 0;
  uint32_t iters = 0;
  napi_get_value_int32(env, argv[1], &solver);
  napi_get_value_double(env, argv[2], &tol);
  napi_get_value_uint32(env, argv[3], &iters);
  int r = sl_set_pressure_solver(s, solver, tol, iters);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}
// pressureSolverStats(sim) -> {iters, breakdown, residPerS, resid0PerS}
static napi_value PressureSolverStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  uint32_t it = 0;
  int br = 0;
  double r1 = 0, r0 = 0;
  int r = sl_pressure_solver_stats(s, &it, &br, &r1, &r0);
  if (r) return throw_sl(env, s, r);
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "iters", num(env, it));
  napi_set_named_property(env, o, "breakdown", boolean(env, br != 0));
  napi_set_named_property(env, o, "residPerS", num(env, r1));
  napi_set_named_property(env, o, "resid0PerS", num(env, r0));
  return o;
}
static napi_value SetRippleMask(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  bool b = false;
  napi_get_value_bool(env, argv[1], &b);
  int r = sl_set_ripple_mask(s, b ? 1 : 0);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}
static napi_value SaveWater(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  uint32_t id = 0;
  int r = sl_save_water(s, &id);
  if (r) return throw_sl(env, s, r);
  return num(env, id);
}
static napi_value HasWaterCopy(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  uint32_t id = 0;
  napi_get_value_uint32(env, argv[1], &id);
  return boolean(env, sl_has_water_copy(s, id) != 0);
}
static napi_value RestoreWater(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  uint32_t id = 0;
  napi_get_value_uint32(env, argv[1], &id);
  int r = sl_restore_water(s, id);
  if (r < 0) return throw_sl(env, s, r);
  return boolean(env, r == 1);
}
static napi_value PollStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  int u = 0;
  int r = sl_poll_stats(s, &u);
  if (r) return throw_sl(env, s, r);
  return boolean(env, u != 0);
}
static napi_value stats_js(napi_env env, const sl_stats_t &st) {
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "count", st.count); set_num(env, o, "ww", st.ww_count); set_num(env, o, "impacts", st.impacts);
  set_num(env, o, "frames", st.frames); set_num(env, o, "vmax", st.vmax); set_num(env, o, "time", st.time);
  set_num(env, o, "liters", st.liters); set_num(env, o, "gpuErrors", st.gpu_errors);
  set_num(env, o, "substeps", st.substeps_last);
  napi_set_named_property(env, o, "full", boolean(env, st.full != 0));
  napi_value grid;
  napi_create_array_with_length(env, 3, &grid);
  for (uint32_t i = 0; i < 3; i++) napi_set_element(env, grid, i, num(env, st.grid[i]));
  napi_set_named_property(env, o, "grid", grid);
  return o;
}
static napi_value Stats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  sl_stats_t st;
  int r = sl_stats(s, &st);
  if (r) return throw_sl(env, s, r);
  return stats_js(env, st);
}

// statsAsync(h) -> Promise of the stats object: a worker thread waits for the work submitted so far
// (sl_wait_idle, which touches no simulation state), then the JS thread reads the counters
// (suiriLab.stats() after device.queue.onSubmittedWorkDone(), 90-main.js:159-165, :196)
struct StatsJob {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref handle_ref = nullptr;  // keeps the simulation alive while the worker waits
  sl_sim *sim = nullptr;
  int wait_rc = 0;
};
static void stats_execute(napi_env, void *data) {
  StatsJob *j = (StatsJob *)data;
  j->wait_rc = sl_wait_idle(j->sim);
}
static void stats_complete(napi_env env, napi_status status, void *data) {
  StatsJob *j = (StatsJob *)data;
  napi_value handle = nullptr;
  napi_get_reference_value(env, j->handle_ref, &handle);
  Handle *h = nullptr;
  napi_get_value_external(env, handle, (void **)&h);
  if (h) h->pending--;
  if (h && h->doomed && !h->pending && h->sim) {  // destroy() was called while the worker waited
    sl_destroy(h->sim);
    h->sim = nullptr;
  }
  sl_stats_t st;
  int r = j->wait_rc;
  if (status != napi_ok) r = SL_ESTATE;
  else if (!h || !h->sim) r = SL_ESTATE;
  else if (!r) r = sl_stats(h->sim, &st);
  if (r) {
    std::string m = "suiri error " + std::to_string(r) + ": " + (h && h->sim ? sl_last_error(h->sim) : "simulation destroyed");
    napi_value msg, err;
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  } else {
    napi_resolve_deferred(env, j->deferred, stats_js(env, st));
  }
  napi_delete_reference(env, j->handle_ref);
  napi_delete_async_work(env, j->work);
  delete j;
}
static napi_value StatsAsync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  StatsJob *j = new StatsJob();
  j->sim = s;
  Handle *hh = nullptr;
  napi_get_value_external(env, argv[0], (void **)&hh);
  hh->pending++;
  napi_value promise, name;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_reference(env, argv[0], 1, &j->handle_ref));
  napi_create_string_utf8(env, "suiri.statsAsync", NAPI_AUTO_LENGTH, &name);
  NAPI_CALL(env, napi_create_async_work(env, nullptr, name, stats_execute, stats_complete, j, &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}
// hostState(h) -> {count, time, frame, vmax, full, impacts, wwCount}: no device wait
static napi_value HostState(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  napi_value o;
  napi_create_object(env, &o);
  const char *names[] = {"count", "time", "frame", "vmax", "full", "impacts", "ww_count"};
  const char *keys[] = {"count", "time", "frame", "vmax", "full", "impacts", "wwCount"};
  for (int i = 0; i < 7; i++) {
    double v = 0;
    sl_get_scalar(s, names[i], &v);
    set_num(env, o, keys[i], v);
  }
  return o;
}
static napi_value SetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  char name[32];
  size_t len = 0;
  napi_get_value_string_utf8(env, argv[1], name, sizeof name, &len);
  double v = 0;
  napi_get_value_double(env, argv[2], &v);
  int r = sl_set_scalar(s, name, v);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}

// readback(h, field) -> ArrayBuffer (reference layout)
static napi_value Readback(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  int32_t f = 0;
  napi_get_value_int32(env, argv[1], &f);
  size_t n = sl_field_bytes(s, f);
  if (!n) { napi_throw_range_error(env, nullptr, "readback: field not present"); return nullptr; }
  void *data = nullptr;
  napi_value ab;
  NAPI_CALL(env, napi_create_arraybuffer(env, n, &data, &ab));
  int r = sl_readback(s, f, data, n);
  if (r) return throw_sl(env, s, r);
  return ab;
}
// upload(h, field, bytes)
static napi_value Upload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  int32_t f = 0;
  napi_get_value_int32(env, argv[1], &f);
  void *data = nullptr;
  size_t len = 0;
  if (!get_bytes(env, argv[2], &data, &len)) { napi_throw_type_error(env, nullptr, "upload: bytes"); return nullptr; }
  int r = sl_upload(s, f, data, len);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}
// bakeFrame(h, dir, frame): frame_%05d.npz (+ surface_%05d.exr) of the current state
static napi_value BakeFrame(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  char dir[4096];
  size_t len = 0;
  NAPI_CALL(env, napi_get_value_string_utf8(env, argv[1], dir, sizeof dir, &len));
  double f = 0;
  napi_get_value_double(env, argv[2], &f);
  int r = sl_bake_frame(s, dir, (uint32_t)f);
  if (r) return throw_sl(env, s, r);
  return nullptr;
}

// jobBytes(config): device bytes of a tank of this configuration
static napi_value JobBytes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_config c;
  sl_config_default(&c, "standard");
  config_from_js(env, argv[0], c);
  return num(env, (double)sl_job_bytes(&c));
}

// ---- serving loop (sl_server_*): jobs with an optional JS source called during tick ----
struct JobCb {
  napi_env env;
  napi_ref fn;
};
struct ServerHandle {
  sl_server *srv = nullptr;
  napi_env env = nullptr;
  std::map<uint64_t, JobCb *> cbs;
};
static void drop_cb(ServerHandle *h, uint64_t id) {
  auto it = h->cbs.find(id);
  if (it == h->cbs.end()) return;
  napi_delete_reference(h->env, it->second->fn);
  delete it->second;
  h->cbs.erase(it);
}
static void server_finalize(napi_env, void *data, void *) {
  ServerHandle *h = (ServerHandle *)data;
  if (h->srv) sl_server_destroy(h->srv);
  for (auto &kv : h->cbs) { napi_delete_reference(h->env, kv.second->fn); delete kv.second; }
  delete h;
}
static ServerHandle *get_server(napi_env env, napi_value v) {
  ServerHandle *h = nullptr;
  if (napi_get_value_external(env, v, (void **)&h) != napi_ok || !h || !h->srv) {
    napi_throw_error(env, nullptr, "suiri: invalid or destroyed server handle");
    return nullptr;
  }
  return h;
}
static napi_value throw_srv(napi_env env, const sl_server *s, int code) {
  std::string m = "suiri server error " + std::to_string(code) + ": " + sl_server_last_error(s);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}
// the job's source(dt, t) -> emission records, called on the JS thread inside serverTick
static uint32_t js_source(void *user, sl_sim *, double dt, double t, sl_emit *out, uint32_t max) {
  JobCb *cb = (JobCb *)user;
  napi_env env = cb->env;
  bool pending = false;
  napi_is_exception_pending(env, &pending);
  if (pending) return 0;
  napi_value fn, global, res, argv[2] = {num(env, dt), num(env, t)};
  if (napi_get_reference_value(env, cb->fn, &fn) != napi_ok || napi_get_global(env, &global) != napi_ok) return 0;
  if (napi_call_function(env, global, fn, 2, argv, &res) != napi_ok) return 0;
  std::vector<sl_emit> list;
  if (!read_emits(env, res, list)) return 0;
  const uint32_t n = list.size() < max ? (uint32_t)list.size() : max;
  memcpy(out, list.data(), n * sizeof(sl_emit));
  return n;
}

// serverCreate({budget, maxJobs, device, autotuneAfter, graphAfter}) -> server
static napi_value ServerCreate(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  size_t argc = 1;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  sl_server_config c;
  sl_server_config_default(&c);
  if (argc >= 1) {
    c.hbm_budget = (uint64_t)get_num(env, argv[0], "budget", (double)c.hbm_budget);
    c.max_jobs = (uint32_t)get_num(env, argv[0], "maxJobs", c.max_jobs);
    c.device = (int32_t)get_num(env, argv[0], "device", c.device);
    c.autotune_after = (uint32_t)get_num(env, argv[0], "autotuneAfter", c.autotune_after);
    c.graph_after = (uint32_t)get_num(env, argv[0], "graphAfter", c.graph_after);
    c.frames_per_tick = (uint32_t)get_num(env, argv[0], "framesPerTick", c.frames_per_tick);
  }
  sl_server *srv = nullptr;
  int r = sl_server_create(&c, &srv);
  if (r) return throw_sl(env, nullptr, r);
  ServerHandle *h = new ServerHandle();
  h->srv = srv;
  h->env = env;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, server_finalize, nullptr, &ext));
  return ext;
}

static bool read_items(napi_env env, napi_value list, std::vector<sl_item> &out);

// serverSubmit(srv, {config, objects | items, fill, steps, dt, source, keep, bakeDir, bakeEvery}) -> id
static napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  if (!h) return nullptr;
  napi_value o = argv[1], v;
  sl_job j;
  memset(&j, 0, sizeof j);
  sl_config_default(&j.cfg, "standard");
  bool has = false;
  napi_has_named_property(env, o, "config", &has);
  if (has) { napi_get_named_property(env, o, "config", &v); config_from_js(env, v, j.cfg); }
  std::vector<float> rows;
  napi_has_named_property(env, o, "objects", &has);
  if (has) {
    napi_get_named_property(env, o, "objects", &v);
    bool ta = false;
    napi_is_typedarray(env, v, &ta);
    if (ta) {
      napi_typedarray_type t;
      size_t n = 0;
      void *data = nullptr;
      napi_get_typedarray_info(env, v, &t, &n, &data, nullptr, nullptr);
      if (t != napi_float32_array || n % 28) { napi_throw_type_error(env, nullptr, "objects: Float32Array of 28-float rows"); return nullptr; }
      rows.assign((float *)data, (float *)data + n);
    }
  }
  std::vector<sl_item> items;
  napi_has_named_property(env, o, "items", &has);
  if (has) {
    napi_get_named_property(env, o, "items", &v);
    if (!read_items(env, v, items)) return nullptr;
  }
  std::vector<sl_emit> fill;
  napi_has_named_property(env, o, "fill", &has);
  if (has) {
    napi_get_named_property(env, o, "fill", &v);
    if (!read_emits(env, v, fill)) { napi_throw_type_error(env, nullptr, "fill: list of emission records"); return nullptr; }
  }
  j.objects = rows.data();
  j.n_objects = (uint32_t)(rows.size() / 28);
  j.fill = fill.data();
  j.n_fill = (uint32_t)fill.size();
  j.items = items.data();
  j.n_items = (uint32_t)items.size();
  j.steps = (uint32_t)get_num(env, o, "steps", 100);
  j.dt = get_num(env, o, "dt", 1.0 / 120);
  j.keep = get_num(env, o, "keep", 0) != 0 ? 1u : 0u;
  j.bake_every = (uint32_t)get_num(env, o, "bakeEvery", 0);
  char dir[4096] = {0};
  napi_has_named_property(env, o, "bakeDir", &has);
  if (has) {
    napi_get_named_property(env, o, "bakeDir", &v);
    size_t len = 0;
    if (napi_get_value_string_utf8(env, v, dir, sizeof dir, &len) == napi_ok && len) j.bake_dir = dir;
  }
  JobCb *cb = nullptr;
  napi_has_named_property(env, o, "source", &has);
  if (has) {
    napi_get_named_property(env, o, "source", &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_function) {
      cb = new JobCb{env, nullptr};
      napi_create_reference(env, v, 1, &cb->fn);
      j.source = js_source;
      j.user = cb;
    }
  }
  uint64_t id = 0;
  int r = sl_server_submit(h->srv, &j, &id);
  if (cb) h->cbs[id] = cb;
  if (r && r != SL_EBUDGET) return throw_srv(env, h->srv, r);
  return num(env, (double)id);  // a refused job is reported by serverPoll
}

static napi_value info_to_js(napi_env env, const sl_job_info &i) {
  static const char *st[] = {"queued", "running", "done", "cancelled", "refused"};
  napi_value o, s;
  napi_create_object(env, &o);
  set_num(env, o, "id", (double)i.id);
  napi_create_string_utf8(env, i.status >= 0 && i.status < 5 ? st[i.status] : "?", NAPI_AUTO_LENGTH, &s);
  napi_set_named_property(env, o, "status", s);
  set_num(env, o, "frames", i.frames);
  set_num(env, o, "steps", i.steps);
  set_num(env, o, "tSubmit", i.t_submit);
  set_num(env, o, "tAdmit", i.t_admit);
  set_num(env, o, "tDone", i.t_done);
  set_num(env, o, "particleSteps", i.particle_steps);
  set_num(env, o, "bytes", (double)i.bytes);
  set_num(env, o, "count", i.count);
  set_num(env, o, "wwCount", i.ww_count);
  return o;
}

// serverTick(srv) -> number of jobs stepped
static napi_value ServerTick(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  if (!h) return nullptr;
  uint32_t running = 0;
  int r = sl_server_tick(h->srv, &running);
  bool pending = false;
  napi_is_exception_pending(env, &pending);
  if (pending) return nullptr;  // a source threw
  if (r) return throw_srv(env, h->srv, r);
  return num(env, running);
}

// serverPoll(srv) -> [info] of the jobs retired since the last poll
static napi_value ServerPoll(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  if (!h) return nullptr;
  napi_value arr;
  napi_create_array(env, &arr);
  sl_job_info buf[64];
  uint32_t k = 0, n = 0;
  do {
    sl_server_poll(h->srv, buf, 64, &n);
    for (uint32_t i = 0; i < n; i++) {
      napi_set_element(env, arr, k++, info_to_js(env, buf[i]));
      drop_cb(h, buf[i].id);
    }
  } while (n == 64);
  return arr;
}

static bool get_id(napi_env env, napi_value v, uint64_t *id) {
  double d = 0;
  if (napi_get_value_double(env, v, &d) != napi_ok) return false;
  *id = (uint64_t)d;
  return true;
}
// serverInfo(srv, id) / serverCancel(srv, id) / serverRelease(srv, id) / serverJobSim(srv, id)
static napi_value ServerInfo(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  uint64_t id;
  if (!h || !get_id(env, argv[1], &id)) return nullptr;
  sl_job_info i;
  if (sl_server_info(h->srv, id, &i)) { napi_throw_error(env, nullptr, "serverInfo: unknown job"); return nullptr; }
  return info_to_js(env, i);
}
static napi_value ServerCancel(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  uint64_t id;
  if (!h || !get_id(env, argv[1], &id)) return nullptr;
  int r = sl_server_cancel(h->srv, id);
  if (r) return throw_srv(env, h->srv, r);
  return nullptr;
}
static napi_value ServerRelease(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  uint64_t id;
  if (!h || !get_id(env, argv[1], &id)) return nullptr;
  int r = sl_server_release(h->srv, id);
  if (r) return throw_srv(env, h->srv, r);
  return nullptr;
}
static napi_value ServerJobSim(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  uint64_t id;
  if (!h || !get_id(env, argv[1], &id)) return nullptr;
  sl_sim *s = sl_server_job_sim(h->srv, id);
  napi_value r;
  if (!s) { napi_get_null(env, &r); return r; }
  Handle *hh = new Handle();
  hh->sim = s;
  hh->owned = false;
  NAPI_CALL(env, napi_create_external(env, hh, finalize, nullptr, &r));
  return r;
}
// serverStats(srv) -> object
static napi_value ServerStats(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  ServerHandle *h = get_server(env, argv[0]);
  if (!h) return nullptr;
  sl_server_stats_t st;
  sl_server_stats(h->srv, &st);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "budget", (double)st.budget);
  set_num(env, o, "used", (double)st.used);
  set_num(env, o, "queued", st.queued);
  set_num(env, o, "running", st.running);
  set_num(env, o, "resident", st.resident);
  set_num(env, o, "done", st.done);
  set_num(env, o, "ticks", (double)st.ticks);
  set_num(env, o, "frames", (double)st.frames);
  set_num(env, o, "particleSteps", st.particle_steps);
  set_num(env, o, "uptime", st.uptime);
  set_num(env, o, "frameMsP50", st.step_ms_p50);
  set_num(env, o, "frameMsP99", st.step_ms_p99);
  set_num(env, o, "latencyP50", st.latency_p50);
  set_num(env, o, "latencyP99", st.latency_p99);
  return o;
}
static napi_value ServerDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  ServerHandle *h = nullptr;
  napi_get_value_external(env, argv[0], (void **)&h);
  if (h && h->srv) {
    sl_server_destroy(h->srv);
    h->srv = nullptr;
    for (auto &kv : h->cbs) { napi_delete_reference(h->env, kv.second->fn); delete kv.second; }
    h->cbs.clear();
  }
  return nullptr;
}

// ---- scene items (sl_item) ----
// JS items of the reference's scene model: {id, kind, shape | mode, pos, rot, dims | r, enabled, and
// the source fields}. Names map to the ids of include/suiri.h.
static const char *const kKinds[] = {"obstacle", "drain", "source", "block"};
static const char *const kShapes[] = {"box", "plate", "sphere", "cylinder", "cone", "torus", "cup", "container", "gutter"};
static const char *const kModes[] = {"stream", "drops", "shower", "burst", "rain"};
// dims keys of each obstacle shape, in sl_item.dims order
static const char *const kShapeDims[][4] = {
    {"w", "h", "d", nullptr}, {"w", "d", "t", nullptr}, {"r", nullptr, nullptr, nullptr}, {"r", "h", nullptr, nullptr},
    {"r", "h", nullptr, nullptr}, {"R", "r", nullptr, nullptr}, {"r", "h", "t", nullptr}, {"w", "h", "d", "t"},
    {"len", "r", "t", nullptr}};

static std::string get_str(napi_env env, napi_value obj, const char *key) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) return "";
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  char buf[64] = {0};
  size_t len = 0;
  if (napi_get_value_string_utf8(env, v, buf, sizeof buf, &len) != napi_ok) return "";
  return std::string(buf, len);
}
static int name_index(const std::string &s, const char *const *names, int n) {
  for (int i = 0; i < n; i++) if (s == names[i]) return i;
  return -1;
}
static void get_vec3(napi_env env, napi_value obj, const char *key, double out[3]) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) return;
  napi_value arr;
  napi_get_named_property(env, obj, key, &arr);
  bool isarr = false;
  napi_is_array(env, arr, &isarr);
  if (!isarr) return;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e;
    if (napi_get_element(env, arr, i, &e) == napi_ok) napi_get_value_double(env, e, &out[i]);
  }
}
static bool item_from_js(napi_env env, napi_value it, sl_item &o, std::string &err) {
  memset(&o, 0, sizeof o);
  const std::string kind = get_str(env, it, "kind");
  const int k = name_index(kind, kKinds, 4);
  if (k < 0) { err = "unknown item kind '" + kind + "'"; return false; }
  o.kind = (uint32_t)k;
  o.id = (uint32_t)(int64_t)get_num(env, it, "id", 0);
  get_vec3(env, it, "pos", o.pos);
  get_vec3(env, it, "rot", o.rot);
  napi_value dims = nullptr;
  bool has = false;
  napi_has_named_property(env, it, "dims", &has);
  if (has) napi_get_named_property(env, it, "dims", &dims);
  if (o.kind == SL_ITEM_OBSTACLE) {
    const std::string shape = get_str(env, it, "shape");
    const int sh = name_index(shape, kShapes, SL_SHAPE_COUNT_);
    if (sh < 0) { err = "unknown obstacle shape '" + shape + "'"; return false; }
    o.shape = (uint32_t)sh;
    for (int q = 0; q < 4 && kShapeDims[sh][q]; q++) o.dims[q] = dims ? get_num(env, dims, kShapeDims[sh][q], 0) : 0;
  } else if (o.kind == SL_ITEM_BLOCK) {
    const char *keys[3] = {"w", "h", "d"};
    for (int q = 0; q < 3; q++) o.dims[q] = dims ? get_num(env, dims, keys[q], 0) : 0;
  } else if (o.kind == SL_ITEM_DRAIN) {
    o.dims[0] = get_num(env, it, "r", 0.025);
  } else {
    const std::string mode = get_str(env, it, "mode");
    const int m = name_index(mode, kModes, SL_MODE_COUNT_);
    if (m < 0) { err = "unknown source mode '" + mode + "'"; return false; }
    o.shape = (uint32_t)m;
    bool en = true;
    napi_has_named_property(env, it, "enabled", &has);
    if (has) { napi_value v; napi_get_named_property(env, it, "enabled", &v); napi_get_value_bool(env, v, &en); }
    o.enabled = en ? 1u : 0u;
    o.pitch = get_num(env, it, "pitch", 0); o.yaw = get_num(env, it, "yaw", 0);
    o.flow = get_num(env, it, "flow", 0); o.speed = get_num(env, it, "speed", 0);
    o.drop_d = get_num(env, it, "dropD", 0); o.drop_rate = get_num(env, it, "dropRate", 0);
    o.spread = get_num(env, it, "spread", 0); o.burst = get_num(env, it, "burst", 0);
    o.burst_every = get_num(env, it, "burstEvery", 0); o.rain_rate = get_num(env, it, "rainRate", 0);
    o.rain_size = get_num(env, it, "rainSize", 0); o.total = get_num(env, it, "total", 0);
  }
  return true;
}
static bool read_items(napi_env env, napi_value list, std::vector<sl_item> &out) {
  bool isarr = false;
  napi_is_array(env, list, &isarr);
  if (!isarr) { napi_throw_type_error(env, nullptr, "items: an array of scene items"); return false; }
  uint32_t n = 0;
  napi_get_array_length(env, list, &n);
  out.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value it;
    napi_get_element(env, list, i, &it);
    std::string err;
    if (!item_from_js(env, it, out[i], err)) { napi_throw_type_error(env, nullptr, err.c_str()); return false; }
  }
  return true;
}
static napi_value vec_js(napi_env env, const float *v) {
  napi_value a;
  napi_create_array_with_length(env, 4, &a);
  for (uint32_t i = 0; i < 4; i++) napi_set_element(env, a, i, num(env, v[i]));
  return a;
}
static napi_value records_js(napi_env env, const sl_emit *e, uint32_t n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t i = 0; i < n; i++) {
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "shape", e[i].shape);
    napi_set_named_property(env, o, "a", vec_js(env, e[i].a));
    napi_set_named_property(env, o, "b", vec_js(env, e[i].b));
    napi_set_named_property(env, o, "c", vec_js(env, e[i].c));
    napi_set_named_property(env, o, "d", vec_js(env, e[i].d));
    set_num(env, o, "count", e[i].count);
    set_num(env, o, "seed", e[i].seed);
    napi_set_element(env, arr, i, o);
  }
  return arr;
}

// a scene handle: a standalone sl_scene (owned) or the one a simulation keeps (borrowed)
struct SceneHandle {
  sl_scene *sc = nullptr;
  bool owned = true;
  Handle *sim = nullptr;       // the simulation of a borrowed scene (its scene dies with it)
  napi_ref sim_ref = nullptr;  // keeps that simulation's handle alive
  napi_env env = nullptr;
};
static void scene_finalize(napi_env, void *data, void *) {
  SceneHandle *h = (SceneHandle *)data;
  if (h->sc && h->owned) sl_scene_destroy(h->sc);
  if (h->sim_ref) napi_delete_reference(h->env, h->sim_ref);
  delete h;
}
static sl_scene *get_scene(napi_env env, napi_value v) {
  SceneHandle *h = nullptr;
  if (napi_get_value_external(env, v, (void **)&h) != napi_ok || !h || !h->sc || (!h->owned && (!h->sim || !h->sim->sim))) {
    napi_throw_error(env, nullptr, "suiri: invalid scene handle (or its simulation was destroyed)");
    return nullptr;
  }
  return h->sc;
}
static napi_value throw_scene(napi_env env, int code, const char *what) {
  std::string m = std::string("suiri scene error ") + std::to_string(code) + ": " + what;
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

// sceneCreate(config) -> scene (host only: no device is touched)
static napi_value SceneCreate(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_config c;
  sl_config_default(&c, "standard");
  config_from_js(env, argv[0], c);
  sl_scene *sc = nullptr;
  int r = sl_scene_create(&c, &sc);
  if (r) return throw_scene(env, r, "sceneCreate: bad configuration");
  SceneHandle *h = new SceneHandle();
  h->sc = sc;
  h->env = env;
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, scene_finalize, nullptr, &ext));
  return ext;
}
// simScene(h) -> the scene state the simulation keeps
static napi_value SimScene(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  SceneHandle *h = new SceneHandle();
  h->sc = sl_sim_scene(s);
  h->owned = false;
  napi_get_value_external(env, argv[0], (void **)&h->sim);
  h->env = env;
  napi_create_reference(env, argv[0], 1, &h->sim_ref);
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, scene_finalize, nullptr, &ext));
  return ext;
}
// packObjects(scene, items, dtMove) -> {rows: Float32Array, n, moving}
static napi_value PackObjects(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_scene *sc = get_scene(env, argv[0]);
  if (!sc) return nullptr;
  std::vector<sl_item> items;
  if (!read_items(env, argv[1], items)) return nullptr;
  double dtm = 0;
  napi_get_value_double(env, argv[2], &dtm);
  float rows[SL_MAX_OBJ * 28];
  uint32_t n = 0;
  int moving = 0;
  int r = sl_scene_pack_objects(sc, items.data(), (uint32_t)items.size(), dtm, rows, &n, &moving);
  if (r) return throw_scene(env, r, "packObjects");
  napi_value ab, ta, o;
  void *data = nullptr;
  NAPI_CALL(env, napi_create_arraybuffer(env, (size_t)n * 28 * 4, &data, &ab));
  if (n) memcpy(data, rows, (size_t)n * 28 * 4);
  NAPI_CALL(env, napi_create_typedarray(env, napi_float32_array, (size_t)n * 28, ab, 0, &ta));
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "rows", ta);
  set_num(env, o, "n", n);
  napi_set_named_property(env, o, "moving", boolean(env, moving != 0));
  return o;
}
// setItems(h, items, dtMove) -> moving (FluidSim.setObjects)
static napi_value SetItems(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  std::vector<sl_item> items;
  if (!read_items(env, argv[1], items)) return nullptr;
  double dtm = 0;
  napi_get_value_double(env, argv[2], &dtm);
  int moving = 0;
  int r = sl_set_items(s, items.data(), (uint32_t)items.size(), dtm, &moving);
  if (r) return throw_sl(env, s, r);
  return boolean(env, moving != 0);
}
// sourceEmits(scene, items, dt) -> records (sourceEmits, 55-emit.js:13-69)
static napi_value SourceEmits(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_scene *sc = get_scene(env, argv[0]);
  if (!sc) return nullptr;
  std::vector<sl_item> items;
  if (!read_items(env, argv[1], items)) return nullptr;
  double dt = 0;
  napi_get_value_double(env, argv[2], &dt);
  std::vector<sl_emit> out(items.size() + 1);
  uint32_t n = 0;
  int r = sl_scene_source_emits(sc, items.data(), (uint32_t)items.size(), dt, out.data(), (uint32_t)out.size(), &n);
  if (r) return throw_scene(env, r, "sourceEmits");
  return records_js(env, out.data(), n);
}
// blockEmits(scene, items, fill) -> records (blockEmits, 55-emit.js:83-106)
static napi_value BlockEmits(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  size_t argc = 3;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  if (argc < 2) { napi_throw_type_error(env, nullptr, "blockEmits(scene, items, fill)"); return nullptr; }
  sl_scene *sc = get_scene(env, argv[0]);
  if (!sc) return nullptr;
  std::vector<sl_item> items;
  if (!read_items(env, argv[1], items)) return nullptr;
  bool fill = false;
  if (argc >= 3) napi_get_value_bool(env, argv[2], &fill);
  std::vector<sl_emit> out(items.size() + 1);
  uint32_t n = 0;
  int r = sl_scene_block_emits(sc, items.data(), (uint32_t)items.size(), fill ? 1 : 0, out.data(),
                               (uint32_t)out.size(), &n);
  if (r) return throw_scene(env, r, "blockEmits");
  return records_js(env, out.data(), n);
}
// blockLiters(block, tank{w, h, d}) -> litres inside the tank (55-emit.js:108-113)
static napi_value BlockLiters(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  sl_item it;
  std::string err;
  if (!item_from_js(env, argv[0], it, err)) { napi_throw_type_error(env, nullptr, err.c_str()); return nullptr; }
  return num(env, sl_block_liters(&it, get_num(env, argv[1], "w", 0), get_num(env, argv[1], "h", 0),
                                  get_num(env, argv[1], "d", 0)));
}
// stepItems(h, frameDt, items): FluidSim.step with the sources of `items` as the emitter
static napi_value StepItems(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  sl_sim *s = get_sim(env, argv[0]);
  if (!s) return nullptr;
  double fdt = 0;
  napi_get_value_double(env, argv[1], &fdt);
  std::vector<sl_item> items;
  if (!read_items(env, argv[2], items)) return nullptr;
  int r = sl_step_items(s, fdt, items.data(), (uint32_t)items.size());
  if (r) return throw_sl(env, s, r);
  return nullptr;
}

static napi_value Version(napi_env env, napi_callback_info) { return num(env, sl_version()); }
static napi_value BuildInfo(napi_env env, napi_callback_info) {
  napi_value v;
  napi_create_string_utf8(env, sl_build_info(), NAPI_AUTO_LENGTH, &v);
  return v;
}

static napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"configDefault", 0, ConfigDefault, 0, 0, 0, napi_default, 0},
      {"create", 0, Create, 0, 0, 0, napi_default, 0},
      {"destroy", 0, Destroy, 0, 0, 0, napi_default, 0},
      {"gridInfo", 0, GridInfo, 0, 0, 0, napi_default, 0},
      {"setObjects", 0, SetObjects, 0, 0, 0, napi_default, 0},
      {"emitNow", 0, EmitNow, 0, 0, 0, napi_default, 0},
      {"stepBegin", 0, StepBegin, 0, 0, 0, napi_default, 0},
      {"substep", 0, Substep, 0, 0, 0, napi_default, 0},
      {"stepEnd", 0, StepEnd, 0, 0, 0, napi_default, 0},
      {"setSubsteps", 0, SetSubsteps, 0, 0, 0, napi_default, 0},
      {"setPressureSolver", 0, SetPressureSolver, 0, 0, 0, napi_default, 0},
      {"pressureSolverStats", 0, PressureSolverStats, 0, 0, 0, napi_default, 0},
      {"clearWater", 0, ClearWater, 0, 0, 0, napi_default, 0},
      {"saveWater", 0, SaveWater, 0, 0, 0, napi_default, 0},
      {"hasWaterCopy", 0, HasWaterCopy, 0, 0, 0, napi_default, 0},
      {"restoreWater", 0, RestoreWater, 0, 0, 0, napi_default, 0},
      {"pollStats", 0, PollStats, 0, 0, 0, napi_default, 0},
      {"stats", 0, Stats, 0, 0, 0, napi_default, 0},
      {"statsAsync", 0, StatsAsync, 0, 0, 0, napi_default, 0},
      {"hostState", 0, HostState, 0, 0, 0, napi_default, 0},
      {"setScalar", 0, SetScalar, 0, 0, 0, napi_default, 0},
      {"sync", 0, Sync, 0, 0, 0, napi_default, 0},
      {"reorder", 0, Reorder, 0, 0, 0, napi_default, 0},
      {"setRippleMask", 0, SetRippleMask, 0, 0, 0, napi_default, 0},
      {"readback", 0, Readback, 0, 0, 0, napi_default, 0},
      {"upload", 0, Upload, 0, 0, 0, napi_default, 0},
      {"version", 0, Version, 0, 0, 0, napi_default, 0},
      {"buildInfo", 0, BuildInfo, 0, 0, 0, napi_default, 0},
      {"bakeFrame", 0, BakeFrame, 0, 0, 0, napi_default, 0},
      {"jobBytes", 0, JobBytes, 0, 0, 0, napi_default, 0},
      {"serverCreate", 0, ServerCreate, 0, 0, 0, napi_default, 0},
      {"serverSubmit", 0, ServerSubmit, 0, 0, 0, napi_default, 0},
      {"serverTick", 0, ServerTick, 0, 0, 0, napi_default, 0},
      {"serverPoll", 0, ServerPoll, 0, 0, 0, napi_default, 0},
      {"serverInfo", 0, ServerInfo, 0, 0, 0, napi_default, 0},
      {"serverCancel", 0, ServerCancel, 0, 0, 0, napi_default, 0},
      {"serverRelease", 0, ServerRelease, 0, 0, 0, napi_default, 0},
      {"serverJobSim", 0, ServerJobSim, 0, 0, 0, napi_default, 0},
      {"serverStats", 0, ServerStats, 0, 0, 0, napi_default, 0},
      {"serverDestroy", 0, ServerDestroy, 0, 0, 0, napi_default, 0},
      {"sceneCreate", 0, SceneCreate, 0, 0, 0, napi_default, 0},
      {"simScene", 0, SimScene, 0, 0, 0, napi_default, 0},
      {"packObjects", 0, PackObjects, 0, 0, 0, napi_default, 0},
      {"setItems", 0, SetItems, 0, 0, 0, napi_default, 0},
      {"sourceEmits", 0, SourceEmits, 0, 0, 0, napi_default, 0},
      {"blockEmits", 0, BlockEmits, 0, 0, 0, napi_default, 0},
      {"blockLiters", 0, BlockLiters, 0, 0, 0, napi_default, 0},
      {"stepItems", 0, StepItems, 0, 0, 0, napi_default, 0},
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
