// Scene items -> device records on the host: obstacle and drain packing with velocities from the
// previous poses (FluidSim.setObjects, src/js/50-sim.js:165-211) and the emitters that turn water
// sources and blocks into emission records (src/js/55-emit.js:13-113). Plain C++ over the C ABI of
// include/suiri.h; every host (Node addon, Python binding, serving loop) calls these functions, so
// the packing and the emission lists have one implementation.
//
// Arithmetic is in double precision, like the reference's JavaScript, and each record field is
// rounded once to f32 where the reference writes its Float32Array.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <new>

#include "../../include/suiri.h"
#include "sl_scene.h"

namespace {

const double kPi = 3.14159265358979323846;
const double kDeg = kPi / 180.0;
const double kRainSpeed = 6.5;  // fall speed of rain [m/s] (55-emit.js:10)

double clampd(double x, double lo, double hi) { return std::min(hi, std::max(lo, x)); }
// JavaScript Math.round: halves round up
double js_round(double x) { return floor(x + 0.5); }

struct Vec {
  double x, y, z;
};
Vec operator+(Vec a, Vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec operator-(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec operator*(Vec a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double length(Vec a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
Vec unit_or_self(Vec a) {  // a / |a|, or a when |a| = 0
  const double l = length(a);
  return l > 0 ? a * (1.0 / l) : a;
}
Vec cross(Vec a, Vec b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec vec_of(const double *p) { return {p[0], p[1], p[2]}; }

// Row-major rotation from Euler angles in degrees, applied about x, then y, then z: R = Rz Ry Rx
// (the convention of the reference's editor, 00-util.js:22-30).
void euler_xyz(const double deg[3], double R[9]) {
  const double cx = cos(deg[0] * kDeg), sx = sin(deg[0] * kDeg);
  const double cy = cos(deg[1] * kDeg), sy = sin(deg[1] * kDeg);
  const double cz = cos(deg[2] * kDeg), sz = sin(deg[2] * kDeg);
  // columns of Rz Ry Rx applied to the unit vectors
  R[0] = cz * cy; R[1] = cz * sy * sx - sz * cx; R[2] = cz * sy * cx + sz * sx;
  R[3] = sz * cy; R[4] = sz * sy * sx + cz * cx; R[5] = sz * sy * cx - cz * sx;
  R[6] = -sy;     R[7] = cy * sx;                R[8] = cy * cx;
}
// rotation vector (axis times angle) of the rotation A B^T, the turn from pose B to pose A
Vec relative_rotation(const double A[9], const double B[9]) {
  double M[9];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) M[3 * r + c] = A[3 * r] * B[3 * c] + A[3 * r + 1] * B[3 * c + 1] + A[3 * r + 2] * B[3 * c + 2];
  const double cosang = clampd((M[0] + M[4] + M[8] - 1) / 2, -1, 1);
  const double ang = acos(cosang);
  const double s2 = 2 * sin(ang);
  // no axis to speak of for a vanishing angle, nor a stable one for a half turn (00-util.js:53-60)
  if (ang < 1e-6 || fabs(s2) < 1e-6) return {0, 0, 0};
  return Vec{(M[7] - M[5]) / s2, (M[2] - M[6]) / s2, (M[3] - M[1]) / s2} * ang;
}
Vec limit_length(Vec v, double cap) {
  const double l = length(v);
  return l > cap ? unit_or_self(v) * cap : v;
}

// The signed-distance record of an obstacle: the shape type that the device's shape_sdf switches on
// and its four dimensions (the table of 10-scene.js:143-156, a contract with the SDF code).
struct ShapeRec {
  int type;
  double d[4];
};
ShapeRec shape_record(const sl_item &it) {
  const double *x = it.dims;
  switch (it.shape) {
    case SL_SHAPE_BOX:  // half extents, edges rounded by 3 mm at most (a quarter of the thinnest side)
      return {0, {x[0] / 2, x[1] / 2, x[2] / 2, std::min(std::min(0.003, x[0] / 4), std::min(x[1] / 4, x[2] / 4))}};
    case SL_SHAPE_PLATE:  // (w, d, t): a box t thick, edges rounded by 2 mm at most
      return {0, {x[0] / 2, x[2] / 2, x[1] / 2, std::min(0.002, x[2] / 4)}};
    case SL_SHAPE_SPHERE: return {1, {x[0], 0, 0, 0}};
    case SL_SHAPE_CYLINDER: return {2, {x[0], x[1] / 2, 0, 0}};
    case SL_SHAPE_CONE: return {3, {x[0], x[1] / 2, 0, 0}};
    case SL_SHAPE_TORUS: return {4, {x[0], x[1], 0, 0}};
    case SL_SHAPE_CUP:  // wall thickness at most 0.8 r
      return {5, {x[0], x[1] / 2, std::min(x[2], x[0] * 0.8), 0}};
    case SL_SHAPE_CONTAINER:  // wall thickness at most a third of the footprint and 0.8 of the height
      return {6, {x[0] / 2, x[1] / 2, x[2] / 2, std::min(std::min(x[3], x[0] / 3), std::min(x[2] / 3, x[1] * 0.8))}};
    case SL_SHAPE_GUTTER: return {7, {x[0] / 2, x[1], x[2], 0}};
    default: return {0, {0.05, 0.05, 0.05, 0}};
  }
}
// radius of a sphere about the obstacle's position that holds its solid (10-scene.js:177-188)
double bound_radius(const ShapeRec &s) {
  const double *d = s.d;
  switch (s.type) {
    case 0: case 6: return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    case 1: return d[0];
    case 2: case 3: case 5: return hypot(d[0], d[1]);
    case 4: return d[0] + d[1];
    case 7: return hypot(d[0], d[1] + d[2]);
    default: return 0.1;
  }
}

// outlet axis of a source from its pitch (up from horizontal) and yaw (about y from +z)
Vec source_axis(const sl_item &s) {
  const double p = s.pitch * kDeg, y = s.yaw * kDeg;
  return {cos(p) * sin(y), sin(p), cos(p) * cos(y)};
}
// two unit vectors across the axis: u = axis x ref (ref = +y unless the axis is nearly vertical,
// then +x), w = axis x u (55-emit.js:4-9)
void cross_basis(Vec axis, Vec &u, Vec &w) {
  const Vec ref = fabs(axis.y) < 0.95 ? Vec{0, 1, 0} : Vec{1, 0, 0};
  u = unit_or_self(cross(axis, ref));
  w = cross(axis, u);
}
// diameter of the outlet that carries `flow` L/min at `speed` (10-scene.js:91-95)
double nozzle_diameter(const sl_item &s) {
  const double q = s.flow / 60000.0;
  return sqrt(4 * q / (kPi * std::max(s.speed, 0.05)));
}

uint32_t next_seed(sl_scene &sc) {
  // a PCG-style state advance and an output hash (lowbias32)
  sc.seed_state = sc.seed_state * 747796405u + 2891336453u;
  uint32_t x = sc.seed_state;
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

void set4(float *dst, double a, double b, double c, double d) {
  dst[0] = (float)a; dst[1] = (float)b; dst[2] = (float)c; dst[3] = (float)d;
}

// the record emitted by one release of a burst source: a ball of water just in front of the outlet
// (55-emit.js:72-79)
sl_emit burst_record(const sl_scene &sc, const sl_item &s) {
  const Vec axis = source_axis(s);
  const double vol = clampd(s.burst, 0.01, 50) / 1000;
  const double radius = cbrt(3 * vol / (4 * kPi));
  const Vec at = vec_of(s.pos) + axis * (radius + 0.004);
  const Vec vel = axis * s.speed;
  sl_emit e;
  memset(&e, 0, sizeof e);
  e.shape = 1;
  e.count = (uint32_t)std::max(1.0, js_round(vol / sc.vp));
  set4(e.a, at.x, at.y, at.z, radius);
  set4(e.b, vel.x, vel.y, vel.z, 0);
  return e;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// scene state
// ------------------------------------------------------------------------------------------------
int sl_scene_init(sl_scene &sc, const sl_config &cfg) {
  int n[3];
  double h, o[3];
  int r = sl_grid_of(cfg, n, h, o);
  if (r) return r;
  sc.h = h;
  sc.origin[0] = o[0]; sc.origin[1] = o[1]; sc.origin[2] = o[2];
  sc.vp = pow(h / 2, 3);  // particle volume (50-sim.js:44)
  sc.tank[0] = cfg.tank_w; sc.tank[1] = cfg.tank_h; sc.tank[2] = cfg.tank_d;
  sc.sources.clear();
  sc.poses.clear();
  sc.seed_state = cfg.seed;
  return SL_OK;
}

extern "C" int sl_scene_create(const sl_config *cfg, sl_scene **out) {
  if (!cfg || !out) return SL_EARG;
  *out = nullptr;
  sl_scene *sc = new (std::nothrow) sl_scene();
  if (!sc) return SL_ENOMEM;
  int r = sl_scene_init(*sc, *cfg);
  if (r) { delete sc; return r; }
  *out = sc;
  return SL_OK;
}
extern "C" void sl_scene_destroy(sl_scene *sc) { delete sc; }
extern "C" const char *sl_scene_last_error(const sl_scene *sc) { return sc ? sc->err.c_str() : ""; }

// ------------------------------------------------------------------------------------------------
// objects, 50-sim.js:165-201
// ------------------------------------------------------------------------------------------------
extern "C" int sl_scene_pack_objects(sl_scene *sc, const sl_item *items, uint32_t n, double dt_move, float *rows,
                                     uint32_t *n_rows, int *moving) {
  if (!sc || !rows || (n && !items)) return SL_EARG;
  memset(rows, 0, sizeof(float) * SL_MAX_OBJ * 28);
  std::map<uint32_t, SlPose> poses;
  uint32_t k = 0;
  bool any_moving = false;
  for (uint32_t i = 0; i < n && k < SL_MAX_OBJ; i++) {
    const sl_item &it = items[i];
    if (it.kind != SL_ITEM_OBSTACLE && it.kind != SL_ITEM_DRAIN) continue;
    SlPose pose;
    memcpy(pose.pos, it.pos, sizeof pose.pos);
    euler_xyz(it.rot, pose.R);
    int type;
    double dims[4], bound, flag;
    if (it.kind == SL_ITEM_OBSTACLE) {
      const ShapeRec s = shape_record(it);
      type = s.type;
      memcpy(dims, s.d, sizeof dims);
      bound = bound_radius(s) + 0.002;
      flag = 1;
    } else {  // a drain: a cylinder of radius r, 12 mm high, shape id 9 (no distance of its own)
      type = 9;
      dims[0] = it.dims[0]; dims[1] = 0.012; dims[2] = 0; dims[3] = 0;
      bound = it.dims[0] + 0.02;
      flag = 2;
    }
    Vec lin{0, 0, 0}, ang{0, 0, 0};
    auto prev = sc->poses.find(it.id);
    if (prev != sc->poses.end() && dt_move > 0) {
      lin = (vec_of(pose.pos) - vec_of(prev->second.pos)) * (1 / dt_move);
      ang = relative_rotation(pose.R, prev->second.R) * (1 / dt_move);
      if (length(lin) > 1e-5 || length(ang) > 1e-4) any_moving = true;
      // a typed-in jump is a teleport, not a motion that should fling the water
      lin = limit_length(lin, 4.0);
      ang = limit_length(ang, 40.0);
    }
    poses[it.id] = pose;
    float *o = rows + 28 * k;
    const double *R = pose.R;
    set4(o + 0, R[0], R[3], R[6], 0);  // rows of R^T: world -> local
    set4(o + 4, R[1], R[4], R[7], 0);
    set4(o + 8, R[2], R[5], R[8], 0);
    set4(o + 12, it.pos[0], it.pos[1], it.pos[2], type);
    set4(o + 16, dims[0], dims[1], dims[2], dims[3]);
    set4(o + 20, lin.x, lin.y, lin.z, bound);
    set4(o + 24, ang.x, ang.y, ang.z, flag);
    k++;
  }
  sc->poses.swap(poses);
  if (n_rows) *n_rows = k;
  if (moving) *moving = any_moving ? 1 : 0;
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// sources, 55-emit.js:13-69
// ------------------------------------------------------------------------------------------------
extern "C" int sl_scene_source_emits(sl_scene *sc, const sl_item *items, uint32_t n, double dt, sl_emit *out,
                                     uint32_t max, uint32_t *n_out) {
  if (!sc || (n && !items) || (max && !out)) return SL_EARG;
  uint32_t k = 0;
  const double vp = sc->vp;
  for (uint32_t i = 0; i < n; i++) {
    const sl_item &s = items[i];
    if (s.kind != SL_ITEM_SOURCE || !s.enabled) continue;
    SlSourceState &st = sc->sources[s.id];
    if (s.total > 0 && st.liters >= s.total) continue;  // the source has run dry
    const Vec axis = source_axis(s);
    sl_emit e;
    memset(&e, 0, sizeof e);
    switch (s.shape) {
      case SL_MODE_STREAM:
      case SL_MODE_SHOWER: {
        // a disc outlet releasing flow * dt of water per substep; the fraction of a particle carries over
        st.carry += s.flow / 60000.0 * dt / vp;
        const double whole = floor(st.carry);
        st.carry -= whole;
        if (whole <= 0) continue;
        const bool shower = s.shape == SL_MODE_SHOWER;
        const double radius = std::max(nozzle_diameter(s) / 2, shower ? 0.022 : 0.0005);
        const double cone = shower ? clampd(s.spread, 0, 80) * kDeg : 0.0;
        Vec u, w;
        cross_basis(axis, u, w);
        const Vec at = vec_of(s.pos) + axis * 0.002;  // just outside the nozzle
        e.shape = 0;
        e.count = (uint32_t)whole;
        set4(e.a, at.x, at.y, at.z, radius);
        set4(e.b, axis.x, axis.y, axis.z, s.speed);
        set4(e.c, u.x, u.y, u.z, cone);
        set4(e.d, w.x, w.y, w.z, s.speed * dt);
        break;
      }
      case SL_MODE_DROPS: {
        // one drop every 1 / rate seconds
        st.wait -= dt;
        if (st.wait > 0) continue;
        st.wait = std::max(0.0, st.wait + 1 / clampd(s.drop_rate, 0.05, 60));
        const double dia = clampd(s.drop_d, 1, 40) / 1000;
        const double parts = std::max(1.0, js_round(kPi / 6 * dia * dia * dia / vp));
        const double radius = std::max(dia / 2, 0.9 * cbrt(3 * parts * vp / (4 * kPi)));
        const Vec at = vec_of(s.pos) + axis * (radius + 0.002);
        const Vec vel = axis * s.speed;
        e.shape = 1;
        e.count = (uint32_t)parts;
        // a one-particle drop sits at the centre; drops of up to 8 particles start in free flight
        set4(e.a, at.x, at.y, at.z, parts == 1 ? 0.0 : radius);
        set4(e.b, vel.x, vel.y, vel.z, parts <= 8 ? 1.0 : 0.0);
        break;
      }
      case SL_MODE_BURST: {
        // once, or every burst_every seconds
        if (st.fired) {
          if (!(s.burst_every > 0)) continue;
          st.wait -= dt;
          if (st.wait > 0) continue;
        }
        st.fired = true;
        st.wait = s.burst_every > 0 ? s.burst_every : 0;
        e = burst_record(*sc, s);
        break;
      }
      case SL_MODE_RAIN: {
        // rate [mm/h] over a square of the given side
        const double side = clampd(s.rain_size, 0.05, 2);
        st.carry += clampd(s.rain_rate, 1, 500) / 3.6e6 * side * side * dt / vp;
        const double whole = floor(st.carry);
        st.carry -= whole;
        if (whole <= 0) continue;
        e.shape = 3;
        e.count = (uint32_t)whole;
        set4(e.a, s.pos[0], s.pos[1], s.pos[2], side / 2);
        set4(e.b, 0, -kRainSpeed, 0, 0);
        set4(e.d, 0, 0, 0, kRainSpeed * dt);
        break;
      }
      default: continue;
    }
    st.liters += e.count * vp * 1000;
    if (k < max) {
      e.seed = next_seed(*sc);
      out[k++] = e;
    }
  }
  if (n_out) *n_out = k;
  return SL_OK;
}

// ------------------------------------------------------------------------------------------------
// blocks, 55-emit.js:83-113
// ------------------------------------------------------------------------------------------------
extern "C" int sl_scene_block_emits(sl_scene *sc, const sl_item *items, uint32_t n, int fill, sl_emit *out,
                                    uint32_t max, uint32_t *n_out) {
  if (!sc || (n && !items) || (max && !out)) return SL_EARG;
  uint32_t k = 0;
  const double sp = sc->h / 2;  // lattice spacing: 8 particles per cell
  // the lattice stays a quarter spacing inside the side walls and 0.5 mm above the floor
  const double in_lo[3] = {sc->origin[0] + sp / 2, 0.0005, sc->origin[2] + sp / 2};
  const double in_hi[3] = {-sc->origin[0] - sp / 2, INFINITY, -sc->origin[2] - sp / 2};
  for (uint32_t i = 0; i < n; i++) {
    const sl_item &b = items[i];
    if (b.kind != SL_ITEM_BLOCK) continue;
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = b.pos[a] - b.dims[a] / 2; hi[a] = b.pos[a] + b.dims[a] / 2; }
    // a block that overlaps the tank's footprint is clipped to its inside (sides and floor)
    const bool overlaps = lo[0] < in_hi[0] && hi[0] > in_lo[0] && lo[2] < in_hi[2] && hi[2] > in_lo[2];
    if (overlaps) {
      for (int a : {0, 2}) { lo[a] = std::max(lo[a], in_lo[a]); hi[a] = std::min(hi[a], in_hi[a]); }
      lo[1] = std::max(lo[1], in_lo[1]);
    }
    double cnt[3], start[3];
    for (int a = 0; a < 3; a++) {
      cnt[a] = std::max(0.0, floor((hi[a] - lo[a]) / sp));
      start[a] = lo[a] + ((hi[a] - lo[a]) - cnt[a] * sp) / 2;  // the lattice centred in the box
    }
    const double total = cnt[0] * cnt[1] * cnt[2];
    if (total <= 0) continue;
    if (k >= max) continue;
    sl_emit e;
    memset(&e, 0, sizeof e);
    e.shape = 2;
    e.count = (uint32_t)std::min(total, 4294967295.0);
    set4(e.a, start[0], start[1], start[2], sp);
    set4(e.b, cnt[0], cnt[1], cnt[2], fill ? 1.0 : 0.0);
    e.seed = next_seed(*sc);
    out[k++] = e;
  }
  if (n_out) *n_out = k;
  return SL_OK;
}

extern "C" double sl_block_liters(const sl_item *b, double tank_w, double tank_h, double tank_d) {
  if (!b) return 0;
  // overlap of [c - half, c + half] with [lo, hi]
  auto overlap = [](double c, double half, double lo, double hi) {
    return std::max(0.0, std::min(c + half, hi) - std::max(c - half, lo));
  };
  return overlap(b->pos[0], b->dims[0] / 2, -tank_w / 2, tank_w / 2) * overlap(b->pos[1], b->dims[1] / 2, 0, tank_h) *
         overlap(b->pos[2], b->dims[2] / 2, -tank_d / 2, tank_d / 2) * 1000;
}
