/*
 * Oracle (test infrastructure only) — plain-C fp32 restatement of the reference liquid step.
 * See suiri_oracle.h for the scope and the numeric conventions. Every function cites the WGSL or
 * JavaScript it restates (paths relative to the reference repository root).
 */
#include "suiri_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------------
 * WGSL scalar semantics
 * ---------------------------------------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;
typedef struct { int32_t x, y, z; } v3i;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline float len2(float x, float y) { return sqrtf(x * x + y * y); }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float fmaxw(float a, float b) { return fmaxf(a, b); }
static inline float fminw(float a, float b) { return fminf(a, b); }
static inline float clampw(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float mixw(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float sel(float f, float t, int c) { return c ? t : f; }
/* i32(f32) / u32(f32): truncate toward zero, saturate, NaN -> 0 */
static inline int32_t f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}
static inline uint32_t f2u(float f) {
  if (f != f || f <= 0.0f) return 0u;
  if (f >= 4294967296.0f) return UINT32_MAX;
  return (uint32_t)f;
}
/* WGSL transcendentals as correctly rounded f32 values: the double function rounded once. This is
   the most accurate result the WGSL error bounds allow and is independent of a library's last ulp. */
static inline float cr_expf(float x) { return (float)exp((double)x); }
static inline float cr_powf(float x, float y) { return (float)pow((double)x, (double)y); }
static inline float cr_sinf(float x) { return (float)sin((double)x); }
static inline float cr_cosf(float x) { return (float)cos((double)x); }
static inline float cr_sinhf(float x) { return (float)sinh((double)x); }
static inline float cr_logf(float x) { return (float)log((double)x); }
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline v3i v2i(v3 g) { v3i r = {f2i(g.x), f2i(g.y), f2i(g.z)}; return r; }
static inline float ld4(const float *a, uint32_t n, int c) { return a[(size_t)n * 4 + c]; }
static inline v3 pos3(const float *p) { return V(p[0], p[1], p[2]); }

/* ------------------------------------------------------------------------------------------------
 * Common helpers, src/js/30-wgsl-sim.js:22-37
 * ---------------------------------------------------------------------------------------------- */
#define NX(s) ((int32_t)(s)->P.dims[0])
#define NY(s) ((int32_t)(s)->P.dims[1])
#define NZ(s) ((int32_t)(s)->P.dims[2])
static inline uint32_t nodeIdx(const or_sim *s, int32_t i, int32_t j, int32_t k) {
  return (uint32_t)(i + (j + k * (NY(s) + 1)) * (NX(s) + 1));
}
static inline uint32_t cellIdx(const or_sim *s, int32_t i, int32_t j, int32_t k) {
  return (uint32_t)(i + (j + k * NY(s)) * NX(s));
}
static inline v3i nodeCoord(const or_sim *s, uint32_t n) {
  uint32_t sx = s->P.dims[0] + 1u, sy = s->P.dims[1] + 1u;
  v3i r = {(int32_t)(n % sx), (int32_t)((n / sx) % sy), (int32_t)(n / (sx * sy))};
  return r;
}
static inline v3i cellCoord(const or_sim *s, uint32_t c) {
  v3i r = {(int32_t)(c % s->P.dims[0]), (int32_t)((c / s->P.dims[0]) % s->P.dims[1]),
           (int32_t)(c / (s->P.dims[0] * s->P.dims[1]))};
  return r;
}
static inline v3 gridPos(const or_sim *s, v3 x) {
  return vmul(vsub(x, V(s->P.origin[0], s->P.origin[1], s->P.origin[2])), s->P.phys[3]);
}
static inline int inGrid(const or_sim *s, v3 g) {
  return g.x >= 0.0f && g.y >= 0.0f && g.z >= 0.0f && g.x < (float)s->P.dims[0] &&
         g.y < (float)s->P.dims[1] && g.z < (float)s->P.dims[2];
}
/* lowbias32-style hash, 30-wgsl-sim.js:36 */
uint32_t or_hashU(uint32_t a) {
  uint32_t x = a;
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
static inline float rnd(uint32_t *st) { *st = or_hashU(*st); return (float)(*st >> 8) / 16777216.0f; }

/* ------------------------------------------------------------------------------------------------
 * SDF library, src/js/30-wgsl-sim.js:40-114
 * ---------------------------------------------------------------------------------------------- */
static inline v3 toLocal(const or_obj *o, v3 p) {
  v3 d = vsub(p, pos3(o->pos));
  return V(vdot(pos3(o->r0), d), vdot(pos3(o->r1), d), vdot(pos3(o->r2), d));
}
static float sdRoundBox(v3 p, v3 b, float r) {
  v3 q = V(fabsf(p.x) - b.x + r, fabsf(p.y) - b.y + r, fabsf(p.z) - b.z + r);
  float l = vlen(V(fmaxw(q.x, 0.0f), fmaxw(q.y, 0.0f), fmaxw(q.z, 0.0f)));
  return l + fminw(fmaxw(q.x, fmaxw(q.y, q.z)), 0.0f) - r;
}
static float sdCyl(v3 p, float r, float hh) {
  float dx = len2(p.x, p.z) - r, dy = fabsf(p.y) - hh;
  return fminw(fmaxw(dx, dy), 0.0f) + len2(fmaxw(dx, 0.0f), fmaxw(dy, 0.0f));
}
static float sdCone(v3 p, float r, float hh) {
  float qx = len2(p.x, p.z), qy = p.y;
  float k1x = 0.0f, k1y = hh, k2x = -r, k2y = 2.0f * hh;
  float cax = qx - fminw(qx, sel(0.0f, r, qy < 0.0f)), cay = fabsf(qy) - hh;
  float t = clampw(((k1x - qx) * k2x + (k1y - qy) * k2y) / (k2x * k2x + k2y * k2y), 0.0f, 1.0f);
  float cbx = qx - k1x + k2x * t, cby = qy - k1y + k2y * t;
  float sg = sel(1.0f, -1.0f, cbx < 0.0f && cay < 0.0f);
  return sg * sqrtf(fminw(cax * cax + cay * cay, cbx * cbx + cby * cby));
}
static float shapeSdf(const or_obj *o, v3 p) {
  v3 q = toLocal(o, p);
  const float *d = o->dims;
  uint32_t t = f2u(o->pos[3] + 0.5f);
  float r = 1e9f;
  switch (t) {
    case 0: r = sdRoundBox(q, V(d[0], d[1], d[2]), d[3]); break;
    case 1: r = vlen(q) - d[0]; break;
    case 2: r = sdCyl(q, d[0], d[1]); break;
    case 3: r = sdCone(q, d[0], d[1]); break;
    case 4: r = len2(len2(q.x, q.z) - d[0], q.y) - d[1]; break;
    case 5: r = fmaxw(sdCyl(q, d[0], d[1]), -sdCyl(vsub(q, V(0.0f, d[2], 0.0f)), d[0] - d[2], d[1])); break;
    case 6:
      r = fmaxw(sdRoundBox(q, V(d[0], d[1], d[2]), 0.002f),
                -sdRoundBox(vsub(q, V(0.0f, d[3], 0.0f)), V(d[0] - d[3], d[1] - 0.0f, d[2] - d[3]), 0.001f));
      break;
    case 7: r = fmaxw(fmaxw(fabsf(len2(q.y, q.z) - d[1]) - 0.5f * d[2], q.y), fabsf(q.x) - d[0]); break;
    default: break;
  }
  return r;
}
static float sceneSdf(const or_sim *s, v3 p) {
  float best = 1e9f;
  for (uint32_t i = 0; i < s->P.ext[2]; i++) {
    const or_obj *o = &s->objs[i];
    if (o->angv[3] != 1.0f) continue;
    if (vlen(vsub(p, pos3(o->pos))) - o->vel[3] > best) continue;
    best = fminw(best, shapeSdf(o, p));
  }
  return best;
}
static v3 sceneNormal(const or_sim *s, v3 p) {
  const float e = 0.0006f;
  v3 n = V(sceneSdf(s, vadd(p, V(e, 0, 0))) - sceneSdf(s, vsub(p, V(e, 0, 0))),
           sceneSdf(s, vadd(p, V(0, e, 0))) - sceneSdf(s, vsub(p, V(0, e, 0))),
           sceneSdf(s, vadd(p, V(0, 0, e))) - sceneSdf(s, vsub(p, V(0, 0, e))));
  float l = vlen(n);
  return l > 1e-9f ? V(n.x / l, n.y / l, n.z / l) : V(0.0f, 1.0f, 0.0f);
}
static v3 solidVelAt(const or_sim *s, v3 p) {
  float best = 1e9f;
  v3 v = V(0, 0, 0);
  for (uint32_t i = 0; i < s->P.ext[2]; i++) {
    const or_obj *o = &s->objs[i];
    if (o->angv[3] != 1.0f) continue;
    if (vlen(vsub(p, pos3(o->pos))) - o->vel[3] > best) continue;
    float d = shapeSdf(o, p);
    if (d < best) { best = d; v = vadd(pos3(o->vel), vcross(pos3(o->angv), vsub(p, pos3(o->pos)))); }
  }
  return v;
}
static int inDrain(const or_sim *s, v3 p) {
  for (uint32_t i = 0; i < s->P.ext[2]; i++) {
    const or_obj *o = &s->objs[i];
    if (o->angv[3] != 2.0f) continue;
    if (sdCyl(toLocal(o, p), o->dims[0], o->dims[1]) < 0.0f) return 1;
  }
  return 0;
}
float or_scene_sdf(const or_sim *s, const float p[3]) { return sceneSdf(s, V(p[0], p[1], p[2])); }
void or_scene_normal(const or_sim *s, const float p[3], float out[3]) {
  v3 n = sceneNormal(s, V(p[0], p[1], p[2]));
  out[0] = n.x; out[1] = n.y; out[2] = n.z;
}

/* ------------------------------------------------------------------------------------------------
 * GEOM, src/js/30-wgsl-sim.js:118-157
 * ---------------------------------------------------------------------------------------------- */
static float faceOpen(const or_sim *s, v3 c, uint32_t axis) {
  float a = 0.5f * s->P.origin[3];
  v3 o1 = V(0, 0, 0), o2 = V(0, 0, 0);
  if (axis == 0) { o1 = V(0.0f, a, 0.0f); o2 = V(0.0f, 0.0f, a); }
  else if (axis == 1) { o1 = V(a, 0.0f, 0.0f); o2 = V(0.0f, 0.0f, a); }
  else { o1 = V(a, 0.0f, 0.0f); o2 = V(0.0f, a, 0.0f); }
  float s0 = sceneSdf(s, vsub(vsub(c, o1), o2));
  float s1 = sceneSdf(s, vsub(vadd(c, o1), o2));
  float s2 = sceneSdf(s, vadd(vadd(c, o1), o2));
  float s3 = sceneSdf(s, vadd(vsub(c, o1), o2));
  float sc = sceneSdf(s, c);
  float pos = fmaxw(s0, 0.0f) + fmaxw(s1, 0.0f) + fmaxw(s2, 0.0f) + fmaxw(s3, 0.0f) + 2.0f * fmaxw(sc, 0.0f);
  float tot = fabsf(s0) + fabsf(s1) + fabsf(s2) + fabsf(s3) + 2.0f * fabsf(sc);
  if (tot < 1e-12f) return 1.0f;
  return clampw(pos / tot, 0.0f, 1.0f);
}
float or_face_open(const or_sim *s, const float c[3], uint32_t axis) { return faceOpen(s, V(c[0], c[1], c[2]), axis); }

void or_geom(or_sim *s) {
  const uint32_t N = s->P.dims[3];
  const float h = s->P.origin[3];
  const v3 o = V(s->P.origin[0], s->P.origin[1], s->P.origin[2]);
#pragma omp parallel for schedule(dynamic, 256) num_threads(s->nthreads)
  for (uint32_t n = 0; n < N; n++) {
    v3i c = nodeCoord(s, n);
    float fr[4] = {0, 0, 0, 0}, sv[4] = {0, 0, 0, 0};
    if (c.y < NY(s) && c.z < NZ(s) && c.x > 0 && c.x < NX(s)) {
      v3 p = vadd(o, vmul(V((float)c.x, (float)c.y + 0.5f, (float)c.z + 0.5f), h));
      fr[0] = faceOpen(s, p, 0); sv[0] = solidVelAt(s, p).x;
    }
    if (c.x < NX(s) && c.z < NZ(s) && c.y > 0) {
      if (c.y >= NY(s)) fr[1] = 1.0f;
      else {
        v3 p = vadd(o, vmul(V((float)c.x + 0.5f, (float)c.y, (float)c.z + 0.5f), h));
        fr[1] = faceOpen(s, p, 1); sv[1] = solidVelAt(s, p).y;
      }
    }
    if (c.x < NX(s) && c.y < NY(s) && c.z > 0 && c.z < NZ(s)) {
      v3 p = vadd(o, vmul(V((float)c.x + 0.5f, (float)c.y + 0.5f, (float)c.z), h));
      fr[2] = faceOpen(s, p, 2); sv[2] = solidVelAt(s, p).z;
    }
    memcpy(&s->frac[(size_t)n * 4], fr, 16);
    memcpy(&s->svel[(size_t)n * 4], sv, 16);
  }
}

/* ------------------------------------------------------------------------------------------------
 * P2G, src/js/30-wgsl-sim.js:159-200  (fixed-point i32 atomics: order-independent)
 * ---------------------------------------------------------------------------------------------- */
#define VS 262144.0f
static void splat(or_sim *s, v3 q, float val, uint32_t comp) {
  v3i b = v2i(V(floorf(q.x), floorf(q.y), floorf(q.z)));
  v3 f = vsub(q, V((float)b.x, (float)b.y, (float)b.z));
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = b.z + dz;
    if (k < 0 || k > NZ(s)) continue;
    float wz = sel(1.0f - f.z, f.z, dz == 1);
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = b.y + dy;
      if (j < 0 || j > NY(s)) continue;
      float wy = sel(1.0f - f.y, f.y, dy == 1);
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = b.x + dx;
        if (i < 0 || i > NX(s)) continue;
        float w = sel(1.0f - f.x, f.x, dx == 1) * wy * wz;
        if (w <= 0.0f) continue;
        size_t n = (size_t)nodeIdx(s, i, j, k) * 6u + comp * 2u;
        __atomic_fetch_add(&s->acc[n], f2i(rintf(w * val * VS)), __ATOMIC_RELAXED);
        __atomic_fetch_add(&s->acc[n + 1], f2i(rintf(w * VS)), __ATOMIC_RELAXED);
      }
    }
  }
}
void or_p2g(or_sim *s) {
  const uint32_t cnt = s->C.count;
  or_particle *parts = s->parts[s->cur];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t id = 0; id < cnt; id++) {
    const or_particle *p = &parts[id];
    if (p->pos[3] < 0.5f) continue;
    v3 g = gridPos(s, pos3(p->pos));
    if (!inGrid(s, g)) continue;
    v3i c = v2i(g);
    __atomic_fetch_add(&s->cellCount[cellIdx(s, c.x, c.y, c.z)], 1u, __ATOMIC_RELAXED);
    v3 v = V(clampw(p->vel[0], -30.0f, 30.0f), clampw(p->vel[1], -30.0f, 30.0f), clampw(p->vel[2], -30.0f, 30.0f));
    splat(s, vsub(g, V(0.0f, 0.5f, 0.5f)), v.x, 0);
    splat(s, vsub(g, V(0.5f, 0.0f, 0.5f)), v.y, 1);
    splat(s, vsub(g, V(0.5f, 0.5f, 0.0f)), v.z, 2);
  }
}

/* NORMALIZE, src/js/30-wgsl-sim.js:203-222 */
void or_normalize(or_sim *s) {
  const uint32_t N = s->P.dims[3], Cc = s->P.ext[0];
  const uint32_t L = N > Cc ? N : Cc;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t n = 0; n < L; n++) {
    if (n < N) {
      float v[4] = {0, 0, 0, 0};
      uint32_t m = 0;
      for (uint32_t c = 0; c < 3; c++) {
        float wt = (float)s->acc[(size_t)n * 6 + c * 2 + 1];
        if (wt > 1.0f) { v[c] = (float)s->acc[(size_t)n * 6 + c * 2] / wt; m |= 1u << c; }
      }
      memcpy(&s->vel[(size_t)n * 4], v, 16);
      float vo[4] = {v[0], v[1], v[2], (float)m};
      memcpy(&s->velOld[(size_t)n * 4], vo, 16);
    }
    if (n < Cc) s->cellType[n] = s->cellCount[n] > 0u ? 1u : 0u;
  }
}

/* FORCES, src/js/30-wgsl-sim.js:224-238 */
void or_forces(or_sim *s) {
  const uint32_t N = s->P.dims[3];
  const float gdt = s->P.phys[1] * s->P.phys[0];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t n = 0; n < N; n++) {
    float *v = &s->vel[(size_t)n * 4];
    const float *f = &s->frac[(size_t)n * 4], *sv = &s->svel[(size_t)n * 4];
    v[1] -= gdt;
    if (f[0] <= 0.0f) v[0] = sv[0];
    if (f[1] <= 0.0f) v[1] = sv[1];
    if (f[2] <= 0.0f) v[2] = sv[2];
  }
}

/* DIVERGENCE, src/js/30-wgsl-sim.js:240-264 */
static inline float flux(float f, float v, float sv) { return f * v + (1.0f - f) * sv; }
void or_divergence(or_sim *s) {
  const uint32_t Cc = s->P.ext[0];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t c = 0; c < Cc; c++) {
    if (s->cellType[c] != 1u) { s->divg[c] = 0.0f; continue; }
    v3i q = cellCoord(s, c);
    uint32_t n0 = nodeIdx(s, q.x, q.y, q.z), nx = nodeIdx(s, q.x + 1, q.y, q.z);
    uint32_t ny = nodeIdx(s, q.x, q.y + 1, q.z), nz = nodeIdx(s, q.x, q.y, q.z + 1);
    float d = flux(ld4(s->frac, nx, 0), ld4(s->vel, nx, 0), ld4(s->svel, nx, 0)) -
              flux(ld4(s->frac, n0, 0), ld4(s->vel, n0, 0), ld4(s->svel, n0, 0));
    d += flux(ld4(s->frac, ny, 1), ld4(s->vel, ny, 1), ld4(s->svel, ny, 1)) -
         flux(ld4(s->frac, n0, 1), ld4(s->vel, n0, 1), ld4(s->svel, n0, 1));
    d += flux(ld4(s->frac, nz, 2), ld4(s->vel, nz, 2), ld4(s->svel, nz, 2)) -
         flux(ld4(s->frac, n0, 2), ld4(s->vel, n0, 2), ld4(s->svel, n0, 2));
    d *= s->P.phys[3];
    float crowd = (float)s->cellCount[c] / s->P.phys2[0] - 1.0f;
    s->divg[c] = d - s->P.phys2[1] * fmaxw(crowd, 0.0f);
  }
}

/* PRESSURE, one colour of red-black SOR, src/js/30-wgsl-sim.js:267-297 */
static inline void nb(const or_sim *s, uint32_t c, float f, int ok, float *num, float *den) {
  if (f <= 0.0f) return;
  *den += f;
  if (ok && s->cellType[c] == 1u) *num += f * s->pres[c];
}
void or_pressure(or_sim *s, uint32_t parity) {
  const uint32_t Cc = s->P.ext[0];
  const uint32_t sx = 1u, sy = s->P.dims[0], sz = s->P.dims[0] * s->P.dims[1];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t c = 0; c < Cc; c++) {
    v3i q = cellCoord(s, c);
    if ((uint32_t)((q.x + q.y + q.z) & 1) != parity) continue;
    if (s->cellType[c] != 1u) { s->pres[c] = 0.0f; continue; }
    float num = 0.0f, den = 0.0f;
    nb(s, c - (q.x > 0 ? sx : 0u), ld4(s->frac, nodeIdx(s, q.x, q.y, q.z), 0), q.x > 0, &num, &den);
    nb(s, c + sx, ld4(s->frac, nodeIdx(s, q.x + 1, q.y, q.z), 0), q.x < NX(s) - 1, &num, &den);
    nb(s, c - (q.y > 0 ? sy : 0u), ld4(s->frac, nodeIdx(s, q.x, q.y, q.z), 1), q.y > 0, &num, &den);
    nb(s, c + sy, ld4(s->frac, nodeIdx(s, q.x, q.y + 1, q.z), 1), q.y < NY(s) - 1, &num, &den);
    nb(s, c - (q.z > 0 ? sz : 0u), ld4(s->frac, nodeIdx(s, q.x, q.y, q.z), 2), q.z > 0, &num, &den);
    nb(s, c + sz, ld4(s->frac, nodeIdx(s, q.x, q.y, q.z + 1), 2), q.z < NZ(s) - 1, &num, &den);
    if (den > 0.0f) {
      float h = s->P.origin[3];
      float pn = (num - h * h * s->divg[c]) / den;
      s->pres[c] = mixw(s->pres[c], pn, s->P.phys2[2]);
    }
  }
}

/* PROJECT, src/js/30-wgsl-sim.js:299-341 */
static inline uint32_t ctype(const or_sim *s, int32_t i, int32_t j, int32_t k) {
  if (j >= NY(s)) return 0u;
  if (i < 0 || j < 0 || k < 0 || i >= NX(s) || k >= NZ(s)) return 2u;
  return s->cellType[cellIdx(s, i, j, k)];
}
static inline float cpres(const or_sim *s, int32_t i, int32_t j, int32_t k) {
  if (ctype(s, i, j, k) != 1u) return 0.0f;
  return s->pres[cellIdx(s, i, j, k)];
}
void or_project(or_sim *s) {
  const uint32_t N = s->P.dims[3];
  const float ih = s->P.phys[3];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (uint32_t n = 0; n < N; n++) {
    v3i c = nodeCoord(s, n);
    float *v = &s->vel[(size_t)n * 4];
    const float *f = &s->frac[(size_t)n * 4], *sv = &s->svel[(size_t)n * 4];
    if (c.y < NY(s) && c.z < NZ(s)) {
      if (f[0] <= 0.0f) v[0] = sv[0];
      else if (ctype(s, c.x - 1, c.y, c.z) == 1u || ctype(s, c.x, c.y, c.z) == 1u)
        v[0] -= (cpres(s, c.x, c.y, c.z) - cpres(s, c.x - 1, c.y, c.z)) * ih;
    }
    if (c.x < NX(s) && c.z < NZ(s)) {
      if (f[1] <= 0.0f) v[1] = sv[1];
      else if (ctype(s, c.x, c.y - 1, c.z) == 1u || ctype(s, c.x, c.y, c.z) == 1u)
        v[1] -= (cpres(s, c.x, c.y, c.z) - cpres(s, c.x, c.y - 1, c.z)) * ih;
    }
    if (c.x < NX(s) && c.y < NY(s)) {
      if (f[2] <= 0.0f) v[2] = sv[2];
      else if (ctype(s, c.x, c.y, c.z - 1) == 1u || ctype(s, c.x, c.y, c.z) == 1u)
        v[2] -= (cpres(s, c.x, c.y, c.z) - cpres(s, c.x, c.y, c.z - 1)) * ih;
    }
  }
}

/* ------------------------------------------------------------------------------------------------
 * G2P and its helpers, src/js/31-wgsl-sim2.js:8-251
 * ---------------------------------------------------------------------------------------------- */
/* transfer1, 31-wgsl-sim2.js:9-36: returns (PIC, FLIP, found) */
static v3 transfer1(const or_sim *s, v3 q, uint32_t comp, float vp) {
  v3i b = v2i(V(floorf(q.x), floorf(q.y), floorf(q.z)));
  v3 f = vsub(q, V((float)b.x, (float)b.y, (float)b.z));
  float sn = 0.0f, so = 0.0f, sw = 0.0f;
  uint32_t bit = 1u << comp;
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = b.z + dz;
    if (k < 0 || k > NZ(s)) continue;
    float wz = sel(1.0f - f.z, f.z, dz == 1);
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = b.y + dy;
      if (j < 0 || j > NY(s)) continue;
      float wy = sel(1.0f - f.y, f.y, dy == 1);
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = b.x + dx;
        if (i < 0 || i > NX(s)) continue;
        uint32_t n = nodeIdx(s, i, j, k);
        if ((f2u(ld4(s->velOld, n, 3)) & bit) == 0u) continue;
        float w = sel(1.0f - f.x, f.x, dx == 1) * wy * wz;
        sn += w * ld4(s->vel, n, comp);
        so += w * ld4(s->velOld, n, comp);
        sw += w;
      }
    }
  }
  if (sw < 1e-6f) return V(vp, vp, 0.0f);
  return V(sn / sw, vp + (sn - so) / sw, 1.0f);
}
/* sample1 / velNew, 31-wgsl-sim2.js:39-59 */
static float sample1(const or_sim *s, v3 q, uint32_t comp) {
  v3i b = v2i(V(floorf(q.x), floorf(q.y), floorf(q.z)));
  v3 f = vsub(q, V((float)b.x, (float)b.y, (float)b.z));
  float acc = 0.0f;
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = b.z + dz; k = k < 0 ? 0 : (k > NZ(s) ? NZ(s) : k);
    float wz = sel(1.0f - f.z, f.z, dz == 1);
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = b.y + dy; j = j < 0 ? 0 : (j > NY(s) ? NY(s) : j);
      float wy = sel(1.0f - f.y, f.y, dy == 1);
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = b.x + dx; i = i < 0 ? 0 : (i > NX(s) ? NX(s) : i);
        acc += sel(1.0f - f.x, f.x, dx == 1) * wy * wz * ld4(s->vel, nodeIdx(s, i, j, k), comp);
      }
    }
  }
  return acc;
}
static v3 velNew(const or_sim *s, v3 g) {
  return V(sample1(s, vsub(g, V(0.0f, 0.5f, 0.5f)), 0), sample1(s, vsub(g, V(0.5f, 0.0f, 0.5f)), 1),
           sample1(s, vsub(g, V(0.5f, 0.5f, 0.0f)), 2));
}
/* tank helpers, 31-wgsl-sim2.js:62-80 */
static inline void fpMin(const or_sim *s, float *lx, float *lz) { *lx = s->P.origin[0]; *lz = s->P.origin[2]; }
static inline void fpMax(const or_sim *s, float *hx, float *hz) {
  *hx = s->P.origin[0] + (float)s->P.dims[0] * s->P.origin[3];
  *hz = s->P.origin[2] + (float)s->P.dims[2] * s->P.origin[3];
}
static inline int inFootprint(const or_sim *s, v3 x) {
  float lx, lz, hx, hz;
  fpMin(s, &lx, &lz); fpMax(s, &hx, &hz);
  return x.x > lx && x.z > lz && x.x < hx && x.z < hz;
}
static void tankClamp(const or_sim *s, v3 x0, v3 *x, v3 *v, float r) {
  float lox, loz, hix, hiz;
  fpMin(s, &lox, &loz); fpMax(s, &hix, &hiz);
  int wasIn = inFootprint(s, x0);
  if (wasIn && x->y < s->P.tank[0]) {
    if (x->x < lox + r) { x->x = lox + r; v->x = fmaxw(v->x, 0.0f); }
    if (x->x > hix - r) { x->x = hix - r; v->x = fminw(v->x, 0.0f); }
    if (x->z < loz + r) { x->z = loz + r; v->z = fmaxw(v->z, 0.0f); }
    if (x->z > hiz - r) { x->z = hiz - r; v->z = fminw(v->z, 0.0f); }
  }
  if (wasIn && x->y < s->P.origin[1] + r) { x->y = s->P.origin[1] + r; v->y = fmaxw(v->y, 0.0f); }
  if (!wasIn && x->y < s->P.tank[0] && inFootprint(s, *x)) {
    *x = V(x0.x, x->y, x0.z);
    *v = V(-0.2f * v->x, v->y, -0.2f * v->z);
  }
}
/* countAt / airAt / nearAir, 31-wgsl-sim2.js:93-109 */
static inline uint32_t countAt(const or_sim *s, v3i c) {
  if (c.y >= NY(s) || c.x < 0 || c.y < 0 || c.z < 0 || c.x >= NX(s) || c.z >= NZ(s)) return 0u;
  return s->cellCount[cellIdx(s, c.x, c.y, c.z)];
}
static inline int airAt(const or_sim *s, v3i c) {
  if (c.y >= NY(s)) return 1;
  if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= NX(s) || c.z >= NZ(s)) return 0;
  return s->cellCount[cellIdx(s, c.x, c.y, c.z)] == 0u;
}
static inline v3i off3(v3i c, int32_t dx, int32_t dy, int32_t dz) {
  v3i r = {wadd(c.x, dx), wadd(c.y, dy), wadd(c.z, dz)};
  return r;
}
static int nearAir(const or_sim *s, v3i c) {
  for (int d = 0; d < 6; d++) {
    int32_t o[3] = {0, 0, 0};
    o[d / 2] = (d & 1) == 1 ? 1 : -1;
    if (airAt(s, off3(c, o[0], o[1], o[2]))) return 1;
  }
  return 0;
}
/* depositRipple, 31-wgsl-sim2.js:110-114 */
static void depositRipple(or_sim *s, float x, float z, float impulse, uint32_t R) {
  float lx, lz, hx, hz;
  fpMin(s, &lx, &lz); fpMax(s, &hx, &hz);
  float u = (x - lx) / (hx - lx), w = (z - lz) / (hz - lz);
  uint32_t i = f2u(clampw(rintf(u * (float)R), 0.0f, (float)R));
  uint32_t j = f2u(clampw(rintf(w * (float)R), 0.0f, (float)R));
  __atomic_fetch_add(&s->rsrc[j * (R + 1u) + i], f2i(clampw(impulse, 0.0f, 2e-3f) * 1e9f), __ATOMIC_RELAXED);
}
/* splashCrown, 31-wgsl-sim2.js:119-148. The reference's liquid constants (998, 0.0728, 0.001) and
   thresholds (700, 2100) are the quality's splash_* values; each constant expression is folded in
   double and rounded once, as the WGSL folds 3.14159265 * 998.0 and 998.0 * 0.0728 */
static void splashCrown(or_sim *s, v3 x, float U, float m, uint32_t id, or_particle *ww) {
  if (s->P.tank[1] <= 0.0f) return;
  const float PIRHO = (float)(3.14159265 * s->q.splash_rho);
  const float RS = (float)(s->q.splash_rho * s->q.splash_sigma);
  const float RHO = (float)s->q.splash_rho, SIGMA = (float)s->q.splash_sigma, MU = (float)s->q.splash_mu;
  const float KJ = s->q.splash_k_jet, KC = s->q.splash_k_crown;
  float D = cr_powf(6.0f * m / PIRHO, (float)(1.0 / 3.0));
  float We = RHO * U * U * D / SIGMA;
  float Oh = MU / sqrtf(RS * D);
  float K = We * cr_powf(Oh, -0.4f);
  uint32_t n = f2u(clampw(rintf(4.0f * (K / KC - 1.0f)), 0.0f, 24.0f));
  uint32_t nj = K > KJ ? (K > KC ? 2u : 1u) : 0u;
  if (n + nj == 0u) return;
  uint32_t st = or_hashU(id * 2654435761u + s->P.ext[3] * 97u + 3u);
  uint32_t base = __atomic_fetch_add(&s->C.wwCount, n + nj, __ATOMIC_RELAXED);
  uint32_t cap = f2u(s->P.tank[2]);
  float r0 = 0.6f * D;
  float ph0 = rnd(&st) * 6.2831853f;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t k = base + i;
    if (k >= cap) return;
    float ph = ph0 + 6.2831853f * ((float)i + 0.3f * rnd(&st)) / (float)n;
    float el = 0.9f + 0.45f * rnd(&st);
    float sp = U * (0.15f + 0.2f * rnd(&st));
    v3 d = V(cr_cosf(ph), 0.0f, cr_sinf(ph));
    or_particle q;
    v3 pp = vadd(x, vmul(d, r0));
    q.pos[0] = pp.x; q.pos[1] = pp.y; q.pos[2] = pp.z; q.pos[3] = 0.5f + 0.4f * rnd(&st);
    float ce = cr_cosf(el), se = cr_sinf(el);
    v3 vv = vadd(vmul(vmul(d, ce), sp), V(0.0f, se * sp, 0.0f));
    q.vel[0] = vv.x; q.vel[1] = vv.y; q.vel[2] = vv.z; q.vel[3] = 0.0f;
    ww[k] = q;
  }
  for (uint32_t j = 0; j < nj; j++) {
    uint32_t k = base + n + j;
    if (k >= cap) return;
    float sp = U * (0.22f + 0.04f * (float)j) * (0.9f + 0.2f * rnd(&st));
    or_particle q;
    q.pos[0] = x.x + 0.0f; q.pos[1] = x.y + 0.001f * (float)j; q.pos[2] = x.z + 0.0f; q.pos[3] = 0.7f;
    float a = (rnd(&st) - 0.5f) * 0.05f;
    float c = (rnd(&st) - 0.5f) * 0.05f;
    q.vel[0] = a; q.vel[1] = sp; q.vel[2] = c; q.vel[3] = 0.0f;
    ww[k] = q;
  }
}
/* Sub-grid droplets of a frame (no counterpart in the reference; the droplet layer of the build).
   Droplet d of frame f: hash stream seeded from (d, f, drop_seed); z = sqrt(-2 ln(1 - u1)) cos(2 pi u2)
   (Box-Muller), D = d50 exp(sigma_ln z), v = vmin + (vmax - vmin) u5, impact point
   (lx + (hx - lx) u3, lz + (hz - lz) u4) on the footprint; mass m = f32(pi/6 rho) D^3; it deposits
   2 m v through depositRipple and splashes through splashCrown at the water level (origin y + the
   frame's ripple depth). */
static void dropletDraw(const or_sim *s, uint32_t d, float *x, float *z, float *D, float *v) {
  uint32_t st = or_hashU(d * 0x9e3779b1u + s->P.ext[3] * 0x85ebca77u + s->q.drop_seed * 0xc2b2ae3du + 7u);
  float u1 = rnd(&st), u2 = rnd(&st), u3 = rnd(&st), u4 = rnd(&st), u5 = rnd(&st);
  float g = sqrtf(-2.0f * cr_logf(1.0f - u1)) * cr_cosf(6.2831853f * u2);
  *D = s->q.drop_d50 * cr_expf(s->q.drop_sigma_ln * g);
  *v = s->q.drop_vmin + (s->q.drop_vmax - s->q.drop_vmin) * u5;
  float lx, lz, hx, hz;
  fpMin(s, &lx, &lz); fpMax(s, &hx, &hz);
  *x = lx + (hx - lx) * u3;
  *z = lz + (hz - lz) * u4;
}
void or_droplets(or_sim *s, double depth) {
  const float mc = (float)(3.141592653589793 / 6.0 * s->q.splash_rho);
  for (uint32_t d = 0; d < s->q.droplets; d++) {
    float x, z, D, v;
    dropletDraw(s, d, &x, &z, &D, &v);
    float m = mc * D * D * D;
    if (s->R > 0u) depositRipple(s, x, z, 2.0f * m * v, s->R);
    splashCrown(s, V(x, s->P.origin[1] + (float)depth, z), v, m, 0x80000000u + d, s->ww[s->ww_cur]);
  }
}
/* test hook: the draws (x, z, D, v) of the current frame's droplets */
void or_droplet_draws(const or_sim *s, float *out) {
  for (uint32_t d = 0; d < s->q.droplets; d++)
    dropletDraw(s, d, out + 4 * d, out + 4 * d + 1, out + 4 * d + 2, out + 4 * d + 3);
}
/* test hook: one splashCrown at x into the current whitewater buffer (appends at wwCount) */
void or_crown(or_sim *s, const float *x, float U, float m, uint32_t id) {
  splashCrown(s, V(x[0], x[1], x[2]), U, m, id, s->ww[s->ww_cur]);
}
static inline void atomic_max_u32(uint32_t *p, uint32_t v) {
  uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

void or_g2p(or_sim *s) {
  const uint32_t cnt = s->C.count;
  or_particle *parts = s->parts[s->cur];
  or_particle *ww = s->ww[s->ww_cur];
#pragma omp parallel for schedule(dynamic, 1024) num_threads(s->nthreads)
  for (uint32_t id = 0; id < cnt; id++) {
    or_particle p = parts[id];
    if (p.pos[3] < 0.5f) continue;
    const float dt = s->P.phys[0];
    const float h = s->P.origin[3];
    v3 x0 = pos3(p.pos);
    v3 v = pos3(p.vel);
    v3 g = gridPos(s, x0);
    int ing = inGrid(s, g);
    if (ing) {
      v3 tx = transfer1(s, vsub(g, V(0.0f, 0.5f, 0.5f)), 0, v.x);
      v3 ty = transfer1(s, vsub(g, V(0.5f, 0.0f, 0.5f)), 1, v.y);
      v3 tz = transfer1(s, vsub(g, V(0.5f, 0.5f, 0.0f)), 2, v.z);
      float fl = s->P.phys[2];
      v = V(mixw(tx.x, tx.y, fl), mixw(ty.x, ty.y, fl), mixw(tz.x, tz.y, fl));
    } else {
      v.y -= s->P.phys[1] * dt;
    }
    v3 x = vadd(x0, vmul(v, dt));
    float rc = s->P.phys2[3];
    float d = sceneSdf(s, x);
    if (d < rc) {
      v3 nrm = sceneNormal(s, x);
      x = vadd(x, vmul(nrm, rc - d));
      v3 vs = solidVelAt(s, x);
      v3 rel = vsub(v, vs);
      float vn = vdot(rel, nrm);
      if (vn < 0.0f) v = vadd(vs, vmul(vsub(rel, vmul(nrm, vn)), 1.0f - s->P.tank[3]));
    }
    tankClamp(s, x0, &x, &v, rc);
    float alive = 1.0f;
    if (inDrain(s, x)) alive = 0.0f;
    if (!inFootprint(s, x) && x.y < 0.002f) alive = 0.0f;
    if (fabsf(x.x) > 20.0f || fabsf(x.z) > 20.0f || x.y < -1.0f || x.y > 30.0f) alive = 0.0f;
    v3i c0 = v2i(g);
    float airborne = p.vel[3];
    if (ing) {
      uint32_t n0 = countAt(s, c0);
      if (n0 >= 4u) airborne = 0.0f;
      else if (n0 <= 2u && wadd(c0.y, 2) < NY(s)) {
        uint32_t around = 0u, above = 0u;
        for (int dz = -1; dz <= 1; dz++)
          for (int dx = -1; dx <= 1; dx++) {
            for (int dy = -1; dy <= 1; dy++) around += countAt(s, off3(c0, dx, dy, dz));
            above += countAt(s, off3(c0, dx, 1, dz)) + countAt(s, off3(c0, dx, 2, dz));
          }
        if (around <= 4u && above == 0u) airborne = 1.0f;
      }
    }
    v3 g1 = gridPos(s, x);
    v3i c1 = v2i(g1);
    if (airborne > 0.5f && p.vel[1] < -0.3f && inGrid(s, g1)) {
      uint32_t others = countAt(s, c1) - ((c1.x == c0.x && c1.y == c0.y && c1.z == c0.z) ? 1u : 0u);
      if (others >= 3u) {
        v3 vl = V(transfer1(s, vsub(g1, V(0.0f, 0.5f, 0.5f)), 0, v.x).x,
                  transfer1(s, vsub(g1, V(0.5f, 0.0f, 0.5f)), 1, v.y).x,
                  transfer1(s, vsub(g1, V(0.5f, 0.5f, 0.0f)), 2, v.z).x);
        float dv = vl.y - p.vel[1];
        if (dv > 0.3f) {
          v = vl;
          airborne = 0.0f;
          __atomic_fetch_add(&s->C.impacts, 1u, __ATOMIC_RELAXED);
          float m = 998.0f * cr_powf(0.5f * h, 3.0f);
          uint32_t R = f2u(s->P.misc[2]);
          if (R > 0u) depositRipple(s, x.x, x.z, 2.0f * m * fminw(dv, 8.0f), R);
          float level = s->P.origin[1] +
                        ((float)c1.y + fminw((float)countAt(s, c1) / s->P.phys2[0], 1.0f)) * h;
          splashCrown(s, V(x.x, level + 0.25f * h, x.z), -p.vel[1], m, id, ww);
        }
      }
    }
    if (ing) atomic_max_u32(&s->C.vmax, fbits(fminw(vlen(v), 60.0f)));
    if (ing && s->P.tank[1] > 0.0f) {
      float sp = vlen(v);
      if (sp > 0.75f) {
        if (nearAir(s, c0)) {
          uint32_t st = or_hashU(id * 9781u + s->P.ext[3] * 6271u + 17u);
          float e = clampw((sp - 0.75f) / 1.6f, 0.0f, 1.0f);
          if (rnd(&st) < s->P.tank[1] * e * e * dt) {
            uint32_t k = __atomic_fetch_add(&s->C.wwCount, 1u, __ATOMIC_RELAXED);
            if (k < f2u(s->P.tank[2])) {
              float j0 = rnd(&st), j1 = rnd(&st), j2 = rnd(&st);
              v3 j = vsub(V(j0, j1, j2), V(0.5f, 0.5f, 0.5f));
              v3 pp = vadd(x, vmul(vmul(j, h), 0.6f));
              or_particle q;
              q.pos[0] = pp.x; q.pos[1] = pp.y; q.pos[2] = pp.z; q.pos[3] = 1.0f + 2.5f * rnd(&st);
              v3 vv = vadd(vmul(v, 0.8f + 0.35f * rnd(&st)), vmul(j, 0.35f));
              q.vel[0] = vv.x; q.vel[1] = vv.y; q.vel[2] = vv.z; q.vel[3] = 0.0f;
              ww[k] = q;
            }
          }
        }
      }
    }
    or_particle out;
    out.pos[0] = x.x; out.pos[1] = x.y; out.pos[2] = x.z; out.pos[3] = alive;
    out.vel[0] = v.x; out.vel[1] = v.y; out.vel[2] = v.z; out.vel[3] = airborne;
    parts[id] = out;
  }
}

/* ------------------------------------------------------------------------------------------------
 * EMIT, src/js/31-wgsl-sim2.js:262-315 (k emitters, `total` threads)
 * ---------------------------------------------------------------------------------------------- */
void or_emit_kernel(or_sim *s, uint32_t k, uint32_t total) {
  or_particle *parts = s->parts[s->cur];
  const uint32_t base_count = s->C.count;
  for (uint32_t t = 0; t < total; t++) {
    uint32_t e = 0;
    for (uint32_t i = 0; i < k; i++) if (t >= s->emits[i].info[0]) e = i;
    const or_emit *em = &s->emits[e];
    uint32_t local = t - em->info[0];
    uint32_t idx = base_count + t;
    if (idx >= s->P.ext[1]) continue;
    uint32_t st = or_hashU(t * 747796405u + em->info[2] * 2891336453u + 1u);
    v3 pos = V(0, 0, 0), vel = V(0, 0, 0);
    float mark = 0.0f;
    uint32_t shape = em->info[3];
    if (shape == 0u) {
      float rr = em->a[3] * sqrtf(rnd(&st));
      float th = 6.2831853f * rnd(&st);
      v3 u = pos3(em->c), w = pos3(em->d), dir = pos3(em->b);
      v3 ring = vmul(vadd(vmul(u, cr_cosf(th)), vmul(w, cr_sinf(th))), rr);
      float r3 = rnd(&st);
      pos = vadd(vadd(pos3(em->a), ring), vmul(dir, r3 * em->d[3]));
      v3 dv = dir;
      if (em->c[3] > 0.0f) {
        float ca = 1.0f - rnd(&st) * (1.0f - cr_cosf(em->c[3]));
        float sa = sqrtf(fmaxw(1.0f - ca * ca, 0.0f));
        float ph = 6.2831853f * rnd(&st);
        dv = vadd(vmul(dir, ca), vmul(vadd(vmul(u, cr_cosf(ph)), vmul(w, cr_sinf(ph))), sa));
      }
      float j0 = rnd(&st), j1 = rnd(&st), j2 = rnd(&st);
      vel = vadd(vmul(dv, em->b[3]), vmul(vsub(V(j0, j1, j2), V(0.5f, 0.5f, 0.5f)), 0.02f * em->b[3]));
    } else if (shape == 1u) {
      float j0 = rnd(&st), j1 = rnd(&st), j2 = rnd(&st);
      v3 dd = vsub(vmul(V(j0, j1, j2), 2.0f), V(1.0f, 1.0f, 1.0f));
      dd = vadd(dd, V(1e-4f, 1e-4f, 1e-4f));
      float l = vlen(dd);
      dd = V(dd.x / l, dd.y / l, dd.z / l);
      float r = cr_powf(rnd(&st), 0.3333f);
      pos = vadd(pos3(em->a), vmul(vmul(dd, em->a[3]), r));
      vel = pos3(em->b);
      mark = em->b[3];
    } else if (shape == 2u) {
      uint32_t nx = f2u(em->b[0]), ny = f2u(em->b[1]);
      uint32_t i = local % nx, j = (local / nx) % ny, kk = local / (nx * ny);
      float j0 = rnd(&st), j1 = rnd(&st), j2 = rnd(&st);
      v3 jit = vmul(vsub(V(j0, j1, j2), V(0.5f, 0.5f, 0.5f)), 0.3f);
      pos = vadd(pos3(em->a), vmul(vadd(vadd(V((float)i, (float)j, (float)kk), V(0.5f, 0.5f, 0.5f)), jit), em->a[3]));
      vel = pos3(em->c);
      v3 g = gridPos(s, pos);
      if (em->b[3] > 0.5f && inGrid(s, g)) {
        v3i c = v2i(g);
        if ((float)s->cellCount[cellIdx(s, c.x, c.y, c.z)] >= 0.5f * s->P.phys2[0]) {
          or_particle q = {{pos.x, pos.y, pos.z, 0.0f}, {0, 0, 0, 0}};
          parts[idx] = q;
          continue;
        }
      }
    } else {
      float r0 = rnd(&st), r1 = rnd(&st), r2 = rnd(&st);
      pos = vadd(pos3(em->a), V((r0 * 2.0f - 1.0f) * em->a[3], -r1 * em->d[3], (r2 * 2.0f - 1.0f) * em->a[3]));
      vel = pos3(em->b);
      mark = 1.0f;
    }
    float alive = 1.0f;
    if (sceneSdf(s, pos) < 0.0f) alive = 0.0f;
    or_particle q = {{pos.x, pos.y, pos.z, alive}, {vel.x, vel.y, vel.z, mark}};
    parts[idx] = q;
  }
}

/* COMPACT, src/js/31-wgsl-sim2.js:317-328, restated in place (see header) */
void or_compact(or_sim *s) {
  /* in place: survivors below the new count n' keep their slots; the dead slots below n' take, in
     ascending order, the survivors at or above n', in ascending order (a deterministic one of the
     orders the reference's atomic append can produce) */
  or_particle *p = s->parts[s->cur];
  uint32_t lim = s->C.count < s->P.ext[1] ? s->C.count : s->P.ext[1];
  uint32_t n1 = 0;
  for (uint32_t id = 0; id < lim; id++)
    if (p[id].pos[3] >= 0.5f) n1++;
  uint32_t hole = 0, src = n1;
  for (;;) {
    while (hole < n1 && p[hole].pos[3] >= 0.5f) hole++;
    while (src < lim && p[src].pos[3] < 0.5f) src++;
    if (hole >= n1 || src >= lim) break;
    p[hole++] = p[src++];
  }
  s->C.newCount = n1;
}

/* spatial re-ordering: stable counting sort by cell index (x fastest), outside the grid last */
void or_reorder(or_sim *s) {
  const uint32_t n = s->C.count;
  const or_particle *src = s->parts[s->cur];
  or_particle *dst = s->parts[1 - s->cur];
  const uint32_t sentinel = (uint32_t)s->nx * (uint32_t)s->ny * (uint32_t)s->nz;
  uint32_t *key = (uint32_t *)malloc((size_t)(n ? n : 1) * 4);
  uint32_t *hist = (uint32_t *)calloc((size_t)sentinel + 2, 4);
  for (uint32_t i = 0; i < n; i++) {
    v3 g = gridPos(s, pos3(src[i].pos));
    uint32_t k = sentinel;
    if (inGrid(s, g)) {
      v3i q = v2i(g);
      k = (uint32_t)q.x + (uint32_t)s->nx * ((uint32_t)q.y + (uint32_t)s->ny * (uint32_t)q.z);
    }
    key[i] = k;
    hist[k + 1]++;
  }
  for (uint32_t k = 1; k <= sentinel + 1; k++) hist[k] += hist[k - 1];
  for (uint32_t i = 0; i < n; i++) dst[hist[key[i]]++] = src[i];
  free(key);
  free(hist);
  s->cur = 1 - s->cur;
}

/* COUNT, src/js/31-wgsl-sim2.js:332-349 */
void or_count(or_sim *s, uint32_t mode, uint32_t y) {
  or_counters *C = &s->C;
  uint32_t cap = s->P.ext[1];
  uint32_t wcap = f2u(s->P.tank[2]);
  if (mode == 1) { C->count = C->newCount; C->newCount = 0; }
  if (mode == 2) { uint32_t t = C->count + y; C->count = t < cap ? t : cap; }
  if (mode == 3) { C->wwCount = C->wwNew < wcap ? C->wwNew : wcap; C->wwNew = 0; }
  if (mode == 4) { C->count = 0; C->newCount = 0; C->wwCount = 0; C->wwNew = 0; C->impacts = 0; }
  uint32_t n = C->count < cap ? C->count : cap;
  uint32_t w = C->wwCount < wcap ? C->wwCount : wcap;
  uint32_t a[14] = {(n + 255u) / 256u, 1, 1, (w + 255u) / 256u, 1, 1, 6, n, 0, 0, 6, w, 0, 0};
  memcpy(s->args, a, sizeof a);
}

/* WW, src/js/31-wgsl-sim2.js:361-409 */
void or_ww(or_sim *s, float dt) {
  const or_particle *src = s->ww[s->ww_cur];
  or_particle *dst = s->ww[1 - s->ww_cur];
  uint32_t wcap = f2u(s->P.tank[2]);
  uint32_t lim = s->C.wwCount < wcap ? s->C.wwCount : wcap;
  for (uint32_t id = 0; id < lim; id++) {
    or_particle p = src[id];
    v3 x = pos3(p.pos), v = pos3(p.vel);
    float life = p.pos[3];
    v3 g = gridPos(s, x);
    float kind = 0.0f;
    if (inGrid(s, g)) {
      v3i c = v2i(g);
      uint32_t n = s->cellCount[cellIdx(s, c.x, c.y, c.z)];
      if (n > 0u) {
        uint32_t up = 0u;
        if (c.y + 1 < NY(s)) up = s->cellCount[cellIdx(s, c.x, c.y + 1, c.z)];
        if (up > 0u) kind = 2.0f;
        else if (g.y - (float)c.y < fminw((float)n / s->P.phys2[0], 1.0f)) kind = 1.0f;
      }
    }
    uint32_t R = f2u(s->P.misc[2]);
    if (p.vel[3] < 0.5f && kind > 0.5f && R > 0u && p.vel[1] < -0.3f) {
      float lx, lz, hx, hz;
      fpMin(s, &lx, &lz); fpMax(s, &hx, &hz);
      float u = (x.x - lx) / (hx - lx), w = (x.z - lz) / (hz - lz);
      uint32_t i = f2u(clampw(rintf(u * (float)R), 0.0f, (float)R));
      uint32_t j = f2u(clampw(rintf(w * (float)R), 0.0f, (float)R));
      s->rsrc[j * (R + 1u) + i] += f2i(clampw(4.2e-6f * -p.vel[1], 0.0f, 1e-4f) * 1e9f);
    }
    if (kind == 0.0f) {
      v.y -= s->P.phys[1] * dt;
      v = vmul(v, 1.0f / (1.0f + 0.6f * dt));
      life -= dt * 0.7f;
    } else {
      v3 vf = velNew(s, g);
      if (kind == 1.0f) { v = V(vf.x, 0.2f * vf.y, vf.z); life -= dt * 0.5f; }
      else { v = vadd(vf, V(0.0f, 0.16f, 0.0f)); life -= dt * 0.3f; }
    }
    v3 xb = x;
    x = vadd(x, vmul(v, dt));
    float d = sceneSdf(s, x);
    if (d < 0.0f) x = vsub(x, vmul(sceneNormal(s, x), d));
    tankClamp(s, xb, &x, &v, 0.001f);
    int outside = !inFootprint(s, x) && x.y < 0.001f;
    if (life > 0.0f && !outside && x.y > -0.5f && !inDrain(s, x)) {
      uint32_t k = s->C.wwNew++;
      if (k < wcap) {
        or_particle q = {{x.x, x.y, x.z, life}, {v.x, v.y, v.z, kind}};
        dst[k] = q;
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------------
 * Ripple layer, src/js/64-ripple.js
 * ---------------------------------------------------------------------------------------------- */
#define RPI 3.14159265358979323846f
/* RSRC, 64-ripple.js:18-31 */
void or_rsrc(or_sim *s, float *out) {
  const uint32_t M = s->M, R = s->R;
  float dx = s->rU_l[0] / (float)R, dz = s->rU_l[1] / (float)R;
  for (uint32_t idx = 0; idx < M * M; idx++) {
    uint32_t i = idx % M, j = idx / M;
    uint32_t pi = i <= R ? i : M - i, pj = j <= R ? j : M - j;
    float v = (float)s->rsrc[pj * (R + 1u) + pi] / 1e9f / (dx * dz);
    out[2 * idx] = v; out[2 * idx + 1] = 0.0f;
  }
}
/* RFFT pass, 64-ripple.js:34-56 */
void or_rfft_pass(or_sim *s, const float *in, float *out, uint32_t ns, uint32_t horiz, float sign) {
  const uint32_t M = s->M;
  for (uint32_t id = 0; id < M * M; id++) {
    uint32_t x = id % M, y = id / M;
    uint32_t idx = horiz == 1u ? x : y;
    uint32_t hs = ns >> 1;
    uint32_t e = (idx / ns) * hs + (idx % hs);
    uint32_t o = e + (M >> 1);
    uint32_t ie = horiz == 1u ? y * M + e : e * M + x;
    uint32_t io = horiz == 1u ? y * M + o : o * M + x;
    float ax = in[2 * ie], ay = in[2 * ie + 1], bx = in[2 * io], by = in[2 * io + 1];
    uint32_t ti = (idx % ns) * (M / ns);
    float wx = s->tw[2 * ti], wy = sign * s->tw[2 * ti + 1];
    out[2 * id] = ax + (wx * bx - wy * by);
    out[2 * id + 1] = ay + (wx * by + wy * bx);
  }
}
/* encodeFFT, 64-ripple.js:205-216: horizontal passes, then vertical; ping-pong a/b; last -> final */
float *or_rfft2d(or_sim *s, float *input, float sign, float *final_out) {
  uint32_t L = 0;
  for (uint32_t ns = 2; ns <= s->M; ns <<= 1) L++;
  uint32_t total = 2 * L, i = 0;
  float *src = input;
  for (uint32_t horiz = 2; horiz-- > 0;) {
    uint32_t hz = horiz; /* 1 first, then 0 */
    for (uint32_t ns = 2; ns <= s->M; ns <<= 1, i++) {
      int last = i == total - 1;
      float *dst = last ? final_out : (src == s->ra ? s->rb : s->ra);
      or_rfft_pass(s, src, dst, ns, hz, sign);
      src = dst;
    }
  }
  return src;
}
/* REVOLVE, 64-ripple.js:58-90 */
void or_revolve(or_sim *s) {
  const uint32_t M = s->M, R = s->R;
  const float g = s->P.phys[1];
  const float sr = s->q.r_sigma_rho, nu = s->q.r_nu, film = s->q.r_film, hpk = s->q.r_hp_k;
  const float Lx = s->rU_l[0], Lz = s->rU_l[1], dt = s->rU_l[2], md = s->rU_l[3];
  const float PI = (float)3.141592653589793;
  for (uint32_t id = 0; id < M * M; id++) {
    int32_t i = (int32_t)(id % M), j = (int32_t)(id / M);
    int32_t mx = i < (int32_t)R ? i : i - (int32_t)M;
    int32_t mz = j < (int32_t)R ? j : j - (int32_t)M;
    float *st = &s->rstate[2 * id];
    if ((mx == 0 && mz == 0) || mx == -(int32_t)R || mz == -(int32_t)R) { st[0] = 0.0f; st[1] = 0.0f; continue; }
    float kx = (float)mx * PI / Lx, kz = (float)mz * PI / Lz;
    float kk = sqrtf(kx * kx + kz * kz);
    float D = fmaxw(md, 0.002f);
    float kd = fminw(kk * D, 15.0f);
    float th = 1.0f - 2.0f / (cr_expf(2.0f * kd) + 1.0f);
    float w2 = (g * kk + sr * kk * kk * kk) * th;
    float w = sqrtf(w2);
    float gam = 2.0f * nu * kk * kk + film * kk * sqrtf(nu * w * 0.125f) +
                kk * sqrtf(nu * w * 0.5f) / fmaxw(cr_sinhf(fminw(2.0f * kd, 30.0f)), 1e-4f);
    float wd = sqrtf(fmaxw(w2 - gam * gam, 1e-6f));
    float sx = st[0], sy = st[1];
    if (dt > 0.0f) {
      float c = cr_cosf(wd * dt), sn = cr_sinf(wd * dt), e = cr_expf(-gam * dt);
      float nx = sx * c + (sy + gam * sx) / wd * sn;
      float ny = sy * c - (w2 * sx + gam * sy) / wd * sn;
      sx = e * nx; sy = e * ny;
    }
    float hp = 1.0f - cr_expf(-(kk * kk) / (hpk * hpk));
    float sw = fmaxw(1.2f * Lx / (float)R, 0.003f);
    float spread = cr_expf(-0.5f * kk * kk * sw * sw);
    sy += -(kk * th / s->q.r_rho_kick) * s->rhat[2 * id] * hp * spread;
    st[0] = sx; st[1] = sy;
  }
}
/* Top-surface height map: the renderer's water-seen-from-the-sun pass (63-wgsl-water.js:5-33, drawn
 * by 66-draw.js:122-126 over the grid footprint F.wmap of 66-draw.js:45-46) and its four separable
 * blur passes (63-wgsl-water.js:35-55, 66-draw.js:127-131). The rasteriser's coverage is restated
 * with explicit texel centres: texel (ix, iy) of the W x W map lies at
 * xz = wmap.xy + ((i + 0.5) / W) * wmap.zw; a live particle at or below tank height + 0.01 covers it
 * when q = r^2 - |xz - p.xz|^2 > 0 (r = 1.25 x render radius) and offers p.y + sqrt(q); the depth
 * test keeps the largest. Texels no particle covers hold -1. */
static void toph_blur(const float *src, float *dst, uint32_t W, int dx, int dy) {
  for (int y = 0; y < (int)W; y++)
    for (int x = 0; x < (int)W; x++) {
      float h0 = src[(size_t)y * W + x];
      if (h0 < 0.0f) { dst[(size_t)y * W + x] = -1.0f; continue; }
      float sum = 0.0f, wsum = 0.0f;
      for (int i = -6; i <= 6; i++) {
        int qx = x + dx * i, qy = y + dy * i;
        qx = qx < 0 ? 0 : (qx > (int)W - 1 ? (int)W - 1 : qx);
        qy = qy < 0 ? 0 : (qy > (int)W - 1 ? (int)W - 1 : qy);
        float h = src[(size_t)qy * W + qx];
        if (h < 0.0f) continue;
        float k = cr_expf(-(float)(i * i) / 18.0f);
        sum += h * k;
        wsum += k;
      }
      dst[(size_t)y * W + x] = sum / wsum;
    }
}
void or_toph_splat(or_sim *s) {
  const uint32_t W = s->q.wmap;
  float *H = s->topH;
  for (size_t i = 0; i < (size_t)W * W; i++) H[i] = -1.0f;
  const float wx = (float)s->origin[0], wz = (float)s->origin[2];
  const float ww = (float)(s->nx * s->h), wd = (float)(s->nz * s->h);
  const float r = (float)s->radius * 1.25f, ylim = (float)s->tank_h + 0.01f, dW = (float)W;
  const or_particle *P = s->parts[s->cur];
  for (uint32_t n = 0; n < s->C.count && n < s->cap; n++) {
    const or_particle *p = &P[n];
    if (p->pos[3] < 0.5f || p->pos[1] > ylim) continue;
    int ix0 = (int)floorf((p->pos[0] - r - wx) / ww * dW - 0.5f), ix1 = (int)ceilf((p->pos[0] + r - wx) / ww * dW - 0.5f);
    int iy0 = (int)floorf((p->pos[2] - r - wz) / wd * dW - 0.5f), iy1 = (int)ceilf((p->pos[2] + r - wz) / wd * dW - 0.5f);
    if (ix0 < 0) ix0 = 0;
    if (iy0 < 0) iy0 = 0;
    if (ix1 > (int)W - 1) ix1 = (int)W - 1;
    if (iy1 > (int)W - 1) iy1 = (int)W - 1;
    for (int iy = iy0; iy <= iy1; iy++)
      for (int ix = ix0; ix <= ix1; ix++) {
        float cx = wx + (((float)ix + 0.5f) / dW) * ww, cz = wz + (((float)iy + 0.5f) / dW) * wd;
        float ddx = cx - p->pos[0], ddz = cz - p->pos[2];
        float q = r * r - (ddx * ddx + ddz * ddz);
        if (q <= 0.0f) continue;
        float y = p->pos[1] + sqrtf(q);
        if (y > H[(size_t)iy * W + ix]) H[(size_t)iy * W + ix] = y;
      }
  }
}
void or_toph_blur(or_sim *s) {
  const uint32_t W = s->q.wmap;
  float *H = s->topH;
  toph_blur(H, s->topB, W, 1, 0);
  toph_blur(s->topB, H, W, 0, 1);
  toph_blur(H, s->topB, W, 1, 0);
  toph_blur(s->topB, H, W, 0, 1);
}
void or_toph(or_sim *s) {
  or_toph_splat(s);
  or_toph_blur(s);
}

/* RMASKBUILD, 64-ripple.js:106-126 */
static float smoothstepw(float e0, float e1, float x) {
  float t = clampw((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
void or_rmaskbuild(or_sim *s) {
  const uint32_t R = s->R, W = s->q.wmap;
  for (uint32_t gy = 0; gy <= R; gy++)
    for (uint32_t gx = 0; gx <= R; gx++) {
      float u = (float)gx / (float)R, w = (float)gy / (float)R;
      float lox = s->P.origin[0], loz = s->P.origin[2];
      float hix = s->P.origin[0] + (float)s->P.dims[0] * s->P.origin[3];
      float hiz = s->P.origin[2] + (float)s->P.dims[2] * s->P.origin[3];
      float xx = lox + u * (hix - lox), zz = loz + w * (hiz - loz);
      float dw = (float)W;
      int32_t tx = f2i(clampw(u * dw, 0.0f, dw - 1.0f)), ty = f2i(clampw(w * dw, 0.0f, dw - 1.0f));
      float hh = s->topH[(size_t)ty * W + tx];
      float m;
      if (hh >= 0.0f) {
        float d = sceneSdf(s, V(xx, hh - 0.002f, zz));
        m = smoothstepw(-0.003f, 0.006f, d);
      } else m = 0.75f;
      s->mask[gy * (R + 1) + gx] = m;
    }
}
/* RMASK, 64-ripple.js:91-104 */
void or_rmask(or_sim *s, const float *field, float *out) {
  const uint32_t M = s->M, R = s->R;
  for (uint32_t idx = 0; idx < M * M; idx++) {
    uint32_t i = idx % M, j = idx / M;
    uint32_t pi = i <= R ? i : M - i, pj = j <= R ? j : M - j;
    float m = s->mask[pj * (R + 1) + pi];
    out[2 * idx] = field[2 * idx] * m; out[2 * idx + 1] = field[2 * idx + 1] * m;
  }
}
/* RSCALE, 64-ripple.js:127-133 */
void or_rscale(or_sim *s) {
  const uint32_t M = s->M;
  const float d = (float)(M * M);
  for (uint32_t i = 0; i < 2 * M * M; i++) s->rstate[i] = s->rstate[i] / d;
}
/* REXTRACT, 64-ripple.js:134-147 (f32 values of the rgba16f texture) */
void or_rextract(or_sim *s, const float *field) {
  const uint32_t M = s->M, R = s->R;
  const float sc = 1.0f / (float)(M * M);
  const float dx = s->rU_l[0] / (float)R, dz = s->rU_l[1] / (float)R;
  for (uint32_t j = 0; j <= R; j++)
    for (uint32_t i = 0; i <= R; i++) {
      float h = field[2 * (j * M + i)] * sc;
      float hl = field[2 * (j * M + (i + M - 1u) % M)] * sc, hr = field[2 * (j * M + (i + 1u) % M)] * sc;
      float hd = field[2 * (((j + M - 1u) % M) * M + i)] * sc, hu = field[2 * (((j + 1u) % M) * M + i)] * sc;
      float *t = &s->tex[4 * (j * (R + 1) + i)];
      t[0] = h; t[1] = (hr - hl) / (2.0f * dx); t[2] = (hu - hd) / (2.0f * dz); t[3] = 0.0f;
    }
}
/* RippleLayer.encodeFrame, 64-ripple.js:222-253 */
void or_ripple_frame(or_sim *s, double dt, double mean_depth) {
  if (s->R == 0) return;
  s->rU_l[0] = (float)(s->nx * s->h); s->rU_l[1] = (float)(s->nz * s->h);
  s->rU_l[2] = (float)dt; s->rU_l[3] = (float)mean_depth;
  or_rsrc(s, s->ra);
  or_rfft2d(s, s->ra, -1.0f, s->rhat);
  or_revolve(s);
  float *out = or_rfft2d(s, s->rstate, 1.0f, s->rb);
  or_rextract(s, out);
  if (s->mask_bound) {
    or_rmaskbuild(s);
    or_rmask(s, s->rb, s->ra);
    or_rfft2d(s, s->ra, -1.0f, s->rstate);
    or_rscale(s);
  }
  memset(s->rsrc, 0, (size_t)(s->R + 1) * (s->R + 1) * 4);
}

/* ------------------------------------------------------------------------------------------------
 * Driver, src/js/50-sim.js
 * ---------------------------------------------------------------------------------------------- */
void or_quality_default(or_quality *q, const char *name) {
  memset(q, 0, sizeof *q);
  /* QUALITY table, 50-sim.js:4-8 */
  q->h = 0.0085; q->cap = 320000; q->ww_cap = 140000; q->iters = 40; q->ripple_R = 256;
  if (name && !strcmp(name, "low")) { q->h = 0.012; q->cap = 140000; q->ww_cap = 60000; q->iters = 28; q->ripple_R = 128; }
  if (name && !strcmp(name, "high")) { q->h = 0.0065; q->cap = 700000; q->ww_cap = 260000; q->iters = 56; q->ripple_R = 512; }
  /* physics constants, 50-sim.js:157-160, 11 */
  q->gravity = 9.81f; q->flip = 0.96f; q->ppc = 8.0f; q->drift = 6.0f; q->omega = 1.6f;
  q->rc_factor = 0.32; q->friction = 0.03f; q->ww_rate = 7.0f;
  /* ripple constants, 64-ripple.js:225, :268, :88 */
  q->r_sigma_rho = (float)(0.0728 / 998.0); q->r_nu = 1.0e-6f; q->r_film = 0.3f;
  q->r_hp_k = (float)(2.0 * 3.141592653589793 / 0.12); q->r_rho_kick = 998.0f;
  q->wmap = 256; /* 65-render.js:161 */
  /* splashCrown, 31-wgsl-sim2.js:121-126 */
  q->splash_rho = 998.0; q->splash_sigma = 0.0728; q->splash_mu = 0.001;
  q->splash_k_jet = 700.0f; q->splash_k_crown = 2100.0f;
  q->ripple_depth = 0.0;
  q->droplets = 0; q->drop_seed = 1;
  q->drop_d50 = 0.0015f; q->drop_sigma_ln = 0.4f; q->drop_vmin = 1.0f; q->drop_vmax = 4.0f;
}

static void *xcalloc(size_t n, size_t sz) { void *p = calloc(n ? n : 1, sz); if (!p) abort(); return p; }

or_sim *or_create(double tank_w, double tank_h, double tank_d, const or_quality *q) {
  or_sim *s = (or_sim *)xcalloc(1, sizeof(or_sim));
  s->q = *q;
  s->reorder_clock = q->reorder_every ? q->reorder_phase % q->reorder_every : 0;
  s->tank_h = tank_h;
  s->nthreads = 1;
  /* grid sizing, 50-sim.js:33-45 */
  s->h = q->h;
  s->nx = (int)fmax(8.0, round(tank_w / q->h));
  s->ny = (int)fmax(8.0, round(tank_h / q->h));
  s->nz = (int)fmax(8.0, round(tank_d / q->h));
  s->h = fmin(fmin(tank_w / s->nx, tank_h / s->ny), tank_d / s->nz);
  s->origin[0] = -s->nx * s->h / 2; s->origin[1] = 0; s->origin[2] = -s->nz * s->h / 2;
  s->rim = s->ny * s->h;
  s->nodes = (uint32_t)((s->nx + 1) * (s->ny + 1) * (s->nz + 1));
  s->cells = (uint32_t)(s->nx * s->ny * s->nz);
  s->cap = q->cap; s->ww_cap = q->ww_cap; s->iters = q->iters;
  s->vp = pow(s->h / 2, 3);
  s->radius = 0.5 * s->h * 1.05;
  s->frame = 0; s->time = 0; s->vmax = 1.0;
  s->geom_dirty = 1;
  for (int i = 0; i < 2; i++) {
    s->parts[i] = (or_particle *)xcalloc(s->cap, sizeof(or_particle));
    s->ww[i] = (or_particle *)xcalloc(s->ww_cap, sizeof(or_particle));
  }
  s->acc = (int32_t *)xcalloc((size_t)s->nodes * 6, 4);
  s->cellCount = (uint32_t *)xcalloc(s->cells, 4);
  s->vel = (float *)xcalloc((size_t)s->nodes * 4, 4);
  s->velOld = (float *)xcalloc((size_t)s->nodes * 4, 4);
  s->frac = (float *)xcalloc((size_t)s->nodes * 4, 4);
  s->svel = (float *)xcalloc((size_t)s->nodes * 4, 4);
  s->cellType = (uint32_t *)xcalloc(s->cells, 4);
  s->divg = (float *)xcalloc(s->cells, 4);
  s->pres = (float *)xcalloc(s->cells, 4);
  /* ripple layer, 64-ripple.js:150-197 */
  s->R = q->ripple_R; s->M = 2 * q->ripple_R;
  if (s->R) {
    size_t M2 = (size_t)s->M * s->M;
    s->rsrc = (int32_t *)xcalloc((size_t)(s->R + 1) * (s->R + 1), 4);
    s->rstate = (float *)xcalloc(M2 * 2, 4);
    s->ra = (float *)xcalloc(M2 * 2, 4);
    s->rb = (float *)xcalloc(M2 * 2, 4);
    s->rhat = (float *)xcalloc(M2 * 2, 4);
    s->tw = (float *)xcalloc((size_t)s->M * 2, 4);
    for (uint32_t t = 0; t < s->M; t++) {
      s->tw[2 * t] = (float)cos((2 * 3.141592653589793 * t) / s->M);
      s->tw[2 * t + 1] = (float)sin((2 * 3.141592653589793 * t) / s->M);
    }
    s->mask = (float *)xcalloc((size_t)(s->R + 1) * (s->R + 1), 4);
    s->tex = (float *)xcalloc((size_t)(s->R + 1) * (s->R + 1) * 4, 4);
    s->topH = (float *)xcalloc((size_t)q->wmap * q->wmap, 4);
    s->topB = (float *)xcalloc((size_t)q->wmap * q->wmap, 4);
    s->mask_bound = 1;
  }
  or_write_params(s, 1.0 / 60.0);
  or_count(s, 4, 0);
  return s;
}

void or_destroy(or_sim *s) {
  if (!s) return;
  for (int i = 0; i < 2; i++) { free(s->parts[i]); free(s->ww[i]); }
  free(s->acc); free(s->cellCount); free(s->vel); free(s->velOld); free(s->frac); free(s->svel);
  free(s->cellType); free(s->divg); free(s->pres);
  free(s->rsrc); free(s->rstate); free(s->ra); free(s->rb); free(s->rhat); free(s->tw);
  free(s->mask); free(s->tex); free(s->topH); free(s->topB);
  free(s);
}

void or_set_threads(or_sim *s, int n) { s->nthreads = n < 1 ? 1 : n; }

/* writeParams, 50-sim.js:151-162 */
void or_write_params(or_sim *s, double dt) {
  or_params *P = &s->P;
  memset(P, 0, sizeof *P);
  P->dims[0] = (uint32_t)s->nx; P->dims[1] = (uint32_t)s->ny; P->dims[2] = (uint32_t)s->nz; P->dims[3] = s->nodes;
  P->ext[0] = s->cells; P->ext[1] = s->cap; P->ext[2] = s->obj_count; P->ext[3] = s->frame;
  P->origin[0] = (float)s->origin[0]; P->origin[1] = (float)s->origin[1]; P->origin[2] = (float)s->origin[2];
  P->origin[3] = (float)s->h;
  P->phys[0] = (float)dt; P->phys[1] = s->q.gravity; P->phys[2] = s->q.flip; P->phys[3] = (float)(1.0 / s->h);
  P->phys2[0] = s->q.ppc; P->phys2[1] = s->q.drift; P->phys2[2] = s->q.omega;
  P->phys2[3] = (float)(s->q.rc_factor * s->h);
  P->tank[0] = (float)s->rim; P->tank[1] = s->q.ww_rate; P->tank[2] = (float)s->ww_cap; P->tank[3] = s->q.friction;
  P->misc[0] = (float)s->time; P->misc[1] = (float)s->radius; P->misc[2] = (float)s->R; P->misc[3] = 0.0f;
}

/* object upload of setObjects, 50-sim.js:203-209: rows are the packed 28-float Obj records */
static int same_f(float a, float b) { return a == b || (a != a && b != b); }
int or_set_objects(or_sim *s, const float *rows, uint32_t n) {
  if (n > OR_MAX_OBJ) n = OR_MAX_OBJ;
  int changed = n != s->obj_count;
  for (uint32_t i = 0; !changed && i < n; i++) {
    const float *cur = (const float *)&s->objs[i];
    for (int k = 0; k < 28; k++) if (!same_f(cur[k], rows[i * 28 + k])) { changed = 1; break; }
  }
  if (changed) {
    memcpy(s->objs, rows, (size_t)n * 28 * 4);
    s->obj_count = n;
    s->geom_dirty = 1;
  }
  return changed;
}

/* writeEmits, 50-sim.js:222-245. `list[i].info` carries (unused, requested count, seed, shape). */
uint32_t or_write_emits(or_sim *s, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds) {
  uint32_t off = 0, k = 0;
  int64_t room = (int64_t)s->cap - (int64_t)s->count - 64;
  int64_t budget = room > 0 ? room : 0;
  int any = 0;
  for (uint32_t i = 0; i < n; i++) if (counts[i] > 0) any = 1;
  for (uint32_t i = 0; i < n; i++) {
    if (k >= OR_MAX_EMIT || budget <= 0) break;
    int64_t cnt = (int64_t)counts[i] < budget ? (int64_t)counts[i] : budget;
    if (cnt <= 0) continue;
    or_emit e = list[i];
    e.info[0] = off; e.info[1] = (uint32_t)cnt; e.info[2] = seeds[i]; e.info[3] = list[i].info[3];
    s->emits[k] = e;
    off += (uint32_t)cnt; budget -= cnt; k++;
  }
  if (any && budget <= 0) s->full = 1;
  if (off > 0) {
    s->emitU[0] = k; s->emitU[1] = off;
    s->count += off;
  }
  return off;
}

/* emitNow, 50-sim.js:248-265 */
void or_emit_now(or_sim *s, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds, int fill) {
  for (uint32_t i = 0; i < n; i += OR_MAX_EMIT) {
    uint32_t m = n - i < OR_MAX_EMIT ? n - i : OR_MAX_EMIT;
    uint32_t total = or_write_emits(s, list + i, counts + i, m, seeds + i);
    if (!total) continue;
    if (fill) {
      or_count(s, 0, 0);
      memset(s->acc, 0, (size_t)s->nodes * 24);
      memset(s->cellCount, 0, (size_t)s->cells * 4);
      or_p2g(s);
    }
    or_emit_kernel(s, s->emitU[0], total);
    or_count(s, 2, total);
  }
}

/* encodeSubstep, 50-sim.js:276-297 */
void or_substep(or_sim *s) {
  memset(s->acc, 0, (size_t)s->nodes * 24);
  memset(s->cellCount, 0, (size_t)s->cells * 4);
  if (s->geom_dirty) { or_geom(s); s->geom_dirty = 0; }
  or_p2g(s);
  or_normalize(s);
  or_forces(s);
  or_divergence(s);
  for (uint32_t i = 0; i < s->iters; i++) { or_pressure(s, 0); or_pressure(s, 1); }
  or_project(s);
  or_g2p(s);
}

int or_substep_count(const or_sim *s, double frame_dt) {
  double vmax = s->vmax > 0.4 ? s->vmax : 0.4;
  double n = ceil(frame_dt / (1.4 * s->h / vmax));
  if (n < 1) n = 1;
  if (n > 6) n = 6;
  return (int)n;
}

void or_step_begin(or_sim *s) { s->C.vmax = 0; }

/* the body of one iteration of step()'s substep loop, 50-sim.js:307-318 */
void or_step_substep(or_sim *s, double dt, const or_emit *list, const uint32_t *counts, uint32_t n, const uint32_t *seeds) {
  s->frame++;
  or_write_params(s, dt);
  uint32_t total = n ? or_write_emits(s, list, counts, n, seeds) : 0;
  if (total > 0) { or_emit_kernel(s, s->emitU[0], total); or_count(s, 2, total); }
  or_count(s, 0, 0);
  or_substep(s);
  s->time += dt;
}

/* frame tail of step(), 50-sim.js:319-335 */
void or_frame_tail(or_sim *s, double frame_dt) {
  or_compact(s);
  or_count(s, 1, 0);
  if (s->q.reorder_every && ++s->reorder_clock >= s->q.reorder_every) {
    or_reorder(s);
    s->reorder_clock = 0;
  }
  double liters = s->count * s->vp * 1000;
  double md = liters / 1000 / (s->nx * s->nz * s->h * s->h);
  md = md < 0.003 ? 0.003 : (md > s->rim ? s->rim : md);
  if (s->q.ripple_depth > 0) md = s->q.ripple_depth;  /* a fixed depth instead of the count's */
  if (s->R) or_droplets(s, md);  /* before WW: the same frame's whitewater pass moves their splashes */
  or_ww(s, (float)frame_dt);
  s->ww_cur = 1 - s->ww_cur;
  or_count(s, 3, 0);
  or_ripple_frame(s, frame_dt, md);
  if (s->R && s->q.top_height) or_toph(s);  /* drawn after the step; the next frame's mask reads it */
}

/* pollStats, 50-sim.js:346-363 */
void or_poll_stats(or_sim *s) {
  s->count = s->C.count;
  s->impacts = s->C.impacts;
  s->ww_count = s->C.wwCount < s->ww_cap ? s->C.wwCount : s->ww_cap;
  float vm; memcpy(&vm, &s->C.vmax, 4);
  if (isfinite(vm)) s->vmax = vm;
  if (s->count < s->cap * 0.9) s->full = 0;
}

/* clearWater, 50-sim.js:365-373 */
void or_clear_water(or_sim *s) {
  or_count(s, 4, 0);
  memset(s->pres, 0, (size_t)s->cells * 4);
  if (s->R) {
    memset(s->rstate, 0, (size_t)s->M * s->M * 8);
    memset(s->rsrc, 0, (size_t)(s->R + 1) * (s->R + 1) * 4);
  }
  s->count = 0; s->ww_count = 0; s->impacts = 0; s->full = 0; s->time = 0; s->vmax = 1;
}

/* ------------------------------------------------------------------------------------------------
 * Raw access for the Python checker
 * ---------------------------------------------------------------------------------------------- */
void *or_field(or_sim *s, const char *name, size_t *bytes) {
  size_t N = s->nodes, Cc = s->cells, M2 = (size_t)s->M * s->M, R1 = (size_t)(s->R + 1) * (s->R + 1);
#define F(nm, ptr, sz) if (!strcmp(name, nm)) { *bytes = (sz); return (void *)(ptr); }
  F("parts", s->parts[s->cur], (size_t)s->cap * 32)
  F("parts_other", s->parts[1 - s->cur], (size_t)s->cap * 32)
  F("ww", s->ww[s->ww_cur], (size_t)s->ww_cap * 32)
  F("ww_other", s->ww[1 - s->ww_cur], (size_t)s->ww_cap * 32)
  F("counters", &s->C, 32)
  F("args", s->args, 64)
  F("acc", s->acc, N * 24)
  F("cellCount", s->cellCount, Cc * 4)
  F("vel", s->vel, N * 16)
  F("velOld", s->velOld, N * 16)
  F("frac", s->frac, N * 16)
  F("svel", s->svel, N * 16)
  F("cellType", s->cellType, Cc * 4)
  F("divg", s->divg, Cc * 4)
  F("pres", s->pres, Cc * 4)
  F("objs", s->objs, sizeof s->objs)
  F("emits", s->emits, sizeof s->emits)
  F("params", &s->P, sizeof s->P)
  F("rsrc", s->rsrc, s->R ? R1 * 4 : 0)
  F("rstate", s->rstate, s->R ? M2 * 8 : 0)
  F("ra", s->ra, s->R ? M2 * 8 : 0)
  F("rb", s->rb, s->R ? M2 * 8 : 0)
  F("rhat", s->rhat, s->R ? M2 * 8 : 0)
  F("tw", s->tw, s->R ? (size_t)s->M * 8 : 0)
  F("mask", s->mask, s->R ? R1 * 4 : 0)
  F("tex", s->tex, s->R ? R1 * 16 : 0)
  F("topH", s->topH, s->R ? (size_t)s->q.wmap * s->q.wmap * 4 : 0)
#undef F
  *bytes = 0;
  return NULL;
}

void or_info(const or_sim *s, double *o) {
  o[0] = s->nx; o[1] = s->ny; o[2] = s->nz; o[3] = s->h; o[4] = s->origin[0]; o[5] = s->origin[1];
  o[6] = s->origin[2]; o[7] = s->rim; o[8] = s->vp; o[9] = s->nodes; o[10] = s->cells; o[11] = s->cap;
  o[12] = s->ww_cap; o[13] = s->frame; o[14] = s->obj_count; o[15] = s->count; o[16] = s->ww_count;
  o[17] = s->impacts; o[18] = s->time; o[19] = s->vmax; o[20] = s->full; o[21] = s->cur; o[22] = s->ww_cur;
  o[23] = s->R; o[24] = s->M; o[25] = s->iters;
}

void or_set_scalar(or_sim *s, const char *name, double v) {
  if (!strcmp(name, "count")) s->count = (uint32_t)v;
  else if (!strcmp(name, "time")) s->time = v;
  else if (!strcmp(name, "vmax")) s->vmax = v;
  else if (!strcmp(name, "frame")) s->frame = (uint32_t)v;
  else if (!strcmp(name, "iters")) s->iters = (uint32_t)v;
  else if (!strcmp(name, "geom_dirty")) s->geom_dirty = (int)v;
  else if (!strcmp(name, "mask_bound")) s->mask_bound = (int)v;
  else if (!strcmp(name, "cur")) s->cur = (int)v;
  else if (!strcmp(name, "ww_cur")) s->ww_cur = (int)v;
  else if (!strcmp(name, "full")) s->full = (int)v;
  else if (!strcmp(name, "ww_count")) s->ww_count = (uint32_t)v;
  else if (!strcmp(name, "impacts")) s->impacts = (uint32_t)v;
}
