// Particle-stage kernels: grid-to-particle transfer + advection + collisions + drop merging +
// whitewater spawning (G2P), emission (EMIT), counter commits (COUNT), compaction, whitewater.
#include "sl_internal.h"


// ------------------------------------------------------------------------------------------------
// particle count of an in-grid cell
__device__ __forceinline__ uint32_t cell_count(const SlTank &T, int32_t x, int32_t y, int32_t z) {
  return ld32(T.cellCount, cell_idx(T, x, y, z));
}

// the nodes (n, n + 1) of one array as a pair (one 8-byte load; the arrays are 4-byte aligned)
__device__ __forceinline__ float2 ld_pair(const float *base, uint32_t n) {
  float2 r;
  __builtin_memcpy(&r, __builtin_assume_aligned((const char *)base + n * 4u, 4), 8);
  return r;
}
// masked PIC/FLIP gather, 31-wgsl-sim2.js:9-36, in two parts so that G2P can issue the loads of one
// component before it sums another: tgather finds the 8 corners of q on component `comp`'s grid
// and loads vel and velOld (whose bit pattern carries the mask) of each row (y, z clamped into the
// grid) as the node pair at b0 = clamp(bx, 0, nx - 1), whose second node is the corner at i = nx and
// whose first is the corner at i = 0 reached from bx = -1; tsum drops the corners outside the grid
// or masked out by select, so the sums take the reference's terms in the reference's order, and
// returns (PIC, FLIP, found).
struct TGather {
  float fx, fy, fz;
  bool ix0, ix1, iy0, iy1, iz0, iz1, hi0, lo1;
  float2 a[4], b[4];  // rows (y0,z0), (y1,z0), (y0,z1), (y1,z1): vel and velOld of nodes b0, b0 + 1
};
__device__ __forceinline__ void tgather(const SlTank &T, f3 q, int comp, TGather &G) {
  const int32_t bx = f2i_hw(floorf(q.x)), by = f2i_hw(floorf(q.y)), bz = f2i_hw(floorf(q.z));
  G.fx = q.x - (float)bx; G.fy = q.y - (float)by; G.fz = q.z - (float)bz;
  G.ix0 = bx >= 0 && bx <= T.nx; G.ix1 = bx + 1 >= 0 && bx + 1 <= T.nx;
  G.iy0 = by >= 0 && by <= T.ny; G.iy1 = by + 1 >= 0 && by + 1 <= T.ny;
  G.iz0 = bz >= 0 && bz <= T.nz; G.iz1 = bz + 1 >= 0 && bz + 1 <= T.nz;
  const int32_t cy[2] = {min(max(by, 0), T.ny), min(max(by + 1, 0), T.ny)};
  const int32_t cz[2] = {min(max(bz, 0), T.nz), min(max(bz + 1, 0), T.nz)};
  const int32_t b0 = min(max(bx, 0), T.nx - 1);
  G.hi0 = bx == T.nx; G.lo1 = bx == -1;
  const float *vn = T.vel + comp * (size_t)T.nodes, *vo = T.velOld + comp * (size_t)T.nodes;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t n = node_idx(T, b0, cy[r & 1], cz[r >> 1]);
    G.a[r] = ld_pair(vn, n);
    G.b[r] = ld_pair(vo, n);
  }
}
__device__ __forceinline__ f3 tsum(const TGather &G, float vp) {
  bool mk[8];
  float a[8], b[8];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float2 pa = G.a[r], pb = G.b[r];
    a[2 * r] = G.hi0 ? pa.y : pa.x; b[2 * r] = G.hi0 ? pb.y : pb.x;
    a[2 * r + 1] = G.lo1 ? pa.x : pa.y; b[2 * r + 1] = G.lo1 ? pb.x : pb.y;
    mk[2 * r] = __float_as_uint(b[2 * r]) != SL_MASKED_OUT;
    mk[2 * r + 1] = __float_as_uint(b[2 * r + 1]) != SL_MASKED_OUT;
  }
  const bool ix[2] = {G.ix0, G.ix1}, iy[2] = {G.iy0, G.iy1}, iz[2] = {G.iz0, G.iz1};
  float sn = 0.0f, so = 0.0f, sw = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 2; dz++) {
    const float wz = dz == 1 ? G.fz : 1.0f - G.fz;
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
      const float wy = dy == 1 ? G.fy : 1.0f - G.fy;
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        const int c = dz * 4 + dy * 2 + dx;
        const bool on = iz[dz] && iy[dy] && ix[dx] && mk[c];
        const float w = (dx == 1 ? G.fx : 1.0f - G.fx) * wy * wz;
        sn = on ? sn + w * a[c] : sn;
        so = on ? so + w * b[c] : so;
        sw = on ? sw + w : sw;
      }
    }
  }
  if (sw < 1e-6f) return F3(vp, vp, 0.0f);
  return F3(sn / sw, vp + (sn - so) / sw, 1.0f);
}
__device__ __forceinline__ f3 transfer1(const SlTank &T, f3 q, int comp, float vp) {
  TGather G;
  tgather(T, q, comp, G);
  return tsum(G, vp);
}
// unmasked clamped sample, 31-wgsl-sim2.js:39-59
__device__ __forceinline__ float sample1(const SlTank &T, f3 q, int comp) {
  int32_t bx = f2i(floorf(q.x)), by = f2i(floorf(q.y)), bz = f2i(floorf(q.z));
  float fx = q.x - (float)bx, fy = q.y - (float)by, fz = q.z - (float)bz;
  const float *vn = T.vel + comp * (size_t)T.nodes;
  float s = 0.0f;
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = min(max(bz + dz, 0), T.nz);
    float wz = dz == 1 ? fz : 1.0f - fz;
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = min(max(by + dy, 0), T.ny);
      float wy = dy == 1 ? fy : 1.0f - fy;
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = min(max(bx + dx, 0), T.nx);
        s += (dx == 1 ? fx : 1.0f - fx) * wy * wz * vn[node_idx(T, i, j, k)];
      }
    }
  }
  return s;
}
// tank footprint + glass-wall clamp, 31-wgsl-sim2.js:62-80
__device__ __forceinline__ void fp_max(const SlTank &T, float &hx, float &hz) {
  hx = T.ox + (float)T.nx * T.h;
  hz = T.oz + (float)T.nz * T.h;
}
__device__ __forceinline__ bool in_footprint(const SlTank &T, f3 x) {
  float hx, hz;
  fp_max(T, hx, hz);
  return x.x > T.ox && x.z > T.oz && x.x < hx && x.z < hz;
}
__device__ __forceinline__ void tank_clamp(const SlTank &T, f3 x0, f3 &x, f3 &v, float r) {
  float hx, hz;
  fp_max(T, hx, hz);
  const float lx = T.ox, lz = T.oz;
  bool wasIn = in_footprint(T, x0);
  if (wasIn && x.y < T.rim) {
    if (x.x < lx + r) { x.x = lx + r; v.x = fmaxf(v.x, 0.0f); }
    if (x.x > hx - r) { x.x = hx - r; v.x = fminf(v.x, 0.0f); }
    if (x.z < lz + r) { x.z = lz + r; v.z = fmaxf(v.z, 0.0f); }
    if (x.z > hz - r) { x.z = hz - r; v.z = fminf(v.z, 0.0f); }
  }
  if (wasIn && x.y < T.oy + r) { x.y = T.oy + r; v.y = fmaxf(v.y, 0.0f); }
  if (!wasIn && x.y < T.rim && in_footprint(T, x)) {
    x = F3(x0.x, x.y, x0.z);
    v = F3(-0.2f * v.x, v.y, -0.2f * v.z);
  }
}
// 31-wgsl-sim2.js:93-109
__device__ __forceinline__ uint32_t count_at(const SlTank &T, int32_t x, int32_t y, int32_t z) {
  if (y >= T.ny || x < 0 || y < 0 || z < 0 || x >= T.nx || z >= T.nz) return 0u;
  return cell_count(T, x, y, z);
}
__device__ __forceinline__ bool air_at(const SlTank &T, int32_t x, int32_t y, int32_t z) {
  if (y >= T.ny) return true;
  if (x < 0 || y < 0 || z < 0 || x >= T.nx || z >= T.nz) return false;
  return cell_count(T, x, y, z) == 0u;
}
__device__ __forceinline__ bool near_air(const SlTank &T, i3 c) {
  return air_at(T, wadd(c.x, -1), c.y, c.z) || air_at(T, wadd(c.x, 1), c.y, c.z) ||
         air_at(T, c.x, wadd(c.y, -1), c.z) || air_at(T, c.x, wadd(c.y, 1), c.z) ||
         air_at(T, c.x, c.y, wadd(c.z, -1)) || air_at(T, c.x, c.y, wadd(c.z, 1));
}
// 31-wgsl-sim2.js:110-114
__device__ __forceinline__ void deposit_ripple(const SlTank &T, float x, float z, float impulse) {
  float hx, hz;
  fp_max(T, hx, hz);
  const uint32_t R = T.R;
  float u = (x - T.ox) / (hx - T.ox), w = (z - T.oz) / (hz - T.oz);
  uint32_t i = f2u(clampf(rintf(u * (float)R), 0.0f, (float)R));
  uint32_t j = f2u(clampf(rintf(w * (float)R), 0.0f, (float)R));
  atomicAdd(&T.rsrc[j * (R + 1u) + i], f2i(clampf(impulse, 0.0f, 2e-3f) * 1e9f));
}
__device__ __forceinline__ void ww_store(const SlTank &T, uint32_t k, f3 p, float life, f3 v, float kind) {
  const int wc = T.ww_cur;
  T.wp[wc][k] = make_float4(p.x, p.y, p.z, life);
  T.wv[wc][k] = make_float4(v.x, v.y, v.z, kind);
}
// splash crown, 31-wgsl-sim2.js:119-148, with the liquid's density, surface tension, viscosity and
// the two K thresholds taken from the tank (the reference's constants by default). Split in two: the
// drop's diameter and the ring and jet counts, then the records written from slot `base` on.
struct SplashCount { float D; uint32_t n, nj; };
__device__ __forceinline__ SplashCount splash_count(const SlTank &T, float U, float m) {
  SplashCount c;
  c.D = sl_pow(6.0f * m / T.sp_pirho, (float)(1.0 / 3.0));
  float We = T.sp_rho * U * U * c.D / T.sp_sigma;
  float Oh = T.sp_mu / sqrtf(T.sp_rs * c.D);
  float K = We * sl_pow(Oh, -0.4f);
  c.n = f2u(clampf(rintf(4.0f * (K / T.sp_kcrown - 1.0f)), 0.0f, 24.0f));
  c.nj = K > T.sp_kjet ? (K > T.sp_kcrown ? 2u : 1u) : 0u;
  if (T.ww_rate <= 0.0f) c.n = c.nj = 0u;
  return c;
}
__device__ __forceinline__ void splash_write(const SlTank &T, f3 x, float U, const SplashCount &c, uint32_t id,
                                             uint32_t base) {
  const uint32_t n = c.n, nj = c.nj;
  uint32_t s = hashU(id * 2654435761u + T.frame * 97u + 3u);
  uint32_t cap = f2u(T.ww_cap_f);
  float r0 = 0.6f * c.D;
  float ph0 = rndu(s) * 6.2831853f;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t k = base + i;
    if (k >= cap) return;
    float ph = ph0 + 6.2831853f * ((float)i + 0.3f * rndu(s)) / (float)n;
    float el = 0.9f + 0.45f * rndu(s);
    float sp = U * (0.15f + 0.2f * rndu(s));
    f3 d = F3(sl_cos(ph), 0.0f, sl_sin(ph));
    f3 p = add3(x, mul3(d, r0));
    float life = 0.5f + 0.4f * rndu(s);
    f3 v = add3(mul3(mul3(d, sl_cos(el)), sp), F3(0.0f, sl_sin(el) * sp, 0.0f));
    ww_store(T, k, p, life, v, 0.0f);
  }
  for (uint32_t j = 0; j < nj; j++) {
    uint32_t k = base + n + j;
    if (k >= cap) return;
    float sp = U * (0.22f + 0.04f * (float)j) * (0.9f + 0.2f * rndu(s));
    f3 p = F3(x.x + 0.0f, x.y + 0.001f * (float)j, x.z + 0.0f);
    float a = (rndu(s) - 0.5f) * 0.05f;
    float c2 = (rndu(s) - 0.5f) * 0.05f;
    ww_store(T, k, p, 0.7f, F3(a, sp, c2), 0.0f);
  }
}
// one splash with its own append (a rare event per particle in G2P)
__device__ __forceinline__ void splash_crown(const SlTank &T, f3 x, float U, float m, uint32_t id) {
  const SplashCount c = splash_count(T, U, m);
  if (c.n + c.nj == 0u) return;
  splash_write(T, x, U, c, id, atomicAdd(&T.C->wwCount, c.n + c.nj));
}

// ------------------------------------------------------------------------------------------------
// Sub-grid droplets (no counterpart in the reference, whose drops are particles of its sources):
// T.drops per frame. Droplet d of frame f draws from its own hash stream: a standard normal z by the
// Box-Muller transform of two uniforms, diameter D = d50 exp(sigma_ln z) (lognormal with median d50),
// speed v uniform in [vmin, vmax], impact point uniform over the tank footprint at the water level
// (the ripple depth of the frame). It deposits the impulse 2 m v, the merge impulse of G2P
// (31-wgsl-sim2.js:226), through depositRipple (:110-114), and splashes as splashCrown (:119-148).
// One thread per droplet; the whitewater it spawns is advanced by the same frame's WW pass.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void droplet_draw(const SlTank &T, uint32_t d, float &x, float &z, float &D, float &v) {
  uint32_t s = hashU(d * 0x9e3779b1u + T.frame * 0x85ebca77u + T.drop_seed * 0xc2b2ae3du + 7u);
  const float u1 = rndu(s), u2 = rndu(s), u3 = rndu(s), u4 = rndu(s), u5 = rndu(s);
  const float g = sqrtf(-2.0f * sl_log(1.0f - u1)) * sl_cos(6.2831853f * u2);  // 1 - u1 in (0, 1]
  D = T.drop_d50 * sl_exp(T.drop_sig * g);
  v = T.drop_vmin + (T.drop_vmax - T.drop_vmin) * u5;
  float hx, hz;
  fp_max(T, hx, hz);
  x = T.ox + (hx - T.ox) * u3;
  z = T.oz + (hz - T.oz) * u4;
}
// One wavefront per workgroup (64 droplets) so that a frame's 4096 droplets spread over 64 CUs; the
// wavefront reserves its splashes' whitewater slots with one atomic (lane prefix sums of the counts)
// instead of one atomic per droplet on the same counter.
template <class A>
__global__ __launch_bounds__(64) void k_droplets(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t d = lb * 64u + threadIdx.x;
  const bool on = d < T.drops;
  float x = 0.0f, z = 0.0f, D = 0.0f, v = 0.0f, m = 0.0f;
  SplashCount c = {0.0f, 0u, 0u};
  if (on) {
    droplet_draw(T, d, x, z, D, v);
    m = T.drop_mc * D * D * D;
    if (T.R > 0u) deposit_ripple(T, x, z, 2.0f * m * v);
    c = splash_count(T, v, m);
  }
  // inclusive scan of the record counts over the wavefront
  const uint32_t mine = c.n + c.nj;
  uint32_t incl = mine;
  const int lane = (int)threadIdx.x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 64);
    if (lane >= o) incl += y;
  }
  const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
  uint32_t base = 0;
  if (lane == 63 && total) base = atomicAdd(&T.C->wwCount, total);
  base = (uint32_t)__shfl((int)base, 63, 64);
  if (mine) splash_write(T, F3(x, T.oy + T.rmd, z), v, c, 0x80000000u + d, base + incl - mine);
}
void sl_k_droplets(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_droplets, dim3(L.w.total), dim3(64), 0, L);
}

// ------------------------------------------------------------------------------------------------
// G2P, 31-wgsl-sim2.js:149-251
// ------------------------------------------------------------------------------------------------
// one particle; returns the bit pattern of its speed candidate for vmax (0 when it has none)
__device__ __forceinline__ uint32_t g2p_one(const SlTank &T, uint32_t id) {
  const int cur = T.cur;
  // id < count <= capacity: the state loads go out together with the flags
  const uint32_t pf = ld32(T.pf[cur], id);
  const f3 x0 = F3(ld32(T.px[cur], id), ld32(T.py[cur], id), ld32(T.pz[cur], id));
  const f3 v0 = F3(ld32(T.vx[cur], id), ld32(T.vy[cur], id), ld32(T.vz[cur], id));
  if (!(pf & 1u)) return 0u;
  const float dt = T.dt, h = T.h;
  f3 v = v0;
  const f3 g = grid_pos(T, x0);
  const bool ing = in_grid(T, g);
  // the particle's own cell count and its cell's object-mask word go out with the gathers (the
  // free-flight test and the object culling below read them)
  const i3 c0 = toi3_hw(g);
  const uint32_t ci0 = cell_idx(T, c0.x, c0.y, c0.z);
  const uint32_t n0 = ing ? ld32(T.cellCount, ci0) : 0u;
  const uint32_t mword = ing ? ld32(T.objMask, ci0 >> 5) : 0xffffffffu;
  if (ing) {
    // two components' loads in flight, the third's issued while the first is summed: all three up
    // front need 87 VGPRs (5 waves per SIMD), this order 79 (6 waves without spills; G2P 0.587 ->
    // 0.525 ms at 256x128x128)
    TGather gx, gy, gz;
    tgather(T, sub3(g, F3(0.0f, 0.5f, 0.5f)), 0, gx);
    tgather(T, sub3(g, F3(0.5f, 0.0f, 0.5f)), 1, gy);
    const f3 tx = tsum(gx, v.x);
    tgather(T, sub3(g, F3(0.5f, 0.5f, 0.0f)), 2, gz);
    const f3 ty = tsum(gy, v.y), tz = tsum(gz, v.z);
    v = F3(mixf(tx.x, tx.y, T.flip), mixf(ty.x, ty.y, T.flip), mixf(tz.x, tz.y, T.flip));
  } else {
    v.y -= T.g * dt;
  }
  f3 x = add3(x0, mul3(v, dt));
  const float rc = T.rc;
  // object loops only where the mask of the start cell allows an object within reach (it covers a
  // particle from inside the footprint that moved at most h along every axis; others take the loops)
  const bool objs = !ing || !in_footprint(T, x0) || !(fabsf(x.x - x0.x) <= h && fabsf(x.y - x0.y) <= h &&
                                                      fabsf(x.z - x0.z) <= h) || ((mword >> (ci0 & 31u)) & 1u);
  // only d < rc matters; no obstacle within reach means scene_sdf(x) > rc
  const float d = objs && obstacle_within(T, x, rc) ? scene_sdf(T, x) : rc;
  if (d < rc) {
    f3 nrm = scene_normal(T, x);
    x = add3(x, mul3(nrm, rc - d));
    f3 vs = solid_vel_at(T, x);
    f3 rel = sub3(v, vs);
    float vn = dot3(rel, nrm);
    if (vn < 0.0f) v = add3(vs, mul3(sub3(rel, mul3(nrm, vn)), 1.0f - T.friction));
  }
  tank_clamp(T, x0, x, v, rc);
  bool alive = true;
  if (objs && in_drain(T, x)) alive = false;
  if (!in_footprint(T, x) && x.y < 0.002f) alive = false;
  if (fabsf(x.x) > 20.0f || fabsf(x.z) > 20.0f || x.y < -1.0f || x.y > 30.0f) alive = false;
  bool airborne = (pf & 2u) != 0u;
  if (ing) {
    if (n0 >= 4u) airborne = false;
    else if (n0 <= 2u && wadd(c0.y, 2) < T.ny) {
      // 3x3 columns of rows y-1 .. y+2 (rows y+1 and y+2 lie in the grid here, row y-1 when y > 0):
      // the same integer sums as the reference's 27 + 18 count lookups
      uint32_t around = 0u, above = 0u;
#pragma unroll
      for (int dz = -1; dz <= 1; dz++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int32_t xx = wadd(c0.x, dx), zz = wadd(c0.z, dz);
          const bool col = xx >= 0 && zz >= 0 && xx < T.nx && zz < T.nz;
          const int32_t xc = min(max(xx, 0), T.nx - 1), zc = min(max(zz, 0), T.nz - 1);
          const uint32_t lo = cell_count(T, xc, max(c0.y - 1, 0), zc), mid = cell_count(T, xc, c0.y, zc);
          const uint32_t up1 = cell_count(T, xc, c0.y + 1, zc), up2 = cell_count(T, xc, c0.y + 2, zc);
          around += col ? (c0.y > 0 ? lo : 0u) + mid + up1 : 0u;
          above += col ? up1 + up2 : 0u;
        }
      if (around <= 4u && above == 0u) airborne = true;
    }
  }
  const f3 g1 = grid_pos(T, x);
  const i3 c1 = toi3_hw(g1);
  if (airborne && v0.y < -0.3f && in_grid(T, g1)) {
    uint32_t others = count_at(T, c1.x, c1.y, c1.z) - ((c1.x == c0.x && c1.y == c0.y && c1.z == c0.z) ? 1u : 0u);
    if (others >= 3u) {
      f3 vl = F3(transfer1(T, sub3(g1, F3(0.0f, 0.5f, 0.5f)), 0, v.x).x,
                 transfer1(T, sub3(g1, F3(0.5f, 0.0f, 0.5f)), 1, v.y).x,
                 transfer1(T, sub3(g1, F3(0.5f, 0.5f, 0.0f)), 2, v.z).x);
      float dv = vl.y - v0.y;
      if (dv > 0.3f) {
        v = vl;
        airborne = false;
        atomicAdd(&T.C->impacts, 1u);
        float m = 998.0f * sl_pow(0.5f * h, 3.0f);
        if (T.R > 0u) deposit_ripple(T, x.x, x.z, 2.0f * m * fminf(dv, 8.0f));
        float level = T.oy + ((float)c1.y + fminf((float)count_at(T, c1.x, c1.y, c1.z) / T.ppc, 1.0f)) * h;
        splash_crown(T, F3(x.x, level + 0.25f * h, x.z), -v0.y, m, id);
      }
    }
  }
  const uint32_t vm = ing ? __float_as_uint(fminf(len3(v), 60.0f)) : 0u;
  if (ing && T.ww_rate > 0.0f) {
    float sp = len3(v);
    if (sp > 0.75f && near_air(T, c0)) {
      uint32_t s = hashU(id * 9781u + T.frame * 6271u + 17u);
      float e = clampf((sp - 0.75f) / 1.6f, 0.0f, 1.0f);
      if (rndu(s) < T.ww_rate * e * e * dt) {
        uint32_t k = atomicAdd(&T.C->wwCount, 1u);
        if (k < f2u(T.ww_cap_f)) {
          float j0 = rndu(s), j1 = rndu(s), j2 = rndu(s);
          f3 j = sub3(F3(j0, j1, j2), F3(0.5f, 0.5f, 0.5f));
          f3 p = add3(x, mul3(mul3(j, h), 0.6f));
          float life = 1.0f + 2.5f * rndu(s);
          f3 wv = add3(mul3(v, 0.8f + 0.35f * rndu(s)), mul3(j, 0.35f));
          ww_store(T, k, p, life, wv, 0.0f);
        }
      }
    }
  }
  st32(T.px[cur], id, x.x); st32(T.py[cur], id, x.y); st32(T.pz[cur], id, x.z);
  st32(T.vx[cur], id, v.x); st32(T.vy[cur], id, v.y); st32(T.vz[cur], id, v.z);
  st32(T.pf[cur], id, (uint8_t)((alive ? 1u : 0u) | (airborne ? 2u : 0u)));
  return vm;
}
// vmax: atomicMax on the bit pattern of a non-negative float (31-wgsl-sim2.js:233), reduced over
// the workgroup first; the workgroup's atomic goes to its slot (workgroup % SL_VSLOTS), which COUNT
// folds into the counter: tens of thousands of workgroups on one address serialise there. MINW is
// the occupancy target handed to the register allocator (waves per SIMD); the variants compute
// identical results.
template <int MINW, class A>
__global__ __launch_bounds__(SL_WG, MINW) void k_g2p(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t id = lb * SL_WG + threadIdx.x;
  uint32_t vm = id < T.C->count ? g2p_one(T, id) : 0u;
  for (int o = 32; o > 0; o >>= 1) vm = max(vm, (uint32_t)__shfl_xor((int)vm, o, 64));
  __shared__ uint32_t red[SL_WG / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vm;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = red[0];
    for (int w = 1; w < SL_WG / 64; w++) m = max(m, red[w]);
    if (m) atomicMax(&T.vslot[(blockIdx.x % SL_VSLOTS) * SL_VSLOT_STRIDE], m);
  }
}
template <int MINW>
static void launch_g2p(const SlLaunch &L) {
  if (L.one) {
    SlOne o;
    o.T = *L.one;
    o.total = L.w.total;
    hipLaunchKernelGGL((k_g2p<MINW, SlOne>), dim3(L.w.total), dim3(SL_WG), 0, L.stream, o);
  } else {
    hipLaunchKernelGGL((k_g2p<MINW, SlWork>), dim3(L.w.total), dim3(SL_WG), 0, L.stream, L.w);
  }
}
void sl_k_g2p(const SlLaunch &L, int variant) {
  if (!L.w.total) return;
  switch (variant & 3) {
    case 0: launch_g2p<1>(L); break;
    case 1: launch_g2p<5>(L); break;
    case 2: launch_g2p<6>(L); break;
    default: launch_g2p<8>(L); break;
  }
}

// ------------------------------------------------------------------------------------------------
// EMIT, 31-wgsl-sim2.js:262-315 (one thread per emitted particle, `emit_total` of them)
// ------------------------------------------------------------------------------------------------
// `reserve`: slots kept free below the capacity (0 here; a batch passes the 64 that the host-side
// budget of writeEmits keeps, 50-sim.js:226, so that the budget is applied on the device count)
template <class A>
__global__ __launch_bounds__(SL_WG) void k_emit(A W, uint32_t reserve) {
  SL_TANK_OF(W, blockIdx.x, tk, lb);
  const SlTank &T = sl_tank(W, tk);
  uint32_t t = lb * SL_WG + threadIdx.x;
  if (t >= T.emit_total) return;
  uint32_t e = 0;
  for (uint32_t i = 0; i < T.emit_k; i++) if (t >= T.emits[i].info[0]) e = i;
  const SlEmitDev &em = T.emits[e];
  uint32_t local = t - em.info[0];
  uint32_t idx = T.C->count + t;
  if (idx >= T.cap || idx + reserve >= T.cap) return;
  uint32_t s = hashU(t * 747796405u + em.info[2] * 2891336453u + 1u);
  f3 pos = F3(0, 0, 0), vel = F3(0, 0, 0);
  bool mark = false;
  const uint32_t shape = em.info[3];
  const int cur = T.cur;
  if (shape == 0u) {
    float rr = em.a[3] * sqrtf(rndu(s));
    float th = 6.2831853f * rndu(s);
    f3 u = F3(em.c[0], em.c[1], em.c[2]), w = F3(em.d[0], em.d[1], em.d[2]), dir = F3(em.b[0], em.b[1], em.b[2]);
    f3 ring = mul3(add3(mul3(u, sl_cos(th)), mul3(w, sl_sin(th))), rr);
    float r3 = rndu(s);
    pos = add3(add3(F3(em.a[0], em.a[1], em.a[2]), ring), mul3(dir, r3 * em.d[3]));
    f3 dv = dir;
    if (em.c[3] > 0.0f) {
      float ca = 1.0f - rndu(s) * (1.0f - sl_cos(em.c[3]));
      float sa = sqrtf(fmaxf(1.0f - ca * ca, 0.0f));
      float ph = 6.2831853f * rndu(s);
      dv = add3(mul3(dir, ca), mul3(add3(mul3(u, sl_cos(ph)), mul3(w, sl_sin(ph))), sa));
    }
    float j0 = rndu(s), j1 = rndu(s), j2 = rndu(s);
    vel = add3(mul3(dv, em.b[3]), mul3(sub3(F3(j0, j1, j2), F3(0.5f, 0.5f, 0.5f)), 0.02f * em.b[3]));
  } else if (shape == 1u) {
    float j0 = rndu(s), j1 = rndu(s), j2 = rndu(s);
    f3 dd = sub3(mul3(F3(j0, j1, j2), 2.0f), F3(1.0f, 1.0f, 1.0f));
    dd = add3(dd, F3(1e-4f, 1e-4f, 1e-4f));
    float l = len3(dd);
    dd = F3(dd.x / l, dd.y / l, dd.z / l);
    float r = sl_pow(rndu(s), 0.3333f);
    pos = add3(F3(em.a[0], em.a[1], em.a[2]), mul3(mul3(dd, em.a[3]), r));
    vel = F3(em.b[0], em.b[1], em.b[2]);
    mark = em.b[3] > 0.5f;
  } else if (shape == 2u) {
    uint32_t nx = f2u(em.b[0]), ny = f2u(em.b[1]);
    uint32_t i = local % nx, j = (local / nx) % ny, k = local / (nx * ny);
    float j0 = rndu(s), j1 = rndu(s), j2 = rndu(s);
    f3 jit = mul3(sub3(F3(j0, j1, j2), F3(0.5f, 0.5f, 0.5f)), 0.3f);
    pos = add3(F3(em.a[0], em.a[1], em.a[2]),
               mul3(add3(add3(F3((float)i, (float)j, (float)k), F3(0.5f, 0.5f, 0.5f)), jit), em.a[3]));
    vel = F3(em.c[0], em.c[1], em.c[2]);
    f3 g = grid_pos(T, pos);
    if (em.b[3] > 0.5f && in_grid(T, g)) {
      i3 c = toi3(g);
      if ((float)T.cellCount[cell_idx(T, c.x, c.y, c.z)] >= 0.5f * T.ppc) {
        T.px[cur][idx] = pos.x; T.py[cur][idx] = pos.y; T.pz[cur][idx] = pos.z;
        T.vx[cur][idx] = 0.0f; T.vy[cur][idx] = 0.0f; T.vz[cur][idx] = 0.0f;
        T.pf[cur][idx] = 0;
        return;
      }
    }
  } else {
    float r0 = rndu(s), r1 = rndu(s), r2 = rndu(s);
    pos = add3(F3(em.a[0], em.a[1], em.a[2]), F3((r0 * 2.0f - 1.0f) * em.a[3], -r1 * em.d[3], (r2 * 2.0f - 1.0f) * em.a[3]));
    vel = F3(em.b[0], em.b[1], em.b[2]);
    mark = true;
  }
  bool alive = !(scene_sdf(T, pos) < 0.0f);
  T.px[cur][idx] = pos.x; T.py[cur][idx] = pos.y; T.pz[cur][idx] = pos.z;
  T.vx[cur][idx] = vel.x; T.vy[cur][idx] = vel.y; T.vz[cur][idx] = vel.z;
  T.pf[cur][idx] = (uint8_t)((alive ? 1u : 0u) | (mark ? 2u : 0u));
}
void sl_k_emit(const SlLaunch &L, uint32_t reserve) {
  if (!L.w.total) return;
  SL_LAUNCH(k_emit, dim3(L.w.total), dim3(SL_WG), 0, L, reserve);
}

// ------------------------------------------------------------------------------------------------
// COUNT, 31-wgsl-sim2.js:332-349 (the indirect arguments are not needed: kernels read the counters)
// mode 0: only fold the G2P speed slots into vmax; 1: count = newCount (and fold); 2: count =
// min(count + y, cap); 3: wwCount = min(wwNew, wcap); 4: clear; 5: vmax = 0 and the slots cleared
// (the writeBuffer of 50-sim.js:306); 6: flip the particle buffer; 7: count += the emitted
// particles that fit below cap - 64 (a batch's emission, whose budget is applied on the device
// count); 8: modes 1 and 3 in one launch (the compaction's commit deferred to the whitewater
// commit, when no kernel between them reads the particle count). One thread per tank (per listed
// slot when `list` is given).
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void vmax_fold(const SlTank &T, bool clear) {
  uint32_t v = T.C->vmax, q[SL_VSLOTS];
#pragma unroll
  for (uint32_t k = 0; k < SL_VSLOTS; k++) q[k] = T.vslot[k * SL_VSLOT_STRIDE];
#pragma unroll
  for (uint32_t k = 0; k < SL_VSLOTS; k++) { v = max(v, q[k]); T.vslot[k * SL_VSLOT_STRIDE] = 0u; }
  T.C->vmax = clear ? 0u : v;
}
__global__ void k_count(SlTank *tanks, const uint32_t *list, int ntanks, uint32_t mode, uint32_t y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntanks) return;
  SlTank &T = tanks[list ? list[i] : i];
  SlCounters *C = T.C;
  uint32_t cap = T.cap, wcap = f2u(T.ww_cap_f);
  if (mode == 0 || mode == 1 || mode == 5 || mode == 8) vmax_fold(T, mode == 5);
  if (mode == 1 || mode == 8) {  // compaction is in place: the buffer stays current
    C->count = C->newCount;
    C->newCount = 0;
  }
  if (mode == 2) { uint32_t yy = y ? y : T.emit_total; uint32_t s = C->count + yy; C->count = s < cap ? s : cap; }
  if (mode == 3 || mode == 8) { C->wwCount = C->wwNew < wcap ? C->wwNew : wcap; C->wwNew = 0; T.ww_cur = 1 - T.ww_cur; }
  if (mode == 4) { C->count = 0; C->newCount = 0; C->wwCount = 0; C->wwNew = 0; C->impacts = 0; }
  if (mode == 6) T.cur = 1 - T.cur;
  if (mode == 7 && T.emit_total) {
    const uint32_t lim = cap > 64u ? cap - 64u : 0u, c = C->count;
    if (c < lim) C->count = c + min(T.emit_total, lim - c);
  }
  // dispatch / draw-indirect arguments for the particle and whitewater counts (31-wgsl-sim2.js:343-348)
  const uint32_t n = min(C->count, cap), w = min(C->wwCount, wcap);
  uint32_t *a = T.args;
  a[0] = (n + 255u) / 256u; a[1] = 1u; a[2] = 1u;
  a[3] = (w + 255u) / 256u; a[4] = 1u; a[5] = 1u;
  a[6] = 6u; a[7] = n; a[8] = 0u; a[9] = 0u;
  a[10] = 6u; a[11] = w; a[12] = 0u; a[13] = 0u;
}
void sl_k_count(const SlLaunch &L, uint32_t mode, uint32_t y) {
  int nt = L.w.ntanks;
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_count, dim3((nt + 63) / 64), dim3(64), 0, L.stream, L.w.tanks, L.w.list, nt, mode, y);
}

// ------------------------------------------------------------------------------------------------
// COMPACT, 31-wgsl-sim2.js:317-328, in place and deterministic. The reference appends survivors with
// an atomic counter, so any order of the survivors is one of its outcomes; here the survivors below
// the new count n' keep their slots and the dead slots below n' are filled, in ascending order, by
// the survivors at or above n', in ascending order -- only as many particles move as died there.
// (1) live count per workgroup of SL_COMPACT_ITEMS particles, (2) exclusive scan of the workgroup
// counts per tank (+ newCount = n'), (3) the dead slots below n' into a list (the P2G binning's
// particle list, unused at this point), (4) the survivors at or above n' into those slots. (3) and
// (4) return at once when nothing died.
// ------------------------------------------------------------------------------------------------
template <class A>
__global__ __launch_bounds__(SL_WG) void k_compact_count(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t lim = min(T.C->count, T.cap);
  const uint32_t base = lb * SL_COMPACT_ITEMS;
  const uint8_t *pf = T.pf[T.cur];
  static_assert(SL_COMPACT_ITEMS == 4 * SL_WG, "a thread counts four consecutive flags");
  // four flags per thread as one word where all four lie below the count (the word then lies inside
  // the flag array, whose length is the capacity)
  const uint32_t id = base + 4u * threadIdx.x;
  uint32_t c = 0;
  if (id + 4u <= lim) {
    const uint32_t w = ld32((const uint32_t *)pf, id / 4u) & 0x01010101u;
    c = __builtin_popcount(w);
  } else {
    for (uint32_t q = 0; q < 4u; q++)
      if (id + q < lim && (pf[id + q] & 1u)) c++;
  }
  __shared__ uint32_t red[SL_WG / 64];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < SL_WG / 64; w++) s += red[w];
    T.scanBlocks[lb] = s;
  }
}
void sl_k_compact_count(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_compact_count, dim3(L.w.total), dim3(SL_WG), 0, L);
}
// one workgroup per tank scans its per-workgroup counts in place (exclusive) and sets newCount
template <class A>
__global__ __launch_bounds__(1024) void k_compact_scan(A W) {
  const int t = W.list ? (int)W.list[blockIdx.x] : (int)blockIdx.x;
  const SlTank &T = sl_tank(W, t);
  const uint32_t lim = min(T.C->count, T.cap);
  const uint32_t nb = (lim + SL_COMPACT_ITEMS - 1) / SL_COMPACT_ITEMS;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry_s;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t carry = 0;
  // 16 consecutive counts per thread (four 16-byte loads, all in flight together; the allocation
  // has 64 bytes beyond the last count, and counts at or beyond nb are taken as 0), so 16 Ki counts
  // take one pass with two barriers instead of sixteen rounds of loads and barriers
  for (uint32_t base = 0; base < nb; base += 16384u) {
    const uint32_t i0 = base + 16u * threadIdx.x;
    uint32_t v[16];
    if (i0 < nb) {
      const uint4 *src = (const uint4 *)(T.scanBlocks + i0);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint4 u = src[q];
        v[4 * q] = u.x; v[4 * q + 1] = u.y; v[4 * q + 2] = u.z; v[4 * q + 3] = u.w;
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (i0 + (uint32_t)q >= nb) v[q] = 0u;
    uint32_t tot = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) tot += v[q];
    uint32_t x = tot;  // inclusive wave scan of the thread totals
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (threadIdx.x < 16) {
      const uint32_t ws = wsum[threadIdx.x];
      uint32_t acc = ws;
      for (int o = 1; o < 16; o <<= 1) {
        const uint32_t y = __shfl_up(acc, o, 16);
        if ((int)threadIdx.x >= o) acc += y;
      }
      wsum[threadIdx.x] = acc - ws;  // exclusive prefix of the wave totals
    }
    __syncthreads();
    uint32_t run = carry + wsum[wv] + x - tot;  // exclusive prefix of this thread's first count
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (i0 + (uint32_t)q < nb) T.scanBlocks[i0 + q] = run;
      run += v[q];
    }
    if (threadIdx.x == 1023) carry_s = run;  // carry + every count of the chunk
    __syncthreads();
    carry = carry_s;
    __syncthreads();  // wsum and carry_s are rewritten by the next chunk
  }
  if (threadIdx.x == 0) T.C->newCount = carry;
}
void sl_k_compact_scan(const SlLaunch &L, int ntanks) {
  if (ntanks <= 0) return;
  SL_LAUNCH(k_compact_scan, dim3(ntanks), dim3(1024), 0, L);
}
// live ranks inside a compaction workgroup: for each of its Q items per thread, the number of live
// particles before it in the workgroup (positions base + q * SL_WG + thread, in position order)
template <int Q>
__device__ __forceinline__ void compact_ranks(const uint8_t (&f)[Q], uint32_t (&rank)[Q]) {
  __shared__ uint32_t wtot[SL_COMPACT_ITEMS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t masks[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    masks[q] = __ballot((f[q] & 1u) != 0u);
    if (lane == 0) wtot[q * (SL_WG / 64) + wv] = (uint32_t)__popcll(masks[q]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int c = 0; c < SL_COMPACT_ITEMS / 64; c++) { uint32_t v = wtot[c]; wtot[c] = s; s += v; }
  }
  __syncthreads();
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int q = 0; q < Q; q++) rank[q] = wtot[q * (SL_WG / 64) + wv] + (uint32_t)__popcll(masks[q] & below);
}
// the slot after the scan entries that holds the number of survivors below n'
__device__ __forceinline__ uint32_t compact_slot_L(const SlTank &T) {
  return (T.cap + SL_COMPACT_ITEMS - 1u) / SL_COMPACT_ITEMS;
}
// (3): dead slot p < n' is the (p - live before p)-th hole
template <class A>
__global__ __launch_bounds__(SL_WG) void k_compact_holes(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t lim = min(T.C->count, T.cap), n1 = T.C->newCount;
  const uint32_t base = lb * SL_COMPACT_ITEMS;
  if (n1 >= lim || base > n1) return;  // nothing died, or no slot below n' here
  // a full workgroup below n' whose particles all live has no hole (the one that holds n' runs: it
  // records the survivors below n')
  const uint32_t nb = (lim + SL_COMPACT_ITEMS - 1) / SL_COMPACT_ITEMS;
  if (lb + 1u < nb && base + SL_COMPACT_ITEMS <= n1 && T.scanBlocks[lb + 1u] - T.scanBlocks[lb] == SL_COMPACT_ITEMS)
    return;
  constexpr int Q = SL_COMPACT_ITEMS / SL_WG;
  uint8_t f[Q];
  uint32_t rank[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const uint32_t id = base + q * SL_WG + threadIdx.x;
    f[q] = id < lim ? ld32(T.pf[T.cur], id) : (uint8_t)0;
  }
  compact_ranks<Q>(f, rank);
  const uint32_t off0 = T.scanBlocks[lb];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const uint32_t id = base + q * SL_WG + threadIdx.x, live_before = off0 + rank[q];
    if (id < n1 && !(f[q] & 1u)) T.perm[id - live_before] = id;
    if (id == n1) T.scanBlocks[compact_slot_L(T)] = live_before;
  }
}
// (4): survivor p >= n' with r survivors in [n', p) moves into the r-th hole
template <class A>
__global__ __launch_bounds__(SL_WG) void k_compact_fill(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t lim = min(T.C->count, T.cap), n1 = T.C->newCount;
  const uint32_t base = lb * SL_COMPACT_ITEMS;
  if (n1 >= lim || base >= lim || base + SL_COMPACT_ITEMS <= n1) return;
  constexpr int Q = SL_COMPACT_ITEMS / SL_WG;
  const int cur = T.cur;
  uint8_t f[Q];
  uint32_t rank[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const uint32_t id = base + q * SL_WG + threadIdx.x;
    f[q] = id < lim ? ld32(T.pf[cur], id) : (uint8_t)0;
  }
  compact_ranks<Q>(f, rank);
  const uint32_t off0 = T.scanBlocks[lb], L = T.scanBlocks[compact_slot_L(T)];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const uint32_t id = base + q * SL_WG + threadIdx.x;
    if (id < n1 || !(f[q] & 1u)) continue;  // f = 0 beyond lim
    const uint32_t d = T.perm[off0 + rank[q] - L];
    st32(T.px[cur], d, ld32(T.px[cur], id)); st32(T.py[cur], d, ld32(T.py[cur], id));
    st32(T.pz[cur], d, ld32(T.pz[cur], id)); st32(T.vx[cur], d, ld32(T.vx[cur], id));
    st32(T.vy[cur], d, ld32(T.vy[cur], id)); st32(T.vz[cur], d, ld32(T.vz[cur], id));
    st32(T.pf[cur], d, f[q]);
  }
}
void sl_k_compact_move(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_compact_holes, dim3(L.w.total), dim3(SL_WG), 0, L);
  SL_LAUNCH(k_compact_fill, dim3(L.w.total), dim3(SL_WG), 0, L);
}

// ------------------------------------------------------------------------------------------------
// WW, 31-wgsl-sim2.js:361-409: spray / foam / bubbles; survivors appended to the other buffer
// ------------------------------------------------------------------------------------------------
// dt = 0: the tank's substep dt (batched steps, where every tank runs one substep per frame)
template <class A>
__global__ __launch_bounds__(SL_WG) void k_ww(A W, float dt) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  if (!(dt > 0.0f)) dt = T.dt;
  const uint32_t wcap = f2u(T.ww_cap_f);
  uint32_t id = lb * SL_WG + threadIdx.x;
  if (id >= min(T.C->wwCount, wcap)) return;
  const int wc = T.ww_cur;
  const float4 P = T.wp[wc][id], Vv = T.wv[wc][id];
  f3 x = F3(P.x, P.y, P.z), v = F3(Vv.x, Vv.y, Vv.z);
  float life = P.w;
  const f3 g = grid_pos(T, x);
  float kind = 0.0f;
  if (in_grid(T, g)) {
    i3 c = toi3(g);
    uint32_t n = T.cellCount[cell_idx(T, c.x, c.y, c.z)];
    if (n > 0u) {
      uint32_t up = 0u;
      if (c.y + 1 < T.ny) up = T.cellCount[cell_idx(T, c.x, c.y + 1, c.z)];
      if (up > 0u) kind = 2.0f;
      else if (g.y - (float)c.y < fminf((float)n / T.ppc, 1.0f)) kind = 1.0f;
    }
  }
  if (Vv.w < 0.5f && kind > 0.5f && T.R > 0u && Vv.y < -0.3f) {
    float hx, hz;
    fp_max(T, hx, hz);
    const uint32_t R = T.R;
    float u = (x.x - T.ox) / (hx - T.ox), w = (x.z - T.oz) / (hz - T.oz);
    uint32_t i = f2u(clampf(rintf(u * (float)R), 0.0f, (float)R));
    uint32_t j = f2u(clampf(rintf(w * (float)R), 0.0f, (float)R));
    atomicAdd(&T.rsrc[j * (R + 1u) + i], f2i(clampf(4.2e-6f * -Vv.y, 0.0f, 1e-4f) * 1e9f));
  }
  if (kind == 0.0f) {
    v.y -= T.g * dt;
    v = mul3(v, 1.0f / (1.0f + 0.6f * dt));
    life -= dt * 0.7f;
  } else {
    f3 vf = F3(sample1(T, sub3(g, F3(0.0f, 0.5f, 0.5f)), 0), sample1(T, sub3(g, F3(0.5f, 0.0f, 0.5f)), 1),
               sample1(T, sub3(g, F3(0.5f, 0.5f, 0.0f)), 2));
    if (kind == 1.0f) { v = F3(vf.x, 0.2f * vf.y, vf.z); life -= dt * 0.5f; }
    else { v = add3(vf, F3(0.0f, 0.16f, 0.0f)); life -= dt * 0.3f; }
  }
  const f3 xb = x;
  x = add3(x, mul3(v, dt));
  float d = scene_sdf(T, x);
  if (d < 0.0f) x = sub3(x, mul3(scene_normal(T, x), d));
  tank_clamp(T, xb, x, v, 0.001f);
  bool outside = !in_footprint(T, x) && x.y < 0.001f;
  const bool keep = life > 0.0f && !outside && x.y > -0.5f && !in_drain(T, x);
  const uint32_t k = wave_append(&T.C->wwNew, keep);  // the survivors of the wave, one atomic
  if (keep && k < wcap) {
    T.wp[1 - wc][k] = make_float4(x.x, x.y, x.z, life);
    T.wv[1 - wc][k] = make_float4(v.x, v.y, v.z, kind);
  }
}
void sl_k_ww(const SlLaunch &L, float dt) {
  if (!L.w.total) return;
  SL_LAUNCH(k_ww, dim3(L.w.total), dim3(SL_WG), 0, L, dt);
}
