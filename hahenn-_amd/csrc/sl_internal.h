// Internal definitions shared by the HIP kernels and the host driver of the suiri library.
//
// Device layout (structure-of-arrays, one tank = one simulation):
//   particles : px,py,pz,vx,vy,vz (f32) + pf (u8: bit0 alive, bit1 free-flight mark), ping-pong x2
//   whitewater: wp (float4 pos.xyz, life), wv (float4 vel.xyz, kind), ping-pong x2
//   MAC nodes : N = (nx+1)(ny+1)(nz+1), x-fastest; per component planes of N floats:
//               acc (3N u64: momentum sum << 32 | weight sum, per axis), vel (3N), velOld (3N, the P2G
//               mask as a bit pattern), frac (3N), svel (3N)
//   cells     : C = nx*ny*nz, x-fastest: cellCount (u32), divg (f32)
//   SOR arrays: red/black split ("packed"): cell (i,j,k) has colour (i+j+k)&1 and slot
//               (i>>1) + hw*(j + ny*k) with hw = (nx+1)/2, in array `colour`; pres, rhs (f32) and
//               sflag (u16 per-cell stencil flags) each hold 2*hw*ny*nz entries.
// A batch is a table of SlTank records in device memory; every stage kernel maps its workgroups onto
// the tanks through a prefix table of workgroup counts, so one launch per stage serves any number of
// tanks (a plain simulation is a batch of one).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef SL_MAX_OBJ
#define SL_MAX_OBJ 64   // as include/suiri.h
#define SL_MAX_EMIT 48
#endif
#define SL_WG 256
// particles per workgroup of the binned P2G's per-particle kernels (k_bin_count, k_bin_scatter): each
// wavefront takes SL_BIN_CHUNKS consecutive runs of 64 particles and issues all their loads at once
#define SL_BIN_CHUNKS 4
#define SL_BIN_PER_WG (SL_WG * SL_BIN_CHUNKS)
#define SL_SOR_WG 256  // slots per SOR workgroup (one per thread; 128 and 1024 measured no better)

// per-cell SOR stencil flags (sflag)
#define SF_OPEN0 0       // bits 0..5: face open (f > 0), faces -x,+x,-y,+y,-z,+z
#define SF_NUM0 6        // bits 6..11: face contributes to the numerator (open, in range, fluid neighbour)
#define SF_FLUID (1u << 12)
#define SF_SLOW (1u << 13)  // an open face has 0 < f < 1: the kernel re-reads the fractions

struct SlCounters { uint32_t count, newCount, wwCount, wwNew, vmax, impacts, pad0, pad1; };
// G2P's speed maxima go to SL_VSLOTS slots SL_VSLOT_STRIDE words apart (workgroup b to slot
// b % SL_VSLOTS), so that the per-workgroup atomics of one launch spread over many addresses; COUNT
// folds the slots into SlCounters::vmax
#define SL_VSLOTS 64u
// a signalling-NaN pattern that no float operation produces (their NaNs are quiet): velOld at a node
// without weight
#define SL_MASKED_OUT 0x7fa5a5a5u
#define SL_VSLOT_STRIDE 64u
struct SlObj { float r0[4], r1[4], r2[4], pos[4], dims[4], vel[4], angv[4]; };
struct SlEmitDev { float a[4], b[4], c[4], d[4]; uint32_t info[4]; };

struct SlTank {
  // sizes
  int32_t nx, ny, nz, hw;
  uint32_t nodes, cells, hcells, iters;  // hcells = hw*ny*nz
  // division (udiv_mag) by hw and hw*ny
  uint32_t hw_mag[3], rowl_mag[3];
  // the rest of what a SOR pass reads: with the sizes above, the record's first 88 bytes (two
  // scalar-cache lines; a batch's passes read one record per tank)
  float omega, omega1;
  float *pres, *rhs;
  uint16_t *sflag;
  uint32_t cap, ww_cap, R, M, logM, obj_count, frame, emit_k, emit_total, cur, ww_cur, wmap;
  int32_t mask_bound, whitewater_on, top_height;
  // division (udiv_mag) by nx, nx*ny, nx+1, (nx+1)*(ny+1)
  uint32_t nx_mag[3], nxy_mag[3], nx1_mag[3], nxy1_mag[3];
  float tankH;  // scene tank height (the renderer's F.tank.y)
  // params as the reference uniform (f32)
  float ox, oy, oz, h, invh, hh, dt, g, gdt, flip, ppc, drift, rc, rim;
  float ww_rate, ww_cap_f, friction, time_f, radius, R_f;
  // ripple uniform (64-ripple.js:223-226)
  float rLx, rLz, rdt, rmd, rsr, rnu, rfilm, rhpk, rrho;
  float rdepth;  // fixed ripple depth (> 0), else the mean depth of the count
  // splashCrown constants (31-wgsl-sim2.js:121-126), each the f32 rounding of the reference's
  // constant expression: pi * rho, rho * sigma (folded in double), rho, sigma, mu, K thresholds
  float sp_pirho, sp_rs, sp_rho, sp_sigma, sp_mu, sp_kjet, sp_kcrown;
  // sub-grid droplets per frame: count, seed, lognormal median and sigma, speed range, mass factor
  // f32(pi / 6 * rho)
  uint32_t drops, drop_seed;
  float drop_d50, drop_sig, drop_vmin, drop_vmax, drop_mc;
  // host-precision values for the device-side frame advance of batched steps
  double time_d, h_d, vp_d, rim_d;
  double dt_frame;  // a batched tank's own frame length (used when the batch step passes dt = 0)
  // particles
  float *px[2], *py[2], *pz[2], *vx[2], *vy[2], *vz[2];
  uint8_t *pf[2];
  float4 *wp[2], *wv[2];
  SlCounters *C;
  uint32_t *vslot;  // SL_VSLOTS speed-maximum slots (bit patterns), folded into C->vmax by COUNT
  uint32_t *args;  // 16 u32 dispatch / draw-indirect arguments (COUNT)
  // grid
  int32_t *acc;
  uint32_t *cellCount;
  // velOld carries P2G's mask: bits SL_MASKED_OUT where the node had no weight (NORMALIZE sets velOld
  // to 0 there and transfer1 skips the node), so G2P takes a corner's mask from its velOld load
  float *vel, *velOld, *frac, *svel;
  float *divg;
  // per SOR slot of a cut cell (SF_SLOW), written by the divergence set-up: its six face fractions in
  // face order and their sum over the open faces, (f0, f1, f2, f3), (f4, f5, den, 0), so a pass
  // reads one 32-byte record instead of six fractions from three node arrays
  float4 *scut;
  SlObj *objs;
  // one bit per cell (x fastest): some object's cull sphere may reach a particle that starts in the
  // cell and moves at most one cell in a substep (k_objmask); G2P skips its object loops elsewhere
  uint32_t *objMask;
  SlEmitDev *emits;
  uint32_t *scanBlocks;   // compaction per-workgroup counts / offsets
  // binned P2G: per particle its cell and rank in the cell, the particle indices in cell order,
  // per cell the start inside its scan block, per scan block its start
  uint32_t *binCell, *binRank, *perm, *cellStart, *binTop;
  // ripple
  int32_t *rsrc;
  float2 *rstate, *ra, *rb, *rhat, *tw;
  float *rmask, *topH, *topB;  // topB: blur ping-pong of the top-height map
  float4 *rtex;
};

// Work map of one launch: workgroup b belongs to tank t with start[t] <= b < start[t+1].
struct SlWork {
  SlTank *tanks;
  const uint32_t *start;  // ntanks+1 prefix of workgroup counts (device)
  int32_t ntanks;
  uint32_t total;         // start[ntanks]
  // per-tank kernels (one thread or workgroup per tank): the table slots to process, in order;
  // nullptr means slots 0 .. ntanks-1
  const uint32_t *list = nullptr;
  // per workgroup of the launch: (table slot, workgroup index within the slot), the prefix table
  // resolved once per map on the device (nullptr: resolve by binary search over `start`)
  const uint2 *wtab = nullptr;
};

// One tank by value in the kernel arguments (a launch for a single simulation): a kernel reads its
// fields one scalar-load level from the kernel arguments, against two (the table pointer, then the
// record) for a tank of the batch table. The members SL_TANK_OF reads are constants, so the lookup
// folds away.
struct SlOne {
  SlTank T;
  uint32_t total;
  static constexpr const uint2 *wtab = nullptr;
  static constexpr const uint32_t *start = nullptr;
  static constexpr const uint32_t *list = nullptr;
  static constexpr int32_t ntanks = 1;
};

#ifdef __HIPCC__
__device__ __forceinline__ const SlTank &sl_tank(const SlOne &W, int) { return W.T; }
// (the work map by value: through a reference the compiler no longer proves the record's pointers
// global, and the batch kernels fell back to flat loads -- G2P over 64 tanks 1.51 -> 1.95 ms)
__device__ __forceinline__ const SlTank &sl_tank(SlWork W, int t) { return W.tanks[t]; }
// ----------------------------------------------------------------------------------------------
// device helpers with WGSL semantics
// ----------------------------------------------------------------------------------------------
struct f3 { float x, y, z; };
struct i3 { int32_t x, y, z; };
__device__ __forceinline__ f3 F3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float len2f(float x, float y) { return sqrtf(x * x + y * y); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
// WGSL transcendentals (exp, pow, sin, cos, sinh) evaluated as the correctly rounded f32 value: the
// double-precision function rounded once to f32. That is the most accurate result the WGSL error
// bounds allow, and it does not depend on a math library's last-ulp behaviour.
__device__ __forceinline__ float sl_exp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float sl_pow(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float sl_sin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float sl_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float sl_sinh(float x) { return (float)sinh((double)x); }
__device__ __forceinline__ float sl_log(float x) { return (float)log((double)x); }
__device__ __forceinline__ int32_t f2i(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}
__device__ __forceinline__ uint32_t f2u(float f) {
  if (f != f || f <= 0.0f) return 0u;
  if (f >= 4294967296.0f) return UINT32_MAX;
  return (uint32_t)f;
}
// f2i by the hardware conversion: v_cvt_i32_f32 truncates toward zero, saturates values beyond the
// int32 range (infinities included) and maps NaN to 0 -- f2i's cases, in one instruction instead of
// the compiler's three branches (tests/probes/cvt_probe.hip checks it on the GPU over 4 Mi random
// bit patterns and the edge cases)
__device__ __forceinline__ int32_t f2i_hw(float f) {
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}
__device__ __forceinline__ i3 toi3_hw(f3 g) { i3 r; r.x = f2i_hw(g.x); r.y = f2i_hw(g.y); r.z = f2i_hw(g.z); return r; }
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ i3 toi3(f3 g) { i3 r; r.x = f2i(g.x); r.y = f2i(g.y); r.z = f2i(g.z); return r; }
__device__ __forceinline__ uint32_t hashU(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
__device__ __forceinline__ float rndu(uint32_t &s) { s = hashU(s); return (float)(s >> 8) / 16777216.0f; }

// element i of an array of fewer than 2^32 bytes at a wave-uniform base: a 32-bit byte offset, so
// the load takes the scalar-base + vector-offset form without 64-bit address arithmetic
// Every base is a device allocation of a tank: the global address space makes the access a global
// (not flat) instruction also where the compiler cannot prove it, as for pointers read from a batch's
// tank table -- a flat access counts on the LDS wait counter too, so it also waits out LDS work.
#define SL_GLOBAL __attribute__((address_space(1)))
template <class E>
__device__ __forceinline__ E ld32(const E *base, uint32_t i) {
  return *(const SL_GLOBAL E *)((const SL_GLOBAL char *)base + i * (uint32_t)sizeof(E));
}
template <class E>
__device__ __forceinline__ void st32(E *base, uint32_t i, E v) {
  *(SL_GLOBAL E *)((SL_GLOBAL char *)base + i * (uint32_t)sizeof(E)) = v;
}
// rows of an M x M ripple field (M a power of two) per workgroup of the row transform: one for
// M >= 1024, else 1024 / M (at most M), so that a workgroup of 256 threads holds about 1024 points;
// the row work map of a tank is M / sl_fft_rows_per_wg(M) workgroups
__host__ __device__ __forceinline__ uint32_t sl_fft_rows_per_wg(uint32_t M) {
  if (M >= 1024u || M == 0u) return 1u;
  const uint32_t r = 1024u / M;
  return r < M ? r : M;
}
// n / d from the triple (m, s1, s2) that sl_magic_u32 computes for d: Granlund & Montgomery,
// "Division by Invariant Integers using Multiplication" (PLDI 1994), fig. 4.1, exact for every
// 32-bit n and 1 <= d < 2^31
__device__ __forceinline__ uint32_t udiv_mag(uint32_t n, const uint32_t g[3]) {
  const uint32_t t = __umulhi(g[0], n);
  return (t + ((n - t) >> g[1])) >> g[2];
}
__device__ __forceinline__ uint32_t node_idx(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  return (uint32_t)(i + (j + k * (T.ny + 1)) * (T.nx + 1));
}
__device__ __forceinline__ uint32_t cell_idx(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  return (uint32_t)(i + (j + k * T.ny) * T.nx);
}
// slot of cell (i,j,k) in its colour array
__device__ __forceinline__ uint32_t packed_idx(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  return (uint32_t)((i >> 1) + T.hw * (j + T.ny * k));
}
__device__ __forceinline__ uint32_t colour_base(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  return ((i + j + k) & 1) ? T.hcells : 0u;
}
__device__ __forceinline__ f3 grid_pos(const SlTank &T, f3 x) {
  return mul3(sub3(x, F3(T.ox, T.oy, T.oz)), T.invh);
}
__device__ __forceinline__ bool in_grid(const SlTank &T, f3 g) {
  return g.x >= 0.0f && g.y >= 0.0f && g.z >= 0.0f && g.x < (float)T.nx && g.y < (float)T.ny &&
         g.z < (float)T.nz;
}

// key of particles outside the grid: above every cell index
__device__ __forceinline__ uint32_t sl_key_sentinel(const SlTank &T) { return T.cells; }

// Append with one atomic per wavefront: the lanes with `keep` get consecutive slots in lane order
// after the counter's old value (the reference appends with one atomicAdd per item; the order of
// appended items is unspecified there too). Every active lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool keep) {
  const uint64_t m = __ballot(keep);
  if (!m) return 0xffffffffu;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  return keep ? base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : 0xffffffffu;
}

// largest t with start[t] <= b
__device__ __forceinline__ int tank_of(const uint32_t *start, int ntanks, uint32_t b) {
  int lo = 0, hi = ntanks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}
// The table slot t of workgroup b of a launch over work map W, and b's index lb among t's workgroups:
// one scalar load from the map's resolved table, else a binary search over the prefix table.
#define SL_TANK_OF(W, b, t, lb)                                         \
  int t = 0;                                                            \
  uint32_t lb = (b);                                                    \
  if ((W).wtab) {                                                       \
    const uint2 e_ = (W).wtab[b];                                       \
    t = (int)e_.x;                                                      \
    lb = e_.y;                                                          \
  } else if ((W).ntanks > 1) {                                          \
    t = tank_of((W).start, (W).ntanks, b);                              \
    lb = (b) - (W).start[t];                                            \
  }

// Workgroup i is dispatched to XCD i % 8. Re-number workgroups so that each XCD works on one
// contiguous range (its L2 then holds the halos its neighbours in that range re-read).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t G) {
  const uint32_t X = 8;
  uint32_t xcd = i % X, pos = i / X, q = G / X, r = G % X;
  return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// ----------------------------------------------------------------------------------------------
// signed distances (30-wgsl-sim.js:40-114)
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ f3 obj_pos(const SlObj &o) { return F3(o.pos[0], o.pos[1], o.pos[2]); }
__device__ __forceinline__ f3 to_local(const SlObj &o, f3 p) {
  f3 d = sub3(p, obj_pos(o));
  return F3(dot3(F3(o.r0[0], o.r0[1], o.r0[2]), d), dot3(F3(o.r1[0], o.r1[1], o.r1[2]), d),
            dot3(F3(o.r2[0], o.r2[1], o.r2[2]), d));
}
__device__ __forceinline__ float sd_round_box(f3 p, f3 b, float r) {
  f3 q = F3(fabsf(p.x) - b.x + r, fabsf(p.y) - b.y + r, fabsf(p.z) - b.z + r);
  float l = len3(F3(fmaxf(q.x, 0.0f), fmaxf(q.y, 0.0f), fmaxf(q.z, 0.0f)));
  return l + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) - r;
}
__device__ __forceinline__ float sd_cyl(f3 p, float r, float hh) {
  float dx = len2f(p.x, p.z) - r, dy = fabsf(p.y) - hh;
  return fminf(fmaxf(dx, dy), 0.0f) + len2f(fmaxf(dx, 0.0f), fmaxf(dy, 0.0f));
}
__device__ __forceinline__ float sd_cone(f3 p, float r, float hh) {
  float qx = len2f(p.x, p.z), qy = p.y;
  float k2x = -r, k2y = 2.0f * hh;
  float cax = qx - fminf(qx, qy < 0.0f ? r : 0.0f), cay = fabsf(qy) - hh;
  float t = clampf(((0.0f - qx) * k2x + (hh - qy) * k2y) / (k2x * k2x + k2y * k2y), 0.0f, 1.0f);
  float cbx = qx - 0.0f + k2x * t, cby = qy - hh + k2y * t;
  float sg = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
  return sg * sqrtf(fminf(cax * cax + cay * cay, cbx * cbx + cby * cby));
}
__device__ __forceinline__ float shape_sdf(const SlObj &o, f3 p) {
  f3 q = to_local(o, p);
  const float *d = o.dims;
  uint32_t t = f2u(o.pos[3] + 0.5f);
  float r = 1e9f;
  switch (t) {
    case 0: r = sd_round_box(q, F3(d[0], d[1], d[2]), d[3]); break;
    case 1: r = len3(q) - d[0]; break;
    case 2: r = sd_cyl(q, d[0], d[1]); break;
    case 3: r = sd_cone(q, d[0], d[1]); break;
    case 4: r = len2f(len2f(q.x, q.z) - d[0], q.y) - d[1]; break;
    case 5: r = fmaxf(sd_cyl(q, d[0], d[1]), -sd_cyl(sub3(q, F3(0.0f, d[2], 0.0f)), d[0] - d[2], d[1])); break;
    case 6:
      r = fmaxf(sd_round_box(q, F3(d[0], d[1], d[2]), 0.002f),
                -sd_round_box(sub3(q, F3(0.0f, d[3], 0.0f)), F3(d[0] - d[3], d[1] - 0.0f, d[2] - d[3]), 0.001f));
      break;
    case 7: r = fmaxf(fmaxf(fabsf(len2f(q.y, q.z) - d[1]) - 0.5f * d[2], q.y), fabsf(q.x) - d[0]); break;
    default: break;
  }
  return r;
}
__device__ __forceinline__ float scene_sdf(const SlTank &T, f3 p) {
  float best = 1e9f;
  for (uint32_t i = 0; i < T.obj_count; i++) {
    const SlObj &o = T.objs[i];
    if (o.angv[3] != 1.0f) continue;
    if (len3(sub3(p, obj_pos(o))) - o.vel[3] > best) continue;
    best = fminf(best, shape_sdf(o, p));
  }
  return best;
}
__device__ __forceinline__ f3 scene_normal(const SlTank &T, f3 p) {
  const float e = 0.0006f;
  f3 n = F3(scene_sdf(T, add3(p, F3(e, 0, 0))) - scene_sdf(T, sub3(p, F3(e, 0, 0))),
            scene_sdf(T, add3(p, F3(0, e, 0))) - scene_sdf(T, sub3(p, F3(0, e, 0))),
            scene_sdf(T, add3(p, F3(0, 0, e))) - scene_sdf(T, sub3(p, F3(0, 0, e))));
  float l = len3(n);
  return l > 1e-9f ? F3(n.x / l, n.y / l, n.z / l) : F3(0.0f, 1.0f, 0.0f);
}
__device__ __forceinline__ f3 solid_vel_at(const SlTank &T, f3 p) {
  float best = 1e9f;
  f3 v = F3(0, 0, 0);
  for (uint32_t i = 0; i < T.obj_count; i++) {
    const SlObj &o = T.objs[i];
    if (o.angv[3] != 1.0f) continue;
    if (len3(sub3(p, obj_pos(o))) - o.vel[3] > best) continue;
    float d = shape_sdf(o, p);
    if (d < best) {
      best = d;
      v = add3(F3(o.vel[0], o.vel[1], o.vel[2]), cross3(F3(o.angv[0], o.angv[1], o.angv[2]), sub3(p, obj_pos(o))));
    }
  }
  return v;
}
// The host's cull radius of object o (row pad r0.w, see cull_radius in capi.hip) is farther than
// `reach` from p, with slack for the rotation rows' rounding: then o's signed distance at p exceeds
// `reach` (for a drain, p is outside its cylinder). Compared squared: |d|^2 > c^2 in f32 gives
// |d| > c (1 - 2^-21), far inside the 2e-4 slack; a negative c (radius -inf) culls everything.
__device__ __forceinline__ bool beyond_cull(const SlObj &o, f3 p, float reach) {
  const f3 d = sub3(p, obj_pos(o));
  const float c = (o.r0[3] + reach) * 1.0002f + 1e-4f;
  return c < 0.0f || d.x * d.x + d.y * d.y + d.z * d.z > c * c;
}
// some obstacle could be within `reach` of p (otherwise scene_sdf(p) > reach)
__device__ __forceinline__ bool obstacle_within(const SlTank &T, f3 p, float reach) {
  for (uint32_t i = 0; i < T.obj_count; i++) {
    const SlObj &o = T.objs[i];
    if (o.angv[3] == 1.0f && !beyond_cull(o, p, reach)) return true;
  }
  return false;
}
__device__ __forceinline__ bool in_drain(const SlTank &T, f3 p) {
  for (uint32_t i = 0; i < T.obj_count; i++) {
    const SlObj &o = T.objs[i];
    if (o.angv[3] != 2.0f || beyond_cull(o, p, 0.0f)) continue;
    if (sd_cyl(to_local(o, p), o.dims[0], o.dims[1]) < 0.0f) return true;
  }
  return false;
}
#endif  // __HIPCC__

// ----------------------------------------------------------------------------------------------
// host-side launchers (defined in the kernel translation units)
// ----------------------------------------------------------------------------------------------
struct SlLaunch {
  SlWork w;         // tanks + work map
  hipStream_t stream;
  // a single simulation's host mirror of its record, equal to the device record for every launch
  // made through this SlLaunch: the kernels that SL_LAUNCH starts take it by value (SlOne)
  const SlTank *one = nullptr;
};
// launches kernel template K (a template over its work argument: SlWork or SlOne) over L, with the
// tank by value when L carries one
#define SL_LAUNCH(K, grid, block, lds, L, ...)                                                   \
  do {                                                                                            \
    if ((L).one) {                                                                                \
      SlOne o_;                                                                                   \
      o_.T = *(L).one;                                                                            \
      o_.total = (L).w.total;                                                                     \
      hipLaunchKernelGGL(K<SlOne>, grid, block, lds, (L).stream, o_, ##__VA_ARGS__);              \
    } else {                                                                                      \
      hipLaunchKernelGGL(K<SlWork>, grid, block, lds, (L).stream, (L).w, ##__VA_ARGS__);          \
    }                                                                                             \
  } while (0)
// grid stages (work map in units of SL_WG-wide workgroups over the stage's index space)
void sl_k_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
// cell counts cleared, and count = min(count + add, cap) (a substep's deferred emission commit)
void sl_k_clear_cells_commit(hipStream_t s, uint32_t *cc, size_t n, SlCounters *C, uint32_t add, uint32_t cap);
void sl_k_geom(const SlLaunch &L);
void sl_k_objmask(const SlLaunch &L);  // over the cells
void sl_k_p2g(const SlLaunch &L, int count_only, int variant);
int sl_p2g_chunk(int variant);
// binned P2G: particles (SL_WG per workgroup), scan blocks (one workgroup each), tanks (one workgroup
// each), tiles (one workgroup each)
#define SL_P2G_BINNED 8
#define BIN_SCAN 4096  // cells per scan block (the block prefix lives in binTop)
#define BIN_TX 8       // cell tile of the binned P2G
#define BIN_TY 8
#define BIN_TZ 4
uint32_t sl_bin_tiles(int nx, int ny, int nz);
uint32_t sl_bin_blocks(uint32_t cells);
void sl_k_p2g_binned(const SlLaunch &parts, const SlLaunch &blocks, const SlLaunch &tanks, const SlLaunch &tiles);
void sl_k_normforce(const SlLaunch &L);
void sl_k_divsetup(const SlLaunch &L);
void sl_k_sor(const SlLaunch &L, uint32_t parity, uint32_t pass);
// what a SOR colour pass of one tank reads, passed by value (the single-tank path)
struct SlSorRec {
  float *pres;
  const float *rhs;
  const uint16_t *sflag;
  const float4 *scut;
  int32_t nx, ny, hw;
  uint32_t nodes, hcells;
  uint32_t hw_mag[3], rowl_mag[3];
  float omega, omega1;
};
SlSorRec sl_sor_rec(const SlTank &T);  // from the host mirror (its pointers are device addresses)
void sl_k_sor_rec(hipStream_t stream, const SlSorRec &R, uint32_t parity);
void sl_k_project(const SlLaunch &L);

// Converged pressure solve (k_mgpcg.hip): multigrid-preconditioned conjugate gradients on the system
// the SOR sweeps relax. One level of the hierarchy: X x Y x Z cells (x fastest), per cell the couplings
// on its lower x, y, z faces and its diagonal (0: not an unknown), the level's right-hand side and
// solution (single precision).
#define SL_MG_MAXLEV 16
#define SL_MG_COARSE 1024  // cells of the coarsest level at most (solved in one workgroup)
struct SlMgLevel {
  int32_t X, Y, Z;
  uint32_t cells;
  uint32_t magX[3], magXY[3];  // udiv_mag by X and by X*Y
  float4 *A;
  uint16_t *F;  // per cell: MGF_* bits
  float *b, *x;
};
// per-cell flags of a level: coupling to the neighbour across face q (order -x, +x, -y, +y, -z, +z),
// open face q (level 0: diag = number of open faces when no open face is fractional), fractional
// faces present (use the records), unknown, red ((i + j + k) even)
#define MGF_OPEN0 0
#define MGF_NUM0 6
#define MGF_SLOW (1u << 12)
#define MGF_ACTIVE (1u << 13)
#define MGF_RED (1u << 14)
struct SlMg {
  int nlev = 0;
  SlMgLevel lev[SL_MG_MAXLEV];
  // level-0 conjugate-gradient vectors (double): solution (kept between substeps as the warm start),
  // right-hand side, residual, direction, matrix times direction
  double *xs = nullptr, *bs = nullptr, *r = nullptr, *p[2] = {nullptr, nullptr}, *q = nullptr;
  double *part = nullptr;   // per-workgroup partial sums / maxima
  float *dlo = nullptr;     // per cut cell: exact sum of its open fractions minus the level-0 record's diagonal
  double *scal = nullptr;   // rho, alpha, beta, max|r| of the last check, max|r| at the start
  uint32_t *ctl = nullptr;  // done, iterations, breakdown
  double rtol = 0;          // tolerance on max |b - A x| (h^2 times the divergence tolerance in 1/s)
  uint32_t max_iters = 0;
  int coarse_sweeps = 8;
  float over = 1.6f;  // coarse-correction factor of the V-cycle
};
uint32_t sl_mg_parts(uint32_t cells);
// enqueues the whole solve on `st`: set-up, coarsening, max_iters CG iterations (early returns once
// converged), and the solution into the SOR pressure layout; `seed`: start from that pressure
void sl_mg_solve(hipStream_t st, const SlTank *d_tank, const SlMg &mg, int seed);
// PROJECT with the double-precision solution as the pressure
void sl_k_project_mg(hipStream_t st, const SlTank *d_tank, uint32_t nodes, const double *xs);
// particle stages
void sl_k_g2p(const SlLaunch &L, int variant);
void sl_k_emit(const SlLaunch &L, uint32_t reserve = 0);
void sl_k_count(const SlLaunch &L, uint32_t mode, uint32_t y);
void sl_k_vmax_reset(const SlLaunch &L);
void sl_k_compact_count(const SlLaunch &L);
void sl_k_compact_scan(const SlLaunch &L, int ntanks);
void sl_k_compact_move(const SlLaunch &L);
void sl_k_ww(const SlLaunch &L, float dt);
void sl_k_droplets(const SlLaunch &L);  // one thread per droplet (64 per workgroup)
void sl_k_clear_i32(const SlLaunch &L, int which);
// spatial re-ordering (rocPRIM radix sort of cell-index keys)
size_t sl_sort_temp_bytes(uint32_t n, int end_bit);
void sl_k_sort_keys(const SlLaunch &L, uint32_t n_ub, uint32_t *keys, uint32_t *vals, uint32_t key_hi,
                    uint32_t val_off);
int sl_sort_pairs(void *temp, size_t temp_bytes, uint32_t *keys, uint32_t *vals, uint32_t n, int end_bit,
                  hipStream_t stream);
void sl_k_sort_gather(const SlLaunch &L, uint32_t n_ub, const uint32_t *sorted_vals, uint32_t val_off);
int sl_k_reorder(const SlLaunch &L, uint32_t n_ub, uint32_t *keys, uint32_t *vals, void *temp, size_t temp_bytes,
                 int end_bit);
// ripple: rows = M workgroups per tank, cols = M/CW, elems = M*M/256, tex = (R+1)^2/256
// top-surface height map of the tanks with top_height set: particles (256 per workgroup), texels
// (wmap^2, 256 per workgroup)
void sl_k_toph(const SlWork &parts, const SlWork &texels, hipStream_t s, const SlTank *one = nullptr);
// maxM: the largest ripple M among the tanks of the launch (sizes the row transforms' LDS)
void sl_k_ripple_frame_ex(const SlWork &rows, const SlWork &cols, const SlWork &elems, const SlWork &tex,
                          hipStream_t s, int mask, uint32_t maxM, const SlTank *one = nullptr);
// workgroups of compaction per tank (each compaction workgroup covers SL_COMPACT_ITEMS particles)
#define SL_COMPACT_ITEMS 1024
