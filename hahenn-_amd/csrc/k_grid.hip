// Grid-stage kernels of the liquid step: face fractions, particle-to-grid transfer, velocity
// normalisation + body forces, divergence + SOR set-up, red-black SOR, pressure projection.
// Each kernel follows one (or two fused) WGSL kernels of the reference; the arithmetic of every
// output value is evaluated in the same order as the WGSL expression it restates.
#include "sl_internal.h"

// Resolve (tank, local workgroup) of workgroup b of a launch.

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
void sl_k_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n) {
  if (!n) return;
  size_t blocks = (n + SL_WG - 1) / SL_WG;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(SL_WG), 0, s, p, v, n);
}
// the cell counts cleared for P2G; the first thread also adds a substep's emitted particles to the
// count, as COUNT mode 2 would right after the emission (no kernel between them reads the count)
__global__ void k_clear_cells_commit(uint32_t *cc, size_t n, SlCounters *C, uint32_t add, uint32_t cap) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && add) {
    const uint32_t s = C->count + add;
    C->count = s < cap ? s : cap;
  }
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) cc[i] = 0u;
}
void sl_k_clear_cells_commit(hipStream_t s, uint32_t *cc, size_t n, SlCounters *C, uint32_t add, uint32_t cap) {
  size_t blocks = (n + SL_WG - 1) / SL_WG;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_clear_cells_commit, dim3((unsigned)blocks), dim3(SL_WG), 0, s, cc, n, C, add, cap);
}

// ------------------------------------------------------------------------------------------------
// GEOM: open fraction and solid velocity of every staggered face, 30-wgsl-sim.js:118-157
// ------------------------------------------------------------------------------------------------
__device__ float face_open(const SlTank &T, f3 c, int axis) {
  float a = 0.5f * T.h;
  f3 o1 = F3(0, 0, 0), o2 = F3(0, 0, 0);
  if (axis == 0) { o1 = F3(0.0f, a, 0.0f); o2 = F3(0.0f, 0.0f, a); }
  else if (axis == 1) { o1 = F3(a, 0.0f, 0.0f); o2 = F3(0.0f, 0.0f, a); }
  else { o1 = F3(a, 0.0f, 0.0f); o2 = F3(0.0f, a, 0.0f); }
  float s0 = scene_sdf(T, sub3(sub3(c, o1), o2));
  float s1 = scene_sdf(T, sub3(add3(c, o1), o2));
  float s2 = scene_sdf(T, add3(add3(c, o1), o2));
  float s3 = scene_sdf(T, add3(sub3(c, o1), o2));
  float sc = scene_sdf(T, c);
  float pos = fmaxf(s0, 0.0f) + fmaxf(s1, 0.0f) + fmaxf(s2, 0.0f) + fmaxf(s3, 0.0f) + 2.0f * fmaxf(sc, 0.0f);
  float tot = fabsf(s0) + fabsf(s1) + fabsf(s2) + fabsf(s3) + 2.0f * fabsf(sc);
  if (tot < 1e-12f) return 1.0f;
  return clampf(pos / tot, 0.0f, 1.0f);
}

template <class A>
__global__ __launch_bounds__(SL_WG) void k_geom(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t n = lb * SL_WG + threadIdx.x;
  if (n >= T.nodes) return;
  uint32_t sx = T.nx + 1, sy = T.ny + 1;
  int32_t cx = (int32_t)(n % sx), cy = (int32_t)((n / sx) % sy), cz = (int32_t)(n / (sx * sy));
  const float h = T.h;
  const f3 o = F3(T.ox, T.oy, T.oz);
  float fr[3] = {0.0f, 0.0f, 0.0f}, sv[3] = {0.0f, 0.0f, 0.0f};
  if (cy < T.ny && cz < T.nz && cx > 0 && cx < T.nx) {
    f3 p = add3(o, mul3(F3((float)cx, (float)cy + 0.5f, (float)cz + 0.5f), h));
    fr[0] = face_open(T, p, 0); sv[0] = solid_vel_at(T, p).x;
  }
  if (cx < T.nx && cz < T.nz && cy > 0) {
    if (cy >= T.ny) fr[1] = 1.0f;
    else {
      f3 p = add3(o, mul3(F3((float)cx + 0.5f, (float)cy, (float)cz + 0.5f), h));
      fr[1] = face_open(T, p, 1); sv[1] = solid_vel_at(T, p).y;
    }
  }
  if (cx < T.nx && cy < T.ny && cz > 0 && cz < T.nz) {
    f3 p = add3(o, mul3(F3((float)cx + 0.5f, (float)cy + 0.5f, (float)cz), h));
    fr[2] = face_open(T, p, 2); sv[2] = solid_vel_at(T, p).z;
  }
  const size_t N = T.nodes;
  for (int c = 0; c < 3; c++) { T.frac[c * N + n] = fr[c]; T.svel[c * N + n] = sv[c]; }
}
void sl_k_geom(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_geom, dim3(L.w.total), dim3(SL_WG), 0, L);
}

// Object mask for G2P's culling (no counterpart in the reference; it only skips work whose outcome
// it fixes). G2P reads the bit of the cell a particle starts in; when the particle started inside the
// footprint and moved at most h along every axis, its obstacle test point (the advected position) and
// its drain test point (after the wall clamp, which then moves it by less than rc along each axis)
// lie in cell c's box grown by 2 h + 2 rc on every side. Bit c is clear only when every object fails
// there: an obstacle or drain whose cull sphere (grown by rc for an obstacle) stays more than 1 mm
// from the grown box fails G2P's own cull test (beyond_cull, which the margin covers with room for
// rounding); one whose cull sphere reaches the box fails when its distance function at the box centre
// exceeds rc (obstacle; 0 for a drain's cylinder) by the box's half-diagonal and 1 mm -- every shape's
// distance changes by at most the distance moved (exact distances, or the max of such), with a rotation
// the host verified orthonormal, which is also when it gives a finite cull radius -- so G2P's collision
// test d < rc and its drain test sd < 0 fail at every test point. An object without a finite cull
// radius sets every bit. Types other than obstacle (1) and drain (2) never matter to G2P.
template <class A>
__global__ __launch_bounds__(SL_WG) void k_objmask(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t c = lb * SL_WG + threadIdx.x;
  const bool in = c < T.cells;
  bool near = false;
  if (in) {
    const uint32_t nx = (uint32_t)T.nx;
    const uint32_t kq = udiv_mag(c, T.nxy_mag), rq = c - kq * (nx * (uint32_t)T.ny), jq = udiv_mag(rq, T.nx_mag);
    const float h = T.h, pad = 2.0f * h + 2.0f * T.rc, half = 0.5f * h + pad;
    const f3 ctr = F3(T.ox + ((float)(rq - jq * nx) + 0.5f) * h, T.oy + ((float)jq + 0.5f) * h, T.oz + ((float)kq + 0.5f) * h);
    const float diag = 1.7320508f * half;  // half-diagonal of the grown box
    for (uint32_t q = 0; q < T.obj_count && !near; q++) {
      const SlObj &o = T.objs[q];
      const bool obst = o.angv[3] == 1.0f;
      if (!obst && o.angv[3] != 2.0f) continue;
      if (!(o.r0[3] < INFINITY)) { near = true; break; }
      const float reach = obst ? T.rc : 0.0f;
      const float cr = (o.r0[3] + reach) * 1.0002f + 1e-4f;
      if (cr < 0.0f) continue;
      const float dx = fmaxf(fabsf(ctr.x - o.pos[0]) - half, 0.0f);
      const float dy = fmaxf(fabsf(ctr.y - o.pos[1]) - half, 0.0f);
      const float dz = fmaxf(fabsf(ctr.z - o.pos[2]) - half, 0.0f);
      const float rr = cr + 1e-3f;
      if (dx * dx + dy * dy + dz * dz > rr * rr) continue;
      const float sd = obst ? shape_sdf(o, ctr) : sd_cyl(to_local(o, ctr), o.dims[0], o.dims[1]);
      near = !(sd > reach + diag + 1e-3f);
    }
  }
  // the wave's 64 cells start at a multiple of 64: lanes 0 and 32 store its two words
  const uint64_t m = __ballot(near);
  const int lane = __lane_id();
  if ((lane & 31) == 0 && in) T.objMask[c >> 5] = (uint32_t)(m >> lane);
}
void sl_k_objmask(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_objmask, dim3(L.w.total), dim3(SL_WG), 0, L);
}

// ------------------------------------------------------------------------------------------------
// P2G: trilinear splat of each velocity component to its 8 face nodes in 2^18 fixed point,
// 30-wgsl-sim.js:159-200. Integer atomics make the sums independent of the particle order.
// ------------------------------------------------------------------------------------------------
#define VSCALE 262144.0f
// One accumulator per (node, component): a u64 holding the momentum sum in the high word and the
// weight sum in the low word. Weight terms are non-negative, so the low word never carries into the
// high one (until 2^32, far past the i32 range the reference itself keeps); the high word then wraps
// exactly like the reference's i32 momentum sum. One 64-bit atomic replaces the reference's two.
__device__ __forceinline__ unsigned long long pack_mw(int32_t mom, int32_t wt) {
  return ((unsigned long long)(uint32_t)mom << 32) + (unsigned long long)(uint32_t)wt;
}
template <bool LDS>
__device__ __forceinline__ void splat(unsigned long long *acc, const SlTank &T, int32_t i0, int32_t j0, int32_t k0,
                                      int32_t ni, int32_t nj, f3 q, float val) {
  int32_t bx = f2i(floorf(q.x)), by = f2i(floorf(q.y)), bz = f2i(floorf(q.z));
  float fx = q.x - (float)bx, fy = q.y - (float)by, fz = q.z - (float)bz;
#pragma unroll
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = bz + dz;
    if (k < 0 || k > T.nz) continue;
    float wz = dz == 1 ? fz : 1.0f - fz;
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = by + dy;
      if (j < 0 || j > T.ny) continue;
      float wy = dy == 1 ? fy : 1.0f - fy;
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = bx + dx;
        if (i < 0 || i > T.nx) continue;
        float w = (dx == 1 ? fx : 1.0f - fx) * wy * wz;
        if (w <= 0.0f) continue;
        unsigned long long v = pack_mw(f2i(rintf(w * val * VSCALE)), f2i(rintf(w * VSCALE)));
        uint32_t e = LDS ? (uint32_t)(((k - k0) * nj + (j - j0)) * ni + (i - i0)) : node_idx(T, i, j, k);
        atomicAdd(&acc[e], v);
      }
    }
  }
}
// Workgroup-privatised P2G (30-wgsl-sim.js:159-200): a workgroup takes a chunk of CHUNK consecutive
// particles, finds the box of MAC nodes they touch, accumulates all their fixed-point contributions
// (and cell counts) in LDS with LDS atomics, then adds the box to the global accumulators with one
// global 64-bit atomic per non-zero (node, component). Integer sums are associative, so the result
// equals per-particle global atomics. Chunks whose box does not fit in LDS (scattered particles) use
// global atomics directly. Variants (chosen by measurement per workload, see sl_p2g_variant):
//   CHUNK     particles per workgroup (512, 1024, 2048)
//   U64       LDS accumulates packed u64 (one LDS atomic per node) or two i32 words (momentum, weight)
//   STRIDED   lane t takes particles base + t*PER + q (spread over the chunk) instead of base + q*256 + t
template <bool LDS64>
__device__ __forceinline__ void splat_box(void *boxp, int32_t vol, int32_t i0, int32_t j0, int32_t k0, int32_t ni,
                                          int32_t nj, const SlTank &T, f3 q, float val, int comp) {
  int32_t bx = f2i(floorf(q.x)), by = f2i(floorf(q.y)), bz = f2i(floorf(q.z));
  float fx = q.x - (float)bx, fy = q.y - (float)by, fz = q.z - (float)bz;
#pragma unroll
  for (int dz = 0; dz < 2; dz++) {
    int32_t k = bz + dz;
    if (k < 0 || k > T.nz) continue;
    float wz = dz == 1 ? fz : 1.0f - fz;
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
      int32_t j = by + dy;
      if (j < 0 || j > T.ny) continue;
      float wy = dy == 1 ? fy : 1.0f - fy;
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        int32_t i = bx + dx;
        if (i < 0 || i > T.nx) continue;
        float w = (dx == 1 ? fx : 1.0f - fx) * wy * wz;
        if (w <= 0.0f) continue;
        int32_t mo = f2i(rintf(w * val * VSCALE)), wi = f2i(rintf(w * VSCALE));
        int32_t e = ((k - k0) * nj + (j - j0)) * ni + (i - i0);
        if (LDS64) {
          atomicAdd(&((unsigned long long *)boxp)[comp * vol + e], pack_mw(mo, wi));
        } else {
          atomicAdd(&((int32_t *)boxp)[(2 * comp) * vol + e], mo);
          atomicAdd(&((int32_t *)boxp)[(2 * comp + 1) * vol + e], wi);
        }
      }
    }
  }
}

#define P2G_LDS_BYTES (60 * 1024)

template <int COUNT_ONLY, int CHUNK, bool LDS64, bool STRIDED, class A>
__global__ __launch_bounds__(SL_WG) void k_p2g(A W) {
  constexpr int PER = CHUNK / SL_WG;
  constexpr int LDS_BYTES = CHUNK <= 512 ? 44 * 1024 : P2G_LDS_BYTES;
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  __shared__ unsigned long long box[LDS_BYTES / 8];
  __shared__ int32_t bb[6];
  const uint32_t cnt = T.C->count;
  const uint32_t base = lb * CHUNK;
  if (base >= cnt) return;
  const int cur = T.cur;
  const int32_t nx = T.nx, ny = T.ny, nz = T.nz;
  f3 g[PER];
  bool ok[PER];
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
#pragma unroll
  for (int q = 0; q < PER; q++) {
    uint32_t id = STRIDED ? base + threadIdx.x * PER + q : base + q * SL_WG + threadIdx.x;
    ok[q] = false;
    if (id < cnt && (T.pf[cur][id] & 1u)) {
      g[q] = grid_pos(T, F3(T.px[cur][id], T.py[cur][id], T.pz[cur][id]));
      if (in_grid(T, g[q])) {
        ok[q] = true;
        i3 c = toi3(g[q]);
        mn[0] = min(mn[0], c.x); mn[1] = min(mn[1], c.y); mn[2] = min(mn[2], c.z);
        mx[0] = max(mx[0], c.x); mx[1] = max(mx[1], c.y); mx[2] = max(mx[2], c.z);
      }
    }
  }
  if (threadIdx.x < 6) bb[threadIdx.x] = threadIdx.x < 3 ? INT32_MAX : INT32_MIN;
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int32_t lo = mn[a], hi = mx[a];
    for (int o = 32; o > 0; o >>= 1) {
      lo = min(lo, __shfl_xor(lo, o, 64));
      hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) { atomicMin(&bb[a], lo); atomicMax(&bb[3 + a], hi); }
  }
  __syncthreads();
  if (bb[0] > bb[3]) return;  // no particle of the chunk is in the grid (uniform across the workgroup)
  // node box [c-1, c+1] around the cell box, clamped to the node range; cell box for the counts
  const int32_t i0 = max(bb[0] - 1, 0), j0 = max(bb[1] - 1, 0), k0 = max(bb[2] - 1, 0);
  const int32_t i1 = min(bb[3] + 1, nx), j1 = min(bb[4] + 1, ny), k1 = min(bb[5] + 1, nz);
  const int32_t ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
  const int64_t vol = (int64_t)ni * nj * nk;
  const int32_t ci = bb[3] - bb[0] + 1, cj = bb[4] - bb[1] + 1, ck = bb[5] - bb[2] + 1;
  const int64_t cvol = (int64_t)ci * cj * ck;
  const int64_t need = (COUNT_ONLY ? 0 : 3 * vol * 8) + cvol * 4;
  const bool use_lds = need <= LDS_BYTES;
  const size_t N = T.nodes;
  unsigned long long *gacc = (unsigned long long *)T.acc;
  uint32_t *cbox = (uint32_t *)box + (COUNT_ONLY ? 0 : 6 * vol);
  if (use_lds) {
    for (int32_t e = threadIdx.x; e < (int32_t)(need / 4); e += SL_WG) ((uint32_t *)box)[e] = 0;
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    if (!ok[q]) continue;
    uint32_t id = STRIDED ? base + threadIdx.x * PER + q : base + q * SL_WG + threadIdx.x;
    i3 c = toi3(g[q]);
    if (use_lds) atomicAdd(&cbox[((c.z - bb[2]) * cj + (c.y - bb[1])) * ci + (c.x - bb[0])], 1u);
    else atomicAdd(&T.cellCount[cell_idx(T, c.x, c.y, c.z)], 1u);
    if (COUNT_ONLY) continue;
    float vx = clampf(T.vx[cur][id], -30.0f, 30.0f);
    float vy = clampf(T.vy[cur][id], -30.0f, 30.0f);
    float vz = clampf(T.vz[cur][id], -30.0f, 30.0f);
    f3 qx = sub3(g[q], F3(0.0f, 0.5f, 0.5f)), qy = sub3(g[q], F3(0.5f, 0.0f, 0.5f)), qz = sub3(g[q], F3(0.5f, 0.5f, 0.0f));
    if (use_lds) {
      splat_box<LDS64>(box, (int32_t)vol, i0, j0, k0, ni, nj, T, qx, vx, 0);
      splat_box<LDS64>(box, (int32_t)vol, i0, j0, k0, ni, nj, T, qy, vy, 1);
      splat_box<LDS64>(box, (int32_t)vol, i0, j0, k0, ni, nj, T, qz, vz, 2);
    } else {
      splat<false>(gacc, T, 0, 0, 0, 0, 0, qx, vx);
      splat<false>(gacc + N, T, 0, 0, 0, 0, 0, qy, vy);
      splat<false>(gacc + 2 * N, T, 0, 0, 0, 0, 0, qz, vz);
    }
  }
  if (!use_lds) return;
  __syncthreads();
  if (!COUNT_ONLY) {
    for (int32_t e = threadIdx.x; e < (int32_t)(3 * vol); e += SL_WG) {
      int32_t comp = e / (int32_t)vol, l = e - comp * (int32_t)vol;
      unsigned long long v;
      if (LDS64) v = box[e];
      else v = pack_mw(((int32_t *)box)[(2 * comp) * vol + l], ((int32_t *)box)[(2 * comp + 1) * vol + l]);
      if (v == 0ull) continue;
      int32_t kk = l / (ni * nj), r = l - kk * ni * nj, jj = r / ni, ii = r - jj * ni;
      atomicAdd(&gacc[(size_t)comp * N + node_idx(T, i0 + ii, j0 + jj, k0 + kk)], v);
    }
  }
  for (int32_t e = threadIdx.x; e < (int32_t)cvol; e += SL_WG) {
    uint32_t v = cbox[e];
    if (v == 0) continue;
    int32_t kk = e / (ci * cj), r = e - kk * ci * cj, jj = r / ci, ii = r - jj * ci;
    atomicAdd(&T.cellCount[cell_idx(T, bb[0] + ii, bb[1] + jj, bb[2] + kk)], v);
  }
}

// variant table: id -> (chunk, u64 LDS, strided)
static const int kP2gChunk[8] = {512, 512, 1024, 1024, 2048, 2048, 1024, 512};
int sl_p2g_chunk(int variant) { return kP2gChunk[variant & 7]; }
template <int C, int CH, bool U, bool S>
static void launch_p2g(const SlLaunch &L) {
  if (L.one) {
    SlOne o;
    o.T = *L.one;
    o.total = L.w.total;
    hipLaunchKernelGGL((k_p2g<C, CH, U, S, SlOne>), dim3(L.w.total), dim3(SL_WG), 0, L.stream, o);
  } else {
    hipLaunchKernelGGL((k_p2g<C, CH, U, S, SlWork>), dim3(L.w.total), dim3(SL_WG), 0, L.stream, L.w);
  }
}
void sl_k_p2g(const SlLaunch &L, int count_only, int variant) {
  if (!L.w.total) return;
#define P2G_LAUNCH(C, CH, U, S) launch_p2g<C, CH, U, S>(L)
  if (count_only) { P2G_LAUNCH(1, 512, false, false); return; }
  switch (variant & 7) {
    case 0: P2G_LAUNCH(0, 512, false, false); break;
    case 1: P2G_LAUNCH(0, 512, true, false); break;
    case 2: P2G_LAUNCH(0, 1024, false, false); break;
    case 3: P2G_LAUNCH(0, 1024, true, false); break;
    case 4: P2G_LAUNCH(0, 2048, false, false); break;
    case 5: P2G_LAUNCH(0, 2048, true, false); break;
    case 6: P2G_LAUNCH(0, 1024, false, true); break;
    default: P2G_LAUNCH(0, 512, false, true); break;
  }
#undef P2G_LAUNCH
}

// ------------------------------------------------------------------------------------------------
// Binned P2G (variant SL_P2G_BINNED). The particles of every cell are listed in cell order each
// substep (count with a per-cell rank, scan, scatter of indices; the particle arrays themselves
// are not moved), then one thread per cell sums the fixed-point terms of its particles for the
// face nodes its stencil reaches in registers, adds them into an LDS window of its workgroup's
// 8x8x4-cell tile, and the window is added to the global accumulators once. Every (particle, node)
// term is the reference's round(w * v * 2^18) and round(w * 2^18) (30-wgsl-sim.js:159-200), and
// integer sums do not depend on the order, so the result equals per-particle atomics; the cost no
// longer depends on how scattered the particle order is.
// ------------------------------------------------------------------------------------------------
#define WIN_X (BIN_TX + 2)
#define WIN_Y (BIN_TY + 2)
#define WIN_Z (BIN_TZ + 2)
#define WIN (WIN_X * WIN_Y * WIN_Z)

// The cell of each live in-grid particle (P2G's own test) and its rank among the cell's particles.
// Lanes of a wavefront that hold the same cell in a run (the usual case in a spatially ordered
// particle array) take their ranks with one atomic. A wavefront takes SL_BIN_CHUNKS consecutive runs
// of 64 particles: their flag and position loads are issued together, and so are their atomics, so
// the kernel waits on memory a few times per wavefront rather than a few times per 64 particles.
template <class A>
__global__ __launch_bounds__(SL_WG) void k_bin_count(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t n = min(T.C->count, T.cap);
  const int cur = T.cur;
  const int lane = __lane_id();
  const uint32_t base_id = lb * SL_BIN_PER_WG + (threadIdx.x >> 6) * (64u * SL_BIN_CHUNKS) + (uint32_t)lane;
  uint32_t cell[SL_BIN_CHUNKS];
#pragma unroll
  for (int q = 0; q < SL_BIN_CHUNKS; q++) {
    const uint32_t id = base_id + 64u * q;
    const bool in = id < n;
    // loaded for every particle below n (the flag decides afterwards), so no load waits on another
    const uint8_t pf = in ? T.pf[cur][id] : (uint8_t)0;
    const f3 p = in ? F3(T.px[cur][id], T.py[cur][id], T.pz[cur][id]) : F3(0.0f, 0.0f, 0.0f);
    cell[q] = 0xffffffffu;
    if (pf & 1u) {
      const f3 g = grid_pos(T, p);
      if (in_grid(T, g)) {
        const i3 c = toi3(g);
        cell[q] = cell_idx(T, c.x, c.y, c.z);
      }
    }
  }
  // runs: a lane starts one unless it holds the same valid cell as the lane before it, so every run
  // is either one lane without a cell or lanes of one cell; the run's first lane takes the ranks
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  int head[SL_BIN_CHUNKS];
  uint32_t base[SL_BIN_CHUNKS];
#pragma unroll
  for (int q = 0; q < SL_BIN_CHUNKS; q++) {
    const uint32_t prev = __shfl_up(cell[q], 1, 64);
    const bool valid = cell[q] != 0xffffffffu;
    const uint64_t starts = __ballot(!valid || lane == 0 || prev != cell[q]);
    head[q] = 63 - __clzll((long long)(starts & upto));
    const uint64_t above = starts & ~upto;
    const int next = above ? __ffsll((unsigned long long)above) - 1 : 64;
    base[q] = 0;
    if (valid && head[q] == lane) base[q] = atomicAdd(&T.cellCount[cell[q]], (uint32_t)(next - lane));
  }
#pragma unroll
  for (int q = 0; q < SL_BIN_CHUNKS; q++) {
    const uint32_t b = __shfl(base[q], head[q], 64);
    const uint32_t id = base_id + 64u * q;
    if (id < n) {
      T.binCell[id] = cell[q];
      T.binRank[id] = b + (uint32_t)(lane - head[q]);
    }
  }
}

// exclusive prefix of the cell counts inside each block of BIN_SCAN cells; the block total goes to
// binTop[block] (256 threads x 16 consecutive cells)
template <class A>
__global__ __launch_bounds__(256) void k_bin_scan(A W) {
  SL_TANK_OF(W, blockIdx.x, t, blk);
  const SlTank &T = sl_tank(W, t);
  const uint32_t C = T.cells, base = blk * BIN_SCAN + threadIdx.x * 16u;
  uint32_t v[16], s = 0;
  if (base + 16u <= C) {
    const uint4 *p = (const uint4 *)(T.cellCount + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint4 u = p[q];
      v[4 * q] = u.x; v[4 * q + 1] = u.y; v[4 * q + 2] = u.z; v[4 * q + 3] = u.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = base + q < C ? T.cellCount[base + q] : 0u;
  }
#pragma unroll
  for (int q = 0; q < 16; q++) s += v[q];
  // exclusive scan of the per-thread sums over the workgroup
  __shared__ uint32_t wsum[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = s;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wv; w++) off += wsum[w];
  uint32_t run = off + x - s;
  if (threadIdx.x == 255) T.binTop[blk] = run + s;
  if (base + 16u <= C) {
    uint4 *p = (uint4 *)(T.cellStart + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      uint4 u;
      u.x = run; run += v[4 * q];
      u.y = run; run += v[4 * q + 1];
      u.z = run; run += v[4 * q + 2];
      u.w = run; run += v[4 * q + 3];
      p[q] = u;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      if (base + q < C) T.cellStart[base + q] = run;
      run += v[q];
    }
  }
}

// exclusive scan of the block totals of one tank, in place (one workgroup per tank)
template <class A>
__global__ __launch_bounds__(1024) void k_bin_top(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  (void)lb;
  const SlTank &T = sl_tank(W, t);
  const uint32_t nb = (T.cells + BIN_SCAN - 1) / BIN_SCAN;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nb ? T.binTop[i] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x < 16) {
      uint32_t ws = wsum[threadIdx.x], acc = ws;
      for (int o = 1; o < 16; o <<= 1) {
        uint32_t y = __shfl_up(acc, o, 16);
        if ((int)threadIdx.x >= o) acc += y;
      }
      wsum[threadIdx.x] = acc - ws;
    }
    __syncthreads();
    const uint32_t excl = carry + wsum[threadIdx.x >> 6] + x - v;
    if (i < nb) T.binTop[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
}

// particle index -> its slot in cell order
// SL_BIN_CHUNKS runs of 64 particles per wavefront, as k_bin_count
template <class A>
__global__ __launch_bounds__(SL_WG) void k_bin_scatter(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t n = min(T.C->count, T.cap);
  const uint32_t base_id = lb * SL_BIN_PER_WG + (threadIdx.x >> 6) * (64u * SL_BIN_CHUNKS) + (uint32_t)__lane_id();
  uint32_t cell[SL_BIN_CHUNKS], rank[SL_BIN_CHUNKS];
#pragma unroll
  for (int q = 0; q < SL_BIN_CHUNKS; q++) {
    const uint32_t id = base_id + 64u * q;
    cell[q] = id < n ? T.binCell[id] : 0xffffffffu;
    rank[q] = id < n ? T.binRank[id] : 0u;
  }
#pragma unroll
  for (int q = 0; q < SL_BIN_CHUNKS; q++) {
    if (cell[q] == 0xffffffffu) continue;
    T.perm[T.cellStart[cell[q]] + T.binTop[cell[q] / BIN_SCAN] + rank[q]] = base_id + 64u * q;
  }
}

// round(x) to i32 for |x| <= 30 * 2^18 and x not NaN: the splat terms below, whose velocity is clamped
// to [-30, 30] and whose weight lies in [0, 1], so f2i's NaN and range cases cannot occur
__device__ __forceinline__ int32_t fix_term(float x) { return (int32_t)rintf(x); }

// one particle's splat into the LDS window: one 64-bit add per corner of the packed (momentum << 32) +
// weight term, the global accumulators' format (the weight sums stay below 2^32, so no carry reaches
// the momentum half)
__device__ __forceinline__ void p2g_splat(f3 g, const float v3[3], int32_t X0, int32_t Y0, int32_t Z0,
                                          unsigned long long *wmw) {
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
    const f3 q = sub3(g, comp == 0 ? F3(0.0f, 0.5f, 0.5f) : (comp == 1 ? F3(0.5f, 0.0f, 0.5f) : F3(0.5f, 0.5f, 0.0f)));
    const float val = v3[comp];
    const int32_t bx = f2i(floorf(q.x)), by = f2i(floorf(q.y)), bz = f2i(floorf(q.z));
    const float fx = q.x - (float)bx, fy = q.y - (float)by, fz = q.z - (float)bz;
    const int e0 = ((bz - Z0) * WIN_Y + (by - Y0)) * WIN_X + (bx - X0) + comp * WIN;
#pragma unroll
    for (int dz = 0; dz < 2; dz++) {
      const float wz = dz == 1 ? fz : 1.0f - fz;
#pragma unroll
      for (int dy = 0; dy < 2; dy++) {
        const float wy = dy == 1 ? fy : 1.0f - fy;
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          const float w = (dx == 1 ? fx : 1.0f - fx) * wy * wz;
          const int e = e0 + (dz * WIN_Y + dy) * WIN_X + dx;
          atomicAdd(&wmw[e], pack_mw(fix_term(w * val * VSCALE), fix_term(w * VSCALE)));
        }
      }
    }
  }
}

// One cell's particles, the reference's splat (30-wgsl-sim.js:159-200) into the LDS node window
// of the workgroup's tile: 8 corners per component with the reference's weight products and
// rounding. One thread walks one cell, so the lanes of a wave add into different cells' corners.
// Particles are read P2G_BATCH at a time (indices clamped to the cell's last particle), so their
// loads are in flight together.
#define P2G_BATCH 4
__device__ __forceinline__ void p2g_cell(const SlTank &T, int cur, uint32_t start, uint32_t n, int32_t X0, int32_t Y0,
                                         int32_t Z0, unsigned long long *wmw) {
  for (uint32_t p0 = 0; p0 < n; p0 += P2G_BATCH) {
    uint32_t ids[P2G_BATCH];
#pragma unroll
    for (int u = 0; u < P2G_BATCH; u++) ids[u] = ld32(T.perm, start + min(p0 + (uint32_t)u, n - 1u));
    f3 g[P2G_BATCH];
    float v3[P2G_BATCH][3];
#pragma unroll
    for (int u = 0; u < P2G_BATCH; u++) {
      const uint32_t id = ids[u];
      g[u] = grid_pos(T, F3(ld32(T.px[cur], id), ld32(T.py[cur], id), ld32(T.pz[cur], id)));
      v3[u][0] = clampf(ld32(T.vx[cur], id), -30.0f, 30.0f);
      v3[u][1] = clampf(ld32(T.vy[cur], id), -30.0f, 30.0f);
      v3[u][2] = clampf(ld32(T.vz[cur], id), -30.0f, 30.0f);
    }
#pragma unroll
    for (int u = 0; u < P2G_BATCH; u++)
      if (p0 + (uint32_t)u < n) p2g_splat(g[u], v3[u], X0, Y0, Z0, wmw);
  }
}

#define P2G_CELL_MAX 12   // particles one thread walks directly; the rest of a crowded cell is chunked
#define P2G_QUEUE 512     // chunk queue of a workgroup
#define P2G_TILE_THREADS (BIN_TX * BIN_TY * BIN_TZ)  // one thread per cell of the tile
static_assert(BIN_TX == 8 && BIN_TY == 8, "the tile kernel decodes a thread as 8 x 8 cells per z layer");

template <class A>
__global__ __launch_bounds__(P2G_TILE_THREADS) void k_p2g_tiles(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  __shared__ unsigned long long wmw[3 * WIN];
  __shared__ uint32_t qstart[P2G_QUEUE], qinfo[P2G_QUEUE], qn;
  for (int e = threadIdx.x; e < 3 * WIN; e += P2G_TILE_THREADS) wmw[e] = 0ull;
  if (threadIdx.x == 0) qn = 0;
  const int ntx = (T.nx + BIN_TX - 1) / BIN_TX, nty = (T.ny + BIN_TY - 1) / BIN_TY;
  const int tx = (int)(lb % (uint32_t)ntx), ty = (int)((lb / (uint32_t)ntx) % (uint32_t)nty);
  const int tz = (int)(lb / (uint32_t)(ntx * nty));
  const int lx = threadIdx.x & 7, ly = (threadIdx.x >> 3) & 7, lz = threadIdx.x >> 6;
  i3 c;
  c.x = tx * BIN_TX + lx; c.y = ty * BIN_TY + ly; c.z = tz * BIN_TZ + lz;
  uint32_t start = 0, n = 0;
  if (c.x < T.nx && c.y < T.ny && c.z < T.nz) {
    const uint32_t cell = cell_idx(T, c.x, c.y, c.z);
    n = T.cellCount[cell];
    if (n) start = T.cellStart[cell] + T.binTop[cell / BIN_SCAN];
  }
  __syncthreads();
  const int32_t X0 = tx * BIN_TX - 1, Y0 = ty * BIN_TY - 1, Z0 = tz * BIN_TZ - 1;  // window origin
  const int cur = T.cur;
  // one thread per cell for the first P2G_CELL_MAX particles; a crowded cell queues the rest in
  // chunks that the workgroup's threads splat the same way afterwards, so a cell with hundreds of
  // particles (under an emitter) does not hold up the launch. (Summing a chunk's terms in registers
  // first takes fewer LDS atomics but 87 VGPRs for the whole kernel, 5 waves per SIMD instead of 8:
  // 0.63 against 0.57 ms per step at 256x128x128.)
  const uint32_t nmain = min(n, (uint32_t)P2G_CELL_MAX);
  if (nmain) p2g_cell(T, cur, start, nmain, X0, Y0, Z0, wmw);
  for (uint32_t o = nmain; o < n; o += P2G_CELL_MAX) {
    const uint32_t m = min(n - o, (uint32_t)P2G_CELL_MAX), k = atomicAdd(&qn, 1u);
    if (k < P2G_QUEUE) { qstart[k] = start + o; qinfo[k] = m; }
    else p2g_cell(T, cur, start + o, m, X0, Y0, Z0, wmw);  // queue full: the cell's own thread
  }
  __syncthreads();
  const uint32_t nq = min(qn, (uint32_t)P2G_QUEUE);
  for (uint32_t k = threadIdx.x; k < nq; k += P2G_TILE_THREADS) {
    p2g_cell(T, cur, qstart[k], qinfo[k], X0, Y0, Z0, wmw);
  }
  __syncthreads();
  unsigned long long *gacc = (unsigned long long *)T.acc;
  const size_t N = T.nodes;
  for (int e = threadIdx.x; e < 3 * WIN; e += P2G_TILE_THREADS) {
    const unsigned long long v = wmw[e];
    if (!v) continue;
    const int comp = e / WIN, l = e - comp * WIN;
    const int wz = l / (WIN_X * WIN_Y), r = l - wz * (WIN_X * WIN_Y), wy = r / WIN_X, wx = r - wy * WIN_X;
    const int32_t i = X0 + wx, j = Y0 + wy, k = Z0 + wz;
    if (i < 0 || i > T.nx || j < 0 || j > T.ny || k < 0 || k > T.nz) continue;  // the reference's node range
    atomicAdd(&gacc[(size_t)comp * N + node_idx(T, i, j, k)], v);
  }
}

uint32_t sl_bin_tiles(int nx, int ny, int nz) {
  return (uint32_t)((nx + BIN_TX - 1) / BIN_TX) * (uint32_t)((ny + BIN_TY - 1) / BIN_TY) *
         (uint32_t)((nz + BIN_TZ - 1) / BIN_TZ);
}
uint32_t sl_bin_blocks(uint32_t cells) { return (cells + BIN_SCAN - 1) / BIN_SCAN; }
void sl_k_p2g_binned(const SlLaunch &parts, const SlLaunch &blocks, const SlLaunch &tanks, const SlLaunch &tiles) {
  if (!parts.w.total || !tiles.w.total) return;
  SL_LAUNCH(k_bin_count, dim3(parts.w.total), dim3(SL_WG), 0, parts);
  SL_LAUNCH(k_bin_scan, dim3(blocks.w.total), dim3(256), 0, blocks);
  SL_LAUNCH(k_bin_top, dim3(tanks.w.total), dim3(1024), 0, tanks);
  SL_LAUNCH(k_bin_scatter, dim3(parts.w.total), dim3(SL_WG), 0, parts);
  SL_LAUNCH(k_p2g_tiles, dim3(tiles.w.total), dim3(P2G_TILE_THREADS), 0, tiles);
}

// ------------------------------------------------------------------------------------------------
// NORMALIZE (30-wgsl-sim.js:203-222) fused with FORCES (30-wgsl-sim.js:224-238), per node.
// The accumulators are cleared after they are read, ready for the next substep's P2G.
// ------------------------------------------------------------------------------------------------
template <class A>
__global__ __launch_bounds__(SL_WG) void k_normforce(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t n = lb * SL_WG + threadIdx.x;
  if (n >= T.nodes) return;
  const size_t N = T.nodes;
  float v[3] = {0.0f, 0.0f, 0.0f};
  unsigned long long *acc = (unsigned long long *)T.acc;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    unsigned long long a = acc[c * N + n];
    int32_t mo = (int32_t)(uint32_t)(a >> 32), wi = (int32_t)(uint32_t)a;
    float wt = (float)wi;
    float old = __uint_as_float(SL_MASKED_OUT);  // velOld: 0 and masked out, kept as the mask pattern
    if (wt > 1.0f) { v[c] = (float)mo / wt; old = v[c]; }
    T.velOld[c * N + n] = old;
    acc[c * N + n] = 0ull;
  }
  v[1] -= T.gdt;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (T.frac[c * N + n] <= 0.0f) v[c] = T.svel[c * N + n];
    T.vel[c * N + n] = v[c];
  }
}
void sl_k_normforce(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_normforce, dim3(L.w.total), dim3(SL_WG), 0, L);
}

// ------------------------------------------------------------------------------------------------
// DIVERGENCE (30-wgsl-sim.js:240-264) fused with the per-substep SOR set-up: the right-hand side
// h*h*div, the 6-face stencil flags of every cell (open faces, contributing fluid neighbours) and
// p = 0 for non-fluid cells (what every colour pass of 30-wgsl-sim.js:283 writes), in the
// red/black split layout.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ float flux(float f, float v, float s) { return f * v + (1.0f - f) * s; }

template <class A>
__global__ __launch_bounds__(SL_WG) void k_divsetup(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t c = lb * SL_WG + threadIdx.x;
  if (c >= T.cells) return;
  const int32_t nx = T.nx, ny = T.ny, nz = T.nz;
  const uint32_t kq = udiv_mag(c, T.nxy_mag), rq = c - kq * (uint32_t)(nx * ny), jq = udiv_mag(rq, T.nx_mag);
  const int32_t i = (int32_t)(rq - jq * (uint32_t)nx), j = (int32_t)jq, k = (int32_t)kq;
  const uint32_t slot = colour_base(T, i, j, k) + packed_idx(T, i, j, k);
  const uint32_t cnt = T.cellCount[c];
  if (cnt == 0u) {
    T.divg[c] = 0.0f;
    T.rhs[slot] = 0.0f;
    T.sflag[slot] = 0;
    if (T.iters > 0) T.pres[slot] = 0.0f;
    return;
  }
  const size_t N = T.nodes;
  const float *fx = T.frac, *fy = T.frac + N, *fz = T.frac + 2 * N;
  const float *vx = T.vel, *vy = T.vel + N, *vz = T.vel + 2 * N;
  const float *sx = T.svel, *sy = T.svel + N, *sz = T.svel + 2 * N;
  uint32_t n0 = node_idx(T, i, j, k), nxp = n0 + 1, nyp = n0 + (uint32_t)(nx + 1);
  uint32_t nzp = n0 + (uint32_t)((nx + 1) * (ny + 1));
  float f0x = fx[n0], f0y = fy[n0], f0z = fz[n0], f1x = fx[nxp], f1y = fy[nyp], f1z = fz[nzp];
  float d = flux(f1x, vx[nxp], sx[nxp]) - flux(f0x, vx[n0], sx[n0]);
  d += flux(f1y, vy[nyp], sy[nyp]) - flux(f0y, vy[n0], sy[n0]);
  d += flux(f1z, vz[nzp], sz[nzp]) - flux(f0z, vz[n0], sz[n0]);
  d *= T.invh;
  float crowd = (float)cnt / T.ppc - 1.0f;
  float dv = d - T.drift * fmaxf(crowd, 0.0f);
  T.divg[c] = dv;
  T.rhs[slot] = T.hh * dv;
  // stencil flags in the face order of 30-wgsl-sim.js:286-291
  const float f[6] = {f0x, f1x, f0y, f1y, f0z, f1z};
  const bool ok[6] = {i > 0, i < nx - 1, j > 0, j < ny - 1, k > 0, k < nz - 1};
  const int32_t off[6] = {-1, 1, -nx, nx, -nx * ny, nx * ny};
  uint32_t fl = SF_FLUID;
  float den = 0.0f;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    if (f[q] <= 0.0f) continue;
    den += f[q];
    fl |= 1u << (SF_OPEN0 + q);
    if (f[q] != 1.0f) fl |= SF_SLOW;
    if (ok[q] && T.cellCount[(int64_t)c + off[q]] > 0u) fl |= 1u << (SF_NUM0 + q);
  }
  T.sflag[slot] = (uint16_t)fl;
  if (fl & SF_SLOW) {
    T.scut[2 * (size_t)slot] = make_float4(f0x, f1x, f0y, f1y);
    T.scut[2 * (size_t)slot + 1] = make_float4(f0z, f1z, den, 0.0f);
  }
}
void sl_k_divsetup(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_divsetup, dim3(L.w.total), dim3(SL_WG), 0, L);
}

// ------------------------------------------------------------------------------------------------
// One colour pass of red-black SOR, 30-wgsl-sim.js:267-297, over the slots of colour `parity`.
// Neighbours of slot m in row (j,k) are slots m-1+r, m+r (x) of the same row and slot m of rows
// j+-1 and planes k+-1 in the other colour's array, with r = (parity+j+k)&1 (i = 2m + r).
// Fast path (all open faces have f = 1): num = sum of contributing neighbour pressures in face
// order, den = number of open faces -- bit-identical to the reference's f*p and f sums.
// ------------------------------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void sor_slot(const R &T, uint32_t lb, uint32_t parity) {
  const uint32_t s = lb * SL_SOR_WG + threadIdx.x;
  const uint32_t H = T.hcells;
  if (s >= H) return;
  const uint32_t hw = (uint32_t)T.hw, rowl = hw * (uint32_t)T.ny;
  const uint32_t k = udiv_mag(s, T.rowl_mag), rem = s - k * rowl, j = udiv_mag(rem, T.hw_mag), m = rem - j * hw;
  const uint32_t r = (parity + j + k) & 1u;
  const int32_t i = (int32_t)(2u * m + r);
  if (i >= T.nx) return;
  const uint16_t *sf = T.sflag + parity * H;
  const uint32_t fl = ld32(sf, s);
  if (!(fl & SF_FLUID)) return;
  float *pm = T.pres + parity * H;
  // neighbours in the other colour's array, read from a base rowl + 1 slots lower: the pressure
  // allocation has that many guard slots on both sides, so every read is inside it (slots that do
  // not contribute are masked by the flags)
  const uint32_t G = rowl + 1u, g = s + G;
  const float *pb = T.pres + (1u - parity) * H - G;
  const float p_n[6] = {ld32(pb, g + r - 1u), ld32(pb, g + r), ld32(pb, g - hw),
                        ld32(pb, g + hw), ld32(pb, g - rowl), ld32(pb, g + rowl)};
  const float b = ld32(T.rhs + parity * H, s), p = ld32((const float *)pm, s);
  // one path for cut and uncut cells: an uncut cell weighs its contributing faces by 1 (1 * p = p,
  // so its sum is the count-weighted one bit for bit) and divides by its number of open faces; a cut
  // cell takes its face fractions and their sum from its 32-byte record (written by the divergence
  // set-up in face order). A numerator bit is set only on an open face.
  float f[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
  float den = (float)__builtin_popcount(fl & 0x3fu);
  if (fl & SF_SLOW) {
    const float4 *rec = T.scut + 2 * ((size_t)parity * H + s);
    const float4 a = rec[0], c = rec[1];
    f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w; f[4] = c.x; f[5] = c.y;
    den = c.z;
  }
  float num = 0.0f;
#pragma unroll
  for (int q = 0; q < 6; q++)
    if (fl & (1u << (SF_NUM0 + q))) num += f[q] * p_n[q];
  if (den > 0.0f) {
    const float pn = (num - b) / den;
    st32(pm, s, p * T.omega1 + pn * T.omega);
  }
}
// a batch: each workgroup finds its tank in the table (passes beyond a tank's own count return)
__global__ __launch_bounds__(SL_SOR_WG) void k_sor(SlWork W, uint32_t parity, uint32_t pass) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  if (pass >= 2u * T.iters) return;
  sor_slot(T, lb, parity);
}
// one tank: what a pass reads travels in the kernel arguments, one scalar-load level less than the
// tank record in device memory (7.52 -> 6.59 us per pass at 256x128x128)
__global__ __launch_bounds__(SL_SOR_WG) void k_sor_rec(SlSorRec R, uint32_t parity) {
  sor_slot(R, xcd_remap(blockIdx.x, gridDim.x), parity);
}
void sl_k_sor(const SlLaunch &L, uint32_t parity, uint32_t pass) {
  if (!L.w.total) return;
  hipLaunchKernelGGL(k_sor, dim3(L.w.total), dim3(SL_SOR_WG), 0, L.stream, L.w, parity, pass);
}
SlSorRec sl_sor_rec(const SlTank &T) {
  SlSorRec R;
  R.pres = T.pres; R.rhs = T.rhs; R.sflag = T.sflag; R.scut = T.scut;
  R.nx = T.nx; R.ny = T.ny; R.hw = T.hw; R.nodes = T.nodes; R.hcells = T.hcells;
  for (int q = 0; q < 3; q++) { R.hw_mag[q] = T.hw_mag[q]; R.rowl_mag[q] = T.rowl_mag[q]; }
  R.omega = T.omega; R.omega1 = T.omega1;
  return R;
}
void sl_k_sor_rec(hipStream_t stream, const SlSorRec &R, uint32_t parity) {
  const uint32_t total = (R.hcells + SL_SOR_WG - 1) / SL_SOR_WG;
  if (!total) return;
  hipLaunchKernelGGL(k_sor_rec, dim3(total), dim3(SL_SOR_WG), 0, stream, R, parity);
}

// ------------------------------------------------------------------------------------------------
// PROJECT, 30-wgsl-sim.js:299-341: outside the side walls and floor is solid, above the grid is air.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ctype(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  if (j >= T.ny) return 0u;
  if (i < 0 || j < 0 || k < 0 || i >= T.nx || k >= T.nz) return 2u;
  return T.cellCount[cell_idx(T, i, j, k)] > 0u ? 1u : 0u;
}
__device__ __forceinline__ float cpres(const SlTank &T, int32_t i, int32_t j, int32_t k, uint32_t ct) {
  if (ct != 1u) return 0.0f;
  return T.pres[colour_base(T, i, j, k) + packed_idx(T, i, j, k)];
}
template <class A>
__global__ __launch_bounds__(SL_WG) void k_project(A W) {
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t n = lb * SL_WG + threadIdx.x;
  if (n >= T.nodes) return;
  const uint32_t sxn = T.nx + 1, syn = T.ny + 1;
  const uint32_t kq = udiv_mag(n, T.nxy1_mag), rq = n - kq * (sxn * syn), jq = udiv_mag(rq, T.nx1_mag);
  const int32_t cx = (int32_t)(rq - jq * sxn), cy = (int32_t)jq, cz = (int32_t)kq;
  const size_t N = T.nodes;
  const float ih = T.invh;
  const uint32_t t0 = ctype(T, cx, cy, cz);
  const float p0 = cpres(T, cx, cy, cz, t0);
  if (cy < T.ny && cz < T.nz) {
    float v = T.vel[n];
    if (T.frac[n] <= 0.0f) v = T.svel[n];
    else {
      uint32_t tm = ctype(T, cx - 1, cy, cz);
      if (tm == 1u || t0 == 1u) v -= (p0 - cpres(T, cx - 1, cy, cz, tm)) * ih;
    }
    T.vel[n] = v;
  }
  if (cx < T.nx && cz < T.nz) {
    float v = T.vel[N + n];
    if (T.frac[N + n] <= 0.0f) v = T.svel[N + n];
    else {
      uint32_t tm = ctype(T, cx, cy - 1, cz);
      if (tm == 1u || t0 == 1u) v -= (p0 - cpres(T, cx, cy - 1, cz, tm)) * ih;
    }
    T.vel[N + n] = v;
  }
  if (cx < T.nx && cy < T.ny) {
    float v = T.vel[2 * N + n];
    if (T.frac[2 * N + n] <= 0.0f) v = T.svel[2 * N + n];
    else {
      uint32_t tm = ctype(T, cx, cy, cz - 1);
      if (tm == 1u || t0 == 1u) v -= (p0 - cpres(T, cx, cy, cz - 1, tm)) * ih;
    }
    T.vel[2 * N + n] = v;
  }
}
void sl_k_project(const SlLaunch &L) {
  if (!L.w.total) return;
  SL_LAUNCH(k_project, dim3(L.w.total), dim3(SL_WG), 0, L);
}
