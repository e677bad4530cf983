// Capillary-gravity ripple layer, src/js/64-ripple.js:8-253.
//
// The reference runs each 2-D FFT as 2*log2(M) global-memory radix-2 Stockham passes (horizontal
// first, then vertical). Here one workgroup keeps a whole row (or a tile of CW columns) in LDS and
// runs all log2(M) passes of that direction there, so each 2-D transform reads and writes the field
// once per direction. Each output is the same butterfly a + w*b with the same f32 twiddle table, so
// the result equals the pass-by-pass transform. The elementwise kernels RSRC, RMASK and RSCALE are
// folded into the loads and stores of the transforms they feed.
#include "sl_internal.h"


#define FFT_LDS_COMPLEX 4096  // complex values per LDS buffer (two buffers = 64 KiB)
#define FFT_LOADS (FFT_LDS_COMPLEX / 256)  // loads per thread of a 256-thread FFT workgroup, at most
// the column kernel's workgroup: 1024 threads, so that its two workgroups per CU (72 KB of LDS each)
// keep 8 waves per SIMD (256 threads: ripple frame 0.118 ms, 1024: 0.107 ms at M = 1024)
#define FFT_COL_THREADS 1024
#define FFT_COL_LOADS (FFT_LDS_COMPLEX / FFT_COL_THREADS)

enum { IN_BUF = 0, IN_SRC = 1, IN_MASK = 2 };
enum { OUT_BUF = 0, OUT_SCALE = 1 };

// butterfly of one output of the Stockham pass of span ns = 2^lns over M = 2^lM points
// (64-ripple.js:44-55): e = (idx / ns) * ns/2 + idx % (ns/2), o = e + M/2, twiddle (idx % ns) * (M / ns)
__device__ __forceinline__ float2 bfly(const float2 *A, uint32_t idx, uint32_t lns, uint32_t lM, uint32_t stride,
                                       const float2 *tw, float sign) {
  const uint32_t e = ((idx >> lns) << (lns - 1u)) + (idx & ((1u << (lns - 1u)) - 1u));
  const uint32_t o = e + (1u << (lM - 1u));
  float2 a = A[e * stride], b = A[o * stride];
  float2 t = tw[(idx & ((1u << lns) - 1u)) << (lM - lns)];
  float wx = t.x, wy = sign * t.y;
  return make_float2(a.x + (wx * b.x - wy * b.y), a.y + (wx * b.y + wy * b.x));
}

// the butterfly a + w*b of 64-ripple.js:52-54 with twiddle t (its imaginary part times sign)
__device__ __forceinline__ float2 bf(float2 a, float2 b, float2 t, float sign) {
  const float wx = t.x, wy = sign * t.y;
  return make_float2(a.x + (wx * b.x - wy * b.y), a.y + (wx * b.y + wy * b.x));
}
// Two consecutive Stockham passes (spans ns = 2^lns and 2 ns, lns + 1 <= lM) in registers. Group g
// of [0, M/4) takes the first pass's inputs g + {0, 1, 2, 3} M/4: they give its outputs E, E + ns/2
// (from g, g + M/2) and E + M/2, E + M/2 + ns/2 (from g + M/4, g + 3M/4), E = (g / (ns/2)) ns +
// g % (ns/2), which are all the inputs of the second pass's outputs j, j + ns/2, j + ns, j + 3ns/2,
// j = (g / (ns/2)) 2ns + g % (ns/2). Each output is the same butterfly with the same table twiddle
// as in the pass-by-pass transform, so the result is unchanged; LDS passes and barriers halve.
__device__ __forceinline__ void bfly4(const float2 *src, float2 *dst, uint32_t g, uint32_t lns, uint32_t lM,
                                      uint32_t stride, const float2 *tw, float sign) {
  const uint32_t hs = 1u << (lns - 1u), ns = hs << 1, Q = 1u << (lM - 2u);
  const uint32_t blk = g >> (lns - 1u), r = g & (hs - 1u);
  const uint32_t E = blk * ns + r, j = blk * 2u * ns + r;
  const float2 i0 = src[g * stride], i1 = src[(g + Q) * stride], i2 = src[(g + 2u * Q) * stride];
  const float2 i3 = src[(g + 3u * Q) * stride];
  const uint32_t s1 = lM - lns, s2 = lM - lns - 1u;
  const float2 fE = bf(i0, i2, tw[(E & (ns - 1u)) << s1], sign);
  const float2 fEh = bf(i0, i2, tw[((E + hs) & (ns - 1u)) << s1], sign);
  const float2 fM = bf(i1, i3, tw[((E + 2u * Q) & (ns - 1u)) << s1], sign);
  const float2 fMh = bf(i1, i3, tw[((E + 2u * Q + hs) & (ns - 1u)) << s1], sign);
  const uint32_t m2 = 2u * ns - 1u;
  dst[j * stride] = bf(fE, fM, tw[(j & m2) << s2], sign);
  dst[(j + hs) * stride] = bf(fEh, fMh, tw[((j + hs) & m2) << s2], sign);
  dst[(j + ns) * stride] = bf(fE, fM, tw[((j + ns) & m2) << s2], sign);
  dst[(j + ns + hs) * stride] = bf(fEh, fMh, tw[((j + ns + hs) & m2) << s2], sign);
}

__device__ __forceinline__ float2 fft_input(const SlTank &T, int mode, const float2 *in, uint32_t x, uint32_t y) {
  const uint32_t M = T.M, R = T.R;
  if (mode == IN_BUF) return in[y * M + x];
  uint32_t pi = x <= R ? x : M - x, pj = y <= R ? y : M - y;
  if (mode == IN_SRC) {  // RSRC, 64-ripple.js:21-30
    float dx = T.rLx / (float)R, dz = T.rLz / (float)R;
    float v = (float)T.rsrc[pj * (R + 1u) + pi] / 1e9f / (dx * dz);
    return make_float2(v, 0.0f);
  }
  float m = T.rmask[pj * (R + 1u) + pi];  // RMASK, 64-ripple.js:95-103
  float2 f = in[y * M + x];
  return make_float2(f.x * m, f.y * m);
}

// all horizontal passes of sl_fft_rows_per_wg(M) consecutive rows (one row for M >= 1024, else
// about 1024 points, so that small fields keep every thread of the workgroup busy); dynamic LDS of
// two buffers of ldsP points (the most points of a workgroup in the launch) and the M twiddles
template <class A>
__global__ __launch_bounds__(256) void k_fft_rows(A W, int which_in, int which_out, int mode, float sign,
                                                  uint32_t ldsP) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t M = T.M, lM = T.logM, y0 = lb * sl_fft_rows_per_wg(M);
  if (y0 >= M) return;
  if (mode == IN_MASK && !T.mask_bound) return;  // tanks of a batch without the absorbing mask
  const uint32_t n = sl_fft_rows_per_wg(M) * M;  // points of the workgroup: rows y0 .. y0 + n/M - 1
  extern __shared__ float2 lds[];  // buffer b at lds + b * ldsP, then the M twiddles at lds + 2 * ldsP
  float2 *stw = lds + 2u * ldsP;
  const float2 *bufs[5] = {T.rstate, T.ra, T.rb, T.rhat, nullptr};
  float2 *obufs[5] = {T.rstate, T.ra, T.rb, T.rhat, nullptr};
  const float2 *in = bufs[which_in];
  float2 *out = obufs[which_out];
  // every load of the thread is issued before the first LDS write (n <= FFT_LDS_COMPLEX = 16 * 256)
  {
    float2 v[FFT_LOADS];
    float2 w[FFT_LOADS];
#pragma unroll
    for (int k = 0; k < FFT_LOADS; k++) {
      const uint32_t q = threadIdx.x + 256u * k;
      if (q < n) v[k] = fft_input(T, mode, in, q & (M - 1u), y0 + (q >> lM));
      if (q < M) w[k] = T.tw[q];
    }
#pragma unroll
    for (int k = 0; k < FFT_LOADS; k++) {
      const uint32_t q = threadIdx.x + 256u * k;
      if (q < n) lds[q] = v[k];
      if (q < M) stw[q] = w[k];
    }
  }
  __syncthreads();
  uint32_t cur = 0, lns = 1;
  for (; lns + 1u <= lM; lns += 2u) {  // row r of the workgroup at offset r * M of each buffer
    const float2 *src = lds + cur * ldsP;
    float2 *dst = lds + (cur ^ 1u) * ldsP;
    for (uint32_t g = threadIdx.x; g < (n >> 2); g += blockDim.x) {
      const uint32_t o = (g >> (lM - 2u)) << lM;
      bfly4(src + o, dst + o, g & ((M >> 2) - 1u), lns, lM, 1u, stw, sign);
    }
    __syncthreads();
    cur ^= 1u;
  }
  if (lns == lM) {  // odd log2 M: the last pass alone
    const float2 *src = lds + cur * ldsP;
    float2 *dst = lds + (cur ^ 1u) * ldsP;
    for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) {
      const uint32_t o = (q >> lM) << lM;
      dst[q] = bfly(src + o, q & (M - 1u), lns, lM, 1, stw, sign);
    }
    __syncthreads();
    cur ^= 1u;
  }
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) out[(size_t)y0 * M + q] = lds[cur * ldsP + q];
}

// all vertical passes of a tile of CW = 4096/M adjacent columns
template <class A>
__global__ __launch_bounds__(FFT_COL_THREADS) void k_fft_cols(A W, int which_in, int which_out, int omode,
                                                              float sign) {
  // adjacent column tiles share the 128-byte lines of every row: XCD-aware order keeps them in one L2
  SL_TANK_OF(W, xcd_remap(blockIdx.x, gridDim.x), t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t M = T.M, lM = T.logM;
  const uint32_t lCW = lM <= 6u ? lM : 12u - lM, CW = 1u << lCW;  // FFT_LDS_COMPLEX / M, at most M
  const uint32_t x0 = lb * CW;
  if (x0 >= M) return;
  if (omode == OUT_SCALE && !T.mask_bound) return;  // tanks of a batch without the absorbing mask
  __shared__ float2 lds[2][FFT_LDS_COMPLEX];
  extern __shared__ float2 stw[];  // the M twiddles (dynamic LDS sized to the largest M of the launch)
  const float2 *bufs[4] = {T.rstate, T.ra, T.rb, T.rhat};
  float2 *obufs[4] = {T.rstate, T.ra, T.rb, T.rhat};
  const float2 *in = bufs[which_in];
  float2 *out = obufs[which_out];
  const uint32_t n = CW * M;
  // LDS layout: column c at offset c, row y at y*CW (stride CW); every load of the thread is issued
  // before the first LDS write (n <= FFT_LDS_COMPLEX = FFT_COL_LOADS * FFT_COL_THREADS)
  {
    float2 v[FFT_COL_LOADS];
#pragma unroll
    for (int k = 0; k < FFT_COL_LOADS; k++) {
      const uint32_t q = threadIdx.x + FFT_COL_THREADS * k;
      if (q < n) v[k] = in[(q >> lCW) * M + x0 + (q & (CW - 1u))];
    }
    float2 w[FFT_COL_LOADS];
#pragma unroll
    for (int k = 0; k < FFT_COL_LOADS; k++) {
      const uint32_t x = threadIdx.x + FFT_COL_THREADS * k;
      if (x < M) w[k] = T.tw[x];
    }
#pragma unroll
    for (int k = 0; k < FFT_COL_LOADS; k++) {
      const uint32_t q = threadIdx.x + FFT_COL_THREADS * k;
      if (q < n) lds[0][q] = v[k];
      if (q < M) stw[q] = w[k];
    }
  }
  __syncthreads();
  int cur = 0;
  uint32_t lns = 1;
  for (; lns + 1u <= lM; lns += 2u) {
    for (uint32_t q = threadIdx.x; q < (n >> 2); q += blockDim.x) {
      const uint32_t c = q & (CW - 1u), g = q >> lCW;
      bfly4(lds[cur] + c, lds[cur ^ 1] + c, g, lns, lM, CW, stw, sign);
    }
    __syncthreads();
    cur ^= 1;
  }
  if (lns == lM) {  // odd log2 M: the last pass alone
    for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) {
      const uint32_t c = q & (CW - 1u), y = q >> lCW;
      lds[cur ^ 1][q] = bfly(lds[cur] + c, y, lns, lM, CW, stw, sign);
    }
    __syncthreads();
    cur ^= 1;
  }
  const float sc = (float)(M * M);
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) {
    const uint32_t c = q & (CW - 1u), y = q >> lCW;
    float2 v = lds[cur][q];
    if (omode == OUT_SCALE) v = make_float2(v.x / sc, v.y / sc);  // RSCALE, 64-ripple.js:127-133
    out[y * M + x0 + c] = v;
  }
}

// REVOLVE, 64-ripple.js:58-90
template <class A>
__global__ __launch_bounds__(256) void k_revolve(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t M = T.M, R = T.R;
  uint32_t id = lb * 256 + threadIdx.x;
  if (id >= M * M) return;
  int32_t i = (int32_t)(id % M), j = (int32_t)(id / M);
  int32_t mx = i < (int32_t)R ? i : i - (int32_t)M;
  int32_t mz = j < (int32_t)R ? j : j - (int32_t)M;
  if ((mx == 0 && mz == 0) || mx == -(int32_t)R || mz == -(int32_t)R) { T.rstate[id] = make_float2(0.0f, 0.0f); return; }
  const float PI = (float)3.141592653589793;
  float kx = (float)mx * PI / T.rLx, kz = (float)mz * PI / T.rLz;
  float kk = sqrtf(kx * kx + kz * kz);
  float D = fmaxf(T.rmd, 0.002f);
  float kd = fminf(kk * D, 15.0f);
  float th = 1.0f - 2.0f / (sl_exp(2.0f * kd) + 1.0f);
  float w2 = (T.g * kk + T.rsr * kk * kk * kk) * th;
  float w = sqrtf(w2);
  float gam = 2.0f * T.rnu * kk * kk + T.rfilm * kk * sqrtf(T.rnu * w * 0.125f) +
              kk * sqrtf(T.rnu * w * 0.5f) / fmaxf(sl_sinh(fminf(2.0f * kd, 30.0f)), 1e-4f);
  float wd = sqrtf(fmaxf(w2 - gam * gam, 1e-6f));
  float2 s = T.rstate[id];
  float sx = s.x, sy = s.y;
  if (T.rdt > 0.0f) {
    float c = sl_cos(wd * T.rdt), sn = sl_sin(wd * T.rdt), e = sl_exp(-gam * T.rdt);
    float nx = sx * c + (sy + gam * sx) / wd * sn;
    float ny = sy * c - (w2 * sx + gam * sy) / wd * sn;
    sx = e * nx; sy = e * ny;
  }
  float hp = 1.0f - sl_exp(-(kk * kk) / (T.rhpk * T.rhpk));
  float sw = fmaxf(1.2f * T.rLx / (float)R, 0.003f);
  float spread = sl_exp(-0.5f * kk * kk * sw * sw);
  sy += -(kk * th / T.rrho) * T.rhat[id].x * hp * spread;
  T.rstate[id] = make_float2(sx, sy);
}

// REXTRACT, 64-ripple.js:134-147 (f32 values; the f16 texture is produced on readback)
__device__ __forceinline__ void rextract_one(const SlTank &T, uint32_t id) {
  const uint32_t M = T.M, R = T.R, R1 = R + 1;
  uint32_t i = id % R1, j = id / R1;
  const float2 *f = T.rb;
  const float s = 1.0f / (float)(M * M);
  float h = f[j * M + i].x * s;
  float hl = f[j * M + (i + M - 1u) % M].x * s, hr = f[j * M + (i + 1u) % M].x * s;
  float hd = f[((j + M - 1u) % M) * M + i].x * s, hu = f[((j + 1u) % M) * M + i].x * s;
  float dx = T.rLx / (float)R, dz = T.rLz / (float)R;
  T.rtex[id] = make_float4(h, (hr - hl) / (2.0f * dx), (hu - hd) / (2.0f * dz), 0.0f);
}

// RMASKBUILD, 64-ripple.js:106-126
__device__ __forceinline__ void rmaskbuild_one(const SlTank &T, uint32_t id) {
  const uint32_t R = T.R, R1 = R + 1;
  uint32_t gx = id % R1, gy = id / R1;
  float u = (float)gx / (float)R, w = (float)gy / (float)R;
  float hix = T.ox + (float)T.nx * T.h, hiz = T.oz + (float)T.nz * T.h;
  float xx = T.ox + u * (hix - T.ox), zz = T.oz + w * (hiz - T.oz);
  float dw = (float)T.wmap;
  int32_t tx = f2i(clampf(u * dw, 0.0f, dw - 1.0f)), ty = f2i(clampf(w * dw, 0.0f, dw - 1.0f));
  float hh = T.topH[(size_t)ty * T.wmap + tx];
  float m;
  if (hh >= 0.0f) {
    float d = scene_sdf(T, F3(xx, hh - 0.002f, zz));
    float tt = clampf((d - -0.003f) / (0.006f - -0.003f), 0.0f, 1.0f);
    m = tt * tt * (3.0f - 2.0f * tt);
  } else {
    m = 0.75f;
  }
  T.rmask[id] = m;
}

// The frame's texel-space tail in one launch over the (R+1)^2 texels: REXTRACT (64-ripple.js:134-147),
// RMASKBUILD where the absorbing mask is on (:106-126; the masked transform that follows reads it), and
// the source field cleared for the next frame's deposits (:252; this frame's forward transform has
// read it already)
template <class A>
__global__ __launch_bounds__(256) void k_rtail(A W, int mask) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t R1 = T.R + 1;
  const uint32_t id = lb * 256 + threadIdx.x;
  if (id >= R1 * R1) return;
  rextract_one(T, id);
  if (mask && T.mask_bound) rmaskbuild_one(T, id);
  T.rsrc[id] = 0;
}

static SlLaunch launch_over(const SlWork &w, hipStream_t s, const SlTank *one) {
  SlLaunch L;
  L.w = w;
  L.stream = s;
  L.one = one;
  return L;
}
// buffers: 0 state, 1 a, 2 b, 3 hat
void sl_k_ripple_frame_ex(const SlWork &rows_w, const SlWork &cols_w, const SlWork &elems_w, const SlWork &tex_w,
                          hipStream_t s, int mask, uint32_t maxM, const SlTank *one) {
  if (!rows_w.total) return;
  const SlLaunch rows = launch_over(rows_w, s, one), cols = launch_over(cols_w, s, one);
  const SlLaunch elems = launch_over(elems_w, s, one), tex = launch_over(tex_w, s, one);
  // row LDS: two buffers of the most points a workgroup takes (M for M >= 1024, else M times the rows
  // per workgroup, at most 1024) and the twiddles
  const uint32_t ldsP = maxM > 1024u ? maxM : 1024u;
  const size_t rl = ((size_t)2 * ldsP + maxM) * sizeof(float2);
  const size_t cl = (size_t)maxM * sizeof(float2);      // column LDS beside the static buffers: the twiddles
  // forward transform of the mirrored source -> hat (64-ripple.js:229-232)
  SL_LAUNCH(k_fft_rows, dim3(rows_w.total), dim3(256), rl, rows, 0, 1, (int)IN_SRC, -1.0f, ldsP);
  SL_LAUNCH(k_fft_cols, dim3(cols_w.total), dim3(FFT_COL_THREADS), cl, cols, 1, 3, (int)OUT_BUF, -1.0f);
  // evolve the modes and add the kick (64-ripple.js:233)
  SL_LAUNCH(k_revolve, dim3(elems_w.total), dim3(256), 0, elems);
  // inverse transform of the state -> b (64-ripple.js:234-235)
  SL_LAUNCH(k_fft_rows, dim3(rows_w.total), dim3(256), rl, rows, 0, 1, (int)IN_BUF, 1.0f, ldsP);
  SL_LAUNCH(k_fft_cols, dim3(cols_w.total), dim3(FFT_COL_THREADS), cl, cols, 1, 2, (int)OUT_BUF, 1.0f);
  // height and slopes (64-ripple.js:236-238), the mask, the source cleared (64-ripple.js:252)
  SL_LAUNCH(k_rtail, dim3(tex_w.total), dim3(256), 0, tex, mask);
  if (mask) {
    // absorb: mask the spatial field, transform back into the state, rescale (64-ripple.js:239-250)
    SL_LAUNCH(k_fft_rows, dim3(rows_w.total), dim3(256), rl, rows, 2, 1, (int)IN_MASK, -1.0f, ldsP);
    SL_LAUNCH(k_fft_cols, dim3(cols_w.total), dim3(FFT_COL_THREADS), cl, cols, 1, 0, (int)OUT_SCALE, -1.0f);
  }
}

// ------------------------------------------------------------------------------------------------
// Top-surface height map: the renderer's water-seen-from-the-sun pass (63-wgsl-water.js:5-33, drawn by
// 66-draw.js:122-126 over the grid footprint of 66-draw.js:45-46) and its four separable blur passes
// (63-wgsl-water.js:35-55, 66-draw.js:127-131). The rasteriser's coverage is stated with explicit
// texel centres: texel (ix, iy) of the W x W map lies at xz = wmap.xy + ((i + 0.5) / W) * wmap.zw; a
// live particle at or below tank height + 0.01 covers it when q = r^2 - |xz - p.xz|^2 > 0
// (r = 1.25 x render radius) and offers p.y + sqrt(q); the depth test keeps the largest (an atomic
// max on an order-preserving integer image of the float). Texels no particle covers hold -1.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ord_key(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord_val(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
template <class A>
__global__ __launch_bounds__(256) void k_toph_init(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t n = T.wmap * T.wmap, i = lb * 256 + threadIdx.x;
  if (!T.top_height || i >= n) return;
  ((uint32_t *)T.topH)[i] = ord_key(-1.0f);
}
template <class A>
__global__ __launch_bounds__(256) void k_toph_splat(A W) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const uint32_t id = lb * 256 + threadIdx.x;
  if (!T.top_height || id >= min(T.C->count, T.cap)) return;
  const int cur = T.cur;
  if (!(T.pf[cur][id] & 1u)) return;
  const float px = T.px[cur][id], py = T.py[cur][id], pz = T.pz[cur][id];
  if (py > T.tankH + 0.01f) return;
  const uint32_t Wm = T.wmap;
  const float dW = (float)Wm, wx = T.ox, wz = T.oz, ww = T.rLx, wd = T.rLz;
  const float r = T.radius * 1.25f;
  int ix0 = (int)floorf((px - r - wx) / ww * dW - 0.5f), ix1 = (int)ceilf((px + r - wx) / ww * dW - 0.5f);
  int iy0 = (int)floorf((pz - r - wz) / wd * dW - 0.5f), iy1 = (int)ceilf((pz + r - wz) / wd * dW - 0.5f);
  ix0 = max(ix0, 0); iy0 = max(iy0, 0);
  ix1 = min(ix1, (int)Wm - 1); iy1 = min(iy1, (int)Wm - 1);
  uint32_t *key = (uint32_t *)T.topH;
  for (int iy = iy0; iy <= iy1; iy++)
    for (int ix = ix0; ix <= ix1; ix++) {
      const float cx = wx + (((float)ix + 0.5f) / dW) * ww, cz = wz + (((float)iy + 0.5f) / dW) * wd;
      const float ddx = cx - px, ddz = cz - pz;
      const float q = r * r - (ddx * ddx + ddz * ddz);
      if (q <= 0.0f) continue;
      const uint32_t k = ord_key(py + sqrtf(q));
      uint32_t *dst = &key[(size_t)iy * Wm + ix];
      if (k > *dst) atomicMax(dst, k);  // the read only skips offers that cannot win
    }
}
// one blur pass along (dx, dy) (63-wgsl-water.js:40-55); pass 0 reads the splat's integer image
template <class A>
__global__ __launch_bounds__(256) void k_toph_blur(A W, int pass) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  const int Wm = (int)T.wmap;
  const uint32_t i = lb * 256 + threadIdx.x;
  if (!T.top_height || i >= (uint32_t)(Wm * Wm)) return;
  const float *src = (pass & 1) ? T.topB : T.topH;
  float *dst = (pass & 1) ? T.topH : T.topB;
  const int dx = (pass & 1) ? 0 : 1, dy = (pass & 1) ? 1 : 0;
  const int x = (int)(i % (uint32_t)Wm), y = (int)(i / (uint32_t)Wm);
  auto load = [&](int qx, int qy) {
    const float v = src[(size_t)qy * Wm + qx];
    return pass == 0 ? ord_val(__float_as_uint(v)) : v;
  };
  const float h0 = load(x, y);
  if (h0 < 0.0f) { dst[i] = -1.0f; return; }
  float sum = 0.0f, wsum = 0.0f;
  for (int k = -6; k <= 6; k++) {
    const int qx = min(max(x + dx * k, 0), Wm - 1), qy = min(max(y + dy * k, 0), Wm - 1);
    const float h = load(qx, qy);
    if (h < 0.0f) continue;
    const float w = sl_exp(-(float)(k * k) / 18.0f);
    sum += h * w;
    wsum += w;
  }
  dst[i] = sum / wsum;
}
void sl_k_toph(const SlWork &parts_w, const SlWork &texels_w, hipStream_t s, const SlTank *one) {
  if (!texels_w.total) return;
  const SlLaunch parts = launch_over(parts_w, s, one), texels = launch_over(texels_w, s, one);
  SL_LAUNCH(k_toph_init, dim3(texels_w.total), dim3(256), 0, texels);
  if (parts_w.total) SL_LAUNCH(k_toph_splat, dim3(parts_w.total), dim3(256), 0, parts);
  for (int pass = 0; pass < 4; pass++) SL_LAUNCH(k_toph_blur, dim3(texels_w.total), dim3(256), 0, texels, pass);
}
