// Converged pressure solve (opt-in; no counterpart in the reference, whose PRESSURE pass is a fixed
// number of red-black SOR sweeps, src/js/30-wgsl-sim.js:267-297).
//
// The system is the one those sweeps relax: for every liquid cell c with an open face (den_c > 0)
//   den_c p_c - sum_q f_q p_{n_q} = -h^2 div_c
// where den_c is the sum of c's open face fractions, and the sum runs over open faces whose neighbour
// is a liquid cell inside the grid (air and the space above the grid hold p = 0). It is symmetric and
// positive (semi-)definite, so it is solved by conjugate gradients preconditioned with one multigrid
// V-cycle:
//   * every level stores per cell its couplings on the lower x, y, z faces and its diagonal (a float4
//     record; diagonal 0: not an unknown) and a flag word (which of the six neighbours it couples to,
//     unknown, colour); on level 0 a cell whose open faces all have fraction 1 (no cut face) is
//     marked fast: its diagonal is the number of open faces and every coupling is 1, so the kernels
//     read only the flag word, and the values are exactly those of the records;
//   * level l+1 halves every dimension (rounding up) and is the exact Galerkin product P^T A_l P for
//     the piecewise-constant prolongation P onto the unknowns of the 2x2x2 children: couplings are the
//     sums of the child couplings across a coarse face, the diagonal the children's diagonals minus
//     twice their internal couplings;
//   * the V-cycle smooths with one red-black Gauss-Seidel sweep before (red, black) and one after
//     (black, red) the coarse correction, which it scales by 1.6; restriction is P^T, and the coarsest
//     level (<= 1024 cells) runs symmetric sweeps in one workgroup, so the preconditioner is a fixed
//     symmetric operator. After the black half-sweep the black cells' residuals are zero, so the
//     restriction sums the red children only; the prolongation is folded into the black half-sweep
//     that follows it (black cells read their red neighbours plus the correction) -- the red
//     half-sweep after it does not read the red values it replaces;
//   * the CG vectors (solution, residual, direction, product) are double precision and the V-cycle
//     single precision: the solution is not limited by the f32 spacing of the pressure, and PROJECT
//     reads it in double precision.
// The solution persists between substeps as the warm start. Scalars (rho, alpha, beta) stay on the
// device; a done flag, set when the largest residual is at most the tolerance, turns every later
// kernel of the solve into an early return, so the host enqueues a fixed number of iterations and
// never waits.
#include "sl_internal.h"

namespace {

constexpr uint32_t MG_WG = 256;
constexpr uint32_t MG_CWG = 1024;

__device__ __forceinline__ bool mg_done(const uint32_t *ctl) { return ctl[0] != 0u; }

__device__ __forceinline__ void cell_ijk(const SlMgLevel &L, uint32_t c, uint32_t &i, uint32_t &j, uint32_t &k) {
  k = udiv_mag(c, L.magXY);
  const uint32_t r = c - k * (uint32_t)L.X * (uint32_t)L.Y;
  j = udiv_mag(r, L.magX);
  i = r - j * (uint32_t)L.X;
}

// diagonal of an unknown
__device__ __forceinline__ float diag_of(const SlMgLevel &L, uint32_t c, uint32_t fl) {
  return (fl & MGF_SLOW) ? L.A[c].w : (float)__builtin_popcount(fl & 0x3fu);
}
// the level-0 diagonal of the double-precision operator: the single-precision record plus the
// residue that makes it the exact sum of the open faces' fractions
__device__ __forceinline__ double diag0(const SlMgLevel &L, const float *dlo, uint32_t c, uint32_t fl) {
  return (fl & MGF_SLOW) ? (double)L.A[c].w + (double)dlo[c] : (double)__builtin_popcount(fl & 0x3fu);
}

// sum over the coupled neighbours n of c: w_cn * v(n), in the type R
template <typename R, typename Get>
__device__ __forceinline__ R nbr_sum(const SlMgLevel &L, uint32_t c, uint32_t fl, Get v) {
  const uint32_t X = (uint32_t)L.X, XY = X * (uint32_t)L.Y;
  R s = (R)0;
  if (!(fl & MGF_SLOW)) {
    if (fl & (1u << (MGF_NUM0 + 0))) s += v(c - 1u);
    if (fl & (1u << (MGF_NUM0 + 1))) s += v(c + 1u);
    if (fl & (1u << (MGF_NUM0 + 2))) s += v(c - X);
    if (fl & (1u << (MGF_NUM0 + 3))) s += v(c + X);
    if (fl & (1u << (MGF_NUM0 + 4))) s += v(c - XY);
    if (fl & (1u << (MGF_NUM0 + 5))) s += v(c + XY);
    return s;
  }
  const float4 a = L.A[c];
  if (fl & (1u << (MGF_NUM0 + 0))) s += (R)a.x * v(c - 1u);
  if (fl & (1u << (MGF_NUM0 + 1))) s += (R)L.A[c + 1u].x * v(c + 1u);
  if (fl & (1u << (MGF_NUM0 + 2))) s += (R)a.y * v(c - X);
  if (fl & (1u << (MGF_NUM0 + 3))) s += (R)L.A[c + X].y * v(c + X);
  if (fl & (1u << (MGF_NUM0 + 4))) s += (R)a.z * v(c - XY);
  if (fl & (1u << (MGF_NUM0 + 5))) s += (R)L.A[c + XY].z * v(c + XY);
  return s;
}

template <class T>
__device__ __forceinline__ T wg_reduce(T v, T *sh, uint32_t n, bool is_max) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t o = n / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + o]) : sh[threadIdx.x] + sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}

// ------------------------------------------------------------------------------------------------
// set-up
// ------------------------------------------------------------------------------------------------
// Level 0 records and flags from the face fractions and the cell counts; right-hand side -h^2 div; the
// warm start (the previous solution, or the SOR pressure when `seed`), 0 on cells that are not unknowns.
__global__ __launch_bounds__(MG_WG) void k_mg_setup0(const SlTank *Tp, SlMgLevel L, double *xs, double *bs,
                                                     float *dlo, uint32_t *ctl, int seed) {
  const SlTank &T = *Tp;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c == 0u) { ctl[0] = 0u; ctl[1] = 0u; ctl[2] = 0u; }
  if (c >= L.cells) return;
  uint32_t i, j, k;
  cell_ijk(L, c, i, j, k);
  uint32_t fl = ((i + j + k) & 1u) ? 0u : MGF_RED;
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (T.cellCount[c] > 0u) {
    const size_t N = T.nodes;
    const uint32_t n0 = node_idx(T, (int32_t)i, (int32_t)j, (int32_t)k);
    const float f[6] = {T.frac[n0], T.frac[n0 + 1], T.frac[N + n0], T.frac[N + n0 + (uint32_t)(T.nx + 1)],
                        T.frac[2 * N + n0], T.frac[2 * N + n0 + (uint32_t)((T.nx + 1) * (T.ny + 1))]};
    const uint32_t X = (uint32_t)L.X, XY = X * (uint32_t)L.Y;
    const bool in[6] = {i > 0u, i + 1u < X, j > 0u, j + 1u < (uint32_t)L.Y, k > 0u, k + 1u < (uint32_t)L.Z};
    const uint32_t off[6] = {c - 1u, c + 1u, c - X, c + X, c - XY, c + XY};
    double den_d = 0.0;  // exact: six single-precision fractions in [0, 1]
#pragma unroll
    for (int q = 0; q < 6; q++) {
      if (!(f[q] > 0.0f)) continue;
      den_d += (double)f[q];
      fl |= 1u << (MGF_OPEN0 + q);
      if (f[q] != 1.0f) fl |= MGF_SLOW;
      // an open face to a liquid neighbour makes that neighbour an unknown too (its den >= f > 0)
      if (in[q] && T.cellCount[off[q]] > 0u) fl |= 1u << (MGF_NUM0 + q);
    }
    const float den = (float)den_d;
    if (den > 0.0f) {
      fl |= MGF_ACTIVE;
      dlo[c] = (float)(den_d - (double)den);
      a = make_float4((fl & (1u << (MGF_NUM0 + 0))) ? f[0] : 0.0f, (fl & (1u << (MGF_NUM0 + 2))) ? f[2] : 0.0f,
                      (fl & (1u << (MGF_NUM0 + 4))) ? f[4] : 0.0f, den);
    } else {
      fl &= MGF_RED;
    }
  }
  L.A[c] = a;
  L.F[c] = (uint16_t)fl;
  if (fl & MGF_ACTIVE) {
    // the divergence of DIVERGENCE (30-wgsl-sim.js:240-264) evaluated in double from the stored
    // velocities, fractions and solid velocities: the single-precision one carries rounding of the
    // order of ulp(flux) / h from its subtractions, above the tolerances this solve is asked for
    const size_t N = T.nodes;
    const uint32_t n0 = node_idx(T, (int32_t)i, (int32_t)j, (int32_t)k);
    const uint32_t nn[6] = {n0, n0 + 1u, n0, n0 + (uint32_t)(T.nx + 1), n0,
                            n0 + (uint32_t)((T.nx + 1) * (T.ny + 1))};
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const size_t e = (size_t)(q >> 1) * N + nn[q];
      const double fq = (double)T.frac[e];
      const double fl_q = fq * (double)T.vel[e] + (1.0 - fq) * (double)T.svel[e];
      d += (q & 1) ? fl_q : -fl_q;
    }
    const double crowd = (double)T.cellCount[c] / (double)T.ppc - 1.0;
    const double dv = d / T.h_d - (double)T.drift * fmax(crowd, 0.0);
    bs[c] = -(T.h_d * T.h_d) * dv;
    if (seed) xs[c] = (double)T.pres[colour_base(T, (int32_t)i, (int32_t)j, (int32_t)k) + packed_idx(T, (int32_t)i, (int32_t)j, (int32_t)k)];
  } else {
    bs[c] = 0.0;
    xs[c] = 0.0;
  }
}

// Galerkin coarsening: level C records from level F (C = ceil(F / 2) in every dimension).
__global__ __launch_bounds__(MG_WG) void k_mg_coarsen(SlMgLevel F, SlMgLevel C) {
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c >= C.cells) return;
  uint32_t I, J, K;
  cell_ijk(C, c, I, J, K);
  const uint32_t FX = (uint32_t)F.X, FXY = FX * (uint32_t)F.Y;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f, diag = 0.0f, inner = 0.0f;
  for (uint32_t q = 0; q < 8u; q++) {
    const uint32_t a = q & 1u, b = (q >> 1) & 1u, d = q >> 2;
    const uint32_t i = 2u * I + a, j = 2u * J + b, k = 2u * K + d;
    if (i >= FX || j >= (uint32_t)F.Y || k >= (uint32_t)F.Z) continue;
    const float4 r = F.A[i + FX * j + FXY * k];
    diag += r.w;
    if (a) inner += r.x; else cx += r.x;
    if (b) inner += r.y; else cy += r.y;
    if (d) inner += r.z; else cz += r.z;
  }
  const float dc = diag - 2.0f * inner;
  // a coarse cell whose unknowns couple only among themselves (a closed pocket) has dc = 0 up to
  // rounding, and then no couplings outside either: it takes no coarse correction
  C.A[c] = dc > 1e-5f * diag ? make_float4(cx, cy, cz, dc) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// flags of a coarse level from its records (every cell takes the record path)
__global__ __launch_bounds__(MG_WG) void k_mg_flags(SlMgLevel L) {
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c >= L.cells) return;
  uint32_t i, j, k;
  cell_ijk(L, c, i, j, k);
  const uint32_t X = (uint32_t)L.X, XY = X * (uint32_t)L.Y;
  uint32_t fl = (((i + j + k) & 1u) ? 0u : MGF_RED) | MGF_SLOW;
  const float4 a = L.A[c];
  if (a.w > 0.0f) {
    fl |= MGF_ACTIVE;
    if (a.x != 0.0f) fl |= 1u << (MGF_NUM0 + 0);
    if (i + 1u < X && L.A[c + 1u].x != 0.0f) fl |= 1u << (MGF_NUM0 + 1);
    if (a.y != 0.0f) fl |= 1u << (MGF_NUM0 + 2);
    if (j + 1u < (uint32_t)L.Y && L.A[c + X].y != 0.0f) fl |= 1u << (MGF_NUM0 + 3);
    if (a.z != 0.0f) fl |= 1u << (MGF_NUM0 + 4);
    if (k + 1u < (uint32_t)L.Z && L.A[c + XY].z != 0.0f) fl |= 1u << (MGF_NUM0 + 5);
  }
  L.F[c] = (uint16_t)fl;
}

// ------------------------------------------------------------------------------------------------
// V-cycle kernels (single precision)
// ------------------------------------------------------------------------------------------------
// Gauss-Seidel half-sweep over the cells of one colour (black: `red` = 0)
__global__ __launch_bounds__(MG_WG) void k_mg_sweep(SlMgLevel L, int red, const uint32_t *ctl) {
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c >= L.cells) return;
  const uint32_t fl = L.F[c];
  if (!(fl & MGF_ACTIVE) || ((fl & MGF_RED) != 0u) != (red != 0)) return;
  const float *x = L.x;
  const float s = L.b[c] + nbr_sum<float>(L, c, fl, [&](uint32_t n) { return x[n]; });
  L.x[c] = s / diag_of(L, c, fl);
}

// b_C = sum over the red unknown children f of C of (b_f - (A x)_f) (the black children's residuals
// are zero after the black half-sweep), and the first red half-sweep of level C from zero
__global__ __launch_bounds__(MG_WG) void k_mg_restrict(SlMgLevel F, SlMgLevel C, const uint32_t *ctl) {
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c >= C.cells) return;
  const uint32_t cf = C.F[c];
  float sum = 0.0f;
  if (cf & MGF_ACTIVE) {
    uint32_t I, J, K;
    cell_ijk(C, c, I, J, K);
    const uint32_t FX = (uint32_t)F.X, FXY = FX * (uint32_t)F.Y;
    const float *x = F.x;
    // the red children of the 2x2x2 block: offsets (a, b, d) with a + b + d even
    constexpr uint32_t RA[4] = {0u, 1u, 1u, 0u}, RB[4] = {0u, 1u, 0u, 1u}, RD[4] = {0u, 0u, 1u, 1u};
    float part[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t i = 2u * I + RA[q], j = 2u * J + RB[q], k = 2u * K + RD[q];
      part[q] = 0.0f;
      if (i < FX && j < (uint32_t)F.Y && k < (uint32_t)F.Z) {
        const uint32_t f = i + FX * j + FXY * k;
        const uint32_t fl = F.F[f];
        if (fl & MGF_ACTIVE)
          part[q] = F.b[f] + nbr_sum<float>(F, f, fl, [&](uint32_t n) { return x[n]; }) - diag_of(F, f, fl) * F.x[f];
      }
    }
    sum = (part[0] + part[1]) + (part[2] + part[3]);
  }
  C.b[c] = sum;
  C.x[c] = (cf & (MGF_ACTIVE | MGF_RED)) == (MGF_ACTIVE | MGF_RED) ? sum / C.A[c].w : 0.0f;
}

// coarse correction x_f += over * x_C(f), folded into the black half-sweep: a black unknown reads each
// coupled (red) neighbour n as x_n + over * x_C(n)
__global__ __launch_bounds__(MG_WG) void k_mg_black_prolong(SlMgLevel F, SlMgLevel C, float over, const uint32_t *ctl) {
  if (mg_done(ctl)) return;
  const uint32_t f = blockIdx.x * MG_WG + threadIdx.x;
  if (f >= F.cells) return;
  const uint32_t fl = F.F[f];
  if (!(fl & MGF_ACTIVE) || (fl & MGF_RED)) return;
  uint32_t i, j, k;
  cell_ijk(F, f, i, j, k);
  const uint32_t FX = (uint32_t)F.X, FXY = FX * (uint32_t)F.Y;
  const uint32_t CX = (uint32_t)C.X, CXY = CX * (uint32_t)C.Y;
  const uint32_t P = (i >> 1) + CX * (j >> 1) + CXY * (k >> 1);
  // parent of each neighbour: the same coarse cell or the next one along that axis
  const uint32_t pn[6] = {P - ((i & 1u) ? 0u : 1u), P + (i & 1u), P - ((j & 1u) ? 0u : CX), P + ((j & 1u) ? CX : 0u),
                          P - ((k & 1u) ? 0u : CXY), P + ((k & 1u) ? CXY : 0u)};
  const uint32_t nn[6] = {f - 1u, f + 1u, f - FX, f + FX, f - FXY, f + FXY};
  const float *x = F.x, *xc = C.x;
  float s = F.b[f];
  if (!(fl & MGF_SLOW)) {
#pragma unroll
    for (int q = 0; q < 6; q++)
      if (fl & (1u << (MGF_NUM0 + q))) s += x[nn[q]] + over * xc[pn[q]];
  } else {
    const float4 a = F.A[f];
    const float w[6] = {a.x, (fl & (1u << (MGF_NUM0 + 1))) ? F.A[f + 1u].x : 0.0f, a.y,
                        (fl & (1u << (MGF_NUM0 + 3))) ? F.A[f + FX].y : 0.0f, a.z,
                        (fl & (1u << (MGF_NUM0 + 5))) ? F.A[f + FXY].z : 0.0f};
#pragma unroll
    for (int q = 0; q < 6; q++)
      if (fl & (1u << (MGF_NUM0 + q))) s += w[q] * (x[nn[q]] + over * xc[pn[q]]);
  }
  F.x[f] = s / diag_of(F, f, fl);
}

// Coarsest level in one workgroup: `sweeps` red-black sweeps then `sweeps` black-red sweeps from zero.
__global__ __launch_bounds__(MG_CWG) void k_mg_coarsest(SlMgLevel L, int sweeps, const uint32_t *ctl) {
  if (mg_done(ctl)) return;
  __shared__ float4 sA[MG_CWG];
  __shared__ float sx[MG_CWG], sb[MG_CWG];
  __shared__ uint16_t sf[MG_CWG];
  const uint32_t c = threadIdx.x;
  const bool in = c < L.cells;
  sA[c] = in ? L.A[c] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  sb[c] = in ? L.b[c] : 0.0f;
  sf[c] = in ? L.F[c] : (uint16_t)0;
  sx[c] = 0.0f;
  __syncthreads();
  const uint32_t X = (uint32_t)L.X, XY = X * (uint32_t)L.Y;
  const uint32_t fl = sf[c];
  const bool act = (fl & MGF_ACTIVE) != 0u;
  const uint32_t col = (fl & MGF_RED) ? 0u : 1u;
  const float dg = (fl & MGF_SLOW) ? sA[c].w : (float)__builtin_popcount(fl & 0x3fu);
  for (int pass = 0; pass < 4 * sweeps; pass++) {
    // passes 0 .. 2 sweeps-1: red, black, ...; then black, red, ...
    const uint32_t want = pass < 2 * sweeps ? (uint32_t)(pass & 1) : (uint32_t)((pass + 1) & 1);
    if (act && col == want) {
      float s = sb[c];
      if (!(fl & MGF_SLOW)) {
        if (fl & (1u << (MGF_NUM0 + 0))) s += sx[c - 1u];
        if (fl & (1u << (MGF_NUM0 + 1))) s += sx[c + 1u];
        if (fl & (1u << (MGF_NUM0 + 2))) s += sx[c - X];
        if (fl & (1u << (MGF_NUM0 + 3))) s += sx[c + X];
        if (fl & (1u << (MGF_NUM0 + 4))) s += sx[c - XY];
        if (fl & (1u << (MGF_NUM0 + 5))) s += sx[c + XY];
      } else {
        const float4 a = sA[c];
        if (fl & (1u << (MGF_NUM0 + 0))) s += a.x * sx[c - 1u];
        if (fl & (1u << (MGF_NUM0 + 1))) s += sA[c + 1u].x * sx[c + 1u];
        if (fl & (1u << (MGF_NUM0 + 2))) s += a.y * sx[c - X];
        if (fl & (1u << (MGF_NUM0 + 3))) s += sA[c + X].y * sx[c + X];
        if (fl & (1u << (MGF_NUM0 + 4))) s += a.z * sx[c - XY];
        if (fl & (1u << (MGF_NUM0 + 5))) s += sA[c + XY].z * sx[c + XY];
      }
      sx[c] = s / dg;
    }
    __syncthreads();
  }
  if (in) L.x[c] = sx[c];
}

// ------------------------------------------------------------------------------------------------
// conjugate gradients on level 0 (double precision vectors)
// ------------------------------------------------------------------------------------------------
// residual r (double), its single-precision copy as the V-cycle's right-hand side, the first red
// half-sweep from zero, and the per-workgroup max |r|
__device__ __forceinline__ void store_residual(const SlMgLevel &L, uint32_t c, uint32_t fl, double *r, double rc) {
  r[c] = rc;
  const float rf = (float)rc;
  L.b[c] = rf;
  L.x[c] = (fl & (MGF_ACTIVE | MGF_RED)) == (MGF_ACTIVE | MGF_RED) ? rf / diag_of(L, c, fl) : 0.0f;
}

// r = b - A x
__global__ __launch_bounds__(MG_WG) void k_cg_resid(SlMgLevel L, const double *xs, const double *bs, double *r,
                                                    double *part, const float *dlo) {
  __shared__ double sh[MG_WG];
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  double m = 0.0;
  if (c < L.cells) {
    const uint32_t fl = L.F[c];
    double rc = 0.0;
    if (fl & MGF_ACTIVE)
      rc = bs[c] + nbr_sum<double>(L, c, fl, [&](uint32_t n) { return xs[n]; }) - diag0(L, dlo, c, fl) * xs[c];
    // every cell once per solve: cells that stopped being unknowns drop what they held
    store_residual(L, c, fl, r, rc);
    m = fabs(rc);
  }
  m = wg_reduce(m, sh, MG_WG, true);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// max over the partials; done when it is at most rtol. Records the iteration count.
__global__ __launch_bounds__(MG_CWG) void k_cg_check(const double *part, uint32_t n, double *scal, uint32_t *ctl,
                                                     double rtol, uint32_t it) {
  __shared__ double sh[MG_CWG];
  if (ctl[0]) return;
  double m = 0.0;
  for (uint32_t q = threadIdx.x; q < n; q += MG_CWG) m = fmax(m, part[q]);
  m = wg_reduce(m, sh, MG_CWG, true);
  if (threadIdx.x == 0) {
    scal[3] = m;
    if (it == 0u) scal[4] = m;
    ctl[1] = it;
    if (!(m > rtol)) ctl[0] = 1u;  // also stops on NaN
  }
}

// the last red half-sweep of the V-cycle on level 0 and the per-workgroup partial of sum r z
__global__ __launch_bounds__(MG_WG) void k_mg_red_dot(SlMgLevel L, const double *r, double *part, const uint32_t *ctl) {
  __shared__ double sh[MG_WG];
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  double v = 0.0;
  if (c < L.cells) {
    const uint32_t fl = L.F[c];
    if (fl & MGF_ACTIVE) {
      float z;
      if (fl & MGF_RED) {
        const float *x = L.x;
        z = (L.b[c] + nbr_sum<float>(L, c, fl, [&](uint32_t n) { return x[n]; })) / diag_of(L, c, fl);
        L.x[c] = z;
      } else {
        z = L.x[c];
      }
      v = r[c] * (double)z;
    }
  }
  v = wg_reduce(v, sh, MG_WG, false);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// per-workgroup partial of sum r z (when level 0 is the coarsest level)
__global__ __launch_bounds__(MG_WG) void k_cg_dot(SlMgLevel L, const double *r, double *part, const uint32_t *ctl) {
  __shared__ double sh[MG_WG];
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  double v = c < L.cells && (L.F[c] & MGF_ACTIVE) ? r[c] * (double)L.x[c] : 0.0;
  v = wg_reduce(v, sh, MG_WG, false);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// rho' = sum r z; beta = rho' / rho (0 on the first call); rho = rho'
__global__ __launch_bounds__(MG_CWG) void k_cg_beta(const double *part, uint32_t n, double *scal, uint32_t *ctl,
                                                    int first) {
  __shared__ double sh[MG_CWG];
  if (ctl[0]) return;
  double v = 0.0;
  for (uint32_t q = threadIdx.x; q < n; q += MG_CWG) v += part[q];
  const double rho = wg_reduce(v, sh, MG_CWG, false);
  if (threadIdx.x == 0) {
    if (!(rho > 0.0)) { ctl[0] = 1u; ctl[2] = 1u; return; }
    scal[2] = first ? 0.0 : rho / scal[0];
    scal[0] = rho;
  }
}

// p' = z + beta p (in a second buffer, so that neighbours read the old p), q = A p', and the
// per-workgroup partial of sum p' q
__global__ __launch_bounds__(MG_WG) void k_cg_spmv(SlMgLevel L, const double *pold, double *pnew, double *q, double *part,
                                                   const double *scal, const uint32_t *ctl, const float *dlo) {
  __shared__ double sh[MG_WG];
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  double v = 0.0;
  if (c < L.cells) {
    const uint32_t fl = L.F[c];
    if (fl & MGF_ACTIVE) {
      const double beta = scal[2];
      const float *z = L.x;
      auto pv = [&](uint32_t n) { return (double)z[n] + beta * pold[n]; };
      const double pc = pv(c);
      const double qc = diag0(L, dlo, c, fl) * pc - nbr_sum<double>(L, c, fl, pv);
      pnew[c] = pc;
      q[c] = qc;
      v = pc * qc;
    }
  }
  v = wg_reduce(v, sh, MG_WG, false);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// alpha = rho / sum p q
__global__ __launch_bounds__(MG_CWG) void k_cg_alpha(const double *part, uint32_t n, double *scal, uint32_t *ctl) {
  __shared__ double sh[MG_CWG];
  if (ctl[0]) return;
  double v = 0.0;
  for (uint32_t q = threadIdx.x; q < n; q += MG_CWG) v += part[q];
  const double pq = wg_reduce(v, sh, MG_CWG, false);
  if (threadIdx.x == 0) {
    if (!(pq > 0.0)) { ctl[0] = 1u; ctl[2] = 1u; return; }
    scal[1] = scal[0] / pq;
  }
}

// x += alpha p, r -= alpha q, then as k_cg_resid
__global__ __launch_bounds__(MG_WG) void k_cg_update(SlMgLevel L, double *xs, double *r, const double *p,
                                                     const double *q, double *part, const double *scal,
                                                     const uint32_t *ctl) {
  __shared__ double sh[MG_WG];
  if (mg_done(ctl)) return;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  double m = 0.0;
  if (c < L.cells) {
    const uint32_t fl = L.F[c];
    if (fl & MGF_ACTIVE) {
      const double al = scal[1];
      xs[c] += al * p[c];
      const double rc = r[c] - al * q[c];
      store_residual(L, c, fl, r, rc);
      m = fabs(rc);
    }
  }
  m = wg_reduce(m, sh, MG_WG, true);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// the solution into the SOR pressure layout (single precision) of the unknowns, for readback and the
// warm start of the reference solver
__global__ __launch_bounds__(MG_WG) void k_mg_finish(const SlTank *Tp, SlMgLevel L, const double *xs) {
  const SlTank &T = *Tp;
  const uint32_t c = blockIdx.x * MG_WG + threadIdx.x;
  if (c >= L.cells) return;
  if (!(L.F[c] & MGF_ACTIVE)) return;
  uint32_t i, j, k;
  cell_ijk(L, c, i, j, k);
  T.pres[colour_base(T, (int32_t)i, (int32_t)j, (int32_t)k) + packed_idx(T, (int32_t)i, (int32_t)j, (int32_t)k)] = (float)xs[c];
}

// PROJECT (30-wgsl-sim.js:299-341) with the double-precision solution as the pressure: the same face
// rules; the pressure difference and its scaling by 1/h in double precision, rounded once into v.
__device__ __forceinline__ uint32_t ctype_mg(const SlTank &T, int32_t i, int32_t j, int32_t k) {
  if (j >= T.ny) return 0u;
  if (i < 0 || j < 0 || k < 0 || i >= T.nx || k >= T.nz) return 2u;
  return T.cellCount[cell_idx(T, i, j, k)] > 0u ? 1u : 0u;
}
__device__ __forceinline__ double cpres_mg(const SlTank &T, const double *xs, int32_t i, int32_t j, int32_t k,
                                           uint32_t ct) {
  return ct == 1u ? xs[cell_idx(T, i, j, k)] : 0.0;
}
__global__ __launch_bounds__(MG_WG) void k_project_mg(const SlTank *Tp, const double *xs) {
  const SlTank &T = *Tp;
  const uint32_t n = blockIdx.x * MG_WG + threadIdx.x;
  if (n >= T.nodes) return;
  const uint32_t sxn = (uint32_t)T.nx + 1u, syn = (uint32_t)T.ny + 1u;
  const uint32_t kq = udiv_mag(n, T.nxy1_mag), rq = n - kq * (sxn * syn), jq = udiv_mag(rq, T.nx1_mag);
  const int32_t cx = (int32_t)(rq - jq * sxn), cy = (int32_t)jq, cz = (int32_t)kq;
  const size_t N = T.nodes;
  const double ih = 1.0 / T.h_d;
  const uint32_t t0 = ctype_mg(T, cx, cy, cz);
  const double p0 = cpres_mg(T, xs, cx, cy, cz, t0);
  if (cy < T.ny && cz < T.nz) {
    float v = T.vel[n];
    if (T.frac[n] <= 0.0f) v = T.svel[n];
    else {
      const uint32_t tm = ctype_mg(T, cx - 1, cy, cz);
      if (tm == 1u || t0 == 1u) v = (float)((double)v - (p0 - cpres_mg(T, xs, cx - 1, cy, cz, tm)) * ih);
    }
    T.vel[n] = v;
  }
  if (cx < T.nx && cz < T.nz) {
    float v = T.vel[N + n];
    if (T.frac[N + n] <= 0.0f) v = T.svel[N + n];
    else {
      const uint32_t tm = ctype_mg(T, cx, cy - 1, cz);
      if (tm == 1u || t0 == 1u) v = (float)((double)v - (p0 - cpres_mg(T, xs, cx, cy - 1, cz, tm)) * ih);
    }
    T.vel[N + n] = v;
  }
  if (cx < T.nx && cy < T.ny) {
    float v = T.vel[2 * N + n];
    if (T.frac[2 * N + n] <= 0.0f) v = T.svel[2 * N + n];
    else {
      const uint32_t tm = ctype_mg(T, cx, cy, cz - 1);
      if (tm == 1u || t0 == 1u) v = (float)((double)v - (p0 - cpres_mg(T, xs, cx, cy, cz - 1, tm)) * ih);
    }
    T.vel[2 * N + n] = v;
  }
}

inline uint32_t blocks(uint64_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

// the V-cycle from level l on, entered with the first red half-sweep of level l already done (its
// right-hand side in lev[l].b); on return lev[l].x holds the result except for level 0's last red
// half-sweep, which the caller fuses with the dot product
void vcycle(hipStream_t st, const SlMg &mg, int l) {
  const SlMgLevel &L = mg.lev[l];
  const SlMgLevel &C = mg.lev[l + 1];
  const uint32_t nb = blocks(L.cells, MG_WG);
  hipLaunchKernelGGL(k_mg_sweep, dim3(nb), dim3(MG_WG), 0, st, L, 0, mg.ctl);
  hipLaunchKernelGGL(k_mg_restrict, dim3(blocks(C.cells, MG_WG)), dim3(MG_WG), 0, st, L, C, mg.ctl);
  if (l + 1 == mg.nlev - 1) hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(MG_CWG), 0, st, C, mg.coarse_sweeps, mg.ctl);
  else vcycle(st, mg, l + 1);
  hipLaunchKernelGGL(k_mg_black_prolong, dim3(nb), dim3(MG_WG), 0, st, L, C, mg.over, mg.ctl);
  if (l > 0) hipLaunchKernelGGL(k_mg_sweep, dim3(nb), dim3(MG_WG), 0, st, L, 1, mg.ctl);
}

// z = V-cycle(r) and the partials of sum r z
void precondition(hipStream_t st, const SlMg &mg) {
  const SlMgLevel &L0 = mg.lev[0];
  const uint32_t nb = blocks(L0.cells, MG_WG);
  if (mg.nlev == 1) {
    hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(MG_CWG), 0, st, L0, mg.coarse_sweeps, mg.ctl);
    hipLaunchKernelGGL(k_cg_dot, dim3(nb), dim3(MG_WG), 0, st, L0, mg.r, mg.part, mg.ctl);
    return;
  }
  vcycle(st, mg, 0);
  hipLaunchKernelGGL(k_mg_red_dot, dim3(nb), dim3(MG_WG), 0, st, L0, mg.r, mg.part, mg.ctl);
}

}  // namespace

uint32_t sl_mg_parts(uint32_t cells) { return blocks(cells, MG_WG); }

void sl_mg_solve(hipStream_t st, const SlTank *d_tank, const SlMg &mg, int seed) {
  const SlMgLevel &L0 = mg.lev[0];
  const uint32_t nb = blocks(L0.cells, MG_WG);
  hipLaunchKernelGGL(k_mg_setup0, dim3(nb), dim3(MG_WG), 0, st, d_tank, L0, mg.xs, mg.bs, mg.dlo, mg.ctl, seed);
  for (int l = 0; l + 1 < mg.nlev; l++) {
    const uint32_t nc = blocks(mg.lev[l + 1].cells, MG_WG);
    hipLaunchKernelGGL(k_mg_coarsen, dim3(nc), dim3(MG_WG), 0, st, mg.lev[l], mg.lev[l + 1]);
    hipLaunchKernelGGL(k_mg_flags, dim3(nc), dim3(MG_WG), 0, st, mg.lev[l + 1]);
  }
  hipLaunchKernelGGL(k_cg_resid, dim3(nb), dim3(MG_WG), 0, st, L0, mg.xs, mg.bs, mg.r, mg.part, mg.dlo);
  hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(MG_CWG), 0, st, mg.part, nb, mg.scal, mg.ctl, mg.rtol, 0u);
  precondition(st, mg);
  hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(MG_CWG), 0, st, mg.part, nb, mg.scal, mg.ctl, 1);
  for (uint32_t it = 1; it <= mg.max_iters; it++) {
    double *pold = mg.p[(it + 1) & 1u], *pnew = mg.p[it & 1u];
    hipLaunchKernelGGL(k_cg_spmv, dim3(nb), dim3(MG_WG), 0, st, L0, pold, pnew, mg.q, mg.part, mg.scal, mg.ctl, mg.dlo);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(MG_CWG), 0, st, mg.part, nb, mg.scal, mg.ctl);
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(MG_WG), 0, st, L0, mg.xs, mg.r, pnew, mg.q, mg.part, mg.scal,
                       mg.ctl);
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(MG_CWG), 0, st, mg.part, nb, mg.scal, mg.ctl, mg.rtol, it);
    if (it == mg.max_iters) break;
    precondition(st, mg);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(MG_CWG), 0, st, mg.part, nb, mg.scal, mg.ctl, 0);
  }
  hipLaunchKernelGGL(k_mg_finish, dim3(nb), dim3(MG_WG), 0, st, d_tank, L0, mg.xs);
}

void sl_k_project_mg(hipStream_t st, const SlTank *d_tank, uint32_t nodes, const double *xs) {
  hipLaunchKernelGGL(k_project_mg, dim3(blocks(nodes, MG_WG)), dim3(MG_WG), 0, st, d_tank, xs);
}
