// Spatial re-ordering of the particle arrays: a stable sort of the live particles by the index of
// the cell that holds them, x fastest (particles outside the grid last, in their order). The
// particles of one cell end up next to each other and a row of cells is one contiguous range: the
// binning of P2G then takes a cell's ranks with one atomic per run of lanes, a P2G tile reads its
// cell rows as contiguous ranges, and the lanes of a G2P wavefront gather from nodes along x.
//
// The reference appends survivors with an atomic counter (31-wgsl-sim2.js:317-328), so its particle
// order is arbitrary; any permutation is a valid state.
#include <string.h>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>
#include "sl_internal.h"


// key: the cell index (the sentinel outside the grid) plus `key_hi`; value: the index plus `val_off`
// (a batch sorts several tanks in one pass, each tank's rank in the high key bits)
template <class A>
__global__ __launch_bounds__(SL_WG) void k_sort_keys(A W, uint32_t *keys, uint32_t *vals, uint32_t n_ub,
                                                      uint32_t key_hi, uint32_t val_off) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t id = lb * SL_WG + threadIdx.x;
  if (id >= n_ub) return;
  uint32_t key = sl_key_sentinel(T);
  if (id < T.C->count) {
    const int cur = T.cur;
    f3 g = grid_pos(T, F3(T.px[cur][id], T.py[cur][id], T.pz[cur][id]));
    if (in_grid(T, g)) {
      i3 c = toi3(g);
      key = cell_idx(T, c.x, c.y, c.z);
    }
  }
  keys[id] = key_hi | key;
  vals[id] = val_off + id;
}

template <class A>
__global__ __launch_bounds__(SL_WG) void k_sort_gather(A W, const uint32_t *vals, uint32_t val_off) {
  SL_TANK_OF(W, blockIdx.x, t, lb);
  const SlTank &T = sl_tank(W, t);
  uint32_t d = lb * SL_WG + threadIdx.x;
  if (d >= T.C->count) return;
  const int cur = T.cur, oth = 1 - cur;
  uint32_t s = vals[d] - val_off;
  T.px[oth][d] = T.px[cur][s]; T.py[oth][d] = T.py[cur][s]; T.pz[oth][d] = T.pz[cur][s];
  T.vx[oth][d] = T.vx[cur][s]; T.vy[oth][d] = T.vy[cur][s]; T.vz[oth][d] = T.vz[cur][s];
  T.pf[oth][d] = T.pf[cur][s];
}

size_t sl_sort_temp_bytes(uint32_t n, int end_bit) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                  (uint32_t *)nullptr, n, 0, end_bit, (hipStream_t)0);
  return bytes;
}

// keys/vals: 2 x n_ub each; after the call the particle arrays are re-ordered into the other
// buffer (the caller flips `cur`)
int sl_k_reorder(const SlLaunch &L, uint32_t n_ub, uint32_t *keys, uint32_t *vals, void *temp, size_t temp_bytes,
                 int end_bit) {
  if (!n_ub) return 0;
  SlLaunch K = L;
  K.w.total = (n_ub + SL_WG - 1) / SL_WG;
  SL_LAUNCH(k_sort_keys, dim3(K.w.total), dim3(SL_WG), 0, K, keys, vals, n_ub, 0u, 0u);
  size_t tb = temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(temp, tb, keys, keys + n_ub, vals, vals + n_ub, n_ub, 0, end_bit,
                                           L.stream);
  if (e != hipSuccess) return (int)e;
  SL_LAUNCH(k_sort_gather, dim3(K.w.total), dim3(SL_WG), 0, K, vals + n_ub, 0u);
  return 0;
}

// the three parts of a re-ordering for a batch that sorts the particles of several tanks together:
// keys and values of one tank at its offset, one stable sort of the n keys (values and keys in the
// first halves of their buffers, results in the second halves), the gather of one tank
void sl_k_sort_keys(const SlLaunch &L, uint32_t n_ub, uint32_t *keys, uint32_t *vals, uint32_t key_hi,
                    uint32_t val_off) {
  if (!n_ub) return;
  SlLaunch K = L;
  K.w.total = (n_ub + SL_WG - 1) / SL_WG;
  SL_LAUNCH(k_sort_keys, dim3(K.w.total), dim3(SL_WG), 0, K, keys + val_off, vals + val_off,
                     n_ub, key_hi, val_off);
}
int sl_sort_pairs(void *temp, size_t temp_bytes, uint32_t *keys, uint32_t *vals, uint32_t n, int end_bit,
                  hipStream_t stream) {
  size_t tb = temp_bytes;
  return (int)rocprim::radix_sort_pairs(temp, tb, keys, keys + n, vals, vals + n, n, 0, end_bit, stream);
}
void sl_k_sort_gather(const SlLaunch &L, uint32_t n_ub, const uint32_t *sorted_vals, uint32_t val_off) {
  if (!n_ub) return;
  SlLaunch K = L;
  K.w.total = (n_ub + SL_WG - 1) / SL_WG;
  SL_LAUNCH(k_sort_gather, dim3(K.w.total), dim3(SL_WG), 0, K, sorted_vals + val_off, val_off);
}
