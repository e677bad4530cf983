// A ripple frame on caller-owned device buffers that allocates no device memory
// (sl_ripple_frame_noalloc): the kernels of the ripple stage (k_ripple.hip) run over a tank record,
// a twiddle table and three M x M complex work fields that live in a file-scope device array,
// `scratch`, of fixed size. For callers that cannot allocate (code under graph capture, tight serving
// loops). Every GPU the library runs on holds its own copy of the array until the process ends; its
// address on a device is taken once, with that device current, and kept per device.
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
#include <string.h>
#include <vector>

#include "../../include/suiri.h"
#include "sl_internal.h"

// layout of `scratch`: the tank record, the twiddle table (M float2), then the work fields a, b and
// hat (M x M float2 each) of the largest supported M
#define NA_REC 0
#define NA_TW 2048
#define NA_FIELDS (NA_TW + SL_RIPPLE_NOALLOC_MAX_M * 8)
#define NA_FIELD_BYTES ((size_t)SL_RIPPLE_NOALLOC_MAX_M * SL_RIPPLE_NOALLOC_MAX_M * 8)
#define NA_BYTES (NA_FIELDS + 3 * NA_FIELD_BYTES)
static_assert(sizeof(SlTank) <= NA_TW, "the tank record fits its slot of the scratch array");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

void sl_set_error(const char *msg);  // capi.hip: the message of sl_last_error(NULL)

namespace {

std::mutex g_mu;
constexpr int kMaxDevices = 64;
unsigned char *g_scratch[kMaxDevices];  // device address of `scratch` per device ordinal
std::vector<float> g_tw[16];            // host twiddle tables, per log2 M

// writes the tank record into the scratch array (a kernel argument, so no host buffer has to outlive
// the call)
__global__ void k_put_record(SlTank *dst, SlTank rec) {
  if (threadIdx.x == 0) *dst = rec;
}

SlWork work_of(SlTank *rec, uint64_t items, uint32_t per_wg) {
  SlWork w;
  w.tanks = rec;
  w.start = nullptr;
  w.ntanks = 1;
  w.total = (uint32_t)((items + per_wg - 1) / per_wg);
  w.list = nullptr;
  return w;
}

int fail(int code, const char *msg) {
  sl_set_error(msg);
  return code;
}

}  // namespace

extern "C" int sl_ripple_frame_noalloc(const sl_config *cfg, uint32_t R, double Lx, double Lz, double dt,
                                       double depth, void *state, int32_t *src, void *tex, void *stream) {
  if (!cfg || !state || !src || !tex) return fail(SL_EARG, "sl_ripple_frame_noalloc: null argument");
  if (R == 0 || (R & (R - 1)) || 2 * (uint64_t)R > SL_RIPPLE_NOALLOC_MAX_M)
    return fail(SL_ECAP, "sl_ripple_frame_noalloc: R must be a power of two with M = 2R <= 512");
  if (!(Lx > 0) || !(Lz > 0) || !(dt >= 0) || !(depth > 0))
    return fail(SL_EARG, "sl_ripple_frame_noalloc: Lx, Lz and depth must be positive, dt >= 0");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(SL_EHIP, "sl_ripple_frame_noalloc: no current device");
  const uint32_t M = 2 * R;
  uint32_t lm = 0;
  while ((1u << lm) < M) lm++;
  unsigned char *base = nullptr;
  const float *tw = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return fail(SL_EHIP, "sl_ripple_frame_noalloc: hipGetSymbolAddress(scratch) failed");
      g_scratch[dev] = (unsigned char *)p;
    }
    base = g_scratch[dev];
    std::vector<float> &t = g_tw[lm];
    if (t.empty()) {  // e^(2 pi i j / M) in double, rounded to f32 (64-ripple.js:175-178), as sl_create
      t.resize(2 * (size_t)M);
      for (uint32_t j = 0; j < M; j++) {
        t[2 * j] = (float)cos((2 * M_PI * j) / M);
        t[2 * j + 1] = (float)sin((2 * M_PI * j) / M);
      }
    }
    tw = t.data();
  }
  hipStream_t st = (hipStream_t)stream;
  SlTank rec;
  memset(&rec, 0, sizeof rec);
  rec.R = R; rec.M = M; rec.logM = lm;
  rec.g = cfg->gravity;
  rec.rLx = (float)Lx; rec.rLz = (float)Lz; rec.rdt = (float)dt; rec.rmd = (float)depth;
  rec.rsr = cfg->r_sigma_rho; rec.rnu = cfg->r_nu; rec.rfilm = cfg->r_film; rec.rhpk = cfg->r_hp_k;
  rec.rrho = cfg->r_rho_kick;
  rec.mask_bound = 0;
  rec.rstate = (float2 *)state;
  rec.rsrc = src;
  rec.rtex = (float4 *)tex;
  rec.tw = (float2 *)(base + NA_TW);
  rec.ra = (float2 *)(base + NA_FIELDS);
  rec.rb = (float2 *)(base + NA_FIELDS + NA_FIELD_BYTES);
  rec.rhat = (float2 *)(base + NA_FIELDS + 2 * NA_FIELD_BYTES);
  SlTank *drec = (SlTank *)(base + NA_REC);
  hipLaunchKernelGGL(k_put_record, dim3(1), dim3(64), 0, st, drec, rec);
  if (hipMemcpyAsync(base + NA_TW, tw, (size_t)M * 8, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(SL_EHIP, "sl_ripple_frame_noalloc: twiddle copy failed");
  // the work maps of encode_ripple (capi.hip): row groups, column tiles of 4096 / M, elements, texels
  uint32_t CW = 4096 / M;
  if (CW > M) CW = M;
  if (CW < 1) CW = 1;
  const uint32_t R1 = R + 1;
  sl_k_ripple_frame_ex(work_of(drec, M / sl_fft_rows_per_wg(M), 1), work_of(drec, M, CW), work_of(drec, (uint64_t)M * M, SL_WG),
                       work_of(drec, (uint64_t)R1 * R1, SL_WG), st, 0, M);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SL_EHIP, hipGetErrorString(e));
  return SL_OK;
}
