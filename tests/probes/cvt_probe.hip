// Checks the library's f2i_hw (v_cvt_i32_f32) against f2i's definition -- truncation toward zero,
// saturation at the int32 range, NaN to 0 -- over 4 Mi random bit patterns and the edge cases, on
// the GPU. Exit code 0 when every value agrees.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "sl_internal.h"

static int32_t f2i_ref(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}

__global__ void k_probe(const float *in, int32_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) out[i] = f2i_hw(in[i]);
}

int main() {
  const uint32_t n = 1u << 22;  // a multiple of 256: the grid covers exactly n threads
  std::vector<float> h(n);
  std::mt19937 rng(7);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t b = rng();
    memcpy(&h[i], &b, 4);
  }
  const float edge[] = {0.0f, -0.0f, 0.5f, -0.5f, 0.99999994f, -0.99999994f, 1.0f, -1.0f, 1.5f, -1.5f,
                        2147483520.0f, 2147483648.0f, -2147483648.0f, -2147483904.0f, 1e30f, -1e30f,
                        __builtin_inff(), -__builtin_inff(), __builtin_nanf(""), -__builtin_nanf(""), 1e-40f,
                        -1e-40f, 16777217.0f, 8388608.5f, -8388608.5f};
  for (size_t k = 0; k < sizeof(edge) / sizeof(edge[0]); k++) h[k] = edge[k];
  float *din = nullptr;
  int32_t *dout = nullptr;
  if (hipMalloc(&din, n * 4) != hipSuccess || hipMalloc(&dout, n * 4) != hipSuccess) return 2;
  if (hipMemcpy(din, h.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess) return 2;
  hipLaunchKernelGGL(k_probe, dim3(n / 256), dim3(256), 0, 0, din, dout, n);
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  std::vector<int32_t> o(n);
  if (hipMemcpy(o.data(), dout, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return 2;
  uint64_t bad = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (o[i] == f2i_ref(h[i])) continue;
    if (bad < 5) printf("mismatch at %a: %d, want %d\n", h[i], o[i], f2i_ref(h[i]));
    bad++;
  }
  printf("f2i_hw: %u inputs, %llu mismatches\n", n, (unsigned long long)bad);
  (void)hipFree(din);
  (void)hipFree(dout);
  return bad ? 1 : 0;
}
