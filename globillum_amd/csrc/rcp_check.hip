// Exhaustive check of the traversal's reciprocal: gi::dev::rcp_rn(x) must
// equal the IEEE division 1.0f / x bit for bit for every one of the 2^32
// float bit patterns (NaN inputs: any NaN).  Prints the number of
// mismatches per input class and exits 0 only if there are none.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "gi_device.h"

namespace {

// 0: normal, 1: subnormal, 2: zero, 3: infinity, 4: NaN
__device__ __forceinline__ int input_class(uint32_t b) {
  const uint32_t e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
  if (e == 0xffu) return m ? 4 : 3;
  if (e == 0u) return m ? 1 : 2;
  return 0;
}

__global__ void k_rcp_check(unsigned long long *mismatch, uint32_t *first) {
  unsigned long long local[5] = {0, 0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    const uint32_t b = (uint32_t)i;
    const float x = __uint_as_float(b);
    const float want = 1.0f / x, got = gi::dev::rcp_rn(x);
    const int c = input_class(b);
    const bool ok = c == 4 ? got != got : __float_as_uint(got) == __float_as_uint(want);
    if (!ok) {
      ++local[c];
      atomicMin(first, b);
    }
  }
  for (int k = 0; k < 5; ++k)
    if (local[k]) atomicAdd(mismatch + k, local[k]);
}

}  // namespace

int main() {
  unsigned long long *d_mis = nullptr;
  uint32_t *d_first = nullptr;
  if (hipMalloc(&d_mis, 5 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&d_first, sizeof(uint32_t)) != hipSuccess) {
    std::fprintf(stderr, "hipMalloc failed\n");
    return 2;
  }
  const uint32_t none = 0xffffffffu;
  if (hipMemset(d_mis, 0, 5 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemcpy(d_first, &none, sizeof(none), hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "initialisation failed\n");
    return 2;
  }
  hipLaunchKernelGGL(k_rcp_check, dim3(2048), dim3(256), 0, 0, d_mis, d_first);
  unsigned long long mis[5];
  uint32_t first = none;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(mis, d_mis, sizeof(mis), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost) != hipSuccess) {
    std::fprintf(stderr, "kernel failed\n");
    return 2;
  }
  const char *names[5] = {"normal", "subnormal", "zero", "infinity", "nan"};
  unsigned long long total = 0;
  for (int k = 0; k < 5; ++k) {
    std::printf("%-9s mismatches %llu\n", names[k], mis[k]);
    total += mis[k];
  }
  std::printf("total mismatches %llu", total);
  if (total) std::printf(" (first input 0x%08x)", first);
  std::printf("\n");
  (void)hipFree(d_mis);
  (void)hipFree(d_first);
  return total ? 1 : 0;
}
