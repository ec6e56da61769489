// Device evaluation of the sampler for a list of (index, dimension) pairs,
// so that tests can compare the kernels' chunked evaluation (gi_device.h,
// halton) with the plain digit loop, dimension by dimension.
//   gi_halton_check SEED IN OUT
// IN holds uint32 pairs (n, dim) with dim < 4 + 7 * 32 (the dimensions a
// render of depth <= 32 uses); OUT receives one float32 per pair.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gi_device.h"
#include "gi_internal.h"

namespace {

constexpr int32_t kDims = 4 + 7 * 32;

__global__ void k_halton_eval(const uint32_t *__restrict__ pairs, uint32_t count, gi::dev::HaltonDev H,
                              float *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = gi::dev::halton(pairs[2 * i], (int)pairs[2 * i + 1], H);
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: gi_halton_check SEED IN OUT\n");
    return 2;
  }
  const uint32_t seed = (uint32_t)std::strtoul(argv[1], nullptr, 10);
  std::FILE *fi = std::fopen(argv[2], "rb");
  if (!fi) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 2;
  }
  std::vector<uint32_t> pairs;
  uint32_t buf[2];
  while (std::fread(buf, sizeof(uint32_t), 2, fi) == 2) {
    if (buf[1] >= (uint32_t)kDims) {
      std::fprintf(stderr, "dimension %u out of range\n", buf[1]);
      std::fclose(fi);
      return 2;
    }
    pairs.push_back(buf[0]);
    pairs.push_back(buf[1]);
  }
  std::fclose(fi);
  const uint32_t count = (uint32_t)(pairs.size() / 2);
  std::vector<gi::HaltonDim> dims;
  std::vector<uint16_t> chunks;
  gi::halton_tables(seed, kDims, dims, chunks);
  void *d_pairs = nullptr, *d_dims = nullptr, *d_chunks = nullptr, *d_out = nullptr;
  const size_t nb = std::max<size_t>(1, count);
  if (!ok(hipMalloc(&d_pairs, nb * 2 * sizeof(uint32_t)), "hipMalloc") ||
      !ok(hipMalloc(&d_out, nb * sizeof(float)), "hipMalloc") ||
      !ok(hipMalloc(&d_dims, dims.size() * sizeof(gi::HaltonDim)), "hipMalloc") ||
      !ok(hipMalloc(&d_chunks, chunks.size() * sizeof(uint16_t)), "hipMalloc"))
    return 2;
  if (!ok(hipMemcpy(d_pairs, pairs.data(), pairs.size() * sizeof(uint32_t), hipMemcpyHostToDevice), "copy") ||
      !ok(hipMemcpy(d_dims, dims.data(), dims.size() * sizeof(gi::HaltonDim), hipMemcpyHostToDevice), "copy") ||
      !ok(hipMemcpy(d_chunks, chunks.data(), chunks.size() * sizeof(uint16_t), hipMemcpyHostToDevice), "copy"))
    return 2;
  const gi::dev::HaltonDev H{static_cast<const gi::HaltonDim *>(d_dims), static_cast<const uint16_t *>(d_chunks), 0u};
  if (count > 0)
    hipLaunchKernelGGL(k_halton_eval, dim3((count + 255) / 256), dim3(256), 0, 0,
                       static_cast<const uint32_t *>(d_pairs), count, H, static_cast<float *>(d_out));
  std::vector<float> out(count);
  if (!ok(hipGetLastError(), "launch") || !ok(hipDeviceSynchronize(), "kernel") ||
      !ok(hipMemcpy(out.data(), d_out, count * sizeof(float), hipMemcpyDeviceToHost), "copy"))
    return 2;
  std::FILE *fo = std::fopen(argv[3], "wb");
  if (!fo || std::fwrite(out.data(), sizeof(float), count, fo) != count) {
    std::fprintf(stderr, "cannot write %s\n", argv[3]);
    if (fo) std::fclose(fo);
    return 2;
  }
  std::fclose(fo);
  (void)hipFree(d_pairs);
  (void)hipFree(d_out);
  (void)hipFree(d_dims);
  (void)hipFree(d_chunks);
  std::printf("%u values\n", count);
  return 0;
}
