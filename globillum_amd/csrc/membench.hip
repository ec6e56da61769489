// Streaming-read microbenchmark of the memory hierarchy (measurement only:
// no part of the rendering path uses it).  It reads a device buffer of a
// given size several times and reports the read bandwidth of the level that
// holds it.  On gfx950 workgroups are dealt to the 8 XCDs round robin and
// every XCD has its own L2, so a block's share of the buffer (a grid-stride
// slice) stays on its XCD from pass to pass: a buffer of 8 x (under 4 MB)
// is read from L2, one of 8 x (4 MB .. 32 MB) from the 256 MB Infinity
// Cache, and a larger one from HBM.  One launch sweeps the buffer until it
// has read at least 1 GiB (so launch gaps do not count); sweep r hands
// block b the slice of block (b + 8r) mod G, on the same XCD but another
// CU, so the per-CU L1 does not serve repeated reads.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_stream_read(const float4 *__restrict__ buf, uint64_t n4, uint32_t first,
                                                        uint32_t sweeps, float *__restrict__ out) {
  const uint32_t G = gridDim.x;
  const uint64_t step = (uint64_t)G * kBlock;
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0, a2 = a0, a3 = a0;
  for (uint32_t r = 0; r < sweeps; ++r) {
    const uint64_t vb = (blockIdx.x + 8ull * (first + r)) % G;
    uint64_t i = vb * kBlock + threadIdx.x;
    for (; i + 3 * step < n4; i += 4 * step) {  // four independent loads in flight per lane
      const float4 v0 = buf[i], v1 = buf[i + step], v2 = buf[i + 2 * step], v3 = buf[i + 3 * step];
      a0.x += v0.x; a0.y += v0.y; a0.z += v0.z; a0.w += v0.w;
      a1.x += v1.x; a1.y += v1.y; a1.z += v1.z; a1.w += v1.w;
      a2.x += v2.x; a2.y += v2.y; a2.z += v2.z; a2.w += v2.w;
      a3.x += v3.x; a3.y += v3.y; a3.z += v3.z; a3.w += v3.w;
    }
    for (; i < n4; i += step) {
      const float4 v = buf[i];
      a0.x += v.x; a0.y += v.y; a0.z += v.z; a0.w += v.w;
    }
  }
  const float s = (a0.x + a1.x + a2.x + a3.x) + (a0.y + a1.y + a2.y + a3.y) + (a0.z + a1.z + a2.z + a3.z) +
                  (a0.w + a1.w + a2.w + a3.w);
  // the buffer holds zeros: the sum is 0 and nothing is stored, but the
  // loads cannot be dropped
  if (s == 1.0f) out[blockIdx.x * kBlock + threadIdx.x] = s;
}

// One pass of a plain copy, dst[i] = src[i]: exactly `n4` * 16 bytes read
// and written, the known-size kernel that scales HBM traffic counters
// (FETCH_SIZE / WRITE_SIZE) of other kernels in the same profiled run.
__global__ __launch_bounds__(kBlock) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                 uint64_t n4) {
  const uint64_t step = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += step) dst[i] = src[i];
}

}  // namespace

extern "C" {

// One launch of k_copy over `bytes` (a multiple of 16) between two new
// buffers on `device`, after both were filled by hipMemsetAsync; returns 0
// or the HIP error code.  Allocates and frees its buffers, synchronises.
int gi_membench_copy(int32_t device, int64_t bytes) {
  if (bytes < 16 || bytes % 16 != 0) return (int)hipErrorInvalidValue;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return (int)e;
  const uint32_t G = (uint32_t)prop.multiProcessorCount * 4u;
  float4 *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  if ((e = hipMalloc(&src, (size_t)bytes)) != hipSuccess) goto done;
  if ((e = hipMalloc(&dst, (size_t)bytes)) != hipSuccess) goto done;
  if ((e = hipStreamCreate(&s)) != hipSuccess) goto done;
  if ((e = hipMemsetAsync(src, 0, (size_t)bytes, s)) != hipSuccess) goto done;
  if ((e = hipMemsetAsync(dst, 0, (size_t)bytes, s)) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_copy, dim3(G), dim3(kBlock), 0, s, src, dst, (uint64_t)bytes / 16u);
  if ((e = hipGetLastError()) != hipSuccess) goto done;
  e = hipStreamSynchronize(s);
done:
  if (s) (void)hipStreamDestroy(s);
  if (dst) (void)hipFree(dst);
  if (src) (void)hipFree(src);
  return (int)e;
}

// Read bandwidth (GB/s, 1e9 bytes) of `passes` launches, each reading a
// zeroed buffer of `bytes` (a multiple of 16, at least 16 * 1024) on
// `device` max(1, 2^30 / bytes) times, timed with HIP events after one
// warm launch.  Returns 0 or the HIP error code.  Makes `device` the
// current device, allocates and frees its own buffers, and synchronises.
int gi_membench_read(int32_t device, int64_t bytes, int32_t passes, double *gbs) {
  if (bytes < 16 * 1024 || bytes % 16 != 0 || passes < 1 || !gbs) return (int)hipErrorInvalidValue;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return (int)e;
  const uint32_t G = (uint32_t)prop.multiProcessorCount * 4u;  // 4 blocks per CU; a multiple of 8
  float4 *buf = nullptr;
  float *out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  float ms = 0.0f;
  const uint64_t n4 = (uint64_t)bytes / 16u;
  const uint32_t sweeps = bytes >= ((int64_t)1 << 30) ? 1u : (uint32_t)(((int64_t)1 << 30) / bytes);
  if ((e = hipMalloc(&buf, (size_t)bytes)) != hipSuccess) goto done;
  if ((e = hipMalloc(&out, (size_t)G * kBlock * sizeof(float))) != hipSuccess) goto done;
  if ((e = hipStreamCreate(&s)) != hipSuccess) goto done;
  if ((e = hipEventCreate(&t0)) != hipSuccess) goto done;
  if ((e = hipEventCreate(&t1)) != hipSuccess) goto done;
  if ((e = hipMemsetAsync(buf, 0, (size_t)bytes, s)) != hipSuccess) goto done;
  hipLaunchKernelGGL(k_stream_read, dim3(G), dim3(kBlock), 0, s, buf, n4, 0u, sweeps, out);
  if ((e = hipGetLastError()) != hipSuccess) goto done;
  if ((e = hipEventRecord(t0, s)) != hipSuccess) goto done;
  for (int32_t p = 0; p < passes; ++p)
    hipLaunchKernelGGL(k_stream_read, dim3(G), dim3(kBlock), 0, s, buf, n4, (uint32_t)(p + 1) * sweeps % G, sweeps,
                       out);
  if ((e = hipGetLastError()) != hipSuccess) goto done;
  if ((e = hipEventRecord(t1, s)) != hipSuccess) goto done;
  if ((e = hipEventSynchronize(t1)) != hipSuccess) goto done;
  if ((e = hipEventElapsedTime(&ms, t0, t1)) != hipSuccess) goto done;
  *gbs = ms > 0.0f ? (double)bytes * sweeps * passes / ((double)ms * 1e6) : 0.0;
done:
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (s) (void)hipStreamDestroy(s);
  if (out) (void)hipFree(out);
  if (buf) (void)hipFree(buf);
  return (int)e;
}

}  // extern "C"
