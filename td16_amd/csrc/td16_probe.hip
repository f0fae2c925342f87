// int32 VALU throughput probe: the integer-op roofline of the search kernel
// (which issues no MFMA) is priced against this sustained rate and against
// the published 256 CU x 64 lanes x 2.4 GHz figure.
#include <hip/hip_runtime.h>

#include <chrono>

#include "../../include/td16_abi.h"

namespace td16 {
int set_msg(char *msg, int cap, int status, const char *fmt, ...);

namespace {

constexpr int kChains = 8;
constexpr int kOpsPerRound = 2 * kChains;  // one add and one xor per chain

__global__ void __launch_bounds__(256) valu_probe_kernel(uint32_t *out, int rounds) {
  uint32_t a[kChains], b[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    a[c] = threadIdx.x + c;
    b[c] = blockIdx.x * 7u + c;
  }
  for (int r = 0; r < rounds; ++r) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b[c]));
      asm volatile("v_xor_b32 %0, %0, %1" : "+v"(b[c]) : "v"(a[c]));
    }
  }
  uint32_t acc = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc ^= a[c] + b[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace
}  // namespace td16

extern "C" int td16_probe_valu_peak(int device, int iterations, double *tops, char *msg,
                                    int cap) {
  using namespace td16;
  if (tops == nullptr || iterations <= 0)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid probe arguments");
  if (hipSetDevice(device) != hipSuccess)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid device %d", device);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return set_msg(msg, cap, TD16_DEVICE_ERROR, "device query failed");
  const int blocks = cus * 8;  // 8 blocks of 4 waves per CU: 8 waves per SIMD
  uint32_t *out = nullptr;
  if (hipMalloc(&out, sizeof(uint32_t) * 256 * (size_t)blocks) != hipSuccess)
    return set_msg(msg, cap, TD16_OUT_OF_MEMORY, "probe allocation failed");
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, out, 64);
  (void)hipEventRecord(e0, nullptr);
  hipLaunchKernelGGL(valu_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, out, iterations);
  (void)hipEventRecord(e1, nullptr);
  const hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (e != hipSuccess) return set_msg(msg, cap, TD16_DEVICE_ERROR, "probe failed");
  const double ops = (double)blocks * 256.0 * (double)iterations * kOpsPerRound;
  *tops = ops / (ms * 1e-3) / 1e12;
  return TD16_OK;
}

namespace td16 {
namespace {
// Grid-stride 16-byte copy: a kernel of known HBM traffic (bytes read and
// written) against which hardware byte counters are calibrated.
__global__ void __launch_bounds__(256) copy_probe_kernel(const uint4 *src, uint4 *dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace
}  // namespace td16

extern "C" int td16_probe_copy(int device, int64_t bytes, double *gbps, char *msg, int cap) {
  using namespace td16;
  if (gbps == nullptr || bytes < 16)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid probe arguments");
  if (hipSetDevice(device) != hipSuccess)
    return set_msg(msg, cap, TD16_INVALID_ARGUMENT, "invalid device %d", device);
  const uint64_t n = (uint64_t)bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return set_msg(msg, cap, TD16_OUT_OF_MEMORY, "probe allocation failed");
  }
  (void)hipMemset(a, 1, n * 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const unsigned blocks = 256 * 16;
  hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
  (void)hipEventRecord(e0, nullptr);
  hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
  (void)hipEventRecord(e1, nullptr);
  const hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return set_msg(msg, cap, TD16_DEVICE_ERROR, "probe failed");
  *gbps = 2.0 * (double)(n * 16) / (ms * 1e-3) / 1e9;
  return TD16_OK;
}
