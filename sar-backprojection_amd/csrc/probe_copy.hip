// Calibration probe for hardware counters: a plain streaming copy of a known number of bytes, so
// that FETCH_SIZE / WRITE_SIZE read on it give the counter's scale (bytes counted per byte moved).
// Built into libsarbp_probe.so; used only by measurement scripts, not by the focusing library.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {
__global__ __launch_bounds__(256) void k_probe_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" int sarbp_probe_copy(int64_t bytes, int32_t reps, float *ms_out) {
  if (bytes <= 0 || bytes % 16 || reps < 1) return -1;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)bytes) != hipSuccess) return -2;
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
    (void)hipFree(a);
    return -2;
  }
  (void)hipMemset(a, 1, (size_t)bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int64_t n = bytes / 16;
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_probe_copy, dim3(4096), dim3(256), 0, nullptr, (const float4 *)a, (float4 *)b, n);
  (void)hipEventRecord(e1, nullptr);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (ms_out) *ms_out = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return e == hipSuccess ? 0 : -4;
}
