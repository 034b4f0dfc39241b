// A plain device copy of known size, built as its own small library
// (libtsph_copy_probe.so).  It is the calibration for the HBM traffic
// counters: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE report, per dispatch of
// k_copy_probe, what the counters read for exactly `mib` MiB read and `mib` MiB
// written, and the ratio scales the counter values of the pipeline's kernels.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ src,
                                                    float4 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

// Copies `mib` MiB device-to-device `reps` times with k_copy_probe; returns 0 on
// success and the average kernel time in *avg_ms (HIP events).
int tsph_copy_probe(int mib, int reps, double *avg_ms) {
  if (mib <= 0 || mib > 16384 || reps <= 0 || !avg_ms) return -1;
  const size_t bytes = (size_t)mib << 20, n = bytes / sizeof(float4);
  float4 *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -2;
  if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) goto done;
  if (hipMemset(src, 0, bytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      goto done;
    const dim3 grid(cus * 8), block(256);
    hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, grid, block, 0, nullptr, src, dst, n);
    hipEventRecord(e1, nullptr);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    *avg_ms = ms / reps;
    rc = 0;
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

}  // extern "C"
