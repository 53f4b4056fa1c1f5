// Copy kernel of known traffic (reads `bytes`, writes `bytes`), used to scale
// hardware byte counters (FETCH_SIZE / WRITE_SIZE) into HBM bytes.
#include <hip/hip_runtime.h>

#include "ent_bfs.h"
#include "ent_common.h"

namespace ent {

__global__ __launch_bounds__(256) void k_calib_copy(const int4 *__restrict__ src, int4 *__restrict__ dst,
                                                    long long n) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nt = (long long)gridDim.x * 256;
  for (long long i = t; i < n; i += nt) dst[i] = src[i];
}

}  // namespace ent

extern "C" int ent_calib_copy(int device, int64_t bytes, int32_t reps, double *ms_per_copy) {
  using namespace ent;
  clear_error();
  if (bytes < 16 || bytes % 16 || reps < 1) return set_error(ENT_EINVAL, "bytes must be a positive multiple of 16");
  if (hipSetDevice(device) != hipSuccess) return set_error(ENT_EDEVICE, "hipSetDevice(%d) failed", device);
  int4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, size_t(bytes)) != hipSuccess || hipMalloc(&b, size_t(bytes)) != hipSuccess) {
    if (a) (void)hipFree(a);
    return set_error(ENT_ENOMEM, "calibration buffers");
  }
  int rc = ENT_OK;
  hipEvent_t e0, e1;
  if (hipMemset(a, 1, size_t(bytes)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = set_error(ENT_EDEVICE, "calibration setup failed");
  } else {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const long long n = bytes / 16;
    hipLaunchKernelGGL(k_calib_copy, dim3(cus * 16), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_calib_copy, dim3(cus * 16), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) rc = set_error(ENT_EDEVICE, "calibration copy failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
