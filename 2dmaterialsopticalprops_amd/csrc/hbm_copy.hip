// Reference copy kernel for HBM traffic counters (measurement only, not part
// of librpa_hip.so): copies n bytes between two device buffers with 16-byte
// loads and stores, so FETCH_SIZE / WRITE_SIZE of a known-size transfer can
// scale the same counters read on the RPA kernels.
#include <hip/hip_runtime.h>

#include <cstddef>

__global__ __launch_bounds__(256) void k_hbm_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                  size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" {

// Copies `mib` MiB three times on device `device`; returns 0 on success and
// the HIP error code otherwise.  Allocates and frees its two buffers.
int hbm_copy_calibration(int device, int mib) {
  if (mib <= 0) return (int)hipErrorInvalidValue;
  hipError_t e = hipSetDevice(device);
  const size_t bytes = (size_t)mib << 20;
  float4 *a = nullptr, *b = nullptr;
  if (e == hipSuccess) e = hipMalloc((void **)&a, bytes);
  if (e == hipSuccess) e = hipMalloc((void **)&b, bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, bytes);
  const size_t n = bytes / sizeof(float4);
  for (int r = 0; r < 3 && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(k_hbm_copy, dim3(8192), dim3(256), 0, nullptr, a, b, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return (int)e;
}
}
