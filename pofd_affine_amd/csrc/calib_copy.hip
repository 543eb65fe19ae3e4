// Plain device-to-device copy of known size, used only to calibrate the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters (bytes the counters report for
// a kernel that reads and writes exactly 8n bytes each).  Not part of the
// likelihood library.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__global__ __launch_bounds__(256) void calib_copy_kernel(const double2* __restrict__ src,
                                                         double2* __restrict__ dst,
                                                         uint64_t n2) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n2;
       i += stride)
    dst[i] = src[i];
}

}  // namespace

// n: doubles (even); src and dst 16-byte aligned device buffers of n doubles.
extern "C" int pofd_calib_copy(const double* src, double* dst, uint64_t n, void* stream) {
  if (!src || !dst || (n & 1)) return 5;
  const uint64_t n2 = n / 2;
  hipLaunchKernelGGL(calib_copy_kernel, dim3(4096), dim3(256), 0,
                     static_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst), n2);
  return hipGetLastError() == hipSuccess ? 0 : 6;
}
