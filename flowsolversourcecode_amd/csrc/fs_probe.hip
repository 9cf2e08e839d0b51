// Calibration kernel for the HBM byte counters: a plain streaming copy of a
// known size.  FETCH_SIZE / WRITE_SIZE read on it convert the same counters
// read on any other kernel into bytes.
#include <hip/hip_runtime.h>

#include <string>

namespace fs {
void set_error(const std::string &msg);

namespace {
// One 16-byte element per thread in address order: the streaming form that
// reaches the highest bandwidth (a grid-stride loop over the same bytes runs
// about a quarter slower).
__global__ void __launch_bounds__(256) k_copy16(long long n16, const double2 *__restrict__ src,
                                                double2 *__restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}
} // namespace
} // namespace fs

extern "C" int fs_time_device_copy(long long bytes, int reps, double *avg_ms) {
  if (bytes <= 0 || bytes % 16 || reps <= 0 || !avg_ms) {
    fs::set_error("fs_time_device_copy: need bytes > 0 and a multiple of 16, reps > 0, avg_ms");
    return -1;
  }
  void *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  auto fail = [&](hipError_t e, const char *what) {
    fs::set_error(std::string("fs_time_device_copy: ") + what + ": " + hipGetErrorString(e));
    rc = -2;
  };
  hipError_t e;
  if ((e = hipMalloc(&src, bytes)) != hipSuccess) fail(e, "hipMalloc");
  else if ((e = hipMalloc(&dst, bytes)) != hipSuccess) fail(e, "hipMalloc");
  else if ((e = hipStreamCreate(&s)) != hipSuccess) fail(e, "hipStreamCreate");
  else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
    fail(e, "hipEventCreate");
  const long long n16 = bytes / 16;
  const long long nblk = (n16 + 255) / 256;
  if (rc == 0 && nblk > 0x7fffffffLL) fail(hipErrorInvalidValue, "size");
  if (rc == 0) {
    const int grid = (int)nblk;
    (void)hipMemsetAsync(src, 0, bytes, s);
    hipLaunchKernelGGL(fs::k_copy16, dim3(grid), dim3(256), 0, s, n16, (const double2 *)src, (double2 *)dst);
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(fs::k_copy16, dim3(grid), dim3(256), 0, s, n16, (const double2 *)src, (double2 *)dst);
    (void)hipEventRecord(e1, s);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) fail(e, "kernel");
    float ms = 0.f;
    if (rc == 0 && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) fail(e, "hipEventElapsedTime");
    *avg_ms = (double)ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}
