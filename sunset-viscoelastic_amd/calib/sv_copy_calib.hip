// sv_copy_calib.hip -- plain device-to-device copy of a known size, used to
// calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch)
// against bytes actually moved.  Copies `mib` MiB (default 256) of doubles
// `reps` times (default 3) with a grid-stride double2 loop and prints the
// algorithmic bytes per dispatch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void k_copy_calib(const double2 *__restrict__ a,
                                                    double2 *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_fill_calib(double2 *__restrict__ a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = double2{(double)i, 1.0};
}

#define CK(x)                                                                     \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));              \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
  const int reps = argc > 2 ? atoi(argv[2]) : 3;
  if (mib == 0 || mib > 16384 || reps < 1 || reps > 100) {
    fprintf(stderr, "usage: sv_copy_calib [MiB 1..16384] [reps 1..100]\n");
    return 2;
  }
  const size_t bytes = mib << 20;
  const size_t n = bytes / sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  const int grid = 256 * 8;
  k_fill_calib<<<grid, 256>>>(a, n);
  CK(hipGetLastError());
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float tot = 0.0f;
  for (int r = 0; r < reps; r++) {
    CK(hipEventRecord(e0));
    k_copy_calib<<<grid, 256>>>(a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    tot += ms;
  }
  printf("k_copy_calib: read %zu B + write %zu B per dispatch (%zu KiB each), %d reps, "
         "mean %.4f ms, %.1f GB/s\n",
         bytes, bytes, bytes >> 10, reps, tot / reps, 2.0 * bytes / (tot / reps * 1e-3) / 1e9);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
