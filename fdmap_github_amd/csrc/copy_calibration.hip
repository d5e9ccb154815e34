// copy_calibration.hip -- a plain FP64 device-to-device copy of known size.
//
// Run under `rocprofv3 --pmc FETCH_SIZE` (and separately WRITE_SIZE), it
// gives the factor between what the hardware byte counters report and the
// bytes a kernel really moves; the bulk kernel's HBM traffic per launch is
// its counter reading times that factor.  Usage: copy_calibration [MiB]
// (default 512 MiB per buffer); prints bytes and the achieved rate.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
  const size_t mib = argc > 1 ? (size_t)atol(argv[1]) : 512;
  if (mib == 0 || mib > 16384) {
    fprintf(stderr, "copy_calibration: size must be 1..16384 MiB\n");
    return 2;
  }
  const size_t n = mib * (1u << 20) / sizeof(double);
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double)) != hipSuccess || hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
    fprintf(stderr, "copy_calibration: hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, n * sizeof(double)) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  for (int r = 0; r < 5; r++) {
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy %zu bytes read + %zu bytes written: %.4f ms, %.1f GB/s\n", n * sizeof(double),
           n * sizeof(double), ms, 2.0 * n * sizeof(double) / (ms * 1e-3) / 1e9);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
