// nb_hbmcopy.hip -- HBM traffic calibration: copies a buffer of known size with a plain
// grid-stride float4 kernel, so that rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch can be
// compared with the bytes actually moved (the scale factor for the force kernel's traffic).
// Usage: nb_hbmcopy [MiB] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? strtoul(argv[1], nullptr, 10) : 1024;
  const int reps = argc > 2 ? atoi(argv[2]) : 3;
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int r = 0; r < reps; r++) {
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(k_copy, dim3(256 * 16), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / ms / 1e6);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
