// copy_probe.hip -- calibration for HBM traffic counters: a plain device copy of exactly 256 MiB
// (float4 loads/stores, grid-stride), five times.  rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of it
// gives the factor between counter readings and bytes moved.
#include <hip/hip_runtime.h>
#include <stdio.h>
__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}
int main() {
  const size_t bytes = 256ull << 20, n = bytes / 16;
  float4 *a, *b;
  if (hipMalloc(&a, bytes) || hipMalloc(&b, bytes)) return 1;
  if (hipMemset(a, 1, bytes) || hipMemset(b, 0, bytes)) return 1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  for (int r = 0; r < 5; ++r) {
    (void)hipEventRecord(e0);
    k_copy<<<4096, 256>>>(a, b, n);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms; (void)hipEventElapsedTime(&ms, e0, e1);
    printf("copy 256 MiB: %.3f ms  %.1f GB/s (read+write)\n", ms, 2.0 * bytes / ms / 1e6);
  }
  return hipDeviceSynchronize() ? 1 : 0;
}
