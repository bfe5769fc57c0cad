// copy_probe.hip -- HBM bandwidth reference for the roofline numbers.
//
// Plain copy / read / write kernels over 1 GiB device buffers, timed with
// HIP events.  The copy kernel is also the known-size workload that
// calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters (1 GiB read and
// 1 GiB written per launch).  Usage: copy_probe
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
template <bool NT>
__global__ void copy_k(const u4* __restrict__ a, u4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    u4 v = a[i];
    if (NT) __builtin_nontemporal_store(v, b + i); else b[i] = v;
  }
}
__global__ void copy_unroll(const u4* __restrict__ a, u4* __restrict__ b, size_t n) {
  // each thread 4 vectors, block-contiguous
  size_t base = (blockIdx.x * (size_t)blockDim.x * 4) + threadIdx.x;
  u4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { size_t i = base + k * blockDim.x; v[k] = i < n ? a[i] : u4{0,0,0,0}; }
#pragma unroll
  for (int k = 0; k < 4; ++k) { size_t i = base + k * blockDim.x; if (i < n) __builtin_nontemporal_store(v[k], b + i); }
}
__global__ void read_k(const u4* __restrict__ a, u4* __restrict__ out, size_t n) {
  u4 acc = {0,0,0,0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) acc ^= a[i];
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x12345678u) out[0] = acc;
}
__global__ void write_k(u4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) __builtin_nontemporal_store(u4{1,2,3,4}, b + i);
}
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while(0)
int main() {
  const size_t bytes = 1ull << 30, n = bytes / 16;
  u4 *a, *b; CK(hipMalloc(&a, bytes)); CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 1, bytes)); CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  auto timeit = [&](const char* name, double moved, auto launch) {
    launch(); CK(hipDeviceSynchronize());
    const int reps = 20; CK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
    float ms; CK(hipEventElapsedTime(&ms, e0, e1));
    printf("%-28s %8.1f us  %6.2f TB/s\n", name, ms * 1e3 / reps, moved / (ms * 1e-3 / reps) / 1e12);
  };
  for (int grid : {1024, 2048, 4096, 8192, 16384}) {
    char nm[64]; snprintf(nm, 64, "copy grid=%d", grid);
    timeit(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(copy_k<false>, dim3(grid), dim3(256), 0, 0, a, b, n); });
    snprintf(nm, 64, "copy NT grid=%d", grid);
    timeit(nm, 2.0 * bytes, [&] { hipLaunchKernelGGL(copy_k<true>, dim3(grid), dim3(256), 0, 0, a, b, n); });
  }
  timeit("copy unroll4 NT", 2.0 * bytes, [&] { hipLaunchKernelGGL(copy_unroll, dim3((n + 1023) / 1024), dim3(256), 0, 0, a, b, n); });
  timeit("read grid=4096", 1.0 * bytes, [&] { hipLaunchKernelGGL(read_k, dim3(4096), dim3(256), 0, 0, a, b, n); });
  timeit("read grid=16384", 1.0 * bytes, [&] { hipLaunchKernelGGL(read_k, dim3(16384), dim3(256), 0, 0, a, b, n); });
  timeit("write NT grid=4096", 1.0 * bytes, [&] { hipLaunchKernelGGL(write_k, dim3(4096), dim3(256), 0, 0, b, n); });
  timeit("hipMemcpyDtoD", 2.0 * bytes, [&] { CK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0)); });
  return 0;
}
