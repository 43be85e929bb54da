// HBM calibration probe: plain copies of a known size, timed with HIP events.
// Its rocprofv3 FETCH_SIZE / WRITE_SIZE readings give the counter-to-bytes factor
// used to turn the sweeps' counters into HBM bytes, and the best variant's GB/s is
// the practical streaming ceiling of the device.
//   usage: hbm_copy [MiB per buffer (default 1024)] [repetitions (default 20)] [variant (-1 = all)]
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

// 0: grid-stride float4 loop over a fixed grid
__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// 1: one float4 per thread, one block per 256 float4s
__global__ void k_copy1(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// 2: four float4 per thread (all loads issued before the stores)
__global__ void k_copy4(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
  size_t base = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    size_t i = base + (size_t)k * blockDim.x;
    if (i < n) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    size_t i = base + (size_t)k * blockDim.x;
    if (i < n) dst[i] = v[k];
  }
}

// 3: as 2 with non-temporal loads and stores
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ void k_copy4nt(const f4v *__restrict__ src, f4v *__restrict__ dst, size_t n)
{
  size_t base = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
  f4v v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    size_t i = base + (size_t)k * blockDim.x;
    if (i < n) v[k] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    size_t i = base + (size_t)k * blockDim.x;
    if (i < n) __builtin_nontemporal_store(v[k], dst + i);
  }
}

// 4: read-only stream (sum), to see the read ceiling alone
__global__ void k_read4(const float4 *__restrict__ src, float *__restrict__ out, size_t n)
{
  size_t base = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    size_t i = base + (size_t)k * blockDim.x;
    if (i < n) {
      float4 v = src[i];
      s += v.x + v.y + v.z + v.w;
    }
  }
  if (s == 1234.5f) out[0] = s;  // keeps the loads live; never true for zeroed input
}

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));          \
      return 1;                                                               \
    }                                                                         \
  } while (0)

int main(int argc, char **argv)
{
  size_t mib = argc > 1 ? strtoul(argv[1], 0, 10) : 1024;
  int reps = argc > 2 ? atoi(argv[2]) : 20;
  int only = argc > 3 ? atoi(argv[3]) : -1;
  size_t bytes = mib << 20, n = bytes / sizeof(float4);
  float4 *a, *b;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const char *names[5] = {"gridstride_f4", "one_f4_per_thread", "four_f4_per_thread", "four_f4_nontemporal",
                          "read_only_four_f4"};
  for (int v = 0; v < 5; v++) {
    if (only >= 0 && v != only) continue;
    for (int r = -1; r < reps; r++) {
      if (r == 0) CK(hipEventRecord(e0, 0));
      switch (v) {
      case 0: hipLaunchKernelGGL(k_copy, dim3(256 * 16), dim3(256), 0, 0, a, b, n); break;
      case 1: hipLaunchKernelGGL(k_copy1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n); break;
      case 2: hipLaunchKernelGGL(k_copy4, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, b, n); break;
      case 3: hipLaunchKernelGGL(k_copy4nt, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, (const f4v *)a, (f4v *)b, n); break;
      case 4: hipLaunchKernelGGL(k_read4, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, a, (float *)b, n); break;
      }
    }
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    double per = ms / reps;
    double moved = (v == 4 ? 1.0 : 2.0) * bytes;
    printf("{\"variant\": \"%s\", \"bytes_read_per_launch\": %zu, \"bytes_written_per_launch\": %zu, "
           "\"ms_per_launch\": %.5f, \"GBps\": %.1f}\n",
           names[v], bytes, v == 4 ? (size_t)0 : bytes, per, moved / (per * 1e-3) / 1e9);
  }
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
