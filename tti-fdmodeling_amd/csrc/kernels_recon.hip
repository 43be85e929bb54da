// Kernels of the boundary-saving reconstruction of the forward wavefield: compact
// copies of a frame of grid points (the state band around the undamped interior and
// the history outside it, saved every forward step and put back during the backward
// pass), a negated copy of an array (coef2 and the source series of time-reversed
// steps) and the check that a box of the grid is undamped.
#include "fdmod_internal.h"

namespace fdm {

__global__ void k_frame_io(FrameIO a)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = a.off[a.nrect];
  if (i >= n) return;
  int r = 0;
  while (r + 1 < a.nrect && i >= a.off[r + 1]) r++;
  const unsigned j = (unsigned)(i - a.off[r]), h = (unsigned)a.h[r];
  const size_t o = (size_t)(a.x0[r] + (int)(j / h)) * a.ldz + (a.z0[r] + (int)(j % h));
  float *pk = a.pack + i;
  if (a.save) {
    for (int f = 0; f < a.nf; f++) pk[(size_t)f * n] = a.grid[f][o];
  } else {
    for (int f = 0; f < a.nf; f++) a.grid[f][o] = pk[(size_t)f * n];
  }
}

__global__ void k_negate(const float *__restrict__ in, float *__restrict__ out, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = -in[i];
}

__global__ void k_count_damped(const float *__restrict__ coef1, const float *__restrict__ coef2, float dt, int ldz,
                               int x0, int z0, int z1, unsigned *count)
{
  const int z = z0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= z1) return;
  const size_t o = (size_t)(x0 + blockIdx.y) * ldz + z;
  if (coef1[o] != 1.0f || coef2[o] != dt) atomicAdd(count, 1u);
}

void launch_frame_io(hipStream_t s, const FrameIO &a)
{
  const long long n = a.off[a.nrect];
  if (n <= 0 || a.nf <= 0) return;
  hipLaunchKernelGGL(k_frame_io, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}

void launch_negate(hipStream_t s, const float *in, float *out, size_t n)
{
  if (n == 0) return;
  hipLaunchKernelGGL(k_negate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
}

void launch_count_damped(hipStream_t s, const float *coef1, const float *coef2, float dt, int ldz, int x0, int x1,
                         int z0, int z1, unsigned *count)
{
  if (x1 <= x0 || z1 <= z0) return;
  hipLaunchKernelGGL(k_count_damped, dim3((unsigned)((z1 - z0 + 255) / 256), (unsigned)(x1 - x0)), dim3(256), 0, s,
                     coef1, coef2, dt, ldz, x0, z0, z1, count);
}

}  // namespace fdm
