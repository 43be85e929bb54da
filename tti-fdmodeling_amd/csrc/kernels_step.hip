// Time-step kernels: the two RSG sweeps (velocity from stress, stress from
// velocity) for the three rotation forms, source injection, receiver gathers,
// the pressure snapshot field and the energy diagnostic.
//
// Arithmetic follows the reference's float expressions term by term with FP
// contraction disabled at build time, so each updated value is the one the
// reference's C computes.
#include "fdmod_internal.h"

namespace fdm {

constexpr int LZ = TZ + 2 * kN - 1;  // staged rows per tile column (halo 4/5)
constexpr int LX = TX + 2 * kN - 1;  // staged columns per tile

// Stage one field's halo tile into LDS: columns gx0 .. gx0+LX-1, rows gz0 .. gz0+LZ-1.
// Points outside the padded grid are never read by a valid output; they are set to 0.
__device__ __forceinline__ void stage(float (*sh)[LZ], const float *__restrict__ f, int gx0, int gz0, const Grid &g)
{
  for (int idx = threadIdx.x; idx < LX * LZ; idx += kThreads) {
    int cx = idx / LZ;
    int rz = idx - cx * LZ;
    int gx = gx0 + cx, gz = gz0 + rz;
    float v = 0.0f;
    if (gx < g.Nx && gz < g.Nz) v = f[(size_t)gx * g.ldz + gz];
    sh[cx][rz] = v;
  }
}

// A1: stress -> velocity point derivative pair at LDS position (X, Z) of (ix, iz),
// d1 along the (+x,+z) diagonal and d2 along the (+x,-z) diagonal (BT:661-699).
__device__ __forceinline__ void rsg_fwd(const float (*S)[LZ], int X, int Z, float &d1, float &d2)
{
  d1 = kC1 * (S[X + 1][Z + 1] - S[X][Z]) + kC2 * (S[X + 2][Z + 2] - S[X - 1][Z - 1]) +
       kC3 * (S[X + 3][Z + 3] - S[X - 2][Z - 2]) + kC4 * (S[X + 4][Z + 4] - S[X - 3][Z - 3]) +
       kC5 * (S[X + 5][Z + 5] - S[X - 4][Z - 4]);
  d2 = kC1 * (S[X + 1][Z] - S[X][Z + 1]) + kC2 * (S[X + 2][Z - 1] - S[X - 1][Z + 2]) +
       kC3 * (S[X + 3][Z - 2] - S[X - 2][Z + 3]) + kC4 * (S[X + 4][Z - 3] - S[X - 3][Z + 4]) +
       kC5 * (S[X + 5][Z - 4] - S[X - 4][Z + 5]);
}

// A2: velocity -> stress derivative pair (BT:843-861).
__device__ __forceinline__ void rsg_bwd(const float (*V)[LZ], int X, int Z, float &d1, float &d2)
{
  d1 = kC1 * (V[X][Z] - V[X - 1][Z - 1]) + kC2 * (V[X + 1][Z + 1] - V[X - 2][Z - 2]) +
       kC3 * (V[X + 2][Z + 2] - V[X - 3][Z - 3]) + kC4 * (V[X + 3][Z + 3] - V[X - 4][Z - 4]) +
       kC5 * (V[X + 4][Z + 4] - V[X - 5][Z - 5]);
  d2 = kC1 * (V[X][Z - 1] - V[X - 1][Z]) + kC2 * (V[X + 1][Z - 2] - V[X - 2][Z + 1]) +
       kC3 * (V[X + 2][Z - 3] - V[X - 3][Z + 2]) + kC4 * (V[X + 3][Z - 4] - V[X - 4][Z + 3]) +
       kC5 * (V[X + 4][Z - 5] - V[X - 5][Z + 4]);
}

// Velocity sweep (A1 + A3).  Bond and tensor: BT:654-704 / TR:617-666.
// Coordinate (COORD): CR:614-671, with the reference's invdx on d(sxx)/dz.
template <bool COORD>
__global__ __launch_bounds__(kThreads) void k_vsweep(SweepArgs a)
{
  __shared__ float sh[3][LX][LZ];
  const Grid &g = a.g;
  const int tz = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int z0 = g.izStart + blockIdx.x * TZ;
  const int x0 = g.ixStart + blockIdx.y * TX;
  stage(sh[0], a.sxx, x0 - 4, z0 - 4, g);
  stage(sh[1], a.szz, x0 - 4, z0 - 4, g);
  stage(sh[2], a.sxz, x0 - 4, z0 - 4, g);
  __syncthreads();
  const bool und = a.undamped[blockIdx.y * g.tilesZ + blockIdx.x] != 0;
  const int iz = z0 + tz;
  const float invdx = g.invdx, invdz = g.invdz;
#pragma unroll 1
  for (int j = 0; j < kColsPerWave; j++) {
    const int cx = w * kColsPerWave + j;
    const int ix = x0 + cx;
    if (ix >= g.ixEnd || iz >= g.izEnd) continue;
    const int X = cx + 4, Z = tz + 4;
    const size_t o = (size_t)ix * g.ldz + iz;
    float d1, d2;
    rsg_fwd(sh[0], X, Z, d1, d2);
    float sxxdx = 0.5f * (d1 + d2) * invdx;
    float sxxdz = 0.0f;
    if (COORD) sxxdz = 0.5f * (d1 - d2) * invdx;
    rsg_fwd(sh[1], X, Z, d1, d2);
    float szzdx = 0.0f;
    if (COORD) szzdx = 0.5f * (d1 + d2) * invdx;
    float szzdz = 0.5f * (d1 - d2) * invdz;
    rsg_fwd(sh[2], X, Z, d1, d2);
    float sxzdx = 0.5f * (d1 + d2) * invdx;
    float sxzdz = 0.5f * (d1 - d2) * invdz;
    const float k1 = und ? 1.0f : a.coef1[o];
    const float k2 = und ? g.dt : a.coef2[o];
    const float ir = a.invrho[o];
    const float kr = k2 * ir;
    if (!COORD) {
      a.vx[o] = k1 * a.vx[o] + kr * (sxxdx + sxzdz);
      a.vz[o] = k1 * a.vz[o] + kr * (sxzdx + szzdz);
    } else {
      const float cv = a.csv[o], sv = a.snv[o];
      a.vx[o] = k1 * a.vx[o] + kr * (cv * sxxdx - sv * sxxdz + sv * sxzdx + cv * sxzdz);
      a.vz[o] = k1 * a.vz[o] + kr * (cv * sxzdx - sv * sxzdz + sv * szzdx + cv * szzdz);
    }
  }
}

// Stress sweep (A2 + A4/A5/A6), with the free-surface stress imaging (BT:875-891)
// done by the top row of tiles, whose first rows are izStart .. izStart+5.
//   FORM_BOND   BT:836-873
//   FORM_TENSOR TR:799-836 and the sigma <- R(tau) sweep TR:839-852, fused pointwise
//   FORM_COORD  CR:804-845
template <int FORM>
__global__ __launch_bounds__(kThreads) void k_ssweep(SweepArgs a)
{
  __shared__ float sh[2][LX][LZ];
  const Grid &g = a.g;
  const int tz = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int z0 = g.izStart + blockIdx.x * TZ;
  const int x0 = g.ixStart + blockIdx.y * TX;
  stage(sh[0], a.vx, x0 - 5, z0 - 5, g);
  stage(sh[1], a.vz, x0 - 5, z0 - 5, g);
  __syncthreads();
  const bool und = a.undamped[blockIdx.y * g.tilesZ + blockIdx.x] != 0;
  const bool fs_top = a.freesurface && blockIdx.x == 0;
  const int iz = z0 + tz;
  const float invdx = g.invdx, invdz = g.invdz;
#pragma unroll 1
  for (int j = 0; j < kColsPerWave; j++) {
    const int cx = w * kColsPerWave + j;
    const int ix = x0 + cx;
    if (ix >= g.ixEnd) break;  // wave-uniform
    const bool valid = iz < g.izEnd;
    const size_t o = (size_t)ix * g.ldz + iz;
    float nxx = 0.0f, nzz = 0.0f, nxz = 0.0f;
    if (valid) {
      const int X = cx + 5, Z = tz + 5;
      float d1, d2;
      rsg_bwd(sh[0], X, Z, d1, d2);
      float vxdx = 0.5f * (d1 + d2) * invdx;
      float vxdz = 0.5f * (d1 - d2) * invdz;
      rsg_bwd(sh[1], X, Z, d1, d2);
      float vzdx = 0.5f * (d1 + d2) * invdx;
      float vzdz = 0.5f * (d1 - d2) * invdz;
      const float k1 = und ? 1.0f : a.coef1[o];
      const float k2 = und ? g.dt : a.coef2[o];
      if (FORM == FDMOD_FORM_BOND) {
        const float c11 = a.c11[o], c13 = a.c13[o], c15 = a.c15[o];
        const float c33 = a.c33[o], c35 = a.c35[o], c55 = a.c55[o];
        nxx = k1 * a.sxx[o] + k2 * (c11 * vxdx + c13 * vzdz + c15 * (vzdx + vxdz));
        nzz = k1 * a.szz[o] + k2 * (c13 * vxdx + c33 * vzdz + c35 * (vzdx + vxdz));
        nxz = k1 * a.sxz[o] + k2 * (c15 * vxdx + c35 * vzdz + c55 * (vzdx + vxdz));
      } else if (FORM == FDMOD_FORM_COORD) {
        const float sn = a.sn[o], cs = a.cs[o];
        float t = sn * vxdx + cs * vxdz;
        vxdx = cs * vxdx - sn * vxdz;
        vxdz = t;
        t = sn * vzdx + cs * vzdz;
        vzdx = cs * vzdx - sn * vzdz;
        vzdz = t;
        const float c11 = a.c11[o], c13 = a.c13[o], c33 = a.c33[o], c55 = a.c55[o];
        nxx = k1 * a.sxx[o] + k2 * (c11 * vxdx + c13 * vzdz);
        nzz = k1 * a.szz[o] + k2 * (c13 * vxdx + c33 * vzdz);
        nxz = k1 * a.sxz[o] + k2 * c55 * (vzdx + vxdz);
      } else {
        const float S2 = a.sin2[o], K2 = a.cos2[o], S2T = a.sin2t[o];
        const float exx = K2 * vxdx + S2 * vzdz - 0.5f * S2T * (vxdz + vzdx);
        const float ezz = S2 * vxdx + K2 * vzdz + 0.5f * S2T * (vxdz + vzdx);
        const float exz = S2T * (vxdx - vzdz) + (K2 - S2) * (vxdz + vzdx);
        const float c11 = a.c11[o], c13 = a.c13[o], c33 = a.c33[o], c55 = a.c55[o];
        const float t1 = k1 * a.txx[o] + k2 * (c11 * exx + c13 * ezz);
        const float t2 = k1 * a.tzz[o] + k2 * (c13 * exx + c33 * ezz);
        const float t3 = k1 * a.txz[o] + k2 * c55 * exz;
        a.txx[o] = t1;
        a.tzz[o] = t2;
        a.txz[o] = t3;
        nxx = K2 * t1 + S2 * t2 + S2T * t3;
        nzz = S2 * t1 + K2 * t2 - S2T * t3;
        nxz = -0.5f * S2T * (t1 - t2) + (K2 - S2) * t3;
      }
    } else if (fs_top && iz < g.Nz) {
      // rows past izEnd are not updated; their current values feed the images
      nxx = a.sxx[o];
      nzz = a.szz[o];
      nxz = a.sxz[o];
    }
    if (fs_top) {
      if (tz == 0) {
        nxz = 0.0f;
        nzz = 0.0f;
      }
      const float s0 = __shfl(nxx, 0, 64);
      if (tz >= 1 && tz <= kN) {
        const size_t m = (size_t)ix * g.ldz + (z0 - tz);
        a.sxx[m] = 2.0f * s0 - nxx;
        a.szz[m] = -nzz;
        a.sxz[m] = -nxz;
      }
    }
    if (valid) {
      a.sxx[o] = nxx;
      a.szz[o] = nzz;
      a.sxz[o] = nxz;
    }
  }
}

void launch_vsweep(hipStream_t s, int form, const SweepArgs &a)
{
  dim3 grid(a.g.tilesZ, a.g.tilesX);
  if (form == FDMOD_FORM_COORD) hipLaunchKernelGGL(k_vsweep<true>, grid, dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL(k_vsweep<false>, grid, dim3(kThreads), 0, s, a);
}

void launch_ssweep(hipStream_t s, int form, const SweepArgs &a)
{
  dim3 grid(a.g.tilesZ, a.g.tilesX);
  if (form == FDMOD_FORM_COORD) hipLaunchKernelGGL(k_ssweep<FDMOD_FORM_COORD>, grid, dim3(kThreads), 0, s, a);
  else if (form == FDMOD_FORM_TENSOR) hipLaunchKernelGGL(k_ssweep<FDMOD_FORM_TENSOR>, grid, dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL(k_ssweep<FDMOD_FORM_BOND>, grid, dim3(kThreads), 0, s, a);
}

// Source injection (BT:706-834 on velocity, BT:893-914 on stress).  Each pass is a
// table of taps; one thread per tap adds ca*src[it] to fa (and fc) and cb*src[it]
// to fb (and fd) at the tap's cell when it lies inside the grid (the reference's
// MAX/MIN clipping).  Passes run in order (stype 4: divergence, then curl), and
// within a pass the sources are walked in order inside one block when their
// footprints may overlap, so every cell sees the additions in the reference's order.
__global__ void k_source(SourceArgs a)
{
  const Grid &g = a.g;
  const int first = a.sequential ? 0 : blockIdx.x;
  const int last = a.sequential ? a.ns : blockIdx.x + 1;
  for (int p = 0; p < a.npass; p++) {
    for (int is = first; is < last; is++) {
      const float sv = a.src[(size_t)is * a.nt + a.it];
      for (int t = threadIdx.x; t < a.ntap; t += blockDim.x) {
        const SrcTap tp = a.taps[p * a.ntap + t];
        const int ix = a.isx[is] + tp.ox, iz = a.isz[is] + tp.oz;
        if (ix < 0 || ix >= g.Nx || iz < 0 || iz >= g.Nz) continue;
        const size_t o = (size_t)ix * g.ldz + iz;
        const float da = tp.ca * sv;
        a.fa[o] = a.fa[o] + da;
        if (a.fc) a.fc[o] = a.fc[o] + da;
        if (a.fb) {
          const float db = tp.cb * sv;
          a.fb[o] = a.fb[o] + db;
          if (a.fd) a.fd[o] = a.fd[o] + db;
        }
      }
      __syncthreads();
    }
  }
}

void launch_source(hipStream_t s, const SourceArgs &a)
{
  dim3 grid(a.sequential ? 1 : a.ns);
  hipLaunchKernelGGL(k_source, grid, dim3(128), 0, s, a);
}

// elastic2dReceiver (L/fdmod/isoelastic.c:103-144): nearest-point gathers into a
// device buffer laid out [sample][trace].
__global__ void k_receivers(RecvArgs a)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nr) return;
  size_t o = (size_t)a.gx[r] * a.ldz + a.gz[r];
  size_t d = (size_t)a.isample * a.nr + r;
  if (a.want_v) {
    a.out_vx[d] = a.vx[o];
    a.out_vz[d] = a.vz[o];
  }
  if (a.want_p) a.out_p[d] = 0.5f * (a.sxx[o] + a.szz[o]);
}

void launch_receivers(hipStream_t s, const RecvArgs &a)
{
  hipLaunchKernelGGL(k_receivers, dim3((a.nr + 255) / 256), dim3(256), 0, s, a);
}

// Pressure snapshot field p = 0.5*(sxx + szz) (isoelastic.c:259-266).
__global__ void k_pressure(const float *sxx, const float *szz, float *p, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.5f * (sxx[i] + szz[i]);
}

void launch_pressure(hipStream_t s, const float *sxx, const float *szz, float *p, int Nx, int ldz)
{
  size_t n = (size_t)Nx * ldz;
  hipLaunchKernelGGL(k_pressure, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sxx, szz, p, n);
}

// Energy diagnostic (BT:926-952).  Per-block float partial sums of the kinetic and
// potential densities over the unpadded model, then one fixed-order final sum.
__global__ void k_energy_partial(Grid g, int nx, int nz, DevArrays d, float *partial)
{
  __shared__ float sk[256], sp[256];
  float ek = 0.0f, ep = 0.0f;
  const float *vx = d.a[FDMOD_A_VX], *vz = d.a[FDMOD_A_VZ], *ir = d.a[FDMOD_A_INVRHO];
  const float *c11 = d.a[FDMOD_A_C11], *c55 = d.a[FDMOD_A_C55];
  const float *sxx = d.a[FDMOD_A_SXX], *szz = d.a[FDMOD_A_SZZ], *sxz = d.a[FDMOD_A_SXZ];
  size_t n = (size_t)nx * nz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int ix = (int)(i / nz), iz = (int)(i - (size_t)ix * nz);
    size_t o = (size_t)(ix + g.e) * g.ldz + (iz + g.e);
    float a11 = c11[o], a55 = c55[o], s1 = sxx[o], s2 = szz[o], s3 = sxz[o];
    ek += vx[o] * vx[o] / ir[o] + vz[o] * vz[o] / ir[o];
    ep += 0.25f * (((a11 + 2.0f * a55) * s1 - a11 * s2) * s1 + (a11 * s1 - (a11 + 2.0f * a55) * s2) * s2) /
              (a55 * (a11 + a55)) +
          s3 * s3 / a55;
  }
  sk[threadIdx.x] = ek;
  sp[threadIdx.x] = ep;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      sk[threadIdx.x] += sk[threadIdx.x + st];
      sp[threadIdx.x] += sp[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sk[0];
    partial[2 * blockIdx.x + 1] = sp[0];
  }
}

__global__ void k_energy_final(const float *partial, int nb, float dx, float dz, float *E, float *Ek, float *Ep,
                               int it)
{
  if (threadIdx.x != 0) return;
  double ek = 0.0, ep = 0.0;
  for (int b = 0; b < nb; b++) {
    ek += partial[2 * b];
    ep += partial[2 * b + 1];
  }
  float fk = 0.5f * (float)ek * dx * dz;
  float fp = 0.5f * (float)ep * dx * dz;
  Ek[it] = fk;
  Ep[it] = fp;
  E[it] = fk + fp;
}

void launch_energy_partial(hipStream_t s, const Grid &g, int nx, int nz, const DevArrays &d, float *partial,
                           int nblocks)
{
  hipLaunchKernelGGL(k_energy_partial, dim3(nblocks), dim3(256), 0, s, g, nx, nz, d, partial);
}

void launch_energy_final(hipStream_t s, const float *partial, int nblocks, float dx, float dz, float *E, float *Ek,
                         float *Ep, int it)
{
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, s, partial, nblocks, dx, dz, E, Ek, Ep, it);
}

}  // namespace fdm
