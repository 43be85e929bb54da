// Adjoint kernels of the modelling loop and of the L2 trace-misfit gradient
// (Bond form).  No reference counterpart: the reference only runs the forward loop.
//
// One forward step is u_{k+1} = FS(S(V(u_k) + s^v_k)) + s^sigma_k and d_k = R u_{k+1}.
// Going backwards through step k the adjoint state mu (same five fields) takes
//   mu += R^T r_k;  mu <- FS^T mu;  G += mu_sigma (x) E(v_{k+1}) (in blocks of steps);
//   mu_v <- coef1 mu_v + E^T(coef2 C mu_sigma);  mu_sigma <- coef1 mu_sigma + D^T(coef2 invrho mu_v)
// where D (stress -> velocity, offsets -4..+5) and E (velocity -> strain, -5..+4) are
// the two RSG derivative operators.  On the grid D^T = -E-shaped and E^T = -D-shaped
// (d1 and d2 of one operator are minus the transposes of the other's), so:
//   adjoint of the S sweep = a velocity-sweep-shaped stencil over M = coef2 C mu_sigma
//   adjoint of the V sweep = a stress-sweep-shaped stencil over N = coef2 invrho mu_v
// Both are gathers, like the forward sweeps, and reuse their staged-column rings.
// The coef1 scaling of each pass is applied by the *other* sweep (mu_sigma's by the
// V-adjoint, mu_v's by the next step's S-adjoint), so that no sweep reads and writes
// the same field: a sweep stages one field family and updates only the other.
//
// Update ranges.  Adjoint fields outside the update range U stay zero when read:
// mu_v is only written inside U, and the image rows above a free surface, which the
// V-adjoint does write, are folded back and cleared by FS^T before the S-adjoint
// stages mu_sigma.  So M and N need no explicit masks.
#include "sweep_common.h"

namespace fdm {

// ---------------------------------------------------------------- S-sweep adjoint
// Staged per row: mu_sigma (3), the six stiffness values and coef2.
struct MRow {
  float s1, s2, s3, c11, c13, c15, c33, c35, c55, k2;
};
struct MCol {
  MRow lo, hi;
};

__device__ __forceinline__ void load_mrow(MRow &r, const AdjArgs &a, size_t o)
{
  r.s1 = a.mxx[o]; r.s2 = a.mzz[o]; r.s3 = a.mxz[o];
  r.c11 = a.c11[o]; r.c13 = a.c13[o]; r.c15 = a.c15[o];
  r.c33 = a.c33[o]; r.c35 = a.c35[o]; r.c55 = a.c55[o];
  r.k2 = a.coef2[o];
}

__device__ __forceinline__ void zero_mrow(MRow &r) { r = MRow{0, 0, 0, 0, 0, 0, 0, 0, 0, 0}; }

__device__ __forceinline__ void load_mcol(MCol &c, const AdjArgs &a, int gx, int rs, int lane, const Grid &g)
{
  const size_t base = (size_t)gx * g.ldz;
  const int r0 = rs + lane, r1 = rs + TZ + lane;
  if (r0 >= 0 && r0 < g.Nz) load_mrow(c.lo, a, base + r0);
  else zero_mrow(c.lo);
  if (lane < kHalo && r1 >= 0 && r1 < g.Nz) load_mrow(c.hi, a, base + r1);
  else zero_mrow(c.hi);
}

// M = coef2 * C * mu_sigma: the adjoints of (vxdx, vzdz, vzdx + vxdz) at a stress point.
__device__ __forceinline__ void m_of(const MRow &r, float &mxx, float &mzz, float &mg)
{
  mxx = r.k2 * (r.c11 * r.s1 + r.c13 * r.s2 + r.c15 * r.s3);
  mzz = r.k2 * (r.c13 * r.s1 + r.c33 * r.s2 + r.c35 * r.s3);
  mg = r.k2 * (r.c15 * r.s1 + r.c35 * r.s2 + r.c55 * r.s3);
}

__device__ __forceinline__ void store_mcol(float (*ring)[kRing][LZ], int slot, const MCol &c, int lane)
{
  float x, z, y;
  m_of(c.lo, x, z, y);
  ring[0][slot][lane] = x;
  ring[1][slot][lane] = z;
  ring[2][slot][lane] = y;
  if (lane < kHalo) {
    m_of(c.hi, x, z, y);
    ring[0][slot][TZ + lane] = x;
    ring[1][slot][TZ + lane] = z;
    ring[2][slot][TZ + lane] = y;
  }
}

struct VAdjOps {
  float mvx, mvz, k1;
};

// mu_v(p) <- coef1(p) mu_v(p) - (Dx M_xx + Dz M_g, Dx M_g + Dz M_zz)(p), p in U.
template <int PD>
__device__ __forceinline__ void adj_s_tile(const AdjArgs &a, int btz, int btx)
{
  __shared__ float ring[3][kRing][LZ];
  const Grid &g = a.g;
  const int segL = g.segLV;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * TZ;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const int iz = z0 + lane;
  const bool zok = iz < g.izEnd;
  const int rs = z0 - 4;
  // prologue in groups of three columns (a staged row is 10 floats)
#pragma unroll
  for (int m0 = 0; m0 < 9; m0 += 3) {
    MCol c[3];
#pragma unroll
    for (int m = 0; m < 3; m++) load_mcol(c[m], a, xa - 4 + m0 + m, rs, lane, g);
#pragma unroll
    for (int m = 0; m < 3; m++) store_mcol(ring, m0 + m, c[m], lane);
  }
  MCol colq[PD];
  VAdjOps opq[PD];
  auto load_ops = [&](VAdjOps &p, int col) {
    const size_t o = (size_t)col * g.ldz + iz;
    p.mvx = a.mvx[o];
    p.mvz = a.mvz[o];
    p.k1 = a.coef1[o];
  };
#pragma unroll
  for (int i = 0; i < PD; i++) {
    if (i <= nc - 1) load_mcol(colq[i], a, xa + 5 + i, rs, lane, g);
    if (zok && i < nc) load_ops(opq[i], xa + i);
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = xa + k;
      MCol cn;
      VAdjOps on;
      if (k <= nc - PD - 1) load_mcol(cn, a, c + 5 + PD, rs, lane, g);
      if (zok && k + PD < nc) load_ops(on, c + PD);
      store_mcol(ring, (u + 9) % kRing, colq[0], lane);
      __syncthreads();
      if (zok) {
        float d1, d2;
        rsg_fwd([&](int dx, int dz) { return ring[0][(u + dx + 4 + kRing) % kRing][lane + 4 + dz]; }, d1, d2);
        const float mxxdx = 0.5f * (d1 + d2) * invdx;
        rsg_fwd([&](int dx, int dz) { return ring[1][(u + dx + 4 + kRing) % kRing][lane + 4 + dz]; }, d1, d2);
        const float mzzdz = 0.5f * (d1 - d2) * invdz;
        rsg_fwd([&](int dx, int dz) { return ring[2][(u + dx + 4 + kRing) % kRing][lane + 4 + dz]; }, d1, d2);
        const float mgdx = 0.5f * (d1 + d2) * invdx;
        const float mgdz = 0.5f * (d1 - d2) * invdz;
        const size_t o = (size_t)c * g.ldz + iz;
        const VAdjOps &op = opq[0];
        a.mvx[o] = op.k1 * op.mvx - (mxxdx + mgdz);
        a.mvz[o] = op.k1 * op.mvz - (mgdx + mzzdz);
      }
#pragma unroll
      for (int i = 0; i + 1 < PD; i++) {
        colq[i] = colq[i + 1];
        opq[i] = opq[i + 1];
      }
      colq[PD - 1] = cn;
      opq[PD - 1] = on;
    });
  }
}

// ---------------------------------------------------------------- V-sweep adjoint
// Staged per row: mu_v (2), coef2, invrho.
struct NRow {
  float vx, vz, k2, ir;
};
struct NCol {
  NRow lo, hi;
};

__device__ __forceinline__ void load_ncol(NCol &c, const AdjArgs &a, int gx, int rs, int lane, const Grid &g)
{
  const size_t base = (size_t)gx * g.ldz;
  const int r0 = rs + lane, r1 = rs + TZ + lane;
  if (r0 >= 0 && r0 < g.Nz) {
    const size_t o = base + r0;
    c.lo = NRow{a.mvx[o], a.mvz[o], a.coef2[o], a.invrho[o]};
  } else {
    c.lo = NRow{0, 0, 0, 0};
  }
  if (lane < kHalo && r1 >= 0 && r1 < g.Nz) {
    const size_t o = base + r1;
    c.hi = NRow{a.mvx[o], a.mvz[o], a.coef2[o], a.invrho[o]};
  } else {
    c.hi = NRow{0, 0, 0, 0};
  }
}

__device__ __forceinline__ void store_ncol(float (*ring)[kRing][LZ], int slot, const NCol &c, int lane)
{
  float kr = c.lo.k2 * c.lo.ir;
  ring[0][slot][lane] = kr * c.lo.vx;
  ring[1][slot][lane] = kr * c.lo.vz;
  if (lane < kHalo) {
    kr = c.hi.k2 * c.hi.ir;
    ring[0][slot][TZ + lane] = kr * c.hi.vx;
    ring[1][slot][TZ + lane] = kr * c.hi.vz;
  }
}

struct SAdjOps {
  float s1, s2, s3, k1;
};

// mu_sigma(q) <- k1(q) mu_sigma(q) - (Dx Nx, Dz Nz, Dz Nx + Dx Nz)(q), with k1 = coef1
// inside U and 1 on the image rows above a free surface (rows [g.izStart, a.izU)).
// The incoming mu_sigma is the one the imaging of this step uses; with a.snap set it
// is also written there (points of U) for the blocked imaging kernel.
template <int PD>
__device__ __forceinline__ void adj_v_tile(const AdjArgs &a, int btz, int btx)
{
  __shared__ float ring[2][kRing][LZ];
  const Grid &g = a.g;
  const int segL = g.segLS;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * TZ;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const int iz = z0 + lane;
  const bool zok = iz < g.izEnd;
  const bool inU = iz >= a.izU;
  const int rs = z0 - 5;
  {
    NCol c[9];
#pragma unroll
    for (int m = 0; m < 9; m++) load_ncol(c[m], a, xa - 5 + m, rs, lane, g);
#pragma unroll
    for (int m = 0; m < 9; m++) store_ncol(ring, m, c[m], lane);
  }
  NCol colq[PD];
  SAdjOps opq[PD];
  auto load_ops = [&](SAdjOps &p, int col) {
    const size_t o = (size_t)col * g.ldz + iz;
    p.s1 = a.mxx[o];
    p.s2 = a.mzz[o];
    p.s3 = a.mxz[o];
    p.k1 = inU ? a.coef1[o] : 1.0f;
  };
#pragma unroll
  for (int i = 0; i < PD; i++) {
    if (i <= nc - 1) load_ncol(colq[i], a, xa + 4 + i, rs, lane, g);
    if (zok && i < nc) load_ops(opq[i], xa + i);
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = xa + k;
      NCol cn;
      SAdjOps on;
      if (k <= nc - PD - 1) load_ncol(cn, a, c + 4 + PD, rs, lane, g);
      if (zok && k + PD < nc) load_ops(on, c + PD);
      store_ncol(ring, (u + 9) % kRing, colq[0], lane);
      __syncthreads();
      if (zok) {
        float d1, d2;
        rsg_bwd([&](int dx, int dz) { return ring[0][(u + dx + 5 + kRing) % kRing][lane + 5 + dz]; }, d1, d2);
        const float nxdx = 0.5f * (d1 + d2) * invdx;
        const float nxdz = 0.5f * (d1 - d2) * invdz;
        rsg_bwd([&](int dx, int dz) { return ring[1][(u + dx + 5 + kRing) % kRing][lane + 5 + dz]; }, d1, d2);
        const float nzdx = 0.5f * (d1 + d2) * invdx;
        const float nzdz = 0.5f * (d1 - d2) * invdz;
        const size_t o = (size_t)c * g.ldz + iz;
        const SAdjOps &op = opq[0];
        a.mxx[o] = op.k1 * op.s1 - nxdx;
        a.mzz[o] = op.k1 * op.s2 - nzdz;
        a.mxz[o] = op.k1 * op.s3 - (nxdz + nzdx);
        if (a.snap[0] && inU) {  // mu_sigma as imaged at this step (after FS^T)
          a.snap[0][o] = op.s1;
          a.snap[1][o] = op.s2;
          a.snap[2][o] = op.s3;
        }
      }
#pragma unroll
      for (int i = 0; i + 1 < PD; i++) {
        colq[i] = colq[i + 1];
        opq[i] = opq[i + 1];
      }
      colq[PD - 1] = cn;
      opq[PD - 1] = on;
    });
  }
}

// ---------------------------------------------------------------- FP16 storage
// Stored wavefields (checkpoints, the velocity history) in FP16 with one power-of-two
// scale per field, column and 64-row block: a block with maximum |v| < 2^E is stored
// as v * 2^(14-E) (exact scaling, |value| < 2^14 < 65504) and the block keeps the
// inverse scale 2^(E-14).  Arithmetic on the decoded values stays in FP32.
__device__ __forceinline__ float wave_max(float m)
{
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  return m;
}

// One wave per (column, block); blockIdx.y = field.  Rows [Nz, ldz) hold zeros.
__global__ __launch_bounds__(256) void k_pack16(Pack16 p)
{
  const int lane = threadIdx.x & 63;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nblk = p.ldz / TZ;
  if (w >= (long)p.nx * nblk) return;
  const int f = blockIdx.y;
  const int ix = (int)(w / nblk), b = (int)(w - (long)ix * nblk);
  const size_t o = (size_t)ix * p.ldz + (size_t)b * TZ + lane;
  const float v = p.src[f][o];
  const float m = wave_max(fabsf(v));
  int E = 0;
  if (m > 0.0f) (void)frexpf(m, &E);  // m < 2^E
  E = max(E, -100);                   // keeps 2^(14-E) and 2^(E-14) normal floats
  p.dst[f][o] = (_Float16)(v * ldexpf(1.0f, 14 - E));
  if (lane == 0) p.inv[f][(size_t)ix * nblk + b] = ldexpf(1.0f, E - 14);
}

__global__ __launch_bounds__(256) void k_unpack16(Pack16 p)
{
  const int lane = threadIdx.x & 63;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nblk = p.ldz / TZ;
  if (w >= (long)p.nx * nblk) return;
  const int f = blockIdx.y;
  const int ix = (int)(w / nblk), b = (int)(w - (long)ix * nblk);
  const size_t o = (size_t)ix * p.ldz + (size_t)b * TZ + lane;
  p.src[f][o] = (float)p.dst[f][o] * p.inv[f][(size_t)ix * nblk + b];
}

void launch_pack16(hipStream_t s, const Pack16 &p, bool pack)
{
  const long waves = (long)p.nx * (p.ldz / TZ);
  dim3 grid((unsigned)((waves + 3) / 4), p.nf);
  if (pack) hipLaunchKernelGGL(k_pack16, grid, dim3(256), 0, s, p);
  else hipLaunchKernelGGL(k_unpack16, grid, dim3(256), 0, s, p);
}

// ---------------------------------------------------------------- imaging
// G_j(q) += sum over a block of nk steps of (mu_sigma (x) E(v_{k+1}))_j, q in U, for
// the six Bond stiffness arrays in the order c11, c13, c15, c33, c35, c55, where
// E = (vxdx, vzdz, vzdx + vxdz) is the strain the forward stress sweep computed at q
// and stored, and mu_sigma the snapshot the V-adjoint wrote.  Pointwise: one thread
// per point keeps the six sums of the block in FP64 registers, so the FP64 arrays are
// read and written once per block of steps.  coef2 is time-independent and
// multiplies the sums in the chain-rule kernel.
template <bool F16>
__global__ __launch_bounds__(256) void k_image(AdjArgs a, ImgBlock b)
{
  const Grid &g = a.g;
  const int iz = g.izStart + blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = g.ixStart + blockIdx.y;
  if (iz >= g.izEnd) return;
  const size_t o = (size_t)ix * g.ldz + iz;
  const size_t so = (size_t)ix * g.tilesZ + (iz - g.izStart) / TZ;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < kImgK; j++) {
    if (j >= b.nk) break;
    const double X = b.lam[j][0][o], Z = b.lam[j][1][o], Y = b.lam[j][2][o];
    double ex, ez, eg;
    if (F16) {
      const float sc = b.esc[j][0][so];
      ex = (float)b.e16[j][0][o] * sc;
      ez = (float)b.e16[j][1][o] * sc;
      eg = (float)b.e16[j][2][o] * sc;
    } else {
      ex = b.e32[j][0][o];
      ez = b.e32[j][1][o];
      eg = b.e32[j][2][o];
    }
    acc[0] += X * ex;
    acc[1] += X * ez + Z * ex;
    acc[2] += X * eg + Y * ex;
    acc[3] += Z * ez;
    acc[4] += Z * eg + Y * ez;
    acc[5] += Y * eg;
  }
#pragma unroll
  for (int q = 0; q < 6; q++) a.G[q][o] += acc[q];
}

constexpr int kPipeAdj = 2;

__global__ __launch_bounds__(64) void k_adj_ssweep(AdjArgs a)
{
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLV, btz, btx)) return;
  adj_s_tile<kPipeAdj>(a, btz, btx);
}

__global__ __launch_bounds__(64) void k_adj_vsweep(AdjArgs a)
{
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLS, btz, btx)) return;
  adj_v_tile<kPipeAdj>(a, btz, btx);
}

static dim3 tiles_grid(const Grid &g, int segL)
{
  const int n = g.tilesZ * ((g.ixEnd - g.ixStart + segL - 1) / segL);
  return dim3((n + kXCD - 1) / kXCD * kXCD);
}

void launch_adj_ssweep(hipStream_t s, const AdjArgs &a)
{
  hipLaunchKernelGGL(k_adj_ssweep, tiles_grid(a.g, a.g.segLV), dim3(64), 0, s, a);
}

void launch_adj_vsweep(hipStream_t s, const AdjArgs &a)
{
  hipLaunchKernelGGL(k_adj_vsweep, tiles_grid(a.g, a.g.segLS), dim3(64), 0, s, a);
}

void launch_image(hipStream_t s, const AdjArgs &a, const ImgBlock &b)
{
  dim3 grid((a.g.izEnd - a.g.izStart + 255) / 256, a.g.ixEnd - a.g.ixStart);
  if (b.e16[0][0]) hipLaunchKernelGGL(k_image<true>, grid, dim3(256), 0, s, a, b);
  else hipLaunchKernelGGL(k_image<false>, grid, dim3(256), 0, s, a, b);
}

// ---------------------------------------------------------------- free surface
// Transpose of the stress imaging (BT:875-891): row izU-k was set to
// (2 sxx(izU) - sxx(izU+k), -szz(izU+k), -sxz(izU+k)) and szz, sxz of row izU to 0.
__global__ void k_adj_fs(AdjArgs a)
{
  const int ix = a.g.ixStart + blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= a.g.ixEnd) return;
  const size_t b = (size_t)ix * a.g.ldz;
  const int z0 = a.izU;
  float acc = 0.0f;
  for (int k = kN; k >= 1; k--) {
    const size_t m = b + z0 - k, p = b + z0 + k;
    const float x = a.mxx[m], z = a.mzz[m], y = a.mxz[m];
    acc += 2.0f * x;
    a.mxx[p] -= x;
    a.mzz[p] -= z;
    a.mxz[p] -= y;
    a.mxx[m] = 0.0f;
    a.mzz[m] = 0.0f;
    a.mxz[m] = 0.0f;
  }
  a.mxx[b + z0] += acc;
  a.mzz[b + z0] = 0.0f;
  a.mxz[b + z0] = 0.0f;
}

void launch_adj_fs(hipStream_t s, const AdjArgs &a)
{
  const int n = a.g.ixEnd - a.g.ixStart;
  hipLaunchKernelGGL(k_adj_fs, dim3((n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------- data-space sources
// Transpose of elastic2dReceiver (isoelastic.c:103-144) for one sample: pressure
// residuals go half to mu_sxx and half to mu_szz, velocity residuals to mu_vx, mu_vz.
// Receivers may share a cell, hence the atomics.
__global__ void k_adj_inject(int nr, int ldz, const int *gx, const int *gz, const float *r1, const float *r2,
                             float *f1, float *f2, float scale)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const size_t o = (size_t)gx[r] * ldz + gz[r];
  atomicAdd(f1 + o, scale * r1[r]);
  atomicAdd(f2 + o, scale * r2[r]);
}

void launch_adj_inject(hipStream_t s, int nr, int ldz, const int *gx, const int *gz, const float *r1,
                       const float *r2, float *f1, float *f2, float scale)
{
  hipLaunchKernelGGL(k_adj_inject, dim3((nr + 255) / 256), dim3(256), 0, s, nr, ldz, gx, gz, r1, r2, f1, f2, scale);
}

// Transpose of the source injection of one step (source_body): the adjoint of
// src[is][it] is the tap-weighted sum of the adjoint fields the taps wrote into.
// One block per source; the sum is in double.
__global__ void k_adj_source(SourceArgs a, const float *ma, const float *mb, double *out)
{
  __shared__ double red[128];
  const Grid &g = a.g;
  const int is = blockIdx.x;
  double acc = 0.0;
  for (int p = 0; p < a.npass; p++)
    for (int t = threadIdx.x; t < a.ntap; t += blockDim.x) {
      const SrcTap tp = a.taps[p * a.ntap + t];
      const int ix = a.isx[is] + tp.ox, iz = a.isz[is] + tp.oz;
      if (ix < 0 || ix >= g.Nx || iz < 0 || iz >= g.Nz) continue;
      const size_t o = (size_t)ix * g.ldz + iz;
      acc += (double)tp.ca * ma[o];
      if (mb) acc += (double)tp.cb * mb[o];
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 64; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(size_t)is * a.nt + a.it] = red[0];
}

void launch_adj_source(hipStream_t s, const SourceArgs &a, const float *ma, const float *mb, double *out)
{
  hipLaunchKernelGGL(k_adj_source, dim3(a.ns), dim3(128), 0, s, a, ma, mb, out);
}

// ---------------------------------------------------------------- FP64 reductions
// r = d - obs (in place over obs) and per-block partial sums of r^2 / 2.
__global__ void k_residual(const float *d, float *r, size_t n, double *partial)
{
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = d[i] - r[i];
    r[i] = v;
    acc += 0.5 * (double)v * (double)v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// partial sums of a . b with a float and b float (bd == null) or double
__global__ void k_dot(const float *a, const float *bf, const double *bd, size_t n, double *partial)
{
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (bd ? bd[i] : (double)bf[i]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] += sum of nb partials, in a fixed order
__global__ void k_sum_partials(const double *partial, int nb, double *out)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += partial[b];
  out[0] += s;
}

void launch_residual(hipStream_t s, const float *d, float *r, size_t n, double *partial, int nb, double *out)
{
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, d, r, n, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nb, out);
}

void launch_dot(hipStream_t s, const float *a, const float *bf, const double *bd, size_t n, double *partial, int nb,
                double *out)
{
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(256), 0, s, a, bf, bd, n, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nb, out);
}

// ---------------------------------------------------------------- chain rule
// Forward-mode dual numbers over the four model directions (vp, eps, delta, theta).
struct D4 {
  double v, d[4];
};
__device__ __forceinline__ D4 cst(double v) { return D4{v, {0, 0, 0, 0}}; }
__device__ __forceinline__ D4 seed(double v, int k)
{
  D4 r = cst(v);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ D4 operator+(D4 a, D4 b)
{
  for (int k = 0; k < 4; k++) a.d[k] += b.d[k];
  a.v += b.v;
  return a;
}
__device__ __forceinline__ D4 operator-(D4 a, D4 b)
{
  for (int k = 0; k < 4; k++) a.d[k] -= b.d[k];
  a.v -= b.v;
  return a;
}
__device__ __forceinline__ D4 operator*(D4 a, D4 b)
{
  D4 r;
  r.v = a.v * b.v;
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D4 operator*(double s, D4 a)
{
  a.v *= s;
  for (int k = 0; k < 4; k++) a.d[k] *= s;
  return a;
}
__device__ __forceinline__ D4 operator/(D4 a, D4 b)
{
  D4 r;
  r.v = a.v / b.v;
  for (int k = 0; k < 4; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
__device__ __forceinline__ D4 dsqrt(D4 a)
{
  D4 r;
  r.v = sqrt(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
__device__ __forceinline__ D4 dsin(D4 a)
{
  D4 r;
  r.v = sin(a.v);
  const double c = cos(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = c * a.d[k];
  return r;
}
__device__ __forceinline__ D4 dcos(D4 a)
{
  D4 r;
  r.v = cos(a.v);
  const double s = -sin(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = s * a.d[k];
  return r;
}
__device__ __forceinline__ bool bad(double v) { return isnan(v) || isinf(v); }

// Thomsen -> Bond-rotated stiffness (BT:493-526) with the Bond free-surface fix-up
// (BT:543-568) on the surface row, as exact functions of (vp, eps, delta, theta).
__device__ void bond_stiffness(D4 vp, double vs, double rho, D4 eps, D4 del, D4 t, bool fixup, D4 c[6])
{
  const D4 r = cst(vs) / vp;
  const D4 f = cst(1.0) - r * r;
  const D4 C11 = rho * (vp * vp) * (cst(1.0) + 2.0 * eps);
  const D4 C13 = rho * (vp * vp * dsqrt(f * (f + 2.0 * del)) - cst(vs * vs));
  const D4 C33 = rho * (vp * vp);
  const D4 C55 = cst(rho * vs * vs);
  const D4 s = dsin(t), co = dcos(t);
  const D4 s2 = s * s, c2 = co * co, s4 = s2 * s2, c4 = c2 * c2;
  const D4 sn2 = dsin(2.0 * t), cs2 = dcos(2.0 * t), sn22 = sn2 * sn2, cs22 = cs2 * cs2;
  c[0] = C11 * c4 + C33 * s4 + (0.5 * C13 + C55) * sn22;
  c[1] = (0.25 * (C11 + C33) - C55) * sn22 + C13 * (s4 + c4);
  c[2] = 0.5 * ((cst(0.0) - C11 * c2) + C33 * s2 + (C13 + 2.0 * C55) * cs2) * sn2;
  c[3] = C11 * s4 + C33 * c4 + (0.5 * C13 + C55) * sn22;
  c[4] = 0.5 * ((cst(0.0) - C11 * s2) + C33 * c2 - (C13 + 2.0 * C55) * cs2) * sn2;
  c[5] = 0.25 * (C11 + C33 - 2.0 * C13) * sn22 + C55 * cs22;
  if (fixup) {
    const D4 c11 = c[0], c13 = c[1], c15 = c[2], c33 = c[3], c35 = c[4], c55 = c[5];
    D4 n35 = (cst(0.0) - (c13 * c55 - c15 * c35)) / (c33 * c55 - c35 * c35);
    const D4 n55 = (c13 * n35 - c15 * c33) / (c33 * c55 - n35 * n35);
    if (bad(n35.v)) n35 = (cst(0.0) - c13) / c33;
    if (bad(n35.v)) n35 = cst(0.0);
    c[0] = c11 - c13 * n35 + c15 * n55;
    for (int j = 1; j < 6; j++) c[j] = cst(0.0);
  }
}

// dJ/d(vp, eps, delta, theta[deg]) on the unpadded model.  Each unpadded cell sums
// the chain rule over the padded cells that replicate it (the transpose of padModel2d,
// L/absorbc/padmod.c:33-58).  G holds sum_t mu (x) E; coef2 multiplies it here.
__global__ void k_chain_bond(Grid g, int nx, int nz, int fs_row, DevArrays d, const double *G, size_t gstride,
                             double theta_scale, double *grad)
{
  const int jz = blockIdx.x * blockDim.x + threadIdx.x;
  const int jx = blockIdx.y;
  if (jz >= nz || jx >= nx) return;
  const int x0 = jx == 0 ? 0 : jx + g.e, x1 = jx == nx - 1 ? g.Nx - 1 : jx + g.e;
  const int z0 = jz == 0 ? 0 : jz + g.e, z1 = jz == nz - 1 ? g.Nz - 1 : jz + g.e;
  double acc[4] = {0, 0, 0, 0};
  for (int ix = x0; ix <= x1; ix++)
    for (int iz = z0; iz <= z1; iz++) {
      const size_t o = (size_t)ix * g.ldz + iz;
      double gc[6];
      bool any = false;
      const double k2 = d.a[FDMOD_A_COEF2][o];
      for (int j = 0; j < 6; j++) {
        gc[j] = k2 * G[j * gstride + o];
        any = any || gc[j] != 0.0;
      }
      if (!any) continue;
      D4 c[6];
      bond_stiffness(seed(d.a[FDMOD_A_VP][o], 0), d.a[FDMOD_A_VS][o], d.a[FDMOD_A_RHO][o],
                     seed(d.a[FDMOD_A_EPS][o], 1), seed(d.a[FDMOD_A_DELTA][o], 2), seed(d.a[FDMOD_A_THETA][o], 3),
                     iz == fs_row, c);
      const double beta = d.a[FDMOD_A_BETA][o];
      for (int m = 0; m < 4; m++) {
        double s = 0.0;
        for (int j = 0; j < 6; j++) s += gc[j] * c[j].d[m];
        acc[m] += m == 0 ? s / beta : (m == 3 ? s * theta_scale : s);
      }
    }
  for (int m = 0; m < 4; m++) grad[((size_t)m * nx + jx) * nz + jz] = acc[m];
}

void launch_chain_bond(hipStream_t s, const Grid &g, int nx, int nz, int fs_row, const DevArrays &d,
                       const double *G, size_t gstride, double theta_scale, double *grad)
{
  dim3 grid((nz + 127) / 128, nx);
  hipLaunchKernelGGL(k_chain_bond, grid, dim3(128), 0, s, g, nx, nz, fs_row, d, G, gstride, theta_scale, grad);
}

// Add one shot's crop gradient [4][nx][nz] into the model gradient [4][nx_full][nz]
// at column offset x0 (shots are added one after another, in shot order).
__global__ void k_accum_grad(const double *src, int nx, int nz, int x0, int nx_full, double *dst)
{
  const int jz = blockIdx.x * blockDim.x + threadIdx.x;
  const int jx = blockIdx.y, m = blockIdx.z;
  if (jz >= nz || jx >= nx) return;
  dst[((size_t)m * nx_full + x0 + jx) * nz + jz] += src[((size_t)m * nx + jx) * nz + jz];
}

void launch_accum_grad(hipStream_t s, const double *src, int nx, int nz, int x0, int nx_full, double *dst)
{
  hipLaunchKernelGGL(k_accum_grad, dim3((nz + 255) / 256, nx, 4), dim3(256), 0, s, src, nx, nz, x0, nx_full, dst);
}

// Frozen absorbing layers: a new padded model is rescaled by the stored Beta as in
// mealFDCoefs (meal.c:454-482: vp/=Beta, vs/=Beta, rho*=Beta), D and coefs kept.
__global__ void k_rescale_frozen(Grid g, float *vp, float *vs, float *rho, const float *Beta)
{
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  const size_t o = (size_t)ix * g.ldz + iz;
  const float b = Beta[o];
  vp[o] /= b;
  vs[o] /= b;
  rho[o] *= b;
}

void launch_rescale_frozen(hipStream_t s, const Grid &g, float *vp, float *vs, float *rho, const float *Beta)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_rescale_frozen, grid, dim3(256), 0, s, g, vp, vs, rho, Beta);
}

}  // namespace fdm
