// Setup kernels: model padding, MEAL damping, Thomsen -> Cij(theta), staggered
// parameter averaging, the Bond free-surface fix-up and the per-tile undamped flags.
//
// These run once per model.  Where the reference evaluates an expression in double
// (double literals promote the float operands) the kernels do the same, with FP
// contraction off, so the stored float parameters equal the reference's.
#include "fdmod_internal.h"

namespace fdm {

// padModel2d, L/absorbc/padmod.c:33-58.  Top/bottom replication on interior
// columns followed by left/right replication of full columns is the same as
// clamping both indices into the interior.  Rows in [Nz, ldz) are zeroed.
__global__ void k_pad(const float *__restrict__ raw, float *__restrict__ dst, int nz, int nx, Grid g)
{
  int iz = blockIdx.x * blockDim.x + threadIdx.x;
  int ix = blockIdx.y;
  if (iz >= g.ldz || ix >= g.Nx) return;
  float v = 0.0f;
  if (iz < g.Nz) {
    int sx = min(max(ix - g.e, 0), nx - 1);
    int sz = min(max(iz - g.e, 0), nz - 1);
    v = raw[(size_t)sx * nz + sz];
  }
  dst[(size_t)ix * g.ldz + iz] = v;
}

void launch_pad(hipStream_t s, const float *raw, float *dst, int nz, int nx, const Grid &g)
{
  dim3 grid((g.ldz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_pad, grid, dim3(256), 0, s, raw, dst, nz, nx, g);
}

// mealCoefs2d (L/absorbc/meal.c:51-201) evaluated per point, followed by
// mealFDCoefs order 1 (meal.c:450-491).  The layer weight w(l) = absorbWeight
// (meal.c:23-29) is a host-computed table of nlayer+1 floats.  Per point:
//   top/bottom layer l (top only without free surface): Dz = d0z*w, Bz = 1+(bz-1)*w
//   left/right layer l (rows iz >= izStart):            Dx = d0x*w, Bx = 1+(bx-1)*w
//   d0 = ((-1*(1+pd))*vp)*ln(R) / (2*n*h),  b = (2*(Vs/(h*fpeak)))/10
//   D = sqrt(Dx^2+Dz^2), Beta = sqrt((Bx-1)^2+(Bz-1)^2)+1
//   vp/=Beta, vs/=Beta, rho*=Beta, D/=Beta, coef1/coef2 in double.
__global__ void k_meal(Grid g, float *vp, float *vs, float *rho, float *D, float *Beta, float *coef1,
                       float *coef2, const float *__restrict__ wtab, int nlayer, int freesurface, float k0,
                       float lnR, float twoThx, float twoThz, float dxf, float dzf, float sqrt3, double dt)
{
  int iz = blockIdx.x * blockDim.x + threadIdx.x;
  int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  size_t o = (size_t)ix * g.ldz + iz;
  float p = vp[o], s = vs[o];
  float Vs = (s == 0.0f) ? p / sqrt3 : s;
  float Dx = 0.0f, Dz = 0.0f, Bx = 1.0f, Bz = 1.0f;
  int lz = 0;
  if (freesurface == 0 && iz < nlayer) lz = nlayer - iz;
  if (iz >= g.Nz - nlayer) lz = iz - (g.Nz - nlayer - 1);
  if (lz > 0) {
    float w = wtab[lz];
    float d0 = k0 * p * lnR / twoThz;
    float ppw = Vs / dzf;
    float b = 2.0f * ppw / 10.0f;
    Dz = d0 * w;
    Bz = 1.0f + (b - 1.0f) * w;
  }
  int izs = freesurface == 0 ? 0 : nlayer;
  int lx = 0;
  if (iz >= izs) {
    if (ix < nlayer) lx = nlayer - ix;
    else if (ix >= g.Nx - nlayer) lx = ix - (g.Nx - nlayer - 1);
  }
  if (lx > 0) {
    float w = wtab[lx];
    float d0 = k0 * p * lnR / twoThx;
    float ppw = Vs / dxf;
    float b = 2.0f * ppw / 10.0f;
    Dx = d0 * w;
    Bx = 1.0f + (b - 1.0f) * w;
  }
  float d = sqrtf(Dx * Dx + Dz * Dz);
  float bx = Bx - 1.0f, bz = Bz - 1.0f;
  float beta = sqrtf(bx * bx + bz * bz) + 1.0f;
  vp[o] = p / beta;
  vs[o] = s / beta;
  rho[o] = rho[o] * beta;
  d = d / beta;
  D[o] = d;
  Beta[o] = beta;
  double hd = 0.5 * dt * (double)d;
  coef1[o] = (float)((1.0 - hd) / (1.0 + hd));
  coef2[o] = (float)(dt / (1.0 + hd));
}

void launch_meal(hipStream_t s, const Grid &g, float *vp, float *vs, float *rho, float *D, float *Beta,
                 float *coef1, float *coef2, const float *wtab, int nlayer, int freesurface, float k0, float lnR,
                 float twoThx, float twoThz, float dxf, float dzf, float sqrt3, double dt)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_meal, grid, dim3(256), 0, s, g, vp, vs, rho, D, Beta, coef1, coef2, wtab, nlayer,
                     freesurface, k0, lnR, twoThx, twoThz, dxf, dzf, sqrt3, dt);
}

// Thomsen -> Cij(theta): Bond BT:493-526, tensor TR:483-503, coordinate CR:497-514.
// rho*vp*vp, vp*vp, vs*vs and rho*vs*vs are float products in the reference; the
// rest is double (pow(r, 2.0) of a float r is exactly r*r in double).  theta is
// converted to radians and stored back as float before the trigonometry.
__global__ void k_cij(Grid g, int form, DevArrays d, float scalar)
{
  int iz = blockIdx.x * blockDim.x + threadIdx.x;
  int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  size_t o = (size_t)ix * g.ldz + iz;
  float vp = d.a[FDMOD_A_VP][o], vs = d.a[FDMOD_A_VS][o], rho = d.a[FDMOD_A_RHO][o];
  float eps = d.a[FDMOD_A_EPS][o], del = d.a[FDMOD_A_DELTA][o];
  float r = vs / vp;
  double f = 1.0 - (double)r * (double)r;
  float rvv = rho * vp * vp;
  float vv = vp * vp;
  float ss = vs * vs;
  float rss = rho * vs * vs;
  double C11 = (double)rvv * (1.0 + 2.0 * (double)eps);
  double C13 = (double)rho * ((double)vv * sqrt(f * (f + 2.0 * (double)del)) - (double)ss);
  double C33 = (double)rvv;
  double C55 = (double)rss;
  float ths = scalar * d.a[FDMOD_A_THETA][o];
  float th = (float)((double)ths / 180.0 * kPI);
  d.a[FDMOD_A_THETA][o] = th;
  double t = (double)th;
  double sn = sin(t), cs = cos(t);
  if (form == FDMOD_FORM_BOND) {
    double s2 = sn * sn, c2 = cs * cs, s4 = s2 * s2, c4 = c2 * c2;
    double sn2 = sin(2.0 * t), cs2 = cos(2.0 * t);
    double sn22 = sn2 * sn2, cs22 = cs2 * cs2;
    d.a[FDMOD_A_C11][o] = (float)(C11 * c4 + C33 * s4 + (0.5 * C13 + C55) * sn22);
    d.a[FDMOD_A_C13][o] = (float)((0.25 * (C11 + C33) - C55) * sn22 + C13 * (s4 + c4));
    d.a[FDMOD_A_C15][o] = (float)(0.5 * (-1.0 * C11 * c2 + C33 * s2 + (C13 + 2.0 * C55) * cs2) * sn2);
    d.a[FDMOD_A_C33][o] = (float)(C11 * s4 + C33 * c4 + (0.5 * C13 + C55) * sn22);
    d.a[FDMOD_A_C35][o] = (float)(0.5 * (-1.0 * C11 * s2 + C33 * c2 - (C13 + 2.0 * C55) * cs2) * sn2);
    d.a[FDMOD_A_C55][o] = (float)(0.25 * (C11 + C33 - 2.0 * C13) * sn22 + C55 * cs22);
  } else {
    d.a[FDMOD_A_C11][o] = (float)C11;
    d.a[FDMOD_A_C13][o] = (float)C13;
    d.a[FDMOD_A_C33][o] = rvv;
    d.a[FDMOD_A_C55][o] = rss;
    if (form == FDMOD_FORM_TENSOR) {
      d.a[FDMOD_A_SIN2][o] = (float)(sn * sn);
      d.a[FDMOD_A_COS2][o] = (float)(cs * cs);
      d.a[FDMOD_A_SIN2T][o] = (float)sin(2.0 * t);
    } else {
      d.a[FDMOD_A_SIN][o] = (float)sn;
      d.a[FDMOD_A_COS][o] = (float)cs;
    }
  }
}

void launch_cij(hipStream_t s, const Grid &g, int form, const DevArrays &d, float scalar)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_cij, grid, dim3(256), 0, s, g, form, d, scalar);
}

// invrho = 4/(sum of the 4 surrounding rho) at velocity points, 1/rho on the last
// row and column (BT:527-542); coordinate form: 4-point means of sin/cos theta
// (CR:516-535).  The 4-term sums are float; 4.0/x, 1.0/x and 0.25*x are double.
__global__ void k_stagger(Grid g, int coord, DevArrays d)
{
  int iz = blockIdx.x * blockDim.x + threadIdx.x;
  int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  size_t L = g.ldz;
  size_t o = (size_t)ix * L + iz;
  const float *rho = d.a[FDMOD_A_RHO];
  const float *sn = d.a[FDMOD_A_SIN], *cs = d.a[FDMOD_A_COS];
  if (ix < g.Nx - 1 && iz < g.Nz - 1) {
    float sr = rho[o] + rho[o + L] + rho[o + 1] + rho[o + L + 1];
    d.a[FDMOD_A_INVRHO][o] = (float)(4.0 / (double)sr);
    if (coord) {
      float ss = sn[o] + sn[o + L] + sn[o + 1] + sn[o + L + 1];
      float sc = cs[o] + cs[o + L] + cs[o + 1] + cs[o + L + 1];
      d.a[FDMOD_A_SINV][o] = (float)(0.25 * (double)ss);
      d.a[FDMOD_A_COSV][o] = (float)(0.25 * (double)sc);
    }
  } else {
    d.a[FDMOD_A_INVRHO][o] = (float)(1.0 / (double)rho[o]);
    if (coord) {
      d.a[FDMOD_A_SINV][o] = sn[o];
      d.a[FDMOD_A_COSV][o] = cs[o];
    }
  }
}

void launch_stagger(hipStream_t s, const Grid &g, int form, const DevArrays &d)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_stagger, grid, dim3(256), 0, s, g, form == FDMOD_FORM_COORD ? 1 : 0, d);
}

// Bond free-surface parameter fix-up on row iz = ealThick (BT:543-568), in float,
// with the reference's use of the new c35 inside c55 and its c11 expression.
__global__ void k_fs_fixup(Grid g, DevArrays d)
{
  int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= g.Nx) return;
  size_t o = (size_t)ix * g.ldz + g.e;
  float *c11 = d.a[FDMOD_A_C11], *c13 = d.a[FDMOD_A_C13], *c15 = d.a[FDMOD_A_C15];
  float *c33 = d.a[FDMOD_A_C33], *c35 = d.a[FDMOD_A_C35], *c55 = d.a[FDMOD_A_C55];
  float a11 = c11[o], a13 = c13[o], a15 = c15[o], a33 = c33[o], a35 = c35[o], a55 = c55[o];
  a35 = -(a13 * a55 - a15 * a35) / (a33 * a55 - a35 * a35);
  a55 = (a13 * a35 - a15 * a33) / (a33 * a55 - a35 * a35);
  if (isnan(a35) || isinf(a35)) a35 = -a13 / a33;
  if (isnan(a35) || isinf(a35)) a35 = 0.0f;
  a11 = a11 - a13 * a35 + a15 * a55;
  c11[o] = a11;
  c13[o] = 0.0f;
  c15[o] = 0.0f;
  c33[o] = 0.0f;
  c35[o] = 0.0f;
  c55[o] = 0.0f;
}

void launch_fs_fixup(hipStream_t s, const Grid &g, const DevArrays &d)
{
  hipLaunchKernelGGL(k_fs_fixup, dim3((g.Nx + 255) / 256), dim3(256), 0, s, g, d);
}

// A sweep tile is "undamped" when coef1 == 1 and coef2 == dt at every output
// point; the sweeps then use those constants, which gives bit-identical results
// (1*x == x, and coef2*invrho == dt*invrho) while skipping two parameter loads.
__global__ void k_tile_flags(Grid g, const float *__restrict__ coef1, const float *__restrict__ coef2,
                             uint8_t *flags)
{
  int tz = blockIdx.x, tx = blockIdx.y;
  int z0 = g.izStart + tz * TZ, x0 = g.ixStart + tx * kFlagW;
  int ok = 1;
  for (int idx = threadIdx.x; idx < TZ * kFlagW; idx += blockDim.x) {
    int ix = x0 + idx / TZ, iz = z0 + idx % TZ;
    if (ix < g.ixEnd && iz < g.izEnd) {
      size_t o = (size_t)ix * g.ldz + iz;
      if (!(coef1[o] == 1.0f && coef2[o] == g.dt)) ok = 0;
    }
  }
  ok = __syncthreads_and(ok);
  if (threadIdx.x == 0) flags[tx * g.tilesZ + tz] = (uint8_t)ok;
}

void launch_tile_flags(hipStream_t s, const Grid &g, const float *coef1, const float *coef2, uint8_t *flags)
{
  hipLaunchKernelGGL(k_tile_flags, dim3(g.tilesZ, g.flagsX), dim3(256), 0, s, g, coef1, coef2, flags);
}

}  // namespace fdm
