/*
 * fdmod_oracle.c -- CPU restatement of the TTI-FDModeling rotated-staggered-grid
 * (RSG) 2-D TTI elastic modelling loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for libfdmod_hip: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * libraries built from it.  Nothing in the product links or calls it.
 *
 * Built twice from this one source:
 *   liboracle_f32.so  (no ORACLE_F64)  FP32 restatement that follows the reference's
 *                     float/double evaluation order operation by operation, so that
 *                     gcc -O3 -fopenmp (no FP contraction on x86-64) reproduces the
 *                     reference's OpenMP program.  It is the "reference OpenMP path".
 *   liboracle_f64.so  (-DORACLE_F64)   every real quantity in double: the FP64 oracle.
 *
 * Citations are relative to the reference tree:
 *   BT = src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_bond_transformation.c
 *   TR = src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_tensor_rotation.c
 *   CR = src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_coordinate_rotation.c
 *
 * Pinning (each function below is checked in tests/test_oracle_*.py by):
 *   pad_model          numpy edge padding (np.pad mode="edge")
 *   stability          the closed form on a constant model (hand value)
 *   meal_coefs         interior D=0/Beta=1 exactly, Rcoef/d0 hand values at the
 *                      outermost layer, every inner layer against 2^((l/n)^alpha) - 1,
 *                      monotone profile, corner = hypot of sides
 *   cij_*              Bond: 2-D Bond matrix R C R^T in numpy; Thomsen inversion
 *                      (eps, delta recovered from C11..C55); theta=0 identity
 *   invrho / trig avg  numpy slicing restatement on random tiny grids
 *   fs_fixup (Bond)    hand-computed VTI special case only -- parity unpinned
 *                      beyond it (the reference cannot be built here)
 *   v/s sweeps         exactness on linear fields (gain sum (2k-1)C_k), plane-wave
 *                      symbol, cross-form equality at constant theta, analytic 2-D
 *                      Green's function for an isotropic explosive source; the
 *                      coordinate form's invdx on d(sxx)/dz at dx != dz (CR:633)
 *   sources            closed-form Gaussian weights (also at dx != dz), linearity
 *   receivers          bit equality with the restated fields
 *   energy             hand values on constant states; bounded, oscillating kinetic
 *                      part without damping or source
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_F64
typedef double real;
#define RF(x) ((double)(x))
#define R_EXP exp
#define R_LOG log
#define R_POW pow
#define R_SQRT sqrt
#define R_LOG10 log10
#define PFX(name) oracle64_##name
#else
typedef float real;
#define RF(x) ((float)(x))
#define R_EXP expf
#define R_LOG logf
#define R_POW powf
#define R_SQRT sqrtf
#define R_LOG10 log10f
#define PFX(name) oracle32_##name
#endif

/* PI as defined by Seismic Unix cwp.h (not vendored in the reference). */
#define ORACLE_PI (3.141592653589793)

enum { FORM_COORD = 0, FORM_TENSOR = 1, FORM_BOND = 2 };

/* array ids for get/set (shared with the Python wrapper and libfdmod_hip) */
enum {
  A_VP = 0, A_VS, A_RHO, A_EPS, A_DEL, A_THETA,
  A_INVRHO, A_C11, A_C13, A_C15, A_C33, A_C35, A_C55,
  A_SIN2, A_COS2, A_SIN2T, A_SIN, A_COS, A_SINV, A_COSV,
  A_D, A_BETA, A_COEF1, A_COEF2,
  A_VX, A_VZ, A_SXX, A_SZZ, A_SXZ, A_TXX, A_TZZ, A_TXZ,
  A_COUNT
};

typedef struct {
  int form, nz, nx, nt;
  float dx, dz, dt, eal_alpha, fpeak, scalar;
  int ealThick; /* effective: user value + N (BT:307-308) */
  int freesurface, mode;
  int N, Nz, Nx, Np;
  int ixStart, ixEnd, izStart, izEnd;
  real *a[A_COUNT];
  /* source */
  int stype, ns;
  int *sposx, *sposz;
  real *src; /* [ns][nt] */
  /* receivers */
  int seistype, mt, nr, ntsample;
  int *rposx, *rposz;
  real *rec_vx, *rec_vz, *rec_p; /* [nr][ntsample] */
  /* energy */
  int energy;
  real *E, *Ek, *Ep;
  char err[256];
} octx;

static const float LSC[5] = {1.24362739770907f, -0.114021164233934f, 0.0272965993810630f,
                             -0.00699199493927320f, 0.00132022981661381f}; /* BT:464-465 */

/* index into a padded [Nx][Nz] array; every user declares a local Nz */
#define AT(arr, ix, iz) (arr)[(size_t)(ix) * (size_t)Nz + (size_t)(iz)]

/* ------------------------------------------------------------------ */
octx *PFX(create)(int form, int nz, int nx, int nt, float dx, float dz, float dt, int ealThick,
                  float eal_alpha, int freesurface, float fpeak, float scalar, int mode)
{
  octx *c = (octx *)calloc(1, sizeof(octx));
  c->form = form; c->nz = nz; c->nx = nx; c->nt = nt;
  c->dx = dx; c->dz = dz; c->dt = dt;
  c->eal_alpha = eal_alpha; c->fpeak = fpeak; c->scalar = scalar;
  c->freesurface = freesurface; c->mode = mode;
  c->N = 5;
  c->ealThick = ealThick + c->N;               /* BT:307-308 */
  c->Nz = nz + 2 * c->ealThick;                /* BT:309-311 */
  c->Nx = nx + 2 * c->ealThick;
  c->Np = c->Nz * c->Nx;
  c->ixStart = c->N;                           /* BT:641-644 */
  c->ixEnd = c->Nx - c->N;
  c->izStart = (freesurface == 0 ? c->N : c->ealThick);
  c->izEnd = c->Nz - c->N;
  for (int i = 0; i < A_COUNT; i++) c->a[i] = (real *)calloc((size_t)c->Np, sizeof(real));
  return c;
}

void PFX(destroy)(octx *c)
{
  if (!c) return;
  for (int i = 0; i < A_COUNT; i++) free(c->a[i]);
  free(c->sposx); free(c->sposz); free(c->src);
  free(c->rposx); free(c->rposz); free(c->rec_vx); free(c->rec_vz); free(c->rec_p);
  free(c->E); free(c->Ek); free(c->Ep);
  free(c);
}

const char *PFX(last_error)(octx *c) { return c->err; }
int PFX(dims)(octx *c, int *out)
{
  out[0] = c->Nz; out[1] = c->Nx; out[2] = c->ealThick; out[3] = c->izStart; out[4] = c->izEnd;
  out[5] = c->ixStart; out[6] = c->ixEnd;
  return 0;
}

/* padModel2d, L/absorbc/padmod.c:33-58: top/bottom on interior columns, then
 * left/right on full columns, both by edge replication. */
static void pad_model(octx *c, real *m)
{
  int e = c->ealThick, Nz = c->Nz, Nx = c->Nx;
#pragma omp parallel for
  for (int ix = e; ix < Nx - e; ix++)
    for (int l = 1; l <= e; l++) {
      AT(m, ix, e - l) = AT(m, ix, e);
      AT(m, ix, Nz - e - 1 + l) = AT(m, ix, Nz - e - 1);
    }
#pragma omp parallel for
  for (int l = 1; l <= e; l++)
    for (int iz = 0; iz < Nz; iz++) {
      AT(m, e - l, iz) = AT(m, e, iz);
      AT(m, Nx - e - 1 + l, iz) = AT(m, Nx - e - 1, iz);
    }
}

/* staggered_grid_stable2d, L/fdmod/stable.c:217-242 (called BT:472 on the
 * padded, pre-MEAL vp).  Returns 0 if stable, -1 otherwise. */
static int stability(octx *c, double *dt_max_out)
{
  double vp_max = __FLT_MIN__, C_sum, dt_max;
  real *vp = c->a[A_VP];
  for (int i = 0; i < c->Np; i++) vp_max = fmaxf((float)vp[i], (float)vp_max);
  C_sum = 0.0f;
  for (int i = 0; i < c->N; i += 2) C_sum += LSC[i] * pow(-1.0, 1.0 * i);
  dt_max = 1.0f / (vp_max * sqrt((1.0 / (c->dx * c->dx) + 1.0 / (c->dz * c->dz)) * C_sum));
  *dt_max_out = dt_max;
  return (c->dt > dt_max) ? -1 : 0;
}

/* absorbWeight, L/absorbc/meal.c:23-29 */
static real absorb_weight(real alpha, int l, int nlayer)
{
  real d = (RF(0.1f) * l) / (RF(0.1f) * nlayer);
  return R_EXP(R_LOG(RF(2.0f)) * R_POW(d, alpha)) - RF(1.0f);
}

/* mealCoefs2d, L/absorbc/meal.c:51-201, then mealFDCoefs meal.c:450-491 (order 1). */
static void meal_coefs(octx *c)
{
  int Nz = c->Nz, Nx = c->Nx, nlayer = c->ealThick;
  real *D = c->a[A_D], *Beta = c->a[A_BETA], *vp = c->a[A_VP], *vs = c->a[A_VS];
  real alpha = c->eal_alpha, pd = 2, PPWfd = 10.0f;
  int izStart = c->freesurface == 0 ? 0 : nlayer;
  real Rcoef = R_POW(RF(10.0f), RF(-1.0f) * ((R_LOG10((real)nlayer) - RF(1.0f)) / (R_LOG10(RF(2.0f)))) -
                                    RF(3.0f) - RF(0.5f) * pd);
  real thx = nlayer * (real)c->dx, thz = nlayer * (real)c->dz;
  real dx = c->dx, dz = c->dz, fpeak = c->fpeak;
  size_t Np = (size_t)c->Np;
  real *Vs = (real *)malloc(Np * sizeof(real)), *Dx = (real *)malloc(Np * sizeof(real));
  real *Dz = (real *)malloc(Np * sizeof(real)), *Bx = (real *)malloc(Np * sizeof(real));
  real *Bz = (real *)malloc(Np * sizeof(real));
  if (alpha == RF(-1.0f)) alpha = RF(2.5f);
  for (size_t i = 0; i < Np; i++) {
    Vs[i] = (vs[i] == RF(0.0f)) ? vp[i] / R_SQRT(RF(3.0f)) : vs[i];
    D[i] = 0; Dx[i] = 0; Dz[i] = 0; Beta[i] = 1; Bx[i] = 1; Bz[i] = 1;
  }
  real lnR = R_LOG(Rcoef);
  if (c->freesurface == 0) {
    for (int ix = 0; ix < Nx; ix++)
      for (int l = 1; l <= nlayer; l++) {
        real w = absorb_weight(alpha, l, nlayer);
        int iz = nlayer - l;
        real d0z = RF(-1.0f) * (RF(1.0f) + pd) * AT(vp, ix, iz) * lnR / (RF(2.0f) * thz);
        real ppw = AT(Vs, ix, iz) / (dz * fpeak);
        real betaz = RF(2.0f) * ppw / PPWfd;
        AT(Dz, ix, iz) = d0z * w;
        AT(Bz, ix, iz) = RF(1.0f) + (betaz - RF(1.0f)) * w;
      }
  }
  for (int ix = 0; ix < Nx; ix++)
    for (int l = 1; l <= nlayer; l++) {
      real w = absorb_weight(alpha, l, nlayer);
      int iz = Nz - nlayer - 1 + l;
      real d0z = RF(-1.0f) * (RF(1.0f) + pd) * AT(vp, ix, iz) * lnR / (RF(2.0f) * thz);
      real ppw = AT(Vs, ix, iz) / (dz * fpeak);
      real betaz = RF(2.0f) * ppw / PPWfd;
      AT(Dz, ix, iz) = d0z * w;
      AT(Bz, ix, iz) = RF(1.0f) + (betaz - RF(1.0f)) * w;
    }
  for (int l = 1; l <= nlayer; l++) {
    real w = absorb_weight(alpha, l, nlayer);
    for (int iz = izStart; iz < Nz; iz++) {
      int ix = nlayer - l;
      real d0x = RF(-1.0f) * (RF(1.0f) + pd) * AT(vp, ix, iz) * lnR / (RF(2.0f) * thx);
      real ppw = AT(Vs, ix, iz) / (dx * fpeak);
      real betax = RF(2.0f) * ppw / PPWfd;
      AT(Dx, ix, iz) = d0x * w;
      AT(Bx, ix, iz) = RF(1.0f) + (betax - RF(1.0f)) * w;
      ix = Nx - nlayer - 1 + l;
      d0x = RF(-1.0f) * (RF(1.0f) + pd) * AT(vp, ix, iz) * lnR / (RF(2.0f) * thx);
      ppw = AT(Vs, ix, iz) / (dx * fpeak);
      betax = RF(2.0f) * ppw / PPWfd;
      AT(Dx, ix, iz) = d0x * w;
      AT(Bx, ix, iz) = RF(1.0f) + (betax - RF(1.0f)) * w;
    }
  }
  for (size_t i = 0; i < Np; i++) {
    D[i] = R_SQRT(R_POW(Dx[i], RF(2.0f)) + R_POW(Dz[i], RF(2.0f)));
    Beta[i] = R_SQRT(R_POW(Bx[i] - RF(1.0f), RF(2.0f)) + R_POW(Bz[i] - RF(1.0f), RF(2.0f))) + RF(1.0f);
  }
  free(Vs); free(Dx); free(Dz); free(Bx); free(Bz);
  /* mealFDCoefs, meal.c:450-491: rescale vp, vs, rho and D in place; order-1 coefs in double */
  real *rho = c->a[A_RHO], *c1 = c->a[A_COEF1], *c2 = c->a[A_COEF2];
  double dt = c->dt;
  for (size_t i = 0; i < Np; i++) {
    vp[i] /= Beta[i];
    vs[i] /= Beta[i];
    rho[i] *= Beta[i];
    D[i] /= Beta[i];
    c1[i] = (real)((1.0 - 0.5 * dt * D[i]) / (1.0 + 0.5 * dt * D[i]));
    c2[i] = (real)(dt / (1.0 + 0.5 * dt * D[i]));
  }
}

/* Thomsen -> Cij(theta).  Bond BT:493-526, tensor TR:483-503, coordinate CR:497-514.
 * f, C11..C55 follow the reference's mixed precision: rho*vp*vp is a float
 * product, the rest is double, results stored as float. */
static void cij(octx *c)
{
  real *vp = c->a[A_VP], *vs = c->a[A_VS], *rho = c->a[A_RHO];
  real *eps = c->a[A_EPS], *del = c->a[A_DEL], *th = c->a[A_THETA];
  int form = c->form;
#pragma omp parallel for
  for (int i = 0; i < c->Np; i++) {
    real r = vs[i] / vp[i];
    double f = 1.0 - (double)r * (double)r; /* pow(r, 2.0) is exact: r*r fits in double */
    real rvv = rho[i] * vp[i] * vp[i];
    real vv = vp[i] * vp[i];
    real ss = vs[i] * vs[i];
    double C11 = (double)rvv * (1.0 + 2.0 * (double)eps[i]);
    double C13 = (double)rho[i] * ((double)vv * sqrt(f * (f + 2.0 * (double)del[i])) - (double)ss);
    double C33 = (double)rvv;
    real rss = rho[i] * vs[i] * vs[i];
    double C55 = (double)rss;
    real ths = c->scalar * th[i];                      /* scalar * theta: a float product */
    th[i] = (real)((double)ths / 180.0 * ORACLE_PI); /* BT:508, stored back as float */
    double t = (double)th[i];
    double s = sin(t), co = cos(t);
    if (form == FORM_BOND) {
      double s2 = s * s, c2 = co * co, s4 = s2 * s2, c4 = c2 * c2;
      double sn2 = sin(2.0 * t), cs2 = cos(2.0 * t);
      double sn22 = sn2 * sn2, cs22 = cs2 * cs2;
      c->a[A_C11][i] = (real)(C11 * c4 + C33 * s4 + (0.5 * C13 + C55) * sn22);
      c->a[A_C13][i] = (real)((0.25 * (C11 + C33) - C55) * sn22 + C13 * (s4 + c4));
      c->a[A_C15][i] = (real)(0.5 * (-1.0 * C11 * c2 + C33 * s2 + (C13 + 2.0 * C55) * cs2) * sn2);
      c->a[A_C33][i] = (real)(C11 * s4 + C33 * c4 + (0.5 * C13 + C55) * sn22);
      c->a[A_C35][i] = (real)(0.5 * (-1.0 * C11 * s2 + C33 * c2 - (C13 + 2.0 * C55) * cs2) * sn2);
      c->a[A_C55][i] = (real)(0.25 * (C11 + C33 - 2.0 * C13) * sn22 + C55 * cs22);
    } else {
      c->a[A_C11][i] = (real)C11;
      c->a[A_C13][i] = (real)C13;
      c->a[A_C33][i] = rvv;
      c->a[A_C55][i] = rss;
      if (form == FORM_TENSOR) { /* TR:515-517 */
        c->a[A_SIN2][i] = (real)(s * s);
        c->a[A_COS2][i] = (real)(co * co);
        c->a[A_SIN2T][i] = (real)sin(2.0 * t);
      } else { /* CR:512-513 */
        c->a[A_SIN][i] = (real)s;
        c->a[A_COS][i] = (real)co;
      }
    }
  }
}

/* invrho BT:527-542; coordinate form trig averages CR:516-535. */
static void stagger_params(octx *c)
{
  real *rho = c->a[A_RHO], *ir = c->a[A_INVRHO];
  real *sn = c->a[A_SIN], *cs = c->a[A_COS], *snv = c->a[A_SINV], *csv = c->a[A_COSV];
  int Nx = c->Nx, Nz = c->Nz, coord = (c->form == FORM_COORD);
#pragma omp parallel for
  for (int ix = 0; ix < Nx; ix++)
    for (int iz = 0; iz < Nz; iz++) {
      if (ix < Nx - 1 && iz < Nz - 1) {
        if (coord) {
          AT(snv, ix, iz) = (real)(0.25 * (AT(sn, ix, iz) + AT(sn, ix + 1, iz) + AT(sn, ix, iz + 1) + AT(sn, ix + 1, iz + 1)));
          AT(csv, ix, iz) = (real)(0.25 * (AT(cs, ix, iz) + AT(cs, ix + 1, iz) + AT(cs, ix, iz + 1) + AT(cs, ix + 1, iz + 1)));
        }
        AT(ir, ix, iz) = (real)(4.0 / (AT(rho, ix, iz) + AT(rho, ix + 1, iz) + AT(rho, ix, iz + 1) + AT(rho, ix + 1, iz + 1)));
      } else {
        if (coord) {
          AT(snv, ix, iz) = AT(sn, ix, iz);
          AT(csv, ix, iz) = AT(cs, ix, iz);
        }
        AT(ir, ix, iz) = (real)(1.0 / AT(rho, ix, iz));
      }
    }
}

/* Bond free-surface parameter fix-up, BT:543-568 (float arithmetic, including
 * the reference's use of the updated c35 in c55 and its c11 sign). */
static void fs_fixup(octx *c)
{
  const int Nz = c->Nz;
  int iz = c->ealThick;
  real *c11 = c->a[A_C11], *c13 = c->a[A_C13], *c15 = c->a[A_C15];
  real *c33 = c->a[A_C33], *c35 = c->a[A_C35], *c55 = c->a[A_C55];
  for (int ix = 0; ix < c->Nx; ix++) {
    AT(c35, ix, iz) = -(AT(c13, ix, iz) * AT(c55, ix, iz) - AT(c15, ix, iz) * AT(c35, ix, iz)) /
                      (AT(c33, ix, iz) * AT(c55, ix, iz) - AT(c35, ix, iz) * AT(c35, ix, iz));
    AT(c55, ix, iz) = (AT(c13, ix, iz) * AT(c35, ix, iz) - AT(c15, ix, iz) * AT(c33, ix, iz)) /
                      (AT(c33, ix, iz) * AT(c55, ix, iz) - AT(c35, ix, iz) * AT(c35, ix, iz));
    if (isnan(AT(c35, ix, iz)) || isinf(AT(c35, ix, iz))) AT(c35, ix, iz) = -AT(c13, ix, iz) / AT(c33, ix, iz);
    if (isnan(AT(c35, ix, iz)) || isinf(AT(c35, ix, iz))) AT(c35, ix, iz) = RF(0.0f);
    AT(c11, ix, iz) = AT(c11, ix, iz) - AT(c13, ix, iz) * AT(c35, ix, iz) + AT(c15, ix, iz) * AT(c55, ix, iz);
    AT(c13, ix, iz) = 0; AT(c15, ix, iz) = 0; AT(c33, ix, iz) = 0; AT(c35, ix, iz) = 0; AT(c55, ix, iz) = 0;
  }
}

/* Model input: six unpadded nz*nx float arrays in file order (nx records of nz).
 * vs and rho must be given (the host applies the default conversions).  The
 * sequence is the reference's: read into the padded interior (BT:345-450), pad
 * (BT:452-457), stability (BT:472), MEAL (BT:474-475), Cij (BT:493-526),
 * invrho (BT:527-542), Bond fix-up (BT:543-568). */
int PFX(set_model)(octx *c, const float *vp, const float *vs, const float *rho, const float *eps,
                   const float *del, const float *theta, double *dt_max)
{
  const float *in[6] = {vp, vs, rho, eps, del, theta};
  const int e = c->ealThick, Nz = c->Nz;
  for (int k = 0; k < 6; k++) {
    real *m = c->a[A_VP + k];
    memset(m, 0, (size_t)c->Np * sizeof(real));
    for (int ix = 0; ix < c->nx; ix++)
      for (int iz = 0; iz < c->nz; iz++) AT(m, ix + e, iz + e) = in[k][(size_t)ix * c->nz + iz];
    pad_model(c, m);
  }
  if (stability(c, dt_max) != 0) {
    snprintf(c->err, sizeof c->err, "unstable: dt=%g > dt_max=%g", c->dt, *dt_max);
    return -2;
  }
  meal_coefs(c);
  cij(c);
  stagger_params(c);
  if (c->form == FORM_BOND && c->freesurface == 1) fs_fixup(c);
  for (int k = A_VX; k <= A_TXZ; k++) memset(c->a[k], 0, (size_t)c->Np * sizeof(real));
  return 0;
}

int PFX(set_source)(octx *c, int stype, int ns, const int *posx, const int *posz, const float *src)
{
  c->stype = stype; c->ns = ns;
  free(c->sposx); free(c->sposz); free(c->src);
  c->sposx = (int *)malloc(sizeof(int) * ns);
  c->sposz = (int *)malloc(sizeof(int) * ns);
  c->src = (real *)malloc(sizeof(real) * (size_t)ns * c->nt);
  for (int i = 0; i < ns; i++) { c->sposx[i] = posx[i]; c->sposz[i] = posz[i]; }
  for (size_t i = 0; i < (size_t)ns * c->nt; i++) c->src[i] = src[i];
  return 0;
}

int PFX(set_receivers)(octx *c, int seistype, int mt, int nr, const int *posx, const int *posz)
{
  c->seistype = seistype; c->mt = mt; c->nr = nr;
  c->ntsample = (c->nt - 1) / mt + 1; /* L/fdmod/obsinfo.c:199 */
  free(c->rposx); free(c->rposz); free(c->rec_vx); free(c->rec_vz); free(c->rec_p);
  c->rposx = (int *)malloc(sizeof(int) * (nr > 0 ? nr : 1));
  c->rposz = (int *)malloc(sizeof(int) * (nr > 0 ? nr : 1));
  for (int i = 0; i < nr; i++) { c->rposx[i] = posx[i]; c->rposz[i] = posz[i]; }
  size_t n = (size_t)(nr > 0 ? nr : 1) * c->ntsample;
  c->rec_vx = (real *)calloc(n, sizeof(real));
  c->rec_vz = (real *)calloc(n, sizeof(real));
  c->rec_p = (real *)calloc(n, sizeof(real));
  return 0;
}

int PFX(set_energy)(octx *c, int on)
{
  c->energy = on;
  free(c->E); free(c->Ek); free(c->Ep);
  c->E = (real *)calloc((size_t)c->nt, sizeof(real));
  c->Ek = (real *)calloc((size_t)c->nt, sizeof(real));
  c->Ep = (real *)calloc((size_t)c->nt, sizeof(real));
  return 0;
}

/* ---------------- time loop pieces ---------------- */

#define SWEEP_FORM FORM_COORD
#define SWEEP_NAME(n) n##_coord
#include "fdmod_oracle_sweeps.inc"
#undef SWEEP_FORM
#undef SWEEP_NAME
#define SWEEP_FORM FORM_TENSOR
#define SWEEP_NAME(n) n##_tensor
#include "fdmod_oracle_sweeps.inc"
#undef SWEEP_FORM
#undef SWEEP_NAME
#define SWEEP_FORM FORM_BOND
#define SWEEP_NAME(n) n##_bond
#include "fdmod_oracle_sweeps.inc"
#undef SWEEP_FORM
#undef SWEEP_NAME

static void v_sweep(octx *c)
{
  if (c->form == FORM_COORD) v_sweep_coord(c);
  else if (c->form == FORM_TENSOR) v_sweep_tensor(c);
  else v_sweep_bond(c);
}
static void s_sweep(octx *c)
{
  if (c->form == FORM_COORD) s_sweep_coord(c);
  else if (c->form == FORM_TENSOR) s_sweep_tensor(c);
  else s_sweep_bond(c);
}

/* Free surface by stress imaging, BT:875-891 (TR:853-870, CR:846-862). */
static void free_surface(octx *c)
{
  const int Nz = c->Nz;
  int z0 = c->izStart, N = c->N;
  real *sxx = c->a[A_SXX], *szz = c->a[A_SZZ], *sxz = c->a[A_SXZ];
  for (int ix = c->ixStart; ix < c->ixEnd; ix++) {
    AT(sxz, ix, z0) = 0;
    AT(szz, ix, z0) = 0;
    for (int k = 1; k <= N; k++) {
      AT(sxx, ix, z0 - k) = RF(2.0f) * AT(sxx, ix, z0) - AT(sxx, ix, z0 + k);
      AT(szz, ix, z0 - k) = -AT(szz, ix, z0 + k);
      AT(sxz, ix, z0 - k) = -AT(sxz, ix, z0 + k);
    }
  }
}

/* Source footprints, BT:706-834 (velocity) and BT:893-914 (stress). */
static void sources(octx *c, int it, int on_stress)
{
  const int N = c->N, e = c->ealThick, Nx = c->Nx, Nz = c->Nz, st = c->stype;
  real dx = c->dx, dz = c->dz, dt = c->dt;
  real invdx = (real)(1.0 / c->dx), invdz = (real)(1.0 / c->dz);
  real alpha = R_LOG(RF(0.8f)) / (dx * dx + dz * dz); /* BT:647 */
  real *vx = c->a[A_VX], *vz = c->a[A_VZ];
  if (on_stress) {
    if (st != 1) return;
    for (int is = 0; is < c->ns; is++) {
      int isx = c->sposx[is] + e, isz = c->sposz[is] + e;
      real s = c->src[(size_t)is * c->nt + it];
      for (int ix = (isx - N > 0 ? isx - N : 0); ix <= (isx + N < Nx - 1 ? isx + N : Nx - 1); ix++) {
        real rx = (ix - isx) * dx;
        for (int iz = (isz - N > 0 ? isz - N : 0); iz <= (isz + N < Nz - 1 ? isz + N : Nz - 1); iz++) {
          real rz = (iz - isz) * dz;
          real w = R_EXP(alpha * (rx * rx + rz * rz)) * invdx * invdz * dt;
          AT(c->a[A_SXX], ix, iz) += RF(0.5f) * w * s;
          AT(c->a[A_SZZ], ix, iz) += RF(0.5f) * w * s;
          if (c->form == FORM_TENSOR) { /* TR:888-891 */
            AT(c->a[A_TXX], ix, iz) += RF(0.5f) * w * s;
            AT(c->a[A_TZZ], ix, iz) += RF(0.5f) * w * s;
          }
        }
      }
    }
    return;
  }
  /* velocity sources: stype 2 and 4 (divergence), 3 and 4 (curl), 5, 6, 7, 8 */
  for (int pass = 0; pass < 2; pass++) {
    int kind;
    if (pass == 0) kind = (st == 2 || st == 4) ? 2 : (st >= 5 && st <= 7) ? st : 0;
    else kind = (st == 3 || st == 4) ? 3 : 0;
    if (kind == 0) continue;
    for (int is = 0; is < c->ns; is++) {
      int isx = c->sposx[is] + e, isz = c->sposz[is] + e;
      real s = c->src[(size_t)is * c->nt + it];
      for (int ix = (isx - N > 0 ? isx - N : 0); ix <= (isx + N < Nx - 1 ? isx + N : Nx - 1); ix++) {
        real rx = (ix - isx) * dx;
        for (int iz = (isz - N > 0 ? isz - N : 0); iz <= (isz + N < Nz - 1 ? isz + N : Nz - 1); iz++) {
          real rz = (iz - isz) * dz;
          real w = R_EXP(alpha * (rx * rx + rz * rz)) * invdx * invdz * dt;
          if (kind == 2) {
            AT(vx, ix, iz) += RF(2.0f) * alpha * rx * w * s;
            AT(vz, ix, iz) += RF(2.0f) * alpha * rz * w * s;
          } else if (kind == 3) {
            AT(vx, ix, iz) -= RF(2.0f) * alpha * rz * w * s;
            AT(vz, ix, iz) += RF(2.0f) * alpha * rx * w * s;
          } else if (kind == 5) {
            AT(vx, ix, iz) += w * s;
          } else if (kind == 6) {
            AT(vz, ix, iz) += w * s;
          } else { /* 7 */
            AT(vx, ix, iz) += RF(0.5f) * R_SQRT(RF(2.0f)) * w * s;
            AT(vz, ix, iz) += RF(0.5f) * R_SQRT(RF(2.0f)) * w * s;
          }
        }
      }
    }
  }
  if (st == 8) { /* BT:814-833, sequential over sources */
    real q = RF(0.125f) * R_SQRT(RF(2.0f));
    for (int is = 0; is < c->ns; is++) {
      int isx = c->sposx[is] + e, isz = c->sposz[is] + e;
      real s = c->src[(size_t)is * c->nt + it];
      AT(vx, isx, isz) -= q * s;
      AT(vz, isx, isz) += q * s;
      AT(vx, isx - 1, isz) -= q * s;
      AT(vz, isx - 1, isz) -= q * s;
      AT(vx, isx - 1, isz - 1) += q * s;
      AT(vz, isx - 1, isz - 1) -= q * s;
      AT(vx, isx, isz - 1) += q * s;
      AT(vz, isx, isz - 1) += q * s;
    }
  }
}

/* elastic2dReceiver, L/fdmod/isoelastic.c:103-144 */
static void receivers(octx *c, int it)
{
  const int Nz = c->Nz;
  if (c->seistype <= 0 || c->nr <= 0 || it % c->mt != 0) return;
  int ir = it / c->mt, e = c->ealThick;
  for (int r = 0; r < c->nr; r++) {
    int gx = c->rposx[r] + e, gz = c->rposz[r] + e;
    size_t o = (size_t)r * c->ntsample + ir;
    if (c->seistype == 1 || c->seistype == 3) {
      c->rec_vx[o] = AT(c->a[A_VX], gx, gz);
      c->rec_vz[o] = AT(c->a[A_VZ], gx, gz);
    }
    if (c->seistype == 2 || c->seistype == 3)
      c->rec_p[o] = RF(0.5f) * (AT(c->a[A_SXX], gx, gz) + AT(c->a[A_SZZ], gx, gz));
  }
}

/* Energy diagnostic, BT:926-952 (float reduction; summation order differs
 * from an OpenMP reduction, so it is compared with a tolerance). */
static void energy(octx *c, int it)
{
  const int Nz = c->Nz;
  real ek = 0, ep = 0;
  int e = c->ealThick;
  real *vx = c->a[A_VX], *vz = c->a[A_VZ], *ir = c->a[A_INVRHO];
  real *c11 = c->a[A_C11], *c55 = c->a[A_C55];
  real *sxx = c->a[A_SXX], *szz = c->a[A_SZZ], *sxz = c->a[A_SXZ];
  for (int ix = 0; ix < c->nx; ix++)
    for (int iz = 0; iz < c->nz; iz++) {
      int X = ix + e, Z = iz + e;
      real a11 = AT(c11, X, Z), a55 = AT(c55, X, Z), s1 = AT(sxx, X, Z), s2 = AT(szz, X, Z), s3 = AT(sxz, X, Z);
      ek += R_POW(AT(vx, X, Z), RF(2.0f)) / AT(ir, X, Z) + R_POW(AT(vz, X, Z), RF(2.0f)) / AT(ir, X, Z);
      ep += RF(0.25f) * (((a11 + RF(2.0f) * a55) * s1 - a11 * s2) * s1 + (a11 * s1 - (a11 + RF(2.0f) * a55) * s2) * s2) /
                (a55 * (a11 + a55)) +
            s3 * s3 / a55;
    }
  c->Ek[it] = RF(0.5f) * ek * c->dx * c->dz;
  c->Ep[it] = RF(0.5f) * ep * c->dx * c->dz;
  c->E[it] = c->Ek[it] + c->Ep[it];
}

/* One time step in the reference's order (A21, BT:651-954). */
static void step(octx *c, int it)
{
  v_sweep(c);
  if (c->stype != 1 && c->ns > 0) sources(c, it, 0);
  s_sweep(c);
  if (c->freesurface == 1) free_surface(c);
  if (c->stype == 1 && c->ns > 0) sources(c, it, 1);
  receivers(c, it);
  if (c->energy) energy(c, it);
}

/* Progress lines on stderr every `every` steps when ORACLE_PROGRESS is set (long
 * FP64 runs that would otherwise print nothing for minutes). */
static void progress(const char *what, int k, int nt)
{
  static int every = -1;
  if (every < 0) {
    const char *env = getenv("ORACLE_PROGRESS");
    every = env ? atoi(env) : 0;
  }
  if (every > 0 && k % every == 0) {
    fprintf(stderr, "oracle: %s step %d of %d\n", what, k, nt);
    fflush(stderr);
  }
}

/* Run loop iterations k in [k0, k0+count); it = k (mode 1) or nt-1-k (mode 2),
 * BT:651.  Returns the loop wall time in seconds (omp_get_wtime, BT:650,955). */
double PFX(run)(octx *c, int k0, int count)
{
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#endif
  for (int k = k0; k < k0 + count && k < c->nt; k++) {
    progress("forward", k, c->nt);
    step(c, c->mode == 1 ? k : c->nt - 1 - k);
  }
#ifdef _OPENMP
  return omp_get_wtime() - t0;
#else
  return 0.0;
#endif
}

/* Individual phases, for operator tests. */
void PFX(v_sweep)(octx *c) { v_sweep(c); }
void PFX(s_sweep)(octx *c) { s_sweep(c); }
void PFX(free_surface)(octx *c) { free_surface(c); }
void PFX(sources)(octx *c, int it, int on_stress) { sources(c, it, on_stress); }

int PFX(get_array)(octx *c, int which, double *out)
{
  if (which < 0 || which >= A_COUNT) return -1;
  for (int i = 0; i < c->Np; i++) out[i] = (double)c->a[which][i];
  return 0;
}
int PFX(set_array)(octx *c, int which, const double *in)
{
  if (which < 0 || which >= A_COUNT) return -1;
  for (int i = 0; i < c->Np; i++) c->a[which][i] = (real)in[i];
  return 0;
}
int PFX(get_traces)(octx *c, int comp, double *out)
{
  real *r = comp == 1 ? c->rec_vx : comp == 3 ? c->rec_vz : c->rec_p;
  if (!r) return -1;
  for (size_t i = 0; i < (size_t)c->nr * c->ntsample; i++) out[i] = (double)r[i];
  return 0;
}
int PFX(get_energy)(octx *c, double *E, double *Ek, double *Ep)
{
  if (!c->E) return -1;
  for (int i = 0; i < c->nt; i++) { E[i] = c->E[i]; Ek[i] = c->Ek[i]; Ep[i] = c->Ep[i]; }
  return 0;
}
int PFX(set_threads)(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ---------------- adjoint and gradient ---------------- */
#include "fdmod_oracle_adj.inc"
