/* suio.c -- Seismic Unix trace files (see suio.h). */
#include "suio.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* byte offsets in the 240-byte trace header (SEG-Y rev0 + SU extension) */
enum {
  O_TRACL = 0, O_TRACR = 4, O_FLDR = 8, O_CDP = 20, O_TRID = 28, O_OFFSET = 36, O_SDEPTH = 48, O_GDEL = 52,
  O_SDEL = 56, O_SCALEL = 68, O_SCALCO = 70, O_SX = 72, O_GX = 80, O_NS = 114, O_DT = 116, O_D1 = 180,
  O_F1 = 184, O_NTR = 204
};

#define GET(T, off) (*(const T *)(h + (off)))
#define PUT(T, off, v) (*(T *)(h + (off)) = (T)(v))

void su_init(su_trace *tr) { memset(tr, 0, sizeof *tr); }

void su_free(su_trace *tr)
{
  free(tr->data);
  tr->data = NULL;
  tr->cap = 0;
}

void su_reserve(su_trace *tr, int ns)
{
  if (ns > tr->cap) {
    tr->data = (float *)realloc(tr->data, sizeof(float) * (size_t)(ns > 0 ? ns : 1));
    tr->cap = ns;
  }
}

int su_read(FILE *fp, su_trace *tr)
{
  unsigned char h[SU_HDRBYTES];
  size_t got = fread(h, 1, SU_HDRBYTES, fp);
  if (got == 0) return 0;
  if (got != SU_HDRBYTES) return -1;
  tr->tracl = GET(int32_t, O_TRACL);
  tr->tracr = GET(int32_t, O_TRACR);
  tr->fldr = GET(int32_t, O_FLDR);
  tr->cdp = GET(int32_t, O_CDP);
  tr->trid = GET(int16_t, O_TRID);
  tr->offset = GET(int32_t, O_OFFSET);
  tr->sdepth = GET(int32_t, O_SDEPTH);
  tr->gdel = GET(int32_t, O_GDEL);
  tr->sdel = GET(int32_t, O_SDEL);
  tr->scalel = GET(int16_t, O_SCALEL);
  tr->scalco = GET(int16_t, O_SCALCO);
  tr->sx = GET(int32_t, O_SX);
  tr->gx = GET(int32_t, O_GX);
  tr->ns = GET(uint16_t, O_NS);
  tr->dt = GET(uint16_t, O_DT);
  tr->d1 = GET(float, O_D1);
  tr->f1 = GET(float, O_F1);
  tr->ntr = GET(int32_t, O_NTR);
  su_reserve(tr, tr->ns);
  if (tr->ns && fread(tr->data, sizeof(float), tr->ns, fp) != tr->ns) return -1;
  return 1;
}

int su_write(FILE *fp, const su_trace *tr)
{
  unsigned char h[SU_HDRBYTES];
  memset(h, 0, sizeof h);
  PUT(int32_t, O_TRACL, tr->tracl);
  PUT(int32_t, O_TRACR, tr->tracr);
  PUT(int32_t, O_FLDR, tr->fldr);
  PUT(int32_t, O_CDP, tr->cdp);
  PUT(int16_t, O_TRID, tr->trid);
  PUT(int32_t, O_OFFSET, tr->offset);
  PUT(int32_t, O_SDEPTH, tr->sdepth);
  PUT(int32_t, O_GDEL, tr->gdel);
  PUT(int32_t, O_SDEL, tr->sdel);
  PUT(int16_t, O_SCALEL, tr->scalel);
  PUT(int16_t, O_SCALCO, tr->scalco);
  PUT(int32_t, O_SX, tr->sx);
  PUT(int32_t, O_GX, tr->gx);
  PUT(uint16_t, O_NS, tr->ns);
  PUT(uint16_t, O_DT, tr->dt);
  PUT(float, O_D1, tr->d1);
  PUT(float, O_F1, tr->f1);
  PUT(int32_t, O_NTR, tr->ntr);
  if (fwrite(h, 1, SU_HDRBYTES, fp) != SU_HDRBYTES) return -1;
  if (tr->ns && fwrite(tr->data, sizeof(float), tr->ns, fp) != tr->ns) return -1;
  return 0;
}

int head_scalar_read(float key, int scalar)
{
  if (scalar == 0) return (int)key;
  int temp = (int)((scalar < 0 ? -1.0 : 1.0) * log10((double)abs(scalar)));
  return (int)(key * pow(10.0, temp));
}

int head_scalar_write(float key, int scalar) { return head_scalar_read(key, -scalar); }

/* 8-point least-squares optimal interpolation, the method of Seismic Unix's ints8r
 * (version unpinned; called by read2dsource, src/lib/fdmod/obsinfo.c:82, with zero
 * outside the trace).  For a fractional offset d in (0, 1) the 8 coefficients w_j
 * (taps at samples i-3 .. i+4 around x = i + d) minimise the mean square error of
 * the interpolation over signals band-limited to fmax of the Nyquist frequency, with
 * fmax = 0.066 + 0.265 ln 8: the normal equations are the symmetric Toeplitz system
 * sum_k R(j-k) w_k = R(3 - j + d), R(x) = sinc(fmax x).  Offsets are tabulated at
 * 1/512 of a sample and the nearest entry is used; offsets 0 and 1 take the sample. */
#define LS8_NTAB 513

static double sincpi(double x) { return x == 0.0 ? 1.0 : sin(M_PI * x) / (M_PI * x); }

/* Levinson recursion for the symmetric Toeplitz system T w = c, T[j][k] = r[|j-k|]. */
static void toeplitz_solve(int n, const double *r, const double *c, double *w)
{
  double a[8], tmp[8];
  double e = r[0];
  a[0] = 1.0;
  w[0] = c[0] / r[0];
  for (int m = 1; m < n; m++) {
    double k = 0.0;  /* reflection coefficient of the prediction-error filter a */
    for (int j = 0; j < m; j++) k -= a[j] * r[m - j];
    k /= e;
    for (int j = 0; j <= m; j++) tmp[j] = (j < m ? a[j] : 0.0) + k * (m - j < m ? a[m - j] : 0.0);
    for (int j = 0; j <= m; j++) a[j] = tmp[j];
    e *= 1.0 - k * k;
    double q = c[m];  /* extend the solution with the backward filter */
    for (int j = 0; j < m; j++) q -= w[j] * r[m - j];
    q /= e;
    w[m] = 0.0;
    for (int j = 0; j <= m; j++) w[j] += q * a[m - j];
  }
}

static float ls8_table[LS8_NTAB][8];
static int ls8_ready = 0;

static void ls8_build(void)
{
  const double fmax = fmin(0.066 + 0.265 * log(8.0), 1.0);
  double r[8], c[8], w[8];
  for (int j = 0; j < 8; j++) r[j] = sincpi(fmax * j);
  for (int t = 0; t < LS8_NTAB; t++) {
    for (int j = 0; j < 8; j++) ls8_table[t][j] = 0.0f;
    if (t == 0) { ls8_table[t][3] = 1.0f; continue; }
    if (t == LS8_NTAB - 1) { ls8_table[t][4] = 1.0f; continue; }
    double d = (double)t / (LS8_NTAB - 1);
    for (int j = 0; j < 8; j++) c[j] = sincpi(fmax * (3.0 - j + d));
    toeplitz_solve(8, r, c, w);
    for (int j = 0; j < 8; j++) ls8_table[t][j] = (float)w[j];
  }
  ls8_ready = 1;
}

void resample8(int nin, float din, float fin, const float *yin, int nout, const float *tout, float *yout)
{
  if (!ls8_ready) ls8_build();
  for (int i = 0; i < nout; i++) {
    double xi = ((double)tout[i] - fin) / din;
    double fl = floor(xi);
    long k0 = (long)fl;
    int t = (int)((xi - fl) * (LS8_NTAB - 1) + 0.5);
    const float *w = ls8_table[t];
    double acc = 0.0;
    for (int j = 0; j < 8; j++) {
      long k = k0 - 3 + j;
      if (k >= 0 && k < nin) acc += (double)w[j] * yin[k];
    }
    yout[i] = (float)acc;
  }
}
