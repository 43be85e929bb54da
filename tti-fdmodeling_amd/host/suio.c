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

/* Kaiser window (beta = 6) over the 8-point support |x| < 4 */
static double kaiser8(double x)
{
  const double beta = 6.0;
  double r = x / 4.0;
  if (fabs(r) >= 1.0) return 0.0;
  double a = beta * sqrt(1.0 - r * r), s = 1.0, t = 1.0, b = 1.0;
  /* I0 by its power series, for the argument and for beta */
  for (int k = 1; k < 30; k++) {
    t *= (a / (2.0 * k)) * (a / (2.0 * k));
    s += t;
  }
  double tb = 1.0;
  for (int k = 1; k < 30; k++) {
    tb *= (beta / (2.0 * k)) * (beta / (2.0 * k));
    b += tb;
  }
  return s / b;
}

void resample8(int nin, float din, float fin, const float *yin, int nout, const float *tout, float *yout)
{
  for (int i = 0; i < nout; i++) {
    double xi = ((double)tout[i] - fin) / din;
    double r = floor(xi + 0.5);
    if (fabs(xi - r) < 1.0 / 1024.0) {
      long k = (long)r;
      yout[i] = (k >= 0 && k < nin) ? yin[k] : 0.0f;
      continue;
    }
    long k0 = (long)floor(xi);
    double acc = 0.0;
    for (long k = k0 - 3; k <= k0 + 4; k++) {
      if (k < 0 || k >= nin) continue;
      double x = xi - k;
      double sinc = x == 0.0 ? 1.0 : sin(M_PI * x) / (M_PI * x);
      acc += yin[k] * sinc * kaiser8(x);
    }
    yout[i] = (float)acc;
  }
}
