/*
 * suio.h -- Seismic Unix trace files for the drop-in programs: the 240-byte SEG-Y
 * rev0 trace header with the SU extension fields, followed by ns native-endian
 * float32 samples.  Only the header fields the modelling programs read or write
 * are exposed; the rest of the header is written as zeros.
 */
#ifndef SUIO_H
#define SUIO_H
#include <stdio.h>

#define SU_HDRBYTES 240

typedef struct {
  int tracl, tracr, fldr, cdp;
  short trid;
  int offset, sdepth, gdel, sdel;
  short scalel, scalco;
  int sx, gx;
  unsigned short ns, dt; /* dt in microseconds */
  float d1, f1;
  int ntr;
  float *data;           /* ns samples (owned) */
  int cap;
} su_trace;

void su_init(su_trace *tr);
void su_free(su_trace *tr);
/* Read one trace; returns 1 on success, 0 at end of file, -1 on a truncated trace. */
int su_read(FILE *fp, su_trace *tr);
int su_write(FILE *fp, const su_trace *tr);
void su_reserve(su_trace *tr, int ns);

/* Coordinate scalars: value = key * 10^(sgn(scalar) * floor(log10|scalar|)),
 * truncated to int (head_scalar_read, src/lib/supro/head.c:37-51); the write form
 * applies the opposite scalar (head.c:60). */
int head_scalar_read(float key, int scalar);
int head_scalar_write(float key, int scalar);

/* Resample a uniformly sampled trace (nin samples, interval din, first sample fin)
 * at the nout times tout[], with zero outside the trace: the 8-point least-squares
 * optimal interpolator of Seismic Unix's ints8r (read2dsource, obsinfo.c:82), with
 * offsets tabulated at 1/512 of a sample.  Times on a sample (to 1/1024 of one) take
 * that sample exactly, so equal intervals copy the trace. */
void resample8(int nin, float din, float fin, const float *yin, int nout, const float *tout, float *yout);

#endif
