/*
 * sumksrc2d -- write a one-trace SU source file (wavelet + position in cells), the
 * input of the modelling programs' sourcefile.  Same parameters and header fields
 * as the reference utility (src/main/wavelet/sumksrc2d.c:52-160); the wavelets
 * follow src/lib/wavelet/wavelet.c (Gaussian :21-37, first derivative of Gaussian
 * :51-67, Ricker :81-99, derivative of Ricker :113-130): t = it*dt - delay, zero once
 * t > delay, and the first sample forced to zero for every type (wavelet.c:37, 67, 99,
 * 130).  The diagnostics for missing or defaulted parameters are the reference's lines
 * (sumksrc2d.c:61-104).
 *   sumksrc2d nt= dt= sdepth= sx= [wtype=1 A=1 fpeak=10 delay=1.5/fpeak wfile=] > src.su
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parlite.h"
#include "suio.h"

static const char *sdoc[] = {
    "",
    " sumksrc2d - one-trace SU source file",
    " sumksrc2d nt= dt= sdepth= sx= [wtype=1 A=1 fpeak=10 delay=1.5/fpeak wfile=] > stdout",
    "   sdepth, sx   source position in cells (depth, distance)",
    "   wtype        1 Ricker, 2 first derivative of Gaussian, 3 derivative of Ricker, 4 Gaussian",
    "   wfile        read nt float32 samples from this file instead",
    0};

static void wavelet(float *w, int wtype, float A, float fpeak, float delay, float dt, int nt)
{
  const float pi = 3.141592654f;
  for (int it = 0; it < nt; it++) {
    float t = it * dt - delay;
    float a = powf(pi * fpeak * t, 2.0f);
    float v = 0.0f;
    if (t <= delay) {
      switch (wtype) {
      case 1: v = A * (1.0f - 2.0f * a) * expf(-1.0f * a); break;
      case 2: v = A * t * expf(-a); break;
      case 3: v = A * (2.0f * powf(pi * fpeak, 2.0f) * t) * (2.0f * a - 3.0f) * expf(-1.0f * a); break;
      case 4: v = 0.5f / powf(pi * fpeak, 2.0f) * expf(-a); break;
      }
    }
    w[it] = v;
  }
  if (nt > 0) w[0] = 0.0f;
}

int main(int argc, char **argv)
{
  initargs(argc, argv);
  set_selfdoc(sdoc);
  requestdoc(0);
  int nt, wtype, sdepth, sx;
  float dt, A, fpeak, delay;
  char *wfile = NULL;
  if (!getparstring("wfile", &wfile)) {
    wfile = NULL;
    warn("<warning>: didn't specify a wavelet file! The program will use one of wavelet in the document!");
  }
  if (!getparint("nt", &nt)) err("<error>: must specify number of samples in time dimension!");
  if (!getparint("wtype", &wtype)) {
    wtype = 1;
    warn("<warning>: didn't specify type of wavelet! The program will set it %d!", wtype);
  }
  if (!getparint("sdepth", &sdepth)) err("<error>: must specify source position in 2nd dimension!");
  if (!getparint("sx", &sx)) err("<error>: must specify source position in 1st dimension!");
  if (!getparfloat("dt", &dt)) err("<error>: must specify sampleing interval in time dimension!");
  if (!getparfloat("A", &A)) {
    A = 1.0f;
    warn("<warning>: didn't specify peak amplitude of wavelet! The program will set it as %10.5f.", A);
  }
  if (!getparfloat("fpeak", &fpeak)) {
    fpeak = 10.0f;
    warn("<warning>: didn't specify fpeak frequency of wavelet! The program will set it as %10.5f Hz.", fpeak);
  }
  if (!getparfloat("delay", &delay)) {
    delay = 1.5f / fpeak;
    warn("<warning>: didn't specify delay time of wavelet! The program will set delay = 1.5 / fpeak =%10.5f Hz!",
         delay);
  }
  if (nt < 1 || nt > 65535) err("<error>: nt must be in 1..65535 (SU ns field)!");
  su_trace tr;
  su_init(&tr);
  su_reserve(&tr, nt);
  if (wfile) {
    FILE *fp = fopen(wfile, "rb");
    if (!fp) err("<error>: failed to open wavelet file (%s)!", wfile);
    if (fread(tr.data, sizeof(float), nt, fp) != (size_t)nt) err("<error>: short wavelet file (%s)!", wfile);
    fclose(fp);
  } else {
    wavelet(tr.data, wtype, A, fpeak, delay, dt, nt);
  }
  tr.tracl = 1;
  tr.trid = 1;
  tr.scalel = -1000;
  tr.scalco = -1000;
  tr.sdepth = head_scalar_write((float)sdepth, tr.scalco);
  tr.sx = head_scalar_write((float)sx, tr.scalco);
  tr.ns = (unsigned short)nt;
  tr.dt = (unsigned short)(dt * 1000000);
  tr.f1 = 0.0f;
  tr.d1 = dt;
  tr.ntr = 1;
  if (su_write(stdout, &tr) != 0) err("<error>: write failed!");
  su_free(&tr);
  return 0;
}
