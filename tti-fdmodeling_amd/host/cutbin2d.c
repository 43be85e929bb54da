/*
 * cutbin2d -- crop a 2-D float32 binary read from stdin (n2 records of n1 samples)
 * to nx2 records of nx1 samples starting at (fx1, fx2), written to stdout; the
 * aperture-crop step of the demo pipeline (src/main/wavelet/cutbin2d.c:30-94).
 *   cutbin2d < in.bin n1= n2= [fx1=0 fx2=0 nx1=n1 nx2=n2] > out.bin
 */
#include <stdio.h>
#include <stdlib.h>

#include "parlite.h"

static const char *sdoc[] = {"", " cutbin2d < in n1= n2= [fx1=0 fx2=0 nx1=n1 nx2=n2] > out", 0};

int main(int argc, char **argv)
{
  initargs(argc, argv);
  set_selfdoc(sdoc);
  requestdoc(0);
  int n1, n2, fx1, fx2, nx1, nx2;
  if (!getparint("n1", &n1)) err("<error>: must specify number of samples in 1st (fast) dimension!");
  if (!getparint("n2", &n2)) err("<error>: must specify number of samples in 2nd (slow) dimension!");
  if (!getparint("fx1", &fx1)) fx1 = 0;
  if (!getparint("fx2", &fx2)) fx2 = 0;
  if (!getparint("nx1", &nx1)) nx1 = n1;
  if (!getparint("nx2", &nx2)) nx2 = n2;
  if (fx1 < 0 || fx2 < 0 || nx1 < 1 || nx2 < 1 || fx1 + nx1 > n1 || fx2 + nx2 > n2)
    err("<error>: the window (%d+%d, %d+%d) does not fit the %d x %d input!", fx1, nx1, fx2, nx2, n1, n2);
  float *rec = (float *)malloc(sizeof(float) * (size_t)n1);
  for (int i2 = 0; i2 < n2; i2++) {
    if (fread(rec, sizeof(float), n1, stdin) != (size_t)n1) err("<error>: input shorter than n1*n2 samples!");
    if (i2 >= fx2 && i2 < fx2 + nx2 && fwrite(rec + fx1, sizeof(float), nx1, stdout) != (size_t)nx1)
      err("<error>: write failed!");
  }
  free(rec);
  return 0;
}
