/*
 * sumkrcev2d -- write the SU receiver-position file of the modelling programs:
 * trace 1 holds the receiver x cells, trace 2 the receiver z cells (gdel), for a
 * spread around the source from offsetmin to offsetmax every dtr cells, left
 * (mode=1), right (mode=2) or both sides (mode=0; the zero offset once).  Same
 * parameters and header fields as src/main/wavelet/sumkrcev2d.c:41-163.
 *   sumkrcev2d sx= offsetmax= nx= [gdel=0 offsetmin=0 dtr=1 mode=0] > rec.su
 */
#include <stdio.h>
#include <stdlib.h>

#include "parlite.h"
#include "suio.h"

static const char *sdoc[] = {
    "",
    " sumkrcev2d - SU receiver position file (x cells in trace 1, z cells in trace 2)",
    " sumkrcev2d sx= offsetmax= nx= [gdel=0 offsetmin=0 dtr=1 mode=0] > stdout",
    "   mode   0 both sides of the source, 1 left only, 2 right only",
    0};

int main(int argc, char **argv)
{
  initargs(argc, argv);
  set_selfdoc(sdoc);
  requestdoc(0);
  int sx, gdel, offsetmin, dtr, offsetmax, nx, mode;
  if (!getparint("sx", &sx)) err("<error>: didn't specify source position in 2nd dimension (unit: cell)!");
  if (!getparint("gdel", &gdel)) gdel = 0;
  if (!getparint("offsetmin", &offsetmin)) offsetmin = 0;
  if (!getparint("dtr", &dtr)) dtr = 1;
  if (!getparint("offsetmax", &offsetmax)) err("<error>: didn't specify maximum offset(unit: cell).");
  if (!getparint("nx", &nx)) err("<error>: didn't specify number of samples in 2nd dimension (unit: cell).");
  if (!getparint("mode", &mode)) mode = 0;
  if (dtr < 1) err("<error>: dtr must be positive!");
  int nleft = 0, nright = 0, gl = 0, gr = 0;
  if (mode == 0 || mode == 1) { /* walk left from sx-offsetmin */
    gl = sx - offsetmin;
    while (gl >= 0 && sx - gl <= offsetmax) {
      nleft++;
      gl -= dtr;
    }
    gl += dtr;
  }
  if (mode == 0 || mode == 2) { /* walk right from sx+offsetmin */
    gr = sx + offsetmin;
    while (gr < nx && gr - sx <= offsetmax) {
      nright++;
      gr += dtr;
    }
    gr -= dtr;
  }
  int ntr = nleft + nright;
  if (mode == 0 && offsetmin == 0) { /* the source column is in both walks */
    ntr--;
    nleft--;
  }
  if (ntr < 1 || ntr > 65535) err("<error>: the number of traces (%d) does not fit an SU trace.", ntr);
  su_trace tr;
  su_init(&tr);
  su_reserve(&tr, ntr);
  tr.tracl = 1;
  tr.trid = 1;
  tr.scalel = -1000;
  tr.scalco = -1000;
  tr.gdel = head_scalar_write((float)gdel, tr.scalel);
  tr.sx = head_scalar_write((float)sx, tr.scalco);
  tr.ns = (unsigned short)ntr;
  tr.ntr = 2;
  tr.dt = (unsigned short)(dtr * 10000000);
  tr.f1 = 0.0f;
  tr.d1 = (float)dtr;
  if (mode == 0 || mode == 1)
    for (int i = 0; i < nleft; i++) tr.data[i] = (float)(gl + i * dtr);
  if (mode == 0 || mode == 2)
    for (int i = 0; i < nright; i++) tr.data[ntr - 1 - i] = (float)(gr - i * dtr);
  if (su_write(stdout, &tr) != 0) err("<error>: write failed!");
  for (int i = 0; i < ntr; i++) tr.data[i] = (float)gdel;
  if (su_write(stdout, &tr) != 0) err("<error>: write failed!");
  su_free(&tr);
  return 0;
}
