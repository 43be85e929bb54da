/*
 * sufdmod2dTTIelastic.c -- the three drop-in modelling programs, built once per
 * rotation form with PROGRAM_FORM set to 0 (coordinate rotation), 1 (tensor
 * rotation) or 2 (Bond transformation); the executables carry the reference names
 * sufdmod2dTTIelastic_{coordinate_rotation,tensor_rotation,bond_transformation}.
 */
#include "../../include/fdmod_hip.h"

int fdmodel_main(int argc, char **argv, int form, const char **sdoc);

#ifndef PROGRAM_FORM
#define PROGRAM_FORM FDMOD_FORM_BOND
#endif

#if PROGRAM_FORM == FDMOD_FORM_BOND
#define TITLE " 2-D TTI elastic modelling, rotated staggered grid, Bond-transformed stiffness (GPU)."
#elif PROGRAM_FORM == FDMOD_FORM_TENSOR
#define TITLE " 2-D TTI elastic modelling, rotated staggered grid, stress tensor rotation (GPU)."
#else
#define TITLE " 2-D TTI elastic modelling, rotated staggered grid, rotated derivative operators (GPU)."
#endif

static const char *sdoc[] = {
    "",
    TITLE,
    "",
    " usage: <program> nz= nx= nt= dx= dz= dt= fpeak= vpfile= epsilonfile= deltafile=",
    "                  thetafile= sourcefile= [optional parameters]",
    "",
    " Required:",
    "   nz, nx, nt      samples in depth (1st, fastest), distance (2nd) and time",
    "   dz, dx, dt      sampling intervals (m, m, s)",
    "   fpeak           peak frequency used to size the absorbing layers (Hz)",
    "   vpfile          P velocity, nx records of nz float32 samples",
    "   epsilonfile     Thomsen epsilon, same layout",
    "   deltafile       Thomsen delta, same layout",
    "   thetafile       tilt of the symmetry axis in degrees, same layout",
    "   sourcefile      SU file: one trace per source, positions in sdepth/sx (cells,",
    "                   scaled by scalel/scalco), ntr of the first trace = source count",
    " Optional:",
    "   vsfile          S velocity (default 0.577350269189626*vp, 0 where vp <= 1500)",
    "   rhofile         density (default 1741*(vp/1000)^0.25, 1000 where vp <= 1500)",
    "   scalar=1        factor applied to theta",
    "   stype=1         1 stress divergence; 2 velocity divergence; 3 curl; 4 = 2 and 3;",
    "                   5 x force; 6 z force; 7 tilted force; 8 four-point curl",
    "   recposfile      SU file: trace 1 = receiver x cells, trace 2 = receiver z cells",
    "   seistype=0      0 none; 1 vx and vz; 2 pressure; 3 all",
    "   mt=1            receiver sampling in time steps",
    "   ealThick=30     absorbing layers (MEAL); eal_alpha=-1 selects the default order",
    "   freesurface=0   1 applies the stress-image free surface",
    "   mode=1          1 forward time order; 2 reversed time order",
    "   energy=0        1 writes src.txt, Eng.txt, Ek.txt, Ep.txt",
    "   savew=0         0 crops snapshots to the model; 1 keeps the absorbing layers",
    "   wsave=1         snapshots: 0 none; 1 vx,vz; 2 pressure; 3 all",
    "   snap_sample     snapshot interval in steps (default 500/(dt*1000))",
    "   fx=0 fz=0       coordinates of the first sample (trace headers)",
    "   shotID=1        shot number in file names and headers",
    "   snap_prefix=snap  trace_prefix=traces",
    "   dumps=1         write the D, Beta, coef1, coef2, invrho and c?? parameter files",
    "",
    " Parameters may also come from files named by par=<file>.",
    0};

int main(int argc, char **argv) { return fdmodel_main(argc, argv, PROGRAM_FORM, sdoc); }
