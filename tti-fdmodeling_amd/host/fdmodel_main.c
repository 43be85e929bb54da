/*
 * fdmodel_main.c -- shared driver of the three drop-in modelling programs
 * sufdmod2dTTIelastic_{coordinate_rotation,tensor_rotation,bond_transformation}.
 *
 * It keeps the reference programs' command line (getpar names and defaults,
 * src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_bond_transformation.c:134-294),
 * their input files (nx records of nz float32 model samples, BT:345-450; SU source
 * and receiver files, L/fdmod/obsinfo.c:34-230) and their outputs (side-effect
 * parameter dumps BT:476-491/569-596, snapshots L/fdmod/isoelastic.c:240-289 and
 * L/fdmod/snapshot.c:28-39, SU traces L/fdmod/traces.c:36-72/146-196, energy text
 * files BT:970-996, the two timing lines BT:955-958/1031-1033).  Everything between
 * reading the inputs and writing the outputs runs in libfdmod_hip on the GPU.
 *
 * Every check that can fail (missing parameters, unreadable files, unstable dt,
 * positions outside the model) runs before the first device call.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/fdmod_hip.h"
#include "parlite.h"
#include "suio.h"

static double wtime(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static float *read_model(const char *path, const char *what, int nz, int nx)
{
  FILE *fp = fopen(path, "rb");
  if (!fp) err("<error>: fail to open %s file: %s!", what, path);
  float *a = (float *)malloc(sizeof(float) * (size_t)nz * nx);
  if (fread(a, sizeof(float), (size_t)nz * nx, fp) != (size_t)nz * nx)
    err("<error>: %s file %s holds fewer than nz*nx = %d samples!", what, path, nz * nx);
  fclose(fp);
  return a;
}

static void write_bin(const char *name, const float *a, size_t n)
{
  FILE *fp = fopen(name, "wb");
  if (!fp) err("<error>: cannot create %s!", name);
  if (fwrite(a, sizeof(float), n, fp) != n) err("<error>: short write on %s!", name);
  fclose(fp);
}

typedef struct {
  int ntrace;
  int *posx, *posz;
  float *src; /* [ntrace][nt] */
} source_t;

typedef struct {
  int ntrace, ntsample;
  int *posx, *posz;
} recv_t;

/* read2dsource (L/fdmod/obsinfo.c:34-87): the first trace's ntr gives the number of
 * source traces; positions come from sdepth/scalel and sx/scalco in cells; the
 * wavelet is resampled onto t = it*dt. */
static void read_source(source_t *s, const char *path, float dt, int nx, int nz, int nt)
{
  FILE *fp = fopen(path, "rb");
  if (!fp) err("<error>: failed to open sourcefile (%s)!", path);
  su_trace tr;
  su_init(&tr);
  if (su_read(fp, &tr) != 1) err("<error>: no trace in sourcefile (%s)!", path);
  s->ntrace = tr.ntr > 0 ? tr.ntr : 1;
  s->posx = (int *)calloc(s->ntrace, sizeof(int));
  s->posz = (int *)calloc(s->ntrace, sizeof(int));
  s->src = (float *)calloc((size_t)s->ntrace * nt, sizeof(float));
  float *tm = (float *)malloc(sizeof(float) * nt);
  for (int it = 0; it < nt; it++) tm[it] = it * dt;
  warn("<warning>: the number of trace is %d in sourcefile (%s)!", s->ntrace, path);
  int itrace = -1, rc = 1;
  do {
    itrace++;
    if (itrace >= s->ntrace) err("<error>: the number of traces in sourcefile (%s) is out of %d!", path, s->ntrace);
    if (tr.scalel == 0 || tr.scalco == 0) err("<error> : scalar must not be 0!");
    s->posz[itrace] = head_scalar_read((float)tr.sdepth, tr.scalel);
    if (s->posz[itrace] < 0 || s->posz[itrace] > nz)
      err("<error>: the position of source in 1st dimension (%d) must be between 0 and %d!", s->posz[itrace], nz);
    s->posx[itrace] = head_scalar_read((float)tr.sx, tr.scalco);
    if (s->posx[itrace] < 0 || s->posx[itrace] > nx)
      err("<error>: the position of source in 2nd dimension (%d) must be between 0 and %d!", s->posx[itrace], nx);
    resample8(tr.ns, tr.dt / 1000000.0f, tr.f1, tr.data, nt, tm, s->src + (size_t)itrace * nt);
  } while ((rc = su_read(fp, &tr)) == 1);
  if (rc < 0) err("<error>: truncated trace in sourcefile (%s)!", path);
  if (itrace + 1 < s->ntrace)
    warn("<warning>: sourcefile (%s) holds %d of the %d announced traces; the rest are silent!", path, itrace + 1,
         s->ntrace);
  fclose(fp);
  free(tm);
  su_free(&tr);
}

/* read2dreceiver (L/fdmod/obsinfo.c:175-230): trace 1 holds the x cells, trace 2
 * the z cells; ntsample = (nt-1)/mt + 1. */
static void read_receivers(recv_t *r, const char *path, int nx, int nz, int nt, int mt)
{
  if (!path) err("<error>: must specify recposfile to record traces!");
  FILE *fp = fopen(path, "rb");
  if (!fp) err("<error>: failed to open recposfile (%s)!", path);
  su_trace trx, trz;
  su_init(&trx);
  su_init(&trz);
  if (su_read(fp, &trx) != 1 || su_read(fp, &trz) != 1)
    err("<error>: failed to get traces in recposfile (%s)!", path);
  if (trx.ns != trz.ns)
    err("<error>: the number of Group coordinate's samples are not match, Group X coordinate : %d, Group Z "
        "coordinate : %d.", trx.ns, trz.ns);
  r->ntrace = trx.ns;
  r->ntsample = (nt - 1) / mt + 1;
  r->posx = (int *)calloc(r->ntrace > 0 ? r->ntrace : 1, sizeof(int));
  r->posz = (int *)calloc(r->ntrace > 0 ? r->ntrace : 1, sizeof(int));
  for (int i = 0; i < r->ntrace; i++) {
    r->posz[i] = (int)trz.data[i];
    if (r->posz[i] < 0 || r->posz[i] > nz)
      err("<error>: the position of receiver in 1st dimension (%d) must be between 0 and %d!", r->posz[i], nz);
    r->posx[i] = (int)trx.data[i];
    if (r->posx[i] < 0 || r->posx[i] > nx)
      err("<error>: the position of receiver in 2nd dimension (%d) must be between 0 and %d!", r->posx[i], nx);
  }
  warn("the number of traces is %d!", r->ntrace);
  warn("the number of samples is %d in traces!", r->ntsample);
  fclose(fp);
  su_free(&trx);
  su_free(&trz);
}

/* traces2d_output / putraces_2d (L/fdmod/traces.c:36-72, 146-196) */
static void write_traces(const char *prefix, const char *comp, const source_t *s, const recv_t *r, const float *rec,
                         float fx, float fz, float dx, float dz, float dt, int mt, int shotID)
{
  char name[512];
  dt = mt * dt;
  snprintf(name, sizeof name, "%s_%s_%04d_%05d_%05d_%05.2fms.su", prefix, comp, shotID, r->ntsample, r->ntrace,
           dt * 1000.0f);
  FILE *fp = fopen(name, "wb");
  if (!fp) err("<error>: cannot create %s!", name);
  su_trace tr;
  su_init(&tr);
  tr.tracl = 1;
  tr.ns = (unsigned short)r->ntsample;
  tr.dt = (unsigned short)(dt * 1000000);
  tr.ntr = r->ntrace;
  tr.scalel = -1000;
  tr.scalco = -1000;
  tr.f1 = 0.0f;
  tr.d1 = dt;
  tr.fldr = shotID;
  tr.sx = head_scalar_write(fx + s->posx[0] * dx, tr.scalco);
  tr.sdel = head_scalar_write(fz + s->posz[0] * dz, tr.scalel);
  for (int i = 0; i < r->ntrace; i++) {
    tr.tracr = i + 1;
    tr.trid = (short)r->posx[i];
    tr.gx = head_scalar_write(fx + r->posx[i] * dx, tr.scalco);
    tr.gdel = head_scalar_write(fz + r->posz[i] * dz, tr.scalel);
    tr.offset = tr.gx - tr.sx;
    tr.cdp = (int)(0.5f * (tr.gx + tr.sx));
    tr.data = (float *)(rec + (size_t)i * r->ntsample);
    if (su_write(fp, &tr) != 0) err("<error>: short write on %s!", name);
  }
  tr.data = NULL;
  fclose(fp);
}

typedef struct {
  const char *prefix;
  float dt;
} snapctx_t;

/* snapshot sink: file names of elastic2dSnapshot (L/fdmod/isoelastic.c:249-288) */
static int snapshot_sink(void *user, int it, const char *kind, const float *data, int n1, int n2)
{
  snapctx_t *sc = (snapctx_t *)user;
  char name[512];
  snprintf(name, sizeof name, "%s_%s_%05d_%05d_%05dms.bin", sc->prefix, kind, n1, n2, (int)(1000 * it * sc->dt));
  FILE *fp = fopen(name, "wb");
  if (!fp) return 1;
  size_t n = (size_t)n1 * n2;
  int bad = fwrite(data, sizeof(float), n, fp) != n;
  fclose(fp);
  return bad;
}

#define CHECK(ctx, call)                                                                         \
  do {                                                                                          \
    int rc_ = (call);                                                                           \
    if (rc_ != FDMOD_OK) err("<error>: %s: %s (%s)", #call, fdmod_strerror(rc_),                \
                             (ctx) ? fdmod_last_error(ctx) : "");                               \
  } while (0)

int fdmodel_main(int argc, char **argv, int form, const char **sdoc)
{
  initargs(argc, argv);
  set_selfdoc(sdoc);
  requestdoc(0);
  double timeStart = wtime();

  int nz, nx, nt, freesurface, ealThick, mode, savew, wsave, stype, seistype, mt, snap_sample, shotID, energy, dumps;
  float dx, dz, dt, fx, fz, scalar, fpeak, eal_alpha;
  char *vpfile = NULL, *vsfile = NULL, *rhofile = NULL, *thetafile = NULL, *epsilonfile = NULL, *deltafile = NULL;
  char *sourcefile = NULL, *recposfile = NULL, *trace_prefix = NULL, *snap_prefix = NULL;

  /* parameters, defaults and diagnostic lines of the reference programs (BT:134-294,
   * the same in TR and CR), including their wording */
  if (!getparfloat("dx", &dx)) err("<error>: must specify sampling interval in 1st dimension!");
  if (!getparfloat("dz", &dz)) err("<error>: must specify sampling interval in 2nd dimension!");
  if (!getparfloat("dt", &dt)) err("<error>: must specify sampling interval in time dimension!");
  if (!getparfloat("fx", &fx)) {
    fx = 0.0f;
    warn("<warning>: didn't specify first sample in 2nd dimension. The program will set it as %f.", fx);
  }
  if (!getparfloat("fz", &fz)) {
    fz = 0.0f;
    warn("<warning>: didn't specify first sample in 1st dimension. The program will set it as %f.", fz);
  }
  if (!getparfloat("scalar", &scalar)) {
    scalar = 1.0f;
    warn("<warning>: didn't specify the scalar to be apllied to dip angle. The program will set it as %f.", scalar);
  }
  if (!getparfloat("fpeak", &fpeak)) err("<error>: must specify peak frequency of source wavelet!");
  if (!getparfloat("eal_alpha", &eal_alpha)) {
    eal_alpha = -1.0f;
    warn("<warning>: didn't specify the order of gaussian weighting function! The program will set it as 0.05 * "
         "ealThick + 1.5.");
  }
  if (!getparint("nz", &nz)) err("<error>: must specify the number of samples in 1st dimension!");
  if (!getparint("nx", &nx)) err("<error>: must specify the number of samples in 2nd dimension!!");
  if (!getparint("nt", &nt)) err("<error>: must specify the number of samples in time dimension!!");
  if (!getparint("freesurface", &freesurface)) {
    freesurface = 0;
    warn("<warning>: didn't specify whether to use free surface boundary conditions. The program will not use it.");
  }
  if (!getparint("ealThick", &ealThick)) {
    ealThick = 30;
    warn("<warning>: didn't specify the number of absorbing layers! The program will set it %d.", ealThick);
  }
  if (!getparint("mode", &mode)) {
    mode = 1;
    warn("<warning>: didn't specify the mode of modelng! The program will it as %d which means the program will "
         "execute the forward modeling.",
         mode);
  }
  if (!getparint("savew", &savew)) {
    savew = 0;
    warn("<warning>: didn't specify whether to save the wavefiled. The program will set it as %d which means the "
         "program will not save the wavefiled of absoring boundary area.",
         savew);
  }
  if (!getparint("wsave", &wsave)) {
    wsave = 1;
    warn("<warning>: didn't specify which wavefiled will be saved. The program will set it as %d which means the "
         "program will not save the wavefiled.",
         wsave);
  }
  if (!getparint("stype", &stype)) {
    stype = 1;
    warn("<warning>: didn't specify the type of source,  and will set stype as %d!", stype);
  }
  if (!getparint("seistype", &seistype)) {
    seistype = 0;
    /* the reference prints savew's value in this line */
    warn("<warning>: didn't specify which wavefiled will be recorded. The program will set it as %d which means the "
         "program will not save the seismic record.",
         savew);
  }
  if (!getparint("mt", &mt)) {
    mt = 1;
    warn("<warning>: didn't specify the sampling interval of seismic receiver_p. The program will set it as %d.", mt);
  }
  if (!getparint("snap_sample", &snap_sample)) {
    snap_sample = 500 / (dt * 1000);
    warn("<warning>: didn't specify the sampling interval of snapshot. The program will set it as 500 / (dt * 1000) "
         "= %d.",
         snap_sample);
  }
  if (!getparint("shotID", &shotID)) {
    shotID = 1;
    warn("<warning>: didn't specify the ID number of shot. The program will set it as %d.", shotID);
  }
  if (!getparint("energy", &energy)) {
    energy = 0;
    warn("<warning>: didn't specify whether to calculate the energy of wavefiled. The program will set it as %d.",
         energy);
  }
  if (!getparint("dumps", &dumps)) dumps = 1; /* parameter/model side-effect files, as the reference writes */
  if (!getparstring("vpfile", &vpfile)) err("<error>: must specify P-wave velocity!");
  if (!getparstring("vsfile", &vsfile))
    warn("<warning>: didn't specify S-wave velocity! The program will set S-wave velocity as vp / sqrt(3)!");
  if (!getparstring("rhofile", &rhofile))
    warn("<warning>: didn't specify density file! The program will set density as 310 * pow(vp, 0.25)!");
  if (!getparstring("epsilonfile", &epsilonfile)) err("<error>: must specify Thomsen parameter (epsilon) file!");
  if (!getparstring("deltafile", &deltafile)) err("<error>: must specify Thomsen parameter (delta) file!");
  if (!getparstring("thetafile", &thetafile)) err("<error>: must specify dip angle (theta) file!");
  if (!getparstring("sourcefile", &sourcefile)) err("<error>: must specify sourcefile!");
  if (!getparstring("recposfile", &recposfile)) warn("<warning>: didn't specify receiver_p position file!");
  if (!getparstring("snap_prefix", &snap_prefix)) {
    snap_prefix = "snap";
    warn("<warning>: didn't specify prefix of snapshot file! The program will set it as %s.", snap_prefix);
  }
  if (!getparstring("trace_prefix", &trace_prefix)) {
    trace_prefix = "traces";
    warn("<warning>: didn't specify the prefix of traces file. The program will set it as %s", trace_prefix);
  }
  if (nz < 1 || nx < 1 || nt < 1) err("<error>: nz, nx and nt must be positive!");
  if (mt < 1) err("<error>: mt must be positive!");
  if (wsave > 0 && snap_sample < 1) err("<error>: snap_sample must be positive!");

  /* model files (BT:345-450); default vs and rho (L/rockmod/modparconvert.c:42-61, 128-143) */
  size_t np = (size_t)nz * nx;
  char fileName[512];
  float *vp = read_model(vpfile, "P-wave velocity", nz, nx);
  float *vs, *rho;
  if (!vsfile) {
    vs = (float *)malloc(sizeof(float) * np);
    for (size_t i = 0; i < np; i++) vs[i] = vp[i] <= 1500.0 ? 0.0f : (float)(0.577350269189626 * vp[i]);
    snprintf(fileName, sizeof fileName, "vs_%05d_%05d.bin", nz, nx);
    write_bin(fileName, vs, np);
  } else {
    vs = read_model(vsfile, "S-wave velocity", nz, nx);
  }
  if (!rhofile) {
    rho = (float *)malloc(sizeof(float) * np);
    for (size_t i = 0; i < np; i++) rho[i] = vp[i] <= 1500.0 ? 1000.0f : (float)(1741.0 * pow(vp[i] / 1000.0, 0.25));
    snprintf(fileName, sizeof fileName, "rho_%05d_%05d.bin", nz, nx);
    write_bin(fileName, rho, np);
  } else {
    rho = read_model(rhofile, "density", nz, nx);
  }
  float *eps = read_model(epsilonfile, "Thomsen parameter (epsilon)", nz, nx);
  float *del = read_model(deltafile, "Thomsen parameter (delta)", nz, nx);
  float *theta = read_model(thetafile, "dip angle", nz, nx);

  fdmod_params P;
  P.form = form; P.nz = nz; P.nx = nx; P.nt = nt; P.dx = dx; P.dz = dz; P.dt = dt;
  P.ealThick = ealThick; P.eal_alpha = eal_alpha; P.freesurface = freesurface; P.fpeak = fpeak;
  P.scalar = scalar; P.mode = mode;
  double dt_max = 0.0;
  int rc = fdmod_check_stability(&P, vp, &dt_max);
  if (rc == FDMOD_EUNSTABLE)
    err("<error>: the finite differenct scheme is unstable! the time step size must smaller than %5.5f!", dt_max);
  if (rc != FDMOD_OK) err("<error>: invalid grid parameters (%s)!", fdmod_strerror(rc));

  source_t src;
  recv_t rvx = {0}, rp = {0};
  read_source(&src, sourcefile, dt, nx, nz, nt);
  if (seistype == 1 || seistype == 3) read_receivers(&rvx, recposfile, nx, nz, nt, mt);
  if (seistype == 2 || seistype == 3) read_receivers(&rp, recposfile, nx, nz, nt, mt);
  recv_t *rr = (seistype == 2) ? &rp : &rvx;

  /* ---- device work starts here ---- */
  fdmod_ctx *ctx = NULL;
  rc = fdmod_create(&P, &ctx);
  if (rc != FDMOD_OK) err("<error>: cannot create the GPU modelling context: %s", fdmod_strerror(rc));
  CHECK(ctx, fdmod_set_model(ctx, vp, vs, rho, eps, del, theta, &dt_max));
  int dims[7];
  fdmod_dims(ctx, dims);
  int Nz = dims[0], Nx = dims[1], ealEff = dims[2];
  size_t Np = (size_t)Nz * Nx;
  if (dumps) { /* BT:476-491 and 569-596; c15/c35 exist for the Bond form only */
    static const struct { int id; const char *name; } dl[] = {
        {FDMOD_A_D, "D"}, {FDMOD_A_BETA, "Beta"}, {FDMOD_A_COEF1, "coef1"}, {FDMOD_A_COEF2, "coef2"},
        {FDMOD_A_INVRHO, "invrho"}, {FDMOD_A_C11, "c11"}, {FDMOD_A_C13, "c13"}, {FDMOD_A_C15, "c15"},
        {FDMOD_A_C33, "c33"}, {FDMOD_A_C35, "c35"}, {FDMOD_A_C55, "c55"}};
    float *buf = (float *)malloc(sizeof(float) * Np);
    for (size_t k = 0; k < sizeof dl / sizeof dl[0]; k++) {
      if ((dl[k].id == FDMOD_A_C15 || dl[k].id == FDMOD_A_C35) && form != FDMOD_FORM_BOND) continue;
      CHECK(ctx, fdmod_get_array(ctx, dl[k].id, buf));
      snprintf(fileName, sizeof fileName, "%s_%05d_%05d.bin", dl[k].name, Nz, Nx);
      write_bin(fileName, buf, Np);
    }
    free(buf);
  }
  CHECK(ctx, fdmod_set_source(ctx, stype, src.ntrace, src.posx, src.posz, src.src));
  if (seistype > 0) CHECK(ctx, fdmod_set_receivers(ctx, seistype, mt, rr->ntrace, rr->posx, rr->posz));
  snapctx_t sc = {snap_prefix, dt};
  if (wsave > 0) CHECK(ctx, fdmod_set_snapshots(ctx, wsave, savew, snap_sample, snapshot_sink, &sc));
  if (energy == 1) CHECK(ctx, fdmod_set_energy(ctx, 1));

  fprintf(stderr, "%s: Input Parameters:\n", argv[0]);
  fprintf(stderr, "%s: nx = %d, nz = %d, nt = %d\n", argv[0], nx, nz, nt);
  fprintf(stderr, "%s: Nx = %d, Nz = %d\n", argv[0], Nx, Nz);
  fprintf(stderr, "%s: dx = %6.5f, dz = %6.5f, dt = %6.5f\n", argv[0], dx, dz, dt);
  fprintf(stderr, "%s: ixStart = %d, ixEnd = %d, izStart = %d, izEnd = %d\n", argv[0], dims[5], dims[6], dims[3],
          dims[4]);

  double tStart = wtime();
  CHECK(ctx, fdmod_run(ctx, 0, nt, NULL));
  double tCost = wtime() - tStart;
  fprintf(stderr, "\n");
  fprintf(stdout, "Finite difference modeling takes %lf s.\n", tCost);

  if (seistype > 0) {
    float *rec = (float *)malloc(sizeof(float) * (size_t)rr->ntrace * rr->ntsample);
    if (seistype == 1 || seistype == 3) {
      CHECK(ctx, fdmod_get_traces(ctx, FDMOD_TRACE_VX, rec));
      write_traces(trace_prefix, "vx", &src, &rvx, rec, fx, fz, dx, dz, dt, mt, shotID);
      CHECK(ctx, fdmod_get_traces(ctx, FDMOD_TRACE_VZ, rec));
      write_traces(trace_prefix, "vz", &src, &rvx, rec, fx, fz, dx, dz, dt, mt, shotID);
    }
    if (seistype == 2 || seistype == 3) {
      CHECK(ctx, fdmod_get_traces(ctx, FDMOD_TRACE_P, rec));
      write_traces(trace_prefix, "p", &src, &rp, rec, fx, fz, dx, dz, dt, mt, shotID);
    }
    free(rec);
  }
  if (energy == 1) { /* BT:970-996 */
    float *E = (float *)malloc(sizeof(float) * nt), *Ek = (float *)malloc(sizeof(float) * nt);
    float *Ep = (float *)malloc(sizeof(float) * nt);
    CHECK(ctx, fdmod_get_energy(ctx, E, Ek, Ep));
    const struct { const char *name; const float *a; } out[] = {
        {"src.txt", src.src}, {"Eng.txt", E}, {"Ek.txt", Ek}, {"Ep.txt", Ep}};
    for (int k = 0; k < 4; k++) {
      FILE *fp = fopen(out[k].name, "w");
      if (!fp) err("<error>: cannot create %s!", out[k].name);
      for (int it = 0; it < nt; it++) fprintf(fp, "%10.5e\n", out[k].a[it]);
      fclose(fp);
    }
    free(E); free(Ek); free(Ep);
  }
  (void)ealEff;
  fdmod_destroy(ctx);
  free(vp); free(vs); free(rho); free(eps); free(del); free(theta);
  free(src.posx); free(src.posz); free(src.src);
  free(rvx.posx); free(rvx.posz); free(rp.posx); free(rp.posz);
  fprintf(stdout, "Total time takes %lf s.\n", wtime() - timeStart);
  return 0;
}
