// libfdmod_hip adjoint entry points: the misfit gradient, the adjoint of the
// modelling operator, the device dot-product test and the frozen-absorber model
// update, for the three rotation forms.  Declarations are in include/fdmod_hip.h.
//
// Gradient schedule:
//   1. forward from rest, recording traces; with checkpoints every seg-th state
//      (5 fields) is stored, with the full history the strain E(v_{k+1}) of every
//      step, written by the stress sweep itself;
//   2. residuals r = d - obs and J = sum r^2 / 2, reduced in FP64 on the device;
//   3. for each checkpoint segment, last first: restore its state and recompute its
//      steps storing the strain of each (skipped with the full history), then run
//      the adjoint steps of the segment backwards, imaging with the stored strain;
//      with boundary-saving reconstruction, each adjoint step is preceded by the
//      time-reversed loop iteration, which rebuilds the state and stores the strain;
//   4. chain rule from the six stiffness sums to (vp, eps, delta, theta) and the
//      transpose of the edge padding.
// Stored fields are FP32 copies or FP16 with per-(column, 64-row block) scales.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "fdmod_ctx.h"

using namespace fdm;

namespace {

constexpr int kRedBlocks = 256;  // partial sums of the FP64 reductions
constexpr int kMaxState = 8;     // vx, vz, sxx, szz, sxz (+ txx, tzz, txz for the tensor form)
constexpr int kNG = 8;           // imaging sums
constexpr int kAdjSegS = 44, kAdjSegV = 44;  // longest adjoint segments (columns)
// blocks per adjoint sweep below which they shorten: the stress adjoint counts strips of
// kTZS = 55 output rows, the velocity adjoint strips of 64
constexpr int kAdjMinTilesS = 1700, kAdjMinTilesV = 1450;

size_t field_floats(const fdmod_ctx *c) { return (size_t)c->g.Nx * c->g.ldz; }

// Gradient field arrays (adjoint fields, FP64 sums, snapshots, history and checkpoint
// slots) keep the row alignment of the modelling pool: a block of same-shape fields
// holds kSpare extra elements, and its field k starts at k * field_floats + the
// context's shift (zsh for 4- and 8-byte elements, zsh16 for FP16), so that row
// izStart of every field begins a 128-byte line.
constexpr size_t kSpare = 2 * kZAlign;
template <typename T>
T *fld(const fdmod_ctx *c, T *base, size_t k)
{
  return base + (sizeof(T) == 2 ? c->zsh16 : c->zsh) + k * field_floats(c);
}

// The adjoint stress rotates through kMusSlots slots: the V-adjoint of each step reads
// mu_sigma from the current slot and writes the updated one into the next, so the slot
// it read is the imaging snapshot of that step, kept until the block's imaging has run
// (at most kImgK steps later; the slot is rewritten kImgK + 1 steps later).
constexpr int kMusSlots = kImgK + 1;
float *mus_field(const fdmod_ctx *c, int slot, int q) { return fld(c, c->d_mus, (size_t)slot * 3 + q); }

const int kStateIds[kMaxState] = {FDMOD_A_VX,  FDMOD_A_VZ,  FDMOD_A_SXX, FDMOD_A_SZZ,
                                  FDMOD_A_SXZ, FDMOD_A_TXX, FDMOD_A_TZZ, FDMOD_A_TXZ};

// Per-form counts: state (= adjoint) fields, imaging snapshots, history fields.
int n_state(int form) { return form == FDMOD_FORM_TENSOR ? 8 : 5; }
int n_snap(int form) { return form == FDMOD_FORM_BOND ? 3 : (form == FDMOD_FORM_TENSOR ? 6 : 5); }
int n_hist(int form) { return form == FDMOD_FORM_BOND ? 3 : (form == FDMOD_FORM_TENSOR ? 6 : 8); }

int check_ready(fdmod_ctx *c, const char *what)
{
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "%s before fdmod_set_model", what);
  if (c->seistype == 0 || c->nr == 0) return fail(c, FDMOD_ESTATE, "%s needs receivers", what);
  return FDMOD_OK;
}

int workspace(fdmod_ctx *c, bool need_G)
{
  const size_t n = field_floats(c);
  if (c->d_mu_n < n) {  // (re)size the per-field workspace, which a batch may lend
    dfree(c->d_mu); dfree(c->d_mus); dfree(c->d_G); dfree(c->d_snap);
    c->d_mu = nullptr; c->d_mus = nullptr; c->d_G = nullptr; c->d_snap = nullptr;
    c->d_mu_n = 0;
  }
  if (!c->d_mu) HIPCALL(c, hipMalloc(&c->d_mu, (kMaxState * n + kSpare) * sizeof(float)));
  if (!c->d_mus) HIPCALL(c, hipMalloc(&c->d_mus, ((size_t)kMusSlots * 3 * n + kSpare) * sizeof(float)));
  c->d_mu_n = n;
  if (!c->d_red) HIPCALL(c, hipMalloc(&c->d_red, (kRedBlocks + 8) * sizeof(double)));
  if (need_G && !c->d_G) HIPCALL(c, hipMalloc(&c->d_G, (kNG * n + kSpare) * sizeof(double)));
  if (need_G && !c->d_snap) HIPCALL(c, hipMalloc(&c->d_snap, ((size_t)kImgK * 6 * n + kSpare) * sizeof(float)));
  dfree(c->d_res);
  c->d_res = nullptr;
  HIPCALL(c, hipMalloc(&c->d_res, 3 * (size_t)c->nr * c->ntsample * sizeof(float)));
  return FDMOD_OK;
}

int reserve(fdmod_ctx *c, void **p, size_t &have, size_t need)
{
  if (have >= need) return FDMOD_OK;
  dfree(*p);
  *p = nullptr;
  have = 0;
  if (hipMalloc(p, need) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, FDMOD_ENOMEM, "cannot allocate %.2f GB of gradient storage", need / 1e9);
  }
  have = need;
  return FDMOD_OK;
}

size_t scales_per_field(const fdmod_ctx *c) { return (size_t)c->g.Nx * (c->g.ldz / TZ); }

// Capacity for nslots slots of nf fields.
int store_reserve(fdmod_ctx *c, FieldStore &st, size_t nslots, int nf, bool f16)
{
  const size_t n = field_floats(c) * nslots * nf + kSpare;
  int rc = reserve(c, &st.data, st.cap_data, n * (f16 ? sizeof(_Float16) : sizeof(float)));
  if (rc || !f16) return rc;
  return reserve(c, (void **)&st.scale, st.cap_scale, scales_per_field(c) * nslots * nf * sizeof(float));
}

// Copy (FP32) or pack (FP16) the nf fields into / out of a slot.
int store_io(fdmod_ctx *c, FieldStore &st, size_t slot, int nf, float *const *fields, bool f16, bool save)
{
  const size_t n = field_floats(c);
  if (!f16) {
    for (int f = 0; f < nf; f++) {
      float *slotf = fld(c, (float *)st.data, slot * nf + f);
      HIPCALL(c, hipMemcpyAsync(save ? slotf : fields[f], save ? fields[f] : slotf, n * sizeof(float),
                                hipMemcpyDeviceToDevice, c->stream));
    }
    return FDMOD_OK;
  }
  Pack16 pk{};
  pk.nx = c->g.Nx;
  pk.ldz = c->g.ldz;
  pk.nf = nf;
  const size_t ns = scales_per_field(c);
  for (int f = 0; f < nf; f++) {
    pk.src[f] = fields[f];
    pk.dst[f] = fld(c, (_Float16 *)st.data, slot * nf + f);
    pk.inv[f] = st.scale + (slot * nf + f) * ns;
    pk.mate[f] = -1;
  }
  // normal stresses as sum and difference (kStateIds: sxx, szz = 2, 3; txx, tzz = 5, 6)
  for (int a = 2; a + 1 < nf && a <= 5; a += 3) {
    pk.mate[a] = a + 1;
    pk.mate[a + 1] = a;
  }
  launch_pack16(c->stream, pk, save);
  return FDMOD_OK;
}

// History slots hold the forward fields one step's imaging needs (n_hist fields; in
// FP16, n_hist scale arrays of scales_per_field each).
SweepArgs with_hist(fdmod_ctx *c, const SweepArgs &sa, size_t slot, bool f16)
{
  SweepArgs r = sa;
  const int nh = n_hist(c->p.form);
  const size_t ns = scales_per_field(c);
  r.estore = f16 ? 2 : 1;
  for (int f = 0; f < nh; f++) {
    if (f16) r.h16[f] = fld(c, (_Float16 *)c->hist.data, slot * nh + f);
    else r.h32[f] = fld(c, (float *)c->hist.data, slot * nh + f);
  }
  if (f16)
    for (int f = 0; f < nh; f++) r.hsc[f] = c->hist.scale + (slot * nh + f) * ns;
  return r;
}

// Imaging block entry j: snapshot slot j and history slot `slot`.
void add_to_block(fdmod_ctx *c, ImgBlock &b, int j, size_t slot, bool f16)
{
  const int nh = n_hist(c->p.form);
  const size_t ns = scales_per_field(c);
  for (int q = 0; q < 6; q++) b.lam[j][q] = q < 3 ? mus_field(c, c->mus_last, q) : fld(c, c->d_snap, (size_t)j * 6 + q);
  for (int f = 0; f < 8; f++) {
    b.h32[j][f] = nullptr;
    b.h16[j][f] = nullptr;
    if (f >= nh) continue;
    if (f16) b.h16[j][f] = fld(c, (const _Float16 *)c->hist.data, slot * nh + f);
    else b.h32[j][f] = fld(c, (const float *)c->hist.data, slot * nh + f);
  }
  b.hsc[j] = f16 ? c->hist.scale + slot * nh * ns : nullptr;
  b.hsc_stride = ns;
  b.inj[j] = c->inj_last;  // set by the adjoint step just run
  b.inj_x0 = c->rline_x0;
  b.inj_z = c->rline_z;
  b.inj_n = c->nr;
}

// Segment lengths of the two adjoint sweeps, chosen once per context.  The adjoint
// sweeps stage 2-6 fields computed from up to 10 loaded operands, so a segment's
// halo columns cost them more than they cost the forward sweeps: they keep longer
// segments (fewer, longer blocks) down to a smaller block count.  On 1126-deep crops
// 1500, 2000 and 3000 columns wide the fastest pairs were 22/22, 22/22 and 33/33
// columns, 3-7% faster backward passes than 11-column segments; the longest segment
// giving at least kAdjMinTilesS / kAdjMinTilesV blocks picks each of them.
void adj_segments(fdmod_ctx *c)
{
  if (c->adj_segS > 0) return;
  const Grid &g = c->g;
  // strips: the stress adjoint outputs kTZS rows per strip, the velocity adjoint TZ
  auto pick = [&](int longest, int min_tiles, int strips) {
    int L = longest;
    while (L > kFlagW && strips * ((g.ixEnd - g.ixStart + L - 1) / L) < min_tiles) L -= kFlagW;
    return L;
  };
  c->adj_segS = pick(kAdjSegS, kAdjMinTilesS, (g.izEnd - g.izStart + kTZS - 1) / kTZS);
  c->adj_segV = pick(kAdjSegV, kAdjMinTilesV, g.tilesZ);
  if (const char *e = getenv("FDMOD_ADJ_SEGS")) {  // tile-shape overrides (tuning experiments)
    const int v = atoi(e);
    if (v >= 1 && v <= 4096) c->adj_segS = v;
  }
  if (const char *e = getenv("FDMOD_ADJ_SEGV")) {
    const int v = atoi(e);
    if (v >= 1 && v <= 4096) c->adj_segV = v;
  }
}

AdjArgs adj_args(fdmod_ctx *c, bool image_rows)
{
  AdjArgs a{};
  adj_segments(c);
  a.g = c->g;
  a.g.segLV = c->adj_segS;  // the stress adjoint is tiled like the velocity sweep
  a.g.segLS = c->adj_segV;  // and the velocity adjoint like the stress sweep
  a.izU = c->g.izStart;
  a.undamped = c->flags;
  a.ftz = c->g.tilesZ;
  if (image_rows && c->p.freesurface) {
    a.g.izStart -= kN;
    a.g.tilesZ = (a.g.izEnd - a.g.izStart + TZ - 1) / TZ;
  }
  float *const *d = c->d.a;
  a.form = c->p.form;
  a.mvx = fld(c, c->d_mu, 0); a.mvz = fld(c, c->d_mu, 1);
  a.mxx = mus_field(c, c->mus_cur, 0); a.mzz = mus_field(c, c->mus_cur, 1); a.mxz = mus_field(c, c->mus_cur, 2);
  a.oxx = a.mxx; a.ozz = a.mzz; a.oxz = a.mxz;
  a.mtxx = fld(c, c->d_mu, 5); a.mtzz = fld(c, c->d_mu, 6); a.mtxz = fld(c, c->d_mu, 7);
  a.coef1 = d[FDMOD_A_COEF1]; a.coef2 = d[FDMOD_A_COEF2]; a.invrho = d[FDMOD_A_INVRHO];
  a.c11 = d[FDMOD_A_C11]; a.c13 = d[FDMOD_A_C13]; a.c15 = d[FDMOD_A_C15];
  a.c33 = d[FDMOD_A_C33]; a.c35 = d[FDMOD_A_C35]; a.c55 = d[FDMOD_A_C55];
  a.sin2 = d[FDMOD_A_SIN2]; a.cos2 = d[FDMOD_A_COS2]; a.sin2t = d[FDMOD_A_SIN2T];
  a.sn = d[FDMOD_A_SIN]; a.cs = d[FDMOD_A_COS]; a.snv = d[FDMOD_A_SINV]; a.csv = d[FDMOD_A_COSV];
  for (int q = 0; q < 6; q++) a.snap[q] = nullptr;
  for (int j = 0; j < kNG; j++) a.G[j] = c->d_G ? fld(c, c->d_G, j) : nullptr;
  return a;
}

SourceArgs source_args(fdmod_ctx *c, int it)
{
  SourceArgs s{};
  s.g = c->g;
  s.isx = c->d_isx; s.isz = c->d_isz; s.src = c->d_src;
  s.nt = c->p.nt; s.ns = c->ns; s.it = it;
  s.taps = c->d_taps; s.ntap = c->ntap; s.npass = c->npass; s.sequential = c->seq;
  return s;
}

// Data-space adjoint sources on the device, [comp][ntsample][nr] with comp 0 = vx,
// 1 = vz, 2 = p; null where a component is absent.
struct DataSrc {
  const float *vx = nullptr, *vz = nullptr, *p = nullptr;
};

// Upload host series [nr][ntsample] into the device layout [ntsample][nr].
int upload_series(fdmod_ctx *c, const float *h, float *d)
{
  std::vector<float> t((size_t)c->nr * c->ntsample);
  for (int r = 0; r < c->nr; r++)
    for (int s = 0; s < c->ntsample; s++) t[(size_t)s * c->nr + r] = h[(size_t)r * c->ntsample + s];
  HIPCALL(c, hipMemcpy(d, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  return FDMOD_OK;
}

int upload_data(fdmod_ctx *c, const float *hvx, const float *hvz, const float *hp, DataSrc &ds)
{
  const size_t nrec = (size_t)c->nr * c->ntsample;
  const bool v = c->seistype == 1 || c->seistype == 3, p = c->seistype == 2 || c->seistype == 3;
  int rc;
  if (v && hvx && hvz) {
    if ((rc = upload_series(c, hvx, c->d_res)) || (rc = upload_series(c, hvz, c->d_res + nrec))) return rc;
    ds.vx = c->d_res;
    ds.vz = c->d_res + nrec;
  }
  if (p && hp) {
    if ((rc = upload_series(c, hp, c->d_res + 2 * nrec))) return rc;
    ds.p = c->d_res + 2 * nrec;
  }
  if (!ds.vx && !ds.p) return fail(c, FDMOD_EINVAL, "no data series for the recorded components (seistype %d)",
                                   c->seistype);
  return FDMOD_OK;
}

// One adjoint step: the transpose of loop iteration k.  The V-adjoint moves mu_sigma to
// the next ring slot; the slot it read (mus_last) holds this step's imaged mu_sigma.
// snap_slot >= 0: it also stores the other imaged fields (tensor and coordinate forms)
// into that snapshot slot, and the step is imaged: pressure receivers on one row of
// consecutive columns then have their residuals added by the sweeps and the imaging
// (c->inj_last) instead of an injection launch.
void adj_step(fdmod_ctx *c, int k, const DataSrc &ds, int snap_slot, double *srcadj)
{
  const fdmod_params &p = c->p;
  hipStream_t s = c->stream;
  const int it = p.mode == 1 ? k : p.nt - 1 - k;
  AdjArgs a = adj_args(c, false);
  const bool rec = it % c->mt == 0;
  const size_t so = (size_t)(it / c->mt) * c->nr;
  // Receivers on one row of consecutive columns: the two sweeps add the pressure
  // residuals where they read mu_sigma.  Not with a free surface (FS^T runs between the
  // injection and the stress adjoint) or the source adjoint (it reads the injected
  // mu_sigma before the sweeps).
  const bool fuse_inj = snap_slot >= 0 && rec && ds.p && c->rline && !p.freesurface && !srcadj;
  c->inj_last = fuse_inj ? ds.p + so : nullptr;
  if (rec && ds.p && !fuse_inj)
    launch_adj_inject(s, c->nr, c->g.ldz, c->d_gx, c->d_gz, ds.p + so, ds.p + so, a.mxx, a.mzz, 0.5f);
  if (srcadj && c->ns > 0 && c->stype == 1) {
    const bool tensor = p.form == FDMOD_FORM_TENSOR;  // TR:888-891: the source also feeds tau
    launch_adj_source(s, source_args(c, it), a.mxx, a.mzz, tensor ? a.mtxx : nullptr, tensor ? a.mtzz : nullptr,
                      srcadj);
  }
  if (p.freesurface) launch_adj_fs(s, a);
  AdjArgs as = a;  // the stress adjoint's strips output kTZS rows each
  as.g.tilesZ = (c->g.izEnd - c->g.izStart + kTZS - 1) / kTZS;
  if (fuse_inj) {
    as.inj = ds.p + so;
    as.inj_x0 = c->rline_x0;
    as.inj_z = c->rline_z;
    as.inj_n = c->nr;
  }
  launch_adj_ssweep(s, as);
  if (rec && ds.vx)
    launch_adj_inject(s, c->nr, c->g.ldz, c->d_gx, c->d_gz, ds.vx + so, ds.vz + so, a.mvx, a.mvz, 1.0f);
  if (srcadj && c->ns > 0 && c->stype != 1) {
    const float *ma = c->stype == 6 ? a.mvz : a.mvx;
    const float *mb = (c->stype == 5 || c->stype == 6) ? nullptr : a.mvz;
    launch_adj_source(s, source_args(c, it), ma, mb, nullptr, nullptr, srcadj);
  }
  AdjArgs av = adj_args(c, true);
  const int nxt = (c->mus_cur + 1) % kMusSlots;
  av.oxx = mus_field(c, nxt, 0);
  av.ozz = mus_field(c, nxt, 1);
  av.oxz = mus_field(c, nxt, 2);
  if (snap_slot >= 0)
    for (int q = 3; q < n_snap(p.form); q++) av.snap[q] = fld(c, c->d_snap, (size_t)snap_slot * 6 + q);
  if (fuse_inj) {
    av.inj = as.inj;
    av.inj_x0 = as.inj_x0;
    av.inj_z = as.inj_z;
    av.inj_n = as.inj_n;
  }
  launch_adj_vsweep(s, av);
  c->mus_last = c->mus_cur;
  c->mus_cur = nxt;
}

int zero_adjoint(fdmod_ctx *c)
{
  const size_t n = field_floats(c);
  HIPCALL(c, hipMemsetAsync(c->d_mu, 0, (kMaxState * n + kSpare) * sizeof(float), c->stream));
  HIPCALL(c, hipMemsetAsync(c->d_mus, 0, ((size_t)kMusSlots * 3 * n + kSpare) * sizeof(float), c->stream));
  c->mus_cur = c->mus_last = 0;
  if (c->d_G) HIPCALL(c, hipMemsetAsync(c->d_G, 0, (kNG * n + kSpare) * sizeof(double), c->stream));
  return FDMOD_OK;
}

int state_io(fdmod_ctx *c, size_t slot, bool f16, bool save)
{
  const int ns = n_state(c->p.form);
  float *f[kMaxState];
  for (int i = 0; i < ns; i++) f[i] = c->d.a[kStateIds[i]];
  return store_io(c, c->ckpt, slot, ns, f, f16, save);
}


// ---- boundary-saving reconstruction ----
// R is the undamped interior (coef1 = 1, coef2 = dt; under a free surface, the rows
// below the surface row), where a loop iteration can be undone: with dt and coef2
// negated the two sweeps subtract what they added, and the sources come off with
// negated series.  A point of R reads the other family of fields up to kN points
// away, so the state on the band of width kN around R is saved at every step and put
// back after each reversed sweep.  Outside R the reversed fields are not the
// forward's, so the history of the frame U \ R (U = the update range) is saved by
// the forward and written over the history the reversed sweeps store there.
// Interior variant: U \ R is the absorbing layers (plus the surface row under a free
// surface), whose imaging adds, through the transpose of the edge padding, only to
// the model's outermost cells.  It saves the surface row's history alone, stores the
// reversed history inside R only and images zero elsewhere: the gradient then leaves
// out the absorbing layers' contributions to those cells.
struct Recon {
  bool interior = false;
  int rx0 = 0, rx1 = 0, rz0 = 0, rz1 = 0;
  FrameIO band{}, frame{};  // rectangles only; fields, pack and direction set per use
  size_t nband = 0, nframe = 0;
};

void set_rects(FrameIO &f, const int (*r)[4], int n, int ldz)
{
  f = FrameIO{};
  f.ldz = ldz;
  long long off = 0;
  for (int i = 0; i < n; i++) {
    const int x0 = r[i][0], x1 = r[i][1], z0 = r[i][2], z1 = r[i][3];
    if (x1 <= x0 || z1 <= z0) continue;
    f.x0[f.nrect] = x0;
    f.z0[f.nrect] = z0;
    f.h[f.nrect] = z1 - z0;
    f.off[f.nrect] = off;
    off += (long long)(x1 - x0) * (z1 - z0);
    f.nrect++;
  }
  f.off[f.nrect] = off;
}

// False when the grid has no undamped interior with room for the band.
bool recon_shape(const fdmod_ctx *c, Recon &rg, bool interior)
{
  const Grid &g = c->g;
  rg.interior = interior;
  rg.rx0 = g.e;
  rg.rx1 = g.Nx - g.e;
  rg.rz0 = c->p.freesurface ? g.e + 1 : g.e;
  rg.rz1 = g.Nz - g.e;
  if (rg.rx1 <= rg.rx0 || rg.rz1 <= rg.rz0 || rg.rx0 < g.ixStart || rg.rz0 < g.izStart || rg.rx0 < kN ||
      rg.rz0 < kN || rg.rx1 + kN > g.Nx || rg.rz1 + kN > g.Nz)
    return false;
  const int bx0 = rg.rx0 - kN, bx1 = rg.rx1 + kN, bz0 = rg.rz0 - kN, bz1 = rg.rz1 + kN;
  const int band[4][4] = {{bx0, bx1, bz0, rg.rz0},
                          {bx0, bx1, rg.rz1, bz1},
                          {bx0, rg.rx0, rg.rz0, rg.rz1},
                          {rg.rx1, bx1, rg.rz0, rg.rz1}};
  const int frame[4][4] = {{g.ixStart, g.ixEnd, g.izStart, rg.rz0},
                           {g.ixStart, g.ixEnd, rg.rz1, g.izEnd},
                           {g.ixStart, rg.rx0, rg.rz0, rg.rz1},
                           {rg.rx1, g.ixEnd, rg.rz0, rg.rz1}};
  const int surface[1][4] = {{rg.rx0, rg.rx1, g.izStart, rg.rz0}};  // empty without a free surface
  set_rects(rg.band, band, 4, g.ldz);
  if (interior) set_rects(rg.frame, surface, 1, g.ldz);
  else set_rects(rg.frame, frame, 4, g.ldz);
  rg.nband = (size_t)rg.band.off[rg.band.nrect];
  rg.nframe = (size_t)rg.frame.off[rg.frame.nrect];
  return true;
}

int recon_geometry(fdmod_ctx *c, Recon &rg, bool interior)
{
  if (!recon_shape(c, rg, interior))
    return fail(c, FDMOD_EINVAL, "boundary-saving reconstruction: the grid has no undamped interior");
  return FDMOD_OK;
}

// R must be undamped for the reversal to be exact: count the points where it is not.
int check_undamped(fdmod_ctx *c, const Recon &rg)
{
  unsigned *cnt = (unsigned *)(c->d_red + kRedBlocks + 4);
  unsigned h = 0;
  HIPCALL(c, hipMemsetAsync(cnt, 0, sizeof(unsigned), c->stream));
  launch_count_damped(c->stream, c->d.a[FDMOD_A_COEF1], c->d.a[FDMOD_A_COEF2], c->g.dt, c->g.ldz, rg.rx0, rg.rx1,
                      rg.rz0, rg.rz1, cnt);
  HIPCALL(c, hipMemcpyAsync(&h, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (h) return fail(c, FDMOD_EINVAL, "boundary-saving reconstruction: %u damped points inside the interior", h);
  return FDMOD_OK;
}

// -coef2 and -src for the reversed steps.
int recon_negated(fdmod_ctx *c)
{
  const size_t n = field_floats(c), ns = (size_t)c->ns * c->p.nt;
  if (!c->d_coef2n) HIPCALL(c, hipMalloc(&c->d_coef2n, (n + kSpare) * sizeof(float)));
  if (ns && c->d_srcn_n < ns) {
    dfree(c->d_srcn);
    c->d_srcn = nullptr;
    c->d_srcn_n = 0;
    HIPCALL(c, hipMalloc(&c->d_srcn, ns * sizeof(float)));
    c->d_srcn_n = ns;
  }
  launch_negate(c->stream, c->d.a[FDMOD_A_COEF2], fld(c, c->d_coef2n, 0), n);
  if (ns) launch_negate(c->stream, c->d_src, c->d_srcn, ns);
  return FDMOD_OK;
}

// Copy nf fields between the grid and a compact frame store.
void frame_io(fdmod_ctx *c, const FrameIO &shape, float *const *fields, int nf, float *pack, bool save)
{
  FrameIO f = shape;
  f.nf = nf;
  f.save = save ? 1 : 0;
  for (int i = 0; i < nf; i++) f.grid[i] = fields[i];
  f.pack = pack;
  launch_frame_io(c->stream, f);
}

// Run the whole adjoint loop with the given data-space sources (no imaging).
int adjoint_only(fdmod_ctx *c, const DataSrc &ds, double *srcadj)
{
  int rc = zero_adjoint(c);
  if (rc) return rc;
  for (int k = c->p.nt - 1; k >= 0; k--) adj_step(c, k, ds, -1, srcadj);
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipGetLastError());
  return FDMOD_OK;
}

int reduce_to_host(fdmod_ctx *c, double *out)
{
  HIPCALL(c, hipMemcpyAsync(out, c->d_red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return FDMOD_OK;
}


// One gradient computation as resumable phases, so that a batch can interleave the
// launches of several shots (each on its own stream).
struct GradRun {
  fdmod_ctx *c = nullptr;
  bool f16 = false, ck16 = false, history = false;  // FP16 strain / FP16 checkpoints
  bool recon = false;                               // boundary-saving reconstruction
  int ncp = 1, seg = 1;
  int nanch = 0;                                    // reconstruction: full states kept
  Recon rg;
  float *band = nullptr, *frame = nullptr;          // reconstruction: per-step stores
  DataSrc ds;
  SweepArgs sa, rsa;                                // forward / time-reversed sweeps
  int sgi = 0, k = 0, phase = 0;  // backward cursor: segment, step, 0 = recompute / 1 = adjoint
  ImgBlock blk{};
  bool done = false;
};

int grad_begin(GradRun &r, fdmod_ctx *c, const float *obs_vx, const float *obs_vz, const float *obs_p,
               const fdmod_grad_opts *opts)
{
  if (!opts) return fail(c, FDMOD_EINVAL, "fdmod_gradient: opts must not be NULL");
  if (opts->storage < FDMOD_STORE_F32 || opts->storage > FDMOD_STORE_F16_ALL)
    return fail(c, FDMOD_EINVAL, "fdmod_gradient: storage must be FDMOD_STORE_F32, _F16 or _F16_ALL");
  if (opts->history < FDMOD_HIST_CHECKPOINTS || opts->history > FDMOD_HIST_RECONSTRUCT_INTERIOR)
    return fail(c, FDMOD_EINVAL,
                "fdmod_gradient: history must be FDMOD_HIST_CHECKPOINTS, _FULL, _RECONSTRUCT or _RECONSTRUCT_INTERIOR");
  if (opts->history >= FDMOD_HIST_RECONSTRUCT && opts->storage != FDMOD_STORE_F32)
    return fail(c, FDMOD_EINVAL, "fdmod_gradient: boundary-saving reconstruction stores FP32 only");
  int rc = check_ready(c, "fdmod_gradient");
  if (rc) return rc;
  const fdmod_params &p = c->p;
  r = GradRun{};
  r.c = c;
  r.f16 = opts->storage != FDMOD_STORE_F32;
  r.ck16 = opts->storage == FDMOD_STORE_F16_ALL;
  r.history = opts->history == FDMOD_HIST_FULL;
  r.recon = opts->history >= FDMOD_HIST_RECONSTRUCT;
  if ((rc = workspace(c, true))) return rc;
  if ((rc = upload_data(c, obs_vx, obs_vz, obs_p, r.ds))) return rc;
  // storage: the whole velocity history, or ncp states one every seg steps plus the
  // velocity history of one segment
  int ncp = opts->ncheckpoints > 0 ? opts->ncheckpoints : (int)ceil(sqrt((double)p.nt));
  ncp = std::min(std::max(ncp, 1), p.nt);
  r.seg = (p.nt + ncp - 1) / ncp;
  r.ncp = (p.nt + r.seg - 1) / r.seg;
  if (r.recon) {
    // [nanch full states][nt bands of n_state fields][nt frames of n_hist fields];
    // kImgK history slots (slot 0 doubles as the forward's scratch)
    const bool interior = opts->history == FDMOD_HIST_RECONSTRUCT_INTERIOR;
    if ((rc = recon_geometry(c, r.rg, interior)) || (rc = check_undamped(c, r.rg)) || (rc = recon_negated(c)))
      return rc;
    r.nanch = opts->ncheckpoints > 0 ? std::min(opts->ncheckpoints, p.nt) : 0;
    r.seg = r.nanch > 0 ? (p.nt + r.nanch) / (r.nanch + 1) : p.nt + 1;
    r.nanch = r.nanch > 0 ? (p.nt - 1) / r.seg : 0;  // states at k = seg, 2 seg, ... < nt
    const int ns = n_state(p.form), nh = n_hist(p.form);
    const size_t st = (size_t)r.nanch * ns * field_floats(c) + kSpare;  // restart states, then band and frame
    const size_t fl = st + (size_t)p.nt * (ns * r.rg.nband + nh * r.rg.nframe);
    if ((rc = reserve(c, &c->ckpt.data, c->ckpt.cap_data, std::max<size_t>(fl, 1) * sizeof(float)))) return rc;
    if ((rc = store_reserve(c, c->hist, kImgK, nh, false))) return rc;
    if (interior)  // the reversed sweeps store inside R only: the rest images zero
      HIPCALL(c, hipMemsetAsync(c->hist.data, 0, ((size_t)kImgK * nh * field_floats(c) + kSpare) * sizeof(float),
                                c->stream));
    r.band = (float *)c->ckpt.data + st;
    r.frame = r.band + (size_t)p.nt * ns * r.rg.nband;
  } else if (r.history) {
    r.seg = p.nt;
    r.ncp = 1;
    if ((rc = store_reserve(c, c->hist, p.nt, n_hist(p.form), r.f16))) return rc;
  } else {
    if ((rc = store_reserve(c, c->ckpt, r.ncp, n_state(p.form), r.ck16))) return rc;
    if ((rc = store_reserve(c, c->hist, r.seg, n_hist(p.form), r.f16))) return rc;
  }
  if ((rc = fdmod_reset(c))) return rc;
  r.sa = sweep_args(c);
  r.rsa = r.sa;
  r.rsa.g.dt = -r.sa.g.dt;
  r.rsa.coef2 = fld(c, c->d_coef2n, 0);
  r.rsa.rev = 1;
  if (r.recon && r.rg.interior) {
    r.rsa.hskip[0] = r.rg.rx0;
    r.rsa.hskip[1] = r.rg.rx1;
    r.rsa.hskip[2] = r.rg.rz0;
    r.rsa.hskip[3] = r.rg.rz1;
    r.rsa.hinside = 1;
  }
  return zero_adjoint(c);
}

float *const *state_fields(fdmod_ctx *c, float **f)
{
  for (int i = 0; i < n_state(c->p.form); i++) f[i] = c->d.a[kStateIds[i]];
  return f;
}

float *const *hist_fields(fdmod_ctx *c, size_t slot, float **f)
{
  const int nh = n_hist(c->p.form);
  for (int i = 0; i < nh; i++) f[i] = fld(c, (float *)c->hist.data, slot * nh + i);
  return f;
}

// Reconstruction, forward step k: the loop iteration with the frame's history stored
// into slot 0 (nothing inside R), then the frame history and the band state saved.
int recon_forward(GradRun &r, int k, int it)
{
  fdmod_ctx *c = r.c;
  const int ns = n_state(c->p.form), nh = n_hist(c->p.form);
  int rc;
  if (r.nanch > 0 && k > 0 && k % r.seg == 0 && (rc = state_io(c, k / r.seg - 1, false, true))) return rc;
  float *f[kMaxState];
  if (r.rg.nframe == 0) {  // interior variant without a free surface: no history to keep
    if ((rc = step(c, it, r.sa, true))) return rc;
  } else {
    SweepArgs sh = with_hist(c, r.sa, 0, false);
    if (r.rg.interior) {  // the surface row
      sh.hskip[0] = r.rg.rx0;
      sh.hskip[1] = r.rg.rx1;
      sh.hskip[2] = c->g.izStart;
      sh.hskip[3] = r.rg.rz0;
      sh.hinside = 1;
    } else {
      sh.hskip[0] = r.rg.rx0;
      sh.hskip[1] = r.rg.rx1;
      sh.hskip[2] = r.rg.rz0;
      sh.hskip[3] = r.rg.rz1;
    }
    if ((rc = step(c, it, sh, true))) return rc;
    frame_io(c, r.rg.frame, hist_fields(c, 0, f), nh, r.frame + (size_t)k * nh * r.rg.nframe, true);
  }
  frame_io(c, r.rg.band, state_fields(c, f), ns, r.band + (size_t)k * ns * r.rg.nband, true);
  return FDMOD_OK;
}

// Reconstruction, backward step k: undo loop iteration k, from (v_{k+1/2},
// sigma_{k+1}) to (v_{k-1/2}, sigma_k), storing its history into slot `slot`.
// The reversed stress sweep recomputes the strain of v_{k+1/2}; the tensor form's
// history takes the sweep's input t, the forward's t before the stress source.
int recon_backward(GradRun &r, int k, int slot)
{
  fdmod_ctx *c = r.c;
  const fdmod_params &p = c->p;
  hipStream_t s = c->stream;
  const int ns = n_state(p.form), nh = n_hist(p.form);
  const int it = p.mode == 1 ? k : p.nt - 1 - k;
  const SweepArgs rs = with_hist(c, r.rsa, slot, false);
  SourceArgs src = source_args(c, it);
  src.src = c->d_srcn;
  if (c->ns > 0 && c->stype == 1) {
    src.fa = rs.sxx;
    src.fb = rs.szz;
    src.fc = p.form == FDMOD_FORM_TENSOR ? rs.txx : nullptr;
    src.fd = p.form == FDMOD_FORM_TENSOR ? rs.tzz : nullptr;
    launch_source(s, src);
  }
  launch_ssweep(s, p.form, rs);
  float *f[kMaxState];
  state_fields(c, f);
  if (k > 0) {
    float *band = r.band + (size_t)(k - 1) * ns * r.rg.nband;  // saved by forward step k - 1
    frame_io(c, r.rg.band, f + 2, ns - 2, band + 2 * r.rg.nband, false);  // stresses at step k
    if (c->ns > 0 && c->stype != 1) {
      src.fa = c->stype == 6 ? rs.vz : rs.vx;
      src.fb = (c->stype == 5 || c->stype == 6) ? nullptr : rs.vz;
      src.fc = src.fd = nullptr;
      launch_source(s, src);
    }
    launch_vsweep(s, p.form, rs);
    frame_io(c, r.rg.band, f, 2, band, false);  // velocities at k - 1/2
  } else if (p.form == FDMOD_FORM_COORD) {
    // the velocity-sweep history of step 0 is that of the zero initial stress
    float *h[kMaxState];
    hist_fields(c, slot, h);
    for (int i = 4; i < 8; i++) HIPCALL(c, hipMemsetAsync(h[i], 0, field_floats(c) * sizeof(float), s));
  }
  float *h[kMaxState];
  frame_io(c, r.rg.frame, hist_fields(c, slot, h), nh, r.frame + (size_t)k * nh * r.rg.nframe, false);
  if (r.nanch > 0 && k > 0 && k % r.seg == 0) return state_io(c, k / r.seg - 1, false, false);
  return FDMOD_OK;
}

// Forward step k from rest: checkpoint or history store, then the loop iteration.
int grad_forward(GradRun &r, int k)
{
  fdmod_ctx *c = r.c;
  const fdmod_params &p = c->p;
  int rc;
  const int it = p.mode == 1 ? k : p.nt - 1 - k;
  if (r.recon) return recon_forward(r, k, it);
  if (!r.history && k % r.seg == 0 && (rc = state_io(c, k / r.seg, r.ck16, true))) return rc;
  if (r.history) return step(c, it, with_hist(c, r.sa, k, r.f16), true);
  return step(c, it, r.sa, true);
}

// Residuals (in place over the uploaded observations); J accumulates on the device.
int grad_residual(GradRun &r)
{
  fdmod_ctx *c = r.c;
  const size_t nrec = (size_t)c->nr * c->ntsample;
  HIPCALL(c, hipMemsetAsync(c->d_red + kRedBlocks, 0, sizeof(double), c->stream));
  if (r.ds.vx) {
    launch_residual(c->stream, c->d_rvx, c->d_res, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
    launch_residual(c->stream, c->d_rvz, c->d_res + nrec, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  }
  if (r.ds.p)
    launch_residual(c->stream, c->d_rp, c->d_res + 2 * nrec, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  r.sgi = r.ncp - 1;
  r.phase = r.history ? 1 : 0;
  r.k = r.history ? r.c->p.nt - 1 : r.sgi * r.seg;
  if (r.recon) {
    r.phase = 2;
    r.k = r.c->p.nt - 1;
  }
  return FDMOD_OK;
}

// One unit of the backward pass: a recomputed forward step (checkpoint mode) or an
// adjoint step.  Segments run last first; the imaging runs once per kImgK adjoint
// steps and at each segment's end, before the next recomputation reuses the slots.
int grad_backward_next(GradRun &r)
{
  fdmod_ctx *c = r.c;
  const fdmod_params &p = c->p;
  const int k0 = r.sgi * r.seg, k1 = std::min(p.nt, k0 + r.seg);
  int rc;
  if (r.phase == 2) {  // reconstruction: reversed iteration, then the adjoint step
    const int j = r.blk.nk;
    if ((rc = recon_backward(r, r.k, j))) return rc;
    adj_step(c, r.k, r.ds, j, nullptr);
    add_to_block(c, r.blk, j, j, false);
    if (++r.blk.nk == kImgK || r.k == 0) {
      launch_image(c->stream, adj_args(c, false), r.blk);
      r.blk = ImgBlock{};
    }
    if (r.k > 0) r.k--;
    else r.done = true;
    return FDMOD_OK;
  }
  if (r.phase == 0) {
    if (r.k == k0 && (rc = state_io(c, r.sgi, r.ck16, false))) return rc;
    if ((rc = step(c, p.mode == 1 ? r.k : p.nt - 1 - r.k, with_hist(c, r.sa, r.k - k0, r.f16), false)))
      return rc;
    if (++r.k == k1) {
      r.phase = 1;
      r.k = k1 - 1;
    }
    return FDMOD_OK;
  }
  adj_step(c, r.k, r.ds, r.blk.nk, nullptr);
  add_to_block(c, r.blk, r.blk.nk, r.k - k0, r.f16);
  if (++r.blk.nk == kImgK || r.k == k0) {
    launch_image(c->stream, adj_args(c, false), r.blk);
    r.blk = ImgBlock{};
  }
  if (r.k > k0) {
    r.k--;
  } else if (r.sgi > 0) {
    r.sgi--;
    r.phase = 0;
    r.k = r.sgi * r.seg;
  } else {
    r.done = true;
  }
  return FDMOD_OK;
}

// Exchange a context's gradient memory with a pool entry.
void swap_grad_mem(fdmod_ctx *c, GradMem &m)
{
  std::swap(c->d_mu, m.mu);
  std::swap(c->d_mus, m.mus);
  std::swap(c->d_G, m.G);
  std::swap(c->d_snap, m.snap);
  std::swap(c->d_mu_n, m.n);
  std::swap(c->ckpt, m.ckpt);
  std::swap(c->hist, m.hist);
}

size_t held_bytes(const GradMem &m, size_t n)
{
  size_t b = m.ckpt.cap_data + m.ckpt.cap_scale + m.hist.cap_data + m.hist.cap_scale;
  if (m.mu) b += n * (kMaxState * 4 + kMusSlots * 3 * 4 + kNG * 8 + kImgK * 6 * 4);
  return b;
}

// Device bytes one shot's gradient needs beyond its modelling arrays.
size_t grad_bytes(const fdmod_ctx *c, const fdmod_grad_opts &o)
{
  const size_t n = field_floats(c), nt = c->p.nt;
  const size_t esz = o.storage == FDMOD_STORE_F32 ? 4 : 2, cksz = o.storage == FDMOD_STORE_F16_ALL ? 2 : 4;
  const int form = c->p.form;
  size_t b = n * (kMaxState * 4 + kMusSlots * 3 * 4 + kNG * 8 + kImgK * 6 * 4);
  Recon rg;
  if (o.history >= FDMOD_HIST_RECONSTRUCT && recon_shape(c, rg, o.history == FDMOD_HIST_RECONSTRUCT_INTERIOR)) {
    const size_t nanch = o.ncheckpoints > 0 ? std::min<size_t>(o.ncheckpoints, nt) : 0;
    b += (kImgK * n_hist(form) + nanch * n_state(form)) * n * 4 +
         nt * (n_state(form) * rg.nband + n_hist(form) * rg.nframe) * 4 + n * 4;
  } else if (o.history == FDMOD_HIST_FULL) {
    b += nt * n_hist(form) * n * esz;
  } else {
    int ncp = o.ncheckpoints > 0 ? o.ncheckpoints : (int)ceil(sqrt((double)nt));
    ncp = std::min(std::max(ncp, 1), (int)nt);
    const size_t seg = (nt + ncp - 1) / ncp;
    b += (nt + seg - 1) / seg * n_state(form) * n * cksz + seg * n_hist(form) * n * esz;
  }
  return b + b / 16;  // block scales and slack
}

// Chain rule and unpadding into c->d_grad ([4][nx][nz], device).
int grad_finish(GradRun &r)
{
  fdmod_ctx *c = r.c;
  const fdmod_params &p = c->p;
  if (!c->d_grad) HIPCALL(c, hipMalloc(&c->d_grad, 4 * (size_t)p.nx * p.nz * sizeof(double)));
  const int fs_row = p.freesurface && p.form == FDMOD_FORM_BOND ? c->g.e : -1;
  if (p.form == FDMOD_FORM_COORD) launch_fold_trig(c->stream, c->g, fld(c, c->d_G, 0), field_floats(c));
  launch_chain(c->stream, c->g, p.form, p.nx, p.nz, fs_row, c->d, fld(c, c->d_G, 0), field_floats(c),
               (double)p.scalar / 180.0 * kPI, c->d_grad);
  HIPCALL(c, hipGetLastError());
  return FDMOD_OK;
}

}  // namespace

extern "C" int fdmod_gradient_opts(fdmod_ctx *c, const float *obs_vx, const float *obs_vz, const float *obs_p,
                                   const fdmod_grad_opts *opts, double *misfit, double *grad)
{
  if (!c) return FDMOD_EINVAL;
  if (!misfit || !grad) return fail(c, FDMOD_EINVAL, "fdmod_gradient: misfit and grad must not be NULL");
  GradRun r;
  int rc = grad_begin(r, c, obs_vx, obs_vz, obs_p, opts);
  if (rc) return rc;
  for (int k = 0; k < c->p.nt; k++)
    if ((rc = grad_forward(r, k))) return rc;
  if ((rc = grad_residual(r))) return rc;
  while (!r.done)
    if ((rc = grad_backward_next(r))) return rc;
  if ((rc = grad_finish(r))) return rc;
  HIPCALL(c, hipMemcpyAsync(misfit, c->d_red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipMemcpyAsync(grad, c->d_grad, 4 * (size_t)c->p.nx * c->p.nz * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipGetLastError());
  return FDMOD_OK;
}

extern "C" int fdmod_gradient(fdmod_ctx *c, const float *obs_vx, const float *obs_vz, const float *obs_p,
                              int ncheckpoints, double *misfit, double *grad)
{
  fdmod_grad_opts o{ncheckpoints, FDMOD_STORE_F32, 0};
  return fdmod_gradient_opts(c, obs_vx, obs_vz, obs_p, &o, misfit, grad);
}

extern "C" int fdmod_adjoint_source(fdmod_ctx *c, const float *y_vx, const float *y_vz, const float *y_p,
                                    double *out)
{
  if (!c) return FDMOD_EINVAL;
  if (!out) return fail(c, FDMOD_EINVAL, "fdmod_adjoint_source: out must not be NULL");
  int rc = check_ready(c, "fdmod_adjoint_source");
  if (rc) return rc;
  if (c->ns == 0) return fail(c, FDMOD_ESTATE, "fdmod_adjoint_source needs a source");
  if ((rc = workspace(c, false))) return rc;
  DataSrc ds;
  if ((rc = upload_data(c, y_vx, y_vz, y_p, ds))) return rc;
  const size_t nsrc = (size_t)c->ns * c->p.nt;
  dfree(c->d_srcadj);
  c->d_srcadj = nullptr;
  HIPCALL(c, hipMalloc(&c->d_srcadj, nsrc * sizeof(double)));
  HIPCALL(c, hipMemsetAsync(c->d_srcadj, 0, nsrc * sizeof(double), c->stream));
  if ((rc = adjoint_only(c, ds, c->d_srcadj))) return rc;
  HIPCALL(c, hipMemcpy(out, c->d_srcadj, nsrc * sizeof(double), hipMemcpyDeviceToHost));
  return FDMOD_OK;
}

extern "C" int fdmod_adjoint_dot(fdmod_ctx *c, const float *y_vx, const float *y_vz, const float *y_p, double *lhs,
                                 double *rhs)
{
  if (!c) return FDMOD_EINVAL;
  if (!lhs || !rhs) return fail(c, FDMOD_EINVAL, "fdmod_adjoint_dot: lhs and rhs must not be NULL");
  int rc = check_ready(c, "fdmod_adjoint_dot");
  if (rc) return rc;
  if (c->ns == 0) return fail(c, FDMOD_ESTATE, "fdmod_adjoint_dot needs a source");
  if ((rc = workspace(c, false))) return rc;
  DataSrc ds;
  if ((rc = upload_data(c, y_vx, y_vz, y_p, ds))) return rc;
  // <F s, y>: forward from rest with the current source
  if ((rc = fdmod_reset(c))) return rc;
  if ((rc = fdmod_run(c, 0, c->p.nt, nullptr))) return rc;
  const size_t nrec = (size_t)c->nr * c->ntsample;
  HIPCALL(c, hipMemsetAsync(c->d_red + kRedBlocks, 0, sizeof(double), c->stream));
  if (ds.vx) {
    launch_dot(c->stream, c->d_rvx, ds.vx, nullptr, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
    launch_dot(c->stream, c->d_rvz, ds.vz, nullptr, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  }
  if (ds.p) launch_dot(c->stream, c->d_rp, ds.p, nullptr, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  if ((rc = reduce_to_host(c, lhs))) return rc;
  // <s, F^T y>
  const size_t nsrc = (size_t)c->ns * c->p.nt;
  dfree(c->d_srcadj);
  c->d_srcadj = nullptr;
  HIPCALL(c, hipMalloc(&c->d_srcadj, nsrc * sizeof(double)));
  HIPCALL(c, hipMemsetAsync(c->d_srcadj, 0, nsrc * sizeof(double), c->stream));
  if ((rc = adjoint_only(c, ds, c->d_srcadj))) return rc;
  HIPCALL(c, hipMemsetAsync(c->d_red + kRedBlocks, 0, sizeof(double), c->stream));
  launch_dot(c->stream, c->d_src, nullptr, c->d_srcadj, nsrc, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  return reduce_to_host(c, rhs);
}

extern "C" int fdmod_update_model(fdmod_ctx *c, const float *vp, const float *vs, const float *rho, const float *eps,
                                  const float *delta, const float *theta)
{
  if (!c) return FDMOD_EINVAL;
  if (!vp || !vs || !rho || !eps || !delta || !theta) return fail(c, FDMOD_EINVAL, "model array is NULL");
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "fdmod_update_model before fdmod_set_model");
  const fdmod_params &p = c->p;
  double lim = 0.0;
  if (fdmod_check_stability(&p, vp, &lim) == FDMOD_EUNSTABLE)
    return fail(c, FDMOD_EUNSTABLE,
                "the finite differenct scheme is unstable! the time step size must smaller than %5.5f!", lim);
  const size_t n = (size_t)p.nz * p.nx;
  float *stage = nullptr;
  HIPCALL(c, hipMalloc(&stage, 6 * n * sizeof(float)));
  const float *in[6] = {vp, vs, rho, eps, delta, theta};
  for (int k = 0; k < 6; k++) {
    hipError_t e = hipMemcpyAsync(stage + k * n, in[k], n * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      dfree(stage);
      return fail(c, FDMOD_EDEVICE, "model upload: %s", hipGetErrorString(e));
    }
    launch_pad(c->stream, stage + k * n, c->d.a[FDMOD_A_VP + k], p.nz, p.nx, c->g);
  }
  launch_rescale_frozen(c->stream, c->g, c->d.a[FDMOD_A_VP], c->d.a[FDMOD_A_VS], c->d.a[FDMOD_A_RHO],
                        c->d.a[FDMOD_A_BETA]);
  launch_cij(c->stream, c->g, p.form, c->d, p.scalar);
  launch_stagger(c->stream, c->g, p.form, c->d);
  if (p.form == FDMOD_FORM_BOND && p.freesurface == 1) launch_fs_fixup(c->stream, c->g, c->d);
  hipError_t e = hipStreamSynchronize(c->stream);
  dfree(stage);
  if (e != hipSuccess || (e = hipGetLastError()) != hipSuccess)
    return fail(c, FDMOD_EDEVICE, "model update: %s", hipGetErrorString(e));
  return fdmod_reset(c);
}

extern "C" int fdmod_batch_gradient(fdmod_batch *b, const float *const *obs_vx, const float *const *obs_vz,
                                    const float *const *obs_p, const fdmod_grad_opts *opts, const int *x0,
                                    int nx_full, double *misfit, double *grad)
{
  if (!b) return FDMOD_EINVAL;
  if (!opts || !misfit || !grad) return bfail(b, FDMOD_EINVAL, "fdmod_batch_gradient: NULL opts/misfit/grad");
  const int n = (int)b->shots.size(), nz = b->p.nz;
  size_t nout = 0;
  std::vector<size_t> off(n, 0);
  for (int s = 0; s < n; s++) {
    const int nxs = b->shots[s]->p.nx;
    if (x0 && (x0[s] < 0 || x0[s] + nxs > nx_full))
      return bfail(b, FDMOD_EINVAL, "shot %d: crop [%d, %d) outside the model (nx = %d)", s, x0[s], x0[s] + nxs,
                   nx_full);
    off[s] = nout;
    if (!x0) nout += 4 * (size_t)nxs * nz;
  }
  if (x0) nout = 4 * (size_t)nx_full * nz;
  double *d_out = nullptr;
  BHIP(b, hipMalloc(&d_out, nout * sizeof(double)));
  BHIP(b, hipMemsetAsync(d_out, 0, nout * sizeof(double), b->stream));
  // shots run in groups whose gradient storage fits in free device memory
  size_t free_b = 0, total_b = 0;
  BHIP(b, hipMemGetInfo(&free_b, &total_b));
  size_t per = 1, nmax = 1;
  for (fdmod_ctx *c : b->shots) {
    per = std::max(per, grad_bytes(c, *opts));
    nmax = std::max(nmax, field_floats(c));
  }
  for (const GradMem &m : b->gmem) free_b += held_bytes(m, nmax);  // pooled memory is reusable
  int group = (int)std::max<size_t>(1, std::min<size_t>(n, (size_t)(0.85 * free_b) / per));
  if (const char *e = getenv("FDMOD_GRAD_GROUP")) group = std::max(1, std::min(group, atoi(e)));  // tuning cap
  if ((int)b->gmem.size() < group) b->gmem.resize(group);
  double Jsum = 0.0;
  int rc = FDMOD_OK;
  for (int g0 = 0; g0 < n && rc == FDMOD_OK; g0 += group) {
    const int g1 = std::min(n, g0 + group);
    std::vector<GradRun> runs(g1 - g0);
    int bad = -1;
    auto check = [&](int s, int r) {
      if (r && rc == FDMOD_OK) {
        rc = r;
        bad = s;
      }
    };
    for (int s = g0; s < g1; s++) swap_grad_mem(b->shots[s], b->gmem[s - g0]);
    for (int s = g0; s < g1 && rc == FDMOD_OK; s++)
      check(s, grad_begin(runs[s - g0], b->shots[s], obs_vx ? obs_vx[s] : nullptr, obs_vz ? obs_vz[s] : nullptr,
                          obs_p ? obs_p[s] : nullptr, opts));
    // the shots advance in lockstep, each on its own stream
    for (int k = 0; k < b->p.nt && rc == FDMOD_OK; k++)
      for (int s = g0; s < g1; s++) check(s, grad_forward(runs[s - g0], k));
    for (int s = g0; s < g1 && rc == FDMOD_OK; s++) check(s, grad_residual(runs[s - g0]));
    while (rc == FDMOD_OK && !runs[0].done)
      for (int s = g0; s < g1; s++) check(s, grad_backward_next(runs[s - g0]));
    for (int s = g0; s < g1 && rc == FDMOD_OK; s++) check(s, grad_finish(runs[s - g0]));
    for (int s = g0; s < g1 && rc == FDMOD_OK; s++) {
      fdmod_ctx *c = b->shots[s];
      double J = 0.0;
      hipError_t e = hipMemcpyAsync(&J, c->d_red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        rc = bfail(b, FDMOD_EDEVICE, "shot %d: %s", s, hipGetErrorString(e));
        break;
      }
      Jsum += J;
      launch_accum_grad(b->stream, c->d_grad, c->p.nx, nz, x0 ? x0[s] : 0, x0 ? nx_full : c->p.nx,
                        d_out + (x0 ? 0 : off[s]));
      e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) rc = bfail(b, FDMOD_EDEVICE, "gradient sum: %s", hipGetErrorString(e));
    }
    if (bad >= 0) bfail(b, rc, "shot %d: %s", bad, b->shots[bad]->err.c_str());
    for (int s = g0; s < g1; s++) {
      (void)hipStreamSynchronize(b->shots[s]->stream);
      swap_grad_mem(b->shots[s], b->gmem[s - g0]);
    }
  }
  if (rc == FDMOD_OK) {
    hipError_t e = hipMemcpy(grad, d_out, nout * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = bfail(b, FDMOD_EDEVICE, "gradient download: %s", hipGetErrorString(e));
    *misfit = Jsum;
  }
  dfree(d_out);
  return rc;
}
