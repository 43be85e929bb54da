// libfdmod_hip adjoint entry points: the misfit gradient, the adjoint of the
// modelling operator, the device dot-product test and the frozen-absorber model
// update (Bond form).  Declarations are in include/fdmod_hip.h.
//
// Gradient schedule (checkpointed recomputation):
//   1. forward from rest, recording traces; every seg-th state (5 fields) is copied
//      into the checkpoint store;
//   2. residuals r = d - obs and J = sum r^2 / 2, reduced in FP64 on the device;
//   3. for each checkpoint segment, last first: restore its state, recompute its
//      steps storing v_{k+1} of each, then run the adjoint steps of the segment
//      backwards, imaging with the stored velocities;
//   4. chain rule from the six stiffness sums to (vp, eps, delta, theta) and the
//      transpose of the edge padding.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fdmod_ctx.h"

using namespace fdm;

namespace {

constexpr int kRedBlocks = 256;  // partial sums of the FP64 reductions
constexpr int kNState = 5;       // vx, vz, sxx, szz, sxz

size_t field_floats(const fdmod_ctx *c) { return (size_t)c->g.Nx * c->g.ldz; }

const int kStateIds[kNState] = {FDMOD_A_VX, FDMOD_A_VZ, FDMOD_A_SXX, FDMOD_A_SZZ, FDMOD_A_SXZ};

int check_ready(fdmod_ctx *c, const char *what)
{
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "%s before fdmod_set_model", what);
  if (c->p.form != FDMOD_FORM_BOND) return fail(c, FDMOD_EINVAL, "%s: only the Bond form has an adjoint", what);
  if (c->seistype == 0 || c->nr == 0) return fail(c, FDMOD_ESTATE, "%s needs receivers", what);
  return FDMOD_OK;
}

int workspace(fdmod_ctx *c, bool need_G)
{
  const size_t n = field_floats(c);
  if (!c->d_mu) HIPCALL(c, hipMalloc(&c->d_mu, kNState * n * sizeof(float)));
  if (!c->d_red) HIPCALL(c, hipMalloc(&c->d_red, (kRedBlocks + 8) * sizeof(double)));
  if (need_G && !c->d_G) HIPCALL(c, hipMalloc(&c->d_G, 6 * n * sizeof(double)));
  dfree(c->d_res);
  c->d_res = nullptr;
  HIPCALL(c, hipMalloc(&c->d_res, 3 * (size_t)c->nr * c->ntsample * sizeof(float)));
  return FDMOD_OK;
}

int grow(fdmod_ctx *c, float **p, size_t &have, size_t need)
{
  if (have >= need) return FDMOD_OK;
  dfree(*p);
  *p = nullptr;
  have = 0;
  if (hipMalloc(p, need * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, FDMOD_ENOMEM, "cannot allocate %.2f GB of adjoint workspace", need * 4.0 / 1e9);
  }
  have = need;
  return FDMOD_OK;
}

AdjArgs adj_args(fdmod_ctx *c, bool image_rows)
{
  AdjArgs a{};
  a.g = c->g;
  a.izU = c->g.izStart;
  if (image_rows && c->p.freesurface) {
    a.g.izStart -= kN;
    a.g.tilesZ = (a.g.izEnd - a.g.izStart + TZ - 1) / TZ;
  }
  const size_t n = field_floats(c);
  float *const *d = c->d.a;
  a.mvx = c->d_mu; a.mvz = c->d_mu + n; a.mxx = c->d_mu + 2 * n; a.mzz = c->d_mu + 3 * n; a.mxz = c->d_mu + 4 * n;
  a.coef1 = d[FDMOD_A_COEF1]; a.coef2 = d[FDMOD_A_COEF2]; a.invrho = d[FDMOD_A_INVRHO];
  a.c11 = d[FDMOD_A_C11]; a.c13 = d[FDMOD_A_C13]; a.c15 = d[FDMOD_A_C15];
  a.c33 = d[FDMOD_A_C33]; a.c35 = d[FDMOD_A_C35]; a.c55 = d[FDMOD_A_C55];
  for (int j = 0; j < 6; j++) a.G[j] = c->d_G ? c->d_G + j * n : nullptr;
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

// One adjoint step: the transpose of loop iteration k.
void adj_step(fdmod_ctx *c, int k, const DataSrc &ds, const float *fvx, const float *fvz, double *srcadj)
{
  const fdmod_params &p = c->p;
  hipStream_t s = c->stream;
  const int it = p.mode == 1 ? k : p.nt - 1 - k;
  AdjArgs a = adj_args(c, false);
  const bool rec = it % c->mt == 0;
  const size_t so = (size_t)(it / c->mt) * c->nr;
  if (rec && ds.p) launch_adj_inject(s, c->nr, c->g.ldz, c->d_gx, c->d_gz, ds.p + so, ds.p + so, a.mxx, a.mzz, 0.5f);
  if (srcadj && c->ns > 0 && c->stype == 1) launch_adj_source(s, source_args(c, it), a.mxx, a.mzz, srcadj);
  if (p.freesurface) launch_adj_fs(s, a);
  if (fvx) {
    a.fvx = fvx;
    a.fvz = fvz;
    launch_image(s, a);
  }
  launch_adj_ssweep(s, a);
  if (rec && ds.vx)
    launch_adj_inject(s, c->nr, c->g.ldz, c->d_gx, c->d_gz, ds.vx + so, ds.vz + so, a.mvx, a.mvz, 1.0f);
  if (srcadj && c->ns > 0 && c->stype != 1) {
    const float *ma = c->stype == 6 ? a.mvz : a.mvx;
    const float *mb = (c->stype == 5 || c->stype == 6) ? nullptr : a.mvz;
    launch_adj_source(s, source_args(c, it), ma, mb, srcadj);
  }
  launch_adj_vsweep(s, adj_args(c, true));
}

int zero_adjoint(fdmod_ctx *c)
{
  const size_t n = field_floats(c);
  HIPCALL(c, hipMemsetAsync(c->d_mu, 0, kNState * n * sizeof(float), c->stream));
  if (c->d_G) HIPCALL(c, hipMemsetAsync(c->d_G, 0, 6 * n * sizeof(double), c->stream));
  return FDMOD_OK;
}

int copy_state(fdmod_ctx *c, float *store, bool save)
{
  const size_t n = field_floats(c);
  for (int f = 0; f < kNState; f++) {
    float *field = c->d.a[kStateIds[f]];
    float *slot = store + f * n;
    HIPCALL(c, hipMemcpyAsync(save ? slot : field, save ? field : slot, n * sizeof(float), hipMemcpyDeviceToDevice,
                              c->stream));
  }
  return FDMOD_OK;
}

// Run the whole adjoint loop with the given data-space sources (no imaging).
int adjoint_only(fdmod_ctx *c, const DataSrc &ds, double *srcadj)
{
  int rc = zero_adjoint(c);
  if (rc) return rc;
  for (int k = c->p.nt - 1; k >= 0; k--) adj_step(c, k, ds, nullptr, nullptr, srcadj);
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

}  // namespace

extern "C" int fdmod_gradient(fdmod_ctx *c, const float *obs_vx, const float *obs_vz, const float *obs_p,
                              int ncheckpoints, double *misfit, double *grad)
{
  if (!c) return FDMOD_EINVAL;
  if (!misfit || !grad) return fail(c, FDMOD_EINVAL, "fdmod_gradient: misfit and grad must not be NULL");
  int rc = check_ready(c, "fdmod_gradient");
  if (rc) return rc;
  const fdmod_params &p = c->p;
  if ((rc = workspace(c, true))) return rc;
  DataSrc ds;
  if ((rc = upload_data(c, obs_vx, obs_vz, obs_p, ds))) return rc;
  // checkpoint layout: ncp states, one every seg steps
  int ncp = ncheckpoints > 0 ? ncheckpoints : (int)ceil(sqrt((double)p.nt));
  ncp = std::min(std::max(ncp, 1), p.nt);
  const int seg = (p.nt + ncp - 1) / ncp;
  ncp = (p.nt + seg - 1) / seg;
  const size_t n = field_floats(c);
  if ((rc = grow(c, &c->d_ckpt, c->n_ckpt, (size_t)ncp * kNState * n))) return rc;
  if ((rc = grow(c, &c->d_stack, c->n_stack, (size_t)seg * 2 * n))) return rc;
  // 1. forward with checkpoints
  if ((rc = fdmod_reset(c))) return rc;
  SweepArgs sa = sweep_args(c);
  for (int k = 0; k < p.nt; k++) {
    if (k % seg == 0 && (rc = copy_state(c, c->d_ckpt + (size_t)(k / seg) * kNState * n, true))) return rc;
    if ((rc = step(c, p.mode == 1 ? k : p.nt - 1 - k, sa, true))) return rc;
  }
  // 2. residuals (in place over the uploaded observations) and the misfit
  const size_t nrec = (size_t)c->nr * c->ntsample;
  HIPCALL(c, hipMemsetAsync(c->d_red + kRedBlocks, 0, sizeof(double), c->stream));
  if (ds.vx) {
    launch_residual(c->stream, c->d_rvx, c->d_res, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
    launch_residual(c->stream, c->d_rvz, c->d_res + nrec, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  }
  if (ds.p) launch_residual(c->stream, c->d_rp, c->d_res + 2 * nrec, nrec, c->d_red, kRedBlocks, c->d_red + kRedBlocks);
  if ((rc = reduce_to_host(c, misfit))) return rc;
  // 3. segments backwards: recompute, then adjoint with imaging
  if ((rc = zero_adjoint(c))) return rc;
  for (int sgi = ncp - 1; sgi >= 0; sgi--) {
    const int k0 = sgi * seg, k1 = std::min(p.nt, k0 + seg);
    if ((rc = copy_state(c, c->d_ckpt + (size_t)sgi * kNState * n, false))) return rc;
    for (int k = k0; k < k1; k++) {
      if ((rc = step(c, p.mode == 1 ? k : p.nt - 1 - k, sa, false))) return rc;
      float *slot = c->d_stack + (size_t)(k - k0) * 2 * n;
      HIPCALL(c, hipMemcpyAsync(slot, c->d.a[FDMOD_A_VX], n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
      HIPCALL(c, hipMemcpyAsync(slot + n, c->d.a[FDMOD_A_VZ], n * sizeof(float), hipMemcpyDeviceToDevice,
                                c->stream));
    }
    for (int k = k1 - 1; k >= k0; k--) {
      const float *slot = c->d_stack + (size_t)(k - k0) * 2 * n;
      adj_step(c, k, ds, slot, slot + n, nullptr);
    }
  }
  // 4. chain rule and unpadding
  if (!c->d_grad) HIPCALL(c, hipMalloc(&c->d_grad, 4 * (size_t)p.nx * p.nz * sizeof(double)));
  const int fs_row = p.freesurface ? c->g.e : -1;
  launch_chain_bond(c->stream, c->g, p.nx, p.nz, fs_row, c->d, c->d_G, n, (double)p.scalar / 180.0 * kPI, c->d_grad);
  HIPCALL(c, hipMemcpyAsync(grad, c->d_grad, 4 * (size_t)p.nx * p.nz * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipGetLastError());
  return FDMOD_OK;
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
