// libfdmod_hip C-ABI: context management, host-side validation, setup sequence
// and the time-loop driver.  Declarations and the reference code each entry
// point replaces are in include/fdmod_hip.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "fdmod_ctx.h"

using namespace fdm;

int fail(fdmod_ctx *c, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}


void dfree(void *p)
{
  if (p) (void)hipFree(p);
}

// which padded arrays a form needs
static bool form_uses(int form, int a)
{
  switch (a) {
  case FDMOD_A_C15: case FDMOD_A_C35: return form == FDMOD_FORM_BOND;
  case FDMOD_A_SIN2: case FDMOD_A_COS2: case FDMOD_A_SIN2T:
  case FDMOD_A_TXX: case FDMOD_A_TZZ: case FDMOD_A_TXZ: return form == FDMOD_FORM_TENSOR;
  case FDMOD_A_SIN: case FDMOD_A_COS: case FDMOD_A_SINV: case FDMOD_A_COSV: return form == FDMOD_FORM_COORD;
  default: return true;
  }
}

extern "C" const char *fdmod_strerror(int code)
{
  switch (code) {
  case FDMOD_OK: return "ok";
  case FDMOD_EINVAL: return "invalid argument";
  case FDMOD_EUNSTABLE: return "finite-difference scheme unstable for this dt";
  case FDMOD_ERANGE: return "position outside the model";
  case FDMOD_EDEVICE: return "HIP runtime error";
  case FDMOD_ENOMEM: return "out of memory";
  case FDMOD_ESTATE: return "call out of order";
  default: return "unknown error";
  }
}

extern "C" const char *fdmod_last_error(const fdmod_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int fdmod_dims(const fdmod_ctx *c, int *o)
{
  if (!c || !o) return FDMOD_EINVAL;
  o[0] = c->g.Nz; o[1] = c->g.Nx; o[2] = c->g.e; o[3] = c->g.izStart; o[4] = c->g.izEnd;
  o[5] = c->g.ixStart; o[6] = c->g.ixEnd;
  return FDMOD_OK;
}

static void release(fdmod_ctx *c)
{
  dfree(c->pool); dfree(c->flags);
  dfree(c->d_isx); dfree(c->d_isz); dfree(c->d_src); dfree(c->d_taps);
  dfree(c->d_gx); dfree(c->d_gz); dfree(c->d_rvx); dfree(c->d_rvz); dfree(c->d_rp);
  dfree(c->d_rc_start); dfree(c->d_rc_z); dfree(c->d_rc_id); dfree(c->d_fprc);
  dfree(c->d_p); dfree(c->d_E); dfree(c->d_Ek); dfree(c->d_Ep); dfree(c->d_partial);
  dfree(c->d_mu); dfree(c->d_mus); dfree(c->d_G); dfree(c->d_snap); dfree(c->d_res); dfree(c->d_red);
  dfree(c->ckpt.data); dfree(c->ckpt.scale); dfree(c->hist.data); dfree(c->hist.scale);
  dfree(c->d_srcadj); dfree(c->d_grad); dfree(c->d_coef2n); dfree(c->d_srcn);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

extern "C" void fdmod_destroy(fdmod_ctx *c)
{
  if (!c) return;
  release(c);
  delete c;
}

// Velocity-sweep blocks the device holds at once: kVBlocksPerCU per CU.
static long vsweep_slots()
{
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess || cus < 1)
    cus = 256;
  return (long)cus * kVBlocksPerCU;
}

extern "C" int fdmod_create(const fdmod_params *p, fdmod_ctx **out)
{
  if (!p || !out) return FDMOD_EINVAL;
  *out = nullptr;
  if (p->form < 0 || p->form > 2) return FDMOD_EINVAL;
  if (p->nz < 1 || p->nx < 1 || p->nt < 1 || p->ealThick < 0) return FDMOD_EINVAL;
  if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0) || !(p->fpeak > 0)) return FDMOD_EINVAL;
  if (p->freesurface != 0 && p->freesurface != 1) return FDMOD_EINVAL;
  if (p->mode != 1 && p->mode != 2) return FDMOD_EINVAL;
  fdmod_ctx *c = new fdmod_ctx();
  c->p = *p;
  Grid &g = c->g;
  g.e = p->ealThick + kN;  // BT:307-308
  g.Nz = p->nz + 2 * g.e;
  g.Nx = p->nx + 2 * g.e;
  g.ldz = (g.Nz + 63) / 64 * 64;
  g.ixStart = kN;
  g.ixEnd = g.Nx - kN;
  g.izStart = p->freesurface == 0 ? kN : g.e;
  g.izEnd = g.Nz - kN;
  g.tilesZ = (g.izEnd - g.izStart + TZ - 1) / TZ;
  // Velocity-sweep segment: the shortest (most blocks, most waves per SIMD) whose blocks
  // still all run in one generation -- kVBlocksPerCU per CU, which its LDS ring allows --
  // between kFlagW and kDefaultSegLV columns.  On 1256 x 8002 (256 CUs) that is 38
  // columns, 4009 blocks: Bond 57.4-58.3 Gpts/s against 56.6-57.6 with 44 columns.
  {
    const long slots = vsweep_slots();
    const long want = ((long)g.tilesZ * (g.ixEnd - g.ixStart) + slots - 1) / slots;
    g.segLV = (int)std::min<long>(kDefaultSegLV, std::max<long>(kFlagW, want));
  }
  g.segLS = default_segLS(p->form, g.ixEnd - g.ixStart);
  if (const char *sl = getenv("FDMOD_SEGLV")) {  // tile-shape overrides (tuning experiments)
    int v = atoi(sl);
    if (v >= 1 && v <= 4096) g.segLV = v;
  }
  if (const char *sl = getenv("FDMOD_SEGLS")) {
    int v = atoi(sl);
    if (v >= 1 && v <= 4096) g.segLS = v;
  }
  g.flagsX = (g.ixEnd - g.ixStart + kFlagW - 1) / kFlagW;
  g.invdx = (float)(1.0 / p->dx);
  g.invdz = (float)(1.0 / p->dz);
  g.dt = p->dt;
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
    fdmod_destroy(c);
    return FDMOD_EDEVICE;
  }
  int nused = 0;
  for (int a = 0; a < FDMOD_A_COUNT; a++) nused += form_uses(p->form, a);
  // Every array starts kZAlign - (izStart mod kZAlign) floats into its slot, so that row
  // izStart -- the first row of every 64-row strip -- begins a 128-byte line: a wave's
  // 64-row access then covers two whole lines instead of straddling three that the
  // neighbouring strips also fetch.  The slot pitch stays Nx*ldz floats; the pool keeps
  // kZAlign spare floats at its end for the shift of the last slot.
  c->zsh = (kZAlign - g.izStart % kZAlign) % kZAlign;
  c->zsh16 = (2 * kZAlign - g.izStart % (2 * kZAlign)) % (2 * kZAlign);
  if (const char *za = getenv("FDMOD_ZALIGN"))  // 0 = no shift (tuning experiments)
    if (atoi(za) == 0) c->zsh = c->zsh16 = 0;
  size_t per = (size_t)g.Nx * g.ldz;
  const size_t pool_n = per * nused + kZAlign;
  if (hipMalloc(&c->pool, pool_n * sizeof(float)) != hipSuccess) {
    fdmod_destroy(c);
    return FDMOD_ENOMEM;
  }
  int k = 0;
  for (int a = 0; a < FDMOD_A_COUNT; a++)
    c->d.a[a] = form_uses(p->form, a) ? c->pool + c->zsh + per * (k++) : nullptr;
  if (hipMemsetAsync(c->pool, 0, pool_n * sizeof(float), c->stream) != hipSuccess ||
      hipMalloc(&c->flags, (size_t)g.tilesZ * g.flagsX + 1) != hipSuccess) {
    fdmod_destroy(c);
    return FDMOD_ENOMEM;
  }
  *out = c;
  return FDMOD_OK;
}

// staggered_grid_stable2d (L/fdmod/stable.c:217-242) on the padded pre-MEAL vp;
// the padded maximum equals the maximum of the unpadded input.
static double stability_limit(const fdmod_params &p, const float *vp)
{
  static const float C[5] = {kC1, kC2, kC3, kC4, kC5};
  double vp_max = __FLT_MIN__;
  size_t n = (size_t)p.nz * p.nx;
  for (size_t i = 0; i < n; i++) vp_max = fmaxf(vp[i], (float)vp_max);
  double C_sum = 0.0f;
  for (int i = 0; i < kN; i += 2) C_sum += C[i] * pow(-1.0, 1.0 * i);
  float dx = p.dx, dz = p.dz;
  return 1.0f / (vp_max * sqrt((1.0 / (dx * dx) + 1.0 / (dz * dz)) * C_sum));
}

extern "C" int fdmod_check_stability(const fdmod_params *p, const float *vp, double *dt_max)
{
  if (!p || !vp || p->nz < 1 || p->nx < 1 || !(p->dx > 0) || !(p->dz > 0)) return FDMOD_EINVAL;
  double lim = stability_limit(*p, vp);
  if (dt_max) *dt_max = lim;
  return p->dt > lim ? FDMOD_EUNSTABLE : FDMOD_OK;
}

extern "C" int fdmod_set_model(fdmod_ctx *c, const float *vp, const float *vs, const float *rho, const float *eps,
                               const float *delta, const float *theta, double *dt_max)
{
  if (!c) return FDMOD_EINVAL;
  if (!vp || !vs || !rho || !eps || !delta || !theta) return fail(c, FDMOD_EINVAL, "model array is NULL");
  const fdmod_params &p = c->p;
  Grid &g = c->g;
  double lim = stability_limit(p, vp);
  if (dt_max) *dt_max = lim;
  if (p.dt > lim)  // stable.c:236-241
    return fail(c, FDMOD_EUNSTABLE,
                "the finite differenct scheme is unstable! the time step size must smaller than %5.5f!", lim);
  size_t n = (size_t)p.nz * p.nx;
  float *stage = nullptr;
  HIPCALL(c, hipMalloc(&stage, 6 * n * sizeof(float) + 256 * sizeof(float)));
  const float *in[6] = {vp, vs, rho, eps, delta, theta};
  for (int k = 0; k < 6; k++) {
    hipError_t e = hipMemcpyAsync(stage + k * n, in[k], n * sizeof(float), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      dfree(stage);
      return fail(c, FDMOD_EDEVICE, "model upload: %s", hipGetErrorString(e));
    }
    launch_pad(c->stream, stage + k * n, c->d.a[FDMOD_A_VP + k], p.nz, p.nx, g);
  }
  // MEAL constants on the host, in the reference's float order (meal.c:63-77, 23-29)
  int nlayer = g.e;
  float alpha = p.eal_alpha == -1.0f ? 2.5f : p.eal_alpha;
  float pd = 2;
  float Rcoef = powf(10.0f, -1.0f * ((log10f((float)nlayer) - 1.0f) / (log10f(2.0f))) - 3.0f - 0.5f * pd);
  std::vector<float> wtab(nlayer + 1, 0.0f);
  for (int l = 1; l <= nlayer; l++) {
    float d = (0.1f * l) / (0.1f * nlayer);
    wtab[l] = expf(logf(2.0f) * powf(d, alpha)) - 1.0f;
  }
  float *d_wtab = stage + 6 * n;
  int rc = FDMOD_OK;
  if (nlayer + 1 > 256) {
    dfree(stage);
    return fail(c, FDMOD_EINVAL, "ealThick too large (%d)", p.ealThick);
  }
  if (hipMemcpyAsync(d_wtab, wtab.data(), wtab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream) !=
      hipSuccess)
    rc = FDMOD_EDEVICE;
  float k0 = -1.0f * (1.0f + pd);
  float thx = nlayer * p.dx, thz = nlayer * p.dz;
  if (rc == FDMOD_OK) {
    launch_meal(c->stream, g, c->d.a[FDMOD_A_VP], c->d.a[FDMOD_A_VS], c->d.a[FDMOD_A_RHO], c->d.a[FDMOD_A_D],
                c->d.a[FDMOD_A_BETA], c->d.a[FDMOD_A_COEF1], c->d.a[FDMOD_A_COEF2], d_wtab, nlayer, p.freesurface,
                k0, logf(Rcoef), 2.0f * thx, 2.0f * thz, p.dx * p.fpeak, p.dz * p.fpeak, sqrtf(3.0f), (double)p.dt);
    launch_cij(c->stream, g, p.form, c->d, p.scalar);
    launch_stagger(c->stream, g, p.form, c->d);
    if (p.form == FDMOD_FORM_BOND && p.freesurface == 1) launch_fs_fixup(c->stream, g, c->d);
    launch_tile_flags(c->stream, g, c->d.a[FDMOD_A_COEF1], c->d.a[FDMOD_A_COEF2], c->flags);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  dfree(stage);
  if (rc != FDMOD_OK || e != hipSuccess || (e = hipGetLastError()) != hipSuccess)
    return fail(c, FDMOD_EDEVICE, "model setup: %s", hipGetErrorString(e));
  c->model_set = true;
  return fdmod_reset(c);
}

extern "C" int fdmod_reset(fdmod_ctx *c)
{
  if (!c) return FDMOD_EINVAL;
  size_t per = (size_t)c->g.Nx * c->g.ldz * sizeof(float);
  for (int a = FDMOD_A_VX; a <= FDMOD_A_TXZ; a++)
    if (c->d.a[a]) HIPCALL(c, hipMemsetAsync(c->d.a[a], 0, per, c->stream));
  size_t nrec = (size_t)c->nr * c->ntsample * sizeof(float);
  if (c->d_rvx) HIPCALL(c, hipMemsetAsync(c->d_rvx, 0, nrec, c->stream));
  if (c->d_rvz) HIPCALL(c, hipMemsetAsync(c->d_rvz, 0, nrec, c->stream));
  if (c->d_rp) HIPCALL(c, hipMemsetAsync(c->d_rp, 0, nrec, c->stream));
  if (c->d_E) {
    HIPCALL(c, hipMemsetAsync(c->d_E, 0, c->p.nt * sizeof(float), c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_Ek, 0, c->p.nt * sizeof(float), c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_Ep, 0, c->p.nt * sizeof(float), c->stream));
  }
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return FDMOD_OK;
}

// Source taps in the reference's float order (BT:706-834, 893-914):
// alpha = logf(0.8)/(dx^2+dz^2), w = expf(alpha*(rx^2+rz^2))*invdx*invdz*dt.
static void build_taps(const fdmod_params &p, int stype, std::vector<SrcTap> &taps, int &ntap, int &npass)
{
  float dx = p.dx, dz = p.dz, dt = p.dt;
  float invdx = (float)(1.0 / dx), invdz = (float)(1.0 / dz);
  float alpha = logf(0.8f) / (dx * dx + dz * dz);
  taps.clear();
  if (stype == 8) {  // BT:814-833
    float q = 0.125f * sqrtf(2.0f);
    taps = {{0, 0, -q, q}, {-1, 0, -q, -q}, {-1, -1, q, -q}, {0, -1, q, q}};
    ntap = 4;
    npass = 1;
    return;
  }
  ntap = (2 * kN + 1) * (2 * kN + 1);
  npass = stype == 4 ? 2 : 1;
  for (int pass = 0; pass < npass; pass++) {
    int kind = stype == 4 ? (pass == 0 ? 2 : 3) : stype;
    for (int ox = -kN; ox <= kN; ox++) {
      float rx = ox * dx;
      for (int oz = -kN; oz <= kN; oz++) {
        float rz = oz * dz;
        float w = expf(alpha * (rx * rx + rz * rz)) * invdx * invdz * dt;
        SrcTap t{ox, oz, 0.0f, 0.0f};
        switch (kind) {
        case 1: t.ca = 0.5f * w; t.cb = 0.5f * w; break;
        case 2: t.ca = 2.0f * alpha * rx * w; t.cb = 2.0f * alpha * rz * w; break;
        case 3: t.ca = -(2.0f * alpha * rz * w); t.cb = 2.0f * alpha * rx * w; break;
        case 5: case 6: t.ca = w; break;
        case 7: t.ca = 0.5f * sqrtf(2.0f) * w; t.cb = 0.5f * sqrtf(2.0f) * w; break;
        }
        taps.push_back(t);
      }
    }
  }
}

// The receivers inside a single stype-1 source's footprint (the cells its taps write),
// for the combined source + receiver launch; nfprc = -1 where that launch does not apply.
static int fp_update(fdmod_ctx *c)
{
  dfree(c->d_fprc);
  c->d_fprc = nullptr;
  c->nfprc = -1;
  if (c->ns != 1 || c->stype != 1 || c->seistype == 0 || c->nr == 0 || c->h_gx.size() != (size_t)c->nr ||
      c->h_isx.size() != 1)
    return FDMOD_OK;
  std::vector<int> in;
  for (int i = 0; i < c->nr; i++)
    if (abs(c->h_gx[i] - c->h_isx[0]) <= kN && abs(c->h_gz[i] - c->h_isz[0]) <= kN) in.push_back(i);
  if (!in.empty()) {
    HIPCALL(c, hipMalloc(&c->d_fprc, in.size() * sizeof(int)));
    HIPCALL(c, hipMemcpy(c->d_fprc, in.data(), in.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  c->nfprc = (int)in.size();
  return FDMOD_OK;
}

extern "C" int fdmod_set_source(fdmod_ctx *c, int stype, int ns, const int *posx, const int *posz, const float *src)
{
  if (!c) return FDMOD_EINVAL;
  if (stype < 1 || stype > 8) return fail(c, FDMOD_EINVAL, "stype must be 1..8 (got %d)", stype);
  if (ns < 0 || (ns > 0 && (!posx || !posz || !src))) return fail(c, FDMOD_EINVAL, "bad source arrays");
  const fdmod_params &p = c->p;
  for (int i = 0; i < ns; i++) {  // obsinfo.c:69-80
    if (posz[i] < 0 || posz[i] > p.nz)
      return fail(c, FDMOD_ERANGE, "the position of source in 1st dimension (%d) must be between 0 and %d!", posz[i],
                  p.nz);
    if (posx[i] < 0 || posx[i] > p.nx)
      return fail(c, FDMOD_ERANGE, "the position of source in 2nd dimension (%d) must be between 0 and %d!", posx[i],
                  p.nx);
  }
  std::vector<SrcTap> taps;
  build_taps(p, stype, taps, c->ntap, c->npass);
  std::vector<int> isx(ns), isz(ns);
  for (int i = 0; i < ns; i++) {
    isx[i] = posx[i] + c->g.e;
    isz[i] = posz[i] + c->g.e;
  }
  c->h_isx = isx;
  c->h_isz = isz;
  c->nfprc = -1;
  // sources may run in parallel only if no two footprints share a cell
  int rad = stype == 8 ? 1 : kN;
  std::vector<int> order(ns);
  for (int i = 0; i < ns; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return isx[a] < isx[b]; });
  bool overlap = false;
  for (int a = 0; a < ns && !overlap; a++)
    for (int b = a + 1; b < ns && isx[order[b]] - isx[order[a]] <= 2 * rad; b++)
      if (abs(isz[order[b]] - isz[order[a]]) <= 2 * rad) {
        overlap = true;
        break;
      }
  c->seq = overlap ? 1 : 0;
  dfree(c->d_isx); dfree(c->d_isz); dfree(c->d_src); dfree(c->d_taps);
  c->d_isx = c->d_isz = nullptr; c->d_src = nullptr; c->d_taps = nullptr;
  c->stype = stype;
  c->ns = ns;
  // the stress sweep can add a single explosive source whose footprint it writes:
  // rows down from the image rows (free surface) or from izStart, columns of U
  {
    const Grid &g = c->g;
    const int top = g.izStart - (p.freesurface ? kN : 0);
    c->fuse_src = ns == 1 && stype == 1 && p.form != FDMOD_FORM_TENSOR && isx[0] - kN >= g.ixStart &&
                  isx[0] + kN < g.ixEnd && isz[0] - kN >= top && isz[0] + kN < g.izEnd;
    if (c->fuse_src) {
      c->src_cell_x = isx[0];
      c->src_cell_z = isz[0];
    }
  }
  if (ns == 0) return FDMOD_OK;
  HIPCALL(c, hipMalloc(&c->d_isx, ns * sizeof(int)));
  HIPCALL(c, hipMalloc(&c->d_isz, ns * sizeof(int)));
  HIPCALL(c, hipMalloc(&c->d_src, (size_t)ns * p.nt * sizeof(float)));
  HIPCALL(c, hipMalloc(&c->d_taps, taps.size() * sizeof(SrcTap)));
  HIPCALL(c, hipMemcpy(c->d_isx, isx.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  HIPCALL(c, hipMemcpy(c->d_isz, isz.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  HIPCALL(c, hipMemcpy(c->d_src, src, (size_t)ns * p.nt * sizeof(float), hipMemcpyHostToDevice));
  HIPCALL(c, hipMemcpy(c->d_taps, taps.data(), taps.size() * sizeof(SrcTap), hipMemcpyHostToDevice));
  return fp_update(c);
}

extern "C" int fdmod_set_receivers(fdmod_ctx *c, int seistype, int mt, int nr, const int *posx, const int *posz)
{
  if (!c) return FDMOD_EINVAL;
  if (seistype < 0 || seistype > 3) return fail(c, FDMOD_EINVAL, "seistype must be 0..3 (got %d)", seistype);
  if (mt < 1) return fail(c, FDMOD_EINVAL, "mt must be >= 1");
  if (nr < 0 || (nr > 0 && (!posx || !posz))) return fail(c, FDMOD_EINVAL, "bad receiver arrays");
  const fdmod_params &p = c->p;
  for (int i = 0; i < nr; i++) {  // obsinfo.c:206-217
    if (posz[i] < 0 || posz[i] > p.nz)
      return fail(c, FDMOD_ERANGE, "the position of receiver in 1st dimension (%d) must be between 0 and %d!",
                  posz[i], p.nz);
    if (posx[i] < 0 || posx[i] > p.nx)
      return fail(c, FDMOD_ERANGE, "the position of receiver in 2nd dimension (%d) must be between 0 and %d!",
                  posx[i], p.nx);
  }
  dfree(c->d_gx); dfree(c->d_gz); dfree(c->d_rvx); dfree(c->d_rvz); dfree(c->d_rp);
  dfree(c->d_rc_start); dfree(c->d_rc_z); dfree(c->d_rc_id);
  c->d_gx = c->d_gz = nullptr; c->d_rvx = c->d_rvz = c->d_rp = nullptr;
  c->d_rc_start = c->d_rc_z = c->d_rc_id = nullptr;
  c->seistype = seistype;
  c->mt = mt;
  c->nr = nr;
  c->rline = false;
  c->ntsample = (p.nt - 1) / mt + 1;  // obsinfo.c:199
  c->nfprc = -1;
  c->h_gx.clear();
  c->h_gz.clear();
  if (nr == 0 || seistype == 0) return FDMOD_OK;
  std::vector<int> gx(nr), gz(nr);
  for (int i = 0; i < nr; i++) {
    gx[i] = posx[i] + c->g.e;
    gz[i] = posz[i] + c->g.e;
  }
  c->h_gx = gx;
  c->h_gz = gz;
  size_t nrec = (size_t)nr * c->ntsample * sizeof(float);
  HIPCALL(c, hipMalloc(&c->d_gx, nr * sizeof(int)));
  HIPCALL(c, hipMalloc(&c->d_gz, nr * sizeof(int)));
  HIPCALL(c, hipMemcpy(c->d_gx, gx.data(), nr * sizeof(int), hipMemcpyHostToDevice));
  HIPCALL(c, hipMemcpy(c->d_gz, gz.data(), nr * sizeof(int), hipMemcpyHostToDevice));
  if (seistype == 1 || seistype == 3) {
    HIPCALL(c, hipMalloc(&c->d_rvx, nrec));
    HIPCALL(c, hipMalloc(&c->d_rvz, nrec));
    HIPCALL(c, hipMemsetAsync(c->d_rvx, 0, nrec, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_rvz, 0, nrec, c->stream));
  }
  if (seistype == 2 || seistype == 3) {
    HIPCALL(c, hipMalloc(&c->d_rp, nrec));
    HIPCALL(c, hipMemsetAsync(c->d_rp, 0, nrec, c->stream));
  }
  if (seistype == 2) {
    // one row, consecutive columns in receiver order (each receiver its own cell)
    bool line = true;
    for (int i = 1; i < nr && line; i++) line = gz[i] == gz[0] && gx[i] == gx[0] + i;
    c->rline = line;
    c->rline_x0 = gx[0];
    c->rline_z = gz[0];
  }
  if (seistype == 2) {  // pressure receivers by padded column, for the stress sweep to record
    const int Nx = c->g.Nx;
    std::vector<int> start(Nx + 1, 0), rz(nr), rid(nr);
    for (int i = 0; i < nr; i++) start[gx[i] + 1]++;
    for (int x = 0; x < Nx; x++) start[x + 1] += start[x];
    std::vector<int> fill(start.begin(), start.end() - 1);
    c->rc_zlo = INT32_MAX;
    c->rc_zhi = -1;
    for (int i = 0; i < nr; i++) {
      const int k = fill[gx[i]]++;
      rz[k] = gz[i];
      rid[k] = i;
      c->rc_zlo = std::min(c->rc_zlo, gz[i]);
      c->rc_zhi = std::max(c->rc_zhi, gz[i]);
    }
    HIPCALL(c, hipMalloc(&c->d_rc_start, (Nx + 1) * sizeof(int)));
    HIPCALL(c, hipMalloc(&c->d_rc_z, nr * sizeof(int)));
    HIPCALL(c, hipMalloc(&c->d_rc_id, nr * sizeof(int)));
    HIPCALL(c, hipMemcpy(c->d_rc_start, start.data(), (Nx + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCALL(c, hipMemcpy(c->d_rc_z, rz.data(), nr * sizeof(int), hipMemcpyHostToDevice));
    HIPCALL(c, hipMemcpy(c->d_rc_id, rid.data(), nr * sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return fp_update(c);
}

extern "C" int fdmod_set_snapshots(fdmod_ctx *c, int wsave, int savew, int snap_sample, fdmod_snapshot_fn fn,
                                   void *user)
{
  if (!c) return FDMOD_EINVAL;
  if (wsave < 0 || wsave > 5 || (savew != 0 && savew != 1)) return fail(c, FDMOD_EINVAL, "bad wsave/savew");
  if (wsave > 0 && (snap_sample < 1 || !fn)) return fail(c, FDMOD_EINVAL, "snapshots need snap_sample>=1 and a sink");
  c->wsave = wsave;
  c->savew = savew;
  c->snap_sample = snap_sample < 1 ? 1 : snap_sample;
  c->snap_fn = fn;
  c->snap_user = user;
  if (wsave >= 2 && !c->d_p) HIPCALL(c, hipMalloc(&c->d_p, (size_t)c->g.Nx * c->g.ldz * sizeof(float)));
  return FDMOD_OK;
}

extern "C" int fdmod_set_energy(fdmod_ctx *c, int on)
{
  if (!c) return FDMOD_EINVAL;
  c->energy = on ? 1 : 0;
  if (c->energy && !c->d_E) {
    c->eblocks = 512;
    HIPCALL(c, hipMalloc(&c->d_E, c->p.nt * sizeof(float)));
    HIPCALL(c, hipMalloc(&c->d_Ek, c->p.nt * sizeof(float)));
    HIPCALL(c, hipMalloc(&c->d_Ep, c->p.nt * sizeof(float)));
    HIPCALL(c, hipMalloc(&c->d_partial, 2 * c->eblocks * sizeof(float)));
    HIPCALL(c, hipMemsetAsync(c->d_E, 0, c->p.nt * sizeof(float), c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_Ek, 0, c->p.nt * sizeof(float), c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_Ep, 0, c->p.nt * sizeof(float), c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
  }
  return FDMOD_OK;
}

SweepArgs sweep_args(fdmod_ctx *c)
{
  SweepArgs a{};
  float *const *d = c->d.a;
  a.g = c->g;
  a.vx = d[FDMOD_A_VX]; a.vz = d[FDMOD_A_VZ];
  a.sxx = d[FDMOD_A_SXX]; a.szz = d[FDMOD_A_SZZ]; a.sxz = d[FDMOD_A_SXZ];
  a.txx = d[FDMOD_A_TXX]; a.tzz = d[FDMOD_A_TZZ]; a.txz = d[FDMOD_A_TXZ];
  a.coef1 = d[FDMOD_A_COEF1]; a.coef2 = d[FDMOD_A_COEF2]; a.invrho = d[FDMOD_A_INVRHO];
  a.c11 = d[FDMOD_A_C11]; a.c13 = d[FDMOD_A_C13]; a.c15 = d[FDMOD_A_C15];
  a.c33 = d[FDMOD_A_C33]; a.c35 = d[FDMOD_A_C35]; a.c55 = d[FDMOD_A_C55];
  a.sin2 = d[FDMOD_A_SIN2]; a.cos2 = d[FDMOD_A_COS2]; a.sin2t = d[FDMOD_A_SIN2T];
  a.sn = d[FDMOD_A_SIN]; a.cs = d[FDMOD_A_COS]; a.snv = d[FDMOD_A_SINV]; a.csv = d[FDMOD_A_COSV];
  a.undamped = c->flags;
  a.freesurface = c->p.freesurface;
  // the state and parameter arrays one tensor-form step streams (20 of Nx x ldz floats)
  // against twice the last-level cache: 1256 x 8002 -> non-temporal; a 1126 x 2000 crop
  // (1246 x 2120) -> cached (its gradient ran 1.3% slower with non-temporal tau)
  a.tau_nt = 20.0 * 4.0 * c->g.Nx * c->g.ldz > 512.0 * (1 << 20);
  // The same rule for the per-point operands, counting the arrays each form streams per
  // step (Bond 14, coordinate 16, tensor 20).  Measured on MI355X: the full 1126 x 7872
  // grid (1256 x 8002) streams 573-819 MB and runs faster with non-temporal operand
  // loads (cached, the stress sweep took 9 us more per launch); 1126-deep crops 1500-4000
  // columns wide stream 120-420 MB and ran their forward steps 5-7% faster with cached
  // loads, the 1126 x 2000 gradients 3-6% faster (Bond FP16 history 0.54 -> 0.525 s).
  const int arrays = c->p.form == FDMOD_FORM_TENSOR ? 20 : c->p.form == FDMOD_FORM_COORD ? 16 : 14;
  a.op_nt = arrays * 4.0 * c->g.Nx * c->g.ldz > 512.0 * (1 << 20);
  return a;
}

static int snapshot(fdmod_ctx *c, int it)
{
  const Grid &g = c->g;
  int e = c->savew == 0 ? g.e : 0;
  int n1 = g.Nz - 2 * e, n2 = g.Nx - 2 * e;
  c->h_snap.resize((size_t)n1 * n2);
  auto emit = [&](const float *dev, const char *kind) -> int {
    HIPCALL(c, hipMemcpy2DAsync(c->h_snap.data(), n1 * sizeof(float), dev + (size_t)e * g.ldz + e,
                                g.ldz * sizeof(float), n1 * sizeof(float), n2, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (c->snap_fn(c->snap_user, it, kind, c->h_snap.data(), n1, n2) != 0)
      return fail(c, FDMOD_EINVAL, "snapshot sink failed at it=%d", it);
    return FDMOD_OK;
  };
  int rc = FDMOD_OK;
  int ws = c->wsave;
  // isoelastic.c:249-288: 1 = vx,vz; 2 = p; 3 = both; 4/5 = vx,vz under P/S names
  if (ws == 1 || ws == 3) {
    if ((rc = emit(c->d.a[FDMOD_A_VX], "vx"))) return rc;
    if ((rc = emit(c->d.a[FDMOD_A_VZ], "vz"))) return rc;
  }
  if (ws == 2 || ws == 3) {
    launch_pressure(c->stream, c->d.a[FDMOD_A_SXX], c->d.a[FDMOD_A_SZZ], c->d_p, g.Nx, g.ldz);
    if ((rc = emit(c->d_p, "p"))) return rc;
  }
  if (ws == 4 || ws == 5) {
    if ((rc = emit(c->d.a[FDMOD_A_VX], ws == 4 ? "vxp" : "vxs"))) return rc;
    if ((rc = emit(c->d.a[FDMOD_A_VZ], ws == 4 ? "vzp" : "vzs"))) return rc;
  }
  return rc;
}

// One iteration of the reference loop (BT:651-954): V sweep, velocity sources,
// S sweep (+ free surface), stress source, snapshot, receivers, energy.
int step(fdmod_ctx *c, int it, const SweepArgs &sa, bool record, const hipEvent_t *mark)
{
  const fdmod_params &p = c->p;
  if (mark) HIPCALL(c, hipEventRecord(mark[0], c->stream));
  launch_vsweep(c->stream, p.form, sa);
  SourceArgs src;
  if (c->ns > 0) {
    src.g = c->g;
    src.isx = c->d_isx; src.isz = c->d_isz; src.src = c->d_src;
    src.nt = p.nt; src.ns = c->ns; src.it = it;
    src.taps = c->d_taps; src.ntap = c->ntap; src.npass = c->npass; src.sequential = c->seq;
    src.fc = src.fd = nullptr;
  }
  if (c->ns > 0 && c->stype != 1) {
    src.fa = c->stype == 6 ? sa.vz : sa.vx;
    src.fb = (c->stype == 5 || c->stype == 6) ? nullptr : sa.vz;
    launch_source(c->stream, src);
  }
  if (mark) HIPCALL(c, hipEventRecord(mark[1], c->stream));
  const bool snap_now = record && c->wsave > 0 && it % c->snap_sample == 0;
  const bool rec_now = record && c->seistype > 0 && c->nr > 0 && it % c->mt == 0;
  const bool fsrc = c->fuse_src && c->ns > 0;
  // The receivers sample after the stress source (BT:893-924), so the sweep records them
  // only when it also adds that source, or when there is none.
  const bool src1 = c->ns > 0 && c->stype == 1;
  const bool frec = rec_now && c->d_rc_start && p.form != FDMOD_FORM_TENSOR && (fsrc || !src1);
  // otherwise one launch can add a single stype-1 source and record the receivers
  const bool combine = rec_now && !frec && src1 && !fsrc && c->nfprc >= 0;
  if (fsrc || frec) {  // the stress sweep adds the source and/or records the receivers
    SweepArgs sf = sa;
    if (fsrc) {
      sf.fuse |= 1;
      sf.fsx = c->src_cell_x;
      sf.fsz = c->src_cell_z;
      sf.ftap = c->d_taps;
      sf.fsrc = c->d_src;
      sf.fit = it;
    }
    if (frec) {
      sf.fuse |= 2;
      sf.rc_start = c->d_rc_start;
      sf.rc_z = c->d_rc_z;
      sf.rc_id = c->d_rc_id;
      sf.rc_zlo = c->rc_zlo;
      sf.rc_zhi = c->rc_zhi;
      sf.rec_p = c->d_rp;
      sf.rec_off = (size_t)(it / c->mt) * c->nr;
    }
    launch_ssweep(c->stream, p.form, sf);
  } else {
    launch_ssweep(c->stream, p.form, sa);
  }
  if (mark) HIPCALL(c, hipEventRecord(mark[2], c->stream));
  RecvArgs r{};
  if (rec_now && !frec) {
    r.nr = c->nr; r.ldz = c->g.ldz; r.gx = c->d_gx; r.gz = c->d_gz;
    r.vx = sa.vx; r.vz = sa.vz; r.sxx = sa.sxx; r.szz = sa.szz;
    r.out_vx = c->d_rvx; r.out_vz = c->d_rvz; r.out_p = c->d_rp;
    r.isample = it / c->mt;
    r.want_v = c->seistype == 1 || c->seistype == 3;
    r.want_p = c->seistype == 2 || c->seistype == 3;
    r.mt = c->mt;
  }
  if (src1 && !fsrc) {
    src.fa = sa.sxx;
    src.fb = sa.szz;
    if (p.form == FDMOD_FORM_TENSOR) {  // TR:888-891
      src.fc = sa.txx;
      src.fd = sa.tzz;
    }
    if (combine) launch_source_receivers(c->stream, src, r, c->d_fprc, c->nfprc);
    else launch_source(c->stream, src);
  }
  if (!record) return FDMOD_OK;
  if (snap_now) {
    int rc = snapshot(c, it);
    if (rc) return rc;
  }
  if (rec_now && !frec && !combine) launch_receivers(c->stream, r);
  if (c->energy) {
    launch_energy_partial(c->stream, c->g, p.nx, p.nz, c->d, c->d_partial, c->eblocks);
    launch_energy_final(c->stream, c->d_partial, c->eblocks, p.dx, p.dz, c->d_E, c->d_Ek, c->d_Ep, it);
  }
  return FDMOD_OK;
}

extern "C" int fdmod_run(fdmod_ctx *c, int k0, int count, double *seconds)
{
  if (!c) return FDMOD_EINVAL;
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "fdmod_run before fdmod_set_model");
  if (k0 < 0 || count < 0) return fail(c, FDMOD_EINVAL, "bad iteration range");
  const fdmod_params &p = c->p;
  int k1 = std::min(k0 + count, p.nt);
  SweepArgs sa = sweep_args(c);
  HIPCALL(c, hipEventRecord(c->ev0, c->stream));
  for (int k = k0; k < k1; k++) {
    int it = p.mode == 1 ? k : p.nt - 1 - k;
    int rc = step(c, it, sa);
    if (rc) return rc;
  }
  HIPCALL(c, hipEventRecord(c->ev1, c->stream));
  HIPCALL(c, hipEventSynchronize(c->ev1));
  HIPCALL(c, hipGetLastError());
  float ms = 0.0f;
  HIPCALL(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return FDMOD_OK;
}

extern "C" int fdmod_get_traces(fdmod_ctx *c, int comp, float *out)
{
  if (!c || !out) return FDMOD_EINVAL;
  float *d = comp == FDMOD_TRACE_VX ? c->d_rvx : comp == FDMOD_TRACE_VZ ? c->d_rvz : comp == FDMOD_TRACE_P ? c->d_rp
                                                                                                           : nullptr;
  if (!d) return fail(c, FDMOD_EINVAL, "component %d was not recorded", comp);
  std::vector<float> tmp((size_t)c->nr * c->ntsample);
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipMemcpy(tmp.data(), d, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->ntsample; s++)
    for (int r = 0; r < c->nr; r++) out[(size_t)r * c->ntsample + s] = tmp[(size_t)s * c->nr + r];
  return FDMOD_OK;
}

extern "C" int fdmod_get_energy(fdmod_ctx *c, float *E, float *Ek, float *Ep)
{
  if (!c || !E || !Ek || !Ep) return FDMOD_EINVAL;
  if (!c->d_E) return fail(c, FDMOD_ESTATE, "energy was not enabled");
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipMemcpy(E, c->d_E, c->p.nt * sizeof(float), hipMemcpyDeviceToHost));
  HIPCALL(c, hipMemcpy(Ek, c->d_Ek, c->p.nt * sizeof(float), hipMemcpyDeviceToHost));
  HIPCALL(c, hipMemcpy(Ep, c->d_Ep, c->p.nt * sizeof(float), hipMemcpyDeviceToHost));
  return FDMOD_OK;
}

extern "C" int fdmod_get_array(fdmod_ctx *c, int which, float *out)
{
  if (!c || !out || which < 0 || which >= FDMOD_A_COUNT) return FDMOD_EINVAL;
  if (!c->d.a[which]) return fail(c, FDMOD_EINVAL, "array %d is not used by this form", which);
  const Grid &g = c->g;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipMemcpy2D(out, g.Nz * sizeof(float), c->d.a[which], g.ldz * sizeof(float), g.Nz * sizeof(float), g.Nx,
                         hipMemcpyDeviceToHost));
  return FDMOD_OK;
}

extern "C" int fdmod_set_array(fdmod_ctx *c, int which, const float *in)
{
  if (!c || !in || which < 0 || which >= FDMOD_A_COUNT) return FDMOD_EINVAL;
  if (!c->d.a[which]) return fail(c, FDMOD_EINVAL, "array %d is not used by this form", which);
  const Grid &g = c->g;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipMemcpy2D(c->d.a[which], g.ldz * sizeof(float), in, g.Nz * sizeof(float), g.Nz * sizeof(float), g.Nx,
                         hipMemcpyHostToDevice));
  if (which == FDMOD_A_COEF1 || which == FDMOD_A_COEF2) {
    launch_tile_flags(c->stream, g, c->d.a[FDMOD_A_COEF1], c->d.a[FDMOD_A_COEF2], c->flags);
    HIPCALL(c, hipStreamSynchronize(c->stream));
  }
  return FDMOD_OK;
}

extern "C" int fdmod_time_sweeps(fdmod_ctx *c, int reps, double *ms_v, double *ms_s)
{
  if (!c || reps < 1) return FDMOD_EINVAL;
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "fdmod_time_sweeps before fdmod_set_model");
  const fdmod_params &p = c->p;
  SweepArgs sa = sweep_args(c);
  // reps loop iterations from iteration 0, each sweep bracketed by events: the sweeps
  // alternate and carry the loop's fused source and receiver work, as in fdmod_run.
  // (Timed back to back, a sweep would find part of its read-only operands still in the
  // last-level cache, which the other sweep evicts in the loop.)
  const int n = std::min(reps, p.nt);
  std::vector<hipEvent_t> ev(3 * (size_t)n, nullptr);
  hipError_t e = hipSuccess;
  int rc = FDMOD_OK;
  for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
  const int wsave = c->wsave;  // no snapshot files while timing
  c->wsave = 0;
  for (int k = 0; k < n && e == hipSuccess && rc == FDMOD_OK; k++)
    rc = step(c, p.mode == 1 ? k : p.nt - 1 - k, sa, true, &ev[3 * (size_t)k]);
  c->wsave = wsave;
  if (e == hipSuccess && rc == FDMOD_OK) e = hipEventSynchronize(ev[3 * (size_t)n - 1]);
  if (e == hipSuccess && rc == FDMOD_OK) e = hipGetLastError();
  double tv = 0.0, ts = 0.0;
  for (int k = 0; k < n && e == hipSuccess && rc == FDMOD_OK; k++) {
    float a = 0.0f, b = 0.0f;
    e = hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
    tv += a;
    ts += b;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc != FDMOD_OK) return rc;
  if (e != hipSuccess) return fail(c, FDMOD_EDEVICE, "fdmod_time_sweeps: %s", hipGetErrorString(e));
  if (ms_v) *ms_v = tv / n;
  if (ms_s) *ms_s = ts / n;
  return FDMOD_OK;
}

// ---------------------------------------------------------------------------------
// Batched shots.  Each shot is a full modelling context (its own padded grid, MEAL,
// Cij, source and receivers, set up exactly as a single shot); the time loop runs
// all shots together: one launch per sweep covers every shot's tiles, and K steps
// of launches are captured once into a hipGraph and replayed.  Kernels that depend
// on the time index read the loop counter from device memory, so one captured
// chunk serves the whole run.


int bfail(fdmod_batch *b, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}


static void batch_unprepare(fdmod_batch *b)
{
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  dfree(b->d_sweep); dfree(b->d_startV); dfree(b->d_startS);
  dfree(b->d_srcV); dfree(b->d_srcS); dfree(b->d_rec);
  b->d_sweep = nullptr; b->d_startV = b->d_startS = nullptr;
  b->d_srcV = b->d_srcS = nullptr; b->d_rec = nullptr;
  b->prepared = false;
}

extern "C" void fdmod_batch_destroy(fdmod_batch *b)
{
  if (!b) return;
  batch_unprepare(b);
  dfree(b->d_step);
  for (GradMem &m : b->gmem) {
    dfree(m.mu); dfree(m.mus); dfree(m.G); dfree(m.snap);
    dfree(m.ckpt.data); dfree(m.ckpt.scale); dfree(m.hist.data); dfree(m.hist.scale);
  }
  for (fdmod_ctx *c : b->shots) fdmod_destroy(c);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" const char *fdmod_batch_last_error(const fdmod_batch *b) { return b ? b->err.c_str() : "null batch"; }

extern "C" int fdmod_batch_create(const fdmod_params *p, int nshots, const int *nx, fdmod_batch **out)
{
  if (!p || !out || nshots < 1 || !nx) return FDMOD_EINVAL;
  *out = nullptr;
  fdmod_batch *b = new fdmod_batch();
  b->p = *p;
  b->segLS = default_segLS(p->form, 0);  // the tuner may choose another length
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipMalloc(&b->d_step, sizeof(int)) != hipSuccess) {
    fdmod_batch_destroy(b);
    return FDMOD_EDEVICE;
  }
  if (const char *gc = getenv("FDMOD_BATCH_PTS")) {  // shot-group size override (tuning experiments)
    const long long v = atoll(gc);
    if (v >= 1) b->group_pts = (size_t)v;
  }
  for (int s = 0; s < nshots; s++) {
    fdmod_params q = *p;
    q.nx = nx[s];
    fdmod_ctx *c = nullptr;
    int rc = fdmod_create(&q, &c);
    if (rc != FDMOD_OK) {
      fdmod_batch_destroy(b);
      return rc;
    }
    b->shots.push_back(c);
  }
  *out = b;
  return FDMOD_OK;
}

extern "C" fdmod_ctx *fdmod_batch_shot(fdmod_batch *b, int shot)
{
  if (!b || shot < 0 || shot >= (int)b->shots.size()) return nullptr;
  b->prepared = false;  // the caller may change the shot through the single-shot API
  return b->shots[shot];
}

static int batch_shot_rc(fdmod_batch *b, int shot, int rc)
{
  if (rc != FDMOD_OK) bfail(b, rc, "shot %d: %s", shot, b->shots[shot]->err.c_str());
  b->prepared = false;
  return rc;
}

extern "C" int fdmod_batch_set_model(fdmod_batch *b, int shot, const float *vp, const float *vs, const float *rho,
                                     const float *eps, const float *delta, const float *theta, double *dt_max)
{
  if (!b || shot < 0 || shot >= (int)b->shots.size()) return FDMOD_EINVAL;
  return batch_shot_rc(b, shot, fdmod_set_model(b->shots[shot], vp, vs, rho, eps, delta, theta, dt_max));
}

extern "C" int fdmod_batch_set_source(fdmod_batch *b, int shot, int stype, int ns, const int *posx, const int *posz,
                                      const float *src)
{
  if (!b || shot < 0 || shot >= (int)b->shots.size()) return FDMOD_EINVAL;
  return batch_shot_rc(b, shot, fdmod_set_source(b->shots[shot], stype, ns, posx, posz, src));
}

extern "C" int fdmod_batch_set_receivers(fdmod_batch *b, int shot, int seistype, int mt, int nr, const int *posx,
                                         const int *posz)
{
  if (!b || shot < 0 || shot >= (int)b->shots.size()) return FDMOD_EINVAL;
  return batch_shot_rc(b, shot, fdmod_set_receivers(b->shots[shot], seistype, mt, nr, posx, posz));
}

static SourceArgs shot_source(fdmod_ctx *c, const SweepArgs &sa, bool stress_phase)
{
  SourceArgs s;
  memset(&s, 0, sizeof s);
  s.g = c->g;
  bool on = c->ns > 0 && (stress_phase ? c->stype == 1 : c->stype != 1);
  if (!on) return s;  // ns = 0: nothing to do in this phase
  s.isx = c->d_isx; s.isz = c->d_isz; s.src = c->d_src;
  s.nt = c->p.nt; s.ns = c->ns; s.it = 0;
  s.taps = c->d_taps; s.ntap = c->ntap; s.npass = c->npass; s.sequential = c->seq;
  if (stress_phase) {
    s.fa = sa.sxx;
    s.fb = sa.szz;
    if (c->p.form == FDMOD_FORM_TENSOR) {
      s.fc = sa.txx;
      s.fd = sa.tzz;
    }
  } else {
    s.fa = c->stype == 6 ? sa.vz : sa.vx;
    s.fb = (c->stype == 5 || c->stype == 6) ? nullptr : sa.vz;
  }
  return s;
}

// Loop iteration (step counter + off) of shot group g.
static int batch_step(fdmod_batch *b, int g, int off)
{
  const int s0 = b->grp[g], n = b->grp[g + 1] - s0;
  launch_vsweep_batch(b->stream, b->p.form, b->d_sweep + s0, b->d_startV + s0 + g, n, b->totV[g]);
  if (b->srcBlocksV)
    launch_source_batch(b->stream, b->d_srcV + s0, n, b->srcBlocksV, b->d_step, off, b->p.nt, b->p.mode);
  launch_ssweep_batch(b->stream, b->p.form, b->d_sweep + s0, b->d_startS + s0 + g, n, b->totS[g], b->d_step, off,
                      b->p.nt, b->p.mode, b->grp_tau_nt[g] != 0);
  if (b->srcBlocksS)
    launch_source_batch(b->stream, b->d_srcS + s0, n, b->srcBlocksS, b->d_step, off, b->p.nt, b->p.mode);
  if (b->maxnr) launch_receivers_batch(b->stream, b->d_rec + s0, n, b->maxnr, b->d_step, off, b->p.nt, b->p.mode);
  return FDMOD_OK;
}

// count iterations of every shot, group after group, then the step counter advances.
// Shots are independent, so the order changes no result, only the cache's reuse: a
// group's fields are about those of one full-size shot, and its consecutive sweeps
// find part of each other's fields still in the device's last-level cache, which a
// sweep over every shot's fields would have evicted.
static void batch_steps(fdmod_batch *b, int count)
{
  for (size_t g = 0; g + 1 < b->grp.size(); g++)
    for (int j = 0; j < count; j++) batch_step(b, (int)g, j);
  launch_advance(b->stream, b->d_step, count);
}

static int batch_prepare(fdmod_batch *b, bool with_graph)
{
  batch_unprepare(b);
  const int n = (int)b->shots.size();
  std::vector<SweepArgs> sw(n);
  std::vector<SourceArgs> sv(n), ss(n);
  std::vector<RecvArgs> rv(n);
  b->srcBlocksV = b->srcBlocksS = b->maxnr = 0;
  // shot groups of up to group_pts padded points (at least one shot each)
  b->grp.assign(1, 0);
  size_t pts = 0;
  std::vector<size_t> gpts;  // padded points per group
  for (int s = 0; s < n; s++) {
    const size_t w = (size_t)b->shots[s]->g.Nx * b->shots[s]->g.ldz;
    if (s > b->grp.back() && pts + w > b->group_pts) {
      b->grp.push_back(s);
      gpts.push_back(pts);
      pts = 0;
    }
    pts += w;
  }
  gpts.push_back(pts);
  b->grp.push_back(n);
  b->grp_tau_nt.assign(gpts.size(), 0);
  const int ng = (int)b->grp.size() - 1;
  std::vector<int> startV(n + ng, 0), startS(n + ng, 0);
  b->totV.assign(ng, 0);
  b->totS.assign(ng, 0);
  b->grpLV.assign(ng, b->segLV);
  if (b->segLV == 0) {
    // per group, as for a single shot: the shortest velocity segment (at least kFlagW
    // columns) whose blocks, over all the group's shots, run in one generation
    const long slots = vsweep_slots();
    for (int g = 0; g < ng; g++) {
      int L = kFlagW;
      for (;; L++) {
        long tiles = 0;
        for (int s = b->grp[g]; s < b->grp[g + 1]; s++) tiles += sweep_tiles(b->shots[s]->g, L);
        if (tiles <= slots || L >= kMaxSegLVBatch) break;
      }
      b->grpLV[g] = L;
    }
  }
  for (int s = 0; s < n; s++) {
    fdmod_ctx *c = b->shots[s];
    if (!c->model_set) return bfail(b, FDMOD_ESTATE, "shot %d has no model", s);
    if (c->wsave > 0 || c->energy) return bfail(b, FDMOD_EINVAL, "shot %d: snapshots/energy are single-shot only", s);
    const int g = (int)(std::upper_bound(b->grp.begin(), b->grp.end(), s) - b->grp.begin()) - 1;
    // the batch's tiles go into the launch arguments only: the shot's own context keeps
    // the single-shot tiles for its gradient and single-shot runs
    sw[s] = sweep_args(c);
    sw[s].g.segLV = b->grpLV[g];
    sw[s].g.segLS = b->segLS;
    sw[s].tau_nt = 20.0 * 4.0 * gpts[g] > 512.0 * (1 << 20);  // tau is reused after the group's sweeps
    b->grp_tau_nt[g] = (char)sw[s].tau_nt;
    const int k = s + g;  // this shot's entry in its group's prefix array
    startV[k + 1] = startV[k] + sweep_tiles(c->g, b->grpLV[g]);
    startS[k + 1] = startS[k] + sweep_tiles(c->g, b->segLS);
    b->totV[g] = startV[k + 1];
    b->totS[g] = startS[k + 1];
    sv[s] = shot_source(c, sw[s], false);
    ss[s] = shot_source(c, sw[s], true);
    // as in the single-shot loop, the stress sweep adds a single covered stype-1 source
    // and records seistype-2 receivers; the sweep reads the iteration on the device
    const bool fsrc = c->fuse_src && c->ns > 0;
    const bool frec = c->seistype == 2 && c->nr > 0 && c->d_rc_start && b->p.form != FDMOD_FORM_TENSOR &&
                      (fsrc || !(c->ns > 0 && c->stype == 1));  // receivers sample after the stress source
    if (fsrc) {
      sw[s].fuse |= 1;
      sw[s].fsx = c->src_cell_x;
      sw[s].fsz = c->src_cell_z;
      sw[s].ftap = c->d_taps;
      sw[s].fsrc = c->d_src;
      ss[s].ns = 0;
    }
    if (frec) {
      sw[s].fuse |= 2;
      sw[s].rc_start = c->d_rc_start;
      sw[s].rc_z = c->d_rc_z;
      sw[s].rc_id = c->d_rc_id;
      sw[s].rc_zlo = c->rc_zlo;
      sw[s].rc_zhi = c->rc_zhi;
      sw[s].rec_p = c->d_rp;
      sw[s].rc_mt = c->mt;
      sw[s].rc_nr = c->nr;
    }
    if (sv[s].ns) b->srcBlocksV = std::max(b->srcBlocksV, sv[s].sequential ? 1 : sv[s].ns);
    if (ss[s].ns) b->srcBlocksS = std::max(b->srcBlocksS, ss[s].sequential ? 1 : ss[s].ns);
    RecvArgs &r = rv[s];
    memset(&r, 0, sizeof r);
    if (c->seistype > 0 && c->nr > 0 && !frec) {
      r.nr = c->nr; r.ldz = c->g.ldz; r.gx = c->d_gx; r.gz = c->d_gz;
      r.vx = sw[s].vx; r.vz = sw[s].vz; r.sxx = sw[s].sxx; r.szz = sw[s].szz;
      r.out_vx = c->d_rvx; r.out_vz = c->d_rvz; r.out_p = c->d_rp;
      r.want_v = c->seistype == 1 || c->seistype == 3;
      r.want_p = c->seistype == 2 || c->seistype == 3;
      r.mt = c->mt;
      b->maxnr = std::max(b->maxnr, c->nr);
    } else {
      r.mt = 1;
    }
  }
  b->totalV = b->totalS = 0;
  for (int g = 0; g < ng; g++) {
    b->totalV += b->totV[g];
    b->totalS += b->totS[g];
  }
  BHIP(b, hipMalloc(&b->d_sweep, n * sizeof(SweepArgs)));
  BHIP(b, hipMalloc(&b->d_startV, (n + ng) * sizeof(int)));
  BHIP(b, hipMalloc(&b->d_startS, (n + ng) * sizeof(int)));
  BHIP(b, hipMalloc(&b->d_srcV, n * sizeof(SourceArgs)));
  BHIP(b, hipMalloc(&b->d_srcS, n * sizeof(SourceArgs)));
  BHIP(b, hipMalloc(&b->d_rec, n * sizeof(RecvArgs)));
  BHIP(b, hipMemcpy(b->d_sweep, sw.data(), n * sizeof(SweepArgs), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_startV, startV.data(), (n + ng) * sizeof(int), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_startS, startS.data(), (n + ng) * sizeof(int), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_srcV, sv.data(), n * sizeof(SourceArgs), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_srcS, ss.data(), n * sizeof(SourceArgs), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_rec, rv.data(), n * sizeof(RecvArgs), hipMemcpyHostToDevice));
  for (fdmod_ctx *c : b->shots) BHIP(b, hipStreamSynchronize(c->stream));
  if (with_graph) {
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    batch_steps(b, b->chunk);
    hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
    if (e != hipSuccess) return bfail(b, FDMOD_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    BHIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  }
  b->prepared = true;
  return FDMOD_OK;
}

static int batch_run_eager(fdmod_batch *b, int count)
{
  batch_steps(b, count);
  return FDMOD_OK;
}

extern "C" int fdmod_batch_run(fdmod_batch *b, int k0, int count, double *seconds)
{
  if (!b) return FDMOD_EINVAL;
  if (k0 < 0 || count < 0) return bfail(b, FDMOD_EINVAL, "bad iteration range");
  count = std::max(0, std::min(count, b->p.nt - k0));
  if (!b->prepared) {
    int rc = batch_prepare(b, true);
    if (rc) return rc;
  }
  BHIP(b, hipMemcpyAsync(b->d_step, &k0, sizeof(int), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  int done = 0;
  while (count - done >= b->chunk) {
    BHIP(b, hipGraphLaunch(b->gexec, b->stream));
    done += b->chunk;
  }
  if (count > done) batch_run_eager(b, count - done);
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  BHIP(b, hipEventSynchronize(b->ev1));
  BHIP(b, hipGetLastError());
  float ms = 0.0f;
  BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return FDMOD_OK;
}

extern "C" int fdmod_batch_reset(fdmod_batch *b)
{
  if (!b) return FDMOD_EINVAL;
  for (size_t s = 0; s < b->shots.size(); s++) {
    int rc = fdmod_reset(b->shots[s]);
    if (rc) return bfail(b, rc, "shot %zu: %s", s, b->shots[s]->err.c_str());
  }
  return FDMOD_OK;
}

extern "C" int fdmod_batch_get_traces(fdmod_batch *b, int shot, int comp, float *out)
{
  if (!b || shot < 0 || shot >= (int)b->shots.size()) return FDMOD_EINVAL;
  BHIP(b, hipStreamSynchronize(b->stream));
  int rc = fdmod_get_traces(b->shots[shot], comp, out);
  if (rc) return bfail(b, rc, "shot %d: %s", shot, b->shots[shot]->err.c_str());
  return FDMOD_OK;
}

// Autotuner: times candidate segment lengths of both sweeps on this batch and keeps the
// fastest of each.  Each candidate runs min(reps, nt) whole loop iterations from
// iteration 0, as the run launches them (HIP events on the batch stream): a sweep timed
// back to back would find operands in the last-level cache that the loop evicts.  The
// wavefields advance during tuning and are reset afterwards.
extern "C" int fdmod_batch_tune(fdmod_batch *b, int reps, int *segLV, int *segLS)
{
  if (!b || reps < 1) return FDMOD_EINVAL;
  static const int candV[] = {0, 22, 33, 44, 55, 66, 88};  // 0: per group (one generation)
  static const int candS[] = {4, 5, 8, 11, 22};
  constexpr int kTuneWarmSteps = 64;
  const int n = std::min(reps, b->p.nt);
  const int zero = 0;
  // Bring the GPU out of idle first: candidates timed straight after the setup read
  // slow in the order they run (on 64 crops of 1126 x 7872 a cold tuner chose 55/11
  // columns, 54.4 Gpts/s, a warm one 22/11, 54.7).
  {
    int rc = batch_prepare(b, false);
    if (rc) return rc;
    for (int done = 0; done < kTuneWarmSteps; done += n) {
      BHIP(b, hipMemcpyAsync(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
      batch_steps(b, n);
    }
  }
  auto time_one = [&](bool vsweep, int L, double &ms_out) -> int {
    if (vsweep) b->segLV = L;
    else b->segLS = L;
    int rc = batch_prepare(b, false);
    if (rc) return rc;
    BHIP(b, hipMemcpyAsync(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
    batch_steps(b, std::min(2, n));  // warm-up iterations
    BHIP(b, hipMemcpyAsync(b->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    batch_steps(b, n);
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    BHIP(b, hipEventSynchronize(b->ev1));
    BHIP(b, hipGetLastError());
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    ms_out = ms / n;
    return FDMOD_OK;
  };
  double best = 1e30;
  int bestV = b->segLV, bestS = b->segLS;
  for (int L : candV) {
    double ms;
    int rc = time_one(true, L, ms);
    if (rc) return rc;
    if (ms < best) best = ms, bestV = L;
  }
  b->segLV = bestV;
  best = 1e30;
  for (int L : candS) {
    double ms;
    int rc = time_one(false, L, ms);
    if (rc) return rc;
    if (ms < best) best = ms, bestS = L;
  }
  b->segLS = bestS;
  batch_unprepare(b);
  if (segLV) *segLV = bestV;
  if (segLS) *segLS = bestS;
  return fdmod_batch_reset(b);
}

extern "C" int fdmod_batch_set_tiles(fdmod_batch *b, int segLV, int segLS)
{
  if (!b || segLV < 0 || segLS < 1 || segLV > 4096 || segLS > 4096) return FDMOD_EINVAL;
  b->segLV = segLV;
  b->segLS = segLS;
  b->prepared = false;
  return FDMOD_OK;
}
