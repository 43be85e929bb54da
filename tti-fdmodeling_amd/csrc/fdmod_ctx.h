// The modelling context behind the C-ABI handle, shared by the translation units
// of the host library (fdmod_api.hip: forward loop and batches; fdmod_grad.hip:
// adjoint and gradient).
#pragma once
#include <stdarg.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "fdmod_internal.h"

// Device slots of padded fields, stored as FP32 or as FP16 with per-(column, 64-row
// block) inverse scales.
struct FieldStore {
  void *data = nullptr;
  float *scale = nullptr;
  size_t cap_data = 0, cap_scale = 0;  // bytes allocated
};

struct fdmod_ctx {
  fdmod_params p;
  fdm::Grid g;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  fdm::DevArrays d;
  float *pool = nullptr;
  int zsh = 0;    // elements each pool and gradient field is shifted by (row izStart on a 128-byte line)
  int zsh16 = 0;  // the same for FP16 gradient storage
  uint8_t *flags = nullptr;
  bool model_set = false;
  // source
  int stype = 0, ns = 0, npass = 0, ntap = 0, seq = 1;
  int *d_isx = nullptr, *d_isz = nullptr;
  float *d_src = nullptr;
  fdm::SrcTap *d_taps = nullptr;
  int fuse_src = 0;  // the stress sweep adds the stress source (one stype-1 source it covers)
  int src_cell_x = 0, src_cell_z = 0;  // its padded cell
  // receivers
  int seistype = 0, mt = 1, nr = 0, ntsample = 0;
  int *d_rc_start = nullptr, *d_rc_z = nullptr, *d_rc_id = nullptr;  // receivers by column (CSR)
  int rc_zlo = 0, rc_zhi = -1;  // padded rows spanned by the receivers
  int *d_gx = nullptr, *d_gz = nullptr;
  float *d_rvx = nullptr, *d_rvz = nullptr, *d_rp = nullptr;
  // pressure receivers i = 0..nr-1 at padded cell (rline_x0 + i, rline_z): the adjoint
  // sweeps then add the pressure residuals themselves (no injection launch)
  bool rline = false;
  int rline_x0 = 0, rline_z = 0;
  const float *inj_last = nullptr;  // the last adjoint step's fused residuals (or null)
  // one unfused stype-1 source: its receivers inside the source's footprint (indices),
  // which the combined source + receiver launch records after the injection; nfprc < 0:
  // no combined launch.  Host copies of the padded source and receiver cells build it.
  int *d_fprc = nullptr;
  int nfprc = -1;
  std::vector<int> h_isx, h_isz, h_gx, h_gz;
  // snapshots
  int wsave = 0, savew = 0, snap_sample = 1;
  fdmod_snapshot_fn snap_fn = nullptr;
  void *snap_user = nullptr;
  float *d_p = nullptr;
  std::vector<float> h_snap;
  // energy
  int energy = 0, eblocks = 0;
  float *d_E = nullptr, *d_Ek = nullptr, *d_Ep = nullptr, *d_partial = nullptr;
  // adjoint / gradient workspace (allocated on first use, kept until destroy)
  float *d_mu = nullptr;       // five adjoint fields, each Nx*ldz
  double *d_G = nullptr;       // six imaging sums, each Nx*ldz
  float *d_mus = nullptr;      // adjoint stress ring: kImgK + 1 slots x (mxx, mzz, mxz), Nx*ldz each
  int mus_cur = 0, mus_last = 0;  // slot the next adjoint step starts from; slot the last one imaged
  float *d_snap = nullptr;     // mu_sigma snapshots of one imaging block (kImgK x 3 fields)
  size_t d_mu_n = 0;           // field size (floats) d_mu, d_G and d_snap were allocated for
  FieldStore ckpt;             // checkpointed states (5 fields per slot)
  FieldStore hist;             // v_{k+1} per step (2 fields per slot): one segment or the whole run
  float *d_res = nullptr;      // data-space adjoint sources [3][ntsample][nr]
  double *d_red = nullptr;     // reduction partials (kRedBlocks) and scalars
  double *d_srcadj = nullptr;  // adjoint of the source time functions [ns][nt]
  double *d_grad = nullptr;    // gradient [4][nx][nz]
  int adj_segS = 0, adj_segV = 0;  // adjoint sweeps' segment lengths (0: not chosen yet)
  float *d_coef2n = nullptr;   // -coef2 (Nx*ldz) and -src ([ns][nt]) for time-reversed steps
  float *d_srcn = nullptr;
  size_t d_srcn_n = 0;
  std::string err;
};

// Record a printf-style message as the context's last error and return code.
int fail(fdmod_ctx *c, int code, const char *fmt, ...);

#define HIPCALL(c, x)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail((c), FDMOD_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

void dfree(void *p);  // hipFree that ignores null

// Gradient memory of one shot, pooled by a batch and lent to the shot being worked on
// (re-allocating tens of GB for every group of shots costs more than the gradient).
struct GradMem {
  float *mu = nullptr;
  float *mus = nullptr;
  double *G = nullptr;
  float *snap = nullptr;
  size_t n = 0;  // field size the three buffers above were allocated for
  FieldStore ckpt, hist;
};

// A batch of shots (fdmod_api.hip: batched forward loop; fdmod_grad.hip: batched
// gradients).  Each shot is a full modelling context with its own stream.
struct fdmod_batch {
  fdmod_params p;
  std::vector<fdmod_ctx *> shots;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // segLV 0: each shot group's velocity segment is chosen so that its blocks run in one
  // generation (grpLV); otherwise every group uses segLV
  int segLV = 0, segLS = fdm::kDefaultSegLS;
  bool prepared = false;
  fdm::SweepArgs *d_sweep = nullptr;
  int *d_startV = nullptr, *d_startS = nullptr, totalV = 0, totalS = 0;
  // Shot groups run one after another inside each captured chunk of steps: group g is
  // shots grp[g] .. grp[g+1]-1, its tile prefix sums start at d_startV/S + grp[g] + g,
  // its tile totals are totV/totS[g].  Groups hold up to group_pts padded points.
  std::vector<int> grp, totV, totS;
  std::vector<char> grp_tau_nt;  // per group: tensor-form tau stored non-temporally
  std::vector<int> grpLV;        // per group: velocity-sweep segment length
  size_t group_pts = fdm::kBatchGroupPts;
  fdm::SourceArgs *d_srcV = nullptr, *d_srcS = nullptr;
  int srcBlocksV = 0, srcBlocksS = 0;
  fdm::RecvArgs *d_rec = nullptr;
  int maxnr = 0;
  int *d_step = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int chunk = 32;
  std::vector<GradMem> gmem;  // gradient memory pool, one entry per shot of a group
  std::string err;
};

int bfail(fdmod_batch *b, int code, const char *fmt, ...);

#define BHIP(b, x)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) return bfail((b), FDMOD_EDEVICE, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)


fdm::SweepArgs sweep_args(fdmod_ctx *c);
// One iteration of the loop; record = false skips snapshots, receivers and energy.
// One loop iteration.  mark, if given, holds three events recorded before the velocity
// sweep, between the sweeps (after any velocity source) and after the stress sweep.
int step(fdmod_ctx *c, int it, const fdm::SweepArgs &sa, bool record = true, const hipEvent_t *mark = nullptr);
