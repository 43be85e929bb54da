// Internal definitions shared by the libfdmod_hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/fdmod_hip.h"

namespace fdm {

// Radius of the RSG stencil (N in the reference, BT:307).
constexpr int kN = 5;

// Least-squares RSG coefficients (BT:464-465).
constexpr float kC1 = 1.24362739770907f;
constexpr float kC2 = -0.114021164233934f;
constexpr float kC3 = 0.0272965993810630f;
constexpr float kC4 = -0.00699199493927320f;
constexpr float kC5 = 0.00132022981661381f;

// PI as defined by Seismic Unix cwp.h, used by the theta conversion (BT:508).
constexpr double kPI = 3.141592653589793;

// Sweep decomposition: one wavefront per block owns a strip of TZ = 64 rows (one
// row per lane) and marches along x over a segment of columns (segLV columns in
// the velocity sweep, segLS in the stress sweep).  "Undamped" flags are kept per
// fine tile of TZ x kFlagW points; a block ANDs the flags its segment covers.
constexpr int TZ = 64;
constexpr int kRing = 11;          // LDS ring of staged columns (10 live + 1 in flight)
constexpr int kFlagW = 11;         // columns per fine flag tile
constexpr int kDefaultSegLV = 44;  // longest velocity-sweep segment
constexpr int kVBlocksPerCU = 16;  // velocity-sweep blocks per CU: 160 KB LDS / 9.6 KB ring
constexpr int kMaxSegLVBatch = 88;  // longest per-group velocity segment of a batched loop
constexpr int kTZS = 55;           // stress adjoint: output rows per strip (64 staged rows, halo 4 + 5)
// Stress-sweep segments: 11 columns on every grid.  Odd segments walk right to left
// and share their halo through L2.  Measured on 1256 x 8002 with 128-byte-aligned
// strips (us per launch, segments of 4 / 6 / 8 / 11 / 16 columns): Bond 99.7 / 95.3 /
// 93.5 / 92.9 / 93.6; coordinate 99.1 / 94.9 / 92.7 / 91.7; tensor 142.8 / 140.9 /
// 137.1 / 137.6.  On a 1256 x 2130 crop 11 columns ran the forward 10% faster than 8.
constexpr int kDefaultSegLS = 11;
inline int default_segLS(int /*form*/, int /*cols*/) { return kDefaultSegLS; }
constexpr int kZAlign = 32;    // floats per 128-byte line: strips start on a line
// Padded grid points per shot group of a batched loop: about one full-size shot
// (10240 columns of 1280 rows); measured on 64 crops 1126 deep, groups of 9-10k
// columns ran 8-10% faster than one group (48.3 vs 44.5 Gpts/s).
constexpr size_t kBatchGroupPts = (size_t)10240 * 1280;

// Device-side geometry of a padded grid.
struct Grid {
  int Nz, Nx, ldz;         // padded sizes; ldz = column pitch in floats (>= Nz)
  int ixStart, ixEnd, izStart, izEnd;
  int e;                   // effective ealThick (user value + kN)
  int segLV, segLS;        // columns per x segment of the velocity / stress sweep
  int tilesZ, flagsX;      // 64-row strips; fine flag tiles along x
  float invdx, invdz, dt;  // float(1.0/dx), float(1.0/dz), dt
};

// Device pointers of one modelling context (all padded Nx*ldz floats).
struct DevArrays {
  float *a[FDMOD_A_COUNT];
};

struct SweepArgs {
  Grid g;
  // fields
  float *vx, *vz, *sxx, *szz, *sxz, *txx, *tzz, *txz;
  // parameters
  const float *coef1, *coef2, *invrho;
  const float *c11, *c13, *c15, *c33, *c35, *c55;
  const float *sin2, *cos2, *sin2t, *sn, *cs, *snv, *csv;
  const uint8_t *undamped;  // per tile: coef1 == 1 and coef2 == dt everywhere in the tile
  int freesurface;
  // Tensor form: tau stored non-temporally (grids whose fields overflow the 256 MB
  // last-level cache; on smaller grids tau is still cached when the next stress sweep
  // reads it).  Read on the host: it selects the kernel.
  int tau_nt;
  // Per-point operands (old state, parameters) loaded non-temporally: grids whose arrays
  // one step streams overflow twice the 256 MB last-level cache.  On smaller grids (the
  // gradient's crops) they load cached, and part of them is still there the next step.
  int op_nt;
  // History output for the gradient, stored at every updated point as FP32 (estore 1)
  // or as FP16 with one power-of-two scale per column and 64-row strip for the fields
  // one sweep stores together (copied into each field's scale array)
  // (estore 2; scales [Nx][tilesZ] in hsc[field]):
  //   Bond        stress sweep (vxdx, vzdz, vzdx + vxdz) -> h[0..2]
  //   tensor      stress sweep (vxdx + vzdz, vxdx - vzdz, vzdx + vxdz) -> h[0..2] and
  //               the new tau (before the stress source) as (txx + tzz, txx - tzz,
  //               txz) -> h[3..5]
  //   coordinate  stress sweep (vxdx, vxdz, vzdx, vzdz) before rotation -> h[0..3];
  //               velocity sweep (-sxxdz + sxzdx, -sxzdz + szzdx, sxxdx + sxzdz,
  //               sxzdx + szzdz) -> h[4..7]
  int estore;
  float *h32[8];
  _Float16 *h16[8];
  float *hsc[8];
  // Time-reversed runs (dt and coef2 negated by the caller): rev = 1 makes the
  // tensor-form history hold the stress sweep's input t, which is the forward's t
  // before the stress source once the caller has taken the source off.  FP32
  // history is not stored inside the box hskip = {x0, x1, z0, z1} (columns x0..x1-1,
  // rows z0..z1-1; an empty box stores everywhere), or with hinside = 1 only inside it.
  int rev;
  int hskip[4];
  int hinside;
  // Single-shot loop, Bond and coordinate forms: the stress sweep also adds the stress
  // source (fuse bit 1: stype 1, one source whose 11 x 11 footprint the sweep writes;
  // ftap = its tap table, ordered ox-major) and records the pressure receivers (bit 2:
  // rc_* = receivers by padded column, CSR; rec_p[rec_off + id]; rows rc_zlo..rc_zhi
  // hold every receiver, so other strips skip the lookup).
  int fuse;
  int fsx, fsz, fit;
  const struct SrcTap *ftap;
  const float *fsrc;
  const int *rc_start, *rc_z, *rc_id;
  int rc_zlo, rc_zhi;
  int rc_mt, rc_nr;  // batched loop: the sweep takes fit and rec_off from the step counter
  float *rec_p;
  size_t rec_off;
};

// Source injection table entry: offset from the source cell and the two
// coefficients that multiply src[it] (already including the reference's
// weight products, computed on the host in the reference's float order).
struct SrcTap {
  int ox, oz;
  float ca, cb;
};

struct SourceArgs {
  Grid g;
  float *fa, *fb;           // targets: vx/vz, or sxx/szz
  float *fc, *fd;           // tensor form stress source also adds to txx/tzz (else null)
  const int *isx, *isz;     // padded source cells, per source
  const float *src;         // [ns][nt]
  int nt, ns, it;
  const SrcTap *taps;       // [npass][ntap]
  int ntap, npass;
  int sequential;           // 1: one block walks the sources in order
};

struct RecvArgs {
  int nr, ldz;
  const int *gx, *gz;       // padded receiver cells
  const float *vx, *vz, *sxx, *szz;
  float *out_vx, *out_vz, *out_p;  // device [ntsample][nr]
  int isample, want_v, want_p;
  int mt;                   // receiver sampling (batched launches derive isample from it)
};

// Arguments of the adjoint sweeps.  g is the grid of the launch: for the
// velocity-sweep adjoint under a free surface its izStart is lowered by kN so the
// image rows receive their adjoint; izU is the first row of the update range.
struct AdjArgs {
  Grid g;
  int izU, form;
  float *mvx, *mvz, *mxx, *mzz, *mxz;  // adjoint state
  float *mtxx, *mtzz, *mtxz;           // tensor form: adjoint of tau
  const float *coef1, *coef2, *invrho;
  const float *c11, *c13, *c15, *c33, *c35, *c55;
  const float *sin2, *cos2, *sin2t, *sn, *cs, *snv, *csv;
  const uint8_t *undamped;             // the forward's per-tile flags (coef1 == 1, coef2 == dt), tiles
  int ftz;                             // of 64 rows from row izU and kFlagW columns; ftz strips
  float *oxx, *ozz, *oxz;              // V-adjoint: where the updated mu_sigma goes (mxx... in place)
  float *snap[6];                      // V-adjoint: imaging snapshots of computed fields, [3..] (or null)
  double *G[8];                        // imaging sums (per-form meaning, kernels_adjoint.hip)
  // Pressure residuals of this step for receivers at padded cells (inj_x0 + i, inj_z),
  // i < inj_n (null: none).  Both adjoint sweeps then add 0.5 inj[i] to mu_sxx and mu_szz
  // where they read them (the imaging does the same with the step's slot).
  const float *inj;
  int inj_x0, inj_z, inj_n;
};

// One imaging launch over nk <= kImgK steps: the V-adjoint's snapshots of each step
// (Bond 3, tensor 6, coordinate 5 fields) and the forward history of the step
// (Bond 3, tensor 6, coordinate 8 fields; FP32 h32, or FP16 h16 with per-(field,
// column, strip) scales: field f's scales of step j start at hsc[j] + f * hsc_stride).
constexpr int kImgK = 16;
struct ImgBlock {
  int nk;
  const float *lam[kImgK][6];
  const float *h32[kImgK][8];
  const _Float16 *h16[kImgK][8];
  const float *hsc[kImgK];
  size_t hsc_stride;
  // steps whose pressure residuals the adjoint sweeps added themselves: the residuals
  // (receivers at padded cells (inj_x0 + i, inj_z), i < inj_n), else null
  const float *inj[kImgK];
  int inj_x0, inj_z, inj_n;
};

// FP16 packing of nf padded fields: dst = src * 2^(14-E) per (column, 64-row block),
// inv = 2^(E-14) per block ([nx][ldz/64]); unpacking writes src back.  A pair of fields
// (a, b) with mate[a] = b, mate[b] = a and a < b is packed as a + b in slot a and a - b in
// slot b, and unpacked as (S + D) / 2 and (S - D) / 2 (mate -1: the field alone).
struct Pack16 {
  int nx, ldz, nf;  // nf <= 8
  float *src[8];
  _Float16 *dst[8];
  float *inv[8];
  int mate[8];
};

// Compact copy of a set of up to 4 rectangles of grid points (rows fastest inside
// each rectangle) for nf fields: pack[f * npts + i] <-> grid[f][x * ldz + z], where
// i enumerates the rectangles' points in order and npts = off[nrect].
struct FrameIO {
  int nrect, ldz, nf, save;  // save = 1: grid -> pack, 0: pack -> grid
  int x0[4], z0[4], h[4];    // rectangle r: columns x0 .., rows z0 .. z0 + h - 1
  long long off[5];          // prefix sums of the rectangles' point counts
  float *grid[8];
  float *pack;
};

// ---- launchers (kernels_setup.hip) ----
void launch_pad(hipStream_t s, const float *raw, float *dst, int nz, int nx, const Grid &g);
void launch_meal(hipStream_t s, const Grid &g, float *vp, float *vs, float *rho, float *D, float *Beta,
                 float *coef1, float *coef2, const float *wtab, int nlayer, int freesurface, float k0, float lnR,
                 float twoThx, float twoThz, float dxf, float dzf, float sqrt3, double dt);
void launch_cij(hipStream_t s, const Grid &g, int form, const DevArrays &d, float scalar);
void launch_stagger(hipStream_t s, const Grid &g, int form, const DevArrays &d);
void launch_fs_fixup(hipStream_t s, const Grid &g, const DevArrays &d);
void launch_tile_flags(hipStream_t s, const Grid &g, const float *coef1, const float *coef2, uint8_t *flags);

// ---- launchers (kernels_step.hip) ----
void launch_vsweep(hipStream_t s, int form, const SweepArgs &a);
void launch_ssweep(hipStream_t s, int form, const SweepArgs &a);
int sweep_tiles(const Grid &g, int segL);
// batched launches over many shots: device arrays of per-shot args, tile prefix sums
void launch_vsweep_batch(hipStream_t s, int form, const SweepArgs *shots, const int *start, int nshots, int total);
void launch_ssweep_batch(hipStream_t s, int form, const SweepArgs *shots, const int *start, int nshots, int total,
                         const int *step, int off, int nt, int mode, bool tau_nt);
void launch_source_batch(hipStream_t s, const SourceArgs *shots, int nshots, int maxblocks, const int *step, int off,
                         int nt, int mode);
void launch_receivers_batch(hipStream_t s, const RecvArgs *shots, int nshots, int maxnr, const int *step, int off,
                            int nt, int mode);
void launch_advance(hipStream_t s, int *step, int n);
void launch_source(hipStream_t s, const SourceArgs &a);
void launch_receivers(hipStream_t s, const RecvArgs &a);
// One stype-1 source (ns = 1) and the receivers in one launch: block 0 injects the source,
// then records the nfp receivers fprc[] inside its footprint; the other blocks record
// every other receiver.
void launch_source_receivers(hipStream_t s, const SourceArgs &src, const RecvArgs &r, const int *fprc, int nfp);
void launch_pressure(hipStream_t s, const float *sxx, const float *szz, float *p, int Nx, int ldz);
void launch_energy_partial(hipStream_t s, const Grid &g, int nx, int nz, const DevArrays &d, float *partial,
                           int nblocks);
void launch_energy_final(hipStream_t s, const float *partial, int nblocks, float dx, float dz, float *E, float *Ek,
                         float *Ep, int it);

// ---- launchers (kernels_adjoint.hip) ----
void launch_adj_ssweep(hipStream_t s, const AdjArgs &a);
void launch_adj_vsweep(hipStream_t s, const AdjArgs &a);
void launch_image(hipStream_t s, const AdjArgs &a, const ImgBlock &b);
void launch_adj_fs(hipStream_t s, const AdjArgs &a);
void launch_adj_inject(hipStream_t s, int nr, int ldz, const int *gx, const int *gz, const float *r1,
                       const float *r2, float *f1, float *f2, float scale);
void launch_adj_source(hipStream_t s, const SourceArgs &a, const float *ma, const float *mb, const float *mc,
                       const float *md, double *out);
void launch_residual(hipStream_t s, const float *d, float *r, size_t n, double *partial, int nb, double *out);
void launch_dot(hipStream_t s, const float *a, const float *bf, const double *bd, size_t n, double *partial, int nb,
                double *out);
// Chain rule into grad [4][nx][nz]; overwrites G[0..3] (the imaging sums are spent).
void launch_chain(hipStream_t s, const Grid &g, int form, int nx, int nz, int fs_row, const DevArrays &d,
                  double *G, size_t gstride, double theta_scale, double *grad);
void launch_fold_trig(hipStream_t s, const Grid &g, double *G, size_t gstride);
void launch_rescale_frozen(hipStream_t s, const Grid &g, float *vp, float *vs, float *rho, const float *Beta);
void launch_pack16(hipStream_t s, const Pack16 &p, bool pack);
void launch_accum_grad(hipStream_t s, const double *src, int nx, int nz, int x0, int nx_full, double *dst);

// ---- launchers (kernels_recon.hip) ----
void launch_frame_io(hipStream_t s, const FrameIO &a);
void launch_negate(hipStream_t s, const float *in, float *out, size_t n);
// Adds to *count the points of columns x0..x1-1, rows z0..z1-1 where coef1 != 1 or
// coef2 != dt.
void launch_count_damped(hipStream_t s, const float *coef1, const float *coef2, float dt, int ldz, int x0, int x1,
                         int z0, int z1, unsigned *count);

}  // namespace fdm
