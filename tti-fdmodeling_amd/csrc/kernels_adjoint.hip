// Adjoint kernels of the modelling loop and of the L2 trace-misfit gradient, for the
// three rotation forms.  No reference counterpart: the reference only runs the
// forward loop.
//
// One forward step is u_{k+1} = FS(S(V(u_k) + s^v_k)) + s^sigma_k and d_k = R u_{k+1}.
// Going backwards through step k the adjoint state mu (same five fields) takes
//   mu += R^T r_k;  mu <- FS^T mu;  G += mu_sigma (x) E(v_{k+1}) (in blocks of steps);
//   mu_v <- coef1 mu_v + E^T(coef2 C mu_sigma);  mu_sigma <- coef1 mu_sigma + D^T(coef2 invrho mu_v)
// where D (stress -> velocity, offsets -4..+5) and E (velocity -> strain, -5..+4) are
// the two RSG derivative operators.  On the grid D^T = -E-shaped and E^T = -D-shaped
// (d1 and d2 of one operator are minus the transposes of the other's), so:
//   adjoint of the S sweep = a velocity-sweep-shaped stencil over M = coef2 C mu_sigma
//   adjoint of the V sweep = a stress-sweep-shaped stencil over N = coef2 invrho mu_v
// Both are gathers, like the forward sweeps, and reuse their staged-column rings.
// The coef1 scaling of each pass is applied by the *other* sweep (mu_sigma's by the
// V-adjoint, mu_v's by the next step's S-adjoint), so that no sweep reads and writes
// the same field: a sweep stages one field family and updates only the other.
//
// Update ranges.  Adjoint fields outside the update range U stay zero when read:
// mu_v is only written inside U, and the image rows above a free surface, which the
// V-adjoint does write, are folded back and cleared by FS^T before the S-adjoint
// stages mu_sigma.  So M and N need no explicit masks.
#include "sweep_common.h"

namespace fdm {

// ---------------------------------------------------------------- S-sweep adjoint
// The staged quantity M is the adjoint of the velocity gradients the stress sweep
// reads at a stress point, computed pointwise from the adjoint stress and the
// parameters (M vanishes outside U because the adjoint stress does there):
//   Bond    M = coef2 C mu_sigma for (vxdx, vzdz, vzdx + vxdz)                 3 fields
//   tensor  a = Q^T mu_sigma + mu_tau, e-adjoint = coef2 C_vti a, then the
//           transpose of the strain rotation, for (vxdx, vzdz, vxdz + vzdx)    3 fields
//   coord   VTI adjoint of the rotated gradients, rotated back, for
//           (vxdx, vxdz, vzdx, vzdz)                                            4 fields
// True when every point of the update range U inside rows [z_lo, z_hi) and columns
// [x_lo, x_hi) is undamped (coef1 == 1.0f, coef2 == dt): a block whose staged and output
// points all qualify uses those constants and loads neither coefficient.  Points outside
// U need no check, the adjoint fields the coefficients multiply there being zero.
__device__ __forceinline__ bool region_undamped(const AdjArgs &a, int z_lo, int z_hi, int x_lo, int x_hi)
{
  const Grid &g = a.g;
  z_lo = max(z_lo, a.izU);
  z_hi = min(z_hi, g.izEnd);
  x_lo = max(x_lo, g.ixStart);
  x_hi = min(x_hi, g.ixEnd);
  if (z_lo >= z_hi || x_lo >= x_hi) return true;
  bool und = true;
  for (int tz = (z_lo - a.izU) / TZ; tz <= (z_hi - 1 - a.izU) / TZ; tz++)
    for (int f = (x_lo - g.ixStart) / kFlagW; f <= (x_hi - 1 - g.ixStart) / kFlagW; f++)
      und = und && a.undamped[f * a.ftz + tz] != 0;
  return und;
}

constexpr int kRowMax = 14;
struct MRow {
  float v[kRowMax];
  float r;  // fused injection: the pressure residual of the row's column
};

template <int FORM>
__device__ __forceinline__ void load_mrow(MRow &r, const AdjArgs &a, size_t o, bool und)
{
  const float k2 = und ? a.g.dt : a.coef2[o];
  r.v[0] = a.mxx[o]; r.v[1] = a.mzz[o]; r.v[2] = a.mxz[o];
  r.v[3] = a.c11[o]; r.v[4] = a.c13[o];
  if (FORM == FDMOD_FORM_BOND) {
    r.v[5] = a.c15[o]; r.v[6] = a.c33[o]; r.v[7] = a.c35[o]; r.v[8] = a.c55[o]; r.v[9] = k2;
  } else if (FORM == FDMOD_FORM_TENSOR) {
    r.v[5] = a.c33[o]; r.v[6] = a.c55[o]; r.v[7] = k2;
    r.v[8] = a.mtxx[o]; r.v[9] = a.mtzz[o]; r.v[10] = a.mtxz[o];
    r.v[11] = a.sin2[o]; r.v[12] = a.cos2[o]; r.v[13] = a.sin2t[o];
  } else {
    r.v[5] = a.c33[o]; r.v[6] = a.c55[o]; r.v[7] = k2; r.v[8] = a.sn[o]; r.v[9] = a.cs[o];
  }
}

template <int FORM>
constexpr int m_fields()
{
  return FORM == FDMOD_FORM_COORD ? 4 : 3;
}

template <int FORM>
__device__ __forceinline__ void m_of(const MRow &r, float (&m)[4])
{
  const float *v = r.v;
  if (FORM == FDMOD_FORM_BOND) {
    const float k2 = v[9];
    m[0] = k2 * (v[3] * v[0] + v[4] * v[1] + v[5] * v[2]);
    m[1] = k2 * (v[4] * v[0] + v[6] * v[1] + v[7] * v[2]);
    m[2] = k2 * (v[5] * v[0] + v[7] * v[1] + v[8] * v[2]);
    m[3] = 0.0f;
  } else if (FORM == FDMOD_FORM_TENSOR) {
    const float S2 = v[11], K2 = v[12], S2T = v[13], k2 = v[7];
    const float a1 = K2 * v[0] + S2 * v[1] - 0.5f * S2T * v[2] + v[8];
    const float a2 = S2 * v[0] + K2 * v[1] + 0.5f * S2T * v[2] + v[9];
    const float a3 = S2T * v[0] - S2T * v[1] + (K2 - S2) * v[2] + v[10];
    const float mexx = k2 * (v[3] * a1 + v[4] * a2);
    const float mezz = k2 * (v[4] * a1 + v[5] * a2);
    const float mexz = k2 * v[6] * a3;
    m[0] = K2 * mexx + S2 * mezz + S2T * mexz;
    m[1] = S2 * mexx + K2 * mezz - S2T * mexz;
    m[2] = -0.5f * S2T * mexx + 0.5f * S2T * mezz + (K2 - S2) * mexz;
    m[3] = 0.0f;
  } else {
    const float k2 = v[7], sn = v[8], cs = v[9];
    const float rxdx = k2 * (v[3] * v[0] + v[4] * v[1]);
    const float rzdz = k2 * (v[4] * v[0] + v[5] * v[1]);
    const float rg = k2 * v[6] * v[2];
    m[0] = cs * rxdx + sn * rg;   // vxdx
    m[1] = -sn * rxdx + cs * rg;  // vxdz
    m[2] = cs * rg + sn * rzdz;   // vzdx
    m[3] = -sn * rg + cs * rzdz;  // vzdz
  }
}

struct VAdjOps {
  float mvx, mvz, k1;
};

// mu_v(p) <- coef1(p) mu_v(p) - (adjoint of the gradients of v read by the stress
// sweep), p in U.  Bond/tensor: (Dx M0 + Dz M2, Dx M2 + Dz M1); coordinate:
// (Dx M0 + Dz M1, Dx M2 + Dz M3).
// Staged column of the stress adjoint: lane l holds row rs + l only (the strip's 64
// staged rows are its kTZS = 55 output rows and their 4 + 5 halo rows).  The caller
// passes the row clamped into the grid: rows past the grid feed no output row.  No
// branch surrounds the loads, so that the look-ahead stays in flight (a load under a
// branch makes the compiler wait for every load before the next use of an older one).
template <int FORM, bool INJ>
__device__ __forceinline__ void load_mrow_lane(MRow &r, const AdjArgs &a, int gx, int rowc, const Grid &g, bool und)
{
  load_mrow<FORM>(r, a, (size_t)gx * g.ldz + rowc, und);
  r.r = INJ ? a.inj[min(max(gx - a.inj_x0, 0), a.inj_n - 1)] : 0.0f;  // one address per wave
}

// M of staged column col into the ring.  INJ: lane ilane holds the receiver row, whose
// mu_sxx and mu_szz take 0.5 r first (the transpose of the pressure receiver, as the
// separate injection adds it).
template <int FORM, bool INJ>
__device__ __forceinline__ void store_mrow(float (*ring)[kRing][TZ], int slot, const MRow &r, int lane, int col,
                                           int ilane, const AdjArgs &a)
{
  constexpr int NM = m_fields<FORM>();
  float m[4];
  if (INJ && lane == ilane && (unsigned)(col - a.inj_x0) < (unsigned)a.inj_n) {
    MRow t = r;
    const float d = 0.5f * r.r;
    t.v[0] = r.v[0] + d;
    t.v[1] = r.v[1] + d;
    m_of<FORM>(t, m);
  } else {
    m_of<FORM>(r, m);
  }
#pragma unroll
  for (int f = 0; f < NM; f++) ring[f][slot][lane] = m[f];
}

// The strip of block row btz outputs rows z0 .. z0+54 (lanes 4..58) and stages rows
// z0-4 .. z0+59, one per lane: each staged value is loaded and turned into M once,
// where 64-row outputs need a second, mostly idle load of the 9 halo rows per field.
template <int FORM, int PD, bool INJ>
__device__ __forceinline__ void adj_s_tile(const AdjArgs &a, int btz, int btx)
{
  constexpr int NM = m_fields<FORM>();
  __shared__ float ring[NM][kRing][TZ];
  const Grid &g = a.g;
  const int segL = g.segLV;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * kTZS;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const int rs = z0 - 4;
  const int iz = rs + lane;
  const int izc = min(max(iz, 0), g.Nz - 1);  // in-grid row for the loads
  const int xe = xa + nc - 1;
  const bool zok = lane >= 4 && lane < 4 + kTZS && iz < g.izEnd;
  const int ilane = INJ ? a.inj_z - rs : -1;  // the lane staging the receiver row (may be none)
  // staged rows rs .. rs+63 of columns xa-4 .. xa+nc+4
  const bool und = region_undamped(a, rs, rs + TZ, xa - 4, xa + nc + 5);
  // prologue in groups of three columns (a staged row is up to 14 floats)
#pragma unroll
  for (int m0 = 0; m0 < 9; m0 += 3) {
    MRow c[3];
#pragma unroll
    for (int m = 0; m < 3; m++) load_mrow_lane<FORM, INJ>(c[m], a, xa - 4 + m0 + m, izc, g, und);
#pragma unroll
    for (int m = 0; m < 3; m++) store_mrow<FORM, INJ>(ring, m0 + m, c[m], lane, xa - 4 + m0 + m, ilane, a);
  }
  MRow colq[PD];
  VAdjOps opq[PD];
  auto load_ops = [&](VAdjOps &p, int col) {
    const size_t o = (size_t)col * g.ldz + izc;
    p.mvx = a.mvx[o];
    p.mvz = a.mvz[o];
    p.k1 = und ? 1.0f : a.coef1[o];
  };
  // look-ahead loads from clamped columns (values past the segment are not used)
#pragma unroll
  for (int i = 0; i < PD; i++) {
    load_mrow_lane<FORM, INJ>(colq[i], a, min(xa + 5 + i, xe + 5), izc, g, und);
    load_ops(opq[i], min(xa + i, xe));
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = xa + k;
      MRow cn;
      VAdjOps on;
      load_mrow_lane<FORM, INJ>(cn, a, min(c + 5 + PD, xe + 5), izc, g, und);
      load_ops(on, min(c + PD, xe));
      store_mrow<FORM, INJ>(ring, (u + 9) % kRing, colq[0], lane, c + 5, ilane, a);
      __syncthreads();
      if (zok) {
        float d1, d2;
        rsg_fwd([&](int dx, int dz) { return ring[0][(u + dx + 4 + kRing) % kRing][lane + dz]; }, d1, d2);
        const float m0dx = 0.5f * (d1 + d2) * invdx;
        rsg_fwd([&](int dx, int dz) { return ring[1][(u + dx + 4 + kRing) % kRing][lane + dz]; }, d1, d2);
        const float m1dz = 0.5f * (d1 - d2) * invdz;
        rsg_fwd([&](int dx, int dz) { return ring[2][(u + dx + 4 + kRing) % kRing][lane + dz]; }, d1, d2);
        const float m2dx = 0.5f * (d1 + d2) * invdx;
        const float m2dz = 0.5f * (d1 - d2) * invdz;
        const size_t o = (size_t)c * g.ldz + iz;
        const VAdjOps &op = opq[0];
        if (FORM != FDMOD_FORM_COORD) {
          a.mvx[o] = op.k1 * op.mvx - (m0dx + m2dz);
          a.mvz[o] = op.k1 * op.mvz - (m2dx + m1dz);
        } else {
          rsg_fwd([&](int dx, int dz) { return ring[NM - 1][(u + dx + 4 + kRing) % kRing][lane + dz]; }, d1, d2);
          const float m3dz = 0.5f * (d1 - d2) * invdz;
          a.mvx[o] = op.k1 * op.mvx - (m0dx + m1dz);
          a.mvz[o] = op.k1 * op.mvz - (m2dx + m3dz);
        }
      }
#pragma unroll
      for (int i = 0; i + 1 < PD; i++) {
        colq[i] = colq[i + 1];
        opq[i] = opq[i + 1];
      }
      colq[PD - 1] = cn;
      opq[PD - 1] = on;
    });
  }
}

// ---------------------------------------------------------------- V-sweep adjoint
// Staged per row: mu_v (2), coef2, invrho and, for the coordinate form, the averaged
// trig (cosv, sinv).  The staged fields N are the adjoints of the stress gradients
// the velocity sweep reads, with a = coef2 invrho mu_vx, b = coef2 invrho mu_vz:
//   Bond, tensor  (a, b)                                                  2 fields
//   coordinate    (P, Q, R, S) = (cv a, sv a, sv b, cv b)                  4 fields
// The coordinate form's six adjoints, of (sxxdx, sxxdz, sxzdx, sxzdz, szzdx, szzdz),
// are (P, -Q, Q + S, P - R, R, S); the stencil is linear, so they are formed from the
// four fields' derivatives.  Four staged fields instead of six cut the LDS ring from
// 19.3 to 12.8 KB per wave: 3 waves per SIMD instead of 2.
struct NRow {
  float vx, vz, k2, ir, cv, sv;
};
struct NCol {
  NRow lo, hi;
};

template <int FORM>
__device__ __forceinline__ void load_nrow(NRow &r, const AdjArgs &a, size_t o, bool und)
{
  r = NRow{a.mvx[o], a.mvz[o], und ? a.g.dt : a.coef2[o], a.invrho[o], 0.0f, 0.0f};
  if (FORM == FDMOD_FORM_COORD) {
    r.cv = a.csv[o];
    r.sv = a.snv[o];
  }
}

// Staged column rows rs+lane (lo) and, for lanes < kHalo, rs+64+lane (hi), loaded with no
// branch around them from rows clamped into the grid (r0c, r1c); store_ncol zeroes the
// rows outside the grid (ok0, ok1).
template <int FORM>
__device__ __forceinline__ void load_ncol(NCol &c, const AdjArgs &a, int gx, int r0c, int r1c, const Grid &g, bool und)
{
  const size_t base = (size_t)gx * g.ldz;
  load_nrow<FORM>(c.lo, a, base + r0c, und);
  load_nrow<FORM>(c.hi, a, base + r1c, und);
}

template <int FORM>
constexpr int n_fields()
{
  return FORM == FDMOD_FORM_COORD ? 4 : 2;
}

template <int FORM>
__device__ __forceinline__ void n_of(const NRow &r, float (&n)[6])
{
  const float kr = r.k2 * r.ir;
  const float av = kr * r.vx, bv = kr * r.vz;
  if (FORM != FDMOD_FORM_COORD) {
    n[0] = av;
    n[1] = bv;
    return;
  }
  n[0] = r.cv * av;  // P
  n[1] = r.sv * av;  // Q
  n[2] = r.sv * bv;  // R
  n[3] = r.cv * bv;  // S
}

template <int FORM>
__device__ __forceinline__ void store_ncol(float (*ring)[kRing][LZ], int slot, const NCol &c, int lane, bool ok0,
                                           bool ok1)
{
  constexpr int NN = n_fields<FORM>();
  float n[6];
  n_of<FORM>(c.lo, n);
#pragma unroll
  for (int f = 0; f < NN; f++) ring[f][slot][lane] = ok0 ? n[f] : 0.0f;
  if (lane < kHalo) {
    n_of<FORM>(c.hi, n);
#pragma unroll
    for (int f = 0; f < NN; f++) ring[f][slot][TZ + lane] = ok1 ? n[f] : 0.0f;
  }
}

// Per-point operands of the velocity adjoint as loaded; anything computed from them is
// computed where they are used, so that no instruction waits on a look-ahead load in the
// iteration that issues it.
struct SAdjOps {
  float s1, s2, s3, k1, t1, t2, t3, S2, K2, S2T, c2, ir, mvx, mvz, r;
};

// mu_sigma(q) <- k1(q) mu_sigma(q) - (stress-gradient adjoints)(q), with k1 = coef1 in U
// and 1 on the image rows above a free surface (rows [g.izStart, a.izU)).  Tensor form:
// sigma is overwritten on U, so there mu_sigma <- -(...), and mu_tau <- coef1 a with
// a = Q^T mu_sigma + mu_tau.  With a.snap set, the incoming values the imaging of this
// step needs are copied there (points of U): mu_sigma, then a (tensor) or (a, b)
// (coordinate).
template <int FORM, int PD, bool INJ>
__device__ __forceinline__ void adj_v_tile(const AdjArgs &a, int btz, int btx)
{
  constexpr int NN = n_fields<FORM>();
  __shared__ float ring[NN][kRing][LZ];
  const Grid &g = a.g;
  const int segL = g.segLS;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * TZ;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const int iz = z0 + lane;
  const bool zok = iz < g.izEnd;
  const bool inU = iz >= a.izU;
  const int rs = z0 - 5;
  const int xe = xa + nc - 1;
  const int izc = min(iz, g.Nz - 1);  // in-grid row for the per-point loads
  const bool irow = INJ && iz == a.inj_z;  // the receiver row (fused injection)
  // staged rows rs+lane and rs+64+lane, clamped into the grid for the loads
  const int r0 = rs + lane, r1 = rs + TZ + lane;
  const bool ok0 = r0 >= 0 && r0 < g.Nz, ok1 = lane < kHalo && r1 >= 0 && r1 < g.Nz;
  const int r0c = min(max(r0, 0), g.Nz - 1), r1c = ok1 ? r1 : r0c;
  // staged rows rs .. rs+72 of columns xa-5 .. xa+nc+3
  const bool und = region_undamped(a, rs, rs + LZ, xa - 5, xa + nc + 4);
#pragma unroll
  for (int m0 = 0; m0 < 9; m0 += 3) {
    NCol c[3];
#pragma unroll
    for (int m = 0; m < 3; m++) load_ncol<FORM>(c[m], a, xa - 5 + m0 + m, r0c, r1c, g, und);
#pragma unroll
    for (int m = 0; m < 3; m++) store_ncol<FORM>(ring, m0 + m, c[m], lane, ok0, ok1);
  }
  NCol colq[PD];
  SAdjOps opq[PD];
  // per-point operands, loaded with no branch around them (selected where used)
  auto load_ops = [&](SAdjOps &p, int col) {
    const size_t o = (size_t)col * g.ldz + izc;
    p = SAdjOps{};
    p.s1 = a.mxx[o];
    p.s2 = a.mzz[o];
    p.s3 = a.mxz[o];
    p.k1 = und ? 1.0f : a.coef1[o];
    if (INJ) p.r = a.inj[min(max(col - a.inj_x0, 0), a.inj_n - 1)];  // one address per wave
    if (FORM == FDMOD_FORM_TENSOR) {
      p.t1 = a.mtxx[o];
      p.t2 = a.mtzz[o];
      p.t3 = a.mtxz[o];
      p.S2 = a.sin2[o];
      p.K2 = a.cos2[o];
      p.S2T = a.sin2t[o];
    }
    if (FORM == FDMOD_FORM_COORD && a.snap[3]) {
      p.c2 = und ? g.dt : a.coef2[o];
      p.ir = a.invrho[o];
      p.mvx = a.mvx[o];
      p.mvz = a.mvz[o];
    }
  };
  // look-ahead loads from clamped columns (values past the segment are not used)
#pragma unroll
  for (int i = 0; i < PD; i++) {
    load_ncol<FORM>(colq[i], a, min(xa + 4 + i, xe + 4), r0c, r1c, g, und);
    load_ops(opq[i], min(xa + i, xe));
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = xa + k;
      NCol cn;
      SAdjOps on;
      load_ncol<FORM>(cn, a, min(c + 4 + PD, xe + 4), r0c, r1c, g, und);
      load_ops(on, min(c + PD, xe));
      store_ncol<FORM>(ring, (u + 9) % kRing, colq[0], lane, ok0, ok1);
      __syncthreads();
      if (zok) {
        // d1, d2 of staged field f; then x- and z-derivative forms
        float dd1[NN], dd2[NN];
#pragma unroll
        for (int f = 0; f < NN; f++)
          rsg_bwd([&](int dx, int dz) { return ring[f][(u + dx + 5 + kRing) % kRing][lane + 5 + dz]; }, dd1[f],
                  dd2[f]);
        const size_t o = (size_t)c * g.ldz + iz;
        const SAdjOps &op = opq[0];
        const float k1 = inU ? op.k1 : 1.0f;  // 1 on the image rows above U
        // fused injection: the receiver's mu_sxx and mu_szz take 0.5 r, as the stress adjoint
        // staged them (the slot keeps the values without it; the imaging adds it again)
        float s1 = op.s1, s2 = op.s2;
        if (INJ) {  // a select, not a branch: a branch here makes the compiler wait on every load
          const bool hit = irow && (unsigned)(c - a.inj_x0) < (unsigned)a.inj_n;
          const float d = 0.5f * op.r;
          s1 = hit ? op.s1 + d : op.s1;
          s2 = hit ? op.s2 + d : op.s2;
        }
        float gxx, gzz, gxz;  // adjoint contributions to subtract
        if constexpr (FORM != FDMOD_FORM_COORD) {
          const float nxdx = 0.5f * (dd1[0] + dd2[0]) * invdx, nxdz = 0.5f * (dd1[0] - dd2[0]) * invdz;
          const float nzdx = 0.5f * (dd1[1] + dd2[1]) * invdx, nzdz = 0.5f * (dd1[1] - dd2[1]) * invdz;
          gxx = nxdx;
          gzz = nzdz;
          gxz = nxdz + nzdx;
        } else {
          // fields P, Q, R, S; sxxdz carries the reference's invdx (CR:633)
          gxx = 0.5f * (dd1[0] + dd2[0]) * invdx - 0.5f * (dd1[1] - dd2[1]) * invdx;
          gxz = 0.5f * ((dd1[1] + dd1[3]) + (dd2[1] + dd2[3])) * invdx +
                0.5f * ((dd1[0] - dd1[2]) - (dd2[0] - dd2[2])) * invdz;
          gzz = 0.5f * (dd1[2] + dd2[2]) * invdx + 0.5f * (dd1[3] - dd2[3]) * invdz;
        }
        if (FORM == FDMOD_FORM_TENSOR && inU) {
          const float a1 = op.K2 * s1 + op.S2 * s2 - 0.5f * op.S2T * op.s3 + op.t1;
          const float a2 = op.S2 * s1 + op.K2 * s2 + 0.5f * op.S2T * op.s3 + op.t2;
          const float a3 = op.S2T * s1 - op.S2T * s2 + (op.K2 - op.S2) * op.s3 + op.t3;
          a.mtxx[o] = k1 * a1;
          a.mtzz[o] = k1 * a2;
          a.mtxz[o] = k1 * a3;
          a.oxx[o] = -gxx;
          a.ozz[o] = -gzz;
          a.oxz[o] = -gxz;
          if (a.snap[3]) {  // read by the imaging up to kImgK steps later
            st_far(a.snap[3] + o, a1);
            st_far(a.snap[4] + o, a2);
            st_far(a.snap[5] + o, a3);
          }
        } else {
          a.oxx[o] = k1 * s1 - gxx;
          a.ozz[o] = k1 * s2 - gzz;
          a.oxz[o] = k1 * op.s3 - gxz;
          if (FORM == FDMOD_FORM_COORD && a.snap[3] && inU) {
            const float kr = op.c2 * op.ir;
            st_far(a.snap[3] + o, kr * op.mvx);
            st_far(a.snap[4] + o, kr * op.mvz);
          }
        }
      }
#pragma unroll
      for (int i = 0; i + 1 < PD; i++) {
        colq[i] = colq[i + 1];
        opq[i] = opq[i + 1];
      }
      colq[PD - 1] = cn;
      opq[PD - 1] = on;
    });
  }
}

// ---------------------------------------------------------------- FP16 storage
// Stored wavefields (checkpoints) in FP16 with one power-of-two scale per field,
// column and 64-row block: a block with maximum |v| < 2^E is stored as v * 2^(14-E)
// (exact scaling, |value| < 2^14 < 65504) and the block keeps the inverse scale
// 2^(E-14).  Arithmetic on the decoded values stays in FP32.  The normal stresses go in
// as their sum and difference (pairs, Pack16::mate): in a P-dominated wavefield sxx and
// szz (and the tensor form's txx and tzz) nearly cancel in the difference that the
// shear and rotation terms of the recomputed segment read, and the difference then
// keeps its own 11 significant bits (tensor form, theta gradient on the 1126 x 2000 crop
// against the FP32 path: 2.4e-3 with the fields stored alone).

// One wave per (column, block); blockIdx.y = field.  Rows [Nz, ldz) hold zeros.
__global__ __launch_bounds__(256) void k_pack16(Pack16 p)
{
  const int lane = threadIdx.x & 63;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nblk = p.ldz / TZ;
  if (w >= (long)p.nx * nblk) return;
  const int f = blockIdx.y;
  const int ix = (int)(w / nblk), b = (int)(w - (long)ix * nblk);
  const size_t o = (size_t)ix * p.ldz + (size_t)b * TZ + lane;
  const int g = p.mate[f];
  float v = p.src[f][o];
  if (g >= 0) v = f < g ? v + p.src[g][o] : p.src[g][o] - v;
  const float m = wave_max64(fabsf(v));  // every lane of the wave is active here
  int E = 0;
  if (m > 0.0f) (void)frexpf(m, &E);  // m < 2^E
  E = max(E, -100);                   // keeps 2^(14-E) and 2^(E-14) normal floats
  p.dst[f][o] = (_Float16)(v * ldexpf(1.0f, 14 - E));
  if (lane == 0) p.inv[f][(size_t)ix * nblk + b] = ldexpf(1.0f, E - 14);
}

__global__ __launch_bounds__(256) void k_unpack16(Pack16 p)
{
  const int lane = threadIdx.x & 63;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nblk = p.ldz / TZ;
  if (w >= (long)p.nx * nblk) return;
  const int f = blockIdx.y;
  const int ix = (int)(w / nblk), b = (int)(w - (long)ix * nblk);
  const size_t o = (size_t)ix * p.ldz + (size_t)b * TZ + lane;
  const size_t sb = (size_t)ix * nblk + b;
  const float v = (float)p.dst[f][o] * p.inv[f][sb];
  const int g = p.mate[f];
  if (g < 0) {
    p.src[f][o] = v;
    return;
  }
  const float u = (float)p.dst[g][o] * p.inv[g][sb];
  p.src[f][o] = f < g ? 0.5f * (v + u) : 0.5f * (u - v);
}

void launch_pack16(hipStream_t s, const Pack16 &p, bool pack)
{
  const long waves = (long)p.nx * (p.ldz / TZ);
  dim3 grid((unsigned)((waves + 3) / 4), p.nf);
  if (pack) hipLaunchKernelGGL(k_pack16, grid, dim3(256), 0, s, p);
  else hipLaunchKernelGGL(k_unpack16, grid, dim3(256), 0, s, p);
}

// ---------------------------------------------------------------- imaging
// Per point of U, sum over a block of nk steps the products of the adjoint snapshots
// (written by the V-adjoint) and the stored forward history (written by the forward
// sweeps), into the FP64 imaging sums G.  One thread per point keeps the block's sums
// in FP64 registers, so G is read and written once per block of steps.
//   Bond    G = c11, c13, c15, c33, c35, c55 sums of mu_sigma (x) E, E = (vxdx, vzdz,
//           vzdx + vxdz); coef2 multiplies them in the chain-rule kernel.
//   tensor  G = c11, c13, c33, c55, sin^2, cos^2, sin 2theta (TR:826-852), from the
//           snapshots mu_sigma and a = Q^T mu_sigma + mu_tau, the strain and tau_{k+1}.
//   coord   G = c11, c13, c33, c55, sin, cos at stress points (CR:821-843) and sinv,
//           cosv at velocity points (CR:663-670).
template <int FORM, bool F16, bool INJ>
__global__ __launch_bounds__(256) void k_image(AdjArgs a, ImgBlock b)
{
  const Grid &g = a.g;
  const int iz = g.izStart + blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = g.ixStart + blockIdx.y;
  if (iz >= g.izEnd) return;
  const size_t o = (size_t)ix * g.ldz + iz;
  const size_t so = (size_t)ix * g.tilesZ + (iz - g.izStart) / TZ;
  constexpr int NH = FORM == FDMOD_FORM_BOND ? 3 : (FORM == FDMOD_FORM_TENSOR ? 6 : 8);
  constexpr int NL = FORM == FDMOD_FORM_BOND ? 3 : (FORM == FDMOD_FORM_TENSOR ? 6 : 5);
  constexpr int NG = FORM == FDMOD_FORM_BOND ? 6 : (FORM == FDMOD_FORM_TENSOR ? 7 : 8);
  double acc[NG];
#pragma unroll
  for (int q = 0; q < NG; q++) acc[q] = 0.0;
  float P[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // time-independent parameters of the point
  if (FORM == FDMOD_FORM_TENSOR) {
    P[0] = a.c11[o]; P[1] = a.c13[o]; P[2] = a.c33[o]; P[3] = a.c55[o]; P[4] = a.coef2[o];
    P[5] = a.sin2[o]; P[6] = a.cos2[o]; P[7] = a.sin2t[o];
  } else if (FORM == FDMOD_FORM_COORD) {
    P[0] = a.c11[o]; P[1] = a.c13[o]; P[2] = a.c33[o]; P[3] = a.c55[o]; P[4] = a.coef2[o];
    P[5] = a.sn[o]; P[6] = a.cs[o];
  }
  // fully unrolled (compile-time j): the block's pointers then come from fixed kernel-argument offsets
  static_for<0, kImgK>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j >= b.nk) return;
    double L[NL], H[NH];
    float l[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) l[q] = b.lam[j][q][o];
    if (INJ) {
      // steps whose pressure residuals the sweeps added themselves (the slot is without
      // them): selects and an unconditional load, no branch, so that the unrolled steps'
      // loads stay in flight together
      const bool on = b.inj[j] != nullptr;
      const float *ip = on ? b.inj[j] : b.lam[j][0];
      const float r = ip[on ? min(max(ix - b.inj_x0, 0), b.inj_n - 1) : 0];
      const bool hit = on && iz == b.inj_z && (unsigned)(ix - b.inj_x0) < (unsigned)b.inj_n;
      const float d = 0.5f * r;
      l[0] = hit ? l[0] + d : l[0];
      l[1] = hit ? l[1] + d : l[1];
    }
#pragma unroll
    for (int q = 0; q < NL; q++) L[q] = l[q];
    if (F16) {
#pragma unroll
      for (int q = 0; q < NH; q++) H[q] = (float)b.h16[j][q][o] * b.hsc[j][q * b.hsc_stride + so];
    } else {
#pragma unroll
      for (int q = 0; q < NH; q++) H[q] = b.h32[j][q][o];
    }
    if (FORM == FDMOD_FORM_BOND) {
      const double X = L[0], Z = L[1], Y = L[2], ex = H[0], ez = H[1], eg = H[2];
      acc[0] += X * ex;
      acc[1] += X * ez + Z * ex;
      acc[2] += X * eg + Y * ex;
      acc[3] += Z * ez;
      acc[4] += Z * eg + Y * ez;
      acc[5] += Y * eg;
    } else if (FORM == FDMOD_FORM_TENSOR) {
      const double lx = L[0], lz = L[1], lc = L[2], a1 = L[3], a2 = L[4], a3 = L[5];
      // the history holds (vxdx + vzdz, vxdx - vzdz, shear, txx + tzz, txx - tzz, txz)
      const double d = 0.5 * (H[0] + H[1]), z = 0.5 * (H[0] - H[1]), gg = H[2];
      const double t1 = 0.5 * (H[3] + H[4]), t2 = 0.5 * (H[3] - H[4]), t3 = H[5];
      const double k2 = P[4], S2 = P[5], K2 = P[6], S2T = P[7];
      const double exx = K2 * d + S2 * z - 0.5 * S2T * gg, ezz = S2 * d + K2 * z + 0.5 * S2T * gg;
      const double exz = S2T * (d - z) + (K2 - S2) * gg;
      const double mexx = k2 * ((double)P[0] * a1 + (double)P[1] * a2);
      const double mezz = k2 * ((double)P[1] * a1 + (double)P[2] * a2);
      const double mexz = k2 * (double)P[3] * a3;
      acc[0] += k2 * a1 * exx;
      acc[1] += k2 * (a1 * ezz + a2 * exx);
      acc[2] += k2 * a2 * ezz;
      acc[3] += k2 * a3 * exz;
      acc[4] += mexx * z + mezz * d - mexz * gg + lx * t2 + lz * t1 - lc * t3;
      acc[5] += mexx * d + mezz * z + mexz * gg + lx * t1 + lz * t2 + lc * t3;
      acc[6] += -0.5 * mexx * gg + 0.5 * mezz * gg + mexz * (d - z) + lx * t3 - lz * t3 - 0.5 * lc * (t1 - t2);
    } else {
      const double lx = L[0], lz = L[1], lc = L[2];
      const double X = H[0], XZ = H[1], ZX = H[2], Z = H[3];
      const double k2 = P[4], sn = P[5], cs = P[6];
      const double pxdx = cs * X - sn * XZ, pxdz = sn * X + cs * XZ;
      const double pzdx = cs * ZX - sn * Z, pzdz = sn * ZX + cs * Z;
      const double rxdx = k2 * ((double)P[0] * lx + (double)P[1] * lz);
      const double rzdz = k2 * ((double)P[1] * lx + (double)P[2] * lz);
      const double rg = k2 * (double)P[3] * lc;
      acc[0] += k2 * lx * pxdx;
      acc[1] += k2 * (lx * pzdz + lz * pxdx);
      acc[2] += k2 * lz * pzdz;
      acc[3] += k2 * lc * (pzdx + pxdz);
      acc[4] += rxdx * (-XZ) + rg * X + rg * (-Z) + rzdz * ZX;
      acc[5] += rxdx * X + rg * XZ + rg * ZX + rzdz * Z;
      acc[6] += L[3] * H[4] + L[4] * H[5];  // sinv: a (-sxxdz + sxzdx) + b (-sxzdz + szzdx)
      acc[7] += L[3] * H[6] + L[4] * H[7];  // cosv: a (sxxdx + sxzdz) + b (sxzdx + szzdz)
    }
  });
#pragma unroll
  for (int q = 0; q < NG; q++) a.G[q][o] += acc[q];
}

// Prefetch depth (iterations) of both adjoint sweeps.  Measured on the 1126 x 2000 crop,
// nt = 4000, FP16-history gradient: depth 1 took 0.556 / 0.800 / 0.730 s (Bond / tensor /
// coordinate) against 0.565 / 0.809 / 0.729 s at depth 2 and 0.569 s (Bond) at depth 3.
#ifndef FDMOD_PIPE_ADJ
#define FDMOD_PIPE_ADJ 1
#endif
constexpr int kPipeAdj = FDMOD_PIPE_ADJ;

template <int FORM, bool INJ>
__global__ __launch_bounds__(64) void k_adj_ssweep(AdjArgs a)
{
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLV, btz, btx)) return;
  adj_s_tile<FORM, kPipeAdj, INJ>(a, btz, btx);
}

template <int FORM, bool INJ>
__global__ __launch_bounds__(64) void k_adj_vsweep(AdjArgs a)
{
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLS, btz, btx)) return;
  adj_v_tile<FORM, kPipeAdj, INJ>(a, btz, btx);
}

static dim3 tiles_grid(const Grid &g, int segL)
{
  const int n = g.tilesZ * ((g.ixEnd - g.ixStart + segL - 1) / segL);
  return dim3((n + kXCD - 1) / kXCD * kXCD);
}

template <bool INJ>
static void launch_adj_s(hipStream_t s, dim3 grid, const AdjArgs &a)
{
  if (a.form == FDMOD_FORM_COORD) hipLaunchKernelGGL((k_adj_ssweep<FDMOD_FORM_COORD, INJ>), grid, dim3(64), 0, s, a);
  else if (a.form == FDMOD_FORM_TENSOR)
    hipLaunchKernelGGL((k_adj_ssweep<FDMOD_FORM_TENSOR, INJ>), grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL((k_adj_ssweep<FDMOD_FORM_BOND, INJ>), grid, dim3(64), 0, s, a);
}

template <bool INJ>
static void launch_adj_v(hipStream_t s, dim3 grid, const AdjArgs &a)
{
  if (a.form == FDMOD_FORM_COORD) hipLaunchKernelGGL((k_adj_vsweep<FDMOD_FORM_COORD, INJ>), grid, dim3(64), 0, s, a);
  else if (a.form == FDMOD_FORM_TENSOR)
    hipLaunchKernelGGL((k_adj_vsweep<FDMOD_FORM_TENSOR, INJ>), grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL((k_adj_vsweep<FDMOD_FORM_BOND, INJ>), grid, dim3(64), 0, s, a);
}

void launch_adj_ssweep(hipStream_t s, const AdjArgs &a)
{
  const dim3 grid = tiles_grid(a.g, a.g.segLV);
  if (a.inj) launch_adj_s<true>(s, grid, a);
  else launch_adj_s<false>(s, grid, a);
}

void launch_adj_vsweep(hipStream_t s, const AdjArgs &a)
{
  const dim3 grid = tiles_grid(a.g, a.g.segLS);
  if (a.inj) launch_adj_v<true>(s, grid, a);
  else launch_adj_v<false>(s, grid, a);
}

void launch_image(hipStream_t s, const AdjArgs &a, const ImgBlock &b)
{
  dim3 grid((a.g.izEnd - a.g.izStart + 255) / 256, a.g.ixEnd - a.g.ixStart);
  const bool f16 = b.h16[0][0] != nullptr;
  bool inj = false;
  for (int j = 0; j < b.nk; j++) inj = inj || b.inj[j] != nullptr;
#define FDMOD_IMG(F)                                                                                \
  do {                                                                                             \
    if (f16 && inj) hipLaunchKernelGGL((k_image<F, true, true>), grid, dim3(256), 0, s, a, b);      \
    else if (f16) hipLaunchKernelGGL((k_image<F, true, false>), grid, dim3(256), 0, s, a, b);      \
    else if (inj) hipLaunchKernelGGL((k_image<F, false, true>), grid, dim3(256), 0, s, a, b);      \
    else hipLaunchKernelGGL((k_image<F, false, false>), grid, dim3(256), 0, s, a, b);              \
  } while (0)
  if (a.form == FDMOD_FORM_COORD) FDMOD_IMG(FDMOD_FORM_COORD);
  else if (a.form == FDMOD_FORM_TENSOR) FDMOD_IMG(FDMOD_FORM_TENSOR);
  else FDMOD_IMG(FDMOD_FORM_BOND);
#undef FDMOD_IMG
}

// ---------------------------------------------------------------- free surface
// Transpose of the stress imaging (BT:875-891): row izU-k was set to
// (2 sxx(izU) - sxx(izU+k), -szz(izU+k), -sxz(izU+k)) and szz, sxz of row izU to 0.
__global__ void k_adj_fs(AdjArgs a)
{
  const int ix = a.g.ixStart + blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= a.g.ixEnd) return;
  const size_t b = (size_t)ix * a.g.ldz;
  const int z0 = a.izU;
  float acc = 0.0f;
  for (int k = kN; k >= 1; k--) {
    const size_t m = b + z0 - k, p = b + z0 + k;
    const float x = a.mxx[m], z = a.mzz[m], y = a.mxz[m];
    acc += 2.0f * x;
    a.mxx[p] -= x;
    a.mzz[p] -= z;
    a.mxz[p] -= y;
    a.mxx[m] = 0.0f;
    a.mzz[m] = 0.0f;
    a.mxz[m] = 0.0f;
  }
  a.mxx[b + z0] += acc;
  a.mzz[b + z0] = 0.0f;
  a.mxz[b + z0] = 0.0f;
}

void launch_adj_fs(hipStream_t s, const AdjArgs &a)
{
  const int n = a.g.ixEnd - a.g.ixStart;
  hipLaunchKernelGGL(k_adj_fs, dim3((n + 255) / 256), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------- data-space sources
// Transpose of elastic2dReceiver (isoelastic.c:103-144) for one sample: pressure
// residuals go half to mu_sxx and half to mu_szz, velocity residuals to mu_vx, mu_vz.
// Receivers may share a cell, hence the atomics.
__global__ void k_adj_inject(int nr, int ldz, const int *gx, const int *gz, const float *r1, const float *r2,
                             float *f1, float *f2, float scale)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const size_t o = (size_t)gx[r] * ldz + gz[r];
  atomicAdd(f1 + o, scale * r1[r]);
  atomicAdd(f2 + o, scale * r2[r]);
}

void launch_adj_inject(hipStream_t s, int nr, int ldz, const int *gx, const int *gz, const float *r1,
                       const float *r2, float *f1, float *f2, float scale)
{
  hipLaunchKernelGGL(k_adj_inject, dim3((nr + 255) / 256), dim3(256), 0, s, nr, ldz, gx, gz, r1, r2, f1, f2, scale);
}

// Transpose of the source injection of one step (source_body): the adjoint of
// src[is][it] is the tap-weighted sum of the adjoint fields the taps wrote into.
// One block per source; the sum is in double.
__global__ void k_adj_source(SourceArgs a, const float *ma, const float *mb, const float *mc, const float *md,
                             double *out)
{
  __shared__ double red[128];
  const Grid &g = a.g;
  const int is = blockIdx.x;
  double acc = 0.0;
  for (int p = 0; p < a.npass; p++)
    for (int t = threadIdx.x; t < a.ntap; t += blockDim.x) {
      const SrcTap tp = a.taps[p * a.ntap + t];
      const int ix = a.isx[is] + tp.ox, iz = a.isz[is] + tp.oz;
      if (ix < 0 || ix >= g.Nx || iz < 0 || iz >= g.Nz) continue;
      const size_t o = (size_t)ix * g.ldz + iz;
      acc += (double)tp.ca * ma[o];
      if (mb) acc += (double)tp.cb * mb[o];
      if (mc) acc += (double)tp.ca * mc[o];  // tensor form: the stress source also fed tau
      if (md) acc += (double)tp.cb * md[o];
    }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 64; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(size_t)is * a.nt + a.it] = red[0];
}

void launch_adj_source(hipStream_t s, const SourceArgs &a, const float *ma, const float *mb, const float *mc,
                       const float *md, double *out)
{
  hipLaunchKernelGGL(k_adj_source, dim3(a.ns), dim3(128), 0, s, a, ma, mb, mc, md, out);
}

// ---------------------------------------------------------------- FP64 reductions
// r = d - obs (in place over obs) and per-block partial sums of r^2 / 2.
__global__ void k_residual(const float *d, float *r, size_t n, double *partial)
{
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = d[i] - r[i];
    r[i] = v;
    acc += 0.5 * (double)v * (double)v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// partial sums of a . b with a float and b float (bd == null) or double
__global__ void k_dot(const float *a, const float *bf, const double *bd, size_t n, double *partial)
{
  __shared__ double red[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (bd ? bd[i] : (double)bf[i]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] += sum of nb partials, in a fixed order
__global__ void k_sum_partials(const double *partial, int nb, double *out)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += partial[b];
  out[0] += s;
}

void launch_residual(hipStream_t s, const float *d, float *r, size_t n, double *partial, int nb, double *out)
{
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, d, r, n, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nb, out);
}

void launch_dot(hipStream_t s, const float *a, const float *bf, const double *bd, size_t n, double *partial, int nb,
                double *out)
{
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(256), 0, s, a, bf, bd, n, partial);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, nb, out);
}

// ---------------------------------------------------------------- chain rule
// Forward-mode dual numbers over the four model directions (vp, eps, delta, theta).
struct D4 {
  double v, d[4];
};
__device__ __forceinline__ D4 cst(double v) { return D4{v, {0, 0, 0, 0}}; }
__device__ __forceinline__ D4 seed(double v, int k)
{
  D4 r = cst(v);
  r.d[k] = 1.0;
  return r;
}
__device__ __forceinline__ D4 operator+(D4 a, D4 b)
{
  for (int k = 0; k < 4; k++) a.d[k] += b.d[k];
  a.v += b.v;
  return a;
}
__device__ __forceinline__ D4 operator-(D4 a, D4 b)
{
  for (int k = 0; k < 4; k++) a.d[k] -= b.d[k];
  a.v -= b.v;
  return a;
}
__device__ __forceinline__ D4 operator*(D4 a, D4 b)
{
  D4 r;
  r.v = a.v * b.v;
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D4 operator*(double s, D4 a)
{
  a.v *= s;
  for (int k = 0; k < 4; k++) a.d[k] *= s;
  return a;
}
__device__ __forceinline__ D4 operator/(D4 a, D4 b)
{
  D4 r;
  r.v = a.v / b.v;
  for (int k = 0; k < 4; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
__device__ __forceinline__ D4 dsqrt(D4 a)
{
  D4 r;
  r.v = sqrt(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
__device__ __forceinline__ D4 dsin(D4 a)
{
  D4 r;
  r.v = sin(a.v);
  const double c = cos(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = c * a.d[k];
  return r;
}
__device__ __forceinline__ D4 dcos(D4 a)
{
  D4 r;
  r.v = cos(a.v);
  const double s = -sin(a.v);
  for (int k = 0; k < 4; k++) r.d[k] = s * a.d[k];
  return r;
}
__device__ __forceinline__ bool bad(double v) { return isnan(v) || isinf(v); }

// Thomsen -> Bond-rotated stiffness (BT:493-526) with the Bond free-surface fix-up
// (BT:543-568) on the surface row, as exact functions of (vp, eps, delta, theta).
__device__ void bond_stiffness(D4 vp, double vs, double rho, D4 eps, D4 del, D4 t, bool fixup, D4 c[6])
{
  const D4 r = cst(vs) / vp;
  const D4 f = cst(1.0) - r * r;
  const D4 C11 = rho * (vp * vp) * (cst(1.0) + 2.0 * eps);
  const D4 C13 = rho * (vp * vp * dsqrt(f * (f + 2.0 * del)) - cst(vs * vs));
  const D4 C33 = rho * (vp * vp);
  const D4 C55 = cst(rho * vs * vs);
  const D4 s = dsin(t), co = dcos(t);
  const D4 s2 = s * s, c2 = co * co, s4 = s2 * s2, c4 = c2 * c2;
  const D4 sn2 = dsin(2.0 * t), cs2 = dcos(2.0 * t), sn22 = sn2 * sn2, cs22 = cs2 * cs2;
  c[0] = C11 * c4 + C33 * s4 + (0.5 * C13 + C55) * sn22;
  c[1] = (0.25 * (C11 + C33) - C55) * sn22 + C13 * (s4 + c4);
  c[2] = 0.5 * ((cst(0.0) - C11 * c2) + C33 * s2 + (C13 + 2.0 * C55) * cs2) * sn2;
  c[3] = C11 * s4 + C33 * c4 + (0.5 * C13 + C55) * sn22;
  c[4] = 0.5 * ((cst(0.0) - C11 * s2) + C33 * c2 - (C13 + 2.0 * C55) * cs2) * sn2;
  c[5] = 0.25 * (C11 + C33 - 2.0 * C13) * sn22 + C55 * cs22;
  if (fixup) {
    const D4 c11 = c[0], c13 = c[1], c15 = c[2], c33 = c[3], c35 = c[4], c55 = c[5];
    D4 n35 = (cst(0.0) - (c13 * c55 - c15 * c35)) / (c33 * c55 - c35 * c35);
    const D4 n55 = (c13 * n35 - c15 * c33) / (c33 * c55 - n35 * n35);
    if (bad(n35.v)) n35 = (cst(0.0) - c13) / c33;
    if (bad(n35.v)) n35 = cst(0.0);
    c[0] = c11 - c13 * n35 + c15 * n55;
    for (int j = 1; j < 6; j++) c[j] = cst(0.0);
  }
}

// Tensor and coordinate forms: the per-point arrays the loop reads (TR:483-517,
// CR:497-513), VTI stiffness then trig, as exact functions of (vp, eps, delta, theta).
__device__ int vti_params(int form, D4 vp, double vs, double rho, D4 eps, D4 del, D4 t, D4 c[8])
{
  const D4 r = cst(vs) / vp;
  const D4 f = cst(1.0) - r * r;
  c[0] = rho * (vp * vp) * (cst(1.0) + 2.0 * eps);
  c[1] = rho * (vp * vp * dsqrt(f * (f + 2.0 * del)) - cst(vs * vs));
  c[2] = rho * (vp * vp);
  c[3] = cst(rho * vs * vs);
  if (form == FDMOD_FORM_TENSOR) {
    const D4 sn = dsin(t), co = dcos(t);
    c[4] = sn * sn;
    c[5] = co * co;
    c[6] = dsin(2.0 * t);
    return 7;
  }
  c[4] = dsin(t);
  c[5] = dcos(t);
  return 6;
}

// dJ/d(vp, eps, delta, theta[deg]) on the unpadded model, in two passes.  Pass 1: at
// every padded point, the chain rule of the point's imaging sums through its parameters
// (Bond: G holds the sums without coef2, multiplied here), written over G[0..3] of the
// point once the point's sums are read.  Pass 2: each unpadded cell adds the values of
// the padded points that replicate it (the transpose of padModel2d,
// L/absorbc/padmod.c:33-58), columns outer, rows inner.  (One pass with the replicas
// summed in place left the corner cells' threads evaluating the FP64 trig of 66 x 66
// points in series: 5.3 ms of a 0.52 s gradient on the 1126 x 2000 crop.)
__global__ void k_chain_point(Grid g, int form, int fs_row, DevArrays d, double *G, size_t gstride,
                              double theta_scale)
{
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= g.Nz) return;
  const int ng = form == FDMOD_FORM_BOND ? 6 : (form == FDMOD_FORM_TENSOR ? 7 : 6);
  const size_t o = (size_t)ix * g.ldz + iz;
  double gc[8];
  bool any = false;
  const double k2 = form == FDMOD_FORM_BOND ? (double)d.a[FDMOD_A_COEF2][o] : 1.0;
  for (int j = 0; j < ng; j++) {
    gc[j] = k2 * G[j * gstride + o];
    any = any || gc[j] != 0.0;
  }
  double out[4] = {0.0, 0.0, 0.0, 0.0};
  if (any) {
    D4 c[8];
    const D4 vp = seed(d.a[FDMOD_A_VP][o], 0), eps = seed(d.a[FDMOD_A_EPS][o], 1);
    const D4 del = seed(d.a[FDMOD_A_DELTA][o], 2), th = seed(d.a[FDMOD_A_THETA][o], 3);
    const double vs = d.a[FDMOD_A_VS][o], rho = d.a[FDMOD_A_RHO][o];
    if (form == FDMOD_FORM_BOND) bond_stiffness(vp, vs, rho, eps, del, th, iz == fs_row, c);
    else vti_params(form, vp, vs, rho, eps, del, th, c);
    const double beta = d.a[FDMOD_A_BETA][o];
    for (int m = 0; m < 4; m++) {
      double s = 0.0;
      for (int j = 0; j < ng; j++) s += gc[j] * c[j].d[m];
      out[m] = m == 0 ? s / beta : (m == 3 ? s * theta_scale : s);
    }
  }
  for (int m = 0; m < 4; m++) G[m * gstride + o] = out[m];
}

__global__ void k_chain_gather(Grid g, int nx, int nz, const double *P, size_t gstride, double *grad)
{
  const int jz = blockIdx.x * blockDim.x + threadIdx.x;
  const int jx = blockIdx.y;
  if (jz >= nz || jx >= nx) return;
  const int x0 = jx == 0 ? 0 : jx + g.e, x1 = jx == nx - 1 ? g.Nx - 1 : jx + g.e;
  const int z0 = jz == 0 ? 0 : jz + g.e, z1 = jz == nz - 1 ? g.Nz - 1 : jz + g.e;
  double acc[4] = {0, 0, 0, 0};
  for (int ix = x0; ix <= x1; ix++)
    for (int iz = z0; iz <= z1; iz++) {
      const size_t o = (size_t)ix * g.ldz + iz;
      for (int m = 0; m < 4; m++) acc[m] += P[m * gstride + o];
    }
  for (int m = 0; m < 4; m++) grad[((size_t)m * nx + jx) * nz + jz] = acc[m];
}

void launch_chain(hipStream_t s, const Grid &g, int form, int nx, int nz, int fs_row, const DevArrays &d,
                  double *G, size_t gstride, double theta_scale, double *grad)
{
  hipLaunchKernelGGL(k_chain_point, dim3((g.Nz + 127) / 128, g.Nx), dim3(128), 0, s, g, form, fs_row, d, G, gstride,
                     theta_scale);
  hipLaunchKernelGGL(k_chain_gather, dim3((nz + 127) / 128, nx), dim3(128), 0, s, g, nx, nz, G, gstride, grad);
}

// Coordinate form: the velocity-point sums of sinv, cosv (G6, G7) go back to the
// stress-point sin, cos (G4, G5) through the transpose of the 4-point average of
// CR:516-535 (sinv(x,z) = (sin(x,z)+sin(x+1,z)+sin(x,z+1)+sin(x+1,z+1))/4 below the
// last row and column, sin(x,z) on them).  Gather form, one thread per point.
__global__ void k_fold_trig(Grid g, double *G, size_t gstride)
{
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  const size_t o = (size_t)ix * g.ldz + iz;
  for (int q = 0; q < 2; q++) {
    const double *gv = G + (6 + q) * gstride;
    double acc = 0.0;
    for (int dx = 0; dx <= 1; dx++)
      for (int dz = 0; dz <= 1; dz++) {
        const int px = ix - dx, pz = iz - dz;
        if (px >= 0 && pz >= 0 && px < g.Nx - 1 && pz < g.Nz - 1) acc += 0.25 * gv[(size_t)px * g.ldz + pz];
      }
    if (ix == g.Nx - 1 || iz == g.Nz - 1) acc += gv[o];
    G[(4 + q) * gstride + o] += acc;
  }
}

void launch_fold_trig(hipStream_t s, const Grid &g, double *G, size_t gstride)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_fold_trig, grid, dim3(256), 0, s, g, G, gstride);
}

// Add one shot's crop gradient [4][nx][nz] into the model gradient [4][nx_full][nz]
// at column offset x0 (shots are added one after another, in shot order).
__global__ void k_accum_grad(const double *src, int nx, int nz, int x0, int nx_full, double *dst)
{
  const int jz = blockIdx.x * blockDim.x + threadIdx.x;
  const int jx = blockIdx.y, m = blockIdx.z;
  if (jz >= nz || jx >= nx) return;
  dst[((size_t)m * nx_full + x0 + jx) * nz + jz] += src[((size_t)m * nx + jx) * nz + jz];
}

void launch_accum_grad(hipStream_t s, const double *src, int nx, int nz, int x0, int nx_full, double *dst)
{
  hipLaunchKernelGGL(k_accum_grad, dim3((nz + 255) / 256, nx, 4), dim3(256), 0, s, src, nx, nz, x0, nx_full, dst);
}

// Frozen absorbing layers: a new padded model is rescaled by the stored Beta as in
// mealFDCoefs (meal.c:454-482: vp/=Beta, vs/=Beta, rho*=Beta), D and coefs kept.
__global__ void k_rescale_frozen(Grid g, float *vp, float *vs, float *rho, const float *Beta)
{
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= g.Nz || ix >= g.Nx) return;
  const size_t o = (size_t)ix * g.ldz + iz;
  const float b = Beta[o];
  vp[o] /= b;
  vs[o] /= b;
  rho[o] *= b;
}

void launch_rescale_frozen(hipStream_t s, const Grid &g, float *vp, float *vs, float *rho, const float *Beta)
{
  dim3 grid((g.Nz + 255) / 256, g.Nx);
  hipLaunchKernelGGL(k_rescale_frozen, grid, dim3(256), 0, s, g, vp, vs, rho, Beta);
}

}  // namespace fdm
