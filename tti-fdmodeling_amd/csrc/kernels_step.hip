// Time-step kernels: the two RSG sweeps (velocity from stress, stress from
// velocity) for the three rotation forms, source injection, receiver gathers,
// the pressure snapshot field and the energy diagnostic.
//
// Arithmetic follows the reference's float expressions term by term with FP
// contraction disabled at build time, so each updated value is the one the
// reference's C computes.
//
// Sweep structure.  One wavefront per block owns a strip of 64 rows (lane = row)
// and marches along x over a segment of columns.  The staged field(s) of the
// stencil live in an LDS ring of kRing columns of 73 rows (64 + the 9-row halo).
// Each iteration stores the column loaded one iteration earlier into the ring,
// issues the loads of the next column and of the next column's per-point
// operands, then evaluates one output column from the ring.  Column loads use a
// wave-uniform base (scalar address) plus the lane's row, so staging costs almost
// no vector ALU work, and memory traffic overlaps the stencil arithmetic of the
// column in flight.  The ring index of every stencil tap is a compile-time
// constant inside a kRing-times unrolled iteration block.
#include "sweep_common.h"

namespace fdm {

// Store one column of NF history fields f0 .. f0+NF-1 for the gradient (all lanes call
// it; rows past izEnd pass zeros).  FP16: the NF fields are stored as v * 2^(14-E) with
// E the exponent of the largest |v| of any of them in the column's 64-row strip
// (|stored| < 2^14), and the strip keeps 2^(E-14) in each field's scale array.  One
// scale for the fields stored together costs no precision: FP16 keeps 11 significant
// bits for every value within 2^27 of the largest, and one wave reduction per column
// instead of NF took 2.6% off the coordinate form's FP16-history gradient.
template <int ESTORE, int NF>
__device__ __forceinline__ void store_hist(const SweepArgs &a, size_t o, int c, int btz, bool zok, const float (&e)[NF],
                                           int f0)
{
  if (ESTORE == 1) {
    const int iz = (int)(o - (size_t)c * a.g.ldz);
    const bool in = c >= a.hskip[0] && c < a.hskip[1] && iz >= a.hskip[2] && iz < a.hskip[3];
    if (zok && in == (a.hinside != 0))
#pragma unroll
      for (int f = 0; f < NF; f++) st_far(a.h32[f0 + f] + o, e[f]);
    return;
  }
  float mx = 0.0f;
#pragma unroll
  for (int f = 0; f < NF; f++) mx = fmaxf(mx, fabsf(e[f]));
  const float m = wave_max64(mx);
  int E = 0;
  if (m > 0.0f) (void)frexpf(m, &E);  // m < 2^E
  E = max(E, -100);                   // keeps 2^(14-E) and 2^(E-14) normal floats
  const float sc = ldexpf(1.0f, 14 - E), inv = ldexpf(1.0f, E - 14);
#pragma unroll
  for (int f = 0; f < NF; f++) {
    if (zok) st_far(a.h16[f0 + f] + o, (_Float16)(e[f] * sc));
    if (threadIdx.x == 0) a.hsc[f0 + f][(size_t)c * a.g.tilesZ + btz] = inv;
  }
}

// ---------------------------------------------------------------- velocity sweep
struct VOps {
  float vx, vz, ir, k1, k2, cv, sv;
};

// OPNT: the per-point operands load non-temporally (grids whose streamed arrays overflow
// the last-level cache); on smaller grids they load cached.
template <bool OPNT>
__device__ __forceinline__ float ldop(const float *p) { return OPNT ? ldnt(p) : ldst(p); }

template <bool COORD, bool OPNT>
__device__ __forceinline__ void load_vops(VOps &p, const SweepArgs &a, size_t o, bool und, float dt)
{
  p.vx = ldop<OPNT>(a.vx + o);
  p.vz = ldop<OPNT>(a.vz + o);
  p.ir = ldop<OPNT>(a.invrho + o);
  p.k1 = und ? 1.0f : ldop<OPNT>(a.coef1 + o);
  p.k2 = und ? dt : ldop<OPNT>(a.coef2 + o);
  if (COORD) {
    p.cv = ldop<OPNT>(a.csv + o);
    p.sv = ldop<OPNT>(a.snv + o);
  } else {
    p.cv = p.sv = 0.0f;
  }
}

// Velocity sweep (A1 + A3).  Bond and tensor: BT:654-704 / TR:617-666.
// Coordinate (COORD): CR:614-671, with the reference's invdx on d(sxx)/dz.
// The loads of the staged column and of the per-point operands run PD iterations
// ahead of their use.  REV walks the segment right to left (odd segments), as the
// stress sweep does, so that neighbouring segments stage the stress columns they
// share at the same time and the second load finds them in L2.  Column xa+j
// (forward) or xe-j (reversed) lives in ring slot (j+4) or (j+5) mod kRing.  The
// ring is the caller's, so that both directions share one LDS array.
template <bool COORD, int PD, int ESTORE = 0, bool OPNT = true, bool REV = false>
__device__ __forceinline__ void vsweep_tile(const SweepArgs &a, int btz, int btx, float (*ring)[kRing][LZ])
{
  const Grid &g = a.g;
  const int segL = g.segLV;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * TZ;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const bool und = segment_undamped(g, a.undamped, btz, xa, nc);
  const int iz = z0 + lane;
  const bool zok = iz < g.izEnd;
  const int izc = min(iz, g.Nz - 1);  // an in-grid row for the look-ahead loads
  const int xe = xa + nc - 1;
  const int rs = z0 - 4;  // first staged row
  const float *const fs[3] = {a.sxx, a.szz, a.sxz};
  // prologue: columns xa-4 .. xa+4 (reversed: xe+5 down to xe-3) into slots 0..8; the
  // next PD columns and operands in flight
  {
    Col<3> c[9];
#pragma unroll
    for (int m = 0; m < 9; m++) load_col<3>(c[m], fs, REV ? xe + 5 - m : xa - 4 + m, rs, lane, g);
#pragma unroll
    for (int m = 0; m < 9; m++) store_col<3>(ring, m, c[m], lane);
  }
  Col<3> colq[PD];
  VOps opq[PD];
#pragma unroll
  for (int i = 0; i < PD; i++) {
    if (i <= nc - 1) load_col<3>(colq[i], fs, REV ? xe - 4 - i : xa + 5 + i, rs, lane, g);
    if (zok && i < nc) load_vops<COORD, OPNT>(opq[i], a, (size_t)(REV ? xe - i : xa + i) * g.ldz + iz, und, g.dt);
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = REV ? xe - k : xa + k;
      Col<3> cn;
      VOps on;
      // look-ahead loads with no branch around them (clamped column and row; the values
      // past the segment or the grid are not used)
      load_col_nb<3>(cn, fs, REV ? max(c - 4 - PD, xa - 4) : min(c + 5 + PD, xe + 5), rs, lane, g);
      load_vops<COORD, OPNT>(on, a, (size_t)(REV ? max(c - PD, xa) : min(c + PD, xe)) * g.ldz + izc, und, g.dt);
      store_col<3>(ring, (u + 9) % kRing, colq[0], lane);
      __syncthreads();
      auto slot = [](int dx) { return REV ? (u - dx + 5 + kRing) % kRing : (u + dx + 4 + kRing) % kRing; };
      float hv[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // coordinate-form history (ESTORE)
      if (zok) {
        const VOps &op = opq[0];
        float d1, d2;
        rsg_fwd([&](int dx, int dz) { return ring[0][slot(dx)][lane + 4 + dz]; }, d1, d2);
        const float sxxdx = 0.5f * (d1 + d2) * invdx;
        const float sxxdz = COORD ? 0.5f * (d1 - d2) * invdx : 0.0f;
        rsg_fwd([&](int dx, int dz) { return ring[1][slot(dx)][lane + 4 + dz]; }, d1, d2);
        const float szzdx = COORD ? 0.5f * (d1 + d2) * invdx : 0.0f;
        const float szzdz = 0.5f * (d1 - d2) * invdz;
        rsg_fwd([&](int dx, int dz) { return ring[2][slot(dx)][lane + 4 + dz]; }, d1, d2);
        const float sxzdx = 0.5f * (d1 + d2) * invdx;
        const float sxzdz = 0.5f * (d1 - d2) * invdz;
        const size_t o = (size_t)c * g.ldz + iz;
        const float kr = op.k2 * op.ir;
        if (!COORD) {
          st_out(a.vx + o, op.k1 * op.vx + kr * (sxxdx + sxzdz));
          st_out(a.vz + o, op.k1 * op.vz + kr * (sxzdx + szzdz));
        } else {
          const float cv = op.cv, sv = op.sv;
          st_out(a.vx + o, op.k1 * op.vx + kr * (cv * sxxdx - sv * sxxdz + sv * sxzdx + cv * sxzdz));
          st_out(a.vz + o, op.k1 * op.vz + kr * (cv * sxzdx - sv * sxzdz + sv * szzdx + cv * szzdz));
          if (ESTORE) {
            hv[0] = -sxxdz + sxzdx;
            hv[1] = -sxzdz + szzdx;
            hv[2] = sxxdx + sxzdz;
            hv[3] = sxzdx + szzdz;
          }
        }
      }
      if (ESTORE) store_hist<ESTORE, 4>(a, (size_t)c * g.ldz + iz, c, btz, zok, hv, 4);
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

// ---------------------------------------------------------------- stress sweep
// Per-form operands of one stress-sweep point: old state, stiffness, trig, coefs.
struct SOps {
  float s1, s2, s3;  // old sxx, szz, sxz (Bond/coord) or txx, tzz, txz (tensor)
  float c11, c13, c33, c55;
  float a, b, c;     // Bond: c15, c35, -; coord: sin, cos, -; tensor: sin2, cos2, sin2t
  float k1, k2;
};

template <int FORM, bool OPNT>
__device__ __forceinline__ void load_sops(SOps &p, const SweepArgs &a, size_t o, bool und, float dt)
{
  if (FORM == FDMOD_FORM_TENSOR) {
    p.s1 = ldop<OPNT>(a.txx + o); p.s2 = ldop<OPNT>(a.tzz + o); p.s3 = ldop<OPNT>(a.txz + o);
  } else {
    p.s1 = ldop<OPNT>(a.sxx + o); p.s2 = ldop<OPNT>(a.szz + o); p.s3 = ldop<OPNT>(a.sxz + o);
  }
  p.c11 = ldop<OPNT>(a.c11 + o); p.c13 = ldop<OPNT>(a.c13 + o); p.c33 = ldop<OPNT>(a.c33 + o); p.c55 = ldop<OPNT>(a.c55 + o);
  if (FORM == FDMOD_FORM_BOND) { p.a = ldop<OPNT>(a.c15 + o); p.b = ldop<OPNT>(a.c35 + o); p.c = 0.0f; }
  else if (FORM == FDMOD_FORM_COORD) { p.a = ldop<OPNT>(a.sn + o); p.b = ldop<OPNT>(a.cs + o); p.c = 0.0f; }
  else { p.a = ldop<OPNT>(a.sin2 + o); p.b = ldop<OPNT>(a.cos2 + o); p.c = ldop<OPNT>(a.sin2t + o); }
  p.k1 = und ? 1.0f : ldop<OPNT>(a.coef1 + o);
  p.k2 = und ? dt : ldop<OPNT>(a.coef2 + o);
}

// Stress sweep (A2 + A4/A5/A6), with the free-surface stress imaging (BT:875-891)
// done by the top strip, whose first rows are izStart .. izStart+5.
//   FORM_BOND   BT:836-873
//   FORM_TENSOR TR:799-836 and the sigma <- R(tau) sweep TR:839-852, fused pointwise
//   FORM_COORD  CR:804-845
// REV walks the segment right to left (odd segments): two neighbouring segments then
// stage the velocity columns they share at the same time, at both their starts or both
// their ends, so the second load finds them in L2 (6% fewer HBM fetches on a 1256 x 8002
// grid).  Column xa+j (forward) or xe+j (reversed) lives in ring slot (j+5) or (4-j)
// mod kRing.  The ring is the caller's, so that both directions share one LDS array.
template <int FORM, int PD, int ESTORE, bool REV, bool TNT = false, bool OPNT = true>
__device__ __forceinline__ void ssweep_tile(const SweepArgs &a, int btz, int btx, float (*ring)[kRing][LZ])
{
  const Grid &g = a.g;
  const int segL = g.segLS;
  const int lane = threadIdx.x;
  const int z0 = g.izStart + btz * TZ;
  const int xa = g.ixStart + btx * segL;
  const int nc = min(segL, g.ixEnd - xa);
  const bool und = segment_undamped(g, a.undamped, btz, xa, nc);
  const bool fs_top = a.freesurface && btz == 0;
  const int iz = z0 + lane;
  const bool zok = iz < g.izEnd;
  const int izc = min(iz, g.Nz - 1);  // an in-grid row for the look-ahead loads
  const int rs = z0 - 5;
  const float *const fv[2] = {a.vx, a.vz};
  const int xe = xa + nc - 1;
  // The fused source's sample and tap weights, and the segment's fused receivers, are
  // loaded here, once: a load inside the column loop that is used in the same iteration
  // would make the compiler wait for every load in flight, look-ahead included.
  __shared__ float tapca[(2 * kN + 1) * (2 * kN + 1)];
  const bool src_here = (a.fuse & 1) && a.fsx + kN >= xa && a.fsx - kN <= xe;
  float fsv = 0.0f;
  if (src_here) {
    fsv = a.fsrc[a.fit];
    for (int t = lane; t < (2 * kN + 1) * (2 * kN + 1); t += TZ) tapca[t] = a.ftap[t].ca;
  }
  // receivers: for segments of fewer than 64 columns lane l holds rc_start[xa + l]
  // (l <= nc) and, when the segment has at most 64 receivers, the row and id of its l-th
  // receiver (CSR order); longer segments read the CSR in the loop
  const bool rec_strip = (a.fuse & 2) && z0 <= a.rc_zhi && z0 + TZ > a.rc_zlo;
  int rcs = 0, rcz = -1, rid = 0, j00 = 0;
  bool rpre = false;
  if (rec_strip && nc < TZ) {
    rcs = lane <= nc ? a.rc_start[xa + lane] : 0;
    j00 = __shfl(rcs, 0, 64);
    const int cnt = __shfl(rcs, nc, 64) - j00;
    rpre = cnt <= TZ;
    if (rpre && lane < cnt) {
      rcz = a.rc_z[j00 + lane];
      rid = a.rc_id[j00 + lane];
    }
  }
  {
    Col<2> c[9];
#pragma unroll
    for (int m = 0; m < 9; m++) load_col<2>(c[m], fv, REV ? xe + 4 - m : xa - 5 + m, rs, lane, g);
#pragma unroll
    for (int m = 0; m < 9; m++) store_col<2>(ring, m, c[m], lane);
  }
  Col<2> colq[PD];
  SOps opq[PD];
#pragma unroll
  for (int i = 0; i < PD; i++) {
    if (i <= nc - 1) load_col<2>(colq[i], fv, REV ? xe - 5 - i : xa + 4 + i, rs, lane, g);
    if (zok && i < nc) load_sops<FORM, OPNT>(opq[i], a, (size_t)(REV ? xe - i : xa + i) * g.ldz + iz, und, g.dt);
  }
  const float invdx = g.invdx, invdz = g.invdz;
  for (int k0 = 0; k0 < nc; k0 += kRing) {
    static_for<0, kRing>([&](auto uc) {
      constexpr int u = decltype(uc)::value;
      const int k = k0 + u;
      if (k >= nc) return;
      const int c = REV ? xe - k : xa + k;
      Col<2> cn;
      SOps on;
      // look-ahead loads with no branch around them (clamped column and row; the values
      // past the segment or the grid are not used)
      load_col_nb<2>(cn, fv, REV ? max(c - 5 - PD, xa - 5) : min(c + 4 + PD, xe + 4), rs, lane, g);
      load_sops<FORM, OPNT>(on, a, (size_t)(REV ? max(c - PD, xa) : min(c + PD, xe)) * g.ldz + izc, und, g.dt);
      store_col<2>(ring, (u + 9) % kRing, colq[0], lane);
      __syncthreads();
      auto slot = [](int dx) { return REV ? (u - dx + 4 + kRing) % kRing : (u + dx + 5 + kRing) % kRing; };
      const size_t o = (size_t)c * g.ldz + iz;
      float nxx = 0.0f, nzz = 0.0f, nxz = 0.0f;
      float hv[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};  // history values (ESTORE)
      if (zok) {
        const SOps &op = opq[0];
        float d1, d2;
        rsg_bwd([&](int dx, int dz) { return ring[0][slot(dx)][lane + 5 + dz]; }, d1, d2);
        float vxdx = 0.5f * (d1 + d2) * invdx;
        float vxdz = 0.5f * (d1 - d2) * invdz;
        rsg_bwd([&](int dx, int dz) { return ring[1][slot(dx)][lane + 5 + dz]; }, d1, d2);
        float vzdx = 0.5f * (d1 + d2) * invdx;
        float vzdz = 0.5f * (d1 - d2) * invdz;
        if (ESTORE && FORM == FDMOD_FORM_BOND) {
          hv[0] = vxdx;
          hv[1] = vzdz;
          hv[2] = vzdx + vxdz;
        } else if (ESTORE && FORM == FDMOD_FORM_TENSOR) {
          // sum and difference of the normal strains: the theta imaging uses their
          // difference, which the stored precision then carries directly (for
          // P-dominated wavefields vxdx and vzdz nearly cancel there)
          hv[0] = vxdx + vzdz;
          hv[1] = vxdx - vzdz;
          hv[2] = vzdx + vxdz;
        } else if (ESTORE) {
          hv[0] = vxdx;
          hv[1] = vxdz;
          hv[2] = vzdx;
          hv[3] = vzdz;
        }
        const float k1 = op.k1, k2 = op.k2;
        if (FORM == FDMOD_FORM_BOND) {
          const float c15 = op.a, c35 = op.b;
          nxx = k1 * op.s1 + k2 * (op.c11 * vxdx + op.c13 * vzdz + c15 * (vzdx + vxdz));
          nzz = k1 * op.s2 + k2 * (op.c13 * vxdx + op.c33 * vzdz + c35 * (vzdx + vxdz));
          nxz = k1 * op.s3 + k2 * (c15 * vxdx + c35 * vzdz + op.c55 * (vzdx + vxdz));
        } else if (FORM == FDMOD_FORM_COORD) {
          const float sn = op.a, cs = op.b;
          float t = sn * vxdx + cs * vxdz;
          vxdx = cs * vxdx - sn * vxdz;
          vxdz = t;
          t = sn * vzdx + cs * vzdz;
          vzdx = cs * vzdx - sn * vzdz;
          vzdz = t;
          nxx = k1 * op.s1 + k2 * (op.c11 * vxdx + op.c13 * vzdz);
          nzz = k1 * op.s2 + k2 * (op.c13 * vxdx + op.c33 * vzdz);
          nxz = k1 * op.s3 + k2 * op.c55 * (vzdx + vxdz);
        } else {
          const float S2 = op.a, K2 = op.b, S2T = op.c;
          const float exx = K2 * vxdx + S2 * vzdz - 0.5f * S2T * (vxdz + vzdx);
          const float ezz = S2 * vxdx + K2 * vzdz + 0.5f * S2T * (vxdz + vzdx);
          const float exz = S2T * (vxdx - vzdz) + (K2 - S2) * (vxdz + vzdx);
          const float t1 = k1 * op.s1 + k2 * (op.c11 * exx + op.c13 * ezz);
          const float t2 = k1 * op.s2 + k2 * (op.c13 * exx + op.c33 * ezz);
          const float t3 = k1 * op.s3 + k2 * op.c55 * exz;
          // tau is next read by the next stress sweep, after the velocity sweep has streamed
          // its fields through the caches: on large grids (TNT) non-temporal stores leave the
          // caches to sigma, which the velocity sweep reads next (tensor stress sweep 133.1
          // -> 129.5 us, 45.9 -> 47.4 Gpts/s on the 1126 x 7872 grid).  A compile-time choice:
          // with a run-time branch the compiler merges the two stores and drops the hint.
          if constexpr (TNT) {
            st_far(a.txx + o, t1);
            st_far(a.tzz + o, t2);
            st_far(a.txz + o, t3);
          } else {
            st_out(a.txx + o, t1);
            st_out(a.tzz + o, t2);
            st_out(a.txz + o, t3);
          }
          // tau stored as (txx + tzz, txx - tzz, txz), for the same reason
          hv[3] = a.rev ? op.s1 + op.s2 : t1 + t2;
          hv[4] = a.rev ? op.s1 - op.s2 : t1 - t2;
          hv[5] = a.rev ? op.s3 : t3;
          nxx = K2 * t1 + S2 * t2 + S2T * t3;
          nzz = S2 * t1 + K2 * t2 - S2T * t3;
          nxz = -0.5f * S2T * (t1 - t2) + (K2 - S2) * t3;
        }
      } else if (FORM != FDMOD_FORM_TENSOR && fs_top && iz < g.Nz) {
        // rows past izEnd are not updated; their current values feed the images.  The
        // Bond and coordinate forms have them from the look-ahead loads (a load here
        // would make the compiler wait for every load in flight, in every strip).
        nxx = opq[0].s1;
        nzz = opq[0].s2;
        nxz = opq[0].s3;
      }
      // the tensor form's look-ahead holds tau, not sigma: a top strip that reaches past
      // izEnd (grids less than 64 rows deep) loads sigma there, under a uniform branch
      if (FORM == FDMOD_FORM_TENSOR && fs_top && z0 + TZ > g.izEnd) {
        if (!zok && iz < g.Nz) {
          nxx = a.sxx[o];
          nzz = a.szz[o];
          nxz = a.sxz[o];
        }
      }
      // free-surface images of row izStart+lane into row izStart-lane (lanes 1..kN)
      const bool img = fs_top && lane >= 1 && lane <= kN;
      float ixx = 0.0f, izz = 0.0f, ixz = 0.0f;
      if (fs_top) {
        if (lane == 0) {
          nxz = 0.0f;
          nzz = 0.0f;
        }
        const float s0 = __shfl(nxx, 0, 64);
        if (img) {
          ixx = 2.0f * s0 - nxx;
          izz = -nzz;
          ixz = -nxz;
        }
      }
      if (src_here) {  // stress source (BT:893-914) after the images, tap by tap as k_source adds it
        const int ox = c - a.fsx;
        if (ox >= -kN && ox <= kN) {
          const float *tp = tapca + (ox + kN) * (2 * kN + 1);
          const int oz = iz - a.fsz, om = z0 - lane - a.fsz;
          if (zok && oz >= -kN && oz <= kN) {
            const float da = tp[oz + kN] * fsv;
            nxx = nxx + da;
            nzz = nzz + da;
          }
          if (img && om >= -kN && om <= kN) {
            const float da = tp[om + kN] * fsv;
            ixx = ixx + da;
            izz = izz + da;
          }
        }
      }
      if (img) {
        const size_t m = (size_t)c * g.ldz + (z0 - lane);
        a.sxx[m] = ixx;
        a.szz[m] = izz;
        a.sxz[m] = ixz;
      }
      if (zok) {
        st_out(a.sxx + o, nxx);
        st_out(a.szz + o, nzz);
        st_out(a.sxz + o, nxz);
      }
      // pressure receivers of this column (isoelastic.c:103-144), in strips that hold some
      if (rec_strip) {
        const int kc = c - xa;
        const int j0 = nc < TZ ? __shfl(rcs, kc, 64) : a.rc_start[c];
        const int j1 = nc < TZ ? __shfl(rcs, kc + 1, 64) : a.rc_start[c + 1];
        if (rpre) {  // the segment's receivers are in the lanes
          for (int j = j0; j < j1; j++) {
            const int z = __shfl(rcz, j - j00, 64), id = __shfl(rid, j - j00, 64);
            if (zok && z == iz) a.rec_p[a.rec_off + id] = 0.5f * (nxx + nzz);
          }
        } else {
          for (int j = j0; j < j1; j++)
            if (zok && a.rc_z[j] == iz) a.rec_p[a.rec_off + a.rc_id[j]] = 0.5f * (nxx + nzz);
        }
      }
      if (ESTORE) {
        if (FORM == FDMOD_FORM_BOND) {
          const float e[3] = {hv[0], hv[1], hv[2]};
          store_hist<ESTORE, 3>(a, o, c, btz, zok, e, 0);
        } else if (FORM == FDMOD_FORM_TENSOR) {
          const float e[3] = {hv[0], hv[1], hv[2]}, t[3] = {hv[3], hv[4], hv[5]};
          store_hist<ESTORE, 3>(a, o, c, btz, zok, e, 0);
          store_hist<ESTORE, 3>(a, o, c, btz, zok, t, 3);
        } else {
          const float e[4] = {hv[0], hv[1], hv[2], hv[3]};
          store_hist<ESTORE, 4>(a, o, c, btz, zok, e, 0);
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

#ifndef FDMOD_PIPE_V
#define FDMOD_PIPE_V 2
#endif
#ifndef FDMOD_PIPE_S
#define FDMOD_PIPE_S 1
#endif
constexpr int kPipeV = FDMOD_PIPE_V;  // prefetch depth (iterations) of the velocity sweep
// Single-shot velocity sweep of the Bond and tensor forms: three columns ahead.  It loads
// two operand words per point fewer than the coordinate form's (no velocity-point trig),
// so it needs more iterations in flight to keep as many bytes in flight.  Measured on
// the 1256 x 8002 grid (Gpts/s of the loop, depth 2 / 3 / 4): Bond 61.0 / 62.3 / 61.9,
// tensor 47.3 / 47.9 / 47.6, coordinate 60.8 / 60.2 / 59.6.
template <bool COORD>
constexpr int pipe_v() { return COORD ? kPipeV : kPipeV + 1; }
// Stress sweep: one column ahead for every form (Bond and coordinate 89 VGPRs, tensor 91,
// five waves per SIMD; the tensor form at depth 2 needs 103, four waves).  Measured on
// the 1256 x 8002 grid with branch-free look-ahead loads: tensor stress sweep 134.7 us at
// depth 1 against 140.8 us at depth 2.
template <int FORM>
constexpr int pipe_s() { return FDMOD_PIPE_S; }

// Single grid: blocks map to tiles of one shot.
template <bool COORD, int PD, int ESTORE = 0, bool OPNT = true>
__global__ __launch_bounds__(64) void k_vsweep(SweepArgs a)
{
  __shared__ float ring[3][kRing][LZ];
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLV, btz, btx)) return;
  // right-to-left odd segments only on grids that stream past the last-level cache:
  // there neighbours stage the columns they share together (1256 x 8002, coordinate form:
  // velocity sweep 71.6-72.1 -> 70.1-70.6 us, loop 59.2-59.4 -> 59.7-60.0 Gpts/s; Bond
  // 60.6-61.0 -> 60.9-61.2, tensor unchanged); on cache-resident grids they
  // saved nothing and the coordinate FP16-history gradient on a 1126 x 2000 crop ran 1%
  // slower (0.629 -> 0.635 s)
  if (OPNT && (btx & 1)) vsweep_tile<COORD, PD, ESTORE, OPNT, true>(a, btz, btx, ring);
  else vsweep_tile<COORD, PD, ESTORE, OPNT, false>(a, btz, btx, ring);
}

template <int FORM, int PD, int ESTORE = 0, bool TNT = false, bool OPNT = true>
__global__ __launch_bounds__(64) void k_ssweep(SweepArgs a)
{
  __shared__ float ring[2][kRing][LZ];
  int btz, btx;
  if (!tile_of_block(a.g, a.g.segLS, btz, btx)) return;
  if (btx & 1) ssweep_tile<FORM, PD, ESTORE, true, TNT, OPNT>(a, btz, btx, ring);
  else ssweep_tile<FORM, PD, ESTORE, false, TNT, OPNT>(a, btz, btx, ring);
}

// Time index of loop iteration k = *step + off (BT:651): it = k or nt-1-k.
__device__ __forceinline__ int step_it(const int *step, int off, int nt, int mode)
{
  const int k = *step + off;
  return mode == 1 ? k : nt - 1 - k;
}

// Batched grids: one launch covers the tiles of every shot.  start[s] is the first
// global tile of shot s (prefix sum over shots, start[nshots] = total); the global
// tile index is XCD-aware as in the single grid, then located by binary search.
__device__ __forceinline__ bool batch_tile(const int *__restrict__ start, int nshots, int &shot, int &lt)
{
  const int total = start[nshots];
  const int chunk = (total + kXCD - 1) / kXCD;
  const int b = blockIdx.x;
  const int t = (b % kXCD) * chunk + b / kXCD;
  if (t >= total) return false;
  int lo = 0, hi = nshots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  shot = lo;
  lt = t - start[lo];
  return true;
}

template <bool COORD, int PD>
__global__ __launch_bounds__(64) void k_vsweep_batch(const SweepArgs *__restrict__ shots, const int *__restrict__ start,
                                                     int nshots)
{
  __shared__ float ring[3][kRing][LZ];
  int s, lt;
  if (!batch_tile(start, nshots, s, lt)) return;
  const SweepArgs a = shots[s];
  const int btx = lt / a.g.tilesZ;
  if (btx & 1) vsweep_tile<COORD, PD, 0, true, true>(a, lt - btx * a.g.tilesZ, btx, ring);
  else vsweep_tile<COORD, PD, 0, true, false>(a, lt - btx * a.g.tilesZ, btx, ring);
}

// Shots with a fused stress source or pressure receivers (as in the single-shot loop)
// take the iteration from the device step counter: the source sample and, when
// it % mt == 0, the receivers' output row.
template <int FORM, int PD, bool TNT = false>
__global__ __launch_bounds__(64) void k_ssweep_batch(const SweepArgs *__restrict__ shots, const int *__restrict__ start,
                                                     int nshots, const int *step, int off, int nt, int mode)
{
  __shared__ float ring[2][kRing][LZ];
  int s, lt;
  if (!batch_tile(start, nshots, s, lt)) return;
  SweepArgs a = shots[s];
  if (a.fuse) {
    const int it = step_it(step, off, nt, mode);
    a.fit = it;
    if (a.fuse & 2) {  // rc_mt >= 1 whenever bit 2 is set
      if (it % a.rc_mt == 0) a.rec_off = (size_t)(it / a.rc_mt) * a.rc_nr;
      else a.fuse &= ~2;
    }
  }
  const int btx = lt / a.g.tilesZ;
  if (btx & 1) ssweep_tile<FORM, PD, 0, true, TNT>(a, lt - btx * a.g.tilesZ, btx, ring);
  else ssweep_tile<FORM, PD, 0, false, TNT>(a, lt - btx * a.g.tilesZ, btx, ring);
}

static dim3 sweep_grid(const Grid &g, int segL)
{
  const int n = g.tilesZ * ((g.ixEnd - g.ixStart + segL - 1) / segL);
  return dim3((n + kXCD - 1) / kXCD * kXCD);
}

int sweep_tiles(const Grid &g, int segL) { return g.tilesZ * ((g.ixEnd - g.ixStart + segL - 1) / segL); }

template <bool COORD, int ESTORE>
static void launch_v(hipStream_t s, dim3 grid, const SweepArgs &a)
{
  // cache-resident grids keep two columns in flight for every form: on the 1126 x 2000 crop the Bond
  // FP16-history gradient took 0.515 / 0.515 s against 0.516 / 0.521 s with three
  if (a.op_nt) hipLaunchKernelGGL((k_vsweep<COORD, pipe_v<COORD>(), ESTORE, true>), grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL((k_vsweep<COORD, kPipeV, ESTORE, false>), grid, dim3(64), 0, s, a);
}

// The tensor form's tau_nt and op_nt hold on the same grids (both count its 20 streamed
// arrays), so its sweeps come in two variants, as do the other forms'.
template <int FORM, int ESTORE>
static void launch_s(hipStream_t s, dim3 grid, const SweepArgs &a)
{
  constexpr bool T = FORM == FDMOD_FORM_TENSOR;
  if (a.op_nt) hipLaunchKernelGGL((k_ssweep<FORM, pipe_s<FORM>(), ESTORE, T, true>), grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL((k_ssweep<FORM, pipe_s<FORM>(), ESTORE, false, false>), grid, dim3(64), 0, s, a);
}

void launch_vsweep(hipStream_t s, int form, const SweepArgs &a)
{
  dim3 grid = sweep_grid(a.g, a.g.segLV);
  if (form == FDMOD_FORM_COORD) {
    if (a.estore == 1) launch_v<true, 1>(s, grid, a);
    else if (a.estore == 2) launch_v<true, 2>(s, grid, a);
    else launch_v<true, 0>(s, grid, a);
  } else {
    launch_v<false, 0>(s, grid, a);
  }
}

void launch_ssweep(hipStream_t s, int form, const SweepArgs &a)
{
  dim3 grid = sweep_grid(a.g, a.g.segLS);
#define FDMOD_SSWEEP(F)                                      \
  do {                                                     \
    if (a.estore == 1) launch_s<F, 1>(s, grid, a);          \
    else if (a.estore == 2) launch_s<F, 2>(s, grid, a);     \
    else launch_s<F, 0>(s, grid, a);                        \
  } while (0)
  if (form == FDMOD_FORM_COORD) FDMOD_SSWEEP(FDMOD_FORM_COORD);
  else if (form == FDMOD_FORM_TENSOR) FDMOD_SSWEEP(FDMOD_FORM_TENSOR);
  else FDMOD_SSWEEP(FDMOD_FORM_BOND);
#undef FDMOD_SSWEEP
}

void launch_vsweep_batch(hipStream_t s, int form, const SweepArgs *shots, const int *start, int nshots, int total)
{
  dim3 grid((total + kXCD - 1) / kXCD * kXCD);
  if (form == FDMOD_FORM_COORD) hipLaunchKernelGGL((k_vsweep_batch<true, kPipeV>), grid, dim3(64), 0, s, shots, start, nshots);
  else hipLaunchKernelGGL((k_vsweep_batch<false, kPipeV>), grid, dim3(64), 0, s, shots, start, nshots);
}

void launch_ssweep_batch(hipStream_t s, int form, const SweepArgs *shots, const int *start, int nshots, int total,
                         const int *step, int off, int nt, int mode, bool tau_nt)
{
  dim3 grid((total + kXCD - 1) / kXCD * kXCD);
  if (form == FDMOD_FORM_COORD)
    hipLaunchKernelGGL((k_ssweep_batch<FDMOD_FORM_COORD, pipe_s<FDMOD_FORM_COORD>()>), grid, dim3(64), 0, s, shots, start,
                       nshots, step, off, nt, mode);
  else if (form == FDMOD_FORM_TENSOR && tau_nt)
    hipLaunchKernelGGL((k_ssweep_batch<FDMOD_FORM_TENSOR, pipe_s<FDMOD_FORM_TENSOR>(), true>), grid, dim3(64), 0, s,
                       shots, start, nshots, step, off, nt, mode);
  else if (form == FDMOD_FORM_TENSOR)
    hipLaunchKernelGGL((k_ssweep_batch<FDMOD_FORM_TENSOR, pipe_s<FDMOD_FORM_TENSOR>()>), grid, dim3(64), 0, s, shots,
                       start, nshots, step, off, nt, mode);
  else
    hipLaunchKernelGGL((k_ssweep_batch<FDMOD_FORM_BOND, pipe_s<FDMOD_FORM_BOND>()>), grid, dim3(64), 0, s, shots, start,
                       nshots, step, off, nt, mode);
}

// Source injection (BT:706-834 on velocity, BT:893-914 on stress).  Each pass is a
// table of taps; one thread per tap adds ca*src[it] to fa (and fc) and cb*src[it]
// to fb (and fd) at the tap's cell when it lies inside the grid (the reference's
// MAX/MIN clipping).  Passes run in order (stype 4: divergence, then curl), and
// within a pass the sources are walked in order inside one block when their
// footprints may overlap, so every cell sees the additions in the reference's order.
__device__ __forceinline__ void source_body(const SourceArgs &a, int it, int blk)
{
  const Grid &g = a.g;
  const int first = a.sequential ? 0 : blk;
  const int last = a.sequential ? a.ns : min(blk + 1, a.ns);
  for (int p = 0; p < a.npass; p++) {
    for (int is = first; is < last; is++) {
      const float sv = a.src[(size_t)is * a.nt + it];
      for (int t = threadIdx.x; t < a.ntap; t += blockDim.x) {
        const SrcTap tp = a.taps[p * a.ntap + t];
        const int ix = a.isx[is] + tp.ox, iz = a.isz[is] + tp.oz;
        if (ix < 0 || ix >= g.Nx || iz < 0 || iz >= g.Nz) continue;
        const size_t o = (size_t)ix * g.ldz + iz;
        const float da = tp.ca * sv;
        a.fa[o] = a.fa[o] + da;
        if (a.fc) a.fc[o] = a.fc[o] + da;
        if (a.fb) {
          const float db = tp.cb * sv;
          a.fb[o] = a.fb[o] + db;
          if (a.fd) a.fd[o] = a.fd[o] + db;
        }
      }
      __syncthreads();
    }
  }
}

__global__ void k_source(SourceArgs a) { source_body(a, a.it, blockIdx.x); }

void launch_source(hipStream_t s, const SourceArgs &a)
{
  dim3 grid(a.sequential ? 1 : a.ns);
  hipLaunchKernelGGL(k_source, grid, dim3(128), 0, s, a);
}

// Batched sources: blockIdx.y = shot; shots without a source in this phase have ns = 0.
__global__ void k_source_batch(const SourceArgs *__restrict__ shots, const int *step, int off, int nt, int mode)
{
  const SourceArgs a = shots[blockIdx.y];
  if (a.ns == 0 || (!a.sequential && (int)blockIdx.x >= a.ns) || (a.sequential && blockIdx.x > 0)) return;
  source_body(a, step_it(step, off, nt, mode), blockIdx.x);
}

void launch_source_batch(hipStream_t s, const SourceArgs *shots, int nshots, int maxblocks, const int *step, int off,
                         int nt, int mode)
{
  hipLaunchKernelGGL(k_source_batch, dim3(maxblocks, nshots), dim3(128), 0, s, shots, step, off, nt, mode);
}

// elastic2dReceiver (L/fdmod/isoelastic.c:103-144): nearest-point gathers into a
// device buffer laid out [sample][trace].
__device__ __forceinline__ void receiver_body(const RecvArgs &a, int isample, int r)
{
  if (r >= a.nr) return;
  size_t o = (size_t)a.gx[r] * a.ldz + a.gz[r];
  size_t d = (size_t)isample * a.nr + r;
  if (a.want_v) {
    a.out_vx[d] = a.vx[o];
    a.out_vz[d] = a.vz[o];
  }
  if (a.want_p) a.out_p[d] = 0.5f * (a.sxx[o] + a.szz[o]);
}

__global__ void k_receivers(RecvArgs a) { receiver_body(a, a.isample, blockIdx.x * blockDim.x + threadIdx.x); }

void launch_receivers(hipStream_t s, const RecvArgs &a)
{
  hipLaunchKernelGGL(k_receivers, dim3((a.nr + 255) / 256), dim3(256), 0, s, a);
}

// Stress source and receivers in one launch (one stype-1 source that the stress sweep does
// not add itself).  Block 0 adds the source tap by tap as k_source does, then, after the
// barrier that ends the injection, records the receivers inside the source's footprint
// (listed on the host).  Blocks 1.. record every other receiver: their cells are outside
// the footprint, so the injection leaves them unchanged and the two run concurrently.
__global__ void k_source_receivers(SourceArgs src, RecvArgs r, const int *__restrict__ fprc, int nfp)
{
  if (blockIdx.x == 0) {
    source_body(src, src.it, 0);
    for (int j = threadIdx.x; j < nfp; j += blockDim.x) receiver_body(r, r.isample, fprc[j]);
    return;
  }
  const int i = (blockIdx.x - 1) * blockDim.x + threadIdx.x;
  if (i >= r.nr) return;
  const int fx = src.isx[0], fz = src.isz[0];
  if (abs(r.gx[i] - fx) <= kN && abs(r.gz[i] - fz) <= kN) return;  // block 0's
  receiver_body(r, r.isample, i);
}

void launch_source_receivers(hipStream_t s, const SourceArgs &src, const RecvArgs &r, const int *fprc, int nfp)
{
  hipLaunchKernelGGL(k_source_receivers, dim3(1 + (r.nr + 255) / 256), dim3(256), 0, s, src, r, fprc, nfp);
}

// Batched receivers: blockIdx.y = shot; records when it % mt == 0 (isoelastic.c:107).
__global__ void k_receivers_batch(const RecvArgs *__restrict__ shots, const int *step, int off, int nt, int mode)
{
  const RecvArgs a = shots[blockIdx.y];
  const int it = step_it(step, off, nt, mode);
  if (a.nr == 0 || it % a.mt != 0) return;
  receiver_body(a, it / a.mt, blockIdx.x * blockDim.x + threadIdx.x);
}

void launch_receivers_batch(hipStream_t s, const RecvArgs *shots, int nshots, int maxnr, const int *step, int off,
                            int nt, int mode)
{
  hipLaunchKernelGGL(k_receivers_batch, dim3((maxnr + 255) / 256, nshots), dim3(256), 0, s, shots, step, off, nt,
                     mode);
}

// Advance the device-side loop counter by n iterations (last node of a captured chunk).
__global__ void k_advance(int *step, int n)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += n;
}

void launch_advance(hipStream_t s, int *step, int n) { hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, step, n); }

// Pressure snapshot field p = 0.5*(sxx + szz) (isoelastic.c:259-266).
__global__ void k_pressure(const float *sxx, const float *szz, float *p, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.5f * (sxx[i] + szz[i]);
}

void launch_pressure(hipStream_t s, const float *sxx, const float *szz, float *p, int Nx, int ldz)
{
  size_t n = (size_t)Nx * ldz;
  hipLaunchKernelGGL(k_pressure, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sxx, szz, p, n);
}

// Energy diagnostic (BT:926-952).  Per-block float partial sums of the kinetic and
// potential densities over the unpadded model, then one fixed-order final sum.
__global__ void k_energy_partial(Grid g, int nx, int nz, DevArrays d, float *partial)
{
  __shared__ float sk[256], sp[256];
  float ek = 0.0f, ep = 0.0f;
  const float *vx = d.a[FDMOD_A_VX], *vz = d.a[FDMOD_A_VZ], *ir = d.a[FDMOD_A_INVRHO];
  const float *c11 = d.a[FDMOD_A_C11], *c55 = d.a[FDMOD_A_C55];
  const float *sxx = d.a[FDMOD_A_SXX], *szz = d.a[FDMOD_A_SZZ], *sxz = d.a[FDMOD_A_SXZ];
  size_t n = (size_t)nx * nz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    int ix = (int)(i / nz), iz = (int)(i - (size_t)ix * nz);
    size_t o = (size_t)(ix + g.e) * g.ldz + (iz + g.e);
    float a11 = c11[o], a55 = c55[o], s1 = sxx[o], s2 = szz[o], s3 = sxz[o];
    ek += vx[o] * vx[o] / ir[o] + vz[o] * vz[o] / ir[o];
    ep += 0.25f * (((a11 + 2.0f * a55) * s1 - a11 * s2) * s1 + (a11 * s1 - (a11 + 2.0f * a55) * s2) * s2) /
              (a55 * (a11 + a55)) +
          s3 * s3 / a55;
  }
  sk[threadIdx.x] = ek;
  sp[threadIdx.x] = ep;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      sk[threadIdx.x] += sk[threadIdx.x + st];
      sp[threadIdx.x] += sp[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sk[0];
    partial[2 * blockIdx.x + 1] = sp[0];
  }
}

__global__ void k_energy_final(const float *partial, int nb, float dx, float dz, float *E, float *Ek, float *Ep,
                               int it)
{
  if (threadIdx.x != 0) return;
  double ek = 0.0, ep = 0.0;
  for (int b = 0; b < nb; b++) {
    ek += partial[2 * b];
    ep += partial[2 * b + 1];
  }
  float fk = 0.5f * (float)ek * dx * dz;
  float fp = 0.5f * (float)ep * dx * dz;
  Ek[it] = fk;
  Ep[it] = fp;
  E[it] = fk + fp;
}

void launch_energy_partial(hipStream_t s, const Grid &g, int nx, int nz, const DevArrays &d, float *partial,
                           int nblocks)
{
  hipLaunchKernelGGL(k_energy_partial, dim3(nblocks), dim3(256), 0, s, g, nx, nz, d, partial);
}

void launch_energy_final(hipStream_t s, const float *partial, int nblocks, float dx, float dz, float *E, float *Ek,
                         float *Ep, int it)
{
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, s, partial, nblocks, dx, dz, E, Ek, Ep, it);
}

}  // namespace fdm
