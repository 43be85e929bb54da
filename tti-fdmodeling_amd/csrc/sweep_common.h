// Device helpers shared by the forward and adjoint sweep kernels: the staged
// column ring, the RSG stencil taps and the XCD-aware tile mapping.
#pragma once
#include <type_traits>
#include <utility>

#include "fdmod_internal.h"

namespace fdm {

constexpr int LZ = TZ + 2 * kN - 1;  // staged rows per column: 64 + 9 halo
constexpr int kHalo = LZ - TZ;       // rows loaded by the first kHalo lanes a second time

// Cache policy (measured on MI355X in the time loop, 1256 x 8002 grid, strips on
// 128-byte lines): default caching for the staged stencil columns and all outputs;
// both sweeps' per-point operands load non-temporally.  Alternatives measured:
// caching the stress sweep's operands, +9 us per stress launch (95 -> 104 us); caching
// the velocity sweep's operands, coordinate form +6 us per velocity launch; non-temporal
// stores of either sweep, 1-4% slower steps; every stream non-temporal, slowest.
__device__ __forceinline__ float ldnt(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ldst(const float *p) { return *p; }
__device__ __forceinline__ void st_out(float *p, float v) { *p = v; }
// Outputs whose next reader comes after other sweeps have streamed through the caches
// (the tensor form's tau, the gradient's stored history): non-temporal stores.
template <typename T>
__device__ __forceinline__ void st_far(T *p, T v) { __builtin_nontemporal_store(v, p); }

// Maximum of v over the 64 lanes of the wave, the same in every lane.  DPP moves:
// within each row of 16 lanes a butterfly (quad_perm [1,0,3,2] and [2,3,0,1], then
// row_ror 4 and 8), then row_bcast 15 into rows 1 and 3 and row_bcast 31 into rows 2
// and 3, so lane 63 holds the wave's maximum; lanes a move leaves out keep their own
// value.  fmaxf is exact, so the order of the reduction does not matter.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_max(float v)
{
  const int w = __float_as_int(v);
  return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(w, w, CTRL, ROW_MASK, 0xF, false)));
}

__device__ __forceinline__ float wave_max64(float v)
{
  v = dpp_max<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_max<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_max<0x124, 0xF>(v);  // row_ror 4
  v = dpp_max<0x128, 0xF>(v);  // row_ror 8
  v = dpp_max<0x142, 0xA>(v);  // row_bcast 15 -> rows 1, 3
  v = dpp_max<0x143, 0xC>(v);  // row_bcast 31 -> rows 2, 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// AND of the fine "undamped" flags covering columns [xa, xa+nc) of strip tz.
__device__ __forceinline__ bool segment_undamped(const Grid &g, const uint8_t *flags, int tz, int xa, int nc)
{
  bool und = true;
  const int f0 = (xa - g.ixStart) / kFlagW, f1 = (xa + nc - 1 - g.ixStart) / kFlagW;
  for (int f = f0; f <= f1; f++) und = und && flags[f * g.tilesZ + tz] != 0;
  return und;
}

template <int U, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
  if constexpr (U < N) {
    f(std::integral_constant<int, U>{});
    static_for<U + 1, N>(f);
  }
}

// Map a 1-D block id to (strip, segment) so that each XCD works on a contiguous
// run of tiles (strips fastest, then segments).  Blocks are dealt to the 8 XCDs
// round-robin, so block b runs on XCD b % 8; giving XCD k the k-th contiguous
// chunk keeps neighbouring tiles -- and the halo they share -- in one L2.
constexpr int kXCD = 8;
__device__ __forceinline__ int tiles_x(const Grid &g, int segL) { return (g.ixEnd - g.ixStart + segL - 1) / segL; }

__device__ __forceinline__ bool tile_of_block(const Grid &g, int segL, int &tz, int &tx)
{
  const int n = g.tilesZ * tiles_x(g, segL);
  const int chunk = (n + kXCD - 1) / kXCD;
  const int b = blockIdx.x;
  const int t = (b % kXCD) * chunk + b / kXCD;
  if (t >= n) return false;
  tx = t / g.tilesZ;
  tz = t - tx * g.tilesZ;
  return true;
}

// One staged column of NF fields: lane l holds row rs+l and, for l < kHalo, row rs+64+l.
template <int NF>
struct Col {
  float lo[NF], hi[NF];
};

template <int NF>
__device__ __forceinline__ void load_col(Col<NF> &c, const float *const (&f)[NF], int gx, int rs, int lane,
                                         const Grid &g)
{
  const size_t base = (size_t)gx * g.ldz;
  const int r0 = rs + lane, r1 = rs + TZ + lane;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const float *col = f[k] + base;
    c.lo[k] = r0 < g.Nz ? ldst(col + r0) : 0.0f;
    c.hi[k] = (lane < kHalo && r1 < g.Nz) ? ldst(col + r1) : 0.0f;
  }
}

// The same column load with no branch around the loads: rows past the grid read row
// Nz-1 and yield zero, and lanes >= kHalo re-read their own row for the unused hi value.
// The sweeps issue their look-ahead loads this way (from clamped columns too): a load
// under a branch makes the compiler's wait before the next use of an older load count
// the path that skipped it, which waits for every load in flight -- the look-ahead is
// then lost.
template <int NF>
__device__ __forceinline__ void load_col_nb(Col<NF> &c, const float *const (&f)[NF], int gx, int rs, int lane,
                                            const Grid &g)
{
  const size_t base = (size_t)gx * g.ldz;
  const int r0 = rs + lane, r1 = rs + TZ + lane;
  const bool ok0 = r0 < g.Nz, ok1 = lane < kHalo && r1 < g.Nz;
  const int a0 = ok0 ? r0 : g.Nz - 1, a1 = ok1 ? r1 : a0;
#pragma unroll
  for (int k = 0; k < NF; k++) {
    const float *col = f[k] + base;
    const float v0 = ldst(col + a0), v1 = ldst(col + a1);
    c.lo[k] = ok0 ? v0 : 0.0f;
    c.hi[k] = ok1 ? v1 : 0.0f;
  }
}

template <int NF>
__device__ __forceinline__ void store_col(float (*ring)[kRing][LZ], int slot, const Col<NF> &c, int lane)
{
#pragma unroll
  for (int k = 0; k < NF; k++) {
    ring[k][slot][lane] = c.lo[k];
    if (lane < kHalo) ring[k][slot][TZ + lane] = c.hi[k];
  }
}

// A1 (BT:661-699): stress -> velocity pair at the output point, reading column
// offset dx and row offset dz through T(dx, dz).  d1 runs along (+x,+z), d2 along (+x,-z).
template <typename T>
__device__ __forceinline__ void rsg_fwd(T S, float &d1, float &d2)
{
  d1 = kC1 * (S(1, 1) - S(0, 0)) + kC2 * (S(2, 2) - S(-1, -1)) + kC3 * (S(3, 3) - S(-2, -2)) +
       kC4 * (S(4, 4) - S(-3, -3)) + kC5 * (S(5, 5) - S(-4, -4));
  d2 = kC1 * (S(1, 0) - S(0, 1)) + kC2 * (S(2, -1) - S(-1, 2)) + kC3 * (S(3, -2) - S(-2, 3)) +
       kC4 * (S(4, -3) - S(-3, 4)) + kC5 * (S(5, -4) - S(-4, 5));
}

// A2 (BT:843-861): velocity -> stress pair.
template <typename T>
__device__ __forceinline__ void rsg_bwd(T V, float &d1, float &d2)
{
  d1 = kC1 * (V(0, 0) - V(-1, -1)) + kC2 * (V(1, 1) - V(-2, -2)) + kC3 * (V(2, 2) - V(-3, -3)) +
       kC4 * (V(3, 3) - V(-4, -4)) + kC5 * (V(4, 4) - V(-5, -5));
  d2 = kC1 * (V(0, -1) - V(-1, 0)) + kC2 * (V(1, -2) - V(-2, 1)) + kC3 * (V(2, -3) - V(-3, 2)) +
       kC4 * (V(3, -4) - V(-4, 3)) + kC5 * (V(4, -5) - V(-5, 4));
}

}  // namespace fdm
