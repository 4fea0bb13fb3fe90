// DFTs of length N = 16 M (M = 1..4: edges 16, 32, 48, 64) on the 16 lanes of a
// DPP row, for the y and z passes of the single-RVE spectral preconditioner.
//
// Lane l of a row holds x[l + 16 j], j < M, in registers.  With n = 16 j + l and
// k = k1 + M k2 (one Cooley-Tukey level, N1 = M in registers, N2 = 16 across
// lanes):
//     X[k1 + M k2] = sum_l W16^(l k2) [ W_N^(l k1) sum_j x[16 j + l] W_M^(j k1) ].
// Forward: the M-point DFT over j in registers, the twiddle W_N^(l k1), then per
// k1 a 16-point radix-2 decimation-in-frequency transform across the row whose
// exchanges are DPP moves (no LDS, no barrier); lane l ends with
// X[k1 + M bitrev4(l)] in register k1.  Inverse: decimation in time from that
// bit-reversed order with conjugate twiddles, the twiddle conj W_N^(l k1) and the
// inverse M-point DFT, back to x[l + 16 j] (unnormalised both ways).  Every lane
// of the row must be active.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "dft.hpp"

namespace cia {

// Value of lane l ^ H of the same 16-lane row.
template <int H>
__device__ __forceinline__ double row_xor(double v) {
  if constexpr (H == 1) return dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
  if constexpr (H == 2) return dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
  // row_half_mirror (i -> 7 - i in each half row), then quad_perm [3,2,1,0]:
  // i -> 8a + 4(1 - b) + c = i ^ 4 for i = 8a + 4b + c
  if constexpr (H == 4) return dpp_f64<0x1B>(dpp_f64<0x141>(v));
  static_assert(H == 1 || H == 2 || H == 4 || H == 8, "row exchange distance");
  return dpp_f64<0x128>(v);  // row_ror:8 = i ^ 8 in a row of 16
}

__device__ __forceinline__ int bitrev4(int l) {
  return ((l & 1) << 3) | ((l & 2) << 1) | ((l & 4) >> 1) | ((l & 8) >> 3);
}

// Per-lane twiddles of one axis of edge N = 16 M: the stage factors of the
// 16-point transform (W16^((l mod h) 8 / h) on the upper lane of each pair, 1 on
// the lower) and W_N^(l k1), k1 = 1 .. M - 1, from the axis table (W_N at offset
// n1 + n2; W16^m = W_N^(M m)).
template <int M>
struct LaneTw {
  double2 s8, s4, s2;
  double2 n[M > 1 ? M - 1 : 1];
};

template <int M>
__device__ __forceinline__ void lane_twiddles(const DftAxis& ax, int l, LaneTw<M>& t) {
  const double2* wn = ax.w + ax.n1 + ax.n2;
  const double2 one = make_double2(1.0, 0.0);
  t.s8 = (l & 8) ? wn[M * (l & 7)] : one;
  t.s4 = (l & 4) ? wn[M * 2 * (l & 3)] : one;
  t.s2 = (l & 2) ? wn[M * 4 * (l & 1)] : one;
#pragma unroll
  for (int k = 1; k < M; ++k) t.n[k - 1] = wn[l * k];
}

// One radix-2 stage across the row.  DIF (forward): lower lane a + p, upper lane
// (p - a) w.  DIT (inverse): the upper lane first takes a conj(w), then lower
// a + p, upper p - a.  w is 1 on lower lanes.
template <int H>
__device__ __forceinline__ double2 dif_stage(double2 a, double2 w, bool up) {
  const double2 p = make_double2(row_xor<H>(a.x), row_xor<H>(a.y));
  const double sg = up ? -1.0 : 1.0;
  const double2 d = make_double2(fma(sg, a.x, p.x), fma(sg, a.y, p.y));
  return H > 1 ? cmul(d, w, false) : d;
}

template <int H>
__device__ __forceinline__ double2 dit_stage_inv(double2 a, double2 w, bool up) {
  if (H > 1) a = cmul(a, w, true);
  const double2 p = make_double2(row_xor<H>(a.x), row_xor<H>(a.y));
  const double sg = up ? -1.0 : 1.0;
  return make_double2(fma(sg, a.x, p.x), fma(sg, a.y, p.y));
}

// M-point DFT in registers (forward sign, or conjugate when INV).
template <int M, bool INV>
__device__ __forceinline__ void dft_regs(double2 (&v)[M]) {
  if constexpr (M == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = make_double2(a.x + b.x, a.y + b.y);
    v[1] = make_double2(a.x - b.x, a.y - b.y);
  } else if constexpr (M == 3) {
    // W3 = c + i s, c = -1/2, s = -sqrt(3)/2 (conjugated: s > 0)
    const double s = INV ? 0.86602540378443864676 : -0.86602540378443864676;
    const double2 t = make_double2(v[1].x + v[2].x, v[1].y + v[2].y);
    const double2 u = make_double2(v[1].x - v[2].x, v[1].y - v[2].y);
    const double2 m = make_double2(fma(-0.5, t.x, v[0].x), fma(-0.5, t.y, v[0].y));
    v[0] = make_double2(v[0].x + t.x, v[0].y + t.y);
    v[1] = make_double2(fma(-s, u.y, m.x), fma(s, u.x, m.y));  // m + i s u
    v[2] = make_double2(fma(s, u.y, m.x), fma(-s, u.x, m.y));  // m - i s u
  } else if constexpr (M == 4) {
    const double2 a0 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
    const double2 a1 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
    const double2 b0 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
    const double2 b1 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
    // i b1 = (-b1.y, b1.x); forward X1 = a1 - i b1, X3 = a1 + i b1
    const double2 p = make_double2(a1.x + b1.y, a1.y - b1.x);  // a1 - i b1
    const double2 q = make_double2(a1.x - b1.y, a1.y + b1.x);  // a1 + i b1
    v[0] = make_double2(a0.x + b0.x, a0.y + b0.y);
    v[2] = make_double2(a0.x - b0.x, a0.y - b0.y);
    v[1] = INV ? q : p;
    v[3] = INV ? p : q;
  }
}

// Forward transform of NV independent lines (one per array row) held as above.
template <int M, int NV>
__device__ __forceinline__ void lane_dft_fwd(double2 (&v)[NV][M], const LaneTw<M>& t, int l) {
  const bool u8 = l & 8, u4 = l & 4, u2 = l & 2, u1 = l & 1;
  const double2 one = make_double2(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < NV; ++r) {
    dft_regs<M, false>(v[r]);
#pragma unroll
    for (int k = 1; k < M; ++k) v[r][k] = cmul(v[r][k], t.n[k - 1], false);
  }
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dif_stage<8>(v[r][k], t.s8, u8);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dif_stage<4>(v[r][k], t.s4, u4);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dif_stage<2>(v[r][k], t.s2, u2);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dif_stage<1>(v[r][k], one, u1);
}

template <int M, int NV>
__device__ __forceinline__ void lane_dft_inv(double2 (&v)[NV][M], const LaneTw<M>& t, int l) {
  const bool u8 = l & 8, u4 = l & 4, u2 = l & 2, u1 = l & 1;
  const double2 one = make_double2(1.0, 0.0);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dit_stage_inv<1>(v[r][k], one, u1);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dit_stage_inv<2>(v[r][k], t.s2, u2);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dit_stage_inv<4>(v[r][k], t.s4, u4);
#pragma unroll
  for (int r = 0; r < NV; ++r)
#pragma unroll
    for (int k = 0; k < M; ++k) v[r][k] = dit_stage_inv<8>(v[r][k], t.s8, u8);
#pragma unroll
  for (int r = 0; r < NV; ++r) {
#pragma unroll
    for (int k = 1; k < M; ++k) v[r][k] = cmul(v[r][k], t.n[k - 1], true);
    dft_regs<M, true>(v[r]);
  }
}

}  // namespace cia
