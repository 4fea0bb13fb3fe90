// Small-size 3-D DFTs of the spectral preconditioner (periodic RVE grids of edge
// n <= 64), in LDS, unnormalised, forward kernel exp(-2 pi i j k / n).
//
// Each 1-D transform uses one level of the Cooley-Tukey factorisation n = n1 n2
// with direct sub-DFTs: with j = n2 j1 + j2 and k = k1 + n1 k2,
//     X[k1 + n1 k2] = sum_j2 W_n2^(j2 k2) [ W_n^(j2 k1) sum_j1 x[n2 j1 + j2] W_n1^(j1 k1) ],
// i.e. (n1 + n2) complex multiply-adds per output instead of n.  A prime n uses
// n1 = 1 (a direct DFT).  k_dft_strided transforms y or z lines read as tiles
// of 16 consecutive x positions, so every load row is 256 contiguous bytes; it
// returns at once when the solver's converged flag is set.  The x-lines are
// transformed inside the fused passes of spectral.hpp and the batched kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace cia {

struct DftAxis {
  int n, n1, n2, pad;
  const double2* w;   // W_n1[n1] | W_n2[n2] | W_n[n], forward sign
  const float2* wf;   // the same in FP32 (FP32 spectra), or null
};

inline void dft_factor(int n, int& n1, int& n2) {
  n1 = 1;
  for (int d = 1; d * d <= n; ++d)
    if (n % d == 0) n1 = d;
  n2 = n / n1;
}

// Host twiddle table of one axis (appended to `out`); returns its offset.
inline size_t dft_twiddles(int n, std::vector<double2>& out) {
  int n1, n2;
  dft_factor(n, n1, n2);
  const size_t off = out.size();
  const double two_pi = 6.283185307179586476925286766559;
  for (int len : {n1, n2, n})
    for (int m = 0; m < len; ++m) {
      const double a = two_pi * (double)m / (double)len;
      out.push_back(make_double2(std::cos(a), -std::sin(a)));
    }
  return off;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b, bool conj_b) {
  const double bi = conj_b ? -b.y : b.y;
  return make_double2(a.x * b.x - a.y * bi, a.x * bi + a.y * b.x);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b, bool conj_b) {
  const float bi = conj_b ? -b.y : b.y;
  return make_float2(a.x * b.x - a.y * bi, a.x * bi + a.y * b.x);
}

// q = floor(t / n) for t < 2^16, n <= 64 without an integer divide: the float
// quotient's fractional part stays >= 1/(2n) away from an integer, far above its
// rounding error.
__device__ __forceinline__ int fdiv(int t, float inv_n) { return (int)(((float)t + 0.5f) * inv_n); }

// The factorisation rule of dft_factor at compile time.
constexpr int ct_n1(int n) {
  int n1 = 1;
  for (int d = 1; d * d <= n; ++d)
    if (n % d == 0) n1 = d;
  return n1;
}

// a * w where w = W_R^idx (W_R = exp(-2 pi i / R), conjugated when INV); idx is a
// constant once the caller's loops are unrolled, so the trivial factors 1, -1
// and -+i fold into moves and sign changes.
template <class C2>
__device__ __forceinline__ C2 tw_mul(C2 a, C2 w, int idx, int R, bool inv) {
  if (idx == 0) return a;
  if (2 * idx == R) return C2{-a.x, -a.y};
  if (4 * idx == R) return inv ? C2{-a.y, a.x} : C2{a.y, -a.x};      // -i (forward)
  if (4 * idx == 3 * R) return inv ? C2{a.y, -a.x} : C2{-a.y, a.x};  // +i (forward)
  return cmul(a, w, inv);
}

// dft_stage for an edge N fixed at compile time: one thread per small DFT of a
// line (STEP_A: the n1-point DFT of the inputs j2, j2 + n2, ...; STEP_B: the
// n2-point DFT of the inputs k1, k1 + n1, ...), its twiddles W_R held in
// registers, every index constant.  The same sums as dft_stage, in another order.
template <int N, bool STEP_A, bool INV, class C2>
__device__ __forceinline__ void dft_stage_ct(const C2* __restrict__ in, C2* __restrict__ out,
                                             int n_lines, int ls, int es,
                                             const C2* __restrict__ w) {
  constexpr int n1 = ct_n1(N), n2 = N / n1;
  constexpr int R = STEP_A ? n1 : n2;  // points of one small DFT
  constexpr int M = STEP_A ? n2 : n1;  // small DFTs per line
  const C2* wr = STEP_A ? w : w + n1;
  const C2* wn = w + n1 + n2;
  C2 tw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) tw[r] = wr[r];
  const int total = n_lines * M;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / M, q = t - line * M;
    const C2* src = in + line * ls;
    C2* dst = out + line * ls;
    C2 a[R];
#pragma unroll
    for (int j = 0; j < R; ++j) a[j] = STEP_A ? src[(n2 * j + q) * es] : src[(j * n1 + q) * es];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      C2 acc = a[0];
#pragma unroll
      for (int j = 1; j < R; ++j) {
        const int idx = (j * k) % R;
        const C2 p = tw_mul(a[j], tw[idx], idx, R, INV);
        acc.x += p.x;
        acc.y += p.y;
      }
      if (STEP_A) {
        if (k != 0) acc = cmul(acc, wn[q * k], INV);  // W_N^(j2 k1), j2 = q
        dst[(q * n1 + k) * es] = acc;
      } else {
        dst[(q + n1 * k) * es] = acc;
      }
    }
  }
}

// One stage over n_lines lines of length n: element e of line l at l*ls + e*es.
// STEP_A: out[(j2 n1 + k1)] = W_n^(j2 k1) sum_j1 in[n2 j1 + j2] W_n1^(j1 k1)
// STEP_B: out[(k1 + n1 k2)] = sum_j2 in[j2 n1 + k1] W_n2^(j2 k2)
// (j1 k1 mod n1 and j2 k2 mod n2 advance incrementally; j2 k1 < n needs no mod.)
// w: twiddles of the axis in LDS.  Edges 12, 16, 24, 32 and 48 run the
// compile-time specialisation above (dft_stage_ct); a stage of odd length R >= 3
// (a prime edge's direct DFT, or an odd factor) runs dft_stage_pairs.
template <class C2>
struct RealOf;
template <>
struct RealOf<float2> {
  using type = float;
};
template <>
struct RealOf<double2> {
  using type = double;
};

// Outputs k and R - k (1 <= k <= (R - 1) / 2) of the R-point DFT (R odd) of
// x_j = in[j st], from the sums and differences of the inputs j and R - j:
//   X_k = x_0 + A -/+ i B,  X_(R-k) = x_0 + A +/- i B  (upper signs forward),
//   A = sum_j (x_j + x_(R-j)) cos(2 pi j k / R),  B = sum_j (x_j - x_(R-j)) sin(2 pi j k / R),
// j = 1 .. (R - 1) / 2: half the products and a quarter of the twiddle reads of the
// two direct sums.  wr[m] = W_R^m = (cos, -sin)(2 pi m / R).
template <bool INV, class C2>
__device__ __forceinline__ void dft_pair(const C2* __restrict__ in, int st, int R, int k,
                                         const C2* __restrict__ wr, C2& xk, C2& xrk) {
  using F = typename RealOf<C2>::type;
  const C2 x0 = in[0];
  F ax = 0, ay = 0, bx = 0, by = 0;
  const int h = (R - 1) >> 1;
  int m = 0;
  for (int j = 1; j <= h; ++j) {
    m += k;
    if (m >= R) m -= R;
    const C2 xj = in[j * st], xm = in[(R - j) * st];
    const C2 wm = wr[m];
    const F c = wm.x, sn = -wm.y;
    ax += (xj.x + xm.x) * c;
    ay += (xj.y + xm.y) * c;
    bx += (xj.x - xm.x) * sn;
    by += (xj.y - xm.y) * sn;
  }
  // -i B = (by, -bx), +i B = (-by, bx)
  const F sx = INV ? -by : by, sy = INV ? bx : -bx;
  xk.x = x0.x + ax + sx;
  xk.y = x0.y + ay + sy;
  xrk.x = x0.x + ax - sx;
  xrk.y = x0.y + ay - sy;
}

// dft_stage for a stage whose small DFTs have odd length R >= 3: one thread per
// output pair (k, R - k) or per k = 0 of a small DFT (dft_pair).  Same stage
// definition as dft_stage.
template <bool STEP_A, bool INV, class C2>
__device__ __forceinline__ void dft_stage_pairs(const C2* __restrict__ in, C2* __restrict__ out,
                                                int n_lines, int ls, int es, const DftAxis& ax,
                                                const C2* __restrict__ w) {
  const int n1 = ax.n1, n2 = ax.n2;
  const int R = STEP_A ? n1 : n2;  // points of one small DFT
  const int M = STEP_A ? n2 : n1;  // small DFTs per line
  const int H = (R + 1) >> 1;      // tasks per small DFT: k = 0 and the (R - 1) / 2 pairs
  const int per_line = M * H;
  const C2* wr = STEP_A ? w : w + n1;
  const C2* wn = w + n1 + n2;
  const int st = (STEP_A ? n2 : n1) * es;
  const float inv_pl = 1.0f / (float)per_line, inv_h = 1.0f / (float)H;
  const int total = n_lines * per_line;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = fdiv(t, inv_pl), rem = t - line * per_line;
    const int q = fdiv(rem, inv_h), k = rem - q * H;
    const C2* src = in + line * ls + q * es;  // STEP_A: j2 = q; STEP_B: k1 = q
    C2* dst = out + line * ls;
    if (k == 0) {
      C2 acc = src[0];
      for (int j = 1; j < R; ++j) {
        const C2 v = src[j * st];
        acc.x += v.x;
        acc.y += v.y;
      }
      dst[(STEP_A ? q * n1 : q) * es] = acc;
    } else {
      C2 xa, xb;
      dft_pair<INV>(src, st, R, k, wr, xa, xb);
      if (STEP_A) {  // outputs k1 = k, R - k of j2 = q, times W_n^(j2 k1)
        dst[(q * n1 + k) * es] = cmul(xa, wn[q * k], INV);
        dst[(q * n1 + R - k) * es] = cmul(xb, wn[q * (R - k)], INV);
      } else {  // outputs k2 = k, R - k of k1 = q
        dst[(q + n1 * k) * es] = xa;
        dst[(q + n1 * (R - k)) * es] = xb;
      }
    }
  }
}

template <bool STEP_A, bool INV, class C2>
__device__ __forceinline__ void dft_stage(const C2* __restrict__ in, C2* __restrict__ out,
                                          int n_lines, int ls, int es, const DftAxis& ax,
                                          const C2* __restrict__ w) {
  switch (ax.n) {
    case 12: dft_stage_ct<12, STEP_A, INV>(in, out, n_lines, ls, es, w); return;
    case 16: dft_stage_ct<16, STEP_A, INV>(in, out, n_lines, ls, es, w); return;
    case 24: dft_stage_ct<24, STEP_A, INV>(in, out, n_lines, ls, es, w); return;
    case 32: dft_stage_ct<32, STEP_A, INV>(in, out, n_lines, ls, es, w); return;
    case 48: dft_stage_ct<48, STEP_A, INV>(in, out, n_lines, ls, es, w); return;
    default: break;
  }
  const int n = ax.n, n1 = ax.n1, n2 = ax.n2;
  if (((STEP_A ? n1 : n2) & 1) && (STEP_A ? n1 : n2) >= 3) {
    dft_stage_pairs<STEP_A, INV>(in, out, n_lines, ls, es, ax, w);
    return;
  }
  const C2* w1 = w;
  const C2* w2 = w + n1;
  const C2* wn = w + n1 + n2;
  const float inv_n = 1.0f / (float)n, inv_n1 = 1.0f / (float)n1;
  const int total = n_lines * n;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = fdiv(t, inv_n), r = t - line * n;
    const C2* src = in + line * ls;
    C2 acc;
    acc.x = 0;
    acc.y = 0;
    if (STEP_A) {
      const int j2 = fdiv(r, inv_n1), k1 = r - j2 * n1;
      int m = 0;
#pragma unroll 4
      for (int j1 = 0; j1 < n1; ++j1) {
        const C2 p = cmul(src[(n2 * j1 + j2) * es], w1[m], INV);
        acc.x += p.x;
        acc.y += p.y;
        m += k1;
        if (m >= n1) m -= n1;
      }
      acc = cmul(acc, wn[j2 * k1], INV);
    } else {
      const int k2 = fdiv(r, inv_n1), k1 = r - k2 * n1;
      int m = 0;
#pragma unroll 4
      for (int j2 = 0; j2 < n2; ++j2) {
        const C2 p = cmul(src[(j2 * n1 + k1) * es], w2[m], INV);
        acc.x += p.x;
        acc.y += p.y;
        m += k2;
        if (m >= n2) m -= n2;
      }
    }
    out[line * ls + r * es] = acc;
  }
}

__device__ __forceinline__ void load_twiddles(const DftAxis& ax, double2* w) {
  const int nt = ax.n1 + ax.n2 + ax.n;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) w[t] = ax.w[t];
}

__device__ __forceinline__ void load_twiddles(const DftAxis& ax, float2* w) {
  const int nt = ax.n1 + ax.n2 + ax.n;
  for (int t = threadIdx.x; t < nt; t += blockDim.x) w[t] = ax.wf[t];
}

constexpr int kDftTile = 16;
constexpr int kDftMaxTw = 64 + 16 + 64;

// Strided pass (y: stride nx, z: stride nx ny) over 16 consecutive x positions:
// blockIdx.x = x tile, blockIdx.y = the other grid index (z for the y pass, y for
// the z pass), blockIdx.z = component.  Dynamic LDS 2 * 16 (n | 1) double2.
template <bool INV, bool ZPASS>
__global__ void __launch_bounds__(256)
    k_dft_strided(int nx, int ny, int nz, int64_t N, DftAxis ax, double2* __restrict__ data,
                  const int32_t* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ double2 dft_sm[];
  __shared__ double2 w[kDftMaxTw];
  const int n = ax.n, ld = n | 1;  // one LDS row per line, odd pitch (distinct banks a row apart)
  double2* A = dft_sm;
  double2* T = dft_sm + kDftTile * ld;
  const int i0 = blockIdx.x * kDftTile, o = blockIdx.y;
  const int64_t plane = (int64_t)nx * ny;
  const int64_t stride = ZPASS ? plane : nx;
  double2* g = data + (int64_t)blockIdx.z * N + i0 + (ZPASS ? (int64_t)nx * o : plane * o);
  const int width = min(kDftTile, nx - i0);
  load_twiddles(ax, w);
  for (int t = threadIdx.x; t < kDftTile * n; t += blockDim.x) {
    const int e = t / kDftTile, ii = t - e * kDftTile;
    A[ii * ld + e] = ii < width ? g[ii + stride * e] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  dft_stage<true, INV>(A, T, kDftTile, ld, 1, ax, w);
  __syncthreads();
  dft_stage<false, INV>(T, A, kDftTile, ld, 1, ax, w);
  __syncthreads();
  for (int t = threadIdx.x; t < kDftTile * n; t += blockDim.x) {
    const int e = t / kDftTile, ii = t - e * kDftTile;
    if (ii < width) g[ii + stride * e] = A[ii * ld + e];
  }
}

}  // namespace cia
