// Energy-based RVE recurrent neural operator: one explicit update of a batch of
// material points (the reference's EnergyDissipationRVERNO.update_batch and
// _constitutive_tensors, src/concrete_impact/nn/models/rve_rno.py:249-418), with
// the exact input gradients of its two scalar energies (the SiLU chain of
// ScalarMLP.value_and_input_gradient, rve_rno.py:110-129) and, on request, the
// algorithmic tangent d stress / d strain (rve_rno.py:290-302, there six autograd
// passes; here the forward-mode derivative of the same chain).
//
// Networks: 4 hidden layers of width 128 with SiLU, a linear output (no bias for
// the energies).  With e = strain / strain_input_scale and z the latent state:
//   eq  energy psi_eq(e)           (6 inputs)
//   neq energy psi_neq([e, z])     (6 + m inputs)
//   evolution  evo([e, z])          (m outputs, direct rate; m(m+1)/2 for the
//                                    lower-triangular mobility factor)
//
// Matrix cores, FP32: v_mfma_f32_16x16x4_f32, operand layout per 64-lane wave
// (lane l, column c = l % 16, lane group g = l / 16):
//   A (16 x 4): lane l holds row c, k = g;  B (4 x 16): column c, k = g
//   D (16 x 16): lane l holds column c, rows 4 g .. 4 g + 3.
// Every layer is computed transposed, out^T = W act^T, so that a wave's 16
// columns are 16 material points (or, in tangent mode, 2 points x 8 directions)
// and one layer's D registers are the next layer's B operands without a trip
// through LDS: D register r of output tile t holds row 16 t + 4 g + r, and the
// next product consumes it in k-step (t, r), where lane group g supplies row
// 16 t + 4 g + r; the A operand (the weights) is read with the same permutation,
// one 16-byte load per lane covering the four k-steps r = 0..3.
//
// Tangent mode: the columns of a point are its value (slot 0) and the
// directional derivatives along the six strain components (slots 1..6; slot 7
// carries a zero direction).  Forward: d a_l = W_l (silu'(a_{l-1}) d a_{l-1});
// backward: d t_l = d g_l silu'(a_l) + g_l silu''(a_l) d a_l, where the value
// column's silu'(a_l) and g_l reach the direction columns by a lane swizzle
// (ds_swizzle, lane & ~7 within the wave half).
#include <math.h>

#include "common.hpp"

namespace cia {
namespace rno {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = 128;      // hidden width
constexpr int NT = H / 16;  // 16-row tiles per hidden layer
constexpr int KIN = 16;     // padded input width (6 + m <= 16)
constexpr int EVO_MAX = 48; // padded evolution outputs (3 tiles)
constexpr int kMaxLatent = 8;
constexpr int kAnchors = 14;  // psi_eq(0), d psi_eq/de (6), psi_neq(0), d psi_neq/de (6)

// Float offsets of one network in the packed weight buffer.  Layout, in order:
//   w1  [H][KIN]  input layer, input columns >= n_in zero
//   b1  [H]
//   wh  3 x ([H][H] weight, [H] bias)          hidden layers 2..4
//   wht 3 x [H][H]                              their transposes
//   w1t [KIN][H]                                transpose of w1
//   wo  [R][H], bo [R]   output layer, R = 16 ceil(n_out / 16), zero beyond n_out
//                        (bo zero for the bias-free energies)
struct Net {
  int64_t w1, b1, wh[3], bh[3], wht[3], w1t, wo, bo;
  int n_out, out_tiles;
};

struct Model {
  Net eq, neq, evo;
  int m;
  int mobility;       // 0 direct_rate, 1 mobility_gradient
  int current_state;  // 1 current_state_stress_forward_euler, 0 cs_semi_implicit_forward_euler
  float tau_inv;
  float inv_scale[6];
};

struct Outputs {
  float *stress, *latent, *rate, *drate, *diss, *free_energy, *violation, *increment, *tangent;
};

inline int64_t net_floats(int n_out) {
  const int64_t r = 16 * ((n_out + 15) / 16);
  return (int64_t)H * KIN + H + 3 * ((int64_t)H * H + H) + 3 * (int64_t)H * H + (int64_t)KIN * H +
         r * H + r;
}

inline Net net_at(int64_t base, int n_out) {
  Net n;
  int64_t o = base;
  n.w1 = o; o += (int64_t)H * KIN;
  n.b1 = o; o += H;
  for (int j = 0; j < 3; ++j) { n.wh[j] = o; o += (int64_t)H * H; n.bh[j] = o; o += H; }
  for (int j = 0; j < 3; ++j) { n.wht[j] = o; o += (int64_t)H * H; }
  n.w1t = o; o += (int64_t)KIN * H;
  n.out_tiles = (n_out + 15) / 16;
  n.wo = o; o += (int64_t)16 * n.out_tiles * H;
  n.bo = o;
  n.n_out = n_out;
  return n;
}

inline int evo_outputs(int m, int mobility) { return mobility ? m * (m + 1) / 2 : m; }

inline int64_t packed_floats(int m, int mobility) {
  return net_floats(1) * 2 + net_floats(evo_outputs(m, mobility));
}

inline Model make_model(int m, int mobility, int current_state, double tau, const double* scale) {
  Model M;
  M.eq = net_at(0, 1);
  M.neq = net_at(net_floats(1), 1);
  M.evo = net_at(2 * net_floats(1), evo_outputs(m, mobility));
  M.m = m;
  M.mobility = mobility;
  M.current_state = current_state;
  M.tau_inv = (float)(1.0 / tau);
  for (int i = 0; i < 6; ++i) M.inv_scale[i] = (float)(1.0 / scale[i]);
  return M;
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Value of lane (l & ~7) within each 32-lane half: the value column of a point.
__device__ __forceinline__ float from_value_lane(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x18));
}

__device__ __forceinline__ float logistic(float a) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-a));
}

// out[ot] rows 16 ot + 4 g + r = sum_k W[row][k] in(k) over the H rows of ``in``;
// the weights of k-step t + 1 are loaded while the products of k-step t issue.
__device__ __forceinline__ void gemm_hh(const float* __restrict__ W, const f32x4 (&in)[NT],
                                        f32x4 (&out)[NT], int c, int g) {
#pragma unroll
  for (int ot = 0; ot < NT; ++ot) out[ot] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const float* w = W + c * H + 4 * g;
  f32x4 a[NT];
#pragma unroll
  for (int ot = 0; ot < NT; ++ot) a[ot] = *(const f32x4*)&w[16 * ot * H];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    f32x4 nx[NT];
    if (t + 1 < NT) {
#pragma unroll
      for (int ot = 0; ot < NT; ++ot) nx[ot] = *(const f32x4*)&w[16 * ot * H + 16 * (t + 1)];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int ot = 0; ot < NT; ++ot) out[ot] = mfma4(a[ot][r], in[t][r], out[ot]);
    if (t + 1 < NT) {
#pragma unroll
      for (int ot = 0; ot < NT; ++ot) a[ot] = nx[ot];
    }
  }
}

// Input layer: out = W1 x, x from LDS (x[c * KIN + k], k-step s: k = 4 s + g).
__device__ __forceinline__ void gemm_in(const float* __restrict__ W1, const float* x,
                                        f32x4 (&out)[NT], int c, int g) {
#pragma unroll
  for (int ot = 0; ot < NT; ++ot) out[ot] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < KIN / 4; ++s) {
    const float b = x[c * KIN + 4 * s + g];
#pragma unroll
    for (int ot = 0; ot < NT; ++ot) out[ot] = mfma4(W1[(16 * ot + c) * KIN + 4 * s + g], b, out[ot]);
  }
}

// One 16-row output tile: d rows 4 g + r = sum_k A[c-th row of the tile][k] in(k).
__device__ __forceinline__ f32x4 gemm_tile(const float* __restrict__ A, const f32x4 (&in)[NT], int c,
                                           int g) {
  f32x4 d = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const f32x4 a = *(const f32x4*)&A[c * H + 16 * t + 4 * g];
#pragma unroll
    for (int r = 0; r < 4; ++r) d = mfma4(a[r], in[t][r], d);
  }
  return d;
}

// Hidden-layer activation.  Value columns: a = acc + b, h = silu(a), S = silu'(a).
// Direction columns (TAN): d a = acc, h = silu'(a) d a, S = silu''(a) d a, with a
// the value column's pre-activation.
template <bool TAN, bool STORE>
__device__ __forceinline__ void act_fwd(const f32x4 (&acc)[NT], const float* __restrict__ b,
                                        f32x4 (&S)[NT], f32x4 (&h)[NT], int g, bool dir) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a = acc[t][r] + (dir ? 0.0f : b[16 * t + 4 * g + r]);
      const float sg = logistic(a);
      const float s1 = sg * (1.0f + a * (1.0f - sg));
      if (!TAN) {
        h[t][r] = a * sg;
        if (STORE) S[t][r] = s1;
      } else {
        const float s2 = sg * (1.0f - sg) * (2.0f + a * (1.0f - 2.0f * sg));
        const float s1v = from_value_lane(s1), s2v = from_value_lane(s2);
        h[t][r] = dir ? s1v * a : a * sg;
        if (STORE) S[t][r] = dir ? s2v * a : s1;
      }
    }
}

// Backward activation: value columns t = g S (S = silu'(a)); direction columns
// d t = d g silu'(a) + g silu''(a) d a, i.e. acc x (value column's S) + (value
// column's acc) x S.
template <bool TAN>
__device__ __forceinline__ void act_bwd(const f32x4 (&acc)[NT], const f32x4 (&S)[NT],
                                        f32x4 (&out)[NT], bool dir) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float gv = acc[t][r], sv = S[t][r];
      if (!TAN) {
        out[t][r] = gv * sv;
      } else {
        const float s1v = from_value_lane(sv), gvv = from_value_lane(gv);
        out[t][r] = dir ? gv * s1v + gvv * sv : gv * sv;
      }
    }
}

// Scalar energy network on the wave's 16 columns of x (LDS [16][KIN]): the value
// of every value column in v[c] (if v), and the input gradient (value columns) or
// its directional derivative (direction columns) in grad[c * KIN + k], k < KIN.
template <bool TAN>
__device__ __forceinline__ void scalar_net(const float* __restrict__ P, const Net& N, const float* x,
                                           float* grad, float* v, int c, int g, bool dir) {
  f32x4 S[4][NT], h[NT], acc[NT];
  gemm_in(P + N.w1, x, acc, c, g);
  act_fwd<TAN, true>(acc, P + N.b1, S[0], h, g, dir);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    gemm_hh(P + N.wh[j], h, acc, c, g);
    act_fwd<TAN, true>(acc, P + N.bh[j], S[j + 1], h, g, dir);
  }
  const float* wo = P + N.wo;
  if (v) {
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) part += wo[16 * t + 4 * g + r] * h[t][r];
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    if (g == 0) v[c] = part;
  }
  // gradient chain from the output weights: t_4 = w_out S_4 (for direction
  // columns S_4 = silu'' d a and the value column's g is w_out, its d g zero)
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) h[t][r] = wo[16 * t + 4 * g + r] * S[3][t][r];
#pragma unroll
  for (int j = 2; j >= 0; --j) {
    gemm_hh(P + N.wht[j], h, acc, c, g);
    act_bwd<TAN>(acc, S[j], h, dir);
  }
  const f32x4 d = gemm_tile(P + N.w1t, h, c, g);
#pragma unroll
  for (int r = 0; r < 4; ++r) grad[c * KIN + 4 * g + r] = d[r];
}

// Vector network (state evolution): outputs (value columns, with bias) or their
// directional derivatives (direction columns) in out[c * EVO_MAX + o], o < n_out.
template <bool TAN>
__device__ __forceinline__ void vector_net(const float* __restrict__ P, const Net& N, const float* x,
                                           float* out, int c, int g, bool dir) {
  f32x4 h[NT], acc[NT];
  gemm_in(P + N.w1, x, acc, c, g);
  act_fwd<TAN, false>(acc, P + N.b1, h, h, g, dir);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    gemm_hh(P + N.wh[j], h, acc, c, g);
    act_fwd<TAN, false>(acc, P + N.bh[j], h, h, g, dir);
  }
  for (int ot = 0; ot < N.out_tiles; ++ot) {
    const f32x4 d = gemm_tile(P + N.wo + (int64_t)16 * ot * H, h, c, g);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = 16 * ot + 4 * g + r;
      if (o < N.n_out) out[c * EVO_MAX + o] = d[r] + (dir ? 0.0f : P[N.bo + o]);
    }
  }
}

struct WaveLds {
  float x[16 * KIN];
  float geq[16 * KIN];
  float gp[16 * KIN];
  float gc[16 * KIN];
  float ev[16 * EVO_MAX];
  float v[3][16];
};

// Lower-triangular factor entry (i, j), i >= j, row-major over the triangle
// (torch.tril_indices order).
__device__ __forceinline__ int tril(int i, int j) { return i * (i + 1) / 2 + j; }

template <bool TAN>
__global__ void __launch_bounds__(256) k_rno_update(Model M, const float* __restrict__ P,
                                                    const float* __restrict__ anchors, int64_t n,
                                                    const float* strain, const float* latent,
                                                    const float* __restrict__ time_step, Outputs o) {
  __shared__ WaveLds lds[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  WaveLds& W = lds[wave];
  const int per_wave = TAN ? 2 : 16;
  const int slot = TAN ? (c & 7) : 0;
  const bool dir = TAN && slot != 0;
  const int k_dir = slot - 1;  // strain component of a direction column (slots 1..6)
  const int64_t p = ((int64_t)blockIdx.x * 4 + wave) * per_wave + (TAN ? (c >> 3) : c);
  const bool valid = p < n;
  const int m = M.m, nin = 6 + m;
  const float tau_inv = M.tau_inv;

  float e[6], z[kMaxLatent];
#pragma unroll
  for (int i = 0; i < 6; ++i) e[i] = valid ? strain[p * 6 + i] * M.inv_scale[i] : 0.0f;
#pragma unroll
  for (int j = 0; j < kMaxLatent; ++j) z[j] = (valid && j < m) ? latent[p * m + j] : 0.0f;
  const float h_tau = valid ? time_step[p] * tau_inv : 0.0f;

  // equilibrium energy at e (direction columns: unit strain directions)
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < KIN; ++k)
      W.x[c * KIN + k] = dir ? (slot <= 6 && k == k_dir ? 1.0f : 0.0f) : (k < 6 ? e[k] : 0.0f);
  }
  __syncthreads();
  scalar_net<TAN>(P, M.eq, W.x, W.geq, W.v[0], c, g, dir);
  __syncthreads();
  // non-equilibrium energy and evolution at the previous state [e, z]
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < KIN; ++k)
      W.x[c * KIN + k] = dir ? (slot <= 6 && k == k_dir ? 1.0f : 0.0f)
                             : (k < 6 ? e[k] : (k < nin ? z[k - 6 < kMaxLatent ? k - 6 : 0] : 0.0f));
  }
  __syncthreads();
  scalar_net<TAN>(P, M.neq, W.x, W.gp, W.v[1], c, g, dir);
  vector_net<TAN>(P, M.evo, W.x, W.ev, c, g, dir);
  __syncthreads();

  // dimensionless rate r (value columns) or its strain derivative (directions)
  float r[kMaxLatent], pf[kMaxLatent];
  if (!M.mobility) {
#pragma unroll
    for (int j = 0; j < kMaxLatent; ++j) {
      r[j] = j < m ? W.ev[c * EVO_MAX + j] : 0.0f;
      pf[j] = 0.0f;
    }
  } else {
    // value column: f = -beta, p = L^T f, r = L p;  direction: d p = dL^T f + L^T df,
    // d r = dL p + L d p, with L, f, p of the point's value column c0
    const int c0 = TAN ? (c & ~7) : c;
    float f[kMaxLatent], p0[kMaxLatent];
#pragma unroll
    for (int j = 0; j < kMaxLatent; ++j) f[j] = j < m ? -W.gp[c0 * KIN + 6 + j] : 0.0f;
#pragma unroll
    for (int i = 0; i < kMaxLatent; ++i) {  // p_i = sum_{j >= i} L_ji f_j
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < kMaxLatent; ++j)
        if (j >= i && j < m && i < m) s += W.ev[c0 * EVO_MAX + tril(j, i)] * f[j];
      p0[i] = s;
    }
    if (!dir) {
#pragma unroll
      for (int i = 0; i < kMaxLatent; ++i) {  // r_i = sum_{j <= i} L_ij p_j
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < kMaxLatent; ++j)
          if (j <= i && i < m) s += W.ev[c0 * EVO_MAX + tril(i, j)] * p0[j];
        r[i] = s;
        pf[i] = p0[i];
      }
    } else {
      float df[kMaxLatent], dp[kMaxLatent];
#pragma unroll
      for (int j = 0; j < kMaxLatent; ++j) df[j] = j < m ? -W.gp[c * KIN + 6 + j] : 0.0f;
#pragma unroll
      for (int i = 0; i < kMaxLatent; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < kMaxLatent; ++j)
          if (j >= i && j < m && i < m)
            s += W.ev[c * EVO_MAX + tril(j, i)] * f[j] + W.ev[c0 * EVO_MAX + tril(j, i)] * df[j];
        dp[i] = s;
      }
#pragma unroll
      for (int i = 0; i < kMaxLatent; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < kMaxLatent; ++j)
          if (j <= i && i < m)
            s += W.ev[c * EVO_MAX + tril(i, j)] * p0[j] + W.ev[c0 * EVO_MAX + tril(i, j)] * dp[j];
        r[i] = s;
        pf[i] = 0.0f;
      }
    }
  }
  // candidate latent z + (dt / tau) r; directions carry (dt / tau) dr
  float zc[kMaxLatent];
#pragma unroll
  for (int j = 0; j < kMaxLatent; ++j) zc[j] = (dir ? 0.0f : z[j]) + h_tau * r[j];
  __syncthreads();  // every lane has read ev / gp of its value column
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < KIN; ++k) {
      float v = 0.0f;
      if (k < 6)
        v = dir ? (slot <= 6 && k == k_dir ? 1.0f : 0.0f) : e[k];
      else if (k < nin)
        v = (dir && slot > 6) ? 0.0f : zc[k - 6 < kMaxLatent ? k - 6 : 0];
      W.x[c * KIN + k] = v;
    }
  }
  __syncthreads();
  scalar_net<TAN>(P, M.neq, W.x, W.gc, W.v[2], c, g, dir);
  __syncthreads();

  if (!valid || g != 0) return;
  const float* gn = M.current_state ? W.gc : W.gp;
  if (!dir) {
    const float a_eq_v = anchors[0], a_ne_v = anchors[7];
    float stress[6], e_aeq = 0.0f, e_ane = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const float aeq = anchors[1 + i], ane = anchors[8 + i];
      stress[i] = (W.geq[c * KIN + i] - aeq) * M.inv_scale[i] + (gn[c * KIN + i] - ane) * M.inv_scale[i];
      e_aeq += e[i] * aeq;
      e_ane += e[i] * ane;
    }
    const float eq = W.v[0][c] - a_eq_v - e_aeq;
    const float prev = W.v[1][c] - a_ne_v - e_ane;
    const float cand = W.v[2][c] - a_ne_v - e_ane;
    float diss = 0.0f;
#pragma unroll
    for (int j = 0; j < kMaxLatent; ++j)
      if (j < m) diss += M.mobility ? pf[j] * pf[j] : -W.gp[c * KIN + 6 + j] * r[j];
    diss *= tau_inv;
#pragma unroll
    for (int i = 0; i < 6; ++i) o.stress[p * 6 + i] = stress[i];
#pragma unroll
    for (int j = 0; j < kMaxLatent; ++j)
      if (j < m) {
        o.latent[p * m + j] = zc[j];
        o.rate[p * m + j] = r[j] * tau_inv;
        o.drate[p * m + j] = r[j];
      }
    o.diss[p] = diss;
    o.free_energy[p] = eq + (M.current_state ? cand : prev);
    o.violation[p] = fmaxf(-diss, 0.0f);
    o.increment[p] = cand - prev;
  } else if (TAN && slot <= 6) {
#pragma unroll
    for (int i = 0; i < 6; ++i)
      o.tangent[p * 36 + i * 6 + k_dir] =
          (W.geq[c * KIN + i] + gn[c * KIN + i]) * M.inv_scale[i] * M.inv_scale[k_dir];
  }
}

// Zero-state values and strain gradients of the two energies, by the same device
// routine as the update (so a zero state gives zero stress to the last bit).
__global__ void __launch_bounds__(64) k_rno_anchors(Model M, const float* __restrict__ P,
                                                    float* __restrict__ anchors) {
  __shared__ WaveLds W;
  const int lane = threadIdx.x, c = lane & 15, g = lane >> 4;
  for (int i = lane; i < 16 * KIN; i += 64) W.x[i] = 0.0f;
  __syncthreads();
  scalar_net<false>(P, M.eq, W.x, W.geq, W.v[0], c, g, false);
  scalar_net<false>(P, M.neq, W.x, W.gp, W.v[1], c, g, false);
  __syncthreads();
  if (lane == 0) {
    anchors[0] = W.v[0][0];
    anchors[7] = W.v[1][0];
    for (int i = 0; i < 6; ++i) {
      anchors[1 + i] = W.geq[i];
      anchors[8 + i] = W.gp[i];
    }
  }
}

}  // namespace rno
}  // namespace cia

using namespace cia;

extern "C" {

cia_status cia_rno_packed_size(int32_t latent_dimension, int32_t evolution, int64_t* floats) {
  CIA_ARG(floats, "cia_rno_packed_size: null output");
  CIA_ARG(latent_dimension >= 1 && latent_dimension <= rno::kMaxLatent,
          "cia_rno_packed_size: latent dimension must be 1..8");
  CIA_ARG(evolution == CIA_RNO_DIRECT_RATE || evolution == CIA_RNO_MOBILITY_GRADIENT,
          "cia_rno_packed_size: unknown evolution");
  *floats = rno::packed_floats(latent_dimension, evolution == CIA_RNO_MOBILITY_GRADIENT);
  return CIA_OK;
}

cia_status cia_rno_anchors(void* stream_, const float* packed, int32_t latent_dimension,
                           int32_t evolution, float* anchors) {
  hipStream_t stream = (hipStream_t)stream_;
  CIA_ARG(packed && anchors, "cia_rno_anchors: null pointer");
  CIA_ARG(latent_dimension >= 1 && latent_dimension <= rno::kMaxLatent,
          "cia_rno_anchors: latent dimension must be 1..8");
  CIA_ARG(evolution == CIA_RNO_DIRECT_RATE || evolution == CIA_RNO_MOBILITY_GRADIENT,
          "cia_rno_anchors: unknown evolution");
  const double ones[6] = {1, 1, 1, 1, 1, 1};
  const rno::Model M = rno::make_model(latent_dimension, evolution == CIA_RNO_MOBILITY_GRADIENT, 1,
                                       1.0, ones);
  hipLaunchKernelGGL(rno::k_rno_anchors, dim3(1), dim3(64), 0, stream, M, packed, anchors);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_rno_update(void* stream_, const float* packed, int32_t latent_dimension, int32_t evolution,
                          int32_t integrator, double reference_time_scale,
                          const double* strain_input_scale, const float* anchors, int64_t n,
                          const float* strain, const float* latent, const float* time_step,
                          float* stress, float* latent_out, float* state_rate,
                          float* dimensionless_state_rate, float* dissipation, float* free_energy,
                          float* thermodynamic_violation, float* nonequilibrium_energy_increment,
                          float* tangent) {
  hipStream_t stream = (hipStream_t)stream_;
  CIA_ARG(n >= 0, "cia_rno_update: negative point count");
  CIA_ARG(latent_dimension >= 1 && latent_dimension <= rno::kMaxLatent,
          "cia_rno_update: latent dimension must be 1..8");
  CIA_ARG(evolution == CIA_RNO_DIRECT_RATE || evolution == CIA_RNO_MOBILITY_GRADIENT,
          "cia_rno_update: unknown evolution");
  CIA_ARG(integrator == CIA_RNO_CS_SEMI_IMPLICIT || integrator == CIA_RNO_CURRENT_STATE,
          "cia_rno_update: unknown integrator");
  CIA_ARG(reference_time_scale > 0.0, "cia_rno_update: reference time scale must be positive");
  CIA_ARG(strain_input_scale, "cia_rno_update: null strain input scale");
  for (int i = 0; i < 6; ++i)
    CIA_ARG(strain_input_scale[i] > 0.0, "cia_rno_update: strain input scales must be positive");
  if (n == 0) return CIA_OK;
  CIA_ARG(packed && anchors && strain && latent && time_step && stress && latent_out && state_rate &&
              dimensionless_state_rate && dissipation && free_energy && thermodynamic_violation &&
              nonequilibrium_energy_increment,
          "cia_rno_update: null pointer");
  const rno::Model M = rno::make_model(latent_dimension, evolution == CIA_RNO_MOBILITY_GRADIENT,
                                       integrator == CIA_RNO_CURRENT_STATE, reference_time_scale,
                                       strain_input_scale);
  const rno::Outputs o{stress, latent_out, state_rate, dimensionless_state_rate, dissipation,
                       free_energy, thermodynamic_violation, nonequilibrium_energy_increment, tangent};
  if (tangent) {
    const int64_t blocks = (n + 7) / 8;  // 2 points per wave
    CIA_ARG(blocks <= INT32_MAX, "cia_rno_update: too many points");
    hipLaunchKernelGGL(rno::k_rno_update<true>, dim3((unsigned)blocks), dim3(256), 0, stream, M,
                       packed, anchors, n, strain, latent, time_step, o);
  } else {
    const int64_t blocks = (n + 63) / 64;  // 16 points per wave
    CIA_ARG(blocks <= INT32_MAX, "cia_rno_update: too many points");
    hipLaunchKernelGGL(rno::k_rno_update<false>, dim3((unsigned)blocks), dim3(256), 0, stream, M,
                       packed, anchors, n, strain, latent, time_step, o);
  }
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

}  // extern "C"
