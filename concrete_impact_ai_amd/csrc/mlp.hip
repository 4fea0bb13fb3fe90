// Recurrent MLP material-point surrogate: BF16 MFMA GEMMs with FP32 accumulation
// and fused bias / GELU / GELU' epilogues, split-K weight gradients, Adam, and a
// fused all-layer inference kernel.
//
// The surrogate plays the role of the reference's recurrent material-point
// network (EnergyDissipationRVERNO, src/concrete_impact/nn/models/rve_rno.py:162-476,
// trained by BPTT in src/concrete_impact/nn/training/rve_rno.py:113-415) with the
// architecture 13 -> 256 -> 256 -> 256 -> 256 -> 12 (GELU): input = strain
// increment (6), log strain-rate feature (1), tanh of the latent state (6); output = stress (6),
// next latent state (6).
//
// Matrix cores: v_mfma_f32_16x16x32_bf16.  Operand layout per 64-lane wave:
//   A (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7
//   B (32 x 16): lane l holds col l % 16, k = 8 (l / 16) .. + 7
//   D (16 x 16): lane l holds col l % 16, rows 4 (l / 16) .. + 3
// (checked against a float64 reference by tests/test_gpu_mlp.py).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "common.hpp"

namespace cia {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GM = 128, GN = 128, GK = 32, GLD = GK + 8;  // LDS row stride 80 B

enum { EPI_LINEAR = 0, EPI_GELU = 1, EPI_DGELU = 2 };

struct GemmArgs {
  int64_t M, N, K;
  const __bf16* A;
  int64_t sAm, sAk;  // A(m, k) = A[m sAm + k sAk]
  const __bf16* B;
  int64_t sBk, sBn;  // B(k, n) = B[k sBk + n sBn]
  int64_t ones_col;  // column n of B that reads 1 (bias gradient), -1 none
  const float* bias; // [N] added in the epilogue (split 0 only), or null
  const float* zin;  // EPI_DGELU: pre-activation z(m, n) = zin[m ldzin + n]
  int64_t ldzin;
  float* Zf;         // EPI_GELU: pre-activation out (fp32), or null
  int64_t ldzf;
  float* Cf;         // fp32 out, or null
  int64_t ldcf;
  __bf16* Cb;        // bf16 out, or null
  int64_t ldcb;
  int64_t k_split;   // K range per blockIdx.z (multiple of GK)
  int64_t split_stride;  // Cf offset per split
};

// GELU in its tanh form (Hendrycks & Gimpel, "Gaussian Error Linear Units"):
// 0.5 x (1 + tanh(k (x + 0.044715 x^3))), k = sqrt(2 / pi), evaluated as
// x sigmoid(2 k (x + 0.044715 x^3)) with the hardware exp2 and reciprocal: about
// eight VALU instructions, where the epilogue of every hidden layer evaluates 32
// of them per lane and MFMA subtile.
constexpr float GELU_2K = 1.5957691216057308f;  // 2 sqrt(2 / pi)
constexpr float GELU_C = 0.044715f;
__device__ __forceinline__ float gelu(float x) {
  const float v = GELU_2K * x * (1.0f + GELU_C * x * x);
  return x * __builtin_amdgcn_rcpf(1.0f + __expf(-v));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float v = GELU_2K * x * (1.0f + GELU_C * x * x);
  const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-v));
  const float dv = GELU_2K * (1.0f + 3.0f * GELU_C * x * x);
  return sg + x * sg * (1.0f - sg) * dv;
}

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 As[GM * GLD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[GN * GLD];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t m0 = (int64_t)blockIdx.x * GM, n0 = (int64_t)blockIdx.y * GN;
  const int64_t kb = (int64_t)blockIdx.z * g.k_split;
  const int64_t ke = (kb + g.k_split < g.K) ? kb + g.k_split : g.K;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const __bf16 zero = (__bf16)0.0f, one = (__bf16)1.0f;
  for (int64_t k0 = kb; k0 < ke; k0 += GK) {
    if (g.sAk == 1) {
      for (int i = tid; i < GM * GK; i += 256) {
        const int r = i / GK, c = i % GK;
        const int64_t m = m0 + r, k = k0 + c;
        As[r * GLD + c] = (m < g.M && k < ke) ? g.A[m * g.sAm + k] : zero;
      }
    } else {
      for (int i = tid; i < GM * GK; i += 256) {
        const int r = i % GM, c = i / GM;
        const int64_t m = m0 + r, k = k0 + c;
        As[r * GLD + c] = (m < g.M && k < ke) ? g.A[m * g.sAm + k * g.sAk] : zero;
      }
    }
    if (g.sBk == 1) {
      for (int i = tid; i < GN * GK; i += 256) {
        const int r = i / GK, c = i % GK;
        const int64_t n = n0 + r, k = k0 + c;
        __bf16 v = zero;
        if (n < g.N && k < ke) v = (n == g.ones_col) ? one : g.B[k + n * g.sBn];
        Bs[r * GLD + c] = v;
      }
    } else {
      for (int i = tid; i < GN * GK; i += 256) {
        const int r = i % GN, c = i / GN;
        const int64_t n = n0 + r, k = k0 + c;
        __bf16 v = zero;
        if (n < g.N && k < ke) v = (n == g.ones_col) ? one : g.B[k * g.sBk + n * g.sBn];
        Bs[r * GLD + c] = v;
      }
    }
    __syncthreads();
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    bf16x8 af[4], bfr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) af[i] = *(const bf16x8*)&As[(wm + 16 * i + lr) * GLD + lk];
#pragma unroll
    for (int j = 0; j < 4; ++j) bfr[j] = *(const bf16x8*)&Bs[(wn + 16 * j + lr) * GLD + lk];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  float* Cf = g.Cf ? g.Cf + (int64_t)blockIdx.z * g.split_stride : nullptr;
  const bool add_bias = g.bias && blockIdx.z == 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t n = n0 + wn + 16 * j + (lane & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        if (m >= g.M) continue;
        float v = acc[i][j][r];
        if (add_bias) v += g.bias[n];
        if (EPI == EPI_GELU) {
          if (g.Zf) g.Zf[m * g.ldzf + n] = v;
          v = gelu(v);
        } else if (EPI == EPI_DGELU) {
          v *= gelu_grad(g.zin[m * g.ldzin + n]);
        }
        if (Cf) Cf[m * g.ldcf + n] = v;
        if (g.Cb) g.Cb[m * g.ldcb + n] = (__bf16)v;
      }
    }
}

constexpr int64_t kMaxWgradSplits = 64;

// Static scratch of cia_mlp_linear_backward_weight_noalloc: the split partials of
// a 256 x 257 weight gradient at the full split count, i.e. exactly the
// workspace cia_mlp_backward_weight_workspace(256, 256) asks for (17.1 MB per GPU).
constexpr int64_t kNoAllocScratchFloats = (kMaxWgradSplits + 1) * 256 * 257;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

// Sum the split partials [S][N][Kx] into dW [N][K] (column K -> db), fixed order.
__global__ void k_wgrad_reduce(int S, int64_t N, int64_t K, int64_t Kx, const float* __restrict__ part,
                               float* __restrict__ dW, float* __restrict__ db, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = N * Kx;
  if (i >= per) return;
  float a = 0.f;
  for (int s = 0; s < S; ++s) a += part[s * per + i];
  const int64_t n = i / Kx, k = i - n * Kx;
  float* dst = (k < K) ? &dW[n * K + k] : &db[n];
  *dst = accumulate ? *dst + a : a;
}

// Adam (Kingma & Ba 2015) with bias correction; refreshes the bf16 copy.
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ grad,
                       float* __restrict__ m1, float* __restrict__ m2, __bf16* __restrict__ pb,
                       float lr, float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i] + wd * p[i];
  const float a = b1 * m1[i] + (1.f - b1) * gr;
  const float b = b2 * m2[i] + (1.f - b2) * gr * gr;
  m1[i] = a;
  m2[i] = b;
  const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[i] = v;
  if (pb) pb[i] = (__bf16)v;
}

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ x, __bf16* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// ---------------------------------------------------------------------------
// Fused inference: 13 -> 256 -> 256 -> 256 -> 256 -> 12 for a tile of FM points
// per workgroup, the activations (bf16, row-major [point][feature]) in one LDS
// buffer of 66 KB, so two workgroups share a CU and one wave's GELU epilogue
// overlaps the other's matrix-core work.  Each of the 4 waves computes a
// 64-feature slice of a layer as the transposed product out^T = W act^T: the A
// operand is 16 weight rows (one 16-byte load per lane from the L2-resident
// weights, prefetched one k-step ahead), the B operand 16 points' activations
// (one 16-byte LDS load per lane), and the accumulator holds 4 consecutive
// features of one point per lane, written back with one 8-byte LDS store.
constexpr int FM = 128;          // points per workgroup
constexpr int FH = 256;          // hidden width
constexpr int FIN = 13, FOUT = 12;
constexpr int FLD = FH + 8;      // LDS row stride (bf16), 528 B: conflict-free 16-byte reads
constexpr size_t FUSED_ACT_BYTES = (size_t)FM * FLD * sizeof(__bf16);

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// A operand of a hidden layer: rows n of W (layer 0: W1 [256][13], k >= 13 read as zero).
__device__ __forceinline__ bf16x8 weight_frag(const __bf16* __restrict__ Wb, int layer, int n,
                                              int k0) {
  bf16x8 f;
  if (layer == 0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) f[t] = (k0 + t) < FIN ? Wb[n * FIN + k0 + t] : (__bf16)0.0f;
  } else {
    f = *(const bf16x8*)&Wb[FH * FIN + (int64_t)(layer - 1) * FH * FH + n * FH + k0];
  }
  return f;
}

// Hidden layers 1..4 on one FM-point tile: act holds the 13 input features per row
// (columns 13..31 zero) on entry and the last hidden activations on exit.
__device__ __forceinline__ void fused_hidden(__bf16* act, const __bf16* __restrict__ Wb,
                                             const float* __restrict__ bias, int wave, int lane) {
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const int wn = wave * 64;
  for (int layer = 0; layer < 4; ++layer) {
    f32x4 acc[4][FM / 16];  // [feature subtile j][point subtile i]
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < FM / 16; ++i) acc[j][i] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int ksteps = layer == 0 ? 1 : FH / 32;
    bf16x8 wf[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) wf[j] = weight_frag(Wb, layer, wn + 16 * j + lr, lk);
    for (int ks = 0; ks < ksteps; ++ks) {
      bf16x8 wn_next[4];
      if (ks + 1 < ksteps) {
#pragma unroll
        for (int j = 0; j < 4; ++j) wn_next[j] = weight_frag(Wb, layer, wn + 16 * j + lr, 32 * (ks + 1) + lk);
      }
#pragma unroll
      for (int i = 0; i < FM / 16; ++i) {
        const bf16x8 af = *(const bf16x8*)&act[(16 * i + lr) * FLD + 32 * ks + lk];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[j], af, acc[j][i], 0, 0, 0);
      }
      if (ks + 1 < ksteps) {
#pragma unroll
        for (int j = 0; j < 4; ++j) wf[j] = wn_next[j];
      }
    }
    __syncthreads();  // every wave has read this layer's input
    const float* b = bias + layer * FH;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n0 = wn + 16 * j + 4 * (lane >> 4);
      const float b0 = b[n0], b1 = b[n0 + 1], b2 = b[n0 + 2], b3 = b[n0 + 3];
#pragma unroll
      for (int i = 0; i < FM / 16; ++i) {
        bf16x4 o;
        o[0] = (__bf16)gelu(acc[j][i][0] + b0);
        o[1] = (__bf16)gelu(acc[j][i][1] + b1);
        o[2] = (__bf16)gelu(acc[j][i][2] + b2);
        o[3] = (__bf16)gelu(acc[j][i][3] + b3);
        *(bf16x4*)&act[(16 * i + lr) * FLD + n0] = o;
      }
    }
    __syncthreads();
  }
}

// Output layer on the activations left by fused_hidden: wave w takes points
// 32 w .. 32 w + 31 (two 16-point subtiles); out(point, column, value) receives the
// FOUT outputs of every point.
template <class Out>
__device__ __forceinline__ void fused_output(const __bf16* act, const __bf16* __restrict__ Wb,
                                             const float* __restrict__ bias, int wave, int lane,
                                             Out out) {
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const __bf16* W5 = Wb + FH * FIN + 3 * FH * FH;  // [12][256]
  f32x4 acc[2] = {(f32x4){0.f, 0.f, 0.f, 0.f}, (f32x4){0.f, 0.f, 0.f, 0.f}};
  for (int ks = 0; ks < FH / 32; ++ks) {
    bf16x8 wf;
    if (lr < FOUT)
      wf = *(const bf16x8*)&W5[lr * FH + 32 * ks + lk];
    else
#pragma unroll
      for (int t = 0; t < 8; ++t) wf[t] = (__bf16)0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const bf16x8 af = *(const bf16x8*)&act[(32 * wave + 16 * i + lr) * FLD + 32 * ks + lk];
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, af, acc[i], 0, 0, 0);
    }
  }
  const int n0 = 4 * (lane >> 4);  // outputs n0 .. n0 + 3 of point lr
  if (n0 < FOUT) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out(32 * wave + 16 * i + lr, n0 + r, acc[i][r] + bias[4 * FH + n0 + r]);
  }
}

__global__ void __launch_bounds__(256, 2) k_mlp_fused(int64_t M, const float* __restrict__ X,
                                                   const __bf16* __restrict__ Wb,
                                                   const float* __restrict__ bias,
                                                   float* __restrict__ Y) {
  extern __shared__ __attribute__((aligned(16))) __bf16 act[];  // [FM][FLD]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t m0 = (int64_t)blockIdx.x * FM;
  // inputs -> bf16, columns 13..31 zero
  for (int i = tid; i < FM * 32; i += 256) {
    const int r = i >> 5, c = i & 31;
    const int64_t m = m0 + r;
    act[r * FLD + c] = (__bf16)((m < M && c < FIN) ? X[m * FIN + c] : 0.0f);
  }
  __syncthreads();
  fused_hidden(act, Wb, bias, wave, lane);
  fused_output(act, Wb, bias, wave, lane, [&](int row, int col, float v) {
    const int64_t m = m0 + row;
    if (m < M) Y[m * FOUT + col] = v;
  });
}

// ---------------------------------------------------------------------------
// Surrogate material-point update (the material protocol of the reference's
// RVERNOMaterialAdapter.update, src/concrete_impact/nn/deployment.py:97-165, for
// this build's network).  The network predicts inelastic strain increments; the
// stress is the homogenised elastic stiffness C times the elastic strain.  Each
// point keeps an anchor: the strain, inelastic strain, latent and elapsed time of
// its last committed network step.  Every call evaluates the network once from the
// anchor,
//   d = eps - eps_a,  tau = age + dt,  rate = |d|_2 / tau,
//   x = [d / eps_scale, clamp_lo(log10(rate / rate_ref) / rate_scale), tanh(z_a)],
//   ein' = ein_a + inelastic_step eps_scale o[0:6],  sigma = C (eps - ein'),
//   z' = z_a + o[6:12],
// and commits (anchor <- eps, ein', z', age 0) only when |d|_inf reaches
// commit_strain (or tau reaches commit_time > 0); otherwise the anchor is kept and
// its age advanced.  The network therefore always sees increments of the size it
// was trained on, whatever the caller's time step.  Rates below the training
// range are clamped to its lower end (the quasi-static limit of the viscoplastic
// response); a point trips the gate when |d|_inf / eps_scale exceeds gate_margin
// d_eps_max, the rate feature exceeds rate_hi + rate_margin, |z'|_inf exceeds
// gate_margin latent_max, or a result is not finite.

struct SurGate {
  double eps_scale, rate_ref, rate_scale, inelastic_step;
  double d_eps_max, rate_lo, rate_hi, latent_max, gate_margin, rate_margin;
  double commit_strain, commit_time;
  double C[36];  // row-major, engineering Voigt
};

struct SurPoint {
  double d[6];
  double dinf, tau, feat_raw, feat;
};

__device__ __forceinline__ SurPoint sur_inputs(int64_t n, int64_t m, const double* __restrict__ strain,
                                               double dt, const SurGate& g,
                                               const double* __restrict__ a_strain,
                                               const float* __restrict__ a_age) {
  SurPoint p;
  double n2 = 0.0;
  p.dinf = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    p.d[c] = strain[c * n + m] - a_strain[c * n + m];
    n2 += p.d[c] * p.d[c];
    p.dinf = fmax(p.dinf, fabs(p.d[c]));
  }
  p.tau = (double)a_age[m] + dt;
  const double rate = sqrt(n2) / p.tau;
  p.feat_raw = log10(fmax(rate, 1e-30) / g.rate_ref) / g.rate_scale;
  p.feat = fmax(p.feat_raw, g.rate_lo);
  return p;
}

__global__ void __launch_bounds__(256, 2)
    k_sur_update(int64_t n, const double* __restrict__ strain, double dt,
                 const __bf16* __restrict__ Wb, const float* __restrict__ bias, SurGate g,
                 const double* __restrict__ a_strain, const double* __restrict__ a_inel,
                 const float* __restrict__ a_lat, const float* __restrict__ a_age,
                 const uint8_t* __restrict__ fb_in, double* __restrict__ stress,
                 double* __restrict__ n_strain, double* __restrict__ n_inel,
                 float* __restrict__ n_lat, float* __restrict__ n_age, uint8_t* __restrict__ fb_out,
                 unsigned int* __restrict__ trips, const double* __restrict__ cur_in,
                 double* __restrict__ cur_out, double* __restrict__ psi,
                 double* __restrict__ diss) {
  extern __shared__ __attribute__((aligned(16))) __bf16 act[];  // [FM][FLD], then o_s
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t m0 = (int64_t)blockIdx.x * FM;
  float* o_s = (float*)((char*)act + FUSED_ACT_BYTES);  // [FM][FOUT + 1] network outputs
  // features: one thread per point of the tile
  if (tid < FM) {
    const int64_t m = m0 + tid;
    __bf16* row = act + tid * FLD;
    if (m < n) {
      const SurPoint p = sur_inputs(n, m, strain, dt, g, a_strain, a_age);
#pragma unroll
      for (int c = 0; c < 6; ++c) row[c] = (__bf16)(float)(p.d[c] / g.eps_scale);
      row[6] = (__bf16)(float)p.feat;
#pragma unroll
      for (int c = 0; c < 6; ++c) row[7 + c] = (__bf16)tanhf(a_lat[c * n + m]);
    } else {
#pragma unroll
      for (int c = 0; c < FIN; ++c) row[c] = (__bf16)0.0f;
    }
#pragma unroll
    for (int c = FIN; c < 32; ++c) row[c] = (__bf16)0.0f;
  }
  __syncthreads();
  fused_hidden(act, Wb, bias, wave, lane);
  fused_output(act, Wb, bias, wave, lane,
               [&](int row, int col, float v) { o_s[row * (FOUT + 1) + col] = v; });
  __syncthreads();
  if (tid < FM) {  // waves 0 and 1, whole
    const int64_t m = m0 + tid;
    bool new_trip = false;
    if (m < n) {
      const SurPoint p = sur_inputs(n, m, strain, dt, g, a_strain, a_age);
      const float* o = o_s + tid * (FOUT + 1);
      double ein[6], eel[6], sig[6];
      float z[6];
      bool finite = true;
      double zinf = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        ein[c] = a_inel[c * n + m] + g.inelastic_step * g.eps_scale * (double)o[c];
        eel[c] = strain[c * n + m] - ein[c];
        z[c] = a_lat[c * n + m] + o[6 + c];
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) v += g.C[6 * r + c] * eel[c];
        sig[r] = v;
        finite = finite && isfinite(v) && isfinite(z[r]);
        zinf = fmax(zinf, fabs((double)z[r]));
      }
      const bool trip = p.dinf / g.eps_scale > g.gate_margin * g.d_eps_max ||
                        p.feat_raw > g.rate_hi + g.rate_margin ||
                        zinf > g.gate_margin * g.latent_max || !finite;
      const bool commit = p.dinf >= g.commit_strain || (g.commit_time > 0.0 && p.tau >= g.commit_time);
#pragma unroll
      for (int c = 0; c < 6; ++c) stress[c * n + m] = sig[c];
      // energies: psi = 1/2 (strain - ein') . sigma; D = sigma . (ein' - ein_prev) / dt with
      // ein_prev the previous call's inelastic strain (cur_in)
      if (psi) {
        double e = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) e += eel[c] * sig[c];
        psi[m] = 0.5 * e;
      }
      if (diss) {
        double e = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) e += sig[c] * (ein[c] - cur_in[c * n + m]);
        diss[m] = e / dt;
      }
      if (cur_out) {
#pragma unroll
        for (int c = 0; c < 6; ++c) cur_out[c * n + m] = ein[c];
      }
      if (commit) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          n_strain[c * n + m] = strain[c * n + m];
          n_inel[c * n + m] = ein[c];
          n_lat[c * n + m] = z[c];
        }
        n_age[m] = 0.0f;
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          n_strain[c * n + m] = a_strain[c * n + m];
          n_inel[c * n + m] = a_inel[c * n + m];
          n_lat[c * n + m] = a_lat[c * n + m];
        }
        n_age[m] = (float)p.tau;
      }
      const uint8_t was = fb_in[m];
      fb_out[m] = (uint8_t)(was | (trip ? 1 : 0));
      new_trip = trip && !was;
    }
    const unsigned long long mask = __ballot(new_trip);
    if (lane == 0 && mask) atomicAdd(trips, (unsigned int)__popcll(mask));
  }
}

// ---------------------------------------------------------------------------
// Fused BPTT training step (the batch step of train_rve_rno,
// src/concrete_impact/nn/training/rve_rno.py:200-415, for this build's network and
// its path-normalised stress loss).  One call computes the loss and the weight and
// bias gradients of a batch of B paths over a window of L steps from step 0:
//
//   k_bptt_prep    training weight blob (bf16, rows padded to multiples of 32
//                  elements): the forward layers and the transposed layers of the
//                  backward pass, from the fp32 master weights
//   k_bptt_fwd     one workgroup per tile of 16 paths runs the L-step recurrence
//                  (network, inelastic strain, latent, stress, loss) and stores the
//                  bf16 network inputs x_t (transposed) and q_t = Cs^T dL/dsigma_t
//   k_bptt_bwd     the same tiles in reverse time: the hidden layers are recomputed
//                  from x_t (GELU' kept in registers), dL/do_t is formed from the
//                  adjoints of the inelastic strain and the latent, and the layers are
//                  back-propagated; layer inputs and pre-activation gradients go to
//                  HBM transposed ([feature][point-step], bf16)
//   k_wgrad_nt     dW = sum over point-steps of dz a^T, split over point-steps into
//                  fp32 partials, with the bias row sums
//   k_bptt_reduce  partials -> gradient blobs, loss partials -> loss
//
// Point-step index m = t Bp + b (Bp = B rounded up to the tile).  Padding paths
// have zero inputs and zero loss weight, so they contribute nothing.
//
// Recurrence (scaled units): x_t = [f_t (6), r_t, tanh(z_t) (6)], o_t = MLP(x_t),
// ein_t = ein_{t-1} + c o_t[0:6], sigma_t = Cs (eps_t - ein_t), z_{t+1} = z_t + o_t[6:12],
// loss = sum_b w_b sum_{t,i} (sigma_t - target_t)_i^2 / (B L 6).
// Adjoints in reverse time: G_t = G_{t+1} - q_t (dL/dein_t), do_t = [c G_t, gz_{t+1}],
// gz_t = gz_{t+1} + dL/dx_t[7:13] (1 - tanh^2 z_t) (dL/dz_t).
//
// Every layer is the transposed product out^T = W act^T on the matrix cores: the A
// operand is 16 weight rows (16-byte loads from the L2-resident blob, issued one
// layer ahead), the B operand 16 points' activations from LDS; 4 waves (one per
// SIMD, up to 512 registers each) share the 256 rows of a layer (64 each).  The
// 16-row layers (output layer, input gradient) split their 8 k-steps over the waves
// and sum the partials through LDS.

constexpr int kTrainThreads = 256;
constexpr int TLD = FH + 8;                 // LDS row stride of activations (bf16)
constexpr int DLD = 40;                     // LDS row stride of dL/do (bf16)
// training weight blob (bf16 element offsets)
constexpr int64_t oF0 = 0;                  // W1 [256][32], columns >= 13 zero
constexpr int64_t oF1 = oF0 + 256 * 32;     // W2 [256][256]
constexpr int64_t oF2 = oF1 + 65536;        // W3
constexpr int64_t oF3 = oF2 + 65536;        // W4
constexpr int64_t oF4 = oF3 + 65536;        // W5 [16][256], rows >= 12 zero
constexpr int64_t oT4 = oF4 + 16 * 256;     // W5^T [256][32], columns >= 12 zero
constexpr int64_t oT3 = oT4 + 256 * 32;     // W4^T [256][256]
constexpr int64_t oT2 = oT3 + 65536;        // W3^T
constexpr int64_t oT1 = oT2 + 65536;        // W2^T
constexpr int64_t oT0 = oT1 + 65536;        // W1^T [16][256], rows >= 13 zero
constexpr int64_t kTrainBlob = oT0 + 16 * 256;
// fp32 master blob: W1 [256][13] | W2..W4 [256][256] | W5 [12][256]
constexpr int64_t mW1 = 0, mW2 = 256 * 13, mW3 = mW2 + 65536, mW4 = mW3 + 65536,
                  mW5 = mW4 + 65536;

// Training blob in fragment order: a matrix of R rows (multiple of 16) and K columns
// (multiple of 32) is stored as 16-row blocks rb x 32-column k-steps ks x 64 lanes x 8
// elements, element e of lane l being M[16 rb + l % 16][32 ks + 8 (l / 16) + e] -- the
// A operand of v_mfma_f32_16x16x32_bf16 -- so that every fragment load of a wave is
// one contiguous 1 KB read.
__device__ __forceinline__ float master_value(const float* __restrict__ w, int seg, int r, int c) {
  switch (seg) {
    case 0: return c < FIN ? w[mW1 + r * FIN + c] : 0.f;            // W1 [256][32]
    case 1: return w[mW2 + r * 256 + c];                             // W2
    case 2: return w[mW3 + r * 256 + c];                             // W3
    case 3: return w[mW4 + r * 256 + c];                             // W4
    case 4: return r < FOUT ? w[mW5 + r * 256 + c] : 0.f;            // W5 [16][256]
    case 5: return c < FOUT ? w[mW5 + c * 256 + r] : 0.f;            // W5^T [256][32]
    case 6: return w[mW4 + c * 256 + r];                             // W4^T
    case 7: return w[mW3 + c * 256 + r];                             // W3^T
    case 8: return w[mW2 + c * 256 + r];                             // W2^T
    default: return r < FIN ? w[mW1 + c * FIN + r] : 0.f;            // W1^T [16][256]
  }
}

__global__ void k_bptt_prep(const float* __restrict__ w, __bf16* __restrict__ blob) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTrainBlob) return;
  const int64_t off[11] = {oF0, oF1, oF2, oF3, oF4, oT4, oT3, oT2, oT1, oT0, kTrainBlob};
  const int cols[10] = {32, 256, 256, 256, 256, 32, 256, 256, 256, 256};
  int seg = 0;
  while (i >= off[seg + 1]) ++seg;
  const int64_t j = i - off[seg];
  const int e = (int)(j % 8), l = (int)((j / 8) % 64);
  const int64_t f = j / 512;
  const int KS = cols[seg] / 32;
  const int ks = (int)(f % KS), rb = (int)(f / KS);
  const int r = 16 * rb + (l & 15), c = 32 * ks + 8 * (l >> 4) + e;
  blob[i] = (__bf16)master_value(w, seg, r, c);
}

struct BpttArgs {
  int B, Bp, L, T, P, s0;    // ..., window start step
  int64_t M;                 // L Bp point-steps
  const int32_t* idx;        // [B] path of each batch row
  const float* feat;         // [P][T][8]: d_eps / eps_scale (6), rate feature, unused
  const float* eps;          // [P][T][6]: strain / eps_scale
  const float* target;       // [P][T][6]: stress / stress_scale
  const float* pw;           // [P] path weights
  float Cs[36];              // stiffness eps_scale / stress_scale, row-major
  float cstep, inv_count;    // inelastic step; 1 / (B L 6)
  const __bf16* blob;        // training weight blob
  const float* bias;         // b1..b4 [256] | b5 [12]
  __bf16* xT;                // inputs x_t, 16 features, tiled transposed (tt_index)
  float* q;                  // [M][12]: Cs^T dL/dsigma_t (6), tanh(z_t) (6)
  __bf16* aT;                // 4 x 256 M: hidden activations h1..h4, tiled transposed
  __bf16* dzT;               // 4 x 256 M: dL/dz1..dz4, tiled transposed
  __bf16* doT;               // dL/do, 16 features, tiled transposed
  float* loss_part;          // [tiles]
  const float* state_in;     // [B][12] latent and inelastic strain at the window start, or null
  float* state_out;          // [B][12] the same at the window end, or null
};

// A-operand fragments of the 64 rows 64 wave .. 64 wave + 63 of a layer: 4 row
// blocks j of 16 rows, KS k-steps of 32.
template <int KS>
struct Frags {
  bf16x8 f[KS][4];
};

// The weight blob is read through a buffer resource: one per-lane byte offset for
// every fragment of a layer, the layer, row-block and k-step offsets in scalar
// registers (flat addressing needed a 64-bit address per fragment).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t blob_rsrc(const __bf16* blob) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)blob, (short)0, (int)(kTrainBlob * 2), 0x00020000);
}

__device__ __forceinline__ bf16x8 blob_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// Fragments of the wave's 64 rows of the fragment-ordered matrix at blob element offset
// base with ld columns: one lane offset for all of them, the rest in scalar registers.
template <int KS>
__device__ __forceinline__ void load_frags(Frags<KS>& w, __amdgpu_buffer_rsrc_t r, int64_t base,
                                           int ld, int wave, int lane) {
  const int kst = ld / 32;
  // an opaque zero keeps the scalar offsets from being hoisted out of the step loop
  // (192 loop-invariant offsets would not fit the scalar registers)
  int z0;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z0));
  const int b = z0 + (int)(base * 2) + 4 * wave * kst * 1024;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j) w.f[ks][j] = blob_load(r, lane * 16, b + (j * kst + ks) * 1024);
}

// Fragments loaded once per kernel are passed through an empty asm statement: the
// compiler then keeps them in registers instead of re-issuing their loads at every
// use inside the step loop (where the load latency would be exposed).
template <int KS>
__device__ __forceinline__ void pin(Frags<KS>& w) {
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j) asm volatile("" : "+v"(w.f[ks][j]));
}

__device__ __forceinline__ void pin(bf16x8 (&w)[2]) {
  asm volatile("" : "+v"(w[0]));
  asm volatile("" : "+v"(w[1]));
}

// acc[j][r] = (W act^T)[64 wave + 16 j + 4 (lane / 16) + r][lane % 16] over the 16
// points whose LD-strided bf16 rows are `in`.
template <int KS, int LD>
__device__ __forceinline__ void tile_mm(const Frags<KS>& w, const __bf16* in, int lane,
                                        f32x4 (&acc)[4]) {
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  // all B fragments are read before the products: with one wave per SIMD nothing
  // else hides an LDS round trip per k-step
  bf16x8 af[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) af[ks] = *(const bf16x8*)&in[lr * LD + 32 * ks + lk];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w.f[ks][j], af[ks], acc[j], 0, 0, 0);
}

// 16-row layer [16][256] at blob offset base: k-steps 2 wave and 2 wave + 1; the
// wave's partial sums go to red[wave][row][point].
__device__ __forceinline__ void tile_mm16(const bf16x8 (&wf)[2], const __bf16* in, float* red,
                                          int wave, int lane) {
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  f32x4 o = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const bf16x8 af = *(const bf16x8*)&in[lr * TLD + 32 * (2 * wave + s) + lk];
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[s], af, o, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) red[(wave * 16 + 4 * (lane >> 4) + r) * 16 + lr] = o[r];
}

__device__ __forceinline__ void load_frags16(bf16x8 (&wf)[2], __amdgpu_buffer_rsrc_t r,
                                             int64_t base, int wave, int lane) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
    wf[s] = blob_load(r, lane * 16, (int)((base + (int64_t)(2 * wave + s) * 512) * 2));
}

__device__ __forceinline__ void gelu_and_grad(float x, float& h, float& d) {
  const float v = GELU_2K * x * (1.0f + GELU_C * x * x);
  const float sg = __builtin_amdgcn_rcpf(1.0f + __expf(-v));
  h = x * sg;
  d = sg + x * sg * (1.0f - sg) * GELU_2K * (1.0f + 3.0f * GELU_C * x * x);
}

// Forward epilogue: GELU(acc + b) -> bf16 rows of `out`.
__device__ __forceinline__ void epi_fwd(const f32x4 (&acc)[4], const float* __restrict__ bias,
                                        __bf16* out, int wave, int lane) {
  const int lr = lane & 15;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n0 = 64 * wave + 16 * j + 4 * (lane >> 4);
    bf16x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = (__bf16)gelu(acc[j][r] + bias[n0 + r]);
    *(bf16x4*)&out[lr * TLD + n0] = o;
  }
}

// Recompute epilogue: as epi_fwd; GELU'(z) goes to the lane's slots of gd (LDS).
__device__ __forceinline__ void epi_rec(const f32x4 (&acc)[4], const float* __restrict__ bias,
                                        __bf16* out, f32x4* gd, int wave, int lane) {
  const int lr = lane & 15;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n0 = 64 * wave + 16 * j + 4 * (lane >> 4);
    bf16x4 o;
    f32x4 d;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float h, g;
      gelu_and_grad(acc[j][r] + bias[n0 + r], h, g);
      d[r] = g;
      o[r] = (__bf16)h;
    }
    gd[(4 * wave + j) * 64 + lane] = d;
    *(bf16x4*)&out[lr * TLD + n0] = o;
  }
}

// Backward epilogue: dz = acc GELU'(z) -> bf16 rows of `out`.
__device__ __forceinline__ void epi_bwd(const f32x4 (&acc)[4], const f32x4* gd, __bf16* out,
                                        int wave, int lane) {
  const int lr = lane & 15;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n0 = 64 * wave + 16 * j + 4 * (lane >> 4);
    const f32x4 d = gd[(4 * wave + j) * 64 + lane];
    bf16x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = (__bf16)(acc[j][r] * d[r]);
    *(bf16x4*)&out[lr * TLD + n0] = o;
  }
}

// Point-step arrays of the weight gradients are "tiled transposed": element (feature
// n, point-step m) of an array of NF features sits at ((m / 16) NF + n) 16 + m % 16,
// so the 16 point-steps of one tile and step form one contiguous block of NF x 16
// (written with coalesced stores) and 8 consecutive point-steps of one feature are one
// 16-byte load (the MFMA operand of the weight-gradient GEMM).
__device__ __forceinline__ int64_t tt_index(int64_t m, int n, int NF) {
  return ((m >> 4) * NF + n) * 16 + (m & 15);
}

// LDS tile [point][feature] of the 16 points -> the step's block of dst (256 features):
// thread t gathers feature pair 2 (t % 128), 2 (t % 128) + 1 for points 8 (t / 128) ..
// + 7 with eight 4-byte LDS reads and writes two 16-byte chunks.
__device__ __forceinline__ void copy_out_t(const __bf16* buf, __bf16* __restrict__ dst,
                                           int64_t mrow, int tid) {
  const int np = tid & 127, h = tid >> 7;
  const uint32_t* b32 = (const uint32_t*)buf;
  uint32_t w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = b32[(8 * h + k) * (TLD / 2) + np];
  uint32_t lo[4], hi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    lo[i] = (w[2 * i] & 0xffffu) | (w[2 * i + 1] << 16);
    hi[i] = (w[2 * i] >> 16) | (w[2 * i + 1] & 0xffff0000u);
  }
  __bf16* blk = dst + (mrow >> 4) * (256 * 16);
  *(uint4*)&blk[(2 * np) * 16 + 8 * h] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
  *(uint4*)&blk[(2 * np + 1) * 16 + 8 * h] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
}

// Workgroup barrier that orders LDS only: __syncthreads() also fences global memory,
// which makes every barrier wait for all of the wave's outstanding global loads (the
// weight fragments prefetched for the next layer).  No thread of these kernels reads
// global data another thread of the same launch writes.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// acc[j][r] as tile_mm, with the A fragments of an 8-k-step layer read from LDS in
// fragment order (lds: [16-row block][k-step][lane][8]): conflict-free 16-byte reads.
__device__ __forceinline__ void tile_mm_lds(const __bf16* wl, const __bf16* in, int wave,
                                            int lane, f32x4 (&acc)[4]) {
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  bf16x8 af[8];
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) af[ks] = *(const bf16x8*)&in[lr * TLD + 32 * ks + lk];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  // A fragments two k-steps ahead of their products
  bf16x8 wf[3][4];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) wf[p][j] = *(const bf16x8*)&wl[(((4 * wave + j) * 8 + p) * 64 + lane) * 8];
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    if (ks + 2 < 8) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        wf[(ks + 2) % 3][j] = *(const bf16x8*)&wl[(((4 * wave + j) * 8 + ks + 2) * 64 + lane) * 8];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[ks % 3][j], af[ks], acc[j], 0, 0, 0);
  }
}

constexpr size_t kFwdLds = (size_t)65536 * 2 + 2 * 16 * TLD * 2 + 4 * 16 * 16 * 4 + (4 * FH + 16) * 4;

// Forward through the window: one workgroup (4 waves, one per SIMD, each owning 64
// rows of every layer) per 16 batch rows.  Row b's state (latent, inelastic strain,
// loss) lives in thread b of wave 0, which loads its next inputs a step ahead.
// Weights: W1 and W5 fragments stay in registers, W2 in registers too, W3 in LDS for
// the whole window (fragment order), W4 streams from the L2-resident blob into a
// register set refilled right after its product, a step ahead of its next use.
// Biases live in LDS, so no epilogue waits on a global load.
__global__ void __launch_bounds__(kTrainThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
    k_bptt_fwd(BpttArgs a) {
  constexpr int TP = 16;
  extern __shared__ __attribute__((aligned(16))) __bf16 fwd_sm[];
  __bf16* w3l = fwd_sm;                                  // W3 [65536], fragment order
  __bf16 (*act)[TP * TLD] = (__bf16 (*)[TP * TLD])(fwd_sm + 65536);
  float* red = (float*)(fwd_sm + 65536 + 2 * TP * TLD);  // [4][16][16]
  float* bias = red + 4 * 16 * TP;                       // [4 FH + 16]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar blob offsets
  const int b0 = blockIdx.x * TP;
  for (int i = tid; i < 4 * FH + FOUT; i += kTrainThreads) bias[i] = a.bias[i];
  for (int i = tid; i < 65536 / 8; i += kTrainThreads)
    *(bf16x8*)&w3l[8 * i] = *(const bf16x8*)&a.blob[oF2 + 8 * i];
  float z[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, ein[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float lacc = 0.f, wgt = 0.f;
  int path = 0;
  bool valid = tid < TP && b0 + tid < a.B;
  if (valid) {
    path = a.idx[b0 + tid];
    valid = path >= 0 && path < a.P;  // rejected by the host wrapper; never read out of range
    if (valid) wgt = a.pw[path];
  }
  const int64_t prow = (int64_t)path * a.T + a.s0;  // the window's first step of the path
  if (valid && a.state_in) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      z[c] = a.state_in[(int64_t)(b0 + tid) * 12 + c];
      ein[c] = a.state_in[(int64_t)(b0 + tid) * 12 + 6 + c];
    }
  }
  const __amdgpu_buffer_rsrc_t rs = blob_rsrc(a.blob);
  Frags<1> w0;
  load_frags<1>(w0, rs, oF0, 32, wave, lane);
  bf16x8 w4[2];
  load_frags16(w4, rs, oF4, wave, lane);
  Frags<8> W2r, W4s;
  load_frags<8>(W2r, rs, oF1, 256, wave, lane);
  load_frags<8>(W4s, rs, oF3, 256, wave, lane);
  pin(w0);
  pin(w4);
  pin(W2r);
  float f_nx[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (valid) {
#pragma unroll
    for (int c = 0; c < 7; ++c) f_nx[c] = a.feat[prow * 8 + c];
  }
  for (int t = 0; t < a.L; ++t) {
    const int64_t mrow = (int64_t)t * a.Bp + b0;
    float e_t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, g_t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (tid < TP) {
      float x[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) x[c] = 0.f;
      if (valid) {
#pragma unroll
        for (int c = 0; c < 7; ++c) x[c] = f_nx[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) x[7 + c] = tanhf(z[c]);  // the latent through tanh
        // this step's strain and target, the next step's features: consumed after
        // the layers
        const float* e = a.eps + (prow + t) * 6;
        const float* g = a.target + (prow + t) * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          e_t[c] = e[c];
          g_t[c] = g[c];
        }
        if (t + 1 < a.L) {
#pragma unroll
          for (int c = 0; c < 7; ++c) f_nx[c] = a.feat[(prow + t + 1) * 8 + c];
        }
      }
      // tanh(z) for the backward pass, zeros on rows past the batch (the backward
      // kernel reads every row of the tile)
#pragma unroll
      for (int c = 0; c < 6; ++c) a.q[(mrow + tid) * 12 + 6 + c] = x[7 + c];
      __bf16* row = act[0] + tid * TLD;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const __bf16 h = (__bf16)x[c];
        row[c] = h;
        a.xT[tt_index(mrow + tid, c, 16)] = h;
      }
#pragma unroll
      for (int c = 16; c < 32; ++c) row[c] = (__bf16)0.0f;
    }
    lds_barrier();
    f32x4 acc[4];
    tile_mm<1, TLD>(w0, act[0], lane, acc);
    epi_fwd(acc, bias, act[1], wave, lane);
    lds_barrier();
    tile_mm<8, TLD>(W2r, act[1], lane, acc);
    epi_fwd(acc, bias + FH, act[0], wave, lane);
    lds_barrier();
    tile_mm_lds(w3l, act[0], wave, lane, acc);
    epi_fwd(acc, bias + 2 * FH, act[1], wave, lane);
    lds_barrier();
    tile_mm<8, TLD>(W4s, act[1], lane, acc);
    __builtin_amdgcn_sched_barrier(0);  // the refill reuses the registers just read
    if (t + 1 < a.L) load_frags<8>(W4s, rs, oF3, 256, wave, lane);
    epi_fwd(acc, bias + 3 * FH, act[0], wave, lane);
    lds_barrier();
    tile_mm16(w4, act[0], red, wave, lane);
    lds_barrier();
    if (tid < TP) {
      float o[12];
#pragma unroll
      for (int n = 0; n < 12; ++n) {
        float sum = bias[4 * FH + n];
#pragma unroll
        for (int w = 0; w < 4; ++w) sum += red[(w * 16 + n) * TP + tid];
        o[n] = sum;
      }
      float qv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (valid) {
        float err[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) ein[c] += a.cstep * o[c];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          float sum = 0.f;
#pragma unroll
          for (int c = 0; c < 6; ++c) sum += a.Cs[6 * r + c] * (e_t[c] - ein[c]);
          err[r] = sum - g_t[r];
          lacc += wgt * err[r] * err[r];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) z[c] += o[6 + c];
        const float sc = 2.0f * wgt * a.inv_count;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          float sum = 0.f;
#pragma unroll
          for (int r = 0; r < 6; ++r) sum += a.Cs[6 * r + c] * err[r];
          qv[c] = sc * sum;
        }
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) a.q[(mrow + tid) * 12 + c] = qv[c];
    }
  }
  if (tid < TP && b0 + tid < a.B && a.state_out) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      a.state_out[(int64_t)(b0 + tid) * 12 + c] = z[c];
      a.state_out[(int64_t)(b0 + tid) * 12 + 6 + c] = ein[c];
    }
  }
  // tile loss partial: threads 0 .. TP - 1 are lanes of wave 0
  if (wave == 0) {
    float v = lane < TP ? lacc : 0.f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) a.loss_part[blockIdx.x] = v * a.inv_count;
  }
}

// Backward through the window, same tiles, reverse time.  The big layers' fragments
// rotate through P and Q, each refilled one layer ahead (W2 in Q on entry and exit);
// GELU' of the recomputed layers is kept in LDS; the next step's q_t and x_t are
// loaded a step ahead.
__global__ void __launch_bounds__(kTrainThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
    k_bptt_bwd(BpttArgs a) {
  constexpr int TP = 16;
  __shared__ __attribute__((aligned(16))) __bf16 act[2][TP * TLD];
  __shared__ __attribute__((aligned(16))) __bf16 dos[TP * DLD];
  __shared__ float red[4 * 16 * TP];
  __shared__ f32x4 gds[4 * 16 * 64];  // [layer][wave][j][lane]
  __shared__ float bias[4 * FH];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar blob offsets
  const int b0 = blockIdx.x * TP;
  for (int i = tid; i < 4 * FH; i += kTrainThreads) bias[i] = a.bias[i];
  const int64_t M = a.M;
  float G[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gz[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const __amdgpu_buffer_rsrc_t rs = blob_rsrc(a.blob);
  Frags<1> w0, t4;
  load_frags<1>(w0, rs, oF0, 32, wave, lane);
  load_frags<1>(t4, rs, oT4, 32, wave, lane);
  bf16x8 t0[2];
  load_frags16(t0, rs, oT0, wave, lane);
  pin(w0);
  pin(t4);
  pin(t0);
  Frags<8> P, Q;
  load_frags<8>(Q, rs, oF1, 256, wave, lane);
  // x_t element of this thread: feature xc = tid / 16 (16 features), point xp
  const int xc = tid >> 4, xp = tid & 15;
  int64_t mrow = (int64_t)(a.L - 1) * a.Bp + b0;
  float q_nx[12] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (tid < TP) {
#pragma unroll
    for (int c = 0; c < 12; ++c) q_nx[c] = a.q[(mrow + tid) * 12 + c];
  }
  __bf16 x_nx = a.xT[tt_index(mrow + xp, xc, 16)];
  float th[6];  // tanh(z_t), the latent input of step t
  for (int t = a.L - 1; t >= 0; --t) {
    mrow = (int64_t)t * a.Bp + b0;
    if (tid < TP) {
      float d[16];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        th[c] = q_nx[6 + c];
        G[c] -= q_nx[c];
        d[c] = a.cstep * G[c];
        d[6 + c] = gz[c];
      }
#pragma unroll
      for (int c = 12; c < 16; ++c) d[c] = 0.f;
      __bf16* row = dos + tid * DLD;
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const __bf16 h = (__bf16)d[c];
        row[c] = h;
        a.doT[tt_index(mrow + tid, c, 16)] = h;
      }
#pragma unroll
      for (int c = 16; c < 32; ++c) row[c] = (__bf16)0.0f;
      if (t > 0) {
#pragma unroll
        for (int c = 0; c < 12; ++c) q_nx[c] = a.q[(mrow - a.Bp + tid) * 12 + c];
      }
    }
    act[0][xp * TLD + xc] = x_nx;
    act[0][xp * TLD + 16 + xc] = (__bf16)0.0f;
    if (t > 0) x_nx = a.xT[tt_index(mrow - a.Bp + xp, xc, 16)];
    lds_barrier();
    f32x4 acc[4];
    // recompute h1..h4; each tile also goes to HBM (tiled transposed) once the
    // barrier after its epilogue has passed (the next layer only reads it)
    load_frags<8>(P, rs, oF2, 256, wave, lane);
    tile_mm<1, TLD>(w0, act[0], lane, acc);
    epi_rec(acc, bias, act[1], gds, wave, lane);
    lds_barrier();
    copy_out_t(act[1], a.aT, mrow, tid);
    tile_mm<8, TLD>(Q, act[1], lane, acc);
    __builtin_amdgcn_sched_barrier(0);  // the refill reuses the registers just read
    load_frags<8>(Q, rs, oF3, 256, wave, lane);
    epi_rec(acc, bias + FH, act[0], gds + 1024, wave, lane);
    lds_barrier();
    copy_out_t(act[0], a.aT + 256 * M, mrow, tid);
    tile_mm<8, TLD>(P, act[0], lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    load_frags<8>(P, rs, oT3, 256, wave, lane);
    epi_rec(acc, bias + 2 * FH, act[1], gds + 2048, wave, lane);
    lds_barrier();
    copy_out_t(act[1], a.aT + 512 * M, mrow, tid);
    tile_mm<8, TLD>(Q, act[1], lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    load_frags<8>(Q, rs, oT2, 256, wave, lane);
    epi_rec(acc, bias + 3 * FH, act[0], gds + 3072, wave, lane);
    lds_barrier();
    copy_out_t(act[0], a.aT + 768 * M, mrow, tid);
    // back-propagate: dL/dh4 = W5^T do, dz_k = dL/dh_k GELU'(z_k), dL/dh_{k-1} = W_k^T dz_k
    tile_mm<1, DLD>(t4, dos, lane, acc);
    epi_bwd(acc, gds + 3072, act[1], wave, lane);
    lds_barrier();
    copy_out_t(act[1], a.dzT + 768 * M, mrow, tid);
    tile_mm<8, TLD>(P, act[1], lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    load_frags<8>(P, rs, oT1, 256, wave, lane);
    epi_bwd(acc, gds + 2048, act[0], wave, lane);
    lds_barrier();
    copy_out_t(act[0], a.dzT + 512 * M, mrow, tid);
    tile_mm<8, TLD>(Q, act[0], lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    load_frags<8>(Q, rs, oF1, 256, wave, lane);  // the next step's first big layer
    epi_bwd(acc, gds + 1024, act[1], wave, lane);
    lds_barrier();
    copy_out_t(act[1], a.dzT + 256 * M, mrow, tid);
    tile_mm<8, TLD>(P, act[1], lane, acc);
    epi_bwd(acc, gds, act[0], wave, lane);
    lds_barrier();
    copy_out_t(act[0], a.dzT, mrow, tid);
    // dL/dx = W1^T dz1; its latent part feeds gz
    tile_mm16(t0, act[0], red, wave, lane);
    lds_barrier();
    if (tid < TP) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        float sum = 0.f;
#pragma unroll
        for (int w = 0; w < 4; ++w) sum += red[(w * 16 + 7 + c) * TP + tid];
        gz[c] += sum * (1.0f - th[c] * th[c]);  // through the tanh of the latent input
      }
    }
  }
}

// part[s][n][k] = sum over point-steps m of split s of A(n, m) Bm(k, m) (k < K) and
// part[s][n][K] = sum of A(n, m) (the bias gradient); A (N features) and Bm (K
// features) bf16 point-step arrays in the tiled transposed layout of tt_index (M a
// multiple of 16).  4 waves in AN x AK, each a (16 WN) x (16 WK) tile.
template <int WN, int WK, int AN, int AK>
__global__ void __launch_bounds__(256) k_wgrad_nt(int64_t M, int N, int K,
                                                  const __bf16* __restrict__ A,
                                                  const __bf16* __restrict__ Bm, int64_t chunk,
                                                  float* __restrict__ part) {
  static_assert(AN * AK == 4, "four waves per workgroup");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = (lane >> 4) * 8;
  const int tilesN = (N + 16 * AN * WN - 1) / (16 * AN * WN);
  const int tn = blockIdx.x % tilesN, tk = blockIdx.x / tilesN;
  const int n0 = (AN * tn + wave / AK) * 16 * WN, k0 = (AK * tk + wave % AK) * 16 * WK;
  if (n0 >= N || k0 >= K) return;
  const int64_t mb = (int64_t)blockIdx.y * chunk;
  const int64_t me = mb + chunk < M ? mb + chunk : M;
  f32x4 acc[WN][WK];
#pragma unroll
  for (int i = 0; i < WN; ++i)
#pragma unroll
    for (int j = 0; j < WK; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  float rs[WN];
#pragma unroll
  for (int i = 0; i < WN; ++i) rs[i] = 0.f;
  const bool with_bias = k0 == 0;
  bf16x8 zero8;
#pragma unroll
  for (int e = 0; e < 8; ++e) zero8[e] = (__bf16)0.0f;
  for (int64_t m0 = mb; m0 < me; m0 += 32) {
    const bool ok = m0 + lk + 8 <= me;
    bf16x8 af[WN], bfr[WK];
#pragma unroll
    for (int i = 0; i < WN; ++i)
      af[i] = ok ? *(const bf16x8*)&A[tt_index(m0 + lk, n0 + 16 * i + lr, N)] : zero8;
#pragma unroll
    for (int j = 0; j < WK; ++j)
      bfr[j] = ok ? *(const bf16x8*)&Bm[tt_index(m0 + lk, k0 + 16 * j + lr, K)] : zero8;
#pragma unroll
    for (int i = 0; i < WN; ++i)
#pragma unroll
      for (int j = 0; j < WK; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (with_bias) {
#pragma unroll
      for (int i = 0; i < WN; ++i)
#pragma unroll
        for (int e = 0; e < 8; ++e) rs[i] += (float)af[i][e];
    }
  }
  const int Kx = K + 1;
  float* P = part + (int64_t)blockIdx.y * N * Kx;
#pragma unroll
  for (int i = 0; i < WN; ++i)
#pragma unroll
    for (int j = 0; j < WK; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        P[(int64_t)(n0 + 16 * i + 4 * (lane >> 4) + r) * Kx + k0 + 16 * j + lr] = acc[i][j][r];
  if (with_bias) {
#pragma unroll
    for (int i = 0; i < WN; ++i) {
      float v = rs[i];
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      if (lane < 16) P[(int64_t)(n0 + 16 * i + lane) * Kx + K] = v;
    }
  }
}

struct WgradLayer {
  const float* part;  // [S][N][K + 1]
  int S, N, K;        // padded partial dimensions
  int Nr, Kr;         // real weight shape
  int64_t w_off, b_off;
};
struct WgradLayers {
  WgradLayer l[5];
};

// Gradient blobs from the split partials (fixed summation order); the loss from the
// tile partials.  A workgroup reduces 16 consecutive outputs: thread (r, o) sums splits
// r, r + 16, ... of output o (16 outputs read as one 64-byte segment per split), and
// the 16 partial sums of each output are added in order through LDS.
__global__ void __launch_bounds__(256) k_bptt_reduce(WgradLayers L, float* __restrict__ gw,
                                                     float* __restrict__ gb,
                                                     const float* __restrict__ loss_part,
                                                     int tiles, float* __restrict__ loss,
                                                     int64_t total, float* __restrict__ sq_part) {
  __shared__ float red[16][17];
  const int ol = threadIdx.x & 15, r = threadIdx.x >> 4;
  int64_t e = (int64_t)blockIdx.x * 16 + ol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < tiles; ++i) s += loss_part[i];
    *loss = s;
  }
  int li = 0;
  float acc = 0.f;
  int n = 0, k = 0;
  const bool valid = e < total;
  if (valid) {
    while (li < 4 && e >= (int64_t)L.l[li].Nr * (L.l[li].Kr + 1)) {
      e -= (int64_t)L.l[li].Nr * (L.l[li].Kr + 1);
      ++li;
    }
    const WgradLayer& w = L.l[li];
    n = (int)(e / (w.Kr + 1));
    k = (int)(e % (w.Kr + 1));
    const int kk = k < w.Kr ? k : w.K;
    const int64_t per = (int64_t)w.N * (w.K + 1);
    const float* src = w.part + (int64_t)n * (w.K + 1) + kk;
    for (int sp = r; sp < w.S; sp += 16) acc += src[sp * per];
  }
  red[r][ol] = acc;
  __syncthreads();
  float sq = 0.f;
  if (r == 0 && valid) {
    float sum = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) sum += red[i][ol];
    const WgradLayer& w = L.l[li];
    if (k < w.Kr)
      gw[w.w_off + (int64_t)n * w.Kr + k] = sum;
    else
      gb[w.b_off + n] = sum;
    sq = sum * sum;
  }
  // the block's sum of squares (the 16 outputs sit in lanes 0 .. 15 of wave 0)
  if (threadIdx.x < 64) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
    if (threadIdx.x == 0) sq_part[blockIdx.x] = sq;
  }
}

// Global gradient norm from the block sums of squares (fixed order) and the clip scale:
// min(1, max_norm / norm) when max_norm > 0, 0 when the norm is not finite (the step
// is skipped), 1 otherwise.  norm_scale[0] = norm, norm_scale[1] = scale.
__global__ void __launch_bounds__(256) k_grad_norm(const float* __restrict__ sq_part, int n,
                                                   float max_norm, float* __restrict__ norm_scale,
                                                   float* __restrict__ norm_out) {
  __shared__ float red[256];
  float v = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) v += sq_part[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float norm = sqrtf(red[0]);
    float scale = 1.f;
    if (!isfinite(norm))
      scale = 0.f;
    else if (max_norm > 0.f && norm > max_norm)
      scale = max_norm / norm;
    norm_scale[0] = norm;
    norm_scale[1] = scale;
    if (norm_out) *norm_out = norm;
  }
}

__global__ void k_grad_scale(int64_t nw, float* __restrict__ gw, int64_t nb,
                             float* __restrict__ gb, const float* __restrict__ norm_scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float s = norm_scale[1];
  if (s == 1.f) return;
  // a zero scale (non-finite norm) writes zeros rather than 0 * inf
  if (i < nw) gw[i] = s == 0.f ? 0.f : gw[i] * s;
  else if (i < nw + nb) gb[i - nw] = s == 0.f ? 0.f : gb[i - nw] * s;
}

}  // namespace cia

extern "C" {

using cia::GemmArgs;

static cia_status gemm_launch(hipStream_t st, int epi, GemmArgs g, int splits) {
  using namespace cia;
  const unsigned gx = (unsigned)((g.M + GM - 1) / GM), gy = (unsigned)((g.N + GN - 1) / GN);
  dim3 grid(gx, gy, (unsigned)splits);
  if (epi == EPI_GELU)
    hipLaunchKernelGGL(k_gemm_bf16<EPI_GELU>, grid, dim3(256), 0, st, g);
  else if (epi == EPI_DGELU)
    hipLaunchKernelGGL(k_gemm_bf16<EPI_DGELU>, grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(k_gemm_bf16<EPI_LINEAR>, grid, dim3(256), 0, st, g);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_mlp_linear_forward(void* stream, int64_t M, int64_t N, int64_t K, const void* X,
                                  const void* W, const float* bias, int32_t activation,
                                  float* Z, float* Y, void* Y_bf16) {
  using namespace cia;
  CIA_ARG(X && W && M >= 0 && N > 0 && K > 0, "invalid arguments to cia_mlp_linear_forward");
  CIA_ARG(Y || Y_bf16, "cia_mlp_linear_forward needs an output");
  if (M == 0) return CIA_OK;
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = (const __bf16*)X; g.sAm = K; g.sAk = 1;
  g.B = (const __bf16*)W; g.sBk = 1; g.sBn = K;  // B(k, n) = W[n][k]
  g.ones_col = -1;
  g.bias = bias;
  g.Zf = activation ? Z : nullptr; g.ldzf = N;
  g.Cf = Y; g.ldcf = N;
  g.Cb = (__bf16*)Y_bf16; g.ldcb = N;
  g.k_split = ((K + GK - 1) / GK) * GK;
  return gemm_launch((hipStream_t)stream, activation ? EPI_GELU : EPI_LINEAR, g, 1);
}

cia_status cia_mlp_linear_backward_data(void* stream, int64_t M, int64_t N, int64_t K,
                                        const void* dZ, const void* W, const float* Zprev,
                                        float* dX, void* dX_bf16) {
  using namespace cia;
  CIA_ARG(dZ && W && M >= 0 && N > 0 && K > 0, "invalid arguments to cia_mlp_linear_backward_data");
  CIA_ARG(dX || dX_bf16, "cia_mlp_linear_backward_data needs an output");
  if (M == 0) return CIA_OK;
  GemmArgs g{};
  // dX[M][K] = dZ[M][N] W[N][K]: GEMM with reduction over N
  g.M = M; g.N = K; g.K = N;
  g.A = (const __bf16*)dZ; g.sAm = N; g.sAk = 1;
  g.B = (const __bf16*)W; g.sBk = K; g.sBn = 1;  // B(n', k') = W[n'][k']
  g.ones_col = -1;
  g.zin = Zprev; g.ldzin = K;
  g.Cf = dX; g.ldcf = K;
  g.Cb = (__bf16*)dX_bf16; g.ldcb = K;
  g.k_split = ((N + GK - 1) / GK) * GK;
  return gemm_launch((hipStream_t)stream, Zprev ? EPI_DGELU : EPI_LINEAR, g, 1);
}

// dW[N][K] (and db as an extra column of ones in X) = dZ^T X: a GEMM whose
// reduction runs over the M rows, split along M into fp32 partials held in
// `workspace` (every partial the reduction reads is written by the GEMM first,
// so the workspace needs no clearing).
static cia_status wgrad_launch(hipStream_t st, int64_t M, int64_t N, int64_t K, const void* dZ,
                               const void* X, float* dW, float* db, float* workspace,
                               int64_t workspace_floats, int32_t accumulate) {
  using namespace cia;
  const int64_t Kx = K + (db ? 1 : 0);
  const int64_t per = N * Kx;
  const int64_t tiles = ((N + GM - 1) / GM) * ((Kx + GN - 1) / GN);
  int64_t splits = std::min<int64_t>(kMaxWgradSplits, std::max<int64_t>(1, 512 / tiles));
  splits = std::min<int64_t>(splits, (M + GK - 1) / GK);
  splits = std::min<int64_t>(splits, workspace_floats / per - 1);
  CIA_ARG(splits >= 1, "workspace too small for cia_mlp_linear_backward_weight");
  int64_t ks = (M + splits - 1) / splits;
  ks = ((ks + GK - 1) / GK) * GK;
  splits = (M + ks - 1) / ks;
  GemmArgs g{};
  g.M = N; g.N = Kx; g.K = M;
  g.A = (const __bf16*)dZ; g.sAm = 1; g.sAk = N;  // A(n, m) = dZ[m][n]
  g.B = (const __bf16*)X; g.sBk = K; g.sBn = 1;   // B(m, k) = X[m][k]
  g.ones_col = db ? K : -1;
  g.Cf = workspace; g.ldcf = Kx;
  g.k_split = ks;
  g.split_stride = per;
  CIA_TRY(gemm_launch(st, EPI_LINEAR, g, (int)splits));
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid_for(per)), dim3(kBlock), 0, st, (int)splits, N, K,
                     Kx, (const float*)workspace, dW, db, accumulate);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_mlp_linear_backward_weight(void* stream, int64_t M, int64_t N, int64_t K,
                                          const void* dZ, const void* X, float* dW, float* db,
                                          float* workspace, int64_t workspace_floats,
                                          int32_t accumulate) {
  CIA_ARG(dZ && X && dW && M > 0 && N > 0 && K > 0 && workspace,
          "invalid arguments to cia_mlp_linear_backward_weight");
  return wgrad_launch((hipStream_t)stream, M, N, K, dZ, X, dW, db, workspace, workspace_floats,
                      accumulate);
}

// Allocation-free variant: the split partials live in the file-scope device
// array `scratch` (one copy per GPU, held until the process ends).  Host code
// never passes `scratch` itself to a kernel: its address on each device is
// taken once with hipGetSymbolAddress, with that device current, and kept.
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[cia::kNoAllocMaxDevices] = {};

static cia_status scratch_on_current_device(hipStream_t st, float** out) {
  int dev = -1;
  CIA_HIP(hipGetDevice(&dev));
  CIA_ARG(dev >= 0 && dev < cia::kNoAllocMaxDevices,
          "cia_mlp_linear_backward_weight_noalloc: device index beyond the supported range");
  if (st) {
    int sdev = -1;
    CIA_HIP(hipStreamGetDevice(st, &sdev));
    CIA_ARG(sdev == dev,
            "cia_mlp_linear_backward_weight_noalloc: the stream's device must be the current device");
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    CIA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(cia::scratch)));
    CIA_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (float*)p;
  }
  *out = g_scratch_ptr[dev];
  return CIA_OK;
}

cia_status cia_mlp_linear_backward_weight_noalloc(void* stream, int64_t M, int64_t N, int64_t K,
                                                  const void* dZ, const void* X, float* dW,
                                                  float* db, int32_t accumulate) {
  using namespace cia;
  CIA_ARG(dZ && X && dW && M > 0 && N > 0 && K > 0,
          "invalid arguments to cia_mlp_linear_backward_weight_noalloc");
  // at least one split partial plus the reserve the split rule keeps
  CIA_ARG(N * (K + 1) <= kNoAllocScratchFloats / 2,
          "cia_mlp_linear_backward_weight_noalloc: N * (K + 1) exceeds the static scratch "
          "(N * (K + 1) <= 2138240 floats, e.g. N = K = 1024 is too large; use "
          "cia_mlp_linear_backward_weight with a workspace)");
  hipStream_t st = (hipStream_t)stream;
  float* ws = nullptr;
  CIA_TRY(scratch_on_current_device(st, &ws));
  return wgrad_launch(st, M, N, K, dZ, X, dW, db, ws, kNoAllocScratchFloats, accumulate);
}

int64_t cia_mlp_backward_weight_workspace(int64_t N, int64_t K) {
  return (cia::kMaxWgradSplits + 1) * N * (K + 1);
}

cia_status cia_mlp_adam_step(void* stream, int64_t n, float* params, const float* grad, float* m1,
                             float* m2, void* params_bf16, float lr, float beta1, float beta2,
                             float eps, int32_t step, float weight_decay) {
  using namespace cia;
  CIA_ARG(params && grad && m1 && m2 && n > 0 && step >= 1, "invalid arguments to cia_mlp_adam_step");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, n, params,
                     grad, m1, m2, (__bf16*)params_bf16, lr, beta1, beta2, eps, c1, c2,
                     weight_decay);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_f32_to_bf16(void* stream, int64_t n, const float* x, void* y) {
  using namespace cia;
  CIA_ARG(x && y && n >= 0, "invalid arguments to cia_f32_to_bf16");
  if (n == 0) return CIA_OK;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, n, x,
                     (__bf16*)y);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_mlp_forward_fused(void* stream, int64_t M, const float* X, const void* weights,
                                 const float* biases, float* Y) {
  using namespace cia;
  CIA_ARG(X && weights && biases && Y && M >= 0, "invalid arguments to cia_mlp_forward_fused");
  if (M == 0) return CIA_OK;
  const size_t lds = FUSED_ACT_BYTES;
  static bool attr_set = false;
  if (!attr_set) {
    CIA_HIP(hipFuncSetAttribute((const void*)k_mlp_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_mlp_fused, dim3((unsigned)((M + FM - 1) / FM)), dim3(256), lds,
                     (hipStream_t)stream, M, X, (const __bf16*)weights, biases, Y);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_surrogate_update(void* stream, int64_t n, const double* strain, double time_step,
                                const void* weights, const float* biases,
                                const cia_surrogate_gate* gate, const double* anchor_strain,
                                const double* anchor_inelastic, const float* anchor_latent,
                                const float* anchor_age, const uint8_t* fallback, double* stress,
                                double* next_anchor_strain, double* next_anchor_inelastic,
                                float* next_latent, float* next_age, uint8_t* next_fallback,
                                uint32_t* trip_counter, const double* inelastic,
                                double* next_inelastic, double* free_energy,
                                double* dissipation) {
  using namespace cia;
  CIA_ARG(strain && weights && biases && gate && anchor_strain && anchor_inelastic &&
              anchor_latent && anchor_age && fallback && stress && next_anchor_strain &&
              next_anchor_inelastic && next_latent && next_age && next_fallback && trip_counter &&
              n >= 0,
          "invalid arguments to cia_surrogate_update");
  CIA_ARG(time_step > 0.0 && gate->eps_scale > 0.0 && gate->rate_ref > 0.0 &&
              gate->rate_scale > 0.0 && gate->commit_strain >= 0.0,
          "cia_surrogate_update: time step and scales must be positive");
  CIA_ARG(n < ((int64_t)1 << 31), "cia_surrogate_update: at most 2^31 - 1 points");
  CIA_ARG(!dissipation || inelastic,
          "cia_surrogate_update: the dissipation needs the previous call's inelastic strain");
  if (n == 0) return CIA_OK;
  const size_t lds = FUSED_ACT_BYTES + (size_t)FM * (FOUT + 1) * sizeof(float);
  static bool attr_set = false;
  if (!attr_set) {
    CIA_HIP(hipFuncSetAttribute((const void*)k_sur_update,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  SurGate g;
  g.eps_scale = gate->eps_scale;
  g.rate_ref = gate->rate_ref;
  g.rate_scale = gate->rate_scale;
  g.inelastic_step = gate->inelastic_step;
  g.d_eps_max = gate->d_eps_max;
  g.rate_lo = gate->rate_lo;
  g.rate_hi = gate->rate_hi;
  g.latent_max = gate->latent_max;
  g.gate_margin = gate->gate_margin;
  g.rate_margin = gate->rate_margin;
  g.commit_strain = gate->commit_strain;
  g.commit_time = gate->commit_time;
  for (int i = 0; i < 36; ++i) g.C[i] = gate->stiffness[i];
  hipLaunchKernelGGL(k_sur_update, dim3((unsigned)((n + FM - 1) / FM)), dim3(256), lds,
                     (hipStream_t)stream, n, strain, time_step, (const __bf16*)weights, biases, g,
                     anchor_strain, anchor_inelastic, anchor_latent, anchor_age, fallback, stress,
                     next_anchor_strain, next_anchor_inelastic, next_latent, next_age,
                     next_fallback, (unsigned int*)trip_counter, inelastic, next_inelastic,
                     free_energy, dissipation);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Fused BPTT training step: workspace layout and entry points.
namespace {

// gradient entries of the network: weights and biases
constexpr int64_t kGradTotal = cia::mW5 + 12 * 256 + 4 * cia::FH + cia::FOUT;

struct BpttLayout {
  int tp, Bp, tiles;
  int64_t M, chunk[5], S[5];
  size_t xT, q, aT, dzT, doT, blob, loss, part[5], sq, norm, total;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

BpttLayout bptt_layout(int64_t paths, int64_t window) {
  BpttLayout L{};
  L.tp = 16;
  L.Bp = (int)((paths + L.tp - 1) / L.tp) * L.tp;
  L.tiles = L.Bp / L.tp;
  L.M = window * L.Bp;
  // point-step splits of the weight gradients: 64 for the 256 x 256 layers (4
  // workgroups each), 256 for the thin input and output layers (2 workgroups each)
  const int64_t target[5] = {256, 64, 64, 64, 256};
  for (int l = 0; l < 5; ++l) {
    L.chunk[l] = std::max<int64_t>(256, ((L.M + target[l] - 1) / target[l] + 31) / 32 * 32);
    L.S[l] = (L.M + L.chunk[l] - 1) / L.chunk[l];
  }
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align256(bytes);
    return at;
  };
  L.xT = take((size_t)16 * L.M * 2);
  L.q = take((size_t)L.M * 12 * 4);
  L.aT = take((size_t)4 * 256 * L.M * 2);
  L.dzT = take((size_t)4 * 256 * L.M * 2);
  L.doT = take((size_t)16 * L.M * 2);
  L.blob = take((size_t)cia::kTrainBlob * 2);
  L.loss = take((size_t)L.tiles * 4);
  const int pn[5] = {256, 256, 256, 256, 16}, pk[5] = {16, 256, 256, 256, 256};
  for (int l = 0; l < 5; ++l) L.part[l] = take((size_t)L.S[l] * pn[l] * (pk[l] + 1) * 4);
  L.sq = take((size_t)(kGradTotal + 15) / 16 * 4);
  L.norm = take(2 * 4);
  L.total = o;
  return L;
}

}  // namespace

extern "C" {

int64_t cia_surrogate_bptt_workspace(int64_t paths, int64_t window) {
  if (paths <= 0 || window <= 0) return 0;
  return (int64_t)bptt_layout(paths, window).total;
}

cia_status cia_surrogate_bptt(void* stream, int64_t n_paths, int64_t path_length, int64_t paths,
                              int64_t window_start, int64_t window, const int32_t* path_index,
                              const float* features, const float* strain, const float* target,
                              const float* path_weight, const float* stiffness,
                              float inelastic_step, const float* weights,
                              const float* biases, float* grad_weights, float* grad_biases,
                              float* loss, float* grad_norm, float max_grad_norm,
                              const float* state_in, float* state_out, void* workspace,
                              int64_t workspace_bytes) {
  using namespace cia;
  CIA_ARG(path_index && features && strain && target && path_weight && stiffness && weights &&
              biases && grad_weights && grad_biases && loss && workspace,
          "invalid arguments to cia_surrogate_bptt");
  CIA_ARG(paths > 0 && n_paths > 0 && window > 0 && window_start >= 0 &&
              window_start + window <= path_length && paths < (1 << 30) && n_paths < (1 << 30) &&
              path_length < (1 << 30),
          "cia_surrogate_bptt: need 0 <= window_start, 0 < window, window_start + window <= "
          "path_length, paths > 0 and n_paths > 0");
  CIA_ARG(!state_out || state_out != state_in,
          "cia_surrogate_bptt: state_out must not alias state_in");
  const BpttLayout L = bptt_layout(paths, window);
  CIA_ARG(L.M <= ((int64_t)1 << 22),
          "cia_surrogate_bptt: paths x window (rounded up to 16 paths) must not exceed 4194304");
  CIA_ARG(workspace_bytes >= (int64_t)L.total,
          "cia_surrogate_bptt: workspace smaller than cia_surrogate_bptt_workspace()");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  BpttArgs a{};
  a.B = (int)paths;
  a.Bp = L.Bp;
  a.L = (int)window;
  a.T = (int)path_length;
  a.P = (int)n_paths;
  a.s0 = (int)window_start;
  a.state_in = state_in;
  a.state_out = state_out;
  a.M = L.M;
  a.idx = path_index;
  a.feat = features;
  a.eps = strain;
  a.target = target;
  a.pw = path_weight;
  for (int i = 0; i < 36; ++i) a.Cs[i] = stiffness[i];
  a.cstep = inelastic_step;
  a.inv_count = (float)(1.0 / ((double)paths * (double)window * 6.0));
  a.blob = (const __bf16*)(ws + L.blob);
  a.bias = biases;
  a.xT = (__bf16*)(ws + L.xT);
  a.q = (float*)(ws + L.q);
  a.aT = (__bf16*)(ws + L.aT);
  a.dzT = (__bf16*)(ws + L.dzT);
  a.doT = (__bf16*)(ws + L.doT);
  a.loss_part = (float*)(ws + L.loss);
  hipLaunchKernelGGL(k_bptt_prep, dim3(grid_for(kTrainBlob)), dim3(kBlock), 0, st, weights,
                     (__bf16*)(ws + L.blob));
  static bool fwd_attr = false;
  if (!fwd_attr) {
    CIA_HIP(hipFuncSetAttribute((const void*)k_bptt_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kFwdLds));
    fwd_attr = true;
  }
  hipLaunchKernelGGL(k_bptt_fwd, dim3(L.tiles), dim3(kTrainThreads), kFwdLds, st, a);
  hipLaunchKernelGGL(k_bptt_bwd, dim3(L.tiles), dim3(kTrainThreads), 0, st, a);
  CIA_HIP(hipGetLastError());
  const int64_t M = L.M;
  float* part[5];
  for (int l = 0; l < 5; ++l) part[l] = (float*)(ws + L.part[l]);
  // layer 1: dz1 (256 features) x inputs (16): 4 x 1 waves of 64 x 16
  hipLaunchKernelGGL((k_wgrad_nt<4, 1, 4, 1>), dim3(1, (unsigned)L.S[0]), dim3(256), 0, st, M,
                     256, 16, a.dzT, a.xT, L.chunk[0], part[0]);
  for (int l = 1; l < 4; ++l)
    hipLaunchKernelGGL((k_wgrad_nt<4, 4, 2, 2>), dim3(4, (unsigned)L.S[l]), dim3(256), 0, st, M,
                       256, 256, a.dzT + (int64_t)l * 256 * M,
                       a.aT + (int64_t)(l - 1) * 256 * M, L.chunk[l], part[l]);
  // output layer: dL/do (16) x h4 (256): 1 x 4 waves of 16 x 64
  hipLaunchKernelGGL((k_wgrad_nt<1, 4, 1, 4>), dim3(1, (unsigned)L.S[4]), dim3(256), 0, st, M, 16,
                     256, a.doT, a.aT + (int64_t)3 * 256 * M, L.chunk[4], part[4]);
  WgradLayers wl{};
  const int Ns[5] = {256, 256, 256, 256, 16}, Ks[5] = {16, 256, 256, 256, 256};
  const int Nr[5] = {256, 256, 256, 256, FOUT}, Kr[5] = {FIN, 256, 256, 256, 256};
  const int64_t woff[5] = {mW1, mW2, mW3, mW4, mW5};
  int64_t total = 0;  // == kGradTotal
  for (int l = 0; l < 5; ++l) {
    wl.l[l] = WgradLayer{part[l], (int)L.S[l], Ns[l], Ks[l], Nr[l], Kr[l], woff[l], (int64_t)l * FH};
    total += (int64_t)Nr[l] * (Kr[l] + 1);
  }
  const unsigned nblk = (unsigned)((total + 15) / 16);
  hipLaunchKernelGGL(k_bptt_reduce, dim3(nblk), dim3(256), 0, st, wl, grad_weights, grad_biases,
                     (const float*)(ws + L.loss), L.tiles, loss, total, (float*)(ws + L.sq));
  hipLaunchKernelGGL(k_grad_norm, dim3(1), dim3(256), 0, st, (const float*)(ws + L.sq), (int)nblk,
                     max_grad_norm, (float*)(ws + L.norm), grad_norm);
  if (max_grad_norm > 0.f)
    hipLaunchKernelGGL(k_grad_scale, dim3(grid_for(kGradTotal)), dim3(kBlock), 0, st, mW5 + 12 * 256,
                       grad_weights, (int64_t)(4 * FH + FOUT), grad_biases,
                       (const float*)(ws + L.norm));
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

}  // extern "C"
