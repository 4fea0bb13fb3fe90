// Recurrent MLP material-point surrogate: BF16 MFMA GEMMs with FP32 accumulation
// and fused bias / GELU / GELU' epilogues, split-K weight gradients, Adam, and a
// fused all-layer inference kernel.
//
// The surrogate plays the role of the reference's recurrent material-point
// network (EnergyDissipationRVERNO, src/concrete_impact/nn/models/rve_rno.py:162-476,
// trained by BPTT in src/concrete_impact/nn/training/rve_rno.py:113-415) with the
// architecture 13 -> 256 -> 256 -> 256 -> 256 -> 12 (GELU): input = strain
// increment (6), log strain-rate feature (1), latent state (6); output = stress (6),
// next latent state (6).
//
// Matrix cores: v_mfma_f32_16x16x32_bf16.  Operand layout per 64-lane wave:
//   A (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7
//   B (32 x 16): lane l holds col l % 16, k = 8 (l / 16) .. + 7
//   D (16 x 16): lane l holds col l % 16, rows 4 (l / 16) .. + 3
// (checked against a float64 reference by tests/test_gpu_mlp.py).
#include <math.h>
#include <string.h>

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
//   x = [d / eps_scale, clamp_lo(log10(rate / rate_ref) / rate_scale), z_a],
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
      for (int c = 0; c < 6; ++c) row[7 + c] = (__bf16)a_lat[c * n + m];
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
