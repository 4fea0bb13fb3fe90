// Host/device helpers shared by the libcia_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/cia.h"
#include "point.hpp"

namespace cia {

// Last HIP error message of the calling thread (exposed by cia_last_error).
void set_error(const char* what, hipError_t err, const char* file, int line);
void set_error_msg(const std::string& msg);

#define CIA_HIP(x)                                        \
  do {                                                    \
    hipError_t e_ = (x);                                  \
    if (e_ != hipSuccess) {                               \
      ::cia::set_error(#x, e_, __FILE__, __LINE__);       \
      return CIA_ERR_HIP;                                 \
    }                                                     \
  } while (0)

#define CIA_ARG(cond, msg)                                \
  do {                                                    \
    if (!(cond)) {                                        \
      ::cia::set_error_msg(msg);                          \
      return CIA_ERR_INVALID_ARGUMENT;                    \
    }                                                     \
  } while (0)

inline PhaseConst make_phase(const cia_material& m) {
  PhaseConst p{};
  p.kind = m.kind;
  const double E = m.young_modulus, nu = m.poisson_ratio;
  p.G = E / (2.0 * (1.0 + nu));
  p.K = E / (3.0 * (1.0 - 2.0 * nu));
  p.lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  p.sy = m.yield_stress;
  p.H = m.hardening_modulus;
  p.tau = m.time_scale;
  p.sref = m.reference_stress;
  p.m = m.rate_exponent;
  p.hs = 3.0 * p.G + p.H;
  p.density = m.density;
  return p;
}

inline Settings make_settings(const cia_update_settings& s) {
  Settings o{};
  o.max_iterations = s.max_iterations;
  o.tol_rel_yield = s.yield_relative_tolerance;
  o.res_abs = s.residual_absolute_tolerance;
  o.res_rel = s.residual_relative_tolerance;
  return o;
}

// Per-device scratch kept for the life of the process: phase table and the
// failure reporting slots of the standalone material update.
struct DeviceCtx {
  PhaseTable* tab = nullptr;
  unsigned long long* first_fail = nullptr;
  cia_point_failure* rec = nullptr;
};
cia_status device_ctx(DeviceCtx** out);

#define CIA_TRY(x)               \
  do {                           \
    cia_status s_ = (x);         \
    if (s_ != CIA_OK) return s_; \
  } while (0)

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return (unsigned)((n + block - 1) / block);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD
// b mod 8), and each XCD has its own L2.  Stencil kernels renumber their blocks so
// that each XCD walks one contiguous range of the grid: the neighbour rows and
// planes that adjacent elements share then come from that XCD's own L2.
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// ---- deterministic block reductions --------------------------------------

// DPP move of a double (both halves) with control CTRL, full row and bank masks.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wavefront: DPP moves on the vector ALU inside each
// row of 16 lanes (swap neighbours, swap pairs, mirror the halves of 8, mirror the
// row) leave every lane with its row sum; the four row sums are read from lanes
// 0, 16, 32, 48.  No LDS traffic, a fixed summation order; every lane returns the
// total.  All 64 lanes must be active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum over a block of up to 1024 threads; result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) r += lds[w];
  }
  __syncthreads();
  return r;
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = -INFINITY;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) r = fmax(r, lds[w]);
  }
  __syncthreads();
  return r;
}

template <int BLOCK>
__device__ __forceinline__ double block_min(double v, double* lds) {
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = INFINITY;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) r = fmin(r, lds[w]);
  }
  __syncthreads();
  return r;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v);

}  // namespace cia
