// Spectral reference-medium preconditioner of the single-RVE PCG, in fused
// passes.
//
// The preconditioner is z = P K0^+ (r - (sum r) delta_0): the exact inverse of
// the Hex8 stiffness K0 of a homogeneous isotropic medium on the periodic grid
// (block-circulant, so the 3-D DFT turns it into 3x3 Hermitian blocks per
// mode), applied to the residual with the pinned-class reaction removed; the
// shift P (z - z(class 0)) is applied by the consumer.  Each residual component
// is transformed as its own complex field so that the z-pass needs no mode
// pairing and fuses with the symbol solve:
//
//   xlines_update_fwd  a group of x-lines: the PCG vector update (x += alpha p,
//                      r -= alpha y; or the initialisation) of the lines'
//                      classes, then the x-DFT of each component into S[m]
//   (y pass)           k_dft_strided of dft.hpp over the three components
//   zcol_solve         tile_w consecutive kx at one ky, all z: forward z-DFT,
//                      the 3x3 symbol solve per mode, r.z by Parseval, inverse
//                      z-DFT, in place
//   (inverse y pass)
//   xlines_inv         a group of x-lines: inverse x-DFT, real part = z, written
//                      interleaved per class
//
// Five launches per application (with the update) instead of eight: the single
// RVE's PCG iteration is bound by launch latency, not by bytes.  The 1-D
// transforms are dft_stage (one Cooley-Tukey level, twiddles in LDS); edges up to
// 64 elements.  Grids whose edges are all 16, 32, 48 or 64 (y and z edges for the
// y and z passes) run the row-lane forms of these passes instead (k_sp_xlines_fwd,
// k_sp_ylines, k_sp_zlines, k_sp_xlines_inv: lanefft.hpp transforms in registers,
// no LDS staging).
#pragma once

#include <hip/hip_complex.h>
#include <hip/hip_runtime.h>

#include "dft.hpp"
#include "symbol.hpp"

namespace cia {

constexpr int kZTile = 16;  // largest z-column tile (kx columns)

// Dynamic LDS of the z-column pass.
inline size_t zcol_lds_bytes(int nz, int tile_w = kZTile) {
  return 6 * (size_t)tile_w * (nz + 1) * sizeof(double2);
}
// x-line groups: lines per workgroup and dynamic LDS of the line passes.
__host__ __device__ inline int xline_group(int nx) { return nx >= 256 ? 1 : 256 / nx; }
inline size_t xline_lds_bytes(int nx) { return 6 * (size_t)xline_group(nx) * (nx + 1) * sizeof(double2); }

// z-column tile (kx in [tile_w xt, tile_w (xt + 1)), ky) of an RVE's spectra: forward
// z-DFT, Z = Khat^-1 (R - sum r) / N per mode, inverse z-DFT, in place.
//
// Split in two so that the kernel can issue every global load of the tile before
// it waits for the PCG scalars: zcol_prefetch reads the first kZcolPf spectral
// values of each thread and the symbols of its first kZcolPfModes modes into
// registers; zcol_solve stores them into LDS (loading the rest of a large tile
// directly), transforms, solves and writes back, and returns this thread's part
// of r.z = sum_k Re(conj(R - sum r) Z).  The caller loads the z twiddles into wz
// between the two.
constexpr int kZcolPf = 4, kZcolPfModes = 2;

struct ZcolTile {
  int ldz, TW, width, per;
  int64_t plane, col0;
  double2 pre[kZcolPf];
  hipFloatComplex q[kZcolPfModes][6];
};

__device__ __forceinline__ void zcol_prefetch(int nx, int ny, int nz, int tile_w, int xt, int ky,
                                              const double2* __restrict__ S, int64_t C,
                                              const hipFloatComplex* __restrict__ kinv,
                                              ZcolTile& t) {
  t.ldz = nz | 1;  // odd row pitch: lanes a row apart fall in distinct LDS banks
  t.TW = tile_w;
  const int i0 = xt * tile_w;
  t.width = min(tile_w, nx - i0);
  t.plane = (int64_t)nx * ny;
  t.col0 = i0 + (int64_t)nx * ky;
  t.per = tile_w * nz;
#pragma unroll
  for (int j = 0; j < kZcolPf; ++j) {
    const int u = threadIdx.x + j * blockDim.x;
    t.pre[j] = make_double2(0.0, 0.0);
    if (u < 3 * t.per) {
      const int m = u / t.per, rem = u - m * t.per;
      const int z = rem / t.TW, ii = rem - z * t.TW;
      if (ii < t.width) t.pre[j] = S[(int64_t)m * C + t.col0 + ii + t.plane * z];
    }
  }
#pragma unroll
  for (int j = 0; j < kZcolPfModes; ++j) {
    const int u = threadIdx.x + j * blockDim.x;
    if (u < t.per) {
      const int kz = u / t.TW, ii = u - kz * t.TW;
      if (ii < t.width) {
        const int64_t k = t.col0 + ii + t.plane * kz;
#pragma unroll
        for (int c = 0; c < 6; ++c) t.q[j][c] = kinv[6 * k + c];
      }
    }
  }
}

__device__ __forceinline__ double zcol_solve(const ZcolTile& t, double2* __restrict__ S, int64_t C,
                                             const hipFloatComplex* __restrict__ kinv,
                                             const double sr[3], double inv_n, const DftAxis& az,
                                             double2* sm, const double2* wz) {
  const int ldz = t.ldz, TW = t.TW, width = t.width, per = t.per;
  double2* A = sm;                   // [3][TW][ldz]
  double2* T = sm + 3 * TW * ldz;    // [3][TW][ldz]
#pragma unroll
  for (int j = 0; j < kZcolPf; ++j) {
    const int u = threadIdx.x + j * blockDim.x;
    if (u < 3 * per) {
      const int m = u / per, rem = u - m * per;
      const int z = rem / TW, ii = rem - z * TW;
      A[(m * TW + ii) * ldz + z] = t.pre[j];
    }
  }
  for (int u = threadIdx.x + kZcolPf * blockDim.x; u < 3 * per; u += blockDim.x) {
    const int m = u / per, rem = u - m * per;
    const int z = rem / TW, ii = rem - z * TW;
    A[(m * TW + ii) * ldz + z] =
        ii < width ? S[(int64_t)m * C + t.col0 + ii + t.plane * z] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  // the three components' lines in one pass per stage (lines past `width` hold zeros)
  dft_stage<true, false>(A, T, 3 * TW, ldz, 1, az, wz);
  __syncthreads();
  dft_stage<false, false>(T, A, 3 * TW, ldz, 1, az, wz);
  __syncthreads();
  double rz = 0.0;
  for (int u = threadIdx.x, j = 0; u < per; u += blockDim.x, ++j) {
    const int kz = u / TW, ii = u - kz * TW;
    if (ii >= width) continue;
    const int64_t k = t.col0 + ii + t.plane * kz;
    double ar[3], ai[3], zr[3], zi[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double2 v = A[(m * TW + ii) * ldz + kz];
      ar[m] = v.x - sr[m];
      ai[m] = v.y;
    }
    if (j < kZcolPfModes) {
      hipFloatComplex q[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) q[c] = j == 0 ? t.q[0][c] : t.q[kZcolPfModes - 1][c];
      spec_solve(q, false, ar, ai, zr, zi);
    } else {
      spec_solve(kinv + 6 * k, false, ar, ai, zr, zi);
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      zr[m] *= inv_n;
      zi[m] *= inv_n;
      rz += ar[m] * zr[m] + ai[m] * zi[m];
      A[(m * TW + ii) * ldz + kz] = make_double2(zr[m], zi[m]);
    }
  }
  __syncthreads();
  dft_stage<true, true>(A, T, 3 * TW, ldz, 1, az, wz);
  __syncthreads();
  dft_stage<false, true>(T, A, 3 * TW, ldz, 1, az, wz);
  __syncthreads();
  for (int u = threadIdx.x; u < 3 * per; u += blockDim.x) {
    const int m = u / per, rem = u - m * per;
    const int z = rem / TW, ii = rem - z * TW;
    if (ii < width) S[(int64_t)m * C + t.col0 + ii + t.plane * z] = A[(m * TW + ii) * ldz + z];
  }
  return rz;
}

// x-line group [l0, l0 + nl) of an RVE (line l = y + n_y z holds classes
// l n_x .. l n_x + n_x - 1; nl n_x <= blockDim.x, one class per thread; class
// vectors interleaved [3 c + m], spectra S = [3][C] complex).  INIT: r = b_scale b,
// x = 0, p_old = 0; otherwise x += alpha p, r -= alpha y.  Then the x-DFT of each
// residual component into S[m] (y and z untransformed); acc receives this
// thread's (r.r, sum r_x, sum r_y, sum r_z).  wx loaded by the caller.
template <bool INIT>
__device__ __forceinline__ void xlines_update_fwd(int nx, int64_t l0, int nl, double alpha,
                                                  double b_scale, const double* __restrict__ b,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ y,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ p_old,
                                                  double2* __restrict__ S, int64_t C,
                                                  const DftAxis& ax, double2* sm,
                                                  const double2* wx, double acc[4]) {
  const int ld = nx | 1, L = xline_group(nx);  // odd row pitch
  const int ncl = nl * nx;
  const int64_t c0 = l0 * nx;
  const float inv_nx = 1.0f / (float)nx;
  double2* A = sm;               // [3][L][ld]
  double2* T = sm + 3 * L * ld;  // [3][L][ld]
#pragma unroll
  for (int s = 0; s < 4; ++s) acc[s] = 0.0;
  const int t = threadIdx.x;
  if (t < ncl) {
    const int64_t c = c0 + t;
    const int l = fdiv(t, inv_nx), i = t - l * nx;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double v;
      if (INIT) {
        v = b_scale * b[3 * c + m];
        x[3 * c + m] = 0.0;
        p_old[3 * c + m] = 0.0;
      } else {
        x[3 * c + m] += alpha * p[3 * c + m];
        v = r[3 * c + m] - alpha * y[3 * c + m];
      }
      r[3 * c + m] = v;
      acc[0] += v * v;
      acc[1 + m] += v;
      A[(m * L + l) * ld + i] = make_double2(v, 0.0);
    }
  }
  __syncthreads();
  // the three components' lines in one pass per stage (lines past nl are not stored)
  dft_stage<true, false>(A, T, 3 * L, ld, 1, ax, wx);
  __syncthreads();
  dft_stage<false, false>(T, A, 3 * L, ld, 1, ax, wx);
  __syncthreads();
  for (int u = threadIdx.x; u < 3 * ncl; u += blockDim.x) {
    const int m = u / ncl, tt = u - m * ncl;
    const int l = fdiv(tt, inv_nx);
    S[(int64_t)m * C + c0 + tt] = A[(m * L + l) * ld + (tt - l * nx)];
  }
}

// x-line group: inverse x-DFT of each component; z[3 c + m] = real part.
__device__ __forceinline__ void xlines_inv(int nx, int64_t l0, int nl,
                                           const double2* __restrict__ S, int64_t C,
                                           double* __restrict__ zout, const DftAxis& ax,
                                           double2* sm, const double2* wx) {
  const int ld = nx | 1, L = xline_group(nx);  // odd row pitch
  const int ncl = nl * nx;
  const int64_t c0 = l0 * nx;
  const float inv_nx = 1.0f / (float)nx;
  double2* A = sm;
  double2* T = sm + 3 * L * ld;
  for (int u = threadIdx.x; u < 3 * ncl; u += blockDim.x) {
    const int m = u / ncl, tt = u - m * ncl;
    const int l = fdiv(tt, inv_nx);
    A[(m * L + l) * ld + (tt - l * nx)] = S[(int64_t)m * C + c0 + tt];
  }
  __syncthreads();
  dft_stage<true, true>(A, T, 3 * L, ld, 1, ax, wx);
  __syncthreads();
  dft_stage<false, true>(T, A, 3 * L, ld, 1, ax, wx);
  __syncthreads();
  const int t = threadIdx.x;
  if (t < ncl) {
    const int l = fdiv(t, inv_nx), i = t - l * nx;
#pragma unroll
    for (int m = 0; m < 3; ++m) zout[3 * (c0 + t) + m] = A[(m * L + l) * ld + i].x;
  }
}

}  // namespace cia
