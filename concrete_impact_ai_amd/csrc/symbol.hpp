// Inverse Fourier symbol of the spectral PCG preconditioner, shared by the
// single-RVE and the batched periodic solvers.
#pragma once

#include <hip/hip_complex.h>

#include <complex>
#include <vector>

#include "hex8.hpp"
#include "point.hpp"

namespace cia {

// Inverse Fourier symbol of the reference-medium Hex8 stiffness on the full
// spectrum (kx fastest, then ky, kz), Hermitian 3x3 blocks stored in FP32 as
// (00, 01, 02, 11, 12, 22); the zero mode maps to 0.  FP32 storage keeps every
// block Hermitian positive definite, so the preconditioner stays SPD.
inline void spectral_symbol(int nx, int ny, int nz, const BoxGeom& g, double K0, double G0,
                            std::vector<hipFloatComplex>& out) {
  Tangent<double> t;
  t.K = K0;
  t.d = 2.0 * G0;
  t.r = 0.0;
  for (int m = 0; m < 6; ++m) t.n[m] = 0.0;
  double ke[8][3][8][3] = {};
  for (int q = 0; q < 8; ++q) {
    for (int b = 0; b < 8; ++b) {
      double cb[3][6];
      b_columns(q, b, g, cb);
      double sb[3][6];
      for (int m = 0; m < 3; ++m) tangent_apply(t, cb[m], sb[m]);
      for (int a = 0; a < 8; ++a) {
        double ca[3][6];
        b_columns(q, a, g, ca);
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m) {
            double v = 0.0;
            for (int z = 0; z < 6; ++z) v += ca[l][z] * sb[m][z];
            ke[a][l][b][m] += g.jw * v;
          }
      }
    }
  }
  double S[3][3][3][3][3] = {};  // [ox+1][oy+1][oz+1][l][m]
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      const int ox = (b & 1) - (a & 1), oy = ((b >> 1) & 1) - ((a >> 1) & 1),
                oz = ((b >> 2) & 1) - ((a >> 2) & 1);
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) S[ox + 1][oy + 1][oz + 1][l][m] += ke[a][l][b][m];
    }
  out.assign((size_t)6 * nx * ny * nz, make_hipFloatComplex(0.f, 0.f));
  const double two_pi = 6.283185307179586476925286766559;
  typedef std::complex<double> cd;
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        const size_t mode = kx + (size_t)nx * (ky + (size_t)ny * kz);
        if (kx == 0 && ky == 0 && kz == 0) continue;
        const double th[3] = {two_pi * kx / nx, two_pi * ky / ny, two_pi * kz / nz};
        cd A[3][3] = {};
        for (int ox = -1; ox <= 1; ++ox)
          for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
              const cd ph = std::exp(cd(0.0, th[0] * ox + th[1] * oy + th[2] * oz));
              for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) A[l][m] += S[ox + 1][oy + 1][oz + 1][l][m] * ph;
            }
        const cd c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        const cd c01 = A[0][2] * A[2][1] - A[0][1] * A[2][2];
        const cd c02 = A[0][1] * A[1][2] - A[0][2] * A[1][1];
        const cd det = A[0][0] * c00 + A[1][0] * c01 + A[2][0] * c02;
        const cd inv[6] = {c00 / det,
                           c01 / det,
                           c02 / det,
                           (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det,
                           (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det,
                           (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det};
        for (int u = 0; u < 6; ++u)
          out[6 * mode + u] = make_hipFloatComplex((float)inv[u].real(), (float)inv[u].imag());
      }
}

// Mode -k of mode k on the nx x ny x nz spectrum (kx fastest).
__device__ __forceinline__ int64_t neg_mode(int64_t k, int nx, int ny, int nz) {
  const int kx = (int)(k % nx), ky = (int)((k / nx) % ny), kz = (int)(k / ((int64_t)nx * ny));
  const int mx = kx ? nx - kx : 0, my = ky ? ny - ky : 0, mz = kz ? nz - kz : 0;
  return mx + (int64_t)nx * (my + (int64_t)ny * mz);
}

__device__ __forceinline__ void spec_solve(const hipFloatComplex* __restrict__ q, bool conj_sym,
                                           const double ar[3], const double ai[3], double zr[3],
                                           double zi[3]) {
  // Hermitian inverse stored as (00, 01, 02, 11, 12, 22); K^-1(-k) = conj K^-1(k).
  const double s = conj_sym ? -1.0 : 1.0;
  const double h[3][3][2] = {
      {{q[0].x, s * q[0].y}, {q[1].x, s * q[1].y}, {q[2].x, s * q[2].y}},
      {{q[1].x, -s * q[1].y}, {q[3].x, s * q[3].y}, {q[4].x, s * q[4].y}},
      {{q[2].x, -s * q[2].y}, {q[4].x, -s * q[4].y}, {q[5].x, s * q[5].y}}};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double xr = 0.0, xi = 0.0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      xr += h[m][l][0] * ar[l] - h[m][l][1] * ai[l];
      xi += h[m][l][0] * ai[l] + h[m][l][1] * ar[l];
    }
    zr[m] = xr;
    zi[m] = xi;
  }
}

}  // namespace cia
