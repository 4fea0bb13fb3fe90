// Structured periodic Hex8 RVE on one MI355X: Newton microequilibrium with a
// matrix-free block-Jacobi PCG linear solve, Armijo line search and the
// homogenised response.
//
// Reference: src/fem/rve/solver.py:62-518 (solve_rve_microequilibrium,
// _build_rve_response, _compute_effective_tangent), src/fem/rve/periodic.py
// (periodic classes and affine map), src/fem/solvers/line_search.py:40-67,
// src/fem/assembly/nonlinear_solid.py:113-180,428-493.
//
// Periodic structure.  On an nx x ny x nz box the periodic classes are the
// nodes of an nx x ny x nz periodic grid, class c = i + nx (j + ny k), which is
// the reference's class numbering (periodic.py:32-37).  Class 0 is the
// eliminated reference class.  The reduced unknown vector of the reference is
// w[3:] of the class vector w used here, and T^T f is the class-summed element
// force, which a periodic node gather produces directly.  u = T w + A eps is
// never formed: B (A eps) = eps exactly on a box element, so every Gauss-point
// strain is eps_macro + B_q w_e.
//
// Kernels (one thread per element or per node):
//   assemble   strain -> constitutive update -> element force rows F[24][E]
//   gather     node sum of F over the 8 elements of each class (fixed order,
//              no atomics: bitwise deterministic), residual norm partials
//   pcg_elem   p = z + beta p_old fused into the element gather, C_q eps, F
//   pcg_node   y = gather(F), p.y partials; last block computes alpha
//   pcg_update x, r, z = M^-1 r, r.z and r.r partials; last block computes beta
// Every reduction is a fixed-order block tree plus a last-block pass over the
// block partials.
#include <string.h>


#include <algorithm>
#include <cstddef>
#include <cmath>
#include <complex>
#include <vector>

#include "common.hpp"
#include "dft.hpp"
#include <hip/hip_complex.h>
#include "hex8.hpp"
#include "lanefft.hpp"
#include "spectral.hpp"
#include "symbol.hpp"

namespace cia {

struct Scalars {
  double rz, pq, alpha, beta, rr, tol2;
  double rzr[2];   // r.z of the last two spectral PCG iterations (ring by parity)
  double norm2[4];
  double sumr[3];  // per-component sum of the PCG residual (spectral preconditioner)
  int32_t done;    // 0 running, 1 converged, 2 breakdown
  int32_t iters;
  int32_t first;   // next spectral apply is the initial one (beta = 0)
  int32_t it_cur;  // index of the spectral PCG iteration in progress (set by the element pass)
};

struct Red {
  double* partials;    // [slots][maxblocks]
  unsigned* counter;
  int maxblocks;
};

// Every block sums the NV partial rows of the previous kernel into out[0..NV) in
// one LDS round: each thread sums its strided partials of all rows, each wave
// reduces them (wave_sum), and after one barrier every thread adds the wave sums
// in wave order.  The order is the same in every block (and the same as one
// block_sum per row), so all blocks obtain bitwise-identical scalars.  A kernel
// calls it at most once (its LDS array is not reused).
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int P,
                                                double out[NV], double* /*lds*/) {
  constexpr int NW = kBlock / 64;
  __shared__ double rp[NV][NW];
  double a[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) a[s] = 0.0;
  for (int b = threadIdx.x; b < P; b += blockDim.x) {
#pragma unroll
    for (int s = 0; s < NV; ++s) a[s] += part[s * P + b];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const double w = wave_sum(a[s]);
    if (lane == 0) rp[s][wid] = w;
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) r += rp[s][w];
    out[s] = r;
  }
}

// This block's NV partial sums, one LDS round (the same summation order as one
// block_sum per value); thread s writes row s.  A kernel calls it at most once.
template <int NV>
__device__ __forceinline__ void write_partials(const double (&v)[NV], double* __restrict__ part,
                                               double* /*lds*/) {
  constexpr int NW = kBlock / 64;
  __shared__ double wp[NV][NW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    const double w = wave_sum(v[s]);
    if (lane == 0) wp[s][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) r += wp[threadIdx.x][w];
    part[threadIdx.x * gridDim.x + blockIdx.x] = r;
  }
}

// One block: out[s] = sum of the NV partial rows written by write_partials.
template <int NV>
__global__ void __launch_bounds__(kBlock) k_finalize(const double* __restrict__ part, int P,
                                                 double* __restrict__ out) {
  __shared__ double lds[kBlock / 64];
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    double a = 0.0;
    for (int b = threadIdx.x; b < P; b += blockDim.x) a += part[s * P + b];
    a = block_sum<kBlock>(a, lds);
    if (threadIdx.x == 0) out[s] = a;
  }
}

template <int NV>
__device__ __forceinline__ void finalize_sums(const double (&v)[NV], double* partials,
                                              unsigned* counter, double* out, int nblocks,
                                              double* lds, bool* is_last) {
  double tot[NV];
#pragma unroll
  for (int s = 0; s < NV; ++s) tot[s] = block_sum<kBlock>(v[s], lds);
  __shared__ bool last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int s = 0; s < NV; ++s) partials[s * nblocks + blockIdx.x] = tot[s];
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    last = prev == (unsigned)(nblocks - 1);
  }
  __syncthreads();
  *is_last = last;
  if (!last) return;
  __threadfence();
#pragma unroll
  for (int s = 0; s < NV; ++s) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
      acc += __builtin_nontemporal_load(&partials[s * nblocks + b]);
    acc = block_sum<kBlock>(acc, lds);
    if (threadIdx.x == 0) out[s] = acc;
  }
  if (threadIdx.x == 0) *counter = 0u;
}

struct Grid {
  int nx, ny, nz;
  int pad;
  int64_t E;  // elements = classes
  BoxGeom g;
};

struct Macro {
  double e[6];   // requested macro strain
  double ec[6];  // committed macro strain
};

// -------------------------------------------------------------------------
// Assembly: strain -> constitutive update -> element forces.  One lane per
// Gauss point (8 lanes per element): every per-point row is read and written
// by consecutive lanes, and the element force is reduced across the 8 lanes.
// The Newton passes (EXTRA false) write only what the Newton step reads: the
// compact tangent and the force rows.  The final pass at the converged
// fluctuation (EXTRA) evaluates the points once more and writes the stresses,
// the next state and the response rows.
template <class T, bool EXTRA>
__global__ void __launch_bounds__(kBlock)
    k_assemble(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
               Settings S, double dt, const double* __restrict__ w, const double* __restrict__ wc,
               Macro M, const T* __restrict__ state_c, T* __restrict__ state_t,
               T* __restrict__ sig, T* __restrict__ tan, T* __restrict__ F,
               T* __restrict__ strain_out, T* __restrict__ rate_out, T* __restrict__ psi_out,
               T* __restrict__ diss_out, T* __restrict__ diag_out,
               unsigned long long* __restrict__ first_fail) {
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  const int64_t np = G.E * 8;
  // Lane q loads the fluctuation at corner q only; the element's 8 lanes share
  // their corners through LDS.
  __shared__ double cw[kBlock][EXTRA ? 6 : 3];
  if (e < G.E) {
    const int64_t cq = periodic_corner(e, q, G.nx, G.ny, G.nz);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const double wv = w[3 * cq + m];
      cw[threadIdx.x][m] = wv;
      if (EXTRA) cw[threadIdx.x][3 + m] = wv - wc[3 * cq + m];
    }
  }
  __syncthreads();
  if (e >= G.E) return;  // after the block's only barrier; whole 8-lane groups leave
  T bx[8], by[8], bz[8];
  gp_factors<T>(q, G.g, bx, by, bz);
  T eps[6], rate[6];
  {
    double we[8][3];
    const int base = threadIdx.x & ~7;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) we[a][m] = cw[base + a][m];
    strain_from(we, bx, by, bz, eps);
    if (EXTRA) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m) we[a][m] = cw[base + a][3 + m];
      strain_from(we, bx, by, bz, rate);
    }
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) eps[m] += (T)M.e[m];
  if (EXTRA) {
#pragma unroll
    for (int m = 0; m < 6; ++m) rate[m] = (rate[m] + (T)(M.e[m] - M.ec[m])) / (T)dt;
  }
  const PhaseConst P = tab->p[ephase[e]];
  const bool j2 = P.kind == CIA_MAT_J2_VISCOPLASTIC;
  T epsp[6];
  T qb = (T)0, ga = (T)0;
  if (j2) {
#pragma unroll
    for (int m = 0; m < 6; ++m) epsp[m] = state_c[m * np + p];
    qb = state_c[6 * np + p];
    ga = state_c[7 * np + p];
  } else {
#pragma unroll
    for (int m = 0; m < 6; ++m) epsp[m] = (T)0;
  }
  PointOut<T> o;
  PointExtra<T> ex;
  const int st = material_point<T, EXTRA>(P, S, dt, eps, EXTRA ? rate : nullptr, epsp, qb, ga, o,
                                          EXTRA ? &ex : nullptr, nullptr);
  if (st != CIA_PT_OK) atomicMin(first_fail, (unsigned long long)p);
  if (EXTRA) {
#pragma unroll
    for (int m = 0; m < 6; ++m) sig[m * np + p] = o.sig[m];
    if (j2) {
#pragma unroll
      for (int m = 0; m < 6; ++m) state_t[m * np + p] = o.epsp[m];
      state_t[6 * np + p] = o.qbar;
      state_t[7 * np + p] = o.gam;
    } else {
      // elastic points carry no internal state: every row of the output is written
#pragma unroll
      for (int m = 0; m < 8; ++m) state_t[m * np + p] = (T)0;
    }
  }
  tan[0 * np + p] = o.tan.d;
  tan[1 * np + p] = o.tan.r;
  if (o.tan.r != (T)0) {
#pragma unroll
    for (int m = 0; m < 6; ++m) tan[(2 + m) * np + p] = o.tan.n[m];
  }
  if (EXTRA) {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      strain_out[m * np + p] = eps[m];
      rate_out[m * np + p] = rate[m];
    }
    psi_out[p] = ex.psi;
    diss_out[p] = ex.diss;
#pragma unroll
    for (int m = 0; m < CIA_N_DIAG; ++m) diag_out[m * np + p] = ex.d[m];
  }
  T f[8][3];
  force_from(o.sig, (T)G.g.jw, bx, by, bz, f);
  T fin[3];
  element_reduce8(q, f, fin);
#pragma unroll
  for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = fin[m];
}

// Failure record of point p (re-evaluated with the record attached).
template <class T>
__global__ void k_diagnose(Grid G, const PhaseTable* __restrict__ tab,
                           const uint8_t* __restrict__ ephase, Settings S, double dt,
                           const double* __restrict__ w, const double* __restrict__ wc, Macro M,
                           const T* __restrict__ state_c, int64_t p, cia_point_failure* rec) {
  const int64_t e = p / 8;
  const int q = (int)(p % 8);
  const int64_t np = G.E * 8;
  int64_t c[8];
  int i, j, k;
  periodic_corners(e, G.nx, G.ny, G.nz, c, i, j, k);
  T we[8][3], dwe[8][3];
  for (int a = 0; a < 8; ++a)
    for (int m = 0; m < 3; ++m) {
      we[a][m] = (T)w[3 * c[a] + m];
      dwe[a][m] = (T)(w[3 * c[a] + m] - wc[3 * c[a] + m]);
    }
  T eps[6], rate[6];
  switch (q) {
    case 0: strain_at<0>(we, G.g, eps); strain_at<0>(dwe, G.g, rate); break;
    case 1: strain_at<1>(we, G.g, eps); strain_at<1>(dwe, G.g, rate); break;
    case 2: strain_at<2>(we, G.g, eps); strain_at<2>(dwe, G.g, rate); break;
    case 3: strain_at<3>(we, G.g, eps); strain_at<3>(dwe, G.g, rate); break;
    case 4: strain_at<4>(we, G.g, eps); strain_at<4>(dwe, G.g, rate); break;
    case 5: strain_at<5>(we, G.g, eps); strain_at<5>(dwe, G.g, rate); break;
    case 6: strain_at<6>(we, G.g, eps); strain_at<6>(dwe, G.g, rate); break;
    default: strain_at<7>(we, G.g, eps); strain_at<7>(dwe, G.g, rate); break;
  }
  for (int m = 0; m < 6; ++m) {
    eps[m] += (T)M.e[m];
    rate[m] = (rate[m] + (T)(M.e[m] - M.ec[m])) / (T)dt;
  }
  const PhaseConst P = tab->p[ephase[e]];
  T epsp[6] = {0, 0, 0, 0, 0, 0};
  T qb = 0, ga = 0;
  if (P.kind == CIA_MAT_J2_VISCOPLASTIC) {
    for (int m = 0; m < 6; ++m) epsp[m] = state_c[m * np + p];
    qb = state_c[6 * np + p];
    ga = state_c[7 * np + p];
  }
  cia_point_failure r;
  memset(&r, 0, sizeof(r));
  PointOut<T> o;
  PointExtra<T> ex;
  r.reason = material_point<T, true>(P, S, dt, eps, rate, epsp, qb, ga, o, &ex, &r);
  r.point_id = p;
  *rec = r;
}

// Periodic node gather of the element force rows.  Class 0 is zeroed after its
// raw sum is kept in class0[3] (for the class reaction diagnostic).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_gather(Grid G, const T* __restrict__ F, double* __restrict__ r, double* __restrict__ class0,
             double* partials) {
  __shared__ double lds[kBlock / 64];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (c < G.E) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t e = periodic_corner_element(i, j, k, a, G.nx, G.ny, G.nz);
#pragma unroll
      for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * G.E + e];
    }
    if (c == 0) {
      if (class0) {
        class0[0] = s[0];
        class0[1] = s[1];
        class0[2] = s[2];
      }
      s[0] = s[1] = s[2] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      r[3 * c + m] = s[m];
      acc += s[m] * s[m];
    }
  }
  const double v[1] = {acc};
  write_partials<1>(v, partials, lds);
}

// ||x||^2, ||y||^2 (y may be null) -> block partials [2][gridDim]
__global__ void __launch_bounds__(kBlock)
    k_norms(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
            double* partials) {
  __shared__ double lds[kBlock / 64];
  double a = 0.0, b = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double xv = x[t];
    a += xv * xv;
    if (y) {
      const double yv = y[t];
      b += yv * yv;
    }
  }
  const double v[2] = {a, b};
  write_partials<2>(v, partials, lds);
}

// out = alpha * x
__global__ void k_scale(int64_t n, double alpha, const double* __restrict__ x,
                        double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = alpha * x[t];
}

// out = x + alpha * y
__global__ void k_axpy(int64_t n, const double* __restrict__ x, double alpha,
                       const double* __restrict__ y, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = x[t] + alpha * y[t];
}

// -------------------------------------------------------------------------
// Block-Jacobi preconditioner: inverse of the 3x3 diagonal block per class.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_precond(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
              const T* __restrict__ tan, double* __restrict__ minv) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.E) return;
  const int64_t np = G.E * 8;
  const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
  double kb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a = 0; a < 8; ++a) {
    const int64_t e = periodic_corner_element(i, j, k, a, G.nx, G.ny, G.nz);
    const double K = tab->p[ephase[e]].K;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t p = 8 * e + q;
      Tangent<double> tc;
      tc.K = K;
      tc.d = (double)tan[p];
      tc.r = (double)tan[np + p];
#pragma unroll
      for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != 0.0 ? (double)tan[(2 + m) * np + p] : 0.0;
      double col[3][6];
      b_columns(q, a, G.g, col);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        double s[6];
        tangent_apply(tc, col[m], s);
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          double v = 0.0;
#pragma unroll
          for (int z = 0; z < 6; ++z) v += col[l][z] * s[z];
          kb[l][m] += G.g.jw * v;
        }
      }
    }
  }
  double inv[6];
  if (c == 0) {
    inv[0] = inv[3] = inv[5] = 1.0;
    inv[1] = inv[2] = inv[4] = 0.0;
  } else {
    const double a00 = kb[0][0], a01 = 0.5 * (kb[0][1] + kb[1][0]), a02 = 0.5 * (kb[0][2] + kb[2][0]);
    const double a11 = kb[1][1], a12 = 0.5 * (kb[1][2] + kb[2][1]), a22 = kb[2][2];
    const double c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const double det = a00 * c00 + a01 * c01 + a02 * c02;
    const double id = 1.0 / det;
    inv[0] = c00 * id;
    inv[1] = c01 * id;
    inv[2] = c02 * id;
    inv[3] = (a00 * a22 - a02 * a02) * id;
    inv[4] = (a01 * a02 - a00 * a12) * id;
    inv[5] = (a00 * a11 - a01 * a01) * id;
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) minv[m * G.E + c] = inv[m];
}

__device__ __forceinline__ void apply_minv(const double* __restrict__ minv, int64_t C, int64_t c,
                                           const double r[3], double z[3]) {
  const double m0 = minv[c], m1 = minv[C + c], m2 = minv[2 * C + c], m3 = minv[3 * C + c],
               m4 = minv[4 * C + c], m5 = minv[5 * C + c];
  z[0] = m0 * r[0] + m1 * r[1] + m2 * r[2];
  z[1] = m1 * r[0] + m3 * r[1] + m4 * r[2];
  z[2] = m2 * r[0] + m4 * r[1] + m5 * r[2];
}

// -------------------------------------------------------------------------
// PCG.  x = 0, r = b, z = M^-1 r, p_old = 0; rz, rr and the tolerance.
__global__ void __launch_bounds__(kBlock)
    k_pcg_init(int64_t C, const double* __restrict__ b, const double* __restrict__ minv,
               double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
               double* __restrict__ p_old, double rtol, double atol, Scalars* sc, double* partials,
               unsigned* counter) {
  __shared__ double lds[kBlock / 64];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (c < C) {
    double rv[3], zv[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) rv[m] = b[3 * c + m];
    apply_minv(minv, C, c, rv, zv);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      x[3 * c + m] = 0.0;
      r[3 * c + m] = rv[m];
      z[3 * c + m] = zv[m];
      p_old[3 * c + m] = 0.0;
      rz += rv[m] * zv[m];
      rr += rv[m] * rv[m];
    }
  }
  const double v[2] = {rz, rr};
  double out[2];
  bool last;
  finalize_sums<2>(v, partials, counter, out, (int)gridDim.x, lds, &last);
  if (last && threadIdx.x == 0) {
    sc->rz = out[0];
    sc->rr = out[1];
    sc->beta = 0.0;
    const double tol = fmax(rtol * sqrt(out[1]), atol);
    sc->tol2 = tol * tol;
    sc->iters = 0;
    sc->done = (out[1] <= sc->tol2) ? 1 : 0;
    if (!(out[0] > 0.0) && sc->done == 0) sc->done = 2;
  }
}

// Element part of y = K p with p = z + beta p_old fused in.  One lane per
// Gauss point; p_new is written for the class whose (0,0,0) corner is this
// element (lane 0 of the group: one writer per class).
template <class T, bool SHIFT>
__global__ void __launch_bounds__(kBlock)
    k_pcg_elem(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
               const T* __restrict__ tan, const double* __restrict__ z,
               const double* __restrict__ p_old, double* __restrict__ p_new,
               T* __restrict__ F, const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  if (e >= G.E) return;
  const double beta = sc->beta;
  const int64_t np = G.E * 8;
  // Spectral preconditioner: z arrives packed as complex (z_x + i z_y | z_z + i 0)
  // and is defined up to the translation that pins class 0.
  const int64_t C = G.E;
  const hipDoubleComplex* zc = reinterpret_cast<const hipDoubleComplex*>(z);
  double z0[3] = {0.0, 0.0, 0.0};
  if (SHIFT) {
    z0[0] = zc[0].x;
    z0[1] = zc[0].y;
    z0[2] = zc[C].x;
  }
  int64_t c[8];
  int i, j, k;
  periodic_corners(e, G.nx, G.ny, G.nz, c, i, j, k);
  T eps[6];
  T bx[8], by[8], bz[8];
  gp_factors<T>(q, G.g, bx, by, bz);
  {
    double pe[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      double zv[3];
      if (SHIFT) {
        const hipDoubleComplex t01 = zc[c[a]], t2 = zc[C + c[a]];
        zv[0] = t01.x - z0[0];
        zv[1] = t01.y - z0[1];
        zv[2] = t2.x - z0[2];
      } else {
#pragma unroll
        for (int m = 0; m < 3; ++m) zv[m] = z[3 * c[a] + m];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) pe[a][m] = zv[m] + beta * p_old[3 * c[a] + m];
    }
    if (q == 0) {
#pragma unroll
      for (int m = 0; m < 3; ++m) p_new[3 * e + m] = pe[0][m];
    }
    strain_from(pe, bx, by, bz, eps);
  }
  Tangent<T> tc;
  tc.K = (T)tab->p[ephase[e]].K;
  tc.d = tan[p];
  tc.r = tan[np + p];
#pragma unroll
  for (int m = 0; m < 6; ++m) tc.n[m] = (T)0;
  if (tc.r != (T)0) {
#pragma unroll
    for (int m = 0; m < 6; ++m) tc.n[m] = tan[(2 + m) * np + p];
  }
  T s[6];
  tangent_apply(tc, eps, s);
  T f[8][3];
  force_from(s, (T)G.g.jw, bx, by, bz, f);
  T fin[3];
  element_reduce8(q, f, fin);
#pragma unroll
  for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = fin[m];
}

// y = gather(F) (class 0 zeroed), p.y -> alpha.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pcg_node(Grid G, const T* __restrict__ F, const double* __restrict__ p,
               double* __restrict__ y, Scalars* sc, double* partials, unsigned* counter) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double pq = 0.0;
  if (c < G.E) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t e = periodic_corner_element(i, j, k, a, G.nx, G.ny, G.nz);
#pragma unroll
      for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * G.E + e];
    }
    if (c == 0) s[0] = s[1] = s[2] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      y[3 * c + m] = s[m];
      pq += p[3 * c + m] * s[m];
    }
  }
  const double v[1] = {pq};
  double out[1];
  bool last;
  finalize_sums<1>(v, partials, counter, out, (int)gridDim.x, lds, &last);
  if (last && threadIdx.x == 0) {
    sc->pq = out[0];
    if (!(out[0] > 0.0) || !isfinite(out[0])) {
      sc->done = 2;
      sc->alpha = 0.0;
    } else {
      sc->alpha = sc->rz / out[0];
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    k_pcg_update(int64_t C, const double* __restrict__ minv, const double* __restrict__ p,
                 const double* __restrict__ y, double* __restrict__ x, double* __restrict__ r,
                 double* __restrict__ z, Scalars* sc, double* partials, unsigned* counter) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const double alpha = sc->alpha;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (c < C) {
    double rv[3], zv[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      x[3 * c + m] += alpha * p[3 * c + m];
      rv[m] = r[3 * c + m] - alpha * y[3 * c + m];
      r[3 * c + m] = rv[m];
    }
    apply_minv(minv, C, c, rv, zv);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      z[3 * c + m] = zv[m];
      rz += rv[m] * zv[m];
      rr += rv[m] * rv[m];
    }
  }
  const double v[2] = {rz, rr};
  double out[2];
  bool last;
  finalize_sums<2>(v, partials, counter, out, (int)gridDim.x, lds, &last);
  if (last && threadIdx.x == 0) {
    sc->beta = out[0] / sc->rz;
    sc->rz = out[0];
    sc->rr = out[1];
    sc->iters += 1;
    if (out[1] <= sc->tol2) sc->done = 1;
    else if (!isfinite(out[1]) || !(out[0] > 0.0)) sc->done = 2;
  }
}


// -------------------------------------------------------------------------
// Spectral (FFT) reference-medium preconditioner.
//
// K0 is the Hex8 stiffness of a homogeneous isotropic reference medium on the
// periodic grid: a block-circulant operator, diagonalised by the 3-D DFT into
// 3x3 Hermitian blocks Khat(k).  For the pinned-class system (class 0 fixed)
// its exact inverse is  z = P Kinv0 (r - (sum r) delta_0),  P = shift so that
// z(class 0) = 0.  The spectrum of r - (sum r) delta_0 is rhat(k) - sum r, the
// shift is applied by the consumers (z(c) - z(0)), and r.z is evaluated by
// Parseval on the half spectrum.

// Packed spectra: ua = FFT(r_x + i r_y), ub = FFT(r_z) over the full grid of N
// modes.  For each pair (k, -k) one thread splits the component spectra
// (Rx = (ua(k) + conj ua(-k)) / 2, Ry = (ua(k) - conj ua(-k)) / 2i, Rz = ub(k)),
// removes sum(r) (the reaction at the pinned class), applies Khat^-1 / N and
// writes va = Zx + i Zy, vb = Zz in place, so that the inverse transforms give
// z_x + i z_y and z_z.  r.z by Parseval over the full spectrum.
// -------------------------------------------------------------------------
// Spectral PCG with deferred reductions.  A kernel that produces a dot product
// writes one partial per block; every block of the next kernel sums those
// partials itself, in the same fixed order, so all blocks see bitwise-identical
// scalars and no kernel needs a grid-wide fence (the kernel boundary makes the
// partials visible across XCDs).  Block 0 records the scalars for the host.


// Spectral PCG kernels (the passes of spectral.hpp, in their x-line form: a
// single RVE needs many small workgroups to fill the chip).
//   k_sp_xfwd<INIT>  one block per group of x-lines: alpha from the p.y partials,
//                    the vector update (INIT: x = 0, r = b, p_old = 0) and the
//                    forward x-DFT of r; partials (r.r, sum r)
//   k_dft_strided    y passes (dft.hpp), three components
//   k_sp_zcol        convergence test on the update partials, then z-DFT, symbol
//                    solve, inverse z-DFT per z-column tile; r.z partials
//   k_sp_xinv        inverse x-DFT per group of x-lines: z (real, interleaved)
template <bool INIT>
__global__ void __launch_bounds__(kBlock)
    k_sp_xfwd(int nx, int64_t n_lines, DftAxis ax, const double* __restrict__ b,
              const double* __restrict__ p, const double* __restrict__ y, double* __restrict__ x,
              double* __restrict__ r, double* __restrict__ p_old, double2* __restrict__ S, int64_t C,
              const double* __restrict__ part_pq, int P, double* __restrict__ part_upd,
              Scalars* __restrict__ sc) {
  extern __shared__ double2 dft_sm[];
  __shared__ double2 wx[kDftMaxTw];
  __shared__ double lds[kBlock / 64];
  double alpha = 0.0;
  int it = 0;
  if (!INIT) {
    if (sc->done) return;
    double pq[1];
    reduce_partials<1>(part_pq, P, pq, lds);
    if (!(pq[0] > 0.0) || !isfinite(pq[0])) {
      if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 2;
      return;
    }
    it = sc->it_cur;
    alpha = sc->rzr[it & 1] / pq[0];
  }
  load_twiddles(ax, wx);
  const int L = xline_group(nx);
  const int64_t l0 = (int64_t)blockIdx.x * L;
  const int nl = (int)min<int64_t>(L, n_lines - l0);
  double acc[4];
  xlines_update_fwd<INIT>(nx, l0, nl, alpha, 1.0, b, p, y, x, r, p_old, S, C, ax, dft_sm, wx, acc);
  write_partials<4>(acc, part_upd, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (INIT) {
      sc->done = 0;
      sc->iters = 0;
    } else {
      sc->iters = it + 1;
      sc->alpha = alpha;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    k_sp_zcol(int nx, int ny, int nz, int tile_w, double inv_n, DftAxis az,
              const hipFloatComplex* __restrict__ kinv, double2* __restrict__ S, int64_t C,
              const double* __restrict__ part_upd, int P, double* __restrict__ part_rz,
              Scalars* __restrict__ sc, int first, double rtol, double atol) {
  extern __shared__ double2 dft_sm[];
  __shared__ double2 wz[kDftMaxTw];
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  // every global load of the tile is issued before the wait for the scalars
  const int tiles_x = (nx + tile_w - 1) / tile_w;
  const int bz = (int)xcd_block();  // neighbouring tiles share the XCD's L2
  ZcolTile tile;
  zcol_prefetch(nx, ny, nz, tile_w, bz % tiles_x, bz / tiles_x, S, C, kinv, tile);
  load_twiddles(az, wz);
  double u[4];
  reduce_partials<4>(part_upd, P, u, lds);
  double tol2;
  if (first) {
    const double tol = fmax(rtol * sqrt(u[0]), atol);
    tol2 = tol * tol;
  } else {
    tol2 = sc->tol2;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rr = u[0];
    if (first) sc->tol2 = tol2;
  }
  if (u[0] <= tol2 || !isfinite(u[0])) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = isfinite(u[0]) ? 1 : 2;
    return;
  }
  const double sr[3] = {u[1], u[2], u[3]};
  const double v[1] = {zcol_solve(tile, S, C, kinv, sr, inv_n, az, dft_sm, wz)};
  write_partials<1>(v, part_rz, lds);
}

__global__ void __launch_bounds__(kBlock)
    k_sp_xinv(int nx, int64_t n_lines, DftAxis ax, const double2* __restrict__ S, int64_t C,
              double* __restrict__ zout, const Scalars* __restrict__ sc) {
  extern __shared__ double2 dft_sm[];
  __shared__ double2 wx[kDftMaxTw];
  if (sc->done) return;
  load_twiddles(ax, wx);
  const int L = xline_group(nx);
  const int64_t l0 = (int64_t)blockIdx.x * L;
  const int nl = (int)min<int64_t>(L, n_lines - l0);
  xlines_inv(nx, l0, nl, S, C, zout, ax, dft_sm, wx);
}

// Row-lane passes (lanefft.hpp) for y and z edges of 16, 32, 48 or 64: a 16-lane row
// holds one line of one x mode of one component, a wave the lines of 4 consecutive
// x modes of one component, and a block of kLaneBlock threads (6 waves) those of 8
// x modes (128-byte rows per load) for the three components.  No LDS staging and
// no barrier inside a transform.
constexpr int kLaneWaves = 6;
constexpr int kLaneBlock = 64 * kLaneWaves;
constexpr int kLaneKx = 8;

__host__ __device__ inline int lane_tiles_x(int nx) { return (nx + kLaneKx - 1) / kLaneKx; }
constexpr int kZlineKx = 16;  // x modes per block of the z pass

// Position of this thread in a row-lane block: component, x mode, lane in the row.
struct LanePos {
  int m, kx, l;
};

__device__ __forceinline__ LanePos lane_pos(int xt) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  return {wave % 3, xt * kLaneKx + 4 * (wave / 3) + (lane >> 4), lane & 15};
}

// y pass (INV: inverse).  The forward pass stores its modes in natural order (lane l,
// register k1 -> mode k1 + M bitrev4(l)); the inverse pass reads them back the same
// way.  Each launch has one block more: the forward one sums the update partials of
// the x pass once into sc (r.r, sum r per component) and sets the PCG tolerance on
// the first application; the inverse one sums the z pass's r.z partials and does the
// iteration's r.z bookkeeping and beta, which the next element pass reads.
template <int M, bool INV>
__global__ void __launch_bounds__(kLaneBlock)
    k_sp_ylines(int nx, int ny, int nz, int64_t C, DftAxis ay, double2* __restrict__ S,
                Scalars* __restrict__ sc, const double* __restrict__ part, int P, int first,
                double rtol, double atol) {
  // part: the forward launch's 4 rows of P update partials (r.r, sum r per component)
  // from the x pass, or the inverse launch's P r.z partials from the z pass
  const int tiles_x = lane_tiles_x(nx);
  const int nb = tiles_x * nz;
  if (INV && (int)blockIdx.x == nb) {
    // r.z from the z pass's block partials, once; the next element pass reads beta
    if (sc->done) return;
    __shared__ double wr[kLaneWaves];
    double a = 0.0;
    for (int b = threadIdx.x; b < P; b += kLaneBlock) a += part[b];
    const double w = wave_sum(a);
    if ((threadIdx.x & 63) == 0) wr[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      double rz = 0.0;
#pragma unroll
      for (int u = 0; u < kLaneWaves; ++u) rz += wr[u];
      if (!(rz > 0.0) || !isfinite(rz)) {
        sc->done = 2;
      } else {
        const int it = sc->iters;  // completed iterations: written only by the x pass
        const double beta = it == 0 ? 0.0 : rz / sc->rzr[(it + 1) & 1];
        sc->it_cur = it;
        sc->rzr[it & 1] = rz;
        sc->rz = rz;
        sc->beta = beta;
      }
    }
    return;
  }
  if (!INV && (int)blockIdx.x == nb) {
    if (sc->done) return;
    __shared__ double ws[kLaneWaves][4];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < P; b += kLaneBlock) {
#pragma unroll
      for (int s = 0; s < 4; ++s) a[s] += part[s * P + b];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double w = wave_sum(a[s]);
      if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][s] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double u[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kLaneWaves; ++w) t += ws[w][s];
        u[s] = t;
      }
      sc->rr = u[0];
      sc->sumr[0] = u[1];
      sc->sumr[1] = u[2];
      sc->sumr[2] = u[3];
      if (first) {
        const double tol = fmax(rtol * sqrt(u[0]), atol);
        sc->tol2 = tol * tol;
      }
    }
    return;
  }
  const int blk = (int)blockIdx.x;
  const LanePos lp = lane_pos(blk % tiles_x);
  const int z = blk / tiles_x, l = lp.l;
  const bool on = lp.kx < nx;
  double2* line = S + lp.m * C + lp.kx + (int64_t)nx * ny * z;
  const int br = bitrev4(l);
  double2 v[1][M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int y = INV ? j + M * br : l + 16 * j;
    v[0][j] = on ? line[(int64_t)nx * y] : make_double2(0.0, 0.0);
  }
  LaneTw<M> tw;
  lane_twiddles<M>(ay, l, tw);
  // the converged flag is read after the loads (one memory round trip)
  if (sc->done) return;
  if (INV)
    lane_dft_inv<M, 1>(v, tw, l);
  else
    lane_dft_fwd<M, 1>(v, tw, l);
  if (on) {
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const int y = INV ? l + 16 * j : j + M * br;
      line[(int64_t)nx * y] = v[0][j];
    }
  }
}

// Forward x pass with the PCG vector update for x edges of 16, 32, 48 or 64 (the
// row-lane form of k_sp_xfwd): a 16-lane row holds one (y, z) line for the three
// components (24 contiguous bytes per lane and position), a block of 256 threads
// 16 lines.  INIT: r = b, x = 0, p_old = 0; otherwise alpha from the p.y partials,
// x += alpha p, r -= alpha y.  Then the x-DFT of r into S (modes in natural order)
// and the block partials (r.r, sum r per component).  Every vector load is issued
// before the wait for the solver's flag and partials.
template <int M, bool INIT>
__global__ void __launch_bounds__(256)
    k_sp_xlines_fwd(int nx, int64_t n_lines, DftAxis ax, const double* __restrict__ b,
                    const double* __restrict__ p, const double* __restrict__ y,
                    double* __restrict__ x, double* __restrict__ r, double* __restrict__ p_old,
                    double2* __restrict__ S, int64_t C, const double* __restrict__ part_pq, int P,
                    double* __restrict__ part_upd, Scalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  const int64_t line = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int l = threadIdx.x & 15;
  const bool on = line < n_lines;
  const int64_t c0 = line * nx;
  double a0[M][3], a1[M][3], a2[M][3], a3[M][3];  // INIT: b; otherwise p, y, r, x
#pragma unroll
  for (int j = 0; j < M; ++j)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * (c0 + l + 16 * j) + m;
      if (INIT) {
        a0[j][m] = on ? b[d] : 0.0;
      } else {
        a0[j][m] = on ? p[d] : 0.0;
        a1[j][m] = on ? y[d] : 0.0;
        a2[j][m] = on ? r[d] : 0.0;
        a3[j][m] = on ? x[d] : 0.0;
      }
    }
  LaneTw<M> tw;
  lane_twiddles<M>(ax, l, tw);
  double alpha = 0.0;
  int it = 0;
  if (!INIT) {
    if (sc->done) return;
    double pq[1];
    reduce_partials<1>(part_pq, P, pq, lds);
    if (!(pq[0] > 0.0) || !isfinite(pq[0])) {
      if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 2;
      return;
    }
    it = sc->it_cur;
    alpha = sc->rzr[it & 1] / pq[0];
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  double2 v[3][M];
#pragma unroll
  for (int j = 0; j < M; ++j)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * (c0 + l + 16 * j) + m;
      double rv;
      if (INIT) {
        rv = a0[j][m];
        if (on) {
          x[d] = 0.0;
          p_old[d] = 0.0;
        }
      } else {
        rv = a2[j][m] - alpha * a1[j][m];
        if (on) x[d] = a3[j][m] + alpha * a0[j][m];
      }
      if (on) r[d] = rv;
      acc[0] += rv * rv;
      acc[1 + m] += rv;
      v[m][j] = make_double2(rv, 0.0);
    }
  lane_dft_fwd<M, 3>(v, tw, l);
  if (on) {
    const int br = bitrev4(l);
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < M; ++k) S[m * C + c0 + k + M * br] = v[m][k];
  }
  write_partials<4>(acc, part_upd, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (INIT) {
      sc->done = 0;
      sc->iters = 0;
    } else {
      sc->iters = it + 1;
      sc->alpha = alpha;
    }
  }
}

// Inverse x pass for x edges of 16, 32, 48 or 64: a 16-lane row holds one (y, z) line
// for the three components, a block of 256 threads 16 lines; z = Re(inverse x-DFT),
// written interleaved per class (24 contiguous bytes per lane and position).
template <int M>
__global__ void __launch_bounds__(256)
    k_sp_xlines_inv(int nx, int64_t n_lines, int64_t C, DftAxis ax, const double2* __restrict__ S,
                    double* __restrict__ zout, const Scalars* __restrict__ sc) {
  const int64_t line = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int l = threadIdx.x & 15;
  const bool on = line < n_lines;
  const int br = bitrev4(l);
  const int64_t c0 = line * nx;
  double2 v[3][M];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int j = 0; j < M; ++j) v[m][j] = on ? S[m * C + c0 + j + M * br] : make_double2(0.0, 0.0);
  LaneTw<M> tw;
  lane_twiddles<M>(ax, l, tw);
  // the converged flag is read after the loads (one memory round trip)
  if (sc->done) return;
  lane_dft_inv<M, 3>(v, tw, l);
  if (on) {
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
      for (int m = 0; m < 3; ++m) zout[3 * (c0 + l + 16 * j) + m] = v[m][j].x;
  }
}

// z pass with the symbol solve: convergence test on the sums left by the forward y
// pass, forward z-DFT, Z = Khat^-1 (R - sum r) / N per mode (the lane's modes
// k1 + M bitrev4(l)), r.z by Parseval, inverse z-DFT, in place; one r.z partial per
// block.  Here a 16-lane row holds one column of one x mode for all three
// components (the symbol solve stays in registers), a wave 4 consecutive x modes
// and a block of 16 KX threads KX x modes.
template <int M, int KX>
__global__ void __launch_bounds__(16 * KX)
    k_sp_zlines(int nx, int ny, int nz, int64_t C, double inv_n, DftAxis az,
                const hipFloatComplex* __restrict__ kinv, double2* __restrict__ S,
                double* __restrict__ part_rz, Scalars* __restrict__ sc) {
  constexpr int NW = KX / 4;
  __shared__ double ws[NW];
  const int tiles_x = (nx + KX - 1) / KX;
  const int blk = (int)blockIdx.x;
  const int xt = blk % tiles_x, ky = blk / tiles_x;
  const int l = threadIdx.x & 15;
  const int kx = xt * KX + (threadIdx.x >> 4);
  const bool on = kx < nx;
  const int64_t plane = (int64_t)nx * ny;
  const int64_t col = kx + (int64_t)nx * ky;
  const int br = bitrev4(l);
  // every global load is issued before the wait for the PCG scalars
  double2 v[3][M];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int j = 0; j < M; ++j)
      v[m][j] = on ? S[m * C + col + plane * (l + 16 * j)] : make_double2(0.0, 0.0);
  float4 q[M][3];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const float4* src = reinterpret_cast<const float4*>(kinv + 6 * (col + plane * (k + M * br)));
#pragma unroll
    for (int c = 0; c < 3; ++c) q[k][c] = on ? src[c] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  LaneTw<M> tw;
  lane_twiddles<M>(az, l, tw);
  // the solver's flag and scalars are read after the loads (one memory round trip)
  if (sc->done) return;
  const double rr = sc->rr, tol2 = sc->tol2;
  if (rr <= tol2 || !isfinite(rr)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = isfinite(rr) ? 1 : 2;
    return;
  }
  const double sr[3] = {sc->sumr[0], sc->sumr[1], sc->sumr[2]};
  lane_dft_fwd<M, 3>(v, tw, l);
  double rz = 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const hipFloatComplex qq[6] = {
        make_hipFloatComplex(q[k][0].x, q[k][0].y), make_hipFloatComplex(q[k][0].z, q[k][0].w),
        make_hipFloatComplex(q[k][1].x, q[k][1].y), make_hipFloatComplex(q[k][1].z, q[k][1].w),
        make_hipFloatComplex(q[k][2].x, q[k][2].y), make_hipFloatComplex(q[k][2].z, q[k][2].w)};
    double ar[3], ai[3], zr[3], zi[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      ar[m] = v[m][k].x - sr[m];
      ai[m] = v[m][k].y;
    }
    spec_solve(qq, false, ar, ai, zr, zi);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      zr[m] *= inv_n;
      zi[m] *= inv_n;
      rz += ar[m] * zr[m] + ai[m] * zi[m];
      v[m][k] = make_double2(zr[m], zi[m]);
    }
  }
  lane_dft_inv<M, 3>(v, tw, l);
  if (on) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < M; ++j) S[m * C + col + plane * (l + 16 * j)] = v[m][j];
  }
  // out-of-range lanes hold zero symbols and add nothing
  const double w = wave_sum(rz);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < NW; ++u) t += ws[u];
    part_rz[blockIdx.x] = t;
  }
}

// y = gather(F) (class 0 zeroed); p.y partials.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sp_node(Grid G, const T* __restrict__ F, const double* __restrict__ p, double* __restrict__ y,
              double* __restrict__ part_pq, const Scalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const int64_t c = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double pq = 0.0;
  if (c < G.E) {
    const unsigned cu = (unsigned)c, rr = cu / (unsigned)G.nx;
    const int i = (int)(cu - rr * (unsigned)G.nx);
    const int k = (int)(rr / (unsigned)G.ny);
    const int j = (int)(rr - (unsigned)k * (unsigned)G.ny);
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t e = periodic_corner_element(i, j, k, a, G.nx, G.ny, G.nz);
#pragma unroll
      for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * G.E + e];
    }
    if (c == 0) s[0] = s[1] = s[2] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      y[3 * c + m] = s[m];
      pq += p[3 * c + m] * s[m];
    }
  }
  const double v[1] = {pq};
  write_partials<1>(v, part_pq, lds);
}

// Element pass: beta from the r.z partials, p = (z - z0) + beta p_old, forces.
// One lane per element: the lane forms p at its 8 corners, the strains of its 8
// Gauss points by sum factorisation (box_gradients), applies the compact
// tangents and spreads the traction sums back to the corners (box_forces) -- no
// corner exchange and no cross-lane force reduction.  At least two waves per SIMD:
// in FP64 the pass needs ~280 registers per lane, and a second resident wave
// (a few spilled values) beats one wave without spills (21.9 -> 19.5 us at 48^3).
//
// PRE: beta and the r.z bookkeeping were done once by the extra block of the inverse
// y launch (row-lane passes); otherwise every block reduces the r.z partials itself.
template <class T, bool PRE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
    k_sp_elem(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
              const T* __restrict__ tan, const double* __restrict__ zr,
              const double* __restrict__ p_old, double* __restrict__ p_new, T* __restrict__ F,
              const double* __restrict__ part_rz, int P, Scalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  double beta;
  if (PRE) {
    beta = sc->beta;
  } else {
    double rz[1];
    reduce_partials<1>(part_rz, P, rz, lds);
    if (!(rz[0] > 0.0) || !isfinite(rz[0])) {
      if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 2;
      return;
    }
    const int it = sc->iters;  // completed iterations: written only by k_sp_update
    beta = it == 0 ? 0.0 : rz[0] / sc->rzr[(it + 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->it_cur = it;
      sc->rzr[it & 1] = rz[0];
      sc->rz = rz[0];
      sc->beta = beta;
    }
  }
  const int64_t e = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  int64_t cq[8];
  int i, j, k;
  periodic_corners(e, G.nx, G.ny, G.nz, cq, i, j, k);
  T u[8][3];  // corner values of p, rounded to T (the strains are formed in T)
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    double o[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) o[m] = (zr[3 * cq[a] + m] - zr[m]) + beta * p_old[3 * cq[a] + m];
    if (a == 0) {
#pragma unroll
      for (int m = 0; m < 3; ++m) p_new[3 * e + m] = o[m];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) u[a][m] = (T)o[m];
  }
  T gX[2][2][3], gY[2][2][3], gZ[2][2][3];
  box_gradients<T>(u, G.g, gX, gY, gZ);
  Tangent<T> tc;
  tc.K = (T)tab->p[ephase[e]].K;
  T dd[8], rr[8];
  load8(tan + 8 * e, dd);
  load8(tan + np + 8 * e, rr);
  bool plastic = false;
#pragma unroll
  for (int q = 0; q < 8; ++q) plastic |= rr[q] != (T)0;
  T SX[2][2][3], SY[2][2][3], SZ[2][2][3];  // traction sums (box_traction_add)
#pragma unroll
  for (int qx = 0; qx < 2; ++qx) {
    T nn[6][4];  // the compact tangents' directions of the 4 points with this qx
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int h = 0; h < 4; ++h) nn[m][h] = (T)0;
    if (plastic) {
#pragma unroll
      for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int h = 0; h < 4; ++h) nn[m][h] = tan[(2 + m) * np + 8 * e + 4 * qx + h];
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int qy = h >> 1, qz = h & 1;
      T eps[6];
      box_strain(gX, gY, gZ, qx, qy, qz, eps);
      tc.d = dd[4 * qx + h];
      tc.r = rr[4 * qx + h];
#pragma unroll
      for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != (T)0 ? nn[m][h] : (T)0;
      T sg[6];
      tangent_apply(tc, eps, sg);
      box_traction_add(sg, qx, qy, qz, SX, SY, SZ);
    }
  }
  T f[8][3];
  box_forces<T>(SX, SY, SZ, G.g, f);
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m) F[(3 * a + m) * G.E + e] = f[a][m];
}

// -------------------------------------------------------------------------
// Effective tangent: element forces of the uniform strain e_j (K A e_j).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_affine_force(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
                   const T* __restrict__ tan, int jdir, T* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  Tangent<T> tc;
  tc.K = (T)tab->p[ephase[e]].K;
  T f[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) f[a][0] = f[a][1] = f[a][2] = (T)0;
  T ej[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) ej[m] = (m == jdir) ? (T)1 : (T)0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t p = 8 * e + q;
    tc.d = tan[p];
    tc.r = tan[np + p];
#pragma unroll
    for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != (T)0 ? tan[(2 + m) * np + p] : (T)0;
    T s[6];
    tangent_apply(tc, ej, s);
    switch (q) {
      case 0: force_at<0>(s, G.g, f); break;
      case 1: force_at<1>(s, G.g, f); break;
      case 2: force_at<2>(s, G.g, f); break;
      case 3: force_at<3>(s, G.g, f); break;
      case 4: force_at<4>(s, G.g, f); break;
      case 5: force_at<5>(s, G.g, f); break;
      case 6: force_at<6>(s, G.g, f); break;
      default: force_at<7>(s, G.g, f); break;
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m) F[(3 * a + m) * G.E + e] = f[a][m];
}

// Column jdir of sum_q C_q (e_j - B_q s) (times jw, summed per block).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_ceff_column(Grid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
                  const T* __restrict__ tan, const double* __restrict__ s, int jdir,
                  double* __restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (e < G.E) {
    const int64_t np = G.E * 8;
    int64_t c[8];
    int i, j, k;
    periodic_corners(e, G.nx, G.ny, G.nz, c, i, j, k);
    double se[8][3];
    for (int a = 0; a < 8; ++a)
      for (int m = 0; m < 3; ++m) se[a][m] = s[3 * c[a] + m];
    Tangent<double> tc;
    tc.K = tab->p[ephase[e]].K;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t p = 8 * e + q;
      tc.d = (double)tan[p];
      tc.r = (double)tan[np + p];
#pragma unroll
      for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != 0.0 ? (double)tan[(2 + m) * np + p] : 0.0;
      double eps[6], sg[6];
      switch (q) {
        case 0: strain_at<0>(se, G.g, eps); break;
        case 1: strain_at<1>(se, G.g, eps); break;
        case 2: strain_at<2>(se, G.g, eps); break;
        case 3: strain_at<3>(se, G.g, eps); break;
        case 4: strain_at<4>(se, G.g, eps); break;
        case 5: strain_at<5>(se, G.g, eps); break;
        case 6: strain_at<6>(se, G.g, eps); break;
        default: strain_at<7>(se, G.g, eps); break;
      }
#pragma unroll
      for (int m = 0; m < 6; ++m) eps[m] = (m == jdir ? 1.0 : 0.0) - eps[m];
      tangent_apply(tc, eps, sg);
#pragma unroll
      for (int m = 0; m < 6; ++m) acc[m] += sg[m];
    }
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const double t = block_sum<kBlock>(acc[m], lds);
    if (threadIdx.x == 0) partials[m * gridDim.x + blockIdx.x] = t;
  }
}

// -------------------------------------------------------------------------
// Response reductions over elements (sums per block, summed on the host).
enum {
  R_SIG = 0,       // 6
  R_PSI = 6,
  R_DISS = 7,
  R_MICRO_WORK = 8,
  R_POS_WORK = 9,
  R_NEG_WORK = 10,
  R_UNLOAD = 11,
  R_ACTIVE = 12,
  R_MARGIN_N = 13,
  R_VIRTUAL = 14,  // 6
  R_PHASE = 20,    // 9 per phase: count, sig(6), psi, diss
  R_N = R_PHASE + 9 * CIA_MAX_PHASES
};

// NPH >= number of phases: the per-phase sums use compile-time register indices.
// One lane per Gauss point, so that every point row is read in coalesced lines: a
// block takes the 8 kBlock points of the elements [kBlock b, kBlock (b + 1)) in 8
// rounds (one partial column per block, as for the element-per-lane kernels), and
// the lane of point q of element e also adds corner a = q of the virtual-work sum.
template <class T, int NPH>
__global__ void __launch_bounds__(kBlock)
    k_response(Grid G, const uint8_t* __restrict__ ephase, double dt, const T* __restrict__ sig,
               const T* __restrict__ rate, const T* __restrict__ psi, const T* __restrict__ diss,
               const T* __restrict__ diag, const T* __restrict__ F, double* __restrict__ partials,
               double* __restrict__ mm_partials) {
  constexpr int RN = R_PHASE + 9 * NPH;
  __shared__ double lds[kBlock / 64];
  __shared__ double red[kBlock / 64][RN];
  const int64_t np = G.E * 8;
  double v[RN];
#pragma unroll
  for (int s = 0; s < RN; ++s) v[s] = 0.0;
  double mn = INFINITY, mx = -INFINITY;
  const double hx = 1.0 / G.g.ihx, hy = 1.0 / G.g.ihy, hz = 1.0 / G.g.ihz;
  const int64_t p0 = (int64_t)blockIdx.x * (8 * kBlock) + threadIdx.x;
  for (int rd = 0; rd < 8; ++rd) {
    const int64_t p = p0 + rd * kBlock;
    if (p >= np) break;
    const int64_t e = p >> 3;
    const int a = (int)(p & 7);
    const int ph = ephase[e];
    double sg[6], mw = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      sg[m] = (double)sig[m * np + p];
      v[R_SIG + m] += sg[m];
      mw += sg[m] * (dt * (double)rate[m * np + p]);
    }
    v[R_MICRO_WORK] += mw;
    const double ps = (double)psi[p], ds = (double)diss[p];
    v[R_PSI] += ps;
    v[R_DISS] += ds;
    const double pw = (double)diag[CIA_D_INCREMENTAL_WORK_DENSITY * np + p];
    v[R_POS_WORK] += fmax(pw, 0.0);
    v[R_NEG_WORK] += fmin(pw, 0.0);
    v[R_UNLOAD] += pw < 0.0 ? 1.0 : 0.0;
    v[R_ACTIVE] += (double)diag[CIA_D_VISCOPLASTIC_ACTIVE * np + p] > 0.5 ? 1.0 : 0.0;
    const double mg = (double)diag[CIA_D_YIELD_ACTIVATION_MARGIN * np + p];
    if (isfinite(mg)) {
      v[R_MARGIN_N] += 1.0;
      mn = fmin(mn, mg);
      mx = fmax(mx, mg);
    }
#pragma unroll
    for (int P = 0; P < NPH; ++P) {
      if (ph == P) {
        v[R_PHASE + 9 * P] += 1.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) v[R_PHASE + 9 * P + 1 + m] += sg[m];
        v[R_PHASE + 9 * P + 7] += ps;
        v[R_PHASE + 9 * P + 8] += ds;
      }
    }
    // virtual work stress: A(x_a)^T F_a with x_a the unwrapped position of corner a
    const unsigned eu = (unsigned)e, rw = eu / (unsigned)G.nx;
    const int i = (int)(eu - rw * (unsigned)G.nx);
    const int k = (int)(rw / (unsigned)G.ny);
    const int j = (int)(rw - (unsigned)k * (unsigned)G.ny);
    const double x = (i + (a & 1)) * hx, y = (j + ((a >> 1) & 1)) * hy, z = (k + ((a >> 2) & 1)) * hz;
    const double fx = (double)F[(3 * a + 0) * G.E + e], fy = (double)F[(3 * a + 1) * G.E + e],
                 fz = (double)F[(3 * a + 2) * G.E + e];
    v[R_VIRTUAL + 0] += x * fx;
    v[R_VIRTUAL + 1] += y * fy;
    v[R_VIRTUAL + 2] += z * fz;
    v[R_VIRTUAL + 3] += 0.5 * z * fy + 0.5 * y * fz;
    v[R_VIRTUAL + 4] += 0.5 * z * fx + 0.5 * x * fz;
    v[R_VIRTUAL + 5] += 0.5 * y * fx + 0.5 * x * fy;
  }
  // one pass of wave sums, then a single LDS exchange
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < RN; ++s) {
    const double t = wave_sum(v[s]);
    if (lane == 0) red[wid][s] = t;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < RN; s += blockDim.x) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += red[w][s];
    partials[s * gridDim.x + blockIdx.x] = t;
  }
  const double tmn = block_min<kBlock>(mn, lds);
  const double tmx = block_max<kBlock>(mx, lds);
  if (threadIdx.x == 0) {
    mm_partials[2 * blockIdx.x] = tmn;
    mm_partials[2 * blockIdx.x + 1] = tmx;
  }
}

// Full (non-periodic) node quantities: displacement and internal force of
// every mesh node, face resultants, ||f||^2 and the periodic fluctuation error.
enum { N_F2 = 0, N_FACE = 1 /* 18 */, N_N = 19 };

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_full_nodes(Grid G, const T* __restrict__ F, const double* __restrict__ w, Macro M,
                 double* __restrict__ u_full, double* __restrict__ f_full,
                 double* __restrict__ partials, double* __restrict__ max_partials) {
  __shared__ double lds[kBlock / 64];
  const int nx1 = G.nx + 1, ny1 = G.ny + 1, nz1 = G.nz + 1;
  const int64_t N = (int64_t)nx1 * ny1 * nz1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[N_N];
#pragma unroll
  for (int s = 0; s < N_N; ++s) v[s] = 0.0;
  double fl_err = 0.0;
  if (t < N) {
    const int i = (int)(t % nx1), j = (int)((t / nx1) % ny1), k = (int)(t / ((int64_t)nx1 * ny1));
    double f[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; ++a) {
      const int ei = i - (a & 1), ej = j - ((a >> 1) & 1), ek = k - ((a >> 2) & 1);
      if (ei < 0 || ej < 0 || ek < 0 || ei >= G.nx || ej >= G.ny || ek >= G.nz) continue;
      const int64_t e = ei + (int64_t)G.nx * (ej + (int64_t)G.ny * ek);
      for (int m = 0; m < 3; ++m) f[m] += (double)F[(3 * a + m) * G.E + e];
    }
    const double hx = 1.0 / G.g.ihx, hy = 1.0 / G.g.ihy, hz = 1.0 / G.g.ihz;
    const int ci = i % G.nx, cj = j % G.ny, ck = k % G.nz;
    const int64_t c = ci + (int64_t)G.nx * (cj + (int64_t)G.ny * ck);
    auto affine = [&](double x, double y, double z, double out[3]) {
      out[0] = x * M.e[0] + 0.5 * z * M.e[4] + 0.5 * y * M.e[5];
      out[1] = y * M.e[1] + 0.5 * z * M.e[3] + 0.5 * x * M.e[5];
      out[2] = z * M.e[2] + 0.5 * y * M.e[3] + 0.5 * x * M.e[4];
    };
    double au[3], ar[3];
    affine(i * hx, j * hy, k * hz, au);
    affine(ci * hx, cj * hy, ck * hz, ar);
    double mism = 0.0;
    for (int m = 0; m < 3; ++m) {
      const double wc = w[3 * c + m];
      const double u = wc + au[m];
      const double ur = wc + ar[m];
      u_full[3 * t + m] = u;
      f_full[3 * t + m] = f[m];
      const double d = (u - au[m]) - (ur - ar[m]);
      mism += d * d;
      v[N_F2] += f[m] * f[m];
      if (i == 0) v[N_FACE + 0 + m] += f[m];
      if (i == G.nx) v[N_FACE + 3 + m] += f[m];
      if (j == 0) v[N_FACE + 6 + m] += f[m];
      if (j == G.ny) v[N_FACE + 9 + m] += f[m];
      if (k == 0) v[N_FACE + 12 + m] += f[m];
      if (k == G.nz) v[N_FACE + 15 + m] += f[m];
    }
    fl_err = sqrt(mism);
  }
  for (int s = 0; s < N_N; ++s) {
    const double tt = block_sum<kBlock>(v[s], lds);
    if (threadIdx.x == 0) partials[s * gridDim.x + blockIdx.x] = tt;
  }
  const double m = block_max<kBlock>(fl_err, lds);
  if (threadIdx.x == 0) max_partials[blockIdx.x] = m;
}

// max over classes of ||sum of element forces|| (class reaction resultant).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_class_max(Grid G, const T* __restrict__ F, double* __restrict__ max_partials) {
  __shared__ double lds[kBlock / 64];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0.0;
  if (c < G.E) {
    const int i = (int)(c % G.nx), j = (int)((c / G.nx) % G.ny), k = (int)(c / ((int64_t)G.nx * G.ny));
    double s[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 8; ++a) {
      const int64_t e = periodic_corner_element(i, j, k, a, G.nx, G.ny, G.nz);
      for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * G.E + e];
    }
    nrm = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  }
  const double m = block_max<kBlock>(nrm, lds);
  if (threadIdx.x == 0) max_partials[blockIdx.x] = m;
}

// The response partials reduced on the device, one block per output (fixed order):
// out[0, nrows) row sums of `part` (R_N rows of ge columns; out[nrows, R_N) = 0),
// out[R_N], out[R_N + 1] min / max of the margin partials, out[R_N + 2 + s] the N_N
// node sums, then the fluctuation and class-reaction maxima.
constexpr int kRespOut = R_N + 4 + N_N;

__global__ void __launch_bounds__(kBlock)
    k_response_rows(const double* __restrict__ part, int ge, int nrows, const double* __restrict__ mm,
                    const double* __restrict__ np_part, int gfull, const double* __restrict__ np_max,
                    const double* __restrict__ cl_max, int gc, double* __restrict__ out) {
  __shared__ double lds[kBlock / 64];
  const int b = (int)blockIdx.x;
  const int t = (int)threadIdx.x;
  if (b < R_N) {
    double a = 0.0;
    if (b < nrows)
      for (int c = t; c < ge; c += kBlock) a += part[(int64_t)b * ge + c];
    a = block_sum<kBlock>(a, lds);
    if (t == 0) out[b] = a;
  } else if (b < R_N + 2) {
    const int hi = b - R_N;
    double a = hi ? -INFINITY : INFINITY;
    for (int c = t; c < ge; c += kBlock) a = hi ? fmax(a, mm[2 * c + 1]) : fmin(a, mm[2 * c]);
    a = hi ? block_max<kBlock>(a, lds) : block_min<kBlock>(a, lds);
    if (t == 0) out[b] = a;
  } else if (b < R_N + 2 + N_N) {
    const int s = b - R_N - 2;
    double a = 0.0;
    for (int c = t; c < gfull; c += kBlock) a += np_part[(int64_t)s * gfull + c];
    a = block_sum<kBlock>(a, lds);
    if (t == 0) out[b] = a;
  } else {
    const bool fl = b == R_N + 2 + N_N;
    const double* src = fl ? np_max : cl_max;
    const int n = fl ? gfull : gc;
    double a = 0.0;
    for (int c = t; c < n; c += kBlock) a = fmax(a, src[c]);
    a = block_max<kBlock>(a, lds);
    if (t == 0) out[b] = a;
  }
}

template <class T>
__global__ void k_copy_cast(int64_t n, const T* __restrict__ src, double* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (double)src[t];
}


// -------------------------------------------------------------------------
// Optional kernel timing with HIP events on the launch stream (used by the
// benchmark to report the average duration of the dominant kernel).
namespace prof {
enum { ASSEMBLE = 0, PCG_ELEM, PCG_NODE, PCG_UPDATE, PRECOND, GATHER, N_IDS };
struct State {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ids;
  std::vector<int> tags;  // PCG iteration index of the launch, -1 for other launches
  size_t used = 0;
  double total_ms[N_IDS] = {0};
  int64_t count[N_IDS] = {0};
};
static State g;
static void begin(int id, hipStream_t st, int tag = -1) {
  if (!g.on) return;
  if (g.ev.size() < 2 * (g.used + 1)) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    g.ev.push_back(a);
    g.ev.push_back(b);
    g.ids.push_back(0);
    g.tags.push_back(-1);
  }
  g.ids[g.used] = id;
  g.tags[g.used] = tag;
  (void)hipEventRecord(g.ev[2 * g.used], st);
}
static void end(hipStream_t st) {
  if (!g.on) return;
  (void)hipEventRecord(g.ev[2 * g.used + 1], st);
  ++g.used;
}
// Call only after the stream has been synchronised.  Launches tagged with a PCG
// iteration index >= active_iters exited at once on the converged flag and are
// not counted.
static void flush(int active_iters = 1 << 30) {
  if (!g.on) return;
  for (size_t i = 0; i < g.used; ++i) {
    float ms = 0.f;
    if (g.tags[i] >= active_iters) continue;
    if (hipEventSynchronize(g.ev[2 * i + 1]) == hipSuccess &&
        hipEventElapsedTime(&ms, g.ev[2 * i], g.ev[2 * i + 1]) == hipSuccess) {
      g.total_ms[g.ids[i]] += ms;
      g.count[g.ids[i]] += 1;
    }
  }
  g.used = 0;
}
}  // namespace prof

}  // namespace cia

// =========================================================================
// Host object
// =========================================================================

struct cia_rve {
  int nx, ny, nz;
  int64_t E, np, C, N;
  cia::Grid G;
  double volume;
  int precision;
  size_t tsize;
  int n_phases;
  std::vector<cia::PhaseConst> phases;
  std::vector<int64_t> phase_points;
  cia_newton_settings newton;
  cia_linear_settings lin;
  double hm_scale;
  // device
  cia::PhaseTable* d_tab = nullptr;
  uint8_t* d_ephase = nullptr;
  void* sig[2] = {nullptr, nullptr};
  void* tan[2] = {nullptr, nullptr};
  void* F = nullptr;
  void* strain_out = nullptr;
  void* rate_out = nullptr;
  void* psi_out = nullptr;
  void* diss_out = nullptr;
  void* diag_out = nullptr;
  double* w = nullptr;
  double* w_cand = nullptr;
  double* dw = nullptr;
  double* r = nullptr;
  double* r_cand = nullptr;
  double* wc_buf = nullptr;
  double* pr = nullptr;
  double* pz = nullptr;
  double* pp[2] = {nullptr, nullptr};
  double* py = nullptr;
  double* minv = nullptr;
  double* sens = nullptr;
  double* u_full = nullptr;
  double* f_full = nullptr;
  double* partials = nullptr;
  double* norm_part = nullptr;  // [2][1024] partials of the vector norms (kept apart from the
                                // PCG partials, which stay live across PCG batches)
  unsigned* counter = nullptr;
  cia::Scalars* sc = nullptr;
  double* class0 = nullptr;
  unsigned long long* first_fail = nullptr;
  cia_point_failure* rec = nullptr;
  double* red = nullptr;  // response partials
  int64_t red_len = 0;
  // pinned host staging of the read-backs: a pageable destination adds a staged copy
  // to every host round trip of the Newton and PCG control.  The device scalar block
  // has the same layout up to `ff` (first_fail lives right after the scalars), so the
  // scalars and the failure flag come back in one copy.
  struct Staging {
    cia::Scalars sc;
    unsigned long long ff;
    double resp[cia::kRespOut];
  }* hst = nullptr;
  int pcg_total = 0;
  static constexpr int kHints = 32;
  int its_hint[kHints] = {0};  // PCG iterations of Newton iteration k in the last increment
  int cur = 0;
  bool have_fields = false;
  // spectral preconditioner
  int prec = CIA_PCG_BLOCK_JACOBI;
  bool fft_ready = false;
  double2* tw = nullptr;             // DFT twiddles of the three axes
  hipStream_t cap_stream = nullptr;  // private stream for graph capture
  struct PairGraph {
    int slot;
    double* x;
    hipGraphExec_t exec;
  };
  std::vector<PairGraph> graphs;     // captured PCG iteration pairs
  cia::DftAxis ax{}, ay{}, az{};
  hipFloatComplex* kinv = nullptr;   // [N][6] inverse symbol
  hipDoubleComplex* uab = nullptr;   // [3][N] residual / preconditioned spectra per component
  double ref_bulk = 0.0, ref_shear = 0.0;
};

namespace cia {

// The scalar block (and, with ff, the first failing point) read back through the
// pinned staging copy, with one synchronisation.
static cia_status fetch_scalars(cia_rve* R, hipStream_t st, Scalars* hs,
                                unsigned long long* ff = nullptr) {
  const size_t bytes = ff ? offsetof(cia_rve::Staging, ff) + sizeof(unsigned long long)
                          : sizeof(Scalars);
  CIA_HIP(hipMemcpyAsync(&R->hst->sc, R->sc, bytes, hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  if (hs) *hs = R->hst->sc;
  if (ff) *ff = R->hst->ff;
  return CIA_OK;
}

static cia_status dmalloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc", e, __FILE__, __LINE__);
    return CIA_ERR_OUT_OF_MEMORY;
  }
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) {
    set_error("hipMemset", e, __FILE__, __LINE__);
    return CIA_ERR_HIP;
  }
  return CIA_OK;
}


static cia_status norm2_async(cia_rve* R, hipStream_t st, int64_t n, const double* x, int slot);

// Capture two spectral PCG iterations (even: p_old = pp[1], odd: p_old = pp[0])
// on the library's private stream and cache the executable graph per (tangent
// slot, solution vector).
template <class F>
static cia_status pcg_pair_graph(cia_rve* R, int slot, double* x, F&& iteration,
                                 hipGraphExec_t* out) {
  for (auto& e : R->graphs)
    if (e.slot == slot && e.x == x) {
      *out = e.exec;
      return CIA_OK;
    }
  hipGraph_t g = nullptr;
  CIA_HIP(hipStreamBeginCapture(R->cap_stream, hipStreamCaptureModeThreadLocal));
  iteration(R->cap_stream, 0, -1);
  iteration(R->cap_stream, 1, -1);
  CIA_HIP(hipStreamEndCapture(R->cap_stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CIA_HIP(e);
  R->graphs.push_back({slot, x, exec});
  *out = exec;
  return CIA_OK;
}

template <class T>
struct Ops {
  static cia_status assemble(cia_rve* R, hipStream_t st, bool extra, int slot, const double* w,
                             const double* wc, const Macro& M, const void* state_c, void* state_t,
                             const Settings& S, double dt) {
    const unsigned grid = grid_for(R->np);
    prof::begin(prof::ASSEMBLE, st);
    if (extra)
      hipLaunchKernelGGL((k_assemble<T, true>), dim3(grid), dim3(kBlock), 0, st, R->G, R->d_tab,
                         R->d_ephase, S, dt, w, wc, M, (const T*)state_c, (T*)state_t,
                         (T*)R->sig[slot], (T*)R->tan[slot], (T*)R->F, (T*)R->strain_out,
                         (T*)R->rate_out, (T*)R->psi_out, (T*)R->diss_out, (T*)R->diag_out,
                         R->first_fail);
    else
      hipLaunchKernelGGL((k_assemble<T, false>), dim3(grid), dim3(kBlock), 0, st, R->G, R->d_tab,
                         R->d_ephase, S, dt, w, wc, M, (const T*)state_c, (T*)nullptr,
                         (T*)nullptr, (T*)R->tan[slot], (T*)R->F, nullptr, nullptr, nullptr,
                         nullptr, nullptr, R->first_fail);
    prof::end(st);
    CIA_HIP(hipGetLastError());
    return CIA_OK;
  }

  static cia_status gather(cia_rve* R, hipStream_t st, double* r, double* class0, int slot_norm) {
    prof::begin(prof::GATHER, st);
    const unsigned g = grid_for(R->C);
    hipLaunchKernelGGL(k_gather<T>, dim3(g), dim3(kBlock), 0, st, R->G, (const T*)R->F, r, class0,
                       R->partials);
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kBlock), 0, st, (const double*)R->partials,
                       (int)g, &R->sc->norm2[slot_norm]);
    prof::end(st);
    CIA_HIP(hipGetLastError());
    return CIA_OK;
  }

  // Solve K x = b with the spectral reference-medium preconditioner.
  static cia_status pcg_spectral(cia_rve* R, hipStream_t st, int slot, const double* b, double* x,
                                 int* iters_out, int* flag_out, double rtol, int hint,
                                 double* x_norm2) {
    const int64_t C = R->C;
    const unsigned gn = grid_for(C), ge = grid_for(R->np), gel = grid_for(R->E);
    const int nx = R->nx, ny = R->ny, nz = R->nz;
    const int64_t n_lines = (int64_t)ny * nz;
    const unsigned g_xl = (unsigned)((n_lines + xline_group(nx) - 1) / xline_group(nx));
    // y and z edges of 16, 32, 48 or 64 take the row-lane passes; other edges the
    // LDS passes (z-column tiles narrowed until the pass has enough workgroups)
    const int My = (ny % 16 == 0 && ny <= 64) ? ny / 16 : 0;
    const int Mz = (nz % 16 == 0 && nz <= 64) ? nz / 16 : 0;
    const int Mx = (nx % 16 == 0 && nx <= 64) ? nx / 16 : 0;
    const bool lanes = My > 0 && Mz > 0;
    const bool xrow = lanes && Mx > 0;  // x passes in row-lane form too
    const unsigned g_xr = (unsigned)((n_lines + 15) / 16);
    const int p_upd = xrow ? (int)g_xr : (int)g_xl;  // update partials per row
    int tile_w = kZTile;
    while (tile_w > 2 && (int64_t)((nx + tile_w - 1) / tile_w) * ny < 512) tile_w /= 2;
    const unsigned g_zc = lanes ? (unsigned)(((nx + kZlineKx - 1) / kZlineKx) * ny)
                                : (unsigned)(((nx + tile_w - 1) / tile_w) * ny);
    const unsigned g_yl = (unsigned)(lane_tiles_x(nx) * nz);
    const double inv_n = 1.0 / (double)C;
    // block partials: p.y (node pass), (r.r, sum r) (x-line pass), r.z (z-column pass)
    double* part_pq = R->partials;
    double* part_upd = R->partials + gn;
    double* part_rz = part_upd + 4 * g_xl;
    const size_t lds_xl = xline_lds_bytes(nx), lds_zc = zcol_lds_bytes(nz, tile_w);
    const size_t lds_y = 2 * (size_t)kDftTile * (ny + 1) * sizeof(double2);
    const dim3 g_y((unsigned)((nx + kDftTile - 1) / kDftTile), (unsigned)nz, 3);
    double2* S = (double2*)R->uab;
    const int32_t* dflag = &R->sc->done;
    const double atol = R->lin.absolute_tolerance;
    auto lane_passes = [&](auto my, auto mz, hipStream_t s, int first) {
      constexpr int MY = decltype(my)::value, MZ = decltype(mz)::value;
      hipLaunchKernelGGL((k_sp_ylines<MY, false>), dim3(g_yl + 1), dim3(kLaneBlock), 0, s, nx, ny,
                         nz, C, R->ay, S, R->sc, (const double*)part_upd, p_upd, first, rtol,
                         atol);
      hipLaunchKernelGGL((k_sp_zlines<MZ, kZlineKx>), dim3(g_zc), dim3(16 * kZlineKx), 0, s, nx,
                         ny, nz, C,
                         inv_n, R->az, (const hipFloatComplex*)R->kinv, S, part_rz, R->sc);
      hipLaunchKernelGGL((k_sp_ylines<MY, true>), dim3(g_yl + 1), dim3(kLaneBlock), 0, s, nx, ny,
                         nz, C, R->ay, S, R->sc, (const double*)part_rz, (int)g_zc, 0, rtol, atol);
    };
    auto lane_z = [&](auto my, hipStream_t s, int first) {
      switch (Mz) {
        case 1: lane_passes(my, std::integral_constant<int, 1>{}, s, first); break;
        case 2: lane_passes(my, std::integral_constant<int, 2>{}, s, first); break;
        case 3: lane_passes(my, std::integral_constant<int, 3>{}, s, first); break;
        default: lane_passes(my, std::integral_constant<int, 4>{}, s, first); break;
      }
    };
    auto precond = [&](hipStream_t s, int first) {
      if (lanes) {
        switch (My) {
          case 1: lane_z(std::integral_constant<int, 1>{}, s, first); break;
          case 2: lane_z(std::integral_constant<int, 2>{}, s, first); break;
          case 3: lane_z(std::integral_constant<int, 3>{}, s, first); break;
          default: lane_z(std::integral_constant<int, 4>{}, s, first); break;
        }
        switch (xrow ? Mx : 0) {
          case 1: hipLaunchKernelGGL(k_sp_xlines_inv<1>, dim3(g_xr), dim3(256), 0, s, nx, n_lines, C,
                                     R->ax, (const double2*)S, R->pz, (const Scalars*)R->sc); break;
          case 2: hipLaunchKernelGGL(k_sp_xlines_inv<2>, dim3(g_xr), dim3(256), 0, s, nx, n_lines, C,
                                     R->ax, (const double2*)S, R->pz, (const Scalars*)R->sc); break;
          case 3: hipLaunchKernelGGL(k_sp_xlines_inv<3>, dim3(g_xr), dim3(256), 0, s, nx, n_lines, C,
                                     R->ax, (const double2*)S, R->pz, (const Scalars*)R->sc); break;
          case 4: hipLaunchKernelGGL(k_sp_xlines_inv<4>, dim3(g_xr), dim3(256), 0, s, nx, n_lines, C,
                                     R->ax, (const double2*)S, R->pz, (const Scalars*)R->sc); break;
          default:
            hipLaunchKernelGGL(k_sp_xinv, dim3(g_xl), dim3(kBlock), lds_xl, s, nx, n_lines, R->ax,
                               (const double2*)S, C, R->pz, (const Scalars*)R->sc);
        }
        return;
      }
      hipLaunchKernelGGL((k_dft_strided<false, false>), g_y, dim3(256), lds_y, s, nx, ny, nz, C,
                         R->ay, S, dflag);
      hipLaunchKernelGGL(k_sp_zcol, dim3(g_zc), dim3(kBlock), lds_zc, s, nx, ny, nz, tile_w, inv_n,
                         R->az, (const hipFloatComplex*)R->kinv, S, C, (const double*)part_upd,
                         (int)g_xl, part_rz, R->sc, first, rtol, R->lin.absolute_tolerance);
      hipLaunchKernelGGL((k_dft_strided<true, false>), g_y, dim3(256), lds_y, s, nx, ny, nz, C,
                         R->ay, S, dflag);
      hipLaunchKernelGGL(k_sp_xinv, dim3(g_xl), dim3(kBlock), lds_xl, s, nx, n_lines, R->ax,
                         (const double2*)S, C, R->pz, (const Scalars*)R->sc);
    };
    // forward x pass with the vector update (INIT: r = b, x = 0, p_old = 0)
    auto xfwd = [&](auto init, hipStream_t s, const double* p_new, double* p_old) {
      constexpr bool I = decltype(init)::value;
      const double* bb = I ? b : nullptr;
      const double* yy = I ? nullptr : (const double*)R->py;
#define CIA_XROW(MM)                                                                            \
  hipLaunchKernelGGL((k_sp_xlines_fwd<MM, I>), dim3(g_xr), dim3(256), 0, s, nx, n_lines, R->ax,  \
                     bb, p_new, yy, x, R->pr, p_old, S, C, (const double*)part_pq, (int)gn,      \
                     part_upd, R->sc)
      switch (xrow ? Mx : 0) {
        case 1: CIA_XROW(1); break;
        case 2: CIA_XROW(2); break;
        case 3: CIA_XROW(3); break;
        case 4: CIA_XROW(4); break;
        default:
          hipLaunchKernelGGL(k_sp_xfwd<I>, dim3(g_xl), dim3(kBlock), lds_xl, s, nx, n_lines, R->ax,
                             bb, p_new, yy, x, R->pr, p_old, S, C, (const double*)part_pq, (int)gn,
                             part_upd, R->sc);
      }
#undef CIA_XROW
    };
    prof::begin(prof::PRECOND, st);
    xfwd(std::true_type{}, st, (const double*)nullptr, R->pp[1]);
    precond(st, 1);
    prof::end(st);
    CIA_HIP(hipGetLastError());
    int done = 0, it = 0;
    // One PCG iteration; the iteration index lives on the device (k_sp_elem reads
    // the completed count, k_sp_xfwd advances it), so a launch sequence does
    // not depend on it and two iterations (p_old / p_new swapped) form a graph.
    auto iteration = [&](hipStream_t s, int pi, int tag) {
      double* p_old = R->pp[1 - pi];
      double* p_new = R->pp[pi];
      prof::begin(prof::PCG_ELEM, s, tag);
      if (lanes)
        hipLaunchKernelGGL((k_sp_elem<T, true>), dim3(gel), dim3(kBlock), 0, s, R->G, R->d_tab,
                           R->d_ephase, (const T*)R->tan[slot], (const double*)R->pz, p_old, p_new,
                           (T*)R->F, (const double*)part_rz, (int)g_zc, R->sc);
      else
        hipLaunchKernelGGL((k_sp_elem<T, false>), dim3(gel), dim3(kBlock), 0, s, R->G, R->d_tab,
                           R->d_ephase, (const T*)R->tan[slot], (const double*)R->pz, p_old, p_new,
                           (T*)R->F, (const double*)part_rz, (int)g_zc, R->sc);
      prof::end(s);
      prof::begin(prof::PCG_NODE, s, tag);
      hipLaunchKernelGGL(k_sp_node<T>, dim3(gn), dim3(kBlock), 0, s, R->G, (const T*)R->F, p_new,
                         R->py, part_pq, R->sc);
      prof::end(s);
      prof::begin(prof::PCG_UPDATE, s, tag);
      xfwd(std::false_type{}, s, (const double*)p_new, (double*)nullptr);
      prof::end(s);
      prof::begin(prof::PRECOND, s, tag);
      precond(s, 0);
      prof::end(s);
    };
    hipGraphExec_t pair = nullptr;
    if (!prof::g.on && R->cap_stream) CIA_TRY(pcg_pair_graph(R, slot, x, iteration, &pair));
    // Launches after convergence exit at once, so the first batch is sized by the
    // iteration count of the same Newton index in the previous increment; later
    // batches are short.  Graph batches are whole pairs of iterations.
    const int interval = std::max(1, R->lin.check_interval);
    const bool hinted = hint >= 0 && hint < cia_rve::kHints && R->its_hint[hint] > 0;
    int chunk = hinted ? R->its_hint[hint] : interval;
    Scalars hs;
    while (it < R->lin.max_iterations) {
      int n = std::min(chunk, R->lin.max_iterations - it);
      chunk = hinted ? std::max(1, interval / 4) : interval;
      if (pair) {
        // whole pairs; an odd final iteration below the limit runs on its own
        int pairs = std::min((n + 1) / 2, (R->lin.max_iterations - it) / 2);
        for (int s = 0; s < pairs; ++s, it += 2) CIA_HIP(hipGraphLaunch(pair, st));
        if (pairs == 0) {
          iteration(st, it & 1, it);
          ++it;
        }
      } else {
        for (int s = 0; s < n; ++s, ++it) iteration(st, it & 1, it);
      }
      CIA_HIP(hipGetLastError());
      CIA_TRY(norm2_async(R, st, 3 * C, x, 3));  // ||x||^2 read with the convergence flag
      CIA_TRY(fetch_scalars(R, st, &hs));
      done = hs.done;
      prof::flush(done ? hs.iters : 1 << 30);
      if (done) break;
    }
    if (!done) {
      CIA_TRY(fetch_scalars(R, st, &hs));
    }
    *iters_out = hs.iters;
    *flag_out = hs.done;
    if (x_norm2) *x_norm2 = hs.norm2[3];
    R->pcg_total += hs.iters;
    if (hint >= 0 && hint < cia_rve::kHints && hs.done == 1) R->its_hint[hint] = hs.iters;
    return CIA_OK;
  }

  // Solve K x = b (K from tan[slot]); returns iterations.  `hint` indexes the
  // per-Newton-iteration history of PCG counts (-1: none).
  static cia_status pcg(cia_rve* R, hipStream_t st, int slot, const double* b, double* x,
                        int* iters_out, int* flag_out, double rtol = -1.0, int hint = -1,
                        double* x_norm2 = nullptr) {
    if (rtol <= 0.0) rtol = R->lin.relative_tolerance;
    if (R->prec == CIA_PCG_SPECTRAL && R->fft_ready)
      return pcg_spectral(R, st, slot, b, x, iters_out, flag_out, rtol, hint, x_norm2);
    const int64_t C = R->C;
    const unsigned gn = grid_for(C), ge = grid_for(R->np);
    prof::begin(prof::PRECOND, st);
    hipLaunchKernelGGL(k_precond<T>, dim3(gn), dim3(kBlock), 0, st, R->G, R->d_tab, R->d_ephase,
                       (const T*)R->tan[slot], R->minv);
    prof::end(st);
    hipLaunchKernelGGL(k_pcg_init, dim3(gn), dim3(kBlock), 0, st, C, b, R->minv, x, R->pr, R->pz,
                       R->pp[1], rtol, R->lin.absolute_tolerance, R->sc,
                       R->partials, R->counter);
    CIA_HIP(hipGetLastError());
    int done = 0;
    int it = 0;
    int pi = 0;  // pp[1 - pi] is p_old at the next element pass
    const int chunk = std::max(1, R->lin.check_interval);
    Scalars hs;
    while (it < R->lin.max_iterations) {
      const int n = std::min(chunk, R->lin.max_iterations - it);
      for (int s = 0; s < n; ++s) {
        double* p_old = R->pp[1 - pi];
        double* p_new = R->pp[pi];
        prof::begin(prof::PCG_ELEM, st);
        hipLaunchKernelGGL((k_pcg_elem<T, false>), dim3(ge), dim3(kBlock), 0, st, R->G, R->d_tab,
                           R->d_ephase, (const T*)R->tan[slot], R->pz, p_old, p_new, (T*)R->F,
                           R->sc);
        prof::end(st);
        prof::begin(prof::PCG_NODE, st);
        hipLaunchKernelGGL(k_pcg_node<T>, dim3(gn), dim3(kBlock), 0, st, R->G, (const T*)R->F,
                           p_new, R->py, R->sc, R->partials, R->counter);
        prof::end(st);
        prof::begin(prof::PCG_UPDATE, st);
        hipLaunchKernelGGL(k_pcg_update, dim3(gn), dim3(kBlock), 0, st, C, R->minv, p_new, R->py,
                           x, R->pr, R->pz, R->sc, R->partials, R->counter);
        prof::end(st);
        pi = 1 - pi;
      }
      CIA_HIP(hipGetLastError());
      it += n;
      CIA_TRY(fetch_scalars(R, st, &hs));
      prof::flush();
      done = hs.done;
      if (done) break;
    }
    if (!done) {
      CIA_TRY(fetch_scalars(R, st, &hs));
    }
    *iters_out = hs.iters;
    *flag_out = hs.done;
    R->pcg_total += hs.iters;
    if (x_norm2) {
      CIA_TRY(norm2_async(R, st, 3 * C, x, 3));
      Scalars h2;
      CIA_TRY(fetch_scalars(R, st, &h2));
      *x_norm2 = h2.norm2[3];
    }
    return CIA_OK;
  }
};

static cia_status read_norms(cia_rve* R, hipStream_t st, double out[4]) {
  Scalars hs;
  CIA_TRY(fetch_scalars(R, st, &hs));
  prof::flush();
  for (int i = 0; i < 4; ++i) out[i] = hs.norm2[i];
  return CIA_OK;
}

// ||x||^2 -> sc->norm2[slot] without a host synchronisation.
static cia_status norm2_async(cia_rve* R, hipStream_t st, int64_t n, const double* x, int slot) {
  const unsigned g = (unsigned)std::min<int64_t>(grid_for(n), 1024);
  hipLaunchKernelGGL(k_norms, dim3(g), dim3(kBlock), 0, st, n, x, (const double*)nullptr,
                     R->norm_part);
  hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kBlock), 0, st, (const double*)R->norm_part,
                     (int)g, &R->sc->norm2[slot]);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

// One synchronisation: the scalar block (norms) and the first failing point.
static cia_status read_norms_fail(cia_rve* R, hipStream_t st, double out[4],
                                  unsigned long long* ff) {
  Scalars hs;
  CIA_TRY(fetch_scalars(R, st, &hs, ff));
  prof::flush();
  for (int i = 0; i < 4; ++i) out[i] = hs.norm2[i];
  return CIA_OK;
}

static cia_status norm2_of(cia_rve* R, hipStream_t st, int64_t n, const double* x, double* out) {
  const unsigned g = (unsigned)std::min<int64_t>(grid_for(n), 1024);
  hipLaunchKernelGGL(k_norms, dim3(g), dim3(kBlock), 0, st, n, x, (const double*)nullptr,
                     R->norm_part);
  hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kBlock), 0, st, (const double*)R->norm_part,
                     (int)g, &R->sc->norm2[2]);
  CIA_HIP(hipGetLastError());
  double v[4];
  CIA_TRY(read_norms(R, st, v));
  *out = v[2];
  return CIA_OK;
}

static cia_status check_fail(cia_rve* R, hipStream_t st, unsigned long long* ff) {
  CIA_HIP(hipMemcpyAsync(&R->hst->ff, R->first_fail, sizeof(*ff), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  *ff = R->hst->ff;
  return CIA_OK;
}

template <class T>
static cia_status solve_impl(cia_rve* R, hipStream_t st, const void* state_c, const double* cmacro,
                             const double* cw, const double* macro, double dt,
                             const Settings& S, int need_tangent, void* next_state,
                             double* next_w, cia_rve_result* res) {
  const int64_t C3 = 3 * R->C;
  Macro M;
  for (int m = 0; m < 6; ++m) {
    M.e[m] = macro[m];
    M.ec[m] = cmacro[m];
  }
  R->pcg_total = 0;
  R->have_fields = false;
  res->gp_updates = 0;
  CIA_HIP(hipMemcpyAsync(R->w, cw, C3 * sizeof(double), hipMemcpyDeviceToDevice, st));
  CIA_HIP(hipMemcpyAsync(R->wc_buf, cw, C3 * sizeof(double), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, R->first_fail, ~0ull);
  const cia_newton_settings& ns = R->newton;
  int cur = 0;
  unsigned long long ff = 0;

  auto point_failure = [&](const double* wv) -> cia_status {
    hipLaunchKernelGGL(k_diagnose<T>, dim3(1), dim3(1), 0, st, R->G, R->d_tab, R->d_ephase, S, dt,
                       wv, R->wc_buf, M, (const T*)state_c, (int64_t)ff, R->rec);
    CIA_HIP(hipGetLastError());
    CIA_HIP(hipMemcpyAsync(&res->point_failure, R->rec, sizeof(cia_point_failure),
                           hipMemcpyDeviceToHost, st));
    CIA_HIP(hipStreamSynchronize(st));
    res->reason = CIA_RVE_MATERIAL_POINT_FAILURE;
    return CIA_ERR_POINT_FAILURE;
  };

  // assembly at the committed fluctuation (iteration 0); ||w||^2 -> norm2[2]
  CIA_TRY(Ops<T>::assemble(R, st, false, cur, R->w, R->wc_buf, M, state_c, nullptr, S, dt));
  res->gp_updates += R->np;
  CIA_TRY(Ops<T>::gather(R, st, R->r, nullptr, 0));
  CIA_TRY(norm2_async(R, st, C3, R->w, 2));
  double nv[4];
  CIA_TRY(read_norms_fail(R, st, nv, &ff));
  if (ff != ~0ull) return point_failure(R->w);
  double rn = sqrt(nv[0]);
  double wn2 = nv[2];
  double r0 = -1.0;
  double last_inc = INFINITY;
  int backtracks = 0;
  int it = 0;
  double rtol = 0.0;
  for (it = 0; it <= ns.max_iterations; ++it) {
    if (r0 < 0.0) r0 = rn;
    rtol = std::max(ns.residual_absolute_tolerance, ns.residual_relative_tolerance * r0);
    const double itol = std::max(ns.increment_absolute_tolerance,
                                 ns.increment_relative_tolerance * std::max(sqrt(wn2), 1.0));
    if (rn <= rtol && (it == 0 || last_inc <= itol)) break;
    if (it == ns.max_iterations) {
      res->reason = CIA_RVE_MAXIMUM_ITERATIONS_EXCEEDED;
      res->failure_iteration = it;
      res->failure_residual_norm = rn;
      res->failure_residual_tolerance = rtol;
      res->failure_increment_norm = last_inc;
      res->failure_increment_tolerance = itol;
      res->newton_iterations = it;
      res->linear_iterations = R->pcg_total;
      return CIA_ERR_RVE_FAILURE;
    }
    // Newton direction: K dw = -r
    hipLaunchKernelGGL(k_scale, dim3(grid_for(C3)), dim3(kBlock), 0, st, C3, -1.0, R->r, R->r_cand);
    int pit = 0, pflag = 0;
    double lin_rtol = R->lin.relative_tolerance;
    if (R->lin.adaptive_forcing) {
      // inexact Newton: solve only as accurately as the current residual warrants
      const double eta = std::min(R->lin.forcing_max > 0 ? R->lin.forcing_max : 0.1, rn / r0);
      lin_rtol = std::max(lin_rtol, std::max(eta, 0.5 * rtol / std::max(rn, 1e-300)));
      lin_rtol = std::min(lin_rtol, 0.5);
    }
    double dn2 = 0.0;
    CIA_TRY(Ops<T>::pcg(R, st, cur, R->r_cand, R->dw, &pit, &pflag, lin_rtol, it, &dn2));
    const double inc_n = sqrt(dn2);
    if (!std::isfinite(inc_n)) {
      res->reason = CIA_RVE_NON_FINITE_INCREMENT;
      res->failure_iteration = it;
      res->failure_residual_norm = rn;
      res->failure_residual_tolerance = rtol;
      res->failure_increment_norm = inc_n;
      res->failure_increment_tolerance = itol;
      return CIA_ERR_RVE_FAILURE;
    }
    if (rn <= rtol && inc_n <= itol) break;
    if (ns.armijo_enabled) {
      const int cand = 1 - cur;
      bool accepted = false;
      double alpha = 1.0;
      for (int bt = 0; bt <= ns.armijo_max_backtracks; ++bt) {
        alpha = std::pow(ns.armijo_reduction_factor, (double)bt);
        hipLaunchKernelGGL(k_axpy, dim3(grid_for(C3)), dim3(kBlock), 0, st, C3, R->w, alpha, R->dw,
                           R->w_cand);
        CIA_TRY(norm2_async(R, st, C3, R->w_cand, 2));
        CIA_TRY(Ops<T>::assemble(R, st, false, cand, R->w_cand, R->wc_buf, M, state_c,
                                 nullptr, S, dt));
        res->gp_updates += R->np;
        CIA_TRY(Ops<T>::gather(R, st, R->r_cand, nullptr, 1));
        CIA_TRY(read_norms_fail(R, st, nv, &ff));
        if (ff != ~0ull) return point_failure(R->w_cand);
        const double cn = sqrt(nv[1]);
        if (0.5 * cn * cn <= 0.5 * rn * rn - ns.armijo_sufficient_decrease * alpha * rn * rn) {
          std::swap(R->w, R->w_cand);
          std::swap(R->r, R->r_cand);
          cur = cand;
          rn = cn;
          wn2 = nv[2];
          last_inc = alpha * inc_n;
          backtracks += bt;
          accepted = true;
          break;
        }
      }
      if (!accepted) {
        res->reason = CIA_RVE_ARMIJO_EXHAUSTED;
        res->failure_iteration = it;
        res->failure_residual_norm = rn;
        res->failure_residual_tolerance = rtol;
        res->failure_increment_norm = inc_n;
        res->failure_increment_tolerance = itol;
        return CIA_ERR_RVE_FAILURE;
      }
    } else {
      hipLaunchKernelGGL(k_axpy, dim3(grid_for(C3)), dim3(kBlock), 0, st, C3, R->w, 1.0, R->dw,
                         R->w_cand);
      std::swap(R->w, R->w_cand);
      CIA_TRY(norm2_async(R, st, C3, R->w, 2));
      CIA_TRY(Ops<T>::assemble(R, st, false, cur, R->w, R->wc_buf, M, state_c, nullptr, S, dt));
      res->gp_updates += R->np;
      CIA_TRY(Ops<T>::gather(R, st, R->r, nullptr, 0));
      CIA_TRY(read_norms_fail(R, st, nv, &ff));
      if (ff != ~0ull) return point_failure(R->w);
      rn = sqrt(nv[0]);
      wn2 = nv[2];
      last_inc = inc_n;
    }
  }
  res->reason = CIA_RVE_OK;
  res->newton_iterations = it;
  res->armijo_backtracks = backtracks;
  res->micro_residual_norm = rn;
  res->micro_residual_tolerance = rtol;
  R->cur = cur;

  // ---- converged: full evaluation at w (energies, diagnostics, next state)
  CIA_TRY(Ops<T>::assemble(R, st, true, cur, R->w, R->wc_buf, M, state_c, next_state, S, dt));
  CIA_TRY(Ops<T>::gather(R, st, R->r, R->class0, 0));
  CIA_HIP(hipMemcpyAsync(next_w, R->w, C3 * sizeof(double), hipMemcpyDeviceToDevice, st));
  {
    const unsigned ge = grid_for(R->E);
    double* part = R->red;
    double* mm = R->red + (int64_t)R_N * ge;
#define CIA_RESPONSE(NPH)                                                                    \
  hipLaunchKernelGGL((k_response<T, NPH>), dim3(ge), dim3(kBlock), 0, st, R->G, R->d_ephase, dt, \
                     (const T*)R->sig[cur], (const T*)R->rate_out, (const T*)R->psi_out,         \
                     (const T*)R->diss_out, (const T*)R->diag_out, (const T*)R->F, part, mm)
    if (R->n_phases <= 1)
      CIA_RESPONSE(1);
    else if (R->n_phases <= 2)
      CIA_RESPONSE(2);
    else if (R->n_phases <= 4)
      CIA_RESPONSE(4);
    else
      CIA_RESPONSE(8);
#undef CIA_RESPONSE
    CIA_HIP(hipGetLastError());
    const unsigned gfull = grid_for(R->N);
    double* np_part = mm + 2 * ge;
    double* np_max = np_part + (int64_t)N_N * gfull;
    double* cl_max = np_max + gfull;
    hipLaunchKernelGGL(k_full_nodes<T>, dim3(gfull), dim3(kBlock), 0, st, R->G, (const T*)R->F,
                       R->w, M, R->u_full, R->f_full, np_part, np_max);
    const unsigned gc = grid_for(R->C);
    hipLaunchKernelGGL(k_class_max<T>, dim3(gc), dim3(kBlock), 0, st, R->G, (const T*)R->F, cl_max);
    CIA_HIP(hipGetLastError());
    const int64_t total = (int64_t)R_N * ge + 2 * ge + (int64_t)N_N * gfull + gfull + gc;
    const int nrows = R_PHASE + 9 * std::min(R->n_phases, CIA_MAX_PHASES);
    double* dout = R->red + total;
    hipLaunchKernelGGL(k_response_rows, dim3(kRespOut), dim3(kBlock), 0, st, part, (int)ge, nrows,
                       mm, np_part, (int)gfull, np_max, cl_max, (int)gc, dout);
    CIA_HIP(hipGetLastError());
    const double* h = R->hst->resp;
    CIA_HIP(hipMemcpyAsync(R->hst->resp, dout, kRespOut * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    CIA_HIP(hipStreamSynchronize(st));
    const double* sums = h;
    const double mn = h[R_N], mx = h[R_N + 1];
    const double* nsum = h + R_N + 2;
    const double flmax = h[R_N + 2 + N_N], clmax = h[R_N + 3 + N_N];

    const double jw = R->G.g.jw, V = R->volume;
    for (int m = 0; m < 6; ++m) res->macro_stress[m] = sums[R_SIG + m] * jw / V;
    res->free_energy_density = sums[R_PSI] * jw / V;
    res->dissipation_density = sums[R_DISS] * jw / V;
    res->volume = V;
    double vdiff = 0.0;
    for (int m = 0; m < 6; ++m) {
      const double d = sums[R_VIRTUAL + m] / V - res->macro_stress[m];
      vdiff += d * d;
    }
    res->virtual_work_stress_error = sqrt(vdiff);
    const double micro = sums[R_MICRO_WORK] * jw / V;
    double macro_w = 0.0;
    for (int m = 0; m < 6; ++m) macro_w += res->macro_stress[m] * (macro[m] - cmacro[m]);
    res->macro_work_increment = macro_w;
    res->micro_work_increment = micro;
    res->hill_mandel_residual = macro_w - micro;
    res->hill_mandel_error =
        fabs(macro_w - micro) / (fabs(macro_w) + fabs(micro) + R->hm_scale * dt);
    res->positive_work_density = sums[R_POS_WORK] * jw / V;
    res->negative_work_density = sums[R_NEG_WORK] * jw / V;
    res->unloading_volume_fraction = sums[R_UNLOAD] * jw / V;
    res->viscoplastic_active_volume_fraction = sums[R_ACTIVE] * jw / V;
    res->yield_activation_margin_available = sums[R_MARGIN_N] > 0.0 ? 1 : 0;
    res->minimum_yield_activation_margin = sums[R_MARGIN_N] > 0.0 ? mn : NAN;
    res->maximum_yield_activation_margin = sums[R_MARGIN_N] > 0.0 ? mx : NAN;
    res->periodic_fluctuation_error = flmax;
    double face_err = 0.0;
    for (int ax = 0; ax < 3; ++ax) {
      double a2 = 0.0, b2 = 0.0, s2 = 0.0;
      for (int m = 0; m < 3; ++m) {
        const double a = nsum[N_FACE + 6 * ax + m], b = nsum[N_FACE + 6 * ax + 3 + m];
        a2 += a * a;
        b2 += b * b;
        s2 += (a + b) * (a + b);
      }
      const double den = std::max(sqrt(a2) + sqrt(b2), 1.0);
      face_err = std::max(face_err, sqrt(s2) / den);
    }
    res->reaction_antiperiodicity_error = face_err;
    res->periodic_class_reaction_error = clmax / std::max(sqrt(nsum[N_F2]), 1.0);
    for (int p = 0; p < R->n_phases && p < 8; ++p) {
      const double cnt = sums[R_PHASE + 9 * p];
      const double pv = cnt * jw;
      res->phase_volume[p] = pv;
      for (int m = 0; m < 6; ++m)
        res->phase_average_stress[p][m] = pv > 0 ? sums[R_PHASE + 9 * p + 1 + m] * jw / pv : NAN;
      res->phase_free_energy_density[p] = pv > 0 ? sums[R_PHASE + 9 * p + 7] * jw / pv : NAN;
      res->phase_dissipation_density[p] = pv > 0 ? sums[R_PHASE + 9 * p + 8] * jw / pv : NAN;
    }
  }
  R->have_fields = true;

  // ---- effective tangent (Schur complement), solver.py:438-461
  if (need_tangent) {
    const unsigned ge = grid_for(R->E);
    std::vector<double> h(6 * ge);
    for (int j = 0; j < 6; ++j) {
      hipLaunchKernelGGL(k_affine_force<T>, dim3(ge), dim3(kBlock), 0, st, R->G, R->d_tab,
                         R->d_ephase, (const T*)R->tan[cur], j, (T*)R->F);
      CIA_TRY(Ops<T>::gather(R, st, R->r_cand, nullptr, 1));
      int pit = 0, pflag = 0;
      CIA_TRY(Ops<T>::pcg(R, st, cur, R->r_cand, R->sens, &pit, &pflag));
      double sn2 = 0.0;
      CIA_TRY(norm2_of(R, st, C3, R->sens, &sn2));
      if (!std::isfinite(sn2)) {
        res->reason = CIA_RVE_NON_FINITE_EFFECTIVE_TANGENT_SENSITIVITY;
        return CIA_ERR_RVE_FAILURE;
      }
      hipLaunchKernelGGL(k_ceff_column<T>, dim3(ge), dim3(kBlock), 0, st, R->G, R->d_tab,
                         R->d_ephase, (const T*)R->tan[cur], R->sens, j, R->red);
      CIA_HIP(hipGetLastError());
      CIA_HIP(hipMemcpyAsync(h.data(), R->red, 6 * ge * sizeof(double), hipMemcpyDeviceToHost, st));
      CIA_HIP(hipStreamSynchronize(st));
      for (int m = 0; m < 6; ++m) {
        double a = 0.0;
        for (unsigned b = 0; b < ge; ++b) a += h[(int64_t)m * ge + b];
        res->effective_tangent[m * 6 + j] = a * R->G.g.jw / R->volume;
      }
    }
    // The F rows now hold affine loads; the final internal force is kept in
    // f_full, which is what CIA_FIELD_INTERNAL_FORCE returns.
  }
  res->linear_iterations = R->pcg_total;
  return CIA_OK;
}

}  // namespace cia

extern "C" {

cia_status cia_rve_create(const cia_rve_desc* d, cia_rve** out) {
  using namespace cia;
  CIA_ARG(d && out, "desc/out is NULL");
  CIA_ARG(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "divisions must be positive");
  CIA_ARG(d->n_phases >= 1 && d->n_phases <= CIA_MAX_PHASES && d->phases, "invalid phases");
  CIA_ARG(d->lengths[0] > 0 && d->lengths[1] > 0 && d->lengths[2] > 0, "lengths must be positive");
  CIA_ARG(d->precision == CIA_PREC_FP64 || d->precision == CIA_PREC_FP32_REFINED, "bad precision");
  auto* R = new cia_rve();
  R->nx = d->nx;
  R->ny = d->ny;
  R->nz = d->nz;
  R->E = (int64_t)d->nx * d->ny * d->nz;
  R->C = R->E;
  R->np = 8 * R->E;
  R->N = (int64_t)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
  R->G.nx = d->nx;
  R->G.ny = d->ny;
  R->G.nz = d->nz;
  R->G.E = R->E;
  const double hx = d->lengths[0] / d->nx, hy = d->lengths[1] / d->ny, hz = d->lengths[2] / d->nz;
  R->G.g.ihx = 1.0 / hx;
  R->G.g.ihy = 1.0 / hy;
  R->G.g.ihz = 1.0 / hz;
  R->G.g.jw = hx * hy * hz / 8.0;
  R->volume = R->G.g.jw * (double)R->np;
  R->precision = d->precision;
  R->tsize = d->precision == CIA_PREC_FP64 ? sizeof(double) : sizeof(float);
  R->n_phases = d->n_phases;
  R->newton = d->newton;
  R->lin = d->linear;
  if (R->lin.max_iterations <= 0) R->lin.max_iterations = 20000;
  if (R->lin.check_interval <= 0) R->lin.check_interval = 16;
  R->hm_scale = d->hill_mandel_power_scale;
  PhaseTable tab{};
  tab.n = d->n_phases;
  for (int i = 0; i < d->n_phases; ++i) {
    tab.p[i] = make_phase(d->phases[i]);
    R->phases.push_back(tab.p[i]);
  }
  std::vector<uint8_t> eph(R->E, 0);
  if (d->element_phase) {
    for (int64_t e = 0; e < R->E; ++e) {
      const int32_t v = d->element_phase[e];
      if (v < 0 || v >= d->n_phases) {
        delete R;
        set_error_msg("element_phase out of range");
        return CIA_ERR_INVALID_ARGUMENT;
      }
      eph[e] = (uint8_t)v;
    }
  }
  const int64_t np = R->np, C3 = 3 * R->C;
  const size_t ts = R->tsize;
  cia_status s = CIA_OK;
#define ALLOC(ptr, bytes)                                   \
  do {                                                      \
    s = dmalloc((void**)&(ptr), (bytes));                   \
    if (s != CIA_OK) {                                      \
      cia_rve_destroy(R);                                   \
      return s;                                             \
    }                                                       \
  } while (0)
  ALLOC(R->d_tab, sizeof(PhaseTable));
  ALLOC(R->d_ephase, R->E);
  for (int k = 0; k < 2; ++k) {
    ALLOC(R->sig[k], 6 * np * ts);
    ALLOC(R->tan[k], 8 * np * ts);
  }
  ALLOC(R->F, 24 * R->E * ts);
  ALLOC(R->strain_out, 6 * np * ts);
  ALLOC(R->rate_out, 6 * np * ts);
  ALLOC(R->psi_out, np * ts);
  ALLOC(R->diss_out, np * ts);
  ALLOC(R->diag_out, (int64_t)CIA_N_DIAG * np * ts);
  ALLOC(R->w, C3 * 8);
  ALLOC(R->w_cand, C3 * 8);
  ALLOC(R->dw, C3 * 8);
  ALLOC(R->r, C3 * 8);
  ALLOC(R->r_cand, C3 * 8);
  ALLOC(R->wc_buf, C3 * 8);
  ALLOC(R->pr, C3 * 8);
  ALLOC(R->pz, C3 * 8);
  ALLOC(R->pp[0], C3 * 8);
  ALLOC(R->pp[1], C3 * 8);
  ALLOC(R->py, C3 * 8);
  ALLOC(R->sens, C3 * 8);
  ALLOC(R->minv, 6 * R->C * 8);
  ALLOC(R->u_full, 3 * R->N * 8);
  ALLOC(R->f_full, 3 * R->N * 8);
  const int64_t maxb = std::max<int64_t>(grid_for(R->E), grid_for(R->N)) + 64;
  // spectral PCG partials: p.y per class block, 4 per x-line group, r.z per z-column tile
  const int64_t sp_parts = grid_for(R->E) + 4 * ((int64_t)d->ny * d->nz / xline_group(d->nx) + 1) +
                           (int64_t)d->nx * d->ny + 64;
  ALLOC(R->partials, std::max(8 * maxb, sp_parts) * 8);
  ALLOC(R->norm_part, 2 * 1024 * sizeof(double));
  ALLOC(R->counter, 64);
  // the scalar block, then first_fail at the offset the staging copy has for it
  ALLOC(R->sc, offsetof(cia_rve::Staging, ff) + sizeof(unsigned long long));
  R->first_fail = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(R->sc) +
                                                         offsetof(cia_rve::Staging, ff));
  ALLOC(R->class0, 3 * 8);
  ALLOC(R->rec, sizeof(cia_point_failure));
  if (hipHostMalloc((void**)&R->hst, sizeof(cia_rve::Staging)) != hipSuccess) {
    set_error_msg("hipHostMalloc of the read-back staging failed");
    cia_rve_destroy(R);
    return CIA_ERR_OUT_OF_MEMORY;
  }
  R->red_len = (int64_t)R_N * grid_for(R->E) + 2 * grid_for(R->E) + (int64_t)N_N * grid_for(R->N) +
               grid_for(R->N) + grid_for(R->C) + kRespOut + 64;
  ALLOC(R->red, R->red_len * 8);
  R->prec = d->linear.preconditioner;
  if (R->prec == CIA_PCG_SPECTRAL && R->C > 1) {
    double kmin = INFINITY, kmax = 0.0, gmin = INFINITY, gmax = 0.0;
    for (int i = 0; i < d->n_phases; ++i) {
      kmin = std::min(kmin, tab.p[i].K);
      kmax = std::max(kmax, tab.p[i].K);
      gmin = std::min(gmin, tab.p[i].G);
      gmax = std::max(gmax, tab.p[i].G);
    }
    R->ref_bulk = d->linear.reference_bulk > 0 ? d->linear.reference_bulk : std::sqrt(kmin * kmax);
    R->ref_shear = d->linear.reference_shear > 0 ? d->linear.reference_shear : std::sqrt(gmin * gmax);
    ALLOC(R->kinv, 6 * R->C * sizeof(hipFloatComplex));
    ALLOC(R->uab, 3 * R->C * sizeof(hipDoubleComplex));
    std::vector<hipFloatComplex> sym;
    spectral_symbol(R->nx, R->ny, R->nz, R->G.g, R->ref_bulk, R->ref_shear, sym);
    if (hipMemcpy(R->kinv, sym.data(), sym.size() * sizeof(hipFloatComplex),
                  hipMemcpyHostToDevice) != hipSuccess) {
      cia_rve_destroy(R);
      set_error_msg("hipMemcpy of the spectral symbol failed");
      return CIA_ERR_HIP;
    }
    if (R->nx > 64 || R->ny > 64 || R->nz > 64) {
      cia_rve_destroy(R);
      set_error_msg("the spectral preconditioner supports RVE edges up to 64 elements");
      return CIA_ERR_INVALID_ARGUMENT;
    }
    std::vector<double2> tw;
    const size_t ox = dft_twiddles(R->nx, tw), oy = dft_twiddles(R->ny, tw),
                 oz = dft_twiddles(R->nz, tw);
    ALLOC(R->tw, tw.size() * sizeof(double2));
    if (hipMemcpy(R->tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice) !=
        hipSuccess) {
      cia_rve_destroy(R);
      set_error_msg("hipMemcpy of the DFT twiddles failed");
      return CIA_ERR_HIP;
    }
    auto axis = [&](int n, size_t off) {
      DftAxis a{};
      a.n = n;
      dft_factor(n, a.n1, a.n2);
      a.w = R->tw + off;
      return a;
    };
    R->ax = axis(R->nx, ox);
    R->ay = axis(R->ny, oy);
    R->az = axis(R->nz, oz);
    if (hipStreamCreateWithFlags(&R->cap_stream, hipStreamNonBlocking) != hipSuccess)
      R->cap_stream = nullptr;  // graphs off: per-kernel launches
    R->fft_ready = true;
  }
#undef ALLOC
  if (hipMemcpy(R->d_tab, &tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(R->d_ephase, eph.data(), R->E, hipMemcpyHostToDevice) != hipSuccess) {
    cia_rve_destroy(R);
    set_error_msg("hipMemcpy of RVE tables failed");
    return CIA_ERR_HIP;
  }
  *out = R;
  return CIA_OK;
}

cia_status cia_rve_destroy(cia_rve* R) {
  if (!R) return CIA_OK;
  void* ptrs[] = {R->d_tab, R->d_ephase, R->sig[0], R->sig[1],
                  R->tan[0], R->tan[1], R->F, R->strain_out, R->rate_out, R->psi_out, R->diss_out,
                  R->diag_out, R->w, R->w_cand, R->dw, R->r, R->r_cand, R->wc_buf, R->pr, R->pz,
                  R->pp[0], R->pp[1], R->py, R->sens, R->minv, R->u_full, R->f_full, R->partials,
                  R->counter, R->sc, R->class0, R->rec, R->red};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (R->kinv) (void)hipFree(R->kinv);
  if (R->uab) (void)hipFree(R->uab);
  if (R->tw) (void)hipFree(R->tw);
  if (R->norm_part) (void)hipFree(R->norm_part);
  if (R->hst) (void)hipHostFree(R->hst);
  for (auto& g : R->graphs) (void)hipGraphExecDestroy(g.exec);
  if (R->cap_stream) (void)hipStreamDestroy(R->cap_stream);
  delete R;
  return CIA_OK;
}

cia_status cia_rve_sizes(const cia_rve* R, int64_t* s) {
  if (!R || !s) return CIA_ERR_INVALID_ARGUMENT;
  s[0] = R->E;
  s[1] = R->np;
  s[2] = R->C;
  s[3] = R->N;
  return CIA_OK;
}

cia_status cia_rve_solve(cia_rve* R, void* stream, const void* committed_state,
                         const double* committed_macro, const double* committed_w,
                         const double* macro_strain, double time_step,
                         const cia_update_settings* settings, int32_t need_tangent,
                         void* next_state, double* next_w, cia_rve_result* result) {
  using namespace cia;
  CIA_ARG(R && committed_state && committed_macro && committed_w && macro_strain && settings &&
              next_state && next_w && result,
          "NULL argument to cia_rve_solve");
  CIA_ARG(time_step > 0.0, "RVE time step must be positive.");
  CIA_ARG((const void*)next_state != (const void*)committed_state, "next_state aliases committed_state");
  memset(result, 0, sizeof(*result));
  result->point_failure.point_id = -1;
  const Settings S = make_settings(*settings);
  hipStream_t st = (hipStream_t)stream;
  if (R->precision == CIA_PREC_FP64)
    return solve_impl<double>(R, st, committed_state, committed_macro, committed_w, macro_strain,
                              time_step, S, need_tangent, next_state, next_w, result);
  return solve_impl<float>(R, st, committed_state, committed_macro, committed_w, macro_strain,
                           time_step, S, need_tangent, next_state, next_w, result);
}

cia_status cia_profile_enable(int32_t on) {
  cia::prof::g.on = on != 0;
  return CIA_OK;
}

cia_status cia_profile_read(double* total_ms, int64_t* counts, int32_t n) {
  using namespace cia;
  CIA_ARG(total_ms && counts && n >= 0, "NULL argument");
  CIA_HIP(hipDeviceSynchronize());
  prof::flush();
  for (int i = 0; i < n && i < prof::N_IDS; ++i) {
    total_ms[i] = prof::g.total_ms[i];
    counts[i] = prof::g.count[i];
    prof::g.total_ms[i] = 0.0;
    prof::g.count[i] = 0;
  }
  return CIA_OK;
}

cia_status cia_rve_copy_field(cia_rve* R, void* stream, int32_t field, double* dst) {
  using namespace cia;
  CIA_ARG(R && dst, "NULL argument");
  CIA_ARG(R->have_fields, "no converged increment to copy from");
  hipStream_t st = (hipStream_t)stream;
  const int64_t np = R->np;
  const void* src = nullptr;
  int64_t n = 0;
  switch (field) {
    case CIA_FIELD_MICRO_STRAIN: src = R->strain_out; n = 6 * np; break;
    case CIA_FIELD_MICRO_STRAIN_RATE: src = R->rate_out; n = 6 * np; break;
    case CIA_FIELD_MICRO_STRESS: src = R->sig[R->cur]; n = 6 * np; break;
    case CIA_FIELD_FREE_ENERGY: src = R->psi_out; n = np; break;
    case CIA_FIELD_DISSIPATION: src = R->diss_out; n = np; break;
    case CIA_FIELD_DIAGNOSTICS: src = R->diag_out; n = (int64_t)CIA_N_DIAG * np; break;
    case CIA_FIELD_DISPLACEMENT:
      CIA_HIP(hipMemcpyAsync(dst, R->u_full, 3 * R->N * 8, hipMemcpyDeviceToDevice, st));
      return CIA_OK;
    case CIA_FIELD_INTERNAL_FORCE:
      CIA_HIP(hipMemcpyAsync(dst, R->f_full, 3 * R->N * 8, hipMemcpyDeviceToDevice, st));
      return CIA_OK;
    default:
      set_error_msg("unknown field");
      return CIA_ERR_INVALID_ARGUMENT;
  }
  if (R->precision == CIA_PREC_FP64) {
    CIA_HIP(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, st));
  } else {
    hipLaunchKernelGGL(k_copy_cast<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, n,
                       (const float*)src, dst);
    CIA_HIP(hipGetLastError());
  }
  return CIA_OK;
}

}  // extern "C"
