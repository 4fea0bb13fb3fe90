// Explicit velocity-Verlet material dynamics on a structured Hex8 box.
//
// Reference: solve_material_explicit (src/fem/dynamics/material.py:89-297) with
// the energy bookkeeping of _record_step (material.py:781-848), the lumped
// kinetic energy (:851-856) and homogeneous Dirichlet zeroing (:1154-1162).
//
// One step is two kernels:
//   k_dyn_elem  one lane per Gauss point: Verlet predictor u+ = u + dt v + dt^2/2 a
//               evaluated for the 8 corners on the fly (Dirichlet dofs zeroed),
//               strain and strain rate (u+ - u)/dt, constitutive update from the
//               committed state written back in place (explicit: every step is
//               committed), element force rows F[24][E], energy partials.
//   k_dyn_node  one lane per node: f_int = gather(F), f_ext = s(t+) f_unit,
//               a+ = (f_ext - f_int)/m, v+ = v + dt/2 (a + a+), Dirichlet zeroing,
//               kinetic-energy and external-work partials.
// Each kernel's block 0 folds the previous kernel's partials into the per-step
// history rows, so the time loop needs no host synchronisation.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "hex8.hpp"

namespace cia {

struct BoxGrid {
  int nx, ny, nz;
  int pad;
  int64_t E, N;  // elements, nodes
  BoxGeom g;
};

struct NmScalars {
  double rr;
  int32_t done;  // 0 running, 1 converged, 2 breakdown
  int32_t iters;
};

// energy history columns
enum {
  H_KINETIC = 0,
  H_FREE = 1,
  H_DISS_RATE = 2,    // sum D w detJ at the step (times dt = dissipated increment)
  H_EXT_DU = 3,       // f_unit . (u+ - u)
  H_POS_WORK = 4,
  H_NEG_WORK = 5,
  H_UNLOAD = 6,       // volume with negative incremental work
  H_ACTIVE = 7,       // viscoplastic-active volume
  H_MARGIN_MIN = 8,
  H_MARGIN_MAX = 9,
  H_N = 10
};
enum { PE_N = 8, PN_N = 2 };  // element partial sums (incl. min/max), node partial sums

__device__ __forceinline__ void box_corners(int64_t e64, int nx, int ny, int64_t nd[8]) {
  const unsigned e = (unsigned)e64;
  const unsigned r = e / (unsigned)nx;
  const int i = (int)(e - r * (unsigned)nx);
  const int k = (int)(r / (unsigned)ny);
  const int j = (int)(r - (unsigned)k * (unsigned)ny);
  const int64_t sx = 1, sy = nx + 1, sz = (int64_t)(nx + 1) * (ny + 1);
  const int64_t base = i + sy * j + sz * k;
#pragma unroll
  for (int a = 0; a < 8; ++a) nd[a] = base + ((a & 1) ? sx : 0) + ((a & 2) ? sy : 0) + ((a & 4) ? sz : 0);
}

// Node of corner a (a = ax + 2 ay + 4 az) of box element e.
__device__ __forceinline__ int64_t box_corner(int64_t e64, int a, int nx, int ny) {
  const unsigned e = (unsigned)e64;
  const unsigned r = e / (unsigned)nx;
  const int i = (int)(e - r * (unsigned)nx);
  const int k = (int)(r / (unsigned)ny);
  const int j = (int)(r - (unsigned)k * (unsigned)ny);
  return (i + (a & 1)) + (int64_t)(nx + 1) * ((j + ((a >> 1) & 1)) + (int64_t)(ny + 1) * (k + ((a >> 2) & 1)));
}

template <int BLOCK>
__device__ __forceinline__ void block_partials(const double* v, int nsum, const double* mins,
                                               const double* maxs, double* part, int P,
                                               double* lds) {
  for (int s = 0; s < nsum; ++s) {
    const double t = block_sum<BLOCK>(v[s], lds);
    if (threadIdx.x == 0) part[s * P + blockIdx.x] = t;
  }
  if (mins) {
    const double a = block_min<BLOCK>(mins[0], lds);
    const double b = block_max<BLOCK>(maxs[0], lds);
    if (threadIdx.x == 0) {
      part[nsum * P + blockIdx.x] = a;
      part[(nsum + 1) * P + blockIdx.x] = b;
    }
  }
}

// Fold the element partials of step `row` into hist[row][...] (block 0 only).
__device__ __forceinline__ void fold_elem(const double* pe, int P, double jw, double* hist_row,
                                          double* lds) {
  __shared__ double bc[PE_N];
  for (int s = 0; s < PE_N; ++s) {
    double a = (s == 6) ? INFINITY : (s == 7 ? -INFINITY : 0.0);
    for (int b = threadIdx.x; b < P; b += blockDim.x) {
      const double v = pe[s * P + b];
      a = (s == 6) ? fmin(a, v) : (s == 7 ? fmax(a, v) : a + v);
    }
    if (s == 6) a = block_min<kBlock>(a, lds);
    else if (s == 7) a = block_max<kBlock>(a, lds);
    else a = block_sum<kBlock>(a, lds);
    if (threadIdx.x == 0) bc[s] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    hist_row[H_FREE] = bc[0] * jw;
    hist_row[H_DISS_RATE] = bc[1] * jw;
    hist_row[H_POS_WORK] = bc[2] * jw;
    hist_row[H_NEG_WORK] = bc[3] * jw;
    hist_row[H_UNLOAD] = bc[4] * jw;
    hist_row[H_ACTIVE] = bc[5] * jw;
    hist_row[H_MARGIN_MIN] = bc[6];
    hist_row[H_MARGIN_MAX] = bc[7];
  }
}

__device__ __forceinline__ void fold_node(const double* pn, int P, double* hist_row, double* lds) {
  __shared__ double bc[PN_N];
  for (int s = 0; s < PN_N; ++s) {
    double a = 0.0;
    for (int b = threadIdx.x; b < P; b += blockDim.x) a += pn[s * P + b];
    a = block_sum<kBlock>(a, lds);
    if (threadIdx.x == 0) bc[s] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    hist_row[H_KINETIC] = bc[0];
    hist_row[H_EXT_DU] = bc[1];
  }
}

// predictor = false: evaluate at u itself (initial assembly, rate 0).
template <class T, bool PREDICT>
__global__ void __launch_bounds__(kBlock)
    k_dyn_elem(BoxGrid G, const PhaseTable* __restrict__ tab, const uint8_t* __restrict__ ephase,
               Settings S, double dt, const double* __restrict__ u, const double* __restrict__ v,
               const double* __restrict__ a, const uint8_t* __restrict__ fixed,
               T* __restrict__ state, T* __restrict__ sig_out, double* __restrict__ F,
               double* __restrict__ part_e, const double* __restrict__ part_n_prev, int Pn,
               double* __restrict__ hist_prev_row, unsigned long long* __restrict__ first_fail) {
  __shared__ double lds[kBlock / 64];
  if (hist_prev_row && blockIdx.x == 0) fold_node(part_n_prev, Pn, hist_prev_row, lds);
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  double vals[6] = {0, 0, 0, 0, 0, 0};
  double mn = INFINITY, mx = -INFINITY;
  // Lane q forms the predictor at corner q only; the element's 8 lanes exchange
  // their corners through LDS.
  __shared__ double cu[kBlock][6];
  {
    double own[6] = {0, 0, 0, 0, 0, 0};
    if (e < G.E) {
      const int64_t nq = box_corner(e, q, G.nx, G.ny);
      const double h2 = 0.5 * dt * dt;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int64_t d = 3 * nq + m;
        const double u0 = u[d];
        double u1 = u0;
        if (PREDICT) u1 = fixed[d] ? 0.0 : u0 + dt * v[d] + h2 * a[d];
        own[m] = u1;
        own[3 + m] = u1 - u0;
      }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) cu[threadIdx.x][m] = own[m];
  }
  __syncthreads();
  if (e < G.E) {
    const int64_t np = G.E * 8;
    T eps[6], rate[6];
    strain_pair_from_rows<T, 6>(&cu[threadIdx.x & ~7], 3, q, G.g, eps, rate);
#pragma unroll
    for (int m = 0; m < 6; ++m) rate[m] /= (T)dt;
    const PhaseConst P = tab->p[ephase[e]];
    const bool j2 = P.kind == CIA_MAT_J2_VISCOPLASTIC;
    T epsp[6];
    T qb = (T)0, ga = (T)0;
    if (j2) {
#pragma unroll
      for (int m = 0; m < 6; ++m) epsp[m] = state[m * np + p];
      qb = state[6 * np + p];
      ga = state[7 * np + p];
    } else {
#pragma unroll
      for (int m = 0; m < 6; ++m) epsp[m] = (T)0;
    }
    PointOut<T> o;
    PointExtra<T> ex;
    const int st = material_point<T, true>(P, S, dt, eps, rate, epsp, qb, ga, o, &ex, nullptr);
    if (st != CIA_PT_OK) atomicMin(first_fail, (unsigned long long)p);
    if (j2) {
#pragma unroll
      for (int m = 0; m < 6; ++m) state[m * np + p] = o.epsp[m];
      state[6 * np + p] = o.qbar;
      state[7 * np + p] = o.gam;
    }
    if (sig_out) {
#pragma unroll
      for (int m = 0; m < 6; ++m) sig_out[m * np + p] = o.sig[m];
    }
    vals[0] = (double)ex.psi;
    vals[1] = (double)ex.diss;
    // increment diagnostics exist for J2 points and for elastic points of a phased model
    const double pw = (double)ex.d[CIA_D_INCREMENTAL_WORK_DENSITY];
    if (isfinite(pw)) {
      vals[2] = fmax(pw, 0.0);
      vals[3] = fmin(pw, 0.0);
      vals[4] = pw < 0.0 ? 1.0 : 0.0;
    }
    const double act = (double)ex.d[CIA_D_VISCOPLASTIC_ACTIVE];
    vals[5] = act > 0.5 ? 1.0 : 0.0;
    const double mg = (double)ex.d[CIA_D_YIELD_ACTIVATION_MARGIN];
    if (isfinite(mg)) {
      mn = mg;
      mx = mg;
    }
    T f[8][3];
    force_at_gp(o.sig, (T)G.g.jw, q, G.g, f);
    T fin[3];
    element_reduce8(q, f, fin);
#pragma unroll
    for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = (double)fin[m];
  }
  block_partials<kBlock>(vals, 6, &mn, &mx, part_e, gridDim.x, lds);
}

// Node pass.  INIT: a = (f_ext - f_int)/m at the initial state only.
template <bool INIT>
__global__ void __launch_bounds__(kBlock)
    k_dyn_node(BoxGrid G, double dt, const double* __restrict__ F, const double* __restrict__ mass,
               const uint8_t* __restrict__ fixed, const double* __restrict__ f_unit, double s_ext,
               double* __restrict__ u, double* __restrict__ v, double* __restrict__ a,
               double* __restrict__ fint_out, double* __restrict__ part_n,
               const double* __restrict__ part_e, int Pe, double* __restrict__ hist_row,
               double jw) {
  __shared__ double lds[kBlock / 64];
  if (hist_row && blockIdx.x == 0) fold_elem(part_e, Pe, jw, hist_row, lds);
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double vals[2] = {0.0, 0.0};
  if (n < G.N) {
    const int nx1 = G.nx + 1, ny1 = G.ny + 1;
    const unsigned nu = (unsigned)n, r = nu / (unsigned)nx1;
    const int i = (int)(nu - r * (unsigned)nx1);
    const int k = (int)(r / (unsigned)ny1);
    const int j = (int)(r - (unsigned)k * (unsigned)ny1);
    double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ei = i - (c & 1), ej = j - ((c >> 1) & 1), ek = k - ((c >> 2) & 1);
      if (ei < 0 || ej < 0 || ek < 0 || ei >= G.nx || ej >= G.ny || ek >= G.nz) continue;
      const int64_t e = ei + (int64_t)G.nx * (ej + (int64_t)G.ny * ek);
#pragma unroll
      for (int m = 0; m < 3; ++m) f[m] += F[(3 * c + m) * G.E + e];
    }
    const double h2 = 0.5 * dt * dt;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      const bool fx = fixed[d] != 0;
      const double fe = s_ext * f_unit[d];
      const double an = fx ? 0.0 : (fe - f[m]) / mass[d];
      if (fint_out) fint_out[d] = f[m];
      if (INIT) {
        a[d] = an;
        const double vv = v[d];
        vals[0] += 0.5 * mass[d] * vv * vv;
      } else {
        const double u0 = u[d];
        const double u1 = fx ? 0.0 : u0 + dt * v[d] + h2 * a[d];
        const double v1 = fx ? 0.0 : v[d] + 0.5 * dt * (a[d] + an);
        vals[0] += 0.5 * mass[d] * v1 * v1;
        vals[1] += f_unit[d] * (u1 - u0);
        u[d] = u1;
        v[d] = v1;
        a[d] = an;
      }
    }
  }
  block_partials<kBlock>(vals, 2, nullptr, nullptr, part_n, gridDim.x, lds);
}

// History folds with one block per partial row (the element pass has one partial per
// block of 256 Gauss points, tens of thousands on a large slab; folding them in one
// block of the next kernel would serialise that kernel behind it).  Same summation
// order as fold_elem / fold_node.
constexpr int kFoldBlock = 1024;

__global__ void __launch_bounds__(kFoldBlock)
    k_fold_elem_rows(const double* __restrict__ pe, int P, double jw, double* __restrict__ hist_row) {
  __shared__ double lds[kFoldBlock / 64];
  const int s = blockIdx.x;
  const double init = (s == 6) ? INFINITY : (s == 7 ? -INFINITY : 0.0);
  // four independent accumulators per thread keep several loads in flight
  double acc[4] = {init, init, init, init};
  const double* row = pe + (int64_t)s * P;
  int b = threadIdx.x;
  for (; b + 3 * kFoldBlock < P; b += 4 * kFoldBlock) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double v = row[b + u * kFoldBlock];
      acc[u] = (s == 6) ? fmin(acc[u], v) : (s == 7 ? fmax(acc[u], v) : acc[u] + v);
    }
  }
  for (; b < P; b += kFoldBlock) {
    const double v = row[b];
    acc[0] = (s == 6) ? fmin(acc[0], v) : (s == 7 ? fmax(acc[0], v) : acc[0] + v);
  }
  double a;
  if (s == 6) a = block_min<kFoldBlock>(fmin(fmin(acc[0], acc[1]), fmin(acc[2], acc[3])), lds);
  else if (s == 7) a = block_max<kFoldBlock>(fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3])), lds);
  else a = block_sum<kFoldBlock>((acc[0] + acc[1]) + (acc[2] + acc[3]), lds);
  if (threadIdx.x == 0) {
    const int col[PE_N] = {H_FREE, H_DISS_RATE, H_POS_WORK, H_NEG_WORK, H_UNLOAD, H_ACTIVE,
                           H_MARGIN_MIN, H_MARGIN_MAX};
    hist_row[col[s]] = s < 6 ? a * jw : a;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_fold_node_rows(const double* __restrict__ pn, int P, double* __restrict__ hist_row) {
  __shared__ double lds[kBlock / 64];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < P; b += blockDim.x) a += pn[s * P + b];
  a = block_sum<kBlock>(a, lds);
  if (threadIdx.x == 0) hist_row[s == 0 ? H_KINETIC : H_EXT_DU] = a;
}

// ---------------------------------------------------------------------------
// Generic-material step: the caller evaluates the material between the strain
// pass and the force pass (assemble_bundle_material_response,
// src/fem/assembly/nonlinear_solid.py:222-300, with any MaterialModel).

// Strains (and rates) at the Verlet predictor u+ (PREDICT) or at u itself.
template <bool PREDICT>
__global__ void __launch_bounds__(kBlock)
    k_box_strain(BoxGrid G, double dt, const double* __restrict__ u, const double* __restrict__ v,
                 const double* __restrict__ a, const uint8_t* __restrict__ fixed,
                 double* __restrict__ strain, double* __restrict__ rate) {
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  int64_t nd[8];
  box_corners(e, G.nx, G.ny, nd);
  double bx[8], by[8], bz[8];
  gp_factors<double>(q, G.g, bx, by, bz);
  double un[8][3], du[8][3];
  const double h2 = 0.5 * dt * dt;
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * nd[c] + m;
      const double u0 = u[d];
      double u1 = u0;
      if (PREDICT) u1 = fixed[d] ? 0.0 : u0 + dt * v[d] + h2 * a[d];
      un[c][m] = u1;
      du[c][m] = u1 - u0;
    }
  double eps[6], rt[6];
  strain_from(un, bx, by, bz, eps);
  strain_from(du, bx, by, bz, rt);
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    strain[m * np + p] = eps[m];
    if (rate) rate[m * np + p] = rt[m] / dt;
  }
}

// Element force rows from caller stresses [6][n_gp] and the energy partials of
// k_dyn_elem from caller point arrays (each may be NULL: contributes nothing).
__global__ void __launch_bounds__(kBlock)
    k_box_force(BoxGrid G, const double* __restrict__ stress, const double* __restrict__ psi,
                const double* __restrict__ diss, const double* __restrict__ work,
                const double* __restrict__ active, const double* __restrict__ margin,
                double* __restrict__ F, double* __restrict__ part_e) {
  __shared__ double lds[kBlock / 64];
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  double vals[6] = {0, 0, 0, 0, 0, 0};
  double mn = INFINITY, mx = -INFINITY;
  if (e < G.E) {
    const int64_t np = G.E * 8;
    double bx[8], by[8], bz[8];
    gp_factors<double>(q, G.g, bx, by, bz);
    double sg[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) sg[m] = stress[m * np + p];
    if (psi) vals[0] = psi[p];
    if (diss) vals[1] = diss[p];
    if (work) {
      const double pw = work[p];
      if (isfinite(pw)) {
        vals[2] = fmax(pw, 0.0);
        vals[3] = fmin(pw, 0.0);
        vals[4] = pw < 0.0 ? 1.0 : 0.0;
      }
    }
    if (active) {
      const double ac = active[p];  // a flag for point materials, a volume fraction for RVEs
      if (isfinite(ac)) vals[5] = ac;
    }
    if (margin) {
      const double mg = margin[p];
      if (isfinite(mg)) {
        mn = mg;
        mx = mg;
      }
    }
    double f[8][3];
    force_from(sg, G.g.jw, bx, by, bz, f);
    double fin[3];
    element_reduce8(q, f, fin);
#pragma unroll
    for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = fin[m];
  }
  block_partials<kBlock>(vals, 6, &mn, &mx, part_e, gridDim.x, lds);
}


// ---------------------------------------------------------------------------
// Implicit Newmark pieces (solve_material_implicit_newmark, material.py:300-701):
// the trial residual R = M a + f_int - f_ext with a = (u - u*)/(beta dt^2), and a
// Jacobi-preconditioned CG for (K_t + c M) x = b on the free dofs.  The mass is
// the consistent Hex8 mass sum_q rho N^T N w detJ (elasticity.py:109-126), applied
// per Gauss-point lane like the stiffness.

// Corner shape values at Gauss point q: N_a = Wx[ax] Wy[ay] Wz[az].
__device__ __forceinline__ void gp_shape(int q, double N[8]) {
  const int qx = q >> 2, qy = (q >> 1) & 1, qz = q & 1;
#pragma unroll
  for (int a = 0; a < 8; ++a)
    N[a] = wq(qx, a & 1) * wq(qy, (a >> 1) & 1) * wq(qz, (a >> 2) & 1);
}

// Sum of P partials, identical in every block (fixed order), broadcast to all threads.
__device__ __forceinline__ double sum_parts(const double* part, int P, double* lds) {
  __shared__ double bc;
  double a = 0.0;
  for (int b = threadIdx.x; b < P; b += blockDim.x) a += part[b];
  a = block_sum<kBlock>(a, lds);
  if (threadIdx.x == 0) bc = a;
  __syncthreads();
  return bc;
}

// Gather element rows F[24][E] to the node of lane n (non-periodic box).
__device__ __forceinline__ void gather_node(const BoxGrid& G, const double* F, int64_t n,
                                            double f[3]) {
  const int nx1 = G.nx + 1, ny1 = G.ny + 1;
  const unsigned nu = (unsigned)n, r = nu / (unsigned)nx1;
  const int i = (int)(nu - r * (unsigned)nx1);
  const int k = (int)(r / (unsigned)ny1);
  const int j = (int)(r - (unsigned)k * (unsigned)ny1);
  f[0] = f[1] = f[2] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ei = i - (c & 1), ej = j - ((c >> 1) & 1), ek = k - ((c >> 2) & 1);
    if (ei < 0 || ej < 0 || ek < 0 || ei >= G.nx || ej >= G.ny || ek >= G.nz) continue;
    const int64_t e = ei + (int64_t)G.nx * (ej + (int64_t)G.ny * ek);
#pragma unroll
    for (int m = 0; m < 3; ++m) f[m] += F[(3 * c + m) * G.E + e];
  }
}

// Strains at u_new and rates (u_new - u_old)/dt.
__global__ void __launch_bounds__(kBlock)
    k_box_strain_between(BoxGrid G, double dt, const double* __restrict__ un,
                         const double* __restrict__ uo, double* __restrict__ strain,
                         double* __restrict__ rate) {
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  int64_t nd[8];
  box_corners(e, G.nx, G.ny, nd);
  double bx[8], by[8], bz[8];
  gp_factors<double>(q, G.g, bx, by, bz);
  double u1[8][3], du[8][3];
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * nd[c] + m;
      u1[c][m] = un[d];
      du[c][m] = un[d] - uo[d];
    }
  double eps[6], rt[6];
  strain_from(u1, bx, by, bz, eps);
  strain_from(du, bx, by, bz, rt);
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    strain[m * np + p] = eps[m];
    if (rate) rate[m * np + p] = rt[m] / dt;
  }
}

// Element rows of f_int (from stresses) and of M a with a = (u - u*) * inv_b.
__global__ void __launch_bounds__(kBlock)
    k_nm_residual_elem(BoxGrid G, double rho, double inv_b, const double* __restrict__ stress,
                       const double* __restrict__ u, const double* __restrict__ ustar,
                       double* __restrict__ F, double* __restrict__ Fm) {
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  int64_t nd[8];
  box_corners(e, G.nx, G.ny, nd);
  double bx[8], by[8], bz[8], N[8];
  gp_factors<double>(q, G.g, bx, by, bz);
  gp_shape(q, N);
  double sg[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) sg[m] = stress[m * np + p];
  double f[8][3], fin[3];
  force_from(sg, G.g.jw, bx, by, bz, f);
  element_reduce8(q, f, fin);
#pragma unroll
  for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = fin[m];
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * nd[b] + m;
      acc[m] += N[b] * ((u[d] - ustar[d]) * inv_b);
    }
  const double s = rho * G.g.jw;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m) f[a][m] = s * N[a] * acc[m];
  element_reduce8(q, f, fin);
#pragma unroll
  for (int m = 0; m < 3; ++m) Fm[(3 * q + m) * G.E + e] = fin[m];
}

// Node pass: a, v, f_int, residual (fixed dofs zero), partials of |R_free|^2, |u_free|^2.
__global__ void __launch_bounds__(kBlock)
    k_nm_residual_node(BoxGrid G, const double* __restrict__ F, const double* __restrict__ Fm,
                       double s_ext, const double* __restrict__ f_unit,
                       const uint8_t* __restrict__ fixed, const double* __restrict__ u,
                       const double* __restrict__ ustar, const double* __restrict__ vstar,
                       double inv_b, double gdt, double* __restrict__ a_out,
                       double* __restrict__ v_out, double* __restrict__ fint_out,
                       double* __restrict__ r_out, double* __restrict__ part) {
  __shared__ double lds[kBlock / 64];
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double vals[2] = {0.0, 0.0};
  if (n < G.N) {
    double fi[3], fm[3];
    gather_node(G, F, n, fi);
    gather_node(G, Fm, n, fm);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      const double acc = (u[d] - ustar[d]) * inv_b;
      a_out[d] = acc;
      v_out[d] = vstar[d] + gdt * acc;
      fint_out[d] = fi[m];
      const double r = fixed[d] ? 0.0 : fm[m] + fi[m] - s_ext * f_unit[d];
      r_out[d] = r;
      vals[0] += r * r;
      if (!fixed[d]) vals[1] += u[d] * u[d];
    }
  }
  block_partials<kBlock>(vals, 2, nullptr, nullptr, part, gridDim.x, lds);
}

// y rows of (K_t + cm M) p per Gauss-point lane; tangent [36][n_gp] row-major per
// point, or NULL for the mass term alone.
__global__ void __launch_bounds__(kBlock)
    k_nm_apply_elem(BoxGrid G, const double* __restrict__ tan, double rho_cm,
                    const double* __restrict__ pv, double* __restrict__ F,
                    const NmScalars* __restrict__ sc) {
  if (sc && sc->done) return;
  const int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int64_t e = p >> 3;
  const int q = (int)(p & 7);
  if (e >= G.E) return;
  const int64_t np = G.E * 8;
  int64_t nd[8];
  box_corners(e, G.nx, G.ny, nd);
  double bx[8], by[8], bz[8], N[8];
  gp_factors<double>(q, G.g, bx, by, bz);
  gp_shape(q, N);
  double pe[8][3];
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int m = 0; m < 3; ++m) pe[c][m] = pv[3 * nd[c] + m];
  double f[8][3];
  if (tan) {
    double eps[6], sg[6];
    strain_from(pe, bx, by, bz, eps);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) t += tan[(6 * i + j) * np + p] * eps[j];
      sg[i] = t;
    }
    force_from(sg, G.g.jw, bx, by, bz, f);
  } else {
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a][0] = f[a][1] = f[a][2] = 0.0;
  }
  double mp[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int m = 0; m < 3; ++m) mp[m] += N[b] * pe[b][m];
  const double s = rho_cm * G.g.jw;
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int m = 0; m < 3; ++m) f[a][m] += s * N[a] * mp[m];
  double fin[3];
  element_reduce8(q, f, fin);
#pragma unroll
  for (int m = 0; m < 3; ++m) F[(3 * q + m) * G.E + e] = fin[m];
}

// Jacobi diagonal of K_t + cm M at every free dof (1 at fixed dofs), inverted.
__global__ void __launch_bounds__(kBlock)
    k_nm_diag(BoxGrid G, const double* __restrict__ tan, double rho_cm,
              const uint8_t* __restrict__ fixed, double* __restrict__ dinv) {
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  const int64_t np = G.E * 8;
  const int nx1 = G.nx + 1, ny1 = G.ny + 1;
  const unsigned nu = (unsigned)n, r = nu / (unsigned)nx1;
  const int i = (int)(nu - r * (unsigned)nx1);
  const int k = (int)(r / (unsigned)ny1);
  const int j = (int)(r - (unsigned)k * (unsigned)ny1);
  double dg[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 8; ++c) {
    const int ei = i - (c & 1), ej = j - ((c >> 1) & 1), ek = k - ((c >> 2) & 1);
    if (ei < 0 || ej < 0 || ek < 0 || ei >= G.nx || ej >= G.ny || ek >= G.nz) continue;
    const int64_t e = ei + (int64_t)G.nx * (ej + (int64_t)G.ny * ek);
    for (int q = 0; q < 8; ++q) {
      const int64_t p = 8 * e + q;
      double col[3][6], N[8];
      b_columns<double>(q, c, G.g, col);
      gp_shape(q, N);
      for (int m = 0; m < 3; ++m) {
        double kd = 0.0;
        if (tan) {
          for (int a = 0; a < 6; ++a) {
            double t = 0.0;
            for (int b = 0; b < 6; ++b) t += tan[(6 * a + b) * np + p] * col[m][b];
            kd += col[m][a] * t;
          }
        }
        dg[m] += G.g.jw * (kd + rho_cm * N[c] * N[c]);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int64_t d = 3 * n + m;
    dinv[d] = (fixed[d] || !(dg[m] > 0.0)) ? 1.0 : 1.0 / dg[m];
  }
}

__global__ void __launch_bounds__(kBlock)
    k_nm_residual_node_free(BoxGrid G, const double* __restrict__ F, double* __restrict__ y) {
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (n >= G.N) return;
  double f[3];
  gather_node(G, F, n, f);
#pragma unroll
  for (int m = 0; m < 3; ++m) y[3 * n + m] = f[m];
}

// PCG state.  Partials: rz[2][P] (double-buffered by iteration parity), rr[P],
// rr0[P], py[P].
__global__ void __launch_bounds__(kBlock)
    k_nm_init(BoxGrid G, const double* __restrict__ b, const double* __restrict__ dinv,
              const uint8_t* __restrict__ fixed, double* __restrict__ x, double* __restrict__ r,
              double* __restrict__ z, double* __restrict__ pv, double* __restrict__ part_rz,
              double* __restrict__ part_rr0, NmScalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double vals[2] = {0.0, 0.0};
  if (n < G.N) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      const double rv = fixed[d] ? 0.0 : b[d];
      const double zv = rv * dinv[d];
      x[d] = 0.0;
      r[d] = rv;
      z[d] = zv;
      pv[d] = zv;
      vals[0] += rv * zv;
      vals[1] += rv * rv;
    }
  }
  const double a0 = block_sum<kBlock>(vals[0], lds);
  const double a1 = block_sum<kBlock>(vals[1], lds);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = a0;
    part_rr0[blockIdx.x] = a1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->done = 0;
    sc->iters = 0;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_nm_apply_node(BoxGrid G, const double* __restrict__ F, const uint8_t* __restrict__ fixed,
                    const double* __restrict__ pv, double* __restrict__ y,
                    double* __restrict__ part_py, const NmScalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double py = 0.0;
  if (n < G.N) {
    double f[3];
    gather_node(G, F, n, f);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      const double yv = fixed[d] ? 0.0 : f[m];
      y[d] = yv;
      py += pv[d] * yv;
    }
  }
  py = block_sum<kBlock>(py, lds);
  if (threadIdx.x == 0) part_py[blockIdx.x] = py;
}

// alpha = rz_it / p.y; x += alpha p; r -= alpha y; z = D^-1 r; partials rz_{it+1}, rr.
__global__ void __launch_bounds__(kBlock)
    k_nm_update(BoxGrid G, int it, const double* __restrict__ dinv, const double* __restrict__ pv,
                const double* __restrict__ y, double* __restrict__ x, double* __restrict__ r,
                double* __restrict__ z, const double* __restrict__ part_py,
                double* __restrict__ part_rz, double* __restrict__ part_rr, int P,
                NmScalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const double py = sum_parts(part_py, P, lds);
  const double rz = sum_parts(part_rz + (it & 1) * P, P, lds);
  if (rz == 0.0) {  // r = 0: x is exact
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 1;
    return;
  }
  if (!(py > 0.0) || !isfinite(py)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = 2;
    return;
  }
  const double alpha = rz / py;
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  double vals[2] = {0.0, 0.0};
  if (n < G.N) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      x[d] += alpha * pv[d];
      const double rv = r[d] - alpha * y[d];
      const double zv = rv * dinv[d];
      r[d] = rv;
      z[d] = zv;
      vals[0] += rv * zv;
      vals[1] += rv * rv;
    }
  }
  const double a0 = block_sum<kBlock>(vals[0], lds);
  const double a1 = block_sum<kBlock>(vals[1], lds);
  if (threadIdx.x == 0) {
    part_rz[((it + 1) & 1) * P + blockIdx.x] = a0;
    part_rr[blockIdx.x] = a1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->iters = it + 1;
}

// Convergence test |r| <= max(rtol |b|, atol); else p = z + (rz_{it+1}/rz_it) p.
__global__ void __launch_bounds__(kBlock)
    k_nm_pdir(BoxGrid G, int it, double rtol, double atol, const double* __restrict__ z,
              double* __restrict__ pv, const double* __restrict__ part_rz,
              const double* __restrict__ part_rr, const double* __restrict__ part_rr0, int P,
              NmScalars* __restrict__ sc) {
  __shared__ double lds[kBlock / 64];
  if (sc->done) return;
  const double rr = sum_parts(part_rr, P, lds);
  const double rr0 = sum_parts(part_rr0, P, lds);
  const double rz0 = sum_parts(part_rz + (it & 1) * P, P, lds);
  const double rz1 = sum_parts(part_rz + ((it + 1) & 1) * P, P, lds);
  const double tol = fmax(rtol * sqrt(rr0), atol);
  if (rr <= tol * tol || !isfinite(rr)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->done = isfinite(rr) ? 1 : 2;
      sc->rr = rr;
    }
    return;
  }
  const double beta = rz1 / rz0;
  const int64_t n = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (n < G.N) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t d = 3 * n + m;
      pv[d] = z[d] + beta * pv[d];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->rr = rr;
}

}  // namespace cia

struct cia_box {
  cia::BoxGrid G;
  int precision;
  int n_phases;
  cia::PhaseTable* d_tab = nullptr;
  uint8_t* d_ephase = nullptr;
  double* F = nullptr;
  double* part_e = nullptr;
  double* part_n = nullptr;
  double* hist = nullptr;  // [steps+1][H_N], grown on demand
  int64_t hist_rows = 0;
  unsigned long long* first_fail = nullptr;
  // Newmark (allocated on first use)
  double* Fm = nullptr;          // [24][E] mass rows
  double* nm_vec = nullptr;      // r, z, p, y, dinv: 5 x 3N
  double* nm_part = nullptr;     // 5 x P partials
  cia::NmScalars* nm_sc = nullptr;
};

namespace cia {
static cia_status dyn_alloc(void** p, size_t bytes) {
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
}  // namespace cia

extern "C" {

cia_status cia_box_create(const cia_box_desc* d, cia_box** out) {
  using namespace cia;
  CIA_ARG(d && out, "NULL argument");
  CIA_ARG(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "divisions must be positive");
  CIA_ARG(d->n_phases >= 1 && d->n_phases <= CIA_MAX_PHASES && d->phases, "invalid phases");
  CIA_ARG(d->precision == CIA_PREC_FP64 || d->precision == CIA_PREC_FP32_REFINED, "bad precision");
  auto* B = new cia_box();
  B->G.nx = d->nx;
  B->G.ny = d->ny;
  B->G.nz = d->nz;
  B->G.E = (int64_t)d->nx * d->ny * d->nz;
  B->G.N = (int64_t)(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
  const double hx = d->lengths[0] / d->nx, hy = d->lengths[1] / d->ny, hz = d->lengths[2] / d->nz;
  B->G.g.ihx = 1.0 / hx;
  B->G.g.ihy = 1.0 / hy;
  B->G.g.ihz = 1.0 / hz;
  B->G.g.jw = hx * hy * hz / 8.0;
  B->precision = d->precision;
  B->n_phases = d->n_phases;
  PhaseTable tab{};
  tab.n = d->n_phases;
  for (int i = 0; i < d->n_phases; ++i) tab.p[i] = make_phase(d->phases[i]);
  std::vector<uint8_t> eph(B->G.E, 0);
  if (d->element_phase) {
    for (int64_t e = 0; e < B->G.E; ++e) {
      const int32_t v = d->element_phase[e];
      if (v < 0 || v >= d->n_phases) {
        delete B;
        set_error_msg("element_phase out of range");
        return CIA_ERR_INVALID_ARGUMENT;
      }
      eph[e] = (uint8_t)v;
    }
  }
  const int64_t Pe = grid_for(8 * B->G.E), Pn = grid_for(B->G.N);
  cia_status s;
#define ALLOC(ptr, bytes)                    \
  do {                                       \
    s = dyn_alloc((void**)&(ptr), (bytes));  \
    if (s != CIA_OK) {                       \
      cia_box_destroy(B);                    \
      return s;                              \
    }                                        \
  } while (0)
  ALLOC(B->d_tab, sizeof(PhaseTable));
  ALLOC(B->d_ephase, B->G.E);
  ALLOC(B->F, 24 * B->G.E * sizeof(double));
  ALLOC(B->part_e, PE_N * Pe * sizeof(double));
  ALLOC(B->part_n, PN_N * Pn * sizeof(double));
  ALLOC(B->first_fail, 8);
#undef ALLOC
  if (hipMemcpy(B->d_tab, &tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(B->d_ephase, eph.data(), B->G.E, hipMemcpyHostToDevice) != hipSuccess) {
    cia_box_destroy(B);
    set_error_msg("hipMemcpy of box tables failed");
    return CIA_ERR_HIP;
  }
  *out = B;
  return CIA_OK;
}

cia_status cia_box_destroy(cia_box* B) {
  if (!B) return CIA_OK;
  void* ptrs[] = {B->d_tab, B->d_ephase, B->F, B->part_e, B->part_n, B->hist, B->first_fail,
                  B->Fm, B->nm_vec, B->nm_part, B->nm_sc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete B;
  return CIA_OK;
}

cia_status cia_box_explicit_steps(cia_box* B, void* stream, int32_t initialize, double time_step,
                                  int32_t n_steps, const double* load_scale,
                                  const cia_update_settings* settings, double* u, double* v,
                                  double* a, void* state, const double* mass,
                                  const uint8_t* fixed, const double* f_unit, double* stress,
                                  double* internal_force, double* history,
                                  cia_point_failure* failure) {
  using namespace cia;
  CIA_ARG(B && settings && u && v && a && mass && fixed && f_unit && history && load_scale,
          "NULL argument to cia_box_explicit_steps");
  CIA_ARG(time_step > 0.0 && n_steps >= 0, "invalid time stepping");
  hipStream_t st = (hipStream_t)stream;
  if (failure) {
    memset(failure, 0, sizeof(*failure));
    failure->point_id = -1;
  }
  const Settings S = make_settings(*settings);
  const int64_t rows = (int64_t)n_steps + 1;
  if (B->hist_rows < rows) {
    if (B->hist) (void)hipFree(B->hist);
    B->hist = nullptr;
    cia_status s = dyn_alloc((void**)&B->hist, rows * H_N * sizeof(double));
    if (s != CIA_OK) return s;
    B->hist_rows = rows;
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, B->first_fail, ~0ull);
  const unsigned ge = grid_for(8 * B->G.E), gn = grid_for(B->G.N);
  const double jw = B->G.g.jw;
  const bool f64 = B->precision == CIA_PREC_FP64;
  auto elem = [&](bool predict, double* hist_prev, bool last) {
    double* sig = last ? stress : nullptr;
    if (f64) {
      if (predict)
        hipLaunchKernelGGL((k_dyn_elem<double, true>), dim3(ge), dim3(kBlock), 0, st, B->G, B->d_tab,
                           B->d_ephase, S, time_step, u, v, a, fixed, (double*)state, (double*)sig,
                           B->F, B->part_e, B->part_n, (int)gn, hist_prev, B->first_fail);
      else
        hipLaunchKernelGGL((k_dyn_elem<double, false>), dim3(ge), dim3(kBlock), 0, st, B->G,
                           B->d_tab, B->d_ephase, S, time_step, u, v, a, fixed, (double*)state,
                           (double*)sig, B->F, B->part_e, B->part_n, (int)gn, hist_prev,
                           B->first_fail);
    } else {
      if (predict)
        hipLaunchKernelGGL((k_dyn_elem<float, true>), dim3(ge), dim3(kBlock), 0, st, B->G, B->d_tab,
                           B->d_ephase, S, time_step, u, v, a, fixed, (float*)state, (float*)sig,
                           B->F, B->part_e, B->part_n, (int)gn, hist_prev, B->first_fail);
      else
        hipLaunchKernelGGL((k_dyn_elem<float, false>), dim3(ge), dim3(kBlock), 0, st, B->G,
                           B->d_tab, B->d_ephase, S, time_step, u, v, a, fixed, (float*)state,
                           (float*)sig, B->F, B->part_e, B->part_n, (int)gn, hist_prev,
                           B->first_fail);
    }
  };
  // history row of step s is base + s; row 0 is the initial state when initialising
  const int64_t base = initialize ? 1 : 0;
  auto fold = [&](double* row) {
    hipLaunchKernelGGL(k_fold_elem_rows, dim3(PE_N), dim3(kFoldBlock), 0, st,
                       (const double*)B->part_e, (int)ge, jw, row);
  };
  auto fold_n = [&](double* row) {
    hipLaunchKernelGGL(k_fold_node_rows, dim3(PN_N), dim3(kBlock), 0, st, (const double*)B->part_n,
                       (int)gn, row);
  };
  if (initialize) {
    elem(false, nullptr, n_steps == 0);
    fold(B->hist);
    hipLaunchKernelGGL((k_dyn_node<true>), dim3(gn), dim3(kBlock), 0, st, B->G, time_step, B->F,
                       mass, fixed, f_unit, load_scale[0], u, v, a,
                       n_steps == 0 ? internal_force : nullptr, B->part_n, B->part_e, (int)ge,
                       (double*)nullptr, jw);
    fold_n(B->hist);
  }
  for (int32_t s = 0; s < n_steps; ++s) {
    const bool last = s + 1 == n_steps;
    double* row = B->hist + (base + s) * H_N;
    elem(true, nullptr, last);
    fold(row);
    hipLaunchKernelGGL((k_dyn_node<false>), dim3(gn), dim3(kBlock), 0, st, B->G, time_step, B->F,
                       mass, fixed, f_unit, load_scale[s + 1], u, v, a,
                       last ? internal_force : nullptr, B->part_n, B->part_e, (int)ge,
                       (double*)nullptr, jw);
    fold_n(row);
  }
  CIA_HIP(hipGetLastError());
  const int64_t nrows = base + n_steps;
  if (nrows > 0)
    CIA_HIP(hipMemcpyAsync(history, B->hist, nrows * H_N * sizeof(double), hipMemcpyDeviceToHost, st));
  unsigned long long ff = 0;
  CIA_HIP(hipMemcpyAsync(&ff, B->first_fail, sizeof(ff), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  if (ff != ~0ull) {
    if (failure) {
      failure->point_id = (int64_t)ff;
      failure->reason = -1;  // reason re-derived by the host on the last committed state
    }
    return CIA_ERR_POINT_FAILURE;
  }
  return CIA_OK;
}

cia_status cia_box_generic_strains(cia_box* B, void* stream, int32_t predict, double time_step,
                                   const double* u, const double* v, const double* a,
                                   const uint8_t* fixed, double* strain, double* rate) {
  using namespace cia;
  CIA_ARG(B && u && strain, "NULL argument to cia_box_generic_strains");
  CIA_ARG(!predict || (v && a && fixed), "the predictor needs v, a and the fixed mask");
  CIA_ARG(time_step > 0.0, "time step must be positive");
  hipStream_t st = (hipStream_t)stream;
  const unsigned ge = grid_for(8 * B->G.E);
  if (predict)
    hipLaunchKernelGGL((k_box_strain<true>), dim3(ge), dim3(kBlock), 0, st, B->G, time_step, u, v,
                       a, fixed, strain, rate);
  else
    hipLaunchKernelGGL((k_box_strain<false>), dim3(ge), dim3(kBlock), 0, st, B->G, time_step, u,
                       v, a, fixed, strain, rate);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_box_generic_correct(cia_box* B, void* stream, int32_t initialize, double time_step,
                                   double load_scale, const double* stress, const double* psi,
                                   const double* diss, const double* work, const double* active,
                                   const double* margin, double* u, double* v, double* a,
                                   const double* mass, const uint8_t* fixed, const double* f_unit,
                                   double* internal_force, double* history_row) {
  using namespace cia;
  CIA_ARG(B && stress && u && v && a && mass && fixed && f_unit && history_row,
          "NULL argument to cia_box_generic_correct");
  CIA_ARG(time_step > 0.0, "time step must be positive");
  hipStream_t st = (hipStream_t)stream;
  const unsigned ge = grid_for(8 * B->G.E), gn = grid_for(B->G.N);
  hipLaunchKernelGGL(k_box_force, dim3(ge), dim3(kBlock), 0, st, B->G, stress, psi, diss, work,
                     active, margin, B->F, B->part_e);
  hipLaunchKernelGGL(k_fold_elem_rows, dim3(PE_N), dim3(kFoldBlock), 0, st,
                     (const double*)B->part_e, (int)ge, B->G.g.jw, history_row);
  if (initialize)
    hipLaunchKernelGGL((k_dyn_node<true>), dim3(gn), dim3(kBlock), 0, st, B->G, time_step, B->F,
                       mass, fixed, f_unit, load_scale, u, v, a, internal_force, B->part_n,
                       B->part_e, (int)ge, (double*)nullptr, B->G.g.jw);
  else
    hipLaunchKernelGGL((k_dyn_node<false>), dim3(gn), dim3(kBlock), 0, st, B->G, time_step, B->F,
                       mass, fixed, f_unit, load_scale, u, v, a, internal_force, B->part_n,
                       B->part_e, (int)ge, (double*)nullptr, B->G.g.jw);
  hipLaunchKernelGGL(k_fold_node_rows, dim3(PN_N), dim3(kBlock), 0, st, (const double*)B->part_n,
                     (int)gn, history_row);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

static cia_status nm_buffers(cia_box* B) {
  using namespace cia;
  if (B->nm_sc) return CIA_OK;
  const int64_t n3 = 3 * B->G.N, P = grid_for(B->G.N);
  cia_status s;
  if ((s = dyn_alloc((void**)&B->Fm, 24 * B->G.E * sizeof(double))) != CIA_OK) return s;
  if ((s = dyn_alloc((void**)&B->nm_vec, 5 * n3 * sizeof(double))) != CIA_OK) return s;
  if ((s = dyn_alloc((void**)&B->nm_part, 5 * P * sizeof(double))) != CIA_OK) return s;
  if ((s = dyn_alloc((void**)&B->nm_sc, sizeof(NmScalars))) != CIA_OK) return s;
  return CIA_OK;
}

cia_status cia_box_strains_between(cia_box* B, void* stream, double time_step, const double* u_new,
                                   const double* u_old, double* strain, double* rate) {
  using namespace cia;
  CIA_ARG(B && u_new && u_old && strain, "NULL argument to cia_box_strains_between");
  CIA_ARG(time_step > 0.0, "time step must be positive");
  hipLaunchKernelGGL(k_box_strain_between, dim3(grid_for(8 * B->G.E)), dim3(kBlock), 0,
                     (hipStream_t)stream, B->G, time_step, u_new, u_old, strain, rate);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_box_newmark_residual(cia_box* B, void* stream, double density, double beta,
                                    double gamma, double time_step, const double* stress,
                                    double load_scale, const double* f_unit, const uint8_t* fixed,
                                    const double* u, const double* u_pred, const double* v_pred,
                                    double* a, double* v, double* internal_force,
                                    double* residual, double* norms) {
  using namespace cia;
  CIA_ARG(B && stress && f_unit && fixed && u && u_pred && v_pred && a && v && internal_force &&
              residual && norms,
          "NULL argument to cia_box_newmark_residual");
  CIA_ARG(beta > 0.0 && time_step > 0.0, "invalid Newmark parameters");
  CIA_TRY(nm_buffers(B));
  hipStream_t st = (hipStream_t)stream;
  const unsigned ge = grid_for(8 * B->G.E), gn = grid_for(B->G.N);
  const double inv_b = 1.0 / (beta * time_step * time_step);
  hipLaunchKernelGGL(k_nm_residual_elem, dim3(ge), dim3(kBlock), 0, st, B->G, density, inv_b,
                     stress, u, u_pred, B->F, B->Fm);
  hipLaunchKernelGGL(k_nm_residual_node, dim3(gn), dim3(kBlock), 0, st, B->G, B->F, B->Fm,
                     load_scale, f_unit, fixed, u, u_pred, v_pred, inv_b, gamma * time_step, a, v,
                     internal_force, residual, B->nm_part);
  CIA_HIP(hipGetLastError());
  std::vector<double> h(2 * (size_t)gn);
  CIA_HIP(hipMemcpyAsync(h.data(), B->nm_part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         st));
  CIA_HIP(hipStreamSynchronize(st));
  double s0 = 0.0, s1 = 0.0;
  for (unsigned b = 0; b < gn; ++b) {
    s0 += h[b];
    s1 += h[gn + b];
  }
  norms[0] = sqrt(s0);
  norms[1] = sqrt(s1);
  return CIA_OK;
}

cia_status cia_box_mass_apply(cia_box* B, void* stream, double density, const double* x,
                              double* y) {
  using namespace cia;
  CIA_ARG(B && x && y, "NULL argument to cia_box_mass_apply");
  CIA_TRY(nm_buffers(B));
  hipStream_t st = (hipStream_t)stream;
  const unsigned ge = grid_for(8 * B->G.E), gn = grid_for(B->G.N);
  hipLaunchKernelGGL(k_nm_apply_elem, dim3(ge), dim3(kBlock), 0, st, B->G, (const double*)nullptr,
                     density, x, B->F, (const NmScalars*)nullptr);
  hipLaunchKernelGGL(k_nm_residual_node_free, dim3(gn), dim3(kBlock), 0, st, B->G, B->F, y);
  CIA_HIP(hipGetLastError());
  return CIA_OK;
}

cia_status cia_box_newmark_solve(cia_box* B, void* stream, const double* tangent, double density,
                                 double mass_coefficient, const uint8_t* fixed, const double* rhs,
                                 double* x, int32_t max_iterations, int32_t check_interval,
                                 double relative_tolerance, double absolute_tolerance,
                                 int32_t* iterations, double* residual_norm) {
  using namespace cia;
  CIA_ARG(B && fixed && rhs && x && iterations && residual_norm,
          "NULL argument to cia_box_newmark_solve");
  CIA_ARG(max_iterations > 0, "max_iterations must be positive");
  CIA_TRY(nm_buffers(B));
  hipStream_t st = (hipStream_t)stream;
  const int64_t n3 = 3 * B->G.N;
  const unsigned ge = grid_for(8 * B->G.E), gn = grid_for(B->G.N);
  const int P = (int)gn;
  double* r = B->nm_vec;
  double* z = r + n3;
  double* pv = z + n3;
  double* y = pv + n3;
  double* dinv = y + n3;
  double* part_rz = B->nm_part;  // [2][P]
  double* part_rr = part_rz + 2 * P;
  double* part_rr0 = part_rr + P;
  double* part_py = part_rr0 + P;
  const double rho_cm = density * mass_coefficient;
  hipLaunchKernelGGL(k_nm_diag, dim3(gn), dim3(kBlock), 0, st, B->G, tangent, rho_cm, fixed, dinv);
  hipLaunchKernelGGL(k_nm_init, dim3(gn), dim3(kBlock), 0, st, B->G, rhs, dinv, fixed, x, r, z, pv,
                     part_rz, part_rr0, B->nm_sc);
  CIA_HIP(hipGetLastError());
  const int chunk = std::max(1, (int)check_interval);
  NmScalars hs{};
  int it = 0;
  while (it < max_iterations) {
    const int n = std::min(chunk, max_iterations - it);
    for (int s = 0; s < n; ++s, ++it) {
      hipLaunchKernelGGL(k_nm_apply_elem, dim3(ge), dim3(kBlock), 0, st, B->G, tangent, rho_cm,
                         (const double*)pv, B->F, (const NmScalars*)B->nm_sc);
      hipLaunchKernelGGL(k_nm_apply_node, dim3(gn), dim3(kBlock), 0, st, B->G, (const double*)B->F,
                         fixed, (const double*)pv, y, part_py, (const NmScalars*)B->nm_sc);
      hipLaunchKernelGGL(k_nm_update, dim3(gn), dim3(kBlock), 0, st, B->G, it, dinv, pv, y, x, r, z,
                         part_py, part_rz, part_rr, P, B->nm_sc);
      hipLaunchKernelGGL(k_nm_pdir, dim3(gn), dim3(kBlock), 0, st, B->G, it, relative_tolerance,
                         absolute_tolerance, z, pv, part_rz, part_rr, part_rr0, P, B->nm_sc);
    }
    CIA_HIP(hipGetLastError());
    CIA_HIP(hipMemcpyAsync(&hs, B->nm_sc, sizeof(hs), hipMemcpyDeviceToHost, st));
    CIA_HIP(hipStreamSynchronize(st));
    if (hs.done) break;
  }
  *iterations = hs.iters;
  *residual_norm = sqrt(hs.rr);
  if (hs.done != 1) {
    set_error_msg(hs.done == 2 ? "Newmark PCG breakdown" : "Newmark PCG did not converge");
    return CIA_ERR_NOT_CONVERGED;
  }
  return CIA_OK;
}

}  // extern "C"
