// Device-side constitutive point updates for gfx950.
//
// j2vp_point() is the per-Gauss-point J2 Perzyna radial return of
// src/fem/materials/plasticity.py:526-681 with the multiplier solve of
// plasticity.py:711-854, the local modulus and consistent tangent of
// plasticity.py:907-963, the energies of plasticity.py:1300-1328 and the
// increment diagnostics of plasticity.py:684-708.  elastic_point() is
// linear_elastic.py:74-100 plus the elastic-phase diagnostics of
// phased.py:171-190.
//
// Storage type T is double (parity path) or float (FP32 element path); the
// yield-consistency scalar solve always runs in double.
//
// The consistent tangent is returned in compact form
//     C eps = K tr(eps) [1,1,1,0,0,0] + d dev_t(eps) + r n (n . eps)
// with dev_t(eps) = (e0 - tr/3, e1 - tr/3, e2 - tr/3, e3/2, e4/2, e5/2),
// which is exactly 3K E2T P_vol + d E2T P_dev + r n(x)n of plasticity.py:944-963
// (d = 2G, r = 0 on the elastic branch).  Eight numbers per point instead of 36.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/cia.h"

#define CIA_MAX_PHASES 8

namespace cia {

struct PhaseConst {
  int32_t kind;
  int32_t pad;
  double G, K, lam;   // shear, bulk, Lame lambda
  double sy, H, tau, sref, m, hs;  // hs = 3G + H
  double density;
};

struct PhaseTable {
  int32_t n;
  int32_t pad;
  PhaseConst p[CIA_MAX_PHASES];
};

struct Settings {
  int32_t max_iterations;
  int32_t pad;
  double tol_rel_yield, res_abs, res_rel;
};

// Compact tangent (see header comment).
template <class T>
struct Tangent {
  T K, d, r;
  T n[6];
};

template <class T>
struct PointOut {
  T sig[6];
  T epsp[6];
  T qbar, gam;
  Tangent<T> tan;
};

// Extra outputs (energies and the reference diagnostics), only when asked for.
template <class T>
struct PointExtra {
  T psi, diss;
  T d[CIA_N_DIAG];
};

__host__ __device__ __forceinline__ bool finite(double x) { return isfinite(x); }

template <class T>
__host__ __device__ __forceinline__ void elastic_stress(const PhaseConst& P, const T e[6], T s[6]) {
  const T lam = (T)P.lam, g = (T)P.G, g2 = (T)(2.0 * P.G);
  const T tr = e[0] + e[1] + e[2];
  s[0] = lam * tr + g2 * e[0];
  s[1] = lam * tr + g2 * e[1];
  s[2] = lam * tr + g2 * e[2];
  s[3] = g * e[3];
  s[4] = g * e[4];
  s[5] = g * e[5];
}

// 1.5 s:s with stress weights (1,1,1,2,2,2) of the deviator of s.
template <class T>
__host__ __device__ __forceinline__ T j2_radicand(const T s[6], T dev[6]) {
  const T p = (s[0] + s[1] + s[2]) / (T)3;
  dev[0] = s[0] - p;
  dev[1] = s[1] - p;
  dev[2] = s[2] - p;
  dev[3] = s[3];
  dev[4] = s[4];
  dev[5] = s[5];
  return (T)1.5 * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                   (T)2 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5]));
}

template <class T>
__host__ __device__ __forceinline__ void tangent_apply(const Tangent<T>& c, const T e[6], T s[6]) {
  const T tr = e[0] + e[1] + e[2];
  const T t3 = tr / (T)3;
  const T kt = c.K * tr;
  const T ne = c.n[0] * e[0] + c.n[1] * e[1] + c.n[2] * e[2] + c.n[3] * e[3] + c.n[4] * e[4] +
               c.n[5] * e[5];
  const T rn = c.r * ne;
  s[0] = kt + c.d * (e[0] - t3) + rn * c.n[0];
  s[1] = kt + c.d * (e[1] - t3) + rn * c.n[1];
  s[2] = kt + c.d * (e[2] - t3) + rn * c.n[2];
  s[3] = c.d * (T)0.5 * e[3] + rn * c.n[3];
  s[4] = c.d * (T)0.5 * e[4] + rn * c.n[4];
  s[5] = c.d * (T)0.5 * e[5] + rn * c.n[5];
}

// Entry (i, j) of the compact tangent as a 6x6 matrix.
template <class T>
__host__ __device__ __forceinline__ T tangent_entry(const Tangent<T>& c, int i, int j) {
  T v = c.r * c.n[i] * c.n[j];
  if (i < 3 && j < 3) v += c.K + c.d * ((i == j ? (T)1 : (T)0) - (T)1 / (T)3);
  if (i >= 3 && i == j) v += c.d * (T)0.5;
  return v;
}

// Yield-consistency scalar solve, plasticity.py:711-854, always FP64.  M1: the
// caller guarantees rate exponent m = 1 (closed form only; the Newton iteration
// for other exponents is not compiled, which keeps the point update lean).
template <bool M1 = false>
__host__ __device__ inline int solve_multiplier(const PhaseConst& P, const Settings& S, double dt,
                                       double f, double& dg, int& iters, double& res_out,
                                       cia_point_failure* rec) {
  const double hs = P.hs;
  const double ub = f / hs;
  const double tol = fmax(S.res_abs, S.res_rel * ub);
  const double ts = P.tau * P.sref;
  if (M1 || P.m == 1.0) {
    const double den = ts / dt + hs;
    const double mult = f / den;
    const double corr = f - hs * mult;
    const double res = mult - dt * corr / ts;
    const double der = 1.0 + dt * hs / ts;
    if (rec) {
      rec->iteration = 0;
      rec->plastic_multiplier = mult;
      rec->corrected_overstress = corr;
      rec->residual = res;
      rec->derivative = der;
      rec->newton_candidate = mult;
      rec->physical_upper_bound = ub;
      rec->residual_tolerance = tol;
    }
    if (!(finite(den) && finite(mult) && finite(corr) && finite(res) && finite(der)))
      return CIA_PT_NON_FINITE_CLOSED_FORM;
    if (!(0.0 < mult && mult < ub) || corr <= 0.0) return CIA_PT_CLOSED_FORM_SOLUTION_OUT_OF_BOUNDS;
    dg = mult;
    iters = 0;
    res_out = res;
    return CIA_PT_OK;
  }
  if (rec) {
    rec->physical_upper_bound = ub;
    rec->residual_tolerance = tol;
  }
  if (S.max_iterations <= 0) return CIA_PT_INVALID_ITERATION_LIMIT;
  if (!(finite(ub) && finite(tol))) return CIA_PT_NON_FINITE_PROBLEM_SCALE;
  double x = 0.5 * ub;
  for (int it = 1; it <= S.max_iterations; ++it) {
    const double over = f - hs * x;
    const double nover = over / P.sref;
    if (rec) {
      rec->iteration = it;
      rec->plastic_multiplier = x;
      rec->corrected_overstress = over;
      rec->residual = NAN;
      rec->derivative = NAN;
      rec->newton_candidate = NAN;
    }
    if (!finite(nover) || nover <= 0.0) return CIA_PT_INVALID_CORRECTED_OVERSTRESS;
    const double rate = pow(nover, P.m) / P.tau;
    const double res = x - dt * rate;
    const double der = 1.0 + dt * P.m * hs * pow(nover, P.m - 1.0) / ts;
    if (rec) {
      rec->residual = res;
      rec->derivative = der;
    }
    if (!(finite(rate) && finite(res) && finite(der))) return CIA_PT_NON_FINITE_ITERATION_VALUE;
    if (der <= 0.0) return CIA_PT_NON_POSITIVE_DERIVATIVE;
    if (fabs(res) <= tol) {
      if (!(0.0 < x && x < ub) || over <= 0.0) return CIA_PT_CONVERGED_SOLUTION_OUT_OF_BOUNDS;
      dg = x;
      iters = it;
      res_out = res;
      return CIA_PT_OK;
    }
    const double xn = x - res / der;
    if (rec) rec->newton_candidate = xn;
    if (!finite(xn)) return CIA_PT_NON_FINITE_NEWTON_CANDIDATE;
    if (!(0.0 < xn && xn < ub)) return CIA_PT_NEWTON_CANDIDATE_OUT_OF_BOUNDS;
    if (it == S.max_iterations) return CIA_PT_MAXIMUM_ITERATIONS_EXCEEDED;
    x = xn;
  }
  return CIA_PT_MAXIMUM_ITERATIONS_EXCEEDED;
}

// Full J2 viscoplastic point update.  Returns a CIA_PT_* status.
// rate == nullptr disables the increment diagnostics (strain_rates None).
template <class T, bool EXTRA, bool M1 = false>
__host__ __device__ inline int j2vp_point(const PhaseConst& P, const Settings& S, double dt, const T eps[6],
                                 const T* rate, const T epsp0[6], T qbar0, T gam0,
                                 PointOut<T>& o, PointExtra<T>* ex, cia_point_failure* rec) {
  T e[6], trial[6], dev[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) e[i] = eps[i] - epsp0[i];
  elastic_stress(P, e, trial);
  const T rad = j2_radicand(trial, dev);
  if (rec) {
    rec->radicand = (double)rad;
    for (int i = 0; i < 6; ++i) {
      rec->trial_stress[i] = (double)trial[i];
      rec->strain[i] = (double)eps[i];
    }
    rec->equivalent_plastic_strain = (double)qbar0;
  }
  if (!isfinite((double)rad)) return CIA_PT_NON_FINITE_J2_INVARIANT;
  if (rad < (T)0) return CIA_PT_NEGATIVE_J2_INVARIANT;
  const T q = sqrt(rad);
  const T radius = (T)P.sy + (T)P.H * qbar0;
  const T f = q - radius;
  const T scale = fmax(fmax((T)P.sy, (T)P.sref), q);
  const T thr = (T)S.tol_rel_yield * scale;
  const T margin = (f - thr) / scale;
  if (rec) {
    rec->trial_yield = (double)f;
    rec->trial_equivalent_stress = (double)q;
    rec->yield_radius = (double)radius;
  }
  if (!isfinite(S.tol_rel_yield) || S.tol_rel_yield < 0.0 || !isfinite((double)thr))
    return CIA_PT_INVALID_YIELD_ACTIVATION_TOLERANCE;

  o.tan.K = (T)P.K;
  const bool elastic = f <= thr;
  T dep[6];
  double dg = 0.0, lres = 0.0, hv = 0.0;
  int iters = 0;
  if (elastic) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      o.sig[i] = trial[i];
      o.epsp[i] = epsp0[i];
      o.tan.n[i] = (T)0;
      dep[i] = (T)0;
    }
    o.qbar = qbar0;
    o.gam = gam0;
    o.tan.d = (T)(2.0 * P.G);
    o.tan.r = (T)0;
  } else {
    T n[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) n[i] = (T)1.5 * dev[i] / q;
    const int st = solve_multiplier<M1>(P, S, dt, (double)f, dg, iters, lres, rec);
    if (st != CIA_PT_OK) return st;
    const T dgT = (T)dg;
    const T g2dg = (T)(2.0 * P.G) * dgT;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      o.sig[i] = trial[i] - g2dg * n[i];
      dep[i] = dgT * (i < 3 ? n[i] : (T)2 * n[i]);
      o.epsp[i] = epsp0[i] + dep[i];
      o.tan.n[i] = n[i];
    }
    o.qbar = qbar0 + dgT;
    o.gam = gam0 + dgT;
    if (M1 || P.m == 1.0) {
      hv = P.tau * P.sref / dt + P.hs;
    } else {
      T fdev[6];
      const T fq = sqrt(fmax(j2_radicand(o.sig, fdev), (T)0));
      const double fy = (double)(fq - ((T)P.sy + (T)P.H * o.qbar));
      if (fy <= 0.0) return CIA_PT_NON_POSITIVE_TANGENT_OVERSTRESS;
      const double fder = P.m * pow(fy / P.sref, P.m - 1.0) / (P.tau * P.sref);
      hv = 1.0 / (dt * fder) + P.hs;
    }
    const double G = P.G;
    o.tan.d = (T)(2.0 * G - 6.0 * G * G * dg / (double)q);
    o.tan.r = (T)(4.0 * G * G * (dg / (double)q - 1.0 / hv));
  }

  if (EXTRA) {
    // energies, plasticity.py:1312-1315 + 588/662-672
    T ee[6], cee[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) ee[i] = eps[i] - o.epsp[i];
    elastic_stress(P, ee, cee);
    T psi = (T)0;
#pragma unroll
    for (int i = 0; i < 6; ++i) psi += cee[i] * ee[i];
    psi = (T)0.5 * psi + (T)0.5 * (T)P.H * o.qbar * o.qbar;
    T pwork = (T)0;
#pragma unroll
    for (int i = 0; i < 6; ++i) pwork += o.sig[i] * dep[i];
    T fdev[6];
    const T fq = sqrt(fmax(j2_radicand(o.sig, fdev), (T)0));
    const T fy = fq - ((T)P.sy + (T)P.H * o.qbar);
    const T pp = elastic ? (T)0 : pwork / (T)dt;
    const T hr = elastic ? (T)0 : (T)P.H * o.qbar * (T)dg / (T)dt;
    ex->psi = psi;
    ex->diss = elastic ? (T)0 : pp - hr;
    T* d = ex->d;
    d[CIA_D_YIELD_VALUE] = elastic ? f : fy;
    d[CIA_D_PLASTIC_WORK] = elastic ? (T)0 : pwork;
    d[CIA_D_LOCAL_ITERATIONS] = (T)iters;
    d[CIA_D_ACTIVE_SURFACE] = elastic ? (T)0 : (T)4;
    d[CIA_D_PLASTIC_MULTIPLIER] = (T)dg;
    d[CIA_D_VISCOPLASTIC_MULTIPLIER] = (T)dg;
    d[CIA_D_EQUIVALENT_STRESS] = elastic ? q : fq;
    d[CIA_D_TRIAL_YIELD_VALUE] = f;
    d[CIA_D_YIELD_SCALE] = scale;
    d[CIA_D_YIELD_ACTIVATION_THRESHOLD] = thr;
    d[CIA_D_YIELD_ACTIVATION_MARGIN] = margin;
    d[CIA_D_VISCOPLASTIC_ACTIVE] = elastic ? (T)0 : (T)1;
    d[CIA_D_TIME_STEP] = (T)dt;
    d[CIA_D_EQUIVALENT_PLASTIC_STRAIN] = o.qbar;
    d[CIA_D_VISCOPLASTIC_HV] = elastic ? (T)0 : (T)hv;
    d[CIA_D_PLASTIC_POWER] = pp;
    d[CIA_D_HARDENING_STORAGE_RATE] = hr;
    d[CIA_D_LOCAL_RESIDUAL] = elastic ? (T)0 : (T)lres;
    d[CIA_D_ROOT_LOWER_BOUND] = (T)0;
    d[CIA_D_ROOT_UPPER_BOUND] = elastic ? (T)0 : (T)((double)f / P.hs);
    if (rate) {
      T de[6], pe[6], ps[6], pdev[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        de[i] = (T)dt * rate[i];
        pe[i] = eps[i] - de[i] - epsp0[i];
      }
      elastic_stress(P, pe, ps);
      const T prad = j2_radicand(ps, pdev);
      if (!isfinite((double)prad)) return CIA_PT_NON_FINITE_J2_INVARIANT;
      T w = (T)0;
#pragma unroll
      for (int i = 0; i < 6; ++i) w += (ps[i] + o.sig[i]) * de[i];
      d[CIA_D_INCREMENTAL_WORK_DENSITY] = (T)0.5 * w;
      d[CIA_D_EQUIVALENT_STRESS_INCREMENT] = fq - sqrt(fmax(prad, (T)0));
    } else {
      d[CIA_D_INCREMENTAL_WORK_DENSITY] = (T)NAN;
      d[CIA_D_EQUIVALENT_STRESS_INCREMENT] = (T)NAN;
    }
  }
  return CIA_PT_OK;
}

// Linear elastic point (phase member): stress, compact tangent, energies and
// the elastic-phase diagnostics of phased.py:171-190 (NaN elsewhere, as the
// NaN-filled union of phased.py:125-134).
template <class T, bool EXTRA>
__host__ __device__ inline int elastic_point(const PhaseConst& P, double dt, const T eps[6], const T* rate,
                                    const T epsp0[6], T qbar0, T gam0, PointOut<T>& o,
                                    PointExtra<T>* ex) {
  elastic_stress(P, eps, o.sig);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    o.epsp[i] = epsp0[i];
    o.tan.n[i] = (T)0;
  }
  o.qbar = qbar0;
  o.gam = gam0;
  o.tan.K = (T)P.K;
  o.tan.d = (T)(2.0 * P.G);
  o.tan.r = (T)0;
  if (EXTRA) {
    T psi = (T)0;
#pragma unroll
    for (int i = 0; i < 6; ++i) psi += eps[i] * o.sig[i];
    ex->psi = (T)0.5 * psi;
    ex->diss = (T)0;
#pragma unroll
    for (int i = 0; i < CIA_N_DIAG; ++i) ex->d[i] = (T)NAN;
    if (rate) {
      T de[6], dsig[6], prev[6], dv[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) de[i] = (T)dt * rate[i];
      elastic_stress(P, de, dsig);
      T w = (T)0;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        prev[i] = o.sig[i] - dsig[i];
        w += (prev[i] + o.sig[i]) * de[i];
      }
      ex->d[CIA_D_INCREMENTAL_WORK_DENSITY] = (T)0.5 * w;
      const T qc = sqrt(j2_radicand(o.sig, dv));
      const T qp = sqrt(j2_radicand(prev, dv));
      ex->d[CIA_D_EQUIVALENT_STRESS_INCREMENT] = qc - qp;
      ex->d[CIA_D_VISCOPLASTIC_ACTIVE] = (T)0;
    }
  }
  return CIA_PT_OK;
}

// M1: every J2 phase the caller passes has rate exponent m = 1.
template <class T, bool EXTRA, bool M1 = false>
__host__ __device__ __forceinline__ int material_point(const PhaseConst& P, const Settings& S, double dt,
                                              const T eps[6], const T* rate, const T epsp0[6],
                                              T qbar0, T gam0, PointOut<T>& o, PointExtra<T>* ex,
                                              cia_point_failure* rec) {
  if (P.kind == CIA_MAT_J2_VISCOPLASTIC)
    return j2vp_point<T, EXTRA, M1>(P, S, dt, eps, rate, epsp0, qbar0, gam0, o, ex, rec);
  return elastic_point<T, EXTRA>(P, dt, eps, rate, epsp0, qbar0, gam0, o, ex);
}

}  // namespace cia
