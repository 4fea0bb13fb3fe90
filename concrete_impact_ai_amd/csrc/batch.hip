// Batched periodic-RVE microequilibrium: many independent structured RVEs of
// different sizes solved together on one MI355X.
//
// Reference: the per-point RVE solves behind RVEHomogenizedMaterial.update
// (src/fem/rve/material.py:104-191, worker loop :267-291) and the path driver
// iterate_prescribed_rve_path (src/fem/rve/path.py:47-100), each RVE increment
// being solve_rve_microequilibrium (src/fem/rve/solver.py:62-227) with the
// Armijo search of src/fem/solvers/line_search.py:40-67.  The reference runs
// the RVEs one after another (or in a process pool); here they share every
// kernel launch.
//
// Layout.  The RVEs are concatenated in one arena: RVE r owns elements (=
// periodic classes) [e0, e0 + E) and points [8 e0, 8 (e0 + E)).  Per-point rows
// are component-major over the whole arena ([c][NPT]), class vectors are
// class-major ([3 CT]), the packed spectra of RVE r occupy [2 e0, 2 e0 + 2E).
//
// Tiles.  Every kernel works on tiles of one family (Gauss-point lanes, classes
// -- one per element, the lanes of the element pass -- or modes, DFT x-lines, DFT
// y/z line groups); a tile never straddles two RVEs,
// so a tile's block reduction is a partial of exactly one RVE, stored at the
// tile's index.  A kernel's blocks walk the family's tiles in rounds of
// gridDim.x tiles; within a round each XCD takes one contiguous range (the
// neighbour rows of a stencil stay in that XCD's L2).  The grid of each kernel
// is chosen by the autotuner (one tile per block up to persistent grids of a
// few blocks per CU).
//
// Control.  Each RVE carries a Newton state machine on the device (NEWTON ->
// PCG -> LINE_SEARCH -> NEWTON ... -> DONE | FAILED).  Small control kernels
// (one wavefront per RVE) reduce that RVE's tile partials in a fixed order and
// advance its state with the reference's tolerance formulas and acceptance
// tests; tiles of RVEs in another state exit at once.  All RVEs take their
// Newton iterations in lockstep, so the host only counts states between phases.
// PCG iterations (spectral preconditioner) are captured as a HIP graph of two
// iterations for the whole batch and replayed until every RVE has converged.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "dft.hpp"
#include "hex8.hpp"
#include "symbol.hpp"

namespace cia {
namespace bt {

enum { F_GP = 0, F_CL = 1, F_DX = 2, F_DY = 3, F_DZ = 4, NF = 5 };
enum { S_NEWTON = 0, S_PCG = 1, S_LS = 2, S_DONE = 3, S_FAILED = 4 };

// Rows of the tile-partial array.
enum {
  P_RR = 0,    // gather: ||r||^2
  P_WW = 1,    // gather: ||w||^2
  P_NORM = 2,  // ||dw||^2
  P_PQ = 3,    // p.y
  P_UPD = 4,   // update: r.r, sum r (3)
  P_RZ = 8,    // r.z
  P_RESP = 9,  // response sums (B_NRESP rows)
  B_NRESP = 13,
  P_PHASE = P_RESP + B_NRESP,  // per-phase sums (B_NPHASE_ROWS rows)
  B_PHASES = 4,                // phases with per-phase sums
  B_NPHASE_ROWS = 9 * B_PHASES,
  P_ROWS = P_PHASE + B_NPHASE_ROWS
};

// Response sums per RVE (final pass): stress (6), free energy, dissipation,
// micro work, positive / negative incremental work, unloading and active point
// counts.
enum { BR_SIG = 0, BR_PSI = 6, BR_DISS = 7, BR_MICRO = 8, BR_POS = 9, BR_NEG = 10, BR_UNLOAD = 11,
       BR_ACTIVE = 12 };
// Per-phase sums, 9 per phase: points, stress (6), free energy, dissipation.

struct Desc {
  int nx, ny, nz, lines_x;
  int64_t e0, E;
  BoxGeom g;
  DftAxis ax, ay, az;
  const hipFloatComplex* kinv;
  int t0[NF];
  int nt[NF];
  int j2mask;  // bit p set: phase p is J2 viscoplastic (carries state rows)
  int pad;
  double inv_n;
  double volume;
};

struct Req {
  double e[6];   // requested macro strain
  double ec[6];  // committed macro strain
  double dt;
  double pad;
};

struct Ctl {
  int32_t state, reason, it, backtracks;
  int32_t bt, cur, pcg_its, p_done;
  int32_t p_iters, fail_it, fail_slot, pad0;
  unsigned long long first_fail;  // lowest failing local point id of the last assembly
  int64_t gp_updates;
  double r0, rn, wn2, last_inc;
  double rtol, itol, inc_n, alpha;
  double lin_rtol, cn;
  double p_tol2, p_rz, p_rz_old, p_alpha, p_beta, p_rr;
  double p_sumr[3];
  double fail_rn, fail_rtol, fail_inc, fail_itol;
};

// Everything a batched kernel needs, passed by value.
struct Arena {
  const Desc* desc;
  const uint16_t* t2r[NF];
  int T[NF];
  int R;
  int Tmax;
  int force;  // autotuner: every RVE is active, whatever its state
  int pad;
  double lin_atol;  // absolute tolerance of the PCG residual
  Ctl* ctl;
  const Req* req;
  const PhaseTable* tabs;
  const uint8_t* ephase;  // [ET]
  int64_t NPT, ET;        // total points, elements (= classes)
  void* st[2];
  void* sig[2];
  void* tan[2];
  void* F;  // [24][ET]
  void* pd;  // [2][NPT] free energy and dissipation of the final pass
  double* w[2];
  double* r[2];
  double* dw;
  double* pr;
  double* pp[2];
  double* py;
  void* uab;     // packed spectra [2 ET] (double2 for FP64 batches, float2 for FP32)
  double* part;  // [P_ROWS][Tmax]
  int* act[NF];  // active-tile lists of the PCG kernels, [T[f]] each
  int* n_act;    // [NF] list lengths
  int* act_off;  // [NF][R] offset of every RVE's tiles in the lists
  const int* grp;     // leader tiles of the final pass's Gauss-point tile groups
  const int* grp_cl;  // leader tiles of the per-phase sums' class tile groups
  int n_grp, n_grp_cl;
};

struct NewtonP {
  cia_newton_settings ns;
  double lin_rtol, lin_atol, forcing_max;
  int adaptive;
  int pad;
};

// Blocks walk the tiles in rounds of gridDim.x; xcd_block() hands each XCD one
// contiguous range of every round.
#define CIA_TILES(TOT)                                                           \
  for (int base_ = 0; base_ < (TOT); base_ += (int)gridDim.x)                    \
    for (int t = base_ + (int)xcd_block(), once_ = 0; once_ == 0 && t < (TOT); once_ = 1)

__device__ __forceinline__ bool pcg_active(const Arena& A, const Ctl& c) {
  return A.force || (c.state == S_PCG && c.p_done == 0);
}

// The PCG, assembly, gather and line-search kernels walk only the listed tiles:
// those of the RVEs whose PCG was running (or that were in the kernel's Newton
// state) at the last compaction, every tile while the autotuner forces all RVEs
// active, in rounds of gridDim.x as CIA_TILES walks all tiles.  RVEs only stop
// during a PCG solve, so the list is a superset of the active tiles and a listed
// tile still checks its RVE; results do not depend on which block takes a tile.
#define CIA_ACTIVE_TILES(FAM)                                                             \
  for (int base_ = 0, n_ = A.n_act[FAM]; base_ < n_; base_ += (int)gridDim.x)            \
    for (int i_ = base_ + (int)xcd_block(), once_ = 0; once_ == 0 && i_ < n_; once_ = 1) \
      for (int t = A.act[FAM][i_], one_ = 0; one_ == 0; one_ = 1)

// RVEs whose tiles a list holds: want < 0, those with a running PCG; otherwise
// those in Newton state `want` (the assembly, gather and line-search kernels).
__device__ __forceinline__ bool listed(const Arena& A, const Ctl& c, int want) {
  return want < 0 ? pcg_active(A, c) : (A.force || c.state == want);
}

// Active-tile lists, step 1 (one block): per family, the offset of every RVE's
// tiles (exclusive scan over the listed RVEs, in RVE order) and the list lengths.
__global__ void __launch_bounds__(1024) k_b_compact_scan(Arena A, int want) {
  __shared__ int wsum[16][NF];
  __shared__ int carry[NF];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid < NF) carry[tid] = 0;
  __syncthreads();
  for (int r0 = 0; r0 < A.R; r0 += 1024) {
    const int r = r0 + tid;
    const bool on = r < A.R && listed(A, A.ctl[r], want);
    int v[NF], x[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) x[f] = v[f] = on ? A.desc[r].nt[f] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int y = __shfl_up(x[f], d, 64);
        if (lane >= d) x[f] += y;
      }
    }
    if (lane == 63) {
#pragma unroll
      for (int f = 0; f < NF; ++f) wsum[wid][f] = x[f];
    }
    __syncthreads();
    if (r < A.R) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        int pre = carry[f];
        for (int u = 0; u < wid; ++u) pre += wsum[u][f];
        A.act_off[f * A.R + r] = pre + x[f] - v[f];
      }
    }
    __syncthreads();
    if (tid < NF) {
      int tot = 0;
      for (int u = 0; u < 16; ++u) tot += wsum[u][tid];
      carry[tid] += tot;
    }
    __syncthreads();
  }
  if (tid < NF) A.n_act[tid] = carry[tid];
}

// Active-tile lists in one launch for batches of at most 1024 RVEs (one block per
// RVE): every block sums the tile counts of the listed RVEs before it itself and
// writes its tile ids; the last block writes the list lengths.  Same lists as
// k_b_compact_scan + k_b_compact_fill.
__global__ void __launch_bounds__(1024) k_b_compact_small(Arena A, int want) {
  __shared__ int wsum[16][NF];
  __shared__ int off[NF];
  const int r = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int v[NF];
  const bool mine = listed(A, A.ctl[r], want);
  const bool on = tid < r && listed(A, A.ctl[tid], want);  // RVE tid precedes r
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    int x = on ? A.desc[tid].nt[f] : 0;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    v[f] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int f = 0; f < NF; ++f) wsum[wid][f] = v[f];
  }
  __syncthreads();
  if (tid < NF) {
    int o = 0;
    for (int u = 0; u < 16; ++u) o += wsum[u][tid];
    off[tid] = o;
    if (r == A.R - 1) A.n_act[tid] = o + (mine ? A.desc[r].nt[tid] : 0);
  }
  __syncthreads();
  if (!mine) return;
  const Desc& D = A.desc[r];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    int* out = A.act[f] + off[f];
    for (int j = tid; j < D.nt[f]; j += blockDim.x) out[j] = D.t0[f] + j;
  }
}

// Active-tile lists, step 2 (one block per RVE): the tile ids of the RVE.
__global__ void __launch_bounds__(kBlock) k_b_compact_fill(Arena A, int want) {
  const int r = blockIdx.x;
  if (!listed(A, A.ctl[r], want)) return;
  const Desc& D = A.desc[r];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    int* out = A.act[f] + A.act_off[f * A.R + r];
    for (int j = threadIdx.x; j < D.nt[f]; j += blockDim.x) out[j] = D.t0[f] + j;
  }
}

// ---------------------------------------------------------------------------
// Assembly: strain -> constitutive update -> element force rows, one lane per
// Gauss point (as k_assemble of the single-RVE solver).  `want` selects the
// RVEs: S_NEWTON (first assembly, slot cur) or S_LS (line-search candidate,
// slot 1 - cur).  M1: every J2 phase of the batch has rate exponent 1.
template <class T, bool M1>
__global__ void __launch_bounds__(kBlock)
    k_b_assemble(Arena A, Settings S, int want, const T* __restrict__ state_c,
                 const double* __restrict__ wc) {
  __shared__ double cw[kBlock][3];  // corner fluctuations (differenced in FP64, gp_strain_sf)
  __shared__ T cs[kBlock][8];       // the points' stresses, for the corner forces (rows padded)
  const int64_t NPT = A.NPT, ET = A.ET;
  CIA_ACTIVE_TILES(F_GP) {
    const int r = A.t2r[F_GP][t];
    const Desc& D = A.desc[r];
    Ctl& c = A.ctl[r];
    if (!(A.force || c.state == want)) continue;
    const int slot = want == S_LS ? 1 - c.cur : c.cur;
    const double* w = A.w[slot];
    const Req& M = A.req[r];
    const double dt = M.dt;
    const int64_t p = (int64_t)(t - D.t0[F_GP]) * kBlock + threadIdx.x;  // local point
    const int64_t e = p >> 3;
    const int q = (int)(p & 7);
    const bool valid = e < D.E;
    const int64_t gp = 8 * D.e0 + p;
    const int64_t ge = D.e0 + e;
    // the committed state is loaded before the corner exchange, so its latency
    // overlaps the corner gather's
    T epsp[6];
    T qb = (T)0, ga = (T)0;
    bool j2 = false;
    if (valid) {
      const int64_t cq = D.e0 + periodic_corner(e, q, D.nx, D.ny, D.nz);
#pragma unroll
      for (int m = 0; m < 3; ++m) cw[threadIdx.x][m] = w[3 * cq + m];
      j2 = (D.j2mask >> A.ephase[ge]) & 1;
      if (j2) {
#pragma unroll
        for (int m = 0; m < 6; ++m) epsp[m] = state_c[m * NPT + gp];
        qb = state_c[6 * NPT + gp];
        ga = state_c[7 * NPT + gp];
      } else {
#pragma unroll
        for (int m = 0; m < 6; ++m) epsp[m] = (T)0;
      }
    }
    __syncthreads();
    if (valid) {
      T eps[6];
      gp_strain_sf<T, double, 3>(&cw[threadIdx.x & ~7], 0, q, D.g, eps);
#pragma unroll
      for (int m = 0; m < 6; ++m) eps[m] += (T)M.e[m];
      const PhaseConst& P = A.tabs[r].p[A.ephase[ge]];  // read where used: fewer live registers
      PointOut<T> o;
      const int st = material_point<T, false, M1>(P, S, dt, eps, nullptr, epsp, qb, ga, o, nullptr,
                                                  nullptr);
      if (st != CIA_PT_OK && !A.force) atomicMin(&c.first_fail, (unsigned long long)p);
      T* sig = (T*)A.sig[slot];
      T* tan = (T*)A.tan[slot];
      T* st_t = (T*)A.st[slot];
#pragma unroll
      for (int m = 0; m < 6; ++m) sig[m * NPT + gp] = o.sig[m];
      if (j2) {
#pragma unroll
        for (int m = 0; m < 6; ++m) st_t[m * NPT + gp] = o.epsp[m];
        st_t[6 * NPT + gp] = o.qbar;
        st_t[7 * NPT + gp] = o.gam;
      }
      tan[0 * NPT + gp] = o.tan.d;
      tan[1 * NPT + gp] = o.tan.r;
      if (o.tan.r != (T)0) {
#pragma unroll
        for (int m = 0; m < 6; ++m) tan[(2 + m) * NPT + gp] = o.tan.n[m];
      }
#pragma unroll
      for (int m = 0; m < 6; ++m) cs[threadIdx.x][m] = o.sig[m];
    }
    __syncthreads();
    if (valid) {
      // lane q: the element force at corner q from the 8 points' stresses
      T fin[3];
      corner_force<T>(&cs[threadIdx.x & ~7], q, D.g, fin);
#pragma unroll
      for (int m = 0; m < 3; ++m) ((T*)A.F)[(3 * q + m) * ET + ge] = fin[m];
    }
    __syncthreads();  // cw and cs are reused by the next tile
  }
}

// Final pass of the converged RVEs, without a second constitutive update: the
// assembly at the converged fluctuation (slot cur) already holds the stresses,
// the compact tangent and the next state, so this pass only forms the strain
// and strain rate, the energies and the increment diagnostics of the reference's
// material point (plasticity.py:1312-1315, 588, 662-672) from the stored values
// -- the plastic strain increment and multiplier as differences of the next and
// the committed state, the viscoplastic flag from the tangent's rank-one
// coefficient (nonzero exactly on the plastic branch) -- copies the next state
// out (elastic points: zero rows) and reduces the response sums.  A block takes
// groups of kFinalGroup consecutive Gauss-point tiles of one RVE (the leader
// tiles listed in A.grp): each lane accumulates its points over the group and
// the group's sums go to the leader tile's partials (k_b_resp reads those only),
// one reduction per group instead of per tile, independent of the grid.
constexpr int kFinalGroup = 16;

template <class T>
__global__ void __launch_bounds__(kBlock, 2)
    k_b_final(Arena A, const T* __restrict__ state_c, T* __restrict__ next_state,
              const double* __restrict__ wc) {
  __shared__ double cw[kBlock][6];  // corner fluctuation and its increment (as the assembly)
  __shared__ double red[kBlock / 64][B_NRESP];
  const int64_t NPT = A.NPT;
  for (int base_ = 0; base_ < A.n_grp; base_ += (int)gridDim.x) {
    const int gi = base_ + (int)xcd_block();
    if (gi >= A.n_grp) break;
    const int tl = A.grp[gi];
    const int r = A.t2r[F_GP][tl];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_DONE) continue;
    const int tend = min(tl + kFinalGroup, D.t0[F_GP] + D.nt[F_GP]);
    const int slot = c.cur;
    const double* w = A.w[slot];
    const Req& M = A.req[r];
    const double dt = M.dt;
    const T* sig = (const T*)A.sig[slot];
    const T* stn = (const T*)A.st[slot];
    const T* tanr = (const T*)A.tan[slot] + NPT;
    double acc[B_NRESP];
#pragma unroll
    for (int s2 = 0; s2 < B_NRESP; ++s2) acc[s2] = 0.0;
    for (int t = tl; t < tend; ++t) {
      const int64_t p = (int64_t)(t - D.t0[F_GP]) * kBlock + threadIdx.x;
      const int64_t e = p >> 3;
      const int q = (int)(p & 7);
      const bool valid = e < D.E;
      const int64_t gp = 8 * D.e0 + p;
      // the point's stored values are loaded before the corner exchange
      T s[6], ep[6], ep0[6], qb = (T)0, qb0 = (T)0, ga = (T)0;
      bool j2 = false, active = false;
      if (valid) {
        const int64_t cq = D.e0 + periodic_corner(e, q, D.nx, D.ny, D.nz);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const double wv = w[3 * cq + m];
          cw[threadIdx.x][m] = wv;
          cw[threadIdx.x][3 + m] = wv - wc[3 * cq + m];
        }
#pragma unroll
        for (int m = 0; m < 6; ++m) s[m] = sig[m * NPT + gp];
        j2 = (D.j2mask >> A.ephase[D.e0 + e]) & 1;
        if (j2) {
#pragma unroll
          for (int m = 0; m < 6; ++m) {
            ep[m] = stn[m * NPT + gp];
            ep0[m] = state_c[m * NPT + gp];
          }
          qb = stn[6 * NPT + gp];
          ga = stn[7 * NPT + gp];
          qb0 = state_c[6 * NPT + gp];
          active = tanr[gp] != (T)0;
        } else {
#pragma unroll
          for (int m = 0; m < 6; ++m) ep[m] = ep0[m] = (T)0;
        }
      }
      __syncthreads();
      if (valid) {
        T eps[6], rate[6];
        // the assembly's strain (same corner values, same operations) and the
        // strain increment of the step
        gp_strain_sf<T, double, 6>(&cw[threadIdx.x & ~7], 0, q, D.g, eps);
        gp_strain_sf<T, double, 6>(&cw[threadIdx.x & ~7], 3, q, D.g, rate);
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          eps[m] += (T)M.e[m];
          rate[m] = (rate[m] + (T)(M.e[m] - M.ec[m])) / (T)dt;
        }
        const PhaseConst& P = A.tabs[r].p[A.ephase[D.e0 + e]];
        if (j2) {
#pragma unroll
          for (int m = 0; m < 6; ++m) next_state[m * NPT + gp] = ep[m];
          next_state[6 * NPT + gp] = qb;
          next_state[7 * NPT + gp] = ga;
        } else {
#pragma unroll
          for (int m = 0; m < 8; ++m) next_state[m * NPT + gp] = (T)0;
        }
        // free energy: elastic strain energy + hardening storage
        T ee[6], cee[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) ee[m] = eps[m] - ep[m];
        elastic_stress(P, ee, cee);
        T psi = (T)0;
#pragma unroll
        for (int m = 0; m < 6; ++m) psi += cee[m] * ee[m];
        psi = (T)0.5 * psi + (T)0.5 * (T)P.H * qb * qb;
        // dissipation: plastic power minus hardening storage rate
        T diss = (T)0;
        if (active) {
          T pwork = (T)0;
#pragma unroll
          for (int m = 0; m < 6; ++m) pwork += s[m] * (ep[m] - ep0[m]);
          diss = pwork / (T)dt - (T)P.H * qb * (qb - qb0) / (T)dt;
        }
        // incremental work density: trapezoid between the elastic trial at the
        // previous strain (committed plastic strain) and the converged stress
        T de[6], pe[6], ps[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          de[m] = (T)dt * rate[m];
          pe[m] = eps[m] - de[m] - ep0[m];
        }
        elastic_stress(P, pe, ps);
        T wsum = (T)0;
#pragma unroll
        for (int m = 0; m < 6; ++m) wsum += (ps[m] + s[m]) * de[m];
        const double pw = (double)((T)0.5 * wsum);
        ((T*)A.pd)[gp] = psi;
        ((T*)A.pd)[NPT + gp] = diss;
        double mw = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          acc[BR_SIG + m] += (double)s[m];
          mw += (double)s[m] * (dt * (double)rate[m]);
        }
        acc[BR_PSI] += (double)psi;
        acc[BR_DISS] += (double)diss;
        acc[BR_MICRO] += mw;
        acc[BR_POS] += fmax(pw, 0.0);
        acc[BR_NEG] += fmin(pw, 0.0);
        acc[BR_UNLOAD] += pw < 0.0 ? 1.0 : 0.0;
        acc[BR_ACTIVE] += active ? 1.0 : 0.0;
      }
      __syncthreads();  // cw is reused by the next tile
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s2 = 0; s2 < B_NRESP; ++s2) {
      const double v = wave_sum(acc[s2]);
      if (lane == 0) red[wid][s2] = v;
    }
    __syncthreads();
    if (threadIdx.x < B_NRESP) {
      double v = 0.0;
#pragma unroll
      for (int u = 0; u < kBlock / 64; ++u) v += red[u][threadIdx.x];
      A.part[(P_RESP + threadIdx.x) * A.Tmax + tl] = v;
    }
    __syncthreads();  // red is reused by the next group
  }
}

// Failure record of local point p of RVE r (re-evaluated with the record).
template <class T>
__global__ void k_b_diagnose(Arena A, Settings S, int r, int slot, const T* __restrict__ state_c,
                             const double* __restrict__ wc, int64_t p, cia_point_failure* rec) {
  const Desc& D = A.desc[r];
  const Req& M = A.req[r];
  const int64_t e = p / 8;
  const int q = (int)(p % 8);
  const int64_t NPT = A.NPT;
  const int64_t gp = 8 * D.e0 + p;
  const double* w = A.w[slot];
  double we[8][3], dwe[8][3];
  for (int a = 0; a < 8; ++a) {
    const int64_t c = D.e0 + periodic_corner(e, a, D.nx, D.ny, D.nz);
    for (int m = 0; m < 3; ++m) {
      we[a][m] = w[3 * c + m];
      dwe[a][m] = w[3 * c + m] - wc[3 * c + m];
    }
  }
  T eps[6], rate[6];  // as the assembly and the final pass form them
  gp_strain_sf<T, double, 3>(we, 0, q, D.g, eps);
  gp_strain_sf<T, double, 3>(dwe, 0, q, D.g, rate);
  for (int m = 0; m < 6; ++m) {
    eps[m] += (T)M.e[m];
    rate[m] = (rate[m] + (T)(M.e[m] - M.ec[m])) / (T)M.dt;
  }
  const PhaseConst P = A.tabs[r].p[A.ephase[D.e0 + e]];
  T epsp[6] = {0, 0, 0, 0, 0, 0};
  T qb = 0, ga = 0;
  if (P.kind == CIA_MAT_J2_VISCOPLASTIC) {
    for (int m = 0; m < 6; ++m) epsp[m] = state_c[m * NPT + gp];
    qb = state_c[6 * NPT + gp];
    ga = state_c[7 * NPT + gp];
  }
  cia_point_failure out;
  memset(&out, 0, sizeof(out));
  PointOut<T> o;
  PointExtra<T> ex;
  out.reason = material_point<T, true>(P, S, M.dt, eps, rate, epsp, qb, ga, o, &ex, &out);
  out.point_id = p;
  *rec = out;
}

// Periodic node gather of the element force rows into r[slot] (class 0
// zeroed); partials ||r||^2 and ||w[slot]||^2.
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_gather(Arena A, int want) {
  __shared__ double lds[kBlock / 64];
  const int64_t ET = A.ET;
  const T* F = (const T*)A.F;
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (!(A.force || c.state == want)) continue;
    const int slot = want == S_LS ? 1 - c.cur : c.cur;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double rr = 0.0, ww = 0.0;
    if (cl < D.E) {
      const unsigned cu = (unsigned)cl, rw = cu / (unsigned)D.nx;
      const int i = (int)(cu - rw * (unsigned)D.nx);
      const int k = (int)(rw / (unsigned)D.ny);
      const int j = (int)(rw - (unsigned)k * (unsigned)D.ny);
      double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int64_t e = D.e0 + periodic_corner_element(i, j, k, a, D.nx, D.ny, D.nz);
#pragma unroll
        for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * ET + e];
      }
      if (cl == 0) s[0] = s[1] = s[2] = 0.0;
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        A.r[slot][3 * gc + m] = s[m];
        rr += s[m] * s[m];
        const double wv = A.w[slot][3 * gc + m];
        ww += wv * wv;
      }
    }
    const double a = block_sum<kBlock>(rr, lds);
    const double b = block_sum<kBlock>(ww, lds);
    if (threadIdx.x == 0) {
      A.part[P_RR * A.Tmax + t] = a;
      A.part[P_WW * A.Tmax + t] = b;
    }
  }
}

// ---------------------------------------------------------------------------
// Class-vector kernels.

// w[0] = the Newton start: the committed fluctuation, or the caller's initial
// guess w0 (every RVE).
__global__ void __launch_bounds__(kBlock) k_b_begin(Arena A, const double* __restrict__ wc) {
  CIA_TILES(A.T[F_CL]) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) A.w[0][3 * gc + m] = wc[3 * gc + m];
    }
  }
}

// Line-search candidate w[1 - cur] = w[cur] + alpha dw.
__global__ void __launch_bounds__(kBlock) k_b_axpy(Arena A) {
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_LS) continue;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
      const double al = c.alpha;
      const double* src = A.w[c.cur];
      double* dst = A.w[1 - c.cur];
#pragma unroll
      for (int m = 0; m < 3; ++m) dst[3 * gc + m] = src[3 * gc + m] + al * A.dw[3 * gc + m];
    }
  }
}

// ||dw||^2 partials of the RVEs that just solved their Newton system.
__global__ void __launch_bounds__(kBlock) k_b_norm_dw(Arena A) {
  __shared__ double lds[kBlock / 64];
  CIA_TILES(A.T[F_CL]) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    if (A.ctl[r].state != S_PCG) continue;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double s = 0.0;
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double v = A.dw[3 * gc + m];
        s += v * v;
      }
    }
    const double a = block_sum<kBlock>(s, lds);
    if (threadIdx.x == 0) A.part[P_NORM * A.Tmax + t] = a;
  }
}

// next_w = w[cur] of the converged RVEs.
__global__ void __launch_bounds__(kBlock) k_b_copy_out(Arena A, double* __restrict__ next_w) {
  CIA_TILES(A.T[F_CL]) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_DONE) continue;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) next_w[3 * gc + m] = A.w[c.cur][3 * gc + m];
    }
  }
}

// ---------------------------------------------------------------------------
// Spectral PCG, batched (see the single-RVE solver for the scheme): the scalar
// steps between the vector kernels run in one wavefront per RVE.

// Complex type of the preconditioner spectra: FP32 batches keep them in FP32
// (the preconditioner only has to be a fixed SPD operator; the PCG vectors, dot
// products and Newton residuals stay FP64).
template <class T>
struct SpecOf {
  using type = double2;
};
template <>
struct SpecOf<float> {
  using type = float2;
};

template <class SC>
__device__ __forceinline__ SC mk2(double a, double b) {
  SC v;
  v.x = a;
  v.y = b;
  return v;
}

// x = 0, r = -r_newton, p_old = 0, packed r; partials (r.r, sum r).
template <class SC>
__global__ void __launch_bounds__(kBlock) k_b_sp_init(Arena A) {
  __shared__ double red[kBlock / 64][4];
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (!pcg_active(A, c)) continue;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
      const double* rn = A.r[c.cur];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double rv = -rn[3 * gc + m];
        A.dw[3 * gc + m] = 0.0;
        A.pr[3 * gc + m] = rv;
        A.pp[1][3 * gc + m] = 0.0;
        v[0] += rv * rv;
        v[1 + m] = rv;
      }
      SC* uab = (SC*)A.uab;
      uab[2 * D.e0 + cl] = mk2<SC>(v[1], v[2]);
      uab[2 * D.e0 + D.E + cl] = mk2<SC>(v[3], 0.0);
    }
    // the four block sums in one LDS round
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double a = wave_sum(v[s]);
      if (lane == 0) red[wid][s] = a;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double a = 0.0;
#pragma unroll
      for (int u = 0; u < kBlock / 64; ++u) a += red[u][threadIdx.x];
      A.part[(P_UPD + threadIdx.x) * A.Tmax + t] = a;
    }
    __syncthreads();
  }
}

// Sum of one partial row over the tiles [t0, t0 + nt) of an RVE, by one
// wavefront in a fixed order (every lane returns the total).
__device__ __forceinline__ double rve_sum(const Arena& A, int row, int t0, int nt) {
  const int lane = threadIdx.x & 63;
  double a = 0.0;
  for (int u = lane; u < nt; u += 64) a += A.part[row * A.Tmax + t0 + u];
  return wave_sum(a);
}

#define CIA_RVE_WAVE()                                   \
  const int r = (int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)); \
  if (r >= A.R) return;                                  \
  const int lane = threadIdx.x & 63;                     \
  Ctl& c = A.ctl[r];                                     \
  const Desc& D = A.desc[r];                             \
  (void)lane;                                            \
  (void)D

// After an update (FIRST: after the initialisation): r.r, sum r, convergence.
template <bool FIRST>
__global__ void __launch_bounds__(kBlock) k_b_scal_upd(Arena A) {
  CIA_RVE_WAVE();
  if (!pcg_active(A, c)) return;
  double u[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) u[s] = rve_sum(A, P_UPD + s, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0 || A.force) return;
  if (FIRST) {
    const double tol = fmax(c.lin_rtol * sqrt(u[0]), A.lin_atol);
    c.p_tol2 = tol * tol;
  } else {
    c.p_iters += 1;
  }
  c.p_rr = u[0];
  c.p_sumr[0] = u[1];
  c.p_sumr[1] = u[2];
  c.p_sumr[2] = u[3];
  if (u[0] <= c.p_tol2 || !isfinite(u[0])) c.p_done = isfinite(u[0]) ? 1 : 2;
}

// r.z: beta of the next element pass.
__global__ void __launch_bounds__(kBlock) k_b_scal_rz(Arena A) {
  CIA_RVE_WAVE();
  if (!pcg_active(A, c)) return;
  const double rz = rve_sum(A, P_RZ, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0 || A.force) return;
  if (!(rz > 0.0) || !isfinite(rz)) {
    c.p_done = 2;
    return;
  }
  c.p_rz_old = c.p_rz;
  c.p_rz = rz;
  c.p_beta = c.p_iters == 0 ? 0.0 : rz / c.p_rz_old;
}

// p.y: alpha of the update.
__global__ void __launch_bounds__(kBlock) k_b_scal_pq(Arena A) {
  CIA_RVE_WAVE();
  if (!pcg_active(A, c)) return;
  const double pq = rve_sum(A, P_PQ, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0 || A.force) return;
  if (!(pq > 0.0) || !isfinite(pq)) {
    c.p_done = 2;
    return;
  }
  c.p_alpha = c.p_rz / pq;
}

// Symbol solve of the packed spectra per mode pair (k, -k); r.z partials.
template <class SC>
__global__ void __launch_bounds__(kBlock) k_b_apply(Arena A) {
  __shared__ double lds[kBlock / 64];
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (!pcg_active(A, c)) continue;
    const int64_t k = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    SC* ua = (SC*)A.uab + 2 * D.e0;
    SC* ub = ua + D.E;
    const double s0 = c.p_sumr[0], s1 = c.p_sumr[1], s2 = c.p_sumr[2];
    double rz = 0.0;
    if (k < D.E) {
      const int64_t mk = neg_mode(k, D.nx, D.ny, D.nz);
      if (k <= mk) {
        const SC a = ua[k], am = ua[mk], b = ub[k];
        double ar[3] = {0.5 * (a.x + am.x) - s0, 0.5 * (a.y + am.y) - s1, b.x - s2};
        double ai[3] = {0.5 * (a.y - am.y), -0.5 * (a.x - am.x), b.y};
        double zr[3], zi[3];
        spec_solve(D.kinv + 6 * k, false, ar, ai, zr, zi);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          zr[m] *= D.inv_n;
          zi[m] *= D.inv_n;
        }
        const double wk = (k == mk) ? 1.0 : 2.0;
        rz += wk * (ar[0] * zr[0] + ai[0] * zi[0] + ar[1] * zr[1] + ai[1] * zi[1] +
                    ar[2] * zr[2] + ai[2] * zi[2]);
        ua[k] = mk2<SC>(zr[0] - zi[1], zi[0] + zr[1]);
        ub[k] = mk2<SC>(zr[2], zi[2]);
        if (mk != k) {
          ua[mk] = mk2<SC>(zr[0] + zi[1], -zi[0] + zr[1]);
          ub[mk] = mk2<SC>(zr[2], -zi[2]);
        }
      }
    }
    const double a = block_sum<kBlock>(rz, lds);
    if (threadIdx.x == 0) A.part[P_RZ * A.Tmax + t] = a;
  }
}

// Element pass: p = (z - z0) + beta p_old at the element corners, forces of
// K p.  One lane per element: the lane gathers its 8 corners, forms the strain
// of all 8 Gauss points by sum factorisation (box_gradients), applies the
// compact tangents and spreads the traction sums back to the corners
// (box_forces).  No corner exchange and no cross-lane force reduction: about a
// quarter of the vector instructions of one lane per Gauss point with an
// 8-lane exchange and reduction, whose issue rate (not HBM) bounded the pass.
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_elem(Arena A, int pi) {
  using SC = typename SpecOf<T>::type;
  const int64_t NPT = A.NPT, ET = A.ET;
  const double* p_old = A.pp[1 - pi];
  double* p_new = A.pp[pi];
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (!pcg_active(A, c)) continue;
    const int64_t e = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    if (e >= D.E) continue;
    const double beta = c.p_beta;
    const T* tan = (const T*)A.tan[c.cur];
    const SC* ua = (const SC*)A.uab + 2 * D.e0;
    const SC* ub = ua + D.E;
    const double z0[3] = {ua[0].x, ua[0].y, ub[0].x};
    const int nx = D.nx, ny = D.ny, nz = D.nz;
    const unsigned eu = (unsigned)e, rw = eu / (unsigned)nx;
    const int i = (int)(eu - rw * (unsigned)nx);
    const int k = (int)(rw / (unsigned)ny);
    const int j = (int)(rw - (unsigned)k * (unsigned)ny);
    const int ip = i + 1 == nx ? 0 : i + 1;
    const int jp = j + 1 == ny ? 0 : j + 1;
    const int kp = k + 1 == nz ? 0 : k + 1;
    // corner values of p = (z - z0) + beta p_old, rounded to T (the strains are formed in T)
    T u[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int ii = (a & 1) ? ip : i, jj = (a & 2) ? jp : j, kk = (a & 4) ? kp : k;
      const int64_t cq = ii + (int64_t)nx * (jj + (int64_t)ny * kk);
      const SC t01 = ua[cq], t2 = ub[cq];
      const int64_t gq = D.e0 + cq;
      const double o0 = (t01.x - z0[0]) + beta * p_old[3 * gq + 0];
      const double o1 = (t01.y - z0[1]) + beta * p_old[3 * gq + 1];
      const double o2 = (t2.x - z0[2]) + beta * p_old[3 * gq + 2];
      if (a == 0) {
        p_new[3 * gq + 0] = o0;
        p_new[3 * gq + 1] = o1;
        p_new[3 * gq + 2] = o2;
      }
      u[a][0] = (T)o0;
      u[a][1] = (T)o1;
      u[a][2] = (T)o2;
    }
    T gX[2][2][3], gY[2][2][3], gZ[2][2][3];
    box_gradients<T>(u, D.g, gX, gY, gZ);
    Tangent<T> tc;
    tc.K = (T)A.tabs[r].p[A.ephase[D.e0 + e]].K;
    const int64_t g8 = 8 * (D.e0 + e);
    T dd[8], rr[8];
    load8(tan + g8, dd);
    load8(tan + NPT + g8, rr);
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
        for (int m = 0; m < 6; ++m) {
          const T* src = tan + (2 + m) * NPT + g8 + 4 * qx;
#pragma unroll
          for (int h = 0; h < 4; ++h) nn[m][h] = src[h];
        }
      }
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int qy = h >> 1, qz = h & 1, q = 4 * qx + h;
        T eps[6];
        box_strain(gX, gY, gZ, qx, qy, qz, eps);
        tc.d = dd[q];
        tc.r = rr[q];
#pragma unroll
        for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != (T)0 ? nn[m][h] : (T)0;
        T sg[6];
        tangent_apply(tc, eps, sg);
        box_traction_add(sg, qx, qy, qz, SX, SY, SZ);
      }
    }
    T f[8][3];
    box_forces<T>(SX, SY, SZ, D.g, f);
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) ((T*)A.F)[(3 * a + m) * ET + D.e0 + e] = f[a][m];
  }
}

// y = gather(F) (class 0 zeroed); p.y partials.
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_node(Arena A, int pi) {
  __shared__ double lds[kBlock / 64];
  const int64_t ET = A.ET;
  const T* F = (const T*)A.F;
  const double* p = A.pp[pi];
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    if (!pcg_active(A, A.ctl[r])) continue;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double pq = 0.0;
    if (cl < D.E) {
      const unsigned cu = (unsigned)cl, rw = cu / (unsigned)D.nx;
      const int i = (int)(cu - rw * (unsigned)D.nx);
      const int k = (int)(rw / (unsigned)D.ny);
      const int j = (int)(rw - (unsigned)k * (unsigned)D.ny);
      double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int64_t e = D.e0 + periodic_corner_element(i, j, k, a, D.nx, D.ny, D.nz);
#pragma unroll
        for (int m = 0; m < 3; ++m) s[m] += (double)F[(3 * a + m) * ET + e];
      }
      if (cl == 0) s[0] = s[1] = s[2] = 0.0;
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        A.py[3 * gc + m] = s[m];
        pq += p[3 * gc + m] * s[m];
      }
    }
    const double a = block_sum<kBlock>(pq, lds);
    if (threadIdx.x == 0) A.part[P_PQ * A.Tmax + t] = a;
  }
}

// x += alpha p, r -= alpha y, packed r; partials (r.r, sum r).
template <class SC>
__global__ void __launch_bounds__(kBlock) k_b_update(Arena A, int pi) {
  __shared__ double red[kBlock / 64][4];
  const double* p = A.pp[pi];
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (!pcg_active(A, c)) continue;
    const double alpha = c.p_alpha;
    const int64_t cl = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (cl < D.E) {
      const int64_t gc = D.e0 + cl;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        A.dw[3 * gc + m] += alpha * p[3 * gc + m];
        const double rv = A.pr[3 * gc + m] - alpha * A.py[3 * gc + m];
        A.pr[3 * gc + m] = rv;
        v[0] += rv * rv;
        v[1 + m] = rv;
      }
      SC* uab = (SC*)A.uab;
      uab[2 * D.e0 + cl] = mk2<SC>(v[1], v[2]);
      uab[2 * D.e0 + D.E + cl] = mk2<SC>(v[3], 0.0);
    }
    // the four block sums in one LDS round
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double a = wave_sum(v[s]);
      if (lane == 0) red[wid][s] = a;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      double a = 0.0;
#pragma unroll
      for (int u = 0; u < kBlock / 64; ++u) a += red[u][threadIdx.x];
      A.part[(P_UPD + threadIdx.x) * A.Tmax + t] = a;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Batched 3-D DFT passes of the packed spectra (one tile = a group of x-lines,
// or 16 consecutive x positions of y- or z-lines, of one RVE).

template <bool INV, class SC>
__global__ void __launch_bounds__(256) k_b_dft_x(Arena A) {
  extern __shared__ __align__(16) unsigned char dft_raw[];
  SC* dft_sm = reinterpret_cast<SC*>(dft_raw);
  __shared__ SC wt[kDftMaxTw];
  CIA_ACTIVE_TILES(F_DX) {
    const int r = A.t2r[F_DX][t];
    const Desc& D = A.desc[r];
    if (!pcg_active(A, A.ctl[r])) continue;
    const int nx = D.nx, lines = D.lines_x;
    const int64_t total_lines = 2 * (int64_t)D.ny * D.nz;
    const int64_t l0 = (int64_t)(t - D.t0[F_DX]) * lines;
    const int nl = (int)min<int64_t>(lines, total_lines - l0);
    const int ld = nx | 1;  // odd row pitch: lanes a row apart fall in distinct LDS banks
    SC* X = dft_sm;
    SC* Y = dft_sm + lines * ld;
    SC* g = (SC*)A.uab + 2 * D.e0 + l0 * nx;
    const float inv_nx = 1.0f / (float)nx;
    load_twiddles(D.ax, wt);
    for (int u = threadIdx.x; u < nl * nx; u += blockDim.x) {
      const int l = fdiv(u, inv_nx);
      X[l * ld + (u - l * nx)] = g[u];
    }
    __syncthreads();
    dft_stage<true, INV>(X, Y, nl, ld, 1, D.ax, wt);
    __syncthreads();
    dft_stage<false, INV>(Y, X, nl, ld, 1, D.ax, wt);
    __syncthreads();
    for (int u = threadIdx.x; u < nl * nx; u += blockDim.x) {
      const int l = fdiv(u, inv_nx);
      g[u] = X[l * ld + (u - l * nx)];
    }
    __syncthreads();
  }
}

// x positions per tile of the strided passes: up to 32, so an RVE edge of up to
// 32 is one tile per (line row, component) and every synchronisation of a tile
// covers twice the lines of a 16-wide tile
constexpr int kBatchDftTile = 32;

template <bool INV, bool ZPASS, class SC>
__global__ void __launch_bounds__(256) k_b_dft_strided(Arena A) {
  extern __shared__ __align__(16) unsigned char dft_raw[];
  SC* dft_sm = reinterpret_cast<SC*>(dft_raw);
  __shared__ SC wt[kDftMaxTw];
  constexpr int FAM = ZPASS ? F_DZ : F_DY;
  CIA_ACTIVE_TILES(FAM) {
    const int r = A.t2r[FAM][t];
    const Desc& D = A.desc[r];
    if (!pcg_active(A, A.ctl[r])) continue;
    const DftAxis& ax = ZPASS ? D.az : D.ay;
    const int n = ax.n, ld = n | 1;  // odd row pitch (distinct LDS banks a row apart)
    const int tiles_x = (D.nx + kBatchDftTile - 1) / kBatchDftTile;
    const int lt = t - D.t0[FAM];
    const int xt = lt % tiles_x, rest = lt / tiles_x;
    const int no = ZPASS ? D.ny : D.nz;
    const int o = rest % no, comp = rest / no;
    SC* X = dft_sm;
    SC* Y = dft_sm + kBatchDftTile * ld;
    // the tiles_x x-tiles of a row share the nx positions evenly (at most
    // kBatchDftTile each), so no tile is left near-empty
    const int tw = (D.nx + tiles_x - 1) / tiles_x;
    const int i0 = xt * tw;
    const int64_t plane = (int64_t)D.nx * D.ny;
    const int64_t stride = ZPASS ? plane : D.nx;
    SC* g = (SC*)A.uab + 2 * D.e0 + (int64_t)comp * D.E + i0 + (ZPASS ? (int64_t)D.nx * o : plane * o);
    const int width = min(tw, D.nx - i0);
    const float inv_w = 1.0f / (float)width;
    load_twiddles(ax, wt);
    for (int u = threadIdx.x; u < width * n; u += blockDim.x) {
      const int e = fdiv(u, inv_w), ii = u - e * width;
      X[ii * ld + e] = g[ii + stride * e];
    }
    __syncthreads();
    dft_stage<true, INV>(X, Y, width, ld, 1, ax, wt);
    __syncthreads();
    dft_stage<false, INV>(Y, X, width, ld, 1, ax, wt);
    __syncthreads();
    for (int u = threadIdx.x; u < width * n; u += blockDim.x) {
      const int e = fdiv(u, inv_w), ii = u - e * width;
      g[ii + stride * e] = X[ii * ld + e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Effective tangent of the converged RVEs of the last solve (the Schur complement
// of _compute_effective_tangent, src/fem/rve/solver.py:438-461, as the single-RVE
// solver forms it): for every macro strain direction e_j the affine force rows of
// e_j through the converged compact tangents, gathered into the Newton residual
// slot; the batched PCG solves K dw = -T^T K A e_j; then column j of
// C_eff = (1/V) sum_q w C_q (e_j + B_q dw).  RVEs in this phase are in S_PCG with
// Ctl::pad0 = 1 (tangent mode).

__global__ void __launch_bounds__(kBlock) k_b_tan_begin(Arena A, double lin_rtol) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= A.R) return;
  Ctl& c = A.ctl[r];
  if (c.state != S_DONE) return;
  c.state = S_PCG;
  c.pad0 = 1;
  c.p_done = 0;
  c.p_iters = 0;
  c.lin_rtol = lin_rtol;
}

__global__ void __launch_bounds__(kBlock) k_b_tan_end(Arena A) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= A.R) return;
  Ctl& c = A.ctl[r];
  if (c.pad0 != 1) return;
  if (c.state == S_PCG) c.state = S_DONE;
  c.pad0 = 0;
}

// Affine force rows of e_j (one lane per element), tangent-mode RVEs.
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_affine(Arena A, int jdir) {
  const int64_t NPT = A.NPT, ET = A.ET;
  CIA_ACTIVE_TILES(F_CL) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_PCG || c.pad0 != 1) continue;
    const int64_t e = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    if (e >= D.E) continue;
    const int64_t ge = D.e0 + e;
    const T* tan = (const T*)A.tan[c.cur];
    Tangent<T> tc;
    tc.K = (T)A.tabs[r].p[A.ephase[ge]].K;
    T f[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a][0] = f[a][1] = f[a][2] = (T)0;
    T ej[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) ej[m] = m == jdir ? (T)1 : (T)0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t p = 8 * ge + q;
      tc.d = tan[p];
      tc.r = tan[NPT + p];
#pragma unroll
      for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != (T)0 ? tan[(2 + m) * NPT + p] : (T)0;
      T sg[6];
      tangent_apply(tc, ej, sg);
      switch (q) {
        case 0: force_at<0>(sg, D.g, f); break;
        case 1: force_at<1>(sg, D.g, f); break;
        case 2: force_at<2>(sg, D.g, f); break;
        case 3: force_at<3>(sg, D.g, f); break;
        case 4: force_at<4>(sg, D.g, f); break;
        case 5: force_at<5>(sg, D.g, f); break;
        case 6: force_at<6>(sg, D.g, f); break;
        default: force_at<7>(sg, D.g, f); break;
      }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) ((T*)A.F)[(3 * a + m) * ET + ge] = f[a][m];
  }
}

// Tile partials of sum_q C_q (e_j + B_q dw) (rows P_RESP .. P_RESP + 5; the response
// rows of the final pass are read before this phase), every tile of the
// tangent-mode RVEs.
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_ceff(Arena A, int jdir) {
  __shared__ double red[kBlock / 64][6];
  const int64_t NPT = A.NPT;
  CIA_TILES(A.T[F_CL]) {
    const int r = A.t2r[F_CL][t];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_PCG || c.pad0 != 1) continue;
    const int64_t e = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (e < D.E) {
      const int64_t ge = D.e0 + e;
      int64_t cq[8];
      int i, j, k;
      periodic_corners(e, D.nx, D.ny, D.nz, cq, i, j, k);
      double se[8][3];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m) se[a][m] = A.dw[3 * (D.e0 + cq[a]) + m];
      const T* tan = (const T*)A.tan[c.cur];
      Tangent<double> tc;
      tc.K = A.tabs[r].p[A.ephase[ge]].K;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t p = 8 * ge + q;
        tc.d = (double)tan[p];
        tc.r = (double)tan[NPT + p];
#pragma unroll
        for (int m = 0; m < 6; ++m) tc.n[m] = tc.r != 0.0 ? (double)tan[(2 + m) * NPT + p] : 0.0;
        double eps[6], sg[6];
        switch (q) {
          case 0: strain_at<0>(se, D.g, eps); break;
          case 1: strain_at<1>(se, D.g, eps); break;
          case 2: strain_at<2>(se, D.g, eps); break;
          case 3: strain_at<3>(se, D.g, eps); break;
          case 4: strain_at<4>(se, D.g, eps); break;
          case 5: strain_at<5>(se, D.g, eps); break;
          case 6: strain_at<6>(se, D.g, eps); break;
          default: strain_at<7>(se, D.g, eps); break;
        }
        eps[jdir] += 1.0;
        tangent_apply(tc, eps, sg);
#pragma unroll
        for (int m = 0; m < 6; ++m) acc[m] += sg[m];
      }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double a = wave_sum(acc[m]);
      if (lane == 0) red[wid][m] = a;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      double a = 0.0;
#pragma unroll
      for (int u = 0; u < kBlock / 64; ++u) a += red[u][threadIdx.x];
      A.part[(P_RESP + threadIdx.x) * A.Tmax + t] = a;
    }
    __syncthreads();
  }
}

// Column j of every tangent-mode RVE's C_eff (out [R][36], row-major), its PCG
// iterations counted, the PCG reset for the next direction.
__global__ void __launch_bounds__(kBlock) k_b_ceff_sum(Arena A, int jdir, double* __restrict__ out) {
  CIA_RVE_WAVE();
  if (c.state != S_PCG || c.pad0 != 1) return;
  double col[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) col[m] = rve_sum(A, P_RESP + m, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0) return;
  bool fin = true;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    out[36 * r + 6 * m + jdir] = col[m] * D.g.jw / D.volume;
    fin = fin && isfinite(col[m]);
  }
  c.pcg_its += c.p_iters;
  c.p_done = 0;
  c.p_iters = 0;
  if (!fin) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_NON_FINITE_EFFECTIVE_TANGENT_SENSITIVITY;
  }
}

// ---------------------------------------------------------------------------
// Newton control (one wavefront per RVE).

__device__ void newton_test(Ctl& c, const NewtonP& N) {
  const cia_newton_settings& ns = N.ns;
  c.rtol = fmax(ns.residual_absolute_tolerance, ns.residual_relative_tolerance * c.r0);
  c.itol = fmax(ns.increment_absolute_tolerance,
                ns.increment_relative_tolerance * fmax(sqrt(c.wn2), 1.0));
  if (c.rn <= c.rtol && (c.it == 0 || c.last_inc <= c.itol)) {
    c.state = S_DONE;
    return;
  }
  if (c.it == ns.max_iterations) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_MAXIMUM_ITERATIONS_EXCEEDED;
    c.fail_it = c.it;
    c.fail_rn = c.rn;
    c.fail_rtol = c.rtol;
    c.fail_inc = c.last_inc;
    c.fail_itol = c.itol;
    return;
  }
  double lin = N.lin_rtol;
  if (N.adaptive) {
    // inexact Newton: solve only as accurately as the current residual warrants
    const double eta = fmin(N.forcing_max > 0 ? N.forcing_max : 0.1, c.rn / c.r0);
    lin = fmax(lin, fmax(eta, 0.5 * c.rtol / fmax(c.rn, 1e-300)));
    lin = fmin(lin, 0.5);
  }
  c.lin_rtol = lin;
  c.state = S_PCG;
  c.p_done = 0;
  c.p_iters = 0;
}

__global__ void __launch_bounds__(kBlock) k_b_ctl_init(Arena A) {
  const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= A.R) return;
  Ctl c;
  memset(&c, 0, sizeof(c));
  c.state = S_NEWTON;
  c.p_done = 1;
  c.first_fail = ~0ull;
  c.r0 = -1.0;
  c.last_inc = INFINITY;
  c.alpha = 1.0;
  c.fail_slot = -1;
  A.ctl[r] = c;
}

// First assembly: residual and fluctuation norms, then the Newton test.
__global__ void __launch_bounds__(kBlock) k_b_ctl_first(Arena A, NewtonP N) {
  CIA_RVE_WAVE();
  if (c.state != S_NEWTON) return;
  const double rn2 = rve_sum(A, P_RR, D.t0[F_CL], D.nt[F_CL]);
  const double wn2 = rve_sum(A, P_WW, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0) return;
  c.gp_updates = 8 * D.E;
  if (c.first_fail != ~0ull) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_MATERIAL_POINT_FAILURE;
    c.fail_slot = c.cur;
    return;
  }
  c.rn = sqrt(rn2);
  c.wn2 = wn2;
  c.r0 = c.rn;
  newton_test(c, N);
}

// After the Newton system solve: increment norm, convergence, line search.
__global__ void __launch_bounds__(kBlock) k_b_ctl_pcg(Arena A, NewtonP N) {
  CIA_RVE_WAVE();
  if (c.state != S_PCG) return;
  const double dn2 = rve_sum(A, P_NORM, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0) return;
  c.pcg_its += c.p_iters;
  const double inc = sqrt(dn2);
  c.inc_n = inc;
  if (!isfinite(inc)) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_NON_FINITE_INCREMENT;
    c.fail_it = c.it;
    c.fail_rn = c.rn;
    c.fail_rtol = c.rtol;
    c.fail_inc = inc;
    c.fail_itol = c.itol;
    return;
  }
  if (c.rn <= c.rtol && inc <= c.itol) {
    c.state = S_DONE;
    return;
  }
  c.state = S_LS;
  c.bt = 0;
  c.alpha = 1.0;
}

// Line-search candidate evaluated: Armijo test, accept (then the next Newton
// test) or backtrack.
__global__ void __launch_bounds__(kBlock) k_b_ctl_ls(Arena A, NewtonP N) {
  CIA_RVE_WAVE();
  if (c.state != S_LS) return;
  const double cn2 = rve_sum(A, P_RR, D.t0[F_CL], D.nt[F_CL]);
  const double wn2 = rve_sum(A, P_WW, D.t0[F_CL], D.nt[F_CL]);
  if (lane != 0) return;
  const cia_newton_settings& ns = N.ns;
  c.gp_updates += 8 * D.E;
  if (c.first_fail != ~0ull) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_MATERIAL_POINT_FAILURE;
    c.fail_slot = 1 - c.cur;
    return;
  }
  const double cn = sqrt(cn2);
  const double rn = c.rn;
  const bool accept = !ns.armijo_enabled ||
                      0.5 * cn * cn <= 0.5 * rn * rn - ns.armijo_sufficient_decrease * c.alpha * rn * rn;
  if (accept) {
    c.cur = 1 - c.cur;
    c.rn = cn;
    c.wn2 = wn2;
    c.last_inc = c.alpha * c.inc_n;
    c.backtracks += c.bt;
    c.it += 1;
    newton_test(c, N);
    return;
  }
  c.bt += 1;
  if (c.bt > ns.armijo_max_backtracks) {
    c.state = S_FAILED;
    c.reason = CIA_RVE_ARMIJO_EXHAUSTED;
    c.fail_it = c.it;
    c.fail_rn = c.rn;
    c.fail_rtol = c.rtol;
    c.fail_inc = c.inc_n;
    c.fail_itol = c.itol;
    return;
  }
  c.alpha = pow(ns.armijo_reduction_factor, (double)c.bt);
}

// Per-phase sums of the converged RVEs (element per lane; an element has one
// phase), over groups of kFinalGroup class tiles of one RVE (leader tiles in
// A.grp_cl): each lane accumulates its elements over the group, and the group's
// B_NPHASE_ROWS sums go to the leader tile's partials (k_b_resp reads those only).
template <class T>
__global__ void __launch_bounds__(kBlock) k_b_phase_sums(Arena A) {
  __shared__ double red[kBlock / 64][B_NPHASE_ROWS];
  const int64_t NPT = A.NPT;
  for (int base_ = 0; base_ < A.n_grp_cl; base_ += (int)gridDim.x) {
    const int gi = base_ + (int)xcd_block();
    if (gi >= A.n_grp_cl) break;
    const int tl = A.grp_cl[gi];
    const int r = A.t2r[F_CL][tl];
    const Desc& D = A.desc[r];
    const Ctl& c = A.ctl[r];
    if (c.state != S_DONE) continue;
    const int tend = min(tl + kFinalGroup, D.t0[F_CL] + D.nt[F_CL]);
    const T* sig = (const T*)A.sig[c.cur];
    const T* pd = (const T*)A.pd;
    double acc[B_NPHASE_ROWS];
#pragma unroll
    for (int s = 0; s < B_NPHASE_ROWS; ++s) acc[s] = 0.0;
    for (int t = tl; t < tend; ++t) {
      const int64_t e = (int64_t)(t - D.t0[F_CL]) * kBlock + threadIdx.x;
      if (e < D.E) {
        double v[9] = {8.0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 8; ++q) {
          const int64_t gp = 8 * (D.e0 + e) + q;
#pragma unroll
          for (int m = 0; m < 6; ++m) v[1 + m] += (double)sig[m * NPT + gp];
          v[7] += (double)pd[gp];
          v[8] += (double)pd[NPT + gp];
        }
        const int ph = A.ephase[D.e0 + e];
#pragma unroll
        for (int P = 0; P < B_PHASES; ++P) {
          if (ph == P) {
#pragma unroll
            for (int m = 0; m < 9; ++m) acc[9 * P + m] += v[m];
          }
        }
      }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < B_NPHASE_ROWS; ++s) {
      const double w = wave_sum(acc[s]);
      if (lane == 0) red[wid][s] = w;
    }
    __syncthreads();
    if (threadIdx.x < B_NPHASE_ROWS) {
      double w = 0.0;
#pragma unroll
      for (int u = 0; u < kBlock / 64; ++u) w += red[u][threadIdx.x];
      A.part[(P_PHASE + threadIdx.x) * A.Tmax + tl] = w;
    }
    __syncthreads();
  }
}

// Response of the converged RVEs: out[r][B_NRESP + B_NPHASE_ROWS] (raw sums).
__global__ void __launch_bounds__(kBlock) k_b_resp(Arena A, double* __restrict__ out) {
  CIA_RVE_WAVE();
  if (c.state != S_DONE) return;
  constexpr int NR = B_NRESP + B_NPHASE_ROWS;
  for (int s = 0; s < B_NRESP; ++s) {
    // the final pass writes one partial per group of kFinalGroup tiles (at the
    // group's leader tile): sum those only
    double a = 0.0;
    for (int u = lane * kFinalGroup; u < D.nt[F_GP]; u += 64 * kFinalGroup)
      a += A.part[(P_RESP + s) * A.Tmax + D.t0[F_GP] + u];
    const double v = wave_sum(a);
    if (lane == 0) out[r * NR + s] = v;
  }
  for (int s = 0; s < B_NPHASE_ROWS; ++s) {
    double a = 0.0;  // one partial per group of kFinalGroup class tiles
    for (int u = lane * kFinalGroup; u < D.nt[F_CL]; u += 64 * kFinalGroup)
      a += A.part[(P_PHASE + s) * A.Tmax + D.t0[F_CL] + u];
    const double v = wave_sum(a);
    if (lane == 0) out[r * NR + B_NRESP + s] = v;
  }
}

// Counts of the RVE states into cnt[0..4], of RVEs with a running PCG into
// cnt[5], and the largest PCG iteration count of the RVEs in PCG into cnt[6].
__global__ void __launch_bounds__(kBlock) k_b_count(Arena A, int* __restrict__ cnt) {
  __shared__ int c[7];
  if (threadIdx.x < 7) c[threadIdx.x] = 0;
  __syncthreads();
  for (int r = threadIdx.x; r < A.R; r += blockDim.x) {
    const Ctl& x = A.ctl[r];
    atomicAdd(&c[x.state], 1);
    if (x.state == S_PCG && x.p_done == 0) atomicAdd(&c[5], 1);
    if (x.state == S_PCG) atomicMax(&c[6], x.p_iters);
  }
  __syncthreads();
  if (threadIdx.x < 7) cnt[threadIdx.x] = c[threadIdx.x];
}

}  // namespace bt
}  // namespace cia

// ===========================================================================
// Host object
// ===========================================================================

using namespace cia;
using namespace cia::bt;

// Kernels whose grid the autotuner chooses.
enum { K_ASM = 0, K_GATHER, K_ELEM, K_NODE, K_UPD, K_APPLY, K_DX, K_DY, K_DZ, K_N };

struct cia_rve_batch {
  int R = 0;
  int precision = CIA_PREC_FP64;
  size_t tsize = 8;
  cia_newton_settings newton{};
  cia_linear_settings lin{};
  double hm_scale = 1.0;
  int64_t NPT = 0, ET = 0;
  bool all_m1 = true;  // every J2 phase has rate exponent 1 (closed-form multiplier)
  std::vector<Desc> h_desc;
  std::vector<int64_t> e0;
  int T[NF] = {0};
  int Tmax = 0;
  size_t lds_x = 0, lds_y = 0, lds_z = 0;
  Arena A{};
  // device allocations owned by the batch
  std::vector<void*> allocs;
  Ctl* ctl = nullptr;
  Req* req = nullptr;
  double* resp = nullptr;
  double* tan_out = nullptr;  // [R][36] effective tangents (allocated on first use)
  bool solved = false;        // a solve has run: its converged states can give tangents
  int* cnt = nullptr;
  cia_point_failure* rec = nullptr;
  // pinned host staging
  Req* h_req = nullptr;
  int* h_cnt = nullptr;
  // autotuned grids
  int grid[K_N] = {0};
  float tuned_us[K_N] = {0};
  bool tuned = false;
  bool autotune = true;
  int cus = 256;
  // graph of two PCG iterations
  hipStream_t cap = nullptr;
  hipGraphExec_t pair = nullptr;
  static constexpr int kHints = 32;
  int its_hint[kHints] = {0};
  // counters since creation: solves, Newton rounds, PCG iterations launched (whole
  // batch), sum over RVEs of their PCG iterations, host synchronisations
  int64_t stats[5] = {0, 0, 0, 0, 0};
};

namespace {

template <class SC>
hipError_t set_dft_lds(int bytes) {
  const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
  hipError_t e = hipFuncSetAttribute((const void*)k_b_dft_strided<false, false, SC>, a, bytes);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)k_b_dft_strided<false, true, SC>, a, bytes);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)k_b_dft_strided<true, false, SC>, a, bytes);
  if (e == hipSuccess) e = hipFuncSetAttribute((const void*)k_b_dft_strided<true, true, SC>, a, bytes);
  return e;
}

cia_status balloc(cia_rve_batch* B, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc", e, __FILE__, __LINE__);
    return CIA_ERR_OUT_OF_MEMORY;
  }
  B->allocs.push_back(*p);
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) {
    set_error("hipMemset", e, __FILE__, __LINE__);
    return CIA_ERR_HIP;
  }
  return CIA_OK;
}

unsigned kgrid(const cia_rve_batch* B, int k, int fam) {
  const int g = B->grid[k];
  const int T = B->T[fam];
  return (unsigned)std::max(1, (g > 0 && g < T) ? g : T);
}

template <class T>
struct BOps {
  using SC = typename SpecOf<T>::type;
  static void assemble(cia_rve_batch* B, hipStream_t st, const Settings& S, int want,
                       const void* state_c, const double* wc) {
    if (B->all_m1)
      hipLaunchKernelGGL((k_b_assemble<T, true>), dim3(kgrid(B, K_ASM, F_GP)), dim3(kBlock), 0, st,
                         B->A, S, want, (const T*)state_c, wc);
    else
      hipLaunchKernelGGL((k_b_assemble<T, false>), dim3(kgrid(B, K_ASM, F_GP)), dim3(kBlock), 0, st,
                         B->A, S, want, (const T*)state_c, wc);
  }
  static void gather(cia_rve_batch* B, hipStream_t st, int want) {
    hipLaunchKernelGGL(k_b_gather<T>, dim3(kgrid(B, K_GATHER, F_CL)), dim3(kBlock), 0, st, B->A,
                       want);
  }
  static void dft(cia_rve_batch* B, hipStream_t st, bool inverse) {
    const unsigned gx = kgrid(B, K_DX, F_DX), gy = kgrid(B, K_DY, F_DY), gz = kgrid(B, K_DZ, F_DZ);
    if (!inverse) {
      hipLaunchKernelGGL((k_b_dft_x<false, SC>), dim3(gx), dim3(256), B->lds_x, st, B->A);
      hipLaunchKernelGGL((k_b_dft_strided<false, false, SC>), dim3(gy), dim3(256), B->lds_y, st, B->A);
      hipLaunchKernelGGL((k_b_dft_strided<false, true, SC>), dim3(gz), dim3(256), B->lds_z, st, B->A);
    } else {
      hipLaunchKernelGGL((k_b_dft_strided<true, true, SC>), dim3(gz), dim3(256), B->lds_z, st, B->A);
      hipLaunchKernelGGL((k_b_dft_strided<true, false, SC>), dim3(gy), dim3(256), B->lds_y, st, B->A);
      hipLaunchKernelGGL((k_b_dft_x<true, SC>), dim3(gx), dim3(256), B->lds_x, st, B->A);
    }
  }
  static unsigned waves(cia_rve_batch* B) { return (unsigned)((B->R + 3) / 4); }
  // Rebuild the active-tile lists: want < 0 for the PCG kernels, else the
  // Newton state of the next assembly / gather / line-search launches.
  static void compact(cia_rve_batch* B, hipStream_t st, int want = -1) {
    if (B->R <= 1024) {
      hipLaunchKernelGGL(k_b_compact_small, dim3((unsigned)B->R), dim3(1024), 0, st, B->A, want);
      return;
    }
    hipLaunchKernelGGL(k_b_compact_scan, dim3(1), dim3(1024), 0, st, B->A, want);
    hipLaunchKernelGGL(k_b_compact_fill, dim3((unsigned)B->R), dim3(kBlock), 0, st, B->A, want);
  }
  // One PCG iteration (pi selects p_new = pp[pi], p_old = pp[1 - pi]).
  static void elem(cia_rve_batch* B, hipStream_t st, int pi) {
    hipLaunchKernelGGL(k_b_elem<T>, dim3(kgrid(B, K_ELEM, F_CL)), dim3(kBlock), 0, st, B->A, pi);
  }
  static void iteration(cia_rve_batch* B, hipStream_t st, int pi) {
    const unsigned gw = waves(B);
    elem(B, st, pi);
    hipLaunchKernelGGL(k_b_node<T>, dim3(kgrid(B, K_NODE, F_CL)), dim3(kBlock), 0, st, B->A, pi);
    hipLaunchKernelGGL(k_b_scal_pq, dim3(gw), dim3(kBlock), 0, st, B->A);
    hipLaunchKernelGGL(k_b_update<SC>, dim3(kgrid(B, K_UPD, F_CL)), dim3(kBlock), 0, st, B->A, pi);
    hipLaunchKernelGGL(k_b_scal_upd<false>, dim3(gw), dim3(kBlock), 0, st, B->A);
    compact(B, st);
    dft(B, st, false);
    hipLaunchKernelGGL(k_b_apply<SC>, dim3(kgrid(B, K_APPLY, F_CL)), dim3(kBlock), 0, st, B->A);
    hipLaunchKernelGGL(k_b_scal_rz, dim3(gw), dim3(kBlock), 0, st, B->A);
    dft(B, st, true);
  }
  static void pcg_start(cia_rve_batch* B, hipStream_t st) {
    const unsigned gw = waves(B);
    compact(B, st);
    hipLaunchKernelGGL(k_b_sp_init<SC>, dim3(kgrid(B, K_UPD, F_CL)), dim3(kBlock), 0, st, B->A);
    hipLaunchKernelGGL(k_b_scal_upd<true>, dim3(gw), dim3(kBlock), 0, st, B->A);
    compact(B, st);
    dft(B, st, false);
    hipLaunchKernelGGL(k_b_apply<SC>, dim3(kgrid(B, K_APPLY, F_CL)), dim3(kBlock), 0, st, B->A);
    hipLaunchKernelGGL(k_b_scal_rz, dim3(gw), dim3(kBlock), 0, st, B->A);
    dft(B, st, true);
  }
};

cia_status read_counts(cia_rve_batch* B, hipStream_t st, int out[7]) {
  B->stats[4] += 1;
  hipLaunchKernelGGL(k_b_count, dim3(1), dim3(kBlock), 0, st, B->A, B->cnt);
  CIA_HIP(hipGetLastError());
  CIA_HIP(hipMemcpyAsync(B->h_cnt, B->cnt, 7 * sizeof(int), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 7; ++i) out[i] = B->h_cnt[i];
  return CIA_OK;
}

// Time `launch` with grid candidates; keep the fastest.
template <class L>
cia_status tune_one(cia_rve_batch* B, hipStream_t st, int k, int fam, L&& launch,
                    bool persistent = false) {
  // PCG kernels (persistent = true) keep to persistent grids over the active-tile
  // list: most of their launches come when part of the RVEs have converged, and a
  // one-tile-per-block grid would dispatch every tile of the batch every time.
  const int T = B->T[fam];
  std::vector<int> cand;
  if (!persistent || T <= 32 * B->cus) cand.push_back(T);  // one tile per block
  for (int m : {1, 2, 4, 8, 16, 32}) {
    const int g = m * B->cus;
    if (g < T && (persistent || m <= 16)) cand.push_back(g);
  }
  hipEvent_t e0, e1;
  CIA_HIP(hipEventCreate(&e0));
  CIA_HIP(hipEventCreate(&e1));
  float best = INFINITY;
  int best_g = T;
  for (int g : cand) {
    B->grid[k] = g;
    launch(st);  // warm-up
    CIA_HIP(hipEventRecord(e0, st));
    const int reps = 3;
    for (int i = 0; i < reps; ++i) launch(st);
    CIA_HIP(hipEventRecord(e1, st));
    CIA_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CIA_HIP(hipEventElapsedTime(&ms, e0, e1));
    const float us = 1000.f * ms / reps;
    if (us < best) {
      best = us;
      best_g = g;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CIA_HIP(hipGetLastError());
  B->grid[k] = best_g;
  B->tuned_us[k] = best;
  return CIA_OK;
}

// Autotune the grid of every batched kernel on the live operator: called after
// the first assembly (slot 0 holds valid tangents) and before the first PCG.
// Only kernels whose outputs are rewritten before they are read are timed on
// the live data; the assembly and gather recompute their own results.
template <class T>
cia_status autotune(cia_rve_batch* B, hipStream_t st, const Settings& S, const void* state_c,
                    const double* wc) {
  using SC = typename SpecOf<T>::type;
  B->A.force = 1;
  BOps<T>::compact(B, st);  // every tile listed while all RVEs are forced active
  cia_status s = CIA_OK;
  auto run = [&](cia_status x) {
    if (s == CIA_OK) s = x;
  };
  run(tune_one(B, st, K_ASM, F_GP, [&](hipStream_t q) {
    BOps<T>::assemble(B, q, S, S_NEWTON, state_c, wc);
  }));
  run(tune_one(B, st, K_GATHER, F_CL, [&](hipStream_t q) { BOps<T>::gather(B, q, S_NEWTON); }));
  run(tune_one(B, st, K_ELEM, F_CL, [&](hipStream_t q) { BOps<T>::elem(B, q, 0); },
               true));
  run(tune_one(B, st, K_NODE, F_CL, [&](hipStream_t q) {
    hipLaunchKernelGGL(k_b_node<T>, dim3(kgrid(B, K_NODE, F_CL)), dim3(kBlock), 0, q, B->A, 0);
  }, true));
  run(tune_one(B, st, K_UPD, F_CL, [&](hipStream_t q) {
    hipLaunchKernelGGL(k_b_update<SC>, dim3(kgrid(B, K_UPD, F_CL)), dim3(kBlock), 0, q, B->A, 0);
  }, true));
  run(tune_one(B, st, K_APPLY, F_CL, [&](hipStream_t q) {
    hipLaunchKernelGGL(k_b_apply<SC>, dim3(kgrid(B, K_APPLY, F_CL)), dim3(kBlock), 0, q, B->A);
  }, true));
  run(tune_one(B, st, K_DX, F_DX, [&](hipStream_t q) {
    hipLaunchKernelGGL((k_b_dft_x<false, SC>), dim3(kgrid(B, K_DX, F_DX)), dim3(256), B->lds_x, q, B->A);
  }, true));
  run(tune_one(B, st, K_DY, F_DY, [&](hipStream_t q) {
    hipLaunchKernelGGL((k_b_dft_strided<false, false, SC>), dim3(kgrid(B, K_DY, F_DY)), dim3(256),
                       B->lds_y, q, B->A);
  }, true));
  run(tune_one(B, st, K_DZ, F_DZ, [&](hipStream_t q) {
    hipLaunchKernelGGL((k_b_dft_strided<false, true, SC>), dim3(kgrid(B, K_DZ, F_DZ)), dim3(256),
                       B->lds_z, q, B->A);
  }, true));
  B->A.force = 0;
  if (s != CIA_OK) return s;
  // The timed assembly and gather recomputed slot 0, r[0] and the gather
  // partials bit for bit; the other timed kernels only touched buffers that the
  // PCG initialisation rewrites (z spectra, p, y, dw, the PCG residual, F).
  B->tuned = true;
  return CIA_OK;
}

template <class T>
cia_status solve_batch(cia_rve_batch* B, hipStream_t st, const void* state_c, const double* wc,
                       const double* w0, const double* cmacro, const double* macro,
                       const double* dt, const Settings& S, void* next_state, double* next_w,
                       cia_rve_result* res) {
  const int R = B->R;
  for (int r = 0; r < R; ++r) {
    Req q{};
    for (int m = 0; m < 6; ++m) {
      q.e[m] = macro[6 * r + m];
      q.ec[m] = cmacro[6 * r + m];
    }
    q.dt = dt[r];
    B->h_req[r] = q;
  }
  CIA_HIP(hipMemcpyAsync(B->req, B->h_req, R * sizeof(Req), hipMemcpyHostToDevice, st));
  const unsigned gcl = kgrid(B, K_UPD, F_CL);
  hipLaunchKernelGGL(k_b_begin, dim3(gcl), dim3(kBlock), 0, st, B->A, w0 ? w0 : wc);
  hipLaunchKernelGGL(k_b_ctl_init, dim3((R + kBlock - 1) / kBlock), dim3(kBlock), 0, st, B->A);
  BOps<T>::compact(B, st, S_NEWTON);
  BOps<T>::assemble(B, st, S, S_NEWTON, state_c, wc);
  BOps<T>::gather(B, st, S_NEWTON);
  CIA_HIP(hipGetLastError());
  if (!B->tuned) {
    if (B->autotune) {
      CIA_TRY(autotune<T>(B, st, S, state_c, wc));
    } else {
      B->tuned = true;
    }
  }
  if (!B->pair && B->cap) {
    hipGraph_t g = nullptr;
    CIA_HIP(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
    BOps<T>::iteration(B, B->cap, 0);
    BOps<T>::iteration(B, B->cap, 1);
    CIA_HIP(hipStreamEndCapture(B->cap, &g));
    const hipError_t e = hipGraphInstantiate(&B->pair, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CIA_HIP(e);
  }
  NewtonP N{};
  N.ns = B->newton;
  N.lin_rtol = B->lin.relative_tolerance;
  N.lin_atol = B->lin.absolute_tolerance;
  N.forcing_max = B->lin.forcing_max;
  N.adaptive = B->lin.adaptive_forcing;
  const unsigned gw = (unsigned)((R + 3) / 4);
  hipLaunchKernelGGL(k_b_ctl_first, dim3(gw), dim3(kBlock), 0, st, B->A, N);
  CIA_HIP(hipGetLastError());
  int cnt[7];
  const int interval = std::max(1, B->lin.check_interval);
  CIA_TRY(read_counts(B, st, cnt));
  for (int round = 0;; ++round) {
    if (cnt[S_PCG] == 0) break;
    B->stats[1] += 1;
    // Newton systems of every RVE in this round, solved together
    BOps<T>::pcg_start(B, st);
    int it = 0;
    const bool hinted = round < cia_rve_batch::kHints && B->its_hint[round] > 0;
    int chunk = hinted ? B->its_hint[round] : interval;
    while (it < B->lin.max_iterations) {
      const int n = std::min(chunk, B->lin.max_iterations - it);
      chunk = hinted ? std::max(2, interval / 4) : interval;
      if (B->pair) {
        // whole pairs; an odd final iteration below the limit runs on its own
        const int pairs = std::min((n + 1) / 2, (B->lin.max_iterations - it) / 2);
        for (int s = 0; s < pairs; ++s, it += 2) CIA_HIP(hipGraphLaunch(B->pair, st));
        if (pairs == 0) {
          BOps<T>::iteration(B, st, it & 1);
          ++it;
        }
      } else {
        for (int s = 0; s < n; ++s, ++it) BOps<T>::iteration(B, st, it & 1);
      }
      CIA_HIP(hipGetLastError());
      CIA_TRY(read_counts(B, st, cnt));
      if (cnt[5] == 0) break;
    }
    if (round < cia_rve_batch::kHints) B->its_hint[round] = std::max(1, cnt[6]);
    B->stats[2] += it;
    hipLaunchKernelGGL(k_b_norm_dw, dim3(gcl), dim3(kBlock), 0, st, B->A);
    hipLaunchKernelGGL(k_b_ctl_pcg, dim3(gw), dim3(kBlock), 0, st, B->A, N);
    CIA_HIP(hipGetLastError());
    // Line search: one step at once (empty lists when no RVE entered it), then
    // until no RVE is left in it; the last count also opens the next round.
    do {
      BOps<T>::compact(B, st, S_LS);
      hipLaunchKernelGGL(k_b_axpy, dim3(gcl), dim3(kBlock), 0, st, B->A);
      BOps<T>::assemble(B, st, S, S_LS, state_c, wc);
      BOps<T>::gather(B, st, S_LS);
      hipLaunchKernelGGL(k_b_ctl_ls, dim3(gw), dim3(kBlock), 0, st, B->A, N);
      CIA_HIP(hipGetLastError());
      CIA_TRY(read_counts(B, st, cnt));
    } while (cnt[S_LS] > 0);
  }
  // converged RVEs: final pass (energies, next state, response sums), next w
  hipLaunchKernelGGL(k_b_final<T>, dim3(kgrid(B, K_ASM, F_GP)), dim3(kBlock), 0, st, B->A,
                     (const T*)state_c, (T*)next_state, wc);
  hipLaunchKernelGGL(k_b_phase_sums<T>, dim3(gcl), dim3(kBlock), 0, st, B->A);
  hipLaunchKernelGGL(k_b_resp, dim3(gw), dim3(kBlock), 0, st, B->A, B->resp);
  hipLaunchKernelGGL(k_b_copy_out, dim3(gcl), dim3(kBlock), 0, st, B->A, next_w);
  CIA_HIP(hipGetLastError());
  std::vector<Ctl> hc(R);
  constexpr int NR = B_NRESP + B_NPHASE_ROWS;
  std::vector<double> hr((size_t)R * NR);
  CIA_HIP(hipMemcpyAsync(hc.data(), B->ctl, R * sizeof(Ctl), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipMemcpyAsync(hr.data(), B->resp, hr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipStreamSynchronize(st));
  bool all_ok = true;
  B->stats[0] += 1;
  B->solved = true;
  for (int r = 0; r < R; ++r) B->stats[3] += hc[r].pcg_its;
  for (int r = 0; r < R; ++r) {
    cia_rve_result& o = res[r];
    memset(&o, 0, sizeof(o));
    o.point_failure.point_id = -1;
    const Ctl& c = hc[r];
    const Desc& D = B->h_desc[r];
    o.newton_iterations = c.it;
    o.armijo_backtracks = c.backtracks;
    o.linear_iterations = c.pcg_its;
    o.gp_updates = c.gp_updates;
    o.micro_residual_norm = c.rn;
    o.micro_residual_tolerance = c.rtol;
    o.volume = D.volume;
    if (c.state != S_DONE) {
      all_ok = false;
      o.reason = c.state == S_FAILED ? c.reason : CIA_RVE_MAXIMUM_ITERATIONS_EXCEEDED;
      o.failure_iteration = c.fail_it;
      o.failure_residual_norm = c.fail_rn;
      o.failure_residual_tolerance = c.fail_rtol;
      o.failure_increment_norm = c.fail_inc;
      o.failure_increment_tolerance = c.fail_itol;
      if (c.reason == CIA_RVE_MATERIAL_POINT_FAILURE && c.fail_slot >= 0) {
        hipLaunchKernelGGL(k_b_diagnose<T>, dim3(1), dim3(1), 0, st, B->A, S, r, c.fail_slot,
                           (const T*)state_c, wc, (int64_t)c.first_fail, B->rec);
        CIA_HIP(hipGetLastError());
        CIA_HIP(hipMemcpyAsync(&o.point_failure, B->rec, sizeof(cia_point_failure),
                               hipMemcpyDeviceToHost, st));
        CIA_HIP(hipStreamSynchronize(st));
      }
      continue;
    }
    o.reason = CIA_RVE_OK;
    const double* s = &hr[(size_t)r * NR];
    const double jw = D.g.jw, V = D.volume;
    for (int m = 0; m < 6; ++m) o.macro_stress[m] = s[BR_SIG + m] * jw / V;
    o.free_energy_density = s[BR_PSI] * jw / V;
    o.dissipation_density = s[BR_DISS] * jw / V;
    const double micro = s[BR_MICRO] * jw / V;
    double macro_w = 0.0;
    for (int m = 0; m < 6; ++m) macro_w += o.macro_stress[m] * (macro[6 * r + m] - cmacro[6 * r + m]);
    o.macro_work_increment = macro_w;
    o.micro_work_increment = micro;
    o.hill_mandel_residual = macro_w - micro;
    o.hill_mandel_error = fabs(macro_w - micro) / (fabs(macro_w) + fabs(micro) + B->hm_scale * dt[r]);
    o.positive_work_density = s[BR_POS] * jw / V;
    o.negative_work_density = s[BR_NEG] * jw / V;
    o.unloading_volume_fraction = s[BR_UNLOAD] * jw / V;
    o.viscoplastic_active_volume_fraction = s[BR_ACTIVE] * jw / V;
    o.minimum_yield_activation_margin = NAN;
    o.maximum_yield_activation_margin = NAN;
    o.virtual_work_stress_error = NAN;
    o.periodic_fluctuation_error = NAN;
    o.reaction_antiperiodicity_error = NAN;
    o.periodic_class_reaction_error = NAN;
    for (int p = 0; p < 8; ++p) {
      const double* ps = s + B_NRESP + 9 * p;
      const double pv = p < B_PHASES ? ps[0] * jw : NAN;
      o.phase_volume[p] = p < B_PHASES ? pv : NAN;
      for (int m = 0; m < 6; ++m)
        o.phase_average_stress[p][m] = p < B_PHASES && pv > 0 ? ps[1 + m] * jw / pv : NAN;
      o.phase_free_energy_density[p] = p < B_PHASES && pv > 0 ? ps[7] * jw / pv : NAN;
      o.phase_dissipation_density[p] = p < B_PHASES && pv > 0 ? ps[8] * jw / pv : NAN;
    }
  }
  return all_ok ? CIA_OK : CIA_ERR_RVE_FAILURE;
}

// Effective tangents of the RVEs that converged in the last solve: six batched PCG
// solves with the converged tangents, one per macro strain direction (kernels
// k_b_tan_begin .. k_b_tan_end above).  tangents: host [R][36], row-major; RVEs that
// did not converge read NaN; linear_iterations (if given, host [R]) receives each
// RVE's PCG iterations of the six solves.
template <class T>
cia_status tangent_batch(cia_rve_batch* B, hipStream_t st, double* tangents, int32_t* its) {
  const int R = B->R;
  if (!B->tan_out) CIA_TRY(balloc(B, (void**)&B->tan_out, (size_t)R * 36 * sizeof(double)));
  const unsigned gr = (unsigned)((R + kBlock - 1) / kBlock), gw = (unsigned)((R + 3) / 4);
  std::vector<Ctl> h0(R);
  CIA_HIP(hipMemcpyAsync(h0.data(), B->ctl, R * sizeof(Ctl), hipMemcpyDeviceToHost, st));
  hipLaunchKernelGGL(k_b_tan_begin, dim3(gr), dim3(kBlock), 0, st, B->A, B->lin.relative_tolerance);
  CIA_HIP(hipGetLastError());
  int cnt[7];
  CIA_TRY(read_counts(B, st, cnt));
  const int interval = std::max(1, B->lin.check_interval);
  for (int j = 0; j < 6 && cnt[S_PCG] > 0; ++j) {
    BOps<T>::compact(B, st);
    hipLaunchKernelGGL(k_b_affine<T>, dim3(kgrid(B, K_ELEM, F_CL)), dim3(kBlock), 0, st, B->A, j);
    BOps<T>::gather(B, st, S_PCG);
    BOps<T>::pcg_start(B, st);
    int it = 0;
    while (it < B->lin.max_iterations) {
      const int n = std::min(interval, B->lin.max_iterations - it);
      if (B->pair) {
        const int pairs = std::min((n + 1) / 2, (B->lin.max_iterations - it) / 2);
        for (int s2 = 0; s2 < pairs; ++s2, it += 2) CIA_HIP(hipGraphLaunch(B->pair, st));
        if (pairs == 0) {
          BOps<T>::iteration(B, st, it & 1);
          ++it;
        }
      } else {
        for (int s2 = 0; s2 < n; ++s2, ++it) BOps<T>::iteration(B, st, it & 1);
      }
      CIA_HIP(hipGetLastError());
      CIA_TRY(read_counts(B, st, cnt));
      if (cnt[5] == 0) break;
    }
    B->stats[2] += it;
    hipLaunchKernelGGL(k_b_ceff<T>, dim3(kgrid(B, K_UPD, F_CL)), dim3(kBlock), 0, st, B->A, j);
    hipLaunchKernelGGL(k_b_ceff_sum, dim3(gw), dim3(kBlock), 0, st, B->A, j, B->tan_out);
    CIA_HIP(hipGetLastError());
    if (j < 5) CIA_TRY(read_counts(B, st, cnt));
  }
  hipLaunchKernelGGL(k_b_tan_end, dim3(gr), dim3(kBlock), 0, st, B->A);
  CIA_HIP(hipGetLastError());
  std::vector<Ctl> hc(R);
  std::vector<double> ht((size_t)R * 36);
  CIA_HIP(hipMemcpyAsync(hc.data(), B->ctl, R * sizeof(Ctl), hipMemcpyDeviceToHost, st));
  CIA_HIP(hipMemcpyAsync(ht.data(), B->tan_out, ht.size() * sizeof(double), hipMemcpyDeviceToHost,
                         st));
  CIA_HIP(hipStreamSynchronize(st));
  bool ok = true;
  for (int r = 0; r < R; ++r) {
    const bool done = h0[r].state == S_DONE && hc[r].state == S_DONE;
    ok = ok && done;
    for (int k = 0; k < 36; ++k) tangents[36 * r + k] = done ? ht[36 * r + k] : NAN;
    if (its) its[r] = hc[r].pcg_its - h0[r].pcg_its;
  }
  return ok ? CIA_OK : CIA_ERR_RVE_FAILURE;
}

}  // namespace

extern "C" {

cia_status cia_rve_batch_destroy(cia_rve_batch* B) {
  if (!B) return CIA_OK;
  for (void* p : B->allocs) (void)hipFree(p);
  if (B->h_req) (void)hipHostFree(B->h_req);
  if (B->h_cnt) (void)hipHostFree(B->h_cnt);
  if (B->pair) (void)hipGraphExecDestroy(B->pair);
  if (B->cap) (void)hipStreamDestroy(B->cap);
  delete B;
  return CIA_OK;
}

cia_status cia_rve_batch_create(const cia_rve_desc* descs, int32_t n, int32_t autotune,
                                cia_rve_batch** out) {
  CIA_ARG(descs && out && n >= 1 && n <= 65535, "invalid batch descriptor array");
  const cia_rve_desc& d0 = descs[0];
  for (int r = 0; r < n; ++r) {
    const cia_rve_desc& d = descs[r];
    CIA_ARG(d.nx >= 2 && d.ny >= 2 && d.nz >= 2 && d.nx <= 64 && d.ny <= 64 && d.nz <= 64,
            "batched RVE edges must lie in [2, 64] elements");
    CIA_ARG(d.n_phases >= 1 && d.n_phases <= CIA_MAX_PHASES && d.phases, "invalid phases");
    CIA_ARG(d.lengths[0] > 0 && d.lengths[1] > 0 && d.lengths[2] > 0, "lengths must be positive");
    CIA_ARG(d.precision == d0.precision && d.hill_mandel_power_scale == d0.hill_mandel_power_scale &&
                memcmp(&d.newton, &d0.newton, sizeof(d.newton)) == 0 &&
                memcmp(&d.linear, &d0.linear, sizeof(d.linear)) == 0,
            "all RVEs of a batch share precision, Newton and linear-solver settings");
  }
  CIA_ARG(d0.precision == CIA_PREC_FP64 || d0.precision == CIA_PREC_FP32_REFINED, "bad precision");
  CIA_ARG(d0.linear.preconditioner == CIA_PCG_SPECTRAL,
          "the batched solver uses the spectral preconditioner");
  auto* B = new cia_rve_batch();
  B->R = n;
  B->precision = d0.precision;
  B->tsize = d0.precision == CIA_PREC_FP64 ? sizeof(double) : sizeof(float);
  B->newton = d0.newton;
  B->lin = d0.linear;
  if (B->lin.max_iterations <= 0) B->lin.max_iterations = 20000;
  if (B->lin.check_interval <= 0) B->lin.check_interval = 16;
  B->hm_scale = d0.hill_mandel_power_scale;
  B->autotune = autotune != 0;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      B->cus = prop.multiProcessorCount;
  }
  // sizes and tiles
  B->h_desc.resize(n);
  int64_t e0 = 0;
  int tcount[NF] = {0};
  size_t lx = 0, ly = 0, lz = 0;
  for (int r = 0; r < n; ++r) {
    const cia_rve_desc& d = descs[r];
    Desc& D = B->h_desc[r];
    D.nx = d.nx;
    D.ny = d.ny;
    D.nz = d.nz;
    D.lines_x = std::max(1, 1024 / d.nx);  // x-lines per tile of the x pass
    D.e0 = e0;
    D.E = (int64_t)d.nx * d.ny * d.nz;
    const double hx = d.lengths[0] / d.nx, hy = d.lengths[1] / d.ny, hz = d.lengths[2] / d.nz;
    D.g.ihx = 1.0 / hx;
    D.g.ihy = 1.0 / hy;
    D.g.ihz = 1.0 / hz;
    D.g.jw = hx * hy * hz / 8.0;
    D.volume = D.g.jw * 8.0 * (double)D.E;
    D.inv_n = 1.0 / (double)D.E;
    D.j2mask = 0;
    for (int i = 0; i < d.n_phases; ++i)
      if (d.phases[i].kind == CIA_MAT_J2_VISCOPLASTIC) D.j2mask |= 1 << i;
    const int nt[NF] = {
        (int)((8 * D.E + kBlock - 1) / kBlock), (int)((D.E + kBlock - 1) / kBlock),
        (int)((2 * (int64_t)d.ny * d.nz + D.lines_x - 1) / D.lines_x),
        ((d.nx + kBatchDftTile - 1) / kBatchDftTile) * d.nz * 2,
        ((d.nx + kBatchDftTile - 1) / kBatchDftTile) * d.ny * 2};
    for (int f = 0; f < NF; ++f) {
      D.t0[f] = tcount[f];
      D.nt[f] = nt[f];
      tcount[f] += nt[f];
    }
    // spectra in FP64 for FP64 batches, in FP32 for FP32 batches
    const size_t csz = d0.precision == CIA_PREC_FP64 ? sizeof(double2) : sizeof(float2);
    lx = std::max(lx, 2 * (size_t)D.lines_x * (d.nx + 1) * csz);
    ly = std::max(ly, 2 * (size_t)kBatchDftTile * (d.ny + 1) * csz);
    lz = std::max(lz, 2 * (size_t)kBatchDftTile * (d.nz + 1) * csz);
    e0 += D.E;
  }
  B->ET = e0;
  B->NPT = 8 * e0;
  B->lds_x = lx;
  B->lds_y = ly;
  B->lds_z = lz;
  {
    // strided passes of edges above 32 in FP64 take more than 64 KB of LDS
    const int big = (int)std::max(ly, lz);
    hipError_t e = hipSuccess;
    if (big > 65536) {
      if (d0.precision == CIA_PREC_FP64) {
        e = set_dft_lds<double2>(big);
      } else {
        e = set_dft_lds<float2>(big);
      }
    }
    if (e != hipSuccess) {
      set_error("hipFuncSetAttribute", e, __FILE__, __LINE__);
      delete B;
      return CIA_ERR_HIP;
    }
  }
  for (int f = 0; f < NF; ++f) {
    B->T[f] = tcount[f];
    B->Tmax = std::max(B->Tmax, tcount[f]);
  }
  const size_t ts = B->tsize;
  const int64_t NPT = B->NPT, ET = B->ET;
  cia_status s = CIA_OK;
#define BALLOC(ptr, bytes)                         \
  do {                                             \
    s = balloc(B, (void**)&(ptr), (bytes));        \
    if (s != CIA_OK) {                             \
      cia_rve_batch_destroy(B);                    \
      return s;                                    \
    }                                              \
  } while (0)
  Arena& A = B->A;
  A.R = n;
  A.lin_atol = B->lin.absolute_tolerance;
  A.Tmax = B->Tmax;
  A.NPT = NPT;
  A.ET = ET;
  for (int f = 0; f < NF; ++f) A.T[f] = B->T[f];
  for (int f = 0; f < NF; ++f) BALLOC(A.act[f], std::max(1, B->T[f]) * sizeof(int));
  for (int fam : {F_GP, F_CL}) {
    std::vector<int> hg;
    for (int r = 0; r < n; ++r)
      for (int u = 0; u < B->h_desc[r].nt[fam]; u += kFinalGroup) hg.push_back(B->h_desc[r].t0[fam] + u);
    int* grp = nullptr;
    BALLOC(grp, std::max<size_t>(1, hg.size()) * sizeof(int));
    if (!hg.empty() &&
        hipMemcpy(grp, hg.data(), hg.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      cia_rve_batch_destroy(B);
      set_error_msg("hipMemcpy of the batch tile groups failed");
      return CIA_ERR_HIP;
    }
    if (fam == F_GP) {
      A.grp = grp;
      A.n_grp = (int)hg.size();
    } else {
      A.grp_cl = grp;
      A.n_grp_cl = (int)hg.size();
    }
  }
  BALLOC(A.n_act, NF * sizeof(int));
  BALLOC(A.act_off, (size_t)NF * n * sizeof(int));
  Desc* d_desc = nullptr;
  BALLOC(d_desc, n * sizeof(Desc));
  for (int f = 0; f < NF; ++f) {
    uint16_t* t2r = nullptr;
    BALLOC(t2r, std::max(1, B->T[f]) * sizeof(uint16_t));
    std::vector<uint16_t> h(B->T[f]);
    for (int r = 0; r < n; ++r)
      for (int u = 0; u < B->h_desc[r].nt[f]; ++u) h[B->h_desc[r].t0[f] + u] = (uint16_t)r;
    if (hipMemcpy(t2r, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess) {
      cia_rve_batch_destroy(B);
      set_error_msg("hipMemcpy of the batch tile table failed");
      return CIA_ERR_HIP;
    }
    A.t2r[f] = t2r;
  }
  PhaseTable* tabs = nullptr;
  BALLOC(tabs, n * sizeof(PhaseTable));
  uint8_t* eph = nullptr;
  BALLOC(eph, ET);
  {
    std::vector<PhaseTable> ht(n);
    std::vector<uint8_t> he(ET, 0);
    for (int r = 0; r < n; ++r) {
      const cia_rve_desc& d = descs[r];
      ht[r] = PhaseTable{};
      ht[r].n = d.n_phases;
      for (int i = 0; i < d.n_phases; ++i) {
        ht[r].p[i] = make_phase(d.phases[i]);
        if (ht[r].p[i].kind == CIA_MAT_J2_VISCOPLASTIC && ht[r].p[i].m != 1.0) B->all_m1 = false;
      }
      if (d.element_phase) {
        for (int64_t e = 0; e < B->h_desc[r].E; ++e) {
          const int32_t v = d.element_phase[e];
          if (v < 0 || v >= d.n_phases) {
            cia_rve_batch_destroy(B);
            set_error_msg("element_phase out of range");
            return CIA_ERR_INVALID_ARGUMENT;
          }
          he[B->h_desc[r].e0 + e] = (uint8_t)v;
        }
      }
    }
    if (hipMemcpy(tabs, ht.data(), n * sizeof(PhaseTable), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(eph, he.data(), ET, hipMemcpyHostToDevice) != hipSuccess) {
      cia_rve_batch_destroy(B);
      set_error_msg("hipMemcpy of the batch phase tables failed");
      return CIA_ERR_HIP;
    }
  }
  A.tabs = tabs;
  A.ephase = eph;
  // Spectral symbols and DFT twiddles, shared by RVEs of equal geometry and medium.
  {
    std::map<std::tuple<int, int, int, double, double, double, double, double>, size_t> sym_of;
    std::vector<hipFloatComplex> sym_all;
    std::map<int, size_t> tw_of;
    std::vector<double2> tw_all;
    std::vector<size_t> sym_off(n);
    std::vector<size_t> tw_off(3 * n);
    for (int r = 0; r < n; ++r) {
      const cia_rve_desc& d = descs[r];
      double kmin = INFINITY, kmax = 0.0, gmin = INFINITY, gmax = 0.0;
      for (int i = 0; i < d.n_phases; ++i) {
        const PhaseConst p = make_phase(d.phases[i]);
        kmin = std::min(kmin, p.K);
        kmax = std::max(kmax, p.K);
        gmin = std::min(gmin, p.G);
        gmax = std::max(gmax, p.G);
      }
      const double K0 = d.linear.reference_bulk > 0 ? d.linear.reference_bulk : std::sqrt(kmin * kmax);
      const double G0 = d.linear.reference_shear > 0 ? d.linear.reference_shear : std::sqrt(gmin * gmax);
      const auto key = std::make_tuple(d.nx, d.ny, d.nz, d.lengths[0], d.lengths[1], d.lengths[2], K0, G0);
      auto it = sym_of.find(key);
      if (it == sym_of.end()) {
        std::vector<hipFloatComplex> sym;
        spectral_symbol(d.nx, d.ny, d.nz, B->h_desc[r].g, K0, G0, sym);
        it = sym_of.emplace(key, sym_all.size()).first;
        sym_all.insert(sym_all.end(), sym.begin(), sym.end());
      }
      sym_off[r] = it->second;
      const int edges[3] = {d.nx, d.ny, d.nz};
      for (int a = 0; a < 3; ++a) {
        auto jt = tw_of.find(edges[a]);
        if (jt == tw_of.end()) jt = tw_of.emplace(edges[a], dft_twiddles(edges[a], tw_all)).first;
        tw_off[3 * r + a] = jt->second;
      }
    }
    hipFloatComplex* kinv = nullptr;
    double2* tw = nullptr;
    float2* twf = nullptr;
    std::vector<float2> twf_all(tw_all.size());
    for (size_t i = 0; i < tw_all.size(); ++i)
      twf_all[i] = make_float2((float)tw_all[i].x, (float)tw_all[i].y);
    BALLOC(kinv, sym_all.size() * sizeof(hipFloatComplex));
    BALLOC(tw, tw_all.size() * sizeof(double2));
    BALLOC(twf, twf_all.size() * sizeof(float2));
    if (hipMemcpy(kinv, sym_all.data(), sym_all.size() * sizeof(hipFloatComplex),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(tw, tw_all.data(), tw_all.size() * sizeof(double2), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(twf, twf_all.data(), twf_all.size() * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess) {
      cia_rve_batch_destroy(B);
      set_error_msg("hipMemcpy of the batch spectral tables failed");
      return CIA_ERR_HIP;
    }
    for (int r = 0; r < n; ++r) {
      Desc& D = B->h_desc[r];
      D.kinv = kinv + sym_off[r];
      auto axis = [&](int nn, size_t off) {
        DftAxis a{};
        a.n = nn;
        dft_factor(nn, a.n1, a.n2);
        a.w = tw + off;
        a.wf = twf + off;
        return a;
      };
      D.ax = axis(D.nx, tw_off[3 * r + 0]);
      D.ay = axis(D.ny, tw_off[3 * r + 1]);
      D.az = axis(D.nz, tw_off[3 * r + 2]);
    }
  }
  if (hipMemcpy(d_desc, B->h_desc.data(), n * sizeof(Desc), hipMemcpyHostToDevice) != hipSuccess) {
    cia_rve_batch_destroy(B);
    set_error_msg("hipMemcpy of the batch descriptors failed");
    return CIA_ERR_HIP;
  }
  A.desc = d_desc;
  for (int k = 0; k < 2; ++k) {
    BALLOC(A.st[k], 8 * NPT * ts);
    BALLOC(A.sig[k], 6 * NPT * ts);
    BALLOC(A.tan[k], 8 * NPT * ts);
    BALLOC(A.w[k], 3 * ET * 8);
    BALLOC(A.r[k], 3 * ET * 8);
    BALLOC(A.pp[k], 3 * ET * 8);
  }
  BALLOC(A.F, 24 * ET * ts);
  BALLOC(A.pd, 2 * NPT * ts);
  BALLOC(A.dw, 3 * ET * 8);
  BALLOC(A.pr, 3 * ET * 8);
  BALLOC(A.py, 3 * ET * 8);
  BALLOC(A.uab, 2 * ET * (d0.precision == CIA_PREC_FP64 ? sizeof(double2) : sizeof(float2)));
  BALLOC(A.part, (size_t)P_ROWS * B->Tmax * 8);
  BALLOC(B->ctl, n * sizeof(Ctl));
  BALLOC(B->req, n * sizeof(Req));
  BALLOC(B->resp, (size_t)n * (B_NRESP + B_NPHASE_ROWS) * 8);
  BALLOC(B->cnt, 64);
  BALLOC(B->rec, sizeof(cia_point_failure));
#undef BALLOC
  A.ctl = B->ctl;
  A.req = B->req;
  if (hipHostMalloc((void**)&B->h_req, n * sizeof(Req)) != hipSuccess ||
      hipHostMalloc((void**)&B->h_cnt, 64) != hipSuccess) {
    cia_rve_batch_destroy(B);
    set_error_msg("hipHostMalloc of the batch staging buffers failed");
    return CIA_ERR_OUT_OF_MEMORY;
  }
  if (hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking) != hipSuccess) B->cap = nullptr;
  *out = B;
  return CIA_OK;
}

cia_status cia_rve_batch_layout(const cia_rve_batch* B, int64_t* element_offsets) {
  CIA_ARG(B && element_offsets, "NULL argument");
  for (int r = 0; r < B->R; ++r) element_offsets[r] = B->h_desc[r].e0;
  element_offsets[B->R] = B->ET;
  return CIA_OK;
}

cia_status cia_rve_batch_solve(cia_rve_batch* B, void* stream, const void* committed_state,
                               const double* committed_w, const double* initial_w,
                               const double* committed_macro, const double* macro_strain,
                               const double* time_step, const cia_update_settings* settings,
                               void* next_state, double* next_w, cia_rve_result* results) {
  CIA_ARG(B && committed_state && committed_w && committed_macro && macro_strain && time_step &&
              settings && next_state && next_w && results,
          "NULL argument to cia_rve_batch_solve");
  CIA_ARG((const void*)next_state != committed_state, "next_state aliases committed_state");
  CIA_ARG((const void*)next_w != (const void*)committed_w, "next_w aliases committed_w");
  CIA_ARG(initial_w == nullptr || (const void*)initial_w != (const void*)next_w,
          "next_w aliases initial_w");
  for (int r = 0; r < B->R; ++r) CIA_ARG(time_step[r] > 0.0, "RVE time step must be positive.");
  const Settings S = make_settings(*settings);
  hipStream_t st = (hipStream_t)stream;
  if (B->precision == CIA_PREC_FP64)
    return solve_batch<double>(B, st, committed_state, committed_w, initial_w, committed_macro,
                               macro_strain, time_step, S, next_state, next_w, results);
  return solve_batch<float>(B, st, committed_state, committed_w, initial_w, committed_macro,
                            macro_strain, time_step, S, next_state, next_w, results);
}

cia_status cia_rve_batch_effective_tangent(cia_rve_batch* B, void* stream, double* tangents,
                                           int32_t* linear_iterations) {
  CIA_ARG(B && tangents, "NULL argument to cia_rve_batch_effective_tangent");
  CIA_ARG(B->solved, "cia_rve_batch_effective_tangent needs a preceding cia_rve_batch_solve");
  hipStream_t st = (hipStream_t)stream;
  if (B->precision == CIA_PREC_FP64) return tangent_batch<double>(B, st, tangents, linear_iterations);
  return tangent_batch<float>(B, st, tangents, linear_iterations);
}

cia_status cia_rve_batch_tuning(const cia_rve_batch* B, int32_t* grids, double* microseconds,
                                int32_t n) {
  CIA_ARG(B && grids && n >= 0, "NULL argument");
  for (int k = 0; k < n && k < K_N; ++k) {
    grids[k] = B->grid[k];
    if (microseconds) microseconds[k] = B->tuned_us[k];
  }
  return CIA_OK;
}

cia_status cia_rve_batch_stats(const cia_rve_batch* B, int64_t* counters, int32_t n) {
  CIA_ARG(B && counters && n >= 0, "NULL argument");
  for (int i = 0; i < n && i < 5; ++i) counters[i] = B->stats[i];
  return CIA_OK;
}

cia_status cia_rve_batch_set_tuning(cia_rve_batch* B, const int32_t* grids, int32_t n) {
  CIA_ARG(B && grids && n == K_N, "grids must hold one entry per tuned kernel");
  CIA_ARG(!B->pair, "the tuning of a batch is fixed once its first solve captured the PCG graph");
  for (int k = 0; k < K_N; ++k) {
    CIA_ARG(grids[k] >= 0, "grid sizes must be non-negative (0: one tile per block)");
    B->grid[k] = grids[k];
  }
  B->tuned = true;
  return CIA_OK;
}

}  // extern "C"
