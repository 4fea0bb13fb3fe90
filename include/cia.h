/*
 * cia.h -- C ABI of libcia_hip.so, the MI355X (gfx950) FEM hot path.
 *
 * Every pointer documented "device" is device memory owned by the caller
 * (for example a torch tensor's data_ptr()).  Arrays of per-point data are
 * structure-of-arrays, component-major: component c of point p lives at
 * [c * n + p].  Voigt order is xx, yy, zz, yz, xz, xy; strains carry
 * engineering shear, stresses are tensor components.  The "stream" argument is
 * a hipStream_t (NULL = the default stream).  Calls return cia_status; a
 * material-point failure fills the caller's cia_point_failure record.
 *
 * The Python host classes that mirror the reference interface bind these
 * entry points through ctypes (concrete_impact_ai_amd/_lib.py).
 */
#ifndef CIA_H
#define CIA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cia_status;

enum {
  CIA_OK = 0,
  CIA_ERR_INVALID_ARGUMENT = 1,
  CIA_ERR_HIP = 2,
  CIA_ERR_POINT_FAILURE = 3, /* a material point failed; see cia_point_failure */
  CIA_ERR_RVE_FAILURE = 4,   /* RVE Newton failed; see cia_rve_result.reason */
  CIA_ERR_OUT_OF_MEMORY = 5,
  CIA_ERR_NOT_CONVERGED = 6
};

/* Material-point failure reasons.  Each maps 1:1 onto the reason string of
 * MaterialPointConvergenceError in src/fem/materials/plasticity.py:543-854,1441-1465. */
enum {
  CIA_PT_OK = 0,
  CIA_PT_NON_FINITE_J2_INVARIANT = 1,
  CIA_PT_NEGATIVE_J2_INVARIANT = 2,
  CIA_PT_INVALID_YIELD_ACTIVATION_TOLERANCE = 3,
  CIA_PT_NON_FINITE_CLOSED_FORM = 4,
  CIA_PT_CLOSED_FORM_SOLUTION_OUT_OF_BOUNDS = 5,
  CIA_PT_INVALID_ITERATION_LIMIT = 6,
  CIA_PT_NON_FINITE_PROBLEM_SCALE = 7,
  CIA_PT_INVALID_CORRECTED_OVERSTRESS = 8,
  CIA_PT_NON_FINITE_ITERATION_VALUE = 9,
  CIA_PT_NON_POSITIVE_DERIVATIVE = 10,
  CIA_PT_CONVERGED_SOLUTION_OUT_OF_BOUNDS = 11,
  CIA_PT_NON_FINITE_NEWTON_CANDIDATE = 12,
  CIA_PT_NEWTON_CANDIDATE_OUT_OF_BOUNDS = 13,
  CIA_PT_MAXIMUM_ITERATIONS_EXCEEDED = 14,
  CIA_PT_NON_POSITIVE_TANGENT_OVERSTRESS = 15 /* plasticity.py:918 RuntimeError */
};

/* Material kinds of the phase table. */
enum { CIA_MAT_LINEAR_ELASTIC = 0, CIA_MAT_J2_VISCOPLASTIC = 1 };

/* One phase: J2ViscoplasticMaterial (plasticity.py:79-91) or
 * LinearElasticMaterial (linear_elastic.py:25-32). */
typedef struct {
  int32_t kind;
  int32_t pad;
  double young_modulus;
  double poisson_ratio;
  double yield_stress;
  double hardening_modulus;
  double time_scale;
  double reference_stress;
  double rate_exponent;
  double density;
} cia_material;

/* MaterialUpdateSettings (src/fem/materials/data.py:16-24). */
typedef struct {
  int32_t max_iterations;
  int32_t pad;
  double yield_relative_tolerance;
  double residual_absolute_tolerance;
  double residual_relative_tolerance;
} cia_update_settings;

/* Output selection flags of cia_material_update_*. */
enum {
  CIA_OUT_TANGENT = 1,     /* full 6x6 algorithmic tangent, [36][n] row-major per point */
  CIA_OUT_ENERGY = 2,      /* free energy and dissipation, [n] each */
  CIA_OUT_DIAGNOSTICS = 4  /* CIA_N_DIAG diagnostic rows, [CIA_N_DIAG][n] */
};

/* Diagnostic rows written with CIA_OUT_DIAGNOSTICS, in this order.  Names and
 * meaning follow plasticity.py:578-595,641-678,704-708,1321-1327. */
enum {
  CIA_D_YIELD_VALUE = 0,
  CIA_D_PLASTIC_WORK,
  CIA_D_LOCAL_ITERATIONS,
  CIA_D_ACTIVE_SURFACE,
  CIA_D_PLASTIC_MULTIPLIER,
  CIA_D_VISCOPLASTIC_MULTIPLIER,
  CIA_D_EQUIVALENT_STRESS,
  CIA_D_TRIAL_YIELD_VALUE,
  CIA_D_YIELD_SCALE,
  CIA_D_YIELD_ACTIVATION_THRESHOLD,
  CIA_D_YIELD_ACTIVATION_MARGIN,
  CIA_D_VISCOPLASTIC_ACTIVE,
  CIA_D_TIME_STEP,
  CIA_D_EQUIVALENT_PLASTIC_STRAIN,
  CIA_D_VISCOPLASTIC_HV,
  CIA_D_PLASTIC_POWER,
  CIA_D_HARDENING_STORAGE_RATE,
  CIA_D_LOCAL_RESIDUAL,
  CIA_D_ROOT_LOWER_BOUND,
  CIA_D_ROOT_UPPER_BOUND,
  CIA_D_INCREMENTAL_WORK_DENSITY,
  CIA_D_EQUIVALENT_STRESS_INCREMENT,
  CIA_N_DIAG
};

/* First failing point of a material update (lowest point id), with the
 * scalars of the reference's failure diagnostics (plasticity.py:857-895). */
typedef struct {
  int64_t point_id;        /* -1 when no point failed */
  int32_t reason;          /* CIA_PT_* */
  int32_t iteration;
  double plastic_multiplier;
  double corrected_overstress;
  double residual;
  double derivative;
  double newton_candidate;
  double trial_yield;
  double physical_upper_bound;
  double residual_tolerance;
  double trial_equivalent_stress;
  double yield_radius;
  double equivalent_plastic_strain;
  double radicand;
  double trial_stress[6];
  double strain[6];
} cia_point_failure;

/* Library / device information.  No reference counterpart (host plumbing). */
int32_t cia_abi_version(void);
cia_status cia_device_name(char* buffer, int32_t capacity);
/* Message of the last failed HIP call or rejected argument on this thread. */
cia_status cia_last_error(char* buffer, int32_t capacity);

/*
 * Material-point update -- replaces J2ViscoplasticMaterial.update /
 * LinearElasticMaterial.update / PhasedMaterialModel.update
 * (src/fem/materials/plasticity.py:110-117,343-431;
 *  src/fem/materials/linear_elastic.py:44-52,74-100;
 *  src/fem/materials/phased.py:67-143).
 *
 *   n            number of points
 *   phases       host array of n_phases phase descriptors
 *   phase_id     device int32 [n] phase of each point, or NULL (all phase 0)
 *   strain       device [6][n] total strain (engineering shear)
 *   strain_rate  device [6][n] or NULL (enables the increment diagnostics)
 *   state_in     device [8][n] committed state rows: plastic_strain (6),
 *                equivalent_plastic_strain, viscoplastic_multiplier; ignored
 *                for elastic points; not modified
 *   stress       device [6][n]
 *   state_out    device [8][n] next state (elastic points copy state_in)
 *   tangent      device [36][n] if CIA_OUT_TANGENT, else may be NULL
 *   free_energy, dissipation  device [n] if CIA_OUT_ENERGY
 *   diagnostics  device [CIA_N_DIAG][n] if CIA_OUT_DIAGNOSTICS
 *   failure      host record, filled on CIA_ERR_POINT_FAILURE (point_id = -1 otherwise)
 *
 * The call synchronises the stream to report failures.
 */
cia_status cia_material_update_f64(
    void* stream, int64_t n, const cia_material* phases, int32_t n_phases,
    const int32_t* phase_id, const double* strain, const double* strain_rate,
    const double* state_in, double time_step, const cia_update_settings* settings,
    uint32_t flags, double* stress, double* state_out, double* tangent,
    double* free_energy, double* dissipation, double* diagnostics,
    cia_point_failure* failure);

/* FP32 variant: strain, stress, state and tangent in float; the
 * yield-consistency multiplier solve stays FP64 (north-star precision mode). */
cia_status cia_material_update_f32(
    void* stream, int64_t n, const cia_material* phases, int32_t n_phases,
    const int32_t* phase_id, const float* strain, const float* strain_rate,
    const float* state_in, double time_step, const cia_update_settings* settings,
    uint32_t flags, float* stress, float* state_out, float* tangent,
    float* free_energy, float* dissipation, float* diagnostics,
    cia_point_failure* failure);

/* ------------------------------------------------------------------------ */
/* Structured periodic Hex8 RVE (src/fem/rve/solver.py, periodic.py, path.py) */

/* NonlinearNewtonSettings + ArmijoSettings (src/fem/solvers/data.py:32-95). */
typedef struct {
  int32_t max_iterations;
  int32_t armijo_enabled;
  double residual_absolute_tolerance;
  double residual_relative_tolerance;
  double increment_absolute_tolerance;
  double increment_relative_tolerance;
  double armijo_sufficient_decrease;
  double armijo_reduction_factor;
  int32_t armijo_max_backtracks;
  int32_t pad;
} cia_newton_settings;

/* Linear solver of the Newton step: block-Jacobi PCG (no reference
 * counterpart; replaces spsolve, src/fem/solvers/linear.py:15-22). */
typedef struct {
  int32_t max_iterations;   /* per solve */
  int32_t check_interval;   /* iterations between host convergence checks */
  double relative_tolerance; /* ||r_k|| <= rtol * ||b|| */
  double absolute_tolerance;
  int32_t preconditioner;   /* CIA_PCG_* */
  int32_t adaptive_forcing; /* 1: inexact Newton, rtol_k = clamp(||r_k|| / ||r_0||) */
  double reference_bulk;    /* spectral reference medium (<= 0: geometric mean of phases) */
  double reference_shear;
  double forcing_max;       /* upper clamp of the adaptive relative tolerance (e.g. 0.1) */
} cia_linear_settings;

/* PCG preconditioners.  CIA_PCG_SPECTRAL is the exact inverse of the Hex8
 * stiffness of a homogeneous reference medium on the periodic grid, applied
 * with the library's own 3-D DFT kernels; CIA_PCG_BLOCK_JACOBI inverts the 3x3 nodal
 * blocks. */
enum { CIA_PCG_SPECTRAL = 0, CIA_PCG_BLOCK_JACOBI = 1 };

enum { CIA_PREC_FP64 = 0, CIA_PREC_FP32_REFINED = 1 };

typedef struct {
  int32_t nx, ny, nz;
  int32_t n_phases;
  double lengths[3];
  const cia_material* phases;     /* host [n_phases] */
  const int32_t* element_phase;   /* host [nx*ny*nz] element phase ids, NULL = phase 0 */
  cia_newton_settings newton;
  cia_linear_settings linear;
  double hill_mandel_power_scale;
  int32_t precision;              /* CIA_PREC_* */
  int32_t pad;
} cia_rve_desc;

/* RVE failure reasons (src/fem/rve/solver.py:129-141,143-158,214-218,455-459). */
enum {
  CIA_RVE_OK = 0,
  CIA_RVE_MAXIMUM_ITERATIONS_EXCEEDED = 1,
  CIA_RVE_NON_FINITE_INCREMENT = 2,
  CIA_RVE_ARMIJO_EXHAUSTED = 3,
  CIA_RVE_NON_FINITE_EFFECTIVE_TANGENT_SENSITIVITY = 4,
  CIA_RVE_MATERIAL_POINT_FAILURE = 5
};

/* Homogenised response and the scalar diagnostics of _build_rve_response
 * (src/fem/rve/solver.py:253-364). */
typedef struct {
  int32_t reason;                 /* CIA_RVE_* */
  int32_t newton_iterations;
  int32_t armijo_backtracks;
  int32_t linear_iterations;      /* PCG iterations summed over the increment */
  int64_t gp_updates;             /* constitutive point evaluations performed */
  double macro_stress[6];
  double effective_tangent[36];   /* row-major, valid when requested */
  double free_energy_density;
  double dissipation_density;
  double micro_residual_norm;
  double micro_residual_tolerance;
  double volume;
  double virtual_work_stress_error;
  double periodic_fluctuation_error;
  double reaction_antiperiodicity_error;
  double periodic_class_reaction_error;
  double macro_work_increment;
  double micro_work_increment;
  double hill_mandel_residual;
  double hill_mandel_error;
  double positive_work_density;
  double negative_work_density;
  double unloading_volume_fraction;
  double viscoplastic_active_volume_fraction;
  int32_t yield_activation_margin_available;
  int32_t pad;
  double minimum_yield_activation_margin;
  double maximum_yield_activation_margin;
  /* failure snapshot (solver.py:497-518) */
  double failure_residual_norm;
  double failure_residual_tolerance;
  double failure_increment_norm;
  double failure_increment_tolerance;
  int32_t failure_iteration;
  int32_t pad2;
  cia_point_failure point_failure;
  /* phase diagnostics (solver.py:391-435), up to 8 phases */
  double phase_volume[8];
  double phase_average_stress[8][6];
  double phase_free_energy_density[8];
  double phase_dissipation_density[8];
} cia_rve_result;

typedef struct cia_rve cia_rve;

/* Build the device operator of one structured periodic RVE
 * (StructuredHex8RVE + build_rve_operator_cache, src/fem/rve/data.py:88-113,
 *  src/fem/rve/solver.py:367-388). */
cia_status cia_rve_create(const cia_rve_desc* desc, cia_rve** out);
cia_status cia_rve_destroy(cia_rve* rve);
/* Sizes: n_elements, n_points (= 8 n_elements), n_classes, n_nodes. */
cia_status cia_rve_sizes(const cia_rve* rve, int64_t* sizes4);

/*
 * One strain-driven increment: solve_rve_microequilibrium
 * (src/fem/rve/solver.py:62-227).
 *
 *   committed_state  device [8][n_points] state rows (see cia_material_update_f64);
 *                    element type double for CIA_PREC_FP64, float for
 *                    CIA_PREC_FP32_REFINED
 *   committed_macro  host [6] macro strain of the committed state
 *   committed_w      device FP64 [3 * n_classes] periodic fluctuation per class
 *                    (class 0, the eliminated reference class, holds zeros;
 *                    the reference's reduced vector is entries 3..end)
 *   macro_strain     host [6] requested macro strain
 *   next_state       device [8][n_points], same element type (must not alias)
 *   next_w           device FP64 [3 * n_classes]
 *   result           host, filled on return
 *
 * Micro fields of the converged state stay inside the RVE object and are
 * copied out with cia_rve_copy_field.
 */
cia_status cia_rve_solve(
    cia_rve* rve, void* stream, const void* committed_state,
    const double* committed_macro, const double* committed_w,
    const double* macro_strain, double time_step,
    const cia_update_settings* settings, int32_t need_tangent,
    void* next_state, double* next_w, cia_rve_result* result);

/* Micro fields of the last converged increment. */
enum {
  CIA_FIELD_MICRO_STRAIN = 0,      /* [6][n_points] */
  CIA_FIELD_MICRO_STRAIN_RATE = 1, /* [6][n_points] */
  CIA_FIELD_MICRO_STRESS = 2,      /* [6][n_points] */
  CIA_FIELD_FREE_ENERGY = 3,       /* [n_points] */
  CIA_FIELD_DISSIPATION = 4,       /* [n_points] */
  CIA_FIELD_DIAGNOSTICS = 5,       /* [CIA_N_DIAG][n_points] */
  CIA_FIELD_DISPLACEMENT = 6,      /* [3 * n_nodes] full node-major displacement */
  CIA_FIELD_INTERNAL_FORCE = 7     /* [3 * n_nodes] full node-major internal force */
};
cia_status cia_rve_copy_field(cia_rve* rve, void* stream, int32_t field, double* dst);


/* ------------------------------------------------------------------------ */
/* Batched periodic RVEs: many independent structured RVEs of any sizes (edges
 * 2..64 elements) solved together, each increment being
 * solve_rve_microequilibrium (src/fem/rve/solver.py:62-227).  Replaces the
 * serial / process-pool loop over RVE solves of RVEHomogenizedMaterial.update
 * (src/fem/rve/material.py:104-191, workers :267-291) and of several
 * iterate_prescribed_rve_path drivers (src/fem/rve/path.py:47-100).
 *
 * All descriptors share precision, Newton and linear-solver settings
 * (spectral preconditioner) and the Hill-Mandel power scale; phases,
 * microstructure, lengths and divisions are per RVE.  With autotune = 1 the
 * first solve times each batched kernel over a set of grid sizes (one tile per
 * workgroup up to persistent grids of 1-16 workgroups per CU) and keeps the
 * fastest.
 *
 * Arena layout: RVE r owns elements (= periodic classes) [off[r], off[r+1])
 * from cia_rve_batch_layout (off has n + 1 entries) and points
 * [8 off[r], 8 off[r+1]).  State rows are [8][8 off[n]] (component-major over
 * the whole arena, element type by precision), fluctuations FP64 [3 off[n]]
 * (class 0 of every RVE holds zeros).
 *
 * cia_rve_batch_solve: committed_macro, macro_strain host [n][6], time_step
 * host [n]; results host [n].  initial_w (device, like committed_w, or NULL) is
 * the Newton start; NULL starts at the committed fluctuation as the reference
 * does.  Strain rates are always taken relative to committed_w.  Returns CIA_OK when every RVE converged, else
 * CIA_ERR_RVE_FAILURE with each RVE's reason in its result (next_state and
 * next_w rows of failed RVEs are not written).  The homogenised stress,
 * energies, work terms, Hill-Mandel residual, unloading and active fractions,
 * the iteration counts and the phase diagnostics of phases 0-3 are filled;
 * field diagnostics that need the full node set (periodic, reaction and
 * virtual-work errors, activation margins) and phases 4-7 read NaN, and micro
 * fields are not retained.
 */
typedef struct cia_rve_batch cia_rve_batch;
cia_status cia_rve_batch_create(const cia_rve_desc* descs, int32_t n, int32_t autotune,
                                cia_rve_batch** out);
cia_status cia_rve_batch_destroy(cia_rve_batch* batch);
cia_status cia_rve_batch_layout(const cia_rve_batch* batch, int64_t* element_offsets);
cia_status cia_rve_batch_solve(cia_rve_batch* batch, void* stream, const void* committed_state,
                               const double* committed_w, const double* initial_w,
                               const double* committed_macro, const double* macro_strain,
                               const double* time_step, const cia_update_settings* settings,
                               void* next_state, double* next_w, cia_rve_result* results);
/* cia_rve_batch_effective_tangent: the effective (consistent, homogenised) tangent
 * of every RVE that converged in the preceding cia_rve_batch_solve, by the Schur
 * complement of _compute_effective_tangent (src/fem/rve/solver.py:438-461): six
 * batched PCG solves K dw_j = -T^T K A e_j with the converged compact tangents, then
 * C_eff e_j = (1/V) sum_q w C_q (e_j + B_q dw_j).  tangents host [n][36] row-major
 * (NaN rows for RVEs that did not converge); linear_iterations host [n] or NULL.
 * Replaces the per-point single-RVE tangent solves of RVEHomogenizedMaterial.update
 * with requirements.tangent (src/fem/rve/material.py:104-191). */
cia_status cia_rve_batch_effective_tangent(cia_rve_batch* batch, void* stream, double* tangents,
                                           int32_t* linear_iterations);
/* Autotuned grid per batched kernel (0 = one tile per workgroup) and its time
 * in microseconds; order: assemble, gather, PCG element pass, PCG node pass,
 * PCG update, symbol solve, DFT x, y, z passes (CIA_BATCH_N_KERNELS entries).
 * No reference counterpart (the autotuner of the batched path). */
enum { CIA_BATCH_N_KERNELS = 9 };
cia_status cia_rve_batch_tuning(const cia_rve_batch* batch, int32_t* grids, double* microseconds,
                                int32_t n);
cia_status cia_rve_batch_set_tuning(cia_rve_batch* batch, const int32_t* grids, int32_t n);
/* Counters since creation (no reference counterpart; measurement): solves, Newton
 * rounds, PCG iterations launched for the whole batch, PCG iterations summed over
 * the RVEs, host synchronisations. */
cia_status cia_rve_batch_stats(const cia_rve_batch* batch, int64_t* counters, int32_t n);

/* ------------------------------------------------------------------------ */
/* Explicit material dynamics on a structured Hex8 box
 * (solve_material_explicit, src/fem/dynamics/material.py:89-297). */

typedef struct {
  int32_t nx, ny, nz;
  int32_t n_phases;
  double lengths[3];
  const cia_material* phases;     /* host [n_phases] */
  const int32_t* element_phase;   /* host [nx*ny*nz], NULL = phase 0 */
  int32_t precision;              /* CIA_PREC_*: element/material work and state type */
  int32_t pad;
} cia_box_desc;

typedef struct cia_box cia_box;

cia_status cia_box_create(const cia_box_desc* desc, cia_box** out);
cia_status cia_box_destroy(cia_box* box);

/* Energy-history columns written per step by cia_box_explicit_steps. */
enum {
  CIA_H_KINETIC = 0,        /* 1/2 sum m v^2 (material.py:851-856) */
  CIA_H_FREE_ENERGY = 1,    /* sum psi w detJ */
  CIA_H_DISSIPATION_RATE = 2, /* sum D w detJ (times dt: dissipated increment) */
  CIA_H_EXT_DU = 3,         /* f_unit . (u_{n+1} - u_n) (external work = mean load scale times this) */
  CIA_H_POSITIVE_WORK = 4,
  CIA_H_NEGATIVE_WORK = 5,
  CIA_H_UNLOADING_VOLUME = 6,
  CIA_H_ACTIVE_VOLUME = 7,
  CIA_H_MARGIN_MIN = 8,
  CIA_H_MARGIN_MAX = 9,
  CIA_N_HISTORY = 10
};

/*
 * Velocity-Verlet steps of solve_material_explicit (material.py:164-284), all on
 * the device.  With initialize = 1 the initial assembly at u (zero rate) sets
 * a = (f_ext(0) - f_int)/m and commits the state first (material.py:124-148).
 *   load_scale   host [n_steps + 1]: f_ext(t_k) = load_scale[k] * f_unit
 *   u, v, a      device [3 n_nodes] (node-major dofs), updated in place
 *   state        device [8][8 n_elements] (element type by precision), in place
 *   mass         device [3 n_nodes] lumped mass
 *   fixed        device uint8 [3 n_nodes]: homogeneous Dirichlet dofs
 *   f_unit       device [3 n_nodes]
 *   stress       device [6][8 n_elements] of the final state, or NULL
 *   internal_force device [3 n_nodes] of the final state, or NULL
 *   history      host [(initialize + n_steps)][CIA_N_HISTORY]
 */
cia_status cia_box_explicit_steps(cia_box* box, void* stream, int32_t initialize,
                                  double time_step, int32_t n_steps, const double* load_scale,
                                  const cia_update_settings* settings, double* u, double* v,
                                  double* a, void* state, const double* mass,
                                  const uint8_t* fixed, const double* f_unit, double* stress,
                                  double* internal_force, double* history,
                                  cia_point_failure* failure);

/*
 * Generic-material explicit step (solve_material_explicit with any
 * MaterialModel, material.py:164-284; the assembly pieces of
 * assemble_bundle_material_response, nonlinear_solid.py:222-300).  The caller
 * evaluates the material between the two calls.  FP64 throughout.
 *   cia_box_generic_strains: strain (and rate, may be NULL) [6][8 n_elements] at
 *     the Verlet predictor u + dt v + dt^2/2 a with fixed dofs zeroed (predict = 1)
 *     or at u itself with zero rate (predict = 0, the initial assembly).
 *   cia_box_generic_correct: internal force of the caller's stresses, then the
 *     node pass of cia_box_explicit_steps (a, v, u updated in place; with
 *     initialize = 1 only a = (f_ext - f_int)/m).  psi, diss, work (incremental
 *     work density), active (viscoplastic-active flag) and margin (yield
 *     activation margin) are [8 n_elements] device arrays or NULL; history_row is
 *     a device row of CIA_N_HISTORY doubles receiving the step's energy terms.
 */
cia_status cia_box_generic_strains(cia_box* box, void* stream, int32_t predict, double time_step,
                                   const double* u, const double* v, const double* a,
                                   const uint8_t* fixed, double* strain, double* rate);
cia_status cia_box_generic_correct(cia_box* box, void* stream, int32_t initialize,
                                   double time_step, double load_scale, const double* stress,
                                   const double* psi, const double* diss, const double* work,
                                   const double* active, const double* margin, double* u,
                                   double* v, double* a, const double* mass,
                                   const uint8_t* fixed, const double* f_unit,
                                   double* internal_force, double* history_row);

/*
 * Implicit Newmark pieces (solve_material_implicit_newmark,
 * src/fem/dynamics/material.py:300-701), FP64, with the consistent Hex8 mass
 * rho sum_q N^T N w detJ of the bundle material's density (elasticity.py:109-126).
 *   cia_box_strains_between: strain at u_new, rate (u_new - u_old)/dt, [6][8 n_el].
 *   cia_box_newmark_residual (_evaluate_implicit_trial, material.py:665-701):
 *     a = (u - u_pred)/(beta dt^2), v = v_pred + gamma dt a, f_int from the
 *     caller's stresses, R = M a + f_int - load_scale f_unit with fixed dofs zero;
 *     norms (host [2]) = |R_free|, |u_free|.
 *   cia_box_mass_apply: y = M x (all dofs).
 *   cia_box_newmark_solve: Jacobi-preconditioned CG for
 *     (K_t + mass_coefficient M) x = rhs (M of the given density) on the free dofs (x = 0 on fixed dofs),
 *     K_t = sum_q B^T C_q B w detJ from tangent [36][8 n_el] (row-major 6x6 per
 *     point; NULL = mass only); stops at |r| <= max(rtol |rhs|, atol); returns
 *     CIA_ERR_NOT_CONVERGED on breakdown or iteration limit.  Replaces the
 *     SuperLU solve of the effective tangent (material.py:540-543).
 */
cia_status cia_box_strains_between(cia_box* box, void* stream, double time_step,
                                   const double* u_new, const double* u_old, double* strain,
                                   double* rate);
cia_status cia_box_newmark_residual(cia_box* box, void* stream, double density, double beta,
                                    double gamma, double time_step, const double* stress,
                                    double load_scale, const double* f_unit,
                                    const uint8_t* fixed, const double* u, const double* u_pred,
                                    const double* v_pred, double* a, double* v,
                                    double* internal_force, double* residual, double* norms);
cia_status cia_box_mass_apply(cia_box* box, void* stream, double density, const double* x,
                              double* y);
cia_status cia_box_newmark_solve(cia_box* box, void* stream, const double* tangent,
                                 double density, double mass_coefficient, const uint8_t* fixed,
                                 const double* rhs, double* x, int32_t max_iterations,
                                 int32_t check_interval, double relative_tolerance,
                                 double absolute_tolerance, int32_t* iterations,
                                 double* residual_norm);

/* ------------------------------------------------------------------------ */
/* Recurrent MLP surrogate (the material-point network role of
 * EnergyDissipationRVERNO, src/concrete_impact/nn/models/rve_rno.py:162-476, and
 * its BPTT training, nn/training/rve_rno.py:113-415) on BF16 matrix cores with
 * FP32 accumulation.  Matrices are row-major; bf16 buffers are passed as void*.
 *   cia_mlp_linear_forward:  Y = act(X W^T + b), X [M][K] bf16, W [N][K] bf16,
 *     b [N] fp32 or NULL, activation 0 = identity, 1 = GELU in its tanh form,
 *     0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))); Z [M][N] fp32 receives the
 *     pre-activation when activation = 1 (or NULL); Y [M][N] fp32 and/or Y_bf16 [M][N].
 *   cia_mlp_linear_backward_data: dX = dZ W, dZ [M][N] bf16, W [N][K] bf16; with
 *     Zprev [M][K] fp32 the result is multiplied by GELU'(Zprev) (the gradient
 *     with respect to the previous layer's pre-activation).
 *   cia_mlp_linear_backward_weight: dW [N][K] = dZ^T X and db [N] = column sums
 *     of dZ (db may be NULL), reduced over the M rows through fp32 split partials
 *     in a workspace of cia_mlp_backward_weight_workspace(N, K) floats; with
 *     accumulate = 1 the results are added to dW, db.
 *   cia_mlp_adam_step: Adam with bias correction (step >= 1) and decoupled weight
 *     decay folded into the gradient; refreshes the bf16 copy (may be NULL).
 *   cia_mlp_forward_fused: the whole 13 -> 256^4 -> 12 network for M points in one
 *     kernel (activations stay in LDS); X [M][13] fp32, Y [M][12] fp32; weights as
 *     one bf16 blob W1 [256][13] | W2..W4 [256][256] | W5 [12][256] and biases as
 *     one fp32 blob b1..b4 [256] | b5 [12].
 */
cia_status cia_mlp_linear_forward(void* stream, int64_t M, int64_t N, int64_t K, const void* X,
                                  const void* W, const float* bias, int32_t activation, float* Z,
                                  float* Y, void* Y_bf16);
cia_status cia_mlp_linear_backward_data(void* stream, int64_t M, int64_t N, int64_t K,
                                        const void* dZ, const void* W, const float* Zprev,
                                        float* dX, void* dX_bf16);
cia_status cia_mlp_linear_backward_weight(void* stream, int64_t M, int64_t N, int64_t K,
                                          const void* dZ, const void* X, float* dW, float* db,
                                          float* workspace, int64_t workspace_floats,
                                          int32_t accumulate);
int64_t cia_mlp_backward_weight_workspace(int64_t N, int64_t K);
/* Allocation-free variant of cia_mlp_linear_backward_weight (no counterpart in
 * the reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight training or serving loops, harnesses that forbid
 * hipMalloc).  The split partials live in a static device array of the library
 * (17.1 MB per GPU, held until the process ends); requires N * (K + 1) <=
 * 2138240 floats and rejects larger shapes with CIA_ERR_INVALID_ARGUMENT.  The
 * stream must belong to the current device.  It calls no hipMalloc / hipFree.
 * Every call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams or host threads) race and must be serialised by the
 * caller.  Results equal cia_mlp_linear_backward_weight given a workspace of
 * cia_mlp_backward_weight_workspace(N, K) floats, bit for bit, for N, K <= 256. */
cia_status cia_mlp_linear_backward_weight_noalloc(void* stream, int64_t M, int64_t N, int64_t K,
                                                  const void* dZ, const void* X, float* dW,
                                                  float* db, int32_t accumulate);
cia_status cia_mlp_adam_step(void* stream, int64_t n, float* params, const float* grad,
                             float* m1, float* m2, void* params_bf16, float lr, float beta1,
                             float beta2, float eps, int32_t step, float weight_decay);
cia_status cia_f32_to_bf16(void* stream, int64_t n, const float* x, void* y);
cia_status cia_mlp_forward_fused(void* stream, int64_t M, const float* X, const void* weights,
                                 const float* biases, float* Y);

/* Fused BPTT training step of the surrogate (the batch step of train_rve_rno,
 * src/concrete_impact/nn/training/rve_rno.py:200-415, and its path-normalised loss,
 * :540-638, for this build's network): loss and gradients of a batch of `paths`
 * rows over the steps window_start .. window_start + window - 1, in one call of six
 * kernels.  The recurrence starts from state_in [paths][12] (latent 6, inelastic
 * strain 6 per row; NULL = zeros, the state at step 0) and, when state_out is not
 * NULL, leaves the state after the window there, so that consecutive windows of a
 * path are trained by truncated BPTT with the state carried between them.
 * The training set holds n_paths paths of path_length steps; batch row b uses path
 * p = path_index[b] (int32, 0 <= p < n_paths; a row with an index outside that range
 * contributes nothing), which supplies at step t:
 *   features [P][path_length][8] fp32: d_eps / eps_scale (6), rate feature, unused;
 *   strain   [P][path_length][6] fp32: cumulative strain / eps_scale;
 *   target   [P][path_length][6] fp32: stress / stress_scale;
 *   path_weight [P] fp32.
 * Recurrence (the rollout of the surrogate material, t from window_start):
 * x_t = [features_t[0:7], tanh(z_t)],
 * o_t = MLP(x_t), ein_t = ein_{t-1} + inelastic_step o_t[0:6], z_{t+1} = z_t + o_t[6:12],
 * sigma_t = Cs (strain_t - ein_t), Cs = stiffness (host, 36 floats row-major, the
 * stiffness in scaled units); loss = sum_b w_p sum_{t,i} (sigma_t - target_t)_i^2 /
 * (paths window 6).  weights / biases: the fp32 master blobs (layout of
 * cia_mlp_forward_fused); grad_weights / grad_biases receive the gradients (same
 * layouts, overwritten) and loss (device, 1 float) the loss.  grad_norm (device, 1
 * float, or NULL) receives the global norm of the gradients as computed; with
 * max_grad_norm > 0 the gradients are then scaled by min(1, max_grad_norm / norm), and
 * set to zero when the norm is not finite (no reference counterpart: the reference
 * rejects such a step, nn/training/rve_rno.py:1040-1067).  Network inputs and
 * activations are bf16, products accumulate in fp32 (as the layer-wise entry
 * points).  workspace: at least cia_surrogate_bptt_workspace(paths, window) bytes of
 * device memory. */
int64_t cia_surrogate_bptt_workspace(int64_t paths, int64_t window);
cia_status cia_surrogate_bptt(void* stream, int64_t n_paths, int64_t path_length, int64_t paths,
                              int64_t window_start, int64_t window, const int32_t* path_index,
                              const float* features, const float* strain, const float* target,
                              const float* path_weight, const float* stiffness,
                              float inelastic_step, const float* weights,
                              const float* biases, float* grad_weights, float* grad_biases,
                              float* loss, float* grad_norm, float max_grad_norm,
                              const float* state_in, float* state_out, void* workspace,
                              int64_t workspace_bytes);

/* Surrogate material-point update: the material protocol of
 * RVERNOMaterialAdapter.update (src/concrete_impact/nn/deployment.py:97-165) for this
 * build's network, one kernel per call (features, the fused network, stress and
 * state update, gate).  The network predicts inelastic strain increments; the stress
 * is the homogenised elastic stiffness times the elastic strain.  Every point keeps
 * an anchor (strain, inelastic strain, latent, age = time since the anchor); the
 * network is evaluated once from the anchor,
 *   d = strain - anchor_strain, tau = age + time_step, rate = |d|_2 / tau,
 *   x = [d / eps_scale, max(log10(rate / rate_ref) / rate_scale, rate_lo), tanh(latent)],
 *   ein' = anchor_inelastic + inelastic_step eps_scale o[0:6],
 *   stress = stiffness (strain - ein'),  z' = latent + o[6:12],
 * and the anchor moves to (strain, ein', z', 0) only when |d|_inf >= commit_strain
 * (or tau >= commit_time when commit_time > 0); otherwise it is kept with age tau.
 * A point trips when |d|_inf / eps_scale > gate_margin d_eps_max, the unclamped rate
 * feature > rate_hi + rate_margin, |z'|_inf > gate_margin latent_max, or a result is
 * not finite: next_fallback = fallback | trip, and trip_counter (device, never
 * reset here) is incremented by the number of points that trip with fallback = 0.
 * Layouts: strain, stress, anchor strain / inelastic strain [6][n] fp64, latents
 * [6][n] fp32, age [n] fp32, fallback [n] uint8; stiffness 6 x 6 row-major
 * (engineering Voigt, Pa); weights / biases as for cia_mlp_forward_fused.
 * Stresses of fallback points are the network's; the caller replaces them.
 * Optional energies (any of the last four pointers may be NULL): next_inelastic
 * [6][n] receives this call's inelastic strain ein'; free_energy [n] the stored
 * elastic energy 1/2 (strain - ein') . stress; dissipation [n] the inelastic power
 * stress . (ein' - inelastic) / time_step, with inelastic [6][n] the previous
 * call's ein' (zeros at the start); dissipation requires inelastic. */
typedef struct cia_surrogate_gate {
  double eps_scale, rate_ref, rate_scale, inelastic_step;
  double d_eps_max, rate_lo, rate_hi, latent_max, gate_margin, rate_margin;
  double commit_strain, commit_time;
  double stiffness[36];
} cia_surrogate_gate;
cia_status cia_surrogate_update(void* stream, int64_t n, const double* strain, double time_step,
                                const void* weights, const float* biases,
                                const cia_surrogate_gate* gate, const double* anchor_strain,
                                const double* anchor_inelastic, const float* anchor_latent,
                                const float* anchor_age, const uint8_t* fallback, double* stress,
                                double* next_anchor_strain, double* next_anchor_inelastic,
                                float* next_latent, float* next_age, uint8_t* next_fallback,
                                uint32_t* trip_counter, const double* inelastic,
                                double* next_inelastic, double* free_energy, double* dissipation);

/* Kernel timing with HIP events on the launch stream (no reference
 * counterpart; measurement plumbing).  Slots: 0 assemble, 1 PCG element
 * pass, 2 PCG node gather, 3 PCG vector update, 4 preconditioner (spectral: FFT,
 * symbol solve, inverse FFT; block Jacobi: block inversion), 5 residual gather.  cia_profile_read returns and clears the accumulated totals. */
cia_status cia_profile_enable(int32_t on);
cia_status cia_profile_read(double* total_ms, int64_t* counts, int32_t n);

/* Energy-based RVE recurrent neural operator (the reference's trained surrogate
 * family): one explicit update of n material points,
 * EnergyDissipationRVERNO.update_batch / _constitutive_tensors
 * (src/concrete_impact/nn/models/rve_rno.py:249-418), FP32 like the reference's
 * artifacts, with the algorithmic tangent of rve_rno.py:290-302 when tangent is not
 * NULL.  Architecture: 4 SiLU hidden layers of width 128 (the artifact's), latent
 * dimension m = 1..8, evolution CIA_RNO_DIRECT_RATE or CIA_RNO_MOBILITY_GRADIENT,
 * integrator CIA_RNO_CS_SEMI_IMPLICIT ("cs_semi_implicit_forward_euler") or
 * CIA_RNO_CURRENT_STATE ("current_state_stress_forward_euler").
 * packed: device FP32 weights, cia_rno_packed_size floats, three networks in the
 * order equilibrium_energy, nonequilibrium_energy, state_evolution, each laid out as
 *   w1 [128][16] (input columns beyond the input width zero), b1 [128],
 *   3 x (w [128][128], b [128]) for hidden layers 2..4, their 3 transposes
 *   [128][128], w1 transposed [16][128], wout [R][128] and bout [R],
 *   R = 16 ceil(n_out / 16), rows beyond n_out and the energies' bout zero
 * (nn.Linear weights are [out][in]).  anchors: device [14] from cia_rno_anchors
 * (prepare_energy_anchors, rve_rno.py:224-247).  Per point, row-major: strain [n][6]
 * (engineering shear, Voigt xx yy zz yz xz xy), latent [n][m], time_step [n];
 * outputs stress [n][6], latent_out [n][m] (may alias latent), state_rate and
 * dimensionless_state_rate [n][m], dissipation, free_energy,
 * thermodynamic_violation, nonequilibrium_energy_increment [n], tangent [n][6][6]
 * (d stress_i / d strain_j at [i][j]) or NULL.  strain_input_scale: host [6]. */
enum { CIA_RNO_DIRECT_RATE = 0, CIA_RNO_MOBILITY_GRADIENT = 1 };
enum { CIA_RNO_CS_SEMI_IMPLICIT = 0, CIA_RNO_CURRENT_STATE = 1 };
cia_status cia_rno_packed_size(int32_t latent_dimension, int32_t evolution, int64_t* floats);
cia_status cia_rno_anchors(void* stream, const float* packed, int32_t latent_dimension,
                           int32_t evolution, float* anchors);
cia_status cia_rno_update(void* stream, const float* packed, int32_t latent_dimension,
                          int32_t evolution, int32_t integrator, double reference_time_scale,
                          const double* strain_input_scale, const float* anchors, int64_t n,
                          const float* strain, const float* latent, const float* time_step,
                          float* stress, float* latent_out, float* state_rate,
                          float* dimensionless_state_rate, float* dissipation, float* free_energy,
                          float* thermodynamic_violation, float* nonequilibrium_energy_increment,
                          float* tangent);

#ifdef __cplusplus
}
#endif

#endif /* CIA_H */
