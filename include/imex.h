/*
 * imex.h — C ABI of the MI355X-native IMEX-MHD per-step hot path.
 *
 * The reference solver (ZefengChen-Kairos/IMEX) exposes its hot path as free
 * C++ functions on AoS `Grid`/`Array` containers driven by mutable globals
 * (include/declarations.hpp, include/config.hpp).  This header replaces that
 * surface with an opaque context that owns every device buffer, scalar and
 * the HIP stream.  Host arrays crossing the boundary use the reference's own
 * layout so a caller can hand over its `Grid::data` / `Array::data` unchanged:
 *
 *   Grid  (include/types.hpp:37-63): double[(ny+6)*(nx+6)*8],
 *         element (i,j,k) at 8*(i + j*(nx+6)) + k, 3 ghost layers.
 *   Array (include/types.hpp:65-91): double[(ny+6)*(nx+6)], (i,j) at i + j*(nx+6).
 *   Krylov vectors (M_P / M_B, src/implicit_operators.cpp:5-139):
 *         compact, (i-3) + nx*(j-3) + c*nx*ny.
 *
 * Host pointers are borrowed for the duration of the call only.  Device
 * buffers stay resident between calls; the stepping entry points never
 * allocate on the device after imex_create.
 *
 * Return codes: 0 = ok; > 0 = warning bitmask (IMEX_WARN_*); < 0 = error
 * (IMEX_ERR_*).  One context per HIP stream; contexts share no state.
 */
#ifndef IMEX_H
#define IMEX_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations (values follow the reference enums in include/config.hpp:27-41) ---- */
enum { IMEX_BC_TRANSMISSIVE = 0, IMEX_BC_REFLECTIVE = 1,           /* x/y_boundary_condition (config.hpp:36-37) */
       IMEX_BC_PERIODIC = 2, IMEX_BC_DIRICHLET = 3,
       IMEX_BC_JET_INFLOW = 4 };                                     /* JET_INFLOW: y only (y_JetInflow) */
enum { IMEX_RECON_NONE = 0, IMEX_RECON_MINMOD = 1, IMEX_RECON_VANLEER = 2, IMEX_RECON_ENO3 = 3,
       IMEX_RECON_WENO5 = 4 };                                        /* Reconstruction (config.hpp:30) */
enum { IMEX_RECONSTATE_PRIM = 0, IMEX_RECONSTATE_CONS = 1 };         /* ReconState (config.hpp:41) */
enum { IMEX_EOS_IDEAL = 0, IMEX_EOS_RK = 1 };                       /* EOSType (config.hpp:118) */
enum { IMEX_CT_CONSISTENT = 0, IMEX_CT_PLAIN = 1 };                 /* Staggered_CT_Consistent / Staggered_CT */
enum { IMEX_PREC_FP64 = 0, IMEX_PREC_MIXED = 1 };
enum { IMEX_FORMULA_ENERGY = 0, IMEX_FORMULA_PRESSURE = 1 };       /* Formula (config.hpp:27) */
enum { IMEX_MAGFIELD_STAGGER = 0, IMEX_MAGFIELD_UNSTAGGER = 1, IMEX_MAGFIELD_POTENTIAL_A = 2,
       IMEX_MAGFIELD_NODIVFREE = 3 };                              /* MagField (config.hpp:28) */
enum { IMEX_OP_MAG = 0, IMEX_OP_P_CARRIER = 1, IMEX_OP_P_JACOBIAN = 2, IMEX_OP_ENERGY = 3, IMEX_OP_POTENTIAL = 4 };
enum { IMEX_CASE_VORTEX = 0, IMEX_CASE_ORSZAG_TANG = 1, IMEX_CASE_BLAST = 2, IMEX_CASE_JET = 3 };
/* How a step is scheduled.  All four give identical results.
 *   GRAPH:  every solve gets an iteration budget learned from earlier steps; the whole step is
 *           launched as one captured hipGraph with no host synchronisation inside it; a step whose
 *           solve outran its budget is discarded and rerun with HOST scheduling.
 *   BUDGET: as GRAPH, launched kernel by kernel (no graph).
 *   HOST:   the host polls GMRES/Newton convergence between launches, as the reference's loops do.
 *   AUTO:   (default) a single instance times a few GRAPH and HOST steps now and then and runs the
 *           faster of the two; a batch runs GRAPH. */
enum { IMEX_SCHED_GRAPH = 0, IMEX_SCHED_BUDGET = 1, IMEX_SCHED_HOST = 2, IMEX_SCHED_AUTO = 3 };

enum {
  IMEX_WARN_GMRES_MAXOUTER = 1,   /* src/implicit_operators.cpp:918 */
  IMEX_WARN_NEWTON_MAXITER = 2,   /* src/implicit_operators.cpp:720-721 */
  IMEX_WARN_POSITIVITY     = 4,   /* src/implicit_operators.cpp:734-736 */
  IMEX_WARN_EOS_MAXITER    = 8    /* RK temperature Newton hit its cap */
};
enum {
  IMEX_ERR_ARG = -1, IMEX_ERR_HIP = -2, IMEX_ERR_NONFINITE = -3,
  IMEX_ERR_UNSUPPORTED = -4, IMEX_ERR_NOMEM = -5
};

/* Run-time replacement of the compile-time globals of include/config.hpp:43-125. */
typedef struct imex_params {
  int nx, ny;                     /* interior cells (config.hpp:52-53) */
  double x0, x1, y0, y1;          /* domain (main.cpp:9-224) */
  double gamma;                   /* gamma_plasma (config.hpp:122) */
  int bc_x, bc_y;                 /* IMEX_BC_* */
  int recon;                      /* IMEX_RECON_* (config.hpp:60): NoRecon, Minmod, VanLeer, ENO3 or WENO5
                                     (reconstruction.cpp:40-266); the mixed form reconstructs in FP32,
                                     choosing ENO3's stencil in FP64 */
  int minmod_central_y;           /* 1 = the vortex test's central y-slope (reconstruction.cpp:70-72) */
  int eos;                        /* IMEX_EOS_* */
  double rk_alfa, rk_b, rk_R, rk_cv;  /* Redlich-Kwong constants (config.hpp:125) */
  double cfl;                     /* CFL (config.hpp:57) */
  int floor_on;                   /* config.hpp:100 */
  double p_floor, theta, mbig_fac;/* config.hpp:101-103 */
  int newton_max;                 /* Pnewton_max (config.hpp:106) */
  double newton_tol;              /* Pnewton_tol (config.hpp:107) */
  double gmres_tol;               /* GMREStol (config.hpp:71) */
  int gmres_restart;              /* JMAX (implicit_operators.cpp:801) */
  int gmres_max_outer;            /* outmax; <= 0 means dim (implicit_operators.cpp:802) */
  double shock_kappa1;            /* config.hpp:77 */
  double time_stop;               /* dt = min(dt, time_stop - t) (main.cpp:341) */
  int ct_mode;                    /* IMEX_CT_* */
  int precision;                  /* IMEX_PREC_* */
  double mixed_inner_tol;         /* mixed form: each FP32 inner solve reduces its FP64 residual by this factor */
  double mixed_outer_tol;         /* mixed form: refinement stops when the FP64 residual is below this
                                     factor of the (preconditioned) right-hand side norm, plus gmres_tol */
  int schedule;                   /* IMEX_SCHED_* */
  int budget_margin;              /* Arnoldi steps granted per GMRES solve beyond the largest count of the
                                     last 8 steps (budgeted schedules; default 1) */
  int mixed_pressure;             /* mixed form: 1 = the pressure solves also run FP32 inner solves; 0 (default)
                                     = they stay FP64 (the pressure system converges in a few FP64 steps, fewer
                                     than the refinement needs) while the magnetic solves and fluxes run FP32 */
  int formula;                    /* IMEX_FORMULA_* (config.hpp:44, default PRESSURE): which explicit flux and
                                     which implicit subsystem a step runs after the magnetic subsystem
                                     (main.cpp:470-486); in the mixed form the energy solve follows
                                     mixed_pressure */
  int recon_state;                /* IMEX_RECONSTATE_* (config.hpp:61; default CONS): reconstruct the conservative
                                     state, or the primitive state with the face states converted back
                                     (reconstruction.cpp:11-24, 272-281).  PRIM keeps the explicit fluxes
                                     in FP64 in the mixed form */
  int magfield;                   /* IMEX_MAGFIELD_* (config.hpp:45; default STAGGER): the magnetic subsystem and
                                     divergence treatment of a stage (main.cpp:424-434, 497-510).  UNSTAGGER and
                                     POTENTIAL_A evolve the vector potential A, which the caller sets with
                                     imex_set_potential before stepping; with STAGGER and NODIVFREE A is not used */
  double jet_rho, jet_v, jet_B0, jet_p;  /* bc_y = JET_INFLOW (config.hpp:253-273): the inflow state (rho, 0, v, 0,
                                     p, 0, B0, 0) held in the bottom ghost cells of the strip |x_c| < jet_halfwidth;
                                     outflow (transmissive) elsewhere.  Defaults: config.hpp:257-260 */
  double jet_halfwidth;           /* 0.05 (config.hpp:261-269) */
  double rho_amb, p_amb;          /* bc = DIRICHLET (config.hpp:247): the ambient density and pressure its ghosts
                                     hold (velocity 0, B copied from the boundary cell); the pressure solve pins
                                     its boundary cells to p_amb.  Defaults 1, 0.1 */
  int mixed_flux;                 /* mixed form: 1 = the explicit fluxes (TVDReconstruction + RusanovC_Update /
                                     Pressure_flux_Update) also run in FP32 (pressure formula, conservative
                                     reconstruction only); 0 (default) = they stay FP64.  FP32 flux rounding
                                     (1e-7 relative, cell to cell) enters the right-hand sides of the magnetic and
                                     pressure solves, whose tolerances then cost more Arnoldi steps than the FP32
                                     fluxes save */
  int adjoint_recompute;          /* imex_adjoint_forward / _backward: 0 (default) = keep every step's tape in
                                     device memory when the tapes fit in half of the free memory (the backward pass
                                     then recomputes nothing), else keep checkpoints only; 1 = checkpoints only,
                                     the backward pass re-runs each segment of `every` steps (least memory) */
} imex_params;

/* Step-global scalars: the mutable globals read by the operators (config.hpp:83-94). */
typedef struct imex_scalars {
  double t, dt, adt;
  double a_max, B_max_x, B_max_y;
  double vx_max, vy_max, v_max;
  double ch, ch_x, ch_y;
  double prev_dt;                 /* the function-static of caldt (timestep.cpp:58) */
  int timestep;
  int rks;                        /* current IMEX stage */
} imex_scalars;

/* Per-stage working set, host side, in the reference layout (Grid / Array). */
typedef struct imex_stage_host {
  double *u_con, *Qe, *Qs;        /* Grid */
  double *bx, *by;                /* Bx_p0, By_0p (Array) */
  double *bx_qs, *by_qs;          /* Bx_p0_Qs, By_0p_Qs (Array) */
  double *fx, *fy;                /* gFluxBy_p0, gFluxBx_0p (Array) */
} imex_stage_host;

typedef struct imex_stats {
  long long steps;
  long long mag_solves, mag_gmres_iters, mag_gmres_restarts;
  long long p_solves, p_gmres_iters, p_gmres_restarts;
  long long newton_iters;
  long long op_applies_mag, op_applies_p;
  long long mixed_refinements;
  int warnings;                   /* OR of IMEX_WARN_* */
  int max_floored;                /* g_max_floored (config.hpp:109) */
  double min_p;                   /* g_min_p_ever (config.hpp:110) */
  double div_b;                   /* cal_divergence_error after the last step */
  long long budget_steps;         /* steps that ran on learned iteration budgets */
  long long budget_reruns;        /* of those, steps rerun because a solve outran its budget */
  long long graph_builds;         /* step graphs captured and instantiated */
} imex_stats;

typedef struct imex_solve_info {
  int iters;                      /* Arnoldi steps (GMRES) or Newton iterations */
  int restarts;
  double res0, res;               /* first and last residual norm */
  int status;
} imex_solve_info;

/* Per-kernel-class timing gathered with HIP events on the context stream. */
typedef struct imex_kernel_profile {
  const char* name;
  long long launches;
  double total_ms;
  double bytes;                   /* algorithmic bytes (distinct per-cell fields read + written) */
} imex_kernel_profile;

typedef struct imex_ctx imex_ctx;

/* ---- parameters and initial data ---- */
int imex_default_params(imex_params* p);
/* Primitive initial data + staggered faces for one of the IMEX_CASE_* problems.
 * Replaces u_initialization (src/initial_conditions.cpp:5-732; vortex :641-669,
 * 698-717; Orszag-Tang :426-435; blast :436-449; faces :718-725).  `amp`,
 * `u0`, `v0` scale the vortex (amplitude, background velocity); `smooth` selects
 * the erf-smoothed blast profile.  Outputs are host arrays (Grid / Array). */
int imex_init_case(const imex_params* p, int case_id, double amp, double u0, double v0,
                   int smooth, double* u_pri, double* bx, double* by);

/* ---- context ---- */
int imex_create(const imex_params* p, imex_ctx** out);
int imex_destroy(imex_ctx* ctx);
const char* imex_last_error(void);
/* Device-resident footprint of the context in bytes. */
long long imex_device_bytes(const imex_ctx* ctx);

/* ---- whole-state exchange (main.cpp:231-243, 561) ---- */
int imex_set_state(imex_ctx* ctx, const double* u_pri, const double* bx, const double* by,
                   double t, int timestep, double prev_dt);
int imex_get_state(imex_ctx* ctx, double* u_con, double* u_pri, double* bx, double* by,
                   imex_scalars* sc);

/* ---- the step loop body (main.cpp:334-595 without I/O and diagnostics) ---- */
int imex_step(imex_ctx* ctx, int nsteps, imex_stats* stats);
/* Same as imex_step but leaves the stream running; imex_sync waits for it. */
int imex_step_async(imex_ctx* ctx, int nsteps);
int imex_sync(imex_ctx* ctx);
int imex_get_stats(imex_ctx* ctx, imex_stats* stats);
int imex_reset_stats(imex_ctx* ctx);
double imex_div_b(imex_ctx* ctx);   /* cal_divergence_error (imex_subsystems.cpp:630-660): the face divergence with
                                      STAGGER, the cell-centred one (one-sided at the domain edges) otherwise; max over a batch */

/* ---- batches: independent instances of uneven grid size stepped together ----
 * The reference runs one grid per process (src/main.cpp:231-250, one Grid built from
 * config.hpp's nx, ny); a batch holds n such runs, sharing every parameter of `p` except
 * the grid size (nx[i], ny[i]), each with its own state, time step and solver history.
 * Every launch of a step covers all instances; the step runs on iteration budgets (the
 * largest any instance needs) and is replayed as a captured graph.  imex_create is a
 * batch of one.  Stage- and operator-level calls take a single-instance context.
 * Stats sum iteration counts over instances (min_p is the minimum, max_floored the
 * maximum).  Instance indices run from 0 to n-1. */
int imex_batch_create(const imex_params* p, int n, const int* nx, const int* ny, imex_ctx** out);
int imex_batch_size(const imex_ctx* ctx);
int imex_batch_set_state(imex_ctx* ctx, int i, const double* u_pri, const double* bx, const double* by,
                         double t, int timestep, double prev_dt);
int imex_batch_get_state(imex_ctx* ctx, int i, double* u_con, double* u_pri, double* bx, double* by,
                         imex_scalars* sc);
double imex_batch_div_b(imex_ctx* ctx, int i);

/* ---- autotuner (no counterpart in the reference, whose loops have no launch shape) ----
 * Times candidate launch shapes on the context's current state (set it first) and keeps the
 * fastest: IMEX_TUNE_EXPLICIT picks the explicit-flux tile among 16x16, 32x8 and 64x4 cells
 * (results identical bit for bit); IMEX_TUNE_KRYLOV picks the grid cap of the magnetic Krylov
 * kernels among 1024, 2048 and 4096 blocks per instance and that of the pressure-operator kernels
 * among 1024, 1536, 2048 and 4096 (a different cap reorders the dot-product sums, so results
 * change by rounding).  The state, time and statistics are left as they were. */
enum { IMEX_TUNE_EXPLICIT = 1, IMEX_TUNE_KRYLOV = 2 };
typedef struct imex_tune_info {
  int explicit_tile_x, explicit_tile_y;   /* chosen explicit tile (cells) */
  int krylov_blocks;                      /* chosen Krylov grid cap (blocks per instance) */
  double explicit_us[5];                  /* both explicit kernels, per candidate 16x16, 32x8, 64x4, 32x16, 64x8 */
  double krylov_us[3];                    /* one Arnoldi step (apply + update), per cap 1024, 2048, 4096 */
  int krylov_p_blocks;                    /* chosen grid cap of the pressure-operator kernels */
  double krylov_p_us[4];                  /* pressure Arnoldi apply, per cap 1024, 1536, 2048, 4096 */
} imex_tune_info;
int imex_autotune(imex_ctx* ctx, int flags, imex_tune_info* out);

/* ---- vector potential (UNSTAGGER and POTENTIAL_A) ----
 * A is a Grid in the reference layout with A_x, A_y, A_z in slots 0..2 (slots 3..7 are ignored on input
 * and zero on output), as the reference's `Grid A` (main.cpp:230). */
/* A_initial (constrained_transport.cpp:224-398) for the built-in cases of imex_init_case (vortex
 * A_z = amp/(2 pi) exp((1 - r^2)/2), Orszag-Tang A_z = cos y + cos(2x)/2, blast A_z = b0 (y - x)), at cell
 * centres, then bound_cond_A (boundary.cpp:177-292). */
int imex_init_potential(const imex_params* p, int case_id, double amp, double* A);
/* The step-start vector potential (main.cpp:244, 347-348); ghosts are recomputed by bound_cond_A. */
int imex_set_potential(imex_ctx* ctx, const double* A);
int imex_get_potential(imex_ctx* ctx, double* A);
int imex_batch_set_potential(imex_ctx* ctx, int i, const double* A);
int imex_batch_get_potential(imex_ctx* ctx, int i, double* A);

/* ---- stage-level drop-ins, one per reference entry point ---- */
int imex_set_stage(imex_ctx* ctx, const imex_stage_host* st, const imex_scalars* sc);
int imex_get_stage(imex_ctx* ctx, imex_stage_host* st, imex_scalars* sc);
/* calch + caldt + MagneticEigenvalues on the current u_pri (timestep.cpp:5-67, 89-120, 254-273). */
int imex_cfl(imex_ctx* ctx, imex_scalars* sc);
/* Qe/Qs/faces_Qs from u_n and the stored stage increments (main.cpp:368-414). */
int imex_stage_combine(imex_ctx* ctx, int rks);
/* TVDReconstruction + RusanovC_Update + bound_cond_u (main.cpp:417-421). */
int imex_explicit_conv(imex_ctx* ctx);
/* MagSubsystem_UpdateB (imex_subsystems.cpp:537-628, incl. CollectMagFluxes). */
int imex_mag_update_b(imex_ctx* ctx);
/* B slots <- cell B~ then Pressure_flux_Update (main.cpp:448-477). */
int imex_explicit_press(imex_ctx* ctx);
/* PreSubsystem_Pressure (imex_subsystems.cpp:76-302, DefinitionofU). */
int imex_pre_subsystem_pressure(imex_ctx* ctx);
/* B slots <- cell B~ then Energy_flux_Update (main.cpp:448-474, flux.cpp:78-200). */
int imex_explicit_energy(imex_ctx* ctx);
/* PreSubsystem_Energy (imex_subsystems.cpp:415-534): implicit total-energy solve with the Jacobi
 * preconditioner precE, then the momentum update by (gamma - 1) grad E. */
int imex_pre_subsystem_energy(imex_ctx* ctx);
/* Staggered_CT_Consistent (ct_flux.cpp:150-191) or Staggered_CT (constrained_transport.cpp:66-143). */
int imex_ct_update(imex_ctx* ctx);
/* The stage's A and A_s (main.cpp:371-412: A_s is the implicit stage combination of A, and A starts the
 * stage as a copy of it), for the two stage-level calls below.  Grids as imex_set_potential. */
int imex_set_stage_potential(imex_ctx* ctx, const double* A, const double* A_s);
int imex_get_stage_potential(imex_ctx* ctx, double* A, double* A_s);
/* MagSubsystem_UpdateA (imex_subsystems.cpp:5-74): A from the potential system (GMRES on
 * Magnetic_Subsystem_potential_AX), then the cell B = curl A. */
int imex_mag_update_a(imex_ctx* ctx);
/* Unstaggered_CT_Consistent (constrained_transport.cpp:502-540): A_z from the consistent corner EMFs
 * of the magnetic subsystem's face fluxes, then B_x, B_y = curl A_z. */
int imex_unstaggered_ct(imex_ctx* ctx);

/* ---- operator level (tests, adjoint) on the current stage working set ---- */
/* Magnetic_Subsystem_AX / Pressure_Subsystem_AX / Energy_Subsystem_AX / Magnetic_Subsystem_potential_AX
 * (implicit_operators.cpp:141-353, 505-616, 741-790, 355-503).  For IMEX_OP_P_JACOBIAN, p_freeze (N) and
 * mbig select V_slope.  The energy operator's coefficients come from the stage's Qe and u_con, the
 * potential operator's from Qe, u_con and (its viscosity) Qs.  Host arrays. */
int imex_apply_op(imex_ctx* ctx, int op, const double* x, double* Ax,
                  const double* p_freeze, double mbig);
/* Transposed operators for the discrete adjoint: ATx = A^T x for the magnetic and pressure
 * operators imex_apply_op applies (periodic grids, ideal gas; the pressure Jacobian only where
 * p_freeze is above the floor).  Same arguments as imex_apply_op. */
int imex_apply_op_T(imex_ctx* ctx, int op, const double* x, double* ATx,
                    const double* p_freeze, double mbig);
/* GMRES_Solver (implicit_operators.cpp:793-919). precond: 0 none, 1 Jacobi (precP, or precE for
 * IMEX_OP_ENERGY, imex_subsystems.cpp:479-500). x in/out. */
int imex_gmres(imex_ctx* ctx, int op, const double* b, double* x, int precond,
               const double* p_freeze, double mbig, imex_solve_info* info);
/* Pressure_NestedNewton (implicit_operators.cpp:625-738). p in/out. */
int imex_pressure_newton(imex_ctx* ctx, const double* b, double* p, imex_solve_info* info);

/* ---- discrete adjoint of the step (no counterpart in the reference, which has no derivative) ----
 * Supported: single-instance contexts, periodic grids, ideal gas, Minmod or VanLeer of the conservative
 * state (recon_state CONS), EMF-consistent
 * CT (magfield STAGGER), pressure formula, FP64 form (other settings return IMEX_ERR_UNSUPPORTED).  Usage: imex_set_state (the initial
 * state x0), imex_adjoint_forward (runs nsteps, keeping a checkpoint every `every` steps; the
 * context then holds the final state), read the final state and form the cotangent of the
 * objective J with respect to it (gu_con: host Grid of conservative variables; gbx, gby: face
 * Arrays, may be null), imex_adjoint_backward: returns dJ/dx0 with respect to the primitive state
 * and the faces given to imex_set_state (interior cells; ghosts are zero).  The backward pass
 * recomputes each checkpoint segment once, taping its intermediate states, and solves the
 * transposed magnetic and pressure systems with GMRES.  The context is left at the final state.
 * Exact up to the solver tolerances where every step is differentiable: the CFL maxima are
 * differentiated through the cells that attain them, limiters through their active branch. */
typedef struct imex_adjoint_info {
  int steps, checkpoints, recomputed_steps;
  long long transposed_solves, mag_iters, p_iters;   /* GMRES iterations of the transposed solves */
  double worst_solve_residual;                       /* largest relative residual left by a transposed solve */
  double backward_ms;                                /* device time of the backward pass */
} imex_adjoint_info;
int imex_adjoint_forward(imex_ctx* ctx, int nsteps, int every);
int imex_adjoint_backward(imex_ctx* ctx, const double* gu_con, const double* gbx, const double* gby,
                          double* adj_u_pri, double* adj_bx, double* adj_by, imex_adjoint_info* info);

/* ---- output files and per-step diagnostics (src/output.cpp:127-326, src/main.cpp:562-720) ---- */
/* One step's diagnostics, as the reference prints or records them after each step. */
typedef struct imex_diag {
  int timestep;
  double t, dt;
  double div_error;               /* cal_divergence_error (main.cpp:562, 577) */
  double B_inf;                   /* max |central div B| of the cell field over the interior (main.cpp:580-590) */
  double dt_comp;                 /* imex_dt / fully_dt of the step (timestep.cpp:56-57, main.cpp:592) */
  double emin, emin_x, emin_y;    /* DIAG line (main.cpp:563-575): smallest internal energy density and its cell centre */
  double mass, intE, magE_max;    /* sum rho dx dy, sum e dx dy, max |B|^2/2 */
} imex_diag;
/* Record one imex_diag per instance after every committed step into a device ring of `capacity`
 * entries (0 disables); recording adds two small launches per step and no host synchronisation. */
int imex_diag_enable(imex_ctx* ctx, int capacity);
/* The recorded diagnostics of instance i, oldest first (at most max); returns the count. */
int imex_diag_read(imex_ctx* ctx, int i, imex_diag* out, int max);
/* The diagnostics of instance i's current state, now (dt_comp from the last step's scalars). */
int imex_diagnose(imex_ctx* ctx, int i, imex_diag* out);
/* Diagnose_CT_EMF (constrained_transport.cpp:181-222) on instance i's current state: CT_EMF_Stats
 * (:165-178) of the corner EMF's central, Rusanov, Ez~ and total parts, 4 rows of
 * {L2 of Ez, max |Ez|, L2 of the face injection dB/dt, its max}. */
int imex_diagnose_ct_emf(imex_ctx* ctx, int i, double out[4][4]);
/* printAllVariablesToVTK's file (output.cpp:127-319), with its layout as is (the field loops start at
 * padded index 0); from host Grids, or from instance i's current state. */
int imex_write_vtk_host(const imex_params* p, const double* u_con, const double* u_pri, double t, int nstep,
                        const char* file);
int imex_write_vtk(imex_ctx* ctx, int i, const char* file, int nstep);
/* The end-of-run field files RK_rho, RK_vx, RK_vy, RK_E, RK_p, RK_Bx, RK_By, RK_Bz, RK_v_mag (.dat), and
 * RK_T.dat with the Redlich-Kwong EOS, into directory `dir` (main.cpp:598-720). */
int imex_write_fields_host(const imex_params* p, const double* u_con, const double* u_pri, const char* dir);
int imex_write_fields(imex_ctx* ctx, int i, const char* dir);
/* A per-step history file ("step value" lines: RK_div_error.dat, RK_B_inf.dat, RK_dt_comp.dat). */
int imex_write_history(const char* file, int n, const int* steps, const double* values);
/* The reference's output directory for a run (output.cpp:286-288) under `root`; returns its length, or
 * an error if it does not fit in `cap` bytes with the terminating zero. */
int imex_output_path(const imex_params* p, const char* root, const char* testname, char* out, int cap);

/* ---- no allocation (no counterpart in the reference) ----
 * cal_divergence_error of device-resident staggered faces, for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops.  Same metric and summation order as
 * imex_div_b (StaggerCT branch, imex_subsystems.cpp:630-660) without a context.  bx, by: device arrays in
 * the Array layout ((ny+6)*(nx+6) doubles, face i+1/2 of row j at i + j*(nx+6)); out: device pointer,
 * receives sqrt(sum over interior cells of dx dy (div B)^2); p supplies nx, ny and the domain; stream: a
 * hipStream_t (null = the default stream).  Enqueues two kernels and returns without synchronising.
 * Allocates nothing: its partial sums live in a fixed 256-double device array of the library (one copy per
 * GPU, held until the process ends, its address resolved on the first call on each GPU, which should
 * therefore not itself be captured).  Calls that overlap on one GPU share that array and race.  Grids up
 * to 32768 cells per side; larger ones return IMEX_ERR_ARG. */
int imex_div_b_noalloc(const imex_params* p, const double* bx, const double* by, double* out, void* stream);

/* ---- profiling ---- */
/* Enable (1) or disable (0) per-launch event timing; enabling clears the totals. */
int imex_profile_enable(imex_ctx* ctx, int on);
/* Fills up to `max` entries, one per kernel class; returns the number written. */
int imex_profile_read(imex_ctx* ctx, imex_kernel_profile* out, int max);

#ifdef __cplusplus
}
#endif
#endif /* IMEX_H */
