// kernels.hpp — internal declarations shared by the kernel files and the host driver.
//
// Every kernel takes the device instance table (`const Inst*`) and a block map (`BMap`) and
// addresses its buffers by id (batch.hpp), so one launch covers every instance of a context.
#pragma once
#include <string>
#include "device_common.hpp"
#include "batch.hpp"

#define TX 64              // operator tile (cells) in x
#define TY 4               // operator tile (cells) in y

// Reduces nblk per-block pairs (a, b) at part[2k], part[2k+1] with every thread of the block (up to
// 1024): min of a, and the sum (SUM) or max of b, all exact in any order; thread 0 gets them in out.
// Many threads keep each thread's chain of loads short when the producer ran thousands of blocks.
template <bool SUM>
__device__ __forceinline__ void pair_reduce(const double* part, int nblk, double* out) {
  __shared__ double sa[16], sb[16];
  double a = 1e300, b = SUM ? 0.0 : -1e300;
  for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
    a = fmin(a, part[2 * k]);
    b = SUM ? b + part[2 * k + 1] : fmax(b, part[2 * k + 1]);
  }
  a = wave_min(a);
  b = SUM ? wave_sum(b) : wave_max(b);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < nw; q++) { a = fmin(a, sa[q]); b = SUM ? b + sb[q] : fmax(b, sb[q]); }
    out[0] = a; out[1] = b;
  }
}

// Positivity statistics of one Newton solution (implicit_operators.cpp:724-737), accumulated on the
// device: pst[0] = min p ever, pst[1] = max floored count, pst[2] = 1 if min p < theta p_floor - 1e-14.
// The nblk per-block (min p, floored count) pairs in partial2 come from k_pstats or from the last
// k_newton_residual of the solve.
__device__ __forceinline__ void pstats_reduce(const Inst& I, int nblk, double thr) {
  double mc[2];
  pair_reduce<true>(gmem(I.partial2), nblk, mc);
  if (threadIdx.x == 0) {
    double* acc = gmem(I.pst);
    acc[0] = fmin(acc[0], mc[0]);
    acc[1] = fmax(acc[1], mc[1]);
    if (mc[0] < thr) acc[2] = 1.0;
  }
}

// Device-resident GMRES / Newton control block.
struct KrylovDev {
  double H[(MAXM + 2) * (MAXM + 1)];   // H[i*(MAXM+1) + j]
  double c[MAXM + 2], s[MAXM + 2], gam[MAXM + 2];
  double tol, tol_eff, hcur, res0, res;
  double newton_res0, newton_res;
  double tol_outer;                    // mixed form: FP64 residual target of the refinement
  int outer, done, converged, jlast, steps, x_is_zero;
  int newton_it, newton_conv, newton_first;
  int steps_total, outer_conv;         // Arnoldi steps over all sweeps; refinement converged
  int sweep_steps[MAXSWEEP];           // Arnoldi steps of each refinement sweep
  unsigned cnt_a, cnt_c, cnt_r, cnt_n, cnt_m;
};

// Operator descriptors, built per instance inside the kernels.  The magnetic operator has 3
// unknowns per cell, the pressure and energy operators 1 (the energy operator reads its
// coefficient from `hhat` and its precE divisor from `Tdiag`).
struct MagOp {
  Dims d;
  const double* Qe;   // padded planes (B in 5..7)
  const double* u;    // padded planes (rho in 0)
  const double* S;
};
struct POp {
  Dims d; Phys P;
  const double* hhat; const double* eta;   // padded planes
  const double* u;                         // padded planes (rho)
  const double* visc;                      // padded plane Vbase_slope(p^e, rho) (RK only, else null)
  const double* pfreeze;                   // compact (Jacobian mode)
  const double* vslope;                    // compact V_slope(p_freeze) (RK only, else null)
  const double* Tdiag;                     // compact
  const double* S;
  int jacobian, precond;
};
// Magnetic_Subsystem_potential_AX's coefficients: B of Qe, rho of u, and the viscosity maxima in S.
struct PotOp {
  Dims d;
  const double* Qe;   // padded planes (B in 5..7)
  const double* u;    // padded planes (rho in 0)
  const double* S;
};
// What the host passes: buffer ids and flags (pot: the vector-potential operator on the magnetic
// slots, 3 unknowns per cell).
struct MagIds { int Qe, u, pot, refl; };   // refl: the magnetic operator with reflective-wall flags
struct PIds { int u, jacobian, precond, energy; };   // energy: Energy_Subsystem_AX on the same slots

__device__ __forceinline__ MagOp make_mag_op(const Inst& I, const MagIds& m) {
  MagOp o; o.d = I.d; o.Qe = gmem(I.buf[m.Qe]); o.u = gmem(I.buf[m.u]); o.S = gmem(I.S); return o;
}
__device__ __forceinline__ PotOp make_pot_op(const Inst& I, const MagIds& m) {
  PotOp o; o.d = I.d; o.Qe = gmem(I.buf[m.Qe]); o.u = gmem(I.buf[m.u]); o.S = gmem(I.S); return o;
}
__device__ __forceinline__ POp make_p_op(const Inst& I, const Phys& P, const PIds& q) {
  POp o; o.d = I.d; o.P = P; o.hhat = gmem(I.buf[B_HHAT]); o.eta = gmem(I.buf[B_ETA]); o.u = gmem(I.buf[q.u]);
  const bool rk = P.eos != IMEX_EOS_IDEAL;
  o.visc = rk ? gmem(I.buf[B_VISC]) : nullptr; o.pfreeze = gmem(I.buf[B_P]); o.vslope = rk ? gmem(I.buf[B_VSLOPE]) : nullptr;
  o.Tdiag = gmem(I.buf[B_TDIAG]); o.S = gmem(I.S); o.jacobian = q.jacobian; o.precond = q.precond;
  return o;
}

// skip kinds: launches of a Jacobian solve whose Newton iteration has already converged
enum { SKIP_NONE = 0, SKIP_NEWTON = 1 };
__device__ __forceinline__ bool skipped(const Inst& I, int kind) { return kind == SKIP_NEWTON && gmem(I.K)->newton_conv; }

// The host side of a launch: the instance table, the block map of the launch's shape and the
// number of instances (grid = total blocks), and the stream.
struct Launch {
  const Inst* IT;
  BMap bm[NSHAPE];
  int nblocks[NSHAPE];
  hipStream_t st;
};

// ---- stage kernels (kernels_stage.hip) ----
// ghost-frame fill of n planes: plane k[p] of buffer b[p]; kind 0 cell, 1 x-normal faces, 2 y-normal faces
void launch_bc_fill(const Launch& L, int n, const int* b, const int* k, int kind, const signed char* jv = nullptr);
void launch_step_start(const Launch& L, const Phys& P, int mode, int iin);
void launch_stage_combine(const Launch& L, int ibxn, int ibyn);
void launch_qe_prep(const Launch& L, const Phys& P, int iQe);
void launch_shock(const Launch& L, const Phys& P, int iQe);
// tile: width of the explicit tile (16, 32 or 64 cells; 256 cells per tile); SH_EXPL is mapped for it
void launch_explicit(const Launch& L, const Phys& P, int kind, int iQe, int iQs, int iu, int mixed, int tile, int rows);
void launch_mag_rhs(const Launch& L, const Phys& P, int iu, int iQe, int iQs, int ib, int ix0);
void launch_mag_writeback(const Launch& L, int iu);
void launch_emf_flux(const Launch& L, int iu, int iQe);
void launch_p_rhs(const Launch& L, const Phys& P, int pass, int iu);
void launch_p_momentum(const Launch& L, int iu);
void launch_p_energy(const Launch& L, const Phys& P, int iu);
// energy formula: RHS, coefficient plane, initial guess and precE divisor; then E and momentum update
void launch_e_rhs(const Launch& L, const Phys& P, int iu, int iQe);
void launch_e_update(const Launch& L, const Phys& P, int iu);
// EMF-consistent corner assembly into B_EZPP from the magnetic subsystem's face fluxes (B_FX, B_FY)
void launch_ct_corners(const Launch& L, int iu);

// ---- per-step diagnostics (kernels_stage.hip) ----
#define DIAG_REC 12          // doubles per record: timestep, t, dt, div_error, B_inf, dt_comp, emin, x, y, mass, intE, magE_max
#define DIAG_SCRATCH 32      // doubles after the ring: one record written on demand, then 16 CT-EMF statistics
// the step's diagnostics from state iu and faces ibx/iby (face div B when face_div, else cell), into
// ring slot diag_n % diag_cap (append) or the scratch record
void launch_diag(const Launch& L, int iu, int ibx, int iby, int face_div, double cfl, int append);
// Diagnose_CT_EMF on state iu: 16 statistics into the scratch area (after its record)
void launch_ct_emf_diag(const Launch& L, const Phys& P, int iu);

// ---- vector potential (kernels_potential.hip) ----
void launch_bc_a(const Launch& L, int ia);                       // bound_cond_A on 3 planes
void launch_a_combine(const Launch& L, int ian, int ia1, int ias);
// MagSubsystem_UpdateA RHS b (compact) and initial guess from A `ia0`
void launch_a_rhs(const Launch& L, int iu, int iQe, int ias, int ia0, int ib, int ix0);
void launch_pot_visc(const Launch& L, int iQs);                  // S_GBX, S_GBY from Qs
void launch_a_update(const Launch& L, int ix, int ia, int iu);    // A <- x, bound_cond_A, B = curl A
void launch_unstaggered_ct(const Launch& L, int iu, int ias, int ia01, int ia);
void launch_copy_faces(const Launch& L, int ibx, int iby, int obx, int oby);
void launch_divb_cell(const Launch& L, int iu);
void launch_ct(const Launch& L, const Phys& P, int plain, int iu, int ibxqs, int ibyqs, int ibx, int iby);
void launch_con2pri(const Launch& L, const Phys& P, int icon, int ipri);
void launch_pri2con(const Launch& L, const Phys& P, int ipri, int icon);
void launch_divb(const Launch& L, int ibx, int iby);
void launch_visc_rk(const Launch& L, const Phys& P, int iu);

// ---- Krylov (kernels_krylov.hip) ----
// op = 0 magnetic or, with mo.pot, vector potential (3 unknowns per cell), 1 pressure or, with
// po.energy, energy (1 unknown per cell); prec = 0 FP64 basis and
// operator, 1 FP32 basis and operator (inner solves of the mixed form; dots, norms, Givens and
// the residual of the refinement stay FP64)
void gmres_residual(const Launch& L, int op, int prec, const MagIds& mo, const PIds& po, const Phys& P, int ix, int ib,
                    int outer, int x_is_zero, int skip);
void gmres_residual_fin(const Launch& L, int op, int prec, int outer, int skip, double outer_tol, double inner_tol);
void arnoldi_apply(const Launch& L, int op, int prec, const MagIds& mo, const PIds& po, const Phys& P, int j);
void arnoldi_apply_fin(const Launch& L, int op, int j);
void arnoldi_update(const Launch& L, int op, int prec, int j, int m);
void arnoldi_update_fin(const Launch& L, int op, int prec, int j, int m);
void gmres_xupdate(const Launch& L, int op, int prec, int ix, int x_zero, int skip);
void op_apply(const Launch& L, int op, const MagIds& mo, const PIds& po, const Phys& P, int ix, int iy);
// blocks of the Newton residual (ideal-gas or Redlich-Kwong instantiation) resident on the current
// device at once, 0 if the occupancy query fails
int newton_residual_resident(bool rk);
void newton_residual(const Launch& L, const PIds& po, const Phys& P, int ip, int ib, int it, double ntol);
void newton_residual_fin(const Launch& L, int it, double ntol);
void newton_mbig(const Launch& L, const Phys& P, int ib, int iu);
void newton_mbig_fin_prhs(const Launch& L, const Phys& P);   // Mbig from k_p_rhs's partials
void vslope_launch(const Launch& L, const Phys& P, int ipf, int iu);
void axpy_launch(const Launch& L, int iy, int ix, int skip);
void pstats_launch(const Launch& L, int ip, double pfloor, double thr);
void pstats_fin_newton(const Launch& L, double thr);           // from the last Newton residual's partials

// Records the message returned by imex_last_error and returns `code` (entry points outside driver.hip).
int imex_internal_error(int code, const std::string& msg);
