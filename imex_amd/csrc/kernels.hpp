// kernels.hpp — internal declarations shared by the kernel files and the host driver.
#pragma once
#include "device_common.hpp"

#define MAXM 32            // largest GMRES restart length supported
#define TX 64              // operator tile (cells) in x
#define TY 4               // operator tile (cells) in y
#define MAXSWEEP 16        // refinement sweeps recorded per mixed-form solve

struct PlaneSet { double* p[8]; int n; };

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

// Operator descriptors.  The magnetic operator has 3 unknowns per cell, the pressure operator 1.
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

// ---- launchers (kernels_stage.hip) ----
void launch_bc_fill(const Dims& d, const PlaneSet& ps, int kind, hipStream_t st);

__global__ void k_step_start(Dims d, Phys P, int mode, const double* uin, double* un, double* partial, int* eflag);
__global__ void k_step_finish(Dims d, Phys P, const double* partial, int nblk, double* S);
__global__ void k_stage_combine(Dims d, const double* un, const double* u1, double* Qe, double* Qs,
                                const double* bxn, const double* bx1, const double* byn, const double* by1,
                                double* bxqs, double* byqs, const double* S);
__global__ void k_qe_prep(Dims d, Phys P, const double* Qe, double* pe, double* cf, int* eflag);
__global__ void k_shock_chi(Dims d, Phys P, const double* Qe, const double* cf, double* chi);
__global__ void k_shock_eta(Dims d, const double* chi, double* eta);
template <class T>
__global__ void k_explicit(Dims d, Phys P, int kind, const double* Qe, const double* Qs, double* u, const double* S);
__global__ void k_mag_rhs(Dims d, Phys P, const double* u, const double* Qe, const double* Qs, const double* pe,
                          double* b, double* x0, double* hhat, const double* S, int* eflag);
__global__ void k_mag_writeback(Dims d, const double* x, double* u);
__global__ void k_emf_flux(Dims d, const double* u, const double* Qe, const double* pe, double* Fx, double* Fy, const double* S);
__global__ void k_p_rhs(Dims d, int pass, const double* u, const double* hhat, double* mxs, double* mys,
                        const double* pe, double* b, double* p0, double* Tdiag, const double* S);
__global__ void k_p_momentum(Dims d, const double* p, const double* mxs, const double* mys, double* u, const double* S);
__global__ void k_p_energy(Dims d, Phys P, const double* p, double* u, int* eflag);
__global__ void k_ct_corner(Dims d, const double* ezx, const double* ezy, double sx, double sy, const double* u, double* Ezpp);
__global__ void k_ct_faces(Dims d, const double* bxqs, const double* byqs, const double* Ezpp, double* bx, double* by,
                           double* u, const double* S);
__global__ void k_ct_plain_emf(Dims d, Phys P, const double* u, double* ezx, double* ezy);
__global__ void k_con2pri(Dims d, Phys P, const double* con, double* pri, int* eflag);
__global__ void k_pri2con(Dims d, Phys P, const double* pri, double* con, int* eflag);
__global__ void k_divb(Dims d, const double* bx, const double* by, double* partial);

// ---- Krylov (kernels_krylov.hip) ----
struct KrylovBufs {
  double* w;           // unnormalised next basis vector / residual (FP64 form)
  double* Av;          // preconditioned operator image of the current basis vector (FP64 form)
  float* w32;          // the same two vectors for the FP32-stored inner solves of the mixed form
  float* Av32;
  double* partial;     // per-block partial sums, MAXM+2 per block
  int nblk_apply, nblk_vec;
};

// op = 0 magnetic (3 unknowns per cell), 1 pressure (1 unknown per cell); prec = 0 FP64 basis and
// operator, 1 FP32 basis and operator (inner solves of the mixed form; dots, norms, Givens and
// the residual of the refinement stay FP64)
void gmres_residual(int op, int prec, const MagOp& mo, const POp& po, const KrylovBufs& kb, const double* x, const double* b,
                    KrylovDev* K, int outer, int x_is_zero, const int* skip, hipStream_t st);
void gmres_residual_fin(const Dims& d, int prec, const KrylovBufs& kb, KrylovDev* K, int outer, const int* skip,
                        double outer_tol, double inner_tol, hipStream_t st);
void arnoldi_apply(int op, int prec, const MagOp& mo, const POp& po, const KrylovBufs& kb, double* const* V64,
                   float* const* V32, int j, KrylovDev* K, hipStream_t st);
void arnoldi_apply_fin(const Dims& d, const KrylovBufs& kb, int j, KrylovDev* K, hipStream_t st);
void arnoldi_update(int nc, int prec, const Dims& d, const KrylovBufs& kb, double* const* V64, float* const* V32, int j,
                    int m, KrylovDev* K, hipStream_t st);
void arnoldi_update_fin(int nc, int prec, const Dims& d, const KrylovBufs& kb, int j, int m, KrylovDev* K, hipStream_t st);
void gmres_xupdate(int nc, int prec, const Dims& d, const KrylovBufs& kb, double* const* V64, float* const* V32, double* x,
                   KrylovDev* K, hipStream_t st);
void op_apply(int op, const MagOp& mo, const POp& po, const double* x, double* Ax, hipStream_t st);
void newton_residual(const POp& po, const KrylovBufs& kb, const double* p, const double* b, double* ng, KrylovDev* K,
                     int it, double ntol, int* eflag, hipStream_t st);
void newton_residual_fin(const Dims& d, const KrylovBufs& kb, KrylovDev* K, int it, double ntol, hipStream_t st);
void newton_mbig(const Dims& d, const Phys& P, const double* b, const double* u, double* S, double* partial, KrylovDev* K,
                 int* eflag, hipStream_t st);
void vslope_launch(const Dims& d, const Phys& P, const double* pf, const double* u, const double* S, double* vs, int* eflag,
                   hipStream_t st);
void axpy_launch(long long n, double* y, const double* x, const int* skip, hipStream_t st);
