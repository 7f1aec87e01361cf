// adjoint.hpp — host-visible launchers of the discrete-adjoint kernels (adjoint_kernels.hip).
//
// Every launcher works on one instance's padded planes (component k of a field at + k*PS) on a
// periodic grid, reading interior cells through a wrapped index.  "bar" arguments are cotangent
// buffers that the launchers add into (atomics); callers zero them first.
#pragma once
#include "adjoint_dev.hpp"

struct AdjScal { double s[SC_N]; };   // h, a_max, B_max_x, B_max_y, vx_max, vy_max

// explicit flux update u[k0..k1) = Qs - h div F(recon(Qe), B of uB)  (flux.cpp:290-326, 409-455)
// fbx / fby: 5 planes each of face cotangents (scratch); Qebar gets components 0..4, uBbar 5..7.
void adj_explicit(hipStream_t st, const AD& a, const Phys& P, int kind, const AdjScal& sc,
                  const double* Qe, const double* Qs, const double* uB, const double* uout, const double* ubar,
                  double* Qebar, double* Qsbar, double* uBbar, double* fbx, double* fby, double* sbar);
// magnetic right-hand side (imex_subsystems.cpp:549-613): b (3 planes) from u, Qe, pe, Qs
void adj_mag_rhs(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* Qe, const double* pe,
                 const double* Qs, const double* bbar, double* ubar, double* Qebar, double* pebar, double* Qsbar,
                 double* sbar);
// magnetic operator (implicit_operators.cpp:141-353): mode 0 adds A^T y into xbar; mode 1 adds the
// cotangents of the coefficients (Qe B, rho of u, scalars) of y . A x
void adj_mag_op(hipStream_t st, const AD& a, const AdjScal& sc, int mode, const double* x, const double* Qe,
                const double* u, const double* y, double* xbar, double* Qebar, double* ubar, double* sbar);
// the magnetic operator's Jacobian in the unknown (81 planes of nx ny, compact interior order) and
// its transpose applied to y (3 planes, written to out)
#define ADJ_MAG_JAC 81
void adj_mag_jac(hipStream_t st, const AD& a, const AdjScal& sc, const double* Qe, const double* u, double* J);
void adj_mag_op_t(hipStream_t st, const AD& a, const double* J, const double* y, double* out);
// EMF fluxes (ct_flux.cpp:5-127): obar = (Fx bar, Fy bar) planes
void adj_emf(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* Qe, const double* pe,
             const double* fbar, double* ubar, double* Qebar, double* pebar, double* sbar);
// p(Qe) and hhat = h(Qe) rho^e / rho (implicit_operators.cpp:506-531): obar = (pe bar, hhat bar)
void adj_pe_hhat(hipStream_t st, const AD& a, const Phys& P, const double* Qe, const double* u, const double* obar,
                 double* Qebar, double* ubar);
// shock indicator (flux.cpp:5-76): chi plane from Qe, eta bar -> chi bar -> Qe bar
void adj_shock(hipStream_t st, const AD& a, const Phys& P, const double* Qe, double* chi, const double* etabar,
               double* chibar, double* Qebar);
// pressure right-hand side (imex_subsystems.cpp:123-175, 223-284), pass 0 or 1
void adj_p_rhs(hipStream_t st, const AD& a, const AdjScal& sc, int pass, const double* ucur, const double* min,
               const double* hhat, const double* bbar, double* ucurbar, double* minbar, double* hhatbar, double* sbar);
// pressure operator (implicit_operators.cpp:505-616), carrier mode: mode 0 A^T y into xbar; mode 1
// coefficient cotangents (hhat, eta, the Redlich-Kwong viscosity plane visc, scalars) of y . A x; visc null
// for the ideal gas
void adj_p_op(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, int mode, const double* x,
              const double* hhat, const double* eta, const double* visc, const double* y, double* xbar, double* hhatbar,
              double* etabar, double* viscbar, double* sbar);
// Jacobi diagonal of the pressure operator (V' + T diag, implicit_operators.cpp:656-668)
void adj_p_diag(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, const double* hhat, double* diag);
// The nested Newton's solution (V_mod(p, rho) + T p = b): dv = dV_mod/dp - 1/(gamma-1) at the solution
// p (mb: the solve's Mbig, one device double); its rho cotangent ubar[rho] -= lam dV_mod/drho; out += dv * in
// (in null: out += dv); and, with the floor on, the Mbig path of the cotangents of b (into lam, the cotangent
// of b from the transposed solve) and, for Redlich-Kwong, of rho.
void adj_newton_dv(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* mb, double* dv);
void adj_newton_rho(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* mb,
                    const double* lam, double* ubar);
void adj_diag_add(hipStream_t st, const AD& a, const double* dv, const double* in, double* out);
// dv = vs - 1/(gamma-1) for the forward's Jacobian-mode diagonal vs (a padded plane)
void adj_slope_dv(hipStream_t st, const AD& a, const Phys& P, const double* vs, double* dv);
void adj_floor_mbig(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* b, const double* u,
                    double* lam, double* ubar, double* work);   // work: >= 256 doubles of scratch
// Redlich-Kwong viscosity scale of the pressure operator, Vbase_slope(p(Qe), rho), and its cotangents
void adj_visc(hipStream_t st, const AD& a, const Phys& P, const double* pe, const double* u, double* visc);
void adj_visc_vjp(hipStream_t st, const AD& a, const Phys& P, const double* pe, const double* u, const double* viscbar,
                  double* pebar, double* ubar);
// momentum update m = m_in - h grad p (imex_subsystems.cpp:186-215)
void adj_momentum(hipStream_t st, const AD& a, const AdjScal& sc, const double* umid, const double* uin,
                  const double* ubar, double* pbar, double* sbar);
// energy u4 = Vbase(p) + ke + |B|^2/2 (imex_subsystems.cpp:286-302)
void adj_energy(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* u4bar,
                double* pbar, double* ubar);
// CT (ct_flux.cpp:150-191, constrained_transport.cpp:147-164): from the cotangents of the output
// faces and cell B (ubar planes 5, 6, consumed and zeroed) to those of the Qs faces, EMF fluxes
// and the post-pressure state (through Et)
void adj_ct(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* bxo, const double* byo,
            const double* bxqs, const double* byqs, double* ubar, const double* bxbar, const double* bybar,
            double* bxqsbar, double* byqsbar, double* fxbar, double* fybar, double* scratch, double* sbar);
// start-of-step maxima (timestep.cpp:5-67, 89-120, 254-273): value planes (6) of the per-cell
// quantities whose maxima give ch, v_max, B_max_x, B_max_y, vx_max, vy_max
void adj_max_values(hipStream_t st, const AD& a, const Phys& P, const double* U, double* vals);
// index of the largest value of plane q (lowest index on ties), for q < 6; out: 6 ints (device)
void adj_argmax(hipStream_t st, const AD& a, const double* vals, int* out, double* work);
// U bar at the argmax cells += qbar[q] d value_q / d U (qbar: 6 device doubles)
void adj_max_grad(hipStream_t st, const AD& a, const Phys& P, const double* U, const int* argmax, const double* qbar,
                  double* Ubar);
// primitive cotangent from a conservative one: pribar = (d con / d pri)^T conbar
void adj_pri2con(hipStream_t st, const AD& a, const Phys& P, const double* pri, const double* conbar, double* pribar);

// ---------------------------------------------------------------- transposed GMRES, device side
// Restart length of the transposed solves and their device control block: the Hessenberg column
// work, the Givens rotations and the convergence decision live on the device, so the host queues
// Arnoldi steps without waiting on each dot product (it polls `cdone` a step behind).
#define ADJ_M 30
struct AdjKry {
  double H[(ADJ_M + 1) * ADJ_M];  // rotated Hessenberg columns: H[i * ADJ_M + j], i <= j
  double cs[ADJ_M], sn[ADJ_M], g[ADJ_M + 1], y[ADJ_M];
  double d1[ADJ_M + 2], d2[ADJ_M + 2];   // dot sums of the two Gram-Schmidt passes of the current step
  double bnorm, tol, res, hn, beta;
  int jlast, cdone, solved, iters;   // last column, cycle ended, solve converged, Arnoldi steps so far
};
struct AdjVecs { const double* p[ADJ_M + 1]; };
// Per-block partial sums: `part` holds akr_blocks x (ADJ_M + 1) doubles (dot products), `pn`
// akr_blocks1 doubles (norms).
int akr_blocks(const AD& a, int nc);
int akr_blocks1(const AD& a, int nc);
// bnorm = ||b||, res = bnorm; solved when b = 0
void akr_init(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* b, double* pn, double tol);
// start of a cycle: r = b - t, beta = ||r||; v0 = r / beta (zero once the solve is done)
void akr_residual(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* b, const double* t, double* r,
                  double* v0, double* pn);
// one classical Gram-Schmidt pass against V_0..V_{nv-1}: d (0: d1, 1: d2) = V^T w, w -= V d; with
// pn, the blocks' shares of ||w||^2 after it
void akr_gs_pass(hipStream_t st, const AD& a, int nc, int nv, const AdjVecs& V, double* w, AdjKry* K, int d,
                 double* part, double* pn);
// column j: H = d1 + d2, hn = ||w||, rotations, residual estimate, end of cycle
void akr_column(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* pn, int j, int m);
// v_{j+1} = w / hn (zero once the cycle has ended)
void akr_next(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* w, double* v);
// back substitution and x += sum_i y_i V_i over the cycle's columns
void akr_update_x(hipStream_t st, const AD& a, int nc, const AdjVecs& V, AdjKry* K, double* x);

// vector helpers over interior cells of nc planes
void adj_dots(hipStream_t st, const AD& a, int nc, int nv, const double* const* v, const double* w, double* out);
void adj_axpys(hipStream_t st, const AD& a, int nc, int nv, const double* coef, const double* const* v, double* w,
               double sign);
void adj_scale_copy(hipStream_t st, const AD& a, int nc, const double* in, double s, double* out);
void adj_div(hipStream_t st, const AD& a, int nc, const double* in, const double* diag, double* out);
void adj_zero(hipStream_t st, const AD& a, int nc, double* p);
void adj_add(hipStream_t st, const AD& a, int nc, const double* in, double s, double* out);
// compact Krylov vector (M_P / M_B order) <-> padded planes
void adj_from_compact(hipStream_t st, const AD& a, int nc, const double* in, double* out);
void adj_to_compact(hipStream_t st, const AD& a, int nc, const double* in, double* out);
