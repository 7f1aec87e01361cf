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
// pressure operator (implicit_operators.cpp:505-616), ideal gas: mode 0 A^T y into xbar; mode 1
// coefficient cotangents (hhat, eta, scalars) of y . A x
void adj_p_op(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, int mode, const double* x,
              const double* hhat, const double* eta, const double* y, double* xbar, double* hhatbar, double* etabar,
              double* sbar);
// Jacobi diagonal of the pressure operator (V' + T diag, implicit_operators.cpp:656-668)
void adj_p_diag(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, const double* hhat, double* diag);
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
