/*
 * imex_oracle.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference's per-step IMEX hot path
 * (ZefengChen-Kairos/IMEX), written from scratch and kept in the reference's
 * arithmetic order (operand order, left-to-right evaluation, no FMA) so that it
 * can serve as the parity oracle for the HIP implementation under imex_amd/.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  The product path never calls it.
 *
 * Scope: GLF = true, add_pvisc = MaterialSpeedp, add_bvisc = true, PressureFormula
 * (DefinitionofU, all_correction) and EnergyFormula, StaggerCT (EMF-consistent and plain),
 * UnStaggerCT, PotentialA and NoDivFree, NoRecon / Minmod / VanLeer / ENO3 / WENO5 of the
 * conservative or the primitive state, ideal gas and Redlich-Kwong, IMEXdt, RKs = 2;
 * periodic, transmissive, reflective, Dirichlet-ambient and jet-inflow boundaries
 * (include/config.hpp:27-65, 100-125, 246-273).
 *
 * How each part is pinned is listed in the headers of tests/test_oracle.py and
 * tests/test_oracle_pins.py.  The reference itself cannot be compiled in this image
 * (Eigen 3 is absent, Makefile:37), so the oracle is pinned by closed forms, Fourier
 * symbols, invariants and brute force, the composed subsystem algebra included.
 *
 * Two timing modes: faithful = 1 recomputes the per-apply invariants exactly as
 * the reference does (ucon2upri(Qe), ShockDetector, 5 EOS enthalpies per cell
 * inside every pressure-operator apply, ucon2upri(Qe) inside every magnetic
 * apply); faithful = 0 hoists them.  Both produce identical bits.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>
#include <functional>
#include <omp.h>
#include "../include/imex.h"

namespace {

constexpr int NV = 8;
constexpr int GC = 3;   // ghost_cells (config.hpp:55)

struct Ctx {
  using P_t = imex_params;
  imex_params P;
  int nx, ny, NX, NY;       // interior, padded
  double dx, dy;
  double Mbig = 0.0;
  int faithful = 0;
  int eos_warn = 0;
  imex_stats st{};
};

// ---------------------------------------------------------------- indexing
inline double& G(std::vector<double>& a, const Ctx& C, int i, int j, int k) { return a[NV * (i + j * C.NX) + k]; }
inline double G(const std::vector<double>& a, const Ctx& C, int i, int j, int k) { return a[NV * (i + j * C.NX) + k]; }
inline double& A(std::vector<double>& a, const Ctx& C, int i, int j) { return a[i + j * C.NX]; }
inline double A(const std::vector<double>& a, const Ctx& C, int i, int j) { return a[i + j * C.NX]; }
using Vec = std::vector<double>;

// ---------------------------------------------------------------- MHD jet inflow strip (config.hpp:253-273)
inline bool jet_on(const Ctx& C) { return C.P.bc_y == IMEX_BC_JET_INFLOW; }
inline bool jet_inlet(const Ctx& C, int i) {                         // cell centre |x_c| < 0.05 (:261-265)
  if (!jet_on(C)) return false;
  double xc = C.P.x0 + (i - GC + 0.5) * C.dx;
  return std::fabs(xc) < C.P.jet_halfwidth;
}
inline bool jet_inlet_face(const Ctx& C, int i) {                    // x face i+1/2 (:266-270)
  if (!jet_on(C)) return false;
  double xf = C.P.x0 + (i - GC + 1.0) * C.dx;
  return std::fabs(xf) < C.P.jet_halfwidth - 1e-12;
}
// Dirichlet ambient (config.hpp:246-250): boundary cells of the pressure solve and ghost cells
inline bool dir_x(const Ctx& C) { return C.P.bc_x == IMEX_BC_DIRICHLET; }
inline bool dir_y(const Ctx& C) { return C.P.bc_y == IMEX_BC_DIRICHLET; }
inline bool dir_bnd(const Ctx& C, int i, int j) {                    // imex_subsystems.cpp:161, implicit_operators.cpp:550
  return (dir_x(C) && (i == GC || i == C.nx + GC - 1)) || (dir_y(C) && (j == GC || j == C.ny + GC - 1));
}
inline double jet_E_val(const Ctx& C) {                              // :271-273 (ideal-gas form, as written)
  return C.P.jet_p / (C.P.gamma - 1.0) + 0.5 * C.P.jet_rho * C.P.jet_v * C.P.jet_v + 0.5 * C.P.jet_B0 * C.P.jet_B0;
}

// ---------------------------------------------------------------- EOS (include/config.hpp:130-239)
// Redlich-Kwong temperature Newton solves: the reference loops without a cap
// (config.hpp:141-142, 157-158); the oracle caps at 1000 iterations and flags.
double rk_T_from_rho_p(Ctx& C, double rho, double p) {              // config.hpp:137-144
  const double mR = C.P.rk_R, ma = C.P.rk_alfa, mb = C.P.rk_b;
  double v = 1.0 / rho;
  double T_old, T_new = (1.0 / rho - mb) * (p + ma * rho * rho) / mR;
  int it = 0;
  do {
    T_old = T_new;
    double f = mR * T_old / (v - mb) - ma / std::sqrt(T_old) / v / (v + mb) - p;           // :130
    double df = mR / (v - mb) + 0.5 * ma / T_old / std::sqrt(T_old) / v / (v + mb);         // :131
    T_new = T_old - f / df;
    if (++it > 1000) { __atomic_store_n(&C.eos_warn, 1, __ATOMIC_RELAXED); break; }
  } while (std::fabs(T_new - T_old) / std::fabs(T_old) > 1e-14);
  return T_new;
}
double rk_e_from_rho_p(Ctx& C, double rho, double p) {              // config.hpp:146-151
  double v = 1.0 / rho;
  double T = rk_T_from_rho_p(C, rho, p);
  double U = std::log(v / (v + C.P.rk_b));
  return C.P.rk_cv * T + 1.5 * C.P.rk_alfa / C.P.rk_b * U / std::sqrt(T);
}
double rk_p_from_rho_e(Ctx& C, double rho, double e) {              // config.hpp:153-161
  const double mR = C.P.rk_R, ma = C.P.rk_alfa, mb = C.P.rk_b, mcv = C.P.rk_cv;
  double v = 1.0 / rho;
  double T_old, T_new = (e + rho * ma) / mcv;
  int it = 0;
  do {
    T_old = T_new;
    double U = std::log(v / (v + mb));
    double f = mcv * T_old + 1.5 * ma / mb * U / std::sqrt(T_old) - e;                      // :132
    double df = mcv - 0.75 * ma / mb * U / T_old / std::sqrt(T_old);                        // :133
    T_new = T_old - f / df;
    if (++it > 1000) { __atomic_store_n(&C.eos_warn, 1, __ATOMIC_RELAXED); break; }
  } while (std::fabs(T_new - T_old) / std::fabs(T_old) > 1e-14);
  double a = ma / std::sqrt(T_new);
  return mR * T_new / (v - mb) - a / v / (v + mb);
}
double rk_sound(Ctx& C, double rho, double p) {                     // config.hpp:163-173
  const double mR = C.P.rk_R, ma = C.P.rk_alfa, mb = C.P.rk_b, mcv = C.P.rk_cv;
  double v = 1.0 / rho;
  double T = rk_T_from_rho_p(C, rho, p);
  double U = std::log(v / (v + mb));
  double p_T = mR / (v - mb) + ma / (2.0 * T * std::sqrt(T) * v * (v + mb));
  double e_T = mcv - 0.75 * ma / mb * U / (T * std::sqrt(T));
  double p_v = -mR * T / ((v - mb) * (v - mb)) + ma * (2.0 * v + mb) / (std::sqrt(T) * v * v * ((v + mb) * (v + mb)));
  double e_v = 1.5 * ma / mb * (1.0 / v - 1.0 / (v + mb)) / std::sqrt(T);
  double a2 = p / (rho * rho) * (p_T / e_T) + v * v * (e_v - e_T * (p_v / p_T));
  return std::sqrt(a2);
}
inline double eos_p_from_rho_e(Ctx& C, double rho, double e) {      // config.hpp:176-179
  return C.P.eos == IMEX_EOS_IDEAL ? (C.P.gamma - 1.0) * rho * e : rk_p_from_rho_e(C, rho, e);
}
inline double eos_e_from_rho_p(Ctx& C, double rho, double p) {      // config.hpp:180-183
  return C.P.eos == IMEX_EOS_IDEAL ? p / ((C.P.gamma - 1.0) * rho) : rk_e_from_rho_p(C, rho, p);
}
inline double eos_sound_a(Ctx& C, double rho, double p) {           // config.hpp:184-187
  return C.P.eos == IMEX_EOS_IDEAL ? std::sqrt(C.P.gamma * p / rho) : rk_sound(C, rho, p);
}
inline double eos_sound_a2(Ctx& C, double rho, double p) { double a = eos_sound_a(C, rho, p); return a * a; }
inline double eos_fast_speed(Ctx& C, double rho, double p, double Bn, double B2) {  // config.hpp:192-198
  double a2 = eos_sound_a2(C, rho, p);
  double t = a2 + B2 / rho;
  double disc = t * t - 4.0 * a2 * Bn * Bn / rho;
  if (disc < 0) disc = 0;
  return std::sqrt(0.5 * (t + std::sqrt(disc)));
}
inline double Vbase(Ctx& C, double p, double rho) {                 // config.hpp:206-209
  return C.P.eos == IMEX_EOS_IDEAL ? p / (C.P.gamma - 1.0) : rho * rk_e_from_rho_p(C, rho, p);
}
inline double Vbase_slope(Ctx& C, double p, double rho) {           // config.hpp:211-217
  if (C.P.eos == IMEX_EOS_IDEAL) return 1.0 / (C.P.gamma - 1.0);
  const double d = 1e-8;
  double ep = rk_e_from_rho_p(C, rho, p + d);
  double em = rk_e_from_rho_p(C, rho, p - d);
  return rho * (ep - em) / (2.0 * d);
}
inline double V_mod(Ctx& C, double p, double rho) {                 // config.hpp:231-235
  if (!C.P.floor_on) return Vbase(C, p, rho);
  return (p >= C.P.p_floor) ? Vbase(C, p, rho) : Vbase(C, C.P.p_floor, rho) + C.Mbig * (p - C.P.p_floor);
}
inline double V_slope(Ctx& C, double p, double rho) {               // config.hpp:236-239
  if (!C.P.floor_on) return Vbase_slope(C, p, rho);
  return (p >= C.P.p_floor) ? Vbase_slope(C, p, rho) : C.Mbig;
}

// ---------------------------------------------------------------- pointwise state (src/state.cpp, reconstruction.cpp:291-309, timestep.cpp:224-247)
inline double calecon(const double* con) {                          // reconstruction.cpp:291-304
  double rho = con[0];
  double vx = con[1] / con[0], vy = con[2] / con[0], vz = con[3] / con[0];
  double U = con[4], Bx = con[5], By = con[6], Bz = con[7];
  double B_square = Bx * Bx + By * By + Bz * Bz;
  double v_square = vx * vx + vy * vy + vz * vz;
  return (U - 0.5 * rho * v_square - 0.5 * B_square) / rho;
}
inline void con2pri(Ctx& C, double* pri, const double* con) {       // state.cpp:206-218
  double e = calecon(con);
  pri[0] = con[0];
  pri[1] = con[1] / con[0];
  pri[2] = con[2] / con[0];
  pri[3] = con[3] / con[0];
  pri[4] = eos_p_from_rho_e(C, con[0], e);
  pri[5] = con[5]; pri[6] = con[6]; pri[7] = con[7];
}
inline void pri2con(Ctx& C, double* con, const double* pri) {       // state.cpp:220-240
  double e = eos_e_from_rho_p(C, pri[0], pri[4]);
  double vx = pri[1], vy = pri[2], vz = pri[3];
  double Bx = pri[5], By = pri[6], Bz = pri[7];
  double B_square = Bx * Bx + By * By + Bz * Bz;
  double v_square = vx * vx + vy * vy + vz * vz;
  con[0] = pri[0];
  con[1] = pri[0] * pri[1];
  con[2] = pri[0] * pri[2];
  con[3] = pri[0] * pri[3];
  con[4] = pri[0] * e + 0.5 * pri[0] * v_square + 0.5 * B_square;
  con[5] = pri[5]; con[6] = pri[6]; con[7] = pri[7];
}
inline double calenthalpy(Ctx& C, const double* con) {              // timestep.cpp:224-230
  double e = calecon(con);
  double pri[NV];
  con2pri(C, pri, con);
  return e + pri[4] / pri[0];
}
inline double calkinetic(const double* con) {                       // timestep.cpp:239-247
  double rho = con[0];
  double vx = con[1] / con[0], vy = con[2] / con[0], vz = con[3] / con[0];
  return 0.5 * rho * (vx * vx + vy * vy + vz * vz);
}
void grid_con2pri(Ctx& C, Vec& pri, const Vec& con) {               // state.cpp:257-270
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++) con2pri(C, &pri[NV * (i + j * C.NX)], &con[NV * (i + j * C.NX)]);
}
void grid_pri2con(Ctx& C, Vec& con, const Vec& pri) {               // state.cpp:242-255
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++) pri2con(C, &con[NV * (i + j * C.NX)], &pri[NV * (i + j * C.NX)]);
}

// ---------------------------------------------------------------- boundary conditions
// bound_cond_u (boundary.cpp:5-175), periodic and transmissive branches: x pass over all rows, then y pass over all columns.
// bound_cond_u (boundary.cpp:5-175); prim = g_bc_prim (the grid holds primitives: slot 4 is p)
// Reflective (:17-54, 89-126): the mirror cell, momenta and B negated (primitive velocities alike).
inline double refl_sign(int k) { return (k == 0 || k == 4) ? 1.0 : -1.0; }
void bc_grid(const Ctx& C, Vec& u, bool prim = false) {
  const int nx = C.nx, ny = C.ny;
  for (int j = 0; j < C.NY; j++)
    for (int g = 0; g < GC; g++)
      for (int k = 0; k < NV; k++) {
        if (C.P.bc_x == IMEX_BC_DIRICHLET) {
          continue;                                                   // below, per cell
        } else if (C.P.bc_x == IMEX_BC_REFLECTIVE) {
          G(u, C, g, j, k) = refl_sign(k) * G(u, C, 2 * GC - 1 - g, j, k);
          G(u, C, nx + GC + g, j, k) = refl_sign(k) * G(u, C, nx + GC - 1 - g, j, k);
        } else if (C.P.bc_x == IMEX_BC_PERIODIC) {
          G(u, C, g, j, k) = G(u, C, nx + g, j, k);
          G(u, C, nx + GC + g, j, k) = G(u, C, GC + g, j, k);
        } else {
          G(u, C, g, j, k) = G(u, C, GC, j, k);
          G(u, C, nx + GC + g, j, k) = G(u, C, nx + GC - 1, j, k);
        }
      }
  // Dirichlet (:56-76, 128-147): rho_amb, v = 0, B copied from the boundary cell, slot 4 = p_amb (primitive
  // grid) or E = p_amb/(gamma-1) + |B|^2/2 of the ghost's B
  auto dirichlet = [&](int gi, int gj, int si, int sj) {
    G(u, C, gi, gj, 0) = C.P.rho_amb;
    G(u, C, gi, gj, 1) = 0; G(u, C, gi, gj, 2) = 0; G(u, C, gi, gj, 3) = 0;
    for (int k = 5; k < NV; k++) G(u, C, gi, gj, k) = G(u, C, si, sj, k);
    G(u, C, gi, gj, 4) = prim ? C.P.p_amb
                              : C.P.p_amb / (C.P.gamma - 1.0) + 0.5 * (G(u, C, gi, gj, 5) * G(u, C, gi, gj, 5) +
                                                                       G(u, C, gi, gj, 6) * G(u, C, gi, gj, 6) +
                                                                       G(u, C, gi, gj, 7) * G(u, C, gi, gj, 7));
  };
  if (dir_x(C))
    for (int j = 0; j < C.NY; j++)
      for (int g = 0; g < GC; g++) { dirichlet(g, j, GC, j); dirichlet(nx + GC + g, j, nx + GC - 1, j); }
  for (int i = 0; i < C.NX; i++)
    for (int g = 0; g < GC; g++)
      for (int k = 0; k < NV; k++) {
        if (C.P.bc_y == IMEX_BC_DIRICHLET) {
          continue;
        } else if (C.P.bc_y == IMEX_BC_REFLECTIVE) {
          G(u, C, i, g, k) = refl_sign(k) * G(u, C, i, 2 * GC - 1 - g, k);
          G(u, C, i, ny + GC + g, k) = refl_sign(k) * G(u, C, i, ny + GC - 1 - g, k);
        } else if (C.P.bc_y == IMEX_BC_PERIODIC) {
          G(u, C, i, g, k) = G(u, C, i, ny + g, k);
          G(u, C, i, ny + GC + g, k) = G(u, C, i, GC + g, k);
        } else {
          G(u, C, i, g, k) = G(u, C, i, GC, k);
          G(u, C, i, ny + GC + g, k) = G(u, C, i, ny + GC - 1, k);
        }
      }
  if (dir_y(C))
    for (int i = 0; i < C.NX; i++)
      for (int g = 0; g < GC; g++) { dirichlet(i, g, i, GC); dirichlet(i, ny + GC + g, i, ny + GC - 1); }
  if (jet_on(C))                                                      // y_JetInflow (:148-172): the strip
    for (int i = 0; i < C.NX; i++) {
      if (!jet_inlet(C, i)) continue;
      for (int g = 0; g < GC; g++) {
        const Ctx::P_t& P = C.P;
        double s[NV] = {P.jet_rho, 0.0, prim ? P.jet_v : P.jet_rho * P.jet_v, 0.0, prim ? P.jet_p : jet_E_val(C),
                        0.0, P.jet_B0, 0.0};
        for (int k = 0; k < NV; k++) G(u, C, i, g, k) = s[k];
      }
    }
}
// bound_cond_array (timestep.cpp:166-222); reflective: the mirror value negated, for every array
void bc_array(const Ctx& C, Vec& u) {
  const int nx = C.nx, ny = C.ny;
  for (int j = 0; j < C.NY; j++)
    for (int g = 0; g < GC; g++) {
      if (C.P.bc_x == IMEX_BC_REFLECTIVE) { A(u, C, g, j) = -A(u, C, 2 * GC - 1 - g, j); A(u, C, nx + GC + g, j) = -A(u, C, nx + GC - 1 - g, j); }
      else if (C.P.bc_x == IMEX_BC_PERIODIC) { A(u, C, g, j) = A(u, C, nx + g, j); A(u, C, nx + GC + g, j) = A(u, C, GC + g, j); }
      else { A(u, C, g, j) = A(u, C, GC, j); A(u, C, nx + GC + g, j) = A(u, C, nx + GC - 1, j); }
    }
  for (int i = 0; i < C.NX; i++)
    for (int g = 0; g < GC; g++) {
      if (C.P.bc_y == IMEX_BC_REFLECTIVE) { A(u, C, i, g) = -A(u, C, i, 2 * GC - 1 - g); A(u, C, i, ny + GC + g) = -A(u, C, i, ny + GC - 1 - g); }
      else if (C.P.bc_y == IMEX_BC_PERIODIC) { A(u, C, i, g) = A(u, C, i, ny + g); A(u, C, i, ny + GC + g) = A(u, C, i, GC + g); }
      else { A(u, C, i, g) = A(u, C, i, GC); A(u, C, i, ny + GC + g) = A(u, C, i, ny + GC - 1); }
    }
}
// bound_cond_staggered (constrained_transport.cpp:11-64)
void bc_staggered(const Ctx& C, Vec& u, int axis) {
  const int nx = C.nx, ny = C.ny;
  if (C.P.bc_x == IMEX_BC_PERIODIC) {
    for (int j = 0; j < C.NY; j++)
      for (int g = 0; g < GC; g++) { A(u, C, g, j) = A(u, C, nx + g, j); A(u, C, nx + GC + g, j) = A(u, C, GC + g, j); }
  } else if (axis == 0) {
    for (int j = 0; j < C.NY; j++) {
      for (int k = 0; k < GC - 1; k++) A(u, C, k, j) = A(u, C, GC - 1, j);
      for (int k = 0; k < GC; k++) A(u, C, nx + GC + k, j) = A(u, C, nx + GC - 1, j);
    }
  } else {
    for (int j = 0; j < C.NY; j++)
      for (int k = 0; k < GC; k++) { A(u, C, k, j) = A(u, C, GC, j); A(u, C, nx + GC + k, j) = A(u, C, nx + GC - 1, j); }
  }
  if (C.P.bc_y == IMEX_BC_PERIODIC) {
    for (int i = 0; i < C.NX; i++)
      for (int g = 0; g < GC; g++) { A(u, C, i, g) = A(u, C, i, ny + g); A(u, C, i, ny + GC + g) = A(u, C, i, GC + g); }
  } else if (axis == 1) {
    for (int i = 0; i < C.NX; i++) {
      for (int k = 0; k < GC - 1; k++) A(u, C, i, k) = A(u, C, i, GC - 1);
      for (int k = 0; k < GC; k++) A(u, C, i, ny + GC + k) = A(u, C, i, ny + GC - 1);
    }
  } else {
    for (int i = 0; i < C.NX; i++)
      for (int k = 0; k < GC; k++) { A(u, C, i, k) = A(u, C, i, GC); A(u, C, i, ny + GC + k) = A(u, C, i, ny + GC - 1); }
  }
  if (jet_on(C)) {                                                    // the inflow strip (:50-63)
    if (axis == 1) {
      for (int i = 0; i < C.NX; i++)
        if (jet_inlet(C, i))
          for (int k = 0; k < GC - 1; k++) A(u, C, i, k) = C.P.jet_B0;
    } else {
      for (int i = 0; i < C.NX - 1; i++)
        if (jet_inlet_face(C, i))
          for (int k = 0; k < GC; k++) A(u, C, i, k) = 0.0;
    }
  }
}
// M_P / M_B index maps (implicit_operators.cpp:5-139): periodic wrap, reflective mirror, clamp otherwise.
inline int mapx(const Ctx& C, int i) {
  if (i < GC) return C.P.bc_x == IMEX_BC_PERIODIC ? C.nx + i : (C.P.bc_x == IMEX_BC_REFLECTIVE ? 2 * GC - 1 - i : GC);
  if (i > C.nx + GC - 1)
    return C.P.bc_x == IMEX_BC_PERIODIC ? i - C.nx : (C.P.bc_x == IMEX_BC_REFLECTIVE ? 2 * (C.nx + GC) - i - 1 : C.nx + GC - 1);
  return i;
}
inline int mapy(const Ctx& C, int j) {
  if (j < GC) return C.P.bc_y == IMEX_BC_PERIODIC ? C.ny + j : (C.P.bc_y == IMEX_BC_REFLECTIVE ? 2 * GC - 1 - j : GC);
  if (j > C.ny + GC - 1)
    return C.P.bc_y == IMEX_BC_PERIODIC ? j - C.ny : (C.P.bc_y == IMEX_BC_REFLECTIVE ? 2 * (C.ny + GC) - j - 1 : C.ny + GC - 1);
  return j;
}
inline int M_P(const Ctx& C, int i, int j) { return (mapx(C, i) - GC) + C.nx * (mapy(C, j) - GC); }
inline int M_B(const Ctx& C, int i, int j, int k) { return M_P(C, i, j) + (k - 5) * C.nx * C.ny; }

// ---------------------------------------------------------------- reconstruction (reconstruction.cpp:5-287)
struct Recon { Vec xL, xR, yL, yR; };

void tvd_reconstruction(Ctx& C, const Vec& ucon, Recon& w) {
  const int nx = C.nx, ny = C.ny;
  const bool prim = C.P.recon_state == IMEX_RECONSTATE_PRIM;        // reconstate (:11-24, 272-281)
  Vec upri;
  if (prim) { upri.assign(ucon.size(), 0.0); grid_con2pri(C, upri, ucon); }
  const Vec& src = prim ? upri : ucon;
  Vec u(ucon.size(), 0.0);
  for (int j = GC; j < ny + GC; j++)
    for (int i = GC - 1; i < nx + GC + 1; i++)
      for (int k = 0; k < NV; k++) G(u, C, i, j, k) = G(src, C, i, j, k);
  bc_grid(C, u);
  w.xL.assign(ucon.size(), 0.0); w.xR.assign(ucon.size(), 0.0);
  w.yL.assign(ucon.size(), 0.0); w.yR.assign(ucon.size(), 0.0);
  #pragma omp parallel for schedule(static)
  for (int j = GC; j < ny + GC; j++)
    for (int i = GC - 1; i < nx + GC + 1; i++)
      for (int k = 0; k < NV; k++) {
        auto U = [&](int a, int b) { return G(u, C, a, b, k); };
        double c = 0.0;
        double s_xR = U(i + 1, j) - U(i, j), s_xL = U(i, j) - U(i - 1, j);
        double s_yR = U(i, j + 1) - U(i, j), s_yL = U(i, j) - U(i, j - 1);
        if (C.P.recon == IMEX_RECON_NONE) {                         // :40-46
          G(w.xL, C, i, j, k) = U(i, j); G(w.xR, C, i, j, k) = U(i, j);
          G(w.yL, C, i, j, k) = U(i, j); G(w.yR, C, i, j, k) = U(i, j);
        } else if (C.P.recon == IMEX_RECON_ENO3) {                  // :104-183
          for (int dir = 0; dir < 2; dir++) {
            double um2, um1, u0, up1, up2;
            if (dir == 0) { um2 = U(i - 2, j); um1 = U(i - 1, j); u0 = U(i, j); up1 = U(i + 1, j); up2 = U(i + 2, j); }
            else { um2 = U(i, j - 2); um1 = U(i, j - 1); u0 = U(i, j); up1 = U(i, j + 1); up2 = U(i, j + 2); }
            double q0L = -1.0 / 6.0 * um2 + 5.0 / 6.0 * um1 + 1.0 / 3.0 * u0;
            double q1L = 1.0 / 3.0 * um1 + 5.0 / 6.0 * u0 - 1.0 / 6.0 * up1;
            double q2L = 11.0 / 6.0 * u0 - 7.0 / 6.0 * up1 + 1.0 / 3.0 * up2;
            double q0R = 1.0 / 3.0 * um2 - 7.0 / 6.0 * um1 + 11.0 / 6.0 * u0;
            double q1R = -1.0 / 6.0 * um1 + 5.0 / 6.0 * u0 + 1.0 / 3.0 * up1;
            double q2R = 1.0 / 3.0 * u0 + 5.0 / 6.0 * up1 - 1.0 / 6.0 * up2;
            double diff1 = std::fabs(u0 - um1), diff2 = std::fabs(up1 - u0), diff3, diff4;
            double wL, wR;
            if (diff1 >= diff2) {                                     // the smoother side is the right one
              diff3 = std::fabs(std::fabs(up2 - up1) - std::fabs(up1 - u0));
              diff4 = std::fabs(std::fabs(up1 - u0) - std::fabs(u0 - um1));
              if (diff3 >= diff4) { wL = q1L; wR = q1R; } else { wL = q2L; wR = q2R; }
            } else {
              diff3 = std::fabs(std::fabs(u0 - um1) - std::fabs(um1 - um2));
              diff4 = std::fabs(std::fabs(up1 - u0) - std::fabs(u0 - um1));
              if (diff4 >= diff3) { wL = q0L; wR = q0R; } else { wL = q1L; wR = q1R; }
            }
            if (dir == 0) { G(w.xL, C, i, j, k) = wL; G(w.xR, C, i, j, k) = wR; }
            else { G(w.yL, C, i, j, k) = wL; G(w.yR, C, i, j, k) = wR; }
          }
        } else if (C.P.recon == IMEX_RECON_MINMOD) {                // :46-75
          if (s_xL * s_xR <= 0) c = 0;
          else if (std::fabs(s_xR) <= std::fabs(s_xL)) c = s_xR;
          else if (std::fabs(s_xL) < std::fabs(s_xR)) c = s_xL;
          G(w.xL, C, i, j, k) = U(i, j) - 0.5 * c;
          G(w.xR, C, i, j, k) = U(i, j) + 0.5 * c;
          if (s_yL * s_yR <= 0) c = 0;
          else if (std::fabs(s_yR) <= std::fabs(s_yL)) c = s_yR;
          else if (std::fabs(s_yL) < std::fabs(s_yR)) c = s_yL;
          if (C.P.minmod_central_y) c = 0.5 * (s_yL + s_yR);       // vortex quirk :70-72
          G(w.yL, C, i, j, k) = U(i, j) - 0.5 * c;
          G(w.yR, C, i, j, k) = U(i, j) + 0.5 * c;
        } else if (C.P.recon == IMEX_RECON_VANLEER) {               // :76-103
          if (std::fabs(s_xR) < 1e-16) c = 0;
          else { double r = s_xL / s_xR; c = (r < 0) ? 0 : std::fmin(2 / (1 + r), (2 * r) / (1 + r)); }
          G(w.xL, C, i, j, k) = U(i, j) - 0.5 * c * s_xL;
          G(w.xR, C, i, j, k) = U(i, j) + 0.5 * c * s_xR;
          if (std::fabs(s_yR) < 1e-16) c = 0;
          else { double r = s_yL / s_yR; c = (r < 0) ? 0 : std::fmin(2 / (1 + r), (2 * r) / (1 + r)); }
          G(w.yL, C, i, j, k) = U(i, j) - 0.5 * c * s_yL;
          G(w.yR, C, i, j, k) = U(i, j) + 0.5 * c * s_yR;
        } else {                                                    // WENO5 :184-266
          for (int dir = 0; dir < 2; dir++) {
            double um2, um1, u0, up1, up2;
            if (dir == 0) { um2 = U(i - 2, j); um1 = U(i - 1, j); u0 = U(i, j); up1 = U(i + 1, j); up2 = U(i + 2, j); }
            else { um2 = U(i, j - 2); um1 = U(i, j - 1); u0 = U(i, j); up1 = U(i, j + 1); up2 = U(i, j + 2); }
            double t0 = um2 - 2 * um1 + u0, t1 = um2 - 4 * um1 + 3 * u0;
            double beta0 = 13.0 / 12.0 * (t0 * t0) + 1.0 / 4.0 * (t1 * t1);
            double t2 = um1 - 2 * u0 + up1, t3 = um1 - up1;
            double beta1 = 13.0 / 12.0 * (t2 * t2) + 1.0 / 4.0 * (t3 * t3);
            double t4 = u0 - 2 * up1 + up2, t5 = 3 * u0 - 4 * up1 + up2;
            double beta2 = 13.0 / 12.0 * (t4 * t4) + 1.0 / 4.0 * (t5 * t5);
            double e0 = 1e-6 + beta0, e1 = 1e-6 + beta1, e2 = 1e-6 + beta2;
            // left face value (i-1/2)
            double q0 = -1.0 / 6.0 * um2 + 5.0 / 6.0 * um1 + 1.0 / 3.0 * u0;
            double q1 = 1.0 / 3.0 * um1 + 5.0 / 6.0 * u0 - 1.0 / 6.0 * up1;
            double q2 = 11.0 / 6.0 * u0 - 7.0 / 6.0 * up1 + 1.0 / 3.0 * up2;
            double a0 = 3.0 / 10.0 / (e0 * e0), a1 = 6.0 / 10.0 / (e1 * e1), a2 = 1.0 / 10.0 / (e2 * e2);
            double wL = a0 / (a0 + a1 + a2) * q0 + a1 / (a0 + a1 + a2) * q1 + a2 / (a0 + a1 + a2) * q2;
            // right face value (i+1/2)
            q0 = 1.0 / 3.0 * um2 - 7.0 / 6.0 * um1 + 11.0 / 6.0 * u0;
            q1 = -1.0 / 6.0 * um1 + 5.0 / 6.0 * u0 + 1.0 / 3.0 * up1;
            q2 = 1.0 / 3.0 * u0 + 5.0 / 6.0 * up1 - 1.0 / 6.0 * up2;
            a0 = 1.0 / 10.0 / (e0 * e0); a1 = 6.0 / 10.0 / (e1 * e1); a2 = 3.0 / 10.0 / (e2 * e2);
            double wR = a0 / (a0 + a1 + a2) * q0 + a1 / (a0 + a1 + a2) * q1 + a2 / (a0 + a1 + a2) * q2;
            if (dir == 0) { G(w.xL, C, i, j, k) = wL; G(w.xR, C, i, j, k) = wR; }
            else { G(w.yL, C, i, j, k) = wL; G(w.yR, C, i, j, k) = wR; }
          }
        }
      }
  if (prim) {     // upri2ucon of the face states (:272-281); only the cells bound_cond_u keeps are converted
    Vec* ws[4] = {&w.xL, &w.xR, &w.yL, &w.yR};
    for (Vec* wv : ws)
      for (int j = GC; j < ny + GC; j++)
        for (int i = GC; i < nx + GC; i++) {
          double pri[NV], con[NV];
          for (int k = 0; k < NV; k++) pri[k] = G(*wv, C, i, j, k);
          pri2con(C, con, pri);
          for (int k = 0; k < NV; k++) G(*wv, C, i, j, k) = con[k];
        }
  }
  bc_grid(C, w.xL); bc_grid(C, w.xR); bc_grid(C, w.yL); bc_grid(C, w.yR);
}

// ---------------------------------------------------------------- explicit fluxes (flux.cpp:202-455, state.cpp:27-44, 86-106, 298-322)
void f_convective(double* f, const double* con) {
  double rho = con[0], vx = con[1] / con[0], vy = con[2] / con[0], vz = con[3] / con[0];
  double k = calkinetic(con);
  f[0] = rho * vx; f[1] = rho * vx * vx; f[2] = rho * vy * vx; f[3] = rho * vz * vx; f[4] = k * vx;
}
void g_convective(double* g, const double* con) {
  double rho = con[0], vx = con[1] / con[0], vy = con[2] / con[0], vz = con[3] / con[0];
  double k = calkinetic(con);
  g[0] = rho * vy; g[1] = rho * vx * vy; g[2] = rho * vy * vy; g[3] = rho * vz * vy; g[4] = k * vy;
}
void mag_flux_p_x(double* f, const double* Q) {
  double B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
  double vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) / Q[0];
  f[0] = 0; f[1] = B2 * 0.5 - Q[5] * Q[5]; f[2] = -Q[5] * Q[6]; f[3] = -Q[5] * Q[7];
  f[4] = Q[1] / Q[0] * B2 - Q[5] * vB;
}
void mag_flux_p_y(double* f, const double* Q) {
  double B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
  double vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) / Q[0];
  f[0] = 0; f[1] = -Q[5] * Q[6]; f[2] = B2 * 0.5 - Q[6] * Q[6]; f[3] = -Q[6] * Q[7];
  f[4] = Q[2] / Q[0] * B2 - Q[6] * vB;
}
// Energy-formula pieces (state.cpp:272-296, 325-346): magnetic flux with the (2 - gamma) B^2/2
// momentum term and v B^2/2 energy term, and the -(gamma - 1) |m|^2 / (2 rho) normal-momentum term.
void mag_flux_e_x(double* f, const double* Q, double gam) {
  double B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
  double vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) / Q[0];
  f[0] = 0; f[1] = (2 - gam) * B2 * 0.5 - Q[5] * Q[5]; f[2] = -Q[5] * Q[6]; f[3] = -Q[5] * Q[7];
  f[4] = Q[1] / Q[0] * B2 * 0.5 - Q[5] * vB;
}
void mag_flux_e_y(double* f, const double* Q, double gam) {
  double B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
  double vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) / Q[0];
  f[0] = 0; f[1] = -Q[5] * Q[6]; f[2] = (2 - gam) * B2 * 0.5 - Q[6] * Q[6]; f[3] = -Q[6] * Q[7];
  f[4] = Q[2] / Q[0] * B2 * 0.5 - Q[6] * vB;
}
inline double kin_flux(const double* Q, double gam) {
  return -(gam - 1) * 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) / Q[0];
}
// ---------------------------------------------------------------- shock detector (flux.cpp:5-76)
void shock_detector(Ctx& C, Vec& eta, const Vec& Qe) {
  Vec pri(Qe.size());
  grid_con2pri(C, pri, Qe);
  Vec et(C.NX * C.NY, 0.0);
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double cf_min = 1.0e300;
      for (int di = -1; di <= 1; di++)
        for (int dj = -1; dj <= 1; dj++) {
          double rho = G(pri, C, i + di, j + dj, 0), p = G(pri, C, i + di, j + dj, 4);
          double Bx = G(pri, C, i + di, j + dj, 5), By = G(pri, C, i + di, j + dj, 6), Bz = G(pri, C, i + di, j + dj, 7);
          double a2 = eos_sound_a2(C, rho, p);
          double cA2 = (Bx * Bx + By * By + Bz * Bz) / rho;
          double cf = std::sqrt(a2 + cA2);
          if (cf < cf_min) cf_min = cf;
        }
      double div_u = (G(pri, C, i + 1, j, 1) - G(pri, C, i - 1, j, 1)) / (2.0 * C.dx)
                   + (G(pri, C, i, j + 1, 2) - G(pri, C, i, j - 1, 2)) / (2.0 * C.dy);
      double h = std::min(C.dx, C.dy);
      double num = h * div_u + C.P.shock_kappa1 * cf_min;
      double den = C.P.shock_kappa1 * cf_min;
      A(et, C, i, j) = std::min(1.0, std::max(0.0, -num / den));
    }
  bc_array(C, et);
  eta.assign(C.NX * C.NY, 0.0);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double m = 0.0;
      for (int di = -1; di <= 1; di++)
        for (int dj = -1; dj <= 1; dj++)
          if (A(et, C, i + di, j + dj) > m) m = A(et, C, i + di, j + dj);
      A(eta, C, i, j) = m;
    }
  bc_array(C, eta);
}

}  // namespace

// The remaining pieces need the scalars; they live in a second anonymous block.
namespace {

struct Stage {
  Vec u_con, Qe, Qs;
  Vec bx, by, bx_qs, by_qs, fx, fy;
  Vec A, As;                 // vector potential of the stage and its implicit combination (Grid, slots 0..2)
};

// Rusanov-type face flux with GLF dissipation s_max = a_max (flux.cpp:202-243, 329-367).
// kind 0 convective (RusanovC), 1 pressure formula (Pressure_flux), 2 energy formula (Energy_flux,
// flux.cpp:78-163: convective with no energy part, plus magnetic, plus kinetic-energy terms).
void rusanov(const imex_scalars& S, double gam, int dir, int kind, const double* QL, const double* QR, double* flux) {
  double fL[5], fR[5];
  if (dir == 0) { f_convective(fL, QL); f_convective(fR, QR); }
  else { g_convective(fL, QL); g_convective(fR, QR); }
  if (kind == 2) {
    fL[4] = 0; fR[4] = 0;
    double bL[5], bR[5];
    if (dir == 0) { mag_flux_e_x(bL, QL, gam); mag_flux_e_x(bR, QR, gam); }
    else { mag_flux_e_y(bL, QL, gam); mag_flux_e_y(bR, QR, gam); }
    double kL[5] = {0, 0, 0, 0, 0}, kR[5] = {0, 0, 0, 0, 0};
    kL[1 + dir] = kin_flux(QL, gam); kR[1 + dir] = kin_flux(QR, gam);
    for (int k = 0; k < 5; k++) { fL[k] = fL[k] + bL[k] + kL[k]; fR[k] = fR[k] + bR[k] + kR[k]; }
  }
  if (kind == 1) {
    double bL[5], bR[5];
    if (dir == 0) { mag_flux_p_x(bL, QL); mag_flux_p_x(bR, QR); }
    else { mag_flux_p_y(bL, QL); mag_flux_p_y(bR, QR); }
    for (int k = 0; k < 5; k++) { fL[k] = fL[k] + bL[k]; fR[k] = fR[k] + bR[k]; }
  }
  double s_max = S.a_max;
  for (int k = 0; k < 5; k++) flux[k] = 0.5 * (fL[k] + fR[k]) - 0.5 * s_max * (QR[k] - QL[k]);
}

// RusanovC_Update (flux.cpp:409-455), Pressure_flux_Update (flux.cpp:290-326) and
// Energy_flux_Update (flux.cpp:165-200).
void flux_update(Ctx& C, const imex_scalars& S, Vec& u_con, const Vec& Qs, const Recon& w, int kind) {
  const double h = S.adt;
  #pragma omp parallel for schedule(static)
  for (int j = GC; j < C.ny + GC; j++)
    for (int i = GC; i < C.nx + GC; i++) {
      double QL[NV], QR[NV], fp[5], fm[5], gp[5], gm[5];
      for (int k = 0; k < NV; k++) { QL[k] = G(w.xR, C, i, j, k); QR[k] = G(w.xL, C, i + 1, j, k); }
      rusanov(S, C.P.gamma, 0, kind, QL, QR, fp);
      for (int k = 0; k < NV; k++) { QL[k] = G(w.yR, C, i, j, k); QR[k] = G(w.yL, C, i, j + 1, k); }
      rusanov(S, C.P.gamma, 1, kind, QL, QR, gp);
      for (int k = 0; k < NV; k++) { QL[k] = G(w.xR, C, i - 1, j, k); QR[k] = G(w.xL, C, i, j, k); }
      rusanov(S, C.P.gamma, 0, kind, QL, QR, fm);
      for (int k = 0; k < NV; k++) { QL[k] = G(w.yR, C, i, j - 1, k); QR[k] = G(w.yL, C, i, j, k); }
      rusanov(S, C.P.gamma, 1, kind, QL, QR, gm);
      for (int k = 0; k < 5; k++)
        G(u_con, C, i, j, k) = G(Qs, C, i, j, k) - h / C.dx * (fp[k] - fm[k]) - h / C.dy * (gp[k] - gm[k]);
    }
}

// ---------------------------------------------------------------- magnetic operator (implicit_operators.cpp:141-353), GLF branch
void mag_apply(Ctx& C, const imex_scalars& S, const Vec& u_con, const Vec& Qe, const double* x, double* Ax) {
  const double h = S.adt, dx = C.dx, dy = C.dy;
  if (C.faithful) { Vec pri(Qe.size()); grid_con2pri(C, pri, Qe); }   // loop-invariant work the reference repeats (:143-144)
  // reflective boundaries (:146-211): a neighbour's unknown is negated when the clamped neighbour index
  // is the first or last interior cell -- at the boundary cell and at the cell next to it
  const bool rx = C.P.bc_x == IMEX_BC_REFLECTIVE, ry = C.P.bc_y == IMEX_BC_REFLECTIVE;
  #pragma omp parallel
  {
  int i = 0, j = 0;                   // per thread: the lambdas below read the current cell
  double fxl = 1, fxr = 1, fyl = 1, fyr = 1;
  auto X = [&](int a, int b, int k) {
    double f = (a < i ? fxl : a > i ? fxr : 1.0) * (b < j ? fyl : b > j ? fyr : 1.0);
    return x[M_B(C, a, b, k)] * f;
  };
  auto Q = [&](int a, int b, int k) { return G(Qe, C, a, b, k); };
  auto R = [&](int a, int b) { return G(u_con, C, a, b, 0); };
  // the worksharing loop runs over its own index: the lambdas above hold references to this thread's i, j
  #pragma omp for schedule(static)
  for (int ii = GC; ii < C.nx + GC; ii++)
    for (i = ii, j = GC; j < C.ny + GC; j++) {
      fxl = (rx && std::max(GC, i - 1) == GC) ? -1.0 : 1.0;
      fxr = (rx && std::min(C.nx + GC - 1, i + 1) == C.nx + GC - 1) ? -1.0 : 1.0;
      fyl = (ry && std::max(GC, j - 1) == GC) ? -1.0 : 1.0;
      fyr = (ry && std::min(C.ny + GC - 1, j + 1) == C.ny + GC - 1) ? -1.0 : 1.0;
      double rho_p0 = 0.5 * (R(i, j) + R(i + 1, j)), rho_m0 = 0.5 * (R(i, j) + R(i - 1, j));
      double rho_0p = 0.5 * (R(i, j) + R(i, j + 1)), rho_0m = 0.5 * (R(i, j) + R(i, j - 1));
      double n_p0[3], n_m0[3], n_0p[3], n_0m[3], n_pp[3], n_pm[3], n_mp[3], n_mm[3];
      double b_pp[3], b_pm[3], b_mp[3], b_mm[3];
      for (int c = 0; c < 3; c++) {
        int k = 5 + c;
        n_p0[c] = 0.5 * (Q(i, j, k) + Q(i + 1, j, k));
        n_m0[c] = 0.5 * (Q(i, j, k) + Q(i - 1, j, k));
        n_0p[c] = 0.5 * (Q(i, j, k) + Q(i, j + 1, k));
        n_0m[c] = 0.5 * (Q(i, j, k) + Q(i, j - 1, k));
        n_pp[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j + 1, k) + Q(i + 1, j + 1, k));
        n_pm[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j - 1, k) + Q(i + 1, j - 1, k));
        n_mp[c] = 0.25 * (Q(i, j, k) + Q(i - 1, j, k) + Q(i, j + 1, k) + Q(i - 1, j + 1, k));
        n_mm[c] = 0.25 * (Q(i, j, k) + Q(i - 1, j, k) + Q(i, j - 1, k) + Q(i - 1, j - 1, k));
        b_pp[c] = 0.25 * (X(i, j, k) + X(i + 1, j, k) + X(i, j + 1, k) + X(i + 1, j + 1, k));
        b_pm[c] = 0.25 * (X(i, j, k) + X(i + 1, j, k) + X(i, j - 1, k) + X(i + 1, j - 1, k));
        b_mp[c] = 0.25 * (X(i, j, k) + X(i - 1, j, k) + X(i, j + 1, k) + X(i - 1, j + 1, k));
        b_mm[c] = 0.25 * (X(i, j, k) + X(i - 1, j, k) + X(i, j - 1, k) + X(i - 1, j - 1, k));
      }
      // linearised magnetic pressure 0.5*(B^n . x) at a cell
      auto PM = [&](int a, int b) { return 0.5 * (Q(a, b, 5) * X(a, b, 5) + Q(a, b, 6) * X(a, b, 6) + Q(a, b, 7) * X(a, b, 7)); };
      // corner magnetic-pressure-like term 0.5*(B^n_c . b_c) - B^n_c[m] b_c[m]
      auto PC = [&](const double* n, const double* b, int m) { return 0.5 * (n[0] * b[0] + n[1] * b[1] + n[2] * b[2]) - n[m] * b[m]; };
      const double Bxn_0p = n_0p[0], Byn_0p = n_0p[1], Bzn_0p = n_0p[2];
      const double Bxn_0m = n_0m[0], Byn_0m = n_0m[1], Bzn_0m = n_0m[2];
      const double Bxn_p0 = n_p0[0], Byn_p0 = n_p0[1], Bzn_p0 = n_p0[2];
      const double Bxn_m0 = n_m0[0], Byn_m0 = n_m0[1], Bzn_m0 = n_m0[2];

      double Bx_1 = Bxn_0p / rho_0p * (1 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy * ((PM(i, j + 1) - Q(i, j + 1, 6) * X(i, j + 1, 6)) - (PM(i, j) - Q(i, j, 6) * X(i, j, 6))));
      double Bx_2 = Byn_0p / rho_0p * (1 / dx * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1 / dy * (Q(i, j + 1, 6) * X(i, j + 1, 5) - Q(i, j, 6) * X(i, j, 5)));
      double Bx_3 = Bxn_0m / rho_0m * (1 / dx * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy * ((PM(i, j) - Q(i, j, 6) * X(i, j, 6)) - (PM(i, j - 1) - Q(i, j - 1, 6) * X(i, j - 1, 6))));
      double Bx_4 = Byn_0m / rho_0m * (1 / dx * (PC(n_pm, b_pm, 0) - PC(n_mm, b_mm, 0)) - 1 / dy * (Q(i, j, 6) * X(i, j, 5) - Q(i, j - 1, 6) * X(i, j - 1, 5)));
      double ax5 = X(i, j, 5) + (h * h / dy) * ((Bx_1 + Bx_2) - (Bx_3 + Bx_4));

      double By_1 = Byn_p0 / rho_p0 * (-1 / dx * ((PM(i + 1, j) - Q(i + 1, j, 5) * X(i + 1, j, 5)) - (PM(i, j) - Q(i, j, 5) * X(i, j, 5))) + 1 / dy * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
      double By_2 = Bxn_p0 / rho_p0 * (-1 / dx * (Q(i + 1, j, 5) * X(i + 1, j, 6) - Q(i, j, 5) * X(i, j, 6)) + 1 / dy * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
      double By_3 = Byn_m0 / rho_m0 * (-1 / dx * ((PM(i, j) - Q(i, j, 5) * X(i, j, 5)) - (PM(i - 1, j) - Q(i - 1, j, 5) * X(i - 1, j, 5))) + 1 / dy * (n_mp[0] * b_mp[1] - n_mm[0] * b_mm[1]));
      double By_4 = Bxn_m0 / rho_m0 * (-1 / dx * (Q(i, j, 5) * X(i, j, 6) - Q(i - 1, j, 5) * X(i - 1, j, 6)) + 1 / dy * (PC(n_mp, b_mp, 1) - PC(n_mm, b_mm, 1)));
      double ax6 = X(i, j, 6) + (h * h / dx) * ((By_1 + By_2) - (By_3 + By_4));

      double Bz_1 = Bzn_p0 / rho_p0 * (-1 / dx * ((PM(i + 1, j) - Q(i + 1, j, 5) * X(i + 1, j, 5)) - (PM(i, j) - Q(i, j, 5) * X(i, j, 5))) + 1 / dy * (n_pp[1] * b_pp[0] - n_pm[1] * b_pm[0]));
      double Bz_2 = Bxn_p0 / rho_p0 * (-1 / dx * (Q(i + 1, j, 5) * X(i + 1, j, 7) - Q(i, j, 5) * X(i, j, 7)) - 1 / dy * (n_pp[1] * b_pp[2] - n_pm[1] * b_pm[2]));
      double Bz_3 = Bzn_m0 / rho_m0 * (-1 / dx * ((PM(i, j) - Q(i, j, 5) * X(i, j, 5)) - (PM(i - 1, j) - Q(i - 1, j, 5) * X(i - 1, j, 5))) + 1 / dy * (n_mp[1] * b_mp[0] - n_mm[1] * b_mm[0]));
      double Bz_4 = Bxn_m0 / rho_m0 * (-1 / dx * (Q(i, j, 5) * X(i, j, 7) - Q(i - 1, j, 5) * X(i - 1, j, 7)) - 1 / dy * (n_mp[1] * b_mp[2] - n_mm[1] * b_mm[2]));
      double Bz_5 = Bzn_0p / rho_0p * (1 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy * ((PM(i, j + 1) - Q(i, j + 1, 6) * X(i, j + 1, 6)) - (PM(i, j) - Q(i, j, 6) * X(i, j, 6))));
      double Bz_6 = Byn_0p / rho_0p * (-1 / dx * (n_pp[0] * b_pp[2] - n_mp[0] * b_mp[2]) - 1 / dy * (Q(i, j + 1, 6) * X(i, j + 1, 7) - Q(i, j, 6) * X(i, j, 7)));
      double Bz_7 = Bzn_0m / rho_0m * (1 / dx * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy * ((PM(i, j) - Q(i, j, 6) * X(i, j, 6)) - (PM(i, j - 1) - Q(i, j - 1, 6) * X(i, j - 1, 6))));
      double Bz_8 = Byn_0m / rho_0m * (-1 / dx * (n_pm[0] * b_pm[2] - n_mm[0] * b_mm[2]) - 1 / dy * (Q(i, j, 6) * X(i, j, 7) - Q(i, j - 1, 6) * X(i, j - 1, 7)));
      double ax7 = X(i, j, 7) + h * h / dx * ((Bz_1 + Bz_2) - (Bz_3 + Bz_4)) + h * h / dy * ((Bz_5 + Bz_6) - (Bz_7 + Bz_8));

      // GLF Laplacian, nu = B_max * d (:311-322)
      double nu_x = S.B_max_x * dx, nu_y = S.B_max_y * dy;
      double lap[3];
      for (int c = 0; c < 3; c++) {
        int k = 5 + c;
        lap[c] = nu_x / (dx * dx) * (X(i - 1, j, k) - 2.0 * X(i, j, k) + X(i + 1, j, k))
               + nu_y / (dy * dy) * (X(i, j - 1, k) - 2.0 * X(i, j, k) + X(i, j + 1, k));
      }
      Ax[M_B(C, i, j, 5)] = ax5 - h * lap[0];
      Ax[M_B(C, i, j, 6)] = ax6 - h * lap[1];
      Ax[M_B(C, i, j, 7)] = ax7 - h * lap[2];
    }
  }
  C.st.op_applies_mag++;
}

// ---------------------------------------------------------------- pressure operator (implicit_operators.cpp:505-616), GLF + MaterialSpeedp
struct PCoef { Vec hhat; Vec eta; Vec pe; };   // hoisted enthalpy ratio, shock indicator, p(Qe)

void p_coef(Ctx& C, const Vec& u_con, const Vec& Qe, PCoef& pc) {
  Vec pri(Qe.size());
  grid_con2pri(C, pri, Qe);
  shock_detector(C, pc.eta, Qe);
  pc.hhat.assign(C.NX * C.NY, 0.0);
  pc.pe.assign(C.NX * C.NY, 0.0);
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++) {
      A(pc.hhat, C, i, j) = calenthalpy(C, &Qe[NV * (i + j * C.NX)]) * G(Qe, C, i, j, 0) / G(u_con, C, i, j, 0);
      A(pc.pe, C, i, j) = G(pri, C, i, j, 4);
    }
}

void p_apply(Ctx& C, const imex_scalars& S, const Vec& u_con, const Vec& Qe, PCoef& pc, const double* x, double* Ax,
             int jacobian, const double* p_freeze) {
  if (C.faithful) p_coef(C, u_con, Qe, pc);   // the reference recomputes these inside every apply (:506-531)
  const double h = S.adt, dx = C.dx, dy = C.dy;
  auto H = [&](int a, int b) { return A(pc.hhat, C, a, b); };
  auto E = [&](int a, int b) { return A(pc.eta, C, a, b); };
  auto X = [&](int a, int b) { return x[M_P(C, a, b)]; };
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double h00 = H(i, j), hp0 = H(i + 1, j), hm0 = H(i - 1, j), h0p = H(i, j + 1), h0m = H(i, j - 1);
      double pii = X(i, j);
      double rho = G(u_con, C, i, j, 0);
      double Vdiag = jacobian ? V_slope(C, p_freeze[M_P(C, i, j)], rho) : (1.0 / (C.P.gamma - 1.0));
      if (dir_bnd(C, i, j)) { Ax[M_P(C, i, j)] = pii; continue; }       // strong Dirichlet: identity row (:549-550)
      double p_0m = (j == GC && jet_inlet(C, i)) ? 0.0 : X(i, j - 1);   // jet strip: homogeneous ghost (:559-561)
      double ax = Vdiag * pii - h * h * (1 / (dx * dx) * (0.5 * (hp0 + h00) * (X(i + 1, j) - X(i, j)) - 0.5 * (h00 + hm0) * (X(i, j) - X(i - 1, j)))
                                       + 1 / (dy * dy) * (0.5 * (h0p + h00) * (X(i, j + 1) - X(i, j)) - 0.5 * (h0m + h00) * (X(i, j) - p_0m)));
      double nu_x = S.vx_max * dx, nu_y = S.vy_max * dy;
      // getp (:533-547): a Dirichlet ghost is the negated mirror value in the viscosity term
      auto GP = [&](int a, int b) {
        int II = a, JJ = b; double sg = 1.0;
        if (dir_x(C)) {
          if (a < GC) { II = 2 * GC - 1 - a; sg = -sg; }
          else if (a >= C.nx + GC) { II = 2 * (C.nx + GC) - 1 - a; sg = -sg; }
        }
        if (dir_y(C)) {
          if (b < GC) { JJ = 2 * GC - 1 - b; sg = -sg; }
          else if (b >= C.ny + GC) { JJ = 2 * (C.ny + GC) - 1 - b; sg = -sg; }
        }
        return sg * X(II, JJ);
      };
      double lap_p = nu_x / (dx * dx) * (0.5 * (E(i + 1, j) + E(i, j)) * (GP(i + 1, j) - pii) - 0.5 * (E(i, j) + E(i - 1, j)) * (pii - GP(i - 1, j)))
                   + nu_y / (dy * dy) * (0.5 * (E(i, j + 1) + E(i, j)) * (GP(i, j + 1) - pii) - 0.5 * (E(i, j) + E(i, j - 1)) * (pii - GP(i, j - 1)));
      double visc_scale = Vbase_slope(C, A(pc.pe, C, i, j), rho);
      Ax[M_P(C, i, j)] = ax - visc_scale * h * lap_p;
    }
  C.st.op_applies_p++;
}

// ---------------------------------------------------------------- GMRES (implicit_operators.cpp:793-919)
// Sums in fixed chunks of 4096 elements, each in index order, then the chunk sums in order: the same
// bits for any thread count.
double dotv(const Vec& a, const Vec& b) {
  const long long n = (long long)a.size(), CH = 4096, nc = (n + CH - 1) / CH;
  std::vector<double> part(nc, 0.0);
  #pragma omp parallel for schedule(static)
  for (long long c = 0; c < nc; c++) {
    double s = 0;
    for (long long i = c * CH; i < std::min(n, (c + 1) * CH); i++) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0;
  for (long long c = 0; c < nc; c++) s += part[c];
  return s;
}
double normv(const Vec& a) { return std::sqrt(dotv(a, a)); }

int gmres(Ctx& C, Vec& x, const Vec& b, const std::function<void(const double*, double*)>& Aop,
          const std::function<void(double*, const double*)>* precond, imex_solve_info* info, long long* iters, long long* restarts) {
  const int JMAX = C.P.gmres_restart > 0 ? C.P.gmres_restart : 25;
  const int dim = (int)x.size();
  const int outmax = C.P.gmres_max_outer > 0 ? C.P.gmres_max_outer : dim;
  auto applyM = [&](Vec& out, const Vec& in) { if (precond) (*precond)(out.data(), in.data()); else out = in; };
  std::vector<Vec> v(JMAX + 1, Vec(dim)), Av(JMAX + 1, Vec(dim));
  Vec w(dim), Ax(dim), r0(dim), tmp(dim);
  std::vector<std::vector<double>> H(JMAX + 2, std::vector<double>(JMAX + 2, 0.0));
  std::vector<double> alpha(JMAX + 2, 0.0), gam(JMAX + 2, 0.0), s(JMAX + 2, 0.0), c(JMAX + 2, 0.0);
  double tol_eff = C.P.gmres_tol;
  int total = 0;
  if (info) { info->iters = 0; info->restarts = 0; info->res0 = 0; info->res = 0; info->status = 0; }
  for (int outer = 0; outer < outmax; outer++) {
    std::fill(Ax.begin(), Ax.end(), 0.0);
    Aop(x.data(), Ax.data());
    #pragma omp parallel for schedule(static)
    for (int n = 0; n < dim; n++) tmp[n] = b[n] - Ax[n];
    applyM(r0, tmp);
    gam[0] = normv(r0);
    if (outer == 0) { tol_eff = C.P.gmres_tol * gam[0] + C.P.gmres_tol; if (info) info->res0 = gam[0]; }
    if (info) info->res = gam[0];
    if (gam[0] <= tol_eff) { if (iters) *iters += total; if (restarts) *restarts += outer; if (info) { info->iters = total; info->restarts = outer; } return 0; }
    #pragma omp parallel for schedule(static)
    for (int n = 0; n < dim; n++) v[0][n] = r0[n] / gam[0];
    std::fill(Ax.begin(), Ax.end(), 0.0);
    Aop(v[0].data(), Ax.data());
    applyM(Av[0], Ax);
    for (auto& row : H) std::fill(row.begin(), row.end(), 0.0);
    for (int j = 0; j < JMAX; j++) {
      total++;
      for (int i = 0; i <= j; i++) H[i][j] = dotv(v[i], Av[j]);         // classical Gram-Schmidt (:850-852)
      w = Av[j];
      for (int i = 0; i <= j; i++) {
        double hij = H[i][j];
        #pragma omp parallel for schedule(static)
        for (int n = 0; n < dim; n++) w[n] -= hij * v[i][n];
      }
      H[j + 1][j] = normv(w);
      for (int i = 0; i < j; i++) {
        double t1 = c[i + 1] * H[i][j] + s[i + 1] * H[i + 1][j];
        double t2 = s[i + 1] * H[i][j] - c[i + 1] * H[i + 1][j];
        H[i][j] = t1; H[i + 1][j] = t2;
      }
      double beta = std::sqrt(H[j][j] * H[j][j] + H[j + 1][j] * H[j + 1][j]);
      s[j + 1] = H[j + 1][j] / beta;
      c[j + 1] = H[j][j] / beta;
      H[j][j] = beta;
      gam[j + 1] = s[j + 1] * gam[j];
      gam[j] = c[j + 1] * gam[j];
      if (info) info->res = std::fabs(gam[j + 1]);
      if (gam[j + 1] > tol_eff) {
        double hn = H[j + 1][j];
        #pragma omp parallel for schedule(static)
        for (int n = 0; n < dim; n++) v[j + 1][n] = w[n] / hn;
        std::fill(Ax.begin(), Ax.end(), 0.0);
        Aop(v[j + 1].data(), Ax.data());
        applyM(Av[j + 1], Ax);
      } else {
        for (int i = j; i >= 0; i--) {
          alpha[i] = gam[i];
          for (int k = i + 1; k <= j; k++) alpha[i] -= H[i][k] * alpha[k];
          alpha[i] /= H[i][i];
        }
        for (int i = 0; i <= j; i++) { double a = alpha[i]; for (int n = 0; n < dim; n++) x[n] += a * v[i][n]; }
        if (iters) *iters += total;
        if (restarts) *restarts += outer;
        if (info) { info->iters = total; info->restarts = outer; }
        return 0;
      }
    }
    std::fill(alpha.begin(), alpha.end(), 0.0);
    for (int i = JMAX - 1; i >= 0; i--) {
      alpha[i] = gam[i];
      for (int k = i + 1; k < JMAX; k++) alpha[i] -= H[i][k] * alpha[k];
      alpha[i] /= H[i][i];
    }
    for (int i = 0; i < JMAX; i++) { double a = alpha[i]; for (int n = 0; n < dim; n++) x[n] += a * v[i][n]; }
  }
  if (iters) *iters += total;
  if (restarts) *restarts += outmax;
  if (info) { info->iters = total; info->restarts = outmax; info->status = IMEX_WARN_GMRES_MAXOUTER; }
  C.st.warnings |= IMEX_WARN_GMRES_MAXOUTER;
  return IMEX_WARN_GMRES_MAXOUTER;
}

// ---------------------------------------------------------------- nested Newton (implicit_operators.cpp:625-738)
int pressure_newton(Ctx& C, const imex_scalars& S, const Vec& u_con, const Vec& Qe, PCoef& pc, Vec& p, const Vec& b, imex_solve_info* info) {
  const int N = C.nx * C.ny;
  const double h = S.adt, dx = C.dx, dy = C.dy;
  const double a = 1.0 / (C.P.gamma - 1.0);
  Vec g(N), dp(N), neg_g(N), Lp(N), Tdiag(N), p_freeze;
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double h00 = A(pc.hhat, C, i, j), hp0 = A(pc.hhat, C, i + 1, j), hm0 = A(pc.hhat, C, i - 1, j);
      double h0p = A(pc.hhat, C, i, j + 1), h0m = A(pc.hhat, C, i, j - 1);
      Tdiag[M_P(C, i, j)] = h * h * ((0.5 * (hp0 + h00) + 0.5 * (hm0 + h00)) / (dx * dx) + (0.5 * (h0p + h00) + 0.5 * (h0m + h00)) / (dy * dy));
    }
  bool jac = false;
  std::function<void(double*, const double*)> precP = [&](double* out, const double* in) {
    #pragma omp parallel for schedule(static)
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++) {
        int idx = M_P(C, i, j);
        if (dir_bnd(C, i, j)) { out[idx] = in[idx]; continue; }     // identity row (:652-655)
        double Vd = jac ? V_slope(C, p_freeze[idx], G(u_con, C, i, j, 0)) : a;
        out[idx] = in[idx] / (Vd + Tdiag[idx]);
      }
  };
  std::function<void(const double*, double*)> Aop = [&](const double* xin, double* out) {
    p_apply(C, S, u_con, Qe, pc, xin, out, jac ? 1 : 0, p_freeze.data());
  };
  int rc = 0;
  if (info) { info->iters = 0; info->restarts = 0; info->status = 0; }
  if (!C.P.floor_on && C.P.eos == IMEX_EOS_IDEAL) {
    jac = false;
    rc |= gmres(C, p, b, Aop, &precP, nullptr, &C.st.p_gmres_iters, &C.st.p_gmres_restarts);
    C.st.p_solves++;
    return rc;
  }
  if (C.P.floor_on) {
    double bmin = *std::min_element(b.begin(), b.end());
    double Vpe = C.P.p_floor / (C.P.gamma - 1.0);
    if (C.P.eos != IMEX_EOS_IDEAL) {
      for (int i = GC; i < C.nx + GC; i++)
        for (int j = GC; j < C.ny + GC; j++) {
          double v = Vbase(C, C.P.p_floor, G(u_con, C, i, j, 0));
          if (std::isfinite(v) && v > Vpe) Vpe = v;
        }
    }
    double M_cor = std::max(0.0, Vpe - bmin) / ((1.0 - C.P.theta) * C.P.p_floor);
    double M_base = C.P.mbig_fac / (C.P.gamma - 1.0);
    C.Mbig = std::max(M_base, M_cor);
  }
  p_freeze = p;
  int it = 0;
  double res0 = -1.0;
  for (it = 0; it < C.P.newton_max; it++) {
    jac = false;
    Aop(p_freeze.data(), Lp.data());
    #pragma omp parallel for schedule(static)
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++) {
        int idx = M_P(C, i, j);
        if (dir_bnd(C, i, j)) g[idx] = Lp[idx] - b[idx];             // identity row: p - p_amb (:697-699)
        else g[idx] = Lp[idx] - a * p_freeze[idx] + V_mod(C, p_freeze[idx], G(u_con, C, i, j, 0)) - b[idx];
      }
    double res = normv(g);
    if (it == 0) res0 = res;
    if (info) { info->res = res; if (it == 0) info->res0 = res; }
    if (res < C.P.newton_tol * res0 + C.P.newton_tol) break;
    for (int n = 0; n < N; n++) { neg_g[n] = -g[n]; dp[n] = 0.0; }
    jac = true;
    rc |= gmres(C, dp, neg_g, Aop, &precP, nullptr, &C.st.p_gmres_iters, &C.st.p_gmres_restarts);
    C.st.p_solves++;
    for (int n = 0; n < N; n++) p_freeze[n] += dp[n];
  }
  C.st.newton_iters += it;
  if (info) info->iters = it;
  if (it == C.P.newton_max) { rc |= IMEX_WARN_NEWTON_MAXITER; C.st.warnings |= IMEX_WARN_NEWTON_MAXITER; }
  p = p_freeze;
  if (C.P.floor_on) {
    int nfl = 0; double pmin = 1e300;
    for (int n = 0; n < N; n++) { if (p[n] < C.P.p_floor) ++nfl; if (p[n] < pmin) pmin = p[n]; }
    if (nfl > C.st.max_floored) C.st.max_floored = nfl;
    if (pmin < C.st.min_p) C.st.min_p = pmin;
    if (pmin < C.P.theta * C.P.p_floor - 1e-14) { rc |= IMEX_WARN_POSITIVITY; C.st.warnings |= IMEX_WARN_POSITIVITY; }
  }
  if (info) info->status = rc;
  return rc;
}

// ---------------------------------------------------------------- magnetic subsystem (imex_subsystems.cpp:537-628)
void collect_mag_fluxes(Ctx& C, const imex_scalars& S, const Vec& u, const Vec& Qe, const Vec& pe, Vec& Fx, Vec& Fy) {
  const double h = S.adt, dx = C.dx, dy = C.dy;
  Fx.assign(C.NX * C.NY, 0.0); Fy.assign(C.NX * C.NY, 0.0);
  auto Q = [&](int a, int b, int k) { return G(Qe, C, a, b, k); };
  auto U = [&](int a, int b, int k) { return G(u, C, a, b, k); };
  auto PE = [&](int a, int b) { return A(pe, C, a, b); };
  #pragma omp parallel for schedule(static)
  for (int i = 1; i < C.NX - 1; i++)
    for (int j = 1; j < C.NY - 1; j++) {       // ct_flux.cpp:14-127
      double rho_p0 = 0.5 * (U(i, j, 0) + U(i + 1, j, 0)), rho_0p = 0.5 * (U(i, j, 0) + U(i, j + 1, 0));
      double rvx_p0 = 0.5 * (U(i, j, 1) + U(i + 1, j, 1)), rvy_p0 = 0.5 * (U(i, j, 2) + U(i + 1, j, 2));
      double rvx_0p = 0.5 * (U(i, j, 1) + U(i, j + 1, 1)), rvy_0p = 0.5 * (U(i, j, 2) + U(i, j + 1, 2));
      double p_pp = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j + 1) + PE(i + 1, j + 1));
      double p_pm = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j - 1) + PE(i + 1, j - 1));
      double p_mp = 0.25 * (PE(i, j) + PE(i - 1, j) + PE(i, j + 1) + PE(i - 1, j + 1));
      double Bxn_p0 = 0.5 * (Q(i, j, 5) + Q(i + 1, j, 5)), Byn_p0 = 0.5 * (Q(i, j, 6) + Q(i + 1, j, 6));
      double Bxn_0p = 0.5 * (Q(i, j, 5) + Q(i, j + 1, 5)), Byn_0p = 0.5 * (Q(i, j, 6) + Q(i, j + 1, 6));
      double n_pp[3], n_pm[3], n_mp[3], b_pp[3], b_pm[3], b_mp[3];
      for (int c = 0; c < 3; c++) {
        int k = 5 + c;
        n_pp[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j + 1, k) + Q(i + 1, j + 1, k));
        n_pm[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j - 1, k) + Q(i + 1, j - 1, k));
        n_mp[c] = 0.25 * (Q(i, j, k) + Q(i - 1, j, k) + Q(i, j + 1, k) + Q(i - 1, j + 1, k));
        b_pp[c] = 0.25 * (U(i, j, k) + U(i + 1, j, k) + U(i, j + 1, k) + U(i + 1, j + 1, k));
        b_pm[c] = 0.25 * (U(i, j, k) + U(i + 1, j, k) + U(i, j - 1, k) + U(i + 1, j - 1, k));
        b_mp[c] = 0.25 * (U(i, j, k) + U(i - 1, j, k) + U(i, j + 1, k) + U(i - 1, j + 1, k));
      }
      auto PM = [&](int a, int b) { return 0.5 * (Q(a, b, 5) * U(a, b, 5) + Q(a, b, 6) * U(a, b, 6) + Q(a, b, 7) * U(a, b, 7)); };
      auto PC = [&](const double* n, const double* bb, int m) { return 0.5 * (n[0] * bb[0] + n[1] * bb[1] + n[2] * bb[2]) - n[m] * bb[m]; };
      // x face (i+1/2, j): flux of By
      double Fn_x = Byn_p0 / rho_p0 * (rvx_p0 - (h / dx) * (PE(i + 1, j) - PE(i, j))) - Bxn_p0 / rho_p0 * (rvy_p0 - (h / dy) * (p_pp - p_pm));
      double By_1 = Byn_p0 / rho_p0 * (-1.0 / dx * ((PM(i + 1, j) - Q(i + 1, j, 5) * U(i + 1, j, 5)) - (PM(i, j) - Q(i, j, 5) * U(i, j, 5))) + 1.0 / dy * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
      double By_2 = Bxn_p0 / rho_p0 * (-1.0 / dx * (Q(i + 1, j, 5) * U(i + 1, j, 6) - Q(i, j, 5) * U(i, j, 6)) + 1.0 / dy * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
      double nu_x = S.B_max_x * dx;
      double Fdiff_x = -nu_x * (U(i + 1, j, 6) - U(i, j, 6)) / dx;
      A(Fx, C, i, j) = Fn_x + h * (By_1 + By_2) + Fdiff_x;
      // y face (i, j+1/2): flux of Bx
      double Fn_y = Bxn_0p / rho_0p * (rvy_0p - (h / dy) * (PE(i, j + 1) - PE(i, j))) - Byn_0p / rho_0p * (rvx_0p - (h / dx) * (p_pp - p_mp));
      double Bx_1 = Bxn_0p / rho_0p * (1.0 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1.0 / dy * ((PM(i, j + 1) - Q(i, j + 1, 6) * U(i, j + 1, 6)) - (PM(i, j) - Q(i, j, 6) * U(i, j, 6))));
      double Bx_2 = Byn_0p / rho_0p * (1.0 / dx * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1.0 / dy * (Q(i, j + 1, 6) * U(i, j + 1, 5) - Q(i, j, 6) * U(i, j, 5)));
      double nu_y = S.B_max_y * dy;
      double Fdiff_y = -nu_y * (U(i, j + 1, 5) - U(i, j, 5)) / dy;
      A(Fy, C, i, j) = Fn_y + h * (Bx_1 + Bx_2) + Fdiff_y;
    }
  bc_array(C, Fx); bc_array(C, Fy);
}

int mag_update_b(Ctx& C, const imex_scalars& S, Stage& W) {
  const double h = S.adt, dx = C.dx, dy = C.dy;
  const int N = C.nx * C.ny;
  Vec pri(W.Qe.size());
  grid_con2pri(C, pri, W.Qe);
  Vec pe(C.NX * C.NY);
  for (int j = 0; j < C.NY; j++) for (int i = 0; i < C.NX; i++) A(pe, C, i, j) = G(pri, C, i, j, 4);
  Vec b(3 * N), x(3 * N);
  const Vec& u = W.u_con;
  auto Q = [&](int a, int bb, int k) { return G(W.Qe, C, a, bb, k); };
  auto U = [&](int a, int bb, int k) { return G(u, C, a, bb, k); };
  auto PE = [&](int a, int bb) { return A(pe, C, a, bb); };
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {      // imex_subsystems.cpp:549-613
      double rho_p0 = 0.5 * (U(i, j, 0) + U(i + 1, j, 0)), rho_m0 = 0.5 * (U(i, j, 0) + U(i - 1, j, 0));
      double rho_0p = 0.5 * (U(i, j, 0) + U(i, j + 1, 0)), rho_0m = 0.5 * (U(i, j, 0) + U(i, j - 1, 0));
      double m[3][4];   // [component][p0, m0, 0p, 0m] face averages of rho v*
      double n[3][4];   // face averages of B^n (Qe)
      for (int c = 0; c < 3; c++) {
        m[c][0] = 0.5 * (U(i, j, 1 + c) + U(i + 1, j, 1 + c)); m[c][1] = 0.5 * (U(i, j, 1 + c) + U(i - 1, j, 1 + c));
        m[c][2] = 0.5 * (U(i, j, 1 + c) + U(i, j + 1, 1 + c)); m[c][3] = 0.5 * (U(i, j, 1 + c) + U(i, j - 1, 1 + c));
        n[c][0] = 0.5 * (Q(i, j, 5 + c) + Q(i + 1, j, 5 + c)); n[c][1] = 0.5 * (Q(i, j, 5 + c) + Q(i - 1, j, 5 + c));
        n[c][2] = 0.5 * (Q(i, j, 5 + c) + Q(i, j + 1, 5 + c)); n[c][3] = 0.5 * (Q(i, j, 5 + c) + Q(i, j - 1, 5 + c));
      }
      double p_pp = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j + 1) + PE(i + 1, j + 1));
      double p_pm = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j - 1) + PE(i + 1, j - 1));
      double p_mp = 0.25 * (PE(i, j) + PE(i - 1, j) + PE(i, j + 1) + PE(i - 1, j + 1));
      double p_mm = 0.25 * (PE(i, j) + PE(i - 1, j) + PE(i, j - 1) + PE(i - 1, j - 1));
      double Bx_star = G(W.Qs, C, i, j, 5) - (h / dy) * ((n[0][2] / rho_0p * (m[1][2] - (h / dy) * (PE(i, j + 1) - PE(i, j))) - n[1][2] / rho_0p * (m[0][2] - h / dx * (p_pp - p_mp)))
                                                 - (n[0][3] / rho_0m * (m[1][3] - (h / dy) * (PE(i, j) - PE(i, j - 1))) - n[1][3] / rho_0m * (m[0][3] - h / dx * (p_pm - p_mm))));
      double By_star = G(W.Qs, C, i, j, 6) - (h / dx) * ((n[1][0] / rho_p0 * (m[0][0] - (h / dx) * (PE(i + 1, j) - PE(i, j))) - n[0][0] / rho_p0 * (m[1][0] - h / dy * (p_pp - p_pm)))
                                                       - (n[1][1] / rho_m0 * (m[0][1] - (h / dx) * (PE(i, j) - PE(i - 1, j))) - n[0][1] / rho_m0 * (m[1][1] - h / dy * (p_mp - p_mm))));
      double Bz_1 = ((n[2][0] / rho_p0 * (m[0][0] - h / dx * (PE(i + 1, j) - PE(i, j))) - n[0][0] / rho_p0 * (m[2][0]))
                   - (n[2][1] / rho_m0 * (m[0][1] - h / dx * (PE(i, j) - PE(i - 1, j))) - n[0][1] / rho_m0 * (m[2][1])));
      double Bz_2 = ((n[2][2] / rho_0p * (m[1][2] - h / dy * (PE(i, j + 1) - PE(i, j))) - n[1][2] / rho_0p * (m[2][2]))
                   - (n[2][3] / rho_0m * (m[1][3] - h / dy * (PE(i, j) - PE(i, j - 1))) - n[1][3] / rho_0m * (m[2][3])));
      double Bz_star = G(W.Qs, C, i, j, 7) - (h / dx) * Bz_1 - (h / dy) * Bz_2;
      int idx = (i - GC) + C.nx * (j - GC);
      b[idx] = Bx_star; b[idx + N] = By_star; b[idx + 2 * N] = Bz_star;
      x[idx] = Q(i, j, 5); x[idx + N] = Q(i, j, 6); x[idx + 2 * N] = Q(i, j, 7);
    }
  std::function<void(const double*, double*)> Aop = [&](const double* xin, double* out) { mag_apply(C, S, W.u_con, W.Qe, xin, out); };
  int rc = gmres(C, x, b, Aop, nullptr, nullptr, &C.st.mag_gmres_iters, &C.st.mag_gmres_restarts);
  C.st.mag_solves++;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      int idx = (i - GC) + C.nx * (j - GC);
      G(W.u_con, C, i, j, 5) = x[idx]; G(W.u_con, C, i, j, 6) = x[idx + N]; G(W.u_con, C, i, j, 7) = x[idx + 2 * N];
    }
  bc_grid(C, W.u_con);
  collect_mag_fluxes(C, S, W.u_con, W.Qe, pe, W.fx, W.fy);
  return rc;
}

// ---------------------------------------------------------------- pressure subsystem (imex_subsystems.cpp:76-302)
int pre_subsystem_pressure(Ctx& C, const imex_scalars& S, Stage& W) {
  const double h = S.adt, dx = C.dx, dy = C.dy;
  const int N = C.nx * C.ny;
  bc_grid(C, W.u_con);
  Vec mx(C.NX * C.NY, 0.0), my(C.NX * C.NY, 0.0), mz(C.NX * C.NY, 0.0);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      A(mx, C, i, j) = G(W.u_con, C, i, j, 1); A(my, C, i, j) = G(W.u_con, C, i, j, 2); A(mz, C, i, j) = G(W.u_con, C, i, j, 3);
    }
  bc_array(C, mx); bc_array(C, my); bc_array(C, mz);
  if (jet_on(C))                                                      // the inflow momentum in the strip (:107-118)
    for (int i = 0; i < C.NX; i++) {
      if (!jet_inlet(C, i)) continue;
      for (int g = 0; g < GC; g++) { A(mx, C, i, g) = 0.0; A(my, C, i, g) = C.P.jet_rho * C.P.jet_v; A(mz, C, i, g) = 0.0; }
    }
  PCoef pc;
  p_coef(C, W.u_con, W.Qe, pc);
  Vec b(N), p(N);
  auto rhs = [&](int pass) {
    #pragma omp parallel for schedule(static)
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++) {
        const Vec& u = W.u_con;
        double B2 = G(u, C, i, j, 5) * G(u, C, i, j, 5) + G(u, C, i, j, 6) * G(u, C, i, j, 6) + G(u, C, i, j, 7) * G(u, C, i, j, 7);
        double h00 = A(pc.hhat, C, i, j), hp0 = A(pc.hhat, C, i + 1, j), hm0 = A(pc.hhat, C, i - 1, j);
        double h0p = A(pc.hhat, C, i, j + 1), h0m = A(pc.hhat, C, i, j - 1);
        double p_star = G(u, C, i, j, 4) - 0.5 * B2;
        double ke;
        if (pass == 0) ke = 0.5 * (A(mx, C, i, j) * A(mx, C, i, j) + A(my, C, i, j) * A(my, C, i, j) + A(mz, C, i, j) * A(mz, C, i, j)) / G(u, C, i, j, 0);
        else ke = 0.5 * (G(u, C, i, j, 1) * G(u, C, i, j, 1) + G(u, C, i, j, 2) * G(u, C, i, j, 2) + G(u, C, i, j, 3) * G(u, C, i, j, 3)) / G(u, C, i, j, 0);
        double RHS = p_star - ke - h * (1 / dx * (0.5 * (hp0 + h00) * 0.5 * (A(mx, C, i + 1, j) + A(mx, C, i, j)) - 0.5 * (hm0 + h00) * 0.5 * (A(mx, C, i - 1, j) + A(mx, C, i, j)))
                                        + (1 / dy * (0.5 * (h0p + h00) * 0.5 * (A(my, C, i, j + 1) + A(my, C, i, j)) - 0.5 * (h0m + h00) * 0.5 * (A(my, C, i, j - 1) + A(my, C, i, j)))));
        b[M_P(C, i, j)] = RHS;
        if (dir_bnd(C, i, j)) b[M_P(C, i, j)] = C.P.p_amb;          // strong Dirichlet (:160-161, 263-264)
        if (j == GC && jet_inlet(C, i)) b[M_P(C, i, j)] += h * h / (dy * dy) * 0.5 * (h00 + h0m) * C.P.jet_p;   // lifting (:163-165, 266-267)
        if (pass == 0) p[M_P(C, i, j)] = A(pc.pe, C, i, j);
      }
  };
  rhs(0);
  int rc = pressure_newton(C, S, W.u_con, W.Qe, pc, p, b, nullptr);
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {      // momentum update :186-215
      double flux_xm = 0.5 * (p[M_P(C, i - 1, j)] + p[M_P(C, i, j)]);
      double flux_xp = 0.5 * (p[M_P(C, i, j)] + p[M_P(C, i + 1, j)]);
      G(W.u_con, C, i, j, 1) = A(mx, C, i, j) - h / dx * (flux_xp - flux_xm);
      double gL = (j == GC && jet_inlet(C, i)) ? C.P.jet_p : p[M_P(C, i, j - 1)];   // jet strip ghost p (:202-203)
      double flux_ym = 0.5 * (gL + p[M_P(C, i, j)]);
      double flux_yp = 0.5 * (p[M_P(C, i, j)] + p[M_P(C, i, j + 1)]);
      G(W.u_con, C, i, j, 2) = A(my, C, i, j) - h / dy * (flux_yp - flux_ym);
    }
  bc_grid(C, W.u_con);
  rhs(1);                                       // all_correction second pass :223-284
  rc |= pressure_newton(C, S, W.u_con, W.Qe, pc, p, b, nullptr);
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {      // DefinitionofU :286-302
      Vec& u = W.u_con;
      G(u, C, i, j, 4) = Vbase(C, p[M_P(C, i, j)], G(u, C, i, j, 0))
                       + 0.5 * (G(u, C, i, j, 1) * G(u, C, i, j, 1) + G(u, C, i, j, 2) * G(u, C, i, j, 2) + G(u, C, i, j, 3) * G(u, C, i, j, 3)) / G(u, C, i, j, 0)
                       + 0.5 * (G(u, C, i, j, 5) * G(u, C, i, j, 5) + G(u, C, i, j, 6) * G(u, C, i, j, 6) + G(u, C, i, j, 7) * G(u, C, i, j, 7));
    }
  bc_grid(C, W.u_con);
  return rc;
}

// ---------------------------------------------------------------- energy subsystem (imex_subsystems.cpp:415-534, implicit_operators.cpp:741-790)
inline double caltotalenthalpy(Ctx& C, const double* con) {        // timestep.cpp:232-237
  double pri[NV];
  con2pri(C, pri, con);
  return (con[4] + pri[4]) / pri[0];
}
// h = caltotalenthalpy(Qe) rho^e / rho on every cell (the reference evaluates it at the cell and
// its four neighbours, ghost cells included, inside each use)
void e_coef(Ctx& C, const Vec& u_con, const Vec& Qe, Vec& hE) {
  hE.assign(C.NX * C.NY, 0.0);
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++)
      A(hE, C, i, j) = caltotalenthalpy(C, &Qe[NV * (i + j * C.NX)]) * G(Qe, C, i, j, 0) / G(u_con, C, i, j, 0);
}
// Energy_Subsystem_AX (implicit_operators.cpp:741-790), periodic / transmissive getE through M_P.
void e_apply(Ctx& C, const imex_scalars& S, const Vec& u_con, const Vec& Qe, Vec& hE, const double* x, double* Ax) {
  if (C.faithful) e_coef(C, u_con, Qe, hE);   // the reference recomputes these inside every apply (:742-767)
  const double h = S.adt, dx = C.dx, dy = C.dy, g = C.P.gamma;
  auto X = [&](int a, int b) { return x[M_P(C, a, b)]; };
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double h00 = A(hE, C, i, j), hp0 = A(hE, C, i + 1, j), hm0 = A(hE, C, i - 1, j);
      double h0p = A(hE, C, i, j + 1), h0m = A(hE, C, i, j - 1);
      double hhp_x = 0.5 * (h00 + hp0), hhm_x = 0.5 * (h00 + hm0);
      double hhp_y = 0.5 * (h00 + h0p), hhm_y = 0.5 * (h00 + h0m);
      double Ei = X(i, j);
      // getE (:772-779): Dirichlet ghosts and the jet strip's ghost are homogeneous (zero)
      auto GE = [&](int a, int b) {
        bool xg = a < GC || a >= C.nx + GC, yg = b < GC || b >= C.ny + GC;
        if ((dir_x(C) && xg) || (dir_y(C) && yg) || (b < GC && jet_inlet(C, a))) return 0.0;
        return X(a, b);
      };
      double ax = Ei - (g - 1) * h * h * (1.0 / (dx * dx) * (hhp_x * (GE(i + 1, j) - Ei) - hhm_x * (Ei - GE(i - 1, j)))
                                        + 1.0 / (dy * dy) * (hhp_y * (GE(i, j + 1) - Ei) - hhm_y * (Ei - GE(i, j - 1))));
      Ax[M_P(C, i, j)] = ax * dx * dy;
    }
  C.st.op_applies_p++;
}
// precE (imex_subsystems.cpp:479-500): diag (dx dy) with diag = 1 + (gamma-1) h^2 (sum of face coefficients / d^2)
void e_diag(Ctx& C, const imex_scalars& S, const Vec& hE, Vec& D) {
  const double h = S.adt, dx = C.dx, dy = C.dy, g = C.P.gamma;
  D.assign(C.nx * C.ny, 0.0);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double h00 = A(hE, C, i, j), hp0 = A(hE, C, i + 1, j), hm0 = A(hE, C, i - 1, j);
      double h0p = A(hE, C, i, j + 1), h0m = A(hE, C, i, j - 1);
      double hhp_x = 0.5 * (h00 + hp0), hhm_x = 0.5 * (h00 + hm0);
      double hhp_y = 0.5 * (h00 + h0p), hhm_y = 0.5 * (h00 + h0m);
      double diag = 1.0 + (g - 1.0) * h * h * ((hhp_x + hhm_x) / (dx * dx) + (hhp_y + hhm_y) / (dy * dy));
      D[M_P(C, i, j)] = diag * dx * dy;
    }
}

int pre_subsystem_energy(Ctx& C, const imex_scalars& S, Stage& W) {
  const double h = S.adt, dx = C.dx, dy = C.dy, g = C.P.gamma;
  const int N = C.nx * C.ny;
  Vec hE;
  e_coef(C, W.u_con, W.Qe, hE);
  Vec b(N), U(N), D;
  const Vec& u = W.u_con;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {       // :427-452, 472
      double hp0 = A(hE, C, i + 1, j), hm0 = A(hE, C, i - 1, j), h0p = A(hE, C, i, j + 1), h0m = A(hE, C, i, j - 1);
      double U_star = G(u, C, i, j, 4) - 0.5 * h / dx * (hp0 * G(u, C, i + 1, j, 1) - hm0 * G(u, C, i - 1, j, 1))
                    - 0.5 * h / dy * (h0p * G(u, C, i, j + 1, 2) - h0m * G(u, C, i, j - 1, 2));
      b[M_P(C, i, j)] = U_star * dx * dy;
      if (jet_on(C) || dir_x(C) || dir_y(C)) {                        // Dirichlet lifting (:453-475): getEc
        double h00 = A(hE, C, i, j);
        double hhp_x = 0.5 * (h00 + hp0), hhm_x = 0.5 * (h00 + hm0), hhp_y = 0.5 * (h00 + h0p), hhm_y = 0.5 * (h00 + h0m);
        auto getEc = [&](int ii, int jj) {
          bool xg = ii < GC || ii >= C.nx + GC, yg = jj < GC || jj >= C.ny + GC;
          if (jj < GC && jet_inlet(C, ii)) return jet_E_val(C);
          if (!((dir_x(C) && xg) || (dir_y(C) && yg))) return 0.0;
          return C.P.p_amb / (C.P.gamma - 1.0) + 0.5 * (G(u, C, ii, jj, 5) * G(u, C, ii, jj, 5) + G(u, C, ii, jj, 6) * G(u, C, ii, jj, 6) +
                                                       G(u, C, ii, jj, 7) * G(u, C, ii, jj, 7));
        };
        double L_const = -(g - 1) * h * h * (1.0 / (dx * dx) * (hhp_x * getEc(i + 1, j) + hhm_x * getEc(i - 1, j))
                                           + 1.0 / (dy * dy) * (hhp_y * getEc(i, j + 1) + hhm_y * getEc(i, j - 1)));
        L_const *= dx * dy;
        b[M_P(C, i, j)] -= L_const;
      }
      U[M_P(C, i, j)] = G(u, C, i, j, 4);
    }
  e_diag(C, S, hE, D);
  std::function<void(const double*, double*)> Aop = [&](const double* xin, double* out) { e_apply(C, S, W.u_con, W.Qe, hE, xin, out); };
  std::function<void(double*, const double*)> precE = [&](double* out, const double* in) {
    if (C.faithful) { Vec hf; e_coef(C, W.u_con, W.Qe, hf); }   // the reference's precE recomputes h (:481-492)
    for (int n = 0; n < N; n++) out[n] = in[n] / D[n];
  };
  int rc = gmres(C, U, b, Aop, &precE, nullptr, &C.st.p_gmres_iters, &C.st.p_gmres_restarts);
  C.st.p_solves++;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) G(W.u_con, C, i, j, 4) = U[M_P(C, i, j)];
  bc_grid(C, W.u_con);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {       // :514-519
      Vec& v = W.u_con;
      G(v, C, i, j, 1) -= h * (1 / (2 * dx)) * (g - 1) * (G(v, C, i + 1, j, 4) - G(v, C, i - 1, j, 4));
      G(v, C, i, j, 2) -= h * (1 / (2 * dy)) * (g - 1) * (G(v, C, i, j + 1, 4) - G(v, C, i, j - 1, 4));
    }
  bc_grid(C, W.u_con);
  return rc;
}

// ---------------------------------------------------------------- CT (ct_flux.cpp:150-191, constrained_transport.cpp:66-164)
void corner_assemble(const Ctx& C, Vec& Ez_pp, const Vec& Ez_p0, const Vec& Ez_0p, const Vec& Et) {
  Vec bar(C.NX * C.NY, 0.0);
  for (int j = 0; j < C.NY - 2; j++)
    for (int i = 0; i < C.NX - 2; i++)
      A(bar, C, i, j) = 0.25 * (A(Ez_p0, C, i, j) + A(Ez_p0, C, i, j + 1) + A(Ez_0p, C, i, j) + A(Ez_0p, C, i + 1, j));
  bc_array(C, bar);
  const double dx = C.dx, dy = C.dy;
  #pragma omp parallel for schedule(static)
  for (int j = 1; j < C.NY - 2; j++)
    for (int i = 1; i < C.NX - 2; i++) {
      double t1 = 0.5 * ((A(Ez_p0, C, i, j) - A(Et, C, i, j)) / (dx / 2.0) + (A(Ez_p0, C, i, j + 1) - A(Et, C, i, j + 1)) / (dx / 2.0));
      double t2 = 0.5 * ((A(Et, C, i + 1, j) - A(Ez_p0, C, i, j)) / (dx / 2.0) + (A(Et, C, i + 1, j + 1) - A(Ez_p0, C, i, j + 1)) / (dx / 2.0));
      double t3 = 0.5 * ((A(Ez_0p, C, i, j) - A(Et, C, i, j)) / (dy / 2.0) + (A(Ez_0p, C, i + 1, j) - A(Et, C, i + 1, j)) / (dy / 2.0));
      double t4 = 0.5 * ((A(Et, C, i, j + 1) - A(Ez_0p, C, i, j)) / (dy / 2.0) + (A(Et, C, i + 1, j + 1) - A(Ez_0p, C, i + 1, j)) / (dy / 2.0));
      A(Ez_pp, C, i, j) = A(bar, C, i, j) + dx / 8.0 * (t1 - t2) + dy / 8.0 * (t3 - t4);
    }
}

void ct_face_update(Ctx& C, const imex_scalars& S, Stage& W, const Vec& Ez_pp) {
  const double h = S.adt;
  for (int j = GC - 1; j < C.ny + GC; j++)
    for (int i = GC - 1; i < C.nx + GC; i++) {
      A(W.bx, C, i, j) = A(W.bx_qs, C, i, j) - h / C.dy * (A(Ez_pp, C, i, j) - A(Ez_pp, C, i, j - 1));
      A(W.by, C, i, j) = A(W.by_qs, C, i, j) + h / C.dx * (A(Ez_pp, C, i, j) - A(Ez_pp, C, i - 1, j));
    }
  bc_staggered(C, W.bx, 0);
  bc_staggered(C, W.by, 1);
  for (int j = GC; j < C.ny + GC; j++)
    for (int i = GC; i < C.nx + GC; i++) {
      G(W.u_con, C, i, j, 5) = 0.5 * (A(W.bx, C, i - 1, j) + A(W.bx, C, i, j));
      G(W.u_con, C, i, j, 6) = 0.5 * (A(W.by, C, i, j - 1) + A(W.by, C, i, j));
    }
  bc_grid(C, W.u_con);
}

void ct_consistent(Ctx& C, const imex_scalars& S, Stage& W) {
  Vec Ez_p0(C.NX * C.NY), Ez_0p(C.NX * C.NY), Et(C.NX * C.NY), Ez_pp(C.NX * C.NY, 0.0);
  for (int n = 0; n < C.NX * C.NY; n++) { Ez_p0[n] = -W.fx[n]; Ez_0p[n] = +W.fy[n]; }
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++)
      A(Et, C, i, j) = G(W.u_con, C, i, j, 2) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 5)
                     - G(W.u_con, C, i, j, 1) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 6);
  corner_assemble(C, Ez_pp, Ez_p0, Ez_0p, Et);
  ct_face_update(C, S, W, Ez_pp);
}

// Plain Staggered_CT (constrained_transport.cpp:66-143): central face EMFs from reconstructed u.
void ct_plain(Ctx& C, const imex_scalars& S, Stage& W) {
  Recon w;
  tvd_reconstruction(C, W.u_con, w);
  Vec Ez_p0(C.NX * C.NY, 0.0), Ez_0p(C.NX * C.NY, 0.0), Et(C.NX * C.NY), Ez_pp(C.NX * C.NY, 0.0), bar(C.NX * C.NY, 0.0);
  auto ez_x = [&](const double* Q) {  // f_full_con[6] = By vx - vy Bx (state.cpp:152-177)
    double vx = Q[1] / Q[0], vy = Q[2] / Q[0];
    return Q[6] * vx - vy * Q[5];
  };
  auto ez_y = [&](const double* Q) {  // g_full_con[5] = Bx vy - vx By (state.cpp:179-203)
    double vx = Q[1] / Q[0], vy = Q[2] / Q[0];
    return Q[5] * vy - vx * Q[6];
  };
  for (int i = 0; i < C.NX - 1; i++)
    for (int j = 0; j < C.NY - 1; j++) {
      A(Ez_p0, C, i, j) = -(0.5 * (ez_x(&w.xR[NV * (i + j * C.NX)]) + ez_x(&w.xL[NV * (i + 1 + j * C.NX)])));
      A(Ez_0p, C, i, j) = (0.5 * (ez_y(&w.yR[NV * (i + j * C.NX)]) + ez_y(&w.yL[NV * (i + (j + 1) * C.NX)])));
    }
  bc_array(C, Ez_p0); bc_array(C, Ez_0p);
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++)
      A(Et, C, i, j) = (G(W.u_con, C, i, j, 2) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 5)
                      - G(W.u_con, C, i, j, 1) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 6));
  corner_assemble(C, Ez_pp, Ez_p0, Ez_0p, Et);
  ct_face_update(C, S, W, Ez_pp);
}

// ---------------------------------------------------------------- vector potential (UnStaggerCT, PotentialA)
// bound_cond_A (boundary.cpp:177-292), periodic and transmissive branches: x first over every row,
// then y over every column; transmissive ghosts are linear extrapolations of the two edge cells.
void bc_A(const Ctx& C, Vec& Av) {
  auto Ag = [&](int i, int j, int k) -> double& { return Av[NV * (i + j * C.NX) + k]; };
  for (int j = 0; j < C.NY; j++)
    for (int k = 0; k < 3; k++) {
      if (C.P.bc_x == IMEX_BC_PERIODIC) {
        for (int g = 0; g < GC; g++) { Ag(g, j, k) = Ag(C.nx + g, j, k); Ag(C.nx + GC + g, j, k) = Ag(GC + g, j, k); }
      } else {
        double sL = Ag(GC + 1, j, k) - Ag(GC, j, k);
        for (int g = 0; g < GC; g++) Ag(GC - 1 - g, j, k) = Ag(GC, j, k) - (g + 1) * sL;
        int iR = C.nx + GC - 1;
        double sR = Ag(iR, j, k) - Ag(iR - 1, j, k);
        for (int g = 0; g < GC; g++) Ag(iR + 1 + g, j, k) = Ag(iR, j, k) + (g + 1) * sR;
      }
    }
  for (int i = 0; i < C.NX; i++)
    for (int k = 0; k < 3; k++) {
      if (C.P.bc_y == IMEX_BC_PERIODIC) {
        for (int g = 0; g < GC; g++) { Ag(i, g, k) = Ag(i, C.ny + g, k); Ag(i, C.ny + GC + g, k) = Ag(i, GC + g, k); }
      } else {
        double sL = Ag(i, GC + 1, k) - Ag(i, GC, k);
        for (int g = 0; g < GC; g++) Ag(i, GC - 1 - g, k) = Ag(i, GC, k) - (g + 1) * sL;
        int jR = C.ny + GC - 1;
        double sR = Ag(i, jR, k) - Ag(i, jR - 1, k);
        for (int g = 0; g < GC; g++) Ag(i, jR + 1 + g, k) = Ag(i, jR, k) + (g + 1) * sR;
      }
    }
  if (jet_on(C))                                                      // even reflection in the strip (:279-291)
    for (int i = 0; i < C.NX; i++) {
      if (!jet_inlet(C, i)) continue;
      for (int k = 0; k < 3; k++)
        for (int g = 0; g < GC; g++) Ag(i, GC - 1 - g, k) = Ag(i, GC + g, k);
    }
}

// GLF viscosity of the potential operator: directional maxima of the magnetic eigenvalues of Qs over
// the interior (implicit_operators.cpp:359-374)
void pot_visc(Ctx& C, const Vec& Qs, double& gx, double& gy) {
  Vec pri(Qs.size());
  grid_con2pri(C, pri, Qs);
  gx = 0.0; gy = 0.0;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      const double* q = &pri[NV * (i + j * C.NX)];
      double Bsq = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
      double bxv = 0.5 * (std::fabs(q[1]) + std::sqrt(q[1] * q[1] + 4.0 * Bsq / q[0]));
      double byv = 0.5 * (std::fabs(q[2]) + std::sqrt(q[2] * q[2] + 4.0 * Bsq / q[0]));
      if (bxv > gx) gx = bxv;
      if (byv > gy) gy = byv;
    }
}

// Magnetic_Subsystem_potential_AX (implicit_operators.cpp:355-503), GLF branch with add_bvisc.
void pot_apply(Ctx& C, const imex_scalars& S, const Vec& u_con, const Vec& Qe, const Vec& Qs, double gx, double gy,
               const double* x, double* Ax) {
  if (C.faithful) pot_visc(C, Qs, gx, gy);   // the reference recomputes these inside every apply (:359-374)
  const double h = S.adt, dx = C.dx, dy = C.dy;
  const int N = C.nx * C.ny;
  Vec AG(NV * C.NX * C.NY, 0.0);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++)
      for (int k = 0; k < 3; k++) AG[NV * (i + j * C.NX) + k] = x[(i - GC) + C.nx * (j - GC) + k * N];
  bc_A(C, AG);
  auto a = [&](int ii, int jj, int k) { return AG[NV * (ii + jj * C.NX) + k]; };
  auto Q = [&](int ii, int jj, int k) { return G(Qe, C, ii, jj, k); };
  #pragma omp parallel for schedule(static)
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double rho_00 = G(u_con, C, i, j, 0);
      double bx_P0 = (a(i + 1, j + 1, 2) - a(i + 1, j - 1, 2)) / (2.0 * dy);
      double by_P0 = -(a(i + 2, j, 2) - a(i, j, 2)) / (2.0 * dx);
      double bz_P0 = (a(i + 2, j, 1) - a(i, j, 1)) / (2.0 * dx) - (a(i + 1, j + 1, 0) - a(i + 1, j - 1, 0)) / (2.0 * dy);
      double bx_M0 = (a(i - 1, j + 1, 2) - a(i - 1, j - 1, 2)) / (2.0 * dy);
      double by_M0 = -(a(i, j, 2) - a(i - 2, j, 2)) / (2.0 * dx);
      double bz_M0 = (a(i, j, 1) - a(i - 2, j, 1)) / (2.0 * dx) - (a(i - 1, j + 1, 0) - a(i - 1, j - 1, 0)) / (2.0 * dy);
      double bx_0P = (a(i, j + 2, 2) - a(i, j, 2)) / (2.0 * dy);
      double by_0P = -(a(i + 1, j + 1, 2) - a(i - 1, j + 1, 2)) / (2.0 * dx);
      double bz_0P = (a(i + 1, j + 1, 1) - a(i - 1, j + 1, 1)) / (2.0 * dx) - (a(i, j + 2, 0) - a(i, j, 0)) / (2.0 * dy);
      double bx_0M = (a(i, j, 2) - a(i, j - 2, 2)) / (2.0 * dy);
      double by_0M = -(a(i + 1, j - 1, 2) - a(i - 1, j - 1, 2)) / (2.0 * dx);
      double bz_0M = (a(i + 1, j - 1, 1) - a(i - 1, j - 1, 1)) / (2.0 * dx) - (a(i, j, 0) - a(i, j - 2, 0)) / (2.0 * dy);
      // magnetic stress of the linearised field: T(alpha, beta) = (B^n . b)/2 delta - B^n_beta b_alpha
      double bd_P0 = Q(i + 1, j, 5) * bx_P0 + Q(i + 1, j, 6) * by_P0 + Q(i + 1, j, 7) * bz_P0;
      double Txx_P0 = bd_P0 / 2.0 - Q(i + 1, j, 5) * bx_P0, Tyx_P0 = -Q(i + 1, j, 5) * by_P0, Tzx_P0 = -Q(i + 1, j, 5) * bz_P0;
      double bd_M0 = Q(i - 1, j, 5) * bx_M0 + Q(i - 1, j, 6) * by_M0 + Q(i - 1, j, 7) * bz_M0;
      double Txx_M0 = bd_M0 / 2.0 - Q(i - 1, j, 5) * bx_M0, Tyx_M0 = -Q(i - 1, j, 5) * by_M0, Tzx_M0 = -Q(i - 1, j, 5) * bz_M0;
      double bd_0P = Q(i, j + 1, 5) * bx_0P + Q(i, j + 1, 6) * by_0P + Q(i, j + 1, 7) * bz_0P;
      double Txy_0P = -Q(i, j + 1, 6) * bx_0P, Tyy_0P = bd_0P / 2.0 - Q(i, j + 1, 6) * by_0P, Tzy_0P = -Q(i, j + 1, 6) * bz_0P;
      double bd_0M = Q(i, j - 1, 5) * bx_0M + Q(i, j - 1, 6) * by_0M + Q(i, j - 1, 7) * bz_0M;
      double Txy_0M = -Q(i, j - 1, 6) * bx_0M, Tyy_0M = bd_0M / 2.0 - Q(i, j - 1, 6) * by_0M, Tzy_0M = -Q(i, j - 1, 6) * bz_0M;
      double divT_x = (Txx_P0 - Txx_M0) / (2.0 * dx) + (Txy_0P - Txy_0M) / (2.0 * dy);
      double divT_y = (Tyx_P0 - Tyx_M0) / (2.0 * dx) + (Tyy_0P - Tyy_0M) / (2.0 * dy);
      double divT_z = (Tzx_P0 - Tzx_M0) / (2.0 * dx) + (Tzy_0P - Tzy_0M) / (2.0 * dy);
      double Fx = h / rho_00 * divT_x, Fy = h / rho_00 * divT_y, Fz = h / rho_00 * divT_z;
      double cross[3] = {Q(i, j, 6) * Fz - Q(i, j, 7) * Fy, Q(i, j, 7) * Fx - Q(i, j, 5) * Fz, Q(i, j, 5) * Fy - Q(i, j, 6) * Fx};
      double nu_x = gx * dx, nu_y = gy * dy;
      int idx = (i - GC) + C.nx * (j - GC);
      for (int k = 0; k < 3; k++) {
        double lap = nu_x / (dx * dx) * (a(i - 1, j, k) - 2.0 * a(i, j, k) + a(i + 1, j, k))
                   + nu_y / (dy * dy) * (a(i, j - 1, k) - 2.0 * a(i, j, k) + a(i, j + 1, k));
        Ax[idx + k * N] = a(i, j, k) - h * cross[k] - h * lap;
      }
    }
  C.st.op_applies_mag++;
}

// MagSubsystem_UpdateA (imex_subsystems.cpp:5-74): W.A in (the initial guess, a copy of W.As) and out.
int mag_update_a(Ctx& C, const imex_scalars& S, Stage& W) {
  const double h = S.adt, dx = C.dx, dy = C.dy;
  const int N = C.nx * C.ny;
  Vec pri(W.Qe.size());
  grid_con2pri(C, pri, W.Qe);
  Vec b(3 * N), x(3 * N);
  const Vec& u = W.u_con;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double Bxn = G(W.Qe, C, i, j, 5), Byn = G(W.Qe, C, i, j, 6), Bzn = G(W.Qe, C, i, j, 7);
      double rho = G(u, C, i, j, 0);
      double Sx = G(u, C, i, j, 1) / rho - h / rho * (G(pri, C, i + 1, j, 4) - G(pri, C, i - 1, j, 4)) / (2.0 * dx);
      double Sy = G(u, C, i, j, 2) / rho - h / rho * (G(pri, C, i, j + 1, 4) - G(pri, C, i, j - 1, 4)) / (2.0 * dy);
      double Sz = G(u, C, i, j, 3) / rho;
      int idx = (i - GC) + C.nx * (j - GC);
      b[idx] = G(W.As, C, i, j, 0) - h * (Byn * Sz - Bzn * Sy);
      b[idx + N] = G(W.As, C, i, j, 1) - h * (Bzn * Sx - Bxn * Sz);
      b[idx + 2 * N] = G(W.As, C, i, j, 2) - h * (Bxn * Sy - Byn * Sx);
      for (int k = 0; k < 3; k++) x[idx + k * N] = G(W.A, C, i, j, k);
    }
  double gx, gy;
  pot_visc(C, W.Qs, gx, gy);
  std::function<void(const double*, double*)> Aop = [&](const double* xin, double* out) { pot_apply(C, S, W.u_con, W.Qe, W.Qs, gx, gy, xin, out); };
  int rc = gmres(C, x, b, Aop, nullptr, nullptr, &C.st.mag_gmres_iters, &C.st.mag_gmres_restarts);
  C.st.mag_solves++;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++)
      for (int k = 0; k < 3; k++) G(W.A, C, i, j, k) = x[(i - GC) + C.nx * (j - GC) + k * N];
  bc_A(C, W.A);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      auto a = [&](int ii, int jj, int k) { return G(W.A, C, ii, jj, k); };
      G(W.u_con, C, i, j, 5) = 1 / (2.0 * dy) * (a(i, j + 1, 2) - a(i, j - 1, 2));
      G(W.u_con, C, i, j, 6) = -1 / (2.0 * dx) * (a(i + 1, j, 2) - a(i - 1, j, 2));
      G(W.u_con, C, i, j, 7) = 1 / (2.0 * dx) * (a(i + 1, j, 1) - a(i - 1, j, 1)) - 1 / (2.0 * dy) * (a(i, j + 1, 0) - a(i, j - 1, 0));
    }
  bc_grid(C, W.u_con);
  return rc;
}

// Unstaggered_CT_Consistent (constrained_transport.cpp:502-540): W.A in (components 0, 1 and the
// initial A_z are the stage's copy of W.As) and out.
void unstaggered_ct(Ctx& C, const imex_scalars& S, Stage& W) {
  const double h = S.adt;
  Vec Ez_p0(C.NX * C.NY), Ez_0p(C.NX * C.NY), Et(C.NX * C.NY), Ez_pp(C.NX * C.NY, 0.0);
  for (int n = 0; n < C.NX * C.NY; n++) { Ez_p0[n] = -W.fx[n]; Ez_0p[n] = +W.fy[n]; }
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++)
      A(Et, C, i, j) = G(W.u_con, C, i, j, 2) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 5)
                     - G(W.u_con, C, i, j, 1) / G(W.u_con, C, i, j, 0) * G(W.u_con, C, i, j, 6);
  corner_assemble(C, Ez_pp, Ez_p0, Ez_0p, Et);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double Ez_cell = 0.25 * (A(Ez_pp, C, i, j) + A(Ez_pp, C, i - 1, j) + A(Ez_pp, C, i, j - 1) + A(Ez_pp, C, i - 1, j - 1));
      G(W.A, C, i, j, 2) = G(W.As, C, i, j, 2) - h * Ez_cell;
    }
  bc_A(C, W.A);
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      G(W.u_con, C, i, j, 5) = (G(W.A, C, i, j + 1, 2) - G(W.A, C, i, j - 1, 2)) / (2.0 * C.dy);
      G(W.u_con, C, i, j, 6) = -(G(W.A, C, i + 1, j, 2) - G(W.A, C, i - 1, j, 2)) / (2.0 * C.dx);
    }
  bc_grid(C, W.u_con);
}

double div_b(const Ctx& C, const Vec& bx, const Vec& by);
double div_b_cell(const Ctx& C, const Vec& u);

// ---------------------------------------------------------------- output files and diagnostics
// printVTKHeader / printVTKVariable / printAllVariablesToVTK (output.cpp:127-319); the field loops run
// over padded indices 0..n-1 as the reference's do.
void write_vtk(const Ctx& C, const Vec& u_con, const Vec& u_pri, double t, int nstep, const std::string& file) {
  std::ofstream o(file.c_str());
  o << "# vtk DataFile Version 3.0" << std::endl << "vtk output" << std::endl << "ASCII" << std::endl;
  o << "DATASET RECTILINEAR_GRID" << std::endl << "FIELD FieldData 2" << std::endl << "CYCLE 1 1 int" << std::endl;
  o << nstep << std::endl << "TIME 1 1 double" << std::endl << t << std::endl;
  o << "DIMENSIONS " << C.nx + 1 << " " << C.ny + 1 << " 1" << std::endl;
  o << "X_COORDINATES " << C.nx + 1 << " double" << std::endl;
  for (int i = GC; i < C.nx + GC; i++) o << C.P.x0 + (((double)i) - GC + 0.5) * C.dx << " ";
  o << std::endl << "Y_COORDINATES " << C.ny + 1 << " double" << std::endl;
  for (int j = GC; j < C.ny + GC; j++) o << C.P.y0 + (((double)j) - GC + 0.5) * C.dy << " ";
  o << std::endl << "Z_COORDINATES 1 double" << std::endl << "0" << std::endl;
  o << "CELL_DATA " << C.nx * C.ny << std::endl;
  const char* names[13] = {"rho", "mom_X", "mom_Y", "mom_Z", "U", "B_X", "B_Y", "B_Z", "velX", "velY", "velZ", "p", "B"};
  o << "FIELD FieldData " << 13 << std::endl;
  for (int var = 0; var < 13; var++) {
    o << names[var] << " 1 " << C.nx * C.ny << " double" << std::endl;
    for (int j = 0; j < C.ny; j++) {
      for (int i = 0; i < C.nx; i++) {
        double v;
        if (var < NV) v = G(u_con, C, i, j, var);
        else if (var < NV + 4) v = G(u_pri, C, i, j, var - NV + 1);
        else v = std::sqrt(G(u_con, C, i, j, 5) * G(u_con, C, i, j, 5) + G(u_con, C, i, j, 6) * G(u_con, C, i, j, 6) + G(u_con, C, i, j, 7) * G(u_con, C, i, j, 7));
        o << std::setw(24) << std::setprecision(std::numeric_limits<double>::digits10 + 1) << v << " ";
      }
      o << std::endl;
    }
    o << std::endl;
  }
  o << "POINT_DATA " << (C.nx + 1) * (C.ny + 1) << std::endl;
}

// the end-of-run field files (main.cpp:598-720)
void write_fields(Ctx& C, const Vec& u_con, const Vec& u_pri, const std::string& dir) {
  auto one = [&](const char* name, const std::function<double(int, int)>& f) {
    std::ofstream o((dir + "/" + name).c_str());
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++)
        o << C.P.x0 + C.dx * (i - GC + 0.5) << " " << C.P.y0 + C.dy * (j - GC + 0.5) << " " << f(i, j) << std::endl;
  };
  one("RK_rho.dat", [&](int i, int j) { return G(u_con, C, i, j, 0); });
  one("RK_vx.dat", [&](int i, int j) { return G(u_pri, C, i, j, 1); });
  one("RK_vy.dat", [&](int i, int j) { return G(u_pri, C, i, j, 2); });
  one("RK_E.dat", [&](int i, int j) { return G(u_con, C, i, j, 4); });
  one("RK_p.dat", [&](int i, int j) { return G(u_pri, C, i, j, 4); });
  one("RK_Bx.dat", [&](int i, int j) { return G(u_pri, C, i, j, 5); });
  one("RK_By.dat", [&](int i, int j) { return G(u_pri, C, i, j, 6); });
  one("RK_Bz.dat", [&](int i, int j) { return G(u_pri, C, i, j, 7); });
  one("RK_v_mag.dat", [&](int i, int j) { return std::sqrt(G(u_pri, C, i, j, 1) * G(u_pri, C, i, j, 1) + G(u_pri, C, i, j, 2) * G(u_pri, C, i, j, 2)); });
  if (C.P.eos == IMEX_EOS_RK)
    one("RK_T.dat", [&](int i, int j) { return rk_T_from_rho_p(C, G(u_pri, C, i, j, 0), G(u_pri, C, i, j, 4)); });
}

// output directory (output.cpp:190-288): root/eos/positivity/test/formula/div/update/NXxNY_dt_RK_Order_CFL
std::string output_path(const imex_params& P, const std::string& root, const std::string& test) {
  std::string eos = P.eos == IMEX_EOS_IDEAL ? "Ideal Gas" : "Redlich Kwong";
  std::string pp = P.floor_on ? "PreservingON" : "PreservingOFF";
  std::string ft = P.formula == IMEX_FORMULA_ENERGY ? "EnergyFormula" : "PressureFormula";
  std::string ut = P.formula == IMEX_FORMULA_ENERGY ? "" : "definition";
  const char* dts[4] = {"StaggerCT", "UnStaggerCT", "Potential", "NoDivFree"};
  std::string dv = (P.nx == 1 || P.ny == 1) ? "" : dts[P.magfield];
  int order = P.recon == IMEX_RECON_WENO5 ? 4 : (P.recon == IMEX_RECON_ENO3 ? 3 : (P.recon == IMEX_RECON_NONE ? 1 : 2));
  std::ostringstream ss;
  ss << std::fixed << std::setprecision(0) << 100 * P.cfl;
  return root + "/" + eos + "/" + pp + "/" + test + "/" + ft + "/" + dv + "/" + ut + "/" + std::to_string(P.nx) + "x" +
         std::to_string(P.ny) + "_" + "IMEXdt" + std::to_string(1) + "_RK" + std::to_string(2) + "_Order" + std::to_string(order) +
         "_CFL0" + ss.str();
}

// per-step diagnostics (main.cpp:562-592, timestep.cpp:56-57)
void diagnose(Ctx& C, const imex_scalars& S, const Vec& u, const Vec& bx, const Vec& by, imex_diag& d) {
  d.timestep = S.timestep; d.t = S.t; d.dt = S.dt;
  d.div_error = C.P.magfield == IMEX_MAGFIELD_STAGGER ? div_b(C, bx, by) : div_b_cell(C, u);
  double emin = 1e30, mass = 0, intE = 0, mmax = 0; int im = -1, jm = -1;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double rho = G(u, C, i, j, 0);
      double k = 0.5 * (G(u, C, i, j, 1) * G(u, C, i, j, 1) + G(u, C, i, j, 2) * G(u, C, i, j, 2) + G(u, C, i, j, 3) * G(u, C, i, j, 3)) / rho;
      double m = 0.5 * (G(u, C, i, j, 5) * G(u, C, i, j, 5) + G(u, C, i, j, 6) * G(u, C, i, j, 6) + G(u, C, i, j, 7) * G(u, C, i, j, 7));
      double e = G(u, C, i, j, 4) - k - m;
      mass += rho; intE += e; if (m > mmax) mmax = m; if (e < emin) { emin = e; im = i; jm = j; }
    }
  d.emin = emin; d.emin_x = C.P.x0 + (im - GC + 0.5) * C.dx; d.emin_y = C.P.y0 + (jm - GC + 0.5) * C.dy;
  d.mass = mass * C.dx * C.dy; d.intE = intE * C.dx * C.dy; d.magE_max = mmax;
  double binf = 0;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double dv = (G(u, C, i + 1, j, 5) - G(u, C, i - 1, j, 5)) / (2.0 * C.dx) + (G(u, C, i, j + 1, 6) - G(u, C, i, j - 1, 6)) / (2.0 * C.dy);
      if (std::fabs(dv) > binf) binf = std::fabs(dv);
    }
  d.B_inf = binf;
  double dmin = std::fmin(C.dx, C.dy);
  double fully_dt = C.P.cfl * 1 / (2 * S.ch / dmin), imex_dt = C.P.cfl * 1 / (2 * S.a_max / dmin);
  d.dt_comp = imex_dt / fully_dt;
}

// Diagnose_CT_EMF (constrained_transport.cpp:181-222) with CT_EMF_Stats (:165-178)
void ct_emf_diag(Ctx& C, const Vec& u, double out[16]) {
  const size_t NA = (size_t)C.NX * C.NY;
  Vec Cp0(NA, 0.0), C0p(NA, 0.0), Rp0(NA, 0.0), R0p(NA, 0.0), Et(NA, 0.0), Zero(NA, 0.0);
  Vec EppC(NA, 0.0), EppR(NA, 0.0), EppT(NA, 0.0), EppTot(NA, 0.0);
  Recon w;
  tvd_reconstruction(C, u, w);
  auto fx6 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[6] * vx - vy * Q[5]; };   // f_full_con[6]
  auto gy5 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[5] * vy - vx * Q[6]; };   // g_full_con[5]
  for (int i = 0; i < C.NX - 1; i++)
    for (int j = 0; j < C.NY - 1; j++) {
      const double* QL = &w.xR[NV * (i + j * C.NX)];
      const double* QR = &w.xL[NV * (i + 1 + j * C.NX)];
      A(Cp0, C, i, j) = -(0.5 * (fx6(QL) + fx6(QR)));
      A(Rp0, C, i, j) = 0.5 * std::fmax(std::fabs(QL[1] / QL[0]), std::fabs(QR[1] / QR[0])) * (QR[6] - QL[6]);
      const double* PL = &w.yR[NV * (i + j * C.NX)];
      const double* PR = &w.yL[NV * (i + (j + 1) * C.NX)];
      A(C0p, C, i, j) = 0.5 * (gy5(PL) + gy5(PR));
      A(R0p, C, i, j) = -0.5 * std::fmax(std::fabs(PL[2] / PL[0]), std::fabs(PR[2] / PR[0])) * (PR[5] - PL[5]);
    }
  bc_array(C, Cp0); bc_array(C, C0p); bc_array(C, Rp0); bc_array(C, R0p);
  for (int j = 0; j < C.NY; j++)
    for (int i = 0; i < C.NX; i++)
      A(Et, C, i, j) = G(u, C, i, j, 2) / G(u, C, i, j, 0) * G(u, C, i, j, 5) - G(u, C, i, j, 1) / G(u, C, i, j, 0) * G(u, C, i, j, 6);
  corner_assemble(C, EppC, Cp0, C0p, Zero);
  corner_assemble(C, EppR, Rp0, R0p, Zero);
  corner_assemble(C, EppT, Zero, Zero, Et);
  for (int j = 1; j < C.NY - 2; j++)
    for (int i = 1; i < C.NX - 2; i++) A(EppTot, C, i, j) = A(EppC, C, i, j) + A(EppR, C, i, j) + A(EppT, C, i, j);
  const Vec* parts[4] = {&EppC, &EppR, &EppT, &EppTot};
  for (int q = 0; q < 4; q++) {
    const Vec& E = *parts[q];
    double l2E = 0, mxE = 0, l2B = 0, mxB = 0;
    for (int j = GC; j < C.ny + GC; j++)
      for (int i = GC; i < C.nx + GC; i++) {
        double e = A(E, C, i, j);
        l2E += C.dx * C.dy * e * e; if (std::fabs(e) > mxE) mxE = std::fabs(e);
        double dBx = -(A(E, C, i, j) - A(E, C, i, j - 1)) / C.dy;
        double dBy = (A(E, C, i, j) - A(E, C, i - 1, j)) / C.dx;
        double m2 = dBx * dBx + dBy * dBy;
        l2B += C.dx * C.dy * m2; if (std::sqrt(m2) > mxB) mxB = std::sqrt(m2);
      }
    out[4 * q] = std::sqrt(l2E); out[4 * q + 1] = mxE; out[4 * q + 2] = std::sqrt(l2B); out[4 * q + 3] = mxB;
  }
}

// ---------------------------------------------------------------- time step (timestep.cpp:5-67, 89-120, 254-273)
void cfl(Ctx& C, imex_scalars& S, const Vec& u_pri) {
  double ch_x = 1e-8, ch_y = 1e-8, ch = 1e-8;
  double bmax_x = 0, bmax_y = 0;
  for (int i = 0; i < C.NX; i++)
    for (int j = 0; j < C.NY; j++) {
      const double* q = &u_pri[NV * (i + j * C.NX)];
      double rho = q[0], p = q[4], Bx = q[5], By = q[6], Bz = q[7];
      double B2 = Bx * Bx + By * By + Bz * Bz;
      double B = std::sqrt(B2);
      double v = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      double cf_x = eos_fast_speed(C, rho, p, Bx, B2), cf_y = eos_fast_speed(C, rho, p, By, B2), cf = eos_fast_speed(C, rho, p, B, B2);
      if (ch_x < q[1] + cf_x) ch_x = q[1] + cf_x;
      if (ch_y < q[2] + cf_y) ch_y = q[2] + cf_y;
      if (ch < v + cf) ch = v + cf;
      double b_x = 0.5 * (std::fabs(q[1]) + std::sqrt(q[1] * q[1] + 4 * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / q[0]));
      double b_y = 0.5 * (std::fabs(q[2]) + std::sqrt(q[2] * q[2] + 4 * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / q[0]));
      if (std::fabs(bmax_x) < std::fabs(b_x)) bmax_x = b_x;
      if (std::fabs(bmax_y) < std::fabs(b_y)) bmax_y = b_y;
    }
  S.ch_x = ch_x; S.ch_y = ch_y; S.ch = ch;
  double vx_max = 0, vy_max = 0, v_max = 0;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      const double* q = &u_pri[NV * (i + j * C.NX)];
      if (std::fabs(q[1]) > std::fabs(vx_max)) vx_max = std::fabs(q[1]);
      if (std::fabs(q[2]) > std::fabs(vy_max)) vy_max = std::fabs(q[2]);
      double vm = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      if (vm > std::fabs(v_max)) v_max = vm;
    }
  S.vx_max = vx_max; S.vy_max = vy_max; S.v_max = v_max;
  double a_max = (S.timestep < 1 + 1) ? ch : v_max;                   // IMEXdt, full_step = 1
  double dt = C.P.cfl * 1 / (2 * a_max / std::fmin(C.dx, C.dy));
  if (S.timestep > 1 && S.prev_dt > 0.0) dt = std::fmin(dt, 2.0 * S.prev_dt);
  S.prev_dt = dt;
  if (S.timestep > 1) a_max = v_max;
  S.a_max = a_max;
  S.B_max_x = bmax_x; S.B_max_y = bmax_y;
  S.dt = std::fmin(dt, C.P.time_stop - S.t);                         // main.cpp:341
}

// ---------------------------------------------------------------- whole run
struct Run {
  Ctx C;
  imex_scalars S{};
  Vec u_pri, u_con, u_n, Hs[2];
  Vec bx, by, bx_n, by_n, HsB[2][2];
  Vec Apot, A_n, HsA[2];       // vector potential (Grid, slots 0..2), its step-start copy and stage increments
  Stage W;
  int cap_point = -1;          // capture: 10*rks + sub (sub: 0 before conv, 1 before mag, 2 before press-flux, 3 before pressure, 4 before CT, 5 end of stage)
  Stage cap; imex_scalars cap_S{};
};

void maybe_capture(Run& R, int point) {
  if (R.cap_point == point) { R.cap = R.W; R.cap_S = R.S; }
}

int run_step(Run& R) {
  Ctx& C = R.C;
  imex_scalars& S = R.S;
  int rc = 0;
  S.timestep++;
  bc_grid(C, R.u_pri, true);                                          // g_bc_prim around bound_cond_u(u_pri) (main.cpp:336)
  cfl(C, S, R.u_pri);
  S.t = S.t + S.dt;
  grid_pri2con(C, R.u_con, R.u_pri);
  R.u_n = R.u_con;
  R.bx_n = R.bx; R.by_n = R.by;
  const int mf = C.P.magfield;
  const bool useA = mf == IMEX_MAGFIELD_UNSTAGGER || mf == IMEX_MAGFIELD_POTENTIAL_A;
  if (useA) R.A_n = R.Apot;                                          // main.cpp:347-348
  const double bb = 1.0 - 1.0 / std::sqrt(2.0);
  const double Ae10 = 1.0 / (2.0 * bb);
  const double Ai[2][2] = {{bb, 0.0}, {1 - bb, bb}};
  const double Ae[2][2] = {{0.0, 0.0}, {Ae10, 0.0}};
  const size_t NG = R.u_con.size(), NA = R.bx.size();
  for (int rks = 0; rks < 2; rks++) {
    S.rks = rks;
    Stage& W = R.W;
    W.Qe = R.u_n; W.Qs = R.u_n;
    W.bx_qs = R.bx_n; W.by_qs = R.by_n;
    for (int r = 0; r < rks; r++) {
      double ce = Ae[rks][r] * S.dt, ci = Ai[rks][r] * S.dt;
      #pragma omp parallel for schedule(static)
      for (size_t n = 0; n < NG; n++) { W.Qe[n] = W.Qe[n] + ce * R.Hs[r][n]; W.Qs[n] = W.Qs[n] + ci * R.Hs[r][n]; }
      for (size_t n = 0; n < NA; n++) { W.bx_qs[n] += ci * R.HsB[r][0][n]; W.by_qs[n] += ci * R.HsB[r][1][n]; }
    }
    if (useA) {                                                      // main.cpp:371-412
      W.As = R.A_n;
      for (int r = 0; r < rks; r++) {
        double ci = Ai[rks][r] * S.dt;
        for (size_t n = 0; n < NG; n += NV)
          for (int k = 0; k < 3; k++) W.As[n + k] += ci * R.HsA[r][n + k];
      }
      W.A = W.As;
    }
    S.adt = Ai[rks][rks] * S.dt;
    W.u_con = R.u_con; W.bx = R.bx; W.by = R.by;
    if (W.fx.size() != NA) { W.fx.assign(NA, 0.0); W.fy.assign(NA, 0.0); }
    maybe_capture(R, 10 * rks + 0);
    Recon w;
    tvd_reconstruction(C, W.Qe, w);
    flux_update(C, S, W.u_con, W.Qs, w, 0);
    bc_grid(C, W.u_con);
    maybe_capture(R, 10 * rks + 1);
    rc |= mf == IMEX_MAGFIELD_POTENTIAL_A ? mag_update_a(C, S, W) : mag_update_b(C, S, W);   // main.cpp:424-434
    bc_grid(C, W.u_con);
    maybe_capture(R, 10 * rks + 2);
    for (size_t n = 0; n < NG; n += NV)
      for (int k = 5; k < 8; k++) { w.xL[n + k] = W.u_con[n + k]; w.xR[n + k] = W.u_con[n + k]; w.yL[n + k] = W.u_con[n + k]; w.yR[n + k] = W.u_con[n + k]; }
    const int energy = C.P.formula == IMEX_FORMULA_ENERGY;     // main.cpp:470-486
    flux_update(C, S, W.u_con, W.Qs, w, energy ? 2 : 1);
    bc_grid(C, W.u_con);
    maybe_capture(R, 10 * rks + 3);
    rc |= energy ? pre_subsystem_energy(C, S, W) : pre_subsystem_pressure(C, S, W);
    bc_grid(C, W.u_con);
    maybe_capture(R, 10 * rks + 4);
    if (mf == IMEX_MAGFIELD_STAGGER) {                                 // main.cpp:497-510
      if (C.P.ct_mode == IMEX_CT_PLAIN) ct_plain(C, S, W); else ct_consistent(C, S, W);
    } else if (mf == IMEX_MAGFIELD_UNSTAGGER) {
      unstaggered_ct(C, S, W);
    }
    bc_grid(C, W.u_con);
    R.u_con = W.u_con; R.bx = W.bx; R.by = W.by;
    if (useA) {                                                      // main.cpp:546-551
      R.Apot = W.A;
      if (R.HsA[rks].size() != NG) R.HsA[rks].assign(NG, 0.0);
      for (size_t n = 0; n < NG; n += NV)
        for (int k = 0; k < 3; k++) R.HsA[rks][n + k] = (W.A[n + k] - W.As[n + k]) / S.adt;
    }
    grid_con2pri(C, R.u_pri, R.u_con);
    if (R.Hs[rks].size() != NG) R.Hs[rks].assign(NG, 0.0);
    #pragma omp parallel for schedule(static)
    for (size_t n = 0; n < NG; n++) R.Hs[rks][n] = (R.u_con[n] - W.Qs[n]) / S.adt;
    for (int c = 0; c < 2; c++) if (R.HsB[rks][c].size() != NA) R.HsB[rks][c].assign(NA, 0.0);
    for (size_t n = 0; n < NA; n++) { R.HsB[rks][0][n] = (R.bx[n] - W.bx_qs[n]) / S.adt; R.HsB[rks][1][n] = (R.by[n] - W.by_qs[n]) / S.adt; }
    maybe_capture(R, 10 * rks + 5);
  }
  grid_con2pri(C, R.u_pri, R.u_con);
  C.st.steps++;
  if (C.eos_warn) { rc |= IMEX_WARN_EOS_MAXITER; C.st.warnings |= IMEX_WARN_EOS_MAXITER; }
  return rc;
}

double div_b(const Ctx& C, const Vec& bx, const Vec& by) {   // imex_subsystems.cpp:630-660, StaggerCT branch
  double err = 0;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double d = (A(bx, C, i, j) - A(bx, C, i - 1, j)) / C.dx + (A(by, C, i, j) - A(by, C, i, j - 1)) / C.dy;
      err += C.dx * C.dy * d * d;
    }
  return std::sqrt(err);
}

// cal_divergence_error for UnStaggerCT / PotentialA / NoDivFree (imex_subsystems.cpp:640-655): the cell
// B's divergence, central inside, one-sided at the domain edges.
double div_b_cell(const Ctx& C, const Vec& u) {
  double err = 0;
  for (int i = GC; i < C.nx + GC; i++)
    for (int j = GC; j < C.ny + GC; j++) {
      double dBxdx, dBydy;
      if (i == GC) dBxdx = (G(u, C, i + 1, j, 5) - G(u, C, i, j, 5)) / C.dx;
      else if (i == C.nx + GC - 1) dBxdx = (G(u, C, i, j, 5) - G(u, C, i - 1, j, 5)) / C.dx;
      else dBxdx = (G(u, C, i + 1, j, 5) - G(u, C, i - 1, j, 5)) / (2.0 * C.dx);
      if (j == GC) dBydy = (G(u, C, i, j + 1, 6) - G(u, C, i, j, 6)) / C.dy;
      else if (j == C.ny + GC - 1) dBydy = (G(u, C, i, j, 6) - G(u, C, i, j - 1, 6)) / C.dy;
      else dBydy = (G(u, C, i, j + 1, 6) - G(u, C, i, j - 1, 6)) / (2.0 * C.dy);
      double d = dBxdx + dBydy;
      err += C.dx * C.dy * d * d;
    }
  return std::sqrt(err);
}

void ctx_init(Ctx& C, const imex_params* p) {
  C.P = *p;
  C.nx = p->nx; C.ny = p->ny; C.NX = p->nx + 2 * GC; C.NY = p->ny + 2 * GC;
  C.dx = (p->x1 - p->x0) / ((p->nx + 1) - 1);                          // main.cpp:226-227
  C.dy = (p->y1 - p->y0) / ((p->ny + 1) - 1);
  C.st = imex_stats{}; C.st.min_p = 1e300;
}

void stage_from_host(const Ctx& C, Stage& W, const imex_stage_host* h) {
  size_t NG = (size_t)C.NX * C.NY * NV, NA = (size_t)C.NX * C.NY;
  W.u_con.assign(h->u_con, h->u_con + NG); W.Qe.assign(h->Qe, h->Qe + NG); W.Qs.assign(h->Qs, h->Qs + NG);
  W.bx.assign(h->bx, h->bx + NA); W.by.assign(h->by, h->by + NA);
  W.bx_qs.assign(h->bx_qs, h->bx_qs + NA); W.by_qs.assign(h->by_qs, h->by_qs + NA);
  W.fx.assign(h->fx, h->fx + NA); W.fy.assign(h->fy, h->fy + NA);
}
void stage_to_host(const Stage& W, imex_stage_host* h) {
  auto cp = [](const Vec& v, double* d) { if (d && !v.empty()) std::memcpy(d, v.data(), v.size() * sizeof(double)); };
  cp(W.u_con, h->u_con); cp(W.Qe, h->Qe); cp(W.Qs, h->Qs); cp(W.bx, h->bx); cp(W.by, h->by);
  cp(W.bx_qs, h->bx_qs); cp(W.by_qs, h->by_qs); cp(W.fx, h->fx); cp(W.fy, h->fy);
}

}  // namespace

// ================================================================ C API (test infrastructure)
extern "C" {

void* orc_create(const imex_params* p, int faithful) {
  Run* R = new Run();
  ctx_init(R->C, p);
  R->C.faithful = faithful;
  size_t NG = (size_t)R->C.NX * R->C.NY * NV, NA = (size_t)R->C.NX * R->C.NY;
  R->u_pri.assign(NG, 0.0); R->u_con.assign(NG, 0.0);
  R->bx.assign(NA, 0.0); R->by.assign(NA, 0.0);
  R->Apot.assign(NG, 0.0);
  return R;
}

// the step-start vector potential (Grid, slots 0..2)
int orc_set_potential(void* h, const double* A) {
  Run& R = *(Run*)h;
  std::memcpy(R.Apot.data(), A, R.Apot.size() * sizeof(double));
  return 0;
}
int orc_get_potential(void* h, double* A) {
  Run& R = *(Run*)h;
  std::memcpy(A, R.Apot.data(), R.Apot.size() * sizeof(double));
  return 0;
}
// A and A_s of the last capture (orc_step_capture)
int orc_capture_potential(void* h, double* A, double* As) {
  Run& R = *(Run*)h;
  size_t n = (size_t)R.C.NX * R.C.NY * NV;
  if (R.cap.A.size() != n) return -1;
  std::memcpy(A, R.cap.A.data(), n * sizeof(double));
  std::memcpy(As, R.cap.As.data(), n * sizeof(double));
  return 0;
}
void orc_destroy(void* h) { delete (Run*)h; }

// Threads of the per-cell loops (results do not depend on it); returns the count now in force.
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_set_state(void* h, const double* u_pri, const double* bx, const double* by, double t, int timestep, double prev_dt) {
  Run& R = *(Run*)h;
  std::memcpy(R.u_pri.data(), u_pri, R.u_pri.size() * sizeof(double));
  std::memcpy(R.bx.data(), bx, R.bx.size() * sizeof(double));
  std::memcpy(R.by.data(), by, R.by.size() * sizeof(double));
  grid_pri2con(R.C, R.u_con, R.u_pri);
  R.S = imex_scalars{}; R.S.t = t; R.S.timestep = timestep; R.S.prev_dt = prev_dt;
  return 0;
}

int orc_step(void* h, int nsteps) {
  Run& R = *(Run*)h;
  int rc = 0;
  for (int s = 0; s < nsteps; s++) rc |= run_step(R);
  return rc;
}

// Run one step, capturing the stage working set at `point` (10*rks + sub).
int orc_step_capture(void* h, int point, imex_stage_host* out, imex_scalars* sc) {
  Run& R = *(Run*)h;
  R.cap_point = point;
  int rc = run_step(R);
  R.cap_point = -1;
  stage_to_host(R.cap, out);
  if (sc) *sc = R.cap_S;
  return rc;
}

int orc_get_state(void* h, double* u_con, double* u_pri, double* bx, double* by, imex_scalars* sc) {
  Run& R = *(Run*)h;
  if (u_con) std::memcpy(u_con, R.u_con.data(), R.u_con.size() * sizeof(double));
  if (u_pri) std::memcpy(u_pri, R.u_pri.data(), R.u_pri.size() * sizeof(double));
  if (bx) std::memcpy(bx, R.bx.data(), R.bx.size() * sizeof(double));
  if (by) std::memcpy(by, R.by.data(), R.by.size() * sizeof(double));
  if (sc) *sc = R.S;
  return 0;
}

int orc_get_stats(void* h, imex_stats* st) {
  Run& R = *(Run*)h;
  *st = R.C.st;
  st->div_b = R.C.P.magfield == IMEX_MAGFIELD_STAGGER ? div_b(R.C, R.bx, R.by) : div_b_cell(R.C, R.u_con);
  return 0;
}

double orc_div_b(const imex_params* p, const double* bx, const double* by) {
  Ctx C; ctx_init(C, p);
  size_t NA = (size_t)C.NX * C.NY;
  Vec vx(bx, bx + NA), vy(by, by + NA);
  return div_b(C, vx, vy);
}

// ---- stage-level entry points on caller-owned host arrays ----
static int run_sub(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, int which, imex_stats* stats) {
  Ctx C; ctx_init(C, p);
  Stage W; stage_from_host(C, W, st);
  imex_scalars S = *sc;
  int rc = 0;
  if (which == 0) {
    Recon w; tvd_reconstruction(C, W.Qe, w); flux_update(C, S, W.u_con, W.Qs, w, 0); bc_grid(C, W.u_con);
  } else if (which == 1) {
    rc = mag_update_b(C, S, W); bc_grid(C, W.u_con);
  } else if (which == 2 || which == 5) {
    Recon w; tvd_reconstruction(C, W.Qe, w);
    for (size_t n = 0; n < W.u_con.size(); n += NV)
      for (int k = 5; k < 8; k++) { w.xL[n + k] = W.u_con[n + k]; w.xR[n + k] = W.u_con[n + k]; w.yL[n + k] = W.u_con[n + k]; w.yR[n + k] = W.u_con[n + k]; }
    flux_update(C, S, W.u_con, W.Qs, w, which == 2 ? 1 : 2); bc_grid(C, W.u_con);
  } else if (which == 6) {
    rc = pre_subsystem_energy(C, S, W); bc_grid(C, W.u_con);
  } else if (which == 3) {
    rc = pre_subsystem_pressure(C, S, W); bc_grid(C, W.u_con);
  } else if (which == 4) {
    if (C.P.ct_mode == IMEX_CT_PLAIN) ct_plain(C, S, W); else ct_consistent(C, S, W);
    bc_grid(C, W.u_con);
  }
  stage_to_host(W, st);
  if (stats) *stats = C.st;
  return rc;
}
int orc_explicit_conv(const imex_params* p, const imex_scalars* sc, imex_stage_host* st) { return run_sub(p, sc, st, 0, nullptr); }
int orc_mag_update_b(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, imex_stats* s) { return run_sub(p, sc, st, 1, s); }
int orc_explicit_press(const imex_params* p, const imex_scalars* sc, imex_stage_host* st) { return run_sub(p, sc, st, 2, nullptr); }
int orc_pre_subsystem_pressure(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, imex_stats* s) { return run_sub(p, sc, st, 3, s); }
int orc_ct_update(const imex_params* p, const imex_scalars* sc, imex_stage_host* st) { return run_sub(p, sc, st, 4, nullptr); }
int orc_explicit_energy(const imex_params* p, const imex_scalars* sc, imex_stage_host* st) { return run_sub(p, sc, st, 5, nullptr); }
int orc_pre_subsystem_energy(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, imex_stats* s) { return run_sub(p, sc, st, 6, s); }

// MagSubsystem_UpdateA / Unstaggered_CT_Consistent on a stage snapshot; A in/out, A_s in (Grids).
static int run_sub_a(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, double* A, const double* As,
                     int which, imex_stats* stats) {
  Ctx C; ctx_init(C, p);
  Stage W; stage_from_host(C, W, st);
  size_t NG = (size_t)C.NX * C.NY * NV;
  W.A.assign(A, A + NG); W.As.assign(As, As + NG);
  int rc = 0;
  if (which == 0) rc = mag_update_a(C, *sc, W); else unstaggered_ct(C, *sc, W);
  bc_grid(C, W.u_con);
  stage_to_host(W, st);
  std::memcpy(A, W.A.data(), NG * sizeof(double));
  if (stats) *stats = C.st;
  return rc;
}
int orc_mag_update_a(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, double* A, const double* As, imex_stats* s) {
  return run_sub_a(p, sc, st, A, As, 0, s);
}
int orc_unstaggered_ct(const imex_params* p, const imex_scalars* sc, imex_stage_host* st, double* A, const double* As) {
  return run_sub_a(p, sc, st, A, As, 1, nullptr);
}
int orc_bc_potential(const imex_params* p, double* A) {
  Ctx C; ctx_init(C, p);
  Vec v(A, A + (size_t)C.NX * C.NY * NV);
  bc_A(C, v);
  std::memcpy(A, v.data(), v.size() * sizeof(double));
  return 0;
}
double orc_div_b_cell(const imex_params* p, const double* u_con) {
  Ctx C; ctx_init(C, p);
  Vec v(u_con, u_con + (size_t)C.NX * C.NY * NV);
  return div_b_cell(C, v);
}

// One face flux of the explicit updates: kind 0 RusanovC, 1 Pressure_flux, 2 Energy_flux; dir 0 x, 1 y.
int orc_flux(const imex_params* p, int kind, int dir, const double* QL, const double* QR, double s_max, double* out) {
  imex_scalars S{};
  S.a_max = s_max;
  rusanov(S, p->gamma, dir, kind, QL, QR, out);
  return 0;
}

// ---- output files and diagnostics ----
int orc_write_vtk(const imex_params* p, const double* u_con, const double* u_pri, double t, int nstep, const char* file) {
  Ctx C; ctx_init(C, p);
  size_t NG = (size_t)C.NX * C.NY * NV;
  write_vtk(C, Vec(u_con, u_con + NG), Vec(u_pri, u_pri + NG), t, nstep, file);
  return 0;
}
int orc_write_fields(const imex_params* p, const double* u_con, const double* u_pri, const char* dir) {
  Ctx C; ctx_init(C, p);
  size_t NG = (size_t)C.NX * C.NY * NV;
  write_fields(C, Vec(u_con, u_con + NG), Vec(u_pri, u_pri + NG), dir);
  return 0;
}
int orc_output_path(const imex_params* p, const char* root, const char* test, char* out, int cap) {
  std::string s = output_path(*p, root, test);
  if ((int)s.size() + 1 > cap) return -1;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}
int orc_diagnose(const imex_params* p, const imex_scalars* sc, const double* u_con, const double* bx, const double* by, imex_diag* d) {
  Ctx C; ctx_init(C, p);
  size_t NG = (size_t)C.NX * C.NY * NV, NA = (size_t)C.NX * C.NY;
  diagnose(C, *sc, Vec(u_con, u_con + NG), Vec(bx, bx + NA), Vec(by, by + NA), *d);
  return 0;
}
int orc_ct_emf_diag(const imex_params* p, const double* u_con, double* out16) {
  Ctx C; ctx_init(C, p);
  size_t NG = (size_t)C.NX * C.NY * NV;
  ct_emf_diag(C, Vec(u_con, u_con + NG), out16);
  return 0;
}

// TVDReconstruction on a Grid: outputs four Grids.
int orc_reconstruct(const imex_params* p, const double* u, double* wxL, double* wxR, double* wyL, double* wyR) {
  Ctx C; ctx_init(C, p);
  size_t NG = (size_t)C.NX * C.NY * NV;
  Vec uv(u, u + NG);
  Recon w; tvd_reconstruction(C, uv, w);
  std::memcpy(wxL, w.xL.data(), NG * 8); std::memcpy(wxR, w.xR.data(), NG * 8);
  std::memcpy(wyL, w.yL.data(), NG * 8); std::memcpy(wyR, w.yR.data(), NG * 8);
  return 0;
}

// Operators on a stage snapshot. op: IMEX_OP_*.
int orc_apply_op(const imex_params* p, const imex_scalars* sc, const imex_stage_host* st, int op,
                 const double* x, double* Ax, const double* p_freeze, double mbig) {
  Ctx C; ctx_init(C, p); C.Mbig = mbig;
  Stage W; stage_from_host(C, W, st);
  if (op == IMEX_OP_MAG) { mag_apply(C, *sc, W.u_con, W.Qe, x, Ax); return 0; }
  if (op == IMEX_OP_ENERGY) { Vec hE; e_coef(C, W.u_con, W.Qe, hE); e_apply(C, *sc, W.u_con, W.Qe, hE, x, Ax); return 0; }
  if (op == IMEX_OP_POTENTIAL) { double gx, gy; pot_visc(C, W.Qs, gx, gy); pot_apply(C, *sc, W.u_con, W.Qe, W.Qs, gx, gy, x, Ax); return 0; }
  PCoef pc; p_coef(C, W.u_con, W.Qe, pc);
  p_apply(C, *sc, W.u_con, W.Qe, pc, x, Ax, op == IMEX_OP_P_JACOBIAN, p_freeze);
  return 0;
}

int orc_gmres(const imex_params* p, const imex_scalars* sc, const imex_stage_host* st, int op,
              const double* b, double* x, int precond, const double* p_freeze, double mbig, imex_solve_info* info) {
  Ctx C; ctx_init(C, p); C.Mbig = mbig;
  Stage W; stage_from_host(C, W, st);
  const int N = C.nx * C.ny, dim = (op == IMEX_OP_MAG || op == IMEX_OP_POTENTIAL) ? 3 * N : N;
  Vec xv(x, x + dim), bv(b, b + dim);
  const imex_scalars S = *sc;
  if (op == IMEX_OP_POTENTIAL) {
    double gx, gy;
    pot_visc(C, W.Qs, gx, gy);
    std::function<void(const double*, double*)> Ap = [&](const double* xin, double* out) { pot_apply(C, S, W.u_con, W.Qe, W.Qs, gx, gy, xin, out); };
    int rc = gmres(C, xv, bv, Ap, nullptr, info, nullptr, nullptr);
    std::memcpy(x, xv.data(), dim * sizeof(double));
    return rc;
  }
  if (op == IMEX_OP_ENERGY) {
    Vec hE, D;
    e_coef(C, W.u_con, W.Qe, hE);
    e_diag(C, S, hE, D);
    std::function<void(const double*, double*)> Ae = [&](const double* xin, double* out) { e_apply(C, S, W.u_con, W.Qe, hE, xin, out); };
    std::function<void(double*, const double*)> precE = [&](double* out, const double* in) { for (int n = 0; n < N; n++) out[n] = in[n] / D[n]; };
    int rc = gmres(C, xv, bv, Ae, precond ? &precE : nullptr, info, nullptr, nullptr);
    std::memcpy(x, xv.data(), dim * sizeof(double));
    return rc;
  }
  PCoef pc;
  if (op != IMEX_OP_MAG) p_coef(C, W.u_con, W.Qe, pc);
  std::function<void(const double*, double*)> Aop = [&](const double* xin, double* out) {
    if (op == IMEX_OP_MAG) mag_apply(C, S, W.u_con, W.Qe, xin, out);
    else p_apply(C, S, W.u_con, W.Qe, pc, xin, out, op == IMEX_OP_P_JACOBIAN, p_freeze);
  };
  Vec Tdiag(N);
  const double h = S.adt, a = 1.0 / (C.P.gamma - 1.0);
  if (op != IMEX_OP_MAG)
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++) {
        double h00 = A(pc.hhat, C, i, j), hp0 = A(pc.hhat, C, i + 1, j), hm0 = A(pc.hhat, C, i - 1, j);
        double h0p = A(pc.hhat, C, i, j + 1), h0m = A(pc.hhat, C, i, j - 1);
        Tdiag[M_P(C, i, j)] = h * h * ((0.5 * (hp0 + h00) + 0.5 * (hm0 + h00)) / (C.dx * C.dx) + (0.5 * (h0p + h00) + 0.5 * (h0m + h00)) / (C.dy * C.dy));
      }
  std::function<void(double*, const double*)> precP = [&](double* out, const double* in) {
    for (int i = GC; i < C.nx + GC; i++)
      for (int j = GC; j < C.ny + GC; j++) {
        int idx = M_P(C, i, j);
        if (dir_bnd(C, i, j)) { out[idx] = in[idx]; continue; }     // identity rows (implicit_operators.cpp:652-655)
        double Vd = op == IMEX_OP_P_JACOBIAN ? V_slope(C, p_freeze[idx], G(W.u_con, C, i, j, 0)) : a;
        out[idx] = in[idx] / (Vd + Tdiag[idx]);
      }
  };
  int rc = gmres(C, xv, bv, Aop, (precond && op != IMEX_OP_MAG) ? &precP : nullptr, info, nullptr, nullptr);
  std::memcpy(x, xv.data(), dim * sizeof(double));
  return rc;
}

// The energy solve's preconditioner divisor diag dx dy per unknown (M_P order).
int orc_energy_diag(const imex_params* p, const imex_scalars* sc, const imex_stage_host* st, double* D) {
  Ctx C; ctx_init(C, p);
  Stage W; stage_from_host(C, W, st);
  Vec hE, Dv;
  e_coef(C, W.u_con, W.Qe, hE);
  e_diag(C, *sc, hE, Dv);
  std::memcpy(D, Dv.data(), Dv.size() * sizeof(double));
  return 0;
}

int orc_pressure_newton(const imex_params* p, const imex_scalars* sc, const imex_stage_host* st,
                        const double* b, double* pv, imex_solve_info* info, double* mbig_out) {
  Ctx C; ctx_init(C, p);
  Stage W; stage_from_host(C, W, st);
  const int N = C.nx * C.ny;
  PCoef pc; p_coef(C, W.u_con, W.Qe, pc);
  Vec pvec(pv, pv + N), bv(b, b + N);
  int rc = pressure_newton(C, *sc, W.u_con, W.Qe, pc, pvec, bv, info);
  std::memcpy(pv, pvec.data(), N * sizeof(double));
  if (mbig_out) *mbig_out = C.Mbig;
  return rc;
}

// Pointwise EOS / state functions for unit pins.
double orc_eos(const imex_params* p, int which, double a, double b) {
  Ctx C; ctx_init(C, p);
  switch (which) {
    case 0: return eos_p_from_rho_e(C, a, b);
    case 1: return eos_e_from_rho_p(C, a, b);
    case 2: return eos_sound_a(C, a, b);
    case 3: return Vbase(C, a, b);
    case 4: return Vbase_slope(C, a, b);
    case 5: return C.P.eos == IMEX_EOS_IDEAL ? 0.0 : rk_T_from_rho_p(C, a, b);
  }
  return NAN;
}
int orc_con2pri(const imex_params* p, const double* con, double* pri, int ncells) {
  Ctx C; ctx_init(C, p);
  for (int n = 0; n < ncells; n++) con2pri(C, pri + NV * n, con + NV * n);
  return 0;
}
int orc_pri2con(const imex_params* p, const double* pri, double* con, int ncells) {
  Ctx C; ctx_init(C, p);
  for (int n = 0; n < ncells; n++) pri2con(C, con + NV * n, pri + NV * n);
  return 0;
}
int orc_shock(const imex_params* p, const double* Qe, double* eta) {
  Ctx C; ctx_init(C, p);
  Vec q(Qe, Qe + (size_t)C.NX * C.NY * NV), e;
  shock_detector(C, e, q);
  std::memcpy(eta, e.data(), e.size() * sizeof(double));
  return 0;
}
int orc_cfl(const imex_params* p, const double* u_pri, imex_scalars* sc) {
  Ctx C; ctx_init(C, p);
  Vec u(u_pri, u_pri + (size_t)C.NX * C.NY * NV);
  cfl(C, *sc, u);
  return 0;
}
// kind 0 grid, 1 array, 2 staggered axis 0, 3 staggered axis 1, 4 grid of primitives
int orc_bc(const imex_params* p, double* a, int kind) {
  Ctx C; ctx_init(C, p);
  size_t n = (size_t)C.NX * C.NY * (kind == 0 || kind == 4 ? NV : 1);
  Vec v(a, a + n);
  if (kind == 0 || kind == 4) bc_grid(C, v, kind == 4); else if (kind == 1) bc_array(C, v); else bc_staggered(C, v, kind - 2);
  std::memcpy(a, v.data(), n * sizeof(double));
  return 0;
}
int orc_map_index(const imex_params* p, int i, int j) { Ctx C; ctx_init(C, p); return M_P(C, i, j); }

}  // extern "C"
