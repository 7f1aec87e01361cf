// device_common.hpp — shared device-side definitions for the IMEX-MHD hot path on gfx950.
//
// Layout: every cell field is a separate FP64 plane of (nx+6) x (ny+6) cells
// (SoA), row-major with the x index fastest, the reference's 3-cell ghost frame
// kept (include/types.hpp:37-91).  Plane p of a grid starts at base + p*PS.
// Krylov vectors are compact (nx*ny per component) in the reference's M_P/M_B
// order (src/implicit_operators.cpp:5-139).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "../../include/imex.h"

#define GC 3
#define NVAR 8

// Step-global scalars kept in device memory so that no kernel needs a host round trip.
enum ScalarSlot {
  S_T = 0, S_DT, S_ADT, S_AMAX, S_BMAXX, S_BMAXY, S_VXMAX, S_VYMAX, S_VMAX,
  S_CH, S_CHX, S_CHY, S_PREVDT, S_TIMESTEP, S_RKS, S_MBIG, S_ADT0,
  S_GBX, S_GBY,        // viscosity maxima of the vector-potential operator (magnetic eigenvalues of Qs)
  S_NSLOTS = 32
};

struct Dims {
  int nx, ny, NX, NY;     // interior and padded extents
  long long PS;           // plane stride (doubles)
  int N;                  // nx*ny
  double dx, dy;
  int perx, pery;         // 1 = periodic, 0 = transmissive
  int refx, refy;         // 1 = reflective (x_Reflective / y_Reflective)
  int dirx, diry;         // 1 = Dirichlet ambient (x_Dirichlet / y_Dirichlet)
  double rho_amb, p_amb, e_amb;   // ambient density and pressure; e_amb = p_amb / (gamma - 1)
  double x0, y0;          // domain origin (cell centres at x0 + (i - g + 1/2) dx)
  int jet;                // 1 = bottom boundary is the MHD jet inflow (y_JetInflow)
  double jet_hw, jet_rho, jet_v, jet_B0, jet_p, jet_E;   // strip half-width and inflow state; jet_E = jet_E_val()
};

// The jet inflow strip of y_JetInflow (config.hpp:261-270): cell centres |x_c| < half-width; x faces
// i+1/2 with |x_f| < half-width - 1e-12 (the staggered Bx rule).
__host__ __device__ inline bool jet_inlet(const Dims& d, int i) {
  return d.jet && fabs(d.x0 + (i - GC + 0.5) * d.dx) < d.jet_hw;
}
__host__ __device__ inline bool jet_inlet_face(const Dims& d, int i) {
  return d.jet && fabs(d.x0 + (i - GC + 1.0) * d.dx) < d.jet_hw - 1e-12;
}
// Conservative (prim = 0) or primitive (prim = 1) inflow state, slot k
__host__ __device__ inline double jet_state(const Dims& d, int prim, int k) {
  switch (k) {
    case 0: return d.jet_rho;
    case 2: return prim ? d.jet_v : d.jet_rho * d.jet_v;
    case 4: return prim ? d.jet_p : d.jet_E;
    case 6: return d.jet_B0;
    default: return 0.0;
  }
}

struct Phys {
  double gamma;
  int eos;                // IMEX_EOS_*
  double rk_alfa, rk_b, rk_R, rk_cv;
  int floor_on;
  double p_floor, theta, mbig_fac;
  double kappa1;
  double cfl, time_stop;
  int recon, central_y;
  int prim;               // reconstruct primitives (ReconState PrimState)
  int bnd;                // a jet inflow, reflective or Dirichlet boundary is present
};

// ---------------------------------------------------------------- index maps
// bound_cond_u / bound_cond_array / M_P map of a padded index to its interior source.
// ghost -> image cell (M_P, implicit_operators.cpp:5-139): periodic wrap, reflective mirror, clamp otherwise
__device__ __forceinline__ int map_x(const Dims& d, int i) {
  if (i < GC) return d.perx ? d.nx + i : (d.refx ? 2 * GC - 1 - i : GC);
  if (i > d.nx + GC - 1) return d.perx ? i - d.nx : (d.refx ? 2 * (d.nx + GC) - i - 1 : d.nx + GC - 1);
  return i;
}
__device__ __forceinline__ int map_y(const Dims& d, int j) {
  if (j < GC) return d.pery ? d.ny + j : (d.refy ? 2 * GC - 1 - j : GC);
  if (j > d.ny + GC - 1) return d.pery ? j - d.ny : (d.refy ? 2 * (d.ny + GC) - j - 1 : d.ny + GC - 1);
  return j;
}
// A ghost cell on a Dirichlet side (x_Dirichlet / y_Dirichlet): it holds the ambient state
__host__ __device__ inline bool dir_ghost(const Dims& d, int i, int j) {
  return (d.dirx && (i < GC || i > d.nx + GC - 1)) || (d.diry && (j < GC || j > d.ny + GC - 1));
}
// A boundary cell of the pressure solve's strong Dirichlet rows (imex_subsystems.cpp:161, implicit_operators.cpp:550)
__host__ __device__ inline bool dir_bnd(const Dims& d, int i, int j) {
  return (d.dirx && (i == GC || i == d.nx + GC - 1)) || (d.diry && (j == GC || j == d.ny + GC - 1));
}
// Reflective ghost fills negate (boundary.cpp:17-54, 89-126; timestep.cpp:175-214): the factor a ghost
// value of sign class `neg` (1: momenta, B and every bound_cond_array array) takes at padded (i, j).
__device__ __forceinline__ double refl_factor(const Dims& d, int i, int j, bool neg) {
  if (!neg) return 1.0;
  bool gx = d.refx && (i < GC || i > d.nx + GC - 1), gy = d.refy && (j < GC || j > d.ny + GC - 1);
  return (gx != gy) ? -1.0 : 1.0;
}
__device__ __forceinline__ long long pidx(const Dims& d, int i, int j) { return (long long)i + (long long)j * d.NX; }
// compact Krylov index of padded cell (i,j) after the boundary map (M_P)
__device__ __forceinline__ int kidx(const Dims& d, int i, int j) { return (map_x(d, i) - GC) + d.nx * (map_y(d, j) - GC); }

// ---------------------------------------------------------------- EOS (include/config.hpp:130-239)
// The reference's Redlich-Kwong temperature solves loop without a cap; here
// they stop after RK_MAXIT iterations and raise a flag in *eflag.
#define RK_MAXIT 200
__host__ __device__ inline double rk_T_from_rho_p(const Phys& P, double rho, double p, int* eflag) {
  const double mR = P.rk_R, ma = P.rk_alfa, mb = P.rk_b;
  double v = 1.0 / rho;
  double T_old, T_new = (1.0 / rho - mb) * (p + ma * rho * rho) / mR;
  int it = 0;
  do {
    T_old = T_new;
    double f = mR * T_old / (v - mb) - ma / sqrt(T_old) / v / (v + mb) - p;
    double df = mR / (v - mb) + 0.5 * ma / T_old / sqrt(T_old) / v / (v + mb);
    T_new = T_old - f / df;
    if (++it > RK_MAXIT) { if (eflag) *eflag = 1; break; }
  } while (fabs(T_new - T_old) / fabs(T_old) > 1e-14);
  return T_new;
}
__device__ inline double rk_e_from_rho_p(const Phys& P, double rho, double p, int* eflag) {
  double v = 1.0 / rho;
  double T = rk_T_from_rho_p(P, rho, p, eflag);
  double U = log(v / (v + P.rk_b));
  return P.rk_cv * T + 1.5 * P.rk_alfa / P.rk_b * U / sqrt(T);
}
__device__ inline double rk_p_from_rho_e(const Phys& P, double rho, double e, int* eflag) {
  const double mR = P.rk_R, ma = P.rk_alfa, mb = P.rk_b, mcv = P.rk_cv;
  double v = 1.0 / rho;
  double T_old, T_new = (e + rho * ma) / mcv;
  double U = log(v / (v + mb));
  int it = 0;
  do {
    T_old = T_new;
    double f = mcv * T_old + 1.5 * ma / mb * U / sqrt(T_old) - e;
    double df = mcv - 0.75 * ma / mb * U / T_old / sqrt(T_old);
    T_new = T_old - f / df;
    if (++it > RK_MAXIT) { if (eflag) *eflag = 1; break; }
  } while (fabs(T_new - T_old) / fabs(T_old) > 1e-14);
  double a = ma / sqrt(T_new);
  return mR * T_new / (v - mb) - a / v / (v + mb);
}
__device__ inline double rk_sound(const Phys& P, double rho, double p, int* eflag) {
  const double mR = P.rk_R, ma = P.rk_alfa, mb = P.rk_b, mcv = P.rk_cv;
  double v = 1.0 / rho;
  double T = rk_T_from_rho_p(P, rho, p, eflag);
  double U = log(v / (v + mb));
  double p_T = mR / (v - mb) + ma / (2.0 * T * sqrt(T) * v * (v + mb));
  double e_T = mcv - 0.75 * ma / mb * U / (T * sqrt(T));
  double p_v = -mR * T / ((v - mb) * (v - mb)) + ma * (2.0 * v + mb) / (sqrt(T) * v * v * ((v + mb) * (v + mb)));
  double e_v = 1.5 * ma / mb * (1.0 / v - 1.0 / (v + mb)) / sqrt(T);
  double a2 = p / (rho * rho) * (p_T / e_T) + v * v * (e_v - e_T * (p_v / p_T));
  return sqrt(a2);
}
__device__ __forceinline__ double eos_p(const Phys& P, double rho, double e, int* ef) {
  return P.eos == IMEX_EOS_IDEAL ? (P.gamma - 1.0) * rho * e : rk_p_from_rho_e(P, rho, e, ef);
}
__device__ __forceinline__ double eos_e(const Phys& P, double rho, double p, int* ef) {
  return P.eos == IMEX_EOS_IDEAL ? p / ((P.gamma - 1.0) * rho) : rk_e_from_rho_p(P, rho, p, ef);
}
__device__ __forceinline__ double eos_a(const Phys& P, double rho, double p, int* ef) {
  return P.eos == IMEX_EOS_IDEAL ? sqrt(P.gamma * p / rho) : rk_sound(P, rho, p, ef);
}
__device__ __forceinline__ double eos_fast(const Phys& P, double rho, double p, double Bn, double B2, int* ef) {
  double a = eos_a(P, rho, p, ef);
  double a2 = a * a;
  double t = a2 + B2 / rho;
  double disc = t * t - 4.0 * a2 * Bn * Bn / rho;
  if (disc < 0) disc = 0;
  return sqrt(0.5 * (t + sqrt(disc)));
}
__device__ __forceinline__ double Vbase(const Phys& P, double p, double rho, int* ef) {
  return P.eos == IMEX_EOS_IDEAL ? p / (P.gamma - 1.0) : rho * rk_e_from_rho_p(P, rho, p, ef);
}
__device__ __forceinline__ double Vbase_slope(const Phys& P, double p, double rho, int* ef) {
  if (P.eos == IMEX_EOS_IDEAL) return 1.0 / (P.gamma - 1.0);
  const double d = 1e-8;
  double ep = rk_e_from_rho_p(P, rho, p + d, ef);
  double em = rk_e_from_rho_p(P, rho, p - d, ef);
  return rho * (ep - em) / (2.0 * d);
}
__device__ __forceinline__ double V_mod(const Phys& P, double p, double rho, double Mbig, int* ef) {
  if (!P.floor_on) return Vbase(P, p, rho, ef);
  return (p >= P.p_floor) ? Vbase(P, p, rho, ef) : Vbase(P, P.p_floor, rho, ef) + Mbig * (p - P.p_floor);
}
__device__ __forceinline__ double V_slope(const Phys& P, double p, double rho, double Mbig, int* ef) {
  if (!P.floor_on) return Vbase_slope(P, p, rho, ef);
  return (p >= P.p_floor) ? Vbase_slope(P, p, rho, ef) : Mbig;
}

// ---------------------------------------------------------------- pointwise state (state.cpp:206-240, reconstruction.cpp:291-304)
__device__ __forceinline__ double calecon(const double* c) {
  double rho = c[0];
  double vx = c[1] / c[0], vy = c[2] / c[0], vz = c[3] / c[0];
  double B2 = c[5] * c[5] + c[6] * c[6] + c[7] * c[7];
  double v2 = vx * vx + vy * vy + vz * vz;
  return (c[4] - 0.5 * rho * v2 - 0.5 * B2) / rho;
}
__device__ __forceinline__ double pressure_of(const Phys& P, const double* c, int* ef) {
  return eos_p(P, c[0], calecon(c), ef);
}
__device__ __forceinline__ void con2pri(const Phys& P, double* pri, const double* c, int* ef) {
  double e = calecon(c);
  pri[0] = c[0];
  pri[1] = c[1] / c[0]; pri[2] = c[2] / c[0]; pri[3] = c[3] / c[0];
  pri[4] = eos_p(P, c[0], e, ef);
  pri[5] = c[5]; pri[6] = c[6]; pri[7] = c[7];
}
__device__ __forceinline__ void pri2con(const Phys& P, double* c, const double* pri, int* ef) {
  double e = eos_e(P, pri[0], pri[4], ef);
  double B2 = pri[5] * pri[5] + pri[6] * pri[6] + pri[7] * pri[7];
  double v2 = pri[1] * pri[1] + pri[2] * pri[2] + pri[3] * pri[3];
  c[0] = pri[0];
  c[1] = pri[0] * pri[1]; c[2] = pri[0] * pri[2]; c[3] = pri[0] * pri[3];
  c[4] = pri[0] * e + 0.5 * pri[0] * v2 + 0.5 * B2;
  c[5] = pri[5]; c[6] = pri[6]; c[7] = pri[7];
}
// calenthalpy (timestep.cpp:224-230): e + p/rho
__device__ __forceinline__ double enthalpy_of(const Phys& P, const double* c, int* ef) {
  double e = calecon(c);
  double p = eos_p(P, c[0], e, ef);
  return e + p / c[0];
}

// ---------------------------------------------------------------- block reductions (wave64)
// Wave-wide sum with DPP lane moves (register to register, no LDS): an inclusive scan within each
// row of 16 lanes (row_shr 1, 2, 4, 8, shifted-in lanes read zero), then the row totals carried
// across rows (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3); lane 63 ends with
// the total, which is returned to every lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0x111, 0xf>(v);      // row_shr:1
  v += dpp_move<0x112, 0xf>(v);      // row_shr:2
  v += dpp_move<0x114, 0xf>(v);      // row_shr:4
  v += dpp_move<0x118, 0xf>(v);      // row_shr:8
  v += dpp_move<0x142, 0xa>(v);      // row_bcast:15
  v += dpp_move<0x143, 0xc>(v);      // row_bcast:31
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_down(v, o, 64));
  return v;
}
// Sum over a block of up to 1024 threads; result valid in thread 0. `sh` holds >= 16 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) for (int w = 0; w < nw; w++) r += sh[w];
  return r;
}

// This thread's share of cal_divergence_error's sum (imex_subsystems.cpp:630-660, StaggerCT branch):
// sum of dx dy (div B)^2 over interior cells t = b * blockDim + tid, t += g * blockDim, from the staggered
// faces in the Array layout (face i+1/2 of row j at i + j * NX).
__device__ __forceinline__ double divb_thread_sum(const double* __restrict__ bx, const double* __restrict__ by, int nx, int NX,
                                                  long long N, double dx, double dy, long long b, long long g) {
  double acc = 0.0;
  for (long long t = b * blockDim.x + threadIdx.x; t < N; t += g * blockDim.x) {
    int i = (int)(t % nx) + GC, j = (int)(t / nx) + GC;
    long long c = (long long)i + (long long)j * NX;
    double v = (bx[c] - bx[c - 1]) / dx + (by[c] - by[c - NX]) / dy;
    acc += dx * dy * v * v;
  }
  return acc;
}

// "Last block" detection for single-pass grid reductions.  Every block calls
// it after writing its partial; the block that returns true sees all partials.
__device__ __forceinline__ bool last_block_arrive(unsigned* counter) {
  __shared__ int s_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (s_last) __threadfence();
  return s_last != 0;
}
