// adjoint_dev.hpp — device side of the discrete adjoint of the IMEX-RK2 step.
//
// Scope: periodic grids, ideal gas or Redlich-Kwong, Minmod, VanLeer or WENO5 reconstruction of the
// conservative state, the nested Newton with the pressure floor active or not, EMF-consistent CT, FP64.
// On a periodic grid every ghost value is a copy of an interior value, so each field is treated
// as a field on the torus of interior cells, read through a wrapped index (0-based interior
// coordinates).
//
// Each nonlinear piece of the stage is written once more as an element function templated on
// its scalar type: with T = double it evaluates the same formula as the forward kernels (and the
// reference lines they cite); with T = DN<K> (a forward-mode dual number carrying K partial
// derivatives) it yields the derivatives of the element's outputs with respect to the K inputs of
// one seeded cell.  A vector-Jacobian product then loops, per element, over the cells of the
// element's stencil, seeds one cell at a time and adds (output cotangent . derivative) into that
// cell's input cotangent with atomics.  Scalars (h, a_max, B_max, v_max) get their cotangents
// from a second pass that seeds them instead, reduced over the grid.
#pragma once
#include "kernels.hpp"

// ---------------------------------------------------------------- dual numbers
template <int K>
struct DN {
  double v, g[K];
  __device__ DN() {}
  __device__ DN(double x) : v(x) {
#pragma unroll
    for (int k = 0; k < K; k++) g[k] = 0.0;
  }
};
template <class T> struct IsDual { static constexpr bool value = false; };
template <int K> struct IsDual<DN<K>> { static constexpr bool value = true; };

__device__ __forceinline__ double val(double x) { return x; }
template <int K> __device__ __forceinline__ double val(const DN<K>& x) { return x.v; }

template <int K> __device__ __forceinline__ DN<K> operator+(const DN<K>& a, const DN<K>& b) {
  DN<K> r; r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] + b.g[k];
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator-(const DN<K>& a, const DN<K>& b) {
  DN<K> r; r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] - b.g[k];
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator-(const DN<K>& a) {
  DN<K> r; r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = -a.g[k];
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator*(const DN<K>& a, const DN<K>& b) {
  DN<K> r; r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator/(const DN<K>& a, const DN<K>& b) {
  DN<K> r; r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = (a.g[k] - r.v * b.g[k]) * ib;
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator+(const DN<K>& a, double b) { DN<K> r = a; r.v += b; return r; }
template <int K> __device__ __forceinline__ DN<K> operator+(double a, const DN<K>& b) { DN<K> r = b; r.v += a; return r; }
template <int K> __device__ __forceinline__ DN<K> operator-(const DN<K>& a, double b) { DN<K> r = a; r.v -= b; return r; }
template <int K> __device__ __forceinline__ DN<K> operator-(double a, const DN<K>& b) {
  DN<K> r; r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = -b.g[k];
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator*(const DN<K>& a, double b) {
  DN<K> r; r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] * b;
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator*(double a, const DN<K>& b) { return b * a; }
template <int K> __device__ __forceinline__ DN<K> operator/(const DN<K>& a, double b) {
  DN<K> r; r.v = a.v / b;
  const double ib = 1.0 / b;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] * ib;
  return r;
}
template <int K> __device__ __forceinline__ DN<K> operator/(double a, const DN<K>& b) {
  DN<K> r; r.v = a / b.v;
  const double s = -r.v / b.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = b.g[k] * s;
  return r;
}
template <int K> __device__ __forceinline__ DN<K> sqrt(const DN<K>& a) {
  DN<K> r; r.v = ::sqrt(a.v);
  const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] * s;
  return r;
}
template <int K> __device__ __forceinline__ DN<K> fabs(const DN<K>& a) { return a.v < 0.0 ? -a : a; }
template <int K> __device__ __forceinline__ DN<K> log(const DN<K>& a) {
  DN<K> r; r.v = ::log(a.v);
  const double s = 1.0 / a.v;
#pragma unroll
  for (int k = 0; k < K; k++) r.g[k] = a.g[k] * s;
  return r;
}

// branch helpers that work for both scalar types
template <class T> __device__ __forceinline__ T tmax(const T& a, const T& b) { return val(a) >= val(b) ? a : b; }
template <class T> __device__ __forceinline__ T tmin(const T& a, const T& b) { return val(a) <= val(b) ? a : b; }

// ---------------------------------------------------------------- EOS for both scalar types
// (include/config.hpp:130-198): ideal gas, or Redlich-Kwong with its temperature solves (Newton to a
// relative step of 1e-14, as the forward's).  With dual numbers the Newton iteration carries the
// derivatives along; they converge with the value (each step solves the linearised equation again).
template <class T>
__device__ __forceinline__ T rkT_from_rho_p(const Phys& P, const T& rho, const T& p) {
  const double mR = P.rk_R, ma = P.rk_alfa, mb = P.rk_b;
  const T v = 1.0 / rho;
  T T_new = (1.0 / rho - mb) * (p + ma * rho * rho) / mR, T_old;
  int it = 0;
  do {
    T_old = T_new;
    const T f = mR * T_old / (v - mb) - ma / sqrt(T_old) / v / (v + mb) - p;
    const T df = mR / (v - mb) + 0.5 * ma / T_old / sqrt(T_old) / v / (v + mb);
    T_new = T_old - f / df;
  } while (++it <= RK_MAXIT && ::fabs(val(T_new) - val(T_old)) / ::fabs(val(T_old)) > 1e-14);
  return T_new;
}
template <class T>
__device__ __forceinline__ T rkT_from_rho_e(const Phys& P, const T& rho, const T& e) {
  const double ma = P.rk_alfa, mb = P.rk_b, mcv = P.rk_cv;
  const T v = 1.0 / rho;
  const T U = log(v / (v + mb));
  T T_new = (e + rho * ma) / mcv, T_old;
  int it = 0;
  do {
    T_old = T_new;
    const T f = mcv * T_old + 1.5 * ma / mb * U / sqrt(T_old) - e;
    const T df = mcv - 0.75 * ma / mb * U / T_old / sqrt(T_old);
    T_new = T_old - f / df;
  } while (++it <= RK_MAXIT && ::fabs(val(T_new) - val(T_old)) / ::fabs(val(T_old)) > 1e-14);
  return T_new;
}
template <class T>
__device__ __forceinline__ T eos_e_t(const Phys& P, const T& rho, const T& p) {
  if (P.eos == IMEX_EOS_IDEAL) return p / ((P.gamma - 1.0) * rho);
  const T v = 1.0 / rho;
  const T Tm = rkT_from_rho_p(P, rho, p);
  const T U = log(v / (v + P.rk_b));
  return P.rk_cv * Tm + 1.5 * P.rk_alfa / P.rk_b * U / sqrt(Tm);
}
template <class T>
__device__ __forceinline__ T eos_p_t(const Phys& P, const T& rho, const T& e) {
  if (P.eos == IMEX_EOS_IDEAL) return (P.gamma - 1.0) * rho * e;
  const T v = 1.0 / rho;
  const T Tm = rkT_from_rho_e(P, rho, e);
  const T a = P.rk_alfa / sqrt(Tm);
  return P.rk_R * Tm / (v - P.rk_b) - a / v / (v + P.rk_b);
}
template <class T>
__device__ __forceinline__ T eos_a_t(const Phys& P, const T& rho, const T& p) {
  if (P.eos == IMEX_EOS_IDEAL) return sqrt(P.gamma * p / rho);
  const double mR = P.rk_R, ma = P.rk_alfa, mb = P.rk_b, mcv = P.rk_cv;
  const T v = 1.0 / rho;
  const T Tm = rkT_from_rho_p(P, rho, p);
  const T U = log(v / (v + mb));
  const T p_T = mR / (v - mb) + ma / (2.0 * Tm * sqrt(Tm) * v * (v + mb));
  const T e_T = mcv - 0.75 * ma / mb * U / (Tm * sqrt(Tm));
  const T p_v = -mR * Tm / ((v - mb) * (v - mb)) + ma * (2.0 * v + mb) / (sqrt(Tm) * v * v * ((v + mb) * (v + mb)));
  const T e_v = 1.5 * ma / mb * (1.0 / v - 1.0 / (v + mb)) / sqrt(Tm);
  const T a2 = p / (rho * rho) * (p_T / e_T) + v * v * (e_v - e_T * (p_v / p_T));
  return sqrt(a2);
}
// Vbase (config.hpp:206-217): rho e(rho, p); its slope is the forward's central difference in p
template <class T>
__device__ __forceinline__ T vbase_t(const Phys& P, const T& p, const T& rho) {
  if (P.eos == IMEX_EOS_IDEAL) return p / (P.gamma - 1.0);
  return rho * eos_e_t(P, rho, p);
}
template <class T>
__device__ __forceinline__ T vbase_slope_t(const Phys& P, const T& p, const T& rho) {
  if (P.eos == IMEX_EOS_IDEAL) return T(1.0 / (P.gamma - 1.0));
  const double d = 1e-8;
  return rho * (eos_e_t(P, rho, p + d) - eos_e_t(P, rho, p - d)) / (2.0 * d);
}
// V_mod (config.hpp:231-234) with the floor on
template <class T>
__device__ __forceinline__ T vmod_t(const Phys& P, const T& p, const T& rho, double Mbig) {
  if (val(p) >= P.p_floor) return vbase_t(P, p, rho);
  return vbase_t(P, T(P.p_floor), rho) + Mbig * (p - P.p_floor);
}

// ---------------------------------------------------------------- torus geometry and fields
struct AD {
  int nx, ny, NX;
  long long PS;
  double dx, dy;
};
__device__ __forceinline__ int wrapi(int i, int n) { i %= n; return i < 0 ? i + n : i; }
__device__ __forceinline__ long long tix(const AD& a, int i, int j) {
  return (long long)(wrapi(i, a.nx) + GC) + (long long)(wrapi(j, a.ny) + GC) * a.NX;
}

#define ADJ_NF 6
// Input fields of an element function (plane 0 of each; component k at + k*PS) and the
// cotangent buffers they accumulate into (same layout; null: no cotangent wanted).
struct AFields {
  const double* f[ADJ_NF];
  double* bar[ADJ_NF];
};

// Step scalars as seen by the element functions.
enum { SC_H, SC_AMAX, SC_BMX, SC_BMY, SC_VXM, SC_VYM, SC_N };
template <class T>
struct Sc { T s[SC_N]; };

// Reader of field values for element functions: seeds the inputs of one cell (si, sj) when T is a
// dual number; Fn::slot(fid, k) gives the derivative slot of component k of field fid (-1: none).
template <class T, class Fn>
struct Rd {
  const AD* a;
  const AFields* F;
  int si, sj;
  __device__ __forceinline__ T operator()(int fid, int k, int i, int j) const {
    const int wi = wrapi(i, a->nx), wj = wrapi(j, a->ny);
    T r(F->f[fid][k * a->PS + (long long)(wi + GC) + (long long)(wj + GC) * a->NX]);
    if constexpr (IsDual<T>::value) {
      if (wi == si && wj == sj) {
        const int s = Fn::slot(fid, k);
        if (s >= 0) r.g[s] = 1.0;
      }
    }
    return r;
  }
};

// ---------------------------------------------------------------- generic kernels
// Input cotangents: for every element e (interior cell (i, j)) with output cotangents obar[o]
// (plane o of `obar`), and every cell c of its stencil, bar[c] += sum_o obar[o] d out_o / d in_c.
template <class Fn>
__global__ void __launch_bounds__(128) k_vjp(Fn fn, AD a, AFields F, const double* __restrict__ obar, Sc<double> sc) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  const long long e = tix(a, i, j);
  double ob[Fn::NOUT];
  bool any = false;
#pragma unroll
  for (int o = 0; o < Fn::NOUT; o++) { ob[o] = obar[o * a.PS + e]; any |= ob[o] != 0.0; }
  if (!any) return;
  typedef DN<Fn::K> T;
  Sc<T> st;
#pragma unroll
  for (int q = 0; q < SC_N; q++) st.s[q] = T(sc.s[q]);
  for (int s = 0; s < Fn::NS; s++) {
    int oi, oj;
    Fn::off(s, oi, oj);
    Rd<T, Fn> rd{&a, &F, wrapi(i + oi, a.nx), wrapi(j + oj, a.ny)};
    T out[Fn::NOUT];
    fn.eval(rd, st, i, j, out);
    const long long c = tix(a, i + oi, j + oj);
    for (int q = 0; q < Fn::K; q++) {
      double g = 0.0;
#pragma unroll
      for (int o = 0; o < Fn::NOUT; o++) g += ob[o] * out[o].g[q];
      if (g != 0.0) {
        int fid, k;
        Fn::which(q, fid, k);
        double* b = F.bar[fid];
        if (b) atomicAdd(b + k * a.PS + c, g);
      }
    }
  }
}

// Jacobian of an element function that is linear in the inputs it differentiates (the magnetic
// operator in the unknown B): J[((s NOUT + o) K + q) N + e] = d out_o(e) / d input_q(cell s of e's
// stencil), N = nx ny, e the compact interior index.  Formed once per solve by the same dual-number
// evaluations as one k_vjp; each transposed application is then the gather below.
template <class Fn>
__global__ void __launch_bounds__(128) k_jac(Fn fn, AD a, AFields F, Sc<double> sc, double* __restrict__ J) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.nx * a.ny;
  if (n >= N) return;
  const int i = n % a.nx, j = n / a.nx;
  typedef DN<Fn::K> T;
  Sc<T> st;
#pragma unroll
  for (int q = 0; q < SC_N; q++) st.s[q] = T(sc.s[q]);
  for (int s = 0; s < Fn::NS; s++) {
    int oi, oj;
    Fn::off(s, oi, oj);
    Rd<T, Fn> rd{&a, &F, wrapi(i + oi, a.nx), wrapi(j + oj, a.ny)};
    T out[Fn::NOUT];
    fn.eval(rd, st, i, j, out);
#pragma unroll
    for (int o = 0; o < Fn::NOUT; o++)
#pragma unroll
      for (int q = 0; q < Fn::K; q++) J[((long long)(s * Fn::NOUT + o) * Fn::K + q) * N + n] = out[o].g[q];
  }
}

// (J^T y) at cell c: the elements e whose stencil holds c (c = e + off(s)), each contributing
// sum_o J_e[s][o][q] y_o(e); written to plane q of `out` (not added).  No atomics: one thread per
// cell, a fixed summation order.
template <class Fn>
__global__ void __launch_bounds__(128) k_jac_t(AD a, const double* __restrict__ J, const double* __restrict__ y,
                                               double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.nx * a.ny;
  if (n >= N) return;
  const int i = n % a.nx, j = n / a.nx;
  double acc[Fn::K];
#pragma unroll
  for (int q = 0; q < Fn::K; q++) acc[q] = 0.0;
  for (int s = 0; s < Fn::NS; s++) {
    int oi, oj;
    Fn::off(s, oi, oj);
    const int ei = wrapi(i - oi, a.nx), ej = wrapi(j - oj, a.ny);
    const int e = ei + a.nx * ej;
    const long long te = tix(a, ei, ej);
#pragma unroll
    for (int o = 0; o < Fn::NOUT; o++) {
      const double yo = y[o * a.PS + te];
#pragma unroll
      for (int q = 0; q < Fn::K; q++) acc[q] += J[((long long)(s * Fn::NOUT + o) * Fn::K + q) * N + e] * yo;
    }
  }
  const long long c = tix(a, i, j);
#pragma unroll
  for (int q = 0; q < Fn::K; q++) out[q * a.PS + c] = acc[q];
}

// Scalar cotangents: sbar[q] += sum_e sum_o obar[o] d out_o / d scalar_q.
template <class Fn>
__global__ void __launch_bounds__(128) k_vjp_scalars(Fn fn, AD a, AFields F, const double* __restrict__ obar, Sc<double> sc,
                                                     double* sbar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[SC_N];
#pragma unroll
  for (int q = 0; q < SC_N; q++) acc[q] = 0.0;
  if (n < a.nx * a.ny) {
    const int i = n % a.nx, j = n / a.nx;
    const long long e = tix(a, i, j);
    typedef DN<SC_N> T;
    Sc<T> st;
#pragma unroll
    for (int q = 0; q < SC_N; q++) { st.s[q] = T(sc.s[q]); st.s[q].g[q] = 1.0; }
    Rd<T, Fn> rd{&a, &F, -1, -1};
    T out[Fn::NOUT];
    fn.eval(rd, st, i, j, out);
#pragma unroll
    for (int o = 0; o < Fn::NOUT; o++) {
      const double ob = obar[o * a.PS + e];
#pragma unroll
      for (int q = 0; q < SC_N; q++) acc[q] += ob * out[o].g[q];
    }
  }
  __shared__ double sh[SC_N][2];
  for (int q = 0; q < SC_N; q++) {
    double v = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < SC_N) {
    double v = sh[threadIdx.x][0] + sh[threadIdx.x][1];
    if (v != 0.0) atomicAdd(sbar + threadIdx.x, v);
  }
}

// Plain evaluation (T = double) into NOUT planes of `out`.
template <class Fn>
__global__ void __launch_bounds__(128) k_eval(Fn fn, AD a, AFields F, double* __restrict__ out, Sc<double> sc) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  Rd<double, Fn> rd{&a, &F, -1, -1};
  double o[Fn::NOUT];
  fn.eval(rd, sc, i, j, o);
  const long long e = tix(a, i, j);
#pragma unroll
  for (int q = 0; q < Fn::NOUT; q++) out[q * a.PS + e] = o[q];
}
