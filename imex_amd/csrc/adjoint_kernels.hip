// adjoint_kernels.hip — element functions and kernels of the discrete adjoint (see adjoint_dev.hpp).
//
// Each element function restates the forward formula of one piece of the stage (the line cited
// with it) for a generic scalar type; the forward kernels and the oracle compute the same
// formulas.  Linear pieces (flux divergence, momentum update, CT, stage combination) have their
// transposes written out directly.
#include "adjoint.hpp"

#define AB 128
static inline int nblk(const AD& a) { return (a.nx * a.ny + AB - 1) / AB; }

// ================================================================ element functions
// ---- reconstruction of one cell (reconstruction.cpp:46-103): left / right face values
template <class T>
__device__ __forceinline__ void rc_cell(int recon, bool central, const T& um, const T& u0, const T& up, T& wl, T& wr) {
  const T sL = u0 - um, sR = up - u0;
  T c(0.0);
  if (recon == IMEX_RECON_MINMOD) {
    if (central) c = 0.5 * (sL + sR);
    else if (val(sL) * val(sR) <= 0.0) c = T(0.0);
    else if (::fabs(val(sR)) <= ::fabs(val(sL))) c = sR;
    else c = sL;
    wl = u0 - 0.5 * c;
    wr = u0 + 0.5 * c;
  } else {
    if (::fabs(val(sR)) < 1e-16) c = T(0.0);
    else {
      const T r = sL / sR;
      if (val(r) < 0.0) c = T(0.0);
      else c = tmin(2.0 / (1.0 + r), (2.0 * r) / (1.0 + r));
    }
    wl = u0 - 0.5 * c * sL;
    wr = u0 + 0.5 * c * sR;
  }
}

// ---- WENO5 reconstruction of one cell (reconstruction.cpp:184-266, Jiang-Shu weights, eps 1e-6,
// pow(x, 2) = x x): left / right face values from the five cells um2 .. up2
template <class T>
__device__ __forceinline__ void rc5_cell(const T& um2, const T& um1, const T& u0, const T& up1, const T& up2, T& wl, T& wr) {
  const T t0 = um2 - 2.0 * um1 + u0, t1 = um2 - 4.0 * um1 + 3.0 * u0;
  const T t2 = um1 - 2.0 * u0 + up1, t3 = um1 - up1;
  const T t4 = u0 - 2.0 * up1 + up2, t5 = 3.0 * u0 - 4.0 * up1 + up2;
  const T b0 = (13.0 / 12.0) * (t0 * t0) + 0.25 * (t1 * t1);
  const T b1 = (13.0 / 12.0) * (t2 * t2) + 0.25 * (t3 * t3);
  const T b2 = (13.0 / 12.0) * (t4 * t4) + 0.25 * (t5 * t5);
  const T e0 = (1e-6 + b0) * (1e-6 + b0), e1 = (1e-6 + b1) * (1e-6 + b1), e2 = (1e-6 + b2) * (1e-6 + b2);
  // left face: linear weights 0.3, 0.6, 0.1 on the three candidate stencils
  T a0 = 0.3 / e0, a1 = 0.6 / e1, a2 = 0.1 / e2;
  T q0 = (-1.0 / 6.0) * um2 + (5.0 / 6.0) * um1 + (1.0 / 3.0) * u0;
  T q1 = (1.0 / 3.0) * um1 + (5.0 / 6.0) * u0 - (1.0 / 6.0) * up1;
  T q2 = (11.0 / 6.0) * u0 - (7.0 / 6.0) * up1 + (1.0 / 3.0) * up2;
  wl = (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
  // right face: linear weights 0.1, 0.6, 0.3
  a0 = 0.1 / e0; a1 = 0.6 / e1; a2 = 0.3 / e2;
  q0 = (1.0 / 3.0) * um2 - (7.0 / 6.0) * um1 + (11.0 / 6.0) * u0;
  q1 = (-1.0 / 6.0) * um1 + (5.0 / 6.0) * u0 + (1.0 / 3.0) * up1;
  q2 = (1.0 / 3.0) * u0 + (5.0 / 6.0) * up1 - (1.0 / 6.0) * up2;
  wr = (a0 * q0 + a1 * q1 + a2 * q2) / (a0 + a1 + a2);
}

// ---- physical flux (state.cpp:27-44, 86-106, 298-322)
template <int DIR, class T>
__device__ __forceinline__ void phys_flux(int kind, const T* Q, T* f) {
  const T rho = Q[0], vx = Q[1] / Q[0], vy = Q[2] / Q[0], vz = Q[3] / Q[0];
  const T kin = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
  const T vn = DIR == 0 ? vx : vy;
  f[0] = rho * vn; f[1] = rho * vx * vn; f[2] = rho * vy * vn; f[3] = rho * vz * vn; f[4] = kin * vn;
  if (kind == 1) {
    const T B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
    const T vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) / Q[0];
    if (DIR == 0) {
      f[1] = f[1] + (B2 * 0.5 - Q[5] * Q[5]); f[2] = f[2] - Q[5] * Q[6]; f[3] = f[3] - Q[5] * Q[7];
      f[4] = f[4] + (Q[1] / Q[0] * B2 - Q[5] * vB);
    } else {
      f[1] = f[1] - Q[5] * Q[6]; f[2] = f[2] + (B2 * 0.5 - Q[6] * Q[6]); f[3] = f[3] - Q[6] * Q[7];
      f[4] = f[4] + (Q[2] / Q[0] * B2 - Q[6] * vB);
    }
  }
}

// ---- Rusanov face flux (flux.cpp:202-243, 329-367); element (i, j) = face between cells
// (i-1, j) and (i, j) (DIR 0) or (i, j-1) and (i, j) (DIR 1).  Fields: 0 Qe (0..4), 1 u (B 5..7).
// W5: WENO5 reconstruction (stencil i-3 .. i+2), else Minmod / VanLeer (i-2 .. i+1).
template <int DIR, bool W5 = false>
struct FnFlux {
  static constexpr int K = 8, NOUT = 5, NS = W5 ? 6 : 4;
  int kind, recon, central;
  __device__ static void off(int s, int& oi, int& oj) {
    const int o = W5 ? s - 3 : s - 2;
    oi = DIR == 0 ? o : 0; oj = DIR == 0 ? 0 : o;
  }
  __device__ static int slot(int fid, int k) { return fid == 0 ? (k < 5 ? k : -1) : (k >= 5 ? k : -1); }
  __device__ static void which(int q, int& fid, int& k) { fid = q < 5 ? 0 : 1; k = q; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* F) const {
    const int nrec = kind == 0 ? 4 : 5;
    const int di = DIR == 0 ? 1 : 0, dj = DIR == 0 ? 0 : 1;
    const bool cen = DIR == 1 && central;
    T QL[8], QR[8];
    for (int k = 0; k < 5; k++) {
      if (k < nrec) {
        T wl, wr, wl2, wr2;
        if constexpr (W5) {
          T a[6];
          for (int o = 0; o < 6; o++) a[o] = rd(0, k, i + (o - 3) * di, j + (o - 3) * dj);
          rc5_cell(a[0], a[1], a[2], a[3], a[4], wl, wr);
          rc5_cell(a[1], a[2], a[3], a[4], a[5], wl2, wr2);
        } else {
          const T a0 = rd(0, k, i - 2 * di, j - 2 * dj), a1 = rd(0, k, i - di, j - dj), a2 = rd(0, k, i, j),
                  a3 = rd(0, k, i + di, j + dj);
          rc_cell(recon, cen, a0, a1, a2, wl, wr);
          rc_cell(recon, cen, a1, a2, a3, wl2, wr2);
        }
        QL[k] = wr; QR[k] = wl2;
      } else { QL[k] = T(0.0); QR[k] = T(0.0); }
    }
    for (int k = 5; k < 8; k++) {
      if (kind == 1) { QL[k] = rd(1, k, i - di, j - dj); QR[k] = rd(1, k, i, j); }
      else { QL[k] = T(0.0); QR[k] = T(0.0); }
    }
    T fL[5], fR[5];
    phys_flux<DIR>(kind, QL, fL);
    phys_flux<DIR>(kind, QR, fR);
    const T s = sc.s[SC_AMAX];
    for (int k = 0; k < 5; k++) F[k] = 0.5 * (fL[k] + fR[k]) - 0.5 * s * (QR[k] - QL[k]);
  }
};

// ---- magnetic right-hand side (imex_subsystems.cpp:549-613).  Fields: 0 u (0..3), 1 Qe (B), 2 p(Qe), 3 Qs (B).
struct FnMagRHS {
  static constexpr int K = 11, NOUT = 3, NS = 9;
  double dx, dy;
  __device__ static void off(int s, int& oi, int& oj) { oi = s % 3 - 1; oj = s / 3 - 1; }
  __device__ static int slot(int fid, int k) {
    switch (fid) {
      case 0: return k < 4 ? k : -1;
      case 1: return k >= 5 ? 4 + k - 5 : -1;
      case 2: return k == 0 ? 7 : -1;
      default: return k >= 5 ? 8 + k - 5 : -1;
    }
  }
  __device__ static void which(int q, int& fid, int& k) {
    if (q < 4) { fid = 0; k = q; } else if (q < 7) { fid = 1; k = q + 1; } else if (q == 7) { fid = 2; k = 0; } else { fid = 3; k = q - 3; }
  }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* out) const {
    const T h = sc.s[SC_H];
    auto U = [&](int a, int b, int k) { return rd(0, k, a, b); };
    auto Q = [&](int a, int b, int k) { return rd(1, k, a, b); };
    auto PE = [&](int a, int b) { return rd(2, 0, a, b); };
    const T u0 = U(i, j, 0);
    const T rho_p0 = 0.5 * (u0 + U(i + 1, j, 0)), rho_m0 = 0.5 * (u0 + U(i - 1, j, 0));
    const T rho_0p = 0.5 * (u0 + U(i, j + 1, 0)), rho_0m = 0.5 * (u0 + U(i, j - 1, 0));
    T m[3][4], n[3][4];
    for (int c = 0; c < 3; c++) {
      const T uc = U(i, j, 1 + c), qc = Q(i, j, 5 + c);
      m[c][0] = 0.5 * (uc + U(i + 1, j, 1 + c)); m[c][1] = 0.5 * (uc + U(i - 1, j, 1 + c));
      m[c][2] = 0.5 * (uc + U(i, j + 1, 1 + c)); m[c][3] = 0.5 * (uc + U(i, j - 1, 1 + c));
      n[c][0] = 0.5 * (qc + Q(i + 1, j, 5 + c)); n[c][1] = 0.5 * (qc + Q(i - 1, j, 5 + c));
      n[c][2] = 0.5 * (qc + Q(i, j + 1, 5 + c)); n[c][3] = 0.5 * (qc + Q(i, j - 1, 5 + c));
    }
    const T p00 = PE(i, j), pp0 = PE(i + 1, j), pm0 = PE(i - 1, j), p0p = PE(i, j + 1), p0m = PE(i, j - 1);
    const T p_pp = 0.25 * (p00 + pp0 + p0p + PE(i + 1, j + 1));
    const T p_pm = 0.25 * (p00 + pp0 + p0m + PE(i + 1, j - 1));
    const T p_mp = 0.25 * (p00 + pm0 + p0p + PE(i - 1, j + 1));
    const T p_mm = 0.25 * (p00 + pm0 + p0m + PE(i - 1, j - 1));
    const double dx_ = dx, dy_ = dy;
    out[0] = rd(3, 5, i, j) - (h / dy_) * ((n[0][2] / rho_0p * (m[1][2] - (h / dy_) * (p0p - p00)) - n[1][2] / rho_0p * (m[0][2] - h / dx_ * (p_pp - p_mp)))
                                         - (n[0][3] / rho_0m * (m[1][3] - (h / dy_) * (p00 - p0m)) - n[1][3] / rho_0m * (m[0][3] - h / dx_ * (p_pm - p_mm))));
    out[1] = rd(3, 6, i, j) - (h / dx_) * ((n[1][0] / rho_p0 * (m[0][0] - (h / dx_) * (pp0 - p00)) - n[0][0] / rho_p0 * (m[1][0] - h / dy_ * (p_pp - p_pm)))
                                         - (n[1][1] / rho_m0 * (m[0][1] - (h / dx_) * (p00 - pm0)) - n[0][1] / rho_m0 * (m[1][1] - h / dy_ * (p_mp - p_mm))));
    const T Bz_1 = (n[2][0] / rho_p0 * (m[0][0] - h / dx_ * (pp0 - p00)) - n[0][0] / rho_p0 * m[2][0])
                 - (n[2][1] / rho_m0 * (m[0][1] - h / dx_ * (p00 - pm0)) - n[0][1] / rho_m0 * m[2][1]);
    const T Bz_2 = (n[2][2] / rho_0p * (m[1][2] - h / dy_ * (p0p - p00)) - n[1][2] / rho_0p * m[2][2])
                 - (n[2][3] / rho_0m * (m[1][3] - h / dy_ * (p00 - p0m)) - n[1][3] / rho_0m * m[2][3]);
    out[2] = rd(3, 7, i, j) - (h / dx_) * Bz_1 - (h / dy_) * Bz_2;
  }
};

// ---- magnetic operator row (implicit_operators.cpp:141-353, GLF branch).  Fields: 0 x (3 planes),
// 1 Qe (B), 2 u (rho).  MODE 0 seeds x (transpose); MODE 1 seeds the coefficients.
template <int MODE>
struct FnMagOp {
  static constexpr int K = MODE == 0 ? 3 : 4, NOUT = 3, NS = 9;
  double dx, dy;
  __device__ static void off(int s, int& oi, int& oj) { oi = s % 3 - 1; oj = s / 3 - 1; }
  __device__ static int slot(int fid, int k) {
    if (MODE == 0) return fid == 0 ? k : -1;
    if (fid == 1) return k >= 5 ? k - 5 : -1;
    if (fid == 2) return k == 0 ? 3 : -1;
    return -1;
  }
  __device__ static void which(int q, int& fid, int& k) {
    if (MODE == 0) { fid = 0; k = q; }
    else if (q < 3) { fid = 1; k = q + 5; } else { fid = 2; k = 0; }
  }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* out) const {
    const T h = sc.s[SC_H];
    const double dx_ = dx, dy_ = dy;
    auto X = [&](int a, int b, int k) { return rd(0, k - 5, a, b); };
    auto Q = [&](int a, int b, int k) { return rd(1, k, a, b); };
    auto Rr = [&](int a, int b) { return rd(2, 0, a, b); };
    const T r00 = Rr(i, j);
    const T rho_p0 = 0.5 * (r00 + Rr(i + 1, j)), rho_m0 = 0.5 * (r00 + Rr(i - 1, j));
    const T rho_0p = 0.5 * (r00 + Rr(i, j + 1)), rho_0m = 0.5 * (r00 + Rr(i, j - 1));
    T n_p0[3], n_m0[3], n_0p[3], n_0m[3], n_pp[3], n_pm[3], n_mp[3], n_mm[3];
    T b_pp[3], b_pm[3], b_mp[3], b_mm[3];
    T q[3][3][3], x[3][3][3];     // [di+1][dj+1][component] cell values of Qe B and x
    for (int c = 0; c < 3; c++)
      for (int di = 0; di < 3; di++)
        for (int dj = 0; dj < 3; dj++) { q[di][dj][c] = Q(i + di - 1, j + dj - 1, 5 + c); x[di][dj][c] = X(i + di - 1, j + dj - 1, 5 + c); }
    for (int c = 0; c < 3; c++) {
      n_p0[c] = 0.5 * (q[1][1][c] + q[2][1][c]); n_m0[c] = 0.5 * (q[1][1][c] + q[0][1][c]);
      n_0p[c] = 0.5 * (q[1][1][c] + q[1][2][c]); n_0m[c] = 0.5 * (q[1][1][c] + q[1][0][c]);
      n_pp[c] = 0.25 * (q[1][1][c] + q[2][1][c] + q[1][2][c] + q[2][2][c]);
      n_pm[c] = 0.25 * (q[1][1][c] + q[2][1][c] + q[1][0][c] + q[2][0][c]);
      n_mp[c] = 0.25 * (q[1][1][c] + q[0][1][c] + q[1][2][c] + q[0][2][c]);
      n_mm[c] = 0.25 * (q[1][1][c] + q[0][1][c] + q[1][0][c] + q[0][0][c]);
      b_pp[c] = 0.25 * (x[1][1][c] + x[2][1][c] + x[1][2][c] + x[2][2][c]);
      b_pm[c] = 0.25 * (x[1][1][c] + x[2][1][c] + x[1][0][c] + x[2][0][c]);
      b_mp[c] = 0.25 * (x[1][1][c] + x[0][1][c] + x[1][2][c] + x[0][2][c]);
      b_mm[c] = 0.25 * (x[1][1][c] + x[0][1][c] + x[1][0][c] + x[0][0][c]);
    }
    auto PM = [&](int di, int dj) {
      return 0.5 * (q[di][dj][0] * x[di][dj][0] + q[di][dj][1] * x[di][dj][1] + q[di][dj][2] * x[di][dj][2]);
    };
    auto PC = [&](const T* nn, const T* bb, int m) { return 0.5 * (nn[0] * bb[0] + nn[1] * bb[1] + nn[2] * bb[2]) - nn[m] * bb[m]; };
    // Q(.,.,6) * X(.,.,6) etc.: component indices 0..2 stand for 5..7
    auto QX = [&](int di, int dj, int cq, int cx) { return q[di][dj][cq] * x[di][dj][cx]; };
    const T Bx_1 = n_0p[0] / rho_0p * (1 / dx_ * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy_ * ((PM(1, 2) - QX(1, 2, 1, 1)) - (PM(1, 1) - QX(1, 1, 1, 1))));
    const T Bx_2 = n_0p[1] / rho_0p * (1 / dx_ * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1 / dy_ * (QX(1, 2, 1, 0) - QX(1, 1, 1, 0)));
    const T Bx_3 = n_0m[0] / rho_0m * (1 / dx_ * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy_ * ((PM(1, 1) - QX(1, 1, 1, 1)) - (PM(1, 0) - QX(1, 0, 1, 1))));
    const T Bx_4 = n_0m[1] / rho_0m * (1 / dx_ * (PC(n_pm, b_pm, 0) - PC(n_mm, b_mm, 0)) - 1 / dy_ * (QX(1, 1, 1, 0) - QX(1, 0, 1, 0)));
    const T ax5 = x[1][1][0] + (h * h / dy_) * ((Bx_1 + Bx_2) - (Bx_3 + Bx_4));
    const T By_1 = n_p0[1] / rho_p0 * (-1 / dx_ * ((PM(2, 1) - QX(2, 1, 0, 0)) - (PM(1, 1) - QX(1, 1, 0, 0))) + 1 / dy_ * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
    const T By_2 = n_p0[0] / rho_p0 * (-1 / dx_ * (QX(2, 1, 0, 1) - QX(1, 1, 0, 1)) + 1 / dy_ * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
    const T By_3 = n_m0[1] / rho_m0 * (-1 / dx_ * ((PM(1, 1) - QX(1, 1, 0, 0)) - (PM(0, 1) - QX(0, 1, 0, 0))) + 1 / dy_ * (n_mp[0] * b_mp[1] - n_mm[0] * b_mm[1]));
    const T By_4 = n_m0[0] / rho_m0 * (-1 / dx_ * (QX(1, 1, 0, 1) - QX(0, 1, 0, 1)) + 1 / dy_ * (PC(n_mp, b_mp, 1) - PC(n_mm, b_mm, 1)));
    const T ax6 = x[1][1][1] + (h * h / dx_) * ((By_1 + By_2) - (By_3 + By_4));
    const T Bz_1 = n_p0[2] / rho_p0 * (-1 / dx_ * ((PM(2, 1) - QX(2, 1, 0, 0)) - (PM(1, 1) - QX(1, 1, 0, 0))) + 1 / dy_ * (n_pp[1] * b_pp[0] - n_pm[1] * b_pm[0]));
    const T Bz_2 = n_p0[0] / rho_p0 * (-1 / dx_ * (QX(2, 1, 0, 2) - QX(1, 1, 0, 2)) - 1 / dy_ * (n_pp[1] * b_pp[2] - n_pm[1] * b_pm[2]));
    const T Bz_3 = n_m0[2] / rho_m0 * (-1 / dx_ * ((PM(1, 1) - QX(1, 1, 0, 0)) - (PM(0, 1) - QX(0, 1, 0, 0))) + 1 / dy_ * (n_mp[1] * b_mp[0] - n_mm[1] * b_mm[0]));
    const T Bz_4 = n_m0[0] / rho_m0 * (-1 / dx_ * (QX(1, 1, 0, 2) - QX(0, 1, 0, 2)) - 1 / dy_ * (n_mp[1] * b_mp[2] - n_mm[1] * b_mm[2]));
    const T Bz_5 = n_0p[2] / rho_0p * (1 / dx_ * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy_ * ((PM(1, 2) - QX(1, 2, 1, 1)) - (PM(1, 1) - QX(1, 1, 1, 1))));
    const T Bz_6 = n_0p[1] / rho_0p * (-1 / dx_ * (n_pp[0] * b_pp[2] - n_mp[0] * b_mp[2]) - 1 / dy_ * (QX(1, 2, 1, 2) - QX(1, 1, 1, 2)));
    const T Bz_7 = n_0m[2] / rho_0m * (1 / dx_ * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy_ * ((PM(1, 1) - QX(1, 1, 1, 1)) - (PM(1, 0) - QX(1, 0, 1, 1))));
    const T Bz_8 = n_0m[1] / rho_0m * (-1 / dx_ * (n_pm[0] * b_pm[2] - n_mm[0] * b_mm[2]) - 1 / dy_ * (QX(1, 1, 1, 2) - QX(1, 0, 1, 2)));
    const T ax7 = x[1][1][2] + h * h / dx_ * ((Bz_1 + Bz_2) - (Bz_3 + Bz_4)) + h * h / dy_ * ((Bz_5 + Bz_6) - (Bz_7 + Bz_8));
    const T nu_x = sc.s[SC_BMX] * dx_, nu_y = sc.s[SC_BMY] * dy_;
    const T ax[3] = {ax5, ax6, ax7};
    for (int c = 0; c < 3; c++) {
      const T lap = nu_x / (dx_ * dx_) * (x[0][1][c] - 2.0 * x[1][1][c] + x[2][1][c])
                  + nu_y / (dy_ * dy_) * (x[1][0][c] - 2.0 * x[1][1][c] + x[1][2][c]);
      out[c] = ax[c] - h * lap;
    }
  }
};

// ---- EMF fluxes (ct_flux.cpp:14-127): element (i, j) gives Fx at face (i+1/2, j) and Fy at
// (i, j+1/2).  Fields: 0 u (rho, rho v, B), 1 Qe (B), 2 p(Qe).
struct FnEMF {
  static constexpr int K = 10, NOUT = 2, NS = 9;
  double dx, dy;
  __device__ static void off(int s, int& oi, int& oj) { oi = s % 3 - 1; oj = s / 3 - 1; }
  __device__ static int slot(int fid, int k) {
    if (fid == 0) return k <= 2 ? k : (k >= 5 ? k - 2 : -1);
    if (fid == 1) return k >= 5 ? 6 + k - 5 : -1;
    return k == 0 ? 9 : -1;
  }
  __device__ static void which(int q, int& fid, int& k) {
    if (q < 3) { fid = 0; k = q; } else if (q < 6) { fid = 0; k = q + 2; } else if (q < 9) { fid = 1; k = q - 1; } else { fid = 2; k = 0; }
  }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* out) const {
    const T h = sc.s[SC_H];
    const double dx_ = dx, dy_ = dy;
    auto U = [&](int a, int b, int k) { return rd(0, k, a, b); };
    auto Q = [&](int a, int b, int k) { return rd(1, k, a, b); };
    auto PE = [&](int a, int b) { return rd(2, 0, a, b); };
    const T rho_p0 = 0.5 * (U(i, j, 0) + U(i + 1, j, 0)), rho_0p = 0.5 * (U(i, j, 0) + U(i, j + 1, 0));
    const T rvx_p0 = 0.5 * (U(i, j, 1) + U(i + 1, j, 1)), rvy_p0 = 0.5 * (U(i, j, 2) + U(i + 1, j, 2));
    const T rvx_0p = 0.5 * (U(i, j, 1) + U(i, j + 1, 1)), rvy_0p = 0.5 * (U(i, j, 2) + U(i, j + 1, 2));
    const T p_pp = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j + 1) + PE(i + 1, j + 1));
    const T p_pm = 0.25 * (PE(i, j) + PE(i + 1, j) + PE(i, j - 1) + PE(i + 1, j - 1));
    const T p_mp = 0.25 * (PE(i, j) + PE(i - 1, j) + PE(i, j + 1) + PE(i - 1, j + 1));
    const T Bxn_p0 = 0.5 * (Q(i, j, 5) + Q(i + 1, j, 5)), Byn_p0 = 0.5 * (Q(i, j, 6) + Q(i + 1, j, 6));
    const T Bxn_0p = 0.5 * (Q(i, j, 5) + Q(i, j + 1, 5)), Byn_0p = 0.5 * (Q(i, j, 6) + Q(i, j + 1, 6));
    T n_pp[3], n_pm[3], n_mp[3], b_pp[3], b_pm[3], b_mp[3];
    for (int c = 0; c < 3; c++) {
      const int k = 5 + c;
      n_pp[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j + 1, k) + Q(i + 1, j + 1, k));
      n_pm[c] = 0.25 * (Q(i, j, k) + Q(i + 1, j, k) + Q(i, j - 1, k) + Q(i + 1, j - 1, k));
      n_mp[c] = 0.25 * (Q(i, j, k) + Q(i - 1, j, k) + Q(i, j + 1, k) + Q(i - 1, j + 1, k));
      b_pp[c] = 0.25 * (U(i, j, k) + U(i + 1, j, k) + U(i, j + 1, k) + U(i + 1, j + 1, k));
      b_pm[c] = 0.25 * (U(i, j, k) + U(i + 1, j, k) + U(i, j - 1, k) + U(i + 1, j - 1, k));
      b_mp[c] = 0.25 * (U(i, j, k) + U(i - 1, j, k) + U(i, j + 1, k) + U(i - 1, j + 1, k));
    }
    auto PM = [&](int a, int b) { return 0.5 * (Q(a, b, 5) * U(a, b, 5) + Q(a, b, 6) * U(a, b, 6) + Q(a, b, 7) * U(a, b, 7)); };
    auto PC = [&](const T* nn, const T* bb, int m) { return 0.5 * (nn[0] * bb[0] + nn[1] * bb[1] + nn[2] * bb[2]) - nn[m] * bb[m]; };
    const T Fn_x = Byn_p0 / rho_p0 * (rvx_p0 - (h / dx_) * (PE(i + 1, j) - PE(i, j))) - Bxn_p0 / rho_p0 * (rvy_p0 - (h / dy_) * (p_pp - p_pm));
    const T By_1 = Byn_p0 / rho_p0 * (-1.0 / dx_ * ((PM(i + 1, j) - Q(i + 1, j, 5) * U(i + 1, j, 5)) - (PM(i, j) - Q(i, j, 5) * U(i, j, 5))) + 1.0 / dy_ * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
    const T By_2 = Bxn_p0 / rho_p0 * (-1.0 / dx_ * (Q(i + 1, j, 5) * U(i + 1, j, 6) - Q(i, j, 5) * U(i, j, 6)) + 1.0 / dy_ * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
    const T nu_x = sc.s[SC_BMX] * dx_;
    const T Fdiff_x = -nu_x * (U(i + 1, j, 6) - U(i, j, 6)) / dx_;
    out[0] = Fn_x + h * (By_1 + By_2) + Fdiff_x;
    const T Fn_y = Bxn_0p / rho_0p * (rvy_0p - (h / dy_) * (PE(i, j + 1) - PE(i, j))) - Byn_0p / rho_0p * (rvx_0p - (h / dx_) * (p_pp - p_mp));
    const T Bx_1 = Bxn_0p / rho_0p * (1.0 / dx_ * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1.0 / dy_ * ((PM(i, j + 1) - Q(i, j + 1, 6) * U(i, j + 1, 6)) - (PM(i, j) - Q(i, j, 6) * U(i, j, 6))));
    const T Bx_2 = Byn_0p / rho_0p * (1.0 / dx_ * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1.0 / dy_ * (Q(i, j + 1, 6) * U(i, j + 1, 5) - Q(i, j, 6) * U(i, j, 5)));
    const T nu_y = sc.s[SC_BMY] * dy_;
    const T Fdiff_y = -nu_y * (U(i, j + 1, 5) - U(i, j, 5)) / dy_;
    out[1] = Fn_y + h * (Bx_1 + Bx_2) + Fdiff_y;
  }
};

// ---- p(Qe) and hhat = calenthalpy(Qe) rho^e / rho (timestep.cpp:224-230, implicit_operators.cpp:506-531).
// Fields: 0 Qe, 1 u (rho).
struct FnPeHhat {
  static constexpr int K = 9, NOUT = 2, NS = 1;
  Phys P;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) { return fid == 0 ? k : (k == 0 ? 8 : -1); }
  __device__ static void which(int q, int& fid, int& k) { if (q < 8) { fid = 0; k = q; } else { fid = 1; k = 0; } }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    T q[8];
    for (int k = 0; k < 8; k++) q[k] = rd(0, k, i, j);
    const T rho = q[0], vx = q[1] / q[0], vy = q[2] / q[0], vz = q[3] / q[0];
    const T B2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
    const T v2 = vx * vx + vy * vy + vz * vz;
    const T e = (q[4] - 0.5 * rho * v2 - 0.5 * B2) / rho;
    const T p = eos_p_t(P, rho, e);
    out[0] = p;
    out[1] = (e + p / rho) * q[0] / rd(1, 0, i, j);
  }
};

// ---- shock indicator before the 3x3 max (flux.cpp:5-51).  Field: 0 Qe.
struct FnChi {
  static constexpr int K = 8, NOUT = 1, NS = 9;
  Phys P;
  double kappa, hmin, dx, dy;
  __device__ static void off(int s, int& oi, int& oj) { oi = s % 3 - 1; oj = s / 3 - 1; }
  __device__ static int slot(int fid, int k) { return fid == 0 ? k : -1; }
  __device__ static void which(int q, int& fid, int& k) { fid = 0; k = q; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    T cf_min(1.0e300);
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++) {
        T q[8];
        for (int k = 0; k < 8; k++) q[k] = rd(0, k, i + di, j + dj);
        const T rho = q[0], vx = q[1] / q[0], vy = q[2] / q[0], vz = q[3] / q[0];
        const T B2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
        const T e = (q[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * B2) / rho;
        const T p = eos_p_t(P, rho, e);
        const T a = eos_a_t(P, rho, p);
        const T cf = sqrt(a * a + B2 / rho);
        if (val(cf) < val(cf_min)) cf_min = cf;
      }
    const T div_u = (rd(0, 1, i + 1, j) / rd(0, 0, i + 1, j) - rd(0, 1, i - 1, j) / rd(0, 0, i - 1, j)) / (2.0 * dx)
                  + (rd(0, 2, i, j + 1) / rd(0, 0, i, j + 1) - rd(0, 2, i, j - 1) / rd(0, 0, i, j - 1)) / (2.0 * dy);
    const T num = hmin * div_u + kappa * cf_min;
    const T den = kappa * cf_min;
    const T x = -num / den;
    out[0] = val(x) < 0.0 ? T(0.0) : (val(x) > 1.0 ? T(1.0) : x);
  }
};

// ---- pressure right-hand side (imex_subsystems.cpp:123-175 pass 0, 223-284 pass 1).
// Fields: 0 current u, 1 momenta before the update (planes 1..3 of a grid), 2 hhat.
struct FnPRHS {
  static constexpr int K = 12, NOUT = 1, NS = 5;
  int pass;
  double dx, dy;
  __device__ static void off(int s, int& oi, int& oj) {
    const int ti[5] = {0, 1, -1, 0, 0}, tj[5] = {0, 0, 0, 1, -1};
    oi = ti[s]; oj = tj[s];
  }
  __device__ static int slot(int fid, int k) {
    if (fid == 0) return k;
    if (fid == 1) return (k >= 1 && k <= 3) ? 8 + k - 1 : -1;
    return k == 0 ? 11 : -1;
  }
  __device__ static void which(int q, int& fid, int& k) {
    if (q < 8) { fid = 0; k = q; } else if (q < 11) { fid = 1; k = q - 7; } else { fid = 2; k = 0; }
  }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* out) const {
    const T h = sc.s[SC_H];
    const double dx_ = dx, dy_ = dy;
    auto H = [&](int a, int b) { return rd(2, 0, a, b); };
    auto MX = [&](int a, int b) { return rd(1, 1, a, b); };
    auto MY = [&](int a, int b) { return rd(1, 2, a, b); };
    const T u5 = rd(0, 5, i, j), u6 = rd(0, 6, i, j), u7 = rd(0, 7, i, j);
    const T B2 = u5 * u5 + u6 * u6 + u7 * u7;
    const T h00 = H(i, j), hp0 = H(i + 1, j), hm0 = H(i - 1, j), h0p = H(i, j + 1), h0m = H(i, j - 1);
    const T p_star = rd(0, 4, i, j) - 0.5 * B2;
    const T rho = rd(0, 0, i, j);
    T ke;
    if (pass == 0) {
      const T a1 = MX(i, j), a2 = MY(i, j), a3 = rd(1, 3, i, j);
      ke = 0.5 * (a1 * a1 + a2 * a2 + a3 * a3) / rho;
    } else {
      const T a1 = rd(0, 1, i, j), a2 = rd(0, 2, i, j), a3 = rd(0, 3, i, j);
      ke = 0.5 * (a1 * a1 + a2 * a2 + a3 * a3) / rho;
    }
    const T mx = MX(i, j), my = MY(i, j);
    out[0] = p_star - ke - h * (1 / dx_ * (0.5 * (hp0 + h00) * 0.5 * (MX(i + 1, j) + mx) - 0.5 * (hm0 + h00) * 0.5 * (MX(i - 1, j) + mx))
                              + (1 / dy_ * (0.5 * (h0p + h00) * 0.5 * (MY(i, j + 1) + my) - 0.5 * (h0m + h00) * 0.5 * (MY(i, j - 1) + my))));
  }
};

// ---- pressure operator row (implicit_operators.cpp:505-616), carrier mode (V' = 1/(gamma-1)); the
// viscosity scale is 1/(gamma-1) for the ideal gas, the plane Vbase_slope(p(Qe), rho) for Redlich-Kwong.
// Fields: 0 x, 1 hhat, 2 eta, 3 visc (rk).  MODE 0 seeds x; MODE 1 seeds hhat, eta and visc.
template <int MODE>
struct FnPOp {
  static constexpr int K = MODE == 0 ? 1 : 3, NOUT = 1, NS = 5;
  double a, dx, dy;
  int rk;
  __device__ static void off(int s, int& oi, int& oj) {
    const int ti[5] = {0, 1, -1, 0, 0}, tj[5] = {0, 0, 0, 1, -1};
    oi = ti[s]; oj = tj[s];
  }
  __device__ static int slot(int fid, int k) {
    if (k != 0) return -1;
    if (MODE == 0) return fid == 0 ? 0 : -1;
    return fid >= 1 && fid <= 3 ? fid - 1 : -1;
  }
  __device__ static void which(int q, int& fid, int& k) { k = 0; fid = MODE == 0 ? 0 : 1 + q; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>& sc, int i, int j, T* out) const {
    const T h = sc.s[SC_H];
    const double dx_ = dx, dy_ = dy;
    auto H = [&](int a_, int b) { return rd(1, 0, a_, b); };
    auto E = [&](int a_, int b) { return rd(2, 0, a_, b); };
    auto X = [&](int a_, int b) { return rd(0, 0, a_, b); };
    const T h00 = H(i, j), hp0 = H(i + 1, j), hm0 = H(i - 1, j), h0p = H(i, j + 1), h0m = H(i, j - 1);
    const T pii = X(i, j), xp0 = X(i + 1, j), xm0 = X(i - 1, j), x0p = X(i, j + 1), x0m = X(i, j - 1);
    const T ax = a * pii - h * h * (1 / (dx_ * dx_) * (0.5 * (hp0 + h00) * (xp0 - pii) - 0.5 * (h00 + hm0) * (pii - xm0))
                                  + 1 / (dy_ * dy_) * (0.5 * (h0p + h00) * (x0p - pii) - 0.5 * (h0m + h00) * (pii - x0m)));
    const T nu_x = sc.s[SC_VXM] * dx_, nu_y = sc.s[SC_VYM] * dy_;
    const T e00 = E(i, j);
    const T lap = nu_x / (dx_ * dx_) * (0.5 * (E(i + 1, j) + e00) * (xp0 - pii) - 0.5 * (e00 + E(i - 1, j)) * (pii - xm0))
                + nu_y / (dy_ * dy_) * (0.5 * (E(i, j + 1) + e00) * (x0p - pii) - 0.5 * (e00 + E(i, j - 1)) * (pii - x0m));
    const T visc = rk ? rd(3, 0, i, j) : T(a);
    out[0] = ax - visc * h * lap;
  }
};

// ---- viscosity scale of the Redlich-Kwong pressure operator, Vbase_slope(p(Qe), rho) (config.hpp:211-217).
// Fields: 0 p(Qe), 1 u (rho).
struct FnVisc {
  static constexpr int K = 2, NOUT = 1, NS = 1;
  Phys P;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) { return k == 0 && fid <= 1 ? fid : -1; }
  __device__ static void which(int q, int& fid, int& k) { fid = q; k = 0; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    out[0] = vbase_slope_t(P, rd(0, 0, i, j), rd(1, 0, i, j));
  }
};

// ---- energy from the pressure (imex_subsystems.cpp:286-302): Vbase(p, rho) + kinetic + magnetic.
// Fields: 0 p, 1 u.
struct FnEnergy {
  static constexpr int K = 8, NOUT = 1, NS = 1;
  Phys P;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) {
    if (fid == 0) return k == 0 ? 0 : -1;
    return k < 4 ? 1 + k : (k >= 5 ? k : -1);
  }
  __device__ static void which(int q, int& fid, int& k) {
    if (q == 0) { fid = 0; k = 0; } else if (q < 5) { fid = 1; k = q - 1; } else { fid = 1; k = q; }
  }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    const T rho = rd(1, 0, i, j), m1 = rd(1, 1, i, j), m2 = rd(1, 2, i, j), m3 = rd(1, 3, i, j);
    const T b5 = rd(1, 5, i, j), b6 = rd(1, 6, i, j), b7 = rd(1, 7, i, j);
    out[0] = vbase_t(P, rd(0, 0, i, j), rho) + 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);
  }
};

// ---- cell EMF for CT, Ez~ = vy Bx - vx By (ct_flux.cpp:166-172).  Field: 0 u.
struct FnEt {
  static constexpr int K = 5, NOUT = 1, NS = 1;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) { return fid != 0 ? -1 : (k <= 2 ? k : (k == 5 || k == 6 ? k - 2 : -1)); }
  __device__ static void which(int q, int& fid, int& k) { fid = 0; k = q <= 2 ? q : q + 2; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    const T rho = rd(0, 0, i, j);
    out[0] = rd(0, 2, i, j) / rho * rd(0, 5, i, j) - rd(0, 1, i, j) / rho * rd(0, 6, i, j);
  }
};

// ---- per-cell quantities of the start-of-step maxima (timestep.cpp:5-67, 89-120, 254-273), from
// the conservative state: v + c_f(|B|), |v|, the two magnetic eigenvalue bounds, |vx|, |vy|.
struct FnMax {
  static constexpr int K = 8, NOUT = 6, NS = 1;
  Phys P;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) { return fid == 0 ? k : -1; }
  __device__ static void which(int q, int& fid, int& k) { fid = 0; k = q; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* out) const {
    T q[8];
    for (int k = 0; k < 8; k++) q[k] = rd(0, k, i, j);
    const T rho = q[0], vx = q[1] / q[0], vy = q[2] / q[0], vz = q[3] / q[0];
    const T B2 = q[5] * q[5] + q[6] * q[6] + q[7] * q[7];
    const T e = (q[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * B2) / rho;
    const T p = eos_p_t(P, rho, e);
    const T Bm = sqrt(B2);
    const T v = sqrt(vx * vx + vy * vy + vz * vz);
    const T a = eos_a_t(P, rho, p);
    const T a2 = a * a;
    const T t = a2 + B2 / rho;
    T disc = t * t - 4.0 * a2 * Bm * Bm / rho;
    if (val(disc) < 0.0) disc = T(0.0);
    const T cf = sqrt(0.5 * (t + sqrt(disc)));
    out[0] = v + cf;
    out[1] = v;
    out[2] = 0.5 * (fabs(vx) + sqrt(vx * vx + 4.0 * B2 / rho));
    out[3] = 0.5 * (fabs(vy) + sqrt(vy * vy + 4.0 * B2 / rho));
    out[4] = fabs(vx);
    out[5] = fabs(vy);
  }
};

// ---- primitive -> conservative (state.cpp:242-255).  Field: 0 primitives.
struct FnPri2Con {
  static constexpr int K = 8, NOUT = 8, NS = 1;
  Phys P;
  __device__ static void off(int, int& oi, int& oj) { oi = 0; oj = 0; }
  __device__ static int slot(int fid, int k) { return fid == 0 ? k : -1; }
  __device__ static void which(int q, int& fid, int& k) { fid = 0; k = q; }
  template <class T, class R>
  __device__ void eval(const R& rd, const Sc<T>&, int i, int j, T* c) const {
    T p[8];
    for (int k = 0; k < 8; k++) p[k] = rd(0, k, i, j);
    const T e = eos_e_t(P, p[0], p[4]);
    const T B2 = p[5] * p[5] + p[6] * p[6] + p[7] * p[7];
    const T v2 = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    c[0] = p[0]; c[1] = p[0] * p[1]; c[2] = p[0] * p[2]; c[3] = p[0] * p[3];
    c[4] = p[0] * e + 0.5 * p[0] * v2 + 0.5 * B2;
    c[5] = p[5]; c[6] = p[6]; c[7] = p[7];
  }
};

// ================================================================ launch helpers
static Sc<double> to_sc(const AdjScal& s) { Sc<double> r; for (int q = 0; q < SC_N; q++) r.s[q] = s.s[q]; return r; }
static AFields fields(std::initializer_list<const double*> f, std::initializer_list<double*> b) {
  AFields F;
  for (int q = 0; q < ADJ_NF; q++) { F.f[q] = nullptr; F.bar[q] = nullptr; }
  int q = 0;
  for (const double* p : f) F.f[q++] = p;
  q = 0;
  for (double* p : b) F.bar[q++] = p;
  return F;
}
template <class Fn>
static void vjp(hipStream_t st, const Fn& fn, const AD& a, const AFields& F, const double* obar, const AdjScal& sc, double* sbar) {
  k_vjp<Fn><<<nblk(a), AB, 0, st>>>(fn, a, F, obar, to_sc(sc));
  if (sbar) k_vjp_scalars<Fn><<<nblk(a), AB, 0, st>>>(fn, a, F, obar, to_sc(sc), sbar);
}

// ================================================================ linear pieces
// flux divergence u = Qs - h (dF/dx + dG/dy): face cotangents, Qs cotangent, h cotangent
__global__ void __launch_bounds__(AB) k_adj_fluxdiv(AD a, int k0, int k1, double h, const double* __restrict__ ubar,
                                                    const double* __restrict__ uout, const double* __restrict__ Qs,
                                                    double* __restrict__ Qsbar, double* __restrict__ fbx,
                                                    double* __restrict__ fby, double* sbar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double hacc = 0.0;
  if (n < a.nx * a.ny) {
    const int i = n % a.nx, j = n / a.nx;
    const long long c = tix(a, i, j), cl = tix(a, i - 1, j), cd = tix(a, i, j - 1);
    for (int k = 0; k < 5; k++) {
      double fx = 0.0, fy = 0.0;
      if (k >= k0 && k < k1) {
        const double ub = ubar[k * a.PS + c];
        fx = (h / a.dx) * (ub - ubar[k * a.PS + cl]);
        fy = (h / a.dy) * (ub - ubar[k * a.PS + cd]);
        Qsbar[k * a.PS + c] += ub;
        hacc += ub * (uout[k * a.PS + c] - Qs[k * a.PS + c]) / h;
      }
      fbx[k * a.PS + c] = fx;
      fby[k * a.PS + c] = fy;
    }
  }
  hacc = wave_sum(hacc);
  if ((threadIdx.x & 63) == 0 && hacc != 0.0) atomicAdd(sbar + SC_H, hacc);
}

__global__ void __launch_bounds__(AB) k_adj_momentum(AD a, double h, const double* __restrict__ umid,
                                                     const double* __restrict__ uin, const double* __restrict__ ubar,
                                                     double* __restrict__ pbar, double* sbar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double hacc = 0.0;
  if (n < a.nx * a.ny) {
    const int i = n % a.nx, j = n / a.nx;
    const long long c = tix(a, i, j);
    const double* u1 = ubar + a.PS;
    const double* u2 = ubar + 2 * a.PS;
    pbar[c] += (h / (2.0 * a.dx)) * (u1[tix(a, i + 1, j)] - u1[tix(a, i - 1, j)])
             + (h / (2.0 * a.dy)) * (u2[tix(a, i, j + 1)] - u2[tix(a, i, j - 1)]);
    hacc = u1[c] * (umid[a.PS + c] - uin[a.PS + c]) / h + u2[c] * (umid[2 * a.PS + c] - uin[2 * a.PS + c]) / h;
  }
  hacc = wave_sum(hacc);
  if ((threadIdx.x & 63) == 0 && hacc != 0.0) atomicAdd(sbar + SC_H, hacc);
}

// CT, part 1: total face cotangents (output faces + cell averages), Qs-face cotangents, h
__global__ void __launch_bounds__(AB) k_adj_ct_faces(AD a, double h, const double* __restrict__ ubar,
                                                     const double* __restrict__ bxbar, const double* __restrict__ bybar,
                                                     const double* __restrict__ bxo, const double* __restrict__ byo,
                                                     const double* __restrict__ bxqs, const double* __restrict__ byqs,
                                                     double* __restrict__ bxt, double* __restrict__ byt,
                                                     double* __restrict__ bxqsbar, double* __restrict__ byqsbar, double* sbar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double hacc = 0.0;
  if (n < a.nx * a.ny) {
    const int i = n % a.nx, j = n / a.nx;
    const long long c = tix(a, i, j);
    const double* u5 = ubar + 5 * a.PS;
    const double* u6 = ubar + 6 * a.PS;
    const double tx = (bxbar ? bxbar[c] : 0.0) + 0.5 * (u5[c] + u5[tix(a, i + 1, j)]);
    const double ty = (bybar ? bybar[c] : 0.0) + 0.5 * (u6[c] + u6[tix(a, i, j + 1)]);
    bxt[c] = tx; byt[c] = ty;
    bxqsbar[c] += tx; byqsbar[c] += ty;
    hacc = tx * (bxo[c] - bxqs[c]) / h + ty * (byo[c] - byqs[c]) / h;
  }
  hacc = wave_sum(hacc);
  if ((threadIdx.x & 63) == 0 && hacc != 0.0) atomicAdd(sbar + SC_H, hacc);
}
// CT, part 2: corner-EMF cotangent; the cell-B cotangents are consumed
__global__ void __launch_bounds__(AB) k_adj_ct_ezpp(AD a, double h, const double* __restrict__ bxt, const double* __restrict__ byt,
                                                    double* __restrict__ ezbar, double* __restrict__ ubar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  const long long c = tix(a, i, j);
  ezbar[c] = -(h / a.dy) * (bxt[c] - bxt[tix(a, i, j + 1)]) + (h / a.dx) * (byt[c] - byt[tix(a, i + 1, j)]);
  ubar[5 * a.PS + c] = 0.0;
  ubar[6 * a.PS + c] = 0.0;
}
// CT, part 3: Ez_pp = (Ezp0(i,j) + Ezp0(i,j+1))/2 + (Ez0p(i,j) + Ez0p(i+1,j))/2 - (sum of the 4 Et)/4,
// the closed form of CT_CornerAssemble (constrained_transport.cpp:147-164), with Ezp0 = -Fx, Ez0p = Fy
__global__ void __launch_bounds__(AB) k_adj_ct_emf(AD a, const double* __restrict__ ezbar, double* __restrict__ fxbar,
                                                   double* __restrict__ fybar, double* __restrict__ etbar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  const long long c = tix(a, i, j);
  const double e00 = ezbar[c], em0 = ezbar[tix(a, i - 1, j)], e0m = ezbar[tix(a, i, j - 1)], emm = ezbar[tix(a, i - 1, j - 1)];
  fxbar[c] += -0.5 * (e00 + e0m);
  fybar[c] += 0.5 * (e00 + em0);
  etbar[c] = -0.25 * (e00 + em0 + e0m + emm);
}

// eta = max over 3x3 of chi (flux.cpp:52-76): the cotangent goes to the first strictly largest
// positive chi in the reference's scan order
__global__ void __launch_bounds__(AB) k_adj_eta(AD a, const double* __restrict__ chi, const double* __restrict__ etabar,
                                                double* __restrict__ chibar) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  const double eb = etabar[tix(a, i, j)];
  if (eb == 0.0) return;
  double m = 0.0;
  long long arg = -1;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++) {
      const long long c = tix(a, i + di, j + dj);
      if (chi[c] > m) { m = chi[c]; arg = c; }
    }
  if (arg >= 0) atomicAdd(chibar + arg, eb);
}

__global__ void __launch_bounds__(AB) k_adj_pdiag(AD a, double h, double av, const double* __restrict__ hh, double* __restrict__ diag) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const int i = n % a.nx, j = n / a.nx;
  const double h00 = hh[tix(a, i, j)], hp0 = hh[tix(a, i + 1, j)], hm0 = hh[tix(a, i - 1, j)];
  const double h0p = hh[tix(a, i, j + 1)], h0m = hh[tix(a, i, j - 1)];
  diag[tix(a, i, j)] = av + h * h * ((0.5 * (hp0 + h00) + 0.5 * (hm0 + h00)) / (a.dx * a.dx) + (0.5 * (h0p + h00) + 0.5 * (h0m + h00)) / (a.dy * a.dy));
}

// argmax of 6 planes over interior cells: block partials, then one block
__global__ void __launch_bounds__(256) k_argmax_part(AD a, const double* __restrict__ vals, double* __restrict__ pv, int* __restrict__ pi) {
  __shared__ double sv[256];
  __shared__ int si[256];
  for (int q = 0; q < 6; q++) {
    double best = -1e300;
    int bi = 0x7fffffff;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.nx * a.ny; n += gridDim.x * blockDim.x) {
      const double v = vals[q * a.PS + tix(a, n % a.nx, n / a.nx)];
      if (v > best || (v == best && n < bi)) { best = v; bi = n; }
    }
    sv[threadIdx.x] = best; si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        const double v = sv[threadIdx.x + s];
        const int ii = si[threadIdx.x + s];
        if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && ii < si[threadIdx.x])) { sv[threadIdx.x] = v; si[threadIdx.x] = ii; }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) { pv[q * gridDim.x + blockIdx.x] = sv[0]; pi[q * gridDim.x + blockIdx.x] = si[0]; }
    __syncthreads();
  }
}
__global__ void k_argmax_fin(int nb, const double* __restrict__ pv, const int* __restrict__ pi, int* __restrict__ out) {
  const int q = threadIdx.x;
  if (q >= 6) return;
  double best = -1e300;
  int bi = 0x7fffffff;
  for (int b = 0; b < nb; b++) {
    const double v = pv[q * nb + b];
    const int ii = pi[q * nb + b];
    if (v > best || (v == best && ii < bi)) { best = v; bi = ii; }
  }
  out[q] = bi;
}
__global__ void k_max_grad(FnMax fn, AD a, AFields F, const int* __restrict__ arg, const double* __restrict__ qbar,
                           double* __restrict__ Ubar) {
  const int q = threadIdx.x;
  if (q >= 6) return;
  const double qb = qbar[q];
  if (qb == 0.0) return;
  const int n = arg[q];
  const int i = n % a.nx, j = n / a.nx;
  typedef DN<8> T;
  Sc<T> st;
  for (int s = 0; s < SC_N; s++) st.s[s] = T(0.0);
  Rd<T, FnMax> rd{&a, &F, i, j};
  T out[6];
  fn.eval(rd, st, i, j, out);
  const long long c = tix(a, i, j);
  for (int k = 0; k < 8; k++) atomicAdd(Ubar + k * a.PS + c, qb * out[q].g[k]);
}

// vector helpers (interior cells of nc planes)
struct VPtrs { const double* p[MAXM + 2]; };
struct VCoef { double c[MAXM + 2]; };
__global__ void __launch_bounds__(256) k_adj_dots(AD a, int nc, int nv, VPtrs v, const double* __restrict__ w, double* out) {
  __shared__ double sh[MAXM + 2][4];
  double acc[MAXM + 2];
  for (int q = 0; q < nv; q++) acc[q] = 0.0;
  const int N = a.nx * a.ny;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N * nc; n += gridDim.x * blockDim.x) {
    const int comp = n / N, m = n % N;
    const long long c = comp * a.PS + tix(a, m % a.nx, m / a.nx);
    const double wv = w[c];
    for (int q = 0; q < nv; q++) acc[q] += v.p[q][c] * wv;
  }
  for (int q = 0; q < nv; q++) {
    const double s = wave_sum(acc[q]);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    const double s = sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
    atomicAdd(out + threadIdx.x, s);
  }
}
__global__ void __launch_bounds__(AB) k_adj_axpys(AD a, int nc, int nv, VCoef cf, VPtrs v, double* __restrict__ w, double sign) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.nx * a.ny;
  if (n >= N * nc) return;
  const int comp = n / N, m = n % N;
  const long long c = comp * a.PS + tix(a, m % a.nx, m / a.nx);
  double s = 0.0;
  for (int q = 0; q < nv; q++) s += cf.c[q] * v.p[q][c];
  w[c] += sign * s;
}
// one grid row of one plane per block row (x over the row's cells): no division per element
__global__ void __launch_bounds__(AB) k_adj_lin(AD a, int nc, int op, const double* __restrict__ in, const double* __restrict__ d,
                                                double s, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y;
  if (i >= a.nx || row >= nc * a.ny) return;
  const int comp = row / a.ny, jr = row - comp * a.ny;
  const long long c = comp * a.PS + (long long)(i + GC) + (long long)(jr + GC) * a.NX;
  switch (op) {
    case 0: out[c] = s * in[c]; break;            // scale-copy
    case 1: out[c] = in[c] / d[c]; break;         // divide by a diagonal (one plane)
    case 2: out[c] = 0.0; break;                  // zero
    default: out[c] += s * in[c]; break;          // add
  }
}
// ---- the nested Newton's converged solution (implicit_operators.cpp:625-738, config.hpp:206-239):
// V_mod(p, rho) + T p = b, T the carrier operator less its 1/(gamma-1) diagonal.  Its Jacobian in p
// is T + diag dV_mod/dp: the transposed solve adds dv = dV_mod/dp - 1/(gamma-1) to the carrier's
// diagonal (Mbig below the floor; for Redlich-Kwong the derivative of rho e(rho, p), not the forward's
// difference quotient, which only steers the Newton iteration).  mode 0: dv; mode 1: the cotangent of
// rho, rhobar -= lam dV_mod/drho.
__global__ void __launch_bounds__(AB) k_adj_newton(AD a, Phys P, int mode, const double* __restrict__ p,
                                                   const double* __restrict__ u, const double* __restrict__ mb,
                                                   const double* __restrict__ lam, double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const long long c = tix(a, n % a.nx, n / a.nx);
  typedef DN<2> T;
  T pv(p[c]), rv(u[c]);
  pv.g[0] = 1.0; rv.g[1] = 1.0;
  const T V = P.floor_on ? vmod_t(P, pv, rv, mb[0]) : vbase_t(P, pv, rv);
  if (mode == 0) out[c] = V.g[0] - 1.0 / (P.gamma - 1.0);
  else out[c] -= lam[c] * V.g[1];
}
// dv = vs - 1/(gamma-1) for a given slope plane vs (the forward's Jacobian-mode diagonal)
__global__ void __launch_bounds__(AB) k_adj_slope_dv(AD a, double av, const double* __restrict__ vs, double* __restrict__ dv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const long long c = tix(a, n % a.nx, n / a.nx);
  dv[c] = vs[c] - av;
}
// out += dv * in (the diagonal correction of the transposed Jacobian; in null: out += dv)
__global__ void __launch_bounds__(AB) k_adj_diag_add(AD a, const double* __restrict__ dv, const double* __restrict__ in,
                                                     double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.nx * a.ny) return;
  const long long c = tix(a, n % a.nx, n / a.nx);
  out[c] += dv[c] * (in ? in[c] : 1.0);
}
// Mbig = max(M_base, M_cor) with M_cor = max(0, Vpe - min b) / ((1 - theta) p_floor) and Vpe =
// max(p_floor/(gamma-1), max rho e(rho, p_floor)) (the second for Redlich-Kwong; :670-686), and V_mod =
// Vbase(p_floor) + Mbig (p - p_floor) below the floor: the cotangent of Mbig is -sum_c lam_c (p_c - p_floor)
// over the floored cells (lam = J^-T pbar).  Where M_cor decides Mbig it reaches b at the cell that attains
// min b (lam[argmin] += Mbig_bar dM_cor/d(min b)) and, when the Redlich-Kwong term decides Vpe, rho at the
// cell that attains it (lowest indices on ties).  One block.
#define FLB 64          // blocks of the floor reduction
__global__ void __launch_bounds__(256) k_adj_floor_part(AD a, Phys P, const double* __restrict__ p, const double* __restrict__ b,
                                                        const double* __restrict__ u, const double* __restrict__ lam,
                                                        double* __restrict__ work) {
  __shared__ double sm[4], sb[4], sv[4];
  __shared__ int si[4], sj[4];
  const int N = a.nx * a.ny;
  const double pf = P.p_floor;
  const bool rk = P.eos != IMEX_EOS_IDEAL;
  double acc = 0.0, bmin = 1e300, vmax = -1e300;
  int imin = N, imax = N;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    const long long c = tix(a, n % a.nx, n / a.nx);
    if (p[c] < pf) acc += lam[c] * (p[c] - pf);
    if (b[c] < bmin) { bmin = b[c]; imin = n; }
    if (rk) {
      const double v = vbase_t(P, pf, u[c]);
      if (v > vmax) { vmax = v; imax = n; }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_down(acc, o, 64);
    const double ob = __shfl_down(bmin, o, 64), ov = __shfl_down(vmax, o, 64);
    const int oi = __shfl_down(imin, o, 64), oj = __shfl_down(imax, o, 64);
    if (ob < bmin || (ob == bmin && oi < imin)) { bmin = ob; imin = oi; }
    if (ov > vmax || (ov == vmax && oj < imax)) { vmax = ov; imax = oj; }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sm[wid] = acc; sb[wid] = bmin; si[wid] = imin; sv[wid] = vmax; sj[wid] = imax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      acc += sm[w];
      if (sb[w] < bmin || (sb[w] == bmin && si[w] < imin)) { bmin = sb[w]; imin = si[w]; }
      if (sv[w] > vmax || (sv[w] == vmax && sj[w] < imax)) { vmax = sv[w]; imax = sj[w]; }
    }
    int* wi = (int*)(work + 3 * FLB);
    work[blockIdx.x] = acc; work[FLB + blockIdx.x] = bmin; work[2 * FLB + blockIdx.x] = vmax;
    wi[blockIdx.x] = imin; wi[FLB + blockIdx.x] = imax;
  }
}
__global__ void k_adj_floor_fin(AD a, Phys P, const double* __restrict__ u, const double* __restrict__ work,
                                double* __restrict__ lam, double* __restrict__ ubar) {
  if (threadIdx.x != 0) return;
  const int N = a.nx * a.ny;
  const int* wi = (const int*)(work + 3 * FLB);
  double acc = 0.0, bmin = 1e300, vmax = -1e300;
  int imin = N, imax = N;
  for (int q = 0; q < FLB; q++) {                 // block order: the same sums on every run
    acc += work[q];
    if (work[FLB + q] < bmin || (work[FLB + q] == bmin && wi[q] < imin)) { bmin = work[FLB + q]; imin = wi[q]; }
    if (work[2 * FLB + q] > vmax || (work[2 * FLB + q] == vmax && wi[FLB + q] < imax)) { vmax = work[2 * FLB + q]; imax = wi[FLB + q]; }
  }
  const double pf = P.p_floor, av = 1.0 / (P.gamma - 1.0);
  const double mbar = -acc;
  double vpe = pf * av;
  const bool vrk = P.eos != IMEX_EOS_IDEAL && vmax > vpe;
  if (vrk) vpe = vmax;
  const double mcor = fmax(0.0, vpe - bmin) / ((1.0 - P.theta) * pf);
  if (mcor > P.mbig_fac * av && vpe - bmin > 0.0 && imin < N) {
    const double s = mbar / ((1.0 - P.theta) * pf);      // Mbar dM_cor/dVpe; dM_cor/d(min b) = -that
    lam[tix(a, imin % a.nx, imin / a.nx)] -= s;
    if (vrk && imax < N) {
      const long long c = tix(a, imax % a.nx, imax / a.nx);
      DN<1> r(u[c]);
      r.g[0] = 1.0;
      ubar[c] += s * vbase_t(P, DN<1>(pf), r).g[0];
    }
  }
}

__global__ void __launch_bounds__(AB) k_adj_compact(AD a, int nc, int dir, const double* __restrict__ in, double* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.nx * a.ny;
  if (n >= N * nc) return;
  const int comp = n / N, m = n % N;
  const long long c = comp * a.PS + tix(a, m % a.nx, m / a.nx);
  if (dir == 0) out[c] = in[n];
  else out[n] = in[c];
}

// ---------------------------------------------------------------- transposed GMRES, device side
// The vector kernels run on a 2-D grid: x over AB-cell slabs of a row, y over groups of AKR_ROWS
// rows of all nc planes, so an element's padded index needs no division.  Sums are written per
// block (no atomics) and reduced in a fixed order by the one-block kernels that use them.
#define AKR_ROWS 8
// element r of the thread in a grid of `rows` rows per block
template <int ROWS = AKR_ROWS>
__device__ __forceinline__ long long akr_index(const AD& a, int nc, int r, bool& ok) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, row = blockIdx.y * ROWS + r;
  ok = i < a.nx && row < nc * a.ny;
  const int rr = ok ? row : 0, comp = rr / a.ny, j = rr - comp * a.ny;
  return comp * a.PS + (long long)((ok ? i : 0) + GC) + (long long)(j + GC) * a.NX;
}
__device__ __forceinline__ int akr_block() { return blockIdx.x + gridDim.x * blockIdx.y; }
static inline dim3 akr_grid(const AD& a, int nc) { return dim3((a.nx + AB - 1) / AB, (nc * a.ny + AKR_ROWS - 1) / AKR_ROWS); }
// one row per block: the elementwise kernels, whose many blocks keep more loads in flight
static inline dim3 akr_grid1(const AD& a, int nc) { return dim3((a.nx + AB - 1) / AB, nc * a.ny); }

// the block's sum of v (all threads call it); thread 0 gets the result
__device__ __forceinline__ double akr_block_sum(double v) {
  __shared__ double sh[AB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < AB / 64; q++) t += sh[q];
  __syncthreads();
  return t;
}
// sum of part[b * stride] over b < nb by one block (fixed order per lane, then over the waves)
__device__ __forceinline__ double akr_fin_sum(const double* part, int stride, int nb) {
  __shared__ double sh[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long long)b * stride];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[q];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(1024) k_akr_init(AdjKry* K, const double* __restrict__ pn, int nb, double tol) {
  const double b2 = akr_fin_sum(pn, 1, nb);
  if (threadIdx.x != 0) return;
  K->bnorm = sqrt(b2);
  K->tol = tol;
  K->res = K->bnorm;
  K->solved = K->bnorm == 0.0 ? 1 : 0;
  K->cdone = K->solved;
  K->iters = 0;
  K->jlast = -1;
}

__global__ void __launch_bounds__(1024) k_akr_cycle(AdjKry* K, const double* __restrict__ pn, int nb) {
  const int t = threadIdx.x;
  const double r2 = akr_fin_sum(pn, 1, nb);
  if (t == 0) K->jlast = -1;
  if (K->solved) { if (t == 0) K->cdone = 1; return; }
  for (int i = t; i <= ADJ_M; i += blockDim.x) K->g[i] = 0.0;
  __syncthreads();
  if (t == 0) {
    const double beta = sqrt(r2);
    K->beta = beta;
    K->res = beta;
    if (beta <= K->tol * K->bnorm) { K->solved = 1; K->cdone = 1; }
    else { K->cdone = 0; K->g[0] = beta; }
  }
}

// out = in / s (s = beta for the first basis vector, hn for the next ones); zero once the cycle has ended
__global__ void __launch_bounds__(AB) k_akr_scale(AD a, int nc, const AdjKry* __restrict__ K, int which,
                                                  const double* __restrict__ in, double* __restrict__ out) {
  const double sc = 1.0 / (which == 0 ? K->beta : K->hn);
  bool ok;
  const long long c = akr_index<1>(a, nc, 0, ok);
  if (ok) out[c] = K->cdone ? 0.0 : in[c] * sc;
}

// part[block][q] = the block's share of V_q . w, four basis vectors per pass so that the
// accumulators stay in registers (past nv the last vector is read again and its product dropped)
__global__ void __launch_bounds__(AB) k_akr_dots(AD a, int nc, int nv, AdjVecs V, const double* __restrict__ w,
                                                 const AdjKry* __restrict__ K, double* __restrict__ part) {
  if (K->cdone) return;
  constexpr int CH = 4;
  long long cs[AKR_ROWS];
  double wv[AKR_ROWS];
#pragma unroll
  for (int r = 0; r < AKR_ROWS; r++) {
    bool ok;
    cs[r] = akr_index(a, nc, r, ok);
    wv[r] = ok ? w[cs[r]] : 0.0;
  }
  double* pb = part + (long long)akr_block() * (ADJ_M + 1);
  for (int q0 = 0; q0 < nv; q0 += CH) {
    double acc[CH] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const double* vp = V.p[min(q0 + u, nv - 1)];
#pragma unroll
      for (int r = 0; r < AKR_ROWS; r++) acc[u] += vp[cs[r]] * wv[r];
    }
#pragma unroll
    for (int u = 0; u < CH; u++) {
      const double t = akr_block_sum(acc[u]);
      if (threadIdx.x == 0 && q0 + u < nv) pb[q0 + u] = t;
    }
  }
}

// d (d1 or d2)[q] = sum over the blocks of part[.][q], q < nv: one wave per q, lanes over blocks
__global__ void __launch_bounds__(1024) k_akr_dsum(AdjKry* K, const double* __restrict__ part, int nb, int nv, int d) {
  if (K->cdone) return;
  double* dst = d == 0 ? K->d1 : K->d2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int q = wv; q < nv; q += nw) {
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += part[(long long)b * (ADJ_M + 1) + q];
    s = wave_sum(s);
    if (lane == 0) dst[q] = s;
  }
}

// w -= sum_q d[q] V_q (q < nv); with norm, pn[block] = the block's share of ||w||^2
__global__ void __launch_bounds__(AB) k_akr_project(AD a, int nc, int nv, AdjVecs V, double* __restrict__ w,
                                                    const AdjKry* __restrict__ K, int d, double* __restrict__ pn) {
  if (K->cdone) return;
  __shared__ double cf[ADJ_M + 1];
  const double* src = d == 0 ? K->d1 : K->d2;
  for (int q = threadIdx.x; q < nv; q += blockDim.x) cf[q] = src[q];
  __syncthreads();
  bool ok;
  const long long c = akr_index<1>(a, nc, 0, ok);
  double sq = 0.0;
  if (ok) {
    double s = 0.0;
    for (int q = 0; q < nv; q++) s += cf[q] * V.p[q][c];
    const double wv = w[c] - s;
    w[c] = wv;
    sq = wv * wv;
  }
  if (pn) {
    const double t = akr_block_sum(sq);
    if (threadIdx.x == 0) pn[akr_block()] = t;
  }
}

// Column j of the Hessenberg matrix from the two Gram-Schmidt passes, the rotations, and whether
// the cycle ends: converged (|g_{j+1}| <= tol ||b||), breakdown (hn = 0) or restart (j + 1 = m).
__global__ void __launch_bounds__(1024) k_akr_column(AdjKry* K, const double* __restrict__ pn, int nb, int j, int m) {
  if (K->cdone) return;
  __shared__ double hcol[ADJ_M + 2];
  const int t = threadIdx.x;
  const double w2 = akr_fin_sum(pn, 1, nb);
  for (int i = t; i <= j; i += blockDim.x) hcol[i] = K->d1[i] + K->d2[i];
  __syncthreads();
  if (t == 0) {
    const double hn = sqrt(w2);
    for (int i = 0; i < j; i++) {
      const double a0 = K->cs[i] * hcol[i] + K->sn[i] * hcol[i + 1];
      hcol[i + 1] = -K->sn[i] * hcol[i] + K->cs[i] * hcol[i + 1];
      hcol[i] = a0;
    }
    const double den = sqrt(hcol[j] * hcol[j] + hn * hn);
    K->cs[j] = den > 0 ? hcol[j] / den : 1.0;
    K->sn[j] = den > 0 ? hn / den : 0.0;
    hcol[j] = den;
    K->g[j + 1] = -K->sn[j] * K->g[j];
    K->g[j] = K->cs[j] * K->g[j];
    for (int i = 0; i <= j; i++) K->H[i * ADJ_M + j] = hcol[i];
    K->hn = hn;
    K->iters += 1;
    K->jlast = j;
    K->res = fabs(K->g[j + 1]);
    const bool conv = K->res <= K->tol * K->bnorm;
    if (conv) K->solved = 1;
    if (conv || hn == 0.0 || j + 1 == m) K->cdone = 1;
  }
}

// y = R^-1 g over the cycle's columns (each block repeats the small back substitution), x += V y.
__global__ void __launch_bounds__(AB) k_akr_update_x(AD a, int nc, AdjVecs V, const AdjKry* __restrict__ K,
                                                     double* __restrict__ x) {
  const int jl = K->jlast;
  if (jl < 0) return;
  __shared__ double y[ADJ_M], Hs[ADJ_M * ADJ_M], gs[ADJ_M];   // H and g staged by parallel loads
  for (int q = threadIdx.x; q < (jl + 1) * ADJ_M; q += blockDim.x) Hs[q] = K->H[q];
  for (int q = threadIdx.x; q <= jl; q += blockDim.x) gs[q] = K->g[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = jl; i >= 0; i--) {
      double s = gs[i];
      for (int k = i + 1; k <= jl; k++) s -= Hs[i * ADJ_M + k] * y[k];
      y[i] = s / Hs[i * ADJ_M + i];
    }
  }
  __syncthreads();
  bool ok;
  const long long c = akr_index<1>(a, nc, 0, ok);
  if (!ok) return;
  double s = 0.0;
  for (int q = 0; q <= jl; q++) s += y[q] * V.p[q][c];
  x[c] += s;
}

// pn[block] = the block's share of ||v||^2
__global__ void __launch_bounds__(AB) k_akr_norm(AD a, int nc, const double* __restrict__ v, double* __restrict__ pn) {
  bool ok;
  const long long c = akr_index<1>(a, nc, 0, ok);
  double sq = 0.0;
  if (ok) { const double vv = v[c]; sq = vv * vv; }
  const double t = akr_block_sum(sq);
  if (threadIdx.x == 0) pn[akr_block()] = t;
}

// r = b - t, and pn[block] = the block's share of ||r||^2
__global__ void __launch_bounds__(AB) k_akr_sub(AD a, int nc, const double* __restrict__ b, const double* __restrict__ t,
                                                double* __restrict__ r, double* __restrict__ pn) {
  bool ok;
  const long long c = akr_index<1>(a, nc, 0, ok);
  double sq = 0.0;
  if (ok) { const double rv = b[c] - t[c]; r[c] = rv; sq = rv * rv; }
  const double s = akr_block_sum(sq);
  if (threadIdx.x == 0) pn[akr_block()] = s;
}

// ================================================================ launchers
void adj_explicit(hipStream_t st, const AD& a, const Phys& P, int kind, const AdjScal& sc,
                  const double* Qe, const double* Qs, const double* uB, const double* uout, const double* ubar,
                  double* Qebar, double* Qsbar, double* uBbar, double* fbx, double* fby, double* sbar) {
  const int k0 = kind == 0 ? 0 : 1, k1 = kind == 0 ? 4 : 5;
  k_adj_fluxdiv<<<nblk(a), AB, 0, st>>>(a, k0, k1, sc.s[SC_H], ubar, uout, Qs, Qsbar, fbx, fby, sbar);
  const AFields F = fields({Qe, uB}, {Qebar, uBbar});
  if (P.recon == IMEX_RECON_WENO5) {
    FnFlux<0, true> fx{kind, P.recon, P.central_y};
    FnFlux<1, true> fy{kind, P.recon, P.central_y};
    vjp(st, fx, a, F, fbx, sc, sbar);
    vjp(st, fy, a, F, fby, sc, sbar);
  } else {
    FnFlux<0> fx{kind, P.recon, P.central_y};
    FnFlux<1> fy{kind, P.recon, P.central_y};
    vjp(st, fx, a, F, fbx, sc, sbar);
    vjp(st, fy, a, F, fby, sc, sbar);
  }
}
void adj_mag_rhs(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* Qe, const double* pe,
                 const double* Qs, const double* bbar, double* ubar, double* Qebar, double* pebar, double* Qsbar,
                 double* sbar) {
  FnMagRHS fn{a.dx, a.dy};
  vjp(st, fn, a, fields({u, Qe, pe, Qs}, {ubar, Qebar, pebar, Qsbar}), bbar, sc, sbar);
}
void adj_mag_op(hipStream_t st, const AD& a, const AdjScal& sc, int mode, const double* x, const double* Qe,
                const double* u, const double* y, double* xbar, double* Qebar, double* ubar, double* sbar) {
  if (mode == 0) {
    FnMagOp<0> fn{a.dx, a.dy};
    vjp(st, fn, a, fields({x, Qe, u}, {xbar, nullptr, nullptr}), y, sc, sbar);
  } else {
    FnMagOp<1> fn{a.dx, a.dy};
    vjp(st, fn, a, fields({x, Qe, u}, {nullptr, Qebar, ubar}), y, sc, sbar);
  }
}
void adj_mag_jac(hipStream_t st, const AD& a, const AdjScal& sc, const double* Qe, const double* u, double* J) {
  FnMagOp<0> fn{a.dx, a.dy};
  // field 0 (the unknown) only needs finite values: the operator is linear in it
  k_jac<FnMagOp<0>><<<nblk(a), AB, 0, st>>>(fn, a, fields({Qe + 5 * a.PS, Qe, u}, {nullptr, nullptr, nullptr}), to_sc(sc), J);
}
void adj_mag_op_t(hipStream_t st, const AD& a, const double* J, const double* y, double* out) {
  k_jac_t<FnMagOp<0>><<<nblk(a), AB, 0, st>>>(a, J, y, out);
}
void adj_emf(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* Qe, const double* pe,
             const double* fbar, double* ubar, double* Qebar, double* pebar, double* sbar) {
  FnEMF fn{a.dx, a.dy};
  vjp(st, fn, a, fields({u, Qe, pe}, {ubar, Qebar, pebar}), fbar, sc, sbar);
}
void adj_pe_hhat(hipStream_t st, const AD& a, const Phys& P, const double* Qe, const double* u, const double* obar,
                 double* Qebar, double* ubar) {
  FnPeHhat fn{P};
  AdjScal z{};
  vjp(st, fn, a, fields({Qe, u}, {Qebar, ubar}), obar, z, nullptr);
}
void adj_shock(hipStream_t st, const AD& a, const Phys& P, const double* Qe, double* chi, const double* etabar,
               double* chibar, double* Qebar) {
  FnChi fn{P, P.kappa1, fmin(a.dx, a.dy), a.dx, a.dy};
  const AFields F = fields({Qe}, {Qebar});
  k_eval<FnChi><<<nblk(a), AB, 0, st>>>(fn, a, F, chi, Sc<double>{});
  k_adj_lin<<<akr_grid1(a, 1), AB, 0, st>>>(a, 1, 2, nullptr, nullptr, 0.0, chibar);
  k_adj_eta<<<nblk(a), AB, 0, st>>>(a, chi, etabar, chibar);
  AdjScal z{};
  vjp(st, fn, a, F, chibar, z, nullptr);
}
void adj_p_rhs(hipStream_t st, const AD& a, const AdjScal& sc, int pass, const double* ucur, const double* min,
               const double* hhat, const double* bbar, double* ucurbar, double* minbar, double* hhatbar, double* sbar) {
  FnPRHS fn{pass, a.dx, a.dy};
  vjp(st, fn, a, fields({ucur, min, hhat}, {ucurbar, minbar, hhatbar}), bbar, sc, sbar);
}
void adj_p_op(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, int mode, const double* x,
              const double* hhat, const double* eta, const double* visc, const double* y, double* xbar, double* hhatbar,
              double* etabar, double* viscbar, double* sbar) {
  const double av = 1.0 / (P.gamma - 1.0);
  const int rk = visc != nullptr;
  if (mode == 0) {
    FnPOp<0> fn{av, a.dx, a.dy, rk};
    vjp(st, fn, a, fields({x, hhat, eta, visc}, {xbar, nullptr, nullptr, nullptr}), y, sc, sbar);
  } else {
    FnPOp<1> fn{av, a.dx, a.dy, rk};
    vjp(st, fn, a, fields({x, hhat, eta, visc}, {nullptr, hhatbar, etabar, rk ? viscbar : nullptr}), y, sc, sbar);
  }
}
void adj_visc(hipStream_t st, const AD& a, const Phys& P, const double* pe, const double* u, double* visc) {
  FnVisc fn{P};
  k_eval<FnVisc><<<nblk(a), AB, 0, st>>>(fn, a, fields({pe, u}, {}), visc, Sc<double>{});
}
void adj_visc_vjp(hipStream_t st, const AD& a, const Phys& P, const double* pe, const double* u, const double* viscbar,
                  double* pebar, double* ubar) {
  FnVisc fn{P};
  AdjScal z{};
  vjp(st, fn, a, fields({pe, u}, {pebar, ubar}), viscbar, z, nullptr);
}
void adj_p_diag(hipStream_t st, const AD& a, const Phys& P, const AdjScal& sc, const double* hhat, double* diag) {
  k_adj_pdiag<<<nblk(a), AB, 0, st>>>(a, sc.s[SC_H], 1.0 / (P.gamma - 1.0), hhat, diag);
}
void adj_momentum(hipStream_t st, const AD& a, const AdjScal& sc, const double* umid, const double* uin,
                  const double* ubar, double* pbar, double* sbar) {
  k_adj_momentum<<<nblk(a), AB, 0, st>>>(a, sc.s[SC_H], umid, uin, ubar, pbar, sbar);
}
void adj_energy(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* u4bar,
                double* pbar, double* ubar) {
  FnEnergy fn{P};
  AdjScal z{};
  vjp(st, fn, a, fields({p, u}, {pbar, ubar}), u4bar, z, nullptr);
}
void adj_ct(hipStream_t st, const AD& a, const AdjScal& sc, const double* u, const double* bxo, const double* byo,
            const double* bxqs, const double* byqs, double* ubar, const double* bxbar, const double* bybar,
            double* bxqsbar, double* byqsbar, double* fxbar, double* fybar, double* scratch, double* sbar) {
  double* bxt = scratch;
  double* byt = scratch + a.PS;
  double* ez = scratch + 2 * a.PS;
  double* et = scratch + 3 * a.PS;
  const double h = sc.s[SC_H];
  k_adj_ct_faces<<<nblk(a), AB, 0, st>>>(a, h, ubar, bxbar, bybar, bxo, byo, bxqs, byqs, bxt, byt, bxqsbar, byqsbar, sbar);
  k_adj_ct_ezpp<<<nblk(a), AB, 0, st>>>(a, h, bxt, byt, ez, ubar);
  k_adj_ct_emf<<<nblk(a), AB, 0, st>>>(a, ez, fxbar, fybar, et);
  FnEt fn;
  AdjScal z{};
  vjp(st, fn, a, fields({u}, {ubar}), et, z, nullptr);
}
void adj_max_values(hipStream_t st, const AD& a, const Phys& P, const double* U, double* vals) {
  FnMax fn{P};
  k_eval<FnMax><<<nblk(a), AB, 0, st>>>(fn, a, fields({U}, {}), vals, Sc<double>{});
}
void adj_argmax(hipStream_t st, const AD& a, const double* vals, int* out, double* work) {
  const int nb = 64;
  double* pv = work;
  int* pi = (int*)(work + 6 * nb);
  k_argmax_part<<<nb, 256, 0, st>>>(a, vals, pv, pi);
  k_argmax_fin<<<1, 64, 0, st>>>(nb, pv, pi, out);
}
void adj_max_grad(hipStream_t st, const AD& a, const Phys& P, const double* U, const int* argmax, const double* qbar,
                  double* Ubar) {
  FnMax fn{P};
  k_max_grad<<<1, 64, 0, st>>>(fn, a, fields({U}, {}), argmax, qbar, Ubar);
}
void adj_pri2con(hipStream_t st, const AD& a, const Phys& P, const double* pri, const double* conbar, double* pribar) {
  FnPri2Con fn{P};
  AdjScal z{};
  vjp(st, fn, a, fields({pri}, {pribar}), conbar, z, nullptr);
}

static inline int nblkc(const AD& a, int nc);
int akr_blocks(const AD& a, int nc) { const dim3 g = akr_grid(a, nc); return (int)(g.x * g.y); }
int akr_blocks1(const AD& a, int nc) { const dim3 g = akr_grid1(a, nc); return (int)(g.x * g.y); }
void akr_init(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* b, double* pn, double tol) {
  k_akr_norm<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, b, pn);
  k_akr_init<<<1, 1024, 0, st>>>(K, pn, akr_blocks1(a, nc), tol);
}
void akr_residual(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* b, const double* t, double* r,
                  double* v0, double* pn) {
  k_akr_sub<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, b, t, r, pn);
  k_akr_cycle<<<1, 1024, 0, st>>>(K, pn, akr_blocks1(a, nc));
  k_akr_scale<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, K, 0, r, v0);
}
void akr_gs_pass(hipStream_t st, const AD& a, int nc, int nv, const AdjVecs& V, double* w, AdjKry* K, int d,
                 double* part, double* pn) {
  k_akr_dots<<<akr_grid(a, nc), AB, 0, st>>>(a, nc, nv, V, w, K, part);
  k_akr_dsum<<<1, 1024, 0, st>>>(K, part, akr_blocks(a, nc), nv, d);
  k_akr_project<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, nv, V, w, K, d, pn);
}
void akr_column(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* pn, int j, int m) {
  k_akr_column<<<1, 1024, 0, st>>>(K, pn, akr_blocks1(a, nc), j, m);
}
void akr_next(hipStream_t st, const AD& a, int nc, AdjKry* K, const double* w, double* v) {
  k_akr_scale<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, K, 1, w, v);
}
void akr_update_x(hipStream_t st, const AD& a, int nc, const AdjVecs& V, AdjKry* K, double* x) {
  k_akr_update_x<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, V, K, x);
}

void adj_dots(hipStream_t st, const AD& a, int nc, int nv, const double* const* v, const double* w, double* out) {
  VPtrs vp;
  for (int q = 0; q < nv; q++) vp.p[q] = v[q];
  k_adj_dots<<<256, 256, 0, st>>>(a, nc, nv, vp, w, out);
}
void adj_axpys(hipStream_t st, const AD& a, int nc, int nv, const double* coef, const double* const* v, double* w,
               double sign) {
  VPtrs vp;
  VCoef cf;
  for (int q = 0; q < nv; q++) { vp.p[q] = v[q]; cf.c[q] = coef[q]; }
  k_adj_axpys<<<(a.nx * a.ny * nc + AB - 1) / AB, AB, 0, st>>>(a, nc, nv, cf, vp, w, sign);
}
static inline int nblkc(const AD& a, int nc) { return (a.nx * a.ny * nc + AB - 1) / AB; }
void adj_scale_copy(hipStream_t st, const AD& a, int nc, const double* in, double s, double* out) {
  k_adj_lin<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, 0, in, nullptr, s, out);
}
void adj_div(hipStream_t st, const AD& a, int nc, const double* in, const double* diag, double* out) {
  k_adj_lin<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, 1, in, diag, 0.0, out);
}
void adj_newton_dv(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* mb, double* dv) {
  k_adj_newton<<<nblk(a), AB, 0, st>>>(a, P, 0, p, u, mb, nullptr, dv);
}
void adj_newton_rho(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* u, const double* mb,
                    const double* lam, double* ubar) {
  k_adj_newton<<<nblk(a), AB, 0, st>>>(a, P, 1, p, u, mb, lam, ubar);
}
void adj_slope_dv(hipStream_t st, const AD& a, const Phys& P, const double* vs, double* dv) {
  k_adj_slope_dv<<<nblk(a), AB, 0, st>>>(a, 1.0 / (P.gamma - 1.0), vs, dv);
}
void adj_diag_add(hipStream_t st, const AD& a, const double* dv, const double* in, double* out) {
  k_adj_diag_add<<<nblk(a), AB, 0, st>>>(a, dv, in, out);
}
void adj_floor_mbig(hipStream_t st, const AD& a, const Phys& P, const double* p, const double* b, const double* u,
                    double* lam, double* ubar, double* work) {
  k_adj_floor_part<<<FLB, 256, 0, st>>>(a, P, p, b, u, lam, work);
  k_adj_floor_fin<<<1, 64, 0, st>>>(a, P, u, work, lam, ubar);
}
void adj_zero(hipStream_t st, const AD& a, int nc, double* p) {
  k_adj_lin<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, 2, nullptr, nullptr, 0.0, p);
}
void adj_add(hipStream_t st, const AD& a, int nc, const double* in, double s, double* out) {
  k_adj_lin<<<akr_grid1(a, nc), AB, 0, st>>>(a, nc, 3, in, nullptr, s, out);
}
void adj_from_compact(hipStream_t st, const AD& a, int nc, const double* in, double* out) {
  k_adj_compact<<<nblkc(a, nc), AB, 0, st>>>(a, nc, 0, in, out);
}
void adj_to_compact(hipStream_t st, const AD& a, int nc, const double* in, double* out) {
  k_adj_compact<<<nblkc(a, nc), AB, 0, st>>>(a, nc, 1, in, out);
}
