// kernels_krylov.hip — matrix-free implicit operators and the device-resident GMRES / Newton pieces.
//
// Operators (Magnetic_Subsystem_AX, Pressure_Subsystem_AX, Energy_Subsystem_AX,
// Magnetic_Subsystem_potential_AX; src/implicit_operators.cpp:141-353, 505-616, 741-790, 355-503,
// GLF / MaterialSpeedp branches) are
// evaluated on 64x4 cell tiles whose input vector, with a one-cell halo mapped through M_P / M_B,
// is staged in LDS.  In Arnoldi mode the staged values are the unnormalised vector divided by its
// norm, so the normalisation v = w / ||w|| (implicit_operators.cpp:876) happens once per element
// inside the operator kernel and v is written back in the same pass.
//
// GMRES (implicit_operators.cpp:793-919): classical Gram-Schmidt as in the reference —
// all dots of step j are formed in the operator kernel from the freshly computed
// M^-1 A v_j, then one update pass forms w = Av_j - sum H_ij v_i and ||w||.  Each block
// writes its partial sums packed per block; a one-block finish kernel after each of the
// two reduces them in a fixed order (so results do not depend on block timing), and the
// second finish runs the Givens / Hessenberg bookkeeping in FP64 and raises `done` on
// convergence or at the restart length; later kernels of the same cycle see `done` and
// exit at once.
#include "kernels.hpp"

// Waves per SIMD the operator kernels are compiled for (their register budget: 512 / waves VGPRs):
// the corner-staged magnetic operator fits 128 VGPRs, the vector-potential and reflective-wall
// magnetic operators need up to 168, the scalar operators fit 128 (FP32 pressure Arnoldi: 64).
template <int OPK, class T> constexpr int op_waves() {
  return OPK == 0 ? 4 : (OPK == 3 || OPK == 4) ? 3 : (sizeof(T) == 4 ? 8 : 4);
}

// Krylov vectors of one storage type (double for the FP64 form, float for the inner solves of
// the mixed form): the basis V_i, the unnormalised next vector w and the operator image Av.
template <class VT> struct KV;
template <> struct KV<double> {
  __device__ static double* V(const Inst& I, int i) { return gmem(I.buf[B_V0 + i]); }
  __device__ static double* w(const Inst& I) { return gmem(I.buf[B_W]); }
  __device__ static double* Av(const Inst& I) { return gmem(I.buf[B_AV]); }
};
template <> struct KV<float> {
  __device__ static float* V(const Inst& I, int i) { return gmem(I.fbuf[F_V0 + i]); }
  __device__ static float* w(const Inst& I) { return gmem(I.fbuf[F_W32]); }
  __device__ static float* Av(const Inst& I) { return gmem(I.fbuf[F_AV32]); }
};

// compact-coordinate boundary map for one-cell halos (M_P with i in [g-1, n+g])
__device__ __forceinline__ int cmap(int g, int n, int per) {
  if (g < 0) return per ? g + n : 0;
  if (g >= n) return per ? (g - n < n ? g - n : n - 1) : n - 1;
  return g;
}

// Tiles are TX x TY cells with a one-cell frame, loaded by a block of TX*TY threads: thread t
// loads interior cell (t % TX, t / TX) and, for t < NHALO, frame element t (top row, bottom
// row, then the left and right columns).
#define NHALO (2 * (TX + 2) + 2 * TY)
__device__ __forceinline__ void halo_coord(int hidx, int& lx, int& ly) {
  if (hidx < TX + 2) { lx = hidx; ly = 0; }
  else if (hidx < 2 * (TX + 2)) { lx = hidx - (TX + 2); ly = TY + 1; }
  else { int k = hidx - 2 * (TX + 2); lx = (k & 1) ? TX + 1 : 0; ly = 1 + (k >> 1); }
}

template <int NC, class T, class VT>
__device__ __forceinline__ void load_tile(T* xs, const Dims& d, const VT* __restrict__ x, double h, bool scale,
                                          int gx0, int gy0) {
  const int W = TX + 2, HH = TY + 2;
  auto put = [&](int lx, int ly) {
    int gx = cmap(gx0 + lx - 1, d.nx, d.perx), gy = cmap(gy0 + ly - 1, d.ny, d.pery);
    long long e = gx + (long long)d.nx * gy;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      double v = (double)x[(long long)k * d.N + e];
      xs[k * W * HH + ly * W + lx] = (T)(scale ? v / h : v);
    }
  };
  const int t = threadIdx.x;
  put(t % TX + 1, t / TX + 1);
  if (t < NHALO) { int lx, ly; halo_coord(t, lx, ly); put(lx, ly); }
}

// ---------------------------------------------------------------- magnetic operator
// Magnetic_Subsystem_AX (implicit_operators.cpp:141-353, GLF branch) in face-flux form.
// The reference's Bx_1+Bx_2 and Bz_5+Bz_6 are fluxes through the cell's upper y-face and
// Bx_3+Bx_4, Bz_7+Bz_8 the same fluxes through its lower y-face; By_1+By_2, Bz_1+Bz_2 (and
// By_3+By_4, Bz_3+Bz_4) likewise through its right (left) x-face.  The four-cell corner
// averages are summed in the order lower-left, lower-right, upper-left, upper-right.  Apart
// from that summation order the arithmetic is the reference's expression by expression.
template <class T> struct MagSmem {
  T x[3][TY + 2][TX + 2];     // unknown B (scaled), one-cell halo through M_B
  T q[3][TY + 2][TX + 2];     // B of Qe (raw ghost frame)
  T r[TY + 2][TX + 2];        // rho of u (raw ghost frame)
};

template <class T, class VT>
__device__ __forceinline__ void mag_load(const MagOp& op, MagSmem<T>& sm, const VT* __restrict__ x, double h, bool scale,
                                         int gx0, int gy0) {
  const Dims& d = op.d;
  auto put = [&](int lx, int ly) {
    int cx = gx0 + lx - 1, cy = gy0 + ly - 1;
    int gx = cmap(cx, d.nx, d.perx), gy = cmap(cy, d.ny, d.pery);
    long long e = gx + (long long)d.nx * gy;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double v = (double)x[(long long)f * d.N + e];
      sm.x[f][ly][lx] = (T)(scale ? v / h : v);
    }
    int pi = min(cx + GC, d.NX - 1), pj = min(cy + GC, d.NY - 1);
    long long c = pidx(d, pi, pj);
#pragma unroll
    for (int f = 0; f < 3; f++) sm.q[f][ly][lx] = (T)op.Qe[(long long)(5 + f) * d.PS + c];
    sm.r[ly][lx] = (T)op.u[c];
  };
  const int t = threadIdx.x;
  // interior element, then (for t < NHALO) one frame element; kept as two passes so that the
  // seven loads of only one element are in flight at a time
#pragma unroll 1
  for (int pass = 0; pass < 2; pass++) {
    int lx = t % TX + 1, ly = t / TX + 1;
    if (pass == 1) {
      if (t >= NHALO) break;
      halo_coord(t, lx, ly);
    }
    put(lx, ly);
  }
}

// Corner (a, b) lies between LDS cells a..a+1 and b..b+1.  Its six terms: n0 b1, n0 b2, n1 b0,
// n1 b2, PC0 = |n.b|/2 - n0 b0, PC1 = |n.b|/2 - n1 b1 with n, b the four-cell averages of B^e
// and of the unknown.
template <class T, class SM>
__device__ __forceinline__ void mag_corner(const SM& sm, int a, int b, T* cn) {
  T n[3], bb[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    n[c] = T(0.25) * (sm.q[c][b][a] + sm.q[c][b][a + 1] + sm.q[c][b + 1][a] + sm.q[c][b + 1][a + 1]);
    bb[c] = T(0.25) * (sm.x[c][b][a] + sm.x[c][b][a + 1] + sm.x[c][b + 1][a] + sm.x[c][b + 1][a + 1]);
  }
  T half = T(0.5) * (n[0] * bb[0] + n[1] * bb[1] + n[2] * bb[2]);
  cn[0] = n[0] * bb[1];
  cn[1] = n[0] * bb[2];
  cn[2] = n[1] * bb[0];
  cn[3] = n[1] * bb[2];
  cn[4] = half - n[0] * bb[0];
  cn[5] = half - n[1] * bb[1];
}
// Per-cell products of LDS cell (a, b): PM - q5 x5, PM - q6 x6, q6 x5, q6 x7, q5 x6, q5 x7
// with PM = (q . x)/2.
template <class T, class SM>
__device__ __forceinline__ void mag_dv(const SM& sm, int a, int b, T* dv) {
  T q5 = sm.q[0][b][a], q6 = sm.q[1][b][a], q7 = sm.q[2][b][a];
  T x5 = sm.x[0][b][a], x6 = sm.x[1][b][a], x7 = sm.x[2][b][a];
  T q5x5 = q5 * x5, q6x6 = q6 * x6;
  T PM = T(0.5) * (q5x5 + q6x6 + q7 * x7);
  dv[0] = PM - q5x5;
  dv[1] = PM - q6x6;
  dv[2] = q6 * x5;
  dv[3] = q6 * x7;
  dv[4] = q5 * x6;
  dv[5] = q5 * x7;
}
// Fluxes (gx, gz) through the y-face between LDS rows yb and yb+1 at column a; cL, cR are the
// corners (a-1, yb), (a, yb) and D, U the cell products of (a, yb), (a, yb+1).
template <class T, class SM>
__device__ __forceinline__ void mag_yface(const SM& sm, int a, int yb, const T* cL, const T* cR,
                                          const T* D, const T* U, T idx, T idy, T midx, T& gx, T& gz) {
  T rf = T(0.5) * (sm.r[yb][a] + sm.r[yb + 1][a]);
  T nf0 = T(0.5) * (sm.q[0][yb][a] + sm.q[0][yb + 1][a]);
  T nf1 = T(0.5) * (sm.q[1][yb][a] + sm.q[1][yb + 1][a]);
  T nf2 = T(0.5) * (sm.q[2][yb][a] + sm.q[2][yb + 1][a]);
  T T1 = idx * (cR[0] - cL[0]) - idy * (U[1] - D[1]);
  gx = nf0 / rf * T1 + nf1 / rf * (idx * (cR[4] - cL[4]) - idy * (U[2] - D[2]));
  gz = nf2 / rf * T1 + nf1 / rf * (midx * (cR[1] - cL[1]) - idy * (U[3] - D[3]));
}
// Fluxes (fy, fz) through the x-face between LDS columns xa and xa+1 at row b; cB, cT are the
// corners (xa, b-1), (xa, b) and L, R the cell products of (xa, b), (xa+1, b).
template <class T, class SM>
__device__ __forceinline__ void mag_xface(const SM& sm, int xa, int b, const T* cB, const T* cT,
                                          const T* L, const T* R, T idy, T midx, T& fy, T& fz) {
  T rf = T(0.5) * (sm.r[b][xa] + sm.r[b][xa + 1]);
  T nf0 = T(0.5) * (sm.q[0][b][xa] + sm.q[0][b][xa + 1]);
  T nf1 = T(0.5) * (sm.q[1][b][xa] + sm.q[1][b][xa + 1]);
  T nf2 = T(0.5) * (sm.q[2][b][xa] + sm.q[2][b][xa + 1]);
  T dpm = R[0] - L[0];
  T T2 = midx * dpm + idy * (cT[0] - cB[0]);
  fy = nf1 / rf * T2 + nf0 / rf * (midx * (R[4] - L[4]) + idy * (cT[5] - cB[5]));
  fz = nf2 / rf * (midx * dpm + idy * (cT[2] - cB[2])) + nf0 / rf * (midx * (R[5] - L[5]) - idy * (cT[3] - cB[3]));
}

// A x for the cell of a valid thread.  The periodic / outflow operator stages the six corner terms
// of every corner of the tile once (a corner is shared by four cells) between two block barriers;
// each thread then forms the four face fluxes of its own cell from the staged tiles (a face shared
// by two cells is formed by both with the same arithmetic).  T is the arithmetic type (double, or
// float in the mixed form).
template <class T> struct MagSmemC : MagSmem<T> {
  T c[6][TY + 1][TX + 1];     // corner (a, b) between LDS cells a..a+1, b..b+1
  T k[8];                     // the operator's uniform coefficients (mag_coef), formed once per block
};
enum { MK_IDX, MK_IDY, MK_MIDX, MK_HHDX, MK_HHDY, MK_LX, MK_LY, MK_H };
// 1/dx, 1/dy, -1/dx, h^2/dx, h^2/dy, nu_x/dx^2, nu_y/dy^2 and h, each formed as mag_compute forms it
template <class T>
__device__ __forceinline__ void mag_coef(const MagOp& op, T* k) {
  const Dims& d = op.d;
  const T h = (T)op.S[S_ADT], dx = (T)d.dx, dy = (T)d.dy;
  const T nu_x = (T)op.S[S_BMAXX] * dx, nu_y = (T)op.S[S_BMAXY] * dy;
  k[MK_IDX] = T(1) / dx; k[MK_IDY] = T(1) / dy; k[MK_MIDX] = T(-1) / dx;
  k[MK_HHDX] = h * h / dx; k[MK_HHDY] = h * h / dy;
  k[MK_LX] = nu_x / (dx * dx); k[MK_LY] = nu_y / (dy * dy);
  k[MK_H] = h;
}
// Corner terms of the (TX+1) x (TY+1) corners of the tile: thread t takes corner (t % TX + 1,
// t / TX + 1); threads t < TX + TY + 1 also take a corner of the left column or the bottom row.
template <class T>
__device__ __forceinline__ void mag_corners(MagSmemC<T>& sm) {
  const int t = threadIdx.x;
  auto put = [&](int a, int b) {
    T cn[6];
    mag_corner(sm, a, b, cn);
#pragma unroll
    for (int k = 0; k < 6; k++) sm.c[k][b][a] = cn[k];
  };
  put(t % TX + 1, t / TX + 1);
  if (t < TX + TY + 1) {
    if (t <= TY) put(0, t);
    else put(t - TY, 0);
  }
}
template <class T>
__device__ __forceinline__ void mag_compute_c(const MagSmemC<T>& sm, int tx, int ty, bool valid, T* out) {
  if (!valid) return;
  const T idx = sm.k[MK_IDX], idy = sm.k[MK_IDY], midx = sm.k[MK_MIDX], h = sm.k[MK_H];
  const int a = tx + 1, b = ty + 1;
  T c00[6], c10[6], c01[6], c11[6];      // corners (a-1, b-1), (a, b-1), (a-1, b), (a, b)
#pragma unroll
  for (int k = 0; k < 6; k++) {
    c00[k] = sm.c[k][b - 1][a - 1];
    c10[k] = sm.c[k][b - 1][a];
    c01[k] = sm.c[k][b][a - 1];
    c11[k] = sm.c[k][b][a];
  }
  T dC[6], dN[6];
  mag_dv(sm, a, b, dC);
  T gxu, gzu, gxl, gzl, fyr, fzr, fyl, fzl;
  mag_dv(sm, a, b + 1, dN);
  mag_yface(sm, a, b, c01, c11, dC, dN, idx, idy, midx, gxu, gzu);
  mag_dv(sm, a, b - 1, dN);
  mag_yface(sm, a, b - 1, c00, c10, dN, dC, idx, idy, midx, gxl, gzl);
  mag_dv(sm, a + 1, b, dN);
  mag_xface(sm, a, b, c10, c11, dC, dN, idy, midx, fyr, fzr);
  mag_dv(sm, a - 1, b, dN);
  mag_xface(sm, a - 1, b, c00, c01, dN, dC, idy, midx, fyl, fzl);
  auto X = [&](int c, int di, int dj) { return sm.x[c][b + dj][a + di]; };
  const T hhdx = sm.k[MK_HHDX], hhdy = sm.k[MK_HHDY];
  T ax5 = X(0, 0, 0) + hhdy * (gxu - gxl);
  T ax6 = X(1, 0, 0) + hhdx * (fyr - fyl);
  T ax7 = X(2, 0, 0) + hhdx * (fzr - fzl) + hhdy * (gzu - gzl);
  const T lx = sm.k[MK_LX], ly = sm.k[MK_LY];
  T lap[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    lap[c] = lx * (X(c, -1, 0) - T(2) * X(c, 0, 0) + X(c, 1, 0)) + ly * (X(c, 0, -1) - T(2) * X(c, 0, 0) + X(c, 0, 1));
  out[0] = ax5 - h * lap[0];
  out[1] = ax6 - h * lap[1];
  out[2] = ax7 - h * lap[2];
}

// The same for a tile type SM (the reflective path's local neighbourhood).
template <class T, class SM>
__device__ __forceinline__ void mag_compute_at(const MagOp& op, const SM& sm, int a, int b, T* out) {
  const Dims& d = op.d;
  const T h = (T)op.S[S_ADT], dx = (T)d.dx, dy = (T)d.dy;
  const T idx = T(1) / dx, idy = T(1) / dy, midx = T(-1) / dx;
  T c00[6], c10[6], c01[6], c11[6];      // corners (a-1, b-1), (a, b-1), (a-1, b), (a, b)
  mag_corner(sm, a - 1, b - 1, c00);
  mag_corner(sm, a, b - 1, c10);
  mag_corner(sm, a - 1, b, c01);
  mag_corner(sm, a, b, c11);
  T dC[6], dN[6];
  mag_dv(sm, a, b, dC);
  T gxu, gzu, gxl, gzl, fyr, fzr, fyl, fzl;
  mag_dv(sm, a, b + 1, dN);
  mag_yface(sm, a, b, c01, c11, dC, dN, idx, idy, midx, gxu, gzu);
  mag_dv(sm, a, b - 1, dN);
  mag_yface(sm, a, b - 1, c00, c10, dN, dC, idx, idy, midx, gxl, gzl);
  mag_dv(sm, a + 1, b, dN);
  mag_xface(sm, a, b, c10, c11, dC, dN, idy, midx, fyr, fzr);
  mag_dv(sm, a - 1, b, dN);
  mag_xface(sm, a - 1, b, c00, c01, dN, dC, idy, midx, fyl, fzl);
  auto X = [&](int c, int di, int dj) { return sm.x[c][b + dj][a + di]; };
  T ax5 = X(0, 0, 0) + (h * h / dy) * (gxu - gxl);
  T ax6 = X(1, 0, 0) + (h * h / dx) * (fyr - fyl);
  T ax7 = X(2, 0, 0) + h * h / dx * (fzr - fzl) + h * h / dy * (gzu - gzl);
  const T nu_x = (T)op.S[S_BMAXX] * dx, nu_y = (T)op.S[S_BMAXY] * dy;
  T lap[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    lap[c] = nu_x / (dx * dx) * (X(c, -1, 0) - T(2) * X(c, 0, 0) + X(c, 1, 0)) + nu_y / (dy * dy) * (X(c, 0, -1) - T(2) * X(c, 0, 0) + X(c, 0, 1));
  out[0] = ax5 - h * lap[0];
  out[1] = ax6 - h * lap[1];
  out[2] = ax7 - h * lap[2];
}

// 3x3 neighbourhood of one cell, the unknown's neighbours multiplied by the reflective flags
template <class T> struct MagLocal {
  T x[3][3][3];
  T q[3][3][3];
  T r[3][3];
};
// Reflective boundaries (implicit_operators.cpp:146-211): the reference negates a neighbour's unknown
// whenever the clamped neighbour index is the first (last) interior cell, i.e. at the boundary cell and
// at the cell next to it; the centre is never negated.  Cells away from reflective walls take the tile
// as it is.
template <class T>
__device__ __forceinline__ void mag_compute_refl(const MagOp& op, const MagSmem<T>& sm, int tx, int ty, int i, int j, bool valid,
                                                 T* out) {
  if (!valid) return;
  const Dims& d = op.d;
  const int a = tx + 1, b = ty + 1;
  const T fxl = (d.refx && i <= GC + 1) ? T(-1) : T(1), fxr = (d.refx && i >= d.nx + GC - 2) ? T(-1) : T(1);
  const T fyl = (d.refy && j <= GC + 1) ? T(-1) : T(1), fyr = (d.refy && j >= d.ny + GC - 2) ? T(-1) : T(1);
  if (fxl > T(0) && fxr > T(0) && fyl > T(0) && fyr > T(0)) { mag_compute_at(op, sm, a, b, out); return; }
  MagLocal<T> L;
  const T fx[3] = {fxl, T(1), fxr}, fy[3] = {fyl, T(1), fyr};
  for (int dj = 0; dj < 3; dj++)
    for (int di = 0; di < 3; di++) {
      for (int c = 0; c < 3; c++) {
        L.x[c][dj][di] = sm.x[c][b + dj - 1][a + di - 1] * (fx[di] * fy[dj]);
        L.q[c][dj][di] = sm.q[c][b + dj - 1][a + di - 1];
      }
      L.r[dj][di] = sm.r[b + dj - 1][a + di - 1];
    }
  mag_compute_at(op, L, 1, 1, out);
}

// ---------------------------------------------------------------- pressure operator
__device__ __forceinline__ double p_vslope(const POp& op, int cidx, double rho) {
  if (op.vslope) return op.vslope[cidx];
  const Phys& P = op.P;
  double p = op.pfreeze[cidx];
  if (!P.floor_on) return 1.0 / (P.gamma - 1.0);
  return (p >= P.p_floor) ? 1.0 / (P.gamma - 1.0) : op.S[S_MBIG];
}
template <class T>
__device__ __forceinline__ void p_apply(const POp& op, const T* xs, int lx, int ly, int i, int j, int cidx, T* out) {
  const Dims& d = op.d;
  const int W = TX + 2;
  const T h = (T)op.S[S_ADT], dx = (T)d.dx, dy = (T)d.dy;
  auto X = [&](int di, int dj) { return xs[(ly + dj) * W + lx + di]; };
  auto H = [&](int di, int dj) { return (T)op.hhat[pidx(d, i + di, j + dj)]; };
  auto E = [&](int di, int dj) { return (T)op.eta[pidx(d, i + di, j + dj)]; };
  T h00 = H(0, 0), hp0 = H(1, 0), hm0 = H(-1, 0), h0p = H(0, 1), h0m = H(0, -1);
  T pii = X(0, 0), xp0 = X(1, 0), xm0 = X(-1, 0), x0p = X(0, 1), x0m = X(0, -1);
  if (dir_bnd(d, i, j)) { out[0] = pii; return; }     // strong Dirichlet: identity row (implicit_operators.cpp:549-550)
  // jet inflow strip: the ghost below the first row is a homogeneous Dirichlet unknown in the enthalpy
  // term only; its known value is lifted into the right-hand side (implicit_operators.cpp:559-561)
  const T p_0m = (j == GC && jet_inlet(d, i)) ? T(0) : x0m;
  double rho = op.u[pidx(d, i, j)];
  T Vdiag = (T)(op.jacobian ? p_vslope(op, cidx, rho) : (1.0 / (op.P.gamma - 1.0)));
  T ax = Vdiag * pii - h * h * (T(1) / (dx * dx) * (T(0.5) * (hp0 + h00) * (xp0 - pii) - T(0.5) * (h00 + hm0) * (pii - xm0))
                              + T(1) / (dy * dy) * (T(0.5) * (h0p + h00) * (x0p - pii) - T(0.5) * (h0m + h00) * (pii - p_0m)));
  T e00 = E(0, 0);
  T nu_x = (T)op.S[S_VXMAX] * dx, nu_y = (T)op.S[S_VYMAX] * dy;
  T lap = nu_x / (dx * dx) * (T(0.5) * (E(1, 0) + e00) * (xp0 - pii) - T(0.5) * (e00 + E(-1, 0)) * (pii - xm0))
        + nu_y / (dy * dy) * (T(0.5) * (E(0, 1) + e00) * (x0p - pii) - T(0.5) * (e00 + E(0, -1)) * (pii - x0m));
  T visc = (T)(op.visc ? op.visc[pidx(d, i, j)] : 1.0 / (op.P.gamma - 1.0));
  out[0] = ax - visc * h * lap;
}
template <class T>
__device__ __forceinline__ void p_precond(const POp& op, int cidx, int i, int j, T* y) {
  if (!op.precond || dir_bnd(op.d, i, j)) return;     // identity rows are left alone (implicit_operators.cpp:652-655)
  double rho = op.u[pidx(op.d, i, j)];
  double Vd = op.jacobian ? p_vslope(op, cidx, rho) : 1.0 / (op.P.gamma - 1.0);
  y[0] = y[0] / (T)(Vd + op.Tdiag[cidx]);
}

// ---------------------------------------------------------------- energy operator
// Energy_Subsystem_AX (implicit_operators.cpp:741-790): (E - (gamma-1) h^2 div(hh grad E)) dx dy with
// face coefficients hh the means of the cell coefficients; precE divides by diag dx dy.
template <class T>
__device__ __forceinline__ void e_apply(const POp& op, const T* xs, int lx, int ly, int i, int j, T* out) {
  const Dims& d = op.d;
  const int W = TX + 2;
  const T h = (T)op.S[S_ADT], dx = (T)d.dx, dy = (T)d.dy, gm1 = (T)(op.P.gamma - 1.0);
  auto X = [&](int di, int dj) { return xs[(ly + dj) * W + lx + di]; };
  auto H = [&](int di, int dj) { return (T)op.hhat[pidx(d, i + di, j + dj)]; };
  T h00 = H(0, 0);
  T hhp_x = T(0.5) * (h00 + H(1, 0)), hhm_x = T(0.5) * (h00 + H(-1, 0));
  T hhp_y = T(0.5) * (h00 + H(0, 1)), hhm_y = T(0.5) * (h00 + H(0, -1));
  T Ei = X(0, 0);
  // getE (:772-779): Dirichlet ghosts and the jet strip's ghost are homogeneous
  auto GE = [&](int di, int dj) {
    return (dir_ghost(d, i + di, j + dj) || (dj < 0 && j == GC && jet_inlet(d, i))) ? T(0) : X(di, dj);
  };
  T ax = Ei - gm1 * h * h * (T(1) / (dx * dx) * (hhp_x * (GE(1, 0) - Ei) - hhm_x * (Ei - GE(-1, 0)))
                           + T(1) / (dy * dy) * (hhp_y * (GE(0, 1) - Ei) - hhm_y * (Ei - GE(0, -1))));
  out[0] = ax * dx * dy;
}
template <class T>
__device__ __forceinline__ void e_precond(const POp& op, int cidx, T* y) {
  if (op.precond) y[0] = y[0] / (T)op.Tdiag[cidx];
}

// Tile interface shared by the GMRES kernels: load the input vector tile (plus whatever the
// operator stages), then compute M^-1 A x for the block's cells in arithmetic type T.
template <class T> struct PSmem { T x[(TY + 2) * (TX + 2)]; };
template <int OPK, class T> struct OpSel;
// Every operator: init (once per block, before the first tile's barrier), load (stage the tile),
// prep (a second pass over the staged tile between two block barriers, when PREP), compute
// (M^-1 A x of one cell).
template <class T> struct OpSel<0, T> {
  static constexpr int NC = 3;
  static constexpr bool PREP = true;
  typedef MagOp Op;
  typedef MagIds Ids;
  __device__ static Op make(const Inst& I, const Phys&, const Ids& ids) { return make_mag_op(I, ids); }
  typedef MagSmemC<T> Smem;
  template <class VT>
  __device__ static void load(const MagOp& op, Smem& sm, const VT* x, double h, bool scale, int gx0, int gy0) {
    mag_load(op, sm, x, h, scale, gx0, gy0);
  }
  __device__ static void init(const MagOp& op, Smem& sm) { if (threadIdx.x == 0) mag_coef(op, sm.k); }
  __device__ static void prep(const MagOp&, Smem& sm) { mag_corners(sm); }
  __device__ static void compute(const MagOp&, Smem& sm, int tx, int ty, int, int, int, bool valid, T* y) {
    mag_compute_c(sm, tx, ty, valid, y);
  }
  __device__ static T center(const Smem& sm, int c, int tx, int ty) { return sm.x[c][ty + 1][tx + 1]; }
  __device__ static void precond(const MagOp&, int, int, int, T*) {}
  __device__ static const Dims& dims(const MagOp& op) { return op.d; }
};
// The magnetic operator with the reflective-wall neighbour flags: its own instantiation, so that the
// flag-free operator keeps its register allocation; each cell forms its own corners (a cell next to
// a wall sees its neighbours with the wall's signs)
template <class T> struct OpSel<4, T> : OpSel<0, T> {
  static constexpr bool PREP = false;
  typedef MagSmem<T> Smem;
  template <class VT>
  __device__ static void load(const MagOp& op, Smem& sm, const VT* x, double h, bool scale, int gx0, int gy0) {
    mag_load(op, sm, x, h, scale, gx0, gy0);
  }
  __device__ static void init(const MagOp&, Smem&) {}
  __device__ static void prep(const MagOp&, Smem&) {}
  __device__ static void compute(const MagOp& op, Smem& sm, int tx, int ty, int i, int j, int, bool valid, T* y) {
    mag_compute_refl(op, sm, tx, ty, i, j, valid, y);
  }
  __device__ static T center(const Smem& sm, int c, int tx, int ty) { return sm.x[c][ty + 1][tx + 1]; }
};
template <class T> struct OpSel<1, T> {
  static constexpr int NC = 1;
  static constexpr bool PREP = false;
  typedef POp Op;
  typedef PIds Ids;
  __device__ static Op make(const Inst& I, const Phys& P, const Ids& ids) { return make_p_op(I, P, ids); }
  typedef PSmem<T> Smem;
  template <class VT>
  __device__ static void load(const POp& op, Smem& sm, const VT* x, double h, bool scale, int gx0, int gy0) {
    load_tile<1>(sm.x, op.d, x, h, scale, gx0, gy0);
  }
  __device__ static void compute(const POp& op, Smem& sm, int tx, int ty, int i, int j, int cidx, bool valid, T* y) {
    if (valid) p_apply(op, sm.x, tx + 1, ty + 1, i, j, cidx, y);
  }
  __device__ static void init(const POp&, Smem&) {}
  __device__ static void prep(const POp&, Smem&) {}
  __device__ static T center(const Smem& sm, int, int tx, int ty) { return sm.x[(ty + 1) * (TX + 2) + tx + 1]; }
  __device__ static void precond(const POp& op, int cidx, int i, int j, T* y) { p_precond(op, cidx, i, j, y); }
  __device__ static const Dims& dims(const POp& op) { return op.d; }
};

template <class T> struct OpSel<2, T> : OpSel<1, T> {
  typedef PSmem<T> Smem;
  __device__ static void compute(const POp& op, Smem& sm, int tx, int ty, int i, int j, int, bool valid, T* y) {
    if (valid) e_apply(op, sm.x, tx + 1, ty + 1, i, j, y);
  }
  __device__ static void precond(const POp& op, int cidx, int, int, T* y) { e_precond(op, cidx, y); }
};

// ---------------------------------------------------------------- vector-potential operator
// Magnetic_Subsystem_potential_AX (implicit_operators.cpp:355-503), GLF branch with add_bvisc:
// A - h B^n x (h / rho) div T(b) - h nu Laplacian(A), with b = curl A at the four face neighbours and
// T(alpha, beta) = (B^n . b)/2 delta - B^n_beta b_alpha.  The unknown's ghost values come from
// bound_cond_A applied to the vector (periodic copy or linear extrapolation), evaluated here in
// the loader on a two-cell frame; in Arnoldi mode the extrapolation acts on the scaled values,
// as the reference's acts on the stored basis vector.
template <class T> struct PotSmem { T x[3][TY + 4][TX + 4]; };

template <class VT>
__device__ __forceinline__ double pot_ghost(const Dims& d, const VT* __restrict__ x, int f, double h, bool scale,
                                            int cx, int cy) {
  auto V = [&](int a, int b) {
    double v = (double)x[(long long)f * d.N + a + (long long)d.nx * b];
    return scale ? v / h : v;
  };
  auto X = [&](int a, int b) -> double {      // the x rule on row b (interior)
    if (a >= 0 && a < d.nx) return V(a, b);
    if (d.perx) return V(a < 0 ? a + d.nx : a - d.nx, b);
    if (a < 0) { double v0 = V(0, b), sL = V(1, b) - v0; return v0 - (-a) * sL; }
    double vr = V(d.nx - 1, b), sR = vr - V(d.nx - 2, b);
    return vr + (a - (d.nx - 1)) * sR;
  };
  if (cy >= 0 && cy < d.ny) return X(cx, cy);
  if (cy < 0 && jet_inlet(d, cx + GC)) return X(cx, -1 - cy);      // even reflection in the jet strip
  if (d.pery) return X(cx, cy < 0 ? cy + d.ny : cy - d.ny);
  if (cy < 0) { double x0 = X(cx, 0), sL = X(cx, 1) - x0; return x0 - (-cy) * sL; }
  double xr = X(cx, d.ny - 1), sR = xr - X(cx, d.ny - 2);
  return xr + (cy - (d.ny - 1)) * sR;
}

template <class T, class VT>
__device__ __forceinline__ void pot_load(const PotOp& op, PotSmem<T>& sm, const VT* __restrict__ x, double h, bool scale,
                                         int gx0, int gy0) {
  const Dims& d = op.d;
  constexpr int W = TX + 4, H = TY + 4;
  for (int e = threadIdx.x; e < W * H; e += blockDim.x) {
    int lx = e % W, ly = e / W;
    int cx = gx0 + lx - 2, cy = gy0 + ly - 2;
    bool used = cx <= d.nx + 1 && cy <= d.ny + 1;        // frames of a partial tile beyond it are never read
#pragma unroll
    for (int f = 0; f < 3; f++) sm.x[f][ly][lx] = used ? (T)pot_ghost(d, x, f, h, scale, cx, cy) : T(0);
  }
}

template <class T>
__device__ __forceinline__ void pot_compute(const PotOp& op, const PotSmem<T>& sm, int tx, int ty, int i, int j, bool valid, T* out) {
  if (!valid) return;
  const Dims& d = op.d;
  const T h = (T)op.S[S_ADT], dx = (T)d.dx, dy = (T)d.dy;
  const int a0 = tx + 2, b0 = ty + 2;
  auto a = [&](int di, int dj, int k) { return sm.x[k][b0 + dj][a0 + di]; };
  auto Q = [&](int di, int dj, int k) { return (T)op.Qe[(long long)k * d.PS + pidx(d, i + di, j + dj)]; };
  const T two_dx = T(2.0) * dx, two_dy = T(2.0) * dy;
  T bx_P0 = (a(1, 1, 2) - a(1, -1, 2)) / two_dy;
  T by_P0 = -(a(2, 0, 2) - a(0, 0, 2)) / two_dx;
  T bz_P0 = (a(2, 0, 1) - a(0, 0, 1)) / two_dx - (a(1, 1, 0) - a(1, -1, 0)) / two_dy;
  T bx_M0 = (a(-1, 1, 2) - a(-1, -1, 2)) / two_dy;
  T by_M0 = -(a(0, 0, 2) - a(-2, 0, 2)) / two_dx;
  T bz_M0 = (a(0, 0, 1) - a(-2, 0, 1)) / two_dx - (a(-1, 1, 0) - a(-1, -1, 0)) / two_dy;
  T bx_0P = (a(0, 2, 2) - a(0, 0, 2)) / two_dy;
  T by_0P = -(a(1, 1, 2) - a(-1, 1, 2)) / two_dx;
  T bz_0P = (a(1, 1, 1) - a(-1, 1, 1)) / two_dx - (a(0, 2, 0) - a(0, 0, 0)) / two_dy;
  T bx_0M = (a(0, 0, 2) - a(0, -2, 2)) / two_dy;
  T by_0M = -(a(1, -1, 2) - a(-1, -1, 2)) / two_dx;
  T bz_0M = (a(1, -1, 1) - a(-1, -1, 1)) / two_dx - (a(0, 0, 0) - a(0, -2, 0)) / two_dy;
  T qx, qy, qz;
  qx = Q(1, 0, 5); qy = Q(1, 0, 6); qz = Q(1, 0, 7);
  T bd = qx * bx_P0 + qy * by_P0 + qz * bz_P0;
  T Txx_P0 = bd / T(2.0) - qx * bx_P0, Tyx_P0 = -qx * by_P0, Tzx_P0 = -qx * bz_P0;
  qx = Q(-1, 0, 5); qy = Q(-1, 0, 6); qz = Q(-1, 0, 7);
  bd = qx * bx_M0 + qy * by_M0 + qz * bz_M0;
  T Txx_M0 = bd / T(2.0) - qx * bx_M0, Tyx_M0 = -qx * by_M0, Tzx_M0 = -qx * bz_M0;
  qx = Q(0, 1, 5); qy = Q(0, 1, 6); qz = Q(0, 1, 7);
  bd = qx * bx_0P + qy * by_0P + qz * bz_0P;
  T Txy_0P = -qy * bx_0P, Tyy_0P = bd / T(2.0) - qy * by_0P, Tzy_0P = -qy * bz_0P;
  qx = Q(0, -1, 5); qy = Q(0, -1, 6); qz = Q(0, -1, 7);
  bd = qx * bx_0M + qy * by_0M + qz * bz_0M;
  T Txy_0M = -qy * bx_0M, Tyy_0M = bd / T(2.0) - qy * by_0M, Tzy_0M = -qy * bz_0M;
  T divT_x = (Txx_P0 - Txx_M0) / two_dx + (Txy_0P - Txy_0M) / two_dy;
  T divT_y = (Tyx_P0 - Tyx_M0) / two_dx + (Tyy_0P - Tyy_0M) / two_dy;
  T divT_z = (Tzx_P0 - Tzx_M0) / two_dx + (Tzy_0P - Tzy_0M) / two_dy;
  const T hr = h / (T)op.u[pidx(d, i, j)];
  T Fx = hr * divT_x, Fy = hr * divT_y, Fz = hr * divT_z;
  qx = Q(0, 0, 5); qy = Q(0, 0, 6); qz = Q(0, 0, 7);
  T cr[3] = {qy * Fz - qz * Fy, qz * Fx - qx * Fz, qx * Fy - qy * Fx};
  const T nu_x = (T)op.S[S_GBX] * dx, nu_y = (T)op.S[S_GBY] * dy;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    T lap = nu_x / (dx * dx) * (a(-1, 0, k) - T(2.0) * a(0, 0, k) + a(1, 0, k))
          + nu_y / (dy * dy) * (a(0, -1, k) - T(2.0) * a(0, 0, k) + a(0, 1, k));
    out[k] = a(0, 0, k) - h * cr[k] - h * lap;
  }
}

template <class T> struct OpSel<3, T> {
  static constexpr int NC = 3;
  static constexpr bool PREP = false;
  typedef PotOp Op;
  typedef MagIds Ids;
  __device__ static Op make(const Inst& I, const Phys&, const Ids& ids) { return make_pot_op(I, ids); }
  typedef PotSmem<T> Smem;
  template <class VT>
  __device__ static void load(const PotOp& op, Smem& sm, const VT* x, double h, bool scale, int gx0, int gy0) {
    pot_load(op, sm, x, h, scale, gx0, gy0);
  }
  __device__ static void init(const PotOp&, Smem&) {}
  __device__ static void prep(const PotOp&, Smem&) {}
  __device__ static void compute(const PotOp& op, Smem& sm, int tx, int ty, int i, int j, int, bool valid, T* y) {
    pot_compute(op, sm, tx, ty, i, j, valid, y);
  }
  __device__ static T center(const Smem& sm, int c, int tx, int ty) { return sm.x[c][ty + 2][tx + 2]; }
  __device__ static void precond(const PotOp&, int, int, int, T*) {}
  __device__ static const Dims& dims(const PotOp& op) { return op.d; }
};

// ---------------------------------------------------------------- reductions
// Sum `nv` per-thread values over the block; results in thread v < nv, written to out[v].
template <int MAXV>
__device__ __forceinline__ void block_sums(const double* acc, int nv, double* red, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int v = 0; v < MAXV; v++) {
    if (v < nv) {
      double s = wave_sum(acc[v]);
      if (lane == 0) red[wid * MAXV + v] = s;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[w * MAXV + threadIdx.x];
    out[threadIdx.x] = s;
  }
}
// The one-block finishes run FIN_THREADS threads: their cost is the chain of dependent loads each
// thread makes over the partial rows (nblk / FIN_THREADS of them), so wide blocks keep it short.
constexpr int FIN_THREADS = 1024;

// One block: out[v] = sum over blocks b < nblk of partial[b*stride + v], v < nv <= NV, summed in
// a fixed order, so the result does not depend on timing.
//  NV <= 2 (norms): thread t sums rows t, t + blockDim.x, ... in that order, then the waves.
//  NV >= 4 (Arnoldi dots): VW lanes share a row (lane l reads value l % VW), so one load
//  instruction of a wave covers 64 / VW rows in contiguous pieces instead of 64 scattered rows;
//  each thread loads eight of its rows before adding any (every load unconditional: rows past
//  nblk repeat the first, values past nv read value 0 and are not added) and adds them in row
//  order; the per-lane sums are reduced over the lanes of equal v, then over the waves.
// (Measured: the row-sharing form takes the dots finish from 6 to 4.6 us at j < 8 and from 14.5
// to 4.8 us at 8 <= j < 16, but makes the one- and two-value finishes slower.)
template <int NV>
__device__ __forceinline__ void final_sums(const double* partial, int stride, int nblk, int nv, double* out) {
  if constexpr (NV <= 2) {
    __shared__ double red[FIN_THREADS / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0.0;
#pragma unroll 4
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
      const double* row = partial + (long long)b * stride;
#pragma unroll
      for (int v = 0; v < NV; v++) if (v < nv) acc[v] += row[v];
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      if (v < nv) {
        double s = wave_sum(acc[v]);
        if (lane == 0) red[wid][v] = s;
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
      double s = 0.0;
      for (int w = 0; w < nw; w++) s += red[w][threadIdx.x];
      out[threadIdx.x] = s;
    }
    return;
  }
  constexpr int VW = NV <= 1 ? 1 : NV <= 2 ? 2 : NV <= 4 ? 4 : NV <= 8 ? 8 : NV <= 16 ? 16 : 32;
  constexpr int RB = 8;
  __shared__ double red[FIN_THREADS / 64][VW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int v = lane % VW;
  const bool on = v < nv;
  const int vv = on ? v : 0;
  const int R = nw * (64 / VW);                       // row slots of the block
  const int slot = wid * (64 / VW) + lane / VW;
  double acc = 0.0;
  for (int b0 = slot; b0 < nblk; b0 += RB * R) {
    double x[RB];
#pragma unroll
    for (int q = 0; q < RB; q++) {
      const int b = b0 + q * R;
      x[q] = partial[(long long)(b < nblk ? b : b0) * stride + vv];
    }
#pragma unroll
    for (int q = 0; q < RB; q++)
      if (on && b0 + q * R < nblk) acc += x[q];
  }
  for (int o = 32; o >= VW; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane < VW) red[wid][lane] = acc;
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------- GMRES residual r0 = M^-1 (b - A x), gam0 = ||r0||
// Always in FP64 (it is the outer residual of the mixed form's iterative refinement); r0 is
// stored in the basis type VT, its norm is taken in FP64 before rounding.
template <int OPK, class VT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(op_waves<OPK, double>())))
k_gmres_residual(typename OpSel<OPK, double>::Ids ids, Phys P, const Inst* __restrict__ IT, BMap bm, int ix, int ib,
                 int outer, int x_is_zero, int skip, int mixed) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (skipped(I, skip)) {        // the caller's outer iteration has already converged: no solve
    if (bx_ == 0 && threadIdx.x == 0) {
      K->done = 1; K->converged = 1; K->jlast = -1; K->steps = 0; K->outer = outer; K->outer_conv = 1;
      if (outer == 0) K->steps_total = 0;
    }
    return;
  }
  if (outer > 0 && K->outer_conv) {      // solved in an earlier sweep or restart cycle: nothing to update
    if (bx_ == 0 && threadIdx.x == 0) K->jlast = -1;
    return;
  }
  typedef OpSel<OPK, double> S;
  constexpr int NC = S::NC;
  const typename S::Op op = S::make(I, P, ids);
  const double* __restrict__ x = gmem(I.buf[ix]);
  const double* __restrict__ b = gmem(I.buf[ib]);
  VT* __restrict__ w = KV<VT>::w(I);
  __shared__ typename S::Smem sm;
  __shared__ double red[8];
  __shared__ double res[2];
  const Dims& d = S::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc[2] = {0.0, 0.0};     // ||M^-1 r||^2, and ||M^-1 b||^2 on the first sweep of the mixed form
  S::init(op, sm);
  for (int tile = bx_; tile < ntx * nty; tile += gx_) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    double y[NC];
    if (!x_is_zero) {
      __syncthreads();
      S::load(op, sm, x, 1.0, false, gx0, gy0);
      __syncthreads();
      if constexpr (S::PREP) { S::prep(op, sm); __syncthreads(); }
      S::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    }
    if (valid) {
      if (mixed && outer == 0) {
        double bb[NC];
        for (int c = 0; c < NC; c++) bb[c] = b[(long long)c * d.N + cidx];
        S::precond(op, cidx, cx + GC, cy + GC, bb);
        for (int c = 0; c < NC; c++) acc[1] += bb[c] * bb[c];
      }
      if (x_is_zero) for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx];
      else for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx] - y[c];
      S::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) { w[(long long)c * d.N + cidx] = (VT)y[c]; acc[0] += y[c] * y[c]; }
    }
  }
  block_sums<2>(acc, 2, red, res);
  __syncthreads();
  if (threadIdx.x < 2) gmem(I.partial)[bx_ * 2 + threadIdx.x] = res[threadIdx.x];   // packed pairs: coalesced finish
}

// local block count of instance i in the map of the kernel whose partials a finish reduces
#define SRC_BLOCKS(src) bmap_count(src, inst_)

// One-block finish of the residual norm: gam0, tolerance, convergence (implicit_operators.cpp:829-836).
// Mixed form: the solve is converged when the FP64 residual is below outer_tol ||M^-1 b|| +
// gmres_tol; otherwise the inner solve of this sweep runs to inner_tol relative to this residual.
__global__ void __launch_bounds__(FIN_THREADS) k_gmres_residual_fin(const Inst* __restrict__ IT, BMap bm, BMap src, int outer,
                                                            int skip, int mixed, double outer_tol, double inner_tol) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (skipped(I, skip)) return;
  if (outer > 0 && K->outer_conv) return;
  __shared__ double tot[2];
  final_sums<2>(gmem(I.partial), 2, SRC_BLOCKS(src), 2, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    double g0 = sqrt(tot[0]);
    if (outer == 0) {
      K->tol_eff = K->tol * g0 + K->tol; K->res0 = g0; K->steps_total = 0; K->outer_conv = 0;
      K->tol_outer = outer_tol * sqrt(tot[1]) + K->tol;
    }
    K->res = g0;
    K->gam[0] = g0;
    K->outer = outer;
    K->steps = 0;
    if (outer < MAXSWEEP) K->sweep_steps[outer] = 0;
    const double tol = mixed ? K->tol_outer : K->tol_eff;
    if (g0 <= tol) { K->done = 1; K->converged = 1; K->jlast = -1; K->outer_conv = 1; }
    else {
      K->done = 0; K->converged = 0; K->jlast = -1; K->hcur = g0;
      if (mixed) K->tol_eff = inner_tol * g0;
    }
  }
}

// ---------------------------------------------------------------- Arnoldi step j, part 1:
// v_j = w / h (written), Av_j = M^-1 A v_j (written), H[i][j] = v_i . Av_j for i <= j.
// T is the operator's arithmetic type, VT the basis storage type; dots are summed in FP64.
// The dots of each tile are reduced over the wave right away (a uniform loop over i, so no
// accumulator array is live across the operator) and summed per wave in LDS, in tile order.
template <int OPK, class T, class VT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(op_waves<OPK, T>())))
k_arnoldi_apply(typename OpSel<OPK, T>::Ids ids, Phys P, const Inst* __restrict__ IT, BMap bm, int j) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (K->done) return;
  typedef OpSel<OPK, T> S;
  constexpr int NC = S::NC;
  const typename S::Op op = S::make(I, P, ids);
  const VT* __restrict__ w = KV<VT>::w(I);
  VT* __restrict__ Av = KV<VT>::Av(I);
  __shared__ typename S::Smem sm;
  __shared__ double wsum[4][MAXM + 1];
  const Dims& d = S::dims(op);
  const double h = K->hcur;
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = lane; i <= j; i += 64) wsum[wid][i] = 0.0;
  VT* vj = KV<VT>::V(I, j);
  S::init(op, sm);
  for (int tile = bx_; tile < ntx * nty; tile += gx_) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    __syncthreads();
    S::load(op, sm, w, h, true, gx0, gy0);
    __syncthreads();
    if constexpr (S::PREP) { S::prep(op, sm); __syncthreads(); }
    T y[NC];
    double vc[NC];
    S::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    if (valid) {
      S::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) {
        long long e = (long long)c * d.N + cidx;
        VT vs = (VT)S::center(sm, c, tx, ty);
        VT as = (VT)y[c];
        vj[e] = vs;
        Av[e] = as;
        vc[c] = (double)vs;
        y[c] = (T)as;
      }
    }
    if constexpr (NC == 1) {
      // Scalar operators: the stored v_i are fetched eight at a time before any of them is used,
      // so their loads are in flight together instead of one HBM round trip per i.
      constexpr int CH = 8;
      const long long cl = valid ? cidx : 0;
      for (int i0 = 0; i0 <= j; i0 += CH) {
        VT vv[CH];
#pragma unroll
        for (int q = 0; q < CH; q++)
          if (i0 + q < j) vv[q] = KV<VT>::V(I, i0 + q)[cl];
#pragma unroll
        for (int q = 0; q < CH; q++) {
          const int i = i0 + q;
          if (i > j) break;
          double part = 0.0;
          if (valid) part += (i < j ? (double)vv[q] : vc[0]) * (double)y[0];
          double s = wave_sum(part);
          if (lane == 0) wsum[wid][i] += s;
        }
      }
    } else {
      // the same for the three-component operators, four stored vectors (all components) at a time
      constexpr int CH = 4;
      const long long cl = valid ? cidx : 0;
      for (int i0 = 0; i0 <= j; i0 += CH) {
        VT vv[CH][NC];
#pragma unroll
        for (int q = 0; q < CH; q++)
          if (i0 + q < j) {
            const VT* vi = KV<VT>::V(I, i0 + q);
#pragma unroll
            for (int c = 0; c < NC; c++) vv[q][c] = vi[(long long)c * d.N + cl];
          }
#pragma unroll
        for (int q = 0; q < CH; q++) {
          const int i = i0 + q;
          if (i > j) break;
          double part = 0.0;
          if (valid)
            for (int c = 0; c < NC; c++) part += (i < j ? (double)vv[q][c] : vc[c]) * (double)y[c];
          double s = wave_sum(part);
          if (lane == 0) wsum[wid][i] += s;
        }
      }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x <= j) {
    const int nw = blockDim.x >> 6;
    double s = 0.0;
    for (int ww = 0; ww < nw; ww++) s += wsum[ww][threadIdx.x];
    gmem(I.partial)[bx_ * (MAXM + 2) + threadIdx.x] = s;
  }
}

// One-block finish of the dots: H[i][j] for i <= j.
template <int JB>
__global__ void __launch_bounds__(FIN_THREADS) k_arnoldi_dots_fin(const Inst* __restrict__ IT, BMap bm, BMap src, int j) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (K->done) return;
  __shared__ double tot[MAXM + 1];
  final_sums<JB>(gmem(I.partial), MAXM + 2, SRC_BLOCKS(src), j + 1, tot);
  __syncthreads();
  if ((int)threadIdx.x <= j) K->H[threadIdx.x * (MAXM + 1) + j] = tot[threadIdx.x];
}

// ---------------------------------------------------------------- Arnoldi step j, part 2:
// w = Av_j - sum_i H[i][j] v_i (FP64 arithmetic, stored as VT), ||w|| of the stored w.
template <int JB, class VT>
__global__ void __launch_bounds__(256) k_arnoldi_update(const Inst* __restrict__ IT, BMap bm, int j, int m, int nc) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (K->done) return;
  const long long n = (long long)nc * I.d.N;
  VT* __restrict__ w = KV<VT>::w(I);
  const VT* __restrict__ Av = KV<VT>::Av(I);
  __shared__ double Hc[JB];
  __shared__ double red[4];
  __shared__ double res[1];
  if ((int)threadIdx.x <= j) Hc[threadIdx.x] = K->H[threadIdx.x * (MAXM + 1) + j];
  __syncthreads();
  double hr[JB];
  const VT* vp[JB];
#pragma unroll
  for (int i = 0; i < JB; i++) { hr[i] = i <= j ? Hc[i] : 0.0; vp[i] = KV<VT>::V(I, i <= j ? i : 0); }
  double acc = 0.0;
  for (long long e = bx_ * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gx_ * blockDim.x) {
    double wv = (double)Av[e];
#pragma unroll
    for (int i = 0; i < JB; i++) if (i <= j) wv -= hr[i] * (double)vp[i][e];
    VT ws = (VT)wv;
    w[e] = ws;
    acc += (double)ws * (double)ws;
  }
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) gmem(I.partial)[bx_] = res[0];   // packed: the finish's loads coalesce
}

// One-block finish of the update: H[j+1][j] = ||w||, Givens rotations, convergence (implicit_operators.cpp:857-897).
// In the mixed form an inner solve that reaches its tolerance ends the sweep but not the solve:
// convergence is decided by the next FP64 residual.
__global__ void __launch_bounds__(FIN_THREADS) k_arnoldi_givens_fin(const Inst* __restrict__ IT, BMap bm, BMap src, int j, int m,
                                                            int mixed) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (K->done) return;
  __shared__ double tot[1];
  // Column j of H, the rotations so far and gam[j] are staged in LDS by parallel loads, so the
  // rotation chain below runs on LDS instead of one global round trip per rotation.
  __shared__ double hc[MAXM + 2], cs[MAXM + 2], sn[MAXM + 2], g0[1];
  const int L = MAXM + 1;
  double* H = K->H;
  const int t = threadIdx.x;
  if (t <= j) hc[t] = H[t * L + j];
  if (t >= 1 && t <= j) { cs[t] = K->c[t]; sn[t] = K->s[t]; }
  if (t == 0) g0[0] = K->gam[j];
  final_sums<1>(gmem(I.partial), 1, SRC_BLOCKS(src), 1, tot);
  __syncthreads();
  if (t == 0) {
    hc[j + 1] = sqrt(tot[0]);
    for (int i = 0; i < j; i++) {
      double t1 = cs[i + 1] * hc[i] + sn[i + 1] * hc[i + 1];
      double t2 = sn[i + 1] * hc[i] - cs[i + 1] * hc[i + 1];
      hc[i] = t1;
      hc[i + 1] = t2;
    }
    double beta = sqrt(hc[j] * hc[j] + hc[j + 1] * hc[j + 1]);
    const double sj = hc[j + 1] / beta, cj = hc[j] / beta;
    K->s[j + 1] = sj;
    K->c[j + 1] = cj;
    hc[j] = beta;
    const double gj1 = sj * g0[0];
    K->gam[j + 1] = gj1;
    K->gam[j] = cj * g0[0];
    K->res = fabs(gj1);
    K->steps = j + 1;
    K->steps_total += 1;
    if (K->outer < MAXSWEEP) K->sweep_steps[K->outer] = j + 1;
    if (gj1 > K->tol_eff) {
      K->hcur = hc[j + 1];
      if (j + 1 == m) { K->done = 1; K->converged = 0; K->jlast = j; }
    } else {
      K->done = 1; K->converged = mixed ? 0 : 1; K->jlast = j;
      if (!mixed) K->outer_conv = 1;       // later restart cycles of a budgeted solve find it solved
    }
  }
  __syncthreads();
  if (t <= j + 1) H[t * L + j] = hc[t];
}

// ---------------------------------------------------------------- x += sum alpha_i v_i (back substitution per block)
template <class VT>
// x_zero: the first sweep of a solve from x = 0, whose x is written (0 + V alpha, or 0 when the
// sweep took no Arnoldi step) rather than read, so the caller need not clear it beforehand.
__global__ void __launch_bounds__(256) k_gmres_xupdate(const Inst* __restrict__ IT, BMap bm, int ix, int nc, int x_zero,
                                                       int skip) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (skipped(I, skip)) return;
  __shared__ double alpha[MAXM + 1];
  __shared__ double Hs[(MAXM + 1) * (MAXM + 1)], gs[MAXM + 1];   // H and gam staged by parallel loads
  const int jl = K->jlast;
  if (jl < 0 && !x_zero) return;
  const int L = MAXM + 1;
  for (int q = threadIdx.x; q < (jl + 1) * (jl + 1); q += blockDim.x) {
    const int i = q / (jl + 1), k = q % (jl + 1);
    Hs[i * L + k] = K->H[i * L + k];
  }
  if ((int)threadIdx.x <= jl) gs[threadIdx.x] = K->gam[threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = jl; i >= 0; i--) {
      double a = gs[i];
      for (int k = i + 1; k <= jl; k++) a -= Hs[i * L + k] * alpha[k];
      alpha[i] = a / Hs[i * L + i];
    }
  }
  __syncthreads();
  double* __restrict__ x = gmem(I.buf[ix]);
  const long long n = (long long)nc * I.d.N;
  for (long long e = bx_ * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gx_ * blockDim.x) {
    double xv = x_zero ? 0.0 : x[e];
    for (int i = 0; i <= jl; i++) xv += alpha[i] * (double)KV<VT>::V(I, i)[e];
    x[e] = xv;
  }
}

// ---------------------------------------------------------------- plain operator application (tests, adjoint)
template <int OPK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(op_waves<OPK, double>())))
k_op_apply(typename OpSel<OPK, double>::Ids ids, Phys P, const Inst* __restrict__ IT, BMap bm, int ix, int iy) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  typedef OpSel<OPK, double> S;
  constexpr int NC = S::NC;
  const typename S::Op op = S::make(I, P, ids);
  const double* __restrict__ x = gmem(I.buf[ix]);
  double* __restrict__ Ax = gmem(I.buf[iy]);
  __shared__ typename S::Smem sm;
  const Dims& d = S::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  S::init(op, sm);
  for (int tile = bx_; tile < ntx * nty; tile += gx_) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    __syncthreads();
    S::load(op, sm, x, 1.0, false, gx0, gy0);
    __syncthreads();
    if constexpr (S::PREP) { S::prep(op, sm); __syncthreads(); }
    double y[NC];
    S::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    if (valid) for (int c = 0; c < NC; c++) Ax[(long long)c * d.N + cidx] = y[c];
  }
}

// ---------------------------------------------------------------- Newton residual (implicit_operators.cpp:690-710)
// g = L p - a p + V_mod(p, rho) - b with L the carrier operator; writes -g and ||g||.
// RK: the Redlich-Kwong instantiation; the ideal-gas one carries no temperature solves (fewer registers)
template <bool RK>
__global__ void __launch_bounds__(256) k_newton_residual(PIds ids, Phys P_, const Inst* __restrict__ IT, BMap bm, int ip,
                                                         int ib, int it, double ntol) {
  BATCH_BLOCK(bm);
  Phys P = P_;
  if constexpr (!RK) P.eos = IMEX_EOS_IDEAL;
  const Inst& I = IT[inst_];
  if (it > 0 && gmem(I.K)->newton_conv) return;     // an earlier iteration of this Newton solve has converged
  const POp op = make_p_op(I, P, ids);
  const double* __restrict__ p = gmem(I.buf[ip]);
  const double* __restrict__ b = gmem(I.buf[ib]);
  double* __restrict__ ng = gmem(I.buf[B_NG]);
  __shared__ double xs[(TY + 2) * (TX + 2)];
  __shared__ double red[4];
  __shared__ double res[1];
  const Dims& d = op.d;
  const double a = 1.0 / (P.gamma - 1.0);
  const double Mbig = op.S[S_MBIG];
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc = 0.0;
  double pmin = 1e300, nfl = 0.0;     // positivity statistics of p (with the floor active)
  int ef = 0;
  for (int tile = bx_; tile < ntx * nty; tile += gx_) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    __syncthreads();
    load_tile<1, double, double>(xs, d, p, 1.0, false, gx0, gy0);
    __syncthreads();
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    if (cx < d.nx && cy < d.ny) {
      int cidx = cx + d.nx * cy;
      double Lp;
      p_apply(op, xs, tx + 1, ty + 1, cx + GC, cy + GC, cidx, &Lp);
      double pf = xs[(ty + 1) * (TX + 2) + tx + 1];
      double rho = op.u[pidx(d, cx + GC, cy + GC)];
      double g = dir_bnd(d, cx + GC, cy + GC) ? Lp - b[cidx] : Lp - a * pf + V_mod(P, pf, rho, Mbig, &ef) - b[cidx];   // :697-699
      ng[cidx] = -g;
      acc += g * g;
      pmin = fmin(pmin, pf);
      if (pf < P.p_floor) nfl += 1;
    }
  }
  if (ef) atomicOr(gmem(I.eflag), 1);
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) gmem(I.partial)[bx_] = res[0];   // packed: the finish's loads coalesce
  if (P.floor_on) {
    // min p and the floored count of this p (exact in any order), for the positivity statistics:
    // the last residual a Newton solve runs is the one of its final p
    __shared__ double sp[8];
    pmin = wave_min(pmin); nfl = wave_sum(nfl);
    if ((threadIdx.x & 63) == 0) { sp[threadIdx.x >> 6] = pmin; sp[4 + (threadIdx.x >> 6)] = nfl; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < (int)(blockDim.x >> 6); w++) { pmin = fmin(pmin, sp[w]); nfl += sp[4 + w]; }
      gmem(I.partial2)[2 * bx_] = pmin; gmem(I.partial2)[2 * bx_ + 1] = nfl;
    }
  }
}

__global__ void __launch_bounds__(FIN_THREADS) k_newton_fin(const Inst* __restrict__ IT, BMap bm, BMap src, int it, double ntol) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  KrylovDev* K = gmem(I.K);
  if (it > 0 && K->newton_conv) return;
  __shared__ double tot[1];
  final_sums<1>(gmem(I.partial), 1, SRC_BLOCKS(src), 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sqrt(tot[0]);
    if (it == 0) { K->newton_res0 = r; K->newton_first = -1; }
    K->newton_res = r;
    K->newton_it = it;
    K->newton_conv = (r < ntol * K->newton_res0 + ntol) ? 1 : 0;
    if (K->newton_conv) K->newton_first = it;   // Jacobian solves this Newton solve needed
  }
}

// Mbig from Cor. 1 (implicit_operators.cpp:670-686): min(b) and, for RK, max rho e(rho, p_floor).
__global__ void __launch_bounds__(256) k_newton_mbig(Phys P, const Inst* __restrict__ IT, BMap bm, int ib, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const Dims& d = I.d;
  const double* __restrict__ b = gmem(I.buf[ib]);
  const double* __restrict__ u = gmem(I.buf[iu]);
  __shared__ double sh[8];
  double bmin = 1e300, vmax = -1e300;
  int ef = 0;
  for (long long e = bx_ * (long long)blockDim.x + threadIdx.x; e < d.N; e += (long long)gx_ * blockDim.x) {
    bmin = fmin(bmin, b[e]);
    if (P.eos != IMEX_EOS_IDEAL) {
      int i = (int)(e % d.nx) + GC, j = (int)(e / d.nx) + GC;
      double v = Vbase(P, P.p_floor, u[pidx(d, i, j)], &ef);
      if (isfinite(v)) vmax = fmax(vmax, v);
    }
  }
  if (ef) atomicOr(gmem(I.eflag), 1);
  bmin = wave_min(bmin); vmax = wave_max(vmax);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) { sh[threadIdx.x >> 6] = bmin; sh[4 + (threadIdx.x >> 6)] = vmax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0], c = sh[4];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) { a = fmin(a, sh[w]); c = fmax(c, sh[4 + w]); }
    gmem(I.partial2)[bx_ * 2] = a; gmem(I.partial2)[bx_ * 2 + 1] = c;
  }
}

// Reduces the (min b, max rho e) pairs in partial2 of the grid src: k_newton_mbig's, or k_p_rhs's
// when that kernel has just formed b.
__global__ void __launch_bounds__(FIN_THREADS) k_newton_mbig_fin(Phys P, const Inst* __restrict__ IT, BMap bm, BMap src) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  double ac[2];
  pair_reduce<false>(gmem(I.partial2), SRC_BLOCKS(src), ac);
  if (threadIdx.x == 0) {
    const double a = ac[0], c = ac[1];
    double Vpe = P.p_floor / (P.gamma - 1.0);
    if (P.eos != IMEX_EOS_IDEAL && c > Vpe) Vpe = c;
    double M_cor = fmax(0.0, Vpe - a) / ((1.0 - P.theta) * P.p_floor);
    double M_base = P.mbig_fac / (P.gamma - 1.0);
    gmem(I.S)[S_MBIG] = fmax(M_base, M_cor);
  }
}

__global__ void k_vslope(Phys P, const Inst* __restrict__ IT, BMap bm, int ipf, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const Dims& d = I.d;
  long long e = bx_ * (long long)blockDim.x + threadIdx.x;
  if (e >= d.N) return;
  int i = (int)(e % d.nx) + GC, j = (int)(e / d.nx) + GC;
  int ef = 0;
  gmem(I.buf[B_VSLOPE])[e] = V_slope(P, gmem(I.buf[ipf])[e], gmem(I.buf[iu])[pidx(d, i, j)], gmem(I.S)[S_MBIG], &ef);
  if (ef) atomicOr(gmem(I.eflag), 1);
}

__global__ void k_axpy1(const Inst* __restrict__ IT, BMap bm, int iy, int ix, int skip) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  if (skipped(I, skip)) return;
  long long e = bx_ * (long long)blockDim.x + threadIdx.x;
  if (e < I.d.N) gmem(I.buf[iy])[e] += gmem(I.buf[ix])[e];
}


// Per-block (min p, floored count) of p for pstats_reduce, when no Newton residual has seen the
// final p (the host-polled Newton loop stopped at its iteration cap after an update).
__global__ void k_pstats(const Inst* __restrict__ IT, BMap bm, int ip, double pfloor) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const double* __restrict__ p = gmem(I.buf[ip]);
  const long long n = I.d.N;
  __shared__ double smin[4], scnt[4];
  double mn = 1e300, cnt = 0;
  for (long long e = bx_ * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gx_ * blockDim.x) {
    double v = p[e];
    mn = fmin(mn, v);
    if (v < pfloor) cnt += 1;
  }
  mn = wave_min(mn); cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) { smin[threadIdx.x >> 6] = mn; scnt[threadIdx.x >> 6] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mn = fmin(mn, smin[w]); cnt += scnt[w]; }
    gmem(I.partial2)[2 * bx_] = smin[0] < mn ? smin[0] : mn;
    gmem(I.partial2)[2 * bx_ + 1] = cnt;
  }
}
__global__ void __launch_bounds__(FIN_THREADS) k_pstats_fin(const Inst* __restrict__ IT, BMap bm, BMap src, double thr) {
  BATCH_BLOCK(bm);
  pstats_reduce(IT[inst_], SRC_BLOCKS(src), thr);
}

// ================================================================ launchers
#define GRID(L, sh) dim3((unsigned)(L).nblocks[sh])
// magnetic-operator kernels and the Newton residual on SH_TILES, pressure-operator kernels on
// SH_TILES_P: their occupancies differ, so their grid caps are set apart
static inline int tiles_shape(int op) { return op == 0 ? SH_TILES : SH_TILES_P; }
static inline int vec_shape(int op) { return op == 0 ? SH_VEC3 : SH_VEC1; }

void gmres_residual(const Launch& L, int op, int prec, const MagIds& mo, const PIds& po, const Phys& P, int ix, int ib,
                    int outer, int x_is_zero, int skip) {
  const int sh = tiles_shape(op);
  const BMap& bm = L.bm[sh];
  if (prec == 0) {
    if (op == 0 && mo.pot) k_gmres_residual<3, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 0);
    else if (op == 0 && mo.refl) k_gmres_residual<4, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 0);
    else if (op == 0) k_gmres_residual<0, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 0);
    else if (po.energy) k_gmres_residual<2, double><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 0);
    else k_gmres_residual<1, double><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 0);
  } else {
    if (op == 0 && mo.pot) k_gmres_residual<3, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 1);
    else if (op == 0 && mo.refl) k_gmres_residual<4, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 1);
    else if (op == 0) k_gmres_residual<0, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 1);
    else if (po.energy) k_gmres_residual<2, float><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 1);
    else k_gmres_residual<1, float><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, ix, ib, outer, x_is_zero, skip, 1);
  }
}
void gmres_residual_fin(const Launch& L, int op, int prec, int outer, int skip, double outer_tol, double inner_tol) {
  k_gmres_residual_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[tiles_shape(op)], outer, skip, prec, outer_tol, inner_tol);
}

void arnoldi_apply(const Launch& L, int op, int prec, const MagIds& mo, const PIds& po, const Phys& P, int j) {
  const int sh = tiles_shape(op);
  const BMap& bm = L.bm[sh];
  if (prec == 0) {
    if (op == 0 && mo.pot) k_arnoldi_apply<3, double, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (op == 0 && mo.refl) k_arnoldi_apply<4, double, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (op == 0) k_arnoldi_apply<0, double, double><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (po.energy) k_arnoldi_apply<2, double, double><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, j);
    else k_arnoldi_apply<1, double, double><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, j);
  } else {
    if (op == 0 && mo.pot) k_arnoldi_apply<3, float, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (op == 0 && mo.refl) k_arnoldi_apply<4, float, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (op == 0) k_arnoldi_apply<0, float, float><<<GRID(L, sh), 256, 0, L.st>>>(mo, P, L.IT, bm, j);
    else if (po.energy) k_arnoldi_apply<2, float, float><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, j);
    else k_arnoldi_apply<1, float, float><<<GRID(L, sh), 256, 0, L.st>>>(po, P, L.IT, bm, j);
  }
}
void arnoldi_apply_fin(const Launch& L, int op, int j) {
  const BMap& one = L.bm[SH_ONE];
  const BMap& src = L.bm[tiles_shape(op)];
  if (j < 8) k_arnoldi_dots_fin<8><<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, one, src, j);
  else if (j < 16) k_arnoldi_dots_fin<16><<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, one, src, j);
  else k_arnoldi_dots_fin<32><<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, one, src, j);
}

template <class VT>
static void launch_update(const Launch& L, int sh, int j, int m, int nc) {
  const BMap& bm = L.bm[sh];
  if (j < 8) k_arnoldi_update<8, VT><<<GRID(L, sh), 256, 0, L.st>>>(L.IT, bm, j, m, nc);
  else if (j < 16) k_arnoldi_update<16, VT><<<GRID(L, sh), 256, 0, L.st>>>(L.IT, bm, j, m, nc);
  else k_arnoldi_update<32, VT><<<GRID(L, sh), 256, 0, L.st>>>(L.IT, bm, j, m, nc);
}
void arnoldi_update(const Launch& L, int op, int prec, int j, int m) {
  const int nc = op == 0 ? 3 : 1;
  if (prec == 0) launch_update<double>(L, vec_shape(op), j, m, nc);
  else launch_update<float>(L, vec_shape(op), j, m, nc);
}
void arnoldi_update_fin(const Launch& L, int op, int prec, int j, int m) {
  k_arnoldi_givens_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[vec_shape(op)], j, m, prec);
}

void gmres_xupdate(const Launch& L, int op, int prec, int ix, int x_zero, int skip) {
  const int nc = op == 0 ? 3 : 1;
  const int sh = vec_shape(op);
  if (prec == 0) k_gmres_xupdate<double><<<GRID(L, sh), 256, 0, L.st>>>(L.IT, L.bm[sh], ix, nc, x_zero, skip);
  else k_gmres_xupdate<float><<<GRID(L, sh), 256, 0, L.st>>>(L.IT, L.bm[sh], ix, nc, x_zero, skip);
}

void op_apply(const Launch& L, int op, const MagIds& mo, const PIds& po, const Phys& P, int ix, int iy) {
  if (op == 0 && mo.pot) k_op_apply<3><<<GRID(L, SH_TILES_ALL), 256, 0, L.st>>>(mo, P, L.IT, L.bm[SH_TILES_ALL], ix, iy);
  else if (op == 0 && mo.refl) k_op_apply<4><<<GRID(L, SH_TILES_ALL), 256, 0, L.st>>>(mo, P, L.IT, L.bm[SH_TILES_ALL], ix, iy);
  else if (op == 0) k_op_apply<0><<<GRID(L, SH_TILES_ALL), 256, 0, L.st>>>(mo, P, L.IT, L.bm[SH_TILES_ALL], ix, iy);
  else if (po.energy) k_op_apply<2><<<GRID(L, SH_TILES_ALL), 256, 0, L.st>>>(po, P, L.IT, L.bm[SH_TILES_ALL], ix, iy);
  else k_op_apply<1><<<GRID(L, SH_TILES_ALL), 256, 0, L.st>>>(po, P, L.IT, L.bm[SH_TILES_ALL], ix, iy);
}

void newton_residual(const Launch& L, const PIds& po, const Phys& P, int ip, int ib, int it, double ntol) {
  if (P.eos != IMEX_EOS_IDEAL) k_newton_residual<true><<<GRID(L, SH_TILES_N), 256, 0, L.st>>>(po, P, L.IT, L.bm[SH_TILES_N], ip, ib, it, ntol);
  else k_newton_residual<false><<<GRID(L, SH_TILES_N), 256, 0, L.st>>>(po, P, L.IT, L.bm[SH_TILES_N], ip, ib, it, ntol);
}
int newton_residual_resident(bool rk) {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const hipError_t e = rk ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_newton_residual<true>, 256, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_newton_residual<false>, 256, 0);
  return e == hipSuccess ? per * ncu : 0;
}
void newton_residual_fin(const Launch& L, int it, double ntol) {
  k_newton_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_TILES_N], it, ntol);
}

void newton_mbig(const Launch& L, const Phys& P, int ib, int iu) {
  k_newton_mbig<<<GRID(L, SH_MBIG), 256, 0, L.st>>>(P, L.IT, L.bm[SH_MBIG], ib, iu);
  k_newton_mbig_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(P, L.IT, L.bm[SH_ONE], L.bm[SH_MBIG]);
}
void newton_mbig_fin_prhs(const Launch& L, const Phys& P) {
  k_newton_mbig_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(P, L.IT, L.bm[SH_ONE], L.bm[SH_INT64x4]);
}

void vslope_launch(const Launch& L, const Phys& P, int ipf, int iu) {
  k_vslope<<<GRID(L, SH_LIN_N), 256, 0, L.st>>>(P, L.IT, L.bm[SH_LIN_N], ipf, iu);
}

void axpy_launch(const Launch& L, int iy, int ix, int skip) {
  k_axpy1<<<GRID(L, SH_LIN_N), 256, 0, L.st>>>(L.IT, L.bm[SH_LIN_N], iy, ix, skip);
}


void pstats_launch(const Launch& L, int ip, double pfloor, double thr) {
  k_pstats<<<GRID(L, SH_PSTATS), 256, 0, L.st>>>(L.IT, L.bm[SH_PSTATS], ip, pfloor);
  k_pstats_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_PSTATS], thr);
}
void pstats_fin_newton(const Launch& L, double thr) {
  k_pstats_fin<<<GRID(L, SH_ONE), FIN_THREADS, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_TILES_N], thr);
}
