// kernels_krylov.hip — matrix-free implicit operators and the device-resident GMRES / Newton pieces.
//
// Operators (Magnetic_Subsystem_AX, Pressure_Subsystem_AX; src/implicit_operators.cpp:141-353,
// 505-616, GLF / MaterialSpeedp branches) are evaluated on 64x4 cell tiles whose input
// vector, with a one-cell halo mapped through M_P / M_B, is staged in LDS.  In Arnoldi
// mode the staged values are the unnormalised vector divided by its norm, so the
// normalisation v = w / ||w|| (implicit_operators.cpp:876) happens once per element
// inside the operator kernel and v is written back in the same pass.
//
// GMRES (implicit_operators.cpp:793-919): classical Gram-Schmidt as in the reference —
// all dots of step j are formed in the operator kernel from the freshly computed
// M^-1 A v_j, then one update pass forms w = Av_j - sum H_ij v_i and ||w||.  Each of
// the two kernels finishes its grid reduction in the last block to arrive, which also
// runs the Givens / Hessenberg bookkeeping in FP64 and raises `done` on convergence or at
// the restart length; later kernels of the same cycle see `done` and exit at once.
#include "kernels.hpp"

#ifndef IMEX_MAG_WAVES
#define IMEX_MAG_WAVES 3
#endif

struct VPtrs { double* v[MAXM + 1]; };

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

template <int NC>
__device__ __forceinline__ void load_tile(double* xs, const Dims& d, const double* __restrict__ x, double h, bool scale,
                                          int gx0, int gy0) {
  const int W = TX + 2, HH = TY + 2;
  auto put = [&](int lx, int ly) {
    int gx = cmap(gx0 + lx - 1, d.nx, d.perx), gy = cmap(gy0 + ly - 1, d.ny, d.pery);
    long long e = gx + (long long)d.nx * gy;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      double v = x[(long long)k * d.N + e];
      xs[k * W * HH + ly * W + lx] = scale ? v / h : v;
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
struct MagSmem {
  double x[3][TY + 2][TX + 2];     // unknown B (scaled), one-cell halo through M_B
  double q[3][TY + 2][TX + 2];     // B of Qe (raw ghost frame)
  double r[TY + 2][TX + 2];        // rho of u (raw ghost frame)
};

__device__ __forceinline__ void mag_load(const MagOp& op, MagSmem& sm, const double* __restrict__ x, double h, bool scale,
                                         int gx0, int gy0) {
  const Dims& d = op.d;
  auto put = [&](int lx, int ly) {
    int cx = gx0 + lx - 1, cy = gy0 + ly - 1;
    int gx = cmap(cx, d.nx, d.perx), gy = cmap(cy, d.ny, d.pery);
    long long e = gx + (long long)d.nx * gy;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double v = x[(long long)f * d.N + e];
      sm.x[f][ly][lx] = scale ? v / h : v;
    }
    int pi = min(cx + GC, d.NX - 1), pj = min(cy + GC, d.NY - 1);
    long long c = pidx(d, pi, pj);
#pragma unroll
    for (int f = 0; f < 3; f++) sm.q[f][ly][lx] = op.Qe[(long long)(5 + f) * d.PS + c];
    sm.r[ly][lx] = op.u[c];
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
__device__ __forceinline__ void mag_corner(const MagSmem& sm, int a, int b, double* cn) {
  double n[3], bb[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    n[c] = 0.25 * (sm.q[c][b][a] + sm.q[c][b][a + 1] + sm.q[c][b + 1][a] + sm.q[c][b + 1][a + 1]);
    bb[c] = 0.25 * (sm.x[c][b][a] + sm.x[c][b][a + 1] + sm.x[c][b + 1][a] + sm.x[c][b + 1][a + 1]);
  }
  double half = 0.5 * (n[0] * bb[0] + n[1] * bb[1] + n[2] * bb[2]);
  cn[0] = n[0] * bb[1];
  cn[1] = n[0] * bb[2];
  cn[2] = n[1] * bb[0];
  cn[3] = n[1] * bb[2];
  cn[4] = half - n[0] * bb[0];
  cn[5] = half - n[1] * bb[1];
}
// Per-cell products of LDS cell (a, b): PM - q5 x5, PM - q6 x6, q6 x5, q6 x7, q5 x6, q5 x7
// with PM = (q . x)/2.
__device__ __forceinline__ void mag_dv(const MagSmem& sm, int a, int b, double* dv) {
  double q5 = sm.q[0][b][a], q6 = sm.q[1][b][a], q7 = sm.q[2][b][a];
  double x5 = sm.x[0][b][a], x6 = sm.x[1][b][a], x7 = sm.x[2][b][a];
  double q5x5 = q5 * x5, q6x6 = q6 * x6;
  double PM = 0.5 * (q5x5 + q6x6 + q7 * x7);
  dv[0] = PM - q5x5;
  dv[1] = PM - q6x6;
  dv[2] = q6 * x5;
  dv[3] = q6 * x7;
  dv[4] = q5 * x6;
  dv[5] = q5 * x7;
}
// Fluxes (gx, gz) through the y-face between LDS rows yb and yb+1 at column a; cL, cR are the
// corners (a-1, yb), (a, yb) and D, U the cell products of (a, yb), (a, yb+1).
__device__ __forceinline__ void mag_yface(const MagSmem& sm, int a, int yb, const double* cL, const double* cR,
                                          const double* D, const double* U, double idx, double idy, double midx,
                                          double& gx, double& gz) {
  double rf = 0.5 * (sm.r[yb][a] + sm.r[yb + 1][a]);
  double nf0 = 0.5 * (sm.q[0][yb][a] + sm.q[0][yb + 1][a]);
  double nf1 = 0.5 * (sm.q[1][yb][a] + sm.q[1][yb + 1][a]);
  double nf2 = 0.5 * (sm.q[2][yb][a] + sm.q[2][yb + 1][a]);
  double T1 = idx * (cR[0] - cL[0]) - idy * (U[1] - D[1]);
  gx = nf0 / rf * T1 + nf1 / rf * (idx * (cR[4] - cL[4]) - idy * (U[2] - D[2]));
  gz = nf2 / rf * T1 + nf1 / rf * (midx * (cR[1] - cL[1]) - idy * (U[3] - D[3]));
}
// Fluxes (fy, fz) through the x-face between LDS columns xa and xa+1 at row b; cB, cT are the
// corners (xa, b-1), (xa, b) and L, R the cell products of (xa, b), (xa+1, b).
__device__ __forceinline__ void mag_xface(const MagSmem& sm, int xa, int b, const double* cB, const double* cT,
                                          const double* L, const double* R, double idy, double midx,
                                          double& fy, double& fz) {
  double rf = 0.5 * (sm.r[b][xa] + sm.r[b][xa + 1]);
  double nf0 = 0.5 * (sm.q[0][b][xa] + sm.q[0][b][xa + 1]);
  double nf1 = 0.5 * (sm.q[1][b][xa] + sm.q[1][b][xa + 1]);
  double nf2 = 0.5 * (sm.q[2][b][xa] + sm.q[2][b][xa + 1]);
  double dpm = R[0] - L[0];
  double T2 = midx * dpm + idy * (cT[0] - cB[0]);
  fy = nf1 / rf * T2 + nf0 / rf * (midx * (R[4] - L[4]) + idy * (cT[5] - cB[5]));
  fz = nf2 / rf * (midx * dpm + idy * (cT[2] - cB[2])) + nf0 / rf * (midx * (R[5] - L[5]) - idy * (cT[3] - cB[3]));
}

// A x for the cell of a valid thread (no block synchronisation: each thread forms the four face
// fluxes of its own cell from the staged tiles; a face shared by two cells is formed by both
// with the same arithmetic).
__device__ __forceinline__ void mag_compute(const MagOp& op, const MagSmem& sm, int tx, int ty, bool valid, double* out) {
  if (!valid) return;
  const Dims& d = op.d;
  const double h = op.S[S_ADT], dx = d.dx, dy = d.dy;
  const double idx = 1 / dx, idy = 1 / dy, midx = -1 / dx;
  const int a = tx + 1, b = ty + 1;
  double c00[6], c10[6], c01[6], c11[6];      // corners (a-1, b-1), (a, b-1), (a-1, b), (a, b)
  mag_corner(sm, a - 1, b - 1, c00);
  mag_corner(sm, a, b - 1, c10);
  mag_corner(sm, a - 1, b, c01);
  mag_corner(sm, a, b, c11);
  double dC[6], dN[6];
  mag_dv(sm, a, b, dC);
  double gxu, gzu, gxl, gzl, fyr, fzr, fyl, fzl;
  mag_dv(sm, a, b + 1, dN);
  mag_yface(sm, a, b, c01, c11, dC, dN, idx, idy, midx, gxu, gzu);
  mag_dv(sm, a, b - 1, dN);
  mag_yface(sm, a, b - 1, c00, c10, dN, dC, idx, idy, midx, gxl, gzl);
  mag_dv(sm, a + 1, b, dN);
  mag_xface(sm, a, b, c10, c11, dC, dN, idy, midx, fyr, fzr);
  mag_dv(sm, a - 1, b, dN);
  mag_xface(sm, a - 1, b, c00, c01, dN, dC, idy, midx, fyl, fzl);
  auto X = [&](int c, int di, int dj) { return sm.x[c][b + dj][a + di]; };
  double ax5 = X(0, 0, 0) + (h * h / dy) * (gxu - gxl);
  double ax6 = X(1, 0, 0) + (h * h / dx) * (fyr - fyl);
  double ax7 = X(2, 0, 0) + h * h / dx * (fzr - fzl) + h * h / dy * (gzu - gzl);
  const double nu_x = op.S[S_BMAXX] * dx, nu_y = op.S[S_BMAXY] * dy;
  double lap[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    lap[c] = nu_x / (dx * dx) * (X(c, -1, 0) - 2.0 * X(c, 0, 0) + X(c, 1, 0)) + nu_y / (dy * dy) * (X(c, 0, -1) - 2.0 * X(c, 0, 0) + X(c, 0, 1));
  out[0] = ax5 - h * lap[0];
  out[1] = ax6 - h * lap[1];
  out[2] = ax7 - h * lap[2];
}

// ---------------------------------------------------------------- pressure operator
__device__ __forceinline__ double p_vslope(const POp& op, int cidx, double rho) {
  if (op.vslope) return op.vslope[cidx];
  const Phys& P = op.P;
  double p = op.pfreeze[cidx];
  if (!P.floor_on) return 1.0 / (P.gamma - 1.0);
  return (p >= P.p_floor) ? 1.0 / (P.gamma - 1.0) : op.S[S_MBIG];
}
__device__ __forceinline__ void p_apply(const POp& op, const double* xs, int lx, int ly, int i, int j, int cidx, double* out) {
  const Dims& d = op.d;
  const int W = TX + 2;
  const double h = op.S[S_ADT], dx = d.dx, dy = d.dy;
  auto X = [&](int di, int dj) { return xs[(ly + dj) * W + lx + di]; };
  auto H = [&](int di, int dj) { return op.hhat[pidx(d, i + di, j + dj)]; };
  auto E = [&](int di, int dj) { return op.eta[pidx(d, i + di, j + dj)]; };
  double h00 = H(0, 0), hp0 = H(1, 0), hm0 = H(-1, 0), h0p = H(0, 1), h0m = H(0, -1);
  double pii = X(0, 0), xp0 = X(1, 0), xm0 = X(-1, 0), x0p = X(0, 1), x0m = X(0, -1);
  double rho = op.u[pidx(d, i, j)];
  double Vdiag = op.jacobian ? p_vslope(op, cidx, rho) : (1.0 / (op.P.gamma - 1.0));
  double ax = Vdiag * pii - h * h * (1 / (dx * dx) * (0.5 * (hp0 + h00) * (xp0 - pii) - 0.5 * (h00 + hm0) * (pii - xm0))
                                   + 1 / (dy * dy) * (0.5 * (h0p + h00) * (x0p - pii) - 0.5 * (h0m + h00) * (pii - x0m)));
  double e00 = E(0, 0);
  double nu_x = op.S[S_VXMAX] * dx, nu_y = op.S[S_VYMAX] * dy;
  double lap = nu_x / (dx * dx) * (0.5 * (E(1, 0) + e00) * (xp0 - pii) - 0.5 * (e00 + E(-1, 0)) * (pii - xm0))
             + nu_y / (dy * dy) * (0.5 * (E(0, 1) + e00) * (x0p - pii) - 0.5 * (e00 + E(0, -1)) * (pii - x0m));
  double visc = op.visc ? op.visc[pidx(d, i, j)] : 1.0 / (op.P.gamma - 1.0);
  out[0] = ax - visc * h * lap;
}
__device__ __forceinline__ void p_precond(const POp& op, int cidx, int i, int j, double* y) {
  if (!op.precond) return;
  double rho = op.u[pidx(op.d, i, j)];
  double Vd = op.jacobian ? p_vslope(op, cidx, rho) : 1.0 / (op.P.gamma - 1.0);
  y[0] = y[0] / (Vd + op.Tdiag[cidx]);
}

// Tile interface shared by the GMRES kernels: load the input vector tile (plus whatever the
// operator stages), then compute M^-1 A x for the block's cells.
struct PSmem { double x[(TY + 2) * (TX + 2)]; };
template <int OPK> struct OpSel;
template <> struct OpSel<0> {
  static constexpr int NC = 3;
  typedef MagOp T;
  typedef MagSmem Smem;
  __device__ static void load(const MagOp& op, Smem& sm, const double* x, double h, bool scale, int gx0, int gy0) {
    mag_load(op, sm, x, h, scale, gx0, gy0);
  }
  __device__ static void compute(const MagOp& op, Smem& sm, int tx, int ty, int, int, int, bool valid, double* y) {
    mag_compute(op, sm, tx, ty, valid, y);
  }
  __device__ static double center(const Smem& sm, int c, int tx, int ty) { return sm.x[c][ty + 1][tx + 1]; }
  __device__ static void precond(const MagOp&, int, int, int, double*) {}
  __device__ static const Dims& dims(const MagOp& op) { return op.d; }
};
template <> struct OpSel<1> {
  static constexpr int NC = 1;
  typedef POp T;
  typedef PSmem Smem;
  __device__ static void load(const POp& op, Smem& sm, const double* x, double h, bool scale, int gx0, int gy0) {
    load_tile<1>(sm.x, op.d, x, h, scale, gx0, gy0);
  }
  __device__ static void compute(const POp& op, Smem& sm, int tx, int ty, int i, int j, int cidx, bool valid, double* y) {
    if (valid) p_apply(op, sm.x, tx + 1, ty + 1, i, j, cidx, y);
  }
  __device__ static double center(const Smem& sm, int, int tx, int ty) { return sm.x[(ty + 1) * (TX + 2) + tx + 1]; }
  __device__ static void precond(const POp& op, int cidx, int i, int j, double* y) { p_precond(op, cidx, i, j, y); }
  __device__ static const Dims& dims(const POp& op) { return op.d; }
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
// One block: out[v] = sum over blocks b < nblk of partial[b*stride + v], v < nv <= NV.  Thread t
// sums rows t, t + 256, ... with all their values in flight at once; the per-thread sums are
// then reduced wave by wave in a fixed order, so the result does not depend on timing.
template <int NV>
__device__ __forceinline__ void final_sums(const double* partial, int stride, int nblk, int nv, double* out) {
  __shared__ double red[4][NV];
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
}

// ---------------------------------------------------------------- GMRES residual r0 = M^-1 (b - A x), gam0 = ||r0||
template <int OPK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OPK == 0 ? IMEX_MAG_WAVES : 4))) k_gmres_residual(typename OpSel<OPK>::T op, KrylovBufs kb, const double* __restrict__ x,
                                                        const double* __restrict__ b, KrylovDev* K, int outer, int x_is_zero,
                                                        const int* skip) {
  if (skip && *skip) {           // the caller's outer iteration has already converged: no solve
    if (blockIdx.x == 0 && threadIdx.x == 0) { K->done = 1; K->converged = 1; K->jlast = -1; K->steps = 0; K->outer = outer; }
    return;
  }
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ typename OpSel<OPK>::Smem sm;
  __shared__ double red[4];
  __shared__ double res[1];
  const Dims& d = OpSel<OPK>::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc = 0.0;
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    double y[NC];
    if (!x_is_zero) {
      __syncthreads();
      OpSel<OPK>::load(op, sm, x, 1.0, false, gx0, gy0);
      __syncthreads();
      OpSel<OPK>::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    }
    if (valid) {
      if (x_is_zero) for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx];
      else for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx] - y[c];
      OpSel<OPK>::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) { kb.w[(long long)c * d.N + cidx] = y[c]; acc += y[c] * y[c]; }
    }
  }
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) kb.partial[blockIdx.x * (MAXM + 2)] = res[0];
}

// One-block finish of the residual norm: gam0, tolerance, convergence (implicit_operators.cpp:829-836).
__global__ void __launch_bounds__(256) k_gmres_residual_fin(const double* partial, int nblk, KrylovDev* K, int outer,
                                                            const int* skip) {
  if (skip && *skip) return;
  __shared__ double tot[1];
  final_sums<1>(partial, MAXM + 2, nblk, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    double g0 = sqrt(tot[0]);
    if (outer == 0) { K->tol_eff = K->tol * g0 + K->tol; K->res0 = g0; }
    K->res = g0;
    K->gam[0] = g0;
    K->outer = outer;
    K->steps = 0;
    if (g0 <= K->tol_eff) { K->done = 1; K->converged = 1; K->jlast = -1; }
    else { K->done = 0; K->converged = 0; K->jlast = -1; K->hcur = g0; }
  }
}

// ---------------------------------------------------------------- Arnoldi step j, part 1:
// v_j = w / h (written), Av_j = M^-1 A v_j (written), H[i][j] = v_i . Av_j for i <= j.
// The dots of each tile are reduced over the wave right away (a uniform loop over i, so no
// accumulator array is live across the operator) and summed per wave in LDS, in tile order.
template <int OPK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OPK == 0 ? IMEX_MAG_WAVES : 4)))
k_arnoldi_apply(typename OpSel<OPK>::T op, KrylovBufs kb, VPtrs V, int j, KrylovDev* K) {
  if (K->done) return;
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ typename OpSel<OPK>::Smem sm;
  __shared__ double wsum[4][MAXM + 1];
  const Dims& d = OpSel<OPK>::dims(op);
  const double h = K->hcur;
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = lane; i <= j; i += 64) wsum[wid][i] = 0.0;
  double* vj = V.v[j];
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    __syncthreads();
    OpSel<OPK>::load(op, sm, kb.w, h, true, gx0, gy0);
    __syncthreads();
    double y[NC], vc[NC];
    OpSel<OPK>::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    if (valid) {
      OpSel<OPK>::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) {
        long long e = (long long)c * d.N + cidx;
        vc[c] = OpSel<OPK>::center(sm, c, tx, ty);
        vj[e] = vc[c];
        kb.Av[e] = y[c];
      }
    }
    for (int i = 0; i <= j; i++) {
      double part = 0.0;
      if (valid) {
        const double* vi = V.v[i];
        for (int c = 0; c < NC; c++) part += (i < j ? vi[(long long)c * d.N + cidx] : vc[c]) * y[c];
      }
      double s = wave_sum(part);
      if (lane == 0) wsum[wid][i] += s;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x <= j) {
    const int nw = blockDim.x >> 6;
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += wsum[w][threadIdx.x];
    kb.partial[blockIdx.x * (MAXM + 2) + threadIdx.x] = s;
  }
}

// One-block finish of the dots: H[i][j] for i <= j.
template <int JB>
__global__ void __launch_bounds__(256) k_arnoldi_dots_fin(const double* partial, int nblk, int j, KrylovDev* K) {
  if (K->done) return;
  __shared__ double tot[MAXM + 1];
  final_sums<JB>(partial, MAXM + 2, nblk, j + 1, tot);
  __syncthreads();
  if ((int)threadIdx.x <= j) K->H[threadIdx.x * (MAXM + 1) + j] = tot[threadIdx.x];
}

// ---------------------------------------------------------------- Arnoldi step j, part 2:
// w = Av_j - sum_i H[i][j] v_i, H[j+1][j] = ||w||, Givens rotations, convergence.
template <int JB>
__global__ void __launch_bounds__(256) k_arnoldi_update(KrylovBufs kb, VPtrs V, int j, int m, long long n, KrylovDev* K) {
  if (K->done) return;
  __shared__ double Hc[JB];
  __shared__ double red[4];
  __shared__ double res[1];
  if ((int)threadIdx.x <= j) Hc[threadIdx.x] = K->H[threadIdx.x * (MAXM + 1) + j];
  __syncthreads();
  double hr[JB];
#pragma unroll
  for (int i = 0; i < JB; i++) hr[i] = i <= j ? Hc[i] : 0.0;
  double acc = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double w = kb.Av[e];
#pragma unroll
    for (int i = 0; i < JB; i++) if (i <= j) w -= hr[i] * V.v[i][e];
    kb.w[e] = w;
    acc += w * w;
  }
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) kb.partial[blockIdx.x * (MAXM + 2)] = res[0];
}

// One-block finish of the update: H[j+1][j] = ||w||, Givens rotations, convergence (implicit_operators.cpp:857-897).
__global__ void __launch_bounds__(256) k_arnoldi_givens_fin(const double* partial, int nblk, int j, int m, KrylovDev* K) {
  if (K->done) return;
  __shared__ double tot[1];
  final_sums<1>(partial, MAXM + 2, nblk, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int L = MAXM + 1;
    double* H = K->H;
    H[(j + 1) * L + j] = sqrt(tot[0]);
    for (int i = 0; i < j; i++) {
      double t1 = K->c[i + 1] * H[i * L + j] + K->s[i + 1] * H[(i + 1) * L + j];
      double t2 = K->s[i + 1] * H[i * L + j] - K->c[i + 1] * H[(i + 1) * L + j];
      H[i * L + j] = t1;
      H[(i + 1) * L + j] = t2;
    }
    double beta = sqrt(H[j * L + j] * H[j * L + j] + H[(j + 1) * L + j] * H[(j + 1) * L + j]);
    K->s[j + 1] = H[(j + 1) * L + j] / beta;
    K->c[j + 1] = H[j * L + j] / beta;
    H[j * L + j] = beta;
    K->gam[j + 1] = K->s[j + 1] * K->gam[j];
    K->gam[j] = K->c[j + 1] * K->gam[j];
    K->res = fabs(K->gam[j + 1]);
    K->steps = j + 1;
    if (K->gam[j + 1] > K->tol_eff) {
      K->hcur = H[(j + 1) * L + j];
      if (j + 1 == m) { K->done = 1; K->converged = 0; K->jlast = j; }
    } else {
      K->done = 1; K->converged = 1; K->jlast = j;
    }
  }
}

// ---------------------------------------------------------------- x += sum alpha_i v_i (back substitution per block)
__global__ void __launch_bounds__(256) k_gmres_xupdate(VPtrs V, double* __restrict__ x, long long n, KrylovDev* K) {
  __shared__ double alpha[MAXM + 1];
  const int jl = K->jlast;
  if (jl < 0) return;
  if (threadIdx.x == 0) {
    const int L = MAXM + 1;
    for (int i = jl; i >= 0; i--) {
      double a = K->gam[i];
      for (int k = i + 1; k <= jl; k++) a -= K->H[i * L + k] * alpha[k];
      alpha[i] = a / K->H[i * L + i];
    }
  }
  __syncthreads();
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double xv = x[e];
#pragma unroll
    for (int i = 0; i <= MAXM; i++) if (i <= jl) xv += alpha[i] * V.v[i][e];
    x[e] = xv;
  }
}

// ---------------------------------------------------------------- plain operator application (tests, adjoint)
template <int OPK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OPK == 0 ? IMEX_MAG_WAVES : 4))) k_op_apply(typename OpSel<OPK>::T op, const double* __restrict__ x, double* __restrict__ Ax) {
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ typename OpSel<OPK>::Smem sm;
  const Dims& d = OpSel<OPK>::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    bool valid = cx < d.nx && cy < d.ny;
    int cidx = cx + d.nx * cy;
    __syncthreads();
    OpSel<OPK>::load(op, sm, x, 1.0, false, gx0, gy0);
    __syncthreads();
    double y[NC];
    OpSel<OPK>::compute(op, sm, tx, ty, cx + GC, cy + GC, cidx, valid, y);
    if (valid) for (int c = 0; c < NC; c++) Ax[(long long)c * d.N + cidx] = y[c];
  }
}

// ---------------------------------------------------------------- Newton residual (implicit_operators.cpp:690-710)
// g = L p - a p + V_mod(p, rho) - b with L the carrier operator; writes -g and ||g||.
__global__ void __launch_bounds__(256) k_newton_residual(POp op, KrylovBufs kb, const double* __restrict__ p,
                                                         const double* __restrict__ b, double* __restrict__ ng,
                                                         KrylovDev* K, int it, double ntol, int* eflag, const int* skip) {
  if (skip && *skip) return;     // an earlier iteration of this Newton solve has converged
  __shared__ double xs[(TY + 2) * (TX + 2)];
  __shared__ double red[4];
  __shared__ double res[1];
  const Dims& d = op.d;
  const Phys& P = op.P;
  const double a = 1.0 / (P.gamma - 1.0);
  const double Mbig = op.S[S_MBIG];
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc = 0.0;
  int ef = 0;
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    __syncthreads();
    load_tile<1>(xs, d, p, 1.0, false, gx0, gy0);
    __syncthreads();
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    if (cx < d.nx && cy < d.ny) {
      int cidx = cx + d.nx * cy;
      double Lp;
      p_apply(op, xs, tx + 1, ty + 1, cx + GC, cy + GC, cidx, &Lp);
      double pf = xs[(ty + 1) * (TX + 2) + tx + 1];
      double rho = op.u[pidx(d, cx + GC, cy + GC)];
      double g = Lp - a * pf + V_mod(P, pf, rho, Mbig, &ef) - b[cidx];
      ng[cidx] = -g;
      acc += g * g;
    }
  }
  if (ef) atomicOr(eflag, 1);
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) kb.partial[blockIdx.x * (MAXM + 2)] = res[0];
}

__global__ void __launch_bounds__(256) k_newton_fin(const double* partial, int nblk, KrylovDev* K, int it, double ntol) {
  if (it > 0 && K->newton_conv) return;
  __shared__ double tot[1];
  final_sums<1>(partial, MAXM + 2, nblk, 1, tot);
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
__global__ void __launch_bounds__(256) k_newton_mbig(Dims d, Phys P, const double* __restrict__ b, const double* __restrict__ u,
                                                     double* S, double* partial, KrylovDev* K, int* eflag) {
  __shared__ double sh[8];
  double bmin = 1e300, vmax = -1e300;
  int ef = 0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < d.N; e += (long long)gridDim.x * blockDim.x) {
    bmin = fmin(bmin, b[e]);
    if (P.eos != IMEX_EOS_IDEAL) {
      int i = (int)(e % d.nx) + GC, j = (int)(e / d.nx) + GC;
      double v = Vbase(P, P.p_floor, u[pidx(d, i, j)], &ef);
      if (isfinite(v)) vmax = fmax(vmax, v);
    }
  }
  if (ef) atomicOr(eflag, 1);
  bmin = wave_min(bmin); vmax = wave_max(vmax);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) { sh[threadIdx.x >> 6] = bmin; sh[4 + (threadIdx.x >> 6)] = vmax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0], c = sh[4];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) { a = fmin(a, sh[w]); c = fmax(c, sh[4 + w]); }
    partial[blockIdx.x * 2] = a; partial[blockIdx.x * 2 + 1] = c;
  }
}

__global__ void k_newton_mbig_fin(Phys P, const double* partial, int nblk, double* S) {
  // min and max are order independent: one wave reduces the block partials
  double a = 1e300, c = -1e300;
  for (int bl = threadIdx.x; bl < nblk; bl += 64) { a = fmin(a, partial[bl * 2]); c = fmax(c, partial[bl * 2 + 1]); }
  a = wave_min(a); c = wave_max(c);
  if (threadIdx.x == 0) {
    double Vpe = P.p_floor / (P.gamma - 1.0);
    if (P.eos != IMEX_EOS_IDEAL && c > Vpe) Vpe = c;
    double M_cor = fmax(0.0, Vpe - a) / ((1.0 - P.theta) * P.p_floor);
    double M_base = P.mbig_fac / (P.gamma - 1.0);
    S[S_MBIG] = fmax(M_base, M_cor);
  }
}

__global__ void k_vslope(Dims d, Phys P, const double* __restrict__ pf, const double* __restrict__ u, const double* S,
                         double* __restrict__ vs, int* eflag) {
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= d.N) return;
  int i = (int)(e % d.nx) + GC, j = (int)(e / d.nx) + GC;
  int ef = 0;
  vs[e] = V_slope(P, pf[e], u[pidx(d, i, j)], S[S_MBIG], &ef);
  if (ef) atomicOr(eflag, 1);
}

__global__ void k_axpy1(long long n, double* __restrict__ y, const double* __restrict__ x, const int* skip) {
  if (skip && *skip) return;
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e < n) y[e] += x[e];
}

// ================================================================ launchers
static inline int grid_for_tiles(const Dims& d, int cap) {
  int nt = ((d.nx + TX - 1) / TX) * ((d.ny + TY - 1) / TY);
  return nt < cap ? nt : cap;
}
static VPtrs make_vptrs(const KrylovBufs&, double* const* hostV) {
  VPtrs V;
  for (int i = 0; i <= MAXM; i++) V.v[i] = hostV[i];
  return V;
}

void gmres_residual(int op, const MagOp& mo, const POp& po, const KrylovBufs& kb, const double* x, const double* b,
                    KrylovDev* K, int outer, int x_is_zero, const int* skip, hipStream_t st) {
  if (op == 0) {
    int g = grid_for_tiles(mo.d, kb.nblk_apply);
    k_gmres_residual<0><<<g, 256, 0, st>>>(mo, kb, x, b, K, outer, x_is_zero, skip);
  } else {
    int g = grid_for_tiles(po.d, kb.nblk_apply);
    k_gmres_residual<1><<<g, 256, 0, st>>>(po, kb, x, b, K, outer, x_is_zero, skip);
  }
}
void gmres_residual_fin(const Dims& d, const KrylovBufs& kb, KrylovDev* K, int outer, const int* skip, hipStream_t st) {
  k_gmres_residual_fin<<<1, 256, 0, st>>>(kb.partial, grid_for_tiles(d, kb.nblk_apply), K, outer, skip);
}

void arnoldi_apply(int op, const MagOp& mo, const POp& po, const KrylovBufs& kb, double* const* hostV, int j,
                   KrylovDev* K, hipStream_t st) {
  VPtrs V = make_vptrs(kb, hostV);
  const Dims& d = op == 0 ? mo.d : po.d;
  int g = grid_for_tiles(d, kb.nblk_apply);
  if (op == 0) k_arnoldi_apply<0><<<g, 256, 0, st>>>(mo, kb, V, j, K);
  else k_arnoldi_apply<1><<<g, 256, 0, st>>>(po, kb, V, j, K);
}
void arnoldi_apply_fin(const Dims& d, const KrylovBufs& kb, int j, KrylovDev* K, hipStream_t st) {
  int g = grid_for_tiles(d, kb.nblk_apply);
  if (j < 8) k_arnoldi_dots_fin<8><<<1, 256, 0, st>>>(kb.partial, g, j, K);
  else if (j < 16) k_arnoldi_dots_fin<16><<<1, 256, 0, st>>>(kb.partial, g, j, K);
  else k_arnoldi_dots_fin<32><<<1, 256, 0, st>>>(kb.partial, g, j, K);
}

void arnoldi_update(int nc, const Dims& d, const KrylovBufs& kb, double* const* hostV, int j, int m, KrylovDev* K,
                    hipStream_t st) {
  VPtrs V = make_vptrs(kb, hostV);
  long long n = (long long)nc * d.N;
  long long nb = (n + 255) / 256;
  int g = (int)(nb < kb.nblk_vec ? nb : kb.nblk_vec);
  if (j < 8) k_arnoldi_update<8><<<g, 256, 0, st>>>(kb, V, j, m, n, K);
  else if (j < 16) k_arnoldi_update<16><<<g, 256, 0, st>>>(kb, V, j, m, n, K);
  else k_arnoldi_update<32><<<g, 256, 0, st>>>(kb, V, j, m, n, K);
}
void arnoldi_update_fin(int nc, const Dims& d, const KrylovBufs& kb, int j, int m, KrylovDev* K, hipStream_t st) {
  long long n = (long long)nc * d.N;
  long long nb = (n + 255) / 256;
  int g = (int)(nb < kb.nblk_vec ? nb : kb.nblk_vec);
  k_arnoldi_givens_fin<<<1, 256, 0, st>>>(kb.partial, g, j, m, K);
}

void gmres_xupdate(int nc, const Dims& d, const KrylovBufs& kb, double* const* hostV, double* x, KrylovDev* K, hipStream_t st) {
  VPtrs V = make_vptrs(kb, hostV);
  long long n = (long long)nc * d.N;
  long long nb = (n + 255) / 256;
  int g = (int)(nb < kb.nblk_vec ? nb : kb.nblk_vec);
  k_gmres_xupdate<<<g, 256, 0, st>>>(V, x, n, K);
}

void op_apply(int op, const MagOp& mo, const POp& po, const double* x, double* Ax, hipStream_t st) {
  if (op == 0) k_op_apply<0><<<grid_for_tiles(mo.d, 4096), 256, 0, st>>>(mo, x, Ax);
  else k_op_apply<1><<<grid_for_tiles(po.d, 4096), 256, 0, st>>>(po, x, Ax);
}

void newton_residual(const POp& po, const KrylovBufs& kb, const double* p, const double* b, double* ng, KrylovDev* K,
                     int it, double ntol, int* eflag, hipStream_t st) {
  int g = grid_for_tiles(po.d, kb.nblk_apply);
  k_newton_residual<<<g, 256, 0, st>>>(po, kb, p, b, ng, K, it, ntol, eflag, it > 0 ? &K->newton_conv : nullptr);
}
void newton_residual_fin(const Dims& d, const KrylovBufs& kb, KrylovDev* K, int it, double ntol, hipStream_t st) {
  k_newton_fin<<<1, 256, 0, st>>>(kb.partial, grid_for_tiles(d, kb.nblk_apply), K, it, ntol);
}

void newton_mbig(const Dims& d, const Phys& P, const double* b, const double* u, double* S, double* partial, KrylovDev* K,
                 int* eflag, hipStream_t st) {
  long long nb = (d.N + 255) / 256;
  int g = (int)(nb < 512 ? nb : 512);
  k_newton_mbig<<<g, 256, 0, st>>>(d, P, b, u, S, partial, K, eflag);
  k_newton_mbig_fin<<<1, 64, 0, st>>>(P, partial, g, S);
}

void vslope_launch(const Dims& d, const Phys& P, const double* pf, const double* u, const double* S, double* vs, int* eflag,
                   hipStream_t st) {
  k_vslope<<<(d.N + 255) / 256, 256, 0, st>>>(d, P, pf, u, S, vs, eflag);
}

void axpy_launch(long long n, double* y, const double* x, const int* skip, hipStream_t st) {
  k_axpy1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, y, x, skip);
}
