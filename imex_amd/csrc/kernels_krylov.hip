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

struct VPtrs { double* v[MAXM + 1]; };

// compact-coordinate boundary map for one-cell halos (M_P with i in [g-1, n+g])
__device__ __forceinline__ int cmap(int g, int n, int per) {
  if (g < 0) return per ? g + n : 0;
  if (g >= n) return per ? (g - n < n ? g - n : n - 1) : n - 1;
  return g;
}

template <int NC>
__device__ __forceinline__ void load_tile(double* xs, const Dims& d, const double* __restrict__ x, double h, bool scale,
                                          int gx0, int gy0) {
  const int W = TX + 2, HH = TY + 2;
  for (int t = threadIdx.x; t < NC * W * HH; t += blockDim.x) {
    int k = t / (W * HH), r = t % (W * HH);
    int ly = r / W, lx = r % W;
    int gx = cmap(gx0 + lx - 1, d.nx, d.perx), gy = cmap(gy0 + ly - 1, d.ny, d.pery);
    double v = x[(long long)k * d.N + gx + (long long)d.nx * gy];
    xs[t] = scale ? v / h : v;
  }
}

// ---------------------------------------------------------------- magnetic operator
__device__ __forceinline__ void mag_apply(const MagOp& op, const double* xs, int lx, int ly, int i, int j, double* out) {
  const Dims& d = op.d;
  const int W = TX + 2, HH = TY + 2;
  const double h = op.S[S_ADT], dx = d.dx, dy = d.dy;
  auto X = [&](int di, int dj, int k) { return xs[((k - 5) * HH + ly + dj) * W + lx + di]; };
  auto Q = [&](int di, int dj, int k) { return op.Qe[k * d.PS + pidx(d, i + di, j + dj)]; };
  auto R = [&](int di, int dj) { return op.u[pidx(d, i + di, j + dj)]; };
  double rho_p0 = 0.5 * (R(0, 0) + R(1, 0)), rho_m0 = 0.5 * (R(0, 0) + R(-1, 0));
  double rho_0p = 0.5 * (R(0, 0) + R(0, 1)), rho_0m = 0.5 * (R(0, 0) + R(0, -1));
  double n_p0[3], n_m0[3], n_0p[3], n_0m[3], n_pp[3], n_pm[3], n_mp[3], n_mm[3];
  double b_pp[3], b_pm[3], b_mp[3], b_mm[3];
  for (int c = 0; c < 3; c++) {
    int k = 5 + c;
    double q00 = Q(0, 0, k), qp0 = Q(1, 0, k), qm0 = Q(-1, 0, k), q0p = Q(0, 1, k), q0m = Q(0, -1, k);
    n_p0[c] = 0.5 * (q00 + qp0); n_m0[c] = 0.5 * (q00 + qm0);
    n_0p[c] = 0.5 * (q00 + q0p); n_0m[c] = 0.5 * (q00 + q0m);
    n_pp[c] = 0.25 * (q00 + qp0 + q0p + Q(1, 1, k));
    n_pm[c] = 0.25 * (q00 + qp0 + q0m + Q(1, -1, k));
    n_mp[c] = 0.25 * (q00 + qm0 + q0p + Q(-1, 1, k));
    n_mm[c] = 0.25 * (q00 + qm0 + q0m + Q(-1, -1, k));
    double x00 = X(0, 0, k), xp0 = X(1, 0, k), xm0 = X(-1, 0, k), x0p = X(0, 1, k), x0m = X(0, -1, k);
    b_pp[c] = 0.25 * (x00 + xp0 + x0p + X(1, 1, k));
    b_pm[c] = 0.25 * (x00 + xp0 + x0m + X(1, -1, k));
    b_mp[c] = 0.25 * (x00 + xm0 + x0p + X(-1, 1, k));
    b_mm[c] = 0.25 * (x00 + xm0 + x0m + X(-1, -1, k));
  }
  auto PM = [&](int di, int dj) { return 0.5 * (Q(di, dj, 5) * X(di, dj, 5) + Q(di, dj, 6) * X(di, dj, 6) + Q(di, dj, 7) * X(di, dj, 7)); };
  auto PC = [&](const double* n, const double* b, int m) { return 0.5 * (n[0] * b[0] + n[1] * b[1] + n[2] * b[2]) - n[m] * b[m]; };
  const double pm00 = PM(0, 0), pmp0 = PM(1, 0), pmm0 = PM(-1, 0), pm0p = PM(0, 1), pm0m = PM(0, -1);
  double Bx_1 = n_0p[0] / rho_0p * (1 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy * ((pm0p - Q(0, 1, 6) * X(0, 1, 6)) - (pm00 - Q(0, 0, 6) * X(0, 0, 6))));
  double Bx_2 = n_0p[1] / rho_0p * (1 / dx * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1 / dy * (Q(0, 1, 6) * X(0, 1, 5) - Q(0, 0, 6) * X(0, 0, 5)));
  double Bx_3 = n_0m[0] / rho_0m * (1 / dx * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy * ((pm00 - Q(0, 0, 6) * X(0, 0, 6)) - (pm0m - Q(0, -1, 6) * X(0, -1, 6))));
  double Bx_4 = n_0m[1] / rho_0m * (1 / dx * (PC(n_pm, b_pm, 0) - PC(n_mm, b_mm, 0)) - 1 / dy * (Q(0, 0, 6) * X(0, 0, 5) - Q(0, -1, 6) * X(0, -1, 5)));
  double ax5 = X(0, 0, 5) + (h * h / dy) * ((Bx_1 + Bx_2) - (Bx_3 + Bx_4));
  double By_1 = n_p0[1] / rho_p0 * (-1 / dx * ((pmp0 - Q(1, 0, 5) * X(1, 0, 5)) - (pm00 - Q(0, 0, 5) * X(0, 0, 5))) + 1 / dy * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
  double By_2 = n_p0[0] / rho_p0 * (-1 / dx * (Q(1, 0, 5) * X(1, 0, 6) - Q(0, 0, 5) * X(0, 0, 6)) + 1 / dy * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
  double By_3 = n_m0[1] / rho_m0 * (-1 / dx * ((pm00 - Q(0, 0, 5) * X(0, 0, 5)) - (pmm0 - Q(-1, 0, 5) * X(-1, 0, 5))) + 1 / dy * (n_mp[0] * b_mp[1] - n_mm[0] * b_mm[1]));
  double By_4 = n_m0[0] / rho_m0 * (-1 / dx * (Q(0, 0, 5) * X(0, 0, 6) - Q(-1, 0, 5) * X(-1, 0, 6)) + 1 / dy * (PC(n_mp, b_mp, 1) - PC(n_mm, b_mm, 1)));
  double ax6 = X(0, 0, 6) + (h * h / dx) * ((By_1 + By_2) - (By_3 + By_4));
  double Bz_1 = n_p0[2] / rho_p0 * (-1 / dx * ((pmp0 - Q(1, 0, 5) * X(1, 0, 5)) - (pm00 - Q(0, 0, 5) * X(0, 0, 5))) + 1 / dy * (n_pp[1] * b_pp[0] - n_pm[1] * b_pm[0]));
  double Bz_2 = n_p0[0] / rho_p0 * (-1 / dx * (Q(1, 0, 5) * X(1, 0, 7) - Q(0, 0, 5) * X(0, 0, 7)) - 1 / dy * (n_pp[1] * b_pp[2] - n_pm[1] * b_pm[2]));
  double Bz_3 = n_m0[2] / rho_m0 * (-1 / dx * ((pm00 - Q(0, 0, 5) * X(0, 0, 5)) - (pmm0 - Q(-1, 0, 5) * X(-1, 0, 5))) + 1 / dy * (n_mp[1] * b_mp[0] - n_mm[1] * b_mm[0]));
  double Bz_4 = n_m0[0] / rho_m0 * (-1 / dx * (Q(0, 0, 5) * X(0, 0, 7) - Q(-1, 0, 5) * X(-1, 0, 7)) - 1 / dy * (n_mp[1] * b_mp[2] - n_mm[1] * b_mm[2]));
  double Bz_5 = n_0p[2] / rho_0p * (1 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1 / dy * ((pm0p - Q(0, 1, 6) * X(0, 1, 6)) - (pm00 - Q(0, 0, 6) * X(0, 0, 6))));
  double Bz_6 = n_0p[1] / rho_0p * (-1 / dx * (n_pp[0] * b_pp[2] - n_mp[0] * b_mp[2]) - 1 / dy * (Q(0, 1, 6) * X(0, 1, 7) - Q(0, 0, 6) * X(0, 0, 7)));
  double Bz_7 = n_0m[2] / rho_0m * (1 / dx * (n_pm[0] * b_pm[1] - n_mm[0] * b_mm[1]) - 1 / dy * ((pm00 - Q(0, 0, 6) * X(0, 0, 6)) - (pm0m - Q(0, -1, 6) * X(0, -1, 6))));
  double Bz_8 = n_0m[1] / rho_0m * (-1 / dx * (n_pm[0] * b_pm[2] - n_mm[0] * b_mm[2]) - 1 / dy * (Q(0, 0, 6) * X(0, 0, 7) - Q(0, -1, 6) * X(0, -1, 7)));
  double ax7 = X(0, 0, 7) + h * h / dx * ((Bz_1 + Bz_2) - (Bz_3 + Bz_4)) + h * h / dy * ((Bz_5 + Bz_6) - (Bz_7 + Bz_8));
  const double nu_x = op.S[S_BMAXX] * dx, nu_y = op.S[S_BMAXY] * dy;
  double lap[3];
  for (int c = 0; c < 3; c++) {
    int k = 5 + c;
    lap[c] = nu_x / (dx * dx) * (X(-1, 0, k) - 2.0 * X(0, 0, k) + X(1, 0, k)) + nu_y / (dy * dy) * (X(0, -1, k) - 2.0 * X(0, 0, k) + X(0, 1, k));
  }
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

template <int OPK> struct OpSel;
template <> struct OpSel<0> {
  static constexpr int NC = 3;
  typedef MagOp T;
  __device__ static void apply(const MagOp& op, const double* xs, int lx, int ly, int i, int j, int, double* out) { mag_apply(op, xs, lx, ly, i, j, out); }
  __device__ static void precond(const MagOp&, int, int, int, double*) {}
  __device__ static const Dims& dims(const MagOp& op) { return op.d; }
};
template <> struct OpSel<1> {
  static constexpr int NC = 1;
  typedef POp T;
  __device__ static void apply(const POp& op, const double* xs, int lx, int ly, int i, int j, int cidx, double* out) { p_apply(op, xs, lx, ly, i, j, cidx, out); }
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
// In the last block: out[v] = sum over blocks b of partial[b*stride + v], v < nv (fixed order per wave).
__device__ __forceinline__ void final_sums(const double* partial, int stride, int nblk, int nv, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int v = wid; v < nv; v += nw) {
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s += partial[(long long)b * stride + v];
    s = wave_sum(s);
    if (lane == 0) out[v] = s;
  }
}

// ---------------------------------------------------------------- GMRES residual r0 = M^-1 (b - A x), gam0 = ||r0||
template <int OPK>
__global__ void __launch_bounds__(256) k_gmres_residual(typename OpSel<OPK>::T op, KrylovBufs kb, const double* __restrict__ x,
                                                        const double* __restrict__ b, KrylovDev* K, int outer, int x_is_zero,
                                                        const int* skip) {
  if (skip && *skip) {           // the caller's outer iteration has already converged: no solve
    if (blockIdx.x == 0 && threadIdx.x == 0) { K->done = 1; K->converged = 1; K->jlast = -1; K->steps = 0; K->outer = outer; }
    return;
  }
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ double xs[NC * (TY + 2) * (TX + 2)];
  __shared__ double red[4];
  __shared__ double res[1];
  const Dims& d = OpSel<OPK>::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc = 0.0;
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    __syncthreads();
    if (!x_is_zero) load_tile<NC>(xs, d, x, 1.0, false, gx0, gy0);
    __syncthreads();
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    if (cx < d.nx && cy < d.ny) {
      int cidx = cx + d.nx * cy;
      double y[NC];
      if (!x_is_zero) OpSel<OPK>::apply(op, xs, tx + 1, ty + 1, cx + GC, cy + GC, cidx, y);
      if (x_is_zero) for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx];
      else for (int c = 0; c < NC; c++) y[c] = b[(long long)c * d.N + cidx] - y[c];
      OpSel<OPK>::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) { kb.w[(long long)c * d.N + cidx] = y[c]; acc += y[c] * y[c]; }
    }
  }
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) kb.partial[blockIdx.x * (MAXM + 2)] = res[0];
  if (!last_block_arrive(&K->cnt_r)) return;
  __shared__ double tot[1];
  final_sums(kb.partial, MAXM + 2, gridDim.x, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    K->cnt_r = 0;
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
template <int OPK>
__global__ void __launch_bounds__(256) k_arnoldi_apply(typename OpSel<OPK>::T op, KrylovBufs kb, VPtrs V, int j, KrylovDev* K) {
  if (K->done) return;
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ double xs[NC * (TY + 2) * (TX + 2)];
  __shared__ double red[4 * (MAXM + 1)];
  __shared__ double outv[MAXM + 1];
  const Dims& d = OpSel<OPK>::dims(op);
  const double h = K->hcur;
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  double acc[MAXM + 1];
#pragma unroll
  for (int i = 0; i <= MAXM; i++) acc[i] = 0.0;
  double* vj = V.v[j];
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    __syncthreads();
    load_tile<NC>(xs, d, kb.w, h, true, gx0, gy0);
    __syncthreads();
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    if (cx < d.nx && cy < d.ny) {
      int cidx = cx + d.nx * cy;
      double y[NC];
      OpSel<OPK>::apply(op, xs, tx + 1, ty + 1, cx + GC, cy + GC, cidx, y);
      OpSel<OPK>::precond(op, cidx, cx + GC, cy + GC, y);
      for (int c = 0; c < NC; c++) {
        long long e = (long long)c * d.N + cidx;
        double vc = xs[(c * (TY + 2) + ty + 1) * (TX + 2) + tx + 1];
        vj[e] = vc;
        kb.Av[e] = y[c];
#pragma unroll
        for (int i = 0; i < MAXM; i++) if (i < j) acc[i] += V.v[i][e] * y[c];
#pragma unroll
        for (int i = 0; i <= MAXM; i++) if (i == j) acc[i] += vc * y[c];
      }
    }
  }
  block_sums<MAXM + 1>(acc, j + 1, red, outv);
  __syncthreads();
  if ((int)threadIdx.x <= j) kb.partial[blockIdx.x * (MAXM + 2) + threadIdx.x] = outv[threadIdx.x];
  if (!last_block_arrive(&K->cnt_a)) return;
  __shared__ double tot[MAXM + 1];
  final_sums(kb.partial, MAXM + 2, gridDim.x, j + 1, tot);
  __syncthreads();
  if ((int)threadIdx.x <= j) K->H[threadIdx.x * (MAXM + 1) + j] = tot[threadIdx.x];
  if (threadIdx.x == 0) K->cnt_a = 0;
}

// ---------------------------------------------------------------- Arnoldi step j, part 2:
// w = Av_j - sum_i H[i][j] v_i, H[j+1][j] = ||w||, Givens rotations, convergence.
__global__ void __launch_bounds__(256) k_arnoldi_update(KrylovBufs kb, VPtrs V, int j, int m, long long n, KrylovDev* K) {
  if (K->done) return;
  __shared__ double Hc[MAXM + 1];
  __shared__ double red[4];
  __shared__ double res[1];
  if ((int)threadIdx.x <= j) Hc[threadIdx.x] = K->H[threadIdx.x * (MAXM + 1) + j];
  __syncthreads();
  double acc = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double w = kb.Av[e];
#pragma unroll
    for (int i = 0; i <= MAXM; i++) if (i <= j) w -= Hc[i] * V.v[i][e];
    kb.w[e] = w;
    acc += w * w;
  }
  block_sums<1>(&acc, 1, red, res);
  if (threadIdx.x == 0) kb.partial[blockIdx.x * (MAXM + 2)] = res[0];
  if (!last_block_arrive(&K->cnt_c)) return;
  __shared__ double tot[1];
  final_sums(kb.partial, MAXM + 2, gridDim.x, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    K->cnt_c = 0;
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
__global__ void __launch_bounds__(256) k_op_apply(typename OpSel<OPK>::T op, const double* __restrict__ x, double* __restrict__ Ax) {
  constexpr int NC = OpSel<OPK>::NC;
  __shared__ double xs[NC * (TY + 2) * (TX + 2)];
  const Dims& d = OpSel<OPK>::dims(op);
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
  for (int tile = blockIdx.x; tile < ntx * nty; tile += gridDim.x) {
    int gx0 = (tile % ntx) * TX, gy0 = (tile / ntx) * TY;
    __syncthreads();
    load_tile<NC>(xs, d, x, 1.0, false, gx0, gy0);
    __syncthreads();
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int cx = gx0 + tx, cy = gy0 + ty;
    if (cx < d.nx && cy < d.ny) {
      int cidx = cx + d.nx * cy;
      double y[NC];
      OpSel<OPK>::apply(op, xs, tx + 1, ty + 1, cx + GC, cy + GC, cidx, y);
      for (int c = 0; c < NC; c++) Ax[(long long)c * d.N + cidx] = y[c];
    }
  }
}

// ---------------------------------------------------------------- Newton residual (implicit_operators.cpp:690-710)
// g = L p - a p + V_mod(p, rho) - b with L the carrier operator; writes -g and ||g||.
__global__ void __launch_bounds__(256) k_newton_residual(POp op, KrylovBufs kb, const double* __restrict__ p,
                                                         const double* __restrict__ b, double* __restrict__ ng,
                                                         KrylovDev* K, int it, double ntol, int* eflag) {
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
  if (!last_block_arrive(&K->cnt_n)) return;
  __shared__ double tot[1];
  final_sums(kb.partial, MAXM + 2, gridDim.x, 1, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    K->cnt_n = 0;
    double r = sqrt(tot[0]);
    if (it == 0) K->newton_res0 = r;
    K->newton_res = r;
    K->newton_it = it;
    K->newton_conv = (r < ntol * K->newton_res0 + ntol) ? 1 : 0;
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
  if (!last_block_arrive(&K->cnt_m)) return;
  if (threadIdx.x == 0) {
    K->cnt_m = 0;
    double a = 1e300, c = -1e300;
    for (int bl = 0; bl < (int)gridDim.x; bl++) { a = fmin(a, partial[bl * 2]); c = fmax(c, partial[bl * 2 + 1]); }
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

void arnoldi_apply(int op, const MagOp& mo, const POp& po, const KrylovBufs& kb, double* const* hostV, int j,
                   KrylovDev* K, hipStream_t st) {
  VPtrs V = make_vptrs(kb, hostV);
  const Dims& d = op == 0 ? mo.d : po.d;
  if (op == 0) k_arnoldi_apply<0><<<grid_for_tiles(d, kb.nblk_apply), 256, 0, st>>>(mo, kb, V, j, K);
  else k_arnoldi_apply<1><<<grid_for_tiles(d, kb.nblk_apply), 256, 0, st>>>(po, kb, V, j, K);
}

void arnoldi_update(int nc, const Dims& d, const KrylovBufs& kb, double* const* hostV, int j, int m, KrylovDev* K,
                    hipStream_t st) {
  VPtrs V = make_vptrs(kb, hostV);
  long long n = (long long)nc * d.N;
  long long nb = (n + 255) / 256;
  int g = (int)(nb < kb.nblk_vec ? nb : kb.nblk_vec);
  k_arnoldi_update<<<g, 256, 0, st>>>(kb, V, j, m, n, K);
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
  k_newton_residual<<<grid_for_tiles(po.d, kb.nblk_apply), 256, 0, st>>>(po, kb, p, b, ng, K, it, ntol, eflag);
}

void newton_mbig(const Dims& d, const Phys& P, const double* b, const double* u, double* S, double* partial, KrylovDev* K,
                 int* eflag, hipStream_t st) {
  long long nb = (d.N + 255) / 256;
  int g = (int)(nb < 512 ? nb : 512);
  k_newton_mbig<<<g, 256, 0, st>>>(d, P, b, u, S, partial, K, eflag);
}

void vslope_launch(const Dims& d, const Phys& P, const double* pf, const double* u, const double* S, double* vs, int* eflag,
                   hipStream_t st) {
  k_vslope<<<(d.N + 255) / 256, 256, 0, st>>>(d, P, pf, u, S, vs, eflag);
}

void axpy_launch(long long n, double* y, const double* x, const int* skip, hipStream_t st) {
  k_axpy1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, y, x, skip);
}
