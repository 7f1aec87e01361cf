// kernels_potential.hip — the vector-potential treatments of the magnetic field (UnStaggerCT and
// PotentialA, main.cpp:424-434, 497-510) and the NoDivFree face pass-through.
//
// A is held as 3 SoA planes (A_x, A_y, A_z) with the ghost frame.  Only interior values of the
// stage combinations are ever read; ghost values are recomputed by bound_cond_A whenever the
// reference recomputes them (after each update, and inside every potential-operator apply, where
// the operator kernel evaluates the same rule on the fly: kernels_krylov.hip).
#include "kernels.hpp"

// ================================================================ bound_cond_A (boundary.cpp:177-292)
// Periodic: copy.  Transmissive: linear extrapolation from the two edge cells.  The reference fills
// x over every row, then y over every column (reading the x-filled values), so a corner ghost is the
// y-rule applied to x-rule values; X() below is the x-filled value of an interior row.
__global__ void k_bc_a(const Inst* __restrict__ IT, BMap bm, int ia) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const int nbot = GC * d.NX, ntop = GC * d.NX, nside = 2 * GC;
  const int total = nbot + ntop + d.ny * nside;
  int t = bx_ * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int i, j;
  if (t < nbot) { j = t / d.NX; i = t % d.NX; }
  else if (t < nbot + ntop) { int u = t - nbot; j = d.ny + GC + u / d.NX; i = u % d.NX; }
  else { int u = t - nbot - ntop; j = GC + u / nside; int c = u % nside; i = c < GC ? c : d.nx + GC + (c - GC); }
  const int iR = d.nx + GC - 1, jR = d.ny + GC - 1;
  for (int k = 0; k < 3; k++) {
    double* a = gmem(I_.buf[ia]) + k * d.PS;
    auto X = [&](int ii, int jj) -> double {
      if (ii >= GC && ii <= iR) return a[pidx(d, ii, jj)];
      if (d.perx) return a[pidx(d, ii < GC ? ii + d.nx : ii - d.nx, jj)];
      if (ii < GC) {
        double sL = a[pidx(d, GC + 1, jj)] - a[pidx(d, GC, jj)];
        return a[pidx(d, GC, jj)] - (GC - ii) * sL;
      }
      double sR = a[pidx(d, iR, jj)] - a[pidx(d, iR - 1, jj)];
      return a[pidx(d, iR, jj)] + (ii - iR) * sR;
    };
    double v;
    if (j >= GC && j <= jR) v = X(i, j);
    else if (j < GC && jet_inlet(d, i)) v = a[pidx(d, i, 2 * GC - 1 - j)];   // even reflection in the jet strip (:279-291)
    else if (d.pery) v = X(i, j < GC ? j + d.ny : j - d.ny);
    else if (j < GC) { double x0 = X(i, GC), sL = X(i, GC + 1) - x0; v = x0 - (GC - j) * sL; }
    else { double xr = X(i, jR), sR = xr - X(i, jR - 1); v = xr + (j - jR) * sR; }
    a[pidx(d, i, j)] = v;
  }
}

// ================================================================ stage combination of A (main.cpp:397-412)
// Stage 2: A_s = A_n + Ai(1,0) dt HsA_0 with HsA_0 = (A_1 - A_n) / adt_0 (stage 1's A_s is A_n).
__global__ void k_a_combine(const Inst* __restrict__ IT, BMap bm, int ian, int ia1, int ias) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= (long long)d.NX * d.NY) return;
  const double* S = gmem(I_.S);
  const double bb = 1.0 - 1.0 / sqrt(2.0);
  const double dt = S[S_DT], adt0 = bb * dt, ci = (1 - bb) * dt;
  const double* an = gmem(I_.buf[ian]);
  const double* a1 = gmem(I_.buf[ia1]);
  double* as = gmem(I_.buf[ias]);
  for (int k = 0; k < 3; k++) {
    double n = an[k * d.PS + c];
    double hs = (a1[k * d.PS + c] - n) / adt0;
    as[k * d.PS + c] = n + ci * hs;
  }
}

// ================================================================ MagSubsystem_UpdateA (imex_subsystems.cpp:5-74)
// b = A_s - h B^n x S with S = v** - (h / rho) grad p^e (central), x0 = A (compact, M_B order).
__global__ void __launch_bounds__(256) k_a_rhs(const Inst* __restrict__ IT, BMap bm, int iu, int iQe, int ias, int ia0,
                                               int ib, int ix0) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ pe = gmem(I_.buf[B_PE]);
  const double* __restrict__ as = gmem(I_.buf[ias]);
  const double* __restrict__ a0 = gmem(I_.buf[ia0]);
  double* __restrict__ b = gmem(I_.buf[ib]);
  double* __restrict__ x0 = gmem(I_.buf[ix0]);
  const double* S = gmem(I_.S);
  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double h = S[S_ADT], dx = d.dx, dy = d.dy;
  long long c = pidx(d, i, j);
  double Bxn = Qe[5 * d.PS + c], Byn = Qe[6 * d.PS + c], Bzn = Qe[7 * d.PS + c];
  double rho = u[c];
  double Sx = u[1 * d.PS + c] / rho - h / rho * (pe[pidx(d, i + 1, j)] - pe[pidx(d, i - 1, j)]) / (2.0 * dx);
  double Sy = u[2 * d.PS + c] / rho - h / rho * (pe[pidx(d, i, j + 1)] - pe[pidx(d, i, j - 1)]) / (2.0 * dy);
  double Sz = u[3 * d.PS + c] / rho;
  int idx = (i - GC) + d.nx * (j - GC);
  b[idx] = as[c] - h * (Byn * Sz - Bzn * Sy);
  b[idx + d.N] = as[d.PS + c] - h * (Bzn * Sx - Bxn * Sz);
  b[idx + 2 * d.N] = as[2 * d.PS + c] - h * (Bxn * Sy - Byn * Sx);
  for (int k = 0; k < 3; k++) x0[idx + k * d.N] = a0[k * d.PS + c];
}

// Viscosity maxima of the potential operator (implicit_operators.cpp:359-374): over the interior,
// max of 0.5 (|v_x| + sqrt(v_x^2 + 4 |B|^2 / rho)) and the y analogue, from Qs.  Per-block partial
// maxima, then one block per instance writes S_GBX, S_GBY (maxima are exact in any order).
__global__ void k_pot_visc(const Inst* __restrict__ IT, BMap bm, int iQs) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ q = gmem(I_.buf[iQs]);
  __shared__ double sx[4], sy[4];
  double mx = 0.0, my = 0.0;
  for (long long t = bx_ * (long long)blockDim.x + threadIdx.x; t < d.N; t += (long long)gx_ * blockDim.x) {
    int i = (int)(t % d.nx) + GC, j = (int)(t / d.nx) + GC;
    long long c = pidx(d, i, j);
    double rho = q[c], vx = q[d.PS + c] / rho, vy = q[2 * d.PS + c] / rho;
    double B5 = q[5 * d.PS + c], B6 = q[6 * d.PS + c], B7 = q[7 * d.PS + c];
    double Bsq = B5 * B5 + B6 * B6 + B7 * B7;
    double bxv = 0.5 * (fabs(vx) + sqrt(vx * vx + 4.0 * Bsq / rho));
    double byv = 0.5 * (fabs(vy) + sqrt(vy * vy + 4.0 * Bsq / rho));
    mx = fmax(mx, bxv); my = fmax(my, byv);
  }
  mx = wave_max(mx); my = wave_max(my);
  if ((threadIdx.x & 63) == 0) { sx[threadIdx.x >> 6] = mx; sy[threadIdx.x >> 6] = my; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mx = fmax(mx, sx[w]); my = fmax(my, sy[w]); }
    mx = fmax(mx, sx[0]); my = fmax(my, sy[0]);
    gmem(I_.partial)[2 * bx_] = mx;
    gmem(I_.partial)[2 * bx_ + 1] = my;
  }
}
__global__ void k_pot_visc_fin(const Inst* __restrict__ IT, BMap bm, BMap src) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const int nblk = bmap_count(src, inst_);
  double mx = 0.0, my = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) { mx = fmax(mx, gmem(I_.partial)[2 * b]); my = fmax(my, gmem(I_.partial)[2 * b + 1]); }
  mx = wave_max(mx); my = wave_max(my);
  if (threadIdx.x == 0) { gmem(I_.S)[S_GBX] = mx; gmem(I_.S)[S_GBY] = my; }
}

// A interior <- the solution (compact, M_B order)
__global__ void k_a_from_compact(const Inst* __restrict__ IT, BMap bm, int ix, int ia) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double* x = gmem(I_.buf[ix]);
  double* a = gmem(I_.buf[ia]);
  int idx = (i - GC) + d.nx * (j - GC);
  for (int k = 0; k < 3; k++) a[k * d.PS + pidx(d, i, j)] = x[idx + k * d.N];
}

// Cell B = curl A (central).  mode 0: MagSubsystem_UpdateA's form, B_x, B_y, B_z
// (imex_subsystems.cpp:63-70); mode 1: Unstaggered_CT_Consistent's form, B_x, B_y
// (constrained_transport.cpp:533-538).  The two write the same quotients in a different order.
__global__ void k_a_curl(const Inst* __restrict__ IT, BMap bm, int ia, int iu, int mode) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double* a = gmem(I_.buf[ia]);
  double* u = gmem(I_.buf[iu]);
  auto A = [&](int ii, int jj, int k) { return a[k * d.PS + pidx(d, ii, jj)]; };
  const double dx = d.dx, dy = d.dy;
  long long c = pidx(d, i, j);
  if (mode == 0) {
    u[5 * d.PS + c] = 1 / (2.0 * dy) * (A(i, j + 1, 2) - A(i, j - 1, 2));
    u[6 * d.PS + c] = -1 / (2.0 * dx) * (A(i + 1, j, 2) - A(i - 1, j, 2));
    u[7 * d.PS + c] = 1 / (2.0 * dx) * (A(i + 1, j, 1) - A(i - 1, j, 1)) - 1 / (2.0 * dy) * (A(i, j + 1, 0) - A(i, j - 1, 0));
  } else {
    u[5 * d.PS + c] = (A(i, j + 1, 2) - A(i, j - 1, 2)) / (2.0 * dy);
    u[6 * d.PS + c] = -(A(i + 1, j, 2) - A(i - 1, j, 2)) / (2.0 * dx);
  }
}

// ================================================================ Unstaggered_CT_Consistent (constrained_transport.cpp:502-540)
// With the consistent corner EMFs Ez_pp (k_ct_corner, as for the staggered CT): A_z = A_s,z - h times
// the mean of the cell's four corners; A_x, A_y keep the stage's values (read from `ia01`).
__global__ void k_a_ct(const Inst* __restrict__ IT, BMap bm, int ias, int ia01, int ia) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double* ez = gmem(I_.buf[B_EZPP]);
  const double* as = gmem(I_.buf[ias]);
  const double* a01 = gmem(I_.buf[ia01]);
  double* a = gmem(I_.buf[ia]);
  const double h = gmem(I_.S)[S_ADT];
  long long c = pidx(d, i, j);
  double Ez_cell = 0.25 * (ez[c] + ez[pidx(d, i - 1, j)] + ez[pidx(d, i, j - 1)] + ez[pidx(d, i - 1, j - 1)]);
  a[c] = a01[c];
  a[d.PS + c] = a01[d.PS + c];
  a[2 * d.PS + c] = as[2 * d.PS + c] - h * Ez_cell;
}

// The staggered faces are left alone by UnStaggerCT, PotentialA and NoDivFree: the stage output faces
// are the stage input faces.
__global__ void k_copy_faces(const Inst* __restrict__ IT, BMap bm, int ibx, int iby, int obx, int oby) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= (long long)d.NX * d.NY) return;
  gmem(I_.buf[obx])[c] = gmem(I_.buf[ibx])[c];
  gmem(I_.buf[oby])[c] = gmem(I_.buf[iby])[c];
}

// cal_divergence_error for the cell-centred treatments (imex_subsystems.cpp:640-655): central inside,
// one-sided at the domain edges; per-block partial sums like k_divb.
__global__ void k_divb_cell(const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  __shared__ double sh[16];
  double acc = 0.0;
  auto U = [&](int ii, int jj, int k) { return u[k * d.PS + pidx(d, ii, jj)]; };
  for (long long t = bx_ * (long long)blockDim.x + threadIdx.x; t < d.N; t += (long long)gx_ * blockDim.x) {
    int i = (int)(t % d.nx) + GC, j = (int)(t / d.nx) + GC;
    double dBxdx, dBydy;
    if (i == GC) dBxdx = (U(i + 1, j, 5) - U(i, j, 5)) / d.dx;
    else if (i == d.nx + GC - 1) dBxdx = (U(i, j, 5) - U(i - 1, j, 5)) / d.dx;
    else dBxdx = (U(i + 1, j, 5) - U(i - 1, j, 5)) / (2.0 * d.dx);
    if (j == GC) dBydy = (U(i, j + 1, 6) - U(i, j, 6)) / d.dy;
    else if (j == d.ny + GC - 1) dBydy = (U(i, j, 6) - U(i, j - 1, 6)) / d.dy;
    else dBydy = (U(i, j + 1, 6) - U(i, j - 1, 6)) / (2.0 * d.dy);
    double v = dBxdx + dBydy;
    acc += d.dx * d.dy * v * v;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) gmem(I_.partial)[bx_] = s;
}

// ================================================================ launchers
#define GRID(L, sh) dim3((unsigned)(L).nblocks[sh])
static const dim3 B64x4(64, 4);

void launch_bc_a(const Launch& L, int ia) {
  k_bc_a<<<GRID(L, SH_FILL0), 256, 0, L.st>>>(L.IT, L.bm[SH_FILL0], ia);
}
void launch_a_combine(const Launch& L, int ian, int ia1, int ias) {
  k_a_combine<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(L.IT, L.bm[SH_LIN_TOT], ian, ia1, ias);
}
void launch_a_rhs(const Launch& L, int iu, int iQe, int ias, int ia0, int ib, int ix0) {
  k_a_rhs<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], iu, iQe, ias, ia0, ib, ix0);
}
void launch_pot_visc(const Launch& L, int iQs) {
  k_pot_visc<<<GRID(L, SH_DIVB), 256, 0, L.st>>>(L.IT, L.bm[SH_DIVB], iQs);
  k_pot_visc_fin<<<GRID(L, SH_ONE), 64, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_DIVB]);
}
void launch_a_update(const Launch& L, int ix, int ia, int iu) {
  k_a_from_compact<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], ix, ia);
  launch_bc_a(L, ia);
  k_a_curl<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], ia, iu, 0);
}
void launch_unstaggered_ct(const Launch& L, int iu, int ias, int ia01, int ia) {
  launch_ct_corners(L, iu);
  k_a_ct<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], ias, ia01, ia);
  launch_bc_a(L, ia);
  k_a_curl<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], ia, iu, 1);
}
void launch_copy_faces(const Launch& L, int ibx, int iby, int obx, int oby) {
  k_copy_faces<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(L.IT, L.bm[SH_LIN_TOT], ibx, iby, obx, oby);
}
void launch_divb_cell(const Launch& L, int iu) {
  k_divb_cell<<<GRID(L, SH_DIVB), 256, 0, L.st>>>(L.IT, L.bm[SH_DIVB], iu);
}
