// kernels_stage.hip — stencil and pointwise kernels of one IMEX stage (FP64 form).
//
// Every kernel works on SoA FP64 planes with the reference's ghost frame.  Kernels
// that the reference runs on the interior followed by a ghost fill are either
// written that way (interior kernel + k_bc_fill) or, where cheaper, evaluated at
// every padded cell through the boundary index map, which produces the same
// values as the reference's copy-type ghost fill.
#include "kernels.hpp"
#include <type_traits>

// ================================================================ ghost fill
// bound_cond_u / bound_cond_array (boundary.cpp:5-175, timestep.cpp:166-222) and
// bound_cond_staggered (constrained_transport.cpp:11-64): ghost (i,j) <- (mx(i), my(j)).
struct FillMap { int lox, hix, loy, hiy; };   // identity range of the map
// planes: buffer id b[p], plane k[p] of that buffer; jv[p] the plane's value in the jet inflow strip's
// ghost rows: -1 = the copy (arrays), -2 = left as the producing kernel wrote it, 0..7 = conservative
// slot, 8..15 = primitive slot (k - 8)
// rn[p] = 1: the plane's reflective ghost is the negated mirror value (momenta, B, bound_cond_array arrays)
struct FillPlanes { int b[8], k[8], n; signed char jv[8], rn[8]; };

__device__ __forceinline__ int fill_mx(const Dims& d, int kind, int i) {
  if (d.perx || (kind == 0 && d.refx)) return map_x(d, i);
  int lo = (kind == 1) ? GC - 1 : GC;           // staggered x-normal keeps face g-1
  if (i < lo) return lo;
  if (i > d.nx + GC - 1) return d.nx + GC - 1;
  return i;
}
__device__ __forceinline__ int fill_my(const Dims& d, int kind, int j) {
  if (d.pery || (kind == 0 && d.refy)) return map_y(d, j);
  int lo = (kind == 2) ? GC - 1 : GC;           // staggered y-normal keeps face g-1
  if (j < lo) return lo;
  if (j > d.ny + GC - 1) return d.ny + GC - 1;
  return j;
}

__global__ void k_bc_fill(const Inst* __restrict__ IT, BMap bm, FillPlanes fp, int kind) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  FillMap fm;
  fm.lox = (!d.perx && kind == 1) ? GC - 1 : GC;
  fm.hix = d.nx + GC - 1;
  fm.loy = (!d.pery && kind == 2) ? GC - 1 : GC;
  fm.hiy = d.ny + GC - 1;
  const int nbot = fm.loy * d.NX;
  const int ntop = (d.NY - 1 - fm.hiy) * d.NX;
  const int nside_l = fm.lox, nside_r = d.NX - 1 - fm.hix;
  const int total = nbot + ntop + (fm.hiy - fm.loy + 1) * (nside_l + nside_r);

  int t = bx_ * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int i, j;
  int rows_mid = fm.hiy - fm.loy + 1;
  if (t < nbot) { j = t / d.NX; i = t % d.NX; }
  else if (t < nbot + ntop) { int u = t - nbot; j = fm.hiy + 1 + u / d.NX; i = u % d.NX; }
  else {
    int u = t - nbot - ntop;
    int per_row = nside_l + nside_r;
    int r = u / per_row, c = u % per_row;
    j = fm.loy + r;
    i = (c < nside_l) ? c : (fm.hix + 1 + (c - nside_l));
    if (r >= rows_mid) return;
  }
  int si = fill_mx(d, kind, i), sj = fill_my(d, kind, j);
  long long dst = pidx(d, i, j), src = pidx(d, si, sj);
  // y_JetInflow overrides (boundary.cpp:148-172, constrained_transport.cpp:50-63)
  int jet_mode = 0;
  if (d.jet) {
    if (kind == 0 && j < GC && jet_inlet(d, i)) jet_mode = 1;
    else if (kind == 1 && j < GC && i < d.NX - 1 && jet_inlet_face(d, i)) jet_mode = 2;
    else if (kind == 2 && j < GC - 1 && jet_inlet(d, i)) jet_mode = 3;
  }
  const bool dg = kind == 0 && dir_ghost(d, i, j);   // Dirichlet ambient ghost (boundary.cpp:56-76, 128-147)
  for (int p = 0; p < fp.n; p++) {
    double* a = gmem(I_.buf[fp.b[p]]) + fp.k[p] * d.PS;
    if ((jet_mode == 1 || dg) && fp.jv[p] == -2) continue;            // formed from the ghost's own state already
    if (dg && fp.jv[p] >= 0) {
      const int slot = fp.jv[p] & 7;
      const double rf = refl_factor(d, i, j, true);
      double v;
      if (slot == 0) v = d.rho_amb;
      else if (slot < 4) v = 0.0;
      else if (slot > 4) v = rf * a[src];                           // B: the boundary cell's
      else if (fp.jv[p] >= 8) v = d.p_amb;
      else {
        const double* g0 = gmem(I_.buf[fp.b[p]]);
        double b5 = rf * g0[5 * d.PS + src], b6 = rf * g0[6 * d.PS + src], b7 = rf * g0[7 * d.PS + src];
        v = d.e_amb + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);
      }
      a[dst] = v;
      continue;
    }
    if (jet_mode == 1 && fp.jv[p] == -2) continue;                    // formed from the strip's state already
    if (jet_mode == 1 && fp.jv[p] >= 0) a[dst] = jet_state(d, fp.jv[p] >= 8, fp.jv[p] & 7);
    else if (jet_mode == 2) a[dst] = 0.0;
    else if (jet_mode == 3) a[dst] = d.jet_B0;
    else a[dst] = kind == 0 ? refl_factor(d, i, j, fp.rn[p]) * a[src] : a[src];
  }
}



// ================================================================ start of step
// calch + caldt + MagneticEigenvalues (timestep.cpp:5-67, 89-120, 254-273) fused with
// the con -> pri -> con round trip of main.cpp:356/561: u_n = pri2con(pri).
// mode 0: `uin` holds primitives; mode 1: `uin` holds conservatives (pri = con2pri).
// RK: the Redlich-Kwong instantiation; the ideal-gas one carries no temperature solves (fewer registers)
template <bool RK>
__global__ void __launch_bounds__(256) k_step_start(Phys P_, const Inst* __restrict__ IT, BMap bm, int mode, int iin) {
  BATCH_BLOCK(bm);
  Phys P = P_;
  if constexpr (!RK) P.eos = IMEX_EOS_IDEAL;
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ uin = gmem(I_.buf[iin]);
  double* __restrict__ un = gmem(I_.buf[B_UN]);
  double* partial = gmem(I_.partial);
  int* eflag = gmem(I_.eflag);

  __shared__ double sh[16];
  double r[8] = {1e-8, 1e-8, 1e-8, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long long total = (long long)d.NX * d.NY;
  int ef = 0;
  for (long long c = bx_ * (long long)blockDim.x + threadIdx.x; c < total; c += (long long)gx_ * blockDim.x) {
    int i = (int)(c % d.NX), j = (int)(c / d.NX);
    double q[NVAR], pri[NVAR];
    for (int k = 0; k < NVAR; k++) q[k] = uin[k * d.PS + c];
    if (mode == 0) { for (int k = 0; k < NVAR; k++) pri[k] = q[k]; }
    else if (j < GC && jet_inlet(d, i)) { for (int k = 0; k < NVAR; k++) pri[k] = jet_state(d, 1, k); }   // g_bc_prim fill
    else if (dir_ghost(d, i, j)) {                       // g_bc_prim fill of a Dirichlet ghost
      pri[0] = d.rho_amb; pri[1] = pri[2] = pri[3] = 0.0; pri[4] = d.p_amb;
      pri[5] = q[5]; pri[6] = q[6]; pri[7] = q[7];
    }
    else con2pri(P, pri, q, &ef);
    double rho = pri[0], p = pri[4], Bx = pri[5], By = pri[6], Bz = pri[7];
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double B = sqrt(B2);
    double v = sqrt(pri[1] * pri[1] + pri[2] * pri[2] + pri[3] * pri[3]);
    double cfx = eos_fast(P, rho, p, Bx, B2, &ef), cfy = eos_fast(P, rho, p, By, B2, &ef), cf = eos_fast(P, rho, p, B, B2, &ef);
    r[0] = fmax(r[0], pri[1] + cfx);
    r[1] = fmax(r[1], pri[2] + cfy);
    r[2] = fmax(r[2], v + cf);
    double bsq = pri[5] * pri[5] + pri[6] * pri[6] + pri[7] * pri[7];
    double bx = 0.5 * (fabs(pri[1]) + sqrt(pri[1] * pri[1] + 4 * bsq / pri[0]));
    double by = 0.5 * (fabs(pri[2]) + sqrt(pri[2] * pri[2] + 4 * bsq / pri[0]));
    r[3] = fmax(r[3], bx);
    r[4] = fmax(r[4], by);
    if (i >= GC && i < d.nx + GC && j >= GC && j < d.ny + GC) {
      r[5] = fmax(r[5], fabs(pri[1]));
      r[6] = fmax(r[6], fabs(pri[2]));
      r[7] = fmax(r[7], v);
    }
    double con[NVAR];
    pri2con(P, con, pri, &ef);
    for (int k = 0; k < NVAR; k++) un[k * d.PS + c] = con[k];
  }
  if (ef) atomicOr(eflag, 1);
  // block max of the 8 values
  for (int k = 0; k < 8; k++) {
    double v = wave_max(r[k]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = sh[0];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sh[w]);
      partial[bx_ * 8 + k] = m;
    }
  }
}

// One-block finish of the start-of-step maxima and the dt rule (timestep.cpp:22-66, main.cpp:341, 353).
__global__ void k_step_finish(Phys P, const Inst* __restrict__ IT, BMap bm, BMap src) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* partial = gmem(I_.partial);
  const int nblk = bmap_count(src, inst_);
  double* S = gmem(I_.S);

  __shared__ double mm[8];
  {   // 8 maxima over the block partials, one wave per value (max is order independent)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int k = wid; k < 8; k += nw) {
      double v = partial[k];
      for (int b = lane; b < nblk; b += 64) v = fmax(v, partial[b * 8 + k]);
      v = wave_max(v);
      if (lane == 0) mm[k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m[8];
    for (int k = 0; k < 8; k++) m[k] = mm[k];
    int ts = (int)S[S_TIMESTEP] + 1;
    double ch = m[2], v_max = m[7];
    S[S_CHX] = m[0]; S[S_CHY] = m[1]; S[S_CH] = ch;
    S[S_VXMAX] = m[5]; S[S_VYMAX] = m[6]; S[S_VMAX] = v_max;
    S[S_BMAXX] = m[3]; S[S_BMAXY] = m[4];
    double a_max = (ts < 2) ? ch : v_max;
    double dt = P.cfl * 1 / (2 * a_max / fmin(d.dx, d.dy));
    double prev = S[S_PREVDT];
    if (ts > 1 && prev > 0.0) dt = fmin(dt, 2.0 * prev);
    S[S_PREVDT] = dt;
    if (ts > 1) a_max = v_max;
    S[S_AMAX] = a_max;
    dt = fmin(dt, P.time_stop - S[S_T]);
    S[S_DT] = dt;
    S[S_T] = S[S_T] + dt;
    S[S_TIMESTEP] = (double)ts;
    const double bb = 1.0 - 1.0 / sqrt(2.0);
    S[S_ADT] = bb * dt;
  }
}

// ================================================================ stage combine (main.cpp:368-414)
// Stage 1 of IMEX-RK2: Hs0 = (u1 - u_n)/adt0, Qe = u_n + Ae10 dt Hs0, Qs = u_n + Ai10 dt Hs0.
__global__ void k_stage_combine(const Inst* __restrict__ IT, BMap bm, int ibxn, int ibyn) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ un = gmem(I_.buf[B_UN]);
  const double* __restrict__ u1 = gmem(I_.buf[B_U1]);
  double* __restrict__ Qe = gmem(I_.buf[B_QE]);
  double* __restrict__ Qs = gmem(I_.buf[B_QS]);
  const double* __restrict__ bxn = gmem(I_.buf[ibxn]);
  const double* __restrict__ byn = gmem(I_.buf[ibyn]);
  const double* __restrict__ bx1 = gmem(I_.buf[B_BX1]);
  const double* __restrict__ by1 = gmem(I_.buf[B_BY1]);
  double* __restrict__ bxqs = gmem(I_.buf[B_BXQS]);
  double* __restrict__ byqs = gmem(I_.buf[B_BYQS]);
  const double* S = gmem(I_.S);

  long long total = (long long)d.NX * d.NY;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= total) return;
  const double bb = 1.0 - 1.0 / sqrt(2.0);
  const double dt = S[S_DT], adt0 = bb * dt;
  const double ce = (1.0 / (2.0 * bb)) * dt, ci = (1 - bb) * dt;
  for (int k = 0; k < NVAR; k++) {
    double n = un[k * d.PS + c];
    double hs = (u1[k * d.PS + c] - n) / adt0;
    Qe[k * d.PS + c] = n + ce * hs;
    Qs[k * d.PS + c] = n + ci * hs;
  }
  double hbx = (bx1[c] - bxn[c]) / adt0, hby = (by1[c] - byn[c]) / adt0;
  bxqs[c] = bxn[c] + ci * hbx;
  byqs[c] = byn[c] + ci * hby;
}

// ================================================================ per-stage Qe preparation
// p(Qe) at every padded cell (Qe_pri(.,.,4)) and the fast speed sqrt(a^2 + |B|^2/rho)
// used by ShockDetector (flux.cpp:22-37).
__global__ void k_qe_prep(Phys P, const Inst* __restrict__ IT, BMap bm, int iQe) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  double* __restrict__ pe = gmem(I_.buf[B_PE]);
  double* __restrict__ cf = gmem(I_.buf[B_CF]);
  int* eflag = gmem(I_.eflag);

  long long total = (long long)d.NX * d.NY;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= total) return;
  double q[NVAR];
  for (int k = 0; k < NVAR; k++) q[k] = Qe[k * d.PS + c];
  int ef = 0;
  double p = pressure_of(P, q, &ef);
  pe[c] = p;
  double rho = q[0];
  double a = eos_a(P, rho, p, &ef);
  double a2 = a * a;
  double cA2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho;
  cf[c] = sqrt(a2 + cA2);
  if (ef) atomicOr(eflag, 1);
}

// ShockDetector step 1 (flux.cpp:14-53): chi~ on the interior.
__global__ void k_shock_chi(Phys P, const Inst* __restrict__ IT, BMap bm, int iQe) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ cf = gmem(I_.buf[B_CF]);
  double* __restrict__ chi = gmem(I_.buf[B_CHI]);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  double cmin = 1.0e300;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++) {
      double v = cf[pidx(d, i + di, j + dj)];
      if (v < cmin) cmin = v;
    }
  const double* r0 = Qe; const double* m1 = Qe + d.PS; const double* m2 = Qe + 2 * d.PS;
  double vxp = m1[pidx(d, i + 1, j)] / r0[pidx(d, i + 1, j)], vxm = m1[pidx(d, i - 1, j)] / r0[pidx(d, i - 1, j)];
  double vyp = m2[pidx(d, i, j + 1)] / r0[pidx(d, i, j + 1)], vym = m2[pidx(d, i, j - 1)] / r0[pidx(d, i, j - 1)];
  double div_u = (vxp - vxm) / (2.0 * d.dx) + (vyp - vym) / (2.0 * d.dy);
  double h = fmin(d.dx, d.dy);
  double num = h * div_u + P.kappa1 * cmin;
  double den = P.kappa1 * cmin;
  chi[pidx(d, i, j)] = fmin(1.0, fmax(0.0, -num / den));
}

// ShockDetector step 2 (flux.cpp:60-75): eta = 3x3 max of chi~.
__global__ void k_shock_eta(const Inst* __restrict__ IT, BMap bm) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ chi = gmem(I_.buf[B_CHI]);
  double* __restrict__ eta = gmem(I_.buf[B_ETA]);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  double m = 0.0;
  for (int di = -1; di <= 1; di++)
    for (int dj = -1; dj <= 1; dj++) {
      double v = chi[pidx(d, i + di, j + dj)];
      if (v > m) m = v;
    }
  eta[pidx(d, i, j)] = m;
}

// ================================================================ reconstruction + explicit fluxes
// TVDReconstruction (reconstruction.cpp:5-287) at one cell from the 5-point stencil a[0..4].
// FP32 reciprocal for the mixed form's flux kernels (v_rcp_f32, about 1 ulp)
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }

// ENO3 stencil choice (reconstruction.cpp:104-183): 0 = cells i-2..i, 1 = i-1..i+1, 2 = i..i+2, by
// comparing first and second differences.
template <class U>
__device__ __forceinline__ int eno3_stencil(const U* a) {
  const U um2 = a[0], um1 = a[1], u0 = a[2], up1 = a[3], up2 = a[4];
  U d1 = fabs(u0 - um1), d2 = fabs(up1 - u0), d3, d4;
  if (d1 >= d2) {
    d3 = fabs(fabs(up2 - up1) - fabs(up1 - u0));
    d4 = fabs(fabs(up1 - u0) - fabs(u0 - um1));
    return d3 >= d4 ? 1 : 2;
  }
  d3 = fabs(fabs(u0 - um1) - fabs(um1 - um2));
  d4 = fabs(fabs(up1 - u0) - fabs(u0 - um1));
  return d4 >= d3 ? 0 : 1;
}

// eno_sel >= 0 gives the ENO3 stencil chosen beforehand (in FP64 by the mixed form's flux kernel: the
// choice is discontinuous in the data, so choosing on FP32-rounded values would flip it near ties).
template <class T>
__device__ __forceinline__ void recon5(int recon, int central, const T* a, T& wL, T& wR, int eno_sel = -1) {
  const T um2 = a[0], um1 = a[1], u0 = a[2], up1 = a[3], up2 = a[4];
  if (recon == IMEX_RECON_NONE) {
    wL = u0; wR = u0;
  } else if (recon == IMEX_RECON_ENO3) {
    // the chosen third-order stencil's two face values
    const int s = eno_sel >= 0 ? eno_sel : eno3_stencil(a);
    if (s == 0) {
      wL = T(-1.0 / 6.0) * um2 + T(5.0 / 6.0) * um1 + T(1.0 / 3.0) * u0;
      wR = T(1.0 / 3.0) * um2 - T(7.0 / 6.0) * um1 + T(11.0 / 6.0) * u0;
    } else if (s == 1) {
      wL = T(1.0 / 3.0) * um1 + T(5.0 / 6.0) * u0 - T(1.0 / 6.0) * up1;
      wR = T(-1.0 / 6.0) * um1 + T(5.0 / 6.0) * u0 + T(1.0 / 3.0) * up1;
    } else {
      wL = T(11.0 / 6.0) * u0 - T(7.0 / 6.0) * up1 + T(1.0 / 3.0) * up2;
      wR = T(1.0 / 3.0) * u0 + T(5.0 / 6.0) * up1 - T(1.0 / 6.0) * up2;
    }
  } else if (recon == IMEX_RECON_MINMOD) {
    T sR = up1 - u0, sL = u0 - um1, c;
    if (sL * sR <= T(0)) c = T(0);
    else if (fabs(sR) <= fabs(sL)) c = sR;
    else c = sL;
    if (central) c = T(0.5) * (sL + sR);
    wL = u0 - T(0.5) * c; wR = u0 + T(0.5) * c;
  } else if (recon == IMEX_RECON_VANLEER) {
    // c = min(2/(1+r), 2r/(1+r)) = 2 min(1, r)/(1+r) for r >= 0
    T sR = up1 - u0, sL = u0 - um1, c;
    if (fabs(sR) < T(1e-16)) c = T(0);
    else {
      T r = std::is_same<T, float>::value ? sL * rcp_t(sR) : sL / sR;
      c = (r < T(0)) ? T(0) : (std::is_same<T, float>::value ? T(2) * fmin(T(1), r) * rcp_t(T(1) + r)
                                                               : T(2) * fmin(T(1), r) / (T(1) + r));
    }
    wL = u0 - T(0.5) * c * sL; wR = u0 + T(0.5) * c * sR;
  } else {
    // Jiang-Shu weights w_k = (d_k/e_k) / sum_m (d_m/e_m), e_k = (eps + beta_k)^2.  FP64: written
    // with the common denominator e0 e1 e2 so that each face value costs one division.  FP32:
    // hardware reciprocals (the products e_i e_j could overflow the FP32 range).
    // Multiply-adds are fused here (the weights are smooth in the data, so rounding moves the face
    // values by an ulp and decides nothing); the limiter and ENO branches above, whose choices
    // switch at ties, keep separate roundings.
#pragma clang fp contract(fast)
    T t0 = um2 - T(2) * um1 + u0, t1 = um2 - T(4) * um1 + T(3) * u0;
    T beta0 = T(13.0 / 12.0) * (t0 * t0) + T(1.0 / 4.0) * (t1 * t1);
    T t2 = um1 - T(2) * u0 + up1, t3 = um1 - up1;
    T beta1 = T(13.0 / 12.0) * (t2 * t2) + T(1.0 / 4.0) * (t3 * t3);
    T t4 = u0 - T(2) * up1 + up2, t5 = T(3) * u0 - T(4) * up1 + up2;
    T beta2 = T(13.0 / 12.0) * (t4 * t4) + T(1.0 / 4.0) * (t5 * t5);
    T e0 = T(1e-6) + beta0, e1 = T(1e-6) + beta1, e2 = T(1e-6) + beta2;
    e0 = e0 * e0; e1 = e1 * e1; e2 = e2 * e2;
    T P0, P1, P2;
    if (std::is_same<T, float>::value) { P0 = rcp_t(e0); P1 = rcp_t(e1); P2 = rcp_t(e2); }
    else { P0 = e1 * e2; P1 = e0 * e2; P2 = e0 * e1; }
    T q0 = T(-1.0 / 6.0) * um2 + T(5.0 / 6.0) * um1 + T(1.0 / 3.0) * u0;
    T q1 = T(1.0 / 3.0) * um1 + T(5.0 / 6.0) * u0 - T(1.0 / 6.0) * up1;
    T q2 = T(11.0 / 6.0) * u0 - T(7.0 / 6.0) * up1 + T(1.0 / 3.0) * up2;
    T a0 = T(0.3) * P0, a1 = T(0.6) * P1, a2 = T(0.1) * P2;
    const T numL = a0 * q0 + a1 * q1 + a2 * q2, denL = a0 + a1 + a2;
    q0 = T(1.0 / 3.0) * um2 - T(7.0 / 6.0) * um1 + T(11.0 / 6.0) * u0;
    q1 = T(-1.0 / 6.0) * um1 + T(5.0 / 6.0) * u0 + T(1.0 / 3.0) * up1;
    q2 = T(1.0 / 3.0) * u0 + T(5.0 / 6.0) * up1 - T(1.0 / 6.0) * up2;
    a0 = T(0.1) * P0; a1 = T(0.6) * P1; a2 = T(0.3) * P2;
    const T numR = a0 * q0 + a1 * q1 + a2 * q2, denR = a0 + a1 + a2;
    if (std::is_same<T, float>::value) {
      wL = numL * rcp_t(denL);
      wR = numR * rcp_t(denR);
    } else {                    // one division for both faces (denL denR stays far inside the FP64 range)
      const T inv = T(1) / (denL * denR);
      wL = numL * (denR * inv);
      wR = numR * (denL * inv);
    }
  }
}

// Rusanov-type flux with GLF dissipation (flux.cpp:78-163, 202-243, 329-367; state.cpp:27-44, 86-106,
// 272-346).  kind 0: convective only; kind 1: convective + magnetic pressure-formula terms; kind 2:
// energy formula — convective without its energy part, magnetic terms with the (2 - gamma) B^2/2
// normal-momentum and v_n B^2/2 energy parts, and -(gamma - 1) |m|^2 / (2 rho) on the normal momentum.
template <class T>
__device__ __forceinline__ void gl_flux(int dir, int kind, T gam, const T* QL, const T* QR, T s_max, T* F) {
  // Multiply-adds are fused (a smooth function of the face states: no branch depends on its rounding).
#pragma clang fp contract(fast)
  T fl[5], fr[5];
  const T* Qs2[2] = {QL, QR};
  T* fo[2] = {fl, fr};
  for (int side = 0; side < 2; side++) {
    const T* Q = Qs2[side];
    T* f = fo[side];
    const T ir = rcp_t(Q[0]);
    T rho = Q[0], vx = Q[1] * ir, vy = Q[2] * ir, vz = Q[3] * ir;
    T kin = T(0.5) * rho * (vx * vx + vy * vy + vz * vz);
    T vn = dir == 0 ? vx : vy;
    f[0] = rho * vn;
    f[1] = dir == 0 ? rho * vx * vx : rho * vx * vy;
    f[2] = dir == 0 ? rho * vy * vx : rho * vy * vy;
    f[3] = rho * vz * vn;
    f[4] = kind == 2 ? T(0) : kin * vn;
    if (kind == 2) {
      T B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
      T vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) * ir;
      T fk = -(gam - T(1)) * T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * ir;
      if (dir == 0) {
        f[1] = f[1] + ((T(2) - gam) * B2 * T(0.5) - Q[5] * Q[5]) + fk;
        f[2] = f[2] + (-Q[5] * Q[6]);
        f[3] = f[3] + (-Q[5] * Q[7]);
        f[4] = f[4] + (vx * B2 * T(0.5) - Q[5] * vB);
      } else {
        f[1] = f[1] + (-Q[5] * Q[6]);
        f[2] = f[2] + ((T(2) - gam) * B2 * T(0.5) - Q[6] * Q[6]) + fk;
        f[3] = f[3] + (-Q[6] * Q[7]);
        f[4] = f[4] + (vy * B2 * T(0.5) - Q[6] * vB);
      }
    }
    if (kind == 1) {
      T B2 = Q[5] * Q[5] + Q[6] * Q[6] + Q[7] * Q[7];
      T vB = (Q[1] * Q[5] + Q[2] * Q[6] + Q[3] * Q[7]) * ir;
      if (dir == 0) {
        f[1] = f[1] + (B2 * T(0.5) - Q[5] * Q[5]);
        f[2] = f[2] + (-Q[5] * Q[6]);
        f[3] = f[3] + (-Q[5] * Q[7]);
        f[4] = f[4] + (vx * B2 - Q[5] * vB);
      } else {
        f[1] = f[1] + (-Q[5] * Q[6]);
        f[2] = f[2] + (B2 * T(0.5) - Q[6] * Q[6]);
        f[3] = f[3] + (-Q[6] * Q[7]);
        f[4] = f[4] + (vy * B2 - Q[6] * vB);
      }
    }
  }
  for (int k = 0; k < 5; k++) F[k] = T(0.5) * (fl[k] + fr[k]) - T(0.5) * s_max * (QR[k] - QL[k]);
}

// RusanovC_Update (kind 0, flux.cpp:409-455), Pressure_flux_Update (kind 1, flux.cpp:290-326) and
// Energy_flux_Update (kind 2, flux.cpp:165-200) fused with TVDReconstruction on 32x8 cell tiles.  Per direction: every cell of the tile
// (plus one halo cell each side) is reconstructed once into LDS, every face flux is formed
// once into LDS, then each cell applies its flux difference.  Face states of a cell c are
// the reconstruction at the ghost-map image of c, as the reference's bound_cond_u on w_*
// produces.  kind 0 writes u[0..3] (the energy slot of the convective predictor is never
// read); kinds 1 and 2 write u[1..4] (their density equals the convective one bit for bit).
// The tile is EX_TX x EX_TY cells, one thread per cell (16x16, 32x8 and 64x4 with 256 threads, 32x16
// and 64x8 with 512, chosen by the autotuner); the tile shape changes which lane forms which face
// value, not the value.  Larger tiles reconstruct fewer frame cells and form fewer tile-edge faces
// twice per cell of the grid.
template <int EX_TX, int EX_TY> struct ExTile {
  static constexpr int ROWS = EX_TY, NT = EX_TX * EX_TY;
  static constexpr int NWX = (EX_TX + 2) * ROWS, NWY = EX_TX * (ROWS + 2);
  static constexpr int NFX = (EX_TX + 1) * ROWS, NFY = EX_TX * (ROWS + 1);
  static constexpr int NW = NWX > NWY ? NWX : NWY;     // cells reconstructed per direction, at most
  static constexpr int NF = NFX > NFY ? NFX : NFY;     // faces per direction, at most
};

// Conservative face states of one cell from the primitive reconstruction (ReconState PrimState):
// con -> pri at the five stencil cells q[(o - 2) step], reconstruct each primitive, pri -> con of both
// face states.  NC = 4 forms rho and the momenta only (rho v needs no EOS); NC = 8 the whole state.
// Loop bounds are compile-time so the stencil stays in registers.
// Stencil cells o < raw are ghost cells of the jet inflow strip: the reference fills the primitive grid's
// ghosts with bound_cond_u outside g_bc_prim, so they hold the conservative inflow state, taken as is.
// The ghosts of a Dirichlet side of the stencil (`amb`) hold, for the same reason, the conservative ambient
// state: rho_amb, 0, 0, 0, e_amb + |B|^2/2 and the boundary cell's B.
// Held as scalars and formed per component on demand, so no array in it is indexed at run time.
struct DirAmb {
  int lo, hi;                  // stencil cells beyond the low / high boundary
  double rho, e, lb[3], hb[3];  // rho_amb, e_amb, and the B of the low / high boundary cell
  __device__ __forceinline__ double val(bool low, int k) const {
    const double b5 = low ? lb[0] : hb[0], b6 = low ? lb[1] : hb[1], b7 = low ? lb[2] : hb[2];
    return k == 0 ? rho : k < 4 ? 0.0 : k == 4 ? e + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7) : k == 5 ? b5 : k == 6 ? b6 : b7;
  }
};
template <int NC>
__device__ __forceinline__ void prim_faces(const Phys& P, int central, const double* __restrict__ q, long long PS,
                                           long long step, int raw, const DirAmb& amb, double* Lc, double* Rc, int* ef) {
  double s[5][NC];
#pragma unroll
  for (int o = 0; o < 5; o++) {
    double c[NVAR];
#pragma unroll
    for (int k = 0; k < NC; k++) c[k] = q[k * PS + (o - 2) * step];
    if (o < amb.lo || o >= 5 - amb.hi) {
#pragma unroll
      for (int k = 0; k < NC; k++) s[o][k] = amb.val(o < amb.lo, k);
    } else if (o < raw) {
#pragma unroll
      for (int k = 0; k < NC; k++) s[o][k] = c[k];
    } else if constexpr (NC == 4) {
      s[o][0] = c[0];
#pragma unroll
      for (int k = 1; k < 4; k++) s[o][k] = c[k] / c[0];
    } else {
      con2pri(P, s[o], c, ef);
    }
  }
  double Lp[NC], Rp[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    double a[5];
#pragma unroll
    for (int o = 0; o < 5; o++) a[o] = s[o][k];
    recon5(P.recon, central, a, Lp[k], Rp[k]);
  }
  if constexpr (NC == 4) {
    Lc[0] = Lp[0]; Rc[0] = Rp[0];
#pragma unroll
    for (int k = 1; k < 4; k++) { Lc[k] = Lp[0] * Lp[k]; Rc[k] = Rp[0] * Rp[k]; }
  } else {
    pri2con(P, Lc, Lp, ef);
    pri2con(P, Rc, Rp, ef);
  }
}

// One direction of the fused update: reconstruct, form the face fluxes, apply their difference.
// T is the arithmetic type of the reconstruction and the fluxes (float in the mixed form); the
// conservative update u = Qs - adt div F is always formed in FP64.  PRIM (FP64 only) reconstructs
// the primitives of the five stencil cells and converts the two face states back (ReconState
// PrimState, reconstruction.cpp:11-24, 272-281); kind 0 needs rho and v only, kinds 1 and 2 the
// whole state, as the face energy rho e(rho, p) + kinetic + |B|^2/2 uses the reconstructed p and B.
template <int DIR, int EX_TX, int EX_TYT, class T, bool PRIM, bool BND>
__device__ __forceinline__ void explicit_dir(const Dims& d, const Phys& P, int kind, const double* __restrict__ Qe,
                                             const double* __restrict__ Qs, double* __restrict__ u, double h, T s_max,
                                             int i0, int j0, T* wl, T* wr, T* F, double* acc, int* eflag) {
  constexpr int EX_TY = ExTile<EX_TX, EX_TYT>::ROWS, EX_NW = ExTile<EX_TX, EX_TYT>::NW, EX_NF = ExTile<EX_TX, EX_TYT>::NF;
  constexpr int NT = ExTile<EX_TX, EX_TYT>::NT;
  constexpr int W = DIR == 0 ? EX_TX + 2 : EX_TX;                  // reconstructed cells per row
  constexpr int NITEMS = DIR == 0 ? (EX_TX + 2) * EX_TY : EX_TX * (EX_TY + 2);
  constexpr int NF = DIR == 0 ? (EX_TX + 1) * EX_TY : EX_TX * (EX_TY + 1);
  constexpr int FW = DIR == 0 ? EX_TX + 1 : EX_TX;                 // faces per row
  const int nrec = kind == 0 ? 4 : 5;       // kind 0 fluxes need rho and momenta only
  const int tx = threadIdx.x % EX_TX, ty = threadIdx.x / EX_TX;
  const int ci = i0 + tx, cj = j0 + ty;
  const bool own = ci < d.nx + GC && cj < d.ny + GC;
  const int k0 = kind == 0 ? 0 : 1, k1 = kind == 0 ? 4 : 5;
  __syncthreads();
  for (int it = threadIdx.x; it < NITEMS; it += blockDim.x) {
    int lx = it % W, ly = it / W;
    int gi = DIR == 0 ? i0 + lx - 1 : i0 + lx;
    int gj = DIR == 0 ? j0 + ly : j0 + ly - 1;
    if (gi > d.nx + GC || gj > d.ny + GC) continue;      // beyond a partial tile: unused
    if (BND && DIR == 1 && gj < GC && jet_inlet(d, gi)) {  // bound_cond_u of the face states: the jet inflow
      for (int k = 0; k < 5; k++)
        if (k < nrec) { wl[k * EX_NW + it] = (T)jet_state(d, 0, k); wr[k * EX_NW + it] = (T)jet_state(d, 0, k); }
      continue;
    }
    int mi = DIR == 0 ? map_x(d, gi) : gi;
    int mj = DIR == 0 ? gj : map_y(d, gj);
    const long long base = pidx(d, mi, mj);
    const long long step = DIR == 0 ? 1 : d.NX;
    // Dirichlet sides (boundary.cpp:56-76, 128-147): TVDReconstruction's ghost fill puts the conservative
    // ambient state, with the boundary cell's B, in the stencil cells beyond the boundary, and its face-state
    // fill puts the ambient state, with E from the face state's own B, in the ghost items
    DirAmb amb;
    amb.lo = amb.hi = 0;
    amb.rho = d.rho_amb; amb.e = d.e_amb;
    amb.lb[0] = amb.lb[1] = amb.lb[2] = amb.hb[0] = amb.hb[1] = amb.hb[2] = 0.0;
    const bool dpass = BND && (DIR == 0 ? d.dirx : d.diry);
    if (dpass) {
      const int t0 = DIR == 0 ? mi : mj, n = DIR == 0 ? d.nx : d.ny;
      amb.lo = max(0, GC + 2 - t0);
      amb.hi = max(0, t0 + 2 - (n + GC - 1));
      if (amb.lo) {
        const long long cb = DIR == 0 ? pidx(d, GC, mj) : pidx(d, mi, GC);
        amb.lb[0] = Qe[5 * d.PS + cb]; amb.lb[1] = Qe[6 * d.PS + cb]; amb.lb[2] = Qe[7 * d.PS + cb];
      }
      if (amb.hi) {
        const long long cb = DIR == 0 ? pidx(d, n + GC - 1, mj) : pidx(d, mi, n + GC - 1);
        amb.hb[0] = Qe[5 * d.PS + cb]; amb.hb[1] = Qe[6 * d.PS + cb]; amb.hb[2] = Qe[7 * d.PS + cb];
      }
    }
    const bool dgi = dpass && (DIR == 0 ? (gi < GC || gi > d.nx + GC - 1) : (gj < GC || gj > d.ny + GC - 1));
    if (dgi) {
      T Lr[5] = {(T)d.rho_amb, T(0), T(0), T(0), T(0)}, Rr[5] = {(T)d.rho_amb, T(0), T(0), T(0), T(0)};
      if (nrec == 5) {                                    // E of each face state from its own reconstructed B
        double bl[3], br[3];
        for (int c = 0; c < 3; c++) {
          double a[5];
#pragma unroll
          for (int o = 0; o < 5; o++) a[o] = o < amb.lo ? amb.val(true, 5 + c) : (o >= 5 - amb.hi ? amb.val(false, 5 + c) : Qe[(5 + c) * d.PS + base + (o - 2) * step]);
          recon5(P.recon, DIR == 1 ? P.central_y : 0, a, bl[c], br[c]);
        }
        Lr[4] = (T)(d.e_amb + 0.5 * (bl[0] * bl[0] + bl[1] * bl[1] + bl[2] * bl[2]));
        Rr[4] = (T)(d.e_amb + 0.5 * (br[0] * br[0] + br[1] * br[1] + br[2] * br[2]));
      }
      for (int k = 0; k < 5; k++)
        if (k < nrec) { wl[k * EX_NW + it] = Lr[k]; wr[k * EX_NW + it] = Rr[k]; }
      continue;
    }
    if constexpr (PRIM) {
      static_assert(std::is_same<T, double>::value, "primitive reconstruction is FP64 only");
      int ef = 0;
      double Lc[NVAR], Rc[NVAR];
      const int raw = (BND && DIR == 1 && jet_inlet(d, gi)) ? GC + 2 - gj : 0;   // stencil rows below GC in the strip
      if (kind == 0) prim_faces<4>(P, DIR == 1 ? P.central_y : 0, Qe + base, d.PS, step, raw, amb, Lc, Rc, &ef);
      else prim_faces<NVAR>(P, DIR == 1 ? P.central_y : 0, Qe + base, d.PS, step, raw, amb, Lc, Rc, &ef);
      for (int k = 0; k < 5; k++)
        if (k < nrec) {
          const double rf = BND ? refl_factor(d, gi, gj, k != 0 && k != 4) : 1.0;   // bound_cond_u of the face states
          wl[k * EX_NW + it] = rf * Lc[k]; wr[k * EX_NW + it] = rf * Rc[k];
        }
      if (ef) atomicOr(eflag, 1);
    } else {
      // all stencil loads are issued before any reconstruction (the fifth variable is loaded for
      // kind 0 too, where it goes unused), so their latencies overlap
      double s64[5][5];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const double* q = Qe + k * d.PS + base;
#pragma unroll
        for (int o = 0; o < 5; o++) {
          if constexpr (BND) s64[k][o] = o < amb.lo ? amb.val(true, k) : (o >= 5 - amb.hi ? amb.val(false, k) : q[(o - 2) * step]);
          else s64[k][o] = q[(o - 2) * step];
        }
      }
#pragma unroll
      for (int k = 0; k < 5; k++) {
        if (k >= nrec) continue;
        const double* a64 = s64[k];
        T a[5];
#pragma unroll
        for (int o = 0; o < 5; o++) a[o] = (T)a64[o];
        const int sel = std::is_same<T, float>::value && P.recon == IMEX_RECON_ENO3 ? eno3_stencil(a64) : -1;
        T L, R;
        recon5(P.recon, DIR == 1 ? P.central_y : 0, a, L, R, sel);
        if constexpr (BND) {
          const T rf = (T)refl_factor(d, gi, gj, k != 0 && k != 4);      // bound_cond_u of the face states
          L = rf * L; R = rf * R;
        }
        wl[k * EX_NW + it] = L;
        wr[k * EX_NW + it] = R;
      }
    }
  }
  __syncthreads();
  // at most two faces per thread; their fluxes are held in registers until every face state has
  // been read, then written over the face states (F aliases wl: a smaller LDS footprint, more
  // blocks per CU)
  static_assert(EX_NF <= EX_NW && NF <= 2 * NT, "flux tile layout");
  T Fo[2][5];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int f = threadIdx.x + r * NT;
    if (f >= NF) continue;
    int fx = f % FW, fy = f / FW;
    // face between local cells (left, right) along DIR; item indices into wl/wr
    int itL, itR, gli, glj, gri, grj;
    if (DIR == 0) { itL = fy * W + fx; itR = fy * W + fx + 1; gli = i0 + fx - 1; gri = i0 + fx; glj = grj = j0 + fy; }
    else { itL = fy * W + fx; itR = (fy + 1) * W + fx; gli = gri = i0 + fx; glj = j0 + fy - 1; grj = j0 + fy; }
    if (gri > d.nx + GC || grj > d.ny + GC) continue;
    T QL[NVAR], QR[NVAR];
    for (int k = 0; k < 5; k++) {
      QL[k] = k < nrec ? wr[k * EX_NW + itL] : T(0);
      QR[k] = k < nrec ? wl[k * EX_NW + itR] : T(0);
    }
    if (kind != 0) {
      for (int k = 5; k < 8; k++) { QL[k] = (T)u[k * d.PS + pidx(d, gli, glj)]; QR[k] = (T)u[k * d.PS + pidx(d, gri, grj)]; }
    } else {
      for (int k = 5; k < 8; k++) { QL[k] = T(0); QR[k] = T(0); }
    }
    gl_flux(DIR, kind, (T)P.gamma, QL, QR, s_max, Fo[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int f = threadIdx.x + r * NT;
    if (f >= NF) continue;
    for (int k = 0; k < 5; k++) F[k * EX_NF + f] = Fo[r][k];
  }
  __syncthreads();
  if (own) {
#pragma clang fp contract(fast)
    long long c = pidx(d, ci, cj);
    if (DIR == 0) {
      int fm = ty * FW + tx, fp = fm + 1;
      for (int k = k0; k < k1; k++) acc[k] = Qs[k * d.PS + c] - h / d.dx * ((double)F[k * EX_NF + fp] - (double)F[k * EX_NF + fm]);
    } else {
      int fm = ty * FW + tx, fp = fm + FW;
      for (int k = k0; k < k1; k++) u[k * d.PS + c] = acc[k] - h / d.dy * ((double)F[k * EX_NF + fp] - (double)F[k * EX_NF + fm]);
    }
  }
}

// BND: the boundary treatments other than periodic and outflow (jet inflow, reflective, Dirichlet) are
// compiled in; without them the kernel carries none of their code.
template <int EX_TX, int EX_TYT, class T, bool PRIM, bool BND>
__global__ void __launch_bounds__(EX_TX * EX_TYT) k_explicit(Phys P, const Inst* __restrict__ IT, BMap bm, int kind, int iQe, int iQs, int iu) {
  constexpr int EX_TY = ExTile<EX_TX, EX_TYT>::ROWS, EX_NW = ExTile<EX_TX, EX_TYT>::NW;
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims d = I_.d;
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ Qs = gmem(I_.buf[iQs]);
  double* __restrict__ u = gmem(I_.buf[iu]);
  const double* S = gmem(I_.S);

  __shared__ T wl[5 * EX_NW], wr[5 * EX_NW];
  T* F = wl;                  // face fluxes, written once the face states have been read
  const int i0 = GC + bx_ * EX_TX, j0 = GC + by_ * EX_TY;
  const double h = S[S_ADT];
  const T s_max = (T)S[S_AMAX];
  int* eflag = gmem(I_.eflag);
  double acc[5];
  explicit_dir<0, EX_TX, EX_TYT, T, PRIM, BND>(d, P, kind, Qe, Qs, u, h, s_max, i0, j0, wl, wr, F, acc, eflag);
  explicit_dir<1, EX_TX, EX_TYT, T, PRIM, BND>(d, P, kind, Qe, Qs, u, h, s_max, i0, j0, wl, wr, F, acc, eflag);
}

// ================================================================ magnetic subsystem
// RHS of MagSubsystem_UpdateB (imex_subsystems.cpp:549-613), initial guess x0 = B(Qe),
// and the enthalpy ratio hhat = calenthalpy(Qe) rho^e / rho (used by the pressure subsystem).
template <bool RK>
__global__ void __launch_bounds__(256) k_mag_rhs(Phys P_, const Inst* __restrict__ IT, BMap bm, int iu, int iQe, int iQs,
                                                 int ib, int ix0) {
  BATCH_BLOCK(bm);
  Phys P = P_;
  if constexpr (!RK) P.eos = IMEX_EOS_IDEAL;
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ Qs = gmem(I_.buf[iQs]);
  const double* __restrict__ pe = gmem(I_.buf[B_PE]);
  double* __restrict__ b = gmem(I_.buf[ib]);
  double* __restrict__ x0 = gmem(I_.buf[ix0]);
  double* __restrict__ hhat = gmem(I_.buf[B_HHAT]);
  const double* S = gmem(I_.S);
  int* eflag = gmem(I_.eflag);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double h = S[S_ADT], dx = d.dx, dy = d.dy;
  auto U = [&](int a, int bb, int k) { return u[k * d.PS + pidx(d, a, bb)]; };
  auto Q = [&](int a, int bb, int k) { return Qe[k * d.PS + pidx(d, a, bb)]; };
  auto PE = [&](int a, int bb) { return pe[pidx(d, a, bb)]; };
  double rho_p0 = 0.5 * (U(i, j, 0) + U(i + 1, j, 0)), rho_m0 = 0.5 * (U(i, j, 0) + U(i - 1, j, 0));
  double rho_0p = 0.5 * (U(i, j, 0) + U(i, j + 1, 0)), rho_0m = 0.5 * (U(i, j, 0) + U(i, j - 1, 0));
  double m[3][4], n[3][4];
  for (int c = 0; c < 3; c++) {
    m[c][0] = 0.5 * (U(i, j, 1 + c) + U(i + 1, j, 1 + c)); m[c][1] = 0.5 * (U(i, j, 1 + c) + U(i - 1, j, 1 + c));
    m[c][2] = 0.5 * (U(i, j, 1 + c) + U(i, j + 1, 1 + c)); m[c][3] = 0.5 * (U(i, j, 1 + c) + U(i, j - 1, 1 + c));
    n[c][0] = 0.5 * (Q(i, j, 5 + c) + Q(i + 1, j, 5 + c)); n[c][1] = 0.5 * (Q(i, j, 5 + c) + Q(i - 1, j, 5 + c));
    n[c][2] = 0.5 * (Q(i, j, 5 + c) + Q(i, j + 1, 5 + c)); n[c][3] = 0.5 * (Q(i, j, 5 + c) + Q(i, j - 1, 5 + c));
  }
  double p00 = PE(i, j), pp0 = PE(i + 1, j), pm0 = PE(i - 1, j), p0p = PE(i, j + 1), p0m = PE(i, j - 1);
  double p_pp = 0.25 * (p00 + pp0 + p0p + PE(i + 1, j + 1));
  double p_pm = 0.25 * (p00 + pp0 + p0m + PE(i + 1, j - 1));
  double p_mp = 0.25 * (p00 + pm0 + p0p + PE(i - 1, j + 1));
  double p_mm = 0.25 * (p00 + pm0 + p0m + PE(i - 1, j - 1));
  long long c0 = pidx(d, i, j);
  double Bx_star = Qs[5 * d.PS + c0] - (h / dy) * ((n[0][2] / rho_0p * (m[1][2] - (h / dy) * (p0p - p00)) - n[1][2] / rho_0p * (m[0][2] - h / dx * (p_pp - p_mp)))
                                                 - (n[0][3] / rho_0m * (m[1][3] - (h / dy) * (p00 - p0m)) - n[1][3] / rho_0m * (m[0][3] - h / dx * (p_pm - p_mm))));
  double By_star = Qs[6 * d.PS + c0] - (h / dx) * ((n[1][0] / rho_p0 * (m[0][0] - (h / dx) * (pp0 - p00)) - n[0][0] / rho_p0 * (m[1][0] - h / dy * (p_pp - p_pm)))
                                                 - (n[1][1] / rho_m0 * (m[0][1] - (h / dx) * (p00 - pm0)) - n[0][1] / rho_m0 * (m[1][1] - h / dy * (p_mp - p_mm))));
  double Bz_1 = ((n[2][0] / rho_p0 * (m[0][0] - h / dx * (pp0 - p00)) - n[0][0] / rho_p0 * (m[2][0]))
               - (n[2][1] / rho_m0 * (m[0][1] - h / dx * (p00 - pm0)) - n[0][1] / rho_m0 * (m[2][1])));
  double Bz_2 = ((n[2][2] / rho_0p * (m[1][2] - h / dy * (p0p - p00)) - n[1][2] / rho_0p * (m[2][2]))
               - (n[2][3] / rho_0m * (m[1][3] - h / dy * (p00 - p0m)) - n[1][3] / rho_0m * (m[2][3])));
  double Bz_star = Qs[7 * d.PS + c0] - (h / dx) * Bz_1 - (h / dy) * Bz_2;
  int idx = (i - GC) + d.nx * (j - GC);
  b[idx] = Bx_star; b[idx + d.N] = By_star; b[idx + 2 * d.N] = Bz_star;
  x0[idx] = Q(i, j, 5); x0[idx + d.N] = Q(i, j, 6); x0[idx + 2 * d.N] = Q(i, j, 7);
  double q[NVAR];
  for (int k = 0; k < NVAR; k++) q[k] = Qe[k * d.PS + c0];
  int ef = 0;
  hhat[c0] = enthalpy_of(P, q, &ef) * q[0] / U(i, j, 0);
  if (j == GC && jet_inlet(d, i)) {   // the jet strip's ghost holds the inflow state: its ratio is formed from it
    const long long cg = pidx(d, i, j - 1);
    for (int k = 0; k < NVAR; k++) q[k] = Qe[k * d.PS + cg];
    hhat[cg] = enthalpy_of(P, q, &ef) * q[0] / U(i, j - 1, 0);
  }
  if (ef) atomicOr(eflag, 1);
}

// u_con[5..7] <- B~ at every padded cell through the M_B map (imex_subsystems.cpp:618-625).
__global__ void k_mag_writeback(const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ x = gmem(I_.buf[B_XK]);
  double* __restrict__ u = gmem(I_.buf[iu]);

  long long total = (long long)d.NX * d.NY;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= total) return;
  int i = (int)(c % d.NX), j = (int)(c / d.NX);
  if (j < GC && jet_inlet(d, i)) {                    // the bound_cond_u that follows: the jet inflow's B
    u[5 * d.PS + c] = 0.0; u[6 * d.PS + c] = d.jet_B0; u[7 * d.PS + c] = 0.0;
    return;
  }
  int id = kidx(d, i, j);
  const double rf = refl_factor(d, i, j, true);        // the bound_cond_u that follows: B negated
  const double b5 = rf * x[id], b6 = rf * x[id + d.N], b7 = rf * x[id + 2 * d.N];
  u[5 * d.PS + c] = b5;
  u[6 * d.PS + c] = b6;
  u[7 * d.PS + c] = b7;
  if (dir_ghost(d, i, j)) u[4 * d.PS + c] = d.e_amb + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);   // and E of the ambient
}

// CollectMagFluxes (ct_flux.cpp:5-130): Fx at x faces (i+1/2, j), Fy at y faces (i, j+1/2),
// evaluated at the ghost-map image of every padded index (= compute on [1, n+2g-1) then bound_cond_array).
// The ten input planes (rho, rho v_x, rho v_y and B~ of u; B of Qe; p(Qe)) of a 64x8 tile with a
// one-cell frame are staged in LDS (each of the 64x4 threads forms two positions, rows ty and
// ty + 4); positions whose ghost-map image is another cell (the ghost frame) read their stencil
// from global memory.  The formula is written once for both readers.
#define EMF_NF 10
#define EMF_TY 8
struct EmfSmem { double v[EMF_NF][EMF_TY + 2][66]; };
// LDS plane of each input: u slots 0, 1, 2, 5, 6, 7, Qe slots 5, 6, 7, p(Qe)
__device__ __forceinline__ int emf_plane_u(int k) { return k < 3 ? k : k - 2; }
template <class RU, class RQ, class RP>
__device__ __forceinline__ void emf_fluxes(const RU& U, const RQ& Q, const RP& PE, int i, int j, double h, double dx, double dy,
                                           double nu_x, double nu_y, double& fx, double& fy) {
  double rho_p0 = 0.5 * (U(i, j, 0) + U(i + 1, j, 0)), rho_0p = 0.5 * (U(i, j, 0) + U(i, j + 1, 0));
  // one division per face density; the quotients below multiply by its reciprocal
  const double irp0 = 1.0 / rho_p0, ir0p = 1.0 / rho_0p;
  double rvx_p0 = 0.5 * (U(i, j, 1) + U(i + 1, j, 1)), rvy_p0 = 0.5 * (U(i, j, 2) + U(i + 1, j, 2));
  double rvx_0p = 0.5 * (U(i, j, 1) + U(i, j + 1, 1)), rvy_0p = 0.5 * (U(i, j, 2) + U(i, j + 1, 2));
  double p00 = PE(i, j);
  double p_pp = 0.25 * (p00 + PE(i + 1, j) + PE(i, j + 1) + PE(i + 1, j + 1));
  double p_pm = 0.25 * (p00 + PE(i + 1, j) + PE(i, j - 1) + PE(i + 1, j - 1));
  double p_mp = 0.25 * (p00 + PE(i - 1, j) + PE(i, j + 1) + PE(i - 1, j + 1));
  double Bxn_p0 = 0.5 * (Q(i, j, 5) + Q(i + 1, j, 5)), Byn_p0 = 0.5 * (Q(i, j, 6) + Q(i + 1, j, 6));
  double Bxn_0p = 0.5 * (Q(i, j, 5) + Q(i, j + 1, 5)), Byn_0p = 0.5 * (Q(i, j, 6) + Q(i, j + 1, 6));
  double n_pp[3], n_pm[3], n_mp[3], b_pp[3], b_pm[3], b_mp[3];
  for (int c = 0; c < 3; c++) {
    int k = 5 + c;
    double q00 = Q(i, j, k), qp0 = Q(i + 1, j, k), qm0 = Q(i - 1, j, k), q0p = Q(i, j + 1, k), q0m = Q(i, j - 1, k);
    n_pp[c] = 0.25 * (q00 + qp0 + q0p + Q(i + 1, j + 1, k));
    n_pm[c] = 0.25 * (q00 + qp0 + q0m + Q(i + 1, j - 1, k));
    n_mp[c] = 0.25 * (q00 + qm0 + q0p + Q(i - 1, j + 1, k));
    double u00 = U(i, j, k), up0 = U(i + 1, j, k), um0 = U(i - 1, j, k), u0p = U(i, j + 1, k), u0m = U(i, j - 1, k);
    b_pp[c] = 0.25 * (u00 + up0 + u0p + U(i + 1, j + 1, k));
    b_pm[c] = 0.25 * (u00 + up0 + u0m + U(i + 1, j - 1, k));
    b_mp[c] = 0.25 * (u00 + um0 + u0p + U(i - 1, j + 1, k));
  }
  auto PM = [&](int a, int b) { return 0.5 * (Q(a, b, 5) * U(a, b, 5) + Q(a, b, 6) * U(a, b, 6) + Q(a, b, 7) * U(a, b, 7)); };
  auto PC = [&](const double* nn, const double* bb, int mm) { return 0.5 * (nn[0] * bb[0] + nn[1] * bb[1] + nn[2] * bb[2]) - nn[mm] * bb[mm]; };
  double Fn_x = Byn_p0 * irp0 * (rvx_p0 - (h / dx) * (PE(i + 1, j) - p00)) - Bxn_p0 * irp0 * (rvy_p0 - (h / dy) * (p_pp - p_pm));
  double By_1 = Byn_p0 * irp0 * (-1.0 / dx * ((PM(i + 1, j) - Q(i + 1, j, 5) * U(i + 1, j, 5)) - (PM(i, j) - Q(i, j, 5) * U(i, j, 5))) + 1.0 / dy * (n_pp[0] * b_pp[1] - n_pm[0] * b_pm[1]));
  double By_2 = Bxn_p0 * irp0 * (-1.0 / dx * (Q(i + 1, j, 5) * U(i + 1, j, 6) - Q(i, j, 5) * U(i, j, 6)) + 1.0 / dy * (PC(n_pp, b_pp, 1) - PC(n_pm, b_pm, 1)));
  double Fdiff_x = -nu_x * (U(i + 1, j, 6) - U(i, j, 6)) * (1.0 / dx);
  fx = Fn_x + h * (By_1 + By_2) + Fdiff_x;
  double Fn_y = Bxn_0p * ir0p * (rvy_0p - (h / dy) * (PE(i, j + 1) - p00)) - Byn_0p * ir0p * (rvx_0p - (h / dx) * (p_pp - p_mp));
  double Bx_1 = Bxn_0p * ir0p * (1.0 / dx * (n_pp[0] * b_pp[1] - n_mp[0] * b_mp[1]) - 1.0 / dy * ((PM(i, j + 1) - Q(i, j + 1, 6) * U(i, j + 1, 6)) - (PM(i, j) - Q(i, j, 6) * U(i, j, 6))));
  double Bx_2 = Byn_0p * ir0p * (1.0 / dx * (PC(n_pp, b_pp, 0) - PC(n_mp, b_mp, 0)) - 1.0 / dy * (Q(i, j + 1, 6) * U(i, j + 1, 5) - Q(i, j, 6) * U(i, j, 5)));
  double Fdiff_y = -nu_y * (U(i, j + 1, 5) - U(i, j, 5)) * (1.0 / dy);
  fy = Fn_y + h * (Bx_1 + Bx_2) + Fdiff_y;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_emf_flux(const Inst* __restrict__ IT, BMap bm, int iu, int iQe) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ pe = gmem(I_.buf[B_PE]);
  double* __restrict__ Fx = gmem(I_.buf[B_FX]);
  double* __restrict__ Fy = gmem(I_.buf[B_FY]);
  const double* S = gmem(I_.S);
  __shared__ EmfSmem sm;

  // stage the tile [i0 - 1, i0 + 64] x [j0 - 1, j0 + EMF_TY] (clamped to the padded grid; clamped
  // entries are read only by positions that take the global path)
  const int i0 = bx_ * 64, j0 = by_ * EMF_TY;
  const int t = threadIdx.x + 64 * threadIdx.y;
  // every load of the thread's (up to three) staged positions is issued before the first LDS store
  constexpr int NSTG = 66 * (EMF_TY + 2), NPT = (NSTG + 255) / 256;
  double v[NPT][EMF_NF];
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int e = t + 256 * q;
    if (e >= NSTG) break;
    const int lx = e % 66, ly = e / 66;
    const int a = min(max(i0 + lx - 1, 0), d.NX - 1), b = min(max(j0 + ly - 1, 0), d.NY - 1);
    const long long c = pidx(d, a, b);
    v[q][0] = u[c]; v[q][1] = u[d.PS + c]; v[q][2] = u[2 * d.PS + c];
    v[q][3] = u[5 * d.PS + c]; v[q][4] = u[6 * d.PS + c]; v[q][5] = u[7 * d.PS + c];
    v[q][6] = Qe[5 * d.PS + c]; v[q][7] = Qe[6 * d.PS + c]; v[q][8] = Qe[7 * d.PS + c];
    v[q][9] = pe[c];
  }
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int e = t + 256 * q;
    if (e >= NSTG) break;
    const int lx = e % 66, ly = e / 66;
#pragma unroll
    for (int f = 0; f < EMF_NF; f++) sm.v[f][ly][lx] = v[q][f];
  }
  __syncthreads();

  const double h = S[S_ADT], dx = d.dx, dy = d.dy;
  const double nu_x = S[S_BMAXX] * dx, nu_y = S[S_BMAXY] * dy;
#pragma unroll 1
  for (int r = 0; r < EMF_TY; r += 4) {
    const int ii = i0 + threadIdx.x, jj = j0 + threadIdx.y + r;
    if (ii >= d.NX || jj >= d.NY) break;
    const int i = map_x(d, ii), j = map_y(d, jj);
    double fx, fy;
    if (i == ii && j == jj) {
      const int ox = i0 - 1, oy = j0 - 1;
      auto U = [&](int a, int b, int k) { return sm.v[emf_plane_u(k)][b - oy][a - ox]; };
      auto Q = [&](int a, int b, int k) { return sm.v[k + 1][b - oy][a - ox]; };
      auto PE = [&](int a, int b) { return sm.v[9][b - oy][a - ox]; };
      emf_fluxes(U, Q, PE, i, j, h, dx, dy, nu_x, nu_y, fx, fy);
    } else {
      auto U = [&](int a, int b, int k) { return u[k * d.PS + pidx(d, a, b)]; };
      auto Q = [&](int a, int b, int k) { return Qe[k * d.PS + pidx(d, a, b)]; };
      auto PE = [&](int a, int b) { return pe[pidx(d, a, b)]; };
      emf_fluxes(U, Q, PE, i, j, h, dx, dy, nu_x, nu_y, fx, fy);
    }
    const double rf = refl_factor(d, ii, jj, true);      // bound_cond_array of the flux arrays: reflective negation
    Fx[pidx(d, ii, jj)] = rf * fx;
    Fy[pidx(d, ii, jj)] = rf * fy;
  }
}

// ================================================================ pressure subsystem pointwise/5-point kernels
// RHS of PreSubsystem_Pressure (imex_subsystems.cpp:123-168 first pass, :226-270 second pass).
// pass 0 also stores rho v* (mxs, mys), the initial pressure p0 = p(Qe) and the Jacobi diagonal
// part Tdiag (implicit_operators.cpp:631-649).
// With the pressure floor active it also leaves, per block, min b and (Redlich-Kwong) max rho
// e(rho, p_floor) in partial2 for the Mbig finish of the Newton solve that follows
// (implicit_operators.cpp:670-686): the values k_newton_mbig would reduce from the same b.
__global__ void __launch_bounds__(256) k_p_rhs(Phys P, const Inst* __restrict__ IT, BMap bm, int pass, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ hhat = gmem(I_.buf[B_HHAT]);
  double* __restrict__ mxs = gmem(I_.buf[B_MXS]);
  double* __restrict__ mys = gmem(I_.buf[B_MYS]);
  const double* __restrict__ pe = gmem(I_.buf[B_PE]);
  double* __restrict__ b = gmem(I_.buf[B_PB]);
  double* __restrict__ p0 = gmem(I_.buf[B_P]);
  double* __restrict__ Tdiag = gmem(I_.buf[B_TDIAG]);
  const double* S = gmem(I_.S);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  const bool valid = i < d.nx + GC && j < d.ny + GC;
  double bmin = 1e300, vmax = -1e300;
  if (valid) {
    const double h = S[S_ADT], dx = d.dx, dy = d.dy;
    long long c = pidx(d, i, j);
    auto H = [&](int a, int bb) { return hhat[pidx(d, a, bb)]; };
    const double* MX = pass == 0 ? u + 1 * d.PS : mxs;
    const double* MY = pass == 0 ? u + 2 * d.PS : mys;
    auto mx = [&](int a, int bb) { return MX[pidx(d, a, bb)]; };
    auto my = [&](int a, int bb) { return MY[pidx(d, a, bb)]; };
    double u5 = u[5 * d.PS + c], u6 = u[6 * d.PS + c], u7 = u[7 * d.PS + c], rho = u[c];
    double B2 = u5 * u5 + u6 * u6 + u7 * u7;
    double h00 = H(i, j), hp0 = H(i + 1, j), hm0 = H(i - 1, j), h0p = H(i, j + 1), h0m = H(i, j - 1);
    double p_star = u[4 * d.PS + c] - 0.5 * B2;
    double u1 = u[1 * d.PS + c], u2 = u[2 * d.PS + c], u3 = u[3 * d.PS + c];
    double ke = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3) / rho;   // pass 0: rho v* == u; pass 1: updated momentum
    double RHS = p_star - ke - h * (1 / dx * (0.5 * (hp0 + h00) * 0.5 * (mx(i + 1, j) + mx(i, j)) - 0.5 * (hm0 + h00) * 0.5 * (mx(i - 1, j) + mx(i, j)))
                                   + (1 / dy * (0.5 * (h0p + h00) * 0.5 * (my(i, j + 1) + my(i, j)) - 0.5 * (h0m + h00) * 0.5 * (my(i, j - 1) + my(i, j)))));
    int idx = (i - GC) + d.nx * (j - GC);
    double bv = dir_bnd(d, i, j) ? d.p_amb : RHS;      // strong Dirichlet row (imex_subsystems.cpp:160-161, 263-264)
    // jet inflow strip: the Dirichlet lifting of the ghost pressure (imex_subsystems.cpp:163-165, 266-267)
    if (j == GC && jet_inlet(d, i)) bv += h * h / (dy * dy) * 0.5 * (h00 + h0m) * d.jet_p;
    b[idx] = bv;
    if (pass == 0) {
      mxs[c] = u1; mys[c] = u2;
      p0[idx] = pe[c];
      Tdiag[idx] = h * h * ((0.5 * (hp0 + h00) + 0.5 * (hm0 + h00)) / (dx * dx) + (0.5 * (h0p + h00) + 0.5 * (h0m + h00)) / (dy * dy));
    }
    if (P.floor_on) {
      bmin = bv;
      if (P.eos != IMEX_EOS_IDEAL) {
        int ef = 0;
        double v = Vbase(P, P.p_floor, rho, &ef);
        if (isfinite(v)) vmax = v;
        if (ef) atomicOr(gmem(I_.eflag), 1);
      }
    }
  }
  if (!P.floor_on) return;
  // min and max are order independent: the same Mbig as a reduction over the whole plane
  __shared__ double sh[8];
  bmin = wave_min(bmin); vmax = wave_max(vmax);
  if (threadIdx.x == 0) { sh[threadIdx.y] = bmin; sh[4 + threadIdx.y] = vmax; }
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    double a = sh[0], v = sh[4];
    for (int w = 1; w < (int)blockDim.y; w++) { a = fmin(a, sh[w]); v = fmax(v, sh[4 + w]); }
    gmem(I_.partial2)[2 * lb_] = a; gmem(I_.partial2)[2 * lb_ + 1] = v;
  }
}

// Momentum update (imex_subsystems.cpp:186-215).
__global__ void k_p_momentum(const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ p = gmem(I_.buf[B_P]);
  const double* __restrict__ mxs = gmem(I_.buf[B_MXS]);
  const double* __restrict__ mys = gmem(I_.buf[B_MYS]);
  double* __restrict__ u = gmem(I_.buf[iu]);
  const double* S = gmem(I_.S);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double h = S[S_ADT];
  double pc = p[kidx(d, i, j)];
  double fxm = 0.5 * (p[kidx(d, i - 1, j)] + pc), fxp = 0.5 * (pc + p[kidx(d, i + 1, j)]);
  const double gL = (j == GC && jet_inlet(d, i)) ? d.jet_p : p[kidx(d, i, j - 1)];   // jet strip ghost p (:202-203)
  double fym = 0.5 * (gL + pc), fyp = 0.5 * (pc + p[kidx(d, i, j + 1)]);
  long long c = pidx(d, i, j);
  u[1 * d.PS + c] = mxs[c] - h / d.dx * (fxp - fxm);
  u[2 * d.PS + c] = mys[c] - h / d.dy * (fyp - fym);
}

// Energy from the definition (imex_subsystems.cpp:286-302).
__global__ void k_p_energy(Phys P, const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ p = gmem(I_.buf[B_P]);
  double* __restrict__ u = gmem(I_.buf[iu]);
  int* eflag = gmem(I_.eflag);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  long long c = pidx(d, i, j);
  double q[NVAR];
  for (int k = 0; k < NVAR; k++) q[k] = u[k * d.PS + c];
  int ef = 0;
  u[4 * d.PS + c] = Vbase(P, p[kidx(d, i, j)], q[0], &ef) + 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]
                  + 0.5 * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
  if (ef) atomicOr(eflag, 1);
}

// ================================================================ energy subsystem (energy formula)
// RHS of PreSubsystem_Energy (imex_subsystems.cpp:427-472) with the coefficient
// h = caltotalenthalpy(Qe) rho^e / rho (timestep.cpp:232-237) formed at the cell and its four
// neighbours from Qe, p(Qe) and rho of u (ghost cells hold boundary copies of all three, so the
// neighbour values equal what the reference computes there).  Writes the cell's h (the operator's
// coefficient plane, B_HHAT), b = U* dx dy, the initial guess E (B_P) and the precE divisor
// diag dx dy (B_TDIAG, :479-500).
__global__ void __launch_bounds__(256) k_e_rhs(Phys P, const Inst* __restrict__ IT, BMap bm, int iu, int iQe) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ Qe = gmem(I_.buf[iQe]);
  const double* __restrict__ pe = gmem(I_.buf[B_PE]);
  double* __restrict__ hE = gmem(I_.buf[B_HHAT]);
  double* __restrict__ b = gmem(I_.buf[B_PB]);
  double* __restrict__ x0 = gmem(I_.buf[B_P]);
  double* __restrict__ D = gmem(I_.buf[B_TDIAG]);
  const double* S = gmem(I_.S);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double h = S[S_ADT], dx = d.dx, dy = d.dy, g = P.gamma;
  auto H = [&](int a, int bb) {
    long long c = pidx(d, a, bb);
    double re = Qe[c];
    return (Qe[4 * d.PS + c] + pe[c]) / re * re / u[c];
  };
  double h00 = H(i, j), hp0 = H(i + 1, j), hm0 = H(i - 1, j), h0p = H(i, j + 1), h0m = H(i, j - 1);
  long long c = pidx(d, i, j);
  auto U = [&](int a, int bb, int k) { return u[k * d.PS + pidx(d, a, bb)]; };
  double e = u[4 * d.PS + c];
  double U_star = e - 0.5 * h / dx * (hp0 * U(i + 1, j, 1) - hm0 * U(i - 1, j, 1))
                - 0.5 * h / dy * (h0p * U(i, j + 1, 2) - h0m * U(i, j - 1, 2));
  int idx = (i - GC) + d.nx * (j - GC);
  b[idx] = U_star * dx * dy;
  x0[idx] = e;
  hE[c] = h00;
  // ghost coefficients the operator reads that are not copies: the jet strip's and the Dirichlet sides'
  if (j == GC && jet_inlet(d, i)) hE[pidx(d, i, j - 1)] = h0m;
  if (dir_ghost(d, i + 1, j)) hE[pidx(d, i + 1, j)] = hp0;
  if (dir_ghost(d, i - 1, j)) hE[pidx(d, i - 1, j)] = hm0;
  if (dir_ghost(d, i, j + 1)) hE[pidx(d, i, j + 1)] = h0p;
  if (dir_ghost(d, i, j - 1)) hE[pidx(d, i, j - 1)] = h0m;
  double hhp_x = 0.5 * (h00 + hp0), hhm_x = 0.5 * (h00 + hm0), hhp_y = 0.5 * (h00 + h0p), hhm_y = 0.5 * (h00 + h0m);
  if (d.jet || d.dirx || d.diry) {   // Dirichlet lifting (imex_subsystems.cpp:453-475): getEc of the ghosts
    auto ec = [&](int a, int bb) {
      if (bb < GC && jet_inlet(d, a)) return d.jet_E;
      if (!dir_ghost(d, a, bb)) return 0.0;
      return d.e_amb + 0.5 * (U(a, bb, 5) * U(a, bb, 5) + U(a, bb, 6) * U(a, bb, 6) + U(a, bb, 7) * U(a, bb, 7));
    };
    double L_const = -(g - 1) * h * h * (1.0 / (dx * dx) * (hhp_x * ec(i + 1, j) + hhm_x * ec(i - 1, j))
                                       + 1.0 / (dy * dy) * (hhp_y * ec(i, j + 1) + hhm_y * ec(i, j - 1)));
    L_const *= dx * dy;
    b[idx] -= L_const;
  }
  double diag = 1.0 + (g - 1.0) * h * h * ((hhp_x + hhm_x) / (dx * dx) + (hhp_y + hhm_y) / (dy * dy));
  D[idx] = diag * dx * dy;
}

// After the solve (imex_subsystems.cpp:508-519): E <- the solution, then
// m_x -= h / (2 dx) (gamma - 1) (E(i+1) - E(i-1)) and likewise m_y, reading E through the ghost map.
__global__ void k_e_update(Phys P, const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ E = gmem(I_.buf[B_P]);
  double* __restrict__ u = gmem(I_.buf[iu]);
  const double* S = gmem(I_.S);

  int i = bx_ * blockDim.x + threadIdx.x + GC, j = by_ * blockDim.y + threadIdx.y + GC;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const double h = S[S_ADT], g = P.gamma;
  long long c = pidx(d, i, j);
  // a Dirichlet ghost holds the ambient energy of its B after bound_cond_u (boundary.cpp:66-70)
  auto EG = [&](int a, int bb) {
    if (!dir_ghost(d, a, bb)) return E[kidx(d, a, bb)];
    const long long q = pidx(d, a, bb);
    return d.e_amb + 0.5 * (u[5 * d.PS + q] * u[5 * d.PS + q] + u[6 * d.PS + q] * u[6 * d.PS + q] + u[7 * d.PS + q] * u[7 * d.PS + q]);
  };
  u[4 * d.PS + c] = E[kidx(d, i, j)];
  u[1 * d.PS + c] -= h * (1 / (2 * d.dx)) * (g - 1) * (EG(i + 1, j) - EG(i - 1, j));
  // the ghost below the jet strip holds the inflow energy after bound_cond_u (imex_subsystems.cpp:508-519)
  const double E_0m = (j == GC && jet_inlet(d, i)) ? d.jet_E : EG(i, j - 1);
  u[2 * d.PS + c] -= h * (1 / (2 * d.dy)) * (g - 1) * (EG(i, j + 1) - E_0m);
}

// ================================================================ constrained transport
// CT_CornerAssemble (constrained_transport.cpp:147-164) with Ez_p0 = -Fx, Ez_0p = +Fy and the
// cell-centre Ez~ = v_y B_x - v_x B_y of the post-pressure state (ct_flux.cpp:155-171).
// Corners (i,j) on [g-2, n+g) — the range the face update reads.
__device__ __forceinline__ double ez_tilde(const Dims& d, const double* u, int i, int j) {
  long long c = pidx(d, i, j);
  double r = u[c];
  return u[2 * d.PS + c] / r * u[5 * d.PS + c] - u[1 * d.PS + c] / r * u[6 * d.PS + c];
}
// Corner EMFs of a 64x4 tile: the cell EMFs Ez~ (each formed once, not once per corner that uses
// it) and the face EMFs of the tile's 65x5 cells are staged in LDS; the ghost frame's corners read
// their averaged term at the ghost-map image from global memory.
__global__ void __launch_bounds__(256) k_ct_corner(const Inst* __restrict__ IT, BMap bm, int iex, int iey, double sx, double sy, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ ezx = gmem(I_.buf[iex]);
  const double* __restrict__ ezy = gmem(I_.buf[iey]);
  const double* __restrict__ u = gmem(I_.buf[iu]);
  double* __restrict__ Ezpp = gmem(I_.buf[B_EZPP]);
  __shared__ double se[3][5][65];     // Ez~, sx ezx, sy ezy

  // Ez_p0 = sx * ezx, Ez_0p = sy * ezy (sx = -1, sy = +1 for the EMF-consistent form)
  const int i0 = bx_ * 64 + GC - 2, j0 = by_ * 4 + GC - 2;
  const int t = threadIdx.x + 64 * threadIdx.y;
  double v[2][7];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = t + 256 * q;
    if (e >= 65 * 5) break;
    const int a = min(i0 + e % 65, d.NX - 1), b = min(j0 + e / 65, d.NY - 1);
    const long long c = pidx(d, a, b);
    v[q][0] = u[c]; v[q][1] = u[d.PS + c]; v[q][2] = u[2 * d.PS + c]; v[q][3] = u[5 * d.PS + c]; v[q][4] = u[6 * d.PS + c];
    v[q][5] = ezx[c]; v[q][6] = ezy[c];
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = t + 256 * q;
    if (e >= 65 * 5) break;
    const int lx = e % 65, ly = e / 65;
    const double r = v[q][0];
    se[0][ly][lx] = v[q][2] / r * v[q][3] - v[q][1] / r * v[q][4];       // ez_tilde
    se[1][ly][lx] = sx * v[q][5];
    se[2][ly][lx] = sy * v[q][6];
  }
  __syncthreads();
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const int lx = threadIdx.x, ly = threadIdx.y;
  auto EX = [&](int a, int b) { return sx * ezx[pidx(d, a, b)]; };
  auto EY = [&](int a, int b) { return sy * ezy[pidx(d, a, b)]; };
  int I = map_x(d, i), J = map_y(d, j);
  // Ez_bar is formed inside and filled by bound_cond_array (constrained_transport.cpp:148-154)
  const double sum = (I == i && J == j) ? se[1][ly][lx] + se[1][ly + 1][lx] + se[2][ly][lx] + se[2][ly][lx + 1]
                                        : EX(I, J) + EX(I, J + 1) + EY(I, J) + EY(I + 1, J);
  double bar = refl_factor(d, i, j, true) * (0.25 * sum);
  const double dx = d.dx, dy = d.dy;
  const double e00 = se[0][ly][lx], e10 = se[0][ly][lx + 1], e01 = se[0][ly + 1][lx], e11 = se[0][ly + 1][lx + 1];
  const double x00 = se[1][ly][lx], x01 = se[1][ly + 1][lx], y00 = se[2][ly][lx], y10 = se[2][ly][lx + 1];
  // the differences over half a cell multiply by the reciprocals of dx/2, dy/2 (one division each per block)
  const double rhx = 1.0 / (dx / 2.0), rhy = 1.0 / (dy / 2.0);
  double t1 = 0.5 * ((x00 - e00) * rhx + (x01 - e01) * rhx);
  double t2 = 0.5 * ((e10 - x00) * rhx + (e11 - x01) * rhx);
  double t3 = 0.5 * ((y00 - e00) * rhy + (y10 - e10) * rhy);
  double t4 = 0.5 * ((e01 - y00) * rhy + (e11 - y10) * rhy);
  Ezpp[pidx(d, i, j)] = bar + dx / 8.0 * (t1 - t2) + dy / 8.0 * (t3 - t4);
}

// Face update (ct_flux.cpp:174-181) + bound_cond_staggered (:180-181) + cell B = face average
// (:184-190) + bound_cond_u, all at every padded index through the maps.
__device__ __forceinline__ int smap_x(const Dims& d, int axis, int i) {
  if (d.perx) return map_x(d, i);
  int lo = axis == 0 ? GC - 1 : GC;
  return i < lo ? lo : (i > d.nx + GC - 1 ? d.nx + GC - 1 : i);
}
__device__ __forceinline__ int smap_y(const Dims& d, int axis, int j) {
  if (d.pery) return map_y(d, j);
  int lo = axis == 1 ? GC - 1 : GC;
  return j < lo ? lo : (j > d.ny + GC - 1 ? d.ny + GC - 1 : j);
}
// The new face values of the 64x4 tile (and of the faces just below and left of it) are kept in
// LDS, so each is formed once; the cell average of an interior cell reads them there, the ghost
// frame's cells form theirs at the ghost-map image.
__global__ void __launch_bounds__(256) k_ct_faces(const Inst* __restrict__ IT, BMap bm, int ibxqs, int ibyqs, int ibx, int iby, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ bxqs = gmem(I_.buf[ibxqs]);
  const double* __restrict__ byqs = gmem(I_.buf[ibyqs]);
  const double* __restrict__ Ezpp = gmem(I_.buf[B_EZPP]);
  double* __restrict__ bx = gmem(I_.buf[ibx]);
  double* __restrict__ by = gmem(I_.buf[iby]);
  double* __restrict__ u = gmem(I_.buf[iu]);
  const double* S = gmem(I_.S);
  __shared__ double sbx[4][65], sby[5][64];   // BX at faces i0-1 .. i0+63, BY at faces j0-1 .. j0+3

  const int i0 = bx_ * blockDim.x, j0 = by_ * blockDim.y;
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const double h = S[S_ADT];
  auto BX = [&](int a, int b) { int I = smap_x(d, 0, a), J = smap_y(d, 0, b); return bxqs[pidx(d, I, J)] - h / d.dy * (Ezpp[pidx(d, I, J)] - Ezpp[pidx(d, I, J - 1)]); };
  auto BY = [&](int a, int b) { int I = smap_x(d, 1, a), J = smap_y(d, 1, b); return byqs[pidx(d, I, J)] + h / d.dx * (Ezpp[pidx(d, I, J)] - Ezpp[pidx(d, I - 1, J)]); };
  const bool in = i < d.NX && j < d.NY;
  if (in) {
    const double fbx = BX(i, j), fby = BY(i, j);
    sbx[threadIdx.y][threadIdx.x + 1] = fbx;
    sby[threadIdx.y + 1][threadIdx.x] = fby;
    const long long c = pidx(d, i, j);
    // the jet inflow strip's ghost faces and cells (constrained_transport.cpp:50-63, boundary.cpp:148-172)
    bx[c] = (j < GC && i < d.NX - 1 && jet_inlet_face(d, i)) ? 0.0 : fbx;
    by[c] = (j < GC - 1 && jet_inlet(d, i)) ? d.jet_B0 : fby;
    if (threadIdx.x == 0 && i > 0) sbx[threadIdx.y][0] = BX(i - 1, j);
    if (threadIdx.y == 0 && j > 0) sby[0][threadIdx.x] = BY(i, j - 1);
  }
  __syncthreads();
  if (!in) return;
  const long long c = pidx(d, i, j);
  if (j < GC && jet_inlet(d, i)) { u[5 * d.PS + c] = 0.0; u[6 * d.PS + c] = d.jet_B0; return; }
  int I = map_x(d, i), J = map_y(d, j);
  const double rf = refl_factor(d, i, j, true);        // bound_cond_u: B negated across a reflective wall
  double b5, b6;
  if (I == i && J == j) {
    b5 = rf * (0.5 * (sbx[threadIdx.y][threadIdx.x] + sbx[threadIdx.y][threadIdx.x + 1]));
    b6 = rf * (0.5 * (sby[threadIdx.y][threadIdx.x] + sby[threadIdx.y + 1][threadIdx.x]));
  } else {
    b5 = rf * (0.5 * (BX(I - 1, J) + BX(I, J)));
    b6 = rf * (0.5 * (BY(I, J - 1) + BY(I, J)));
  }
  u[5 * d.PS + c] = b5;
  u[6 * d.PS + c] = b6;
  if (dir_ghost(d, i, j)) {                             // the Dirichlet ambient's E follows its B
    const double b7 = u[7 * d.PS + c];
    u[4 * d.PS + c] = d.e_amb + 0.5 * (b5 * b5 + b6 * b6 + b7 * b7);
  }
}

// Reconstructed face states of rho, m_x, m_y, B_x, B_y (the EMF terms read no other component) on
// the faces (i+1/2, j) and (i, j+1/2) of the interior cell (i, j): w[dir][0] is the right state of
// (i, j), w[dir][1] the left state of its neighbour, taken at the neighbour's ghost-map image.  With
// the primitive reconstruction the stencil holds rho, v_x, v_y, B_x, B_y and the face momenta are
// rho v, as upri2ucon forms them (reconstruction.cpp:11-24, 272-281).
__device__ __forceinline__ void face_states_5(const Dims& d, const Phys& P, const double* __restrict__ u, int i, int j,
                                              double (*w)[2][5]) {
  const int ks[5] = {0, 1, 2, 5, 6};
  for (int dir = 0; dir < 2; dir++)
    for (int side = 0; side < 2; side++) {
      int ci = dir == 0 ? map_x(d, i + side) : i, cj = dir == 0 ? j : map_y(d, j + side);
      long long at[5];
      for (int o = 0; o < 5; o++) at[o] = dir == 0 ? pidx(d, ci - 2 + o, cj) : pidx(d, ci, cj - 2 + o);
      // jet strip ghosts of the primitive grid hold the conservative inflow state (see prim_faces)
      const int raw = (dir == 1 && jet_inlet(d, ci)) ? GC + 2 - cj : 0;
      for (int q = 0; q < 5; q++) {
        const double* src = u + ks[q] * d.PS;
        double a[5];
        for (int o = 0; o < 5; o++) a[o] = (P.prim && (q == 1 || q == 2) && o >= raw) ? src[at[o]] / u[at[o]] : src[at[o]];
        double wl, wr;
        recon5(P.recon, dir == 1 ? P.central_y : 0, a, wl, wr);
        w[dir][side][q] = side == 0 ? wr : wl;
      }
      if (P.prim) { w[dir][side][1] *= w[dir][side][0]; w[dir][side][2] *= w[dir][side][0]; }
    }
}

// Plain Staggered_CT face EMFs (constrained_transport.cpp:73-100): central averages of
// f_full_con[6] / g_full_con[5] over the reconstructed states of u_con.
__global__ void __launch_bounds__(256) k_ct_plain_emf(Phys P, const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  double* __restrict__ ezx = gmem(I_.buf[B_EZX]);
  double* __restrict__ ezy = gmem(I_.buf[B_EZY]);

  int ii = bx_ * blockDim.x + threadIdx.x, jj = by_ * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  int i = map_x(d, ii), j = map_y(d, jj);
  double w[2][2][5];
  face_states_5(d, P, u, i, j, w);
  // components in w: 0 rho, 1 m_x, 2 m_y, 3 B_x, 4 B_y
  auto fx6 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[4] * vx - vy * Q[3]; };
  auto gy5 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[3] * vy - vx * Q[4]; };
  const double rf = refl_factor(d, ii, jj, true);              // bound_cond_array of the face EMFs
  ezx[pidx(d, ii, jj)] = rf * (0.5 * (fx6(w[0][0]) + fx6(w[0][1])));   // Ez_p0 = -ezx
  ezy[pidx(d, ii, jj)] = rf * (0.5 * (gy5(w[1][0]) + gy5(w[1][1])));   // Ez_0p = +ezy
}

// ================================================================ per-step diagnostics (main.cpp:562-592)
// Per block over the interior: sum rho, sum e, the div B sum of squares (face or cell metric), max |B|^2/2,
// max |central div B| and the smallest e with the lowest (i-major) index among equal minima, which
// is the cell the reference's first-strictly-smaller scan keeps.
__global__ void k_diag(const Inst* __restrict__ IT, BMap bm, int iu, int ibx, int iby, int face_div) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  const double* __restrict__ bx = gmem(I_.buf[ibx]);
  const double* __restrict__ by = gmem(I_.buf[iby]);
  auto U = [&](int i, int j, int k) { return u[k * d.PS + pidx(d, i, j)]; };
  double mass = 0, intE = 0, dsq = 0, mmax = 0, binf = 0, emin = 1e30, key = 1e300;
  for (long long t = bx_ * (long long)blockDim.x + threadIdx.x; t < d.N; t += (long long)gx_ * blockDim.x) {
    int i = (int)(t % d.nx) + GC, j = (int)(t / d.nx) + GC;
    double rho = U(i, j, 0);
    double k = 0.5 * (U(i, j, 1) * U(i, j, 1) + U(i, j, 2) * U(i, j, 2) + U(i, j, 3) * U(i, j, 3)) / rho;
    double m = 0.5 * (U(i, j, 5) * U(i, j, 5) + U(i, j, 6) * U(i, j, 6) + U(i, j, 7) * U(i, j, 7));
    double e = U(i, j, 4) - k - m;
    mass += rho; intE += e; mmax = fmax(mmax, m);
    double kk = (double)(i - GC) * d.ny + (j - GC);
    if (e < emin || (e == emin && kk < key)) { emin = e; key = kk; }
    double v;
    if (face_div) v = (bx[pidx(d, i, j)] - bx[pidx(d, i - 1, j)]) / d.dx + (by[pidx(d, i, j)] - by[pidx(d, i, j - 1)]) / d.dy;
    else {
      double a, b;
      if (i == GC) a = (U(i + 1, j, 5) - U(i, j, 5)) / d.dx;
      else if (i == d.nx + GC - 1) a = (U(i, j, 5) - U(i - 1, j, 5)) / d.dx;
      else a = (U(i + 1, j, 5) - U(i - 1, j, 5)) / (2.0 * d.dx);
      if (j == GC) b = (U(i, j + 1, 6) - U(i, j, 6)) / d.dy;
      else if (j == d.ny + GC - 1) b = (U(i, j, 6) - U(i, j - 1, 6)) / d.dy;
      else b = (U(i, j + 1, 6) - U(i, j - 1, 6)) / (2.0 * d.dy);
      v = a + b;
    }
    dsq += d.dx * d.dy * v * v;
    double c = (U(i + 1, j, 5) - U(i - 1, j, 5)) / (2.0 * d.dx) + (U(i, j + 1, 6) - U(i, j - 1, 6)) / (2.0 * d.dy);
    binf = fmax(binf, fabs(c));
  }
  __shared__ double sh[7][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  mass = wave_sum(mass); intE = wave_sum(intE); dsq = wave_sum(dsq); mmax = wave_max(mmax); binf = wave_max(binf);
  for (int off = 32; off > 0; off >>= 1) {          // (emin, key) lexicographic minimum over the wave
    double oe = __shfl_down(emin, off), ok = __shfl_down(key, off);
    if (oe < emin || (oe == emin && ok < key)) { emin = oe; key = ok; }
  }
  if (lane == 0) { sh[0][w] = mass; sh[1][w] = intE; sh[2][w] = dsq; sh[3][w] = mmax; sh[4][w] = binf; sh[5][w] = emin; sh[6][w] = key; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* row = gmem(I_.partial) + bx_ * 8;
    double r[7] = {sh[0][0], sh[1][0], sh[2][0], sh[3][0], sh[4][0], sh[5][0], sh[6][0]};
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
      r[0] += sh[0][q]; r[1] += sh[1][q]; r[2] += sh[2][q]; r[3] = fmax(r[3], sh[3][q]); r[4] = fmax(r[4], sh[4][q]);
      if (sh[5][q] < r[5] || (sh[5][q] == r[5] && sh[6][q] < r[6])) { r[5] = sh[5][q]; r[6] = sh[6][q]; }
    }
    for (int q = 0; q < 7; q++) row[q] = r[q];
  }
}
__global__ void k_diag_fin(const Inst* __restrict__ IT, BMap bm, BMap src, double cfl, int append) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  if (threadIdx.x != 0) return;
  const int nblk = bmap_count(src, inst_);
  const double* P = gmem(I_.partial);
  double r[7] = {P[0], P[1], P[2], P[3], P[4], P[5], P[6]};
  for (int b = 1; b < nblk; b++) {
    const double* row = P + b * 8;
    r[0] += row[0]; r[1] += row[1]; r[2] += row[2]; r[3] = fmax(r[3], row[3]); r[4] = fmax(r[4], row[4]);
    if (row[5] < r[5] || (row[5] == r[5] && row[6] < r[6])) { r[5] = row[5]; r[6] = row[6]; }
  }
  const double* S = gmem(I_.S);
  const int key = (int)r[6], im = key / d.ny + GC, jm = key % d.ny + GC;
  const double dmin = fmin(d.dx, d.dy);
  const double fully_dt = cfl * 1 / (2 * S[S_CH] / dmin), imex_dt = cfl * 1 / (2 * S[S_AMAX] / dmin);
  double* rec;
  if (append) { int n = *gmem(I_.diag_n); rec = gmem(I_.diag) + (long long)(n % I_.diag_cap) * DIAG_REC; *gmem(I_.diag_n) = n + 1; }
  else rec = gmem(I_.diag) + (long long)I_.diag_cap * DIAG_REC;
  rec[0] = S[S_TIMESTEP]; rec[1] = S[S_T]; rec[2] = S[S_DT];
  rec[3] = sqrt(r[2]); rec[4] = r[4]; rec[5] = imex_dt / fully_dt;
  rec[6] = r[5]; rec[7] = d.x0 + (im - GC + 0.5) * d.dx; rec[8] = d.y0 + (jm - GC + 0.5) * d.dy;
  rec[9] = r[0] * d.dx * d.dy; rec[10] = r[1] * d.dx * d.dy; rec[11] = r[3];
}

// ---------------------------------------------------------------- Diagnose_CT_EMF (constrained_transport.cpp:181-222)
// Face parts of the corner EMF from the reconstruction of the state: central -(f6_L + f6_R)/2 and
// Rusanov max(|v_n|)/2 (B_t,R - B_t,L) on x faces (B_EZX, B_MXS), likewise on y faces (B_EZY, B_MYS),
// at every padded index through the ghost maps (the reference's bound_cond_array of the face arrays).
__global__ void __launch_bounds__(256) k_ctd_faces(Phys P, const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  int ii = bx_ * blockDim.x + threadIdx.x, jj = by_ * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  int i = map_x(d, ii), j = map_y(d, jj);
  double w[2][2][5];
  face_states_5(d, P, u, i, j, w);
  // components in w: 0 rho, 1 m_x, 2 m_y, 3 B_x, 4 B_y
  auto fx6 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[4] * vx - vy * Q[3]; };
  auto gy5 = [](const double* Q) { double vx = Q[1] / Q[0], vy = Q[2] / Q[0]; return Q[3] * vy - vx * Q[4]; };
  const double *L = w[0][0], *R = w[0][1];
  long long c = pidx(d, ii, jj);
  const double rf = refl_factor(d, ii, jj, true);              // bound_cond_array of the face arrays
  gmem(I_.buf[B_EZX])[c] = rf * (-(0.5 * (fx6(L) + fx6(R))));
  gmem(I_.buf[B_MXS])[c] = rf * (0.5 * fmax(fabs(L[1] / L[0]), fabs(R[1] / R[0])) * (R[4] - L[4]));
  L = w[1][0]; R = w[1][1];
  gmem(I_.buf[B_EZY])[c] = rf * (0.5 * (gy5(L) + gy5(R)));
  gmem(I_.buf[B_MYS])[c] = rf * (-0.5 * fmax(fabs(L[2] / L[0]), fabs(R[2] / R[0])) * (R[3] - L[3]));
}

// CT_CornerAssemble (constrained_transport.cpp:147-164) of the three parts: central and Rusanov faces
// with Ez~ = 0 (B_EZPP, B_CHI), and Ez~ with zero faces (B_VISC); corners [g-2, n+g).
__global__ void k_ctd_corners(const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ u = gmem(I_.buf[iu]);
  int i = bx_ * blockDim.x + threadIdx.x + GC - 2, j = by_ * blockDim.y + threadIdx.y + GC - 2;
  if (i >= d.nx + GC || j >= d.ny + GC) return;
  const int I = map_x(d, i), J = map_y(d, j);
  const double dx = d.dx, dy = d.dy;
  auto corner = [&](const double* ex, const double* ey, bool tilde) {
    auto EX = [&](int a, int b) { return ex ? ex[pidx(d, a, b)] : 0.0; };
    auto EY = [&](int a, int b) { return ey ? ey[pidx(d, a, b)] : 0.0; };
    auto ET = [&](int a, int b) { return tilde ? ez_tilde(d, u, a, b) : 0.0; };
    double bar = refl_factor(d, i, j, true) * (0.25 * (EX(I, J) + EX(I, J + 1) + EY(I, J) + EY(I + 1, J)));
    double e00 = ET(i, j), e10 = ET(i + 1, j), e01 = ET(i, j + 1), e11 = ET(i + 1, j + 1);
    double t1 = 0.5 * ((EX(i, j) - e00) / (dx / 2.0) + (EX(i, j + 1) - e01) / (dx / 2.0));
    double t2 = 0.5 * ((e10 - EX(i, j)) / (dx / 2.0) + (e11 - EX(i, j + 1)) / (dx / 2.0));
    double t3 = 0.5 * ((EY(i, j) - e00) / (dy / 2.0) + (EY(i + 1, j) - e10) / (dy / 2.0));
    double t4 = 0.5 * ((e01 - EY(i, j)) / (dy / 2.0) + (e11 - EY(i + 1, j)) / (dy / 2.0));
    return bar + dx / 8.0 * (t1 - t2) + dy / 8.0 * (t3 - t4);
  };
  long long c = pidx(d, i, j);
  gmem(I_.buf[B_EZPP])[c] = corner(gmem(I_.buf[B_EZX]), gmem(I_.buf[B_EZY]), false);
  gmem(I_.buf[B_CHI])[c] = corner(gmem(I_.buf[B_MXS]), gmem(I_.buf[B_MYS]), false);
  gmem(I_.buf[B_VISC])[c] = corner(nullptr, nullptr, true);
}

// CT_EMF_Stats (constrained_transport.cpp:165-178) of the three parts and their sum: per block sums
// of dx dy Ez^2 and dx dy |dB/dt|^2 and maxima of |Ez|, |dB/dt| (16 values per block row).
__global__ void k_ctd_stats(const Inst* __restrict__ IT, BMap bm) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* E3[3] = {gmem(I_.buf[B_EZPP]), gmem(I_.buf[B_CHI]), gmem(I_.buf[B_VISC])};
  double acc[16];
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  for (long long t = bx_ * (long long)blockDim.x + threadIdx.x; t < d.N; t += (long long)gx_ * blockDim.x) {
    int i = (int)(t % d.nx) + GC, j = (int)(t / d.nx) + GC;
    long long c = pidx(d, i, j), cw = pidx(d, i - 1, j), cs = pidx(d, i, j - 1);
    for (int q = 0; q < 4; q++) {
      double e, ew, es;
      if (q < 3) { e = E3[q][c]; ew = E3[q][cw]; es = E3[q][cs]; }
      else {
        e = E3[0][c] + E3[1][c] + E3[2][c];
        ew = E3[0][cw] + E3[1][cw] + E3[2][cw];
        es = E3[0][cs] + E3[1][cs] + E3[2][cs];
      }
      double dBx = -(e - es) / d.dy, dBy = (e - ew) / d.dx;
      double m2 = dBx * dBx + dBy * dBy;
      acc[4 * q] += d.dx * d.dy * e * e; acc[4 * q + 1] = fmax(acc[4 * q + 1], fabs(e));
      acc[4 * q + 2] += d.dx * d.dy * m2; acc[4 * q + 3] = fmax(acc[4 * q + 3], sqrt(m2));
    }
  }
  __shared__ double sh[16][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < 16; q++) {
    double v = (q & 1) ? wave_max(acc[q]) : wave_sum(acc[q]);
    if (lane == 0) sh[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int q = threadIdx.x;
    double v = sh[q][0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) v = (q & 1) ? fmax(v, sh[q][k]) : v + sh[q][k];
    gmem(I_.partial)[bx_ * 16 + q] = v;
  }
}
__global__ void k_ctd_fin(const Inst* __restrict__ IT, BMap bm, BMap src) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const int q = threadIdx.x;
  if (q >= 16) return;
  const int nblk = bmap_count(src, inst_);
  const double* P = gmem(I_.partial);
  double v = P[q];
  for (int b = 1; b < nblk; b++) v = (q & 1) ? fmax(v, P[b * 16 + q]) : v + P[b * 16 + q];
  gmem(I_.diag)[(long long)I_.diag_cap * DIAG_REC + DIAG_REC + q] = (q & 1) ? v : sqrt(v);
}

// ================================================================ state conversion, diagnostics
__global__ void k_con2pri(Phys P, const Inst* __restrict__ IT, BMap bm, int icon, int ipri) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ con = gmem(I_.buf[icon]);
  double* __restrict__ pri = gmem(I_.buf[ipri]);
  int* eflag = gmem(I_.eflag);

  long long total = (long long)d.NX * d.NY;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= total) return;
  double q[NVAR], r[NVAR];
  for (int k = 0; k < NVAR; k++) q[k] = con[k * d.PS + c];
  int ef = 0;
  con2pri(P, r, q, &ef);
  for (int k = 0; k < NVAR; k++) pri[k * d.PS + c] = r[k];
  if (ef) atomicOr(eflag, 1);
}
__global__ void k_pri2con(Phys P, const Inst* __restrict__ IT, BMap bm, int ipri, int icon) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ pri = gmem(I_.buf[ipri]);
  double* __restrict__ con = gmem(I_.buf[icon]);
  int* eflag = gmem(I_.eflag);

  long long total = (long long)d.NX * d.NY;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= total) return;
  double q[NVAR], r[NVAR];
  for (int k = 0; k < NVAR; k++) q[k] = pri[k * d.PS + c];
  int ef = 0;
  pri2con(P, r, q, &ef);
  for (int k = 0; k < NVAR; k++) con[k * d.PS + c] = r[k];
  if (ef) atomicOr(eflag, 1);
}
// cal_divergence_error, StaggerCT branch (imex_subsystems.cpp:630-660): per-block partial sums.
__global__ void k_divb(const Inst* __restrict__ IT, BMap bm, int ibx, int iby) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  const double* __restrict__ bx = gmem(I_.buf[ibx]);
  const double* __restrict__ by = gmem(I_.buf[iby]);
  double* partial = gmem(I_.partial);

  __shared__ double sh[16];
  double acc = divb_thread_sum(bx, by, d.nx, d.NX, d.N, d.dx, d.dy, bx_, gx_);
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[bx_] = s;
}

// Vbase_slope(p^e, rho) of the RK equation of state at every padded cell (the viscosity factor of
// the pressure operator, implicit_operators.cpp:584-600).
__global__ void k_visc_rk(Phys P, const Inst* __restrict__ IT, BMap bm, int iu) {
  BATCH_BLOCK(bm);
  const Inst& I_ = IT[inst_];
  const Dims& d = I_.d;
  long long c = bx_ * (long long)blockDim.x + threadIdx.x;
  if (c >= (long long)d.NX * d.NY) return;
  int ef = 0;
  gmem(I_.buf[B_VISC])[c] = Vbase_slope(P, gmem(I_.buf[B_PE])[c], gmem(I_.buf[iu])[c], &ef);
  if (ef) atomicOr(gmem(I_.eflag), 1);
}

// ================================================================ launchers
#define GRID(L, sh) dim3((unsigned)(L).nblocks[sh])
static const dim3 B64x4(64, 4);

void launch_bc_fill(const Launch& L, int n, const int* b, const int* k, int kind, const signed char* jv) {
  FillPlanes fp;
  fp.n = n;
  for (int p = 0; p < n; p++) {
    fp.b[p] = b[p]; fp.k[p] = k[p];
    // Grids take the inflow state (primitive in B_UPRI), arrays the copy, unless the caller says otherwise
    fp.jv[p] = jv ? jv[p] : (b[p] >= B_UPRI && b[p] <= B_QS ? (signed char)(k[p] + (b[p] == B_UPRI ? 8 : 0)) : (signed char)-1);
    // reflective sign class: grid slots other than rho and E (p); arrays negate; a plane formed from each
    // cell's own state (jv -2: the enthalpy ratio, even in B and v) does not
    const bool grid = b[p] >= B_UPRI && b[p] <= B_QS;
    fp.rn[p] = fp.jv[p] == -2 ? 0 : (grid ? (k[p] != 0 && k[p] != 4) : 1);
  }
  const int sh = kind == 0 ? SH_FILL0 : (kind == 1 ? SH_FILL1 : SH_FILL2);
  k_bc_fill<<<GRID(L, sh), 256, 0, L.st>>>(L.IT, L.bm[sh], fp, kind);
}
void launch_step_start(const Launch& L, const Phys& P, int mode, int iin) {
  if (P.eos != IMEX_EOS_IDEAL) k_step_start<true><<<GRID(L, SH_START), 256, 0, L.st>>>(P, L.IT, L.bm[SH_START], mode, iin);
  else k_step_start<false><<<GRID(L, SH_START), 256, 0, L.st>>>(P, L.IT, L.bm[SH_START], mode, iin);
  k_step_finish<<<GRID(L, SH_ONE), 256, 0, L.st>>>(P, L.IT, L.bm[SH_ONE], L.bm[SH_START]);
}
void launch_stage_combine(const Launch& L, int ibxn, int ibyn) {
  k_stage_combine<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(L.IT, L.bm[SH_LIN_TOT], ibxn, ibyn);
}
void launch_qe_prep(const Launch& L, const Phys& P, int iQe) {
  k_qe_prep<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(P, L.IT, L.bm[SH_LIN_TOT], iQe);
}
void launch_shock(const Launch& L, const Phys& P, int iQe) {
  k_shock_chi<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iQe);
  const int chi[1] = {B_CHI}, k0[1] = {0};
  launch_bc_fill(L, 1, chi, k0, 0);
  k_shock_eta<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4]);
  const int eta[1] = {B_ETA};
  launch_bc_fill(L, 1, eta, k0, 0);
}
template <int TXE, int TYE>
static void launch_explicit_t(const Launch& L, const Phys& P, int kind, int iQe, int iQs, int iu, int mixed) {
  constexpr int NT = TXE * TYE;
  if (P.bnd) {
    if (P.prim) k_explicit<TXE, TYE, double, true, true><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
    else if (mixed) k_explicit<TXE, TYE, float, false, true><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
    else k_explicit<TXE, TYE, double, false, true><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
  } else {
    if (P.prim) k_explicit<TXE, TYE, double, true, false><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
    else if (mixed) k_explicit<TXE, TYE, float, false, false><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
    else k_explicit<TXE, TYE, double, false, false><<<GRID(L, SH_EXPL), NT, 0, L.st>>>(P, L.IT, L.bm[SH_EXPL], kind, iQe, iQs, iu);
  }
}
// (tile, rows): the explicit tile (16x16, 32x8, 64x4, 32x16 or 64x8 cells); the SH_EXPL block map must
// be built for it
void launch_explicit(const Launch& L, const Phys& P, int kind, int iQe, int iQs, int iu, int mixed, int tile, int rows) {
  if (tile == 16) launch_explicit_t<16, 16>(L, P, kind, iQe, iQs, iu, mixed);
  else if (tile == 64 && rows == 8) launch_explicit_t<64, 8>(L, P, kind, iQe, iQs, iu, mixed);
  else if (tile == 64) launch_explicit_t<64, 4>(L, P, kind, iQe, iQs, iu, mixed);
  else if (rows == 16) launch_explicit_t<32, 16>(L, P, kind, iQe, iQs, iu, mixed);
  else launch_explicit_t<32, 8>(L, P, kind, iQe, iQs, iu, mixed);
}
void launch_mag_rhs(const Launch& L, const Phys& P, int iu, int iQe, int iQs, int ib, int ix0) {
  if (P.eos != IMEX_EOS_IDEAL) k_mag_rhs<true><<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iu, iQe, iQs, ib, ix0);
  else k_mag_rhs<false><<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iu, iQe, iQs, ib, ix0);
}
void launch_mag_writeback(const Launch& L, int iu) {
  k_mag_writeback<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(L.IT, L.bm[SH_LIN_TOT], iu);
}
void launch_emf_flux(const Launch& L, int iu, int iQe) {
  k_emf_flux<<<GRID(L, SH_PAD64x8), B64x4, 0, L.st>>>(L.IT, L.bm[SH_PAD64x8], iu, iQe);
}
void launch_p_rhs(const Launch& L, const Phys& P, int pass, int iu) {
  k_p_rhs<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], pass, iu);
}
void launch_p_momentum(const Launch& L, int iu) {
  k_p_momentum<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT64x4], iu);
}
void launch_e_rhs(const Launch& L, const Phys& P, int iu, int iQe) {
  k_e_rhs<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iu, iQe);
}
void launch_e_update(const Launch& L, const Phys& P, int iu) {
  k_e_update<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iu);
}
void launch_p_energy(const Launch& L, const Phys& P, int iu) {
  k_p_energy<<<GRID(L, SH_INT64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_INT64x4], iu);
}
void launch_ct(const Launch& L, const Phys& P, int plain, int iu, int ibxqs, int ibyqs, int ibx, int iby) {
  int iex = B_FX, iey = B_FY;
  if (plain) {
    k_ct_plain_emf<<<GRID(L, SH_PAD64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_PAD64x4], iu);
    iex = B_EZX; iey = B_EZY;
  }
  k_ct_corner<<<GRID(L, SH_INT2_64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT2_64x4], iex, iey, -1.0, 1.0, iu);
  k_ct_faces<<<GRID(L, SH_PAD64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_PAD64x4], ibxqs, ibyqs, ibx, iby, iu);
}
void launch_diag(const Launch& L, int iu, int ibx, int iby, int face_div, double cfl, int append) {
  k_diag<<<GRID(L, SH_DIVB), 256, 0, L.st>>>(L.IT, L.bm[SH_DIVB], iu, ibx, iby, face_div);
  k_diag_fin<<<GRID(L, SH_ONE), 64, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_DIVB], cfl, append);
}
void launch_ct_emf_diag(const Launch& L, const Phys& P, int iu) {
  k_ctd_faces<<<GRID(L, SH_PAD64x4), B64x4, 0, L.st>>>(P, L.IT, L.bm[SH_PAD64x4], iu);
  k_ctd_corners<<<GRID(L, SH_INT2_64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT2_64x4], iu);
  k_ctd_stats<<<GRID(L, SH_DIVB), 256, 0, L.st>>>(L.IT, L.bm[SH_DIVB]);
  k_ctd_fin<<<GRID(L, SH_ONE), 64, 0, L.st>>>(L.IT, L.bm[SH_ONE], L.bm[SH_DIVB]);
}
void launch_ct_corners(const Launch& L, int iu) {
  k_ct_corner<<<GRID(L, SH_INT2_64x4), B64x4, 0, L.st>>>(L.IT, L.bm[SH_INT2_64x4], B_FX, B_FY, -1.0, 1.0, iu);
}
void launch_con2pri(const Launch& L, const Phys& P, int icon, int ipri) {
  k_con2pri<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(P, L.IT, L.bm[SH_LIN_TOT], icon, ipri);
}
void launch_pri2con(const Launch& L, const Phys& P, int ipri, int icon) {
  k_pri2con<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(P, L.IT, L.bm[SH_LIN_TOT], ipri, icon);
}
void launch_divb(const Launch& L, int ibx, int iby) {
  k_divb<<<GRID(L, SH_DIVB), 256, 0, L.st>>>(L.IT, L.bm[SH_DIVB], ibx, iby);
}
void launch_visc_rk(const Launch& L, const Phys& P, int iu) {
  k_visc_rk<<<GRID(L, SH_LIN_TOT), 256, 0, L.st>>>(P, L.IT, L.bm[SH_LIN_TOT], iu);
}
