// batch.hpp — instance table and block map shared by every kernel.
//
// A context holds one or more independent instances (grids of their own size, with their own
// scalars, Krylov control block and buffers).  Kernels receive the device table of instances
// and small buffer ids instead of raw pointers; one launch covers every instance: workgroup b
// belongs to the instance i with pre[i] <= b < pre[i+1] and is local block b - pre[i] of that
// instance's grid (gx[i] blocks per row); a per-workgroup table holds (i, b - pre[i]).  A
// single-instance context is a batch of one.
#pragma once

#define MAXM 32            // largest GMRES restart length supported
#define MAXSWEEP 16        // refinement sweeps recorded per mixed-form solve

// FP64 buffers of an instance.  Grids (B_UPRI..B_QS) are 8 planes; the vector potentials (B_A0..B_AS,
// allocated only when the magnetic treatment evolves A) 3 planes; the others one plane or one
// compact Krylov vector (3 components for the magnetic system).
enum BufId {
  B_UPRI, B_UN, B_U1, B_U2, B_U3, B_QE, B_QS,
  B_BX0, B_BY0, B_BX1, B_BY1, B_BX2, B_BY2, B_BXQS, B_BYQS, B_FX, B_FY,
  B_PE, B_CF, B_CHI, B_ETA, B_HHAT, B_MXS, B_MYS, B_EZPP, B_EZX, B_EZY, B_VISC,
  B_W, B_AV, B_XK, B_BK, B_TMPN, B_P, B_PB, B_NG, B_DP, B_TDIAG, B_VSLOPE,
  B_A0, B_A1, B_A2, B_AS,
  B_V0, B_VEND = B_V0 + MAXM + 1,
  NBUF = B_VEND
};
// FP32 buffers (inner solves of the mixed form)
enum FBufId { F_W32, F_AV32, F_V0, F_VEND = F_V0 + MAXM + 1, NFBUF = F_VEND };

struct KrylovDev;

struct Inst {
  Dims d;
  double* buf[NBUF];
  float* fbuf[NFBUF];
  double* S;             // step scalars (ScalarSlot)
  double* partial;       // per-block partial sums
  double* partial2;      // two per-block values a producer leaves for a later finish (Mbig, positivity)
  double* pst;           // positivity statistics (min p, max floored, violation)
  int* eflag;            // EOS iteration cap reached
  KrylovDev* K;          // GMRES / Newton control block
  int* ctl;              // budgeted-schedule failure flags and iteration log
  double* bsave;         // scalar snapshot of the budgeted step
  int* esave;
  double* diag;          // per-step diagnostics: ring of diag_cap records, then a scratch record area
  int* diag_n;           // records written so far
  int diag_cap;
};

// Buffers of every instance live in device global memory, but a pointer read from the instance
// table is generic to the compiler, which would then address it with flat instructions (these
// count against both the memory and the LDS wait counters).  gmem() states the address space:
// going through an integer keeps the compiler from folding the cast away.
template <class V>
__device__ __forceinline__ V* gmem(V* p) {
  return (V*)(__attribute__((address_space(1))) V*)(unsigned long long)p;
}

struct BMap {
  const int2* blk;       // per workgroup: (instance, local block)
  const int* pre;        // ni + 1 prefix sums of the per-instance block counts
  const int* gx;         // blocks per grid row of each instance
  int ni;
  int gx0, nb0;          // row width and block count when ni == 1 (no table reads)
};

// Local block count of instance i.
__device__ __forceinline__ int bmap_count(const BMap& m, int i) {
  return m.ni == 1 ? m.nb0 : m.pre[i + 1] - m.pre[i];
}

// Opens a kernel body: inst_ is the instance, bx_/by_ the local block indices, gx_/gy_ the local
// grid (for 1D grids gy_ = 1 and gx_ is the local block count).  A single-instance map reads no
// table, so the kernel's first loads do not wait on a lookup.
#define BATCH_BLOCK(bm)                                                     \
  int inst_, lb_, gx_;                                                      \
  if ((bm).ni == 1) { inst_ = 0; lb_ = blockIdx.x; gx_ = (bm).gx0; }        \
  else {                                                                    \
    const int2 e_ = (bm).blk[blockIdx.x];                                   \
    inst_ = e_.x; lb_ = e_.y; gx_ = (bm).gx[inst_];                         \
  }                                                                         \
  const int gy_ = bmap_count(bm, inst_) / gx_;                              \
  const int bx_ = lb_ % gx_, by_ = lb_ / gx_;                               \
  (void)gy_; (void)bx_; (void)by_

// launch shapes (per-instance grids, computed once per context)
enum Shape {
  SH_INT64x4,     // (nx, ny) in 64x4 tiles
  SH_EXPL,        // (nx, ny) in explicit tiles (16x16, 32x8 or 64x4, autotuned)
  SH_PAD64x4,     // (nx+6, ny+6) in 64x4 tiles
  SH_PAD64x8,     // (nx+6, ny+6) in 64x8 tiles
  SH_INT2_64x4,   // (nx+2, ny+2) in 64x4 tiles
  SH_LIN_TOT,     // (nx+6)(ny+6) cells, 256 per block
  SH_LIN_N,       // nx ny cells, 256 per block
  SH_START,       // start-of-step reduction grid
  SH_TILES,       // operator tiles (64x4), capped: magnetic operator
  SH_TILES_P,     // operator tiles (64x4), capped: pressure operator
  SH_TILES_N,     // operator tiles (64x4), one round of resident blocks: Newton residual
  SH_TILES_ALL,   // operator tiles, one block each up to 4096
  SH_VEC3,        // 3 nx ny vector elements, capped
  SH_VEC1,        // nx ny vector elements, capped
  SH_MBIG,        // Mbig reduction grid
  SH_PSTATS,      // positivity statistics grid
  SH_DIVB,        // div B reduction grid
  SH_ONE,         // one block per instance
  SH_FILL0, SH_FILL1, SH_FILL2,   // ghost-frame fills (cell / x-normal faces / y-normal faces)
  NSHAPE
};
