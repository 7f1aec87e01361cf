// driver.hip — C ABI (include/imex.h) and the C++ host driver of the IMEX-MHD step on one GPU.
//
// The context replaces the reference's mutable globals (include/config.hpp:43-125) and
// its caller-owned Grid/Array objects (src/main.cpp:231-250): all fields live on the
// device for the lifetime of the context, in SoA planes with the 3-cell ghost frame.
// The step loop body of src/main.cpp:334-595 is reproduced kernel by kernel; host code
// only sequences launches and reads the GMRES / Newton convergence flags.
//
// A context holds one instance (imex_create) or a batch of independent instances of
// uneven grid size (imex_batch_create).  Every launch covers all instances: kernels find
// their instance and local block through a per-shape block map (batch.hpp).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include "kernels.hpp"
#include "adjoint.hpp"

static thread_local std::string g_err;
static int set_err(int code, const std::string& m) { g_err = m; return code; }
int imex_internal_error(int code, const std::string& msg) { return set_err(code, msg); }

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return set_err(IMEX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)

#define CTLN 2048          // ints of budget log per instance (bounds the solves of one step)

// ---------------------------------------------------------------- budgeted-schedule bookkeeping
// A budgeted step runs every GMRES solve for a fixed number of Arnoldi launches and every
// nested-Newton solve for a fixed number of Jacobian solves; kernels queued past convergence
// exit at once.  ctl[0] of each instance collects failures (a solve that had not converged when
// its budget ran out), ctl[1..] the iteration count of each solve, in launch order.  The
// scalars, positivity statistics and EOS flag are saved at the start so that a failed step can
// be rerun exactly.
__global__ void k_budget_begin(const Inst* __restrict__ IT, BMap bm) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const int t = threadIdx.x;
  if (t < S_NSLOTS) gmem(I.bsave)[t] = gmem(I.S)[t];
  if (t < 4) gmem(I.bsave)[S_NSLOTS + t] = gmem(I.pst)[t];
  if (t == 0) { gmem(I.esave)[0] = gmem(I.eflag)[0]; gmem(I.ctl)[0] = 0; }
}
__global__ void k_budget_restore(const Inst* __restrict__ IT, BMap bm) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const int t = threadIdx.x;
  if (t < S_NSLOTS) gmem(I.S)[t] = gmem(I.bsave)[t];
  if (t < 4) gmem(I.pst)[t] = gmem(I.bsave)[S_NSLOTS + t];
  if (t == 0) gmem(I.eflag)[0] = gmem(I.esave)[0];
}
// GMRES log: ctl[slot] = Arnoldi steps over all sweeps, ctl[slot+1] = sweeps, then the steps of
// each of the nsw budgeted sweeps.  A FP64-form solve has one sweep and converges inside it; a
// mixed-form solve converges at the FP64 residual that opens sweep `sweeps`.
__global__ void k_budget_gmres_check(const Inst* __restrict__ IT, BMap bm, int slot, int mixed, int nsw) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  const KrylovDev* K = gmem(I.K);
  int* ctl = gmem(I.ctl);
  const int t = threadIdx.x;
  // sweeps: the mixed form's refinement sweeps (the residual of sweep `outer` found it converged),
  // or the FP64 form's restart cycles (cycle `outer` converged)
  const int sweeps = mixed ? K->outer : K->outer + 1;
  if (t == 0) {
    ctl[slot] = K->steps_total;
    ctl[slot + 1] = sweeps;
    if (!K->outer_conv) ctl[0] |= 1;
  }
  if (t < nsw) ctl[slot + 2 + t] = t < sweeps && t < MAXSWEEP ? K->sweep_steps[t] : 0;
}
// Newton log: ctl[slot] = Jacobian solves the solve needed, or -1 when it did not converge; with the
// floor active also the positivity statistics of the final p, from the last Newton residual's
// partials (grid src).
__global__ void __launch_bounds__(1024) k_budget_newton_check(const Inst* __restrict__ IT, BMap bm, int slot, BMap src, int pst, double thr) {
  BATCH_BLOCK(bm);
  const Inst& I = IT[inst_];
  if (threadIdx.x == 0) {
    gmem(I.ctl)[slot] = gmem(I.K)->newton_conv ? gmem(I.K)->newton_first : -1;
    if (!gmem(I.K)->newton_conv) gmem(I.ctl)[0] |= 2;
  }
  if (pst) pstats_reduce(I, bmap_count(src, inst_), thr);
}

// ---------------------------------------------------------------- per-kernel-class event profiler
// When enabled, every launch on the step path is bracketed by two HIP events on the
// context stream and tagged with its kernel class and algorithmic byte count (distinct
// per-cell fields read once plus fields written once, times cells, times 8 bytes).
enum KClass {
  KC_START, KC_COMBINE, KC_QEPREP, KC_SHOCK, KC_EXPL_CONV, KC_EXPL_PRESS, KC_MAG_RHS, KC_MAG_APPLY, KC_MAG_UPDATE,
  KC_MAG_RES, KC_MAG_XUPD, KC_P_APPLY, KC_P_UPDATE, KC_P_RES, KC_P_XUPD, KC_NEWTON_RES, KC_EMF, KC_P_RHS, KC_P_MOM,
  KC_P_ENERGY, KC_CT, KC_BC, KC_MISC, KC_KFIN, KC_EARLY_EXIT, KC_EXPL_ENERGY, KC_E_RHS, KC_E_APPLY, KC_E_UPDATE,
  KC_E_RES, KC_E_XUPD, KC_E_MOM, KC_A_UPDATE, KC_N
};
static const char* kc_names[KC_N] = {
  "step_start", "stage_combine", "qe_prep", "shock", "explicit_conv", "explicit_press", "mag_rhs", "mag_arnoldi_apply",
  "mag_arnoldi_update", "mag_gmres_residual", "mag_gmres_xupdate", "p_arnoldi_apply", "p_arnoldi_update",
  "p_gmres_residual", "p_gmres_xupdate", "newton_residual", "emf_flux", "p_rhs", "p_momentum", "p_energy", "ct",
  "ghost_fill", "misc", "krylov_finish", "early_exit", "explicit_energy", "e_rhs", "e_arnoldi_apply",
  "e_arnoldi_update", "e_gmres_residual", "e_gmres_xupdate", "e_momentum", "a_update"};

struct ProfRec { int cls; int e0, e1; double bytes; };
struct Prof {
  bool on = false;
  std::vector<hipEvent_t> ev;
  int next = 0;
  std::vector<ProfRec> recs;
  long long gen = 0;          // bumped whenever recs are flushed
  double ms[KC_N] = {0}, bytes[KC_N] = {0};
  long long n[KC_N] = {0};
};

// Solve positions within a step: GMRES solve key = (stage*3 + sub)*NIT + Newton iteration with sub 0
// the magnetic solve, 1/2 the two pressure solves; Newton key = stage*3 + sub.
#define NIT 20
#define NGKEY (6 * NIT)
#define NNKEY 6
// Budget table of a step: Arnoldi steps per sweep and sweeps per GMRES key, Jacobian solves per
// Newton key.  It is also the identity of the step's graph.
struct BudTab { int B[NGKEY], S[NGKEY], N[NNKEY]; };

// One budgeted GMRES (kind 0) or nested-Newton (kind 1) solve of the queued step.
struct BudgetRec {
  int kind, op, slot, key = 0;
  int parent = -1, it = -1;    // enclosing Newton solve (index into brecs) and its iteration
  int nsw = 1, B = 0;                     // budgeted sweeps and Arnoldi steps per sweep
  std::vector<int> ra, ru, rfa, rfu;      // profiler records of the Arnoldi launches and finishes, sweep-major
  std::vector<int> rres, rresf, rx;       // per sweep (+ the closing residual of a mixed-form solve)
  std::vector<int> rnres, rnresf, raxpy;          // Newton: residual (+finish) / update per iteration
};
// Launches of a captured step whose execution depends on the budgets: Arnoldi step j of sweep
// sw of GMRES key `key` (4 kernels), the residual (+finish) opening sweep sw, the x update of
// sweep sw, the log kernel of a GMRES solve, and, per Newton key and iteration j, the
// iteration's work (vslope, dp fill, update) and its residual (+finish).
enum { TG_ARN, TG_RES, TG_XUPD, TG_GCHK, TG_NIT, TG_NRES };
struct NodeTag { hipGraphNode_t node; unsigned char kind; short key, sw, j; };
// A step graph is captured with a capacity budget table; each replay enables the nodes that
// the current budgets use and disables the rest (disabled nodes are skipped, not launched).
struct GraphEntry {
  BudTab cap, applied;
  int parity, state_mode;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<NodeTag> tags;
  std::vector<BudgetRec> brecs;
  int nslot = 1;
  long long used = 0;
  bool gk[NGKEY] = {};       // the GMRES budget positions that have nodes in the graph
};
#define BHIST 8
#define MAX_GRAPHS 6
#define NBK 5

// buffer ids of the stage being queued
struct StageBind { int u, Qe, Qs, bxqs, byqs, bxo, byo, As, Ao; };   // As/Ao: the stage's A_s and output A

// Points of a host-scheduled step at which a recording hook (the adjoint's tape) is called.
enum HookPoint { HK_START, HK_IN, HK_CONV, HK_MAG, HK_PRESS, HK_P0, HK_MOM, HK_P1, HK_PR, HK_CT };
struct imex_ctx;
typedef void (*StepHook)(imex_ctx*, int);
struct AdjState;
static void adj_free(imex_ctx* c);

struct imex_ctx {
  Prof prof;
  imex_params prm;             // shared parameters (nx, ny of instance 0)
  Phys P;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  long long bytes = 0;
  // instances
  int ni = 1;
  std::vector<imex_params> iprm;     // per-instance parameters (grid size)
  std::vector<Inst> hI;              // host copy of the instance table
  Inst* IT = nullptr;                // device instance table
  int* bmap_dev = nullptr;           // block maps of all shapes
  int2* blk_dev = nullptr;           // per-workgroup table of a batch (all shapes)
  long long blk_cap = 0;             // its capacity (the largest caps the autotuner may choose)
  Launch L;
  long long cells = 0;               // sum of nx ny over instances
  // grid caps of the Krylov kernels (autotuned): a cap that is a multiple of the blocks a kernel
  // keeps resident avoids a partly filled last round (magnetic FP64 apply: 3 waves/SIMD, 768
  // blocks; pressure FP64: 6, 1536; pressure FP32 apply: 8, 2048)
  int cap_tiles = 2048, cap_tiles_p = 2048, cap_vec = 2048;
  int cap_tiles_n = 2048;            // Newton residual: the blocks resident at once (one round)
  int ex_tile = 32, ex_rows = 8;          // explicit-kernel tile (16x16, 32x8, 64x4, 32x16 or 64x8; autotuned)
  // state buffers: the step reads Us (or u_pri before the first step) and its last stage writes Uo;
  // faces likewise (bxn/byn read, bx2/by2 written)
  int iUs = B_U2, iUo = B_U3, iBxn = B_BX0, iByn = B_BY0, iBx2 = B_BX2, iBy2 = B_BY2;
  // vector potential (UNSTAGGER / POTENTIAL_A): the step reads An, stage 1 writes A1, stage 2 writes Ao
  bool needA = false;
  int iAn = B_A0, iAo = B_A2;
  std::vector<char> a_set;           // per instance: imex_set_potential was called
  KrylovDev* Kh = nullptr;     // pinned host mirrors of instance 0's control block (2 slots)
  KrylovDev Klast;             // status of the last GMRES cycle read by the host
  double* Sh = nullptr;        // pinned host copy of one instance's scalars
  int* ctlh = nullptr;         // pinned host copy of every instance's budget log
  int* ctl_all = nullptr;      // device: ni * CTLN
  hipEvent_t ev[2];
  int state_mode = 0;          // 0: u_pri holds the state; 1: Us holds the conservative state
  StageBind cur;
  bool qe_ready = false, hhat_ready = false;
  imex_stats stats{};
  int pred[2] = {6, 6};         // Arnoldi steps of the last magnetic / pressure solve
  int pred_newton = 1;         // Jacobian solves of the last nested-Newton solve
  // budgeted schedule
  bool budget = false;         // the step being queued runs on budgets
  int nslot = 1;
  std::vector<BudgetRec> brecs;
  // budget history, per kind: 0/1 Arnoldi steps per sweep of a magnetic / pressure
  // solve, 2 Jacobian solves per Newton solve, 3/4 sweeps per magnetic / pressure solve
  int hist[NBK][BHIST];        // per-step maxima of the recent steps
  int hist_n = 0, hist_pos = 0;
  // extra budgets after an overrun: Arnoldi steps (+2 per GMRES overrun of a batch, +1 of a
  // single instance), refinement sweeps (a batch) and Newton iterations (+1 per overrun of their
  // kind); each decays by 1 every 16 steps that fit
  int margin_extra = 0, sweep_extra = 0, newton_extra = 0, fit_streak = 0;
  int step_mx[NBK] = {0, 0, 0, 0, 0};   // maxima of the step being run under host scheduling
  // the same per solve position (-1: no such solve in that step)
  int khist[BHIST][NGKEY][2], nhist[BHIST][NNKEY];
  int kcur[NGKEY][2], ncur[NNKEY];
  BudTab tab;
  int stage_idx = 0, sub_idx = 0, newton_it = 0;
  // mixed form: FP32 fluxes (pressure formula only, mixed_flux); per operator (bit 0 magnetic,
  // bit 1 pressure / energy) FP32 inner solves
  int mixed = 0, mixed_ops = 0, mixed_flux = 0;
  // IMEX_SCHED_AUTO (single instance): timed trials of host-polled and budgeted (graph) steps;
  // phase 0 host trial, 1 budget trial, 2 the faster schedule for AUTO_RUN steps
  int auto_phase = 2, auto_left = 0;
  bool auto_budget = false, auto_win = true;
  double auto_ms[2] = {0.0, 0.0};
  std::chrono::steady_clock::time_point auto_t0;
  std::vector<GraphEntry> graphs;
  long long graph_clock = 0;
  bool capturing = false;      // launches are being captured into a step graph
  bool tag_err = false;
  bool no_toggle = false;      // node enabling failed once: graphs are captured per exact budget table
  std::vector<NodeTag> tags;
  StepHook hook = nullptr;     // called at HookPoint's while a step is queued host-scheduled
  bool taping = false;         // steps are host-scheduled (the hook sees every point)
  AdjState* adj = nullptr;     // discrete-adjoint checkpoints, tape and work buffers
  int diag_cap = 0;            // per-step diagnostics ring capacity (0: not recording)
  bool diag_alloc = false;     // the diagnostics buffers exist (their capacity is fixed at allocation)
  const Dims& d0() const { return hI[0].d; }
};

static Dims make_dims(const imex_params& p) {
  Dims d;
  d.nx = p.nx; d.ny = p.ny; d.NX = p.nx + 2 * GC; d.NY = p.ny + 2 * GC;
  long long ps = (long long)d.NX * d.NY;
  d.PS = (ps + 31) / 32 * 32;
  d.N = p.nx * p.ny;
  d.dx = (p.x1 - p.x0) / ((p.nx + 1) - 1);     // main.cpp:226-227
  d.dy = (p.y1 - p.y0) / ((p.ny + 1) - 1);
  d.perx = p.bc_x == IMEX_BC_PERIODIC; d.pery = p.bc_y == IMEX_BC_PERIODIC;
  d.refx = p.bc_x == IMEX_BC_REFLECTIVE; d.refy = p.bc_y == IMEX_BC_REFLECTIVE;
  d.dirx = p.bc_x == IMEX_BC_DIRICHLET; d.diry = p.bc_y == IMEX_BC_DIRICHLET;
  d.rho_amb = p.rho_amb; d.p_amb = p.p_amb; d.e_amb = p.p_amb / (p.gamma - 1.0);   // config.hpp:247, boundary.cpp:67-69
  d.x0 = p.x0; d.y0 = p.y0;
  d.jet = p.bc_y == IMEX_BC_JET_INFLOW;
  d.jet_hw = p.jet_halfwidth; d.jet_rho = p.jet_rho; d.jet_v = p.jet_v; d.jet_B0 = p.jet_B0; d.jet_p = p.jet_p;
  d.jet_E = p.jet_p / (p.gamma - 1.0) + 0.5 * p.jet_rho * p.jet_v * p.jet_v + 0.5 * p.jet_B0 * p.jet_B0;   // config.hpp:271-273
  return d;
}
static Phys make_phys(const imex_params& p) {
  Phys P;
  P.gamma = p.gamma; P.eos = p.eos;
  P.rk_alfa = p.rk_alfa; P.rk_b = p.rk_b; P.rk_R = p.rk_R; P.rk_cv = p.rk_cv;
  P.floor_on = p.floor_on; P.p_floor = p.p_floor; P.theta = p.theta; P.mbig_fac = p.mbig_fac;
  P.kappa1 = p.shock_kappa1; P.cfl = p.cfl; P.time_stop = p.time_stop;
  P.recon = p.recon; P.central_y = p.minmod_central_y;
  P.prim = p.recon_state == IMEX_RECONSTATE_PRIM;
  P.bnd = p.bc_y == IMEX_BC_JET_INFLOW || p.bc_x == IMEX_BC_REFLECTIVE || p.bc_y == IMEX_BC_REFLECTIVE ||
          p.bc_x == IMEX_BC_DIRICHLET || p.bc_y == IMEX_BC_DIRICHLET;
  return P;
}

static void prof_collect(imex_ctx* c) {
  Prof& P = c->prof;
  if (P.recs.empty()) { P.next = 0; return; }
  (void)hipStreamSynchronize(c->st);
  for (const ProfRec& r : P.recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, P.ev[r.e0], P.ev[r.e1]) != hipSuccess) continue;
    P.ms[r.cls] += ms; P.bytes[r.cls] += r.bytes; P.n[r.cls] += 1;
  }
  P.recs.clear();
  P.next = 0;
  P.gen++;
}
static int prof_begin(imex_ctx* c) {
  Prof& P = c->prof;
  if (!P.on) return -1;
  if (P.next + 2 > 65536) prof_collect(c);
  while ((int)P.ev.size() < P.next + 2) { hipEvent_t e; if (hipEventCreate(&e) != hipSuccess) return -1; P.ev.push_back(e); }
  int e0 = P.next;
  P.next += 2;
  (void)hipEventRecord(P.ev[e0], c->st);
  return e0;
}
static int prof_end(imex_ctx* c, int e0, int cls, double fields) {
  if (e0 < 0) return -1;
  Prof& P = c->prof;
  (void)hipEventRecord(P.ev[e0 + 1], c->st);
  P.recs.push_back(ProfRec{cls, e0, e0 + 1, fields * (double)c->cells * 8.0});
  return (int)P.recs.size() - 1;
}
#define LAUNCH(cls, fields, stmt) do { int e0_ = prof_begin(c); stmt; prof_end(c, e0_, (cls), (double)(fields)); } while (0)

template <class T>
static int dalloc(imex_ctx* c, T** out, size_t count) {
  void* ptr = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  if (hipMalloc(&ptr, bytes) != hipSuccess) return set_err(IMEX_ERR_NOMEM, "hipMalloc failed");
  c->allocs.push_back(ptr);
  c->bytes += (long long)bytes;
  *out = (T*)ptr;
  return 0;
}

static inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

// explicit-kernel tile candidates of the autotuner (cells: width, rows), in imex_tune_info order
constexpr int NEXT = 5;
static const int EX_TILES[NEXT][2] = {{16, 16}, {32, 8}, {64, 4}, {32, 16}, {64, 8}};

// Per-instance grid of a launch shape: (blocks per row, blocks).
static void shape_grid(const imex_ctx* c, int sh, const Dims& d, int& gx, int& nb) {
  long long tiles = cdiv(d.nx, TX) * cdiv(d.ny, TY);
  long long tot = (long long)d.NX * d.NY;
  gx = 1;
  switch (sh) {
    case SH_INT64x4: gx = (int)cdiv(d.nx, 64); nb = gx * (int)cdiv(d.ny, 4); break;
    case SH_EXPL: {
      const int tx = c->ex_tile, ty = c->ex_rows;
      gx = (int)cdiv(d.nx, tx); nb = gx * (int)cdiv(d.ny, ty); break;
    }
    case SH_PAD64x4: gx = (int)cdiv(d.NX, 64); nb = gx * (int)cdiv(d.NY, 4); break;
    case SH_PAD64x8: gx = (int)cdiv(d.NX, 64); nb = gx * (int)cdiv(d.NY, 8); break;
    case SH_INT2_64x4: gx = (int)cdiv(d.nx + 2, 64); nb = gx * (int)cdiv(d.ny + 2, 4); break;
    case SH_LIN_TOT: nb = (int)cdiv(tot, 256); gx = nb; break;
    case SH_LIN_N: nb = (int)cdiv(d.N, 256); gx = nb; break;
    case SH_START: nb = (int)std::min<long long>(cdiv(tot, 256), 1024); gx = nb; break;
    case SH_TILES: nb = (int)std::min<long long>(tiles, c->cap_tiles); gx = nb; break;
    case SH_TILES_P: nb = (int)std::min<long long>(tiles, c->cap_tiles_p); gx = nb; break;
    case SH_TILES_N: nb = (int)std::min<long long>(tiles, c->cap_tiles_n); gx = nb; break;
    case SH_TILES_ALL: nb = (int)std::min<long long>(tiles, 4096); gx = nb; break;
    case SH_VEC3: nb = (int)std::min<long long>(cdiv(3LL * d.N, 256), c->cap_vec); gx = nb; break;
    case SH_VEC1: nb = (int)std::min<long long>(cdiv(d.N, 256), c->cap_vec); gx = nb; break;
    case SH_MBIG: nb = (int)std::min<long long>(cdiv(d.N, 256), 512); gx = nb; break;
    case SH_PSTATS: nb = (int)std::min<long long>(cdiv(d.N, 256), 256); gx = nb; break;
    case SH_DIVB: nb = (int)std::min<long long>(cdiv(d.N, 256), 256); gx = nb; break;
    case SH_ONE: nb = 1; gx = 1; break;
    default: {   // ghost fills
      const int kind = sh - SH_FILL0;
      int lox = (!d.perx && kind == 1) ? GC - 1 : GC, hix = d.nx + GC - 1;
      int loy = (!d.pery && kind == 2) ? GC - 1 : GC, hiy = d.ny + GC - 1;
      long long total = (long long)loy * d.NX + (long long)(d.NY - 1 - hiy) * d.NX + (long long)(hiy - loy + 1) * (lox + d.NX - 1 - hix);
      nb = (int)std::max<long long>(1, cdiv(total, 256)); gx = nb; break;
    }
  }
  if (nb < 1) { nb = 1; gx = 1; }
}

// Block maps of every shape for the current instance sizes and caps: per-instance prefix sums and
// row widths, and for a batch the per-workgroup (instance, local block) table.
static long long map_blocks(imex_ctx* c) {
  long long tot = 0;
  for (int sh = 0; sh < NSHAPE; sh++)
    for (int i = 0; i < c->ni; i++) { int g, nb; shape_grid(c, sh, c->hI[i].d, g, nb); tot += nb; }
  return tot;
}
static int build_maps(imex_ctx* c) {
  const int ni = c->ni;
  HIPCHK(hipStreamSynchronize(c->st));      // no launch may be reading the maps being replaced
  std::vector<int> h((size_t)NSHAPE * (2 * ni + 1));
  std::vector<int2> blk;
  std::vector<size_t> boff(NSHAPE);
  for (int sh = 0; sh < NSHAPE; sh++) {
    int* pre = &h[(size_t)sh * (2 * ni + 1)];
    int* gx = pre + ni + 1;
    pre[0] = 0;
    boff[sh] = blk.size();
    for (int i = 0; i < ni; i++) {
      int g, nb;
      shape_grid(c, sh, c->hI[i].d, g, nb);
      pre[i + 1] = pre[i] + nb;
      gx[i] = g;
      if (ni > 1) for (int b = 0; b < nb; b++) blk.push_back(make_int2(i, b));
    }
    BMap& m = c->L.bm[sh];
    c->L.nblocks[sh] = pre[ni];
    m.pre = c->bmap_dev + (size_t)sh * (2 * ni + 1);
    m.gx = m.pre + ni + 1;
    m.ni = ni;
    m.gx0 = gx[0];
    m.nb0 = pre[1];
    m.blk = nullptr;
  }
  HIPCHK(hipMemcpy(c->bmap_dev, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
  if (ni > 1) {
    if ((long long)blk.size() > c->blk_cap) return set_err(IMEX_ERR_ARG, "block table capacity exceeded");
    for (int sh = 0; sh < NSHAPE; sh++) c->L.bm[sh].blk = c->blk_dev + boff[sh];
    HIPCHK(hipMemcpy(c->blk_dev, blk.data(), blk.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  return 0;
}


// precision of the inner GMRES of operator op (0 FP64, 1 FP32 with FP64 refinement), and of the
// solve at budget key k
static inline int op_prec(const imex_ctx* c, int op) { return (c->mixed_ops >> op) & 1; }
static inline int key_prec(const imex_ctx* c, int k) { return op_prec(c, (k / NIT) % 3 == 0 ? 0 : 1); }
static inline MagIds mag_ids(const imex_ctx* c) {
  MagIds m; m.Qe = c->cur.Qe; m.u = c->cur.u; m.pot = 0;
  m.refl = c->prm.bc_x == IMEX_BC_REFLECTIVE || c->prm.bc_y == IMEX_BC_REFLECTIVE;
  return m;
}
static inline MagIds pot_ids(const imex_ctx* c) { MagIds m = mag_ids(c); m.pot = 1; m.refl = 0; return m; }
static inline PIds p_ids(const imex_ctx* c, int jacobian, int precond) {
  PIds q; q.u = c->cur.u; q.jacobian = jacobian; q.precond = precond; q.energy = 0; return q;
}
static inline PIds e_ids(const imex_ctx* c, int precond) { PIds q = p_ids(c, 0, precond); q.energy = 1; return q; }
// profiler class of a Krylov launch: magnetic, pressure or energy operator
static inline int kcls(int op, const PIds& po, int mag, int p, int e) { return op == 0 ? mag : (po.energy ? e : p); }
// algorithmic fields of one residual / Arnoldi apply of the single-unknown operators (bench bytes)
static inline double p_res_fields(const PIds& po, int xz) {
  return po.energy ? (xz ? 2 : 4) + po.precond : (xz ? 2 : 5) + po.precond + po.jacobian;
}
static inline double p_apply_fields(const PIds& po, double vb, int j) {
  return po.energy ? 1 + po.precond + vb * (3 + j) : 5 + po.jacobian + vb * (1 + j);
}

// the enthalpy-ratio plane: the reference forms a neighbour's ratio from the neighbour's own state, which in
// the jet strip's ghost row is the inflow state (the producing kernels write that cell); elsewhere a copy
static const signed char hhat_jv[1] = {-2};
static void fill_planes(imex_ctx* c, std::initializer_list<std::pair<int, int>> planes, int kind = 0,
                        const signed char* jv = nullptr) {
  int b[8], k[8], n = 0;
  for (auto& pk : planes) { b[n] = pk.first; k[n] = pk.second; n++; }
  // a Dirichlet ghost's E follows its B (bound_cond_u refills the whole cell): a fill of a Grid's B
  // planes refills its E plane too
  const bool dirichlet = c->prm.bc_x == IMEX_BC_DIRICHLET || c->prm.bc_y == IMEX_BC_DIRICHLET;
  if (dirichlet && !jv && n < 8 && b[0] >= B_UPRI && b[0] <= B_QS) {
    bool hasB = false, hasE = false;
    for (int q = 0; q < n; q++) { hasB |= k[q] >= 5; hasE |= k[q] == 4; }
    if (hasB && !hasE) { b[n] = b[0]; k[n] = 4; n++; }
  }
  int e0 = prof_begin(c);
  launch_bc_fill(c->L, n, b, k, kind, jv);
  prof_end(c, e0, KC_BC, 0.0);
}

static inline int gkey(const imex_ctx* c) {
  return (c->stage_idx * 3 + c->sub_idx) * NIT + std::min(std::max(c->newton_it, 0), NIT - 1);
}
static inline int nkey(const imex_ctx* c) { return c->stage_idx * 3 + c->sub_idx; }

// Records the node that the launch just captured (single-stream capture: it is the only
// dependency of the next launch).
static void tag_last(imex_ctx* c, int kind, int key, int sw, int j) {
  if (!c->capturing) return;
  hipStreamCaptureStatus stt;
  const hipGraphNode_t* deps = nullptr;
  size_t nd = 0;
  if (hipStreamGetCaptureInfo_v2(c->st, &stt, nullptr, nullptr, &deps, &nd) != hipSuccess || nd != 1 || !deps) {
    c->tag_err = true;
    return;
  }
  c->tags.push_back(NodeTag{deps[0], (unsigned char)kind, (short)key, (short)sw, (short)j});
}

static void cur_reset(imex_ctx* c) {
  for (int k = 0; k < NGKEY; k++) c->kcur[k][0] = c->kcur[k][1] = -1;
  for (int k = 0; k < NNKEY; k++) c->ncur[k] = -1;
}
static void cur_gmres(imex_ctx* c, int key, int steps, int sweeps) {
  c->kcur[key][0] = std::max(c->kcur[key][0], steps);
  c->kcur[key][1] = std::max(c->kcur[key][1], sweeps);
}

// ---------------------------------------------------------------- GMRES, budgeted schedule
// GMRES_Solver (implicit_operators.cpp:823-911) with a fixed number of launches: the FP64 form
// runs nsw restart cycles (residual, B Arnoldi steps, x update; implicit_operators.cpp:899-910),
// the last of which must have converged; the mixed form runs nsw refinement sweeps (FP64
// residual, B FP32 Arnoldi steps, x update) and the FP64 residual that must have converged.
// The kernels are the ones the host-polled loop below uses, so a solve that converges within
// its budget gives the same result bit for bit.  Launches past convergence exit at once; a solve
// still unconverged at the end of its budget flags the step for a rerun.
static int run_gmres_budget(imex_ctx* c, int op, const MagIds& mo, const PIds& po, int ix, int ib,
                            int x_is_zero, int skip, int parent, int it) {
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  const int nc = op == 0 ? 3 : 1;
  const int prec = op_prec(c, op);
  const Launch& L = c->L;
  BudgetRec r;
  r.kind = 0; r.op = op; r.parent = parent; r.it = it; r.key = gkey(c);
  r.B = std::min(m, std::max(1, c->tab.B[r.key]));
  r.nsw = std::min(MAXSWEEP - 1, std::max(1, c->tab.S[r.key]));
  r.slot = c->nslot; c->nslot += 2 + r.nsw;
  for (int sw = 0; sw <= r.nsw; sw++) {
    if (sw == r.nsw && !prec) break;
    const int xz = (x_is_zero && sw == 0) ? 1 : 0;
    const double res_fields = op == 0 ? (xz ? 6 : 13) : p_res_fields(po, xz);
    const int k = r.key;
    int e0 = prof_begin(c);
    gmres_residual(L, op, prec, mo, po, c->P, ix, ib, sw, xz, skip);
    r.rres.push_back(prof_end(c, e0, kcls(op, po, KC_MAG_RES, KC_P_RES, KC_E_RES), res_fields));
    tag_last(c, TG_RES, k, sw, 0);
    e0 = prof_begin(c);
    gmres_residual_fin(L, op, prec, sw, skip, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol);
    r.rresf.push_back(prof_end(c, e0, KC_KFIN, 0));
    tag_last(c, TG_RES, k, sw, 0);
    if (sw == r.nsw) break;
    const double vb = prec ? 0.5 : 1.0;       // bytes of a basis field relative to an FP64 field
    for (int j = 0; j < r.B; j++) {
      e0 = prof_begin(c);
      arnoldi_apply(L, op, prec, mo, po, c->P, j);
      r.ra.push_back(prof_end(c, e0, kcls(op, po, KC_MAG_APPLY, KC_P_APPLY, KC_E_APPLY),
                              op == 0 ? 4 + vb * (9 + 3 * j) : p_apply_fields(po, vb, j)));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_apply_fin(L, op, j);
      r.rfa.push_back(prof_end(c, e0, KC_KFIN, 0));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_update(L, op, prec, j, m);
      r.ru.push_back(prof_end(c, e0, kcls(op, po, KC_MAG_UPDATE, KC_P_UPDATE, KC_E_UPDATE), vb * (op == 0 ? 9 + 3 * j : 3 + j)));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_update_fin(L, op, prec, j, m);
      r.rfu.push_back(prof_end(c, e0, KC_KFIN, 0));
      tag_last(c, TG_ARN, k, sw, j);
    }
    e0 = prof_begin(c);
    gmres_xupdate(L, op, prec, ix, x_is_zero && sw == 0, skip);
    r.rx.push_back(prof_end(c, e0, kcls(op, po, KC_MAG_XUPD, KC_P_XUPD, KC_E_XUPD), (double)nc * (2 + vb * r.B)));
    tag_last(c, TG_XUPD, k, sw, 0);
  }
  k_budget_gmres_check<<<dim3(L.nblocks[SH_ONE]), 64, 0, c->st>>>(L.IT, L.bm[SH_ONE], r.slot, prec, r.nsw);
  tag_last(c, TG_GCHK, r.key, 0, 0);
  c->brecs.push_back(r);
  return 0;
}

// ---------------------------------------------------------------- GMRES host loop
// GMRES_Solver (implicit_operators.cpp:793-919).  Arnoldi steps are queued in chunks; the
// convergence flag of chunk k is read (async copy + event) while chunk k+1 is already queued.
// Host polling reads the control block of the context's only instance.
static int run_gmres(imex_ctx* c, int op, const MagIds& mo, const PIds& po, int ix, int ib, int x_is_zero,
                     imex_solve_info* info, long long* it_acc, long long* rs_acc, int skip = SKIP_NONE,
                     int first_chunk = 0) {
  if (c->budget) return run_gmres_budget(c, op, mo, po, ix, ib, x_is_zero, skip, -1, -1);
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  const int nc = op == 0 ? 3 : 1;
  const long long dim = (long long)nc * c->d0().N;
  const Launch& L = c->L;
  KrylovDev* Kd = c->hI[0].K;
  // The reference allows dim restart cycles (implicit_operators.cpp:802); the host loop is
  // additionally capped at 4000 cycles (100k Arnoldi steps at m = 25) and raises
  // IMEX_WARN_GMRES_MAXOUTER when it gives up.
  const long long outmax = std::min<long long>(c->prm.gmres_max_outer > 0 ? c->prm.gmres_max_outer : dim, 4000);
  const int chunk = 2;
  const int prec = op_prec(c, op);
  const double vb = prec ? 0.5 : 1.0;       // bytes of a basis field relative to an FP64 field
  if (first_chunk <= 0) first_chunk = c->pred[op];
  int total = 0, rc = 0, maxsweep = 0;
  int converged = 0;
  long long outer = 0;
  double res0 = 0, res = 0;
  for (outer = 0; outer < outmax; outer++) {
    const int xz = (x_is_zero && outer == 0) ? 1 : 0;
    const double res_fields = op == 0 ? (xz ? 6 : 13) : p_res_fields(po, xz);
    const int sk = (outer == 0 || prec) ? skip : SKIP_NONE;
    LAUNCH(kcls(op, po, KC_MAG_RES, KC_P_RES, KC_E_RES), res_fields, gmres_residual(L, op, prec, mo, po, c->P, ix, ib, (int)outer, xz, sk));
    LAUNCH(KC_KFIN, 0, gmres_residual_fin(L, op, prec, (int)outer, sk, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol));
    const long long gen0 = c->prof.gen;
    int recs_a[MAXM], recs_u[MAXM], recs_fa[MAXM], recs_fu[MAXM];
    int j = 0, slot = 0, last = -1;
    bool stop = false;
    while (!stop) {
      if (j >= m) break;                       // whole cycle queued; the last chunk decides
      int jend = std::min(m, j + (j == 0 ? std::max(first_chunk, 1) : chunk));
      for (; j < jend; j++) {
        int e0 = prof_begin(c);
        arnoldi_apply(L, op, prec, mo, po, c->P, j);
        recs_a[j] = prof_end(c, e0, kcls(op, po, KC_MAG_APPLY, KC_P_APPLY, KC_E_APPLY),
                             op == 0 ? 4 + vb * (9 + 3 * j) : p_apply_fields(po, vb, j));
        e0 = prof_begin(c);
        arnoldi_apply_fin(L, op, j);
        recs_fa[j] = prof_end(c, e0, KC_KFIN, 0);
        e0 = prof_begin(c);
        arnoldi_update(L, op, prec, j, m);
        recs_u[j] = prof_end(c, e0, kcls(op, po, KC_MAG_UPDATE, KC_P_UPDATE, KC_E_UPDATE), vb * (op == 0 ? 9 + 3 * j : 3 + j));
        e0 = prof_begin(c);
        arnoldi_update_fin(L, op, prec, j, m);
        recs_fu[j] = prof_end(c, e0, KC_KFIN, 0);
      }
      HIPCHK(hipMemcpyAsync(&c->Kh[slot].tol, &Kd->tol, sizeof(KrylovDev) - offsetof(KrylovDev, tol), hipMemcpyDeviceToHost, c->st));
      HIPCHK(hipEventRecord(c->ev[slot], c->st));
      int prev = last;
      last = slot;
      slot ^= 1;
      if (prev >= 0) {                         // the chunk before this one has had time to finish
        HIPCHK(hipEventSynchronize(c->ev[prev]));
        if (c->Kh[prev].done) stop = true;
      }
    }
    int pending = last;
    HIPCHK(hipEventSynchronize(c->ev[pending]));
    KrylovDev& ks = c->Kh[pending];
    int rx;
    {
      int e0 = prof_begin(c);
      gmres_xupdate(L, op, prec, ix, xz, sk);
      rx = prof_end(c, e0, kcls(op, po, KC_MAG_XUPD, KC_P_XUPD, KC_E_XUPD), (double)nc * (2 + vb * (ks.jlast + 1)));
    }
    if (c->prof.on && c->prof.gen == gen0) {      // launches queued past convergence exited at once
      auto idle = [&](int r) { if (r >= 0) { c->prof.recs[r].cls = KC_EARLY_EXIT; c->prof.recs[r].bytes = 0; } };
      for (int jj = ks.steps; jj < j; jj++) { idle(recs_a[jj]); idle(recs_u[jj]); idle(recs_fa[jj]); idle(recs_fu[jj]); }
      if (ks.jlast < 0 && (!xz || (sk == SKIP_NEWTON && ks.newton_conv))) idle(rx);   // x = 0 is written
    }
    total += ks.steps;
    maxsweep = std::max(maxsweep, ks.steps);
    c->Klast = ks;
    if (outer == 0) res0 = ks.res0;
    res = ks.res;
    if (ks.converged) { converged = 1; break; }
  }
  if (!converged) { rc |= IMEX_WARN_GMRES_MAXOUTER; c->stats.warnings |= IMEX_WARN_GMRES_MAXOUTER; }
  if (it_acc) *it_acc += total;
  if (prec) c->stats.mixed_refinements += outer;
  else if (rs_acc) *rs_acc += outer;
  if (total > 0 && outer == 0) c->pred[op] = total;
  // budgets learn the longest sweep (or restart cycle) and the number of sweeps (cycles)
  const int nsweeps = prec ? (int)outer : (int)std::min<long long>(outer + 1, MAXSWEEP);
  c->step_mx[op] = std::max(c->step_mx[op], maxsweep);
  c->step_mx[3 + op] = std::max(c->step_mx[3 + op], nsweeps);
  cur_gmres(c, gkey(c), maxsweep, nsweeps);
  if (info) { info->iters = total; info->restarts = (int)outer; info->res0 = res0; info->res = res; info->status = rc; }
  return rc;
}

// ---------------------------------------------------------------- Newton helpers
// Mbig (Cor. 1) before a Newton solve: from k_p_rhs's per-block partials when that kernel has just
// formed b (mbig_ready), else from a reduction over b.
static void mbig_launch(imex_ctx* c, int ib, bool mbig_ready) {
  if (mbig_ready) newton_mbig_fin_prhs(c->L, c->P);
  else newton_mbig(c->L, c->P, ib, c->cur.u);
}
static inline double pst_threshold(const imex_ctx* c) { return c->prm.theta * c->prm.p_floor - 1e-14; }

// ---------------------------------------------------------------- nested Newton, budgeted schedule
// Pressure_NestedNewton (implicit_operators.cpp:625-738) with a fixed number of Jacobian solves:
// residual, solve, update per iteration, then the residual that must have converged.  Once a
// residual has converged, the later residuals, fills, solves and updates of this Newton solve
// skip on the device.
static int run_newton_budget(imex_ctx* c, int ib, int ip, bool mbig_ready) {
  const Launch& L = c->L;
  if (!c->prm.floor_on && c->prm.eos == IMEX_EOS_IDEAL)
    return run_gmres_budget(c, 1, mag_ids(c), p_ids(c, 0, 1), ip, ib, 0, SKIP_NONE, -1, -1);
  if (c->prm.floor_on) LAUNCH(KC_MISC, 1, mbig_launch(c, ib, mbig_ready));
  const int NB = std::min(c->prm.newton_max, std::max(1, c->tab.N[nkey(c)]));
  const int self = (int)c->brecs.size();
  {
    BudgetRec nr;
    nr.kind = 1; nr.op = 1; nr.slot = c->nslot++; nr.key = nkey(c);
    c->brecs.push_back(nr);
  }
  for (int it = 0; it <= NB; it++) {
    c->newton_it = it;
    const int nk = c->brecs[self].key;
    int e0 = prof_begin(c);
    newton_residual(L, p_ids(c, 0, 0), c->P, ip, ib, it, c->prm.newton_tol);
    int rr = prof_end(c, e0, KC_NEWTON_RES, 6);
    c->brecs[self].rnres.push_back(rr);
    tag_last(c, TG_NRES, nk, 0, it);
    e0 = prof_begin(c);
    newton_residual_fin(L, it, c->prm.newton_tol);
    rr = prof_end(c, e0, KC_KFIN, 0);
    c->brecs[self].rnresf.push_back(rr);
    tag_last(c, TG_NRES, nk, 0, it);
    if (it == NB) break;
    if (c->P.eos != IMEX_EOS_IDEAL) {
      vslope_launch(L, c->P, ip, c->cur.u);
      tag_last(c, TG_NIT, nk, 0, it);
    }
    // the solve starts from dp = 0: its first x update writes dp instead of reading it
    run_gmres_budget(c, 1, mag_ids(c), p_ids(c, 1, 1), B_DP, B_NG, 1, SKIP_NEWTON, self, it);
    c->newton_it = 0;
    e0 = prof_begin(c);
    axpy_launch(L, ip, B_DP, SKIP_NEWTON);
    rr = prof_end(c, e0, KC_MISC, 3);
    c->brecs[self].raxpy.push_back(rr);
    tag_last(c, TG_NIT, nk, 0, it);
  }
  // the last residual that ran (a converged one, or the closing residual NB) saw the final p
  k_budget_newton_check<<<dim3(L.nblocks[SH_ONE]), c->prm.floor_on ? 1024 : 64, 0, c->st>>>(L.IT, L.bm[SH_ONE], c->brecs[self].slot, L.bm[SH_TILES_N],
                                                                   c->prm.floor_on, pst_threshold(c));
  return 0;
}

// ---------------------------------------------------------------- nested Newton (implicit_operators.cpp:625-738)
static int run_newton(imex_ctx* c, int ib, int ip, imex_solve_info* info, bool mbig_ready) {
  if (c->budget) return run_newton_budget(c, ib, ip, mbig_ready);
  int rc = 0;
  const Launch& L = c->L;
  if (!c->prm.floor_on && c->prm.eos == IMEX_EOS_IDEAL) {
    rc |= run_gmres(c, 1, mag_ids(c), p_ids(c, 0, 1), ip, ib, 0, info, &c->stats.p_gmres_iters, &c->stats.p_gmres_restarts);
    c->stats.p_solves++;
    return rc;
  }
  if (c->prm.floor_on) LAUNCH(KC_MISC, 1, mbig_launch(c, ib, mbig_ready));
  int it;
  double r0 = 0, r = 0;
  for (it = 0; it < c->prm.newton_max; it++) {
    c->newton_it = it;
    // residual, then the Jacobian solve; the solve is skipped on the device when the residual
    // has converged, and the host learns it from the GMRES status it polls anyway
    LAUNCH(KC_NEWTON_RES, 6, newton_residual(L, p_ids(c, 0, 0), c->P, ip, ib, it, c->prm.newton_tol));
    LAUNCH(KC_KFIN, 0, newton_residual_fin(L, it, c->prm.newton_tol));
    if (c->P.eos != IMEX_EOS_IDEAL) vslope_launch(L, c->P, ip, c->cur.u);
    long long its = 0, rss = 0;     // dp = 0 is written by the solve's first x update
    // Newton usually needs pred_newton Jacobian solves; past that the solve is expected to be
    // skipped, so only a one-step chunk is queued speculatively
    int first = it < c->pred_newton ? c->pred[1] : 1;
    rc |= run_gmres(c, 1, mag_ids(c), p_ids(c, 1, 1), B_DP, B_NG, 1, nullptr, &its, &rss, SKIP_NEWTON, first);
    if (it == 0) r0 = c->Klast.newton_res0;
    r = c->Klast.newton_res;
    if (c->Klast.newton_conv) { rc &= ~IMEX_WARN_GMRES_MAXOUTER; break; }
    c->stats.p_gmres_iters += its; c->stats.p_gmres_restarts += rss;
    c->stats.p_solves++;
    LAUNCH(KC_MISC, 3, axpy_launch(L, ip, B_DP, SKIP_NEWTON));
  }
  c->newton_it = 0;
  c->stats.newton_iters += it;
  c->ncur[nkey(c)] = std::max(c->ncur[nkey(c)], it);
  c->pred_newton = std::max(1, it);
  c->step_mx[2] = std::max(c->step_mx[2], it);
  if (it == c->prm.newton_max) { rc |= IMEX_WARN_NEWTON_MAXITER; c->stats.warnings |= IMEX_WARN_NEWTON_MAXITER; }
  // a converged solve's last residual saw the final p; at the iteration cap p moved after it
  if (c->prm.floor_on) {
    if (it == c->prm.newton_max) LAUNCH(KC_MISC, 1, pstats_launch(L, ip, c->prm.p_floor, pst_threshold(c)));
    else LAUNCH(KC_MISC, 0, pstats_fin_newton(L, pst_threshold(c)));
  }
  if (info) { info->iters = it; info->res0 = r0; info->res = r; info->status = rc; info->restarts = 0; }
  return rc;
}

// ---------------------------------------------------------------- stage pieces
static void ensure_qe(imex_ctx* c) {
  if (c->qe_ready) return;
  LAUNCH(KC_QEPREP, 10, launch_qe_prep(c->L, c->P, c->cur.Qe));
  LAUNCH(KC_SHOCK, 7, launch_shock(c->L, c->P, c->cur.Qe));
  c->qe_ready = true;
}

static int sub_explicit_conv(imex_ctx* c) {
  LAUNCH(KC_EXPL_CONV, 13, launch_explicit(c->L, c->P, 0, c->cur.Qe, c->cur.Qs, c->cur.u, c->mixed_flux, c->ex_tile, c->ex_rows));
  const int u = c->cur.u;
  fill_planes(c, {{u, 0}, {u, 1}, {u, 2}, {u, 3}});
  c->hhat_ready = false;
  return 0;
}

// emf: CollectMagFluxes (ct_flux.cpp:5-147), whose fluxes only the CT updates read; a whole step with
// PotentialA or NoDivFree leaves it out (MagSubsystem_UpdateB still computes it when called alone)
static int sub_mag_update_b(imex_ctx* c, bool emf = true) {
  c->sub_idx = 0;
  ensure_qe(c);
  LAUNCH(KC_MAG_RHS, 23, launch_mag_rhs(c->L, c->P, c->cur.u, c->cur.Qe, c->cur.Qs, B_BK, B_XK));
  fill_planes(c, {{B_HHAT, 0}}, 0, hhat_jv);
  c->hhat_ready = true;
  int rc = run_gmres(c, 0, mag_ids(c), p_ids(c, 0, 0), B_XK, B_BK, 0, nullptr, &c->stats.mag_gmres_iters, &c->stats.mag_gmres_restarts);
  if (!c->budget) c->stats.mag_solves++;
  LAUNCH(KC_MISC, 6, launch_mag_writeback(c->L, c->cur.u));
  if (emf) LAUNCH(KC_EMF, 12, launch_emf_flux(c->L, c->cur.u, c->cur.Qe));
  return rc;
}

// MagSubsystem_UpdateA (imex_subsystems.cpp:5-74): RHS and initial guess (from `ia0`: the stage's A,
// which a step starts as a copy of A_s), the potential operator's viscosity from Qs, one GMRES solve
// on the magnetic slot, then A, bound_cond_A and B = curl A.
static int sub_mag_update_a(imex_ctx* c, int ia0) {
  c->sub_idx = 0;
  ensure_qe(c);
  LAUNCH(KC_MAG_RHS, 17, launch_a_rhs(c->L, c->cur.u, c->cur.Qe, c->cur.As, ia0, B_BK, B_XK));
  LAUNCH(KC_MISC, 5, launch_pot_visc(c->L, c->cur.Qs));
  c->hhat_ready = false;
  int rc = run_gmres(c, 0, pot_ids(c), p_ids(c, 0, 0), B_XK, B_BK, 0, nullptr, &c->stats.mag_gmres_iters, &c->stats.mag_gmres_restarts);
  if (!c->budget) c->stats.mag_solves++;
  LAUNCH(KC_A_UPDATE, 12, launch_a_update(c->L, B_XK, c->cur.Ao, c->cur.u));
  const int u = c->cur.u;
  fill_planes(c, {{u, 5}, {u, 6}, {u, 7}});
  return rc;
}

// Unstaggered_CT_Consistent (constrained_transport.cpp:502-540); A_x, A_y of the stage from `ia01`
static int sub_unstaggered_ct(imex_ctx* c, int ia01) {
  LAUNCH(KC_CT, 16, launch_unstaggered_ct(c->L, c->cur.u, c->cur.As, ia01, c->cur.Ao));
  const int u = c->cur.u;
  fill_planes(c, {{u, 5}, {u, 6}});
  return 0;
}

static int sub_explicit_press(imex_ctx* c) {
  // algorithmic planes: Qe 0..4 (all eight with PrimState), Qs 1..4, the cell B, u 1..4 written
  LAUNCH(KC_EXPL_PRESS, c->P.prim ? 19 : 16, launch_explicit(c->L, c->P, 1, c->cur.Qe, c->cur.Qs, c->cur.u, c->mixed_flux, c->ex_tile, c->ex_rows));
  const int u = c->cur.u;
  fill_planes(c, {{u, 1}, {u, 2}, {u, 3}, {u, 4}});
  return 0;
}

static void ensure_hhat(imex_ctx* c) {
  if (c->hhat_ready) return;
  // hhat is produced by the magnetic RHS kernel; run it for its side output when the
  // pressure subsystem is called on its own (its b and x0 go to scratch)
  ensure_qe(c);
  launch_mag_rhs(c->L, c->P, c->cur.u, c->cur.Qe, c->cur.Qs, B_TMPN, B_W);
  fill_planes(c, {{B_HHAT, 0}}, 0, hhat_jv);
  c->hhat_ready = true;
}

static int sub_pressure(imex_ctx* c) {
  ensure_qe(c);
  ensure_hhat(c);
  const int u = c->cur.u;
  if (c->P.eos != IMEX_EOS_IDEAL) launch_visc_rk(c->L, c->P, u);
  LAUNCH(KC_P_RHS, 15, launch_p_rhs(c->L, c->P, 0, u));
  static const signed char mom_jv[2] = {1, 2};       // rhov*_star in the jet strip: (0, rho v) (imex_subsystems.cpp:107-118)
  fill_planes(c, {{B_MXS, 0}, {B_MYS, 0}}, 0, mom_jv);
  c->sub_idx = 1;
  int rc = run_newton(c, B_PB, B_P, nullptr, true);      // k_p_rhs left the Mbig partials of b
  if (c->hook) c->hook(c, HK_P0);
  LAUNCH(KC_P_MOM, 5, launch_p_momentum(c->L, u));
  fill_planes(c, {{u, 1}, {u, 2}});
  if (c->hook) c->hook(c, HK_MOM);
  LAUNCH(KC_P_RHS, 12, launch_p_rhs(c->L, c->P, 1, u));
  c->sub_idx = 2;
  rc |= run_newton(c, B_PB, B_P, nullptr, true);
  if (c->hook) c->hook(c, HK_P1);
  LAUNCH(KC_P_ENERGY, 9, launch_p_energy(c->L, c->P, u));
  fill_planes(c, {{u, 4}});
  if (c->hook) c->hook(c, HK_PR);
  return rc;
}

// B slots <- cell B~, then Energy_flux_Update (main.cpp:448-474, flux.cpp:165-200)
static int sub_explicit_energy(imex_ctx* c) {
  LAUNCH(KC_EXPL_ENERGY, c->P.prim ? 19 : 16, launch_explicit(c->L, c->P, 2, c->cur.Qe, c->cur.Qs, c->cur.u, c->mixed_flux, c->ex_tile, c->ex_rows));
  const int u = c->cur.u;
  fill_planes(c, {{u, 1}, {u, 2}, {u, 3}, {u, 4}});
  return 0;
}

// PreSubsystem_Energy (imex_subsystems.cpp:415-534): coefficient plane, RHS, initial guess and precE
// divisor, one preconditioned GMRES solve (the pressure solve's slot: budgets, precision), then E
// and the momentum update.  The coefficient overwrites the pressure enthalpy plane.
static int prep_energy(imex_ctx* c) {
  ensure_qe(c);
  LAUNCH(KC_E_RHS, 14, launch_e_rhs(c->L, c->P, c->cur.u, c->cur.Qe));
  fill_planes(c, {{B_HHAT, 0}}, 0, hhat_jv);
  c->hhat_ready = false;
  return 0;
}
static int sub_energy(imex_ctx* c) {
  prep_energy(c);
  c->sub_idx = 1;
  int rc = run_gmres(c, 1, mag_ids(c), e_ids(c, 1), B_P, B_PB, 0, nullptr, &c->stats.p_gmres_iters, &c->stats.p_gmres_restarts);
  if (!c->budget) c->stats.p_solves++;
  const int u = c->cur.u;
  LAUNCH(KC_E_MOM, 6, launch_e_update(c->L, c->P, u));
  fill_planes(c, {{u, 1}, {u, 2}, {u, 4}});
  return rc;
}

static int sub_ct(imex_ctx* c) {
  LAUNCH(KC_CT, 15, launch_ct(c->L, c->P, c->prm.ct_mode == IMEX_CT_PLAIN, c->cur.u, c->cur.bxqs, c->cur.byqs, c->cur.bxo, c->cur.byo));
  return 0;
}

static int run_stage(imex_ctx* c) {
  int rc = 0;
  c->qe_ready = false; c->hhat_ready = false;
  ensure_qe(c);
  if (c->hook) c->hook(c, HK_IN);
  rc |= sub_explicit_conv(c);
  if (c->hook) c->hook(c, HK_CONV);
  const int mf = c->prm.magfield;                  // main.cpp:424-434
  if (mf == IMEX_MAGFIELD_POTENTIAL_A) rc |= sub_mag_update_a(c, c->cur.As);
  else rc |= sub_mag_update_b(c, mf == IMEX_MAGFIELD_STAGGER || mf == IMEX_MAGFIELD_UNSTAGGER);
  if (c->hook) c->hook(c, HK_MAG);
  if (c->prm.formula == IMEX_FORMULA_ENERGY) {     // main.cpp:470-486
    rc |= sub_explicit_energy(c);
    rc |= sub_energy(c);
  } else {
    rc |= sub_explicit_press(c);
    if (c->hook) c->hook(c, HK_PRESS);
    rc |= sub_pressure(c);
  }
  if (mf == IMEX_MAGFIELD_STAGGER) rc |= sub_ct(c);   // main.cpp:497-510
  else {
    if (mf == IMEX_MAGFIELD_UNSTAGGER) rc |= sub_unstaggered_ct(c, c->cur.As);
    LAUNCH(KC_MISC, 4, launch_copy_faces(c->L, c->cur.bxqs, c->cur.byqs, c->cur.bxo, c->cur.byo));
  }
  if (c->hook) c->hook(c, HK_CT);
  return rc;
}

static void bind_stage0(imex_ctx* c) {
  c->stage_idx = 0;
  c->cur.u = B_U1; c->cur.Qe = B_UN; c->cur.Qs = B_UN;
  c->cur.bxqs = c->iBxn; c->cur.byqs = c->iByn; c->cur.bxo = B_BX1; c->cur.byo = B_BY1;
  c->cur.As = c->iAn; c->cur.Ao = B_A1;
}
static void bind_stage1(imex_ctx* c) {
  c->stage_idx = 1;
  c->cur.u = c->iUo; c->cur.Qe = B_QE; c->cur.Qs = B_QS;
  c->cur.bxqs = B_BXQS; c->cur.byqs = B_BYQS; c->cur.bxo = c->iBx2; c->cur.byo = c->iBy2;
  c->cur.As = B_AS; c->cur.Ao = c->iAo;
}

static int step_start(imex_ctx* c) {
  const int in = c->state_mode ? c->iUs : B_UPRI;
  LAUNCH(KC_START, 16, launch_step_start(c->L, c->P, c->state_mode, in));
  return 0;
}

// Everything a step queues; the step's input (Us or u_pri, bxn, byn) is only read, so a
// discarded step can be queued again from the same input.
static int step_body(imex_ctx* c) {
  int rc = 0;
  step_start(c);
  if (c->hook) c->hook(c, HK_START);
  bind_stage0(c);
  rc |= run_stage(c);
  LAUNCH(KC_COMBINE, 38, launch_stage_combine(c->L, c->iBxn, c->iByn));
  if (c->needA) LAUNCH(KC_COMBINE, 9, launch_a_combine(c->L, c->iAn, B_A1, B_AS));
  bind_stage1(c);
  rc |= run_stage(c);
  return rc;
}

static void step_commit(imex_ctx* c) {
  std::swap(c->iBxn, c->iBx2); std::swap(c->iByn, c->iBy2); std::swap(c->iUs, c->iUo); std::swap(c->iAn, c->iAo);
  c->state_mode = 1;
  c->stats.steps++;
}


static void hist_push(imex_ctx* c, const int mx[NBK]) {
  for (int k = 0; k < NBK; k++) c->hist[k][c->hist_pos] = mx[k];
  std::memcpy(c->khist[c->hist_pos], c->kcur, sizeof(c->kcur));
  std::memcpy(c->nhist[c->hist_pos], c->ncur, sizeof(c->ncur));
  c->hist_pos = (c->hist_pos + 1) % BHIST;
  c->hist_n = std::min(c->hist_n + 1, BHIST);
}

// Budgets, per solve position: the largest count that position needed in the recent steps
// (or, for a position not seen lately, the largest of its kind), plus budget_margin Arnoldi
// steps per GMRES sweep (restart cycle), plus the extra margins that recent overruns have raised.
// Returns false (host scheduling) until there is history, or when a recent solve needed more
// sweeps or restart cycles than a budget can hold.  A batch has no host scheduling: it starts
// from generous budgets and grows them when a step outruns them.
static bool choose_budgets(imex_ctx* c) {
  if (c->taping) return false;
  if (c->ni == 1 && (c->prm.schedule == IMEX_SCHED_HOST || c->hist_n == 0)) return false;
  if (c->ni == 1 && c->prm.schedule == IMEX_SCHED_AUTO && !c->auto_budget) return false;
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  int mx[NBK] = {0, 0, 0, 0, 0};
  if (c->hist_n == 0) { mx[0] = mx[1] = std::min(m - 1, 11); mx[2] = 2; mx[3] = mx[4] = 3; }
  for (int i = 0; i < c->hist_n; i++) for (int k = 0; k < NBK; k++) mx[k] = std::max(mx[k], c->hist[k][i]);
  if (c->ni == 1 && (mx[3] >= MAXSWEEP - 1 || mx[4] >= MAXSWEEP - 1)) return false;   // too many sweeps or cycles
  for (int k = 0; k < NGKEY; k++) {
    const int op = (k / NIT) % 3 == 0 ? 0 : 1;
    int st = -1, sw = -1;
    for (int i = 0; i < c->hist_n; i++) { st = std::max(st, c->khist[i][k][0]); sw = std::max(sw, c->khist[i][k][1]); }
    if (st < 0) {            // a Newton iteration not reached lately: the largest of its Newton solve
      const int k0 = k - k % NIT;
      for (int i = 0; i < c->hist_n; i++)
        for (int q = k0; q < k0 + NIT; q++) { st = std::max(st, c->khist[i][q][0]); sw = std::max(sw, c->khist[i][q][1]); }
    }
    if (st < 0) { st = mx[op]; sw = mx[3 + op]; }
    c->tab.B[k] = std::min(m, std::max(1, st + c->prm.budget_margin + c->margin_extra));
    c->tab.S[k] = std::min(MAXSWEEP - 1, std::max(1, sw + c->sweep_extra));
  }
  for (int k = 0; k < NNKEY; k++) {
    int nb = -1;
    for (int i = 0; i < c->hist_n; i++) nb = std::max(nb, c->nhist[i][k]);
    c->tab.N[k] = std::min(std::min(c->prm.newton_max, NIT - 1), std::max(1, (nb < 0 ? mx[2] : nb) + c->newton_extra));
  }
  return true;
}

// Reads the step's logs (all instances): returns the failure bits (1 GMRES, 2 Newton) if a solve
// outran its budget, else 0 after booking the iteration counts (sums over instances, maxima for
// the budgets) and, with the profiler on, moving launches that found their solve converged to
// early_exit.
static int budget_finish(imex_ctx* c, long long gen0) {
  HIPCHK(hipMemcpyAsync(c->ctlh, c->ctl_all, (size_t)c->ni * CTLN * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipEventRecord(c->ev[0], c->st));
  HIPCHK(hipEventSynchronize(c->ev[0]));
  int fail = 0;
  for (int i = 0; i < c->ni; i++) fail |= c->ctlh[(size_t)i * CTLN];
  if (fail) return fail;
  const bool fix = c->prof.on && c->prof.gen == gen0 && c->ni == 1;
  auto idle = [&](int r) { if (fix && r >= 0) { c->prof.recs[r].cls = KC_EARLY_EXIT; c->prof.recs[r].bytes = 0; } };
  int mx[NBK] = {0, 0, 0, 0, 0};
  for (int i = 0; i < c->ni; i++) {
    const int* ctl = c->ctlh + (size_t)i * CTLN;
    for (const BudgetRec& r : c->brecs) {
      if (r.kind == 1) {
        const int nf = ctl[r.slot];
        c->stats.newton_iters += nf;
        mx[2] = std::max(mx[2], nf);
        c->ncur[r.key] = std::max(c->ncur[r.key], nf);
        for (size_t q = 0; q < r.rnres.size(); q++) if ((int)q > nf) { idle(r.rnres[q]); idle(r.rnresf[q]); }
        for (size_t q = 0; q < r.raxpy.size(); q++) if ((int)q >= nf) idle(r.raxpy[q]);
        continue;
      }
      const int steps = ctl[r.slot], sweeps = ctl[r.slot + 1];
      const bool skipped = r.parent >= 0 && r.it >= ctl[c->brecs[r.parent].slot];
      const int rp0 = op_prec(c, r.op);
      for (int sw = 0; sw < (int)r.rres.size(); sw++) {
        const bool live = !skipped && (rp0 ? sw <= sweeps : sw < sweeps);   // residuals that ran
        if (!live) { idle(r.rres[sw]); idle(r.rresf[sw]); }
      }
      for (int sw = 0; sw < r.nsw; sw++) {
        const int ss = (!skipped && sw < sweeps) ? ctl[r.slot + 2 + sw] : 0;   // Arnoldi steps of this sweep
        for (int j = ss; j < r.B; j++) {
          const size_t q = (size_t)sw * r.B + j;
          if (q < r.ra.size()) { idle(r.ra[q]); idle(r.ru[q]); idle(r.rfa[q]); idle(r.rfu[q]); }
        }
        if (ss == 0 && sw < (int)r.rx.size()) idle(r.rx[sw]);
        if (!skipped && sw < sweeps) { mx[r.op] = std::max(mx[r.op], ss); cur_gmres(c, r.key, ss, 0); }
      }
      if (skipped) continue;
      mx[3 + r.op] = std::max(mx[3 + r.op], sweeps);
      const int rp = op_prec(c, r.op);
      cur_gmres(c, r.key, 0, sweeps);
      if (rp) c->stats.mixed_refinements += sweeps;
      else if (r.op == 0) c->stats.mag_gmres_restarts += std::max(0, sweeps - 1);
      else c->stats.p_gmres_restarts += std::max(0, sweeps - 1);
      if (r.op == 0) { c->stats.mag_solves++; c->stats.mag_gmres_iters += steps; }
      else { c->stats.p_solves++; c->stats.p_gmres_iters += steps; }
    }
  }
  hist_push(c, mx);
  return 0;
}

// Whether a tagged node runs under budget table tb.
static bool tag_enabled(const imex_ctx* c, const NodeTag& t, const BudTab& tb) {
  if (t.kind == TG_NIT) return t.j < tb.N[t.key];
  if (t.kind == TG_NRES) return t.j <= tb.N[t.key];
  const int sub = (t.key / NIT) % 3, it = t.key % NIT;
  const bool live = sub == 0 || it < tb.N[t.key / NIT];      // the Newton iteration holding the solve runs
  if (!live) return false;
  const int mixed = key_prec(c, t.key);
  const int S = tb.S[t.key];
  switch (t.kind) {
    case TG_ARN: return t.sw < S && t.j < tb.B[t.key];
    case TG_RES: return mixed ? t.sw <= S : t.sw < S;       // the mixed form's closing residual
    case TG_XUPD: return t.sw < S;
    default: return true;                                     // TG_GCHK
  }
}

// Whether graph e holds every launch that table t enables.  GMRES positions without nodes in
// the graph do not count: the Newton capacities decide which of them the step reaches.
static bool tab_covers(const GraphEntry& e, const BudTab& t) {
  for (int k = 0; k < NGKEY; k++) {
    if (!e.gk[k]) continue;
    if (t.B[k] > e.cap.B[k] || t.S[k] > e.cap.S[k]) return false;
  }
  for (int k = 0; k < NNKEY; k++) if (t.N[k] > e.cap.N[k]) return false;
  return true;
}

static void drop_graphs(imex_ctx* c) {
  for (GraphEntry& g : c->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
  c->graphs.clear();
}

// Queue (or replay) the budgeted step as a captured hipGraph.  A graph is captured for a
// capacity table a little above the current budgets (and for the current state/face buffer set
// and state mode); replays toggle the budget-dependent nodes to the current table.
static int budget_graph_launch(imex_ctx* c) {
  const int parity = c->iUs == B_U2 ? 0 : 1;
  const BudTab cur = c->tab;
  GraphEntry* g = nullptr;
  for (GraphEntry& e : c->graphs) {
    if (e.parity != parity || e.state_mode != c->state_mode) continue;
    if (c->no_toggle ? std::memcmp(&e.cap, &cur, sizeof(BudTab)) == 0 : tab_covers(e, cur)) { g = &e; break; }
  }
  if (!g) {
    if ((int)c->graphs.size() >= MAX_GRAPHS) {     // drop the least recently used graph
      size_t lru = 0;
      for (size_t i = 1; i < c->graphs.size(); i++) if (c->graphs[i].used < c->graphs[lru].used) lru = i;
      (void)hipGraphExecDestroy(c->graphs[lru].exec);
      (void)hipGraphDestroy(c->graphs[lru].graph);
      c->graphs.erase(c->graphs.begin() + (long)lru);
    }
    BudTab cap = cur;
    if (!c->no_toggle) {
      const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
      for (int k = 0; k < NGKEY; k++) {
        // every Arnoldi step of a restart cycle: the nodes past the budget stay disabled, and
        // Arnoldi budgets, which move most from step to step, never force a capture
        cap.B[k] = m;
        // a refinement sweep more for the mixed form; the FP64 form gets a restart cycle more only
        // where one has been needed
        cap.S[k] = (key_prec(c, k) || cur.S[k] > 1) ? std::min(MAXSWEEP - 1, cur.S[k] + 1) : cur.S[k];
      }
      for (int k = 0; k < NNKEY; k++) cap.N[k] = std::min(std::min(c->prm.newton_max, NIT - 1), cur.N[k] + 1);
      // the new graph also covers the graphs of this parity it replaces, so capacities only grow
      // and budgets that move up and down stop causing captures
      for (size_t i = c->graphs.size(); i-- > 0;) {
        const GraphEntry& e = c->graphs[i];
        if (e.parity != parity || e.state_mode != c->state_mode) continue;
        for (int k = 0; k < NGKEY; k++) { cap.B[k] = std::max(cap.B[k], e.cap.B[k]); cap.S[k] = std::max(cap.S[k], e.cap.S[k]); }
        for (int k = 0; k < NNKEY; k++) cap.N[k] = std::max(cap.N[k], e.cap.N[k]);
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
        c->graphs.erase(c->graphs.begin() + (long)i);
      }
    }
    c->tab = cap;
    c->tags.clear();
    c->tag_err = false;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = step_body(c);
    c->capturing = false;
    hipError_t e1 = hipStreamEndCapture(c->st, &graph);
    c->tab = cur;
    if (rc < 0) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e1 != hipSuccess) return set_err(IMEX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e1));
    GraphEntry ne;
    ne.cap = cap; ne.applied = cap; ne.parity = parity; ne.state_mode = c->state_mode;
    hipError_t e2 = hipGraphInstantiate(&ne.exec, graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) { (void)hipGraphDestroy(graph); return set_err(IMEX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2)); }
    ne.graph = graph;
    if (!c->tag_err) ne.tags = c->tags;
    ne.brecs = c->brecs;
    ne.nslot = c->nslot;
    for (const BudgetRec& r : c->brecs) if (r.kind == 0 && r.key >= 0 && r.key < NGKEY) ne.gk[r.key] = true;
    c->graphs.push_back(ne);
    g = &c->graphs.back();
    c->stats.graph_builds++;
    if (c->tag_err && !c->no_toggle) {           // nodes cannot be identified: fall back to exact tables
      c->no_toggle = true;
      return budget_graph_launch(c);
    }
  }
  if (!c->no_toggle && std::memcmp(&g->applied, &cur, sizeof(BudTab)) != 0) {
    for (const NodeTag& t : g->tags) {
      const bool en = tag_enabled(c, t, cur);
      if (en == tag_enabled(c, t, g->applied)) continue;
      if (hipGraphNodeSetEnabled(g->exec, t.node, en ? 1 : 0) != hipSuccess) {
        // enabling is not available here: drop the graph and capture per exact table from now on
        c->no_toggle = true;
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(g->exec);
        (void)hipGraphDestroy(g->graph);
        c->graphs.erase(c->graphs.begin() + (g - c->graphs.data()));
        return budget_graph_launch(c);
      }
    }
    g->applied = cur;
  }
  g->used = ++c->graph_clock;
  c->brecs = g->brecs;
  c->nslot = g->nslot;
  HIPCHK(hipGraphLaunch(g->exec, c->st));
  return 0;
}

static int one_step(imex_ctx* c) {
  int rc = 0;
  const Launch& L = c->L;
  for (int attempt = 0;; attempt++) {
    cur_reset(c);
    if (!choose_budgets(c)) break;
    if (c->ni > 1 && attempt > 0) {               // a batch grows the budgets of a failed step
      const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
      for (int k = 0; k < NGKEY; k++) { c->tab.B[k] = std::min(m, c->tab.B[k] + 4 * attempt); c->tab.S[k] = std::min(MAXSWEEP - 1, c->tab.S[k] + attempt); }
      for (int k = 0; k < NNKEY; k++) c->tab.N[k] = std::min(std::min(c->prm.newton_max, NIT - 1), c->tab.N[k] + attempt);
    }
    const long long gen0 = c->prof.gen;
    c->budget = true;
    c->nslot = 1;
    c->brecs.clear();
    k_budget_begin<<<dim3(L.nblocks[SH_ONE]), 64, 0, c->st>>>(L.IT, L.bm[SH_ONE]);
    if (c->prm.schedule != IMEX_SCHED_BUDGET && !c->prof.on) rc = budget_graph_launch(c);
    else rc = step_body(c);
    c->budget = false;
    if (rc < 0) return rc;
    if (c->nslot > CTLN) return set_err(IMEX_ERR_ARG, "too many solves in one step");
    int f = budget_finish(c, gen0);
    if (f < 0) return f;
    c->stats.budget_steps++;
    if (f == 0) {
      if (++c->fit_streak >= 16) {
        c->fit_streak = 0;
        c->margin_extra = std::max(0, c->margin_extra - 1);
        c->sweep_extra = std::max(0, c->sweep_extra - 1);
        c->newton_extra = std::max(0, c->newton_extra - 1);
      }
      step_commit(c);
      return rc;
    }
    c->fit_streak = 0;
    // the margins grow after an overrun and shrink again after 16 steps that fit: a rerun costs
    // a whole step, a unit of margin the launches of one Arnoldi step (or Newton iteration) that
    // exit at once.  A single instance's rerun is host-scheduled and its counts enter the
    // history, so its sweep (restart cycle) budgets need no extra margin.
    if (f & 1) {
      c->margin_extra = std::min(8, c->margin_extra + (c->ni > 1 ? 2 : 1));
      if (c->ni > 1) c->sweep_extra = std::min(4, c->sweep_extra + 1);
    }
    if (f & 2) c->newton_extra = std::min(4, c->newton_extra + 1);
    // a solve outran its budget: restore the step's scalars and rerun it (host scheduling for a
    // single instance, larger budgets for a batch)
    c->stats.budget_reruns++;
    k_budget_restore<<<dim3(L.nblocks[SH_ONE]), 64, 0, c->st>>>(L.IT, L.bm[SH_ONE]);
    if (c->ni == 1) break;
    if (attempt >= 3) return set_err(IMEX_ERR_UNSUPPORTED, "batch step did not converge within its budgets");
  }
  cur_reset(c);
  for (int k = 0; k < NBK; k++) c->step_mx[k] = 0;
  rc = step_body(c);
  hist_push(c, c->step_mx);
  step_commit(c);
  return rc;
}

// ---------------------------------------------------------------- host helpers: AoS <-> SoA
static void aos_to_soa(const Dims& d, const double* aos, std::vector<double>& soa) {
  soa.assign((size_t)NVAR * d.PS, 0.0);
  for (long long c = 0; c < (long long)d.NX * d.NY; c++)
    for (int k = 0; k < NVAR; k++) soa[(size_t)k * d.PS + c] = aos[NVAR * c + k];
}
static void soa_to_aos(const Dims& d, const std::vector<double>& soa, double* aos) {
  for (long long c = 0; c < (long long)d.NX * d.NY; c++)
    for (int k = 0; k < NVAR; k++) aos[NVAR * c + k] = soa[(size_t)k * d.PS + c];
}
static int up_grid(imex_ctx* c, int i, int b, const double* aos) {
  const Dims& d = c->hI[i].d;
  std::vector<double> s; aos_to_soa(d, aos, s);
  HIPCHK(hipMemcpy(c->hI[i].buf[b], s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int down_grid(imex_ctx* c, int i, int b, double* aos) {
  const Dims& d = c->hI[i].d;
  std::vector<double> s((size_t)NVAR * d.PS);
  HIPCHK(hipMemcpy(s.data(), c->hI[i].buf[b], s.size() * sizeof(double), hipMemcpyDeviceToHost));
  soa_to_aos(d, s, aos);
  return 0;
}
static int up_plane(imex_ctx* c, int i, int b, const double* a) {
  const Dims& d = c->hI[i].d;
  HIPCHK(hipMemcpy(c->hI[i].buf[b], a, (size_t)d.NX * d.NY * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int down_plane(imex_ctx* c, int i, int b, double* a) {
  const Dims& d = c->hI[i].d;
  HIPCHK(hipMemcpy(a, c->hI[i].buf[b], (size_t)d.NX * d.NY * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

static void scalars_from_dev(const double* Sh, imex_scalars* sc) {
  sc->t = Sh[S_T]; sc->dt = Sh[S_DT]; sc->adt = Sh[S_ADT]; sc->a_max = Sh[S_AMAX];
  sc->B_max_x = Sh[S_BMAXX]; sc->B_max_y = Sh[S_BMAXY]; sc->vx_max = Sh[S_VXMAX]; sc->vy_max = Sh[S_VYMAX];
  sc->v_max = Sh[S_VMAX]; sc->ch = Sh[S_CH]; sc->ch_x = Sh[S_CHX]; sc->ch_y = Sh[S_CHY]; sc->prev_dt = Sh[S_PREVDT];
  sc->timestep = (int)Sh[S_TIMESTEP]; sc->rks = (int)Sh[S_RKS];
}
static void scalars_to_dev(const imex_scalars* sc, double* Sh) {
  Sh[S_T] = sc->t; Sh[S_DT] = sc->dt; Sh[S_ADT] = sc->adt; Sh[S_AMAX] = sc->a_max;
  Sh[S_BMAXX] = sc->B_max_x; Sh[S_BMAXY] = sc->B_max_y; Sh[S_VXMAX] = sc->vx_max; Sh[S_VYMAX] = sc->vy_max;
  Sh[S_VMAX] = sc->v_max; Sh[S_CH] = sc->ch; Sh[S_CHX] = sc->ch_x; Sh[S_CHY] = sc->ch_y; Sh[S_PREVDT] = sc->prev_dt;
  Sh[S_TIMESTEP] = sc->timestep; Sh[S_RKS] = sc->rks;
}
static int pull_scalars(imex_ctx* c, int i) {
  HIPCHK(hipMemcpyAsync(c->Sh, c->hI[i].S, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---------------------------------------------------------------- host boundary fills for initial data
static int hmapx(const imex_params& p, int i) {
  const bool per = p.bc_x == IMEX_BC_PERIODIC, ref = p.bc_x == IMEX_BC_REFLECTIVE;
  if (i < GC) return per ? p.nx + i : (ref ? 2 * GC - 1 - i : GC);
  if (i > p.nx + GC - 1) return per ? i - p.nx : (ref ? 2 * (p.nx + GC) - i - 1 : p.nx + GC - 1);
  return i;
}
static int hmapy(const imex_params& p, int j) {
  const bool per = p.bc_y == IMEX_BC_PERIODIC, ref = p.bc_y == IMEX_BC_REFLECTIVE;
  if (j < GC) return per ? p.ny + j : (ref ? 2 * GC - 1 - j : GC);
  if (j > p.ny + GC - 1) return per ? j - p.ny : (ref ? 2 * (p.ny + GC) - j - 1 : p.ny + GC - 1);
  return j;
}
// Ghost fill with the reference's pass order (x over all rows, then y over all columns), then the jet
// inflow strip: the inflow state in a Grid's bottom ghost rows (prim: the Grid holds primitives, as under
// g_bc_prim), the staggered overrides of bound_cond_staggered (constrained_transport.cpp:50-63).
static void hfill(const imex_params& p, double* a, int nvar, int kind, int prim = 0) {
  const int NX = p.nx + 2 * GC, NY = p.ny + 2 * GC;
  // cells mirror across a reflective boundary with momenta and B negated; staggered faces there
  // are filled as for an outflow (bound_cond_staggered's non-periodic branch)
  auto mx = [&](int i) {
    if (p.bc_x == IMEX_BC_PERIODIC || (kind == 0 && p.bc_x == IMEX_BC_REFLECTIVE)) return hmapx(p, i);
    int lo = kind == 1 ? GC - 1 : GC;
    return i < lo ? lo : (i > p.nx + GC - 1 ? p.nx + GC - 1 : i);
  };
  auto my = [&](int j) {
    if (p.bc_y == IMEX_BC_PERIODIC || (kind == 0 && p.bc_y == IMEX_BC_REFLECTIVE)) return hmapy(p, j);
    int lo = kind == 2 ? GC - 1 : GC;
    return j < lo ? lo : (j > p.ny + GC - 1 ? p.ny + GC - 1 : j);
  };
  auto sg = [&](int k, bool ref) { return (ref && kind == 0 && nvar == NVAR && k != 0 && k != 4) ? -1.0 : 1.0; };
  const bool rx = p.bc_x == IMEX_BC_REFLECTIVE, ry = p.bc_y == IMEX_BC_REFLECTIVE;
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      int si = mx(i);
      if (si != i) for (int k = 0; k < nvar; k++) a[(size_t)nvar * (i + j * NX) + k] = sg(k, rx) * a[(size_t)nvar * (si + j * NX) + k];
    }
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++) {
      int sj = my(j);
      if (sj != j) for (int k = 0; k < nvar; k++) a[(size_t)nvar * (i + j * NX) + k] = sg(k, ry) * a[(size_t)nvar * (i + sj * NX) + k];
    }
  const Dims d = make_dims(p);
  if (kind == 0 && nvar == NVAR && (d.dirx || d.diry)) {      // Dirichlet ambient (boundary.cpp:56-76, 128-147)
    auto amb = [&](int i, int j) {
      double* q = &a[(size_t)NVAR * (i + j * NX)];
      q[0] = d.rho_amb; q[1] = q[2] = q[3] = 0.0;
      q[4] = prim ? d.p_amb : d.e_amb + 0.5 * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
    };
    // the passes above copied B (outflow rule); rho, v and slot 4 of the ghosts take the ambient state
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++)
        if (dir_ghost(d, i, j)) amb(i, j);
  }
  if (p.bc_y != IMEX_BC_JET_INFLOW) return;
  for (int i = 0; i < NX; i++) {
    if (kind == 0 && nvar == NVAR && jet_inlet(d, i))
      for (int j = 0; j < GC; j++)
        for (int k = 0; k < NVAR; k++) a[(size_t)NVAR * (i + j * NX) + k] = jet_state(d, prim, k);
    if (kind == 1 && i < NX - 1 && jet_inlet_face(d, i))
      for (int j = 0; j < GC; j++) a[i + j * NX] = 0.0;
    if (kind == 2 && jet_inlet(d, i))
      for (int j = 0; j < GC - 1; j++) a[i + j * NX] = d.jet_B0;
  }
}

// bound_cond_A (boundary.cpp:177-292) on a host Grid (slots 0..2): x over every row, then y over every
// column; periodic copy, transmissive linear extrapolation from the two edge cells.
static void hbc_A(const imex_params& p, double* A) {
  const int NX = p.nx + 2 * GC, NY = p.ny + 2 * GC;
  auto a = [&](int i, int j, int k) -> double& { return A[(size_t)NVAR * (i + (size_t)j * NX) + k]; };
  for (int j = 0; j < NY; j++)
    for (int k = 0; k < 3; k++) {
      if (p.bc_x == IMEX_BC_PERIODIC) {
        for (int g = 0; g < GC; g++) { a(g, j, k) = a(p.nx + g, j, k); a(p.nx + GC + g, j, k) = a(GC + g, j, k); }
      } else {
        double sL = a(GC + 1, j, k) - a(GC, j, k);
        for (int g = 0; g < GC; g++) a(GC - 1 - g, j, k) = a(GC, j, k) - (g + 1) * sL;
        int iR = p.nx + GC - 1;
        double sR = a(iR, j, k) - a(iR - 1, j, k);
        for (int g = 0; g < GC; g++) a(iR + 1 + g, j, k) = a(iR, j, k) + (g + 1) * sR;
      }
    }
  for (int i = 0; i < NX; i++)
    for (int k = 0; k < 3; k++) {
      if (p.bc_y == IMEX_BC_PERIODIC) {
        for (int g = 0; g < GC; g++) { a(i, g, k) = a(i, p.ny + g, k); a(i, p.ny + GC + g, k) = a(i, GC + g, k); }
      } else {
        double sL = a(i, GC + 1, k) - a(i, GC, k);
        for (int g = 0; g < GC; g++) a(i, GC - 1 - g, k) = a(i, GC, k) - (g + 1) * sL;
        int jR = p.ny + GC - 1;
        double sR = a(i, jR, k) - a(i, jR - 1, k);
        for (int g = 0; g < GC; g++) a(i, jR + 1 + g, k) = a(i, jR, k) + (g + 1) * sR;
      }
    }
  if (p.bc_y == IMEX_BC_JET_INFLOW) {                // even reflection in the inflow strip (boundary.cpp:279-291)
    const Dims d = make_dims(p);
    for (int i = 0; i < NX; i++)
      if (jet_inlet(d, i))
        for (int k = 0; k < 3; k++)
          for (int g = 0; g < GC; g++) a(i, GC - 1 - g, k) = a(i, GC + g, k);
  }
}

// 3 planes of a vector-potential buffer <-> slots 0..2 of a host Grid
static int up_pot(imex_ctx* c, int i, int b, const double* A) {
  const Dims& d = c->hI[i].d;
  std::vector<double> s((size_t)3 * d.PS, 0.0);
  for (long long q = 0; q < (long long)d.NX * d.NY; q++)
    for (int k = 0; k < 3; k++) s[(size_t)k * d.PS + q] = A[NVAR * q + k];
  HIPCHK(hipMemcpy(c->hI[i].buf[b], s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int down_pot(imex_ctx* c, int i, int b, double* A) {
  const Dims& d = c->hI[i].d;
  std::vector<double> s((size_t)3 * d.PS);
  HIPCHK(hipMemcpy(s.data(), c->hI[i].buf[b], s.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (long long q = 0; q < (long long)d.NX * d.NY; q++)
    for (int k = 0; k < NVAR; k++) A[NVAR * q + k] = k < 3 ? s[(size_t)k * d.PS + q] : 0.0;
  return 0;
}

// per-step diagnostics of the current state (iUs after a step, u_pri before the first: the B slots and
// rho are the same, but e is not, so a state still in u_pri is converted first)
static void diag_launch(imex_ctx* c, const Launch& L, int append) {
  const bool face = c->prm.magfield == IMEX_MAGFIELD_STAGGER;
  if (c->state_mode == 0) launch_pri2con(L, c->P, B_UPRI, c->iUs);
  launch_diag(L, c->iUs, c->iBxn, c->iByn, face ? 1 : 0, c->prm.cfl, append);
}

// ================================================================ C ABI
extern "C" {

const char* imex_last_error(void) { return g_err.c_str(); }

int imex_default_params(imex_params* p) {
  if (!p) return IMEX_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->nx = 64; p->ny = 64;
  p->x0 = -5; p->x1 = 5; p->y0 = -5; p->y1 = 5;
  p->gamma = 5.0 / 3.0;
  p->bc_x = IMEX_BC_PERIODIC; p->bc_y = IMEX_BC_PERIODIC;
  p->recon = IMEX_RECON_MINMOD; p->minmod_central_y = 0;
  p->eos = IMEX_EOS_IDEAL;
  p->rk_alfa = 0.05; p->rk_b = 0.1; p->rk_R = 0.4; p->rk_cv = 1.0;      // config.hpp:125
  p->cfl = 0.8;
  p->floor_on = 1; p->p_floor = 1e-7; p->theta = 0.9; p->mbig_fac = 1e6;  // config.hpp:100-103
  p->newton_max = 20; p->newton_tol = 1e-8;                              // config.hpp:106-107
  p->gmres_tol = 1e-14; p->gmres_restart = 25; p->gmres_max_outer = 0;  // config.hpp:71, implicit_operators.cpp:801-802
  p->shock_kappa1 = 0.001;                                               // config.hpp:77
  p->time_stop = 1e300;
  p->ct_mode = IMEX_CT_CONSISTENT;
  p->precision = IMEX_PREC_FP64;
  p->mixed_inner_tol = 1e-6;
  p->mixed_outer_tol = 1e-9;
  p->schedule = IMEX_SCHED_AUTO;
  p->budget_margin = 1;
  p->formula = IMEX_FORMULA_PRESSURE;                                    // config.hpp:44
  p->recon_state = IMEX_RECONSTATE_CONS;                                 // config.hpp:61
  p->jet_rho = 1.4; p->jet_v = 10000.0; p->jet_B0 = sqrt(20000.0); p->jet_p = 1.0;   // config.hpp:257-260
  p->jet_halfwidth = 0.05;                                               // config.hpp:261-269
  p->rho_amb = 1.0; p->p_amb = 0.1;                                      // config.hpp:247
  return 0;
}

int imex_init_case(const imex_params* p, int case_id, double amp, double u0, double v0, int smooth,
                   double* u, double* bx, double* by) {
  if (!p || !u || !bx || !by) return set_err(IMEX_ERR_ARG, "null argument");
  const int NX = p->nx + 2 * GC, NY = p->ny + 2 * GC;
  const double dx = (p->x1 - p->x0) / p->nx, dy = (p->y1 - p->y0) / p->ny;
  const double g = p->gamma;
  std::memset(u, 0, sizeof(double) * NVAR * NX * NY);
  std::memset(bx, 0, sizeof(double) * NX * NY);
  std::memset(by, 0, sizeof(double) * NX * NY);
  auto U = [&](int i, int j, int k) -> double& { return u[NVAR * (i + j * NX) + k]; };
  for (int i = GC; i < p->nx + GC; i++)
    for (int j = GC; j < p->ny + GC; j++) {
      double x = p->x0 + (i - GC + 0.5) * dx, y = p->y0 + (j - GC + 0.5) * dy;
      if (case_id == IMEX_CASE_VORTEX) {           // initial_conditions.cpp:641-668, amplitude / background added
        const double eps = amp, mu_tilde = amp * sqrt(4.0 * M_PI);
        double r2 = x * x + y * y;
        double eh = exp(0.5 * (1.0 - r2)), ef = exp(1.0 - r2);
        U(i, j, 0) = 1.0;
        double du = -eps / (2.0 * M_PI) * eh * y, dv = eps / (2.0 * M_PI) * eh * x;
        U(i, j, 1) = u0 != 0.0 ? u0 + du : du;
        U(i, j, 2) = v0 != 0.0 ? v0 + dv : dv;
        U(i, j, 3) = 0.0;
        U(i, j, 4) = 1.0 + (1.0 / (8.0 * M_PI)) * (mu_tilde / (2.0 * M_PI)) * (mu_tilde / (2.0 * M_PI)) * (1.0 - r2) * ef
                   - 0.5 * (eps / (2.0 * M_PI)) * (eps / (2.0 * M_PI)) * ef;
        U(i, j, 5) = -amp / (2.0 * M_PI) * eh * y;
        U(i, j, 6) = amp / (2.0 * M_PI) * eh * x;
        U(i, j, 7) = 0.0;
      } else if (case_id == IMEX_CASE_ORSZAG_TANG) {   // initial_conditions.cpp:426-435 on [0,2pi]^2
        U(i, j, 0) = g * g;
        U(i, j, 1) = -sin(y); U(i, j, 2) = sin(x); U(i, j, 3) = 0;
        U(i, j, 4) = g;
        U(i, j, 5) = -sin(y); U(i, j, 6) = sin(2 * x); U(i, j, 7) = 0;
      } else if (case_id == IMEX_CASE_BLAST) {         // initial_conditions.cpp:436-449, field along (1,1)
        double rad = sqrt(x * x + y * y);
        const double rD = 0.1, prsL = 1000.0, prsR = 0.1;
        U(i, j, 0) = 1.0;
        if (smooth) { double e = 1 * sqrt(dx * dx + dy * dy); U(i, j, 4) = 0.5 * (prsR + prsL) + 0.5 * (prsR - prsL) * erf((rad - rD) / e); }
        else U(i, j, 4) = rad < rD ? prsL : prsR;
        double b0 = amp / sqrt(4.0 * M_PI) / sqrt(2.0);
        U(i, j, 5) = b0; U(i, j, 6) = b0; U(i, j, 7) = 0.0;
      } else if (case_id == IMEX_CASE_JET) {           // testMHDJet ambient (initial_conditions.cpp:671-682)
        U(i, j, 0) = 0.1 * g;
        U(i, j, 1) = 0.0; U(i, j, 2) = 0.0; U(i, j, 3) = 0.0;
        U(i, j, 4) = 1.0;
        U(i, j, 5) = 0.0; U(i, j, 6) = p->jet_B0; U(i, j, 7) = 0.0;
      } else {
        return set_err(IMEX_ERR_ARG, "unknown case");
      }
    }
  hfill(*p, u, NVAR, 0, 1);                          // primitives (initial_conditions.cpp:686-689)
  auto BX = [&](int i, int j) -> double& { return bx[i + j * NX]; };
  auto BY = [&](int i, int j) -> double& { return by[i + j * NX]; };
  if (case_id == IMEX_CASE_VORTEX) {                 // faces from the corner potential (:698-717)
    auto Az = [&](int i, int j) {
      double x = p->x0 + (i - GC + 1.0) * dx, y = p->y0 + (j - GC + 1.0) * dy;
      double r2 = x * x + y * y;
      return (amp / (2.0 * M_PI)) * exp(0.5 * (1.0 - r2));
    };
    for (int i = 0; i < NX; i++) for (int j = 1; j < NY; j++) BX(i, j) = (Az(i, j) - Az(i, j - 1)) / dy;
    for (int i = 1; i < NX; i++) for (int j = 0; j < NY; j++) BY(i, j) = -(Az(i, j) - Az(i - 1, j)) / dx;
    for (int i = 1; i < NX; i++)
      for (int j = 1; j < NY; j++) {
        U(i, j, 5) = 0.5 * (BX(i - 1, j) + BX(i, j));
        U(i, j, 6) = 0.5 * (BY(i, j - 1) + BY(i, j));
      }
    hfill(*p, u, NVAR, 0, 1);
  } else {                                          // faces = cell averages (:718-725)
    for (int i = 0; i < NX - 1; i++)
      for (int j = 0; j < NY - 1; j++) {
        BX(i, j) = 0.5 * (U(i, j, 5) + U(i + 1, j, 5));
        BY(i, j) = 0.5 * (U(i, j, 6) + U(i, j + 1, 6));
      }
  }
  hfill(*p, bx, 1, 1);
  hfill(*p, by, 1, 2);
  return 0;
}

// A_initial (constrained_transport.cpp:224-398) for the cases of imex_init_case: A_z whose curl is the
// case's B, at every cell centre, then bound_cond_A.
int imex_init_potential(const imex_params* p, int case_id, double amp, double* A) {
  if (!p || !A) return set_err(IMEX_ERR_ARG, "null argument");
  const int NX = p->nx + 2 * GC, NY = p->ny + 2 * GC;
  const double dx = (p->x1 - p->x0) / p->nx, dy = (p->y1 - p->y0) / p->ny;
  std::memset(A, 0, sizeof(double) * NVAR * NX * NY);
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++) {
      double x = p->x0 + (i - GC + 0.5) * dx, y = p->y0 + (j - GC + 0.5) * dy;
      double az;
      if (case_id == IMEX_CASE_VORTEX) az = (amp / (2.0 * M_PI)) * exp(0.5 * (1.0 - (x * x + y * y)));   // :387-391
      else if (case_id == IMEX_CASE_ORSZAG_TANG) az = cos(y) + cos(2.0 * x) / 2.0;   // :341-342 on [0,2pi]^2
      else if (case_id == IMEX_CASE_BLAST) { double b0 = amp / sqrt(4.0 * M_PI) / sqrt(2.0); az = b0 * y - b0 * x; }
      else if (case_id == IMEX_CASE_JET) az = -p->jet_B0 * x;                                 // :240-243
      else return set_err(IMEX_ERR_ARG, "unknown case");
      A[(size_t)NVAR * (i + (size_t)j * NX) + 2] = az;
    }
  hbc_A(*p, A);
  return 0;
}


static int check_params(const imex_params* prm) {
  if (prm->gmres_restart > MAXM) return set_err(IMEX_ERR_ARG, "gmres_restart exceeds 32");
  if (prm->bc_x < IMEX_BC_TRANSMISSIVE || prm->bc_x > IMEX_BC_DIRICHLET ||
      prm->bc_y < IMEX_BC_TRANSMISSIVE || prm->bc_y > IMEX_BC_JET_INFLOW)
    return set_err(IMEX_ERR_ARG, "bad boundary (the jet inflow is a y boundary)");
  if ((prm->bc_x == IMEX_BC_DIRICHLET || prm->bc_y == IMEX_BC_DIRICHLET) && !(prm->rho_amb > 0 && prm->p_amb > 0))
    return set_err(IMEX_ERR_ARG, "the Dirichlet ambient needs rho_amb, p_amb > 0");
  if (prm->bc_y == IMEX_BC_JET_INFLOW && !(prm->jet_rho > 0 && prm->jet_p > 0 && prm->jet_halfwidth > 0))
    return set_err(IMEX_ERR_ARG, "jet inflow state needs jet_rho, jet_p, jet_halfwidth > 0");
  if (prm->recon < IMEX_RECON_NONE || prm->recon > IMEX_RECON_WENO5)
    return set_err(IMEX_ERR_ARG, "bad reconstruction");
  if (prm->recon_state != IMEX_RECONSTATE_PRIM && prm->recon_state != IMEX_RECONSTATE_CONS)
    return set_err(IMEX_ERR_ARG, "bad recon_state");
  if (prm->precision != IMEX_PREC_FP64 && prm->precision != IMEX_PREC_MIXED)
    return set_err(IMEX_ERR_ARG, "bad precision");
  if (prm->schedule != IMEX_SCHED_GRAPH && prm->schedule != IMEX_SCHED_BUDGET && prm->schedule != IMEX_SCHED_HOST &&
      prm->schedule != IMEX_SCHED_AUTO)
    return set_err(IMEX_ERR_ARG, "bad schedule");
  if (prm->formula != IMEX_FORMULA_PRESSURE && prm->formula != IMEX_FORMULA_ENERGY)
    return set_err(IMEX_ERR_ARG, "bad formula");
  if (prm->magfield < IMEX_MAGFIELD_STAGGER || prm->magfield > IMEX_MAGFIELD_NODIVFREE)
    return set_err(IMEX_ERR_ARG, "bad magfield");
  return 0;
}

// Device memory of instance i, carved from one allocation (256-byte aligned pieces).
static int alloc_instance(imex_ctx* c, int i) {
  Inst& I = c->hI[i];
  const Dims& d = I.d;
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  const size_t G8 = (size_t)NVAR * d.PS, G1 = (size_t)d.PS, N = (size_t)d.N, n3 = 3 * N;
  // partial rows: the largest local block count of any kernel that writes per-block partials
  const long long tiles = cdiv(d.nx, TX) * cdiv(d.ny, TY);
  const long long rows = std::min<long long>(4096, std::max({tiles, cdiv(3LL * d.N, 256), cdiv((long long)d.NX * d.NY, 256)}));
  std::vector<std::pair<void**, size_t>> parts;
  auto P8 = [&](double** p, size_t n) { parts.push_back({(void**)p, n * sizeof(double)}); };
  for (int b = B_UPRI; b <= B_QS; b++) P8(&I.buf[b], G8);
  for (int b = B_BX0; b <= B_VISC; b++) P8(&I.buf[b], G1);
  for (int b : {B_W, B_AV, B_XK, B_BK}) P8(&I.buf[b], n3);
  P8(&I.buf[B_TMPN], 6 * N);
  for (int b = B_P; b <= B_VSLOPE; b++) P8(&I.buf[b], N);
  for (int b = B_A0; b <= B_AS; b++) I.buf[b] = nullptr;
  if (c->needA) for (int b = B_A0; b <= B_AS; b++) P8(&I.buf[b], 3 * G1);
  for (int k = 0; k <= m; k++) P8(&I.buf[B_V0 + k], n3);
  for (int k = 0; k < NFBUF; k++) I.fbuf[k] = nullptr;
  if (c->mixed) {
    for (int b : {F_W32, F_AV32}) parts.push_back({(void**)&I.fbuf[b], n3 * sizeof(float)});
    for (int k = 0; k <= m; k++) parts.push_back({(void**)&I.fbuf[F_V0 + k], n3 * sizeof(float)});
  }
  P8(&I.S, S_NSLOTS);
  P8(&I.partial, (size_t)rows * (MAXM + 2));
  P8(&I.partial2, 2 * (size_t)std::max<long long>(rows, cdiv(d.nx, 64) * cdiv(d.ny, 4)));
  P8(&I.pst, 4);
  P8(&I.bsave, S_NSLOTS + 4);
  parts.push_back({(void**)&I.eflag, sizeof(int)});
  parts.push_back({(void**)&I.esave, sizeof(int)});
  parts.push_back({(void**)&I.K, sizeof(KrylovDev)});
  size_t off = 0;
  std::vector<size_t> offs;
  for (auto& pt : parts) { off = (off + 255) & ~(size_t)255; offs.push_back(off); off += pt.second; }
  char* base = nullptr;
  int rc = dalloc(c, &base, off);
  if (rc) return rc;
  HIPCHK(hipMemset(base, 0, off));
  for (size_t q = 0; q < parts.size(); q++) *parts[q].first = base + offs[q];
  for (int k = m + 1; k <= MAXM; k++) I.buf[B_V0 + k] = I.buf[B_V0];
  if (c->mixed) for (int k = m + 1; k <= MAXM; k++) I.fbuf[F_V0 + k] = I.fbuf[F_V0];
  I.ctl = c->ctl_all + (size_t)i * CTLN;
  KrylovDev kz; std::memset(&kz, 0, sizeof(kz)); kz.tol = c->prm.gmres_tol;
  HIPCHK(hipMemcpy(I.K, &kz, sizeof(kz), hipMemcpyHostToDevice));
  const double init[4] = {1e300, 0.0, 0.0, 0.0};
  HIPCHK(hipMemcpy(I.pst, init, sizeof(init), hipMemcpyHostToDevice));
  return 0;
}

// Launch descriptor covering instance i alone (host set-up work on one instance of a batch).
static int one_launch(imex_ctx* c, int i, Launch& L1) {
  if (c->ni == 1) { L1 = c->L; return 0; }
  L1.IT = c->IT + i; L1.st = c->st;
  for (int sh = 0; sh < NSHAPE; sh++) {
    int g, nb;
    shape_grid(c, sh, c->hI[i].d, g, nb);
    L1.nblocks[sh] = nb;
    BMap& m = L1.bm[sh];
    m.blk = nullptr; m.pre = nullptr; m.gx = nullptr;
    m.ni = 1; m.gx0 = g; m.nb0 = nb;
  }
  return 0;
}

// ||div B||_2 of every instance
static int div_b_all(imex_ctx* c, std::vector<double>& out) {
  if (c->prm.magfield == IMEX_MAGFIELD_STAGGER) launch_divb(c->L, c->iBxn, c->iByn);
  else launch_divb_cell(c->L, c->state_mode ? c->iUs : B_UPRI);     // B slots are the same in pri and con
  HIPCHK(hipStreamSynchronize(c->st));
  out.assign(c->ni, 0.0);
  std::vector<double> h;
  for (int i = 0; i < c->ni; i++) {
    int g, nb;
    shape_grid(c, SH_DIVB, c->hI[i].d, g, nb);
    h.resize(nb);
    HIPCHK(hipMemcpy(h.data(), c->hI[i].partial, nb * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0; for (double v : h) s += v;
    out[i] = std::sqrt(s);
  }
  return 0;
}

int imex_batch_create(const imex_params* prm, int n, const int* nx, const int* ny, imex_ctx** out) {
  if (!prm || !out || n < 1 || !nx || !ny) return set_err(IMEX_ERR_ARG, "null argument or empty batch");
  int rc = check_params(prm);
  if (rc) return rc;
  for (int i = 0; i < n; i++)
    if (nx[i] < 4 || ny[i] < 4) return set_err(IMEX_ERR_ARG, "grid too small (need nx, ny >= 4)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(IMEX_ERR_HIP, "no HIP device");
  imex_ctx* c = new imex_ctx();
  c->prm = *prm;
  c->prm.nx = nx[0]; c->prm.ny = ny[0];
  c->P = make_phys(*prm);
  c->mixed = prm->precision == IMEX_PREC_MIXED ? 1 : 0;
  c->mixed_ops = c->mixed ? (1 | (prm->mixed_pressure ? 2 : 0)) : 0;
  // FP32 explicit fluxes are opt-in (mixed_flux): their rounding, 1e-7 relative and uncorrelated from
  // cell to cell, enters the right-hand sides of the magnetic and pressure solves, and the solves'
  // tolerances then take more Arnoldi steps than the FP32 fluxes save (Orszag-Tang 1024^2: the FP64
  // pressure solves need 470 instead of 267 Arnoldi steps over 20 steps).  Under the energy formula
  // they stay FP64 in any case: the energy-form discretisation has a slowly growing mode (about x1.1
  // per step on Orszag-Tang) that FP32 flux rounding would seed, while the pressure formula resets E
  // from p each stage and keeps such errors flat.
  c->mixed_flux = c->mixed && prm->mixed_flux && prm->formula == IMEX_FORMULA_PRESSURE &&
                  prm->recon_state == IMEX_RECONSTATE_CONS;
  c->needA = prm->magfield == IMEX_MAGFIELD_UNSTAGGER || prm->magfield == IMEX_MAGFIELD_POTENTIAL_A;
  c->cap_tiles_p = op_prec(c, 1) ? 2048 : 1536;
  // A tile-looping grid larger than the blocks resident at once ends in a partly filled round: the Newton
  // residual runs one round of resident blocks
  {
    const int res = newton_residual_resident(prm->eos != IMEX_EOS_IDEAL);
    c->cap_tiles_n = res > 0 ? std::min(res, 4096) : 2048;
  }
  c->ni = n;
  c->iprm.assign(n, c->prm);
  c->hI.resize(n);
  c->a_set.assign(n, 0);
  for (int i = 0; i < n; i++) {
    c->iprm[i].nx = nx[i]; c->iprm[i].ny = ny[i];
    std::memset(&c->hI[i], 0, sizeof(Inst));
    c->hI[i].d = make_dims(c->iprm[i]);
    c->cells += (long long)nx[i] * ny[i];
  }
  auto fail = [&](int r) { imex_destroy(c); return r; };
  if ((rc = dalloc(c, &c->ctl_all, (size_t)n * CTLN))) return fail(rc);
  if (hipMemset(c->ctl_all, 0, (size_t)n * CTLN * sizeof(int)) != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "memset"));
  for (int i = 0; i < n; i++) if ((rc = alloc_instance(c, i))) return fail(rc);
  if ((rc = dalloc(c, &c->IT, (size_t)n))) return fail(rc);
  if (hipMemcpy(c->IT, c->hI.data(), (size_t)n * sizeof(Inst), hipMemcpyHostToDevice) != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "table"));
  if ((rc = dalloc(c, &c->bmap_dev, (size_t)NSHAPE * (2 * n + 1)))) return fail(rc);
  if (n > 1) {          // block table sized for every launch shape the autotuner may choose
    const int ct = c->cap_tiles, ctp = c->cap_tiles_p, cv = c->cap_vec, et = c->ex_tile, er = c->ex_rows;
    c->cap_tiles = c->cap_tiles_p = c->cap_vec = 4096;
    for (int q = 0; q < NEXT; q++) {
      c->ex_tile = EX_TILES[q][0]; c->ex_rows = EX_TILES[q][1];
      c->blk_cap = std::max(c->blk_cap, map_blocks(c));
    }
    c->cap_tiles = ct; c->cap_tiles_p = ctp; c->cap_vec = cv; c->ex_tile = et; c->ex_rows = er;
    if ((rc = dalloc(c, &c->blk_dev, (size_t)c->blk_cap))) return fail(rc);
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { c->st = nullptr; return fail(set_err(IMEX_ERR_HIP, "stream")); }
  (void)hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming);
  (void)hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming);
  c->L.IT = c->IT; c->L.st = c->st;
  if ((rc = build_maps(c))) return fail(rc);
  if (hipHostMalloc((void**)&c->Kh, 2 * sizeof(KrylovDev), hipHostMallocDefault) != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "pinned"));
  if (hipHostMalloc((void**)&c->Sh, S_NSLOTS * sizeof(double), hipHostMallocDefault) != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "pinned"));
  if (hipHostMalloc((void**)&c->ctlh, (size_t)n * CTLN * sizeof(int), hipHostMallocDefault) != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "pinned"));
  if (hipDeviceSynchronize() != hipSuccess) return fail(set_err(IMEX_ERR_HIP, "init sync"));
  c->stats.min_p = 1e300;
  bind_stage0(c);
  *out = c;
  return 0;
}

int imex_create(const imex_params* prm, imex_ctx** out) {
  if (!prm || !out) return set_err(IMEX_ERR_ARG, "null argument");
  return imex_batch_create(prm, 1, &prm->nx, &prm->ny, out);
}

int imex_batch_size(const imex_ctx* c) { return c ? c->ni : 0; }

int imex_destroy(imex_ctx* c) {
  if (!c) return 0;
  if (c->st) (void)hipStreamSynchronize(c->st);
  adj_free(c);
  drop_graphs(c);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->Kh) (void)hipHostFree(c->Kh);
  if (c->Sh) (void)hipHostFree(c->Sh);
  if (c->ctlh) (void)hipHostFree(c->ctlh);
  for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
  if (c->st) { (void)hipEventDestroy(c->ev[0]); (void)hipEventDestroy(c->ev[1]); (void)hipStreamDestroy(c->st); }
  delete c;
  return 0;
}

long long imex_device_bytes(const imex_ctx* c) { return c ? c->bytes : 0; }

int imex_batch_set_state(imex_ctx* c, int i, const double* u_pri, const double* bx, const double* by,
                         double t, int timestep, double prev_dt) {
  if (!c || !u_pri || !bx || !by) return set_err(IMEX_ERR_ARG, "null argument");
  if (i < 0 || i >= c->ni) return set_err(IMEX_ERR_ARG, "instance index out of range");
  HIPCHK(hipStreamSynchronize(c->st));
  const Dims& d = c->hI[i].d;
  std::vector<double> tmp(u_pri, u_pri + (size_t)NVAR * d.NX * d.NY);
  hfill(c->iprm[i], tmp.data(), NVAR, 0, 1);    // bound_cond_u(u_pri) at the top of the step (main.cpp:336)
  int rc = up_grid(c, i, B_UPRI, tmp.data());
  if (!rc) rc = up_plane(c, i, c->iBxn, bx);
  if (!rc) rc = up_plane(c, i, c->iByn, by);
  if (rc) return rc;
  std::memset(c->Sh, 0, S_NSLOTS * sizeof(double));
  c->Sh[S_T] = t; c->Sh[S_TIMESTEP] = timestep; c->Sh[S_PREVDT] = prev_dt;
  HIPCHK(hipMemcpy(c->hI[i].S, c->Sh, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice));
  if (c->ni == 1) c->state_mode = 0;
  else if (c->state_mode == 1) {        // the batch already holds conservative states: convert this one
    Launch L1;
    if ((rc = one_launch(c, i, L1))) return rc;
    launch_pri2con(L1, c->P, B_UPRI, c->iUs);
    HIPCHK(hipStreamSynchronize(c->st));
  }
  c->hist_n = 0; c->hist_pos = 0;       // iteration budgets are relearned for the new state
  c->margin_extra = c->sweep_extra = c->newton_extra = c->fit_streak = 0;
  return 0;
}

int imex_batch_set_potential(imex_ctx* c, int i, const double* A) {
  if (!c || !A) return set_err(IMEX_ERR_ARG, "null argument");
  if (i < 0 || i >= c->ni) return set_err(IMEX_ERR_ARG, "instance index out of range");
  if (!c->needA) return set_err(IMEX_ERR_UNSUPPORTED, "this magfield does not evolve the vector potential");
  HIPCHK(hipStreamSynchronize(c->st));
  int rc = up_pot(c, i, c->iAn, A);
  if (rc) return rc;
  Launch L1;
  if ((rc = one_launch(c, i, L1))) return rc;
  launch_bc_a(L1, c->iAn);
  HIPCHK(hipStreamSynchronize(c->st));
  c->a_set[i] = 1;
  return 0;
}
int imex_set_potential(imex_ctx* c, const double* A) { return imex_batch_set_potential(c, 0, A); }
int imex_batch_get_potential(imex_ctx* c, int i, double* A) {
  if (!c || !A) return set_err(IMEX_ERR_ARG, "null argument");
  if (i < 0 || i >= c->ni) return set_err(IMEX_ERR_ARG, "instance index out of range");
  if (!c->needA) return set_err(IMEX_ERR_UNSUPPORTED, "this magfield does not evolve the vector potential");
  HIPCHK(hipStreamSynchronize(c->st));
  return down_pot(c, i, c->iAn, A);
}
int imex_get_potential(imex_ctx* c, double* A) { return imex_batch_get_potential(c, 0, A); }

int imex_set_state(imex_ctx* c, const double* u_pri, const double* bx, const double* by, double t, int timestep, double prev_dt) {
  return imex_batch_set_state(c, 0, u_pri, bx, by, t, timestep, prev_dt);
}

int imex_batch_get_state(imex_ctx* c, int i, double* u_con, double* u_pri, double* bx, double* by, imex_scalars* sc) {
  if (!c) return set_err(IMEX_ERR_ARG, "null ctx");
  if (i < 0 || i >= c->ni) return set_err(IMEX_ERR_ARG, "instance index out of range");
  HIPCHK(hipStreamSynchronize(c->st));
  Launch L1;
  int rc = one_launch(c, i, L1);
  if (rc) return rc;
  if (c->state_mode == 1) launch_con2pri(L1, c->P, c->iUs, B_UPRI);
  else launch_pri2con(L1, c->P, B_UPRI, c->iUs);
  HIPCHK(hipStreamSynchronize(c->st));
  if (u_con) rc = down_grid(c, i, c->iUs, u_con);
  if (!rc && u_pri) rc = down_grid(c, i, B_UPRI, u_pri);
  if (!rc && bx) rc = down_plane(c, i, c->iBxn, bx);
  if (!rc && by) rc = down_plane(c, i, c->iByn, by);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c, i))) return rc; scalars_from_dev(c->Sh, sc); }
  return 0;
}

int imex_get_state(imex_ctx* c, double* u_con, double* u_pri, double* bx, double* by, imex_scalars* sc) {
  return imex_batch_get_state(c, 0, u_con, u_pri, bx, by, sc);
}

// ---------------------------------------------------------------- automatic schedule
// Budgeted steps save the host round trips of the polled loop but pay for the launches that
// find their solve converged and for the reruns of steps that outrun their budgets; which wins
// depends on the problem (small grids and steady iteration counts favour budgets, large grids
// whose counts swing favour polling).  The schedules give the same results bit for bit, so a
// single instance can pick by measurement: a trial of AUTO_WARM + AUTO_TIMED host-polled steps
// (the first trial also gives the budgets their history), one of as many budgeted graph steps
// (the untimed ones absorb the other schedule's queued tail and the captures of both state
// parities), then the faster schedule for AUTO_RUN steps, then trials again.  A batch always
// runs budgeted.
#define AUTO_WARM 2
#define AUTO_TIMED 3
#define AUTO_RUN 48
static bool auto_before(imex_ctx* c) {
  if (c->ni != 1 || c->prm.schedule != IMEX_SCHED_AUTO) return false;
  if (c->taping || c->prof.on) { c->auto_budget = c->auto_win; return false; }
  if (c->auto_left == 0) {
    if (c->auto_phase == 1) c->auto_win = c->auto_ms[1] < c->auto_ms[0];
    c->auto_phase = (c->auto_phase + 1) % 3;
    c->auto_left = c->auto_phase < 2 ? AUTO_WARM + AUTO_TIMED : AUTO_RUN;
  }
  c->auto_budget = c->auto_phase == 0 ? false : c->auto_phase == 1 ? true : c->auto_win;
  if (c->auto_phase < 2 && c->auto_left == AUTO_TIMED) c->auto_t0 = std::chrono::steady_clock::now();
  return true;
}
static int auto_after(imex_ctx* c) {
  if (--c->auto_left > 0 || c->auto_phase == 2) return 0;
  HIPCHK(hipStreamSynchronize(c->st));       // the trial's last step has finished on the device
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c->auto_t0).count();
  c->auto_ms[c->auto_phase] = ms / AUTO_TIMED;
  return 0;
}

int imex_step_async(imex_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return set_err(IMEX_ERR_ARG, "bad argument");
  if (c->needA)
    for (int i = 0; i < c->ni; i++)
      if (!c->a_set[i]) return set_err(IMEX_ERR_ARG, "UNSTAGGER / POTENTIAL_A need the vector potential (imex_set_potential)");
  int rc = 0;
  for (int s = 0; s < nsteps; s++) {
    const bool timed = auto_before(c);
    int r = one_step(c);
    if (r < 0) return r;
    if (timed && (r = auto_after(c)) < 0) return r;
    rc |= r;
    if (c->diag_cap > 0) diag_launch(c, c->L, 1);
  }
  return rc;
}

int imex_sync(imex_ctx* c) {
  if (!c) return IMEX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipGetLastError());
  return 0;
}

int imex_get_stats(imex_ctx* c, imex_stats* s) {
  if (!c || !s) return IMEX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->st));
  double mn = 1e300; int mf = 0; bool viol = false, ef_any = false;
  for (int i = 0; i < c->ni; i++) {
    int ef = 0;
    HIPCHK(hipMemcpy(&ef, c->hI[i].eflag, sizeof(int), hipMemcpyDeviceToHost));
    double acc[4];
    HIPCHK(hipMemcpy(acc, c->hI[i].pst, sizeof(acc), hipMemcpyDeviceToHost));
    ef_any |= ef != 0;
    mn = std::min(mn, acc[0]);
    mf = std::max(mf, (int)acc[1]);
    viol |= acc[2] != 0.0;
  }
  if (ef_any) c->stats.warnings |= IMEX_WARN_EOS_MAXITER;
  if (viol) c->stats.warnings |= IMEX_WARN_POSITIVITY;
  c->stats.min_p = mn;
  c->stats.max_floored = mf;
  *s = c->stats;
  s->div_b = imex_div_b(c);
  return 0;
}

int imex_reset_stats(imex_ctx* c) {
  if (!c) return IMEX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->st));
  c->stats = imex_stats{};
  c->stats.min_p = 1e300;
  const double init[4] = {1e300, 0.0, 0.0, 0.0};
  const int zero = 0;
  for (int i = 0; i < c->ni; i++) {
    HIPCHK(hipMemcpy(c->hI[i].pst, init, sizeof(init), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->hI[i].eflag, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  return 0;
}

int imex_step(imex_ctx* c, int nsteps, imex_stats* stats) {
  int rc = imex_step_async(c, nsteps);
  if (rc < 0) return rc;
  int r2 = imex_sync(c);
  if (r2 < 0) return r2;
  if (stats) imex_get_stats(c, stats);
  return rc | c->stats.warnings;
}

double imex_batch_div_b(imex_ctx* c, int i) {
  if (!c || i < 0 || i >= c->ni) return NAN;
  std::vector<double> v;
  if (div_b_all(c, v)) return NAN;
  return v[i];
}

double imex_div_b(imex_ctx* c) {
  if (!c) return NAN;
  std::vector<double> v;
  if (div_b_all(c, v)) return NAN;
  return *std::max_element(v.begin(), v.end());
}

// ---- stage-level drop-ins (single-instance contexts) ----
#define SINGLE(c) do { if (!(c)) return IMEX_ERR_ARG; if ((c)->ni != 1) return set_err(IMEX_ERR_UNSUPPORTED, "stage- and operator-level calls take a single-instance context"); } while (0)

int imex_set_stage(imex_ctx* c, const imex_stage_host* st, const imex_scalars* sc) {
  SINGLE(c);
  if (!st || !sc) return set_err(IMEX_ERR_ARG, "null argument");
  c->cur.u = B_U1; c->cur.Qe = B_QE; c->cur.Qs = B_QS;
  c->cur.bxqs = B_BXQS; c->cur.byqs = B_BYQS; c->cur.bxo = B_BX1; c->cur.byo = B_BY1;
  c->cur.As = B_AS; c->cur.Ao = B_A1;
  int rc = up_grid(c, 0, B_U1, st->u_con);
  if (!rc) rc = up_grid(c, 0, B_QE, st->Qe);
  if (!rc) rc = up_grid(c, 0, B_QS, st->Qs);
  if (!rc && st->bx) rc = up_plane(c, 0, B_BX1, st->bx);
  if (!rc && st->by) rc = up_plane(c, 0, B_BY1, st->by);
  if (!rc && st->bx_qs) rc = up_plane(c, 0, B_BXQS, st->bx_qs);
  if (!rc && st->by_qs) rc = up_plane(c, 0, B_BYQS, st->by_qs);
  if (!rc && st->fx) rc = up_plane(c, 0, B_FX, st->fx);
  if (!rc && st->fy) rc = up_plane(c, 0, B_FY, st->fy);
  if (rc) return rc;
  std::memset(c->Sh, 0, S_NSLOTS * sizeof(double));
  scalars_to_dev(sc, c->Sh);
  HIPCHK(hipMemcpy(c->hI[0].S, c->Sh, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice));
  c->qe_ready = false; c->hhat_ready = false;
  return 0;
}

int imex_get_stage(imex_ctx* c, imex_stage_host* st, imex_scalars* sc) {
  SINGLE(c);
  if (!st) return set_err(IMEX_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  int rc = 0;
  if (st->u_con) rc = down_grid(c, 0, c->cur.u, st->u_con);
  if (!rc && st->Qe) rc = down_grid(c, 0, c->cur.Qe, st->Qe);
  if (!rc && st->Qs) rc = down_grid(c, 0, c->cur.Qs, st->Qs);
  if (!rc && st->bx) rc = down_plane(c, 0, c->cur.bxo, st->bx);
  if (!rc && st->by) rc = down_plane(c, 0, c->cur.byo, st->by);
  if (!rc && st->bx_qs) rc = down_plane(c, 0, c->cur.bxqs, st->bx_qs);
  if (!rc && st->by_qs) rc = down_plane(c, 0, c->cur.byqs, st->by_qs);
  if (!rc && st->fx) rc = down_plane(c, 0, B_FX, st->fx);
  if (!rc && st->fy) rc = down_plane(c, 0, B_FY, st->fy);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c, 0))) return rc; scalars_from_dev(c->Sh, sc); }
  return 0;
}

static int finish(imex_ctx* c, int rc) {
  hipError_t e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return set_err(IMEX_ERR_HIP, hipGetErrorString(e));
  e = hipGetLastError();
  if (e != hipSuccess) return set_err(IMEX_ERR_HIP, hipGetErrorString(e));
  return rc;
}

int imex_cfl(imex_ctx* c, imex_scalars* sc) {
  SINGLE(c);
  step_start(c);
  int rc = finish(c, 0);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c, 0))) return rc; scalars_from_dev(c->Sh, sc); }
  return 0;
}

int imex_stage_combine(imex_ctx* c, int rks) {
  SINGLE(c);
  if (rks == 0) { bind_stage0(c); c->qe_ready = false; return 0; }
  launch_stage_combine(c->L, c->iBxn, c->iByn);
  bind_stage1(c);
  c->qe_ready = false;
  return finish(c, 0);
}

int imex_explicit_conv(imex_ctx* c) { SINGLE(c); ensure_qe(c); return finish(c, sub_explicit_conv(c)); }
int imex_mag_update_b(imex_ctx* c) { SINGLE(c); return finish(c, sub_mag_update_b(c)); }
int imex_explicit_press(imex_ctx* c) { SINGLE(c); return finish(c, sub_explicit_press(c)); }
int imex_pre_subsystem_pressure(imex_ctx* c) { SINGLE(c); return finish(c, sub_pressure(c)); }
int imex_ct_update(imex_ctx* c) { SINGLE(c); return finish(c, sub_ct(c)); }
int imex_explicit_energy(imex_ctx* c) { SINGLE(c); return finish(c, sub_explicit_energy(c)); }

#define NEEDA(c) do { if (!(c)->needA) return set_err(IMEX_ERR_UNSUPPORTED, "this magfield does not evolve the vector potential"); } while (0)
int imex_set_stage_potential(imex_ctx* c, const double* A, const double* As) {
  SINGLE(c); NEEDA(c);
  if (!A || !As) return set_err(IMEX_ERR_ARG, "null argument");
  c->cur.As = B_AS; c->cur.Ao = B_A1;
  int rc = up_pot(c, 0, B_A1, A);
  if (!rc) rc = up_pot(c, 0, B_AS, As);
  return rc;
}
int imex_get_stage_potential(imex_ctx* c, double* A, double* As) {
  SINGLE(c); NEEDA(c);
  HIPCHK(hipStreamSynchronize(c->st));
  int rc = 0;
  if (A) rc = down_pot(c, 0, c->cur.Ao, A);
  if (!rc && As) rc = down_pot(c, 0, c->cur.As, As);
  return rc;
}
// stage level: the initial guess and A_x, A_y come from the stage's A as given
int imex_mag_update_a(imex_ctx* c) { SINGLE(c); NEEDA(c); return finish(c, sub_mag_update_a(c, c->cur.Ao)); }
int imex_unstaggered_ct(imex_ctx* c) { SINGLE(c); NEEDA(c); return finish(c, sub_unstaggered_ct(c, c->cur.Ao)); }
int imex_pre_subsystem_energy(imex_ctx* c) { SINGLE(c); return finish(c, sub_energy(c)); }

// ---- operator level ----
static int prep_pressure_operator(imex_ctx* c, int op, const double* p_freeze, double mbig) {
  ensure_qe(c);
  ensure_hhat(c);
  const Dims& d = c->d0();
  if (c->P.eos != IMEX_EOS_IDEAL) launch_visc_rk(c->L, c->P, c->cur.u);
  // Tdiag via the first-pass RHS kernel (its b and p outputs are overwritten by the caller)
  launch_p_rhs(c->L, c->P, 0, c->cur.u);
  if (op == IMEX_OP_P_JACOBIAN) {
    if (!p_freeze) return set_err(IMEX_ERR_ARG, "p_freeze required");
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(c->hI[0].buf[B_P], p_freeze, d.N * sizeof(double), hipMemcpyHostToDevice));
    double Sm = mbig;
    HIPCHK(hipMemcpy(c->hI[0].S + S_MBIG, &Sm, sizeof(double), hipMemcpyHostToDevice));
    if (c->P.eos != IMEX_EOS_IDEAL) vslope_launch(c->L, c->P, B_P, c->cur.u);
  }
  return 0;
}

int imex_apply_op(imex_ctx* c, int op, const double* x, double* Ax, const double* p_freeze, double mbig) {
  SINGLE(c);
  if (!x || !Ax) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d0();
  const bool three = op == IMEX_OP_MAG || op == IMEX_OP_POTENTIAL;
  const int nc = three ? 3 : 1;
  const size_t n = (size_t)nc * d.N;
  int rc = 0;
  if (op < IMEX_OP_MAG || op > IMEX_OP_POTENTIAL) return set_err(IMEX_ERR_ARG, "bad operator");
  if (op == IMEX_OP_ENERGY) prep_energy(c);
  else if (op == IMEX_OP_POTENTIAL) launch_pot_visc(c->L, c->cur.Qs);
  else if (op != IMEX_OP_MAG && (rc = prep_pressure_operator(c, op, p_freeze, mbig))) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(c->hI[0].buf[B_XK], x, n * sizeof(double), hipMemcpyHostToDevice));
  const PIds po = op == IMEX_OP_ENERGY ? e_ids(c, 0) : p_ids(c, op == IMEX_OP_P_JACOBIAN, 0);
  op_apply(c->L, three ? 0 : 1, op == IMEX_OP_POTENTIAL ? pot_ids(c) : mag_ids(c), po, c->P, B_XK, B_AV);
  if ((rc = finish(c, 0))) return rc;
  HIPCHK(hipMemcpy(Ax, c->hI[0].buf[B_AV], n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}


int imex_gmres(imex_ctx* c, int op, const double* b, double* x, int precond, const double* p_freeze, double mbig,
               imex_solve_info* info) {
  SINGLE(c);
  if (!b || !x) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d0();
  const bool three = op == IMEX_OP_MAG || op == IMEX_OP_POTENTIAL;
  const int nc = three ? 3 : 1;
  const size_t n = (size_t)nc * d.N;
  int rc = 0;
  if (op < IMEX_OP_MAG || op > IMEX_OP_POTENTIAL) return set_err(IMEX_ERR_ARG, "bad operator");
  if (op == IMEX_OP_ENERGY) prep_energy(c);
  else if (op == IMEX_OP_POTENTIAL) launch_pot_visc(c->L, c->cur.Qs);
  else if (op != IMEX_OP_MAG && (rc = prep_pressure_operator(c, op, p_freeze, mbig))) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(c->hI[0].buf[B_BK], b, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->hI[0].buf[B_XK], x, n * sizeof(double), hipMemcpyHostToDevice));
  const int pc = precond && !three;
  const PIds po = op == IMEX_OP_ENERGY ? e_ids(c, pc) : p_ids(c, op == IMEX_OP_P_JACOBIAN, pc);
  rc = run_gmres(c, three ? 0 : 1, op == IMEX_OP_POTENTIAL ? pot_ids(c) : mag_ids(c), po, B_XK, B_BK, 0, info, nullptr, nullptr);
  int r2 = finish(c, rc);
  if (r2 < 0) return r2;
  HIPCHK(hipMemcpy(x, c->hI[0].buf[B_XK], n * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

int imex_pressure_newton(imex_ctx* c, const double* b, double* p, imex_solve_info* info) {
  SINGLE(c);
  if (!b || !p) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d0();
  int rc = prep_pressure_operator(c, IMEX_OP_P_CARRIER, nullptr, 0.0);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(c->hI[0].buf[B_PB], b, d.N * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->hI[0].buf[B_P], p, d.N * sizeof(double), hipMemcpyHostToDevice));
  rc = run_newton(c, B_PB, B_P, info, false);
  int r2 = finish(c, rc);
  if (r2 < 0) return r2;
  HIPCHK(hipMemcpy(p, c->hI[0].buf[B_P], d.N * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

int imex_profile_enable(imex_ctx* c, int on) {
  if (!c) return IMEX_ERR_ARG;
  prof_collect(c);
  Prof& P = c->prof;
  P.on = on != 0;
  for (int k = 0; k < KC_N; k++) { P.ms[k] = 0; P.bytes[k] = 0; P.n[k] = 0; }
  return 0;
}

int imex_profile_read(imex_ctx* c, imex_kernel_profile* out, int max) {
  if (!c) return IMEX_ERR_ARG;
  prof_collect(c);
  int n = 0;
  for (int k = 0; k < KC_N && n < max; k++, n++) {
    out[n].name = kc_names[k];
    out[n].launches = c->prof.n[k];
    out[n].total_ms = c->prof.ms[k];
    out[n].bytes = c->prof.bytes[k];
  }
  return n;
}

// ---- autotuner ----
// Times the candidates of each tunable on the context's current state and keeps the fastest:
// the explicit tile (16x16, 32x8, 64x4 cells; the same results bit for bit) and the grid cap of
// the Krylov kernels (1024, 2048, 4096 blocks per instance; a different cap changes the order of
// the dot-product sums, so results move by rounding).  The step's scalars and statistics are
// snapshotted and restored; the state itself is only read.
int imex_autotune(imex_ctx* c, int flags, imex_tune_info* out) {
  if (!c) return IMEX_ERR_ARG;
  if (c->capturing) return set_err(IMEX_ERR_ARG, "autotune inside a capture");
  HIPCHK(hipStreamSynchronize(c->st));
  imex_tune_info info;
  std::memset(&info, 0, sizeof(info));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_err(IMEX_ERR_HIP, "event"); }
  const Launch& L = c->L;
  auto timed = [&](int reps, auto&& body) -> double {
    body();                                   // warm-up
    (void)hipEventRecord(e0, c->st);
    for (int r = 0; r < reps; r++) body();
    (void)hipEventRecord(e1, c->st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return 1000.0 * ms / reps;
  };
  int rc = 0;
  k_budget_begin<<<dim3(L.nblocks[SH_ONE]), 64, 0, c->st>>>(L.IT, L.bm[SH_ONE]);
  step_start(c);
  bind_stage0(c);
  c->qe_ready = false;
  ensure_qe(c);
  const int caps[3] = {1024, 2048, 4096};
  // Each candidate is timed in ROUNDS interleaved rounds and keeps its fastest, so that a slow
  // moment of the device does not decide between candidates a few percent apart.
  constexpr int ROUNDS = 3;
  auto keep_min = [](double& slot, double t, int round) { slot = round == 0 ? t : std::min(slot, t); };
  if (flags & IMEX_TUNE_EXPLICIT) {
    int best = 1;
    for (int rd = 0; rd < ROUNDS && !rc; rd++)
      for (int q = 0; q < NEXT && !rc; q++) {
        c->ex_tile = EX_TILES[q][0]; c->ex_rows = EX_TILES[q][1];
        if ((rc = build_maps(c))) break;
        keep_min(info.explicit_us[q], timed(5, [&] {
          launch_explicit(L, c->P, 0, c->cur.Qe, c->cur.Qs, c->cur.u, c->mixed_flux, c->ex_tile, c->ex_rows);
          launch_explicit(L, c->P, 1, c->cur.Qe, c->cur.Qs, c->cur.u, c->mixed_flux, c->ex_tile, c->ex_rows);
        }), rd);
      }
    for (int q = 0; q < NEXT; q++) if (info.explicit_us[q] < info.explicit_us[best]) best = q;
    c->ex_tile = EX_TILES[best][0]; c->ex_rows = EX_TILES[best][1];
  }
  if (!rc && (flags & IMEX_TUNE_KRYLOV)) {
    const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
    const int j = std::min(5, m - 1);
    int best = 1;
    launch_mag_rhs(L, c->P, c->cur.u, c->cur.Qe, c->cur.Qs, B_BK, B_XK);
    for (int rd = 0; rd < ROUNDS && !rc; rd++)
      for (int q = 0; q < 3 && !rc; q++) {
        c->cap_tiles = c->cap_vec = caps[q];
        if ((rc = build_maps(c))) break;
        gmres_residual(L, 0, op_prec(c, 0), mag_ids(c), p_ids(c, 0, 0), c->P, B_XK, B_BK, 0, 0, SKIP_NONE);
        gmres_residual_fin(L, 0, op_prec(c, 0), 0, SKIP_NONE, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol);
        keep_min(info.krylov_us[q], timed(10, [&] {
          arnoldi_apply(L, 0, op_prec(c, 0), mag_ids(c), p_ids(c, 0, 0), c->P, j);
          arnoldi_update(L, 0, op_prec(c, 0), j, m);
        }), rd);
      }
    for (int q = 0; q < 3; q++) if (info.krylov_us[q] < info.krylov_us[best]) best = q;
    c->cap_tiles = c->cap_vec = caps[best];
    // pressure operator: its own cap
    const int pcaps[4] = {1024, 1536, 2048, 4096};
    int pbest = 0;
    for (int rd = 0; rd < ROUNDS && !rc; rd++)
      for (int q = 0; q < 4 && !rc; q++) {
        c->cap_tiles_p = pcaps[q];
        if ((rc = build_maps(c))) break;
        gmres_residual(L, 1, op_prec(c, 1), mag_ids(c), p_ids(c, 0, 1), c->P, B_XK, B_BK, 0, 1, SKIP_NONE);
        gmres_residual_fin(L, 1, op_prec(c, 1), 0, SKIP_NONE, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol);
        keep_min(info.krylov_p_us[q], timed(10, [&] {
          arnoldi_apply(L, 1, op_prec(c, 1), mag_ids(c), p_ids(c, 0, 1), c->P, j);
        }), rd);
      }
    for (int q = 0; q < 4; q++) if (info.krylov_p_us[q] < info.krylov_p_us[pbest]) pbest = q;
    c->cap_tiles_p = pcaps[pbest];
  }
  if (!rc) rc = build_maps(c);
  k_budget_restore<<<dim3(L.nblocks[SH_ONE]), 64, 0, c->st>>>(L.IT, L.bm[SH_ONE]);
  (void)hipStreamSynchronize(c->st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graphs(c);                 // captured launches carry the old grids
  c->qe_ready = false; c->hhat_ready = false;
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  info.explicit_tile_x = c->ex_tile;
  info.explicit_tile_y = c->ex_rows;
  info.krylov_blocks = c->cap_tiles;
  info.krylov_p_blocks = c->cap_tiles_p;
  if (out) *out = info;
  return 0;
}

}  // extern "C"

#include "adjoint_host.inc"
#include "output_host.inc"
