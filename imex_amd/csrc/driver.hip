// driver.hip — C ABI (include/imex.h) and the C++ host driver of the IMEX-MHD step on one GPU.
//
// The context replaces the reference's mutable globals (include/config.hpp:43-125) and
// its caller-owned Grid/Array objects (src/main.cpp:231-250): all fields live on the
// device for the lifetime of the context, in SoA planes with the 3-cell ghost frame.
// The step loop body of src/main.cpp:334-595 is reproduced kernel by kernel; host code
// only sequences launches and reads the GMRES / Newton convergence flags.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include "kernels.hpp"

static thread_local std::string g_err;
static int set_err(int code, const std::string& m) { g_err = m; return code; }

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return set_err(IMEX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)

// ---------------------------------------------------------------- small kernels local to the driver
__global__ void k_fill(long long n, double* p, double v) {
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}
__global__ void k_visc_rk(Dims d, Phys P, const double* pe, const double* u, double* visc, int* eflag) {
  long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= (long long)d.NX * d.NY) return;
  int ef = 0;
  visc[c] = Vbase_slope(P, pe[c], u[c], &ef);
  if (ef) atomicOr(eflag, 1);
}
// Positivity statistics of one Newton solution (implicit_operators.cpp:724-737), accumulated on the
// device: acc[0] = min p ever, acc[1] = max floored count, acc[2] = 1 if min p < theta p_floor - 1e-14.
__global__ void k_pstats(long long n, const double* p, double pfloor, double* partial) {
  __shared__ double smin[4], scnt[4];
  double mn = 1e300, cnt = 0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double v = p[e];
    mn = fmin(mn, v);
    if (v < pfloor) cnt += 1;
  }
  mn = wave_min(mn); cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) { smin[threadIdx.x >> 6] = mn; scnt[threadIdx.x >> 6] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mn = fmin(mn, smin[w]); cnt += scnt[w]; }
    partial[2 * blockIdx.x] = smin[0] < mn ? smin[0] : mn;
    partial[2 * blockIdx.x + 1] = cnt;
  }
}
__global__ void k_pstats_fin(const double* partial, int nblk, double thr, double* acc) {
  // min, and a sum of small integers held in doubles: both exact in any order, one wave reduces
  double m = 1e300, c = 0;
  for (int b = threadIdx.x; b < nblk; b += 64) { m = fmin(m, partial[2 * b]); c += partial[2 * b + 1]; }
  m = wave_min(m); c = wave_sum(c);
  if (threadIdx.x == 0) {
    acc[0] = fmin(acc[0], m);
    acc[1] = fmax(acc[1], c);
    if (m < thr) acc[2] = 1.0;
  }
}

__global__ void k_fill_skip(long long n, double* p, double v, const int* skip) {
  if (skip && *skip) return;
  long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

// ---------------------------------------------------------------- budgeted-schedule bookkeeping
// A budgeted step runs every GMRES solve for a fixed number of Arnoldi launches and every
// nested-Newton solve for a fixed number of Jacobian solves; kernels queued past convergence
// exit at once.  ctl[0] collects failures (a solve that had not converged when its budget ran
// out), ctl[1..] the iteration count of each solve, in launch order.  The scalars, positivity
// statistics and EOS flag are saved at the start so that a failed step can be rerun exactly.
__global__ void k_budget_begin(const double* S, const double* pst, const int* eflag, double* save, int* esave, int* ctl) {
  const int t = threadIdx.x;
  if (t < S_NSLOTS) save[t] = S[t];
  if (t < 4) save[S_NSLOTS + t] = pst[t];
  if (t == 0) { esave[0] = eflag[0]; ctl[0] = 0; }
}
__global__ void k_budget_restore(double* S, double* pst, int* eflag, const double* save, const int* esave) {
  const int t = threadIdx.x;
  if (t < S_NSLOTS) S[t] = save[t];
  if (t < 4) pst[t] = save[S_NSLOTS + t];
  if (t == 0) eflag[0] = esave[0];
}
// GMRES log: ctl[slot] = Arnoldi steps over all sweeps, ctl[slot+1] = sweeps, then the steps of
// each of the nsw budgeted sweeps.  A FP64-form solve has one sweep and converges inside it; a
// mixed-form solve converges at the FP64 residual that opens sweep `sweeps`.
__global__ void k_budget_gmres_check(const KrylovDev* K, int* ctl, int slot, int mixed, int nsw) {
  const int t = threadIdx.x;
  if (t == 0) {
    if (mixed) {
      ctl[slot] = K->steps_total;
      ctl[slot + 1] = K->outer;
      if (!K->outer_conv) ctl[0] |= 1;
    } else {
      ctl[slot] = K->steps;
      ctl[slot + 1] = 1;
      if (!K->converged) ctl[0] |= 1;
    }
  }
  if (t < nsw) ctl[slot + 2 + t] = mixed ? (t < K->outer ? K->sweep_steps[t] : 0) : K->steps;
}
__global__ void k_budget_newton_check(const KrylovDev* K, int* ctl, int slot) {
  if (threadIdx.x == 0) {
    ctl[slot] = K->newton_conv ? K->newton_first : -1;
    if (!K->newton_conv) ctl[0] |= 2;
  }
}

// ---------------------------------------------------------------- context
struct StageBind {
  double *u, *Qe, *Qs, *bxqs, *byqs, *bxo, *byo;
};


// ---------------------------------------------------------------- per-kernel-class event profiler
// When enabled, every launch on the step path is bracketed by two HIP events on the
// context stream and tagged with its kernel class and algorithmic byte count (distinct
// per-cell fields read once plus fields written once, times cells, times 8 bytes).
enum KClass {
  KC_START, KC_COMBINE, KC_QEPREP, KC_SHOCK, KC_EXPL_CONV, KC_EXPL_PRESS, KC_MAG_RHS, KC_MAG_APPLY, KC_MAG_UPDATE,
  KC_MAG_RES, KC_MAG_XUPD, KC_P_APPLY, KC_P_UPDATE, KC_P_RES, KC_P_XUPD, KC_NEWTON_RES, KC_EMF, KC_P_RHS, KC_P_MOM,
  KC_P_ENERGY, KC_CT, KC_BC, KC_MISC, KC_KFIN, KC_EARLY_EXIT, KC_N
};
static const char* kc_names[KC_N] = {
  "step_start", "stage_combine", "qe_prep", "shock", "explicit_conv", "explicit_press", "mag_rhs", "mag_arnoldi_apply",
  "mag_arnoldi_update", "mag_gmres_residual", "mag_gmres_xupdate", "p_arnoldi_apply", "p_arnoldi_update",
  "p_gmres_residual", "p_gmres_xupdate", "newton_residual", "emf_flux", "p_rhs", "p_momentum", "p_energy", "ct",
  "ghost_fill", "misc", "krylov_finish", "early_exit"};

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
  std::vector<int> rnres, rnresf, rfill, raxpy;   // Newton: residual (+finish) / dp fill / update per iteration
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
};
#define BHIST 8
#define MAX_GRAPHS 6
#define NBK 5


struct imex_ctx {
  Prof prof;
  imex_params prm;
  Dims d;
  Phys P;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  long long bytes = 0;
  // grids (8 planes)
  double *u_pri = nullptr, *u_n = nullptr, *U1 = nullptr, *U2 = nullptr, *Qe = nullptr, *Qs = nullptr;
  // planes
  double *bx_n, *by_n, *bx1, *by1, *bx2, *by2, *bxqs, *byqs, *fx, *fy;
  double *pe, *cf, *chi, *eta, *hhat, *mxs, *mys, *ezpp, *ezx, *ezy, *visc;
  // krylov
  double* V[MAXM + 1];
  double *w, *Av, *xk, *bk, *partial;
  double *p, *pb, *ng, *dp, *Tdiag, *vslope, *tmpN, *pst;
  KrylovDev* K = nullptr;
  KrylovDev* Kh = nullptr;     // pinned host mirrors (2 slots)
  double* S = nullptr;
  double* Sh = nullptr;        // pinned host copy of S
  int* eflag = nullptr;
  unsigned* cnt_start = nullptr;
  unsigned* cnt_pstats = nullptr;
  KrylovDev Klast;             // status of the last GMRES cycle read by the host
  hipEvent_t ev[2];
  int state_mode = 0;          // 0: u_pri holds the state; 1: Us holds the conservative state
  StageBind cur;
  bool qe_ready = false, hhat_ready = false;
  imex_stats stats{};
  int pred[2] = {6, 6};         // Arnoldi steps of the last magnetic / pressure solve
  int pred_newton = 1;         // Jacobian solves of the last nested-Newton solve
  int nblk_start = 1024;
  KrylovBufs kb;
  // state buffers: the step reads Us (or u_pri before the first step) and its last stage writes Uo
  double *U3 = nullptr, *Us = nullptr, *Uo = nullptr;
  // budgeted schedule
  bool budget = false;         // the step being queued runs on budgets
  double* bsave = nullptr;     // S, positivity statistics at the start of the budgeted step
  int* esave = nullptr;
  int* ctl = nullptr;          // device: failure flags + per-solve iteration log
  int* ctlh = nullptr;         // pinned host copy
  int nslot = 1;
  std::vector<BudgetRec> brecs;
  // budget history, per kind: 0/1 Arnoldi steps per sweep of a magnetic / pressure
  // solve, 2 Jacobian solves per Newton solve, 3/4 sweeps per magnetic / pressure solve
  int hist[NBK][BHIST];        // per-step maxima of the recent steps
  int hist_n = 0, hist_pos = 0;
  int step_mx[NBK] = {0, 0, 0, 0, 0};   // maxima of the step being run under host scheduling
  // the same per solve position (-1: no such solve in that step)
  int khist[BHIST][NGKEY][2], nhist[BHIST][NNKEY];
  int kcur[NGKEY][2], ncur[NNKEY];
  BudTab tab;
  int stage_idx = 0, sub_idx = 0, newton_it = 0;
  // mixed form
  int mixed = 0;
  float* V32[MAXM + 1];
  std::vector<GraphEntry> graphs;
  long long graph_clock = 0;
  bool capturing = false;      // launches are being captured into a step graph
  bool tag_err = false;
  bool no_toggle = false;      // node enabling failed once: graphs are captured per exact budget table
  std::vector<NodeTag> tags;
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
  return d;
}
static Phys make_phys(const imex_params& p) {
  Phys P;
  P.gamma = p.gamma; P.eos = p.eos;
  P.rk_alfa = p.rk_alfa; P.rk_b = p.rk_b; P.rk_R = p.rk_R; P.rk_cv = p.rk_cv;
  P.floor_on = p.floor_on; P.p_floor = p.p_floor; P.theta = p.theta; P.mbig_fac = p.mbig_fac;
  P.kappa1 = p.shock_kappa1; P.cfl = p.cfl; P.time_stop = p.time_stop;
  P.recon = p.recon; P.central_y = p.minmod_central_y;
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
  P.recs.push_back(ProfRec{cls, e0, e0 + 1, fields * (double)c->d.N * 8.0});
  return (int)P.recs.size() - 1;
}
#define LAUNCH(cls, fields, stmt) do { int e0_ = prof_begin(c); stmt; prof_end(c, e0_, (cls), (double)(fields)); } while (0)

template <class T>
static int dalloc(imex_ctx* c, T** out, size_t count) {
  void* ptr = nullptr;
  size_t bytes = count * sizeof(T);
  if (hipMalloc(&ptr, bytes) != hipSuccess) return set_err(IMEX_ERR_NOMEM, "hipMalloc failed");
  c->allocs.push_back(ptr);
  c->bytes += (long long)bytes;
  *out = (T*)ptr;
  return 0;
}

static inline unsigned nb(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }
static inline dim3 grid2(int nx, int ny, int bx = 64, int by = 4) { return dim3((nx + bx - 1) / bx, (ny + by - 1) / by); }

static void fill_planes(imex_ctx* c, std::initializer_list<double*> planes, int kind = 0) {
  PlaneSet ps; ps.n = 0;
  for (double* p : planes) ps.p[ps.n++] = p;
  int e0 = prof_begin(c);
  launch_bc_fill(c->d, ps, kind, c->st);
  prof_end(c, e0, KC_BC, 0.0);
}

static MagOp mag_op(imex_ctx* c) {
  MagOp m; m.d = c->d; m.Qe = c->cur.Qe; m.u = c->cur.u; m.S = c->S; return m;
}
static POp p_op(imex_ctx* c, int jacobian, int precond) {
  POp o; o.d = c->d; o.P = c->P; o.hhat = c->hhat; o.eta = c->eta; o.u = c->cur.u;
  o.visc = c->P.eos == IMEX_EOS_IDEAL ? nullptr : c->visc;
  o.pfreeze = c->p; o.vslope = c->P.eos == IMEX_EOS_IDEAL ? nullptr : c->vslope;
  o.Tdiag = c->Tdiag; o.S = c->S; o.jacobian = jacobian; o.precond = precond;
  return o;
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
// runs one restart cycle of B Arnoldi steps; the mixed form runs nsw refinement sweeps (FP64
// residual, B FP32 Arnoldi steps, x update) and the FP64 residual that must have converged.
// The kernels are the ones the host-polled loop below uses, so a solve that converges within
// its budget gives the same result bit for bit.  Launches past convergence exit at once; a solve
// still unconverged at the end of its budget flags the step for a rerun.
static int run_gmres_budget(imex_ctx* c, int op, const MagOp& mo, const POp& po, double* x, const double* b,
                            int x_is_zero, const int* skip, int parent, int it) {
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  const int nc = op == 0 ? 3 : 1;
  const int prec = c->mixed;
  BudgetRec r;
  r.kind = 0; r.op = op; r.parent = parent; r.it = it; r.key = gkey(c);
  r.B = std::min(m, std::max(1, c->tab.B[r.key]));
  r.nsw = prec ? std::min(MAXSWEEP - 1, std::max(1, c->tab.S[r.key])) : 1;
  r.slot = c->nslot; c->nslot += 2 + r.nsw;
  for (int sw = 0; sw <= r.nsw; sw++) {
    if (sw == r.nsw && !prec) break;
    const int xz = (x_is_zero && sw == 0) ? 1 : 0;
    const double res_fields = op == 0 ? (xz ? 6 : 13) : (xz ? 2 : 5) + po.precond + po.jacobian;
    const int k = r.key;
    int e0 = prof_begin(c);
    gmres_residual(op, prec, mo, po, c->kb, x, b, c->K, sw, xz, skip, c->st);
    r.rres.push_back(prof_end(c, e0, op == 0 ? KC_MAG_RES : KC_P_RES, res_fields));
    tag_last(c, TG_RES, k, sw, 0);
    e0 = prof_begin(c);
    gmres_residual_fin(c->d, prec, c->kb, c->K, sw, skip, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol, c->st);
    r.rresf.push_back(prof_end(c, e0, KC_KFIN, 0));
    tag_last(c, TG_RES, k, sw, 0);
    if (sw == r.nsw) break;
    const double vb = prec ? 0.5 : 1.0;       // bytes of a basis field relative to an FP64 field
    for (int j = 0; j < r.B; j++) {
      e0 = prof_begin(c);
      arnoldi_apply(op, prec, mo, po, c->kb, c->V, c->V32, j, c->K, c->st);
      r.ra.push_back(prof_end(c, e0, op == 0 ? KC_MAG_APPLY : KC_P_APPLY,
                              op == 0 ? 4 + vb * (9 + 3 * j) : 5 + po.jacobian + vb * (1 + j)));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_apply_fin(c->d, c->kb, j, c->K, c->st);
      r.rfa.push_back(prof_end(c, e0, KC_KFIN, 0));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_update(nc, prec, c->d, c->kb, c->V, c->V32, j, m, c->K, c->st);
      r.ru.push_back(prof_end(c, e0, op == 0 ? KC_MAG_UPDATE : KC_P_UPDATE, vb * (op == 0 ? 9 + 3 * j : 3 + j)));
      tag_last(c, TG_ARN, k, sw, j);
      e0 = prof_begin(c);
      arnoldi_update_fin(nc, prec, c->d, c->kb, j, m, c->K, c->st);
      r.rfu.push_back(prof_end(c, e0, KC_KFIN, 0));
      tag_last(c, TG_ARN, k, sw, j);
    }
    e0 = prof_begin(c);
    gmres_xupdate(nc, prec, c->d, c->kb, c->V, c->V32, x, c->K, c->st);
    r.rx.push_back(prof_end(c, e0, op == 0 ? KC_MAG_XUPD : KC_P_XUPD, (double)nc * (2 + vb * r.B)));
    tag_last(c, TG_XUPD, k, sw, 0);
  }
  k_budget_gmres_check<<<1, 64, 0, c->st>>>(c->K, c->ctl, r.slot, prec, r.nsw);
  tag_last(c, TG_GCHK, r.key, 0, 0);
  c->brecs.push_back(r);
  return 0;
}

// ---------------------------------------------------------------- GMRES host loop
// GMRES_Solver (implicit_operators.cpp:793-919).  Arnoldi steps are queued in chunks; the
// convergence flag of chunk k is read (async copy + event) while chunk k+1 is already queued.
static int run_gmres(imex_ctx* c, int op, const MagOp& mo, const POp& po, double* x, const double* b, int x_is_zero,
                     imex_solve_info* info, long long* it_acc, long long* rs_acc, const int* skip = nullptr,
                     int first_chunk = 0) {
  if (c->budget) return run_gmres_budget(c, op, mo, po, x, b, x_is_zero, skip, -1, -1);
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  const int nc = op == 0 ? 3 : 1;
  const long long dim = (long long)nc * c->d.N;
  // The reference allows dim restart cycles (implicit_operators.cpp:802); the host loop is
  // additionally capped at 4000 cycles (100k Arnoldi steps at m = 25) and raises
  // IMEX_WARN_GMRES_MAXOUTER when it gives up.
  const long long outmax = std::min<long long>(c->prm.gmres_max_outer > 0 ? c->prm.gmres_max_outer : dim, 4000);
  const int chunk = 2;
  const int prec = c->mixed;
  const double vb = prec ? 0.5 : 1.0;       // bytes of a basis field relative to an FP64 field
  if (first_chunk <= 0) first_chunk = c->pred[op];
  int total = 0, rc = 0, maxsweep = 0;
  int converged = 0;
  long long outer = 0;
  double res0 = 0, res = 0;
  for (outer = 0; outer < outmax; outer++) {
    const int xz = (x_is_zero && outer == 0) ? 1 : 0;
    const double res_fields = op == 0 ? (xz ? 6 : 13) : (xz ? 2 : 5) + po.precond + po.jacobian;
    const int* sk = (outer == 0 || prec) ? skip : nullptr;
    LAUNCH(op == 0 ? KC_MAG_RES : KC_P_RES, res_fields, gmres_residual(op, prec, mo, po, c->kb, x, b, c->K, (int)outer, xz, sk, c->st));
    LAUNCH(KC_KFIN, 0, gmres_residual_fin(c->d, prec, c->kb, c->K, (int)outer, sk, c->prm.mixed_outer_tol, c->prm.mixed_inner_tol, c->st));
    const long long gen0 = c->prof.gen;
    int recs_a[MAXM], recs_u[MAXM], recs_fa[MAXM], recs_fu[MAXM];
    int j = 0, slot = 0, last = -1;
    bool stop = false;
    while (!stop) {
      if (j >= m) break;                       // whole cycle queued; the last chunk decides
      int jend = std::min(m, j + (j == 0 ? std::max(first_chunk, 1) : chunk));
      for (; j < jend; j++) {
        int e0 = prof_begin(c);
        arnoldi_apply(op, prec, mo, po, c->kb, c->V, c->V32, j, c->K, c->st);
        recs_a[j] = prof_end(c, e0, op == 0 ? KC_MAG_APPLY : KC_P_APPLY,
                             op == 0 ? 4 + vb * (9 + 3 * j) : 5 + po.jacobian + vb * (1 + j));
        e0 = prof_begin(c);
        arnoldi_apply_fin(c->d, c->kb, j, c->K, c->st);
        recs_fa[j] = prof_end(c, e0, KC_KFIN, 0);
        e0 = prof_begin(c);
        arnoldi_update(nc, prec, c->d, c->kb, c->V, c->V32, j, m, c->K, c->st);
        recs_u[j] = prof_end(c, e0, op == 0 ? KC_MAG_UPDATE : KC_P_UPDATE, vb * (op == 0 ? 9 + 3 * j : 3 + j));
        e0 = prof_begin(c);
        arnoldi_update_fin(nc, prec, c->d, c->kb, j, m, c->K, c->st);
        recs_fu[j] = prof_end(c, e0, KC_KFIN, 0);
      }
      HIPCHK(hipMemcpyAsync(&c->Kh[slot].tol, &c->K->tol, sizeof(KrylovDev) - offsetof(KrylovDev, tol), hipMemcpyDeviceToHost, c->st));
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
      gmres_xupdate(nc, prec, c->d, c->kb, c->V, c->V32, x, c->K, c->st);
      rx = prof_end(c, e0, op == 0 ? KC_MAG_XUPD : KC_P_XUPD, (double)nc * (2 + vb * (ks.jlast + 1)));
    }
    if (c->prof.on && c->prof.gen == gen0) {      // launches queued past convergence exited at once
      auto idle = [&](int r) { if (r >= 0) { c->prof.recs[r].cls = KC_EARLY_EXIT; c->prof.recs[r].bytes = 0; } };
      for (int jj = ks.steps; jj < j; jj++) { idle(recs_a[jj]); idle(recs_u[jj]); idle(recs_fa[jj]); idle(recs_fu[jj]); }
      if (ks.jlast < 0) idle(rx);
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
  if (prec) {
    c->step_mx[op] = std::max(c->step_mx[op], maxsweep);
    c->step_mx[3 + op] = std::max(c->step_mx[3 + op], (int)outer);
    cur_gmres(c, gkey(c), maxsweep, (int)outer);
  } else {
    c->step_mx[op] = std::max(c->step_mx[op], outer == 0 ? total : MAXM + 1);
    c->step_mx[3 + op] = std::max(c->step_mx[3 + op], 1);
    cur_gmres(c, gkey(c), outer == 0 ? total : MAXM + 1, 1);
  }
  if (info) { info->iters = total; info->restarts = (int)outer; info->res0 = res0; info->res = res; info->status = rc; }
  return rc;
}

// ---------------------------------------------------------------- nested Newton, budgeted schedule
// Pressure_NestedNewton (implicit_operators.cpp:625-738) with a fixed number of Jacobian solves:
// residual, solve, update per iteration, then the residual that must have converged.  Once a
// residual has converged, the later residuals, fills, solves and updates of this Newton solve
// skip on the device.
static int run_newton_budget(imex_ctx* c, const double* b, double* p) {
  const Dims& d = c->d;
  if (!c->prm.floor_on && c->prm.eos == IMEX_EOS_IDEAL)
    return run_gmres_budget(c, 1, mag_op(c), p_op(c, 0, 1), p, b, 0, nullptr, -1, -1);
  if (c->prm.floor_on) LAUNCH(KC_MISC, 1, newton_mbig(d, c->P, b, c->cur.u, c->S, c->partial, c->K, c->eflag, c->st));
  const int NB = std::min(c->prm.newton_max, std::max(1, c->tab.N[nkey(c)]));
  const int self = (int)c->brecs.size();
  {
    BudgetRec nr;
    nr.kind = 1; nr.op = 1; nr.slot = c->nslot++; nr.key = nkey(c);
    c->brecs.push_back(nr);
  }
  const int* conv = &c->K->newton_conv;
  for (int it = 0; it <= NB; it++) {
    c->newton_it = it;
    POp carrier = p_op(c, 0, 0);
    const int nk = c->brecs[self].key;
    int e0 = prof_begin(c);
    newton_residual(carrier, c->kb, p, b, c->ng, c->K, it, c->prm.newton_tol, c->eflag, c->st);
    int rr = prof_end(c, e0, KC_NEWTON_RES, 6);
    c->brecs[self].rnres.push_back(rr);
    tag_last(c, TG_NRES, nk, 0, it);
    e0 = prof_begin(c);
    newton_residual_fin(d, c->kb, c->K, it, c->prm.newton_tol, c->st);
    rr = prof_end(c, e0, KC_KFIN, 0);
    c->brecs[self].rnresf.push_back(rr);
    tag_last(c, TG_NRES, nk, 0, it);
    if (it == NB) break;
    if (c->P.eos != IMEX_EOS_IDEAL) {
      vslope_launch(d, c->P, p, c->cur.u, c->S, c->vslope, c->eflag, c->st);
      tag_last(c, TG_NIT, nk, 0, it);
    }
    e0 = prof_begin(c);
    k_fill_skip<<<nb(d.N), 256, 0, c->st>>>(d.N, c->dp, 0.0, conv);
    rr = prof_end(c, e0, KC_MISC, 1);
    c->brecs[self].rfill.push_back(rr);
    tag_last(c, TG_NIT, nk, 0, it);
    run_gmres_budget(c, 1, mag_op(c), p_op(c, 1, 1), c->dp, c->ng, 1, conv, self, it);
    c->newton_it = 0;
    e0 = prof_begin(c);
    axpy_launch(d.N, p, c->dp, conv, c->st);
    rr = prof_end(c, e0, KC_MISC, 3);
    c->brecs[self].raxpy.push_back(rr);
    tag_last(c, TG_NIT, nk, 0, it);
  }
  k_budget_newton_check<<<1, 64, 0, c->st>>>(c->K, c->ctl, c->brecs[self].slot);
  if (c->prm.floor_on) {
    long long g = std::min<long long>(nb(d.N), 256);
    LAUNCH(KC_MISC, 1, (k_pstats<<<(unsigned)g, 256, 0, c->st>>>(d.N, p, c->prm.p_floor, c->partial),
                        k_pstats_fin<<<1, 64, 0, c->st>>>(c->partial, (int)g, c->prm.theta * c->prm.p_floor - 1e-14, c->pst)));
  }
  return 0;
}

// ---------------------------------------------------------------- nested Newton (implicit_operators.cpp:625-738)
static int run_newton(imex_ctx* c, const double* b, double* p, imex_solve_info* info) {
  if (c->budget) return run_newton_budget(c, b, p);
  int rc = 0;
  const Dims& d = c->d;
  if (!c->prm.floor_on && c->prm.eos == IMEX_EOS_IDEAL) {
    POp po = p_op(c, 0, 1);
    MagOp mo = mag_op(c);
    rc |= run_gmres(c, 1, mo, po, p, b, 0, info, &c->stats.p_gmres_iters, &c->stats.p_gmres_restarts);
    c->stats.p_solves++;
    return rc;
  }
  if (c->prm.floor_on) LAUNCH(KC_MISC, 1, newton_mbig(d, c->P, b, c->cur.u, c->S, c->partial, c->K, c->eflag, c->st));
  int it;
  double r0 = 0, r = 0;
  const int* conv = &c->K->newton_conv;
  for (it = 0; it < c->prm.newton_max; it++) {
    c->newton_it = it;
    // residual, then the Jacobian solve; the solve is skipped on the device when the residual
    // has converged, and the host learns it from the GMRES status it polls anyway
    POp carrier = p_op(c, 0, 0);
    LAUNCH(KC_NEWTON_RES, 6, newton_residual(carrier, c->kb, p, b, c->ng, c->K, it, c->prm.newton_tol, c->eflag, c->st));
    LAUNCH(KC_KFIN, 0, newton_residual_fin(d, c->kb, c->K, it, c->prm.newton_tol, c->st));
    if (c->P.eos != IMEX_EOS_IDEAL) vslope_launch(d, c->P, p, c->cur.u, c->S, c->vslope, c->eflag, c->st);
    LAUNCH(KC_MISC, 1, (k_fill<<<nb(d.N), 256, 0, c->st>>>(d.N, c->dp, 0.0)));
    POp jac = p_op(c, 1, 1);
    MagOp mo = mag_op(c);
    long long its = 0, rss = 0;
    // Newton usually needs pred_newton Jacobian solves; past that the solve is expected to be
    // skipped, so only a one-step chunk is queued speculatively
    int first = it < c->pred_newton ? c->pred[1] : 1;
    rc |= run_gmres(c, 1, mo, jac, c->dp, c->ng, 1, nullptr, &its, &rss, conv, first);
    if (it == 0) r0 = c->Klast.newton_res0;
    r = c->Klast.newton_res;
    if (c->Klast.newton_conv) { rc &= ~IMEX_WARN_GMRES_MAXOUTER; break; }
    c->stats.p_gmres_iters += its; c->stats.p_gmres_restarts += rss;
    c->stats.p_solves++;
    LAUNCH(KC_MISC, 3, axpy_launch(d.N, p, c->dp, conv, c->st));
  }
  c->newton_it = 0;
  c->stats.newton_iters += it;
  c->ncur[nkey(c)] = std::max(c->ncur[nkey(c)], it);
  c->pred_newton = std::max(1, it);
  c->step_mx[2] = std::max(c->step_mx[2], it);
  if (it == c->prm.newton_max) { rc |= IMEX_WARN_NEWTON_MAXITER; c->stats.warnings |= IMEX_WARN_NEWTON_MAXITER; }
  if (c->prm.floor_on) {
    long long g = std::min<long long>(nb(d.N), 256);
    LAUNCH(KC_MISC, 1, (k_pstats<<<(unsigned)g, 256, 0, c->st>>>(d.N, p, c->prm.p_floor, c->partial),
                        k_pstats_fin<<<1, 64, 0, c->st>>>(c->partial, (int)g, c->prm.theta * c->prm.p_floor - 1e-14, c->pst)));
  }
  if (info) { info->iters = it; info->res0 = r0; info->res = r; info->status = rc; info->restarts = 0; }
  return rc;
}

// ---------------------------------------------------------------- stage pieces
static void ensure_qe(imex_ctx* c) {
  if (c->qe_ready) return;
  const Dims& d = c->d;
  long long tot = (long long)d.NX * d.NY;
  LAUNCH(KC_QEPREP, 10, (k_qe_prep<<<nb(tot), 256, 0, c->st>>>(d, c->P, c->cur.Qe, c->pe, c->cf, c->eflag)));
  LAUNCH(KC_SHOCK, 5, (k_shock_chi<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->P, c->cur.Qe, c->cf, c->chi)));
  fill_planes(c, {c->chi});
  LAUNCH(KC_SHOCK, 2, (k_shock_eta<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->chi, c->eta)));
  fill_planes(c, {c->eta});
  c->qe_ready = true;
}

// TVDReconstruction + RusanovC_Update (kind 0) / Pressure_flux_Update (kind 1); the mixed form
// reconstructs and forms the fluxes in FP32.
static void launch_explicit(imex_ctx* c, int kind) {
  const Dims& d = c->d;
  if (c->mixed) k_explicit<float><<<grid2(d.nx, d.ny, 32, 8), dim3(256), 0, c->st>>>(d, c->P, kind, c->cur.Qe, c->cur.Qs, c->cur.u, c->S);
  else k_explicit<double><<<grid2(d.nx, d.ny, 32, 8), dim3(256), 0, c->st>>>(d, c->P, kind, c->cur.Qe, c->cur.Qs, c->cur.u, c->S);
}

static int sub_explicit_conv(imex_ctx* c) {
  const Dims& d = c->d;
  LAUNCH(KC_EXPL_CONV, 13, launch_explicit(c, 0));
  double* u = c->cur.u;
  fill_planes(c, {u, u + d.PS, u + 2 * d.PS, u + 3 * d.PS});
  c->hhat_ready = false;
  return 0;
}

static int sub_mag_update_b(imex_ctx* c) {
  const Dims& d = c->d;
  c->sub_idx = 0;
  ensure_qe(c);
  LAUNCH(KC_MAG_RHS, 23, (k_mag_rhs<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->P, c->cur.u, c->cur.Qe, c->cur.Qs, c->pe, c->bk, c->xk, c->hhat, c->S, c->eflag)));
  fill_planes(c, {c->hhat});
  c->hhat_ready = true;
  MagOp mo = mag_op(c);
  POp po = p_op(c, 0, 0);
  int rc = run_gmres(c, 0, mo, po, c->xk, c->bk, 0, nullptr, &c->stats.mag_gmres_iters, &c->stats.mag_gmres_restarts);
  if (!c->budget) c->stats.mag_solves++;
  long long tot = (long long)d.NX * d.NY;
  LAUNCH(KC_MISC, 6, (k_mag_writeback<<<nb(tot), 256, 0, c->st>>>(d, c->xk, c->cur.u)));
  LAUNCH(KC_EMF, 12, (k_emf_flux<<<grid2(d.NX, d.NY, 64, 4), dim3(64, 4), 0, c->st>>>(d, c->cur.u, c->cur.Qe, c->pe, c->fx, c->fy, c->S)));
  return rc;
}

static int sub_explicit_press(imex_ctx* c) {
  const Dims& d = c->d;
  LAUNCH(KC_EXPL_PRESS, 16, launch_explicit(c, 1));
  double* u = c->cur.u;
  fill_planes(c, {u + d.PS, u + 2 * d.PS, u + 3 * d.PS, u + 4 * d.PS});
  return 0;
}

static void ensure_hhat(imex_ctx* c) {
  if (c->hhat_ready) return;
  // hhat is produced by the magnetic RHS kernel; run it for its side output when the
  // pressure subsystem is called on its own.
  const Dims& d = c->d;
  ensure_qe(c);
  k_mag_rhs<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->P, c->cur.u, c->cur.Qe, c->cur.Qs, c->pe, c->tmpN, c->tmpN + 3 * (long long)d.N, c->hhat, c->S, c->eflag);
  fill_planes(c, {c->hhat});
  c->hhat_ready = true;
}

static int sub_pressure(imex_ctx* c) {
  const Dims& d = c->d;
  ensure_qe(c);
  ensure_hhat(c);
  double* u = c->cur.u;
  if (c->P.eos != IMEX_EOS_IDEAL) {
    long long tot = (long long)d.NX * d.NY;
    k_visc_rk<<<nb(tot), 256, 0, c->st>>>(d, c->P, c->pe, u, c->visc, c->eflag);
  }
  LAUNCH(KC_P_RHS, 15, (k_p_rhs<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, 0, u, c->hhat, c->mxs, c->mys, c->pe, c->pb, c->p, c->Tdiag, c->S)));
  fill_planes(c, {c->mxs, c->mys});
  c->sub_idx = 1;
  int rc = run_newton(c, c->pb, c->p, nullptr);
  LAUNCH(KC_P_MOM, 5, (k_p_momentum<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->p, c->mxs, c->mys, u, c->S)));
  fill_planes(c, {u + d.PS, u + 2 * d.PS});
  LAUNCH(KC_P_RHS, 12, (k_p_rhs<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, 1, u, c->hhat, c->mxs, c->mys, c->pe, c->pb, c->p, c->Tdiag, c->S)));
  c->sub_idx = 2;
  rc |= run_newton(c, c->pb, c->p, nullptr);
  LAUNCH(KC_P_ENERGY, 9, (k_p_energy<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, c->P, c->p, u, c->eflag)));
  fill_planes(c, {u + 4 * d.PS});
  return rc;
}

static int sub_ct(imex_ctx* c) {
  const Dims& d = c->d;
  const double* ex = c->fx; const double* ey = c->fy;
  double sx = -1.0, sy = 1.0;
  if (c->prm.ct_mode == IMEX_CT_PLAIN) {
    k_ct_plain_emf<<<grid2(d.NX, d.NY), dim3(64, 4), 0, c->st>>>(d, c->P, c->cur.u, c->ezx, c->ezy);
    ex = c->ezx; ey = c->ezy; sx = -1.0; sy = 1.0;
  }
  LAUNCH(KC_CT, 8, (k_ct_corner<<<grid2(d.nx + 2, d.ny + 2), dim3(64, 4), 0, c->st>>>(d, ex, ey, sx, sy, c->cur.u, c->ezpp)));
  LAUNCH(KC_CT, 7, (k_ct_faces<<<grid2(d.NX, d.NY), dim3(64, 4), 0, c->st>>>(d, c->cur.bxqs, c->cur.byqs, c->ezpp, c->cur.bxo, c->cur.byo, c->cur.u, c->S)));
  return 0;
}

static int run_stage(imex_ctx* c) {
  int rc = 0;
  c->qe_ready = false; c->hhat_ready = false;
  ensure_qe(c);
  rc |= sub_explicit_conv(c);
  rc |= sub_mag_update_b(c);
  rc |= sub_explicit_press(c);
  rc |= sub_pressure(c);
  rc |= sub_ct(c);
  return rc;
}

static void bind_stage0(imex_ctx* c) {
  c->stage_idx = 0;
  c->cur.u = c->U1; c->cur.Qe = c->u_n; c->cur.Qs = c->u_n;
  c->cur.bxqs = c->bx_n; c->cur.byqs = c->by_n; c->cur.bxo = c->bx1; c->cur.byo = c->by1;
}
static void bind_stage1(imex_ctx* c) {
  c->stage_idx = 1;
  c->cur.u = c->Uo; c->cur.Qe = c->Qe; c->cur.Qs = c->Qs;
  c->cur.bxqs = c->bxqs; c->cur.byqs = c->byqs; c->cur.bxo = c->bx2; c->cur.byo = c->by2;
}

static int step_start(imex_ctx* c) {
  const Dims& d = c->d;
  long long tot = (long long)d.NX * d.NY;
  long long g = std::min<long long>(nb(tot), c->nblk_start);
  const double* in = c->state_mode ? c->Us : c->u_pri;
  LAUNCH(KC_START, 16, (k_step_start<<<(unsigned)g, 256, 0, c->st>>>(d, c->P, c->state_mode, in, c->u_n, c->partial, c->eflag),
                        k_step_finish<<<1, 256, 0, c->st>>>(d, c->P, c->partial, (int)g, c->S)));
  return 0;
}

// Everything a step queues; the step's input (Us or u_pri, bx_n, by_n) is only read, so a
// discarded step can be queued again from the same input.
static int step_body(imex_ctx* c) {
  int rc = 0;
  const Dims& d = c->d;
  step_start(c);
  bind_stage0(c);
  rc |= run_stage(c);
  long long tot = (long long)d.NX * d.NY;
  LAUNCH(KC_COMBINE, 38, (k_stage_combine<<<nb(tot), 256, 0, c->st>>>(d, c->u_n, c->U1, c->Qe, c->Qs, c->bx_n, c->bx1, c->by_n, c->by1, c->bxqs, c->byqs, c->S)));
  bind_stage1(c);
  rc |= run_stage(c);
  return rc;
}

static void step_commit(imex_ctx* c) {
  std::swap(c->bx_n, c->bx2); std::swap(c->by_n, c->by2); std::swap(c->Us, c->Uo);
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
// steps per GMRES sweep.  Returns false (host scheduling) until there is history, or when a
// recent solve needed a restart.
static bool choose_budgets(imex_ctx* c) {
  if (c->prm.schedule == IMEX_SCHED_HOST || c->hist_n == 0) return false;
  const int m = std::min(std::max(c->prm.gmres_restart, 1), MAXM);
  int mx[NBK] = {0, 0, 0, 0, 0};
  for (int i = 0; i < c->hist_n; i++) for (int k = 0; k < NBK; k++) mx[k] = std::max(mx[k], c->hist[k][i]);
  if (mx[0] >= m || mx[1] >= m) return false;              // a recent solve needed a restart
  if (mx[3] >= MAXSWEEP - 1 || mx[4] >= MAXSWEEP - 1) return false;
  for (int k = 0; k < NGKEY; k++) {
    const int op = (k / NIT) % 3 == 0 ? 0 : 1;
    int st = -1, sw = -1;
    for (int i = 0; i < c->hist_n; i++) { st = std::max(st, c->khist[i][k][0]); sw = std::max(sw, c->khist[i][k][1]); }
    if (st < 0) { st = mx[op]; sw = mx[3 + op]; }
    c->tab.B[k] = std::min(m, std::max(1, st + c->prm.budget_margin));
    c->tab.S[k] = std::max(1, sw);
  }
  for (int k = 0; k < NNKEY; k++) {
    int nb = -1;
    for (int i = 0; i < c->hist_n; i++) nb = std::max(nb, c->nhist[i][k]);
    c->tab.N[k] = std::max(1, nb < 0 ? mx[2] : nb);
  }
  return true;
}

// Reads the step's log: returns 1 if a solve outran its budget, else books the iteration counts
// and, with the profiler on, moves launches that found their solve converged to early_exit.
static int budget_finish(imex_ctx* c, long long gen0) {
  HIPCHK(hipMemcpyAsync(c->ctlh, c->ctl, (size_t)c->nslot * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipEventRecord(c->ev[0], c->st));
  HIPCHK(hipEventSynchronize(c->ev[0]));
  if (c->ctlh[0]) return 1;
  const bool fix = c->prof.on && c->prof.gen == gen0;
  auto idle = [&](int r) { if (fix && r >= 0) { c->prof.recs[r].cls = KC_EARLY_EXIT; c->prof.recs[r].bytes = 0; } };
  int mx[NBK] = {0, 0, 0, 0, 0};
  for (const BudgetRec& r : c->brecs) {
    if (r.kind == 1) {
      const int nf = c->ctlh[r.slot];
      c->stats.newton_iters += nf;
      mx[2] = std::max(mx[2], nf);
      c->ncur[r.key] = std::max(c->ncur[r.key], nf);
      for (size_t i = 0; i < r.rnres.size(); i++) if ((int)i > nf) { idle(r.rnres[i]); idle(r.rnresf[i]); }
      for (size_t i = 0; i < r.rfill.size(); i++) if ((int)i >= nf) { idle(r.rfill[i]); idle(r.raxpy[i]); }
      continue;
    }
    const int steps = c->ctlh[r.slot], sweeps = c->ctlh[r.slot + 1];
    const bool skipped = r.parent >= 0 && r.it >= c->ctlh[c->brecs[r.parent].slot];
    for (int sw = 0; sw < (int)r.rres.size(); sw++) {
      const bool live = !skipped && sw <= sweeps;                       // residuals that ran
      if (!live) { idle(r.rres[sw]); idle(r.rresf[sw]); }
    }
    for (int sw = 0; sw < r.nsw; sw++) {
      const int ss = (!skipped && sw < sweeps) ? c->ctlh[r.slot + 2 + sw] : 0;   // Arnoldi steps of this sweep
      for (int j = ss; j < r.B; j++) {
        const size_t q = (size_t)sw * r.B + j;
        idle(r.ra[q]); idle(r.ru[q]); idle(r.rfa[q]); idle(r.rfu[q]);
      }
      if (ss == 0) idle(r.rx[sw]);
      if (!skipped && sw < sweeps) { mx[r.op] = std::max(mx[r.op], ss); cur_gmres(c, r.key, ss, 0); }
    }
    if (skipped) continue;
    mx[3 + r.op] = std::max(mx[3 + r.op], sweeps);
    cur_gmres(c, r.key, c->mixed ? 0 : steps, sweeps);
    if (c->mixed) c->stats.mixed_refinements += sweeps;
    if (r.op == 0) { c->stats.mag_solves++; c->stats.mag_gmres_iters += steps; }
    else { c->stats.p_solves++; c->stats.p_gmres_iters += steps; }
  }
  hist_push(c, mx);
  return 0;
}

// Whether a tagged node runs under budget table tb.
static bool tag_enabled(const NodeTag& t, const BudTab& tb, int mixed) {
  if (t.kind == TG_NIT) return t.j < tb.N[t.key];
  if (t.kind == TG_NRES) return t.j <= tb.N[t.key];
  const int sub = (t.key / NIT) % 3, it = t.key % NIT;
  const bool live = sub == 0 || it < tb.N[t.key / NIT];      // the Newton iteration holding the solve runs
  if (!live) return false;
  const int S = mixed ? tb.S[t.key] : 1;
  switch (t.kind) {
    case TG_ARN: return t.sw < S && t.j < tb.B[t.key];
    case TG_RES: return !mixed || t.sw <= S;
    case TG_XUPD: return t.sw < S;
    default: return true;                                     // TG_GCHK
  }
}

static bool tab_covers(const BudTab& cap, const BudTab& t, int mixed) {
  for (int k = 0; k < NGKEY; k++) {
    if (t.B[k] > cap.B[k]) return false;
    if (mixed && t.S[k] > cap.S[k]) return false;
  }
  for (int k = 0; k < NNKEY; k++) if (t.N[k] > cap.N[k]) return false;
  return true;
}

// Queue (or replay) the budgeted step as a captured hipGraph.  A graph is captured for a
// capacity table a little above the current budgets (and for the current state/face buffer set
// and state mode); replays toggle the budget-dependent nodes to the current table.
static int budget_graph_launch(imex_ctx* c) {
  const int parity = c->Us == c->U2 ? 0 : 1;
  const BudTab cur = c->tab;
  GraphEntry* g = nullptr;
  for (GraphEntry& e : c->graphs) {
    if (e.parity != parity || e.state_mode != c->state_mode) continue;
    if (c->no_toggle ? std::memcmp(&e.cap, &cur, sizeof(BudTab)) == 0 : tab_covers(e.cap, cur, c->mixed)) { g = &e; break; }
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
        cap.B[k] = std::min(m, cur.B[k] + 2);
        cap.S[k] = c->mixed ? std::min(MAXSWEEP - 1, cur.S[k] + 1) : 1;
      }
      for (int k = 0; k < NNKEY; k++) cap.N[k] = std::min(c->prm.newton_max, cur.N[k] + 1);
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
      const bool en = tag_enabled(t, cur, c->mixed);
      if (en == tag_enabled(t, g->applied, c->mixed)) continue;
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
  cur_reset(c);
  if (choose_budgets(c)) {
    const long long gen0 = c->prof.gen;
    c->budget = true;
    c->nslot = 1;
    c->brecs.clear();
    k_budget_begin<<<1, 64, 0, c->st>>>(c->S, c->pst, c->eflag, c->bsave, c->esave, c->ctl);
    if (c->prm.schedule == IMEX_SCHED_GRAPH && !c->prof.on) rc = budget_graph_launch(c);
    else rc = step_body(c);
    c->budget = false;
    if (rc < 0) return rc;
    if (c->nslot > 256) return set_err(IMEX_ERR_ARG, "too many solves in one step");
    int f = budget_finish(c, gen0);
    if (f < 0) return f;
    c->stats.budget_steps++;
    if (f == 0) { step_commit(c); return rc; }
    // a solve outran its budget: restore the step's scalars and rerun it under host scheduling
    c->stats.budget_reruns++;
    k_budget_restore<<<1, 64, 0, c->st>>>(c->S, c->pst, c->eflag, c->bsave, c->esave);
    cur_reset(c);
  }
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
static int up_grid(imex_ctx* c, double* dst, const double* aos) {
  std::vector<double> s; aos_to_soa(c->d, aos, s);
  HIPCHK(hipMemcpy(dst, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int down_grid(imex_ctx* c, const double* src, double* aos) {
  std::vector<double> s((size_t)NVAR * c->d.PS);
  HIPCHK(hipMemcpy(s.data(), src, s.size() * sizeof(double), hipMemcpyDeviceToHost));
  soa_to_aos(c->d, s, aos);
  return 0;
}
static int up_plane(imex_ctx* c, double* dst, const double* a) {
  HIPCHK(hipMemcpy(dst, a, (size_t)c->d.NX * c->d.NY * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int down_plane(imex_ctx* c, const double* src, double* a) {
  HIPCHK(hipMemcpy(a, src, (size_t)c->d.NX * c->d.NY * sizeof(double), hipMemcpyDeviceToHost));
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
static int pull_scalars(imex_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->Sh, c->S, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// ---------------------------------------------------------------- host boundary fills for initial data
static int hmapx(const imex_params& p, int i) {
  if (i < GC) return p.bc_x == IMEX_BC_PERIODIC ? p.nx + i : GC;
  if (i > p.nx + GC - 1) return p.bc_x == IMEX_BC_PERIODIC ? i - p.nx : p.nx + GC - 1;
  return i;
}
static int hmapy(const imex_params& p, int j) {
  if (j < GC) return p.bc_y == IMEX_BC_PERIODIC ? p.ny + j : GC;
  if (j > p.ny + GC - 1) return p.bc_y == IMEX_BC_PERIODIC ? j - p.ny : p.ny + GC - 1;
  return j;
}
// Ghost fill with the reference's pass order (x over all rows, then y over all columns).
static void hfill(const imex_params& p, double* a, int nvar, int kind) {
  const int NX = p.nx + 2 * GC, NY = p.ny + 2 * GC;
  auto mx = [&](int i) {
    if (p.bc_x == IMEX_BC_PERIODIC) return hmapx(p, i);
    int lo = kind == 1 ? GC - 1 : GC;
    return i < lo ? lo : (i > p.nx + GC - 1 ? p.nx + GC - 1 : i);
  };
  auto my = [&](int j) {
    if (p.bc_y == IMEX_BC_PERIODIC) return hmapy(p, j);
    int lo = kind == 2 ? GC - 1 : GC;
    return j < lo ? lo : (j > p.ny + GC - 1 ? p.ny + GC - 1 : j);
  };
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      int si = mx(i);
      if (si != i) for (int k = 0; k < nvar; k++) a[(size_t)nvar * (i + j * NX) + k] = a[(size_t)nvar * (si + j * NX) + k];
    }
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY; j++) {
      int sj = my(j);
      if (sj != j) for (int k = 0; k < nvar; k++) a[(size_t)nvar * (i + j * NX) + k] = a[(size_t)nvar * (i + sj * NX) + k];
    }
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
  p->schedule = IMEX_SCHED_GRAPH;
  p->budget_margin = 1;
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
      } else {
        return set_err(IMEX_ERR_ARG, "unknown case");
      }
    }
  hfill(*p, u, NVAR, 0);
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
    hfill(*p, u, NVAR, 0);
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

int imex_create(const imex_params* prm, imex_ctx** out) {
  if (!prm || !out) return set_err(IMEX_ERR_ARG, "null argument");
  if (prm->nx < 4 || prm->ny < 4) return set_err(IMEX_ERR_ARG, "grid too small (need nx, ny >= 4)");
  if (prm->gmres_restart > MAXM) return set_err(IMEX_ERR_ARG, "gmres_restart exceeds 32");
  if ((prm->bc_x != IMEX_BC_PERIODIC && prm->bc_x != IMEX_BC_TRANSMISSIVE) ||
      (prm->bc_y != IMEX_BC_PERIODIC && prm->bc_y != IMEX_BC_TRANSMISSIVE))
    return set_err(IMEX_ERR_UNSUPPORTED, "only periodic and transmissive boundaries are supported");
  if (prm->recon != IMEX_RECON_MINMOD && prm->recon != IMEX_RECON_VANLEER && prm->recon != IMEX_RECON_WENO5)
    return set_err(IMEX_ERR_UNSUPPORTED, "reconstruction must be Minmod, VanLeer or WENO5");
  if (prm->precision != IMEX_PREC_FP64 && prm->precision != IMEX_PREC_MIXED)
    return set_err(IMEX_ERR_ARG, "bad precision");
  if (prm->schedule != IMEX_SCHED_GRAPH && prm->schedule != IMEX_SCHED_BUDGET && prm->schedule != IMEX_SCHED_HOST)
    return set_err(IMEX_ERR_ARG, "bad schedule");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(IMEX_ERR_HIP, "no HIP device");
  imex_ctx* c = new imex_ctx();
  c->prm = *prm;
  c->d = make_dims(*prm);
  c->P = make_phys(*prm);
  const Dims& d = c->d;
  int rc = 0;
  auto A = [&](double** p, size_t n) { if (!rc) rc = dalloc(c, p, n); };
  size_t G8 = (size_t)NVAR * d.PS, G1 = (size_t)d.PS;
  A(&c->u_pri, G8); A(&c->u_n, G8); A(&c->U1, G8); A(&c->U2, G8); A(&c->U3, G8); A(&c->Qe, G8); A(&c->Qs, G8);
  double** planes[] = {&c->bx_n, &c->by_n, &c->bx1, &c->by1, &c->bx2, &c->by2, &c->bxqs, &c->byqs, &c->fx, &c->fy,
                       &c->pe, &c->cf, &c->chi, &c->eta, &c->hhat, &c->mxs, &c->mys, &c->ezpp, &c->ezx, &c->ezy, &c->visc};
  for (double** pp : planes) A(pp, G1);
  size_t n3 = (size_t)3 * d.N;
  int m = std::min(std::max(prm->gmres_restart, 1), MAXM);
  for (int i = 0; i <= MAXM; i++) {
    if (i <= m) A(&c->V[i], n3); else c->V[i] = c->V[0];
  }
  A(&c->w, n3); A(&c->Av, n3); A(&c->xk, n3); A(&c->bk, n3); A(&c->tmpN, 6 * (size_t)d.N);
  c->mixed = prm->precision == IMEX_PREC_MIXED ? 1 : 0;
  c->kb.w32 = nullptr; c->kb.Av32 = nullptr;
  for (int i = 0; i <= MAXM; i++) c->V32[i] = nullptr;
  if (c->mixed) {       // FP32 basis of the inner solves
    auto AF = [&](float** pp, size_t n) { if (!rc) rc = dalloc(c, pp, n); };
    for (int i = 0; i <= m; i++) AF(&c->V32[i], n3);
    for (int i = m + 1; i <= MAXM; i++) c->V32[i] = c->V32[0];
    AF(&c->kb.w32, n3); AF(&c->kb.Av32, n3);
  }
  c->nblk_start = 1024;
  c->kb.nblk_apply = 2048; c->kb.nblk_vec = 2048;
  A(&c->partial, (size_t)4096 * (MAXM + 2));
  A(&c->p, d.N); A(&c->pb, d.N); A(&c->ng, d.N); A(&c->dp, d.N); A(&c->Tdiag, d.N); A(&c->vslope, d.N); A(&c->pst, 4);
  if (!rc) rc = dalloc(c, &c->K, 1);
  if (!rc) rc = dalloc(c, &c->S, S_NSLOTS);
  if (!rc) rc = dalloc(c, &c->eflag, 1);
  if (!rc) rc = dalloc(c, &c->cnt_start, 1);
  if (!rc) rc = dalloc(c, &c->cnt_pstats, 1);
  if (!rc) rc = dalloc(c, &c->bsave, S_NSLOTS + 4);
  if (!rc) rc = dalloc(c, &c->esave, 1);
  if (!rc) rc = dalloc(c, &c->ctl, 256);
  if (rc) { imex_destroy(c); return rc; }
  c->Us = c->U2; c->Uo = c->U3;
  c->kb.w = c->w; c->kb.Av = c->Av; c->kb.partial = c->partial;     // w32 / Av32 set above
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { imex_destroy(c); return set_err(IMEX_ERR_HIP, "stream"); }
  if (hipHostMalloc((void**)&c->Kh, 2 * sizeof(KrylovDev), hipHostMallocDefault) != hipSuccess) { imex_destroy(c); return set_err(IMEX_ERR_HIP, "pinned"); }
  if (hipHostMalloc((void**)&c->Sh, S_NSLOTS * sizeof(double), hipHostMallocDefault) != hipSuccess) { imex_destroy(c); return set_err(IMEX_ERR_HIP, "pinned"); }
  if (hipHostMalloc((void**)&c->ctlh, 256 * sizeof(int), hipHostMallocDefault) != hipSuccess) { imex_destroy(c); return set_err(IMEX_ERR_HIP, "pinned"); }
  (void)hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming);
  (void)hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming);
  // zero every buffer once (outside any captured region)
  KrylovDev kz; std::memset(&kz, 0, sizeof(kz)); kz.tol = prm->gmres_tol;
  (void)hipMemcpy(c->K, &kz, sizeof(kz), hipMemcpyHostToDevice);
  std::memset(c->Sh, 0, S_NSLOTS * sizeof(double));
  (void)hipMemcpy(c->S, c->Sh, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemset(c->eflag, 0, sizeof(int));
  (void)hipMemset(c->cnt_start, 0, sizeof(unsigned));
  (void)hipMemset(c->cnt_pstats, 0, sizeof(unsigned));
  (void)hipMemset(c->ctl, 0, 256 * sizeof(int));
  for (double** pp : planes) (void)hipMemset(*pp, 0, G1 * sizeof(double));
  for (double* g8 : {c->U1, c->U2, c->U3, c->Qe, c->Qs, c->u_n, c->u_pri}) (void)hipMemset(g8, 0, G8 * sizeof(double));
  if (hipDeviceSynchronize() != hipSuccess) { imex_destroy(c); return set_err(IMEX_ERR_HIP, "init sync"); }
  c->stats.min_p = 1e300;
  { double init[4] = {1e300, 0.0, 0.0, 0.0}; (void)hipMemcpy(c->pst, init, sizeof(init), hipMemcpyHostToDevice); }
  bind_stage0(c);
  *out = c;
  return 0;
}

int imex_destroy(imex_ctx* c) {
  if (!c) return 0;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->Kh) (void)hipHostFree(c->Kh);
  if (c->Sh) (void)hipHostFree(c->Sh);
  if (c->ctlh) (void)hipHostFree(c->ctlh);
  for (GraphEntry& g : c->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
  for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
  if (c->st) { (void)hipEventDestroy(c->ev[0]); (void)hipEventDestroy(c->ev[1]); (void)hipStreamDestroy(c->st); }
  delete c;
  return 0;
}

long long imex_device_bytes(const imex_ctx* c) { return c ? c->bytes : 0; }

int imex_set_state(imex_ctx* c, const double* u_pri, const double* bx, const double* by, double t, int timestep, double prev_dt) {
  if (!c || !u_pri || !bx || !by) return set_err(IMEX_ERR_ARG, "null argument");
  std::vector<double> tmp(u_pri, u_pri + (size_t)NVAR * c->d.NX * c->d.NY);
  hfill(c->prm, tmp.data(), NVAR, 0);           // bound_cond_u(u_pri) at the top of the step (main.cpp:336)
  int rc = up_grid(c, c->u_pri, tmp.data());
  if (!rc) rc = up_plane(c, c->bx_n, bx);
  if (!rc) rc = up_plane(c, c->by_n, by);
  if (rc) return rc;
  std::memset(c->Sh, 0, S_NSLOTS * sizeof(double));
  c->Sh[S_T] = t; c->Sh[S_TIMESTEP] = timestep; c->Sh[S_PREVDT] = prev_dt;
  HIPCHK(hipMemcpy(c->S, c->Sh, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice));
  c->state_mode = 0;
  c->hist_n = 0; c->hist_pos = 0;       // iteration budgets are relearned for the new state
  return 0;
}

int imex_get_state(imex_ctx* c, double* u_con, double* u_pri, double* bx, double* by, imex_scalars* sc) {
  if (!c) return set_err(IMEX_ERR_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->st));
  const Dims& d = c->d;
  long long tot = (long long)d.NX * d.NY;
  if (c->state_mode == 1) {
    k_con2pri<<<nb(tot), 256, 0, c->st>>>(d, c->P, c->Us, c->u_pri, c->eflag);
  } else {
    k_pri2con<<<nb(tot), 256, 0, c->st>>>(d, c->P, c->u_pri, c->Us, c->eflag);
  }
  HIPCHK(hipStreamSynchronize(c->st));
  int rc = 0;
  if (u_con) rc = down_grid(c, c->Us, u_con);
  if (!rc && u_pri) rc = down_grid(c, c->u_pri, u_pri);
  if (!rc && bx) rc = down_plane(c, c->bx_n, bx);
  if (!rc && by) rc = down_plane(c, c->by_n, by);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c))) return rc; scalars_from_dev(c->Sh, sc); }
  return 0;
}

int imex_step_async(imex_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return set_err(IMEX_ERR_ARG, "bad argument");
  int rc = 0;
  for (int s = 0; s < nsteps; s++) {
    int r = one_step(c);
    if (r < 0) return r;
    rc |= r;
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
  int ef = 0;
  HIPCHK(hipMemcpy(&ef, c->eflag, sizeof(int), hipMemcpyDeviceToHost));
  if (ef) c->stats.warnings |= IMEX_WARN_EOS_MAXITER;
  double acc[4];
  HIPCHK(hipMemcpy(acc, c->pst, sizeof(acc), hipMemcpyDeviceToHost));
  c->stats.min_p = acc[0];
  c->stats.max_floored = (int)acc[1];
  if (acc[2] != 0.0) c->stats.warnings |= IMEX_WARN_POSITIVITY;
  *s = c->stats;
  s->div_b = imex_div_b(c);
  return 0;
}

int imex_reset_stats(imex_ctx* c) {
  if (!c) return IMEX_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->st));
  c->stats = imex_stats{};
  c->stats.min_p = 1e300;
  double init[4] = {1e300, 0.0, 0.0, 0.0};
  HIPCHK(hipMemcpy(c->pst, init, sizeof(init), hipMemcpyHostToDevice));
  int zero = 0;
  HIPCHK(hipMemcpy(c->eflag, &zero, sizeof(int), hipMemcpyHostToDevice));
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

double imex_div_b(imex_ctx* c) {
  if (!c) return NAN;
  const Dims& d = c->d;
  int g = 256;
  k_divb<<<g, 256, 0, c->st>>>(d, c->bx_n, c->by_n, c->partial);
  std::vector<double> h(g);
  if (hipMemcpyAsync(h.data(), c->partial, g * sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess) return NAN;
  if (hipStreamSynchronize(c->st) != hipSuccess) return NAN;
  double s = 0; for (double v : h) s += v;
  return sqrt(s);
}

// ---- stage-level drop-ins ----
int imex_set_stage(imex_ctx* c, const imex_stage_host* st, const imex_scalars* sc) {
  if (!c || !st || !sc) return set_err(IMEX_ERR_ARG, "null argument");
  c->cur.u = c->U1; c->cur.Qe = c->Qe; c->cur.Qs = c->Qs;
  c->cur.bxqs = c->bxqs; c->cur.byqs = c->byqs; c->cur.bxo = c->bx1; c->cur.byo = c->by1;
  int rc = up_grid(c, c->U1, st->u_con);
  if (!rc) rc = up_grid(c, c->Qe, st->Qe);
  if (!rc) rc = up_grid(c, c->Qs, st->Qs);
  if (!rc && st->bx) rc = up_plane(c, c->bx1, st->bx);
  if (!rc && st->by) rc = up_plane(c, c->by1, st->by);
  if (!rc && st->bx_qs) rc = up_plane(c, c->bxqs, st->bx_qs);
  if (!rc && st->by_qs) rc = up_plane(c, c->byqs, st->by_qs);
  if (!rc && st->fx) rc = up_plane(c, c->fx, st->fx);
  if (!rc && st->fy) rc = up_plane(c, c->fy, st->fy);
  if (rc) return rc;
  std::memset(c->Sh, 0, S_NSLOTS * sizeof(double));
  scalars_to_dev(sc, c->Sh);
  HIPCHK(hipMemcpy(c->S, c->Sh, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice));
  c->qe_ready = false; c->hhat_ready = false;
  return 0;
}

int imex_get_stage(imex_ctx* c, imex_stage_host* st, imex_scalars* sc) {
  if (!c || !st) return set_err(IMEX_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(c->st));
  int rc = 0;
  if (st->u_con) rc = down_grid(c, c->cur.u, st->u_con);
  if (!rc && st->Qe) rc = down_grid(c, c->cur.Qe, st->Qe);
  if (!rc && st->Qs) rc = down_grid(c, c->cur.Qs, st->Qs);
  if (!rc && st->bx) rc = down_plane(c, c->cur.bxo, st->bx);
  if (!rc && st->by) rc = down_plane(c, c->cur.byo, st->by);
  if (!rc && st->bx_qs) rc = down_plane(c, c->cur.bxqs, st->bx_qs);
  if (!rc && st->by_qs) rc = down_plane(c, c->cur.byqs, st->by_qs);
  if (!rc && st->fx) rc = down_plane(c, c->fx, st->fx);
  if (!rc && st->fy) rc = down_plane(c, c->fy, st->fy);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c))) return rc; scalars_from_dev(c->Sh, sc); }
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
  if (!c) return IMEX_ERR_ARG;
  step_start(c);
  int rc = finish(c, 0);
  if (rc) return rc;
  if (sc) { if ((rc = pull_scalars(c))) return rc; scalars_from_dev(c->Sh, sc); }
  return 0;
}

int imex_stage_combine(imex_ctx* c, int rks) {
  if (!c) return IMEX_ERR_ARG;
  if (rks == 0) { bind_stage0(c); c->qe_ready = false; return 0; }
  const Dims& d = c->d;
  long long tot = (long long)d.NX * d.NY;
  k_stage_combine<<<nb(tot), 256, 0, c->st>>>(d, c->u_n, c->U1, c->Qe, c->Qs, c->bx_n, c->bx1, c->by_n, c->by1, c->bxqs, c->byqs, c->S);
  bind_stage1(c);
  c->qe_ready = false;
  return finish(c, 0);
}

int imex_explicit_conv(imex_ctx* c) { if (!c) return IMEX_ERR_ARG; ensure_qe(c); return finish(c, sub_explicit_conv(c)); }
int imex_mag_update_b(imex_ctx* c) { if (!c) return IMEX_ERR_ARG; return finish(c, sub_mag_update_b(c)); }
int imex_explicit_press(imex_ctx* c) { if (!c) return IMEX_ERR_ARG; return finish(c, sub_explicit_press(c)); }
int imex_pre_subsystem_pressure(imex_ctx* c) { if (!c) return IMEX_ERR_ARG; return finish(c, sub_pressure(c)); }
int imex_ct_update(imex_ctx* c) { if (!c) return IMEX_ERR_ARG; return finish(c, sub_ct(c)); }

// ---- operator level ----
static int prep_pressure_operator(imex_ctx* c, int op, const double* p_freeze, double mbig) {
  ensure_qe(c);
  ensure_hhat(c);
  const Dims& d = c->d;
  if (c->P.eos != IMEX_EOS_IDEAL) {
    long long tot = (long long)d.NX * d.NY;
    k_visc_rk<<<nb(tot), 256, 0, c->st>>>(d, c->P, c->pe, c->cur.u, c->visc, c->eflag);
  }
  // Tdiag via the first-pass RHS kernel (its b / p outputs go to scratch)
  k_p_rhs<<<grid2(d.nx, d.ny), dim3(64, 4), 0, c->st>>>(d, 0, c->cur.u, c->hhat, c->mxs, c->mys, c->pe, c->tmpN, c->tmpN + d.N, c->Tdiag, c->S);
  if (op == IMEX_OP_P_JACOBIAN) {
    if (!p_freeze) return set_err(IMEX_ERR_ARG, "p_freeze required");
    HIPCHK(hipMemcpy(c->p, p_freeze, d.N * sizeof(double), hipMemcpyHostToDevice));
    double Sm = mbig;
    HIPCHK(hipMemcpy(c->S + S_MBIG, &Sm, sizeof(double), hipMemcpyHostToDevice));
    if (c->P.eos != IMEX_EOS_IDEAL) vslope_launch(d, c->P, c->p, c->cur.u, c->S, c->vslope, c->eflag, c->st);
  }
  return 0;
}

int imex_apply_op(imex_ctx* c, int op, const double* x, double* Ax, const double* p_freeze, double mbig) {
  if (!c || !x || !Ax) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d;
  int nc = op == IMEX_OP_MAG ? 3 : 1;
  size_t n = (size_t)nc * d.N;
  int rc = 0;
  if (op != IMEX_OP_MAG && (rc = prep_pressure_operator(c, op, p_freeze, mbig))) return rc;
  HIPCHK(hipMemcpy(c->tmpN + 3 * (size_t)d.N, x, n * sizeof(double), hipMemcpyHostToDevice));
  MagOp mo = mag_op(c);
  POp po = p_op(c, op == IMEX_OP_P_JACOBIAN, 0);
  op_apply(op == IMEX_OP_MAG ? 0 : 1, mo, po, c->tmpN + 3 * (size_t)d.N, c->tmpN, c->st);
  if ((rc = finish(c, 0))) return rc;
  HIPCHK(hipMemcpy(Ax, c->tmpN, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int imex_apply_op_T(imex_ctx* c, int op, const double* x, double* ATx, const double* p_freeze, double mbig) {
  (void)c; (void)op; (void)x; (void)ATx; (void)p_freeze; (void)mbig;
  return set_err(IMEX_ERR_UNSUPPORTED, "transposed operators are not built yet");
}

int imex_gmres(imex_ctx* c, int op, const double* b, double* x, int precond, const double* p_freeze, double mbig,
               imex_solve_info* info) {
  if (!c || !b || !x) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d;
  int nc = op == IMEX_OP_MAG ? 3 : 1;
  size_t n = (size_t)nc * d.N;
  int rc = 0;
  if (op != IMEX_OP_MAG && (rc = prep_pressure_operator(c, op, p_freeze, mbig))) return rc;
  HIPCHK(hipMemcpy(c->bk, b, n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->xk, x, n * sizeof(double), hipMemcpyHostToDevice));
  MagOp mo = mag_op(c);
  POp po = p_op(c, op == IMEX_OP_P_JACOBIAN, precond && op != IMEX_OP_MAG);
  rc = run_gmres(c, op == IMEX_OP_MAG ? 0 : 1, mo, po, c->xk, c->bk, 0, info, nullptr, nullptr);
  int r2 = finish(c, rc);
  if (r2 < 0) return r2;
  HIPCHK(hipMemcpy(x, c->xk, n * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

int imex_pressure_newton(imex_ctx* c, const double* b, double* p, imex_solve_info* info) {
  if (!c || !b || !p) return set_err(IMEX_ERR_ARG, "null argument");
  const Dims& d = c->d;
  int rc = prep_pressure_operator(c, IMEX_OP_P_CARRIER, nullptr, 0.0);
  if (rc) return rc;
  HIPCHK(hipMemcpy(c->pb, b, d.N * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->p, p, d.N * sizeof(double), hipMemcpyHostToDevice));
  rc = run_newton(c, c->pb, c->p, info);
  int r2 = finish(c, rc);
  if (r2 < 0) return r2;
  HIPCHK(hipMemcpy(p, c->p, d.N * sizeof(double), hipMemcpyDeviceToHost));
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

}  // extern "C"
