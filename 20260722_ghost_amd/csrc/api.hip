// Host side of libghost_bor: C ABI, problem set-up, device buffers and the
// launch schedule of the BoR-MoM hot path (fill -> batched LU/solve -> far field).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fill.hip"
#include "excite.hip"
#include "lu.hip"
#include "adjoint.hip"
#include "dfill.hip"
#include "../../include/ghost_bor.h"

using namespace ghost;

// ─────────────────────────────────────────────────────────────────────────────
// errors
// ─────────────────────────────────────────────────────────────────────────────

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(GHOST_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                              \
  } while (0)

#define LAUNCHCHK()                                                                  \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return fail(GHOST_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                               \
  } while (0)

extern "C" const char* ghost_last_error(void) { return g_err.c_str(); }

// ─────────────────────────────────────────────────────────────────────────────
// quadrature rules
// ─────────────────────────────────────────────────────────────────────────────

// Gauss-Legendre on [-1, 1] by Newton iteration on P_n, ascending nodes.
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int j = 2; j <= n; ++j) {
        const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      double pn = n == 1 ? z : p1;
      double pm = n == 1 ? 1.0 : p0;
      if (n == 1) { pn = z; pm = 1.0; }
      pp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    // recompute derivative at the converged node
    double p0 = 1.0, p1 = z;
    for (int j = 2; j <= n; ++j) {
      const double p2 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p0) / j;
      p0 = p1;
      p1 = p2;
    }
    pp = n * (z * p1 - p0) / (z * z - 1.0);
    const double wt = 2.0 / ((1.0 - z * z) * pp * pp);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[i] = wt;
    w[n - 1 - i] = wt;
  }
  if (n & 1) x[n / 2] = 0.0;
}

// Graded cells of [0,1]^2 toward s = s' (diag) or toward the corner
// (s, s') = (ca, cb) (bor_solver.py:54-79), depth 4.
static void graded_cells(bool diag, double ca, double cb, double a0, double a1, double b0, double b1, int lev,
                         std::vector<double>& cells) {
  const bool hit = diag ? !(a1 <= b0 || b1 <= a0) : (a0 <= ca && ca <= a1 && b0 <= cb && cb <= b1);
  if (lev >= 4 || !hit) {
    cells.push_back(a0); cells.push_back(a1); cells.push_back(b0); cells.push_back(b1);
    return;
  }
  const double am = 0.5 * (a0 + a1), bm = 0.5 * (b0 + b1);
  graded_cells(diag, ca, cb, a0, am, b0, bm, lev + 1, cells);
  graded_cells(diag, ca, cb, a0, am, bm, b1, lev + 1, cells);
  graded_cells(diag, ca, cb, am, a1, b0, bm, lev + 1, cells);
  graded_cells(diag, ca, cb, am, a1, bm, b1, lev + 1, cells);
}

// Tensor Gauss points on each graded cell: order 4 on the test axis, 5 on the
// source axis (bor_solver.py:85-110).
static void graded_points(bool diag, double ca, double cb, std::vector<double>& s, std::vector<double>& sp,
                          std::vector<double>& w) {
  std::vector<double> cells;
  graded_cells(diag, ca, cb, 0.0, 1.0, 0.0, 1.0, 0, cells);
  std::vector<double> x4, w4, x5, w5;
  gauss_legendre(4, x4, w4);
  gauss_legendre(5, x5, w5);
  for (size_t c = 0; c < cells.size(); c += 4) {
    const double a0 = cells[c], a1 = cells[c + 1], b0 = cells[c + 2], b1 = cells[c + 3];
    const double ha = a1 - a0, hb = b1 - b0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 5; ++j) {
        s.push_back(a0 + 0.5 * (x4[i] + 1.0) * ha);
        sp.push_back(b0 + 0.5 * (x5[j] + 1.0) * hb);
        w.push_back((0.5 * w4[i] * ha) * (0.5 * w5[j] * hb));
      }
  }
}

// Same-surface near kinds: 0 the element with itself, 1 f = e + 1 (shared
// corner s = 1, s' = 0), 2 f = e - 1 (s = 0, s' = 1).
static void cell_points(int kind, std::vector<double>& s, std::vector<double>& sp, std::vector<double>& w) {
  graded_points(kind == 0, kind == 1 ? 1.0 : 0.0, kind == 1 ? 0.0 : 1.0, s, sp, w);
}

// n x n Gauss tensor grid on [0,1]^2, test axis slowest (bor_solver.py:1203-1209).
static void tensor_points(int n, std::vector<double>& s, std::vector<double>& sp, std::vector<double>& w) {
  std::vector<double> x, wx;
  gauss_legendre(n, x, wx);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      s.push_back(0.5 * (x[i] + 1.0));
      sp.push_back(0.5 * (x[j] + 1.0));
      w.push_back(0.25 * wx[i] * wx[j]);
    }
}

// ─────────────────────────────────────────────────────────────────────────────
// state
// ─────────────────────────────────────────────────────────────────────────────

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Problem {
  int Nn = 0, Ne = 0, P = 0, M = 0, nE = 0, nB = 0, ntheta = 0, nrhs = 0, nmodes = 0;
  int zmodes = 0, zoff = 0;      // far-block mode stride and index of mode 0
  double k = 0, cE = 0, cM = 0;
  // surface impedance (IBC-EFIE): per-element Z_s, the coefficient of the
  // impedance operator, and whether any Z_s is nonzero
  std::vector<double2> zs;
  double cI = 0, cMass = 0;
  bool has_zs = false;
  // homogeneous medium of the fill (ghost_bor_set_medium): k = k + j ki and
  // the wave impedance eta; free space by default
  double ki = 0.0;
  double eta_re = 376.730313668, eta_im = 0.0;
  // penetrable body (PMCHWT): this problem's fill keeps its EFIE and bracket
  // operators apart (d_zfar / d_NB and d_zfar2 / d_NB2); the exterior problem
  // names its interior-medium twin in ghost_bor::inner
  bool pen = false;
  int inner = -1;
  bool standalone = false;       // an operator of a multi-region body (ghost_bor_add_operator)
  void* d_zfar2 = nullptr;
  double2* d_NB2 = nullptr;
  double2* d_zs = nullptr;
  std::vector<double> nodes, theta;
  bool axis0 = false, axis1 = false;
  int sys0 = 0;
  size_t dF_off = 0;             // first element of this problem's far-field block
  // host geometry
  std::vector<double> pts;  // 9 arrays of P
  std::vector<double> L, trho, tz;
  int npp = 0;
  std::vector<NearPairInfo> pairs;
  std::vector<int2> batches;
  std::vector<double> tail_u, tail_w;
  // device
  double* d_pts = nullptr;
  double* d_nodes = nullptr;
  double* d_elem = nullptr;      // L, trho, tz
  double* d_tabs = nullptr;      // sin2E | cosE | cxB | sxB | cosB | sinB
  double* d_theta = nullptr;
  double* d_J = nullptr;
  double2* d_E = nullptr;        // [P][ntheta] w rho e^{j k cos(theta) z}
  double* d_cells = nullptr;     // s | sp | w
  double* d_core = nullptr;      // u | w
  double* d_tail = nullptr;      // u | w
  NearPairInfo* d_pairs = nullptr;
  int2* d_batches = nullptr;
  double2* d_NP = nullptr;
  double2* d_NB = nullptr;
  void* d_zfar = nullptr;
  double* d_alt = nullptr;       // adjoint: perturbed pts | nodes | elem
  int cell_off[8] = {0}, cell_cnt[8] = {0};
  size_t tab_off[6] = {0, 0, 0, 0, 0, 0};
};

// Cross-surface operator: EFIE T and rotated-PV brackets B between the test
// bases of one operator surface and the source bases of another in one
// medium (bor_solver.py:1115-1297), node-based [2 Np][2 Nq] far blocks with
// the near pairs added in.
struct Cross {
  int op_p = -1, op_q = -1;
  int Np = 0, Nq = 0, Nep = 0, Neq = 0, M = 0, nmodes = 0, nE = 0, nB = 0, zmodes = 0, zoff = 0;
  double k = 0, ki = 0, eta_re = 0, eta_im = 0;
  double near_factor = 2.0;
  int near_order = 12;
  std::vector<NearPairInfo> pairs;   // near pairs: kind 0 tensor grid, 1 + 2a + b corner<a><b>
  std::vector<int2> batches;
  int npp = 0;
  std::vector<unsigned char> skip;   // [Nep][Neq]
  std::vector<int> plist;            // pair indices grouped by parity class
  int par_off[5] = {0, 0, 0, 0, 0};
  int cell_off[8] = {0}, cell_cnt[8] = {0};
  std::vector<double> cells, tail_u, tail_w;
  size_t tab_off[6] = {0, 0, 0, 0, 0, 0};
  double* d_tabs = nullptr;
  unsigned char* d_skip = nullptr;
  double* d_cells = nullptr;
  double* d_core = nullptr;
  double* d_tail = nullptr;
  NearPairInfo* d_pairs = nullptr;
  int2* d_batches = nullptr;
  int* d_plist = nullptr;
  double2* d_NP = nullptr;
  double2* d_NB = nullptr;
  double2* d_NB2 = nullptr;
  double2* d_zfar = nullptr;
  double2* d_zfar2 = nullptr;
  double2* d_zs = nullptr;           // unit source weights of the brackets
};

// Multi-region system set (bor_solver.py:1805-2041): per mode a system
// Q^T (sum of coefficient x operator blocks) Q, the right-hand sides of the
// surfaces that bound the exterior, and their far field.
struct Region {
  int M = 0, nmodes = 0, n_full = 0, ntheta = 0, nrhs = 0, sys0 = 0;
  std::vector<double> theta;
  struct Block { int src, idx, which; double cre, cim; int row_off, col_off; };
  std::vector<Block> blocks;
  std::vector<std::vector<int>> red;     // per mode: reduced index of each full index (-1: dropped)
  std::vector<std::vector<double>> cf;   // per mode: its coefficient
  struct Ext { int op, off_J, off_M; };
  std::vector<Ext> ext;
  size_t dF_off = 0;
  double2* d_full = nullptr;             // [nmodes][n_full][n_full]
  double2* d_bfull = nullptr;            // [nmodes][n_full][nrhs]
  double2* d_xfull = nullptr;            // [nmodes][n_full][nrhs]
  int* d_ptr = nullptr;
  int* d_mem = nullptr;
  double* d_cf = nullptr;
  std::vector<SysInfo*> d_vsys;          // per exterior surface: one entry per mode
};

struct Launch {  // one batched kernel launch of the solve schedule
  int kind;      // see enum below
  int ops;       // index into the op arrays
  int nops;      // number of systems (grid.y)
  int gx;        // grid.x
  int arg;       // extra (RPT for panels)
  double flops = 0.0;  // algorithmic real flops of the launch
  int tt[3] = {0, 0, 0};  // flattened block count per tile shape (GEMM launches)
  int fam = -1;        // profile family of a GEMM launch (-1: K_GEMM)
};

// kernel families timed by the optional per-launch HIP-event profile
enum {
  K_ZERO, K_NEAR_MODAL, K_NEAR_BLOCKS, K_FAR_FILL, K_ASSEMBLE, K_BESSEL, K_RHS, K_PANEL, K_SWAP,
  K_TRI_INV, K_GEMM, K_GEMM64, K_VEC, K_FARFIELD, K_GEMM_BACK, K_GEMM_SOLVE, K_TRSV, K_COUNT
};
static const char* kKernelNames[K_COUNT] = {"zero", "near_modal", "near_blocks", "far_fill", "assemble",
                                            "bessel", "rhs", "lu_panel", "lu_swap", "tri_inv", "gemm",
                                            "gemm_fp64_residual", "vec", "farfield", "gemm_back",
                                            "gemm_solve", "probe_solve"};

enum { L_PANEL, L_SWAP, L_TRIL, L_TRIU, L_GEMM_AUG, L_GEMM_AUG_W, L_GEMM_Y, L_GEMM_Y_W, L_GEMM_Y_LU,
       L_GEMM_TW, L_GEMM_TLU };   // transposed-operand solves (adjoint): W^T and factor-block^T

struct ghost_bor {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = GHOST_FP64;
  int nb = 64;
  int residual = 1;
  bool built = false;
  bool solved = false;               // the augmented columns hold the solution of the current fill
  std::vector<Problem> probs;
  std::vector<Problem> inner;        // fill-only problems: interior media of penetrable problems and
                                     // the operators of multi-region bodies
  std::vector<Cross> cross;          // cross-surface operators
  std::vector<Region> regions;       // multi-region system sets
  std::vector<SysInfo> sys;          // one entry per (problem, mode) and per dense system
  std::vector<SysInfo> ss;           // the systems factored and solved (nrhs = all their columns)
  int pairing = 1;                   // solve mode -m with mode m's factors
  SysInfo* d_sys = nullptr;
  size_t a_elems = 0, x_elems = 0, a64_elems = 0;
  int piv_elems = 0;
  int max_blocks = 0;
  void* d_A = nullptr;          // augmented systems, FP64 or FP32 complex
  double2* d_A64 = nullptr;     // FP64 copies (residual / refinement)
  double2* d_B64 = nullptr;
  double2* d_X64 = nullptr;
  double2* d_R64 = nullptr;
  float2* d_Y32 = nullptr;
  int* d_piv = nullptr;
  int* d_perm = nullptr;
  int* d_info = nullptr;
  void* d_W = nullptr;          // I - T^{-1} per (system, block, lower/upper)
  double2* d_dF = nullptr;
  double2* d_partial = nullptr;
  double* d_red = nullptr;
  int ff_chunk = 64;
  size_t dF_elems = 0, partial_elems = 0;
  // schedule
  std::vector<PanelOp> h_panel;
  std::vector<SwapOp> h_swap;
  std::vector<TrsmOp> h_trsm;
  std::vector<GemmOp> h_gemm;
  std::vector<VecOp> h_vec;
  PanelOp* d_panel = nullptr;
  SwapOp* d_swap = nullptr;
  TrsmOp* d_trsm = nullptr;
  GemmOp* d_gemm = nullptr;
  VecOp* d_vec = nullptr;
  // LU + back substitution in up to two interleaved parts of the systems (even
  // / odd solve index), run on two streams so that one part's latency-bound
  // panel phases overlap the other part's trailing updates; schedule 2 covers
  // every system in one launch per step (used while the per-launch profile is
  // on, so that each launch is timed over the whole batch)
  int nparts = 1;
  int lu_streams = 2;                // 1: the whole-batch schedule on one stream
  std::vector<Launch> lu_sched[3], back_sched[3];
  std::vector<Launch> fwdY_sched, backY_sched;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // adjoint (FP64 handles): A^T lambda = G through the LU factors, Q = -Lambda X^T
  std::vector<Launch> adjT_fwd, adjT_back;
  int adj_q_ops = -1;
  int adj_refine = 1;                // FP64 refinement sweeps of lambda on FP32 factors (mixed handles)
  bool adj_alloc = false;
  double2* d_Xs = nullptr;           // forward solution, x layout
  double2* d_lam = nullptr;          // G, then z = P lambda (x layout)
  double2* d_lam2 = nullptr;         // lambda
  double2* d_Q = nullptr;            // [-Lambda X^T | Lambda] in the augmented layout (FP64; FP32 on
                                     // mixed handles, where it is half the bytes)
  float2* d_X32 = nullptr;           // mixed adjoint: FP32 copies of X and Lambda for the Q GEMM
  double2* d_rs = nullptr;           // row or column sums, per system at row_off
  double2* d_acoef = nullptr;        // a' per problem column
  double2* d_ffpart = nullptr;       // per-element far-field partial sums of one problem
  double2* d_ffs = nullptr;          // per-element contraction with a'
  double* d_adj_far = nullptr;       // analytic adjoint: far pair partials [Ne * Ne][4]
  double* d_adj_near = nullptr;      // near batch partials [batches][4]
  double* d_adj_ex = nullptr;        // excitation partials [Ne][ntheta][2]
  double* d_adj_grad = nullptr;      // per-node gradient [Nn]
  double* d_adj_gauss = nullptr;     // element Gauss rule per problem: s [4] | weights [4]
  double* d_objw = nullptr;          // objective weights per problem column (captured adjoint)
  bool has_objective = false;
  int* d_rowoff = nullptr;
  int triu_ops[3] = {0, 0, 0}, triu_n[3] = {0, 0, 0};
  int res_gemm = -1, res_gemm_aug = -1;  // residual GEMMs with X in X64 / in the augmented columns
  int res_tt[3] = {0, 0, 0}, res_aug_tt[3] = {0, 0, 0}, adj_q_tt[3] = {0, 0, 0};
  int vec_all = 0, vec_aug = 0;          // VecOps over all systems: x-blocks / augmented RHS -> x-blocks
  int vec_toaug = 0, vec_probe = 0;      // x-blocks -> augmented RHS; the n x 2 probe blocks
  // mixed precision: power-of-two column scales of B (bscale) and of each
  // refinement residual (rscale), per system at c_off; probe scales at 2 si
  double* d_bscale = nullptr;
  double* d_rscale = nullptr;
  double* d_pscale = nullptr;
  float2* d_P32 = nullptr;               // FP32 probe corrections, n x 2 per system at p_off
  TrsvOp* d_trsv = nullptr;              // the probe's two-column solves, one op per system
  int rpt = 1;
  // error probe (adaptive refinement): two fixed random combinations of the columns
  int probe_ops = -1;                // 3 GemmOps per system: BS = B S, XS = X S, RP = BS - A XS
  double2* d_S = nullptr;            // [nrhs_max][2]
  double2* d_BS = nullptr;           // per system n x 2 (offset p_off = sum n * 2)
  double2* d_XS = nullptr;
  std::vector<long long> p_off;
  std::vector<double> probe_hist;    // max over systems of ||A32^-1 RP|| / ||XS|| before each sweep
  std::vector<double> res_norm;      // per system ||AX - B||_F / ||B||_F
  hipEvent_t ev[8];
  unsigned ev_recorded = 0;          // bit i: phase i (fill, solve, far field) has recorded its event pair
  float t_ms[8] = {0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int refine_iters = 0;
  double probe_tol = 1e-5;
  int gemm_variant = -1;             // see gemm_launch; -1: 5 for FP32, 4 for FP64 (measured at n ~ 2400)
  int lu_super = 0;                  // panels per LU super-block (trailing update K = nb x lu_super);
                                     // 0: 8 for FP32 factors, 4 for FP64
  std::vector<DevBuf> allocs;
  // optional per-launch profile
  bool profile = false;
  bool capturing = false;
  std::vector<hipEvent_t> pev;
  struct Rec { int kind; double flops; int e0; };
  std::vector<Rec> recs;
};

static void prof_begin(ghost_bor* h) {
  if (!h->profile) return;
  const size_t need = 2 * (h->recs.size() + 1);
  while (h->pev.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { h->profile = false; return; }
    h->pev.push_back(e);
  }
  hipEventRecord(h->pev[2 * h->recs.size()], h->stream);
}

static void prof_end(ghost_bor* h, int kind, double flops) {
  if (!h->profile) return;
  const int i = (int)h->recs.size();
  hipEventRecord(h->pev[2 * i + 1], h->stream);
  h->recs.push_back(ghost_bor::Rec{kind, flops, 2 * i});
}

template <typename P> static int dalloc(ghost_bor* h, P** out, size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return fail(GHOST_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  h->allocs.push_back({p, bytes});
  *out = (P*)p;
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// C ABI: lifecycle
// ─────────────────────────────────────────────────────────────────────────────

extern "C" int ghost_bor_create(ghost_bor** out, int precision, int nb) {
  if (!out) return fail(GHOST_ERR_ARG, "null output handle");
  if (precision != GHOST_FP64 && precision != GHOST_MIXED)
    return fail(GHOST_ERR_ARG, "precision must be GHOST_FP64 or GHOST_MIXED");
  if (nb <= 0) nb = 64;
  if (nb % 8 != 0 || nb > 64) return fail(GHOST_ERR_ARG, "block size must be a multiple of 8 and <= 64");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  ghost_bor* h = new ghost_bor();
  h->device = dev;
  h->precision = precision;
  h->nb = nb;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) { delete h; return fail(GHOST_ERR_HIP, "stream create: %s", hipGetErrorString(e)); }
  for (int i = 0; i < 8; ++i) {
    e = hipEventCreate(&h->ev[i]);
    if (e != hipSuccess) { delete h; return fail(GHOST_ERR_HIP, "event create: %s", hipGetErrorString(e)); }
  }
  *out = h;
  return 0;
}

extern "C" int ghost_bor_destroy(ghost_bor* h) {
  if (!h) return 0;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->stream2) hipStreamSynchronize(h->stream2);
  for (auto e : h->pev) hipEventDestroy(e);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  for (auto& a : h->allocs) hipFree(a.p);
  for (int i = 0; i < 8; ++i) hipEventDestroy(h->ev[i]);

  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->stream2) hipStreamDestroy(h->stream2);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

extern "C" int ghost_bor_set_residual(ghost_bor* h, int on) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set_residual must precede build");
  h->residual = on ? 1 : 0;
  return 0;
}

extern "C" int ghost_bor_set_pairing(ghost_bor* h, int on) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set mode pairing before build");
  h->pairing = on ? 1 : 0;
  return 0;
}

// LDS of one k_far_fill work-group: the azimuth tables (dynamic) and the
// per-wavefront order stash (static); far_fill_waves() picks the default
// wavefronts per block, one when the tables leave no room for two stashes,
// and 0 when not even one fits.
static size_t far_table_bytes(int rbytes, int nE, int nB) {
  return (size_t)rbytes * (size_t)(2 * nB + 4 * (nB / 2 + 4) + 2 * (nE / 2 + 4) + nE);
}
static size_t far_stash_bytes(int rbytes, int wpb, int NT) { return (size_t)rbytes * wpb * NT * 8 * 64; }
static int far_fill_waves(int rbytes, int NT, int nE, int nB) {
  const size_t lds = 160 * 1024, tab = far_table_bytes(rbytes, nE, nB);
  const int w = far_waves(rbytes, NT);
  if (tab + far_stash_bytes(rbytes, w, NT) <= lds) return w;
  if (rbytes == 8 && tab + far_stash_bytes(rbytes, 1, NT) <= lds) return 1;
  return 0;
}
extern "C" int ghost_far_fill_waves(int precision, int m_max, int n_xi_efie, int n_xi_bracket) {
  return far_fill_waves(precision == GHOST_MIXED ? 4 : 8, (m_max + 2 + 15) / 16, n_xi_efie, n_xi_bracket);
}
template <typename R, int NT>
static void launch_far_fill(bool split, bool ibc, bool ck, int wpb, int grid, size_t tab_bytes, hipStream_t s,
                            const FarArgs& fa) {
  if constexpr (sizeof(R) == 8) {
    if (split && ibc) {   // cross-surface operators of multi-region bodies (any medium)
      if (wpb == 1) hipLaunchKernelGGL((k_far_fill<R, NT, true, true, 1, true>), dim3(grid), dim3(64), tab_bytes, s, fa);
      else {
        constexpr int W = far_waves(sizeof(R), NT);
        hipLaunchKernelGGL((k_far_fill<R, NT, true, true, W, true>), dim3(grid), dim3(64 * W), tab_bytes, s, fa);
      }
      return;
    }
    if (ck) {   // lossy medium (never split: the adjoint refuses media)
      const dim3 b(64 * wpb);
      if (wpb == 1) {
        if (ibc) hipLaunchKernelGGL((k_far_fill<R, NT, false, true, 1, true>), dim3(grid), b, tab_bytes, s, fa);
        else hipLaunchKernelGGL((k_far_fill<R, NT, false, false, 1, true>), dim3(grid), b, tab_bytes, s, fa);
      } else {
        constexpr int W = far_waves(sizeof(R), NT);
        if (ibc) hipLaunchKernelGGL((k_far_fill<R, NT, false, true, W, true>), dim3(grid), b, tab_bytes, s, fa);
        else hipLaunchKernelGGL((k_far_fill<R, NT, false, false, W, true>), dim3(grid), b, tab_bytes, s, fa);
      }
      return;
    }
    if (wpb == 1) {
      if (split) hipLaunchKernelGGL((k_far_fill<R, NT, true, false, 1>), dim3(grid), dim3(64), tab_bytes, s, fa);
      else if (ibc) hipLaunchKernelGGL((k_far_fill<R, NT, false, true, 1>), dim3(grid), dim3(64), tab_bytes, s, fa);
      else hipLaunchKernelGGL((k_far_fill<R, NT, false, false, 1>), dim3(grid), dim3(64), tab_bytes, s, fa);
      return;
    }
  }
  if (split) hipLaunchKernelGGL((k_far_fill<R, NT, true>), dim3(grid), dim3(64 * wpb), tab_bytes, s, fa);
  else if (ibc) hipLaunchKernelGGL((k_far_fill<R, NT, false, true>), dim3(grid), dim3(64 * wpb), tab_bytes, s, fa);
  else hipLaunchKernelGGL((k_far_fill<R, NT, false>), dim3(grid), dim3(64 * wpb), tab_bytes, s, fa);
}

extern "C" int ghost_bor_add_problem(ghost_bor* h, const double* nodes, int Nn, double k, int m_max,
                                     int n_xi_efie, int n_xi_bracket, double c_efie, double c_mfie,
                                     const double* thetas_deg, int n_theta, int flags, int* idx) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "problems must be added before build");
  if (!nodes || Nn < 3) return fail(GHOST_ERR_ARG, "need at least 3 generatrix nodes");
  if (!(k > 0)) return fail(GHOST_ERR_ARG, "wavenumber must be positive (real k only)");
  if (m_max < 0 || m_max > 126) return fail(GHOST_ERR_ARG, "m_max must lie in [0, 126]");
  auto pow2 = [](int v) { return v >= 4 && (v & (v - 1)) == 0; };
  if (!pow2(n_xi_efie) || !pow2(n_xi_bracket))
    return fail(GHOST_ERR_ARG, "xi grid sizes must be powers of two >= 4");
  if (n_xi_efie > 4096 || n_xi_bracket > 4096)
    return fail(GHOST_ERR_ARG, "xi grid sizes above 4096 samples are not supported");
  if (far_fill_waves(h->precision == GHOST_MIXED ? 4 : 8, (m_max + 2 + 15) / 16, n_xi_efie, n_xi_bracket) == 0)
    return fail(GHOST_ERR_ARG,
                "far fill: the %s azimuth tables of n_xi %d / %d with %d order tiles exceed the 160 KiB of LDS",
                h->precision == GHOST_MIXED ? "FP32" : "FP64", n_xi_efie, n_xi_bracket, (m_max + 2 + 15) / 16);
  if (!thetas_deg || n_theta < 1) return fail(GHOST_ERR_ARG, "need at least one aspect angle");
  Problem p;
  p.Nn = Nn;
  p.Ne = Nn - 1;
  p.P = 4 * p.Ne;
  p.k = k;
  p.M = m_max;
  p.nmodes = 2 * m_max + 1;
  p.nE = n_xi_efie;
  p.nB = n_xi_bracket;
  p.cE = c_efie;
  p.cM = c_mfie;
  p.ntheta = n_theta;
  p.nrhs = 2 * n_theta;
  p.nodes.assign(nodes, nodes + 2 * Nn);
  double rmax = 0;
  for (int i = 0; i < Nn; ++i) {
    if (p.nodes[2 * i] < -1e-12) return fail(GHOST_ERR_ARG, "generatrix rho must be >= 0");
    p.nodes[2 * i] = std::max(p.nodes[2 * i], 0.0);
    rmax = std::max(rmax, p.nodes[2 * i]);
  }
  for (int e = 0; e < p.Ne; ++e) {
    const double dr = p.nodes[2 * e + 2] - p.nodes[2 * e], dz = p.nodes[2 * e + 3] - p.nodes[2 * e + 1];
    const double L = std::hypot(dr, dz);
    if (!(L > 0)) return fail(GHOST_ERR_ARG, "generatrix has a zero-length element (%d)", e);
    p.L.push_back(L);
    p.trho.push_back(dr / L);
    p.tz.push_back(dz / L);
  }
  const double tol = 1e-12 * std::max(1.0, rmax);  // bor_kernels.py:79-80
  p.axis0 = p.nodes[0] <= tol;
  p.axis1 = p.nodes[2 * (Nn - 1)] <= tol;
  for (int i = 0; i < n_theta; ++i) p.theta.push_back(thetas_deg[i] * M_PI / 180.0);
  p.sys0 = (int)h->sys.size();
  // systems m = -M..M
  for (int m = -m_max; m <= m_max; ++m) {
    SysInfo s;
    memset(&s, 0, sizeof(s));
    s.prob = (int)h->probs.size();
    s.m = m;
    // active bases (bor_solver.py:611-622); GHOST_UNMASKED keeps all 2 Nn
    const bool all = (flags & GHOST_UNMASKED) != 0;
    const bool t_first = all || (p.axis0 ? (std::abs(m) == 1) : false);
    const bool t_last = all || (p.axis1 ? (std::abs(m) == 1) : false);
    const bool f_first = all || !p.axis0, f_last = all || !p.axis1;
    s.t0 = t_first ? 0 : 1;
    s.nt = Nn - (t_first ? 0 : 1) - (t_last ? 0 : 1);
    s.f0 = f_first ? 0 : 1;
    const int nf = Nn - (f_first ? 0 : 1) - (f_last ? 0 : 1);
    s.n = s.nt + nf;
    s.nrhs = p.nrhs;
    s.ncols = p.nrhs;
    h->sys.push_back(s);
  }
  if (idx) *idx = (int)h->probs.size();
  h->probs.push_back(std::move(p));
  return 0;
}

static int load_generatrix(Problem& p, const double* nodes, int Nn);
static void geometry_tables(const double* nodes, int Ne, std::vector<double>& pts, std::vector<double>& elem);

// New node coordinates for a built problem (same node count, the same ends
// on the axis): the point, node and element tables are recomputed and copied
// into the problem's device buffers, so a built handle, and a graph captured
// on it, serve a family of shapes without reallocating.  The quadrature
// rules chosen at build stay (n_xi, each near pair's tail order), as in the
// adjoint's derivative.  Plain BoR problems only (no impedance, no medium, no
// penetrable or multi-region surfaces).
extern "C" int ghost_bor_set_nodes(ghost_bor* h, int prob, const double* nodes) {
  if (!h || !nodes) return fail(GHOST_ERR_ARG, "null argument");
  if (!h->built) return fail(GHOST_ERR_STATE, "build first");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index %d", prob);
  Problem& p = h->probs[prob];
  if (p.pen || p.standalone || p.has_zs || p.ki != 0.0 || p.sys0 < 0)
    return fail(GHOST_ERR_ARG, "set_nodes covers plain conductor problems only");
  Problem q;
  q.k = p.k;
  int rc;
  if ((rc = load_generatrix(q, nodes, p.Nn))) return rc;
  if (q.axis0 != p.axis0 || q.axis1 != p.axis1)
    return fail(GHOST_ERR_ARG, "the generatrix ends must stay on (or off) the axis as built");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double> pts, elem;
  geometry_tables(q.nodes.data(), q.Ne, pts, elem);
  HIPCHK(hipMemcpy(p.d_pts, pts.data(), pts.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_nodes, q.nodes.data(), q.nodes.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_elem, elem.data(), elem.size() * 8, hipMemcpyHostToDevice));
  p.nodes = q.nodes;
  p.pts = pts;
  p.L = q.L;
  p.trho = q.trho;
  p.tz = q.tz;
  h->solved = false;
  return 0;
}

// Surface impedance of one problem (IBC-EFIE, bor_solver.py:850-866,
// 929-934): zs complex per element; Z += c_mass (1/2) M_Zs + c_ibc K'_Zs.
extern "C" int ghost_bor_set_impedance(ghost_bor* h, int prob, const double* zs, double c_ibc, double c_mass) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set the surface impedance before build");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index %d", prob);
  if (!zs) return fail(GHOST_ERR_ARG, "null impedance array");
  if (!std::isfinite(c_ibc) || !std::isfinite(c_mass)) return fail(GHOST_ERR_ARG, "impedance coefficients must be finite");
  Problem& p = h->probs[prob];
  if (p.pen) return fail(GHOST_ERR_ARG, "a penetrable body takes no surface impedance");
  std::vector<double2> z(p.Ne);
  bool any = false;
  for (int e = 0; e < p.Ne; ++e) {
    z[e] = make_double2(zs[2 * e], zs[2 * e + 1]);
    if (!std::isfinite(z[e].x) || !std::isfinite(z[e].y))
      return fail(GHOST_ERR_ARG, "surface impedance of element %d is not finite", e);
    any = any || z[e].x != 0.0 || z[e].y != 0.0;
  }
  if (any && c_ibc != 0.0 && p.cM != 0.0)
    return fail(GHOST_ERR_ARG, "a surface impedance needs c_mfie = 0 (the impedance operator takes the MFIE's place)");
  p.zs = z;
  p.has_zs = any;
  p.cI = any ? c_ibc : 0.0;
  p.cMass = any ? c_mass : 0.0;
  return 0;
}

// Homogeneous medium of one problem's fill (bor_solver.py:175-199 with
// medium=(eps, mu)): k = k_real + j k_im (k_im <= 0 for a lossy medium) and
// the wave impedance eta.  FP64 handles only.
extern "C" int ghost_bor_set_medium(ghost_bor* h, int prob, double k_im, double eta_re, double eta_im) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set the medium before build");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index %d", prob);
  if (!std::isfinite(k_im) || !std::isfinite(eta_re) || !std::isfinite(eta_im))
    return fail(GHOST_ERR_ARG, "medium parameters must be finite");
  if (k_im > 0.0) return fail(GHOST_ERR_ARG, "Im k must be <= 0 (a causal, passive medium)");
  if (k_im != 0.0 && h->precision != GHOST_FP64)
    return fail(GHOST_ERR_ARG, "a lossy medium (complex k) needs an FP64 handle");
  Problem& p = h->probs[prob];
  if (p.pen) return fail(GHOST_ERR_ARG, "the exterior of a penetrable body is free space");
  p.ki = k_im;
  p.eta_re = eta_re;
  p.eta_im = eta_im;
  return 0;
}

// A caller-filled dense system (n x n, n_rhs columns) solved in the same batch
// as the mode systems; it belongs to no problem (no fill, no far field).
// Turn problem prob into a homogeneous penetrable body (PMCHWT,
// bor_solver.py:1019-1094): the problem's own medium is the exterior (free
// space), the interior medium has k = k_re + j k_im and wave impedance eta;
// each mode system becomes the 2 x 2 block system in (J, M / eta0) of twice
// the single-surface size.  FP64 handles only.
extern "C" int ghost_bor_set_interior(ghost_bor* h, int prob, double k_re, double k_im, double eta_re,
                                      double eta_im, int n_xi_efie, int n_xi_bracket) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set the interior medium before build");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index %d", prob);
  if (h->precision != GHOST_FP64) return fail(GHOST_ERR_ARG, "a penetrable body needs an FP64 handle");
  if (!std::isfinite(k_re) || !std::isfinite(k_im) || !std::isfinite(eta_re) || !std::isfinite(eta_im))
    return fail(GHOST_ERR_ARG, "interior medium parameters must be finite");
  if (!(k_re > 0.0) || k_im > 0.0) return fail(GHOST_ERR_ARG, "interior k needs Re k > 0 and Im k <= 0");
  if (eta_re == 0.0 && eta_im == 0.0) return fail(GHOST_ERR_ARG, "interior wave impedance must be nonzero");
  auto pow2 = [](int v) { return v >= 4 && (v & (v - 1)) == 0; };
  if (!pow2(n_xi_efie) || !pow2(n_xi_bracket) || n_xi_efie > 4096 || n_xi_bracket > 4096)
    return fail(GHOST_ERR_ARG, "interior xi grid sizes must be powers of two in [4, 4096]");
  Problem& p = h->probs[prob];
  if (p.pen) return fail(GHOST_ERR_STATE, "problem %d already has an interior medium", prob);
  if (p.has_zs) return fail(GHOST_ERR_ARG, "a penetrable body takes no surface impedance");
  if (p.cE != 1.0 || p.cM != 0.0) return fail(GHOST_ERR_ARG, "add a penetrable body with c_efie = 1, c_mfie = 0");
  if (p.ki != 0.0 || p.eta_re != 376.730313668 || p.eta_im != 0.0)
    return fail(GHOST_ERR_ARG, "the exterior of a penetrable body is free space");
  if (far_fill_waves(8, (p.M + 2 + 15) / 16, n_xi_efie, n_xi_bracket) == 0)
    return fail(GHOST_ERR_ARG, "far fill: interior azimuth tables of n_xi %d / %d exceed the 160 KiB of LDS",
                n_xi_efie, n_xi_bracket);
  for (int m = -p.M; m <= p.M; ++m)
    if (2 * h->sys[p.sys0 + m + p.M].n > 8192)
      return fail(GHOST_ERR_ARG, "PMCHWT system size %d exceeds the panel limit 8192", 2 * h->sys[p.sys0 + m + p.M].n);
  // the bracket operators of both media at unit weight, no mass term
  p.zs.assign(p.Ne, make_double2(1.0, 0.0));
  p.has_zs = true;
  p.cI = 1.0;
  p.cMass = 0.0;
  p.pen = true;
  Problem q;
  q.Nn = p.Nn; q.Ne = p.Ne; q.P = p.P; q.M = p.M; q.nmodes = p.nmodes;
  q.nE = n_xi_efie; q.nB = n_xi_bracket;
  q.k = k_re; q.ki = k_im; q.eta_re = eta_re; q.eta_im = eta_im;
  q.cE = 1.0; q.cM = 0.0;
  q.zs = p.zs; q.has_zs = true; q.cI = 1.0; q.cMass = 0.0; q.pen = true;
  q.ntheta = 1; q.nrhs = 0; q.theta.assign(1, 0.0);
  q.nodes = p.nodes; q.L = p.L; q.trho = p.trho; q.tz = p.tz;
  q.axis0 = p.axis0; q.axis1 = p.axis1;
  q.sys0 = -1;
  p.inner = (int)h->inner.size();
  h->inner.push_back(std::move(q));
  for (int m = -p.M; m <= p.M; ++m) {
    SysInfo& sy = h->sys[p.sys0 + m + p.M];
    sy.pen = sy.n;
    sy.n *= 2;
  }
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// C ABI: multi-region bodies (bor_solver.py:1100-1297, 1805-2172)
// ─────────────────────────────────────────────────────────────────────────────

static int pow2_xi(int v) { return v >= 4 && v <= 4096 && (v & (v - 1)) == 0; }

// Generatrix of a fill-only problem: nodes (rho >= 0, snapped), element
// lengths and tangents, axis ends.
static int load_generatrix(Problem& p, const double* nodes, int Nn) {
  if (!nodes || Nn < 3) return fail(GHOST_ERR_ARG, "need at least 3 generatrix nodes");
  p.Nn = Nn;
  p.Ne = Nn - 1;
  p.P = 4 * p.Ne;
  p.nodes.assign(nodes, nodes + 2 * Nn);
  double rmax = 0;
  for (int i = 0; i < Nn; ++i) {
    if (!std::isfinite(p.nodes[2 * i]) || !std::isfinite(p.nodes[2 * i + 1]))
      return fail(GHOST_ERR_ARG, "generatrix node %d is not finite", i);
    if (p.nodes[2 * i] < -1e-12) return fail(GHOST_ERR_ARG, "generatrix rho must be >= 0");
    p.nodes[2 * i] = std::max(p.nodes[2 * i], 0.0);
    rmax = std::max(rmax, p.nodes[2 * i]);
  }
  p.L.clear(); p.trho.clear(); p.tz.clear();
  for (int e = 0; e < p.Ne; ++e) {
    const double dr = p.nodes[2 * e + 2] - p.nodes[2 * e], dz = p.nodes[2 * e + 3] - p.nodes[2 * e + 1];
    const double L = std::hypot(dr, dz);
    if (!(L > 0)) return fail(GHOST_ERR_ARG, "generatrix has a zero-length element (%d)", e);
    p.L.push_back(L);
    p.trho.push_back(dr / L);
    p.tz.push_back(dz / L);
  }
  const double tol = 1e-12 * std::max(1.0, rmax);
  p.axis0 = p.nodes[0] <= tol;
  p.axis1 = p.nodes[2 * (Nn - 1)] <= tol;
  return 0;
}

// One surface in one homogeneous medium, filled as operators only (no
// systems of its own): the EFIE operator T, and with brackets != 0 also the
// unit-weight rotated-PV brackets B that the PMCHWT coupling blocks take
// (bor_solver.py:554-575).  thetas: aspects of the plane-wave excitation and
// far field when the surface bounds the exterior (n_theta may be 0).
extern "C" int ghost_bor_add_operator(ghost_bor* h, const double* nodes, int Nn, double k_re, double k_im,
                                      double eta_re, double eta_im, int m_max, int n_xi_efie, int n_xi_bracket,
                                      int brackets, const double* thetas_deg, int n_theta, int* op) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "add operators before build");
  if (h->precision != GHOST_FP64) return fail(GHOST_ERR_ARG, "multi-region bodies need an FP64 handle");
  if (!std::isfinite(k_re) || !std::isfinite(k_im) || !std::isfinite(eta_re) || !std::isfinite(eta_im))
    return fail(GHOST_ERR_ARG, "medium parameters must be finite");
  if (!(k_re > 0.0) || k_im > 0.0) return fail(GHOST_ERR_ARG, "k needs Re k > 0 and Im k <= 0");
  if (eta_re == 0.0 && eta_im == 0.0) return fail(GHOST_ERR_ARG, "wave impedance must be nonzero");
  if (m_max < 0 || m_max > 126) return fail(GHOST_ERR_ARG, "m_max must lie in [0, 126]");
  if (!pow2_xi(n_xi_efie) || !pow2_xi(n_xi_bracket))
    return fail(GHOST_ERR_ARG, "xi grid sizes must be powers of two in [4, 4096]");
  if (far_fill_waves(8, (m_max + 2 + 15) / 16, n_xi_efie, n_xi_bracket) == 0)
    return fail(GHOST_ERR_ARG, "far fill: azimuth tables of n_xi %d / %d exceed the 160 KiB of LDS", n_xi_efie,
                n_xi_bracket);
  if (n_theta < 0 || (n_theta > 0 && !thetas_deg)) return fail(GHOST_ERR_ARG, "bad aspect list");
  Problem q;
  int rc = load_generatrix(q, nodes, Nn);
  if (rc) return rc;
  q.k = k_re; q.ki = k_im; q.eta_re = eta_re; q.eta_im = eta_im;
  q.M = m_max; q.nmodes = 2 * m_max + 1;
  q.nE = n_xi_efie; q.nB = n_xi_bracket;
  q.cE = 1.0; q.cM = 0.0;
  if (brackets) {
    q.zs.assign(q.Ne, make_double2(1.0, 0.0));
    q.has_zs = true; q.cI = 1.0; q.cMass = 0.0; q.pen = true;
  }
  q.ntheta = n_theta; q.nrhs = 2 * n_theta;
  for (int i = 0; i < n_theta; ++i) q.theta.push_back(thetas_deg[i] * M_PI / 180.0);
  q.standalone = true;
  q.sys0 = -1;
  if (op) *op = (int)h->inner.size();
  h->inner.push_back(std::move(q));
  return 0;
}

static double gauss_rho_max(const Problem& p) {
  std::vector<double> x4, w4;
  gauss_legendre(4, x4, w4);
  double r = 0.0;
  for (int e = 0; e < p.Ne; ++e)
    for (int g = 0; g < 4; ++g) {
      const double s = 0.5 * (x4[g] + 1.0);
      r = std::max(r, p.nodes[2 * e] + s * (p.nodes[2 * e + 2] - p.nodes[2 * e]));
    }
  return r;
}

static double seg_point(const double* c, const double* a, const double* b) {
  const double abx = b[0] - a[0], aby = b[1] - a[1];
  double t = ((c[0] - a[0]) * abx + (c[1] - a[1]) * aby) / std::max(abx * abx + aby * aby, 1e-300);
  t = std::min(1.0, std::max(0.0, t));
  return std::hypot(c[0] - (a[0] + t * abx), c[1] - (a[1] + t * aby));
}

// The cross-surface operator between the test bases of operator op_test and
// the source bases of op_src (both in one medium).  Element pairs sharing an
// endpoint (a junction) are integrated on corner-graded cells, pairs closer
// than near_factor x the longer element on a near_order^2 tensor grid, both
// with the near kernels; the rest on the far tables, whose n_xi follow the
// reference's rule for the two point sets (bor_kernels.py:149-152, 489-492).
extern "C" int ghost_bor_add_cross(ghost_bor* h, int op_test, int op_src, double near_factor, int near_order,
                                   int* idx) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "add cross operators before build");
  const int ni = (int)h->inner.size();
  if (op_test < 0 || op_test >= ni || op_src < 0 || op_src >= ni || !h->inner[op_test].standalone ||
      !h->inner[op_src].standalone)
    return fail(GHOST_ERR_ARG, "bad operator index");
  if (op_test == op_src) return fail(GHOST_ERR_ARG, "a cross operator joins two different surfaces");
  if (!(near_factor > 0.0) || near_order < 1 || near_order > 64)
    return fail(GHOST_ERR_ARG, "near_factor must be > 0 and near_order in [1, 64]");
  const Problem& a = h->inner[op_test];
  const Problem& b = h->inner[op_src];
  if (a.k != b.k || a.ki != b.ki || a.eta_re != b.eta_re || a.eta_im != b.eta_im)
    return fail(GHOST_ERR_ARG, "a cross operator needs both surfaces in the same medium");
  if (a.M != b.M) return fail(GHOST_ERR_ARG, "both operators need the same m_max");
  Cross c;
  c.op_p = op_test; c.op_q = op_src;
  c.Np = a.Nn; c.Nq = b.Nn; c.Nep = a.Ne; c.Neq = b.Ne; c.M = a.M; c.nmodes = a.nmodes;
  c.k = a.k; c.ki = a.ki; c.eta_re = a.eta_re; c.eta_im = a.eta_im;
  c.near_factor = near_factor; c.near_order = near_order;
  const double osc = 2.0 * std::hypot(a.k, a.ki) * std::sqrt(gauss_rho_max(a) * gauss_rho_max(b));
  c.nE = (int)std::pow(2.0, std::ceil(std::log2(std::max({64.0, 4.0 * (c.M + 2), 6.0 * (osc + 4.0)}))));
  c.nB = (int)std::pow(2.0, std::ceil(std::log2(std::max({128.0, 6.0 * (c.M + 2), 8.0 * (osc + 4.0)}))));
  if (c.nE > 4096 || c.nB > 4096 || far_fill_waves(8, (c.M + 2 + 15) / 16, c.nE, c.nB) == 0)
    return fail(GHOST_ERR_ARG, "cross operator: azimuth grids %d / %d beyond the far fill's limits", c.nE, c.nB);
  // near classification (bor_solver.py:1131-1164)
  auto span = [](const Problem& p) {
    double lo = p.nodes[0], hi = p.nodes[0];
    for (double v : p.nodes) { lo = std::min(lo, v); hi = std::max(hi, v); }
    return hi - lo;
  };
  const double tol = 1e-9 * std::max({span(a), span(b), 1e-9});
  c.skip.assign((size_t)c.Nep * c.Neq, 0);
  for (int e = 0; e < c.Nep; ++e)
    for (int f = 0; f < c.Neq; ++f) {
      const double* pe[2] = {&a.nodes[2 * e], &a.nodes[2 * e + 2]};
      const double* qf[2] = {&b.nodes[2 * f], &b.nodes[2 * f + 2]};
      int nshared = 0, sa = 0, sb = 0;
      for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v)
          if (std::hypot(pe[u][0] - qf[v][0], pe[u][1] - qf[v][1]) <= tol) { ++nshared; sa = u; sb = v; }
      if (nshared > 1) return fail(GHOST_ERR_ARG, "surfaces share a whole element (%d, %d)", e, f);
      const double d = std::min({seg_point(pe[0], qf[0], qf[1]), seg_point(pe[1], qf[0], qf[1]),
                                 seg_point(qf[0], pe[0], pe[1]), seg_point(qf[1], pe[0], pe[1])});
      int kind = -1;
      if (nshared == 1) kind = 1 + 2 * sa + sb;
      else if (d <= tol) return fail(GHOST_ERR_ARG, "surfaces cross or touch away from a junction (%d, %d)", e, f);
      else if (d < near_factor * std::max(a.L[e], b.L[f])) kind = 0;
      if (kind < 0) continue;
      NearPairInfo pi;
      pi.e = e; pi.f = f; pi.kind = kind; pi.pp_off = 0; pi.n_tail = 0; pi.tail_off = 0;
      c.pairs.push_back(pi);
      c.skip[(size_t)e * c.Neq + f] = 1;
    }
  if (idx) *idx = (int)h->cross.size();
  h->cross.push_back(std::move(c));
  return 0;
}

// A multi-region system set: one system per mode -m_max..m_max over n_full
// unknowns before the junction map, thetas the aspects.
extern "C" int ghost_bor_add_region(ghost_bor* h, int m_max, int n_full, const double* thetas_deg, int n_theta,
                                    int* set) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "add region systems before build");
  if (h->precision != GHOST_FP64) return fail(GHOST_ERR_ARG, "multi-region bodies need an FP64 handle");
  if (m_max < 0 || m_max > 126) return fail(GHOST_ERR_ARG, "m_max must lie in [0, 126]");
  if (n_full < 1) return fail(GHOST_ERR_ARG, "n_full must be positive");
  if (!thetas_deg || n_theta < 1) return fail(GHOST_ERR_ARG, "need at least one aspect angle");
  Region r;
  r.M = m_max; r.nmodes = 2 * m_max + 1; r.n_full = n_full; r.ntheta = n_theta; r.nrhs = 2 * n_theta;
  for (int i = 0; i < n_theta; ++i) r.theta.push_back(thetas_deg[i] * M_PI / 180.0);
  r.red.assign(r.nmodes, std::vector<int>());
  r.cf.assign(r.nmodes, std::vector<double>());
  r.sys0 = (int)h->sys.size();
  for (int m = -m_max; m <= m_max; ++m) {
    SysInfo sy;
    memset(&sy, 0, sizeof(sy));
    sy.prob = -1;
    sy.m = m;
    sy.nrhs = r.nrhs;
    sy.ncols = r.nrhs;
    h->sys.push_back(sy);
  }
  if (set) *set = (int)h->regions.size();
  h->regions.push_back(std::move(r));
  return 0;
}

// Add coef x block to every mode's full system at (row_off, col_off): src 0
// operator idx (its own surface), 1 cross operator idx; which 0 the EFIE
// operator T, 1 the coupling operator P (brackets, columns remapped).
extern "C" int ghost_bor_region_block(ghost_bor* h, int set, int src, int idx, int which, double c_re, double c_im,
                                      int row_off, int col_off) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "add blocks before build");
  if (set < 0 || set >= (int)h->regions.size()) return fail(GHOST_ERR_ARG, "bad region set");
  Region& r = h->regions[set];
  if (!std::isfinite(c_re) || !std::isfinite(c_im)) return fail(GHOST_ERR_ARG, "block coefficient must be finite");
  if (which != 0 && which != 1) return fail(GHOST_ERR_ARG, "which is 0 (T) or 1 (P)");
  int Np, Nq, M;
  if (src == 0) {
    if (idx < 0 || idx >= (int)h->inner.size() || !h->inner[idx].standalone)
      return fail(GHOST_ERR_ARG, "bad operator index");
    const Problem& o = h->inner[idx];
    if (which == 1 && !o.pen) return fail(GHOST_ERR_ARG, "operator %d was added without brackets", idx);
    Np = Nq = o.Nn;
    M = o.M;
  } else if (src == 1) {
    if (idx < 0 || idx >= (int)h->cross.size()) return fail(GHOST_ERR_ARG, "bad cross operator index");
    Np = h->cross[idx].Np;
    Nq = h->cross[idx].Nq;
    M = h->cross[idx].M;
  } else {
    return fail(GHOST_ERR_ARG, "src is 0 (operator) or 1 (cross operator)");
  }
  if (M != r.M) return fail(GHOST_ERR_ARG, "block m_max %d differs from the set's %d", M, r.M);
  if (row_off < 0 || col_off < 0 || row_off + 2 * Np > r.n_full || col_off + 2 * Nq > r.n_full)
    return fail(GHOST_ERR_ARG, "block outside the full system");
  r.blocks.push_back(Region::Block{src, idx, which, c_re, c_im, row_off, col_off});
  return 0;
}

// Junction map of mode m: full index i goes to reduced index red_of_full[i]
// (-1: dropped) with coefficient coef[i] (x_full = Q x_red; A_red = Q^T A Q).
extern "C" int ghost_bor_region_map(ghost_bor* h, int set, int m, int n_red, const int* red_of_full,
                                    const double* coef) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set maps before build");
  if (set < 0 || set >= (int)h->regions.size()) return fail(GHOST_ERR_ARG, "bad region set");
  Region& r = h->regions[set];
  if (m < -r.M || m > r.M) return fail(GHOST_ERR_ARG, "mode outside range");
  if (n_red < 1 || n_red > 8192) return fail(GHOST_ERR_ARG, "reduced size %d outside [1, 8192]", n_red);
  if (!red_of_full || !coef) return fail(GHOST_ERR_ARG, "null map");
  std::vector<int> used(n_red, 0);
  for (int i = 0; i < r.n_full; ++i) {
    if (red_of_full[i] < -1 || red_of_full[i] >= n_red) return fail(GHOST_ERR_ARG, "map entry %d out of range", i);
    if (red_of_full[i] >= 0) {
      if (!std::isfinite(coef[i])) return fail(GHOST_ERR_ARG, "map coefficient %d is not finite", i);
      used[red_of_full[i]] = 1;
    }
  }
  for (int u = 0; u < n_red; ++u)
    if (!used[u]) return fail(GHOST_ERR_ARG, "reduced unknown %d of mode %d has no full index", u, m);
  r.red[m + r.M].assign(red_of_full, red_of_full + r.n_full);
  r.cf[m + r.M].assign(coef, coef + r.n_full);
  SysInfo& sy = h->sys[r.sys0 + m + r.M];
  sy.n = n_red;
  sy.nt = n_red;
  return 0;
}

// A surface bounding the exterior: operator op (free space, the set's
// aspects) carries J in full rows [off_J, off_J + 2 Nn) and, off_M >= 0, M /
// eta0 right after them (off_M = off_J + 2 Nn).  Its rows take the plane
// wave's <W, E_inc> (and eta0 <W, H_inc>), and its currents radiate.
extern "C" int ghost_bor_region_exterior(ghost_bor* h, int set, int op, int off_J, int off_M) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set exterior surfaces before build");
  if (set < 0 || set >= (int)h->regions.size()) return fail(GHOST_ERR_ARG, "bad region set");
  Region& r = h->regions[set];
  if (op < 0 || op >= (int)h->inner.size() || !h->inner[op].standalone) return fail(GHOST_ERR_ARG, "bad operator");
  const Problem& o = h->inner[op];
  if (o.M != r.M) return fail(GHOST_ERR_ARG, "operator m_max differs from the set's");
  if (o.ntheta != r.ntheta) return fail(GHOST_ERR_ARG, "an exterior operator needs the set's aspects");
  for (int i = 0; i < r.ntheta; ++i)
    if (std::fabs(o.theta[i] - r.theta[i]) > 1e-15) return fail(GHOST_ERR_ARG, "an exterior operator needs the set's aspects");
  if (o.ki != 0.0) return fail(GHOST_ERR_ARG, "an exterior operator is in free space");
  if (off_J < 0 || off_J + 2 * o.Nn > r.n_full) return fail(GHOST_ERR_ARG, "off_J outside the full system");
  if (off_M >= 0 && (off_M != off_J + 2 * o.Nn || off_M + 2 * o.Nn > r.n_full))
    return fail(GHOST_ERR_ARG, "M rows must follow the surface's J rows");
  r.ext.push_back(Region::Ext{op, off_J, off_M});
  return 0;
}

extern "C" int ghost_bor_add_dense(ghost_bor* h, int n, int nrhs, int* sys_idx) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "add systems before build");
  if (n <= 0 || nrhs <= 0) return fail(GHOST_ERR_ARG, "n and nrhs must be positive");
  if (n > 8192) return fail(GHOST_ERR_ARG, "system size %d exceeds the panel limit 8192", n);
  SysInfo s;
  memset(&s, 0, sizeof(s));
  s.prob = -1;
  s.m = 0;
  s.n = n;
  s.nt = n;
  s.nrhs = nrhs;
  s.ncols = nrhs;
  if (sys_idx) *sys_idx = (int)h->sys.size();
  h->sys.push_back(s);
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// build: host tables, device buffers, solve schedule
// ─────────────────────────────────────────────────────────────────────────────

// Gauss points, order 4 (bor_kernels.py:103-125), of a generatrix with Ne
// elements: pts = rho | z | t_rho | t_z | w (x element length) | T0 | T1 |
// D0 | D1 (d(rho T)/dt), each [4 Ne]; elem = L | t_rho | t_z, each [Ne].
static void geometry_tables(const double* nodes, int Ne, std::vector<double>& pts, std::vector<double>& elem) {
  std::vector<double> x4, w4;
  gauss_legendre(4, x4, w4);
  const int P = 4 * Ne;
  pts.assign(9 * (size_t)P, 0.0);
  elem.assign(3 * (size_t)Ne, 0.0);
  double* rho = &pts[0 * P]; double* z = &pts[1 * P]; double* tr = &pts[2 * P];
  double* tzp = &pts[3 * P]; double* w = &pts[4 * P]; double* T0 = &pts[5 * P];
  double* T1 = &pts[6 * P]; double* D0 = &pts[7 * P]; double* D1 = &pts[8 * P];
  for (int e = 0; e < Ne; ++e) {
    const double r0 = nodes[2 * e], z0 = nodes[2 * e + 1];
    const double dr = nodes[2 * e + 2] - r0, dz = nodes[2 * e + 3] - z0;
    const double L = std::hypot(dr, dz);
    elem[e] = L;
    elem[Ne + e] = dr / L;
    elem[2 * Ne + e] = dz / L;
    for (int g = 0; g < 4; ++g) {
      const int i = 4 * e + g;
      const double s = 0.5 * (x4[g] + 1.0);
      rho[i] = r0 + s * dr;
      z[i] = z0 + s * dz;
      tr[i] = dr / L;
      tzp[i] = dz / L;
      w[i] = 0.5 * w4[g] * L;
      T0[i] = 1.0 - s;
      T1[i] = s;
      D0[i] = (dr * (1.0 - s) - rho[i]) / L;
      D1[i] = (dr * s + rho[i]) / L;
    }
  }
}

static int prepare_problem(ghost_bor* h, Problem& p) {
  const int P = p.P, Ne = p.Ne;
  std::vector<double> elem;
  geometry_tables(p.nodes.data(), Ne, p.pts, elem);
  // ── near pairs: cells, tail orders (bor_kernels.py:238-240) ──
  std::vector<double> cs[3], csp[3], cw[3];
  for (int kd = 0; kd < 3; ++kd) cell_points(kd, cs[kd], csp[kd], cw[kd]);
  std::vector<double> cell_all;
  int off = 0;
  for (int kd = 0; kd < 3; ++kd) {
    p.cell_off[kd] = off;
    p.cell_cnt[kd] = (int)cs[kd].size();
    off += p.cell_cnt[kd];
  }
  std::map<int, int> tail_of;   // n_tail -> offset
  p.pairs.assign(3 * (size_t)Ne, NearPairInfo{-1, -1, 0, 0, 0, 0});
  int pp = 0;
  for (int e = 0; e < Ne; ++e)
    for (int d = 0; d < 3; ++d) {
      const int f = e + d - 1;
      if (f < 0 || f >= Ne) continue;
      const int kind = (f == e) ? 0 : (f == e + 1 ? 1 : 2);
      double osc = 0.0;
      const double r0e = p.nodes[2 * e], dre = p.nodes[2 * e + 2] - r0e;
      const double r0f = p.nodes[2 * f], drf = p.nodes[2 * f + 2] - r0f;
      for (size_t q = 0; q < cs[kind].size(); ++q) {
        const double rp = r0e + cs[kind][q] * dre, rq = r0f + csp[kind][q] * drf;
        const double rr4 = 4.0 * rp * rq;
        const double a = std::sqrt(std::max(rr4, 1e-300));
        osc = std::max(osc, std::hypot(p.k, p.ki) * a);
      }
      const double o = osc / M_PI + (p.M + 2);
      const int nt = (int)std::min(1024.0, std::max(64.0, std::ceil(4.0 * o)));
      NearPairInfo pi;
      pi.e = e;
      pi.f = f;
      pi.kind = kind;
      pi.pp_off = pp;
      pi.n_tail = nt;
      if (!tail_of.count(nt)) {
        tail_of[nt] = (int)p.tail_u.size();
        std::vector<double> xt, wt;
        gauss_legendre(nt, xt, wt);
        for (int i = 0; i < nt; ++i) { p.tail_u.push_back(0.5 * (xt[i] + 1.0)); p.tail_w.push_back(0.5 * wt[i]); }
      }
      pi.tail_off = tail_of[nt];
      p.pairs[3 * e + d] = pi;
      for (int b = 0; b < p.cell_cnt[kind]; b += 32) p.batches.push_back(make_int2(3 * e + d, b));
      pp += p.cell_cnt[kind];
    }
  p.npp = pp;

  // ── device buffers ──
  int rc;
  if ((rc = dalloc(h, &p.d_pts, p.pts.size() * 8))) return rc;
  if ((rc = dalloc(h, &p.d_nodes, p.nodes.size() * 8))) return rc;
  if ((rc = dalloc(h, &p.d_elem, 3 * (size_t)Ne * 8))) return rc;
  // far tables with exact values at 0, pi/2, pi, 3pi/2
  std::vector<double> tabs;
  auto cos_exact = [](long long i, int N) {
    i %= N;
    if (i == 0) return 1.0;
    if (2 * i == N) return -1.0;
    if (4 * i == N || 4 * i == 3 * (long long)N) return 0.0;
    return std::cos(2.0 * M_PI * (double)i / N);
  };
  auto sin_exact = [](long long i, int N) {
    i %= N;
    if (i == 0 || 2 * i == N) return 0.0;
    if (4 * i == N) return 1.0;
    if (4 * i == 3 * (long long)N) return -1.0;
    return std::sin(2.0 * M_PI * (double)i / N);
  };
  p.tab_off[0] = tabs.size();
  for (int j = 0; j <= p.nE / 2; ++j) { const double s = std::sin(M_PI * j / p.nE); tabs.push_back(j == p.nE / 2 ? 1.0 : s * s); }
  p.tab_off[1] = tabs.size();
  for (int i = 0; i < p.nE; ++i) tabs.push_back(cos_exact(i, p.nE));
  p.tab_off[2] = tabs.size();
  for (int j = 0; j <= p.nB / 2; ++j) tabs.push_back(cos_exact(j, p.nB));
  p.tab_off[3] = tabs.size();
  for (int j = 0; j <= p.nB / 2; ++j) tabs.push_back(sin_exact(j, p.nB));
  p.tab_off[4] = tabs.size();
  for (int i = 0; i < p.nB; ++i) tabs.push_back(cos_exact(i, p.nB));
  p.tab_off[5] = tabs.size();
  for (int i = 0; i < p.nB; ++i) tabs.push_back(sin_exact(i, p.nB));
  if ((rc = dalloc(h, &p.d_tabs, tabs.size() * 8))) return rc;
  if ((rc = dalloc(h, &p.d_theta, p.theta.size() * 8))) return rc;
  if ((rc = dalloc(h, &p.d_J, (size_t)(p.M + 2) * P * p.ntheta * 8))) return rc;
  if ((rc = dalloc(h, &p.d_E, (size_t)P * p.ntheta * sizeof(double2)))) return rc;
  std::vector<double> cells;
  for (int kd = 0; kd < 3; ++kd) cells.insert(cells.end(), cs[kd].begin(), cs[kd].end());
  for (int kd = 0; kd < 3; ++kd) cells.insert(cells.end(), csp[kd].begin(), csp[kd].end());
  for (int kd = 0; kd < 3; ++kd) cells.insert(cells.end(), cw[kd].begin(), cw[kd].end());
  if ((rc = dalloc(h, &p.d_cells, cells.size() * 8))) return rc;
  std::vector<double> x48, w48, core;
  gauss_legendre(48, x48, w48);
  for (int i = 0; i < 48; ++i) core.push_back(0.5 * (x48[i] + 1.0));
  for (int i = 0; i < 48; ++i) core.push_back(0.5 * w48[i]);
  if ((rc = dalloc(h, &p.d_core, core.size() * 8))) return rc;
  std::vector<double> tail(p.tail_u);
  tail.insert(tail.end(), p.tail_w.begin(), p.tail_w.end());
  if ((rc = dalloc(h, &p.d_tail, tail.size() * 8))) return rc;
  if ((rc = dalloc(h, &p.d_pairs, p.pairs.size() * sizeof(NearPairInfo)))) return rc;
  if ((rc = dalloc(h, &p.d_batches, p.batches.size() * sizeof(int2)))) return rc;
  if ((rc = dalloc(h, &p.d_NP, (size_t)p.npp * (p.M + 2) * 5 * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &p.d_NB, (size_t)3 * Ne * 16 * p.nmodes * sizeof(double2)))) return rc;
  // far blocks [2Nn][2Nn][zmodes]: modes -M..M, or 0..M with pairing, mode 0
  // and the stride at multiples of 16 so that every 16-order tile the far fill
  // reads and writes is whole cache lines (128 B in FP32, 256 B in FP64)
  p.zoff = h->pairing ? 0 : (p.M + 15) / 16 * 16;
  p.zmodes = (p.zoff + p.M + 1 + 15) / 16 * 16;
  const size_t zel = (size_t)(2 * p.Nn) * (2 * p.Nn) * p.zmodes;
  const size_t zsz = h->precision == GHOST_FP64 ? sizeof(double2) : sizeof(float2);
  if ((rc = dalloc(h, &p.d_zfar, zel * zsz))) return rc;
  if (p.pen) {
    if ((rc = dalloc(h, &p.d_zfar2, zel * zsz))) return rc;
    if ((rc = dalloc(h, &p.d_NB2, (size_t)3 * Ne * 16 * p.nmodes * sizeof(double2)))) return rc;
  }

  if (p.has_zs) {
    if ((rc = dalloc(h, &p.d_zs, (size_t)Ne * sizeof(double2)))) return rc;
    HIPCHK(hipMemcpy(p.d_zs, p.zs.data(), (size_t)Ne * sizeof(double2), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(p.d_pts, p.pts.data(), p.pts.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_nodes, p.nodes.data(), p.nodes.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_elem, elem.data(), elem.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_tabs, tabs.data(), tabs.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_theta, p.theta.data(), p.theta.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_cells, cells.data(), cells.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_core, core.data(), core.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_tail, tail.data(), tail.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_pairs, p.pairs.data(), p.pairs.size() * sizeof(NearPairInfo), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_batches, p.batches.data(), p.batches.size() * sizeof(int2), hipMemcpyHostToDevice));
  p.tail_u.clear();  // keep sizes only
  return 0;
}

static double gemm_flops(const std::vector<GemmOp>& ops, int first, int n, bool triangular) {
  double f = 0.0;
  for (int i = first; i < first + n; ++i) {
    const GemmOp& g = ops[i];
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) continue;
    // complex multiply-add = 8 real flops; a triangular block costs half of a full one
    f += (triangular ? 4.0 : 8.0) * (double)g.M * g.N * g.K;
  }
  return f;
}

// Flattened tile starts (GemmOp::t0) of ops [first, first + n) for the three
// tile shapes; tt receives each shape's block count.
static void tile_prefix(std::vector<GemmOp>& g, int first, int n, int* tt) {
  static const int BM[3] = {64, 64, 128}, BN[3] = {64, 128, 128};
  for (int sh = 0; sh < 3; ++sh) {
    int acc = 0;
    for (int i = first; i < first + n; ++i) {
      GemmOp& o = g[i];
      o.t0[sh] = acc;
      if (o.M > 0 && o.N > 0 && o.K > 0) acc += ((o.M + BM[sh] - 1) / BM[sh]) * ((o.N + BN[sh] - 1) / BN[sh]);
    }
    tt[sh] = acc;
  }
}

static void set_flops(ghost_bor* h, std::vector<Launch>& sched) {
  for (Launch& L : sched) {
    if (L.kind == L_GEMM_AUG_W || L.kind == L_GEMM_Y_W || L.kind == L_GEMM_AUG || L.kind == L_GEMM_Y_LU ||
        L.kind == L_GEMM_TW || L.kind == L_GEMM_TLU)
      tile_prefix(h->h_gemm, L.ops, L.nops, L.tt);
    if (L.kind == L_GEMM_AUG_W || L.kind == L_GEMM_Y_W || L.kind == L_GEMM_TW)
      L.flops = gemm_flops(h->h_gemm, L.ops, L.nops, true);
    else if (L.kind == L_GEMM_AUG || L.kind == L_GEMM_Y_LU || L.kind == L_GEMM_TLU)
      L.flops = gemm_flops(h->h_gemm, L.ops, L.nops, false);
  }
}

static int build_schedule(ghost_bor* h) {
  const int nb = h->nb;
  const int nsys = (int)h->ss.size();
  int maxn = 0;
  for (auto& s : h->ss) maxn = std::max(maxn, s.n);
  h->rpt = (maxn + 1023) / 1024;
  if (h->rpt > 8) return fail(GHOST_ERR_ARG, "system size %d exceeds the panel limit 8192", maxn);
  // pivot offsets
  int poff = 0;
  std::vector<int> pofs(nsys);
  for (int si = 0; si < nsys; ++si) { pofs[si] = poff; poff += h->ss[si].n; }
  const int nblk_max = (maxn + nb - 1) / nb;
  h->max_blocks = nblk_max;
  const bool mixed = h->precision == GHOST_MIXED;
  auto wl = [&](int s, int b) { return ((long long)s * nblk_max + b) * 2 * 64 * 64; };
  auto wu = [&](int s, int b) { return ((long long)s * nblk_max + b) * 2 * 64 * 64 + 64 * 64; };
  auto tiles = [](int M, int N) { return (M > 0 && N > 0) ? ((M + 63) / 64) * ((N + 63) / 64) : 0; };
  // blocked forward (unit-lower L, diagonal blocks through W = I - L11^-1) and
  // backward (U, through the upper inverses) substitution on an RHS block of
  // nc(si) columns at element offset off(si), row-major with ld = nc(si)
  auto add_tri_solves = [&](std::vector<Launch>& fwd, std::vector<Launch>& back, auto off, auto nc) {
    for (int b = 0; b < nblk_max; ++b) {
      const int k = b * nb;
      Launch g1{L_GEMM_Y_W, (int)h->h_gemm.size(), nsys, 0, 0};
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        const long long o = off(si);
        const int c = nc(si);
        GemmOp go{wl(si, b), o + (long long)k * c, o + (long long)k * c, 64, c, c, w, c, w};
        h->h_gemm.push_back(go);
        g1.gx = std::max(g1.gx, tiles(go.M, go.N));
      }
      Launch g2{L_GEMM_Y_LU, (int)h->h_gemm.size(), nsys, 0, 0};
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        const long long o = off(si);
        const int c = nc(si);
        GemmOp go{s.a_off + (long long)(k + w) * s.ld + k, o + (long long)k * c, o + (long long)(k + w) * c,
                  s.ld, c, c, s.n - k - w, c, w};
        if (w == 0) go.M = 0;
        h->h_gemm.push_back(go);
        g2.gx = std::max(g2.gx, tiles(go.M, go.N));
      }
      fwd.push_back(g1);
      fwd.push_back(g2);
    }
    for (int st = 0; st < nblk_max; ++st) {
      Launch g1{L_GEMM_Y_W, (int)h->h_gemm.size(), nsys, 0, 0};
      std::vector<GemmOp> upd;
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int nbs = (s.n + nb - 1) / nb;
        const int b = nbs - 1 - st;
        const long long o = off(si);
        const int c = nc(si);
        GemmOp go{0, 0, 0, 64, c, c, 0, 0, 0};
        GemmOp gu{0, 0, 0, s.ld, c, c, 0, 0, 0};
        if (b >= 0) {
          const int kb = b * nb, w = std::min(nb, s.n - kb);
          go = GemmOp{wu(si, b), o + (long long)kb * c, o + (long long)kb * c, 64, c, c, w, c, w};
          gu = GemmOp{s.a_off + kb, o + (long long)kb * c, o, s.ld, c, c, kb, c, w};
        }
        h->h_gemm.push_back(go);
        g1.gx = std::max(g1.gx, tiles(go.M, go.N));
        upd.push_back(gu);
      }
      Launch g2{L_GEMM_Y_LU, (int)h->h_gemm.size(), nsys, 0, 0};
      for (auto& gu : upd) { h->h_gemm.push_back(gu); g2.gx = std::max(g2.gx, tiles(gu.M, gu.N)); }
      back.push_back(g1);
      back.push_back(g2);
    }
  };

  // ── LU steps on the augmented systems, in super-blocks of S panels ──
  // Inside a super-block starting at column k0, panel b (columns [k, k+w))
  // first catches up on the super-block's earlier panels (rows [k, n), K =
  // k - k0), is factored and its interchanges applied, then its row block
  // catches up on the right (rows [k, k+w), columns [k+w, ncols)) and is
  // multiplied by L11^-1.  One trailing update with K = k1 - k0 then covers
  // rows and columns beyond the super-block: the trailing matrix is read and
  // written once per S panels instead of once per panel.
  // deeper super-blocks pay off for the FP32 factors, whose split-bf16 GEMM
  // is bound by the C-tile traffic of short-K updates (measured at n ~ 2400:
  // S = 4 / 8 / 12 give 137.0 / 134.6 / 133.8 ms per step); the FP64 GEMM is
  // MFMA-bound and S = 4..12 are within 2 %
  const int S = h->lu_super > 0 ? h->lu_super : (h->precision == GHOST_FP64 ? 4 : 8);
  h->nparts = nsys >= 2 ? 2 : 1;
  for (int part = 0; part < 3; ++part) {
  if (part == 1 && h->nparts < 2) continue;
  std::vector<int> psys;
  for (int si = part == 2 ? 0 : part; si < nsys; si += part == 2 ? 1 : h->nparts) psys.push_back(si);
  const int np = (int)psys.size();
  auto lu_gemm = [&](int kind, auto make) {
    Launch g{kind, (int)h->h_gemm.size(), np, 0, 0};
    for (int si : psys) {
      GemmOp go = make(h->ss[si]);
      if (go.M <= 0 || go.N <= 0 || go.K <= 0) go.M = go.N = go.K = 0;
      h->h_gemm.push_back(go);
      g.gx = std::max(g.gx, tiles(go.M, go.N));
    }
    h->lu_sched[part].push_back(g);
  };
  // the second part's super-blocks start half a super-block later: its panel
  // chains (latency-bound, one work-group per system) then fall on the first
  // part's trailing updates and the other way round, where two parts in step
  // would run their chains together and their trailing updates together
  const int first_sb = (part == 1 && S >= 2) ? S / 2 : S;
  for (int b0 = 0, sbw = first_sb; b0 < nblk_max; b0 += sbw, sbw = S) {
    const int k0 = b0 * nb, nsb = std::min(sbw, nblk_max - b0);
    for (int sb = 0; sb < nsb; ++sb) {
      const int b = b0 + sb, k = b * nb;
      if (sb > 0)   // panel catch-up: A[k:n, k:k+w] -= L[k:n, k0:k] U[k0:k, k:k+w]
        lu_gemm(L_GEMM_AUG, [&](const SysInfo& s) {
          const int w = k < s.n ? std::min(nb, s.n - k) : 0;
          return GemmOp{s.a_off + (long long)k * s.ld + k0, s.a_off + (long long)k0 * s.ld + k,
                        s.a_off + (long long)k * s.ld + k, s.ld, s.ld, s.ld, s.n - k, w, k - k0};
        });
      // only the systems that still have a panel here (uneven batches); the rows
      // per thread follow the rows left below k, not the largest system
      Launch lp{L_PANEL, (int)h->h_panel.size(), 0, 0, 1};
      Launch ls{L_SWAP, (int)h->h_swap.size(), 0, 0, 0};
      Launch lt{L_TRIL, (int)h->h_trsm.size(), 0, 0, 0};
      for (int si : psys) {
        const SysInfo& s = h->ss[si];
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        if (w <= 0) continue;
        const int ncols = s.n + s.nrhs;
        h->h_panel.push_back(PanelOp{s.a_off, s.n, s.ld, k, w, pofs[si], si});
        lp.arg = std::max(lp.arg, (s.n - k + 1023) / 1024);
        h->h_swap.push_back(SwapOp{s.a_off, s.ld, k, w, k, k + w, ncols, pofs[si]});
        ls.gx = std::max(ls.gx, (k + (ncols - k - w) + 255) / 256);
        h->h_trsm.push_back(TrsmOp{s.a_off + (long long)k * s.ld + k, wl(si, b), s.ld, w});
        ++lp.nops;
        ++ls.nops;
        ++lt.nops;
      }
      lp.gx = lp.nops;
      lt.gx = lt.nops;
      if (lp.nops > 0) {   // a part may have run out of panels before the largest system
        h->lu_sched[part].push_back(lp);
        h->lu_sched[part].push_back(ls);
        h->lu_sched[part].push_back(lt);
      }
      if (sb > 0)   // row catch-up: A[k:k+w, k+w:] -= L[k:k+w, k0:k] U[k0:k, k+w:]
        lu_gemm(L_GEMM_AUG, [&](const SysInfo& s) {
          const int w = k < s.n ? std::min(nb, s.n - k) : 0;
          return GemmOp{s.a_off + (long long)k * s.ld + k0, s.a_off + (long long)k0 * s.ld + k + w,
                        s.a_off + (long long)k * s.ld + k + w, s.ld, s.ld, s.ld, w, s.n + s.nrhs - k - w,
                        k - k0};
        });
      // X = L11^-1 X on rows [k, k+w), columns [k+w, ncols)
      lu_gemm(L_GEMM_AUG_W, [&](const SysInfo& s) {
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        return GemmOp{wl((int)(&s - h->ss.data()), b), s.a_off + (long long)k * s.ld + k + w,
                      s.a_off + (long long)k * s.ld + k + w, 64, s.ld, s.ld, w, s.n + s.nrhs - k - w, w};
      });
    }
    // trailing update: A[k1:n, k1:] -= L[k1:n, k0:k1] U[k0:k1, k1:]
    const int k1 = k0 + nsb * nb;
    lu_gemm(L_GEMM_AUG, [&](const SysInfo& s) {
      return GemmOp{s.a_off + (long long)k1 * s.ld + k0, s.a_off + (long long)k0 * s.ld + k1,
                    s.a_off + (long long)k1 * s.ld + k1, s.ld, s.ld, s.ld, s.n - k1, s.n + s.nrhs - k1, k1 - k0};
    });
  }
  // ── upper diagonal-block inverses (all blocks, one launch) ──
  h->triu_ops[part] = (int)h->h_trsm.size();
  for (int si : psys) {
    const SysInfo& s = h->ss[si];
    for (int b = 0; b < nblk_max; ++b) {
      const int k = b * nb;
      const int w = k < s.n ? std::min(nb, s.n - k) : 0;
      if (w > 0) h->h_trsm.push_back(TrsmOp{s.a_off + (long long)k * s.ld + k, wu(si, b), s.ld, w});
    }
  }
  h->triu_n[part] = (int)h->h_trsm.size() - h->triu_ops[part];
  // ── back substitution on the RHS columns, in super-blocks of S row blocks ──
  // Walking up from the last block: block b catches up on the super-block's
  // blocks below it (K = rows solved so far in the super-block), is solved
  // through W_u, and after the super-block one update with K up to nb x S
  // carries its rows into every row above.
  auto back_gemm = [&](int kind, auto make) {
    Launch g{kind, (int)h->h_gemm.size(), np, 0, 0};
    for (int si : psys) {
      GemmOp go = make(si, h->ss[si]);
      if (go.M <= 0 || go.N <= 0 || go.K <= 0) go.M = go.N = go.K = 0;
      h->h_gemm.push_back(go);
      g.gx = std::max(g.gx, tiles(go.M, go.N));
    }
    h->back_sched[part].push_back(g);
  };
  for (int st0 = 0; st0 < nblk_max; st0 += S) {
    const int nsb = std::min(S, nblk_max - st0);
    // first row past the super-block's already-solved part, per system
    auto k_end = [&](const SysInfo& s) { return std::min(s.n, ((s.n + nb - 1) / nb - st0) * nb); };
    for (int sb = 0; sb < nsb; ++sb) {
      const int st = st0 + sb;
      auto blk = [&](const SysInfo& s, int& kb, int& w) {
        const int b = (s.n + nb - 1) / nb - 1 - st;
        kb = b * nb;
        w = b >= 0 ? std::min(nb, s.n - kb) : 0;
        return b;
      };
      if (sb > 0)   // X_b -= U[kb:kb+w, kb+w:k_end] X[kb+w:k_end]
        back_gemm(L_GEMM_AUG, [&](int, const SysInfo& s) {
          int kb, w;
          if (blk(s, kb, w) < 0) return GemmOp{0, 0, 0, s.ld, s.ld, s.ld, 0, 0, 0};
          return GemmOp{s.a_off + (long long)kb * s.ld + kb + w, s.a_off + (long long)(kb + w) * s.ld + s.n,
                        s.a_off + (long long)kb * s.ld + s.n, s.ld, s.ld, s.ld, w, s.nrhs, k_end(s) - kb - w};
        });
      back_gemm(L_GEMM_AUG_W, [&](int si, const SysInfo& s) {   // X_b = U_bb^-1 X_b
        int kb, w;
        const int b = blk(s, kb, w);
        if (b < 0) return GemmOp{0, 0, 0, 64, s.ld, s.ld, 0, 0, 0};
        return GemmOp{wu(si, b), s.a_off + (long long)kb * s.ld + s.n, s.a_off + (long long)kb * s.ld + s.n, 64,
                      s.ld, s.ld, w, s.nrhs, w};
      });
    }
    // rows above the super-block: X[0:k_lo] -= U[0:k_lo, k_lo:k_end] X[k_lo:k_end]
    back_gemm(L_GEMM_AUG, [&](int, const SysInfo& s) {
      const int b_lo = (s.n + nb - 1) / nb - 1 - (st0 + nsb - 1);
      const int k_lo = std::max(0, b_lo) * nb;
      return GemmOp{s.a_off + k_lo, s.a_off + (long long)k_lo * s.ld + s.n, s.a_off + s.n, s.ld, s.ld, s.ld, k_lo,
                    s.nrhs, k_end(s) - k_lo};
    });
  }
  }  // parts
  // ── VecOps over all systems (copies, casts, norms) ──
  std::vector<int> c_off(nsys);
  for (int si = 0, co = 0; si < nsys; ++si) { c_off[si] = co; co += h->ss[si].nrhs; }
  h->vec_all = (int)h->h_vec.size();
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->ss[si];
    h->h_vec.push_back(VecOp{s.x_off, s.x_off, s.nrhs, s.nrhs, s.n, s.nrhs, 0, c_off[si]});
  }
  h->vec_aug = (int)h->h_vec.size();
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->ss[si];
    h->h_vec.push_back(VecOp{s.a_off + s.n, s.x_off, s.ld, s.nrhs, s.n, s.nrhs, 0, c_off[si]});
  }
  h->vec_toaug = (int)h->h_vec.size();
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->ss[si];
    h->h_vec.push_back(VecOp{s.x_off, s.a_off + s.n, s.nrhs, s.ld, s.n, s.nrhs, 0, c_off[si]});
  }
  // ── residual / refinement GEMM: R = B - A64 X (X in X64 or in the augmented columns) ──
  h->res_gemm = (int)h->h_gemm.size();
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->ss[si];
    h->h_gemm.push_back(GemmOp{s.a64_off, s.x_off, s.x_off, s.n, s.nrhs, s.nrhs, s.n, s.nrhs, s.n});
  }
  h->res_gemm_aug = (int)h->h_gemm.size();
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->ss[si];
    h->h_gemm.push_back(GemmOp{s.a64_off, s.a_off + s.n, s.x_off, s.n, s.ld, s.nrhs, s.n, s.nrhs, s.n});
  }
  if (mixed) {
    // probe: BS = B64 S (K = nrhs), XS = X64 S, then BS -= A64 XS (K = n)
    h->probe_ops = (int)h->h_gemm.size();
    long long po = 0;
    h->p_off.assign(nsys, 0);
    for (int si = 0; si < nsys; ++si) { h->p_off[si] = po; po += 2LL * h->ss[si].n; }
    h->vec_probe = (int)h->h_vec.size();
    for (int si = 0; si < nsys; ++si)
      h->h_vec.push_back(VecOp{h->p_off[si], h->p_off[si], 2, 2, h->ss[si].n, 2, 0, 2 * si});
    for (int si = 0; si < nsys; ++si) {   // C = 0 - B (-S): written into BS as +B S (see ghost_bor_probe)
      const SysInfo& s = h->ss[si];
      h->h_gemm.push_back(GemmOp{s.x_off, 0, h->p_off[si], s.nrhs, 2, 2, s.n, 2, s.nrhs});
    }
    for (int si = 0; si < nsys; ++si) {
      const SysInfo& s = h->ss[si];
      h->h_gemm.push_back(GemmOp{s.x_off, 0, h->p_off[si], s.nrhs, 2, 2, s.n, 2, s.nrhs});
    }
    for (int si = 0; si < nsys; ++si) {
      const SysInfo& s = h->ss[si];
      h->h_gemm.push_back(GemmOp{s.a64_off, h->p_off[si], h->p_off[si], s.n, 2, 2, s.n, 2, s.n});
    }
    // forward / backward solves with the FP32 factors on the refinement
    // corrections Y (n x nrhs at x_off); the probe's n x 2 blocks at p_off
    // take k_trsv2 instead
    add_tri_solves(h->fwdY_sched, h->backY_sched, [&](int si) { return h->ss[si].x_off; },
                   [&](int si) { return h->ss[si].nrhs; });
  }
  {
    // transposed solves on the lambda block (n x nrhs at x_off): with P A = L U,
    // A^T lambda = G is U^T y = G (forward, diagonal blocks (U11^-1)^T = I - W_u^T),
    // then L^T z = y (backward, I - W_l^T), then lambda = P^T z
    for (int b = 0; b < nblk_max; ++b) {
      const int k = b * nb;
      Launch g1{L_GEMM_TW, (int)h->h_gemm.size(), nsys, 0, 0};
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        const long long o = s.x_off;
        const int c = s.nrhs;
        h->h_gemm.push_back(GemmOp{wu(si, b), o + (long long)k * c, o + (long long)k * c, 64, c, c, w, c, w});
      }
      Launch g2{L_GEMM_TLU, (int)h->h_gemm.size(), nsys, 0, 0};
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int w = k < s.n ? std::min(nb, s.n - k) : 0;
        const long long o = s.x_off;
        const int c = s.nrhs;
        GemmOp go{s.a_off + (long long)k * s.ld + k + w, o + (long long)k * c, o + (long long)(k + w) * c, s.ld, c,
                  c, s.n - k - w, c, w};
        if (w == 0) go.M = 0;
        h->h_gemm.push_back(go);
      }
      h->adjT_fwd.push_back(g1);
      h->adjT_fwd.push_back(g2);
    }
    for (int st = 0; st < nblk_max; ++st) {
      Launch g1{L_GEMM_TW, (int)h->h_gemm.size(), nsys, 0, 0};
      std::vector<GemmOp> upd;
      for (int si = 0; si < nsys; ++si) {
        const SysInfo& s = h->ss[si];
        const int b = (s.n + nb - 1) / nb - 1 - st;
        const long long o = s.x_off;
        const int c = s.nrhs;
        GemmOp go{0, 0, 0, 64, c, c, 0, 0, 0};
        GemmOp gu{0, 0, 0, s.ld, c, c, 0, 0, 0};
        if (b >= 0) {
          const int kb = b * nb, w = std::min(nb, s.n - kb);
          go = GemmOp{wl(si, b), o + (long long)kb * c, o + (long long)kb * c, 64, c, c, w, c, w};
          gu = GemmOp{s.a_off + (long long)kb * s.ld, o + (long long)kb * c, o, s.ld, c, c, kb, c, w};
        }
        h->h_gemm.push_back(go);
        upd.push_back(gu);
      }
      Launch g2{L_GEMM_TLU, (int)h->h_gemm.size(), nsys, 0, 0};
      for (auto& gu : upd) h->h_gemm.push_back(gu);
      h->adjT_back.push_back(g1);
      h->adjT_back.push_back(g2);
    }
    // Q[:, :n] = -Lambda X^T (X stored n x nrhs, used transposed)
    h->adj_q_ops = (int)h->h_gemm.size();
    for (int si = 0; si < nsys; ++si) {
      const SysInfo& s = h->ss[si];
      h->h_gemm.push_back(GemmOp{s.x_off, s.x_off, s.a_off, s.nrhs, s.nrhs, s.ld, s.n, s.n, s.nrhs});
    }
  }
  for (auto& sc : h->back_sched)
    for (auto& L : sc) L.fam = K_GEMM_BACK;
  for (auto* sc : {&h->fwdY_sched, &h->backY_sched, &h->adjT_fwd, &h->adjT_back})
    for (auto& L : *sc) L.fam = K_GEMM_SOLVE;
  tile_prefix(h->h_gemm, h->res_gemm, nsys, h->res_tt);
  tile_prefix(h->h_gemm, h->res_gemm_aug, nsys, h->res_aug_tt);
  if (h->adj_q_ops >= 0) tile_prefix(h->h_gemm, h->adj_q_ops, nsys, h->adj_q_tt);
  for (int part = 0; part < 3; ++part) {
    set_flops(h, h->lu_sched[part]);
    set_flops(h, h->back_sched[part]);
  }
  set_flops(h, h->adjT_fwd);
  set_flops(h, h->adjT_back);
  set_flops(h, h->fwdY_sched);
  set_flops(h, h->backY_sched);
  for (size_t i = 0; i < h->h_vec.size(); ++i) h->h_vec[i].piv = pofs[i % nsys];
  h->piv_elems = poff;
  int rc;
  if ((rc = dalloc(h, &h->d_panel, h->h_panel.size() * sizeof(PanelOp)))) return rc;
  if ((rc = dalloc(h, &h->d_swap, h->h_swap.size() * sizeof(SwapOp)))) return rc;
  if ((rc = dalloc(h, &h->d_trsm, h->h_trsm.size() * sizeof(TrsmOp)))) return rc;
  if ((rc = dalloc(h, &h->d_gemm, h->h_gemm.size() * sizeof(GemmOp)))) return rc;
  if ((rc = dalloc(h, &h->d_vec, h->h_vec.size() * sizeof(VecOp)))) return rc;
  HIPCHK(hipMemcpy(h->d_panel, h->h_panel.data(), h->h_panel.size() * sizeof(PanelOp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_swap, h->h_swap.data(), h->h_swap.size() * sizeof(SwapOp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_trsm, h->h_trsm.data(), h->h_trsm.size() * sizeof(TrsmOp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gemm, h->h_gemm.data(), h->h_gemm.size() * sizeof(GemmOp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_vec, h->h_vec.data(), h->h_vec.size() * sizeof(VecOp), hipMemcpyHostToDevice));
  return 0;
}

// Far azimuth tables of n_xi nE / nB with exact values at 0, pi/2, pi, 3pi/2:
// sin^2(pi j/nE) | cos(2 pi i/nE) | cos, sin(2 pi j/nB) half tables | full cos, sin.
static void azimuth_tables(int nE, int nB, std::vector<double>& tabs, size_t* off) {
  auto cos_exact = [](long long i, int N) {
    i %= N;
    if (i == 0) return 1.0;
    if (2 * i == N) return -1.0;
    if (4 * i == N || 4 * i == 3 * (long long)N) return 0.0;
    return std::cos(2.0 * M_PI * (double)i / N);
  };
  auto sin_exact = [](long long i, int N) {
    i %= N;
    if (i == 0 || 2 * i == N) return 0.0;
    if (4 * i == N) return 1.0;
    if (4 * i == 3 * (long long)N) return -1.0;
    return std::sin(2.0 * M_PI * (double)i / N);
  };
  tabs.clear();
  off[0] = tabs.size();
  for (int j = 0; j <= nE / 2; ++j) { const double v = std::sin(M_PI * j / nE); tabs.push_back(j == nE / 2 ? 1.0 : v * v); }
  off[1] = tabs.size();
  for (int i = 0; i < nE; ++i) tabs.push_back(cos_exact(i, nE));
  off[2] = tabs.size();
  for (int j = 0; j <= nB / 2; ++j) tabs.push_back(cos_exact(j, nB));
  off[3] = tabs.size();
  for (int j = 0; j <= nB / 2; ++j) tabs.push_back(sin_exact(j, nB));
  off[4] = tabs.size();
  for (int i = 0; i < nB; ++i) tabs.push_back(cos_exact(i, nB));
  off[5] = tabs.size();
  for (int i = 0; i < nB; ++i) tabs.push_back(sin_exact(i, nB));
}

template <typename T> static int upload(ghost_bor* h, T** d, const std::vector<T>& v) {
  int rc = dalloc(h, d, std::max<size_t>(1, v.size()) * sizeof(T));
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static int prepare_cross(ghost_bor* h, Cross& c) {
  const Problem& a = h->inner[c.op_p];
  const Problem& b = h->inner[c.op_q];
  // cell tables: 0 tensor grid, 1 + 2a + b corner<a><b>
  std::vector<double> cs[5], csp[5], cw[5];
  tensor_points(c.near_order, cs[0], csp[0], cw[0]);
  for (int kd = 1; kd < 5; ++kd) graded_points(false, (double)((kd - 1) >> 1), (double)((kd - 1) & 1), cs[kd], csp[kd], cw[kd]);
  int off = 0;
  for (int kd = 0; kd < 5; ++kd) { c.cell_off[kd] = off; c.cell_cnt[kd] = (int)cs[kd].size(); off += c.cell_cnt[kd]; }
  c.cells.clear();
  for (int kd = 0; kd < 5; ++kd) c.cells.insert(c.cells.end(), cs[kd].begin(), cs[kd].end());
  for (int kd = 0; kd < 5; ++kd) c.cells.insert(c.cells.end(), csp[kd].begin(), csp[kd].end());
  for (int kd = 0; kd < 5; ++kd) c.cells.insert(c.cells.end(), cw[kd].begin(), cw[kd].end());
  // tail orders per pair (bor_kernels.py:238-240), point-pair offsets, batches
  std::map<int, int> tail_of;
  int pp = 0;
  c.batches.clear();
  for (size_t i = 0; i < c.pairs.size(); ++i) {
    NearPairInfo& pi = c.pairs[i];
    const int kd = pi.kind;
    double osc = 0.0;
    const double r0e = a.nodes[2 * pi.e], dre = a.nodes[2 * pi.e + 2] - r0e;
    const double r0f = b.nodes[2 * pi.f], drf = b.nodes[2 * pi.f + 2] - r0f;
    for (size_t q = 0; q < cs[kd].size(); ++q) {
      const double rp = r0e + cs[kd][q] * dre, rq = r0f + csp[kd][q] * drf;
      osc = std::max(osc, std::hypot(c.k, c.ki) * std::sqrt(std::max(4.0 * rp * rq, 1e-300)));
    }
    const int nt = (int)std::min(1024.0, std::max(64.0, std::ceil(4.0 * (osc / M_PI + (c.M + 2)))));
    if (!tail_of.count(nt)) {
      tail_of[nt] = (int)c.tail_u.size();
      std::vector<double> xt, wt;
      gauss_legendre(nt, xt, wt);
      for (int j = 0; j < nt; ++j) { c.tail_u.push_back(0.5 * (xt[j] + 1.0)); c.tail_w.push_back(0.5 * wt[j]); }
    }
    pi.n_tail = nt;
    pi.tail_off = tail_of[nt];
    pi.pp_off = pp;
    for (int q = 0; q < c.cell_cnt[kd]; q += 32) c.batches.push_back(make_int2((int)i, q));
    pp += c.cell_cnt[kd];
  }
  c.npp = pp;
  // pairs by parity class of (e, f): disjoint footprints within a class
  c.plist.clear();
  for (int cls = 0; cls < 4; ++cls) {
    c.par_off[cls] = (int)c.plist.size();
    for (size_t i = 0; i < c.pairs.size(); ++i)
      if (2 * (c.pairs[i].e & 1) + (c.pairs[i].f & 1) == cls) c.plist.push_back((int)i);
  }
  c.par_off[4] = (int)c.plist.size();
  c.zoff = h->pairing ? 0 : (c.M + 15) / 16 * 16;
  c.zmodes = (c.zoff + c.M + 1 + 15) / 16 * 16;
  std::vector<double> tabs;
  azimuth_tables(c.nE, c.nB, tabs, c.tab_off);
  std::vector<double> core;
  {
    std::vector<double> x48, w48;
    gauss_legendre(48, x48, w48);
    for (int i = 0; i < 48; ++i) core.push_back(0.5 * (x48[i] + 1.0));
    for (int i = 0; i < 48; ++i) core.push_back(0.5 * w48[i]);
  }
  std::vector<double> tail(c.tail_u);
  tail.insert(tail.end(), c.tail_w.begin(), c.tail_w.end());
  // unit weights: the brackets take the source element's, the (zero) mass
  // term reads the test element's, so both surfaces' element counts fit
  std::vector<double2> ones(std::max(c.Nep, c.Neq), make_double2(1.0, 0.0));
  int rc;
  if ((rc = upload(h, &c.d_tabs, tabs)) || (rc = upload(h, &c.d_skip, c.skip)) || (rc = upload(h, &c.d_cells, c.cells)) ||
      (rc = upload(h, &c.d_core, core)) || (rc = upload(h, &c.d_tail, tail)) || (rc = upload(h, &c.d_pairs, c.pairs)) ||
      (rc = upload(h, &c.d_batches, c.batches)) || (rc = upload(h, &c.d_plist, c.plist)))
    return rc;
  double2* d_ones = nullptr;
  if ((rc = upload(h, &d_ones, ones))) return rc;
  c.d_zs = d_ones;
  const size_t np = std::max<size_t>(1, c.pairs.size());
  if ((rc = dalloc(h, &c.d_NP, std::max<size_t>(1, (size_t)c.npp) * (c.M + 2) * 5 * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &c.d_NB, np * 16 * c.nmodes * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &c.d_NB2, np * 16 * c.nmodes * sizeof(double2)))) return rc;
  const size_t zel = (size_t)(2 * c.Np) * (2 * c.Nq) * c.zmodes;
  if ((rc = dalloc(h, &c.d_zfar, zel * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &c.d_zfar2, zel * sizeof(double2)))) return rc;
  return 0;
}

// Device side of a region set: full-system, right-hand-side and current
// buffers, the junction map as lists, one virtual system entry per mode for
// each exterior surface (its rows of the full vectors, unmasked).
static int prepare_region(ghost_bor* h, Region& r) {
  int rc;
  const size_t nf = (size_t)r.n_full;
  if ((rc = dalloc(h, &r.d_full, (size_t)r.nmodes * nf * nf * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &r.d_bfull, (size_t)r.nmodes * nf * r.nrhs * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &r.d_xfull, (size_t)r.nmodes * nf * r.nrhs * sizeof(double2)))) return rc;
  std::vector<int> ptr((size_t)r.nmodes * (nf + 1), 0), mem((size_t)r.nmodes * nf, 0);
  std::vector<double> cf((size_t)r.nmodes * nf, 0.0);
  for (int mi = 0; mi < r.nmodes; ++mi) {
    const int nr = h->sys[r.sys0 + mi].n;
    std::vector<std::vector<int>> lists(nr);
    for (int i = 0; i < r.n_full; ++i)
      if (r.red[mi][i] >= 0) lists[r.red[mi][i]].push_back(i);
    int* P = &ptr[(size_t)mi * (nf + 1)];
    int u = 0;
    for (int q = 0; q < nr; ++q) {
      P[q] = u;
      for (int i : lists[q]) {
        mem[(size_t)mi * nf + u] = i;
        cf[(size_t)mi * nf + u] = r.cf[mi][i];
        ++u;
      }
    }
    P[nr] = u;
  }
  if ((rc = upload(h, &r.d_ptr, ptr)) || (rc = upload(h, &r.d_mem, mem)) || (rc = upload(h, &r.d_cf, cf))) return rc;
  r.d_vsys.clear();
  for (const auto& e : r.ext) {
    const Problem& o = h->inner[e.op];
    std::vector<SysInfo> vs(r.nmodes);
    for (int mi = 0; mi < r.nmodes; ++mi) {
      SysInfo& v = vs[mi];
      memset(&v, 0, sizeof(v));
      v.prob = -1;
      v.m = mi - r.M;
      v.nt = o.Nn;
      v.t0 = 0;
      v.f0 = 0;
      v.pen = e.off_M >= 0 ? 2 * o.Nn : 0;
      v.n = e.off_M >= 0 ? 4 * o.Nn : 2 * o.Nn;
      v.nrhs = r.nrhs;
      v.ncols = r.nrhs;
      v.x_off = ((long long)mi * r.n_full + e.off_J) * r.nrhs;
    }
    SysInfo* d = nullptr;
    if ((rc = upload(h, &d, vs))) return rc;
    r.d_vsys.push_back(d);
  }
  return 0;
}

extern "C" int ghost_bor_build(ghost_bor* h) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "already built");
  if (h->sys.empty()) return fail(GHOST_ERR_STATE, "no problems or dense systems added");
  for (size_t ri = 0; ri < h->regions.size(); ++ri) {
    const Region& r = h->regions[ri];
    for (int mi = 0; mi < r.nmodes; ++mi)
      if (h->sys[r.sys0 + mi].n < 1) return fail(GHOST_ERR_STATE, "region set %zu: no map for mode %d", ri, mi - r.M);
    if (r.blocks.empty()) return fail(GHOST_ERR_STATE, "region set %zu has no blocks", ri);
  }
  HIPCHK(hipSetDevice(h->device));
  const bool mixed = h->precision == GHOST_MIXED;
  // systems to solve: every mode, or with pairing modes m >= 0 only (mode -m
  // reads mode m's right-hand sides and currents with the signs of the pair)
  h->ss.clear();
  for (auto& s : h->sys) {
    const bool paired = h->pairing && s.prob >= 0;
    s.col = 0;
    s.flip = 0;
    if (paired && s.m < 0) continue;
    s.ncols = s.nrhs;
    s.owner = (int)h->ss.size();
    SysInfo o = s;
    o.nrhs = o.ncols;
    h->ss.push_back(o);
  }
  for (auto& s : h->sys)
    if (h->pairing && s.prob >= 0 && s.m < 0) {
      const SysInfo& tw = h->sys[h->probs[s.prob].sys0 + h->probs[s.prob].M - s.m];
      if (tw.n != s.n || tw.nt != s.nt || tw.t0 != s.t0 || tw.f0 != s.f0)
        return fail(GHOST_ERR_STATE, "modes %d and %d have different bases", s.m, -s.m);
      s.owner = tw.owner;
      s.ncols = tw.ncols;
      s.col = 0;
      s.flip = 1;
    }
  // offsets of the solved systems, shared by their mode entries
  size_t a = 0, x = 0, a64 = 0;
  for (auto& s : h->ss) {
    s.ld = ((s.n + s.nrhs + 15) / 16) * 16;
    s.a_off = (long long)a;
    s.x_off = (long long)x;
    s.a64_off = (long long)a64;
    a += (size_t)s.n * s.ld;
    x += (size_t)s.n * s.nrhs;
    a64 += (size_t)s.n * s.n;
  }
  for (auto& s : h->sys) {
    const SysInfo& o = h->ss[s.owner];
    s.ld = o.ld;
    s.a_off = o.a_off;
    s.x_off = o.x_off;
    s.a64_off = o.a64_off;
  }
  h->a_elems = a;
  h->x_elems = x;
  h->a64_elems = a64;
  int rc;
  for (auto& p : h->probs)
    if ((rc = prepare_problem(h, p))) return rc;
  for (auto& p : h->inner)
    if ((rc = prepare_problem(h, p))) return rc;
  for (auto& c : h->cross)
    if ((rc = prepare_cross(h, c))) return rc;
  for (auto& r : h->regions)
    if ((rc = prepare_region(h, r))) return rc;
  if ((rc = build_schedule(h))) return rc;
  const size_t csz = mixed ? sizeof(float2) : sizeof(double2);
  if ((rc = dalloc(h, &h->d_sys, h->sys.size() * sizeof(SysInfo)))) return rc;
  HIPCHK(hipMemcpy(h->d_sys, h->sys.data(), h->sys.size() * sizeof(SysInfo), hipMemcpyHostToDevice));
  if ((rc = dalloc(h, &h->d_A, h->a_elems * csz))) return rc;
  if (mixed || h->residual) {
    if ((rc = dalloc(h, &h->d_A64, h->a64_elems * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_B64, h->x_elems * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_R64, h->x_elems * sizeof(double2)))) return rc;
  }
  if (mixed) {
    if ((rc = dalloc(h, &h->d_X64, h->x_elems * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_Y32, h->x_elems * sizeof(float2)))) return rc;
    size_t pe = 0, nrmax = 0;
    for (auto& s : h->ss) { pe += 2 * (size_t)s.n; nrmax = std::max(nrmax, (size_t)s.nrhs); }
    if ((rc = dalloc(h, &h->d_S, nrmax * 2 * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_BS, pe * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_XS, pe * sizeof(double2)))) return rc;
    if ((rc = dalloc(h, &h->d_P32, pe * sizeof(float2)))) return rc;
    std::vector<TrsvOp> tv;
    for (size_t si = 0; si < h->ss.size(); ++si) {
      const SysInfo& s = h->ss[si];
      tv.push_back(TrsvOp{s.a_off, (long long)si * h->max_blocks * 2 * 64 * 64, h->p_off[si], s.ld, s.n});
    }
    if ((rc = dalloc(h, &h->d_trsv, tv.size() * sizeof(TrsvOp)))) return rc;
    HIPCHK(hipMemcpy(h->d_trsv, tv.data(), tv.size() * sizeof(TrsvOp), hipMemcpyHostToDevice));
    size_t ncol = 0;
    for (auto& s : h->ss) ncol += (size_t)s.nrhs;
    if ((rc = dalloc(h, &h->d_bscale, ncol * sizeof(double)))) return rc;
    if ((rc = dalloc(h, &h->d_rscale, ncol * sizeof(double)))) return rc;
    if ((rc = dalloc(h, &h->d_pscale, 2 * h->ss.size() * sizeof(double)))) return rc;
    // S: fixed pseudo-random unit-modulus entries, -1 * S so that k_gemm_update's C -= A B adds B S
    std::vector<double2> S(nrmax * 2);
    unsigned long long st = 0x9E3779B97F4A7C15ULL;
    for (auto& z : S) {
      st ^= st << 13; st ^= st >> 7; st ^= st << 17;
      const double ph = 2.0 * M_PI * (double)(st >> 11) / 9007199254740992.0;
      z = make_double2(-std::cos(ph), -std::sin(ph));
    }
    HIPCHK(hipMemcpy(h->d_S, S.data(), S.size() * sizeof(double2), hipMemcpyHostToDevice));
  }
  if ((rc = dalloc(h, &h->d_piv, (size_t)h->piv_elems * sizeof(int)))) return rc;
  if ((rc = dalloc(h, &h->d_perm, (size_t)h->piv_elems * sizeof(int)))) return rc;
  if ((rc = dalloc(h, &h->d_info, h->ss.size() * sizeof(int)))) return rc;
  if ((rc = dalloc(h, &h->d_W, h->ss.size() * (size_t)h->max_blocks * 2 * 64 * 64 * csz))) return rc;
  h->dF_elems = 0;
  h->partial_elems = 0;
  for (auto& p : h->probs) {
    p.dF_off = h->dF_elems;
    h->dF_elems += (size_t)p.nmodes * p.nrhs;
    const size_t nch = (p.P + h->ff_chunk - 1) / h->ff_chunk;
    h->partial_elems = std::max(h->partial_elems, nch * p.nmodes * p.nrhs);
  }
  for (auto& r : h->regions) {
    r.dF_off = h->dF_elems;
    h->dF_elems += (size_t)r.nmodes * r.nrhs;
    for (const auto& e : r.ext) {
      const size_t nch = (h->inner[e.op].P + h->ff_chunk - 1) / h->ff_chunk;
      h->partial_elems = std::max(h->partial_elems, nch * r.nmodes * r.nrhs);
    }
  }
  if ((rc = dalloc(h, &h->d_dF, h->dF_elems * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_partial, h->partial_elems * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_red, 2 * h->sys.size() * sizeof(double)))) return rc;   // per-system probe ratios
  h->built = true;
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// phases
// ─────────────────────────────────────────────────────────────────────────────

// Device geometry of one side of the operator: Gauss point tables (9 arrays of
// P), nodes [Nn][2] and per-element L | t_rho | t_z.  Test and source sides
// are the problem's own tables except in the adjoint's perturbed fills.
struct Geom {
  const double* pts;
  const double* nodes;
  const double* elem;
};

static PointSet point_set(const double* pts, int P) {
  return PointSet{pts, pts + P, pts + 2 * P, pts + 3 * P, pts + 4 * P, pts + 5 * P, pts + 6 * P, pts + 7 * P,
                  pts + 8 * P};
}

static Geom base_geom(const Problem& p) { return Geom{p.d_pts, p.d_nodes, p.d_elem}; }

template <typename R> static int far_fill(ghost_bor* h, Problem& p, const Geom& tg, const Geom& sg) {
  FarArgs fa;
  fa.pts = point_set(tg.pts, p.P);
  fa.spts = point_set(sg.pts, p.P);
  const bool split = tg.pts != sg.pts;   // only the adjoint's perturbed fills
  fa.sin2E = p.d_tabs + p.tab_off[0];
  fa.cosE = p.d_tabs + p.tab_off[1];
  fa.cxB = p.d_tabs + p.tab_off[2];
  fa.sxB = p.d_tabs + p.tab_off[3];
  fa.cosB = p.d_tabs + p.tab_off[4];
  fa.sinB = p.d_tabs + p.tab_off[5];
  fa.Ne = p.Ne;
  fa.Nn = p.Nn;
  fa.Neq = p.Ne;
  fa.Nnq = p.Nn;
  fa.skip = nullptr;
  fa.M = p.M;
  fa.nE = p.nE;
  fa.nB = p.nB;
  fa.k = p.k;
  // EFIE coefficient cE * j k eta 2 pi (complex for a lossy medium); far MFIE coefficient -2 pi cM
  const double ke_re = p.k * p.eta_re - p.ki * p.eta_im, ke_im = p.k * p.eta_im + p.ki * p.eta_re;
  fa.ce_re = -p.cE * ke_im * 2.0 * M_PI;
  fa.ce_im = p.cE * ke_re * 2.0 * M_PI;
  fa.ki = p.ki;
  const bool ck = p.ki != 0.0;
  if (ck && sizeof(R) != 8) return fail(GHOST_ERR_ARG, "complex k needs an FP64 fill");
  fa.cm = -2.0 * M_PI * p.cM;
  const bool ibc = p.has_zs && p.cI != 0.0;
  fa.zs = p.d_zs;
  fa.cz = 2.0 * M_PI * p.cI;
  if (ibc && split) return fail(GHOST_ERR_ARG, "perturbed (adjoint) fills of an impedance body are not supported");
  fa.Z = p.d_zfar;
  fa.Z2 = p.pen ? p.d_zfar2 : nullptr;
  fa.neg = h->pairing ? 0 : 1;
  fa.zmodes = p.zmodes;
  fa.zoff = p.zoff;
  const int NT = (p.M + 2 + 15) / 16;
  if (p.nE > 4096 || p.nB > 4096) return fail(GHOST_ERR_ARG, "azimuth grid above 4096 samples");
  for (int pe = 0; pe < 2; ++pe)
    for (int pf = 0; pf < 2; ++pf) {
      const int nep = (p.Ne - pe + 1) / 2, nfp = (p.Ne - pf + 1) / 2;
      if (nep <= 0 || nfp <= 0) continue;
      fa.pe = pe;
      fa.pf = pf;
      fa.nfp = nfp;
      if (split) {   // ordered pairs (e, f)
        fa.npairs = nep * nfp;
        fa.delta = 0;
        fa.imax = 0;
      } else {
        // unordered pairs e < f with f - e >= 2: each wavefront also fills
        // (f, e), from the same order samples for the EFIE part
        fa.delta = pf >= pe ? 1 : 2;
        fa.imax = std::min(nep, std::max(0, nfp - fa.delta));
        const long long base = nfp - fa.delta, im = fa.imax;
        fa.npairs = (int)(im * base - im * (im - 1) / 2);
        if (fa.npairs <= 0) continue;
      }
      const size_t tab_bytes = far_table_bytes((int)sizeof(R), p.nE, p.nB);
      const int wpb = far_fill_waves((int)sizeof(R), NT, p.nE, p.nB);
      if (wpb == 0)
        return fail(GHOST_ERR_ARG,
                    "far fill: azimuth tables (%zu B) and a one-wavefront order stash (%zu B) exceed the 160 KiB of LDS"
                    " (n_xi %d / %d, %d order tiles)",
                    tab_bytes, far_stash_bytes((int)sizeof(R), 1, NT), p.nE, p.nB, NT);
      const int grid = (fa.npairs + wpb - 1) / wpb;
      prof_begin(h);
      switch (NT) {
#define FARCASE(n)                                                                                      \
  case n:                                                                                               \
    launch_far_fill<R, n>(split, ibc, ck, wpb, grid, tab_bytes, h->stream, fa);                          \
    break;
        FARCASE(1) FARCASE(2) FARCASE(3) FARCASE(4) FARCASE(5) FARCASE(6) FARCASE(7) FARCASE(8)
#undef FARCASE
        default: return fail(GHOST_ERR_ARG, "m_max too large for the far fill");
      }
      {
        // MFMA flops issued (16x16x4: 2048 per wavefront instruction): per
        // element pair, (n_xi/8 + 1) sample groups x NT order tiles x 2 (re, im)
        // for the EFIE orders and the same per bracket orientation
        const double ge = p.nE / 8 + 1, gb = p.nB / 8 + 1, orient = split ? 1.0 : 2.0;
        prof_end(h, K_FAR_FILL, (double)fa.npairs * (ge + orient * gb) * NT * 2.0 * 2048.0);
      }
      LAUNCHCHK();
    }
  return 0;
}

static int rhs_problem(ghost_bor* h, Problem& p, const Geom& tg);

static ExciteArgs excite_args(const Problem& p, const Geom& tg) {
  ExciteArgs ea;
  ea.pts = point_set(tg.pts, p.P);
  ea.P = p.P;
  ea.Nn = p.Nn;
  ea.M = p.M;
  ea.ntheta = p.ntheta;
  ea.k = p.k;
  ea.cE = p.cE;
  ea.cM = p.cM;
  ea.theta = p.d_theta;
  ea.J = p.d_J;
  ea.E = p.d_E;
  ea.zs = (p.has_zs && !p.pen) ? p.d_zs : nullptr;   // a penetrable body's unit weights radiate nothing
  return ea;
}

// Operators of one problem's medium: far blocks (d_zfar, and d_zfar2 for a
// penetrable body's brackets) and the near band (d_NB, d_NB2).
static int fill_operators(ghost_bor* h, Problem& p, const Geom& tg, const Geom& sg) {
  const bool mixed = h->precision == GHOST_MIXED;
  const int P = p.P;
  if ((p.M + 2 + 7) / 8 > 16 || p.M + 1 > 128) return fail(GHOST_ERR_ARG, "m_max too large for the near band");
  const size_t zel = (size_t)(2 * p.Nn) * (2 * p.Nn) * p.zmodes;
  const size_t zero_n = mixed ? (zel + 1) / 2 : zel;  // in double2 units
  prof_begin(h);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (zero_n + 255) / 256)), dim3(256), 0,
                     h->stream, (double2*)p.d_zfar, zero_n);
  if (p.pen)
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (zero_n + 255) / 256)), dim3(256), 0,
                       h->stream, (double2*)p.d_zfar2, zero_n);
  prof_end(h, K_ZERO, 0.0);

  LAUNCHCHK();
  // near stage 1 + 2 (FP64) on stream2, beside the far fill (disjoint outputs;
  // joined before the assembly): both keep every SIMD busy on a large problem,
  // but a batch of small problems fills each other's launch tails; one stream
  // while the per-launch profile is on
  const bool split = !h->profile;
  hipStream_t ns = split ? h->stream2 : h->stream;
  if (split) {
    HIPCHK(hipEventRecord(h->ev_fork, h->stream));
    HIPCHK(hipStreamWaitEvent(ns, h->ev_fork, 0));
  }
  NearArgs na;
  na.nodes = tg.nodes;
  na.L = tg.elem;
  na.trho = tg.elem + p.Ne;
  na.tz = tg.elem + 2 * p.Ne;
  na.snodes = sg.nodes;
  na.sL = sg.elem;
  na.strho = sg.elem + p.Ne;
  na.stz = sg.elem + 2 * p.Ne;
  const int ncell = p.cell_off[2] + p.cell_cnt[2];
  na.cell_s = p.d_cells;
  na.cell_sp = p.d_cells + ncell;
  na.cell_w = p.d_cells + 2 * ncell;
  for (int i = 0; i < 3; ++i) { na.cell_off[i] = p.cell_off[i]; na.cell_cnt[i] = p.cell_cnt[i]; }
  na.core_u = p.d_core;
  na.core_w = p.d_core + 48;
  na.tail_u = p.d_tail;
  na.tail_w = p.d_tail + p.tail_w.size();
  na.pairs = p.d_pairs;
  na.batches = p.d_batches;
  na.Ne = p.Ne;
  na.M = p.M;
  na.nord = p.M + 2;
  na.ibc = (p.has_zs && p.cI != 0.0) ? 1 : 0;
  na.k = p.k;
  na.ki = p.ki;
  na.NP = p.d_NP;
  if (!p.batches.empty()) {
    const int C = (na.nord + 7) / 8;
    const dim3 g((unsigned)p.batches.size());
    prof_begin(h);
    switch (C) {
#define NEARCASE(c)                                                                  \
  case c:                                                                            \
    if (mixed) hipLaunchKernelGGL((k_near_modal<c, float>), g, dim3(256), 0, ns, na);  \
    else hipLaunchKernelGGL((k_near_modal<c, double>), g, dim3(256), 0, ns, na);      \
    break;
      NEARCASE(1) NEARCASE(2) NEARCASE(3) NEARCASE(4) NEARCASE(5) NEARCASE(6) NEARCASE(7) NEARCASE(8)
      NEARCASE(9) NEARCASE(10) NEARCASE(11) NEARCASE(12) NEARCASE(13) NEARCASE(14) NEARCASE(15) NEARCASE(16)
#undef NEARCASE
      default: return fail(GHOST_ERR_ARG, "m_max too large for the near quadrature");
    }
    {
      // projection flops: 10 real FMAs (G and four brackets, re and im) per
      // (point pair, azimuth node, order)
      double fl = 0.0;
      for (const auto& pi : p.pairs)
        if (pi.e >= 0) fl += (double)p.cell_cnt[pi.kind] * (48 + pi.n_tail) * na.nord * 20.0;
      prof_end(h, K_NEAR_MODAL, fl);
    }
    LAUNCHCHK();
  }
  NearBlockArgs nb;
  nb.na = na;
  nb.gw = tg.pts + 4 * P;
  nb.grho = tg.pts;
  nb.gT0 = tg.pts + 5 * P;
  nb.gT1 = tg.pts + 6 * P;
  {
    const double ke_re = p.k * p.eta_re - p.ki * p.eta_im, ke_im = p.k * p.eta_im + p.ki * p.eta_re;
    nb.ce_re = -p.cE * ke_im * 2.0 * M_PI;
    nb.ce_im = p.cE * ke_re * 2.0 * M_PI;
  }
  nb.cM = p.cM;
  nb.zs = p.d_zs;
  nb.cI = p.cI;
  nb.cMass = p.cMass;
  nb.NB = p.d_NB;
  nb.NB2 = p.pen ? p.d_NB2 : nullptr;
  const int mlanes = ((p.M + 1 + 15) / 16) * 16;
  if (mlanes > 128) return fail(GHOST_ERR_ARG, "m_max too large for the near blocks");
  prof_begin(h);
  hipLaunchKernelGGL(k_near_blocks, dim3(3 * p.Ne), dim3(256), 0, ns, nb, mlanes);
  prof_end(h, K_NEAR_BLOCKS, 0.0);
  LAUNCHCHK();
  // far
  int rc = mixed ? far_fill<float>(h, p, tg, sg) : far_fill<double>(h, p, tg, sg);
  if (split) {
    HIPCHK(hipEventRecord(h->ev_join, ns));
    HIPCHK(hipStreamWaitEvent(h->stream, h->ev_join, 0));
  }
  return rc;
}

// PMCHWT systems of a penetrable problem from its exterior and interior
// operators (both filled).
static int assemble_pen(ghost_bor* h, Problem& p) {
  const Problem& q = h->inner[p.inner];
  AsmPenArgs aa;
  aa.Te = (const double2*)p.d_zfar;
  aa.Be = (const double2*)p.d_zfar2;
  aa.Ti = (const double2*)q.d_zfar;
  aa.Bi = (const double2*)q.d_zfar2;
  aa.NTe = p.d_NB;
  aa.NBe = p.d_NB2;
  aa.NTi = q.d_NB;
  aa.NBi = q.d_NB2;
  aa.sys = h->d_sys + p.sys0;
  aa.Nn = p.Nn;
  aa.Ne = p.Ne;
  aa.M = p.M;
  aa.mlo = h->pairing ? p.M : 0;
  if (q.zmodes != p.zmodes || q.zoff != p.zoff) return fail(GHOST_ERR_STATE, "interior far blocks differ in layout");
  aa.zmodes = p.zmodes;
  aa.zoff = p.zoff;
  aa.eta0 = 376.730313668;
  {
    // (eta0 / eta_i)^2
    const double n2 = q.eta_re * q.eta_re + q.eta_im * q.eta_im;
    const double ir = aa.eta0 * q.eta_re / n2, ii = -aa.eta0 * q.eta_im / n2;
    aa.r2_re = ir * ir - ii * ii;
    aa.r2_im = 2.0 * ir * ii;
  }
  aa.A = (double2*)h->d_A;
  aa.A64 = h->d_A64;
  dim3 g(2 * p.Nn, (2 * p.Nn + 31) / 32, (p.nmodes - aa.mlo + 31) / 32);
  prof_begin(h);
  hipLaunchKernelGGL(k_assemble_pen, g, dim3(256), 0, h->stream, aa);
  prof_end(h, K_ASSEMBLE, 0.0);
  LAUNCHCHK();
  return 0;
}

static int fill_problem(ghost_bor* h, Problem& p, const Geom& tg, const Geom& sg) {
  const bool mixed = h->precision == GHOST_MIXED;
  int rc;
  if (p.pen) {
    if (p.inner < 0 || mixed) return fail(GHOST_ERR_STATE, "penetrable problem without its FP64 interior");
    Problem& q = h->inner[p.inner];
    if ((rc = fill_operators(h, q, base_geom(q), base_geom(q)))) return rc;
    if ((rc = fill_operators(h, p, tg, sg))) return rc;
    if ((rc = assemble_pen(h, p))) return rc;
    return rhs_problem(h, p, tg);
  }
  if ((rc = fill_operators(h, p, tg, sg))) return rc;

  // assemble
  AsmArgs aa;
  aa.Zfar = p.d_zfar;
  aa.NB = p.d_NB;
  aa.sys = h->d_sys + p.sys0;
  aa.Nn = p.Nn;
  aa.Ne = p.Ne;
  aa.M = p.M;
  aa.A = h->d_A;
  aa.A64 = h->d_A64;
  aa.mlo = h->pairing ? p.M : 0;
  aa.zmodes = p.zmodes;
  aa.zoff = p.zoff;
  dim3 g(2 * p.Nn, (2 * p.Nn + 31) / 32, (p.nmodes - aa.mlo + 31) / 32);
  prof_begin(h);
  if (mixed) hipLaunchKernelGGL((k_assemble<float, float>), g, dim3(256), 0, h->stream, aa);
  else hipLaunchKernelGGL((k_assemble<double, double>), g, dim3(256), 0, h->stream, aa);
  prof_end(h, K_ASSEMBLE, 0.0);
  LAUNCHCHK();
  return rhs_problem(h, p, tg);
}

// Bessel table and right-hand sides of one problem (test geometry tg).
static int rhs_problem(ghost_bor* h, Problem& p, const Geom& tg) {
  const bool mixed = h->precision == GHOST_MIXED;
  const int P = p.P;
  ExciteArgs ea = excite_args(p, tg);
  const size_t nbes = (size_t)P * p.ntheta;
  prof_begin(h);
  hipLaunchKernelGGL(k_bessel, dim3((unsigned)std::min<size_t>(65535, (nbes + 127) / 128)), dim3(128), 0,
                     h->stream, ea);
  prof_end(h, K_BESSEL, 0.0);

  LAUNCHCHK();
  // one thread per (node, aspect) and chunk of 8 consecutive modes
  const int mlo = h->pairing ? p.M : 0;   // paired: modes -m share mode m's right-hand sides
  const int nsys = p.nmodes - mlo;
  constexpr int MC = 8;
  dim3 gr((p.ntheta + 127) / 128, p.Nn, (nsys + MC - 1) / MC);
  const SysInfo* dsys = h->d_sys + p.sys0 + mlo;
  prof_begin(h);
  // mixed: B in FP64 only; the solve writes the column-scaled FP32 copy into the augmented columns
  if (mixed) hipLaunchKernelGGL((k_rhs<double, MC>), gr, dim3(128), 0, h->stream, ea, dsys, nsys, nullptr, h->d_B64);
  else hipLaunchKernelGGL((k_rhs<double, MC>), gr, dim3(128), 0, h->stream, ea, dsys, nsys, h->d_A, h->d_B64);
  prof_end(h, K_RHS, 0.0);
  LAUNCHCHK();
  return 0;
}

// Cross-surface operator: near pairs (modal values, then blocks), far pairs
// on the split far fill with a skip mask, then the near blocks added into
// the far blocks by parity class.
static int cross_fill(ghost_bor* h, Cross& c) {
  const Problem& a = h->inner[c.op_p];
  const Problem& b = h->inner[c.op_q];
  const size_t zel = (size_t)(2 * c.Np) * (2 * c.Nq) * c.zmodes;
  prof_begin(h);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (zel + 255) / 256)), dim3(256), 0, h->stream,
                     c.d_zfar, zel);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (zel + 255) / 256)), dim3(256), 0, h->stream,
                     c.d_zfar2, zel);
  prof_end(h, K_ZERO, 0.0);
  LAUNCHCHK();
  NearArgs na;
  memset(&na, 0, sizeof(na));
  na.nodes = a.d_nodes; na.L = a.d_elem; na.trho = a.d_elem + a.Ne; na.tz = a.d_elem + 2 * a.Ne;
  na.snodes = b.d_nodes; na.sL = b.d_elem; na.strho = b.d_elem + b.Ne; na.stz = b.d_elem + 2 * b.Ne;
  const size_t ncell = c.cells.size() / 3;
  na.cell_s = c.d_cells; na.cell_sp = c.d_cells + ncell; na.cell_w = c.d_cells + 2 * ncell;
  for (int i = 0; i < 8; ++i) { na.cell_off[i] = c.cell_off[i]; na.cell_cnt[i] = c.cell_cnt[i]; }
  na.core_u = c.d_core; na.core_w = c.d_core + 48;
  na.tail_u = c.d_tail; na.tail_w = c.d_tail + c.tail_w.size();
  na.pairs = c.d_pairs; na.batches = c.d_batches;
  na.Ne = c.Nep; na.M = c.M; na.nord = c.M + 2; na.ibc = 1; na.k = c.k; na.ki = c.ki; na.NP = c.d_NP;
  if (!c.pairs.empty()) {
    const int C = (na.nord + 7) / 8;
    const dim3 g((unsigned)c.batches.size());
    prof_begin(h);
    switch (C) {
#define NEARCASE(cc) \
  case cc: hipLaunchKernelGGL((k_near_modal<cc, double>), g, dim3(256), 0, h->stream, na); break;
      NEARCASE(1) NEARCASE(2) NEARCASE(3) NEARCASE(4) NEARCASE(5) NEARCASE(6) NEARCASE(7) NEARCASE(8)
      NEARCASE(9) NEARCASE(10) NEARCASE(11) NEARCASE(12) NEARCASE(13) NEARCASE(14) NEARCASE(15) NEARCASE(16)
#undef NEARCASE
      default: return fail(GHOST_ERR_ARG, "m_max too large for the near quadrature");
    }
    prof_end(h, K_NEAR_MODAL, 0.0);
    LAUNCHCHK();
    NearBlockArgs nb;
    memset(&nb, 0, sizeof(nb));
    nb.na = na;
    nb.gw = a.d_pts + 4 * a.P; nb.grho = a.d_pts; nb.gT0 = a.d_pts + 5 * a.P; nb.gT1 = a.d_pts + 6 * a.P;
    const double ke_re = c.k * c.eta_re - c.ki * c.eta_im, ke_im = c.k * c.eta_im + c.ki * c.eta_re;
    nb.ce_re = -ke_im * 2.0 * M_PI;
    nb.ce_im = ke_re * 2.0 * M_PI;
    nb.cM = 0.0; nb.zs = c.d_zs; nb.cI = 1.0; nb.cMass = 0.0;
    nb.NB = c.d_NB; nb.NB2 = c.d_NB2;
    const int mlanes = ((c.M + 1 + 15) / 16) * 16;
    if (mlanes > 128) return fail(GHOST_ERR_ARG, "m_max too large for the near blocks");
    prof_begin(h);
    hipLaunchKernelGGL(k_near_blocks, dim3((unsigned)c.pairs.size()), dim3(256), 0, h->stream, nb, mlanes);
    prof_end(h, K_NEAR_BLOCKS, 0.0);
    LAUNCHCHK();
  }
  FarArgs fa;
  memset(&fa, 0, sizeof(fa));
  fa.pts = point_set(a.d_pts, a.P);
  fa.spts = point_set(b.d_pts, b.P);
  fa.sin2E = c.d_tabs + c.tab_off[0]; fa.cosE = c.d_tabs + c.tab_off[1];
  fa.cxB = c.d_tabs + c.tab_off[2]; fa.sxB = c.d_tabs + c.tab_off[3];
  fa.cosB = c.d_tabs + c.tab_off[4]; fa.sinB = c.d_tabs + c.tab_off[5];
  fa.Ne = c.Nep; fa.Nn = c.Np; fa.Neq = c.Neq; fa.Nnq = c.Nq; fa.skip = c.d_skip;
  fa.M = c.M; fa.nE = c.nE; fa.nB = c.nB; fa.k = c.k; fa.ki = c.ki;
  {
    const double ke_re = c.k * c.eta_re - c.ki * c.eta_im, ke_im = c.k * c.eta_im + c.ki * c.eta_re;
    fa.ce_re = -ke_im * 2.0 * M_PI;
    fa.ce_im = ke_re * 2.0 * M_PI;
  }
  fa.cm = 0.0; fa.zs = c.d_zs; fa.cz = 2.0 * M_PI;
  fa.Z = c.d_zfar; fa.Z2 = c.d_zfar2;
  fa.neg = h->pairing ? 0 : 1; fa.zmodes = c.zmodes; fa.zoff = c.zoff;
  const int NT = (c.M + 2 + 15) / 16;
  const size_t tab_bytes = far_table_bytes(8, c.nE, c.nB);
  const int wpb = far_fill_waves(8, NT, c.nE, c.nB);
  if (wpb == 0) return fail(GHOST_ERR_ARG, "cross far fill exceeds the LDS");
  for (int pe = 0; pe < 2; ++pe)
    for (int pf = 0; pf < 2; ++pf) {
      const int nep = (c.Nep - pe + 1) / 2, nfp = (c.Neq - pf + 1) / 2;
      if (nep <= 0 || nfp <= 0) continue;
      fa.pe = pe; fa.pf = pf; fa.nfp = nfp; fa.npairs = nep * nfp; fa.delta = 0; fa.imax = 0;
      const int grid = (fa.npairs + wpb - 1) / wpb;
      prof_begin(h);
      switch (NT) {
#define FARCASE(n) \
  case n: launch_far_fill<double, n>(true, true, true, wpb, grid, tab_bytes, h->stream, fa); break;
        FARCASE(1) FARCASE(2) FARCASE(3) FARCASE(4) FARCASE(5) FARCASE(6) FARCASE(7) FARCASE(8)
#undef FARCASE
        default: return fail(GHOST_ERR_ARG, "m_max too large for the far fill");
      }
      prof_end(h, K_FAR_FILL, 0.0);
      LAUNCHCHK();
    }
  const int mlo = h->pairing ? c.M : 0;
  for (int cls = 0; cls < 4; ++cls) {
    const int n = c.par_off[cls + 1] - c.par_off[cls];
    if (n == 0) continue;
    hipLaunchKernelGGL(k_cross_near_add, dim3(n), dim3(256), 0, h->stream, c.d_NB, c.d_NB2, c.d_pairs,
                       c.d_plist + c.par_off[cls], c.M, c.Np, c.Nq, c.zmodes, c.zoff, mlo, c.d_zfar, c.d_zfar2);
    LAUNCHCHK();
  }
  return 0;
}

// Region set: the full systems from their blocks, reduced by the junction
// map into the augmented systems, and the right-hand sides of the exterior
// surfaces, reduced the same way.
static int region_fill(ghost_bor* h, Region& r) {
  const size_t nf = (size_t)r.n_full;
  const size_t nA = (size_t)r.nmodes * nf * nf, nB = (size_t)r.nmodes * nf * r.nrhs;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (nA + 255) / 256)), dim3(256), 0, h->stream,
                     r.d_full, nA);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (nB + 255) / 256)), dim3(256), 0, h->stream,
                     r.d_bfull, nB);
  LAUNCHCHK();
  for (const auto& bk : r.blocks) {
    RegionBlockArgs ra;
    memset(&ra, 0, sizeof(ra));
    if (bk.src == 0) {
      const Problem& o = h->inner[bk.idx];
      ra.Z = (const double2*)(bk.which ? o.d_zfar2 : o.d_zfar);
      ra.NB = bk.which ? o.d_NB2 : o.d_NB;
      ra.Np = ra.Nq = o.Nn; ra.Ne = o.Ne; ra.zmodes = o.zmodes; ra.zoff = o.zoff;
    } else {
      const Cross& c = h->cross[bk.idx];
      ra.Z = bk.which ? c.d_zfar2 : c.d_zfar;
      ra.NB = nullptr;
      ra.Np = c.Np; ra.Nq = c.Nq; ra.Ne = c.Nep; ra.zmodes = c.zmodes; ra.zoff = c.zoff;
    }
    ra.M = r.M; ra.paired = h->pairing; ra.remap = bk.which;
    ra.cre = bk.cre; ra.cim = bk.cim; ra.row_off = bk.row_off; ra.col_off = bk.col_off; ra.n_full = r.n_full;
    ra.A = r.d_full;
    dim3 g(2 * ra.Np, (2 * ra.Nq + 31) / 32, (r.nmodes + 31) / 32);
    prof_begin(h);
    hipLaunchKernelGGL(k_region_block, g, dim3(256), 0, h->stream, ra);
    prof_end(h, K_ASSEMBLE, 0.0);
    LAUNCHCHK();
  }
  int nmax = 0;
  for (int mi = 0; mi < r.nmodes; ++mi) nmax = std::max(nmax, h->sys[r.sys0 + mi].n);
  RegionMap q{r.d_ptr, r.d_mem, r.d_cf, r.n_full};
  prof_begin(h);
  hipLaunchKernelGGL(k_region_reduce, dim3(nmax, (nmax + 255) / 256, r.nmodes), dim3(256), 0, h->stream, q,
                     r.d_full, h->d_sys + r.sys0, (double2*)h->d_A, h->d_A64);
  prof_end(h, K_ASSEMBLE, 0.0);
  LAUNCHCHK();
  constexpr int MC = 8;
  for (size_t e = 0; e < r.ext.size(); ++e) {
    const Problem& o = h->inner[r.ext[e].op];
    ExciteArgs ea = excite_args(o, base_geom(o));
    dim3 gr((o.ntheta + 127) / 128, o.Nn, (r.nmodes + MC - 1) / MC);
    prof_begin(h);
    hipLaunchKernelGGL((k_rhs<double, MC>), gr, dim3(128), 0, h->stream, ea, r.d_vsys[e], r.nmodes, nullptr,
                       r.d_bfull);
    prof_end(h, K_RHS, 0.0);
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(k_region_rhs, dim3(nmax, (r.nrhs + 255) / 256, r.nmodes), dim3(256), 0, h->stream, q,
                     r.d_bfull, h->d_sys + r.sys0, (double2*)h->d_A, h->d_B64);
  LAUNCHCHK();
  return 0;
}

static int run_fill(ghost_bor* h) {
  int rc;
  for (auto& o : h->inner) {
    if (!o.standalone) continue;   // interior twins fill with their exterior problem
    if ((rc = fill_operators(h, o, base_geom(o), base_geom(o)))) return rc;
    if (o.ntheta > 0) {   // an exterior surface: the Bessel and phase tables of its aspects
      ExciteArgs ea = excite_args(o, base_geom(o));
      const size_t nbes = (size_t)o.P * o.ntheta;
      prof_begin(h);
      hipLaunchKernelGGL(k_bessel, dim3((unsigned)std::min<size_t>(65535, (nbes + 127) / 128)), dim3(128), 0,
                         h->stream, ea);
      prof_end(h, K_BESSEL, 0.0);
      LAUNCHCHK();
    }
  }
  for (auto& c : h->cross)
    if ((rc = cross_fill(h, c))) return rc;
  for (auto& p : h->probs)
    if ((rc = fill_problem(h, p, base_geom(p), base_geom(p)))) return rc;
  for (auto& r : h->regions)
    if ((rc = region_fill(h, r))) return rc;
  return 0;
}

// Only the right-hand sides of every problem (Bessel table + excitation), for
// callers that want V_m(theta) without the operator.
extern "C" int ghost_bor_fill_rhs(ghost_bor* h) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  h->solved = false;
  for (auto& p : h->probs) {
    int rc = rhs_problem(h, p, base_geom(p));
    if (rc) return rc;
  }
  return 0;
}

// Write a caller's solution X (n x n_rhs, masked rows, complex row-major) of
// mode m of problem prob where the far field reads it.
extern "C" int ghost_bor_set_solution(ghost_bor* h, int prob, int m, const double* X) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (!X) return fail(GHOST_ERR_ARG, "null solution");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index");
  const Problem& p = h->probs[prob];
  if (m < -p.M || m > p.M) return fail(GHOST_ERR_ARG, "mode outside range");
  const SysInfo& s = h->sys[p.sys0 + m + p.M];
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double2> row(s.nrhs);
  for (int r = 0; r < s.n; ++r) {
    for (int c = 0; c < s.nrhs; ++c) {
      // mode -m of a pair shares mode m's storage: x_m = S [x_{-m}(VV), -x_{-m}(HH)]
      const double f = s.flip ? pair_row_sign(s, r) * ((c & 1) ? -1.0 : 1.0) : 1.0;
      row[c] = make_double2(f * X[2 * ((size_t)r * s.nrhs + c)], f * X[2 * ((size_t)r * s.nrhs + c) + 1]);
    }
    double2* dst = h->precision == GHOST_MIXED ? h->d_X64 + s.x_off + (size_t)r * s.ncols + s.col
                                               : (double2*)h->d_A + (size_t)s.a_off + (size_t)r * s.ld + s.n + s.col;
    HIPCHK(hipMemcpy(dst, row.data(), (size_t)s.nrhs * sizeof(double2), hipMemcpyHostToDevice));
  }
  return 0;
}

// Batched complex GEMM C -= A B with the handle's tile variant:
// 0: 64 x 64 single-buffered; 1: 64 x 64, 2: 64 x 128, 3: 128 x 128 (FP32) /
// 64 x 128 (FP64), double-buffered; 4: variant 1 with the C tile loaded
// before the K loop; 5 (FP32 only; FP64 runs 4): 64 x 64 on the bf16 MFMA
// with three-way split operands (k_gemm_bf3).
template <typename T, bool TA = false, bool TB = false>
static void gemm_launch(ghost_bor* h, const GemmOp* ops, int nops, const int* tt, const void* A, const void* B,
                        void* C) {
  int v = h->gemm_variant;
  if (v < 0) v = sizeof(T) == 8 ? 4 : 5;
  if (v == 3 && sizeof(T) == 8) v = 2;
  if (v == 5 && sizeof(T) == 8) v = 4;
  const int sh = (v == 2) ? 1 : (v == 3 ? 2 : 0);
  if (tt[sh] <= 0) return;
  const dim3 grid((unsigned)tt[sh]);
  if constexpr (sizeof(T) == 4) {
    if (v == 5) {
      hipLaunchKernelGGL((k_gemm_bf3<TA, TB>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C);
      return;
    }
  }
  switch (v) {
    case 0: hipLaunchKernelGGL((k_gemm_update<T, TA, TB>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C); break;
    case 1: hipLaunchKernelGGL((k_gemm2<T, 2, 2, TA, TB>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C); break;
    case 2: hipLaunchKernelGGL((k_gemm2<T, 2, 4, TA, TB>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C); break;
    case 4: hipLaunchKernelGGL((k_gemm2<T, 2, 2, TA, TB, true>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C); break;
    default: hipLaunchKernelGGL((k_gemm2<T, 4, 4, TA, TB>), grid, dim3(256), 0, h->stream, ops, nops, A, B, C); break;
  }
}

template <typename T> static int run_sched(ghost_bor* h, const std::vector<Launch>& sched, void* factor,
                                           void* rhs) {
  for (const Launch& L : sched) {
    prof_begin(h);
    switch (L.kind) {
      case L_PANEL:
        // register sub-panel: 8 columns, 4 when FP64 rows need 2+ registers sets
        switch (L.arg) {
          case 1: hipLaunchKernelGGL((k_lu_panel<T, 1, 8>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          case 2: hipLaunchKernelGGL((k_lu_panel<T, 2, sizeof(T) == 8 ? 4 : 8>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          case 3: hipLaunchKernelGGL((k_lu_panel<T, 3, sizeof(T) == 8 ? 4 : 8>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          case 4: hipLaunchKernelGGL((k_lu_panel<T, 4, 4>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          // 4097 .. 8192 rows: 5 .. 8 rows per thread with 4-column (FP32) or
          // 2-column (FP64) register sub-panels, within 1024 threads' registers
          case 5: hipLaunchKernelGGL((k_lu_panel<T, 5, sizeof(T) == 8 ? 2 : 4>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          case 6: hipLaunchKernelGGL((k_lu_panel<T, 6, sizeof(T) == 8 ? 2 : 4>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          case 7: hipLaunchKernelGGL((k_lu_panel<T, 7, sizeof(T) == 8 ? 2 : 4>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
          default: hipLaunchKernelGGL((k_lu_panel<T, 8, sizeof(T) == 8 ? 2 : 4>), dim3(L.nops), dim3(1024), 0, h->stream, h->d_panel + L.ops, factor, h->d_piv, h->d_info); break;
        }
        break;
      case L_SWAP:
        if (L.gx > 0) hipLaunchKernelGGL((k_lu_swap<T>), dim3(L.gx, L.nops), dim3(256), 0, h->stream, h->d_swap + L.ops, factor, h->d_piv);
        break;
      case L_TRIL:
        hipLaunchKernelGGL((k_tri_inv<T, false>), dim3(L.nops), dim3(256), 0, h->stream, h->d_trsm + L.ops, factor, h->d_W);
        break;
      case L_GEMM_AUG_W:
        gemm_launch<T>(h, h->d_gemm + L.ops, L.nops, L.tt, h->d_W, factor, factor);
        break;
      case L_GEMM_AUG:
        gemm_launch<T>(h, h->d_gemm + L.ops, L.nops, L.tt, factor, factor, factor);
        break;
      case L_GEMM_Y_W:
        gemm_launch<T>(h, h->d_gemm + L.ops, L.nops, L.tt, h->d_W, rhs, rhs);
        break;
      case L_GEMM_Y_LU:
        gemm_launch<T>(h, h->d_gemm + L.ops, L.nops, L.tt, factor, rhs, rhs);
        break;
      case L_GEMM_TW:
        gemm_launch<T, true, false>(h, h->d_gemm + L.ops, L.nops, L.tt, h->d_W, rhs, rhs);
        break;
      case L_GEMM_TLU:
        gemm_launch<T, true, false>(h, h->d_gemm + L.ops, L.nops, L.tt, factor, rhs, rhs);
        break;
      default: break;
    }
    {
      int kind = L.fam >= 0 ? L.fam : K_GEMM;
      if (L.kind == L_PANEL) kind = K_PANEL;
      else if (L.kind == L_SWAP) kind = K_SWAP;
      else if (L.kind == L_TRIL || L.kind == L_TRIU) kind = K_TRI_INV;
      prof_end(h, kind, L.flops);
    }
    LAUNCHCHK();
  }
  return 0;
}

template <typename T> static int run_lu(ghost_bor* h) {
  const int nsys = (int)h->ss.size();
  hipLaunchKernelGGL(k_zero_int, dim3((nsys + 255) / 256), dim3(256), 0, h->stream, h->d_info, nsys);
  LAUNCHCHK();
  // the second part on stream2 between a fork and a join event; while the
  // per-launch profile is on (its events time one stream) the whole batch
  // runs as schedule 2 on one stream
  const bool split = h->nparts > 1 && h->lu_streams > 1 && !h->profile;
  hipStream_t s0 = h->stream;
  if (split) {
    HIPCHK(hipEventRecord(h->ev_fork, s0));
    HIPCHK(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
  }
  int rc = 0;
  const int p_lo = split ? 0 : 2, p_hi = split ? 2 : 3;
  for (int part = p_lo; part < p_hi && !rc; ++part) {
    h->stream = (split && part == 1) ? h->stream2 : s0;
    if ((rc = run_sched<T>(h, h->lu_sched[part], h->d_A, nullptr))) break;
    if (h->triu_n[part] > 0) {
      prof_begin(h);
      hipLaunchKernelGGL((k_tri_inv<T, true>), dim3(h->triu_n[part]), dim3(256), 0, h->stream,
                         h->d_trsm + h->triu_ops[part], h->d_A, h->d_W);
      prof_end(h, K_TRI_INV, 0.0);
      if (hipGetLastError() != hipSuccess) { rc = fail(GHOST_ERR_HIP, "k_tri_inv launch failed"); break; }
    }
    rc = run_sched<T>(h, h->back_sched[part], h->d_A, nullptr);
  }
  h->stream = s0;
  if (rc) return rc;
  if (split) {
    HIPCHK(hipEventRecord(h->ev_join, h->stream2));
    HIPCHK(hipStreamWaitEvent(s0, h->ev_join, 0));
  }
  return 0;
}

// R64 = B64 - A64 X with X in X64 (x_in_aug = 0) or in the FP64 augmented columns
static int residual_gemm(ghost_bor* h, int x_in_aug) {
  const int nsys = (int)h->ss.size();
  size_t maxn = 0, maxr = 0;
  int gx = 0;
  for (auto& s : h->ss) {
    maxn = std::max(maxn, (size_t)s.n);
    maxr = std::max(maxr, (size_t)s.nrhs);
    gx = std::max(gx, ((s.n + 63) / 64) * ((s.nrhs + 63) / 64));
  }
  dim3 gv((unsigned)((maxr + 255) / 256), (unsigned)maxn, nsys);
  hipLaunchKernelGGL(k_copy64, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_B64, h->d_R64);
  LAUNCHCHK();
  prof_begin(h);
  (void)gx;
  if (x_in_aug)
    gemm_launch<double>(h, h->d_gemm + h->res_gemm_aug, nsys, h->res_aug_tt, h->d_A64, h->d_A, h->d_R64);
  else
    gemm_launch<double>(h, h->d_gemm + h->res_gemm, nsys, h->res_tt, h->d_A64, h->d_X64, h->d_R64);
  prof_end(h, K_GEMM64, gemm_flops(h->h_gemm, x_in_aug ? h->res_gemm_aug : h->res_gemm, nsys, false));
  LAUNCHCHK();
  return 0;
}

extern "C" int ghost_bor_fill(ghost_bor* h) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  h->solved = false;
  if (!h->capturing) HIPCHK(hipEventRecord(h->ev[0], h->stream));
  int rc = run_fill(h);
  if (rc) return rc;
  if (!h->capturing) {
    HIPCHK(hipEventRecord(h->ev[1], h->stream));
    h->ev_recorded |= 1u;
  }
  return 0;
}

static int frob2(ghost_bor* h, const double2* buf, std::vector<double>& out) {
  // sum |z|^2 per mode entry over its n x nrhs columns of the block laid out by x_off
  const int nsys = (int)h->sys.size();
  out.assign(nsys, 0.0);
  std::vector<double2> host(h->x_elems);
  HIPCHK(hipMemcpyAsync(host.data(), buf, h->x_elems * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int si = 0; si < nsys; ++si) {
    const SysInfo& s = h->sys[si];
    double acc = 0;
    for (int r = 0; r < s.n; ++r)
      for (int c = 0; c < s.nrhs; ++c) {
        const double2 z = host[s.x_off + (size_t)r * s.ncols + s.col + c];
        acc += z.x * z.x + z.y * z.y;
      }
    out[si] = acc;
  }
  return 0;
}

// Error probe of the current X64 (mixed precision only): with S two fixed
// random column combinations, P = B S - A64 X S in FP64 and E = A32^-1 P with
// the FP32 factors, eta[s] = ||E|| / ||X S|| estimates the relative error of
// system s's currents (the correction the next refinement sweep would make).
static int probe_error(ghost_bor* h, std::vector<double>& eta) {
  const int nsys = (int)h->ss.size();
  size_t pe = 0;
  for (auto& s : h->ss) pe += 2 * (size_t)s.n;
  int maxn = 0;
  for (auto& s : h->ss) maxn = std::max(maxn, s.n);
  const dim3 gz((unsigned)std::min<size_t>(65535, (pe + 255) / 256));
  hipLaunchKernelGGL(k_zero, gz, dim3(256), 0, h->stream, h->d_BS, pe);
  hipLaunchKernelGGL(k_zero, gz, dim3(256), 0, h->stream, h->d_XS, pe);
  LAUNCHCHK();
  prof_begin(h);
  // S is stored negated, so C -= A (-S) accumulates +A S
  const dim3 gp((unsigned)((maxn + 3) / 4), nsys);
  hipLaunchKernelGGL(k_gemm_n2, gp, dim3(256), 0, h->stream, h->d_gemm + h->probe_ops, h->d_B64, h->d_S, h->d_BS);
  hipLaunchKernelGGL(k_gemm_n2, gp, dim3(256), 0, h->stream, h->d_gemm + h->probe_ops + nsys, h->d_X64, h->d_S,
                     h->d_XS);
  hipLaunchKernelGGL(k_gemm_n2, gp, dim3(256), 0, h->stream, h->d_gemm + h->probe_ops + 2 * nsys, h->d_A64,
                     h->d_XS, h->d_BS);
  prof_end(h, K_GEMM64, 0.0);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_col_scale, dim3(1, 1, nsys), dim3(64, 4), 0, h->stream, h->d_vec + h->vec_probe,
                     h->d_BS, h->d_pscale);
  hipLaunchKernelGGL(k_scale_cast, dim3(1, maxn, nsys), dim3(64), 0, h->stream, h->d_vec + h->vec_probe,
                     h->d_perm, h->d_BS, h->d_pscale, h->d_P32);
  LAUNCHCHK();
  int rc;
  prof_begin(h);
  hipLaunchKernelGGL(k_trsv2<false>, dim3(nsys), dim3(1024), 0, h->stream, h->d_trsv, (const float2*)h->d_A,
                     (const float2*)h->d_W, h->d_P32);
  hipLaunchKernelGGL(k_trsv2<true>, dim3(nsys), dim3(1024), 0, h->stream, h->d_trsv, (const float2*)h->d_A,
                     (const float2*)h->d_W, h->d_P32);
  prof_end(h, K_TRSV, 0.0);
  LAUNCHCHK();
  // ratios on the device: only nsys doubles come back to decide the sweep
  hipLaunchKernelGGL(k_probe_eta, dim3(nsys), dim3(256), 0, h->stream, h->d_vec + h->vec_probe, h->d_P32, h->d_XS,
                     h->d_pscale, h->d_red);
  LAUNCHCHK();
  eta.assign(nsys, 0.0);
  HIPCHK(hipMemcpyAsync(eta.data(), h->d_red, nsys * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int ghost_bor_solve(ghost_bor* h, int refine_iters) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  const int nsys = (int)h->ss.size();
  const bool mixed = h->precision == GHOST_MIXED;
  if (!h->capturing) HIPCHK(hipEventRecord(h->ev[2], h->stream));
  int rc;
  if (!mixed) {
    if ((rc = run_lu<double>(h))) return rc;
  } else {
    size_t maxn = 0, maxr = 0;
    for (auto& s : h->ss) { maxn = std::max(maxn, (size_t)s.n); maxr = std::max(maxr, (size_t)s.nrhs); }
    dim3 gv((unsigned)((maxr + 255) / 256), (unsigned)maxn, nsys);
    dim3 gs((unsigned)((maxr + 63) / 64), 1, nsys);
    // FP32 right-hand sides = B64 with power-of-two column scales (no underflow
    // for the high orders), into the augmented columns of the factor
    hipLaunchKernelGGL(k_col_scale, gs, dim3(64, 4), 0, h->stream, h->d_vec + h->vec_all, h->d_B64, h->d_bscale);
    hipLaunchKernelGGL(k_scale_cast, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_toaug, nullptr, h->d_B64,
                       h->d_bscale, (float2*)h->d_A);
    LAUNCHCHK();
    if ((rc = run_lu<float>(h))) return rc;
    // X64 = X32 / scale from the augmented columns
    hipLaunchKernelGGL(k_widen_scaled, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_aug, (const float2*)h->d_A,
                       h->d_bscale, h->d_X64, 0);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_perm_from_pivots, dim3(nsys), dim3(256), 0, h->stream, h->d_vec + h->vec_all,
                       nsys, h->d_piv, h->d_perm);
    LAUNCHCHK();
    // refine_iters >= 0: that many FP64-residual sweeps; < 0: adaptive, up to
    // -refine_iters sweeps while the error probe exceeds h->probe_tol
    const bool adaptive = refine_iters < 0;
    const int max_it = adaptive ? -refine_iters : refine_iters;
    h->probe_hist.clear();
    if (adaptive && h->capturing) return fail(GHOST_ERR_STATE, "adaptive refinement cannot be captured");
    for (int it = 0; it < max_it; ++it) {
      if (adaptive) {
        std::vector<double> eta;
        if ((rc = probe_error(h, eta))) return rc;
        const double mx = *std::max_element(eta.begin(), eta.end());
        h->probe_hist.push_back(mx);
        if (mx <= h->probe_tol) break;
      }
      if ((rc = residual_gemm(h, 0))) return rc;      // R = B - A X in FP64
      hipLaunchKernelGGL(k_col_scale, gs, dim3(64, 4), 0, h->stream, h->d_vec + h->vec_all, h->d_R64, h->d_rscale);
      hipLaunchKernelGGL(k_scale_cast, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_perm, h->d_R64,
                         h->d_rscale, h->d_Y32);
      LAUNCHCHK();
      if ((rc = run_sched<float>(h, h->fwdY_sched, h->d_A, h->d_Y32))) return rc;
      if ((rc = run_sched<float>(h, h->backY_sched, h->d_A, h->d_Y32))) return rc;
      hipLaunchKernelGGL(k_widen_scaled, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_Y32, h->d_rscale,
                         h->d_X64, 1);
      LAUNCHCHK();
    }
  }
  if (!h->capturing) {
    HIPCHK(hipEventRecord(h->ev[3], h->stream));
    h->ev_recorded |= 2u;
  }
  h->solved = true;
  h->refine_iters = refine_iters;
  return 0;
}

extern "C" int ghost_bor_set_lu_super(ghost_bor* h, int panels) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (h->built) return fail(GHOST_ERR_STATE, "set the LU super-block before build");
  if (panels < 1 || panels > 16) return fail(GHOST_ERR_ARG, "LU super-block must be 1..16 panels");
  h->lu_super = panels;
  return 0;
}

extern "C" int ghost_bor_set_gemm_variant(ghost_bor* h, int variant) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (variant < -1 || variant > 5) return fail(GHOST_ERR_ARG, "GEMM variant must be -1..5");
  h->gemm_variant = variant;
  return 0;
}

extern "C" int ghost_bor_set_lu_streams(ghost_bor* h, int n) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (n != 1 && n != 2) return fail(GHOST_ERR_ARG, "LU streams must be 1 or 2");
  h->lu_streams = n;
  return 0;
}

extern "C" int ghost_bor_set_probe_tol(ghost_bor* h, double tol) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  h->probe_tol = tol;
  return 0;
}

// Probe residuals measured by the last adaptive solve (one per probe), *n in/out.
extern "C" int ghost_bor_refine_history(ghost_bor* h, double* out, int* n) {
  if (!h || !n) return fail(GHOST_ERR_ARG, "null argument");
  const int k = std::min(*n, (int)h->probe_hist.size());
  for (int i = 0; i < k; ++i) out[i] = h->probe_hist[i];
  *n = (int)h->probe_hist.size();
  return 0;
}

extern "C" int ghost_bor_farfield(ghost_bor* h) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  const bool mixed = h->precision == GHOST_MIXED;
  if (!h->capturing) HIPCHK(hipEventRecord(h->ev[4], h->stream));
  const int nsys = (int)h->sys.size();
  for (auto& p : h->probs) {
    ExciteArgs ea = excite_args(p, base_geom(p));
    const int P = p.P;
    const int nch = (P + h->ff_chunk - 1) / h->ff_chunk;
    // paired modes: only m >= 0 is summed, mode -m's far field is mode m's
    const int mlo = h->pairing ? p.M : 0, nv = p.nmodes - mlo;
    dim3 g((p.nrhs + 255) / 256, nv, nch);
    // partial layout [chunk][system-in-problem][nrhs], reduced into this problem's dF block
    prof_begin(h);
    if (mixed)
      hipLaunchKernelGGL((k_farfield_partial<double>), g, dim3(256), 0, h->stream, ea, h->d_sys + p.sys0 + mlo,
                         (const void*)h->d_X64, 0, h->ff_chunk, h->d_partial, nv);
    else
      hipLaunchKernelGGL((k_farfield_partial<double>), g, dim3(256), 0, h->stream, ea, h->d_sys + p.sys0 + mlo,
                         (const void*)h->d_A, 1, h->ff_chunk, h->d_partial, nv);
    prof_end(h, K_FARFIELD, 0.0);
    LAUNCHCHK();
    const int tot = nv * p.nrhs;
    hipLaunchKernelGGL(k_farfield_reduce, dim3((tot + 63) / 64), dim3(256), 0, h->stream, h->d_partial, nch,
                       nv, p.nrhs, p.k, h->d_dF + p.dF_off + (size_t)mlo * p.nrhs, 0);
    LAUNCHCHK();
    if (mlo > 0) {
      hipLaunchKernelGGL(k_mirror_dF, dim3((p.M * p.nrhs + 255) / 256), dim3(256), 0, h->stream,
                         h->d_dF + p.dF_off, p.M, p.nrhs);
      LAUNCHCHK();
    }
  }
  (void)nsys;
  // region sets: x_full = Q x_red, then every exterior surface's J (and M)
  for (auto& r : h->regions) {
    const size_t nX = (size_t)r.nmodes * r.n_full * r.nrhs;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (nX + 255) / 256)), dim3(256), 0, h->stream,
                       r.d_xfull, nX);
    int nmax = 0;
    for (int mi = 0; mi < r.nmodes; ++mi) nmax = std::max(nmax, h->sys[r.sys0 + mi].n);
    RegionMap q{r.d_ptr, r.d_mem, r.d_cf, r.n_full};
    hipLaunchKernelGGL(k_region_expand, dim3(nmax, (r.nrhs + 255) / 256, r.nmodes), dim3(256), 0, h->stream, q,
                       (const double2*)h->d_A, h->d_sys + r.sys0, r.d_xfull);
    LAUNCHCHK();
    for (size_t e = 0; e < r.ext.size(); ++e) {
      const Problem& o = h->inner[r.ext[e].op];
      ExciteArgs ea = excite_args(o, base_geom(o));
      const int nch = (o.P + h->ff_chunk - 1) / h->ff_chunk;
      dim3 g((r.nrhs + 255) / 256, r.nmodes, nch);
      prof_begin(h);
      hipLaunchKernelGGL(k_farfield_partial<double>, g, dim3(256), 0, h->stream, ea, r.d_vsys[e],
                         (const void*)r.d_xfull, 0, h->ff_chunk, h->d_partial, r.nmodes);
      prof_end(h, K_FARFIELD, 0.0);
      LAUNCHCHK();
      const int tot = r.nmodes * r.nrhs;
      hipLaunchKernelGGL(k_farfield_reduce, dim3((tot + 63) / 64), dim3(256), 0, h->stream, h->d_partial, nch,
                         r.nmodes, r.nrhs, o.k, h->d_dF + r.dF_off, e > 0 ? 1 : 0);
      LAUNCHCHK();
    }
  }
  if (!h->capturing) {
    HIPCHK(hipEventRecord(h->ev[5], h->stream));
    h->ev_recorded |= 4u;
  }
  return 0;
}

extern "C" int ghost_bor_residual(ghost_bor* h, double* out) {
  // ||A X - B||_F / ||B||_F per system (bor_solver.py:754-756); FP64 A and B copies
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (!h->d_A64) return fail(GHOST_ERR_STATE, "residual copies were disabled at build time");
  HIPCHK(hipSetDevice(h->device));
  const bool mixed = h->precision == GHOST_MIXED;
  int rc = residual_gemm(h, mixed ? 0 : 1);
  if (rc) return rc;
  std::vector<double> rn, bn;
  if ((rc = frob2(h, h->d_R64, rn))) return rc;
  if ((rc = frob2(h, h->d_B64, bn))) return rc;
  for (size_t i = 0; i < rn.size(); ++i) out[i] = bn[i] > 0 ? std::sqrt(rn[i] / bn[i]) : 0.0;
  return 0;
}

extern "C" int ghost_bor_run(ghost_bor* h, int refine_iters) {
  int rc = ghost_bor_fill(h);
  if (rc) return rc;
  if ((rc = ghost_bor_solve(h, refine_iters))) return rc;
  return ghost_bor_farfield(h);
}

extern "C" int ghost_bor_set_profile(ghost_bor* h, int on) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  h->profile = on != 0;
  h->recs.clear();
  return 0;
}

// Per kernel family since the last ghost_bor_set_profile: summed ms, launch
// count and algorithmic flops (GEMM families), each out array of length
// ghost_bor_kernel_count().
extern "C" int ghost_bor_kernel_count(void) { return K_COUNT; }

extern "C" const char* ghost_bor_kernel_name(int kind) {
  return (kind >= 0 && kind < K_COUNT) ? kKernelNames[kind] : "";
}

extern "C" int ghost_bor_kernel_times(ghost_bor* h, double* ms, int* count, double* flops) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int k = 0; k < K_COUNT; ++k) { ms[k] = 0.0; count[k] = 0; flops[k] = 0.0; }
  for (const auto& r : h->recs) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, h->pev[r.e0], h->pev[r.e0 + 1]));
    ms[r.kind] += t;
    count[r.kind] += 1;
    flops[r.kind] += r.flops;
  }
  return 0;
}

// Capture fill + solve + far field of the whole batch into one hipGraph
// (kernels only: no allocations, memsets or host synchronisation inside).
extern "C" int ghost_bor_capture(ghost_bor* h, int refine_iters) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->gexec) { hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  if (h->graph) { hipGraphDestroy(h->graph); h->graph = nullptr; }
  const bool prof = h->profile;
  h->profile = false;
  h->capturing = true;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = run_fill(h);
  if (!rc) rc = ghost_bor_solve(h, refine_iters);
  if (!rc) rc = ghost_bor_farfield(h);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->capturing = false;
  h->profile = prof;
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(GHOST_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  h->graph = g;
  HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  return 0;
}

// Replay the captured graph on the handle's stream (asynchronous).
extern "C" int ghost_bor_launch(ghost_bor* h) {
  if (!h || !h->gexec) return fail(GHOST_ERR_STATE, "capture first");
  HIPCHK(hipSetDevice(h->device));
  if (!h->capturing) HIPCHK(hipEventRecord(h->ev[0], h->stream));
  HIPCHK(hipGraphLaunch(h->gexec, h->stream));
  if (!h->capturing) HIPCHK(hipEventRecord(h->ev[6], h->stream));
  h->solved = true;
  return 0;
}

// Milliseconds of the last graph replay.
extern "C" int ghost_bor_graph_ms(ghost_bor* h, float* ms) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipEventElapsedTime(ms, h->ev[0], h->ev[6]));
  return 0;
}

// Device-wide barrier on the current device (every stream, every handle).
extern "C" int ghost_device_name(char* buf, int len) {
  if (!buf || len <= 0) return fail(GHOST_ERR_ARG, "null or empty buffer");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  HIPCHK(hipGetDeviceProperties(&pr, dev));
  snprintf(buf, (size_t)len, "%s|%s|cu%d", pr.name, pr.gcnArchName, pr.multiProcessorCount);
  return 0;
}

extern "C" int ghost_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

extern "C" int ghost_bor_sync(ghost_bor* h) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int ghost_bor_timings(ghost_bor* h, float* ms3) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  // a phase that has not run since build reports -1
  for (int i = 0; i < 3; ++i) {
    ms3[i] = -1.0f;
    if (h->ev_recorded & (1u << i)) HIPCHK(hipEventElapsedTime(&ms3[i], h->ev[2 * i], h->ev[2 * i + 1]));
  }
  return 0;
}

extern "C" int ghost_bor_num_systems(ghost_bor* h, int prob, int* nmodes, int* m_max) {
  if (!h || prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index");
  *nmodes = h->probs[prob].nmodes;
  *m_max = h->probs[prob].M;
  return 0;
}

extern "C" int ghost_bor_system_info(ghost_bor* h, int prob, int m, int* n, int* nt, int* t0, int* f0) {
  if (!h || prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index");
  const Problem& p = h->probs[prob];
  if (m < -p.M || m > p.M) return fail(GHOST_ERR_ARG, "mode %d outside [-%d, %d]", m, p.M, p.M);
  const SysInfo& s = h->sys[p.sys0 + m + p.M];
  *n = s.n; *nt = s.nt; *t0 = s.t0; *f0 = s.f0;
  return 0;
}

extern "C" int ghost_bor_get_dF(ghost_bor* h, int prob, double* out) {
  // out: complex [2M+1][2][n_theta] interleaved (mode index m + M, pol VV = 0, HH = 1)
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index");
  const Problem& p = h->probs[prob];
  std::vector<double2> buf((size_t)p.nmodes * p.nrhs);
  HIPCHK(hipMemcpyAsync(buf.data(), h->d_dF + p.dF_off, buf.size() * sizeof(double2),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int mi = 0; mi < p.nmodes; ++mi)
    for (int it = 0; it < p.ntheta; ++it)
      for (int pol = 0; pol < 2; ++pol) {
        const double2 v = buf[(size_t)mi * p.nrhs + 2 * it + pol];
        const size_t o = (((size_t)mi * 2 + pol) * p.ntheta + it) * 2;
        out[o] = v.x;
        out[o + 1] = v.y;
      }
  return 0;
}

// Copy one mode system out: A (n x n), B (n x nrhs), X (n x nrhs), complex interleaved, row-major.
// which: 0 = A and B as assembled (call after fill, before solve), 1 = X (after solve).
// Mode -m of a pair is read from mode m's storage: A = S A_m S, B and X = S [VV, -HH] of mode m's.
// Far field of a region set: complex [2M+1][2][n_theta] as ghost_bor_get_dF.
extern "C" int ghost_bor_region_get_dF(ghost_bor* h, int set, double* out) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (set < 0 || set >= (int)h->regions.size()) return fail(GHOST_ERR_ARG, "bad region set");
  if (!out) return fail(GHOST_ERR_ARG, "null output");
  const Region& r = h->regions[set];
  std::vector<double2> buf((size_t)r.nmodes * r.nrhs);
  HIPCHK(hipMemcpyAsync(buf.data(), h->d_dF + r.dF_off, buf.size() * sizeof(double2), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int mi = 0; mi < r.nmodes; ++mi)
    for (int it = 0; it < r.ntheta; ++it)
      for (int pol = 0; pol < 2; ++pol) {
        const double2 v = buf[(size_t)mi * r.nrhs + 2 * it + pol];
        const size_t o = (((size_t)mi * 2 + pol) * r.ntheta + it) * 2;
        out[o] = v.x;
        out[o + 1] = v.y;
      }
  return 0;
}

// One mode of a region set: size (n) and, which = 0, the reduced system A [n][n]
// and right-hand sides B [n][nrhs] after fill; which = 1, the currents X
// [n][nrhs] after solve (complex, interleaved).  Any output may be null.
extern "C" int ghost_bor_region_system(ghost_bor* h, int set, int m, int which, int* n, double* out1,
                                       double* out2) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (set < 0 || set >= (int)h->regions.size()) return fail(GHOST_ERR_ARG, "bad region set");
  const Region& r = h->regions[set];
  if (m < -r.M || m > r.M) return fail(GHOST_ERR_ARG, "mode outside range");
  const SysInfo& sy = h->sys[r.sys0 + m + r.M];
  if (n) *n = sy.n;
  if (!out1 && !out2) return 0;
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double2> row((size_t)sy.ld);
  for (int i = 0; i < sy.n; ++i) {
    HIPCHK(hipMemcpy(row.data(), (double2*)h->d_A + sy.a_off + (size_t)i * sy.ld, (size_t)sy.ld * sizeof(double2),
                     hipMemcpyDeviceToHost));
    if (which == 0 && out1)
      for (int c = 0; c < sy.n; ++c) {
        out1[2 * ((size_t)i * sy.n + c)] = row[c].x;
        out1[2 * ((size_t)i * sy.n + c) + 1] = row[c].y;
      }
    double* o = which == 0 ? out2 : out1;
    if (o)
      for (int c = 0; c < sy.nrhs; ++c) {
        o[2 * ((size_t)i * sy.nrhs + c)] = row[sy.n + c].x;
        o[2 * ((size_t)i * sy.nrhs + c) + 1] = row[sy.n + c].y;
      }
  }
  return 0;
}

extern "C" int ghost_bor_get_system(ghost_bor* h, int prob, int m, int which, double* out1, double* out2) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (prob < 0 || prob >= (int)h->probs.size()) return fail(GHOST_ERR_ARG, "bad problem index");
  const Problem& p = h->probs[prob];
  if (m < -p.M || m > p.M) return fail(GHOST_ERR_ARG, "mode outside range");
  const SysInfo& s = h->sys[p.sys0 + m + p.M];
  const bool mixed = h->precision == GHOST_MIXED;
  HIPCHK(hipStreamSynchronize(h->stream));
  const size_t csz = mixed ? sizeof(float2) : sizeof(double2);
  std::vector<char> rowbuf((size_t)s.ld * csz);
  std::vector<double2> row64((size_t)std::max(s.n, s.ncols));
  auto get = [&](int c) -> double2 {
    if (mixed) { float2 v = ((float2*)rowbuf.data())[c]; return make_double2(v.x, v.y); }
    return ((double2*)rowbuf.data())[c];
  };
  auto sgn = [&](int r, int c) { return s.flip ? pair_row_sign(s, r) * pair_row_sign(s, c) : 1.0; };
  // right-hand sides and currents of mode -m: S [VV, -HH] of mode m's
  auto rsgn = [&](int r, int c) { return s.flip ? pair_row_sign(s, r) * ((c & 1) ? -1.0 : 1.0) : 1.0; };
  auto put = [](double* o, size_t i, double2 v, double f) { o[2 * i] = f * v.x; o[2 * i + 1] = f * v.y; };
  for (int r = 0; r < s.n; ++r) {
    const bool aug_row = which == 0 ? !(mixed && h->d_A64) : !mixed;
    if (aug_row)
      HIPCHK(hipMemcpy(rowbuf.data(), (char*)h->d_A + ((size_t)s.a_off + (size_t)r * s.ld) * csz,
                       (size_t)s.ld * csz, hipMemcpyDeviceToHost));
    if (which == 0) {
      if (aug_row) {
        for (int c = 0; c < s.n; ++c) put(out1, (size_t)r * s.n + c, get(c), sgn(r, c));
        if (out2)
          for (int c = 0; c < s.nrhs; ++c) put(out2, (size_t)r * s.nrhs + c, get(s.n + s.col + c), rsgn(r, c));
      } else {   // mixed: the FP64 copies are exact; prefer them
        HIPCHK(hipMemcpy(row64.data(), h->d_A64 + s.a64_off + (size_t)r * s.n, (size_t)s.n * sizeof(double2),
                         hipMemcpyDeviceToHost));
        for (int c = 0; c < s.n; ++c) put(out1, (size_t)r * s.n + c, row64[c], sgn(r, c));
        if (out2) {
          HIPCHK(hipMemcpy(row64.data(), h->d_B64 + s.x_off + (size_t)r * s.ncols + s.col,
                           (size_t)s.nrhs * sizeof(double2), hipMemcpyDeviceToHost));
          for (int c = 0; c < s.nrhs; ++c) put(out2, (size_t)r * s.nrhs + c, row64[c], rsgn(r, c));
        }
      }
    } else if (mixed) {
      HIPCHK(hipMemcpy(row64.data(), h->d_X64 + s.x_off + (size_t)r * s.ncols + s.col,
                       (size_t)s.nrhs * sizeof(double2), hipMemcpyDeviceToHost));
      for (int c = 0; c < s.nrhs; ++c) put(out1, (size_t)r * s.nrhs + c, row64[c], rsgn(r, c));
    } else {
      for (int c = 0; c < s.nrhs; ++c) put(out1, (size_t)r * s.nrhs + c, get(s.n + s.col + c), rsgn(r, c));
    }
  }
  return 0;
}

extern "C" int ghost_bor_info(ghost_bor* h, int* out) {
  // per-system count of zero pivots met by the panel factorisation (a pair's
  // two modes report their shared factorisation)
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  std::vector<int> info(h->ss.size());
  HIPCHK(hipMemcpy(info.data(), h->d_info, info.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < h->sys.size(); ++i) out[i] = info[h->sys[i].owner];
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// compat samplers: the bor_stream_kernel.c entry points, HIP-backed
// (host pointers in and out, synchronous, reentrant)
// ─────────────────────────────────────────────────────────────────────────────

struct ScopedBufs {
  std::vector<void*> p;
  hipStream_t s = nullptr;
  ~ScopedBufs() {
    if (s) hipStreamSynchronize(s);
    for (void* q : p) hipFree(q);
    if (s) hipStreamDestroy(s);
  }
  template <typename T> T* up(const T* src, size_t n) {
    void* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(d);
    if (src && hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) return nullptr;
    return (T*)d;
  }
};

extern "C" int ghost_sample_g(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                              const double* rho_q, const double* z_q, double k, const double* sin2_tab,
                              double* out) {
  if (nr <= 0 || np_ <= 0 || nxi <= 0) return 0;
  if (!rho_p || !z_p || !rho_q || !z_q || !sin2_tab || !out) return fail(GHOST_ERR_ARG, "sample_g: null array");
  ScopedBufs b;
  HIPCHK(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
  const double *a = b.up(rho_p, nr), *bz = b.up(z_p, nr), *c = b.up(rho_q, np_), *d = b.up(z_q, np_);
  const double* t = b.up(sin2_tab, nxi);
  const size_t tot = (size_t)nr * np_ * nxi;
  double2* o = b.up((const double2*)nullptr, tot);
  if (!a || !bz || !c || !d || !t || !o) return fail(GHOST_ERR_OOM, "sample_g: device allocation failed");
  hipLaunchKernelGGL(k_sample_g, dim3((unsigned)std::min<size_t>(65535, (tot + 255) / 256)), dim3(256), 0, b.s,
                     nr, np_, nxi, a, bz, c, d, k, t, o);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(out, o, tot * sizeof(double2), hipMemcpyDeviceToHost, b.s));
  HIPCHK(hipStreamSynchronize(b.s));
  return 0;
}

static int sample_brackets(bool ibc, int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                           const double* tr_p, const double* tz_p, const double* rho_q, const double* z_q,
                           const double* tr_q, const double* tz_q, double k, const double* cx_tab,
                           const double* sx_tab, double* o_tt, double* o_tf, double* o_ft, double* o_ff) {
  if (nr <= 0 || np_ <= 0 || nxi <= 0) return 0;
  if (!rho_p || !z_p || !tr_p || !tz_p || !rho_q || !z_q || !tr_q || !tz_q || !cx_tab || !sx_tab || !o_tt || !o_tf ||
      !o_ft || !o_ff)
    return fail(GHOST_ERR_ARG, "sample brackets: null array");
  ScopedBufs b;
  HIPCHK(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
  const double *rp = b.up(rho_p, nr), *zp = b.up(z_p, nr), *trp = b.up(tr_p, nr), *tzp = b.up(tz_p, nr);
  const double *rq = b.up(rho_q, np_), *zq = b.up(z_q, np_), *trq = b.up(tr_q, np_), *tzq = b.up(tz_q, np_);
  const double *cx = b.up(cx_tab, nxi), *sx = b.up(sx_tab, nxi);
  const size_t tot = (size_t)nr * np_ * nxi;
  double2* o[4];
  for (int i = 0; i < 4; ++i) o[i] = b.up((const double2*)nullptr, tot);
  if (!rp || !zp || !trp || !tzp || !rq || !zq || !trq || !tzq || !cx || !sx || !o[0] || !o[1] || !o[2] || !o[3])
    return fail(GHOST_ERR_OOM, "sample brackets: device allocation failed");
  const dim3 g((unsigned)std::min<size_t>(65535, (tot + 255) / 256));
  if (ibc)
    hipLaunchKernelGGL(k_sample_brackets<true>, g, dim3(256), 0, b.s, nr, np_, nxi, rp, zp, trp, tzp, rq, zq, trq,
                       tzq, k, cx, sx, o[0], o[1], o[2], o[3]);
  else
    hipLaunchKernelGGL(k_sample_brackets<false>, g, dim3(256), 0, b.s, nr, np_, nxi, rp, zp, trp, tzp, rq, zq, trq,
                       tzq, k, cx, sx, o[0], o[1], o[2], o[3]);
  LAUNCHCHK();
  double* outs[4] = {o_tt, o_tf, o_ft, o_ff};
  for (int i = 0; i < 4; ++i)
    HIPCHK(hipMemcpyAsync(outs[i], o[i], tot * sizeof(double2), hipMemcpyDeviceToHost, b.s));
  HIPCHK(hipStreamSynchronize(b.s));
  return 0;
}

extern "C" int ghost_sample_mfie(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                                 const double* tr_p, const double* tz_p, const double* rho_q,
                                 const double* z_q, const double* tr_q, const double* tz_q, double k,
                                 const double* cx_tab, const double* sx_tab, double* o_tt, double* o_tf,
                                 double* o_ft, double* o_ff) {
  return sample_brackets(false, nr, np_, nxi, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, cx_tab, sx_tab,
                         o_tt, o_tf, o_ft, o_ff);
}

extern "C" int ghost_sample_ibc(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                                const double* tr_p, const double* tz_p, const double* rho_q,
                                const double* z_q, const double* tr_q, const double* tz_q, double k,
                                const double* cx_tab, const double* sx_tab, double* o_tt, double* o_tf,
                                double* o_ft, double* o_ff) {
  return sample_brackets(true, nr, np_, nxi, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, cx_tab, sx_tab,
                         o_tt, o_tf, o_ft, o_ff);
}

// void-returning names with the exact bor_stream_kernel.c signatures
// (bor_stream_kernel.c:26-29, 65-72, 108-115): errors surface through
// ghost_last_error() because the reference ABI has no error channel.
// ── no-allocation bracket sampler ──────────────────────────────────────────
// ghost_sample_mfie_noalloc stages its inputs and outputs through `scratch`, a
// file-scope device array, instead of device buffers allocated per call: for
// callers that cannot call hipMalloc (tight loops, harnesses that forbid it).
// Each device holds its own copy of the array at its own address, looked up
// once per device with hipGetSymbolAddress; host code uses only that address.
// The sampler overwrites every staged word before reading it, so nothing in
// the array needs clearing between calls.  Calls that overlap on one device
// share the array and race.
static constexpr size_t kScratchBytes = (size_t)48 << 20;   // GHOST_NOALLOC_SCRATCH_BYTES
static __device__ __align__(256) char scratch[kScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices];

static int scratch_ptr(char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GHOST_ERR_ARG, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GHOST_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" int ghost_sample_mfie_noalloc(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                                         const double* tr_p, const double* tz_p, const double* rho_q,
                                         const double* z_q, const double* tr_q, const double* tz_q, double k,
                                         const double* cx_tab, const double* sx_tab, double* o_tt, double* o_tf,
                                         double* o_ft, double* o_ff) {
  if (nr <= 0 || np_ <= 0 || nxi <= 0) return 0;
  if (!rho_p || !z_p || !tr_p || !tz_p || !rho_q || !z_q || !tr_q || !tz_q || !cx_tab || !sx_tab || !o_tt || !o_tf ||
      !o_ft || !o_ff)
    return fail(GHOST_ERR_ARG, "sample_mfie_noalloc: null array");
  // staging layout, every piece on a 256-byte boundary: 4 test arrays [nr],
  // 4 source arrays [np_], 2 azimuth tables [nxi], 4 outputs [nr np_ nxi] complex
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t tot = (size_t)nr * np_ * nxi;
  const size_t bP = up(8 * (size_t)nr), bQ = up(8 * (size_t)np_), bX = up(8 * (size_t)nxi), bO = up(16 * tot);
  const size_t need = 4 * bP + 4 * bQ + 2 * bX + 4 * bO;
  if (need > kScratchBytes)
    return fail(GHOST_ERR_ARG,
                "sample_mfie_noalloc: %d x %d x %d samples need %zu bytes of staging, above the %zu-byte scratch array"
                " (split the rows)",
                nr, np_, nxi, need, kScratchBytes);
  char* base = nullptr;
  if (int rc = scratch_ptr(&base)) return rc;
  const double* hin[10] = {rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, cx_tab, sx_tab};
  const size_t nin[10] = {(size_t)nr, (size_t)nr, (size_t)nr, (size_t)nr, (size_t)np_, (size_t)np_,
                          (size_t)np_, (size_t)np_, (size_t)nxi, (size_t)nxi};
  const double* din[10];
  size_t off = 0;
  for (int i = 0; i < 10; ++i) {
    din[i] = (const double*)(base + off);
    HIPCHK(hipMemcpy((void*)din[i], hin[i], 8 * nin[i], hipMemcpyHostToDevice));
    off += i < 4 ? bP : (i < 8 ? bQ : bX);
  }
  double2* o[4];
  for (int i = 0; i < 4; ++i) {
    o[i] = (double2*)(base + off);
    off += bO;
  }
  const dim3 g((unsigned)std::min<size_t>(65535, (tot + 255) / 256));
  hipLaunchKernelGGL(k_sample_brackets<false>, g, dim3(256), 0, 0, nr, np_, nxi, din[0], din[1], din[2], din[3],
                     din[4], din[5], din[6], din[7], k, din[8], din[9], o[0], o[1], o[2], o[3]);
  LAUNCHCHK();
  double* outs[4] = {o_tt, o_tf, o_ft, o_ff};
  for (int i = 0; i < 4; ++i) HIPCHK(hipMemcpy(outs[i], o[i], tot * sizeof(double2), hipMemcpyDeviceToHost));
  return 0;
}

// The reference's samplers return nothing, and its loader allocates the
// outputs with np.empty (bor_streaming.py:287, 327) without ever asking for an
// error: on any failure the void forms below fill every output with NaN, so
// the failure surfaces in the caller's fill instead of as uninitialised memory,
// and report the first failure of the process on stderr.
static void compat_failed(const char* name, size_t tot, double* const* outs, int nout) {
  static std::once_flag once;
  std::call_once(once, [&] { fprintf(stderr, "libghost_bor: %s failed: %s (outputs set to NaN)\n", name, g_err.c_str()); });
  const double nan = std::nan("");
  for (int i = 0; i < nout; ++i)
    if (outs[i])
      for (size_t j = 0; j < 2 * tot; ++j) outs[i][j] = nan;
}
static size_t compat_total(int nr, int np_, int nxi) {
  return (nr > 0 && np_ > 0 && nxi > 0) ? (size_t)nr * np_ * nxi : 0;
}
extern "C" void sample_g(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* rho_q,
                         const double* z_q, double k, const double* sin2_tab, double* out) {
  if (ghost_sample_g(nr, np_, nxi, rho_p, z_p, rho_q, z_q, k, sin2_tab, out) != 0)
    compat_failed("sample_g", compat_total(nr, np_, nxi), &out, 1);
}
extern "C" void sample_mfie(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                            const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                            const double* tz_q, double k, const double* cx_tab, const double* sx_tab,
                            double* o_tt, double* o_tf, double* o_ft, double* o_ff) {
  if (ghost_sample_mfie(nr, np_, nxi, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, cx_tab, sx_tab, o_tt,
                        o_tf, o_ft, o_ff) != 0) {
    double* const outs[4] = {o_tt, o_tf, o_ft, o_ff};
    compat_failed("sample_mfie", compat_total(nr, np_, nxi), outs, 4);
  }
}
extern "C" void sample_ibc(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                           const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                           const double* tz_q, double k, const double* cx_tab, const double* sx_tab,
                           double* o_tt, double* o_tf, double* o_ft, double* o_ff) {
  if (ghost_sample_ibc(nr, np_, nxi, rho_p, z_p, tr_p, tz_p, rho_q, z_q, tr_q, tz_q, k, cx_tab, sx_tab, o_tt,
                       o_tf, o_ft, o_ff) != 0) {
    double* const outs[4] = {o_tt, o_tf, o_ft, o_ff};
    compat_failed("sample_ibc", compat_total(nr, np_, nxi), outs, 4);
  }
}

// ─────────────────────────────────────────────────────────────────────────────
// adjoint: gradient of a real far-field objective w.r.t. node radial offsets
// ─────────────────────────────────────────────────────────────────────────────

static int adjoint_alloc(ghost_bor* h) {
  if (h->adj_alloc) return 0;
  int rc;
  size_t nrow = 0, ncol = 0, ffmax = 0, nemax = 0;
  for (auto& s : h->sys) nrow += (size_t)s.n;
  for (auto& p : h->probs) {
    ncol += (size_t)p.nrhs;
    ffmax = std::max(ffmax, (size_t)p.Ne * p.nmodes * p.nrhs);
    nemax = std::max(nemax, (size_t)p.Ne);
  }
  if ((rc = dalloc(h, &h->d_Xs, h->x_elems * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_lam, h->x_elems * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_lam2, h->x_elems * sizeof(double2)))) return rc;
  const bool mixed = h->precision == GHOST_MIXED;
  if ((rc = dalloc(h, &h->d_Q, h->a_elems * (mixed ? sizeof(float2) : sizeof(double2))))) return rc;
  if (mixed && (rc = dalloc(h, &h->d_X32, h->x_elems * sizeof(float2)))) return rc;
  if ((rc = dalloc(h, &h->d_rs, nrow * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_acoef, ncol * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_ffpart, ffmax * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_ffs, nemax * sizeof(double2)))) return rc;
  if ((rc = dalloc(h, &h->d_rowoff, h->sys.size() * sizeof(int)))) return rc;
  std::vector<int> ro(h->sys.size());
  for (size_t i = 0, o = 0; i < h->sys.size(); ++i) { ro[i] = (int)o; o += (size_t)h->sys[i].n; }
  HIPCHK(hipMemcpy(h->d_rowoff, ro.data(), ro.size() * sizeof(int), hipMemcpyHostToDevice));
  for (auto& p : h->probs)
    if ((rc = dalloc(h, &p.d_alt, (9 * (size_t)p.P + 2 * (size_t)p.Nn + 3 * (size_t)p.Ne) * 8))) return rc;
  size_t far_n = 0, near_n = 0, ex_n = 0, nn_max = 0;
  for (auto& p : h->probs) {
    far_n = std::max(far_n, (size_t)p.Ne * p.Ne * 4);
    near_n = std::max(near_n, p.batches.size() * 4);
    ex_n = std::max(ex_n, (size_t)p.Ne * p.ntheta * 2);
    nn_max = std::max(nn_max, (size_t)p.Nn);
  }
  if ((rc = dalloc(h, &h->d_adj_far, std::max<size_t>(far_n, 4) * 8))) return rc;
  if ((rc = dalloc(h, &h->d_adj_near, std::max<size_t>(near_n, 4) * 8))) return rc;
  if ((rc = dalloc(h, &h->d_adj_ex, std::max<size_t>(ex_n, 2) * 8))) return rc;
  size_t nn_all = 0;
  for (auto& p : h->probs) nn_all += (size_t)p.Nn;
  (void)nn_max;
  if ((rc = dalloc(h, &h->d_adj_grad, std::max<size_t>(nn_all, 1) * 8))) return rc;
  if ((rc = dalloc(h, &h->d_adj_gauss, 8 * std::max<size_t>(h->probs.size(), 1) * 8))) return rc;
  {
    // the element Gauss rule (s, weight on [0, 1]) of each problem's point tables
    std::vector<double> gq(8 * h->probs.size());
    for (size_t pi = 0; pi < h->probs.size(); ++pi) {
      const Problem& p = h->probs[pi];
      for (int i = 0; i < 4; ++i) {
        gq[8 * pi + i] = p.pts[6 * (size_t)p.P + i];              // T1 = s
        gq[8 * pi + 4 + i] = p.pts[4 * (size_t)p.P + i] / p.L[0]; // w / L
      }
    }
    HIPCHK(hipMemcpy(h->d_adj_gauss, gq.data(), gq.size() * 8, hipMemcpyHostToDevice));
  }
  h->adj_alloc = true;
  return 0;
}

// Node q of a problem is perturbed in class r when q = r (mod 3) and it lies
// off the axis by more than twice the step.
static bool adj_free(const Problem& p, int q, int r, double step) {
  return q >= 0 && q < p.Nn && q % 3 == r && p.nodes[2 * q] > 2.0 * step;
}

// What every adjoint form needs of the handle (solved: checked by the
// callers that run on an already solved handle).
static int adjoint_checks(ghost_bor* h, bool need_solved) {
  if (!h) return fail(GHOST_ERR_ARG, "null handle");
  if (!h->built) return fail(GHOST_ERR_STATE, "build first");
  for (auto& s : h->sys) {
    if (s.prob < 0) return fail(GHOST_ERR_STATE, "the adjoint covers handles of BoR problems only");
    if (s.ncols != s.nrhs) return fail(GHOST_ERR_STATE, "the adjoint needs a handle built with mode pairing off");
  }
  for (auto& p : h->probs)
    if (p.has_zs || p.ki != 0.0 || p.eta_re != 376.730313668 || p.eta_im != 0.0)
      return fail(GHOST_ERR_STATE, "the adjoint covers PEC bodies in free space only");
  if (!h->regions.empty() || !h->cross.empty())
    return fail(GHOST_ERR_STATE, "the adjoint covers PEC bodies in free space only (no multi-region sets)");
  if (need_solved && !h->solved) return fail(GHOST_ERR_STATE, "fill and solve before the adjoint");
  HIPCHK(hipSetDevice(h->device));
  return 0;
}

// a' = a (-j k eta0 / 4 pi) per problem column, from the caller's a
static int adjoint_host_acoef(ghost_bor* h, const double* acoef, std::vector<size_t>& coff) {
  // a' = a (-j k eta0 / 4 pi) per problem column; G = a' dS/dX
  coff.assign(h->probs.size(), 0);
  std::vector<double2> ac;
  for (size_t pi = 0; pi < h->probs.size(); ++pi) {
    const Problem& p = h->probs[pi];
    coff[pi] = ac.size();
    const double pr = -p.k * 376.730313668 / (4.0 * M_PI);   // pref = j pr
    for (int c = 0; c < p.nrhs; ++c) {
      const double ar = acoef[2 * (coff[pi] + c)], ai = acoef[2 * (coff[pi] + c) + 1];
      ac.push_back(make_double2(-ai * pr, ar * pr));       // (ar + j ai)(j pr)
    }
  }
  HIPCHK(hipMemcpyAsync(h->d_acoef, ac.data(), ac.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  return 0;
}

// lambda = A^-T G through the forward LU factors (G = a' dS/dX with a' in
// d_acoef, coff: each problem's first column) and Q = [-Lambda X^T | Lambda],
// shared by both adjoint forms; kernels only (capturable).
static int adjoint_lambda_q(ghost_bor* h, const std::vector<size_t>& coff) {
  int rc;
  const bool mixed = h->precision == GHOST_MIXED;
  const int nsys = (int)h->sys.size();
  size_t maxn = 0, maxr = 0, nrow = 0;
  for (auto& s : h->sys) {
    maxn = std::max(maxn, (size_t)s.n);
    maxr = std::max(maxr, (size_t)s.nrhs);
    nrow += (size_t)s.n;
  }
  const dim3 gv((unsigned)((maxr + 255) / 256), (unsigned)maxn, nsys);
  // forward solution: out of the FP64 augmented columns, or the refined FP64
  // currents of a mixed-precision solve; the factor's row order
  if (mixed) hipLaunchKernelGGL(k_copy64, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_X64, h->d_Xs);
  else hipLaunchKernelGGL(k_copy64, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_aug, (const double2*)h->d_A, h->d_Xs);
  hipLaunchKernelGGL(k_perm_from_pivots, dim3(nsys), dim3(256), 0, h->stream, h->d_vec + h->vec_all, nsys,
                     h->d_piv, h->d_perm);
  LAUNCHCHK();
  for (size_t pi = 0; pi < h->probs.size(); ++pi) {
    Problem& p = h->probs[pi];
    int mn = 0;
    for (int i = 0; i < p.nmodes; ++i) mn = std::max(mn, h->sys[p.sys0 + i].n);
    hipLaunchKernelGGL(k_ffvec, dim3((p.nrhs + 255) / 256, mn, p.nmodes), dim3(256), 0, h->stream,
                       excite_args(p, base_geom(p)), h->d_sys + p.sys0, h->d_acoef + coff[pi], h->d_lam);
    LAUNCHCHK();
  }
  // lambda = A^-T G through the LU factors
  if (!mixed) {
    if ((rc = run_sched<double>(h, h->adjT_fwd, h->d_A, h->d_lam))) return rc;
    if ((rc = run_sched<double>(h, h->adjT_back, h->d_A, h->d_lam))) return rc;
    hipLaunchKernelGGL(k_unpermute, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_perm, h->d_lam,
                       h->d_lam2);
    LAUNCHCHK();
  } else {
    // FP32 factors under FP64 refinement, as the forward solve: R = G - A64^T
    // lambda in FP64 (the residual GEMM's ops with the operand transposed),
    // column-scaled to FP32, z = (L^T)^-1 (U^T)^-1 R with the FP32 factors,
    // lambda += P^T z; the first pass starts from lambda = 0 (R = G)
    const dim3 gs((unsigned)((maxr + 63) / 64), 1, nsys);
    for (int it = 0; it <= h->adj_refine; ++it) {
      hipLaunchKernelGGL(k_copy64, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_lam, h->d_R64);
      if (it > 0)
        gemm_launch<double, true, false>(h, h->d_gemm + h->res_gemm, nsys, h->res_tt, h->d_A64, h->d_lam2, h->d_R64);
      hipLaunchKernelGGL(k_col_scale, gs, dim3(64, 4), 0, h->stream, h->d_vec + h->vec_all, h->d_R64, h->d_rscale);
      hipLaunchKernelGGL(k_scale_cast, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, nullptr, h->d_R64,
                         h->d_rscale, h->d_Y32);
      LAUNCHCHK();
      if ((rc = run_sched<float>(h, h->adjT_fwd, h->d_A, h->d_Y32))) return rc;
      if ((rc = run_sched<float>(h, h->adjT_back, h->d_A, h->d_Y32))) return rc;
      hipLaunchKernelGGL(k_unpermute_widen, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_perm, h->d_Y32,
                         h->d_rscale, h->d_lam2, it > 0 ? 1 : 0);
      LAUNCHCHK();
    }
  }
  // Q = [-Lambda X^T | Lambda]
  if (!mixed) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (h->a_elems + 255) / 256)), dim3(256), 0,
                       h->stream, h->d_Q, h->a_elems);
    gemm_launch<double, false, true>(h, h->d_gemm + h->adj_q_ops, nsys, h->adj_q_tt, h->d_lam2, h->d_Xs, h->d_Q);
    hipLaunchKernelGGL(k_copy64, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_toaug, h->d_lam2, h->d_Q);
  } else {
    // FP32 Q (half the bytes of the FP64 systems it pairs with): Lambda and X
    // cast to FP32, Q = -Lambda X^T on the FP32 GEMM, Lambda into the
    // augmented columns
    const size_t qz = (h->a_elems + 1) / 2;   // float2 pairs as double2
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(65535, (qz + 255) / 256)), dim3(256), 0, h->stream,
                       h->d_Q, qz);
    hipLaunchKernelGGL(k_cast32, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_lam2, h->d_Y32);
    hipLaunchKernelGGL(k_cast32, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_all, h->d_Xs, h->d_X32);
    gemm_launch<float, false, true>(h, h->d_gemm + h->adj_q_ops, nsys, h->adj_q_tt, h->d_Y32, h->d_X32,
                                    (float2*)h->d_Q);
    hipLaunchKernelGGL(k_cast32, gv, dim3(256), 0, h->stream, h->d_vec + h->vec_toaug, h->d_lam2, (float2*)h->d_Q);
  }
  LAUNCHCHK();
  return 0;
}

extern "C" int ghost_bor_adjoint(ghost_bor* h, const double* acoef, double step, double* grad) {
  if (!h || !acoef || !grad) return fail(GHOST_ERR_ARG, "null argument");
  if (!(step >= 0.0)) return fail(GHOST_ERR_ARG, "step must be >= 0 (0: automatic)");
  if (h->precision != GHOST_FP64) return fail(GHOST_ERR_STATE, "the difference adjoint needs an FP64 handle");
  const bool prof = h->profile;
  h->profile = false;
  std::vector<size_t> coff;
  int rc;
  if ((rc = adjoint_checks(h, true)) || (rc = adjoint_alloc(h)) || (rc = adjoint_host_acoef(h, acoef, coff)) ||
      (rc = adjoint_lambda_q(h, coff))) {
    h->profile = prof;
    return rc;
  }
  const int nsys = (int)h->sys.size();
  size_t maxn = 0, nrow = 0;
  for (auto& s : h->sys) {
    maxn = std::max(maxn, (size_t)s.n);
    nrow += (size_t)s.n;
  }
  HIPCHK(hipSetDevice(h->device));
  // half-width per problem: given, or 1e-6 of the shortest element (test and
  // source sides move separately, so nearly coincident near-band point pairs,
  // a few thousandths of an element apart, set the curvature scale)
  std::vector<double> hstep(h->probs.size(), step);
  for (size_t pi = 0; pi < h->probs.size(); ++pi)
    if (step == 0.0) {
      const Problem& p = h->probs[pi];
      double lmin = INFINITY;
      for (int e = 0; e < p.Ne; ++e)
        lmin = std::min(lmin, std::hypot(p.nodes[2 * e + 2] - p.nodes[2 * e], p.nodes[2 * e + 3] - p.nodes[2 * e + 1]));
      hstep[pi] = 1e-6 * lmin;
    }
  // perturbed fills: side 0 moves the test geometry (rows, right-hand sides and
  // far-field projection), side 1 the source geometry (columns); nodes of one
  // residue class mod 3 move together, whose row / column triples are disjoint
  size_t goff_tot = 0;
  std::vector<size_t> goff(h->probs.size());
  for (size_t pi = 0; pi < h->probs.size(); ++pi) { goff[pi] = goff_tot; goff_tot += (size_t)h->probs[pi].Nn; }
  std::vector<double> g(goff_tot, 0.0);
  std::vector<double2> rs_plus(nrow), rs(nrow);
  std::vector<std::vector<double2>> ff_plus(h->probs.size()), ff(h->probs.size());
  for (int side = 0; side < 2; ++side)
    for (int r = 0; r < 3; ++r)
      for (int sg = 0; sg < 2; ++sg) {
        const double sign = sg == 0 ? 1.0 : -1.0;
        for (size_t pi = 0; pi < h->probs.size(); ++pi) {
          Problem& p = h->probs[pi];
          std::vector<double> nodes(p.nodes);
          for (int q = 0; q < p.Nn; ++q)
            if (adj_free(p, q, r, hstep[pi])) nodes[2 * q] += sign * hstep[pi];
          std::vector<double> pts, elem;
          geometry_tables(nodes.data(), p.Ne, pts, elem);
          double* alt = p.d_alt;
          HIPCHK(hipMemcpyAsync(alt, pts.data(), pts.size() * 8, hipMemcpyHostToDevice, h->stream));
          HIPCHK(hipMemcpyAsync(alt + pts.size(), nodes.data(), nodes.size() * 8, hipMemcpyHostToDevice, h->stream));
          HIPCHK(hipMemcpyAsync(alt + pts.size() + nodes.size(), elem.data(), elem.size() * 8, hipMemcpyHostToDevice,
                                h->stream));
          const Geom ag{alt, alt + pts.size(), alt + pts.size() + nodes.size()};
          if ((rc = fill_problem(h, p, side == 0 ? ag : base_geom(p), side == 1 ? ag : base_geom(p)))) return rc;
          if (side == 0) {
            // explicit far-field term: X fixed, per-element partial sums on the moved points
            hipLaunchKernelGGL((k_farfield_partial<double>), dim3((p.nrhs + 255) / 256, p.nmodes, p.Ne), dim3(256), 0,
                               h->stream, excite_args(p, ag), h->d_sys + p.sys0, (const void*)h->d_Xs, 0, 4,
                               h->d_ffpart, p.nmodes);
            hipLaunchKernelGGL(k_adj_ffsum, dim3(p.Ne), dim3(256), 0, h->stream, h->d_ffpart, h->d_acoef + coff[pi],
                               p.nmodes, p.nrhs, h->d_ffs);
            LAUNCHCHK();
            ff[pi].resize(p.Ne);
            HIPCHK(hipMemcpyAsync(ff[pi].data(), h->d_ffs, p.Ne * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
          }
        }
        if (side == 0)
          hipLaunchKernelGGL(k_adj_rowsum, dim3((unsigned)maxn, nsys), dim3(256), 0, h->stream, h->d_sys, h->d_rowoff,
                             h->d_Q, (const double2*)h->d_A, h->d_rs);
        else
          hipLaunchKernelGGL(k_adj_colsum, dim3((unsigned)((maxn + 63) / 64), 1, nsys), dim3(64, 4), 0, h->stream,
                             h->d_sys, h->d_rowoff, h->d_Q, (const double2*)h->d_A, h->d_rs);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(rs.data(), h->d_rs, nrow * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (sg == 0) {
          rs_plus.swap(rs);
          ff_plus.swap(ff);
          continue;
        }
        // central differences, binned to the moved node of each row / column / element
        size_t o = 0;
        for (int si = 0; si < nsys; ++si) {
          const SysInfo& s = h->sys[si];
          const Problem& p = h->probs[s.prob];
          const double inv2h = 0.5 / hstep[s.prob];
          for (int i = 0; i < s.n; ++i) {
            const int node = i < s.nt ? s.t0 + i : s.f0 + (i - s.nt);
            for (int q = node - 1; q <= node + 1; ++q)
              if (adj_free(p, q, r, hstep[s.prob])) g[goff[s.prob] + q] += (rs_plus[o + i].x - rs[o + i].x) * inv2h;
          }
          o += (size_t)s.n;
        }
        if (side == 0)
          for (size_t pi = 0; pi < h->probs.size(); ++pi) {
            const Problem& p = h->probs[pi];
            const double inv2h = 0.5 / hstep[pi];
            for (int e = 0; e < p.Ne; ++e)
              for (int q = e; q <= e + 1; ++q)
                if (adj_free(p, q, r, hstep[pi])) g[goff[pi] + q] += (ff_plus[pi][e].x - ff[pi][e].x) * inv2h;
          }
      }
  std::copy(g.begin(), g.end(), grad);
  h->solved = false;   // the systems now hold the last perturbed fill
  h->profile = prof;
  return 0;
}

static QView adj_qview(ghost_bor* h) {
  if (h->precision == GHOST_MIXED) return QView{nullptr, (const float2*)h->d_Q};
  return QView{h->d_Q, nullptr};
}

// Analytic form: the fill's shape derivative contracted with Q in the kernels
// of dfill.hip (far pairs, near pairs with the mass term, right-hand sides
// and far-field projection), gathered per node into d_adj_grad (every
// problem's nodes in turn).  Kernels only (capturable); the systems keep the
// forward factors (nothing is refilled).
static int adjoint_analytic_kernels(ghost_bor* h, const std::vector<size_t>& coff) {
  size_t goff = 0;
  for (size_t pi = 0; pi < h->probs.size(); ++pi) {
    Problem& p = h->probs[pi];
    const SysInfo* sys = h->d_sys + p.sys0;
    const double* gauss = h->d_adj_gauss + 8 * pi;   // s [4] | weights [4]
    const double ke = p.k * 376.730313668;
    // far element pairs
    {
      AdjFarArgs fa;
      fa.nodes = p.d_nodes;
      fa.gs = gauss;
      fa.gwq = gauss + 4;
      fa.sys = sys;
      fa.Q = adj_qview(h);
      fa.Ne = p.Ne; fa.M = p.M; fa.nE = p.nE; fa.nB = p.nB;
      fa.k = p.k;
      fa.ce_re = 0.0;
      fa.ce_im = p.cE * ke * 2.0 * M_PI;
      fa.cm = -2.0 * M_PI * p.cM;
      fa.out = h->d_adj_far;
      const size_t lds = (16 * (size_t)(p.M + 1) + (size_t)ADJ_CHUNK * 28) * sizeof(double2);
      hipLaunchKernelGGL(k_adj_far, dim3((unsigned)((size_t)p.Ne * p.Ne)), dim3(256), lds, h->stream, fa);
      LAUNCHCHK();
    }
    // near element pairs (and the mass term)
    if (!p.batches.empty()) {
      AdjNearArgs na;
      na.nodes = p.d_nodes;
      const int ncell = p.cell_off[2] + p.cell_cnt[2];
      na.cell_s = p.d_cells;
      na.cell_sp = p.d_cells + ncell;
      na.cell_w = p.d_cells + 2 * ncell;
      for (int i = 0; i < 8; ++i) { na.cell_off[i] = p.cell_off[i]; na.cell_cnt[i] = p.cell_cnt[i]; }
      na.core_u = p.d_core;
      na.core_w = p.d_core + 48;
      na.tail_u = p.d_tail;
      na.tail_w = p.d_tail + p.tail_w.size();
      static_assert(sizeof(AdjNearPair) == sizeof(NearPairInfo), "near pair layouts");
      na.pairs = reinterpret_cast<const AdjNearPair*>(p.d_pairs);
      na.batches = p.d_batches;
      na.sys = sys;
      na.Q = adj_qview(h);
      na.gs = gauss;
      na.gwq = gauss + 4;
      na.Ne = p.Ne; na.M = p.M;
      na.k = p.k;
      na.ce_re = 0.0;
      na.ce_im = p.cE * ke * 2.0 * M_PI;
      na.cm = -2.0 * M_PI * p.cM;
      na.cmass = M_PI * p.cM;
      na.out = h->d_adj_near;
      // the 16 staged Q entry rows and the 19 contracted coefficient rows of one point pair
      const size_t lds = 35 * (size_t)(p.M + 1) * sizeof(double2);
      hipLaunchKernelGGL(k_adj_near, dim3((unsigned)p.batches.size()), dim3(256), lds, h->stream, na);
      LAUNCHCHK();
    }
    // right-hand sides and far-field projection
    {
      AdjExciteArgs ea;
      ea.nodes = p.d_nodes;
      ea.gs = gauss;
      ea.gwq = gauss + 4;
      ea.sys = sys;
      ea.Q = adj_qview(h);
      ea.X = h->d_Xs;
      ea.acoef = h->d_acoef + coff[pi];
      ea.J = p.d_J;
      ea.theta = p.d_theta;
      ea.Ne = p.Ne; ea.P = p.P; ea.M = p.M; ea.ntheta = p.ntheta;
      ea.k = p.k; ea.cE = p.cE; ea.cM = p.cM;
      ea.out = h->d_adj_ex;
      hipLaunchKernelGGL(k_adj_excite, dim3((unsigned)((p.ntheta + 127) / 128), (unsigned)p.Ne), dim3(128), 0,
                         h->stream, ea);
      LAUNCHCHK();
    }
    double rmax = 0.0;
    for (int q = 0; q < p.Nn; ++q) rmax = std::max(rmax, p.nodes[2 * q]);
    hipLaunchKernelGGL(k_adj_gather, dim3((unsigned)p.Nn), dim3(256), 0, h->stream, h->d_adj_far,
                       h->d_adj_near, p.d_batches, reinterpret_cast<const AdjNearPair*>(p.d_pairs),
                       (int)p.batches.size(), h->d_adj_ex, p.Ne, p.ntheta, p.d_nodes,
                       1e-12 * std::max(1.0, rmax), h->d_adj_grad + goff);
    LAUNCHCHK();
    goff += (size_t)p.Nn;
  }
  return 0;
}

extern "C" int ghost_bor_get_gradient(ghost_bor* h, double* grad);

extern "C" int ghost_bor_adjoint_analytic(ghost_bor* h, const double* acoef, double* grad) {
  if (!h || !acoef || !grad) return fail(GHOST_ERR_ARG, "null argument");
  const bool prof = h->profile;
  h->profile = false;
  std::vector<size_t> coff;
  int rc;
  if ((rc = adjoint_checks(h, true)) || (rc = adjoint_alloc(h)) || (rc = adjoint_host_acoef(h, acoef, coff)) ||
      (rc = adjoint_lambda_q(h, coff)) || (rc = adjoint_analytic_kernels(h, coff))) {
    h->profile = prof;
    return rc;
  }
  h->profile = prof;
  return ghost_bor_get_gradient(h, grad);
}

// Objective J = sum_c w_c sigma_dBsm(c) over every problem column for the
// captured forward + adjoint: a_c = w_c (20 / ln 10) conj(F_c) / |F_c|^2 is
// then formed on the device from the far field of the same launch.
extern "C" int ghost_bor_set_objective(ghost_bor* h, const double* weights) {
  if (!h || !weights) return fail(GHOST_ERR_ARG, "null argument");
  if (!h->built) return fail(GHOST_ERR_STATE, "build first");
  HIPCHK(hipSetDevice(h->device));
  size_t ncol = 0;
  for (auto& p : h->probs) ncol += (size_t)p.nrhs;
  int rc;
  if (!h->d_objw && (rc = dalloc(h, &h->d_objw, std::max<size_t>(ncol, 1) * sizeof(double)))) return rc;
  HIPCHK(hipMemcpy(h->d_objw, weights, ncol * sizeof(double), hipMemcpyHostToDevice));
  h->has_objective = true;
  return 0;
}

// Capture fill + solve (refine_iters fixed sweeps) + far field + the
// objective's a' + lambda, Q and the analytic gradient into the handle's
// graph: ghost_bor_launch then runs the forward RCS and its gradient in one
// graph launch; ghost_bor_get_dF and ghost_bor_get_gradient read the results.
extern "C" int ghost_bor_capture_adjoint(ghost_bor* h, int refine_iters) {
  int rc;
  if ((rc = adjoint_checks(h, false))) return rc;
  if (!h->has_objective) return fail(GHOST_ERR_STATE, "set the objective before capturing the adjoint");
  if (refine_iters < 0) return fail(GHOST_ERR_ARG, "a captured solve takes a fixed number of refinement sweeps");
  if ((rc = adjoint_alloc(h))) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->gexec) { hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  if (h->graph) { hipGraphDestroy(h->graph); h->graph = nullptr; }
  std::vector<size_t> coff(h->probs.size());
  {
    size_t c = 0;
    for (size_t pi = 0; pi < h->probs.size(); ++pi) { coff[pi] = c; c += (size_t)h->probs[pi].nrhs; }
  }
  const bool prof = h->profile;
  h->profile = false;
  h->capturing = true;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  rc = run_fill(h);
  if (!rc) rc = ghost_bor_solve(h, refine_iters);
  if (!rc) rc = ghost_bor_farfield(h);
  if (!rc) {
    for (size_t pi = 0; pi < h->probs.size(); ++pi) {
      const Problem& p = h->probs[pi];
      hipLaunchKernelGGL(k_objective_acoef, dim3((unsigned)((p.nrhs + 127) / 128)), dim3(128), 0, h->stream,
                         h->d_dF + p.dF_off, p.nmodes, p.nrhs, h->d_objw + coff[pi],
                         -p.k * 376.730313668 / (4.0 * M_PI), h->d_acoef + coff[pi]);
    }
    if (hipGetLastError() != hipSuccess) rc = fail(GHOST_ERR_HIP, "objective kernel launch failed");
  }
  if (!rc) rc = adjoint_lambda_q(h, coff);
  if (!rc) rc = adjoint_analytic_kernels(h, coff);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->capturing = false;
  h->profile = prof;
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(GHOST_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  h->graph = g;
  HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  return 0;
}

// dJ/d rho of every node of every problem (problem order) from the last
// analytic adjoint, eager or captured.
extern "C" int ghost_bor_get_gradient(ghost_bor* h, double* grad) {
  if (!h || !grad) return fail(GHOST_ERR_ARG, "null argument");
  if (!h->adj_alloc) return fail(GHOST_ERR_STATE, "no adjoint has run on this handle");
  HIPCHK(hipSetDevice(h->device));
  size_t nn = 0;
  for (auto& p : h->probs) nn += (size_t)p.Nn;
  HIPCHK(hipMemcpyAsync(grad, h->d_adj_grad, nn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

// ─────────────────────────────────────────────────────────────────────────────
// measurement utility: a plain device copy of known size (byte-counter calibration)
// ─────────────────────────────────────────────────────────────────────────────

__global__ void k_copy_bytes(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

extern "C" int ghost_copy_calibration(double bytes, float* ms) {
  if (!ms || !(bytes >= 16.0)) return fail(GHOST_ERR_ARG, "bytes must be >= 16");
  const size_t n = (size_t)(bytes / sizeof(float4));
  float4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
    rc = fail(GHOST_ERR_OOM, "copy buffers");
  } else if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = fail(GHOST_ERR_HIP, "stream / events");
  } else {
    hipMemsetAsync(a, 1, n * sizeof(float4), s);
    hipLaunchKernelGGL(k_copy_bytes, dim3(8192), dim3(256), 0, s, a, b, n);   // warm
    hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_copy_bytes, dim3(8192), dim3(256), 0, s, a, b, n);
    hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      rc = fail(GHOST_ERR_HIP, "copy timing");
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (a) hipFree(a);
  if (b) hipFree(b);
  return rc;
}

// ─────────────────────────────────────────────────────────────────────────────
// caller-filled dense systems
// ─────────────────────────────────────────────────────────────────────────────

static int dense_sys(ghost_bor* h, int idx, const SysInfo** out) {
  if (!h || !h->built) return fail(GHOST_ERR_STATE, "build first");
  if (idx < 0 || idx >= (int)h->sys.size() || h->sys[idx].prob >= 0)
    return fail(GHOST_ERR_ARG, "%d is not a dense system of this handle", idx);
  *out = &h->sys[idx];
  return 0;
}

// A: n x n, B: n x nrhs, complex row-major interleaved doubles.
extern "C" int ghost_bor_set_dense(ghost_bor* h, int idx, const double* A, const double* B) {
  const SysInfo* sp;
  int rc = dense_sys(h, idx, &sp);
  if (rc) return rc;
  if (!A || !B) return fail(GHOST_ERR_ARG, "null matrix");
  const SysInfo& s = *sp;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  const bool mixed = h->precision == GHOST_MIXED;
  const size_t csz = mixed ? sizeof(float2) : sizeof(double2);
  std::vector<char> row((size_t)s.ld * csz, 0);
  for (int r = 0; r < s.n; ++r) {
    for (int c = 0; c < s.n + s.nrhs; ++c) {
      const double* src = c < s.n ? A + 2 * ((size_t)r * s.n + c) : B + 2 * ((size_t)r * s.nrhs + (c - s.n));
      if (mixed) ((float2*)row.data())[c] = make_float2((float)src[0], (float)src[1]);
      else ((double2*)row.data())[c] = make_double2(src[0], src[1]);
    }
    HIPCHK(hipMemcpy((char*)h->d_A + ((size_t)s.a_off + (size_t)r * s.ld) * csz, row.data(), (size_t)s.ld * csz,
                     hipMemcpyHostToDevice));
  }
  if (h->d_A64) {
    HIPCHK(hipMemcpy(h->d_A64 + s.a64_off, A, (size_t)s.n * s.n * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_B64 + s.x_off, B, (size_t)s.n * s.nrhs * sizeof(double2), hipMemcpyHostToDevice));
  }
  h->solved = false;
  return 0;
}

// X: n x nrhs after ghost_bor_solve.
extern "C" int ghost_bor_get_dense(ghost_bor* h, int idx, double* X) {
  const SysInfo* sp;
  int rc = dense_sys(h, idx, &sp);
  if (rc) return rc;
  if (!X) return fail(GHOST_ERR_ARG, "null output");
  const SysInfo& s = *sp;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->precision == GHOST_MIXED) {
    HIPCHK(hipMemcpy(X, h->d_X64 + s.x_off, (size_t)s.n * s.nrhs * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<double2> row(s.ld);
  for (int r = 0; r < s.n; ++r) {
    HIPCHK(hipMemcpy(row.data(), (double2*)h->d_A + (size_t)s.a_off + (size_t)r * s.ld, (size_t)s.ld * sizeof(double2),
                     hipMemcpyDeviceToHost));
    memcpy(X + 2 * (size_t)r * s.nrhs, row.data() + s.n, (size_t)s.nrhs * sizeof(double2));
  }
  return 0;
}

// Row interchanges of system idx's factorization: ipiv[r] = the row swapped
// with row r at step r (0-based within the system), LAPACK getrf's ipiv - 1.
// Mode -m of a pair reports mode m's (S A_m S pivots on the same rows).
extern "C" int ghost_bor_get_pivots(ghost_bor* h, int idx, int* ipiv) {
  if (!h || !ipiv) return fail(GHOST_ERR_ARG, "null argument");
  if (idx < 0 || idx >= (int)h->sys.size()) return fail(GHOST_ERR_ARG, "system index %d out of range", idx);
  if (!h->solved) return fail(GHOST_ERR_STATE, "no factorization yet: solve or run first");
  long long off = 0;   // pivots are stored per solved system (a pair shares mode m's)
  for (int si = 0; si < h->sys[idx].owner; ++si) off += h->ss[si].n;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(ipiv, h->d_piv + off, (size_t)h->sys[idx].n * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

// Number of systems (mode systems of every problem and dense systems): the
// length of the per-system arrays of ghost_bor_residual and ghost_bor_info.
extern "C" int ghost_bor_system_count(ghost_bor* h, int* n) {
  if (!h || !n) return fail(GHOST_ERR_ARG, "null argument");
  *n = (int)h->sys.size();
  return 0;
}
