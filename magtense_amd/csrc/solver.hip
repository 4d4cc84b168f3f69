// Batched device-resident micromagnetic solver: RKSUITE method 'M'
// (Bogacki-Shampine 4(5), FSAL, range mode with the degree-6 interpolant) for
// B independent magnetization states that share one demag tensor.
//
// Reference behaviour restated here:
//   driver        source/NumericalIntegration/NumericalIntegration/ODE_Solvers.f90:124-237
//   RKSUITE 'M'   source/NumericalIntegration/NumericalIntegration/RKSuitef90.f90
//                 setup :596-856, range_integrate :942-1183, step_integrate :1185-1587,
//                 step :1726-1978, stiff :1981-2367, interpolate :2655-2997
//   solver loop   source/MagTenseMicroMag/LandauLifshitzEquationSolver.f90:50-209
//   field storage LandauLifshitzEquationSolver.f90:306-349
//
// Design.  Each state's step control (scalars: t, h, error norms, counters)
// runs on the host as a C++20 coroutine written like the serial algorithm.
// Whenever the algorithm needs device work - a right-hand side at a stage
// argument, a linear combination, a weighted norm - the coroutine suspends
// with one operation.  The scheduler gathers the operations of all states
// into one "tick": every right-hand side of the tick shares ONE pass over the
// demag tensor (a batched MFMA GEMM, or the per-state FP32 mat-vec), and the
// vector work of all states runs in one launch.  States that accept, reject
// or interpolate at different moments simply issue different operations; the
// common case keeps them in lockstep (seven right-hand sides per step).
// Per-state vectors never leave HBM; only the scalars the control needs are
// read back (two small copies per step).

#include <algorithm>
#include <cmath>
#include <complex>
#include <coroutine>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mt_common.h"
#include "mt_kernels.h"
#include "../../include/magtense_hip.h"

namespace mt {



namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kMu0 = 4.0 * kPi * 1e-7;

// ---------------------------------------------------------------------------
// Coroutine plumbing
// ---------------------------------------------------------------------------
template <typename T>
struct Task {
  struct promise_type {
    T value{};
    std::coroutine_handle<> continuation{};
    Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    struct Final {
      bool await_ready() noexcept { return false; }
      std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
        auto c = h.promise().continuation;
        return c ? c : std::noop_coroutine();
      }
      void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void return_value(T v) { value = v; }
    void unhandled_exception() { std::terminate(); }
  };
  std::coroutine_handle<promise_type> h;
  explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
  Task(Task&& o) noexcept : h(o.h) { o.h = nullptr; }
  Task(const Task&) = delete;
  ~Task() {
    if (h) h.destroy();
  }
  bool await_ready() { return false; }
  std::coroutine_handle<> await_suspend(std::coroutine_handle<> parent) {
    h.promise().continuation = parent;
    return h;
  }
  T await_resume() { return h.promise().value; }
};

// ---------------------------------------------------------------------------
// Per-state control data (mirrors rk_comm_real_1d for one state)
// ---------------------------------------------------------------------------
struct Op {
  enum Kind { kNone, kEval, kVec, kPause } kind = kNone;
  VecDesc pre{};      // kEval: combination forming the stage argument (+ optional reduction)
  VecDesc post{};     // kEval: reduction after the right-hand side (rkind may be none)
  double t = 0;       // kEval: stage time
  double* out = nullptr;
  double* fields = nullptr;
  VecDesc vec{};      // kVec
  double res[4] = {0, 0, 0, 0};
  double post_res[4] = {0, 0, 0, 0};
  bool defer_pre = false;  // kEval: the pre-reduction's result may wait for the next synchronising tick
};

struct Sim {
  int id = 0;
  // RKSUITE state
  double t = 0, t_old = 0, t_start = 0, t_end = 0, dir = 1, h = 0, h_old = 0, h_start = 0, h_average = 0, tol = 0;
  double errold = 0, tlast = 0, range_t_end = 0;
  long long f_count = 0, full_f_count = 0, step_count = 0, bad_step_count = 0, stiff_bad_step_count = 0;
  long long hit_t_end_count = 0;
  bool at_t_start = true, at_t_end = false, chkeff = false, phase2 = true;
  int st_range = -1, st_interp = -1;
  // device vectors (3n doubles each); stage[1..7]; p[1..5]
  double *y = nullptr, *yp = nullptr, *y_old = nullptr, *scratch = nullptr, *errest = nullptr, *weights = nullptr;
  double* stage[8] = {};
  double *xstage = nullptr, *ytemp = nullptr, *arg = nullptr, *y_step = nullptr, *y_last = nullptr;
  double* p[6] = {};
  // applied field table for this state (dynamic solver) or rows (explicit solver)
  std::vector<double> ht, hx, hy, hz;
  int hext_row = 0;
  // totals over all field rows
  long long tot_f = 0, tot_steps = 0, tot_bad = 0, flag3 = 0, flag4 = 0;
  long long steps_run = 0;   // accepted steps since creation (for the step budget)
  long long budget = 0;      // pause when steps_run >= budget (0: no limit)
  // scheduling
  Op op;
  std::coroutine_handle<> resume_point{};
  std::unique_ptr<Task<int>> top;
  bool started = false, done = false;
  int error = 0;
  std::string error_msg;
  long long evals = 0;
};

struct OpAwait {
  Sim* s;
  bool await_ready() { return false; }
  void await_suspend(std::coroutine_handle<> h) { s->resume_point = h; }
  void await_resume() {}
};

VecDesc blank() {
  VecDesc d;
  std::memset(&d, 0, sizeof(d));
  d.bcoef = 1.0;
  d.post_scale = 1.0;
  d.pack_col = -1;
  d.rscale = 1.0;
  return d;
}

void add_term(VecDesc& d, const double* v, double c) {
  d.src[d.nterm] = v;
  d.coef[d.nterm] = c;
  d.nterm += 1;
}

VecDesc copy_desc(double* dst, const double* src) {
  VecDesc d = blank();
  d.dst = dst;
  d.base = src;
  return d;
}

// RKSUITE 'M' tables (RKSuitef90.f90:225-412); a(i,j) -> A[i][j] (1-based)
struct Tables {
  double A[12][12] = {};
  double B[9] = {}, Cc[12] = {}, E[8] = {}, R[12][7] = {};
  double cost, safety, expon, stability_radius, tan_angle, rs, rs1, rs2, rs3, rs4, toosml;
  double round_off, sqrrmc, cubrmc, sqtiny;
  int max_stiff_iters;
  Tables() {
    auto& a = A;
    a[2][1] = 1.0 / 6.0;
    a[3][1] = 2.0 / 27.0; a[3][2] = 4.0 / 27.0;
    a[4][1] = 183.0 / 1372.0; a[4][2] = -162.0 / 343.0; a[4][3] = 1053.0 / 1372.0;
    a[5][1] = 68.0 / 297.0; a[5][2] = -4.0 / 11.0; a[5][3] = 42.0 / 143.0; a[5][4] = 1960.0 / 3861.0;
    a[6][1] = 597.0 / 22528.0; a[6][2] = 81.0 / 352.0; a[6][3] = 63099.0 / 585728.0; a[6][4] = 58653.0 / 366080.0;
    a[6][5] = 4617.0 / 20480.0;
    a[7][1] = 174197.0 / 959244.0; a[7][2] = -30942.0 / 79937.0; a[7][3] = 8152137.0 / 19744439.0;
    a[7][4] = 666106.0 / 1039181.0; a[7][5] = -29421.0 / 29068.0; a[7][6] = 482048.0 / 414219.0;
    a[8][1] = 587.0 / 8064.0; a[8][2] = 0.0; a[8][3] = 4440339.0 / 15491840.0; a[8][4] = 24353.0 / 124800.0;
    a[8][5] = 387.0 / 44800.0; a[8][6] = 2152.0 / 5985.0; a[8][7] = 7267.0 / 94080.0;
    B[1] = 2479.0 / 34992.0; B[2] = 0.0; B[3] = 123.0 / 416.0; B[4] = 612941.0 / 3411720.0; B[5] = 43.0 / 1440.0;
    B[6] = 2272.0 / 6561.0; B[7] = 79937.0 / 1113912.0; B[8] = 3293.0 / 556956.0;
    E[1] = -3.0 / 1280.0; E[2] = 0.0; E[3] = 6561.0 / 632320.0; E[4] = -343.0 / 20800.0; E[5] = 243.0 / 12800.0;
    E[6] = -1.0 / 95.0; E[7] = 0.0;
    Cc[1] = 0.0; Cc[2] = 1.0 / 6.0; Cc[3] = 2.0 / 9.0; Cc[4] = 3.0 / 7.0; Cc[5] = 2.0 / 3.0; Cc[6] = 3.0 / 4.0;
    Cc[7] = 1.0; Cc[8] = 1.0; Cc[9] = 1.0 / 2.0; Cc[10] = 5.0 / 6.0; Cc[11] = 1.0 / 9.0;
    a[9][1] = 455.0 / 6144.0; a[10][1] = -837888343715.0 / 13176988637184.0;
    a[11][1] = 98719073263.0 / 1551965184000.0;
    a[9][2] = 0.0; a[10][2] = 30409415.0 / 52955362.0; a[11][2] = 1307.0 / 123552.0;
    a[9][3] = 10256301.0 / 35409920.0; a[10][3] = -48321525963.0 / 759168069632.0;
    a[11][3] = 4632066559387.0 / 70181753241600.0;
    a[9][4] = 2307361.0 / 17971200.0; a[10][4] = 8530738453321.0 / 197654829557760.0;
    a[11][4] = 7828594302389.0 / 382182512025600.0;
    a[9][5] = -387.0 / 102400.0; a[10][5] = 1361640523001.0 / 1626788720640.0; a[11][5] = 40763687.0 / 11070259200.0;
    a[9][6] = 73.0 / 5130.0; a[10][6] = -13143060689.0 / 38604458898.0; a[11][6] = 34872732407.0 / 224610586200.0;
    a[9][7] = -7267.0 / 215040.0; a[10][7] = 18700221969.0 / 379584034816.0; a[11][7] = -2561897.0 / 30105600.0;
    a[9][8] = 1.0 / 32.0; a[10][8] = -5831595.0 / 847285792.0; a[11][8] = 1.0 / 10.0;
    a[10][9] = -5183640.0 / 26477681.0; a[11][9] = -1.0 / 10.0;
    a[11][10] = -1403317093.0 / 11371610250.0;
    auto& r = R;
    r[1][6] = -12134338393.0 / 1050809760.0; r[1][5] = -1620741229.0 / 50038560.0;
    r[1][4] = -2048058893.0 / 59875200.0; r[1][3] = -87098480009.0 / 5254048800.0;
    r[1][2] = -11513270273.0 / 3502699200.0;
    r[3][6] = -33197340367.0 / 1218433216.0; r[3][5] = -539868024987.0 / 6092166080.0;
    r[3][4] = -39991188681.0 / 374902528.0; r[3][3] = -69509738227.0 / 1218433216.0;
    r[3][2] = -29327744613.0 / 2436866432.0;
    r[4][6] = -284800997201.0 / 19905339168.0; r[4][5] = -7896875450471.0 / 165877826400.0;
    r[4][4] = -333945812879.0 / 5671036800.0; r[4][3] = -16209923456237.0 / 497633479200.0;
    r[4][2] = -2382590741699.0 / 331755652800.0;
    r[5][6] = -540919.0 / 741312.0; r[5][5] = -103626067.0 / 43243200.0; r[5][4] = -633779.0 / 211200.0;
    r[5][3] = -32406787.0 / 18532800.0; r[5][2] = -36591193.0 / 86486400.0;
    r[6][6] = 7157998304.0 / 374350977.0; r[6][5] = 30405842464.0 / 623918295.0; r[6][4] = 183022264.0 / 5332635.0;
    r[6][3] = -3357024032.0 / 1871754885.0; r[6][2] = -611586736.0 / 89131185.0;
    r[7][6] = -138073.0 / 9408.0; r[7][5] = -719433.0 / 15680.0; r[7][4] = -1620541.0 / 31360.0;
    r[7][3] = -385151.0 / 15680.0; r[7][2] = -65403.0 / 15680.0;
    r[8][6] = 1245.0 / 64.0; r[8][5] = 3991.0 / 64.0; r[8][4] = 4715.0 / 64.0; r[8][3] = 2501.0 / 64.0;
    r[8][2] = 149.0 / 16.0; r[8][1] = 1.0;
    r[9][6] = 55.0 / 3.0; r[9][5] = 71.0; r[9][4] = 103.0; r[9][3] = 199.0 / 3.0; r[9][2] = 16.0;
    r[10][6] = -1774004627.0 / 75810735.0; r[10][5] = -1774004627.0 / 25270245.0; r[10][4] = -26477681.0 / 359975.0;
    r[10][3] = -11411880511.0 / 379053675.0; r[10][2] = -423642896.0 / 126351225.0;
    r[11][6] = 35.0; r[11][5] = 105.0; r[11][4] = 117.0; r[11][3] = 59.0; r[11][2] = 12.0;
    cost = 7.0;
    safety = 0.8;
    expon = 1.0 / 5.0;
    stability_radius = 3.9;
    tan_angle = 5.2;
    max_stiff_iters = std::min(8, std::max(1, (int)(0.05 * cost * 50.0)));
    rs = 2.0; rs1 = 1.0 / rs; rs2 = rs * rs; rs3 = rs * rs2; rs4 = 1.0 / rs3;
    double cdiff = 1.0;
    for (int i = 1; i <= 7; ++i)
      for (int j = i + 1; j <= 8; ++j) {
        double d = Cc[i] - Cc[j];
        if (d != 0) cdiff = std::min(cdiff, std::fabs(d));
      }
    const double eps = 2.220446049250313e-16;
    round_off = 10.0 * eps;
    sqrrmc = std::sqrt(eps);
    cubrmc = std::pow(eps, 1.0 / 3.0);
    sqtiny = std::sqrt(2.2250738585072014e-308);
    toosml = round_off / cdiff;
  }
};

const Tables& tables() {
  static Tables T;
  return T;
}

// interp1_MagTense / locate (util.f90:39-85); clamps right of the table.
double interp1(const std::vector<double>& xs, const std::vector<double>& ys, double x) {
  int n = (int)xs.size();
  if (n == 0) return 0.0;
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if ((xs[n - 1] >= xs[0]) == (x >= xs[jm - 1])) jl = jm;
    else ju = jm;
  }
  int j;
  if (x == xs[0]) j = 1;
  else if (x == xs[n - 1]) j = n - 1;
  else j = jl;
  if (j == 0) return 0.0;
  if (j >= n) return ys[n - 1];
  double xl = (x - xs[j - 1]) / (xs[j] - xs[j - 1]);
  return (1 - xl) * ys[j - 1] + xl * ys[j];
}

}  // namespace

// ---------------------------------------------------------------------------
// Solver context
// ---------------------------------------------------------------------------
struct Solver {
  // problem
  int nx = 0, ny = 0, nz = 0, n = 0, B = 0;
  double dx = 0, dy = 0, dz = 0;
  int solver = 2;          // 1 explicit, 2 dynamic
  int precision = 0;       // MT_PREC_FP32 / MT_PREC_BF16X2
  double gamma = 0, alpha0 = 0, tol = 1e-4, thres = 1e-6, conv_tol = 1e-4;
  std::vector<double> alpha_t, alpha_v;
  std::vector<double> t_req, t_conv;
  int nt = 0, n_h = 1;
  bool return_fields = false;
  // device data
  hipStream_t stream = nullptr;
  float* K32 = nullptr;
  uint8_t* Ksplit = nullptr;
  uint8_t* Kf32 = nullptr;  // MT_PREC_FP32_MFMA: the dense tensor as FP32 MFMA fragments
  int *xs = nullptr, *xe = nullptr, *xc = nullptr;  // loaded exchange operator (unstructured grids)
  double* xv = nullptr;
  long long exch_nnz_host = 0;
  uint8_t* Ktab = nullptr;    // translation-invariant fragment table (bf16x2 on uniform grids)
  bool defer_early = false;   // read the early error estimate with stage 8's (cheap right-hand sides)
  bool stage_fusion = true;   // small FP32 dense problems: stage argument formed in the demag + LL launch
  bool near_llg = true;       // split product: the LL right-hand side evaluated by the near-band kernel
  uint32_t* near_tab = nullptr;  // near-band correction table (bf16x2 on uniform grids, dense tensor)
  int near_r = 0;             // its radius (0: no correction)
  int demag_approx = 1;       // 1 dense, 2 threshold, 4 threshold fraction
  SparseDemag sparse;
  double *jfact = nullptr, *kfact = nullptr, *aniso = nullptr;
  float* ms32 = nullptr;
  double* pool = nullptr;     // state vectors
  double* yout = nullptr;     // [B][n_h][nt][3n]
  double* fout = nullptr;     // [B][n_h][nt][12n]
  float* packf = nullptr;     // [B][3n]
  uint8_t* packb = nullptr;   // fragment layout, 64 columns
  float* hbuf = nullptr;      // [ksplit][cols][3n]
  int ksplit = 1;
  int hcols = 0;
  double* partials = nullptr;
  double* results = nullptr;  // device [3][B*4]
  double* h_results = nullptr;  // pinned
  uint8_t* arena_h = nullptr;   // pinned descriptor arena
  uint8_t* arena_d = nullptr;
  size_t arena_size = 0, arena_off = 0;
  std::vector<double> t_out_h;  // [B][n_h][nt]
  LlgGrid grid{};
  std::vector<Sim> sims;
  std::vector<double> m0;       // [B][3n]
  // statistics
  long long ticks = 0, eval_ticks = 0, demag_passes = 0;
  double gemm_ms = 0.0;
  double near_ms = 0.0;  // of gemm_ms: the near-band correction
  // timing mode: per-kernel event times and algorithmic bytes of the LL
  // right-hand side and of the stepper's vector operations
  double vec_ms = 0.0, llg_ms = 0.0, vec_bytes = 0.0, llg_bytes = 0.0;
  long long vec_launches = 0, llg_launches = 0;
  bool timing = false;
  long long total_target = 0;   // > 0: stop once the states' accepted steps sum to this
  hipEvent_t ev[8] = {};  // timing: vec-only, pre-vec, GEMM, LL kernel, post-vec boundaries, near band

  ~Solver() { release(); }

  void release() {
    auto f = [](void* p) {
      if (p) (void)hipFree(p);
    };
    f(K32); f(Ksplit); f(Kf32); f(Ktab); f(near_tab); f(jfact); f(kfact); f(aniso); f(ms32); f(pool); f(yout); f(fout); f(packf); f(packb);
    f(hbuf); f(partials); f(results); f(arena_d); f(xs); f(xe); f(xc); f(xv);
    sparse.release();
    if (h_results) (void)hipHostFree(h_results);
    if (arena_h) (void)hipHostFree(arena_h);
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    xs = xe = xc = nullptr; xv = nullptr;
    K32 = nullptr; Ksplit = nullptr; Kf32 = nullptr; Ktab = nullptr; near_tab = nullptr; jfact = kfact = aniso = nullptr; ms32 = nullptr; pool = yout = fout = nullptr;
    packf = nullptr; packb = nullptr; hbuf = nullptr; partials = results = nullptr; arena_d = nullptr;
    h_results = nullptr; arena_h = nullptr; stream = nullptr;
    for (hipEvent_t& e : ev) e = nullptr;
  }

  size_t n3() const { return (size_t)3 * n; }

  // Algorithmic HBM bytes of one state's vector operation: every distinct
  // vector read or written once (a reduction operand that the combination just
  // wrote is not read again), plus the demag input pack (BF16 hi+lo or FP32).
  double vec_desc_bytes(const VecDesc& d, int pack_mode) const {
    const void* rd[kMaxTerms + 16];
    const void* wr[4];
    int nr = 0, nw = 0;
    auto add = [](const void** set, int& k, const void* p) {
      if (!p) return;
      for (int i = 0; i < k; ++i)
        if (set[i] == p) return;
      set[k++] = p;
    };
    if (d.dst) {
      add(wr, nw, d.dst);
      add(rd, nr, d.base);
      for (int k = 0; k < d.nterm; ++k) add(rd, nr, d.src[k]);
    }
    auto rd_red = [&](const void* p) {
      if (p && p != d.dst) add(rd, nr, p);
    };
    switch (d.rkind) {
      case kRedStepB:
      case kRedFinal:
        for (int k = 0; k < d.rn; ++k) rd_red(d.rsrc[k]);
        rd_red(d.ra); rd_red(d.rb);
        if (d.rkind == kRedFinal) { add(wr, nw, d.rd); add(wr, nw, d.rw); }
        break;
      case kRedStepA:
        rd_red(d.ra); rd_red(d.rb); rd_red(d.rc); rd_red(d.rd); rd_red(d.rw); add(wr, nw, d.rw);
        break;
      case kRedYpNorm:
      case kRedWeightsAvg:
        rd_red(d.ra); rd_red(d.rb); add(wr, nw, d.rw);
        break;
      case kRedWdot:
        rd_red(d.rw);
        for (int k = 0; k < 2 * d.rn && k < 8; ++k) rd_red(d.rsrc[k]);
        break;
      case kRedMaxDiff:
        rd_red(d.ra); rd_red(d.rb);
        break;
      default:
        break;
    }
    double bytes = 8.0 * (double)n3() * (nr + nw);
    if (d.dst && d.pack_col >= 0 && pack_mode != kPackNone) bytes += 4.0 * (double)n3();
    return bytes;
  }

  // Algorithmic HBM bytes of one LL-kernel launch over S states: per state the
  // argument, the demag sums (FP32 partials or FP64) and dm/dt (plus the four
  // stored fields when asked); per launch the per-cell coefficient arrays.
  double llg_launch_bytes(const std::vector<LlgDesc>& ld) const {
    const double c3 = (double)n3();
    double bytes = 8.0 * n + 4.0 * n;  // jfact, ms32
    if (grid.aniso) bytes += 48.0 * n + 8.0 * n;
    if (grid.xs) bytes += 8.0 * n + 12.0 * (double)exch_nnz_host;
    for (const LlgDesc& L : ld) {
      bytes += 8.0 * c3;
      bytes += L.hpart64 ? 8.0 * c3 + 8.0 * n : 4.0 * c3 * L.nsplit;
      if (L.out) bytes += 8.0 * c3;
      if (L.fields) bytes += 32.0 * c3;
    }
    return bytes;
  }

  // descriptor arena: pinned host staging + device copy; reset after syncs
  template <typename T>
  int stage_descs(const std::vector<T>& v, const T** dptr) {
    size_t bytes = v.size() * sizeof(T);
    if (bytes > arena_size) {
      set_error("descriptor arena too small for one launch");
      return kErrArg;
    }
    size_t off = round_up((int64_t)arena_off, 256);
    if (off + bytes > arena_size) {
      MT_HIP(hipStreamSynchronize(stream));
      off = 0;
    }
    std::memcpy(arena_h + off, v.data(), bytes);
    MT_HIP(hipMemcpyAsync(arena_d + off, arena_h + off, bytes, hipMemcpyHostToDevice, stream));
    *dptr = reinterpret_cast<const T*>(arena_d + off);
    arena_off = off + bytes;
    return kOk;
  }

  // Several descriptor arrays with one host -> device copy: stage_reserve makes
  // room for `bytes` (plus alignment) without a reset in between, stage_put
  // lays an array out in the pinned arena, stage_flush copies what was put
  // since the reservation.
  size_t stage_from = 0;
  int stage_reserve(size_t bytes) {
    if (bytes > arena_size) {
      set_error("descriptor arena too small for one demag pass");
      return kErrArg;
    }
    if (round_up((int64_t)arena_off, 256) + bytes > arena_size) {
      MT_HIP(hipStreamSynchronize(stream));
      arena_off = 0;
    }
    stage_from = round_up((int64_t)arena_off, 256);
    return kOk;
  }
  template <typename T>
  const T* stage_put(const std::vector<T>& v) {
    const size_t off = round_up((int64_t)arena_off, 256);
    std::memcpy(arena_h + off, v.data(), v.size() * sizeof(T));
    arena_off = off + v.size() * sizeof(T);
    return reinterpret_cast<const T*>(arena_d + off);
  }
  int stage_flush() {
    if (arena_off > stage_from)
      MT_HIP(hipMemcpyAsync(arena_d + stage_from, arena_h + stage_from, arena_off - stage_from,
                            hipMemcpyHostToDevice, stream));
    return kOk;
  }

  double alpha_at(double t) const { return alpha0 == 0 ? interp1(alpha_t, alpha_v, t) : alpha0; }

  void hext_at(const Sim& s, double t, double out[3]) const {
    if (solver == 1) {
      out[0] = s.hx[s.hext_row]; out[1] = s.hy[s.hext_row]; out[2] = s.hz[s.hext_row];
    } else {
      out[0] = interp1(s.ht, s.hx, t); out[1] = interp1(s.ht, s.hy, t); out[2] = interp1(s.ht, s.hz, t);
    }
  }

  // ---- operations issued by the per-state coroutines ----------------------
  OpAwait eval(Sim& s, double t, const VecDesc& pre, double* out, const VecDesc& post, double* fields = nullptr,
               bool defer_pre = false) {
    s.op = Op();
    s.op.kind = Op::kEval;
    s.op.pre = pre;
    s.op.post = post;
    s.op.t = t;
    s.op.out = out;
    s.op.fields = fields;
    s.op.defer_pre = defer_pre;
    return OpAwait{&s};
  }
  OpAwait vec(Sim& s, const VecDesc& d) {
    s.op = Op();
    s.op.kind = Op::kVec;
    s.op.vec = d;
    return OpAwait{&s};
  }
  OpAwait pause(Sim& s) {
    s.op = Op();
    s.op.kind = Op::kPause;
    return OpAwait{&s};
  }

  // ---- RKSUITE 'M' as coroutines ----------------------------------------------
  int setup(Sim& s, double t_start, double t_end) {
    // RKSuitef90.f90:596-856 (method 'M', task 'R', no error assessment, h_start = 0)
    const Tables& T = tables();
    s.h_start = 0.0; s.h_old = 0.0; s.f_count = 0; s.full_f_count = 0; s.step_count = 0; s.bad_step_count = 0;
    s.at_t_start = true; s.at_t_end = false; s.stiff_bad_step_count = 0; s.hit_t_end_count = 0;
    s.errold = 0.0; s.h_average = 0.0; s.chkeff = false; s.phase2 = true;
    s.st_range = -1; s.st_interp = -1;
    if (t_start == t_end) return fail(s, "T_START = T_END");
    s.t_end = t_end; s.t_start = t_start; s.t_old = t_start; s.t = t_start;
    s.dir = (t_end - t_start) >= 0 ? 1.0 : -1.0;
    if (tol > 0.01 || tol < T.round_off) return fail(s, "tolerance outside the permitted range [10*eps, 0.01]");
    s.tol = tol;
    if (thres < T.sqtiny) return fail(s, "threshold below sqrt(tiny)");
    double hmin = std::max(T.sqtiny, T.toosml * std::max(std::fabs(t_start), std::fabs(t_end)));
    if (std::fabs(t_end - t_start) < hmin) return fail(s, "T_END and T_START not clearly distinguishable");
    s.h = std::max(std::fabs(t_end - t_start) / T.rs3, hmin);
    return kOk;
  }

  int fail(Sim& s, const std::string& msg) {
    s.error = kErrSolver;
    s.error_msg = "state " + std::to_string(s.id) + ": " + msg;
    return kErrSolver;
  }

  // step_r1 (RKSuitef90.f90:1726-1978) for the primary integration
  Task<int> step_r1(Sim& s, double tnow, double& htry, double& err, double hmin) {
    const Tables& T = tables();
    for (;;) {  // attempt_step
      if (s.phase2) {
        VecDesc w = blank();
        w.rkind = kRedYpNorm;  // weights = max(|y|, thresh)
        w.ra = s.y; w.rb = s.yp; w.rw = s.weights; w.thresh = thres;
        co_await vec(s, w);
      }
      bool restart = false;
      for (int i = 2; i <= 8; ++i) {
        VecDesc pre = blank();
        pre.dst = s.scratch;  // y_new
        pre.base = s.y;
        add_term(pre, s.yp, htry * T.A[i][1]);
        for (int j = 2; j <= i - 1; ++j)
          if (T.A[i][j] != 0.0) add_term(pre, s.stage[j - 1], htry * T.A[i][j]);
        if (i == 7) {  // stepb (RKSuitef90.f90:1957-1976)
          pre.rkind = kRedStepB;
          pre.rn = 5;
          pre.rsrc[0] = s.yp; pre.rcoef[0] = T.E[1];
          pre.rsrc[1] = s.stage[2]; pre.rcoef[1] = T.E[3];
          pre.rsrc[2] = s.stage[3]; pre.rcoef[2] = T.E[4];
          pre.rsrc[3] = s.stage[4]; pre.rcoef[3] = T.E[5];
          pre.rsrc[4] = s.stage[5]; pre.rcoef[4] = T.E[6];
          pre.ra = s.y; pre.rb = s.scratch; pre.thresh = thres;
        }
        double tstg = tnow + T.Cc[i] * htry;
        if (s.at_t_end && T.Cc[i] == 1.0) tstg = s.t_end;
        const bool check_a = s.phase2 && i <= 3 && std::fabs(htry) > hmin;
        VecDesc post = blank();
        if (check_a) {
          post.rkind = kRedStepA;
          post.ra = s.y; post.rb = s.scratch; post.rc = s.stage[i - 1]; post.rd = s.yp; post.rw = s.weights;
        }
        if (i == 8) {  // fused error estimate of the completed step (RKSuitef90.f90:1861-1874)
          post.rkind = kRedFinal;
          post.rn = 8;
          post.rsrc[0] = s.yp; post.rcoef[0] = T.A[8][1] - T.B[1];
          for (int j = 2; j <= 8; ++j) {
            post.rsrc[j - 1] = s.stage[j - 1];
            post.rcoef[j - 1] = (j <= 7 ? T.A[8][j] : 0.0) - T.B[j];
          }
          post.rscale = htry;
          post.ra = s.y; post.rb = s.scratch; post.rd = s.errest; post.rw = s.weights; post.thresh = thres;
        }
        // Stage 8 is formed from stages 1..7 whatever the early estimate says,
        // so when a right-hand side is cheap next to a host round trip the
        // estimate is read together with stage 8's: a failed estimate then
        // rejects the step exactly as before and the extra stage-8
        // evaluation is discarded uncounted.
        const bool defer = i == 7 && defer_early;
        co_await eval(s, tstg, pre, s.stage[i - 1], post, nullptr, defer);
        if (defer) continue;
        if (i == 7 || (i == 8 && defer_early)) {
          err = std::fabs(s.h) * s.op.res[0];  // stepb scales by |comm%h|
          if (err > s.tol) co_return 0;       // the stage-7 evaluation is not counted
          if (i == 8) s.f_count += 1;         // stage 7
        }
        s.f_count += 1;
        if (i == 8) err = s.op.post_res[0];
        if (check_a) {  // stepa (RKSuitef90.f90:1876-1955)
          double ynrm = s.op.post_res[0], ystgnm = s.op.post_res[1], argdif = s.op.post_res[2];
          double fdiff = s.op.post_res[3];
          double tdiff = tstg - tnow;
          double twt = std::fabs(s.t_end - tnow);
          ynrm = std::max(ynrm, std::fabs(tnow) / twt);
          ystgnm = std::max(ystgnm, std::fabs(tstg) / twt);
          argdif = std::max(argdif, std::fabs(tdiff) / twt);
          bool cutbak = false;
          if (argdif > T.round_off * std::max(ynrm, ystgnm)) {
            if (std::fabs(htry) * fdiff > T.stability_radius * argdif) {
              double scl = std::max(T.rs4, std::min(T.rs1, (T.stability_radius * argdif) / (std::fabs(htry) * fdiff)));
              htry = scl * htry;
              cutbak = true;
            }
          }
          if (cutbak) {
            s.at_t_end = false;
            if (std::fabs(htry) <= hmin) {
              htry = std::copysign(hmin, htry);
              s.phase2 = false;
            }
            restart = true;
            break;
          }
        }
      }
      if (restart) continue;
      co_return 0;
    }
  }

  // approx_jacobian of stiff_r1 (RKSuitef90.f90:2239-2272): out = (h_avg/temp1)(f(t, y + temp1 v) - yp)
  Task<int> approx_jacobian(Sim& s, const double* v, double vdotv, double scale, double* out) {
    double temp1 = scale / std::sqrt(vdotv);
    VecDesc pre = blank();
    pre.dst = s.scratch;  // vtemp shares storage with y_new
    pre.base = s.y;
    add_term(pre, v, temp1);
    co_await eval(s, s.t, pre, out, blank());
    s.f_count += 1;
    VecDesc d = blank();
    d.dst = out; d.base = out; add_term(d, s.yp, -1.0); d.post_scale = s.h_average / temp1;
    co_await vec(s, d);
    co_return 0;
  }

  Task<int> wdot(Sim& s, const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                 const double* b2, double out[3]) {
    VecDesc d = blank();
    d.rkind = kRedWdot;
    d.rw = s.weights;
    int k = 0;
    const double* pr[6] = {a0, b0, a1, b1, a2, b2};
    for (int q = 0; q < 3; ++q)
      if (pr[2 * q]) {
        d.rsrc[2 * k] = pr[2 * q];
        d.rsrc[2 * k + 1] = pr[2 * q + 1];
        ++k;
      }
    d.rn = k;
    co_await vec(s, d);
    out[0] = s.op.res[0]; out[1] = s.op.res[1]; out[2] = s.op.res[2];
    co_return 0;
  }

  // stiff_r1 (RKSuitef90.f90:1981-2367)
  Task<int> stiff(Sim& s, bool toomch, bool& sure_stiff) {
    const Tables& T = tables();
    typedef std::complex<double> cd;
    sure_stiff = false;
    bool lots_of_fails = false;
    if ((s.step_count - 10) % 40 == 0) {
      lots_of_fails = s.stiff_bad_step_count >= 10;
      s.stiff_bad_step_count = 0;
    }
    bool maybe_stiff = toomch || lots_of_fails;
    double v0v0 = 0, scale = 0;
    double* v0 = s.errest;
    double* v1 = s.stage[1];
    double* v2 = s.stage[2];
    double* v3 = s.stage[3];
    double r3[3];
    if (maybe_stiff) {
      VecDesc w = blank();
      w.rkind = kRedWeightsAvg; w.ra = s.y; w.rb = s.y_old; w.rw = s.weights; w.thresh = thres;
      co_await vec(s, w);
      double q = std::fabs(s.h / s.h_average);
      maybe_stiff = 0.2 < q && q < 5.0;
      if (maybe_stiff) {
        maybe_stiff = T.cost * std::fabs((s.t_end - s.t) / s.h_average) > 5000.0;
        if (maybe_stiff) {
          co_await wdot(s, v0, v0, s.y, s.y, nullptr, nullptr, r3);
          v0v0 = r3[0];
          double ynrm = std::sqrt(r3[1]);
          scale = ynrm * T.sqrrmc;
          if (scale == 0.0) {
            scale = v0v0 * T.sqrrmc;
            maybe_stiff = scale > 0.0;
          }
        }
      }
    }
    if (!maybe_stiff) co_return 0;
    if (v0v0 == 0.0) {
      VecDesc d = blank();
      d.dst = v0; d.bconst = 1.0;
      co_await vec(s, d);
      co_await wdot(s, v0, v0, nullptr, nullptr, nullptr, nullptr, r3);
      v0v0 = r3[0];
    }
    double v0nrm = std::sqrt(v0v0);
    {
      VecDesc d = blank();
      d.dst = v0; d.base = v0; d.bcoef = 1.0 / v0nrm;
      co_await vec(s, d);
    }
    v0v0 = 1.0;
    cd root_pair[2], prv_root_pair[2];
    double rold = 0, rho = 0;
    auto dominant = [&](double a11, double a10, double a00) {
      double ratio = a10 / a00;
      rho = std::fabs(ratio);
      double det = a00 * a11 - a10 * a10;
      double res = std::fabs(det / a00);
      bool big = det == 0.0 || (res <= std::fabs(a11) * 0.001 * 0.001 && std::fabs(ratio - rold) <= 0.001 * rho);
      if (big) { root_pair[0] = cd(ratio, 0.0); root_pair[1] = cd(0.0, 0.0); }
      rold = ratio;
      return big;
    };
    auto qroots = [&](double alpha, double beta, cd out[2]) {
      cd temp = alpha / 2.0;
      cd sqdisc = std::sqrt(temp * temp - beta);
      if (sqdisc == cd(0.0, 0.0)) { out[0] = -temp; out[1] = -temp; }
      else {
        cd r1 = -temp + sqdisc, r2 = -temp + sqdisc;  // as written in the reference
        if (std::abs(r1) > std::abs(r2)) { out[0] = r1; out[1] = r2; } else { out[0] = r2; out[1] = r1; }
      }
    };
    int ntry = 1;
    for (;;) {
      co_await approx_jacobian(s, v0, v0v0, scale, v1);
      co_await wdot(s, v1, v1, v1, v0, nullptr, nullptr, r3);
      double v1v1 = r3[0], v1v0 = r3[1];
      if (std::sqrt(v1v1) > 1.0e10 * std::sqrt(v0v0)) co_return 0;
      if (ntry == 1) {
        rold = v1v0 / v0v0;
        if (std::fabs(rold) < T.cubrmc) co_return 0;
      } else {
        if (dominant(v1v1, v1v0, v0v0)) break;
      }
      co_await approx_jacobian(s, v1, v1v1, scale, v2);
      co_await wdot(s, v2, v2, v2, v0, v2, v1, r3);
      double v2v2 = r3[0], v2v0 = r3[1], v2v1 = r3[2];
      if (dominant(v2v2, v2v1, v1v1)) break;
      double det1 = v0v0 * v1v1 - v1v0 * v1v0;
      double alpha1 = (-v0v0 * v2v1 + v1v0 * v2v0) / det1;
      double beta1 = (v1v0 * v2v1 - v1v1 * v2v0) / det1;
      co_await approx_jacobian(s, v2, v2v2, scale, v3);
      co_await wdot(s, v3, v3, v3, v1, v3, v2, r3);
      double v3v3 = r3[0], v3v1 = r3[1], v3v2 = r3[2];
      if (dominant(v3v3, v3v2, v2v2)) break;
      double det2 = v1v1 * v2v2 - v2v1 * v2v1;
      double alpha2 = (-v1v1 * v3v2 + v2v1 * v3v1) / det2;
      double beta2 = (v2v1 * v3v2 - v2v2 * v3v1) / det2;
      double res2 = std::fabs(v3v3 + alpha2 * v3v2 + beta2 * v3v1 + alpha2 * v3v2 + alpha2 * alpha2 * v2v2 +
                              alpha2 * beta2 * v2v1 + beta2 * v3v1 + beta2 * alpha2 * v2v1 + beta2 * beta2 * v1v1);
      if (res2 <= std::fabs(v3v3) * 0.001 * 0.001) {
        qroots(alpha1, beta1, prv_root_pair);
        qroots(alpha2, beta2, root_pair);
        rho = std::abs(prv_root_pair[0]);
        double dist = std::min(std::abs(root_pair[0] - prv_root_pair[0]), std::abs(root_pair[0] - prv_root_pair[1]));
        if (dist <= 0.001 * rho) break;
      }
      ntry += 1;
      if (ntry > T.max_stiff_iters) co_return 0;
      double v3nrm = std::sqrt(v3v3);
      VecDesc d = blank();
      d.dst = v0; d.base = v3; d.bcoef = 1.0 / v3nrm;
      co_await vec(s, d);
      v0v0 = 1.0;
    }
    if (root_pair[0].real() < 0.0) {
      if (!(std::abs(root_pair[1]) >= 0.9 * rho && root_pair[1].real() > 0.0)) {
        if (std::fabs(root_pair[0].imag()) <= std::fabs(root_pair[0].real()) * T.tan_angle)
          sure_stiff = rho >= 0.9 * T.stability_radius;
      }
    }
    co_return 0;
  }

  // step_integrate_r1 (RKSuitef90.f90:1185-1587); returns the flag
  Task<int> step_integrate(Sim& s) {
    const Tables& T = tables();
    int ier = 1;
    double hmin = 0;
    if (s.at_t_start) {
      VecDesc pre = copy_desc(s.arg, s.y);
      VecDesc post = blank();
      post.rkind = kRedYpNorm;  // weights = max(|y|, thresh); ypnorm
      post.ra = s.y; post.rb = s.yp; post.rw = s.weights; post.thresh = thres;
      co_await eval(s, s.t, pre, s.yp, post);
      s.f_count += 1;
      bool phase1 = s.h_start == 0.0;
      s.phase2 = phase1;
      if (phase1) {
        s.h = std::fabs(s.h);
        double ypnorm = s.op.post_res[0];
        double tau = std::pow(s.tol, T.expon);
        if (s.h * ypnorm > tau) s.h = tau / ypnorm;
        hmin = std::max(T.sqtiny, T.toosml * std::max(std::fabs(s.t_start), std::fabs(s.t_end)));
        s.h = s.dir * std::max(s.h, hmin);
      }
    } else {
      if (s.at_t_end) {
        fail(s, "integration already reached T_END");
        co_return -1;
      }
    }
    bool phase3 = true;
    bool failed = false;
    for (;;) {  // take_step
      s.h = std::copysign(std::fabs(s.h), s.dir);
      s.at_t_end = s.dir * ((s.t + s.h) - s.t_end) >= 0.0;
      if (s.at_t_end) s.h = s.t_end - s.t;
      else if (s.dir * ((s.t + 2.0 * s.h) - s.t_end) >= 0.0) s.h = (s.t_end - s.t) / 2.0;
      hmin = std::max(T.sqtiny, T.toosml * std::max(std::fabs(s.t), std::fabs(s.t + s.h)));
      if (std::fabs(s.h) < hmin) { ier = 5; break; }
      if (s.chkeff) {
        s.hit_t_end_count += 1;
        if (s.hit_t_end_count >= 100 && s.hit_t_end_count >= s.step_count / 3) {
          ier = 2; s.hit_t_end_count = 0; break;
        }
      }
      if (!failed) {
        bool toomch = s.f_count > 5000;
        if (toomch) {
          ier = 3;
          s.full_f_count += s.f_count;
          s.f_count = 0;
        }
        bool sure_stiff = false;
        co_await stiff(s, toomch, sure_stiff);
        if (sure_stiff) ier = 4;
        if (ier != 1) break;
      }
      double htry = s.h;
      double err = 0.0;
      co_await step_r1(s, s.t, htry, err, hmin);
      s.h = htry;
      if (err > s.tol) {
        double alpha;
        if (s.at_t_start) { phase3 = false; alpha = T.rs1; }
        else {
          s.bad_step_count += 1;
          s.stiff_bad_step_count += 1;
          if (failed) alpha = T.rs1;
          else { alpha = T.safety * std::pow(s.tol / err, T.expon); alpha = std::max(alpha, T.rs1); }
        }
        s.h = alpha * s.h;
        failed = true;
        continue;
      }
      double beta = std::pow(err / s.tol, T.expon);
      if (!s.at_t_start) {
        double t1 = std::pow(err, T.expon) / s.h;
        double t2 = std::pow(s.errold, T.expon) / s.h_old;
        if (t1 < t2 * 100.0 && t2 < t1 * 100.0) beta = beta * (t1 / t2);
      }
      double alpha = T.rs3;
      if (T.safety < beta * alpha) alpha = T.safety / beta;
      if (s.at_t_start) {
        s.phase2 = false;
        phase3 = phase3 && !s.at_t_end && (alpha > T.rs2);
        if (phase3) { s.h = alpha * s.h; continue; }
      }
      alpha = std::min(alpha, T.rs);
      if (failed) alpha = std::min(alpha, 1.0);
      alpha = std::max(alpha, T.rs1);
      s.h_old = s.h;
      s.h = alpha * s.h;
      if (s.at_t_start) s.h_average = s.h_old;
      else s.h_average = 0.9 * s.h_average + 0.1 * s.h_old;
      s.at_t_start = false;
      s.errold = err;
      s.t_old = s.t;
      if (s.at_t_end) s.t = s.t_end;
      else s.t = s.t + s.h_old;
      s.step_count += 1;
      // y_old <- y; y <- y_new; yp_old (shares y_new's storage) <- yp; yp <- last stage:
      // a permutation of buffers, no copies
      double* old_y = s.y;
      double* old_yold = s.y_old;
      double* old_scratch = s.scratch;
      double* old_yp = s.yp;
      double* old_s7 = s.stage[7];
      s.y_old = old_y;
      s.y = old_scratch;
      s.scratch = old_yp;
      s.yp = old_s7;
      s.stage[7] = old_yold;
      s.steps_run += 1;
      break;
    }
    s.at_t_end = s.t == s.t_end;
    s.chkeff = s.at_t_end;
    if (ier == 1) s.st_interp = -2;  // usable
    if (ier == 3) s.flag3 += 1;
    if (ier == 4) s.flag4 += 1;
    if (ier == 1 && s.budget > 0 && s.steps_run >= s.budget) co_await pause(s);
    co_return ier;
  }

  // interpolate_r1 for METHOD 'M' (RKSuitef90.f90:2655-2997); yp_old lives in scratch
  Task<int> interpolate(Sim& s, double t_want, double* y_want) {
    const Tables& T = tables();
    const double* yp_old = s.scratch;
    if (s.st_interp == -2) {
      // extra_stages (RKSuitef90.f90:2929-2995): stage 9 -> stage[7], 10 -> xstage, 11 -> stage[1]
      for (int i = 9; i <= 11; ++i) {
        VecDesc pre = blank();
        pre.dst = s.ytemp;
        pre.base = s.y_old;
        for (int j = 1; j <= i - 1; ++j) {
          const double* v = j == 1 ? yp_old
                            : j <= 7 ? s.stage[j - 1]
                            : j == 8 ? s.yp
                            : j == 9 ? s.stage[7]
                                     : s.xstage;
          add_term(pre, v, s.h_old * T.A[i][j]);
        }
        double* out = i == 9 ? s.stage[7] : (i == 10 ? s.xstage : s.stage[1]);
        co_await eval(s, s.t_old + T.Cc[i] * s.h_old, pre, out, blank());
        s.f_count += 1;
      }
      // form_intrp (RKSuitef90.f90:2826-2873): p(:,k) = h_old * sum_i r(i,k+1) stage_i
      const double* st[12];
      st[1] = yp_old;
      for (int j = 2; j <= 7; ++j) st[j] = s.stage[j - 1];
      st[8] = s.yp;
      st[9] = s.stage[7];
      st[10] = s.xstage;
      st[11] = s.stage[1];
      for (int k = 1; k <= 5; ++k) {
        VecDesc d = blank();
        d.dst = s.p[k];
        for (int i = 1; i <= 11; ++i)
          if (T.R[i][k + 1] != 0.0 && i != 2) add_term(d, st[i], T.R[i][k + 1]);
        d.post_scale = s.h_old;
        co_await vec(s, d);
      }
      s.st_interp = 1;
    }
    // evaluate_intrp (RKSuitef90.f90:2876-2899), expanded powers of sigma
    double sigma = (t_want - s.t) / s.h_old;
    VecDesc d = blank();
    d.dst = y_want;
    d.base = s.y;
    add_term(d, s.yp, s.h_old * sigma);
    double sp = sigma * sigma;
    for (int k = 1; k <= 5; ++k) {
      add_term(d, s.p[k], sp);
      sp *= sigma;
    }
    co_await vec(s, d);
    co_return 0;
  }

  // range_integrate_r1 (RKSuitef90.f90:942-1183)
  Task<int> range_integrate(Sim& s, double t_want, double& t_got, double* y_got) {
    const Tables& T = tables();
    if (s.st_range == 5 || s.st_range == 6) {
      fail(s, "RANGE_INTEGRATE called after a hard failure (flag 5/6)");
      co_return -1;
    }
    if (s.at_t_start) {
      s.range_t_end = s.t_end;
      s.tlast = s.t_start;
      t_got = s.t_start;
      if (s.h_start == 0.0) {
        s.h = std::min(std::fabs(s.h), std::fabs(t_want - s.t_start));
        double hmin = std::max(T.sqtiny, T.toosml * std::max(std::fabs(s.t_start), std::fabs(s.t_end)));
        s.h = std::max(s.h, hmin);
      }
    } else if (s.tlast == s.range_t_end) {
      fail(s, "RANGE_INTEGRATE called after reaching T_END");
      co_return -1;
    }
    if (s.dir * (t_want - s.tlast) <= 0.0 || s.dir * (t_want - s.range_t_end) > 0.0) {
      fail(s, "T_WANT outside (T_GOT, T_END]");
      co_return -1;
    }
    int ier = 1;
    for (;;) {
      bool goback = s.dir * (s.t - t_want) >= 0.0;
      if (goback) {
        co_await interpolate(s, t_want, y_got);
        t_got = t_want;
        break;
      }
      int flag = co_await step_integrate(s);
      ier = flag;
      if (flag == 1) continue;
      if (flag < 0) co_return flag;
      t_got = s.t;
      co_await vec(s, copy_desc(y_got, s.y));
      break;
    }
    s.tlast = t_got;
    s.st_range = ier;
    co_return ier;
  }

  // MagTense_ODE_RKSuite + the field-row loop of SolveLandauLifshitzEquation
  Task<int> run_state(Sim& s) {
    std::vector<double> comb(t_req);
    for (double v : t_conv) comb.push_back(v);
    std::vector<double> uniq;
    for (double v : comb)
      if (std::find(uniq.begin(), uniq.end(), v) == uniq.end()) uniq.push_back(v);
    std::vector<double> sorted = uniq;
    std::stable_sort(sorted.begin(), sorted.end());
    for (int hi = 0; hi < n_h; ++hi) {
      s.hext_row = hi;
      double* yo = yout + ((size_t)s.id * n_h + hi) * nt * n3();
      double* to = t_out_h.data() + ((size_t)s.id * n_h + hi) * nt;
      // initial state: m0 for the first row, else the last stored state of the previous row
      const double* start = hi == 0 ? s.y_step : yo - n3();
      co_await vec(s, copy_desc(s.y, start));
      co_await vec(s, copy_desc(yo, s.y));
      co_await vec(s, copy_desc(s.y_last, s.y));
      if (setup(s, t_req[0], t_req[nt - 1]) != kOk) co_return -1;
      to[0] = t_req[0];
      int k = 1;
      for (size_t i = 1; i < sorted.size(); ++i) {
        double tw = sorted[i], t_step = 0;
        int flag = co_await range_integrate(s, tw, t_step, s.y_step);
        if (flag < 0) co_return flag;
        bool is_conv = std::find(t_conv.begin(), t_conv.end(), tw) != t_conv.end();
        if (is_conv) {
          VecDesc d = blank();
          d.rkind = kRedMaxDiff; d.ra = s.y_step; d.rb = s.y_last;
          co_await vec(s, d);
          if (s.op.res[0] < conv_tol) {
            if (k < nt) {
              co_await vec(s, copy_desc(yo + (size_t)k * n3(), s.y_step));
              to[k] = t_step;
            }
            break;
          }
          co_await vec(s, copy_desc(s.y_last, s.y_step));
        }
        bool is_out = std::find(t_req.begin(), t_req.end(), tw) != t_req.end();
        if (is_out && k < nt) {
          co_await vec(s, copy_desc(yo + (size_t)k * n3(), s.y_step));
          to[k] = t_step;
          k += 1;
        }
      }
      s.tot_f += s.full_f_count + s.f_count;
      s.tot_steps += s.step_count;
      s.tot_bad += s.bad_step_count;
      if (return_fields) {
        // StoreHeffComponents: fields at every stored magnetization, at the requested times
        double* fo = fout + ((size_t)s.id * n_h + hi) * nt * (size_t)12 * n;
        for (int j = 0; j < nt; ++j) {
          VecDesc pre = copy_desc(s.arg, yo + (size_t)j * n3());
          co_await eval(s, t_req[j], pre, nullptr, blank(), fo + (size_t)j * 12 * n);
        }
      }
      // the next field row starts from y_step <- last stored state
      co_await vec(s, copy_desc(s.y_step, yo + (size_t)(nt - 1) * n3()));
    }
    co_return 0;
  }

  // ---- scheduler ----------------------------------------------------------------
  int run_until_idle();
  int execute_tick(std::vector<Sim*>& evals, std::vector<Sim*>& vecs);
};

int Solver::execute_tick(std::vector<Sim*>& evals, std::vector<Sim*>& vecs) {
  const size_t N3 = n3();
  bool need_sync = false;
  ticks += 1;
  const int nres = B * 4;
  double* res_vec = results;
  double* res_pre = results + nres;
  double* res_post = results + 2 * nres;
  PackSpec nopack{kPackNone, nullptr, 0, nullptr};
  // vector-only operations of this tick
  bool vec_reduce = false;
  if (!vecs.empty()) {
    std::vector<VecDesc> d(vecs.size());
    for (size_t k = 0; k < vecs.size(); ++k) {
      d[k] = vecs[k]->op.vec;
      d[k].rslot = vecs[k]->id;
      if (d[k].rkind != kRedNone && d[k].rkind != kRedWeightsAvg) vec_reduce = true;
    }
    if (timing) MT_HIP(hipEventRecord(ev[0], stream));
    int rc;
    if ((int)d.size() <= kInlineStates) {
      rc = launch_vec_inline(d.data(), (int)d.size(), n, partials, res_vec, nopack, vec_reduce, stream);
    } else {
      const VecDesc* dd = nullptr;
      rc = stage_descs(d, &dd);
      if (rc) return rc;
      rc = launch_vec(dd, (int)d.size(), n, partials, res_vec, nopack, vec_reduce, stream);
    }
    if (rc) return rc;
    need_sync |= vec_reduce;
    if (timing) {
      MT_HIP(hipEventRecord(ev[1], stream));
      MT_HIP(hipEventSynchronize(ev[1]));
      float ms = 0;
      MT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
      vec_ms += ms;
      vec_launches += 1;
      for (const VecDesc& v : d) vec_bytes += vec_desc_bytes(v, kPackNone);
    }
  }
  // right-hand sides: groups of up to 64 columns per demag pass
  bool pre_sync = false, post_reduce_any = false;
  for (size_t g0 = 0; g0 < evals.size(); g0 += kMaxCols) {
    size_t g1 = std::min(evals.size(), g0 + (size_t)kMaxCols);
    int S = (int)(g1 - g0);
    std::vector<VecDesc> pre(S), post(S);
    std::vector<LlgDesc> ld(S);
    bool pre_red = false, post_red = false;
    for (int k = 0; k < S; ++k) {
      Sim& s = *evals[g0 + k];
      pre[k] = s.op.pre;
      pre[k].pack_col = k;
      pre[k].rslot = s.id;  // per-state result slots: a deferred result survives later ticks
      if (pre[k].rkind != kRedNone) {
        pre_red = true;
        if (!s.op.defer_pre) pre_sync = true;
      }
      post[k] = s.op.post;
      post[k].rslot = s.id;
      if (post[k].rkind != kRedNone) post_red = true;
      LlgDesc& L = ld[k];
      std::memset(&L, 0, sizeof(L));
      L.m = pre[k].dst;
      L.out = s.op.out;
      L.fields = s.op.fields;
      L.hpart = hbuf + (size_t)k * N3;
      L.split_stride = (int64_t)hcols * N3;
      L.nsplit = precision == MT_PREC_BF16X2 ? ksplit + (near_r > 0 ? 1 : 0) : (sparse.blocks || Kf32 ? ksplit : 1);
      hext_at(s, s.op.t, L.hext);
      L.alpha = alpha_at(s.op.t);
      L.gamma = gamma;
      s.evals += 1;
    }
    PackSpec pk;
    if (precision == MT_PREC_BF16X2) pk = PackSpec{kPackBf16, nullptr, 0, packb};
    else if (sparse.blocks || Kf32) pk = PackSpec{kPackF32Frag, nullptr, 0, packb};
    else pk = PackSpec{kPackF32, packf, (int64_t)N3, nullptr};
    if (timing) MT_HIP(hipEventRecord(ev[2], stream));
    // Small groups pass their descriptors by value; otherwise the group's
    // three descriptor arrays go over in one copy (the near-band correction
    // reads the LL descriptors from device memory)
    const bool inl = S <= kInlineStates && !(precision == MT_PREC_BF16X2 && near_r > 0);
    // the FP32 dense path runs its demag rows and the LL right-hand side in one kernel
    const bool fused = precision == MT_PREC_FP32 && demag_approx == MT_DEMAG_DENSE;
    // and on small problems forms the stage argument in that kernel as well
    const bool stage_fused = fused && inl && !pre_red && stage_fusion && n <= kStageFusedMaxCells;
    // the split product's near-band kernel evaluates the LL right-hand side
    const bool llg_in_near = precision == MT_PREC_BF16X2 && near_r > 0 && near_llg;
    const VecDesc* dpre = nullptr;
    const LlgDesc* dl = nullptr;
    const VecDesc* dpost = nullptr;
    int rc;
    if (!inl) {
      rc = stage_reserve(3 * 256 + (size_t)S * (2 * sizeof(VecDesc) + sizeof(LlgDesc)));
      if (rc) return rc;
      dpre = stage_put(pre);
      dl = stage_put(ld);
      dpost = post_red ? stage_put(post) : nullptr;
      rc = stage_flush();
      if (rc) return rc;
    }
    double* pre_parts = partials + (size_t)B * vec_blocks_per_state((int)N3) * 4;
    if (!stage_fused) {
      rc = inl ? launch_vec_inline(pre.data(), S, n, pre_parts, res_pre, pk, pre_red, stream)
               : launch_vec(dpre, S, n, pre_parts, res_pre, pk, pre_red, stream);
      if (rc) return rc;
    }
    if (timing) MT_HIP(hipEventRecord(ev[3], stream));
    if (precision == MT_PREC_BF16X2) {
      if (Ktab)
        rc = launch_gemm_bf16x2_toeplitz(Ktab, nx, ny, nz, packb, hbuf, (int64_t)N3, (int64_t)hcols * N3, S, ksplit,
                                         stream);
      else if (sparse.blocks)
        rc = launch_gemm_bsr(sparse.blocks->frags, sparse.blocks->view(), false, packb, n, hbuf, (int64_t)N3,
                                    (int64_t)hcols * N3, S, ksplit, stream);
      else
        rc = launch_gemm_bf16x2(Ksplit, packb, n, hbuf, (int64_t)N3, (int64_t)hcols * N3, S, ksplit, stream);
      // the near band's FP32 correction: straight into each cell's LL right-hand
      // side (llg_in_near), or into one more partial-sum slot for the LL kernel
      if (rc == kOk && near_r > 0 && timing) MT_HIP(hipEventRecord(ev[7], stream));
      if (rc == kOk && near_r > 0)
        rc = launch_near_correction(near_tab, near_r, nx, ny, nz, dl, nullptr, 0, S,
                                    hbuf + (size_t)ksplit * hcols * N3, (int64_t)N3, 1.0f, false, stream,
                                    llg_in_near ? &grid : nullptr);
    } else if (Kf32) {
      rc = launch_gemm_f32(Kf32, packb, n, hbuf, (int64_t)N3, (int64_t)hcols * N3, S, ksplit, stream);
    } else if (sparse.blocks) {
      rc = launch_gemm_bsr(sparse.blocks->frags, sparse.blocks->view(), true, packb, n, hbuf, (int64_t)N3,
                           (int64_t)hcols * N3, S, ksplit, stream);
    } else if (demag_approx != MT_DEMAG_DENSE) {
      rc = sparse.apply(packf, (int64_t)N3, S, 1.0f, hbuf, stream);
    } else if (stage_fused) {
      rc = launch_stage_demag_llg_f32(K32, grid, pre.data(), ld.data(), S, stream);
    } else {
      rc = launch_demag_llg_f32(K32, grid, dl, ld.data(), packf, (int64_t)N3, S, stream);
    }
    if (rc) return rc;
    if (timing) MT_HIP(hipEventRecord(ev[4], stream));
    demag_passes += 1;
    if (!fused && !llg_in_near) {
      rc = inl ? launch_llg_inline(grid, ld.data(), S, stream) : launch_llg(grid, dl, S, stream);
      if (rc) return rc;
    }
    if (timing) MT_HIP(hipEventRecord(ev[5], stream));
    if (post_red) {
      double* post_parts = partials + (size_t)2 * B * vec_blocks_per_state((int)N3) * 4;
      rc = inl ? launch_vec_inline(post.data(), S, n, post_parts, res_post, nopack, true, stream)
               : launch_vec(dpost, S, n, post_parts, res_post, nopack, true, stream);
      if (rc) return rc;
    }
    if (timing) {
      MT_HIP(hipEventRecord(ev[6], stream));
      MT_HIP(hipEventSynchronize(ev[6]));
      float ms[4] = {0, 0, 0, 0};
      MT_HIP(hipEventElapsedTime(&ms[0], ev[2], ev[3]));
      MT_HIP(hipEventElapsedTime(&ms[1], ev[3], ev[4]));
      MT_HIP(hipEventElapsedTime(&ms[2], ev[4], ev[5]));
      MT_HIP(hipEventElapsedTime(&ms[3], ev[5], ev[6]));
      gemm_ms += ms[1];
      if (near_r > 0 && precision == MT_PREC_BF16X2) {
        float mn = 0;
        MT_HIP(hipEventElapsedTime(&mn, ev[7], ev[4]));
        near_ms += mn;
      }
      // fused / llg_in_near: the LL right-hand side ran inside the demag or the
      // near-band kernel (its time is in gemm_ms, and in near_ms for the latter)
      const bool llg_own = !fused && !llg_in_near;
      llg_ms += llg_own ? ms[2] : 0.0f;
      llg_launches += llg_own ? 1 : 0;
      llg_bytes += llg_own ? llg_launch_bytes(ld) : 0.0;
      // stage_fused: the stage combination ran inside the demag kernel (its time is in gemm_ms)
      vec_ms += (stage_fused ? 0.0f : ms[0]) + (post_red ? ms[3] : 0.0f);
      vec_launches += (stage_fused ? 0 : 1) + (post_red ? 1 : 0);
      if (!stage_fused)
        for (const VecDesc& v : pre) vec_bytes += vec_desc_bytes(v, pk.mode);
      if (post_red)
        for (const VecDesc& v : post) vec_bytes += vec_desc_bytes(v, kPackNone);
    }
    post_reduce_any |= post_red;
  }
  if (!evals.empty()) eval_ticks += 1;
  need_sync |= pre_sync || post_reduce_any;
  if (need_sync) {
    MT_HIP(hipMemcpyAsync(h_results, results, sizeof(double) * 3 * nres, hipMemcpyDeviceToHost, stream));
    MT_HIP(hipStreamSynchronize(stream));
    arena_off = 0;
    for (Sim* v : vecs)
      for (int q = 0; q < 4; ++q) v->op.res[q] = h_results[v->id * 4 + q];
    for (Sim* e : evals)
      for (int q = 0; q < 4; ++q) {
        e->op.res[q] = h_results[nres + e->id * 4 + q];
        e->op.post_res[q] = h_results[2 * nres + e->id * 4 + q];
      }
  }
  for (Sim* s : vecs) s->op.kind = Op::kNone;
  for (Sim* s : evals) s->op.kind = Op::kNone;
  return kOk;
}

int Solver::run_until_idle() {
  std::vector<Sim*> evals, vecs;
  for (;;) {
    if (total_target > 0) {
      long long acc = 0;
      for (const Sim& s : sims) acc += s.steps_run;
      if (acc >= total_target) break;
    }
    evals.clear();
    vecs.clear();
    for (Sim& s : sims) {
      if (s.done) continue;
      if (s.op.kind == Op::kPause) {
        if (s.budget > 0 && s.steps_run >= s.budget) continue;
        s.op.kind = Op::kNone;  // budget raised: continue
      }
      if (s.op.kind == Op::kNone) {
        if (!s.started) {
          s.started = true;
          s.top.reset(new Task<int>(run_state(s)));
          s.top->h.resume();
        } else {
          s.resume_point.resume();
        }
        if (s.top->h.done()) {
          s.done = true;
          if (s.error) {
            set_error(s.error_msg);
            return s.error;
          }
          int rv = s.top->h.promise().value;
          if (rv < 0) {
            set_error(s.error_msg.empty() ? "solver failed" : s.error_msg);
            return kErrSolver;
          }
          continue;
        }
      }
      if (s.op.kind == Op::kEval) evals.push_back(&s);
      else if (s.op.kind == Op::kVec) vecs.push_back(&s);
    }
    if (evals.empty() && vecs.empty()) break;
    int rc = execute_tick(evals, vecs);
    if (rc) return rc;
  }
  MT_HIP(hipStreamSynchronize(stream));
  arena_off = 0;
  return kOk;
}

}  // namespace mt

// ============================================================================
// C ABI
// ============================================================================
using mt::Solver;

extern "C" {

int mt_solver_create(const mt_micromag_problem* q, mt_solver** out) {
  using namespace mt;
  if (!q || !out) {
    set_error("null argument");
    return kErrArg;
  }
  *out = nullptr;
  const bool tetra = q->grid_type == 2;
  const bool unstructured = q->grid_type == 3 || tetra;  // loaded points and exchange operator
  MT_CHECK(q->grid_type == 0 || q->grid_type == 1 || unstructured,
           "grid_type must be 1 (uniform), 2 (tetrahedra) or 3 (unstructured prisms)");
  if (tetra) {
    MT_CHECK(q->nx >= 1 && q->ny >= 1 && q->nz >= 1 && q->grid_pts && q->grid_nodes && q->grid_elements &&
                 q->grid_nnod >= 4,
             "tetrahedra need grid_pts, grid_nodes (grid_nnod >= 4) and grid_elements");
  } else if (unstructured) {
    MT_CHECK(q->nx >= 1 && q->ny >= 1 && q->nz >= 1 && q->grid_pts && q->grid_abc,
             "unstructured prisms need grid_pts and grid_abc");
    MT_CHECK(q->n_ave[0] >= 1 && q->n_ave[1] >= 1 && q->n_ave[2] >= 1, "n_ave entries must be >= 1");
  }
  if (unstructured) {
    MT_CHECK(q->exch_nnz >= 0 && q->exch_rows_start && q->exch_rows_end && (q->exch_nnz == 0 || (q->exch_cols && q->exch_vals)),
             "non-uniform grids need the exchange operator (exch_rows_start / rows_end / cols / vals)");
  } else {
    MT_CHECK(q->nx >= 2 && q->ny >= 2 && q->nz >= 1,
             "the uniform-grid exchange operator needs nx >= 2 and ny >= 2 (LandauLifshitzEquationSolver.f90:1215)");
  }
  MT_CHECK(q->n_states >= 1, "n_states must be >= 1");
  MT_CHECK(q->nt >= 2 && q->t, "need at least two output times");
  MT_CHECK(q->nt_hext >= 1 && q->hext, "need an applied-field table");
  MT_CHECK(q->solver == 1 || q->solver == 2, "solver must be 1 (explicit) or 2 (dynamic)");
  MT_CHECK(q->precision == MT_PREC_FP32 || q->precision == MT_PREC_BF16X2 || q->precision == MT_PREC_FP32_MFMA,
           "unknown precision");
  const int approx = q->demag_approx == 0 ? MT_DEMAG_DENSE : q->demag_approx;
  MT_CHECK(approx == MT_DEMAG_DENSE || approx == MT_DEMAG_THRESHOLD || approx == MT_DEMAG_THRESHOLD_FRACTION,
           "demag_approx must be 1 (dense), 2 (threshold) or 4 (threshold fraction)");
  MT_CHECK(q->Ms && q->m0, "Ms and m0 are required");
  {  // the reference integrates NaNs silently; refuse them before any device work
    const size_t n3c = (size_t)3 * q->nx * q->ny * q->nz;
    const size_t nm = (q->m0_per_state ? (size_t)q->n_states : 1) * n3c;
    for (size_t k = 0; k < nm; ++k) MT_CHECK(std::isfinite(q->m0[k]), "m0 holds a NaN or infinite value");
    for (size_t k = 0; k < n3c / 3; ++k)
      MT_CHECK(std::isfinite(q->Ms[k]) && q->Ms[k] != 0.0, "Ms must be finite and non-zero in every cell");
  }
  std::unique_ptr<Solver> S(new Solver());
  Solver& P = *S;
  P.nx = q->nx; P.ny = q->ny; P.nz = q->nz;
  P.n = q->nx * q->ny * q->nz;
  P.B = q->n_states;
  P.solver = q->solver;
  P.precision = q->precision;
  P.demag_approx = approx;
  P.gamma = q->gamma; P.alpha0 = q->alpha0;
  P.tol = q->tol; P.thres = q->thres; P.conv_tol = q->conv_tol;
  for (int k = 0; k < q->nt_alpha; ++k) {
    P.alpha_t.push_back(q->alphat[2 * k]);
    P.alpha_v.push_back(q->alphat[2 * k + 1]);
  }
  P.t_req.assign(q->t, q->t + q->nt);
  P.nt = q->nt;
  if (q->nt_conv > 0 && q->t_conv) P.t_conv.assign(q->t_conv, q->t_conv + q->nt_conv);
  P.n_h = q->solver == 1 ? q->nt_hext : 1;
  P.return_fields = q->return_fields != 0;
  const int n = P.n;
  const size_t n3 = P.n3();
  // grid (LandauLifshitzEquationSolver.f90:815-872)
  std::vector<double> xs(P.nx), ys(P.ny), zs(P.nz);
  auto axis = [](int m, double L, double& d, std::vector<double>& c) {
    if (m > 1) {
      d = L / m;
      for (int i = 0; i < m; ++i) c[i] = -L / 2 + i * d + d / 2;
    } else {
      d = L;
      c[0] = 0.0;
    }
  };
  axis(P.nx, q->Lx, P.dx, xs);
  axis(P.ny, q->Ly, P.dy, ys);
  axis(P.nz, q->Lz, P.dz, zs);
  std::vector<double> pts((size_t)n * 3);
  if (unstructured) {
    pts.assign(q->grid_pts, q->grid_pts + (size_t)n * 3);
  } else {
    for (int k = 0; k < P.nz; ++k)
      for (int j = 0; j < P.ny; ++j)
        for (int i = 0; i < P.nx; ++i) {
          size_t ind = (size_t)i + (size_t)j * P.nx + (size_t)k * P.nx * P.ny;
          pts[ind * 3] = xs[i]; pts[ind * 3 + 1] = ys[j]; pts[ind * 3 + 2] = zs[k];
        }
  }
  // loaded exchange operator: 1-based rows_start / rows_end / cols -> 0-based, checked
  std::vector<int> hxs, hxe, hxc;
  if (unstructured) {
    hxs.resize(n); hxe.resize(n);
    hxc.resize(q->exch_nnz);
    for (int r = 0; r < n; ++r) {
      hxs[r] = q->exch_rows_start[r] - 1;
      hxe[r] = q->exch_rows_end[r] - 1;
      MT_CHECK(hxs[r] >= 0 && hxs[r] <= hxe[r] && hxe[r] <= q->exch_nnz,
               "exchange operator: rows_start / rows_end out of range (1-based, within exch_nnz)");
    }
    P.exch_nnz_host = q->exch_nnz;
    for (int k = 0; k < q->exch_nnz; ++k) {
      hxc[k] = q->exch_cols[k] - 1;
      MT_CHECK(hxc[k] >= 0 && hxc[k] < n, "exchange operator: column index out of range (1-based)");
    }
  }
  MT_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
  for (hipEvent_t& e : P.ev) MT_HIP(hipEventCreate(&e));
  // tensor
  const size_t nn = (size_t)n * n;
  // the threshold modes build their sparse form from the FP32 tensor (CSR in
  // FP32, block-sparse split-BF16 tiles in bf16x2)
  const bool need32 = P.precision == MT_PREC_FP32 || P.precision == MT_PREC_FP32_MFMA || q->keep_fp32_tensor ||
                      approx != MT_DEMAG_DENSE;
  const bool can_toep = !unstructured && P.precision == MT_PREC_BF16X2 && !q->K_in && P.nx >= 32 && P.nx % 32 == 0 &&
                        approx == MT_DEMAG_DENSE;
  MT_CHECK(q->toeplitz != 1 || can_toep,
           "the translation-invariant tensor needs a uniform grid, precision bf16x2, nx % 32 == 0 and no K_in");
  const bool use_toep = can_toep && q->toeplitz >= 0;
  const bool needsplit = P.precision == MT_PREC_BF16X2 && !use_toep && approx == MT_DEMAG_DENSE;
  if (use_toep) {
    MT_HIP(hipMalloc(&P.Ktab, (size_t)toeplitz_tensor_bytes(P.nx, P.ny, P.nz)));
    int rc = launch_toeplitz_assemble(P.nx, P.ny, P.nz, q->Lx, q->Ly, q->Lz, P.Ktab, P.stream);
    if (rc) return rc;
  }
  if (need32) MT_HIP(hipMalloc(&P.K32, 6 * nn * sizeof(float)));
  if (needsplit) MT_HIP(hipMalloc(&P.Ksplit, (size_t)split_tensor_bytes(n)));
  if (q->K_in) {
    DevArr<float> tmp;
    float* Kd = P.K32;
    if (!Kd) {
      MT_HIP(tmp.alloc(6 * nn));
      Kd = tmp.p;
    }
    MT_HIP(hipMemcpyAsync(Kd, q->K_in, 6 * nn * sizeof(float), hipMemcpyHostToDevice, P.stream));
    if (needsplit) {
      int rc = launch_split(Kd, n, P.Ksplit, P.stream);
      if (rc) return rc;
    }
    MT_HIP(hipStreamSynchronize(P.stream));
  } else if (unstructured) {
    // averaged prism tensor / tetrahedron tensor in FP32, split afterwards for the MFMA product
    std::vector<int> elem0;
    const size_t geo = tetra ? (size_t)q->grid_nnod * 3 : pts.size();
    if (tetra) {  // 1-based node indices -> 0-based, checked; degenerate elements refused
      elem0.resize((size_t)n * 4);
      for (size_t k = 0; k < elem0.size(); ++k) {
        elem0[k] = q->grid_elements[k] - 1;
        MT_CHECK(elem0[k] >= 0 && elem0[k] < q->grid_nnod, "grid_elements: node index out of range (1-based)");
      }
      for (int e = 0; e < n; ++e) {
        const double* v0 = q->grid_nodes + (size_t)elem0[4 * e] * 3;
        double d[3][3];
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 3; ++c) d[k][c] = q->grid_nodes[(size_t)elem0[4 * e + 1 + k] * 3 + c] - v0[c];
        const double vol = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) -
                           d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
                           d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        MT_CHECK(vol != 0.0, "grid_elements: degenerate tetrahedron (zero volume)");
      }
    }
    DevArr<float> tmp;
    DevArr<double> dpts, dabc;
    DevArr<int> delem;
    float* Kd = P.K32;
    if (!Kd) {
      MT_HIP(tmp.alloc(6 * nn));
      Kd = tmp.p;
    }
    MT_HIP(dpts.alloc(pts.size()));
    MT_HIP(dabc.alloc(geo));
    MT_HIP(hipMemcpyAsync(dpts.p, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
    MT_HIP(hipMemcpyAsync(dabc.p, tetra ? q->grid_nodes : q->grid_abc, geo * sizeof(double), hipMemcpyHostToDevice,
                          P.stream));
    int rc = kOk;
    if (tetra) {
      MT_HIP(delem.alloc(elem0.size()));
      MT_HIP(hipMemcpyAsync(delem.p, elem0.data(), elem0.size() * sizeof(int), hipMemcpyHostToDevice, P.stream));
      rc = launch_assemble_tetra(dpts.p, dabc.p, delem.p, n, Kd, nullptr, P.stream);
    } else {
      rc = launch_assemble_avg(dpts.p, dabc.p, n, q->n_ave, Kd, nullptr, P.stream);
    }
    if (rc == kOk && P.Ksplit) rc = launch_split(Kd, n, P.Ksplit, P.stream);
    MT_HIP(hipStreamSynchronize(P.stream));
    if (rc) return rc;
  } else if (P.K32 || P.Ksplit) {
    DevArr<double> dpts, dabc;
    MT_HIP(dpts.alloc(pts.size()));
    MT_HIP(dabc.alloc(3));
    double abc[3] = {P.dx, P.dy, P.dz};
    MT_HIP(hipMemcpyAsync(dpts.p, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
    MT_HIP(hipMemcpyAsync(dabc.p, abc, sizeof(abc), hipMemcpyHostToDevice, P.stream));
    int rc = launch_assemble(dpts.p, dabc.p, 0, nullptr, dpts.p, n, n, P.K32, nullptr, P.Ksplit, P.stream);
    if (rc) return rc;
    MT_HIP(hipStreamSynchronize(P.stream));
  }
  if (q->K_out) {
    MT_CHECK(P.K32 != nullptr, "K_out needs the FP32 tensor (precision fp32 or keep_fp32_tensor)");
    MT_HIP(hipMemcpyAsync(q->K_out, P.K32, 6 * nn * sizeof(float), hipMemcpyDeviceToHost, P.stream));
  }
  if (P.precision == MT_PREC_FP32_MFMA && approx == MT_DEMAG_DENSE) {
    // the FP32 planes as FP32 MFMA fragments; the planes go unless kept
    MT_HIP(hipMalloc(&P.Kf32, (size_t)split_tensor_bytes(n)));
    int rc = launch_f32_fragments(P.K32, n, P.Kf32, P.stream);
    if (rc) return rc;
    MT_HIP(hipStreamSynchronize(P.stream));
    if (!q->keep_fp32_tensor) {
      MT_HIP(hipFree(P.K32));
      P.K32 = nullptr;
    }
  }
  if (approx != MT_DEMAG_DENSE) {
    // ApplyThresholdDense (DemagAuxFunctions.f90:20-122) after the tensor (LandauLifshitzEquationSolver.f90:1157)
    float thr = (float)q->demag_threshold;
    if (approx == MT_DEMAG_THRESHOLD_FRACTION) {
      int rc = threshold_for_fraction(P.K32, n, q->demag_threshold, &thr, P.stream);
      if (rc) return rc;
    }
    int rc = kOk;
    // FP32 with five or more states: the kept entries as FP32 tiles on FP32
    // MFMAs (one pass over the tiles for all states; on the 32768-cell film at
    // 8 % density 1.8 ms for up to 32 states and 3.0 ms for 64, against 1.76
    // ms for four states on the CSR and 2.7-9.8 ms from five);
    // MT_F32_TILE_STATES sets the smallest batch that takes them (0: never)
    const char* env_ft = std::getenv("MT_F32_TILE_STATES");
    const int f32_tile_states = env_ft ? std::atoi(env_ft) : 5;
    const bool f32_tiles = P.precision == MT_PREC_FP32_MFMA ||
                           (P.precision == MT_PREC_FP32 && f32_tile_states > 0 && P.B >= f32_tile_states);
    if (P.precision == MT_PREC_BF16X2 || f32_tiles) {
      P.sparse.blocks = new BlockSparse();
      rc = P.sparse.blocks->build(P.K32, n, thr, unstructured ? 0 : P.nx, P.stream, f32_tiles);
      P.sparse.n = n;
      P.sparse.threshold = thr;
      for (int c = 0; c < 6; ++c) P.sparse.nnz[c] = P.sparse.blocks->nnz[c];
    } else {
      rc = P.sparse.build(P.K32, n, thr, P.stream);
    }
    if (rc) return rc;
    if (!q->keep_fp32_tensor) {
      MT_HIP(hipStreamSynchronize(P.stream));
      MT_HIP(hipFree(P.K32));
      P.K32 = nullptr;
    }
  }
  // per-cell factors (LandauLifshitzEquationSolver.f90:416-425)
  std::vector<double> jf(n), kf(n), an;
  std::vector<float> ms(n);
  bool any_k = false;
  for (int i = 0; i < n; ++i) {
    double Ms = q->Ms[i];
    double K0 = q->K0 ? q->K0[i] : 0.0;
    jf[i] = q->A0 / (kMu0 * Ms);
    kf[i] = K0 / (kMu0 * Ms);
    ms[i] = (float)Ms;
    if (K0 != 0.0) any_k = true;
  }
  MT_HIP(hipMalloc(&P.jfact, n * sizeof(double)));
  MT_HIP(hipMalloc(&P.kfact, n * sizeof(double)));
  MT_HIP(hipMalloc(&P.ms32, n * sizeof(float)));
  MT_HIP(hipMemcpyAsync(P.jfact, jf.data(), n * sizeof(double), hipMemcpyHostToDevice, P.stream));
  MT_HIP(hipMemcpyAsync(P.kfact, kf.data(), n * sizeof(double), hipMemcpyHostToDevice, P.stream));
  MT_HIP(hipMemcpyAsync(P.ms32, ms.data(), n * sizeof(float), hipMemcpyHostToDevice, P.stream));
  if (any_k && q->u_ea) {
    an.resize((size_t)n * 6);
    for (int i = 0; i < n; ++i) {
      double ux = q->u_ea[i * 3], uy = q->u_ea[i * 3 + 1], uz = q->u_ea[i * 3 + 2];
      double* A = &an[(size_t)i * 6];
      A[0] = ux * ux; A[1] = ux * uy; A[2] = ux * uz; A[3] = uy * uy; A[4] = uy * uz; A[5] = uz * uz;
    }
    MT_HIP(hipMalloc(&P.aniso, an.size() * sizeof(double)));
    MT_HIP(hipMemcpyAsync(P.aniso, an.data(), an.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
  }
  P.grid.n = n; P.grid.nx = P.nx; P.grid.ny = P.ny; P.grid.nz = P.nz;
  P.grid.idx2 = 1.0 / (P.dx * P.dx); P.grid.idy2 = 1.0 / (P.dy * P.dy); P.grid.idz2 = 1.0 / (P.dz * P.dz);
  P.grid.jfact = P.jfact; P.grid.kfact = P.kfact; P.grid.ms32 = P.ms32; P.grid.aniso = P.aniso;
  if (unstructured) {
    MT_HIP(hipMalloc(&P.xs, n * sizeof(int)));
    MT_HIP(hipMalloc(&P.xe, n * sizeof(int)));
    MT_HIP(hipMalloc(&P.xc, (q->exch_nnz > 0 ? q->exch_nnz : 1) * sizeof(int)));
    MT_HIP(hipMalloc(&P.xv, (q->exch_nnz > 0 ? q->exch_nnz : 1) * sizeof(double)));
    MT_HIP(hipMemcpyAsync(P.xs, hxs.data(), n * sizeof(int), hipMemcpyHostToDevice, P.stream));
    MT_HIP(hipMemcpyAsync(P.xe, hxe.data(), n * sizeof(int), hipMemcpyHostToDevice, P.stream));
    if (q->exch_nnz > 0) {
      MT_HIP(hipMemcpyAsync(P.xc, hxc.data(), q->exch_nnz * sizeof(int), hipMemcpyHostToDevice, P.stream));
      MT_HIP(hipMemcpyAsync(P.xv, q->exch_vals, q->exch_nnz * sizeof(double), hipMemcpyHostToDevice, P.stream));
    }
    P.grid.xs = P.xs; P.grid.xe = P.xe; P.grid.xc = P.xc; P.grid.xv = P.xv;
  }
  // state vectors: 23 per state
  constexpr int kSlots = 23;
  const size_t slot = (size_t)round_up((int64_t)n3, 64);
  MT_HIP(hipMalloc(&P.pool, (size_t)P.B * kSlots * slot * sizeof(double)));
  MT_HIP(hipMemsetAsync(P.pool, 0, (size_t)P.B * kSlots * slot * sizeof(double), P.stream));
  const size_t ysz = (size_t)P.B * P.n_h * P.nt * n3;
  MT_HIP(hipMalloc(&P.yout, ysz * sizeof(double)));
  MT_HIP(hipMemsetAsync(P.yout, 0, ysz * sizeof(double), P.stream));
  if (P.return_fields) {
    MT_HIP(hipMalloc(&P.fout, ysz * 4 * sizeof(double)));
    MT_HIP(hipMemsetAsync(P.fout, 0, ysz * 4 * sizeof(double), P.stream));
  }
  P.hcols = std::min(P.B, kMaxCols);
  // one host round trip per step instead of two where a pass over the tensor
  // is cheap next to it (under 2 GB of FP32 tensor, N < ~9100 cells): the
  // early error estimate is then read with stage 8's; MT_DEFER_EARLY=0/1 overrides
  {
    const char* env = std::getenv("MT_DEFER_EARLY");
    P.defer_early = env ? std::atoi(env) != 0 : (double)n * n * 24.0 < 2e9;
    // MT_STAGE_FUSED=0: the stage argument in its own launch on small problems too
    const char* env_sf = std::getenv("MT_STAGE_FUSED");
    P.stage_fusion = env_sf ? std::atoi(env_sf) != 0 : true;
    // MT_NEAR_LLG=0: the LL kernel after the near-band correction, as its own launch
    const char* env_nl = std::getenv("MT_NEAR_LLG");
    P.near_llg = env_nl ? std::atoi(env_nl) != 0 : true;
  }
  if (P.precision == MT_PREC_BF16X2) {
    // a block-sparse tensor's schedule fixes its partial-sum slots
    P.ksplit = P.sparse.blocks ? P.sparse.blocks->ksplit
                               : (q->ksplit > 0 ? q->ksplit
                                                : (P.Ktab ? gemm_ksplit_table(P.nx, P.ny, P.nz) : gemm_ksplit_dense(n, n)));
    MT_HIP(hipMalloc(&P.packb, (size_t)packed_m_bytes(n)));
    MT_HIP(hipMemsetAsync(P.packb, 0, (size_t)packed_m_bytes(n), P.stream));
    // near-band correction of the split product (uniform grids: dense tensor,
    // table or block-sparse tiles): MT_NEAR_RADIUS cells (default 2, 0 switches it off)
    if (!unstructured) {
      const char* env = std::getenv("MT_NEAR_RADIUS");
      P.near_r = env ? std::atoi(env) : 2;
      if (P.near_r < 0 || P.near_r > 4) P.near_r = 2;
    }
    if (P.near_r > 0) {
      const int R = P.near_r, Rz = near_band_rz(R, P.nz);
      // threshold modes: the entries the threshold drops get no correction either
      const float thr = approx == MT_DEMAG_DENSE ? -1.0f : P.sparse.threshold;
      MT_HIP(hipMalloc(&P.near_tab, (size_t)near_table_words(R, P.nz) * sizeof(uint32_t)));
      if (q->K_in) {
        // the caller's tensor: each offset's values from the first pair of cells that has it
        std::vector<uint32_t> tab((size_t)near_table_words(R, P.nz), 0u);
        const int w = 2 * R + 1;
        for (int o = 0; o < w * w * (2 * Rz + 1); ++o) {
          const int ox = o % w - R, oy = (o / w) % w - R, oz = o / (w * w) - Rz;
          if (std::abs(ox) >= P.nx || std::abs(oy) >= P.ny || std::abs(oz) >= P.nz) continue;
          const size_t i = (size_t)std::max(ox, 0) + (size_t)std::max(oy, 0) * P.nx +
                           (size_t)std::max(oz, 0) * P.nx * P.ny;
          const size_t j = (size_t)std::max(-ox, 0) + (size_t)std::max(-oy, 0) * P.nx +
                           (size_t)std::max(-oz, 0) * P.nx * P.ny;
          float k32[6];
          for (int c = 0; c < 6; ++c) {
            k32[c] = q->K_in[(size_t)c * nn + i * n + j];
            if (!(std::fabs(k32[c]) > thr)) k32[c] = 0.0f;
          }
          near_pack_offset(k32, &tab[(size_t)o * kNearWords]);
        }
        MT_HIP(hipMemcpyAsync(P.near_tab, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice, P.stream));
        MT_HIP(hipStreamSynchronize(P.stream));
      } else {
        int rc = launch_near_table(P.nx, P.ny, P.nz, q->Lx, q->Ly, q->Lz, R, Rz, thr, P.near_tab, P.stream);
        if (rc) return rc;
      }
    }
    MT_HIP(hipMalloc(&P.hbuf, (size_t)(P.ksplit + (P.near_r > 0 ? 1 : 0)) * P.hcols * n3 * sizeof(float)));
  } else if (P.sparse.blocks || P.Kf32) {
    // FP32 tiles or the FP32 MFMA dense tensor: FP32 state fragments, the
    // schedule's (or the dense kernel's) partial-sum slots
    P.ksplit = P.sparse.blocks ? P.sparse.blocks->ksplit : (q->ksplit > 0 ? q->ksplit : gemm_ksplit_dense(n, n));
    MT_HIP(hipMalloc(&P.packb, (size_t)packed_m_bytes(n)));
    MT_HIP(hipMemsetAsync(P.packb, 0, (size_t)packed_m_bytes(n), P.stream));
    MT_HIP(hipMalloc(&P.hbuf, (size_t)P.ksplit * P.hcols * n3 * sizeof(float)));
  } else {
    P.ksplit = 1;
    P.hcols = P.B;
    MT_HIP(hipMalloc(&P.packf, (size_t)P.B * n3 * sizeof(float)));
    MT_HIP(hipMalloc(&P.hbuf, (size_t)P.B * n3 * sizeof(float)));
  }
  const int bps = vec_blocks_per_state((int)n3);
  MT_HIP(hipMalloc(&P.partials, (size_t)3 * P.B * bps * 4 * sizeof(double)));
  MT_HIP(hipMalloc(&P.results, (size_t)3 * P.B * 4 * sizeof(double)));
  MT_HIP(hipHostMalloc(&P.h_results, (size_t)3 * P.B * 4 * sizeof(double), hipHostMallocDefault));
  // one tick stages at most one vector-only launch over all states and one
  // demag pass of up to 64 states (three descriptor arrays)
  P.arena_size = std::max((size_t)8 << 20, (size_t)(P.B + 2 * kMaxCols) * sizeof(VecDesc) +
                                               (size_t)kMaxCols * sizeof(LlgDesc) + 4096);
  MT_HIP(hipHostMalloc(&P.arena_h, P.arena_size, hipHostMallocDefault));
  MT_HIP(hipMalloc(&P.arena_d, P.arena_size));
  P.t_out_h.assign((size_t)P.B * P.n_h * P.nt, 0.0);
  // states
  P.sims.resize(P.B);
  for (int b = 0; b < P.B; ++b) {
    Sim& s = P.sims[b];
    s.id = b;
    double* base = P.pool + (size_t)b * kSlots * slot;
    int k = 0;
    auto next = [&]() { return base + (size_t)(k++) * slot; };
    s.y = next(); s.yp = next(); s.y_old = next(); s.scratch = next(); s.errest = next(); s.weights = next();
    for (int j = 1; j <= 7; ++j) s.stage[j] = next();
    s.xstage = next(); s.ytemp = next(); s.arg = next(); s.y_step = next(); s.y_last = next();
    for (int j = 1; j <= 5; ++j) s.p[j] = next();
    const double* hb = q->hext + (size_t)(q->hext_per_state ? b : 0) * q->nt_hext * 4;
    for (int r = 0; r < q->nt_hext; ++r) {
      s.ht.push_back(hb[r * 4]); s.hx.push_back(hb[r * 4 + 1]);
      s.hy.push_back(hb[r * 4 + 2]); s.hz.push_back(hb[r * 4 + 3]);
    }
    const double* m0 = q->m0 + (size_t)(q->m0_per_state ? b : 0) * n3;
    MT_HIP(hipMemcpyAsync(s.y_step, m0, n3 * sizeof(double), hipMemcpyHostToDevice, P.stream));
    s.budget = q->max_steps > 0 ? q->max_steps : 0;
  }
  MT_HIP(hipStreamSynchronize(P.stream));
  *out = reinterpret_cast<mt_solver*>(S.release());
  return kOk;
}

int mt_solver_advance(mt_solver* h, long long steps) {
  using namespace mt;
  if (!h) {
    set_error("null solver");
    return kErrArg;
  }
  Solver& P = *reinterpret_cast<Solver*>(h);
  for (Sim& s : P.sims) s.budget = steps > 0 ? s.steps_run + steps : 0;
  return P.run_until_idle();
}

int mt_solver_run(mt_solver* h) { return mt_solver_advance(h, 0); }

int mt_solver_advance_total(mt_solver* h, long long total_steps) {
  using namespace mt;
  if (!h || total_steps <= 0) {
    set_error("mt_solver_advance_total: null solver or non-positive step count");
    return kErrArg;
  }
  Solver& P = *reinterpret_cast<Solver*>(h);
  long long acc = 0;
  for (Sim& s : P.sims) {
    s.budget = 0;
    acc += s.steps_run;
  }
  P.total_target = acc + total_steps;
  int rc = P.run_until_idle();
  P.total_target = 0;
  return rc;
}

long long mt_solver_steps_total(mt_solver* h) {
  long long acc = 0;
  for (const mt::Sim& s : reinterpret_cast<Solver*>(h)->sims) acc += s.steps_run;
  return acc;
}

int mt_solver_results(mt_solver* h, double* t_out, double* y_out, double* fields, mt_stats* stats) {
  using namespace mt;
  if (!h) {
    set_error("null solver");
    return kErrArg;
  }
  Solver& P = *reinterpret_cast<Solver*>(h);
  const size_t ysz = (size_t)P.B * P.n_h * P.nt * P.n3();
  if (t_out) std::memcpy(t_out, P.t_out_h.data(), P.t_out_h.size() * sizeof(double));
  if (y_out) MT_HIP(hipMemcpyAsync(y_out, P.yout, ysz * sizeof(double), hipMemcpyDeviceToHost, P.stream));
  if (fields) {
    MT_CHECK(P.fout != nullptr, "fields were not requested at creation");
    MT_HIP(hipMemcpyAsync(fields, P.fout, ysz * 4 * sizeof(double), hipMemcpyDeviceToHost, P.stream));
  }
  MT_HIP(hipStreamSynchronize(P.stream));
  if (stats) {
    for (int b = 0; b < P.B; ++b) {
      const Sim& s = P.sims[b];
      mt_stats& o = stats[b];
      o.f_calls = s.tot_f + (s.done ? 0 : s.full_f_count + s.f_count);
      o.steps = s.tot_steps + (s.done ? 0 : s.step_count);
      o.bad_steps = s.tot_bad + (s.done ? 0 : s.bad_step_count);
      o.flag3 = s.flag3;
      o.flag4 = s.flag4;
      o.h_last = s.h;
      o.t_reached = s.t;
      o.evals = s.evals;
      o.ticks = P.ticks;
      o.demag_passes = P.demag_passes;
      o.demag_ms = P.gemm_ms;
      o.near_ms = P.near_ms;
      o.vec_ms = P.vec_ms;
      o.llg_ms = P.llg_ms;
      o.vec_bytes = P.vec_bytes;
      o.llg_bytes = P.llg_bytes;
      o.vec_launches = P.vec_launches;
      o.llg_launches = P.llg_launches;
    }
  }
  return kOk;
}

// Copies the current state vector y of every state (B x 3n) to the host.
int mt_solver_state(mt_solver* h, double* y) {
  using namespace mt;
  Solver& P = *reinterpret_cast<Solver*>(h);
  for (int b = 0; b < P.B; ++b)
    MT_HIP(hipMemcpyAsync(y + (size_t)b * P.n3(), P.sims[b].y, P.n3() * sizeof(double), hipMemcpyDeviceToHost,
                          P.stream));
  MT_HIP(hipStreamSynchronize(P.stream));
  return kOk;
}

int mt_solver_set_timing(mt_solver* h, int on) {
  mt::Solver& P = *reinterpret_cast<mt::Solver*>(h);
  P.timing = on != 0;
  P.gemm_ms = P.near_ms = P.vec_ms = P.llg_ms = P.vec_bytes = P.llg_bytes = 0.0;
  P.vec_launches = P.llg_launches = 0;
  return 0;
}

int mt_solver_demag_info(mt_solver* h, long long* nnz6, float* threshold) {
  MT_CHECK(h, "null solver");
  mt::Solver& P = *reinterpret_cast<mt::Solver*>(h);
  MT_CHECK(P.demag_approx != MT_DEMAG_DENSE, "the solver uses the dense demag tensor");
  for (int c = 0; c < 6; ++c)
    if (nnz6) nnz6[c] = P.sparse.nnz[c];
  if (threshold) *threshold = P.sparse.threshold;
  return mt::kOk;
}

int mt_solver_demag_form(mt_solver* h, int* form) {
  MT_CHECK(h && form, "null solver or output");
  mt::Solver& P = *reinterpret_cast<mt::Solver*>(h);
  if (P.demag_approx == MT_DEMAG_DENSE) *form = 0;
  else if (!P.sparse.blocks) *form = 1;
  else *form = P.sparse.blocks->f32 ? 3 : 2;
  return mt::kOk;
}

void mt_solver_destroy(mt_solver* h) { delete reinterpret_cast<Solver*>(h); }

int mt_run_micromag(const mt_micromag_problem* q, double* t_out, double* y_out, double* fields, mt_stats* stats) {
  mt_solver* h = nullptr;
  int rc = mt_solver_create(q, &h);
  if (rc) return rc;
  rc = mt_solver_run(h);
  if (rc == 0) rc = mt_solver_results(h, t_out, y_out, fields, stats);
  mt_solver_destroy(h);
  return rc;
}

}  // extern "C"
