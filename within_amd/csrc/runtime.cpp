// runtime.cpp -- the C ABI (include/within_amd.h): device upload of the plan
// and the solver loops (block CG / GMRES, FP64 iterative refinement) that
// drive the kernels.  Mirrors Solver::solve / solve_batch
// (crates/within/src/solver.rs:158-347) for k right-hand sides per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/within_amd.h"
#include "kernels.hpp"
#include "coarse.hpp"
#include "gmres.hpp"
#include "multi.hpp"
#include "plan.hpp"

using namespace wa;

namespace {

thread_local std::string g_err = "";

struct DevError : std::runtime_error {
  explicit DevError(const std::string& m) : std::runtime_error(m) {}
};

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      throw DevError(std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);      \
  } while (0)

int set_error(const std::exception& e) {
  g_err = e.what();
  if (auto* we = dynamic_cast<const Error*>(&e)) return we->code;
  if (dynamic_cast<const DevError*>(&e)) return WA_ERR_DEVICE;
  return WA_ERR_INVALID_ARGUMENT;
}

constexpr int KMAX = 32;
constexpr double F64_EPS = 2.220446049250313e-16;

int pow2_at_least(int k) {
  int p = 1;
  while (p < k) p <<= 1;
  return p;
}

}  // namespace

struct wa_design {
  Design d;
};

struct Timer {
  bool on = false;
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  // classes: 0 implicit operator, 1 preconditioner, 2 vector work,
  // 3 observation space, 4 explicit Gramian, 5 coarse correction
  int64_t launches[6] = {0, 0, 0, 0, 0, 0};
  double ms[6] = {0, 0, 0, 0, 0, 0};
  hipEvent_t get() {
    if (pool.empty()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  void flush() {
    for (auto& r : pending) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.cls] += t;
      launches[r.cls]++;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  ~Timer() {
    for (auto& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct wa_solver {
  // one call at a time uses the handle's stream and work buffers: calls from
  // several host threads on one handle are serialised (the reference's
  // Solver is shared across threads, solver.rs:242-277)
  mutable std::mutex call_mu;
  Plan plan;
  wa_params params{};
  bool host_only = false;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  int64_t n = 0, m = 0;      // observations, internal DOFs (active levels)
  int64_t mu = 0;            // user DOFs (every level of every factor)
  const uint32_t* int_of = nullptr;  // user DOF -> internal (UINT32_MAX = inactive)
  const uint32_t* ext_of = nullptr;  // internal -> user DOF
  std::vector<uint32_t> h_int_of, h_ext_of;  // host copies of the two maps
  int Q = 0;
  int dev = 0;                // the device the solver's state lives on
  // design
  uint32_t* codes = nullptr;  // Q x n (original order)
  uint32_t* offs = nullptr;
  double* wobs = nullptr;     // weights in observation order (null = unit)
  std::vector<DevPass<double>> p64;
  std::vector<DevPass<float>> p32;
  bool has_pre = false, has_gram = false;
  DevPrecond<double> pre64{};
  DevPrecond<float> pre32{};
  DevCsr<double> g64{};
  DevCsr<float> g32{};
  // workspaces (kp <= 32)
  double *Yr = nullptr, *E = nullptr;
  double *B64 = nullptr, *X64 = nullptr, *R64 = nullptr, *T64 = nullptr;
  double *cx64 = nullptr, *cr64 = nullptr, *cz64 = nullptr, *cp64 = nullptr, *cap64 = nullptr;
  float *cx32 = nullptr, *cr32 = nullptr, *cz32 = nullptr, *cp32 = nullptr, *cap32 = nullptr;
  float* rhs32 = nullptr;
  PrecondWork<double> w64{};
  PrecondWork<float> w32{};
  double* part64 = nullptr;  // item partials (FP64; FP32 kernels use it reinterpreted)
  double* redp = nullptr;    // reduction partials
  double* sc = nullptr;      // small per-column scalars
  int* imask = nullptr;
  ColState* cs = nullptr;
  // LSMR (allocated on its first solve): sqrt weights per observation,
  // observation-space partials, per-column state, reorthogonalisation ring
  double* swobs = nullptr;
  double* lpart = nullptr;
  double* lcoef = nullptr;
  LsmrCol* lcs = nullptr;
  double *ring_v = nullptr, *ring_p = nullptr;
  int64_t ring_cap = 0;
  double* hbuf = nullptr;    // pinned host scratch
  Timer timer;
  // GMRES bases, allocated on first use
  template <typename T>
  struct GmresBufs {
    int restart = 0;
    T *V = nullptr, *Z = nullptr, *w = nullptr, *r = nullptr;
    double* hd = nullptr;
  };
  // captured CG iterations, one per (precision, block width, x vector, side
  // x-update, operator representation); replayed with hipGraphLaunch
  struct GraphKey {
    // kind: 0 block CG, 1 per-problem (stacked) CG, 2 flexible block CG with
    // the coarse correction
    int kind, t, kp, op, side;
    const void* x;
  };
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
  bool use_graphs() const;
  void graphs_clear() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (hipEvent_t e : gmd.ev) (void)hipEventDestroy(e);
    graphs.clear();
  }
  template <typename T, typename F>
  hipGraphExec_t cg_graph(int kp, T* x, bool side, F&& body, int kind = 0);
  std::vector<PairSched> pairs;            // per-factor-pair subranges (sweep / overlapped apply)
  // additive apply with the pairs' local solves on two streams: the lead
  // pair (longest level chain) on st, the others on st2
  bool overlap = false;
  size_t lead = 0;
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_upload = nullptr;  // the null stream's uploads, before device casts on st
  double* part2 = nullptr;                 // st2's split-row partials
  // the lead pair's trailing levels as one launch per direction (engine.hpp
  // k_level_flow): counters of the forward and backward runs, the previous
  // apply's ticket totals
  uint32_t* flow_ctr = nullptr;
  int flow_nl = 0;
  bool flow_used = false;
  uint32_t flow_prev_f = 0, flow_prev_b = 0;
  double *zc2_64 = nullptr, *p2_64 = nullptr;  // st2's dense-core scratch
  float *zc2_32 = nullptr, *p2_32 = nullptr;
  double* mw64 = nullptr;                  // r_work, delta, G delta (multiplicative)
  float* mw32 = nullptr;
  double* ones = nullptr;
  // Gramian diagonal of each internal DOF and the internal offset of each
  // factor (inputs of the coarse-space network)
  std::vector<double> h_diag;
  std::vector<int64_t> h_foff;
  // learned coarse-space correction (coarse.hpp), allocated on first use:
  // network inputs and activations (m x 32 FP32), training samples, the
  // parameters with their Adam moments, and the coarse space of the apply
  struct CoarseNet {
    bool ready = false, trained = false, enabled = false;
    uint64_t seed = 0;
    float *X0 = nullptr, *dinv = nullptr, *cvec = nullptr;
    float *H[kGnnLayers] = {}, *S[kGnnLayers] = {}, *T[kGnnLayers] = {};
    float *P = nullptr, *GP = nullptr, *E = nullptr, *R = nullptr, *dP = nullptr, *dZa = nullptr, *dZb = nullptr;
    float *Ys = nullptr, *Tb = nullptr;
    float *theta = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *xpart = nullptr;
    double* W12 = nullptr;
    double ridge = 1e-4;
    bf16_t* pack = nullptr;
    double *xpart64 = nullptr, *Ac = nullptr, *U = nullptr, *V = nullptr, *loss = nullptr;
    double nu = 0.0;
    int kept = 0;  // coarse vectors kept after G-orthonormalisation
    bf16_t *Pb = nullptr, *GPb = nullptr;
    double *Ainv = nullptr, *cpart = nullptr, *cmid = nullptr, *c1 = nullptr, *cd = nullptr, *r2 = nullptr;
    double* dpart = nullptr;
    std::vector<float> h_feat;
  } cn;
  bool fcg = false;  // set by cg() while the coarse correction is part of its preconditioner
  // device GMRES state (allocated on first use): column states, Hessenberg
  // columns, the two Gram-Schmidt passes' coefficients, y, per-column scalars,
  // reduction partials, per-step extension counts and their events
  struct GmDev {
    GmCol* cols = nullptr;
    double *H = nullptr, *h = nullptr, *h2 = nullptr, *y = nullptr, *coef = nullptr, *beta = nullptr;
    double *abs_tol = nullptr, *part = nullptr, *part2 = nullptr;
    int *live = nullptr, *lim = nullptr, *ninc = nullptr;
    std::vector<hipEvent_t> ev;
  } gmd;
  GmDev& gm_dev();
  GmresBufs<double> gm64;
  GmresBufs<float> gm32;
  template <typename T> GmresBufs<T>& gm(int restart);
  // serialized preconditioner (built on first request)
  mutable std::vector<uint8_t> precond_bytes;
  // autotuning of the operator schedules: host row pointers of the Gramian and
  // of the factor passes (internal numbering), capacity of part64 in slots
  std::vector<uint32_t> h_gram_ptr;
  std::vector<std::vector<uint32_t>> h_pass_ptr;
  int64_t part_cap = 0;
  void ensure_part(int64_t slots) {
    if (slots <= part_cap) return;
    part64 = dalloc<double>(slots * KMAX);
    part_cap = slots;
    w64.part = part64;
    w32.part = (float*)part64;
  }
  // device user buffers for host-pointer calls
  double* ubuf = nullptr;
  int64_t ubuf_len = 0;
  double* staging(int64_t need) {
    if (need > ubuf_len) {
      if (ubuf) HIPCHK(hipFree(ubuf));
      ubuf = nullptr;
      ubuf_len = 0;
      HIPCHK(hipMalloc(&ubuf, sizeof(double) * need));
      ubuf_len = need;
    }
    return ubuf;
  }

  template <typename T>
  T* dalloc(int64_t count) {
    if (count <= 0) count = 1;
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(T) * (size_t)count));
    allocs.push_back(p);
    return (T*)p;
  }
  template <typename T>
  T* upload(const std::vector<T>& v) {
    T* p = dalloc<T>((int64_t)v.size());
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return p;
  }
  // an FP32 copy, made on the device, of n doubles already uploaded (on the
  // null stream: st waits for that stream's work first)
  float* narrowed(const double* d, int64_t n) {
    float* p = dalloc<float>(n);
    if (!ev_upload) HIPCHK(hipEventCreateWithFlags(&ev_upload, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev_upload, nullptr));
    HIPCHK(hipStreamWaitEvent(st, ev_upload, 0));
    launch_narrow(st, d, p, n);
    return p;
  }
  // schedule shape on stderr when WITHIN_AMD_VERBOSE is set
  static void sched_report(const char* name, const Schedule& s, const std::vector<uint32_t>& ptr) {
    if (!std::getenv("WITHIN_AMD_VERBOSE")) return;
    int64_t rows = 0, shortr = 0, elems = 0, short_e = 0, whole = 0, split = 0;
    for (const Segment& g : s.segs)
      for (uint32_t r = g.r0; r < g.r1; ++r) {
        const int64_t len = ptr[r + 1] - ptr[r];
        ++rows;
        elems += len;
        if (len <= (int64_t)s.short_max) { ++shortr; short_e += len; }
      }
    for (const Item& it : s.items) (it.slot < 0 ? whole : split)++;
    std::fprintf(stderr,
                 "[within_amd] schedule %-8s segs %6zu rows %9lld (short %9lld, %10lld elems) elems %10lld "
                 "items %8lld whole + %8lld chunks, fixups %6zu\n",
                 name, s.segs.size(), (long long)rows, (long long)shortr, (long long)short_e, (long long)elems,
                 (long long)whole, (long long)split, s.fixups.size());
  }
  DevSched upload_sched(const Schedule& s) {
    DevSched d;
    d.items = upload(s.items);
    d.n_items = (int64_t)s.items.size();
    d.fix = upload(s.fixups);
    d.n_fix = (int64_t)s.fixups.size();
    d.n_slots = s.n_slots;
    d.short_max = s.short_max;
    d.segs = s.segs;
    return d;
  }
  ~wa_solver() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (void* p : allocs) (void)hipFree(p);
    if (ubuf) (void)hipFree(ubuf);
    if (st2) (void)hipStreamSynchronize(st2);
    if (hbuf) (void)hipHostFree(hbuf);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_upload) (void)hipEventDestroy(ev_upload);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }

  // scalar scratch layout (each KMAX doubles)
  double* s_(int i) { return sc + (int64_t)i * KMAX; }

  void tic(int cls, hipEvent_t& a) {
    if (!timer.on) return;
    a = timer.get();
    HIPCHK(hipEventRecord(a, st));
  }
  void toc(int cls, hipEvent_t a) {
    if (!timer.on) return;
    hipEvent_t b = timer.get();
    HIPCHK(hipEventRecord(b, st));
    timer.pending.push_back({cls, a, b});
  }

  void setup_device();
  void setup_overlap();
  void setup_flow();
  void check_flow();
  template <typename T> FlowCall flow_call(int kp);
  template <typename T> void op_apply(const T* X, T* Y, int kp);
  template <typename T> void pre_apply(const T* R, T* Z, int kp, const SideWork* side = nullptr);
  template <typename T> void precondition(const T* R, T* Z, int kp, const SideWork* side = nullptr);
  void rhs64(const double* V, double* out, int kp);
  template <typename T>
  void cg(int k, int kp, const T* rhs, T* x, const std::vector<double>& tol, std::vector<int>& conv,
          std::vector<int64_t>& its);
  template <typename T>
  void gmres_dev(int k, int kp, const T* rhs, T* x, const std::vector<double>& tol, std::vector<int>& conv,
                 std::vector<int64_t>& its);
  template <typename T>
  void gmres(int k, int kp, const T* rhs, T* x, const std::vector<double>& tol, std::vector<int>& conv,
             std::vector<int64_t>& its);
  template <typename T>
  void krylov(int k, int kp, const T* rhs, T* x, const std::vector<double>& tol, std::vector<int>& conv,
              std::vector<int64_t>& its) {
    if (params.krylov == 1 && params.restart < kGmresMaxV) gmres_dev<T>(k, kp, rhs, x, tol, conv, its);
    else if (params.krylov == 1) gmres<T>(k, kp, rhs, x, tol, conv, its);
    else cg<T>(k, kp, rhs, x, tol, conv, its);
  }
  void lsmr(int k, int kp, std::vector<int>& conv, std::vector<int64_t>& its);
  // learned coarse correction
  bool coarse_supported() const { return params.preconditioner == 1 && has_pre && m > 0; }
  void coarse_alloc(uint64_t seed);
  void coarse_init_params(uint64_t seed);
  void coarse_samples(int iters, uint64_t seed);
  void gnn_forward();
  void gnn_loss(bool grad);
  double gnn_read_loss();
  void coarse_finalize();
  double coarse_validate(bool with, std::vector<int64_t>& its);
  template <typename T> void coarse_apply(const T* R, T* Z, int kp, const SideWork* side);
  void solve_block(const double* Y, int64_t ldy, int c0, int kb, double* X, int64_t ldx, double* dem,
                   int64_t ldd, wa_rhs_stats* stats);
  std::vector<double> col_norms(const double* V, int64_t rows, int kp);
  template <typename T> T* vec(int which);
};

template <typename T>
static void alloc_gmres(wa_solver* s, wa_solver::GmresBufs<T>& b, int restart) {
  if (b.restart >= restart) return;
  // bases sized for the largest block (KMAX columns); freed with the solver
  b.V = s->dalloc<T>((int64_t)(restart + 1) * s->m * KMAX);
  b.Z = s->dalloc<T>((int64_t)restart * s->m * KMAX);
  b.w = s->dalloc<T>(s->m * KMAX);
  b.r = s->dalloc<T>(s->m * KMAX);
  b.hd = s->dalloc<double>((int64_t)(restart + 2) * KMAX);
  b.restart = restart;
}
template <> wa_solver::GmresBufs<double>& wa_solver::gm<double>(int restart) {
  alloc_gmres(this, gm64, restart);
  return gm64;
}
template <> wa_solver::GmresBufs<float>& wa_solver::gm<float>(int restart) {
  alloc_gmres(this, gm32, restart);
  return gm32;
}

template <> double* wa_solver::vec<double>(int w) {
  switch (w) { case 0: return cx64; case 1: return cr64; case 2: return cz64; case 3: return cp64; default: return cap64; }
}
template <> float* wa_solver::vec<float>(int w) {
  switch (w) { case 0: return cx32; case 1: return cr32; case 2: return cz32; case 3: return cp32; default: return cap32; }
}

// Smallest relative residual asked of an FP32 inner solve (FP32 recurrences
// stay reliable well above this); WITHIN_AMD_INNER_FLOOR overrides it.
static double inner_floor() {
  static double v = [] {
    const char* e = std::getenv("WITHIN_AMD_INNER_FLOOR");
    const double d = e ? std::atof(e) : 0.0;
    return d > 0.0 && d < 1.0 ? d : 1e-5;
  }();
  return v;
}

// A column of the mixed path whose FP64 residual falls by less than this
// factor over one FP32 correction has stalled (WITHIN_AMD_STALL_RATIO
// overrides; read on every solve)
static double stall_ratio() {
  const char* e = std::getenv("WITHIN_AMD_STALL_RATIO");
  const double d = e ? std::atof(e) : 0.0;
  return d > 0.0 && d < 1.0 ? d : 0.9;
}

static void check_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    throw DevError(std::string("within_amd requires a gfx950 (MI355X) device, found ") + prop.gcnArchName);
}

void wa_solver::setup_device() {
  const Design& d = plan.d;
  const bool report = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
  auto t_ph = std::chrono::steady_clock::now();
  auto phase = [&](const char* what) {
    if (!report) return;
    (void)hipDeviceSynchronize();
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[within_amd] device %s: %.3f s\n", what, std::chrono::duration<double>(t1 - t_ph).count());
    t_ph = t1;
  };
  check_device();
  HIPCHK(hipGetDevice(&dev));
  phase("init");
  DeviceLayout lay;
  // the preconditioner's arrays are laid out (host work, writing lay.pre only)
  // on a thread of their own while the operator's are uploaded; joined before
  // the preconditioner's upload and on every way out
  struct Joiner {
    std::thread t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } pre_lay;
  std::string pre_err;
  int pre_code = 0;
  if (plan.lay_op) {
    // the operator part was laid out while the preconditioner was built
    lay = std::move(*plan.lay_op);
    plan.lay_op.reset();
    pre_lay.t = std::thread([this, &lay, &pre_err, &pre_code] {
      try {
        layout_precond(plan, lay);
      } catch (const Error& e) {
        pre_err = e.what();
        pre_code = e.code;
      } catch (const std::exception& e) {
        pre_err = e.what();
        pre_code = WA_ERR_INVALID_ARGUMENT;
      }
    });
  } else {
    lay = make_device_layout(plan);
  }
  phase("layout");
  n = d.n_obs;
  m = lay.m;
  mu = d.n_dofs;
  Q = d.Q;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  int_of = upload(lay.int_of);
  h_int_of = lay.int_of;
  h_ext_of = lay.ext_of;
  ext_of = upload(lay.ext_of);
  // design (original observation order) for D x and demeaning: the internal
  // DOF of every (factor, observation), so the factor offsets are zero
  {
    codes = upload(lay.dof_of_obs);
    offs = upload(std::vector<uint32_t>(Q, 0u));
    if (!d.weights.empty()) wobs = upload(d.weights);
  }
  int64_t max_slots = 1;
  h_foff = lay.off;
  h_diag.assign((size_t)m, 0.0);
  for (const FactorPass& fp : lay.passes)
    for (size_t j = 0; j < fp.diag.size(); ++j) h_diag[(size_t)lay.off[fp.q] + j] = fp.diag[j];
  for (const FactorPass& fp : lay.passes) {
    DevPass<double> a;
    DevPass<float> b;
    a.row_ptr = b.row_ptr = upload(fp.row_ptr);
    a.other = b.other = upload(fp.other);
    a.obs = b.obs = upload(fp.obs);
    if (!fp.w.empty()) {
      a.ew = upload(fp.w);
      b.ew = narrowed(a.ew, (int64_t)fp.w.size());
    }
    a.diag = upload(fp.diag);
    b.diag = narrowed(a.diag, (int64_t)fp.diag.size());
    a.qm = b.qm = Q - 1;
    a.off = b.off = (uint32_t)lay.off[fp.q];
    a.levels = b.levels = lay.levels[fp.q];
    a.n_elem = b.n_elem = n;
    {
      h_pass_ptr.push_back(fp.row_ptr);
      Schedule sc = make_schedule(fp.row_ptr, kChunk, kShortPass);
      sched_report("pass", sc, fp.row_ptr);
      a.sched = b.sched = upload_sched(sc);
    }
    max_slots = std::max(max_slots, a.sched.n_slots);
    p64.push_back(a);
    p32.push_back(b);
  }
  phase("operator passes");
  if (plan.has_gram) {
    const Csr& g = lay.gram;
    g64.ptr = g32.ptr = upload(g.indptr);
    g64.idx = g32.idx = upload(g.indices);
    g64.val = upload(g.data);
    g32.val = narrowed(g64.val, (int64_t)g.data.size());
    g64.nrows = g32.nrows = g.nrows;
    g64.nnz = g32.nnz = (int64_t)g.indices.size();
    {
      h_gram_ptr = g.indptr;
      Schedule sc = make_schedule(g.indptr, kGramChunk, kShortRow);
      sched_report("gramian", sc, g.indptr);
      g64.sched = g32.sched = upload_sched(sc);
    }
    max_slots = std::max(max_slots, g64.sched.n_slots);
    has_gram = true;
  }
  phase("gramian");
  if (pre_lay.t.joinable()) pre_lay.t.join();
  if (pre_code) throw Error(pre_code, pre_err);
  phase("preconditioner layout");
  if (!plan.meta.empty()) {
    const PrecondFlat& f = lay.pre;
    // dense cores: the large ones are applied by the dense kernels from
    // zero-padded copies; the rest keep the gather path (their matrices packed)
    const size_t n_cores = f.core_ptr.empty() ? 0 : f.core_ptr.size() - 1;
    std::vector<char> big_row(f.core_idx.size(), 0);
    std::vector<double> small_mat;
    std::vector<uint64_t> small_off(n_cores, 0);
    std::vector<size_t> big_cores;
    for (size_t c = 0; c < n_cores; ++c) {
      const uint32_t b = f.core_ptr[c], nc = f.core_ptr[c + 1] - b;
      if (nc >= kDenseCore) {
        big_cores.push_back(c);
        std::fill(big_row.begin() + b, big_row.begin() + b + nc, 1);
      } else {
        const double* src = f.core_mat.data() + f.core_mat_off[c];
        small_off[c] = small_mat.size();
        small_mat.insert(small_mat.end(), src, src + (size_t)nc * nc);
      }
    }
    auto upload_big = [&](auto& P, auto tag) {
      using T = decltype(tag);
      for (size_t c : big_cores) {
        DenseCore<T> d;
        d.b = f.core_ptr[c];
        d.n = f.core_ptr[c + 1] - d.b;
        d.ld = (d.n + 31) / 32 * 32;
        if (const DeviceDense* dv = device_core(f, c)) {
          // factored on the device: padded (and, for FP32, rounded) there,
          // once the inverse's own kernels are done
          dv->ready();
          T* M = dalloc<T>((int64_t)d.n * d.ld);
          launch_pad_core<T>(st, static_cast<const double*>(dv->p), d.n, M, d.ld);
          d.M = M;
        } else {
          std::vector<T> m((size_t)d.n * d.ld, T(0));
          const double* src = f.core_mat.data() + f.core_mat_off[c];
          for (size_t i = 0; i < d.n; ++i)
            for (size_t j = 0; j < d.n; ++j) m[i * d.ld + j] = (T)src[i * d.n + j];
          d.M = upload(m);
        }
        P.big.push_back(d);
      }
    };
    upload_big(pre64, double());
    upload_big(pre32, float());
    auto sizes = [&](auto& P) {
      P.n_loc = f.n_loc;
      P.n_red = f.n_red;
      P.n_sub = f.n_sub;
      P.n_elim = (int64_t)f.elim_loc.size();
      P.n_levels = (int64_t)f.lvl_ptr.size() - 1;
      P.n_core_rows = (int64_t)f.core_idx.size();
    };
    sizes(pre64);
    sizes(pre32);
    // the FP64 values go up; their FP32 copies are made on the device
    pre64.loc_scale = upload(f.loc_scale);
    pre64.elim_inv = upload(f.elim_inv);
    pre64.elim_val = upload(f.elim_val);
    pre64.keep_val = upload(f.keep_val);
    pre64.keep_rowsum = upload(f.keep_rowsum);
    pre64.fin_val = upload(f.fin_val);
    pre64.bout_val = upload(f.bout_val);
    pre64.inv_pivot = upload(f.inv_pivot);
    pre64.core_mat = upload(small_mat);
    pre32.loc_scale = narrowed(pre64.loc_scale, (int64_t)f.loc_scale.size());
    pre32.elim_inv = narrowed(pre64.elim_inv, (int64_t)f.elim_inv.size());
    pre32.elim_val = narrowed(pre64.elim_val, (int64_t)f.elim_val.size());
    pre32.keep_val = narrowed(pre64.keep_val, (int64_t)f.keep_val.size());
    pre32.keep_rowsum = narrowed(pre64.keep_rowsum, (int64_t)f.keep_rowsum.size());
    pre32.fin_val = narrowed(pre64.fin_val, (int64_t)f.fin_val.size());
    pre32.bout_val = narrowed(pre64.bout_val, (int64_t)f.bout_val.size());
    pre32.inv_pivot = narrowed(pre64.inv_pivot, (int64_t)f.inv_pivot.size());
    pre32.core_mat = narrowed(pre64.core_mat, (int64_t)small_mat.size());
    // integer arrays shared by both precisions
    const uint32_t* loc_gidx = upload(f.loc_gidx);
    const uint32_t* loc_ptr = upload(f.loc_ptr);
    const uint32_t* red_ptr = upload(f.red_ptr);
    const uint32_t* loc_sub = upload(f.loc_sub);
    const uint32_t* red_sub = upload(f.red_sub);
    const uint32_t* red_loc = upload(f.red_loc);
    const uint32_t* elim_loc = upload(f.elim_loc);
    const uint32_t* elim_ptr = upload(f.elim_ptr);
    const uint32_t* elim_idx = upload(f.elim_idx);
    const uint32_t* keep_ptr = upload(f.keep_ptr);
    const uint32_t* keep_idx = upload(f.keep_idx);
    const uint32_t* lvl_rows = upload(f.lvl_rows);
    const uint32_t* fin_ptr = upload(f.fin_ptr);
    const uint32_t* fin_idx = upload(f.fin_idx);
    const uint32_t* bout_ptr = upload(f.bout_ptr);
    const uint32_t* bout_idx = upload(f.bout_idx);
    const uint32_t* core_ptr = upload(f.core_ptr);
    const uint32_t* core_idx = upload(f.core_idx);
    const uint32_t* core_row_core = upload(f.core_row_core);
    const uint64_t* core_mat_off = upload(small_off);
    const uint32_t* dof_ptr = upload(f.dof_ptr);
    const uint32_t* dof_loc = upload(f.dof_loc);
    auto ints = [&](auto& P) {
      P.loc_gidx = loc_gidx; P.loc_ptr = loc_ptr; P.red_ptr = red_ptr; P.loc_sub = loc_sub;
      P.red_sub = red_sub; P.red_loc = red_loc; P.elim_loc = elim_loc; P.elim_ptr = elim_ptr;
      P.elim_idx = elim_idx; P.keep_ptr = keep_ptr; P.keep_idx = keep_idx; P.lvl_rows = lvl_rows;
      P.fin_ptr = fin_ptr; P.fin_idx = fin_idx; P.bout_ptr = bout_ptr; P.bout_idx = bout_idx;
      P.core_ptr = core_ptr; P.core_idx = core_idx; P.core_row_core = core_row_core;
      P.core_mat_off = core_mat_off; P.dof_ptr = dof_ptr; P.dof_loc = dof_loc;
    };
    // dense-core rows: row r of core c gathers |c| elements
    std::vector<uint32_t> core_row_ptr(1, 0);
    for (size_t c = 0; c + 1 < f.core_ptr.size(); ++c) {
      uint64_t nc = f.core_ptr[c + 1] - f.core_ptr[c];
      for (uint64_t r = 0; r < nc; ++r) {
        if ((uint64_t)core_row_ptr.back() + nc >= UINT32_MAX) throw Error(WA_ERR_OVERFLOW, "dense cores too large");
        core_row_ptr.push_back(core_row_ptr.back() + (uint32_t)nc);
      }
    }
    const uint32_t* d_core_row_ptr = upload(core_row_ptr);
    auto level_sched = [&](const char* name, const std::vector<uint32_t>& ptr) {
      Schedule sc;
      sc.short_max = kShortLevel;
      for (size_t t = 0; t + 1 < f.lvl_ptr.size(); ++t) append_schedule(sc, ptr, f.lvl_ptr[t], f.lvl_ptr[t + 1], kChunk);
      sched_report(name, sc, ptr);
      return upload_sched(sc);
    };
    auto whole_sched = [&](const char* name, const std::vector<uint32_t>& ptr, const std::vector<char>* skip,
                           uint32_t short_max = kShortRow) {
      Schedule sc = make_schedule(ptr, kChunk, short_max, skip);
      sched_report(name, sc, ptr);
      return upload_sched(sc);
    };
    DevSched s_loc = whole_sched("local", f.loc_ptr, nullptr);
    DevSched s_red = whole_sched("reduced", f.red_ptr, nullptr);
    DevSched s_keep = whole_sched("keep", f.keep_ptr, nullptr);
    DevSched s_elim = whole_sched("elim", f.elim_ptr, nullptr, kShortElim);
    DevSched s_fwd = level_sched("forward", f.fin_ptr);
    DevSched s_bwd = level_sched("backward", f.bout_ptr);
    static_assert(kDenseCore > kShortRow, "dense-core rows must not be short rows");
    DevSched s_core = whole_sched("core", core_row_ptr, &big_row);
    for (const DevSched* d : {&s_loc, &s_red, &s_keep, &s_elim, &s_fwd, &s_bwd, &s_core})
      max_slots = std::max(max_slots, d->n_slots);
    auto scheds = [&](auto& P) {
      P.core_row_ptr = d_core_row_ptr;
      P.s_loc = s_loc; P.s_red = s_red; P.s_keep = s_keep; P.s_elim = s_elim;
      P.s_fwd = s_fwd; P.s_bwd = s_bwd; P.s_core = s_core;
    };
    ints(pre64);
    ints(pre32);
    scheds(pre64);
    scheds(pre32);
    {
      // per-pair subranges (subdomains are ordered by factor pair)
      std::vector<int64_t> elim_ptr_sub(1, 0), core_row_sub(1, 0);
      for (const SubMeta& sd : plan.meta) {
        elim_ptr_sub.push_back(elim_ptr_sub.back() + sd.n_elim);
        core_row_sub.push_back(core_row_sub.back() + sd.n_core);
      }
      const std::vector<SubMeta>& sm = plan.meta;
      size_t s0 = 0;
      while (s0 < sm.size()) {
        size_t s1 = s0;
        while (s1 < sm.size() && sm[s1].q == sm[s0].q && sm[s1].r == sm[s0].r) ++s1;
        PairSched ps;
        ps.loc0 = f.loc_ptr[s0]; ps.loc1 = f.loc_ptr[s1];
        ps.red0 = f.red_ptr[s0]; ps.red1 = f.red_ptr[s1];
        ps.core0 = core_row_sub[s0]; ps.core1 = core_row_sub[s1];
        auto range = [&](const std::vector<uint32_t>& ptr, int64_t r0, int64_t r1,
                         const std::vector<char>* skip = nullptr, uint32_t short_max = kShortRow) {
          Schedule sc;
          sc.short_max = short_max;
          append_schedule(sc, ptr, (u32)r0, (u32)r1, kChunk, skip);
          return upload_sched(sc);
        };
        ps.s_loc = range(f.loc_ptr, (int64_t)s0, (int64_t)s1);
        ps.s_red = range(f.red_ptr, (int64_t)s0, (int64_t)s1);
        ps.s_keep = range(f.keep_ptr, ps.red0, ps.red1);
        ps.s_elim = range(f.elim_ptr, elim_ptr_sub[s0], elim_ptr_sub[s1], nullptr, kShortElim);
        ps.s_core = range(core_row_ptr, ps.core0, ps.core1, &big_row);
        for (size_t bi = 0; bi < pre64.big.size(); ++bi)
          if (pre64.big[bi].b >= ps.core0 && pre64.big[bi].b < ps.core1) ps.big.push_back(bi);
        auto pair_levels = [&](const std::vector<uint32_t>& ptr) {
          Schedule sc;
          sc.short_max = kShortLevel;
          for (size_t t = 0; t + 1 < f.lvl_ptr.size(); ++t) {
            u32 a0 = f.lvl_ptr[t], a1 = f.lvl_ptr[t + 1];
            while (a0 < a1 && f.red_sub[f.lvl_rows[a0]] < s0) ++a0;
            u32 b1 = a0;
            while (b1 < a1 && f.red_sub[f.lvl_rows[b1]] < s1) ++b1;
            append_schedule(sc, ptr, a0, b1, kChunk);
          }
          if (std::getenv("WITHIN_AMD_VERBOSE")) {
            // per level: rows, elements, longest row
            std::fprintf(stderr, "[within_amd] pair %zu %s levels:", pairs.size(), &ptr == &f.fin_ptr ? "fwd" : "bwd");
            for (const Segment& g : sc.segs) {
              int64_t el = 0, mx = 0;
              for (u32 r = g.r0; r < g.r1; ++r) {
                el += ptr[r + 1] - ptr[r];
                mx = std::max<int64_t>(mx, ptr[r + 1] - ptr[r]);
              }
              std::fprintf(stderr, " %u/%lld/%lld", g.r1 - g.r0, (long long)el, (long long)mx);
            }
            std::fprintf(stderr, "\n");
          }
          return upload_sched(sc);
        };
        ps.s_fwd = pair_levels(f.fin_ptr);
        ps.s_bwd = pair_levels(f.bout_ptr);
        for (const DevSched* d : {&ps.s_loc, &ps.s_red, &ps.s_keep, &ps.s_elim, &ps.s_fwd, &ps.s_bwd, &ps.s_core})
          max_slots = std::max(max_slots, d->n_slots);
        pairs.push_back(std::move(ps));
        s0 = s1;
      }
      if (plan.precond == 2) {
        mw64 = dalloc<double>(3 * m * KMAX);
        mw32 = dalloc<float>(3 * m * KMAX);
        std::vector<double> one(KMAX, 1.0);
        ones = upload(one);
      }
    }
    int64_t nsub = std::max<int64_t>(f.n_sub, 1);
    w64.B = dalloc<double>(f.n_loc * KMAX);
    w64.S = dalloc<double>(f.n_loc * KMAX);
    w64.Rr = dalloc<double>(f.n_red * KMAX);
    w64.C = dalloc<double>((int64_t)f.core_idx.size() * KMAX);
    int64_t max_ld = 0;
    for (const DenseCore<double>& d : pre64.big) max_ld = std::max<int64_t>(max_ld, d.ld);
    if (max_ld > 0) {
      w64.Zc = dalloc<double>(max_ld * KMAX);
      w64.P = dalloc<double>(kCoreSplitMax * max_ld * KMAX);
      w32.Zc = dalloc<float>(max_ld * KMAX);
      w32.P = dalloc<float>(kCoreSplitMax * max_ld * KMAX);
    }
    w64.sum1 = dalloc<double>(nsub * KMAX);
    w64.sum2 = dalloc<double>(nsub * KMAX);
    w64.sum3 = dalloc<double>(nsub * KMAX);
    w32.B = dalloc<float>(f.n_loc * KMAX);
    w32.S = dalloc<float>(f.n_loc * KMAX);
    w32.Rr = dalloc<float>(f.n_red * KMAX);
    w32.C = dalloc<float>((int64_t)f.core_idx.size() * KMAX);
    w32.sum1 = dalloc<float>(nsub * KMAX);
    w32.sum2 = dalloc<float>(nsub * KMAX);
    w32.sum3 = dalloc<float>(nsub * KMAX);
    has_pre = true;
    setup_overlap();
    setup_flow();
  }
  phase("preconditioner");
  part64 = dalloc<double>(max_slots * KMAX);
  part_cap = max_slots;
  w64.part = part64;
  w32.part = (float*)part64;
  redp = dalloc<double>((int64_t)kRedBlocks * KMAX);
  sc = dalloc<double>(16 * KMAX);
  imask = dalloc<int>(3 * KMAX);
  cs = dalloc<ColState>(KMAX);
  HIPCHK(hipHostMalloc((void**)&hbuf, sizeof(double) * 16 * KMAX, hipHostMallocDefault));
  // vectors
  Yr = dalloc<double>(n * KMAX);
  E = dalloc<double>(n * KMAX);
  B64 = dalloc<double>(m * KMAX);
  X64 = dalloc<double>(m * KMAX);
  R64 = dalloc<double>(m * KMAX);
  T64 = dalloc<double>(m * KMAX);
  cx64 = dalloc<double>(m * KMAX); cr64 = dalloc<double>(m * KMAX); cz64 = dalloc<double>(m * KMAX);
  cp64 = dalloc<double>(m * KMAX); cap64 = dalloc<double>(m * KMAX);
  cx32 = dalloc<float>(m * KMAX); cr32 = dalloc<float>(m * KMAX); cz32 = dalloc<float>(m * KMAX);
  cp32 = dalloc<float>(m * KMAX); cap32 = dalloc<float>(m * KMAX);
  rhs32 = dalloc<float>(m * KMAX);
  HIPCHK(hipDeviceSynchronize());
  phase("work vectors");
}

// Two-stream additive apply.  The elimination levels of a large reduced
// system are one short launch each, bound by latency rather than bandwidth;
// the other factor pairs' local solves (long Schur-block rows, eliminated
// rows) run beside them on a second stream.  Only when one pair has a level
// chain worth hiding (at least kOverlapLevels forward levels) and the other
// pairs' launches are no more than the lead's; WITHIN_AMD_PRECOND_STREAMS=1
// keeps every apply on one stream.
void wa_solver::setup_overlap() {
  constexpr int kOverlapLevels = 8;
  const char* e = std::getenv("WITHIN_AMD_PRECOND_STREAMS");
  if (plan.precond != 1 || pairs.size() < 2 || (e && std::atoi(e) == 1)) return;
  auto depth = [](const PairSched& ps) {
    int d = 0;
    for (const Segment& g : ps.s_fwd.segs) d += (g.r1 > g.r0) ? 1 : 0;
    return d;
  };
  lead = 0;
  for (size_t i = 1; i < pairs.size(); ++i)
    if (depth(pairs[i]) > depth(pairs[lead]) ||
        (depth(pairs[i]) == depth(pairs[lead]) && pairs[i].red1 - pairs[i].red0 > pairs[lead].red1 - pairs[lead].red0))
      lead = i;
  if (depth(pairs[lead]) < kOverlapLevels) return;
  // the other pairs run as their own launches on st2, where one stream would
  // merge them: only when they fit beside the lead's chain (measured: 3 pairs at
  // 20M obs gain 10 %, 6 pairs at 50M obs lose 4 %)
  auto launches = [](const PairSched& ps) {
    int64_t c = 4;  // mean subtraction, two scatters, the combine's share
    for (const DevSched* d : {&ps.s_loc, &ps.s_loc, &ps.s_keep, &ps.s_red, &ps.s_fwd, &ps.s_bwd, &ps.s_core,
                              &ps.s_elim})
      for (const Segment& g : d->segs) c += (g.r1 > g.r0 || g.i1 > g.i0 ? 1 : 0) + (g.f1 > g.f0 ? 1 : 0);
    return c + 4 * (int64_t)ps.big.size();
  };
  int64_t others = 0;
  for (size_t i = 0; i < pairs.size(); ++i)
    if (i != lead) others += launches(pairs[i]);
  if (others > launches(pairs[lead])) return;
  int64_t slots = 1, ld = 0;
  for (size_t i = 0; i < pairs.size(); ++i) {
    if (i == lead) continue;
    const PairSched& ps = pairs[i];
    for (const DevSched* d : {&ps.s_loc, &ps.s_keep, &ps.s_red, &ps.s_elim, &ps.s_fwd, &ps.s_bwd, &ps.s_core})
      slots = std::max(slots, d->n_slots);
    for (size_t bi : ps.big) ld = std::max<int64_t>(ld, pre64.big[bi].ld);
  }
  part2 = dalloc<double>(slots * KMAX);
  if (ld > 0) {
    zc2_64 = dalloc<double>(ld * KMAX);
    p2_64 = dalloc<double>(kCoreSplitMax * ld * KMAX);
    zc2_32 = dalloc<float>(ld * KMAX);
    p2_32 = dalloc<float>(kCoreSplitMax * ld * KMAX);
  }
  HIPCHK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  overlap = true;
}

// Trailing levels of the lead pair as one launch per direction: the longest
// suffix of its level chain (the dense-core level excluded) whose levels have
// at most kFlowRows rows and no rows split into chunks, in both directions,
// when it has at least kFlowMin levels.  WITHIN_AMD_LEVEL_FLOW=0 keeps one
// launch per level (bitwise the same results).
// Switched off: both GPU runs of the bitwise test stalled in the first level
// run (every wait for a previous level's count went unanswered, with the
// counters first at agent, then at system scope); the cause is not found, so
// no solver takes this path.
constexpr bool kLevelRuns = false;
void wa_solver::setup_flow() {
  constexpr uint32_t kFlowRows = 4096;
  constexpr int kFlowMin = 4;
  const char* e = std::getenv("WITHIN_AMD_LEVEL_FLOW");
  if (!kLevelRuns || !overlap || (e && std::atoi(e) == 0)) return;
  PairSched& ps = pairs[lead];
  const std::vector<Segment>& sf = ps.s_fwd.segs;
  const std::vector<Segment>& sb = ps.s_bwd.segs;
  if (sf.size() != sb.size() || sf.size() < 2) return;
  const int t1 = (int)sf.size() - 1;
  auto ok = [&](const Segment& g) { return g.r1 - g.r0 <= kFlowRows && g.f1 == g.f0; };
  int t0 = t1;
  while (t0 > 0 && ok(sf[t0 - 1]) && ok(sb[t0 - 1])) --t0;
  if (t1 - t0 < kFlowMin) return;
  std::vector<FlowLevel> hf, hb;
  for (int t = t0; t < t1; ++t) hf.push_back({sf[t].r0, sf[t].r1, sf[t].i0, sf[t].i1});
  for (int t = t1 - 1; t >= t0; --t) hb.push_back({sb[t].r0, sb[t].r1, sb[t].i0, sb[t].i1});
  ps.flow_fwd = upload(hf);
  ps.flow_bwd = upload(hb);
  ps.flow_t0 = t0;
  ps.flow_t1 = t1;
  flow_nl = t1 - t0;
  flow_ctr = dalloc<uint32_t>(2 * (2 + flow_nl));
  HIPCHK(hipMemset(flow_ctr, 0, sizeof(uint32_t) * 2 * (2 + flow_nl)));
}

template <typename T>
FlowCall wa_solver::flow_call(int kp) {
  FlowCall fc;
  if (!flow_ctr) return fc;
  fc.ctr_f = flow_ctr;
  fc.ctr_b = flow_ctr + 2 + flow_nl;
  fc.spin = 5000000u;  // 50 ms of the 100 MHz real-time counter per wait
  fc.prev_f = flow_prev_f;
  fc.prev_b = flow_prev_b;
  fc.check = flow_used ? 1 : 0;
  flow_totals<T>(pairs[lead], kp, flow_prev_f, flow_prev_b);
  flow_used = true;
  return fc;
}

// After a synchronised solve or apply: the level runs finished without a
// timed-out wait and handed out every unit; otherwise the results are not
// trusted, the error is reported and the solver goes back to one launch per
// level.
void wa_solver::check_flow() {
  if (!flow_ctr || !flow_used) return;
  uint32_t* h = reinterpret_cast<uint32_t*>(hbuf);
  HIPCHK(hipMemcpyAsync(h, flow_ctr, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h + 2, flow_ctr + 2 + flow_nl, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const uint32_t tick_f = h[0], err = h[1] | h[3], tick_b = h[2];
  if (err == 0 && tick_f >= flow_prev_f && tick_b >= flow_prev_b) return;
  pairs[lead].flow_t0 = pairs[lead].flow_t1 = -1;
  flow_used = false;
  char msg[160];
  std::snprintf(msg, sizeof msg, "preconditioner level run failed (error bits %u, units %u/%u and %u/%u)", err,
                tick_f, flow_prev_f, tick_b, flow_prev_b);
  throw Error(WA_ERR_DEVICE, msg);
}

template <typename T>
static const DevPass<T>& pass_of(wa_solver* s, int q);
template <> const DevPass<double>& pass_of<double>(wa_solver* s, int q) { return s->p64[q]; }
template <> const DevPass<float>& pass_of<float>(wa_solver* s, int q) { return s->p32[q]; }

template <typename T>
void wa_solver::op_apply(const T* X, T* Y, int kp) {
  hipEvent_t a = nullptr;
  if (params.operator_repr == 1 && has_gram) {
    tic(4, a);
    if constexpr (std::is_same<T, double>::value) launch_spmm<double>(st, g64, X, Y, part64, kp);
    else launch_spmm<float>(st, g32, X, Y, (float*)part64, kp);
    toc(4, a);
    return;
  }
  tic(0, a);
  for (int q = 0; q < Q; ++q) launch_pull<T>(st, pass_of<T>(this, q), X, Y, (T*)part64, kp);
  toc(0, a);
}

template <typename T>
void wa_solver::pre_apply(const T* R, T* Z, int kp, const SideWork* side) {
  hipEvent_t a = nullptr;
  tic(1, a);
  if (params.preconditioner == 2) {
    if constexpr (std::is_same<T, double>::value)
      launch_precond_mult<double>(st, pre64, pairs, g64, R, Z, w64, mw64, mw64 + m * KMAX, mw64 + 2 * m * KMAX,
                                  ones, kp, m);
    else
      launch_precond_mult<float>(st, pre32, pairs, g32, R, Z, w32, mw32, mw32 + m * KMAX, mw32 + 2 * m * KMAX,
                                 ones, kp, m);
  } else if (overlap) {
    const FlowCall fc = flow_call<T>(kp);
    const FlowCall* fl = fc.ctr_f && pairs[lead].flow_t0 >= 0 ? &fc : nullptr;
    if constexpr (std::is_same<T, double>::value)
      launch_precond_overlap<double>(st, st2, ev_fork, ev_join, pre64, pairs, lead, R, Z, w64, part2, zc2_64, p2_64,
                                     kp, m, side, fl);
    else
      launch_precond_overlap<float>(st, st2, ev_fork, ev_join, pre32, pairs, lead, R, Z, w32, (float*)part2, zc2_32,
                                    p2_32, kp, m, side, fl);
    side = nullptr;  // done on st2
  } else if constexpr (std::is_same<T, double>::value) {
    launch_precond<double>(st, pre64, R, Z, w64, kp, m);
  } else {
    launch_precond<float>(st, pre32, R, Z, w32, kp, m);
  }
  if (side) side->run(st, side->ctx);
  toc(1, a);
}

// z = M^-1 r.  No preconditioner: z = r.  A Schwarz preconditioner with no
// subdomains (one factor: no factor pairs) sums over nothing, z = 0, as the
// reference's additive and multiplicative applies do
// (schwarz/additive/preconditioner.rs:164-214, multiplicative.rs:128-245).
template <typename T>
void wa_solver::precondition(const T* R, T* Z, int kp, const SideWork* side) {
  if (fcg) {
    coarse_apply<T>(R, Z, kp, side);
    return;
  }
  const int64_t bytes = sizeof(T) * m * kp;
  if (params.preconditioner == 0) HIPCHK(hipMemcpyAsync(Z, R, bytes, hipMemcpyDeviceToDevice, st));
  else if (has_pre) {
    pre_apply<T>(R, Z, kp, side);
    side = nullptr;
  } else HIPCHK(hipMemsetAsync(Z, 0, bytes, st));
  if (side) side->run(st, side->ctx);
}

void wa_solver::rhs64(const double* V, double* out, int kp) {
  hipEvent_t a = nullptr;
  tic(3, a);
  for (int q = 0; q < Q; ++q) launch_scatter_obs(st, p64[q], wobs, V, out, part64, kp);
  toc(3, a);
}

std::vector<double> wa_solver::col_norms(const double* V, int64_t rows, int kp) {
  launch_dot<double>(st, V, V, rows, kp, redp, sc);
  HIPCHK(hipMemcpyAsync(hbuf, sc, sizeof(double) * kp, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<double> r(kp);
  for (int c = 0; c < kp; ++c) r[c] = std::sqrt(hbuf[c]);
  return r;
}

// CG iterations are captured into a HIP graph and replayed, unless
// WITHIN_AMD_GRAPHS=0, kernel timing is on (its event records), or an apply
// would record a memset (a Schwarz preconditioner without subdomains).
bool wa_solver::use_graphs() const {
  const char* e = std::getenv("WITHIN_AMD_GRAPHS");
  if (e && std::atoi(e) == 0) return false;
  if (timer.on) return false;
  if (params.preconditioner == 2) return false;
  if (params.preconditioner == 1 && !has_pre) return false;
  return true;
}

// Most CG iterations replayed per host check (WITHIN_AMD_GRAPH_CHUNK
// overrides): 8 when a block vector holds at most 4M values (launch-bound
// iterations: the 100k-observation 2-factor problem went from 1.64 to 1.34 ms
// per solve), else 2 (20M observations, k = 32: 90.5 ms eager, 90.1 ms with 1
// or 2 iterations per check, 96.1 / 92.1 ms with up to 4 / 8).
static int graph_chunk_max(int64_t values) {
  const char* e = std::getenv("WITHIN_AMD_GRAPH_CHUNK");
  const int v = e ? std::atoi(e) : 0;
  if (v >= 1 && v <= 64) return v;
  return values <= ((int64_t)4 << 20) ? 8 : 2;
}

// Iterations to replay before the next host check: as many as the slowest
// active column still needs at its last per-iteration residual reduction
// (with a margin), capped by graph_chunk_max.  Columns that finish inside a
// chunk are frozen by their masks, so results do not depend on the chunking.
static int replay_chunk(const std::vector<ColState>& hs, const std::vector<double>& rate, int64_t values) {
  double need = 1.0;
  for (size_t c = 0; c < hs.size(); ++c) {
    if (!hs[c].active || !(rate[c] > 0.0 && rate[c] < 1.0) || !(hs[c].rnorm > 0.0)) continue;
    const double goal = hs[c].tol * hs[c].bnorm / hs[c].rnorm;
    if (goal < 1.0) need = std::max(need, std::log(goal) / std::log(rate[c]));
  }
  return (int)std::max(1.0, std::min((double)graph_chunk_max(values), std::floor(0.7 * need - 0.5)));
}

template <typename T, typename F>
hipGraphExec_t wa_solver::cg_graph(int kp, T* x, bool side, F&& body, int kind) {
  const GraphKey key{kind, (int)sizeof(T), kp, params.operator_repr, side ? 1 : 0, x};
  for (auto& g : graphs)
    if (g.first.kind == key.kind && g.first.t == key.t && g.first.kp == key.kp && g.first.op == key.op &&
        g.first.side == key.side && g.first.x == key.x)
      return g.second;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(st, &g));
  hipGraphExec_t ge = nullptr;
  const hipError_t rc = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(rc);
  graphs.push_back({key, ge});
  return ge;
}

// Block CG: k independent recurrences sharing every operator and
// preconditioner apply (solve/cg.rs:45-133 per column).
template <typename T>
void wa_solver::cg(int k, int kp, const T* rhs, T* x, const std::vector<double>& tol, std::vector<int>& conv,
                   std::vector<int64_t>& its) {
  T *r = vec<T>(1), *z = vec<T>(2), *p = vec<T>(3), *ap = vec<T>(4);
  const int64_t bytes = sizeof(T) * m * kp;
  double *bn2 = s_(0), *rz = s_(1), *pap = s_(2), *alpha = s_(3), *rr = s_(4), *rzn = s_(5), *beta = s_(6),
         *tolv = s_(7);
  // with the learned coarse correction the preconditioner is the balancing
  // two-level one (coarse_apply), stored in BF16, so CG runs in its flexible
  // form (beta = -<z, Ap> / <p, Ap>)
  const bool flex = cn.enabled && coarse_supported();
  struct FcgFlag {
    bool& f;
    FcgFlag(bool& x, bool v) : f(x) { f = v; }
    ~FcgFlag() { f = false; }
  } fcg_flag(fcg, flex);
  int* mask = imask;
  int* nact = imask + KMAX;
  std::vector<double> tp(KMAX, 0.0);
  for (int c = 0; c < k; ++c) tp[c] = tol[c];
  HIPCHK(hipMemcpyAsync(tolv, tp.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(x, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(r, rhs, bytes, hipMemcpyDeviceToDevice, st));
  hipEvent_t ev = nullptr;
  precondition<T>(r, z, kp);
  HIPCHK(hipMemcpyAsync(p, z, bytes, hipMemcpyDeviceToDevice, st));
  tic(2, ev);
  launch_dot<T>(st, r, z, m, kp, redp, rz);
  launch_dot<T>(st, rhs, rhs, m, kp, redp, bn2);
  launch_cg_start(st, cs, bn2, rz, tolv, k, kp);
  toc(2, ev);
  // with the two-stream additive apply, x += alpha p (not needed until the
  // end, p unchanged until the next direction update) runs on the second
  // stream beside the level chain instead of in the residual update
  const bool side_x = overlap && params.preconditioner == 1 && has_pre;
  struct XUpd {
    T* x;
    const T* p;
    const double* alpha;
    int64_t m;
    int kp;
  } xu{x, p, alpha, m, kp};
  const SideWork xside{[](hipStream_t s2, void* c) {
                         const XUpd* u = static_cast<const XUpd*>(c);
                         launch_cg_x<T>(s2, u->x, u->p, u->alpha, u->m, u->kp);
                       },
                       &xu};
  // one iteration (cg.rs:75-125 for every column): the same launches whether
  // enqueued directly or replayed from a captured graph
  auto iteration = [&]() {
    op_apply<T>(p, ap, kp);
    tic(2, ev);
    int nb = launch_dot_part<T>(st, p, ap, m, kp, redp);
    launch_cg_alpha(st, cs, redp, nb, pap, alpha, kp);
    nb = launch_cg_update<T>(st, side_x ? nullptr : x, p, r, ap, alpha, m, kp, redp);
    launch_cg_check(st, cs, redp, nb, rr, kp);
    toc(2, ev);
    precondition<T>(r, z, kp, side_x ? &xside : nullptr);
    tic(2, ev);
    if (flex) {
      nb = launch_dot2_part<T>(st, z, r, ap, m, kp, cn.dpart);
      launch_fcg_beta(st, cs, cn.dpart, nb, pap, rzn, beta, mask, nact, kp);
    } else {
      nb = launch_dot_part<T>(st, r, z, m, kp, redp);
      launch_cg_beta(st, cs, redp, nb, rzn, beta, mask, nact, kp);
    }
    launch_xpby<T>(st, p, z, beta, mask, m, kp);
    toc(2, ev);
  };
  hipGraphExec_t graph = use_graphs() ? cg_graph<T>(kp, x, side_x, iteration, flex ? 2 : 0) : nullptr;
  std::vector<ColState> hs(kp);
  auto read_state = [&]() {
    HIPCHK(hipMemcpyAsync(hs.data(), cs, sizeof(ColState) * kp, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int a = 0;
    for (int c = 0; c < kp; ++c) a += hs[c].active;
    return a;
  };
  // all columns may already be done (zero right-hand sides)
  bool any = read_state() > 0;
  int64_t itn = 0;
  std::vector<double> rate(kp, 0.0), before(kp, 0.0);
  while (any && itn < params.maxiter) {
    // with a graph, several iterations are replayed per host check
    int chunk = graph ? replay_chunk(hs, rate, m * kp) : 1;
    chunk = (int)std::min<int64_t>(chunk, params.maxiter - itn);
    for (int c = 0; c < kp; ++c) before[c] = hs[c].rnorm;
    for (int i = 0; i < chunk; ++i) {
      if (graph) HIPCHK(hipGraphLaunch(graph, st));
      else iteration();
    }
    itn += chunk;
    any = read_state() > 0;
    for (int c = 0; c < kp; ++c)
      if (before[c] > 0.0 && hs[c].rnorm > 0.0) rate[c] = std::pow(hs[c].rnorm / before[c], 1.0 / chunk);
  }
  std::vector<ColState> h(kp);
  HIPCHK(hipMemcpyAsync(h.data(), cs, sizeof(ColState) * kp, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int c = 0; c < k; ++c) {
    if (h[c].active || h[c].nan_stop) {  // maxiter reached, or a NaN residual (which would run it out)
      conv[c] = 0;
      its[c] = params.maxiter;
    } else {
      conv[c] = h[c].converged;
      its[c] = h[c].iters;
    }
  }
}

wa_solver::GmDev& wa_solver::gm_dev() {
  GmDev& d = gmd;
  if (d.cols) return d;
  d.cols = dalloc<GmCol>(KMAX);
  d.H = dalloc<double>((int64_t)KMAX * kGmresMaxV * (kGmresMaxV + 1));
  d.h = dalloc<double>((int64_t)kGmresMaxV * KMAX);
  d.h2 = dalloc<double>((int64_t)kGmresMaxV * KMAX);
  d.y = dalloc<double>((int64_t)kGmresMaxV * KMAX);
  d.coef = dalloc<double>(KMAX);
  d.beta = dalloc<double>(KMAX);
  d.abs_tol = dalloc<double>(KMAX);
  d.part = dalloc<double>((int64_t)kGmRedBlocks * kGmresMaxV * KMAX);
  d.part2 = dalloc<double>((int64_t)kGmRedBlocks * KMAX);
  d.live = dalloc<int>(KMAX);
  d.lim = dalloc<int>(KMAX);
  d.ninc = dalloc<int>(kGmresMaxV);
  d.ev.resize(kGmresMaxV);
  for (hipEvent_t& e : d.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return d;
}

// Right-preconditioned GMRES(restart) for a block of columns, one Arnoldi
// process per column (solve/gmres.rs:201-530), with the Arnoldi step on the
// device: two passes of block Gram-Schmidt (all dot products of a pass in one
// launch), the FP64 Hessenberg / Givens update and the stop decision per
// column in one kernel.  The host enqueues step j + 1 once step j - 1 has
// finished (its count of extending columns, copied to pinned memory), so the
// device is never idle waiting for the host; at most one step after the last
// column stops is wasted.  The true residual at each restart is FP64-reduced
// on the device as before.
template <typename T>
void wa_solver::gmres_dev(int k, int kp, const T* b, T* x, const std::vector<double>& tol, std::vector<int>& conv,
                          std::vector<int64_t>& its) {
  const int mr = params.restart;
  const int64_t vb = m * kp, vs = m * KMAX;
  GmresBufs<T>& G = gm<T>(mr);
  GmDev& D = gm_dev();
  T *w = G.w, *r = G.r;
  auto V = [&](int i) { return G.V + (int64_t)i * vs; };
  auto Z = [&](int i) { return G.Z + (int64_t)i * vs; };
  HIPCHK(hipMemsetAsync(x, 0, sizeof(T) * vb, st));
  auto norms = [&](const T* v) {
    launch_dot<T>(st, v, v, m, kp, redp, s_(10));
    HIPCHK(hipMemcpyAsync(hbuf, s_(10), sizeof(double) * kp, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> o(kp);
    for (int c = 0; c < kp; ++c) o[c] = std::sqrt(hbuf[c]);
    return o;
  };
  std::vector<double> bn = norms(b), abs_tol(KMAX, 0.0);
  std::vector<int> live(KMAX, 0);
  for (int c = 0; c < k; ++c) {
    abs_tol[c] = tol[c] * bn[c];
    if (bn[c] < F64_EPS) { conv[c] = 1; its[c] = 0; } else live[c] = 1;
  }
  HIPCHK(hipMemcpyAsync(D.abs_tol, abs_tol.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
  auto any_live = [&]() { for (int c = 0; c < kp; ++c) if (live[c]) return true; return false; };
  int* ninc_h = reinterpret_cast<int*>(hbuf) + 2 * KMAX;  // pinned slots, one per step
  std::vector<GmCol> hc(kp);
  int64_t total = 0;
  hipEvent_t et = nullptr;
  while (any_live()) {
    // true residual r = b - A x (restart), gmres.rs:324-335
    op_apply<T>(x, r, kp);
    launch_rsub<T>(st, r, b, vb);
    std::vector<double> beta = norms(r);
    for (int c = 0; c < kp; ++c) {
      if (!live[c]) continue;
      if (beta[c] <= abs_tol[c]) { live[c] = 0; conv[c] = 1; its[c] = total; }
      else if (total >= params.maxiter) { live[c] = 0; conv[c] = 0; its[c] = total; }
    }
    if (!any_live()) break;
    HIPCHK(hipMemcpyAsync(D.live, live.data(), sizeof(int) * KMAX, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(D.beta, beta.data(), sizeof(double) * kp, hipMemcpyHostToDevice, st));
    launch_gm_cycle_start(st, D.cols, D.H, D.beta, D.live, D.coef, mr, kp);
    launch_scale_col<T>(st, V(0), r, D.coef, m, kp);
    const int64_t t0 = total;
    const int64_t cycle = std::min<int64_t>(params.maxiter - total, mr);
    int j = 0;
    while (j < cycle) {
      if (j >= 2) {  // the count of step j - 2 (step j - 1 is running): stop when no column extends
        HIPCHK(hipEventSynchronize(D.ev[j - 2]));
        if (ninc_h[j - 2] == 0) break;
      }
      precondition<T>(V(j), Z(j), kp);
      op_apply<T>(Z(j), w, kp);
      tic(2, et);
      launch_multi_dot<T>(st, w, V(0), vs, j + 1, m, kp, D.part, D.h, 0);
      launch_multi_axpy<T>(st, w, V(0), vs, j + 1, D.h, nullptr, -1.0, m, kp, nullptr);
      launch_multi_dot<T>(st, w, V(0), vs, j + 1, m, kp, D.part, D.h2, 0);
      const int nb = launch_multi_axpy<T>(st, w, V(0), vs, j + 1, D.h2, nullptr, -1.0, m, kp, D.part2);
      ++total;
      launch_gm_givens(st, D.cols, D.H, D.h, D.h2, D.part2, nb, D.abs_tol, D.coef, D.ninc + j, j, mr, total,
                       params.maxiter, kp);
      launch_scale_col<T>(st, V(j + 1), w, D.coef, m, kp);
      toc(2, et);
      HIPCHK(hipMemcpyAsync(ninc_h + j, D.ninc + j, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(D.ev[j], st));
      ++j;
    }
    total = t0 + j;
    // y = R^-1 g per column (gmres.rs:499-519), x += Z y (:523-530)
    launch_gm_backsolve(st, D.cols, D.H, D.live, D.y, D.lim, mr, kp);
    launch_multi_axpy<T>(st, x, Z(0), vs, j, D.y, D.lim, 1.0, m, kp, nullptr);
    HIPCHK(hipMemcpyAsync(hc.data(), D.cols, sizeof(GmCol) * kp, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool need_true = false;
    for (int c = 0; c < kp; ++c) {
      if (!live[c]) continue;
      if (hc[c].outcome == 0) { live[c] = 0; conv[c] = 1; its[c] = t0 + hc[c].jc; }
      else if (hc[c].outcome == 1) { live[c] = 0; conv[c] = 0; its[c] = t0 + hc[c].jc; }
      else if (hc[c].outcome == 2) need_true = true;
    }
    if (need_true) {  // lucky breakdown: decide on the true residual
      op_apply<T>(x, r, kp);
      launch_rsub<T>(st, r, b, vb);
      std::vector<double> tr = norms(r);
      for (int c = 0; c < kp; ++c)
        if (live[c] && hc[c].outcome == 2) {
          live[c] = 0;
          conv[c] = tr[c] <= abs_tol[c];
          its[c] = t0 + hc[c].jc;
        }
    }
  }
}

// Right-preconditioned GMRES(restart) with modified Gram-Schmidt and Givens
// rotations, one Arnoldi process per column (solve/gmres.rs:201-530).  All
// active columns advance in lockstep (they share the iteration count); the
// Hessenberg / Givens recurrences and the triangular solve stay in FP64 on
// the host, the basis and every vector operation on the device.
template <typename T>
void wa_solver::gmres(int k, int kp, const T* b, T* x, const std::vector<double>& tol, std::vector<int>& conv,
                      std::vector<int64_t>& its) {
  const int mr = params.restart;
  const int64_t vb = m * kp;  // elements per block vector
  GmresBufs<T>& G = gm<T>(mr);
  T *w = G.w, *r = G.r;
  auto V = [&](int i) { return G.V + (int64_t)i * m * KMAX; };
  auto Z = [&](int i) { return G.Z + (int64_t)i * m * KMAX; };
  double* hd = G.hd;  // (mr + 2) x KMAX dots of the current column
  HIPCHK(hipMemsetAsync(x, 0, sizeof(T) * vb, st));
  auto norms = [&](const T* v) {
    launch_dot<T>(st, v, v, m, kp, redp, s_(10));
    HIPCHK(hipMemcpyAsync(hbuf, s_(10), sizeof(double) * kp, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<double> o(kp);
    for (int c = 0; c < kp; ++c) o[c] = std::sqrt(hbuf[c]);
    return o;
  };
  auto upload_coef = [&](const std::vector<double>& v, double* dst) {
    HIPCHK(hipMemcpyAsync(dst, v.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
  };
  std::vector<double> bn = norms(b), abs_tol(kp, 0.0);
  std::vector<char> live(kp, 0);
  for (int c = 0; c < kp; ++c) {
    if (c >= k) continue;
    abs_tol[c] = tol[c] * bn[c];
    if (bn[c] < F64_EPS) { conv[c] = 1; its[c] = 0; } else live[c] = 1;
  }
  auto any_live = [&]() { for (int c = 0; c < kp; ++c) if (live[c]) return true; return false; };
  int64_t total = 0;
  std::vector<double> H((size_t)(mr + 1) * mr * KMAX), cs(mr * KMAX), sn(mr * KMAX), g((mr + 1) * KMAX);
  auto h = [&](int c, int i, int j) -> double& { return H[((size_t)c * mr + j) * (mr + 1) + i]; };
  while (any_live()) {
    // true residual r = b - A x (restart), gmres.rs:324-335
    op_apply<T>(x, r, kp);
    launch_rsub<T>(st, r, b, vb);
    std::vector<double> beta = norms(r);
    for (int c = 0; c < kp; ++c) {
      if (!live[c]) continue;
      if (beta[c] <= abs_tol[c]) { live[c] = 0; conv[c] = 1; its[c] = total; }
      else if (total >= params.maxiter) { live[c] = 0; conv[c] = 0; its[c] = total; }
    }
    if (!any_live()) break;
    std::vector<double> coef(KMAX, 0.0);
    for (int c = 0; c < kp; ++c) coef[c] = live[c] ? 1.0 / beta[c] : 0.0;
    upload_coef(coef, s_(11));
    launch_scale_col<T>(st, V(0), r, s_(11), m, kp);
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(cs.begin(), cs.end(), 0.0);
    std::fill(sn.begin(), sn.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    std::vector<double> hmax(kp, 0.0);
    for (int c = 0; c < kp; ++c) g[(size_t)c * (mr + 1)] = live[c] ? beta[c] : 0.0;
    const int64_t cycle = std::min<int64_t>(params.maxiter - total, mr);
    std::vector<char> inc = live;           // still extending its basis this cycle
    std::vector<int> jc(kp, 0), outcome(kp, 3);  // 0 converged, 1 maxiter, 2 breakdown, 3 restart
    std::vector<double> resid(kp, 0.0);
    int j = 0;
    auto any_inc = [&]() { for (int c = 0; c < kp; ++c) if (inc[c]) return true; return false; };
    while (j < cycle && any_inc()) {
      precondition<T>(V(j), Z(j), kp);
      op_apply<T>(Z(j), w, kp);
      hipEvent_t ev = nullptr;
      tic(2, ev);
      for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt, gmres.rs:201-220
        launch_dot<T>(st, w, V(i), m, kp, redp, hd + (int64_t)i * KMAX);
        launch_axpy_col<T>(st, w, V(i), hd + (int64_t)i * KMAX, -1.0, m, kp);
      }
      launch_dot<T>(st, w, w, m, kp, redp, hd + (int64_t)(j + 1) * KMAX);
      toc(2, ev);
      std::vector<double> hh((size_t)(j + 2) * KMAX);
      HIPCHK(hipMemcpyAsync(hh.data(), hd, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      ++total;
      std::fill(coef.begin(), coef.end(), 0.0);
      for (int c = 0; c < kp; ++c) {
        if (!inc[c]) continue;
        double lmax = 0.0;
        for (int i = 0; i <= j; ++i) {
          h(c, i, j) = hh[(size_t)i * KMAX + c];
          lmax = std::max(lmax, std::fabs(h(c, i, j)));
        }
        const double hj1 = std::sqrt(hh[(size_t)(j + 1) * KMAX + c]);
        h(c, j + 1, j) = hj1;
        hmax[c] = std::max(std::max(hmax[c], lmax), hj1);
        double* csc = &cs[(size_t)c * mr];
        double* snc = &sn[(size_t)c * mr];
        double* gc = &g[(size_t)c * (mr + 1)];
        for (int i = 0; i < j; ++i) {  // previous rotations, gmres.rs:223-231
          const double a = h(c, i, j), bb = h(c, i + 1, j);
          const double t = csc[i] * a + snc[i] * bb;
          h(c, i + 1, j) = -snc[i] * a + csc[i] * bb;
          h(c, i, j) = t;
        }
        const double a = h(c, j, j), bb = h(c, j + 1, j), rr = std::hypot(a, bb);
        const double cc = rr == 0.0 ? 1.0 : a / rr, ss = rr == 0.0 ? 0.0 : bb / rr;
        csc[j] = cc;
        snc[j] = ss;
        h(c, j, j) = cc * a + ss * bb;
        h(c, j + 1, j) = 0.0;
        const double t = cc * gc[j] + ss * gc[j + 1];
        gc[j + 1] = -ss * gc[j] + cc * gc[j + 1];
        gc[j] = t;
        jc[c] = j + 1;
        resid[c] = std::fabs(gc[j + 1]);
        if (resid[c] <= abs_tol[c]) { outcome[c] = 0; inc[c] = 0; }
        else if (total >= params.maxiter) { outcome[c] = 1; inc[c] = 0; }
        else if (hj1 > 1e-14 * hmax[c]) coef[c] = 1.0 / hj1;
        else { outcome[c] = 2; inc[c] = 0; }
      }
      ++j;
      if (any_inc() && j < cycle) {
        upload_coef(coef, s_(11));
        launch_scale_col<T>(st, V(j), w, s_(11), m, kp);
      }
    }
    // y = R^-1 g per column (gmres.rs:499-519), x += Z y (:523-530)
    int jmax = 0;
    std::vector<std::vector<double>> y(kp);
    for (int c = 0; c < kp; ++c) {
      if (!live[c]) continue;
      jmax = std::max(jmax, jc[c]);
      y[c].assign(jc[c], 0.0);
      const double* gc = &g[(size_t)c * (mr + 1)];
      for (int i = jc[c] - 1; i >= 0; --i) {
        double sum = gc[i];
        for (int q = i + 1; q < jc[c]; ++q) sum -= h(c, i, q) * y[c][q];
        const double dg = h(c, i, i);
        y[c][i] = std::fabs(dg) < 1e-14 * (1.0 + std::fabs(sum)) ? 0.0 : sum / dg;
      }
    }
    for (int i = 0; i < jmax; ++i) {
      std::fill(coef.begin(), coef.end(), 0.0);
      for (int c = 0; c < kp; ++c)
        if (live[c] && i < jc[c]) coef[c] = y[c][i];
      upload_coef(coef, s_(11));
      launch_axpy_col<T>(st, x, Z(i), s_(11), 1.0, m, kp);
    }
    bool need_true = false;
    for (int c = 0; c < kp; ++c) {
      if (!live[c]) continue;
      if (outcome[c] == 0) { live[c] = 0; conv[c] = 1; its[c] = total - (j - jc[c]); }
      else if (outcome[c] == 1) { live[c] = 0; conv[c] = 0; its[c] = total - (j - jc[c]); }
      else if (outcome[c] == 2) need_true = true;
    }
    if (need_true) {  // lucky breakdown: decide on the true residual
      op_apply<T>(x, r, kp);
      launch_rsub<T>(st, r, b, vb);
      std::vector<double> tr = norms(r);
      for (int c = 0; c < kp; ++c)
        if (live[c] && outcome[c] == 2) {
          live[c] = 0;
          conv[c] = tr[c] <= abs_tol[c];
          its[c] = total - (j - jc[c]);
        }
    }
  }
}

// One block of up to 32 right-hand sides (Solver::solve per column,
// solver.rs:158-239 + iterative_refinement :296-347).
// LSMR over W^(1/2) D with the Schwarz preconditioner as M ~ D'WD
// (solver.rs:162-207, lsmr.rs:122-316): the modified Golub-Kahan process for
// every column of the block at once, FP64 throughout.  Without a
// preconditioner M = I, which gives the plain Golub-Kahan recurrences in exact
// arithmetic (lsmr/tests.rs:270-309).  Reads Y from Yr; leaves x in X64.
void wa_solver::lsmr(int k, int kp, std::vector<int>& conv, std::vector<int64_t>& its) {
  const Design& d = plan.d;
  if (!std::isfinite(params.tol) || params.tol < 0.0)  // lsmr.rs:262-266
    throw Error(WA_ERR_ITERATIVE_SOLVE, "lsmr: tolerance must be finite and nonnegative");
  const int64_t vbytes = sizeof(double) * m * kp;
  if (!lpart) {
    lpart = dalloc<double>((int64_t)kLsmrBlocks * KMAX);
    lcoef = dalloc<double>(KMAX);
    lcs = dalloc<LsmrCol>(KMAX);
    if (!d.weights.empty()) {
      std::vector<double> sw(n);
      for (int64_t i = 0; i < n; ++i) sw[i] = std::sqrt(d.weights[i]);
      swobs = upload(sw);
    }
  }
  const int64_t cap = std::min<int64_t>(params.local_size, std::min<int64_t>(n, mu));
  if (cap > ring_cap) {
    ring_v = dalloc<double>(cap * m * KMAX);
    ring_p = dalloc<double>(cap * m * KMAX);
    ring_cap = cap;
  }
  double *X = X64, *V = cx64, *P = cr64, *H = cz64, *HB = cp64, *ATU = cap64, *U = E;
  int* nact = imask + KMAX;
  hipEvent_t ev = nullptr;
  auto adjoint = [&]() {  // ATU = D' (W^(1/2) U)
    tic(3, ev);
    for (int q = 0; q < Q; ++q) launch_scatter_obs(st, p64[q], swobs, U, ATU, part64, kp);
    toc(3, ev);
  };
  auto solve_m = [&]() { precondition<double>(P, V, kp); };  // V = M^-1 P
  int64_t count = 0, next = 0;
  auto push = [&]() {
    if (cap == 0) return;
    launch_lsmr_push(st, V, P, ring_v + next * m * KMAX, ring_p + next * m * KMAX, lcs, m, kp);
    next = (next + 1) % cap;
    if (count < cap) ++count;
  };
  // u = W^(1/2) y (unnormalised; 1/beta_1 is carried as beta_prev_inv),
  // p~ = A'u / beta_1, v = M^-1 p~, alpha_1 = sqrt(<v, p~>)
  tic(3, ev);
  int nb = launch_lsmr_rhs(st, n, swobs, Yr, U, kp, lpart);
  toc(3, ev);
  tic(2, ev);
  launch_lsmr_start(st, lcs, lpart, nb, params.tol, k, kp);
  HIPCHK(hipMemsetAsync(P, 0, vbytes, st));
  toc(2, ev);
  adjoint();
  tic(2, ev);
  launch_lsmr_ptilde(st, P, ATU, lcs, m, kp);
  toc(2, ev);
  solve_m();
  tic(2, ev);
  nb = launch_dot_part<double>(st, V, P, m, kp, redp);
  launch_lsmr_alpha1(st, lcs, redp, nb, kp, nact);
  launch_lsmr_init_vec(st, V, H, HB, X, lcs, m, kp);
  toc(2, ev);
  push();
  std::vector<LsmrCol> h(kp);
  HIPCHK(hipMemcpyAsync(h.data(), lcs, sizeof(LsmrCol) * kp, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  bool any = false;
  for (int c = 0; c < k; ++c) {
    if (h[c].bad) throw Error(WA_ERR_ITERATIVE_SOLVE, "lsmr: right-hand side entries must be finite");
    any = any || h[c].active;
  }
  int64_t itn = 0;
  while (any && itn < params.maxiter) {
    ++itn;
    tic(3, ev);
    nb = launch_lsmr_fwd(st, n, Q, codes, swobs, V, U, lcs, kp, lpart);
    toc(3, ev);
    tic(2, ev);
    launch_lsmr_beta(st, lcs, lpart, nb, kp);
    toc(2, ev);
    adjoint();
    tic(2, ev);
    launch_lsmr_ptilde(st, P, ATU, lcs, m, kp);
    toc(2, ev);
    solve_m();
    tic(2, ev);
    const int64_t start = count < cap ? 0 : next;
    for (int64_t i = 0; i < count; ++i) {  // oldest first (bidiag.rs:206-217)
      const int64_t j = (start + i) % cap;
      launch_dot<double>(st, V, ring_p + j * m * KMAX, m, kp, redp, lcoef);
      launch_lsmr_reorth(st, V, P, ring_v + j * m * KMAX, ring_p + j * m * KMAX, lcoef, lcs, m, kp);
    }
    nb = launch_dot_part<double>(st, V, P, m, kp, redp);
    launch_lsmr_step(st, lcs, redp, nb, (int)itn, kp, nact);
    launch_lsmr_update(st, V, H, HB, X, lcs, m, kp);
    toc(2, ev);
    push();
    HIPCHK(hipMemcpyAsync(hbuf, nact, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    any = *(int*)hbuf > 0;
  }
  HIPCHK(hipMemcpyAsync(h.data(), lcs, sizeof(LsmrCol) * kp, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int c = 0; c < k; ++c) {
    if (h[c].active) {  // maxiter reached
      conv[c] = 0;
      its[c] = params.maxiter;
    } else {
      conv[c] = h[c].converged;
      its[c] = h[c].iters;
    }
  }
}

// ---------------------------------------------------------------------------
// Learned coarse-space correction (coarse.hpp).
//
// Network over the factor-level graph (the Gramian's sparsity): H_0 = input
// features (m x 32) and, for layer l,
//   Z_l = H_l Ws_l + (A H_l) Wa_l + b_l,   A = (1/Q) D^-1/2 G D^-1/2,
//   H_{l+1} = ReLU(Z_l) (hidden layers),   P = Z_{L-1} (32 coarse vectors).
// A H is the solver's own FP32 operator applied to D^-1/2 H.  The loss, on
// error samples E (m x S) of the plain solver with residuals R = G E, is
//   L(P) = -<U, (P'GP + eps)^-1 U> / nu,   U = P'R,   nu = sum_s e_s' r_s:
// minus the share of the samples' G-energy that the G-orthogonal projection
// onto span(P) removes, i.e. what a coarse correction with P would take out
// of the errors the one-level method leaves.  With V = (P'GP + eps)^-1 U,
//   dL/dP = -(2/nu) (R - GP V) V'.
// eps (the ridge, relative to the mean diagonal of P'GP) bounds the
// amplification of FP32 rounding in GP by ill-conditioned P'GP.
// ---------------------------------------------------------------------------
static uint64_t smix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// N(0, 1) from the counter i (Box-Muller on two 53-bit uniforms)
static double host_randn(uint64_t seed, uint64_t i) {
  const uint64_t h1 = smix64(seed ^ smix64(2 * i)), h2 = smix64(seed ^ smix64(2 * i + 1));
  const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// Input features of internal DOF i (user DOF u, factor q, Gramian diagonal d):
//   [1, z, one-hot(q mod 4) (4 columns), z^2 - 1, 25 N(0,1) values keyed by u],
// z = standardised log d.  The random columns let the layers form smooth
// functions that differ between regions of the graph.
void wa_solver::coarse_alloc(uint64_t seed) {
  const int64_t mm = m * kCoarse;
  if (!cn.ready) {
    cn.X0 = dalloc<float>(mm);
    cn.dinv = dalloc<float>(m);
    cn.cvec = dalloc<float>(m);
    cn.H[0] = cn.X0;
    for (int l = 1; l < kGnnLayers; ++l) cn.H[l] = dalloc<float>(mm);
    for (int l = 0; l < kGnnLayers; ++l) {
      cn.S[l] = dalloc<float>(mm);
      cn.T[l] = dalloc<float>(mm);
    }
    for (float** b : {&cn.P, &cn.GP, &cn.E, &cn.R, &cn.dP, &cn.dZa, &cn.dZb, &cn.Ys, &cn.Tb}) *b = dalloc<float>(mm);
    for (float** b : {&cn.theta, &cn.grad, &cn.m1, &cn.m2}) *b = dalloc<float>(kGnnParams);
    cn.pack = dalloc<bf16_t>(kGnnLayers * kLayerPack);
    cn.xpart = dalloc<float>((int64_t)kXtyBlocksMax * (kMat + kCoarse));
    cn.W12 = dalloc<double>(2 * kMat);
    cn.xpart64 = dalloc<double>((int64_t)kXtyBlocksMax * kMat);
    for (double** b : {&cn.Ac, &cn.U, &cn.V, &cn.Ainv, &cn.c1, &cn.cd}) *b = dalloc<double>(kMat);
    cn.loss = dalloc<double>(1);
    cn.Pb = dalloc<bf16_t>(mm);
    cn.GPb = dalloc<bf16_t>(mm);
    cn.cpart = dalloc<double>((int64_t)kCprojBlocksMax * kMat);
    cn.cmid = dalloc<double>((int64_t)kCprojMid * kMat);
    cn.r2 = dalloc<double>(mm);
    cn.dpart = dalloc<double>(2 * (int64_t)kRedBlocks * KMAX);
    cn.ready = true;
  }
  cn.seed = seed;
  std::vector<float> dinv((size_t)m), cvec((size_t)m);
  std::vector<double> z((size_t)m);
  double mean = 0.0, var = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    const double d = h_diag[(size_t)i];
    z[(size_t)i] = std::log(std::max(d, 1e-300));
    dinv[(size_t)i] = d > 0.0 ? (float)(1.0 / std::sqrt(d)) : 0.0f;
    cvec[(size_t)i] = d > 0.0 ? (float)(1.0 / std::sqrt(d) / Q) : 0.0f;
    mean += z[(size_t)i];
  }
  mean /= (double)m;
  for (int64_t i = 0; i < m; ++i) var += (z[(size_t)i] - mean) * (z[(size_t)i] - mean);
  const double sd = std::max(std::sqrt(var / (double)m), 1e-12);
  const uint64_t fseed = smix64(seed + 1);
  cn.h_feat.assign((size_t)mm, 0.0f);
  for (int64_t i = 0; i < m; ++i) {
    float* f = cn.h_feat.data() + i * kCoarse;
    int q = 0;
    while (q + 1 < (int)h_foff.size() && i >= h_foff[(size_t)q + 1]) ++q;
    const double zi = (z[(size_t)i] - mean) / sd;
    const uint64_t u = h_ext_of[(size_t)i];
    f[0] = 1.0f;
    f[1] = (float)zi;
    f[2 + (q & 3)] = 1.0f;
    f[6] = (float)(zi * zi - 1.0);
    for (int c = 7; c < kCoarse; ++c) f[c] = (float)host_randn(fseed, u * kCoarse + (uint64_t)c);
  }
  HIPCHK(hipMemcpyAsync(cn.X0, cn.h_feat.data(), sizeof(float) * mm, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cn.dinv, dinv.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cn.cvec, cvec.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
  launch_rowscale(st, m, cn.dinv, cn.X0, cn.S[0]);
  HIPCHK(hipStreamSynchronize(st));
}

// Glorot-uniform weights (fan-in 64: two 32-wide inputs), zero biases, zero
// Adam moments
void wa_solver::coarse_init_params(uint64_t seed) {
  std::vector<float> th((size_t)kGnnParams, 0.0f), zero((size_t)kGnnParams, 0.0f);
  const double a = std::sqrt(6.0 / (64.0 + 32.0));
  const uint64_t ps = smix64(seed + 2);
  for (int l = 0; l < kGnnLayers; ++l)
    for (int i = 0; i < 2 * kMat; ++i) {
      const uint64_t h = smix64(ps ^ smix64((uint64_t)l * kLayerParams + (uint64_t)i));
      const double u = ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      th[(size_t)l * kLayerParams + i] = (float)(a * (2.0 * u - 1.0));
    }
  HIPCHK(hipMemcpyAsync(cn.theta, th.data(), sizeof(float) * kGnnParams, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cn.m1, zero.data(), sizeof(float) * kGnnParams, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(cn.m2, zero.data(), sizeof(float) * kGnnParams, hipMemcpyHostToDevice, st));
  launch_gnn_pack(st, cn.theta, cn.pack);
  HIPCHK(hipStreamSynchronize(st));
}

// Error samples of the plain method: right-hand sides b = G xi with xi ~ N(0,1)
// (so xi solves G x = b up to the null space of G), `iters` iterations of
// FP32 block CG with the one-level preconditioner, then E = xi - x_iters and
// R = G E (32 samples).
void wa_solver::coarse_samples(int iters, uint64_t seed) {
  float* xi = cn.dP;
  launch_randn(st, m * kCoarse, smix64(seed + 3), xi);
  op_apply<float>(xi, cn.Ys, kCoarse);
  const int64_t saved = params.maxiter;
  const bool en = cn.enabled;
  cn.enabled = false;
  params.maxiter = std::max(iters, 1);
  std::vector<double> tol(KMAX, 0.0);
  std::vector<int> conv(KMAX, 0);
  std::vector<int64_t> its(KMAX, 0);
  try {
    cg<float>(KMAX, KMAX, cn.Ys, cx32, tol, conv, its);
  } catch (...) {
    params.maxiter = saved;
    cn.enabled = en;
    throw;
  }
  params.maxiter = saved;
  cn.enabled = en;
  launch_sub(st, m * kCoarse, xi, cx32, cn.E);
  op_apply<float>(cn.E, cn.R, kCoarse);
  launch_dot<float>(st, cn.E, cn.R, m, KMAX, redp, s_(12));
  HIPCHK(hipMemcpyAsync(hbuf, s_(12), sizeof(double) * KMAX, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double nu = 0.0;
  for (int c = 0; c < KMAX; ++c) nu += hbuf[c];
  if (!(nu > 0.0)) throw Error(WA_ERR_ITERATIVE_SOLVE, "coarse training: the error samples carry no G-energy");
  cn.nu = nu;
}

void wa_solver::gnn_forward() {
  for (int l = 0; l < kGnnLayers; ++l) {
    const bf16_t* pk = cn.pack + (size_t)l * kLayerPack;
    const float* b = cn.theta + (size_t)l * kLayerParams + 2 * kMat;
    const bool last = l == kGnnLayers - 1;
    op_apply<float>(cn.S[l], cn.T[l], kCoarse);
    launch_gnn_layer(st, m, cn.H[l], pk, cn.T[l], cn.cvec, pk + kMat, b, nullptr, last ? 0 : 1,
                     last ? cn.P : cn.H[l + 1], last ? nullptr : cn.S[l + 1], cn.dinv);
  }
}

// loss (device scalar cn.loss) and, with grad, its gradient in cn.grad
void wa_solver::gnn_loss(bool grad) {
  op_apply<float>(cn.P, cn.GP, kCoarse);
  launch_xty64(st, m, cn.P, cn.GP, cn.xpart64, cn.cmid, cn.Ac);
  launch_xty64(st, m, cn.P, cn.R, cn.xpart64, cn.cmid, cn.U);
  launch_coarse_small(st, cn.Ac, cn.U, cn.nu, cn.ridge, cn.V, cn.loss, cn.W12, cn.W12 + kMat);
  if (!grad) return;
  launch_tall_small(st, m, cn.R, cn.W12, cn.GP, cn.W12 + kMat, cn.dP);  // dL/dP
  float* dz = cn.dP;
  float* nxt = cn.dZa;
  for (int l = kGnnLayers - 1; l >= 0; --l) {
    float* g = cn.grad + (size_t)l * kLayerParams;
    const bf16_t* pk = cn.pack + (size_t)l * kLayerPack;
    launch_gnn_xty(st, m, cn.H[l], nullptr, dz, cn.xpart, cn.cmid, g, g + 2 * kMat);  // dWs, db
    launch_gnn_xty(st, m, cn.T[l], cn.cvec, dz, cn.xpart, cn.cmid, g + kMat, nullptr);  // dWa (input A H_l)
    if (l == 0) break;
    // dH_l = dZ Ws' + A (dZ Wa') (A is symmetric); dZ_{l-1} = dH_l o [H_l > 0]
    launch_gnn_layer(st, m, dz, pk + 3 * kMat, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, cn.Ys,
                     cn.dinv);
    op_apply<float>(cn.Ys, cn.Tb, kCoarse);
    launch_gnn_layer(st, m, dz, pk + 2 * kMat, cn.Tb, cn.cvec, nullptr, nullptr, cn.H[l], 0, nxt, nullptr, nullptr);
    dz = nxt;
    nxt = nxt == cn.dZa ? cn.dZb : cn.dZa;
  }
}

double wa_solver::gnn_read_loss() {
  HIPCHK(hipMemcpyAsync(hbuf, cn.loss, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return hbuf[0];
}

// Eigen-decomposition of a symmetric n x n matrix by cyclic Jacobi rotations:
// A = V diag(w) V' (columns of V).
static void sym_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  double norm2 = 0.0;
  for (double v : A) norm2 += v * v;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
    if (off <= 1e-32 * norm2) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A J
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // J' (A J)
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  // V J
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  w.assign((size_t)n, 0.0);
  for (int i = 0; i < n; ++i) w[(size_t)i] = A[(size_t)i * n + i];
}

// Pseudo-inverse of a symmetric matrix; eigenvalues below 1e-10 of the
// largest are dropped.
static std::vector<double> sym_pinv(const std::vector<double>& A, int n) {
  std::vector<double> w, V;
  sym_eig(A, n, w, V);
  double wmax = 0.0;
  for (double x : w) wmax = std::max(wmax, std::fabs(x));
  std::vector<double> out((size_t)n * n, 0.0);
  for (int e = 0; e < n; ++e) {
    if (!(w[(size_t)e] > 1e-10 * wmax)) continue;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        out[(size_t)i * n + j] += V[(size_t)i * n + e] * V[(size_t)j * n + e] / w[(size_t)e];
  }
  return out;
}

// P'(GP) in FP64, symmetrised, on the host
static std::vector<double> coarse_matrix(wa_solver* s, const float* P, const float* GP) {
  launch_xty64(s->st, s->m, P, GP, s->cn.xpart64, s->cn.cmid, s->cn.Ac);
  std::vector<double> A((size_t)kMat);
  HIPCHK(hipMemcpyAsync(A.data(), s->cn.Ac, sizeof(double) * kMat, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (int i = 0; i < kCoarse; ++i)
    for (int j = i + 1; j < kCoarse; ++j) {
      const double v = 0.5 * (A[(size_t)i * kCoarse + j] + A[(size_t)j * kCoarse + i]);
      A[(size_t)i * kCoarse + j] = A[(size_t)j * kCoarse + i] = v;
    }
  return A;
}

// The coarse space of the apply.  The network's columns are first made
// G-orthonormal (P <- P V L^-1/2 from P'GP = V L V', dropping directions
// below 1e-6 of the largest eigenvalue: they carry almost no G-energy), so
// the apply's coarse matrix is well conditioned and the rounding of its BF16
// copies is not amplified.  Then P and GP = G P are stored in BF16, and
// (P'GP)^+ is formed from the rounded values (symmetrised), so the apply's
// projections are consistent with each other.
void wa_solver::coarse_finalize() {
  gnn_forward();
  op_apply<float>(cn.P, cn.GP, kCoarse);
  std::vector<double> w, V;
  sym_eig(coarse_matrix(this, cn.P, cn.GP), kCoarse, w, V);
  double wmax = 0.0;
  for (double x : w) wmax = std::max(wmax, x);
  std::vector<double> T((size_t)kMat, 0.0);
  int col = 0;
  for (int e = 0; e < kCoarse; ++e) {
    if (!(w[(size_t)e] > 1e-6 * wmax)) continue;
    const double f = 1.0 / std::sqrt(w[(size_t)e]);
    for (int i = 0; i < kCoarse; ++i) T[(size_t)i * kCoarse + col] = V[(size_t)i * kCoarse + e] * f;
    ++col;
  }
  cn.kept = col;
  HIPCHK(hipMemcpyAsync(cn.W12, T.data(), sizeof(double) * kMat, hipMemcpyHostToDevice, st));
  float* Pf = cn.dP;
  float* GPf = cn.dZa;
  launch_tall_small(st, m, cn.P, cn.W12, nullptr, nullptr, cn.dZb);
  launch_to_bf16(st, m * kCoarse, cn.dZb, cn.Pb, Pf);
  op_apply<float>(Pf, cn.GP, kCoarse);
  launch_to_bf16(st, m * kCoarse, cn.GP, cn.GPb, GPf);
  const std::vector<double> inv = sym_pinv(coarse_matrix(this, Pf, GPf), kCoarse);
  HIPCHK(hipMemcpyAsync(cn.Ainv, inv.data(), sizeof(double) * kMat, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  cn.trained = true;
}

// One FP32 block CG of 32 fresh right-hand sides b = G xi to a relative
// residual of 1e-6, with or without the coarse correction; returns the
// milliseconds of the second of two identical runs (the first captures the
// iteration graph) and the per-column iteration counts.
double wa_solver::coarse_validate(bool with, std::vector<int64_t>& its) {
  float* xi = cn.dZb;
  float* b = cn.Tb;
  launch_randn(st, m * kCoarse, smix64(cn.seed + 4), xi);
  op_apply<float>(xi, b, kCoarse);
  const bool saved = cn.enabled;
  cn.enabled = with;
  graphs_clear();
  std::vector<double> tol(KMAX, 1e-6);
  std::vector<int> conv(KMAX, 0);
  its.assign(KMAX, 0);
  double ms = 0.0;
  try {
    for (int rep = 0; rep < 2; ++rep) {
      HIPCHK(hipStreamSynchronize(st));
      const auto t0 = std::chrono::steady_clock::now();
      cg<float>(KMAX, KMAX, b, cx32, tol, conv, its);
      HIPCHK(hipStreamSynchronize(st));
      ms = 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  } catch (...) {
    cn.enabled = saved;
    graphs_clear();
    throw;
  }
  cn.enabled = saved;
  graphs_clear();
  return ms;
}

// z = Q r + (I - Q G) M (I - G Q) r with Q = P (P'GP)^+ P' (the balancing
// two-level preconditioner): c1 = Ainv P'r, r2 = r - GP c1, w = M r2,
// z = w + P (c1 - Ainv GP'w).
template <typename T>
void wa_solver::coarse_apply(const T* R, T* Z, int kp, const SideWork* side) {
  hipEvent_t a = nullptr;
  T* R2 = reinterpret_cast<T*>(cn.r2);
  tic(5, a);
  int nb = launch_cproj_part<T>(st, m, cn.Pb, R, kp, cn.cpart);
  launch_cproj_fin(st, cn.cpart, nb, kp, cn.cmid, cn.Ainv, nullptr, 1.0, cn.c1);
  launch_cexpand<T>(st, m, cn.GPb, cn.c1, -1.0, R, R2, kp);
  toc(5, a);
  pre_apply<T>(R2, Z, kp, side);
  tic(5, a);
  nb = launch_cproj_part<T>(st, m, cn.GPb, Z, kp, cn.cpart);
  launch_cproj_fin(st, cn.cpart, nb, kp, cn.cmid, cn.Ainv, cn.c1, -1.0, cn.cd);
  launch_cexpand<T>(st, m, cn.Pb, cn.cd, 1.0, Z, Z, kp);
  toc(5, a);
}

void wa_solver::solve_block(const double* Y, int64_t ldy, int c0, int kb, double* X, int64_t ldx, double* dem,
                            int64_t ldd, wa_rhs_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  const int kp = pow2_at_least(kb);
  hipEvent_t ev = nullptr;
  tic(3, ev);
  launch_col_to_row(st, Y + (int64_t)c0 * ldy, ldy, n, kb, Yr, kp);
  toc(3, ev);
  rhs64(Yr, B64, kp);
  std::vector<double> bn = col_norms(B64, m, kp);
  std::vector<double> rhs_norm(kb), abs_tol(kb);
  for (int c = 0; c < kb; ++c) {
    rhs_norm[c] = std::max(bn[c], 1e-15);
    abs_tol[c] = params.tol * rhs_norm[c];
  }
  std::vector<int> conv(kb, 0), cconv(kb, 0);
  std::vector<int64_t> its(kb, 0), cits(kb, 0);
  const int64_t vbytes = sizeof(double) * m * kp;
  const bool mixed = params.precision == 1 && params.krylov != 2;
  bool fresh = false;      // E and R64 hold y - Dx and D'W(y - Dx) for the current X64
  bool dof_final = false;  // R64 holds b - Gx for the current X64 (E is stale)
  if (params.krylov == 2) {
    lsmr(kb, kp, conv, its);
  } else if (!mixed) {
    std::vector<double> tol(kb, params.tol);
    krylov<double>(kb, kp, B64, cx64, tol, conv, its);
    HIPCHK(hipMemcpyAsync(X64, cx64, vbytes, hipMemcpyDeviceToDevice, st));
    std::vector<int> live(kb, 1);
    for (int64_t it = 0; it < params.max_refinements; ++it) {
      launch_demean(st, n, Q, codes, offs, X64, Yr, E, kp);
      rhs64(E, R64, kp);
      fresh = true;
      std::vector<double> cn = col_norms(R64, m, kp);
      std::vector<double> scale(KMAX, 0.0), ctol(kb, 1.0);
      bool anyc = false;
      for (int c = 0; c < kb; ++c) {
        if (live[c] && cn[c] <= abs_tol[c]) live[c] = 0;
        if (live[c]) {
          anyc = true;
          scale[c] = 1.0;
          ctol[c] = std::min(abs_tol[c] / cn[c], 1.0);
        }
      }
      if (!anyc) break;
      HIPCHK(hipMemcpyAsync(s_(8), scale.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
      launch_cast_scale<double, double>(st, R64, T64, s_(8), m, kp);
      krylov<double>(kb, kp, T64, cx64, ctol, cconv, cits);
      std::vector<int> msk(KMAX, 0);
      for (int c = 0; c < kb; ++c) msk[c] = live[c];
      HIPCHK(hipMemcpyAsync(imask + KMAX + 1, msk.data(), sizeof(int) * KMAX, hipMemcpyHostToDevice, st));
      launch_axpy_acc<double>(st, X64, cx64, nullptr, imask + KMAX + 1, m, kp);
      fresh = false;
      for (int c = 0; c < kb; ++c)
        if (live[c]) {
          its[c] += cits[c];
          conv[c] = cconv[c];
        }
    }
  } else {
    // FP32 inner Krylov, FP64 outer residual and accumulation.  With the
    // explicit Gramian every residual, the final one included, is b - G x in
    // FP64: the same vector as D'W(y - Dx) (b = D'W y, G = D'W D), with a
    // rounding of ~eps (|b| + |G||x|), far below any tolerance here.  Without
    // it the residual is D'W(y - Dx) from observation space.
    HIPCHK(hipMemsetAsync(X64, 0, vbytes, st));
    const int64_t max_outer = std::max<int64_t>(params.max_refinements, 0) + 8;
    std::vector<int> live(kb, 1), fb(kb, 0);  // fb: handed to the FP64 refinement below
    std::vector<double> rn_prev(kb, 0.0);
    const double stall = stall_ratio();
    const bool dof_resid = has_gram;
    // r = b - G x in FP64 (explicit Gramian) or D'W(y - Dx) from observation space
    auto residual = [&]() {
      if (dof_resid) {
        hipEvent_t ex = nullptr;
        tic(3, ex);  // counted with the observation-space work it replaces
        launch_spmm<double>(st, g64, X64, R64, part64, kp);
        launch_rsub<double>(st, R64, B64, m * kp);
        toc(3, ex);
        dof_final = true;
      } else {
        launch_demean(st, n, Q, codes, offs, X64, Yr, E, kp);
        rhs64(E, R64, kp);
        fresh = true;
      }
      return col_norms(R64, m, kp);
    };
    for (int64_t outer = 0; outer <= max_outer; ++outer) {
      std::vector<double> rn;
      if (outer == 0) {
        HIPCHK(hipMemcpyAsync(R64, B64, vbytes, hipMemcpyDeviceToDevice, st));
        rn = bn;
      } else {
        rn = residual();
      }
      std::vector<double> scale(KMAX, 0.0), itol(kb, 1.0);
      std::vector<int> msk(KMAX, 0);
      bool anyc = false;
      for (int c = 0; c < kb; ++c) {
        if (live[c] && (rn[c] <= abs_tol[c] || rn[c] < F64_EPS)) live[c] = 0;
        // an FP32 correction that hardly moved the FP64 residual: the column
        // has stalled in FP32 and finishes in FP64
        if (live[c] && outer > 0 && !std::isnan(rn[c]) && rn[c] > stall * rn_prev[c]) {
          live[c] = 0;
          fb[c] = 1;
        }
        rn_prev[c] = rn[c];
        if (live[c] && std::isnan(rn[c])) {
          // a NaN residual: every further pass would run out its budget, as the
          // FP64 path's initial solve and max_refinements corrections do
          live[c] = 0;
          its[c] = (1 + std::max<int64_t>(params.max_refinements, 0)) * params.maxiter;
        }
        if (live[c]) {
          anyc = true;
          scale[c] = 1.0 / rn[c];
          msk[c] = 1;
          itol[c] = std::min(1.0, std::max(abs_tol[c] / rn[c], inner_floor()));
        }
      }
      if (outer == max_outer)
        for (int c = 0; c < kb; ++c)
          if (live[c]) fb[c] = 1;  // the FP32 budget is spent: finish in FP64
      if (!anyc || outer == max_outer) break;
      HIPCHK(hipMemcpyAsync(s_(8), scale.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
      launch_cast_scale<double, float>(st, R64, rhs32, s_(8), m, kp);
      krylov<float>(kb, kp, rhs32, cx32, itol, cconv, cits);
      for (int c = 0; c < kb; ++c) scale[c] = msk[c] ? rn[c] : 0.0;
      HIPCHK(hipMemcpyAsync(s_(9), scale.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(imask + KMAX + 1, msk.data(), sizeof(int) * KMAX, hipMemcpyHostToDevice, st));
      launch_axpy_acc<float>(st, X64, cx32, s_(9), imask + KMAX + 1, m, kp);
      fresh = dof_final = false;
      for (int c = 0; c < kb; ++c)
        if (live[c]) its[c] += cits[c];
    }
    // FP64 refinement of the columns that stalled in FP32 (or ran out of FP32
    // passes): corrections solved in FP64 from the current x, as the FP64
    // path's refinement loop (solver.rs:296-347), at most max_refinements + 1
    bool anyfb = false;
    for (int c = 0; c < kb; ++c) anyfb = anyfb || fb[c];
    for (int64_t it = 0; anyfb && it <= std::max<int64_t>(params.max_refinements, 0); ++it) {
      const std::vector<double> rn = (fresh || dof_final) ? col_norms(R64, m, kp) : residual();
      std::vector<double> scale(KMAX, 0.0), ctol(kb, 1.0);
      std::vector<int> msk(KMAX, 0);
      bool anyc = false;
      for (int c = 0; c < kb; ++c) {
        if (fb[c] && (rn[c] <= abs_tol[c] || rn[c] < F64_EPS || std::isnan(rn[c]))) fb[c] = 0;
        if (fb[c]) {
          anyc = true;
          scale[c] = 1.0;
          msk[c] = 1;
          ctol[c] = std::min(abs_tol[c] / rn[c], 1.0);
        }
      }
      if (!anyc) break;
      HIPCHK(hipMemcpyAsync(s_(8), scale.data(), sizeof(double) * KMAX, hipMemcpyHostToDevice, st));
      launch_cast_scale<double, double>(st, R64, T64, s_(8), m, kp);
      krylov<double>(kb, kp, T64, cx64, ctol, cconv, cits);
      HIPCHK(hipMemcpyAsync(imask + KMAX + 1, msk.data(), sizeof(int) * KMAX, hipMemcpyHostToDevice, st));
      launch_axpy_acc<double>(st, X64, cx64, nullptr, imask + KMAX + 1, m, kp);
      fresh = dof_final = false;
      for (int c = 0; c < kb; ++c)
        if (msk[c]) its[c] += cits[c];
    }
  }
  // final residual for the final X: the observation-space D'W(y - Dx), or,
  // in the mixed mode with the explicit Gramian, the same vector as b - Gx
  // (FP64); the demeaned output is then written straight from Y
  if (!fresh && !dof_final) {
    if (mixed && has_gram) {
      launch_spmm<double>(st, g64, X64, R64, part64, kp);
      launch_rsub<double>(st, R64, B64, m * kp);
      dof_final = true;
    } else {
      launch_demean(st, n, Q, codes, offs, X64, Yr, E, kp);
      rhs64(E, R64, kp);
      fresh = true;
    }
  }
  std::vector<double> fr = col_norms(R64, m, kp);
  tic(3, ev);
  if (dem) {
    if (fresh)
      launch_row_to_col(st, E, kp, n, kb, dem + (int64_t)c0 * ldd, ldd);
    else
      launch_demean_cols(st, n, Q, codes, offs, X64, Y + (int64_t)c0 * ldy, ldy, dem + (int64_t)c0 * ldd, ldd, kb,
                         kp);
  }
  launch_row_to_col(st, X64, kp, mu, kb, X + (int64_t)c0 * ldx, ldx, int_of);
  toc(3, ev);
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  check_flow();
  double tsolve = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int c = 0; c < kb; ++c) {
    wa_rhs_stats& s = stats[c0 + c];
    s.residual = fr[c] / rhs_norm[c];
    s.iterations = its[c];
    s.converged = mixed ? (s.residual <= params.tol) : conv[c];
    s.time_solve = tsolve;
  }
  if (timer.on) timer.flush();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static void need_device(wa_solver* s) {
  if (!s) throw Error(WA_ERR_INVALID_ARGUMENT, "null solver");
  if (s->host_only) throw DevError("solver was built host-only (no device state)");
}
static void need_design(wa_solver* s) {
  need_device(s);
  if (s->plan.d.Q == 0) throw Error(WA_ERR_INVALID_ARGUMENT, "a preconditioner on its own has no design");
}

// Level A: column-major blocks of k vectors (leading dimension ld).  The
// operators run on blocks of at most KMAX columns in the handle's row-major
// work vectors, so a call allocates nothing; host pointers are staged through
// the handle's device buffer, which grows when a call needs more and is kept.
template <typename F>
static wa_status level_a(wa_solver* s, bool device, const double* in, int64_t rows_in, int64_t ldin, double* out,
                         int64_t rows_out, int64_t ldout, int32_t k, F&& body) {
  try {
    need_device(s);
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (k < 0) throw Error(WA_ERR_INVALID_ARGUMENT, "k must be >= 0");
    if (k == 0) return WA_OK;
    if (!in || !out || ldin < rows_in || ldout < rows_out)
      throw Error(WA_ERR_INVALID_ARGUMENT, "bad buffer or leading dimension");
    const double* din = in;
    double* dout = out;
    int64_t li = ldin, lo = ldout;
    if (!device) {
      double* buf = s->staging(rows_in * k + rows_out * k);
      HIPCHK(hipMemcpy2DAsync(buf, sizeof(double) * rows_in, in, sizeof(double) * ldin, sizeof(double) * rows_in,
                              (size_t)k, hipMemcpyHostToDevice, s->st));
      din = buf;
      dout = buf + rows_in * k;
      li = rows_in;
      lo = rows_out;
    }
    for (int c0 = 0; c0 < k; c0 += KMAX) {
      const int kb = std::min(KMAX, k - c0);
      body(din + (int64_t)c0 * li, li, dout + (int64_t)c0 * lo, lo, kb, pow2_at_least(kb));
    }
    if (!device)
      HIPCHK(hipMemcpy2DAsync(out, sizeof(double) * ldout, dout, sizeof(double) * rows_out,
                              sizeof(double) * rows_out, (size_t)k, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipGetLastError());
    s->check_flow();
    if (s->timer.on) s->timer.flush();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

// the Level A operators on one column block (in: ldi, out: ldo, kb columns,
// padded to kp)
static wa_status gram_apply_impl(wa_solver* s, bool device, const double* X, double* Y, int32_t k, int64_t ld) {
  if (s && s->plan.d.Q == 0) {
    g_err = "a preconditioner on its own has no design";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const int64_t mu = s ? s->mu : 0;
  return level_a(s, device, X, mu, ld, Y, mu, ld, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->m, kb, s->cp64, kp, s->ext_of);
                   const int saved = s->params.operator_repr;
                   s->params.operator_repr = 0;
                   s->op_apply<double>(s->cp64, s->cap64, kp);
                   s->params.operator_repr = saved;
                   launch_row_to_col(s->st, s->cap64, kp, s->mu, kb, out, ldo, s->int_of);
                 });
}

static wa_status gram_apply_explicit_impl(wa_solver* s, bool device, const double* X, double* Y, int32_t k,
                                          int64_t ld) {
  if (s && !s->has_gram) {
    g_err = "solver has no explicit Gramian (build with operator_repr = 1 or multiplicative Schwarz)";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const int64_t mu = s ? s->mu : 0;
  return level_a(s, device, X, mu, ld, Y, mu, ld, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->m, kb, s->cp64, kp, s->ext_of);
                   launch_spmm<double>(s->st, s->g64, s->cp64, s->cap64, s->part64, kp);
                   launch_row_to_col(s->st, s->cap64, kp, s->mu, kb, out, ldo, s->int_of);
                 });
}

static wa_status precond_apply_impl(wa_solver* s, bool device, const double* R, double* Z, int32_t k, int64_t ld) {
  if (s && s->params.preconditioner == 0) {
    g_err = "solver has no preconditioner";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const int64_t mu = s ? s->mu : 0;
  return level_a(s, device, R, mu, ld, Z, mu, ld, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->m, kb, s->cr64, kp, s->ext_of);
                   s->precondition<double>(s->cr64, s->cz64, kp);
                   launch_row_to_col(s->st, s->cz64, kp, s->mu, kb, out, ldo, s->int_of);
                 });
}

static wa_status matvec_d_impl(wa_solver* s, bool device, const double* X, int64_t ldx, double* Y, int64_t ldy,
                               int32_t k) {
  if (s && s->plan.d.Q == 0) {
    g_err = "a preconditioner on its own has no design";
    return WA_ERR_INVALID_ARGUMENT;
  }
  return level_a(s, device, X, s ? s->mu : 0, ldx, Y, s ? s->n : 0, ldy, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->m, kb, s->cp64, kp, s->ext_of);
                   launch_demean(s->st, s->n, s->Q, s->codes, s->offs, s->cp64, nullptr, s->E, kp);
                   launch_row_to_col(s->st, s->E, kp, s->n, kb, out, ldo);
                 });
}

static wa_status rmatvec_wdt_impl(wa_solver* s, bool device, const double* R, int64_t ldr, double* X, int64_t ldx,
                                  int32_t k) {
  if (s && s->plan.d.Q == 0) {
    g_err = "a preconditioner on its own has no design";
    return WA_ERR_INVALID_ARGUMENT;
  }
  return level_a(s, device, R, s ? s->n : 0, ldr, X, s ? s->mu : 0, ldx, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->n, kb, s->Yr, kp);
                   s->rhs64(s->Yr, s->B64, kp);
                   launch_row_to_col(s->st, s->B64, kp, s->mu, kb, out, ldo, s->int_of);
                 });
}

// ---------------------------------------------------------------------------
// Many independent problems (wa_multi_*).  The problems with the same number
// of factors are stacked block-diagonally into one design -- factor q's levels
// of problem p follow those of problem p - 1 and the observations are
// concatenated -- and one solver serves the stack: its factor-pair subdomains
// are connected components, which never join two problems (no observation
// links them), so each problem gets the subdomains it would get alone.  The
// Krylov scalars are kept per (problem, column) pair (multi.hip): every
// problem follows the CG recurrence of its own solve (Solver::solve_batch,
// solver.rs:242-277), while the operator and preconditioner applies serve
// all problems in one launch each.
// ---------------------------------------------------------------------------
struct MultiGroup {
  std::unique_ptr<wa_solver> s;
  std::vector<int> probs;                       // global problem index of each stacked problem
  std::vector<int64_t> obs_off, nobs, ndofs, map_off;
  const uint32_t* dmap = nullptr;               // per problem: its user DOF -> stacked internal row
  SegIndex ix;
  double *partA = nullptr, *partB = nullptr, *pairv = nullptr, *alpha = nullptr, *beta = nullptr;
  double *tolv = nullptr, *scale = nullptr, *scale2 = nullptr;
  int *mask = nullptr, *msk = nullptr, *nact = nullptr, *kcols = nullptr;
  ColState* cs = nullptr;
};

struct wa_multi {
  std::mutex mu;
  int n_problems = 0;
  std::vector<MultiGroup> groups;
  std::vector<int64_t> nobs, ndofs;  // per global problem
};

namespace {

template <typename T>
void h2d(hipStream_t st, T* dst, const std::vector<T>& v) {
  if (!v.empty()) HIPCHK(hipMemcpyAsync(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

// per-pair column norms of a stacked DOF block
std::vector<double> seg_norms(wa_solver* s, MultiGroup& g, const double* V, int kp) {
  const int np = g.ix.P * kp;
  launch_seg_dot<double>(s->st, V, nullptr, g.ix, kp, g.partA);
  launch_seg_totals(s->st, g.partA, g.ix, kp, g.pairv);
  std::vector<double> out(np);
  HIPCHK(hipMemcpyAsync(out.data(), g.pairv, sizeof(double) * np, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (double& v : out) v = std::sqrt(v);
  return out;
}

// block CG with one recurrence per (problem, column) pair (cg.rs:45-133 per pair)
template <typename T>
void seg_cg(wa_solver* s, MultiGroup& g, int kp, const T* rhs, T* x, const std::vector<double>& tol,
            std::vector<int>& conv, std::vector<int64_t>& its) {
  hipStream_t st = s->st;
  const int np = g.ix.P * kp;
  T *r = s->vec<T>(1), *z = s->vec<T>(2), *p = s->vec<T>(3), *ap = s->vec<T>(4);
  const int64_t bytes = sizeof(T) * s->m * kp;
  h2d(st, g.tolv, tol);
  HIPCHK(hipMemsetAsync(x, 0, bytes, st));
  HIPCHK(hipMemcpyAsync(r, rhs, bytes, hipMemcpyDeviceToDevice, st));
  s->precondition<T>(r, z, kp);
  HIPCHK(hipMemcpyAsync(p, z, bytes, hipMemcpyDeviceToDevice, st));
  launch_seg_dot<T>(st, r, z, g.ix, kp, g.partA);
  launch_seg_dot<T>(st, rhs, nullptr, g.ix, kp, g.partB);
  launch_seg_start(st, g.cs, g.partB, g.partA, g.ix, g.tolv, g.kcols, kp);
  auto iteration = [&]() {
    s->op_apply<T>(p, ap, kp);
    launch_seg_dot<T>(st, p, ap, g.ix, kp, g.partA);
    launch_seg_alpha(st, g.cs, g.partA, g.ix, g.alpha, kp);
    launch_seg_update<T>(st, x, p, r, ap, g.alpha, g.ix, kp, g.partA);
    launch_seg_check(st, g.cs, g.partA, g.ix, kp);
    s->precondition<T>(r, z, kp);
    launch_seg_dot<T>(st, r, z, g.ix, kp, g.partA);
    launch_seg_beta(st, g.cs, g.partA, g.ix, g.beta, g.mask, g.nact, kp);
    launch_seg_xpby<T>(st, p, z, g.beta, g.mask, g.ix, kp);
  };
  hipGraphExec_t graph = s->use_graphs() ? s->cg_graph<T>(kp, x, false, iteration, 1) : nullptr;
  std::vector<ColState> h(np);
  auto read_state = [&]() {
    HIPCHK(hipMemcpyAsync(h.data(), g.cs, sizeof(ColState) * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int a = 0;
    for (int j = 0; j < np; ++j) a += h[j].active;
    return a;
  };
  int nact = read_state();
  int64_t itn = 0;
  std::vector<double> rate(np, 0.0), before(np, 0.0);
  while (nact > 0 && itn < s->params.maxiter) {
    int chunk = graph ? replay_chunk(h, rate, s->m * kp) : 1;
    chunk = (int)std::min<int64_t>(chunk, s->params.maxiter - itn);
    for (int j = 0; j < np; ++j) before[j] = h[j].rnorm;
    for (int i = 0; i < chunk; ++i) {
      if (graph) HIPCHK(hipGraphLaunch(graph, st));
      else iteration();
    }
    itn += chunk;
    nact = read_state();
    for (int j = 0; j < np; ++j)
      if (before[j] > 0.0 && h[j].rnorm > 0.0) rate[j] = std::pow(h[j].rnorm / before[j], 1.0 / chunk);
  }
  for (int j = 0; j < np; ++j) {
    if (h[j].active || h[j].nan_stop) {
      conv[j] = 0;
      its[j] = s->params.maxiter;
    } else {
      conv[j] = h[j].converged;
      its[j] = h[j].iters;
    }
  }
}

// One column block of the stack (Yr holds it): Solver::solve + refinement
// (solver.rs:158-239, 296-347) per pair.  Leaves x in X64 and the final
// demeaned block in E; the final residual is observation-space.
void seg_solve_block(wa_solver* s, MultiGroup& g, int kp, const std::vector<int>& kc, std::vector<int>& conv,
                     std::vector<int64_t>& its, std::vector<double>& res) {
  hipStream_t st = s->st;
  const wa_params& prm = s->params;
  const int P = g.ix.P, np = P * kp;
  const int64_t m = s->m, vbytes = sizeof(double) * m * kp;
  h2d(st, g.kcols, kc);
  s->rhs64(s->Yr, s->B64, kp);
  const std::vector<double> bn = seg_norms(s, g, s->B64, kp);
  std::vector<double> rhs_norm(np), abs_tol(np);
  std::vector<int> live(np, 0);
  for (int j = 0; j < np; ++j) {
    rhs_norm[j] = std::max(bn[j], 1e-15);
    abs_tol[j] = prm.tol * rhs_norm[j];
    live[j] = (j % kp) < kc[j / kp];
  }
  conv.assign(np, 1);
  its.assign(np, 0);
  std::vector<int> cconv(np, 0);
  std::vector<int64_t> cits(np, 0);
  auto demean_resid = [&]() {
    launch_demean(st, s->n, s->Q, s->codes, s->offs, s->X64, s->Yr, s->E, kp);
    s->rhs64(s->E, s->R64, kp);
  };
  const bool mixed = prm.precision == 1;
  if (!mixed) {
    seg_cg<double>(s, g, kp, s->B64, s->cx64, std::vector<double>(np, prm.tol), conv, its);
    HIPCHK(hipMemcpyAsync(s->X64, s->cx64, vbytes, hipMemcpyDeviceToDevice, st));
    for (int64_t it = 0; it < prm.max_refinements; ++it) {
      demean_resid();
      const std::vector<double> cn = seg_norms(s, g, s->R64, kp);
      std::vector<double> scale(np, 0.0), ctol(np, 1.0);
      bool any = false;
      for (int j = 0; j < np; ++j) {
        if (live[j] && cn[j] <= abs_tol[j]) live[j] = 0;
        if (live[j]) {
          any = true;
          scale[j] = 1.0;
          ctol[j] = std::min(abs_tol[j] / cn[j], 1.0);
        }
      }
      if (!any) break;
      h2d(st, g.scale, scale);
      launch_seg_cast_scale<double, double>(st, s->R64, s->T64, g.scale, g.ix, kp);
      seg_cg<double>(s, g, kp, s->T64, s->cx64, ctol, cconv, cits);
      h2d(st, g.msk, live);
      launch_seg_axpy_acc<double>(st, s->X64, s->cx64, nullptr, g.msk, g.ix, kp);
      for (int j = 0; j < np; ++j)
        if (live[j]) {
          its[j] += cits[j];
          conv[j] = cconv[j];
        }
    }
  } else {
    HIPCHK(hipMemsetAsync(s->X64, 0, vbytes, st));
    const int64_t max_outer = std::max<int64_t>(prm.max_refinements, 0) + 8;
    for (int64_t outer = 0; outer <= max_outer; ++outer) {
      std::vector<double> rn;
      if (outer == 0) {
        HIPCHK(hipMemcpyAsync(s->R64, s->B64, vbytes, hipMemcpyDeviceToDevice, st));
        rn = bn;
      } else {
        demean_resid();
        rn = seg_norms(s, g, s->R64, kp);
      }
      std::vector<double> scale(np, 0.0), itol(np, 1.0);
      std::vector<int> msk(np, 0);
      bool any = false;
      for (int j = 0; j < np; ++j) {
        if (live[j] && (rn[j] <= abs_tol[j] || rn[j] < F64_EPS)) live[j] = 0;
        if (live[j] && std::isnan(rn[j])) {
          live[j] = 0;
          its[j] = (1 + std::max<int64_t>(prm.max_refinements, 0)) * prm.maxiter;
        }
        if (live[j]) {
          any = true;
          scale[j] = 1.0 / rn[j];
          msk[j] = 1;
          itol[j] = std::min(1.0, std::max(abs_tol[j] / rn[j], inner_floor()));
        }
      }
      if (!any || outer == max_outer) break;
      h2d(st, g.scale, scale);
      launch_seg_cast_scale<double, float>(st, s->R64, s->rhs32, g.scale, g.ix, kp);
      seg_cg<float>(s, g, kp, s->rhs32, s->cx32, itol, cconv, cits);
      for (int j = 0; j < np; ++j) scale[j] = msk[j] ? rn[j] : 0.0;
      h2d(st, g.scale2, scale);
      h2d(st, g.msk, msk);
      launch_seg_axpy_acc<float>(st, s->X64, s->cx32, g.scale2, g.msk, g.ix, kp);
      for (int j = 0; j < np; ++j)
        if (msk[j]) its[j] += cits[j];
    }
  }
  demean_resid();
  const std::vector<double> fr = seg_norms(s, g, s->R64, kp);
  res.assign(np, 0.0);
  for (int j = 0; j < np; ++j) {
    res[j] = fr[j] / rhs_norm[j];
    if (mixed) conv[j] = res[j] <= prm.tol;
  }
}

// stack the designs of one group into one block-diagonal design
Design stack_designs(const std::vector<const Design*>& ds, std::vector<std::vector<int64_t>>& lev_off) {
  const int Q = ds[0]->Q;
  Design d;
  d.Q = Q;
  d.codes.resize(Q);
  lev_off.assign(ds.size(), std::vector<int64_t>(Q, 0));
  std::vector<int64_t> lev(Q, 0);
  bool weighted = false;
  for (size_t p = 0; p < ds.size(); ++p) {
    d.n_obs += ds[p]->n_obs;
    weighted = weighted || !ds[p]->weights.empty();
    for (int q = 0; q < Q; ++q) {
      lev_off[p][q] = lev[q];
      lev[q] += ds[p]->levels[q];
    }
  }
  if (d.n_obs >= (int64_t)UINT32_MAX) throw Error(WA_ERR_OVERFLOW, "Overflow: stacked observations exceed u32 indexing");
  int64_t off = 0;
  for (int q = 0; q < Q; ++q) {
    d.codes[q].reserve(d.n_obs);
    for (size_t p = 0; p < ds.size(); ++p)
      for (u32 c : ds[p]->codes[q]) d.codes[q].push_back((u32)(c + lev_off[p][q]));
    d.levels.push_back(lev[q]);
    d.offset.push_back(off);
    off += lev[q];
  }
  if (off >= (int64_t)UINT32_MAX) throw Error(WA_ERR_OVERFLOW, "Overflow: stacked DOFs exceed u32 indexing");
  d.n_dofs = off;
  if (weighted) {
    d.weights.reserve(d.n_obs);
    for (const Design* x : ds)
      for (int64_t i = 0; i < x->n_obs; ++i) d.weights.push_back(x->w(i));
  }
  return d;
}

// tiles of the internal DOF rows (one problem each) and the per-problem maps
void build_group_index(MultiGroup& g, const std::vector<const Design*>& ds,
                       const std::vector<std::vector<int64_t>>& lev_off) {
  wa_solver* s = g.s.get();
  const Design& d = s->plan.d;
  const int P = (int)ds.size(), Q = d.Q;
  auto prob_of = [&](uint32_t u) {
    int q = Q - 1;
    while (q > 0 && (int64_t)u < d.offset[q]) --q;
    const int64_t lvl = (int64_t)u - d.offset[q];
    int lo = 0, hi = P - 1;  // last problem whose level offset is <= lvl
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (lev_off[mid][q] <= lvl) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  std::vector<SegTile> tiles;
  std::vector<std::vector<uint32_t>> per(P);
  for (int64_t i = 0; i < s->m;) {
    const uint32_t p = (uint32_t)prob_of(s->h_ext_of[i]);
    int64_t j = i + 1;
    while (j < s->m && j - i < (int64_t)kSegTileRows && (uint32_t)prob_of(s->h_ext_of[j]) == p) ++j;
    per[p].push_back((uint32_t)tiles.size());
    tiles.push_back({(uint32_t)i, (uint32_t)j, p, 0u});
    i = j;
  }
  std::vector<uint32_t> pt_ptr(1, 0), pt_idx;
  for (int p = 0; p < P; ++p) {
    pt_idx.insert(pt_idx.end(), per[p].begin(), per[p].end());
    pt_ptr.push_back((uint32_t)pt_idx.size());
  }
  std::vector<uint32_t> dmap;
  for (int p = 0; p < P; ++p) {
    g.map_off.push_back((int64_t)dmap.size());
    for (int q = 0; q < Q; ++q)
      for (int64_t l = 0; l < ds[p]->levels[q]; ++l) dmap.push_back(s->h_int_of[d.offset[q] + lev_off[p][q] + l]);
  }
  g.dmap = s->upload(dmap);
  g.ix.tiles = s->upload(tiles);
  g.ix.n_tiles = (int64_t)tiles.size();
  g.ix.pt_ptr = s->upload(pt_ptr);
  g.ix.pt_idx = s->upload(pt_idx);
  g.ix.P = P;
  const int64_t nt = std::max<int64_t>(g.ix.n_tiles, 1), npm = (int64_t)P * KMAX;
  g.partA = s->dalloc<double>(nt * KMAX);
  g.partB = s->dalloc<double>(nt * KMAX);
  g.pairv = s->dalloc<double>(npm);
  g.alpha = s->dalloc<double>(npm);
  g.beta = s->dalloc<double>(npm);
  g.tolv = s->dalloc<double>(npm);
  g.scale = s->dalloc<double>(npm);
  g.scale2 = s->dalloc<double>(npm);
  g.mask = s->dalloc<int>(npm);
  g.msk = s->dalloc<int>(npm);
  g.nact = s->dalloc<int>(1);
  g.kcols = s->dalloc<int>(P);
  g.cs = s->dalloc<ColState>(npm);
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free implicit operator apply (wa_gram_apply_noalloc): its scratch
// -- the row-major copies of a column block of X and of G X, and the
// split-row partial sums of the factor passes -- lives in one file-scope
// device array instead of the solver's work vectors.  Each GPU holds its own
// copy of the array until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for that device.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchDoubles = (int64_t)1 << 23;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device ordinal

// the array's address on device `dev` (which must be current), looked up on
// the first call on that device
double* scratch_on(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DevError("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

int64_t align32(int64_t v) { return (v + 31) / 32 * 32; }
}  // namespace

// ---------------------------------------------------------------------------
// Autotuner of the operator schedules (wa_solver_autotune).  The segmented
// reductions run over host-built schedules whose shape -- the element count of
// a work item (longer rows are split into chunks summed by a fix-up) and the
// longest row one lane group sums alone -- sets how rows map to waves.  The
// reference fixes the equivalent constants (domain.rs:191,199 chunk sizes,
// csr_block.rs:4-6); here each candidate schedule is built, uploaded and timed
// with HIP events on the solver's stream for the block width of the solves,
// and the fastest is kept.  Results do not depend on the choice beyond the
// order of the fix-up sums.
// ---------------------------------------------------------------------------
namespace {

template <typename F>
double time_launches(hipStream_t st, int reps, F&& f) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  f();  // warm
  HIPCHK(hipEventRecord(a, st));
  for (int i = 0; i < reps; ++i) f();
  HIPCHK(hipEventRecord(b, st));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return (double)ms / reps;
}

}  // namespace

static void autotune_impl(wa_solver* s, int kp, int reps, int64_t* out) {
  hipStream_t st = s->st;
  HIPCHK(hipMemsetAsync(s->cp32, 0, sizeof(float) * s->m * kp, st));
  const uint32_t chunks[] = {256, 512, 1024, 2048, 4096};
  const uint32_t shorts[] = {4, 8, 16, 32};
  // explicit Gramian (FP32, the inner Krylov's operator)
  if (s->has_gram) {
    DevSched best = s->g32.sched;
    int64_t best_c = kGramChunk, best_s = kShortRow;
    auto run = [&]() { launch_spmm<float>(st, s->g32, s->cp32, s->cap32, (float*)s->part64, kp); };
    const double t0 = time_launches(st, reps, run);
    double tb = t0;
    for (uint32_t c : chunks)
      for (uint32_t sh : shorts) {
        if (c == kGramChunk && sh == kShortRow) continue;
        Schedule sc = make_schedule(s->h_gram_ptr, c, sh);
        s->ensure_part(sc.n_slots);
        const DevSched keep = s->g32.sched;
        s->g32.sched = s->upload_sched(sc);
        const double t = time_launches(st, reps, run);
        if (t < tb) {
          tb = t;
          best = s->g32.sched;
          best_c = c;
          best_s = sh;
        }
        s->g32.sched = keep;
      }
    s->g32.sched = s->g64.sched = best;
    out[0] = best_c;
    out[1] = best_s;
    out[2] = (int64_t)(t0 * 1e3);
    out[3] = (int64_t)(tb * 1e3);
  }
  // factor passes of the implicit operator (all passes share one shape)
  if (!s->p32.empty()) {
    std::vector<DevSched> best;
    for (auto& p : s->p32) best.push_back(p.sched);
    int64_t best_c = kChunk, best_s = kShortPass;
    auto run = [&]() {
      for (int q = 0; q < s->Q; ++q) launch_pull<float>(st, s->p32[q], s->cp32, s->cap32, (float*)s->part64, kp);
    };
    const double t0 = time_launches(st, reps, run);
    double tb = t0;
    for (uint32_t c : chunks)
      for (uint32_t sh : shorts) {
        if (c == kChunk && sh == kShortPass) continue;
        std::vector<DevSched> keep, cand;
        for (int q = 0; q < s->Q; ++q) {
          Schedule sc = make_schedule(s->h_pass_ptr[q], c, sh);
          s->ensure_part(sc.n_slots);
          keep.push_back(s->p32[q].sched);
          cand.push_back(s->upload_sched(sc));
          s->p32[q].sched = cand.back();
        }
        const double t = time_launches(st, reps, run);
        if (t < tb) {
          tb = t;
          best = cand;
          best_c = c;
          best_s = sh;
        }
        for (int q = 0; q < s->Q; ++q) s->p32[q].sched = keep[q];
      }
    for (int q = 0; q < s->Q; ++q) s->p32[q].sched = s->p64[q].sched = best[q];
    out[4] = best_c;
    out[5] = best_s;
    out[6] = (int64_t)(t0 * 1e3);
    out[7] = (int64_t)(tb * 1e3);
  }
  HIPCHK(hipStreamSynchronize(st));
}

extern "C" {

const char* wa_last_error(void) { return g_err.c_str(); }

void wa_params_default(wa_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->krylov = 0;
  p->restart = 30;
  p->operator_repr = 0;
  p->preconditioner = 1;
  p->tol = 1e-8;
  p->maxiter = 1000;
  p->max_refinements = 2;
  p->precision = 0;
  p->block = 0;
  p->host_only = 0;
  p->local.ac_seed = 0;
  p->local.ac_split = 1;
  p->local.approx_schur = 1;
  p->local.schur_seed = 0;
  p->local.schur_split = 1;
  p->local.dense_threshold = 24;
  p->local.dense_core = 0;
}

wa_status wa_design_create(const uint32_t* codes, int64_t n_obs, int32_t n_factors, int64_t ld,
                           const double* weights, wa_design** out) {
  try {
    if (!out) throw Error(WA_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (n_obs > 0 && !codes) throw Error(WA_ERR_INVALID_ARGUMENT, "null codes");
    auto d = std::make_unique<wa_design>();
    d->d = make_design(codes, n_obs, n_factors, ld, weights);
    *out = d.release();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

void wa_design_destroy(wa_design* d) { delete d; }
int64_t wa_design_n_dofs(const wa_design* d) { return d ? d->d.n_dofs : 0; }
int64_t wa_design_n_obs(const wa_design* d) { return d ? d->d.n_obs : 0; }
int32_t wa_design_n_factors(const wa_design* d) { return d ? d->d.Q : 0; }
int64_t wa_design_n_levels(const wa_design* d, int32_t q) {
  return (d && q >= 0 && q < d->d.Q) ? d->d.levels[q] : 0;
}

static void local_from_cfg(const LocalConfig& c, wa_local_config& l) {
  l.ac_seed = c.ac_seed;
  l.ac_split = c.ac_split;
  l.approx_schur = c.approx_schur ? 1 : 0;
  l.schur_seed = c.schur_seed;
  l.schur_split = c.schur_split;
  l.dense_threshold = c.dense_threshold;
  l.dense_core = c.dense_core;
}

// a solver for a design, optionally with a serialized preconditioner
static std::unique_ptr<wa_solver> create_solver(const Design& dd, const wa_params* p, const uint8_t* pre,
                                                int64_t pre_len) {
  if (p->krylov < 0 || p->krylov > 2) throw Error(WA_ERR_INVALID_ARGUMENT, "unknown Krylov method");
  if (p->krylov == 2 && p->preconditioner == 2)
    throw Error(WA_ERR_INVALID_ARGUMENT, "LSMR requires a symmetric preconditioner; use additive or GMRES");
  if (p->local_size < 0) throw Error(WA_ERR_INVALID_ARGUMENT, "LSMR local_size must be >= 0");
  if (p->preconditioner < 0 || p->preconditioner > 2) throw Error(WA_ERR_INVALID_ARGUMENT, "unknown preconditioner");
  if (p->preconditioner == 2 && p->krylov == 0)
    throw Error(WA_ERR_INVALID_ARGUMENT, "CG requires a symmetric preconditioner; use additive or GMRES");
  if (p->krylov == 1 && p->restart <= 0) throw Error(WA_ERR_INVALID_ARGUMENT, "GMRES restart must be positive");
  if (dd.Q > 8) throw Error(WA_ERR_UNSUPPORTED, "at most 8 factors are supported");
  if (p->local.ac_split < 1 || p->local.schur_split < 1)
    throw Error(WA_ERR_INVALID_ARGUMENT, "split must be >= 1");
  auto s = std::make_unique<wa_solver>();
  s->params = *p;
  s->host_only = p->host_only != 0;
  Plan& pl = s->plan;
  pl.d = dd;
  pl.cfg.ac_seed = p->local.ac_seed;
  pl.cfg.ac_split = p->local.ac_split;
  pl.cfg.approx_schur = p->local.approx_schur != 0;
  pl.cfg.schur_seed = p->local.schur_seed;
  pl.cfg.schur_split = p->local.schur_split;
  pl.cfg.dense_threshold = p->local.dense_threshold;
  pl.cfg.dense_core = p->local.dense_core == 0 ? 16384 : p->local.dense_core;
  pl.cfg.device = !s->host_only;
  pl.precond = p->preconditioner;
  pl.explicit_op = p->operator_repr == 1;
  if (pre) {
    const int64_t nd = deserialize_precond(pre, pre_len, pl);
    if (nd != pl.d.n_dofs)
      throw Error(WA_ERR_INVALID_ARGUMENT, "preconditioner was built for a design with a different number of DOFs");
    if (pl.precond == 2 && p->krylov != 1)
      throw Error(WA_ERR_INVALID_ARGUMENT, "CG and LSMR require a symmetric preconditioner; use additive or GMRES");
    std::vector<char> seen(pl.d.n_dofs, 0);
    for (int q = 0; q < pl.d.Q; ++q)
      for (u32 c : pl.d.codes[q]) seen[pl.d.offset[q] + c] = 1;
    for (u32 g : pl.pre.loc_gidx)
      if (!seen[g])
        throw Error(WA_ERR_INVALID_ARGUMENT,
                    "the preconditioner does not match the design (it covers an unobserved level)");
    if (pl.precond == 2) pl.gram = Csr(), pl.has_gram = false;  // rebuilt from the design
    s->params.preconditioner = pl.precond;
    local_from_cfg(pl.cfg, s->params.local);
  }
  pl.layout_early = !s->host_only;
  build_plan(pl);
  if (!s->host_only) s->setup_device();
  return s;
}

// solver from a design, optionally with a serialized preconditioner
static wa_status create_impl(const wa_design* d, const wa_params* p, const uint8_t* pre, int64_t pre_len,
                             wa_solver** out) {
  try {
    if (!d || !p || !out) throw Error(WA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    *out = create_solver(d->d, p, pre, pre_len).release();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_solver_create(const wa_design* d, const wa_params* p, wa_solver** out) {
  return create_impl(d, p, nullptr, 0, out);
}

wa_status wa_solver_create_with_precond(const wa_design* d, const wa_params* p, const uint8_t* bytes, int64_t len,
                                        wa_solver** out) {
  if (!bytes) {
    g_err = "null preconditioner bytes";
    return WA_ERR_INVALID_ARGUMENT;
  }
  return create_impl(d, p, bytes, len, out);
}

wa_status wa_precond_create(const uint8_t* bytes, int64_t len, int32_t host_only, wa_solver** out) {
  try {
    if (!bytes || !out) throw Error(WA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto s = std::make_unique<wa_solver>();
    wa_params_default(&s->params);
    s->host_only = host_only != 0;
    Plan& pl = s->plan;
    pl.d.n_dofs = deserialize_precond(bytes, len, pl);
    s->params.preconditioner = pl.precond;
    s->params.krylov = pl.precond == 2 ? 1 : 0;
    local_from_cfg(pl.cfg, s->params.local);
    if (!s->host_only) s->setup_device();
    *out = s.release();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_solver_precond_bytes(const wa_solver* s, uint8_t* buf, int64_t* len) {
  try {
    if (!s || !len) throw Error(WA_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (s->plan.precond == 0) throw Error(WA_ERR_INVALID_ARGUMENT, "solver has no preconditioner");
    if (s->precond_bytes.empty()) s->precond_bytes = serialize_precond(s->plan);
    const int64_t need = (int64_t)s->precond_bytes.size();
    if (buf) {
      if (*len < need) throw Error(WA_ERR_INVALID_ARGUMENT, "buffer too small for the preconditioner bytes");
      std::memcpy(buf, s->precond_bytes.data(), (size_t)need);
    }
    *len = need;
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_solver_params(const wa_solver* s, wa_params* out) {
  if (!s || !out) {
    g_err = "null argument";
    return WA_ERR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(s->call_mu);
  *out = s->params;
  return WA_OK;
}

void wa_solver_destroy(wa_solver* s) { delete s; }

wa_status wa_solver_set_params(wa_solver* s, const wa_params* p) {
  try {
    if (!s || !p) throw Error(WA_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lock(s->call_mu);
    const wa_params& o = s->params;
    if (p->preconditioner != o.preconditioner || std::memcmp(&p->local, &o.local, sizeof(wa_local_config)) != 0)
      throw Error(WA_ERR_INVALID_ARGUMENT, "preconditioner settings are fixed when the solver is built");
    if (p->operator_repr == 1 && !s->has_gram)
      throw Error(WA_ERR_INVALID_ARGUMENT, "solver has no explicit Gramian");
    if (p->krylov < 0 || p->krylov > 2) throw Error(WA_ERR_INVALID_ARGUMENT, "unknown Krylov method");
    if (p->krylov != 1 && p->preconditioner == 2)
      throw Error(WA_ERR_INVALID_ARGUMENT, "CG and LSMR require a symmetric preconditioner");
    if (p->local_size < 0) throw Error(WA_ERR_INVALID_ARGUMENT, "LSMR local_size must be >= 0");
    if (p->krylov == 1 && p->restart <= 0) throw Error(WA_ERR_INVALID_ARGUMENT, "GMRES restart must be positive");
    s->params = *p;
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

static wa_status solve_impl(wa_solver* s, const double* Y, int32_t k, int64_t ldy, double* X, int64_t ldx,
                            double* dem, int64_t ldd, wa_rhs_stats* stats, bool device) {
  try {
    need_design(s);
    if (k < 0 || !stats || (k > 0 && (!Y || !X))) throw Error(WA_ERR_INVALID_ARGUMENT, "bad arguments");
    if (ldy < s->n || ldx < s->mu || (dem && ldd < s->n))
      throw Error(WA_ERR_OBS_COUNT_MISMATCH, "ObservationCountMismatch: leading dimension too small");
    if (k == 0) return WA_OK;
    std::lock_guard<std::mutex> lock(s->call_mu);
    const double* dY = Y;
    double *dX = X, *dD = dem;
    if (!device) {
      const int64_t need = ldy * k + s->mu * (int64_t)k + (dem ? s->n * (int64_t)k : 0);
      double* y = s->staging(need);
      HIPCHK(hipMemcpyAsync(y, Y, sizeof(double) * ldy * k, hipMemcpyHostToDevice, s->st));
      dY = y;
      dX = y + ldy * k;
      dD = dem ? dX + s->mu * (int64_t)k : nullptr;
    }
    const int64_t lx = device ? ldx : s->mu, ld = device ? ldd : s->n;
    int blk = s->params.block <= 0 ? KMAX : std::min(s->params.block, KMAX);
    for (int c0 = 0; c0 < k; c0 += blk) {
      int kb = std::min(blk, k - c0);
      s->solve_block(dY, ldy, c0, kb, dX, lx, dD, ld, stats);
    }
    if (!device) {
      for (int c = 0; c < k; ++c)
        HIPCHK(hipMemcpyAsync(X + (int64_t)c * ldx, dX + (int64_t)c * s->mu, sizeof(double) * s->mu,
                              hipMemcpyDeviceToHost, s->st));
      if (dem)
        for (int c = 0; c < k; ++c)
          HIPCHK(hipMemcpyAsync(dem + (int64_t)c * ldd, dD + (int64_t)c * s->n, sizeof(double) * s->n,
                                hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipStreamSynchronize(s->st));
    }
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_solve_batch(wa_solver* s, const double* Y, int32_t k, int64_t ldy, double* X, int64_t ldx,
                         double* demeaned, int64_t ldd, wa_rhs_stats* stats) {
  return solve_impl(s, Y, k, ldy, X, ldx, demeaned, ldd, stats, false);
}
wa_status wa_solve_batch_device(wa_solver* s, const double* Y, int32_t k, int64_t ldy, double* X, int64_t ldx,
                                double* demeaned, int64_t ldd, wa_rhs_stats* stats) {
  return solve_impl(s, Y, k, ldy, X, ldx, demeaned, ldd, stats, true);
}

wa_status wa_gram_apply(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld) {
  return gram_apply_impl(s, false, X, Y, k, ld);
}
wa_status wa_gram_apply_device(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld) {
  return gram_apply_impl(s, true, X, Y, k, ld);
}
wa_status wa_gram_apply_noalloc(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld) {
  try {
    need_design(s);
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (k < 0) throw Error(WA_ERR_INVALID_ARGUMENT, "k must be >= 0");
    if (k == 0) return WA_OK;
    if (!X || !Y || ld < s->mu) throw Error(WA_ERR_INVALID_ARGUMENT, "bad buffer or leading dimension");
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != s->dev) throw Error(WA_ERR_INVALID_ARGUMENT, "wa_gram_apply_noalloc: make the solver's device current");
    int64_t slots = 1;
    for (const DevPass<double>& p : s->p64) slots = std::max(slots, p.sched.n_slots);
    // per block of kp columns: X rows, G X rows, pass partials
    auto need = [&](int kp) { return 2 * align32(s->m * kp) + align32(slots * kp); };
    if (need(1) > kScratchDoubles) {
      char msg[200];
      std::snprintf(msg, sizeof msg,
                    "wa_gram_apply_noalloc: the design needs %lld bytes of scratch per column, more than the "
                    "%lld-byte scratch array",
                    (long long)(need(1) * 8), (long long)(kScratchDoubles * 8));
      throw Error(WA_ERR_INVALID_ARGUMENT, msg);
    }
    int kbmax = KMAX;
    while (need(kbmax) > kScratchDoubles) kbmax /= 2;
    double* sp = scratch_on(dev);
    for (int c0 = 0; c0 < k; c0 += kbmax) {
      const int kb = std::min(kbmax, k - c0), kp = pow2_at_least(kb);
      double* xr = sp;
      double* gx = sp + align32(s->m * kp);
      double* part = gx + align32(s->m * kp);
      launch_col_to_row(s->st, X + (int64_t)c0 * ld, ld, s->m, kb, xr, kp, s->ext_of);
      for (int q = 0; q < s->Q; ++q) launch_pull<double>(s->st, s->p64[q], xr, gx, part, kp);
      launch_row_to_col(s->st, gx, kp, s->mu, kb, Y + (int64_t)c0 * ld, ld, s->int_of);
    }
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipGetLastError());
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}
wa_status wa_gram_apply_explicit(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld) {
  return gram_apply_explicit_impl(s, false, X, Y, k, ld);
}
wa_status wa_gram_apply_explicit_device(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld) {
  return gram_apply_explicit_impl(s, true, X, Y, k, ld);
}
wa_status wa_precond_apply(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld) {
  return precond_apply_impl(s, false, R, Z, k, ld);
}
wa_status wa_precond_apply_device(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld) {
  return precond_apply_impl(s, true, R, Z, k, ld);
}
wa_status wa_matvec_d(wa_solver* s, const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t k) {
  return matvec_d_impl(s, false, X, ldx, Y, ldy, k);
}
wa_status wa_matvec_d_device(wa_solver* s, const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t k) {
  return matvec_d_impl(s, true, X, ldx, Y, ldy, k);
}
wa_status wa_rmatvec_wdt(wa_solver* s, const double* R, int64_t ldr, double* X, int64_t ldx, int32_t k) {
  return rmatvec_wdt_impl(s, false, R, ldr, X, ldx, k);
}
wa_status wa_rmatvec_wdt_device(wa_solver* s, const double* R, int64_t ldr, double* X, int64_t ldx, int32_t k) {
  return rmatvec_wdt_impl(s, true, R, ldr, X, ldx, k);
}

wa_status wa_multi_create(const wa_design* const* designs, int32_t n_problems, const wa_params* p,
                          wa_multi** out) {
  try {
    if (!designs || !p || !out || n_problems <= 0) throw Error(WA_ERR_INVALID_ARGUMENT, "bad arguments");
    *out = nullptr;
    if (p->krylov != 0)
      throw Error(WA_ERR_UNSUPPORTED, "a batch of problems is solved with CG (KrylovMethod::Cg) only");
    if (p->preconditioner == 2)
      throw Error(WA_ERR_INVALID_ARGUMENT, "CG requires a symmetric preconditioner; use additive or GMRES");
    if (p->host_only) throw Error(WA_ERR_INVALID_ARGUMENT, "a batch of problems needs the device");
    auto mp = std::make_unique<wa_multi>();
    mp->n_problems = n_problems;
    std::vector<int> qs;  // factor counts in order of first appearance
    for (int i = 0; i < n_problems; ++i) {
      if (!designs[i]) throw Error(WA_ERR_INVALID_ARGUMENT, "null design");
      mp->nobs.push_back(designs[i]->d.n_obs);
      mp->ndofs.push_back(designs[i]->d.n_dofs);
      if (std::find(qs.begin(), qs.end(), designs[i]->d.Q) == qs.end()) qs.push_back(designs[i]->d.Q);
    }
    for (int Q : qs) {
      MultiGroup g;
      std::vector<const Design*> ds;
      for (int i = 0; i < n_problems; ++i)
        if (designs[i]->d.Q == Q) {
          g.probs.push_back(i);
          ds.push_back(&designs[i]->d);
        }
      std::vector<std::vector<int64_t>> lev_off;
      const Design stacked = stack_designs(ds, lev_off);
      int64_t o = 0;
      for (const Design* d : ds) {
        g.obs_off.push_back(o);
        g.nobs.push_back(d->n_obs);
        g.ndofs.push_back(d->n_dofs);
        o += d->n_obs;
      }
      g.s = create_solver(stacked, p, nullptr, 0);
      build_group_index(g, ds, lev_off);
      mp->groups.push_back(std::move(g));
    }
    *out = mp.release();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

void wa_multi_destroy(wa_multi* mp) { delete mp; }

wa_status wa_multi_info(const wa_multi* mp, int32_t group, int64_t* info16) {
  if (!mp || !info16 || group < 0 || group >= (int32_t)mp->groups.size()) {
    g_err = "bad group index";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const wa_status rc = wa_solver_info(mp->groups[group].s.get(), info16);
  if (rc == WA_OK) info16[15] = (int64_t)mp->groups[group].probs.size();
  return rc;
}

wa_status wa_multi_solve(wa_multi* mp, const double* const* Y, const int32_t* k, const int64_t* ldy,
                         double* const* X, const int64_t* ldx, double* const* demeaned, const int64_t* ldd,
                         wa_rhs_stats* const* stats) {
  try {
    if (!mp || !Y || !k || !ldy || !X || !ldx || !stats) throw Error(WA_ERR_INVALID_ARGUMENT, "bad arguments");
    if (demeaned && !ldd) throw Error(WA_ERR_INVALID_ARGUMENT, "demeaned needs leading dimensions");
    for (int i = 0; i < mp->n_problems; ++i) {
      if (k[i] < 0) throw Error(WA_ERR_INVALID_ARGUMENT, "k must be >= 0");
      if (k[i] == 0) continue;
      if (!Y[i] || !X[i] || !stats[i]) throw Error(WA_ERR_INVALID_ARGUMENT, "null buffer");
      if (ldy[i] < mp->nobs[i] || ldx[i] < mp->ndofs[i] || (demeaned && demeaned[i] && ldd[i] < mp->nobs[i]))
        throw Error(WA_ERR_OBS_COUNT_MISMATCH, "ObservationCountMismatch: leading dimension too small");
    }
    std::lock_guard<std::mutex> lock(mp->mu);
    const auto t0 = std::chrono::steady_clock::now();
    for (MultiGroup& g : mp->groups) {
      wa_solver* s = g.s.get();
      std::lock_guard<std::mutex> lock2(s->call_mu);
      hipStream_t st = s->st;
      const int P = (int)g.probs.size();
      int kmax = 0;
      int64_t out_rows = 0;  // every problem's x and demeaned rows, one after another
      for (int j = 0; j < P; ++j) {
        kmax = std::max(kmax, k[g.probs[j]]);
        out_rows += g.nobs[j] + g.ndofs[j];
      }
      for (int c0 = 0; c0 < kmax; c0 += KMAX) {
        const int kb = std::min(KMAX, kmax - c0), kp = pow2_at_least(kb);
        std::vector<int> kc(P);
        for (int j = 0; j < P; ++j) kc[j] = std::max(0, std::min(kb, k[g.probs[j]] - c0));
        // stage this column block of every problem as one stacked column-major block
        double* stg = s->staging(std::max(s->n, out_rows) * (int64_t)kb);
        HIPCHK(hipMemsetAsync(stg, 0, sizeof(double) * s->n * kb, st));
        for (int j = 0; j < P; ++j) {
          const int i = g.probs[j];
          if (kc[j] == 0) continue;
          HIPCHK(hipMemcpy2DAsync(stg + g.obs_off[j], sizeof(double) * s->n, Y[i] + (int64_t)c0 * ldy[i],
                                  sizeof(double) * ldy[i], sizeof(double) * g.nobs[j], (size_t)kc[j],
                                  hipMemcpyDefault, st));
        }
        launch_col_to_row(st, stg, s->n, s->n, kb, s->Yr, kp);
        std::vector<int> conv;
        std::vector<int64_t> its;
        std::vector<double> res;
        seg_solve_block(s, g, kp, kc, conv, its, res);
        // the outputs reuse the staging block (stream order: the input was consumed)
        double* o_next = stg;
        for (int j = 0; j < P; ++j) {
          const int i = g.probs[j];
          if (kc[j] == 0) continue;
          double* xo = o_next;
          double* dm = xo + g.ndofs[j] * kc[j];
          o_next = dm + g.nobs[j] * kc[j];
          launch_row_to_col(st, s->X64, kp, g.ndofs[j], kc[j], xo, g.ndofs[j], g.dmap + g.map_off[j]);
          HIPCHK(hipMemcpy2DAsync(X[i] + (int64_t)c0 * ldx[i], sizeof(double) * ldx[i], xo,
                                  sizeof(double) * g.ndofs[j], sizeof(double) * g.ndofs[j], (size_t)kc[j],
                                  hipMemcpyDefault, st));
          if (demeaned && demeaned[i]) {
            launch_row_to_col(st, s->E + g.obs_off[j] * kp, kp, g.nobs[j], kc[j], dm, g.nobs[j]);
            HIPCHK(hipMemcpy2DAsync(demeaned[i] + (int64_t)c0 * ldd[i], sizeof(double) * ldd[i], dm,
                                    sizeof(double) * g.nobs[j], sizeof(double) * g.nobs[j], (size_t)kc[j],
                                    hipMemcpyDefault, st));
          }
          for (int c = 0; c < kc[j]; ++c) {
            wa_rhs_stats& o = stats[i][c0 + c];
            o.converged = conv[j * kp + c];
            o.iterations = its[j * kp + c];
            o.residual = res[j * kp + c];
          }
        }
        HIPCHK(hipStreamSynchronize(st));  // the next block reuses the staging block
      }
      HIPCHK(hipGetLastError());
      if (s->timer.on) s->timer.flush();
    }
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < mp->n_problems; ++i)
      for (int c = 0; c < k[i]; ++c) stats[i][c].time_solve = t;
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_solver_autotune(wa_solver* s, int32_t k, int32_t reps, int64_t* out8) {
  try {
    need_design(s);
    if (!out8 || k <= 0 || reps <= 0) throw Error(WA_ERR_INVALID_ARGUMENT, "bad arguments");
    std::lock_guard<std::mutex> lock(s->call_mu);
    for (int i = 0; i < 8; ++i) out8[i] = -1;
    autotune_impl(s, pow2_at_least(std::min(k, KMAX)), reps, out8);
    s->graphs_clear();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

void wa_coarse_params_default(wa_coarse_params* p) {
  if (!p) return;
  *p = wa_coarse_params{};
  p->steps = 200;
  p->sample_iters = 8;
  p->seed = 0;
  p->lr = 3e-3;
  p->force = 0;
  p->ridge = 1e-4;
}

wa_status wa_coarse_train(wa_solver* s, const wa_coarse_params* p, wa_coarse_report* out) {
  try {
    need_design(s);
    if (!p || !out || p->steps < 0 || p->sample_iters < 1 || !(p->lr >= 0.0) || !(p->ridge >= 0.0))
      throw Error(WA_ERR_INVALID_ARGUMENT, "bad coarse training parameters");
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (!s->coarse_supported())
      throw Error(WA_ERR_UNSUPPORTED, "the learned coarse correction needs additive Schwarz with subdomains");
    const auto t0 = std::chrono::steady_clock::now();
    *out = wa_coarse_report{};
    s->cn.enabled = false;
    s->graphs_clear();
    s->coarse_alloc(p->seed);
    s->cn.ridge = p->ridge;
    s->coarse_init_params(p->seed);
    s->coarse_samples(p->sample_iters, p->seed);
    s->gnn_forward();
    s->gnn_loss(p->steps > 0);
    out->loss_first = s->gnn_read_loss();
    double b1t = 1.0, b2t = 1.0;
    for (int step = 1; step <= p->steps; ++step) {
      if (step > 1) {
        s->gnn_forward();
        s->gnn_loss(true);
      }
      b1t *= 0.9;
      b2t *= 0.999;
      launch_gnn_adam(s->st, s->cn.theta, s->cn.grad, s->cn.m1, s->cn.m2, p->lr, 1.0 - b1t, 1.0 - b2t);
      launch_gnn_pack(s->st, s->cn.theta, s->cn.pack);
    }
    s->gnn_forward();
    s->gnn_loss(false);
    out->loss_last = s->gnn_read_loss();
    out->steps = p->steps;
    s->coarse_finalize();
    std::vector<int64_t> ip, ic;
    out->ms_plain = s->coarse_validate(false, ip);
    out->ms_coarse = s->coarse_validate(true, ic);
    double a = 0.0, b = 0.0;
    for (int c = 0; c < KMAX; ++c) {
      a += (double)ip[c];
      b += (double)ic[c];
    }
    out->iters_plain = a / KMAX;
    out->iters_coarse = b / KMAX;
    // plain Schwarz-PCG stays unless the correction cuts the validation
    // block's iterations and its solve time
    const bool on = p->force > 0   ? true
                    : p->force < 0 ? false
                                   : out->iters_coarse < out->iters_plain && out->ms_coarse < out->ms_plain;
    s->cn.enabled = on;
    s->graphs_clear();
    out->enabled = on ? 1 : 0;
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipGetLastError());
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    out->kept = s->cn.kept;
    return WA_OK;
  } catch (const std::exception& e) {
    if (s) s->cn.enabled = false;
    return (wa_status)set_error(e);
  }
}

wa_status wa_coarse_enable(wa_solver* s, int32_t on) {
  try {
    need_design(s);
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (on && !s->cn.trained) throw Error(WA_ERR_INVALID_ARGUMENT, "the coarse correction is not trained");
    s->cn.enabled = on != 0;
    s->graphs_clear();
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_coarse_eval(wa_solver* s, const float* theta, double* loss, float* grad) {
  try {
    need_design(s);
    if (!theta || !loss) throw Error(WA_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (!s->cn.ready || !(s->cn.nu > 0.0)) throw Error(WA_ERR_INVALID_ARGUMENT, "no coarse training set");
    HIPCHK(hipMemcpyAsync(s->cn.theta, theta, sizeof(float) * kGnnParams, hipMemcpyHostToDevice, s->st));
    launch_gnn_pack(s->st, s->cn.theta, s->cn.pack);
    s->gnn_forward();
    s->gnn_loss(grad != nullptr);
    *loss = s->gnn_read_loss();
    if (grad) {
      HIPCHK(hipMemcpyAsync(grad, s->cn.grad, sizeof(float) * kGnnParams, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipStreamSynchronize(s->st));
    }
    HIPCHK(hipGetLastError());
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_coarse_export(const wa_solver* cs, int32_t what, void* buf, int64_t cap, int64_t* len) {
  try {
    wa_solver* s = const_cast<wa_solver*>(cs);
    need_design(s);
    if (!len) throw Error(WA_ERR_INVALID_ARGUMENT, "null length");
    std::lock_guard<std::mutex> lock(s->call_mu);
    if (!s->cn.ready) throw Error(WA_ERR_INVALID_ARGUMENT, "no coarse training set");
    const int64_t mm = s->m * kCoarse;
    const void* src = nullptr;
    int64_t n = 0, esz = 4;
    bool bf = false;
    switch (what) {
      case 0: src = s->cn.X0; n = mm; break;
      case 1: src = s->cn.E; n = mm; break;
      case 2: src = s->cn.R; n = mm; break;
      case 3: src = s->cn.theta; n = kGnnParams; break;
      case 4: src = s->ext_of; n = s->m; break;
      case 5: n = 1; esz = 8; break;
      case 6: src = s->cn.Pb; n = mm; bf = true; break;
      case 7: src = s->cn.GPb; n = mm; bf = true; break;
      case 8: src = s->cn.Ainv; n = kMat; esz = 8; break;
      case 9: src = s->cn.dinv; n = s->m; break;
      case 10: src = s->cn.dP; n = mm; break;
      case 11: src = s->cn.P; n = mm; break;
      case 12: src = s->cn.GP; n = mm; break;
      default: throw Error(WA_ERR_INVALID_ARGUMENT, "unknown export");
    }
    if ((what == 6 || what == 7 || what == 8) && !s->cn.trained)
      throw Error(WA_ERR_INVALID_ARGUMENT, "the coarse correction is not trained");
    *len = n;
    if (!buf) return WA_OK;
    if (cap < n) throw Error(WA_ERR_INVALID_ARGUMENT, "buffer too small");
    if (what == 5) {
      *static_cast<double*>(buf) = s->cn.nu;
    } else if (bf) {
      std::vector<uint16_t> h((size_t)n);
      HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(uint16_t) * n, hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipStreamSynchronize(s->st));
      float* o = static_cast<float*>(buf);
      for (int64_t i = 0; i < n; ++i) {
        const uint32_t u = (uint32_t)h[(size_t)i] << 16;
        std::memcpy(&o[i], &u, 4);
      }
    } else {
      HIPCHK(hipMemcpyAsync(buf, src, (size_t)(esz * n), hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipStreamSynchronize(s->st));
    }
    return WA_OK;
  } catch (const std::exception& e) {
    return (wa_status)set_error(e);
  }
}

wa_status wa_coarse_precond_apply(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld) {
  if (s && (!s->cn.trained || !s->coarse_supported())) {
    g_err = "the coarse correction is not trained";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const int64_t mu = s ? s->mu : 0;
  return level_a(s, false, R, mu, ld, Z, mu, ld, k,
                 [&](const double* in, int64_t ldi, double* out, int64_t ldo, int kb, int kp) {
                   launch_col_to_row(s->st, in, ldi, s->m, kb, s->cr64, kp, s->ext_of);
                   s->coarse_apply<double>(s->cr64, s->cz64, kp, nullptr);
                   launch_row_to_col(s->st, s->cz64, kp, s->mu, kb, out, ldo, s->int_of);
                 });
}

wa_status wa_precond_work(const wa_solver* s, int64_t* work, int64_t* out4) {
  if (!s || !out4) { g_err = "null argument"; return WA_ERR_INVALID_ARGUMENT; }
  if (s->plan.precond == 0) { g_err = "solver has no preconditioner"; return WA_ERR_INVALID_ARGUMENT; }
  int64_t total = 0, mx = 0, scatter = 0;
  for (size_t i = 0; i < s->plan.meta.size(); ++i) {
    const SubMeta& m = s->plan.meta[i];
    const int64_t w = std::max(m.n_q, m.n_r) > 10000 ? 2 * m.c_nnz + m.n_q + m.n_r : 0;
    if (work) work[i] = w;
    total += w;
    mx = std::max(mx, w);
    scatter += m.n_q + m.n_r;
  }
  out4[0] = total;
  out4[1] = mx;
  out4[2] = scatter;
  out4[3] = s->plan.d.n_dofs;
  return WA_OK;
}

wa_status wa_solver_info(const wa_solver* s, int64_t* info) {
  if (!s || !info) { g_err = "null argument"; return WA_ERR_INVALID_ARGUMENT; }
  const PrecondFlat& f = s->plan.pre;
  std::memset(info, 0, sizeof(int64_t) * 16);
  info[0] = (int64_t)s->plan.meta.size();
  info[1] = f.n_loc;
  info[2] = f.n_red;
  info[3] = f.lvl_ptr.empty() ? 0 : (int64_t)f.lvl_ptr.size() - 2;
  info[4] = f.ac_nnz;
  info[5] = (int64_t)f.core_idx.size();
  info[6] = f.core_ptr.empty() ? 0 : (int64_t)f.core_ptr.size() - 1;
  info[7] = f.c_nnz;
  info[8] = s->plan.has_gram ? (int64_t)s->plan.gram.indices.size() : 0;
  info[9] = s->plan.d.n_dofs;
  info[10] = s->plan.d.n_obs;
  info[11] = (int64_t)s->plan.passes.size();
  info[12] = f.dense_subs;
  info[13] = f.max_core;
  info[14] = s->host_only ? 0 : s->m;  // DOFs the device works on (active levels)
  // streams of the additive apply; trailing levels run as one launch per direction
  info[15] = s->host_only ? 0 : (s->overlap ? 2 : 1) + ((int64_t)s->flow_nl << 8);
  return WA_OK;
}

wa_status wa_solver_subdomain(const wa_solver* s, int64_t i, int64_t* meta, uint32_t* gidx, double* pou) {
  if (!s || !meta || i < 0 || i >= (int64_t)s->plan.subs.size()) {
    g_err = "bad subdomain index";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const Subdomain& sd = s->plan.subs[i];
  meta[0] = sd.q; meta[1] = sd.r; meta[2] = sd.ct.n_q(); meta[3] = sd.ct.n_r();
  meta[4] = sd.eliminate_q; meta[5] = sd.dense;
  meta[6] = (int64_t)sd.f.round_ptr.size() - 1;
  meta[7] = (int64_t)sd.f.core.size();
  if (gidx) std::memcpy(gidx, sd.gidx.data(), sizeof(uint32_t) * sd.gidx.size());
  if (pou)
    for (size_t k = 0; k < sd.gidx.size(); ++k) pou[k] = sd.pou.empty() ? 1.0 : sd.pou[k];
  return WA_OK;
}

wa_status wa_solver_subdomain_factor(const wa_solver* s, int64_t i, int64_t* nnz_out, double* pivot,
                                     uint32_t* col_ptr, uint32_t* col_idx, double* col_w, uint32_t* round_ptr,
                                     uint32_t* order, uint32_t* core, double* core_inv) {
  if (!s || i < 0 || i >= (int64_t)s->plan.subs.size()) {
    g_err = "bad subdomain index";
    return WA_ERR_INVALID_ARGUMENT;
  }
  const Factor& f = s->plan.subs[i].f;
  if (nnz_out) *nnz_out = (int64_t)f.col_idx.size();
  if (pivot) std::memcpy(pivot, f.pivot.data(), sizeof(double) * f.pivot.size());
  if (col_ptr) std::memcpy(col_ptr, f.col_ptr.data(), sizeof(uint32_t) * f.col_ptr.size());
  if (col_idx) std::memcpy(col_idx, f.col_idx.data(), sizeof(uint32_t) * f.col_idx.size());
  if (col_w) std::memcpy(col_w, f.col_w.data(), sizeof(double) * f.col_w.size());
  if (round_ptr) std::memcpy(round_ptr, f.round_ptr.data(), sizeof(uint32_t) * f.round_ptr.size());
  if (order) std::memcpy(order, f.order.data(), sizeof(uint32_t) * f.order.size());
  if (core) std::memcpy(core, f.core.data(), sizeof(uint32_t) * f.core.size());
  if (core_inv) {
    if (f.core_dev) {
      const std::vector<double> h = f.core_dev->download();
      std::memcpy(core_inv, h.data(), sizeof(double) * h.size());
    } else {
      std::memcpy(core_inv, f.core_inv.data(), sizeof(double) * f.core_inv.size());
    }
  }
  return WA_OK;
}

void wa_timing_enable(wa_solver* s, int32_t on) {
  if (!s) return;
  std::lock_guard<std::mutex> lock(s->call_mu);
  s->timer.on = on != 0 && !s->host_only;
}
void wa_timing_reset(wa_solver* s) {
  if (!s) return;
  std::lock_guard<std::mutex> lock(s->call_mu);
  for (int i = 0; i < 6; ++i) { s->timer.launches[i] = 0; s->timer.ms[i] = 0; }
}
wa_status wa_timing_get(wa_solver* s, int32_t kclass, int64_t* launches, double* ms) {
  if (!s || kclass < 0 || kclass > 5) { g_err = "bad timing class"; return WA_ERR_INVALID_ARGUMENT; }
  std::lock_guard<std::mutex> lock(s->call_mu);
  if (launches) *launches = s->timer.launches[kclass];
  if (ms) *ms = s->timer.ms[kclass];
  return WA_OK;
}

}  // extern "C"
