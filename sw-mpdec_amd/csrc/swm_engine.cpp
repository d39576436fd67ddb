// swm_engine.cpp -- host side of the solve engine and of the C ABI
// (include/swm.h): per-job device memory, tile tables, the lockstep step of
// each time scheme over all running jobs of that scheme, the device-decided
// Jacobi loop, admission against a device-memory budget, and HIP-event
// kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "swm.h"
#include "swm_engine.hpp"

using namespace swm;

namespace {
thread_local std::string g_error;
}  // namespace

// the calling thread's message for swm_last_error(NULL)
void swm::set_thread_error(const std::string& msg) { g_error = msg; }

namespace {

size_t align_up(size_t b) { return (b + 255) & ~size_t(255); }
int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

const char* const PROF_NAMES[SWM_PROF_NCAT] = {
    "fvtd_x", "fvtd_y", "fvtd_update", "momentum", "patankar_assemble", "jacobi_sweep",
    "timestep", "other"};

// HIP-event timing of kernel launches by category.
struct Profiler {
  unsigned mask = 0;  // bit c: record the launches of category c; 0: off
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<int> cats;

  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  ~Profiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

// Bytes of every device array of one job, in allocation order.
struct JobLayout {
  size_t n, nfx, nfy;
  int nblk;
  size_t bytes;
};

JobLayout job_layout(int nx, int ny, int wb, int bathy) {
  JobLayout L;
  L.n = (size_t)nx * ny;
  L.nfx = (size_t)(nx + 1) * ny;
  L.nfy = (size_t)nx * (ny + 1);
  L.nblk = cdiv((long)L.n, kCellBlock);
  size_t b = 0;
  auto add = [&](size_t cnt) { b += align_up(cnt * sizeof(double)); };
  for (int k = 0; k < 3; ++k) add(L.n);                  // U^n
  for (int k = 0; k < 3 * 5; ++k) add(L.n);              // nodes 1..3: hres, hun, hvn, 2 x ping-pong
  for (int k = 0; k < 4 * 2; ++k) add(L.n);              // FU(2:3) per node
  for (int k = 0; k < 4; ++k) { add(L.nfx); add(L.nfy); }  // qx, qy per node
  for (int k = 0; k < kSlots * 3; ++k) { add(L.nfx); add(L.nfy); }  // face flux scratch
  for (int k = 0; k < kMaxSys * 5; ++k) add(L.n);        // Patankar matrices
  if (wb) {
    for (int k = 0; k < 3; ++k) { add(L.nfx); add(L.nfy); }
    add(L.n);
    add(L.n);
  }
  add((size_t)(kMaxSys + 1) * L.nblk);                   // Jacobi partial sums
  add(3 * (size_t)L.nblk);                               // TimeStep partials
  L.bytes = b;
  (void)bathy;  // the bathymetry tables are a separate allocation (bathy_bytes)
  return L;
}

size_t bathy_bytes(size_t n) { return 32 * n * sizeof(double); }

// A growable device array (owning, move-only).
template <class T>
struct DevVec {
  T* p = nullptr;
  size_t cap = 0;
  DevVec() = default;
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
  DevVec(DevVec&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  DevVec& operator=(DevVec&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  hipError_t reserve(size_t cnt) {
    if (cnt <= cap) return hipSuccess;
    size_t c = std::max(cnt, cap * 2 + 16);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void**)&p, c * sizeof(T));
    if (e == hipSuccess) cap = c;
    return e;
  }
  ~DevVec() {
    if (p) (void)hipFree(p);
  }
};

struct Tables {
  std::vector<TileRef> xt[kNumShapes], yt[kNumShapes], ct, cb;  // face tiles per shape
  std::vector<int> jl;
  DevVec<TileRef> dxt[kNumShapes], dyt[kNumShapes], dct, dcb;
  DevVec<int> djl;
  int nxt[kNumShapes] = {}, nyt[kNumShapes] = {}, nct = 0, ncb = 0, njobs = 0;
};

enum JobState { JOB_RUNNING = 0, JOB_DONE = 1, JOB_FAILED = 2, JOB_FREE = 3 };

struct Job {
  JobDev dev{};        // host copy of the device descriptor
  JobLayout lay{};
  char* slab = nullptr;
  double* bathy_buf = nullptr;  // 32 n doubles, allocated when tables are set
  int state = JOB_FREE;
  int steps_left = 0;   // < 0: run until told otherwise (plain solver)
  int steps_done = 0;
  double t = 0.0;
  long long jacobi_sweeps = 0;
  int nonconv = 0;
  int last_iters[13] = {0};
  int pred[13];
  bool have_pred = false;  // pred holds the sweep counts of an accepted step
  const SchemeDef* pred_sc = nullptr;  // ... of this scheme
  std::vector<StepRec>* log = nullptr;  // receives each committed step while set
  bool wb_ready = false;
  const SchemeDef* scheme = &kSchemes[0];  // time scheme of engine / multi-step runs
  int hp_assoc = SWM_HP_AS_REFERENCE;     // depth view of the two-node schemes
  Tables own;          // tables covering this job alone
  double* vout_buf = nullptr;  // the writer's state (dev.vout), when kept
  bool vout_valid = false;     // vout holds the state of the last step
};

// Jobs step together when they share a scheme and, for the two-node schemes,
// the depth view: group 0 mPDeC5, 1-2 mPEuler, 3-4 mPDeC2 (reference view,
// by-node view), 5-8 the explicit schemes 1-4.
constexpr int kGroups = 9;
// autotuner candidates: each tile shape with (graph 0, margin 1), then
// (graph, margin) = (1,1), (0,1), (1,2), (0,2) with the faster shape; each
// set is timed in kTuneRounds interleaved rounds
constexpr int kTuneExec = 4;
constexpr int kTuneCands = kNumShapes + kTuneExec;
constexpr int kTuneRounds = 2;
constexpr int kTunePhases = kTuneRounds * kTuneCands;
int group_of(const SchemeDef* d, int hp_assoc) {
  const int si = (int)(d - kSchemes);
  if (si == 0) return 0;
  if (si <= 2) return 1 + 2 * (si - 1) + (hp_assoc == SWM_HP_BY_NODE ? 1 : 0);
  return 5 + (si - 3);
}

}  // namespace

struct swm_engine {
  long long budget = 0;     // <= 0: no limit
  long long in_use = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<Job>> jobs;   // index = job id
  DevVec<JobDev> djobs;
  DevVec<JobStatus> dstatus;
  JobStatus* hstatus = nullptr;  // pinned mirror
  size_t hstatus_cap = 0;
  double* hdt = nullptr;         // pinned: host-provided dt of a step
  int* hzero = nullptr;          // pinned: 0, for clearing dt_bad
  int gated = 0;                 // 1 while the launches of a batch are issued
  bool dirty = true;             // device descriptors / tables out of date
  Tables groups[kGroups];        // running jobs, by step group (group_of)
  int max_iter = 1000;
  double tol = 1.e-16;
  long long ambiguous = 0;
  bool track_min = false;        // fold every stage depth into JobStatus::hmin
  // Step execution (results are identical either way; only speed differs):
  // speculative steps launch each correction's predicted sweeps without host
  // reads and check once at the end; they may be replayed from a captured
  // HIP graph.  -1 = chosen by the autotuner.
  int use_graph = -1, margin = -1;
  int tile_shape = -1;           // face-kernel tile shape (kShapes), -1 = autotuned
  int tuned_graph = 1, tuned_margin = 1, tuned_shape = 0;
  int cur_shape = 0;             // the shape the launches use now
  int tune_phase = 0;            // autotuner: timed steps so far
  double tune_ms[kTuneCands];    // per candidate: ms per million cell-steps, the fastest round
  int tune_skips = 0;            // steps of the current candidate not timed (graph capture)
  long long graph_epoch = 0;     // bumped whenever captured arguments change
  int keep_count = 0;            // jobs that keep the writer's state (k_keep_state launched)
  long long spec_steps = 0, redone_steps = 0, graph_launches = 0, graph_captures = 0;
  long long status_reads = 0;    // host waits for the device status (read_status)
  struct Captured {
    long long epoch = -1;
    int device_dt = 0, track = 0;
    int sweeps[5] = {0, 0, 0, 0, 0};
    const void* tables = nullptr;
    int group = -1, shape = -1;
    long long used = 0;
    hipGraphExec_t exec = nullptr;
  };
  Captured caps[kGroups];        // one captured step per table set in use
  long long cap_clock = 0;
  Profiler prof;
  std::string err;
};

struct swm_solver {
  swm_engine* e = nullptr;
  int job = -1;
  // asynchronous output (swm_download_primitive_async): a device snapshot of
  // the primitives, copied to the host on a stream of its own
  double* out_stage = nullptr;
  hipStream_t out_stream = nullptr;
  hipEvent_t out_ready = nullptr, out_done = nullptr;
  bool out_pending = false;
};

#define FAIL(E, code, ...)                                                     \
  do {                                                                         \
    char _b[512];                                                              \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                                     \
    if (E) (E)->err = _b;                                                      \
    g_error = _b;                                                              \
    return code;                                                               \
  } while (0)

#define HIPCK(E, call)                                                         \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess)                                                      \
      FAIL(E, _e == hipErrorOutOfMemory ? SWM_E_NOMEM : SWM_E_HIP, "%s failed: %s", #call, \
           hipGetErrorString(_e));                                             \
  } while (0)

#define LAUNCHCK(E)                                                            \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) FAIL(E, SWM_E_HIP, "kernel launch failed: %s", hipGetErrorString(_e)); \
  } while (0)

namespace {

template <class F>
void timed(swm_engine* e, int cat, F&& launch) {
  if (!((e->prof.mask >> cat) & 1u)) {
    launch();
    return;
  }
  hipEvent_t a = e->prof.next(), b = e->prof.next();
  if (a) (void)hipEventRecord(a, e->stream);
  launch();
  if (b) (void)hipEventRecord(b, e->stream);
  if (a && b)
    e->prof.cats.push_back(cat);
  else
    e->prof.used = 2 * e->prof.cats.size();
}

void add_tiles(Tables& T, int id, const JobDev& d) {
  for (int sh = 0; sh < kNumShapes; ++sh) {
    const FaceShape& F = kShapes[sh];
    const int nxt = cdiv(d.nx + 1, F.xw - 1) * cdiv(d.ny, F.xh);
    const int nyt = cdiv(d.nx, F.yw) * cdiv(d.ny + 1, F.yh - 1);
    for (int t = 0; t < nxt; ++t) T.xt[sh].push_back({id, t});
    for (int t = 0; t < nyt; ++t) T.yt[sh].push_back({id, t});
  }
  const int nct = cdiv(d.nx, kTile) * cdiv(d.ny, kTile);
  for (int t = 0; t < nct; ++t) T.ct.push_back({id, t});
  for (int t = 0; t < d.nblk; ++t) T.cb.push_back({id, t});
  T.jl.push_back(id);
}

int upload_tables(swm_engine* e, Tables& T) {
  for (int sh = 0; sh < kNumShapes; ++sh) {
    T.nxt[sh] = (int)T.xt[sh].size();
    T.nyt[sh] = (int)T.yt[sh].size();
    HIPCK(e, T.dxt[sh].reserve(T.xt[sh].size()));
    HIPCK(e, T.dyt[sh].reserve(T.yt[sh].size()));
    if (T.nxt[sh])
      HIPCK(e, hipMemcpy(T.dxt[sh].p, T.xt[sh].data(), T.nxt[sh] * sizeof(TileRef),
                         hipMemcpyHostToDevice));
    if (T.nyt[sh])
      HIPCK(e, hipMemcpy(T.dyt[sh].p, T.yt[sh].data(), T.nyt[sh] * sizeof(TileRef),
                         hipMemcpyHostToDevice));
  }
  T.nct = (int)T.ct.size();
  T.ncb = (int)T.cb.size();
  T.njobs = (int)T.jl.size();
  HIPCK(e, T.dct.reserve(T.ct.size()));
  HIPCK(e, T.dcb.reserve(T.cb.size()));
  HIPCK(e, T.djl.reserve(T.jl.size()));
  if (T.nct) HIPCK(e, hipMemcpy(T.dct.p, T.ct.data(), T.nct * sizeof(TileRef), hipMemcpyHostToDevice));
  if (T.ncb) HIPCK(e, hipMemcpy(T.dcb.p, T.cb.data(), T.ncb * sizeof(TileRef), hipMemcpyHostToDevice));
  if (T.njobs) HIPCK(e, hipMemcpy(T.djl.p, T.jl.data(), T.njobs * sizeof(int), hipMemcpyHostToDevice));
  return SWM_OK;
}

// Device job descriptors and the all-running-jobs tables.
int refresh(swm_engine* e) {
  if (!e->dirty) return SWM_OK;
  HIPCK(e, hipStreamSynchronize(e->stream));
  const size_t nj = e->jobs.size();
  HIPCK(e, e->djobs.reserve(std::max<size_t>(nj, 1)));
  std::vector<JobDev> host(nj);
  for (size_t i = 0; i < nj; ++i) host[i] = e->jobs[i]->dev;
  if (nj) HIPCK(e, hipMemcpy(e->djobs.p, host.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice));
  for (Tables& T : e->groups) {
    for (int sh = 0; sh < kNumShapes; ++sh) {
      T.xt[sh].clear();
      T.yt[sh].clear();
    }
    T.ct.clear();
    T.cb.clear();
    T.jl.clear();
  }
  for (size_t i = 0; i < nj; ++i) {
    const Job& J = *e->jobs[i];
    if (J.state == JOB_RUNNING) add_tiles(e->groups[group_of(J.scheme, J.hp_assoc)], (int)i, J.dev);
  }
  for (Tables& T : e->groups) {
    int rc = upload_tables(e, T);
    if (rc) return rc;
  }
  e->dirty = false;
  ++e->graph_epoch;  // job descriptors and tables may have moved
  return SWM_OK;
}

EngineArgs eargs(swm_engine* e) {
  EngineArgs a;
  a.jobs = e->djobs.p;
  a.max_iter = e->max_iter;
  a.gated = e->gated;
  a.tol = e->tol;
  return a;
}

int fvtd_launch(swm_engine* e, Tables& T, const NodeList& L) {
  EngineArgs a = eargs(e);
  hipStream_t s = e->stream;
  const int sh = e->cur_shape;
  timed(e, 0, [&] { launch_fvtd(a, sh, T.dxt[sh].p, T.nxt[sh], nullptr, 0, nullptr, 0, L, s); });
  timed(e, 1, [&] { launch_fvtd(a, sh, nullptr, 0, T.dyt[sh].p, T.nyt[sh], nullptr, 0, L, s); });
  timed(e, 2, [&] { launch_fvtd(a, sh, nullptr, 0, nullptr, 0, T.dct.p, T.nct, L, s); });
  LAUNCHCK(e);
  return SWM_OK;
}

int read_status(swm_engine* e) {
  ++e->status_reads;
  const size_t nj = e->jobs.size();
  HIPCK(e, hipMemcpyAsync(e->hstatus, e->dstatus.p, nj * sizeof(JobStatus),
                          hipMemcpyDeviceToHost, e->stream));
  HIPCK(e, hipStreamSynchronize(e->stream));
  return SWM_OK;
}

// ---------------------------------------------------------------------------
// One step of a modified-Patankar DeC scheme of every job in T: mPDeC5
// (src/timediscretization.f90:458-566), mPEuler (:671-764) or mPDeC2
// (:571-664), which differ only in MSteps, KCorr and thetaDeC.
//
// The launch sequence: [TimeStep] -> FVTD(U^n) -> KCorr corrections, each
// [depth view] -> momentum -> assembly (+ first sweep) -> Jacobi sweeps ->
// record -> finalize -> FVTD of the new nodes (all corrections but the last)
// -> (after the host check) commit.  Each sweep is followed by the
// device-side stop decision, and a sweep of a stopped system is a no-op, so
// launching more sweeps than a solve needs changes nothing.
//
// Polling mode reads the status back during each correction until every
// system has stopped.  Speculative mode launches a fixed number of sweeps per
// correction (the last step's counts plus a margin) and reads the status once,
// before the commit; if any solve needed more sweeps ("overrun") U^n is still
// intact and the step is run again in polling mode.  A speculative launch
// sequence depends only on the job set and the sweep counts, so it is captured
// once as a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct StepPlan {
  const SchemeDef* sc;
  int hp_view;    // the reference's depth view of the two-node schemes
  bool device_dt;
  bool speculative;
  int sweeps[5];  // speculative: sweeps per correction, counting the fused one
};

// Systems solved in correction kk (1-based): nodes 1..MSteps-1, or only the
// last node in the last correction.
int systems_in(const SchemeDef& D, int kk) { return kk == D.kcorr ? 1 : D.msteps - 1; }

// ---------------------------------------------------------------------------
// The explicit schemes of the build without PATANKAR
// (src/timediscretization.f90:104-451): FVTD launches that write Ut(1:3) into
// derivative slots, and stage combinations (k_combine) with the reference's
// weights and evaluation order.  States live in slots 0..5 (0 = U^n); the
// result is left in the scheme's final_state slot for the commit.
// ---------------------------------------------------------------------------
struct Term {
  int deriv;  // 1: derivative slot, 0: state slot
  int src;
  double c;
};

CombOp comb_op(int mode, int out, int base, std::initializer_list<Term> terms) {
  CombOp op{};
  op.mode = mode;
  op.nout = 1;
  op.out[0] = out;
  op.base = base;
  for (const Term& t : terms) {
    op.deriv[op.nterm] = t.deriv;
    op.src[op.nterm] = t.src;
    op.c[0][op.nterm] = t.c;
    ++op.nterm;
  }
  return op;
}

// thetaDeC of DeC5 (src/timediscretization.f90:389-404; the same literals as
// mPDeC5's), [m][r] = thetaDeC(m+1, r+1)
constexpr double kThetaDeC5[4][4] = {
    {0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000},
    {0.1103005664791649049, 0.1896994335208351257, -0.0339073642291439076, 0.0103005664791649201},
    {0.0730327668541684294, 0.4505740308958107176, 0.2269672331458314485, -0.0269672331458315692},
    {0.0833333333333333287, 0.4166666666666666852, 0.4166666666666666852, 0.0833333333333333287}};

// RK65 (src/timediscretization.f90:264-269): ARK(i,j) of the column-major
// reshape as [i-1][j-1], and bRK
constexpr double kARK65[6][5] = {
    {0., 0., 0., 0., 0.},
    {0.25, 0., 0., 0., 0.},
    {0.125, 0.125, 0., 0., 0.},
    {0., 0., 0.5, 0., 0.},
    {0.1875, -0.375, 0.375, 0.5625, 0.},
    {-0.42857142857142855, 1.1428571428571428, 0.8571428571428571, -1.7142857142857142,
     1.1428571428571428}};
constexpr double kBRK65[6] = {0.077777777777777778, 0., 0.35555555555555556,
                              0.13333333333333333, 0.35555555555555556, 0.077777777777777778};

int explicit_body(swm_engine* e, Tables& T, const SchemeDef& D) {
  EngineArgs a = eargs(e);
  hipStream_t s = e->stream;
  // FVTimeDerivative of state slots into derivative slots, Ut(1:3)
  auto fvtd = [&](std::initializer_list<std::pair<int, int>> in_out) {
    NodeList L{};
    L.ut1 = 1;
    for (const auto& io : in_out) {
      L.m[L.count] = io.first;
      L.out[L.count] = io.second;
      L.slot[L.count] = L.count;
      ++L.count;
    }
    return fvtd_launch(e, T, L);
  };
  auto comb = [&](const CombOp& op) {
    timed(e, 3, [&] { launch_combine(a, T.dcb.p, T.ncb, op, s); });
  };
  // the writer's state: the input of the step's last FVTD, before a stage
  // combination overwrites its slot
  auto keep = [&](int m) {
    if (e->keep_count) launch_keep_state(a, T.dcb.p, T.ncb, m, s);
  };
  constexpr int X = 0, F = 1;
  int rc = SWM_OK;
  switch (D.timescheme) {
    case 1:  // ForwardEuler (:198-232): U = K0 + Ut*dt
      if ((rc = fvtd({{0, 0}}))) return rc;
      keep(0);
      comb(comb_op(0, 1, 0, {{X, 0, 1.0}, {F, 0, 1.0}}));
      break;
    case 2:  // SSPRK4 (:104-192)
      if ((rc = fvtd({{0, 0}}))) return rc;
      comb(comb_op(0, 1, 0, {{X, 0, 1.00000000000000}, {F, 0, 0.39175222700392}}));
      if ((rc = fvtd({{1, 0}}))) return rc;
      comb(comb_op(0, 2, 0,
                   {{X, 0, 0.44437049406734}, {X, 1, 0.55562950593266}, {F, 0, 0.36841059262959}}));
      if ((rc = fvtd({{2, 0}}))) return rc;
      comb(comb_op(0, 3, 0,
                   {{X, 0, 0.62010185138540}, {X, 2, 0.37989814861460}, {F, 0, 0.25189177424738}}));
      if ((rc = fvtd({{3, 0}}))) return rc;
      comb(comb_op(0, 4, 0,
                   {{X, 0, 0.17807995410773}, {X, 3, 0.82192004589227}, {F, 0, 0.54497475021237}}));
      // K5 = Ut of stage 4 stays in derivative slot 0; stage 5's goes to 1
      if ((rc = fvtd({{4, 1}}))) return rc;
      keep(4);
      comb(comb_op(0, 1, 0,
                   {{X, 0, 0.00683325884039}, {X, 2, 0.51723167208978}, {X, 3, 0.12759831133288},
                    {X, 4, 0.34833675773694}, {F, 0, 0.08460416338212},
                    {F, 1, 0.22600748319395}}));
      break;
    case 3: {  // RK65 (:238-356): K_j in derivative slot j, stage states in slot 1
      if ((rc = fvtd({{0, 0}}))) return rc;
      for (int st = 1; st <= 5; ++st) {
        CombOp op = comb_op(1, 1, 0, {});
        for (int j = 0; j < st; ++j) {
          op.deriv[op.nterm] = F;
          op.src[op.nterm] = j;
          op.c[0][op.nterm] = kARK65[st][j];
          ++op.nterm;
        }
        comb(op);
        if ((rc = fvtd({{1, st}}))) return rc;
      }
      keep(1);
      CombOp op = comb_op(1, 1, 0, {});
      for (int j = 0; j < 6; ++j) {
        op.deriv[op.nterm] = F;
        op.src[op.nterm] = j;
        op.c[0][op.nterm] = kBRK65[j];
        ++op.nterm;
      }
      comb(op);
      break;
    }
    case 4: {  // DeC5 (:362-451): the mPDeC5 loop on all three variables
      if ((rc = fvtd({{0, 0}}))) return rc;
      int src[4] = {0, 0, 0, 0};  // FUp(r): FVTD(U^n) for every r at kk = 1
      for (int kk = 1; kk <= 5; ++kk) {
        // the correction's nodes in one launch: they share Ua(1) and FUp
        CombOp op = comb_op(2, 0, 0, {});
        op.nout = 0;
        for (int m = (kk == 5 ? 3 : 1); m < 4; ++m) {
          op.out[op.nout] = m;
          for (int r = 0; r < 4; ++r) op.c[op.nout][r] = kThetaDeC5[m][r];
          ++op.nout;
        }
        for (int r = 0; r < 4; ++r) {
          op.deriv[r] = F;
          op.src[r] = src[r];
        }
        op.nterm = 4;
        comb(op);
        if (kk != 5) {
          if ((rc = fvtd({{1, 1}, {2, 2}, {3, 3}}))) return rc;
          if (kk == 4) keep(3);
          for (int r = 0; r < 4; ++r) src[r] = r;
        }
      }
      break;
    }
    default:
      FAIL(e, SWM_E_ARG, "internal: timescheme %d is not explicit", D.timescheme);
  }
  LAUNCHCK(e);
  return SWM_OK;
}

// The step's launches; returns after the last finalize / FVTD (no commit).
int step_body(swm_engine* e, Tables& T, const StepPlan& P) {
  EngineArgs a = eargs(e);
  hipStream_t s = e->stream;
  if (P.device_dt) {
    timed(e, 6, [&] { launch_timestep(a, T.dcb.p, T.ncb, T.djl.p, T.njobs, s); });
    LAUNCHCK(e);
  }
  const SchemeDef& D = *P.sc;
  if (is_explicit(D)) return explicit_body(e, T, D);
  // FVTimeDerivative(U^n): node 0, shared by all nodes at kk = 1
  NodeList L0{};
  L0.count = 1;
  L0.m[0] = 0;
  L0.out[0] = 0;
  L0.slot[0] = 0;
  int rc = fvtd_launch(e, T, L0);
  if (rc) return rc;
  // the writer's state: the input of the step's last FVTD (U^n when the
  // scheme has a single correction)
  if (e->keep_count && D.kcorr == 1) launch_keep_state(a, T.dcb.p, T.ncb, 0, s);
  SysList S{};
  for (int r = 0; r < 4; ++r) S.src[r] = 0;
  S.hp_is_un = 1;
  S.track_min = e->track_min ? 1 : 0;
  S.nr = D.msteps;
  S.theta = D.theta;
  S.hp_view = P.hp_view;
  int solve = 0;
  for (int kk = 1; kk <= D.kcorr; ++kk) {
    S.count = 0;
    if (kk == D.kcorr)
      S.m[S.count++] = D.msteps - 1;
    else
      for (int m = 1; m < D.msteps; ++m) S.m[S.count++] = m;
    if (S.hp_view) timed(e, 4, [&] { launch_hp_view(a, T.dcb.p, T.ncb, kk == 1, s); });
    // the DeC momentum update runs inside the assembly kernel
    timed(e, 4, [&] { launch_assemble(a, T.dcb.p, T.ncb, T.djl.p, T.njobs, S, s); });
    LAUNCHCK(e);
    if (P.speculative) {
      for (int b = 1; b < P.sweeps[kk - 1]; ++b)
        timed(e, 5, [&] { launch_sweep(a, T.dcb.p, T.ncb, T.djl.p, T.njobs, S, s); });
      LAUNCHCK(e);
    } else {
      int pred = 1;
      for (int q = 0; q < T.njobs; ++q)
        for (int k = 0; k < S.count; ++k)
          pred = std::max(pred, e->jobs[T.jl[q]]->pred[solve + k]);
      int batch = std::max(0, pred - 1);
      long launched = 1;
      for (;;) {
        for (int b = 0; b < batch; ++b)
          timed(e, 5, [&] { launch_sweep(a, T.dcb.p, T.ncb, T.djl.p, T.njobs, S, s); });
        LAUNCHCK(e);
        launched += batch;
        rc = read_status(e);
        if (rc) return rc;
        int kmin = e->max_iter;
        bool any = false;
        for (int q = 0; q < T.njobs; ++q) {
          const JobStatus& js = e->hstatus[T.jl[q]];
          if (js.halt) continue;  // not in this batch's step
          for (int k = 0; k < S.count; ++k)
            if (js.st[k].active) {
              any = true;
              kmin = std::min(kmin, js.st[k].k);
            }
        }
        if (!any) break;
        if (launched > (long)e->max_iter + 2)
          FAIL(e, SWM_E_STATE, "internal: Jacobi did not stop after %ld sweeps", launched);
        batch = std::max(2, kmin / 4);
      }
    }
    timed(e, 7, [&] { launch_solve_done(a, T.djl.p, T.njobs, S, solve, s); });
    timed(e, 7, [&] { launch_finalize(a, T.dcb.p, T.ncb, T.djl.p, T.njobs, S, s); });
    LAUNCHCK(e);
    solve += S.count;
    if (kk != D.kcorr) {
      NodeList L{};
      L.count = S.count;
      for (int k = 0; k < S.count; ++k) {
        L.m[k] = S.m[k];
        L.out[k] = S.m[k];
        L.slot[k] = k;
      }
      rc = fvtd_launch(e, T, L);
      if (rc) return rc;
      if (e->keep_count && kk == D.kcorr - 1)
        launch_keep_state(a, T.dcb.p, T.ncb, D.msteps - 1, s);
      for (int r = 0; r < 4; ++r) S.src[r] = r;
      S.hp_is_un = 0;
    }
  }
  return SWM_OK;
}

// The step's launches, then the commit decision (k_gate) and the commit.
int step_full(swm_engine* e, Tables& T, const StepPlan& P) {
  int rc = step_body(e, T, P);
  if (rc) return rc;
  EngineArgs a = eargs(e);
  hipStream_t s = e->stream;
  timed(e, 7, [&] { launch_gate(a, T.djl.p, T.njobs, P.sc->nsolve, s); });
  timed(e, 7, [&] { launch_commit(a, T.dcb.p, T.ncb, P.sc->final_state, s); });
  LAUNCHCK(e);
  return SWM_OK;
}

void drop_graph(swm_engine::Captured& c) {
  if (c.exec) (void)hipGraphExecDestroy(c.exec);
  c = swm_engine::Captured{};
}

void drop_graphs(swm_engine* e) {
  for (auto& c : e->caps) drop_graph(c);
}

// Speculative step (body, gate, commit) through a captured graph (capturing it when the job set,
// the scheme or the sweep counts changed).  Each table set keeps its own
// graph, so groups of jobs stepped one after another do not recapture.
int step_body_graph(swm_engine* e, Tables& T, const StepPlan& P) {
  const int group = group_of(P.sc, P.hp_view ? SWM_HP_AS_REFERENCE : SWM_HP_BY_NODE);
  swm_engine::Captured* cp = nullptr;
  for (auto& c : e->caps)
    if (c.exec && c.tables == (const void*)&T) cp = &c;
  if (!cp) {  // a free entry, else the least recently used
    cp = &e->caps[0];
    for (auto& c : e->caps)
      if (!c.exec || (cp->exec && c.used < cp->used)) cp = &c;
  }
  auto& c = *cp;
  bool same = c.exec && c.epoch == e->graph_epoch && c.device_dt == (int)P.device_dt &&
              c.track == (int)e->track_min && c.tables == (const void*)&T && c.group == group &&
              c.shape == e->cur_shape;
  for (int i = 0; same && i < 5; ++i) same = c.sweeps[i] == P.sweeps[i];
  c.used = ++e->cap_clock;
  if (!same) {
    drop_graph(c);
    c.used = e->cap_clock;
    HIPCK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = step_full(e, T, P);
    hipGraph_t g = nullptr;
    hipError_t er = hipStreamEndCapture(e->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (er != hipSuccess) FAIL(e, SWM_E_HIP, "graph capture failed: %s", hipGetErrorString(er));
    hipGraphExec_t ge = nullptr;
    er = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (er != hipSuccess) FAIL(e, SWM_E_HIP, "graph instantiate failed: %s", hipGetErrorString(er));
    c.exec = ge;
    c.epoch = e->graph_epoch;
    c.device_dt = P.device_dt;
    c.track = e->track_min;
    c.tables = &T;
    c.group = group;
    c.shape = e->cur_shape;
    for (int i = 0; i < 5; ++i) c.sweeps[i] = P.sweeps[i];
    ++e->graph_captures;
  }
  HIPCK(e, hipGraphLaunch(c.exec, e->stream));
  ++e->graph_launches;
  return SWM_OK;
}

// Autotuner over (tile shape, graph, margin).  The first speculative steps of
// an engine time the candidates one step each, in interleaved rounds (so that
// a drift of the clocks during tuning favours none of them; a candidate's
// fastest round counts): the face-kernel tile shapes first, without a graph
// and by the face kernels' own time (HIP events), then graph on/off x margin
// 1/2 with the faster shape, by the step's wall time; the fastest of the
// latter is kept.  A step that captured a graph is not timed (the candidate
// is timed on its next step, a replay, unless it captured twice in a row),
// and each time is divided by the cell-steps of the job set it ran on, so
// that candidates timed while jobs join and leave compare fairly.
// Explicit settings (swm_engine_set_execution, swm_engine_set_tile_shape)
// override their dimension.
struct TuneCand {
  int graph, margin, shape;
};

// Candidate id of tuning step `phase`: kTuneRounds rounds over the shapes,
// then kTuneRounds rounds over the execution settings.
int tune_id(int phase) {
  if (phase < kTuneRounds * kNumShapes) return phase % kNumShapes;
  return kNumShapes + (phase - kTuneRounds * kNumShapes) % kTuneExec;
}

TuneCand tune_cand(const swm_engine* e, int id) {
  if (id < kNumShapes) return {0, 1, id};
  int best = 0;  // the faster shape
  for (int i = 1; i < kNumShapes; ++i)
    if (e->tune_ms[i] < e->tune_ms[best]) best = i;
  static const int cg[kTuneExec] = {1, 0, 1, 0}, cm[kTuneExec] = {1, 1, 2, 2};
  return {cg[id - kNumShapes], cm[id - kNumShapes], best};
}

bool tuning(const swm_engine* e) {
  return e->tune_phase < kTunePhases && (e->use_graph < 0 || e->margin < 0 || e->tile_shape < 0);
}

void tune_choice(swm_engine* e, int& graph, int& margin) {
  TuneCand c = tuning(e) ? tune_cand(e, tune_id(e->tune_phase))
                         : TuneCand{e->tuned_graph, e->tuned_margin, e->tuned_shape};
  graph = e->use_graph >= 0 ? e->use_graph : c.graph;
  margin = e->margin >= 0 ? e->margin : c.margin;
  e->cur_shape = e->tile_shape >= 0 ? e->tile_shape : c.shape;
}

void tune_record(swm_engine* e, double ms) {
  if (!tuning(e)) return;
  double& t = e->tune_ms[tune_id(e->tune_phase)];
  t = std::min(t, ms);
  if (++e->tune_phase == kTunePhases) {
    int best = kNumShapes;
    for (int i = kNumShapes + 1; i < kTuneCands; ++i)
      if (e->tune_ms[i] < e->tune_ms[best]) best = i;
    const TuneCand c = tune_cand(e, best);
    e->tuned_graph = c.graph;
    e->tuned_margin = c.margin;
    e->tuned_shape = c.shape;
  }
}

// Consumes the records of a batch: every committed step of every job of T
// (the job's counters, its sweep-count prediction and, while set, its log);
// with `only`, just the jobs it marks.
void consume_batch(swm_engine* e, Tables& T, const SchemeDef* sc,
                   const std::vector<char>* only = nullptr) {
  for (int q = 0; q < T.njobs; ++q) {
    if (only && !(*only)[q]) continue;
    Job& J = *e->jobs[T.jl[q]];
    const JobStatus& js = e->hstatus[T.jl[q]];
    for (int r = 0; r < js.batch_done; ++r) {
      const StepRec& rec = js.rec[r];
      J.steps_done += 1;
      J.t += rec.dt;
      if (J.steps_left > 0) --J.steps_left;
      for (int i = 0; i < sc->nsolve; ++i) J.jacobi_sweeps += rec.iters[i];
      J.nonconv += rec.nonconv;
      e->ambiguous += rec.ambiguous;
      if (J.log) J.log->push_back(rec);
    }
    if (js.batch_done > 0) {
      const StepRec& last = js.rec[js.batch_done - 1];
      for (int i = 0; i < 13; ++i) {
        J.last_iters[i] = last.iters[i];
        J.pred[i] = last.iters[i];
      }
      J.have_pred = true;
      J.pred_sc = sc;
      J.vout_valid = J.vout_buf != nullptr;  // kept by every step body
    }
  }
}

// The device time loop's clipping and stop test (src/timediscretization.f90:
// 58-59, 83-94) for a plain solve.
struct LoopArgs {
  double t, t_an, t_end;
};

// A batch: up to `steps` steps of every job of T with one read of the status
// at the end (the device decides, step by step, which jobs go on; see
// k_gate).  host_dt: one step with this dt for every job (steps = 1).
// Steps run speculatively when every job has a sweep-count prediction for the
// scheme, else one step in polling mode (the explicit schemes, which solve
// nothing, run every step of the batch without host reads).  A job whose
// speculative step overran its launched sweeps has that step redone in
// polling mode, alone, before the batch returns.
int engine_batch(swm_engine* e, Tables& T, const SchemeDef* sc, int hp_assoc,
                 const double* host_dt, int steps, const LoopArgs* loop = nullptr) {
  if (T.njobs == 0) return SWM_OK;
  hipStream_t s = e->stream;
  e->cur_shape = e->tile_shape >= 0 ? e->tile_shape : e->tuned_shape;  // tune_choice may vary it
  StepPlan P{};
  P.sc = sc;
  P.hp_view = (!is_explicit(*sc) && sc->msteps < 4 && hp_assoc == SWM_HP_AS_REFERENCE) ? 1 : 0;
  P.device_dt = host_dt == nullptr;
  bool spec = e->margin != -2 && !is_explicit(*sc);  // nothing to predict without solves
  for (int q = 0; q < T.njobs; ++q) {
    const Job& J = *e->jobs[T.jl[q]];
    spec = spec && J.have_pred && J.pred_sc == sc;
  }
  steps = std::max(1, std::min(steps, kMaxBatch));
  if (host_dt || (!spec && !is_explicit(*sc)) || (spec && tuning(e))) steps = 1;
  BatchArgs B{};
  B.max_steps = steps;
  if (loop) {
    B.loop = 1;
    B.t = loop->t;
    B.t_an = loop->t_an;
    B.t_end = loop->t_end;
  }
  e->gated = 1;
  struct Ungate {
    swm_engine* e;
    ~Ungate() { e->gated = 0; }
  } ungate{e};
  launch_batch_begin(eargs(e), T.djl.p, T.njobs, B, s);
  LAUNCHCK(e);
  if (host_dt) {
    *e->hdt = *host_dt;
    for (int q = 0; q < T.njobs; ++q) {
      JobStatus* st = e->jobs[T.jl[q]]->dev.status;
      HIPCK(e, hipMemcpyAsync(&st->dt, e->hdt, sizeof(double), hipMemcpyHostToDevice, s));
      HIPCK(e, hipMemcpyAsync(&st->dt_bad, e->hzero, sizeof(int), hipMemcpyHostToDevice, s));
    }
  }
  if (spec) {
    int graph, margin;
    tune_choice(e, graph, margin);
    P.speculative = true;
    int first = 0;
    for (int kk = 0; kk < sc->kcorr; ++kk) {
      int mx = 1;
      const int ns = systems_in(*sc, kk + 1);
      for (int q = 0; q < T.njobs; ++q)
        for (int k = 0; k < ns; ++k) mx = std::max(mx, e->jobs[T.jl[q]]->pred[first + k]);
      P.sweeps[kk] = std::min(mx + margin, e->max_iter);
      first += ns;
    }
    // a tile-shape candidate is timed by HIP events around the face kernels
    // (the only launches the shape changes), unless the caller is recording
    // events itself or has set graph replay on (events keep it off); the
    // other candidates by the step's wall time
    const bool face_events =
        tuning(e) && tune_id(e->tune_phase) < kNumShapes && !e->prof.mask && !graph;
    const size_t rec0 = e->prof.cats.size();  // records a caller has not read yet stay
    if (face_events) e->prof.mask = 0x3u;     // fvtd_x, fvtd_y
    const auto t0 = std::chrono::steady_clock::now();
    const long long captures0 = e->graph_captures;
    int rc = SWM_OK;
    for (int b = 0; b < steps && !rc; ++b)
      rc = (graph && !e->prof.mask) ? step_body_graph(e, T, P) : step_full(e, T, P);
    if (!rc) rc = read_status(e);
    double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (face_events) {
      e->prof.mask = 0;
      double face_ms = 0.;  // fvtd_x + fvtd_y of this step's records, which are then dropped
      for (size_t q = rec0; !rc && q < e->prof.cats.size(); ++q) {
        if (e->prof.cats[q] > 1) continue;
        float t = 0.f;
        const hipError_t er =
            hipEventElapsedTime(&t, e->prof.pool[2 * q], e->prof.pool[2 * q + 1]);
        if (er != hipSuccess) {
          e->err = std::string("hipEventElapsedTime failed: ") + hipGetErrorString(er);
          g_error = e->err;
          rc = SWM_E_HIP;
        }
        face_ms += t;
      }
      e->prof.cats.resize(rec0);
      e->prof.used = 2 * rec0;
      ms = face_ms;
    }
    if (rc) return rc;
    e->spec_steps += steps;
    bool overrun = false;
    for (int q = 0; q < T.njobs; ++q) overrun = overrun || e->hstatus[T.jl[q]].halt == kHaltOverrun;
    if (!overrun) {
      const bool captured = e->graph_captures != captures0;
      if (captured && e->tune_skips < 2) {
        ++e->tune_skips;
      } else {
        double cells = 0.;
        for (int q = 0; q < T.njobs; ++q) cells += (double)e->jobs[T.jl[q]]->dev.n;
        tune_record(e, ms / ((double)steps * cells * 1e-6));
        e->tune_skips = 0;
      }
    }
  } else {
    P.speculative = false;
    for (int b = 0; b < steps; ++b) {
      int rc = step_full(e, T, P);
      if (rc) return rc;
    }
    int rc = read_status(e);
    if (rc) return rc;
  }
  consume_batch(e, T, sc);
  std::vector<char> redo(T.njobs, 0);
  bool overrun = false;
  for (int q = 0; q < T.njobs; ++q) {
    redo[q] = e->hstatus[T.jl[q]].halt == kHaltOverrun;
    overrun = overrun || redo[q];
  }
  if (overrun) {
    // U^n of the overrun jobs is intact (their step was not committed): redo
    // that step in polling mode; every other job keeps its stop and is skipped
    B.redo = 1;
    B.max_steps = 1;
    launch_batch_begin(eargs(e), T.djl.p, T.njobs, B, s);
    LAUNCHCK(e);
    for (int q = 0; q < T.njobs; ++q) e->redone_steps += redo[q];
    P.speculative = false;
    int rc = step_full(e, T, P);
    if (rc) return rc;
    rc = read_status(e);
    if (rc) return rc;
    consume_batch(e, T, sc, &redo);
  }
  return SWM_OK;
}

int job_from_id(swm_engine* e, int id, Job** out) {
  if (id < 0 || id >= (int)e->jobs.size() || e->jobs[id]->state == JOB_FREE)
    FAIL(e, SWM_E_ARG, "no job with id %d", id);
  *out = e->jobs[id].get();
  return SWM_OK;
}

int upload_state(swm_engine* e, Job& J, const double* U, int ghosts, double* h, double* hu,
                 double* hv) {
  const int nx = J.dev.nx, ny = J.dev.ny;
  size_t cnt = 3 * (size_t)(nx + 2 * ghosts) * (ny + 2 * ghosts);
  double* tmp = nullptr;
  HIPCK(e, hipMalloc((void**)&tmp, cnt * sizeof(double)));
  hipError_t er = hipMemcpyAsync(tmp, U, cnt * sizeof(double), hipMemcpyHostToDevice, e->stream);
  if (er == hipSuccess) {
    launch_aos_to_soa(nx, ny, ghosts, tmp, h, hu, hv, e->stream);
    er = hipGetLastError();
  }
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (er != hipSuccess) FAIL(e, SWM_E_HIP, "upload failed: %s", hipGetErrorString(er));
  return SWM_OK;
}

// FVTimeDerivative of the well-balanced state held in node 1's arrays into
// FXWB/FYWB/SWB (InitializeWBVariables, src/finitevolume.f90:373-392).
int build_wb(swm_engine* e, int id) {
  Job& J = *e->jobs[id];
  int rc = refresh(e);
  if (rc) return rc;
  NodeList L{};
  L.count = 1;
  L.m[0] = 1;
  L.out[0] = 1;
  L.slot[0] = 0;
  L.wb_init = 1;
  rc = fvtd_launch(e, J.own, L);
  if (rc) return rc;
  launch_copy_wb(eargs(e), J.own.dcb.p, J.own.ncb, 0, e->stream);
  LAUNCHCK(e);
  HIPCK(e, hipStreamSynchronize(e->stream));
  J.wb_ready = true;
  return SWM_OK;
}

int init_wb(swm_engine* e, int id, const double* Uwb, int ghosts) {
  Job& J = *e->jobs[id];
  int rc = upload_state(e, J, Uwb, ghosts, J.dev.hres[1], J.dev.hun[1], J.dev.hvn[1]);
  if (rc) return rc;
  return build_wb(e, id);
}

int status_capacity(swm_engine* e, size_t nj) {
  if (nj <= e->hstatus_cap) return SWM_OK;
  size_t cap = std::max<size_t>(nj, 2 * e->hstatus_cap + 8);
  // the device status array is reallocated: jobs point into it
  HIPCK(e, hipStreamSynchronize(e->stream));
  JobStatus* oldp = e->dstatus.p;
  size_t oldcap = e->dstatus.cap;
  JobStatus* np = nullptr;
  HIPCK(e, hipMalloc((void**)&np, cap * sizeof(JobStatus)));
  HIPCK(e, hipMemset(np, 0, cap * sizeof(JobStatus)));
  if (oldp) HIPCK(e, hipMemcpy(np, oldp, oldcap * sizeof(JobStatus), hipMemcpyDeviceToDevice));
  if (oldp) (void)hipFree(oldp);
  e->dstatus.p = np;
  e->dstatus.cap = cap;
  for (size_t i = 0; i < e->jobs.size(); ++i) e->jobs[i]->dev.status = np + i;
  JobStatus* h = nullptr;
  HIPCK(e, hipHostMalloc((void**)&h, cap * sizeof(JobStatus), hipHostMallocDefault));
  if (e->hstatus) (void)hipHostFree(e->hstatus);
  e->hstatus = h;
  e->hstatus_cap = cap;
  e->dirty = true;
  return SWM_OK;
}

// Allocates a job's device memory and descriptors (no data yet).
int alloc_job(swm_engine* e, int nx, int ny, double dx, double dy, double g, double cfl,
              int wb, const double wbnd[4], int nsteps, long long extra_bytes, int* id_out) {
  *id_out = -1;
  if (nx < 3 || ny < 3) FAIL(e, SWM_E_ARG, "nx, ny must be >= 3 (got %d, %d)", nx, ny);
  if ((long long)nx * ny > (1LL << 30)) FAIL(e, SWM_E_ARG, "grid too large (%d x %d)", nx, ny);
  if (!(dx > 0) || !(dy > 0)) FAIL(e, SWM_E_ARG, "dx, dy must be > 0");
  if (!wbnd) FAIL(e, SWM_E_ARG, "wgp_bnd is NULL");
  JobLayout lay = job_layout(nx, ny, wb, 0);
  if (e->budget > 0 && e->in_use + (long long)lay.bytes + extra_bytes > e->budget)
    FAIL(e, SWM_E_BUDGET, "job of %lld bytes does not fit (%lld of %lld in use)",
         (long long)lay.bytes + extra_bytes, e->in_use, e->budget);
  int id = -1;
  for (size_t i = 0; i < e->jobs.size(); ++i)
    if (e->jobs[i]->state == JOB_FREE) {
      id = (int)i;
      break;
    }
  if (id < 0) {
    int rc = status_capacity(e, e->jobs.size() + 1);
    if (rc) return rc;
    e->jobs.emplace_back(new Job());
    id = (int)e->jobs.size() - 1;
    e->jobs[id]->dev.status = e->dstatus.p + id;
  }
  Job& J = *e->jobs[id];
  JobStatus* keep_status = J.dev.status;
  J.dev = JobDev{};
  J.dev.status = keep_status;
  J.own = Tables{};
  J.lay = lay;
  hipError_t er = hipMalloc((void**)&J.slab, lay.bytes);
  if (er != hipSuccess) {
    J.slab = nullptr;
    FAIL(e, er == hipErrorOutOfMemory ? SWM_E_NOMEM : SWM_E_HIP, "hipMalloc(%zu) failed: %s",
         lay.bytes, hipGetErrorString(er));
  }
  size_t off = 0;
  auto take = [&](size_t cnt) {
    double* p = reinterpret_cast<double*>(J.slab + off);
    off += align_up(cnt * sizeof(double));
    return p;
  };
  JobDev& D = J.dev;
  D.nx = nx;
  D.ny = ny;
  D.n = (int)lay.n;
  D.wb = wb;
  D.bathy = 0;
  D.nblk = lay.nblk;
  D.dx = dx;
  D.dy = dy;
  D.g = g;
  D.cfl = cfl;
  for (int i = 0; i < 4; ++i) D.wbnd[i] = wbnd[i];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) D.wgp[i][j] = wbnd[i] * wbnd[j];  // src/mesh.f90:136
  D.h = take(lay.n);
  D.hu = take(lay.n);
  D.hv = take(lay.n);
  for (int m = 1; m < 4; ++m) {
    D.hres[m] = take(lay.n);
    D.hun[m] = take(lay.n);
    D.hvn[m] = take(lay.n);
    D.xj[m][0] = take(lay.n);
    D.xj[m][1] = take(lay.n);
  }
  for (int r = 0; r < 4; ++r) {
    D.FU2[r] = take(lay.n);
    D.FU3[r] = take(lay.n);
  }
  for (int r = 0; r < 4; ++r) {
    D.qx[r] = take(lay.nfx);
    D.qy[r] = take(lay.nfy);
  }
  for (int sl = 0; sl < kSlots; ++sl)
    for (int v = 0; v < 3; ++v) {
      D.FX[sl][v] = take(lay.nfx);
      D.FY[sl][v] = take(lay.nfy);
    }
  for (int q = 0; q < kMaxSys; ++q)
    for (int k = 0; k < 5; ++k) D.M[q][k] = take(lay.n);
  // state and derivative slots of the explicit schemes (JobDev)
  D.sh[0] = D.h;
  D.shu[0] = D.hu;
  D.shv[0] = D.hv;
  for (int m = 1; m < 4; ++m) {
    D.sh[m] = D.hres[m];
    D.shu[m] = D.hun[m];
    D.shv[m] = D.hvn[m];
  }
  D.sh[4] = D.M[0][0];
  D.shu[4] = D.M[0][1];
  D.shv[4] = D.M[0][2];
  D.sh[5] = D.M[0][3];
  D.shu[5] = D.M[0][4];
  D.shv[5] = D.M[1][0];
  for (int r = 0; r < 4; ++r) {
    D.d1[r] = D.qx[r];
    D.d2[r] = D.FU2[r];
    D.d3[r] = D.FU3[r];
  }
  D.d1[4] = D.M[1][1];
  D.d2[4] = D.M[1][2];
  D.d3[4] = D.M[1][3];
  D.d1[5] = D.M[1][4];
  D.d2[5] = D.M[2][0];
  D.d3[5] = D.M[2][1];
  if (wb) {
    for (int v = 0; v < 3; ++v) {
      D.FXWB[v] = take(lay.nfx);
      D.FYWB[v] = take(lay.nfy);
    }
    D.SWB2 = take(lay.n);
    D.SWB3 = take(lay.n);
  }
  D.part = take((size_t)(kMaxSys + 1) * lay.nblk);
  D.tspart = take(3 * (size_t)lay.nblk);
  if (off > lay.bytes) FAIL(e, SWM_E_STATE, "internal: job slab overflow");
  e->in_use += (long long)lay.bytes;
  J.state = JOB_RUNNING;
  J.steps_left = nsteps;
  J.steps_done = 0;
  J.t = 0.0;
  J.jacobi_sweeps = 0;
  J.nonconv = 0;
  J.wb_ready = false;
  for (int i = 0; i < 13; ++i) J.pred[i] = 4;
  J.have_pred = false;
  J.pred_sc = nullptr;
  J.log = nullptr;
  J.scheme = &kSchemes[0];
  J.hp_assoc = SWM_HP_AS_REFERENCE;
  e->dirty = true;
  add_tiles(J.own, id, D);
  ++e->graph_epoch;
  int rc = upload_tables(e, J.own);
  if (rc) return rc;
  HIPCK(e, hipMemsetAsync(e->dstatus.p + id, 0, sizeof(JobStatus), e->stream));
  // the step-count stop of the device time loop (-1: none)
  launch_set_steps_left(e->dstatus.p + id, nsteps > 0 ? nsteps : -1, e->stream);
  LAUNCHCK(e);
  HIPCK(e, hipStreamSynchronize(e->stream));
  *id_out = id;
  return SWM_OK;
}

// JobDev::recfix for a ReconstructionFix value: 4 (= Reconstruction, the
// reference's setting) -> 0, 1 and 2 -> themselves, anything else -> -1.
int recfix_code(int recfix) {
  if (recfix == 4) return 0;
  if (recfix == 1 || recfix == 2) return recfix;
  return -1;
}

// The job's bathymetry tables, allocated on first use (within the budget).
int ensure_bathy(swm_engine* e, Job& J) {
  const size_t n = J.lay.n;
  if (!J.bathy_buf) {
    const long long b = (long long)bathy_bytes(n);
    if (e->budget > 0 && e->in_use + b > e->budget)
      FAIL(e, SWM_E_BUDGET, "bathymetry tables of %lld bytes do not fit", b);
    HIPCK(e, hipMalloc((void**)&J.bathy_buf, (size_t)b));
    e->in_use += b;
  }
  J.dev.bx = J.bathy_buf;
  J.dev.by = J.bathy_buf + 16 * n;
  return SWM_OK;
}

int job_set_bathy(swm_engine* e, int id, const double* bx, const double* by) {
  Job& J = *e->jobs[id];
  if (!bx != !by) FAIL(e, SWM_E_ARG, "dbdx and dbdy must both be set or both NULL");
  J.wb_ready = false;  // the well-balanced reference depends on the source term
  if (!bx) {
    J.dev.bathy = 0;
    e->dirty = true;
    return SWM_OK;
  }
  const size_t n = J.lay.n;
  int rc = ensure_bathy(e, J);
  if (rc) return rc;
  J.dev.bathy = 1;
  HIPCK(e, hipMemcpyAsync(J.dev.bx, bx, 16 * n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  HIPCK(e, hipMemcpyAsync(J.dev.by, by, 16 * n * sizeof(double), hipMemcpyHostToDevice, e->stream));
  HIPCK(e, hipStreamSynchronize(e->stream));
  e->dirty = true;
  return SWM_OK;
}

// The reference's initial data on the device (src/main.f90:22-26): the state
// of InitialCondition ic, the bathymetry tables of BathymetryFlag bflag (0:
// flat bottom) and, for a well-balanced job, the well-balanced state and its
// FVTD.  The domain [x0[0], x1[0]] x [x0[1], x1[1]] must give the job's
// Mesh_DX (src/mesh.f90:39-41) bitwise.
// ExactFunction's setup for a job: checks the arguments and that the domain
// gives the job's Mesh_DX bitwise.
int exact_setup(swm_engine* e, const Job& J, int ic, int bflag, const double* x0,
                const double* x1, double t, InitialSetup& c) {
  if (ic < 1 || ic > 6) FAIL(e, SWM_E_ARG, "initial condition %d is not one of 1..6", ic);
  if (bflag < 0 || bflag > 2) FAIL(e, SWM_E_ARG, "bathymetry flag %d is not one of 0..2", bflag);
  if (!x0 || !x1) FAIL(e, SWM_E_ARG, "domain corners are NULL");
  c = InitialSetup{};
  c.ic = ic;
  c.bflag = bflag;
  const int nel[2] = {J.dev.nx, J.dev.ny};
  const double hdx[2] = {J.dev.dx, J.dev.dy};
  for (int d = 0; d < 2; ++d) {
    c.x0[d] = x0[d];
    c.sx[d] = std::fabs(x1[d] - x0[d]);
    c.dx[d] = std::fabs(c.sx[d]) / (double)nel[d];
    if (!(c.dx[d] == hdx[d]))
      FAIL(e, SWM_E_ARG, "domain gives Mesh_DX(%d) = %.17g, the handle has %.17g", d + 1,
           c.dx[d], hdx[d]);
  }
  c.g = J.dev.g;
  c.pi = std::acos(-1.0);
  const double q1 = 0.5 * std::sqrt(3. / 7. + 2. / 7. * std::sqrt(6. / 5.));
  const double q2 = 0.5 * std::sqrt(3. / 7. - 2. / 7. * std::sqrt(6. / 5.));
  const double nodes[4] = {-q1, -q2, q2, q1};
  for (int g = 0; g < 4; ++g) {
    c.node[g] = nodes[g];
    c.w[g] = J.dev.wbnd[g];  // WeightsGP(i,j) = quadWeights1D(i)*quadWeights1D(j)
  }
  c.t = t;
  return SWM_OK;
}

int fill_initial_data(swm_engine* e, int id, int ic, int bflag, const double* x0,
                      const double* x1) {
  Job& J = *e->jobs[id];
  InitialSetup c;
  int rc0 = exact_setup(e, J, ic, bflag, x0, x1, 0.0, c);
  if (rc0) return rc0;
  if (J.dev.wb && ic != 2 && ic != 3)
    FAIL(e, SWM_E_ARG,
         "initial condition %d has no well-balanced state (ExactFunctionWB covers 2 and 3)", ic);
  if (bflag) {
    int rc = ensure_bathy(e, J);
    if (rc) return rc;
  }
  const bool wb = J.dev.wb != 0;
  launch_initial_data(c, J.dev.nx, J.dev.ny, J.dev.h, J.dev.hu, J.dev.hv,
                      wb ? J.dev.hres[1] : nullptr, wb ? J.dev.hun[1] : nullptr,
                      wb ? J.dev.hvn[1] : nullptr, bflag ? J.dev.bx : nullptr,
                      bflag ? J.dev.by : nullptr, e->stream);
  LAUNCHCK(e);
  HIPCK(e, hipStreamSynchronize(e->stream));
  J.dev.bathy = bflag ? 1 : 0;
  J.wb_ready = false;
  e->dirty = true;
  return wb ? build_wb(e, id) : SWM_OK;
}

// The benchmark families on the device (swm_fill_family): the state, for a
// lake the bathymetry tables, and for a well-balanced job the well-balanced
// state (the lake's h = 1 - b) and its FVTD.
int fill_family(swm_engine* e, int id, const swm_family* f) {
  Job& J = *e->jobs[id];
  if (!f) FAIL(e, SWM_E_ARG, "family is NULL");
  if (f->kind < SWM_FAMILY_VORTEX || f->kind > SWM_FAMILY_DAM_BREAK)
    FAIL(e, SWM_E_ARG, "family kind %d is not one of SWM_FAMILY_*", f->kind);
  const bool lake = f->kind == SWM_FAMILY_LAKE_BUMPS;
  if (lake && (f->nbumps < 0 || f->nbumps > SWM_FAMILY_MAX_BUMPS))
    FAIL(e, SWM_E_ARG, "nbumps %d is not in 0..%d", f->nbumps, SWM_FAMILY_MAX_BUMPS);
  if (J.dev.wb && !lake)
    FAIL(e, SWM_E_ARG, "family %d has no well-balanced state (only the lake has one)", f->kind);
  if (!(J.dev.dx == 1.0 / (double)J.dev.nx) || !(J.dev.dy == 1.0 / (double)J.dev.ny))
    FAIL(e, SWM_E_ARG, "the families live on the unit square: dx, dy must be 1/nx, 1/ny");
  FamilySetup c{};
  c.kind = f->kind;
  c.nbumps = lake ? f->nbumps : 0;
  c.dx = J.dev.dx;
  c.dy = J.dev.dy;
  c.g = J.dev.g;
  const double q1 = 0.5 * std::sqrt(3. / 7. + 2. / 7. * std::sqrt(6. / 5.));
  const double q2 = 0.5 * std::sqrt(3. / 7. - 2. / 7. * std::sqrt(6. / 5.));
  const double nodes[4] = {-q1, -q2, q2, q1};
  for (int g = 0; g < 4; ++g) {
    c.node[g] = nodes[g];
    c.w[g] = J.dev.wbnd[g];
  }
  c.eps = f->eps;
  for (int k = 0; k < SWM_FAMILY_MAX_BUMPS; ++k) {
    c.A[k] = f->A[k];
    c.cx[k] = f->cx[k];
    c.cy[k] = f->cy[k];
    c.s[k] = f->s[k];
  }
  c.amp = f->amp;
  c.px = f->px;
  c.py = f->py;
  c.ps = f->ps;
  c.radius = f->radius;
  c.h_in = f->h_in;
  c.h_out = f->h_out;
  if (lake) {
    int rc = ensure_bathy(e, J);
    if (rc) return rc;
  }
  const bool wb = J.dev.wb != 0;
  launch_family_data(c, J.dev.nx, J.dev.ny, J.dev.h, J.dev.hu, J.dev.hv,
                     wb ? J.dev.hres[1] : nullptr, wb ? J.dev.hun[1] : nullptr,
                     wb ? J.dev.hvn[1] : nullptr, lake ? J.dev.bx : nullptr,
                     lake ? J.dev.by : nullptr, e->stream);
  LAUNCHCK(e);
  HIPCK(e, hipStreamSynchronize(e->stream));
  J.dev.bathy = lake ? 1 : 0;
  J.wb_ready = false;
  e->dirty = true;
  return wb ? build_wb(e, id) : SWM_OK;
}

// PostProcessing's L1 error of job J's state (swm_l1_error).
int job_l1_error(swm_engine* e, Job& J, int ic, int bflag, double t, const double* x0,
                 const double* x1, double* err) {
  if (!err) FAIL(e, SWM_E_ARG, "err is NULL");
  InitialSetup c;
  int rc = exact_setup(e, J, ic, bflag, x0, x1, t, c);
  if (rc) return rc;
  const size_t n = J.lay.n;
  double* tmp = nullptr;
  HIPCK(e, hipMalloc((void**)&tmp, (3 * n + 3) * sizeof(double)));
  launch_l1_error(c, J.dev.nx, J.dev.ny, J.dev.h, J.dev.hu, J.dev.hv, tmp, tmp + 3 * n,
                  e->stream);
  hipError_t er = hipGetLastError();
  if (er == hipSuccess)
    er = hipMemcpyAsync(err, tmp + 3 * n, 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (er != hipSuccess) FAIL(e, SWM_E_HIP, "L1 error failed: %s", hipGetErrorString(er));
  return SWM_OK;
}

int engine_add(swm_engine* e, const swm_job_desc* d, int* id_out) {
  *id_out = -1;
  if (!d) FAIL(e, SWM_E_ARG, "job description is NULL");
  // initial data made on the device: a benchmark family, or the reference's own
  const bool fam = !d->U0 && d->family != nullptr;
  const bool exact = !d->U0 && !fam && d->initial_condition > 0;
  if (!d->U0 && !exact && !fam)
    FAIL(e, SWM_E_ARG, "U0 is NULL and neither family nor initial_condition is set");
  if (d->wellbalanced && !d->Uwb && !exact && !fam)
    FAIL(e, SWM_E_ARG, "wellbalanced job needs Uwb");
  if (!d->dbdx != !d->dbdy) FAIL(e, SWM_E_ARG, "dbdx and dbdy must both be set or both NULL");
  const SchemeDef* sc = find_scheme(d->timescheme == 0 ? 5 : d->timescheme);
  if (!sc) FAIL(e, SWM_E_ARG, "timescheme %d is not one of 0..7", d->timescheme);
  if (d->hp_association != SWM_HP_AS_REFERENCE && d->hp_association != SWM_HP_BY_NODE)
    FAIL(e, SWM_E_ARG, "hp_association must be SWM_HP_AS_REFERENCE or SWM_HP_BY_NODE");
  const int fix = recfix_code(d->reconstruction_fix == 0 ? 4 : d->reconstruction_fix);
  if (fix < 0)
    FAIL(e, SWM_E_ARG, "reconstruction_fix %d is not 0, 4, 1 or 2", d->reconstruction_fix);
  const bool tables = fam     ? d->family->kind == SWM_FAMILY_LAKE_BUMPS
                      : exact ? d->bathymetry_flag > 0
                              : d->dbdx != nullptr;
  const long long extra = tables ? (long long)bathy_bytes((size_t)d->nx * d->ny) : 0;
  int id;
  int rc = alloc_job(e, d->nx, d->ny, d->dx, d->dy, d->gravity, d->cfl, d->wellbalanced ? 1 : 0,
                     d->wgp_bnd, d->nsteps, extra, &id);
  if (rc) return rc;
  *id_out = id;
  Job& J = *e->jobs[id];
  J.scheme = sc;
  J.hp_assoc = d->hp_association;
  J.dev.recfix = fix;
  if (fam) {
    rc = fill_family(e, id, d->family);
    if (rc) return rc;
  } else if (exact) {
    rc = fill_initial_data(e, id, d->initial_condition, d->bathymetry_flag, d->mesh_x0,
                           d->mesh_x1);
    if (rc) return rc;
  } else {
    if (d->dbdx) {
      rc = job_set_bathy(e, id, d->dbdx, d->dbdy);
      if (rc) return rc;
    }
    if (d->wellbalanced) {
      rc = init_wb(e, id, d->Uwb, 0);
      if (rc) return rc;
    }
    rc = upload_state(e, J, d->U0, 0, J.dev.h, J.dev.hu, J.dev.hv);
    if (rc) return rc;
  }
  if (d->nsteps == 0) J.state = JOB_DONE;  // nothing to run; holds U0 until released
  return SWM_OK;
}

void engine_free_job(swm_engine* e, int id) {
  Job& J = *e->jobs[id];
  if (J.state == JOB_FREE) return;
  (void)hipStreamSynchronize(e->stream);
  if (J.vout_buf) {
    (void)hipFree(J.vout_buf);
    J.vout_buf = nullptr;
    J.vout_valid = false;
    e->in_use -= (long long)(3 * J.lay.n * sizeof(double));
    --e->keep_count;
  }
  if (J.slab) (void)hipFree(J.slab);
  J.slab = nullptr;
  e->in_use -= (long long)J.lay.bytes;
  if (J.bathy_buf) {
    (void)hipFree(J.bathy_buf);
    J.bathy_buf = nullptr;
    e->in_use -= (long long)bathy_bytes(J.lay.n);
  }
  J.state = JOB_FREE;
  J.own = Tables{};
  e->dirty = true;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------

int swm_engine_create(long long budget_bytes, swm_engine_t** out) {
  if (!out) FAIL((swm_engine*)nullptr, SWM_E_ARG, "out is NULL");
  *out = nullptr;
  swm_engine* e = new swm_engine();
  e->budget = budget_bytes;
  std::fill(std::begin(e->tune_ms), std::end(e->tune_ms), HUGE_VAL);
  hipError_t er = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (er != hipSuccess) {
    g_error = std::string("hipStreamCreate failed: ") + hipGetErrorString(er);
    delete e;
    return SWM_E_HIP;
  }
  int rc = status_capacity(e, 8);
  if (rc == SWM_OK && hipHostMalloc((void**)&e->hdt, sizeof(double), hipHostMallocDefault) != hipSuccess) {
    e->hdt = nullptr;
    g_error = "hipHostMalloc failed";
    rc = SWM_E_HIP;
  }
  if (rc == SWM_OK && hipHostMalloc((void**)&e->hzero, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    e->hzero = nullptr;
    g_error = "hipHostMalloc failed";
    rc = SWM_E_HIP;
  }
  if (rc == SWM_OK) *e->hzero = 0;
  if (rc) {
    swm_engine_destroy(e);
    return rc;
  }
  *out = e;
  return SWM_OK;
}

void swm_engine_destroy(swm_engine_t* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  for (size_t i = 0; i < e->jobs.size(); ++i) engine_free_job(e, (int)i);
  e->jobs.clear();
  if (e->hstatus) (void)hipHostFree(e->hstatus);
  if (e->hdt) (void)hipHostFree(e->hdt);
  if (e->hzero) (void)hipHostFree(e->hzero);
  if (e->dstatus.p) (void)hipFree(e->dstatus.p);
  e->dstatus.p = nullptr;
  if (e->djobs.p) (void)hipFree(e->djobs.p);
  e->djobs.p = nullptr;
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* swm_engine_last_error(const swm_engine_t* e) { return e ? e->err.c_str() : g_error.c_str(); }

long long swm_job_footprint_bytes(int nx, int ny, int wellbalanced, int bathymetry) {
  if (nx < 3 || ny < 3) return -1;
  const JobLayout L = job_layout(nx, ny, wellbalanced ? 1 : 0, 0);
  return (long long)(L.bytes + (bathymetry ? bathy_bytes(L.n) : 0));
}

long long swm_footprint_bytes(int nx, int ny, int wellbalanced, int bathymetry) {
  return swm_job_footprint_bytes(nx, ny, wellbalanced, bathymetry);
}

long long swm_engine_bytes_in_use(const swm_engine_t* e) { return e ? e->in_use : -1; }

int swm_engine_submit(swm_engine_t* e, const swm_job_desc* d, int* job_id) {
  if (!e || !job_id) return SWM_E_ARG;
  *job_id = -1;
  int id = -1;
  int rc = engine_add(e, d, &id);
  if (rc) {
    if (id >= 0) engine_free_job(e, id);  // a partially built job keeps no memory
    return rc;
  }
  *job_id = id;
  return SWM_OK;
}

// Up to max_steps steps of every running job, each group of jobs in batches
// of one host read; a job stops on its own at its step count or when its dt
// is not finite (it then leaves the tables before the next batch).
static int engine_run(swm_engine* e, int max_steps, int* n_running, int* steps_run) {
  int rc = refresh(e);
  if (rc) return rc;
  int most = 0;
  for (int g = 0; g < kGroups; ++g) {
    Tables& T = e->groups[g];
    std::vector<std::pair<int, int>> start;  // (job, steps_done) of the group
    for (int q = 0; q < T.njobs; ++q) start.push_back({T.jl[q], e->jobs[T.jl[q]]->steps_done});
    int done = 0;
    while (done < max_steps && T.njobs > 0) {
      const Job& J0 = *e->jobs[T.jl[0]];
      rc = engine_batch(e, T, J0.scheme, J0.hp_assoc, nullptr, max_steps - done);
      if (rc) return rc;
      const bool changed = e->dirty;
      const int done_before = done;
      for (int q = 0; q < T.njobs; ++q) {
        const int id = T.jl[q];
        Job& J = *e->jobs[id];
        if (e->hstatus[id].halt == kHaltDtBad) {  // not committed: keeps its last state
          J.state = JOB_FAILED;
          e->dirty = true;
        } else if (J.steps_left == 0) {
          J.state = JOB_DONE;
          e->dirty = true;
        }
      }
      for (const auto& js : start) done = std::max(done, e->jobs[js.first]->steps_done - js.second);
      if (done == done_before && !e->dirty && !changed)
        FAIL(e, SWM_E_STATE, "internal: a batch of %d jobs committed no step", T.njobs);
      rc = refresh(e);  // finished and failed jobs leave the tables
      if (rc) return rc;
    }
    most = std::max(most, done);
  }
  int running = 0;
  for (const auto& J : e->jobs) running += J->state == JOB_RUNNING;
  if (n_running) *n_running = running;
  if (steps_run) *steps_run = most;
  return SWM_OK;
}

int swm_engine_step(swm_engine_t* e, int* n_running) {
  if (!e) return SWM_E_ARG;
  return engine_run(e, 1, n_running, nullptr);
}

int swm_engine_run(swm_engine_t* e, int max_steps, int* n_running, int* steps_run) {
  if (!e) return SWM_E_ARG;
  if (max_steps < 1) FAIL(e, SWM_E_ARG, "max_steps must be >= 1 (got %d)", max_steps);
  return engine_run(e, max_steps, n_running, steps_run);
}

int swm_engine_job_info(swm_engine_t* e, int job_id, swm_job_info* info) {
  if (!e || !info) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  info->state = J->state;
  info->steps_done = J->steps_done;
  info->steps_left = J->steps_left;
  info->t = J->t;
  info->jacobi_sweeps = J->jacobi_sweeps;
  info->nonconverged = J->nonconv;
  info->bytes = (long long)J->lay.bytes + (J->bathy_buf ? (long long)bathy_bytes(J->lay.n) : 0) +
                (J->vout_buf ? (long long)(3 * J->lay.n * sizeof(double)) : 0);
  return SWM_OK;
}

int swm_engine_track_stage_min(swm_engine_t* e, int on) {
  if (!e) return SWM_E_ARG;
  e->track_min = on != 0;
  for (size_t id = 0; id < e->jobs.size(); ++id)
    if (e->jobs[id]->state != JOB_FREE)
      HIPCK(e, hipMemsetAsync(&e->dstatus.p[id].hmin_set, 0, sizeof(int), e->stream));
  HIPCK(e, hipStreamSynchronize(e->stream));
  return SWM_OK;
}

int swm_engine_stage_min_h(swm_engine_t* e, int job_id, double* hmin) {
  if (!e || !hmin) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  JobStatus st;
  HIPCK(e, hipStreamSynchronize(e->stream));
  HIPCK(e, hipMemcpy(&st, e->dstatus.p + job_id, sizeof(JobStatus), hipMemcpyDeviceToHost));
  if (!st.hmin_set) FAIL(e, SWM_E_STATE, "no DeC stage recorded since tracking was enabled");
  *hmin = st.hmin;
  return SWM_OK;
}

int swm_engine_fetch(swm_engine_t* e, int job_id, double* U) {
  if (!e || !U) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  const size_t n = J->lay.n;
  double* tmp = nullptr;
  HIPCK(e, hipMalloc((void**)&tmp, 3 * n * sizeof(double)));
  launch_soa_to_aos(J->dev.nx, J->dev.ny, J->dev.h, J->dev.hu, J->dev.hv, tmp, e->stream);
  hipError_t er = hipGetLastError();
  if (er == hipSuccess)
    er = hipMemcpyAsync(U, tmp, 3 * n * sizeof(double), hipMemcpyDeviceToHost, e->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (er != hipSuccess) FAIL(e, SWM_E_HIP, "fetch failed: %s", hipGetErrorString(er));
  return SWM_OK;
}

int swm_engine_release(swm_engine_t* e, int job_id) {
  if (!e) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  engine_free_job(e, job_id);
  return SWM_OK;
}

int swm_engine_set_jacobi_limits(swm_engine_t* e, int max_iter, double tol) {
  if (!e) return SWM_E_ARG;
  if (max_iter < 0) FAIL(e, SWM_E_ARG, "max_iter < 0");
  e->max_iter = max_iter;
  e->tol = tol;
  ++e->graph_epoch;  // captured launches carry the limits
  return SWM_OK;
}

int swm_engine_set_execution(swm_engine_t* e, int use_graph, int margin) {
  if (!e) return SWM_E_ARG;
  if (margin > 100 || margin < -2) FAIL(e, SWM_E_ARG, "margin must be in [-2, 100]");
  e->use_graph = use_graph < 0 ? -1 : (use_graph != 0);
  e->margin = margin;
  return SWM_OK;
}

int swm_engine_set_tile_shape(swm_engine_t* e, int shape) {
  if (!e) return SWM_E_ARG;
  if (shape < -1 || shape >= kNumShapes)
    FAIL(e, SWM_E_ARG, "tile shape must be -1 (autotuned) or 0..%d (got %d)", kNumShapes - 1,
         shape);
  e->tile_shape = shape;
  return SWM_OK;
}

int swm_engine_tile_shape(const swm_engine_t* e, int dims[4]) {
  if (!e) return SWM_E_ARG;
  const int sh = e->tile_shape >= 0 ? e->tile_shape : e->tuned_shape;
  if (dims) {
    dims[0] = kShapes[sh].xw;
    dims[1] = kShapes[sh].xh;
    dims[2] = kShapes[sh].yw;
    dims[3] = kShapes[sh].yh;
  }
  return sh;
}

int swm_engine_execution_stats(const swm_engine_t* e, long long out[6]) {
  if (!e || !out) return SWM_E_ARG;
  out[0] = e->spec_steps;
  out[1] = e->redone_steps;
  out[2] = e->graph_launches;
  out[3] = e->graph_captures;
  out[4] = e->use_graph >= 0 ? e->use_graph : e->tuned_graph;
  out[5] = e->margin != -1 ? e->margin : e->tuned_margin;
  return SWM_OK;
}

int swm_engine_status_reads(const swm_engine_t* e, long long* n) {
  if (!e || !n) return SWM_E_ARG;
  *n = e->status_reads;
  return SWM_OK;
}

int swm_engine_profile_enable(swm_engine_t* e, int on) {
  return swm_engine_profile_categories(e, on ? (1u << SWM_PROF_NCAT) - 1u : 0u);
}

int swm_engine_profile_categories(swm_engine_t* e, unsigned mask) {
  if (!e) return SWM_E_ARG;
  if (mask >> SWM_PROF_NCAT)
    FAIL(e, SWM_E_ARG, "profile mask 0x%x has bits beyond the %d categories", mask, SWM_PROF_NCAT);
  e->prof.mask = mask;
  return SWM_OK;
}

int swm_engine_profile_read(swm_engine_t* e, double ms[SWM_PROF_NCAT],
                            long long launches[SWM_PROF_NCAT]) {
  if (!e || !ms || !launches) return SWM_E_ARG;
  HIPCK(e, hipStreamSynchronize(e->stream));
  for (int c = 0; c < SWM_PROF_NCAT; ++c) {
    ms[c] = 0.0;
    launches[c] = 0;
  }
  for (size_t k = 0; k < e->prof.cats.size(); ++k) {
    float t = 0.f;
    HIPCK(e, hipEventElapsedTime(&t, e->prof.pool[2 * k], e->prof.pool[2 * k + 1]));
    ms[e->prof.cats[k]] += t;
    launches[e->prof.cats[k]] += 1;
  }
  e->prof.cats.clear();
  e->prof.used = 0;
  return SWM_OK;
}

const char* swm_profile_name(int cat) {
  return (cat >= 0 && cat < SWM_PROF_NCAT) ? PROF_NAMES[cat] : "";
}

// ---------------------------------------------------------------------------
// plain solver: a one-job engine with the reference's call structure
// ---------------------------------------------------------------------------

static Job* plain_job(swm_solver_t* s) { return s->e->jobs[s->job].get(); }

int swm_create(int nx, int ny, double dx, double dy, double gravity, double cfl, int wellbalanced,
               const double wgp_bnd[4], swm_solver_t** out) {
  if (!out) FAIL((swm_engine*)nullptr, SWM_E_ARG, "out is NULL");
  *out = nullptr;
  if (nx < 3 || ny < 3)
    FAIL((swm_engine*)nullptr, SWM_E_ARG, "nx, ny must be >= 3 (got %d, %d)", nx, ny);
  if (!wgp_bnd) FAIL((swm_engine*)nullptr, SWM_E_ARG, "wgp_bnd is NULL");
  if (!(dx > 0) || !(dy > 0)) FAIL((swm_engine*)nullptr, SWM_E_ARG, "dx, dy must be > 0");
  swm_solver* s = new swm_solver();
  int rc = swm_engine_create(0, &s->e);
  if (rc) {
    delete s;
    return rc;
  }
  rc = alloc_job(s->e, nx, ny, dx, dy, gravity, cfl, wellbalanced ? 1 : 0, wgp_bnd, -1, 0, &s->job);
  if (rc) {
    swm_destroy(s);
    return rc;
  }
  *out = s;
  return SWM_OK;
}

void swm_destroy(swm_solver_t* s) {
  if (!s) return;
  if (s->out_stream) {
    (void)hipStreamSynchronize(s->out_stream);
    (void)hipEventDestroy(s->out_ready);
    (void)hipEventDestroy(s->out_done);
    (void)hipStreamDestroy(s->out_stream);
  }
  if (s->out_stage) (void)hipFree(s->out_stage);
  swm_engine_destroy(s->e);
  delete s;
}

const char* swm_last_error(const swm_solver_t* s) {
  return s ? s->e->err.c_str() : g_error.c_str();
}

int swm_set_bathy_grad(swm_solver_t* s, const double* dbdx, const double* dbdy) {
  if (!s) return SWM_E_ARG;
  return job_set_bathy(s->e, s->job, dbdx, dbdy);
}

int swm_init_wb(swm_solver_t* s, const double* Uwb, int ghosts) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (!Uwb) FAIL(e, SWM_E_ARG, "Uwb is NULL");
  if (ghosts != 0 && ghosts != 3) FAIL(e, SWM_E_ARG, "ghosts must be 0 or 3 (got %d)", ghosts);
  if (!plain_job(s)->dev.wb) FAIL(e, SWM_E_STATE, "handle was created without wellbalanced");
  return init_wb(e, s->job, Uwb, ghosts);
}

int swm_fill_initial_data(swm_solver_t* s, int initial_condition, int bathymetry_flag,
                          const double x0[2], const double x1[2]) {
  if (!s) return SWM_E_ARG;
  return fill_initial_data(s->e, s->job, initial_condition, bathymetry_flag, x0, x1);
}

int swm_fill_family(swm_solver_t* s, const swm_family* family) {
  if (!s) return SWM_E_ARG;
  return fill_family(s->e, s->job, family);
}

int swm_l1_error(swm_solver_t* s, int initial_condition, int bathymetry_flag, double t,
                 const double x0[2], const double x1[2], double err[3]) {
  if (!s) return SWM_E_ARG;
  return job_l1_error(s->e, *plain_job(s), initial_condition, bathymetry_flag, t, x0, x1, err);
}

int swm_engine_l1_error(swm_engine_t* e, int job_id, int initial_condition, int bathymetry_flag,
                        double t, const double x0[2], const double x1[2], double err[3]) {
  if (!e) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  return job_l1_error(e, *J, initial_condition, bathymetry_flag, t, x0, x1, err);
}

int swm_download_primitive(swm_solver_t* s, double* V) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (!V) FAIL(e, SWM_E_ARG, "V is NULL");
  Job& J = *plain_job(s);
  const size_t n = J.lay.n;
  double* tmp = nullptr;
  HIPCK(e, hipMalloc((void**)&tmp, 3 * n * sizeof(double)));
  launch_primitives(n, J.dev.h, J.dev.hu, J.dev.hv, tmp, e->stream);
  hipError_t er = hipGetLastError();
  if (er == hipSuccess)
    er = hipMemcpyAsync(V, tmp, 3 * n * sizeof(double), hipMemcpyDeviceToHost, e->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (er != hipSuccess) FAIL(e, SWM_E_HIP, "primitive download failed: %s", hipGetErrorString(er));
  return SWM_OK;
}

// Keeping (or not) the writer's state of job J (swm_keep_output_state).
static int job_keep_output(swm_engine* e, Job& J, int on) {
  if ((on != 0) == (J.vout_buf != nullptr)) return SWM_OK;
  const long long b = (long long)(3 * J.lay.n * sizeof(double));
  if (on) {
    if (e->budget > 0 && e->in_use + b > e->budget)
      FAIL(e, SWM_E_BUDGET, "the kept output state of %lld bytes does not fit", b);
    HIPCK(e, hipMalloc((void**)&J.vout_buf, (size_t)b));
    e->in_use += b;
    ++e->keep_count;
  } else {
    HIPCK(e, hipStreamSynchronize(e->stream));
    (void)hipFree(J.vout_buf);
    J.vout_buf = nullptr;
    e->in_use -= b;
    --e->keep_count;
  }
  J.dev.vout = J.vout_buf;
  J.vout_valid = false;
  e->dirty = true;  // the descriptor changed; captured steps are rebuilt
  return SWM_OK;
}

// The primitives of job J's kept writer state (swm_download_output_primitive).
static int job_output_primitive(swm_engine* e, Job& J, double* V) {
  if (!V) FAIL(e, SWM_E_ARG, "V is NULL");
  if (!J.vout_valid)
    FAIL(e, SWM_E_STATE,
         "no kept output state: swm_keep_output_state before a step of a Patankar scheme");
  const size_t n = J.lay.n;
  double* tmp = nullptr;
  HIPCK(e, hipMalloc((void**)&tmp, 3 * n * sizeof(double)));
  launch_primitives(n, J.vout_buf, J.vout_buf + n, J.vout_buf + 2 * n, tmp, e->stream);
  hipError_t er = hipGetLastError();
  if (er == hipSuccess)
    er = hipMemcpyAsync(V, tmp, 3 * n * sizeof(double), hipMemcpyDeviceToHost, e->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  (void)hipFree(tmp);
  if (er != hipSuccess) FAIL(e, SWM_E_HIP, "output download failed: %s", hipGetErrorString(er));
  return SWM_OK;
}

int swm_keep_output_state(swm_solver_t* s, int on) {
  if (!s) return SWM_E_ARG;
  return job_keep_output(s->e, *plain_job(s), on);
}

int swm_download_output_primitive(swm_solver_t* s, double* V) {
  if (!s) return SWM_E_ARG;
  return job_output_primitive(s->e, *plain_job(s), V);
}

int swm_engine_keep_output_state(swm_engine_t* e, int job_id, int on) {
  if (!e) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  return job_keep_output(e, *J, on);
}

int swm_engine_fetch_output_primitive(swm_engine_t* e, int job_id, double* V) {
  if (!e) return SWM_E_ARG;
  Job* J;
  int rc = job_from_id(e, job_id, &J);
  if (rc) return rc;
  return job_output_primitive(e, *J, V);
}

int swm_download_primitive_async(swm_solver_t* s, double* V) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (!V) FAIL(e, SWM_E_ARG, "V is NULL");
  Job& J = *plain_job(s);
  const size_t n = J.lay.n;
  if (!s->out_stream) {
    HIPCK(e, hipMalloc((void**)&s->out_stage, 3 * n * sizeof(double)));
    HIPCK(e, hipStreamCreateWithFlags(&s->out_stream, hipStreamNonBlocking));
    HIPCK(e, hipEventCreateWithFlags(&s->out_ready, hipEventDisableTiming));
    HIPCK(e, hipEventCreateWithFlags(&s->out_done, hipEventDisableTiming));
  }
  // the previous copy must have left the snapshot before it is overwritten
  if (s->out_pending) HIPCK(e, hipStreamWaitEvent(e->stream, s->out_done, 0));
  launch_primitives(n, J.dev.h, J.dev.hu, J.dev.hv, s->out_stage, e->stream);
  LAUNCHCK(e);
  HIPCK(e, hipEventRecord(s->out_ready, e->stream));
  HIPCK(e, hipStreamWaitEvent(s->out_stream, s->out_ready, 0));
  HIPCK(e, hipMemcpyAsync(V, s->out_stage, 3 * n * sizeof(double), hipMemcpyDeviceToHost,
                          s->out_stream));
  HIPCK(e, hipEventRecord(s->out_done, s->out_stream));
  s->out_pending = true;
  return SWM_OK;
}

int swm_output_wait(swm_solver_t* s) {
  if (!s) return SWM_E_ARG;
  if (s->out_pending) {
    HIPCK(s->e, hipEventSynchronize(s->out_done));
    s->out_pending = false;
  }
  return SWM_OK;
}

int swm_upload(swm_solver_t* s, const double* U, int ghosts) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (!U) FAIL(e, SWM_E_ARG, "U is NULL");
  if (ghosts != 0 && ghosts != 3) FAIL(e, SWM_E_ARG, "ghosts must be 0 or 3 (got %d)", ghosts);
  Job& J = *plain_job(s);
  return upload_state(e, J, U, ghosts, J.dev.h, J.dev.hu, J.dev.hv);
}

int swm_download(swm_solver_t* s, double* U) {
  if (!s) return SWM_E_ARG;
  return swm_engine_fetch(s->e, s->job, U);
}

int swm_timestep(swm_solver_t* s, double* dt_min, double* lmx, double* lmy) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  int rc = refresh(e);
  if (rc) return rc;
  Tables& T = plain_job(s)->own;
  timed(e, 6, [&] { launch_timestep(eargs(e), T.dcb.p, T.ncb, T.djl.p, T.njobs, e->stream); });
  LAUNCHCK(e);
  rc = read_status(e);
  if (rc) return rc;
  const JobStatus& js = e->hstatus[s->job];
  if (dt_min) *dt_min = js.dt_min;  // TimeStep's value: no time-loop clipping
  if (lmx) *lmx = js.lmx;
  if (lmy) *lmy = js.lmy;
  return SWM_OK;
}

// One step of the given scheme with a caller-provided dt (the TimeDiscretizationBy*
// routines); jacobi_iters receives the scheme's solve count of values.
static int plain_step(swm_solver_t* s, int timescheme, double dt, int* jacobi_iters) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  Job& J = *plain_job(s);
  if (J.dev.wb && !J.wb_ready)
    FAIL(e, SWM_E_STATE,
         "well-balanced state not initialised (swm_init_wb or swm_fill_initial_data)");
  const SchemeDef* sc = find_scheme(timescheme);
  int rc = refresh(e);
  if (rc) return rc;
  const int before = J.nonconv;
  rc = engine_batch(e, J.own, sc, J.hp_assoc, &dt, 1);
  if (rc) return rc;
  if (jacobi_iters)
    for (int i = 0; i < sc->nsolve; ++i) jacobi_iters[i] = J.last_iters[i];
  return J.nonconv - before;
}

int swm_step_mpdec5(swm_solver_t* s, double t, double dt, int jacobi_iters[13]) {
  (void)t;
  return plain_step(s, 5, dt, jacobi_iters);
}

int swm_step_mpeuler(swm_solver_t* s, double t, double dt, int jacobi_iters[1]) {
  (void)t;
  return plain_step(s, 6, dt, jacobi_iters);
}

int swm_step_mpdec2(swm_solver_t* s, double t, double dt, int jacobi_iters[2]) {
  (void)t;
  return plain_step(s, 7, dt, jacobi_iters);
}

int swm_step_forward_euler(swm_solver_t* s, double t, double dt) {
  (void)t;
  return plain_step(s, 1, dt, nullptr);
}

int swm_step_ssprk4(swm_solver_t* s, double t, double dt) {
  (void)t;
  return plain_step(s, 2, dt, nullptr);
}

int swm_step_rk65(swm_solver_t* s, double t, double dt) {
  (void)t;
  return plain_step(s, 3, dt, nullptr);
}

int swm_step_dec5(swm_solver_t* s, double t, double dt) {
  (void)t;
  return plain_step(s, 4, dt, nullptr);
}

int swm_set_timescheme(swm_solver_t* s, int timescheme) {
  if (!s) return SWM_E_ARG;
  const SchemeDef* sc = find_scheme(timescheme);
  if (!sc) FAIL(s->e, SWM_E_ARG, "timescheme %d is not one of 1..7", timescheme);
  plain_job(s)->scheme = sc;
  return SWM_OK;
}

int swm_set_reconstruction_fix(swm_solver_t* s, int recfix) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  const int code = recfix_code(recfix);
  if (code < 0)
    FAIL(e, SWM_E_ARG,
         "reconstruction fix %d is not 4, 1 or 2 (3, WENO3, defines 2 of the 4 Gauss values)",
         recfix);
  Job& J = *plain_job(s);
  if (J.dev.recfix != code) {
    J.dev.recfix = code;
    J.wb_ready = false;  // the well-balanced reference is an FVTD with this setting
    e->dirty = true;
  }
  return SWM_OK;
}

int swm_set_hp_association(swm_solver_t* s, int mode) {
  if (!s) return SWM_E_ARG;
  if (mode != SWM_HP_AS_REFERENCE && mode != SWM_HP_BY_NODE)
    FAIL(s->e, SWM_E_ARG, "hp association must be SWM_HP_AS_REFERENCE or SWM_HP_BY_NODE (got %d)",
         mode);
  plain_job(s)->hp_assoc = mode;
  return SWM_OK;
}

// Sets a job's step log for the duration of one API call.
struct LogScope {
  Job& J;
  std::vector<StepRec> log;
  explicit LogScope(Job& j) : J(j) { J.log = &log; }
  ~LogScope() { J.log = nullptr; }
};

int swm_run_steps(swm_solver_t* s, int nsteps, double* dts, int* jacobi_iters) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (nsteps < 0) FAIL(e, SWM_E_ARG, "nsteps < 0");
  Job& J = *plain_job(s);
  if (J.dev.wb && !J.wb_ready)
    FAIL(e, SWM_E_STATE,
         "well-balanced state not initialised (swm_init_wb or swm_fill_initial_data)");
  int rc = refresh(e);
  if (rc) return rc;
  const int before = J.nonconv;
  const int ns = J.scheme->nsolve;
  LogScope L(J);
  int k = 0;
  while (k < nsteps) {
    // dt from TimeStep on the device, within the step; a step whose dt is
    // not finite is not committed, so the state stays that of step k
    L.log.clear();
    rc = engine_batch(e, J.own, J.scheme, J.hp_assoc, nullptr, nsteps - k);
    if (rc) return rc;
    for (const StepRec& r : L.log) {
      if (dts) dts[k] = r.dt;
      if (jacobi_iters)
        for (int i = 0; i < ns; ++i) jacobi_iters[ns * k + i] = r.iters[i];
      ++k;
    }
    const JobStatus& js = e->hstatus[s->job];
    if (js.halt == kHaltDtBad) {
      if (dts) dts[k] = js.dt;
      FAIL(e, SWM_E_STATE, "time step %d: dt is not finite (%g)", k, js.dt);
    }
    if (L.log.empty()) FAIL(e, SWM_E_STATE, "internal: a batch committed no step (halt %d)", js.halt);
  }
  HIPCK(e, hipStreamSynchronize(e->stream));
  return J.nonconv - before;
}

int swm_advance(swm_solver_t* s, double* t, double t_analyze, double t_end, int max_steps,
                int* steps_done, double* dts) {
  if (!s || !t) return SWM_E_ARG;
  swm_engine* e = s->e;
  if (max_steps < 0) FAIL(e, SWM_E_ARG, "max_steps < 0");
  Job& J = *plain_job(s);
  if (J.dev.wb && !J.wb_ready)
    FAIL(e, SWM_E_STATE,
         "well-balanced state not initialised (swm_init_wb or swm_fill_initial_data)");
  int rc = refresh(e);
  if (rc) return rc;
  const int before = J.nonconv;
  LogScope L(J);
  int k = 0;
  if (steps_done) *steps_done = 0;
  double dt_prev = 0.0;
  while (k < max_steps) {
    // src/timediscretization.f90:58-59 and :83-94 on the device; a batch
    // covers about the steps left to the next stop (k_gate ends it there)
    int batch = 1;
    if (dt_prev > 0.0) {
      const double left = std::min(t_analyze, t_end) - *t;
      batch = (int)std::min(16.0, std::max(1.0, std::floor(left / dt_prev) + 1.0));
    }
    batch = std::min(batch, max_steps - k);
    const LoopArgs la{*t, t_analyze, t_end};
    L.log.clear();
    rc = engine_batch(e, J.own, J.scheme, J.hp_assoc, nullptr, batch, &la);
    if (rc) return rc;
    for (const StepRec& r : L.log) {
      if (dts) dts[k] = r.dt;
      *t = *t + r.dt;  // as k_gate: the loop's t = t + dt
      dt_prev = r.dt;
      ++k;
    }
    if (steps_done) *steps_done = k;
    const JobStatus& js = e->hstatus[s->job];
    if (js.halt == kHaltDtBad)
      FAIL(e, SWM_E_STATE, "dt is not finite (%g) at t = %.17g", js.dt, *t);
    if (js.halt == kHaltReached) break;
    if (L.log.empty()) FAIL(e, SWM_E_STATE, "internal: a batch committed no step (halt %d)", js.halt);
  }
  HIPCK(e, hipStreamSynchronize(e->stream));
  return J.nonconv - before;
}

int swm_fvtd(swm_solver_t* s, double* Ut, double* FX, double* FY, double* S, double* Prod,
             double* Dest) {
  if (!s) return SWM_E_ARG;
  swm_engine* e = s->e;
  Job& J = *plain_job(s);
  if (J.dev.wb && !J.wb_ready) FAIL(e, SWM_E_STATE, "well-balanced state not initialised");
  int rc = refresh(e);
  if (rc) return rc;
  NodeList L{};
  L.count = 1;
  L.m[0] = 0;
  L.out[0] = 0;
  L.slot[0] = 0;
  L.debug_S = 1;
  rc = fvtd_launch(e, J.own, L);
  if (rc) return rc;
  const int nx = J.dev.nx, ny = J.dev.ny;
  const size_t n = J.lay.n, nfx = J.lay.nfx, nfy = J.lay.nfy;
  const JobDev& D = J.dev;
  std::vector<double> ut2(n), ut3(n), s2(n), s3(n), qx(nfx), qy(nfy);
  std::vector<double> fx[3], fy[3];
  auto get = [&](std::vector<double>& v, const double* p, size_t cnt) {
    v.resize(cnt);
    return hipMemcpyAsync(v.data(), p, cnt * sizeof(double), hipMemcpyDeviceToHost, e->stream);
  };
  for (int v = 0; v < 3; ++v) {
    HIPCK(e, get(fx[v], D.FX[0][v], nfx));
    HIPCK(e, get(fy[v], D.FY[0][v], nfy));
  }
  HIPCK(e, get(ut2, D.FU2[0], n));
  HIPCK(e, get(ut3, D.FU3[0], n));
  HIPCK(e, get(s2, D.M[0][0], n));
  HIPCK(e, get(s3, D.M[0][1], n));
  HIPCK(e, get(qx, D.qx[0], nfx));
  HIPCK(e, get(qy, D.qy[0], nfy));
  HIPCK(e, hipStreamSynchronize(e->stream));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t k = (size_t)j * nx + i;
      if (Ut) {
        Ut[3 * k] = 0.0;
        Ut[3 * k + 1] = ut2[k];
        Ut[3 * k + 2] = ut3[k];
      }
      if (S) {
        S[3 * k] = 0.0;  // SourceFunc's S(1) is identically 0
        S[3 * k + 1] = s2[k];
        S[3 * k + 2] = s3[k];
      }
    }
  // the reference's module FX/FY after UpdateTimeDerivative: the face
  // kernels store them with FXWB/FYWB already subtracted
  if (FX)
    for (size_t k = 0; k < nfx; ++k)
      for (int v = 0; v < 3; ++v) FX[3 * k + v] = fx[v][k];
  if (FY)
    for (size_t k = 0; k < nfy; ++k)
      for (int v = 0; v < 3; ++v) FY[3 * k + v] = fy[v][k];
  if (Prod || Dest)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t k = (size_t)j * nx + i;
        double q[5] = {qy[(size_t)j * nx + i], qx[(size_t)j * (nx + 1) + i], 0.0,
                       qx[(size_t)j * (nx + 1) + i + 1], qy[(size_t)(j + 1) * nx + i]};
        for (int sl = 0; sl < 5; ++sl) {
          double P, Dd;
          if (sl == 2) {
            P = 0.;
            Dd = 0.;
          } else if (sl < 2) {
            P = (q[sl] > 0.) ? q[sl] : 0.;
            Dd = (q[sl] < 0.) ? -q[sl] : -0.;
          } else {
            P = (q[sl] < 0.) ? -q[sl] : 0.;
            Dd = (q[sl] > 0.) ? q[sl] : -0.;
          }
          if (Prod) Prod[5 * k + sl] = P;
          if (Dest) Dest[5 * k + sl] = Dd;
        }
      }
  return SWM_OK;
}

int swm_track_stage_min(swm_solver_t* s, int on) {
  if (!s) return SWM_E_ARG;
  return swm_engine_track_stage_min(s->e, on);
}

int swm_stage_min_h(swm_solver_t* s, double* hmin) {
  if (!s) return SWM_E_ARG;
  return swm_engine_stage_min_h(s->e, s->job, hmin);
}

int swm_set_execution(swm_solver_t* s, int use_graph, int margin) {
  if (!s) return SWM_E_ARG;
  return swm_engine_set_execution(s->e, use_graph, margin);
}

int swm_set_tile_shape(swm_solver_t* s, int shape) {
  if (!s) return SWM_E_ARG;
  return swm_engine_set_tile_shape(s->e, shape);
}

int swm_tile_shape(const swm_solver_t* s, int dims[4]) {
  if (!s) return SWM_E_ARG;
  return swm_engine_tile_shape(s->e, dims);
}

int swm_execution_stats(const swm_solver_t* s, long long stats[6]) {
  if (!s) return SWM_E_ARG;
  return swm_engine_execution_stats(s->e, stats);
}

int swm_status_reads(const swm_solver_t* s, long long* n) {
  if (!s) return SWM_E_ARG;
  return swm_engine_status_reads(s->e, n);
}

int swm_set_jacobi_limits(swm_solver_t* s, int max_iter, double tol) {
  if (!s) return SWM_E_ARG;
  return swm_engine_set_jacobi_limits(s->e, max_iter, tol);
}

long long swm_ambiguous_decisions(const swm_solver_t* s) { return s ? s->e->ambiguous : -1; }

int swm_profile_enable(swm_solver_t* s, int on) {
  if (!s) return SWM_E_ARG;
  return swm_engine_profile_enable(s->e, on);
}

int swm_profile_categories(swm_solver_t* s, unsigned mask) {
  if (!s) return SWM_E_ARG;
  return swm_engine_profile_categories(s->e, mask);
}

int swm_profile_read(swm_solver_t* s, double ms[SWM_PROF_NCAT], long long launches[SWM_PROF_NCAT]) {
  if (!s) return SWM_E_ARG;
  return swm_engine_profile_read(s->e, ms, launches);
}

int swm_sync(swm_solver_t* s) {
  if (!s) return SWM_E_ARG;
  HIPCK(s->e, hipStreamSynchronize(s->e->stream));
  return SWM_OK;
}

int swm_debug_divsqrt(int n, const double* a, const double* b, double* q, double* r) {
  if (n <= 0 || !a || !b || !q || !r) return SWM_E_ARG;
  double* da = nullptr;
  size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc((void**)&da, 4 * bytes) != hipSuccess) return SWM_E_NOMEM;
  double *db = da + n, *dq = db + n, *dr = dq + n;
  int rc = SWM_OK;
  if (hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = SWM_E_HIP;
  if (rc == SWM_OK) {
    launch_divsqrt(n, da, db, dq, dr, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(q, dq, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(r, dr, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = SWM_E_HIP;
  }
  (void)hipFree(da);
  return rc;
}

int swm_debug_exp(int n, const double* x, double* y) {
  if (n <= 0 || !x || !y) return SWM_E_ARG;
  double* dx = nullptr;
  size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc((void**)&dx, 2 * bytes) != hipSuccess) return SWM_E_NOMEM;
  double* dy = dx + n;
  int rc = SWM_OK;
  if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SWM_E_HIP;
  if (rc == SWM_OK) {
    launch_det_exp(n, dx, dy, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(y, dy, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = SWM_E_HIP;
  }
  (void)hipFree(dx);
  return rc;
}

int swm_debug_shared_div(int n, const double* a, const double* b, double* q) {
  if (n <= 0 || !a || !b || !q) return SWM_E_ARG;
  double* da = nullptr;
  size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc((void**)&da, 3 * bytes) != hipSuccess) return SWM_E_NOMEM;
  double *db = da + n, *dq = db + n;
  int rc = SWM_OK;
  if (hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = SWM_E_HIP;
  if (rc == SWM_OK) {
    launch_shared_div(n, da, db, dq, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(q, dq, bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = SWM_E_HIP;
  }
  (void)hipFree(da);
  return rc;
}

int swm_debug_weno5(int n, const double* q, double* first, double* second) {
  if (n <= 0 || !q || !first || !second) return SWM_E_ARG;
  double* dq = nullptr;
  const size_t nq = 5 * (size_t)n, nf = 2 * (size_t)n, ns = 4 * (size_t)n;
  if (hipMalloc((void**)&dq, (nq + nf + ns) * sizeof(double)) != hipSuccess) return SWM_E_NOMEM;
  double *df = dq + nq, *ds = df + nf;
  int rc = SWM_OK;
  if (hipMemcpy(dq, q, nq * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = SWM_E_HIP;
  if (rc == SWM_OK) {
    launch_weno5(n, dq, df, ds, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(first, df, nf * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(second, ds, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = SWM_E_HIP;
  }
  (void)hipFree(dq);
  return rc;
}

}  // extern "C"
