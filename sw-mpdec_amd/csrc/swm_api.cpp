// swm_api.cpp -- host side of the C ABI declared in include/swm.h: device
// memory of one solve, the mPDeC5 step orchestration and the Jacobi loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swm.h"
#include "swm_kernels.hpp"

using namespace swm;

namespace {

thread_local std::string g_create_error;

// thetaDeC (src/timediscretization.f90:491-506), THETA[m][r] = thetaDeC(m+1, r+1)
const double THETA[4][4] = {
    {0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000, 0.0000000000000000000},
    {0.1103005664791649049, 0.1896994335208351257, -0.0339073642291439076, 0.0103005664791649201},
    {0.0730327668541684294, 0.4505740308958107176, 0.2269672331458314485, -0.0269672331458315692},
    {0.0833333333333333287, 0.4166666666666666852, 0.4166666666666666852, 0.0833333333333333287}};

size_t align_up(size_t b) { return (b + 255) & ~size_t(255); }

const char* const PROF_NAMES[SWM_PROF_NCAT] = {
    "fvtd_x", "fvtd_y", "fvtd_update", "momentum", "patankar_assemble", "jacobi_sweep",
    "timestep", "other"};

// HIP-event timing of every kernel launch of a handle, by category.
struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<int> cats;  // category of event pair k = (pool[2k], pool[2k+1])

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

struct Layout {
  size_t n, nfx, nfy;
  int nb_jac, nb_ts;
};

Layout layout_of(int nx, int ny) {
  Layout L;
  L.n = (size_t)nx * ny;
  L.nfx = (size_t)(nx + 1) * ny;
  L.nfy = (size_t)nx * (ny + 1);
  GridGeom G{};
  G.nx = nx;
  G.ny = ny;
  L.nb_jac = jacobi_nblocks(G);
  L.nb_ts = timestep_nblocks(G);
  return L;
}

// Byte size of every device array of one handle, in allocation order.
size_t slab_bytes(int nx, int ny, int wb, int bathy) {
  Layout L = layout_of(nx, ny);
  size_t b = 0;
  auto add = [&](size_t cnt) { b += align_up(cnt * sizeof(double)); };
  for (int k = 0; k < 3; ++k) add(L.n);                 // U^n
  for (int k = 0; k < 3 * 3; ++k) add(L.n);             // node depths, 3 roles x 3 nodes
  for (int k = 0; k < 3 * 2; ++k) add(L.n);             // node momenta
  for (int k = 0; k < 4 * 2; ++k) add(L.n);             // FU(2:3) per node
  for (int k = 0; k < 4; ++k) { add(L.nfx); add(L.nfy); }  // qx, qy per node
  for (int k = 0; k < 3; ++k) { add(L.nfx); add(L.nfy); }  // FX, FY scratch
  for (int k = 0; k < kMaxSys * 5; ++k) add(L.n);       // Patankar matrices
  add(3 * L.n);                                         // staging for AoS copies
  if (wb) {
    for (int k = 0; k < 3; ++k) { add(L.nfx); add(L.nfy); }
    add(L.n);
    add(L.n);
  }
  add(((size_t)(kMaxSys + 1) * L.nb_jac));             // Jacobi partial sums
  add(4 * (size_t)L.nb_ts + 4);                         // TimeStep partials + out
  b += align_up(kMaxSys * sizeof(JacState)) + align_up(2 * sizeof(unsigned int));
  return b;
}

// Bathymetry tables live in their own allocation, made on first use.
size_t bathy_bytes(int nx, int ny) { return 32 * (size_t)nx * ny * sizeof(double); }

}  // namespace

struct swm_solver {
  int nx = 0, ny = 0;
  Layout L{};
  GridGeom G{};
  double cfl = 0;
  int wb = 0, bathy = 0;
  bool wb_ready = false;
  hipStream_t stream = nullptr;
  char* slab = nullptr;
  size_t slab_size = 0, slab_used = 0;
  double *h = nullptr, *hu = nullptr, *hv = nullptr;
  double* hbuf[4][3] = {};          // nodes 1..3: three depth buffers whose roles rotate
  double *hun[4] = {}, *hvn[4] = {};  // node momenta
  double *FU2[4] = {}, *FU3[4] = {};
  double *qx[4] = {}, *qy[4] = {};
  double *FX[3] = {}, *FY[3] = {};
  double* M[kMaxSys][5] = {};
  double* stage = nullptr;
  double *FXWB[3] = {}, *FYWB[3] = {}, *SWB2 = nullptr, *SWB3 = nullptr;
  double *bx = nullptr, *by = nullptr;
  double* jac_partial = nullptr;
  double* ts_partial = nullptr;
  double* ts_out = nullptr;
  JacState* st = nullptr;
  unsigned int* counters = nullptr;
  JacState* st_host = nullptr;   // pinned
  double* ts_host = nullptr;     // pinned
  int max_iter = 1000;
  double tol = 1.e-16;
  int pred[13];
  long long ambiguous = 0;
  Profiler prof;
  std::string err;

  double* take(size_t cnt) {
    size_t b = align_up(cnt * sizeof(double));
    double* p = reinterpret_cast<double*>(slab + slab_used);
    slab_used += b;
    return p;
  }
};

#define SWM_FAIL(h, code, ...)                                                 \
  do {                                                                         \
    char _b[512];                                                              \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                                     \
    if (h) (h)->err = _b; else g_create_error = _b;                            \
    return code;                                                               \
  } while (0)

#define HIPCK(h, call)                                                         \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess)                                                      \
      SWM_FAIL(h, _e == hipErrorOutOfMemory ? SWM_E_NOMEM : SWM_E_HIP,         \
               "%s failed: %s", #call, hipGetErrorString(_e));                 \
  } while (0)

#define LAUNCHCK(h)                                                            \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess)                                                      \
      SWM_FAIL(h, SWM_E_HIP, "kernel launch failed: %s", hipGetErrorString(_e)); \
  } while (0)

namespace {

template <class F>
void timed(swm_solver* s, int cat, F&& launch) {
  if (!s->prof.on) {
    launch();
    return;
  }
  hipEvent_t a = s->prof.next(), b = s->prof.next();
  if (a) (void)hipEventRecord(a, s->stream);
  launch();
  if (b) (void)hipEventRecord(b, s->stream);
  if (a && b)
    s->prof.cats.push_back(cat);
  else
    s->prof.used = 2 * s->prof.cats.size();
}

int allocate(swm_solver* s) {
  const Layout& L = s->L;
  s->slab_size = slab_bytes(s->nx, s->ny, s->wb, 0);
  HIPCK(s, hipMalloc((void**)&s->slab, s->slab_size));
  s->h = s->take(L.n);
  s->hu = s->take(L.n);
  s->hv = s->take(L.n);
  for (int m = 1; m < 4; ++m)
    for (int r = 0; r < 3; ++r) s->hbuf[m][r] = s->take(L.n);
  for (int m = 1; m < 4; ++m) {
    s->hun[m] = s->take(L.n);
    s->hvn[m] = s->take(L.n);
  }
  for (int m = 0; m < 4; ++m) {
    s->FU2[m] = s->take(L.n);
    s->FU3[m] = s->take(L.n);
  }
  for (int m = 0; m < 4; ++m) {
    s->qx[m] = s->take(L.nfx);
    s->qy[m] = s->take(L.nfy);
  }
  for (int v = 0; v < 3; ++v) {
    s->FX[v] = s->take(L.nfx);
    s->FY[v] = s->take(L.nfy);
  }
  for (int k = 0; k < kMaxSys; ++k)
    for (int e = 0; e < 5; ++e) s->M[k][e] = s->take(L.n);
  s->stage = s->take(3 * L.n);
  if (s->wb) {
    for (int v = 0; v < 3; ++v) {
      s->FXWB[v] = s->take(L.nfx);
      s->FYWB[v] = s->take(L.nfy);
    }
    s->SWB2 = s->take(L.n);
    s->SWB3 = s->take(L.n);
  }
  s->jac_partial = s->take((size_t)(kMaxSys + 1) * L.nb_jac);
  s->ts_partial = s->take(4 * (size_t)L.nb_ts + 4);
  s->ts_out = s->ts_partial + 4 * (size_t)L.nb_ts;
  s->st = reinterpret_cast<JacState*>(s->slab + s->slab_used);
  s->slab_used += align_up(kMaxSys * sizeof(JacState));
  s->counters = reinterpret_cast<unsigned int*>(s->slab + s->slab_used);
  s->slab_used += align_up(2 * sizeof(unsigned int));
  if (s->slab_used > s->slab_size)
    SWM_FAIL(s, SWM_E_STATE, "internal: slab overflow (%zu > %zu)", s->slab_used, s->slab_size);
  HIPCK(s, hipMemsetAsync(s->slab, 0, s->slab_size, s->stream));
  HIPCK(s, hipHostMalloc((void**)&s->st_host, kMaxSys * sizeof(JacState), hipHostMallocDefault));
  HIPCK(s, hipHostMalloc((void**)&s->ts_host, 4 * sizeof(double), hipHostMallocDefault));
  HIPCK(s, hipStreamSynchronize(s->stream));
  return SWM_OK;
}

// One FVTimeDerivative of state (h, hu, hv) into node r's outputs.
int fvtd(swm_solver* s, const double* h, const double* hu, const double* hv, int r,
         double* S2 = nullptr, double* S3 = nullptr) {
  FvtdIO io{};
  io.h = h;
  io.hu = hu;
  io.hv = hv;
  for (int v = 0; v < 3; ++v) {
    io.FX[v] = s->FX[v];
    io.FY[v] = s->FY[v];
  }
  io.Ut2 = s->FU2[r];
  io.Ut3 = s->FU3[r];
  io.qx = s->qx[r];
  io.qy = s->qy[r];
  io.S2 = S2;
  io.S3 = S3;
  WbArrays wb{};
  if (s->wb) {
    for (int v = 0; v < 3; ++v) {
      wb.FX[v] = s->FXWB[v];
      wb.FY[v] = s->FYWB[v];
    }
    wb.S2 = s->SWB2;
    wb.S3 = s->SWB3;
  }
  Bathy b{};
  if (s->bathy) {
    b.bx = s->bx;
    b.by = s->by;
  }
  timed(s, 0, [&] { launch_fvtd_x(s->G, io, s->stream); });
  timed(s, 1, [&] { launch_fvtd_y(s->G, io, s->stream); });
  timed(s, 2, [&] { launch_fvtd_update(s->G, io, wb, b, false, s->stream); });
  LAUNCHCK(s);
  return SWM_OK;
}

JacCtl jac_ctl(swm_solver* s) {
  JacCtl c{};
  c.st = s->st;
  c.partial = s->jac_partial;
  c.counter = s->counters;
  c.max_iter = s->max_iter;
  c.tol = s->tol;
  return c;
}

// Solves the Patankar systems of nodes ms[0..nsys) (0-based DeC nodes).
// hp_role[m]: -1 = h^n is the previous correction's depth, else hbuf index.
// On return res_role[m] holds the hbuf index of the new depth.
int patankar_solve(swm_solver* s, double dt, const int* ms, int nsys, const int* hp_role,
                   double* const* qx_node, double* const* qy_node, int* res_role,
                   int* iters_out, int& nonconv, int pred) {
  PatSys sys[kMaxSys];
  int pp[kMaxSys][2];
  for (int q = 0; q < nsys; ++q) {
    const int m = ms[q];
    PatSys& P = sys[q];
    P.hp = hp_role[m] < 0 ? s->h : s->hbuf[m][hp_role[m]];
    int a = 0;
    for (int r = 0; r < 3 && a < 2; ++r)
      if (r != hp_role[m]) pp[q][a++] = r;
    P.x[0] = s->hbuf[m][pp[q][0]];
    P.x[1] = s->hbuf[m][pp[q][1]];
    P.MS = s->M[q][0];
    P.MW = s->M[q][1];
    P.ME = s->M[q][2];
    P.MN = s->M[q][3];
    P.D = s->M[q][4];
    for (int r = 0; r < 4; ++r) {
      P.cth[r] = dt * THETA[m][r];
      P.th_pos[r] = THETA[m][r] > 0. ? 1 : 0;
    }
  }
  PatNodes nodes{};
  for (int r = 0; r < 4; ++r) {
    nodes.qx[r] = qx_node[r];
    nodes.qy[r] = qy_node[r];
  }
  JacCtl ctl = jac_ctl(s);
  timed(s, 4, [&] { launch_patankar_assemble(s->G, nodes, sys, nsys, s->h, ctl, s->stream); });
  LAUNCHCK(s);
  int batch = std::max(1, pred - 1);
  long launched = 1;
  for (;;) {
    for (int b = 0; b < batch; ++b)
      timed(s, 5, [&] { launch_jacobi_sweep(s->G, sys, nsys, s->h, ctl, s->stream); });
    LAUNCHCK(s);
    launched += batch;
    HIPCK(s, hipMemcpyAsync(s->st_host, s->st, nsys * sizeof(JacState), hipMemcpyDeviceToHost,
                            s->stream));
    HIPCK(s, hipStreamSynchronize(s->stream));
    bool any = false;
    for (int q = 0; q < nsys; ++q) any = any || s->st_host[q].active;
    if (!any) break;
    if (launched > (long)s->max_iter + 2)
      SWM_FAIL(s, SWM_E_STATE, "internal: Jacobi did not stop after %ld sweeps", launched);
    int kmin = s->max_iter;
    for (int q = 0; q < nsys; ++q)
      if (s->st_host[q].active) kmin = std::min(kmin, s->st_host[q].k);
    batch = std::max(2, kmin / 4);
  }
  bool anyzero = false;
  for (int q = 0; q < nsys; ++q) {
    const JacState& st = s->st_host[q];
    iters_out[q] = st.k;
    nonconv += st.nonconv;
    s->ambiguous += st.ambiguous;
    anyzero = anyzero || st.zero;
    res_role[ms[q]] = pp[q][st.cur];
  }
  if (anyzero) {
    launch_jacobi_zero_fill((int)s->L.n, sys, nsys, ctl, s->stream);
    LAUNCHCK(s);
  }
  return SWM_OK;
}

int check_layout_arg(swm_solver* s, int ghosts) {
  if (ghosts != 0 && ghosts != 3) SWM_FAIL(s, SWM_E_ARG, "ghosts must be 0 or 3 (got %d)", ghosts);
  return SWM_OK;
}

int upload_state(swm_solver* s, const double* U, int ghosts, double* h, double* hu, double* hv) {
  size_t cnt = 3 * (size_t)(s->nx + 2 * ghosts) * (s->ny + 2 * ghosts);
  double* dst = s->stage;
  double* tmp = nullptr;
  if (ghosts) {  // the padded array does not fit the staging area: allocate
    HIPCK(s, hipMalloc((void**)&tmp, cnt * sizeof(double)));
    dst = tmp;
  }
  HIPCK(s, hipMemcpyAsync(dst, U, cnt * sizeof(double), hipMemcpyHostToDevice, s->stream));
  launch_aos_to_soa(s->nx, s->ny, ghosts, dst, h, hu, hv, s->stream);
  LAUNCHCK(s);
  HIPCK(s, hipStreamSynchronize(s->stream));
  if (tmp) HIPCK(s, hipFree(tmp));
  return SWM_OK;
}

}  // namespace

extern "C" {

int swm_create(int nx, int ny, double dx, double dy, double gravity, double cfl, int wellbalanced,
               const double wgp_bnd[4], swm_solver_t** out) {
  if (!out) SWM_FAIL((swm_solver*)nullptr, SWM_E_ARG, "out is NULL");
  *out = nullptr;
  if (nx < 3 || ny < 3) SWM_FAIL((swm_solver*)nullptr, SWM_E_ARG, "nx, ny must be >= 3 (got %d, %d)", nx, ny);
  if ((long long)nx * ny > (1LL << 31) - 1)
    SWM_FAIL((swm_solver*)nullptr, SWM_E_ARG, "grid too large (%d x %d)", nx, ny);
  if (!wgp_bnd) SWM_FAIL((swm_solver*)nullptr, SWM_E_ARG, "wgp_bnd is NULL");
  if (!(dx > 0) || !(dy > 0)) SWM_FAIL((swm_solver*)nullptr, SWM_E_ARG, "dx, dy must be > 0");
  swm_solver* s = new swm_solver();
  s->nx = nx;
  s->ny = ny;
  s->L = layout_of(nx, ny);
  s->G.nx = nx;
  s->G.ny = ny;
  s->G.dx = dx;
  s->G.dy = dy;
  s->G.g = gravity;
  for (int i = 0; i < 4; ++i) s->G.wbnd[i] = wgp_bnd[i];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) s->G.wgp[i][j] = wgp_bnd[i] * wgp_bnd[j];  // src/mesh.f90:136
  s->cfl = cfl;
  s->wb = wellbalanced ? 1 : 0;
  for (int i = 0; i < 13; ++i) s->pred[i] = 4;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("hipStreamCreate failed: ") + hipGetErrorString(e);
    delete s;
    return SWM_E_HIP;
  }
  int rc = allocate(s);
  if (rc != SWM_OK) {
    g_create_error = s->err;
    swm_destroy(s);
    return rc;
  }
  *out = s;
  return SWM_OK;
}

void swm_destroy(swm_solver_t* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->slab) (void)hipFree(s->slab);
  if (s->bx) (void)hipFree(s->bx);
  if (s->st_host) (void)hipHostFree(s->st_host);
  if (s->ts_host) (void)hipHostFree(s->ts_host);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* swm_last_error(const swm_solver_t* s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

long long swm_footprint_bytes(int nx, int ny, int wellbalanced, int bathymetry) {
  if (nx < 3 || ny < 3) return -1;
  return (long long)(slab_bytes(nx, ny, wellbalanced ? 1 : 0, 0) +
                     (bathymetry ? bathy_bytes(nx, ny) : 0));
}

int swm_set_bathy_grad(swm_solver_t* s, const double* dbdx, const double* dbdy) {
  if (!s) return SWM_E_ARG;
  if (!dbdx != !dbdy) SWM_FAIL(s, SWM_E_ARG, "dbdx and dbdy must both be set or both NULL");
  if (!dbdx) {
    s->bathy = 0;
    return SWM_OK;
  }
  size_t b = 16 * s->L.n * sizeof(double);
  if (!s->bx) {
    HIPCK(s, hipMalloc((void**)&s->bx, 2 * b));
    s->by = s->bx + 16 * s->L.n;
  }
  HIPCK(s, hipMemcpyAsync(s->bx, dbdx, b, hipMemcpyHostToDevice, s->stream));
  HIPCK(s, hipMemcpyAsync(s->by, dbdy, b, hipMemcpyHostToDevice, s->stream));
  HIPCK(s, hipStreamSynchronize(s->stream));
  s->bathy = 1;
  s->wb_ready = false;  // the WB reference state depends on the source term
  return SWM_OK;
}

int swm_init_wb(swm_solver_t* s, const double* Uwb, int ghosts) {
  if (!s || !Uwb) return s ? (s->err = "Uwb is NULL", SWM_E_ARG) : SWM_E_ARG;
  if (!s->wb) SWM_FAIL(s, SWM_E_STATE, "handle was created without wellbalanced");
  int rc = check_layout_arg(s, ghosts);
  if (rc) return rc;
  // the WB state goes through the node-1 buffers, which the next step rewrites
  double *h = s->hbuf[1][0], *hu = s->hun[1], *hv = s->hvn[1];
  rc = upload_state(s, Uwb, ghosts, h, hu, hv);
  if (rc) return rc;
  FvtdIO io{};
  io.h = h;
  io.hu = hu;
  io.hv = hv;
  for (int v = 0; v < 3; ++v) {
    io.FX[v] = s->FXWB[v];
    io.FY[v] = s->FYWB[v];
  }
  io.S2 = s->SWB2;
  io.S3 = s->SWB3;
  WbArrays none{};
  Bathy b{};
  if (s->bathy) {
    b.bx = s->bx;
    b.by = s->by;
  }
  launch_fvtd_x(s->G, io, s->stream);
  launch_fvtd_y(s->G, io, s->stream);
  launch_fvtd_update(s->G, io, none, b, true, s->stream);
  LAUNCHCK(s);
  HIPCK(s, hipStreamSynchronize(s->stream));
  s->wb_ready = true;
  return SWM_OK;
}

int swm_upload(swm_solver_t* s, const double* U, int ghosts) {
  if (!s) return SWM_E_ARG;
  if (!U) SWM_FAIL(s, SWM_E_ARG, "U is NULL");
  int rc = check_layout_arg(s, ghosts);
  if (rc) return rc;
  return upload_state(s, U, ghosts, s->h, s->hu, s->hv);
}

int swm_download(swm_solver_t* s, double* U) {
  if (!s) return SWM_E_ARG;
  if (!U) SWM_FAIL(s, SWM_E_ARG, "U is NULL");
  launch_soa_to_aos(s->nx, s->ny, s->h, s->hu, s->hv, s->stage, s->stream);
  LAUNCHCK(s);
  HIPCK(s, hipMemcpyAsync(U, s->stage, 3 * s->L.n * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
  HIPCK(s, hipStreamSynchronize(s->stream));
  return SWM_OK;
}

int swm_timestep(swm_solver_t* s, double* dt_min, double* lmx, double* lmy) {
  if (!s) return SWM_E_ARG;
  TsCtl c{};
  c.partial = s->ts_partial;
  c.counter = s->counters + 1;
  c.out = s->ts_out;
  c.cfl = s->cfl;
  timed(s, 6, [&] { launch_timestep(s->G, s->h, s->hu, s->hv, c, s->stream); });
  LAUNCHCK(s);
  HIPCK(s, hipMemcpyAsync(s->ts_host, s->ts_out, 3 * sizeof(double), hipMemcpyDeviceToHost,
                          s->stream));
  HIPCK(s, hipStreamSynchronize(s->stream));
  if (dt_min) *dt_min = s->ts_host[0];
  if (lmx) *lmx = s->ts_host[1];
  if (lmy) *lmy = s->ts_host[2];
  return SWM_OK;
}

int swm_step_mpdec5(swm_solver_t* s, double t, double dt, int jacobi_iters[13]) {
  (void)t;
  if (!s) return SWM_E_ARG;
  if (s->wb && !s->wb_ready) SWM_FAIL(s, SWM_E_STATE, "well-balanced state not initialised (swm_init_wb)");
  const size_t nb = s->L.n * sizeof(double);
  int iters[13] = {0};
  int nonconv = 0, solve = 0, rc;
  // FVTimeDerivative(U^n) -> node 0, shared by all four nodes at kk = 1
  rc = fvtd(s, s->h, s->hu, s->hv, 0);
  if (rc) return rc;
  const double* FU2p[4] = {s->FU2[0], s->FU2[0], s->FU2[0], s->FU2[0]};
  const double* FU3p[4] = {s->FU3[0], s->FU3[0], s->FU3[0], s->FU3[0]};
  // node PDS pointers (ProdUp/DestUp(m) = PDS(U^n) at kk = 1)
  double* qx_node[4] = {s->qx[0], s->qx[0], s->qx[0], s->qx[0]};
  double* qy_node[4] = {s->qy[0], s->qy[0], s->qy[0], s->qy[0]};
  int hp_role[4] = {-1, -1, -1, -1}, res_role[4] = {-1, -1, -1, -1};
  for (int kk = 1; kk <= 5; ++kk) {
    int ms[3], nsys = 0;
    if (kk == 5)
      ms[nsys++] = 3;
    else
      for (int m = 1; m < 4; ++m) ms[nsys++] = m;
    for (int q = 0; q < nsys; ++q) {
      const int m = ms[q];
      double cth[4];
      for (int r = 0; r < 4; ++r) cth[r] = dt * THETA[m][r];
      timed(s, 3, [&] {
        launch_momentum((int)s->L.n, s->hu, s->hv, FU2p, FU3p, cth, s->hun[m], s->hvn[m],
                        s->stream);
      });
    }
    LAUNCHCK(s);
    int it3[3] = {0, 0, 0};
    int pred = 0;
    for (int q = 0; q < nsys; ++q) pred = std::max(pred, s->pred[solve + q]);
    rc = patankar_solve(s, dt, ms, nsys, hp_role, qx_node, qy_node, res_role, it3, nonconv, pred);
    if (rc) return rc;
    for (int q = 0; q < nsys; ++q) {
      iters[solve + q] = it3[q];
      s->pred[solve + q] = it3[q];
    }
    solve += nsys;
    if (kk != 5) {
      for (int q = 0; q < nsys; ++q) {
        const int m = ms[q];
        rc = fvtd(s, s->hbuf[m][res_role[m]], s->hun[m], s->hvn[m], m);
        if (rc) return rc;
        FU2p[m] = s->FU2[m];
        FU3p[m] = s->FU3[m];
        qx_node[m] = s->qx[m];
        qy_node[m] = s->qy[m];
      }
    }
    for (int q = 0; q < nsys; ++q) hp_role[ms[q]] = res_role[ms[q]];
  }
  // U^{n+1} = Ua(4)
  HIPCK(s, hipMemcpyAsync(s->h, s->hbuf[3][res_role[3]], nb, hipMemcpyDeviceToDevice, s->stream));
  HIPCK(s, hipMemcpyAsync(s->hu, s->hun[3], nb, hipMemcpyDeviceToDevice, s->stream));
  HIPCK(s, hipMemcpyAsync(s->hv, s->hvn[3], nb, hipMemcpyDeviceToDevice, s->stream));
  if (jacobi_iters)
    for (int i = 0; i < 13; ++i) jacobi_iters[i] = iters[i];
  return nonconv;
}

int swm_run_steps(swm_solver_t* s, int nsteps, double* dts, int* jacobi_iters) {
  if (!s) return SWM_E_ARG;
  if (nsteps < 0) SWM_FAIL(s, SWM_E_ARG, "nsteps < 0");
  int nonconv = 0;
  for (int k = 0; k < nsteps; ++k) {
    double dt;
    int rc = swm_timestep(s, &dt, nullptr, nullptr);
    if (rc) return rc;
    if (dts) dts[k] = dt;
    if (!std::isfinite(dt)) SWM_FAIL(s, SWM_E_STATE, "time step %d: dt is not finite (%g)", k, dt);
    rc = swm_step_mpdec5(s, 0.0, dt, jacobi_iters ? jacobi_iters + 13 * k : nullptr);
    if (rc < 0) return rc;
    nonconv += rc;
  }
  HIPCK(s, hipStreamSynchronize(s->stream));
  return nonconv;
}

int swm_fvtd(swm_solver_t* s, double* Ut, double* FX, double* FY, double* S, double* Prod,
             double* Dest) {
  if (!s) return SWM_E_ARG;
  if (s->wb && !s->wb_ready) SWM_FAIL(s, SWM_E_STATE, "well-balanced state not initialised");
  const int nx = s->nx, ny = s->ny;
  const size_t n = s->L.n, nfx = s->L.nfx, nfy = s->L.nfy;
  // S(2:3) goes to the node-2 matrix scratch, which no step keeps between calls
  double* S2 = s->M[0][0];
  double* S3 = s->M[0][1];
  int rc = fvtd(s, s->h, s->hu, s->hv, 0, S2, S3);
  if (rc) return rc;
  std::vector<double> ut2(n), ut3(n), s2(n), s3(n), qx(nfx), qy(nfy);
  std::vector<double> fx[3], fy[3];
  for (int v = 0; v < 3; ++v) {
    fx[v].resize(nfx);
    fy[v].resize(nfy);
    HIPCK(s, hipMemcpyAsync(fx[v].data(), s->FX[v], nfx * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCK(s, hipMemcpyAsync(fy[v].data(), s->FY[v], nfy * 8, hipMemcpyDeviceToHost, s->stream));
  }
  HIPCK(s, hipMemcpyAsync(ut2.data(), s->FU2[0], n * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCK(s, hipMemcpyAsync(ut3.data(), s->FU3[0], n * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCK(s, hipMemcpyAsync(s2.data(), S2, n * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCK(s, hipMemcpyAsync(s3.data(), S3, n * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCK(s, hipMemcpyAsync(qx.data(), s->qx[0], nfx * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCK(s, hipMemcpyAsync(qy.data(), s->qy[0], nfy * 8, hipMemcpyDeviceToHost, s->stream));
  std::vector<double> fxwb[3], fywb[3];
  if (s->wb)
    for (int v = 0; v < 3; ++v) {
      fxwb[v].resize(nfx);
      fywb[v].resize(nfy);
      HIPCK(s, hipMemcpyAsync(fxwb[v].data(), s->FXWB[v], nfx * 8, hipMemcpyDeviceToHost, s->stream));
      HIPCK(s, hipMemcpyAsync(fywb[v].data(), s->FYWB[v], nfy * 8, hipMemcpyDeviceToHost, s->stream));
    }
  HIPCK(s, hipStreamSynchronize(s->stream));
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
  // the reference's module FX/FY after UpdateTimeDerivative (WB subtracted)
  if (FX)
    for (int j = 0; j < ny; ++j)
      for (int f = 0; f <= nx; ++f) {
        size_t k = (size_t)j * (nx + 1) + f;
        for (int v = 0; v < 3; ++v) FX[3 * k + v] = s->wb ? fx[v][k] - fxwb[v][k] : fx[v][k];
      }
  if (FY)
    for (int g = 0; g <= ny; ++g)
      for (int i = 0; i < nx; ++i) {
        size_t k = (size_t)g * nx + i;
        for (int v = 0; v < 3; ++v) FY[3 * k + v] = s->wb ? fy[v][k] - fywb[v][k] : fy[v][k];
      }
  if (Prod || Dest)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t k = (size_t)j * nx + i;
        double q[5] = {qy[(size_t)j * nx + i], qx[(size_t)j * (nx + 1) + i], 0.0,
                       qx[(size_t)j * (nx + 1) + i + 1], qy[(size_t)(j + 1) * nx + i]};
        for (int sl = 0; sl < 5; ++sl) {
          double P, D;
          if (sl == 2) {
            P = 0.;
            D = 0.;
          } else if (sl < 2) {
            P = (q[sl] > 0.) ? q[sl] : 0.;
            D = (q[sl] < 0.) ? -q[sl] : -0.;
          } else {
            P = (q[sl] < 0.) ? -q[sl] : 0.;
            D = (q[sl] > 0.) ? q[sl] : -0.;
          }
          if (Prod) Prod[5 * k + sl] = P;
          if (Dest) Dest[5 * k + sl] = D;
        }
      }
  return SWM_OK;
}

int swm_set_jacobi_limits(swm_solver_t* s, int max_iter, double tol) {
  if (!s) return SWM_E_ARG;
  if (max_iter < 0) SWM_FAIL(s, SWM_E_ARG, "max_iter < 0");
  s->max_iter = max_iter;
  s->tol = tol;
  return SWM_OK;
}

long long swm_ambiguous_decisions(const swm_solver_t* s) { return s ? s->ambiguous : -1; }

int swm_profile_enable(swm_solver_t* s, int on) {
  if (!s) return SWM_E_ARG;
  s->prof.on = on != 0;
  return SWM_OK;
}

int swm_profile_read(swm_solver_t* s, double ms[SWM_PROF_NCAT], long long launches[SWM_PROF_NCAT]) {
  if (!s || !ms || !launches) return SWM_E_ARG;
  HIPCK(s, hipStreamSynchronize(s->stream));
  for (int c = 0; c < SWM_PROF_NCAT; ++c) {
    ms[c] = 0.0;
    launches[c] = 0;
  }
  for (size_t k = 0; k < s->prof.cats.size(); ++k) {
    float t = 0.f;
    HIPCK(s, hipEventElapsedTime(&t, s->prof.pool[2 * k], s->prof.pool[2 * k + 1]));
    ms[s->prof.cats[k]] += t;
    launches[s->prof.cats[k]] += 1;
  }
  s->prof.cats.clear();
  s->prof.used = 0;
  return SWM_OK;
}

const char* swm_profile_name(int cat) {
  return (cat >= 0 && cat < SWM_PROF_NCAT) ? PROF_NAMES[cat] : "";
}

int swm_sync(swm_solver_t* s) {
  if (!s) return SWM_E_ARG;
  HIPCK(s, hipStreamSynchronize(s->stream));
  return SWM_OK;
}

int swm_debug_divsqrt(int n, const double* a, const double* b, double* q, double* r) {
  if (n <= 0 || !a || !b || !q || !r) return SWM_E_ARG;
  double *da = nullptr;
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

}  // extern "C"
