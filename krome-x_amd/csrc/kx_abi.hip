// The C ABI of one network's library (include/krome.h): KROME-X's
// krome_* entry points (runtime/interfaces/C/krome.h, krome.cpp) with the
// batched solves on gfx950.  Compiled once per network with
// -DKX_NET=net_<name> -DKX_NET_HEADER="net_<name>.h".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include KX_NET_HEADER
#include "kx_plain.h"
#include "kx_persist.h"
#include "../../include/krome.h"

using Net = kx::KX_NET;

namespace {

constexpr int NS = Net::NSPEC;
const char* const kVersion = "2.1.0-mi355x";

// network metadata
#include KX_META_HEADER

struct Opts {
  // solver options (SolverOpts defaults, krome.hpp:187-255)
  bool evolve_T = false;
  int jac = 0;
  int linsol = 0;
  double rtol = 1e-8, atol_n = 1e-20, atol_T = 1e-6, atol_u = 1e-10;
  long max_steps = 100000;
  int max_err_fails = 7, max_conv_fails = 10, max_restarts = 5;
  bool clip_negative = true, renormalize_x = false, error_report = true;
  int conserve = 0;
  // batch options (CellsOpts, krome.hpp:272-332)
  int path = 0;
  bool device_cache = true;
  long long device_memory = 0;
  int min_device = 10000, min_interleave = -1;
  // this build's device kernels: 0 the plain FP64 kernel, 1 the
  // persistent kernel; its Jacobian / LU precision (0 FP64, 1 FP32)
  int kernel = 0;
  int mixed = 0;
  // persistent kernel: cells per queue grab, workgroups per CU
  int queue_chunk = 0;
  int waves_per_cu = 0;
};

struct Flag {
  int v;
  const char* name;
  const char* msg;
};
const Flag kFlags[] = {
    {0, "FLAG_OK", "success"},
    {-1, "FLAG_TOO_MUCH_WORK", "max_steps exhausted before t_end"},
    {-2, "FLAG_TOO_MUCH_ACC", "tolerances too small for machine precision"},
    {-3, "FLAG_ERR_FAILURE", "repeated error-test failures"},
    {-4, "FLAG_CONV_FAILURE", "repeated Newton convergence failures"},
    {-6, "FLAG_LSETUP_FAIL", "Newton matrix setup / factorization failed"},
    {-7, "FLAG_LSOLVE_FAIL", "linear solve failed"},
    {-8, "FLAG_RHSFUNC_FAIL", "right-hand side failed"},
    {-9, "FLAG_FIRST_RHSFUNC_ERR",
     "right-hand side failed at the initial state"},
    {-10, "FLAG_REPTD_RHSFUNC_ERR",
     "repeated recoverable right-hand side errors"},
    {-11, "FLAG_UNREC_RHSFUNC_ERR", "right-hand side failed unrecoverably"},
    {-16, "FLAG_NLS_FAIL", "nonlinear solver failure"},
    {-22, "FLAG_ILL_INPUT", "illegal input"},
    {-27, "FLAG_TOO_CLOSE", "t_end too close to the start"},
    {-102, "FLAG_NO_DUST", "cooling DUSTBINS with Tgas evolved needs dust"},
    {-103, "FLAG_NO_FIELD", "photo reactions need a field Phi"},
    {-104, "FLAG_UNSUPPORTED",
     "configuration not supported by the requested path, or backend not "
     "built"},
    {-105, "FLAG_PARAM_UNSET", "a required parameter is unset (NaN)"},
    {-106, "FLAG_CONSERVE_GAS_TOTAL",
     "conservation is not available with dust.gas_total"},
    {-107, "FLAG_SURFACE_BINS", "surface.bins needs dust bins"},
    {-108, "FLAG_NO_LAPACK", "LINSOL_LAPACK in a build without LAPACK"},
    {-109, "FLAG_NO_KOKKOS", "device path used before initialization"},
    {-110, "FLAG_NONPOSITIVE_T",
     "the integration ended at a non-positive or non-finite Tgas (input "
     "untouched)"},
    {-111, "FLAG_DEVICE_MEMORY",
     "device path: not even one cell fits in the device memory"},
    {-112, "FLAG_CONSERVE_UNREACHED",
     "consistent advection: the reference cannot be reached"},
    {1, "FLAG_HOST_RECOVERED", "failed on the device, solved on the host"},
    {2, "FLAG_CONSERVE_LARGE", "consistent advection: a large correction"},
};

kx::SolveOpts solve_opts(const krome_opts* h);

// ---- device state ------------------------------------------------------

std::mutex g_dev_mutex;
bool g_dev_init = false;

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "krome (HIP): %s: %s\n", what, hipGetErrorString(e));
  return false;
}

// device buffers, grown as needed and kept (CellsOpts::device_cache)
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  bool ensure(size_t b) {
    if (b <= bytes) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (!hip_ok(hipMalloc(&p, b), "hipMalloc")) return false;
    bytes = b;
    return true;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};
struct DevCache {
  Buf work, iwork, n, T, dt, par, flags, stats, queue, count;
  void release() {
    for (Buf* b : {&work, &iwork, &n, &T, &dt, &par, &flags, &stats, &queue,
                   &count})
      b->release();
  }
};
DevCache g_cache;
std::mutex g_cache_mutex;

template <bool ET>
constexpr int neq() { return kx::Problem<Net, ET>::N; }

// bytes of the plain kernel's workspace per cell
size_t plain_cell_bytes(bool et) {
  const int N = et ? neq<true>() : neq<false>();
  return sizeof(double) * ((kx::LMAX + 9) * N + 2 * N * N) + sizeof(int) * N;
}

// the plain kernel over a chunk of resident arrays
template <bool ET>
int launch_plain(kx::CellsArgs a, DevCache& dc, hipStream_t s) {
  constexpr int N = neq<ET>();
  const long S = a.ncell;
  if (!dc.work.ensure(sizeof(double) * kx::cell_doubles<N>() * S)) return -1;
  if (!dc.iwork.ensure(sizeof(int) * N * S)) return -1;
  const int tpb = 128;
  const int blocks = (a.ncell + tpb - 1) / tpb;
  hipLaunchKernelGGL((kx::plain_cells_kernel<Net, ET>), dim3(blocks),
                     dim3(tpb), 0, s, a, static_cast<double*>(dc.work.p),
                     static_cast<int*>(dc.iwork.p), S);
  return hip_ok(hipGetLastError(), "plain kernel launch") ? 0 : -1;
}

// HIP events around the last solve's kernel launches, on their stream
hipEvent_t g_ev[2] = {nullptr, nullptr};
bool g_ev_recorded = false;

int solve_resident_(kx::CellsArgs a, const Opts& o, DevCache& dc,
                    hipStream_t s);

// the cells of one device call, in chunks that fit the memory budget:
// arrays already resident; flags always written
int solve_resident(kx::CellsArgs a, const Opts& o, DevCache& dc,
                   hipStream_t s) {
  if (!g_ev[0] && (hipEventCreate(&g_ev[0]) != hipSuccess ||
                   hipEventCreate(&g_ev[1]) != hipSuccess))
    g_ev[0] = g_ev[1] = nullptr;
  if (g_ev[0]) (void)hipEventRecord(g_ev[0], s);
  const int r = solve_resident_(a, o, dc, s);
  if (g_ev[0]) {
    (void)hipEventRecord(g_ev[1], s);
    g_ev_recorded = true;
  }
  return r;
}

int solve_resident_(kx::CellsArgs a, const Opts& o, DevCache& dc,
                    hipStream_t s) {
  if (a.ncell <= 0) return 0;
  if (a.o.evolve_T && !Net::THERMAL) {
    return kx::FLAG_UNSUPPORTED;
  }
  if (o.kernel == 1) {
    if constexpr (Net::THERMAL)
      if (a.o.evolve_T)
        return kx::launch_persistent<Net, true>(a, o.mixed, o.queue_chunk,
                                                o.waves_per_cu, s);
    return kx::launch_persistent<Net, false>(a, o.mixed, o.queue_chunk,
                                             o.waves_per_cu, s);
  }
  // chunking of the plain kernel by the memory budget
  size_t budget;
  if (o.device_memory < 0) {
    budget = (size_t)-1;
  } else if (o.device_memory > 0) {
    budget = (size_t)o.device_memory;
  } else {
    size_t fr = 0, tot = 0;
    if (!hip_ok(hipMemGetInfo(&fr, &tot), "hipMemGetInfo")) return -1;
    budget = (size_t)(0.8 * (double)(fr + dc.work.bytes + dc.iwork.bytes));
  }
  const size_t per = plain_cell_bytes(a.o.evolve_T != 0);
  long chunk = (long)(budget / per);
  if (chunk < 1) return kx::FLAG_DEVICE_MEMORY;
  if (chunk > a.ncell) chunk = a.ncell;
  for (long c0 = 0; c0 < a.ncell; c0 += chunk) {
    kx::CellsArgs b = a;
    b.ncell = (int)((a.ncell - c0) < chunk ? (a.ncell - c0) : chunk);
    b.n = a.n + c0 * a.n_cs;
    b.T = a.T + c0;
    b.dt = a.dt + c0;
    b.par = a.par + c0 * a.par_stride;
    b.flags = a.flags + c0;
    b.stats = a.stats ? a.stats + c0 * 9 : nullptr;
    int r;
    if constexpr (Net::THERMAL)
      r = a.o.evolve_T ? launch_plain<true>(b, dc, s)
                       : launch_plain<false>(b, dc, s);
    else
      r = launch_plain<false>(b, dc, s);
    if (r != 0) return r;
  }
  return 0;
}

}  // namespace

// ---- description -------------------------------------------------------

extern "C" {

const char* krome_flag_message(int flag) {
  for (const Flag& f : kFlags)
    if (f.v == flag) return f.msg;
  if (flag > 0) return "not converged: t_max reached first";
  return flag <= -100 ? "unknown KROME flag" : "integrator failure";
}
const char* krome_flag_name(int flag) {
  for (const Flag& f : kFlags)
    if (f.v == flag) return f.name;
  return "?";
}
const char* krome_version(void) { return kVersion; }
const char* krome_network_name(void) { return Net::NAME; }
int krome_nspec(void) { return Net::NSPEC; }
int krome_nrea(void) { return Net::NREA; }
int krome_npar(void) { return Net::NPAR; }
const char* krome_species_name(int i) {
  return (i >= 0 && i < NS) ? kSpecies[i] : nullptr;
}
int krome_species_index(const char* name) {
  if (!name) return -1;
  for (int i = 0; i < NS; ++i)
    if (!std::strcmp(kSpecies[i], name)) return i;
  return -1;
}
const char* krome_param_name(int i) {
  return (i >= 0 && i < Net::NPAR) ? kParams[i] : nullptr;
}
int krome_param_index(const char* name) {
  if (!name) return -1;
  for (int i = 0; i < Net::NPAR; ++i)
    if (!std::strcmp(kParams[i], name)) return i;
  return -1;
}
void krome_params_default(double* par) {
  for (int i = 0; i < Net::NPAR; ++i) par[i] = kParamDefaults[i];
}
const char* krome_param_required_by(int i) {
  if (i < 0 || i >= Net::NPAR) return nullptr;
  return std::isnan(kParamDefaults[i]) ? kParamRequiredBy[i] : "";
}

// ---- options -----------------------------------------------------------

struct krome_opts : Opts {};

krome_opts* krome_opts_new(void) { return new krome_opts(); }
void krome_opts_free(krome_opts* h) { delete h; }

int krome_opts_set(krome_opts* h, const char* key, double v) {
  if (!h || !key) return -1;
  auto is = [&](const char* k) { return !std::strcmp(key, k); };
  if (is("evolve_T")) { h->evolve_T = v != 0.0; return 0; }
  if (is("clip_negative")) { h->clip_negative = v != 0.0; return 0; }
  if (is("renormalize_x")) { h->renormalize_x = v != 0.0; return 0; }
  if (is("error_report")) { h->error_report = v != 0.0; return 0; }
  if (is("device_cache")) { h->device_cache = v != 0.0; return 0; }
  if (is("rtol")) { h->rtol = v; return 0; }
  if (is("atol_n")) { h->atol_n = v; return 0; }
  if (is("atol_T")) { h->atol_T = v; return 0; }
  if (is("atol_u")) { h->atol_u = v; return 0; }
  if (is("max_steps")) { h->max_steps = (long)v; return 0; }
  if (is("max_err_fails") || is("max_conv_fails") || is("max_restarts") ||
      is("conserve")) {
    if (v < 0.0) return -1;
    if (is("conserve") && v > 3.0) return -1;
    int* slot = is("max_err_fails") ? &h->max_err_fails
                : is("max_conv_fails") ? &h->max_conv_fails
                : is("max_restarts") ? &h->max_restarts : &h->conserve;
    *slot = (int)v;
    return 0;
  }
  if (is("jac")) {
    if (v < 0 || v > 2) return -1;
    h->jac = (int)v;
    return 0;
  }
  if (is("linsol")) {
    if (v < 0 || v > 3) return -1;
    h->linsol = (int)v;
    return 0;
  }
  if (is("path")) {
    if (v < 0 || v > 2) return -1;
    h->path = (int)v;
    return 0;
  }
  if (is("device_memory")) { h->device_memory = (long long)v; return 0; }
  if (is("min_device")) { h->min_device = (int)v; return 0; }
  if (is("min_interleave")) { h->min_interleave = (int)v; return 0; }
  if (is("device_kernel")) {
    if (v < 0 || v > 1) return -1;
    h->kernel = (int)v;
    return 0;
  }
  if (is("mixed_precision")) {
    if (v < 0 || v > 1) return -1;
    h->mixed = (int)v;
    return 0;
  }
  if (is("queue_chunk")) {
    if (v < 0) return -1;
    h->queue_chunk = (int)v;
    return 0;
  }
  if (is("waves_per_cu")) {
    if (v < 0) return -1;
    h->waves_per_cu = (int)v;
    return 0;
  }
  return -1;
}

int krome_opts_get(const krome_opts* h, const char* key, double* v) {
  if (!h || !key || !v) return -1;
  auto is = [&](const char* k) { return !std::strcmp(key, k); };
  if (is("evolve_T")) *v = h->evolve_T;
  else if (is("clip_negative")) *v = h->clip_negative;
  else if (is("renormalize_x")) *v = h->renormalize_x;
  else if (is("error_report")) *v = h->error_report;
  else if (is("device_cache")) *v = h->device_cache;
  else if (is("rtol")) *v = h->rtol;
  else if (is("atol_n")) *v = h->atol_n;
  else if (is("atol_T")) *v = h->atol_T;
  else if (is("atol_u")) *v = h->atol_u;
  else if (is("max_steps")) *v = (double)h->max_steps;
  else if (is("max_err_fails")) *v = h->max_err_fails;
  else if (is("max_conv_fails")) *v = h->max_conv_fails;
  else if (is("max_restarts")) *v = h->max_restarts;
  else if (is("conserve")) *v = h->conserve;
  else if (is("jac")) *v = h->jac;
  else if (is("linsol")) *v = h->linsol;
  else if (is("path")) *v = h->path;
  else if (is("device_memory")) *v = (double)h->device_memory;
  else if (is("min_device")) *v = h->min_device;
  else if (is("min_interleave")) *v = h->min_interleave;
  else if (is("device_kernel")) *v = h->kernel;
  else if (is("mixed_precision")) *v = h->mixed;
  else if (is("queue_chunk")) *v = h->queue_chunk;
  else if (is("waves_per_cu")) *v = h->waves_per_cu;
  else return -1;
  return 0;
}

// ---- device ------------------------------------------------------------

int krome_has_device(void) { return 1; }

int krome_device_initialize_id(int device_id) {
  std::lock_guard<std::mutex> lk(g_dev_mutex);
  if (g_dev_init) return 1;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return kx::FLAG_UNSUPPORTED;
  if (device_id >= 0) {
    if (device_id >= n) return kx::FLAG_UNSUPPORTED;
    if (!hip_ok(hipSetDevice(device_id), "hipSetDevice"))
      return kx::FLAG_UNSUPPORTED;
  }
  g_dev_init = true;
  return 0;
}
int krome_device_initialize(void) { return krome_device_initialize_id(-1); }
int krome_device_finalize(void) {
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  g_cache.release();
  std::lock_guard<std::mutex> lk2(g_dev_mutex);
  g_dev_init = false;
  return 0;
}

// the duration of the last device solve's kernels (HIP events recorded
// around their launches on their stream), waiting for them to finish
int krome_device_last_kernel_ms(double* ms) {
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  if (!ms || !g_ev_recorded) return -1;
  float f = 0.0f;
  if (!hip_ok(hipEventSynchronize(g_ev[1]), "hipEventSynchronize") ||
      !hip_ok(hipEventElapsedTime(&f, g_ev[0], g_ev[1]),
              "hipEventElapsedTime"))
    return -1;
  *ms = (double)f;
  return 0;
}

}  // extern "C"

namespace {

kx::SolveOpts solve_opts(const krome_opts* h) {
  kx::SolveOpts o;
  o.rtol = h->rtol;
  o.atol_n = h->atol_n;
  o.atol_T = h->atol_T;
  o.max_steps = h->max_steps;
  o.max_err_fails = h->max_err_fails;
  o.max_conv_fails = h->max_conv_fails;
  o.max_restarts = h->max_restarts;
  o.evolve_T = h->evolve_T ? 1 : 0;
  o.jac = h->jac;
  o.clip_negative = h->clip_negative ? 1 : 0;
  o.conserve = h->conserve;
  return o;
}

// what the whole call rejects before integrating (every cell flagged)
int reject(const krome_opts* h) {
  if (h->conserve == 2) return kx::FLAG_UNSUPPORTED;
  if (h->linsol == 2 || h->linsol == 3) return kx::FLAG_UNSUPPORTED;
  if (h->path == 1) return kx::FLAG_UNSUPPORTED;  // no host path here
  if (h->evolve_T && !Net::THERMAL) return kx::FLAG_UNSUPPORTED;
  return 0;
}

// the lowest failing cell's flag
int first_flag(const int* fl, int ncell) {
  for (int c = 0; c < ncell; ++c)
    if (fl[c] != 0) return fl[c];
  return 0;
}

}  // namespace

extern "C" {

int krome_solve_cells(int ncell, double* n, double* Tgas, const double* dt,
                      const double* par, int par_stride,
                      const krome_cells* extra, int* flags,
                      const krome_opts* h) {
  (void)extra;  // no photon bins or dust bins in this build's networks
  if (!h || ncell < 0 || !n || !Tgas || !dt) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (ncell == 0) return 0;
  std::vector<int> fl(ncell, 0);
  // parameters: rows, or the defaults; a REQUIRED one unset: -105 for all
  std::vector<double> dflt(Net::NPAR);
  krome_params_default(dflt.data());
  const double* P = par ? par : dflt.data();
  const int pst = par ? par_stride : 0;
  bool missing = false;
  for (int c = 0; c < ncell; ++c) {
    const bool m = Net::params_missing(P + (long)c * pst) >= 0;
    fl[c] = m ? kx::FLAG_PARAM_UNSET : 0;
    missing |= m;
  }
  int rj = missing ? kx::FLAG_PARAM_UNSET : reject(h);
  if (rj != 0) {
    for (int c = 0; c < ncell; ++c)
      fl[c] = missing ? fl[c] : rj;
    if (flags) std::memcpy(flags, fl.data(), sizeof(int) * ncell);
    return rj;
  }
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache local;
  DevCache& dc = h->device_cache ? g_cache : local;
  const size_t nb = sizeof(double) * (size_t)ncell * NS;
  const size_t pb = sizeof(double) * (size_t)(pst ? ncell : 1) * Net::NPAR;
  if (!dc.n.ensure(nb) || !dc.T.ensure(sizeof(double) * ncell) ||
      !dc.dt.ensure(sizeof(double) * ncell) || !dc.par.ensure(pb) ||
      !dc.flags.ensure(sizeof(int) * ncell))
    return kx::FLAG_DEVICE_MEMORY;
  hipStream_t s = nullptr;
  bool ok = hip_ok(hipMemcpy(dc.n.p, n, nb, hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.T.p, Tgas, sizeof(double) * ncell,
                             hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.dt.p, dt, sizeof(double) * ncell,
                             hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.par.p, P, pb, hipMemcpyHostToDevice), "H2D");
  if (!ok) return -1;
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = static_cast<double*>(dc.n.p);
  a.n_cs = NS;
  a.n_ss = 1;
  a.T = static_cast<double*>(dc.T.p);
  a.dt = static_cast<const double*>(dc.dt.p);
  a.par = static_cast<const double*>(dc.par.p);
  a.par_stride = pst;
  a.flags = static_cast<int*>(dc.flags.p);
  a.stats = nullptr;
  a.o = solve_opts(h);
  int r = solve_resident(a, *h, dc, s);
  if (r == 0) {
    ok = hip_ok(hipDeviceSynchronize(), "kernel") &&
         hip_ok(hipMemcpy(n, dc.n.p, nb, hipMemcpyDeviceToHost), "D2H") &&
         hip_ok(hipMemcpy(Tgas, dc.T.p, sizeof(double) * ncell,
                          hipMemcpyDeviceToHost), "D2H") &&
         hip_ok(hipMemcpy(fl.data(), dc.flags.p, sizeof(int) * ncell,
                          hipMemcpyDeviceToHost), "D2H");
    if (!ok) r = -1;
  }
  if (!h->device_cache) local.release();
  if (r != 0) {
    if (r == kx::FLAG_DEVICE_MEMORY && flags)
      for (int c = 0; c < ncell; ++c) flags[c] = r;
    return r;
  }
  if (flags) std::memcpy(flags, fl.data(), sizeof(int) * ncell);
  return first_flag(fl.data(), ncell);
}

int krome_solve_cells_x(int ncell, double* x, const double* rho,
                        double* Tgas, const double* dt, const double* par,
                        int par_stride, const krome_cells* extra, int* flags,
                        const krome_opts* h) {
  if (!h || ncell < 0 || !x || !rho || !Tgas || !dt) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (ncell == 0) return 0;
  std::vector<double> n((size_t)ncell * NS), xin(ncell);
  for (int c = 0; c < ncell; ++c) {
    double s = 0.0;
    for (int i = 0; i < NS; ++i) {
      const double m = Net::mass_g(i);
      const double xi = x[(size_t)c * NS + i];
      if (m > 0.0) s += xi;
      n[(size_t)c * NS + i] = m > 0.0 ? rho[c] * xi / m : 0.0;
    }
    xin[c] = s;
  }
  std::vector<int> fl(ncell, 0);
  int* f = flags ? flags : fl.data();
  const int r = krome_solve_cells(ncell, n.data(), Tgas, dt, par, par_stride,
                                  extra, f, h);
  for (int c = 0; c < ncell; ++c) {
    if (f[c] != 0) continue;
    double* xc = x + (size_t)c * NS;
    double s = 0.0;
    for (int i = 0; i < NS; ++i) {
      xc[i] = Net::mass_g(i) * n[(size_t)c * NS + i] / rho[c];
      if (Net::mass_g(i) > 0.0) s += xc[i];
    }
    if (h->renormalize_x && s > 0.0)
      for (int i = 0; i < NS; ++i) xc[i] *= xin[c] / s;
  }
  return r;
}

// device-resident arrays (krome.h: krome_solve_cells_device): n in the
// species-major (n[i ld + c]) or cell-major layout, everything on the
// device; the call synchronizes before and after
int krome_solve_cells_device(int ncell, double* n, int layout, long ld,
                             double* Tgas, const double* dt,
                             const double* par, int par_stride,
                             const krome_cells* extra, int* flags,
                             krome_device_summary* summary,
                             const krome_opts* h) {
  (void)extra;
  if (!h || ncell < 0 || !n || !Tgas || !dt) return -1;
  if (layout != KROME_LAYOUT_SPECIES_MAJOR &&
      layout != KROME_LAYOUT_CELL_MAJOR)
    return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (layout == KROME_LAYOUT_SPECIES_MAJOR && ld != 0 && ld < ncell)
    return -1;
  if (ncell == 0) return 0;
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  // parameters: the defaults on the device when none given
  const double* P = par;
  int pst = par ? par_stride : 0;
  if (!P) {
    std::vector<double> d(Net::NPAR);
    krome_params_default(d.data());
    if (!dc.par.ensure(sizeof(double) * Net::NPAR)) return kx::FLAG_DEVICE_MEMORY;
    if (!hip_ok(hipMemcpy(dc.par.p, d.data(), sizeof(double) * Net::NPAR,
                          hipMemcpyHostToDevice), "H2D"))
      return -1;
    P = static_cast<const double*>(dc.par.p);
  }
  int* fl = flags;
  if (!fl) {
    if (!dc.flags.ensure(sizeof(int) * ncell)) return kx::FLAG_DEVICE_MEMORY;
    fl = static_cast<int*>(dc.flags.p);
  }
  if (!dc.count.ensure(sizeof(int))) return kx::FLAG_DEVICE_MEMORY;
  const int np = kx::params_check_device<Net>(ncell, P, pst, fl,
                                              static_cast<int*>(dc.count.p));
  if (np < 0) return -1;
  if (np > 0) return kx::FLAG_PARAM_UNSET;
  const int rj = reject(h);
  if (rj != 0) {
    kx::fill_flags(ncell, fl, rj);
    return rj;
  }
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = n;
  if (layout == KROME_LAYOUT_SPECIES_MAJOR) {
    a.n_cs = 1;
    a.n_ss = ld ? ld : ncell;
  } else {
    a.n_cs = NS;
    a.n_ss = 1;
  }
  a.T = Tgas;
  a.dt = dt;
  a.par = P;
  a.par_stride = pst;
  a.flags = fl;
  a.stats = nullptr;
  a.o = solve_opts(h);
  int r = solve_resident(a, *h, dc, nullptr);
  if (r != 0) return r;
  if (!hip_ok(hipDeviceSynchronize(), "kernel")) return -1;
  std::vector<int> hf(ncell);
  if (!hip_ok(hipMemcpy(hf.data(), fl, sizeof(int) * ncell,
                        hipMemcpyDeviceToHost), "D2H"))
    return -1;
  if (summary) {
    summary->device = 0;
    summary->host = 0;
    summary->failed = 0;
    for (int c = 0; c < ncell; ++c) {
      if (hf[c] == 0) summary->device++;
      else summary->failed++;
    }
  }
  return first_flag(hf.data(), ncell);
}

int krome_solve_cells_x_device(int ncell, double* x, const double* rho,
                               int layout, long ld, double* Tgas,
                               const double* dt, const double* par,
                               int par_stride, const krome_cells* extra,
                               int* flags, krome_device_summary* summary,
                               const krome_opts* h) {
  (void)ncell; (void)x; (void)rho; (void)layout; (void)ld; (void)Tgas;
  (void)dt; (void)par; (void)par_stride; (void)extra; (void)flags;
  (void)summary; (void)h;
  return kx::FLAG_UNSUPPORTED;
}

// per-cell statistics of the last kind of call, for the benchmarks and the
// tests (this build's extension: the counters of krome::SolveStats,
// krome.hpp:257-268, 9 per cell: steps, rhs_evals, jac_evals,
// rhs_evals_jac, lin_setups, newton_iters, newton_fails, err_fails,
// restarts); stats on the device, ncell x 9
int krome_solve_cells_device_stats(int ncell, double* n, int layout, long ld,
                                   double* Tgas, const double* dt,
                                   const double* par, int par_stride,
                                   int* flags, long* stats,
                                   const krome_opts* h) {
  if (!h || ncell < 0 || !n || !Tgas || !dt || !flags) return -1;
  if (par == nullptr) return -1;
  if (par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (layout == KROME_LAYOUT_SPECIES_MAJOR && ld != 0 && ld < ncell)
    return -1;
  if (ncell == 0) return 0;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  if (!g_cache.count.ensure(sizeof(int))) return -1;
  const int np = kx::params_check_device<Net>(
      ncell, par, par_stride, flags, static_cast<int*>(g_cache.count.p));
  if (np < 0) return -1;
  if (np > 0) return kx::FLAG_PARAM_UNSET;
  const int rj = reject(h);
  if (rj != 0) {
    kx::fill_flags(ncell, flags, rj);
    return rj;
  }
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = n;
  if (layout == KROME_LAYOUT_SPECIES_MAJOR) {
    a.n_cs = 1;
    a.n_ss = ld ? ld : ncell;
  } else {
    a.n_cs = NS;
    a.n_ss = 1;
  }
  a.T = Tgas;
  a.dt = dt;
  a.par = par;
  a.par_stride = par_stride;
  a.flags = flags;
  a.stats = stats;
  a.o = solve_opts(h);
  return solve_resident(a, *h, g_cache, nullptr);
}

}  // extern "C"
