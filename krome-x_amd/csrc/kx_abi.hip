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
#include "kx_conserve.h"
#include "kx_massfrac.h"
#include "kx_query.h"
#include "kx_photo.h"
#include "kx_adjoint.h"
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
  // persistent kernel: cells per queue grab, waves per CU; the lane
  // queue's lanes per workgroup (device_kernel 2, and 1 above 16
  // equations)
  int queue_chunk = 0;
  int waves_per_cu = 0;
  int lane_block = 128;
  // a wave per cell (networks above 64 species, device_kernel 1): cells
  // per workgroup, 1 to 16 (0: 4)
  int cells_per_workgroup = 0;
  // the adjoint: accepted steps between checkpoints, the checkpoint
  // pool's bytes (0: room for about 1536 steps per cell of a launch)
  int checkpoint_steps = 32;
  long long checkpoint_memory = 0;
  // photon bins (krome_opts_set_photo): edges [eV], the in-bin spectrum
  // exponent and the bin-averaged cross sections sigma[NPHOTO][nbins],
  // followed by their photoelectron-energy averages sigmaE[NPHOTO][nbins]
  int nbins = 0;
  double photo_alpha = 1.0;
  std::vector<double> Elo, Ehi, sigma;
  // per-species tolerances (krome_opts_set_atol_species / rtol_species):
  // NSPEC entries each, empty when unset; a negative entry takes the
  // scalar
  std::vector<double> atol_sp, rtol_sp;
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
double species_tol(const std::vector<double>& v, int i, double scalar);
// the device address of the current device's scratch (the solve that
// allocates nothing), looked up on the device's first use
unsigned char* scratch_of_current_device();

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
  Buf work, iwork, n, T, dt, par, flags, stats, queue, count, cons;
  Buf xn, xin, rho;  // the mass-fraction entry points' densities
  Buf prow, sig, phi;  // photo: kernel rows, cross sections, a host field
  Buf pwork, pord;     // the persistent kernel's workspace and queue order
  Buf awork, aiwork, apool, anext;  // the adjoint's workspace and pool
  Buf tol;  // the per-equation tolerances of a call (SolveOpts::tol)
  void release() {
    for (Buf* b : {&work, &iwork, &n, &T, &dt, &par, &flags, &stats, &queue,
                   &count, &cons, &xn, &xin, &rho, &prow, &sig, &phi, &pwork,
                   &pord, &awork, &aiwork, &apool, &anext, &tol})
      b->release();
  }
};
DevCache g_cache;
std::mutex g_cache_mutex;

// a call's private buffers (device_cache off), released on every exit
struct LocalCache {
  DevCache dc;
  bool on;
  ~LocalCache() {
    if (on) dc.release();
  }
};

// where a solve's kernels take their device memory: the library's buffers
// (dc, grown as needed), or fixed regions (dc null: the scratch of the
// entry point that allocates nothing)
struct SolveMem {
  DevCache* dc = nullptr;
  char* work = nullptr;
  size_t work_bytes = 0;
  char* ord = nullptr;
  size_t ord_bytes = 0;
};

template <bool ET>
constexpr int neq() { return kx::Problem<Net, ET>::N; }

// bytes of the plain kernel's workspace per cell
size_t plain_cell_bytes(bool et) {
  return et ? sizeof(double) * kx::cell_doubles<kx::PlainProblem<Net, true>>() +
                  sizeof(int) * neq<true>()
            : sizeof(double) * kx::cell_doubles<kx::PlainProblem<Net, false>>() +
                  sizeof(int) * neq<false>();
}

// the parameter rows the kernels read, Net::NROW wide: for a network with
// photo reactions the caller's rows P (device, stride pst, 0: one row)
// followed by each cell's photo bin sums over its field Phi (device
// [ncell][nbins]; null: no field, the sums 0), built on the device and
// returned in P, pst; for the others the caller's rows unchanged
int kernel_rows(int ncell, const double*& P, int& pst, const double* Phi,
                const Opts& o, DevCache& dc, hipStream_t s) {
  if constexpr (Net::NPHOTO == 0) {
    (void)ncell; (void)P; (void)pst; (void)Phi; (void)o; (void)dc; (void)s;
    return 0;
  } else {
    const int B = (Phi && o.nbins > 0) ? o.nbins : 0;
    const int rows = (B > 0 || pst != 0) ? ncell : 1;
    if (!dc.prow.ensure(sizeof(double) * (size_t)rows * Net::NROW))
      return kx::FLAG_DEVICE_MEMORY;
    if (B > 0) {
      // sigma, then sigmaE when the network has PHOTO heating
      const size_t sb = sizeof(double) * (size_t)Net::NPHOTO * B *
                        (Net::PHOTO_HEAT ? 2 : 1);
      if (!dc.sig.ensure(sb)) return kx::FLAG_DEVICE_MEMORY;
      if (!hip_ok(hipMemcpy(dc.sig.p, o.sigma.data(), sb,
                            hipMemcpyHostToDevice), "H2D"))
        return -1;
    }
    int block = 256;
    size_t lds = 0;
    if (B > 0) kx::photo_rows_shape(B, block, lds);
    hipLaunchKernelGGL(kx::photo_rows_kernel<Net>,
                       dim3((rows + block - 1) / block), dim3(block), lds, s,
                       rows, P, pst, B ? Phi : nullptr, B,
                       static_cast<const double*>(dc.sig.p),
                       static_cast<double*>(dc.prow.p), lds > 0 ? 1 : 0);
    if (!hip_ok(hipGetLastError(), "photo rows kernel")) return -1;
    P = static_cast<const double*>(dc.prow.p);
    pst = rows == 1 ? 0 : Net::NROW;
    return 0;
  }
}

// the plain kernel over a chunk of resident arrays: its workspace in the
// library's buffers, or (dc null) at work, the doubles of the chunk's
// cells followed by their pivot ints
template <bool ET>
int launch_plain(kx::CellsArgs a, DevCache* dc, char* work, hipStream_t s) {
  constexpr int N = neq<ET>();
  // the workspace in whole blocks of 64 slots (kx::blocked_slot)
  const long S = kx::blocked_slots(a.ncell);
  const size_t db =
      sizeof(double) * kx::cell_doubles<kx::PlainProblem<Net, ET>>() * S;
  double* wd = reinterpret_cast<double*>(work);
  int* wi = reinterpret_cast<int*>(work + db);
  if (dc) {
    if (!dc->work.ensure(db)) return -1;
    if (!dc->iwork.ensure(sizeof(int) * N * S)) return -1;
    wd = static_cast<double*>(dc->work.p);
    wi = static_cast<int*>(dc->iwork.p);
  }
  const int tpb = 128;
  const int blocks = (a.ncell + tpb - 1) / tpb;
  hipLaunchKernelGGL((kx::plain_cells_kernel<Net, ET>), dim3(blocks),
                     dim3(tpb), 0, s, a, wd, wi);
  return hip_ok(hipGetLastError(), "plain kernel launch") ? 0 : -1;
}

// HIP events around the last device call's kernel launches (a solve, or
// a conservation), on their stream
hipEvent_t g_ev[2] = {nullptr, nullptr};
bool g_ev_recorded = false;

void ev_begin(hipStream_t s) {
  if (!g_ev[0] && (hipEventCreate(&g_ev[0]) != hipSuccess ||
                   hipEventCreate(&g_ev[1]) != hipSuccess))
    g_ev[0] = g_ev[1] = nullptr;
  if (g_ev[0]) (void)hipEventRecord(g_ev[0], s);
}
void ev_end(hipStream_t s) {
  if (!g_ev[0]) return;
  (void)hipEventRecord(g_ev[1], s);
  g_ev_recorded = true;
}

int solve_kernels(kx::CellsArgs a, const Opts& o, const SolveMem& m,
                  hipStream_t s);
int solve_resident_(kx::CellsArgs a, const Opts& o, DevCache& dc,
                    hipStream_t s) {
  SolveMem m;
  m.dc = &dc;
  return solve_kernels(a, o, m, s);
}

// the solve with the post-solve conservation of mode 2 (solve_cell.hpp:
// 1484-1488): each cell's input totals taken before the kernels, the
// integrated cells projected onto them after (modes 1 and 3 run inside
// the solve kernels)
int solve_conserved(kx::CellsArgs a, const Opts& o, DevCache& dc,
                    hipStream_t s) {
  if (a.o.conserve != 2 || a.ncell <= 0) return solve_resident_(a, o, dc, s);
  if (!dc.cons.ensure(sizeof(double) * (size_t)a.ncell * (Net::NELEM + 1)))
    return kx::FLAG_DEVICE_MEMORY;
  double* tg = static_cast<double*>(dc.cons.p);
  const int blocks = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(kx::conserve_targets_kernel<Net>, dim3(blocks),
                     dim3(256), 0, s, a, tg);
  kx::CellsArgs b = a;
  b.o.conserve = 0;
  const int r = solve_resident_(b, o, dc, s);
  if (r != 0) return r;
  hipLaunchKernelGGL(kx::conserve_post_kernel<Net>, dim3(blocks), dim3(256),
                     0, s, a, (const double*)tg);
  return hip_ok(hipGetLastError(), "conservation kernels") ? 0 : -1;
}

// the cells of one device call, in chunks that fit the memory budget:
// arrays already resident; flags always written
int solve_resident(kx::CellsArgs a, const Opts& o, DevCache& dc,
                   hipStream_t s) {
  ev_begin(s);
  const int r = solve_conserved(a, o, dc, s);
  ev_end(s);
  return r;
}

// the solve kernels over resident arrays, their device memory from m
int solve_kernels(kx::CellsArgs a, const Opts& o, const SolveMem& m,
                  hipStream_t s) {
  if (a.ncell <= 0) return 0;
  if (a.o.evolve_T && !Net::THERMAL) {
    return kx::FLAG_UNSUPPORTED;
  }
  DevCache* dc = m.dc;
  if (o.kernel == 1 || o.kernel == 2) {
    // the workspace and the queue order: the library's buffers, or the
    // fixed regions when they are large enough
    auto mem = [&](size_t wb, size_t ob, kx::PersistMem& pm) -> int {
      if (dc) {
        if (!dc->pwork.ensure(wb) || !dc->pord.ensure(ob))
          return kx::FLAG_DEVICE_MEMORY;
        pm.work = dc->pwork.p;
        pm.ord = dc->pord.p;
        return 0;
      }
      if (wb > m.work_bytes || ob > m.ord_bytes) return kx::FLAG_UNSUPPORTED;
      pm.work = m.work;
      pm.ord = m.ord;
      return 0;
    };
    if (o.kernel == 2) {  // the lane queue, for any network
      if constexpr (Net::THERMAL)
        if (a.o.evolve_T)
          return kx::launch_lane_queue<Net, true>(a, o.waves_per_cu, s, mem,
                                                  o.lane_block);
      return kx::launch_lane_queue<Net, false>(a, o.waves_per_cu, s, mem,
                                               o.lane_block);
    }
    if constexpr (Net::NSPEC > KX_BIG_N)  // a wave per cell
      if (!a.o.evolve_T)
        return kx::launch_wave_queue<Net>(a, o.waves_per_cu,
                                          o.cells_per_workgroup, s, mem);
    if constexpr (Net::THERMAL)
      if (a.o.evolve_T)
        return kx::launch_persistent<Net, true>(a, o.mixed, o.queue_chunk,
                                                o.waves_per_cu, s, mem);
    return kx::launch_persistent<Net, false>(a, o.mixed, o.queue_chunk,
                                             o.waves_per_cu, s, mem);
  }
  // chunking of the plain kernel by the memory budget (or the fixed
  // region)
  size_t budget;
  if (!dc) {
    budget = m.work_bytes;
  } else if (o.device_memory < 0) {
    budget = (size_t)-1;
  } else if (o.device_memory > 0) {
    budget = (size_t)o.device_memory;
  } else {
    size_t fr = 0, tot = 0;
    if (!hip_ok(hipMemGetInfo(&fr, &tot), "hipMemGetInfo")) return -1;
    budget = (size_t)(0.8 * (double)(fr + dc->work.bytes + dc->iwork.bytes));
  }
  const size_t per = plain_cell_bytes(a.o.evolve_T != 0);
  // whole blocks of 64 slots: at least one
  long chunk = (long)(budget / per) / kx::SLOT_BLOCK * kx::SLOT_BLOCK;
  if (chunk < kx::SLOT_BLOCK) return kx::FLAG_DEVICE_MEMORY;
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
      r = a.o.evolve_T ? launch_plain<true>(b, dc, m.work, s)
                       : launch_plain<false>(b, dc, m.work, s);
    else
      r = launch_plain<false>(b, dc, m.work, s);
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

// photon bins (krome.h:203-228, krome.cpp:313-377): the bins' cross
// sections are averaged here, once, on the host (kx_photo.h)
int krome_opts_set_photo(krome_opts* h, int nbins, const double* Elo,
                         const double* Ehi) {
  if (!h || nbins <= 0 || !Elo || !Ehi) return -1;
  for (int b = 0; b < nbins; ++b)
    if (!(Ehi[b] > Elo[b]) || !(Elo[b] > 0.0)) return -1;
  h->Elo.assign(Elo, Elo + nbins);
  h->Ehi.assign(Ehi, Ehi + nbins);
  h->nbins = nbins;
  kx::photo_bin_xsec<Net>(nbins, Elo, Ehi, h->photo_alpha, h->sigma);
  return 0;
}
int krome_opts_nbins(const krome_opts* h) { return h ? h->nbins : 0; }
int krome_opts_set_photo_alpha(krome_opts* h, double alpha) {
  if (!h || !std::isfinite(alpha)) return -1;
  h->photo_alpha = alpha;
  if (h->nbins > 0)
    kx::photo_bin_xsec<Net>(h->nbins, h->Elo.data(), h->Ehi.data(), alpha,
                            h->sigma);
  return 0;
}
// Doppler broadening of tabulated cross sections and a dust opacity of the
// bins are not built here: only "none" is accepted
int krome_opts_set_photo_broadening(krome_opts* h, double Tgas,
                                    double vturb) {
  if (!h || Tgas < 0.0 || vturb < 0.0) return -1;
  return (Tgas == 0.0 && vturb == 0.0) ? 0 : -1;
}
int krome_opts_set_photo_dust_opacity(krome_opts* h, const double* kappa) {
  if (!h || h->nbins == 0) return -1;
  return kappa ? -1 : 0;
}
int krome_opts_photo_sigma(const krome_opts* h, double* sigma) {
  if (!h || h->nbins == 0 || !sigma) return -1;
  for (int i = 0; i < Net::NPHOTO * h->nbins; ++i) sigma[i] = h->sigma[i];
  return 0;
}
int krome_nphoto(void) { return Net::NPHOTO; }
int krome_photo_reaction(int r) {
  return (r >= 0 && r < Net::NPHOTO) ? Net::photo_reaction[r] : -1;
}
int krome_photo_reactant(int r) {
  return (r >= 0 && r < Net::NPHOTO) ? Net::photo_reactant[r] : -1;
}
double krome_photo_pdiss(int r) {
  return (r >= 0 && r < Net::NPHOTO) ? 1.0 : std::nan("");
}
double krome_photo_xsec(int r, double E) {
  if (r < 0 || r >= Net::NPHOTO) return std::nan("");
  return (E >= Net::photo_Emin[r] && E <= Net::photo_Emax[r])
             ? Net::photo_xsec(r, E)
             : 0.0;
}
int krome_bins_lin(int nbins, double lower, double upper, double* Elo,
                   double* Ehi) {
  if (nbins <= 0 || !(upper > lower) || !(lower > 0.0) || !Elo || !Ehi)
    return -1;
  const double dE = (upper - lower) / nbins;
  for (int b = 0; b < nbins; ++b) {
    Elo[b] = lower + b * dE;
    Ehi[b] = lower + (b + 1) * dE;
  }
  return 0;
}
int krome_bins_log(int nbins, double lower, double upper, double* Elo,
                   double* Ehi) {
  if (nbins <= 0 || !(upper > lower) || !(lower > 0.0) || !Elo || !Ehi)
    return -1;
  const double l0 = std::log10(lower);
  const double dl = (std::log10(upper) - l0) / nbins;
  for (int b = 0; b < nbins; ++b) {
    Elo[b] = std::pow(10.0, l0 + b * dl);
    Ehi[b] = std::pow(10.0, l0 + (b + 1) * dl);
  }
  return 0;
}

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
    if (v < 0 || v > 2) return -1;
    h->kernel = (int)v;
    return 0;
  }
  if (is("lane_block")) {
    if (v != 64 && v != 128) return -1;
    h->lane_block = (int)v;
    return 0;
  }
  if (is("cells_per_workgroup")) {
    if (v != 0 && v != 1 && v != 2 && v != 4 && v != 8 && v != 16)
      return -1;
    h->cells_per_workgroup = (int)v;
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
  if (is("checkpoint_steps")) {
    if (v < 1 || v > 1024) return -1;
    h->checkpoint_steps = (int)v;
    return 0;
  }
  if (is("checkpoint_memory")) {
    if (v < 0) return -1;
    h->checkpoint_memory = (long long)v;
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
  else if (is("lane_block")) *v = h->lane_block;
  else if (is("cells_per_workgroup")) *v = h->cells_per_workgroup;
  else if (is("checkpoint_steps")) *v = h->checkpoint_steps;
  else if (is("checkpoint_memory")) *v = (double)h->checkpoint_memory;
  else return -1;
  return 0;
}

// per-species tolerances (krome.h:175-185, krome.cpp:496-538): v[NSPEC]
// copied into the handle (NULL: the scalar for every species); a NaN, or a
// zero atol, is refused with the handle unchanged; get writes what each
// species gets (its entry, or the scalar where the entry is negative)
int krome_opts_set_atol_species(krome_opts* h, const double* atol) {
  if (!h) return -1;
  if (!atol) {
    h->atol_sp.clear();
    return 0;
  }
  for (int i = 0; i < NS; ++i)
    if (std::isnan(atol[i]) || atol[i] == 0.0) return -1;
  h->atol_sp.assign(atol, atol + NS);
  return 0;
}
int krome_opts_set_rtol_species(krome_opts* h, const double* rtol) {
  if (!h) return -1;
  if (!rtol) {
    h->rtol_sp.clear();
    return 0;
  }
  for (int i = 0; i < NS; ++i)
    if (std::isnan(rtol[i])) return -1;
  h->rtol_sp.assign(rtol, rtol + NS);
  return 0;
}
int krome_opts_get_atol_species(const krome_opts* h, double* atol) {
  if (!h || !atol) return -1;
  for (int i = 0; i < NS; ++i) atol[i] = species_tol(h->atol_sp, i, h->atol_n);
  return 0;
}
int krome_opts_get_rtol_species(const krome_opts* h, double* rtol) {
  if (!h || !rtol) return -1;
  for (int i = 0; i < NS; ++i) rtol[i] = species_tol(h->rtol_sp, i, h->rtol);
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
  // the selected device's scratch address, for the solve that allocates
  // nothing
  if (!scratch_of_current_device()) return kx::FLAG_UNSUPPORTED;
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

// the duration of the last device solve's or conservation's kernels (HIP
// events recorded around their launches on their stream), waiting for
// them to finish
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
  o.linsol = h->linsol;
  return o;
}

// a species' tolerance: its entry, or the scalar where the handle has no
// array or the entry is negative (solve_cell.hpp:866-869)
double species_tol(const std::vector<double>& v, int i, double scalar) {
  return !v.empty() && !(v[i] < 0.0) ? v[i] : scalar;
}

// the per-equation tolerances of a call, SolveOpts::tol's layout (N
// absolute values, then N relative ones, N = NSPEC + evolve_T; Tgas takes
// atol_T and the scalar rtol), into t; false when the handle has none.
// rtol_v: some species' rtol differs from the scalar, so the integrator
// uses per-slot weights (solve_cell.hpp:872-879, 2129-2139)
bool tol_rows(const krome_opts* h, std::vector<double>& t, int& rtol_v) {
  if (h->atol_sp.empty() && h->rtol_sp.empty()) return false;
  const int N = NS + (h->evolve_T ? 1 : 0);
  t.assign(2 * N, 0.0);
  rtol_v = 0;
  for (int i = 0; i < NS; ++i) {
    t[i] = species_tol(h->atol_sp, i, h->atol_n);
    t[N + i] = species_tol(h->rtol_sp, i, h->rtol);
    rtol_v |= t[N + i] != h->rtol;
  }
  if (h->evolve_T) {
    t[NS] = h->atol_T;
    t[N + NS] = h->rtol;
  }
  return true;
}

// the call's per-equation tolerances on the device, in buf (grown) or at
// fixed (room for 2 (NSPEC + 1) doubles), their address in o.tol
int attach_tol(kx::SolveOpts& o, const krome_opts* h, Buf* buf,
               double* fixed) {
  std::vector<double> t;
  int rv = 0;
  o.tol = nullptr;
  o.rtol_v = 0;
  if (!tol_rows(h, t, rv)) return 0;
  double* d = fixed;
  if (buf) {
    if (!buf->ensure(sizeof(double) * t.size())) return kx::FLAG_DEVICE_MEMORY;
    d = static_cast<double*>(buf->p);
  }
  if (!hip_ok(hipMemcpy(d, t.data(), sizeof(double) * t.size(),
                        hipMemcpyHostToDevice), "H2D"))
    return -1;
  o.tol = d;
  o.rtol_v = rv;
  return 0;
}

// what the whole call rejects before integrating (every cell flagged)
int reject(const krome_opts* h) {
  if (h->linsol == 2) return kx::FLAG_UNSUPPORTED;  // no LAPACK here
  // the lockstep persistent kernel (networks of up to KX_PERSIST_MAX_N
  // equations) factorizes the dense Newton matrix only; larger networks
  // take the lane queue, which runs the plain solve with either LU
  if (h->kernel == 1 && h->linsol == 3 &&
      kx::Problem<Net, false>::N + (h->evolve_T ? 1 : 0) <= KX_PERSIST_MAX_N)
    return kx::FLAG_UNSUPPORTED;
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

// rows v in the layout -> (cell stride, species stride)
void strides(int layout, long ld, int ncell, long& cs, long& ss) {
  if (layout == KROME_LAYOUT_SPECIES_MAJOR) {
    cs = 1;
    ss = ld ? ld : ncell;
  } else {
    cs = NS;
    ss = 1;
  }
}

// the mass-fraction form around a resident solve (solve_cells.hpp:
// 176-200, solve_kokkos.hpp:1450-1458, 1519-1528): a.n holds the caller's
// x (in its layout); the densities live in dc.xn (species-major) during
// the solve; the cells with flag 0 are converted back (with skip_dt, only
// those with dt > 0, as the device path does); failed cells keep x
int solve_x_resident(kx::CellsArgs a, const double* rho, bool skip_dt,
                     int renorm, const Opts& o, DevCache& dc, hipStream_t s) {
  const int nc = a.ncell;
  if (!dc.xn.ensure(sizeof(double) * (size_t)nc * NS) ||
      !dc.xin.ensure(sizeof(double) * (size_t)nc))
    return kx::FLAG_DEVICE_MEMORY;
  double* xn = static_cast<double*>(dc.xn.p);
  double* xin = static_cast<double*>(dc.xin.p);
  const dim3 g((nc + 255) / 256), b(256);
  hipLaunchKernelGGL(kx::x_to_n_kernel<Net>, g, b, 0, s, nc,
                     (const double*)a.n, a.n_cs, a.n_ss, rho, xn, xin);
  kx::CellsArgs d = a;
  d.n = xn;
  d.n_cs = 1;
  d.n_ss = nc;
  const int r = solve_resident(d, o, dc, s);
  if (r != 0) return r;
  hipLaunchKernelGGL(kx::n_to_x_kernel<Net>, g, b, 0, s, nc,
                     (const double*)xn, a.n, a.n_cs, a.n_ss, rho,
                     (const double*)xin, (const int*)a.flags,
                     skip_dt ? a.dt : nullptr, renorm);
  return hip_ok(hipGetLastError(), "mass-fraction kernels") ? 0 : -1;
}

// krome_solve_cells / _x on host arrays (krome.cpp:396-470): rows copied
// to the device, solved there, copied back; rho (may be NULL) makes v the
// mass fractions
int solve_host(int ncell, double* v, const double* rho, double* Tgas,
               const double* dt, const double* par, int par_stride,
               const double* Phi, int* flags, const krome_opts* h) {
  if (!h || ncell < 0 || !v || !Tgas || !dt) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (ncell == 0) return 0;
  // photon bins configured and no field (krome.cpp:415-417)
  if (h->nbins > 0 && !Phi) return kx::FLAG_NO_FIELD;
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
  LocalCache local{DevCache(), !h->device_cache};
  DevCache& dc = h->device_cache ? g_cache : local.dc;
  const size_t nb = sizeof(double) * (size_t)ncell * NS;
  const size_t pb = sizeof(double) * (size_t)(pst ? ncell : 1) * Net::NPAR;
  const bool field = Net::NPHOTO > 0 && h->nbins > 0;
  const size_t fb = sizeof(double) * (size_t)ncell * h->nbins;
  if (!dc.n.ensure(nb) || !dc.T.ensure(sizeof(double) * ncell) ||
      !dc.dt.ensure(sizeof(double) * ncell) || !dc.par.ensure(pb) ||
      !dc.flags.ensure(sizeof(int) * ncell) ||
      (rho && !dc.rho.ensure(sizeof(double) * ncell)) ||
      (field && !dc.phi.ensure(fb)))
    return kx::FLAG_DEVICE_MEMORY;
  hipStream_t s = nullptr;
  bool ok = hip_ok(hipMemcpy(dc.n.p, v, nb, hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.T.p, Tgas, sizeof(double) * ncell,
                             hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.dt.p, dt, sizeof(double) * ncell,
                             hipMemcpyHostToDevice), "H2D") &&
            hip_ok(hipMemcpy(dc.par.p, P, pb, hipMemcpyHostToDevice), "H2D") &&
            (!rho || hip_ok(hipMemcpy(dc.rho.p, rho, sizeof(double) * ncell,
                                      hipMemcpyHostToDevice), "H2D")) &&
            (!field || hip_ok(hipMemcpy(dc.phi.p, Phi, fb,
                                        hipMemcpyHostToDevice), "H2D"));
  if (!ok) return -1;
  const double* Pk = static_cast<const double*>(dc.par.p);
  int pk = pst;
  const int rr = kernel_rows(ncell, Pk, pk,
                             field ? static_cast<const double*>(dc.phi.p)
                                   : nullptr, *h, dc, s);
  if (rr != 0) return rr;
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = static_cast<double*>(dc.n.p);
  a.n_cs = NS;
  a.n_ss = 1;
  a.T = static_cast<double*>(dc.T.p);
  a.dt = static_cast<const double*>(dc.dt.p);
  a.par = Pk;
  a.par_stride = pk;
  a.flags = static_cast<int*>(dc.flags.p);
  a.stats = nullptr;
  a.o = solve_opts(h);
  if (const int rt = attach_tol(a.o, h, &dc.tol, nullptr)) return rt;
  int r = rho ? solve_x_resident(a, static_cast<const double*>(dc.rho.p),
                                 false, h->renormalize_x ? 1 : 0, *h, dc, s)
              : solve_resident(a, *h, dc, s);
  if (r == 0) {
    ok = hip_ok(hipDeviceSynchronize(), "kernel") &&
         hip_ok(hipMemcpy(v, dc.n.p, nb, hipMemcpyDeviceToHost), "D2H") &&
         hip_ok(hipMemcpy(Tgas, dc.T.p, sizeof(double) * ncell,
                          hipMemcpyDeviceToHost), "D2H") &&
         hip_ok(hipMemcpy(fl.data(), dc.flags.p, sizeof(int) * ncell,
                          hipMemcpyDeviceToHost), "D2H");
    if (!ok) r = -1;
  }
  if (r != 0) {
    if (r == kx::FLAG_DEVICE_MEMORY && flags)
      for (int c = 0; c < ncell; ++c) flags[c] = r;
    return r;
  }
  if (flags) std::memcpy(flags, fl.data(), sizeof(int) * ncell);
  return first_flag(fl.data(), ncell);
}

// krome_solve_cells(_x)_device (krome.cpp:1237-1304): arrays on the
// device, v in either layout (rho: mass fractions)
int solve_device(int ncell, double* v, const double* rho, int layout,
                 long ld, double* Tgas, const double* dt, const double* par,
                 int par_stride, const double* Phi, int* flags,
                 krome_device_summary* summary, const krome_opts* h) {
  if (summary) *summary = krome_device_summary{0, 0, 0};
  if (!h || ncell < 0 || !v || !Tgas || !dt) return -1;
  if (layout != KROME_LAYOUT_SPECIES_MAJOR &&
      layout != KROME_LAYOUT_CELL_MAJOR)
    return -1;
  if (ld != 0 && ld < ncell) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (ncell == 0) return 0;
  if (h->nbins > 0 && !Phi) return kx::FLAG_NO_FIELD;
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  // parameters: the defaults on the device when none given
  const double* P = par;
  int pst = par ? par_stride : 0;
  if (!P) {
    std::vector<double> d(Net::NPAR);
    krome_params_default(d.data());
    if (!dc.par.ensure(sizeof(double) * Net::NPAR))
      return kx::FLAG_DEVICE_MEMORY;
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
  const int rr = kernel_rows(ncell, P, pst, h->nbins > 0 ? Phi : nullptr,
                             *h, dc, nullptr);
  if (rr != 0) return rr;
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = v;
  strides(layout, ld, ncell, a.n_cs, a.n_ss);
  a.T = Tgas;
  a.dt = dt;
  a.par = P;
  a.par_stride = pst;
  a.flags = fl;
  a.stats = nullptr;
  a.o = solve_opts(h);
  if (const int rt = attach_tol(a.o, h, &dc.tol, nullptr)) return rt;
  const int r = rho ? solve_x_resident(a, rho, true,
                                       h->renormalize_x ? 1 : 0, *h, dc,
                                       nullptr)
                    : solve_resident(a, *h, dc, nullptr);
  if (r != 0) return r;
  if (!hip_ok(hipDeviceSynchronize(), "kernel")) return -1;
  std::vector<int> hf(ncell);
  if (!hip_ok(hipMemcpy(hf.data(), fl, sizeof(int) * ncell,
                        hipMemcpyDeviceToHost), "D2H"))
    return -1;
  if (summary)
    for (int c = 0; c < ncell; ++c) {
      if (hf[c] == 0) summary->device++;
      else summary->failed++;
    }
  return first_flag(hf.data(), ncell);
}

}  // namespace

extern "C" {

int krome_solve_cells(int ncell, double* n, double* Tgas, const double* dt,
                      const double* par, int par_stride,
                      const krome_cells* extra, int* flags,
                      const krome_opts* h) {
  // the photon field of the cells (extra->Phi); no dust bins here
  return solve_host(ncell, n, nullptr, Tgas, dt, par, par_stride,
                    extra ? extra->Phi : nullptr, flags, h);
}

int krome_solve_cells_x(int ncell, double* x, const double* rho,
                        double* Tgas, const double* dt, const double* par,
                        int par_stride, const krome_cells* extra, int* flags,
                        const krome_opts* h) {
  if (!rho) return -1;
  return solve_host(ncell, x, rho, Tgas, dt, par, par_stride,
                    extra ? extra->Phi : nullptr, flags, h);
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
  return solve_device(ncell, n, nullptr, layout, ld, Tgas, dt, par,
                      par_stride, extra ? extra->Phi : nullptr, flags,
                      summary, h);
}

// the same with mass fractions x and the gas density rho[ncell] on the
// device (krome.cpp:1295-1304)
int krome_solve_cells_x_device(int ncell, double* x, const double* rho,
                               int layout, long ld, double* Tgas,
                               const double* dt, const double* par,
                               int par_stride, const krome_cells* extra,
                               int* flags, krome_device_summary* summary,
                               const krome_opts* h) {
  if (!rho) return -1;
  return solve_device(ncell, x, rho, layout, ld, Tgas, dt, par, par_stride,
                      extra ? extra->Phi : nullptr, flags, summary, h);
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
  if (h->nbins > 0) return kx::FLAG_NO_FIELD;  // no field argument here
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
  int pst = par_stride;
  const int rr = kernel_rows(ncell, par, pst, nullptr, *h, g_cache, nullptr);
  if (rr != 0) return rr;
  par_stride = pst;
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = n;
  strides(layout, ld, ncell, a.n_cs, a.n_ss);
  a.T = Tgas;
  a.dt = dt;
  a.par = par;
  a.par_stride = par_stride;
  a.flags = flags;
  a.stats = stats;
  a.o = solve_opts(h);
  if (const int rt = attach_tol(a.o, h, &g_cache.tol, nullptr)) return rt;
  return solve_resident(a, *h, g_cache, nullptr);
}

}  // extern "C"

// ---- the solve that allocates nothing ------------------------------------
// krome_solve_cells_device_noalloc keeps every buffer its kernels need in
// one file-scope device array, scratch, which each GPU holds from the
// library's load until the process ends.  Its regions (bytes):
//   head   256: the parameter-check count, the lowest failing cell, the
//          counts of integrated and failed cells (ints)
//   par    the default parameter row (when the caller gives none), then
//          the per-equation tolerances (2 (NSPEC + 1) doubles)
//   flags  4 x KX_NOALLOC_MAX_CELLS (when the caller gives none)
//   order  the persistent queue's order for KX_NOALLOC_MAX_CELLS cells
//   work   the persistent workspace of this network (FP64, its largest
//          form) when that is at most KX_NOALLOC_WORK_CAP, else 64 MiB; the
//          plain kernel runs in chunks of as many cells as fit there
namespace {

#ifndef KX_NOALLOC_MAX_CELLS
#define KX_NOALLOC_MAX_CELLS (1L << 20)
#endif
#ifndef KX_NOALLOC_WORK_CAP
#define KX_NOALLOC_WORK_CAP (160UL << 20)
#endif

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
template <bool ET>
constexpr size_t na_persist() {
  if constexpr (kx::Problem<Net, ET>::N > KX_PERSIST_MAX_N) return 0;
  else return kx::persist_work_bytes<Net, ET, false>();
}
constexpr size_t na_persist_max() {
  const size_t a = na_persist<false>();
  const size_t b = Net::THERMAL ? na_persist<true>() : 0;
  return a > b ? a : b;
}
constexpr size_t NA_HEAD = 256;
constexpr size_t NA_PAR =
    na_align(sizeof(double) * (Net::NPAR + 1 + 2 * (NS + 1)));
constexpr size_t NA_FLAGS = na_align(sizeof(int) * KX_NOALLOC_MAX_CELLS);
constexpr size_t NA_ORDER =
    na_align(sizeof(int) * (2 * KX_NOALLOC_MAX_CELLS + kx::ORDER_BUCKETS));
constexpr size_t NA_WORK =
    (na_persist_max() > 0 && na_persist_max() <= KX_NOALLOC_WORK_CAP)
        ? na_align(na_persist_max())
        : (64UL << 20);
constexpr size_t NA_BYTES = NA_HEAD + NA_PAR + NA_FLAGS + NA_ORDER + NA_WORK;

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// the device address of each GPU's scratch, looked up once per device with
// that device current (the array's name is only a host-side stand-in)
constexpr int NA_MAX_DEVICES = 64;
unsigned char* g_scratch[NA_MAX_DEVICES] = {};
std::mutex g_scratch_mutex;

unsigned char* scratch_of_current_device() {
  int dev = 0;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice") || dev < 0 ||
      dev >= NA_MAX_DEVICES)
    return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                "hipGetSymbolAddress(scratch)"))
      return nullptr;
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch[dev];
}

// per cell: the lowest failing cell (atomicMin) and the counts of
// integrated and failed cells, into head[1..3]
__global__ void na_summary_kernel(int ncell, const int* flags, int* head) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  if (flags[c] != 0) {
    atomicMin(&head[1], c);
    atomicAdd(&head[3], 1);
  } else {
    atomicAdd(&head[2], 1);
  }
}
// head[1..3] = (ncell, 0, 0) before the summary
__global__ void na_head_kernel(int ncell, int* head) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    head[1] = ncell;
    head[2] = 0;
    head[3] = 0;
  }
}

std::mutex g_noalloc_mutex;

}  // namespace

extern "C" {

int krome_solve_cells_device_noalloc(int ncell, double* n, int layout,
                                     long ld, double* Tgas, const double* dt,
                                     const double* par, int par_stride,
                                     int* flags,
                                     krome_device_summary* summary,
                                     const krome_opts* h) {
  if (summary) *summary = krome_device_summary{0, 0, 0};
  if (!h || ncell < 0 || !n || !Tgas || !dt) return -1;
  if (layout != KROME_LAYOUT_SPECIES_MAJOR &&
      layout != KROME_LAYOUT_CELL_MAJOR)
    return -1;
  if (ld != 0 && ld < ncell) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  if (ncell == 0) return 0;
  if (ncell > KX_NOALLOC_MAX_CELLS) {
    std::fprintf(stderr,
                 "krome: krome_solve_cells_device_noalloc: %d cells exceed "
                 "its maximum of %ld\n",
                 ncell, (long)KX_NOALLOC_MAX_CELLS);
    return -1;
  }
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  unsigned char* base = scratch_of_current_device();
  if (!base) return -1;
  std::lock_guard<std::mutex> lk(g_noalloc_mutex);
  int* head = reinterpret_cast<int*>(base);
  double* dpar = reinterpret_cast<double*>(base + NA_HEAD);
  int* dflags = reinterpret_cast<int*>(base + NA_HEAD + NA_PAR);
  char* ord = reinterpret_cast<char*>(base + NA_HEAD + NA_PAR + NA_FLAGS);
  char* work = ord + NA_ORDER;
  int* fl = flags ? flags : dflags;
  // no photon field and no post-solve projection here (their rows and
  // totals would need buffers of their own)
  const bool unsupported =
      Net::NPHOTO > 0 || h->nbins > 0 || h->conserve == 2;
  const double* P = par;
  int pst = par ? par_stride : 0;
  if (!P) {
    double d[Net::NPAR > 0 ? Net::NPAR : 1];
    krome_params_default(d);
    if (!hip_ok(hipMemcpy(dpar, d, sizeof(double) * Net::NPAR,
                          hipMemcpyHostToDevice), "H2D"))
      return -1;
    P = dpar;
  }
  const int np = kx::params_check_device<Net>(ncell, P, pst, fl, head);
  if (np < 0) return -1;
  if (np > 0) return kx::FLAG_PARAM_UNSET;
  int rj = unsupported ? kx::FLAG_UNSUPPORTED : reject(h);
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = n;
  strides(layout, ld, ncell, a.n_cs, a.n_ss);
  a.T = Tgas;
  a.dt = dt;
  a.par = P;
  a.par_stride = pst;
  a.flags = fl;
  a.stats = nullptr;
  a.o = solve_opts(h);
  if (rj == 0)
    if (const int rt = attach_tol(a.o, h, nullptr, dpar + Net::NPAR + 1))
      return rt;
  SolveMem m;
  m.work = work;
  m.work_bytes = NA_WORK;
  m.ord = ord;
  m.ord_bytes = NA_ORDER;
  int r = rj;
  if (rj == 0) r = solve_kernels(a, *h, m, nullptr);
  if (r == kx::FLAG_UNSUPPORTED || (rj != 0 && r != 0)) {
    kx::fill_flags(ncell, fl, r);
    return r;
  }
  if (r != 0) return r;
  const dim3 g((ncell + 255) / 256), b(256);
  hipLaunchKernelGGL(na_head_kernel, dim3(1), dim3(64), 0, nullptr, ncell,
                     head);
  hipLaunchKernelGGL(na_summary_kernel, g, b, 0, nullptr, ncell,
                     (const int*)fl, head);
  int hh[4] = {0, 0, 0, 0};
  if (!hip_ok(hipGetLastError(), "summary kernel") ||
      !hip_ok(hipMemcpy(hh, head, sizeof(hh), hipMemcpyDeviceToHost), "D2H"))
    return -1;
  if (summary) {
    summary->device = hh[2];
    summary->failed = hh[3];
  }
  if (hh[1] >= ncell) return 0;
  int f = 0;
  if (!hip_ok(hipMemcpy(&f, fl + hh[1], sizeof(int), hipMemcpyDeviceToHost),
              "D2H"))
    return -1;
  return f;
}

}  // extern "C"

// ---- the discrete adjoint --------------------------------------------------
namespace {

// cells of one adjoint launch (its workspace is sized for them)
constexpr long ADJ_CHUNK = 65536;

template <bool ET>
int adjoint_launch(kx::AdjArgs g, const Opts& o, long& peak) {
  using Prob = kx::AdjProblem<Net, ET>;
  constexpr long CKD = kx::adj_ck_doubles<Net, ET>();
  const bool bwd = g.gn || g.gT;
  const int seg = o.checkpoint_steps;
  const int fails = o.max_err_fails + o.max_conv_fails;
  g.seg_steps = seg;
  g.margin = bwd ? 4 * fails + 8 : 0;
  g.cap = bwd ? 4 * seg + g.margin : 0;
  g.ycap = bwd ? seg + 2 * (o.max_err_fails + 2) + 16 : 0;
  const long nc = g.a.ncell;
  const long S = nc < ADJ_CHUNK ? nc : ADJ_CHUNK;
  const size_t wd = sizeof(double) *
                    (size_t)kx::adj_cell_doubles<Net, ET>(g.cap, g.ycap) * S;
  const size_t wi = sizeof(int) * (size_t)Prob::N * S;
  long slots = 0;
  if (bwd) {
    // automatic: room for about 1536 steps per cell on average (the
    // cells share the pool)
    const long per = seg < 32 ? 1536 / seg : 48;
    slots = o.checkpoint_memory > 0
                ? (long)(o.checkpoint_memory / (long long)(CKD * 8))
                : per * S;
    if (slots < 1) return kx::FLAG_DEVICE_MEMORY;
  }
  DevCache& dc = g_cache;
  if (!dc.awork.ensure(wd) || !dc.aiwork.ensure(wi) ||
      (bwd && (!dc.apool.ensure(sizeof(double) * (size_t)slots * CKD) ||
               !dc.anext.ensure(sizeof(unsigned long long)))))
    return kx::FLAG_DEVICE_MEMORY;
  g.work = static_cast<double*>(dc.awork.p);
  g.iwork = static_cast<int*>(dc.aiwork.p);
  g.S = S;
  g.pool = static_cast<double*>(dc.apool.p);
  g.pool_slots = slots;
  g.pool_next = static_cast<unsigned long long*>(dc.anext.p);
  peak = 0;
  for (long c0 = 0; c0 < nc; c0 += S) {
    kx::AdjArgs b = g;
    const long m = (nc - c0) < S ? (nc - c0) : S;
    b.a.ncell = (int)m;
    b.a.n = g.a.n + c0;  // species-major: element i of cell c at i ld + c
    b.a.T = g.a.T + c0;
    b.a.dt = g.a.dt + c0;
    b.a.par = g.a.par + c0 * g.a.par_stride;
    b.a.flags = g.a.flags + c0;
    b.lnk = g.lnk ? g.lnk + c0 * Net::NREA : nullptr;
    b.gn = g.gn ? g.gn + c0 : nullptr;
    b.gT = g.gT ? g.gT + c0 : nullptr;
    b.g_lnk = g.g_lnk ? g.g_lnk + c0 * Net::NREA : nullptr;
    b.g_n0 = g.g_n0 ? g.g_n0 + c0 : nullptr;
    b.g_T0 = g.g_T0 ? g.g_T0 + c0 : nullptr;
    unsigned long long zero = 0;
    if (bwd && !hip_ok(hipMemcpy(b.pool_next, &zero, sizeof(zero),
                                 hipMemcpyHostToDevice), "H2D"))
      return -1;
    hipLaunchKernelGGL((kx::adjoint_cells_kernel<Net, ET>),
                       dim3((m + 63) / 64), dim3(64), 0, nullptr, b);
    if (!hip_ok(hipGetLastError(), "adjoint kernel")) return -1;
    unsigned long long used = 0;
    if (bwd) {
      if (!hip_ok(hipMemcpy(&used, b.pool_next, sizeof(used),
                            hipMemcpyDeviceToHost), "D2H"))
        return -1;
      if ((long)used > slots) used = slots;
    }
    const long p = (long)(wd + wi + used * CKD * sizeof(double));
    if (p > peak) peak = p;
  }
  return hip_ok(hipDeviceSynchronize(), "adjoint kernel") ? 0 : -1;
}

}  // namespace

extern "C" {

int krome_adjoint_cells_device(int ncell, double* n, double* Tgas,
                               const double* dt, const double* par,
                               int par_stride, const double* lnk, int* flags,
                               const double* gn, const double* gT,
                               double* g_lnk, double* g_n0, double* g_T0,
                               double* mem_bytes, const krome_opts* h) {
  if (!h || ncell < 0 || !n || !Tgas || !dt || !par || !flags) return -1;
  if (par_stride != 0 && par_stride != Net::NPAR) return -1;
  const bool bwd = gn || gT;
  if (bwd && (!g_lnk || !g_n0 || (h->evolve_T && !g_T0))) return -1;
  if (ncell == 0) return 0;
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  if (!g_cache.count.ensure(sizeof(int))) return kx::FLAG_DEVICE_MEMORY;
  const int np = kx::params_check_device<Net>(
      ncell, par, par_stride, flags, static_cast<int*>(g_cache.count.p));
  if (np < 0) return -1;
  if (np > 0) return kx::FLAG_PARAM_UNSET;
  // the adjoint differentiates the plain solve: no post-solve conservation
  // (its projection is another map), no photon fields, no host path
  // (device_kernel and mixed_precision do not apply: its own FP64 kernel)
  const int rj = (h->linsol == 2 || h->path == 1 || h->conserve != 0 ||
                  Net::NPHOTO > 0 || h->nbins > 0 ||
                  (h->evolve_T && !Net::THERMAL))
                     ? kx::FLAG_UNSUPPORTED
                     : 0;
  if (rj != 0) {
    kx::fill_flags(ncell, flags, rj);
    return rj;
  }
  kx::AdjArgs g;
  g.a.ncell = ncell;
  g.a.n = n;
  g.a.n_cs = 1;
  g.a.n_ss = ncell;
  g.a.T = Tgas;
  g.a.dt = dt;
  g.a.par = par;
  g.a.par_stride = par_stride;
  g.a.flags = flags;
  g.a.stats = nullptr;
  g.a.o = solve_opts(h);
  if (const int rt = attach_tol(g.a.o, h, &g_cache.tol, nullptr)) return rt;
  g.ld = ncell;
  g.lnk = lnk;
  g.gn = gn;
  g.gT = gT;
  g.g_lnk = g_lnk;
  g.g_n0 = g_n0;
  g.g_T0 = g_T0;
  long peak = 0;
  int r;
  if constexpr (Net::NSPEC > KX_BIG_N)  // its dense per-lane Jacobians
    r = kx::FLAG_UNSUPPORTED;
  else if constexpr (Net::THERMAL)
    r = h->evolve_T ? adjoint_launch<true>(g, *h, peak)
                    : adjoint_launch<false>(g, *h, peak);
  else
    r = adjoint_launch<false>(g, *h, peak);
  if (r == kx::FLAG_UNSUPPORTED) kx::fill_flags(ncell, flags, r);
  if (mem_bytes) *mem_bytes = (double)peak;
  if (r != 0) return r;
  std::vector<int> hf(ncell);
  if (!hip_ok(hipMemcpy(hf.data(), flags, sizeof(int) * ncell,
                        hipMemcpyDeviceToHost), "D2H"))
    return -1;
  return first_flag(hf.data(), ncell);
}

// ---- element totals and consistent advection ----------------------------

const char* krome_element_name(int e) {
  return (e >= 0 && e < Net::NELEM) ? kElements[e] : nullptr;
}
int krome_nelem(void) { return Net::NELEM; }
int krome_species_natoms(int i, int e) {
  if (i < 0 || i >= NS || e < 0 || e >= Net::NELEM) return -1;
  return Net::natoms(i, e);
}
double krome_species_mass(int i) {
  return (i >= 0 && i < NS) ? Net::mass_g(i) : std::nan("");
}
int krome_species_charge(int i) {
  return (i >= 0 && i < NS) ? Net::charge(i) : 0;
}

}  // extern "C"

namespace {

// the element totals of host rows, on the device
int totals_host(int ncell, const double* v, double* tot, bool x) {
  if (ncell < 0 || !v || !tot) return -1;
  if (ncell == 0) return 0;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  const size_t vb = sizeof(double) * (size_t)ncell * NS;
  const size_t tb = sizeof(double) * (size_t)ncell * Net::NELEM;
  if (!dc.n.ensure(vb) || !dc.cons.ensure(tb)) return kx::FLAG_DEVICE_MEMORY;
  if (!hip_ok(hipMemcpy(dc.n.p, v, vb, hipMemcpyHostToDevice), "H2D"))
    return -1;
  const dim3 g((ncell + 255) / 256), b(256);
  const double* dv = static_cast<const double*>(dc.n.p);
  double* dt = static_cast<double*>(dc.cons.p);
  if (x)
    hipLaunchKernelGGL((kx::totals_kernel<Net, true>), g, b, 0, nullptr,
                       ncell, dv, (long)NS, 1L, dt);
  else
    hipLaunchKernelGGL((kx::totals_kernel<Net, false>), g, b, 0, nullptr,
                       ncell, dv, (long)NS, 1L, dt);
  if (!hip_ok(hipGetLastError(), "totals kernel") ||
      !hip_ok(hipMemcpy(tot, dt, tb, hipMemcpyDeviceToHost), "D2H"))
    return -1;
  return 0;
}

// consistent advection of resident rows; the lowest failing cell's flag
int conserve_resident(int ncell, double* v, long cs, long ss,
                      const double* ref, long rs, int* flags, bool x,
                      DevCache& dc) {
  if (!dc.count.ensure(sizeof(int))) return kx::FLAG_DEVICE_MEMORY;
  int* first = static_cast<int*>(dc.count.p);
  if (!hip_ok(hipMemcpy(first, &ncell, sizeof(int), hipMemcpyHostToDevice),
              "H2D"))
    return -1;
  const dim3 g((ncell + 255) / 256), b(256);
  ev_begin(nullptr);
  if (x)
    hipLaunchKernelGGL((kx::conserve_kernel<Net, true>), g, b, 0, nullptr,
                       ncell, v, cs, ss, ref, rs, flags, first);
  else
    hipLaunchKernelGGL((kx::conserve_kernel<Net, false>), g, b, 0, nullptr,
                       ncell, v, cs, ss, ref, rs, flags, first);
  ev_end(nullptr);
  if (!hip_ok(hipGetLastError(), "conserve kernel")) return -1;
  int lo = ncell;
  if (!hip_ok(hipMemcpy(&lo, first, sizeof(int), hipMemcpyDeviceToHost),
              "D2H"))
    return -1;
  if (lo >= ncell) return 0;
  int f = 0;
  if (!hip_ok(hipMemcpy(&f, flags + lo, sizeof(int), hipMemcpyDeviceToHost),
              "D2H"))
    return -1;
  return f;
}

// the query functions on host rows (krome.cpp:106-133, 621-644, 753-770):
// rates, fluxes or the thermal balance of the states n[ncell][nspec] at
// Tgas, evaluated on the device; a cell with an unset REQUIRED parameter
// is skipped (its outputs untouched) and its flag returned
int query_host(int ncell, const double* n, const double* Tgas,
               const double* par, int par_stride, const double* Phi,
               const krome_opts* h, int what, double* out, double* heat,
               double* cool, double* heats, double* cools, double* dTdt) {
  if (ncell < 0 || !n || !Tgas) return -1;
  if (par && par_stride != 0 && par_stride != Net::NPAR) return -1;
  const int B = h ? h->nbins : 0;
  if (B > 0 && !Phi) return kx::FLAG_NO_FIELD;  // krome.cpp:113-116
  if (ncell == 0) return 0;
  std::vector<double> dflt(Net::NPAR);
  krome_params_default(dflt.data());
  const double* P = par ? par : dflt.data();
  const int pst = par ? par_stride : 0;
  const int per = what == kx::Q_THERMAL ? 3 : Net::NREA;
  const int nh = Net::NHEAT > 0 ? Net::NHEAT : 1;
  const int nc = Net::NCOOL > 0 ? Net::NCOOL : 1;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  const size_t nb = sizeof(double) * (size_t)ncell * NS;
  const size_t pb = sizeof(double) * (size_t)(pst ? ncell : 1) * Net::NPAR;
  const size_t ob = sizeof(double) * (size_t)ncell * (per + nh + nc);
  const bool field = Net::NPHOTO > 0 && B > 0;
  const size_t fb = sizeof(double) * (size_t)ncell * B;
  if (!dc.n.ensure(nb) || !dc.T.ensure(sizeof(double) * ncell) ||
      !dc.par.ensure(pb) || !dc.work.ensure(ob) ||
      !dc.flags.ensure(sizeof(int) * ncell) || (field && !dc.phi.ensure(fb)))
    return kx::FLAG_DEVICE_MEMORY;
  if (!hip_ok(hipMemcpy(dc.n.p, n, nb, hipMemcpyHostToDevice), "H2D") ||
      !hip_ok(hipMemcpy(dc.T.p, Tgas, sizeof(double) * ncell,
                        hipMemcpyHostToDevice), "H2D") ||
      !hip_ok(hipMemcpy(dc.par.p, P, pb, hipMemcpyHostToDevice), "H2D") ||
      (field && !hip_ok(hipMemcpy(dc.phi.p, Phi, fb, hipMemcpyHostToDevice),
                        "H2D")))
    return -1;
  const double* Pk = static_cast<const double*>(dc.par.p);
  int pk = pst;
  const Opts none;
  const int rr = kernel_rows(ncell, Pk, pk,
                             field ? static_cast<const double*>(dc.phi.p)
                                   : nullptr,
                             h ? static_cast<const Opts&>(*h) : none, dc,
                             nullptr);
  if (rr != 0) return rr;
  double* o = static_cast<double*>(dc.work.p);
  double* hs = o + (size_t)ncell * per;
  double* cs = hs + (size_t)ncell * nh;
  hipLaunchKernelGGL(kx::query_kernel<Net>, dim3((ncell + 127) / 128),
                     dim3(128), 0, nullptr, ncell,
                     static_cast<const double*>(dc.n.p),
                     static_cast<const double*>(dc.T.p),
                     Pk, pk, what, o, hs,
                     cs, static_cast<int*>(dc.flags.p));
  if (!hip_ok(hipGetLastError(), "query kernel")) return -1;
  std::vector<double> ho((size_t)ncell * (per + nh + nc));
  std::vector<int> fl(ncell);
  if (!hip_ok(hipMemcpy(ho.data(), o, ob, hipMemcpyDeviceToHost), "D2H") ||
      !hip_ok(hipMemcpy(fl.data(), dc.flags.p, sizeof(int) * ncell,
                        hipMemcpyDeviceToHost), "D2H"))
    return -1;
  const double* hh = ho.data() + (size_t)ncell * per;
  const double* hc = hh + (size_t)ncell * nh;
  int first = 0;
  for (int c = 0; c < ncell; ++c) {
    if (fl[c] != 0) {
      if (first == 0) first = fl[c];
      continue;
    }
    const double* oc = ho.data() + (size_t)c * per;
    if (what != kx::Q_THERMAL) {
      for (int r = 0; r < Net::NREA; ++r) out[(size_t)c * Net::NREA + r] = oc[r];
      continue;
    }
    if (heat) heat[c] = oc[0];
    if (cool) cool[c] = oc[1];
    if (dTdt) dTdt[c] = oc[2];
    if (heats)
      for (int i = 0; i < Net::NHEAT; ++i)
        heats[(size_t)c * Net::NHEAT + i] = hh[(size_t)c * Net::NHEAT + i];
    if (cools)
      for (int i = 0; i < Net::NCOOL; ++i)
        cools[(size_t)c * Net::NCOOL + i] = hc[(size_t)c * Net::NCOOL + i];
  }
  return first;
}

// krome_conserve_cells(_x) on host rows (krome.cpp:943-977): the same
// arguments and errors; the rows corrected on the device
int conserve_host(int ncell, double* v, const double* ref, int ref_stride,
                  const double* par, int par_stride, int* flags, bool x) {
  if (ncell < 0 || !v || !ref || ref_stride < 0 || par_stride < 0) return -1;
  if (ref_stride != 0 && ref_stride != Net::NELEM) return -1;
  (void)par;  // the parameters select freeze directives: none here
  if (ncell == 0) return 0;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  const size_t vb = sizeof(double) * (size_t)ncell * NS;
  const size_t rb =
      sizeof(double) * (size_t)(ref_stride ? ncell : 1) * Net::NELEM;
  if (!dc.n.ensure(vb) || !dc.cons.ensure(rb) ||
      !dc.flags.ensure(sizeof(int) * ncell))
    return kx::FLAG_DEVICE_MEMORY;
  if (!hip_ok(hipMemcpy(dc.n.p, v, vb, hipMemcpyHostToDevice), "H2D") ||
      !hip_ok(hipMemcpy(dc.cons.p, ref, rb, hipMemcpyHostToDevice), "H2D"))
    return -1;
  int* df = static_cast<int*>(dc.flags.p);
  const int r = conserve_resident(ncell, static_cast<double*>(dc.n.p),
                                  (long)NS, 1L,
                                  static_cast<const double*>(dc.cons.p),
                                  (long)ref_stride, df, x, dc);
  if (r == -1 || r == kx::FLAG_DEVICE_MEMORY) return r;
  if (!hip_ok(hipMemcpy(v, dc.n.p, vb, hipMemcpyDeviceToHost), "D2H"))
    return -1;
  if (flags && !hip_ok(hipMemcpy(flags, df, sizeof(int) * ncell,
                                 hipMemcpyDeviceToHost), "D2H"))
    return -1;
  return r;
}

// krome_conserve_cells(_x)_device (krome.cpp:1306-1343,
// solve_kokkos.hpp:1787-1822): arrays on the device
int conserve_device(int ncell, double* v, int layout, long ld,
                    const double* ref, int ref_stride, const double* par,
                    int par_stride, int* flags, bool x) {
  if (ncell < 0 || !v || !ref ||
      (ref_stride != 0 && ref_stride != Net::NELEM) ||
      (par && par_stride != 0 && par_stride != Net::NPAR) ||
      (layout != KROME_LAYOUT_SPECIES_MAJOR &&
       layout != KROME_LAYOUT_CELL_MAJOR) ||
      (ld != 0 && ld < ncell))
    return -1;
  if (ncell == 0) return 0;
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  DevCache& dc = g_cache;
  int* fl = flags;
  if (!fl) {
    if (!dc.flags.ensure(sizeof(int) * ncell)) return kx::FLAG_DEVICE_MEMORY;
    fl = static_cast<int*>(dc.flags.p);
  }
  long cs, ss;
  strides(layout, ld, ncell, cs, ss);
  return conserve_resident(ncell, v, cs, ss, ref, (long)ref_stride, fl, x,
                           dc);
}

}  // namespace

extern "C" {

int krome_get_rates(int ncell, const double* n, const double* Tgas,
                    const double* par, int par_stride, const double* Phi,
                    const krome_opts* h, double* k) {
  if (!k) return -1;
  return query_host(ncell, n, Tgas, par, par_stride, Phi, h, kx::Q_RATES, k,
                    nullptr, nullptr, nullptr, nullptr, nullptr);
}
int krome_get_fluxes(int ncell, const double* n, const double* Tgas,
                     const double* par, int par_stride, const double* Phi,
                     const krome_opts* h, double* fluxes) {
  if (!fluxes) return -1;
  return query_host(ncell, n, Tgas, par, par_stride, Phi, h, kx::Q_FLUXES,
                    fluxes, nullptr, nullptr, nullptr, nullptr, nullptr);
}
int krome_nheat(void) { return Net::NHEAT; }
int krome_ncool(void) { return Net::NCOOL; }
const char* krome_heating_name(int i) {
  return (i >= 0 && i < Net::NHEAT) ? kHeatNames[i] : nullptr;
}
const char* krome_cooling_name(int i) {
  return (i >= 0 && i < Net::NCOOL) ? kCoolNames[i] : nullptr;
}
int krome_get_thermal(int ncell, const double* n, const double* Tgas,
                      const double* par, int par_stride, const double* Phi,
                      const krome_opts* h, double* heat, double* cool,
                      double* heats, double* cools, double* dTdt) {
  return query_host(ncell, n, Tgas, par, par_stride, Phi, h, kx::Q_THERMAL,
                    nullptr, heat, cool, heats, cools, dTdt);
}

int krome_get_element_totals(int ncell, const double* n, double* tot) {
  return totals_host(ncell, n, tot, false);
}
int krome_get_element_totals_x(int ncell, const double* x, double* tot) {
  return totals_host(ncell, x, tot, true);
}
int krome_conserve_cells(int ncell, double* n, const double* ref,
                         int ref_stride, const double* par, int par_stride,
                         int* flags) {
  return conserve_host(ncell, n, ref, ref_stride, par, par_stride, flags,
                       false);
}
int krome_conserve_cells_x(int ncell, double* x, const double* ref,
                           int ref_stride, const double* par,
                           int par_stride, int* flags) {
  return conserve_host(ncell, x, ref, ref_stride, par, par_stride, flags,
                       true);
}
int krome_conserve_cells_device(int ncell, double* n, int layout, long ld,
                                const double* ref, int ref_stride,
                                const double* par, int par_stride,
                                int* flags) {
  return conserve_device(ncell, n, layout, ld, ref, ref_stride, par,
                         par_stride, flags, false);
}
int krome_conserve_cells_x_device(int ncell, double* x, int layout, long ld,
                                  const double* ref, int ref_stride,
                                  const double* par, int par_stride,
                                  int* flags) {
  return conserve_device(ncell, x, layout, ld, ref, ref_stride, par,
                         par_stride, flags, true);
}

}  // extern "C"
