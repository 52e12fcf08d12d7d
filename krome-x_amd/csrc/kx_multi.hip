// Several generated networks in one persistent launch: the library
// libkrome_multi.so holds the primordial3 and primordial networks and one
// entry point that solves a batch of each in a single launch of
// persist_pair_kernel.  It stands beside the per-network libraries (whose
// C ABI is the reference's, include/krome.h); the reference has one
// network per generated library, so this entry point has no counterpart
// there.  Every wave solves primordial3 cells until that queue is empty,
// then primordial cells, so the waves that finish the first network early
// fill the launch with the second instead of leaving each launch's tail
// to itself.  A cell's arithmetic is the same as in its own network's
// persistent launch: the results are bitwise those of the per-network
// libraries' persistent kernel with the same options.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "net_primordial.h"
#include "net_primordial3.h"
#include "kx_plain.h"
#include "kx_persist.h"
#include "kx_util.h"
#include "../../include/krome_multi.h"

namespace {

using NetA = kx::net_primordial3;  // first in the launch: the costlier cells
using NetB = kx::net_primordial;
constexpr int NNET = 2;
const char* const kNames[NNET] = {NetB::NAME, NetA::NAME};

std::mutex g_mutex;
hipEvent_t g_ev[2] = {nullptr, nullptr};
int* g_count = nullptr;

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "krome_multi (HIP): %s: %s\n", what,
               hipGetErrorString(e));
  return false;
}

// a device buffer grown as needed and kept (guarded by g_mutex)
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
};
// each network's persistent workspace and queue order
Buf g_work[NNET], g_ord[NNET];

// the memory of network k's persistent launch
kx::PersistMem persist_mem(int k, size_t wb, size_t ob, int& r) {
  kx::PersistMem m;
  r = 0;
  if (!g_work[k].ensure(wb) || !g_ord[k].ensure(ob)) {
    r = kx::FLAG_DEVICE_MEMORY;
    return m;
  }
  m.work = g_work[k].p;
  m.ord = g_ord[k].p;
  return m;
}

int nspec_of(int net) { return net == 0 ? NetB::NSPEC : NetA::NSPEC; }
int npar_of(int net) { return net == 0 ? NetB::NPAR : NetA::NPAR; }

// the batch as the kernels see it: species-major n[i ncell + c]
kx::CellsArgs cells_args(const krome_multi_batch& b) {
  kx::CellsArgs a;
  a.ncell = b.ncell;
  a.n = b.n;
  a.n_cs = 1;
  a.n_ss = b.ncell;
  a.T = b.Tgas;
  a.dt = b.dt;
  a.par = b.par;
  a.par_stride = npar_of(b.network);
  a.flags = b.flags;
  a.stats = nullptr;
  kx::SolveOpts& o = a.o;
  o.rtol = b.rtol;
  o.atol_n = b.atol_n;
  o.atol_T = b.atol_T;
  o.max_steps = b.max_steps;
  o.max_err_fails = 7;
  o.max_conv_fails = 10;
  o.max_restarts = 5;
  o.evolve_T = b.evolve_T;
  o.jac = 0;
  o.clip_negative = 1;
  o.conserve = 0;
  o.linsol = 0;
  return a;
}

// REQUIRED parameters unset in any cell: -105 (its flags set), else 0
int params_check(const krome_multi_batch& b) {
  if (!g_count && !hip_ok(hipMalloc(&g_count, sizeof(int)), "hipMalloc"))
    return -1;
  const int np = b.network == 0
      ? kx::params_check_device<NetB>(b.ncell, b.par, NetB::NPAR, b.flags,
                                      g_count)
      : kx::params_check_device<NetA>(b.ncell, b.par, NetA::NPAR, b.flags,
                                      g_count);
  if (np < 0) return -1;
  return np > 0 ? kx::FLAG_PARAM_UNSET : 0;
}

int first_flag(const krome_multi_batch& b) {
  std::vector<int> f(b.ncell);
  if (!hip_ok(hipMemcpy(f.data(), b.flags, sizeof(int) * b.ncell,
                        hipMemcpyDeviceToHost), "D2H"))
    return -1;
  for (int v : f)
    if (v != 0) return v;
  return 0;
}

}  // namespace

extern "C" {

int krome_multi_nnet(void) { return NNET; }
const char* krome_multi_network_name(int i) {
  return (i >= 0 && i < NNET) ? kNames[i] : nullptr;
}
int krome_multi_nspec(int i) { return (i >= 0 && i < NNET) ? nspec_of(i) : -1; }
int krome_multi_npar(int i) { return (i >= 0 && i < NNET) ? npar_of(i) : -1; }

int krome_multi_solve_device(const krome_multi_batch* b, int nbatch,
                             int mixed_precision, double* kernel_ms) {
  if (!b || nbatch < 1 || nbatch > NNET) return -1;
  const krome_multi_batch* by_net[NNET] = {nullptr, nullptr};
  for (int k = 0; k < nbatch; ++k) {
    const krome_multi_batch& x = b[k];
    if (x.network < 0 || x.network >= NNET || by_net[x.network] ||
        x.ncell < 0 || (x.ncell > 0 && (!x.n || !x.Tgas || !x.dt ||
                                        !x.par || !x.flags)))
      return -1;
    // primordial3 is built with Tgas evolved, primordial at fixed Tgas
    if ((x.network == 1) != (x.evolve_T != 0)) return kx::FLAG_UNSUPPORTED;
    by_net[x.network] = &x;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  for (const krome_multi_batch* x : by_net)
    if (x && x->ncell > 0) {
      const int r = params_check(*x);
      if (r != 0) return r;
    }
  const krome_multi_batch* ba = by_net[1];  // primordial3
  const krome_multi_batch* bb = by_net[0];  // primordial
  const bool has_a = ba && ba->ncell > 0, has_b = bb && bb->ncell > 0;
  if (!has_a && !has_b) return 0;
  if (!g_ev[0] && (!hip_ok(hipEventCreate(&g_ev[0]), "event") ||
                   !hip_ok(hipEventCreate(&g_ev[1]), "event")))
    return -1;
  hipStream_t s = nullptr;
  (void)hipEventRecord(g_ev[0], s);
  int r;
  auto mem_of = [](int k) {
    return [k](size_t wb, size_t ob, kx::PersistMem& m) -> int {
      int r = 0;
      m = persist_mem(k, wb, ob, r);
      return r;
    };
  };
  if (has_a && has_b) {
    // primordial's Jacobian is analytic (Tgas fixed): its Newton matrix
    // stays FP64, as in launch_persistent
    const kx::CellsArgs xa = cells_args(*ba), xb = cells_args(*bb);
    const size_t wa = mixed_precision
                          ? kx::persist_work_bytes<NetA, true, true>()
                          : kx::persist_work_bytes<NetA, true, false>();
    int ra = 0, rb = 0;
    const kx::PersistMem ma =
        persist_mem(1, wa, kx::persist_order_bytes(xa.ncell), ra);
    const kx::PersistMem mb = persist_mem(
        0, kx::persist_work_bytes<NetB, false, false>(),
        kx::persist_order_bytes(xb.ncell), rb);
    if (ra != 0 || rb != 0) return ra != 0 ? ra : rb;
    r = mixed_precision
        ? kx::launch_persistent_pair_t<NetA, true, true, NetB, false, false>(
              xa, xb, 0, 0, s, ma, mb)
        : kx::launch_persistent_pair_t<NetA, true, false, NetB, false, false>(
              xa, xb, 0, 0, s, ma, mb);
  } else if (has_a) {
    r = kx::launch_persistent<NetA, true>(cells_args(*ba), mixed_precision,
                                          0, 0, s, mem_of(1));
  } else {
    r = kx::launch_persistent<NetB, false>(cells_args(*bb), mixed_precision,
                                           0, 0, s, mem_of(0));
  }
  (void)hipEventRecord(g_ev[1], s);
  if (r != 0) return r;
  if (!hip_ok(hipEventSynchronize(g_ev[1]), "kernel")) return -1;
  if (kernel_ms) {
    float f = 0.0f;
    if (!hip_ok(hipEventElapsedTime(&f, g_ev[0], g_ev[1]), "elapsed"))
      return -1;
    *kernel_ms = f;
  }
  for (const krome_multi_batch* x : {ba, bb})
    if (x && x->ncell > 0) {
      const int f = first_flag(*x);
      if (f != 0) return f;
    }
  return 0;
}

}  // extern "C"
