// Several generated networks in one persistent launch: the body of a multi
// library, included by the file krome-x_amd/build.py generates for a
// network list (csrc/generated/multi_<name>.hip), which defines
//   KX_MULTI_NAME          the library's name
//   KX_MULTI_NETS(X)       X(index, network, Tgas evolved) per network, in
//                          launch order (the costliest cells first); index
//                          is the network's number in the C interface
//   KX_MULTI_COUNT         the number of networks
// after including each network's header.  The reference generates one
// library per network, so this entry point (include/krome_multi.h) has no
// counterpart there; each batch is solved as the per-network library's
// persistent kernel would solve it (device_kernel 1).
//
// One launch of multi_kernel, KX_PERSIST_BLOCK lanes per workgroup: every
// wave works through the networks in launch order.  A network of up to
// KX_PERSIST_MAX_N equations runs in the lockstep persistent machine
// (persist_run: the wave's lanes share each rhs evaluation, refilled from
// the network's queue); a larger one in the lane queue (lane_queue_run:
// each lane solves whole cells with the plain kernel's code, taking the
// next from the network's queue), and a lane that finds one lane-queue
// network's queue exhausted goes on to the next network by itself.  So
// the waves that finish a network early fill the launch with the next
// instead of leaving each network's tail to a launch of its own.  The
// queues hand out each network's densest cells first.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "kx_persist.h"
#include "kx_plain.h"
#include "kx_util.h"
#include "../../include/krome_multi.h"

namespace kx {

// one network's part of a multi launch
template <class Net, bool ET>
struct MultiPart {
  static constexpr int N = Problem<Net, ET>::N;
  static constexpr bool LOCKSTEP = N <= KX_PERSIST_MAX_N;
  // a large network at fixed Tgas: a wave per cell
  static constexpr bool WAVE = !LOCKSTEP && Net::NSPEC > KX_BIG_N && !ET;
  // the C interface's mode: 1 lockstep, 2 a wave per cell, 0 lane queue
  static constexpr int MODE = LOCKSTEP ? 1 : (WAVE ? 2 : 0);
  // doubles of the block's LDS its lanes use: the lockstep machine's
  // arena, or each wave's LU row accumulator (its own 64 columns of the
  // block's rows)
  static constexpr long LDS =
      LOCKSTEP ? lds_doubles<Net, ET>()
               : (WAVE ? wave_lds_rows(Net::NSPEC) * KX_PERSIST_BLOCK
                       : 0);
  CellsArgs a;
  int on;  // cells in this launch
  // the lockstep machine: its workspace (FP64 or FP32 matrices) and queue
  PersistWorkOf<Net, ET, false> wd;
  PersistWorkOf<Net, ET, true> wm;
  PersistQueue pq;
  // the lane queue: the queue and the slots' interleaved workspace
  LaneQueue lq;
  double* work;
  int* iwork;

};

// this wave's (lockstep) or this lane's (lane queue) share of one
// network's cells; MIXED: the FP32 Newton matrix where it applies (Tgas
// evolved with the difference-quotient Jacobian).  The part's fields come
// by value, straight from the kernel's arguments, so that the compiler
// keeps the workspace addresses uniform (scalar registers)
template <class Net, bool ET, bool MIXED>
__device__ __attribute__((always_inline)) inline void multi_run(
    CellsArgs a, int on, PersistWorkOf<Net, ET, false> wd,
    PersistWorkOf<Net, ET, true> wm, PersistQueue pq, LaneQueue lq,
    double* work, int* iwork, double* lds, long slot, long lanes) {
  if (!on) return;
  if constexpr (MultiPart<Net, ET>::LOCKSTEP) {
    (void)lq; (void)work; (void)iwork; (void)slot; (void)lanes;
    if constexpr (MIXED && ET) {
      const bool dq = a.o.jac == 2 || a.o.jac == 0;
      if (dq) persist_run<Net, ET, true>(a, wm, pq, lds);
      else persist_run<Net, ET, false>(a, wd, pq, lds);
    } else {
      (void)wm;
      persist_run<Net, ET, false>(a, wd, pq, lds);
    }
  } else if constexpr (MultiPart<Net, ET>::WAVE) {
    (void)wd; (void)wm; (void)pq; (void)lanes;
    const WaveLds lw{lds + (threadIdx.x & ~63), (long)KX_PERSIST_BLOCK};
    wave_queue_run<Net>(a, lq, work, iwork, slot >> 6, lw);
  } else {
    (void)wd; (void)wm; (void)pq; (void)lds;
    lane_queue_run<Net, ET>(a, lq, work, iwork, slot);
  }
}

}  // namespace kx

namespace {

using kx::MultiPart;

struct MultiArgs {
#define KX_M_FIELD(i, nm, et) MultiPart<kx::net_##nm, et> p_##nm;
  KX_MULTI_NETS(KX_M_FIELD)
#undef KX_M_FIELD
  long lanes;
};

constexpr long multi_lds() {
  long m = 1;
#define KX_M_LDS(i, nm, et) \
  m = MultiPart<kx::net_##nm, et>::LDS > m ? MultiPart<kx::net_##nm, et>::LDS : m;
  KX_MULTI_NETS(KX_M_LDS)
#undef KX_M_LDS
  return m;
}

template <bool MIXED>
__global__ void __launch_bounds__(KX_PERSIST_BLOCK, KX_PERSIST_WAVES_PER_EU)
multi_kernel(MultiArgs g) {
  __shared__ double lds[multi_lds()];
  const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
#define KX_M_RUN(i, nm, et)                                             \
  kx::multi_run<kx::net_##nm, et, MIXED>(                               \
      g.p_##nm.a, g.p_##nm.on, g.p_##nm.wd, g.p_##nm.wm, g.p_##nm.pq,   \
      g.p_##nm.lq, g.p_##nm.work, g.p_##nm.iwork, lds, slot, g.lanes);
  KX_MULTI_NETS(KX_M_RUN)
#undef KX_M_RUN
}

// the C interface's view of the networks, by index
struct NetInfo {
  const char* name;
  int nspec, npar, et, mode;
};
const NetInfo kNets[KX_MULTI_COUNT] = {
#define KX_M_INFO(i, nm, et)                                          \
  {kx::net_##nm::NAME, kx::net_##nm::NSPEC, kx::net_##nm::NPAR, et, \
   MultiPart<kx::net_##nm, et>::MODE},
    KX_MULTI_NETS(KX_M_INFO)
#undef KX_M_INFO
};
// the launch-order position of network index i
int order_of(int i) {
  int k = 0;
#define KX_M_ORD(j, nm, et) \
  if (j == i) return k;     \
  ++k;
  KX_MULTI_NETS(KX_M_ORD)
#undef KX_M_ORD
  return -1;
}

std::mutex g_mutex;
hipEvent_t g_ev[2] = {nullptr, nullptr};
int* g_count = nullptr;
int g_wpc = 0, g_chunk = 0;  // krome_multi_set_launch

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "krome_multi %s (HIP): %s: %s\n", KX_MULTI_NAME,
               what, hipGetErrorString(e));
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
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};
// each network's workspace and queue order, by launch position
Buf g_work[KX_MULTI_COUNT], g_ord[KX_MULTI_COUNT];

// a batch as the kernels see it: species-major n[i ncell + c]; the solver
// options the batch names, the others at the per-network defaults
kx::CellsArgs cells_args(const krome_multi_batch& b) {
  kx::CellsArgs a;
  a.ncell = b.ncell;
  a.n = b.n;
  a.n_cs = 1;
  a.n_ss = b.ncell;
  a.T = b.Tgas;
  a.dt = b.dt;
  a.par = b.par;
  a.par_stride = kNets[order_of(b.network)].npar;
  a.flags = b.flags;
  a.stats = b.stats;
  kx::SolveOpts& o = a.o;
  o.rtol = b.rtol;
  o.atol_n = b.atol_n;
  o.atol_T = b.atol_T;
  o.max_steps = b.max_steps;
  o.max_err_fails = 7;
  o.max_conv_fails = 10;
  o.max_restarts = 5;
  o.evolve_T = b.evolve_T;
  o.jac = b.jac;
  o.clip_negative = 1;
  o.conserve = 0;
  o.linsol = 0;
  o.tol = nullptr;
  o.rtol_v = 0;
  return a;
}

// REQUIRED parameters unset in any cell of a batch: -105 (its flags set)
template <class Net>
int params_check(const krome_multi_batch& b) {
  if (!g_count && !hip_ok(hipMalloc(&g_count, sizeof(int)), "hipMalloc"))
    return -1;
  const int np = kx::params_check_device<Net>(b.ncell, b.par, Net::NPAR,
                                              b.flags, g_count);
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

// network part p's workspace and queue in the library's buffers (launch
// position k), for a launch of `lanes` lanes; 0 or a flag
template <class Net, bool ET>
int prepare_part(MultiPart<Net, ET>& p, const krome_multi_batch* b, int k,
                 long lanes, int mixed, hipStream_t s) {
  p.on = b && b->ncell > 0;
  if (!p.on) return 0;
  p.a = cells_args(*b);
  kx::PersistMem m;
  size_t wb;
  if constexpr (MultiPart<Net, ET>::LOCKSTEP) {
    wb = kx::persist_need<Net, ET>(p.a, mixed);
  } else if constexpr (MultiPart<Net, ET>::WAVE) {
    wb = kx::wave_queue_bytes<Net>(lanes / 64);
  } else {
    wb = kx::lane_queue_bytes<Net, ET>(lanes);
  }
  if (!g_work[k].ensure(wb) ||
      !g_ord[k].ensure(kx::persist_order_bytes(p.a.ncell)))
    return kx::FLAG_DEVICE_MEMORY;
  m.work = g_work[k].p;
  m.ord = g_ord[k].p;
  if constexpr (MultiPart<Net, ET>::LOCKSTEP) {
    if (kx::persist_mixed<ET>(p.a, mixed))
      return kx::persist_prepare<Net, ET, true>(p.a, m, g_chunk, s, p.wm,
                                                p.pq);
    return kx::persist_prepare<Net, ET, false>(p.a, m, g_chunk, s, p.wd,
                                               p.pq);
  } else if constexpr (MultiPart<Net, ET>::WAVE) {
    return kx::wave_queue_prepare<Net>(p.a, m, lanes / 64, s, p.lq, p.work,
                                       p.iwork);
  } else {
    return kx::lane_queue_prepare<Net, ET>(p.a, m, lanes, s, p.lq, p.work,
                                           p.iwork);
  }
}

}  // namespace

extern "C" {

int krome_multi_nnet(void) { return KX_MULTI_COUNT; }
const char* krome_multi_network_name(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? kNets[order_of(i)].name : nullptr;
}
int krome_multi_nspec(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? kNets[order_of(i)].nspec : -1;
}
int krome_multi_npar(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? kNets[order_of(i)].npar : -1;
}
int krome_multi_evolve_T(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? kNets[order_of(i)].et : -1;
}
int krome_multi_lockstep(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? (kNets[order_of(i)].mode == 1)
                                        : -1;
}
int krome_multi_mode(int i) {
  return (i >= 0 && i < KX_MULTI_COUNT) ? kNets[order_of(i)].mode : -1;
}

int krome_multi_set_launch(int waves_per_cu, int queue_chunk) {
  if (waves_per_cu < 0 || waves_per_cu > 64 || queue_chunk < 0 ||
      queue_chunk > 1 << 20)
    return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  g_wpc = waves_per_cu;
  g_chunk = queue_chunk;
  return 0;
}

int krome_multi_finalize(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  for (int k = 0; k < KX_MULTI_COUNT; ++k) {
    g_work[k].release();
    g_ord[k].release();
  }
  return 0;
}

int krome_multi_solve_device(const krome_multi_batch* b, int nbatch,
                             int mixed_precision, double* kernel_ms) {
  if (!b || nbatch < 1 || nbatch > KX_MULTI_COUNT) return -1;
  // by launch position
  const krome_multi_batch* at[KX_MULTI_COUNT] = {};
  for (int k = 0; k < nbatch; ++k) {
    const krome_multi_batch& x = b[k];
    if (x.network < 0 || x.network >= KX_MULTI_COUNT || x.ncell < 0 ||
        at[order_of(x.network)] ||
        (x.ncell > 0 && (!x.n || !x.Tgas || !x.dt || !x.par || !x.flags)) ||
        x.jac < 0 || x.jac > 2)
      return -1;
    // each network is built with Tgas fixed or evolved
    if ((x.evolve_T != 0) != (kNets[order_of(x.network)].et != 0))
      return kx::FLAG_UNSUPPORTED;
    at[order_of(x.network)] = &x;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!hip_ok(hipDeviceSynchronize(), "fence")) return -1;
  long most = 0;
  {
    int k = 0, r = 0;
  // the lanes the launch needs: a cell per lane, or a wave per cell
#define KX_M_CHECK(i, nm, et)                                           \
  if (r == 0 && at[k] && at[k]->ncell > 0) {                            \
    r = params_check<kx::net_##nm>(*at[k]);                             \
    const long need = at[k]->ncell *                                    \
        (MultiPart<kx::net_##nm, et>::WAVE ? 64L : 1L);                  \
    if (need > most) most = need;                                       \
  }                                                                     \
  ++k;
    KX_MULTI_NETS(KX_M_CHECK)
#undef KX_M_CHECK
    if (r != 0) return r;
  }
  if (most == 0) return 0;
  long lanes = 0;
  if (kx::persist_lanes(most, g_wpc, lanes) != 0) return -1;
  if (!g_ev[0] && (!hip_ok(hipEventCreate(&g_ev[0]), "event") ||
                   !hip_ok(hipEventCreate(&g_ev[1]), "event")))
    return -1;
  hipStream_t s = nullptr;
  MultiArgs g;
  g.lanes = lanes;
  {
    int k = 0, r = 0;
#define KX_M_PREP(i, nm, et) \
  if (r == 0) r = prepare_part(g.p_##nm, at[k], k, lanes, mixed_precision, s); \
  ++k;
    KX_MULTI_NETS(KX_M_PREP)
#undef KX_M_PREP
    if (r != 0) return r;
  }
  (void)hipEventRecord(g_ev[0], s);
  constexpr long B = KX_PERSIST_BLOCK;
  if (mixed_precision)
    hipLaunchKernelGGL(multi_kernel<true>, dim3(lanes / B), dim3(B), 0, s, g);
  else
    hipLaunchKernelGGL(multi_kernel<false>, dim3(lanes / B), dim3(B), 0, s,
                       g);
  if (!hip_ok(hipGetLastError(), "launch")) return -1;
  (void)hipEventRecord(g_ev[1], s);
  if (!hip_ok(hipEventSynchronize(g_ev[1]), "kernel")) return -1;
  if (kernel_ms) {
    float f = 0.0f;
    if (!hip_ok(hipEventElapsedTime(&f, g_ev[0], g_ev[1]), "elapsed"))
      return -1;
    *kernel_ms = f;
  }
  for (int k = 0; k < KX_MULTI_COUNT; ++k)
    if (at[k] && at[k]->ncell > 0) {
      const int f = first_flag(*at[k]);
      if (f != 0) return f;
    }
  return 0;
}

}  // extern "C"
