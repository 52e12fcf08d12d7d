// The plain FP64 drop-in kernel: one thread per cell, each running the
// whole per-cell solve (KROME-X solve(), runtime/core/solve_cell.hpp:
// 2094-2203, over the integrator of kx_cell.h), the counterpart of the
// reference's Kokkos kernel "krome_bdf_cells" (runtime/core/
// solve_kokkos.hpp:972-1039) with its interleaved workspace layout.
#pragma once
#include "kx_cell.h"
#include "kx_problem.h"

namespace kx {

// where a call's arrays live and how the species array is laid out:
// element (c, i) of n at n[c * n_cs + i * n_ss] (cell-major: n_cs = NSPEC,
// n_ss = 1; species-major: n_cs = 1, n_ss = ld)
struct CellsArgs {
  int ncell;
  double* n;
  long n_cs, n_ss;
  double* T;
  const double* dt;
  const double* par;
  int par_stride;     // 0: one row for every cell
  int* flags;         // ncell, always written
  long* stats;        // ncell x 9 or nullptr
  SolveOpts o;
};

// the integrator options of a cell (solve_cell.hpp:1662-1705: the
// tolerances per slot, the DQ Jacobian with evolved Tgas, the charge
// invariant with the DQ Jacobian and electrons, charge_newton_ 1631-1642);
// false for what the kernels do not support.  Prob: the problem type the
// cell is solved with (Problem, or BigProblem for a large network)
template <class Net, bool ET, class Prob = Problem<Net, ET>>
KX_HD inline bool cell_opts(const SolveOpts& o, CellOpts<Prob::N>& co) {
  constexpr int NS = Net::NSPEC;
  constexpr int N = Prob::N;
  co.rtol = o.rtol;
  co.rtol_v = o.tol != nullptr && o.rtol_v != 0;
  co.max_steps = o.max_steps;
  co.max_err_fails = o.max_err_fails;
  co.max_conv_fails = o.max_conv_fails;
  co.max_restarts = o.max_restarts;
  co.dq_jac = o.jac == 2 || (o.jac == 0 && ET);
  co.inv_row = -1;
  co.inv_n = 0;
  if constexpr (Prob::STRIDED) {
    // a large network: Tgas fixed, the analytic Jacobian into the sparse
    // LU (linsol auto resolves to it: sparse_auto_ picks the sparse LU
    // for these networks), no LAPACK / dense LU
    co.atol_n = o.atol_n;
    co.tol = o.tol;
    co.sparse = true;
    return !ET && !co.dq_jac && o.linsol != 1 && o.linsol != 2;
  } else {
    for (int i = 0; i < NS; ++i) co.atol[i] = o.atol_n;
    if (ET) co.atol[N - 1] = o.atol_T;
    if (o.tol != nullptr)
      for (int i = 0; i < N; ++i) {
        co.atol[i] = o.tol[i];
        co.rtolv[i] = co.rtol_v ? o.tol[N + i] : o.rtol;
      }
    // the linear solver (solve_cell.hpp:1601-1620): 3 sparse; 0 (auto)
    // sparse where its factorization is at most a tenth of the dense one
    // and the border at most 4 equations; else dense
    co.sparse = o.linsol == 3 ||
                (o.linsol == 0 && N - NS <= 4 &&
                 2.0 * NS * NS * NS / 3.0 >= 10.0 * Net::SP_FLOPS);
    for (int i = 0; i < N; ++i) co.inv_c[i] = i < NS ? Net::charge(i) : 0;
    if (co.dq_jac && Net::HAS_E) {
      co.inv_row = Net::I_E;
      co.inv_n = NS;
    }
    // the analytic Jacobian: Problem::jac (every network, Tgas fixed or
    // evolved)
    return true;
  }
}

// the post-solve of a successful cell: clip, conservation (solve_cell.hpp:
// 1458-1490, modes 1 and 3), write back
template <class Net, bool ET>
KX_HD inline void cell_finish(const CellsArgs& a, int c, const double* y) {
  constexpr int NS = Net::NSPEC;
  double n0[NS], n[NS];
  for (int i = 0; i < NS; ++i) {
    n0[i] = a.n[c * a.n_cs + i * a.n_ss];
    n[i] = y[i];
    if (a.o.clip_negative && n[i] < 0.0) n[i] = 0.0;
  }
  if (Net::HAS_E && a.o.conserve == 1) {
    double q0 = 0.0, q = 0.0;
    for (int i = 0; i < NS; ++i) {
      q0 += Net::charge(i) * n0[i];
      q += Net::charge(i) * n[i];
    }
    n[Net::I_E] = ::fmax(n[Net::I_E] + (q - q0), 0.0);
  } else if (Net::HAS_E && a.o.conserve == 3) {
    double qi = 0.0;
    for (int i = 0; i < NS; ++i)
      if (i != Net::I_E) qi += Net::charge(i) * n[i];
    n[Net::I_E] = ::fmax(qi, 0.0);
  }
  for (int i = 0; i < NS; ++i) a.n[c * a.n_cs + i * a.n_ss] = n[i];
  if (ET) a.T[c] = y[NS];
}

// the whole per-cell solve of cell c by one thread, its workspace at work
// and iwork (this thread's slot of an interleaved workspace, stride S);
// host-callable too (the CPU tests run the device's code with it)
template <class Net, bool ET>
KX_HD inline void plain_cell(const CellsArgs& a, int c, double* work,
                             int* iwork, long S) {
  using Prob = PlainProblem<Net, ET>;
  constexpr int NS = Net::NSPEC;
  constexpr int N = Prob::N;
  long* sc = a.stats ? a.stats + (long)c * 9 : nullptr;
  if (!(a.dt[c] > 0.0)) {  // dt <= 0: untouched (solve_cells.hpp:96-101)
    a.flags[c] = 0;
    if (sc)
      for (int i = 0; i < 9; ++i) sc[i] = 0;
    return;
  }
  Prob P;
  P.p = Net::params_from_row(a.par + (long)c * a.par_stride);
  P.Tfix = a.T[c];
  P.hit_T0 = false;
  CellOpts<N> co;
  if (!cell_opts<Net, ET, Prob>(a.o, co)) {
    a.flags[c] = FLAG_UNSUPPORTED;
    return;
  }
  double y[N];
  for (int i = 0; i < NS; ++i) y[i] = a.n[c * a.n_cs + i * a.n_ss];
  if (ET) y[N - 1] = a.T[c];
  Cell<N, Prob> cv(P, work, iwork, S, co);
  if constexpr (Prob::STRIDED) {
    P.k = SView{cv.rates_rows(), S};
    P.k_set = false;
  }
  cv.reinit(0.0, y, 0.0);
  CellStats tot;
  tot.zero();
  double t = 0.0;
  int restarts = 0, flag = 0;
  for (;;) {  // bdf::integrate's restart loop (bdf_cell.hpp:1513-1547)
    const int f = cv.solve(a.dt[c], y, &t);
    tot.add(cv.st);
    if (f >= 0) { flag = 0; break; }
    const bool again = f == ERR_FAILURE || f == CONV_FAILURE ||
                       f == REPTD_RHSFUNC_ERR;
    if (!again || restarts >= co.max_restarts) { flag = f; break; }
    ++restarts;
    const double hc = cv.current_step();
    cv.reinit(t, y, hc > 0.0 ? 0.1 * hc : 0.0);
  }
  if (flag == 0 && ET && !(y[N - 1] > 0.0 && kx::finite(y[N - 1])))
    flag = FLAG_NONPOSITIVE_T;
  if (P.hit_T0 && (flag == REPTD_RHSFUNC_ERR || flag == RHSFUNC_FAIL ||
                   flag == UNREC_RHSFUNC_ERR || flag == FIRST_RHSFUNC_ERR))
    flag = FLAG_NONPOSITIVE_T;
  a.flags[c] = flag;
  if (sc) {
    sc[0] = tot.nst; sc[1] = tot.nfe; sc[2] = tot.nje; sc[3] = tot.nfeDQ;
    sc[4] = tot.nsetups; sc[5] = tot.nni; sc[6] = tot.nnf; sc[7] = tot.netf;
    sc[8] = restarts;
  }
  if (flag == 0) cell_finish<Net, ET>(a, c, y);
}

// The one-lane-per-cell kernels' workspace in blocks of 64 slots: element
// i of slot c at ((c / 64) D + i) 64 + c % 64 (D doubles or ints a slot).
// A wave's 64 lanes still read 64 consecutive values, and one cell's data
// span D x 512 bytes of its block rather than D x (slots x 8) bytes of the
// whole workspace, so a lane left alone at a launch's end touches a few
// pages, not one page per access
constexpr long SLOT_BLOCK = 64;
KX_HD constexpr long blocked_slot(long c, long D) {
  return (c / SLOT_BLOCK) * SLOT_BLOCK * D + c % SLOT_BLOCK;
}
// slots a blocked workspace holds for n cells (whole blocks)
KX_HD constexpr long blocked_slots(long n) {
  return (n + SLOT_BLOCK - 1) / SLOT_BLOCK * SLOT_BLOCK;
}

#if defined(__HIPCC__)
template <class Net, bool ET>
__global__ void __launch_bounds__(128)
plain_cells_kernel(CellsArgs a, double* work, int* iwork) {
  using Prob = PlainProblem<Net, ET>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  plain_cell<Net, ET>(a, c, work + blocked_slot(c, cell_doubles<Prob>()),
                      iwork + blocked_slot(c, Prob::N), SLOT_BLOCK);
}

// The persistent form of the one-thread-per-cell solve, for the networks
// too large for the lockstep kernel's registers (more than
// KX_PERSIST_MAX_N equations): every lane takes the next cell of a device
// work queue (one atomic per cell, over a permutation that hands out the
// densest, costliest cells first), solves it with plain_cell in its own
// workspace slot, and takes another until the queue is exhausted.  A lane
// that finishes early takes the next cell at once, so the lanes do not
// wait for the slowest cell of their first batch; the launch holds only
// the resident lanes' workspaces (slot = the lane's global index, stride
// S = the launch's lanes), whatever the number of cells.
struct LaneQueue {
  int* counter;       // the next queue position (zeroed before the launch)
  const int* order;   // queue position -> cell
};

template <class Net, bool ET>
__device__ __attribute__((always_inline)) inline void lane_queue_run(
    const CellsArgs& a, LaneQueue q, double* work, int* iwork, long slot) {
  using Prob = PlainProblem<Net, ET>;
  double* w = work + blocked_slot(slot, cell_doubles<Prob>());
  int* iw = iwork + blocked_slot(slot, Prob::N);
  for (;;) {
    const int i = atomicAdd(q.counter, 1);
    if (i >= a.ncell) break;
    plain_cell<Net, ET>(a, q.order[i], w, iw, SLOT_BLOCK);
  }
}

// the waves per SIMD the lane queue's build must allow (its registers:
// 1 leaves the compiler all of them)
#ifndef KX_LANE_WAVES_PER_EU
#define KX_LANE_WAVES_PER_EU 1
#endif
template <class Net, bool ET>
__global__ void __launch_bounds__(128)
__attribute__((amdgpu_waves_per_eu(KX_LANE_WAVES_PER_EU)))
lane_queue_kernel(CellsArgs a, LaneQueue q, double* work, int* iwork,
                  long S) {
  const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= S) return;
  lane_queue_run<Net, ET>(a, q, work, iwork, slot);
}

// ---- a wave per cell (networks above KX_BIG_N species) ------------------
// doubles of a wave slot: the cell's workspace (contiguous), then its state
template <class Net>
constexpr long wave_slot_doubles() {
  return cell_doubles<WaveProblem<Net>>() + Net::NSPEC;
}

// the whole per-cell solve of cell c by the calling wave, all 64 lanes
// (Tgas fixed): its workspace at work (wave_slot_doubles), pivots at iwork,
// the LU's row accumulator in lw.  Flags and counters from lane 0; the
// densities read and written by the lane that owns each species
template <class Net>
__device__ __attribute__((always_inline)) inline void wave_cell(const CellsArgs& a, int c, double* work,
                                 int* iwork, WaveLds lw) {
  using Prob = WaveProblem<Net>;
  constexpr int NS = Net::NSPEC;
  constexpr int N = Prob::N;
  const int l = wave_lane();
  long* sc = a.stats ? a.stats + (long)c * 9 : nullptr;
  if (!(a.dt[c] > 0.0)) {  // dt <= 0: untouched (solve_cells.hpp:96-101)
    if (l == 0) {
      a.flags[c] = 0;
      if (sc)
        for (int i = 0; i < 9; ++i) sc[i] = 0;
    }
    return;
  }
  Prob P;
  P.p = Net::params_from_row(a.par + (long)c * a.par_stride);
  P.Tfix = a.T[c];
  P.hit_T0 = false;
  CellOpts<N> co;
  if (!cell_opts<Net, false, Prob>(a.o, co)) {
    if (l == 0) a.flags[c] = FLAG_UNSUPPORTED;
    return;
  }
  double* y = work + cell_doubles<Prob>();
  for (int i = l; i < NS; i += 64) y[i] = a.n[c * a.n_cs + i * a.n_ss];
  Cell<N, Prob> cv(P, work, iwork, 1, co);
  cv.lw = lw;
  P.k = cv.rates_rows();
  P.k_set = false;
  cv.reinit(0.0, y, 0.0);
  CellStats tot;
  tot.zero();
  double t = 0.0;
  int restarts = 0, flag = 0;
  for (;;) {  // bdf::integrate's restart loop (bdf_cell.hpp:1513-1547)
    const int f = cv.solve(a.dt[c], y, &t);
    tot.add(cv.st);
    if (f >= 0) { flag = 0; break; }
    const bool again = f == ERR_FAILURE || f == CONV_FAILURE ||
                       f == REPTD_RHSFUNC_ERR;
    if (!again || restarts >= co.max_restarts) { flag = f; break; }
    ++restarts;
    const double hc = cv.current_step();
    cv.reinit(t, y, hc > 0.0 ? 0.1 * hc : 0.0);
  }
  wave_sync();
  if (l == 0) {
    a.flags[c] = flag;
    if (sc) {
      sc[0] = tot.nst; sc[1] = tot.nfe; sc[2] = tot.nje; sc[3] = tot.nfeDQ;
      sc[4] = tot.nsetups; sc[5] = tot.nni; sc[6] = tot.nnf; sc[7] = tot.netf;
      sc[8] = restarts;
    }
  }
  if (flag != 0) return;
  // cell_finish split over the lanes (solve_cell.hpp:1458-1490, modes 1
  // and 3; the charge sums reduced by the butterfly)
  double q0 = 0.0, q = 0.0, qi = 0.0;
  for (int i = l; i < NS; i += 64) {
    const double n0 = a.n[c * a.n_cs + i * a.n_ss];
    double v = y[i];
    if (a.o.clip_negative && v < 0.0) v = 0.0;
    y[i] = v;
    q0 += Net::charge(i) * n0;
    q += Net::charge(i) * v;
    if (!(Net::HAS_E && i == Net::I_E)) qi += Net::charge(i) * v;
  }
  q0 = wave_sum(q0);
  q = wave_sum(q);
  qi = wave_sum(qi);
  wave_sync();
  if constexpr (Net::HAS_E) {
    if (l == (Net::I_E & 63)) {
      if (a.o.conserve == 1) y[Net::I_E] = ::fmax(y[Net::I_E] + (q - q0), 0.0);
      else if (a.o.conserve == 3) y[Net::I_E] = ::fmax(qi, 0.0);
    }
  }
  wave_sync();
  for (int i = l; i < NS; i += 64) a.n[c * a.n_cs + i * a.n_ss] = y[i];
}

// each wave takes the next cell of the device work queue (lane 0's
// atomic, broadcast), solves it with its 64 lanes and takes another until
// the queue is exhausted; a workgroup holds blockDim / 64 cells at a time
// ("cells per workgroup"), each wave with its own slot of the workspace
// and its own 64 columns of the block's LDS (NSPEC / 64 + 1 rows)
template <class Net>
__device__ __attribute__((always_inline)) inline void wave_queue_run(const CellsArgs& a, LaneQueue q,
                                      double* work, int* iwork, long slot,
                                      WaveLds lw) {
  constexpr long CD = wave_slot_doubles<Net>();
  for (;;) {
    int i = 0;
    if (wave_lane() == 0) i = atomicAdd(q.counter, 1);
    i = __shfl(i, 0, 64);
    if (i >= a.ncell) break;
    wave_cell<Net>(a, q.order[i], work + slot * CD,
                   iwork + slot * Net::NSPEC, lw);
  }
}

// MAXB: the largest workgroup the instantiation serves.  Up to 8 cells
// (512 lanes) a workgroup its waves may hold 256 VGPRs (two waves per
// SIMD); 16 cells (1024 lanes) allow 128 (four per SIMD)
template <class Net, int MAXB>
__global__ void __launch_bounds__(MAXB)
wave_queue_kernel(CellsArgs a, LaneQueue q, double* work, int* iwork,
                  long slots) {
  extern __shared__ double wave_lds[];
  const long slot = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (slot >= slots) return;
  const WaveLds lw{wave_lds + (threadIdx.x & ~63), (long)blockDim.x};
  wave_queue_run<Net>(a, q, work, iwork, slot, lw);
}
#endif

}  // namespace kx
