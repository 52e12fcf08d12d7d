// The discrete adjoint of the per-cell BDF solve: gradients of a cell's
// final state (n, Tgas at t_end) with respect to the logarithms of its
// rate coefficients (k_r -> k_r exp(theta_r)) and to its initial state.
//
// What is differentiated.  KROME-X's integrator (runtime/core/cvode/
// bdf_cell.hpp) is CVODE's variable-order, variable-step BDF in Nordsieck
// form.  Its control (step sizes, orders, which attempts fail, when the
// Jacobian is reused) is taken as it happened in the forward run, frozen;
// what remains is a sequence of maps on the Nordsieck array zn[0..5]:
//   - linear ones: the rescaling zn[j] *= eta^j (bdf_cell.hpp:878-887), the
//     Pascal predictor and its undo (:889-900, the restore of a failed
//     attempt), the order changes (:841-876);
//   - the corrector of each accepted step (bdf_cell.hpp:958-1065,
//     1386-1400): acor solves G(acor) = rl1 zn[1] + acor - gamma f(zn[0] +
//     acor) = 0 (the Newton iteration's limit), then zn[j] += l_j acor, and
//     zn[5] = acor where the step keeps it for an order increase;
//   - zn[1] = h f(zn[0]) at the start of each solve and of each restart
//     (bdf_cell.hpp:566-638, 1513-1547) and after a third error-test
//     failure at order 1 (:1371-1381);
//   - the output y(t_end) = sum_j s^j zn[j] (:1492-1500), then the
//     clipping of negative densities (solve_cell.hpp:2180-2190).
// The adjoint runs these backwards.  A corrector's adjoint, from the
// implicit function theorem on G = 0: with abar = sum_j l_j zbar_j
// (+ zbar_5 where kept), mu = (I - gamma J)^-T abar, then zbar_0 += mu -
// abar (= gamma J^T mu), zbar_1 -= rl1 mu, thetabar += gamma f_theta^T mu;
// J is the exact Jacobian at the step's state (dual numbers, kx_dual.h).
// f_theta^T w, the rates' part: reaction r's flux times the stoichiometric
// sum of w over its species, plus w_T (gamma_ad - 1)/(k_B ntot) k_r
// d(heat - cool)/dk_r for the CHEM heating and cooling (the generated
// chem_dk), the only thermal terms that read the rates.
//
// Memory (the forward is recomputed, not stored).  The forward run keeps
// a checkpoint of the integrator's whole state (kx_cell.h: save) every
// checkpoint_steps accepted steps (32 by default) and at every (re)start,
// in a pool shared by the cells (one atomic per checkpoint, each linked to
// the cell's previous one).  The backward pass walks a cell's checkpoints
// from the last: it reloads one, re-runs the forward to the next (the
// same code on the same state: the same steps, bit for bit), recording the
// segment's operations on a tape of its own, and runs the tape backwards.
// Memory: the pool (checkpoints x (64 + 8N + N^2 + the Newton matrix)
// doubles) plus one segment's tape per cell, instead of every step.
#pragma once
#include "kx_adjmath.h"
#include "kx_cell.h"
#include "kx_plain.h"

namespace kx {

// the tape of one cell's segment: records of TW doubles (the operation,
// the order, its coefficients) at rec[(r TW + f) S], and the states of its
// corrector and zn[1] = h f(zn[0]) records at ys[(k N + i) S].  In the
// forward run it only counts (store false), so that the recomputation of a
// segment ends where the forward run took the next checkpoint.
enum : int {
  OP_RESCALE = 1, OP_PASCAL, OP_ORDER_UP, OP_ORDER_DOWN, OP_STEP, OP_ZN1F
};
constexpr int TW = 12;

template <int N>
struct AdjTape {
  static constexpr bool ACTIVE = true;
  double* rec = nullptr;
  double* ys = nullptr;
  long S = 1;
  int cap = 0, ycap = 0;
  bool store = false;
  int nrec = 0, ny = 0, nsteps = 0, last_step = -1;
  bool overflow = false;

  KX_HD void reset() {
    nrec = ny = nsteps = 0;
    last_step = -1;
  }
  KX_HD double& f(int r, int k) const { return rec[((long)r * TW + k) * S]; }
  // a new record (nullptr-free: -1 when full, the cell then fails)
  KX_HD int put(int op, int q) {
    if (nrec >= cap) {
      overflow = true;
      return -1;
    }
    const int r = nrec++;
    if (store) {
      f(r, 0) = op;
      f(r, 1) = q;
    }
    return r;
  }
  KX_HD int put_y(const double* y) {
    if (ny >= ycap) {
      overflow = true;
      return -1;
    }
    const int k = ny++;
    if (store)
      for (int i = 0; i < N; ++i) ys[((long)k * N + i) * S] = y[i];
    return k;
  }
  KX_HD void rescale(double eta, int q) {
    const int r = put(OP_RESCALE, q);
    if (store && r >= 0) f(r, 2) = eta;
  }
  KX_HD void pascal(int sign, int q) {
    const int r = put(OP_PASCAL, q);
    if (store && r >= 0) f(r, 2) = sign;
  }
  KX_HD void order_up(int q, int L, double A1, const double* l) {
    const int r = put(OP_ORDER_UP, q);
    if (store && r >= 0) {
      f(r, 2) = A1;
      f(r, 3) = L;
      for (int j = 0; j < LMAX; ++j) f(r, 4 + j) = l[j];
    }
  }
  KX_HD void order_down(int q, const double* l) {
    const int r = put(OP_ORDER_DOWN, q);
    if (store && r >= 0)
      for (int j = 0; j < LMAX; ++j) f(r, 4 + j) = l[j];
  }
  KX_HD void step(int q, double gamma, double rl1, const double* l,
                  const double* y, bool saved) {
    const int r = put(OP_STEP, q);
    const int k = put_y(y);
    ++nsteps;
    last_step = r;
    if (store && r >= 0) {
      f(r, 2) = gamma;
      f(r, 3) = rl1;
      for (int j = 0; j < LMAX; ++j) f(r, 4 + j) = l[j];
      f(r, 10) = saved ? 1.0 : 0.0;
      f(r, 11) = k;
    }
  }
  // zn[5] = acor of the last accepted step (an order increase chosen)
  KX_HD void saveq() {
    if (store && last_step >= 0) f(last_step, 10) = 1.0;
  }
  KX_HD void zn1f(double h, const double* y) {
    const int r = put(OP_ZN1F, 0);
    const int k = put_y(y);
    if (store && r >= 0) {
      f(r, 2) = h;
      f(r, 11) = k;
    }
  }
  // the segment has reached its length: steps, or tape room for one more
  // step with every failure its options allow
  KX_HD bool full(int seg_steps, int margin, int ymargin) const {
    return nsteps >= seg_steps || nrec + margin > cap || ny + ymargin > ycap;
  }
};

// the arguments of one adjoint launch (one chunk of cells).  n and the
// cotangents species-major (element i of cell c at i * ld + c), lnk and
// g_lnk cell-major [ncell][NREA].  Without cotangents (gn and gT null)
// the kernel only integrates (the forward of the adjoint's own code).
struct AdjArgs {
  CellsArgs a;
  long ld;
  const double* lnk;
  const double* gn;
  const double* gT;
  double* g_lnk;
  double* g_n0;
  double* g_T0;
  // the workspace: S cells' integrator storage, rate scales, cotangents
  // and tapes (adj_cell_doubles each, interleaved) and their pivots
  double* work;
  int* iwork;
  long S;
  // the checkpoint pool: slots of ck_doubles, its next free slot
  double* pool;
  long pool_slots;
  unsigned long long* pool_next;
  int seg_steps, cap, ycap, margin;
};

// the workspace doubles of a cell: the integrator's, the rate scales, the
// Nordsieck cotangents, the rates' gradient and the tape
template <class Net, bool ET>
KX_HD constexpr long adj_cell_doubles(int cap, int ycap) {
  using Prob = AdjProblem<Net, ET>;
  return cell_doubles<Prob>() + Net::NREA + (long)LMAX * Prob::N +
         Net::NREA + (long)cap * TW + (long)ycap * Prob::N;
}
// a checkpoint: its header (the previous slot + 1, whether it starts a
// solve, the steps of the solve so far) and the integrator's state
constexpr int CK_HEAD = 4;
template <class Net, bool ET>
KX_HD constexpr long adj_ck_doubles() {
  using Prob = AdjProblem<Net, ET>;
  return CK_HEAD + Cell<Prob::N, Prob, AdjTape<Prob::N>>::CK_DOUBLES;
}

// begin() and iter() of the adjoint's integrator, each compiled once: the
// forward run and the recomputation of a segment call the same machine
// code (two inlined copies could contract their multiply-adds differently
// and take different steps)
#if defined(__HIPCC__)
#define KX_ONCE __host__ __device__ __attribute__((noinline))
#else
#define KX_ONCE __attribute__((noinline))
#endif
template <class CellT>
KX_ONCE int adj_begin(CellT& cv, double tout, double* t) {
  return cv.begin(tout, t);
}
template <class CellT>
KX_ONCE int adj_iter(CellT& cv, double tout, double* y, double* t,
                     long* nloc) {
  return cv.iter(tout, y, t, *nloc);
}
// the next free checkpoint slot (on the host, one cell at a time)
KX_HD inline unsigned long long adj_next_slot(unsigned long long* next) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(next, 1ull);
#else
  return (*next)++;
#endif
}

// one cell's adjoint solve (the kernel's thread c; also callable on the
// host, cell by cell, for the CPU checks)
template <class Net, bool ET>
KX_HD inline void adjoint_cell(const AdjArgs& g, int c) {
  using Prob = AdjProblem<Net, ET>;
  using CellT = Cell<Prob::N, Prob, AdjTape<Prob::N>>;
  constexpr int NS = Net::NSPEC;
  constexpr int N = Prob::N;
  constexpr int R = Net::NREA;
  const CellsArgs& a = g.a;
  const long S = g.S;
  const bool bwd = g.gn || g.gT;
  // where a cotangent comes from and a gradient goes (0 without one)
  auto gn_of = [&](int i) { return g.gn ? g.gn[i * g.ld + c] : 0.0; };
  if (!(a.dt[c] > 0.0)) {  // untouched: the identity
    a.flags[c] = 0;
    if (bwd) {
      for (int i = 0; i < NS; ++i) g.g_n0[i * g.ld + c] = gn_of(i);
      if (ET && g.g_T0) g.g_T0[c] = g.gT ? g.gT[c] : 0.0;
      for (int r = 0; r < R; ++r) g.g_lnk[(long)c * R + r] = 0.0;
    }
    return;
  }
  double* w = g.work + c;
  const long celld = cell_doubles<Prob>();
  double* sc = w + celld * S;           // exp(theta_r): sc[r S]
  double* zb = sc + (long)R * S;        // zbar_j[i] at zb[(j N + i) S]
  double* gth = zb + (long)LMAX * N * S;  // thetabar_r at gth[r S]
  double* tape = gth + (long)R * S;
  double* ys = tape + (long)g.cap * TW * S;
  Prob P;
  P.p = Net::params_from_row(a.par + (long)c * a.par_stride);
  P.Tfix = a.T[c];
  P.hit_T0 = false;
  P.sc = nullptr;
  P.scs = S;
  if (g.lnk) {
    for (int r = 0; r < R; ++r) sc[r * S] = ::exp(g.lnk[(long)c * R + r]);
    P.sc = sc;
  }
  CellOpts<N> co;
  if (!cell_opts<Net, ET>(a.o, co)) {
    a.flags[c] = FLAG_UNSUPPORTED;
    return;
  }
  double y[N];
  for (int i = 0; i < NS; ++i) y[i] = a.n[c * a.n_cs + i * a.n_ss];
  if (ET) y[N - 1] = a.T[c];
  CellT cv(P, w, g.iwork + c, S, co);
  cv.tp.rec = tape;
  cv.tp.ys = ys;
  cv.tp.S = S;
  cv.tp.cap = g.cap;
  cv.tp.ycap = g.ycap;
  cv.tp.store = false;
  constexpr long CKD = adj_ck_doubles<Net, ET>();
  const double tout = a.dt[c];
  // ---- the forward run, with checkpoints when a gradient is wanted ----
  long last = -1;        // the cell's latest checkpoint slot
  bool lost = false;     // the pool or a tape ran out
  auto checkpoint = [&](bool at_begin, long nloc) {
    if (!bwd || lost) return;
    const unsigned long long k = adj_next_slot(g.pool_next);
    if ((long)k >= g.pool_slots) {
      lost = true;
      return;
    }
    double* ck = g.pool + (long)k * CKD;
    ck[0] = (double)(last + 1);
    ck[1] = at_begin ? 1.0 : 0.0;
    ck[2] = (double)nloc;
    cv.save(ck + CK_HEAD, 1);
    last = (long)k;
  };
  cv.reinit(0.0, y, 0.0);
  double t = 0.0;
  int restarts = 0, flag = 0;
  for (;;) {  // bdf::integrate's restart loop (bdf_cell.hpp:1513-1547)
    checkpoint(true, 0);
    cv.tp.reset();
    int f = adj_begin(cv, tout, &t);
    if (f == BDF_OK) {
      long nloc = 0;
      for (;;) {
        if (bwd && cv.tp.full(g.seg_steps, g.margin, 16)) {
          checkpoint(false, nloc);
          cv.tp.reset();
        }
        f = adj_iter(cv, tout, y, &t, &nloc);
        if (f != CellT::ITER_CONTINUE) break;
      }
    }
    if (f >= 0) {
      flag = 0;
      break;
    }
    const bool again = f == ERR_FAILURE || f == CONV_FAILURE ||
                       f == REPTD_RHSFUNC_ERR;
    if (!again || restarts >= co.max_restarts) {
      flag = f;
      break;
    }
    ++restarts;
    const double hc = cv.current_step();
    cv.reinit(t, y, hc > 0.0 ? 0.1 * hc : 0.0);
  }
  if (flag == 0 && ET && !(y[N - 1] > 0.0 && kx::finite(y[N - 1])))
    flag = FLAG_NONPOSITIVE_T;
  if (P.hit_T0 && (flag == REPTD_RHSFUNC_ERR || flag == RHSFUNC_FAIL ||
                   flag == UNREC_RHSFUNC_ERR || flag == FIRST_RHSFUNC_ERR))
    flag = FLAG_NONPOSITIVE_T;
  if (flag == 0 && bwd && (lost || cv.tp.overflow)) flag = FLAG_DEVICE_MEMORY;
  a.flags[c] = flag;
  if (flag == 0) cell_finish<Net, ET>(a, c, y);
  if (!bwd) return;
  for (int r = 0; r < R; ++r) gth[r * S] = 0.0;
  if (flag != 0) {  // the inputs kept: the identity, no rate gradient
    for (int i = 0; i < NS; ++i) g.g_n0[i * g.ld + c] = gn_of(i);
    if (ET && g.g_T0) g.g_T0[c] = g.gT ? g.gT[c] : 0.0;
    for (int r = 0; r < R; ++r) g.g_lnk[(long)c * R + r] = 0.0;
    return;
  }
  // ---- the backward pass, segment by segment from the last ------------
  // the cotangent of the output y(t_end) (zero where a density was
  // clipped)
  double ybar[N];
  for (int i = 0; i < NS; ++i)
    ybar[i] = (a.o.clip_negative && y[i] < 0.0) ? 0.0 : gn_of(i);
  if (ET) ybar[NS] = g.gT ? g.gT[c] : 0.0;
  auto Z = [&](int j, int i) -> double& { return zb[((long)j * N + i) * S]; };
  for (int j = 0; j < LMAX; ++j)
    for (int i = 0; i < N; ++i) Z(j, i) = 0.0;
  cv.tp.store = true;
  bool first = true;
  double Jm[N * N];
  for (long cur = last; cur >= 0;) {
    const double* ck = g.pool + cur * CKD;
    const long prev = (long)ck[0] - 1;
    const bool at_begin = ck[1] != 0.0;
    long nloc = (long)ck[2];
    cv.load(ck + CK_HEAD, 1);
    cv.tp.reset();
    int f = BDF_OK;
    if (at_begin) f = adj_begin(cv, tout, &t);
    if (f == BDF_OK)
      while (!cv.tp.full(g.seg_steps, g.margin, 16)) {
        f = adj_iter(cv, tout, y, &t, &nloc);
        if (f != CellT::ITER_CONTINUE) break;
      }
    if (first) {  // the output's interpolation at t_end
      double cj[LMAX];
      cv.dky_weights(tout, cj);
      for (int j = 0; j < LMAX; ++j)
        for (int i = 0; i < N; ++i) Z(j, i) = cj[j] * ybar[i];
      first = false;
    }
    const AdjTape<N>& tp = cv.tp;
    for (int r = tp.nrec - 1; r >= 0; --r) {
      const int op = (int)tp.f(r, 0), q = (int)tp.f(r, 1);
      if (op == OP_RESCALE) {
        const double eta = tp.f(r, 2);
        double cc = eta;
        for (int j = 1; j <= q; j++) {
          for (int i = 0; i < N; ++i) Z(j, i) *= cc;
          cc = eta * cc;
        }
      } else if (op == OP_PASCAL) {
        // zn[j-1] +-= zn[j] for k = 1..q, j = q..k; backwards: zbar_j
        // +-= zbar_{j-1} for k = q..1, j = k..q
        const double sg = tp.f(r, 2);
        for (int k = q; k >= 1; k--)
          for (int j = k; j <= q; j++)
            for (int i = 0; i < N; ++i) Z(j, i) += sg * Z(j - 1, i);
      } else if (op == OP_ORDER_UP) {
        // zn[L] = A1 zn[5]; zn[j] += l_j zn[L] (j = 2..q)
        const double A1 = tp.f(r, 2);
        const int L = (int)tp.f(r, 3);
        for (int i = 0; i < N; ++i) {
          double u = Z(L, i);
          for (int j = 2; j <= q; j++) u += tp.f(r, 4 + j) * Z(j, i);
          if (L == QMAX) {
            Z(QMAX, i) = A1 * u;
          } else {
            Z(L, i) = 0.0;
            Z(QMAX, i) += A1 * u;
          }
        }
      } else if (op == OP_ORDER_DOWN) {
        // zn[j] -= l_j zn[q] (j = 2..q-1)
        for (int i = 0; i < N; ++i) {
          double u = 0.0;
          for (int j = 2; j < q; j++) u += tp.f(r, 4 + j) * Z(j, i);
          Z(q, i) -= u;
        }
      } else if (op == OP_STEP) {
        const double gam = tp.f(r, 2), rl1 = tp.f(r, 3);
        const bool saved = tp.f(r, 10) != 0.0;
        const int k = (int)tp.f(r, 11);
        double yn[N], ab[N], mu[N];
        for (int i = 0; i < N; ++i) yn[i] = ys[((long)k * N + i) * S];
        for (int i = 0; i < N; ++i) {
          double s = saved ? Z(QMAX, i) : 0.0;
          for (int j = 0; j <= q; j++) s += tp.f(r, 4 + j) * Z(j, i);
          ab[i] = s;
          if (saved) Z(QMAX, i) = 0.0;
        }
        jac_exact(P, yn, Jm);
        solve_transposed<N>(Jm, gam, ab, mu);
        for (int i = 0; i < N; ++i) {
          Z(0, i) += mu[i] - ab[i];
          Z(1, i) -= rl1 * mu[i];
        }
        theta_vjp<Net, ET>(P, yn, mu, gam, gth, S);
      } else if (op == OP_ZN1F) {
        const double hh = tp.f(r, 2);
        const int k = (int)tp.f(r, 11);
        double yn[N], v[N];
        for (int i = 0; i < N; ++i) {
          yn[i] = ys[((long)k * N + i) * S];
          v[i] = Z(1, i);
          Z(1, i) = 0.0;
        }
        jac_exact(P, yn, Jm);
        for (int jj = 0; jj < N; ++jj) {
          double s = 0.0;
          for (int i = 0; i < N; ++i) s += Jm[i + jj * N] * v[i];
          Z(0, jj) += hh * s;
        }
        theta_vjp<Net, ET>(P, yn, v, hh, gth, S);
      }
    }
    // a (re)start: zn[1..5] were set there, only zn[0] came from before
    if (at_begin)
      for (int j = 1; j < LMAX; ++j)
        for (int i = 0; i < N; ++i) Z(j, i) = 0.0;
    cur = prev;
  }
  for (int i = 0; i < NS; ++i) g.g_n0[i * g.ld + c] = Z(0, i);
  if (ET && g.g_T0) g.g_T0[c] = Z(0, NS);
  for (int r = 0; r < R; ++r) g.g_lnk[(long)c * R + r] = gth[r * S];
}

#if defined(__HIPCC__)
template <class Net, bool ET>
__global__ void __launch_bounds__(64) adjoint_cells_kernel(AdjArgs g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < g.a.ncell) adjoint_cell<Net, ET>(g, c);
}
#endif

}  // namespace kx
