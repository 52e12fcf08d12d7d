// The persistent batched kernel: KROME-X's per-cell integrator
// (runtime/core/cvode/bdf_cell.hpp, CVODE's BDF; the per-cell solve of
// runtime/core/solve_cell.hpp:2094-2203) restated as a state machine that
// suspends at every right-hand-side evaluation.  A wave's 64 lanes each
// integrate a cell; every iteration of the kernel's loop evaluates ONE rhs
// per lane, at one program point, for all lanes together (lanes in any
// phase of the method: a Newton residual, a difference-quotient Jacobian
// column, the initial-step estimate), then advances each lane's machine
// to its next rhs request.  A lane whose cell is done takes the next cell
// from a device work queue (one atomic per wave), so the waves stay full
// while cells of very different stiffness finish at different times.
//
// The per-lane workspace (Nordsieck array, the method's vectors, the
// saved Jacobian and the factorized Newton matrix) lives in HBM,
// interleaved across the kernel's lanes (element i of lane s at i S + s:
// a wave's accesses are coalesced); the method's scalars live in
// registers.
//
// Precision: FP64 throughout, or (mixed) the Jacobian (FP64 difference
// quotients, stored in FP32), the Newton matrix, its LU factorization and
// the triangular solves in FP32, while the state, the Nordsieck history,
// the Newton residual (an FP64 rhs), the WRMS norms and the local error
// test stay FP64; the Newton iteration refines the FP32 solves.  The
// FP32 matrix is an equivalent form of I - gamma J that FP32 can hold
// (conservation rows, error-weight scaling: form_and_factor).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <utility>

#include "kx_cell.h"
#include "kx_plain.h"
#include "kx_problem.h"
#include "kx_util.h"

// the waves per SIMD the persistent kernel is compiled for (its register
// budget: 2 -> 256 VGPRs + AGPRs per lane)
#ifndef KX_PERSIST_WAVES_PER_EU
#define KX_PERSIST_WAVES_PER_EU 1
#endif
// the persistent kernel's lane slots (the workspace stride): one wave per
// SIMD on the 256 CUs of an MI355X
#ifndef KX_PERSIST_LANES
#define KX_PERSIST_LANES 65536
#endif

namespace kx {

// program counters of a lane's machine
enum Pc : int {
  P_FETCH = 0, P_SOLVE_BEGIN, P_FIRST_RHS, P_HIN_BEGIN, P_HIN_REQ, P_HIN_RET,
  P_HIN_END, P_AFTER_HIN, P_LOOP_TOP, P_STEP_BEGIN, P_ATTEMPT, P_NLS_BEGIN,
  P_NLS_OUTER, P_RESID_REQ, P_RESID_RET, P_LSETUP_BEGIN, P_DQ_BEGIN,
  P_DQ_REQ, P_DQ_RET, P_LSETUP_FORM, P_INNER_ITER,
  P_INNER_EXIT, P_NLS_EXIT, P_HANDLE, P_ERRTEST_RHS, P_STEP_DONE,
  P_FAIL_END, P_SOLVE_END, P_CELL_DONE, P_IDLE
};
// vectors of a lane's workspace (index into the interleaved arrays)
// (the first ones, the Nordsieck array, the error weights and the
// accumulated correction, in LDS when they fit; the others in HBM)
enum Vec : int {
  V_ZN0 = 0, V_EWT = 6, V_ACOR, V_Y, V_TEMPV, V_VT1,
  V_COUNT
};
// the rhs destination that is not a workspace vector: the Newton
// residual's f(y), kept in registers (persist_cells_kernel, fq)
constexpr int RQ_FTEMP = -1;
// the LDS of a workgroup: 160 KB on gfx950, KX_PERSIST_BLOCK lanes
#ifndef KX_PERSIST_BLOCK
#define KX_PERSIST_BLOCK 256
#endif
constexpr long LDS_BYTES = 163840;
// the vectors of a lane that live in LDS
template <int N>
constexpr int lds_vectors() {
  const long v = LDS_BYTES / (8L * KX_PERSIST_BLOCK * N);
  return v > 8 ? 8 : (int)v;
}

template <int N, class MT>
struct PersistWork {
  double* v;     // V_COUNT * N * S doubles
  MT* J;         // N * N * S
  MT* A;         // N * N * S
  int* piv;      // N * S
  long* tot;     // TOT_SLOTS * S: the cells' counter totals
  // the interleave stride: the kernel's lane slots, a compile-time
  // constant, so that the elements of a lane sit at constant offsets from
  // one base (the compiler then sees that distinct elements never alias)
  static constexpr long S = KX_PERSIST_LANES;
  struct LaneState* sp;  // this lane's scalars when kept in registers
  double* lds;    // lds_vectors<N>() * N * KX_PERSIST_BLOCK: this block's
  int tid;        // the lane's slot in it
  KX_HD double& at(int vec, int i, long s) const {
    constexpr int NL = lds_vectors<N>();
    if (vec < NL) return lds[((long)vec * N + i) * KX_PERSIST_BLOCK + tid];
    return v[((long)(vec - NL) * N + i) * S + s];
  }
  // mixed precision keeps J and A as pairs of FP32 elements when N is even
  // (flattened column-major index k in pair k / 2: 8-byte accesses, the
  // matrices moved with half the memory instructions)
  static constexpr bool PAIRS = std::is_same<MT, float>::value && N % 2 == 0;
  KX_HD static long mat_index(int k, long s) {
    return PAIRS ? ((long)(k >> 1) * S + s) * 2 + (k & 1) : (long)k * S + s;
  }
  KX_HD MT& Jat(int k, long s) const { return J[mat_index(k, s)]; }
  KX_HD MT& Aat(int k, long s) const { return A[mat_index(k, s)]; }
  KX_HD float2& Jpair(int p, long s) const {
    return reinterpret_cast<float2*>(J)[(long)p * S + s];
  }
  KX_HD float2& Apair(int p, long s) const {
    return reinterpret_cast<float2*>(A)[(long)p * S + s];
  }
};

// the scalars of a lane's machine (CVodeMem, the loop variables of the
// suspended functions, and the cell bookkeeping), held in registers: the
// integer members are 32-bit (a call's counters stay below max_steps
// times a few, cell indices below 2^31) to leave registers to the stages
struct LaneState {
  // CVodeMem
  double tn, h, hin, hprime, hscale, next_h, eta, etamax, etaq, etaqm1,
      etaqp1, saved_tq5;
  double tau[7], tq[6], l[6];
  double rl1, gamma, gammap, gamrat, crate, delp, delnrm, acnrm;
  int q, qprime, qwait, L, convfail;
  int nscon, nstlp, nstlj;
  int acnrmcur, jcur, nls_jcur;
  // loop variables of cvStep / cvNls / cvHin / the DQ Jacobian
  double saved_t, dsm, tol, cg, inc, minInc, tret;
  double hin_hg, hin_hs, hin_hnew, hin_hub, hin_hlb;
  int ncf, nef, nflag, niters, nconvfails, curiter, first_resid, callSetup,
      jbad, jcol, jr, count1, count2, nls_r;
  int nstloc;
  // counters of this CVode call
  int nst, nfe, nfeDQ, nje, nsetups, nni, nnf, ncfn, netf;
  int restarts, flag;
  // the cell
  int cell;
  double tout;
};
// a member of this lane (st), an element of an array member (sta)
#define ST(f) (w.sp->f)
#define STA(f, j) (w.sp->f[j])
// the cell's counter totals over its restarts, in HBM (touched at a
// call's end only): total k of lane s at k S + s
constexpr int TOT_SLOTS = 8;
#define TOT(k) (w.tot[(long)(k) * w.S + s])

// a per-step record of one chosen cell, in debugging builds only
#ifdef KX_TRACE
__device__ double kx_trace_buf[8 * 32768];
__device__ int kx_trace_n;
__device__ long kx_trace_cell = -1;
#define KX_TRACE_REC2(tag, a, b)                                          \
  if (ST(cell) == kx_trace_cell) {                                        \
    const int k_ = atomicAdd(&kx_trace_n, 1);                             \
    if (k_ < 32768) {                                                     \
      double* r_ = kx_trace_buf + 8 * k_;                                 \
      r_[0] = (tag); r_[1] = ST(tn); r_[2] = ST(h); r_[3] = (double)ST(q); \
      r_[4] = (a); r_[5] = (double)ST(nef); r_[6] = (double)ST(nst);      \
      r_[7] = (b);                                                        \
    }                                                                     \
  }
#define KX_TRACE_REC(tag, a) KX_TRACE_REC2(tag, a, ST(acnrm))
#else
#define KX_TRACE_REC2(tag, a, b)
#define KX_TRACE_REC(tag, a)
#endif

template <class Net, bool ET, bool MIXED>
struct Persist {
  using Prob = Problem<Net, ET>;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = Prob::N;
  using MT = typename std::conditional<MIXED, float, double>::type;
  using W = PersistWork<N, MT>;

  // ---- vector kernels over one lane's slot ------------------------------
  static KX_HD void axpy(const W& w, long s, double a, int x, int y) {
    if (a == 1.0) {
      for (int i = 0; i < N; i++) w.at(y, i, s) += w.at(x, i, s);
    } else if (a == -1.0) {
      for (int i = 0; i < N; i++) w.at(y, i, s) -= w.at(x, i, s);
    } else {
      for (int i = 0; i < N; i++) w.at(y, i, s) += a * w.at(x, i, s);
    }
  }
  // z = a x + b y with N_VLinearSum's special cases (bdf_cell.hpp:308-374)
  static KX_HD void lsum(const W& w, long s, double a, int x, double b, int y,
                         int z) {
    if (b == 1.0 && z == y) return axpy(w, s, a, x, z);
    if (a == 1.0 && z == x) return axpy(w, s, b, y, z);
    double xv[N], yv[N], zv[N];
    for (int i = 0; i < N; i++) {
      xv[i] = w.at(x, i, s);
      yv[i] = w.at(y, i, s);
    }
    if (a == 1.0 && b == 1.0) {
      for (int i = 0; i < N; i++) zv[i] = xv[i] + yv[i];
    } else if ((a == 1.0 && b == -1.0) || (a == -1.0 && b == 1.0)) {
      const bool t = a == 1.0;
      for (int i = 0; i < N; i++) zv[i] = t ? xv[i] - yv[i] : yv[i] - xv[i];
    } else if (a == 1.0 || b == 1.0) {
      const bool t = a == 1.0;
      const double c = t ? b : a;
      for (int i = 0; i < N; i++)
        zv[i] = t ? (c * yv[i]) + xv[i] : (c * xv[i]) + yv[i];
    } else if (a == -1.0 || b == -1.0) {
      const bool t = a == -1.0;
      const double c = t ? b : a;
      for (int i = 0; i < N; i++)
        zv[i] = t ? (c * yv[i]) - xv[i] : (c * xv[i]) - yv[i];
    } else if (a == b) {
      for (int i = 0; i < N; i++) zv[i] = a * (xv[i] + yv[i]);
    } else if (a == -b) {
      for (int i = 0; i < N; i++) zv[i] = a * (xv[i] - yv[i]);
    } else {
      for (int i = 0; i < N; i++) zv[i] = (a * xv[i]) + (b * yv[i]);
    }
    for (int i = 0; i < N; i++) w.at(z, i, s) = zv[i];
  }
  static KX_HD void scal(const W& w, long s, double c, int x, int z) {
    for (int i = 0; i < N; i++) w.at(z, i, s) = c * w.at(x, i, s);
  }
  static KX_HD void zero(const W& w, long s, int z) {
    for (int i = 0; i < N; i++) w.at(z, i, s) = 0.0;
  }
  static KX_HD double wrms(const W& w, long s, int x) {
    double sum = 0.0;
    for (int i = 0; i < N; i++) {
      const double p = w.at(x, i, s) * w.at(V_EWT, i, s);
      sum += p * p;
    }
    return bdfc::rt(sum / N);
  }
  // cvEwtSetSV into vector dst (bdf_cell.hpp:689-713)
  static KX_HD double atol(const SolveOpts& o, int i) {
    if (o.tol != nullptr) return o.tol[i];
    return i < NS ? o.atol_n : o.atol_T;
  }
  static KX_HD int ewt_set(const W& w, long s, const SolveOpts& o,
                           bool atolmin0, int src, int dst) {
    if (o.rtol_v) {  // per-species rtol (o.tol set)
      int bad = 0;
      for (int i = 0; i < N; i++) {
        const double d = ::fma(o.tol[N + i], ::fabs(w.at(src, i, s)),
                               o.tol[i]);
        bad |= !(d > 0.0);
        w.at(dst, i, s) = 1.0 / d;
      }
      return bad ? -1 : 0;
    }
    double v[N];
    for (int i = 0; i < N; i++) v[i] = ::fabs(w.at(src, i, s));
    for (int i = 0; i < N; i++) v[i] = (o.rtol * v[i]) + atol(o, i);
    if (atolmin0) {
      double m = v[0];
      for (int i = 1; i < N; i++) m = v[i] < m ? v[i] : m;
      if (m <= 0.0) return -1;
    }
    for (int i = 0; i < N; i++) w.at(dst, i, s) = 1.0 / v[i];
    return 0;
  }

  // ---- the BDF bookkeeping (bdf_cell.hpp:825-954, 1386-1488) ------------
  // The Nordsieck operations below run element by element with the
  // element's history in registers (each element of the array is read and
  // written once), the same operations in the same order per element as
  // the vector forms of bdf_cell.hpp.
  static KX_HD void rescale(const W& w, long s) {
    double c[QMAX + 1];
    c[1] = ST(eta);
#pragma unroll
    for (int j = 2; j <= QMAX; j++) c[j] = ST(eta) * c[j - 1];
    const int q = ST(q);
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
      for (int j = 1; j <= QMAX; j++)
        if (j <= q) w.at(j, i, s) = c[j] * w.at(j, i, s);
    ST(h) = ST(hscale) * ST(eta);
    ST(next_h) = ST(h);
    ST(hscale) = ST(h);
    ST(nscon) = 0;
  }
  // the Newton residual delta = rl1 zn[1] + acor - gamma ftemp: the two
  // N_VLinearSum calls of cvNlsResid (bdf_cell.hpp:1029-1030), their
  // special cases included, fused element by element
  static KX_HD void residual(const W& w, long s, double rl1, double gamma,
                             const double* fq, double* dl) {
    const double ng = -gamma;
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double x = w.at(1, i, s), y = w.at(V_ACOR, i, s);
      const double f = fq[i];
      const double t = rl1 == 1.0 ? x + y : (rl1 == -1.0 ? y - x : rl1 * x + y);
      dl[i] = ng == 1.0 ? t + f : (ng == -1.0 ? t - f : t + ng * f);
    }
  }
  // zn[j-1] += zn[j] (predict) or -= (restore) for k = 1..q, j = q..k
  template <bool ADD>
  static KX_HD void pascal(const W& w, long s) {
    const int q = ST(q);
#pragma unroll
    for (int i = 0; i < N; i++) {
      double z[QMAX + 1];
#pragma unroll
      for (int j = 0; j <= QMAX; j++) z[j] = j <= q ? w.at(j, i, s) : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; k++)
#pragma unroll
        for (int j = QMAX; j >= k; j--)
          if (k <= q && j <= q) z[j - 1] = ADD ? z[j - 1] + z[j] : z[j - 1] - z[j];
#pragma unroll
      for (int j = 0; j < QMAX; j++)
        if (j < q) w.at(j, i, s) = z[j];
    }
  }
  static KX_HD void predict(const W& w, long s) {
    ST(tn) += ST(h);
    pascal<true>(w, s);
  }
  static KX_HD void restore(const W& w, long s) {
    ST(tn) = ST(saved_t);
    pascal<false>(w, s);
  }
  static KX_HD void order_up(const W& w, long s) {
#pragma unroll
    for (int i = 0; i <= QMAX; i++) STA(l, i) = 0.0;
    double a1 = 1.0, prod = 1.0, xiold = 1.0, a0 = -1.0;
    STA(l, 2) = 1.0;
    double hsum = ST(hscale);
#pragma unroll
    for (int j = 1; j < QMAX; j++)
      if (j < ST(q)) {
        hsum += STA(tau, j + 1);
        const double xi = hsum / ST(hscale);
        prod *= xi;
        a0 -= 1.0 / (j + 1);
        a1 += 1.0 / xi;
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
          if (i <= j + 2) STA(l, i) = STA(l, i) * xiold + STA(l, i - 1);
        xiold = xi;
      }
    const double A1 = (-a0 - a1) / prod;
    scal(w, s, A1, QMAX, ST(L));
#pragma unroll
    for (int j = 2; j <= QMAX; j++)
      if (j <= ST(q)) axpy(w, s, STA(l, j), ST(L), j);
  }
  static KX_HD void order_down(const W& w, long s) {
#pragma unroll
    for (int i = 0; i <= QMAX; i++) STA(l, i) = 0.0;
    STA(l, 2) = 1.0;
    double hsum = 0.0;
#pragma unroll
    for (int j = 1; j <= QMAX - 2; j++)
      if (j <= ST(q) - 2) {
        hsum += STA(tau, j);
        const double xi = hsum / ST(hscale);
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
          if (i <= j + 2) STA(l, i) = STA(l, i) * xi + STA(l, i - 1);
      }
#pragma unroll
    for (int j = 2; j < QMAX; j++)
      if (j < ST(q)) axpy(w, s, -STA(l, j), ST(q), j);
  }
  static KX_HD void order_change(const W& w, long s, int dq) {
    if (ST(q) == 2 && dq != 1) return;
    if (dq == 1) order_up(w, s);
    else if (dq == -1) order_down(w, s);
  }
  static KX_HD void adjust(const W& w, long s) {
    if (ST(qprime) != ST(q)) {
      order_change(w, s, ST(qprime) - ST(q));
      ST(q) = ST(qprime);
      ST(L) = ST(q) + 1;
      ST(qwait) = ST(L);
    }
    rescale(w, s);
  }
  static KX_HD void set_coeffs(const W& w, long s) {
    const int q = ST(q);
    double xi_inv = 1.0, xistar_inv = 1.0, a0 = -1.0, a0h = -1.0;
    STA(l, 0) = STA(l, 1) = 1.0;
#pragma unroll
    for (int i = 2; i <= QMAX; i++)
      if (i <= q) STA(l, i) = 0.0;
    double hsum = ST(h);
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < QMAX; j++)
        if (j < q) {
          hsum += STA(tau, j - 1);
          xi_inv = ST(h) / hsum;
          a0 -= 1.0 / j;
#pragma unroll
          for (int i = QMAX; i >= 1; i--)
            if (i <= j) STA(l, i) += STA(l, i - 1) * xi_inv;
        }
      a0 -= 1.0 / q;
      xistar_inv = -STA(l, 1) - a0;
      double tq1 = 0.0;
#pragma unroll
      for (int j = 1; j <= QMAX; j++)
        if (j == q - 1) tq1 = STA(tau, j);
      hsum += tq1;
      xi_inv = ST(h) / hsum;
      a0h = -STA(l, 1) - xi_inv;
#pragma unroll
      for (int i = QMAX; i >= 1; i--)
        if (i <= q) STA(l, i) += STA(l, i - 1) * xistar_inv;
    }
    double lq = 0.0;
#pragma unroll
    for (int i = 1; i <= QMAX; i++)
      if (i == q) lq = STA(l, i);
    const double A1 = 1.0 - a0h + a0;
    const double A2 = 1.0 + q * A1;
    STA(tq, 2) = ::fabs(A1 / (a0 * A2));
    STA(tq, 5) = ::fabs(A2 * xistar_inv / (lq * xi_inv));
    if (ST(qwait) == 1) {
      if (q > 1) {
        const double C = xistar_inv / lq;
        const double A3 = a0 + 1.0 / q;
        const double A4 = a0h + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        STA(tq, 1) = ::fabs(C * Cpinv);
      } else {
        STA(tq, 1) = 1.0;
      }
      double tqq = 0.0;
#pragma unroll
      for (int j = 1; j <= QMAX; j++)
        if (j == q) tqq = STA(tau, j);
      hsum += tqq;
      xi_inv = ST(h) / hsum;
      const double A5 = a0 - (1.0 / (q + 1));
      const double A6 = a0h - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      STA(tq, 3) = ::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    STA(tq, 4) = bdfc::CORTES / STA(tq, 2);
    ST(rl1) = 1.0 / STA(l, 1);
    ST(gamma) = ST(h) * ST(rl1);
    if (ST(nst) == 0) ST(gammap) = ST(gamma);
    ST(gamrat) = ST(nst) > 0 ? ST(gamma) / ST(gammap) : 1.0;
  }
  static KX_HD void complete(const W& w, long s) {
    ST(nst)++;
    ST(nscon)++;
#pragma unroll
    for (int i = QMAX; i >= 2; i--)
      if (i <= ST(q)) STA(tau, i) = STA(tau, i - 1);
    if (ST(q) == 1 && ST(nst) > 1) STA(tau, 2) = STA(tau, 1);
    STA(tau, 1) = ST(h);
    {
      const int q = ST(q);
#pragma unroll
      for (int i = 0; i < N; i++) {
        const double a = w.at(V_ACOR, i, s);
#pragma unroll
        for (int j = 0; j <= QMAX; j++)
          if (j <= q) w.at(j, i, s) += STA(l, j) * a;
      }
    }
    ST(qwait)--;
    if (ST(qwait) == 1 && ST(q) != QMAX) {
      for (int i = 0; i < N; i++) w.at(QMAX, i, s) = w.at(V_ACOR, i, s);
      ST(saved_tq5) = STA(tq, 5);
    }
  }
  static KX_HD void set_eta(const W& w, long s) {
    using namespace bdfc;
    if (ST(eta) > ETA_MIN_FX && ST(eta) < ETA_MAX_FX) {
      ST(eta) = 1.0;
      ST(hprime) = ST(h);
      return;
    }
    if (ST(eta) >= ETA_MAX_FX) ST(eta) = mn(ST(eta), ST(etamax));
    else ST(eta) = mx(ST(eta), ETA_MIN);
    ST(hprime) = ST(h) * ST(eta);
    if (ST(qprime) < ST(q)) ST(nscon) = 0;
  }
  static KX_HD void next_step(const W& w, long s) {
    using namespace bdfc;
    if (ST(etamax) == 1.0) {
      ST(qwait) = ST(qwait) > 2 ? ST(qwait) : 2;
      ST(qprime) = ST(q);
      ST(hprime) = ST(h);
      ST(eta) = 1.0;
      return;
    }
    ST(etaq) = 1.0 / (kx::pow(BIAS2 * ST(dsm), 1.0 / ST(L)) + ADDON);
    if (ST(qwait) != 0) {
      ST(eta) = ST(etaq);
      ST(qprime) = ST(q);
      set_eta(w, s);
      return;
    }
    ST(qwait) = 2;
    ST(etaqm1) = 0.0;
    if (ST(q) > 1) {
      const double ddn = wrms(w, s, ST(q)) * STA(tq, 1);
      ST(etaqm1) = 1.0 / (kx::pow(BIAS1 * ddn, 1.0 / ST(q)) + ADDON);
    }
    ST(etaqp1) = 0.0;
    if (ST(q) != QMAX && ST(saved_tq5) != 0.0) {
      double p = 1.0;
      const double b = ST(h) / STA(tau, 2);
#pragma unroll
      for (int i = 1; i <= LMAX; i++)
        if (i <= ST(L)) p *= b;
      const double cquot = (STA(tq, 5) / ST(saved_tq5)) * p;
      lsum(w, s, -cquot, QMAX, 1.0, V_ACOR, V_TEMPV);
      const double dup = wrms(w, s, V_TEMPV) * STA(tq, 3);
      ST(etaqp1) = 1.0 / (kx::pow(BIAS3 * dup, 1.0 / (ST(L) + 1)) + ADDON);
    }
    const double etam = mx(ST(etaqm1), mx(ST(etaq), ST(etaqp1)));
    if (etam > ETA_MIN_FX && etam < ETA_MAX_FX) {
      ST(eta) = 1.0;
      ST(qprime) = ST(q);
    } else if (etam == ST(etaq)) {
      ST(eta) = ST(etaq);
      ST(qprime) = ST(q);
    } else if (etam == ST(etaqm1)) {
      ST(eta) = ST(etaqm1);
      ST(qprime) = ST(q) - 1;
    } else {
      ST(eta) = ST(etaqp1);
      ST(qprime) = ST(q) + 1;
      for (int i = 0; i < N; i++) w.at(QMAX, i, s) = w.at(V_ACOR, i, s);
    }
    set_eta(w, s);
  }

  // ---- the Newton matrix -------------------------------------------------
  // A = I - gamma J and its LU with partial pivoting (bdf_cell.hpp:
  // 1099-1108, 402-433), in the matrix precision.
  //
  // Mixed precision factors, in FP32, a matrix with the same solution
  // that FP32 can hold.  Where gamma J is stiff, the 1 of I - gamma J
  // sits below FP32's resolution of the stiff entries, and along the
  // combinations every reaction conserves (c_k^T J = 0: nuclei, charge)
  // the FP32 matrix is singular, so its solutions are garbage there.  For
  // each conserved combination one row, that of the species carrying the
  // most of it, is replaced by c_k^T (its right-hand side by c_k^T b): the
  // exact solution satisfies c_k^T x = c_k^T b since c_k^T A = c_k^T.
  // The other rows i are scaled as D A D^-1 (the similarity with the
  // error weights, D = diag(2^e_i), 2^e_i the power of two at or below
  // component i's weight), so the FP32 rounding is relative to what the
  // weighted convergence test sees; a replaced row is scaled by a power
  // of two to a largest entry in [1, 2).  Powers of two keep the scaling
  // exact.  Per row k, piv[k] packs the pivot row (bits 0-5), e_k + 1024
  // (bits 6-16), the law replacing row k plus one, 0 for none (bits
  // 17-19), and the replaced row's scale exponent + 1024 (bits 20-30).
  static constexpr int NC = MIXED ? Net::NCONS : 0;
  static_assert(!MIXED || (N <= 64 && Net::NCONS <= 7), "piv packing");
  static KX_HD int pack(int pv, int e, int law, int rs) {
    return pv | ((e + 1024) << 6) | (law << 17) | ((rs + 1024) << 20);
  }
  static KX_HD int unpack_pv(int p) { return MIXED ? (p & 63) : p; }
  static KX_HD int unpack_e(int p) { return ((p >> 6) & 2047) - 1024; }
  static KX_HD int unpack_law(int p) { return (p >> 17) & 7; }
  static KX_HD int unpack_rs(int p) { return ((p >> 20) & 2047) - 1024; }
  static KX_HD int clamp_exp(int k) {
    return k < -1000 ? -1000 : (k > 1000 ? 1000 : k);
  }

  // the scaling exponents and the replaced rows of the mixed-precision
  // matrix (see above)
  static KX_HD void mixed_scaling(const W& w, long s, int* e, int* law,
                                  int* rs) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      e[i] = clamp_exp(::ilogb(w.at(V_EWT, i, s)));
      law[i] = rs[i] = 0;
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
      // the row replaced: the species carrying the most of law k
      int r = -1;
      double best = -1.0;
#pragma unroll
      for (int i = 0; i < NS; i++)
        if (Net::cons(k, i) != 0 && law[i] == 0) {
          const double v = ::fabs(Net::cons(k, i) * w.at(V_Y, i, s));
          if (v > best) {
            best = v;
            r = i;
          }
        }
      // its scale: the largest entry c_kj 2^-e_j into [1, 2)
      int top = -2000;
#pragma unroll
      for (int j = 0; j < NS; j++)
        if (Net::cons(k, j) != 0) {
          const int t = ::ilogb((double)Net::cons(k, j)) - e[j];
          top = t > top ? t : top;
        }
#pragma unroll
      for (int i = 0; i < NS; i++)
        if (i == r) {
          law[i] = k + 1;
          rs[i] = clamp_exp(-top);
        }
    }
  }

  // networks of up to REG_LU_N equations factorize in registers; larger
  // ones in place in the workspace (the register arrays would spill)
  static constexpr int REG_LU_N = 10;
  static KX_HD int form_and_factor(const W& w, long s, double gamma) {
    if constexpr (N > REG_LU_N) {
      return form_and_factor_mem(w, s, gamma);
    } else if constexpr (MIXED) {
      return form_and_factor_regs(w, s, gamma);
    } else {
      // FP64: I - gamma J and its LU with the reference's operations in
      // its order, in registers (one load and one store of the matrix)
      double a[N * N];
#pragma unroll
      for (int k = 0; k < N * N; k++) a[k] = w.Jat(k, s);
      const double c = -gamma;
#pragma unroll
      for (int j = 0; j < N; j++) {
#pragma unroll
        for (int i = 0; i < N; i++) a[i + j * N] = a[i + j * N] * c;
        a[j + j * N] += 1.0;
      }
      int fail = 0;
      lu64_steps(a, fail, w, s, std::make_integer_sequence<int, N>{});
#pragma unroll
      for (int k = 0; k < N * N; k++) w.Aat(k, s) = a[k];
      return fail;
    }
  }

  // the Newton matrix formed and factorized in place in the workspace
  // (either precision; mixed: the scaled matrix of form_and_factor_regs)
  static KX_HD int form_and_factor_mem(const W& w, long s, double gamma) {
    int e[N], law[N], rs[N];
#pragma unroll
    for (int i = 0; i < N; i++) e[i] = law[i] = rs[i] = 0;
    if constexpr (MIXED) {
      mixed_scaling(w, s, e, law, rs);
      const float gf = (float)(-gamma);
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const float jv = w.Jat(i + j * N, s);
          float v = i == j ? (float)(1.0 - gamma * (double)jv)
                           : ::ldexpf(gf * jv, e[i] - e[j]);
          for (int k = 0; k < NC; k++)
            if (law[i] == k + 1)
              v = j < NS ? ::ldexpf((float)Net::cons(k, j), rs[i] - e[j])
                         : 0.0f;
          w.Aat(i + j * N, s) = ::fminf(::fmaxf(v, -1e30f), 1e30f);
        }
    } else {
      const MT c = MT(-gamma);
      for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++)
          w.Aat(i + j * N, s) = w.Jat(i + j * N, s) * c;
        w.Aat(j + j * N, s) += MT(1);
      }
    }
    for (int k = 0; k < N; k++) {
      int l_ = k;
      MT best = kx::fabs(w.Aat(k + k * N, s));
      for (int i = k + 1; i < N; i++) {
        const MT v = kx::fabs(w.Aat(i + k * N, s));
        if (v > best) {
          best = v;
          l_ = i;
        }
      }
      w.piv[(long)k * w.S + s] = MIXED ? pack(l_, e[k], law[k], rs[k]) : l_;
      if (w.Aat(l_ + k * N, s) == MT(0)) return k + 1;
      if (l_ != k)
        for (int j = 0; j < N; j++) {
          const MT t = w.Aat(l_ + j * N, s);
          w.Aat(l_ + j * N, s) = w.Aat(k + j * N, s);
          w.Aat(k + j * N, s) = t;
        }
      const MT r = MT(1) / w.Aat(k + k * N, s);
      for (int i = k + 1; i < N; i++) w.Aat(i + k * N, s) *= r;
      for (int j = k + 1; j < N; j++) {
        const MT akj = w.Aat(k + j * N, s);
        if (akj != MT(0))
          for (int i = k + 1; i < N; i++)
            w.Aat(i + j * N, s) -= akj * w.Aat(i + k * N, s);
      }
    }
    return 0;
  }

  // one column step k of the register LU (k a compile-time constant, so
  // every index into a is one: the array stays in registers)
  template <int K>
  static KX_HD void lu_step(float (&a)[N * N], int& fail, const W& w, long s,
                            const int* e, const int* law, const int* rs) {
    int l = K;
    float best = ::fabsf(a[K + K * N]);
#pragma unroll
    for (int i = K + 1; i < N; i++) {
      const float v = ::fabsf(a[i + K * N]);
      if (v > best) {
        best = v;
        l = i;
      }
    }
    w.piv[(long)K * w.S + s] = pack(l, e[K], law[K], rs[K]);
    if (best == 0.0f && fail == 0) fail = K + 1;
    if (l != K) {
#pragma unroll
      for (int j = 0; j < N; j++) {
        const float tk = a[K + j * N];
        float tl = tk;
#pragma unroll
        for (int i = K + 1; i < N; i++) tl = i == l ? a[i + j * N] : tl;
        a[K + j * N] = tl;
#pragma unroll
        for (int i = K + 1; i < N; i++)
          a[i + j * N] = i == l ? tk : a[i + j * N];
      }
    }
    const float r = best != 0.0f ? __builtin_amdgcn_rcpf(a[K + K * N]) : 0.0f;
#pragma unroll
    for (int i = K + 1; i < N; i++) a[i + K * N] *= r;
#pragma unroll
    for (int j = K + 1; j < N; j++) {
      const float akj = a[K + j * N];
#pragma unroll
      for (int i = K + 1; i < N; i++)
        a[i + j * N] = ::fmaf(-akj, a[i + K * N], a[i + j * N]);
    }
  }
  // one column step k of the FP64 register LU (bdf_cell.hpp:402-433: the
  // first largest pivot, the row swap, the column scaled by 1 / pivot,
  // updates skipped for a zero a_kj); a zero pivot ends the factorization
  template <int K>
  static KX_HD void lu64_step(double (&a)[N * N], int& fail, const W& w,
                              long s) {
    if (fail) return;
    int l = K;
    double best = ::fabs(a[K + K * N]);
#pragma unroll
    for (int i = K + 1; i < N; i++) {
      const double v = ::fabs(a[i + K * N]);
      if (v > best) {
        best = v;
        l = i;
      }
    }
    w.piv[(long)K * w.S + s] = l;
    if (best == 0.0) {
      fail = K + 1;
      return;
    }
    if (l != K) {
#pragma unroll
      for (int j = 0; j < N; j++) {
        const double tk = a[K + j * N];
        double tl = tk;
#pragma unroll
        for (int i = K + 1; i < N; i++) tl = i == l ? a[i + j * N] : tl;
        a[K + j * N] = tl;
#pragma unroll
        for (int i = K + 1; i < N; i++)
          a[i + j * N] = i == l ? tk : a[i + j * N];
      }
    }
    const double r = 1.0 / a[K + K * N];
#pragma unroll
    for (int i = K + 1; i < N; i++) a[i + K * N] *= r;
#pragma unroll
    for (int j = K + 1; j < N; j++) {
      const double akj = a[K + j * N];
#pragma unroll
      for (int i = K + 1; i < N; i++)
        a[i + j * N] =
            akj != 0.0 ? a[i + j * N] - akj * a[i + K * N] : a[i + j * N];
    }
  }
  template <int... K>
  static KX_HD void lu64_steps(double (&a)[N * N], int& fail, const W& w,
                               long s, std::integer_sequence<int, K...>) {
    (lu64_step<K>(a, fail, w, s), ...);
  }
  template <int... K>
  static KX_HD void lu_steps(float (&a)[N * N], int& fail, const W& w, long s,
                             const int* e, const int* law, const int* rs,
                             std::integer_sequence<int, K...>) {
    (lu_step<K>(a, fail, w, s, e, law, rs), ...);
  }

  // an FP32 matrix of the workspace into registers and back (pairs when
  // the layout has them)
  static KX_HD void load_mat(const W& w, long s, float* base, float* a) {
    if constexpr (W::PAIRS) {
#pragma unroll
      for (int p = 0; p < N * N / 2; p++) {
        const float2 v = reinterpret_cast<const float2*>(base)[(long)p * W::S + s];
        a[2 * p] = v.x;
        a[2 * p + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int k = 0; k < N * N; k++) a[k] = base[(long)k * W::S + s];
    }
  }
  static KX_HD void store_mat(const W& w, long s, float* base, const float* a) {
    if constexpr (W::PAIRS) {
#pragma unroll
      for (int p = 0; p < N * N / 2; p++)
        reinterpret_cast<float2*>(base)[(long)p * W::S + s] =
            make_float2(a[2 * p], a[2 * p + 1]);
    } else {
#pragma unroll
      for (int k = 0; k < N * N; k++) base[(long)k * W::S + s] = a[k];
    }
  }

  // mixed precision: the scaled FP32 matrix formed, factorized with
  // partial pivoting and stored back from registers (indices fixed at
  // compile time; the pivot row's swap by selects), so the factorization
  // costs one load and one store of the matrix instead of a dependent
  // chain of workspace round trips
  static KX_HD int form_and_factor_regs(const W& w, long s, double gamma) {
    int e[N], law[N], rs[N];
    mixed_scaling(w, s, e, law, rs);
    float a[N * N];
    load_mat(w, s, w.J, a);
    // in FP32: the off-diagonal -gamma J_ij 2^(e_i - e_j) (the scaling
    // exact); the diagonal 1 - gamma J_ii from FP64 (its scaling is 1)
    const float gf = (float)(-gamma);
#pragma unroll
    for (int j = 0; j < N; j++)
#pragma unroll
      for (int i = 0; i < N; i++) {
        float v = i == j ? (float)(1.0 - gamma * (double)a[i + j * N])
                         : ::ldexpf(gf * a[i + j * N], e[i] - e[j]);
#pragma unroll
        for (int k = 0; k < NC; k++)
          if (law[i] == k + 1)
            v = j < NS ? ::ldexpf((float)Net::cons(k, j), rs[i] - e[j]) : 0.0f;
        a[i + j * N] = ::fminf(::fmaxf(v, -1e30f), 1e30f);
      }
    int fail = 0;
    lu_steps(a, fail, w, s, e, law, rs, std::make_integer_sequence<int, N>{});
    store_mat(w, s, w.A, a);
    return fail;
  }

  // delta := A^-1 delta, times 2 / (1 + gamrat) when gamma changed since
  // the setup (bdf_cell.hpp:435-457, 1134-1153).  Mixed: the right-hand
  // side of the factored matrix (rows weighted by 2^e_i, the replaced
  // rows c_k^T b scaled), scaled by its largest magnitude into FP32,
  // solved, and the solution unweighted by 2^-e_i.
  static KX_HD void lsolve(const W& w, long s, double gamrat, double* dl) {
    MT b[N];
    int pv[N], e[N], law[N], rs[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
      const int p = w.piv[(long)k * w.S + s];
      pv[k] = unpack_pv(p);
      e[k] = MIXED ? unpack_e(p) : 0;
      law[k] = MIXED ? unpack_law(p) : 0;
      rs[k] = MIXED ? unpack_rs(p) : 0;
    }
    double sc = 1.0;
    if (MIXED) {
      double bs[N], cb[NC > 0 ? NC : 1], m = 0.0;
#pragma unroll
      for (int k = 0; k < NC; k++) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < NS; j++)
          if (Net::cons(k, j) != 0) t += Net::cons(k, j) * dl[j];
        cb[k] = t;
      }
#pragma unroll
      for (int i = 0; i < N; i++) {
        double v = ::ldexp(dl[i], e[i]);
#pragma unroll
        for (int k = 0; k < NC; k++)
          if (law[i] == k + 1) v = ::ldexp(cb[k], rs[i]);
        bs[i] = v;
        m = ::fmax(m, ::fabs(v));
      }
      sc = m > 0.0 ? m : 1.0;
#pragma unroll
      for (int i = 0; i < N; i++) b[i] = MT(bs[i] / sc);
    } else {
#pragma unroll
      for (int i = 0; i < N; i++) b[i] = MT(dl[i]);
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
      const int pk = pv[k];
      if (pk != k) {
        MT u = MT(0);
        for (int i = 0; i < N; i++) u = i == pk ? b[i] : u;
        const MT t = b[k];
        for (int i = 0; i < N; i++) b[i] = i == pk ? t : b[i];
        b[k] = u;
      }
    }
    if constexpr (MIXED && N <= REG_LU_N) {
      // the factors into registers (pairs), then the two triangular solves
      float am[N * N];
      load_mat(w, s, w.A, am);
#pragma unroll
      for (int k = 0; k < N - 1; k++)
#pragma unroll
        for (int i = k + 1; i < N; i++) b[i] -= am[i + k * N] * b[k];
#pragma unroll
      for (int k = N - 1; k > 0; k--) {
        b[k] *= __builtin_amdgcn_rcpf(am[k + k * N]);
#pragma unroll
        for (int i = 0; i < k; i++) b[i] -= am[i + k * N] * b[k];
      }
      b[0] *= __builtin_amdgcn_rcpf(am[0]);
    } else {
      for (int k = 0; k < N - 1; k++)
        for (int i = k + 1; i < N; i++) b[i] -= w.Aat(i + k * N, s) * b[k];
      for (int k = N - 1; k > 0; k--) {
        b[k] /= w.Aat(k + k * N, s);
        for (int i = 0; i < k; i++) b[i] -= w.Aat(i + k * N, s) * b[k];
      }
      b[0] /= w.Aat(0, s);
    }
    const double c = gamrat != 1.0 ? 2.0 / (1.0 + gamrat) : 1.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
      double x = double(b[i]);
      if (MIXED) x = ::ldexp(sc * x, -e[i]);
      dl[i] = gamrat != 1.0 ? c * x : x;
    }
  }

  // the charge invariant of the corrections (bdf_cell.hpp:1267-1287)
  // (the charges c over the species; the electrons' row r)
  static KX_HD double inv_dot(const double* dl) {
    double sum = 0.0;
    for (int i = 0; i < NS; i++)
      if (Net::charge(i) != 0) sum += Net::charge(i) * dl[i];
    return sum;
  }
  static KX_HD void inv_set(const W& w, long s, double target, double* dl) {
    constexpr int r = Net::I_E;
    double sum = 0.0, sa = 0.0, yr = 0.0, cr = 0.0;
    for (int i = 0; i < NS; i++) {
      if (i != r && Net::charge(i) != 0) {
        sum += Net::charge(i) * dl[i];
        sa += ::fabs(Net::charge(i) * (w.at(0, i, s) + w.at(V_ACOR, i, s)));
      }
      if (i == r) {
        yr = w.at(0, i, s) + w.at(V_ACOR, i, s);
        cr = Net::charge(i);
      }
    }
    const double xr = (target - sum) / cr;
    if (::fabs(cr * yr) >= 1e-6 * sa && yr + xr >= 0.0) dl[r] = xr;
  }
  // cvNlsConvTest (bdf_cell.hpp:1035-1048)
  static KX_HD int conv_test(const W& w, long s, int m, const double* dl) {
    using namespace bdfc;
    {
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < N; i++) {
        const double p = dl[i] * w.at(V_EWT, i, s);
        sum += p * p;
      }
      ST(delnrm) = rt(sum / N);
    }
    if (m > 0) ST(crate) = mx(CRDOWN * ST(crate), ST(delnrm) / ST(delp));
    const double dcon = ST(delnrm) * mn(1.0, ST(crate)) / ST(tol);
    if (dcon <= 1.0) {
      ST(acnrm) = m == 0 ? ST(delnrm) : wrms(w, s, V_ACOR);
      ST(acnrmcur) = true;
      return BDF_OK;
    }
    if (m >= 1 && ST(delnrm) > RDIV * ST(delp)) return NLS_CONV_RECVR;
    ST(delp) = ST(delnrm);
    return NLS_CONTINUE;
  }
  // CVodeGetDky(t, 0) into V_Y (bdf_cell.hpp:1492-1500)
  static KX_HD void dky(const W& w, long s, double t) {
    const double sv = (t - ST(tn)) / ST(h);
    double out[N];
    for (int i = 0; i < N; i++) out[i] = 0.0;
#pragma unroll
    for (int j = QMAX; j >= 0; j--)
      if (j <= ST(q)) {
        double c = 1.0;
#pragma unroll
        for (int i = 0; i < QMAX; i++)
          if (i < j) c *= sv;
        if (j == ST(q))
          for (int i = 0; i < N; i++) out[i] = c * w.at(j, i, s);
        else
          for (int i = 0; i < N; i++) out[i] += c * w.at(j, i, s);
      }
    for (int i = 0; i < N; i++) w.at(V_Y, i, s) = out[i];
  }
  static KX_HD void reinit(const W& w, long s, double t0,
                           double h_init) {
    ST(tn) = t0;
    ST(q) = 1;
    ST(L) = 2;
    ST(qwait) = ST(L);
    ST(etamax) = bdfc::ETA_MAX_FS;
    ST(nst) = ST(nfe) = ST(nfeDQ) = ST(nje) = ST(nsetups) = ST(nni) = ST(nnf) =
        ST(ncfn) = ST(netf) = 0;
    ST(nstlp) = ST(nscon) = ST(nstlj) = 0;
    ST(next_h) = 0.0;
    ST(hin) = h_init;
    ST(acnrmcur) = ST(nls_jcur) = ST(jcur) = false;
    ST(saved_tq5) = 0.0;
#pragma unroll
    for (int i = 0; i <= LMAX; i++) STA(tau, i) = 0.0;
    ST(eta) = 1.0;
    (void)w;
    (void)s;
  }
  static KX_HD void add_stats(const W& w, long s) {
    TOT(0) += ST(nst); TOT(1) += ST(nfe); TOT(2) += ST(nje);
    TOT(3) += ST(nfeDQ); TOT(4) += ST(nsetups); TOT(5) += ST(nni);
    TOT(6) += ST(nnf); TOT(7) += ST(netf);
  }
};

// ---------------------------------------------------------------------------
// the kernel

// the workspace view of one network's kernel
template <class Net, bool ET, bool MIXED>
using PersistWorkOf =
    PersistWork<Problem<Net, ET>::N,
                typename std::conditional<MIXED, float, double>::type>;

// doubles of a block's LDS that one network's lanes use
template <class Net, bool ET>
constexpr long lds_doubles() {
  constexpr int N = Problem<Net, ET>::N;
  return (long)lds_vectors<N>() * N * KX_PERSIST_BLOCK;
}

// a network's work queue: a counter over the permutation order of its
// cells; chunk > 0 makes a wave take chunk cells per atomic into a
// reservoir (base, left) of its own (two ints of res per wave, in the
// workspace's head), which it hands out to its idle lanes before it takes
// the next chunk (chunk 0: exactly the cells its idle lanes need, one
// atomic per fetch)
struct PersistQueue {
  int* counter;
  const int* order;
  int chunk;
  int* res;
};

// one network's cells on this wave: its lanes take cells from the
// network's queue until the queue is exhausted and every lane of the wave
// is idle.  lds_vec: the block's LDS, each lane using only its own column
// (tid) of it, so waves of one block may run different networks at once.
template <class Net, bool ET, bool MIXED>
__device__ __attribute__((always_inline)) inline void persist_run(
    CellsArgs a, PersistWorkOf<Net, ET, MIXED> w0, PersistQueue qu,
    double* lds_vec) {
  using M = Persist<Net, ET, MIXED>;
  using Prob = Problem<Net, ET>;
  constexpr int N = M::N;
  constexpr int NS = Net::NSPEC;
  using namespace bdfc;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  // this wave's reservoir, empty at the start (only the wave's leader
  // lanes touch its two ints, by atomics)
  int* res = qu.res + 2 * (s / 64);
  const int chunk = qu.chunk >= 64 ? qu.chunk : 0;
  if (chunk > 0 && lane == 0) {
    __hip_atomic_store(&res[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&res[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }

  const SolveOpts& o = a.o;
  const bool dq_jac = o.jac == 2 || (o.jac == 0 && ET);
  // the analytic Jacobian here is the mass-action block at fixed Tgas
  const bool supported = dq_jac || !(ET || Net::RATES_DENSITY);
  const bool inv = dq_jac && Net::HAS_E;  // the charge invariant
  bool atolmin0 =
      (ET ? (o.atol_T < o.atol_n ? o.atol_T : o.atol_n) : o.atol_n) == 0.0;
  if (o.tol != nullptr) {
    double m = o.tol[0];
    for (int i = 1; i < N; i++) m = o.tol[i] < m ? o.tol[i] : m;
    atolmin0 = m == 0.0;
  }
  const long mxstep = o.max_steps == 0 ? 500 : o.max_steps;
  LaneState lane_state;

  int pc = P_FETCH;
  Prob P;
  P.hit_T0 = false;
  // the rhs request of this lane: source vector, destination vector,
  // precision (0: FP64, 1: FP32), and its return code
  int rq_dst = 0, rr = 0;
  // the state the rhs is requested at, handed to the rhs in registers;
  // the last Newton-residual rhs f(y), kept in registers across ticks
  // until the next one (the residual and the difference quotients read it)
  double yq[N], fq[N];
  bool want = false;

  for (;;) {
    // ---- advance: run the machine to its next rhs request --------------
    want = false;
    while (!want && pc != P_IDLE) {
      // the workspace's pointers and stride made opaque at every step of
      // the machine: the addresses of the unrolled vector elements are then
      // formed where they are used instead of being held across the loop
      auto w = w0;
      {
        unsigned long long pv = (unsigned long long)w0.v,
                           pj = (unsigned long long)w0.J,
                           pa = (unsigned long long)w0.A,
                           ps = (unsigned long long)w0.tot,
                           pp = (unsigned long long)w0.piv;
        asm volatile("" : "+s"(pv), "+s"(pj), "+s"(pa), "+s"(ps), "+s"(pp));
        w.v = (double*)pv;
        w.J = (decltype(w.J))pj;
        w.A = (decltype(w.A))pa;
        w.tot = (long*)ps;
        w.piv = (int*)pp;
      }
      w.sp = &lane_state;
      w.lds = lds_vec;
      w.tid = threadIdx.x;
      // the machine's stages in the order a step flows through them: a
      // lane runs every stage its transitions reach in one pass (each
      // stage once per wave and pass, masked to the lanes there); the
      // rare backward transitions take another pass
      if (pc == P_FIRST_RHS) do {
          ST(nfe)++;
          if (rr != 0) {
            ST(flag) = rr < 0 ? RHSFUNC_FAIL : FIRST_RHSFUNC_ERR;
            pc = P_SOLVE_END;
            break;
          }
          ST(h) = ST(hin);
          if (ST(h) != 0.0 && (ST(tout) - ST(tn)) * ST(h) < 0.0) {
            ST(flag) = ILL_INPUT;
            pc = P_SOLVE_END;
            break;
          }
          pc = ST(h) == 0.0 ? P_HIN_BEGIN : P_AFTER_HIN;
          break;
      } while (0);
      if (pc == P_HIN_RET) do {
          ST(nfe)++;
          if (rr < 0) {
            ST(flag) = RHSFUNC_FAIL;
            pc = P_SOLVE_END;
            break;
          }
          if (rr > 0) {
            ST(hin_hg) *= 0.2;
            ST(count2)++;
            if (ST(count2) <= MAX_ITERS) {
              pc = P_HIN_REQ;
              break;
            }
            if (ST(count1) <= 2) {
              ST(flag) = REPTD_RHSFUNC_ERR;
              pc = P_SOLVE_END;
              break;
            }
            ST(hin_hnew) = ST(hin_hs);
            pc = P_HIN_END;
            break;
          }
          const double hg = ST(hin_hg);
          M::lsum(w, s, 1.0 / hg, V_TEMPV, -1.0 / hg, 1, V_TEMPV);
          const double ydd = M::wrms(w, s, V_TEMPV);
          const double hub = ST(hin_hub);
          ST(hin_hs) = hg;
          ST(hin_hnew) = (ydd * hub * hub > 2.0) ? rt(2.0 / ydd) : rt(hg * hub);
          if (ST(count1) == MAX_ITERS) {
            pc = P_HIN_END;
            break;
          }
          const double hrat = ST(hin_hnew) / hg;
          if (hrat > 0.5 && hrat < 2.0) {
            pc = P_HIN_END;
            break;
          }
          if (ST(count1) > 1 && hrat > 2.0) {
            ST(hin_hnew) = hg;
            pc = P_HIN_END;
            break;
          }
          ST(hin_hg) = ST(hin_hnew);
          ST(count1)++;
          ST(count2) = 1;
          pc = P_HIN_REQ;
          break;
      } while (0);
      if (pc == P_ERRTEST_RHS) do {
          ST(nfe)++;
          if (rr != 0) {
            ST(flag) = rr < 0 ? RHSFUNC_FAIL : UNREC_RHSFUNC_ERR;
            pc = P_FAIL_END;
            break;
          }
          M::scal(w, s, ST(h), V_TEMPV, 1);
          pc = P_ATTEMPT;
          break;
      } while (0);
      if (pc == P_DQ_RET) do {
          ST(nfeDQ)++;
          const int j = ST(jcol);
          if (rr != 0) {
            ST(jr) = rr;
            pc = P_LSETUP_FORM;
            break;
          }
          // lin_sum(ii, vt1, -ii, ftemp): the a == -b form.  The
          // differences are FP64 in both precisions; mixed precision stores
          // them in FP32 (FP32 differences, with their sqrt(FLT_EPSILON)
          // steps, blur the slow chemistry into the fast and cost the
          // accuracy the FP64 residual cannot buy back)
          const double ii = 1.0 / ST(inc);
          for (int i = 0; i < N; i++)
            w.Jat(i + j * N, s) = typename M::MT(
                ii * (w.at(V_VT1, i, s) - fq[i]));
          ST(jcol)++;
          if (ST(jcol) < N) {
            pc = P_DQ_REQ;
          } else {
            ST(jr) = 0;
            pc = P_LSETUP_FORM;
          }
          break;
      } while (0);
      if (pc == P_RESID_RET) do {
          ST(nfe)++;
          const int r = rr < 0 ? RHSFUNC_FAIL : (rr > 0 ? RHSFUNC_RECVR : BDF_OK);
          // the residual itself is formed where it is used (P_INNER_ITER),
          // from ftemp, zn[1] and acor, which stay as they are until then
          if (ST(first_resid)) {
            if (r != BDF_OK) {
              ST(nls_r) = r;
              pc = P_NLS_EXIT;
            } else {
              pc = ST(callSetup) ? P_LSETUP_BEGIN : P_INNER_ITER;
            }
          } else {
            if (r != BDF_OK) {
              ST(nls_r) = r;
              pc = P_INNER_EXIT;
            } else {
              pc = P_INNER_ITER;
            }
          }
          break;
      } while (0);
      if (pc == P_HIN_BEGIN) do {  // cvHin (bdf_cell.hpp:717-764)
          const double tdist = ::fabs(ST(tout) - ST(tn));
          const double tround = DBL_EPSILON * mx(::fabs(ST(tn)), ::fabs(ST(tout)));
          ST(hin_hlb) = HLB_FACTOR * tround;
          // cvUpperBoundH0 with temp1 = V_TEMPV, temp2 = V_ACOR
          for (int i = 0; i < N; i++) w.at(V_ACOR, i, s) = ::fabs(w.at(0, i, s));
          M::ewt_set(w, s, o, atolmin0, 0, V_TEMPV);
          for (int i = 0; i < N; i++)
            w.at(V_TEMPV, i, s) = 1.0 / w.at(V_TEMPV, i, s);
          M::lsum(w, s, HUB_FACTOR, V_ACOR, 1.0, V_TEMPV, V_TEMPV);
          for (int i = 0; i < N; i++) w.at(V_ACOR, i, s) = ::fabs(w.at(1, i, s));
          double inv = 0.0;
          for (int i = 0; i < N; i++) {
            const double v = w.at(V_ACOR, i, s) / w.at(V_TEMPV, i, s);
            w.at(V_TEMPV, i, s) = v;
            if (::fabs(v) > inv) inv = ::fabs(v);
          }
          double hub = HUB_FACTOR * tdist;
          if (hub * inv > 1.0) hub = 1.0 / inv;
          ST(hin_hub) = hub;
          ST(hin_hg) = rt(ST(hin_hlb) * hub);
          if (hub < ST(hin_hlb)) {
            ST(h) = ST(hin_hg);
            pc = P_AFTER_HIN;
            break;
          }
          ST(hin_hs) = ST(hin_hg);
          ST(hin_hnew) = ST(hin_hg);
          ST(count1) = 1;
          ST(count2) = 1;
          pc = P_HIN_REQ;
          break;
      } while (0);
      if (pc == P_HIN_END) do {
          double h0 = H_BIAS * ST(hin_hnew);
          if (h0 < ST(hin_hlb)) h0 = ST(hin_hlb);
          if (h0 > ST(hin_hub)) h0 = ST(hin_hub);
          ST(h) = h0;
          pc = P_AFTER_HIN;
          break;
      } while (0);
      if (pc == P_HIN_REQ) do {
          // y = hg zn[1] + zn[0] (N_VLinearSum with b = 1)
          {
            const double a = ST(hin_hg);
#pragma unroll
            for (int i = 0; i < N; i++) {
              const double x = w.at(1, i, s), y = w.at(0, i, s);
              yq[i] = a == 1.0 ? x + y : (a == -1.0 ? y - x : a * x + y);
            }
          }
          rq_dst = V_TEMPV; want = true;
          pc = P_HIN_RET;
          break;
      } while (0);
      if (pc == P_AFTER_HIN) do {
          ST(hscale) = ST(h);
          ST(hprime) = ST(h);
          M::scal(w, s, ST(h), 1, 1);
          ST(nstloc) = 0;
          pc = P_LOOP_TOP;
          break;
      } while (0);
      if (pc == P_LSETUP_BEGIN) do {  // cvNlsLSetup + cvLsSetup
          if (ST(jbad)) ST(convfail) = FAIL_BAD_J;
          const double dg = ::fabs((ST(gamma) / ST(gammap)) - 1.0);
          const bool jbad_ls = ST(nst) == 0 || ST(nst) >= ST(nstlj) + MSBJ ||
                               (ST(convfail) == FAIL_BAD_J && dg < DGMAX_JBAD) ||
                               ST(convfail) == FAIL_OTHER;
          ST(jr) = 0;
          if (!jbad_ls) {
            ST(jcur) = false;
            pc = P_LSETUP_FORM;
            break;
          }
          ST(jcur) = true;
          if (!dq_jac) {
            // the analytic Jacobian at y: the mass-action block at fixed
            // Tgas (with Tgas evolved or density-dependent rates the cells
            // are flagged unsupported before they get here; the host sends
            // those options to the lane queue)
            double yr[N], Jr[N * N];
            for (int i = 0; i < N; i++) yr[i] = w.at(V_Y, i, s);
            for (int k = 0; k < N * N; k++) Jr[k] = 0.0;
            ST(jr) = P.jac_block(yr, Jr);
            for (int k = 0; k < N * N; k++) w.Jat(k, s) = typename M::MT(Jr[k]);
            pc = P_LSETUP_FORM;
            break;
          }
          pc = P_DQ_BEGIN;
          break;
      } while (0);
      if (pc == P_DQ_BEGIN) do {  // cvLsDenseDQJac (bdf_cell.hpp:1290-1311)
          double fn = 0.0;
          {
            double sum = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
              const double p = fq[i] * w.at(V_EWT, i, s);
              sum += p * p;
            }
            fn = rt(sum / N);
          }
          ST(minInc) = fn != 0.0 ? (MIN_INC_MULT * ::fabs(ST(h)) *
                                   DBL_EPSILON * (double)N * fn)
                                : 1.0;
          ST(jcol) = 0;
          pc = P_DQ_REQ;
          break;
      } while (0);
      if (pc == P_DQ_REQ) do {
          // the perturbed state in registers only (y itself unchanged)
          const int j = ST(jcol);
          const double srur = rt(DBL_EPSILON);
          double ys = 0.0, ew = 0.0;
#pragma unroll
          for (int i = 0; i < N; i++) {
            yq[i] = w.at(V_Y, i, s);
            if (i == j) {
              ys = yq[i];
              ew = w.at(V_EWT, i, s);
            }
          }
          ST(inc) = mx(srur * ::fabs(ys), ST(minInc) / ew);
#pragma unroll
          for (int i = 0; i < N; i++)
            if (i == j) yq[i] += ST(inc);
          rq_dst = V_VT1; want = true;
          pc = P_DQ_RET;
          break;
      } while (0);
      if (pc == P_LSETUP_FORM) do {
          int ret = ST(jr) < 0 ? -1 : (ST(jr) > 0 ? 1 : 0);
          if (ST(jcur)) {
            ST(nje)++;
            ST(nstlj) = ST(nst);
          }
          if (ret == 0) ret = M::form_and_factor(w, s, ST(gamma)) > 0 ? 1 : 0;
          ST(nsetups)++;
          ST(nls_jcur) = ST(jcur);
          ST(gamrat) = 1.0;
          ST(gammap) = ST(gamma);
          ST(crate) = 1.0;
          ST(delnrm) = 0.0;
          ST(nstlp) = ST(nst);
          if (ret != 0) {
            ST(nls_r) = ret < 0 ? LSETUP_FAIL : NLS_CONV_RECVR;
            pc = P_NLS_EXIT;
          } else {
            pc = P_INNER_ITER;
          }
          break;
      } while (0);
      if (pc == P_INNER_ITER) do {
          ST(niters)++;
          // the Newton correction in registers: residual, solve, charge
          // invariant, accumulation and convergence test in one stage
          double dl[N];
          M::residual(w, s, ST(rl1), ST(gamma), fq, dl);
          ST(cg) = inv ? M::inv_dot(dl) : 0.0;
#pragma unroll
          for (int i = 0; i < N; i++) dl[i] = -1.0 * dl[i];
          M::lsolve(w, s, ST(gamrat), dl);
          if (inv) M::inv_set(w, s, -ST(cg), dl);
#pragma unroll
          for (int i = 0; i < N; i++) w.at(V_ACOR, i, s) += dl[i];
          const int r = M::conv_test(w, s, ST(curiter), dl);
          KX_TRACE_REC2(5.0, ST(delnrm), ST(crate));
          ST(curiter)++;
          if (r == BDF_OK) {
            ST(nls_jcur) = false;
            ST(nls_r) = BDF_OK;
            pc = P_NLS_EXIT;
          } else if (r != NLS_CONTINUE) {
            ST(nls_r) = r;
            pc = P_INNER_EXIT;
          } else if (ST(curiter) >= NLS_MAXCOR) {
            ST(nls_r) = NLS_CONV_RECVR;
            pc = P_INNER_EXIT;
          } else {
            ST(first_resid) = 0;
            pc = P_RESID_REQ;
          }
          break;
      } while (0);
      if (pc == P_INNER_EXIT) do {
          if (ST(nls_r) > 0 && !ST(nls_jcur)) {
            ST(nconvfails)++;
            ST(callSetup) = 1;
            ST(jbad) = 1;
            M::zero(w, s, V_ACOR);
            pc = P_NLS_OUTER;
          } else {
            pc = P_NLS_EXIT;
          }
          break;
      } while (0);
      if (pc == P_NLS_EXIT) do {
          const int r = ST(nls_r);
          if (r != BDF_OK) ST(nconvfails)++;
          ST(nni) += ST(niters);
          ST(nnf) += ST(nconvfails);
          if (r != BDF_OK) {
            ST(nflag) = r;
          } else {
            M::lsum(w, s, 1.0, 0, 1.0, V_ACOR, V_Y);
            if (!ST(acnrmcur)) ST(acnrm) = M::wrms(w, s, V_ACOR);
            ST(jcur) = false;
            ST(nflag) = BDF_OK;
          }
          pc = P_HANDLE;
          break;
      } while (0);
      if (pc == P_HANDLE) do {  // cvHandleNFlag, cvDoErrorTest
          const int f = ST(nflag);
          if (f != BDF_OK) {
            KX_TRACE_REC(1.0, (double)f);
            ST(ncfn)++;
            M::restore(w, s);
            if (f < 0) {
              ST(flag) = (f == LSETUP_FAIL || f == LSOLVE_FAIL ||
                         f == RHSFUNC_FAIL) ? f : NLS_FAIL;
              pc = P_FAIL_END;
              break;
            }
            ST(ncf)++;
            ST(etamax) = 1.0;
            if (::fabs(ST(h)) <= 0.0 || ST(ncf) == o.max_conv_fails) {
              if (f == NLS_CONV_RECVR) {
                ST(flag) = CONV_FAILURE;
                pc = P_FAIL_END;
                break;
              }
              if (f == RHSFUNC_RECVR) {
                ST(flag) = REPTD_RHSFUNC_ERR;
                pc = P_FAIL_END;
                break;
              }
            }
            ST(eta) = mx(ETA_CF, 0.0);
            ST(nflag) = PREV_CONV_FAIL;
            M::rescale(w, s);
            pc = P_ATTEMPT;
            break;
          }
          const double dsm = ST(acnrm) * STA(tq, 2);
          ST(dsm) = dsm;
          KX_TRACE_REC(2.0, dsm);
          if (dsm <= 1.0) {
            pc = P_STEP_DONE;
            break;
          }
          ST(nef)++;
          ST(netf)++;
          ST(nflag) = PREV_ERR_FAIL;
          M::restore(w, s);
          if (::fabs(ST(h)) <= 0.0 || ST(nef) == o.max_err_fails) {
            ST(flag) = ERR_FAILURE;
            pc = P_FAIL_END;
            break;
          }
          ST(etamax) = 1.0;
          if (ST(nef) <= MXNEF1) {
            ST(eta) = 1.0 / (kx::pow(BIAS2 * dsm, 1.0 / ST(L)) + ADDON);
            ST(eta) = mx(ETA_MIN_EF, mx(ST(eta), 0.0));
            if (ST(nef) >= SMALL_NEF) ST(eta) = mn(ST(eta), ETA_MAX_EF);
            M::rescale(w, s);
            pc = P_ATTEMPT;
            break;
          }
          if (ST(q) > 1) {
            ST(eta) = mx(ETA_MIN_EF, 0.0);
            M::order_change(w, s, -1);
            ST(L) = ST(q);
            ST(q)--;
            ST(qwait) = ST(L);
            M::rescale(w, s);
            pc = P_ATTEMPT;
            break;
          }
          ST(eta) = mx(ETA_MIN_EF, 0.0);
          ST(h) *= ST(eta);
          ST(next_h) = ST(h);
          ST(hscale) = ST(h);
          ST(qwait) = LONG_WAIT;
          ST(nscon) = 0;
#pragma unroll
          for (int i = 0; i < N; i++) yq[i] = w.at(0, i, s);
          rq_dst = V_TEMPV; want = true;
          pc = P_ERRTEST_RHS;
          break;
      } while (0);
      if (pc == P_STEP_DONE) do {
          M::complete(w, s);
          M::next_step(w, s);
          KX_TRACE_REC(3.0, ST(eta));
          KX_TRACE_REC2(6.0, w.at(0, N - 1, s), w.at(0, 3, s));
          ST(etamax) = ST(nst) <= SMALL_NST ? ETA_MAX_ES : ETA_MAX_GS;
          M::scal(w, s, STA(tq, 2), V_ACOR, V_ACOR);
          ST(nstloc)++;
          if ((ST(tn) - ST(tout)) * ST(h) >= 0.0) {
            ST(tret) = ST(tout);
            M::dky(w, s, ST(tout));
            ST(next_h) = ST(hprime);
            ST(flag) = BDF_OK;
            pc = P_SOLVE_END;
          } else {
            pc = P_LOOP_TOP;
          }
          break;
      } while (0);
      if (pc == P_LOOP_TOP) do {
          ST(next_h) = ST(h);
          if (ST(nst) > 0 && M::ewt_set(w, s, o, atolmin0, 0, V_EWT) != 0) {
            ST(flag) = ILL_INPUT;
            pc = P_FAIL_END;
            break;
          }
          if (mxstep > 0 && ST(nstloc) >= mxstep) {
            ST(flag) = TOO_MUCH_WORK;
            pc = P_FAIL_END;
            break;
          }
          if (DBL_EPSILON * M::wrms(w, s, 0) > 1.0) {
            ST(flag) = TOO_MUCH_ACC;
            pc = P_FAIL_END;
            break;
          }
          pc = P_STEP_BEGIN;
          break;
      } while (0);
      if (pc == P_STEP_BEGIN) do {
          ST(ncf) = 0;
          ST(nef) = 0;
          ST(dsm) = 0.0;
          if (ST(nst) > 0 && ST(hprime) != ST(h)) M::adjust(w, s);
          ST(saved_t) = ST(tn);
          ST(nflag) = FIRST_CALL;
          pc = P_ATTEMPT;
          break;
      } while (0);
      if (pc == P_ATTEMPT) do {
          M::predict(w, s);
          M::set_coeffs(w, s);
          pc = P_NLS_BEGIN;
          break;
      } while (0);
      if (pc == P_NLS_BEGIN) do {
          const int nf = ST(nflag);
          ST(convfail) = (nf == FIRST_CALL || nf == PREV_ERR_FAIL) ? NO_FAILURES
                                                                  : FAIL_OTHER;
          ST(callSetup) = nf == PREV_CONV_FAIL || nf == PREV_ERR_FAIL ||
                         ST(nst) == 0 || ST(nst) >= ST(nstlp) + MSBP ||
                         ::fabs(ST(gamrat) - 1.0) > DGMAX_LSETUP;
          M::zero(w, s, V_ACOR);
          ST(tol) = STA(tq, 4);
          ST(jbad) = 0;
          ST(niters) = 0;
          ST(nconvfails) = 0;
          pc = P_NLS_OUTER;
          break;
      } while (0);
      if (pc == P_NLS_OUTER) do {
          ST(curiter) = 0;
          ST(first_resid) = 1;
          pc = P_RESID_REQ;
          break;
      } while (0);
      if (pc == P_RESID_REQ) do {
          // y = zn[0] + acor, kept for a Jacobian and handed to the rhs
#pragma unroll
          for (int i = 0; i < N; i++) {
            yq[i] = w.at(0, i, s) + w.at(V_ACOR, i, s);
            w.at(V_Y, i, s) = yq[i];
          }
          rq_dst = RQ_FTEMP; want = true;
          pc = P_RESID_RET;
          break;
      } while (0);
      if (pc == P_FAIL_END) do {
          ST(tret) = ST(tn);
          for (int i = 0; i < N; i++) w.at(V_Y, i, s) = w.at(0, i, s);
          pc = P_SOLVE_END;
          break;
      } while (0);
      if (pc == P_SOLVE_END) do {  // bdf::integrate's restarts
          M::add_stats(w, s);
          const int f = ST(flag);
          KX_TRACE_REC(4.0, (double)f);
          if (f >= 0) {
            ST(flag) = 0;
            pc = P_CELL_DONE;
            break;
          }
          const bool again = f == ERR_FAILURE || f == CONV_FAILURE ||
                             f == REPTD_RHSFUNC_ERR;
          if (f == TOO_CLOSE || f == ILL_INPUT || f == RHSFUNC_FAIL ||
              f == FIRST_RHSFUNC_ERR || !again ||
              ST(restarts) >= o.max_restarts) {
            pc = P_CELL_DONE;
            break;
          }
          ST(restarts)++;
          const double hc = ST(next_h);
          // the reached state (zn[0], = y on a failure) is the new start
          if (ST(tret) != ST(tn))
            for (int i = 0; i < N; i++) w.at(0, i, s) = w.at(V_Y, i, s);
          M::reinit(w, s, ST(tret), hc > 0.0 ? 0.1 * hc : 0.0);
          pc = P_SOLVE_BEGIN;
          break;
      } while (0);
      if (pc == P_CELL_DONE) do {
          const int c = ST(cell);
          int flag = ST(flag);
          double y[N];
          for (int i = 0; i < N; i++) y[i] = w.at(V_Y, i, s);
          if (flag == 0 && ET &&
              !(y[N - 1] > 0.0 && kx::finite(y[N - 1])))
            flag = FLAG_NONPOSITIVE_T;
          if (P.hit_T0 && (flag == REPTD_RHSFUNC_ERR || flag == RHSFUNC_FAIL ||
                           flag == UNREC_RHSFUNC_ERR ||
                           flag == FIRST_RHSFUNC_ERR))
            flag = FLAG_NONPOSITIVE_T;
          a.flags[c] = flag;
          if (a.stats) {
            long* sc = a.stats + (long)c * 9;
            for (int i = 0; i < 8; ++i) sc[i] = TOT(i);
            sc[8] = ST(restarts);
          }
          if (flag == 0) cell_finish<Net, ET>(a, c, y);
          pc = P_FETCH;
          break;
      } while (0);
      if (pc == P_FETCH) do {
          // one atomic per wave for the lanes that need a cell (or, with
          // a chunk, the cells of the wave's reservoir first)
          const unsigned long long need = __ballot(1);
          const int rank = __popcll(need & ((1ull << lane) - 1ull));
          const int cnt = __popcll(need);
          const int leader = __ffsll((long long)need) - 1;
          int idx = 0;
          if (chunk == 0) {
            int base = 0;
            if (lane == leader) base = atomicAdd(qu.counter, cnt);
            idx = __shfl(base, leader) + rank;
          } else {
            // lanes rank < left take the reservoir's cells; the others
            // the start of a new chunk (cnt - left <= 64 <= chunk)
            int rb = 0, rl = 0, nb = 0;
            if (lane == leader) {
              rb = __hip_atomic_load(&res[0], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
              rl = __hip_atomic_load(&res[1], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
              if (rl < cnt) nb = atomicAdd(qu.counter, chunk);
              __hip_atomic_store(&res[0], rl < cnt ? nb + (cnt - rl) : rb + cnt,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              __hip_atomic_store(&res[1],
                                 rl < cnt ? chunk - (cnt - rl) : rl - cnt,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            rb = __shfl(rb, leader);
            rl = __shfl(rl, leader);
            nb = __shfl(nb, leader);
            idx = rank < rl ? rb + rank : nb + (rank - rl);
          }
          if (idx >= a.ncell) {
            pc = P_IDLE;
            break;
          }
          const int c = qu.order[idx];
          ST(cell) = c;
          for (int i = 0; i < TOT_SLOTS; ++i) TOT(i) = 0;
          ST(restarts) = 0;
          if (!(a.dt[c] > 0.0) || !supported) {
            a.flags[c] = supported ? 0 : FLAG_UNSUPPORTED;
            if (a.stats)
              for (int i = 0; i < 9; ++i) a.stats[(long)c * 9 + i] = 0;
            pc = P_FETCH;
            break;
          }
          P.p = Net::params_from_row(a.par + (long)c * a.par_stride);
          P.Tfix = a.T[c];
          P.hit_T0 = false;
          for (int i = 0; i < NS; ++i)
            w.at(0, i, s) = a.n[c * a.n_cs + i * a.n_ss];
          if (ET) w.at(0, N - 1, s) = a.T[c];
          ST(tout) = a.dt[c];
          M::reinit(w, s, 0.0, 0.0);
          pc = P_SOLVE_BEGIN;
          break;
      } while (0);
      if (pc == P_SOLVE_BEGIN) do {
          ST(tret) = ST(tn);
          const double tdist = ::fabs(ST(tout) - ST(tn));
          const double tround = DBL_EPSILON * mx(::fabs(ST(tn)), ::fabs(ST(tout)));
          if (tdist == 0.0 || tdist < 2.0 * tround) {
            ST(flag) = TOO_CLOSE;
            pc = P_SOLVE_END;
            break;
          }
          if (M::ewt_set(w, s, o, atolmin0, 0, V_EWT) != 0) {
            ST(flag) = ILL_INPUT;
            pc = P_SOLVE_END;
            break;
          }
#pragma unroll
          for (int i = 0; i < N; i++) yq[i] = w.at(0, i, s);
          rq_dst = 1; want = true;
          pc = P_FIRST_RHS;
          break;
      } while (0);
    }
    // ---- the rhs of every lane that asked, at one program point ---------
    const unsigned long long busy = __ballot(pc != P_IDLE);
    if (busy == 0ull) break;
    if (want) {
      auto w = w0;
      {
        unsigned long long pv = (unsigned long long)w0.v;
        asm volatile("" : "+s"(pv));
        w.v = (double*)pv;
      }
      w.lds = lds_vec;
      w.tid = threadIdx.x;
      double fr[N];
      rr = P.template rhs_r<double>(yq, fr);
      if (rq_dst == RQ_FTEMP) {
#pragma unroll
        for (int i = 0; i < N; i++) fq[i] = fr[i];
      } else {
        for (int i = 0; i < N; i++) w.at(rq_dst, i, s) = fr[i];
      }
    }
  }
}

template <class Net, bool ET, bool MIXED>
__global__ void __launch_bounds__(KX_PERSIST_BLOCK, KX_PERSIST_WAVES_PER_EU)
persist_cells_kernel(CellsArgs a, PersistWorkOf<Net, ET, MIXED> w0,
                     PersistQueue q) {
  __shared__ double lds_vec[lds_doubles<Net, ET>()];
  persist_run<Net, ET, MIXED>(a, w0, q, lds_vec);
}

// ---- the queue's order: the densest cells first ---------------------------
// A cell's work grows with its density (its reaction rates, and so its
// stiffness, scale with it; for primordial3 collapse cells the correlation
// of log(steps) with log(nH) is 0.86).  Handing the expensive cells out
// first leaves the cheap ones for the end of the launch, where a late
// expensive cell would keep its wave, and the launch, running alone.  A
// counting sort over half-octave buckets of the total number density, and
// within those over quarter-octave buckets of the temperature, so that the
// lanes of a wave start on cells in the same branches of the piecewise
// rate and cooling fits.
constexpr int ORDER_NB = 256, ORDER_TB = 64;
constexpr int ORDER_BUCKETS = ORDER_NB * ORDER_TB;

template <class Net>
__global__ void order_keys_kernel(CellsArgs a, int* key, int* hist) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  double nt = 0.0;
  for (int i = 0; i < Net::NSPEC; i++)
    nt += ::fabs(a.n[(long)c * a.n_cs + i * a.n_ss]);
  int b = (int)::floor(2.0 * ::log2(::fmax(nt, 1e-300))) + 128;
  b = b < 0 ? 0 : (b > ORDER_NB - 1 ? ORDER_NB - 1 : b);
  b = ORDER_NB - 1 - b;  // densest first
  int t = (int)::floor(4.0 * ::log2(::fmax(a.T[c], 1.0)));
  t = t < 0 ? 0 : (t > ORDER_TB - 1 ? ORDER_TB - 1 : t);
  b = b * ORDER_TB + t;
  key[c] = b;
  atomicAdd(&hist[b], 1);
}

// exclusive prefix sums of the bucket counts (one thread)
__global__ inline void order_scan_kernel(int* hist) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int sum = 0;
  for (int b = 0; b < ORDER_BUCKETS; b++) {
    const int v = hist[b];
    hist[b] = sum;
    sum += v;
  }
}

__global__ inline void order_scatter_kernel(int ncell, const int* key,
                                            int* off, int* order) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  order[atomicAdd(&off[key[c]], 1)] = c;
}

// bytes of the persistent kernel's workspace per lane
template <int N, bool MIXED>
constexpr long persist_lane_bytes() {
  return (long)sizeof(double) * ((V_COUNT - lds_vectors<N>()) * N +
                                 TOT_SLOTS) +
         2L * N * N * (MIXED ? sizeof(float) : sizeof(double)) +
         (long)sizeof(int) * N;
}
// the workspace's head: the queue counter, then two ints per wave of the
// lane slots (the queue's reservoirs)
constexpr size_t PERSIST_HEAD = 256 + 2 * sizeof(int) * (KX_PERSIST_LANES / 64);
// bytes of one network's persistent workspace: the head, then the lane
// slots' vectors, matrices, totals and pivots
template <class Net, bool ET, bool MIXED>
constexpr size_t persist_work_bytes() {
  return (size_t)persist_lane_bytes<Problem<Net, ET>::N, MIXED>() *
             KX_PERSIST_LANES + PERSIST_HEAD;
}
// bytes of the queue's order of ncell cells: the bucket counts, the keys
// and the permutation
inline size_t persist_order_bytes(long ncell) {
  return sizeof(int) * (2 * (size_t)ncell + ORDER_BUCKETS);
}

// the device memory of one network's persistent launch, given by the
// caller: the workspace (persist_work_bytes) and the order
// (persist_order_bytes of the call's cells)
struct PersistMem {
  void* work = nullptr;
  void* ord = nullptr;
};

// resident lanes of a launch over ncell cells: waves_per_cu waves on each
// CU (0: automatic), no more than the cells rounded up to whole workgroups
// or the slots
inline int persist_lanes(long ncell, int wpc, long& lanes) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess)
    return -1;
  if (wpc <= 0) wpc = 4 * KX_PERSIST_WAVES_PER_EU;
  lanes = (long)ncu * wpc * 64;
  constexpr long B = KX_PERSIST_BLOCK;
  const long need = (ncell + B - 1) / B * B;
  if (lanes > need) lanes = need;
  lanes = (lanes + B - 1) / B * B;
  if (lanes > KX_PERSIST_LANES) lanes = KX_PERSIST_LANES;
  return 0;
}

// one network's workspace view and queue in the memory m, on stream s:
// the queue counter zeroed (by a kernel: no memset node when the caller
// captures the stream) and the permutation the lanes take cells in
template <class Net, bool ET, bool MIXED>
int persist_prepare(CellsArgs a, PersistMem m, int chunk, hipStream_t s,
                    PersistWorkOf<Net, ET, MIXED>& w, PersistQueue& q) {
  constexpr int N = Problem<Net, ET>::N;
  using MT = typename std::conditional<MIXED, float, double>::type;
  const long slots = KX_PERSIST_LANES;
  char* p = static_cast<char*>(m.work);
  int* queue = reinterpret_cast<int*>(p);
  int* res = reinterpret_cast<int*>(p + 256);
  p += PERSIST_HEAD;
  w.v = reinterpret_cast<double*>(p);
  p += sizeof(double) * (V_COUNT - lds_vectors<N>()) * N * slots;
  w.J = reinterpret_cast<MT*>(p);
  p += sizeof(MT) * N * N * slots;
  w.A = reinterpret_cast<MT*>(p);
  p += sizeof(MT) * N * N * slots;
  w.tot = reinterpret_cast<long*>(p);
  p += sizeof(long) * TOT_SLOTS * slots;
  w.piv = reinterpret_cast<int*>(p);
  int* hist = static_cast<int*>(m.ord);
  int* key = hist + ORDER_BUCKETS;
  int* order = key + a.ncell;
  const int cb = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(fill_flags_kernel, dim3(1), dim3(64), 0, s, 1, queue, 0);
  hipLaunchKernelGGL(fill_flags_kernel, dim3(ORDER_BUCKETS / 256), dim3(256),
                     0, s, ORDER_BUCKETS, hist, 0);
  hipLaunchKernelGGL(order_keys_kernel<Net>, dim3(cb), dim3(256), 0, s, a,
                     key, hist);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(64), 0, s, hist);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(cb), dim3(256), 0, s,
                     a.ncell, key, hist, order);
  q.counter = queue;
  q.order = order;
  q.chunk = chunk;
  q.res = res;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the persistent kernel over resident arrays in the memory m: workgroups
// of 256 lanes, waves_per_cu waves on each of the device's CUs (0:
// automatic); chunk: cells per queue grab (0: as many as the wave's idle
// lanes)
template <class Net, bool ET, bool MIXED>
int launch_persistent_t(CellsArgs a, int wpc, int chunk, hipStream_t s,
                        PersistMem m) {
  long lanes = 0;
  if (persist_lanes(a.ncell, wpc, lanes) != 0) return -1;
  PersistWorkOf<Net, ET, MIXED> w;
  PersistQueue q;
  const int r = persist_prepare<Net, ET, MIXED>(a, m, chunk, s, w, q);
  if (r != 0) return r;
  constexpr long B = KX_PERSIST_BLOCK;
  hipLaunchKernelGGL((persist_cells_kernel<Net, ET, MIXED>),
                     dim3(lanes / B), dim3(B), 0, s, a, w, q);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the persistent kernel factorizes its Newton matrix in registers: for
// networks of up to KX_PERSIST_MAX_N equations; larger ones are refused
// (-104) and take the plain kernel
#ifndef KX_PERSIST_MAX_N
#define KX_PERSIST_MAX_N 16
#endif

// whether a launch factorizes the FP32 Newton matrix: with the
// difference-quotient Jacobian and Tgas evolved; with Tgas fixed the
// matrix stays FP64.  With the analytic Jacobian its FP32 form lost the
// trace ions of cold, dense cells; with difference quotients at fixed Tgas
// it failed cells of coolHCN, primordial2 and primordial_wD (rhs failures,
// convergence failures)
template <bool ET>
inline bool persist_mixed(const CellsArgs& a, int mixed) {
  const bool dq = a.o.jac == 2 || (a.o.jac == 0 && ET);
  return mixed && ET && dq;
}

// ---- the lane queue: networks above KX_PERSIST_MAX_N equations ----------
// bytes of one lane slot of the lane-queue launch (the plain solve's
// interleaved workspace: doubles, then pivots)
template <class Net, bool ET>
constexpr size_t lane_slot_bytes() {
  using Prob = PlainProblem<Net, ET>;
  return sizeof(double) * cell_doubles<Prob>() + sizeof(int) * Prob::N;
}
// resident lanes of a lane-queue launch over ncell cells in workgroups of
// block lanes: waves_per_cu waves of 64 on each CU (0: 4, one per SIMD at
// this kernel's register use), no more than the cells rounded up to a
// workgroup
inline int lane_queue_lanes(long ncell, int wpc, long& lanes,
                            int block = 128) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess)
    return -1;
  if (wpc <= 0) wpc = 4;
  lanes = (long)ncu * wpc * 64;
  const long need = (ncell + block - 1) / block * block;
  if (lanes > need) lanes = need;
  lanes = (lanes + block - 1) / block * block;
  return 0;
}
// bytes of a lane-queue launch's workspace for a given number of lanes:
// the counter's 256 bytes, then the slots
template <class Net, bool ET>
constexpr size_t lane_queue_bytes(long lanes) {
  return 256 + lane_slot_bytes<Net, ET>() * (size_t)lanes;
}
// the queue (counter zeroed by a kernel, the densest-first permutation)
// and the slots' workspace in m, on stream s
template <class Net, bool ET>
int lane_queue_prepare(CellsArgs a, PersistMem m, long lanes, hipStream_t s,
                       LaneQueue& q, double*& work, int*& iwork) {
  using Prob = PlainProblem<Net, ET>;
  char* p = static_cast<char*>(m.work);
  q.counter = reinterpret_cast<int*>(p);
  p += 256;
  work = reinterpret_cast<double*>(p);
  p += sizeof(double) * cell_doubles<Prob>() * (size_t)lanes;
  iwork = reinterpret_cast<int*>(p);
  int* hist = static_cast<int*>(m.ord);
  int* key = hist + ORDER_BUCKETS;
  int* order = key + a.ncell;
  const int cb = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(fill_flags_kernel, dim3(1), dim3(64), 0, s, 1,
                     q.counter, 0);
  hipLaunchKernelGGL(fill_flags_kernel, dim3(ORDER_BUCKETS / 256), dim3(256),
                     0, s, ORDER_BUCKETS, hist, 0);
  hipLaunchKernelGGL(order_keys_kernel<Net>, dim3(cb), dim3(256), 0, s, a,
                     key, hist);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(64), 0, s, hist);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(cb), dim3(256), 0, s,
                     a.ncell, key, hist, order);
  q.order = order;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
// the lane-queue launch with its memory from the caller (as
// launch_persistent): FP64 throughout (mixed precision does not apply);
// block: lanes per workgroup, 64 or 128 (0: 128)
template <class Net, bool ET, class Mem>
inline int launch_lane_queue(CellsArgs a, int wpc, hipStream_t s,
                             Mem&& mem, int block = 128) {
  if (block <= 0) block = 128;
  if (block != 64 && block != 128) return -1;
  long lanes = 0;
  if (lane_queue_lanes(a.ncell, wpc, lanes, block) != 0) return -1;
  PersistMem m;
  const int r = mem(lane_queue_bytes<Net, ET>(lanes),
                    persist_order_bytes(a.ncell), m);
  if (r != 0) return r;
  LaneQueue q;
  double* work = nullptr;
  int* iwork = nullptr;
  if (lane_queue_prepare<Net, ET>(a, m, lanes, s, q, work, iwork) != 0)
    return -1;
  hipLaunchKernelGGL((lane_queue_kernel<Net, ET>), dim3(lanes / block),
                     dim3(block), 0, s, a, q, work, iwork, lanes);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---- a wave per cell: networks above KX_BIG_N species --------------------
// wave slots of a launch over ncell cells: waves_per_cu waves on each CU
// (0: as many as are resident at once, 8 up to 8 cells a workgroup, 16 at
// 16), cells per workgroup cpb, no more than the cells rounded up to a
// workgroup
inline int wave_queue_slots(long ncell, int wpc, int cpb, long& slots) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess)
    return -1;
  if (wpc <= 0) wpc = cpb > 8 ? 16 : 8;
  slots = (long)ncu * wpc;
  if (slots > ncell) slots = ncell;
  slots = (slots + cpb - 1) / cpb * cpb;
  return 0;
}
template <class Net>
constexpr size_t wave_queue_bytes(long slots) {
  return 256 + (sizeof(double) * wave_slot_doubles<Net>() +
                sizeof(int) * Net::NSPEC) * (size_t)slots;
}
// the queue and the slots' workspace in m (as lane_queue_prepare)
template <class Net>
int wave_queue_prepare(CellsArgs a, PersistMem m, long slots, hipStream_t s,
                       LaneQueue& q, double*& work, int*& iwork) {
  char* p = static_cast<char*>(m.work);
  q.counter = reinterpret_cast<int*>(p);
  p += 256;
  work = reinterpret_cast<double*>(p);
  p += sizeof(double) * wave_slot_doubles<Net>() * (size_t)slots;
  iwork = reinterpret_cast<int*>(p);
  int* hist = static_cast<int*>(m.ord);
  int* key = hist + ORDER_BUCKETS;
  int* order = key + a.ncell;
  const int cb = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(fill_flags_kernel, dim3(1), dim3(64), 0, s, 1,
                     q.counter, 0);
  hipLaunchKernelGGL(fill_flags_kernel, dim3(ORDER_BUCKETS / 256), dim3(256),
                     0, s, ORDER_BUCKETS, hist, 0);
  hipLaunchKernelGGL(order_keys_kernel<Net>, dim3(cb), dim3(256), 0, s, a,
                     key, hist);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(64), 0, s, hist);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(cb), dim3(256), 0, s,
                     a.ncell, key, hist, order);
  q.order = order;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
// the wave-per-cell launch with its memory from the caller: cpb cells per
// workgroup (1, 2, 4, 8 or 16; 0: 4), waves_per_cu waves per CU (0: 8, or 16 at 16 cells);
// FP64 (mixed precision does not apply)
template <class Net, class Mem>
inline int launch_wave_queue(CellsArgs a, int wpc, int cpb, hipStream_t s,
                             Mem&& mem) {
  if (cpb <= 0) cpb = 4;
  if (cpb != 1 && cpb != 2 && cpb != 4 && cpb != 8 && cpb != 16) return -1;
  long slots = 0;
  if (wave_queue_slots(a.ncell, wpc, cpb, slots) != 0) return -1;
  PersistMem m;
  const int r = mem(wave_queue_bytes<Net>(slots),
                    persist_order_bytes(a.ncell), m);
  if (r != 0) return r;
  LaneQueue q;
  double* work = nullptr;
  int* iwork = nullptr;
  if (wave_queue_prepare<Net>(a, m, slots, s, q, work, iwork) != 0)
    return -1;
  const size_t lds = sizeof(double) * wave_lds_rows(Net::NSPEC) * 64 * cpb;
  if (cpb > 8)
    hipLaunchKernelGGL((wave_queue_kernel<Net, 1024>), dim3(slots / cpb),
                       dim3(64 * cpb), lds, s, a, q, work, iwork, slots);
  else
    hipLaunchKernelGGL((wave_queue_kernel<Net, 512>), dim3(slots / cpb),
                       dim3(64 * cpb), lds, s, a, q, work, iwork, slots);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// bytes of the workspace a launch needs (0: the network is refused)
template <class Net, bool ET>
inline size_t persist_need(const CellsArgs& a, int mixed) {
  if constexpr (Problem<Net, ET>::N > KX_PERSIST_MAX_N) {
    (void)a;
    (void)mixed;
    return 0;
  } else {
    return persist_mixed<ET>(a, mixed) ? persist_work_bytes<Net, ET, true>()
                                       : persist_work_bytes<Net, ET, false>();
  }
}

// the persistent launch with its memory from the caller: mem(work_bytes,
// order_bytes, m) fills m, returning 0 or a flag
template <class Net, bool ET, class Mem>
inline int launch_persistent(CellsArgs a, int mixed, int chunk, int wpc,
                             hipStream_t s, Mem&& mem) {
  if constexpr (Net::NSPEC > KX_BIG_N && !ET) {
    // a large network: a wave per cell from the device work queue
    (void)mixed; (void)chunk;
    return launch_wave_queue<Net>(a, wpc, 0, s, mem);
  } else if constexpr (Problem<Net, ET>::N > KX_PERSIST_MAX_N) {
    // too large for the lockstep kernel's registers: the lane queue
    (void)mixed; (void)chunk;
    return launch_lane_queue<Net, ET>(a, wpc, s, mem);
  } else {
    // the analytic Jacobian with Tgas evolved or with density-dependent
    // rates (its difference quotients are per lane, not the lockstep
    // machine's shared rhs): the lane queue
    const bool dq = a.o.jac == 2 || (a.o.jac == 0 && ET);
    if (!dq && (ET || Net::RATES_DENSITY)) {
      (void)mixed; (void)chunk;
      return launch_lane_queue<Net, ET>(a, wpc, s, mem);
    }
    PersistMem m;
    const int r = mem(persist_need<Net, ET>(a, mixed),
                      persist_order_bytes(a.ncell), m);
    if (r != 0) return r;
    return persist_mixed<ET>(a, mixed)
               ? launch_persistent_t<Net, ET, true>(a, wpc, chunk, s, m)
               : launch_persistent_t<Net, ET, false>(a, wpc, chunk, s, m);
  }
}

}  // namespace kx
