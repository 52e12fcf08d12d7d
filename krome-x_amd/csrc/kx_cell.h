// The per-cell BDF integrator of the plain FP64 drop-in kernel: one GPU
// thread per cell, every vector and matrix of the cell in a workspace of
// interleaved cells (element i of thread slot c at c + i S: a wave's 64
// lanes read 64 consecutive doubles).  It is KROME-X's integrator
// (runtime/core/cvode/bdf_cell.hpp:491-1547, CVODE 7.9.0's BDF in
// Nordsieck form, orders 1-5, modified Newton, dense LU with partial
// pivoting, local error test, order/step selection, restarts), with the
// same steps, constants and operation order, so that a cell reproduces the
// CPU solve up to the last bits of the math library.  N, the number of
// equations, is a compile-time constant: the loops unroll and the
// problem's rhs works on register arrays.
#pragma once
#include <cfloat>

#include "kx_physics.h"

namespace kx {

constexpr int QMAX = 5;
constexpr int LMAX = QMAX + 1;

// one cell's vector in the interleaved workspace, as an indexable view:
// element i at p[i s]
struct SView {
  double* p;
  long s;
  KX_HD double& operator[](long i) const { return p[i * s]; }
};

// ---- a wave per cell ------------------------------------------------------
// A large network's cell can be solved by the 64 lanes of a wave together
// (Prob::LANES = 64): lane l handles elements l, l + 64, ... of every
// vector, the reductions (norms, minima, maxima) go through a xor
// butterfly, and the rhs, the Jacobian, the sparse LU and its solves are
// split over the lanes (kx_problem.h WaveProblem, Cell's sp_factor /
// sp_solve).  Every lane of the wave computes the same scalars, so the
// control flow is the wave's own; data another lane wrote are read after
// wave_sync().
KX_HD inline int wave_lane() {
#if defined(__HIP_DEVICE_COMPILE__)
  return threadIdx.x & 63;
#else
  return 0;
#endif
}
// the sum, minimum and maximum over the wave's 64 lanes, the same bits on
// every lane: a butterfly within each row of 16 lanes by DPP moves (lane
// pairs 1, 2, then the half-row and row mirrors: each step combines two
// partials in either order, commutative, so both lanes of a pair agree),
// then the four row results read as scalars and combined as (r0 + r1) +
// (r2 + r3).  Called by all 64 lanes together
#if defined(__HIP_DEVICE_COMPILE__)
template <int CTRL>
__device__ inline double wave_dpp(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffu), CTRL,
                                             0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF,
                                             0xF, false);
  return __builtin_bit_cast(
      double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline double wave_read(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffu), lane);
  const unsigned hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <class Op>
__device__ inline double wave_reduce(double v, Op op) {
  v = op(v, wave_dpp<0xB1>(v));   // quad_perm [1, 0, 3, 2]
  v = op(v, wave_dpp<0x4E>(v));   // quad_perm [2, 3, 0, 1]
  v = op(v, wave_dpp<0x141>(v));  // row_half_mirror
  v = op(v, wave_dpp<0x140>(v));  // row_mirror
  return op(op(wave_read(v, 0), wave_read(v, 16)),
            op(wave_read(v, 32), wave_read(v, 48)));
}
#endif
KX_HD inline double wave_sum(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return wave_reduce(v, [](double x, double y) { return x + y; });
#else
  return v;
#endif
}
KX_HD inline double wave_max(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return wave_reduce(v, [](double x, double y) { return ::fmax(x, y); });
#else
  return v;
#endif
}
KX_HD inline double wave_min(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return wave_reduce(v, [](double x, double y) { return ::fmin(x, y); });
#else
  return v;
#endif
}
// what the wave's lanes stored is visible to its other lanes (a wave's
// memory operations complete in order; this orders the compiler's)
KX_HD inline void wave_sync() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
#endif
}
// pointers the device code may name by address space: LDS (ds_read and
// ds_write, which wait only on the LDS counter) and global memory
// (global_load and global_store rather than flat accesses, which make
// every wait cover both counters); on the host both are plain pointers
#if defined(__HIP_DEVICE_COMPILE__)
#define KX_LDS __attribute__((address_space(3)))
#define KX_GLOBAL __attribute__((address_space(1)))
#else
#define KX_LDS
#define KX_GLOBAL
#endif
template <class T>
KX_HD inline KX_LDS T* lds_ptr(T* p) {
  return (KX_LDS T*)p;
}
template <class T>
KX_HD inline KX_GLOBAL T* global_ptr(T* p) {
  return (KX_GLOBAL T*)p;
}

// the wave's row accumulator of the sparse LU in LDS: element k at
// base[(k / 64) stride + k % 64] (the wave's own 64 columns of a block's
// arena of stride doubles per row)
KX_HD constexpr int imin(int a, int b) { return a < b ? a : b; }
KX_HD constexpr int imax(int a, int b) { return a > b ? a : b; }
// rows of 64 doubles a wave's accumulator needs: elements 0 .. NS (NS is
// the spare slot)
constexpr long wave_lds_rows(long ns) { return ns / 64 + 1; }
struct WaveLds {
  KX_LDS double* base;
  long stride;
  KX_HD WaveLds() : base(nullptr), stride(0) {}
  KX_HD WaveLds(double* b, long s) : base(lds_ptr(b)), stride(s) {}
  KX_HD KX_LDS double& operator[](long k) const {
    return base[(k >> 6) * stride + (k & 63)];
  }
};

// CVODE's flags (bdf_cell.hpp:107-122)
enum : int {
  BDF_OK = 0, TOO_MUCH_WORK = -1, TOO_MUCH_ACC = -2, ERR_FAILURE = -3,
  CONV_FAILURE = -4, LSETUP_FAIL = -6, LSOLVE_FAIL = -7, RHSFUNC_FAIL = -8,
  FIRST_RHSFUNC_ERR = -9, REPTD_RHSFUNC_ERR = -10, UNREC_RHSFUNC_ERR = -11,
  NLS_FAIL = -16, ILL_INPUT = -22, TOO_CLOSE = -27
};

// KROME flags (runtime/krome.hpp:95-145)
enum : int {
  FLAG_NO_FIELD = -103, FLAG_UNSUPPORTED = -104, FLAG_PARAM_UNSET = -105,
  FLAG_NONPOSITIVE_T = -110,
  FLAG_DEVICE_MEMORY = -111
};

namespace bdfc {
// cvode.c / cvode_ls.c / cvode_nls.c constants (bdf_cell.hpp:241-272)
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr int MAX_ITERS = 4;
constexpr double CORTES = 0.1, ETA_MIN_FX = 0.0, ETA_MAX_FX = 1.5,
                 ETA_MAX_FS = 10000.0, ETA_MAX_ES = 10.0, ETA_MAX_GS = 10.0,
                 ETA_MIN = 0.1, ETA_MAX_EF = 0.2, ETA_MIN_EF = 0.1,
                 ETA_CF = 0.25;
constexpr long SMALL_NST = 10;
constexpr int SMALL_NEF = 2;
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr int LONG_WAIT = 10, MXNEF1 = 3;
constexpr long MSBP = 20, MSBJ = 51;
constexpr double DGMAX_LSETUP = 0.3, DGMAX_JBAD = 0.2, MIN_INC_MULT = 1000.0;
constexpr int NLS_MAXCOR = 3;
constexpr double CRDOWN = 0.3, RDIV = 2.0;
enum : int {
  DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, TRY_AGAIN = 5, FIRST_CALL = 6,
  PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 9, RHSFUNC_RECVR = 10,
  NLS_CONTINUE = 901, NLS_CONV_RECVR = 902
};
enum : int { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
KX_HD inline double mx(double a, double b) { return a > b ? a : b; }
KX_HD inline double mn(double a, double b) { return a < b ? a : b; }
KX_HD inline double rt(double x) { return x <= 0.0 ? 0.0 : ::sqrt(x); }
}  // namespace bdfc

// options of a cell's integration (bdf_cell.hpp:133-170): the scalar
// rtol, per-equation absolute tolerances, the step budget and failure
// limits, the Jacobian (difference quotients or the problem's) and the
// charge invariant of the Newton corrections (inv_row >= 0: on)
#ifndef KX_BIG_N
// more equations than this: the strided form (BigProblem, kx_problem.h)
#define KX_BIG_N 64
#endif
template <int N, bool BIG = (N > KX_BIG_N)>
struct CellOpts {
  double rtol;
  double atol[N];
  // per-equation relative tolerances (KROME's rtol_species): the weights
  // are then 1 / fma(rtolv_i, |y_i|, atol_i) (bdf_cell.hpp:691-697)
  bool rtol_v;
  double rtolv[N];
  long max_steps;
  int max_err_fails, max_conv_fails, max_restarts;
  bool dq_jac;
  bool sparse;  // the static-pivot sparse LU of the species block
  int inv_row, inv_n;
  int inv_c[N];
  KX_HD double atol_at(int i) const { return atol[i]; }
  KX_HD double rtol_at(int i) const { return rtolv[i]; }
  KX_HD int invc(int i) const { return inv_c[i]; }
};
// the options of a large network (Tgas fixed, the analytic Jacobian and
// the sparse LU): per-equation tolerances read from the call's rows in
// device memory, no charge invariant (it pairs with difference quotients)
template <int N>
struct CellOpts<N, true> {
  double rtol, atol_n;
  const double* tol;  // atol[N] then rtol[N] (device), or nullptr
  bool rtol_v;
  long max_steps;
  int max_err_fails, max_conv_fails, max_restarts;
  bool dq_jac;
  bool sparse;
  int inv_row, inv_n;
  KX_HD double atol_at(int i) const { return tol ? tol[i] : atol_n; }
  KX_HD double rtol_at(int i) const { return tol ? tol[N + i] : rtol; }
  KX_HD int invc(int) const { return 0; }
};

// CVODE's counters (bdf_cell.hpp:177-191)
struct CellStats {
  long nst, nfe, nfeDQ, nje, nsetups, nni, nnf, ncfn, netf;
  KX_HD void zero() {
    nst = nfe = nfeDQ = nje = nsetups = nni = nnf = ncfn = netf = 0;
  }
  KX_HD void add(const CellStats& s) {
    nst += s.nst; nfe += s.nfe; nfeDQ += s.nfeDQ; nje += s.nje;
    nsetups += s.nsetups; nni += s.nni; nnf += s.nnf; ncfn += s.ncfn;
    netf += s.netf;
  }
};

// doubles of a cell's workspace: Nordsieck array (6 N), 9 vectors, the
// saved Jacobian (Prob::J_DOUBLES: N^2, or the sparse pattern's values of
// a large network), the Newton matrix's storage (Prob::NEWTON_DOUBLES:
// N^2, or the sparse factors when larger) and the rates of a large
// network (Prob::K_DOUBLES); pivots are N ints
template <class Prob>
constexpr long cell_doubles() {
  return (LMAX + 9) * Prob::N + Prob::J_DOUBLES + Prob::NEWTON_DOUBLES +
         Prob::K_DOUBLES;
}

// the recorder of the integrator's operations on its Nordsieck array:
// none (the solve kernels).  An active Tape (the adjoint kernel's,
// kx_adjoint.h) receives, in the order they happen, every linear map the
// step applies to zn (rescale, the Pascal predictor and its undo, order
// changes), each accepted step's corrector (order, gamma, l, the new
// state) and each zn[1] = h f(zn[0]).
struct NoTape {
  static constexpr bool ACTIVE = false;
};

// the integrator of one cell on strided storage; Prob provides
//   int rhs(const double* y, double* f)   (registers, N values)
//   int jac(const double* y, double* J)   (registers, N x N col-major,
//                                          zeroed by the caller)
template <int N, class Prob, class Tape = NoTape>
class Cell {
 public:
  CellStats st;
  Tape tp;
  // iter(): the integration goes on
  static constexpr int ITER_CONTINUE = 1;

  KX_HD Cell(Prob& prob, double* work, int* iwork, long stride,
             const CellOpts<N>& opts)
      : P(prob), S(stride), o(opts) {
    double* w = work;
    for (int j = 0; j < LMAX; j++, w += N * S) zn[j] = w;
    ewt = w; w += N * S;
    y = w; w += N * S;
    acor = w; w += N * S;
    tempv = w; w += N * S;
    ftemp = w; w += N * S;
    delta = w; w += N * S;
    vt1 = w; w += N * S;
    vt2 = w; w += N * S;
    vt3 = w; w += N * S;
    A = w; w += Prob::NEWTON_DOUBLES * S;
    J = w; w += Prob::J_DOUBLES * S;
    K = w;
    piv = iwork;
    mxstep = o.max_steps == 0 ? 500 : o.max_steps;
    double m = o.atol_at(0);
    for (int i = 1; i < N; i++) m = o.atol_at(i) < m ? o.atol_at(i) : m;
    atolmin0 = m == 0.0;
  }

  KX_HD double& at(double* v, int i) const { return v[(long)i * S]; }
  // the elements this thread handles: all (one thread per cell), or every
  // 64th from its lane (a wave per cell, Prob::LANES 64); reductions of
  // the per-thread partials over the cell's lanes
  static constexpr int VS = Prob::LANES;
  KX_HD int vb() const {
    if constexpr (VS > 1) return wave_lane();
    else return 0;
  }
  KX_HD double vsum(double v) const {
    if constexpr (VS > 1) return wave_sum(v);
    else return v;
  }
  KX_HD double vmax(double v) const {
    if constexpr (VS > 1) return wave_max(v);
    else return v;
  }
  KX_HD double vmin(double v) const {
    if constexpr (VS > 1) return wave_min(v);
    else return v;
  }
  KX_HD void vsync() const {
    if constexpr (VS > 1) wave_sync();
  }
  // the LU's row accumulator of a wave per cell (set by its kernel)
  WaveLds lw;
  // the view of a workspace vector the strided problems read and write
  KX_HD SView view(const double* v) const {
    return SView{const_cast<double*>(v), S};
  }
  // the rates' rows of a large network (Prob::K_DOUBLES of them)
  KX_HD double* rates_rows() const { return K; }

  KX_HD void reinit(double t0, const double* y0, double h_init) {
    tn = t0;
    for (int i = vb(); i < N; i += VS) at(zn[0], i) = y0[i];
    q = 1; L = 2; qwait = L;
    etamax = bdfc::ETA_MAX_FS;
    qu = 0; hu = 0.0; tolsf = 1.0;
    st.zero();
    nstlp = nscon = nstlj = 0;
    h0u = next_h = 0.0;
    next_q = 0;
    hin = h_init;
    acnrmcur = nls_jcur = jcur = false;
    saved_tq5 = 0.0;
    for (int i = 0; i <= LMAX; i++) tau[i] = 0.0;
    eta = 1.0;
  }

  // CVode(tout, CV_NORMAL) (bdf_cell.hpp:566-638); yout registers
  KX_HD int solve(double tout, double* yout, double* tret) {
    const int r = begin(tout, tret);
    if (r != BDF_OK) return r;
    long nloc = 0;
    for (;;) {
      const int k = iter(tout, yout, tret, nloc);
      if (k != ITER_CONTINUE) return k;
    }
  }
  // the part of CVode before its step loop: TOO_CLOSE, the weights, the
  // first rhs, the initial step, zn[1] = h f(zn[0]); BDF_OK or a flag
  KX_HD int begin(double tout, double* tret) {
    using namespace bdfc;
    *tret = tn;
    const double tdist = ::fabs(tout - tn);
    const double tround = DBL_EPSILON * mx(::fabs(tn), ::fabs(tout));
    if (tdist == 0.0 || tdist < 2.0 * tround) return TOO_CLOSE;
    if (ewt_set(zn[0], ewt) != 0) return ILL_INPUT;
    const int r0 = rhs_v(zn[0], zn[1]);
    st.nfe++;
    if (r0 < 0) return RHSFUNC_FAIL;
    if (r0 > 0) return FIRST_RHSFUNC_ERR;
    h = hin;
    if (h != 0.0 && (tout - tn) * h < 0.0) return ILL_INPUT;
    if (h == 0.0) {
      const int hf = hin_est(tout);
      if (hf != BDF_OK) return hf;
    }
    hscale = h; h0u = h; hprime = h;
    scal(h, zn[1], zn[1]);
    if constexpr (Tape::ACTIVE) tape_zn1f();
    return BDF_OK;
  }
  // one pass of CVode's step loop (nloc: the steps of this call):
  // ITER_CONTINUE, BDF_OK at tout (yout = y(tout)), or a failure flag
  // (yout = zn[0], the last accepted state)
  KX_HD int iter(double tout, double* yout, double* tret, long& nloc) {
    using namespace bdfc;
    int istate = BDF_OK;
    do {
      next_h = h;
      next_q = q;
      if (st.nst > 0 && ewt_set(zn[0], ewt) != 0) { istate = ILL_INPUT; break; }
      if (mxstep > 0 && nloc >= mxstep) { istate = TOO_MUCH_WORK; break; }
      tolsf = DBL_EPSILON * wrms(zn[0], ewt);
      if (tolsf > 1.0) { istate = TOO_MUCH_ACC; tolsf *= 2.0; break; }
      tolsf = 1.0;
      const int kf = step();
      if (kf != BDF_OK) { istate = kf; break; }
      nloc++;
      if ((tn - tout) * h >= 0.0) {
        *tret = tout;
        dky(tout, yout);
        next_q = qprime;
        next_h = hprime;
        return BDF_OK;
      }
      return ITER_CONTINUE;
    } while (0);
    *tret = tn;
    for (int i = vb(); i < N; i += VS) yout[i] = at(zn[0], i);
    return istate;
  }
  // the interpolation weights of CVodeGetDky(t, 0): y(t) = sum_j c[j]
  // zn[j], j <= the returned order
  KX_HD int dky_weights(double t, double* c) const {
    const double s = (t - tn) / h;
    for (int j = 0; j <= QMAX; j++) {
      double v = 1.0;
      for (int i = 0; i < j; i++) v *= s;
      c[j] = j <= q ? v : 0.0;
    }
    return q;
  }

  // the state that persists from one step to the next, as doubles (the
  // integer members exactly): the scalars, the Nordsieck array, the error
  // weights, the saved Jacobian, the Newton matrix's storage and its
  // pivots; element k at ck[k * cs]
  static constexpr int CK_SCALARS = 64;
  static constexpr long CK_DOUBLES = CK_SCALARS + (long)(LMAX + 1) * N +
                                     (long)N * N + Prob::NEWTON_DOUBLES + N;
  KX_HD void save(double* ck, long cs) const { ckpt<true>(ck, cs); }
  KX_HD void load(const double* ck, long cs) {
    ckpt<false>(const_cast<double*>(ck), cs);
  }

  KX_HD double current_step() const { return next_h; }

 private:
  template <bool SAVE, class T>
  KX_HD static void ck1(double* ck, long cs, long& k, T& v) {
    if (SAVE) ck[k * cs] = (double)v;
    else v = (T)ck[k * cs];
    ++k;
  }
  template <bool SAVE>
  KX_HD void ckpt(double* ck, long cs) const {
    Cell& m = const_cast<Cell&>(*this);
    long k = 0;
    double* d[] = {&m.tn, &m.h, &m.hin, &m.hprime, &m.hscale, &m.h0u,
                   &m.next_h, &m.hu, &m.eta, &m.etamax, &m.etaq, &m.etaqm1,
                   &m.etaqp1, &m.tolsf, &m.saved_tq5, &m.rl1, &m.gamma,
                   &m.gammap, &m.gamrat, &m.crate, &m.delp, &m.delnrm,
                   &m.acnrm};
    for (double* v : d) ck1<SAVE>(ck, cs, k, *v);
    for (int i = 0; i <= LMAX; i++) ck1<SAVE>(ck, cs, k, m.tau[i]);
    for (int i = 0; i < 6; i++) ck1<SAVE>(ck, cs, k, m.tq[i]);
    for (int i = 0; i < LMAX; i++) ck1<SAVE>(ck, cs, k, m.l[i]);
    int* iv[] = {&m.q, &m.qprime, &m.next_q, &m.qu, &m.L, &m.qwait,
                 &m.convfail};
    for (int* v : iv) ck1<SAVE>(ck, cs, k, *v);
    long* lv[] = {&m.nscon, &m.nstlp, &m.nstlj, &m.st.nst, &m.st.nfe,
                  &m.st.nfeDQ, &m.st.nje, &m.st.nsetups, &m.st.nni,
                  &m.st.nnf, &m.st.ncfn, &m.st.netf};
    for (long* v : lv) ck1<SAVE>(ck, cs, k, *v);
    bool* bv[] = {&m.acnrmcur, &m.jcur, &m.nls_jcur};
    for (bool* v : bv) ck1<SAVE>(ck, cs, k, *v);
    k = CK_SCALARS;
    for (int j = 0; j < LMAX; j++)
      for (int i = 0; i < N; i++) ck1<SAVE>(ck, cs, k, m.at(m.zn[j], i));
    for (int i = 0; i < N; i++) ck1<SAVE>(ck, cs, k, m.at(m.ewt, i));
    for (long i = 0; i < (long)N * N; i++) ck1<SAVE>(ck, cs, k, m.J[i * S]);
    for (long i = 0; i < Prob::NEWTON_DOUBLES; i++)
      ck1<SAVE>(ck, cs, k, m.A[i * S]);
    for (int i = 0; i < N; i++) ck1<SAVE>(ck, cs, k, m.piv[i * S]);
  }
  // the tape's records of zn[1] = h f(zn[0]) and of an accepted step
  KX_HD void tape_zn1f() {
    double yv[N];
    for (int i = 0; i < N; i++) yv[i] = at(zn[0], i);
    tp.zn1f(h, yv);
  }

  Prob& P;
  const long S;
  const CellOpts<N>& o;
  long mxstep;
  bool atolmin0;
  double* zn[LMAX];
  double *ewt, *y, *acor, *tempv, *ftemp, *delta, *vt1, *vt2, *vt3, *A, *J;
  double* K;
  int* piv;
  double tn = 0, h = 0, hin = 0, hprime = 0, hscale = 0, h0u = 0, next_h = 0,
         hu = 0, eta = 1, etamax = 0, etaq = 0, etaqm1 = 0, etaqp1 = 0,
         tolsf = 1, saved_tq5 = 0;
  double tau[LMAX + 1], tq[6], l[LMAX];
  double rl1 = 0, gamma = 0, gammap = 0, gamrat = 0, crate = 0, delp = 0,
         delnrm = 0, acnrm = 0;
  int q = 1, qprime = 1, next_q = 0, qu = 0, L = 2, qwait = 2, convfail = 0;
  long nscon = 0, nstlp = 0, nstlj = 0;
  bool acnrmcur = false, jcur = false, nls_jcur = false;

  // ---- vector kernels (the serial N_Vector, bdf_cell.hpp:293-400) ----
  KX_HD int rhs_v(const double* yv, double* fv) {
    if constexpr (Prob::STRIDED) {
      return P.rhs_s(view(yv), view(fv));
    } else {
      double yr[N], fr[N];
      for (int i = 0; i < N; i++) yr[i] = at(const_cast<double*>(yv), i);
      const int r = P.rhs(yr, fr);
      for (int i = 0; i < N; i++) at(fv, i) = fr[i];
      return r;
    }
  }
  KX_HD void axpy(double a, const double* x, double* yv) {
    if (a == 1.0) {
      for (int i = vb(); i < N; i += VS) at(yv, i) += at(const_cast<double*>(x), i);
    } else if (a == -1.0) {
      for (int i = vb(); i < N; i += VS) at(yv, i) -= at(const_cast<double*>(x), i);
    } else {
      for (int i = vb(); i < N; i += VS)
        at(yv, i) += a * at(const_cast<double*>(x), i);
    }
  }
  KX_HD void lsum(double a, const double* x_, double b, const double* y_,
                  double* z) {
    double* x = const_cast<double*>(x_);
    double* yv = const_cast<double*>(y_);
    if (b == 1.0 && z == yv) return axpy(a, x, z);
    if (a == 1.0 && z == x) return axpy(b, yv, z);
    if (a == 1.0 && b == 1.0) {
      for (int i = vb(); i < N; i += VS) at(z, i) = at(x, i) + at(yv, i);
      return;
    }
    if ((a == 1.0 && b == -1.0) || (a == -1.0 && b == 1.0)) {
      const bool t = a == 1.0;
      double* p = t ? x : yv;
      double* m = t ? yv : x;
      for (int i = vb(); i < N; i += VS) at(z, i) = at(p, i) - at(m, i);
      return;
    }
    if (a == 1.0 || b == 1.0) {
      const bool t = a == 1.0;
      const double c = t ? b : a;
      double* v1 = t ? yv : x;
      double* v2 = t ? x : yv;
      for (int i = vb(); i < N; i += VS) at(z, i) = (c * at(v1, i)) + at(v2, i);
      return;
    }
    if (a == -1.0 || b == -1.0) {
      const bool t = a == -1.0;
      const double c = t ? b : a;
      double* v1 = t ? yv : x;
      double* v2 = t ? x : yv;
      for (int i = vb(); i < N; i += VS) at(z, i) = (c * at(v1, i)) - at(v2, i);
      return;
    }
    if (a == b) {
      for (int i = vb(); i < N; i += VS) at(z, i) = a * (at(x, i) + at(yv, i));
      return;
    }
    if (a == -b) {
      for (int i = vb(); i < N; i += VS) at(z, i) = a * (at(x, i) - at(yv, i));
      return;
    }
    for (int i = vb(); i < N; i += VS) at(z, i) = (a * at(x, i)) + (b * at(yv, i));
  }
  KX_HD void scal(double c, const double* x, double* z) {
    for (int i = vb(); i < N; i += VS) at(z, i) = c * at(const_cast<double*>(x), i);
  }
  KX_HD double wrms(const double* x, const double* w) {
    double s = 0.0;
    for (int i = vb(); i < N; i += VS) {
      const double p = at(const_cast<double*>(x), i) *
                       at(const_cast<double*>(w), i);
      s += p * p;
    }
    return bdfc::rt(vsum(s) / N);
  }

  // cvEwtSetSV, or KROME's per-species form (bdf_cell.hpp:689-713)
  KX_HD int ewt_set(const double* yc, double* w) {
    if (o.rtol_v) {
      if constexpr (VS > 1) {  // every lane's elements, then one decision
        double bad = 0.0;
        for (int i = vb(); i < N; i += VS) {
          const double d =
              ::fma(o.rtol_at(i), ::fabs(at(const_cast<double*>(yc), i)),
                    o.atol_at(i));
          if (!(d > 0.0)) bad = 1.0;
          at(w, i) = 1.0 / d;
        }
        return vmax(bad) > 0.0 ? -1 : 0;
      } else {
        for (int i = 0; i < N; i++) {
          const double d =
              ::fma(o.rtol_at(i), ::fabs(at(const_cast<double*>(yc), i)),
                    o.atol_at(i));
          if (!(d > 0.0)) return -1;
          at(w, i) = 1.0 / d;
        }
        return 0;
      }
    }
    if constexpr (Prob::STRIDED) {  // the same steps in the workspace
      for (int i = vb(); i < N; i += VS)
        at(w, i) = (o.rtol * ::fabs(at(const_cast<double*>(yc), i))) +
                   o.atol_at(i);
      if (atolmin0) {
        double m = at(w, vb());
        for (int i = vb() + VS; i < N; i += VS)
          m = at(w, i) < m ? at(w, i) : m;
        if (vmin(m) <= 0.0) return -1;
      }
      for (int i = vb(); i < N; i += VS) at(w, i) = 1.0 / at(w, i);
      return 0;
    } else {
      double v[N];
      for (int i = 0; i < N; i++)
        v[i] = ::fabs(at(const_cast<double*>(yc), i));
      for (int i = 0; i < N; i++) v[i] = (o.rtol * v[i]) + o.atol[i];
      if (atolmin0) {
        double m = v[0];
        for (int i = 1; i < N; i++) m = v[i] < m ? v[i] : m;
        if (m <= 0.0) return -1;
      }
      for (int i = 0; i < N; i++) at(w, i) = 1.0 / v[i];
      return 0;
    }
  }

  // cvHin, cvUpperBoundH0, cvYddNorm (bdf_cell.hpp:717-794)
  KX_HD int hin_est(double tout) {
    using namespace bdfc;
    const double tdiff = tout - tn;
    const int sign = tdiff > 0.0 ? 1 : -1;
    const double tdist = ::fabs(tdiff);
    const double tround = DBL_EPSILON * mx(::fabs(tn), ::fabs(tout));
    const double hlb = HLB_FACTOR * tround;
    const double hub = hub_est(tdist);
    double hg = rt(hlb * hub);
    if (hub < hlb) {
      h = sign == -1 ? -hg : hg;
      return BDF_OK;
    }
    double hs = hg, hnew = hg, ydd = 0.0;
    for (int c1 = 1; c1 <= MAX_ITERS; c1++) {
      bool ok = false;
      for (int c2 = 1; c2 <= MAX_ITERS; c2++) {
        const int r = ydd_norm(hg * sign, &ydd);
        if (r < 0) return RHSFUNC_FAIL;
        if (r == BDF_OK) { ok = true; break; }
        hg *= 0.2;
      }
      if (!ok) {
        if (c1 <= 2) return REPTD_RHSFUNC_ERR;
        hnew = hs;
        break;
      }
      hs = hg;
      hnew = (ydd * hub * hub > 2.0) ? rt(2.0 / ydd) : rt(hg * hub);
      if (c1 == MAX_ITERS) break;
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) break;
      if (c1 > 1 && hrat > 2.0) { hnew = hg; break; }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = sign == -1 ? -h0 : h0;
    return BDF_OK;
  }
  KX_HD double hub_est(double tdist) {
    double* t1 = tempv;
    double* t2 = acor;
    for (int i = vb(); i < N; i += VS) at(t2, i) = ::fabs(at(zn[0], i));
    ewt_set(zn[0], t1);
    for (int i = vb(); i < N; i += VS) at(t1, i) = 1.0 / at(t1, i);
    lsum(bdfc::HUB_FACTOR, t2, 1.0, t1, t1);
    for (int i = vb(); i < N; i += VS) at(t2, i) = ::fabs(at(zn[1], i));
    for (int i = vb(); i < N; i += VS) at(t1, i) = at(t2, i) / at(t1, i);
    double inv = 0.0;
    for (int i = vb(); i < N; i += VS)
      if (::fabs(at(t1, i)) > inv) inv = ::fabs(at(t1, i));
    inv = vmax(inv);
    double hub = bdfc::HUB_FACTOR * tdist;
    if (hub * inv > 1.0) hub = 1.0 / inv;
    return hub;
  }
  KX_HD int ydd_norm(double hg, double* out) {
    lsum(hg, zn[1], 1.0, zn[0], y);
    const int r = rhs_v(y, tempv);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return bdfc::RHSFUNC_RECVR;
    lsum(1.0 / hg, tempv, -1.0 / hg, zn[1], tempv);
    *out = wrms(tempv, ewt);
    return BDF_OK;
  }

  // cvStep (bdf_cell.hpp:798-822)
  KX_HD int step() {
    using namespace bdfc;
    int ncf = 0, nef = 0;
    double dsm = 0.0;
    if (st.nst > 0 && hprime != h) adjust();
    const double saved_t = tn;
    int nflag = FIRST_CALL;
    for (;;) {
      predict();
      set_coeffs();
      nflag = nls(nflag);
      const int kf = handle_nflag(&nflag, saved_t, &ncf);
      if (kf == PREDICT_AGAIN) continue;
      if (kf != DO_ERROR_TEST) return kf;
      const int ef = error_test(&nflag, saved_t, &nef, &dsm);
      if (ef == TRY_AGAIN) continue;
      if (ef != BDF_OK) return ef;
      break;
    }
    complete();
    next_step(dsm);
    etamax = st.nst <= SMALL_NST ? ETA_MAX_ES : ETA_MAX_GS;
    scal(tq[2], acor, acor);
    return BDF_OK;
  }

  // order and step size changes (bdf_cell.hpp:825-887)
  KX_HD void adjust() {
    if (qprime != q) {
      order_change(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }
  KX_HD void order_change(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) order_up();
    else if (dq == -1) order_down();
  }
  KX_HD void order_up() {
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    double a1 = 1.0, prod = 1.0, xiold = 1.0, a0 = -1.0;
    l[2] = 1.0;
    double hsum = hscale;
    for (int j = 1; j < q; j++) {
      hsum += tau[j + 1];
      const double xi = hsum / hscale;
      prod *= xi;
      a0 -= 1.0 / (j + 1);
      a1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xiold + l[i - 1];
      xiold = xi;
    }
    const double A1 = (-a0 - a1) / prod;
    if constexpr (Tape::ACTIVE) tp.order_up(q, L, A1, l);
    scal(A1, zn[QMAX], zn[L]);
    for (int j = 2; j <= q; j++) axpy(l[j], zn[L], zn[j]);
  }
  KX_HD void order_down() {
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; j++) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xi + l[i - 1];
    }
    if constexpr (Tape::ACTIVE) tp.order_down(q, l);
    for (int j = 2; j < q; j++) axpy(-l[j], zn[q], zn[j]);
  }
  KX_HD void rescale() {
    if constexpr (Tape::ACTIVE) tp.rescale(eta, q);
    double c = eta;
    for (int j = 1; j <= q; j++) {
      scal(c, zn[j], zn[j]);
      c = eta * c;
    }
    h = hscale * eta;
    next_h = h;
    hscale = h;
    nscon = 0;
  }
  KX_HD void predict() {
    if constexpr (Tape::ACTIVE) tp.pascal(1, q);
    tn += h;
    for (int k = 1; k <= q; k++)
      for (int j = q; j >= k; j--) axpy(1.0, zn[j], zn[j - 1]);
  }
  KX_HD void restore(double saved_t) {
    if constexpr (Tape::ACTIVE) tp.pascal(-1, q);
    tn = saved_t;
    for (int k = 1; k <= q; k++)
      for (int j = q; j >= k; j--) axpy(-1.0, zn[j], zn[j - 1]);
  }

  // cvSet, cvSetBDF, cvSetTqBDF (bdf_cell.hpp:903-954)
  KX_HD void set_coeffs() {
    double xi_inv = 1.0, xistar_inv = 1.0, a0 = -1.0, a0h = -1.0;
    l[0] = l[1] = 1.0;
    for (int i = 2; i <= q; i++) l[i] = 0.0;
    double hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; j++) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        a0 -= 1.0 / j;
        for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xi_inv;
      }
      a0 -= 1.0 / q;
      xistar_inv = -l[1] - a0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      a0h = -l[1] - xi_inv;
      for (int i = q; i >= 1; i--) l[i] += l[i - 1] * xistar_inv;
    }
    const double A1 = 1.0 - a0h + a0;
    const double A2 = 1.0 + q * A1;
    tq[2] = ::fabs(A1 / (a0 * A2));
    tq[5] = ::fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / l[q];
        const double A3 = a0 + 1.0 / q;
        const double A4 = a0h + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = ::fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = a0 - (1.0 / (q + 1));
      const double A6 = a0h - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = ::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = bdfc::CORTES / tq[2];
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (st.nst == 0) gammap = gamma;
    gamrat = st.nst > 0 ? gamma / gammap : 1.0;
  }

  // the Newton iteration (bdf_cell.hpp:958-1065)
  KX_HD int nls(int nflag) {
    using namespace bdfc;
    convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES
                                                               : FAIL_OTHER;
    bool setup = nflag == PREV_CONV_FAIL || nflag == PREV_ERR_FAIL ||
                 st.nst == 0 || st.nst >= nstlp + MSBP ||
                 ::fabs(gamrat - 1.0) > DGMAX_LSETUP;
    for (int i = vb(); i < N; i += VS) at(acor, i) = 0.0;
    const double tol = tq[4];
    bool jbad = false;
    long niters = 0, nfails = 0;
    int r;
    for (;;) {
      int it = 0;
      r = residual();
      if (r != BDF_OK) break;
      if (setup) {
        r = lsetup(jbad, &nls_jcur);
        if (r != BDF_OK) break;
      }
      for (;;) {
        niters++;
        const bool inv = o.dq_jac && o.inv_row >= 0;
        const double cg = inv ? inv_dot(delta) : 0.0;
        scal(-1.0, delta, delta);
        lsolve(delta);
        if (inv) inv_set(delta, -cg);
        lsum(1.0, acor, 1.0, delta, acor);
        r = conv_test(it, tol);
        it++;
        if (r == BDF_OK) { nls_jcur = false; break; }
        if (r != NLS_CONTINUE) break;
        if (it >= NLS_MAXCOR) { r = NLS_CONV_RECVR; break; }
        r = residual();
        if (r != BDF_OK) break;
      }
      if (r == BDF_OK) break;
      if (r > 0 && !nls_jcur) {
        nfails++;
        setup = true;
        jbad = true;
        for (int i = vb(); i < N; i += VS) at(acor, i) = 0.0;
        continue;
      }
      break;
    }
    if (r != BDF_OK) nfails++;
    st.nni += niters;
    st.nnf += nfails;
    if (r != BDF_OK) return r;
    lsum(1.0, zn[0], 1.0, acor, y);
    if (!acnrmcur) acnrm = wrms(acor, ewt);
    jcur = false;
    return BDF_OK;
  }
  KX_HD int residual() {
    lsum(1.0, zn[0], 1.0, acor, y);
    const int r = rhs_v(y, ftemp);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return bdfc::RHSFUNC_RECVR;
    lsum(rl1, zn[1], 1.0, acor, delta);
    lsum(-gamma, ftemp, 1.0, delta, delta);
    return BDF_OK;
  }
  KX_HD int conv_test(int m, double tol) {
    using namespace bdfc;
    delnrm = wrms(delta, ewt);
    if (m > 0) crate = mx(CRDOWN * crate, delnrm / delp);
    const double dcon = delnrm * mn(1.0, crate) / tol;
    if (dcon <= 1.0) {
      acnrm = m == 0 ? delnrm : wrms(acor, ewt);
      acnrmcur = true;
      return BDF_OK;
    }
    if (m >= 1 && delnrm > RDIV * delp) return NLS_CONV_RECVR;
    delp = delnrm;
    return NLS_CONTINUE;
  }
  KX_HD int lsetup(bool jbad, bool* jcur_out) {
    if (jbad) convfail = bdfc::FAIL_BAD_J;
    const int r = ls_setup();
    st.nsetups++;
    *jcur_out = jcur;
    gamrat = 1.0;
    gammap = gamma;
    crate = 1.0;
    delnrm = 0.0;
    nstlp = st.nst;
    if (r < 0) return LSETUP_FAIL;
    if (r > 0) return bdfc::NLS_CONV_RECVR;
    return BDF_OK;
  }
  // J reused or re-evaluated, A = I - gamma J, LU (bdf_cell.hpp:1069-1130)
  KX_HD int ls_setup() {
    using namespace bdfc;
    const double dg = ::fabs((gamma / gammap) - 1.0);
    const bool jbad = st.nst == 0 || st.nst >= nstlj + MSBJ ||
                      (convfail == FAIL_BAD_J && dg < DGMAX_JBAD) ||
                      convfail == FAIL_OTHER;
    int r = 0;
    if (!jbad) {
      jcur = false;
    } else {
      jcur = true;
      const int jr = o.dq_jac ? dq_jac() : an_jac();
      if (jr < 0) r = -1;
      else if (jr > 0) r = 1;
    }
    if (r == 0 && !o.sparse && !Prob::STRIDED) {
      const double c = -gamma;
      for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) at(A, i + j * N) = at(J, i + j * N) * c;
        at(A, j + j * N) += 1.0;
      }
    }
    if (jcur) {
      st.nje++;
      nstlj = st.nst;
    }
    if (r != 0) return r;
    if constexpr (VS > 1) return sp_factor_wave(-gamma);
    if (o.sparse || Prob::STRIDED) return sp_factor(-gamma);
    if constexpr (!Prob::STRIDED) return lu() > 0 ? 1 : 0;
    return 1;
  }

  // ---- the static-pivot sparse LU (bdf_cell.hpp:1155-1264) -------------
  // The species block in the pivot order and with the fill of the
  // network's generated tables (Net::sp_perm, sp_rp, sp_ci, sp_di: uniform
  // across the wave, scalar loads); the NB = N - NS border equations (Tgas)
  // through their Schur complement.  A holds val[SP_NNZ], the row
  // accumulator w[NS], Bt = L^-1 B (column b at b NS), Ct = C U^-1 (row r
  // at r NS) and Z = D - Ct Bt (NB x NB, column-major), all strided by S.
  using Net = typename Prob::NetT;
  static constexpr int NS = Net::NSPEC;
  static constexpr int NB = N - NS;
  KX_HD double& spv(long k) const { return A[k * S]; }
  static constexpr long SP_W = Net::SP_NNZ;
  static constexpr long SP_BT = SP_W + NS;
  static constexpr long SP_CT = SP_BT + (long)NS * NB;
  static constexpr long SP_Z = SP_CT + (long)NB * NS;

  // c J + I gathered and factorized: 0, or 1 (recoverable) at a zero or
  // non-finite pivot
  KX_HD int sp_factor(double c) {
    for (int k = 0; k < NS; k++) {
      const long row = Net::sp_perm[k];
      for (int p = Net::sp_rp[k]; p < Net::sp_rp[k + 1]; p++) {
        // a large network's J holds the pattern's values in this order
        if constexpr (Prob::STRIDED) spv(p) = at(J, p) * c;
        else spv(p) = at(J, row + (long)Net::sp_perm[Net::sp_ci[p]] * N) * c;
      }
      spv(Net::sp_di[k]) += 1.0;
    }
    for (int i = 0; i < NS; i++) {  // row i of L and U
      const int p0 = Net::sp_rp[i], p1 = Net::sp_rp[i + 1];
      for (int p = p0; p < p1; p++) spv(SP_W + Net::sp_ci[p]) = spv(p);
      for (int p = p0; p < Net::sp_di[i]; p++) {
        const int k = Net::sp_ci[p];
        const double lik = spv(SP_W + k) / spv(Net::sp_di[k]);
        spv(SP_W + k) = lik;
        for (int q = Net::sp_di[k] + 1; q < Net::sp_rp[k + 1]; q++)
          spv(SP_W + Net::sp_ci[q]) -= lik * spv(q);
      }
      const double pv = spv(SP_W + i);
      if (pv == 0.0 || !(::fabs(pv) <= DBL_MAX)) return 1;
      for (int p = p0; p < p1; p++) spv(p) = spv(SP_W + Net::sp_ci[p]);
    }
    if constexpr (NB > 0) {
      for (int b = 0; b < NB; b++) {  // Bt = L^-1 B
        const long x = SP_BT + (long)b * NS;
        for (int k = 0; k < NS; k++)
          spv(x + k) = at(J, Net::sp_perm[k] + (long)(NS + b) * N) * c;
        sp_lower(x);
      }
      for (int r = 0; r < NB; r++) {  // Ct = C U^-1
        const long x = SP_CT + (long)r * NS;
        for (int k = 0; k < NS; k++)
          spv(x + k) = at(J, NS + r + (long)Net::sp_perm[k] * N) * c;
        for (int k = 0; k < NS; k++) {
          const double xk = spv(x + k) / spv(Net::sp_di[k]);
          spv(x + k) = xk;
          if (xk != 0.0)
            for (int q = Net::sp_di[k] + 1; q < Net::sp_rp[k + 1]; q++)
              spv(x + Net::sp_ci[q]) -= xk * spv(q);
        }
      }
      for (int b = 0; b < NB; b++)  // Z = D - Ct Bt
        for (int r = 0; r < NB; r++) {
          double z = at(J, NS + r + (long)(NS + b) * N) * c;
          if (r == b) z += 1.0;
          for (int k = 0; k < NS; k++)
            z -= spv(SP_CT + (long)r * NS + k) * spv(SP_BT + (long)b * NS + k);
          spv(SP_Z + r + (long)b * NB) = z;
        }
      return sp_border_lu() > 0 ? 1 : 0;
    }
    return 0;
  }
  // the same factorization by a wave per cell: the rows in order, each
  // row's entries and each U row's update split over the lanes (distinct
  // columns: the updates are independent, each the serial one's
  // arithmetic), the row accumulator in the wave's LDS (lw); Tgas fixed
  // (no border).  Every value an elimination reads from memory (the pivot
  // u_kk, the U row of k) is final before the row starts, so a row's
  // eliminations go WG at a time: all their loads first, then the
  // eliminations in order.  A lane with no U entry of its own updates the
  // accumulator's spare slot NS (the wave's LDS holds NS / 64 + 1 rows)
  static constexpr int WG = 4;
  // the longest U row off the diagonal, and the longest column of the
  // strict lower (L) and upper (U) factors
  static constexpr int sp_max_u_row() {
    int m = 0;
    for (int k = 0; k < NS; ++k)
      if (Net::sp_rp[k + 1] - Net::sp_di[k] - 1 > m)
        m = Net::sp_rp[k + 1] - Net::sp_di[k] - 1;
    return m;
  }
  template <bool UPPER>
  static constexpr int sp_max_col() {
    constexpr const int* CP = UPPER ? Net::LT_UCP : Net::LT_LCP;
    int m = 0;
    for (int k = 0; k < NS; ++k)
      if (CP[k + 1] - CP[k] > m) m = CP[k + 1] - CP[k];
    return m;
  }
  KX_HD void wave_elim_group(const KX_GLOBAL double* a, WaveLds lw, int p,
                             int pd) const {
    int k[WG], c0[WG], c1[WG];
    double piv[WG], u0[WG], u1[WG];
    int dd[WG], ee[WG];
#pragma unroll
    for (int w = 0; w < WG; ++w) {
      const int pp = imin(p + w, pd - 1);
      k[w] = Net::sp_ci[pp];
      dd[w] = Net::LT_PD[pp];
      ee[w] = Net::LT_PE[pp];
    }
    // every load unconditional (indices clamped into the row: a lane
    // without the entry loads one it does not use), the selections after
#pragma unroll
    for (int w = 0; w < WG; ++w) {
      const int d = dd[w], end = ee[w];
      const int q0 = d + 1 + vb(), q1 = q0 + VS;
      const int e0 = imin(q0, end - 1), e1 = imin(q1, end - 1);
      piv[w] = a[d];
      const int ci0 = Net::sp_ci[e0], ci1 = Net::sp_ci[e1];
      u0[w] = a[e0];
      u1[w] = a[e1];
      c0[w] = q0 < end ? ci0 : NS;
      c1[w] = q1 < end ? ci1 : NS;
    }
#pragma unroll
    for (int w = 0; w < WG; ++w) {
      if (p + w < pd) {
        // columns k < c0 < c1 (or the spare slot): the three reads first
        const double xk = lw[k[w]], x0 = lw[c0[w]], x1 = lw[c1[w]];
        const double lik = xk / piv[w];
        lw[k[w]] = lik;
        lw[c0[w]] = x0 - lik * u0[w];
        lw[c1[w]] = x1 - lik * u1[w];
        // U rows longer than two lanes' worth (none in cloud; the loop's
        // table loads would wait on the LDS writes above)
        if constexpr (sp_max_u_row() > 2 * VS) {
          const int end = Net::sp_rp[k[w] + 1];
          for (int q = Net::sp_di[k[w]] + 1 + 2 * VS + vb(); q < end;
               q += VS)
            lw[Net::sp_ci[q]] -= lik * a[q];
        }
      }
    }
  }
  // compiled once and called (KX_HD_CALL): the loop then allocates its
  // registers by itself, with no spill reloads inside it (measured 11 %
  // faster on cloud than inlined into the step)
  KX_HD_CALL int sp_factor_wave(double c) {
    static_assert(NB == 0, "a wave per cell serves fixed-Tgas networks");
    static_assert(VS == 64, "a wave per cell has 64 lanes");
    // the wave's workspace is contiguous (S = 1)
    KX_GLOBAL double* a = global_ptr(A);
    const KX_GLOBAL double* j = global_ptr(J);
    const WaveLds lw = this->lw;  // the pointers in registers
    vsync();
    for (int p = vb(); p < Net::SP_NNZ; p += VS) a[p] = j[p] * c;
    vsync();
    for (int k = vb(); k < NS; k += VS) a[Net::sp_di[k]] += 1.0;
    vsync();
    // the lane's first entry of row i (column cn, value vn; none: the
    // spare slot), loaded a row ahead: a row's values change only when it
    // is written back
    int cn = NS;
    double vn = 0.0;
    {
      const int p = Net::sp_rp[0] + vb(), e = imin(p, Net::sp_rp[1] - 1);
      const int ce = Net::sp_ci[e];
      const double ve = a[e];
      cn = p < Net::sp_rp[1] ? ce : NS;
      vn = ve;
    }
    for (int i = 0; i < NS; i++) {
      const int p0 = Net::sp_rp[i], pd = Net::sp_di[i],
                p1 = Net::sp_rp[i + 1];
      if (p0 == pd) {
        // no eliminations: the row is final as scaled, only its pivot is
        // checked (the accumulator's value of it would be the same)
        const double pv = a[pd];
        if (pv == 0.0 || !(::fabs(pv) <= DBL_MAX)) return 1;
        if (i + 1 < NS) {
          const int p = p1 + vb(), e = imin(p, Net::sp_rp[i + 2] - 1);
          const int ce = Net::sp_ci[e];
          const double ve = a[e];
          cn = p < Net::sp_rp[i + 2] ? ce : NS;
          vn = ve;
        }
        continue;
      }
      lw[cn] = vn;
      for (int p = p0 + VS + vb(); p < p1; p += VS) lw[Net::sp_ci[p]] = a[p];
      if (i + 1 < NS) {
        const int p = p1 + vb(), e = imin(p, Net::sp_rp[i + 2] - 1);
        const int ce = Net::sp_ci[e];
        const double ve = a[e];
        cn = p < Net::sp_rp[i + 2] ? ce : NS;
        vn = ve;
      }
      vsync();
      for (int p = p0; p < pd; p += WG) wave_elim_group(a, lw, p, pd);
      vsync();
      const double pv = lw[i];
      if (pv == 0.0 || !(::fabs(pv) <= DBL_MAX)) return 1;
      for (int p = p0 + vb(); p < p1; p += VS) a[p] = lw[Net::sp_ci[p]];
      vsync();
    }
    // the factors' values again in column order, for the solves
    KX_GLOBAL double* cl = a + SP_WCOL;
    KX_GLOBAL double* cu = cl + Net::LT_LCP[NS];
    for (int e = vb(); e < Net::LT_LCP[NS]; e += VS) cl[e] = a[Net::LT_LCQ[e]];
    for (int e = vb(); e < Net::LT_UCP[NS]; e += VS) cu[e] = a[Net::LT_UCQ[e]];
    vsync();
    return 0;
  }
  // where a wave's factorization keeps its factors' values in column
  // order: after the values and the serial form's row accumulator
  static constexpr long SP_WCOL = Net::SP_NNZ + NS;
  // WG columns of a factor: this lane's first two entries of each (rows
  // r0, r1, values v0, v1; none: the spare slot, 0) and, for U, the pivots
  struct WaveCols {
    int r0[WG], r1[WG];
    double v0[WG], v1[WG], d[WG];
  };
  // columns k0, k0 + s, ... (WG of them, clamped to 0 .. NS - 1) of L
  // (UPPER false) or U; cv: the part's values in column order.  Every load
  // unconditional (indices clamped into the part), the selections after
  template <bool UPPER>
  KX_HD void wave_cols_load(const KX_GLOBAL double* a,
                            const KX_GLOBAL double* cv, int k0, int s,
                            WaveCols& g) const {
    constexpr const int* CP = UPPER ? Net::LT_UCP : Net::LT_LCP;
    constexpr const int* CR = UPPER ? Net::LT_UCR : Net::LT_LCR;
    constexpr int NE = CP[NS];
#pragma unroll
    for (int w = 0; w < WG; ++w) {
      const int k = imax(imin(k0 + s * w, NS - 1), 0);
      const int e1 = CP[k + 1];
      const int q0 = CP[k] + vb(), q1 = q0 + VS;
      const int i0 = imax(imin(q0, e1 - 1), 0), i1 = imax(imin(q1, e1 - 1), 0);
      const int ra = CR[imin(i0, NE - 1)], rb = CR[imin(i1, NE - 1)];
      const double va = cv[imin(i0, NE - 1)], vb_ = cv[imin(i1, NE - 1)];
      g.r0[w] = q0 < e1 ? ra : NS;
      g.v0[w] = q0 < e1 ? va : 0.0;
      g.r1[w] = q1 < e1 ? rb : NS;
      g.v1[w] = q1 < e1 ? vb_ : 0.0;
      if (UPPER) g.d[w] = a[Net::sp_di[k]];
    }
  }
  // column k's updates of the rows below (L) or above (U) it by x_k
  template <bool UPPER>
  KX_HD static void wave_col_apply(WaveLds lw, const KX_GLOBAL double* cv,
                                   const WaveCols& g, int w, int k,
                                   double xk) {
    constexpr const int* CP = UPPER ? Net::LT_UCP : Net::LT_LCP;
    constexpr const int* CR = UPPER ? Net::LT_UCR : Net::LT_LCR;
    // distinct rows in a column: the lanes' updates are independent
    const double y0 = lw[g.r0[w]], y1 = lw[g.r1[w]];
    lw[g.r0[w]] = y0 - g.v0[w] * xk;
    lw[g.r1[w]] = y1 - g.v1[w] * xk;
    // longer columns: four more entries a lane at a time, loads first
    if constexpr (sp_max_col<UPPER>() > 2 * VS) {
    const int end = CP[k + 1];
    for (int e0 = CP[k] + 2 * VS; e0 < end; e0 += 4 * VS) {
      int rr[4];
      double vv[4], yy[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u * VS + wave_lane(), ee = imin(e, end - 1);
        const int re = CR[ee];
        const double ve = cv[ee];
        rr[u] = e < end ? re : NS;
        vv[u] = e < end ? ve : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) yy[u] = lw[rr[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) lw[rr[u]] = yy[u] - vv[u] * xk;
    }
    }
  }
  // b = (c J + I)^-1 b with sp_factor_wave's factors, b in the workspace,
  // by columns.  L y = b forward: column k's entries update their rows by
  // y_k, each row taking its updates in increasing column order, which is
  // the serial row solve's arithmetic.  U x = y backward over the
  // non-empty columns: x_k = y_k / u_kk, then its entries update the rows
  // above (each row's updates in decreasing column order: the last bits
  // differ from the serial solve); the rows of empty columns are divided
  // at the end.  The columns' operands are loaded WG columns at a time
  KX_HD_CALL void sp_solve_wave(double* bv) {
    static_assert(VS == 64, "a wave per cell has 64 lanes");
    const KX_GLOBAL double* a = global_ptr(A);
    const KX_GLOBAL double* cl = a + SP_WCOL;
    const KX_GLOBAL double* cu = cl + Net::LT_LCP[NS];
    KX_GLOBAL double* gb = global_ptr(bv);
    const WaveLds lw = this->lw;
    const int l = vb();
    vsync();
    for (int k = l; k < NS; k += VS) lw[k] = gb[Net::sp_perm[k]];
    vsync();
    for (int k0 = 0; k0 < NS; k0 += WG) {  // L y = b
      WaveCols g;
      wave_cols_load<false>(a, cl, k0, 1, g);
#pragma unroll
      for (int w = 0; w < WG; ++w) {
        const int k = k0 + w;
        if (k < NS && Net::LT_LCP[k] < Net::LT_LCP[k + 1]) {
          const double yk = lw[k];
          wave_col_apply<false>(lw, cl, g, w, k, yk);
          vsync();
        }
      }
    }
    for (int k0 = NS - 1; k0 >= 0; k0 -= WG) {  // U x = y
      WaveCols g;
      wave_cols_load<true>(a, cu, k0, -1, g);
#pragma unroll
      for (int w = 0; w < WG; ++w) {
        const int k = k0 - w;
        if (k >= 0 && Net::LT_UCP[k] < Net::LT_UCP[k + 1]) {
          const double xk = lw[k] / g.d[w];
          vsync();
          lw[k] = xk;
          wave_col_apply<true>(lw, cu, g, w, k, xk);
          vsync();
        }
      }
    }
    // the rows of the empty columns: their values are final now
    for (int k = l; k < NS; k += VS)
      if (Net::LT_UCP[k] == Net::LT_UCP[k + 1]) lw[k] = lw[k] / a[Net::sp_di[k]];
    vsync();
    for (int k = l; k < NS; k += VS) gb[Net::sp_perm[k]] = lw[k];
    vsync();
  }
  // x = L^-1 x (unit lower triangle), x at A + x0 (strided)
  KX_HD void sp_lower(long x0) {
    for (int i = 0; i < NS; i++) {
      double s = spv(x0 + i);
      for (int p = Net::sp_rp[i]; p < Net::sp_di[i]; p++)
        s -= spv(p) * spv(x0 + Net::sp_ci[p]);
      spv(x0 + i) = s;
    }
  }
  // the border's Schur complement: dense LU with partial pivoting
  // (bdf_cell.hpp:402-433), pivots in piv
  KX_HD int sp_border_lu() {
    for (int k = 0; k < NB; k++) {
      int l_ = k;
      for (int i = k + 1; i < NB; i++)
        if (::fabs(spv(SP_Z + i + k * NB)) > ::fabs(spv(SP_Z + l_ + k * NB)))
          l_ = i;
      piv[k * S] = l_;
      if (spv(SP_Z + l_ + k * NB) == 0.0) return k + 1;
      if (l_ != k)
        for (int j = 0; j < NB; j++) {
          const double t = spv(SP_Z + l_ + j * NB);
          spv(SP_Z + l_ + j * NB) = spv(SP_Z + k + j * NB);
          spv(SP_Z + k + j * NB) = t;
        }
      const double rr = 1.0 / spv(SP_Z + k + k * NB);
      for (int i = k + 1; i < NB; i++) spv(SP_Z + i + k * NB) *= rr;
      for (int j = k + 1; j < NB; j++) {
        const double akj = spv(SP_Z + k + j * NB);
        if (akj != 0.0)
          for (int i = k + 1; i < NB; i++)
            spv(SP_Z + i + j * NB) -= akj * spv(SP_Z + i + k * NB);
      }
    }
    return 0;
  }
  // b = (c J + I)^-1 b with the factors of sp_factor (b in registers)
  template <class BV>
  KX_HD void sp_solve(BV& b) {
    for (int k = 0; k < NS; k++) spv(SP_W + k) = b[Net::sp_perm[k]];
    sp_lower(SP_W);
    if constexpr (NB > 0) {
      double bb[NB];
      for (int r = 0; r < NB; r++) {
        double s = b[NS + r];
        for (int k = 0; k < NS; k++)
          s -= spv(SP_CT + (long)r * NS + k) * spv(SP_W + k);
        bb[r] = s;
      }
      for (int k = 0; k < NB; k++) {  // getrs on Z (bdf_cell.hpp:435-457)
        const int pk = piv[k * S];
        if (pk != k) {
          double t = bb[k], u = 0.0;
          for (int i = 0; i < NB; i++) u = i == pk ? bb[i] : u;
          for (int i = 0; i < NB; i++) bb[i] = i == pk ? t : bb[i];
          bb[k] = u;
        }
      }
      for (int k = 0; k < NB - 1; k++)
        for (int i = k + 1; i < NB; i++)
          bb[i] -= spv(SP_Z + i + k * NB) * bb[k];
      for (int k = NB - 1; k > 0; k--) {
        bb[k] /= spv(SP_Z + k + k * NB);
        for (int i = 0; i < k; i++) bb[i] -= spv(SP_Z + i + k * NB) * bb[k];
      }
      bb[0] /= spv(SP_Z);
      for (int q = 0; q < NB; q++) {
        b[NS + q] = bb[q];
        const double v = bb[q];
        if (v != 0.0)
          for (int k = 0; k < NS; k++)
            spv(SP_W + k) -= spv(SP_BT + (long)q * NS + k) * v;
      }
    }
    for (int i = NS - 1; i >= 0; i--) {
      double s = spv(SP_W + i);
      for (int p = Net::sp_di[i] + 1; p < Net::sp_rp[i + 1]; p++)
        s -= spv(p) * spv(SP_W + Net::sp_ci[p]);
      spv(SP_W + i) = s / spv(Net::sp_di[i]);
    }
    for (int k = 0; k < NS; k++) b[Net::sp_perm[k]] = spv(SP_W + k);
  }
  KX_HD int an_jac() {
    if constexpr (Prob::STRIDED) {
      return P.jac_s(view(y), view(J));
    } else {
      double yr[N], Jr[N * N];
      for (int i = 0; i < N; i++) yr[i] = at(y, i);
      for (int k = 0; k < N * N; k++) Jr[k] = 0.0;
      const int r = P.jac(yr, Jr);
      for (int k = 0; k < N * N; k++) at(J, k) = Jr[k];
      return r;
    }
  }
  // cvLsDenseDQJac (bdf_cell.hpp:1290-1311)
  KX_HD int dq_jac() {
    if constexpr (Prob::STRIDED) return -1;  // refused by cell_opts
    const double srur = bdfc::rt(DBL_EPSILON);
    const double fn = wrms(ftemp, ewt);
    const double minInc =
        fn != 0.0 ? (bdfc::MIN_INC_MULT * ::fabs(h) * DBL_EPSILON * (double)N *
                     fn)
                  : 1.0;
    int r = 0;
    for (int j = 0; j < N; j++) {
      const double ys = at(y, j);
      const double inc = bdfc::mx(srur * ::fabs(ys), minInc / at(ewt, j));
      at(y, j) += inc;
      r = rhs_v(y, vt1);
      st.nfeDQ++;
      if (r != 0) break;
      at(y, j) = ys;
      const double ii = 1.0 / inc;
      lsum(ii, vt1, -ii, ftemp, J + (long)j * N * S);
    }
    return r;
  }
  // dense LU with partial pivoting (bdf_cell.hpp:402-433)
  KX_HD int lu() {
    for (int k = 0; k < N; k++) {
      int l_ = k;
      for (int i = k + 1; i < N; i++)
        if (::fabs(at(A, i + k * N)) > ::fabs(at(A, l_ + k * N))) l_ = i;
      piv[k * S] = l_;
      if (at(A, l_ + k * N) == 0.0) return k + 1;
      if (l_ != k)
        for (int j = 0; j < N; j++) {
          const double t = at(A, l_ + j * N);
          at(A, l_ + j * N) = at(A, k + j * N);
          at(A, k + j * N) = t;
        }
      const double rr = 1.0 / at(A, k + k * N);
      for (int i = k + 1; i < N; i++) at(A, i + k * N) *= rr;
      for (int j = k + 1; j < N; j++) {
        const double akj = at(A, k + j * N);
        if (akj != 0.0)
          for (int i = k + 1; i < N; i++)
            at(A, i + j * N) -= akj * at(A, i + k * N);
      }
    }
    return 0;
  }
  // triangular solves and the gamma-ratio scaling (bdf_cell.hpp:435-457,
  // 1134-1153)
  KX_HD void lsolve(double* bv) {
    if constexpr (Prob::STRIDED) {  // in place in the workspace
      SView b = view(bv);
      if constexpr (VS > 1) sp_solve_wave(bv);
      else sp_solve(b);
      if (gamrat != 1.0) {
        const double c = 2.0 / (1.0 + gamrat);
        for (int i = vb(); i < N; i += VS) b[i] = c * b[i];
      }
      return;
    }
    double b[N];
    for (int i = 0; i < N; i++) b[i] = at(bv, i);
    if (o.sparse) {
      sp_solve(b);
      if (gamrat != 1.0) {
        const double c = 2.0 / (1.0 + gamrat);
        for (int i = 0; i < N; i++) b[i] = c * b[i];
      }
      for (int i = 0; i < N; i++) at(bv, i) = b[i];
      return;
    }
    for (int k = 0; k < N; k++) {
      const int pk = piv[k * S];
      if (pk != k) {
        double t = b[k], u = 0.0;
        for (int i = 0; i < N; i++) u = i == pk ? b[i] : u;
        for (int i = 0; i < N; i++) b[i] = i == pk ? t : b[i];
        b[k] = u;
      }
    }
    for (int k = 0; k < N - 1; k++)
      for (int i = k + 1; i < N; i++) b[i] -= at(A, i + k * N) * b[k];
    for (int k = N - 1; k > 0; k--) {
      b[k] /= at(A, k + k * N);
      for (int i = 0; i < k; i++) b[i] -= at(A, i + k * N) * b[k];
    }
    b[0] /= at(A, 0);
    if (gamrat != 1.0) {
      const double c = 2.0 / (1.0 + gamrat);
      for (int i = 0; i < N; i++) b[i] = c * b[i];
    }
    for (int i = 0; i < N; i++) at(bv, i) = b[i];
  }
  // the charge invariant (bdf_cell.hpp:1267-1287)
  KX_HD double inv_dot(double* x) {
    double s = 0.0;
    for (int i = 0; i < N; i++)
      if (i < o.inv_n && o.invc(i) != 0) s += o.invc(i) * at(x, i);
    return s;
  }
  KX_HD void inv_set(double* x, double target) {
    const int r = o.inv_row;
    double s = 0.0, sa = 0.0, yr = 0.0, cr = 0.0;
    for (int i = 0; i < N; i++) {
      if (i < o.inv_n && i != r && o.invc(i) != 0) {
        s += o.invc(i) * at(x, i);
        sa += ::fabs(o.invc(i) * (at(zn[0], i) + at(acor, i)));
      }
      if (i == r) {
        yr = at(zn[0], i) + at(acor, i);
        cr = o.invc(i);
      }
    }
    const double xr = (target - s) / cr;
    if (::fabs(cr * yr) >= 1e-6 * sa && yr + xr >= 0.0) at(x, r) = xr;
  }
  // cvHandleNFlag (bdf_cell.hpp:1315-1336)
  KX_HD int handle_nflag(int* nf, double saved_t, int* ncf) {
    using namespace bdfc;
    const int f = *nf;
    if (f == BDF_OK) return DO_ERROR_TEST;
    st.ncfn++;
    restore(saved_t);
    if (f < 0) {
      if (f == LSETUP_FAIL || f == LSOLVE_FAIL || f == RHSFUNC_FAIL) return f;
      return NLS_FAIL;
    }
    (*ncf)++;
    etamax = 1.0;
    if (::fabs(h) <= 0.0 || *ncf == o.max_conv_fails) {
      if (f == NLS_CONV_RECVR) return CONV_FAILURE;
      if (f == RHSFUNC_RECVR) return REPTD_RHSFUNC_ERR;
    }
    eta = mx(ETA_CF, 0.0);
    *nf = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }
  // cvDoErrorTest (bdf_cell.hpp:1339-1381)
  KX_HD int error_test(int* nf, double saved_t, int* nef, double* dsmp) {
    using namespace bdfc;
    const double dsm = acnrm * tq[2];
    *dsmp = dsm;
    if (dsm <= 1.0) return BDF_OK;
    (*nef)++;
    st.netf++;
    *nf = PREV_ERR_FAIL;
    restore(saved_t);
    if (::fabs(h) <= 0.0 || *nef == o.max_err_fails) return ERR_FAILURE;
    etamax = 1.0;
    if (*nef <= MXNEF1) {
      eta = 1.0 / (::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
      eta = mx(ETA_MIN_EF, mx(eta, 0.0));
      if (*nef >= SMALL_NEF) eta = mn(eta, ETA_MAX_EF);
      rescale();
      return TRY_AGAIN;
    }
    if (q > 1) {
      eta = mx(ETA_MIN_EF, 0.0);
      order_change(-1);
      L = q;
      q--;
      qwait = L;
      rescale();
      return TRY_AGAIN;
    }
    eta = mx(ETA_MIN_EF, 0.0);
    h *= eta;
    next_h = h;
    hscale = h;
    qwait = LONG_WAIT;
    nscon = 0;
    const int r = rhs_v(zn[0], tempv);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return UNREC_RHSFUNC_ERR;
    scal(h, tempv, zn[1]);
    if constexpr (Tape::ACTIVE) tape_zn1f();
    return TRY_AGAIN;
  }
  // cvCompleteStep (bdf_cell.hpp:1386-1400)
  KX_HD void complete() {
    st.nst++;
    nscon++;
    hu = h;
    qu = q;
    for (int i = q; i >= 2; i--) tau[i] = tau[i - 1];
    if (q == 1 && st.nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; j++) axpy(l[j], acor, zn[j]);
    qwait--;
    const bool save_q = qwait == 1 && q != QMAX;
    if (save_q) {
      for (int i = vb(); i < N; i += VS) at(zn[QMAX], i) = at(acor, i);
      saved_tq5 = tq[5];
    }
    if constexpr (Tape::ACTIVE) {
      double yv[N];
      for (int i = 0; i < N; i++) yv[i] = at(zn[0], i);
      tp.step(q, gamma, rl1, l, yv, save_q);
    }
  }
  // cvPrepareNextStep and the order / step choice (bdf_cell.hpp:1403-1488)
  KX_HD void next_step(double dsm) {
    using namespace bdfc;
    if (etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = wrms(zn[q], ewt) * tq[1];
      etaqm1 = 1.0 / (::pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    etaqp1 = 0.0;
    if (q != QMAX && saved_tq5 != 0.0) {
      double p = 1.0;
      const double b = h / tau[2];
      for (int i = 1; i <= L; i++) p *= b;
      const double cquot = (tq[5] / saved_tq5) * p;
      lsum(-cquot, zn[QMAX], 1.0, acor, tempv);
      const double dup = wrms(tempv, ewt) * tq[3];
      etaqp1 = 1.0 / (::pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    const double etam = mx(etaqm1, mx(etaq, etaqp1));
    if (etam > ETA_MIN_FX && etam < ETA_MAX_FX) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
      for (int i = vb(); i < N; i += VS) at(zn[QMAX], i) = at(acor, i);
      if constexpr (Tape::ACTIVE) tp.saveq();
    }
    set_eta();
  }
  KX_HD void set_eta() {
    using namespace bdfc;
    if (eta > ETA_MIN_FX && eta < ETA_MAX_FX) {
      eta = 1.0;
      hprime = h;
      return;
    }
    if (eta >= ETA_MAX_FX) {
      eta = mn(eta, etamax);
    } else {
      eta = mx(eta, ETA_MIN);
    }
    hprime = h * eta;
    if (qprime < q) nscon = 0;
  }
  // CVodeGetDky(t, 0) (bdf_cell.hpp:1492-1500)
  KX_HD void dky(double t, double* out) {
    const double s = (t - tn) / h;
    for (int i = vb(); i < N; i += VS) out[i] = 0.0;
    for (int j = q; j >= 0; j--) {
      double c = 1.0;
      for (int i = 0; i < j; i++) c *= s;
      if (j == q)
        for (int i = vb(); i < N; i += VS) out[i] = c * at(zn[j], i);
      else
        for (int i = vb(); i < N; i += VS) out[i] += c * at(zn[j], i);
    }
  }
};

}  // namespace kx
