// ORACLE (test infrastructure only, never the product path): a CPU
// restatement of the per-cell integrator of KROME-X,
// runtime/core/cvode/bdf_cell.hpp, itself CVODE 7.9.0's BDF method.  Unit
// stride, FP64, one cell.  The steps, constants and the order of the
// floating-point operations are CVODE's (they decide step sizes, orders
// and therefore bitwise results), each function citing the reference
// lines it restates.  The reference header, compiled with the same
// problem (oracle/Makefile, target ref), must give bitwise the same
// states and counters: tests/test_oracle.py checks that.
#pragma once
#include <cfloat>
#include <cmath>

// a hook for per-step tracing in debugging builds; empty otherwise
#ifndef KXO_TRACE
#define KXO_TRACE(tag, a)
#define KXO_TRACE2(tag, a, b)
#endif

namespace kxo {

constexpr int QMAX = 5;            // bdf_cell.hpp:103
constexpr int LMAX = QMAX + 1;

// CVODE's return flags (bdf_cell.hpp:107-122)
enum : int {
  OK = 0, TOO_MUCH_WORK = -1, TOO_MUCH_ACC = -2, ERR_FAILURE = -3,
  CONV_FAILURE = -4, LSETUP_FAIL = -6, LSOLVE_FAIL = -7, RHSFUNC_FAIL = -8,
  FIRST_RHSFUNC_ERR = -9, REPTD_RHSFUNC_ERR = -10, UNREC_RHSFUNC_ERR = -11,
  NLS_FAIL = -16, ILL_INPUT = -22, TOO_CLOSE = -27
};

// the species block's static-pivot sparse LU tables in pivot order
// (bdf_cell.hpp:125-131): perm[k] the equation of pivot k; row k of L + U
// at ci[rp[k] .. rp[k+1]), columns ascending; di[k] its diagonal
struct Sparse {
  int ns = 0, nnz = 0;
  const int *perm = nullptr, *rp = nullptr, *ci = nullptr, *di = nullptr;
};

// options of one integration (bdf_cell.hpp:133-170)
struct Opts {
  double rtol = 1e-4;
  const double* atol = nullptr;
  // per-equation relative tolerances (KROME's rtol_species), or nullptr
  const double* rtol_v = nullptr;
  long max_steps = 500;
  int max_err_fails = 7;
  int max_conv_fails = 10;
  double h0 = 0.0;
  bool dq_jac = false;
  // the linear invariant kept by the Newton corrections (charge)
  int inv_row = -1, inv_n = 0;
  const int* inv_c = nullptr;
  // the static-pivot sparse LU instead of the dense one
  bool sparse = false;
  Sparse sp;
};

// CVODE's counters (bdf_cell.hpp:177-191)
struct Stats {
  long nst = 0, nfe = 0, nfeDQ = 0, nje = 0, nsetups = 0, nni = 0, nnf = 0,
       ncfn = 0, netf = 0;
  void add(const Stats& s) {
    nst += s.nst; nfe += s.nfe; nfeDQ += s.nfeDQ; nje += s.nje;
    nsetups += s.nsetups; nni += s.nni; nnf += s.nnf; ncfn += s.ncfn;
    netf += s.netf;
  }
};

// workspace of an N-equation cell: Nordsieck array, 9 vectors, saved J,
// Newton matrix: N x N, or the sparse factors (nnz), their row
// accumulator (ns), the border's L^-1 B and C U^-1 (ns x m each) and its
// Schur complement (m x m), m = N - ns (bdf_cell.hpp:206-220)
inline long newton_doubles(int N, bool sparse = false, int ns = 0,
                           int nnz = 0) {
  const long m = N - ns;
  return sparse ? nnz + ns + 2L * ns * m + m * m : (long)N * N;
}
inline long work_doubles(int N, bool sparse = false, int ns = 0,
                         int nnz = 0) {
  return (LMAX + 9) * (long)N + (long)N * N +
         newton_doubles(N, sparse, ns, nnz);
}

namespace c_ {
// cvode.c constants (bdf_cell.hpp:241-272)
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr int MAX_ITERS = 4;
constexpr double CORTES = 0.1, ETA_MIN_FX = 0.0, ETA_MAX_FX = 1.5,
                 ETA_MAX_FS = 10000.0, ETA_MAX_ES = 10.0, ETA_MAX_GS = 10.0,
                 ETA_MIN = 0.1, ETA_MAX_EF = 0.2, ETA_MIN_EF = 0.1,
                 ETA_CF = 0.25;
constexpr long SMALL_NST = 10;
constexpr int SMALL_NEF = 2;
constexpr double ONEPSM = 1.000001, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0,
                 BIAS3 = 10.0;
constexpr int LONG_WAIT = 10, MXNEF1 = 3;
constexpr long MSBP = 20, MSBJ = 51;
constexpr double DGMAX_LSETUP = 0.3, DGMAX_JBAD = 0.2, MIN_INC_MULT = 1000.0;
constexpr int NLS_MAXCOR = 3;
constexpr double CRDOWN = 0.3, RDIV = 2.0;
// internal flags (bdf_cell.hpp:228-238)
enum : int {
  DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, TRY_AGAIN = 5, FIRST_CALL = 6,
  PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 9, RHSFUNC_RECVR = 10,
  NLS_CONTINUE = 901, NLS_CONV_RECVR = 902
};
enum : int { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };

inline double mx(double a, double b) { return a > b ? a : b; }
inline double mn(double a, double b) { return a < b ? a : b; }
inline double rt(double x) { return x <= 0.0 ? 0.0 : std::sqrt(x); }

// y += a x, with N_VLinearSum's a = +-1 forms (bdf_cell.hpp:293-307)
inline void axpy(int N, double a, const double* x, double* y) {
  if (a == 1.0) {
    for (int i = 0; i < N; i++) y[i] += x[i];
  } else if (a == -1.0) {
    for (int i = 0; i < N; i++) y[i] -= x[i];
  } else {
    for (int i = 0; i < N; i++) y[i] += a * x[i];
  }
}
// z = a x + b y with the special cases of the serial N_VLinearSum, each
// rounding its own way (bdf_cell.hpp:308-374)
inline void lsum(int N, double a, const double* x, double b, const double* y,
                 double* z) {
  if (b == 1.0 && z == y) return axpy(N, a, x, z);
  if (a == 1.0 && z == x) return axpy(N, b, y, z);
  if (a == 1.0 && b == 1.0) {
    for (int i = 0; i < N; i++) z[i] = x[i] + y[i];
    return;
  }
  if ((a == 1.0 && b == -1.0) || (a == -1.0 && b == 1.0)) {
    const bool t = a == 1.0;  // z = x - y or y - x
    const double* p = t ? x : y;
    const double* m = t ? y : x;
    for (int i = 0; i < N; i++) z[i] = p[i] - m[i];
    return;
  }
  if (a == 1.0 || b == 1.0) {
    const bool t = a == 1.0;
    const double c = t ? b : a;
    const double* v1 = t ? y : x;  // scaled
    const double* v2 = t ? x : y;
    for (int i = 0; i < N; i++) z[i] = (c * v1[i]) + v2[i];
    return;
  }
  if (a == -1.0 || b == -1.0) {
    const bool t = a == -1.0;
    const double c = t ? b : a;
    const double* v1 = t ? y : x;
    const double* v2 = t ? x : y;
    for (int i = 0; i < N; i++) z[i] = (c * v1[i]) - v2[i];
    return;
  }
  if (a == b) {
    for (int i = 0; i < N; i++) z[i] = a * (x[i] + y[i]);
    return;
  }
  if (a == -b) {
    for (int i = 0; i < N; i++) z[i] = a * (x[i] - y[i]);
    return;
  }
  for (int i = 0; i < N; i++) z[i] = (a * x[i]) + (b * y[i]);
}
inline void scal(int N, double c, const double* x, double* z) {
  for (int i = 0; i < N; i++) z[i] = c * x[i];
}
inline double wrms(int N, const double* x, const double* w) {
  double s = 0.0;
  for (int i = 0; i < N; i++) {
    const double p = x[i] * w[i];
    s += p * p;
  }
  return rt(s / N);
}
inline double maxnorm(int N, const double* x) {
  double m = 0.0;
  for (int i = 0; i < N; i++)
    if (std::fabs(x[i]) > m) m = std::fabs(x[i]);
  return m;
}
// dense LU with partial pivoting, column-major (bdf_cell.hpp:402-433):
// 0, or k + 1 at a zero pivot
inline int lu(double* a, int n, int* p) {
  for (int k = 0; k < n; k++) {
    double* ck = a + (long)k * n;
    int l = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(ck[i]) > std::fabs(ck[l])) l = i;
    p[k] = l;
    if (ck[l] == 0.0) return k + 1;
    if (l != k)
      for (int j = 0; j < n; j++) {
        double* cj = a + (long)j * n;
        const double t = cj[l];
        cj[l] = cj[k];
        cj[k] = t;
      }
    const double r = 1.0 / ck[k];
    for (int i = k + 1; i < n; i++) ck[i] *= r;
    for (int j = k + 1; j < n; j++) {
      double* cj = a + (long)j * n;
      const double akj = cj[k];
      if (akj != 0.0)
        for (int i = k + 1; i < n; i++) cj[i] -= akj * ck[i];
    }
  }
  return 0;
}
// the triangular solves (bdf_cell.hpp:435-457)
inline void lus(const double* a, int n, const int* p, double* b) {
  for (int k = 0; k < n; k++) {
    const int pk = p[k];
    if (pk != k) {
      const double t = b[k];
      b[k] = b[pk];
      b[pk] = t;
    }
  }
  for (int k = 0; k < n - 1; k++) {
    const double* ck = a + (long)k * n;
    for (int i = k + 1; i < n; i++) b[i] -= ck[i] * b[k];
  }
  for (int k = n - 1; k > 0; k--) {
    const double* ck = a + (long)k * n;
    b[k] /= ck[k];
    for (int i = 0; i < k; i++) b[i] -= ck[i] * b[k];
  }
  b[0] /= a[0];
}
}  // namespace c_

// one cell's integrator: CVodeReInit + CVode(tout, CV_NORMAL)
// (bdf_cell.hpp:491-1501)
template <class Prob>
class Bdf {
 public:
  Stats st;

  Bdf(Prob& prob, int n, double* work, int* iwork, const Opts& opts)
      : P(prob), N(n), o(opts) {
    double* w = work;
    for (int j = 0; j < LMAX; j++, w += N) zn[j] = w;
    double** v[9] = {&ewt, &y, &acor, &tempv, &ftemp, &delta, &vt1, &vt2,
                     &vt3};
    for (double** q : v) {
      *q = w;
      w += N;
    }
    A = w;
    w += newton_doubles(N, o.sparse, o.sp.ns, o.sp.nnz);
    J = w;
    piv = iwork;
    mxstep = o.max_steps == 0 ? 500 : o.max_steps;
    double m = o.atol[0];
    for (int i = 1; i < N; i++) m = o.atol[i] < m ? o.atol[i] : m;
    atolmin0 = m == 0.0;
  }

  // bdf_cell.hpp:536-562
  void reinit(double t0, const double* y0, double h_init) {
    tn = t0;
    for (int i = 0; i < N; i++) zn[0][i] = y0[i];
    q = 1;
    L = 2;
    qwait = L;
    etamax = c_::ETA_MAX_FS;
    qu = 0;
    hu = 0.0;
    tolsf = 1.0;
    st = Stats();
    nstlp = nscon = nstlj = 0;
    h0u = next_h = 0.0;
    next_q = 0;
    hin = h_init;
    acnrmcur = nls_jcur = jcur = false;
    saved_tq5 = 0.0;
    for (int i = 0; i <= LMAX; i++) tau[i] = 0.0;
    eta = 1.0;
  }

  // CVode(tout, CV_NORMAL) (bdf_cell.hpp:566-638)
  int solve(double tout, double* yout, double* tret) {
    using namespace c_;
    *tret = tn;
    const double tdist = std::fabs(tout - tn);
    const double tround = DBL_EPSILON * c_::mx(std::fabs(tn), std::fabs(tout));
    if (tdist == 0.0 || tdist < 2.0 * tround) return TOO_CLOSE;
    if (ewt_set(zn[0], ewt) != 0) return ILL_INPUT;
    const int r0 = P.rhs(tn, zn[0], zn[1]);
    st.nfe++;
    if (r0 < 0) return RHSFUNC_FAIL;
    if (r0 > 0) return FIRST_RHSFUNC_ERR;
    h = hin;
    if (h != 0.0 && (tout - tn) * h < 0.0) return ILL_INPUT;
    if (h == 0.0) {
      const int hf = hin_est(tout);
      if (hf != OK) return hf;
    }
    // hmax_inv = 0, hmin = 0: no bound applies
    hscale = h;
    h0u = h;
    hprime = h;
    c_::scal(N, h, zn[1], zn[1]);
    long nloc = 0;
    int istate = OK;
    for (;;) {
      next_h = h;
      next_q = q;
      if (st.nst > 0 && ewt_set(zn[0], ewt) != 0) {
        istate = ILL_INPUT;
        break;
      }
      if (mxstep > 0 && nloc >= mxstep) {
        istate = TOO_MUCH_WORK;
        break;
      }
      tolsf = DBL_EPSILON * c_::wrms(N, zn[0], ewt);
      if (tolsf > 1.0) {
        istate = TOO_MUCH_ACC;
        tolsf *= 2.0;
        break;
      }
      tolsf = 1.0;
      const int kf = step();
      if (kf != OK) {
        istate = kf;
        break;
      }
      nloc++;
      if ((tn - tout) * h >= 0.0) {
        *tret = tout;
        dky(tout, yout);
        next_q = qprime;
        next_h = hprime;
        return OK;
      }
    }
    *tret = tn;
    for (int i = 0; i < N; i++) yout[i] = zn[0][i];
    return istate;
  }

  double current_step() const { return next_h; }

 private:
  Prob& P;
  const int N;
  const Opts o;
  long mxstep;
  bool atolmin0;
  double* zn[LMAX];
  double *ewt, *y, *acor, *tempv, *ftemp, *delta, *vt1, *vt2, *vt3, *A, *J;
  int* piv;
  double tn = 0, h = 0, hin = 0, hprime = 0, hscale = 0, h0u = 0, next_h = 0,
         hu = 0, eta = 1, etamax = 0, etaq = 0, etaqm1 = 0, etaqp1 = 0,
         tolsf = 1, saved_tq5 = 0;
  double tau[LMAX + 1] = {}, tq[6] = {}, l[LMAX] = {};
  double rl1 = 0, gamma = 0, gammap = 0, gamrat = 0, crate = 0, delp = 0,
         delnrm = 0, acnrm = 0;
  int q = 1, qprime = 1, next_q = 0, qu = 0, L = 2, qwait = 2, convfail = 0;
  long nscon = 0, nstlp = 0, nstlj = 0;
  bool acnrmcur = false, jcur = false, nls_jcur = false;

  // cvEwtSetSV, or with rtol_v KROME's per-slot weights
  // 1 / fma(rtol_i, |y_i|, atol_i) (bdf_cell.hpp:689-713)
  int ewt_set(const double* yc, double* w) {
    if (o.rtol_v != nullptr) {
      for (int i = 0; i < N; i++) {
        const double d = std::fma(o.rtol_v[i], std::fabs(yc[i]), o.atol[i]);
        if (!(d > 0.0)) return -1;
        w[i] = 1.0 / d;
      }
      return 0;
    }
    for (int i = 0; i < N; i++) w[i] = std::fabs(yc[i]);
    for (int i = 0; i < N; i++) w[i] = (o.rtol * w[i]) + o.atol[i];
    if (atolmin0) {
      double m = w[0];
      for (int i = 1; i < N; i++) m = w[i] < m ? w[i] : m;
      if (m <= 0.0) return -1;
    }
    for (int i = 0; i < N; i++) w[i] = 1.0 / w[i];
    return 0;
  }

  // cvHin (bdf_cell.hpp:717-764)
  int hin_est(double tout) {
    using namespace c_;
    const double tdiff = tout - tn;
    const int sign = tdiff > 0.0 ? 1 : -1;
    const double tdist = std::fabs(tdiff);
    const double tround = DBL_EPSILON * c_::mx(std::fabs(tn), std::fabs(tout));
    const double hlb = HLB_FACTOR * tround;
    const double hub = hub_est(tdist);
    double hg = c_::rt(hlb * hub);
    if (hub < hlb) {
      h = sign == -1 ? -hg : hg;
      return OK;
    }
    double hs = hg, hnew = hg, ydd = 0.0;
    for (int c1 = 1; c1 <= MAX_ITERS; c1++) {
      bool ok = false;
      for (int c2 = 1; c2 <= MAX_ITERS; c2++) {
        const int r = ydd_norm(hg * sign, &ydd);
        if (r < 0) return RHSFUNC_FAIL;
        if (r == OK) {
          ok = true;
          break;
        }
        hg *= 0.2;
      }
      if (!ok) {
        if (c1 <= 2) return REPTD_RHSFUNC_ERR;
        hnew = hs;
        break;
      }
      hs = hg;
      hnew = (ydd * hub * hub > 2.0) ? c_::rt(2.0 / ydd) : c_::rt(hg * hub);
      if (c1 == MAX_ITERS) break;
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) break;
      if (c1 > 1 && hrat > 2.0) {
        hnew = hg;
        break;
      }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = sign == -1 ? -h0 : h0;
    return OK;
  }

  // cvUpperBoundH0 (bdf_cell.hpp:766-780)
  double hub_est(double tdist) {
    double* t1 = tempv;
    double* t2 = acor;
    for (int i = 0; i < N; i++) t2[i] = std::fabs(zn[0][i]);
    ewt_set(zn[0], t1);
    for (int i = 0; i < N; i++) t1[i] = 1.0 / t1[i];
    c_::lsum(N, c_::HUB_FACTOR, t2, 1.0, t1, t1);
    for (int i = 0; i < N; i++) t2[i] = std::fabs(zn[1][i]);
    for (int i = 0; i < N; i++) t1[i] = t2[i] / t1[i];
    const double inv = c_::maxnorm(N, t1);
    double hub = c_::HUB_FACTOR * tdist;
    if (hub * inv > 1.0) hub = 1.0 / inv;
    return hub;
  }

  // cvYddNorm (bdf_cell.hpp:782-794)
  int ydd_norm(double hg, double* out) {
    c_::lsum(N, hg, zn[1], 1.0, zn[0], y);
    const int r = P.rhs(tn + hg, y, tempv);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return c_::RHSFUNC_RECVR;
    c_::lsum(N, 1.0 / hg, tempv, -1.0 / hg, zn[1], tempv);
    *out = c_::wrms(N, tempv, ewt);
    return OK;
  }

  // cvStep (bdf_cell.hpp:798-822)
  int step() {
    using namespace c_;
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
      if (kf != DO_ERROR_TEST) KXO_TRACE(1, (double)kf);
      if (kf == PREDICT_AGAIN) continue;
      if (kf != DO_ERROR_TEST) return kf;
      const int ef = error_test(&nflag, saved_t, &nef, &dsm);
      KXO_TRACE(2, dsm);
      if (ef == TRY_AGAIN) continue;
      if (ef != OK) return ef;
      break;
    }
    complete();
    next_step(dsm);
    KXO_TRACE(3, eta);
    KXO_TRACE2(6, zn[0][N - 1], zn[0][3]);
    etamax = st.nst <= SMALL_NST ? ETA_MAX_ES : ETA_MAX_GS;
    c_::scal(N, tq[2], acor, acor);
    return OK;
  }

  // cvAdjustParams / cvAdjustOrder / cvIncreaseBDF / cvDecreaseBDF /
  // cvRescale (bdf_cell.hpp:825-887)
  void adjust() {
    if (qprime != q) {
      order_change(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }
  void order_change(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) order_up();
    else if (dq == -1) order_down();
  }
  void order_up() {
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
    c_::scal(N, A1, zn[QMAX], zn[L]);
    for (int j = 2; j <= q; j++) c_::axpy(N, l[j], zn[L], zn[j]);
  }
  void order_down() {
    for (int i = 0; i <= QMAX; i++) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; j++) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; i--) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; j++) c_::axpy(N, -l[j], zn[q], zn[j]);
  }
  void rescale() {
    double c = eta;
    for (int j = 1; j <= q; j++) {
      c_::scal(N, c, zn[j], zn[j]);
      c = eta * c;
    }
    h = hscale * eta;
    next_h = h;
    hscale = h;
    nscon = 0;
  }

  // cvPredict / cvRestore (bdf_cell.hpp:890-901)
  void predict() {
    tn += h;
    for (int k = 1; k <= q; k++)
      for (int j = q; j >= k; j--) c_::axpy(N, 1.0, zn[j], zn[j - 1]);
  }
  void restore(double saved_t) {
    tn = saved_t;
    for (int k = 1; k <= q; k++)
      for (int j = q; j >= k; j--) c_::axpy(N, -1.0, zn[j], zn[j - 1]);
  }

  // cvSet, cvSetBDF, cvSetTqBDF (bdf_cell.hpp:903-954)
  void set_coeffs() {
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
    // error constants
    const double A1 = 1.0 - a0h + a0;
    const double A2 = 1.0 + q * A1;
    tq[2] = std::fabs(A1 / (a0 * A2));
    tq[5] = std::fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / l[q];
        const double A3 = a0 + 1.0 / q;
        const double A4 = a0h + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = std::fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = a0 - (1.0 / (q + 1));
      const double A6 = a0h - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = c_::CORTES / tq[2];
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (st.nst == 0) gammap = gamma;
    gamrat = st.nst > 0 ? gamma / gammap : 1.0;
  }

  // cvNls + SUNNonlinSolSolve_Newton (bdf_cell.hpp:958-1019)
  int nls(int nflag) {
    using namespace c_;
    convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES
                                                               : FAIL_OTHER;
    bool setup = nflag == PREV_CONV_FAIL || nflag == PREV_ERR_FAIL ||
                 st.nst == 0 || st.nst >= nstlp + MSBP ||
                 std::fabs(gamrat - 1.0) > DGMAX_LSETUP;
    for (int i = 0; i < N; i++) acor[i] = 0.0;
    const double tol = tq[4];
    bool jbad = false;
    long niters = 0, nfails = 0;
    int r;
    for (;;) {
      int it = 0;
      r = residual();
      if (r != OK) break;
      if (setup) {
        r = lsetup(jbad, &nls_jcur);
        if (r != OK) break;
      }
      for (;;) {
        niters++;
        const bool inv = o.dq_jac && o.inv_row >= 0;
        const double cg = inv ? inv_dot(delta) : 0.0;
        c_::scal(N, -1.0, delta, delta);
        r = lsolve(delta);
        if (r != OK) break;
        if (inv) inv_set(delta, -cg);
        c_::lsum(N, 1.0, acor, 1.0, delta, acor);
        r = conv_test(it, tol);
        KXO_TRACE2(5, delnrm, crate);
        it++;
        if (r == OK) {
          nls_jcur = false;
          break;
        }
        if (r != NLS_CONTINUE) break;
        if (it >= NLS_MAXCOR) {
          r = NLS_CONV_RECVR;
          break;
        }
        r = residual();
        if (r != OK) break;
      }
      if (r == OK) break;
      if (r > 0 && !nls_jcur) {
        nfails++;
        setup = true;
        jbad = true;
        for (int i = 0; i < N; i++) acor[i] = 0.0;
        continue;
      }
      break;
    }
    if (r != OK) nfails++;
    st.nni += niters;
    st.nnf += nfails;
    if (r != OK) return r;
    c_::lsum(N, 1.0, zn[0], 1.0, acor, y);
    if (!acnrmcur) acnrm = c_::wrms(N, acor, ewt);
    jcur = false;
    return OK;
  }

  // cvNlsResidual (bdf_cell.hpp:1022-1032)
  int residual() {
    c_::lsum(N, 1.0, zn[0], 1.0, acor, y);
    const int r = P.rhs(tn, y, ftemp);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return c_::RHSFUNC_RECVR;
    c_::lsum(N, rl1, zn[1], 1.0, acor, delta);
    c_::lsum(N, -gamma, ftemp, 1.0, delta, delta);
    return OK;
  }

  // cvNlsConvTest (bdf_cell.hpp:1035-1048)
  int conv_test(int m, double tol) {
    using namespace c_;
    delnrm = c_::wrms(N, delta, ewt);
    if (m > 0) crate = c_::mx(CRDOWN * crate, delnrm / delp);
    const double dcon = delnrm * c_::mn(1.0, crate) / tol;
    if (dcon <= 1.0) {
      acnrm = m == 0 ? delnrm : c_::wrms(N, acor, ewt);
      acnrmcur = true;
      return OK;
    }
    if (m >= 1 && delnrm > RDIV * delp) return NLS_CONV_RECVR;
    delp = delnrm;
    return NLS_CONTINUE;
  }

  // cvNlsLSetup (bdf_cell.hpp:1051-1065)
  int lsetup(bool jbad, bool* jcur_out) {
    if (jbad) convfail = c_::FAIL_BAD_J;
    const int r = ls_setup();
    st.nsetups++;
    *jcur_out = jcur;
    gamrat = 1.0;
    gammap = gamma;
    crate = 1.0;
    delnrm = 0.0;
    nstlp = st.nst;
    if (r < 0) return LSETUP_FAIL;
    if (r > 0) return c_::NLS_CONV_RECVR;
    return OK;
  }

  // cvLsSetup: J reused or re-evaluated, A = I - gamma J, LU
  // (bdf_cell.hpp:1069-1130)
  int ls_setup() {
    using namespace c_;
    const double dg = std::fabs((gamma / gammap) - 1.0);
    const bool jbad = st.nst == 0 || st.nst >= nstlj + MSBJ ||
                      (convfail == FAIL_BAD_J && dg < DGMAX_JBAD) ||
                      convfail == FAIL_OTHER;
    int r = 0;
    if (!jbad) {
      jcur = false;
    } else {
      jcur = true;
      if (!o.dq_jac)
        for (long k = 0; k < (long)N * N; k++) J[k] = 0.0;
      const int jr = o.dq_jac ? dq_jac(tn, y, ftemp, J, vt1)
                              : P.jac(tn, y, ftemp, J, vt1, vt2, vt3);
      if (jr < 0) r = -1;
      else if (jr > 0) r = 1;
    }
    if (r == 0 && !o.sparse) {
      const double c = -gamma;
      for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) {
          const long k = i + (long)j * N;
          A[k] = J[k] * c;
        }
        A[j + (long)j * N] += 1.0;
      }
    }
    if (jcur) {
      st.nje++;
      nstlj = st.nst;
    }
    if (r != 0) return r;
    if (o.sparse) return sp_factor(-gamma);
    return c_::lu(A, N, piv) > 0 ? 1 : 0;
  }

  // cvLsSolve (bdf_cell.hpp:1134-1153)
  int lsolve(double* b) {
    if (o.sparse) sp_solve(b);
    else c_::lus(A, N, piv, b);
    if (gamrat != 1.0) c_::scal(N, 2.0 / (1.0 + gamrat), b, b);
    return OK;
  }

  // ---- the static-pivot sparse LU (bdf_cell.hpp:1155-1264) -------------
  // A holds the factors val[nnz], a row accumulator w[ns], Bt = L^-1 B
  // (column b at b ns), Ct = C U^-1 (row r at r ns) and the border's Schur
  // complement Z = D - Ct Bt (m x m, column-major, pivots in piv)
  double* sp_w() const { return A + o.sp.nnz; }
  double* sp_Bt() const { return sp_w() + o.sp.ns; }
  double* sp_Ct() const { return sp_Bt() + (long)o.sp.ns * (N - o.sp.ns); }
  double* sp_Z() const {
    return sp_Ct() + (long)(N - o.sp.ns) * o.sp.ns;
  }

  // c J + I gathered in pivot order and factorized: 0, or 1 (recoverable)
  // at a zero or non-finite pivot
  int sp_factor(double c) {
    const Sparse& t = o.sp;
    const int ns = t.ns, m = N - ns;
    double* val = A;
    double* w = sp_w();
    for (int k = 0; k < ns; k++) {
      const long row = t.perm[k];
      for (int p = t.rp[k]; p < t.rp[k + 1]; p++)
        val[p] = J[row + (long)t.perm[t.ci[p]] * N] * c;
      val[t.di[k]] += 1.0;
    }
    for (int i = 0; i < ns; i++) {  // row i of L and U
      for (int p = t.rp[i]; p < t.rp[i + 1]; p++) w[t.ci[p]] = val[p];
      for (int p = t.rp[i]; p < t.di[i]; p++) {
        const int k = t.ci[p];
        const double lik = w[k] / val[t.di[k]];
        w[k] = lik;
        for (int q = t.di[k] + 1; q < t.rp[k + 1]; q++)
          w[t.ci[q]] -= lik * val[q];
      }
      const double pv = w[i];
      if (pv == 0.0 || !(std::fabs(pv) <= DBL_MAX)) return 1;
      for (int p = t.rp[i]; p < t.rp[i + 1]; p++) val[p] = w[t.ci[p]];
    }
    if (m == 0) return 0;
    double* Bt = sp_Bt();
    double* Ct = sp_Ct();
    double* Z = sp_Z();
    for (int b = 0; b < m; b++) {
      double* x = Bt + (long)b * ns;
      for (int k = 0; k < ns; k++) x[k] = J[t.perm[k] + (long)(ns + b) * N] * c;
      sp_lower(x);
    }
    for (int r = 0; r < m; r++) {
      double* x = Ct + (long)r * ns;
      for (int k = 0; k < ns; k++) x[k] = J[ns + r + (long)t.perm[k] * N] * c;
      for (int k = 0; k < ns; k++) {
        const double xk = x[k] / val[t.di[k]];
        x[k] = xk;
        if (xk != 0.0)
          for (int q = t.di[k] + 1; q < t.rp[k + 1]; q++)
            x[t.ci[q]] -= xk * val[q];
      }
    }
    for (int b = 0; b < m; b++)
      for (int r = 0; r < m; r++) {
        double z = J[ns + r + (long)(ns + b) * N] * c;
        if (r == b) z += 1.0;
        for (int k = 0; k < ns; k++)
          z -= Ct[(long)r * ns + k] * Bt[(long)b * ns + k];
        Z[r + (long)b * m] = z;
      }
    return c_::lu(Z, m, piv) > 0 ? 1 : 0;
  }
  // x = L^-1 x (unit lower triangle, pivot order)
  void sp_lower(double* x) const {
    const Sparse& t = o.sp;
    for (int i = 0; i < t.ns; i++) {
      double s = x[i];
      for (int p = t.rp[i]; p < t.di[i]; p++) s -= A[p] * x[t.ci[p]];
      x[i] = s;
    }
  }
  // b = (c J + I)^-1 b with the factors of sp_factor
  void sp_solve(double* b) {
    const Sparse& t = o.sp;
    const int ns = t.ns, m = N - ns;
    double* x = sp_w();
    for (int k = 0; k < ns; k++) x[k] = b[t.perm[k]];
    sp_lower(x);
    if (m > 0) {
      const double* Bt = sp_Bt();
      const double* Ct = sp_Ct();
      for (int r = 0; r < m; r++) {
        double s = b[ns + r];
        for (int k = 0; k < ns; k++) s -= Ct[(long)r * ns + k] * x[k];
        b[ns + r] = s;
      }
      c_::lus(sp_Z(), m, piv, b + ns);
      for (int q = 0; q < m; q++) {
        const double v = b[ns + q];
        if (v != 0.0)
          for (int k = 0; k < ns; k++) x[k] -= Bt[(long)q * ns + k] * v;
      }
    }
    for (int i = ns - 1; i >= 0; i--) {
      double s = x[i];
      for (int p = t.di[i] + 1; p < t.rp[i + 1]; p++) s -= A[p] * x[t.ci[p]];
      x[i] = s / A[t.di[i]];
    }
    for (int k = 0; k < ns; k++) b[t.perm[k]] = x[k];
  }

  // the charge invariant of the corrections (bdf_cell.hpp:1267-1287)
  double inv_dot(const double* x) const {
    double s = 0.0;
    for (int i = 0; i < o.inv_n; i++)
      if (o.inv_c[i] != 0) s += o.inv_c[i] * x[i];
    return s;
  }
  void inv_set(double* x, double target) const {
    const int r = o.inv_row;
    double s = 0.0, sa = 0.0;
    for (int i = 0; i < o.inv_n; i++)
      if (i != r && o.inv_c[i] != 0) {
        s += o.inv_c[i] * x[i];
        sa += std::fabs(o.inv_c[i] * (zn[0][i] + acor[i]));
      }
    const double yr = zn[0][r] + acor[r];
    const double xr = (target - s) / o.inv_c[r];
    if (std::fabs(o.inv_c[r] * yr) >= 1e-6 * sa && yr + xr >= 0.0) x[r] = xr;
  }

  // cvLsDenseDQJac (bdf_cell.hpp:1290-1311)
  int dq_jac(double t, double* yv, const double* fy, double* Jac,
             double* tmp) {
    const double srur = c_::rt(DBL_EPSILON);
    const double fn = c_::wrms(N, fy, ewt);
    const double minInc =
        fn != 0.0 ? (c_::MIN_INC_MULT * std::fabs(h) * DBL_EPSILON * (double)N *
                     fn)
                  : 1.0;
    int r = 0;
    for (int j = 0; j < N; j++) {
      const double ys = yv[j];
      const double inc = c_::mx(srur * std::fabs(ys), minInc / ewt[j]);
      yv[j] += inc;
      r = P.rhs(t, yv, tmp);
      st.nfeDQ++;
      if (r != 0) break;
      yv[j] = ys;
      const double ii = 1.0 / inc;
      c_::lsum(N, ii, tmp, -ii, fy, Jac + (long)j * N);
    }
    return r;
  }

  // cvHandleNFlag (bdf_cell.hpp:1315-1336)
  int handle_nflag(int* nf, double saved_t, int* ncf) {
    using namespace c_;
    const int f = *nf;
    if (f == OK) return DO_ERROR_TEST;
    st.ncfn++;
    restore(saved_t);
    if (f < 0) {
      if (f == LSETUP_FAIL || f == LSOLVE_FAIL || f == RHSFUNC_FAIL) return f;
      return NLS_FAIL;
    }
    (*ncf)++;
    etamax = 1.0;
    if (*ncf == o.max_conv_fails) {  // (|h| <= hmin ONEPSM never: hmin 0)
      if (f == NLS_CONV_RECVR) return CONV_FAILURE;
      if (f == RHSFUNC_RECVR) return REPTD_RHSFUNC_ERR;
    }
    eta = c_::mx(ETA_CF, 0.0 / std::fabs(h));
    *nf = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }

  // cvDoErrorTest (bdf_cell.hpp:1339-1381)
  int error_test(int* nf, double saved_t, int* nef, double* dsmp) {
    using namespace c_;
    const double dsm = acnrm * tq[2];
    *dsmp = dsm;
    if (dsm <= 1.0) return OK;
    (*nef)++;
    st.netf++;
    *nf = PREV_ERR_FAIL;
    restore(saved_t);
    if (*nef == o.max_err_fails) return ERR_FAILURE;
    etamax = 1.0;
    if (*nef <= MXNEF1) {
      eta = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
      eta = c_::mx(ETA_MIN_EF, c_::mx(eta, 0.0 / std::fabs(h)));
      if (*nef >= SMALL_NEF) eta = c_::mn(eta, ETA_MAX_EF);
      rescale();
      return TRY_AGAIN;
    }
    if (q > 1) {
      eta = c_::mx(ETA_MIN_EF, 0.0 / std::fabs(h));
      order_change(-1);
      L = q;
      q--;
      qwait = L;
      rescale();
      return TRY_AGAIN;
    }
    eta = c_::mx(ETA_MIN_EF, 0.0 / std::fabs(h));
    h *= eta;
    next_h = h;
    hscale = h;
    qwait = LONG_WAIT;
    nscon = 0;
    const int r = P.rhs(tn, zn[0], tempv);
    st.nfe++;
    if (r < 0) return RHSFUNC_FAIL;
    if (r > 0) return UNREC_RHSFUNC_ERR;
    c_::scal(N, h, tempv, zn[1]);
    return TRY_AGAIN;
  }

  // cvCompleteStep (bdf_cell.hpp:1386-1400)
  void complete() {
    st.nst++;
    nscon++;
    hu = h;
    qu = q;
    for (int i = q; i >= 2; i--) tau[i] = tau[i - 1];
    if (q == 1 && st.nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; j++) c_::axpy(N, l[j], acor, zn[j]);
    qwait--;
    if (qwait == 1 && q != QMAX) {
      for (int i = 0; i < N; i++) zn[QMAX][i] = acor[i];
      saved_tq5 = tq[5];
    }
  }

  // cvPrepareNextStep, cvSetEta, cvComputeEtaqm1/qp1, cvChooseEta
  // (bdf_cell.hpp:1403-1488)
  void next_step(double dsm) {
    using namespace c_;
    if (etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = c_::wrms(N, zn[q], ewt) * tq[1];
      etaqm1 = 1.0 / (std::pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    etaqp1 = 0.0;
    if (q != QMAX && saved_tq5 != 0.0) {
      double p = 1.0;  // (h / tau[2])^L
      const double b = h / tau[2];
      for (int i = 1; i <= L; i++) p *= b;
      const double cquot = (tq[5] / saved_tq5) * p;
      c_::lsum(N, -cquot, zn[QMAX], 1.0, acor, tempv);
      const double dup = c_::wrms(N, tempv, ewt) * tq[3];
      etaqp1 = 1.0 / (std::pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    // cvChooseEta
    const double etam = c_::mx(etaqm1, c_::mx(etaq, etaqp1));
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
      for (int i = 0; i < N; i++) zn[QMAX][i] = acor[i];
    }
    set_eta();
  }
  void set_eta() {
    using namespace c_;
    if (eta > ETA_MIN_FX && eta < ETA_MAX_FX) {
      eta = 1.0;
      hprime = h;
      return;
    }
    if (eta >= ETA_MAX_FX) {
      eta = c_::mn(eta, etamax);
      eta /= c_::mx(1.0, std::fabs(h) * 0.0 * eta);
    } else {
      eta = c_::mx(eta, ETA_MIN);
      eta = c_::mx(eta, 0.0 / std::fabs(h));
    }
    hprime = h * eta;
    if (qprime < q) nscon = 0;
  }

  // CVodeGetDky(t, 0) (bdf_cell.hpp:1492-1500)
  void dky(double t, double* out) {
    const double s = (t - tn) / h;
    for (int j = q; j >= 0; j--) {
      double c = 1.0;
      for (int i = 0; i < j; i++) c *= s;
      if (j == q) c_::scal(N, c, zn[j], out);
      else c_::axpy(N, c, zn[j], out);
    }
  }
};

// one CVode call from t = 0 with restarts after error-test, convergence
// or repeated-rhs failures, from a tenth of the reduced step, up to
// max_restarts times (bdf_cell.hpp:1513-1547)
template <class Prob>
int integrate(Prob& prob, int N, double* y, double dt, const Opts& o,
              int max_restarts, double* work, int* iwork, Stats* stats,
              int* nrestarts) {
  Bdf<Prob> cv(prob, N, work, iwork, o);
  cv.reinit(0.0, y, o.h0);
  double t = 0.0;
  int restarts = 0;
  for (;;) {
    const int f = cv.solve(dt, y, &t);
    if (stats != nullptr) stats->add(cv.st);
    if (f >= 0) {
      if (nrestarts) *nrestarts = restarts;
      return 0;
    }
    const bool again = f == ERR_FAILURE || f == CONV_FAILURE ||
                       f == REPTD_RHSFUNC_ERR;
    if (!again || restarts >= max_restarts) {
      if (nrestarts) *nrestarts = restarts;
      return f;
    }
    ++restarts;
    const double hc = cv.current_step();
    cv.reinit(t, y, hc > 0.0 ? 0.1 * hc : 0.0);
  }
}

}  // namespace kxo
