// The per-state mathematics of the discrete adjoint (kx_adjoint.h), for
// host and device: the rhs with each rate scaled by exp(theta_r), its
// exact Jacobian by dual numbers (kx_dual.h), the rates' part of its
// derivative applied to a vector (f_theta^T w), and the transposed solve
// with I - gamma J.  The adjoint kernel evaluates them at the states of
// its backward steps; the host build checks them against the reference's
// Jacobians and against difference quotients.
#pragma once
#include "kx_physics.h"
#include "kx_problem.h"

namespace kx {

// the problem of kx_problem.h with each rate scaled by the cell's
// exp(theta_r) (sc: element r at sc[r * scs]; null: unscaled)
template <class Net, bool ET>
struct AdjProblem {
  using NetT = Net;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = NS + (ET ? 1 : 0);
  static constexpr long NEWTON_DOUBLES = Problem<Net, ET>::NEWTON_DOUBLES;
  static constexpr bool STRIDED = false;
  static constexpr int LANES = 1;
  static constexpr long J_DOUBLES = (long)N * N;
  static constexpr long K_DOUBLES = 0;
  typename Net::Params p;
  double Tfix;
  bool hit_T0;
  const double* sc;
  long scs;

  template <class R>
  KX_HD void rates(const R* y, R* k) const {
    const R T = ET ? y[NS] : R(Tfix);
    const R Tr = kx::fmin(kx::fmax(T, R(p.Tcmb)), R(1e9));
    Net::rates(Tr, y, p, k);
    if (sc)
      for (int r = 0; r < Net::NREA; ++r) k[r] = k[r] * R(sc[r * scs]);
  }
  template <class R>
  KX_HD int rhs_r(const R* y, R* f) {
    if (ET && !(y[N - 1] > R(0))) {
      hit_T0 = true;
      return 1;
    }
    R k[Net::NREA];
    rates(y, k);
    Net::rhs(y, k, f);
    if constexpr (ET) {
      const R Tth = kx::fmin(kx::fmax(y[NS], R(2.73)), R(1e9));
      f[NS] = Net::thermal(Tth, y, k, p);
    }
    bool fin = true;
    for (int i = 0; i < N; ++i) fin = fin && kx::finite(f[i]);
    return fin ? 0 : 1;
  }
  KX_HD int rhs(const double* y, double* f) { return rhs_r<double>(y, f); }
  // the analytic species Jacobian at fixed Tgas (J zeroed by the caller)
  KX_HD int jac(const double* y, double* J) {
    double k[Net::NREA];
    rates(y, k);
    Net::jac_species(y, k, J, N);
    return 0;
  }
};

// the exact Jacobian of the rhs at y (column-major, N x N): one dual
// evaluation per column
template <class Prob>
KX_HD inline void jac_exact(Prob& P, const double* y, double* J) {
  constexpr int N = Prob::N;
  for (int j = 0; j < N; ++j) {
    Dual<1> yd[N], fd[N];
    for (int i = 0; i < N; ++i) yd[i] = Dual<1>(y[i]);
    yd[j].d[0] = 1.0;
    const bool t0 = P.hit_T0;
    (void)P.template rhs_r<Dual<1>>(yd, fd);
    P.hit_T0 = t0;
    for (int i = 0; i < N; ++i) J[i + j * N] = fd[i].d[0];
  }
}

// g_r += c (f_theta^T w)_r at y: the rates' part of the rhs's derivative by
// theta_r = log k_r, applied to w (g: element r at g[r * gs])
template <class Net, bool ET>
KX_HD inline void theta_vjp(AdjProblem<Net, ET>& P, const double* y,
                            const double* w, double c, double* g, long gs) {
  constexpr int NS = Net::NSPEC;
  constexpr int R = Net::NREA;
  double k[R];
  P.rates(y, k);
  double fac = 0.0, dk[ET ? R : 1];
  if constexpr (ET) {
    const double Tth = ::fmin(::fmax(y[NS], 2.73), 1e9);
    Net::chem_dk(Tth, y, dk);
    fac = w[NS] * (Net::gamma_index(Tth, y, P.p) - 1.0) /
          (phys::kB * Net::ntot(y));
  }
  for (int r = 0; r < R; ++r) {
    double fl = k[r];
    for (int p = Net::LT_KF_OFF[r]; p < Net::LT_KF_OFF[r + 1]; ++p)
      fl *= y[Net::LT_KF_SP[p]];
    double s = 0.0;
    for (int p = Net::LT_SC_OFF[r]; p < Net::LT_SC_OFF[r + 1]; ++p)
      s += Net::LT_SC_SGN[p] * w[Net::LT_SC_ROW[p]];
    double gr = fl * s;
    if constexpr (ET) gr += fac * k[r] * dk[r];
    g[r * gs] += c * gr;
  }
}

// mu = (I - gamma J)^-T b: LU with partial pivoting of A = I - gamma J
// (P A = L U), then U^T z = b, L^T w = z, mu = P^T w
template <int N>
KX_HD inline void solve_transposed(const double* J, double gamma,
                                   const double* b, double* mu) {
  double a[N * N];
  int pv[N];
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i)
      a[i + j * N] = (i == j ? 1.0 : 0.0) - gamma * J[i + j * N];
  for (int k = 0; k < N; ++k) {
    int p = k;
    for (int i = k + 1; i < N; ++i)
      if (::fabs(a[i + k * N]) > ::fabs(a[p + k * N])) p = i;
    pv[k] = p;
    if (p != k)
      for (int j = 0; j < N; ++j) {
        const double t = a[k + j * N];
        a[k + j * N] = a[p + j * N];
        a[p + j * N] = t;
      }
    const double d = a[k + k * N];
    const double r = d != 0.0 ? 1.0 / d : 0.0;
    for (int i = k + 1; i < N; ++i) a[i + k * N] *= r;
    for (int j = k + 1; j < N; ++j) {
      const double akj = a[k + j * N];
      for (int i = k + 1; i < N; ++i) a[i + j * N] -= a[i + k * N] * akj;
    }
  }
  double z[N];
  for (int i = 0; i < N; ++i) {  // U^T z = b (U^T lower, diagonal u_ii)
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= a[k + i * N] * z[k];
    z[i] = s / a[i + i * N];
  }
  for (int i = N - 1; i >= 0; --i) {  // L^T w = z (unit upper)
    double s = z[i];
    for (int k = i + 1; k < N; ++k) s -= a[k + i * N] * z[k];
    z[i] = s;
  }
  for (int k = N - 1; k >= 0; --k) {  // mu = P^T w
    const int p = pv[k];
    if (p != k) {
      const double t = z[k];
      z[k] = z[p];
      z[p] = t;
    }
  }
  for (int i = 0; i < N; ++i) mu[i] = z[i];
}

}  // namespace kx
