// One cell's problem on the device: the right-hand side and the analytic
// Jacobian of KROME-X's per-cell solver (runtime/core/solve_cell.hpp:
// state_T_ / thermal_T_ 540-557, full_rhs_ 750-789, rhs_host_ 804-813,
// jac_cell_ 1122-1132) over a generated network Net, with the state
// y = [n_0 .. n_{NSPEC-1} | Tgas (when evolved)] in registers.
#pragma once
#include "kx_cell.h"
#include "kx_physics.h"

namespace kx {

// the solve's options as the kernels see them (SolverOpts, krome.hpp:
// 187-255, the keys of this path)
struct SolveOpts {
  double rtol, atol_n, atol_T;
  long max_steps;
  int max_err_fails, max_conv_fails, max_restarts;
  int evolve_T;      // Tgas evolved (a thermal network)
  int jac;           // 0 auto (analytic at fixed T, else DQ), 1 analytic, 2 DQ
  int clip_negative;
  int conserve;      // 0 off, 1 electrons restore the charge, 2 project
                     // onto the input totals, 3 neutralize
  int linsol;        // 0 auto, 1 dense, 3 the static-pivot sparse LU
  // per-equation tolerances (device memory, the call's N equations):
  // tol[0..N) the absolute ones, tol[N..2N) the relative ones (read only
  // with rtol_v); nullptr: atol_n for every species, atol_T for Tgas and
  // the scalar rtol (KROME's atol_species / rtol_species resolved per
  // species on the host, solve_cell.hpp:866-879, 2129-2139)
  const double* tol = nullptr;
  int rtol_v = 0;    // the weights 1 / fma(rtol_i, |y_i|, atol_i)
};

// KC: keep the rates of the rate temperature from one evaluation to the
// next (the reference's rate cache: rates_fixed once per solve at fixed
// Tgas, and per Tgas value when Tgas evolves, so that the difference
// quotients' species columns and the Newton iterations at an unchanged
// Tgas skip them; runtime/core/solve_cell.hpp:594-607, 1302-1329).  The
// cached values are the ones the evaluation would compute: the solve is
// bit for bit the same.  The one-thread-per-cell kernels keep the cache in
// the thread's private memory; the lockstep kernel, whose lanes share one
// rhs evaluation, does not
template <class Net, bool ET, bool KC = false>
struct Problem {
  using NetT = Net;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = NS + (ET ? 1 : 0);
  // the state and the rhs in registers (N values), the saved Jacobian
  // dense (N x N); one thread per cell
  static constexpr bool STRIDED = false;
  static constexpr int LANES = 1;
  static constexpr long J_DOUBLES = (long)N * N;
  static constexpr long K_DOUBLES = 0;
  // doubles of the Newton matrix's storage: N x N dense, or the sparse
  // factors, their row accumulator, the border's two blocks and its Schur
  // complement (m = N - NS border equations)
  static constexpr long SPARSE_DOUBLES =
      Net::SP_NNZ + NS + 2L * NS * (N - NS) + (long)(N - NS) * (N - NS);
  static constexpr long NEWTON_DOUBLES =
      SPARSE_DOUBLES > (long)N * N ? SPARSE_DOUBLES : (long)N * N;
  typename Net::Params p;
  double Tfix;
  bool hit_T0;
  double kc[KC ? Net::NREA : 1];  // the rates of kc_T (KC)
  double kc_T = -1.0;             // their rate temperature; -1: none yet

  KX_HD double rate_T(const double* y) const {
    const double T = ET ? y[NS] : Tfix;
    return ::fmin(::fmax(T, p.Tcmb), 1e9);
  }
  // the rates at rate temperature Tr and state y into k
  template <class R>
  KX_HD void rates_at(R Tr, const R* y, R* k) {
    if constexpr (KC && std::is_same<R, double>::value) {
      if (!(Tr == kc_T)) {
        for (int i = 0; i < Net::NREA; ++i) kc[i] = 0.0;
        Net::rates_T(Tr, p, kc);
        kc_T = Tr;
      }
      for (int i = 0; i < Net::NREA; ++i) k[i] = kc[i];
      Net::rates_n(Tr, y, p, k);
    } else {
      Net::rates(Tr, y, p, k);
    }
  }
  // the rhs at y; 1 (recoverable) for a non-positive Tgas or a non-finite
  // result, the integrator then retries with a smaller step
  template <class R>
  KX_HD int rhs_r(const R* y, R* f) {
    if (ET && !(y[N - 1] > R(0))) {
      hit_T0 = true;
      return 1;
    }
    const R T = ET ? y[N - 1] : R(Tfix);
    const R Tr = kx::fmin(kx::fmax(T, R(p.Tcmb)), R(1e9));
    R k[Net::NREA];
    rates_at(Tr, y, k);
    Net::rhs(y, k, f);
    if constexpr (ET) {
      const R Tth = kx::fmin(kx::fmax(y[NS], R(2.73)), R(1e9));
      f[NS] = Net::thermal(Tth, y, k, p);
    }
    bool fin = true;
    for (int i = 0; i < N; ++i) fin = fin && kx_isfinite(f[i]);
    return fin ? 0 : 1;
  }
  KX_HD int rhs(const double* y, double* f) { return rhs_r<double>(y, f); }
  // the analytic Jacobian at fixed Tgas (mass action at the rates of the
  // state; J zeroed by the caller)
  KX_HD int jac(const double* y, double* J) {
    double k[Net::NREA];
    rates_at(rate_T(y), y, k);
    Net::jac_species(y, k, J, N);
    return 0;
  }
  template <class R>
  KX_HD static bool kx_isfinite(R x) {
    return kx::finite(x);
  }
};

// The problem of a large network (more than KX_BIG_N species: cloud's
// 452), whose vectors do not fit a lane's registers: the state, the rhs
// and the Jacobian stay in the cell's interleaved workspace and are read
// and written through strided views.  Tgas is fixed and no rate reads a
// density (the networks this form serves), so the rates are evaluated once
// per solve into the cell's rows k (the reference's kfix cache, runtime/
// core/solve_cell.hpp:594-607, 1302-1308) and the analytic species
// Jacobian is written straight into the sparse LU's pattern (the
// generated jac_species_sparse): J_DOUBLES = the pattern's nonzeros, no
// dense N x N matrix; the Newton matrix is the sparse factors and their
// row accumulator.
template <class Net>
struct BigProblem {
  using NetT = Net;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = NS;
  static constexpr bool STRIDED = true;
  static constexpr int LANES = 1;
  static constexpr long J_DOUBLES = Net::SP_NNZ;
  static constexpr long NEWTON_DOUBLES = Net::SP_NNZ + NS;
  static constexpr long K_DOUBLES = Net::NREA;
  static_assert(!Net::RATES_DENSITY,
                "the strided form caches the rates: none may read n");
  typename Net::Params p;
  double Tfix;
  bool hit_T0;
  SView k;       // the cell's rates rows (set by the kernel)
  bool k_set;    // evaluated for this solve

  KX_HD double rate_T() const { return ::fmin(::fmax(Tfix, p.Tcmb), 1e9); }
  // once per solve, so a call rather than a copy at each rhs site: the
  // thousands of rate expressions are compiled once
  KX_HD_CALL void rates() {
    for (int r = 0; r < Net::NREA; ++r) k[r] = 0.0;
    Net::rates_T(rate_T(), p, k);
    k_set = true;
  }
  // the rhs at y into f; 1 (recoverable) for a non-finite result
  KX_HD int rhs_s(SView y, SView f) {
    if (!k_set) rates();
    Net::rhs_loop_v(y, k, f);
    bool fin = true;
    for (int i = 0; i < N; ++i) fin = fin && kx::finite(f[i]);
    return fin ? 0 : 1;
  }
  // the analytic Jacobian at y into the pattern's values Jp
  KX_HD int jac_s(SView y, SView Jp) {
    if (!k_set) rates();
    for (long q = 0; q < J_DOUBLES; ++q) Jp[q] = 0.0;
    Net::jac_species_sparse(y, k, Jp);
    return 0;
  }
};

// The same problem solved by the 64 lanes of a wave together (a wave per
// cell, Cell with LANES 64; the cell's workspace contiguous, S = 1): the
// rates once per solve (every lane evaluates them and stores the same
// values), the rhs as the reactions' fluxes split over the lanes, then
// each species' sum of its fluxes in the loop form's order (the generated
// LT_ROW_* tables: the same additions in the same order as the scatter,
// so the same bits), the Jacobian as its terms over the lanes, then each
// pattern position's sum of its terms in loop order (LT_JPOS_*).  The
// workspace after the Cell's: the rates, the fluxes and the terms
template <class Net>
struct WaveProblem {
  using NetT = Net;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = NS;
  static constexpr bool STRIDED = true;
  static constexpr int LANES = 64;
  static constexpr long J_DOUBLES = Net::SP_NNZ;
  static constexpr long NEWTON_DOUBLES = Net::SP_NNZ + NS;
  static constexpr long K_DOUBLES = 2L * Net::NREA + Net::LT_NTERM;
  static_assert(!Net::RATES_DENSITY,
                "the wave form caches the rates: none may read n");
  typename Net::Params p;
  double Tfix;
  bool hit_T0;
  double* k;     // the cell's rates, then the fluxes, then the terms
  bool k_set;

  KX_HD double rate_T() const { return ::fmin(::fmax(Tfix, p.Tcmb), 1e9); }
  KX_HD_CALL void rates() {
    for (int r = wave_lane(); r < Net::NREA; r += LANES) k[r] = 0.0;
    wave_sync();
    Net::rates_T(rate_T(), p, k);
    wave_sync();
    k_set = true;
  }
  // the rhs at y into f (both in the workspace, contiguous); 1
  // (recoverable) for a non-finite result, the same on every lane
  KX_HD int rhs_s(SView y, SView f) {
    if (!k_set) rates();
    double* fl = k + Net::NREA;
    const int l = wave_lane();
    wave_sync();
    for (int r = l; r < Net::NREA; r += LANES) {
      double v = k[r];
      for (int q = Net::LT_KF_OFF[r]; q < Net::LT_KF_OFF[r + 1]; ++q)
        v = v * y[Net::LT_KF_SP[q]];
      fl[r] = v;
    }
    wave_sync();
    double bad = 0.0;
    for (int i = l; i < NS; i += LANES) {
      double v = 0.0;
      for (int e = Net::LT_ROW_OFF[i]; e < Net::LT_ROW_OFF[i + 1]; ++e) {
        if (Net::LT_ROW_SGN[e] > 0) v += fl[Net::LT_ROW_R[e]];
        else v -= fl[Net::LT_ROW_R[e]];
      }
      f[i] = v;
      if (!kx::finite(v)) bad = 1.0;
    }
    wave_sync();
    return wave_max(bad) > 0.0 ? 1 : 0;
  }
  // the analytic Jacobian at y into the pattern's values Jp
  KX_HD int jac_s(SView y, SView Jp) {
    if (!k_set) rates();
    double* dt = k + 2L * Net::NREA;
    const int l = wave_lane();
    wave_sync();
    for (int t = l; t < Net::LT_NTERM; t += LANES) {
      double d = k[Net::LT_JT_R[t]];
      for (int q = Net::LT_JT_FOFF[t]; q < Net::LT_JT_FOFF[t + 1]; ++q)
        d = d * y[Net::LT_JF_SP[q]];
      if (Net::LT_JT_MULT[t] > 1) d = d * double(Net::LT_JT_MULT[t]);
      dt[t] = d;
    }
    wave_sync();
    for (long q = l; q < J_DOUBLES; q += LANES) {
      double v = 0.0;
      for (int e = Net::LT_JPOS_OFF[q]; e < Net::LT_JPOS_OFF[q + 1]; ++e) {
        if (Net::LT_JPOS_SGN[e] > 0) v += dt[Net::LT_JPOS_T[e]];
        else v -= dt[Net::LT_JPOS_T[e]];
      }
      Jp[q] = v;
    }
    wave_sync();
    return 0;
  }
};

// the problem type the plain kernel solves a network with
template <class Net, bool ET, bool BIG = (Net::NSPEC > KX_BIG_N)>
struct PlainProblemOf {
  using type = Problem<Net, ET, true>;
};
template <class Net>
struct PlainProblemOf<Net, false, true> {
  using type = BigProblem<Net>;
};
// (a large network with Tgas evolved: refused before any launch)
template <class Net>
struct PlainProblemOf<Net, true, true> {
  using type = BigProblem<Net>;
};
template <class Net, bool ET>
using PlainProblem = typename PlainProblemOf<Net, ET>::type;

}  // namespace kx
