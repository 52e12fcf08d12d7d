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
  // the analytic Jacobian (runtime/core/solve_cell.hpp:883-1118,
  // JAC_ANALYTIC; J column-major N x N, zeroed by the caller): the
  // mass-action species block at the state's rates; for rates that read
  // the densities, their derivatives d k_r / d n_j by difference quotients
  // of the rates along each species (the step and cut-off of the
  // reference's fd_dk_, which it uses for table and extern rates; its
  // other density-dependent rates it differentiates symbolically), added
  // as the rhs of those derivatives (the rhs is linear in k); with Tgas
  // evolved, the Tgas column by central differences of the full rhs and
  // the Tgas row by difference quotients of the thermal rate with the
  // rates re-evaluated at the perturbed densities (fd_step_n_)
  KX_HD int jac(const double* y, double* J) {
    const double Tr = rate_T(y);
    double k[Net::NREA];
    rates_at(Tr, y, k);
    Net::jac_species(y, k, J, N);
    if constexpr (Net::RATES_DENSITY) rate_derivs(Tr, y, k, J);
    if constexpr (ET) tgas_jac(Tr, y, J);
    return 0;
  }
  // the mass-action species block alone at the state's rates (the whole
  // analytic Jacobian at fixed Tgas for rates that read only Tgas; J
  // zeroed by the caller)
  KX_HD int jac_block(const double* y, double* J) {
    double k[Net::NREA];
    rates_at(rate_T(y), y, k);
    Net::jac_species(y, k, J, N);
    return 0;
  }
  // J += the density-dependent rates' derivatives (a function of its own:
  // its rate arrays stay out of the caller's frame)
  KX_HD_CALL void rate_derivs(double Tr, const double* y, const double* k,
                              double* J) {
    {
      double yp[N], kp[Net::NREA], km[Net::NREA], col[NS];
      for (int i = 0; i < N; ++i) yp[i] = y[i];
      for (int j = 0; j < NS; ++j) {
        const double h = ::fmax(1e-7 * ::fabs(y[j]), 1e-30);
        yp[j] = y[j] + h;
        rates_at(Tr, yp, kp);
        double den = h;
        if (y[j] > h) {
          yp[j] = y[j] - h;
          rates_at(Tr, yp, km);
          den = 2.0 * h;
        } else {
          for (int r = 0; r < Net::NREA; ++r) km[r] = k[r];
        }
        yp[j] = y[j];
        bool any = false;
        for (int r = 0; r < Net::NREA; ++r) {
          const double d = kp[r] - km[r];
          kp[r] = ::fabs(d) <= 1e-15 * ::fabs(k[r]) ? 0.0 : d / den;
          any = any || kp[r] != 0.0;
        }
        if (!any) continue;
        Net::rhs(y, kp, col);
        for (int i = 0; i < NS; ++i)
          if (col[i] != 0.0) J[i + j * N] += col[i];
      }
    }
  }
  // the Tgas column and row (Tgas evolved)
  KX_HD_CALL void tgas_jac(double Tr, const double* y, double* J) {
    {
      double k[Net::NREA];
      double yp[N], fp[N], fm[N];
      for (int i = 0; i < N; ++i) yp[i] = y[i];
      const double h = ::fmax(1e-6 * ::fabs(y[NS]), 1e-8);
      const bool t0 = hit_T0;
      yp[NS] = y[NS] + h;
      rhs(yp, fp);
      yp[NS] = y[NS] - h;
      rhs(yp, fm);
      yp[NS] = y[NS];
      hit_T0 = t0;
      for (int i = 0; i < N; ++i)
        J[i + NS * N] = (fp[i] - fm[i]) / (2.0 * h);
      const double Tth = ::fmin(::fmax(y[NS], 2.73), 1e9);
      double kp[Net::NREA];
      rates_at(Tr, y, k);
      const double f0 = Net::thermal(Tth, y, k, p);
      double s = 0.0;
      for (int i = 0; i < NS; ++i) s += ::fabs(y[i]);
      for (int j = 0; j < NS; ++j) {
        const double hj =
            ::fmax(::fmax(1e-7 * ::fabs(y[j]), 1e-9 * s), 1e-30);
        yp[j] = y[j] + hj;
        rates_at(Tr, yp, kp);
        const double f1 = Net::thermal(Tth, yp, kp, p);
        yp[j] = y[j];
        J[NS + j * N] = (f1 - f0) / hj;
      }
    }
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
// LT_ROW_* tables; a row of up to 32 contributions takes the scatter's
// additions in its order, so the same bits; a longer row is summed in
// runs of 32, LT_SEG_*, whose sums are added in order), the Jacobian as
// its terms over the lanes, then each pattern position's sum of its terms
// in loop order (LT_JPOS_*: the scatter's bits).  The workspace after the
// Cell's: the rates, the fluxes and the terms or runs
template <class Net>
struct WaveProblem {
  using NetT = Net;
  static constexpr int NS = Net::NSPEC;
  static constexpr int N = NS;
  static constexpr bool STRIDED = true;
  static constexpr int LANES = 64;
  static constexpr long J_DOUBLES = Net::SP_NNZ;
  // the factors' values, the serial form's row accumulator, then the
  // factors' strict lower and upper parts again in column order (the
  // wave's triangular solves)
  static constexpr long NEWTON_DOUBLES = 2L * Net::SP_NNZ + NS;
  // the rates, the fluxes, then the Jacobian's terms or the rhs runs'
  // partial sums (never needed together)
  static constexpr long K_DOUBLES =
      2L * Net::NREA +
      (Net::LT_NTERM > Net::LT_NRUN ? Net::LT_NTERM : Net::LT_NRUN);
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
  // reactions a lane takes per batch, and contributions per gather chunk:
  // the loads of a batch or chunk are all issued before its arithmetic, so
  // they wait on memory once, not once each
  static constexpr int RB = 4;
  static constexpr int GC = 8;
  // sum of the contributions [e, e1) of a signed table (i added, ~i
  // subtracted) of the values x, from v, in table order
  KX_HD static double gather(const int* tab, int e, int e1,
                             const KX_GLOBAL double* x, double v) {
    for (; e < e1; e += GC) {
      int ix[GC];
      double xv[GC];
#pragma unroll
      for (int j = 0; j < GC; ++j) ix[j] = tab[imin(e + j, e1 - 1)];
#pragma unroll
      for (int j = 0; j < GC; ++j) xv[j] = x[ix[j] >= 0 ? ix[j] : ~ix[j]];
#pragma unroll
      for (int j = 0; j < GC; ++j)
        if (e + j < e1) v = ix[j] >= 0 ? v + xv[j] : v - xv[j];
    }
    return v;
  }
  // the rhs at y into f (both in the workspace, contiguous); 1
  // (recoverable) for a non-finite result, the same on every lane.  The
  // fluxes k_r prod y over the lanes; then each species row's sum of its
  // contributions in the loop form's order: a row of up to 32
  // contributions is summed whole by one lane (the scatter's arithmetic),
  // a longer one in runs of 32 on several lanes whose partial sums are
  // then added in order
  KX_HD int rhs_s(SView y, SView f) {
    if (!k_set) rates();
    const KX_GLOBAL double* kr = global_ptr(k);
    KX_GLOBAL double* fl = global_ptr(k + Net::NREA);
    KX_GLOBAL double* part = global_ptr(k + 2L * Net::NREA);
    const KX_GLOBAL double* yv = global_ptr(y.p);
    const long ys = y.s;
    const int l = wave_lane();
    constexpr int W = Net::LT_KF_W;
    wave_sync();
    for (int r0 = l; r0 < Net::NREA; r0 += RB * LANES) {
      double kv[RB], xv[RB][W];
      int sv[RB][W];
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const int r = imin(r0 + b * LANES, Net::NREA - 1);
        kv[b] = kr[r];
#pragma unroll
        for (int j = 0; j < W; ++j) sv[b][j] = Net::LT_KFW[r * W + j];
      }
#pragma unroll
      for (int b = 0; b < RB; ++b)
#pragma unroll
        for (int j = 0; j < W; ++j)
          xv[b][j] = yv[(sv[b][j] >= 0 ? sv[b][j] : 0) * ys];
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        if (r0 + b * LANES < Net::NREA) {
          double v = kv[b];
#pragma unroll
          for (int j = 0; j < W; ++j)
            if (sv[b][j] >= 0) v = v * xv[b][j];
          fl[r0 + b * LANES] = v;
        }
      }
    }
    wave_sync();
    for (int s = l; s < Net::LT_NSEG; s += LANES) {
      const int id = Net::LT_SEG_ID[s];
      const double v = gather(Net::LT_ROW_RS, Net::LT_SEG_LO[s],
                              Net::LT_SEG_HI[s], fl, 0.0);
      if (id >= 0) part[id] = v;
    }
    wave_sync();
    double bad = 0.0;
    for (int i = l; i < NS; i += LANES) {
      const int a = Net::LT_RSEG_OFF[i], b = Net::LT_RSEG_OFF[i + 1];
      double v = part[a];
      for (int q = a + 1; q < b; ++q) v += part[q];
      f[i] = v;
      if (!kx::finite(v)) bad = 1.0;
    }
    wave_sync();
    return wave_max(bad) > 0.0 ? 1 : 0;
  }
  // the analytic Jacobian at y into the pattern's values Jp: the terms
  // over the lanes, then each pattern position's sum of its terms in the
  // loop form's order (the scatter's arithmetic)
  KX_HD int jac_s(SView y, SView Jp) {
    if (!k_set) rates();
    const KX_GLOBAL double* kr = global_ptr(k);
    KX_GLOBAL double* dt = global_ptr(k + 2L * Net::NREA);
    const KX_GLOBAL double* yv = global_ptr(y.p);
    KX_GLOBAL double* jv = global_ptr(Jp.p);
    const long ys = y.s, js = Jp.s;
    const int l = wave_lane();
    constexpr int W = Net::LT_JF_W;
    wave_sync();
    for (int t0 = l; t0 < Net::LT_NTERM; t0 += RB * LANES) {
      double kv[RB], xv[RB][W];
      int sv[RB][W], mv[RB];
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const int t = imin(t0 + b * LANES, Net::LT_NTERM - 1);
        kv[b] = kr[Net::LT_JT_R[t]];
        mv[b] = Net::LT_JT_MULT[t];
#pragma unroll
        for (int j = 0; j < W; ++j) sv[b][j] = Net::LT_JFW[t * W + j];
      }
#pragma unroll
      for (int b = 0; b < RB; ++b)
#pragma unroll
        for (int j = 0; j < W; ++j)
          xv[b][j] = yv[(sv[b][j] >= 0 ? sv[b][j] : 0) * ys];
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        if (t0 + b * LANES < Net::LT_NTERM) {
          double d = kv[b];
#pragma unroll
          for (int j = 0; j < W; ++j)
            if (sv[b][j] >= 0) d = d * xv[b][j];
          if (mv[b] > 1) d = d * double(mv[b]);
          dt[t0 + b * LANES] = d;
        }
      }
    }
    wave_sync();
    for (long q = l; q < J_DOUBLES; q += LANES)
      jv[q * js] = gather(Net::LT_JPOS_TS, Net::LT_JPOS_OFF[q],
                          Net::LT_JPOS_OFF[q + 1], dt, 0.0);
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
