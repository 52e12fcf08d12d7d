// Host build of the adjoint (krome-x_amd/csrc/kx_adjmath.h, kx_adjoint.h)
// for the CPU tests: the rate-scaled rhs, its exact Jacobian by dual
// numbers, the rates' part of its derivative applied to a vector, the
// transposed solve with I - gamma J, and the whole per-cell adjoint solve
// (adjoint_cell, the GPU kernel's thread body) run cell by cell.  Built by
// tests/test_adjoint.py with g++ into tests/_build/.  Networks: 0
// primordial (Tgas fixed), 1 primordial3 (Tgas evolved).
#include <cstring>

#include <vector>

#include "kx_adjoint.h"
#include "net_primordial.h"
#include "net_primordial3.h"

namespace {

template <class Net, bool ET>
struct Cells {
  using Prob = kx::AdjProblem<Net, ET>;
  static constexpr int NSPEC = Net::NSPEC;
  static constexpr int N = Prob::N;
  static constexpr int R = Net::NREA;
  // the problem of cell c: its parameter row, fixed Tgas and rate scales
  static void setup(Prob& P, int c, const double* Tfix, const double* par,
                    const double* lnk, double* sc) {
    P.p = Net::params_from_row(par + (long)c * Net::NPAR);
    P.Tfix = Tfix[c];
    P.hit_T0 = false;
    P.sc = nullptr;
    P.scs = 1;
    if (lnk) {
      for (int r = 0; r < R; ++r) sc[r] = std::exp(lnk[(long)c * R + r]);
      P.sc = sc;
    }
  }
  static int rhs(int nc, const double* y, const double* Tfix,
                 const double* par, const double* lnk, double* f) {
    for (int c = 0; c < nc; ++c) {
      Prob P;
      double sc[R];
      setup(P, c, Tfix, par, lnk, sc);
      if (P.rhs(y + (long)c * N, f + (long)c * N) != 0) return 1;
    }
    return 0;
  }
  static int jac(int nc, const double* y, const double* Tfix,
                 const double* par, const double* lnk, double* J) {
    for (int c = 0; c < nc; ++c) {
      Prob P;
      double sc[R];
      setup(P, c, Tfix, par, lnk, sc);
      kx::jac_exact(P, y + (long)c * N, J + (long)c * N * N);
    }
    return 0;
  }
  static int vjp(int nc, const double* y, const double* Tfix,
                 const double* par, const double* lnk, const double* w,
                 double* g) {
    for (int c = 0; c < nc; ++c) {
      Prob P;
      double sc[R];
      setup(P, c, Tfix, par, lnk, sc);
      for (int r = 0; r < R; ++r) g[(long)c * R + r] = 0.0;
      kx::theta_vjp<Net, ET>(P, y + (long)c * N, w + (long)c * N, 1.0,
                             g + (long)c * R, 1);
    }
    return 0;
  }
  // the device entry's arrays (n, gn, g_n0 species-major [NSPEC][nc],
  // lnk, g_lnk [nc][NREA]); o: evolve_T, rtol, atol_n, atol_T, max_steps,
  // checkpoint_steps; gn null: forward only.  Returns the slots used.
  static long adjoint(int nc, double* n, double* T, const double* dt,
                      const double* par, const double* lnk, int* flags,
                      const double* gn, const double* gT, double* g_lnk,
                      double* g_n0, double* g_T0, const double* o) {
    kx::AdjArgs g;
    g.a.ncell = nc;
    g.a.n = n;
    g.a.n_cs = 1;
    g.a.n_ss = nc;
    g.a.T = T;
    g.a.dt = dt;
    g.a.par = par;
    g.a.par_stride = Net::NPAR;
    g.a.flags = flags;
    g.a.stats = nullptr;
    kx::SolveOpts& so = g.a.o;
    so.evolve_T = ET;
    so.rtol = o[1];
    so.atol_n = o[2];
    so.atol_T = o[3];
    so.max_steps = (long)o[4];
    so.max_err_fails = 7;
    so.max_conv_fails = 10;
    so.max_restarts = 5;
    so.jac = 0;
    so.clip_negative = 1;
    so.conserve = 0;
    so.linsol = 1;
    g.ld = nc;
    g.lnk = lnk;
    g.gn = gn;
    g.gT = gT;
    g.g_lnk = g_lnk;
    g.g_n0 = g_n0;
    g.g_T0 = g_T0;
    const bool bwd = gn || gT;
    const int seg = (int)o[5];
    g.seg_steps = seg;
    g.margin = bwd ? 4 * 17 + 8 : 0;
    g.cap = bwd ? 4 * seg + g.margin : 0;
    g.ycap = bwd ? seg + 2 * 9 + 16 : 0;
    // o[6] != 0: every cell's workspace interleaved at stride nc, as in a
    // launch (else each cell alone at stride 1)
    const bool il = o[6] != 0.0;
    g.S = il ? nc : 1;
    const long cd = kx::adj_cell_doubles<Net, ET>(g.cap, g.ycap);
    std::vector<double> work(cd * g.S);
    std::vector<int> iwork((long)N * g.S);
    const long ckd = kx::adj_ck_doubles<Net, ET>();
    const long slots = 1 << 16;
    std::vector<double> pool(bwd ? slots * ckd : 1);
    unsigned long long next = 0;
    g.work = work.data();
    g.iwork = iwork.data();
    g.pool = pool.data();
    g.pool_slots = slots;
    g.pool_next = &next;
    long used = 0;
    constexpr int R = Net::NREA;
    if (il) {  // the launch's layout: cell c at offset c, one pool
      g.pool_slots = slots;
      for (int c = 0; c < nc; ++c) kx::adjoint_cell<Net, ET>(g, c);
      return (long)next;
    }
    for (int c = 0; c < nc; ++c) {
      // one cell at a time, as cell 0 of arrays that start at it (its
      // workspace at stride 1, the pool reused)
      kx::AdjArgs b = g;
      b.a.ncell = 1;
      b.a.n = n + c;
      b.a.T = T + c;
      b.a.dt = dt + c;
      b.a.par = par + (long)c * Net::NPAR;
      b.a.flags = flags + c;
      b.lnk = lnk ? lnk + (long)c * R : nullptr;
      b.gn = gn ? gn + c : nullptr;
      b.gT = gT ? gT + c : nullptr;
      b.g_lnk = g_lnk ? g_lnk + (long)c * R : nullptr;
      b.g_n0 = g_n0 ? g_n0 + c : nullptr;
      b.g_T0 = g_T0 ? g_T0 + c : nullptr;
      next = 0;
      kx::adjoint_cell<Net, ET>(b, 0);
      used = next > (unsigned long long)used ? (long)next : used;
    }
    return used;
  }
  // one cell's whole tape (one segment): the forward run of the adjoint's
  // integrator recording every operation; rec[cap * TW], ys[ycap * N];
  // out: nrec, ny, the final order and the output's interpolation weights
  // c[6].  Returns the solve's flag (its restarts must be 0 for a replay).
  static int tape(const double* n0, double T0, double dt, const double* par,
                  const double* o, double* rec, int cap, double* ys, int ycap,
                  int* counts, double* c) {
    using CellT = kx::Cell<N, Prob, kx::AdjTape<N>>;
    Prob P;
    P.p = Net::params_from_row(par);
    P.Tfix = T0;
    P.hit_T0 = false;
    P.sc = nullptr;
    P.scs = 1;
    kx::SolveOpts so{};
    so.rtol = o[1];
    so.atol_n = o[2];
    so.atol_T = o[3];
    so.max_steps = (long)o[4];
    so.max_err_fails = 7;
    so.max_conv_fails = 10;
    so.max_restarts = 0;
    so.evolve_T = ET;
    so.jac = 0;
    so.linsol = 1;
    kx::CellOpts<N> co;
    kx::cell_opts<Net, ET>(so, co);
    std::vector<double> work(kx::cell_doubles<Prob>());
    std::vector<int> iwork(N);
    CellT cv(P, work.data(), iwork.data(), 1, co);
    cv.tp.rec = rec;
    cv.tp.ys = ys;
    cv.tp.S = 1;
    cv.tp.cap = cap;
    cv.tp.ycap = ycap;
    cv.tp.store = true;
    double y[N];
    for (int i = 0; i < Net::NSPEC; ++i) y[i] = n0[i];
    if (ET) y[N - 1] = T0;
    cv.reinit(0.0, y, 0.0);
    double t = 0.0;
    int f = cv.solve(dt, y, &t);
    counts[0] = cv.tp.nrec;
    counts[1] = cv.tp.ny;
    counts[2] = cv.tp.overflow ? 1 : 0;
    counts[3] = cv.dky_weights(dt, c);
    return f;
  }
  // the discrete scheme of a tape replayed from y0 with rates scaled by
  // exp(lnk): the same rescalings, predictors, order changes, zn[1] =
  // h f(zn[0]) and correctors (order, gamma, l) as recorded, each corrector
  // equation rl1 zn[1] + acor - gamma f(zn[0] + acor) = 0 solved by Newton
  // with the exact Jacobian to convergence; y(t_end) by the recorded
  // interpolation weights, negative densities clipped
  static int replay(const double* rec, int nrec, const double* ys,
                    const double* y0, const double* par, double T0,
                    const double* lnk, const double* c, double* yout) {
    Prob P;
    P.p = Net::params_from_row(par);
    P.Tfix = T0;
    P.hit_T0 = false;
    double sc[R];
    for (int r = 0; r < R; ++r) sc[r] = std::exp(lnk[r]);
    P.sc = sc;
    P.scs = 1;
    double Z[kx::LMAX][N] = {};
    for (int i = 0; i < N; ++i) Z[0][i] = y0[i];
    auto f = [&](const double* y, double* fv) { return P.rhs_r(y, fv); };
    for (int r = 0; r < nrec; ++r) {
      const double* x = rec + (long)r * kx::TW;
      const int op = (int)x[0], q = (int)x[1];
      if (op == kx::OP_RESCALE) {
        double cc = x[2];
        for (int j = 1; j <= q; ++j) {
          for (int i = 0; i < N; ++i) Z[j][i] *= cc;
          cc *= x[2];
        }
      } else if (op == kx::OP_PASCAL) {
        for (int k = 1; k <= q; ++k)
          for (int j = q; j >= k; --j)
            for (int i = 0; i < N; ++i) Z[j - 1][i] += x[2] * Z[j][i];
      } else if (op == kx::OP_ORDER_UP) {
        const int L = (int)x[3];
        for (int i = 0; i < N; ++i) Z[L][i] = x[2] * Z[kx::QMAX][i];
        for (int j = 2; j <= q; ++j)
          for (int i = 0; i < N; ++i) Z[j][i] += x[4 + j] * Z[L][i];
      } else if (op == kx::OP_ORDER_DOWN) {
        for (int j = 2; j < q; ++j)
          for (int i = 0; i < N; ++i) Z[j][i] -= x[4 + j] * Z[q][i];
      } else if (op == kx::OP_ZN1F) {
        double fv[N];
        if (f(Z[0], fv) != 0) return 1;
        for (int i = 0; i < N; ++i) Z[1][i] = x[2] * fv[i];
      } else if (op == kx::OP_STEP) {
        const double gam = x[2], rl1 = x[3];
        const int k = (int)x[11];
        double acor[N], y[N], fv[N], J[N * N], dl[N], Jt[N * N];
        for (int i = 0; i < N; ++i) acor[i] = ys[(long)k * N + i] - Z[0][i];
        for (int it = 0; it < 60; ++it) {
          for (int i = 0; i < N; ++i) y[i] = Z[0][i] + acor[i];
          if (f(y, fv) != 0) return 2;
          double G[N];
          for (int i = 0; i < N; ++i)
            G[i] = rl1 * Z[1][i] + acor[i] - gam * fv[i];
          // (I - gamma J) dl = -G, by the transposed solve of the transpose
          // (J exact at the first iterate, then kept: the recorded state is
          // within the tolerance of the solution)
          if (it == 0) {
            kx::jac_exact(P, y, J);
            for (int a = 0; a < N; ++a)
              for (int b = 0; b < N; ++b) Jt[a + b * N] = J[b + a * N];
          }
          double mG[N];
          for (int i = 0; i < N; ++i) mG[i] = -G[i];
          kx::solve_transposed<N>(Jt, gam, mG, dl);
          double m = 0.0, s = 0.0;
          for (int i = 0; i < N; ++i) {
            acor[i] += dl[i];
            m = std::fmax(m, std::fabs(dl[i]) /
                                 (std::fabs(Z[0][i] + acor[i]) + 1e-300));
            s = std::fmax(s, std::fabs(dl[i]));
          }
          if (m < 1e-15 || s == 0.0) break;
        }
        for (int j = 0; j <= q; ++j)
          for (int i = 0; i < N; ++i) Z[j][i] += x[4 + j] * acor[i];
        if (x[10] != 0.0)
          for (int i = 0; i < N; ++i) Z[kx::QMAX][i] = acor[i];
      }
    }
    for (int i = 0; i < N; ++i) {
      double v = 0.0;
      for (int j = kx::LMAX - 1; j >= 0; --j) v += c[j] * Z[j][i];
      yout[i] = (i < Net::NSPEC && v < 0.0) ? 0.0 : v;
    }
    return 0;
  }
  static int solve_t(int nc, const double* J, const double* gamma,
                     const double* b, double* mu) {
    for (int c = 0; c < nc; ++c)
      kx::solve_transposed<N>(J + (long)c * N * N, gamma[c], b + (long)c * N,
                              mu + (long)c * N);
    return 0;
  }
};

using C0 = Cells<kx::net_primordial, false>;
using C1 = Cells<kx::net_primordial3, true>;

}  // namespace

extern "C" {

int adj_neq(int net) { return net == 0 ? C0::N : C1::N; }
int adj_rhs(int net, int nc, const double* y, const double* Tfix,
            const double* par, const double* lnk, double* f) {
  return net == 0 ? C0::rhs(nc, y, Tfix, par, lnk, f)
                  : C1::rhs(nc, y, Tfix, par, lnk, f);
}
int adj_jac(int net, int nc, const double* y, const double* Tfix,
            const double* par, const double* lnk, double* J) {
  return net == 0 ? C0::jac(nc, y, Tfix, par, lnk, J)
                  : C1::jac(nc, y, Tfix, par, lnk, J);
}
int adj_vjp(int net, int nc, const double* y, const double* Tfix,
            const double* par, const double* lnk, const double* w,
            double* g) {
  return net == 0 ? C0::vjp(nc, y, Tfix, par, lnk, w, g)
                  : C1::vjp(nc, y, Tfix, par, lnk, w, g);
}
long adj_adjoint(int net, int nc, double* n, double* T, const double* dt,
                 const double* par, const double* lnk, int* flags,
                 const double* gn, const double* gT, double* g_lnk,
                 double* g_n0, double* g_T0, const double* o) {
  return net == 0 ? C0::adjoint(nc, n, T, dt, par, lnk, flags, gn, gT, g_lnk,
                                g_n0, g_T0, o)
                  : C1::adjoint(nc, n, T, dt, par, lnk, flags, gn, gT, g_lnk,
                                g_n0, g_T0, o);
}
int adj_tape(int net, const double* n0, double T0, double dt,
             const double* par, const double* o, double* rec, int cap,
             double* ys, int ycap, int* counts, double* c) {
  return net == 0 ? C0::tape(n0, T0, dt, par, o, rec, cap, ys, ycap, counts, c)
                  : C1::tape(n0, T0, dt, par, o, rec, cap, ys, ycap, counts,
                             c);
}
int adj_replay(int net, const double* rec, int nrec, const double* ys,
               const double* y0, const double* par, double T0,
               const double* lnk, const double* c, double* yout) {
  return net == 0 ? C0::replay(rec, nrec, ys, y0, par, T0, lnk, c, yout)
                  : C1::replay(rec, nrec, ys, y0, par, T0, lnk, c, yout);
}
// nb replays of one tape, in parallel: y0s [nb][N], lnks [nb][NREA],
// youts [nb][N]; the number that failed
int adj_replay_batch(int net, const double* rec, int nrec, const double* ys,
                     int nb, const double* y0s, const double* par, double T0,
                     const double* lnks, const double* c, double* youts) {
  const int N = adj_neq(net);
  const int R = net == 0 ? kx::net_primordial::NREA
                         : kx::net_primordial3::NREA;
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
  for (int b = 0; b < nb; ++b)
    bad += adj_replay(net, rec, nrec, ys, y0s + (long)b * N, par, T0,
                      lnks + (long)b * R, c, youts + (long)b * N) != 0;
  return bad;
}
int adj_solve_t(int net, int nc, const double* J, const double* gamma,
                const double* b, double* mu) {
  return net == 0 ? C0::solve_t(nc, J, gamma, b, mu)
                  : C1::solve_t(nc, J, gamma, b, mu);
}

}  // extern "C"
