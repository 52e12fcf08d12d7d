// Fixture writer over a KROME-X project generated by the reference's own
// code generator (tests/golden/make_golden.py runs the generator in this
// container and compiles this file against its output with g++).  It calls
// the reference's public C++ API only: krome::solve_cells (the PER-CELL
// OpenMP path, runtime/core/solve_cells.hpp:47-118) for whole solves, and
// the per-cell building blocks of solve_cell.hpp (init_ctx_, full_rhs_,
// jac_host_) plus the generated rates / thermal functions for the
// function-level vectors.  Nothing of it is part of the product.
//
//   ref_driver solve IN OUT     IN: i64 ncell, nspec, npar, nopt; f64
//                               opts[nopt]; n0[ncell nspec]; T[ncell];
//                               dt[ncell]; par[ncell npar].  OUT: f64
//                               n[ncell nspec], T[ncell], flag[ncell],
//                               stats[ncell 9], seconds
//   ref_driver funcs IN OUT     IN as above (dt unused).  OUT per sample:
//                               k[NREA], f[NEQ] (full rhs), J[NEQ^2]
//                               (column-major, the solver's analytic
//                               Jacobian), and the thermal terms
//                               (8 values, 0 when not configured)
//   ref_driver totals IN OUT    IN as above (n0 the cells).  OUT:
//                               krome::get_element_totals of each row
//                               [ncell NELEM], then get_element_totals_x
//                               of the same rows read as mass fractions
//   ref_driver conserve IN OUT  IN as above, opts[0] 0 (n0 densities) or
//                               1 (mass fractions), followed by
//                               ref[ncell NELEM].  OUT: krome::
//                               conserve_cells(_x) (ref_stride 1, the
//                               cells' params): rows[ncell nspec],
//                               flag[ncell]
//   ref_driver meta IN OUT      OUT: NELEM, species_elements[NSPEC NELEM],
//                               species_mass_g[NSPEC], species_charge
//                               [NSPEC] (f64); OUT.names: the element
//                               names, one a line
//   ref_driver sparse IN OUT    OUT: SPARSE_NNZ, SPARSE_FLOPS, sp_perm
//                               [NSPEC], sp_rp[NSPEC + 1], sp_ci[NNZ],
//                               sp_di[NSPEC] (f64)
//   (solve: opts[4] 1 dense, 3 sparse, else auto; opts[7], when given,
//   is SolverOpts::conserve; with 8 + 2 NSPEC options, opts[8..] are the
//   per-species atol then rtol, SolverOpts::atol_species / rtol_species)
//   ref_driver pbins|prates|psolve IN OUT   photon bins (networks with
//                               photo reactions): IN as above followed by
//                               i64 B; f64 alpha, Elo[B], Ehi[B],
//                               Phi[ncell B]; i64 NE; f64 E[NE].  pbins
//                               OUT: the PhotoModule's bin-averaged
//                               sigma[NPHOTO B], sigmaE[NPHOTO B],
//                               photo_xsec[r](E)
//                               [NPHOTO NE], NPHOTO, photo_reaction,
//                               photo_reactant, photo_Emin, photo_Emax;
//                               prates OUT per cell: krome::get_rates
//                               with the cell's field, k[NREA], then
//                               krome::get_thermal's heat, cool, dTdt,
//                               heats[12], cools[19] (the active terms
//                               first, zeros after) and get_photo_heating
//                               [NPHOTO]; psolve OUT:
//                               as solve, with the bins and Phi
#include "krome.hpp"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct In {
  int64_t ncell = 0, nspec = 0, npar = 0, nopt = 0;
  std::vector<double> opts, n0, T, dt, par;
};

bool read_in(const char* path, In& in) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int64_t h[4];
  if (std::fread(h, sizeof(int64_t), 4, f) != 4) return false;
  in.ncell = h[0];
  in.nspec = h[1];
  in.npar = h[2];
  in.nopt = h[3];
  auto rd = [&](std::vector<double>& v, size_t m) {
    v.resize(m);
    return std::fread(v.data(), sizeof(double), m, f) == m;
  };
  bool ok = rd(in.opts, in.nopt) && rd(in.n0, in.ncell * in.nspec) &&
            rd(in.T, in.ncell) && rd(in.dt, in.ncell) &&
            rd(in.par, in.ncell * in.npar);
  std::fclose(f);
  return ok;
}

// opts: evolve_T, rtol, atol_n, atol_T, linsol, jac, max_steps
krome::SolverOpts solver_opts(const In& in) {
  krome::SolverOpts so;
  so.evolve_T = in.opts[0] != 0.0;
  so.rtol = in.opts[1];
  so.atol_n = in.opts[2];
  so.atol_T = in.opts[3];
  const int ls = static_cast<int>(in.opts[4]);
  so.linsol = ls == 1   ? krome::LINSOL_DENSE
              : ls == 3 ? krome::LINSOL_SPARSE
                        : krome::LINSOL_AUTO;
  const int jc = static_cast<int>(in.opts[5]);
  so.jac = jc == 1 ? krome::JAC_ANALYTIC
                   : (jc == 2 ? krome::JAC_FD : krome::JAC_AUTO);
  so.max_steps = static_cast<long>(in.opts[6]);
  if (in.nopt > 7) so.conserve = static_cast<int>(in.opts[7]);
  // opts[8 .. 8 + NSPEC) per-species atol, then NSPEC per-species rtol
  // (SolverOpts::atol_species / rtol_species; negative: the scalar)
  if (in.nopt >= 8 + 2 * krome::NSPEC) {
    so.atol_species = in.opts.data() + 8;
    so.rtol_species = in.opts.data() + 8 + krome::NSPEC;
  }
  so.error_report = false;
  return so;
}

int do_solve(const In& in, const char* out) {
  if (in.nspec != krome::NSPEC || in.npar != krome::NPAR) return 2;
  krome::SolverOpts so = solver_opts(in);
  krome::CellsOpts co;
  co.path = krome::PATH_PERCELL;
  std::vector<int> flags(in.ncell, 0);
  std::vector<krome::SolveStats> st(in.ncell);
  co.flags = flags.data();
  co.stats = st.data();
  std::vector<krome::Params> P(in.ncell);
  for (int64_t c = 0; c < in.ncell; ++c)
    krome::params_from_row(in.par.data() + c * in.npar, P[c]);
  std::vector<double> n = in.n0, T = in.T;
  const auto t0 = std::chrono::steady_clock::now();
  krome::solve_cells(static_cast<int>(in.ncell), n.data(), T.data(),
                     in.dt.data(), P.data(), so, co);
  const double secs = std::chrono::duration<double>(
                          std::chrono::steady_clock::now() - t0).count();
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::fwrite(n.data(), sizeof(double), n.size(), f);
  std::fwrite(T.data(), sizeof(double), T.size(), f);
  for (int64_t c = 0; c < in.ncell; ++c) {
    const double fl = flags[c];
    std::fwrite(&fl, sizeof(double), 1, f);
  }
  for (int64_t c = 0; c < in.ncell; ++c) {
    const krome::SolveStats& s = st[c];
    const double v[9] = {double(s.steps), double(s.rhs_evals),
                         double(s.jac_evals), double(s.rhs_evals_jac),
                         double(s.lin_setups), double(s.newton_iters),
                         double(s.newton_fails), double(s.err_fails),
                         double(s.restarts)};
    std::fwrite(v, sizeof(double), 9, f);
  }
  std::fwrite(&secs, sizeof(double), 1, f);
  std::fclose(f);
  return 0;
}

// the configured thermal terms at the thermal temperature: H2 cooling
// (with its opacity factor), continuum, chemical cooling, chemical
// heating, compression heating, gamma, dT/dt, the H2 collisional
// dissociation rate at the rate temperature
template <class P>
void thermal_terms(double Tth, double Trate, const double* n, const double* k,
                   const P& p, krome::ThermCtx* tc, double* o) {
  using TC = krome::ThermalConfig;
  for (int i = 0; i < 8; ++i) o[i] = 0.0;
  if constexpr (TC::cool_H2)
    o[0] = krome::cooling_H2(Tth, n) * krome::h2_opacity_(Tth, n, p);
  if constexpr (TC::cool_cont) o[1] = krome::cooling_Continuum(Tth, n);
  if constexpr (TC::cool_chem) o[2] = krome::cooling_chem(Tth, n, k, p);
  if constexpr (TC::heat_chem) o[3] = krome::heating_chem(Tth, n, k, p);
  if constexpr (TC::heat_compress) o[4] = krome::heat_compress_(Tth, n, p);
  o[5] = krome::gamma_index(Tth, n, p);
  double dnx[krome::NSPEC];
  o[6] = krome::thermal_aux(Tth, n, k, tc, dnx);
#ifdef KROME_EXTERN_dissH2_Martin96
  o[7] = krome::dissH2_Martin96(Trate, n, p);
#else
  (void)Trate;
#endif
}

int do_funcs(const In& in, const char* out) {
  if (in.nspec != krome::NSPEC || in.npar != krome::NPAR) return 2;
  krome::SolverOpts so = solver_opts(in);
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  for (int64_t c = 0; c < in.ncell; ++c) {
    krome::Params p;
    krome::params_from_row(in.par.data() + c * in.npar, p);
    const double* n = in.n0.data() + c * in.nspec;
    krome::SolveCtx ctx;
    if (krome::init_ctx_(ctx, in.T[c], p, so) != 0) return 4;
    krome::install_hooks_(ctx, n, so);
    const int NEQ = ctx.NEQ;
    std::vector<double> y(NEQ), fv(NEQ), J(NEQ * NEQ), t1(NEQ), t2(NEQ);
    krome::pack_state_(ctx, n, in.T[c], so, y.data());
    double k[krome::NREA + 1];
    const double Trate = krome::state_T_(y.data(), &ctx);
    krome::rates(Trate, y.data(), ctx.tc.p, k);
    krome::full_rhs_(y.data(), fv.data(), &ctx);
    krome::jac_host_(y.data(), krome::DenseMat_(J.data(), NEQ), &ctx,
                     t1.data(), t2.data());
    double th[8];
    krome::rates(Trate, y.data(), ctx.tc.p, k);
    thermal_terms(krome::thermal_T_(in.T[c]), Trate, y.data(), k, ctx.tc.p,
                  &ctx.tc, th);
    std::fwrite(k, sizeof(double), krome::NREA, f);
    std::fwrite(fv.data(), sizeof(double), NEQ, f);
    std::fwrite(J.data(), sizeof(double), NEQ * NEQ, f);
    std::fwrite(th, sizeof(double), 8, f);
  }
  std::fclose(f);
  return 0;
}

int do_meta(const char* out) {
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::vector<double> v;
  v.push_back(krome::NELEM);
  for (int i = 0; i < krome::NSPEC * krome::NELEM; ++i)
    v.push_back(species_elements[i]);
  for (int i = 0; i < krome::NSPEC; ++i) v.push_back(species_mass_g[i]);
  for (int i = 0; i < krome::NSPEC; ++i) v.push_back(species_charge[i]);
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  const std::string nm = std::string(out) + ".names";
  f = std::fopen(nm.c_str(), "w");
  if (!f) return 3;
  for (int e = 0; e < krome::NELEM; ++e)
    std::fprintf(f, "%s\n", krome::element_names[e]);
  std::fclose(f);
  return 0;
}

int do_sparse(const char* out) {
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::vector<double> v = {double(krome::SPARSE_NNZ),
                           double(krome::SPARSE_FLOPS)};
  for (int i = 0; i < krome::NSPEC; ++i) v.push_back(sp_perm[i]);
  for (int i = 0; i <= krome::NSPEC; ++i) v.push_back(sp_rp[i]);
  for (int i = 0; i < krome::SPARSE_NNZ; ++i) v.push_back(sp_ci[i]);
  for (int i = 0; i < krome::NSPEC; ++i) v.push_back(sp_di[i]);
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return 0;
}

int do_totals(const In& in, const char* out) {
  if (in.nspec != krome::NSPEC) return 2;
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::vector<double> tot(krome::NELEM > 0 ? krome::NELEM : 1);
  for (int64_t c = 0; c < in.ncell; ++c) {
    krome::get_element_totals(in.n0.data() + c * in.nspec, tot.data());
    std::fwrite(tot.data(), sizeof(double), krome::NELEM, f);
  }
  for (int64_t c = 0; c < in.ncell; ++c) {
    krome::get_element_totals_x(in.n0.data() + c * in.nspec, tot.data());
    std::fwrite(tot.data(), sizeof(double), krome::NELEM, f);
  }
  std::fclose(f);
  return 0;
}

// the photon bins and fields after the standard input
struct PhotoIn {
  int64_t B = 0, NE = 0;
  double alpha = 1.0;
  std::vector<double> Elo, Ehi, Phi, E;
};

bool read_photo(const char* path, const In& in, PhotoIn& ph) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  const long skip = sizeof(int64_t) * 4 +
      sizeof(double) * (in.nopt + in.ncell * in.nspec + 2 * in.ncell +
                        in.ncell * in.npar);
  bool ok = std::fseek(f, skip, SEEK_SET) == 0 &&
            std::fread(&ph.B, sizeof(int64_t), 1, f) == 1 &&
            std::fread(&ph.alpha, sizeof(double), 1, f) == 1;
  auto rd = [&](std::vector<double>& v, size_t m) {
    v.resize(m);
    return std::fread(v.data(), sizeof(double), m, f) == m;
  };
  ok = ok && rd(ph.Elo, ph.B) && rd(ph.Ehi, ph.B) &&
       rd(ph.Phi, in.ncell * ph.B) &&
       std::fread(&ph.NE, sizeof(int64_t), 1, f) == 1 && rd(ph.E, ph.NE);
  std::fclose(f);
  return ok;
}

krome::PhotoModule photo_module(const PhotoIn& ph) {
  krome::Radiation rad;
  rad.Elo = ph.Elo;
  rad.Ehi = ph.Ehi;
  rad.Phi.assign(ph.B, 0.0);
  rad.alpha = ph.alpha;
  krome::PhotoModule pm;
  pm.init(rad);
  return pm;
}

int do_photo(const char* mode, const In& in, const PhotoIn& ph,
             const char* out) {
  if (in.nspec != krome::NSPEC || in.npar != krome::NPAR) return 2;
  const krome::PhotoModule pm = photo_module(ph);
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::vector<krome::Params> P(in.ncell);
  for (int64_t c = 0; c < in.ncell; ++c)
    krome::params_from_row(in.par.data() + c * in.npar, P[c]);
  if (!std::strcmp(mode, "pbins")) {
    std::fwrite(pm.sigma.data(), sizeof(double), pm.sigma.size(), f);
    std::fwrite(pm.sigmaE.data(), sizeof(double), pm.sigmaE.size(), f);
    for (int r = 0; r < krome::NPHOTO; ++r)
      for (int64_t e = 0; e < ph.NE; ++e) {
        const double v = krome::photo_xsec[r](ph.E[e]);
        std::fwrite(&v, sizeof(double), 1, f);
      }
    std::vector<double> v = {double(krome::NPHOTO)};
    for (int r = 0; r < krome::NPHOTO; ++r) v.push_back(photo_reaction[r]);
    for (int r = 0; r < krome::NPHOTO; ++r) v.push_back(photo_reactant[r]);
    for (int r = 0; r < krome::NPHOTO; ++r) v.push_back(photo_Emin[r]);
    for (int r = 0; r < krome::NPHOTO; ++r) v.push_back(photo_Emax[r]);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
  } else if (!std::strcmp(mode, "prates")) {
    krome::SolverOpts so = solver_opts(in);
    so.photo = &pm;
    std::vector<double> phi(ph.B), k(krome::NREA);
    for (int64_t c = 0; c < in.ncell; ++c) {
      for (int64_t b = 0; b < ph.B; ++b) phi[b] = ph.Phi[c * ph.B + b];
      so.Phi = phi.data();
      const double* nc = in.n0.data() + c * in.nspec;
      if (krome::get_rates(nc, in.T[c], P[c], so, k.data()) != 0) return 4;
      std::fwrite(k.data(), sizeof(double), krome::NREA, f);
      double th[3 + 12 + 19] = {0.0};
      if (krome::get_thermal(nc, in.T[c], P[c], so, th, th + 1, th + 3,
                             th + 15, th + 2) != 0)
        return 5;
      std::fwrite(th, sizeof(double), 3 + 12 + 19, f);
      std::vector<double> ph(krome::NPHOTO > 0 ? krome::NPHOTO : 1, 0.0);
      if (krome::get_photo_heating(nc, in.T[c], P[c], so, ph.data()) != 0)
        return 6;
      std::fwrite(ph.data(), sizeof(double), krome::NPHOTO, f);
    }
  } else {
    krome::SolverOpts so = solver_opts(in);
    so.photo = &pm;
    krome::CellsOpts co;
    co.path = krome::PATH_PERCELL;
    std::vector<double> Phi = ph.Phi;
    co.Phi = Phi.data();
    std::vector<int> flags(in.ncell, 0);
    std::vector<krome::SolveStats> st(in.ncell);
    co.flags = flags.data();
    co.stats = st.data();
    std::vector<double> n = in.n0, T = in.T;
    const auto t0 = std::chrono::steady_clock::now();
    krome::solve_cells(static_cast<int>(in.ncell), n.data(), T.data(),
                       in.dt.data(), P.data(), so, co);
    const double secs = std::chrono::duration<double>(
                            std::chrono::steady_clock::now() - t0).count();
    std::fwrite(n.data(), sizeof(double), n.size(), f);
    std::fwrite(T.data(), sizeof(double), T.size(), f);
    for (int64_t c = 0; c < in.ncell; ++c) {
      const double fl = flags[c];
      std::fwrite(&fl, sizeof(double), 1, f);
    }
    for (int64_t c = 0; c < in.ncell; ++c) {
      const krome::SolveStats& s = st[c];
      const double v[9] = {double(s.steps), double(s.rhs_evals),
                           double(s.jac_evals), double(s.rhs_evals_jac),
                           double(s.lin_setups), double(s.newton_iters),
                           double(s.newton_fails), double(s.err_fails),
                           double(s.restarts)};
      std::fwrite(v, sizeof(double), 9, f);
    }
    std::fwrite(&secs, sizeof(double), 1, f);
  }
  std::fclose(f);
  return 0;
}

int do_conserve(const In& in, const std::vector<double>& ref,
                const char* out) {
  if (in.nspec != krome::NSPEC || in.npar != krome::NPAR) return 2;
  std::vector<krome::Params> P(in.ncell);
  for (int64_t c = 0; c < in.ncell; ++c)
    krome::params_from_row(in.par.data() + c * in.npar, P[c]);
  std::vector<double> v = in.n0;
  std::vector<int> flags(in.ncell, 0);
  const bool x = in.opts[0] != 0.0;
  if (x)
    krome::conserve_cells_x(static_cast<int>(in.ncell), v.data(), ref.data(),
                            1, P.data(), 1, flags.data());
  else
    krome::conserve_cells(static_cast<int>(in.ncell), v.data(), ref.data(),
                          1, P.data(), 1, flags.data());
  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  for (int64_t c = 0; c < in.ncell; ++c) {
    const double fl = flags[c];
    std::fwrite(&fl, sizeof(double), 1, f);
  }
  std::fclose(f);
  return 0;
}

// the parameter row's field names (krome::param_names), one a line
int do_params(const char* out) {
  FILE* f = std::fopen(out, "w");
  if (!f) return 3;
  for (int i = 0; i < krome::NPAR; ++i)
    std::fprintf(f, "%s\n", krome::param_names[i]);
  std::fclose(f);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr,
                 "usage: %s solve|funcs|totals|conserve|meta|sparse|params "
                 "IN OUT\n",
                 argv[0]);
    return 1;
  }
  if (!std::strcmp(argv[1], "params")) return do_params(argv[3]);
  In in;
  if (!read_in(argv[2], in)) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 1;
  }
  if (!std::strcmp(argv[1], "solve")) return do_solve(in, argv[3]);
  if (!std::strcmp(argv[1], "funcs")) return do_funcs(in, argv[3]);
  if (!std::strcmp(argv[1], "totals")) return do_totals(in, argv[3]);
  if (!std::strcmp(argv[1], "meta")) return do_meta(argv[3]);
  if (!std::strcmp(argv[1], "sparse")) return do_sparse(argv[3]);
  if (!std::strcmp(argv[1], "pbins") || !std::strcmp(argv[1], "prates") ||
      !std::strcmp(argv[1], "psolve")) {
    PhotoIn ph;
    if (!read_photo(argv[2], in, ph)) return 1;
    return do_photo(argv[1], in, ph, argv[3]);
  }
  if (!std::strcmp(argv[1], "conserve")) {
    // ref[ncell NELEM] follows the standard input in the same file
    FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 1;
    const long skip = 8 * (4 + in.nopt + in.ncell * in.nspec + 2 * in.ncell +
                           in.ncell * in.npar);
    std::fseek(f, skip, SEEK_SET);
    std::vector<double> ref(static_cast<size_t>(in.ncell) * krome::NELEM);
    const size_t m = std::fread(ref.data(), sizeof(double), ref.size(), f);
    std::fclose(f);
    if (m != ref.size()) return 1;
    return do_conserve(in, ref, argv[3]);
  }
  return 1;
}
