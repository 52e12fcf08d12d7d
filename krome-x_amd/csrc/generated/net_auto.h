// generated by netgen.py from auto.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_auto {
  static constexpr const char* NAME = "auto";
  static constexpr int NSPEC = 3;
  static constexpr int NREA = 3;
  static constexpr int NPAR = 2;
  static constexpr int IDX_C = 0;
  static constexpr int IDX_Cj = 1;
  static constexpr int IDX_E = 2;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 2;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 1;
  // PHOTO heating: the energy sums sum_b sigmaE_qb Phi_b follow the
  // rate sums in the row (sigmaE: sigma times the photoelectron's
  // energy E - Emin_q, bin-averaged)
  static constexpr bool PHOTO_HEAT = false;
  static constexpr int NROW = NPAR + NPHOTO;
  // parameter row (param_names order), then the photo sums
  struct Params {
    double gamma_ad;
    double Tcmb;
    double ph[NPHOTO];
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.gamma_ad = row[0];
    p.Tcmb = row[1];
    for (int r = 0; r < NPHOTO; ++r) p.ph[r] = row[NPAR + r];
    return p;
  }
  static constexpr int photo_reaction[1] = {0};
  static constexpr int photo_reactant[1] = {0};
  static constexpr double photo_Emin[1] = {11.26};
  static constexpr double photo_Emax[1] = {INFINITY};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return (E_eV >= 291.0 ? kx::sigma_vy95(E_eV, 2.910e+02, 8.655e+01, 7.421e+01, 5.498e+01, 1.503e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.939e+01, 1.026e+01, 4.564e+03, 1.568e+00, 1.085e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.126e+01, 9.435e+00, 1.152e+03, 5.687e+00, 6.336e+00, 4.474e-01, 1) : kx::sigma_v96(E_eV, 2.144e+00, 5.027e+02, 6.216e+01, 5.101e+00, 9.157e-02, 1.133e+00, 1.607e+00));  // C -> C+ + e-
      default: return R(0) * E_eV;
    }
  }
  // index of the first REQUIRED parameter still NaN, or -1
  KX_HD static int params_missing(const double* row) {
    (void)row;
    return -1;
  }
  KX_HD static int charge(int i) {
    switch (i) {
      case 1: return 1;
      case 2: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.99447347289326e-23;
      case 1: return 1.9943823790562555e-23;
      case 2: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 1;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 1 + e) {
      case 0: return 1;
      case 1: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 9 nonzeros, 9 in L + U with the fill, 13 flops per
  // factorization (dense: 18); rows in pivot order
  static constexpr int SP_NNZ = 9;
  static constexpr int SP_FLOPS = 13;
  static constexpr int sp_perm[3] = {0, 1, 2};
  static constexpr int sp_rp[4] = {0, 3, 6, 9};
  static constexpr int sp_ci[9] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  static constexpr int sp_di[3] = {0, 4, 8};
  // conserved combinations c_k . n (the nuclei of C, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 2;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 4: return 1;
      case 5: return -1;
      default: return 0;
    }
  }
  template <class R> KX_HD static R ntot(const R* n) {
    R s = R(0);
    s += n[0];
    s += n[1];
    s += n[2];
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.99447347289326e-23);
    s += n[1] * R(1.9943823790562555e-23);
    s += n[2] * R(9.10938370046373e-28);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    k[0] = R(p.ph[0]);  // C -> C+ + e- (photo)
    k[1] = ((R(2.995e-09) / ((kx::sqrt((Tgas / R(0.00667))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.00667)))), (R(1.0) - (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1943000.0)))), (R(1.0) + (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(6.346e-09) * kx::exp(((-R(12.17)) / Tgas))) + (R(9.793e-09) * kx::exp(((-R(73.8)) / Tgas)))) + (R(1.634e-06) * kx::exp(((-R(15230.0)) / Tgas)))) + (R(0.0008369) * kx::exp(((-R(120700.0)) / Tgas)))) + (R(0.0003355) * kx::exp(((-R(214400.0)) / Tgas))))));  // C+ + e- -> C
    {  // C + e- -> C+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[2] = kx::colion_v96(Tgas, R(11.3), R(0.0), R(6.85e-08), R(0.193), R(0.25));  // C + e- -> C+ + e- + e-
    }
  }
  // density-dependent rates
  template <class R>
  KX_HD static void rates_n(R Trate, const R* n, const Params& p, R* k) {
    (void)p; (void)n;
    const R Tgas = Trate;
    (void)Tgas;
  }
  // every reaction: the T-only ones in k (zeroed) then the rest
  template <class R>
  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {
    for (int i = 0; i < NREA; ++i) k[i] = R(0);
    rates_T(Trate, p, k);
    rates_n(Trate, n, p, k);
  }
  static constexpr int NREA_T = 3;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[4] = {0, 1, 3, 5};
  static constexpr int LT_KF_SP[5] = {0, 1, 2, 0, 2};
  static constexpr int LT_SC_OFF[4] = {0, 3, 6, 11};
  static constexpr int LT_SC_ROW[11] = {0, 1, 2, 1, 2, 0, 0, 2, 1, 2, 2};
  static constexpr int LT_SC_SGN[11] = {-1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1};
  static constexpr int LT_JT_OFF[4] = {0, 1, 3, 5};
  static constexpr int LT_JT_COL[5] = {0, 1, 2, 0, 2};
  static constexpr int LT_JT_MULT[5] = {1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[6] = {0, 0, 1, 2, 3, 4};
  static constexpr int LT_JF_SP[4] = {2, 1, 2, 0};
  // the reaction fluxes k prod(n) [cm^-3 s^-1] (krome_get_fluxes)
  template <class R>
  KX_HD static void fluxes(const R* n, const R* k, R* fl) {
    for (int r = 0; r < NREA; ++r) {
      R f = k[r];
      for (int p = LT_KF_OFF[r]; p < LT_KF_OFF[r + 1]; ++p)
        f = f * n[LT_KF_SP[p]];
      fl[r] = f;
    }
  }
  // dn/dt = sum over reactions of S k prod(n), in reaction order
  template <class R>
  KX_HD static void rhs_loop(const R* n, const R* k, R* dn) {
    for (int i = 0; i < NSPEC; ++i) dn[i] = R(0);
    for (int r = 0; r < NREA; ++r) {
      R fl = k[r];
      for (int p = LT_KF_OFF[r]; p < LT_KF_OFF[r + 1]; ++p)
        fl = fl * n[LT_KF_SP[p]];
      for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q) {
        if (LT_SC_SGN[q] > 0) dn[LT_SC_ROW[q]] += fl;
        else dn[LT_SC_ROW[q]] -= fl;
      }
    }
  }
  // d(dn_i)/d(n_j) at fixed rates into J (column-major, element
  // (i, j) at J[i + j ld]; added to, the caller zeroes it)
  template <class R>
  KX_HD static void jac_species_loop(const R* n, const R* k, R* J,
                                     int ld) {
    for (int r = 0; r < NREA; ++r)
      for (int t = LT_JT_OFF[r]; t < LT_JT_OFF[r + 1]; ++t) {
        R d = k[r];
        for (int p = LT_JT_FOFF[t]; p < LT_JT_FOFF[t + 1]; ++p)
          d = d * n[LT_JF_SP[p]];
        if (LT_JT_MULT[t] > 1) d = d * R(LT_JT_MULT[t]);
        R* col = J + (long)LT_JT_COL[t] * ld;
        for (int q = LT_SC_OFF[r]; q < LT_SC_OFF[r + 1]; ++q) {
          if (LT_SC_SGN[q] > 0) col[LT_SC_ROW[q]] += d;
          else col[LT_SC_ROW[q]] -= d;
        }
      }
  }
  template <class R>
  KX_HD static void rhs_unrolled(const R* n, const R* k, R* dn) {
    for (int i = 0; i < NSPEC; ++i) dn[i] = R(0);
    R fl;
    fl = k[0]*n[0];  // C -> C+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[2] += fl;
    fl = k[1]*n[1]*n[2];  // C+ + e- -> C
    dn[1] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    fl = k[2]*n[0]*n[2];  // C + e- -> C+ + e- + e-
    dn[0] -= fl;
    dn[2] -= fl;
    dn[1] += fl;
    dn[2] += fl;
    dn[2] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[2 + 0 * ld] += d;
    d = k[1]*n[2];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[2 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[1]*n[1];  // C+ + e- -> C, d/d n[2]
    J[1 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    d = k[2]*n[2];  // C + e- -> C+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[2 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[2 + 0 * ld] += d;
    J[2 + 0 * ld] += d;
    d = k[2]*n[0];  // C + e- -> C+ + e- + e-, d/d n[2]
    J[0 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    (void)d; (void)n; (void)k; (void)J; (void)ld;
  }
#if defined(KX_FORCE_RHS_LOOP)
  static constexpr bool RHS_LOOP_USED = true;
#else
  static constexpr bool RHS_LOOP_USED = RHS_LOOP;
#endif
  template <class R>
  KX_HD static void rhs(const R* n, const R* k, R* dn) {
    if constexpr (RHS_LOOP_USED) rhs_loop(n, k, dn);
    else rhs_unrolled(n, k, dn);
  }
  template <class R>
  KX_HD static void jac_species(const R* n, const R* k, R* J, int ld) {
    if constexpr (RHS_LOOP_USED) jac_species_loop(n, k, J, ld);
    else jac_species_unrolled(n, k, J, ld);
  }
};
}  // namespace kx
