// generated by netgen.py from SM.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_SM {
  static constexpr const char* NAME = "SM";
  static constexpr int NSPEC = 6;
  static constexpr int NREA = 12;
  static constexpr int NPAR = 2;
  static constexpr int IDX_H = 0;
  static constexpr int IDX_H2 = 1;
  static constexpr int IDX_H2O = 2;
  static constexpr int IDX_O = 3;
  static constexpr int IDX_O2 = 4;
  static constexpr int IDX_OH = 5;
  static constexpr bool HAS_E = false;
  static constexpr int I_E = -1;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 0;
  // PHOTO heating: the energy sums sum_b sigmaE_qb Phi_b follow the
  // rate sums in the row (sigmaE: sigma times the photoelectron's
  // energy E - Emin_q, bin-averaged)
  static constexpr bool PHOTO_HEAT = false;
  static constexpr int NROW = NPAR + NPHOTO;
  // parameter row (param_names order), then the photo sums
  struct Params {
    double gamma_ad;
    double Tcmb;
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.gamma_ad = row[0];
    p.Tcmb = row[1];
    return p;
  }
  static constexpr int photo_reaction[1] = {0};
  static constexpr int photo_reactant[1] = {0};
  static constexpr double photo_Emin[1] = {0.0};
  static constexpr double photo_Emax[1] = {0.0};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
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
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.6738233791328e-24;
      case 1: return 3.3476467582656e-24;
      case 2: return 2.9914611284799e-23;
      case 3: return 2.65669645265334e-23;
      case 4: return 5.31339290530668e-23;
      case 5: return 2.82407879056662e-23;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 2;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 2 + e) {
      case 0: return 1;
      case 2: return 2;
      case 4: return 2;
      case 5: return 1;
      case 7: return 1;
      case 9: return 2;
      case 10: return 1;
      case 11: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 32 nonzeros, 32 in L + U with the fill, 91 flops per
  // factorization (dense: 144); rows in pivot order
  static constexpr int SP_NNZ = 32;
  static constexpr int SP_FLOPS = 91;
  static constexpr int sp_perm[6] = {4, 0, 1, 2, 3, 5};
  static constexpr int sp_rp[7] = {0, 4, 10, 15, 20, 26, 32};
  static constexpr int sp_ci[32] = {0, 1, 4, 5, 0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 0, 1, 2, 3, 4, 5, 0, 1, 2, 3, 4, 5};
  static constexpr int sp_di[6] = {0, 5, 11, 17, 24, 31};
  // conserved combinations c_k . n (the nuclei of H, O): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 2;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 2;
      case 2: return 2;
      case 5: return 1;
      case 8: return 1;
      case 9: return 1;
      case 10: return 2;
      case 11: return 1;
      default: return 0;
    }
  }
  template <class R> KX_HD static R ntot(const R* n) {
    R s = R(0);
    s += n[0];
    s += n[1];
    s += n[2];
    s += n[3];
    s += n[4];
    s += n[5];
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(1.0);
    s += n[1] * R(2.0);
    s += n[2] * R(2.0);
    s += n[3] * R(0.0);
    s += n[4] * R(0.0);
    s += n[5] * R(1.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.6738233791328e-24);
    s += n[1] * R(3.3476467582656e-24);
    s += n[2] * R(2.9914611284799e-23);
    s += n[3] * R(2.65669645265334e-23);
    s += n[4] * R(5.31339290530668e-23);
    s += n[5] * R(2.82407879056662e-23);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    k[0] = R(2.0);  // H2 -> H + H
    k[1] = R(216.0);  // H + H -> H2
    k[2] = R(1.0);  // O2 -> O + O
    k[3] = R(337.5);  // O + O -> O2
    k[4] = R(1.0);  // H2O -> H + OH
    k[5] = R(1400.0);  // H + OH -> H2O
    k[6] = R(1000.0);  // H2 + O -> H + OH
    k[7] = R(10800.0);  // OH + H -> H2 + O
    k[8] = R(1000.0);  // O2 + H -> O + OH
    k[9] = R(33750.0);  // OH + O -> O2 + H
    k[10] = R(100.0);  // H2 + O -> H2O
    k[11] = R(0.7714);  // H2O -> H2 + O
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
  static constexpr int NREA_T = 12;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[13] = {0, 1, 3, 4, 6, 7, 9, 11, 13, 15, 17, 19, 20};
  static constexpr int LT_KF_SP[20] = {1, 0, 0, 4, 3, 3, 2, 0, 5, 1, 3, 5, 0, 4, 0, 5, 3, 1, 3, 2};
  static constexpr int LT_SC_OFF[13] = {0, 3, 6, 9, 12, 15, 18, 22, 26, 30, 34, 37, 40};
  static constexpr int LT_SC_ROW[40] = {1, 0, 0, 0, 0, 1, 4, 3, 3, 3, 3, 4, 2, 0, 5, 0, 5, 2, 1, 3, 0, 5, 5, 0, 1, 3, 4, 0, 3, 5, 5, 3, 4, 0, 1, 3, 2, 2, 1, 3};
  static constexpr int LT_SC_SGN[40] = {-1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1, 1};
  static constexpr int LT_JT_OFF[13] = {0, 1, 2, 3, 4, 5, 7, 9, 11, 13, 15, 17, 18};
  static constexpr int LT_JT_COL[18] = {1, 0, 4, 3, 2, 0, 5, 1, 3, 0, 5, 0, 4, 3, 5, 1, 3, 2};
  static constexpr int LT_JT_MULT[18] = {1, 2, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[19] = {0, 0, 1, 1, 2, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 14};
  static constexpr int LT_JF_SP[14] = {0, 3, 5, 0, 3, 1, 5, 0, 4, 0, 5, 3, 3, 1};
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
    fl = k[0]*n[1];  // H2 -> H + H
    dn[1] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[1]*n[0]*n[0];  // H + H -> H2
    dn[0] -= fl;
    dn[0] -= fl;
    dn[1] += fl;
    fl = k[2]*n[4];  // O2 -> O + O
    dn[4] -= fl;
    dn[3] += fl;
    dn[3] += fl;
    fl = k[3]*n[3]*n[3];  // O + O -> O2
    dn[3] -= fl;
    dn[3] -= fl;
    dn[4] += fl;
    fl = k[4]*n[2];  // H2O -> H + OH
    dn[2] -= fl;
    dn[0] += fl;
    dn[5] += fl;
    fl = k[5]*n[0]*n[5];  // H + OH -> H2O
    dn[0] -= fl;
    dn[5] -= fl;
    dn[2] += fl;
    fl = k[6]*n[1]*n[3];  // H2 + O -> H + OH
    dn[1] -= fl;
    dn[3] -= fl;
    dn[0] += fl;
    dn[5] += fl;
    fl = k[7]*n[5]*n[0];  // OH + H -> H2 + O
    dn[5] -= fl;
    dn[0] -= fl;
    dn[1] += fl;
    dn[3] += fl;
    fl = k[8]*n[4]*n[0];  // O2 + H -> O + OH
    dn[4] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[5] += fl;
    fl = k[9]*n[5]*n[3];  // OH + O -> O2 + H
    dn[5] -= fl;
    dn[3] -= fl;
    dn[4] += fl;
    dn[0] += fl;
    fl = k[10]*n[1]*n[3];  // H2 + O -> H2O
    dn[1] -= fl;
    dn[3] -= fl;
    dn[2] += fl;
    fl = k[11]*n[2];  // H2O -> H2 + O
    dn[2] -= fl;
    dn[1] += fl;
    dn[3] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0];  // H2 -> H + H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[1]*n[0]*R(2.0);  // H + H -> H2, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    d = k[2];  // O2 -> O + O, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[3 + 4 * ld] += d;
    d = k[3]*n[3]*R(2.0);  // O + O -> O2, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    d = k[4];  // H2O -> H + OH, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[5 + 2 * ld] += d;
    d = k[5]*n[5];  // H + OH -> H2O, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[5 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    d = k[5]*n[0];  // H + OH -> H2O, d/d n[5]
    J[0 + 5 * ld] -= d;
    J[5 + 5 * ld] -= d;
    J[2 + 5 * ld] += d;
    d = k[6]*n[3];  // H2 + O -> H + OH, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[3 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[5 + 1 * ld] += d;
    d = k[6]*n[1];  // H2 + O -> H + OH, d/d n[3]
    J[1 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[5 + 3 * ld] += d;
    d = k[7]*n[5];  // OH + H -> H2 + O, d/d n[0]
    J[5 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[3 + 0 * ld] += d;
    d = k[7]*n[0];  // OH + H -> H2 + O, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[0 + 5 * ld] -= d;
    J[1 + 5 * ld] += d;
    J[3 + 5 * ld] += d;
    d = k[8]*n[4];  // O2 + H -> O + OH, d/d n[0]
    J[4 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[5 + 0 * ld] += d;
    d = k[8]*n[0];  // O2 + H -> O + OH, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[0 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[5 + 4 * ld] += d;
    d = k[9]*n[5];  // OH + O -> O2 + H, d/d n[3]
    J[5 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[9]*n[3];  // OH + O -> O2 + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[3 + 5 * ld] -= d;
    J[4 + 5 * ld] += d;
    J[0 + 5 * ld] += d;
    d = k[10]*n[3];  // H2 + O -> H2O, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[3 + 1 * ld] -= d;
    J[2 + 1 * ld] += d;
    d = k[10]*n[1];  // H2 + O -> H2O, d/d n[3]
    J[1 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[11];  // H2O -> H2 + O, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
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
