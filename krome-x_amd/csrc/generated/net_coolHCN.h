// generated by netgen.py from coolHCN.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_coolHCN {
  static constexpr const char* NAME = "coolHCN";
  static constexpr int NSPEC = 11;
  static constexpr int NREA = 7;
  static constexpr int NPAR = 3;
  static constexpr int IDX_CN = 0;
  static constexpr int IDX_CO = 1;
  static constexpr int IDX_H = 2;
  static constexpr int IDX_Hj = 3;
  static constexpr int IDX_H2 = 4;
  static constexpr int IDX_H2O = 5;
  static constexpr int IDX_HCN = 6;
  static constexpr int IDX_HCO = 7;
  static constexpr int IDX_O = 8;
  static constexpr int IDX_OH = 9;
  static constexpr int IDX_E = 10;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 10;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = true;
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
    double crate;
    double gamma_ad;
    double Tcmb;
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.crate = row[0];
    p.gamma_ad = row[1];
    p.Tcmb = row[2];
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
      case 3: return 1;
      case 10: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 4.3203905434798804e-23;
      case 1: return 4.6511699255465995e-23;
      case 2: return 1.6738233791328e-24;
      case 3: return 1.6729124407627538e-24;
      case 4: return 3.3476467582656e-24;
      case 5: return 2.9914611284799e-23;
      case 6: return 4.4877728813931596e-23;
      case 7: return 4.81855226345988e-23;
      case 8: return 2.65669645265334e-23;
      case 9: return 2.82407879056662e-23;
      case 10: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 4;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 4 + e) {
      case 0: return 1;
      case 2: return 1;
      case 4: return 1;
      case 7: return 1;
      case 9: return 1;
      case 13: return 1;
      case 17: return 2;
      case 21: return 2;
      case 23: return 1;
      case 24: return 1;
      case 25: return 1;
      case 26: return 1;
      case 28: return 1;
      case 29: return 1;
      case 31: return 1;
      case 35: return 1;
      case 37: return 1;
      case 39: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 48 nonzeros, 48 in L + U with the fill, 77 flops per
  // factorization (dense: 887); rows in pivot order
  static constexpr int SP_NNZ = 48;
  static constexpr int SP_FLOPS = 77;
  static constexpr int sp_perm[11] = {3, 10, 5, 1, 6, 0, 2, 4, 7, 8, 9};
  static constexpr int sp_rp[12] = {0, 3, 6, 8, 11, 14, 16, 27, 38, 40, 44, 48};
  static constexpr int sp_ci[48] = {0, 1, 6, 0, 1, 6, 2, 10, 3, 8, 9, 4, 5, 10, 5, 10, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 8, 9, 5, 8, 9, 10, 5, 8, 9, 10};
  static constexpr int sp_di[11] = {0, 4, 6, 8, 11, 14, 22, 34, 38, 42, 47};
  // conserved combinations c_k . n (the nuclei of C, H, N, O, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 5;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 6: return 1;
      case 7: return 1;
      case 13: return 1;
      case 14: return 1;
      case 15: return 2;
      case 16: return 2;
      case 17: return 1;
      case 18: return 1;
      case 20: return 1;
      case 22: return 1;
      case 28: return 1;
      case 34: return 1;
      case 38: return 1;
      case 40: return 1;
      case 41: return 1;
      case 42: return 1;
      case 47: return 1;
      case 54: return -1;
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
    s += n[6];
    s += n[7];
    s += n[8];
    s += n[9];
    s += n[10];
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(1.0);
    s += n[3] * R(1.0);
    s += n[4] * R(2.0);
    s += n[5] * R(2.0);
    s += n[6] * R(1.0);
    s += n[7] * R(1.0);
    s += n[8] * R(0.0);
    s += n[9] * R(1.0);
    s += n[10] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(4.3203905434798804e-23);
    s += n[1] * R(4.6511699255465995e-23);
    s += n[2] * R(1.6738233791328e-24);
    s += n[3] * R(1.6729124407627538e-24);
    s += n[4] * R(3.3476467582656e-24);
    s += n[5] * R(2.9914611284799e-23);
    s += n[6] * R(4.4877728813931596e-23);
    s += n[7] * R(4.81855226345988e-23);
    s += n[8] * R(2.65669645265334e-23);
    s += n[9] * R(2.82407879056662e-23);
    s += n[10] * R(9.10938370046373e-28);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    k[0] = (R(1e-11) * kx::exp(((-R(1000.0)) / Tgas)));  // CN + OH -> O + HCN
    k[1] = R(5e-11);  // HCO + O -> CO + OH
    k[2] = ((R(1.65e-12) * kx::pow((Tgas / R(300.0)), R(1.14))) * kx::exp(((-R(50.0)) / Tgas)));  // OH + OH -> H2O + O
    k[4] = (R(3.5e-12) * kx::pow((Tgas / R(300.0)), (-R(0.7))));  // H+ + e- -> H
    k[5] = (R(0.46) * R(p.crate));  // H -> H+ + e-
    k[6] = (R(0.1) * R(p.crate));  // H2 -> H + H
  }
  // density-dependent rates
  template <class R>
  KX_HD static void rates_n(R Trate, const R* n, const Params& p, R* k) {
    (void)p; (void)n;
    const R Tgas = Trate;
    (void)Tgas;
    {
      const R v_fA = (R(1.0) / (R(1.0) + (R(10000.0) * kx::exp(((-R(600.0)) / Tgas)))));
      k[3] = (((((R(3e-18) * kx::sqrt(Tgas)) * v_fA) / (((R(1.0) + (R(0.04) * kx::sqrt((R(2.0) * Tgas)))) + (R(0.002) * Tgas)) + (R(8e-06) * kx::ipow<2>(Tgas)))) / n[2]) * Hnuclei(n));  // H + H -> H2
    }
  }
  // every reaction: the T-only ones in k (zeroed) then the rest
  template <class R>
  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {
    for (int i = 0; i < NREA; ++i) k[i] = R(0);
    rates_T(Trate, p, k);
    rates_n(Trate, n, p, k);
  }
  static constexpr int NREA_T = 6;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[8] = {0, 2, 4, 6, 8, 10, 11, 12};
  static constexpr int LT_KF_SP[12] = {0, 9, 7, 8, 9, 9, 2, 2, 3, 10, 2, 4};
  static constexpr int LT_SC_OFF[8] = {0, 4, 8, 12, 15, 18, 21, 24};
  static constexpr int LT_SC_ROW[24] = {0, 9, 8, 6, 7, 8, 1, 9, 9, 9, 5, 8, 2, 2, 4, 3, 10, 2, 2, 3, 10, 4, 2, 2};
  static constexpr int LT_SC_SGN[24] = {-1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, 1, 1, -1, 1, 1};
  static constexpr int LT_JT_OFF[8] = {0, 2, 4, 5, 6, 8, 9, 10};
  static constexpr int LT_JT_COL[10] = {0, 9, 7, 8, 9, 2, 3, 10, 2, 4};
  static constexpr int LT_JT_MULT[10] = {1, 1, 1, 1, 2, 2, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 8, 8};
  static constexpr int LT_JF_SP[8] = {9, 0, 8, 7, 9, 2, 10, 3};
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
    fl = k[0]*n[0]*n[9];  // CN + OH -> O + HCN
    dn[0] -= fl;
    dn[9] -= fl;
    dn[8] += fl;
    dn[6] += fl;
    fl = k[1]*n[7]*n[8];  // HCO + O -> CO + OH
    dn[7] -= fl;
    dn[8] -= fl;
    dn[1] += fl;
    dn[9] += fl;
    fl = k[2]*n[9]*n[9];  // OH + OH -> H2O + O
    dn[9] -= fl;
    dn[9] -= fl;
    dn[5] += fl;
    dn[8] += fl;
    fl = k[3]*n[2]*n[2];  // H + H -> H2
    dn[2] -= fl;
    dn[2] -= fl;
    dn[4] += fl;
    fl = k[4]*n[3]*n[10];  // H+ + e- -> H
    dn[3] -= fl;
    dn[10] -= fl;
    dn[2] += fl;
    fl = k[5]*n[2];  // H -> H+ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[10] += fl;
    fl = k[6]*n[4];  // H2 -> H + H
    dn[4] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[9];  // CN + OH -> O + HCN, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[9 + 0 * ld] -= d;
    J[8 + 0 * ld] += d;
    J[6 + 0 * ld] += d;
    d = k[0]*n[0];  // CN + OH -> O + HCN, d/d n[9]
    J[0 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[8 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[1]*n[8];  // HCO + O -> CO + OH, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[8 + 7 * ld] -= d;
    J[1 + 7 * ld] += d;
    J[9 + 7 * ld] += d;
    d = k[1]*n[7];  // HCO + O -> CO + OH, d/d n[8]
    J[7 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[1 + 8 * ld] += d;
    J[9 + 8 * ld] += d;
    d = k[2]*n[9]*R(2.0);  // OH + OH -> H2O + O, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[5 + 9 * ld] += d;
    J[8 + 9 * ld] += d;
    d = k[3]*n[2]*R(2.0);  // H + H -> H2, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    d = k[4]*n[10];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[10 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[4]*n[3];  // H+ + e- -> H, d/d n[10]
    J[3 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[2 + 10 * ld] += d;
    d = k[5];  // H -> H+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[6];  // H2 -> H + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
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
