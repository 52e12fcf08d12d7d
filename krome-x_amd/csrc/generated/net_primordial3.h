// generated by netgen.py from primordial3.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_primordial3 {
  static constexpr const char* NAME = "primordial3";
  static constexpr int NSPEC = 9;
  static constexpr int NREA = 27;
  static constexpr int NPAR = 3;
  static constexpr int IDX_H = 0;
  static constexpr int IDX_Hj = 1;
  static constexpr int IDX_Hk = 2;
  static constexpr int IDX_H2 = 3;
  static constexpr int IDX_H2j = 4;
  static constexpr int IDX_He = 5;
  static constexpr int IDX_Hej = 6;
  static constexpr int IDX_Hejj = 7;
  static constexpr int IDX_E = 8;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 8;
  static constexpr bool THERMAL = true;
  static constexpr int NHEAT = 2;
  static constexpr int NCOOL = 3;
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
    double Tcmb;
    double tff;
    double opratio;
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.Tcmb = row[0];
    p.tff = row[1];
    p.opratio = row[2];
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
    if (row[1] != row[1]) return 1;
    (void)row;
    return -1;
  }
  KX_HD static int charge(int i) {
    switch (i) {
      case 1: return 1;
      case 2: return -1;
      case 4: return 1;
      case 6: return 1;
      case 7: return 2;
      case 8: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.6738233791328e-24;
      case 1: return 1.6729124407627538e-24;
      case 2: return 1.6747343175028463e-24;
      case 3: return 3.3476467582656e-24;
      case 4: return 3.346735819895553e-24;
      case 5: return 6.64647366797316e-24;
      case 6: return 6.645562729603113e-24;
      case 7: return 6.644651791233067e-24;
      case 8: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 2;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 2 + e) {
      case 0: return 1;
      case 2: return 1;
      case 4: return 1;
      case 6: return 2;
      case 8: return 2;
      case 11: return 1;
      case 13: return 1;
      case 15: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 55 nonzeros, 55 in L + U with the fill, 174 flops per
  // factorization (dense: 486); rows in pivot order
  static constexpr int SP_NNZ = 55;
  static constexpr int SP_FLOPS = 174;
  static constexpr int sp_perm[9] = {5, 6, 7, 0, 1, 2, 3, 4, 8};
  static constexpr int sp_rp[10] = {0, 3, 7, 10, 19, 25, 31, 40, 46, 55};
  static constexpr int sp_ci[55] = {0, 1, 8, 0, 1, 2, 8, 1, 2, 8, 0, 1, 2, 3, 4, 5, 6, 7, 8, 3, 4, 5, 6, 7, 8, 3, 4, 5, 6, 7, 8, 0, 1, 2, 3, 4, 5, 6, 7, 8, 3, 4, 5, 6, 7, 8, 0, 1, 2, 3, 4, 5, 6, 7, 8};
  static constexpr int sp_di[9] = {0, 4, 8, 13, 20, 27, 37, 44, 54};
  // conserved combinations c_k . n (the nuclei of H, He, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 3;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 2: return 1;
      case 3: return 2;
      case 4: return 2;
      case 14: return 1;
      case 15: return 1;
      case 16: return 1;
      case 19: return 1;
      case 20: return -1;
      case 22: return 1;
      case 24: return 1;
      case 25: return 2;
      case 26: return -1;
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
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(1.0);
    s += n[1] * R(1.0);
    s += n[2] * R(1.0);
    s += n[3] * R(2.0);
    s += n[4] * R(2.0);
    s += n[5] * R(0.0);
    s += n[6] * R(0.0);
    s += n[7] * R(0.0);
    s += n[8] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.6738233791328e-24);
    s += n[1] * R(1.6729124407627538e-24);
    s += n[2] * R(1.6747343175028463e-24);
    s += n[3] * R(3.3476467582656e-24);
    s += n[4] * R(3.346735819895553e-24);
    s += n[5] * R(6.64647366797316e-24);
    s += n[6] * R(6.645562729603113e-24);
    s += n[7] * R(6.644651791233067e-24);
    s += n[8] * R(9.10938370046373e-28);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[0] = kx::exp((((((((((-R(32.71396786)) + (R(13.536556) * v_lnTe)) - (R(5.73932875) * kx::ipow<2>(v_lnTe))) + (R(1.56315498) * kx::ipow<3>(v_lnTe))) - (R(0.2877056) * kx::ipow<4>(v_lnTe))) + (R(0.0348255977) * kx::ipow<5>(v_lnTe))) - (R(0.00263197617) * kx::ipow<6>(v_lnTe))) + (R(0.000111954395) * kx::ipow<7>(v_lnTe))) - (R(2.03914985e-06) * kx::ipow<8>(v_lnTe))));  // H + e- -> H+ + e- + e-
    }
    {  // H+ + e- -> H: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas < R(5500.0)) k[1] = (R(3.92e-13) * kx::pow(v_invTe, R(0.6353)));  // H+ + e- -> H
      if (Tgas >= R(5500.0)) k[2] = kx::exp(((((((((((-R(28.61303380689232)) - (R(0.7241125657826851) * v_lnTe)) - (R(0.02026044731984691) * kx::ipow<2>(v_lnTe))) - (R(0.002380861877349834) * kx::ipow<3>(v_lnTe))) - (R(0.0003212605213188796) * kx::ipow<4>(v_lnTe))) - (R(1.421502914054107e-05) * kx::ipow<5>(v_lnTe))) + (R(4.989108920299513e-06) * kx::ipow<6>(v_lnTe))) + (R(5.755614137575758e-07) * kx::ipow<7>(v_lnTe))) - (R(1.856767039775261e-08) * kx::ipow<8>(v_lnTe))) - (R(3.071135243196595e-09) * kx::ipow<9>(v_lnTe))));  // H+ + e- -> H
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[3] = kx::exp((((((((((-R(44.09864886)) + (R(23.91596563) * v_lnTe)) - (R(10.7532302) * kx::ipow<2>(v_lnTe))) + (R(3.05803875) * kx::ipow<3>(v_lnTe))) - (R(0.56851189) * kx::ipow<4>(v_lnTe))) + (R(0.0679539123) * kx::ipow<5>(v_lnTe))) - (R(0.0050090561) * kx::ipow<6>(v_lnTe))) + (R(0.000206723616) * kx::ipow<7>(v_lnTe))) - (R(3.64916141e-06) * kx::ipow<8>(v_lnTe))));  // He + e- -> He+ + e- + e-
    }
    {  // He+ + e- -> He: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      if (Tgas < R(9280.0)) k[4] = (R(3.92e-13) * kx::pow(v_invTe, R(0.6353)));  // He+ + e- -> He
      if (Tgas >= R(9280.0)) k[5] = (((R(1.54e-09) * (R(1.0) + (R(0.3) / kx::exp((R(8.099328789667) * v_invTe))))) / (kx::exp((R(40.49664394833662) * v_invTe)) * kx::pow(v_Te, R(1.5)))) + (R(3.92e-13) / kx::pow(v_Te, R(0.6353))));  // He+ + e- -> He
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[6] = kx::exp((((((((((-R(68.71040990212)) + (R(43.93347632635) * v_lnTe)) - (R(18.48066993568) * kx::ipow<2>(v_lnTe))) + (R(4.701626486759002) * kx::ipow<3>(v_lnTe))) - (R(0.7692466334492) * kx::ipow<4>(v_lnTe))) + (R(0.08113042097303) * kx::ipow<5>(v_lnTe))) - (R(0.005324020628287001) * kx::ipow<6>(v_lnTe))) + (R(0.0001975705312221) * kx::ipow<7>(v_lnTe))) - (R(3.165581065665e-06) * kx::ipow<8>(v_lnTe))));  // He+ + e- -> He++ + e- + e-
    }
    k[7] = (R(1.891e-10) / ((kx::sqrt((Tgas / R(9.37))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(9.37)))), R(0.2476))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2774000.0)))), R(1.7524))));  // He++ + e- -> He+
    k[8] = ((R(1.4e-18) * kx::pow(Tgas, R(0.928))) * kx::exp(((-Tgas) / R(16200.0))));  // H + e- -> H-
    {
      const R v_a1 = R(1.35e-09);
      const R v_a10 = R(3.0424);
      const R v_a11 = R(3.2576e-14);
      const R v_a12 = R(3.7741);
      const R v_a2 = R(0.098493);
      const R v_a3 = R(0.32852);
      const R v_a4 = R(0.5561);
      const R v_a5 = R(2.771e-07);
      const R v_a6 = R(2.1826);
      const R v_a7 = R(0.006191);
      const R v_a8 = R(1.0461);
      const R v_a9 = R(8.9712e-11);
      k[9] = ((v_a1 * ((kx::pow(Tgas, v_a2) + (v_a3 * kx::pow(Tgas, v_a4))) + (v_a5 * kx::pow(Tgas, v_a6)))) / (((R(1.0) + (v_a7 * kx::pow(Tgas, v_a8))) + (v_a9 * kx::pow(Tgas, v_a10))) + (v_a11 * kx::pow(Tgas, v_a12))));  // H- + H -> H2 + e-
    }
    {  // H + H+ -> H2+: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(30.0)) k[10] = (R(2.1e-20) * kx::pow((Tgas / R(30.0)), (-R(0.15))));  // H + H+ -> H2+
      if (Tgas >= R(30.0)) k[11] = kx::pow(R(10.0), ((((-R(18.2)) - (R(3.194) * kx::log10(Tgas))) + (R(1.786) * kx::ipow<2>(kx::log10(Tgas)))) - (R(0.2072) * kx::ipow<3>(kx::log10(Tgas)))));  // H + H+ -> H2+
    }
    k[12] = R(6e-10);  // H2+ + H -> H2 + H+
    {  // H2 + H+ -> H2+ + H: 1 piece(s), zero below 100.0, clamped above 30000.0
      R Tgas = Trate;
      if (Tgas > R(30000.0)) Tgas = R(30000.0);
      const R v_asav = R(21237.15);
      const R v_invT = (R(1.0) / Tgas);
      const R v_bsav1 = (-R(3.3232183e-07));
      const R v_bsav2 = R(3.3735382e-07);
      const R v_bsav3 = (-R(1.4491368e-07));
      const R v_bsav4 = R(3.4172805e-08);
      const R v_bsav5 = (-R(4.7813728e-09));
      const R v_bsav6 = R(3.9731542e-10);
      const R v_bsav7 = (-R(1.8171411e-11));
      const R v_bsav8 = R(3.5311932e-13);
      const R v_sumsav = (((((((v_bsav1 + (v_bsav2 * kx::log(Tgas))) + (v_bsav3 * kx::ipow<2>(kx::log(Tgas)))) + (v_bsav4 * kx::ipow<3>(kx::log(Tgas)))) + (v_bsav5 * kx::ipow<4>(kx::log(Tgas)))) + (v_bsav6 * kx::ipow<5>(kx::log(Tgas)))) + (v_bsav7 * kx::ipow<6>(kx::log(Tgas)))) + (v_bsav8 * kx::ipow<7>(kx::log(Tgas))));
      if (Tgas >= R(100.0)) k[13] = (v_sumsav * kx::exp(((-v_asav) * v_invT)));  // H2 + H+ -> H2+ + H
    }
    k[14] = ((R(35.5) * kx::pow(Tgas, (-R(2.28)))) * kx::exp(((-R(46707.0)) / Tgas)));  // H2 + e- -> H + H-
    k[15] = ((R(4.38e-10) * kx::exp(((-R(102000.0)) / Tgas))) * kx::pow(Tgas, R(0.35)));  // H2 + e- -> H + H + e-
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[17] = kx::exp((((((((((-R(18.01849334273)) + (R(2.360852208681) * v_lnTe)) - (R(0.2827443061704) * kx::ipow<2>(v_lnTe))) + (R(0.01623316639567) * kx::ipow<3>(v_lnTe))) - (R(0.03365012031362999) * kx::ipow<4>(v_lnTe))) + (R(0.01178329782711) * kx::ipow<5>(v_lnTe))) - (R(0.001656194699504) * kx::ipow<6>(v_lnTe))) + (R(0.0001068275202678) * kx::ipow<7>(v_lnTe))) - (R(2.631285809207e-06) * kx::ipow<8>(v_lnTe))));  // H- + e- -> H + e- + e-
    }
    {  // H- + H -> H + H + e-: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas < R(1160.0)) k[18] = (R(2.56e-09) * kx::pow(v_Te, R(1.78186)));  // H- + H -> H + H + e-
      if (Tgas >= R(1160.0)) k[19] = kx::exp(((((((((((-R(20.37260896533324)) + (R(1.139449335841631) * v_lnTe)) - (R(0.1421013521554148) * kx::ipow<2>(v_lnTe))) + (R(0.00846445538663) * kx::ipow<3>(v_lnTe))) - (R(0.0014327641212992) * kx::ipow<4>(v_lnTe))) + (R(0.0002012250284791) * kx::ipow<5>(v_lnTe))) + (R(8.66396324309e-05) * kx::ipow<6>(v_lnTe))) - (R(2.585009680264e-05) * kx::ipow<7>(v_lnTe))) + (R(2.4555011970392e-06) * kx::ipow<8>(v_lnTe))) - (R(8.06838246118e-08) * kx::ipow<9>(v_lnTe))));  // H- + H -> H + H + e-
    }
    {  // H- + H+ -> H + H: 1 piece(s), zero below 10.0, clamped above 100000.0
      R Tgas = Trate;
      if (Tgas > R(100000.0)) Tgas = R(100000.0);
      if (Tgas >= R(10.0)) k[20] = ((((R(2.96e-06) / kx::sqrt(Tgas)) - R(1.73e-09)) + (R(2.5e-10) * kx::sqrt(Tgas))) - (R(7.77e-13) * Tgas));  // H- + H+ -> H + H
    }
    k[21] = (R(1e-08) * kx::pow(Tgas, (-R(0.4))));  // H- + H+ -> H2+ + e-
    {  // H2+ + e- -> H + H: 1 piece(s), zero below 0.0, clamped above 10000.0
      R Tgas = Trate;
      if (Tgas > R(10000.0)) Tgas = R(10000.0);
      if (true) k[22] = (R(1000000.0) * (((((R(4.2278e-14) - (R(2.3088e-17) * Tgas)) + (R(7.3428e-21) * kx::ipow<2>(Tgas))) - (R(7.5474e-25) * kx::ipow<3>(Tgas))) + (R(3.3468e-29) * kx::ipow<4>(Tgas))) - (R(5.528e-34) * kx::ipow<5>(Tgas))));  // H2+ + e- -> H + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[23] = (R(5e-07) * kx::sqrt((R(100.0) * v_invT)));  // H2+ + H- -> H + H2
    }
    k[24] = ((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5)))));  // H + H + H -> H2 + H
    k[25] = (((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5))))) / R(8.0));  // H2 + H + H -> H2 + H2
  }
  // density-dependent rates
  template <class R>
  KX_HD static void rates_n(R Trate, const R* n, const Params& p, R* k) {
    (void)p; (void)n;
    const R Tgas = Trate;
    (void)Tgas;
    k[16] = kx::dissH2_Martin96(Tgas, Hnuclei(n));  // H2 + H -> H + H + H
    {
      const R v_T = Tgas;
      const R v_ncr21 = kx::pow(R(10.0), ((R(4.845) - (R(1.3) * kx::log10((v_T * R(0.0001))))) + (R(1.62) * kx::ipow<2>(kx::log10((v_T * R(0.0001)))))));
      const R v_a21 = (R(1.0) / (R(1.0) + (Hnuclei(n) / v_ncr21)));
      const R v_invT = (R(1.0) / Tgas);
      const R v_kh21 = (((R(8.125e-08) * kx::pow(v_T, (-R(0.5)))) * kx::exp(((-R(52000.0)) * v_invT))) * (R(1.0) - kx::exp(((-R(6000.0)) * v_invT))));
      const R v_kl21 = (R(1.18e-10) * kx::exp(((-R(69500.0)) * v_invT)));
      k[26] = (kx::pow(v_kh21, (R(1.0) - v_a21)) * kx::pow(v_kl21, v_a21));  // H2 + H2 -> H + H + H2
    }
  }
  // every reaction: the T-only ones in k (zeroed) then the rest
  template <class R>
  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {
    for (int i = 0; i < NREA; ++i) k[i] = R(0);
    rates_T(Trate, p, k);
    rates_n(Trate, n, p, k);
  }
  static constexpr int NREA_T = 25;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[28] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 51, 54, 56};
  static constexpr int LT_KF_SP[56] = {0, 8, 1, 8, 1, 8, 5, 8, 6, 8, 6, 8, 6, 8, 7, 8, 0, 8, 2, 0, 0, 1, 0, 1, 4, 0, 3, 1, 3, 8, 3, 8, 3, 0, 2, 8, 2, 0, 2, 0, 2, 1, 2, 1, 4, 8, 4, 2, 0, 0, 0, 3, 0, 0, 3, 3};
  static constexpr int LT_SC_OFF[28] = {0, 5, 8, 11, 16, 19, 22, 27, 30, 33, 37, 40, 43, 47, 51, 55, 60, 65, 70, 75, 80, 84, 88, 92, 96, 101, 106, 111};
  static constexpr int LT_SC_ROW[111] = {0, 8, 1, 8, 8, 1, 8, 0, 1, 8, 0, 5, 8, 6, 8, 8, 6, 8, 5, 6, 8, 5, 6, 8, 7, 8, 8, 7, 8, 6, 0, 8, 2, 2, 0, 3, 8, 0, 1, 4, 0, 1, 4, 4, 0, 3, 1, 3, 1, 4, 0, 3, 8, 0, 2, 3, 8, 0, 0, 8, 3, 0, 0, 0, 0, 2, 8, 0, 8, 8, 2, 0, 0, 0, 8, 2, 0, 0, 0, 8, 2, 1, 0, 0, 2, 1, 4, 8, 4, 8, 0, 0, 4, 2, 0, 3, 0, 0, 0, 3, 0, 3, 0, 0, 3, 3, 3, 3, 0, 0, 3};
  static constexpr int LT_SC_SGN[111] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1};
  static constexpr int LT_JT_OFF[28] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 49, 51, 52};
  static constexpr int LT_JT_COL[52] = {0, 8, 1, 8, 1, 8, 5, 8, 6, 8, 6, 8, 6, 8, 7, 8, 0, 8, 0, 2, 0, 1, 0, 1, 0, 4, 1, 3, 3, 8, 3, 8, 0, 3, 2, 8, 0, 2, 0, 2, 1, 2, 1, 2, 4, 8, 2, 4, 0, 0, 3, 3};
  static constexpr int LT_JT_MULT[52] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 2, 1, 2};
  static constexpr int LT_JT_FOFF[53] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 50, 52, 54, 55};
  static constexpr int LT_JF_SP[55] = {8, 0, 8, 1, 8, 1, 8, 5, 8, 6, 8, 6, 8, 6, 8, 7, 8, 0, 2, 0, 1, 0, 1, 0, 4, 0, 3, 1, 8, 3, 8, 3, 3, 0, 8, 2, 2, 0, 2, 0, 2, 1, 2, 1, 8, 4, 4, 2, 0, 0, 3, 0, 0, 0, 3};
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
    fl = k[0]*n[0]*n[8];  // H + e- -> H+ + e- + e-
    dn[0] -= fl;
    dn[8] -= fl;
    dn[1] += fl;
    dn[8] += fl;
    dn[8] += fl;
    fl = k[1]*n[1]*n[8];  // H+ + e- -> H
    dn[1] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    fl = k[2]*n[1]*n[8];  // H+ + e- -> H
    dn[1] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    fl = k[3]*n[5]*n[8];  // He + e- -> He+ + e- + e-
    dn[5] -= fl;
    dn[8] -= fl;
    dn[6] += fl;
    dn[8] += fl;
    dn[8] += fl;
    fl = k[4]*n[6]*n[8];  // He+ + e- -> He
    dn[6] -= fl;
    dn[8] -= fl;
    dn[5] += fl;
    fl = k[5]*n[6]*n[8];  // He+ + e- -> He
    dn[6] -= fl;
    dn[8] -= fl;
    dn[5] += fl;
    fl = k[6]*n[6]*n[8];  // He+ + e- -> He++ + e- + e-
    dn[6] -= fl;
    dn[8] -= fl;
    dn[7] += fl;
    dn[8] += fl;
    dn[8] += fl;
    fl = k[7]*n[7]*n[8];  // He++ + e- -> He+
    dn[7] -= fl;
    dn[8] -= fl;
    dn[6] += fl;
    fl = k[8]*n[0]*n[8];  // H + e- -> H-
    dn[0] -= fl;
    dn[8] -= fl;
    dn[2] += fl;
    fl = k[9]*n[2]*n[0];  // H- + H -> H2 + e-
    dn[2] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[8] += fl;
    fl = k[10]*n[0]*n[1];  // H + H+ -> H2+
    dn[0] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    fl = k[11]*n[0]*n[1];  // H + H+ -> H2+
    dn[0] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    fl = k[12]*n[4]*n[0];  // H2+ + H -> H2 + H+
    dn[4] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[1] += fl;
    fl = k[13]*n[3]*n[1];  // H2 + H+ -> H2+ + H
    dn[3] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    dn[0] += fl;
    fl = k[14]*n[3]*n[8];  // H2 + e- -> H + H-
    dn[3] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    dn[2] += fl;
    fl = k[15]*n[3]*n[8];  // H2 + e- -> H + H + e-
    dn[3] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[8] += fl;
    fl = k[16]*n[3]*n[0];  // H2 + H -> H + H + H
    dn[3] -= fl;
    dn[0] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[17]*n[2]*n[8];  // H- + e- -> H + e- + e-
    dn[2] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    dn[8] += fl;
    dn[8] += fl;
    fl = k[18]*n[2]*n[0];  // H- + H -> H + H + e-
    dn[2] -= fl;
    dn[0] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[8] += fl;
    fl = k[19]*n[2]*n[0];  // H- + H -> H + H + e-
    dn[2] -= fl;
    dn[0] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[8] += fl;
    fl = k[20]*n[2]*n[1];  // H- + H+ -> H + H
    dn[2] -= fl;
    dn[1] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[21]*n[2]*n[1];  // H- + H+ -> H2+ + e-
    dn[2] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    dn[8] += fl;
    fl = k[22]*n[4]*n[8];  // H2+ + e- -> H + H
    dn[4] -= fl;
    dn[8] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[23]*n[4]*n[2];  // H2+ + H- -> H + H2
    dn[4] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[3] += fl;
    fl = k[24]*n[0]*n[0]*n[0];  // H + H + H -> H2 + H
    dn[0] -= fl;
    dn[0] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[0] += fl;
    fl = k[25]*n[3]*n[0]*n[0];  // H2 + H + H -> H2 + H2
    dn[3] -= fl;
    dn[0] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[3] += fl;
    fl = k[26]*n[3]*n[3];  // H2 + H2 -> H + H + H2
    dn[3] -= fl;
    dn[3] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[3] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[8];  // H + e- -> H+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[8 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[8 + 0 * ld] += d;
    J[8 + 0 * ld] += d;
    d = k[0]*n[0];  // H + e- -> H+ + e- + e-, d/d n[8]
    J[0 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[1 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    d = k[1]*n[8];  // H+ + e- -> H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[8 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[1]*n[1];  // H+ + e- -> H, d/d n[8]
    J[1 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    d = k[2]*n[8];  // H+ + e- -> H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[8 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[2]*n[1];  // H+ + e- -> H, d/d n[8]
    J[1 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    d = k[3]*n[8];  // He + e- -> He+ + e- + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[8 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[8 + 5 * ld] += d;
    J[8 + 5 * ld] += d;
    d = k[3]*n[5];  // He + e- -> He+ + e- + e-, d/d n[8]
    J[5 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[6 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    d = k[4]*n[8];  // He+ + e- -> He, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[8 + 6 * ld] -= d;
    J[5 + 6 * ld] += d;
    d = k[4]*n[6];  // He+ + e- -> He, d/d n[8]
    J[6 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[5 + 8 * ld] += d;
    d = k[5]*n[8];  // He+ + e- -> He, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[8 + 6 * ld] -= d;
    J[5 + 6 * ld] += d;
    d = k[5]*n[6];  // He+ + e- -> He, d/d n[8]
    J[6 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[5 + 8 * ld] += d;
    d = k[6]*n[8];  // He+ + e- -> He++ + e- + e-, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[8 + 6 * ld] -= d;
    J[7 + 6 * ld] += d;
    J[8 + 6 * ld] += d;
    J[8 + 6 * ld] += d;
    d = k[6]*n[6];  // He+ + e- -> He++ + e- + e-, d/d n[8]
    J[6 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    d = k[7]*n[8];  // He++ + e- -> He+, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[8 + 7 * ld] -= d;
    J[6 + 7 * ld] += d;
    d = k[7]*n[7];  // He++ + e- -> He+, d/d n[8]
    J[7 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[6 + 8 * ld] += d;
    d = k[8]*n[8];  // H + e- -> H-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[8 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    d = k[8]*n[0];  // H + e- -> H-, d/d n[8]
    J[0 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[2 + 8 * ld] += d;
    d = k[9]*n[2];  // H- + H -> H2 + e-, d/d n[0]
    J[2 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[8 + 0 * ld] += d;
    d = k[9]*n[0];  // H- + H -> H2 + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    d = k[10]*n[1];  // H + H+ -> H2+, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] -= d;
    J[4 + 0 * ld] += d;
    d = k[10]*n[0];  // H + H+ -> H2+, d/d n[1]
    J[0 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    d = k[11]*n[1];  // H + H+ -> H2+, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] -= d;
    J[4 + 0 * ld] += d;
    d = k[11]*n[0];  // H + H+ -> H2+, d/d n[1]
    J[0 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    d = k[12]*n[4];  // H2+ + H -> H2 + H+, d/d n[0]
    J[4 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[1 + 0 * ld] += d;
    d = k[12]*n[0];  // H2+ + H -> H2 + H+, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[0 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[1 + 4 * ld] += d;
    d = k[13]*n[3];  // H2 + H+ -> H2+ + H, d/d n[1]
    J[3 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[13]*n[1];  // H2 + H+ -> H2+ + H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[1 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[14]*n[8];  // H2 + e- -> H + H-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[8 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[14]*n[3];  // H2 + e- -> H + H-, d/d n[8]
    J[3 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    J[2 + 8 * ld] += d;
    d = k[15]*n[8];  // H2 + e- -> H + H + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[8 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[8 + 3 * ld] += d;
    d = k[15]*n[3];  // H2 + e- -> H + H + e-, d/d n[8]
    J[3 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    J[0 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    d = k[16]*n[3];  // H2 + H -> H + H + H, d/d n[0]
    J[3 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    d = k[16]*n[0];  // H2 + H -> H + H + H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[17]*n[8];  // H- + e- -> H + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[8 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    d = k[17]*n[2];  // H- + e- -> H + e- + e-, d/d n[8]
    J[2 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    J[8 + 8 * ld] += d;
    d = k[18]*n[2];  // H- + H -> H + H + e-, d/d n[0]
    J[2 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    J[8 + 0 * ld] += d;
    d = k[18]*n[0];  // H- + H -> H + H + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    d = k[19]*n[2];  // H- + H -> H + H + e-, d/d n[0]
    J[2 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    J[8 + 0 * ld] += d;
    d = k[19]*n[0];  // H- + H -> H + H + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    d = k[20]*n[2];  // H- + H+ -> H + H, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[20]*n[1];  // H- + H+ -> H + H, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    d = k[21]*n[2];  // H- + H+ -> H2+ + e-, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    J[8 + 1 * ld] += d;
    d = k[21]*n[1];  // H- + H+ -> H2+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    J[8 + 2 * ld] += d;
    d = k[22]*n[8];  // H2+ + e- -> H + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[8 + 4 * ld] -= d;
    J[0 + 4 * ld] += d;
    J[0 + 4 * ld] += d;
    d = k[22]*n[4];  // H2+ + e- -> H + H, d/d n[8]
    J[4 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[0 + 8 * ld] += d;
    J[0 + 8 * ld] += d;
    d = k[23]*n[4];  // H2+ + H- -> H + H2, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[23]*n[2];  // H2+ + H- -> H + H2, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[0 + 4 * ld] += d;
    J[3 + 4 * ld] += d;
    d = k[24]*n[0]*n[0]*R(3.0);  // H + H + H -> H2 + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    d = k[25]*n[3]*n[0]*R(2.0);  // H2 + H + H -> H2 + H2, d/d n[0]
    J[3 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[3 + 0 * ld] += d;
    d = k[25]*n[0]*n[0];  // H2 + H + H -> H2 + H2, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[3 + 3 * ld] += d;
    J[3 + 3 * ld] += d;
    d = k[26]*n[3]*R(2.0);  // H2 + H2 -> H + H + H2, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[3 + 3 * ld] += d;
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
  // thermal terms at the thermal temperature Tth (capped to
  // [2.73, 1e9] K by the caller) with the rates k of the state
  template <class R>
  KX_HD static R cooling_chem(const R* n, const R* k) {
    R c = R(0);
    c += k[15] * R(4.48) * n[3]*n[8];  // H2 + e- -> H + H + e-
    c += k[16] * R(4.48) * n[3]*n[0];  // H2 + H -> H + H + H
    c += k[26] * R(4.48) * n[3]*n[3];  // H2 + H2 -> H + H + H2
    return c * R(1.602176634e-12);
  }
  template <class R>
  KX_HD static R heating_chem(R Tth, const R* n, const R* k) {
    R h = R(0);
    const R ffac = kx::h2_heat_fraction(Tth, n[0], n[3], Hnuclei(n));
    h += k[9] * (R(3.53) * ffac) * n[2]*n[0];  // H- + H -> H2 + e-
    h += k[12] * (R(1.83) * ffac) * n[4]*n[0];  // H2+ + H -> H2 + H+
    h += k[24] * (R(4.48) * ffac) * n[0]*n[0]*n[0];  // H + H + H -> H2 + H
    h += k[25] * (R(4.48) * ffac) * n[3]*n[0]*n[0];  // H2 + H + H -> H2 + H2
    (void)Tth;
    return h * R(1.602176634e-12);
  }
  template <class R>
  KX_HD static R gamma_index(R Tth, const R* n, const Params& p) {
    const R iTv = R(1) / kx::fmax(Tth, R(63.3236349736972));
    const R x_H2 = R(6332.36349736972) * iTv;
    const R gi_H2 = R(0.5) * (R(3) + R(2) * kx::gamma_rot_ortho_para(Tth, R(p.opratio) / (R(p.opratio) + R(1))) + R(2) * kx::vib_heat_capacity(x_H2));
    const R gi_H2j = R(0.5) * (R(3) + R(2) * kx::gamma_rot(Tth, R(61.72352804491875)) + R(0));
    const R mosum = n[0] + n[1] + n[2] + n[5] + n[6] + n[7] + n[8];
    const R ysum = mosum + n[3] + n[4];
    const R gsum = R(1.5) * mosum + n[3] * gi_H2 + n[4] * gi_H2j;
    (void)p;
    return gsum > R(0) ? R(1) + ysum / gsum : R(5) / R(3);
  }
  // dTgas/dt: heating minus cooling (the reference's term order:
  // heat CHEM, COMPRESS, PHOTO; cool H2, CHEM, CONT)
  template <class R>
  KX_HD static R thermal(R Tth, const R* n, const R* k, const Params& p) {
    R heat = R(0), cool = R(0);
    heat += heating_chem(Tth, n, k);
    heat += ntot(n) * R(1.380649e-16) * Tth / R(p.tff);
    cool += kx::cooling_H2(Tth, n[0], n[1], n[3], n[8], n[5]) * kx::h2_opacity_ripamonti(ntot(n));
    cool += cooling_chem(n, k);
    cool += kx::cooling_continuum(Tth, rho(n), ntot(n));
    return (gamma_index(Tth, n, p) - R(1)) * (heat - cool)
        / (R(1.380649e-16) * ntot(n));
  }
  // the active thermal terms one by one, in the order of the names
  // (heat: CHEM, COMPRESS; cool: H2, CHEM, CONT), the same expressions as thermal()
  template <class R>
  KX_HD static void thermal_terms(R Tth, const R* n, const R* k,
                                  const Params& p, R* heats, R* cools) {
    heats[0] = heating_chem(Tth, n, k);
    heats[1] = ntot(n) * R(1.380649e-16) * Tth / R(p.tff);
    cools[0] = kx::cooling_H2(Tth, n[0], n[1], n[3], n[8], n[5]) * kx::h2_opacity_ripamonti(ntot(n));
    cools[1] = cooling_chem(n, k);
    cools[2] = kx::cooling_continuum(Tth, rho(n), ntot(n));
    (void)Tth; (void)n; (void)k; (void)p; (void)heats; (void)cools;
  }
  // d(heat - cool)/dk_r [erg cm^-3 s^-1 per unit rate] into c[NREA]:
  // the CHEM terms, the only thermal terms that read the rates
  template <class R>
  KX_HD static void chem_dk(R Tth, const R* n, R* c) {
    for (int r = 0; r < NREA; ++r) c[r] = R(0);
    const R ffac = kx::h2_heat_fraction(Tth, n[0], n[3], Hnuclei(n));
    c[9] += (R(3.53) * ffac) * n[2]*n[0] * R(1.602176634e-12);
    c[12] += (R(1.83) * ffac) * n[4]*n[0] * R(1.602176634e-12);
    c[24] += (R(4.48) * ffac) * n[0]*n[0]*n[0] * R(1.602176634e-12);
    c[25] += (R(4.48) * ffac) * n[3]*n[0]*n[0] * R(1.602176634e-12);
    c[15] -= R(4.48) * n[3]*n[8] * R(1.602176634e-12);
    c[16] -= R(4.48) * n[3]*n[0] * R(1.602176634e-12);
    c[26] -= R(4.48) * n[3]*n[3] * R(1.602176634e-12);
    (void)Tth; (void)n;
  }
};
}  // namespace kx
