// generated by netgen.py from primordial2.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_primordial2 {
  static constexpr const char* NAME = "primordial2";
  static constexpr int NSPEC = 11;
  static constexpr int NREA = 36;
  static constexpr int NPAR = 2;
  static constexpr int IDX_H = 0;
  static constexpr int IDX_Hj = 1;
  static constexpr int IDX_Hk = 2;
  static constexpr int IDX_H2 = 3;
  static constexpr int IDX_H2j = 4;
  static constexpr int IDX_H3j = 5;
  static constexpr int IDX_He = 6;
  static constexpr int IDX_Hej = 7;
  static constexpr int IDX_Hejj = 8;
  static constexpr int IDX_HeHj = 9;
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
      case 1: return 1;
      case 2: return -1;
      case 4: return 1;
      case 5: return 1;
      case 7: return 1;
      case 8: return 2;
      case 9: return 1;
      case 10: return -1;
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
      case 5: return 5.020559199028353e-24;
      case 6: return 6.64647366797316e-24;
      case 7: return 6.645562729603113e-24;
      case 8: return 6.644651791233067e-24;
      case 9: return 8.319386108735913e-24;
      case 10: return 9.10938370046373e-28;
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
      case 10: return 3;
      case 13: return 1;
      case 15: return 1;
      case 17: return 1;
      case 18: return 1;
      case 19: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 86 nonzeros, 90 in L + U with the fill, 398 flops per
  // factorization (dense: 887); rows in pivot order
  static constexpr int SP_NNZ = 90;
  static constexpr int SP_FLOPS = 398;
  static constexpr int sp_perm[11] = {8, 1, 5, 7, 0, 2, 3, 4, 6, 9, 10};
  static constexpr int sp_rp[12] = {0, 3, 10, 16, 23, 34, 44, 55, 64, 72, 79, 90};
  static constexpr int sp_ci[90] = {0, 3, 10, 1, 3, 4, 5, 6, 7, 10, 2, 5, 6, 7, 9, 10, 0, 3, 4, 5, 6, 8, 10, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 1, 3, 4, 5, 6, 7, 8, 9, 10, 3, 4, 5, 6, 7, 8, 9, 10, 4, 5, 6, 7, 8, 9, 10, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
  static constexpr int sp_di[11] = {0, 3, 10, 17, 27, 38, 50, 60, 69, 77, 89};
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
      case 5: return 3;
      case 9: return 1;
      case 17: return 1;
      case 18: return 1;
      case 19: return 1;
      case 20: return 1;
      case 23: return 1;
      case 24: return -1;
      case 26: return 1;
      case 27: return 1;
      case 29: return 1;
      case 30: return 2;
      case 31: return 1;
      case 32: return -1;
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
    s += n[0] * R(1.0);
    s += n[1] * R(1.0);
    s += n[2] * R(1.0);
    s += n[3] * R(2.0);
    s += n[4] * R(2.0);
    s += n[5] * R(3.0);
    s += n[6] * R(0.0);
    s += n[7] * R(0.0);
    s += n[8] * R(0.0);
    s += n[9] * R(1.0);
    s += n[10] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.6738233791328e-24);
    s += n[1] * R(1.6729124407627538e-24);
    s += n[2] * R(1.6747343175028463e-24);
    s += n[3] * R(3.3476467582656e-24);
    s += n[4] * R(3.346735819895553e-24);
    s += n[5] * R(5.020559199028353e-24);
    s += n[6] * R(6.64647366797316e-24);
    s += n[7] * R(6.645562729603113e-24);
    s += n[8] * R(6.644651791233067e-24);
    s += n[9] * R(8.319386108735913e-24);
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
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[0] = kx::exp((((((((((-R(32.71396786)) + (R(13.536556) * v_lnTe)) - (R(5.73932875) * kx::ipow<2>(v_lnTe))) + (R(1.56315498) * kx::ipow<3>(v_lnTe))) - (R(0.2877056) * kx::ipow<4>(v_lnTe))) + (R(0.0348255977) * kx::ipow<5>(v_lnTe))) - (R(0.00263197617) * kx::ipow<6>(v_lnTe))) + (R(0.000111954395) * kx::ipow<7>(v_lnTe))) - (R(2.03914985e-06) * kx::ipow<8>(v_lnTe))));  // H + e- -> H+ + e- + e-
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[1] = kx::exp(((((((((((-R(28.61303380689232)) - (R(0.7241125657826851) * v_lnTe)) - (R(0.02026044731984691) * kx::ipow<2>(v_lnTe))) - (R(0.002380861877349834) * kx::ipow<3>(v_lnTe))) - (R(0.0003212605213188796) * kx::ipow<4>(v_lnTe))) - (R(1.421502914054107e-05) * kx::ipow<5>(v_lnTe))) + (R(4.989108920299513e-06) * kx::ipow<6>(v_lnTe))) + (R(5.755614137575758e-07) * kx::ipow<7>(v_lnTe))) - (R(1.856767039775261e-08) * kx::ipow<8>(v_lnTe))) - (R(3.071135243196595e-09) * kx::ipow<9>(v_lnTe))));  // H+ + e- -> H
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[2] = kx::exp((((((((((-R(44.09864886)) + (R(23.91596563) * v_lnTe)) - (R(10.7532302) * kx::ipow<2>(v_lnTe))) + (R(3.05803875) * kx::ipow<3>(v_lnTe))) - (R(0.56851189) * kx::ipow<4>(v_lnTe))) + (R(0.0679539123) * kx::ipow<5>(v_lnTe))) - (R(0.0050090561) * kx::ipow<6>(v_lnTe))) + (R(0.000206723616) * kx::ipow<7>(v_lnTe))) - (R(3.64916141e-06) * kx::ipow<8>(v_lnTe))));  // He + e- -> He+ + e- + e-
    }
    {
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      k[3] = ((R(3.925e-13) * kx::pow(v_invTe, R(0.6353))) + ((R(1.544e-09) * kx::pow(v_invTe, R(1.5))) * ((R(0.3) * kx::exp(((-R(48.596)) * v_invTe))) + kx::exp(((-R(40.496)) * v_invTe)))));  // He+ + e- -> He
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[4] = kx::exp((((((((((-R(68.71040990212)) + (R(43.93347632635) * v_lnTe)) - (R(18.48066993568) * kx::ipow<2>(v_lnTe))) + (R(4.701626486759002) * kx::ipow<3>(v_lnTe))) - (R(0.7692466334492) * kx::ipow<4>(v_lnTe))) + (R(0.08113042097303) * kx::ipow<5>(v_lnTe))) - (R(0.005324020628287001) * kx::ipow<6>(v_lnTe))) + (R(0.0001975705312221) * kx::ipow<7>(v_lnTe))) - (R(3.165581065665e-06) * kx::ipow<8>(v_lnTe))));  // He+ + e- -> He++ + e- + e-
    }
    {
      const R v_sqrTgas = kx::sqrt(Tgas);
      k[5] = (((R(3.36e-10) / v_sqrTgas) / kx::pow((Tgas * R(0.001)), R(0.2))) / (R(1.0) + kx::pow((Tgas * R(1e-06)), R(0.7))));  // He++ + e- -> He+
    }
    k[6] = ((R(1.4e-18) * kx::pow(Tgas, R(0.928))) * kx::exp(((-Tgas) / R(16200.0))));  // H + e- -> H-
    {  // H- + H -> H2 + e-: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(300.0)) k[7] = R(1.5e-09);  // H- + H -> H2 + e-
      if (Tgas >= R(300.0)) k[8] = (R(4e-09) * kx::pow(Tgas, (-R(0.17))));  // H- + H -> H2 + e-
    }
    {
      const R v_logT = kx::log10(Tgas);
      k[9] = kx::pow(R(10.0), ((((-R(19.38)) - (R(1.523) * v_logT)) + (R(1.118) * kx::ipow<2>(v_logT))) - (R(0.1269) * kx::ipow<3>(v_logT))));  // H + H+ -> H2+
    }
    k[10] = R(6.4e-10);  // H2+ + H -> H2 + H+
    {
      const R v_invT = (R(1.0) / Tgas);
      k[11] = (R(3e-10) * kx::exp(((-R(21050.0)) * v_invT)));  // H2 + H+ -> H2+ + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[12] = ((R(4.4e-10) * kx::pow(Tgas, R(0.35))) * kx::exp(((-R(102000.0)) * v_invT)));  // H2 + e- -> H + H + e-
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[14] = kx::exp((((((((((-R(18.01849334273)) + (R(2.360852208681) * v_lnTe)) - (R(0.2827443061704) * kx::ipow<2>(v_lnTe))) + (R(0.01623316639567) * kx::ipow<3>(v_lnTe))) - (R(0.03365012031362999) * kx::ipow<4>(v_lnTe))) + (R(0.01178329782711) * kx::ipow<5>(v_lnTe))) - (R(0.001656194699504) * kx::ipow<6>(v_lnTe))) + (R(0.0001068275202678) * kx::ipow<7>(v_lnTe))) - (R(2.631285809207e-06) * kx::ipow<8>(v_lnTe))));  // H- + e- -> H + e- + e-
    }
    {
      const R v_invsqrT = (R(1.0) / kx::sqrt(Tgas));
      const R v_sqrTgas = kx::sqrt(Tgas);
      k[15] = ((((R(5.7e-06) * v_invsqrT) + R(6.3e-08)) - (R(9.2e-11) * v_sqrTgas)) + (R(4.4e-13) * Tgas));  // H- + H+ -> H + H
    }
    {  // H- + H+ -> H2+ + e-: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(8000.0)) k[16] = (R(6.9e-09) * kx::pow(Tgas, (-R(0.35))));  // H- + H+ -> H2+ + e-
      if (Tgas >= R(8000.0)) k[17] = (R(9.6e-07) * kx::pow(Tgas, (-R(0.9))));  // H- + H+ -> H2+ + e-
    }
    {
      const R v_invsqrT = (R(1.0) / kx::sqrt(Tgas));
      k[18] = (R(2e-07) * v_invsqrT);  // H2+ + e- -> H + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[19] = (R(2.3e-07) / kx::sqrt(v_T32));  // H2+ + H- -> H + H2
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[20] = (R(5.5e-29) * v_invT);  // H + H + H -> H2 + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[21] = ((R(5.5e-29) * v_invT) * R(0.125));  // H2 + H + H -> H2 + H2
    }
    {
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      k[23] = (R(0.00017) * kx::exp((((((((((-R(32.71396786)) + (R(13.536556) * v_lnTe)) - (R(5.73932875) * kx::ipow<2>(v_lnTe))) + (R(1.56315498) * kx::ipow<3>(v_lnTe))) - (R(0.2877056) * kx::ipow<4>(v_lnTe))) + (R(0.0348255977) * kx::ipow<5>(v_lnTe))) - (R(0.00263197617) * kx::ipow<6>(v_lnTe))) + (R(0.000111954395) * kx::ipow<7>(v_lnTe))) - (R(2.03914985e-06) * kx::ipow<8>(v_lnTe)))));  // H + H -> H+ + e- + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[24] = (R(3.7e-14) * kx::exp(((-R(35.0)) * v_invT)));  // He+ + H2 -> H+ + He + H
    }
    k[25] = R(1.3e-10);  // H2+ + He -> HeH+ + H
    k[26] = R(2.08e-09);  // H2+ + H2 -> H3+ + H
    {
      const R v_invsqrT32 = kx::sqrt((R(300.0) / Tgas));
      k[27] = (R(2.3e-07) * v_invsqrT32);  // H3+ + H- -> H2 + H2
    }
    k[28] = R(1.9e-15);  // He+ + H -> H+ + He
    {
      const R v_invsqrT32 = kx::sqrt((R(300.0) / Tgas));
      k[29] = (R(2.3e-07) * v_invsqrT32);  // He+ + H- -> H + He
    }
    k[30] = R(7.2e-15);  // He+ + H2 -> H2+ + He
    k[31] = R(9.1e-10);  // HeH+ + H -> H2+ + He
    k[32] = R(1.5e-09);  // HeH+ + H2 -> H3+ + He
    {
      const R v_invsqrT32 = kx::sqrt((R(300.0) / Tgas));
      k[33] = (R(1.5e-08) * v_invsqrT32);  // H3+ + e- -> H2 + H
    }
    {
      const R v_invsqrT32 = kx::sqrt((R(300.0) / Tgas));
      k[34] = (R(1.5e-08) * v_invsqrT32);  // H3+ + e- -> H + H + H
    }
    {
      const R v_invsqrT32 = kx::sqrt((R(300.0) / Tgas));
      k[35] = (R(2e-07) * v_invsqrT32);  // HeH+ + e- -> H + He
    }
  }
  // density-dependent rates
  template <class R>
  KX_HD static void rates_n(R Trate, const R* n, const Params& p, R* k) {
    (void)p; (void)n;
    const R Tgas = Trate;
    (void)Tgas;
    {
      const R v_T = Tgas;
      const R v_ncr13 = kx::pow(R(10.0), ((R(4.0) - (R(0.416) * kx::log10((v_T * R(0.0001))))) - (R(0.327) * kx::ipow<2>(kx::log10((v_T * R(0.0001)))))));
      const R v_a13 = (R(1.0) / (R(1.0) + (Hnuclei(n) / v_ncr13)));
      const R v_invT = (R(1.0) / Tgas);
      const R v_invsqrT = (R(1.0) / kx::sqrt(Tgas));
      const R v_kh13 = (((R(6.5e-07) * v_invsqrT) * kx::exp(((-R(52000.0)) * v_invT))) * (R(1.0) - kx::exp(((-R(6000.0)) * v_invT))));
      const R v_kl13 = (R(1.12e-10) * kx::exp(((-R(70350.0)) * v_invT)));
      k[13] = (kx::pow(v_kh13, (R(1.0) - v_a13)) * kx::pow(v_kl13, v_a13));  // H2 + H -> H + H + H
    }
    {
      const R v_T = Tgas;
      const R v_ncr21 = kx::pow(R(10.0), ((R(4.845) - (R(1.3) * kx::log10((v_T * R(0.0001))))) + (R(1.62) * kx::ipow<2>(kx::log10((v_T * R(0.0001)))))));
      const R v_a21 = (R(1.0) / (R(1.0) + (Hnuclei(n) / v_ncr21)));
      const R v_invT = (R(1.0) / Tgas);
      const R v_invsqrT = (R(1.0) / kx::sqrt(Tgas));
      const R v_kh21 = (((R(8.125e-08) * v_invsqrT) * kx::exp(((-R(52000.0)) * v_invT))) * (R(1.0) - kx::exp(((-R(6000.0)) * v_invT))));
      const R v_kl21 = (R(1.18e-10) * kx::exp(((-R(69500.0)) * v_invT)));
      k[22] = (kx::pow(v_kh21, (R(1.0) - v_a21)) * kx::pow(v_kl21, v_a21));  // H2 + H2 -> H + H + H2
    }
  }
  // every reaction: the T-only ones in k (zeroed) then the rest
  template <class R>
  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {
    for (int i = 0; i < NREA; ++i) k[i] = R(0);
    rates_T(Trate, p, k);
    rates_n(Trate, n, p, k);
  }
  static constexpr int NREA_T = 34;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[37] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 43, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74};
  static constexpr int LT_KF_SP[74] = {0, 10, 1, 10, 6, 10, 7, 10, 7, 10, 8, 10, 0, 10, 2, 0, 2, 0, 0, 1, 4, 0, 3, 1, 3, 10, 3, 0, 2, 10, 2, 1, 2, 1, 2, 1, 4, 10, 4, 2, 0, 0, 0, 3, 0, 0, 3, 3, 0, 0, 7, 3, 4, 6, 4, 3, 5, 2, 7, 0, 7, 2, 7, 3, 9, 0, 9, 3, 5, 10, 5, 10, 9, 10};
  static constexpr int LT_SC_OFF[37] = {0, 5, 8, 13, 16, 21, 24, 27, 31, 35, 38, 42, 46, 51, 56, 61, 65, 69, 73, 77, 81, 86, 91, 96, 101, 106, 110, 114, 118, 122, 126, 130, 134, 138, 142, 147, 151};
  static constexpr int LT_SC_ROW[151] = {0, 10, 1, 10, 10, 1, 10, 0, 6, 10, 7, 10, 10, 7, 10, 6, 7, 10, 8, 10, 10, 8, 10, 7, 0, 10, 2, 2, 0, 3, 10, 2, 0, 3, 10, 0, 1, 4, 4, 0, 3, 1, 3, 1, 4, 0, 3, 10, 0, 0, 10, 3, 0, 0, 0, 0, 2, 10, 0, 10, 10, 2, 1, 0, 0, 2, 1, 4, 10, 2, 1, 4, 10, 4, 10, 0, 0, 4, 2, 0, 3, 0, 0, 0, 3, 0, 3, 0, 0, 3, 3, 3, 3, 0, 0, 3, 0, 0, 1, 10, 0, 7, 3, 1, 6, 0, 4, 6, 9, 0, 4, 3, 5, 0, 5, 2, 3, 3, 7, 0, 1, 6, 7, 2, 0, 6, 7, 3, 4, 6, 9, 0, 4, 6, 9, 3, 5, 6, 5, 10, 3, 0, 5, 10, 0, 0, 0, 9, 10, 0, 6};
  static constexpr int LT_SC_SGN[151] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1};
  static constexpr int LT_JT_OFF[37] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 41, 43, 44, 45, 47, 49, 51, 53, 55, 57, 59, 61, 63, 65, 67, 69};
  static constexpr int LT_JT_COL[69] = {0, 10, 1, 10, 6, 10, 7, 10, 7, 10, 8, 10, 0, 10, 0, 2, 0, 2, 0, 1, 0, 4, 1, 3, 3, 10, 0, 3, 2, 10, 1, 2, 1, 2, 1, 2, 4, 10, 2, 4, 0, 0, 3, 3, 0, 3, 7, 4, 6, 3, 4, 2, 5, 0, 7, 2, 7, 3, 7, 0, 9, 3, 9, 5, 10, 5, 10, 9, 10};
  static constexpr int LT_JT_MULT[69] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 2, 1, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[70] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 42, 44, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72};
  static constexpr int LT_JF_SP[72] = {10, 0, 10, 1, 10, 6, 10, 7, 10, 7, 10, 8, 10, 0, 2, 0, 2, 0, 1, 0, 4, 0, 3, 1, 10, 3, 3, 0, 10, 2, 2, 1, 2, 1, 2, 1, 10, 4, 4, 2, 0, 0, 3, 0, 0, 0, 3, 0, 7, 3, 6, 4, 4, 3, 5, 2, 7, 0, 7, 2, 7, 3, 9, 0, 9, 3, 10, 5, 10, 5, 10, 9};
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
    fl = k[0]*n[0]*n[10];  // H + e- -> H+ + e- + e-
    dn[0] -= fl;
    dn[10] -= fl;
    dn[1] += fl;
    dn[10] += fl;
    dn[10] += fl;
    fl = k[1]*n[1]*n[10];  // H+ + e- -> H
    dn[1] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    fl = k[2]*n[6]*n[10];  // He + e- -> He+ + e- + e-
    dn[6] -= fl;
    dn[10] -= fl;
    dn[7] += fl;
    dn[10] += fl;
    dn[10] += fl;
    fl = k[3]*n[7]*n[10];  // He+ + e- -> He
    dn[7] -= fl;
    dn[10] -= fl;
    dn[6] += fl;
    fl = k[4]*n[7]*n[10];  // He+ + e- -> He++ + e- + e-
    dn[7] -= fl;
    dn[10] -= fl;
    dn[8] += fl;
    dn[10] += fl;
    dn[10] += fl;
    fl = k[5]*n[8]*n[10];  // He++ + e- -> He+
    dn[8] -= fl;
    dn[10] -= fl;
    dn[7] += fl;
    fl = k[6]*n[0]*n[10];  // H + e- -> H-
    dn[0] -= fl;
    dn[10] -= fl;
    dn[2] += fl;
    fl = k[7]*n[2]*n[0];  // H- + H -> H2 + e-
    dn[2] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[10] += fl;
    fl = k[8]*n[2]*n[0];  // H- + H -> H2 + e-
    dn[2] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[10] += fl;
    fl = k[9]*n[0]*n[1];  // H + H+ -> H2+
    dn[0] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    fl = k[10]*n[4]*n[0];  // H2+ + H -> H2 + H+
    dn[4] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[1] += fl;
    fl = k[11]*n[3]*n[1];  // H2 + H+ -> H2+ + H
    dn[3] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    dn[0] += fl;
    fl = k[12]*n[3]*n[10];  // H2 + e- -> H + H + e-
    dn[3] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[10] += fl;
    fl = k[13]*n[3]*n[0];  // H2 + H -> H + H + H
    dn[3] -= fl;
    dn[0] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[14]*n[2]*n[10];  // H- + e- -> H + e- + e-
    dn[2] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    dn[10] += fl;
    dn[10] += fl;
    fl = k[15]*n[2]*n[1];  // H- + H+ -> H + H
    dn[2] -= fl;
    dn[1] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[16]*n[2]*n[1];  // H- + H+ -> H2+ + e-
    dn[2] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    dn[10] += fl;
    fl = k[17]*n[2]*n[1];  // H- + H+ -> H2+ + e-
    dn[2] -= fl;
    dn[1] -= fl;
    dn[4] += fl;
    dn[10] += fl;
    fl = k[18]*n[4]*n[10];  // H2+ + e- -> H + H
    dn[4] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[19]*n[4]*n[2];  // H2+ + H- -> H + H2
    dn[4] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[3] += fl;
    fl = k[20]*n[0]*n[0]*n[0];  // H + H + H -> H2 + H
    dn[0] -= fl;
    dn[0] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[0] += fl;
    fl = k[21]*n[3]*n[0]*n[0];  // H2 + H + H -> H2 + H2
    dn[3] -= fl;
    dn[0] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    dn[3] += fl;
    fl = k[22]*n[3]*n[3];  // H2 + H2 -> H + H + H2
    dn[3] -= fl;
    dn[3] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[3] += fl;
    fl = k[23]*n[0]*n[0];  // H + H -> H+ + e- + H
    dn[0] -= fl;
    dn[0] -= fl;
    dn[1] += fl;
    dn[10] += fl;
    dn[0] += fl;
    fl = k[24]*n[7]*n[3];  // He+ + H2 -> H+ + He + H
    dn[7] -= fl;
    dn[3] -= fl;
    dn[1] += fl;
    dn[6] += fl;
    dn[0] += fl;
    fl = k[25]*n[4]*n[6];  // H2+ + He -> HeH+ + H
    dn[4] -= fl;
    dn[6] -= fl;
    dn[9] += fl;
    dn[0] += fl;
    fl = k[26]*n[4]*n[3];  // H2+ + H2 -> H3+ + H
    dn[4] -= fl;
    dn[3] -= fl;
    dn[5] += fl;
    dn[0] += fl;
    fl = k[27]*n[5]*n[2];  // H3+ + H- -> H2 + H2
    dn[5] -= fl;
    dn[2] -= fl;
    dn[3] += fl;
    dn[3] += fl;
    fl = k[28]*n[7]*n[0];  // He+ + H -> H+ + He
    dn[7] -= fl;
    dn[0] -= fl;
    dn[1] += fl;
    dn[6] += fl;
    fl = k[29]*n[7]*n[2];  // He+ + H- -> H + He
    dn[7] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[6] += fl;
    fl = k[30]*n[7]*n[3];  // He+ + H2 -> H2+ + He
    dn[7] -= fl;
    dn[3] -= fl;
    dn[4] += fl;
    dn[6] += fl;
    fl = k[31]*n[9]*n[0];  // HeH+ + H -> H2+ + He
    dn[9] -= fl;
    dn[0] -= fl;
    dn[4] += fl;
    dn[6] += fl;
    fl = k[32]*n[9]*n[3];  // HeH+ + H2 -> H3+ + He
    dn[9] -= fl;
    dn[3] -= fl;
    dn[5] += fl;
    dn[6] += fl;
    fl = k[33]*n[5]*n[10];  // H3+ + e- -> H2 + H
    dn[5] -= fl;
    dn[10] -= fl;
    dn[3] += fl;
    dn[0] += fl;
    fl = k[34]*n[5]*n[10];  // H3+ + e- -> H + H + H
    dn[5] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    dn[0] += fl;
    dn[0] += fl;
    fl = k[35]*n[9]*n[10];  // HeH+ + e- -> H + He
    dn[9] -= fl;
    dn[10] -= fl;
    dn[0] += fl;
    dn[6] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[10];  // H + e- -> H+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[10 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[10 + 0 * ld] += d;
    J[10 + 0 * ld] += d;
    d = k[0]*n[0];  // H + e- -> H+ + e- + e-, d/d n[10]
    J[0 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[1 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    d = k[1]*n[10];  // H+ + e- -> H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[10 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[1]*n[1];  // H+ + e- -> H, d/d n[10]
    J[1 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    d = k[2]*n[10];  // He + e- -> He+ + e- + e-, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[10 + 6 * ld] -= d;
    J[7 + 6 * ld] += d;
    J[10 + 6 * ld] += d;
    J[10 + 6 * ld] += d;
    d = k[2]*n[6];  // He + e- -> He+ + e- + e-, d/d n[10]
    J[6 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[7 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    d = k[3]*n[10];  // He+ + e- -> He, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[10 + 7 * ld] -= d;
    J[6 + 7 * ld] += d;
    d = k[3]*n[7];  // He+ + e- -> He, d/d n[10]
    J[7 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[6 + 10 * ld] += d;
    d = k[4]*n[10];  // He+ + e- -> He++ + e- + e-, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[10 + 7 * ld] -= d;
    J[8 + 7 * ld] += d;
    J[10 + 7 * ld] += d;
    J[10 + 7 * ld] += d;
    d = k[4]*n[7];  // He+ + e- -> He++ + e- + e-, d/d n[10]
    J[7 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[8 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    d = k[5]*n[10];  // He++ + e- -> He+, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[10 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    d = k[5]*n[8];  // He++ + e- -> He+, d/d n[10]
    J[8 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[7 + 10 * ld] += d;
    d = k[6]*n[10];  // H + e- -> H-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[10 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    d = k[6]*n[0];  // H + e- -> H-, d/d n[10]
    J[0 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[2 + 10 * ld] += d;
    d = k[7]*n[2];  // H- + H -> H2 + e-, d/d n[0]
    J[2 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[10 + 0 * ld] += d;
    d = k[7]*n[0];  // H- + H -> H2 + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[8]*n[2];  // H- + H -> H2 + e-, d/d n[0]
    J[2 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[10 + 0 * ld] += d;
    d = k[8]*n[0];  // H- + H -> H2 + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[9]*n[1];  // H + H+ -> H2+, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] -= d;
    J[4 + 0 * ld] += d;
    d = k[9]*n[0];  // H + H+ -> H2+, d/d n[1]
    J[0 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    d = k[10]*n[4];  // H2+ + H -> H2 + H+, d/d n[0]
    J[4 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[1 + 0 * ld] += d;
    d = k[10]*n[0];  // H2+ + H -> H2 + H+, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[0 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[1 + 4 * ld] += d;
    d = k[11]*n[3];  // H2 + H+ -> H2+ + H, d/d n[1]
    J[3 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[11]*n[1];  // H2 + H+ -> H2+ + H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[1 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[12]*n[10];  // H2 + e- -> H + H + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[10 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[10 + 3 * ld] += d;
    d = k[12]*n[3];  // H2 + e- -> H + H + e-, d/d n[10]
    J[3 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    J[0 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    d = k[13]*n[3];  // H2 + H -> H + H + H, d/d n[0]
    J[3 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    d = k[13]*n[0];  // H2 + H -> H + H + H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[14]*n[10];  // H- + e- -> H + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[10 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[14]*n[2];  // H- + e- -> H + e- + e-, d/d n[10]
    J[2 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    J[10 + 10 * ld] += d;
    d = k[15]*n[2];  // H- + H+ -> H + H, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[15]*n[1];  // H- + H+ -> H + H, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    d = k[16]*n[2];  // H- + H+ -> H2+ + e-, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    J[10 + 1 * ld] += d;
    d = k[16]*n[1];  // H- + H+ -> H2+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[17]*n[2];  // H- + H+ -> H2+ + e-, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[4 + 1 * ld] += d;
    J[10 + 1 * ld] += d;
    d = k[17]*n[1];  // H- + H+ -> H2+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    J[10 + 2 * ld] += d;
    d = k[18]*n[10];  // H2+ + e- -> H + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[10 + 4 * ld] -= d;
    J[0 + 4 * ld] += d;
    J[0 + 4 * ld] += d;
    d = k[18]*n[4];  // H2+ + e- -> H + H, d/d n[10]
    J[4 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    J[0 + 10 * ld] += d;
    d = k[19]*n[4];  // H2+ + H- -> H + H2, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[19]*n[2];  // H2+ + H- -> H + H2, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[0 + 4 * ld] += d;
    J[3 + 4 * ld] += d;
    d = k[20]*n[0]*n[0]*R(3.0);  // H + H + H -> H2 + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    d = k[21]*n[3]*n[0]*R(2.0);  // H2 + H + H -> H2 + H2, d/d n[0]
    J[3 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    J[3 + 0 * ld] += d;
    d = k[21]*n[0]*n[0];  // H2 + H + H -> H2 + H2, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[3 + 3 * ld] += d;
    J[3 + 3 * ld] += d;
    d = k[22]*n[3]*R(2.0);  // H2 + H2 -> H + H + H2, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    J[3 + 3 * ld] += d;
    d = k[23]*n[0]*R(2.0);  // H + H -> H+ + e- + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[10 + 0 * ld] += d;
    J[0 + 0 * ld] += d;
    d = k[24]*n[7];  // He+ + H2 -> H+ + He + H, d/d n[3]
    J[7 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[1 + 3 * ld] += d;
    J[6 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[24]*n[3];  // He+ + H2 -> H+ + He + H, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[3 + 7 * ld] -= d;
    J[1 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    J[0 + 7 * ld] += d;
    d = k[25]*n[6];  // H2+ + He -> HeH+ + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[6 + 4 * ld] -= d;
    J[9 + 4 * ld] += d;
    J[0 + 4 * ld] += d;
    d = k[25]*n[4];  // H2+ + He -> HeH+ + H, d/d n[6]
    J[4 + 6 * ld] -= d;
    J[6 + 6 * ld] -= d;
    J[9 + 6 * ld] += d;
    J[0 + 6 * ld] += d;
    d = k[26]*n[4];  // H2+ + H2 -> H3+ + H, d/d n[3]
    J[4 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[5 + 3 * ld] += d;
    J[0 + 3 * ld] += d;
    d = k[26]*n[3];  // H2+ + H2 -> H3+ + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[0 + 4 * ld] += d;
    d = k[27]*n[5];  // H3+ + H- -> H2 + H2, d/d n[2]
    J[5 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[27]*n[2];  // H3+ + H- -> H2 + H2, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[2 + 5 * ld] -= d;
    J[3 + 5 * ld] += d;
    J[3 + 5 * ld] += d;
    d = k[28]*n[7];  // He+ + H -> H+ + He, d/d n[0]
    J[7 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[6 + 0 * ld] += d;
    d = k[28]*n[0];  // He+ + H -> H+ + He, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[0 + 7 * ld] -= d;
    J[1 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    d = k[29]*n[7];  // He+ + H- -> H + He, d/d n[2]
    J[7 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[6 + 2 * ld] += d;
    d = k[29]*n[2];  // He+ + H- -> H + He, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[2 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    d = k[30]*n[7];  // He+ + H2 -> H2+ + He, d/d n[3]
    J[7 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[6 + 3 * ld] += d;
    d = k[30]*n[3];  // He+ + H2 -> H2+ + He, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[3 + 7 * ld] -= d;
    J[4 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    d = k[31]*n[9];  // HeH+ + H -> H2+ + He, d/d n[0]
    J[9 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[4 + 0 * ld] += d;
    J[6 + 0 * ld] += d;
    d = k[31]*n[0];  // HeH+ + H -> H2+ + He, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[0 + 9 * ld] -= d;
    J[4 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[32]*n[9];  // HeH+ + H2 -> H3+ + He, d/d n[3]
    J[9 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[5 + 3 * ld] += d;
    J[6 + 3 * ld] += d;
    d = k[32]*n[3];  // HeH+ + H2 -> H3+ + He, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[3 + 9 * ld] -= d;
    J[5 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[33]*n[10];  // H3+ + e- -> H2 + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[10 + 5 * ld] -= d;
    J[3 + 5 * ld] += d;
    J[0 + 5 * ld] += d;
    d = k[33]*n[5];  // H3+ + e- -> H2 + H, d/d n[10]
    J[5 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[3 + 10 * ld] += d;
    J[0 + 10 * ld] += d;
    d = k[34]*n[10];  // H3+ + e- -> H + H + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[10 + 5 * ld] -= d;
    J[0 + 5 * ld] += d;
    J[0 + 5 * ld] += d;
    J[0 + 5 * ld] += d;
    d = k[34]*n[5];  // H3+ + e- -> H + H + H, d/d n[10]
    J[5 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    J[0 + 10 * ld] += d;
    J[0 + 10 * ld] += d;
    d = k[35]*n[10];  // HeH+ + e- -> H + He, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] -= d;
    J[0 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[35]*n[9];  // HeH+ + e- -> H + He, d/d n[10]
    J[9 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[0 + 10 * ld] += d;
    J[6 + 10 * ld] += d;
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
