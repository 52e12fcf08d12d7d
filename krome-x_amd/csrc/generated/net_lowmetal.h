// generated by netgen.py from lowmetal.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_lowmetal {
  static constexpr const char* NAME = "lowmetal";
  static constexpr int NSPEC = 16;
  static constexpr int NREA = 51;
  static constexpr int NPAR = 2;
  static constexpr int IDX_C = 0;
  static constexpr int IDX_Cj = 1;
  static constexpr int IDX_H = 2;
  static constexpr int IDX_Hj = 3;
  static constexpr int IDX_Hk = 4;
  static constexpr int IDX_H2 = 5;
  static constexpr int IDX_H2j = 6;
  static constexpr int IDX_He = 7;
  static constexpr int IDX_Hej = 8;
  static constexpr int IDX_Hejj = 9;
  static constexpr int IDX_O = 10;
  static constexpr int IDX_Oj = 11;
  static constexpr int IDX_Si = 12;
  static constexpr int IDX_Sij = 13;
  static constexpr int IDX_Sijj = 14;
  static constexpr int IDX_E = 15;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 15;
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
      case 1: return 1;
      case 3: return 1;
      case 4: return -1;
      case 6: return 1;
      case 8: return 1;
      case 9: return 2;
      case 11: return 1;
      case 13: return 1;
      case 14: return 2;
      case 15: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.99447347289326e-23;
      case 1: return 1.9943823790562555e-23;
      case 2: return 1.6738233791328e-24;
      case 3: return 1.6729124407627538e-24;
      case 4: return 1.6747343175028463e-24;
      case 5: return 3.3476467582656e-24;
      case 6: return 3.346735819895553e-24;
      case 7: return 6.64647366797316e-24;
      case 8: return 6.645562729603113e-24;
      case 9: return 6.644651791233067e-24;
      case 10: return 2.65669645265334e-23;
      case 11: return 2.6566053588163353e-23;
      case 12: return 4.6636239685461e-23;
      case 13: return 4.6635328747090956e-23;
      case 14: return 4.663441780872091e-23;
      case 15: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 5;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 5 + e) {
      case 0: return 1;
      case 5: return 1;
      case 11: return 1;
      case 16: return 1;
      case 21: return 1;
      case 26: return 2;
      case 31: return 2;
      case 37: return 1;
      case 42: return 1;
      case 47: return 1;
      case 53: return 1;
      case 58: return 1;
      case 64: return 1;
      case 69: return 1;
      case 74: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 120 nonzeros, 127 in L + U with the fill, 506 flops per
  // factorization (dense: 2730); rows in pivot order
  static constexpr int SP_NNZ = 127;
  static constexpr int SP_FLOPS = 506;
  static constexpr int sp_perm[16] = {9, 14, 4, 5, 6, 7, 11, 10, 0, 1, 2, 3, 8, 12, 13, 15};
  static constexpr int sp_rp[17] = {0, 3, 8, 13, 19, 25, 31, 37, 43, 50, 57, 71, 85, 96, 103, 111, 127};
  static constexpr int sp_ci[127] = {0, 12, 15, 1, 10, 11, 14, 15, 2, 4, 10, 11, 15, 2, 3, 4, 10, 11, 15, 2, 3, 4, 10, 11, 15, 5, 7, 8, 12, 13, 15, 6, 7, 10, 11, 12, 15, 6, 7, 10, 11, 12, 15, 8, 9, 10, 11, 12, 13, 15, 8, 9, 10, 11, 12, 13, 15, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 0, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 9, 10, 11, 12, 13, 14, 15, 1, 9, 10, 11, 12, 13, 14, 15, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
  static constexpr int sp_di[16] = {0, 3, 8, 14, 21, 25, 31, 38, 43, 51, 65, 80, 92, 100, 109, 126};
  // conserved combinations c_k . n (the nuclei of C, H, He, O, Si, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 6;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 18: return 1;
      case 19: return 1;
      case 20: return 1;
      case 21: return 2;
      case 22: return 2;
      case 39: return 1;
      case 40: return 1;
      case 41: return 1;
      case 58: return 1;
      case 59: return 1;
      case 76: return 1;
      case 77: return 1;
      case 78: return 1;
      case 81: return 1;
      case 83: return 1;
      case 84: return -1;
      case 86: return 1;
      case 88: return 1;
      case 89: return 2;
      case 91: return 1;
      case 93: return 1;
      case 94: return 2;
      case 95: return -1;
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
    s += n[11];
    s += n[12];
    s += n[13];
    s += n[14];
    s += n[15];
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(1.0);
    s += n[3] * R(1.0);
    s += n[4] * R(1.0);
    s += n[5] * R(2.0);
    s += n[6] * R(2.0);
    s += n[7] * R(0.0);
    s += n[8] * R(0.0);
    s += n[9] * R(0.0);
    s += n[10] * R(0.0);
    s += n[11] * R(0.0);
    s += n[12] * R(0.0);
    s += n[13] * R(0.0);
    s += n[14] * R(0.0);
    s += n[15] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.99447347289326e-23);
    s += n[1] * R(1.9943823790562555e-23);
    s += n[2] * R(1.6738233791328e-24);
    s += n[3] * R(1.6729124407627538e-24);
    s += n[4] * R(1.6747343175028463e-24);
    s += n[5] * R(3.3476467582656e-24);
    s += n[6] * R(3.346735819895553e-24);
    s += n[7] * R(6.64647366797316e-24);
    s += n[8] * R(6.645562729603113e-24);
    s += n[9] * R(6.644651791233067e-24);
    s += n[10] * R(2.65669645265334e-23);
    s += n[11] * R(2.6566053588163353e-23);
    s += n[12] * R(4.6636239685461e-23);
    s += n[13] * R(4.6635328747090956e-23);
    s += n[14] * R(4.663441780872091e-23);
    s += n[15] * R(9.10938370046373e-28);
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
    {
      const R v_sqrTgas = kx::sqrt(Tgas);
      k[7] = (((R(3.36e-10) / v_sqrTgas) / kx::pow((Tgas / R(1000.0)), R(0.2))) / (R(1.0) + kx::pow((Tgas / R(1000000.0)), R(0.7))));  // He++ + e- -> He+
    }
    {
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      k[8] = (R(6.77e-15) * kx::pow(v_Te, R(0.8779)));  // H + e- -> H-
    }
    {  // H- + H -> H2 + e-: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas < R(1160.0)) k[9] = R(1.43e-09);  // H- + H -> H2 + e-
      if (Tgas >= R(1160.0)) k[10] = kx::exp(((((((((-R(20.06913897587003)) + (R(0.2289800603272916) * v_lnTe)) + (R(0.03599837721023835) * kx::ipow<2>(v_lnTe))) - (R(0.004555120027032095) * kx::ipow<3>(v_lnTe))) - (R(0.0003105115447124016) * kx::ipow<4>(v_lnTe))) + (R(0.0001073294010367247) * kx::ipow<5>(v_lnTe))) - (R(8.36671960467864e-06) * kx::ipow<6>(v_lnTe))) + (R(2.238306228891639e-07) * kx::ipow<7>(v_lnTe))));  // H- + H -> H2 + e-
    }
    {  // H + H+ -> H2+: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(6700.0)) k[11] = (R(1.85e-23) * kx::pow(Tgas, R(1.8)));  // H + H+ -> H2+
      if (Tgas >= R(6700.0)) k[12] = (R(5.81e-16) * kx::pow((Tgas / R(56200.0)), ((-R(0.6657)) * kx::log10((Tgas / R(56200.0))))));  // H + H+ -> H2+
    }
    k[13] = R(6e-10);  // H2+ + H -> H2 + H+
    {  // H2 + H+ -> H2+ + H: 1 piece(s), zero below 3480.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(3480.0)) k[14] = kx::exp((((((((((-R(24.24914687731536)) + (R(3.400824447095291) * v_lnTe)) - (R(3.898003964650152) * kx::ipow<2>(v_lnTe))) + (R(2.045587822403071) * kx::ipow<3>(v_lnTe))) - (R(0.5416182856220388) * kx::ipow<4>(v_lnTe))) + (R(0.0841077503763412) * kx::ipow<5>(v_lnTe))) - (R(0.007879026154483455) * kx::ipow<6>(v_lnTe))) + (R(0.0004138398421504563) * kx::ipow<7>(v_lnTe))) - (R(9.36345888928611e-06) * kx::ipow<8>(v_lnTe))));  // H2 + H+ -> H2+ + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[15] = ((R(5.6e-11) * kx::exp(((-R(102124.0)) * v_invT))) * kx::sqrt(Tgas));  // H2 + e- -> H + H + e-
    }
    {
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      k[16] = (((R(1.0670825e-10) * kx::pow(v_Te, R(2.012))) * kx::exp(((-R(4.463)) * v_invTe))) / kx::pow((R(1.0) + (R(0.2472) * v_Te)), R(3.512)));  // H2 + H -> H + H + H
    }
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
    {
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      k[20] = (R(6.5e-09) / kx::sqrt(v_Te));  // H- + H+ -> H + H
    }
    k[21] = (R(1e-08) * kx::pow(Tgas, (-R(0.4))));  // H- + H+ -> H2+ + e-
    {  // H2+ + e- -> H + H: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(617.0)) k[22] = R(1e-08);  // H2+ + e- -> H + H
      if (Tgas >= R(617.0)) k[23] = (R(1.32e-06) * kx::pow(Tgas, (-R(0.76))));  // H2+ + e- -> H + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[24] = (R(5e-07) * kx::sqrt((R(100.0) * v_invT)));  // H2+ + H- -> H + H2
    }
    k[25] = ((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5)))));  // H + H + H -> H2 + H
    k[26] = (((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5))))) / R(8.0));  // H2 + H + H -> H2 + H2
    {  // C+ + e- -> C: 2 piece(s), zero below 0.0, clamped above 21140.0
      R Tgas = Trate;
      if (Tgas > R(21140.0)) Tgas = R(21140.0);
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas < R(7950.0)) k[27] = (R(4.67e-12) * kx::pow(v_T32, (-R(0.6))));  // C+ + e- -> C
      if (Tgas >= R(7950.0)) k[28] = ((R(1.23e-17) * kx::pow(v_T32, R(2.49))) * kx::exp((R(21845.6) * v_invT)));  // C+ + e- -> C
    }
    {  // Si+ + e- -> Si: 2 piece(s), zero below 0.0, clamped above 10000.0
      R Tgas = Trate;
      if (Tgas > R(10000.0)) Tgas = R(10000.0);
      const R v_T32 = (Tgas / R(300.0));
      if (Tgas < R(2000.0)) k[29] = (R(7.5e-12) * kx::pow(v_T32, (-R(0.55))));  // Si+ + e- -> Si
      if (Tgas >= R(2000.0)) k[30] = (R(4.86e-12) * kx::pow(v_T32, (-R(0.32))));  // Si+ + e- -> Si
    }
    {  // O+ + e- -> O: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas < R(400.0)) k[31] = (R(1.3e-10) * kx::pow(Tgas, (-R(0.64))));  // O+ + e- -> O
      if (Tgas >= R(400.0)) k[32] = ((R(1.41e-10) * kx::pow(Tgas, (-R(0.66)))) + (((R(0.00074) * kx::pow(Tgas, (-R(1.5)))) * kx::exp(((-R(175000.0)) * v_invT))) * (R(1.0) + (R(0.062) * kx::exp(((-R(145000.0)) * v_invT))))));  // O+ + e- -> O
    }
    {
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_u1 = (R(11.26) * v_invTe);
      k[33] = (((R(6.85e-08) * kx::pow(v_u1, R(0.25))) * kx::exp((-v_u1))) / (R(0.193) + v_u1));  // C + e- -> C+ + e- + e-
    }
    {
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_u2 = (R(8.2) * v_invTe);
      k[34] = ((((R(1.88e-07) * (R(1.0) + kx::sqrt(v_u2))) * kx::pow(v_u2, R(0.25))) * kx::exp((-v_u2))) / (R(0.376) + v_u2));  // Si + e- -> Si+ + e- + e-
    }
    {
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_u3 = (R(13.6) * v_invTe);
      k[35] = (((R(3.59e-08) * kx::pow(v_u3, R(0.34))) * kx::exp((-v_u3))) / (R(0.073) + v_u3));  // O + e- -> O+ + e- + e-
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[36] = ((R(4.99e-11) * kx::pow(Tgas, R(0.405))) + (R(7.54e-10) * kx::pow(v_invT, R(0.458))));  // O+ + H -> O + H+
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[37] = (((R(1.08e-11) * kx::pow(Tgas, R(0.517))) + (R(4e-10) * kx::pow(Tgas, R(0.00669)))) * kx::exp(((-R(227.0)) * v_invT)));  // O + H+ -> O+ + H
    }
    k[38] = (((R(4.991e-15) * kx::pow((Tgas / R(10000.0)), R(0.3794))) * kx::exp(((-Tgas) / R(1121000.0)))) + ((R(2.78e-15) * kx::pow((Tgas / R(10000.0)), (-R(0.2163)))) * kx::exp(((-Tgas) / R(815800.0)))));  // O + He+ -> O+ + He
    k[39] = (R(3.9e-16) * kx::pow(Tgas, R(0.213)));  // C + H+ -> C+ + H
    {
      const R v_invT = (R(1.0) / Tgas);
      k[40] = ((R(6.08e-14) * kx::pow((Tgas / R(10000.0)), R(1.96))) * kx::exp(((-R(170000.0)) * v_invT)));  // C+ + H -> C + H+
    }
    {  // C + He+ -> C+ + He: 3 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(200.0)) k[41] = (R(8.58e-17) * kx::pow(Tgas, R(0.757)));  // C + He+ -> C+ + He
      if (Tgas >= R(200.0) && Tgas < R(2000.0)) k[42] = (R(3.25e-17) * kx::pow(Tgas, R(0.968)));  // C + He+ -> C+ + He
      if (Tgas >= R(2000.0)) k[43] = (R(2.77e-19) * kx::pow(Tgas, R(1.597)));  // C + He+ -> C+ + He
    }
    {  // Si + H+ -> Si+ + H: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      if (Tgas < R(10000.0)) k[44] = (R(5.88e-13) * kx::pow(Tgas, R(0.848)));  // Si + H+ -> Si+ + H
      if (Tgas >= R(10000.0)) k[45] = (R(1.45e-13) * Tgas);  // Si + H+ -> Si+ + H
    }
    k[46] = R(3.3e-09);  // Si + He+ -> Si+ + He
    k[47] = R(2.1e-09);  // C+ + Si -> C + Si+
    {
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      k[48] = (((R(4.1e-10) * kx::pow((Tgas / R(10000.0)), R(0.24))) * (R(1.0) + (R(3.17) * kx::exp((Tgas / R(2390000.0)))))) * kx::exp(((-R(3.178)) * v_invTe)));  // Si+ + H+ -> Si++ + H
    }
    k[49] = ((R(1.23e-09) * kx::pow((Tgas / R(10000.0)), R(0.24))) * (R(1.0) + (R(3.17) * kx::exp((Tgas / R(2390000.0))))));  // Si++ + H -> Si+ + H+
    k[50] = (R(1.75e-12) * kx::pow((Tgas / R(10000.0)), (-R(0.6346))));  // Si++ + e- -> Si+
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
  static constexpr int NREA_T = 51;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[52] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 53, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104};
  static constexpr int LT_KF_SP[104] = {2, 15, 3, 15, 3, 15, 7, 15, 8, 15, 8, 15, 8, 15, 9, 15, 2, 15, 4, 2, 4, 2, 2, 3, 2, 3, 6, 2, 5, 3, 5, 15, 5, 2, 4, 15, 4, 2, 4, 2, 4, 3, 4, 3, 6, 15, 6, 15, 6, 4, 2, 2, 2, 5, 2, 2, 1, 15, 1, 15, 13, 15, 13, 15, 11, 15, 11, 15, 0, 15, 12, 15, 10, 15, 11, 2, 10, 3, 10, 8, 0, 3, 1, 2, 0, 8, 0, 8, 0, 8, 12, 3, 12, 3, 12, 8, 1, 12, 13, 3, 14, 2, 14, 15};
  static constexpr int LT_SC_OFF[52] = {0, 5, 8, 11, 16, 19, 22, 27, 30, 33, 37, 41, 44, 47, 51, 55, 60, 65, 70, 75, 80, 84, 88, 92, 96, 100, 105, 110, 113, 116, 119, 122, 125, 128, 133, 138, 143, 147, 151, 155, 159, 163, 167, 171, 175, 179, 183, 187, 191, 195, 199, 202};
  static constexpr int LT_SC_ROW[202] = {2, 15, 3, 15, 15, 3, 15, 2, 3, 15, 2, 7, 15, 8, 15, 15, 8, 15, 7, 8, 15, 7, 8, 15, 9, 15, 15, 9, 15, 8, 2, 15, 4, 4, 2, 5, 15, 4, 2, 5, 15, 2, 3, 6, 2, 3, 6, 6, 2, 5, 3, 5, 3, 6, 2, 5, 15, 2, 2, 15, 5, 2, 2, 2, 2, 4, 15, 2, 15, 15, 4, 2, 2, 2, 15, 4, 2, 2, 2, 15, 4, 3, 2, 2, 4, 3, 6, 15, 6, 15, 2, 2, 6, 15, 2, 2, 6, 4, 2, 5, 2, 2, 2, 5, 2, 5, 2, 2, 5, 5, 1, 15, 0, 1, 15, 0, 13, 15, 12, 13, 15, 12, 11, 15, 10, 11, 15, 10, 0, 15, 1, 15, 15, 12, 15, 13, 15, 15, 10, 15, 11, 15, 15, 11, 2, 10, 3, 10, 3, 11, 2, 10, 8, 11, 7, 0, 3, 1, 2, 1, 2, 0, 3, 0, 8, 1, 7, 0, 8, 1, 7, 0, 8, 1, 7, 12, 3, 13, 2, 12, 3, 13, 2, 12, 8, 13, 7, 1, 12, 0, 13, 13, 3, 14, 2, 14, 2, 13, 3, 14, 15, 13};
  static constexpr int LT_SC_SGN[202] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1};
  static constexpr int LT_JT_OFF[52] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 51, 53, 55, 57, 59, 61, 63, 65, 67, 69, 71, 73, 75, 77, 79, 81, 83, 85, 87, 89, 91, 93, 95, 97, 99, 101};
  static constexpr int LT_JT_COL[101] = {2, 15, 3, 15, 3, 15, 7, 15, 8, 15, 8, 15, 8, 15, 9, 15, 2, 15, 2, 4, 2, 4, 2, 3, 2, 3, 2, 6, 3, 5, 5, 15, 2, 5, 4, 15, 2, 4, 2, 4, 3, 4, 3, 4, 6, 15, 6, 15, 4, 6, 2, 2, 5, 1, 15, 1, 15, 13, 15, 13, 15, 11, 15, 11, 15, 0, 15, 12, 15, 10, 15, 2, 11, 3, 10, 8, 10, 0, 3, 1, 2, 0, 8, 0, 8, 0, 8, 3, 12, 3, 12, 8, 12, 1, 12, 3, 13, 2, 14, 14, 15};
  static constexpr int LT_JT_MULT[101] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[102] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 52, 54, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72, 73, 74, 75, 76, 77, 78, 79, 80, 81, 82, 83, 84, 85, 86, 87, 88, 89, 90, 91, 92, 93, 94, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104};
  static constexpr int LT_JF_SP[104] = {15, 2, 15, 3, 15, 3, 15, 7, 15, 8, 15, 8, 15, 8, 15, 9, 15, 2, 4, 2, 4, 2, 3, 2, 3, 2, 6, 2, 5, 3, 15, 5, 5, 2, 15, 4, 4, 2, 4, 2, 4, 3, 4, 3, 15, 6, 15, 6, 6, 4, 2, 2, 5, 2, 2, 2, 15, 1, 15, 1, 15, 13, 15, 13, 15, 11, 15, 11, 15, 0, 15, 12, 15, 10, 11, 2, 10, 3, 10, 8, 3, 0, 2, 1, 8, 0, 8, 0, 8, 0, 12, 3, 12, 3, 12, 8, 12, 1, 13, 3, 14, 2, 15, 14};
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
    fl = k[0]*n[2]*n[15];  // H + e- -> H+ + e- + e-
    dn[2] -= fl;
    dn[15] -= fl;
    dn[3] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[1]*n[3]*n[15];  // H+ + e- -> H
    dn[3] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    fl = k[2]*n[3]*n[15];  // H+ + e- -> H
    dn[3] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    fl = k[3]*n[7]*n[15];  // He + e- -> He+ + e- + e-
    dn[7] -= fl;
    dn[15] -= fl;
    dn[8] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[4]*n[8]*n[15];  // He+ + e- -> He
    dn[8] -= fl;
    dn[15] -= fl;
    dn[7] += fl;
    fl = k[5]*n[8]*n[15];  // He+ + e- -> He
    dn[8] -= fl;
    dn[15] -= fl;
    dn[7] += fl;
    fl = k[6]*n[8]*n[15];  // He+ + e- -> He++ + e- + e-
    dn[8] -= fl;
    dn[15] -= fl;
    dn[9] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[7]*n[9]*n[15];  // He++ + e- -> He+
    dn[9] -= fl;
    dn[15] -= fl;
    dn[8] += fl;
    fl = k[8]*n[2]*n[15];  // H + e- -> H-
    dn[2] -= fl;
    dn[15] -= fl;
    dn[4] += fl;
    fl = k[9]*n[4]*n[2];  // H- + H -> H2 + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[15] += fl;
    fl = k[10]*n[4]*n[2];  // H- + H -> H2 + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[15] += fl;
    fl = k[11]*n[2]*n[3];  // H + H+ -> H2+
    dn[2] -= fl;
    dn[3] -= fl;
    dn[6] += fl;
    fl = k[12]*n[2]*n[3];  // H + H+ -> H2+
    dn[2] -= fl;
    dn[3] -= fl;
    dn[6] += fl;
    fl = k[13]*n[6]*n[2];  // H2+ + H -> H2 + H+
    dn[6] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[3] += fl;
    fl = k[14]*n[5]*n[3];  // H2 + H+ -> H2+ + H
    dn[5] -= fl;
    dn[3] -= fl;
    dn[6] += fl;
    dn[2] += fl;
    fl = k[15]*n[5]*n[15];  // H2 + e- -> H + H + e-
    dn[5] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[15] += fl;
    fl = k[16]*n[5]*n[2];  // H2 + H -> H + H + H
    dn[5] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[17]*n[4]*n[15];  // H- + e- -> H + e- + e-
    dn[4] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[18]*n[4]*n[2];  // H- + H -> H + H + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[15] += fl;
    fl = k[19]*n[4]*n[2];  // H- + H -> H + H + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[15] += fl;
    fl = k[20]*n[4]*n[3];  // H- + H+ -> H + H
    dn[4] -= fl;
    dn[3] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[21]*n[4]*n[3];  // H- + H+ -> H2+ + e-
    dn[4] -= fl;
    dn[3] -= fl;
    dn[6] += fl;
    dn[15] += fl;
    fl = k[22]*n[6]*n[15];  // H2+ + e- -> H + H
    dn[6] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[23]*n[6]*n[15];  // H2+ + e- -> H + H
    dn[6] -= fl;
    dn[15] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[24]*n[6]*n[4];  // H2+ + H- -> H + H2
    dn[6] -= fl;
    dn[4] -= fl;
    dn[2] += fl;
    dn[5] += fl;
    fl = k[25]*n[2]*n[2]*n[2];  // H + H + H -> H2 + H
    dn[2] -= fl;
    dn[2] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[2] += fl;
    fl = k[26]*n[5]*n[2]*n[2];  // H2 + H + H -> H2 + H2
    dn[5] -= fl;
    dn[2] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[5] += fl;
    fl = k[27]*n[1]*n[15];  // C+ + e- -> C
    dn[1] -= fl;
    dn[15] -= fl;
    dn[0] += fl;
    fl = k[28]*n[1]*n[15];  // C+ + e- -> C
    dn[1] -= fl;
    dn[15] -= fl;
    dn[0] += fl;
    fl = k[29]*n[13]*n[15];  // Si+ + e- -> Si
    dn[13] -= fl;
    dn[15] -= fl;
    dn[12] += fl;
    fl = k[30]*n[13]*n[15];  // Si+ + e- -> Si
    dn[13] -= fl;
    dn[15] -= fl;
    dn[12] += fl;
    fl = k[31]*n[11]*n[15];  // O+ + e- -> O
    dn[11] -= fl;
    dn[15] -= fl;
    dn[10] += fl;
    fl = k[32]*n[11]*n[15];  // O+ + e- -> O
    dn[11] -= fl;
    dn[15] -= fl;
    dn[10] += fl;
    fl = k[33]*n[0]*n[15];  // C + e- -> C+ + e- + e-
    dn[0] -= fl;
    dn[15] -= fl;
    dn[1] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[34]*n[12]*n[15];  // Si + e- -> Si+ + e- + e-
    dn[12] -= fl;
    dn[15] -= fl;
    dn[13] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[35]*n[10]*n[15];  // O + e- -> O+ + e- + e-
    dn[10] -= fl;
    dn[15] -= fl;
    dn[11] += fl;
    dn[15] += fl;
    dn[15] += fl;
    fl = k[36]*n[11]*n[2];  // O+ + H -> O + H+
    dn[11] -= fl;
    dn[2] -= fl;
    dn[10] += fl;
    dn[3] += fl;
    fl = k[37]*n[10]*n[3];  // O + H+ -> O+ + H
    dn[10] -= fl;
    dn[3] -= fl;
    dn[11] += fl;
    dn[2] += fl;
    fl = k[38]*n[10]*n[8];  // O + He+ -> O+ + He
    dn[10] -= fl;
    dn[8] -= fl;
    dn[11] += fl;
    dn[7] += fl;
    fl = k[39]*n[0]*n[3];  // C + H+ -> C+ + H
    dn[0] -= fl;
    dn[3] -= fl;
    dn[1] += fl;
    dn[2] += fl;
    fl = k[40]*n[1]*n[2];  // C+ + H -> C + H+
    dn[1] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[3] += fl;
    fl = k[41]*n[0]*n[8];  // C + He+ -> C+ + He
    dn[0] -= fl;
    dn[8] -= fl;
    dn[1] += fl;
    dn[7] += fl;
    fl = k[42]*n[0]*n[8];  // C + He+ -> C+ + He
    dn[0] -= fl;
    dn[8] -= fl;
    dn[1] += fl;
    dn[7] += fl;
    fl = k[43]*n[0]*n[8];  // C + He+ -> C+ + He
    dn[0] -= fl;
    dn[8] -= fl;
    dn[1] += fl;
    dn[7] += fl;
    fl = k[44]*n[12]*n[3];  // Si + H+ -> Si+ + H
    dn[12] -= fl;
    dn[3] -= fl;
    dn[13] += fl;
    dn[2] += fl;
    fl = k[45]*n[12]*n[3];  // Si + H+ -> Si+ + H
    dn[12] -= fl;
    dn[3] -= fl;
    dn[13] += fl;
    dn[2] += fl;
    fl = k[46]*n[12]*n[8];  // Si + He+ -> Si+ + He
    dn[12] -= fl;
    dn[8] -= fl;
    dn[13] += fl;
    dn[7] += fl;
    fl = k[47]*n[1]*n[12];  // C+ + Si -> C + Si+
    dn[1] -= fl;
    dn[12] -= fl;
    dn[0] += fl;
    dn[13] += fl;
    fl = k[48]*n[13]*n[3];  // Si+ + H+ -> Si++ + H
    dn[13] -= fl;
    dn[3] -= fl;
    dn[14] += fl;
    dn[2] += fl;
    fl = k[49]*n[14]*n[2];  // Si++ + H -> Si+ + H+
    dn[14] -= fl;
    dn[2] -= fl;
    dn[13] += fl;
    dn[3] += fl;
    fl = k[50]*n[14]*n[15];  // Si++ + e- -> Si+
    dn[14] -= fl;
    dn[15] -= fl;
    dn[13] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[15];  // H + e- -> H+ + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[15 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    d = k[0]*n[2];  // H + e- -> H+ + e- + e-, d/d n[15]
    J[2 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[3 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[1]*n[15];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[15 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[1]*n[3];  // H+ + e- -> H, d/d n[15]
    J[3 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    d = k[2]*n[15];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[15 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[2]*n[3];  // H+ + e- -> H, d/d n[15]
    J[3 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    d = k[3]*n[15];  // He + e- -> He+ + e- + e-, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[15 + 7 * ld] -= d;
    J[8 + 7 * ld] += d;
    J[15 + 7 * ld] += d;
    J[15 + 7 * ld] += d;
    d = k[3]*n[7];  // He + e- -> He+ + e- + e-, d/d n[15]
    J[7 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[8 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[4]*n[15];  // He+ + e- -> He, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[15 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    d = k[4]*n[8];  // He+ + e- -> He, d/d n[15]
    J[8 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[7 + 15 * ld] += d;
    d = k[5]*n[15];  // He+ + e- -> He, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[15 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    d = k[5]*n[8];  // He+ + e- -> He, d/d n[15]
    J[8 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[7 + 15 * ld] += d;
    d = k[6]*n[15];  // He+ + e- -> He++ + e- + e-, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[15 + 8 * ld] -= d;
    J[9 + 8 * ld] += d;
    J[15 + 8 * ld] += d;
    J[15 + 8 * ld] += d;
    d = k[6]*n[8];  // He+ + e- -> He++ + e- + e-, d/d n[15]
    J[8 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[9 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[7]*n[15];  // He++ + e- -> He+, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[15 + 9 * ld] -= d;
    J[8 + 9 * ld] += d;
    d = k[7]*n[9];  // He++ + e- -> He+, d/d n[15]
    J[9 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[8 + 15 * ld] += d;
    d = k[8]*n[15];  // H + e- -> H-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[15 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    d = k[8]*n[2];  // H + e- -> H-, d/d n[15]
    J[2 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[4 + 15 * ld] += d;
    d = k[9]*n[4];  // H- + H -> H2 + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    d = k[9]*n[2];  // H- + H -> H2 + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[10]*n[4];  // H- + H -> H2 + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    d = k[10]*n[2];  // H- + H -> H2 + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[11]*n[3];  // H + H+ -> H2+, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] -= d;
    J[6 + 2 * ld] += d;
    d = k[11]*n[2];  // H + H+ -> H2+, d/d n[3]
    J[2 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[6 + 3 * ld] += d;
    d = k[12]*n[3];  // H + H+ -> H2+, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] -= d;
    J[6 + 2 * ld] += d;
    d = k[12]*n[2];  // H + H+ -> H2+, d/d n[3]
    J[2 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[6 + 3 * ld] += d;
    d = k[13]*n[6];  // H2+ + H -> H2 + H+, d/d n[2]
    J[6 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[13]*n[2];  // H2+ + H -> H2 + H+, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[2 + 6 * ld] -= d;
    J[5 + 6 * ld] += d;
    J[3 + 6 * ld] += d;
    d = k[14]*n[5];  // H2 + H+ -> H2+ + H, d/d n[3]
    J[5 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[6 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[14]*n[3];  // H2 + H+ -> H2+ + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[3 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    d = k[15]*n[15];  // H2 + e- -> H + H + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[15 + 5 * ld] -= d;
    J[2 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    J[15 + 5 * ld] += d;
    d = k[15]*n[5];  // H2 + e- -> H + H + e-, d/d n[15]
    J[5 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    J[2 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[16]*n[5];  // H2 + H -> H + H + H, d/d n[2]
    J[5 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    d = k[16]*n[2];  // H2 + H -> H + H + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[2 + 5 * ld] -= d;
    J[2 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    d = k[17]*n[15];  // H- + e- -> H + e- + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[15 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[17]*n[4];  // H- + e- -> H + e- + e-, d/d n[15]
    J[4 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[18]*n[4];  // H- + H -> H + H + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    d = k[18]*n[2];  // H- + H -> H + H + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[19]*n[4];  // H- + H -> H + H + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[15 + 2 * ld] += d;
    d = k[19]*n[2];  // H- + H -> H + H + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[20]*n[4];  // H- + H+ -> H + H, d/d n[3]
    J[4 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[20]*n[3];  // H- + H+ -> H + H, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
    d = k[21]*n[4];  // H- + H+ -> H2+ + e-, d/d n[3]
    J[4 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[6 + 3 * ld] += d;
    J[15 + 3 * ld] += d;
    d = k[21]*n[3];  // H- + H+ -> H2+ + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] -= d;
    J[6 + 4 * ld] += d;
    J[15 + 4 * ld] += d;
    d = k[22]*n[15];  // H2+ + e- -> H + H, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[15 + 6 * ld] -= d;
    J[2 + 6 * ld] += d;
    J[2 + 6 * ld] += d;
    d = k[22]*n[6];  // H2+ + e- -> H + H, d/d n[15]
    J[6 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    J[2 + 15 * ld] += d;
    d = k[23]*n[15];  // H2+ + e- -> H + H, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[15 + 6 * ld] -= d;
    J[2 + 6 * ld] += d;
    J[2 + 6 * ld] += d;
    d = k[23]*n[6];  // H2+ + e- -> H + H, d/d n[15]
    J[6 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[2 + 15 * ld] += d;
    J[2 + 15 * ld] += d;
    d = k[24]*n[6];  // H2+ + H- -> H + H2, d/d n[4]
    J[6 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[5 + 4 * ld] += d;
    d = k[24]*n[4];  // H2+ + H- -> H + H2, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[4 + 6 * ld] -= d;
    J[2 + 6 * ld] += d;
    J[5 + 6 * ld] += d;
    d = k[25]*n[2]*n[2]*R(3.0);  // H + H + H -> H2 + H, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    d = k[26]*n[5]*n[2]*R(2.0);  // H2 + H + H -> H2 + H2, d/d n[2]
    J[5 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[5 + 2 * ld] += d;
    d = k[26]*n[2]*n[2];  // H2 + H + H -> H2 + H2, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[2 + 5 * ld] -= d;
    J[2 + 5 * ld] -= d;
    J[5 + 5 * ld] += d;
    J[5 + 5 * ld] += d;
    d = k[27]*n[15];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[15 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[27]*n[1];  // C+ + e- -> C, d/d n[15]
    J[1 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[0 + 15 * ld] += d;
    d = k[28]*n[15];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[15 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[28]*n[1];  // C+ + e- -> C, d/d n[15]
    J[1 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[0 + 15 * ld] += d;
    d = k[29]*n[15];  // Si+ + e- -> Si, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[15 + 13 * ld] -= d;
    J[12 + 13 * ld] += d;
    d = k[29]*n[13];  // Si+ + e- -> Si, d/d n[15]
    J[13 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[12 + 15 * ld] += d;
    d = k[30]*n[15];  // Si+ + e- -> Si, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[15 + 13 * ld] -= d;
    J[12 + 13 * ld] += d;
    d = k[30]*n[13];  // Si+ + e- -> Si, d/d n[15]
    J[13 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[12 + 15 * ld] += d;
    d = k[31]*n[15];  // O+ + e- -> O, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[15 + 11 * ld] -= d;
    J[10 + 11 * ld] += d;
    d = k[31]*n[11];  // O+ + e- -> O, d/d n[15]
    J[11 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[10 + 15 * ld] += d;
    d = k[32]*n[15];  // O+ + e- -> O, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[15 + 11 * ld] -= d;
    J[10 + 11 * ld] += d;
    d = k[32]*n[11];  // O+ + e- -> O, d/d n[15]
    J[11 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[10 + 15 * ld] += d;
    d = k[33]*n[15];  // C + e- -> C+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[15 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[15 + 0 * ld] += d;
    J[15 + 0 * ld] += d;
    d = k[33]*n[0];  // C + e- -> C+ + e- + e-, d/d n[15]
    J[0 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[1 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[34]*n[15];  // Si + e- -> Si+ + e- + e-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[15 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[15 + 12 * ld] += d;
    J[15 + 12 * ld] += d;
    d = k[34]*n[12];  // Si + e- -> Si+ + e- + e-, d/d n[15]
    J[12 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[13 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[35]*n[15];  // O + e- -> O+ + e- + e-, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[15 + 10 * ld] -= d;
    J[11 + 10 * ld] += d;
    J[15 + 10 * ld] += d;
    J[15 + 10 * ld] += d;
    d = k[35]*n[10];  // O + e- -> O+ + e- + e-, d/d n[15]
    J[10 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[11 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    J[15 + 15 * ld] += d;
    d = k[36]*n[11];  // O+ + H -> O + H+, d/d n[2]
    J[11 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[10 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[36]*n[2];  // O+ + H -> O + H+, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] -= d;
    J[10 + 11 * ld] += d;
    J[3 + 11 * ld] += d;
    d = k[37]*n[10];  // O + H+ -> O+ + H, d/d n[3]
    J[10 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[11 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[37]*n[3];  // O + H+ -> O+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[3 + 10 * ld] -= d;
    J[11 + 10 * ld] += d;
    J[2 + 10 * ld] += d;
    d = k[38]*n[10];  // O + He+ -> O+ + He, d/d n[8]
    J[10 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[11 + 8 * ld] += d;
    J[7 + 8 * ld] += d;
    d = k[38]*n[8];  // O + He+ -> O+ + He, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[8 + 10 * ld] -= d;
    J[11 + 10 * ld] += d;
    J[7 + 10 * ld] += d;
    d = k[39]*n[3];  // C + H+ -> C+ + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[2 + 0 * ld] += d;
    d = k[39]*n[0];  // C + H+ -> C+ + H, d/d n[3]
    J[0 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[1 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[40]*n[2];  // C+ + H -> C + H+, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[2 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[3 + 1 * ld] += d;
    d = k[40]*n[1];  // C+ + H -> C + H+, d/d n[2]
    J[1 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[41]*n[8];  // C + He+ -> C+ + He, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[8 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[7 + 0 * ld] += d;
    d = k[41]*n[0];  // C + He+ -> C+ + He, d/d n[8]
    J[0 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[1 + 8 * ld] += d;
    J[7 + 8 * ld] += d;
    d = k[42]*n[8];  // C + He+ -> C+ + He, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[8 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[7 + 0 * ld] += d;
    d = k[42]*n[0];  // C + He+ -> C+ + He, d/d n[8]
    J[0 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[1 + 8 * ld] += d;
    J[7 + 8 * ld] += d;
    d = k[43]*n[8];  // C + He+ -> C+ + He, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[8 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[7 + 0 * ld] += d;
    d = k[43]*n[0];  // C + He+ -> C+ + He, d/d n[8]
    J[0 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[1 + 8 * ld] += d;
    J[7 + 8 * ld] += d;
    d = k[44]*n[12];  // Si + H+ -> Si+ + H, d/d n[3]
    J[12 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[13 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[44]*n[3];  // Si + H+ -> Si+ + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[3 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[2 + 12 * ld] += d;
    d = k[45]*n[12];  // Si + H+ -> Si+ + H, d/d n[3]
    J[12 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[13 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[45]*n[3];  // Si + H+ -> Si+ + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[3 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[2 + 12 * ld] += d;
    d = k[46]*n[12];  // Si + He+ -> Si+ + He, d/d n[8]
    J[12 + 8 * ld] -= d;
    J[8 + 8 * ld] -= d;
    J[13 + 8 * ld] += d;
    J[7 + 8 * ld] += d;
    d = k[46]*n[8];  // Si + He+ -> Si+ + He, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[8 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[7 + 12 * ld] += d;
    d = k[47]*n[12];  // C+ + Si -> C + Si+, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[12 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    J[13 + 1 * ld] += d;
    d = k[47]*n[1];  // C+ + Si -> C + Si+, d/d n[12]
    J[1 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[0 + 12 * ld] += d;
    J[13 + 12 * ld] += d;
    d = k[48]*n[13];  // Si+ + H+ -> Si++ + H, d/d n[3]
    J[13 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[14 + 3 * ld] += d;
    J[2 + 3 * ld] += d;
    d = k[48]*n[3];  // Si+ + H+ -> Si++ + H, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[3 + 13 * ld] -= d;
    J[14 + 13 * ld] += d;
    J[2 + 13 * ld] += d;
    d = k[49]*n[14];  // Si++ + H -> Si+ + H+, d/d n[2]
    J[14 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[13 + 2 * ld] += d;
    J[3 + 2 * ld] += d;
    d = k[49]*n[2];  // Si++ + H -> Si+ + H+, d/d n[14]
    J[14 + 14 * ld] -= d;
    J[2 + 14 * ld] -= d;
    J[13 + 14 * ld] += d;
    J[3 + 14 * ld] += d;
    d = k[50]*n[15];  // Si++ + e- -> Si+, d/d n[14]
    J[14 + 14 * ld] -= d;
    J[15 + 14 * ld] -= d;
    J[13 + 14 * ld] += d;
    d = k[50]*n[14];  // Si++ + e- -> Si+, d/d n[15]
    J[14 + 15 * ld] -= d;
    J[15 + 15 * ld] -= d;
    J[13 + 15 * ld] += d;
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
