// generated by netgen.py from primordial_photoH2.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_primordial_photoH2 {
  static constexpr const char* NAME = "primordial_photoH2";
  static constexpr int NSPEC = 12;
  static constexpr int NREA = 44;
  static constexpr int NPAR = 3;
  static constexpr int IDX_D = 0;
  static constexpr int IDX_Dj = 1;
  static constexpr int IDX_H = 2;
  static constexpr int IDX_Hj = 3;
  static constexpr int IDX_Hk = 4;
  static constexpr int IDX_H2 = 5;
  static constexpr int IDX_H2j = 6;
  static constexpr int IDX_HD = 7;
  static constexpr int IDX_He = 8;
  static constexpr int IDX_Hej = 9;
  static constexpr int IDX_Hejj = 10;
  static constexpr int IDX_E = 11;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 11;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 8;
  // PHOTO heating: the energy sums sum_b sigmaE_qb Phi_b follow the
  // rate sums in the row (sigmaE: sigma times the photoelectron's
  // energy E - Emin_q, bin-averaged)
  static constexpr bool PHOTO_HEAT = false;
  static constexpr int NROW = NPAR + NPHOTO;
  // parameter row (param_names order), then the photo sums
  struct Params {
    double J21;
    double gamma_ad;
    double Tcmb;
    double ph[NPHOTO];
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.J21 = row[0];
    p.gamma_ad = row[1];
    p.Tcmb = row[2];
    for (int r = 0; r < NPHOTO; ++r) p.ph[r] = row[NPAR + r];
    return p;
  }
  static constexpr int photo_reaction[8] = {34, 35, 36, 37, 38, 39, 40, 41};
  static constexpr int photo_reactant[8] = {2, 0, 8, 9, 5, 5, 5, 4};
  static constexpr double photo_Emin[8] = {13.6, 13.6, 24.59, 54.42, 15.4, 16.5, 17.7, 0.755};
  static constexpr double photo_Emax[8] = {INFINITY, INFINITY, INFINITY, INFINITY, 16.5, 17.7, 30.0, 1000.0};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return kx::sigma_v96(E_eV, 4.298e-01, 5.475e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // H -> H+ + e-
      case 1: return kx::sigma_v96(E_eV, 4.298e-01, 5.475e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // D -> D+ + e-
      case 2: return kx::sigma_v96(E_eV, 1.361e+01, 9.492e+02, 1.469e+00, 3.188e+00, 2.039e+00, 4.434e-01, 2.136e+00);  // He -> He+ + e-
      case 3: return kx::sigma_v96(E_eV, 1.720e+00, 1.369e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // He+ -> He++ + e-
      case 4: return ((R(9.56e-17) * (E_eV / R(15.4))) - R(9.4e-17));  // H2 -> H2+ + e-
      case 5: return ((R(2.16e-17) * (E_eV / R(15.4))) - R(1.48e-17));  // H2 -> H2+ + e-
      case 6: return (R(1.51e-17) * kx::pow((E_eV / R(15.4)), (-R(2.71))));  // H2 -> H2+ + e-
      case 7: return ((R(2.11e-16) * kx::pow((E_eV - R(0.755)), R(1.5))) * kx::ipow<-3>(E_eV));  // H- -> H + e-
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
      case 9: return 1;
      case 10: return 2;
      case 11: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 3.3443256801323995e-24;
      case 1: return 3.343414741762353e-24;
      case 2: return 1.6738233791328e-24;
      case 3: return 1.6729124407627538e-24;
      case 4: return 1.6747343175028463e-24;
      case 5: return 3.3476467582656e-24;
      case 6: return 3.346735819895553e-24;
      case 7: return 5.0181490592651994e-24;
      case 8: return 6.64647366797316e-24;
      case 9: return 6.645562729603113e-24;
      case 10: return 6.644651791233067e-24;
      case 11: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 3;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 3 + e) {
      case 0: return 1;
      case 3: return 1;
      case 7: return 1;
      case 10: return 1;
      case 13: return 1;
      case 16: return 2;
      case 19: return 2;
      case 21: return 1;
      case 22: return 1;
      case 26: return 1;
      case 29: return 1;
      case 32: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 82 nonzeros, 86 in L + U with the fill, 359 flops per
  // factorization (dense: 1152); rows in pivot order
  static constexpr int SP_NNZ = 86;
  static constexpr int SP_FLOPS = 359;
  static constexpr int sp_perm[12] = {8, 9, 10, 6, 1, 0, 2, 3, 4, 5, 7, 11};
  static constexpr int sp_rp[13] = {0, 3, 7, 10, 16, 23, 31, 40, 49, 57, 66, 74, 86};
  static constexpr int sp_ci[86] = {0, 1, 11, 0, 1, 2, 11, 1, 2, 11, 3, 6, 7, 8, 9, 11, 4, 5, 6, 7, 9, 10, 11, 4, 5, 6, 7, 8, 9, 10, 11, 3, 4, 5, 6, 7, 8, 9, 10, 11, 3, 4, 5, 6, 7, 8, 9, 10, 11, 3, 5, 6, 7, 8, 9, 10, 11, 3, 4, 5, 6, 7, 8, 9, 10, 11, 4, 5, 6, 7, 8, 9, 10, 11, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
  static constexpr int sp_di[12] = {0, 4, 8, 10, 16, 24, 34, 44, 53, 63, 72, 85};
  // conserved combinations c_k . n (the nuclei of D, H, He, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 4;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 7: return 1;
      case 14: return 1;
      case 15: return 1;
      case 16: return 1;
      case 17: return 2;
      case 18: return 2;
      case 19: return 1;
      case 32: return 1;
      case 33: return 1;
      case 34: return 1;
      case 37: return 1;
      case 39: return 1;
      case 40: return -1;
      case 42: return 1;
      case 45: return 1;
      case 46: return 2;
      case 47: return -1;
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
    s += n[7] * R(1.0);
    s += n[8] * R(0.0);
    s += n[9] * R(0.0);
    s += n[10] * R(0.0);
    s += n[11] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(3.3443256801323995e-24);
    s += n[1] * R(3.343414741762353e-24);
    s += n[2] * R(1.6738233791328e-24);
    s += n[3] * R(1.6729124407627538e-24);
    s += n[4] * R(1.6747343175028463e-24);
    s += n[5] * R(3.3476467582656e-24);
    s += n[6] * R(3.346735819895553e-24);
    s += n[7] * R(5.0181490592651994e-24);
    s += n[8] * R(6.64647366797316e-24);
    s += n[9] * R(6.645562729603113e-24);
    s += n[10] * R(6.644651791233067e-24);
    s += n[11] * R(9.10938370046373e-28);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    {  // H + e- -> H+ + e- + e-: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(10.0)) k[0] = kx::exp((((((((((-R(32.71396786)) + (R(13.536556) * v_lnTe)) - (R(5.73932875) * kx::ipow<2>(v_lnTe))) + (R(1.56315498) * kx::ipow<3>(v_lnTe))) - (R(0.2877056) * kx::ipow<4>(v_lnTe))) + (R(0.0348255977) * kx::ipow<5>(v_lnTe))) - (R(0.00263197617) * kx::ipow<6>(v_lnTe))) + (R(0.000111954395) * kx::ipow<7>(v_lnTe))) - (R(2.03914985e-06) * kx::ipow<8>(v_lnTe))));  // H + e- -> H+ + e- + e-
    }
    {  // H+ + e- -> H: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(10.0) && Tgas < R(5500.0)) k[1] = (R(3.92e-13) * kx::pow(v_invTe, R(0.6353)));  // H+ + e- -> H
      if (Tgas >= R(5500.0)) k[2] = kx::exp(((((((((((-R(28.61303380689232)) - (R(0.7241125657826851) * v_lnTe)) - (R(0.02026044731984691) * kx::ipow<2>(v_lnTe))) - (R(0.002380861877349834) * kx::ipow<3>(v_lnTe))) - (R(0.0003212605213188796) * kx::ipow<4>(v_lnTe))) - (R(1.421502914054107e-05) * kx::ipow<5>(v_lnTe))) + (R(4.989108920299513e-06) * kx::ipow<6>(v_lnTe))) + (R(5.755614137575758e-07) * kx::ipow<7>(v_lnTe))) - (R(1.856767039775261e-08) * kx::ipow<8>(v_lnTe))) - (R(3.071135243196595e-09) * kx::ipow<9>(v_lnTe))));  // H+ + e- -> H
    }
    {  // He + e- -> He+ + e- + e-: 1 piece(s), zero below 9280.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(9280.0)) k[3] = kx::exp((((((((((-R(44.09864886)) + (R(23.91596563) * v_lnTe)) - (R(10.7532302) * kx::ipow<2>(v_lnTe))) + (R(3.05803875) * kx::ipow<3>(v_lnTe))) - (R(0.56851189) * kx::ipow<4>(v_lnTe))) + (R(0.0679539123) * kx::ipow<5>(v_lnTe))) - (R(0.0050090561) * kx::ipow<6>(v_lnTe))) + (R(0.000206723616) * kx::ipow<7>(v_lnTe))) - (R(3.64916141e-06) * kx::ipow<8>(v_lnTe))));  // He + e- -> He+ + e- + e-
    }
    {  // He+ + e- -> He: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      if (Tgas >= R(10.0) && Tgas < R(9280.0)) k[4] = (R(3.92e-13) * kx::pow(v_invTe, R(0.6353)));  // He+ + e- -> He
      if (Tgas >= R(9280.0)) k[5] = (((R(1.54e-09) * (R(1.0) + (R(0.3) / kx::exp((R(8.099328789667) * v_invTe))))) / (kx::exp((R(40.49664394833662) * v_invTe)) * kx::pow(v_Te, R(1.5)))) + (R(3.92e-13) / kx::pow(v_Te, R(0.6353))));  // He+ + e- -> He
    }
    {  // He+ + e- -> He++ + e- + e-: 1 piece(s), zero below 9280.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(9280.0)) k[6] = kx::exp((((((((((-R(68.71040990212)) + (R(43.93347632635) * v_lnTe)) - (R(18.48066993568) * kx::ipow<2>(v_lnTe))) + (R(4.701626486759002) * kx::ipow<3>(v_lnTe))) - (R(0.7692466334492) * kx::ipow<4>(v_lnTe))) + (R(0.08113042097303) * kx::ipow<5>(v_lnTe))) - (R(0.005324020628287001) * kx::ipow<6>(v_lnTe))) + (R(0.0001975705312221) * kx::ipow<7>(v_lnTe))) - (R(3.165581065665e-06) * kx::ipow<8>(v_lnTe))));  // He+ + e- -> He++ + e- + e-
    }
    {  // He++ + e- -> He+: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_sqrTgas = kx::sqrt(Tgas);
      if (Tgas >= R(10.0)) k[7] = (((R(3.36e-10) / v_sqrTgas) / kx::pow((Tgas / R(1000.0)), R(0.2))) / (R(1.0) + kx::pow((Tgas / R(1000000.0)), R(0.7))));  // He++ + e- -> He+
    }
    {  // H + e- -> H-: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      if (Tgas >= R(10.0)) k[8] = (R(6.77e-15) * kx::pow(v_Te, R(0.8779)));  // H + e- -> H-
    }
    {  // H- + H -> H2 + e-: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(10.0) && Tgas < R(1160.0)) k[9] = R(1.43e-09);  // H- + H -> H2 + e-
      if (Tgas >= R(1160.0)) k[10] = kx::exp(((((((((-R(20.06913897587003)) + (R(0.2289800603272916) * v_lnTe)) + (R(0.03599837721023835) * kx::ipow<2>(v_lnTe))) - (R(0.004555120027032095) * kx::ipow<3>(v_lnTe))) - (R(0.0003105115447124016) * kx::ipow<4>(v_lnTe))) + (R(0.0001073294010367247) * kx::ipow<5>(v_lnTe))) - (R(8.36671960467864e-06) * kx::ipow<6>(v_lnTe))) + (R(2.238306228891639e-07) * kx::ipow<7>(v_lnTe))));  // H- + H -> H2 + e-
    }
    {  // H + H+ -> H2+: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0) && Tgas < R(6700.0)) k[11] = (R(1.85e-23) * kx::pow(Tgas, R(1.8)));  // H + H+ -> H2+
      if (Tgas >= R(6700.0)) k[12] = (R(5.81e-16) * kx::pow((Tgas / R(56200.0)), ((-R(0.6657)) * kx::log10((Tgas / R(56200.0))))));  // H + H+ -> H2+
    }
    {  // H2+ + H -> H2 + H+: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0)) k[13] = R(6e-10);  // H2+ + H -> H2 + H+
    }
    {  // H2 + H+ -> H2+ + H: 1 piece(s), zero below 3480.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(3480.0)) k[14] = kx::exp((((((((((-R(24.24914687731536)) + (R(3.400824447095291) * v_lnTe)) - (R(3.898003964650152) * kx::ipow<2>(v_lnTe))) + (R(2.045587822403071) * kx::ipow<3>(v_lnTe))) - (R(0.5416182856220388) * kx::ipow<4>(v_lnTe))) + (R(0.0841077503763412) * kx::ipow<5>(v_lnTe))) - (R(0.007879026154483455) * kx::ipow<6>(v_lnTe))) + (R(0.0004138398421504563) * kx::ipow<7>(v_lnTe))) - (R(9.36345888928611e-06) * kx::ipow<8>(v_lnTe))));  // H2 + H+ -> H2+ + H
    }
    {  // H2 + e- -> H + H + e-: 1 piece(s), zero below 3480.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(3480.0)) k[15] = ((R(5.6e-11) * kx::exp(((-R(102124.0)) * v_invT))) * kx::sqrt(Tgas));  // H2 + e- -> H + H + e-
    }
    {  // H2 + H -> H + H + H: 1 piece(s), zero below 3480.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(3480.0)) k[16] = ((R(1.0670825e-10) * kx::pow(v_Te, R(2.012))) / (kx::exp((R(4.463) * v_invTe)) * kx::pow((R(1.0) + (R(0.2472) * v_Te)), R(3.512))));  // H2 + H -> H + H + H
    }
    {  // H- + e- -> H + e- + e-: 1 piece(s), zero below 464.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(464.0)) k[17] = kx::exp((((((((((-R(18.01849334273)) + (R(2.360852208681) * v_lnTe)) - (R(0.2827443061704) * kx::ipow<2>(v_lnTe))) + (R(0.01623316639567) * kx::ipow<3>(v_lnTe))) - (R(0.03365012031362999) * kx::ipow<4>(v_lnTe))) + (R(0.01178329782711) * kx::ipow<5>(v_lnTe))) - (R(0.001656194699504) * kx::ipow<6>(v_lnTe))) + (R(0.0001068275202678) * kx::ipow<7>(v_lnTe))) - (R(2.631285809207e-06) * kx::ipow<8>(v_lnTe))));  // H- + e- -> H + e- + e-
    }
    {  // H- + H -> H + H + e-: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(10.0) && Tgas < R(1160.0)) k[18] = (R(2.56e-09) * kx::pow(v_Te, R(1.78186)));  // H- + H -> H + H + e-
      if (Tgas >= R(1160.0)) k[19] = kx::exp(((((((((((-R(20.37260896533324)) + (R(1.139449335841631) * v_lnTe)) - (R(0.1421013521554148) * kx::ipow<2>(v_lnTe))) + (R(0.00846445538663) * kx::ipow<3>(v_lnTe))) - (R(0.0014327641212992) * kx::ipow<4>(v_lnTe))) + (R(0.0002012250284791) * kx::ipow<5>(v_lnTe))) + (R(8.66396324309e-05) * kx::ipow<6>(v_lnTe))) - (R(2.585009680264e-05) * kx::ipow<7>(v_lnTe))) + (R(2.4555011970392e-06) * kx::ipow<8>(v_lnTe))) - (R(8.06838246118e-08) * kx::ipow<9>(v_lnTe))));  // H- + H -> H + H + e-
    }
    {  // H- + H+ -> H + H: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_Te = (Tgas * R(8.617333262145179e-05));
      if (Tgas >= R(10.0)) k[20] = (R(6.5e-09) / kx::sqrt(v_Te));  // H- + H+ -> H + H
    }
    {  // H- + H+ -> H2+ + e-: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0)) k[21] = (R(1e-08) * kx::pow(Tgas, (-R(0.4))));  // H- + H+ -> H2+ + e-
    }
    {  // H2+ + e- -> H + H: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0) && Tgas < R(617.0)) k[22] = R(1e-08);  // H2+ + e- -> H + H
      if (Tgas >= R(617.0)) k[23] = (R(1.32e-06) * kx::pow(Tgas, (-R(0.76))));  // H2+ + e- -> H + H
    }
    {  // H2+ + H- -> H + H2: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(10.0)) k[24] = (R(5e-07) * kx::sqrt((R(100.0) * v_invT)));  // H2+ + H- -> H + H2
    }
    {  // H+ + D -> H + D+: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(10.0)) k[25] = (R(1e-09) * kx::exp(((-R(41.0)) * v_invT)));  // H+ + D -> H + D+
    }
    {  // H + D+ -> H+ + D: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0)) k[26] = R(1e-09);  // H + D+ -> H+ + D
    }
    {  // H2 + D+ -> HD + H+: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(10.0)) k[27] = R(2.1e-09);  // H2 + D+ -> HD + H+
    }
    {  // HD + H+ -> H2 + D+: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(10.0)) k[28] = (R(1e-09) * kx::exp(((-R(457.0)) * v_invT)));  // HD + H+ -> H2 + D+
    }
    {  // H2 + D -> HD + H: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(10.0)) k[29] = (R(7.5e-11) * kx::exp(((-R(3820.0)) * v_invT)));  // H2 + D -> HD + H
    }
    {  // HD + H -> H2 + D: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(10.0)) k[30] = (R(7.5e-11) * kx::exp(((-R(4240.0)) * v_invT)));  // HD + H -> H2 + D
    }
    {  // D + H- -> HD + e-: 1 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_T32 = (Tgas / R(300.0));
      if (Tgas >= R(10.0)) k[31] = (R(1.5e-09) * kx::pow(v_T32, (-R(0.1))));  // D + H- -> HD + e-
    }
    {  // H + H + H -> H2 + H: 2 piece(s), zero below 10.0, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_T32 = (Tgas / R(300.0));
      if (Tgas >= R(10.0) && Tgas < R(300.0)) k[32] = (R(1.3e-32) * kx::pow(v_T32, (-R(0.38))));  // H + H + H -> H2 + H
      if (Tgas >= R(300.0)) k[33] = (R(1.3e-32) * kx::ipow<-1>(v_T32));  // H + H + H -> H2 + H
    }
    k[34] = R(p.ph[0]);  // H -> H+ + e- (photo)
    k[35] = R(p.ph[1]);  // D -> D+ + e- (photo)
    k[36] = R(p.ph[2]);  // He -> He+ + e- (photo)
    k[37] = R(p.ph[3]);  // He+ -> He++ + e- (photo)
    k[38] = R(p.ph[4]);  // H2 -> H2+ + e- (photo)
    k[39] = R(p.ph[5]);  // H2 -> H2+ + e- (photo)
    k[40] = R(p.ph[6]);  // H2 -> H2+ + e- (photo)
    k[41] = R(p.ph[7]);  // H- -> H + e- (photo)
    k[42] = (R(9e-13) * R(p.J21));  // H2 -> H + H
    k[43] = (R(3.6e-12) * kx::pow((Tgas / R(300.0)), (-R(0.75))));  // D+ + e- -> D
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
  static constexpr int NREA_T = 44;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[45] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 67, 70, 71, 72, 73, 74, 75, 76, 77, 78, 79, 81};
  static constexpr int LT_KF_SP[81] = {2, 11, 3, 11, 3, 11, 8, 11, 9, 11, 9, 11, 9, 11, 10, 11, 2, 11, 4, 2, 4, 2, 2, 3, 2, 3, 6, 2, 5, 3, 5, 11, 5, 2, 4, 11, 4, 2, 4, 2, 4, 3, 4, 3, 6, 11, 6, 11, 6, 4, 3, 0, 2, 1, 5, 1, 7, 3, 5, 0, 7, 2, 0, 4, 2, 2, 2, 2, 2, 2, 2, 0, 8, 9, 5, 5, 5, 4, 5, 1, 11};
  static constexpr int LT_SC_OFF[45] = {0, 5, 8, 11, 16, 19, 22, 27, 30, 33, 37, 41, 44, 47, 51, 55, 60, 65, 70, 75, 80, 84, 88, 92, 96, 100, 104, 108, 112, 116, 120, 124, 128, 133, 138, 141, 144, 147, 150, 153, 156, 159, 162, 165, 168};
  static constexpr int LT_SC_ROW[168] = {2, 11, 3, 11, 11, 3, 11, 2, 3, 11, 2, 8, 11, 9, 11, 11, 9, 11, 8, 9, 11, 8, 9, 11, 10, 11, 11, 10, 11, 9, 2, 11, 4, 4, 2, 5, 11, 4, 2, 5, 11, 2, 3, 6, 2, 3, 6, 6, 2, 5, 3, 5, 3, 6, 2, 5, 11, 2, 2, 11, 5, 2, 2, 2, 2, 4, 11, 2, 11, 11, 4, 2, 2, 2, 11, 4, 2, 2, 2, 11, 4, 3, 2, 2, 4, 3, 6, 11, 6, 11, 2, 2, 6, 11, 2, 2, 6, 4, 2, 5, 3, 0, 2, 1, 2, 1, 3, 0, 5, 1, 7, 3, 7, 3, 5, 1, 5, 0, 7, 2, 7, 2, 5, 0, 0, 4, 7, 11, 2, 2, 2, 5, 2, 2, 2, 2, 5, 2, 2, 3, 11, 0, 1, 11, 8, 9, 11, 9, 10, 11, 5, 6, 11, 5, 6, 11, 5, 6, 11, 4, 2, 11, 5, 2, 2, 1, 11, 0};
  static constexpr int LT_SC_SGN[168] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, -1, 1};
  static constexpr int LT_JT_OFF[45] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 65, 66, 67, 68, 69, 70, 71, 72, 73, 74, 75, 77};
  static constexpr int LT_JT_COL[77] = {2, 11, 3, 11, 3, 11, 8, 11, 9, 11, 9, 11, 9, 11, 10, 11, 2, 11, 2, 4, 2, 4, 2, 3, 2, 3, 2, 6, 3, 5, 5, 11, 2, 5, 4, 11, 2, 4, 2, 4, 3, 4, 3, 4, 6, 11, 6, 11, 4, 6, 0, 3, 1, 2, 1, 5, 3, 7, 0, 5, 2, 7, 0, 4, 2, 2, 2, 0, 8, 9, 5, 5, 5, 4, 5, 1, 11};
  static constexpr int LT_JT_MULT[77] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[78] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 66, 68, 68, 68, 68, 68, 68, 68, 68, 68, 68, 69, 70};
  static constexpr int LT_JF_SP[70] = {11, 2, 11, 3, 11, 3, 11, 8, 11, 9, 11, 9, 11, 9, 11, 10, 11, 2, 4, 2, 4, 2, 3, 2, 3, 2, 6, 2, 5, 3, 11, 5, 5, 2, 11, 4, 4, 2, 4, 2, 4, 3, 4, 3, 11, 6, 11, 6, 6, 4, 3, 0, 2, 1, 5, 1, 7, 3, 5, 0, 7, 2, 4, 0, 2, 2, 2, 2, 11, 1};
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
    fl = k[0]*n[2]*n[11];  // H + e- -> H+ + e- + e-
    dn[2] -= fl;
    dn[11] -= fl;
    dn[3] += fl;
    dn[11] += fl;
    dn[11] += fl;
    fl = k[1]*n[3]*n[11];  // H+ + e- -> H
    dn[3] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    fl = k[2]*n[3]*n[11];  // H+ + e- -> H
    dn[3] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    fl = k[3]*n[8]*n[11];  // He + e- -> He+ + e- + e-
    dn[8] -= fl;
    dn[11] -= fl;
    dn[9] += fl;
    dn[11] += fl;
    dn[11] += fl;
    fl = k[4]*n[9]*n[11];  // He+ + e- -> He
    dn[9] -= fl;
    dn[11] -= fl;
    dn[8] += fl;
    fl = k[5]*n[9]*n[11];  // He+ + e- -> He
    dn[9] -= fl;
    dn[11] -= fl;
    dn[8] += fl;
    fl = k[6]*n[9]*n[11];  // He+ + e- -> He++ + e- + e-
    dn[9] -= fl;
    dn[11] -= fl;
    dn[10] += fl;
    dn[11] += fl;
    dn[11] += fl;
    fl = k[7]*n[10]*n[11];  // He++ + e- -> He+
    dn[10] -= fl;
    dn[11] -= fl;
    dn[9] += fl;
    fl = k[8]*n[2]*n[11];  // H + e- -> H-
    dn[2] -= fl;
    dn[11] -= fl;
    dn[4] += fl;
    fl = k[9]*n[4]*n[2];  // H- + H -> H2 + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[11] += fl;
    fl = k[10]*n[4]*n[2];  // H- + H -> H2 + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[11] += fl;
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
    fl = k[15]*n[5]*n[11];  // H2 + e- -> H + H + e-
    dn[5] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[11] += fl;
    fl = k[16]*n[5]*n[2];  // H2 + H -> H + H + H
    dn[5] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[17]*n[4]*n[11];  // H- + e- -> H + e- + e-
    dn[4] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    dn[11] += fl;
    dn[11] += fl;
    fl = k[18]*n[4]*n[2];  // H- + H -> H + H + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[11] += fl;
    fl = k[19]*n[4]*n[2];  // H- + H -> H + H + e-
    dn[4] -= fl;
    dn[2] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    dn[11] += fl;
    fl = k[20]*n[4]*n[3];  // H- + H+ -> H + H
    dn[4] -= fl;
    dn[3] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[21]*n[4]*n[3];  // H- + H+ -> H2+ + e-
    dn[4] -= fl;
    dn[3] -= fl;
    dn[6] += fl;
    dn[11] += fl;
    fl = k[22]*n[6]*n[11];  // H2+ + e- -> H + H
    dn[6] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[23]*n[6]*n[11];  // H2+ + e- -> H + H
    dn[6] -= fl;
    dn[11] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[24]*n[6]*n[4];  // H2+ + H- -> H + H2
    dn[6] -= fl;
    dn[4] -= fl;
    dn[2] += fl;
    dn[5] += fl;
    fl = k[25]*n[3]*n[0];  // H+ + D -> H + D+
    dn[3] -= fl;
    dn[0] -= fl;
    dn[2] += fl;
    dn[1] += fl;
    fl = k[26]*n[2]*n[1];  // H + D+ -> H+ + D
    dn[2] -= fl;
    dn[1] -= fl;
    dn[3] += fl;
    dn[0] += fl;
    fl = k[27]*n[5]*n[1];  // H2 + D+ -> HD + H+
    dn[5] -= fl;
    dn[1] -= fl;
    dn[7] += fl;
    dn[3] += fl;
    fl = k[28]*n[7]*n[3];  // HD + H+ -> H2 + D+
    dn[7] -= fl;
    dn[3] -= fl;
    dn[5] += fl;
    dn[1] += fl;
    fl = k[29]*n[5]*n[0];  // H2 + D -> HD + H
    dn[5] -= fl;
    dn[0] -= fl;
    dn[7] += fl;
    dn[2] += fl;
    fl = k[30]*n[7]*n[2];  // HD + H -> H2 + D
    dn[7] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[0] += fl;
    fl = k[31]*n[0]*n[4];  // D + H- -> HD + e-
    dn[0] -= fl;
    dn[4] -= fl;
    dn[7] += fl;
    dn[11] += fl;
    fl = k[32]*n[2]*n[2]*n[2];  // H + H + H -> H2 + H
    dn[2] -= fl;
    dn[2] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[2] += fl;
    fl = k[33]*n[2]*n[2]*n[2];  // H + H + H -> H2 + H
    dn[2] -= fl;
    dn[2] -= fl;
    dn[2] -= fl;
    dn[5] += fl;
    dn[2] += fl;
    fl = k[34]*n[2];  // H -> H+ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[11] += fl;
    fl = k[35]*n[0];  // D -> D+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[11] += fl;
    fl = k[36]*n[8];  // He -> He+ + e-
    dn[8] -= fl;
    dn[9] += fl;
    dn[11] += fl;
    fl = k[37]*n[9];  // He+ -> He++ + e-
    dn[9] -= fl;
    dn[10] += fl;
    dn[11] += fl;
    fl = k[38]*n[5];  // H2 -> H2+ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[11] += fl;
    fl = k[39]*n[5];  // H2 -> H2+ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[11] += fl;
    fl = k[40]*n[5];  // H2 -> H2+ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[11] += fl;
    fl = k[41]*n[4];  // H- -> H + e-
    dn[4] -= fl;
    dn[2] += fl;
    dn[11] += fl;
    fl = k[42]*n[5];  // H2 -> H + H
    dn[5] -= fl;
    dn[2] += fl;
    dn[2] += fl;
    fl = k[43]*n[1]*n[11];  // D+ + e- -> D
    dn[1] -= fl;
    dn[11] -= fl;
    dn[0] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[11];  // H + e- -> H+ + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[11 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[0]*n[2];  // H + e- -> H+ + e- + e-, d/d n[11]
    J[2 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[3 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    d = k[1]*n[11];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[11 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[1]*n[3];  // H+ + e- -> H, d/d n[11]
    J[3 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    d = k[2]*n[11];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[11 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[2]*n[3];  // H+ + e- -> H, d/d n[11]
    J[3 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    d = k[3]*n[11];  // He + e- -> He+ + e- + e-, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[11 + 8 * ld] -= d;
    J[9 + 8 * ld] += d;
    J[11 + 8 * ld] += d;
    J[11 + 8 * ld] += d;
    d = k[3]*n[8];  // He + e- -> He+ + e- + e-, d/d n[11]
    J[8 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    d = k[4]*n[11];  // He+ + e- -> He, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[11 + 9 * ld] -= d;
    J[8 + 9 * ld] += d;
    d = k[4]*n[9];  // He+ + e- -> He, d/d n[11]
    J[9 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[8 + 11 * ld] += d;
    d = k[5]*n[11];  // He+ + e- -> He, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[11 + 9 * ld] -= d;
    J[8 + 9 * ld] += d;
    d = k[5]*n[9];  // He+ + e- -> He, d/d n[11]
    J[9 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[8 + 11 * ld] += d;
    d = k[6]*n[11];  // He+ + e- -> He++ + e- + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[11 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[11 + 9 * ld] += d;
    J[11 + 9 * ld] += d;
    d = k[6]*n[9];  // He+ + e- -> He++ + e- + e-, d/d n[11]
    J[9 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    d = k[7]*n[11];  // He++ + e- -> He+, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[11 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[7]*n[10];  // He++ + e- -> He+, d/d n[11]
    J[10 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    d = k[8]*n[11];  // H + e- -> H-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[11 + 2 * ld] -= d;
    J[4 + 2 * ld] += d;
    d = k[8]*n[2];  // H + e- -> H-, d/d n[11]
    J[2 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[4 + 11 * ld] += d;
    d = k[9]*n[4];  // H- + H -> H2 + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[9]*n[2];  // H- + H -> H2 + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[10]*n[4];  // H- + H -> H2 + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[10]*n[2];  // H- + H -> H2 + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
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
    d = k[15]*n[11];  // H2 + e- -> H + H + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[11 + 5 * ld] -= d;
    J[2 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    J[11 + 5 * ld] += d;
    d = k[15]*n[5];  // H2 + e- -> H + H + e-, d/d n[11]
    J[5 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    J[2 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
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
    d = k[17]*n[11];  // H- + e- -> H + e- + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[11 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[17]*n[4];  // H- + e- -> H + e- + e-, d/d n[11]
    J[4 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    J[11 + 11 * ld] += d;
    d = k[18]*n[4];  // H- + H -> H + H + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[18]*n[2];  // H- + H -> H + H + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[19]*n[4];  // H- + H -> H + H + e-, d/d n[2]
    J[4 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[19]*n[2];  // H- + H -> H + H + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[2 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
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
    J[11 + 3 * ld] += d;
    d = k[21]*n[3];  // H- + H+ -> H2+ + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] -= d;
    J[6 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[22]*n[11];  // H2+ + e- -> H + H, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[11 + 6 * ld] -= d;
    J[2 + 6 * ld] += d;
    J[2 + 6 * ld] += d;
    d = k[22]*n[6];  // H2+ + e- -> H + H, d/d n[11]
    J[6 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    J[2 + 11 * ld] += d;
    d = k[23]*n[11];  // H2+ + e- -> H + H, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[11 + 6 * ld] -= d;
    J[2 + 6 * ld] += d;
    J[2 + 6 * ld] += d;
    d = k[23]*n[6];  // H2+ + e- -> H + H, d/d n[11]
    J[6 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[2 + 11 * ld] += d;
    J[2 + 11 * ld] += d;
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
    d = k[25]*n[3];  // H+ + D -> H + D+, d/d n[0]
    J[3 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    J[1 + 0 * ld] += d;
    d = k[25]*n[0];  // H+ + D -> H + D+, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[0 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    J[1 + 3 * ld] += d;
    d = k[26]*n[2];  // H + D+ -> H+ + D, d/d n[1]
    J[2 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[3 + 1 * ld] += d;
    J[0 + 1 * ld] += d;
    d = k[26]*n[1];  // H + D+ -> H+ + D, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[1 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    d = k[27]*n[5];  // H2 + D+ -> HD + H+, d/d n[1]
    J[5 + 1 * ld] -= d;
    J[1 + 1 * ld] -= d;
    J[7 + 1 * ld] += d;
    J[3 + 1 * ld] += d;
    d = k[27]*n[1];  // H2 + D+ -> HD + H+, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[1 + 5 * ld] -= d;
    J[7 + 5 * ld] += d;
    J[3 + 5 * ld] += d;
    d = k[28]*n[7];  // HD + H+ -> H2 + D+, d/d n[3]
    J[7 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[5 + 3 * ld] += d;
    J[1 + 3 * ld] += d;
    d = k[28]*n[3];  // HD + H+ -> H2 + D+, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[3 + 7 * ld] -= d;
    J[5 + 7 * ld] += d;
    J[1 + 7 * ld] += d;
    d = k[29]*n[5];  // H2 + D -> HD + H, d/d n[0]
    J[5 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[7 + 0 * ld] += d;
    J[2 + 0 * ld] += d;
    d = k[29]*n[0];  // H2 + D -> HD + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[0 + 5 * ld] -= d;
    J[7 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    d = k[30]*n[7];  // HD + H -> H2 + D, d/d n[2]
    J[7 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[0 + 2 * ld] += d;
    d = k[30]*n[2];  // HD + H -> H2 + D, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[2 + 7 * ld] -= d;
    J[5 + 7 * ld] += d;
    J[0 + 7 * ld] += d;
    d = k[31]*n[4];  // D + H- -> HD + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[4 + 0 * ld] -= d;
    J[7 + 0 * ld] += d;
    J[11 + 0 * ld] += d;
    d = k[31]*n[0];  // D + H- -> HD + e-, d/d n[4]
    J[0 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[7 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[32]*n[2]*n[2]*R(3.0);  // H + H + H -> H2 + H, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    d = k[33]*n[2]*n[2]*R(3.0);  // H + H + H -> H2 + H, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[5 + 2 * ld] += d;
    J[2 + 2 * ld] += d;
    d = k[34];  // H -> H+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[11 + 2 * ld] += d;
    d = k[35];  // D -> D+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[11 + 0 * ld] += d;
    d = k[36];  // He -> He+ + e-, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[9 + 8 * ld] += d;
    J[11 + 8 * ld] += d;
    d = k[37];  // He+ -> He++ + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[11 + 9 * ld] += d;
    d = k[38];  // H2 -> H2+ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[11 + 5 * ld] += d;
    d = k[39];  // H2 -> H2+ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[11 + 5 * ld] += d;
    d = k[40];  // H2 -> H2+ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[11 + 5 * ld] += d;
    d = k[41];  // H- -> H + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[2 + 4 * ld] += d;
    J[11 + 4 * ld] += d;
    d = k[42];  // H2 -> H + H, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[2 + 5 * ld] += d;
    J[2 + 5 * ld] += d;
    d = k[43]*n[11];  // D+ + e- -> D, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[11 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[43]*n[1];  // D+ + e- -> D, d/d n[11]
    J[1 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[0 + 11 * ld] += d;
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
