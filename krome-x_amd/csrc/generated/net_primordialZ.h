// generated by netgen.py from primordialZ.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_primordialZ {
  static constexpr const char* NAME = "primordialZ";
  static constexpr int NSPEC = 27;
  static constexpr int NREA = 74;
  static constexpr int NPAR = 2;
  static constexpr int IDX_C = 0;
  static constexpr int IDX_Cj = 1;
  static constexpr int IDX_CH = 2;
  static constexpr int IDX_CH2 = 3;
  static constexpr int IDX_CH3 = 4;
  static constexpr int IDX_CH4 = 5;
  static constexpr int IDX_CO = 6;
  static constexpr int IDX_COj = 7;
  static constexpr int IDX_CO2 = 8;
  static constexpr int IDX_H = 9;
  static constexpr int IDX_Hj = 10;
  static constexpr int IDX_Hk = 11;
  static constexpr int IDX_H2 = 12;
  static constexpr int IDX_H2j = 13;
  static constexpr int IDX_H2O = 14;
  static constexpr int IDX_H2Oj = 15;
  static constexpr int IDX_H3Oj = 16;
  static constexpr int IDX_He = 17;
  static constexpr int IDX_Hej = 18;
  static constexpr int IDX_Hejj = 19;
  static constexpr int IDX_O = 20;
  static constexpr int IDX_Oj = 21;
  static constexpr int IDX_O2 = 22;
  static constexpr int IDX_O2j = 23;
  static constexpr int IDX_OH = 24;
  static constexpr int IDX_OHj = 25;
  static constexpr int IDX_E = 26;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 26;
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
      case 7: return 1;
      case 10: return 1;
      case 11: return -1;
      case 13: return 1;
      case 15: return 1;
      case 16: return 1;
      case 18: return 1;
      case 19: return 2;
      case 21: return 1;
      case 23: return 1;
      case 25: return 1;
      case 26: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.99447347289326e-23;
      case 1: return 1.9943823790562555e-23;
      case 2: return 2.1618558108065398e-23;
      case 3: return 2.32923814871982e-23;
      case 4: return 2.4966204866331e-23;
      case 5: return 2.6640028245463796e-23;
      case 6: return 4.6511699255465995e-23;
      case 7: return 4.651078831709595e-23;
      case 8: return 7.30786637819994e-23;
      case 9: return 1.6738233791328e-24;
      case 10: return 1.6729124407627538e-24;
      case 11: return 1.6747343175028463e-24;
      case 12: return 3.3476467582656e-24;
      case 13: return 3.346735819895553e-24;
      case 14: return 2.9914611284799e-23;
      case 15: return 2.9913700346428956e-23;
      case 16: return 3.1587523725561754e-23;
      case 17: return 6.64647366797316e-24;
      case 18: return 6.645562729603113e-24;
      case 19: return 6.644651791233067e-24;
      case 20: return 2.65669645265334e-23;
      case 21: return 2.6566053588163353e-23;
      case 22: return 5.31339290530668e-23;
      case 23: return 5.313301811469676e-23;
      case 24: return 2.82407879056662e-23;
      case 25: return 2.823987696729615e-23;
      case 26: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 4;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 4 + e) {
      case 0: return 1;
      case 4: return 1;
      case 8: return 1;
      case 9: return 1;
      case 12: return 1;
      case 13: return 2;
      case 16: return 1;
      case 17: return 3;
      case 20: return 1;
      case 21: return 4;
      case 24: return 1;
      case 27: return 1;
      case 28: return 1;
      case 31: return 1;
      case 32: return 1;
      case 35: return 2;
      case 37: return 1;
      case 41: return 1;
      case 45: return 1;
      case 49: return 2;
      case 53: return 2;
      case 57: return 2;
      case 59: return 1;
      case 61: return 2;
      case 63: return 1;
      case 65: return 3;
      case 67: return 1;
      case 70: return 1;
      case 74: return 1;
      case 78: return 1;
      case 83: return 1;
      case 87: return 1;
      case 91: return 2;
      case 95: return 2;
      case 97: return 1;
      case 99: return 1;
      case 101: return 1;
      case 103: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 620 nonzeros, 636 in L + U with the fill, 8736 flops per
  // factorization (dense: 13122); rows in pivot order
  static constexpr int SP_NNZ = 636;
  static constexpr int SP_FLOPS = 8736;
  static constexpr int sp_perm[27] = {11, 13, 17, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 19, 20, 21, 22, 23, 24, 25, 26};
  static constexpr int sp_rp[28] = {0, 6, 12, 17, 44, 71, 98, 125, 152, 179, 206, 233, 260, 287, 314, 341, 368, 395, 422, 438, 447, 474, 501, 528, 555, 582, 609, 636};
  static constexpr int sp_ci[636] = {0, 1, 12, 13, 14, 26, 0, 1, 12, 13, 14, 26, 2, 12, 13, 18, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 2, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26};
  static constexpr int sp_di[27] = {0, 7, 12, 20, 48, 76, 104, 132, 160, 188, 216, 244, 272, 300, 328, 356, 384, 412, 429, 439, 467, 495, 523, 551, 579, 607, 635};
  // conserved combinations c_k . n (the nuclei of C, H, He, O, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 5;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 2: return 1;
      case 3: return 1;
      case 4: return 1;
      case 5: return 1;
      case 6: return 1;
      case 7: return 1;
      case 8: return 1;
      case 29: return 1;
      case 30: return 2;
      case 31: return 3;
      case 32: return 4;
      case 36: return 1;
      case 37: return 1;
      case 38: return 1;
      case 39: return 2;
      case 40: return 2;
      case 41: return 2;
      case 42: return 2;
      case 43: return 3;
      case 51: return 1;
      case 52: return 1;
      case 71: return 1;
      case 72: return 1;
      case 73: return 1;
      case 87: return 1;
      case 88: return 1;
      case 89: return 2;
      case 95: return 1;
      case 96: return 1;
      case 97: return 1;
      case 101: return 1;
      case 102: return 1;
      case 103: return 2;
      case 104: return 2;
      case 105: return 1;
      case 106: return 1;
      case 109: return 1;
      case 115: return 1;
      case 118: return 1;
      case 119: return -1;
      case 121: return 1;
      case 123: return 1;
      case 124: return 1;
      case 126: return 1;
      case 127: return 2;
      case 129: return 1;
      case 131: return 1;
      case 133: return 1;
      case 134: return -1;
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
    s += n[16];
    s += n[17];
    s += n[18];
    s += n[19];
    s += n[20];
    s += n[21];
    s += n[22];
    s += n[23];
    s += n[24];
    s += n[25];
    s += n[26];
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(1.0);
    s += n[3] * R(2.0);
    s += n[4] * R(3.0);
    s += n[5] * R(4.0);
    s += n[6] * R(0.0);
    s += n[7] * R(0.0);
    s += n[8] * R(0.0);
    s += n[9] * R(1.0);
    s += n[10] * R(1.0);
    s += n[11] * R(1.0);
    s += n[12] * R(2.0);
    s += n[13] * R(2.0);
    s += n[14] * R(2.0);
    s += n[15] * R(2.0);
    s += n[16] * R(3.0);
    s += n[17] * R(0.0);
    s += n[18] * R(0.0);
    s += n[19] * R(0.0);
    s += n[20] * R(0.0);
    s += n[21] * R(0.0);
    s += n[22] * R(0.0);
    s += n[23] * R(0.0);
    s += n[24] * R(1.0);
    s += n[25] * R(1.0);
    s += n[26] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.99447347289326e-23);
    s += n[1] * R(1.9943823790562555e-23);
    s += n[2] * R(2.1618558108065398e-23);
    s += n[3] * R(2.32923814871982e-23);
    s += n[4] * R(2.4966204866331e-23);
    s += n[5] * R(2.6640028245463796e-23);
    s += n[6] * R(4.6511699255465995e-23);
    s += n[7] * R(4.651078831709595e-23);
    s += n[8] * R(7.30786637819994e-23);
    s += n[9] * R(1.6738233791328e-24);
    s += n[10] * R(1.6729124407627538e-24);
    s += n[11] * R(1.6747343175028463e-24);
    s += n[12] * R(3.3476467582656e-24);
    s += n[13] * R(3.346735819895553e-24);
    s += n[14] * R(2.9914611284799e-23);
    s += n[15] * R(2.9913700346428956e-23);
    s += n[16] * R(3.1587523725561754e-23);
    s += n[17] * R(6.64647366797316e-24);
    s += n[18] * R(6.645562729603113e-24);
    s += n[19] * R(6.644651791233067e-24);
    s += n[20] * R(2.65669645265334e-23);
    s += n[21] * R(2.6566053588163353e-23);
    s += n[22] * R(5.31339290530668e-23);
    s += n[23] * R(5.313301811469676e-23);
    s += n[24] * R(2.82407879056662e-23);
    s += n[25] * R(2.823987696729615e-23);
    s += n[26] * R(9.10938370046373e-28);
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
    k[20] = ((((R(2.96e-06) / kx::sqrt(Tgas)) - R(1.73e-09)) + (R(2.5e-10) * kx::sqrt(Tgas))) - (R(7.77e-13) * Tgas));  // H- + H+ -> H + H
    k[21] = (R(1e-08) * kx::pow(Tgas, (-R(0.4))));  // H- + H+ -> H2+ + e-
    k[22] = (R(1000000.0) * (((((R(4.2278e-14) - (R(2.3088e-17) * Tgas)) + (R(7.3428e-21) * kx::ipow<2>(Tgas))) - (R(7.5474e-25) * kx::ipow<3>(Tgas))) + (R(3.3468e-29) * kx::ipow<4>(Tgas))) - (R(5.528e-34) * kx::ipow<5>(Tgas))));  // H2+ + e- -> H + H
    {
      const R v_invT = (R(1.0) / Tgas);
      k[23] = (R(5e-07) * kx::sqrt((R(100.0) * v_invT)));  // H2+ + H- -> H + H2
    }
    k[24] = ((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5)))));  // H + H + H -> H2 + H
    k[25] = (((R(6e-32) * kx::pow(Tgas, (-R(0.25)))) + (R(2e-31) * kx::pow(Tgas, (-R(0.5))))) / R(8.0));  // H2 + H + H -> H2 + H2
    k[27] = (R(1.2e-15) * kx::pow((Tgas / R(300.0)), R(0.25)));  // He+ + H -> He + H+
    {  // He + H+ -> He+ + H: 2 piece(s), zero below 0.0, clamped above INFINITY
      R Tgas = Trate;
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas < R(10000.0)) k[28] = ((R(1.26e-09) * kx::pow(Tgas, (-R(0.75)))) * kx::exp(((-R(127500.0)) * v_invT)));  // He + H+ -> He+ + H
      if (Tgas >= R(10000.0)) k[29] = (R(4e-37) * kx::pow(Tgas, R(4.74)));  // He + H+ -> He+ + H
    }
    k[71] = ((R(6.622e-11) / ((kx::sqrt((Tgas / R(4.136))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4.136)))), (R(1.0) - (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4214000.0)))), (R(1.0) + (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((R(5.629e-08) * kx::exp(((-R(5395.0)) / Tgas))) + (R(2.55e-07) * kx::exp(((-R(17700.0)) / Tgas)))) + (R(0.0006173) * kx::exp(((-R(167100.0)) / Tgas)))) + (R(0.0001627) * kx::exp(((-R(268700.0)) / Tgas))))));  // O+ + e- -> O
    k[72] = R(0.0);  // CO+ + e- -> CO
    k[73] = (R(2.75e-07) * kx::pow((Tgas / R(300.0)), (-R(0.55))));  // CO+ + e- -> C + O
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
    k[30] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(6.8e-10));  // O+ + H -> H+ + O
    {
      const R v_invT = (R(1.0) / Tgas);
      k[31] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(7e-10) * kx::exp(((-R(232.0)) * v_invT))));  // H+ + O -> O+ + H
    }
    k[32] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(1.7e-09));  // O+ + H2 -> OH+ + H
    k[33] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(1.01e-09));  // OH+ + H2 -> H2O+ + H
    k[34] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(8.3e-10));  // H2O+ + H2 -> H3O+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[35] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(2e-07) * kx::pow(v_T32, (-R(0.5)))));  // H2O+ + e- -> O + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[36] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(1.6e-07) * kx::pow(v_T32, (-R(0.5)))));  // H2O+ + e- -> OH + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[37] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(3.5e-07) * kx::pow(v_T32, (-R(0.5)))));  // H3O+ + e- -> H2O + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[38] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(9.9e-19) * kx::pow(v_T32, (-R(0.38)))));  // O + H -> OH
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[39] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(3.43e-13) * kx::pow(v_T32, R(2.67))) * kx::exp(((-R(3160.0)) * v_invT))));  // O + H2 -> OH + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[40] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(1.55e-12) * kx::pow(v_T32, R(1.6))) * kx::exp(((-R(1660.0)) * v_invT))));  // H2 + OH -> H2O + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[41] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(1.65e-12) * kx::pow(v_T32, R(1.14))) * kx::exp(((-R(50.0)) * v_invT))));  // OH + OH -> H2O + O
    }
    k[42] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(2.1e-09));  // H+ + OH -> OH+ + H
    k[43] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(6.9e-09));  // H+ + H2O -> H2O+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[44] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(7e-14) * kx::pow(v_T32, R(2.8))) * kx::exp(((-R(1950.0)) * v_invT))));  // H + OH -> H2 + O
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[45] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(6.83e-12) * kx::pow(v_T32, R(1.6))) * kx::exp(((-R(9720.0)) * v_invT))));  // H + H2O -> OH + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[46] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(4.9e-20) * kx::pow(v_T32, R(1.58))));  // O + O -> O2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[47] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(4.34e-11) * kx::pow(v_T32, (-R(0.5)))) * kx::exp(((-R(30.0)) * v_invT))));  // O + OH -> O2 + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[48] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(3.3e-10) * kx::exp(((-R(8460.0)) * v_invT))));  // H + O2 -> OH + O
    }
    k[49] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(2e-09));  // H+ + O2 -> O2+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[50] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(1.95e-07) * kx::pow(v_T32, (-R(0.7)))));  // O2+ + e- -> O + O
    }
    k[51] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(6.6e-11));  // O + CH -> CO + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[52] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(1.1e-10) * kx::sqrt(v_T32)));  // C + OH -> CO + H
    }
    k[53] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(3.3e-11));  // C + O2 -> CO + O
    k[54] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(6.2e-10));  // C+ + O2 -> O+ + CO
    k[55] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(1e-13));  // OH + CO -> CO2 + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[56] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(4.4e-12) * kx::pow(v_T32, (-R(0.61)))));  // C+ + e- -> C
    }
    k[57] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(7.7e-10));  // C+ + OH -> CO+ + H
    k[58] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(7.5e-10));  // CO+ + H -> H+ + CO
    k[59] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(1e-17));  // C + H -> CH
    {
      const R v_invT = (R(1.0) / Tgas);
      k[60] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(6.64e-10) * kx::exp(((-R(11700.0)) * v_invT))));  // C + H2 -> CH + H
    }
    k[61] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(4.98e-11));  // H + CH -> C + H2
    {
      const R v_invT = (R(1.0) / Tgas);
      k[62] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(2.38e-10) * kx::exp(((-R(1760.0)) * v_invT))));  // H2 + CH -> CH2 + H
    }
    k[63] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(2.7e-10));  // H + CH2 -> CH + H2
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[64] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(5.18e-11) * kx::pow(v_T32, R(0.17))) * kx::exp(((-R(6400.0)) * v_invT))));  // H2 + CH2 -> CH3 + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[65] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(1e-10) * kx::exp(((-R(7600.0)) * v_invT))));  // H + CH3 -> CH2 + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[66] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(6.86e-14) * kx::pow(v_T32, R(2.74))) * kx::exp(((-R(4740.0)) * v_invT))));  // H2 + CH3 -> CH4 + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[67] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : ((R(5.82e-13) * kx::ipow<3>(v_T32)) * kx::exp(((-R(4045.0)) * v_invT))));  // H + CH4 -> H2 + CH3
    }
    k[68] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : R(1e-17));  // H2 + C -> CH2
    {
      const R v_T32 = (Tgas / R(300.0));
      k[69] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(6.5e-07) * kx::pow(v_T32, (-R(0.5)))));  // H3O+ + e- -> OH + H + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[70] = ((ntot(n) > R(100000000000000.0)) ? R(0.0) : (R(6.3e-09) * kx::pow(v_T32, (-R(0.48)))));  // OH+ + e- -> O + H
    }
  }
  // every reaction: the T-only ones in k (zeroed) then the rest
  template <class R>
  KX_HD static void rates(R Trate, const R* n, const Params& p, R* k) {
    for (int i = 0; i < NREA; ++i) k[i] = R(0);
    rates_T(Trate, p, k);
    rates_n(Trate, n, p, k);
  }
  static constexpr int NREA_T = 31;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[75] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 51, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150};
  static constexpr int LT_KF_SP[150] = {9, 26, 10, 26, 10, 26, 17, 26, 18, 26, 18, 26, 18, 26, 19, 26, 9, 26, 11, 9, 9, 10, 9, 10, 13, 9, 12, 10, 12, 26, 12, 26, 12, 9, 11, 26, 11, 9, 11, 9, 11, 10, 11, 10, 13, 26, 13, 11, 9, 9, 9, 12, 9, 9, 12, 12, 18, 9, 17, 10, 17, 10, 21, 9, 10, 20, 21, 12, 25, 12, 15, 12, 15, 26, 15, 26, 16, 26, 20, 9, 20, 12, 12, 24, 24, 24, 10, 24, 10, 14, 9, 24, 9, 14, 20, 20, 20, 24, 9, 22, 10, 22, 23, 26, 20, 2, 0, 24, 0, 22, 1, 22, 24, 6, 1, 26, 1, 24, 7, 9, 0, 9, 0, 12, 9, 2, 12, 2, 9, 3, 12, 3, 9, 4, 12, 4, 9, 5, 12, 0, 16, 26, 25, 26, 21, 26, 7, 26, 7, 26};
  static constexpr int LT_SC_OFF[75] = {0, 5, 8, 11, 16, 19, 22, 27, 30, 33, 37, 40, 43, 47, 51, 55, 60, 65, 70, 75, 80, 84, 88, 92, 96, 101, 106, 111, 115, 119, 123, 127, 131, 135, 139, 143, 147, 151, 155, 158, 162, 166, 170, 174, 178, 182, 186, 189, 193, 197, 201, 205, 209, 213, 217, 221, 225, 228, 232, 236, 239, 243, 247, 251, 255, 259, 263, 267, 271, 274, 279, 283, 286, 289, 293};
  static constexpr int LT_SC_ROW[293] = {9, 26, 10, 26, 26, 10, 26, 9, 10, 26, 9, 17, 26, 18, 26, 26, 18, 26, 17, 18, 26, 17, 18, 26, 19, 26, 26, 19, 26, 18, 9, 26, 11, 11, 9, 12, 26, 9, 10, 13, 9, 10, 13, 13, 9, 12, 10, 12, 10, 13, 9, 12, 26, 9, 11, 12, 26, 9, 9, 26, 12, 9, 9, 9, 9, 11, 26, 9, 26, 26, 11, 9, 9, 9, 26, 11, 9, 9, 9, 26, 11, 10, 9, 9, 11, 10, 13, 26, 13, 26, 9, 9, 13, 11, 9, 12, 9, 9, 9, 12, 9, 12, 9, 9, 12, 12, 12, 12, 9, 9, 12, 18, 9, 17, 10, 17, 10, 18, 9, 17, 10, 18, 9, 21, 9, 10, 20, 10, 20, 21, 9, 21, 12, 25, 9, 25, 12, 15, 9, 15, 12, 16, 9, 15, 26, 20, 12, 15, 26, 24, 9, 16, 26, 14, 9, 20, 9, 24, 20, 12, 24, 9, 12, 24, 14, 9, 24, 24, 14, 20, 10, 24, 25, 9, 10, 14, 15, 9, 9, 24, 12, 20, 9, 14, 24, 12, 20, 20, 22, 20, 24, 22, 9, 9, 22, 24, 20, 10, 22, 23, 9, 23, 26, 20, 20, 20, 2, 6, 9, 0, 24, 6, 9, 0, 22, 6, 20, 1, 22, 21, 6, 24, 6, 8, 9, 1, 26, 0, 1, 24, 7, 9, 7, 9, 10, 6, 0, 9, 2, 0, 12, 2, 9, 9, 2, 0, 12, 12, 2, 3, 9, 9, 3, 2, 12, 12, 3, 4, 9, 9, 4, 3, 12, 12, 4, 5, 9, 9, 5, 12, 4, 12, 0, 3, 16, 26, 24, 9, 9, 25, 26, 20, 9, 21, 26, 20, 7, 26, 6, 7, 26, 0, 20};
  static constexpr int LT_SC_SGN[293] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1};
  static constexpr int LT_JT_OFF[75] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 49, 51, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 81, 83, 85, 87, 89, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144};
  static constexpr int LT_JT_COL[144] = {9, 26, 10, 26, 10, 26, 17, 26, 18, 26, 18, 26, 18, 26, 19, 26, 9, 26, 9, 11, 9, 10, 9, 10, 9, 13, 10, 12, 12, 26, 12, 26, 9, 12, 11, 26, 9, 11, 9, 11, 10, 11, 10, 11, 13, 26, 11, 13, 9, 9, 12, 12, 9, 18, 10, 17, 10, 17, 9, 21, 10, 20, 12, 21, 12, 25, 12, 15, 15, 26, 15, 26, 16, 26, 9, 20, 12, 20, 12, 24, 24, 10, 24, 10, 14, 9, 24, 9, 14, 20, 20, 24, 9, 22, 10, 22, 23, 26, 2, 20, 0, 24, 0, 22, 1, 22, 6, 24, 1, 26, 1, 24, 7, 9, 0, 9, 0, 12, 2, 9, 2, 12, 3, 9, 3, 12, 4, 9, 4, 12, 5, 9, 0, 12, 16, 26, 25, 26, 21, 26, 7, 26, 7, 26};
  static constexpr int LT_JT_MULT[144] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 2, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[145] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 50, 52, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72, 73, 74, 75, 76, 77, 78, 79, 80, 81, 82, 83, 84, 85, 86, 87, 88, 89, 90, 91, 92, 93, 94, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136, 137, 138, 139, 140, 141, 142, 143, 144, 145, 146, 147};
  static constexpr int LT_JF_SP[147] = {26, 9, 26, 10, 26, 10, 26, 17, 26, 18, 26, 18, 26, 18, 26, 19, 26, 9, 11, 9, 10, 9, 10, 9, 13, 9, 12, 10, 26, 12, 26, 12, 12, 9, 26, 11, 11, 9, 11, 9, 11, 10, 11, 10, 26, 13, 13, 11, 9, 9, 12, 9, 9, 9, 12, 18, 9, 17, 10, 17, 10, 21, 9, 20, 10, 21, 12, 25, 12, 15, 12, 26, 15, 26, 15, 26, 16, 20, 9, 20, 12, 24, 12, 24, 24, 10, 14, 10, 24, 9, 14, 9, 20, 24, 20, 22, 9, 22, 10, 26, 23, 20, 2, 24, 0, 22, 0, 22, 1, 24, 6, 26, 1, 24, 1, 9, 7, 9, 0, 12, 0, 9, 2, 12, 2, 9, 3, 12, 3, 9, 4, 12, 4, 9, 5, 12, 0, 26, 16, 26, 25, 26, 21, 26, 7, 26, 7};
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
    fl = k[0]*n[9]*n[26];  // H + e- -> H+ + e- + e-
    dn[9] -= fl;
    dn[26] -= fl;
    dn[10] += fl;
    dn[26] += fl;
    dn[26] += fl;
    fl = k[1]*n[10]*n[26];  // H+ + e- -> H
    dn[10] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    fl = k[2]*n[10]*n[26];  // H+ + e- -> H
    dn[10] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    fl = k[3]*n[17]*n[26];  // He + e- -> He+ + e- + e-
    dn[17] -= fl;
    dn[26] -= fl;
    dn[18] += fl;
    dn[26] += fl;
    dn[26] += fl;
    fl = k[4]*n[18]*n[26];  // He+ + e- -> He
    dn[18] -= fl;
    dn[26] -= fl;
    dn[17] += fl;
    fl = k[5]*n[18]*n[26];  // He+ + e- -> He
    dn[18] -= fl;
    dn[26] -= fl;
    dn[17] += fl;
    fl = k[6]*n[18]*n[26];  // He+ + e- -> He++ + e- + e-
    dn[18] -= fl;
    dn[26] -= fl;
    dn[19] += fl;
    dn[26] += fl;
    dn[26] += fl;
    fl = k[7]*n[19]*n[26];  // He++ + e- -> He+
    dn[19] -= fl;
    dn[26] -= fl;
    dn[18] += fl;
    fl = k[8]*n[9]*n[26];  // H + e- -> H-
    dn[9] -= fl;
    dn[26] -= fl;
    dn[11] += fl;
    fl = k[9]*n[11]*n[9];  // H- + H -> H2 + e-
    dn[11] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[26] += fl;
    fl = k[10]*n[9]*n[10];  // H + H+ -> H2+
    dn[9] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    fl = k[11]*n[9]*n[10];  // H + H+ -> H2+
    dn[9] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    fl = k[12]*n[13]*n[9];  // H2+ + H -> H2 + H+
    dn[13] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[10] += fl;
    fl = k[13]*n[12]*n[10];  // H2 + H+ -> H2+ + H
    dn[12] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    dn[9] += fl;
    fl = k[14]*n[12]*n[26];  // H2 + e- -> H + H-
    dn[12] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    dn[11] += fl;
    fl = k[15]*n[12]*n[26];  // H2 + e- -> H + H + e-
    dn[12] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[26] += fl;
    fl = k[16]*n[12]*n[9];  // H2 + H -> H + H + H
    dn[12] -= fl;
    dn[9] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[17]*n[11]*n[26];  // H- + e- -> H + e- + e-
    dn[11] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    dn[26] += fl;
    dn[26] += fl;
    fl = k[18]*n[11]*n[9];  // H- + H -> H + H + e-
    dn[11] -= fl;
    dn[9] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[26] += fl;
    fl = k[19]*n[11]*n[9];  // H- + H -> H + H + e-
    dn[11] -= fl;
    dn[9] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[26] += fl;
    fl = k[20]*n[11]*n[10];  // H- + H+ -> H + H
    dn[11] -= fl;
    dn[10] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[21]*n[11]*n[10];  // H- + H+ -> H2+ + e-
    dn[11] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    dn[26] += fl;
    fl = k[22]*n[13]*n[26];  // H2+ + e- -> H + H
    dn[13] -= fl;
    dn[26] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[23]*n[13]*n[11];  // H2+ + H- -> H + H2
    dn[13] -= fl;
    dn[11] -= fl;
    dn[9] += fl;
    dn[12] += fl;
    fl = k[24]*n[9]*n[9]*n[9];  // H + H + H -> H2 + H
    dn[9] -= fl;
    dn[9] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[9] += fl;
    fl = k[25]*n[12]*n[9]*n[9];  // H2 + H + H -> H2 + H2
    dn[12] -= fl;
    dn[9] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[12] += fl;
    fl = k[26]*n[12]*n[12];  // H2 + H2 -> H + H + H2
    dn[12] -= fl;
    dn[12] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[12] += fl;
    fl = k[27]*n[18]*n[9];  // He+ + H -> He + H+
    dn[18] -= fl;
    dn[9] -= fl;
    dn[17] += fl;
    dn[10] += fl;
    fl = k[28]*n[17]*n[10];  // He + H+ -> He+ + H
    dn[17] -= fl;
    dn[10] -= fl;
    dn[18] += fl;
    dn[9] += fl;
    fl = k[29]*n[17]*n[10];  // He + H+ -> He+ + H
    dn[17] -= fl;
    dn[10] -= fl;
    dn[18] += fl;
    dn[9] += fl;
    fl = k[30]*n[21]*n[9];  // O+ + H -> H+ + O
    dn[21] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[20] += fl;
    fl = k[31]*n[10]*n[20];  // H+ + O -> O+ + H
    dn[10] -= fl;
    dn[20] -= fl;
    dn[21] += fl;
    dn[9] += fl;
    fl = k[32]*n[21]*n[12];  // O+ + H2 -> OH+ + H
    dn[21] -= fl;
    dn[12] -= fl;
    dn[25] += fl;
    dn[9] += fl;
    fl = k[33]*n[25]*n[12];  // OH+ + H2 -> H2O+ + H
    dn[25] -= fl;
    dn[12] -= fl;
    dn[15] += fl;
    dn[9] += fl;
    fl = k[34]*n[15]*n[12];  // H2O+ + H2 -> H3O+ + H
    dn[15] -= fl;
    dn[12] -= fl;
    dn[16] += fl;
    dn[9] += fl;
    fl = k[35]*n[15]*n[26];  // H2O+ + e- -> O + H2
    dn[15] -= fl;
    dn[26] -= fl;
    dn[20] += fl;
    dn[12] += fl;
    fl = k[36]*n[15]*n[26];  // H2O+ + e- -> OH + H
    dn[15] -= fl;
    dn[26] -= fl;
    dn[24] += fl;
    dn[9] += fl;
    fl = k[37]*n[16]*n[26];  // H3O+ + e- -> H2O + H
    dn[16] -= fl;
    dn[26] -= fl;
    dn[14] += fl;
    dn[9] += fl;
    fl = k[38]*n[20]*n[9];  // O + H -> OH
    dn[20] -= fl;
    dn[9] -= fl;
    dn[24] += fl;
    fl = k[39]*n[20]*n[12];  // O + H2 -> OH + H
    dn[20] -= fl;
    dn[12] -= fl;
    dn[24] += fl;
    dn[9] += fl;
    fl = k[40]*n[12]*n[24];  // H2 + OH -> H2O + H
    dn[12] -= fl;
    dn[24] -= fl;
    dn[14] += fl;
    dn[9] += fl;
    fl = k[41]*n[24]*n[24];  // OH + OH -> H2O + O
    dn[24] -= fl;
    dn[24] -= fl;
    dn[14] += fl;
    dn[20] += fl;
    fl = k[42]*n[10]*n[24];  // H+ + OH -> OH+ + H
    dn[10] -= fl;
    dn[24] -= fl;
    dn[25] += fl;
    dn[9] += fl;
    fl = k[43]*n[10]*n[14];  // H+ + H2O -> H2O+ + H
    dn[10] -= fl;
    dn[14] -= fl;
    dn[15] += fl;
    dn[9] += fl;
    fl = k[44]*n[9]*n[24];  // H + OH -> H2 + O
    dn[9] -= fl;
    dn[24] -= fl;
    dn[12] += fl;
    dn[20] += fl;
    fl = k[45]*n[9]*n[14];  // H + H2O -> OH + H2
    dn[9] -= fl;
    dn[14] -= fl;
    dn[24] += fl;
    dn[12] += fl;
    fl = k[46]*n[20]*n[20];  // O + O -> O2
    dn[20] -= fl;
    dn[20] -= fl;
    dn[22] += fl;
    fl = k[47]*n[20]*n[24];  // O + OH -> O2 + H
    dn[20] -= fl;
    dn[24] -= fl;
    dn[22] += fl;
    dn[9] += fl;
    fl = k[48]*n[9]*n[22];  // H + O2 -> OH + O
    dn[9] -= fl;
    dn[22] -= fl;
    dn[24] += fl;
    dn[20] += fl;
    fl = k[49]*n[10]*n[22];  // H+ + O2 -> O2+ + H
    dn[10] -= fl;
    dn[22] -= fl;
    dn[23] += fl;
    dn[9] += fl;
    fl = k[50]*n[23]*n[26];  // O2+ + e- -> O + O
    dn[23] -= fl;
    dn[26] -= fl;
    dn[20] += fl;
    dn[20] += fl;
    fl = k[51]*n[20]*n[2];  // O + CH -> CO + H
    dn[20] -= fl;
    dn[2] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    fl = k[52]*n[0]*n[24];  // C + OH -> CO + H
    dn[0] -= fl;
    dn[24] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    fl = k[53]*n[0]*n[22];  // C + O2 -> CO + O
    dn[0] -= fl;
    dn[22] -= fl;
    dn[6] += fl;
    dn[20] += fl;
    fl = k[54]*n[1]*n[22];  // C+ + O2 -> O+ + CO
    dn[1] -= fl;
    dn[22] -= fl;
    dn[21] += fl;
    dn[6] += fl;
    fl = k[55]*n[24]*n[6];  // OH + CO -> CO2 + H
    dn[24] -= fl;
    dn[6] -= fl;
    dn[8] += fl;
    dn[9] += fl;
    fl = k[56]*n[1]*n[26];  // C+ + e- -> C
    dn[1] -= fl;
    dn[26] -= fl;
    dn[0] += fl;
    fl = k[57]*n[1]*n[24];  // C+ + OH -> CO+ + H
    dn[1] -= fl;
    dn[24] -= fl;
    dn[7] += fl;
    dn[9] += fl;
    fl = k[58]*n[7]*n[9];  // CO+ + H -> H+ + CO
    dn[7] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[6] += fl;
    fl = k[59]*n[0]*n[9];  // C + H -> CH
    dn[0] -= fl;
    dn[9] -= fl;
    dn[2] += fl;
    fl = k[60]*n[0]*n[12];  // C + H2 -> CH + H
    dn[0] -= fl;
    dn[12] -= fl;
    dn[2] += fl;
    dn[9] += fl;
    fl = k[61]*n[9]*n[2];  // H + CH -> C + H2
    dn[9] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[12] += fl;
    fl = k[62]*n[12]*n[2];  // H2 + CH -> CH2 + H
    dn[12] -= fl;
    dn[2] -= fl;
    dn[3] += fl;
    dn[9] += fl;
    fl = k[63]*n[9]*n[3];  // H + CH2 -> CH + H2
    dn[9] -= fl;
    dn[3] -= fl;
    dn[2] += fl;
    dn[12] += fl;
    fl = k[64]*n[12]*n[3];  // H2 + CH2 -> CH3 + H
    dn[12] -= fl;
    dn[3] -= fl;
    dn[4] += fl;
    dn[9] += fl;
    fl = k[65]*n[9]*n[4];  // H + CH3 -> CH2 + H2
    dn[9] -= fl;
    dn[4] -= fl;
    dn[3] += fl;
    dn[12] += fl;
    fl = k[66]*n[12]*n[4];  // H2 + CH3 -> CH4 + H
    dn[12] -= fl;
    dn[4] -= fl;
    dn[5] += fl;
    dn[9] += fl;
    fl = k[67]*n[9]*n[5];  // H + CH4 -> H2 + CH3
    dn[9] -= fl;
    dn[5] -= fl;
    dn[12] += fl;
    dn[4] += fl;
    fl = k[68]*n[12]*n[0];  // H2 + C -> CH2
    dn[12] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    fl = k[69]*n[16]*n[26];  // H3O+ + e- -> OH + H + H
    dn[16] -= fl;
    dn[26] -= fl;
    dn[24] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[70]*n[25]*n[26];  // OH+ + e- -> O + H
    dn[25] -= fl;
    dn[26] -= fl;
    dn[20] += fl;
    dn[9] += fl;
    fl = k[71]*n[21]*n[26];  // O+ + e- -> O
    dn[21] -= fl;
    dn[26] -= fl;
    dn[20] += fl;
    fl = k[72]*n[7]*n[26];  // CO+ + e- -> CO
    dn[7] -= fl;
    dn[26] -= fl;
    dn[6] += fl;
    fl = k[73]*n[7]*n[26];  // CO+ + e- -> C + O
    dn[7] -= fl;
    dn[26] -= fl;
    dn[0] += fl;
    dn[20] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[26];  // H + e- -> H+ + e- + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[26 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[26 + 9 * ld] += d;
    J[26 + 9 * ld] += d;
    d = k[0]*n[9];  // H + e- -> H+ + e- + e-, d/d n[26]
    J[9 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[10 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    d = k[1]*n[26];  // H+ + e- -> H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[26 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[1]*n[10];  // H+ + e- -> H, d/d n[26]
    J[10 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    d = k[2]*n[26];  // H+ + e- -> H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[26 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[2]*n[10];  // H+ + e- -> H, d/d n[26]
    J[10 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    d = k[3]*n[26];  // He + e- -> He+ + e- + e-, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[26 + 17 * ld] -= d;
    J[18 + 17 * ld] += d;
    J[26 + 17 * ld] += d;
    J[26 + 17 * ld] += d;
    d = k[3]*n[17];  // He + e- -> He+ + e- + e-, d/d n[26]
    J[17 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[18 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    d = k[4]*n[26];  // He+ + e- -> He, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[26 + 18 * ld] -= d;
    J[17 + 18 * ld] += d;
    d = k[4]*n[18];  // He+ + e- -> He, d/d n[26]
    J[18 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[17 + 26 * ld] += d;
    d = k[5]*n[26];  // He+ + e- -> He, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[26 + 18 * ld] -= d;
    J[17 + 18 * ld] += d;
    d = k[5]*n[18];  // He+ + e- -> He, d/d n[26]
    J[18 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[17 + 26 * ld] += d;
    d = k[6]*n[26];  // He+ + e- -> He++ + e- + e-, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[26 + 18 * ld] -= d;
    J[19 + 18 * ld] += d;
    J[26 + 18 * ld] += d;
    J[26 + 18 * ld] += d;
    d = k[6]*n[18];  // He+ + e- -> He++ + e- + e-, d/d n[26]
    J[18 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[19 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    d = k[7]*n[26];  // He++ + e- -> He+, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[26 + 19 * ld] -= d;
    J[18 + 19 * ld] += d;
    d = k[7]*n[19];  // He++ + e- -> He+, d/d n[26]
    J[19 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[18 + 26 * ld] += d;
    d = k[8]*n[26];  // H + e- -> H-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[26 + 9 * ld] -= d;
    J[11 + 9 * ld] += d;
    d = k[8]*n[9];  // H + e- -> H-, d/d n[26]
    J[9 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[11 + 26 * ld] += d;
    d = k[9]*n[11];  // H- + H -> H2 + e-, d/d n[9]
    J[11 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[26 + 9 * ld] += d;
    d = k[9]*n[9];  // H- + H -> H2 + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] -= d;
    J[12 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    d = k[10]*n[10];  // H + H+ -> H2+, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] -= d;
    J[13 + 9 * ld] += d;
    d = k[10]*n[9];  // H + H+ -> H2+, d/d n[10]
    J[9 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    d = k[11]*n[10];  // H + H+ -> H2+, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] -= d;
    J[13 + 9 * ld] += d;
    d = k[11]*n[9];  // H + H+ -> H2+, d/d n[10]
    J[9 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    d = k[12]*n[13];  // H2+ + H -> H2 + H+, d/d n[9]
    J[13 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[10 + 9 * ld] += d;
    d = k[12]*n[9];  // H2+ + H -> H2 + H+, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[9 + 13 * ld] -= d;
    J[12 + 13 * ld] += d;
    J[10 + 13 * ld] += d;
    d = k[13]*n[12];  // H2 + H+ -> H2+ + H, d/d n[10]
    J[12 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[13]*n[10];  // H2 + H+ -> H2+ + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[10 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[14]*n[26];  // H2 + e- -> H + H-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[26 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[11 + 12 * ld] += d;
    d = k[14]*n[12];  // H2 + e- -> H + H-, d/d n[26]
    J[12 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    J[11 + 26 * ld] += d;
    d = k[15]*n[26];  // H2 + e- -> H + H + e-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[26 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[26 + 12 * ld] += d;
    d = k[15]*n[12];  // H2 + e- -> H + H + e-, d/d n[26]
    J[12 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    d = k[16]*n[12];  // H2 + H -> H + H + H, d/d n[9]
    J[12 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[16]*n[9];  // H2 + H -> H + H + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[17]*n[26];  // H- + e- -> H + e- + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[26 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    d = k[17]*n[11];  // H- + e- -> H + e- + e-, d/d n[26]
    J[11 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    J[26 + 26 * ld] += d;
    d = k[18]*n[11];  // H- + H -> H + H + e-, d/d n[9]
    J[11 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[26 + 9 * ld] += d;
    d = k[18]*n[9];  // H- + H -> H + H + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[9 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    d = k[19]*n[11];  // H- + H -> H + H + e-, d/d n[9]
    J[11 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[26 + 9 * ld] += d;
    d = k[19]*n[9];  // H- + H -> H + H + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[9 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    d = k[20]*n[11];  // H- + H+ -> H + H, d/d n[10]
    J[11 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[20]*n[10];  // H- + H+ -> H + H, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[9 + 11 * ld] += d;
    d = k[21]*n[11];  // H- + H+ -> H2+ + e-, d/d n[10]
    J[11 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    J[26 + 10 * ld] += d;
    d = k[21]*n[10];  // H- + H+ -> H2+ + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] -= d;
    J[13 + 11 * ld] += d;
    J[26 + 11 * ld] += d;
    d = k[22]*n[26];  // H2+ + e- -> H + H, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[26 + 13 * ld] -= d;
    J[9 + 13 * ld] += d;
    J[9 + 13 * ld] += d;
    d = k[22]*n[13];  // H2+ + e- -> H + H, d/d n[26]
    J[13 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[9 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[23]*n[13];  // H2+ + H- -> H + H2, d/d n[11]
    J[13 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[12 + 11 * ld] += d;
    d = k[23]*n[11];  // H2+ + H- -> H + H2, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[11 + 13 * ld] -= d;
    J[9 + 13 * ld] += d;
    J[12 + 13 * ld] += d;
    d = k[24]*n[9]*n[9]*R(3.0);  // H + H + H -> H2 + H, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[25]*n[12]*n[9]*R(2.0);  // H2 + H + H -> H2 + H2, d/d n[9]
    J[12 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[25]*n[9]*n[9];  // H2 + H + H -> H2 + H2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[12 + 12 * ld] += d;
    J[12 + 12 * ld] += d;
    d = k[26]*n[12]*R(2.0);  // H2 + H2 -> H + H + H2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[12 + 12 * ld] += d;
    d = k[27]*n[18];  // He+ + H -> He + H+, d/d n[9]
    J[18 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[17 + 9 * ld] += d;
    J[10 + 9 * ld] += d;
    d = k[27]*n[9];  // He+ + H -> He + H+, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[9 + 18 * ld] -= d;
    J[17 + 18 * ld] += d;
    J[10 + 18 * ld] += d;
    d = k[28]*n[17];  // He + H+ -> He+ + H, d/d n[10]
    J[17 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[18 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[28]*n[10];  // He + H+ -> He+ + H, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[10 + 17 * ld] -= d;
    J[18 + 17 * ld] += d;
    J[9 + 17 * ld] += d;
    d = k[29]*n[17];  // He + H+ -> He+ + H, d/d n[10]
    J[17 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[18 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[29]*n[10];  // He + H+ -> He+ + H, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[10 + 17 * ld] -= d;
    J[18 + 17 * ld] += d;
    J[9 + 17 * ld] += d;
    d = k[30]*n[21];  // O+ + H -> H+ + O, d/d n[9]
    J[21 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[20 + 9 * ld] += d;
    d = k[30]*n[9];  // O+ + H -> H+ + O, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[9 + 21 * ld] -= d;
    J[10 + 21 * ld] += d;
    J[20 + 21 * ld] += d;
    d = k[31]*n[20];  // H+ + O -> O+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[20 + 10 * ld] -= d;
    J[21 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[31]*n[10];  // H+ + O -> O+ + H, d/d n[20]
    J[10 + 20 * ld] -= d;
    J[20 + 20 * ld] -= d;
    J[21 + 20 * ld] += d;
    J[9 + 20 * ld] += d;
    d = k[32]*n[21];  // O+ + H2 -> OH+ + H, d/d n[12]
    J[21 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[25 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[32]*n[12];  // O+ + H2 -> OH+ + H, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[12 + 21 * ld] -= d;
    J[25 + 21 * ld] += d;
    J[9 + 21 * ld] += d;
    d = k[33]*n[25];  // OH+ + H2 -> H2O+ + H, d/d n[12]
    J[25 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[15 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[33]*n[12];  // OH+ + H2 -> H2O+ + H, d/d n[25]
    J[25 + 25 * ld] -= d;
    J[12 + 25 * ld] -= d;
    J[15 + 25 * ld] += d;
    J[9 + 25 * ld] += d;
    d = k[34]*n[15];  // H2O+ + H2 -> H3O+ + H, d/d n[12]
    J[15 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[16 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[34]*n[12];  // H2O+ + H2 -> H3O+ + H, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[12 + 15 * ld] -= d;
    J[16 + 15 * ld] += d;
    J[9 + 15 * ld] += d;
    d = k[35]*n[26];  // H2O+ + e- -> O + H2, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[26 + 15 * ld] -= d;
    J[20 + 15 * ld] += d;
    J[12 + 15 * ld] += d;
    d = k[35]*n[15];  // H2O+ + e- -> O + H2, d/d n[26]
    J[15 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[20 + 26 * ld] += d;
    J[12 + 26 * ld] += d;
    d = k[36]*n[26];  // H2O+ + e- -> OH + H, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[26 + 15 * ld] -= d;
    J[24 + 15 * ld] += d;
    J[9 + 15 * ld] += d;
    d = k[36]*n[15];  // H2O+ + e- -> OH + H, d/d n[26]
    J[15 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[24 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[37]*n[26];  // H3O+ + e- -> H2O + H, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[26 + 16 * ld] -= d;
    J[14 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    d = k[37]*n[16];  // H3O+ + e- -> H2O + H, d/d n[26]
    J[16 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[14 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[38]*n[20];  // O + H -> OH, d/d n[9]
    J[20 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[24 + 9 * ld] += d;
    d = k[38]*n[9];  // O + H -> OH, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[9 + 20 * ld] -= d;
    J[24 + 20 * ld] += d;
    d = k[39]*n[20];  // O + H2 -> OH + H, d/d n[12]
    J[20 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[24 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[39]*n[12];  // O + H2 -> OH + H, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[12 + 20 * ld] -= d;
    J[24 + 20 * ld] += d;
    J[9 + 20 * ld] += d;
    d = k[40]*n[24];  // H2 + OH -> H2O + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[24 + 12 * ld] -= d;
    J[14 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[40]*n[12];  // H2 + OH -> H2O + H, d/d n[24]
    J[12 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[14 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[41]*n[24]*R(2.0);  // OH + OH -> H2O + O, d/d n[24]
    J[24 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[14 + 24 * ld] += d;
    J[20 + 24 * ld] += d;
    d = k[42]*n[24];  // H+ + OH -> OH+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[24 + 10 * ld] -= d;
    J[25 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[42]*n[10];  // H+ + OH -> OH+ + H, d/d n[24]
    J[10 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[25 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[43]*n[14];  // H+ + H2O -> H2O+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[14 + 10 * ld] -= d;
    J[15 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[43]*n[10];  // H+ + H2O -> H2O+ + H, d/d n[14]
    J[10 + 14 * ld] -= d;
    J[14 + 14 * ld] -= d;
    J[15 + 14 * ld] += d;
    J[9 + 14 * ld] += d;
    d = k[44]*n[24];  // H + OH -> H2 + O, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[24 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[20 + 9 * ld] += d;
    d = k[44]*n[9];  // H + OH -> H2 + O, d/d n[24]
    J[9 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[12 + 24 * ld] += d;
    J[20 + 24 * ld] += d;
    d = k[45]*n[14];  // H + H2O -> OH + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[14 + 9 * ld] -= d;
    J[24 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[45]*n[9];  // H + H2O -> OH + H2, d/d n[14]
    J[9 + 14 * ld] -= d;
    J[14 + 14 * ld] -= d;
    J[24 + 14 * ld] += d;
    J[12 + 14 * ld] += d;
    d = k[46]*n[20]*R(2.0);  // O + O -> O2, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[20 + 20 * ld] -= d;
    J[22 + 20 * ld] += d;
    d = k[47]*n[24];  // O + OH -> O2 + H, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[24 + 20 * ld] -= d;
    J[22 + 20 * ld] += d;
    J[9 + 20 * ld] += d;
    d = k[47]*n[20];  // O + OH -> O2 + H, d/d n[24]
    J[20 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[22 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[48]*n[22];  // H + O2 -> OH + O, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[22 + 9 * ld] -= d;
    J[24 + 9 * ld] += d;
    J[20 + 9 * ld] += d;
    d = k[48]*n[9];  // H + O2 -> OH + O, d/d n[22]
    J[9 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[24 + 22 * ld] += d;
    J[20 + 22 * ld] += d;
    d = k[49]*n[22];  // H+ + O2 -> O2+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[22 + 10 * ld] -= d;
    J[23 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[49]*n[10];  // H+ + O2 -> O2+ + H, d/d n[22]
    J[10 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[23 + 22 * ld] += d;
    J[9 + 22 * ld] += d;
    d = k[50]*n[26];  // O2+ + e- -> O + O, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[26 + 23 * ld] -= d;
    J[20 + 23 * ld] += d;
    J[20 + 23 * ld] += d;
    d = k[50]*n[23];  // O2+ + e- -> O + O, d/d n[26]
    J[23 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[20 + 26 * ld] += d;
    J[20 + 26 * ld] += d;
    d = k[51]*n[20];  // O + CH -> CO + H, d/d n[2]
    J[20 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[6 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[51]*n[2];  // O + CH -> CO + H, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[2 + 20 * ld] -= d;
    J[6 + 20 * ld] += d;
    J[9 + 20 * ld] += d;
    d = k[52]*n[24];  // C + OH -> CO + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[24 + 0 * ld] -= d;
    J[6 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[52]*n[0];  // C + OH -> CO + H, d/d n[24]
    J[0 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[6 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[53]*n[22];  // C + O2 -> CO + O, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[22 + 0 * ld] -= d;
    J[6 + 0 * ld] += d;
    J[20 + 0 * ld] += d;
    d = k[53]*n[0];  // C + O2 -> CO + O, d/d n[22]
    J[0 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[6 + 22 * ld] += d;
    J[20 + 22 * ld] += d;
    d = k[54]*n[22];  // C+ + O2 -> O+ + CO, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[22 + 1 * ld] -= d;
    J[21 + 1 * ld] += d;
    J[6 + 1 * ld] += d;
    d = k[54]*n[1];  // C+ + O2 -> O+ + CO, d/d n[22]
    J[1 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[21 + 22 * ld] += d;
    J[6 + 22 * ld] += d;
    d = k[55]*n[24];  // OH + CO -> CO2 + H, d/d n[6]
    J[24 + 6 * ld] -= d;
    J[6 + 6 * ld] -= d;
    J[8 + 6 * ld] += d;
    J[9 + 6 * ld] += d;
    d = k[55]*n[6];  // OH + CO -> CO2 + H, d/d n[24]
    J[24 + 24 * ld] -= d;
    J[6 + 24 * ld] -= d;
    J[8 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[56]*n[26];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[26 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[56]*n[1];  // C+ + e- -> C, d/d n[26]
    J[1 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[0 + 26 * ld] += d;
    d = k[57]*n[24];  // C+ + OH -> CO+ + H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[24 + 1 * ld] -= d;
    J[7 + 1 * ld] += d;
    J[9 + 1 * ld] += d;
    d = k[57]*n[1];  // C+ + OH -> CO+ + H, d/d n[24]
    J[1 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[7 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[58]*n[9];  // CO+ + H -> H+ + CO, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[9 + 7 * ld] -= d;
    J[10 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    d = k[58]*n[7];  // CO+ + H -> H+ + CO, d/d n[9]
    J[7 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[59]*n[9];  // C + H -> CH, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[9 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    d = k[59]*n[0];  // C + H -> CH, d/d n[9]
    J[0 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    d = k[60]*n[12];  // C + H2 -> CH + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[12 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[60]*n[0];  // C + H2 -> CH + H, d/d n[12]
    J[0 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[2 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[61]*n[9];  // H + CH -> C + H2, d/d n[2]
    J[9 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[12 + 2 * ld] += d;
    d = k[61]*n[2];  // H + CH -> C + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] -= d;
    J[0 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[62]*n[12];  // H2 + CH -> CH2 + H, d/d n[2]
    J[12 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[62]*n[2];  // H2 + CH -> CH2 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[2 + 12 * ld] -= d;
    J[3 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[63]*n[9];  // H + CH2 -> CH + H2, d/d n[3]
    J[9 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    J[12 + 3 * ld] += d;
    d = k[63]*n[3];  // H + CH2 -> CH + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[3 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[64]*n[12];  // H2 + CH2 -> CH3 + H, d/d n[3]
    J[12 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[9 + 3 * ld] += d;
    d = k[64]*n[3];  // H2 + CH2 -> CH3 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[3 + 12 * ld] -= d;
    J[4 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[65]*n[9];  // H + CH3 -> CH2 + H2, d/d n[4]
    J[9 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[12 + 4 * ld] += d;
    d = k[65]*n[4];  // H + CH3 -> CH2 + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[4 + 9 * ld] -= d;
    J[3 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[66]*n[12];  // H2 + CH3 -> CH4 + H, d/d n[4]
    J[12 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[9 + 4 * ld] += d;
    d = k[66]*n[4];  // H2 + CH3 -> CH4 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[4 + 12 * ld] -= d;
    J[5 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[67]*n[9];  // H + CH4 -> H2 + CH3, d/d n[5]
    J[9 + 5 * ld] -= d;
    J[5 + 5 * ld] -= d;
    J[12 + 5 * ld] += d;
    J[4 + 5 * ld] += d;
    d = k[67]*n[5];  // H + CH4 -> H2 + CH3, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[5 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[4 + 9 * ld] += d;
    d = k[68]*n[12];  // H2 + C -> CH2, d/d n[0]
    J[12 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    d = k[68]*n[0];  // H2 + C -> CH2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[0 + 12 * ld] -= d;
    J[3 + 12 * ld] += d;
    d = k[69]*n[26];  // H3O+ + e- -> OH + H + H, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[26 + 16 * ld] -= d;
    J[24 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    d = k[69]*n[16];  // H3O+ + e- -> OH + H + H, d/d n[26]
    J[16 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[24 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[70]*n[26];  // OH+ + e- -> O + H, d/d n[25]
    J[25 + 25 * ld] -= d;
    J[26 + 25 * ld] -= d;
    J[20 + 25 * ld] += d;
    J[9 + 25 * ld] += d;
    d = k[70]*n[25];  // OH+ + e- -> O + H, d/d n[26]
    J[25 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[20 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[71]*n[26];  // O+ + e- -> O, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[26 + 21 * ld] -= d;
    J[20 + 21 * ld] += d;
    d = k[71]*n[21];  // O+ + e- -> O, d/d n[26]
    J[21 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[20 + 26 * ld] += d;
    d = k[72]*n[26];  // CO+ + e- -> CO, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[26 + 7 * ld] -= d;
    J[6 + 7 * ld] += d;
    d = k[72]*n[7];  // CO+ + e- -> CO, d/d n[26]
    J[7 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[6 + 26 * ld] += d;
    d = k[73]*n[26];  // CO+ + e- -> C + O, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[26 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    J[20 + 7 * ld] += d;
    d = k[73]*n[7];  // CO+ + e- -> C + O, d/d n[26]
    J[7 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[0 + 26 * ld] += d;
    J[20 + 26 * ld] += d;
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
