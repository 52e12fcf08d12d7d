// generated by netgen.py from primordialZ_UV.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_primordialZ_UV {
  static constexpr const char* NAME = "primordialZ_UV";
  static constexpr int NSPEC = 29;
  static constexpr int NREA = 84;
  static constexpr int NPAR = 3;
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
  static constexpr int IDX_H3j = 16;
  static constexpr int IDX_H3Oj = 17;
  static constexpr int IDX_He = 18;
  static constexpr int IDX_Hej = 19;
  static constexpr int IDX_Hejj = 20;
  static constexpr int IDX_HeHj = 21;
  static constexpr int IDX_O = 22;
  static constexpr int IDX_Oj = 23;
  static constexpr int IDX_O2 = 24;
  static constexpr int IDX_O2j = 25;
  static constexpr int IDX_OH = 26;
  static constexpr int IDX_OHj = 27;
  static constexpr int IDX_E = 28;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 28;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = true;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 4;
  // PHOTO heating: the energy sums sum_b sigmaE_qb Phi_b follow the
  // rate sums in the row (sigmaE: sigma times the photoelectron's
  // energy E - Emin_q, bin-averaged)
  static constexpr bool PHOTO_HEAT = false;
  static constexpr int NROW = NPAR + NPHOTO;
  // parameter row (param_names order), then the photo sums
  struct Params {
    double redshift;
    double gamma_ad;
    double Tcmb;
    double ph[NPHOTO];
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.redshift = row[0];
    p.gamma_ad = row[1];
    p.Tcmb = row[2];
    for (int r = 0; r < NPHOTO; ++r) p.ph[r] = row[NPAR + r];
    return p;
  }
  static constexpr int photo_reaction[4] = {36, 37, 38, 39};
  static constexpr int photo_reactant[4] = {9, 18, 0, 22};
  static constexpr double photo_Emin[4] = {13.6, 24.6, 11.26, 13.62};
  static constexpr double photo_Emax[4] = {10000.0, 10000.0, 291.0, 538.0};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return kx::sigma_v96(E_eV, R(0.4298), R(54750.0), R(32.88), R(2.963), R(0.0), R(0.0), R(0.0));  // H -> H+ + e-
      case 1: return kx::sigma_v96(E_eV, R(13.61), R(949.2), R(1.469), R(3.188), R(2.039), R(0.4434), R(2.136));  // He -> He+ + e-
      case 2: return kx::sigma_v96(E_eV, R(2.144), R(502.7), R(62.16), R(5.101), R(0.09157), R(1.133), R(1.607));  // C -> C+ + e-
      case 3: return kx::sigma_v96(E_eV, R(1.24), R(1745.0), R(3.784), R(17.64), R(0.07589), R(8.698), R(0.1271));  // O -> O+ + e-
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
      case 17: return 1;
      case 19: return 1;
      case 20: return 2;
      case 21: return 1;
      case 23: return 1;
      case 25: return 1;
      case 27: return 1;
      case 28: return -1;
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
      case 16: return 5.020559199028353e-24;
      case 17: return 3.1587523725561754e-23;
      case 18: return 6.64647366797316e-24;
      case 19: return 6.645562729603113e-24;
      case 20: return 6.644651791233067e-24;
      case 21: return 8.319386108735913e-24;
      case 22: return 2.65669645265334e-23;
      case 23: return 2.6566053588163353e-23;
      case 24: return 5.31339290530668e-23;
      case 25: return 5.313301811469676e-23;
      case 26: return 2.82407879056662e-23;
      case 27: return 2.823987696729615e-23;
      case 28: return 9.10938370046373e-28;
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
      case 69: return 3;
      case 71: return 1;
      case 74: return 1;
      case 78: return 1;
      case 82: return 1;
      case 85: return 1;
      case 86: return 1;
      case 91: return 1;
      case 95: return 1;
      case 99: return 2;
      case 103: return 2;
      case 105: return 1;
      case 107: return 1;
      case 109: return 1;
      case 111: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 259 nonzeros, 290 in L + U with the fill, 1498 flops per
  // factorization (dense: 16259); rows in pivot order
  static constexpr int SP_NNZ = 290;
  static constexpr int SP_FLOPS = 1498;
  static constexpr int sp_perm[29] = {5, 4, 3, 8, 20, 2, 16, 17, 14, 21, 18, 11, 13, 19, 25, 15, 27, 10, 23, 0, 1, 6, 7, 9, 12, 22, 24, 26, 28};
  static constexpr int sp_rp[30] = {0, 4, 9, 15, 18, 21, 27, 33, 37, 45, 51, 59, 69, 78, 87, 91, 99, 105, 120, 130, 140, 149, 160, 168, 197, 226, 241, 252, 268, 290};
  static constexpr int sp_ci[290] = {0, 1, 23, 24, 0, 1, 2, 23, 24, 1, 2, 5, 19, 23, 24, 3, 21, 27, 4, 13, 28, 2, 5, 19, 23, 24, 25, 6, 9, 11, 12, 24, 28, 7, 15, 24, 28, 7, 8, 15, 17, 23, 24, 27, 28, 9, 10, 12, 23, 24, 28, 9, 10, 11, 12, 13, 23, 24, 28, 6, 9, 10, 11, 12, 13, 17, 23, 24, 28, 9, 10, 11, 12, 13, 17, 23, 24, 28, 4, 10, 11, 12, 13, 17, 23, 24, 28, 14, 17, 26, 28, 8, 15, 16, 17, 23, 24, 27, 28, 16, 17, 18, 24, 27, 28, 8, 11, 12, 13, 15, 16, 17, 18, 22, 23, 24, 25, 26, 27, 28, 17, 18, 20, 22, 23, 24, 25, 26, 27, 28, 5, 19, 20, 22, 23, 24, 25, 26, 27, 28, 19, 20, 22, 23, 24, 25, 26, 27, 28, 5, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 20, 22, 23, 24, 25, 26, 27, 28, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 5, 14, 15, 16, 17, 18, 19, 20, 22, 23, 24, 25, 26, 27, 28, 17, 18, 19, 20, 22, 23, 24, 25, 26, 27, 28, 7, 8, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 4, 6, 7, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 23, 24, 25, 26, 27, 28};
  static constexpr int sp_di[29] = {0, 5, 10, 15, 18, 22, 27, 33, 38, 45, 52, 62, 72, 82, 87, 92, 99, 111, 121, 131, 141, 152, 161, 191, 221, 237, 249, 266, 289};
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
      case 31: return 1;
      case 32: return 2;
      case 33: return 3;
      case 34: return 4;
      case 38: return 1;
      case 39: return 1;
      case 40: return 1;
      case 41: return 2;
      case 42: return 2;
      case 43: return 2;
      case 44: return 2;
      case 45: return 3;
      case 46: return 3;
      case 50: return 1;
      case 55: return 1;
      case 56: return 1;
      case 76: return 1;
      case 77: return 1;
      case 78: return 1;
      case 79: return 1;
      case 93: return 1;
      case 94: return 1;
      case 95: return 2;
      case 101: return 1;
      case 102: return 1;
      case 104: return 1;
      case 109: return 1;
      case 110: return 1;
      case 111: return 2;
      case 112: return 2;
      case 113: return 1;
      case 114: return 1;
      case 117: return 1;
      case 123: return 1;
      case 126: return 1;
      case 127: return -1;
      case 129: return 1;
      case 131: return 1;
      case 132: return 1;
      case 133: return 1;
      case 135: return 1;
      case 136: return 2;
      case 137: return 1;
      case 139: return 1;
      case 141: return 1;
      case 143: return 1;
      case 144: return -1;
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
    s += n[27];
    s += n[28];
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
    s += n[17] * R(3.0);
    s += n[18] * R(0.0);
    s += n[19] * R(0.0);
    s += n[20] * R(0.0);
    s += n[21] * R(1.0);
    s += n[22] * R(0.0);
    s += n[23] * R(0.0);
    s += n[24] * R(0.0);
    s += n[25] * R(0.0);
    s += n[26] * R(1.0);
    s += n[27] * R(1.0);
    s += n[28] * R(0.0);
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
    s += n[16] * R(5.020559199028353e-24);
    s += n[17] * R(3.1587523725561754e-23);
    s += n[18] * R(6.64647366797316e-24);
    s += n[19] * R(6.645562729603113e-24);
    s += n[20] * R(6.644651791233067e-24);
    s += n[21] * R(8.319386108735913e-24);
    s += n[22] * R(2.65669645265334e-23);
    s += n[23] * R(2.6566053588163353e-23);
    s += n[24] * R(5.31339290530668e-23);
    s += n[25] * R(5.313301811469676e-23);
    s += n[26] * R(2.82407879056662e-23);
    s += n[27] * R(2.823987696729615e-23);
    s += n[28] * R(9.10938370046373e-28);
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
    {  // H- + H -> H2 + e-: 2 piece(s), zero below 2.73, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(2.73) && Tgas < R(300.0)) k[7] = R(1.5e-09);  // H- + H -> H2 + e-
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
    {  // H- + H+ -> H2+ + e-: 2 piece(s), zero below 2.73, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(2.73) && Tgas < R(8000.0)) k[16] = (R(6.9e-09) * kx::pow(Tgas, (-R(0.35))));  // H- + H+ -> H2+ + e-
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
    k[36] = R(p.ph[0]);  // H -> H+ + e- (photo)
    k[37] = R(p.ph[1]);  // He -> He+ + e- (photo)
    k[38] = R(p.ph[2]);  // C -> C+ + e- (photo)
    k[39] = R(p.ph[3]);  // O -> O+ + e- (photo)
    k[40] = (R(6.8) * R(1e-10));  // O+ + H -> H+ + O
    {
      const R v_invT = (R(1.0) / Tgas);
      k[41] = ((R(7.0) * R(1e-10)) * kx::exp(((-R(232.0)) * v_invT)));  // H+ + O -> O+ + H
    }
    k[42] = (R(1.7) * R(1e-09));  // O+ + H2 -> OH+ + H
    k[43] = (R(1.01) * R(1e-09));  // OH+ + H2 -> H2O+ + H
    k[44] = (R(8.3) * R(1e-10));  // H2O+ + H2 -> H3O+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[45] = ((R(2.0) * R(1e-07)) * kx::pow(v_T32, (-R(0.5))));  // H2O+ + e- -> O + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[46] = ((R(1.6) * R(1e-07)) * kx::pow(v_T32, (-R(0.5))));  // H2O+ + e- -> OH + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[47] = ((R(3.5) * R(1e-07)) * kx::pow(v_T32, (-R(0.5))));  // H3O+ + e- -> H2O + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[48] = ((R(9.9) * R(1e-19)) * kx::pow(v_T32, (-R(0.38))));  // O + H -> OH
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[49] = (((R(3.43) * R(1e-13)) * kx::pow(v_T32, R(2.67))) * kx::exp(((-R(3160.0)) * v_invT)));  // O + H2 -> OH + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[50] = (((R(1.55) * R(1e-12)) * kx::pow(v_T32, R(1.6))) * kx::exp(((-R(1660.0)) * v_invT)));  // H2 + OH -> H2O + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[51] = (((R(1.65) * R(1e-12)) * kx::pow(v_T32, R(1.14))) * kx::exp(((-R(50.0)) * v_invT)));  // OH + OH -> H2O + O
    }
    k[52] = (R(2.1) * R(1e-09));  // H+ + OH -> OH+ + H
    k[53] = (R(6.9) * R(1e-09));  // H+ + H2O -> H2O+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[54] = (((R(7.0) * R(1e-14)) * kx::pow(v_T32, R(2.8))) * kx::exp(((-R(1950.0)) * v_invT)));  // H + OH -> H2 + O
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[55] = (((R(6.83) * R(1e-12)) * kx::pow(v_T32, R(1.6))) * kx::exp(((-R(9720.0)) * v_invT)));  // H + H2O -> OH + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[56] = ((R(4.9) * R(1e-20)) * kx::pow(v_T32, R(1.58)));  // O + O -> O2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[57] = (((R(4.34) * R(1e-11)) * kx::pow(v_T32, (-R(0.5)))) * kx::exp(((-R(30.0)) * v_invT)));  // O + OH -> O2 + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[58] = ((R(3.3) * R(1e-10)) * kx::exp(((-R(8460.0)) * v_invT)));  // H + O2 -> OH + O
    }
    k[59] = (R(2.0) * R(1e-09));  // H+ + O2 -> O2+ + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[60] = ((R(1.95) * R(1e-07)) * kx::pow(v_T32, (-R(0.7))));  // O2+ + e- -> O + O
    }
    k[61] = (R(6.6) * R(1e-11));  // O + CH -> CO + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[62] = ((R(1.1) * R(1e-10)) * kx::sqrt(v_T32));  // C + OH -> CO + H
    }
    k[63] = (R(3.3) * R(1e-11));  // C + O2 -> CO + O
    k[64] = (R(6.2) * R(1e-10));  // C+ + O2 -> O+ + CO
    k[65] = (R(1.0) * R(1e-13));  // OH + CO -> CO2 + H
    {
      const R v_T32 = (Tgas / R(300.0));
      k[66] = ((R(4.4) * R(1e-12)) * kx::pow(v_T32, (-R(0.61))));  // C+ + e- -> C
    }
    k[67] = (R(7.7) * R(1e-10));  // C+ + OH -> CO+ + H
    k[68] = (R(7.5) * R(1e-10));  // CO+ + H -> H+ + CO
    k[69] = (R(1.0) * R(1e-17));  // C + H -> CH
    {
      const R v_invT = (R(1.0) / Tgas);
      k[70] = ((R(6.64) * R(1e-10)) * kx::exp(((-R(11700.0)) * v_invT)));  // C + H2 -> CH + H
    }
    k[71] = (R(4.98) * R(1e-11));  // H + CH -> C + H2
    {
      const R v_invT = (R(1.0) / Tgas);
      k[72] = ((R(2.38) * R(1e-10)) * kx::exp(((-R(1760.0)) * v_invT)));  // H2 + CH -> CH2 + H
    }
    k[73] = (R(2.7) * R(1e-10));  // H + CH2 -> CH + H2
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[74] = (((R(5.18) * R(1e-11)) * kx::pow(v_T32, R(0.17))) * kx::exp(((-R(6400.0)) * v_invT)));  // H2 + CH2 -> CH3 + H
    }
    {
      const R v_invT = (R(1.0) / Tgas);
      k[75] = ((R(1.0) * R(1e-10)) * kx::exp(((-R(7600.0)) * v_invT)));  // H + CH3 -> CH2 + H2
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[76] = (((R(6.86) * R(1e-14)) * kx::pow(v_T32, R(2.74))) * kx::exp(((-R(4740.0)) * v_invT)));  // H2 + CH3 -> CH4 + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      k[77] = (((R(5.82) * R(1e-13)) * kx::ipow<3>(v_T32)) * kx::exp(((-R(4045.0)) * v_invT)));  // H + CH4 -> H2 + CH3
    }
    k[78] = (R(1.0) * R(1e-17));  // H2 + C -> CH2
    {
      const R v_T32 = (Tgas / R(300.0));
      k[79] = (R(6.5e-07) * kx::pow(v_T32, (-R(0.5))));  // H3O+ + e- -> OH + H + H
    }
    {
      const R v_T32 = (Tgas / R(300.0));
      k[80] = (R(6.3e-09) * kx::pow(v_T32, (-R(0.48))));  // OH+ + e- -> O + H
    }
    k[81] = ((R(6.622e-11) / ((kx::sqrt((Tgas / R(4.136))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4.136)))), (R(1.0) - (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4214000.0)))), (R(1.0) + (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((R(5.629e-08) * kx::exp(((-R(5395.0)) / Tgas))) + (R(2.55e-07) * kx::exp(((-R(17700.0)) / Tgas)))) + (R(0.0006173) * kx::exp(((-R(167100.0)) / Tgas)))) + (R(0.0001627) * kx::exp(((-R(268700.0)) / Tgas))))));  // O+ + e- -> O
    k[82] = R(0.0);  // CO+ + e- -> CO
    k[83] = (R(2.75e-07) * kx::pow((Tgas / R(300.0)), (-R(0.55))));  // CO+ + e- -> C + O
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
  static constexpr int NREA_T = 82;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[85] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 43, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 75, 76, 77, 78, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150, 152, 154, 156, 158, 160, 162, 164, 166};
  static constexpr int LT_KF_SP[166] = {9, 28, 10, 28, 18, 28, 19, 28, 19, 28, 20, 28, 9, 28, 11, 9, 11, 9, 9, 10, 13, 9, 12, 10, 12, 28, 12, 9, 11, 28, 11, 10, 11, 10, 11, 10, 13, 28, 13, 11, 9, 9, 9, 12, 9, 9, 12, 12, 9, 9, 19, 12, 13, 18, 13, 12, 16, 11, 19, 9, 19, 11, 19, 12, 21, 9, 21, 12, 16, 28, 16, 28, 21, 28, 9, 18, 0, 22, 23, 9, 10, 22, 23, 12, 27, 12, 15, 12, 15, 28, 15, 28, 17, 28, 22, 9, 22, 12, 12, 26, 26, 26, 10, 26, 10, 14, 9, 26, 9, 14, 22, 22, 22, 26, 9, 24, 10, 24, 25, 28, 22, 2, 0, 26, 0, 24, 1, 24, 26, 6, 1, 28, 1, 26, 7, 9, 0, 9, 0, 12, 9, 2, 12, 2, 9, 3, 12, 3, 9, 4, 12, 4, 9, 5, 12, 0, 17, 28, 27, 28, 23, 28, 7, 28, 7, 28};
  static constexpr int LT_SC_OFF[85] = {0, 5, 8, 13, 16, 21, 24, 27, 31, 35, 38, 42, 46, 51, 56, 61, 65, 69, 73, 77, 81, 86, 91, 96, 101, 106, 110, 114, 118, 122, 126, 130, 134, 138, 142, 147, 151, 154, 157, 160, 163, 167, 171, 175, 179, 183, 187, 191, 195, 198, 202, 206, 210, 214, 218, 222, 226, 229, 233, 237, 241, 245, 249, 253, 257, 261, 265, 268, 272, 276, 279, 283, 287, 291, 295, 299, 303, 307, 311, 314, 319, 323, 326, 329, 333};
  static constexpr int LT_SC_ROW[333] = {9, 28, 10, 28, 28, 10, 28, 9, 18, 28, 19, 28, 28, 19, 28, 18, 19, 28, 20, 28, 28, 20, 28, 19, 9, 28, 11, 11, 9, 12, 28, 11, 9, 12, 28, 9, 10, 13, 13, 9, 12, 10, 12, 10, 13, 9, 12, 28, 9, 9, 28, 12, 9, 9, 9, 9, 11, 28, 9, 28, 28, 11, 10, 9, 9, 11, 10, 13, 28, 11, 10, 13, 28, 13, 28, 9, 9, 13, 11, 9, 12, 9, 9, 9, 12, 9, 12, 9, 9, 12, 12, 12, 12, 9, 9, 12, 9, 9, 10, 28, 9, 19, 12, 10, 18, 9, 13, 18, 21, 9, 13, 12, 16, 9, 16, 11, 12, 12, 19, 9, 10, 18, 19, 11, 9, 18, 19, 12, 13, 18, 21, 9, 13, 18, 21, 12, 16, 18, 16, 28, 12, 9, 16, 28, 9, 9, 9, 21, 28, 9, 18, 9, 10, 28, 18, 19, 28, 0, 1, 28, 22, 23, 28, 23, 9, 10, 22, 10, 22, 23, 9, 23, 12, 27, 9, 27, 12, 15, 9, 15, 12, 17, 9, 15, 28, 22, 12, 15, 28, 26, 9, 17, 28, 14, 9, 22, 9, 26, 22, 12, 26, 9, 12, 26, 14, 9, 26, 26, 14, 22, 10, 26, 27, 9, 10, 14, 15, 9, 9, 26, 12, 22, 9, 14, 26, 12, 22, 22, 24, 22, 26, 24, 9, 9, 24, 26, 22, 10, 24, 25, 9, 25, 28, 22, 22, 22, 2, 6, 9, 0, 26, 6, 9, 0, 24, 6, 22, 1, 24, 23, 6, 26, 6, 8, 9, 1, 28, 0, 1, 26, 7, 9, 7, 9, 10, 6, 0, 9, 2, 0, 12, 2, 9, 9, 2, 0, 12, 12, 2, 3, 9, 9, 3, 2, 12, 12, 3, 4, 9, 9, 4, 3, 12, 12, 4, 5, 9, 9, 5, 12, 4, 12, 0, 3, 17, 28, 26, 9, 9, 27, 28, 22, 9, 23, 28, 22, 7, 28, 6, 7, 28, 0, 22};
  static constexpr int LT_SC_SGN[333] = {-1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1};
  static constexpr int LT_JT_OFF[85] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 41, 43, 44, 45, 47, 49, 51, 53, 55, 57, 59, 61, 63, 65, 67, 69, 70, 71, 72, 73, 75, 77, 79, 81, 83, 85, 87, 89, 91, 93, 95, 96, 98, 100, 102, 104, 105, 107, 109, 111, 113, 115, 117, 119, 121, 123, 125, 127, 129, 131, 133, 135, 137, 139, 141, 143, 145, 147, 149, 151, 153, 155, 157, 159};
  static constexpr int LT_JT_COL[159] = {9, 28, 10, 28, 18, 28, 19, 28, 19, 28, 20, 28, 9, 28, 9, 11, 9, 11, 9, 10, 9, 13, 10, 12, 12, 28, 9, 12, 11, 28, 10, 11, 10, 11, 10, 11, 13, 28, 11, 13, 9, 9, 12, 12, 9, 12, 19, 13, 18, 12, 13, 11, 16, 9, 19, 11, 19, 12, 19, 9, 21, 12, 21, 16, 28, 16, 28, 21, 28, 9, 18, 0, 22, 9, 23, 10, 22, 12, 23, 12, 27, 12, 15, 15, 28, 15, 28, 17, 28, 9, 22, 12, 22, 12, 26, 26, 10, 26, 10, 14, 9, 26, 9, 14, 22, 22, 26, 9, 24, 10, 24, 25, 28, 2, 22, 0, 26, 0, 24, 1, 24, 6, 26, 1, 28, 1, 26, 7, 9, 0, 9, 0, 12, 2, 9, 2, 12, 3, 9, 3, 12, 4, 9, 4, 12, 5, 9, 0, 12, 17, 28, 27, 28, 23, 28, 7, 28, 7, 28};
  static constexpr int LT_JT_MULT[159] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 2, 1, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[160] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 42, 44, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72, 72, 72, 72, 72, 73, 74, 75, 76, 77, 78, 79, 80, 81, 82, 83, 84, 85, 86, 87, 88, 89, 90, 91, 92, 93, 94, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136, 137, 138, 139, 140, 141, 142, 143, 144, 145, 146, 147, 148, 149, 150, 151, 152, 153, 154, 155, 156, 157, 158};
  static constexpr int LT_JF_SP[158] = {28, 9, 28, 10, 28, 18, 28, 19, 28, 19, 28, 20, 28, 9, 11, 9, 11, 9, 10, 9, 13, 9, 12, 10, 28, 12, 12, 9, 28, 11, 11, 10, 11, 10, 11, 10, 28, 13, 13, 11, 9, 9, 12, 9, 9, 9, 12, 9, 19, 12, 18, 13, 13, 12, 16, 11, 19, 9, 19, 11, 19, 12, 21, 9, 21, 12, 28, 16, 28, 16, 28, 21, 23, 9, 22, 10, 23, 12, 27, 12, 15, 12, 28, 15, 28, 15, 28, 17, 22, 9, 22, 12, 26, 12, 26, 26, 10, 14, 10, 26, 9, 14, 9, 22, 26, 22, 24, 9, 24, 10, 28, 25, 22, 2, 26, 0, 24, 0, 24, 1, 26, 6, 28, 1, 26, 1, 9, 7, 9, 0, 12, 0, 9, 2, 12, 2, 9, 3, 12, 3, 9, 4, 12, 4, 9, 5, 12, 0, 28, 17, 28, 27, 28, 23, 28, 7, 28, 7};
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
    fl = k[0]*n[9]*n[28];  // H + e- -> H+ + e- + e-
    dn[9] -= fl;
    dn[28] -= fl;
    dn[10] += fl;
    dn[28] += fl;
    dn[28] += fl;
    fl = k[1]*n[10]*n[28];  // H+ + e- -> H
    dn[10] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    fl = k[2]*n[18]*n[28];  // He + e- -> He+ + e- + e-
    dn[18] -= fl;
    dn[28] -= fl;
    dn[19] += fl;
    dn[28] += fl;
    dn[28] += fl;
    fl = k[3]*n[19]*n[28];  // He+ + e- -> He
    dn[19] -= fl;
    dn[28] -= fl;
    dn[18] += fl;
    fl = k[4]*n[19]*n[28];  // He+ + e- -> He++ + e- + e-
    dn[19] -= fl;
    dn[28] -= fl;
    dn[20] += fl;
    dn[28] += fl;
    dn[28] += fl;
    fl = k[5]*n[20]*n[28];  // He++ + e- -> He+
    dn[20] -= fl;
    dn[28] -= fl;
    dn[19] += fl;
    fl = k[6]*n[9]*n[28];  // H + e- -> H-
    dn[9] -= fl;
    dn[28] -= fl;
    dn[11] += fl;
    fl = k[7]*n[11]*n[9];  // H- + H -> H2 + e-
    dn[11] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[28] += fl;
    fl = k[8]*n[11]*n[9];  // H- + H -> H2 + e-
    dn[11] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[28] += fl;
    fl = k[9]*n[9]*n[10];  // H + H+ -> H2+
    dn[9] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    fl = k[10]*n[13]*n[9];  // H2+ + H -> H2 + H+
    dn[13] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[10] += fl;
    fl = k[11]*n[12]*n[10];  // H2 + H+ -> H2+ + H
    dn[12] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    dn[9] += fl;
    fl = k[12]*n[12]*n[28];  // H2 + e- -> H + H + e-
    dn[12] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[28] += fl;
    fl = k[13]*n[12]*n[9];  // H2 + H -> H + H + H
    dn[12] -= fl;
    dn[9] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[14]*n[11]*n[28];  // H- + e- -> H + e- + e-
    dn[11] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    dn[28] += fl;
    dn[28] += fl;
    fl = k[15]*n[11]*n[10];  // H- + H+ -> H + H
    dn[11] -= fl;
    dn[10] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[16]*n[11]*n[10];  // H- + H+ -> H2+ + e-
    dn[11] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    dn[28] += fl;
    fl = k[17]*n[11]*n[10];  // H- + H+ -> H2+ + e-
    dn[11] -= fl;
    dn[10] -= fl;
    dn[13] += fl;
    dn[28] += fl;
    fl = k[18]*n[13]*n[28];  // H2+ + e- -> H + H
    dn[13] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[19]*n[13]*n[11];  // H2+ + H- -> H + H2
    dn[13] -= fl;
    dn[11] -= fl;
    dn[9] += fl;
    dn[12] += fl;
    fl = k[20]*n[9]*n[9]*n[9];  // H + H + H -> H2 + H
    dn[9] -= fl;
    dn[9] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[9] += fl;
    fl = k[21]*n[12]*n[9]*n[9];  // H2 + H + H -> H2 + H2
    dn[12] -= fl;
    dn[9] -= fl;
    dn[9] -= fl;
    dn[12] += fl;
    dn[12] += fl;
    fl = k[22]*n[12]*n[12];  // H2 + H2 -> H + H + H2
    dn[12] -= fl;
    dn[12] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[12] += fl;
    fl = k[23]*n[9]*n[9];  // H + H -> H+ + e- + H
    dn[9] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[28] += fl;
    dn[9] += fl;
    fl = k[24]*n[19]*n[12];  // He+ + H2 -> H+ + He + H
    dn[19] -= fl;
    dn[12] -= fl;
    dn[10] += fl;
    dn[18] += fl;
    dn[9] += fl;
    fl = k[25]*n[13]*n[18];  // H2+ + He -> HeH+ + H
    dn[13] -= fl;
    dn[18] -= fl;
    dn[21] += fl;
    dn[9] += fl;
    fl = k[26]*n[13]*n[12];  // H2+ + H2 -> H3+ + H
    dn[13] -= fl;
    dn[12] -= fl;
    dn[16] += fl;
    dn[9] += fl;
    fl = k[27]*n[16]*n[11];  // H3+ + H- -> H2 + H2
    dn[16] -= fl;
    dn[11] -= fl;
    dn[12] += fl;
    dn[12] += fl;
    fl = k[28]*n[19]*n[9];  // He+ + H -> H+ + He
    dn[19] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[18] += fl;
    fl = k[29]*n[19]*n[11];  // He+ + H- -> H + He
    dn[19] -= fl;
    dn[11] -= fl;
    dn[9] += fl;
    dn[18] += fl;
    fl = k[30]*n[19]*n[12];  // He+ + H2 -> H2+ + He
    dn[19] -= fl;
    dn[12] -= fl;
    dn[13] += fl;
    dn[18] += fl;
    fl = k[31]*n[21]*n[9];  // HeH+ + H -> H2+ + He
    dn[21] -= fl;
    dn[9] -= fl;
    dn[13] += fl;
    dn[18] += fl;
    fl = k[32]*n[21]*n[12];  // HeH+ + H2 -> H3+ + He
    dn[21] -= fl;
    dn[12] -= fl;
    dn[16] += fl;
    dn[18] += fl;
    fl = k[33]*n[16]*n[28];  // H3+ + e- -> H2 + H
    dn[16] -= fl;
    dn[28] -= fl;
    dn[12] += fl;
    dn[9] += fl;
    fl = k[34]*n[16]*n[28];  // H3+ + e- -> H + H + H
    dn[16] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[35]*n[21]*n[28];  // HeH+ + e- -> H + He
    dn[21] -= fl;
    dn[28] -= fl;
    dn[9] += fl;
    dn[18] += fl;
    fl = k[36]*n[9];  // H -> H+ + e-
    dn[9] -= fl;
    dn[10] += fl;
    dn[28] += fl;
    fl = k[37]*n[18];  // He -> He+ + e-
    dn[18] -= fl;
    dn[19] += fl;
    dn[28] += fl;
    fl = k[38]*n[0];  // C -> C+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[28] += fl;
    fl = k[39]*n[22];  // O -> O+ + e-
    dn[22] -= fl;
    dn[23] += fl;
    dn[28] += fl;
    fl = k[40]*n[23]*n[9];  // O+ + H -> H+ + O
    dn[23] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[22] += fl;
    fl = k[41]*n[10]*n[22];  // H+ + O -> O+ + H
    dn[10] -= fl;
    dn[22] -= fl;
    dn[23] += fl;
    dn[9] += fl;
    fl = k[42]*n[23]*n[12];  // O+ + H2 -> OH+ + H
    dn[23] -= fl;
    dn[12] -= fl;
    dn[27] += fl;
    dn[9] += fl;
    fl = k[43]*n[27]*n[12];  // OH+ + H2 -> H2O+ + H
    dn[27] -= fl;
    dn[12] -= fl;
    dn[15] += fl;
    dn[9] += fl;
    fl = k[44]*n[15]*n[12];  // H2O+ + H2 -> H3O+ + H
    dn[15] -= fl;
    dn[12] -= fl;
    dn[17] += fl;
    dn[9] += fl;
    fl = k[45]*n[15]*n[28];  // H2O+ + e- -> O + H2
    dn[15] -= fl;
    dn[28] -= fl;
    dn[22] += fl;
    dn[12] += fl;
    fl = k[46]*n[15]*n[28];  // H2O+ + e- -> OH + H
    dn[15] -= fl;
    dn[28] -= fl;
    dn[26] += fl;
    dn[9] += fl;
    fl = k[47]*n[17]*n[28];  // H3O+ + e- -> H2O + H
    dn[17] -= fl;
    dn[28] -= fl;
    dn[14] += fl;
    dn[9] += fl;
    fl = k[48]*n[22]*n[9];  // O + H -> OH
    dn[22] -= fl;
    dn[9] -= fl;
    dn[26] += fl;
    fl = k[49]*n[22]*n[12];  // O + H2 -> OH + H
    dn[22] -= fl;
    dn[12] -= fl;
    dn[26] += fl;
    dn[9] += fl;
    fl = k[50]*n[12]*n[26];  // H2 + OH -> H2O + H
    dn[12] -= fl;
    dn[26] -= fl;
    dn[14] += fl;
    dn[9] += fl;
    fl = k[51]*n[26]*n[26];  // OH + OH -> H2O + O
    dn[26] -= fl;
    dn[26] -= fl;
    dn[14] += fl;
    dn[22] += fl;
    fl = k[52]*n[10]*n[26];  // H+ + OH -> OH+ + H
    dn[10] -= fl;
    dn[26] -= fl;
    dn[27] += fl;
    dn[9] += fl;
    fl = k[53]*n[10]*n[14];  // H+ + H2O -> H2O+ + H
    dn[10] -= fl;
    dn[14] -= fl;
    dn[15] += fl;
    dn[9] += fl;
    fl = k[54]*n[9]*n[26];  // H + OH -> H2 + O
    dn[9] -= fl;
    dn[26] -= fl;
    dn[12] += fl;
    dn[22] += fl;
    fl = k[55]*n[9]*n[14];  // H + H2O -> OH + H2
    dn[9] -= fl;
    dn[14] -= fl;
    dn[26] += fl;
    dn[12] += fl;
    fl = k[56]*n[22]*n[22];  // O + O -> O2
    dn[22] -= fl;
    dn[22] -= fl;
    dn[24] += fl;
    fl = k[57]*n[22]*n[26];  // O + OH -> O2 + H
    dn[22] -= fl;
    dn[26] -= fl;
    dn[24] += fl;
    dn[9] += fl;
    fl = k[58]*n[9]*n[24];  // H + O2 -> OH + O
    dn[9] -= fl;
    dn[24] -= fl;
    dn[26] += fl;
    dn[22] += fl;
    fl = k[59]*n[10]*n[24];  // H+ + O2 -> O2+ + H
    dn[10] -= fl;
    dn[24] -= fl;
    dn[25] += fl;
    dn[9] += fl;
    fl = k[60]*n[25]*n[28];  // O2+ + e- -> O + O
    dn[25] -= fl;
    dn[28] -= fl;
    dn[22] += fl;
    dn[22] += fl;
    fl = k[61]*n[22]*n[2];  // O + CH -> CO + H
    dn[22] -= fl;
    dn[2] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    fl = k[62]*n[0]*n[26];  // C + OH -> CO + H
    dn[0] -= fl;
    dn[26] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    fl = k[63]*n[0]*n[24];  // C + O2 -> CO + O
    dn[0] -= fl;
    dn[24] -= fl;
    dn[6] += fl;
    dn[22] += fl;
    fl = k[64]*n[1]*n[24];  // C+ + O2 -> O+ + CO
    dn[1] -= fl;
    dn[24] -= fl;
    dn[23] += fl;
    dn[6] += fl;
    fl = k[65]*n[26]*n[6];  // OH + CO -> CO2 + H
    dn[26] -= fl;
    dn[6] -= fl;
    dn[8] += fl;
    dn[9] += fl;
    fl = k[66]*n[1]*n[28];  // C+ + e- -> C
    dn[1] -= fl;
    dn[28] -= fl;
    dn[0] += fl;
    fl = k[67]*n[1]*n[26];  // C+ + OH -> CO+ + H
    dn[1] -= fl;
    dn[26] -= fl;
    dn[7] += fl;
    dn[9] += fl;
    fl = k[68]*n[7]*n[9];  // CO+ + H -> H+ + CO
    dn[7] -= fl;
    dn[9] -= fl;
    dn[10] += fl;
    dn[6] += fl;
    fl = k[69]*n[0]*n[9];  // C + H -> CH
    dn[0] -= fl;
    dn[9] -= fl;
    dn[2] += fl;
    fl = k[70]*n[0]*n[12];  // C + H2 -> CH + H
    dn[0] -= fl;
    dn[12] -= fl;
    dn[2] += fl;
    dn[9] += fl;
    fl = k[71]*n[9]*n[2];  // H + CH -> C + H2
    dn[9] -= fl;
    dn[2] -= fl;
    dn[0] += fl;
    dn[12] += fl;
    fl = k[72]*n[12]*n[2];  // H2 + CH -> CH2 + H
    dn[12] -= fl;
    dn[2] -= fl;
    dn[3] += fl;
    dn[9] += fl;
    fl = k[73]*n[9]*n[3];  // H + CH2 -> CH + H2
    dn[9] -= fl;
    dn[3] -= fl;
    dn[2] += fl;
    dn[12] += fl;
    fl = k[74]*n[12]*n[3];  // H2 + CH2 -> CH3 + H
    dn[12] -= fl;
    dn[3] -= fl;
    dn[4] += fl;
    dn[9] += fl;
    fl = k[75]*n[9]*n[4];  // H + CH3 -> CH2 + H2
    dn[9] -= fl;
    dn[4] -= fl;
    dn[3] += fl;
    dn[12] += fl;
    fl = k[76]*n[12]*n[4];  // H2 + CH3 -> CH4 + H
    dn[12] -= fl;
    dn[4] -= fl;
    dn[5] += fl;
    dn[9] += fl;
    fl = k[77]*n[9]*n[5];  // H + CH4 -> H2 + CH3
    dn[9] -= fl;
    dn[5] -= fl;
    dn[12] += fl;
    dn[4] += fl;
    fl = k[78]*n[12]*n[0];  // H2 + C -> CH2
    dn[12] -= fl;
    dn[0] -= fl;
    dn[3] += fl;
    fl = k[79]*n[17]*n[28];  // H3O+ + e- -> OH + H + H
    dn[17] -= fl;
    dn[28] -= fl;
    dn[26] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[80]*n[27]*n[28];  // OH+ + e- -> O + H
    dn[27] -= fl;
    dn[28] -= fl;
    dn[22] += fl;
    dn[9] += fl;
    fl = k[81]*n[23]*n[28];  // O+ + e- -> O
    dn[23] -= fl;
    dn[28] -= fl;
    dn[22] += fl;
    fl = k[82]*n[7]*n[28];  // CO+ + e- -> CO
    dn[7] -= fl;
    dn[28] -= fl;
    dn[6] += fl;
    fl = k[83]*n[7]*n[28];  // CO+ + e- -> C + O
    dn[7] -= fl;
    dn[28] -= fl;
    dn[0] += fl;
    dn[22] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[28];  // H + e- -> H+ + e- + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[28 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    d = k[0]*n[9];  // H + e- -> H+ + e- + e-, d/d n[28]
    J[9 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[10 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    d = k[1]*n[28];  // H+ + e- -> H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[28 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[1]*n[10];  // H+ + e- -> H, d/d n[28]
    J[10 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    d = k[2]*n[28];  // He + e- -> He+ + e- + e-, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[28 + 18 * ld] -= d;
    J[19 + 18 * ld] += d;
    J[28 + 18 * ld] += d;
    J[28 + 18 * ld] += d;
    d = k[2]*n[18];  // He + e- -> He+ + e- + e-, d/d n[28]
    J[18 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[19 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    d = k[3]*n[28];  // He+ + e- -> He, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[28 + 19 * ld] -= d;
    J[18 + 19 * ld] += d;
    d = k[3]*n[19];  // He+ + e- -> He, d/d n[28]
    J[19 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[18 + 28 * ld] += d;
    d = k[4]*n[28];  // He+ + e- -> He++ + e- + e-, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[28 + 19 * ld] -= d;
    J[20 + 19 * ld] += d;
    J[28 + 19 * ld] += d;
    J[28 + 19 * ld] += d;
    d = k[4]*n[19];  // He+ + e- -> He++ + e- + e-, d/d n[28]
    J[19 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[20 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    d = k[5]*n[28];  // He++ + e- -> He+, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[28 + 20 * ld] -= d;
    J[19 + 20 * ld] += d;
    d = k[5]*n[20];  // He++ + e- -> He+, d/d n[28]
    J[20 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[19 + 28 * ld] += d;
    d = k[6]*n[28];  // H + e- -> H-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[28 + 9 * ld] -= d;
    J[11 + 9 * ld] += d;
    d = k[6]*n[9];  // H + e- -> H-, d/d n[28]
    J[9 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[11 + 28 * ld] += d;
    d = k[7]*n[11];  // H- + H -> H2 + e-, d/d n[9]
    J[11 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    d = k[7]*n[9];  // H- + H -> H2 + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] -= d;
    J[12 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    d = k[8]*n[11];  // H- + H -> H2 + e-, d/d n[9]
    J[11 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    d = k[8]*n[9];  // H- + H -> H2 + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] -= d;
    J[12 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    d = k[9]*n[10];  // H + H+ -> H2+, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] -= d;
    J[13 + 9 * ld] += d;
    d = k[9]*n[9];  // H + H+ -> H2+, d/d n[10]
    J[9 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    d = k[10]*n[13];  // H2+ + H -> H2 + H+, d/d n[9]
    J[13 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[10 + 9 * ld] += d;
    d = k[10]*n[9];  // H2+ + H -> H2 + H+, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[9 + 13 * ld] -= d;
    J[12 + 13 * ld] += d;
    J[10 + 13 * ld] += d;
    d = k[11]*n[12];  // H2 + H+ -> H2+ + H, d/d n[10]
    J[12 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[11]*n[10];  // H2 + H+ -> H2+ + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[10 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[12]*n[28];  // H2 + e- -> H + H + e-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[28 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[28 + 12 * ld] += d;
    d = k[12]*n[12];  // H2 + e- -> H + H + e-, d/d n[28]
    J[12 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    d = k[13]*n[12];  // H2 + H -> H + H + H, d/d n[9]
    J[12 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[13]*n[9];  // H2 + H -> H + H + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[14]*n[28];  // H- + e- -> H + e- + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[28 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    d = k[14]*n[11];  // H- + e- -> H + e- + e-, d/d n[28]
    J[11 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    J[28 + 28 * ld] += d;
    d = k[15]*n[11];  // H- + H+ -> H + H, d/d n[10]
    J[11 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[15]*n[10];  // H- + H+ -> H + H, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[9 + 11 * ld] += d;
    d = k[16]*n[11];  // H- + H+ -> H2+ + e-, d/d n[10]
    J[11 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    J[28 + 10 * ld] += d;
    d = k[16]*n[10];  // H- + H+ -> H2+ + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] -= d;
    J[13 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    d = k[17]*n[11];  // H- + H+ -> H2+ + e-, d/d n[10]
    J[11 + 10 * ld] -= d;
    J[10 + 10 * ld] -= d;
    J[13 + 10 * ld] += d;
    J[28 + 10 * ld] += d;
    d = k[17]*n[10];  // H- + H+ -> H2+ + e-, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[10 + 11 * ld] -= d;
    J[13 + 11 * ld] += d;
    J[28 + 11 * ld] += d;
    d = k[18]*n[28];  // H2+ + e- -> H + H, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[28 + 13 * ld] -= d;
    J[9 + 13 * ld] += d;
    J[9 + 13 * ld] += d;
    d = k[18]*n[13];  // H2+ + e- -> H + H, d/d n[28]
    J[13 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[19]*n[13];  // H2+ + H- -> H + H2, d/d n[11]
    J[13 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[12 + 11 * ld] += d;
    d = k[19]*n[11];  // H2+ + H- -> H + H2, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[11 + 13 * ld] -= d;
    J[9 + 13 * ld] += d;
    J[12 + 13 * ld] += d;
    d = k[20]*n[9]*n[9]*R(3.0);  // H + H + H -> H2 + H, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[21]*n[12]*n[9]*R(2.0);  // H2 + H + H -> H2 + H2, d/d n[9]
    J[12 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[21]*n[9]*n[9];  // H2 + H + H -> H2 + H2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[9 + 12 * ld] -= d;
    J[12 + 12 * ld] += d;
    J[12 + 12 * ld] += d;
    d = k[22]*n[12]*R(2.0);  // H2 + H2 -> H + H + H2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[9 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    J[12 + 12 * ld] += d;
    d = k[23]*n[9]*R(2.0);  // H + H -> H+ + e- + H, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[24]*n[19];  // He+ + H2 -> H+ + He + H, d/d n[12]
    J[19 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[10 + 12 * ld] += d;
    J[18 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[24]*n[12];  // He+ + H2 -> H+ + He + H, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[12 + 19 * ld] -= d;
    J[10 + 19 * ld] += d;
    J[18 + 19 * ld] += d;
    J[9 + 19 * ld] += d;
    d = k[25]*n[18];  // H2+ + He -> HeH+ + H, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[18 + 13 * ld] -= d;
    J[21 + 13 * ld] += d;
    J[9 + 13 * ld] += d;
    d = k[25]*n[13];  // H2+ + He -> HeH+ + H, d/d n[18]
    J[13 + 18 * ld] -= d;
    J[18 + 18 * ld] -= d;
    J[21 + 18 * ld] += d;
    J[9 + 18 * ld] += d;
    d = k[26]*n[13];  // H2+ + H2 -> H3+ + H, d/d n[12]
    J[13 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[16 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[26]*n[12];  // H2+ + H2 -> H3+ + H, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[12 + 13 * ld] -= d;
    J[16 + 13 * ld] += d;
    J[9 + 13 * ld] += d;
    d = k[27]*n[16];  // H3+ + H- -> H2 + H2, d/d n[11]
    J[16 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[12 + 11 * ld] += d;
    J[12 + 11 * ld] += d;
    d = k[27]*n[11];  // H3+ + H- -> H2 + H2, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[11 + 16 * ld] -= d;
    J[12 + 16 * ld] += d;
    J[12 + 16 * ld] += d;
    d = k[28]*n[19];  // He+ + H -> H+ + He, d/d n[9]
    J[19 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[18 + 9 * ld] += d;
    d = k[28]*n[9];  // He+ + H -> H+ + He, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[9 + 19 * ld] -= d;
    J[10 + 19 * ld] += d;
    J[18 + 19 * ld] += d;
    d = k[29]*n[19];  // He+ + H- -> H + He, d/d n[11]
    J[19 + 11 * ld] -= d;
    J[11 + 11 * ld] -= d;
    J[9 + 11 * ld] += d;
    J[18 + 11 * ld] += d;
    d = k[29]*n[11];  // He+ + H- -> H + He, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[11 + 19 * ld] -= d;
    J[9 + 19 * ld] += d;
    J[18 + 19 * ld] += d;
    d = k[30]*n[19];  // He+ + H2 -> H2+ + He, d/d n[12]
    J[19 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[18 + 12 * ld] += d;
    d = k[30]*n[12];  // He+ + H2 -> H2+ + He, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[12 + 19 * ld] -= d;
    J[13 + 19 * ld] += d;
    J[18 + 19 * ld] += d;
    d = k[31]*n[21];  // HeH+ + H -> H2+ + He, d/d n[9]
    J[21 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[13 + 9 * ld] += d;
    J[18 + 9 * ld] += d;
    d = k[31]*n[9];  // HeH+ + H -> H2+ + He, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[9 + 21 * ld] -= d;
    J[13 + 21 * ld] += d;
    J[18 + 21 * ld] += d;
    d = k[32]*n[21];  // HeH+ + H2 -> H3+ + He, d/d n[12]
    J[21 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[16 + 12 * ld] += d;
    J[18 + 12 * ld] += d;
    d = k[32]*n[12];  // HeH+ + H2 -> H3+ + He, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[12 + 21 * ld] -= d;
    J[16 + 21 * ld] += d;
    J[18 + 21 * ld] += d;
    d = k[33]*n[28];  // H3+ + e- -> H2 + H, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[28 + 16 * ld] -= d;
    J[12 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    d = k[33]*n[16];  // H3+ + e- -> H2 + H, d/d n[28]
    J[16 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[12 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[34]*n[28];  // H3+ + e- -> H + H + H, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[28 + 16 * ld] -= d;
    J[9 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    J[9 + 16 * ld] += d;
    d = k[34]*n[16];  // H3+ + e- -> H + H + H, d/d n[28]
    J[16 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[35]*n[28];  // HeH+ + e- -> H + He, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[28 + 21 * ld] -= d;
    J[9 + 21 * ld] += d;
    J[18 + 21 * ld] += d;
    d = k[35]*n[21];  // HeH+ + e- -> H + He, d/d n[28]
    J[21 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[9 + 28 * ld] += d;
    J[18 + 28 * ld] += d;
    d = k[36];  // H -> H+ + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[28 + 9 * ld] += d;
    d = k[37];  // He -> He+ + e-, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[19 + 18 * ld] += d;
    J[28 + 18 * ld] += d;
    d = k[38];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[28 + 0 * ld] += d;
    d = k[39];  // O -> O+ + e-, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[23 + 22 * ld] += d;
    J[28 + 22 * ld] += d;
    d = k[40]*n[23];  // O+ + H -> H+ + O, d/d n[9]
    J[23 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[22 + 9 * ld] += d;
    d = k[40]*n[9];  // O+ + H -> H+ + O, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[9 + 23 * ld] -= d;
    J[10 + 23 * ld] += d;
    J[22 + 23 * ld] += d;
    d = k[41]*n[22];  // H+ + O -> O+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[22 + 10 * ld] -= d;
    J[23 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[41]*n[10];  // H+ + O -> O+ + H, d/d n[22]
    J[10 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[23 + 22 * ld] += d;
    J[9 + 22 * ld] += d;
    d = k[42]*n[23];  // O+ + H2 -> OH+ + H, d/d n[12]
    J[23 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[27 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[42]*n[12];  // O+ + H2 -> OH+ + H, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[12 + 23 * ld] -= d;
    J[27 + 23 * ld] += d;
    J[9 + 23 * ld] += d;
    d = k[43]*n[27];  // OH+ + H2 -> H2O+ + H, d/d n[12]
    J[27 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[15 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[43]*n[12];  // OH+ + H2 -> H2O+ + H, d/d n[27]
    J[27 + 27 * ld] -= d;
    J[12 + 27 * ld] -= d;
    J[15 + 27 * ld] += d;
    J[9 + 27 * ld] += d;
    d = k[44]*n[15];  // H2O+ + H2 -> H3O+ + H, d/d n[12]
    J[15 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[17 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[44]*n[12];  // H2O+ + H2 -> H3O+ + H, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[12 + 15 * ld] -= d;
    J[17 + 15 * ld] += d;
    J[9 + 15 * ld] += d;
    d = k[45]*n[28];  // H2O+ + e- -> O + H2, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[28 + 15 * ld] -= d;
    J[22 + 15 * ld] += d;
    J[12 + 15 * ld] += d;
    d = k[45]*n[15];  // H2O+ + e- -> O + H2, d/d n[28]
    J[15 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[22 + 28 * ld] += d;
    J[12 + 28 * ld] += d;
    d = k[46]*n[28];  // H2O+ + e- -> OH + H, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[28 + 15 * ld] -= d;
    J[26 + 15 * ld] += d;
    J[9 + 15 * ld] += d;
    d = k[46]*n[15];  // H2O+ + e- -> OH + H, d/d n[28]
    J[15 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[26 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[47]*n[28];  // H3O+ + e- -> H2O + H, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[28 + 17 * ld] -= d;
    J[14 + 17 * ld] += d;
    J[9 + 17 * ld] += d;
    d = k[47]*n[17];  // H3O+ + e- -> H2O + H, d/d n[28]
    J[17 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[14 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[48]*n[22];  // O + H -> OH, d/d n[9]
    J[22 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[26 + 9 * ld] += d;
    d = k[48]*n[9];  // O + H -> OH, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[9 + 22 * ld] -= d;
    J[26 + 22 * ld] += d;
    d = k[49]*n[22];  // O + H2 -> OH + H, d/d n[12]
    J[22 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[26 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[49]*n[12];  // O + H2 -> OH + H, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[12 + 22 * ld] -= d;
    J[26 + 22 * ld] += d;
    J[9 + 22 * ld] += d;
    d = k[50]*n[26];  // H2 + OH -> H2O + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[26 + 12 * ld] -= d;
    J[14 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[50]*n[12];  // H2 + OH -> H2O + H, d/d n[26]
    J[12 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[14 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[51]*n[26]*R(2.0);  // OH + OH -> H2O + O, d/d n[26]
    J[26 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[14 + 26 * ld] += d;
    J[22 + 26 * ld] += d;
    d = k[52]*n[26];  // H+ + OH -> OH+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[26 + 10 * ld] -= d;
    J[27 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[52]*n[10];  // H+ + OH -> OH+ + H, d/d n[26]
    J[10 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[27 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[53]*n[14];  // H+ + H2O -> H2O+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[14 + 10 * ld] -= d;
    J[15 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[53]*n[10];  // H+ + H2O -> H2O+ + H, d/d n[14]
    J[10 + 14 * ld] -= d;
    J[14 + 14 * ld] -= d;
    J[15 + 14 * ld] += d;
    J[9 + 14 * ld] += d;
    d = k[54]*n[26];  // H + OH -> H2 + O, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[26 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[22 + 9 * ld] += d;
    d = k[54]*n[9];  // H + OH -> H2 + O, d/d n[26]
    J[9 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[12 + 26 * ld] += d;
    J[22 + 26 * ld] += d;
    d = k[55]*n[14];  // H + H2O -> OH + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[14 + 9 * ld] -= d;
    J[26 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[55]*n[9];  // H + H2O -> OH + H2, d/d n[14]
    J[9 + 14 * ld] -= d;
    J[14 + 14 * ld] -= d;
    J[26 + 14 * ld] += d;
    J[12 + 14 * ld] += d;
    d = k[56]*n[22]*R(2.0);  // O + O -> O2, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[22 + 22 * ld] -= d;
    J[24 + 22 * ld] += d;
    d = k[57]*n[26];  // O + OH -> O2 + H, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[26 + 22 * ld] -= d;
    J[24 + 22 * ld] += d;
    J[9 + 22 * ld] += d;
    d = k[57]*n[22];  // O + OH -> O2 + H, d/d n[26]
    J[22 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[24 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[58]*n[24];  // H + O2 -> OH + O, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[24 + 9 * ld] -= d;
    J[26 + 9 * ld] += d;
    J[22 + 9 * ld] += d;
    d = k[58]*n[9];  // H + O2 -> OH + O, d/d n[24]
    J[9 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[26 + 24 * ld] += d;
    J[22 + 24 * ld] += d;
    d = k[59]*n[24];  // H+ + O2 -> O2+ + H, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[24 + 10 * ld] -= d;
    J[25 + 10 * ld] += d;
    J[9 + 10 * ld] += d;
    d = k[59]*n[10];  // H+ + O2 -> O2+ + H, d/d n[24]
    J[10 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[25 + 24 * ld] += d;
    J[9 + 24 * ld] += d;
    d = k[60]*n[28];  // O2+ + e- -> O + O, d/d n[25]
    J[25 + 25 * ld] -= d;
    J[28 + 25 * ld] -= d;
    J[22 + 25 * ld] += d;
    J[22 + 25 * ld] += d;
    d = k[60]*n[25];  // O2+ + e- -> O + O, d/d n[28]
    J[25 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[22 + 28 * ld] += d;
    J[22 + 28 * ld] += d;
    d = k[61]*n[22];  // O + CH -> CO + H, d/d n[2]
    J[22 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[6 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[61]*n[2];  // O + CH -> CO + H, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[2 + 22 * ld] -= d;
    J[6 + 22 * ld] += d;
    J[9 + 22 * ld] += d;
    d = k[62]*n[26];  // C + OH -> CO + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[26 + 0 * ld] -= d;
    J[6 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[62]*n[0];  // C + OH -> CO + H, d/d n[26]
    J[0 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[6 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[63]*n[24];  // C + O2 -> CO + O, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[24 + 0 * ld] -= d;
    J[6 + 0 * ld] += d;
    J[22 + 0 * ld] += d;
    d = k[63]*n[0];  // C + O2 -> CO + O, d/d n[24]
    J[0 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[6 + 24 * ld] += d;
    J[22 + 24 * ld] += d;
    d = k[64]*n[24];  // C+ + O2 -> O+ + CO, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[24 + 1 * ld] -= d;
    J[23 + 1 * ld] += d;
    J[6 + 1 * ld] += d;
    d = k[64]*n[1];  // C+ + O2 -> O+ + CO, d/d n[24]
    J[1 + 24 * ld] -= d;
    J[24 + 24 * ld] -= d;
    J[23 + 24 * ld] += d;
    J[6 + 24 * ld] += d;
    d = k[65]*n[26];  // OH + CO -> CO2 + H, d/d n[6]
    J[26 + 6 * ld] -= d;
    J[6 + 6 * ld] -= d;
    J[8 + 6 * ld] += d;
    J[9 + 6 * ld] += d;
    d = k[65]*n[6];  // OH + CO -> CO2 + H, d/d n[26]
    J[26 + 26 * ld] -= d;
    J[6 + 26 * ld] -= d;
    J[8 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[66]*n[28];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[28 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[66]*n[1];  // C+ + e- -> C, d/d n[28]
    J[1 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[0 + 28 * ld] += d;
    d = k[67]*n[26];  // C+ + OH -> CO+ + H, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[26 + 1 * ld] -= d;
    J[7 + 1 * ld] += d;
    J[9 + 1 * ld] += d;
    d = k[67]*n[1];  // C+ + OH -> CO+ + H, d/d n[26]
    J[1 + 26 * ld] -= d;
    J[26 + 26 * ld] -= d;
    J[7 + 26 * ld] += d;
    J[9 + 26 * ld] += d;
    d = k[68]*n[9];  // CO+ + H -> H+ + CO, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[9 + 7 * ld] -= d;
    J[10 + 7 * ld] += d;
    J[6 + 7 * ld] += d;
    d = k[68]*n[7];  // CO+ + H -> H+ + CO, d/d n[9]
    J[7 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[6 + 9 * ld] += d;
    d = k[69]*n[9];  // C + H -> CH, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[9 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    d = k[69]*n[0];  // C + H -> CH, d/d n[9]
    J[0 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    d = k[70]*n[12];  // C + H2 -> CH + H, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[12 + 0 * ld] -= d;
    J[2 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[70]*n[0];  // C + H2 -> CH + H, d/d n[12]
    J[0 + 12 * ld] -= d;
    J[12 + 12 * ld] -= d;
    J[2 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[71]*n[9];  // H + CH -> C + H2, d/d n[2]
    J[9 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[0 + 2 * ld] += d;
    J[12 + 2 * ld] += d;
    d = k[71]*n[2];  // H + CH -> C + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] -= d;
    J[0 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[72]*n[12];  // H2 + CH -> CH2 + H, d/d n[2]
    J[12 + 2 * ld] -= d;
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[72]*n[2];  // H2 + CH -> CH2 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[2 + 12 * ld] -= d;
    J[3 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[73]*n[9];  // H + CH2 -> CH + H2, d/d n[3]
    J[9 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    J[12 + 3 * ld] += d;
    d = k[73]*n[3];  // H + CH2 -> CH + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[3 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[74]*n[12];  // H2 + CH2 -> CH3 + H, d/d n[3]
    J[12 + 3 * ld] -= d;
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[9 + 3 * ld] += d;
    d = k[74]*n[3];  // H2 + CH2 -> CH3 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[3 + 12 * ld] -= d;
    J[4 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[75]*n[9];  // H + CH3 -> CH2 + H2, d/d n[4]
    J[9 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    J[12 + 4 * ld] += d;
    d = k[75]*n[4];  // H + CH3 -> CH2 + H2, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[4 + 9 * ld] -= d;
    J[3 + 9 * ld] += d;
    J[12 + 9 * ld] += d;
    d = k[76]*n[12];  // H2 + CH3 -> CH4 + H, d/d n[4]
    J[12 + 4 * ld] -= d;
    J[4 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[9 + 4 * ld] += d;
    d = k[76]*n[4];  // H2 + CH3 -> CH4 + H, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[4 + 12 * ld] -= d;
    J[5 + 12 * ld] += d;
    J[9 + 12 * ld] += d;
    d = k[77]*n[9];  // H + CH4 -> H2 + CH3, d/d n[5]
    J[9 + 5 * ld] -= d;
    J[5 + 5 * ld] -= d;
    J[12 + 5 * ld] += d;
    J[4 + 5 * ld] += d;
    d = k[77]*n[5];  // H + CH4 -> H2 + CH3, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[5 + 9 * ld] -= d;
    J[12 + 9 * ld] += d;
    J[4 + 9 * ld] += d;
    d = k[78]*n[12];  // H2 + C -> CH2, d/d n[0]
    J[12 + 0 * ld] -= d;
    J[0 + 0 * ld] -= d;
    J[3 + 0 * ld] += d;
    d = k[78]*n[0];  // H2 + C -> CH2, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[0 + 12 * ld] -= d;
    J[3 + 12 * ld] += d;
    d = k[79]*n[28];  // H3O+ + e- -> OH + H + H, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[28 + 17 * ld] -= d;
    J[26 + 17 * ld] += d;
    J[9 + 17 * ld] += d;
    J[9 + 17 * ld] += d;
    d = k[79]*n[17];  // H3O+ + e- -> OH + H + H, d/d n[28]
    J[17 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[26 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[80]*n[28];  // OH+ + e- -> O + H, d/d n[27]
    J[27 + 27 * ld] -= d;
    J[28 + 27 * ld] -= d;
    J[22 + 27 * ld] += d;
    J[9 + 27 * ld] += d;
    d = k[80]*n[27];  // OH+ + e- -> O + H, d/d n[28]
    J[27 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[22 + 28 * ld] += d;
    J[9 + 28 * ld] += d;
    d = k[81]*n[28];  // O+ + e- -> O, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[28 + 23 * ld] -= d;
    J[22 + 23 * ld] += d;
    d = k[81]*n[23];  // O+ + e- -> O, d/d n[28]
    J[23 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[22 + 28 * ld] += d;
    d = k[82]*n[28];  // CO+ + e- -> CO, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[28 + 7 * ld] -= d;
    J[6 + 7 * ld] += d;
    d = k[82]*n[7];  // CO+ + e- -> CO, d/d n[28]
    J[7 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[6 + 28 * ld] += d;
    d = k[83]*n[28];  // CO+ + e- -> C + O, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[28 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    J[22 + 7 * ld] += d;
    d = k[83]*n[7];  // CO+ + e- -> C + O, d/d n[28]
    J[7 + 28 * ld] -= d;
    J[28 + 28 * ld] -= d;
    J[0 + 28 * ld] += d;
    J[22 + 28 * ld] += d;
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
