// generated by netgen.py from chianti.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_chianti {
  static constexpr const char* NAME = "chianti";
  static constexpr int NSPEC = 10;
  static constexpr int NREA = 18;
  static constexpr int NPAR = 2;
  static constexpr int IDX_C = 0;
  static constexpr int IDX_Cj = 1;
  static constexpr int IDX_Cjj = 2;
  static constexpr int IDX_Cjjj = 3;
  static constexpr int IDX_Cjjjj = 4;
  static constexpr int IDX_Cjjjjj = 5;
  static constexpr int IDX_Cjjjjjj = 6;
  static constexpr int IDX_H = 7;
  static constexpr int IDX_Hj = 8;
  static constexpr int IDX_E = 9;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 9;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 6;
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
  static constexpr int photo_reaction[6] = {0, 1, 2, 3, 4, 5};
  static constexpr int photo_reactant[6] = {0, 1, 2, 3, 4, 5};
  static constexpr double photo_Emin[6] = {11.26, 24.38, 47.89, 64.49, 392.1, 490.0};
  static constexpr double photo_Emax[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return (E_eV >= 291.0 ? kx::sigma_vy95(E_eV, 2.910e+02, 8.655e+01, 7.421e+01, 5.498e+01, 1.503e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.939e+01, 1.026e+01, 4.564e+03, 1.568e+00, 1.085e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.126e+01, 9.435e+00, 1.152e+03, 5.687e+00, 6.336e+00, 4.474e-01, 1) : kx::sigma_v96(E_eV, 2.144e+00, 5.027e+02, 6.216e+01, 5.101e+00, 9.157e-02, 1.133e+00, 1.607e+00));  // C -> C+ + e-
      case 1: return (E_eV >= 307.6 ? kx::sigma_vy95(E_eV, 3.076e+02, 9.113e+01, 6.649e+01, 9.609e+01, 1.338e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.047e+01, 2.991e+00, 1.184e+03, 3.085e+00, 1.480e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.438e+01, 1.094e+01, 1.792e+02, 3.308e+01, 4.150e+00, 5.276e-01, 1) : kx::sigma_v96(E_eV, 4.058e-01, 8.709e+00, 1.261e+02, 8.578e+00, 2.093e+00, 4.929e+01, 3.234e+00));  // C+ -> C++ + e-
      case 2: return (E_eV >= 328.9 ? kx::sigma_vy95(E_eV, 3.289e+02, 8.370e+01, 8.067e+01, 7.471e+01, 1.442e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.789e+01, 3.942e+00, 1.219e+02, 1.499e+01, 7.489e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 4.614e+00, 1.539e+04, 1.737e+00, 1.593e+01, 5.922e+00, 4.378e-03, 2.528e-02));  // C++ -> C+++ + e-
      case 3: return (E_eV >= 352.2 ? kx::sigma_vy95(E_eV, 3.522e+02, 8.412e+01, 8.111e+01, 7.459e+01, 1.428e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.449e+01, 7.843e+00, 7.109e+01, 2.828e+01, 4.754e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 3.506e+00, 1.068e+02, 1.436e+01, 7.457e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // C+++ -> C++++ + e-
      case 4: return kx::sigma_v96(E_eV, 4.624e+01, 2.344e+02, 2.183e+01, 2.581e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // C++++ -> C+++++ + e-
      case 5: return kx::sigma_v96(E_eV, 1.548e+01, 1.521e+03, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // C+++++ -> C++++++ + e-
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
      case 2: return 2;
      case 3: return 3;
      case 4: return 4;
      case 5: return 5;
      case 6: return 6;
      case 8: return 1;
      case 9: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.99447347289326e-23;
      case 1: return 1.9943823790562555e-23;
      case 2: return 1.9942912852192507e-23;
      case 3: return 1.9942001913822462e-23;
      case 4: return 1.994109097545241e-23;
      case 5: return 1.9940180037082366e-23;
      case 6: return 1.9939269098712318e-23;
      case 7: return 1.6738233791328e-24;
      case 8: return 1.6729124407627538e-24;
      case 9: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 2;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 2 + e) {
      case 0: return 1;
      case 2: return 1;
      case 4: return 1;
      case 6: return 1;
      case 8: return 1;
      case 10: return 1;
      case 12: return 1;
      case 15: return 1;
      case 17: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 36 nonzeros, 36 in L + U with the fill, 63 flops per
  // factorization (dense: 666); rows in pivot order
  static constexpr int SP_NNZ = 36;
  static constexpr int SP_FLOPS = 63;
  static constexpr int sp_perm[10] = {7, 8, 0, 1, 2, 3, 4, 5, 6, 9};
  static constexpr int sp_rp[11] = {0, 1, 2, 5, 9, 13, 17, 21, 25, 28, 36};
  static constexpr int sp_ci[36] = {0, 1, 2, 3, 9, 2, 3, 4, 9, 3, 4, 5, 9, 4, 5, 6, 9, 5, 6, 7, 9, 6, 7, 8, 9, 7, 8, 9, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int sp_di[10] = {0, 1, 2, 6, 10, 14, 18, 22, 26, 35};
  // conserved combinations c_k . n (the nuclei of C, H, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 3;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 2: return 1;
      case 3: return 1;
      case 4: return 1;
      case 5: return 1;
      case 6: return 1;
      case 17: return 1;
      case 18: return 1;
      case 21: return 1;
      case 22: return 2;
      case 23: return 3;
      case 24: return 4;
      case 25: return 5;
      case 26: return 6;
      case 28: return 1;
      case 29: return -1;
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
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(0.0);
    s += n[3] * R(0.0);
    s += n[4] * R(0.0);
    s += n[5] * R(0.0);
    s += n[6] * R(0.0);
    s += n[7] * R(1.0);
    s += n[8] * R(1.0);
    s += n[9] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.99447347289326e-23);
    s += n[1] * R(1.9943823790562555e-23);
    s += n[2] * R(1.9942912852192507e-23);
    s += n[3] * R(1.9942001913822462e-23);
    s += n[4] * R(1.994109097545241e-23);
    s += n[5] * R(1.9940180037082366e-23);
    s += n[6] * R(1.9939269098712318e-23);
    s += n[7] * R(1.6738233791328e-24);
    s += n[8] * R(1.6729124407627538e-24);
    s += n[9] * R(9.10938370046373e-28);
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
    k[1] = R(p.ph[1]);  // C+ -> C++ + e- (photo)
    k[2] = R(p.ph[2]);  // C++ -> C+++ + e- (photo)
    k[3] = R(p.ph[3]);  // C+++ -> C++++ + e- (photo)
    k[4] = R(p.ph[4]);  // C++++ -> C+++++ + e- (photo)
    k[5] = R(p.ph[5]);  // C+++++ -> C++++++ + e- (photo)
    k[6] = ((R(2.995e-09) / ((kx::sqrt((Tgas / R(0.00667))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.00667)))), (R(1.0) - (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1943000.0)))), (R(1.0) + (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(6.346e-09) * kx::exp(((-R(12.17)) / Tgas))) + (R(9.793e-09) * kx::exp(((-R(73.8)) / Tgas)))) + (R(1.634e-06) * kx::exp(((-R(15230.0)) / Tgas)))) + (R(0.0008369) * kx::exp(((-R(120700.0)) / Tgas)))) + (R(0.0003355) * kx::exp(((-R(214400.0)) / Tgas))))));  // C+ + e- -> C
    k[7] = ((R(2.067e-09) / ((kx::sqrt((Tgas / R(0.1643))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.1643)))), (R(1.0) - (R(0.8012) + (R(0.0427) * kx::exp(((-R(63410.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2172000.0)))), (R(1.0) + (R(0.8012) + (R(0.0427) * kx::exp(((-R(63410.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(3.489e-06) * kx::exp(((-R(2660.0)) / Tgas))) + (R(2.222e-07) * kx::exp(((-R(3756.0)) / Tgas)))) + (R(1.954e-05) * kx::exp(((-R(25660.0)) / Tgas)))) + (R(0.004212) * kx::exp(((-R(140000.0)) / Tgas)))) + (R(0.0002037) * kx::exp(((-R(1801000.0)) / Tgas)))) + (R(0.0002936) * kx::exp(((-R(4307000.0)) / Tgas))))));  // C++ + e- -> C+
    k[8] = ((R(1.12e-10) / ((kx::sqrt((Tgas / R(111.5))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(111.5)))), (R(1.0) - R(0.6737)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(5938000.0)))), (R(1.0) + R(0.6737))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(4.673e-07) * kx::exp(((-R(723.3)) / Tgas))) + (R(1.887e-05) * kx::exp(((-R(2847.0)) / Tgas)))) + (R(1.305e-05) * kx::exp(((-R(10540.0)) / Tgas)))) + (R(0.003099) * kx::exp(((-R(89150.0)) / Tgas)))) + (R(0.0003001) * kx::exp(((-R(281200.0)) / Tgas)))) + (R(0.002553) * kx::exp(((-R(3254000.0)) / Tgas))))));  // C+++ + e- -> C++
    k[9] = ((R(4.798e-11) / ((kx::sqrt((Tgas / R(1355.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1355.0)))), (R(1.0) - R(0.4834)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(18720000.0)))), (R(1.0) + R(0.4834))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.002646) * kx::exp(((-R(2804000.0)) / Tgas))) + (R(0.01762) * kx::exp(((-R(3485000.0)) / Tgas)))) + ((-R(0.0007843)) * kx::exp(((-R(4324000.0)) / Tgas))))));  // C++++ + e- -> C+++
    k[10] = ((R(2.044e-10) / ((kx::sqrt((Tgas / R(264.7))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(264.7)))), (R(1.0) - R(0.6742)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(27730000.0)))), (R(1.0) + R(0.6742))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.001426) * kx::exp(((-R(3116000.0)) / Tgas))) + (R(0.03046) * kx::exp(((-R(4075000.0)) / Tgas)))) + (R(0.0008373) * kx::exp(((-R(5749000.0)) / Tgas))))));  // C+++++ + e- -> C++++
    k[11] = (R(5.337e-10) / ((kx::sqrt((Tgas / R(95.02))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(95.02)))), (R(1.0) - R(0.7485)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(25170000.0)))), (R(1.0) + R(0.7485)))));  // C++++++ + e- -> C+++++
    {  // C + e- -> C+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[12] = kx::colion_v96(Tgas, R(11.3), R(0.0), R(6.85e-08), R(0.193), R(0.25));  // C + e- -> C+ + e- + e-
    }
    {  // C+ + e- -> C++ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[13] = kx::colion_v96(Tgas, R(24.4), R(1.0), R(1.86e-08), R(0.286), R(0.24));  // C+ + e- -> C++ + e- + e-
    }
    {  // C++ + e- -> C+++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[14] = kx::colion_v96(Tgas, R(47.9), R(1.0), R(6.35e-09), R(0.427), R(0.21));  // C++ + e- -> C+++ + e- + e-
    }
    {  // C+++ + e- -> C++++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[15] = kx::colion_v96(Tgas, R(64.5), R(1.0), R(1.5e-09), R(0.416), R(0.13));  // C+++ + e- -> C++++ + e- + e-
    }
    {  // C++++ + e- -> C+++++ + e- + e-: 1 piece(s), zero below 232090.4, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(232090.4)) k[16] = kx::colion_v96(Tgas, R(392.1), R(1.0), R(2.99e-10), R(0.666), R(0.02));  // C++++ + e- -> C+++++ + e- + e-
    }
    {  // C+++++ + e- -> C++++++ + e- + e-: 1 piece(s), zero below 232090.4, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(232090.4)) k[17] = kx::colion_v96(Tgas, R(490.0), R(1.0), R(1.23e-10), R(0.62), R(0.16));  // C+++++ + e- -> C++++++ + e- + e-
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
  static constexpr int NREA_T = 18;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[19] = {0, 1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30};
  static constexpr int LT_KF_SP[30] = {0, 1, 2, 3, 4, 5, 1, 9, 2, 9, 3, 9, 4, 9, 5, 9, 6, 9, 0, 9, 1, 9, 2, 9, 3, 9, 4, 9, 5, 9};
  static constexpr int LT_SC_OFF[19] = {0, 3, 6, 9, 12, 15, 18, 21, 24, 27, 30, 33, 36, 41, 46, 51, 56, 61, 66};
  static constexpr int LT_SC_ROW[66] = {0, 1, 9, 1, 2, 9, 2, 3, 9, 3, 4, 9, 4, 5, 9, 5, 6, 9, 1, 9, 0, 2, 9, 1, 3, 9, 2, 4, 9, 3, 5, 9, 4, 6, 9, 5, 0, 9, 1, 9, 9, 1, 9, 2, 9, 9, 2, 9, 3, 9, 9, 3, 9, 4, 9, 9, 4, 9, 5, 9, 9, 5, 9, 6, 9, 9};
  static constexpr int LT_SC_SGN[66] = {-1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1};
  static constexpr int LT_JT_OFF[19] = {0, 1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30};
  static constexpr int LT_JT_COL[30] = {0, 1, 2, 3, 4, 5, 1, 9, 2, 9, 3, 9, 4, 9, 5, 9, 6, 9, 0, 9, 1, 9, 2, 9, 3, 9, 4, 9, 5, 9};
  static constexpr int LT_JT_MULT[30] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[31] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24};
  static constexpr int LT_JF_SP[24] = {9, 1, 9, 2, 9, 3, 9, 4, 9, 5, 9, 6, 9, 0, 9, 1, 9, 2, 9, 3, 9, 4, 9, 5};
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
    dn[9] += fl;
    fl = k[1]*n[1];  // C+ -> C++ + e-
    dn[1] -= fl;
    dn[2] += fl;
    dn[9] += fl;
    fl = k[2]*n[2];  // C++ -> C+++ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[9] += fl;
    fl = k[3]*n[3];  // C+++ -> C++++ + e-
    dn[3] -= fl;
    dn[4] += fl;
    dn[9] += fl;
    fl = k[4]*n[4];  // C++++ -> C+++++ + e-
    dn[4] -= fl;
    dn[5] += fl;
    dn[9] += fl;
    fl = k[5]*n[5];  // C+++++ -> C++++++ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    fl = k[6]*n[1]*n[9];  // C+ + e- -> C
    dn[1] -= fl;
    dn[9] -= fl;
    dn[0] += fl;
    fl = k[7]*n[2]*n[9];  // C++ + e- -> C+
    dn[2] -= fl;
    dn[9] -= fl;
    dn[1] += fl;
    fl = k[8]*n[3]*n[9];  // C+++ + e- -> C++
    dn[3] -= fl;
    dn[9] -= fl;
    dn[2] += fl;
    fl = k[9]*n[4]*n[9];  // C++++ + e- -> C+++
    dn[4] -= fl;
    dn[9] -= fl;
    dn[3] += fl;
    fl = k[10]*n[5]*n[9];  // C+++++ + e- -> C++++
    dn[5] -= fl;
    dn[9] -= fl;
    dn[4] += fl;
    fl = k[11]*n[6]*n[9];  // C++++++ + e- -> C+++++
    dn[6] -= fl;
    dn[9] -= fl;
    dn[5] += fl;
    fl = k[12]*n[0]*n[9];  // C + e- -> C+ + e- + e-
    dn[0] -= fl;
    dn[9] -= fl;
    dn[1] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[13]*n[1]*n[9];  // C+ + e- -> C++ + e- + e-
    dn[1] -= fl;
    dn[9] -= fl;
    dn[2] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[14]*n[2]*n[9];  // C++ + e- -> C+++ + e- + e-
    dn[2] -= fl;
    dn[9] -= fl;
    dn[3] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[15]*n[3]*n[9];  // C+++ + e- -> C++++ + e- + e-
    dn[3] -= fl;
    dn[9] -= fl;
    dn[4] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[16]*n[4]*n[9];  // C++++ + e- -> C+++++ + e- + e-
    dn[4] -= fl;
    dn[9] -= fl;
    dn[5] += fl;
    dn[9] += fl;
    dn[9] += fl;
    fl = k[17]*n[5]*n[9];  // C+++++ + e- -> C++++++ + e- + e-
    dn[5] -= fl;
    dn[9] -= fl;
    dn[6] += fl;
    dn[9] += fl;
    dn[9] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[1];  // C+ -> C++ + e-, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[2 + 1 * ld] += d;
    J[9 + 1 * ld] += d;
    d = k[2];  // C++ -> C+++ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[3];  // C+++ -> C++++ + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[9 + 3 * ld] += d;
    d = k[4];  // C++++ -> C+++++ + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[9 + 4 * ld] += d;
    d = k[5];  // C+++++ -> C++++++ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[9 + 5 * ld] += d;
    d = k[6]*n[9];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[9 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[6]*n[1];  // C+ + e- -> C, d/d n[9]
    J[1 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[0 + 9 * ld] += d;
    d = k[7]*n[9];  // C++ + e- -> C+, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[9 + 2 * ld] -= d;
    J[1 + 2 * ld] += d;
    d = k[7]*n[2];  // C++ + e- -> C+, d/d n[9]
    J[2 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[1 + 9 * ld] += d;
    d = k[8]*n[9];  // C+++ + e- -> C++, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[9 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[8]*n[3];  // C+++ + e- -> C++, d/d n[9]
    J[3 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    d = k[9]*n[9];  // C++++ + e- -> C+++, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[9 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    d = k[9]*n[4];  // C++++ + e- -> C+++, d/d n[9]
    J[4 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[3 + 9 * ld] += d;
    d = k[10]*n[9];  // C+++++ + e- -> C++++, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[9 + 5 * ld] -= d;
    J[4 + 5 * ld] += d;
    d = k[10]*n[5];  // C+++++ + e- -> C++++, d/d n[9]
    J[5 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[4 + 9 * ld] += d;
    d = k[11]*n[9];  // C++++++ + e- -> C+++++, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[9 + 6 * ld] -= d;
    J[5 + 6 * ld] += d;
    d = k[11]*n[6];  // C++++++ + e- -> C+++++, d/d n[9]
    J[6 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[5 + 9 * ld] += d;
    d = k[12]*n[9];  // C + e- -> C+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[9 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    J[9 + 0 * ld] += d;
    d = k[12]*n[0];  // C + e- -> C+ + e- + e-, d/d n[9]
    J[0 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[1 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[13]*n[9];  // C+ + e- -> C++ + e- + e-, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[9 + 1 * ld] -= d;
    J[2 + 1 * ld] += d;
    J[9 + 1 * ld] += d;
    J[9 + 1 * ld] += d;
    d = k[13]*n[1];  // C+ + e- -> C++ + e- + e-, d/d n[9]
    J[1 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[2 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[14]*n[9];  // C++ + e- -> C+++ + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[9 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    J[9 + 2 * ld] += d;
    d = k[14]*n[2];  // C++ + e- -> C+++ + e- + e-, d/d n[9]
    J[2 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[3 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[15]*n[9];  // C+++ + e- -> C++++ + e- + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[9 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[9 + 3 * ld] += d;
    J[9 + 3 * ld] += d;
    d = k[15]*n[3];  // C+++ + e- -> C++++ + e- + e-, d/d n[9]
    J[3 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[4 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[16]*n[9];  // C++++ + e- -> C+++++ + e- + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[9 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[9 + 4 * ld] += d;
    J[9 + 4 * ld] += d;
    d = k[16]*n[4];  // C++++ + e- -> C+++++ + e- + e-, d/d n[9]
    J[4 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[5 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    d = k[17]*n[9];  // C+++++ + e- -> C++++++ + e- + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[9 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[9 + 5 * ld] += d;
    J[9 + 5 * ld] += d;
    d = k[17]*n[5];  // C+++++ + e- -> C++++++ + e- + e-, d/d n[9]
    J[5 + 9 * ld] -= d;
    J[9 + 9 * ld] -= d;
    J[6 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
    J[9 + 9 * ld] += d;
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
