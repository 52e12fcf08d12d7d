// generated by netgen.py from photobin.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_photobin {
  static constexpr const char* NAME = "photobin";
  static constexpr int NSPEC = 8;
  static constexpr int NREA = 10;
  static constexpr int NPAR = 3;
  static constexpr int IDX_C = 0;
  static constexpr int IDX_Cj = 1;
  static constexpr int IDX_H = 2;
  static constexpr int IDX_Hj = 3;
  static constexpr int IDX_H2 = 4;
  static constexpr int IDX_He = 5;
  static constexpr int IDX_Hej = 6;
  static constexpr int IDX_E = 7;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 7;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 3;
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
    double ph[NPHOTO];
  };
  KX_HD static Params params_from_row(const double* row) {
    Params p;
    p.crate = row[0];
    p.gamma_ad = row[1];
    p.Tcmb = row[2];
    for (int r = 0; r < NPHOTO; ++r) p.ph[r] = row[NPAR + r];
    return p;
  }
  static constexpr int photo_reaction[3] = {5, 6, 7};
  static constexpr int photo_reactant[3] = {0, 2, 5};
  static constexpr double photo_Emin[3] = {11.26, 13.6, 24.59};
  static constexpr double photo_Emax[3] = {INFINITY, INFINITY, INFINITY};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return (E_eV >= 291.0 ? kx::sigma_vy95(E_eV, 2.910e+02, 8.655e+01, 7.421e+01, 5.498e+01, 1.503e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.939e+01, 1.026e+01, 4.564e+03, 1.568e+00, 1.085e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.126e+01, 9.435e+00, 1.152e+03, 5.687e+00, 6.336e+00, 4.474e-01, 1) : kx::sigma_v96(E_eV, 2.144e+00, 5.027e+02, 6.216e+01, 5.101e+00, 9.157e-02, 1.133e+00, 1.607e+00));  // C -> C+ + e-
      case 1: return kx::sigma_v96(E_eV, 4.298e-01, 5.475e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // H -> H+ + e-
      case 2: return kx::sigma_v96(E_eV, 1.361e+01, 9.492e+02, 1.469e+00, 3.188e+00, 2.039e+00, 4.434e-01, 2.136e+00);  // He -> He+ + e-
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
      case 6: return 1;
      case 7: return -1;
      default: return 0;
    }
  }
  KX_HD static double mass_g(int i) {
    switch (i) {
      case 0: return 1.99447347289326e-23;
      case 1: return 1.9943823790562555e-23;
      case 2: return 1.6738233791328e-24;
      case 3: return 1.6729124407627538e-24;
      case 4: return 3.3476467582656e-24;
      case 5: return 6.64647366797316e-24;
      case 6: return 6.645562729603113e-24;
      case 7: return 9.10938370046373e-28;
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
      case 13: return 2;
      case 17: return 1;
      case 20: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 22 nonzeros, 22 in L + U with the fill, 27 flops per
  // factorization (dense: 341); rows in pivot order
  static constexpr int SP_NNZ = 22;
  static constexpr int SP_FLOPS = 27;
  static constexpr int sp_perm[8] = {4, 6, 5, 0, 1, 2, 3, 7};
  static constexpr int sp_rp[9] = {0, 1, 3, 4, 7, 10, 13, 16, 22};
  static constexpr int sp_ci[22] = {0, 1, 2, 2, 3, 4, 7, 3, 4, 7, 5, 6, 7, 5, 6, 7, 2, 3, 4, 5, 6, 7};
  static constexpr int sp_di[8] = {0, 1, 3, 4, 8, 10, 14, 21};
  // conserved combinations c_k . n (the nuclei of C, H, He, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 4;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 10: return 1;
      case 11: return 1;
      case 12: return 2;
      case 21: return 1;
      case 22: return 1;
      case 25: return 1;
      case 27: return 1;
      case 30: return 1;
      case 31: return -1;
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
    return s;
  }
  template <class R> KX_HD static R Hnuclei(const R* n) {
    R s = R(0);
    s += n[0] * R(0.0);
    s += n[1] * R(0.0);
    s += n[2] * R(1.0);
    s += n[3] * R(1.0);
    s += n[4] * R(2.0);
    s += n[5] * R(0.0);
    s += n[6] * R(0.0);
    s += n[7] * R(0.0);
    return s;
  }
  template <class R> KX_HD static R rho(const R* n) {
    R s = R(0);
    s += n[0] * R(1.99447347289326e-23);
    s += n[1] * R(1.9943823790562555e-23);
    s += n[2] * R(1.6738233791328e-24);
    s += n[3] * R(1.6729124407627538e-24);
    s += n[4] * R(3.3476467582656e-24);
    s += n[5] * R(6.64647366797316e-24);
    s += n[6] * R(6.645562729603113e-24);
    s += n[7] * R(9.10938370046373e-28);
    return s;
  }
  // rates of the rate temperature Trate (and parameters); the
  // caller zeroes k
  template <class R>
  KX_HD static void rates_T(R Trate, const Params& p, R* k) {
    (void)p;
    const R Tgas = Trate;
    (void)Tgas;
    {  // H+ + e- -> H: 2 piece(s), zero below 2.73, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      const R v_invTe = (R(1.0) / (Tgas * R(8.617333262145179e-05)));
      const R v_lnTe = kx::log((Tgas * R(8.617333262145179e-05)));
      if (Tgas >= R(2.73) && Tgas < R(5500.0)) k[0] = (R(3.92e-13) * kx::pow(v_invTe, R(0.6353)));  // H+ + e- -> H
      if (Tgas >= R(5500.0)) k[1] = kx::exp(((((((((((-R(28.61303380689232)) - (R(0.7241125657826851) * v_lnTe)) - (R(0.02026044731984691) * kx::ipow<2>(v_lnTe))) - (R(0.002380861877349834) * kx::ipow<3>(v_lnTe))) - (R(0.0003212605213188796) * kx::ipow<4>(v_lnTe))) - (R(1.421502914054107e-05) * kx::ipow<5>(v_lnTe))) + (R(4.989108920299513e-06) * kx::ipow<6>(v_lnTe))) + (R(5.755614137575758e-07) * kx::ipow<7>(v_lnTe))) - (R(1.856767039775261e-08) * kx::ipow<8>(v_lnTe))) - (R(3.071135243196595e-09) * kx::ipow<9>(v_lnTe))));  // H+ + e- -> H
    }
    {  // C+ + e- -> C: 3 piece(s), zero below 2.73, clamped above 21140.0
      R Tgas = Trate;
      if (Tgas > R(21140.0)) Tgas = R(21140.0);
      const R v_T32 = (Tgas / R(300.0));
      const R v_invT = (R(1.0) / Tgas);
      if (Tgas >= R(2.73) && Tgas < R(7950.0)) k[2] = (R(4.67e-12) * kx::pow(v_T32, (-R(0.6))));  // C+ + e- -> C
      if (Tgas >= R(7950.0) && Tgas < R(21140.0)) k[3] = ((R(1.23e-17) * kx::pow(v_T32, R(2.49))) * kx::exp((R(21845.6) * v_invT)));  // C+ + e- -> C
      if (Tgas >= R(21140.0)) k[4] = ((R(9.62e-08) * kx::pow(v_T32, (-R(1.37)))) * kx::exp((R(115786.2) * v_invT)));  // C+ + e- -> C
    }
    k[5] = R(p.ph[0]);  // C -> C+ + e- (photo)
    k[6] = R(p.ph[1]);  // H -> H+ + e- (photo)
    k[7] = R(p.ph[2]);  // He -> He+ + e- (photo)
    k[8] = (R(0.46) * R(p.crate));  // H -> H+ + e-
    k[9] = (R(1020.0) * R(p.crate));  // C -> C+ + e-
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
  static constexpr int NREA_T = 10;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[11] = {0, 2, 4, 6, 8, 10, 11, 12, 13, 14, 15};
  static constexpr int LT_KF_SP[15] = {3, 7, 3, 7, 1, 7, 1, 7, 1, 7, 0, 2, 5, 2, 0};
  static constexpr int LT_SC_OFF[11] = {0, 3, 6, 9, 12, 15, 18, 21, 24, 27, 30};
  static constexpr int LT_SC_ROW[30] = {3, 7, 2, 3, 7, 2, 1, 7, 0, 1, 7, 0, 1, 7, 0, 0, 1, 7, 2, 3, 7, 5, 6, 7, 2, 3, 7, 0, 1, 7};
  static constexpr int LT_SC_SGN[30] = {-1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1};
  static constexpr int LT_JT_OFF[11] = {0, 2, 4, 6, 8, 10, 11, 12, 13, 14, 15};
  static constexpr int LT_JT_COL[15] = {3, 7, 3, 7, 1, 7, 1, 7, 1, 7, 0, 2, 5, 2, 0};
  static constexpr int LT_JT_MULT[15] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 10, 10, 10, 10, 10};
  static constexpr int LT_JF_SP[10] = {7, 3, 7, 3, 7, 1, 7, 1, 7, 1};
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
    fl = k[0]*n[3]*n[7];  // H+ + e- -> H
    dn[3] -= fl;
    dn[7] -= fl;
    dn[2] += fl;
    fl = k[1]*n[3]*n[7];  // H+ + e- -> H
    dn[3] -= fl;
    dn[7] -= fl;
    dn[2] += fl;
    fl = k[2]*n[1]*n[7];  // C+ + e- -> C
    dn[1] -= fl;
    dn[7] -= fl;
    dn[0] += fl;
    fl = k[3]*n[1]*n[7];  // C+ + e- -> C
    dn[1] -= fl;
    dn[7] -= fl;
    dn[0] += fl;
    fl = k[4]*n[1]*n[7];  // C+ + e- -> C
    dn[1] -= fl;
    dn[7] -= fl;
    dn[0] += fl;
    fl = k[5]*n[0];  // C -> C+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[7] += fl;
    fl = k[6]*n[2];  // H -> H+ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[7] += fl;
    fl = k[7]*n[5];  // He -> He+ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[7] += fl;
    fl = k[8]*n[2];  // H -> H+ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[7] += fl;
    fl = k[9]*n[0];  // C -> C+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[7] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[7];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[7 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[0]*n[3];  // H+ + e- -> H, d/d n[7]
    J[3 + 7 * ld] -= d;
    J[7 + 7 * ld] -= d;
    J[2 + 7 * ld] += d;
    d = k[1]*n[7];  // H+ + e- -> H, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[7 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[1]*n[3];  // H+ + e- -> H, d/d n[7]
    J[3 + 7 * ld] -= d;
    J[7 + 7 * ld] -= d;
    J[2 + 7 * ld] += d;
    d = k[2]*n[7];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[7 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[2]*n[1];  // C+ + e- -> C, d/d n[7]
    J[1 + 7 * ld] -= d;
    J[7 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    d = k[3]*n[7];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[7 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[3]*n[1];  // C+ + e- -> C, d/d n[7]
    J[1 + 7 * ld] -= d;
    J[7 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    d = k[4]*n[7];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[7 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[4]*n[1];  // C+ + e- -> C, d/d n[7]
    J[1 + 7 * ld] -= d;
    J[7 + 7 * ld] -= d;
    J[0 + 7 * ld] += d;
    d = k[5];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[7 + 0 * ld] += d;
    d = k[6];  // H -> H+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[7 + 2 * ld] += d;
    d = k[7];  // He -> He+ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[7 + 5 * ld] += d;
    d = k[8];  // H -> H+ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[7 + 2 * ld] += d;
    d = k[9];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[7 + 0 * ld] += d;
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
