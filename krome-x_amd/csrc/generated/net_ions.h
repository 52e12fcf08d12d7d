// generated by netgen.py from ions.knet -- do not edit
#pragma once
#include "kx_physics.h"

namespace kx {
struct net_ions {
  static constexpr const char* NAME = "ions";
  static constexpr int NSPEC = 55;
  static constexpr int NREA = 137;
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
  static constexpr int IDX_He = 9;
  static constexpr int IDX_Hej = 10;
  static constexpr int IDX_Hejj = 11;
  static constexpr int IDX_Mg = 12;
  static constexpr int IDX_Mgj = 13;
  static constexpr int IDX_Mgjj = 14;
  static constexpr int IDX_Mgjjj = 15;
  static constexpr int IDX_Mgjjjj = 16;
  static constexpr int IDX_Mgjjjjj = 17;
  static constexpr int IDX_Mgjjjjjj = 18;
  static constexpr int IDX_N = 19;
  static constexpr int IDX_Nj = 20;
  static constexpr int IDX_Njj = 21;
  static constexpr int IDX_Njjj = 22;
  static constexpr int IDX_Njjjj = 23;
  static constexpr int IDX_Njjjjj = 24;
  static constexpr int IDX_Njjjjjj = 25;
  static constexpr int IDX_Ne = 26;
  static constexpr int IDX_Nej = 27;
  static constexpr int IDX_Nejj = 28;
  static constexpr int IDX_Nejjj = 29;
  static constexpr int IDX_Nejjjj = 30;
  static constexpr int IDX_Nejjjjj = 31;
  static constexpr int IDX_Nejjjjjj = 32;
  static constexpr int IDX_O = 33;
  static constexpr int IDX_Oj = 34;
  static constexpr int IDX_Ojj = 35;
  static constexpr int IDX_Ojjj = 36;
  static constexpr int IDX_Ojjjj = 37;
  static constexpr int IDX_Ojjjjj = 38;
  static constexpr int IDX_Ojjjjjj = 39;
  static constexpr int IDX_S = 40;
  static constexpr int IDX_Sj = 41;
  static constexpr int IDX_Sjj = 42;
  static constexpr int IDX_Sjjj = 43;
  static constexpr int IDX_Sjjjj = 44;
  static constexpr int IDX_Sjjjjj = 45;
  static constexpr int IDX_Sjjjjjj = 46;
  static constexpr int IDX_Si = 47;
  static constexpr int IDX_Sij = 48;
  static constexpr int IDX_Sijj = 49;
  static constexpr int IDX_Sijjj = 50;
  static constexpr int IDX_Sijjjj = 51;
  static constexpr int IDX_Sijjjjj = 52;
  static constexpr int IDX_Sijjjjjj = 53;
  static constexpr int IDX_E = 54;
  static constexpr bool HAS_E = true;
  static constexpr int I_E = 54;
  static constexpr bool THERMAL = false;
  static constexpr int NHEAT = 0;
  static constexpr int NCOOL = 0;
  static constexpr bool RATES_DENSITY = false;
  // photo reactions (photo(xsec=...)): their rates are the bin sums
  // sum_b sigma_rb Phi_b of the cell's field, which the kernels read
  // after the parameters, in the row's last NPHOTO slots
  static constexpr int NPHOTO = 45;
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
  static constexpr int photo_reaction[45] = {92, 93, 94, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136};
  static constexpr int photo_reactant[45] = {7, 9, 10, 0, 1, 2, 3, 4, 5, 19, 20, 21, 22, 23, 24, 33, 34, 35, 36, 37, 38, 26, 27, 28, 29, 30, 31, 12, 13, 14, 15, 16, 17, 47, 48, 49, 50, 51, 52, 40, 41, 42, 43, 44, 45};
  static constexpr double photo_Emin[45] = {13.6, 24.59, 54.42, 11.26, 24.38, 47.89, 64.49, 392.1, 490.0, 14.53, 29.6, 47.45, 77.47, 97.89, 552.1, 13.62, 35.12, 54.94, 77.41, 113.9, 138.1, 21.56, 40.96, 63.46, 97.12, 126.2, 157.9, 7.646, 15.04, 80.14, 109.3, 141.3, 186.5, 8.152, 16.35, 33.49, 45.14, 166.8, 205.1, 10.36, 23.33, 34.83, 47.31, 72.68, 88.05};
  static constexpr double photo_Emax[45] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  // cross section [cm^2] of photo reaction q at photon energy E_eV
  static double photo_xsec(int q, double E_eV) {
    using R = double;
    switch (q) {
      case 0: return kx::sigma_v96(E_eV, 4.298e-01, 5.475e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // H -> H+ + e-
      case 1: return kx::sigma_v96(E_eV, 1.361e+01, 9.492e+02, 1.469e+00, 3.188e+00, 2.039e+00, 4.434e-01, 2.136e+00);  // He -> He+ + e-
      case 2: return kx::sigma_v96(E_eV, 1.720e+00, 1.369e+04, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // He+ -> He++ + e-
      case 3: return (E_eV >= 291.0 ? kx::sigma_vy95(E_eV, 2.910e+02, 8.655e+01, 7.421e+01, 5.498e+01, 1.503e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.939e+01, 1.026e+01, 4.564e+03, 1.568e+00, 1.085e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.126e+01, 9.435e+00, 1.152e+03, 5.687e+00, 6.336e+00, 4.474e-01, 1) : kx::sigma_v96(E_eV, 2.144e+00, 5.027e+02, 6.216e+01, 5.101e+00, 9.157e-02, 1.133e+00, 1.607e+00));  // C -> C+ + e-
      case 4: return (E_eV >= 307.6 ? kx::sigma_vy95(E_eV, 3.076e+02, 9.113e+01, 6.649e+01, 9.609e+01, 1.338e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.047e+01, 2.991e+00, 1.184e+03, 3.085e+00, 1.480e+01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.438e+01, 1.094e+01, 1.792e+02, 3.308e+01, 4.150e+00, 5.276e-01, 1) : kx::sigma_v96(E_eV, 4.058e-01, 8.709e+00, 1.261e+02, 8.578e+00, 2.093e+00, 4.929e+01, 3.234e+00));  // C+ -> C++ + e-
      case 5: return (E_eV >= 328.9 ? kx::sigma_vy95(E_eV, 3.289e+02, 8.370e+01, 8.067e+01, 7.471e+01, 1.442e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.789e+01, 3.942e+00, 1.219e+02, 1.499e+01, 7.489e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 4.614e+00, 1.539e+04, 1.737e+00, 1.593e+01, 5.922e+00, 4.378e-03, 2.528e-02));  // C++ -> C+++ + e-
      case 6: return (E_eV >= 352.2 ? kx::sigma_vy95(E_eV, 3.522e+02, 8.412e+01, 8.111e+01, 7.459e+01, 1.428e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.449e+01, 7.843e+00, 7.109e+01, 2.828e+01, 4.754e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 3.506e+00, 1.068e+02, 1.436e+01, 7.457e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // C+++ -> C++++ + e-
      case 7: return kx::sigma_v96(E_eV, 4.624e+01, 2.344e+02, 2.183e+01, 2.581e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // C++++ -> C+++++ + e-
      case 8: return kx::sigma_v96(E_eV, 1.548e+01, 1.521e+03, 3.288e+01, 2.963e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // C+++++ -> C++++++ + e-
      case 9: return (E_eV >= 404.8 ? kx::sigma_vy95(E_eV, 4.048e+02, 1.270e+02, 4.748e+01, 1.380e+02, 1.252e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.541e+01, 1.482e+01, 7.722e+02, 2.306e+00, 9.139e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.453e+01, 1.164e+01, 1.029e+04, 2.361e+00, 8.821e+00, 4.239e-01, 1) : kx::sigma_v96(E_eV, 4.034e+00, 8.235e+02, 8.033e+01, 3.928e+00, 9.097e-02, 8.598e-01, 2.325e+00));  // N -> N+ + e-
      case 10: return (E_eV >= 423.6 ? kx::sigma_vy95(E_eV, 4.236e+02, 1.242e+02, 5.002e+01, 9.100e+01, 1.335e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.796e+01, 1.094e+01, 7.483e+02, 2.793e+00, 9.956e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.960e+01, 1.827e+01, 1.724e+02, 8.893e+01, 3.348e+00, 4.209e-01, 1) : kx::sigma_v96(E_eV, 6.128e-02, 1.944e+00, 8.163e+02, 8.773e+00, 1.043e+01, 4.280e+02, 2.030e+01));  // N+ -> N++ + e-
      case 11: return (E_eV >= 447.3 ? kx::sigma_vy95(E_eV, 4.473e+02, 1.220e+02, 5.235e+01, 9.428e+01, 1.335e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 5.545e+01, 5.853e+00, 1.908e+02, 6.264e+00, 9.711e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.745e+01, 1.925e+01, 9.400e+01, 1.152e+02, 3.194e+00, 5.496e-01, 1) : kx::sigma_v96(E_eV, 2.420e-01, 9.375e-01, 2.788e+02, 9.156e+00, 1.850e+00, 1.877e+02, 3.999e+00));  // N++ -> N+++ + e-
      case 12: return (E_eV >= 475.3 ? kx::sigma_vy95(E_eV, 4.753e+02, 1.070e+02, 7.046e+01, 5.342e+01, 1.552e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 7.747e+01, 6.225e+00, 1.110e+02, 1.733e+01, 6.719e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 5.494e+00, 1.690e+04, 1.714e+00, 1.706e+01, 7.904e+00, 6.415e-03, 1.937e-02));  // N+++ -> N++++ + e-
      case 13: return (E_eV >= 504.3 ? kx::sigma_vy95(E_eV, 5.043e+02, 1.060e+02, 7.304e+01, 5.547e+01, 1.538e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 9.789e+01, 1.862e+01, 3.447e+01, 4.231e+01, 3.606e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 4.471e+00, 8.376e+01, 3.297e+01, 6.003e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // N++++ -> N+++++ + e-
      case 14: return kx::sigma_v96(E_eV, 6.943e+01, 1.519e+02, 2.627e+01, 2.315e+00, 0.000e+00, 0.000e+00, 0.000e+00);  // N+++++ -> N++++++ + e-
      case 15: return (E_eV >= 538.0 ? kx::sigma_vy95(E_eV, 5.380e+02, 1.774e+02, 3.237e+01, 3.812e+02, 1.083e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.848e+01, 1.994e+01, 2.415e+02, 3.241e+00, 8.037e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.362e+01, 1.391e+01, 1.220e+05, 1.364e+00, 1.140e+01, 4.103e-01, 1) : kx::sigma_v96(E_eV, 1.240e+00, 1.745e+03, 3.784e+00, 1.764e+01, 7.589e-02, 8.698e+00, 1.271e-01));  // O -> O+ + e-
      case 16: return (E_eV >= 558.1 ? kx::sigma_vy95(E_eV, 5.581e+02, 1.690e+02, 3.584e+01, 1.894e+02, 1.185e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.599e+01, 1.759e+01, 1.962e+02, 4.020e+00, 7.999e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.512e+01, 1.745e+01, 5.186e+02, 1.728e+01, 4.995e+00, 2.182e-02, 1) : kx::sigma_v96(E_eV, 1.386e+00, 5.967e+01, 3.175e+01, 8.943e+00, 1.934e-02, 2.131e+01, 1.503e-02));  // O+ -> O++ + e-
      case 17: return (E_eV >= 584.0 ? kx::sigma_vy95(E_eV, 5.840e+02, 1.620e+02, 3.939e+01, 1.104e+02, 1.289e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.551e+01, 1.594e+01, 1.217e+02, 6.156e+00, 7.271e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 5.494e+01, 2.067e+01, 2.318e+02, 7.136e+01, 3.618e+00, 5.538e-02, 1) : kx::sigma_v96(E_eV, 1.723e-01, 6.753e+02, 3.852e+02, 6.822e+00, 1.191e-01, 3.839e-03, 4.569e-01));  // O++ -> O+++ + e-
      case 18: return (E_eV >= 614.4 ? kx::sigma_vy95(E_eV, 6.144e+02, 1.593e+02, 4.123e+01, 1.141e+02, 1.287e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 8.737e+01, 7.942e+00, 1.063e+02, 9.708e+00, 8.183e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 7.741e+01, 1.937e+01, 1.619e+02, 7.312e+01, 3.648e+00, 3.760e-02, 1) : kx::sigma_v96(E_eV, 2.044e-01, 8.659e-01, 4.931e+02, 8.785e+00, 3.143e+00, 3.328e+02, 4.285e+01));  // O+++ -> O++++ + e-
      case 19: return (E_eV >= 649.1 ? kx::sigma_vy95(E_eV, 6.491e+02, 1.377e+02, 5.735e+01, 5.486e+01, 1.540e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.139e+02, 8.467e+00, 9.641e+01, 2.287e+01, 6.061e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 2.854e+00, 1.642e+04, 1.792e+00, 2.647e+01, 2.836e+01, 3.036e-02, 5.554e-02));  // O++++ -> O+++++ + e-
      case 20: return (E_eV >= 683.7 ? kx::sigma_vy95(E_eV, 6.837e+02, 1.354e+02, 6.029e+01, 5.682e+01, 1.533e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.381e+02, 9.141e+00, 6.896e+01, 3.896e+01, 4.943e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 7.824e+00, 6.864e+01, 3.210e+01, 5.495e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // O+++++ -> O++++++ + e-
      case 21: return (E_eV >= 870.1 ? kx::sigma_vy95(E_eV, 8.701e+02, 3.144e+02, 1.664e+01, 2.042e+05, 8.450e-01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.847e+01, 3.204e+01, 5.615e+01, 5.808e+00, 6.678e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.156e+01, 2.000e+01, 1.691e+04, 2.442e+00, 1.043e+01, 3.345e-01, 1) : kx::sigma_v96(E_eV, 4.870e+00, 4.287e+03, 5.798e+00, 8.355e+00, 2.434e-01, 4.236e-02, 5.873e+00));  // Ne -> Ne+ + e-
      case 22: return (E_eV >= 883.1 ? kx::sigma_vy95(E_eV, 8.831e+02, 2.452e+02, 2.783e+01, 6.075e+01, 1.420e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.374e+01, 3.187e+01, 4.025e+01, 8.495e+00, 6.038e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.096e+01, 3.428e+01, 2.766e+02, 4.179e+01, 4.029e+00, 3.052e-01, 1) : kx::sigma_v96(E_eV, 1.247e+01, 1.583e+03, 3.935e+00, 7.810e+00, 6.558e-02, 1.520e+00, 1.084e-01));  // Ne+ -> Ne++ + e-
      case 23: return (E_eV >= 913.1 ? kx::sigma_vy95(E_eV, 9.131e+02, 2.387e+02, 2.943e+01, 6.525e+01, 1.424e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 8.721e+01, 3.192e+01, 3.797e+01, 1.054e+01, 5.652e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.346e+01, 3.485e+01, 2.451e+02, 6.937e+01, 3.645e+00, 2.042e-01, 1) : kx::sigma_v96(E_eV, 7.753e-01, 5.708e+00, 6.725e+01, 1.005e+01, 4.633e-01, 7.654e+01, 2.023e+00));  // Ne++ -> Ne+++ + e-
      case 24: return (E_eV >= 948.0 ? kx::sigma_vy95(E_eV, 9.480e+02, 2.360e+02, 3.027e+01, 6.734e+01, 1.422e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.132e+02, 3.092e+01, 4.190e+01, 1.139e+01, 5.554e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 9.712e+01, 3.494e+01, 2.214e+02, 7.203e+01, 3.563e+00, 2.573e-01, 1) : kx::sigma_v96(E_eV, 5.566e+00, 1.685e+03, 6.409e+02, 3.056e+00, 8.290e-03, 5.149e+00, 6.687e+00));  // Ne+++ -> Ne++++ + e-
      case 25: return (E_eV >= 987.3 ? kx::sigma_vy95(E_eV, 9.873e+02, 2.343e+02, 3.097e+01, 6.907e+01, 1.416e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.415e+02, 2.615e+01, 5.421e+01, 1.117e+01, 5.907e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.262e+02, 3.316e+01, 1.962e+02, 7.397e+01, 3.563e+00, 2.895e-01, 1) : kx::sigma_v96(E_eV, 1.248e+00, 2.430e+00, 1.066e+02, 8.999e+00, 6.855e-01, 9.169e+01, 3.702e-01));  // Ne++++ -> Ne+++++ + e-
      case 26: return (E_eV >= 1031.0 ? kx::sigma_vy95(E_eV, 1.031e+03, 2.308e+02, 3.232e+01, 7.167e+01, 1.411e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.719e+02, 2.496e+01, 5.575e+01, 1.427e+01, 5.562e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.579e+02, 2.740e+01, 1.717e+02, 6.423e+01, 3.819e+00, 1.680e-01, 1) : kx::sigma_v96(E_eV, 1.499e+00, 9.854e-01, 1.350e+02, 8.836e+00, 1.656e+00, 1.042e+02, 1.435e+00));  // Ne+++++ -> Ne++++++ + e-
      case 27: return (E_eV >= 54.9 ? kx::sigma_vy95(E_eV, 1.311e+03, 2.711e+02, 3.561e+01, 2.374e+01, 1.952e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 9.400e+01, 4.587e+01, 1.671e+01, 2.389e+01, 4.742e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 5.490e+01, 4.937e+01, 2.023e+02, 1.079e+04, 2.960e+00, 1.463e-02, 1) + kx::sigma_vy95(E_eV, 7.646e+00, 9.393e+00, 3.034e+00, 2.625e+07, 3.923e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 1.197e+01, 1.372e+08, 2.228e-01, 1.574e+01, 2.805e-01, 0.000e+00, 0.000e+00));  // Mg -> Mg+ + e-
      case 28: return (E_eV >= 65.69 ? kx::sigma_vy95(E_eV, 1.320e+03, 3.709e+02, 1.767e+01, 1.346e+02, 1.225e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 9.881e+01, 5.142e+01, 1.290e+01, 5.775e+01, 3.903e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 6.569e+01, 4.940e+01, 2.049e+02, 4.112e+03, 2.995e+00, 2.223e-05, 1) + kx::sigma_vy95(E_eV, 1.504e+01, 8.139e+00, 3.278e+00, 4.341e+07, 3.610e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 8.139e+00, 3.278e+00, 4.341e+07, 3.610e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // Mg+ -> Mg++ + e-
      case 29: return (E_eV >= 1336.0 ? kx::sigma_vy95(E_eV, 1.336e+03, 3.607e+02, 1.877e+01, 1.401e+02, 1.238e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.111e+02, 4.804e+01, 1.914e+01, 1.613e+01, 5.050e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 8.014e+01, 5.319e+01, 2.104e+02, 7.813e+01, 3.630e+00, 2.851e-01, 1) : kx::sigma_v96(E_eV, 1.086e+01, 5.377e+02, 9.779e+00, 7.117e+00, 2.604e+00, 4.860e+00, 3.722e+00));  // Mg++ -> Mg+++ + e-
      case 30: return (E_eV >= 1356.0 ? kx::sigma_vy95(E_eV, 1.356e+03, 3.345e+02, 2.212e+01, 7.147e+01, 1.413e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.411e+02, 4.810e+01, 2.118e+01, 1.675e+01, 4.941e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.093e+02, 5.137e+01, 2.223e+02, 9.301e+01, 3.513e+00, 5.276e-03, 1) : kx::sigma_v96(E_eV, 2.912e+01, 1.394e+03, 2.895e+00, 6.487e+00, 4.326e-02, 9.402e-01, 1.135e-01));  // Mg+++ -> Mg++++ + e-
      case 31: return (E_eV >= 1400.0 ? kx::sigma_vy95(E_eV, 1.400e+03, 3.259e+02, 2.350e+01, 6.173e+01, 1.464e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.735e+02, 4.710e+01, 2.448e+01, 1.636e+01, 4.951e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.413e+02, 5.114e+01, 2.134e+02, 9.563e+01, 3.438e+00, 5.419e-03, 1) : kx::sigma_v96(E_eV, 9.762e-01, 1.728e+00, 9.184e+01, 1.006e+01, 8.090e-01, 1.276e+02, 3.979e+00));  // Mg++++ -> Mg+++++ + e-
      case 32: return (E_eV >= 1449.0 ? kx::sigma_vy95(E_eV, 1.449e+03, 2.777e+02, 3.278e+01, 5.000e+01, 1.650e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.076e+02, 4.500e+01, 2.434e+01, 2.600e+01, 4.500e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.865e+02, 4.155e+01, 2.901e+02, 5.000e+01, 4.000e+00, 3.000e-01, 1) : kx::sigma_v96(E_eV, 1.711e+00, 2.185e+00, 9.350e+01, 9.202e+00, 6.325e-01, 1.007e+02, 1.729e+00));  // Mg+++++ -> Mg++++++ + e-
      case 33: return (E_eV >= 106.0 ? kx::sigma_vy95(E_eV, 1.846e+03, 5.322e+02, 1.184e+01, 2.580e+02, 1.102e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.560e+02, 7.017e+01, 1.166e+01, 4.742e+01, 3.933e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.060e+02, 7.808e+01, 1.532e+02, 5.765e+06, 2.639e+00, 2.774e-04, 1) + kx::sigma_vy95(E_eV, 1.517e+01, 1.413e+01, 1.166e+01, 2.288e+01, 5.334e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 8.152e+00, 2.212e+01, 1.845e+02, 3.849e+00, 9.721e+00, 2.921e-01, 1) : kx::sigma_v96(E_eV, 2.317e+01, 2.506e+01, 2.057e+01, 3.546e+00, 2.837e-01, 1.672e-05, 4.207e-01));  // Si -> Si+ + e-
      case 34: return (E_eV >= 118.6 ? kx::sigma_vy95(E_eV, 1.848e+03, 4.580e+02, 1.628e+01, 7.706e+01, 1.385e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.619e+02, 6.738e+01, 1.263e+01, 3.623e+01, 4.172e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.186e+02, 7.154e+01, 1.832e+02, 3.537e+02, 3.133e+00, 2.870e-04, 1) + kx::sigma_vy95(E_eV, 2.240e+01, 1.364e+01, 8.454e+00, 7.489e+01, 4.676e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.635e+01, 2.123e+01, 6.975e+01, 4.907e+00, 9.525e+00, 3.169e-01, 1) : kx::sigma_v96(E_eV, 2.556e+00, 4.140e+00, 1.337e+01, 1.191e+01, 1.570e+00, 6.634e+00, 1.272e-01));  // Si+ -> Si++ + e-
      case 35: return (E_eV >= 131.1 ? kx::sigma_vy95(E_eV, 1.852e+03, 3.201e+02, 3.591e+01, 2.174e+01, 2.123e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.744e+02, 5.843e+01, 1.569e+01, 2.296e+01, 4.803e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.311e+02, 6.652e+01, 2.197e+02, 1.169e+02, 3.529e+00, 8.658e-04, 1) + kx::sigma_vy95(E_eV, 3.349e+01, 1.182e+01, 8.666e+00, 3.532e+02, 4.208e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 1.659e-01, 5.790e-04, 1.474e+02, 1.336e+01, 8.626e-01, 9.613e+01, 6.442e-01));  // Si++ -> Si+++ + e-
      case 36: return (E_eV >= 146.6 ? kx::sigma_vy95(E_eV, 1.868e+03, 3.131e+02, 3.784e+01, 2.017e+01, 2.186e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.899e+02, 5.977e+01, 1.486e+01, 2.786e+01, 4.565e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.466e+02, 6.482e+01, 2.326e+02, 1.188e+02, 3.547e+00, 8.417e-04, 1) + kx::sigma_vy95(E_eV, 4.514e+01, 1.288e+01, 6.083e+00, 1.356e+06, 3.353e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 1.288e+01, 6.083e+00, 1.356e+06, 3.353e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // Si+++ -> Si++++ + e-
      case 37: return (E_eV >= 1887.0 ? kx::sigma_vy95(E_eV, 1.887e+03, 3.460e+02, 3.100e+01, 1.979e+01, 2.094e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.076e+02, 6.653e+01, 1.463e+01, 2.200e+01, 4.611e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.668e+02, 7.021e+01, 2.109e+02, 8.199e+01, 3.588e+00, 6.329e-05, 1) : kx::sigma_v96(E_eV, 7.761e-01, 8.863e-01, 1.541e+02, 9.980e+00, 1.303e+00, 2.009e+02, 4.537e+00));  // Si++++ -> Si+++++ + e-
      case 38: return (E_eV >= 1946.0 ? kx::sigma_vy95(E_eV, 1.946e+03, 3.799e+02, 2.398e+01, 5.000e+01, 1.650e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.468e+02, 6.237e+01, 1.648e+01, 2.600e+01, 4.500e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.051e+02, 6.007e+01, 2.789e+02, 5.000e+01, 4.000e+00, 3.000e-01, 1) : kx::sigma_v96(E_eV, 6.305e+01, 7.293e+01, 1.558e+02, 2.400e+00, 2.989e-03, 1.115e+00, 8.051e-02));  // Si+++++ -> Si++++++ + e-
      case 39: return (E_eV >= 170.0 ? kx::sigma_vy95(E_eV, 2.477e+03, 8.114e+02, 6.649e+00, 3.734e+03, 8.646e-01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.350e+02, 1.047e+02, 8.520e+00, 9.469e+01, 3.346e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.700e+02, 9.152e+01, 1.883e+02, 7.193e+01, 3.633e+00, 2.485e-01, 1) + kx::sigma_vy95(E_eV, 2.130e+01, 1.916e+01, 1.003e+01, 3.296e+01, 5.038e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.036e+01, 2.975e+01, 5.644e+01, 1.321e+01, 7.513e+00, 2.621e-01, 1) : kx::sigma_v96(E_eV, 1.808e+01, 4.564e+04, 1.000e+00, 1.361e+01, 6.385e-01, 9.935e-01, 2.486e-01));  // S -> S+ + e-
      case 40: return (E_eV >= 184.6 ? kx::sigma_vy95(E_eV, 2.478e+03, 7.306e+02, 8.303e+00, 7.483e+02, 9.844e-01, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.387e+02, 9.694e+01, 9.659e+00, 5.188e+01, 3.738e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.846e+02, 9.058e+01, 1.896e+02, 7.538e+01, 3.635e+00, 2.934e-01, 1) + kx::sigma_vy95(E_eV, 3.190e+01, 2.047e+01, 7.824e+00, 2.396e+02, 3.950e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.333e+01, 2.890e+01, 1.054e+02, 7.474e+00, 8.421e+00, 2.840e-01, 1) : kx::sigma_v96(E_eV, 8.787e+00, 3.136e+02, 3.442e+00, 1.281e+01, 7.354e-01, 2.782e+00, 1.788e-01));  // S+ -> S++ + e-
      case 41: return (E_eV >= 199.5 ? kx::sigma_vy95(E_eV, 2.486e+03, 6.434e+02, 1.082e+01, 1.410e+02, 1.221e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.536e+02, 9.233e+01, 1.048e+01, 4.007e+01, 3.968e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 1.995e+02, 9.246e+01, 1.780e+02, 1.498e+02, 3.319e+00, 2.142e-02, 1) + kx::sigma_vy95(E_eV, 4.415e+01, 2.175e+01, 8.208e+00, 7.941e+02, 3.591e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.483e+01, 2.934e+01, 3.702e+01, 1.445e+01, 7.321e+00, 2.989e-01, 1) : kx::sigma_v96(E_eV, 2.027e+00, 6.666e+00, 5.454e+01, 8.611e+00, 4.109e+00, 1.568e+01, 9.421e+00));  // S++ -> S+++ + e-
      case 42: return (E_eV >= 216.4 ? kx::sigma_vy95(E_eV, 2.502e+03, 5.523e+02, 1.504e+01, 5.527e+01, 1.517e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.703e+02, 8.665e+01, 1.155e+01, 3.202e+01, 4.227e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.164e+02, 8.744e+01, 2.037e+02, 9.310e+01, 3.565e+00, 9.497e-03, 1) + kx::sigma_vy95(E_eV, 5.750e+01, 2.143e+01, 9.080e+00, 7.224e+02, 3.618e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 4.731e+01, 3.325e+01, 1.099e+01, 3.639e+01, 5.977e+00, 3.761e-01, 1) : kx::sigma_v96(E_eV, 2.173e+00, 2.606e+00, 6.641e+01, 8.655e+00, 1.863e+00, 1.975e+01, 3.361e+00));  // S+++ -> S++++ + e-
      case 43: return (E_eV >= 235.0 ? kx::sigma_vy95(E_eV, 2.522e+03, 4.813e+02, 2.029e+01, 3.854e+01, 1.736e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.888e+02, 7.355e+01, 1.407e+01, 2.397e+01, 4.754e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.350e+02, 7.411e+01, 2.919e+02, 4.864e+01, 4.142e+00, 2.785e-02, 1) + kx::sigma_vy95(E_eV, 7.268e+01, 1.271e+01, 1.363e+01, 2.570e+02, 4.361e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 1.713e-01, 5.072e-04, 1.986e+02, 1.307e+01, 7.880e-01, 9.424e+01, 6.265e-01));  // S++++ -> S+++++ + e-
      case 44: return (E_eV >= 255.7 ? kx::sigma_vy95(E_eV, 2.544e+03, 4.681e+02, 2.148e+01, 4.092e+01, 1.738e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 3.094e+02, 6.780e+01, 1.495e+01, 2.519e+01, 4.826e+00, 0.000e+00, 0) + kx::sigma_vy95(E_eV, 2.557e+02, 5.747e+01, 4.712e+02, 3.610e+01, 4.742e+00, 2.480e-02, 1) + kx::sigma_vy95(E_eV, 8.805e+01, 1.413e+01, 9.139e+00, 1.656e+03, 3.626e+00, 0.000e+00, 0) : kx::sigma_v96(E_eV, 1.413e+01, 9.139e+00, 1.656e+03, 3.626e+00, 0.000e+00, 0.000e+00, 0.000e+00));  // S+++++ -> S++++++ + e-
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
      case 10: return 1;
      case 11: return 2;
      case 13: return 1;
      case 14: return 2;
      case 15: return 3;
      case 16: return 4;
      case 17: return 5;
      case 18: return 6;
      case 20: return 1;
      case 21: return 2;
      case 22: return 3;
      case 23: return 4;
      case 24: return 5;
      case 25: return 6;
      case 27: return 1;
      case 28: return 2;
      case 29: return 3;
      case 30: return 4;
      case 31: return 5;
      case 32: return 6;
      case 34: return 1;
      case 35: return 2;
      case 36: return 3;
      case 37: return 4;
      case 38: return 5;
      case 39: return 6;
      case 41: return 1;
      case 42: return 2;
      case 43: return 3;
      case 44: return 4;
      case 45: return 5;
      case 46: return 6;
      case 48: return 1;
      case 49: return 2;
      case 50: return 3;
      case 51: return 4;
      case 52: return 5;
      case 53: return 6;
      case 54: return -1;
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
      case 9: return 6.64647366797316e-24;
      case 10: return 6.645562729603113e-24;
      case 11: return 6.644651791233067e-24;
      case 12: return 4.0359402013713e-23;
      case 13: return 4.035849107534295e-23;
      case 14: return 4.035758013697291e-23;
      case 15: return 4.035666919860286e-23;
      case 16: return 4.0355758260232815e-23;
      case 17: return 4.0354847321862764e-23;
      case 18: return 4.0353936383492725e-23;
      case 19: return 2.3259170705866198e-23;
      case 20: return 2.3258259767496153e-23;
      case 21: return 2.3257348829126106e-23;
      case 22: return 2.325643789075606e-23;
      case 23: return 2.3255526952386013e-23;
      case 24: return 2.3254616014015968e-23;
      case 25: return 2.325370507564592e-23;
      case 26: return 3.3509678363988e-23;
      case 27: return 3.350876742561795e-23;
      case 28: return 3.350785648724791e-23;
      case 29: return 3.350694554887786e-23;
      case 30: return 3.3506034610507816e-23;
      case 31: return 3.3505123672137765e-23;
      case 32: return 3.3504212733767726e-23;
      case 33: return 2.65669645265334e-23;
      case 34: return 2.6566053588163353e-23;
      case 35: return 2.656514264979331e-23;
      case 36: return 2.6564231711423263e-23;
      case 37: return 2.656332077305321e-23;
      case 38: return 2.656240983468317e-23;
      case 39: return 2.656149889631312e-23;
      case 40: return 5.3236882475196e-23;
      case 41: return 5.323597153682596e-23;
      case 42: return 5.323506059845591e-23;
      case 43: return 5.323414966008586e-23;
      case 44: return 5.323323872171582e-23;
      case 45: return 5.323232778334578e-23;
      case 46: return 5.323141684497573e-23;
      case 47: return 4.6636239685461e-23;
      case 48: return 4.6635328747090956e-23;
      case 49: return 4.663441780872091e-23;
      case 50: return 4.6633506870350866e-23;
      case 51: return 4.6632595931980815e-23;
      case 52: return 4.663168499361077e-23;
      case 53: return 4.6630774055240725e-23;
      case 54: return 9.10938370046373e-28;
      default: return 0.0;
    }
  }
  static constexpr int NELEM = 9;
  KX_HD static constexpr int natoms(int i, int e) {
    switch (i * 9 + e) {
      case 0: return 1;
      case 9: return 1;
      case 18: return 1;
      case 27: return 1;
      case 36: return 1;
      case 45: return 1;
      case 54: return 1;
      case 64: return 1;
      case 73: return 1;
      case 83: return 1;
      case 92: return 1;
      case 101: return 1;
      case 111: return 1;
      case 120: return 1;
      case 129: return 1;
      case 138: return 1;
      case 147: return 1;
      case 156: return 1;
      case 165: return 1;
      case 175: return 1;
      case 184: return 1;
      case 193: return 1;
      case 202: return 1;
      case 211: return 1;
      case 220: return 1;
      case 229: return 1;
      case 239: return 1;
      case 248: return 1;
      case 257: return 1;
      case 266: return 1;
      case 275: return 1;
      case 284: return 1;
      case 293: return 1;
      case 303: return 1;
      case 312: return 1;
      case 321: return 1;
      case 330: return 1;
      case 339: return 1;
      case 348: return 1;
      case 357: return 1;
      case 367: return 1;
      case 376: return 1;
      case 385: return 1;
      case 394: return 1;
      case 403: return 1;
      case 412: return 1;
      case 421: return 1;
      case 431: return 1;
      case 440: return 1;
      case 449: return 1;
      case 458: return 1;
      case 467: return 1;
      case 476: return 1;
      case 485: return 1;
      default: return 0;
    }
  }
  // static-pivot sparse LU of the species block of I - gamma J:
  // 253 nonzeros, 253 in L + U with the fill, 477 flops per
  // factorization (dense: 110916); rows in pivot order
  static constexpr int SP_NNZ = 253;
  static constexpr int SP_FLOPS = 477;
  static constexpr int sp_perm[55] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54};
  static constexpr int sp_rp[56] = {0, 3, 7, 11, 15, 19, 23, 26, 29, 32, 35, 39, 42, 45, 49, 53, 57, 61, 65, 68, 71, 75, 79, 83, 87, 91, 94, 97, 101, 105, 109, 113, 117, 120, 123, 127, 131, 135, 139, 143, 146, 149, 153, 157, 161, 165, 169, 172, 175, 179, 183, 187, 191, 195, 198, 253};
  static constexpr int sp_ci[253] = {0, 1, 54, 0, 1, 2, 54, 1, 2, 3, 54, 2, 3, 4, 54, 3, 4, 5, 54, 4, 5, 6, 54, 5, 6, 54, 7, 8, 54, 7, 8, 54, 9, 10, 54, 9, 10, 11, 54, 10, 11, 54, 12, 13, 54, 12, 13, 14, 54, 13, 14, 15, 54, 14, 15, 16, 54, 15, 16, 17, 54, 16, 17, 18, 54, 17, 18, 54, 19, 20, 54, 19, 20, 21, 54, 20, 21, 22, 54, 21, 22, 23, 54, 22, 23, 24, 54, 23, 24, 25, 54, 24, 25, 54, 26, 27, 54, 26, 27, 28, 54, 27, 28, 29, 54, 28, 29, 30, 54, 29, 30, 31, 54, 30, 31, 32, 54, 31, 32, 54, 33, 34, 54, 33, 34, 35, 54, 34, 35, 36, 54, 35, 36, 37, 54, 36, 37, 38, 54, 37, 38, 39, 54, 38, 39, 54, 40, 41, 54, 40, 41, 42, 54, 41, 42, 43, 54, 42, 43, 44, 54, 43, 44, 45, 54, 44, 45, 46, 54, 45, 46, 54, 47, 48, 54, 47, 48, 49, 54, 48, 49, 50, 54, 49, 50, 51, 54, 50, 51, 52, 54, 51, 52, 53, 54, 52, 53, 54, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54};
  static constexpr int sp_di[55] = {0, 4, 8, 12, 16, 20, 24, 26, 30, 32, 36, 40, 42, 46, 50, 54, 58, 62, 66, 68, 72, 76, 80, 84, 88, 92, 94, 98, 102, 106, 110, 114, 118, 120, 124, 128, 132, 136, 140, 144, 146, 150, 154, 158, 162, 166, 170, 172, 176, 180, 184, 188, 192, 196, 252};
  // conserved combinations c_k . n (the nuclei of C, H, He, Mg, N, Ne, O, S, Si, the charge): every
  // reaction keeps them, so c_k^T J = 0 over the species
  static constexpr int NCONS = 10;
  KX_HD static constexpr int cons(int k, int i) {
    switch (k * NSPEC + i) {
      case 0: return 1;
      case 1: return 1;
      case 2: return 1;
      case 3: return 1;
      case 4: return 1;
      case 5: return 1;
      case 6: return 1;
      case 62: return 1;
      case 63: return 1;
      case 119: return 1;
      case 120: return 1;
      case 121: return 1;
      case 177: return 1;
      case 178: return 1;
      case 179: return 1;
      case 180: return 1;
      case 181: return 1;
      case 182: return 1;
      case 183: return 1;
      case 239: return 1;
      case 240: return 1;
      case 241: return 1;
      case 242: return 1;
      case 243: return 1;
      case 244: return 1;
      case 245: return 1;
      case 301: return 1;
      case 302: return 1;
      case 303: return 1;
      case 304: return 1;
      case 305: return 1;
      case 306: return 1;
      case 307: return 1;
      case 363: return 1;
      case 364: return 1;
      case 365: return 1;
      case 366: return 1;
      case 367: return 1;
      case 368: return 1;
      case 369: return 1;
      case 425: return 1;
      case 426: return 1;
      case 427: return 1;
      case 428: return 1;
      case 429: return 1;
      case 430: return 1;
      case 431: return 1;
      case 487: return 1;
      case 488: return 1;
      case 489: return 1;
      case 490: return 1;
      case 491: return 1;
      case 492: return 1;
      case 493: return 1;
      case 496: return 1;
      case 497: return 2;
      case 498: return 3;
      case 499: return 4;
      case 500: return 5;
      case 501: return 6;
      case 503: return 1;
      case 505: return 1;
      case 506: return 2;
      case 508: return 1;
      case 509: return 2;
      case 510: return 3;
      case 511: return 4;
      case 512: return 5;
      case 513: return 6;
      case 515: return 1;
      case 516: return 2;
      case 517: return 3;
      case 518: return 4;
      case 519: return 5;
      case 520: return 6;
      case 522: return 1;
      case 523: return 2;
      case 524: return 3;
      case 525: return 4;
      case 526: return 5;
      case 527: return 6;
      case 529: return 1;
      case 530: return 2;
      case 531: return 3;
      case 532: return 4;
      case 533: return 5;
      case 534: return 6;
      case 536: return 1;
      case 537: return 2;
      case 538: return 3;
      case 539: return 4;
      case 540: return 5;
      case 541: return 6;
      case 543: return 1;
      case 544: return 2;
      case 545: return 3;
      case 546: return 4;
      case 547: return 5;
      case 548: return 6;
      case 549: return -1;
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
    s += n[29];
    s += n[30];
    s += n[31];
    s += n[32];
    s += n[33];
    s += n[34];
    s += n[35];
    s += n[36];
    s += n[37];
    s += n[38];
    s += n[39];
    s += n[40];
    s += n[41];
    s += n[42];
    s += n[43];
    s += n[44];
    s += n[45];
    s += n[46];
    s += n[47];
    s += n[48];
    s += n[49];
    s += n[50];
    s += n[51];
    s += n[52];
    s += n[53];
    s += n[54];
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
    s += n[10] * R(0.0);
    s += n[11] * R(0.0);
    s += n[12] * R(0.0);
    s += n[13] * R(0.0);
    s += n[14] * R(0.0);
    s += n[15] * R(0.0);
    s += n[16] * R(0.0);
    s += n[17] * R(0.0);
    s += n[18] * R(0.0);
    s += n[19] * R(0.0);
    s += n[20] * R(0.0);
    s += n[21] * R(0.0);
    s += n[22] * R(0.0);
    s += n[23] * R(0.0);
    s += n[24] * R(0.0);
    s += n[25] * R(0.0);
    s += n[26] * R(0.0);
    s += n[27] * R(0.0);
    s += n[28] * R(0.0);
    s += n[29] * R(0.0);
    s += n[30] * R(0.0);
    s += n[31] * R(0.0);
    s += n[32] * R(0.0);
    s += n[33] * R(0.0);
    s += n[34] * R(0.0);
    s += n[35] * R(0.0);
    s += n[36] * R(0.0);
    s += n[37] * R(0.0);
    s += n[38] * R(0.0);
    s += n[39] * R(0.0);
    s += n[40] * R(0.0);
    s += n[41] * R(0.0);
    s += n[42] * R(0.0);
    s += n[43] * R(0.0);
    s += n[44] * R(0.0);
    s += n[45] * R(0.0);
    s += n[46] * R(0.0);
    s += n[47] * R(0.0);
    s += n[48] * R(0.0);
    s += n[49] * R(0.0);
    s += n[50] * R(0.0);
    s += n[51] * R(0.0);
    s += n[52] * R(0.0);
    s += n[53] * R(0.0);
    s += n[54] * R(0.0);
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
    s += n[9] * R(6.64647366797316e-24);
    s += n[10] * R(6.645562729603113e-24);
    s += n[11] * R(6.644651791233067e-24);
    s += n[12] * R(4.0359402013713e-23);
    s += n[13] * R(4.035849107534295e-23);
    s += n[14] * R(4.035758013697291e-23);
    s += n[15] * R(4.035666919860286e-23);
    s += n[16] * R(4.0355758260232815e-23);
    s += n[17] * R(4.0354847321862764e-23);
    s += n[18] * R(4.0353936383492725e-23);
    s += n[19] * R(2.3259170705866198e-23);
    s += n[20] * R(2.3258259767496153e-23);
    s += n[21] * R(2.3257348829126106e-23);
    s += n[22] * R(2.325643789075606e-23);
    s += n[23] * R(2.3255526952386013e-23);
    s += n[24] * R(2.3254616014015968e-23);
    s += n[25] * R(2.325370507564592e-23);
    s += n[26] * R(3.3509678363988e-23);
    s += n[27] * R(3.350876742561795e-23);
    s += n[28] * R(3.350785648724791e-23);
    s += n[29] * R(3.350694554887786e-23);
    s += n[30] * R(3.3506034610507816e-23);
    s += n[31] * R(3.3505123672137765e-23);
    s += n[32] * R(3.3504212733767726e-23);
    s += n[33] * R(2.65669645265334e-23);
    s += n[34] * R(2.6566053588163353e-23);
    s += n[35] * R(2.656514264979331e-23);
    s += n[36] * R(2.6564231711423263e-23);
    s += n[37] * R(2.656332077305321e-23);
    s += n[38] * R(2.656240983468317e-23);
    s += n[39] * R(2.656149889631312e-23);
    s += n[40] * R(5.3236882475196e-23);
    s += n[41] * R(5.323597153682596e-23);
    s += n[42] * R(5.323506059845591e-23);
    s += n[43] * R(5.323414966008586e-23);
    s += n[44] * R(5.323323872171582e-23);
    s += n[45] * R(5.323232778334578e-23);
    s += n[46] * R(5.323141684497573e-23);
    s += n[47] * R(4.6636239685461e-23);
    s += n[48] * R(4.6635328747090956e-23);
    s += n[49] * R(4.663441780872091e-23);
    s += n[50] * R(4.6633506870350866e-23);
    s += n[51] * R(4.6632595931980815e-23);
    s += n[52] * R(4.663168499361077e-23);
    s += n[53] * R(4.6630774055240725e-23);
    s += n[54] * R(9.10938370046373e-28);
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
      if (Tgas >= R(2.73) && Tgas < R(5500.0)) k[0] = (R(3.92e-13) * kx::pow((R(1.0) / (Tgas * R(8.617333262145179e-05))), R(0.6353)));  // H+ + e- -> H
      if (Tgas >= R(5500.0)) k[1] = kx::exp(((((((((((-R(28.61303380689232)) - (R(0.7241125657826851) * kx::log((Tgas * R(8.617333262145179e-05))))) - (R(0.02026044731984691) * kx::ipow<2>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.002380861877349834) * kx::ipow<3>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.0003212605213188796) * kx::ipow<4>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(1.421502914054107e-05) * kx::ipow<5>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(4.989108920299513e-06) * kx::ipow<6>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(5.755614137575758e-07) * kx::ipow<7>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(1.856767039775261e-08) * kx::ipow<8>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(3.071135243196595e-09) * kx::ipow<9>(kx::log((Tgas * R(8.617333262145179e-05)))))));  // H+ + e- -> H
    }
    {  // He+ + e- -> He: 2 piece(s), zero below 2.73, clamped above 100000000.0
      R Tgas = Trate;
      if (Tgas > R(100000000.0)) Tgas = R(100000000.0);
      if (Tgas >= R(2.73) && Tgas < R(9280.0)) k[2] = (R(3.92e-13) * kx::pow((R(1.0) / (Tgas * R(8.617333262145179e-05))), R(0.6353)));  // He+ + e- -> He
      if (Tgas >= R(9280.0)) k[3] = (((R(1.54e-09) * (R(1.0) + (R(0.3) / kx::exp((R(8.099328789667) * (R(1.0) / (Tgas * R(8.617333262145179e-05)))))))) / (kx::exp((R(40.49664394833662) * (R(1.0) / (Tgas * R(8.617333262145179e-05))))) * kx::pow((Tgas * R(8.617333262145179e-05)), R(1.5)))) + (R(3.92e-13) / kx::pow((Tgas * R(8.617333262145179e-05)), R(0.6353))));  // He+ + e- -> He
    }
    k[4] = (((R(3.36e-10) / kx::sqrt(Tgas)) / kx::pow((Tgas * R(0.001)), R(0.2))) / (R(1.0) + kx::pow((Tgas * R(1e-06)), R(0.7))));  // He++ + e- -> He+
    k[5] = ((R(2.995e-09) / ((kx::sqrt((Tgas / R(0.00667))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.00667)))), (R(1.0) - (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1943000.0)))), (R(1.0) + (R(0.7849) + (R(0.1597) * kx::exp(((-R(49550.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(6.346e-09) * kx::exp(((-R(12.17)) / Tgas))) + (R(9.793e-09) * kx::exp(((-R(73.8)) / Tgas)))) + (R(1.634e-06) * kx::exp(((-R(15230.0)) / Tgas)))) + (R(0.0008369) * kx::exp(((-R(120700.0)) / Tgas)))) + (R(0.0003355) * kx::exp(((-R(214400.0)) / Tgas))))));  // C+ + e- -> C
    k[6] = ((R(2.067e-09) / ((kx::sqrt((Tgas / R(0.1643))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.1643)))), (R(1.0) - (R(0.8012) + (R(0.0427) * kx::exp(((-R(63410.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2172000.0)))), (R(1.0) + (R(0.8012) + (R(0.0427) * kx::exp(((-R(63410.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(3.489e-06) * kx::exp(((-R(2660.0)) / Tgas))) + (R(2.222e-07) * kx::exp(((-R(3756.0)) / Tgas)))) + (R(1.954e-05) * kx::exp(((-R(25660.0)) / Tgas)))) + (R(0.004212) * kx::exp(((-R(140000.0)) / Tgas)))) + (R(0.0002037) * kx::exp(((-R(1801000.0)) / Tgas)))) + (R(0.0002936) * kx::exp(((-R(4307000.0)) / Tgas))))));  // C++ + e- -> C+
    k[7] = ((R(1.12e-10) / ((kx::sqrt((Tgas / R(111.5))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(111.5)))), (R(1.0) - R(0.6737)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(5938000.0)))), (R(1.0) + R(0.6737))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(4.673e-07) * kx::exp(((-R(723.3)) / Tgas))) + (R(1.887e-05) * kx::exp(((-R(2847.0)) / Tgas)))) + (R(1.305e-05) * kx::exp(((-R(10540.0)) / Tgas)))) + (R(0.003099) * kx::exp(((-R(89150.0)) / Tgas)))) + (R(0.0003001) * kx::exp(((-R(281200.0)) / Tgas)))) + (R(0.002553) * kx::exp(((-R(3254000.0)) / Tgas))))));  // C+++ + e- -> C++
    k[8] = ((R(4.798e-11) / ((kx::sqrt((Tgas / R(1355.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1355.0)))), (R(1.0) - R(0.4834)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(18720000.0)))), (R(1.0) + R(0.4834))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.002646) * kx::exp(((-R(2804000.0)) / Tgas))) + (R(0.01762) * kx::exp(((-R(3485000.0)) / Tgas)))) + ((-R(0.0007843)) * kx::exp(((-R(4324000.0)) / Tgas))))));  // C++++ + e- -> C+++
    k[9] = ((R(2.044e-10) / ((kx::sqrt((Tgas / R(264.7))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(264.7)))), (R(1.0) - R(0.6742)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(27730000.0)))), (R(1.0) + R(0.6742))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.001426) * kx::exp(((-R(3116000.0)) / Tgas))) + (R(0.03046) * kx::exp(((-R(4075000.0)) / Tgas)))) + (R(0.0008373) * kx::exp(((-R(5749000.0)) / Tgas))))));  // C+++++ + e- -> C++++
    k[10] = (R(5.337e-10) / ((kx::sqrt((Tgas / R(95.02))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(95.02)))), (R(1.0) - R(0.7485)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(25170000.0)))), (R(1.0) + R(0.7485)))));  // C++++++ + e- -> C+++++
    k[11] = ((R(6.387e-10) / ((kx::sqrt((Tgas / R(0.09467))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.09467)))), (R(1.0) - (R(0.7308) + (R(0.244) * kx::exp(((-R(67390.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2954000.0)))), (R(1.0) + (R(0.7308) + (R(0.244) * kx::exp(((-R(67390.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(1.658e-08) * kx::exp(((-R(12.65)) / Tgas))) + (R(2.76e-08) * kx::exp(((-R(84.25)) / Tgas)))) + (R(2.391e-09) * kx::exp(((-R(296.4)) / Tgas)))) + (R(7.585e-07) * kx::exp(((-R(5923.0)) / Tgas)))) + (R(0.0003012) * kx::exp(((-R(127800.0)) / Tgas)))) + (R(0.0007132) * kx::exp(((-R(218400.0)) / Tgas))))));  // N+ + e- -> N
    k[12] = ((R(2.41e-09) / ((kx::sqrt((Tgas / R(0.1231))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.1231)))), (R(1.0) - (R(0.7948) + (R(0.0774) * kx::exp(((-R(101600.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(3016000.0)))), (R(1.0) + (R(0.7948) + (R(0.0774) * kx::exp(((-R(101600.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(7.712e-08) * kx::exp(((-R(71.13)) / Tgas))) + (R(4.839e-08) * kx::exp(((-R(276.5)) / Tgas)))) + (R(2.218e-06) * kx::exp(((-R(14390.0)) / Tgas)))) + (R(0.001536) * kx::exp(((-R(134700.0)) / Tgas)))) + (R(0.003647) * kx::exp(((-R(249600.0)) / Tgas)))) + (R(4.234e-05) * kx::exp(((-R(2204000.0)) / Tgas))))));  // N++ + e- -> N+
    k[13] = ((R(7.923e-10) / ((kx::sqrt((Tgas / R(3.75))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(3.75)))), (R(1.0) - (R(0.7768) + (R(0.0223) * kx::exp(((-R(72060.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(3468000.0)))), (R(1.0) + (R(0.7768) + (R(0.0223) * kx::exp(((-R(72060.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(3.386e-06) * kx::exp(((-R(1406.0)) / Tgas))) + (R(3.036e-05) * kx::exp(((-R(6965.0)) / Tgas)))) + (R(5.945e-05) * kx::exp(((-R(26040.0)) / Tgas)))) + (R(0.001195) * kx::exp(((-R(130400.0)) / Tgas)))) + (R(0.006462) * kx::exp(((-R(196500.0)) / Tgas)))) + (R(0.001358) * kx::exp(((-R(4466000.0)) / Tgas))))));  // N+++ + e- -> N++
    k[14] = ((R(1.533e-10) / ((kx::sqrt((Tgas / R(182.3))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(182.3)))), (R(1.0) - R(0.6682)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(7751000.0)))), (R(1.0) + R(0.6682))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.04e-06) * kx::exp(((-R(3084.0)) / Tgas))) + (R(6.986e-05) * kx::exp(((-R(13320.0)) / Tgas)))) + (R(0.0003168) * kx::exp(((-R(64750.0)) / Tgas)))) + (R(0.004353) * kx::exp(((-R(118100.0)) / Tgas)))) + (R(0.0007765) * kx::exp(((-R(668700.0)) / Tgas)))) + (R(0.005101) * kx::exp(((-R(4778000.0)) / Tgas))))));  // N++++ + e- -> N+++
    k[15] = ((R(6.245e-11) / ((kx::sqrt((Tgas / R(1957.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1957.0)))), (R(1.0) - R(0.4985)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(21770000.0)))), (R(1.0) + R(0.4985))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.005761) * kx::exp(((-R(3860000.0)) / Tgas))) + (R(0.03434) * kx::exp(((-R(4883000.0)) / Tgas)))) + ((-R(0.00166)) * kx::exp(((-R(6259000.0)) / Tgas))))));  // N+++++ + e- -> N++++
    k[16] = ((R(2.388e-10) / ((kx::sqrt((Tgas / R(396.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(396.0)))), (R(1.0) - R(0.6732)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(35830000.0)))), (R(1.0) + R(0.6732))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.002801) * kx::exp(((-R(4198000.0)) / Tgas))) + (R(0.04362) * kx::exp(((-R(5516000.0)) / Tgas)))) + (R(0.001117) * kx::exp(((-R(8050000.0)) / Tgas))))));  // N++++++ + e- -> N+++++
    k[17] = ((R(6.622e-11) / ((kx::sqrt((Tgas / R(4.136))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4.136)))), (R(1.0) - (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4214000.0)))), (R(1.0) + (R(0.6109) + (R(0.4093) * kx::exp(((-R(87700.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((R(5.629e-08) * kx::exp(((-R(5395.0)) / Tgas))) + (R(2.55e-07) * kx::exp(((-R(17700.0)) / Tgas)))) + (R(0.0006173) * kx::exp(((-R(167100.0)) / Tgas)))) + (R(0.0001627) * kx::exp(((-R(268700.0)) / Tgas))))));  // O+ + e- -> O
    k[18] = ((R(2.096e-09) / ((kx::sqrt((Tgas / R(0.1602))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.1602)))), (R(1.0) - (R(0.7668) + (R(0.107) * kx::exp(((-R(139200.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4377000.0)))), (R(1.0) + (R(0.7668) + (R(0.107) * kx::exp(((-R(139200.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(1.627e-07) * kx::exp(((-R(45.35)) / Tgas))) + (R(1.262e-07) * kx::exp(((-R(284.7)) / Tgas)))) + (R(6.663e-07) * kx::exp(((-R(4166.0)) / Tgas)))) + (R(3.925e-06) * kx::exp(((-R(28770.0)) / Tgas)))) + (R(0.002406) * kx::exp(((-R(195300.0)) / Tgas)))) + (R(0.001146) * kx::exp(((-R(364600.0)) / Tgas))))));  // O++ + e- -> O+
    k[19] = ((R(2.501e-09) / ((kx::sqrt((Tgas / R(0.5235))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.5235)))), (R(1.0) - (R(0.7844) + (R(0.0447) * kx::exp(((-R(164200.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(4470000.0)))), (R(1.0) + (R(0.7844) + (R(0.0447) * kx::exp(((-R(164200.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(3.932e-07) * kx::exp(((-R(150.9)) / Tgas))) + (R(2.523e-07) * kx::exp(((-R(621.1)) / Tgas)))) + (R(3.447e-05) * kx::exp(((-R(15620.0)) / Tgas)))) + (R(0.005776) * kx::exp(((-R(193600.0)) / Tgas)))) + (R(0.005101) * kx::exp(((-R(470000.0)) / Tgas))))));  // O+++ + e- -> O++
    k[20] = ((R(3.955e-09) / ((kx::sqrt((Tgas / R(0.6821))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(0.6821)))), (R(1.0) - R(0.7813)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(5076000.0)))), (R(1.0) + R(0.7813))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(1.615e-05) * kx::exp(((-R(756.9)) / Tgas))) + (R(9.299e-06) * kx::exp(((-R(3659.0)) / Tgas)))) + (R(0.000153) * kx::exp(((-R(19840.0)) / Tgas)))) + (R(0.0006616) * kx::exp(((-R(84290.0)) / Tgas)))) + (R(0.0108) * kx::exp(((-R(229400.0)) / Tgas)))) + (R(0.0007503) * kx::exp(((-R(1161000.0)) / Tgas)))) + (R(0.002892) * kx::exp(((-R(6137000.0)) / Tgas))))));  // O++++ + e- -> O+++
    k[21] = ((R(1.724e-10) / ((kx::sqrt((Tgas / R(337.2))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(337.2)))), (R(1.0) - R(0.6556)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(10300000.0)))), (R(1.0) + R(0.6556))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.389e-05) * kx::exp(((-R(23260.0)) / Tgas))) + (R(0.0001355) * kx::exp(((-R(32090.0)) / Tgas)))) + (R(0.005885) * kx::exp(((-R(131600.0)) / Tgas)))) + (R(0.002163) * kx::exp(((-R(673100.0)) / Tgas)))) + (R(0.0006341) * kx::exp(((-R(1892000.0)) / Tgas)))) + (R(0.01348) * kx::exp(((-R(6150000.0)) / Tgas))))));  // O+++++ + e- -> O++++
    k[22] = ((R(8.193e-11) / ((kx::sqrt((Tgas / R(2392.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2392.0)))), (R(1.0) - R(0.5165)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(24870000.0)))), (R(1.0) + R(0.5165))))) + (kx::pow(Tgas, (-R(1.5))) * ((R(0.06135) * kx::exp(((-R(6113000.0)) / Tgas))) + (R(0.0001968) * kx::exp(((-R(36560000.0)) / Tgas))))));  // O++++++ + e- -> O+++++
    k[23] = ((R(1.295e-11) / ((kx::sqrt((Tgas / R(67.39))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(67.39)))), (R(1.0) - (R(0.3556) + (R(0.6472) * kx::exp(((-R(159800.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(7563000.0)))), (R(1.0) + (R(0.3556) + (R(0.6472) * kx::exp(((-R(159800.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(4.152e-09) * kx::exp(((-R(26.89)) / Tgas))) + (R(4.656e-09) * kx::exp(((-R(202.1)) / Tgas)))) + (R(1.31e-08) * kx::exp(((-R(720.0)) / Tgas)))) + (R(1.417e-09) * kx::exp(((-R(48920.0)) / Tgas)))) + (R(0.0007968) * kx::exp(((-R(314400.0)) / Tgas)))) + (R(1.271e-05) * kx::exp(((-R(673800.0)) / Tgas))))));  // Ne+ + e- -> Ne
    k[24] = ((R(1.773e-10) / ((kx::sqrt((Tgas / R(9.924))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(9.924)))), (R(1.0) - (R(0.6434) + (R(0.2205) * kx::exp(((-R(229200.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(8878000.0)))), (R(1.0) + (R(0.6434) + (R(0.2205) * kx::exp(((-R(229200.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.98e-08) * kx::exp(((-R(45.79)) / Tgas))) + (R(1.257e-07) * kx::exp(((-R(475.3)) / Tgas)))) + (R(1.122e-06) * kx::exp(((-R(14810.0)) / Tgas)))) + (R(0.002626) * kx::exp(((-R(281000.0)) / Tgas)))) + (R(0.0008802) * kx::exp(((-R(476300.0)) / Tgas)))) + (R(1.231) * kx::exp(((-R(467700000.0)) / Tgas))))));  // Ne++ + e- -> Ne+
    k[25] = ((R(8.321e-10) / ((kx::sqrt((Tgas / R(3.332))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(3.332)))), (R(1.0) - (R(0.7254) + (R(0.0921) * kx::exp(((-R(304400.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(8696000.0)))), (R(1.0) + (R(0.7254) + (R(0.0921) * kx::exp(((-R(304400.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.763e-06) * kx::exp(((-R(639.3)) / Tgas))) + (R(1.053e-05) * kx::exp(((-R(1499.0)) / Tgas)))) + (R(4.453e-05) * kx::exp(((-R(32270.0)) / Tgas)))) + (R(0.006244) * kx::exp(((-R(256100.0)) / Tgas)))) + (R(0.0003146) * kx::exp(((-R(450500.0)) / Tgas)))) + (R(0.004465) * kx::exp(((-R(793400.0)) / Tgas))))));  // Ne+++ + e- -> Ne++
    k[26] = ((R(1.861e-09) / ((kx::sqrt((Tgas / R(2.504))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(2.504)))), (R(1.0) - (R(0.7593) + (R(0.0406) * kx::exp(((-R(325500.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(8037000.0)))), (R(1.0) + (R(0.7593) + (R(0.0406) * kx::exp(((-R(325500.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.922e-06) * kx::exp(((-R(205.0)) / Tgas))) + (R(7.144e-06) * kx::exp(((-R(2205.0)) / Tgas)))) + (R(2.836e-05) * kx::exp(((-R(9271.0)) / Tgas)))) + (R(9.82e-05) * kx::exp(((-R(49880.0)) / Tgas)))) + (R(0.008379) * kx::exp(((-R(290400.0)) / Tgas)))) + (R(0.01009) * kx::exp(((-R(878200.0)) / Tgas))))));  // Ne++++ + e- -> Ne+++
    k[27] = ((R(1.127e-09) / ((kx::sqrt((Tgas / R(13.11))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(13.11)))), (R(1.0) - (R(0.7556) + (R(0.025) * kx::exp(((-R(277100.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(8047000.0)))), (R(1.0) + (R(0.7556) + (R(0.025) * kx::exp(((-R(277100.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(5.653e-06) * kx::exp(((-R(628.0)) / Tgas))) + (R(4.344e-05) * kx::exp(((-R(2812.0)) / Tgas)))) + (R(0.0001086) * kx::exp(((-R(13240.0)) / Tgas)))) + (R(0.000598) * kx::exp(((-R(80640.0)) / Tgas)))) + (R(0.01457) * kx::exp(((-R(305200.0)) / Tgas)))) + (R(0.01601) * kx::exp(((-R(1032000.0)) / Tgas)))) + (R(0.0005365) * kx::exp(((-R(2388000.0)) / Tgas))))));  // Ne+++++ + e- -> Ne++++
    k[28] = ((R(2.557e-09) / ((kx::sqrt((Tgas / R(6.293))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(6.293)))), (R(1.0) - R(0.7601)))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(8091000.0)))), (R(1.0) + R(0.7601))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(8.46e-05) * kx::exp(((-R(1049.0)) / Tgas))) + (R(0.0001817) * kx::exp(((-R(3829.0)) / Tgas)))) + (R(0.001176) * kx::exp(((-R(61330.0)) / Tgas)))) + (R(0.01397) * kx::exp(((-R(256800.0)) / Tgas)))) + (R(0.005566) * kx::exp(((-R(460000.0)) / Tgas)))) + (R(0.006229) * kx::exp(((-R(1324000.0)) / Tgas)))) + (R(0.01031) * kx::exp(((-R(9353000.0)) / Tgas))))));  // Ne++++++ + e- -> Ne+++++
    k[29] = ((R(5.452e-11) / ((kx::sqrt((Tgas / R(5.637))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(5.637)))), (R(1.0) - (R(0.6845) + (R(0.3945) * kx::exp(((-R(836000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1551000.0)))), (R(1.0) + (R(0.6845) + (R(0.3945) * kx::exp(((-R(836000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((R(3.871e-08) * kx::exp(((-R(8415.0)) / Tgas))) + (R(4.732e-07) * kx::exp(((-R(16820.0)) / Tgas)))) + (R(0.001599) * kx::exp(((-R(50000.0)) / Tgas)))) + (R(2.628e-05) * kx::exp(((-R(275900.0)) / Tgas))))));  // Mg+ + e- -> Mg
    k[30] = ((R(1.345e-11) / ((kx::sqrt((Tgas / R(787.7))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(787.7)))), (R(1.0) - (R(0.1074) + (R(0.4631) * kx::exp(((-R(502700.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(79250000.0)))), (R(1.0) + (R(0.1074) + (R(0.4631) * kx::exp(((-R(502700.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(6.269e-06) * kx::exp(((-R(410400.0)) / Tgas))) + (R(0.0009181) * kx::exp(((-R(576600.0)) / Tgas)))) + (R(0.0003082) * kx::exp(((-R(731000.0)) / Tgas))))));  // Mg++ + e- -> Mg+
    k[31] = ((R(1.249e-10) / ((kx::sqrt((Tgas / R(77.48))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(77.48)))), (R(1.0) - (R(0.56) + (R(0.1917) * kx::exp(((-R(513900.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(20150000.0)))), (R(1.0) + (R(0.56) + (R(0.1917) * kx::exp(((-R(513900.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(9.4e-08) * kx::exp(((-R(266.0)) / Tgas))) + (R(3.818e-07) * kx::exp(((-R(1141.0)) / Tgas)))) + (R(2.064e-07) * kx::exp(((-R(3210.0)) / Tgas)))) + (R(2.71e-05) * kx::exp(((-R(172700.0)) / Tgas)))) + (R(0.003802) * kx::exp(((-R(452300.0)) / Tgas)))) + (R(0.003086) * kx::exp(((-R(910500.0)) / Tgas))))));  // Mg+++ + e- -> Mg++
    k[32] = ((R(4.031e-10) / ((kx::sqrt((Tgas / R(32.05))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(32.05)))), (R(1.0) - (R(0.6803) + (R(0.0764) * kx::exp(((-R(539900.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(16260000.0)))), (R(1.0) + (R(0.6803) + (R(0.0764) * kx::exp(((-R(539900.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(2.574e-06) * kx::exp(((-R(112.9)) / Tgas))) + (R(2.001e-06) * kx::exp(((-R(1585.0)) / Tgas)))) + (R(1.829e-05) * kx::exp(((-R(11930.0)) / Tgas)))) + (R(2.362e-05) * kx::exp(((-R(41510.0)) / Tgas)))) + (R(0.006413) * kx::exp(((-R(368400.0)) / Tgas)))) + (R(0.004118) * kx::exp(((-R(692500.0)) / Tgas)))) + (R(0.007224) * kx::exp(((-R(1150000.0)) / Tgas))))));  // Mg++++ + e- -> Mg+++
    k[33] = ((R(7.515e-10) / ((kx::sqrt((Tgas / R(25.82))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(25.82)))), (R(1.0) - (R(0.7203) + (R(0.0436) * kx::exp(((-R(569100.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(13550000.0)))), (R(1.0) + (R(0.7203) + (R(0.0436) * kx::exp(((-R(569100.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.407e-05) * kx::exp(((-R(8061.0)) / Tgas))) + (R(0.0001187) * kx::exp(((-R(21540.0)) / Tgas)))) + (R(0.0003845) * kx::exp(((-R(66400.0)) / Tgas)))) + (R(0.01333) * kx::exp(((-R(340600.0)) / Tgas)))) + (R(0.02804) * kx::exp(((-R(1327000.0)) / Tgas)))) + (R(0.0008746) * kx::exp(((-R(3347000.0)) / Tgas))))));  // Mg+++++ + e- -> Mg++++
    k[34] = ((R(9.133e-10) / ((kx::sqrt((Tgas / R(36.74))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(36.74)))), (R(1.0) - (R(0.7353) + (R(0.0211) * kx::exp(((-R(404900.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(12630000.0)))), (R(1.0) + (R(0.7353) + (R(0.0211) * kx::exp(((-R(404900.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(2.931e-05) * kx::exp(((-R(766.7)) / Tgas))) + (R(5.192e-05) * kx::exp(((-R(3444.0)) / Tgas)))) + (R(0.0001769) * kx::exp(((-R(24090.0)) / Tgas)))) + (R(0.0007988) * kx::exp(((-R(119000.0)) / Tgas)))) + (R(0.01488) * kx::exp(((-R(399900.0)) / Tgas)))) + (R(0.04546) * kx::exp(((-R(1502000.0)) / Tgas)))) + (R(0.0001505) * kx::exp(((-R(24640000.0)) / Tgas))))));  // Mg++++++ + e- -> Mg+++++
    k[35] = ((R(3.262e-11) / ((kx::sqrt((Tgas / R(15.9))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(15.9)))), (R(1.0) - (R(0.627) + (R(0.2333) * kx::exp(((-R(58280.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(42370000.0)))), (R(1.0) + (R(0.627) + (R(0.2333) * kx::exp(((-R(58280.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(3.408e-08) * kx::exp(((-R(24.31)) / Tgas))) + (R(1.913e-07) * kx::exp(((-R(129.3)) / Tgas)))) + (R(1.679e-07) * kx::exp(((-R(427.2)) / Tgas)))) + (R(7.523e-07) * kx::exp(((-R(3729.0)) / Tgas)))) + (R(8.386e-05) * kx::exp(((-R(55140.0)) / Tgas)))) + (R(0.004083) * kx::exp(((-R(129500.0)) / Tgas))))));  // Si+ + e- -> Si
    k[36] = ((R(1.964e-10) / ((kx::sqrt((Tgas / R(7.712))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(7.712)))), (R(1.0) - (R(0.6287) + (R(0.1523) * kx::exp(((-R(480400.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(29510000.0)))), (R(1.0) + (R(0.6287) + (R(0.1523) * kx::exp(((-R(480400.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(2.93e-06) * kx::exp(((-R(116.2)) / Tgas))) + (R(2.803e-06) * kx::exp(((-R(5721.0)) / Tgas)))) + (R(9.023e-05) * kx::exp(((-R(34770.0)) / Tgas)))) + (R(0.006909) * kx::exp(((-R(117600.0)) / Tgas)))) + (R(2.582e-05) * kx::exp(((-R(3505000.0)) / Tgas))))));  // Si++ + e- -> Si+
    k[37] = ((R(6.739e-11) / ((kx::sqrt((Tgas / R(216.6))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(216.6)))), (R(1.0) - (R(0.4931) + (R(0.1667) * kx::exp(((-R(904600.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(44910000.0)))), (R(1.0) + (R(0.4931) + (R(0.1667) * kx::exp(((-R(904600.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(3.819e-06) * kx::exp(((-R(3802.0)) / Tgas))) + (R(2.421e-05) * kx::exp(((-R(12800.0)) / Tgas)))) + (R(0.0002283) * kx::exp(((-R(59530.0)) / Tgas)))) + (R(0.008604) * kx::exp(((-R(102600.0)) / Tgas)))) + (R(0.002617) * kx::exp(((-R(1154000.0)) / Tgas))))));  // Si+++ + e- -> Si++
    k[38] = ((R(5.134e-11) / ((kx::sqrt((Tgas / R(1009.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1009.0)))), (R(1.0) - (R(0.3678) + (R(0.1646) * kx::exp(((-R(1084000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(85140000.0)))), (R(1.0) + (R(0.3678) + (R(0.1646) * kx::exp(((-R(1084000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(0.0001422) * kx::exp(((-R(768500.0)) / Tgas))) + (R(0.009474) * kx::exp(((-R(1208000.0)) / Tgas)))) + (R(0.00165) * kx::exp(((-R(1839000.0)) / Tgas))))));  // Si++++ + e- -> Si+++
    k[39] = ((R(2.468e-10) / ((kx::sqrt((Tgas / R(164.9))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(164.9)))), (R(1.0) - (R(0.6113) + (R(0.0636) * kx::exp(((-R(983700.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(32310000.0)))), (R(1.0) + (R(0.6113) + (R(0.0636) * kx::exp(((-R(983700.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(7.163e-07) * kx::exp(((-R(562.5)) / Tgas))) + (R(2.656e-06) * kx::exp(((-R(2952.0)) / Tgas)))) + (R(1.119e-06) * kx::exp(((-R(9682.0)) / Tgas)))) + (R(4.796e-05) * kx::exp(((-R(147300.0)) / Tgas)))) + (R(0.004052) * kx::exp(((-R(506400.0)) / Tgas)))) + (R(0.006101) * kx::exp(((-R(804700.0)) / Tgas)))) + (R(0.02366) * kx::exp(((-R(1623000.0)) / Tgas))))));  // Si+++++ + e- -> Si++++
    k[40] = ((R(4.615e-10) / ((kx::sqrt((Tgas / R(114.3))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(114.3)))), (R(1.0) - (R(0.6753) + (R(0.0356) * kx::exp(((-R(859500.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(23770000.0)))), (R(1.0) + (R(0.6753) + (R(0.0356) * kx::exp(((-R(859500.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(2.086e-06) * kx::exp(((-R(370.8)) / Tgas))) + (R(9.423e-06) * kx::exp(((-R(3870.0)) / Tgas)))) + (R(3.423e-05) * kx::exp(((-R(22260.0)) / Tgas)))) + (R(0.000395) * kx::exp(((-R(131800.0)) / Tgas)))) + (R(0.01535) * kx::exp(((-R(528500.0)) / Tgas)))) + (R(0.04986) * kx::exp(((-R(1742000.0)) / Tgas)))) + (R(0.0004067) * kx::exp(((-R(8392000.0)) / Tgas))))));  // Si++++++ + e- -> Si+++++
    k[41] = ((R(1.384e-10) / ((kx::sqrt((Tgas / R(1.074))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1.074)))), (R(1.0) - (R(0.6886) + (R(0.1845) * kx::exp(((-R(18580.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(715900.0)))), (R(1.0) + (R(0.6886) + (R(0.1845) * kx::exp(((-R(18580.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(7.3e-08) * kx::exp(((-R(507.7)) / Tgas))) + (R(2.577e-07) * kx::exp(((-R(600.7)) / Tgas)))) + (R(4.961e-08) * kx::exp(((-R(2342.0)) / Tgas)))) + (R(9.52e-07) * kx::exp(((-R(7269.0)) / Tgas)))) + (R(9.586e-07) * kx::exp(((-R(21900.0)) / Tgas)))) + (R(0.0006849) * kx::exp(((-R(148300.0)) / Tgas)))) + (R(0.0006539) * kx::exp(((-R(190600.0)) / Tgas))))));  // S+ + e- -> S
    k[42] = ((R(2.478e-11) / ((kx::sqrt((Tgas / R(329.4))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(329.4)))), (R(1.0) - (R(0.4642) + (R(0.3351) * kx::exp(((-R(763000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(21660000.0)))), (R(1.0) + (R(0.4642) + (R(0.3351) * kx::exp(((-R(763000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((((R(3.04e-07) * kx::exp(((-R(50.16)) / Tgas))) + (R(4.393e-07) * kx::exp(((-R(326.6)) / Tgas)))) + (R(1.609e-06) * kx::exp(((-R(3102.0)) / Tgas)))) + (R(4.98e-06) * kx::exp(((-R(12100.0)) / Tgas)))) + (R(3.457e-05) * kx::exp(((-R(49690.0)) / Tgas)))) + (R(0.008617) * kx::exp(((-R(201000.0)) / Tgas)))) + (R(0.0009284) * kx::exp(((-R(257500.0)) / Tgas))))));  // S++ + e- -> S+
    k[43] = ((R(3.043e-10) / ((kx::sqrt((Tgas / R(16.78))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(16.78)))), (R(1.0) - (R(0.6947) + (R(0.0795) * kx::exp(((-R(68680.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(20500000.0)))), (R(1.0) + (R(0.6947) + (R(0.0795) * kx::exp(((-R(68680.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(5.817e-07) * kx::exp(((-R(362.8)) / Tgas))) + (R(1.391e-06) * kx::exp(((-R(1058.0)) / Tgas)))) + (R(1.123e-05) * kx::exp(((-R(7160.0)) / Tgas)))) + (R(0.0001521) * kx::exp(((-R(32600.0)) / Tgas)))) + (R(0.001875) * kx::exp(((-R(123500.0)) / Tgas)))) + (R(0.02097) * kx::exp(((-R(207000.0)) / Tgas))))));  // S+++ + e- -> S++
    k[44] = ((R(2.615e-10) / ((kx::sqrt((Tgas / R(62.38))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(62.38)))), (R(1.0) - (R(0.6343) + (R(0.0773) * kx::exp(((-R(1059000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(28030000.0)))), (R(1.0) + (R(0.6343) + (R(0.0773) * kx::exp(((-R(1059000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((((R(9.571e-06) * kx::exp(((-R(1180.0)) / Tgas))) + (R(6.268e-05) * kx::exp(((-R(6443.0)) / Tgas)))) + (R(0.0003807) * kx::exp(((-R(22640.0)) / Tgas)))) + (R(0.01874) * kx::exp(((-R(153000.0)) / Tgas)))) + (R(0.005526) * kx::exp(((-R(356400.0)) / Tgas))))));  // S++++ + e- -> S+++
    k[45] = ((R(1.588e-10) / ((kx::sqrt((Tgas / R(335.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(335.0)))), (R(1.0) - (R(0.5584) + (R(0.0591) * kx::exp(((-R(1656000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(51880000.0)))), (R(1.0) + (R(0.5584) + (R(0.0591) * kx::exp(((-R(1656000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * ((((((R(2.816e-06) * kx::exp(((-R(7590.0)) / Tgas))) + (R(3.172e-05) * kx::exp(((-R(15580.0)) / Tgas)))) + (R(0.0001832) * kx::exp(((-R(40130.0)) / Tgas)))) + (R(0.00436) * kx::exp(((-R(115600.0)) / Tgas)))) + (R(0.01618) * kx::exp(((-R(160100.0)) / Tgas)))) + (R(0.007707) * kx::exp(((-R(1839000.0)) / Tgas))))));  // S+++++ + e- -> S++++
    k[46] = ((R(9.565e-11) / ((kx::sqrt((Tgas / R(1599.0))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(1599.0)))), (R(1.0) - (R(0.4517) + (R(0.0612) * kx::exp(((-R(1986000.0)) / Tgas))))))) * kx::pow((R(1.0) + kx::sqrt((Tgas / R(92520000.0)))), (R(1.0) + (R(0.4517) + (R(0.0612) * kx::exp(((-R(1986000.0)) / Tgas)))))))) + (kx::pow(Tgas, (-R(1.5))) * (((R(3.931e-05) * kx::exp(((-R(945500.0)) / Tgas))) + (R(0.004431) * kx::exp(((-R(1365000.0)) / Tgas)))) + (R(0.05156) * kx::exp(((-R(2169000.0)) / Tgas))))));  // S++++++ + e- -> S+++++
    k[47] = kx::exp((((((((((-R(32.71396786)) + (R(13.536556) * kx::log((Tgas * R(8.617333262145179e-05))))) - (R(5.73932875) * kx::ipow<2>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(1.56315498) * kx::ipow<3>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.2877056) * kx::ipow<4>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.0348255977) * kx::ipow<5>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.00263197617) * kx::ipow<6>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.000111954395) * kx::ipow<7>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(2.03914985e-06) * kx::ipow<8>(kx::log((Tgas * R(8.617333262145179e-05)))))));  // H + e- -> H+ + e- + e-
    k[48] = kx::exp((((((((((-R(44.09864886)) + (R(23.91596563) * kx::log((Tgas * R(8.617333262145179e-05))))) - (R(10.7532302) * kx::ipow<2>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(3.05803875) * kx::ipow<3>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.56851189) * kx::ipow<4>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.0679539123) * kx::ipow<5>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.0050090561) * kx::ipow<6>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.000206723616) * kx::ipow<7>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(3.64916141e-06) * kx::ipow<8>(kx::log((Tgas * R(8.617333262145179e-05)))))));  // He + e- -> He+ + e- + e-
    k[49] = kx::exp((((((((((-R(68.71040990212)) + (R(43.93347632635) * kx::log((Tgas * R(8.617333262145179e-05))))) - (R(18.48066993568) * kx::ipow<2>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(4.701626486759002) * kx::ipow<3>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.7692466334492) * kx::ipow<4>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.08113042097303) * kx::ipow<5>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(0.005324020628287001) * kx::ipow<6>(kx::log((Tgas * R(8.617333262145179e-05)))))) + (R(0.0001975705312221) * kx::ipow<7>(kx::log((Tgas * R(8.617333262145179e-05)))))) - (R(3.165581065665e-06) * kx::ipow<8>(kx::log((Tgas * R(8.617333262145179e-05)))))));  // He+ + e- -> He++ + e- + e-
    {  // C + e- -> C+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[50] = kx::colion_v96(Tgas, R(11.3), R(0.0), R(6.85e-08), R(0.193), R(0.25));  // C + e- -> C+ + e- + e-
    }
    {  // C+ + e- -> C++ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[51] = kx::colion_v96(Tgas, R(24.4), R(1.0), R(1.86e-08), R(0.286), R(0.24));  // C+ + e- -> C++ + e- + e-
    }
    {  // C++ + e- -> C+++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[52] = kx::colion_v96(Tgas, R(47.9), R(1.0), R(6.35e-09), R(0.427), R(0.21));  // C++ + e- -> C+++ + e- + e-
    }
    {  // C+++ + e- -> C++++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[53] = kx::colion_v96(Tgas, R(64.5), R(1.0), R(1.5e-09), R(0.416), R(0.13));  // C+++ + e- -> C++++ + e- + e-
    }
    {  // C++++ + e- -> C+++++ + e- + e-: 1 piece(s), zero below 232090.4, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(232090.4)) k[54] = kx::colion_v96(Tgas, R(392.1), R(1.0), R(2.99e-10), R(0.666), R(0.02));  // C++++ + e- -> C+++++ + e- + e-
    }
    {  // C+++++ + e- -> C++++++ + e- + e-: 1 piece(s), zero below 232090.4, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(232090.4)) k[55] = kx::colion_v96(Tgas, R(490.0), R(1.0), R(1.23e-10), R(0.62), R(0.16));  // C+++++ + e- -> C++++++ + e- + e-
    }
    {  // N + e- -> N+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[56] = kx::colion_v96(Tgas, R(14.5), R(0.0), R(4.82e-08), R(0.0652), R(0.42));  // N + e- -> N+ + e- + e-
    }
    {  // N+ + e- -> N++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[57] = kx::colion_v96(Tgas, R(29.6), R(0.0), R(2.98e-08), R(0.31), R(0.3));  // N+ + e- -> N++ + e- + e-
    }
    {  // N++ + e- -> N+++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[58] = kx::colion_v96(Tgas, R(47.5), R(1.0), R(8.1e-09), R(0.35), R(0.24));  // N++ + e- -> N+++ + e- + e-
    }
    {  // N+++ + e- -> N++++ + e- + e-: 1 piece(s), zero below 46418.08, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(46418.08)) k[59] = kx::colion_v96(Tgas, R(77.5), R(1.0), R(3.71e-09), R(0.549), R(0.18));  // N+++ + e- -> N++++ + e- + e-
    }
    {  // N++++ + e- -> N+++++ + e- + e-: 1 piece(s), zero below 46418.08, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(46418.08)) k[60] = kx::colion_v96(Tgas, R(97.9), R(0.0), R(1.51e-09), R(0.0167), R(0.74));  // N++++ + e- -> N+++++ + e- + e-
    }
    {  // N+++++ + e- -> N++++++ + e- + e-: 1 piece(s), zero below 348135.6, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(348135.6)) k[61] = kx::colion_v96(Tgas, R(552.1), R(0.0), R(3.71e-10), R(0.546), R(0.29));  // N+++++ + e- -> N++++++ + e- + e-
    }
    {  // O + e- -> O+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[62] = kx::colion_v96(Tgas, R(13.6), R(0.0), R(3.59e-08), R(0.073), R(0.34));  // O + e- -> O+ + e- + e-
    }
    {  // O+ + e- -> O++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[63] = kx::colion_v96(Tgas, R(35.1), R(1.0), R(1.39e-08), R(0.212), R(0.22));  // O+ + e- -> O++ + e- + e-
    }
    {  // O++ + e- -> O+++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[64] = kx::colion_v96(Tgas, R(54.9), R(1.0), R(9.31e-09), R(0.27), R(0.27));  // O++ + e- -> O+++ + e- + e-
    }
    {  // O+++ + e- -> O++++ + e- + e-: 1 piece(s), zero below 46418.08, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(46418.08)) k[65] = kx::colion_v96(Tgas, R(77.4), R(0.0), R(1.02e-08), R(0.614), R(0.27));  // O+++ + e- -> O++++ + e- + e-
    }
    {  // O++++ + e- -> O+++++ + e- + e-: 1 piece(s), zero below 58022.6, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(58022.6)) k[66] = kx::colion_v96(Tgas, R(113.9), R(1.0), R(2.19e-09), R(0.63), R(0.17));  // O++++ + e- -> O+++++ + e- + e-
    }
    {  // O+++++ + e- -> O++++++ + e- + e-: 1 piece(s), zero below 81231.63, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(81231.63)) k[67] = kx::colion_v96(Tgas, R(138.1), R(0.0), R(1.95e-09), R(0.36), R(0.54));  // O+++++ + e- -> O++++++ + e- + e-
    }
    {  // Ne + e- -> Ne+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[68] = kx::colion_v96(Tgas, R(21.6), R(1.0), R(1.5e-08), R(0.0329), R(0.43));  // Ne + e- -> Ne+ + e- + e-
    }
    {  // Ne+ + e- -> Ne++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[69] = kx::colion_v96(Tgas, R(41.0), R(0.0), R(1.98e-08), R(0.295), R(0.2));  // Ne+ + e- -> Ne++ + e- + e-
    }
    {  // Ne++ + e- -> Ne+++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[70] = kx::colion_v96(Tgas, R(63.5), R(1.0), R(7.03e-09), R(0.0677), R(0.39));  // Ne++ + e- -> Ne+++ + e- + e-
    }
    {  // Ne+++ + e- -> Ne++++ + e- + e-: 1 piece(s), zero below 58022.6, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(58022.6)) k[71] = kx::colion_v96(Tgas, R(97.1), R(1.0), R(4.24e-09), R(0.0482), R(0.58));  // Ne+++ + e- -> Ne++++ + e- + e-
    }
    {  // Ne++++ + e- -> Ne+++++ + e- + e-: 1 piece(s), zero below 81231.63, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(81231.63)) k[72] = kx::colion_v96(Tgas, R(126.2), R(1.0), R(2.79e-09), R(0.305), R(0.25));  // Ne++++ + e- -> Ne+++++ + e- + e-
    }
    {  // Ne+++++ + e- -> Ne++++++ + e- + e-: 1 piece(s), zero below 81231.63, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(81231.63)) k[73] = kx::colion_v96(Tgas, R(157.9), R(0.0), R(3.45e-09), R(0.581), R(0.28));  // Ne+++++ + e- -> Ne++++++ + e- + e-
    }
    {  // Mg + e- -> Mg+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[74] = kx::colion_v96(Tgas, R(7.6), R(0.0), R(6.21e-07), R(0.592), R(0.39));  // Mg + e- -> Mg+ + e- + e-
    }
    {  // Mg+ + e- -> Mg++ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[75] = kx::colion_v96(Tgas, R(15.2), R(0.0), R(1.92e-08), R(0.0027), R(0.85));  // Mg+ + e- -> Mg++ + e- + e-
    }
    {  // Mg++ + e- -> Mg+++ + e- + e-: 1 piece(s), zero below 46418.08, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(46418.08)) k[76] = kx::colion_v96(Tgas, R(80.1), R(1.0), R(5.56e-09), R(0.107), R(0.3));  // Mg++ + e- -> Mg+++ + e- + e-
    }
    {  // Mg+++ + e- -> Mg++++ + e- + e-: 1 piece(s), zero below 58022.6, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(58022.6)) k[77] = kx::colion_v96(Tgas, R(109.3), R(1.0), R(4.35e-09), R(0.159), R(0.31));  // Mg+++ + e- -> Mg++++ + e- + e-
    }
    {  // Mg++++ + e- -> Mg+++++ + e- + e-: 1 piece(s), zero below 81231.63, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(81231.63)) k[78] = kx::colion_v96(Tgas, R(141.3), R(0.0), R(7.1e-09), R(0.658), R(0.25));  // Mg++++ + e- -> Mg+++++ + e- + e-
    }
    {  // Mg+++++ + e- -> Mg++++++ + e- + e-: 1 piece(s), zero below 116045.2, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(116045.2)) k[79] = kx::colion_v96(Tgas, R(186.5), R(1.0), R(1.7e-09), R(0.242), R(0.28));  // Mg+++++ + e- -> Mg++++++ + e- + e-
    }
    {  // Si + e- -> Si+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[80] = kx::colion_v96(Tgas, R(8.2), R(1.0), R(1.88e-07), R(0.376), R(0.25));  // Si + e- -> Si+ + e- + e-
    }
    {  // Si+ + e- -> Si++ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[81] = kx::colion_v96(Tgas, R(16.4), R(1.0), R(6.43e-08), R(0.632), R(0.2));  // Si+ + e- -> Si++ + e- + e-
    }
    {  // Si++ + e- -> Si+++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[82] = kx::colion_v96(Tgas, R(33.5), R(1.0), R(2.01e-08), R(0.473), R(0.22));  // Si++ + e- -> Si+++ + e- + e-
    }
    {  // Si+++ + e- -> Si++++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[83] = kx::colion_v96(Tgas, R(54.0), R(1.0), R(4.94e-09), R(0.172), R(0.23));  // Si+++ + e- -> Si++++ + e- + e-
    }
    {  // Si++++ + e- -> Si+++++ + e- + e-: 1 piece(s), zero below 81231.63, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(81231.63)) k[84] = kx::colion_v96(Tgas, R(166.8), R(1.0), R(1.76e-09), R(0.102), R(0.31));  // Si++++ + e- -> Si+++++ + e- + e-
    }
    {  // Si+++++ + e- -> Si++++++ + e- + e-: 1 piece(s), zero below 116045.2, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(116045.2)) k[85] = kx::colion_v96(Tgas, R(205.3), R(1.0), R(1.74e-09), R(0.18), R(0.29));  // Si+++++ + e- -> Si++++++ + e- + e-
    }
    {  // S + e- -> S+ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[86] = kx::colion_v96(Tgas, R(10.4), R(1.0), R(5.49e-08), R(0.1), R(0.25));  // S + e- -> S+ + e- + e-
    }
    {  // S+ + e- -> S++ + e- + e-: 1 piece(s), zero below 11604.52, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(11604.52)) k[87] = kx::colion_v96(Tgas, R(23.3), R(1.0), R(6.81e-08), R(0.693), R(0.21));  // S+ + e- -> S++ + e- + e-
    }
    {  // S++ + e- -> S+++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[88] = kx::colion_v96(Tgas, R(34.8), R(1.0), R(2.14e-08), R(0.353), R(0.24));  // S++ + e- -> S+++ + e- + e-
    }
    {  // S+++ + e- -> S++++ + e- + e-: 1 piece(s), zero below 23209.04, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(23209.04)) k[89] = kx::colion_v96(Tgas, R(47.3), R(1.0), R(1.66e-08), R(1.03), R(0.14));  // S+++ + e- -> S++++ + e- + e-
    }
    {  // S++++ + e- -> S+++++ + e- + e-: 1 piece(s), zero below 34813.56, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(34813.56)) k[90] = kx::colion_v96(Tgas, R(72.6), R(1.0), R(6.12e-09), R(0.58), R(0.19));  // S++++ + e- -> S+++++ + e- + e-
    }
    {  // S+++++ + e- -> S++++++ + e- + e-: 1 piece(s), zero below 116045.2, clamped above 232090400.0
      R Tgas = Trate;
      if (Tgas > R(232090400.0)) Tgas = R(232090400.0);
      if (Tgas >= R(116045.2)) k[91] = kx::colion_v96(Tgas, R(88.1), R(1.0), R(1.33e-09), R(0.0688), R(0.35));  // S+++++ + e- -> S++++++ + e- + e-
    }
    k[92] = R(p.ph[0]);  // H -> H+ + e- (photo)
    k[93] = R(p.ph[1]);  // He -> He+ + e- (photo)
    k[94] = R(p.ph[2]);  // He+ -> He++ + e- (photo)
    k[95] = R(p.ph[3]);  // C -> C+ + e- (photo)
    k[96] = R(p.ph[4]);  // C+ -> C++ + e- (photo)
    k[97] = R(p.ph[5]);  // C++ -> C+++ + e- (photo)
    k[98] = R(p.ph[6]);  // C+++ -> C++++ + e- (photo)
    k[99] = R(p.ph[7]);  // C++++ -> C+++++ + e- (photo)
    k[100] = R(p.ph[8]);  // C+++++ -> C++++++ + e- (photo)
    k[101] = R(p.ph[9]);  // N -> N+ + e- (photo)
    k[102] = R(p.ph[10]);  // N+ -> N++ + e- (photo)
    k[103] = R(p.ph[11]);  // N++ -> N+++ + e- (photo)
    k[104] = R(p.ph[12]);  // N+++ -> N++++ + e- (photo)
    k[105] = R(p.ph[13]);  // N++++ -> N+++++ + e- (photo)
    k[106] = R(p.ph[14]);  // N+++++ -> N++++++ + e- (photo)
    k[107] = R(p.ph[15]);  // O -> O+ + e- (photo)
    k[108] = R(p.ph[16]);  // O+ -> O++ + e- (photo)
    k[109] = R(p.ph[17]);  // O++ -> O+++ + e- (photo)
    k[110] = R(p.ph[18]);  // O+++ -> O++++ + e- (photo)
    k[111] = R(p.ph[19]);  // O++++ -> O+++++ + e- (photo)
    k[112] = R(p.ph[20]);  // O+++++ -> O++++++ + e- (photo)
    k[113] = R(p.ph[21]);  // Ne -> Ne+ + e- (photo)
    k[114] = R(p.ph[22]);  // Ne+ -> Ne++ + e- (photo)
    k[115] = R(p.ph[23]);  // Ne++ -> Ne+++ + e- (photo)
    k[116] = R(p.ph[24]);  // Ne+++ -> Ne++++ + e- (photo)
    k[117] = R(p.ph[25]);  // Ne++++ -> Ne+++++ + e- (photo)
    k[118] = R(p.ph[26]);  // Ne+++++ -> Ne++++++ + e- (photo)
    k[119] = R(p.ph[27]);  // Mg -> Mg+ + e- (photo)
    k[120] = R(p.ph[28]);  // Mg+ -> Mg++ + e- (photo)
    k[121] = R(p.ph[29]);  // Mg++ -> Mg+++ + e- (photo)
    k[122] = R(p.ph[30]);  // Mg+++ -> Mg++++ + e- (photo)
    k[123] = R(p.ph[31]);  // Mg++++ -> Mg+++++ + e- (photo)
    k[124] = R(p.ph[32]);  // Mg+++++ -> Mg++++++ + e- (photo)
    k[125] = R(p.ph[33]);  // Si -> Si+ + e- (photo)
    k[126] = R(p.ph[34]);  // Si+ -> Si++ + e- (photo)
    k[127] = R(p.ph[35]);  // Si++ -> Si+++ + e- (photo)
    k[128] = R(p.ph[36]);  // Si+++ -> Si++++ + e- (photo)
    k[129] = R(p.ph[37]);  // Si++++ -> Si+++++ + e- (photo)
    k[130] = R(p.ph[38]);  // Si+++++ -> Si++++++ + e- (photo)
    k[131] = R(p.ph[39]);  // S -> S+ + e- (photo)
    k[132] = R(p.ph[40]);  // S+ -> S++ + e- (photo)
    k[133] = R(p.ph[41]);  // S++ -> S+++ + e- (photo)
    k[134] = R(p.ph[42]);  // S+++ -> S++++ + e- (photo)
    k[135] = R(p.ph[43]);  // S++++ -> S+++++ + e- (photo)
    k[136] = R(p.ph[44]);  // S+++++ -> S++++++ + e- (photo)
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
  static constexpr int NREA_T = 137;
  // the loop form's index tables (netgen.loop_tables)
  static constexpr bool RHS_LOOP = false;
  static constexpr int LT_KF_OFF[138] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150, 152, 154, 156, 158, 160, 162, 164, 166, 168, 170, 172, 174, 176, 178, 180, 182, 184, 185, 186, 187, 188, 189, 190, 191, 192, 193, 194, 195, 196, 197, 198, 199, 200, 201, 202, 203, 204, 205, 206, 207, 208, 209, 210, 211, 212, 213, 214, 215, 216, 217, 218, 219, 220, 221, 222, 223, 224, 225, 226, 227, 228, 229};
  static constexpr int LT_KF_SP[229] = {8, 54, 8, 54, 10, 54, 10, 54, 11, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 6, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 25, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 39, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 32, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 18, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 53, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45, 54, 46, 54, 7, 54, 9, 54, 10, 54, 0, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 19, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 33, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 26, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 12, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 47, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 40, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45, 54, 7, 9, 10, 0, 1, 2, 3, 4, 5, 19, 20, 21, 22, 23, 24, 33, 34, 35, 36, 37, 38, 26, 27, 28, 29, 30, 31, 12, 13, 14, 15, 16, 17, 47, 48, 49, 50, 51, 52, 40, 41, 42, 43, 44, 45};
  static constexpr int LT_SC_OFF[138] = {0, 3, 6, 9, 12, 15, 18, 21, 24, 27, 30, 33, 36, 39, 42, 45, 48, 51, 54, 57, 60, 63, 66, 69, 72, 75, 78, 81, 84, 87, 90, 93, 96, 99, 102, 105, 108, 111, 114, 117, 120, 123, 126, 129, 132, 135, 138, 141, 146, 151, 156, 161, 166, 171, 176, 181, 186, 191, 196, 201, 206, 211, 216, 221, 226, 231, 236, 241, 246, 251, 256, 261, 266, 271, 276, 281, 286, 291, 296, 301, 306, 311, 316, 321, 326, 331, 336, 341, 346, 351, 356, 361, 366, 369, 372, 375, 378, 381, 384, 387, 390, 393, 396, 399, 402, 405, 408, 411, 414, 417, 420, 423, 426, 429, 432, 435, 438, 441, 444, 447, 450, 453, 456, 459, 462, 465, 468, 471, 474, 477, 480, 483, 486, 489, 492, 495, 498, 501};
  static constexpr int LT_SC_ROW[501] = {8, 54, 7, 8, 54, 7, 10, 54, 9, 10, 54, 9, 11, 54, 10, 1, 54, 0, 2, 54, 1, 3, 54, 2, 4, 54, 3, 5, 54, 4, 6, 54, 5, 20, 54, 19, 21, 54, 20, 22, 54, 21, 23, 54, 22, 24, 54, 23, 25, 54, 24, 34, 54, 33, 35, 54, 34, 36, 54, 35, 37, 54, 36, 38, 54, 37, 39, 54, 38, 27, 54, 26, 28, 54, 27, 29, 54, 28, 30, 54, 29, 31, 54, 30, 32, 54, 31, 13, 54, 12, 14, 54, 13, 15, 54, 14, 16, 54, 15, 17, 54, 16, 18, 54, 17, 48, 54, 47, 49, 54, 48, 50, 54, 49, 51, 54, 50, 52, 54, 51, 53, 54, 52, 41, 54, 40, 42, 54, 41, 43, 54, 42, 44, 54, 43, 45, 54, 44, 46, 54, 45, 7, 54, 8, 54, 54, 9, 54, 10, 54, 54, 10, 54, 11, 54, 54, 0, 54, 1, 54, 54, 1, 54, 2, 54, 54, 2, 54, 3, 54, 54, 3, 54, 4, 54, 54, 4, 54, 5, 54, 54, 5, 54, 6, 54, 54, 19, 54, 20, 54, 54, 20, 54, 21, 54, 54, 21, 54, 22, 54, 54, 22, 54, 23, 54, 54, 23, 54, 24, 54, 54, 24, 54, 25, 54, 54, 33, 54, 34, 54, 54, 34, 54, 35, 54, 54, 35, 54, 36, 54, 54, 36, 54, 37, 54, 54, 37, 54, 38, 54, 54, 38, 54, 39, 54, 54, 26, 54, 27, 54, 54, 27, 54, 28, 54, 54, 28, 54, 29, 54, 54, 29, 54, 30, 54, 54, 30, 54, 31, 54, 54, 31, 54, 32, 54, 54, 12, 54, 13, 54, 54, 13, 54, 14, 54, 54, 14, 54, 15, 54, 54, 15, 54, 16, 54, 54, 16, 54, 17, 54, 54, 17, 54, 18, 54, 54, 47, 54, 48, 54, 54, 48, 54, 49, 54, 54, 49, 54, 50, 54, 54, 50, 54, 51, 54, 54, 51, 54, 52, 54, 54, 52, 54, 53, 54, 54, 40, 54, 41, 54, 54, 41, 54, 42, 54, 54, 42, 54, 43, 54, 54, 43, 54, 44, 54, 54, 44, 54, 45, 54, 54, 45, 54, 46, 54, 54, 7, 8, 54, 9, 10, 54, 10, 11, 54, 0, 1, 54, 1, 2, 54, 2, 3, 54, 3, 4, 54, 4, 5, 54, 5, 6, 54, 19, 20, 54, 20, 21, 54, 21, 22, 54, 22, 23, 54, 23, 24, 54, 24, 25, 54, 33, 34, 54, 34, 35, 54, 35, 36, 54, 36, 37, 54, 37, 38, 54, 38, 39, 54, 26, 27, 54, 27, 28, 54, 28, 29, 54, 29, 30, 54, 30, 31, 54, 31, 32, 54, 12, 13, 54, 13, 14, 54, 14, 15, 54, 15, 16, 54, 16, 17, 54, 17, 18, 54, 47, 48, 54, 48, 49, 54, 49, 50, 54, 50, 51, 54, 51, 52, 54, 52, 53, 54, 40, 41, 54, 41, 42, 54, 42, 43, 54, 43, 44, 54, 44, 45, 54, 45, 46, 54};
  static constexpr int LT_SC_SGN[501] = {-1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1, -1, 1, 1};
  static constexpr int LT_JT_OFF[138] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150, 152, 154, 156, 158, 160, 162, 164, 166, 168, 170, 172, 174, 176, 178, 180, 182, 184, 185, 186, 187, 188, 189, 190, 191, 192, 193, 194, 195, 196, 197, 198, 199, 200, 201, 202, 203, 204, 205, 206, 207, 208, 209, 210, 211, 212, 213, 214, 215, 216, 217, 218, 219, 220, 221, 222, 223, 224, 225, 226, 227, 228, 229};
  static constexpr int LT_JT_COL[229] = {8, 54, 8, 54, 10, 54, 10, 54, 11, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 6, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 25, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 39, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 32, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 18, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 53, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45, 54, 46, 54, 7, 54, 9, 54, 10, 54, 0, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 19, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 33, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 26, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 12, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 47, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 40, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45, 54, 7, 9, 10, 0, 1, 2, 3, 4, 5, 19, 20, 21, 22, 23, 24, 33, 34, 35, 36, 37, 38, 26, 27, 28, 29, 30, 31, 12, 13, 14, 15, 16, 17, 47, 48, 49, 50, 51, 52, 40, 41, 42, 43, 44, 45};
  static constexpr int LT_JT_MULT[229] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int LT_JT_FOFF[230] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72, 73, 74, 75, 76, 77, 78, 79, 80, 81, 82, 83, 84, 85, 86, 87, 88, 89, 90, 91, 92, 93, 94, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104, 105, 106, 107, 108, 109, 110, 111, 112, 113, 114, 115, 116, 117, 118, 119, 120, 121, 122, 123, 124, 125, 126, 127, 128, 129, 130, 131, 132, 133, 134, 135, 136, 137, 138, 139, 140, 141, 142, 143, 144, 145, 146, 147, 148, 149, 150, 151, 152, 153, 154, 155, 156, 157, 158, 159, 160, 161, 162, 163, 164, 165, 166, 167, 168, 169, 170, 171, 172, 173, 174, 175, 176, 177, 178, 179, 180, 181, 182, 183, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184, 184};
  static constexpr int LT_JF_SP[184] = {54, 8, 54, 8, 54, 10, 54, 10, 54, 11, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 6, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 25, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 39, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 32, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 18, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 53, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45, 54, 46, 54, 7, 54, 9, 54, 10, 54, 0, 54, 1, 54, 2, 54, 3, 54, 4, 54, 5, 54, 19, 54, 20, 54, 21, 54, 22, 54, 23, 54, 24, 54, 33, 54, 34, 54, 35, 54, 36, 54, 37, 54, 38, 54, 26, 54, 27, 54, 28, 54, 29, 54, 30, 54, 31, 54, 12, 54, 13, 54, 14, 54, 15, 54, 16, 54, 17, 54, 47, 54, 48, 54, 49, 54, 50, 54, 51, 54, 52, 54, 40, 54, 41, 54, 42, 54, 43, 54, 44, 54, 45};
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
    fl = k[0]*n[8]*n[54];  // H+ + e- -> H
    dn[8] -= fl;
    dn[54] -= fl;
    dn[7] += fl;
    fl = k[1]*n[8]*n[54];  // H+ + e- -> H
    dn[8] -= fl;
    dn[54] -= fl;
    dn[7] += fl;
    fl = k[2]*n[10]*n[54];  // He+ + e- -> He
    dn[10] -= fl;
    dn[54] -= fl;
    dn[9] += fl;
    fl = k[3]*n[10]*n[54];  // He+ + e- -> He
    dn[10] -= fl;
    dn[54] -= fl;
    dn[9] += fl;
    fl = k[4]*n[11]*n[54];  // He++ + e- -> He+
    dn[11] -= fl;
    dn[54] -= fl;
    dn[10] += fl;
    fl = k[5]*n[1]*n[54];  // C+ + e- -> C
    dn[1] -= fl;
    dn[54] -= fl;
    dn[0] += fl;
    fl = k[6]*n[2]*n[54];  // C++ + e- -> C+
    dn[2] -= fl;
    dn[54] -= fl;
    dn[1] += fl;
    fl = k[7]*n[3]*n[54];  // C+++ + e- -> C++
    dn[3] -= fl;
    dn[54] -= fl;
    dn[2] += fl;
    fl = k[8]*n[4]*n[54];  // C++++ + e- -> C+++
    dn[4] -= fl;
    dn[54] -= fl;
    dn[3] += fl;
    fl = k[9]*n[5]*n[54];  // C+++++ + e- -> C++++
    dn[5] -= fl;
    dn[54] -= fl;
    dn[4] += fl;
    fl = k[10]*n[6]*n[54];  // C++++++ + e- -> C+++++
    dn[6] -= fl;
    dn[54] -= fl;
    dn[5] += fl;
    fl = k[11]*n[20]*n[54];  // N+ + e- -> N
    dn[20] -= fl;
    dn[54] -= fl;
    dn[19] += fl;
    fl = k[12]*n[21]*n[54];  // N++ + e- -> N+
    dn[21] -= fl;
    dn[54] -= fl;
    dn[20] += fl;
    fl = k[13]*n[22]*n[54];  // N+++ + e- -> N++
    dn[22] -= fl;
    dn[54] -= fl;
    dn[21] += fl;
    fl = k[14]*n[23]*n[54];  // N++++ + e- -> N+++
    dn[23] -= fl;
    dn[54] -= fl;
    dn[22] += fl;
    fl = k[15]*n[24]*n[54];  // N+++++ + e- -> N++++
    dn[24] -= fl;
    dn[54] -= fl;
    dn[23] += fl;
    fl = k[16]*n[25]*n[54];  // N++++++ + e- -> N+++++
    dn[25] -= fl;
    dn[54] -= fl;
    dn[24] += fl;
    fl = k[17]*n[34]*n[54];  // O+ + e- -> O
    dn[34] -= fl;
    dn[54] -= fl;
    dn[33] += fl;
    fl = k[18]*n[35]*n[54];  // O++ + e- -> O+
    dn[35] -= fl;
    dn[54] -= fl;
    dn[34] += fl;
    fl = k[19]*n[36]*n[54];  // O+++ + e- -> O++
    dn[36] -= fl;
    dn[54] -= fl;
    dn[35] += fl;
    fl = k[20]*n[37]*n[54];  // O++++ + e- -> O+++
    dn[37] -= fl;
    dn[54] -= fl;
    dn[36] += fl;
    fl = k[21]*n[38]*n[54];  // O+++++ + e- -> O++++
    dn[38] -= fl;
    dn[54] -= fl;
    dn[37] += fl;
    fl = k[22]*n[39]*n[54];  // O++++++ + e- -> O+++++
    dn[39] -= fl;
    dn[54] -= fl;
    dn[38] += fl;
    fl = k[23]*n[27]*n[54];  // Ne+ + e- -> Ne
    dn[27] -= fl;
    dn[54] -= fl;
    dn[26] += fl;
    fl = k[24]*n[28]*n[54];  // Ne++ + e- -> Ne+
    dn[28] -= fl;
    dn[54] -= fl;
    dn[27] += fl;
    fl = k[25]*n[29]*n[54];  // Ne+++ + e- -> Ne++
    dn[29] -= fl;
    dn[54] -= fl;
    dn[28] += fl;
    fl = k[26]*n[30]*n[54];  // Ne++++ + e- -> Ne+++
    dn[30] -= fl;
    dn[54] -= fl;
    dn[29] += fl;
    fl = k[27]*n[31]*n[54];  // Ne+++++ + e- -> Ne++++
    dn[31] -= fl;
    dn[54] -= fl;
    dn[30] += fl;
    fl = k[28]*n[32]*n[54];  // Ne++++++ + e- -> Ne+++++
    dn[32] -= fl;
    dn[54] -= fl;
    dn[31] += fl;
    fl = k[29]*n[13]*n[54];  // Mg+ + e- -> Mg
    dn[13] -= fl;
    dn[54] -= fl;
    dn[12] += fl;
    fl = k[30]*n[14]*n[54];  // Mg++ + e- -> Mg+
    dn[14] -= fl;
    dn[54] -= fl;
    dn[13] += fl;
    fl = k[31]*n[15]*n[54];  // Mg+++ + e- -> Mg++
    dn[15] -= fl;
    dn[54] -= fl;
    dn[14] += fl;
    fl = k[32]*n[16]*n[54];  // Mg++++ + e- -> Mg+++
    dn[16] -= fl;
    dn[54] -= fl;
    dn[15] += fl;
    fl = k[33]*n[17]*n[54];  // Mg+++++ + e- -> Mg++++
    dn[17] -= fl;
    dn[54] -= fl;
    dn[16] += fl;
    fl = k[34]*n[18]*n[54];  // Mg++++++ + e- -> Mg+++++
    dn[18] -= fl;
    dn[54] -= fl;
    dn[17] += fl;
    fl = k[35]*n[48]*n[54];  // Si+ + e- -> Si
    dn[48] -= fl;
    dn[54] -= fl;
    dn[47] += fl;
    fl = k[36]*n[49]*n[54];  // Si++ + e- -> Si+
    dn[49] -= fl;
    dn[54] -= fl;
    dn[48] += fl;
    fl = k[37]*n[50]*n[54];  // Si+++ + e- -> Si++
    dn[50] -= fl;
    dn[54] -= fl;
    dn[49] += fl;
    fl = k[38]*n[51]*n[54];  // Si++++ + e- -> Si+++
    dn[51] -= fl;
    dn[54] -= fl;
    dn[50] += fl;
    fl = k[39]*n[52]*n[54];  // Si+++++ + e- -> Si++++
    dn[52] -= fl;
    dn[54] -= fl;
    dn[51] += fl;
    fl = k[40]*n[53]*n[54];  // Si++++++ + e- -> Si+++++
    dn[53] -= fl;
    dn[54] -= fl;
    dn[52] += fl;
    fl = k[41]*n[41]*n[54];  // S+ + e- -> S
    dn[41] -= fl;
    dn[54] -= fl;
    dn[40] += fl;
    fl = k[42]*n[42]*n[54];  // S++ + e- -> S+
    dn[42] -= fl;
    dn[54] -= fl;
    dn[41] += fl;
    fl = k[43]*n[43]*n[54];  // S+++ + e- -> S++
    dn[43] -= fl;
    dn[54] -= fl;
    dn[42] += fl;
    fl = k[44]*n[44]*n[54];  // S++++ + e- -> S+++
    dn[44] -= fl;
    dn[54] -= fl;
    dn[43] += fl;
    fl = k[45]*n[45]*n[54];  // S+++++ + e- -> S++++
    dn[45] -= fl;
    dn[54] -= fl;
    dn[44] += fl;
    fl = k[46]*n[46]*n[54];  // S++++++ + e- -> S+++++
    dn[46] -= fl;
    dn[54] -= fl;
    dn[45] += fl;
    fl = k[47]*n[7]*n[54];  // H + e- -> H+ + e- + e-
    dn[7] -= fl;
    dn[54] -= fl;
    dn[8] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[48]*n[9]*n[54];  // He + e- -> He+ + e- + e-
    dn[9] -= fl;
    dn[54] -= fl;
    dn[10] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[49]*n[10]*n[54];  // He+ + e- -> He++ + e- + e-
    dn[10] -= fl;
    dn[54] -= fl;
    dn[11] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[50]*n[0]*n[54];  // C + e- -> C+ + e- + e-
    dn[0] -= fl;
    dn[54] -= fl;
    dn[1] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[51]*n[1]*n[54];  // C+ + e- -> C++ + e- + e-
    dn[1] -= fl;
    dn[54] -= fl;
    dn[2] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[52]*n[2]*n[54];  // C++ + e- -> C+++ + e- + e-
    dn[2] -= fl;
    dn[54] -= fl;
    dn[3] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[53]*n[3]*n[54];  // C+++ + e- -> C++++ + e- + e-
    dn[3] -= fl;
    dn[54] -= fl;
    dn[4] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[54]*n[4]*n[54];  // C++++ + e- -> C+++++ + e- + e-
    dn[4] -= fl;
    dn[54] -= fl;
    dn[5] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[55]*n[5]*n[54];  // C+++++ + e- -> C++++++ + e- + e-
    dn[5] -= fl;
    dn[54] -= fl;
    dn[6] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[56]*n[19]*n[54];  // N + e- -> N+ + e- + e-
    dn[19] -= fl;
    dn[54] -= fl;
    dn[20] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[57]*n[20]*n[54];  // N+ + e- -> N++ + e- + e-
    dn[20] -= fl;
    dn[54] -= fl;
    dn[21] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[58]*n[21]*n[54];  // N++ + e- -> N+++ + e- + e-
    dn[21] -= fl;
    dn[54] -= fl;
    dn[22] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[59]*n[22]*n[54];  // N+++ + e- -> N++++ + e- + e-
    dn[22] -= fl;
    dn[54] -= fl;
    dn[23] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[60]*n[23]*n[54];  // N++++ + e- -> N+++++ + e- + e-
    dn[23] -= fl;
    dn[54] -= fl;
    dn[24] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[61]*n[24]*n[54];  // N+++++ + e- -> N++++++ + e- + e-
    dn[24] -= fl;
    dn[54] -= fl;
    dn[25] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[62]*n[33]*n[54];  // O + e- -> O+ + e- + e-
    dn[33] -= fl;
    dn[54] -= fl;
    dn[34] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[63]*n[34]*n[54];  // O+ + e- -> O++ + e- + e-
    dn[34] -= fl;
    dn[54] -= fl;
    dn[35] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[64]*n[35]*n[54];  // O++ + e- -> O+++ + e- + e-
    dn[35] -= fl;
    dn[54] -= fl;
    dn[36] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[65]*n[36]*n[54];  // O+++ + e- -> O++++ + e- + e-
    dn[36] -= fl;
    dn[54] -= fl;
    dn[37] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[66]*n[37]*n[54];  // O++++ + e- -> O+++++ + e- + e-
    dn[37] -= fl;
    dn[54] -= fl;
    dn[38] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[67]*n[38]*n[54];  // O+++++ + e- -> O++++++ + e- + e-
    dn[38] -= fl;
    dn[54] -= fl;
    dn[39] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[68]*n[26]*n[54];  // Ne + e- -> Ne+ + e- + e-
    dn[26] -= fl;
    dn[54] -= fl;
    dn[27] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[69]*n[27]*n[54];  // Ne+ + e- -> Ne++ + e- + e-
    dn[27] -= fl;
    dn[54] -= fl;
    dn[28] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[70]*n[28]*n[54];  // Ne++ + e- -> Ne+++ + e- + e-
    dn[28] -= fl;
    dn[54] -= fl;
    dn[29] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[71]*n[29]*n[54];  // Ne+++ + e- -> Ne++++ + e- + e-
    dn[29] -= fl;
    dn[54] -= fl;
    dn[30] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[72]*n[30]*n[54];  // Ne++++ + e- -> Ne+++++ + e- + e-
    dn[30] -= fl;
    dn[54] -= fl;
    dn[31] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[73]*n[31]*n[54];  // Ne+++++ + e- -> Ne++++++ + e- + e-
    dn[31] -= fl;
    dn[54] -= fl;
    dn[32] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[74]*n[12]*n[54];  // Mg + e- -> Mg+ + e- + e-
    dn[12] -= fl;
    dn[54] -= fl;
    dn[13] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[75]*n[13]*n[54];  // Mg+ + e- -> Mg++ + e- + e-
    dn[13] -= fl;
    dn[54] -= fl;
    dn[14] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[76]*n[14]*n[54];  // Mg++ + e- -> Mg+++ + e- + e-
    dn[14] -= fl;
    dn[54] -= fl;
    dn[15] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[77]*n[15]*n[54];  // Mg+++ + e- -> Mg++++ + e- + e-
    dn[15] -= fl;
    dn[54] -= fl;
    dn[16] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[78]*n[16]*n[54];  // Mg++++ + e- -> Mg+++++ + e- + e-
    dn[16] -= fl;
    dn[54] -= fl;
    dn[17] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[79]*n[17]*n[54];  // Mg+++++ + e- -> Mg++++++ + e- + e-
    dn[17] -= fl;
    dn[54] -= fl;
    dn[18] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[80]*n[47]*n[54];  // Si + e- -> Si+ + e- + e-
    dn[47] -= fl;
    dn[54] -= fl;
    dn[48] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[81]*n[48]*n[54];  // Si+ + e- -> Si++ + e- + e-
    dn[48] -= fl;
    dn[54] -= fl;
    dn[49] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[82]*n[49]*n[54];  // Si++ + e- -> Si+++ + e- + e-
    dn[49] -= fl;
    dn[54] -= fl;
    dn[50] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[83]*n[50]*n[54];  // Si+++ + e- -> Si++++ + e- + e-
    dn[50] -= fl;
    dn[54] -= fl;
    dn[51] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[84]*n[51]*n[54];  // Si++++ + e- -> Si+++++ + e- + e-
    dn[51] -= fl;
    dn[54] -= fl;
    dn[52] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[85]*n[52]*n[54];  // Si+++++ + e- -> Si++++++ + e- + e-
    dn[52] -= fl;
    dn[54] -= fl;
    dn[53] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[86]*n[40]*n[54];  // S + e- -> S+ + e- + e-
    dn[40] -= fl;
    dn[54] -= fl;
    dn[41] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[87]*n[41]*n[54];  // S+ + e- -> S++ + e- + e-
    dn[41] -= fl;
    dn[54] -= fl;
    dn[42] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[88]*n[42]*n[54];  // S++ + e- -> S+++ + e- + e-
    dn[42] -= fl;
    dn[54] -= fl;
    dn[43] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[89]*n[43]*n[54];  // S+++ + e- -> S++++ + e- + e-
    dn[43] -= fl;
    dn[54] -= fl;
    dn[44] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[90]*n[44]*n[54];  // S++++ + e- -> S+++++ + e- + e-
    dn[44] -= fl;
    dn[54] -= fl;
    dn[45] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[91]*n[45]*n[54];  // S+++++ + e- -> S++++++ + e- + e-
    dn[45] -= fl;
    dn[54] -= fl;
    dn[46] += fl;
    dn[54] += fl;
    dn[54] += fl;
    fl = k[92]*n[7];  // H -> H+ + e-
    dn[7] -= fl;
    dn[8] += fl;
    dn[54] += fl;
    fl = k[93]*n[9];  // He -> He+ + e-
    dn[9] -= fl;
    dn[10] += fl;
    dn[54] += fl;
    fl = k[94]*n[10];  // He+ -> He++ + e-
    dn[10] -= fl;
    dn[11] += fl;
    dn[54] += fl;
    fl = k[95]*n[0];  // C -> C+ + e-
    dn[0] -= fl;
    dn[1] += fl;
    dn[54] += fl;
    fl = k[96]*n[1];  // C+ -> C++ + e-
    dn[1] -= fl;
    dn[2] += fl;
    dn[54] += fl;
    fl = k[97]*n[2];  // C++ -> C+++ + e-
    dn[2] -= fl;
    dn[3] += fl;
    dn[54] += fl;
    fl = k[98]*n[3];  // C+++ -> C++++ + e-
    dn[3] -= fl;
    dn[4] += fl;
    dn[54] += fl;
    fl = k[99]*n[4];  // C++++ -> C+++++ + e-
    dn[4] -= fl;
    dn[5] += fl;
    dn[54] += fl;
    fl = k[100]*n[5];  // C+++++ -> C++++++ + e-
    dn[5] -= fl;
    dn[6] += fl;
    dn[54] += fl;
    fl = k[101]*n[19];  // N -> N+ + e-
    dn[19] -= fl;
    dn[20] += fl;
    dn[54] += fl;
    fl = k[102]*n[20];  // N+ -> N++ + e-
    dn[20] -= fl;
    dn[21] += fl;
    dn[54] += fl;
    fl = k[103]*n[21];  // N++ -> N+++ + e-
    dn[21] -= fl;
    dn[22] += fl;
    dn[54] += fl;
    fl = k[104]*n[22];  // N+++ -> N++++ + e-
    dn[22] -= fl;
    dn[23] += fl;
    dn[54] += fl;
    fl = k[105]*n[23];  // N++++ -> N+++++ + e-
    dn[23] -= fl;
    dn[24] += fl;
    dn[54] += fl;
    fl = k[106]*n[24];  // N+++++ -> N++++++ + e-
    dn[24] -= fl;
    dn[25] += fl;
    dn[54] += fl;
    fl = k[107]*n[33];  // O -> O+ + e-
    dn[33] -= fl;
    dn[34] += fl;
    dn[54] += fl;
    fl = k[108]*n[34];  // O+ -> O++ + e-
    dn[34] -= fl;
    dn[35] += fl;
    dn[54] += fl;
    fl = k[109]*n[35];  // O++ -> O+++ + e-
    dn[35] -= fl;
    dn[36] += fl;
    dn[54] += fl;
    fl = k[110]*n[36];  // O+++ -> O++++ + e-
    dn[36] -= fl;
    dn[37] += fl;
    dn[54] += fl;
    fl = k[111]*n[37];  // O++++ -> O+++++ + e-
    dn[37] -= fl;
    dn[38] += fl;
    dn[54] += fl;
    fl = k[112]*n[38];  // O+++++ -> O++++++ + e-
    dn[38] -= fl;
    dn[39] += fl;
    dn[54] += fl;
    fl = k[113]*n[26];  // Ne -> Ne+ + e-
    dn[26] -= fl;
    dn[27] += fl;
    dn[54] += fl;
    fl = k[114]*n[27];  // Ne+ -> Ne++ + e-
    dn[27] -= fl;
    dn[28] += fl;
    dn[54] += fl;
    fl = k[115]*n[28];  // Ne++ -> Ne+++ + e-
    dn[28] -= fl;
    dn[29] += fl;
    dn[54] += fl;
    fl = k[116]*n[29];  // Ne+++ -> Ne++++ + e-
    dn[29] -= fl;
    dn[30] += fl;
    dn[54] += fl;
    fl = k[117]*n[30];  // Ne++++ -> Ne+++++ + e-
    dn[30] -= fl;
    dn[31] += fl;
    dn[54] += fl;
    fl = k[118]*n[31];  // Ne+++++ -> Ne++++++ + e-
    dn[31] -= fl;
    dn[32] += fl;
    dn[54] += fl;
    fl = k[119]*n[12];  // Mg -> Mg+ + e-
    dn[12] -= fl;
    dn[13] += fl;
    dn[54] += fl;
    fl = k[120]*n[13];  // Mg+ -> Mg++ + e-
    dn[13] -= fl;
    dn[14] += fl;
    dn[54] += fl;
    fl = k[121]*n[14];  // Mg++ -> Mg+++ + e-
    dn[14] -= fl;
    dn[15] += fl;
    dn[54] += fl;
    fl = k[122]*n[15];  // Mg+++ -> Mg++++ + e-
    dn[15] -= fl;
    dn[16] += fl;
    dn[54] += fl;
    fl = k[123]*n[16];  // Mg++++ -> Mg+++++ + e-
    dn[16] -= fl;
    dn[17] += fl;
    dn[54] += fl;
    fl = k[124]*n[17];  // Mg+++++ -> Mg++++++ + e-
    dn[17] -= fl;
    dn[18] += fl;
    dn[54] += fl;
    fl = k[125]*n[47];  // Si -> Si+ + e-
    dn[47] -= fl;
    dn[48] += fl;
    dn[54] += fl;
    fl = k[126]*n[48];  // Si+ -> Si++ + e-
    dn[48] -= fl;
    dn[49] += fl;
    dn[54] += fl;
    fl = k[127]*n[49];  // Si++ -> Si+++ + e-
    dn[49] -= fl;
    dn[50] += fl;
    dn[54] += fl;
    fl = k[128]*n[50];  // Si+++ -> Si++++ + e-
    dn[50] -= fl;
    dn[51] += fl;
    dn[54] += fl;
    fl = k[129]*n[51];  // Si++++ -> Si+++++ + e-
    dn[51] -= fl;
    dn[52] += fl;
    dn[54] += fl;
    fl = k[130]*n[52];  // Si+++++ -> Si++++++ + e-
    dn[52] -= fl;
    dn[53] += fl;
    dn[54] += fl;
    fl = k[131]*n[40];  // S -> S+ + e-
    dn[40] -= fl;
    dn[41] += fl;
    dn[54] += fl;
    fl = k[132]*n[41];  // S+ -> S++ + e-
    dn[41] -= fl;
    dn[42] += fl;
    dn[54] += fl;
    fl = k[133]*n[42];  // S++ -> S+++ + e-
    dn[42] -= fl;
    dn[43] += fl;
    dn[54] += fl;
    fl = k[134]*n[43];  // S+++ -> S++++ + e-
    dn[43] -= fl;
    dn[44] += fl;
    dn[54] += fl;
    fl = k[135]*n[44];  // S++++ -> S+++++ + e-
    dn[44] -= fl;
    dn[45] += fl;
    dn[54] += fl;
    fl = k[136]*n[45];  // S+++++ -> S++++++ + e-
    dn[45] -= fl;
    dn[46] += fl;
    dn[54] += fl;
    (void)fl;
  }
  template <class R>
  KX_HD static void jac_species_unrolled(const R* n, const R* k, R* J,
                                         int ld) {
    R d;
    d = k[0]*n[54];  // H+ + e- -> H, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[54 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    d = k[0]*n[8];  // H+ + e- -> H, d/d n[54]
    J[8 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[7 + 54 * ld] += d;
    d = k[1]*n[54];  // H+ + e- -> H, d/d n[8]
    J[8 + 8 * ld] -= d;
    J[54 + 8 * ld] -= d;
    J[7 + 8 * ld] += d;
    d = k[1]*n[8];  // H+ + e- -> H, d/d n[54]
    J[8 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[7 + 54 * ld] += d;
    d = k[2]*n[54];  // He+ + e- -> He, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[54 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[2]*n[10];  // He+ + e- -> He, d/d n[54]
    J[10 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[9 + 54 * ld] += d;
    d = k[3]*n[54];  // He+ + e- -> He, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[54 + 10 * ld] -= d;
    J[9 + 10 * ld] += d;
    d = k[3]*n[10];  // He+ + e- -> He, d/d n[54]
    J[10 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[9 + 54 * ld] += d;
    d = k[4]*n[54];  // He++ + e- -> He+, d/d n[11]
    J[11 + 11 * ld] -= d;
    J[54 + 11 * ld] -= d;
    J[10 + 11 * ld] += d;
    d = k[4]*n[11];  // He++ + e- -> He+, d/d n[54]
    J[11 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[10 + 54 * ld] += d;
    d = k[5]*n[54];  // C+ + e- -> C, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[54 + 1 * ld] -= d;
    J[0 + 1 * ld] += d;
    d = k[5]*n[1];  // C+ + e- -> C, d/d n[54]
    J[1 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[0 + 54 * ld] += d;
    d = k[6]*n[54];  // C++ + e- -> C+, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[54 + 2 * ld] -= d;
    J[1 + 2 * ld] += d;
    d = k[6]*n[2];  // C++ + e- -> C+, d/d n[54]
    J[2 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[1 + 54 * ld] += d;
    d = k[7]*n[54];  // C+++ + e- -> C++, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[54 + 3 * ld] -= d;
    J[2 + 3 * ld] += d;
    d = k[7]*n[3];  // C+++ + e- -> C++, d/d n[54]
    J[3 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[2 + 54 * ld] += d;
    d = k[8]*n[54];  // C++++ + e- -> C+++, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[54 + 4 * ld] -= d;
    J[3 + 4 * ld] += d;
    d = k[8]*n[4];  // C++++ + e- -> C+++, d/d n[54]
    J[4 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[3 + 54 * ld] += d;
    d = k[9]*n[54];  // C+++++ + e- -> C++++, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[54 + 5 * ld] -= d;
    J[4 + 5 * ld] += d;
    d = k[9]*n[5];  // C+++++ + e- -> C++++, d/d n[54]
    J[5 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[4 + 54 * ld] += d;
    d = k[10]*n[54];  // C++++++ + e- -> C+++++, d/d n[6]
    J[6 + 6 * ld] -= d;
    J[54 + 6 * ld] -= d;
    J[5 + 6 * ld] += d;
    d = k[10]*n[6];  // C++++++ + e- -> C+++++, d/d n[54]
    J[6 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[5 + 54 * ld] += d;
    d = k[11]*n[54];  // N+ + e- -> N, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[54 + 20 * ld] -= d;
    J[19 + 20 * ld] += d;
    d = k[11]*n[20];  // N+ + e- -> N, d/d n[54]
    J[20 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[19 + 54 * ld] += d;
    d = k[12]*n[54];  // N++ + e- -> N+, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[54 + 21 * ld] -= d;
    J[20 + 21 * ld] += d;
    d = k[12]*n[21];  // N++ + e- -> N+, d/d n[54]
    J[21 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[20 + 54 * ld] += d;
    d = k[13]*n[54];  // N+++ + e- -> N++, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[54 + 22 * ld] -= d;
    J[21 + 22 * ld] += d;
    d = k[13]*n[22];  // N+++ + e- -> N++, d/d n[54]
    J[22 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[21 + 54 * ld] += d;
    d = k[14]*n[54];  // N++++ + e- -> N+++, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[54 + 23 * ld] -= d;
    J[22 + 23 * ld] += d;
    d = k[14]*n[23];  // N++++ + e- -> N+++, d/d n[54]
    J[23 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[22 + 54 * ld] += d;
    d = k[15]*n[54];  // N+++++ + e- -> N++++, d/d n[24]
    J[24 + 24 * ld] -= d;
    J[54 + 24 * ld] -= d;
    J[23 + 24 * ld] += d;
    d = k[15]*n[24];  // N+++++ + e- -> N++++, d/d n[54]
    J[24 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[23 + 54 * ld] += d;
    d = k[16]*n[54];  // N++++++ + e- -> N+++++, d/d n[25]
    J[25 + 25 * ld] -= d;
    J[54 + 25 * ld] -= d;
    J[24 + 25 * ld] += d;
    d = k[16]*n[25];  // N++++++ + e- -> N+++++, d/d n[54]
    J[25 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[24 + 54 * ld] += d;
    d = k[17]*n[54];  // O+ + e- -> O, d/d n[34]
    J[34 + 34 * ld] -= d;
    J[54 + 34 * ld] -= d;
    J[33 + 34 * ld] += d;
    d = k[17]*n[34];  // O+ + e- -> O, d/d n[54]
    J[34 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[33 + 54 * ld] += d;
    d = k[18]*n[54];  // O++ + e- -> O+, d/d n[35]
    J[35 + 35 * ld] -= d;
    J[54 + 35 * ld] -= d;
    J[34 + 35 * ld] += d;
    d = k[18]*n[35];  // O++ + e- -> O+, d/d n[54]
    J[35 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[34 + 54 * ld] += d;
    d = k[19]*n[54];  // O+++ + e- -> O++, d/d n[36]
    J[36 + 36 * ld] -= d;
    J[54 + 36 * ld] -= d;
    J[35 + 36 * ld] += d;
    d = k[19]*n[36];  // O+++ + e- -> O++, d/d n[54]
    J[36 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[35 + 54 * ld] += d;
    d = k[20]*n[54];  // O++++ + e- -> O+++, d/d n[37]
    J[37 + 37 * ld] -= d;
    J[54 + 37 * ld] -= d;
    J[36 + 37 * ld] += d;
    d = k[20]*n[37];  // O++++ + e- -> O+++, d/d n[54]
    J[37 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[36 + 54 * ld] += d;
    d = k[21]*n[54];  // O+++++ + e- -> O++++, d/d n[38]
    J[38 + 38 * ld] -= d;
    J[54 + 38 * ld] -= d;
    J[37 + 38 * ld] += d;
    d = k[21]*n[38];  // O+++++ + e- -> O++++, d/d n[54]
    J[38 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[37 + 54 * ld] += d;
    d = k[22]*n[54];  // O++++++ + e- -> O+++++, d/d n[39]
    J[39 + 39 * ld] -= d;
    J[54 + 39 * ld] -= d;
    J[38 + 39 * ld] += d;
    d = k[22]*n[39];  // O++++++ + e- -> O+++++, d/d n[54]
    J[39 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[38 + 54 * ld] += d;
    d = k[23]*n[54];  // Ne+ + e- -> Ne, d/d n[27]
    J[27 + 27 * ld] -= d;
    J[54 + 27 * ld] -= d;
    J[26 + 27 * ld] += d;
    d = k[23]*n[27];  // Ne+ + e- -> Ne, d/d n[54]
    J[27 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[26 + 54 * ld] += d;
    d = k[24]*n[54];  // Ne++ + e- -> Ne+, d/d n[28]
    J[28 + 28 * ld] -= d;
    J[54 + 28 * ld] -= d;
    J[27 + 28 * ld] += d;
    d = k[24]*n[28];  // Ne++ + e- -> Ne+, d/d n[54]
    J[28 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[27 + 54 * ld] += d;
    d = k[25]*n[54];  // Ne+++ + e- -> Ne++, d/d n[29]
    J[29 + 29 * ld] -= d;
    J[54 + 29 * ld] -= d;
    J[28 + 29 * ld] += d;
    d = k[25]*n[29];  // Ne+++ + e- -> Ne++, d/d n[54]
    J[29 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[28 + 54 * ld] += d;
    d = k[26]*n[54];  // Ne++++ + e- -> Ne+++, d/d n[30]
    J[30 + 30 * ld] -= d;
    J[54 + 30 * ld] -= d;
    J[29 + 30 * ld] += d;
    d = k[26]*n[30];  // Ne++++ + e- -> Ne+++, d/d n[54]
    J[30 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[29 + 54 * ld] += d;
    d = k[27]*n[54];  // Ne+++++ + e- -> Ne++++, d/d n[31]
    J[31 + 31 * ld] -= d;
    J[54 + 31 * ld] -= d;
    J[30 + 31 * ld] += d;
    d = k[27]*n[31];  // Ne+++++ + e- -> Ne++++, d/d n[54]
    J[31 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[30 + 54 * ld] += d;
    d = k[28]*n[54];  // Ne++++++ + e- -> Ne+++++, d/d n[32]
    J[32 + 32 * ld] -= d;
    J[54 + 32 * ld] -= d;
    J[31 + 32 * ld] += d;
    d = k[28]*n[32];  // Ne++++++ + e- -> Ne+++++, d/d n[54]
    J[32 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[31 + 54 * ld] += d;
    d = k[29]*n[54];  // Mg+ + e- -> Mg, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[54 + 13 * ld] -= d;
    J[12 + 13 * ld] += d;
    d = k[29]*n[13];  // Mg+ + e- -> Mg, d/d n[54]
    J[13 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[12 + 54 * ld] += d;
    d = k[30]*n[54];  // Mg++ + e- -> Mg+, d/d n[14]
    J[14 + 14 * ld] -= d;
    J[54 + 14 * ld] -= d;
    J[13 + 14 * ld] += d;
    d = k[30]*n[14];  // Mg++ + e- -> Mg+, d/d n[54]
    J[14 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[13 + 54 * ld] += d;
    d = k[31]*n[54];  // Mg+++ + e- -> Mg++, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[54 + 15 * ld] -= d;
    J[14 + 15 * ld] += d;
    d = k[31]*n[15];  // Mg+++ + e- -> Mg++, d/d n[54]
    J[15 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[14 + 54 * ld] += d;
    d = k[32]*n[54];  // Mg++++ + e- -> Mg+++, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[54 + 16 * ld] -= d;
    J[15 + 16 * ld] += d;
    d = k[32]*n[16];  // Mg++++ + e- -> Mg+++, d/d n[54]
    J[16 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[15 + 54 * ld] += d;
    d = k[33]*n[54];  // Mg+++++ + e- -> Mg++++, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[54 + 17 * ld] -= d;
    J[16 + 17 * ld] += d;
    d = k[33]*n[17];  // Mg+++++ + e- -> Mg++++, d/d n[54]
    J[17 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[16 + 54 * ld] += d;
    d = k[34]*n[54];  // Mg++++++ + e- -> Mg+++++, d/d n[18]
    J[18 + 18 * ld] -= d;
    J[54 + 18 * ld] -= d;
    J[17 + 18 * ld] += d;
    d = k[34]*n[18];  // Mg++++++ + e- -> Mg+++++, d/d n[54]
    J[18 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[17 + 54 * ld] += d;
    d = k[35]*n[54];  // Si+ + e- -> Si, d/d n[48]
    J[48 + 48 * ld] -= d;
    J[54 + 48 * ld] -= d;
    J[47 + 48 * ld] += d;
    d = k[35]*n[48];  // Si+ + e- -> Si, d/d n[54]
    J[48 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[47 + 54 * ld] += d;
    d = k[36]*n[54];  // Si++ + e- -> Si+, d/d n[49]
    J[49 + 49 * ld] -= d;
    J[54 + 49 * ld] -= d;
    J[48 + 49 * ld] += d;
    d = k[36]*n[49];  // Si++ + e- -> Si+, d/d n[54]
    J[49 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[48 + 54 * ld] += d;
    d = k[37]*n[54];  // Si+++ + e- -> Si++, d/d n[50]
    J[50 + 50 * ld] -= d;
    J[54 + 50 * ld] -= d;
    J[49 + 50 * ld] += d;
    d = k[37]*n[50];  // Si+++ + e- -> Si++, d/d n[54]
    J[50 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[49 + 54 * ld] += d;
    d = k[38]*n[54];  // Si++++ + e- -> Si+++, d/d n[51]
    J[51 + 51 * ld] -= d;
    J[54 + 51 * ld] -= d;
    J[50 + 51 * ld] += d;
    d = k[38]*n[51];  // Si++++ + e- -> Si+++, d/d n[54]
    J[51 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[50 + 54 * ld] += d;
    d = k[39]*n[54];  // Si+++++ + e- -> Si++++, d/d n[52]
    J[52 + 52 * ld] -= d;
    J[54 + 52 * ld] -= d;
    J[51 + 52 * ld] += d;
    d = k[39]*n[52];  // Si+++++ + e- -> Si++++, d/d n[54]
    J[52 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[51 + 54 * ld] += d;
    d = k[40]*n[54];  // Si++++++ + e- -> Si+++++, d/d n[53]
    J[53 + 53 * ld] -= d;
    J[54 + 53 * ld] -= d;
    J[52 + 53 * ld] += d;
    d = k[40]*n[53];  // Si++++++ + e- -> Si+++++, d/d n[54]
    J[53 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[52 + 54 * ld] += d;
    d = k[41]*n[54];  // S+ + e- -> S, d/d n[41]
    J[41 + 41 * ld] -= d;
    J[54 + 41 * ld] -= d;
    J[40 + 41 * ld] += d;
    d = k[41]*n[41];  // S+ + e- -> S, d/d n[54]
    J[41 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[40 + 54 * ld] += d;
    d = k[42]*n[54];  // S++ + e- -> S+, d/d n[42]
    J[42 + 42 * ld] -= d;
    J[54 + 42 * ld] -= d;
    J[41 + 42 * ld] += d;
    d = k[42]*n[42];  // S++ + e- -> S+, d/d n[54]
    J[42 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[41 + 54 * ld] += d;
    d = k[43]*n[54];  // S+++ + e- -> S++, d/d n[43]
    J[43 + 43 * ld] -= d;
    J[54 + 43 * ld] -= d;
    J[42 + 43 * ld] += d;
    d = k[43]*n[43];  // S+++ + e- -> S++, d/d n[54]
    J[43 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[42 + 54 * ld] += d;
    d = k[44]*n[54];  // S++++ + e- -> S+++, d/d n[44]
    J[44 + 44 * ld] -= d;
    J[54 + 44 * ld] -= d;
    J[43 + 44 * ld] += d;
    d = k[44]*n[44];  // S++++ + e- -> S+++, d/d n[54]
    J[44 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[43 + 54 * ld] += d;
    d = k[45]*n[54];  // S+++++ + e- -> S++++, d/d n[45]
    J[45 + 45 * ld] -= d;
    J[54 + 45 * ld] -= d;
    J[44 + 45 * ld] += d;
    d = k[45]*n[45];  // S+++++ + e- -> S++++, d/d n[54]
    J[45 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[44 + 54 * ld] += d;
    d = k[46]*n[54];  // S++++++ + e- -> S+++++, d/d n[46]
    J[46 + 46 * ld] -= d;
    J[54 + 46 * ld] -= d;
    J[45 + 46 * ld] += d;
    d = k[46]*n[46];  // S++++++ + e- -> S+++++, d/d n[54]
    J[46 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[45 + 54 * ld] += d;
    d = k[47]*n[54];  // H + e- -> H+ + e- + e-, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[54 + 7 * ld] -= d;
    J[8 + 7 * ld] += d;
    J[54 + 7 * ld] += d;
    J[54 + 7 * ld] += d;
    d = k[47]*n[7];  // H + e- -> H+ + e- + e-, d/d n[54]
    J[7 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[8 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[48]*n[54];  // He + e- -> He+ + e- + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[54 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[54 + 9 * ld] += d;
    J[54 + 9 * ld] += d;
    d = k[48]*n[9];  // He + e- -> He+ + e- + e-, d/d n[54]
    J[9 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[10 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[49]*n[54];  // He+ + e- -> He++ + e- + e-, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[54 + 10 * ld] -= d;
    J[11 + 10 * ld] += d;
    J[54 + 10 * ld] += d;
    J[54 + 10 * ld] += d;
    d = k[49]*n[10];  // He+ + e- -> He++ + e- + e-, d/d n[54]
    J[10 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[11 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[50]*n[54];  // C + e- -> C+ + e- + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[54 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[54 + 0 * ld] += d;
    J[54 + 0 * ld] += d;
    d = k[50]*n[0];  // C + e- -> C+ + e- + e-, d/d n[54]
    J[0 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[1 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[51]*n[54];  // C+ + e- -> C++ + e- + e-, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[54 + 1 * ld] -= d;
    J[2 + 1 * ld] += d;
    J[54 + 1 * ld] += d;
    J[54 + 1 * ld] += d;
    d = k[51]*n[1];  // C+ + e- -> C++ + e- + e-, d/d n[54]
    J[1 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[2 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[52]*n[54];  // C++ + e- -> C+++ + e- + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[54 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[54 + 2 * ld] += d;
    J[54 + 2 * ld] += d;
    d = k[52]*n[2];  // C++ + e- -> C+++ + e- + e-, d/d n[54]
    J[2 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[3 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[53]*n[54];  // C+++ + e- -> C++++ + e- + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[54 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[54 + 3 * ld] += d;
    J[54 + 3 * ld] += d;
    d = k[53]*n[3];  // C+++ + e- -> C++++ + e- + e-, d/d n[54]
    J[3 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[4 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[54]*n[54];  // C++++ + e- -> C+++++ + e- + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[54 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[54 + 4 * ld] += d;
    J[54 + 4 * ld] += d;
    d = k[54]*n[4];  // C++++ + e- -> C+++++ + e- + e-, d/d n[54]
    J[4 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[5 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[55]*n[54];  // C+++++ + e- -> C++++++ + e- + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[54 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[54 + 5 * ld] += d;
    J[54 + 5 * ld] += d;
    d = k[55]*n[5];  // C+++++ + e- -> C++++++ + e- + e-, d/d n[54]
    J[5 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[6 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[56]*n[54];  // N + e- -> N+ + e- + e-, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[54 + 19 * ld] -= d;
    J[20 + 19 * ld] += d;
    J[54 + 19 * ld] += d;
    J[54 + 19 * ld] += d;
    d = k[56]*n[19];  // N + e- -> N+ + e- + e-, d/d n[54]
    J[19 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[20 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[57]*n[54];  // N+ + e- -> N++ + e- + e-, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[54 + 20 * ld] -= d;
    J[21 + 20 * ld] += d;
    J[54 + 20 * ld] += d;
    J[54 + 20 * ld] += d;
    d = k[57]*n[20];  // N+ + e- -> N++ + e- + e-, d/d n[54]
    J[20 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[21 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[58]*n[54];  // N++ + e- -> N+++ + e- + e-, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[54 + 21 * ld] -= d;
    J[22 + 21 * ld] += d;
    J[54 + 21 * ld] += d;
    J[54 + 21 * ld] += d;
    d = k[58]*n[21];  // N++ + e- -> N+++ + e- + e-, d/d n[54]
    J[21 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[22 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[59]*n[54];  // N+++ + e- -> N++++ + e- + e-, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[54 + 22 * ld] -= d;
    J[23 + 22 * ld] += d;
    J[54 + 22 * ld] += d;
    J[54 + 22 * ld] += d;
    d = k[59]*n[22];  // N+++ + e- -> N++++ + e- + e-, d/d n[54]
    J[22 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[23 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[60]*n[54];  // N++++ + e- -> N+++++ + e- + e-, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[54 + 23 * ld] -= d;
    J[24 + 23 * ld] += d;
    J[54 + 23 * ld] += d;
    J[54 + 23 * ld] += d;
    d = k[60]*n[23];  // N++++ + e- -> N+++++ + e- + e-, d/d n[54]
    J[23 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[24 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[61]*n[54];  // N+++++ + e- -> N++++++ + e- + e-, d/d n[24]
    J[24 + 24 * ld] -= d;
    J[54 + 24 * ld] -= d;
    J[25 + 24 * ld] += d;
    J[54 + 24 * ld] += d;
    J[54 + 24 * ld] += d;
    d = k[61]*n[24];  // N+++++ + e- -> N++++++ + e- + e-, d/d n[54]
    J[24 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[25 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[62]*n[54];  // O + e- -> O+ + e- + e-, d/d n[33]
    J[33 + 33 * ld] -= d;
    J[54 + 33 * ld] -= d;
    J[34 + 33 * ld] += d;
    J[54 + 33 * ld] += d;
    J[54 + 33 * ld] += d;
    d = k[62]*n[33];  // O + e- -> O+ + e- + e-, d/d n[54]
    J[33 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[34 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[63]*n[54];  // O+ + e- -> O++ + e- + e-, d/d n[34]
    J[34 + 34 * ld] -= d;
    J[54 + 34 * ld] -= d;
    J[35 + 34 * ld] += d;
    J[54 + 34 * ld] += d;
    J[54 + 34 * ld] += d;
    d = k[63]*n[34];  // O+ + e- -> O++ + e- + e-, d/d n[54]
    J[34 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[35 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[64]*n[54];  // O++ + e- -> O+++ + e- + e-, d/d n[35]
    J[35 + 35 * ld] -= d;
    J[54 + 35 * ld] -= d;
    J[36 + 35 * ld] += d;
    J[54 + 35 * ld] += d;
    J[54 + 35 * ld] += d;
    d = k[64]*n[35];  // O++ + e- -> O+++ + e- + e-, d/d n[54]
    J[35 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[36 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[65]*n[54];  // O+++ + e- -> O++++ + e- + e-, d/d n[36]
    J[36 + 36 * ld] -= d;
    J[54 + 36 * ld] -= d;
    J[37 + 36 * ld] += d;
    J[54 + 36 * ld] += d;
    J[54 + 36 * ld] += d;
    d = k[65]*n[36];  // O+++ + e- -> O++++ + e- + e-, d/d n[54]
    J[36 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[37 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[66]*n[54];  // O++++ + e- -> O+++++ + e- + e-, d/d n[37]
    J[37 + 37 * ld] -= d;
    J[54 + 37 * ld] -= d;
    J[38 + 37 * ld] += d;
    J[54 + 37 * ld] += d;
    J[54 + 37 * ld] += d;
    d = k[66]*n[37];  // O++++ + e- -> O+++++ + e- + e-, d/d n[54]
    J[37 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[38 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[67]*n[54];  // O+++++ + e- -> O++++++ + e- + e-, d/d n[38]
    J[38 + 38 * ld] -= d;
    J[54 + 38 * ld] -= d;
    J[39 + 38 * ld] += d;
    J[54 + 38 * ld] += d;
    J[54 + 38 * ld] += d;
    d = k[67]*n[38];  // O+++++ + e- -> O++++++ + e- + e-, d/d n[54]
    J[38 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[39 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[68]*n[54];  // Ne + e- -> Ne+ + e- + e-, d/d n[26]
    J[26 + 26 * ld] -= d;
    J[54 + 26 * ld] -= d;
    J[27 + 26 * ld] += d;
    J[54 + 26 * ld] += d;
    J[54 + 26 * ld] += d;
    d = k[68]*n[26];  // Ne + e- -> Ne+ + e- + e-, d/d n[54]
    J[26 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[27 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[69]*n[54];  // Ne+ + e- -> Ne++ + e- + e-, d/d n[27]
    J[27 + 27 * ld] -= d;
    J[54 + 27 * ld] -= d;
    J[28 + 27 * ld] += d;
    J[54 + 27 * ld] += d;
    J[54 + 27 * ld] += d;
    d = k[69]*n[27];  // Ne+ + e- -> Ne++ + e- + e-, d/d n[54]
    J[27 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[28 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[70]*n[54];  // Ne++ + e- -> Ne+++ + e- + e-, d/d n[28]
    J[28 + 28 * ld] -= d;
    J[54 + 28 * ld] -= d;
    J[29 + 28 * ld] += d;
    J[54 + 28 * ld] += d;
    J[54 + 28 * ld] += d;
    d = k[70]*n[28];  // Ne++ + e- -> Ne+++ + e- + e-, d/d n[54]
    J[28 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[29 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[71]*n[54];  // Ne+++ + e- -> Ne++++ + e- + e-, d/d n[29]
    J[29 + 29 * ld] -= d;
    J[54 + 29 * ld] -= d;
    J[30 + 29 * ld] += d;
    J[54 + 29 * ld] += d;
    J[54 + 29 * ld] += d;
    d = k[71]*n[29];  // Ne+++ + e- -> Ne++++ + e- + e-, d/d n[54]
    J[29 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[30 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[72]*n[54];  // Ne++++ + e- -> Ne+++++ + e- + e-, d/d n[30]
    J[30 + 30 * ld] -= d;
    J[54 + 30 * ld] -= d;
    J[31 + 30 * ld] += d;
    J[54 + 30 * ld] += d;
    J[54 + 30 * ld] += d;
    d = k[72]*n[30];  // Ne++++ + e- -> Ne+++++ + e- + e-, d/d n[54]
    J[30 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[31 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[73]*n[54];  // Ne+++++ + e- -> Ne++++++ + e- + e-, d/d n[31]
    J[31 + 31 * ld] -= d;
    J[54 + 31 * ld] -= d;
    J[32 + 31 * ld] += d;
    J[54 + 31 * ld] += d;
    J[54 + 31 * ld] += d;
    d = k[73]*n[31];  // Ne+++++ + e- -> Ne++++++ + e- + e-, d/d n[54]
    J[31 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[32 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[74]*n[54];  // Mg + e- -> Mg+ + e- + e-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[54 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[54 + 12 * ld] += d;
    J[54 + 12 * ld] += d;
    d = k[74]*n[12];  // Mg + e- -> Mg+ + e- + e-, d/d n[54]
    J[12 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[13 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[75]*n[54];  // Mg+ + e- -> Mg++ + e- + e-, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[54 + 13 * ld] -= d;
    J[14 + 13 * ld] += d;
    J[54 + 13 * ld] += d;
    J[54 + 13 * ld] += d;
    d = k[75]*n[13];  // Mg+ + e- -> Mg++ + e- + e-, d/d n[54]
    J[13 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[14 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[76]*n[54];  // Mg++ + e- -> Mg+++ + e- + e-, d/d n[14]
    J[14 + 14 * ld] -= d;
    J[54 + 14 * ld] -= d;
    J[15 + 14 * ld] += d;
    J[54 + 14 * ld] += d;
    J[54 + 14 * ld] += d;
    d = k[76]*n[14];  // Mg++ + e- -> Mg+++ + e- + e-, d/d n[54]
    J[14 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[15 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[77]*n[54];  // Mg+++ + e- -> Mg++++ + e- + e-, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[54 + 15 * ld] -= d;
    J[16 + 15 * ld] += d;
    J[54 + 15 * ld] += d;
    J[54 + 15 * ld] += d;
    d = k[77]*n[15];  // Mg+++ + e- -> Mg++++ + e- + e-, d/d n[54]
    J[15 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[16 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[78]*n[54];  // Mg++++ + e- -> Mg+++++ + e- + e-, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[54 + 16 * ld] -= d;
    J[17 + 16 * ld] += d;
    J[54 + 16 * ld] += d;
    J[54 + 16 * ld] += d;
    d = k[78]*n[16];  // Mg++++ + e- -> Mg+++++ + e- + e-, d/d n[54]
    J[16 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[17 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[79]*n[54];  // Mg+++++ + e- -> Mg++++++ + e- + e-, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[54 + 17 * ld] -= d;
    J[18 + 17 * ld] += d;
    J[54 + 17 * ld] += d;
    J[54 + 17 * ld] += d;
    d = k[79]*n[17];  // Mg+++++ + e- -> Mg++++++ + e- + e-, d/d n[54]
    J[17 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[18 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[80]*n[54];  // Si + e- -> Si+ + e- + e-, d/d n[47]
    J[47 + 47 * ld] -= d;
    J[54 + 47 * ld] -= d;
    J[48 + 47 * ld] += d;
    J[54 + 47 * ld] += d;
    J[54 + 47 * ld] += d;
    d = k[80]*n[47];  // Si + e- -> Si+ + e- + e-, d/d n[54]
    J[47 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[48 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[81]*n[54];  // Si+ + e- -> Si++ + e- + e-, d/d n[48]
    J[48 + 48 * ld] -= d;
    J[54 + 48 * ld] -= d;
    J[49 + 48 * ld] += d;
    J[54 + 48 * ld] += d;
    J[54 + 48 * ld] += d;
    d = k[81]*n[48];  // Si+ + e- -> Si++ + e- + e-, d/d n[54]
    J[48 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[49 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[82]*n[54];  // Si++ + e- -> Si+++ + e- + e-, d/d n[49]
    J[49 + 49 * ld] -= d;
    J[54 + 49 * ld] -= d;
    J[50 + 49 * ld] += d;
    J[54 + 49 * ld] += d;
    J[54 + 49 * ld] += d;
    d = k[82]*n[49];  // Si++ + e- -> Si+++ + e- + e-, d/d n[54]
    J[49 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[50 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[83]*n[54];  // Si+++ + e- -> Si++++ + e- + e-, d/d n[50]
    J[50 + 50 * ld] -= d;
    J[54 + 50 * ld] -= d;
    J[51 + 50 * ld] += d;
    J[54 + 50 * ld] += d;
    J[54 + 50 * ld] += d;
    d = k[83]*n[50];  // Si+++ + e- -> Si++++ + e- + e-, d/d n[54]
    J[50 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[51 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[84]*n[54];  // Si++++ + e- -> Si+++++ + e- + e-, d/d n[51]
    J[51 + 51 * ld] -= d;
    J[54 + 51 * ld] -= d;
    J[52 + 51 * ld] += d;
    J[54 + 51 * ld] += d;
    J[54 + 51 * ld] += d;
    d = k[84]*n[51];  // Si++++ + e- -> Si+++++ + e- + e-, d/d n[54]
    J[51 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[52 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[85]*n[54];  // Si+++++ + e- -> Si++++++ + e- + e-, d/d n[52]
    J[52 + 52 * ld] -= d;
    J[54 + 52 * ld] -= d;
    J[53 + 52 * ld] += d;
    J[54 + 52 * ld] += d;
    J[54 + 52 * ld] += d;
    d = k[85]*n[52];  // Si+++++ + e- -> Si++++++ + e- + e-, d/d n[54]
    J[52 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[53 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[86]*n[54];  // S + e- -> S+ + e- + e-, d/d n[40]
    J[40 + 40 * ld] -= d;
    J[54 + 40 * ld] -= d;
    J[41 + 40 * ld] += d;
    J[54 + 40 * ld] += d;
    J[54 + 40 * ld] += d;
    d = k[86]*n[40];  // S + e- -> S+ + e- + e-, d/d n[54]
    J[40 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[41 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[87]*n[54];  // S+ + e- -> S++ + e- + e-, d/d n[41]
    J[41 + 41 * ld] -= d;
    J[54 + 41 * ld] -= d;
    J[42 + 41 * ld] += d;
    J[54 + 41 * ld] += d;
    J[54 + 41 * ld] += d;
    d = k[87]*n[41];  // S+ + e- -> S++ + e- + e-, d/d n[54]
    J[41 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[42 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[88]*n[54];  // S++ + e- -> S+++ + e- + e-, d/d n[42]
    J[42 + 42 * ld] -= d;
    J[54 + 42 * ld] -= d;
    J[43 + 42 * ld] += d;
    J[54 + 42 * ld] += d;
    J[54 + 42 * ld] += d;
    d = k[88]*n[42];  // S++ + e- -> S+++ + e- + e-, d/d n[54]
    J[42 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[43 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[89]*n[54];  // S+++ + e- -> S++++ + e- + e-, d/d n[43]
    J[43 + 43 * ld] -= d;
    J[54 + 43 * ld] -= d;
    J[44 + 43 * ld] += d;
    J[54 + 43 * ld] += d;
    J[54 + 43 * ld] += d;
    d = k[89]*n[43];  // S+++ + e- -> S++++ + e- + e-, d/d n[54]
    J[43 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[44 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[90]*n[54];  // S++++ + e- -> S+++++ + e- + e-, d/d n[44]
    J[44 + 44 * ld] -= d;
    J[54 + 44 * ld] -= d;
    J[45 + 44 * ld] += d;
    J[54 + 44 * ld] += d;
    J[54 + 44 * ld] += d;
    d = k[90]*n[44];  // S++++ + e- -> S+++++ + e- + e-, d/d n[54]
    J[44 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[45 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[91]*n[54];  // S+++++ + e- -> S++++++ + e- + e-, d/d n[45]
    J[45 + 45 * ld] -= d;
    J[54 + 45 * ld] -= d;
    J[46 + 45 * ld] += d;
    J[54 + 45 * ld] += d;
    J[54 + 45 * ld] += d;
    d = k[91]*n[45];  // S+++++ + e- -> S++++++ + e- + e-, d/d n[54]
    J[45 + 54 * ld] -= d;
    J[54 + 54 * ld] -= d;
    J[46 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    J[54 + 54 * ld] += d;
    d = k[92];  // H -> H+ + e-, d/d n[7]
    J[7 + 7 * ld] -= d;
    J[8 + 7 * ld] += d;
    J[54 + 7 * ld] += d;
    d = k[93];  // He -> He+ + e-, d/d n[9]
    J[9 + 9 * ld] -= d;
    J[10 + 9 * ld] += d;
    J[54 + 9 * ld] += d;
    d = k[94];  // He+ -> He++ + e-, d/d n[10]
    J[10 + 10 * ld] -= d;
    J[11 + 10 * ld] += d;
    J[54 + 10 * ld] += d;
    d = k[95];  // C -> C+ + e-, d/d n[0]
    J[0 + 0 * ld] -= d;
    J[1 + 0 * ld] += d;
    J[54 + 0 * ld] += d;
    d = k[96];  // C+ -> C++ + e-, d/d n[1]
    J[1 + 1 * ld] -= d;
    J[2 + 1 * ld] += d;
    J[54 + 1 * ld] += d;
    d = k[97];  // C++ -> C+++ + e-, d/d n[2]
    J[2 + 2 * ld] -= d;
    J[3 + 2 * ld] += d;
    J[54 + 2 * ld] += d;
    d = k[98];  // C+++ -> C++++ + e-, d/d n[3]
    J[3 + 3 * ld] -= d;
    J[4 + 3 * ld] += d;
    J[54 + 3 * ld] += d;
    d = k[99];  // C++++ -> C+++++ + e-, d/d n[4]
    J[4 + 4 * ld] -= d;
    J[5 + 4 * ld] += d;
    J[54 + 4 * ld] += d;
    d = k[100];  // C+++++ -> C++++++ + e-, d/d n[5]
    J[5 + 5 * ld] -= d;
    J[6 + 5 * ld] += d;
    J[54 + 5 * ld] += d;
    d = k[101];  // N -> N+ + e-, d/d n[19]
    J[19 + 19 * ld] -= d;
    J[20 + 19 * ld] += d;
    J[54 + 19 * ld] += d;
    d = k[102];  // N+ -> N++ + e-, d/d n[20]
    J[20 + 20 * ld] -= d;
    J[21 + 20 * ld] += d;
    J[54 + 20 * ld] += d;
    d = k[103];  // N++ -> N+++ + e-, d/d n[21]
    J[21 + 21 * ld] -= d;
    J[22 + 21 * ld] += d;
    J[54 + 21 * ld] += d;
    d = k[104];  // N+++ -> N++++ + e-, d/d n[22]
    J[22 + 22 * ld] -= d;
    J[23 + 22 * ld] += d;
    J[54 + 22 * ld] += d;
    d = k[105];  // N++++ -> N+++++ + e-, d/d n[23]
    J[23 + 23 * ld] -= d;
    J[24 + 23 * ld] += d;
    J[54 + 23 * ld] += d;
    d = k[106];  // N+++++ -> N++++++ + e-, d/d n[24]
    J[24 + 24 * ld] -= d;
    J[25 + 24 * ld] += d;
    J[54 + 24 * ld] += d;
    d = k[107];  // O -> O+ + e-, d/d n[33]
    J[33 + 33 * ld] -= d;
    J[34 + 33 * ld] += d;
    J[54 + 33 * ld] += d;
    d = k[108];  // O+ -> O++ + e-, d/d n[34]
    J[34 + 34 * ld] -= d;
    J[35 + 34 * ld] += d;
    J[54 + 34 * ld] += d;
    d = k[109];  // O++ -> O+++ + e-, d/d n[35]
    J[35 + 35 * ld] -= d;
    J[36 + 35 * ld] += d;
    J[54 + 35 * ld] += d;
    d = k[110];  // O+++ -> O++++ + e-, d/d n[36]
    J[36 + 36 * ld] -= d;
    J[37 + 36 * ld] += d;
    J[54 + 36 * ld] += d;
    d = k[111];  // O++++ -> O+++++ + e-, d/d n[37]
    J[37 + 37 * ld] -= d;
    J[38 + 37 * ld] += d;
    J[54 + 37 * ld] += d;
    d = k[112];  // O+++++ -> O++++++ + e-, d/d n[38]
    J[38 + 38 * ld] -= d;
    J[39 + 38 * ld] += d;
    J[54 + 38 * ld] += d;
    d = k[113];  // Ne -> Ne+ + e-, d/d n[26]
    J[26 + 26 * ld] -= d;
    J[27 + 26 * ld] += d;
    J[54 + 26 * ld] += d;
    d = k[114];  // Ne+ -> Ne++ + e-, d/d n[27]
    J[27 + 27 * ld] -= d;
    J[28 + 27 * ld] += d;
    J[54 + 27 * ld] += d;
    d = k[115];  // Ne++ -> Ne+++ + e-, d/d n[28]
    J[28 + 28 * ld] -= d;
    J[29 + 28 * ld] += d;
    J[54 + 28 * ld] += d;
    d = k[116];  // Ne+++ -> Ne++++ + e-, d/d n[29]
    J[29 + 29 * ld] -= d;
    J[30 + 29 * ld] += d;
    J[54 + 29 * ld] += d;
    d = k[117];  // Ne++++ -> Ne+++++ + e-, d/d n[30]
    J[30 + 30 * ld] -= d;
    J[31 + 30 * ld] += d;
    J[54 + 30 * ld] += d;
    d = k[118];  // Ne+++++ -> Ne++++++ + e-, d/d n[31]
    J[31 + 31 * ld] -= d;
    J[32 + 31 * ld] += d;
    J[54 + 31 * ld] += d;
    d = k[119];  // Mg -> Mg+ + e-, d/d n[12]
    J[12 + 12 * ld] -= d;
    J[13 + 12 * ld] += d;
    J[54 + 12 * ld] += d;
    d = k[120];  // Mg+ -> Mg++ + e-, d/d n[13]
    J[13 + 13 * ld] -= d;
    J[14 + 13 * ld] += d;
    J[54 + 13 * ld] += d;
    d = k[121];  // Mg++ -> Mg+++ + e-, d/d n[14]
    J[14 + 14 * ld] -= d;
    J[15 + 14 * ld] += d;
    J[54 + 14 * ld] += d;
    d = k[122];  // Mg+++ -> Mg++++ + e-, d/d n[15]
    J[15 + 15 * ld] -= d;
    J[16 + 15 * ld] += d;
    J[54 + 15 * ld] += d;
    d = k[123];  // Mg++++ -> Mg+++++ + e-, d/d n[16]
    J[16 + 16 * ld] -= d;
    J[17 + 16 * ld] += d;
    J[54 + 16 * ld] += d;
    d = k[124];  // Mg+++++ -> Mg++++++ + e-, d/d n[17]
    J[17 + 17 * ld] -= d;
    J[18 + 17 * ld] += d;
    J[54 + 17 * ld] += d;
    d = k[125];  // Si -> Si+ + e-, d/d n[47]
    J[47 + 47 * ld] -= d;
    J[48 + 47 * ld] += d;
    J[54 + 47 * ld] += d;
    d = k[126];  // Si+ -> Si++ + e-, d/d n[48]
    J[48 + 48 * ld] -= d;
    J[49 + 48 * ld] += d;
    J[54 + 48 * ld] += d;
    d = k[127];  // Si++ -> Si+++ + e-, d/d n[49]
    J[49 + 49 * ld] -= d;
    J[50 + 49 * ld] += d;
    J[54 + 49 * ld] += d;
    d = k[128];  // Si+++ -> Si++++ + e-, d/d n[50]
    J[50 + 50 * ld] -= d;
    J[51 + 50 * ld] += d;
    J[54 + 50 * ld] += d;
    d = k[129];  // Si++++ -> Si+++++ + e-, d/d n[51]
    J[51 + 51 * ld] -= d;
    J[52 + 51 * ld] += d;
    J[54 + 51 * ld] += d;
    d = k[130];  // Si+++++ -> Si++++++ + e-, d/d n[52]
    J[52 + 52 * ld] -= d;
    J[53 + 52 * ld] += d;
    J[54 + 52 * ld] += d;
    d = k[131];  // S -> S+ + e-, d/d n[40]
    J[40 + 40 * ld] -= d;
    J[41 + 40 * ld] += d;
    J[54 + 40 * ld] += d;
    d = k[132];  // S+ -> S++ + e-, d/d n[41]
    J[41 + 41 * ld] -= d;
    J[42 + 41 * ld] += d;
    J[54 + 41 * ld] += d;
    d = k[133];  // S++ -> S+++ + e-, d/d n[42]
    J[42 + 42 * ld] -= d;
    J[43 + 42 * ld] += d;
    J[54 + 42 * ld] += d;
    d = k[134];  // S+++ -> S++++ + e-, d/d n[43]
    J[43 + 43 * ld] -= d;
    J[44 + 43 * ld] += d;
    J[54 + 43 * ld] += d;
    d = k[135];  // S++++ -> S+++++ + e-, d/d n[44]
    J[44 + 44 * ld] -= d;
    J[45 + 44 * ld] += d;
    J[54 + 44 * ld] += d;
    d = k[136];  // S+++++ -> S++++++ + e-, d/d n[45]
    J[45 + 45 * ld] -= d;
    J[46 + 45 * ld] += d;
    J[54 + 45 * ld] += d;
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
