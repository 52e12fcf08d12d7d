// Runtime physics of the per-cell right-hand side, for host (the CPU
// oracle, g++) and gfx950 device code (hipcc), each a template on the real
// type R (double on every path that runs; the generated rhs shares the
// template).  The formulas are the published fits KROME-X evaluates; their
// coefficients are physical data and are the same numbers:
//   cooling_H2          Glover & Abel 2008 + HM79/Glover 2015 LTE limit
//                       (runtime/core/cooling.hpp:74-177)
//   cooling_continuum   Lenzuni+96 Planck opacity with a Jeans-length
//                       escape fraction (cooling.hpp:283-303)
//   dissH2_Martin96     Martin, Schwarz & Mandy 1996 H2 collisional
//                       dissociation (runtime/core/special_rates.hpp:20-66)
//   h2_heat_fraction    HM79 thermalized fraction of the H2 formation
//                       heat (kromex/codegen.py:5005-5011)
//   h2_opacity_ripamonti  Ripamonti & Abel 2004 (codegen.py:3657-3660)
//   gamma_rot, gamma_rot_ortho_para  exact rotational heat capacities
//                       (codegen.py:3745-3779)
//   H2_solomonLW        Solomon-process H2 photodissociation from J(12.87
//                       eV), k = 1.38e9 J (runtime/core/h2_shielding.hpp:
//                       84-94)
//   H/He_recombination_on_dust  Weingartner & Draine 2001 grain-assisted
//                       recombination fits (dust_rates.hpp:34-66)
// and, on the host only, the photo cross sections the photon bins are
// built from: sigma_v96 (Verner et al. 1996 fit), xsec_interp (a table
// interpolated linearly, zero outside it) and H2_sigmaLW (the Lyman-Werner
// line and continuum fits, h2_shielding.hpp:96-125).
#pragma once
#include <cmath>
#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
// every helper of the per-cell solve is inlined into its kernel: the
// persistent kernel evaluates the rhs at one program point and keeps the
// lanes' state in registers, which calls through the stack would defeat
#define KX_HD __host__ __device__ __attribute__((always_inline))
// a function compiled once and called (large bodies that run rarely)
#define KX_HD_CALL __host__ __device__ __attribute__((noinline))
#else
#define KX_HD
#define KX_HD_CALL
#endif

namespace kx {

// the math of the rhs under one name.  On the GPU every logarithm is
// log2 (scaled by a constant): the rhs takes log, log10 and powers of the
// same few arguments (Tgas, Te), and one log2 per argument then serves
// them all; a logarithm of a constant folds at compile time.  Each form
// differs from the C library's by a rounding or two (relative 1e-15, or
// |y log2 x| 2^-52 for a power: under 1e-14 for the fits' arguments).
// The host build (the oracle) keeps the C library's functions.
#if defined(__HIP_DEVICE_COMPILE__)
KX_HD inline double log2_(double x) {
  return __builtin_constant_p(x) ? __builtin_log2(x) : ::log2(x);
}
#endif
KX_HD inline double exp(double x) { return ::exp(x); }
KX_HD inline double log(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return log2_(x) * 0.69314718055994530942;
#else
  return ::log(x);
#endif
}
KX_HD inline double log10(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return log2_(x) * 0.30102999566398119521;
#else
  return ::log10(x);
#endif
}
KX_HD inline double sqrt(double x) { return ::sqrt(x); }
// x^y, on the GPU as 2^(y log2 x).  The rhs's bases are positive
// (temperatures, densities, 1 + ...); x = 0 gives pow's 0 or inf for
// y != 0, and y = 0 gives 1.
KX_HD inline double pow(double x, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return y == 0.0 ? 1.0 : ::exp2(y * log2_(x));
#else
  return ::pow(x, y);
#endif
}
KX_HD inline double tanh(double x) { return ::tanh(x); }
KX_HD inline double fabs(double x) { return ::fabs(x); }
KX_HD inline double fmin(double x, double y) { return ::fmin(x, y); }
KX_HD inline double fmax(double x, double y) { return ::fmax(x, y); }
// a finite value (not inf, not NaN).  By the builtin, not by x - x == 0:
// with multiply-add contraction, x - x of a product a b compiles to
// fma(a, b, -a b), the product's rounding error, which need not be 0
KX_HD inline bool finite(double x) { return __builtin_isfinite(x); }
KX_HD inline bool finite(float x) { return __builtin_isfinite(x); }
// (FP32: the mixed-precision Newton matrix)
KX_HD inline float fabs(float x) { return ::fabsf(x); }
KX_HD inline float fmin(float x, float y) { return ::fminf(x, y); }
KX_HD inline float fmax(float x, float y) { return ::fmaxf(x, y); }

// x^N for an integer constant N by binary powering
template <int N, class R>
KX_HD inline R ipow(R x) {
  if constexpr (N < 0) {
    return R(1) / ipow<-N>(x);
  } else if constexpr (N == 0) {
    (void)x;
    return R(1);
  } else if constexpr (N == 1) {
    return x;
  } else {
    const R h = ipow<N / 2>(x);
    if constexpr (N % 2 == 1) return h * h * x;
    else return h * h;
  }
}

}  // namespace kx

// the dual numbers of the adjoint's exact Jacobians: declared here, after
// the scalar math and before every template written over it, so that the
// kx:: calls of those templates find the dual overloads too
#include "kx_dual.h"

namespace kx {

namespace phys {
constexpr double kB = 1.380649e-16;           // erg / K
constexpr double p_mass = 1.67262192369e-24;  // g
constexpr double gravity = 6.6743e-08;        // cm^3 g^-1 s^-2
constexpr double stefboltz = 5.6703744191844286e-05;
constexpr double pi = 3.141592653589793;
}  // namespace phys

// smooth window in log T: 10^(200 (s(x, -0.2) s(-x, -1.2) - 1)) with
// s(x, x0) = 10 / (10 + exp(-50 (x - x0))), x the position of log T in
// [lmin, lmax]; 0 below 1e-199
template <class R>
KX_HD inline R cool_window(R logT, R lmin, R lmax) {
  const R x = (logT - lmin) / (lmax - lmin);
  const R s1 = R(10) / (R(10) + kx::exp(R(-50) * (x - R(-0.2))));
  const R s2 = R(10) / (R(10) + kx::exp(R(-50) * (-x - R(-1.2))));
  const R w = kx::pow(R(10), R(200) * (s1 * s2 - R(1)));
  return w < R(1e-199) ? R(0) : w;
}

// H2 line cooling [erg cm^-3 s^-1]: the low-density rates of the
// colliders H, H+, H2, e-, He combined with the LTE limit HDL as
// n(H2) / (1/HDL + 1/L_low)
template <class R>
KX_HD inline R cooling_H2(R T, R nH, R nHj, R nH2, R ne, R nHe) {
  const R T3 = T * R(1e-3);
  const R l3 = kx::log10(T3);
  const R lT = kx::log10(T);
  const R l2 = l3 * l3, l3_ = l2 * l3, l4 = l3_ * l3, l5 = l4 * l3,
          l6 = l5 * l3, l7 = l6 * l3, l8 = l7 * l3;
  const R w14 = cool_window(lT, R(1), R(4));
  const R w24 = cool_window(lT, R(2), R(4));
  R cool = R(0);
  // H2-H
  R f;
  if (T <= R(1e2))
    f = kx::pow(R(10), R(-16.818342) + R(37.383713) * l3 + R(58.145166) * l2
                + R(48.656103) * l3_ + R(20.159831) * l4 + R(3.8479610) * l5);
  else if (T <= R(1e3))
    f = kx::pow(R(10), R(-24.311209) + R(3.5692468) * l3 - R(11.332860) * l2
                - R(27.850082) * l3_ - R(21.328264) * l4 - R(4.2519023) * l5);
  else if (T <= R(6e3))
    f = kx::pow(R(10), R(-24.311209) + R(4.6450521) * l3 - R(3.7209846) * l2
                + R(5.9369081) * l3_ - R(5.5108049) * l4 + R(1.5538288) * l5);
  else
    f = R(1.862314467912518e-22) * cool_window(lT, R(1), kx::log10(R(6e3)));
  cool += f * nH;
  // H2-H+
  if (T > R(1e1) && T <= R(1e4))
    f = kx::pow(R(10), R(-22.089523) + R(1.5714711) * l3 + R(0.015391166) * l2
                - R(0.23619985) * l3_ - R(0.51002221) * l4
                + R(0.32168730) * l5) * nHj;
  else
    f = R(1.182509139382060e-21) * nHj * w14;
  cool += f;
  // H2-H2
  cool += w24 * kx::pow(R(10), R(-23.962112) + R(2.09433740) * l3
                        - R(0.77151436) * l2 + R(0.43693353) * l3_
                        - R(0.14913216) * l4 - R(0.033638326) * l5) * nH2;
  // H2-e
  if (T <= R(5e2))
    f = kx::pow(R(10), kx::fmin(R(-21.928796) + R(16.815730) * l3
                                + R(96.743155) * l2 + R(343.19180) * l3_
                                + R(734.71651) * l4 + R(983.67576) * l5
                                + R(801.81247) * l6 + R(364.14446) * l7
                                + R(70.609154) * l8, R(30))) * ne;
  else
    f = kx::pow(R(10), R(-22.921189) + R(1.6802758) * l3 + R(0.93310622) * l2
                + R(4.0406627) * l3_ - R(4.7274036) * l4 - R(8.8077017) * l5
                + R(8.9167183) * l6 + R(6.4380698) * l7
                - R(6.3701156) * l8) * ne;
  cool += f * w24;
  // H2-He
  if (T > R(1e1) && T <= R(1e4))
    f = kx::pow(R(10), R(-23.689237) + R(2.1892372) * l3 - R(0.81520438) * l2
                + R(0.29036281) * l3_ - R(0.16596184) * l4
                + R(0.19191375) * l5) * nHe;
  else
    f = R(1.002560385050777e-22) * nHe * w14;
  cool += f;
  if (cool <= R(0)) return R(0);
  // LTE limit
  R HDL;
  if (T < R(2e3)) {
    const R HDLR = R(9.5e-22) * kx::pow(T3, R(3.76))
                       / (R(1) + R(0.12) * kx::pow(T3, R(2.1)))
                       * kx::exp(-kx::pow(R(0.13) / T3, R(3)))
                   + R(3e-24) * kx::exp(R(-0.51) / T3);
    const R HDLV = R(6.7e-19) * kx::exp(R(-5.86) / T3)
                   + R(1.6e-18) * kx::exp(R(-11.7) / T3);
    HDL = HDLR + HDLV;
  } else if (T <= R(1e4)) {
    HDL = kx::pow(R(10), R(-20.584225) + R(5.0194035) * l3 - R(1.5738805) * l2
                  - R(4.7155769) * l3_ + R(2.4714161) * l4
                  + R(5.4710750) * l5 - R(3.9467356) * l6
                  - R(2.2148338) * l7 + R(1.8161874) * l8);
  } else {
    const R d = R(1) / (R(1) + kx::exp(kx::fmin((T - R(3e4)) * R(2e-4),
                                                   R(3e2))));
    HDL = R(5.531333679406485e-19) * d;
  }
  if (HDL == R(0)) return R(0);
  return nH2 / (R(1) / HDL + R(1) / cool);
}

// optically thick H2 line cooling factor (Ripamonti & Abel 2004)
template <class R>
KX_HD inline R h2_opacity_ripamonti(R ntot) {
  return kx::fmin(R(1), kx::pow(kx::fmax(R(1.25e-10) * ntot, R(1e-40)),
                                R(-0.45)));
}

// continuum cooling [erg cm^-3 s^-1]: 4 sigma T^4 kappa_P rho, times the
// escape fraction min(1, tau^-2) over a Jeans length
template <class R>
KX_HD inline R cooling_continuum(R T, R rho_in, R ntot) {
  R kgas = R(0);
  if (!(rho_in < R(1e-12)))
    kgas = kx::pow(R(10), R(1.000042) * kx::log10(kx::fmin(rho_in, R(0.5)))
                              + R(2.14989));
  const R rho = kx::fmax(rho_in, R(1e-40));
  const R mu = rho / kx::fmax(ntot, R(1e-40)) / R(phys::p_mass);
  const R lj = kx::sqrt(R(phys::pi) * R(phys::kB) * T / rho
                        / R(phys::p_mass) / R(phys::gravity) / mu);
  const R tau = lj * kgas * rho + R(1e-40);
  const R beta = kx::fmin(R(1), R(1) / (tau * tau));
  return R(4) * R(phys::stefboltz) * T * T * T * T * kgas * rho * beta;
}

// thermalized fraction 1 / (1 + ncr / n_H) of the H2 formation heat (HM79
// critical density; dd the H nuclei density)
template <class R>
KX_HD inline R h2_heat_fraction(R T, R nH, R nH2, R dd) {
  const R a = R(4e2) / T;
  const R ncr = R(1e6) / kx::sqrt(T)
                / (R(1.6) * kx::exp(-(a * a)) * nH / dd
                   + R(1.4) * kx::exp(R(-1.2e4) / (T + R(1.2e3))) * nH2 / dd);
  return R(1) / (R(1) + ncr / dd);
}

// vibrational heat capacity C_vib / k = x^2 e^x / (e^x - 1)^2 of a
// diatomic, x = theta_vib / T
template <class R>
KX_HD inline R vib_heat_capacity(R x) {
  const R e = kx::exp(x);
  return x * x * e / ((e - R(1)) * (e - R(1)));
}

// rotational heat capacity C_rot / k of a diatomic with rotational
// constant brot [K]: the variance of E_j / T over the 11 lowest levels
template <class R>
KX_HD inline R gamma_rot(R T, R brot) {
  const R ib = brot / kx::fmax(T, R(10));
  R z = R(0), e1 = R(0), e2 = R(0);
  for (int j = 0; j <= 10; ++j) {
    const R E = R(j * (j + 1)) * ib;
    const R g = R(2 * j + 1) * kx::exp(-E);
    z += g;
    e1 += g * E;
    e2 += g * E * E;
  }
  const R m = e1 / z;
  return e2 / z - m * m;
}

// the same for H2 as a mixture of para (even J) and ortho (odd J, weight
// 3) with ortho fraction a, rotational constant 85.4 K
template <class R>
KX_HD inline R gamma_rot_ortho_para(R T, R a) {
  const R ib = R(85.4) / kx::fmax(T, R(10));
  R zp = R(0), p1 = R(0), p2 = R(0), zo = R(0), o1 = R(0), o2 = R(0);
  for (int j = 0; j <= 10; j += 2) {
    const int j1 = j + 1;
    const R Ep = R(j * (j + 1)) * ib;
    const R Eo = R(j1 * (j1 + 1)) * ib;
    const R gp = R(2 * j + 1) * kx::exp(-Ep);
    const R go = R(3 * (2 * j1 + 1)) * kx::exp(-Eo);
    zp += gp;
    p1 += gp * Ep;
    p2 += gp * Ep * Ep;
    zo += go;
    o1 += go * Eo;
    o2 += go * Eo * Eo;
  }
  const R b = R(1) - a;
  const R mp = p1 / zp, mo = o1 / zo;
  return b * (p2 / zp - mp * mp) + a * (o2 / zo - mo * mo);
}

// H2 collisional dissociation by H (Martin, Schwarz & Mandy 1996):
// collision-induced dissociation plus dissociative tunnelling, each a
// density interpolation between low- and high-density limits; n_H the H
// nuclei density, the fit evaluated at min(T, 4.5e4 K)
template <class R>
KX_HD inline R dissH2_Martin96(R Tin, R n_H) {
  const R T = kx::fmin(Tin, R(4.5e4));
  const R lT = kx::log10(T);
  const R iT = R(1) / T;
  const R lT2 = lT * lT, lT3 = lT2 * lT;
  R k = R(0);
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    // the fit coefficients of the two channels
    const R c0 = ch == 0 ? R(-178.4239) : R(-142.7664);
    const R c1 = ch == 0 ? R(-68.42243) : R(42.70741);
    const R c2 = ch == 0 ? R(43.20243) : R(-2.027365);
    const R c3 = ch == 0 ? R(-4.633167) : R(-0.2582097);
    const R c4 = ch == 0 ? R(69.70086) : R(21.36094);
    const R c5 = ch == 0 ? R(40870.38) : R(27535.31);
    const R c6 = ch == 0 ? R(-23705.70) : R(-21467.79);
    const R c7 = ch == 0 ? R(128.8953) : R(60.34928);
    const R c8 = ch == 0 ? R(-53.91334) : R(-27.43096);
    const R c9 = ch == 0 ? R(5.315517) : R(2.676150);
    const R c10 = ch == 0 ? R(-19.73427) : R(-11.28215);
    const R c11 = ch == 0 ? R(16780.95) : R(14254.55);
    const R c12 = ch == 0 ? R(-25786.11) : R(-23125.20);
    const R c13 = ch == 0 ? R(14.82123) : R(9.305564);
    const R c14 = ch == 0 ? R(-4.890915) : R(-2.464009);
    const R c15 = ch == 0 ? R(0.4749030) : R(0.1985955);
    const R c16 = ch == 0 ? R(-133.8283) : R(743.0600);
    const R c17 = ch == 0 ? R(-1.164408) : R(-1.174242);
    const R c18 = ch == 0 ? R(0.8227443) : R(0.7502286);
    const R c19 = ch == 0 ? R(0.5864073) : R(0.2358848);
    const R c20 = ch == 0 ? R(-2.056313) : R(2.937507);
    const R kh1 = c0 + c1 * lT + c2 * lT2 + c3 * lT3
                  + c4 * kx::log10(R(1) + c5 * iT);
    const R kh2 = c6 * iT;
    const R kl1 = c7 + c8 * lT + c9 * lT2 + c10 * kx::log10(R(1) + c11 * iT);
    const R kl2 = c12 * iT;
    const R nc1 = c13 + c14 * lT + c15 * lT2 + c16 * iT;
    const R nc2 = c17 + nc1;
    const R pp = c18 + c19 * kx::exp(-T / R(1.850e3))
                 + c20 * kx::exp(-T / R(4.40e2));
    const R n_c1 = kx::pow(R(10), nc1);
    const R n_c2 = kx::pow(R(10), nc2);
    const R lk = kh1 - (kh1 - kl1) / (R(1) + kx::pow(n_H / n_c1, pp)) + kh2
                 - (kh2 - kl2) / (R(1) + kx::pow(n_H / n_c2, pp));
    k += kx::pow(R(10), lk);
  }
  return k;
}

// Solomon-process H2 photodissociation rate [s^-1] from the intensity at
// 12.87 eV (flux_eV in eV cm^-2 s^-1 Hz^-1 sr^-1): 1.38e9 J [erg units]
template <class R>
KX_HD inline R H2_solomonLW(R flux_eV) {
  return R(1.38e9) * flux_eV * R(1.602176634e-12);
}

// Weingartner & Draine 2001 grain-assisted recombination [cm^3 s^-1]:
// k = s / (1 + a psi^b (1 + c T^d psi^(e + f ln T))), s = pre d2g / 0.01,
// psi = G0 sqrt(T) / n_e.  No electrons: 0 (the psi -> inf limit);
// psi <= 0 (no field): s (the psi -> 0 limit)
template <class R>
KX_HD inline R wd01_recombination(R pre, R a, R b, R c, R d, R e, R f, R T,
                                  R ne, R G0, R d2g) {
  if (ne < R(1e-20)) return R(0);
  const R s = pre * (d2g / R(1e-2));
  const R psi = G0 * kx::sqrt(T) / ne;
  if (psi <= R(0)) return s;
  return s / (R(1) + a * kx::pow(psi, b)
                         * (R(1) + c * kx::pow(T, d)
                                       * kx::pow(psi, e + f * kx::log(T))));
}
template <class R>
KX_HD inline R H_recombination_on_dust(R T, R ne, R G0, R d2g) {
  return wd01_recombination(R(1.225e-13), R(8.074e-6), R(1.378), R(5.087e2),
                            R(0.01586), R(-0.4723), R(-1.102e-5), T, ne, G0,
                            d2g);
}
template <class R>
KX_HD inline R He_recombination_on_dust(R T, R ne, R G0, R d2g) {
  return wd01_recombination(R(5.572e-14), R(3.185e-7), R(1.512), R(5.115e3),
                            R(3.903e-7), R(-0.4956), R(-5.494e-7), T, ne, G0,
                            d2g);
}

// H2 self-shielding factor, Wolcott-Green & Haiman 2011: the H2 column
// N(H2) = 2 x 1.87e21 (n(H2) / 1e3)^(2/3) (the default number-to-column
// estimate), x = N(H2) / 5e14, b5 the thermal Doppler width of an H atom
// (proton + electron mass) [km/s],
// f = 0.965 / (1 + x/b5)^1.1 + 0.035 / sqrt(1 + x) exp(-8.5e-4 sqrt(1 + x))
template <class R>
KX_HD inline R fshield_wg11(R T, R nH2) {
  const R N = R(1.87e21) * kx::pow(kx::fmax(kx::fmax(R(1e-40), nH2),
                                            R(1e-40)) * R(1e-3),
                                   R(2.0) / R(3.0));
  const R xN = R(2) * N * R(2e-15);
  const R H_mass = R(1.67262192369e-24) + R(9.1093837015e-28);
  const R b5 = kx::sqrt(R(1.380649e-16) * T / H_mass) * R(1e-5);
  return R(0.965) / kx::pow(R(1) + xN / b5, R(1.1)) +
         R(0.035) / kx::sqrt(R(1) + xN) *
             kx::exp(R(-8.5e-4) * kx::sqrt(R(1) + xN));
}

// ---- the legacy rate-database fits, functions of Tgas ---------------------

// collisional ionization, Voronov 1997: U = dE / (kB T),
// k = A (1 + P sqrt(U)) U^K exp(-U) / (X + U)
template <class R>
KX_HD inline R colion_v96(R T, R dE, R P, R A, R X, R K) {
  const R U = dE / (T * R(8.617333262145179e-05));
  return A * (R(1) + P * kx::sqrt(U)) * kx::pow(U, K) * kx::exp(-U) / (X + U);
}
// radiative recombination, a power law a (1e4 / T)^b
template <class R>
KX_HD inline R recV96(R T, R a, R b) {
  return a * kx::pow(R(1e4) / T, b);
}
// radiative recombination, Verner & Ferland 1996:
// r1 / (t (1 + t)^(1 - r2) (1 + sqrt(T / r4))^(1 + r2)), t = sqrt(T / r3)
template <class R>
KX_HD inline R recNewV96(R T, R r1, R r2, R r3, R r4) {
  const R t = kx::sqrt(T / r3);
  return r1 / (t * kx::pow(t + R(1), R(1) - r2) *
               kx::pow(R(1) + kx::sqrt(T / r4), R(1) + r2));
}
// radiative recombination of Fe ions: r1 / t^(r2 + r3 + log10 t),
// t = sqrt(T / 1e4)
template <class R>
KX_HD inline R recFeV96(R T, R r1, R r2, R r3) {
  const R t = kx::sqrt(T * R(1e-4));
  return r1 / kx::pow(t, r2 + r3 + kx::log10(t));
}

// ---- photo cross sections [cm^2] of photon energy E [eV], host only -----

// Verner et al. 1996 analytic fit of a ground-state photoionization cross
// section: x = E/E0 - y0, y = sqrt(x^2 + y1^2),
// sigma = 1e-18 s0 ((x-1)^2 + yw^2) y^(P/2 - 5.5) (1 + sqrt(y/ya))^-P
inline double sigma_v96(double E, double E0, double s0, double ya, double P,
                        double yw, double y0, double y1) {
  const double x = E / E0 - y0;
  const double y = std::sqrt(x * x + y1 * y1);
  return 1e-18 * s0 * ((x - 1.0) * (x - 1.0) + yw * yw)
         * std::pow(y, 0.5 * P - 5.5) * std::pow(1.0 + std::sqrt(y / ya), -P);
}

// Verner & Yakovlev 1995 fit of one shell, 0 below its threshold Eth:
// y = E/E0, sigma = 1e-18 s0 ((y-1)^2 + yw^2) y^(P/2 - 5.5 - l)
// (1 + sqrt(y/ya))^-P
inline double sigma_vy95(double E, double Eth, double E0, double s0,
                         double ya, double P, double yw, double l) {
  if (E < Eth) return 0.0;
  const double y = E / E0;
  return 1e-18 * s0 * ((y - 1.0) * (y - 1.0) + yw * yw)
         * std::pow(y, 0.5 * P - 5.5 - l) * std::pow(1.0 + std::sqrt(y / ya), -P);
}

// a tabulated cross section (tE ascending), linear between the points and
// zero outside [tE[0], tE[n-1]]
inline double xsec_interp(double E, const double* tE, const double* tS,
                          int n) {
  if (E < tE[0] || E > tE[n - 1]) return 0.0;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int m = (lo + hi) / 2;
    if (tE[m] <= E) lo = m;
    else hi = m;
  }
  const double f = (E - tE[lo]) / (tE[lo + 1] - tE[lo]);
  return tS[lo] + f * (tS[lo + 1] - tS[lo]);
}

// the effective H2 Lyman-Werner cross section: the Lyman and Werner fits of
// para (v=0) and ortho (v=1) H2, each 1e-18 10^(polynomial in E) on its
// window, weighted by the ortho:para ratio 3
inline double H2_sigmaLW(double E) {
  double sL0 = 0.0, sW0 = 0.0, sL1 = 0.0, sW1 = 0.0;
  if (E > 14.675 && E < 16.820)
    sL0 = 1e-18 * std::pow(10.0, 15.1289 - 1.05139 * E);
  else if (E > 16.820 && E < 17.6)
    sL0 = 1e-18 * std::pow(10.0, -31.41 + 1.8042e-2 * E * E * E
                                     - 4.2339e-5 * std::pow(E, 5.0));
  if (E > 14.675 && E < 17.7)
    sW0 = 1e-18 * std::pow(10.0, 13.5311 - 0.9182618 * E);
  if (E > 14.159 && E < 15.302)
    sL1 = 1e-18 * std::pow(10.0, 12.0218406 - 0.819429 * E);
  else if (E > 15.302 && E < 17.2)
    sL1 = 1e-18 * std::pow(10.0, 16.04644 - 1.082438 * E);
  if (E > 14.159 && E < 17.2)
    sW1 = 1e-18 * std::pow(10.0, 12.87367 - 0.85088597 * E);
  const double para = 1.0 / (3.0 + 1.0);
  return para * (sL0 + sW0) + (1.0 - para) * (sL1 + sW1);
}

}  // namespace kx
