// The query functions of the C ABI on gfx950: rate coefficients, reaction
// fluxes and the thermal balance of given states, as the solver sees them
// (KROME-X's get_rates, get_fluxes, get_thermal: runtime/core/query.hpp:
// 37-108, krome.cpp:106-133).  One thread per cell; the rates at the rate
// temperature min(max(Tgas, Tcmb), 1e9), the thermal terms at
// min(max(Tgas, 2.73), 1e9).  A cell whose REQUIRED parameter is unset
// gets FLAG_PARAM_UNSET and no outputs.
#pragma once
#include <hip/hip_runtime.h>

#include "kx_plain.h"

namespace kx {

enum QueryWhat { Q_RATES = 0, Q_FLUXES = 1, Q_THERMAL = 2 };

// n[c NSPEC + i], T[c], par rows (stride pst); outputs per cell:
// Q_RATES / Q_FLUXES out[c NREA + r]; Q_THERMAL out[c 3] = heat, cool,
// dTdt, heats[c NHEAT + h], cools[c NCOOL + q]
template <class Net>
__global__ void __launch_bounds__(128)
query_kernel(int ncell, const double* __restrict__ n,
             const double* __restrict__ T, const double* __restrict__ par,
             int pst, int what, double* __restrict__ out,
             double* __restrict__ heats, double* __restrict__ cools,
             int* __restrict__ flags) {
  constexpr int NS = Net::NSPEC;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double* row = par + (long)c * pst;
  if (Net::params_missing(row) >= 0) {
    flags[c] = FLAG_PARAM_UNSET;
    return;
  }
  flags[c] = 0;
  const typename Net::Params p = Net::params_from_row(row);
  double y[NS];
  for (int i = 0; i < NS; ++i) y[i] = n[(long)c * NS + i];
  double k[Net::NREA];
  const double Tr = ::fmin(::fmax(T[c], p.Tcmb), 1e9);
  Net::rates(Tr, y, p, k);
  if (what == Q_RATES) {
    for (int r = 0; r < Net::NREA; ++r) out[(long)c * Net::NREA + r] = k[r];
    return;
  }
  if (what == Q_FLUXES) {
    double fl[Net::NREA];
    Net::fluxes(y, k, fl);
    for (int r = 0; r < Net::NREA; ++r) out[(long)c * Net::NREA + r] = fl[r];
    return;
  }
  double heat = 0.0, cool = 0.0, dTdt = 0.0;
  if constexpr (Net::THERMAL) {
    const double Tth = ::fmin(::fmax(T[c], 2.73), 1e9);
    double H[Net::NHEAT > 0 ? Net::NHEAT : 1];
    double C[Net::NCOOL > 0 ? Net::NCOOL : 1];
    Net::thermal_terms(Tth, y, k, p, H, C);
    for (int i = 0; i < Net::NHEAT; ++i) {
      heat += H[i];
      heats[(long)c * Net::NHEAT + i] = H[i];
    }
    for (int i = 0; i < Net::NCOOL; ++i) {
      cool += C[i];
      cools[(long)c * Net::NCOOL + i] = C[i];
    }
    dTdt = (Net::gamma_index(Tth, y, p) - 1.0) * (heat - cool) /
           (1.380649e-16 * Net::ntot(y));
  }
  out[(long)c * 3] = heat;
  out[(long)c * 3 + 1] = cool;
  out[(long)c * 3 + 2] = dTdt;
}

}  // namespace kx
