// Mass fractions around a batched solve on gfx950: KROME-X's x_to_n_,
// x_sum_ and n_to_x_out_ (runtime/core/solve_cell.hpp:2210-2232) as the
// device gather / scatter of its mass-fraction entry points
// (runtime/core/solve_cells.hpp:176-200, solve_kokkos.hpp:1450-1458,
// 1519-1528).  n = rho x / m on the way in (0 for a species without a
// mass), x = m n / rho on the way out for the cells that succeeded,
// rescaled to the input sum with renormalize_x.  One thread per cell; the
// caller's x in either layout, v[c cs + i ss]; n species-major.
#pragma once
#include <hip/hip_runtime.h>

#include "kx_physics.h"

namespace kx {

template <class Net>
__global__ void __launch_bounds__(256)
x_to_n_kernel(int ncell, const double* __restrict__ x, long cs, long ss,
              const double* __restrict__ rho, double* __restrict__ n,
              double* __restrict__ xin) {
  constexpr int NS = Net::NSPEC;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double r = rho[c];
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double m = Net::mass_g(i);
    const double v = x[c * cs + i * ss];
    if (m > 0.0) s += v;
    n[(long)i * ncell + c] = m > 0.0 ? r * v / m : 0.0;
  }
  xin[c] = s;
}

// the cells with flag 0 (and, where dt is given, dt > 0) back to x
template <class Net>
__global__ void __launch_bounds__(256)
n_to_x_kernel(int ncell, const double* __restrict__ n,
              double* __restrict__ x, long cs, long ss,
              const double* __restrict__ rho,
              const double* __restrict__ xin, const int* __restrict__ flags,
              const double* __restrict__ dt, int renorm) {
  constexpr int NS = Net::NSPEC;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell || flags[c] != 0) return;
  if (dt != nullptr && !(dt[c] > 0.0)) return;
  const double r = rho[c];
  double v[NS], s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    v[i] = Net::mass_g(i) * n[(long)i * ncell + c] / r;
    if (Net::mass_g(i) > 0.0) s += v[i];
  }
  if (renorm && s > 0.0) {
    const double f = xin[c] / s;
#pragma unroll
    for (int i = 0; i < NS; ++i) v[i] *= f;
  }
#pragma unroll
  for (int i = 0; i < NS; ++i) x[c * cs + i * ss] = v[i];
}

}  // namespace kx
