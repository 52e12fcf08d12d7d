// Photon bins for the photo reactions of a generated network (KROME-X's
// photo module, runtime/core/photo.hpp: PhotoModule::init 135-253,
// photo_rates 357-375).
//
// The rate of photo reaction q in a cell is the bin sum
//   k_q = sum_b sigma_qb Phi_b
// of the cell's photon field Phi_b [photons cm^-2 s^-1] (held fixed during
// a solve) with the bin-averaged cross sections
//   sigma_qb = Int_{bin b ^ [Emin_q, Emax_q]} sigma_q(E) E^-alpha dE
//              / Int_{bin b} E^-alpha dE,
// the photons inside a bin distributed as dN/dE ~ E^-alpha (alpha 1: the
// same number per ln E).
//
//   photo_bin_xsec  the table sigma[NPHOTO][B] of a set of bins, on the
//                   host, once per bin set: the overlap of each bin with
//                   the reaction's window by a 64-cell midpoint rule, in
//                   ln E (alpha != 0) or in E (alpha 0), each cell weighted
//                   by its exact Int E^-alpha dE, so a constant cross
//                   section averages exactly
//   photo_rows_kernel  once per solve, on the device: each cell's kernel
//                   parameter row, the caller's NPAR parameters followed by
//                   its NPHOTO bin sums (one thread per cell, the block's
//                   fields staged through LDS; HBM-bound: 8 (B + NROW) bytes
//                   per cell plus the caller's row)
//
// The generated rates read the sums from the row (Params::ph), so the
// solve kernels see a photo rate like any rate of the parameters: a
// constant of the cell.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

namespace kx {

// Int_a^c E^-alpha dE
inline double photo_weight(double alpha, double a, double c) {
  if (alpha == 1.0) return std::log(c / a);
  if (alpha == 0.0) return c - a;
  const double q = 1.0 - alpha;
  return (std::pow(c, q) - std::pow(a, q)) / q;
}

// sigma[q B + b] of the bins [Elo[b], Ehi[b]] (0 < Elo < Ehi, checked by
// the caller), followed by sigmaE[NPHOTO B] at sigma + NPHOTO B: the same
// average of sigma(E) times the photoelectron's energy max(E - Emin_q, 0)
// [cm^2 eV] (photo.hpp:227-246), the photoheating's table
template <class Net>
void photo_bin_xsec(int B, const double* Elo, const double* Ehi,
                    double alpha, std::vector<double>& sigma) {
  constexpr int QUAD = 64;
  const size_t NB = (size_t)(Net::NPHOTO > 0 ? Net::NPHOTO : 1) * B;
  sigma.assign(2 * NB, 0.0);
  for (int q = 0; q < Net::NPHOTO; ++q) {
    for (int b = 0; b < B; ++b) {
      const double a = std::fmax(Elo[b], Net::photo_Emin[q]);
      const double c = std::fmin(Ehi[b], Net::photo_Emax[q]);
      if (!(c > a)) continue;
      const double dE = (c - a) / QUAD;
      const double dl = std::log(c / a) / QUAD;
      double s = 0.0, se = 0.0;
      for (int i = 0; i < QUAD; ++i) {
        double E, w;
        if (alpha == 0.0) {
          E = a + (i + 0.5) * dE;
          w = dE;
        } else {
          E = a * std::exp((i + 0.5) * dl);
          const double lo = a * std::exp(i * dl);
          const double hi = i + 1 == QUAD ? c : a * std::exp((i + 1) * dl);
          w = photo_weight(alpha, lo, hi);
        }
        const double x = Net::photo_xsec(q, E) * w;
        s += x;
        se += x * std::fmax(E - Net::photo_Emin[q], 0.0);
      }
      const double wb = photo_weight(alpha, Elo[b], Ehi[b]);
      sigma[(size_t)q * B + b] = s / wb;
      sigma[NB + (size_t)q * B + b] = se / wb;
    }
  }
}

// out[c NROW ..]: the parameter row par[c pst ..] (pst 0: one row for
// every cell), then sum_b sigma[q B + b] Phi[c B + b] for each photo
// reaction q (0 without a field: B 0 or Phi null), and with PHOTO heating
// the same sums over sigmaE (at sigma + NPHOTO B).  With a field, a block
// of blockDim.x cells first copies its fields, one contiguous stretch of
// HBM, into LDS with coalesced loads (rows padded to B + 1 doubles, so the
// lanes' reads of one bin fall in different banks); each thread then sums
// its own row from LDS, b ascending per reaction.  lds_rows 0: no LDS
// stage, each thread reads its row from HBM (fields too wide for LDS).
template <class Net>
__global__ void __launch_bounds__(256)
photo_rows_kernel(int rows, const double* __restrict__ par, int pst,
                  const double* __restrict__ Phi, int B,
                  const double* __restrict__ sigma, double* __restrict__ out,
                  int lds_rows) {
  extern __shared__ double tile[];
  const long c0 = (long)blockIdx.x * blockDim.x;
  const int nrow = (int)((rows - c0) < blockDim.x ? (rows - c0) : blockDim.x);
  const bool staged = Phi && lds_rows > 0;
  if (staged) {
    const double* src = Phi + c0 * B;
    for (int i = threadIdx.x; i < nrow * B; i += blockDim.x) {
      const int r = i / B;
      tile[r * (B + 1) + (i - r * B)] = src[i];
    }
  }
  __syncthreads();
  const long c = c0 + threadIdx.x;
  if (c >= rows) return;
  const double* p = par + c * pst;
  double* o = out + c * Net::NROW;
  for (int i = 0; i < Net::NPAR; ++i) o[i] = p[i];
  constexpr int NQ = Net::NPHOTO * (Net::PHOTO_HEAT ? 2 : 1);
  double s[NQ > 0 ? NQ : 1];
  for (int q = 0; q < NQ; ++q) s[q] = 0.0;
  if (Phi) {
    const double* f = staged ? tile + threadIdx.x * (B + 1) : Phi + c * B;
    for (int b = 0; b < B; ++b) {
      const double fb = f[b];
      for (int q = 0; q < NQ; ++q) s[q] += sigma[q * B + b] * fb;
    }
  }
  for (int q = 0; q < NQ; ++q) o[Net::NPAR + q] = s[q];
}

// the launch shape of photo_rows_kernel for B bins: cells per block and
// dynamic LDS bytes (64 KiB at most; 0: unstaged)
inline void photo_rows_shape(int B, int& block, size_t& lds) {
  const size_t cap = 64 * 1024;
  block = 256;
  for (; block >= 64; block /= 2) {
    lds = (size_t)block * (B + 1) * sizeof(double);
    if (lds <= cap) return;
  }
  block = 256;
  lds = 0;
}

}  // namespace kx
