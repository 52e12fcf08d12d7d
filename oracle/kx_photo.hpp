// The oracle's photon bins (test infrastructure only): a CPU restatement of
// KROME-X's PhotoModule::init and photo_rates (runtime/core/photo.hpp:
// 135-253 and 357-375) over the generated network's cross sections, used to
// build the parameter rows of photo networks for the oracle's solves and
// to check the library's bin table.
//
//   sigma[q B + b] = (sum over 64 cells of [a, c] = bin b ^ [Emin_q,
//                     Emax_q], equal in ln E, or in E when alpha is 0, of
//                     sigma_q(E_mid) Int_cell E^-alpha dE)
//                    / Int_bin E^-alpha dE            (photo.hpp:186-250)
//   row[NPAR + q]  = sum_b sigma[q B + b] Phi_b, b ascending
//                                                     (photo.hpp:357-375)
//   sigmaE[q B + b] = the same average of sigma(E) max(E - Emin_q, 0),
//   and with PHOTO heating row[NPAR + NPHOTO + q] = sum_b sigmaE Phi_b
//                                                     (photo.hpp:227-246)
#pragma once
#include <cmath>
#include <vector>

namespace kxo_photo {

// Int_lo^hi E^-alpha dE (the generated photo_wint_)
inline double wint(double alpha, double lo, double hi) {
  if (alpha == 1.0) return std::log(hi / lo);
  if (alpha == 0.0) return hi - lo;
  return (std::pow(hi, 1.0 - alpha) - std::pow(lo, 1.0 - alpha)) /
         (1.0 - alpha);
}

// sigma[NPHOTO B] then sigmaE[NPHOTO B]
template <class Net>
void bins(int B, const double* Elo, const double* Ehi, double alpha,
          double* sigma) {
  const int nq = 64;
  const size_t nb = static_cast<size_t>(Net::NPHOTO) * B;
  for (int q = 0; q < Net::NPHOTO; ++q)
    for (int b = 0; b < B; ++b) {
      double* out = sigma + static_cast<size_t>(q) * B + b;
      double* outE = out + nb;
      *out = 0.0;
      *outE = 0.0;
      const double lo = Elo[b] > Net::photo_Emin[q] ? Elo[b]
                                                    : Net::photo_Emin[q];
      const double hi = Ehi[b] < Net::photo_Emax[q] ? Ehi[b]
                                                    : Net::photo_Emax[q];
      if (hi <= lo) continue;
      const double step = (hi - lo) / nq;
      const double lstep = std::log(hi / lo) / nq;
      double acc = 0.0, accE = 0.0;
      for (int i = 0; i < nq; ++i) {
        double E, x;
        if (alpha == 0.0) {
          E = lo + (i + 0.5) * step;
          x = Net::photo_xsec(q, E) * step;
        } else {
          const double e0 = lo * std::exp(i * lstep);
          const double e1 = (i == nq - 1) ? hi : lo * std::exp((i + 1) * lstep);
          E = lo * std::exp((i + 0.5) * lstep);
          x = Net::photo_xsec(q, E) * wint(alpha, e0, e1);
        }
        acc += x;
        const double Ee = E - Net::photo_Emin[q];
        accE += x * (Ee > 0.0 ? Ee : 0.0);
      }
      *out = acc / wint(alpha, Elo[b], Ehi[b]);
      *outE = accE / wint(alpha, Elo[b], Ehi[b]);
    }
}

// rows out[c NROW]: par[c pst] then the bin sums of Phi[c B]
template <class Net>
void rows(int ncell, const double* par, int pst, const double* Phi, int B,
          const double* sigma, double* out) {
  for (int c = 0; c < ncell; ++c) {
    double* o = out + static_cast<size_t>(c) * Net::NROW;
    for (int i = 0; i < Net::NPAR; ++i)
      o[i] = par[static_cast<size_t>(c) * pst + i];
    const int nq = Net::NPHOTO * (Net::PHOTO_HEAT ? 2 : 1);
    for (int q = 0; q < nq; ++q) {
      double s = 0.0;
      for (int b = 0; b < B; ++b)
        s += sigma[static_cast<size_t>(q) * B + b] *
             Phi[static_cast<size_t>(c) * B + b];
      o[Net::NPAR + q] = s;
    }
  }
}

}  // namespace kxo_photo
