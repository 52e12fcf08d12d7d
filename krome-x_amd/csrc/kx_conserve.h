// Element totals and consistent advection on gfx950: KROME-X's
// conservation API with one thread per cell,
//   totals, totals_x      runtime/core/query.hpp:152-170
//   residual              runtime/core/solve_cell.hpp:1336-1353
//   project               solve_cell.hpp:1355-1432
//   post-solve mode 2     solve_cell.hpp:1469-1489
//   to_n, to_x            solve_cell.hpp:1491-1562
//   the cells' flags      runtime/core/solve_cells.hpp:216-242,
//                         solve_kokkos.hpp:1787-1822 (the lowest failing
//                         cell's flag)
// Every loop has a compile-time trip count and is unrolled, so the cell's
// species, the small normal-equation system (at most NELEM + 1 rows) and
// its pivot row exchanges stay in registers; a row exchange is a
// conditional swap over the rows below the pivot.  The arithmetic is the
// reference's, operation by operation and without contraction (IEEE add,
// multiply, divide, fabs, fmax): the results are those of the reference
// built without contraction, bit for bit.
//
// The species the reference's projections hold fixed are the opaque ones
// and those of freeze directives (solve_cell.hpp:1440-1449): netgen.py
// has no freeze directives; the opaque species (cloud's grains: no atoms,
// not the electron) are held, and the mass-fraction form also holds the
// species without a mass.  `fixed` is a bit set over the species.
//
// Memory: a cell's row is read once and written once, v[c cs + i ss]
// (species-major: cs 1, ss ld, coalesced; cell-major: cs NSPEC, ss 1, a
// wave reads a contiguous 64 NSPEC doubles).  HBM-bound: per cell 2 NSPEC
// doubles, the reference row (NELEM doubles, unless shared) and a flag.
#pragma once
#include <hip/hip_runtime.h>

#include "kx_plain.h"

namespace kx {

constexpr int FLAG_CONSERVE_UNREACHED = -112;
constexpr int FLAG_CONSERVE_LARGE = 2;

template <class Net>
struct Conserve {
  static constexpr int NS = Net::NSPEC;
  static constexpr int NE = Net::NELEM;
  // a bit per species
  struct Mask {
    unsigned long long w[(NS + 63) / 64] = {};
    KX_HD bool test(int i) const { return (w[i >> 6] >> (i & 63)) & 1ull; }
    KX_HD void set(int i) { w[i >> 6] |= 1ull << (i & 63); }
  };
  // the opaque species: no atoms of any element, and not the electron
  KX_HD static bool opaque(int i) {
    if (Net::HAS_E && i == Net::I_E) return false;
    for (int e = 0; e < NE; ++e)
      if (Net::natoms(i, e) != 0) return false;
    return true;
  }
  KX_HD static Mask opaque_mask() {
    Mask m;
    for (int i = 0; i < NS; ++i)
      if (opaque(i)) m.set(i);
    return m;
  }
  static_assert(NE >= 1, "a network without elements has nothing to keep");
  // the species loops unroll fully up to 64 species; a larger network's
  // (cloud: 452) stay loops
  static constexpr int UNR = NS <= 64 ? NS : 1;

  // conserved row k at species i: atoms of element k (k < NE), charge (NE)
  KX_HD static double row(int k, int i) {
    return k < NE ? double(Net::natoms(i, k)) : double(Net::charge(i));
  }
  KX_HD static bool is_fixed(const Mask& fixed, int i) {
    return fixed.test(i);
  }

  KX_HD static void totals(const double* n, double* tot) {
#pragma clang fp contract(off)
#pragma unroll
    for (int e = 0; e < NE; ++e) tot[e] = 0.0;
#pragma unroll UNR
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int e = 0; e < NE; ++e) tot[e] += double(Net::natoms(i, e)) * n[i];
  }

  KX_HD static void totals_x(const double* x, double* tot) {
#pragma clang fp contract(off)
#pragma unroll
    for (int e = 0; e < NE; ++e) tot[e] = 0.0;
#pragma unroll UNR
    for (int i = 0; i < NS; ++i) {
      const double m = Net::mass_g(i);
      if (!(m > 0.0)) continue;
      const double ni = x[i] / m;
#pragma unroll
      for (int e = 0; e < NE; ++e) tot[e] += double(Net::natoms(i, e)) * ni;
    }
  }

  // max over the NR rows of |target - A n| / (|target| + sum |A| n)
  template <int NR>
  KX_HD static double residual(const double* n, const double* target) {
#pragma clang fp contract(off)
    double worst = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      double t = 0.0, sc = ::fabs(target[k]);
#pragma unroll UNR
      for (int i = 0; i < NS; ++i) {
        const double a = row(k, i) * n[i];
        t += a;
        sc += ::fabs(a);
      }
      if (sc > 0.0) worst = ::fmax(worst, ::fabs(target[k] - t) / sc);
    }
    return worst;
  }

  // n moved onto the totals of the first NR rows: n_i (1 + sum_k A_ki
  // lambda_k) over the free species, (A W A^T) lambda = target - A n with
  // W = diag(n); Gaussian elimination with partial pivoting, a pivot below
  // 1e-14 of the largest diagonal skipped (lambda 0); repeated after
  // clipping negatives, at most 20 times, until the residual is below 1e-13
  template <int NR>
  KX_HD static double project(double* n, const double* target,
                              const Mask& fixed) {
#pragma clang fp contract(off)
    for (int it = 0; it < 20; ++it) {
      double r[NR], M[NR][NR], lam[NR];
      bool used[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        double t1 = 0.0;
#pragma unroll UNR
        for (int i = 0; i < NS; ++i) t1 += row(k, i) * n[i];
        r[k] = target[k] - t1;
#pragma unroll
        for (int l = 0; l < NR; ++l) {
          double s = 0.0;
#pragma unroll UNR
          for (int i = 0; i < NS; ++i)
            if (!is_fixed(fixed, i)) s += row(k, i) * row(l, i) * n[i];
          M[k][l] = s;
        }
      }
      double scale = 0.0;
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        lam[k] = 0.0;
        used[k] = false;
        scale = ::fmax(scale, ::fabs(M[k][k]));
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        int piv = -1;
        double best = 1e-14 * scale;
#pragma unroll
        for (int q = c; q < NR; ++q)
          if (::fabs(M[q][c]) > best) {
            best = ::fabs(M[q][c]);
            piv = q;
          }
        if (piv < 0) continue;
#pragma unroll
        for (int q = c + 1; q < NR; ++q)
          if (q == piv) {
#pragma unroll
            for (int l = 0; l < NR; ++l) {
              const double t = M[q][l];
              M[q][l] = M[c][l];
              M[c][l] = t;
            }
            const double t = r[q];
            r[q] = r[c];
            r[c] = t;
          }
        used[c] = true;
#pragma unroll
        for (int q = c + 1; q < NR; ++q) {
          const double f = M[q][c] / M[c][c];
#pragma unroll
          for (int l = c; l < NR; ++l) M[q][l] -= f * M[c][l];
          r[q] -= f * r[c];
        }
      }
#pragma unroll
      for (int c = NR - 1; c >= 0; --c) {
        if (!used[c]) continue;
        double s = r[c];
#pragma unroll
        for (int l = c + 1; l < NR; ++l) s -= M[c][l] * lam[l];
        lam[c] = s / M[c][c];
      }
      bool clipped = false;
#pragma unroll UNR
      for (int i = 0; i < NS; ++i) {
        if (is_fixed(fixed, i)) continue;
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < NR; ++k) d += row(k, i) * lam[k];
        n[i] += n[i] * d;
        if (n[i] < 0.0) {
          n[i] = 0.0;
          clipped = true;
        }
      }
      if (!clipped || residual<NR>(n, target) < 1e-13) break;
    }
    return residual<NR>(n, target);
  }

  // the conserved totals (elements, then the charge) of a cell's input
  KX_HD static void targets(const double* n0, double* target) {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k <= NE; ++k) {
      target[k] = 0.0;
#pragma unroll UNR
      for (int i = 0; i < NS; ++i) target[k] += row(k, i) * n0[i];
    }
  }

  // consistent advection of one cell onto the element totals ref[NE] and
  // zero charge: negatives clipped, the elements projected (the charge
  // kept), the electrons from neutrality, or the charge projected too when
  // the electrons cannot carry it; 0, FLAG_CONSERVE_UNREACHED (residual
  // above 1e-12) or FLAG_CONSERVE_LARGE (a species moved by over 50 %)
  KX_HD static int to_n(double* n, const double* ref, const Mask& fixed) {
#pragma clang fp contract(off)
    double n0[NS];
#pragma unroll UNR
    for (int i = 0; i < NS; ++i) {
      if (!is_fixed(fixed, i) && n[i] < 0.0) n[i] = 0.0;
      n0[i] = n[i];
    }
    double target[NE + 1];
#pragma unroll
    for (int k = 0; k < NE; ++k) target[k] = ref[k];
    target[NE] = 0.0;
    project<NE>(n, target, fixed);
    bool charge_done = false;
    if constexpr (Net::HAS_E) {
      if (!is_fixed(fixed, Net::I_E)) {
        double qi = 0.0;
#pragma unroll UNR
        for (int i = 0; i < NS; ++i)
          if (i != Net::I_E) qi += double(Net::charge(i)) * n[i];
        if (qi >= 0.0) {
          n[Net::I_E] = qi;
          charge_done = true;
        } else {
          n[Net::I_E] = 0.0;
        }
      }
    }
    if (!charge_done) project<NE + 1>(n, target, fixed);
    if (residual<NE + 1>(n, target) > 1e-12) return FLAG_CONSERVE_UNREACHED;
    int f = 0;
#pragma unroll UNR
    for (int i = 0; i < NS; ++i)
      if (n0[i] > 0.0 && ::fabs(n[i] - n0[i]) > 0.5 * n0[i])
        f = FLAG_CONSERVE_LARGE;
    return f;
  }

  // the same on mass fractions x with ref per gram: n = x / m, corrected,
  // x = m n; species without a mass held
  KX_HD static int to_x(double* x, const double* ref) {
#pragma clang fp contract(off)
    Mask fixed = opaque_mask();
    double n[NS];
#pragma unroll UNR
    for (int i = 0; i < NS; ++i) {
      const double m = Net::mass_g(i);
      if (!(m > 0.0)) fixed.set(i);
      n[i] = m > 0.0 ? x[i] / m : 0.0;
    }
    const int f = to_n(n, ref, fixed);
#pragma unroll UNR
    for (int i = 0; i < NS; ++i)
      if (!is_fixed(fixed, i)) x[i] = Net::mass_g(i) * n[i];
    return f;
  }
};

// the lowest cell of the wave with a negative flag into *first (atomic
// min over the waves; *first starts at ncell)
__device__ inline void note_failure(int c, int f, int* first) {
  int m = (f < 0) ? c : 0x7fffffff;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(m, off, 64);
    m = o < m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m != 0x7fffffff) atomicMin(first, m);
}

// element totals tot[c NE + e] of the rows v (X: mass fractions)
template <class Net, bool X>
__global__ void __launch_bounds__(256)
totals_kernel(int ncell, const double* __restrict__ v, long cs, long ss,
              double* __restrict__ tot) {
  using C = Conserve<Net>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double n[C::NS], t[C::NE];
#pragma unroll C::UNR
  for (int i = 0; i < C::NS; ++i) n[i] = v[c * cs + i * ss];
  if (X) C::totals_x(n, t);
  else C::totals(n, t);
#pragma unroll
  for (int e = 0; e < C::NE; ++e) tot[(long)c * C::NE + e] = t[e];
}

// consistent advection of the rows v onto ref + c rs (rs 0: one row for
// all, NE: a row per cell); flags[c]; the lowest failing cell into *first
template <class Net, bool X>
__global__ void __launch_bounds__(256)
conserve_kernel(int ncell, double* __restrict__ v, long cs, long ss,
                const double* __restrict__ ref, long rs,
                int* __restrict__ flags, int* first) {
  using C = Conserve<Net>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  int f = 0;
  if (c < ncell) {
    double n[C::NS], r[C::NE];
#pragma unroll C::UNR
    for (int i = 0; i < C::NS; ++i) n[i] = v[c * cs + i * ss];
#pragma unroll
    for (int e = 0; e < C::NE; ++e) r[e] = ref[c * rs + e];
    f = X ? C::to_x(n, r) : C::to_n(n, r, C::opaque_mask());
#pragma unroll C::UNR
    for (int i = 0; i < C::NS; ++i) v[c * cs + i * ss] = n[i];
    flags[c] = f;
  }
  note_failure(c, f, first);
}

// post-solve conservation, mode 2: before the solve, each cell's conserved
// totals target[c (NE + 1) + k] of its input ...
template <class Net>
__global__ void __launch_bounds__(256)
conserve_targets_kernel(CellsArgs a, double* __restrict__ target) {
  using C = Conserve<Net>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  double n[C::NS], t[C::NE + 1];
#pragma unroll C::UNR
  for (int i = 0; i < C::NS; ++i) n[i] = a.n[c * a.n_cs + i * a.n_ss];
  C::targets(n, t);
#pragma unroll
  for (int k = 0; k <= C::NE; ++k) target[(long)c * (C::NE + 1) + k] = t[k];
}

// ... and after it, each integrated cell (dt > 0, flag 0: clipped already)
// projected onto them, the charge row included
template <class Net>
__global__ void __launch_bounds__(256)
conserve_post_kernel(CellsArgs a, const double* __restrict__ target) {
  using C = Conserve<Net>;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell || !(a.dt[c] > 0.0) || a.flags[c] != 0) return;
  double n[C::NS], t[C::NE + 1];
#pragma unroll C::UNR
  for (int i = 0; i < C::NS; ++i) n[i] = a.n[c * a.n_cs + i * a.n_ss];
#pragma unroll
  for (int k = 0; k <= C::NE; ++k) t[k] = target[(long)c * (C::NE + 1) + k];
  C::template project<C::NE + 1>(n, t, C::opaque_mask());
#pragma unroll C::UNR
  for (int i = 0; i < C::NS; ++i) a.n[c * a.n_cs + i * a.n_ss] = n[i];
}

}  // namespace kx
