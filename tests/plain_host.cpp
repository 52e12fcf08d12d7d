// Host build of the plain kernel's per-cell solve (kx::plain_cell, the
// GPU thread body of krome-x_amd/csrc/kx_plain.h) for the CPU tests: the
// same code the plain kernel and the lane queue run, here cell by cell over
// an interleaved workspace of the batch's cells, so the large-network
// strided form (cloud: the state, the rhs and the sparse Jacobian in the
// workspace, the rates cached per solve) is pinned on the CPU against the
// reference's solves.  Built by tests/test_cloud.py with g++ into
// tests/_build/.  Networks: 0 primordialZ (register form, Tgas fixed), 1
// cloud (strided form).
#include <vector>

#include "kx_plain.h"
#include "net_cloud.h"
#include "net_primordialZ.h"

namespace {

template <class Net>
int solve(int ncell, double* n, double* T, const double* dt,
          const double* par, const double* o, int* flags, long* stats) {
  using Prob = kx::PlainProblem<Net, false>;
  kx::CellsArgs a;
  a.ncell = ncell;
  a.n = n;
  a.n_cs = Net::NSPEC;  // cell-major rows, as the fixtures hold them
  a.n_ss = 1;
  a.T = T;
  a.dt = dt;
  a.par = par;
  a.par_stride = Net::NPAR;
  a.flags = flags;
  a.stats = stats;
  kx::SolveOpts& so = a.o;
  so.rtol = o[0];
  so.atol_n = o[1];
  so.atol_T = o[2];
  so.max_steps = static_cast<long>(o[3]);
  so.max_err_fails = 7;
  so.max_conv_fails = 10;
  so.max_restarts = 5;
  so.evolve_T = 0;
  so.jac = static_cast<int>(o[4]);
  so.clip_negative = 1;
  so.conserve = 0;
  so.linsol = static_cast<int>(o[5]);
  const long S = ncell;
  std::vector<double> work(kx::cell_doubles<Prob>() * S);
  std::vector<int> iwork(Prob::N * S);
  for (int c = 0; c < ncell; ++c)
    kx::plain_cell<Net, false>(a, c, work.data() + c, iwork.data() + c, S);
  for (int c = 0; c < ncell; ++c)
    if (flags[c] != 0) return flags[c];
  return 0;
}

}  // namespace

extern "C" {

// net 0 primordialZ, 1 cloud; n[ncell NSPEC] and T updated in place;
// o = rtol, atol_n, atol_T, max_steps, jac, linsol; stats [ncell 9]
int plain_host_solve(int net, int ncell, double* n, double* T,
                     const double* dt, const double* par, const double* o,
                     int* flags, long* stats) {
  if (net == 0)
    return solve<kx::net_primordialZ>(ncell, n, T, dt, par, o, flags, stats);
  if (net == 1)
    return solve<kx::net_cloud>(ncell, n, T, dt, par, o, flags, stats);
  return -1;
}

}  // extern "C"
