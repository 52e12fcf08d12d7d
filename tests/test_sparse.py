"""The static-pivot sparse LU of the species block (linsol 3, the
reference's LINSOL_SPARSE: bdf_cell.hpp:1155-1264, with its dense border
for Tgas through the Schur complement).

- the generated tables (pivot order by minimum degree, rows of L + U with
  their fill, diagonal positions, flops) equal the reference generator's
  (tests/golden/*_sparse.npz);
- the oracle's restatement is bitwise equal to the reference's own
  integrator header driving the same problem with the same tables (final
  states, Tgas and all 9 counters), on primordial (no border) and
  primordial3 (one border equation);
- the oracle's solve agrees with the reference's solve_cells with linsol 3
  within the solve bars, flags equal;
- on the GPU (plain kernel through the C ABI): the same bars against the
  reference's sparse solve and the oracle; the persistent kernels refuse
  linsol 3 (-104 for every cell, inputs untouched), for the networks it
  serves; a larger one runs in the lane queue, bitwise the plain kernel.
"""
import sys

import numpy as np
import pytest

import cells
from conftest import golden, oracle_opts, rel_err

NETS = ("primordial", "primordial3", "primordialZ")
TOL = {"primordial": 5e-5, "primordial3": 1e-6, "primordialZ": 3e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _opts(d):
    o = d["opts"]
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


@pytest.mark.parametrize("net", NETS)
def test_tables_vs_reference_generator(net):
    g = golden("%s_sparse.npz" % net)
    sp = cells.network_meta(net)["sparse"]
    for k in ("perm", "rp", "ci", "di"):
        assert np.array_equal(np.array(sp[k]), g["sp_" + k]), k
    assert sp["flops"] == int(g["sparse_flops"])
    # a permutation; rows ascending with the diagonal inside
    ns = len(sp["perm"])
    assert sorted(sp["perm"]) == list(range(ns))
    for k in range(ns):
        row = sp["ci"][sp["rp"][k]:sp["rp"][k + 1]]
        assert row == sorted(row) and sp["ci"][sp["di"][k]] == k


@pytest.mark.parametrize("net", NETS)
def test_oracle_bitwise_vs_reference_integrator(oracle, oracle_ref, net):
    d = golden("%s_sparse.npz" % net)
    o = _opts(d)
    assert o[4] == 3.0
    a = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"], o)
    b = oracle_ref.solve(net, d["n0"], d["T0"], d["dt"], d["par"], o)
    assert a[0] == b[0] == 0
    for i in (1, 2, 3, 4):
        assert np.array_equal(a[i], b[i])


@pytest.mark.parametrize("net", NETS)
def test_oracle_vs_reference_sparse_solve(oracle, net):
    d = golden("%s_sparse.npz" % net)
    r, n, T, fl, st = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"],
                                   _opts(d))
    assert r == 0 and np.array_equal(fl, d["flag"])
    assert rel_err(n, d["n"]) < TOL[net]
    assert np.max(np.abs(T - d["T"]) / d["T"]) < TOL[net]
    assert abs(st[:, 0].sum() / d["stats"][:, 0].sum() - 1) < 0.02
    # and the dense LU reaches the same states (the matrix differs only in
    # its rounding: the pattern is exact for these networks)
    o = _opts(d)
    o[4] = 1.0
    r2, n2, T2, fl2, st2 = oracle.solve(net, d["n0"], d["T0"], d["dt"],
                                        d["par"], o)
    assert rel_err(n, n2) < TOL[net]


# ---- the HIP plain kernel --------------------------------------------------

@pytest.fixture(scope="module")
def nets():
    import torch
    assert torch.cuda.is_available()
    import krome
    return {n: krome.load(n) for n in NETS}


@pytest.mark.gpu
@pytest.mark.parametrize("net", NETS)
def test_gpu_sparse_vs_reference(nets, oracle, net):
    d = golden("%s_sparse.npz" % net)
    N = nets[net]
    o = d["opts"]
    say("sparse", net)
    opts = N.options(evolve_T=bool(o[0]), rtol=o[1], atol_n=o[2],
                     atol_T=o[3], linsol=3, jac=int(o[5]),
                     device_kernel="plain", error_report=False)
    res = N.solve(d["n0"], d["T0"], d["dt"], par=d["par"], opts=opts,
                  check=False)
    assert res.ierr == 0 and np.array_equal(res.flags, d["flag"])
    assert rel_err(res.n, d["n"]) < TOL[net]
    assert np.max(np.abs(res.Tgas - d["T"]) / d["T"]) < TOL[net]
    r, n, T, fl, st = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"],
                                   _opts(d))
    assert rel_err(res.n, n) < TOL[net]
    # the lockstep persistent kernel factorizes the dense matrix only:
    # refused; a network above its 16 equations takes the lane queue,
    # the plain kernel's solve (so the same results, bit for bit)
    op = N.options(evolve_T=bool(o[0]), rtol=o[1], atol_n=o[2],
                   atol_T=o[3], linsol=3, jac=int(o[5]),
                   device_kernel="persistent", error_report=False)
    res2 = N.solve(d["n0"], d["T0"], d["dt"], par=d["par"], opts=op,
                   check=False)
    if N.nspec + int(bool(o[0])) <= 16:
        assert res2.ierr == -104 and (res2.flags == -104).all()
        assert np.array_equal(res2.n, d["n0"])
    else:
        assert np.array_equal(res2.flags, res.flags)
        assert np.array_equal(res2.n, res.n)


@pytest.mark.gpu
def test_gpu_sparse_large_batch(nets, oracle):
    """65536 primordial3 cells with linsol 3 against the oracle's sparse
    solve of a sample, and every flag 0"""
    N = nets["primordial3"]
    w = cells.primordial3_cells(65536, first=1 << 20)
    say("sparse large", 65536)
    opts = N.options(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                     linsol=3, device_kernel="plain", error_report=False)
    res = N.solve(w["n"], w["T"], w["dt"], par=w["par"], opts=opts,
                  check=False)
    assert res.ierr == 0 and (res.flags == 0).all()
    idx = np.arange(0, 65536, 683)
    r, n, T, fl, st = oracle.solve("primordial3", w["n"][idx], w["T"][idx],
                                   w["dt"][idx], w["par"][idx],
                                   oracle_opts(evolve_T=1, linsol=3))
    assert rel_err(res.n[idx], n) < TOL["primordial3"]
    assert np.max(np.abs(res.Tgas[idx] - T) / T) < TOL["primordial3"]
