"""jac 1 (the analytic Jacobian, runtime/core/solve_cell.hpp:883-1118)
with Tgas evolved and with density-dependent rates:

- the oracle's Jacobian against the reference's own (jac_host_, the
  solver's analytic Jacobian, in tests/golden/<net>_funcs.npz):
  primordialZ at fixed Tgas (27 species, rates that read the densities:
  their derivatives by difference quotients here, symbolic through the
  aggregates or by difference quotients in the reference) within 1e-6 of
  each column's scale; primordial3 with Tgas evolved (10 equations): the
  Tgas column (central differences of the full rhs) within 1e-5, the
  species block within 1e-6 below 1e6 K (one sample at 3e6 K differs by
  7e-4 in the column of a trace species, where a difference quotient of
  a rate along a species at 7e-10 cm^-3 is roundoff), the Tgas row (the
  thermal rate's difference quotients with the rates re-evaluated) within
  1e-4 plus the quotients' roundoff;
- whole solves with jac 1 against the reference's (tests/golden/
  primordial3_jac1.npz, written by make_golden.py jac1 from the reference's
  solve_cells with jac 1): the oracle, and on the GPU the plain kernel,
  the persistent launch (which runs these options in the lane queue) and
  the lane queue, at the C2 bar (1e-6), flags equal.
"""
import ctypes

import numpy as np
import pytest

import krome
from conftest import NET_ID, golden, oracle_opts, rel_err

D = ctypes.POINTER(ctypes.c_double)
BAR = 1e-6


def full_jac(oracle, net, g, evolve_T):
    n, T, par = g["n"], g["T"], g["par"]
    nc, ns = n.shape
    N = ns + evolve_T
    J = np.zeros((nc, N * N))
    r = oracle.lib.kxo_jac_full(NET_ID[net], nc,
                                np.ascontiguousarray(n).ctypes.data_as(D),
                                np.ascontiguousarray(T).ctypes.data_as(D),
                                np.ascontiguousarray(par).ctypes.data_as(D),
                                evolve_T, J.ctypes.data_as(D))
    assert r == 0
    return J.reshape(nc, N, N).transpose(0, 2, 1)  # [c, i, j]


def test_jacobian_fixed_tgas_density_rates_vs_reference(oracle):
    g = golden("primordialZ_funcs.npz")
    J, Jr = full_jac(oracle, "primordialZ", g, 0), g["J"]
    scale = np.maximum(np.abs(Jr).max(axis=1, keepdims=True), 1e-300)
    assert np.max(np.abs(J - Jr) / scale) < BAR


def test_jacobian_evolved_tgas_vs_reference(oracle):
    g = golden("primordial3_funcs.npz")
    n, T = g["n"], g["T"]
    ns = n.shape[1]
    J, Jr = full_jac(oracle, "primordial3", g, 1), g["J"]
    scale = np.maximum(np.abs(Jr).max(axis=1, keepdims=True), 1e-300)
    d = np.abs(J - Jr) / scale
    # the Tgas column
    assert d[:, :, ns].max() < 1e-5
    # the species block
    cool = T < 1e6
    assert d[cool][:, :ns, :ns].max() < BAR
    assert d[:, :ns, :ns].max() < 1e-3
    # the Tgas row: the thermal rate's quotients, with their roundoff
    fT = np.abs(g["f"][:, ns])
    h = np.maximum(np.maximum(1e-7 * np.abs(n),
                              1e-9 * np.abs(n).sum(axis=1, keepdims=True)),
                   1e-30)
    allow = 1e-4 * np.abs(Jr[:, ns, :ns]) + 1e-10 * fT[:, None] / h
    assert (np.abs(J[:, ns, :ns] - Jr[:, ns, :ns]) <= allow).all()


def _opts_row(o):
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


def _check(n, T, flags, g):
    assert np.array_equal(flags, g["flag_bitwise"])
    assert rel_err(n, g["n_bitwise"]) < BAR
    assert np.max(np.abs(T - g["T_bitwise"]) / g["T_bitwise"]) < BAR


def test_oracle_jac1_solve_vs_reference(oracle):
    g = golden("primordial3_jac1.npz")
    assert g["opts"][5] == 1 and g["opts"][0] == 1
    r, n, T, fl, st = oracle.solve("primordial3", g["n0"], g["T0"], g["dt"],
                                   g["par"], _opts_row(g["opts"]))
    _check(n, T, fl, g)
    # no difference-quotient rhs evaluations: the Jacobian is analytic
    assert st[:, 3].sum() == 0
    # the reference's steps within 20 % per cell, 1 % in total
    sr = g["stats_bitwise"][:, 0]
    assert np.max(np.abs(st[:, 0] - sr) / sr) < 0.2
    assert abs(st[:, 0].sum() - sr.sum()) / sr.sum() < 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["plain", "persistent", "lanes"])
def test_gpu_jac1_vs_reference(kernel):
    g = golden("primordial3_jac1.npz")
    o = g["opts"]
    net = krome.load("primordial3")
    opts = net.options(evolve_T=True, rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=int(o[4]), jac=1, max_steps=int(o[6]),
                       device_kernel=kernel, error_report=False)
    print("primordial3 jac 1:", kernel, flush=True)
    res = net.solve(g["n0"], g["T0"], g["dt"], par=g["par"], opts=opts,
                    check=False)
    assert res.ierr == 0
    _check(res.n, res.Tgas, res.flags, g)
