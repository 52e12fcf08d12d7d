"""The oracle (oracle/: the CPU restatement of KROME-X's per-cell solve)
pinned before it is trusted:

- function level against the reference's own generated code (golden
  vectors written by tests/golden/make_golden.py): rates within 1e-9
  relative, the full rhs within 1e-12 of its scale and the thermal row
  within 1e-12 relative (the reference's netcheck bars, kromex/
  netcheck.py:602-760), the analytic Jacobian within 1e-12 of its scale;
- the integrator bitwise against the reference's own integrator header
  (runtime/core/cvode/bdf_cell.hpp) driving the same problem: equal states
  and equal counters nst/nfe/nje/nfeDQ/nsetups/nni/nnf/netf/restarts;
- whole solves against the reference's PER-CELL solver on the golden
  cells, with equal flags;
- what the mathematics fixes: every reaction conserves charge and the
  nuclei of every element, so the rhs has zero charge and nuclei rates; the analytic
  Jacobian is the derivative of the rhs (central differences, netcheck's
  1e-4 bar); the charge stays constant through a primordial3 solve (the
  Newton iteration's charge invariant); special cases of solve_cells.
"""
import ctypes

import numpy as np
import pytest

import cells
from conftest import golden, oracle_opts, rel_err

NETS = ("primordial", "primordial3", "primordialZ")


def _opts_of(d):
    o = d["opts"]
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


@pytest.mark.parametrize("net", NETS)
def test_rates_and_rhs_vs_reference_generated_code(oracle, net):
    g = golden("%s_funcs.npz" % net)
    evo = net == "primordial3"
    ns = len(cells.network_meta(net)["species"])
    k, f = oracle.funcs(net, g["n"], g["T"], g["par"], evo)
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)  # the same pieces / cuts
    nz = kr != 0.0
    # the rates: netcheck's bar (kromex/netcheck.py:676-717)
    assert np.max(np.abs(k[nz] - kr[nz]) / np.abs(kr[nz])) < 1e-9
    # the species rows: within 1e-10 of the largest (the rates' last-bit
    # differences of up to ~5e-11 carried by the largest fluxes)
    fr = g["f"]
    scale = np.abs(fr[:, :ns]).max(axis=1, keepdims=True)
    assert np.max(np.abs(f[:, :ns] - fr[:, :ns]) / scale) < 1e-10
    if evo:  # the thermal row
        assert np.max(np.abs(f[:, ns] - fr[:, ns]) / np.abs(fr[:, ns])) < 1e-12


def test_analytic_jacobian_vs_reference_and_fd(oracle):
    """primordial at fixed Tgas: the mass-action Jacobian equals the
    reference's analytic one, and is the derivative of the rhs"""
    net = "primordial"
    g = golden("primordial_funcs.npz")
    n, T, par = g["n"], g["T"], g["par"]
    nc, ns = n.shape
    J = np.zeros((nc, ns * ns))
    D = ctypes.POINTER(ctypes.c_double)
    oracle.lib.kxo_jac(0, nc, np.ascontiguousarray(n).ctypes.data_as(D),
                       np.ascontiguousarray(T).ctypes.data_as(D),
                       np.ascontiguousarray(par).ctypes.data_as(D),
                       J.ctypes.data_as(D))
    J = J.reshape(nc, ns, ns).transpose(0, 2, 1)  # [c, i, j]
    Jr = g["J"]
    scale = np.abs(Jr).max(axis=(1, 2), keepdims=True)
    assert np.max(np.abs(J - Jr) / scale) < 1e-12
    # central differences of the rhs (relative step 1e-4; the rates are
    # fixed, the rhs a polynomial), netcheck's 1e-4 bar on each column, with
    # the roundoff of the difference allowed: a row whose flux terms
    # R_i = sum_j |J_ij y_j| nearly cancel resolves only ~1e-16 R_i / h
    for c in range(0, nc, 17):
        y = n[c]
        R = np.abs(J[c]) @ np.abs(y)
        for j in range(ns):
            h = 1e-4 * max(abs(y[j]), 1e-30 * y.sum())
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            _, fp = oracle.funcs(net, yp[None], T[c:c + 1], par[c:c + 1], 0)
            _, fm = oracle.funcs(net, ym[None], T[c:c + 1], par[c:c + 1], 0)
            col = (fp[0, :ns] - fm[0, :ns]) / (2 * h)
            allow = 1e-4 * np.abs(J[c, :, j]).max() + 1e-13 * R / h
            assert (np.abs(col - J[c, :, j]) <= allow).all(), (c, j)


@pytest.mark.parametrize("net", NETS)
def test_rhs_conserves_charge_and_nuclei(oracle, net):
    """every reaction conserves charge and the nuclei of every element: the
    rhs is orthogonal to the charges and to each element's atom counts
    (the compositions parsed from the species names, checked against the
    reference's tables in test_abi and test_conserve)"""
    g = golden("%s_funcs.npz" % net)
    meta = cells.network_meta(net)
    _, f = oracle.funcs(net, g["n"], g["T"], g["par"],
                        net == "primordial3")
    ns = len(meta["species"])
    dn = f[:, :ns]
    q = np.array(meta["charge"], dtype=float)
    atoms = np.array(meta["natoms"], dtype=float)  # [species, element]
    assert atoms.shape == (ns, len(meta["elements"]))
    # by hand, for the names every network here has: H2 holds 2 H, He+ one
    # He, e- nothing
    el = meta["elements"]
    sp = meta["species"]
    assert atoms[sp.index("H2"), el.index("H")] == 2
    assert atoms[sp.index("He+"), el.index("He")] == 1
    assert not atoms[sp.index("e-")].any()
    scale = np.abs(dn).sum(axis=1)
    for w in [q] + [atoms[:, e] for e in range(len(el))]:
        assert np.max(np.abs(dn @ w) / scale) < 1e-13


@pytest.mark.parametrize("net", NETS)
def test_integrator_bitwise_vs_reference_integrator(oracle, oracle_ref, net):
    d = golden("%s_solve.npz" % net)
    o = _opts_of(d)
    a = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"], o)
    b = oracle_ref.solve(net, d["n0"], d["T0"], d["dt"], d["par"], o)
    assert a[0] == b[0] == 0
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[4], b[4])


# whole solves against the reference: C2 (rtol 1e-8) within 1e-6; C1 is
# integrated at rtol 1e-6, where the last-bit differences of the rate
# expressions move the trajectories by up to ~10 rtol: 5e-5 there
TOL = {"primordial": 5e-5, "primordial3": 1e-6, "primordialZ": 1e-4}


@pytest.mark.parametrize("net", NETS)
def test_solve_vs_reference_solver(oracle, net):
    d = golden("%s_solve.npz" % net)
    r, n, T, fl, st = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"],
                                   _opts_of(d))
    assert r == 0
    assert np.array_equal(fl, d["flag_bitwise"])
    assert rel_err(n, d["n_bitwise"]) < TOL[net]
    assert np.max(np.abs(T - d["T_bitwise"]) / d["T_bitwise"]) < TOL[net]
    # the work agrees too (not bitwise: the rate expressions differ in
    # their last bits): total steps within 2 %
    assert abs(st[:, 0].sum() / d["stats_bitwise"][:, 0].sum() - 1) < 0.02


def test_charge_and_nuclei_through_a_solve(oracle):
    net = "primordial3"
    d = golden("primordial3_solve.npz")
    meta = cells.network_meta(net)
    r, n, T, fl, st = oracle.solve(net, d["n0"], d["T0"], d["dt"], d["par"],
                                   _opts_of(d))
    q = np.array(meta["charge"], dtype=float)
    n0 = d["n0"]
    ref = np.abs(n0 * q).sum(axis=1)
    assert np.max(np.abs(n @ q - n0 @ q) / ref) < 1e-8
    he = np.array([1.0 if s.rstrip("+-") == "He" else 0.0
                   for s in meta["species"]])
    assert np.max(np.abs(n @ he - n0 @ he) / (n0 @ he)) < 1e-6


def test_special_cases(oracle):
    net = "primordial3"
    w = cells.primordial3_cells(8)
    o = oracle_opts(evolve_T=1)
    # dt <= 0: untouched, flag 0
    dt = w["dt"].copy()
    dt[2] = 0.0
    dt[5] = -1.0
    r, n, T, fl, st = oracle.solve(net, w["n"], w["T"], dt, w["par"], o)
    assert r == 0 and (fl == 0).all()
    assert np.array_equal(n[[2, 5]], w["n"][[2, 5]])
    assert np.array_equal(T[[2, 5]], w["T"][[2, 5]])
    assert (st[[2, 5]] == 0).all()
    # a REQUIRED parameter unset in one cell: that cell flagged -105, the
    # others 0, nothing integrated (solve_cells.hpp:133-140)
    par = w["par"].copy()
    par[3, 1] = np.nan
    r, n, T, fl, st = oracle.solve(net, w["n"], w["T"], w["dt"], par, o)
    assert r == -105 and fl[3] == -105 and (np.delete(fl, 3) == 0).all()
    assert np.array_equal(n, w["n"])
    # a tiny step budget: TOO_MUCH_WORK after the restarts, inputs untouched
    r, n, T, fl, st = oracle.solve(net, w["n"], w["T"], w["dt"], w["par"],
                                   oracle_opts(evolve_T=1, max_steps=3))
    assert r == -1 and (fl == -1).all()
    assert np.array_equal(n, w["n"]) and np.array_equal(T, w["T"])
    # a non-positive Tgas: the first rhs is refused -> FLAG_NONPOSITIVE_T
    T0 = w["T"].copy()
    T0[1] = -5.0
    r, n, T, fl, st = oracle.solve(net, w["n"], T0, w["dt"], w["par"], o)
    assert fl[1] == -110 and r == -110
    assert (np.delete(fl, 1) == 0).all()
