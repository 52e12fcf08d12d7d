"""The query functions (krome_get_rates, krome_get_fluxes,
krome_get_thermal; runtime/core/query.hpp:37-108): the rates, fluxes and
thermal terms of states as the solver sees them, computed on the device and
checked against the reference's own generated code (tests/golden/
*_funcs.npz: its rates, and its thermal terms at the thermal temperature)
with the oracle's bars (tests/test_oracle.py): rates within 1e-9 relative
and with the same zero pattern; thermal terms within 1e-10 relative."""
import ctypes
import sys

import numpy as np
import pytest

import cells
from conftest import PKG, golden

NETS = ("primordial", "primordial3")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(net):
    so = PKG / ("libkrome_%s.so" % net)
    if not so.exists():
        pytest.skip("%s not built" % so.name)
    import krome
    return krome.load(net)


@pytest.mark.parametrize("net", NETS)
def test_term_names_and_argument_checks(net):
    N = load(net)
    if net == "primordial3":
        assert N.heating_terms == ("CHEM", "COMPRESS")
        assert N.cooling_terms == ("H2", "CHEM", "CONT")
    else:
        assert N.heating_terms == () and N.cooling_terms == ()
    L = N._lib
    assert L.krome_heating_name(-1) is None
    assert L.krome_cooling_name(N._lib.krome_ncool()) is None
    D = ctypes.POINTER(ctypes.c_double)
    n = np.ones((2, N.nspec))
    T = np.ones(2)
    k = np.zeros((2, N.nrea))
    nd, Td, kd = (a.ctypes.data_as(D) for a in (n, T, k))
    for f in (L.krome_get_rates, L.krome_get_fluxes):
        assert f(-1, nd, Td, None, 0, None, None, kd) == -1
        assert f(2, None, Td, None, 0, None, None, kd) == -1
        assert f(2, nd, None, None, 0, None, None, kd) == -1
        assert f(2, nd, Td, None, 0, None, None, None) == -1
        assert f(2, nd, Td, nd, N.npar + 1, None, None, kd) == -1
        assert f(0, nd, Td, None, 0, None, None, kd) == 0
    assert L.krome_get_thermal(2, nd, None, None, 0, None, None, None, None,
                               None, None, None) == -1
    assert L.krome_get_thermal(0, nd, Td, None, 0, None, None, None, None,
                               None, None, None) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("net", NETS)
def test_gpu_rates_and_fluxes_vs_reference(net):
    g = golden("%s_funcs.npz" % net)
    N = load(net)
    say("query rates", net)
    k = N.rates(g["n"], g["T"], par=g["par"])
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)
    nz = kr != 0.0
    assert np.max(np.abs(k[nz] - kr[nz]) / np.abs(kr[nz])) < 1e-9
    # fluxes: k prod(n) over each reaction's kinetic reactants
    fl = N.fluxes(g["n"], g["T"], par=g["par"])
    meta = cells.network_meta(net)
    sp = {s: i for i, s in enumerate(meta["species"])}
    want = np.array(kr, copy=True)
    for r, eq in enumerate(meta["reactions"]):
        for t in eq.split("->")[0].split(" + "):
            t = t.strip()
            if not t.startswith("{"):
                want[:, r] *= g["n"][:, sp[t]]
    nz = want != 0.0
    assert np.array_equal(fl == 0.0, want == 0.0)
    assert np.max(np.abs(fl[nz] - want[nz]) / np.abs(want[nz])) < 1e-9


@pytest.mark.gpu
def test_gpu_thermal_vs_reference():
    """primordial3's thermal terms against the reference's: H2 cooling (with
    its opacity), continuum, chemical cooling and heating, compression
    heating, and dT/dt"""
    net = "primordial3"
    g = golden("%s_funcs.npz" % net)
    N = load(net)
    say("query thermal")
    th = N.thermal(g["n"], g["T"], par=g["par"])
    ref = g["thermal"]
    pairs = ((th["cools"][:, 0], ref[:, 0]), (th["cools"][:, 2], ref[:, 1]),
             (th["cools"][:, 1], ref[:, 2]), (th["heats"][:, 0], ref[:, 3]),
             (th["heats"][:, 1], ref[:, 4]), (th["dTdt"], ref[:, 6]))
    for got, want in pairs:
        scale = np.maximum(np.abs(want), 1e-300)
        assert np.max(np.abs(got - want) / scale) < 1e-10
    assert np.allclose(th["heat"], th["heats"].sum(axis=1), rtol=1e-15)
    assert np.allclose(th["cool"], th["cools"].sum(axis=1), rtol=1e-15)


@pytest.mark.gpu
def test_gpu_query_unset_parameter():
    N = load("primordial3")
    w = cells.primordial3_cells(8)
    par = w["par"].copy()
    par[3, 1] = np.nan  # tff
    D = ctypes.POINTER(ctypes.c_double)
    n = np.ascontiguousarray(w["n"])
    T = np.ascontiguousarray(w["T"])
    k = np.full((8, N.nrea), -7.0)
    r = N._lib.krome_get_rates(8, n.ctypes.data_as(D), T.ctypes.data_as(D),
                               par.ctypes.data_as(D), N.npar, None, None,
                               k.ctypes.data_as(D))
    assert r == -105
    assert (k[3] == -7.0).all() and (np.delete(k, 3, axis=0) != -7.0).any()
