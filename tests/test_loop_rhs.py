"""The loop form of the rhs and the species Jacobian (index tables, the
form of networks above netgen.RHS_LOOP_MIN reactions: the reference's
codegen.py:8151-8223).

- the tables describe the network: each reaction's kinetic reactants, its
  scatter rows (reactants -1, products +1) and its Jacobian terms;
- the oracle evaluates both generated forms at the reference's function
  samples: bitwise equal rhs and Jacobian (same operations, same order);
- on the GPU, libkrome_primordial_rhsloop.so (the primordial library
  compiled with the loop form) solves the golden cells like the unrolled
  library: same flags, states within the solve bar of each other and of
  the reference.
"""
import ctypes
import sys

import numpy as np
import pytest

import cells
from conftest import NET_ID, PKG, golden, rel_err

NETS = ("primordial", "primordial3")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("net", NETS)
def test_tables_describe_the_network(net):
    """the tables, read back, are the network's stoichiometry"""
    import netgen
    hdr = (PKG / "csrc" / "generated" / ("net_%s.h" % net)).read_text()
    meta = cells.network_meta(net)
    tab = {}
    for name in ("KF_OFF", "KF_SP", "SC_OFF", "SC_ROW", "SC_SGN", "JT_OFF",
                 "JT_COL", "JT_MULT", "JT_FOFF", "JF_SP"):
        line = next(x for x in hdr.splitlines()
                    if "int LT_%s[" % name in x)
        tab[name] = [int(v) for v in line.split("{")[1].split("}")[0]
                     .split(",")]
    ns, nr = len(meta["species"]), meta["nrea"]
    assert len(tab["KF_OFF"]) == nr + 1 and len(tab["SC_OFF"]) == nr + 1
    # the net stoichiometry of every reaction, from the scatter rows, is
    # that of its equation
    sp = {s: i for i, s in enumerate(meta["species"])}
    for r, eq in enumerate(meta["reactions"]):
        lhs, rhs = eq.split("->")
        want = np.zeros(ns, dtype=int)
        for t in lhs.split(" + "):
            want[sp[t.strip().strip("{}")]] -= 1
        for t in rhs.split(" + "):
            want[sp[t.strip()]] += 1
        got = np.zeros(ns, dtype=int)
        for q in range(tab["SC_OFF"][r], tab["SC_OFF"][r + 1]):
            got[tab["SC_ROW"][q]] += tab["SC_SGN"][q]
        assert np.array_equal(got, want), eq
        # one Jacobian term per distinct kinetic reactant, multiplicities
        # summing to the reaction order
        kin = tab["KF_SP"][tab["KF_OFF"][r]:tab["KF_OFF"][r + 1]]
        terms = range(tab["JT_OFF"][r], tab["JT_OFF"][r + 1])
        assert sorted(tab["JT_COL"][t] for t in terms) == sorted(set(kin))
        assert sum(tab["JT_MULT"][t] for t in terms) == len(kin)
    assert netgen.RHS_LOOP_MIN == 500


@pytest.mark.parametrize("net", NETS)
def test_oracle_forms_bitwise(oracle, net):
    g = golden("%s_funcs.npz" % net)
    n = np.ascontiguousarray(g["n"])
    T = np.ascontiguousarray(g["T"])
    par = np.ascontiguousarray(g["par"])
    nc, ns = n.shape
    f = np.zeros((2, nc, ns))
    J = np.zeros((2, nc, ns * ns))
    D = ctypes.POINTER(ctypes.c_double)
    r = oracle.lib.kxo_rhs_forms(NET_ID[net], nc, n.ctypes.data_as(D),
                                 T.ctypes.data_as(D), par.ctypes.data_as(D),
                                 f.ctypes.data_as(D), J.ctypes.data_as(D))
    assert r == 0
    assert np.array_equal(f[0], f[1])
    assert np.array_equal(J[0], J[1])
    # and the unrolled species rows are the reference's (the bar of
    # tests/test_oracle.py)
    scale = np.abs(g["f"][:, :ns]).max(axis=1, keepdims=True)
    assert np.max(np.abs(f[1] - g["f"][:, :ns]) / scale) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,mixed", [("plain", 0), ("persistent", 0),
                                          ("persistent", 1)])
def test_gpu_loop_form_library(kernel, mixed):
    import torch
    assert torch.cuda.is_available()
    import krome
    so = PKG / "libkrome_primordial_rhsloop.so"
    assert so.exists(), "build the libraries first (krome-x_amd/build.py)"
    loop = krome.load(path=so)
    base = krome.load("primordial")
    d = golden("primordial_solve.npz")
    o = d["opts"]
    say("loop form", kernel, mixed)
    out = []
    for N in (base, loop):
        opts = N.options(evolve_T=False, rtol=o[1], atol_n=o[2],
                         atol_T=o[3], device_kernel=kernel,
                         mixed_precision=mixed, error_report=False)
        res = N.solve(d["n0"], d["T0"], d["dt"], par=d["par"], opts=opts,
                      check=False)
        assert res.ierr == 0
        out.append(res)
    assert np.array_equal(out[0].flags, out[1].flags)
    tol = 5e-5
    assert rel_err(out[1].n, out[0].n) < tol
    assert rel_err(out[1].n, d["n_bitwise"]) < (max(tol, 1e-5) if mixed
                                                else tol)
