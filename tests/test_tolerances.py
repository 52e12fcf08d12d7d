"""Per-species tolerances (krome_opts_set_atol_species / rtol_species,
runtime/interfaces/C/krome.h:175-185; applied in runtime/core/
solve_cell.hpp:866-879, 2129-2139 and the per-slot error weights
1 / fma(rtol_i, |y_i|, atol_i) of runtime/core/cvode/bdf_cell.hpp:689-697).

Pinned by the reference itself: tests/golden/<net>_tol.npz holds the
reference's solve_cells (PATH_PERCELL) of 256 cells with per-species atol
only, then with per-species atol and rtol (some entries negative: the
scalar), written by tests/golden/make_golden.py tol.

- CPU: the oracle's restatement equals the reference's own integrator
  (oracle/_ref) bitwise with these tolerances; the oracle meets the
  reference's solves at the solve bars with equal flags; tolerance rows
  equal to the scalars reproduce the scalar solve bit for bit; the C ABI's
  set/get semantics (NaN, zero atol, NULL, negative entries).
- GPU: the plain FP64, persistent FP64 and persistent mixed kernels
  against the reference's solves through krome_solve_cells.
"""
import ctypes

import numpy as np
import pytest

import krome
from conftest import golden, oracle_opts, rel_err

NETS = ("primordial", "primordial3")
# the solve bars of the scalar solves (test_oracle.py): C2 at rtol 1e-8
# within 1e-6; C1 at rtol 1e-6 within 5e-5
TOL = {"primordial": 5e-5, "primordial3": 1e-6}
VARIANTS = ("atol", "both")


def _opts(g, tag):
    o = g["opts_" + tag]
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


def _effective(g, tag):
    """the tolerance each species gets in variant tag (negative: the
    scalar)"""
    o = g["opts_" + tag]
    ns = g["n0"].shape[1]
    a, r = o[8:8 + ns], o[8 + ns:8 + 2 * ns]
    return (np.where(a < 0, o[2], a), np.where(r < 0, o[1], r))


@pytest.mark.parametrize("net", NETS)
@pytest.mark.parametrize("tag", VARIANTS)
def test_oracle_bitwise_vs_reference_integrator(oracle, oracle_ref, net,
                                                tag):
    g = golden("%s_tol.npz" % net)
    o, tol = _opts(g, tag), _effective(g, tag)
    a = oracle.solve(net, g["n0"], g["T0"], g["dt"], g["par"], o, tol)
    b = oracle_ref.solve(net, g["n0"], g["T0"], g["dt"], g["par"], o, tol)
    assert a[0] == b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("net", NETS)
@pytest.mark.parametrize("tag", VARIANTS)
def test_oracle_vs_reference_solves(oracle, net, tag):
    g = golden("%s_tol.npz" % net)
    r, n, T, fl, st = oracle.solve(net, g["n0"], g["T0"], g["dt"], g["par"],
                                   _opts(g, tag), _effective(g, tag))
    assert np.array_equal(fl, g["flag_" + tag])
    assert rel_err(n, g["n_" + tag]) < TOL[net]
    assert np.max(np.abs(T - g["T_" + tag]) / g["T_" + tag]) < TOL[net]
    # the steps within 2 %, as the scalar solves
    s, sr = st[:, 0].sum(), g["stats_" + tag][:, 0].sum()
    assert abs(s - sr) < 0.02 * sr


def test_tolerances_change_the_solve(oracle):
    """the per-species rows take effect: with them the reference takes
    other steps than in its scalar solve, and the oracle's steps move off
    its own scalar solve in the same cells"""
    net = "primordial3"
    g = golden("%s_tol.npz" % net)
    d = golden("%s_solve.npz" % net)
    assert np.array_equal(d["n0"], g["n0"])
    ref_steps = g["stats_both"][:, 0].astype(float)
    ref_scalar = d["stats_bitwise"][:, 0].astype(float)
    moved = ref_steps != ref_scalar
    assert moved.mean() > 0.5
    o = _opts(g, "both")
    _, n, _, _, st = oracle.solve(net, g["n0"], g["T0"], g["dt"], g["par"],
                                  o, _effective(g, "both"))
    _, n0, _, _, st0 = oracle.solve(net, g["n0"], g["T0"], g["dt"],
                                    g["par"], o)
    mine = st[:, 0] != st0[:, 0]
    assert mine.mean() > 0.5
    # the same cells move, but for a few whose step counts the last-bit
    # differences of the rate expressions decide
    assert (mine == moved).mean() > 0.95
    assert not np.array_equal(n, n0)


@pytest.mark.parametrize("net", NETS)
def test_scalar_rows_reproduce_the_scalar_solve(oracle, net):
    """rows equal to the scalars: the same weights in the same arithmetic
    (rtol rows equal to rtol keep CVODE's scalar form), so the same solve
    bit for bit; a per-slot rtol form with every rtol equal computes
    fma(rtol, |y|, atol), which differs from (rtol |y|) + atol only in the
    last bit of the weights"""
    g = golden("%s_tol.npz" % net)
    o = _opts(g, "atol")
    ns = g["n0"].shape[1]
    a = oracle.solve(net, g["n0"], g["T0"], g["dt"], g["par"], o)
    b = oracle.solve(net, g["n0"], g["T0"], g["dt"], g["par"], o,
                     (np.full(ns, o[2]), np.full(ns, o[1])))
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_abi_set_get_semantics():
    net = krome.load("primordial3")
    L = net._lib
    D = ctypes.POINTER(ctypes.c_double)
    for f in ("set_atol_species", "set_rtol_species", "get_atol_species",
              "get_rtol_species"):
        getattr(L, "krome_opts_" + f).restype = ctypes.c_int
        getattr(L, "krome_opts_" + f).argtypes = [ctypes.c_void_p, D]
    ns = net.nspec
    opts = net.options(rtol=1e-7, atol_n=1e-19)
    out = np.zeros(ns)
    # unset: the scalars
    assert L.krome_opts_get_atol_species(opts._h, out.ctypes.data_as(D)) == 0
    assert (out == 1e-19).all()
    assert L.krome_opts_get_rtol_species(opts._h, out.ctypes.data_as(D)) == 0
    assert (out == 1e-7).all()
    # negative entries take the scalar
    v = np.full(ns, -1.0)
    v[3] = 1e-25
    assert L.krome_opts_set_atol_species(opts._h, v.ctypes.data_as(D)) == 0
    L.krome_opts_get_atol_species(opts._h, out.ctypes.data_as(D))
    assert out[3] == 1e-25 and (np.delete(out, 3) == 1e-19).all()
    # NaN and a zero atol refused, the handle unchanged
    bad = v.copy()
    bad[0] = np.nan
    assert L.krome_opts_set_atol_species(opts._h, bad.ctypes.data_as(D)) == -1
    assert L.krome_opts_set_rtol_species(opts._h, bad.ctypes.data_as(D)) == -1
    bad[0] = 0.0
    assert L.krome_opts_set_atol_species(opts._h, bad.ctypes.data_as(D)) == -1
    L.krome_opts_get_atol_species(opts._h, out.ctypes.data_as(D))
    assert out[3] == 1e-25
    # a zero rtol is allowed
    assert L.krome_opts_set_rtol_species(opts._h, bad.ctypes.data_as(D)) == 0
    L.krome_opts_get_rtol_species(opts._h, out.ctypes.data_as(D))
    assert out[0] == 0.0 and out[1] == 1e-7
    # NULL: back to the scalars
    assert L.krome_opts_set_atol_species(opts._h, None) == 0
    L.krome_opts_get_atol_species(opts._h, out.ctypes.data_as(D))
    assert (out == 1e-19).all()
    # the Python handle's methods
    opts.set_species_tolerances(atol=v)
    assert opts.species_tolerances()[0][3] == 1e-25


@pytest.mark.gpu
@pytest.mark.parametrize("net", NETS)
@pytest.mark.parametrize("tag", VARIANTS)
@pytest.mark.parametrize("kern,mixed", [("plain", 0), ("persistent", 0),
                                        ("persistent", 1)])
def test_gpu_kernels_vs_reference(net, tag, kern, mixed):
    g = golden("%s_tol.npz" % net)
    o = g["opts_" + tag]
    ns = g["n0"].shape[1]
    lib = krome.load(net)
    opts = lib.options(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=int(o[4]), jac=int(o[5]), max_steps=int(o[6]),
                       device_kernel=kern, mixed_precision=mixed)
    opts.set_species_tolerances(atol=o[8:8 + ns], rtol=o[8 + ns:8 + 2 * ns])
    print("tolerances: %s %s %s mixed=%d" % (net, tag, kern, mixed),
          flush=True)
    res = lib.solve(g["n0"], g["T0"], g["dt"], par=g["par"], opts=opts,
                    check=False)
    assert res.ierr == 0
    assert np.array_equal(res.flags, g["flag_" + tag])
    # the FP64 kernels at the solve bars; the mixed kernel (FP32 Newton
    # matrix with Tgas evolved; FP64 at fixed Tgas) at its 1e-5 bar
    tol = 1e-5 if (mixed and net == "primordial3") else TOL[net]
    assert rel_err(res.n, g["n_" + tag]) < tol
    assert np.max(np.abs(res.Tgas - g["T_" + tag]) / g["T_" + tag]) < tol
