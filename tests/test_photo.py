"""Photon bins and photo reactions (galaxy_ism): KROME-X's photo module,
runtime/core/photo.hpp (PhotoModule::init 135-253, photo_rates 357-375),
the cross sections of photo(xsec=...) that the generator resolves
(kromex/codegen.py:686-812, 7530-7629) and the field argument of the C ABI
(krome.h:203-228, 598-617; krome.cpp:313-377, 415-417, 1121-1131).

Pinned against tests/golden/galaxy_ism_photo.npz, written by the
reference's own generator and PhotoModule in this container
(tests/golden/make_golden.py photo): its cross sections sampled on an
energy grid, its bin tables for three in-bin spectra, its rates with the
fields, and its solves of 256 cells with the fields."""
import ctypes
import math
import sys

import numpy as np
import pytest

import cells
from conftest import PKG, golden, oracle_opts

NET = "galaxy_ism"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load():
    so = PKG / ("libkrome_%s.so" % NET)
    if not so.exists():
        pytest.skip("%s not built" % so.name)
    import krome
    return krome.load(NET)


def oopts(o):
    """the oracle's options from a fixture's (evolve_T, rtol, atol_n,
    atol_T, linsol, jac, max_steps)"""
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    d[(a == 0) & (b == 0)] = 0.0
    return d


def big_species_rel(n, nr):
    """max relative error over species above 1e-12 of the cell total"""
    big = nr > 1e-12 * nr.sum(axis=1, keepdims=True)
    return rel(n, nr)[big].max()


def test_photo_tables_vs_reference():
    """the generator's photo reactions, reactants and energy windows are
    the reference generator's (the windows from the database thresholds,
    the tables' ends and the photo() arguments)"""
    g = golden("%s_photo.npz" % NET)
    ph = cells.network_meta(NET)["photo"]
    assert [p["reaction"] for p in ph] == list(g["photo_reaction"])
    assert [p["reactant"] for p in ph] == list(g["photo_reactant"])
    assert [p["Emin"] for p in ph] == list(g["photo_Emin"])
    assert [p["Emax"] for p in ph] == list(g["photo_Emax"])
    N = load()
    assert N.nphoto == len(ph)
    assert N.photo_reactions() == [(p["reaction"], p["reactant"])
                                   for p in ph]


def test_cross_sections_vs_reference():
    """sigma_q(E) of every source galaxy_ism uses (the database's Verner
    fits, the SWRI and Leiden tables, an expression, H2_sigmaLW) against
    the reference's generated functions on 3000+ energies, the window
    edges included"""
    g = golden("%s_photo.npz" % NET)
    N = load()
    E = g["E"]
    for q in range(N.nphoto):
        mine = N.photo_xsec(q, E)
        win = (E >= g["photo_Emin"][q]) & (E <= g["photo_Emax"][q])
        ref = np.where(win, g["xsec"][q], 0.0)
        assert np.array_equal(mine == 0.0, ref == 0.0), q
        assert rel(mine, ref).max() < 1e-13, q
    assert math.isnan(N._lib.krome_photo_xsec(-1, 10.0))
    assert math.isnan(N._lib.krome_photo_xsec(N.nphoto, 10.0))


@pytest.mark.parametrize("alpha", [1.0, 0.0, 2.0])
def test_bin_tables_vs_reference(alpha, oracle):
    """the library's bin-averaged cross sections (computed on the host when
    the bins are set) and the oracle's restatement against the reference's
    PhotoModule::init, for each in-bin spectrum"""
    g = golden("%s_photo.npz" % NET)
    N = load()
    ref = g["sigma_a%g" % alpha]
    o = N.options().set_photo_alpha(alpha).set_photo(g["Elo"], g["Ehi"])
    assert o.nbins == len(g["Elo"])
    mine = o.photo_sigma()
    assert np.array_equal(mine == 0.0, ref == 0.0)
    assert rel(mine, ref).max() < 1e-14
    orc = oracle.photo_bins(NET, g["Elo"], g["Ehi"], alpha)
    assert rel(orc, ref).max() < 1e-14


def test_bin_average_is_the_integral(oracle):
    """independent of the reference: a bin's sigma times its weight
    Int E^-alpha dE is Int sigma(E) E^-alpha dE over the bin's overlap
    with the window (a smooth fit integrated with scipy's adaptive
    quadrature; the 64-cell midpoint rule is good to 1e-3 over these bins
    of a cross section falling as E^-3, far below what a lost weight or a
    wrong exponent would change), a bin
    outside every window averages to 0, and splitting a bin into halves
    keeps the integral"""
    from scipy.integrate import quad
    N = load()
    elo = np.array([0.2, 11.0, 13.0, 25.0, 60.0, 150.0])
    ehi = np.array([0.5, 14.0, 20.0, 40.0, 90.0, 400.0])
    for alpha in (1.0, 0.0, 2.0):
        sig = oracle.photo_bins(NET, elo, ehi, alpha)
        o = N.options().set_photo_alpha(alpha).set_photo(elo, ehi)
        lib = o.photo_sigma()
        assert rel(lib, sig).max() < 1e-14
        assert (sig[:, 0] == 0.0).all()  # 0.2-0.5 eV: below every window

        def wint(a, b):
            return (math.log(b / a) if alpha == 1.0 else
                    (b - a if alpha == 0.0 else
                     (b ** (1 - alpha) - a ** (1 - alpha)) / (1 - alpha)))
        meta = cells.network_meta(NET)["photo"]
        for q in (0, 1, 2):  # H, He, He+: Verner fits, smooth
            lo_q, hi_q = meta[q]["Emin"], meta[q]["Emax"]
            for b in range(len(elo)):
                a, c = max(elo[b], lo_q), min(ehi[b], hi_q)
                if c <= a:
                    assert sig[q, b] == 0.0
                    continue
                want = quad(lambda e: N.photo_xsec(q, e)[0] * e ** -alpha,
                            a, c, limit=200, epsrel=1e-10)[0]
                assert abs(sig[q, b] * wint(elo[b], ehi[b]) - want) \
                    <= 1e-3 * abs(want), (alpha, q, b)
        # halves of the 25-40 eV bin
        mid = math.sqrt(25.0 * 40.0)
        h = oracle.photo_bins(NET, np.array([25.0, mid]),
                              np.array([mid, 40.0]), alpha)
        whole = sig[:, 3] * wint(25.0, 40.0)
        parts = h[:, 0] * wint(25.0, mid) + h[:, 1] * wint(mid, 40.0)
        nz = whole != 0
        assert rel(parts[nz], whole[nz]).max() < 1e-3


def test_oracle_rates_with_fields_vs_reference(oracle):
    """the oracle's rates over rows with the bin sums against the
    reference's get_rates with the same bins and fields"""
    g = golden("%s_photo.npz" % NET)
    sig = oracle.photo_bins(NET, g["Elo"], g["Ehi"], 1.0)
    rows = oracle.photo_rows(NET, g["fpar"], g["fPhi"], sig)
    k, _ = oracle.funcs(NET, g["fn"], g["fT"], rows, 1)
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)
    nz = kr != 0.0
    assert (rel(k[nz], kr[nz])).max() < 1e-9
    # the photo reactions' rates are the bin sums, nothing else
    ph = cells.network_meta(NET)["photo"]
    for q, p in enumerate(ph):
        s = g["fPhi"] @ sig[q]
        assert rel(k[:, p["reaction"]], s).max() < 1e-13


def test_oracle_solve_vs_reference(oracle):
    """the oracle's solve of the 256 fixture cells with their fields
    against the reference's solve_cells with the same bins (Tgas evolved,
    rtol 1e-8): flags equal, species within 1e-5, steps within 2 %"""
    g = golden("%s_photo.npz" % NET)
    sig = oracle.photo_bins(NET, g["Elo"], g["Ehi"], 1.0)
    rows = oracle.photo_rows(NET, g["par"], g["Phi"], sig)
    r, n, T, fl, st = oracle.solve(NET, g["n0"], g["T0"], g["dt"], rows,
                                   oopts(g["opts"]))
    assert r == 0
    for m in ("bitwise", "production"):
        assert np.array_equal(fl, g["flag_" + m])
        assert big_species_rel(n, g["n_" + m]) < 1e-5
        assert np.allclose(T, g["T_" + m], rtol=1e-12, atol=0)
        steps, ref = st[:, 0].sum(), g["stats_" + m][:, 0].sum()
        assert abs(steps - ref) <= 0.02 * ref


def test_photo_argument_checks():
    """the bins and field arguments (krome.cpp:313-377, 415-417)"""
    N = load()
    L = N._lib
    D = ctypes.POINTER(ctypes.c_double)
    h = L.krome_opts_new()
    try:
        lo, hi = np.array([1.0, 2.0]), np.array([2.0, 1.5])
        assert L.krome_opts_set_photo(h, 2, lo.ctypes.data_as(D),
                                      hi.ctypes.data_as(D)) == -1
        assert L.krome_opts_nbins(h) == 0
        lo, hi = np.array([0.0]), np.array([1.0])
        assert L.krome_opts_set_photo(h, 1, lo.ctypes.data_as(D),
                                      hi.ctypes.data_as(D)) == -1
        assert L.krome_opts_set_photo(h, 0, lo.ctypes.data_as(D),
                                      hi.ctypes.data_as(D)) == -1
        s = np.zeros(8)
        assert L.krome_opts_photo_sigma(h, s.ctypes.data_as(D)) == -1
        assert L.krome_opts_set_photo_dust_opacity(h, None) == -1
        lo, hi = N.bins_log(4, 1.0, 100.0)
        assert np.allclose(lo[1:], hi[:-1], rtol=0, atol=0)
        assert np.isclose(lo[0], 1.0, rtol=1e-15)
        assert np.isclose(hi[-1], 100.0, rtol=1e-14)
        assert np.allclose(np.log(hi / lo), math.log(100.0) / 4, rtol=1e-12)
        assert L.krome_opts_set_photo(h, 4, lo.ctypes.data_as(D),
                                      hi.ctypes.data_as(D)) == 0
        assert L.krome_opts_nbins(h) == 4
        assert L.krome_opts_set_photo_alpha(h, float("nan")) == -1
        assert L.krome_opts_set_photo_broadening(h, 0.0, 0.0) == 0
        assert L.krome_opts_set_photo_broadening(h, 100.0, 0.0) == -1
        assert L.krome_opts_set_photo_broadening(h, -1.0, 0.0) == -1
        assert L.krome_opts_set_photo_dust_opacity(h, None) == 0
        k = np.ones(4)
        assert L.krome_opts_set_photo_dust_opacity(
            h, k.ctypes.data_as(D)) == -1
        # bins set and no field: -103 before anything runs, flags
        # untouched (no GPU work)
        n = np.ones((2, N.nspec))
        T = np.full(2, 100.0)
        dt = np.ones(2)
        fl = np.full(2, 7, dtype=np.intc)
        r = L.krome_solve_cells(2, n.ctypes.data_as(D), T.ctypes.data_as(D),
                                dt.ctypes.data_as(D), None, 0, None,
                                fl.ctypes.data_as(ctypes.POINTER(
                                    ctypes.c_int)), h)
        assert r == -103 and (fl == 7).all()
        kk = np.zeros((2, N.nrea))
        assert L.krome_get_rates(2, n.ctypes.data_as(D), T.ctypes.data_as(D),
                                 None, 0, None, h,
                                 kk.ctypes.data_as(D)) == -103
    finally:
        L.krome_opts_free(h)
    assert N._lib.krome_photo_reaction(-1) == -1
    assert N._lib.krome_photo_reactant(N.nphoto) == -1


# ---- on the GPU ------------------------------------------------------------


def photo_opts(N, g, **kw):
    return N.options(**kw).set_photo(g["Elo"], g["Ehi"])


@pytest.mark.gpu
def test_gpu_rates_with_fields_vs_reference():
    """krome_get_rates with bins and fields, computed on the device (the
    photo rows kernel, then the rates), against the reference's"""
    g = golden("%s_photo.npz" % NET)
    N = load()
    o = photo_opts(N, g)
    say("photo rates", g["fn"].shape)
    k = N.rates(g["fn"], g["fT"], par=g["fpar"], Phi=g["fPhi"], opts=o)
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)
    nz = kr != 0.0
    assert rel(k[nz], kr[nz]).max() < 1e-9
    # no bins: the photo rates are 0 (the reference's rates without a
    # PhotoModule)
    k0 = N.rates(g["fn"], g["fT"], par=g["fpar"])
    ph = [p for p, _ in N.photo_reactions()]
    assert (k0[:, ph] == 0.0).all()
    other = np.setdiff1d(np.arange(N.nrea), ph)
    assert np.array_equal(k0[:, other], k[:, other])


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,mixed", [("plain", 0), ("persistent", 0),
                                          ("persistent", 1)])
def test_gpu_solve_with_fields_vs_reference(kernel, mixed, oracle):
    """the 256 fixture cells with their fields through krome_solve_cells:
    flags equal to the reference's; FP64 within 1e-5 of the reference and
    the oracle, the mixed-precision Newton matrix within 1e-4"""
    g = golden("%s_photo.npz" % NET)
    N = load()
    o = photo_opts(N, g, evolve_T=1, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                   linsol=1, jac=0, max_steps=100000, device_kernel=kernel,
                   mixed_precision=mixed)
    say("photo solve", kernel, mixed)
    res = N.solve(g["n0"], g["T0"], g["dt"], par=g["par"], Phi=g["Phi"],
                  opts=o, check=False)
    assert res.ierr == 0
    sig = oracle.photo_bins(NET, g["Elo"], g["Ehi"], 1.0)
    rows = oracle.photo_rows(NET, g["par"], g["Phi"], sig)
    _, no, To, flo, _ = oracle.solve(NET, g["n0"], g["T0"], g["dt"], rows,
                                     oopts(g["opts"]))
    bar = 1e-4 if mixed else 1e-5
    for m in ("bitwise", "production"):
        assert np.array_equal(res.flags, g["flag_" + m])
        assert big_species_rel(res.n, g["n_" + m]) < bar
    assert np.array_equal(res.flags, flo)
    assert big_species_rel(res.n, no) < bar
    assert np.allclose(res.Tgas, To, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_gpu_device_resident_fields():
    """krome_solve_cells_device with extra->Phi on the device: 65536 cells
    of the workload, the plain FP64 kernel against the persistent FP64
    kernel (flags equal, species within 1e-5: the two evaluate the same
    method in a different order of operations), and the first 512 against
    the oracle"""
    import torch
    N = load()
    w = cells.galaxy_ism_cells(65536)
    elo, ehi = w["bins"]
    dev = torch.device("cuda:0")
    out = {}
    for kernel in ("plain", "persistent"):
        o = N.options(evolve_T=1, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                      linsol=1, device_kernel=kernel).set_photo(elo, ehi)
        n = torch.tensor(w["n"].T.copy(), device=dev)
        T = torch.tensor(w["T"], device=dev)
        dt = torch.tensor(w["dt"], device=dev)
        par = torch.tensor(w["par"], device=dev)
        Phi = torch.tensor(w["Phi"], device=dev)
        fl = torch.zeros(65536, dtype=torch.int32, device=dev)
        say("photo device", kernel)
        r = N.solve_device(n, T, dt, par, opts=o, flags=fl, Phi=Phi)
        assert r == 0
        out[kernel] = (n.cpu().numpy().T, T.cpu().numpy(), fl.cpu().numpy())
    a, b = out["plain"], out["persistent"]
    assert np.array_equal(a[2], b[2]) and (a[2] == 0).all()
    assert big_species_rel(b[0], a[0]) < 1e-5
    import conftest
    orc = conftest.Oracle(conftest.REPO / "oracle" / "_build" /
                          "libkx_oracle.so", "kxo_")
    sig = orc.photo_bins(NET, elo, ehi, 1.0)
    m = 512
    rows = orc.photo_rows(NET, w["par"][:m], w["Phi"][:m], sig)
    _, no, _, flo, _ = orc.solve(NET, w["n"][:m], w["T"][:m], w["dt"][:m],
                                 rows, oracle_opts(**w["opts"]))
    assert np.array_equal(flo, a[2][:m])
    assert big_species_rel(a[0][:m], no) < 1e-5


# ---- photoheating: galaxy_ism with its thermal balance ------------------
# (galaxy_ism_heat: heating CHEM + PHOTO, cooling H2 + CHEM, Tgas evolved;
# tests/golden/galaxy_ism_heat_photo.npz from the reference's project with
# those thermal options)

HEAT = "galaxy_ism_heat"


def load_heat():
    so = PKG / ("libkrome_%s.so" % HEAT)
    if not so.exists():
        pytest.skip("%s not built" % so.name)
    import krome
    return krome.load(HEAT)


@pytest.mark.parametrize("alpha", [1.0, 0.0, 2.0])
def test_photoheating_tables_vs_reference(alpha, oracle):
    """the oracle's sigmaE (sigma times the photoelectron's energy E -
    Emin, bin-averaged) against PhotoModule::init's"""
    g = golden("%s_photo.npz" % HEAT)
    t = oracle.photo_tables(HEAT, g["Elo"], g["Ehi"], alpha)
    nph = t.shape[0] // 2
    assert rel(t[:nph], g["sigma_a%g" % alpha]).max() < 1e-14
    ref = g["sigmaE_a%g" % alpha]
    assert np.array_equal(t[nph:] == 0.0, ref == 0.0)
    assert rel(t[nph:], ref).max() < 1e-14


def test_oracle_thermal_with_fields_vs_reference(oracle):
    """dTgas/dt of the oracle (CHEM heating with the Solomon pumping, the
    photoheating of the fields, H2 and CHEM cooling) against the
    reference's get_thermal, and the photoheating against the sum of its
    get_photo_heating per reaction"""
    g = golden("%s_photo.npz" % HEAT)
    t = oracle.photo_tables(HEAT, g["Elo"], g["Ehi"], 1.0)
    rows = oracle.photo_rows(HEAT, g["fpar"], g["fPhi"], t)
    k, f = oracle.funcs(HEAT, g["fn"], g["fT"], rows, 1)
    th = g["thermal"]
    assert rel(f[:, -1], th[:, 2]).max() < 1e-12
    assert rel(g["photo_heat"].sum(axis=1), th[:, 3 + 1]).max() < 1e-14
    nz = g["k"] != 0.0
    assert rel(k[nz], g["k"][nz]).max() < 1e-9


def heat_solve_check(n, T, fl, g, bar):
    """the cells both solves finish within bar; the reference's cells
    trapped at max_steps (flag -1, Tgas cycling at a break of a rate fit)
    trapped here too, with at most two more"""
    ref_fl = g["flag_bitwise"]
    assert set(np.unique(fl)) <= {0, -1}
    trapped, ref_trapped = set(np.flatnonzero(fl)), set(np.flatnonzero(ref_fl))
    assert len(trapped ^ ref_trapped) <= 2, (trapped, ref_trapped)
    both = (fl == 0) & (ref_fl == 0)
    assert both.sum() >= len(fl) - 8
    nr = g["n_bitwise"][both]
    assert big_species_rel(n[both], nr) < bar
    assert rel(T[both], g["T_bitwise"][both]).max() < bar


def test_oracle_photoheating_solve_vs_reference(oracle):
    g = golden("%s_photo.npz" % HEAT)
    t = oracle.photo_tables(HEAT, g["Elo"], g["Ehi"], 1.0)
    rows = oracle.photo_rows(HEAT, g["par"], g["Phi"], t)
    r, n, T, fl, st = oracle.solve(HEAT, g["n0"], g["T0"], g["dt"], rows,
                                   oopts(g["opts"]))
    heat_solve_check(n, T, fl, g, 1e-5)


@pytest.mark.gpu
def test_gpu_thermal_with_fields_vs_reference():
    """krome_get_thermal with bins and fields, on the device: the active
    terms (heat CHEM, PHOTO; cool H2, CHEM) and dT/dt against the
    reference's get_thermal"""
    g = golden("%s_photo.npz" % HEAT)
    N = load_heat()
    assert N.heating_terms == ("CHEM", "PHOTO")
    assert N.cooling_terms == ("H2", "CHEM")
    o = photo_opts(N, g)
    say("photo thermal")
    out = N.thermal(g["fn"], g["fT"], par=g["fpar"], Phi=g["fPhi"], opts=o)
    th = g["thermal"]
    for got, want in ((out["heats"][:, 0], th[:, 3]),
                      (out["heats"][:, 1], th[:, 4]),
                      (out["cools"][:, 0], th[:, 15]),
                      (out["dTdt"], th[:, 2])):
        assert rel(got, want).max() < 1e-10
    # the CHEM cooling's terms are as small as 1e-214 here: to the scale
    assert np.max(np.abs(out["cools"][:, 1] - th[:, 16])
                  / np.maximum(np.abs(th[:, 15:17]).max(axis=1),
                               1e-300)) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,mixed", [("plain", 0), ("persistent", 0),
                                          ("persistent", 1)])
def test_gpu_photoheating_solve_vs_reference(kernel, mixed):
    g = golden("%s_photo.npz" % HEAT)
    N = load_heat()
    o = photo_opts(N, g, evolve_T=1, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                   linsol=1, jac=0, max_steps=100000, device_kernel=kernel,
                   mixed_precision=mixed)
    say("photoheating solve", kernel, mixed)
    res = N.solve(g["n0"], g["T0"], g["dt"], par=g["par"], Phi=g["Phi"],
                  opts=o, check=False)
    heat_solve_check(res.n, res.Tgas, res.flags, g, 1e-4 if mixed else 1e-5)


def test_solve_signature_as_the_reference_binding():
    """Network.solve takes the reference binding's arguments
    (runtime/interfaces/python/krome.py:458-507): Td and adust, refused
    here before any work (no dust bins); a field without bins is refused
    too"""
    import krome
    N = load()
    n = np.ones(N.nspec)
    with pytest.raises(krome.KromeError):
        N.solve(n, 100.0, 1.0, Td=np.ones(2))
    with pytest.raises(krome.KromeError):
        N.solve(n, 100.0, 1.0, adust=np.ones(2))
    with pytest.raises(krome.KromeError):
        N.solve(n, 100.0, 1.0, Phi=np.ones(32))


@pytest.mark.gpu
def test_gpu_field_forms():
    """the other forms of the field argument on the device: one Phi row for
    every cell equals that row repeated; krome_solve_cells_x with fields
    equals the density solve converted to mass fractions; the fluxes with
    fields are k prod(n) of the rates with fields"""
    g = golden("%s_photo.npz" % NET)
    N = load()
    m = 64
    o = photo_opts(N, g, evolve_T=1, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                   linsol=1, device_kernel="plain")
    row = g["Phi"][5]
    say("field forms")
    a = N.solve(g["n0"][:m], g["T0"][:m], g["dt"][:m], par=g["par"][:m],
                Phi=row, opts=o, check=False)
    b = N.solve(g["n0"][:m], g["T0"][:m], g["dt"][:m], par=g["par"][:m],
                Phi=np.tile(row, (m, 1)), opts=o, check=False)
    assert a.ierr == b.ierr == 0
    assert np.array_equal(a.n, b.n) and np.array_equal(a.Tgas, b.Tgas)
    assert a.Phi is not None and a.Td is None
    # mass fractions: x = m_i n_i / rho, solved with the same fields
    mass = np.array(N.masses)
    n0 = g["n0"][:m]
    rho = n0 @ mass
    x0 = n0 * mass / rho[:, None]
    c = N.solve(n0, g["T0"][:m], g["dt"][:m], par=g["par"][:m],
                Phi=g["Phi"][:m], opts=o, check=False)
    xs = N.solve_x(x0, rho, g["T0"][:m], g["dt"][:m], par=g["par"][:m],
                   Phi=g["Phi"][:m], opts=o, check=False)
    assert c.ierr == xs.ierr == 0
    xn = c.n * mass / rho[:, None]
    big = xn > 1e-12
    assert rel(xs.n[big], xn[big]).max() < 1e-6
    # fluxes with fields
    k = N.rates(g["fn"], g["fT"], par=g["fpar"], Phi=g["fPhi"], opts=o)
    fl = N.fluxes(g["fn"], g["fT"], par=g["fpar"], Phi=g["fPhi"], opts=o)
    meta = cells.network_meta(NET)
    sp = {s: i for i, s in enumerate(meta["species"])}
    want = np.array(k, copy=True)
    for r, eq in enumerate(meta["reactions"]):
        for t in eq.split("->")[0].split(" + "):
            want[:, r] *= g["fn"][:, sp[t.strip()]]
    nz = want != 0.0
    assert np.array_equal(fl == 0.0, want == 0.0)
    assert rel(fl[nz], want[nz]).max() < 1e-12
