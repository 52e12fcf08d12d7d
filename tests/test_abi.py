"""The C ABI (include/krome.h) without a GPU: every declared symbol is
exported by each network library, the libraries load, and the host-side
logic (network description, option keys and their validation, argument
checks that return before any device work) behaves as the reference's
(runtime/interfaces/C/krome.cpp:198-311, 396-436)."""
import ctypes
import math
import re
import subprocess

import numpy as np
import pytest

import cells
from conftest import PKG, REPO

NETS = ("primordial", "primordial3", "galaxy_ism", "primordialZ",
        "galaxy_ism_heat")


def declared():
    text = (REPO / "include" / "krome.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(krome_[a-z_]+)\s*\(", text)))


def lib(net):
    so = PKG / ("libkrome_%s.so" % net)
    if not so.exists():
        pytest.skip("%s not built" % so.name)
    return so


@pytest.mark.parametrize("net", NETS)
def test_exports_every_declared_symbol(net):
    so = lib(net)
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (krome_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    assert len(declared()) >= 25


@pytest.fixture(params=NETS)
def net(request):
    import krome
    lib(request.param)
    return krome.load(request.param)


def test_network_description(net):
    meta = cells.network_meta(net.name)
    assert net.species == meta["species"]
    assert net.params == meta["params"]
    assert net.nrea == meta["nrea"]
    d = net.params_default()
    for a, b in zip(d, meta["param_defaults"]):
        assert (math.isnan(a) and math.isnan(b)) or a == b
    assert net._lib.krome_species_index(b"e-") == net.species.index("e-")
    assert net._lib.krome_species_index(b"nope") == -1
    if net.name == "primordial3":
        assert net.required == {"tff": "thermal.heating = ['COMPRESS']"}
    assert net.flag_message(-1).startswith("max_steps")
    assert net.flag_name(-110) == "FLAG_NONPOSITIVE_T"
    assert net._lib.krome_has_device() == 1


def test_options(net):
    o = net.options()
    assert o["rtol"] == 1e-8 and o["atol_n"] == 1e-20 and o["atol_T"] == 1e-6
    assert o["max_steps"] == 100000 and o["max_restarts"] == 5
    assert o["clip_negative"] == 1 and o["error_report"] == 1
    o["rtol"] = 1e-6
    assert o["rtol"] == 1e-6
    o["jac"] = "fd"
    assert o["jac"] == 2
    L = net._lib
    assert L.krome_opts_set(o._h, b"jac", 3.0) == -1
    assert L.krome_opts_set(o._h, b"conserve", 4.0) == -1
    assert L.krome_opts_set(o._h, b"max_restarts", -1.0) == -1
    assert L.krome_opts_set(o._h, b"no_such_key", 1.0) == -1
    o["device_kernel"] = "persistent"
    o["mixed_precision"] = 1
    assert o["device_kernel"] == 1 and o["mixed_precision"] == 1


def test_argument_checks_without_device_work(net):
    """rejections that return before the device is touched"""
    L = net._lib
    o = net.options()
    D = ctypes.POINTER(ctypes.c_double)
    n = np.ones((2, net.nspec))
    T = np.ones(2)
    dt = np.ones(2)
    par = np.tile(net.params_default(), (2, 1))
    # bad par_stride, negative ncell, null arrays: -1
    assert L.krome_solve_cells(2, n.ctypes.data_as(D), T.ctypes.data_as(D),
                               dt.ctypes.data_as(D), par.ctypes.data_as(D),
                               net.npar + 1, None, None, o._h) == -1
    assert L.krome_solve_cells(-1, n.ctypes.data_as(D), T.ctypes.data_as(D),
                               dt.ctypes.data_as(D), None, 0, None, None,
                               o._h) == -1
    assert L.krome_solve_cells(2, None, T.ctypes.data_as(D),
                               dt.ctypes.data_as(D), None, 0, None, None,
                               o._h) == -1
    assert L.krome_solve_cells(0, n.ctypes.data_as(D), T.ctypes.data_as(D),
                               dt.ctypes.data_as(D), None, 0, None, None,
                               o._h) == 0
    if net.name == "primordial3":
        # the default tff is unset (NaN): -105, the cell flags set, nothing
        # integrated (solve_cells.hpp:133-140)
        fl = np.full(2, 7, dtype=np.intc)
        r = L.krome_solve_cells(2, n.ctypes.data_as(D), T.ctypes.data_as(D),
                                dt.ctypes.data_as(D), None, 0, None,
                                fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                o._h)
        assert r == -105 and (fl == -105).all()
        assert (n == 1).all()
    # the host per-cell path does not exist in this build: -104
    o["path"] = "percell"
    par[:, :] = np.nan_to_num(par, nan=1e10)
    fl = np.zeros(2, dtype=np.intc)
    r = L.krome_solve_cells(2, n.ctypes.data_as(D), T.ctypes.data_as(D),
                            dt.ctypes.data_as(D), par.ctypes.data_as(D),
                            net.npar, None,
                            fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            o._h)
    assert r == -104 and (fl == -104).all()


def test_element_metadata_vs_reference(net):
    """element names, atom counts, masses and charges equal the tables of
    the reference's generated project (tests/golden/*_conserve.npz,
    galaxy_ism(_heat)_tables.npz, primordialZ_tables.npz)"""
    from conftest import golden
    g = golden("%s_%s.npz" % (net.name, "tables" if net.name in (
        "galaxy_ism", "primordialZ", "galaxy_ism_heat") else "conserve"))
    assert net.elements == tuple(g["element_names"])
    assert np.array_equal(net.composition, g["species_elements"])
    assert np.array_equal(net.masses, g["species_mass_g"])
    assert np.array_equal(net.charges, g["species_charge"])
    L = net._lib
    assert L.krome_element_name(-1) is None
    assert L.krome_element_name(len(net.elements)) is None
    assert L.krome_species_natoms(net.nspec, 0) == -1
    assert L.krome_species_natoms(0, len(net.elements)) == -1
    assert math.isnan(L.krome_species_mass(net.nspec))
    assert L.krome_species_charge(-1) == 0


def test_conserve_argument_checks_without_device_work(net):
    """krome.cpp:929-977, 1306-1343: -1 on bad arguments, 0 for no cells"""
    L = net._lib
    D = ctypes.POINTER(ctypes.c_double)
    nel = len(net.elements)
    n = np.ones((2, net.nspec))
    ref = np.ones((2, nel))
    tot = np.zeros((2, nel))
    nd, rd = n.ctypes.data_as(D), ref.ctypes.data_as(D)
    assert L.krome_get_element_totals(-1, nd, tot.ctypes.data_as(D)) == -1
    assert L.krome_get_element_totals(2, None, tot.ctypes.data_as(D)) == -1
    assert L.krome_get_element_totals_x(2, nd, None) == -1
    assert L.krome_get_element_totals(0, nd, tot.ctypes.data_as(D)) == 0
    for f in (L.krome_conserve_cells, L.krome_conserve_cells_x):
        assert f(-1, nd, rd, nel, None, 0, None) == -1
        assert f(2, None, rd, nel, None, 0, None) == -1
        assert f(2, nd, None, nel, None, 0, None) == -1
        assert f(2, nd, rd, nel + 1, None, 0, None) == -1
        assert f(2, nd, rd, -1, None, 0, None) == -1
        assert f(2, nd, rd, nel, None, -1, None) == -1
        assert f(0, nd, rd, nel, None, 0, None) == 0
    p = ctypes.c_void_p(16)  # never dereferenced: rejected or no cells
    for f in (L.krome_conserve_cells_device, L.krome_conserve_cells_x_device):
        assert f(-1, p, 0, 0, p, nel, None, 0, None) == -1
        assert f(2, None, 0, 0, p, nel, None, 0, None) == -1
        assert f(2, p, 0, 0, None, nel, None, 0, None) == -1
        assert f(2, p, 0, 0, p, nel + 1, None, 0, None) == -1
        assert f(2, p, 0, 0, p, nel, p, net.npar + 1, None) == -1
        assert f(2, p, 2, 0, p, nel, None, 0, None) == -1
        assert f(2, p, 0, 1, p, nel, None, 0, None) == -1
        assert f(2, p, 1, 1, p, nel, None, 0, None) == -1
        assert f(0, p, 0, 0, p, nel, None, 0, None) == 0
