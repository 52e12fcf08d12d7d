"""The generator on the other shipped networks it covers (NO, SM, coolHCN,
lowmetal, primordial2, primordial_wD, primordial_photo,
primordial_photoH2, auto, chianti, ions, ionsH2, primordial_UV, xrays,
primordialZ_UV, photobin; krome-x_amd/netgen.py EXTRA), each at the reference's
default thermal options (Tgas fixed), pinned against the reference's own
generated project (tests/golden/<net>_extra.npz, or <net>_photo.npz with
photon bins and fields; tests/golden/make_golden.py extra):

- the rates at 136 states (the generic cells and a Tgas grid across the
  fits' breaks) within 1e-9 relative with the same zeros (netcheck's bar,
  kromex/netcheck.py:676-717), and the species rhs within 1e-10 of its
  row's scale;
- whole solves of 256 generic cells (cells.generic_cells: rtol 1e-6, the
  difference-quotient Jacobian where a rate reads the densities): the
  oracle against the reference's solve_cells, flags equal, species above
  1e-12 of the cell total within 1e-3 (at rtol 1e-6 the last bits of the
  rate expressions move trace species by up to several hundred rtol; NO
  and SM are bitwise, the largest are 5.8e-4 and 4.7e-4, He++ at 5e-11
  and 2.5e-7 of the cell total in xrays and primordialZ_UV),
  steps within 5 %;
- on the GPU, the plain and persistent FP64 kernels against the reference
  and the oracle on the same cells, flags equal, within 1e-3 (measured
  2.6e-5 to 6.2e-4, the largest xrays' He++ as for the oracle); SM within
  2e-3 (measured 1.3e-3 plain, 5.3e-4 persistent); networks above 16
  equations (primordialZ_UV, ions, ionsH2) run in the persistent lane
  queue (the plain kernel's solve, a cell per lane from the device work
  queue).  SM's rates are constants and its cells run up to 45307
  steps into stiff equilibria, where the order and the fusion of the
  floating-point operations (which the GPU kernels do not share with the
  reference's build, while the oracle is bitwise equal to it there) move
  the trace species that far.
"""
import sys

import numpy as np
import pytest

import cells
from conftest import PKG, golden, oracle_opts

NETS = cells.EXTRA_NETWORKS
GPU_BAR = {"SM": 2e-3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture(net):
    if cells.network_meta(net)["photo"]:
        return golden("%s_photo.npz" % net)
    return golden("%s_extra.npz" % net)


def oopts(o):
    return oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                       linsol=o[4], jac=o[5], max_steps=o[6])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    d[(a == 0) & (b == 0)] = 0.0
    return d


def big_species_rel(n, nr):
    big = nr > 1e-12 * nr.sum(axis=1, keepdims=True)
    return rel(n, nr)[big].max()


def rows(oracle, net, g, par_key, phi_key):
    """the oracle's parameter rows (with the photo sums for a photo
    network)"""
    if not cells.network_meta(net)["photo"]:
        return g[par_key]
    t = oracle.photo_tables(net, g["Elo"], g["Ehi"], 1.0)
    return oracle.photo_rows(net, g[par_key], g[phi_key], t)


@pytest.mark.parametrize("net", NETS)
def test_rates_and_rhs_vs_reference(oracle, net):
    g = fixture(net)
    k, f = oracle.funcs(net, g["fn"], g["fT"],
                        rows(oracle, net, g, "fpar", "fPhi"), 0)
    kr = g["k"]
    assert np.array_equal(k == 0.0, kr == 0.0)
    nz = kr != 0.0
    assert rel(k[nz], kr[nz]).max() < 1e-9
    if "f" in g.files:
        ns = len(cells.network_meta(net)["species"])
        fr = g["f"][:, :ns]
        scale = np.maximum(np.abs(fr).max(axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(f[:, :ns] - fr) / scale) < 1e-10


@pytest.mark.parametrize("net", NETS)
def test_oracle_solve_vs_reference(oracle, net):
    g = fixture(net)
    r, n, T, fl, st = oracle.solve(net, g["n0"], g["T0"], g["dt"],
                                   rows(oracle, net, g, "par", "Phi"),
                                   oopts(g["opts"]))
    assert r == 0 and np.array_equal(fl, g["flag_bitwise"])
    assert big_species_rel(n, g["n_bitwise"]) < 1e-3
    steps, ref = st[:, 0].sum(), g["stats_bitwise"][:, 0].sum()
    assert abs(steps - ref) <= 0.05 * ref


@pytest.fixture(scope="module")
def libs():
    import krome
    return {n: krome.load(n) for n in NETS
            if (PKG / ("libkrome_%s.so" % n)).exists()}


@pytest.mark.gpu
@pytest.mark.parametrize("net", ["coolHCN", "primordial2", "primordial_wD"])
def test_gpu_mixed_keeps_fp64_at_fixed_tgas(libs, net):
    """with Tgas fixed the persistent kernel keeps the FP64 Newton matrix
    whatever mixed_precision says (its FP32 form failed cells of these
    networks): the results are the FP64 kernel's, bitwise"""
    g = fixture(net)
    N = libs[net]
    o = g["opts"]
    out = []
    for mixed in (0, 1):
        opts = N.options(evolve_T=False, rtol=o[1], atol_n=o[2], atol_T=o[3],
                         jac=int(o[5]), linsol=1, device_kernel="persistent",
                         mixed_precision=mixed, error_report=False)
        say("fixed-Tgas mixed", net, mixed)
        out.append(N.solve(g["n0"], g["T0"], g["dt"], par=g["par"],
                           opts=opts, check=False))
    assert out[0].ierr == out[1].ierr == 0
    assert np.array_equal(out[0].n, out[1].n)
    assert np.array_equal(out[0].flags, out[1].flags)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["plain", "persistent"])
@pytest.mark.parametrize("net", NETS)
def test_gpu_solve_vs_reference(libs, oracle, net, kernel):
    if net not in libs:
        # ions, ionsH2: built on request only (krome-x_amd/build.py)
        pytest.skip("libkrome_%s.so is built on request only" % net)
    g = fixture(net)
    N = libs[net]
    o = g["opts"]
    opts = N.options(evolve_T=False, rtol=o[1], atol_n=o[2], atol_T=o[3],
                     jac=int(o[5]), linsol=1, device_kernel=kernel,
                     error_report=False)
    phi = None
    if cells.network_meta(net)["photo"]:
        opts.set_photo(g["Elo"], g["Ehi"])
        phi = g["Phi"]
    say("network", net, kernel)
    res = N.solve(g["n0"], g["T0"], g["dt"], par=g["par"], Phi=phi,
                  opts=opts, check=False)
    assert res.ierr == 0 and np.array_equal(res.flags, g["flag_bitwise"])
    bar = GPU_BAR.get(net, 1e-3)
    assert big_species_rel(res.n, g["n_bitwise"]) < bar
    _, no, _, flo, _ = oracle.solve(net, g["n0"], g["T0"], g["dt"],
                                    rows(oracle, net, g, "par", "Phi"),
                                    oopts(g["opts"]))
    assert big_species_rel(res.n, no) < bar
