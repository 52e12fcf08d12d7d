"""Element totals and consistent advection (krome_get_element_totals(_x),
krome_conserve_cells(_x)(_device), the solve's post-solve conservation
mode 2): the oracle pinned, then the HIP kernels against it.

Oracle (CPU): bit for bit against the reference's own outputs
(tests/golden/*_conserve.npz, written by tests/golden/make_golden.py from
the reference built without contraction), and against what the
mathematics fixes: a pass without clipping is the weighted least-change
projection n + W A^T (A W A^T)^-1 (t - A n), W = diag(n), solved here
with numpy; the corrected cells meet their totals and are neutral; the
flags follow from the corrections.

GPU (through the C ABI): the kernels compute the reference's operations
in its order without contraction, so totals, corrected rows and flags are
bit-identical to the reference and to the oracle; the post-solve mode 2
(an integration, not bitwise) within the solve tolerance of the
reference's solve_cells with conserve 2.
"""
import sys

import numpy as np
import pytest

import cells
from conftest import golden, oracle_opts, rel_err

NETS = ("primordial", "primordial3")
# the solve fixtures' agreement bars (tests/test_oracle.py)
TOL = {"primordial": 5e-5, "primordial3": 1e-6}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rows(g):
    """conserved rows A[NELEM + 1, NSPEC]: the elements, then the charge"""
    return np.vstack([g["species_elements"].T,
                      g["species_charge"][None, :]]).astype(float)


def residual(A, n, t):
    """per-row |t - A n| / (|t| + sum |A| n), the reference's measure"""
    tot = n @ A.T
    sc = np.abs(t) + np.abs(n[:, None, :] * A[None]).sum(-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(sc > 0, np.abs(t - tot) / sc, 0.0)
    return r


# ---- the oracle ------------------------------------------------------------

@pytest.mark.parametrize("net", NETS)
def test_oracle_totals_vs_reference(oracle, net):
    g = golden("%s_conserve.npz" % net)
    assert np.array_equal(oracle.totals(net, g["in_n"]), g["tot_p"])
    assert np.array_equal(oracle.totals(net, g["in_n"], x=True), g["totx_p"])


@pytest.mark.parametrize("mode", ["n", "x"])
@pytest.mark.parametrize("net", NETS)
def test_oracle_conserve_vs_reference(oracle, net, mode):
    g = golden("%s_conserve.npz" % net)
    r, v, fl = oracle.conserve(net, g["in_" + mode], g["ref_" + mode],
                               x=mode == "x")
    assert np.array_equal(fl, g["flag_" + mode])
    assert np.array_equal(v, g["out_" + mode])
    neg = fl[fl < 0]
    assert r == (neg[0] if neg.size else 0)
    # every outcome is exercised
    assert set(np.unique(fl)) == {0, 2, -112}


@pytest.mark.parametrize("net", NETS)
def test_oracle_projection_closed_form(oracle, net):
    """cells whose element pass needs no clipping and no skipped pivot:
    the oracle's result is the textbook weighted projection, electrons
    then set to the charge of the other species"""
    g = golden("%s_conserve.npz" % net)
    A = rows(g)
    ne = A.shape[0] - 1
    ie = cells.network_meta(net)["species"].index("e-")
    _, v, fl = oracle.conserve(net, g["in_n"], g["ref_n"])
    checked = 0
    for c in range(v.shape[0]):
        n = np.maximum(g["in_n"][c], 0.0)
        Ae = A[:ne]
        M = (Ae * n[None, :]) @ Ae.T
        if np.min(np.diag(M)) <= 1e-12 * np.max(np.diag(M)):
            continue  # a row near the pivot threshold (or without carriers)
        lam = np.linalg.solve(M, g["ref_n"][c] - Ae @ n)
        n1 = n + n * (Ae.T @ lam)
        if (n1 < 0).any():
            continue
        qi = np.delete(A[ne] * n1, ie).sum()
        if qi < 0:
            continue
        n1[ie] = qi
        scale = n1.sum()
        assert np.max(np.abs(v[c] - n1)) <= 1e-12 * scale, c
        checked += 1
    assert checked >= 100


@pytest.mark.parametrize("mode", ["n", "x"])
@pytest.mark.parametrize("net", NETS)
def test_oracle_conserve_invariants(oracle, net, mode):
    """reached cells meet the element totals and are neutral; the LARGE
    warning exactly where a species moved by over 50 %; cells without any
    helium carrier (perturbation group 5) cannot reach a helium total"""
    g = golden("%s_conserve.npz" % net)
    A = rows(g)
    m = g["species_mass_g"]
    v0 = g["in_" + mode]
    _, v, fl = oracle.conserve(net, v0, g["ref_" + mode], x=mode == "x")
    if mode == "x":  # per gram: densities n = x / m
        v, v0 = v / m, v0 / m
    t = np.hstack([g["ref_" + mode], np.zeros((v.shape[0], 1))])
    res = residual(A, v, t).max(axis=1)
    ok = fl >= 0
    assert (res[ok] <= 1e-12).all() and (res[~ok] > 1e-12).all()
    n0 = np.maximum(v0, 0.0)
    big = ((n0 > 0) & (np.abs(v - n0) > 0.5 * n0)).any(axis=1)
    assert np.array_equal(fl[ok] == 2, big[ok])
    he = [i for i, s in enumerate(cells.network_meta(net)["species"])
          if s.startswith("He")]
    grp5 = (np.arange(v.shape[0]) % 8) == 5
    assert (fl[grp5] == -112).all()
    assert (v[grp5][:, he] == 0).all()


@pytest.mark.parametrize("net", NETS)
def test_oracle_solve_conserve2_vs_reference(oracle, net):
    g = golden("%s_conserve.npz" % net)
    o = g["solve_opts"]
    r, n, T, fl, st = oracle.solve(
        net, g["solve_n0"], g["solve_T0"], g["solve_dt"], g["solve_par"],
        oracle_opts(evolve_T=o[0], rtol=o[1], atol_n=o[2], atol_T=o[3],
                    linsol=o[4], jac=o[5], max_steps=o[6], conserve=o[7]))
    assert r == 0 and np.array_equal(fl, g["solve_flag"])
    assert rel_err(n, g["solve_n"]) < TOL[net]
    assert np.max(np.abs(T - g["solve_T"]) / g["solve_T"]) < TOL[net]
    # the projection leaves the input totals in the same state as the
    # reference's (deuterium at 1e-20 nH stays below the pivot threshold
    # in primordial: that row is left as the integration made it)
    A = rows(g)
    t0 = g["solve_n0"] @ A.T
    ra, rb = residual(A, n, t0), residual(A, g["solve_n"], t0)
    keep = rb.max(axis=0) < 1e-12
    assert (ra[:, keep] < 1e-12).all()
    assert keep.sum() >= A.shape[0] - 1


# ---- the HIP kernels -------------------------------------------------------

@pytest.fixture(scope="module")
def nets():
    import torch
    assert torch.cuda.is_available()
    import krome
    return {n: krome.load(n) for n in NETS}


@pytest.mark.gpu
@pytest.mark.parametrize("net", NETS)
def test_gpu_totals_bitwise(nets, net):
    g = golden("%s_conserve.npz" % net)
    N = nets[net]
    say("totals", net)
    assert np.array_equal(N.element_totals(g["in_n"]), g["tot_p"])
    assert np.array_equal(N.element_totals_x(g["in_n"]), g["totx_p"])
    one = N.element_totals(g["in_n"][3])
    assert one.shape == (len(N.elements),)
    assert np.array_equal(one, g["tot_p"][3])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["n", "x"])
@pytest.mark.parametrize("net", NETS)
def test_gpu_conserve_bitwise(nets, oracle, net, mode):
    g = golden("%s_conserve.npz" % net)
    N = nets[net]
    say("conserve", net, mode)
    fl = np.zeros(g["in_n"].shape[0], dtype=np.intc)
    f = N.conserve_x if mode == "x" else N.conserve
    v = f(g["in_" + mode], g["ref_" + mode], flags=fl, check=False)
    assert np.array_equal(fl, g["flag_" + mode])
    assert np.array_equal(v, g["out_" + mode])
    with pytest.raises(Exception):
        f(g["in_" + mode], g["ref_" + mode])  # unreached cells raise
    # one reference row for every cell (ref_stride 0) = that row repeated
    ref = g["ref_" + mode][0]
    v1 = f(g["in_" + mode], ref, flags=fl, check=False)
    r, vo, flo = oracle.conserve(net, g["in_" + mode],
                                 np.tile(ref, (fl.size, 1)), x=mode == "x")
    assert np.array_equal(v1, vo) and np.array_equal(fl, flo)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("net", NETS)
def test_gpu_conserve_device_layouts(nets, oracle, net, layout):
    """device arrays: species-major with ld > ncell, or cell-major; ragged
    cell counts; flags optional; the return value is the lowest failing
    cell's flag"""
    import torch
    g = golden("%s_conserve.npz" % net)
    N = nets[net]
    L = N._lib
    for nc in (1, 63, 65, g["in_n"].shape[0]):
        say("device layout", net, layout, nc)
        v0 = g["in_n"][:nc]
        ref = torch.tensor(g["ref_n"][:nc], device="cuda")
        ld = nc + 5
        if layout == 0:
            buf = torch.full((N.nspec, ld), -7.0, dtype=torch.float64,
                             device="cuda")
            buf[:, :nc] = torch.tensor(v0.T, device="cuda")
        else:
            buf = torch.tensor(v0, device="cuda")
        fl = torch.zeros(nc, dtype=torch.int32, device="cuda")
        r = L.krome_conserve_cells_device(nc, buf.data_ptr(), layout, ld,
                                          ref.data_ptr(), len(N.elements),
                                          None, 0, fl.data_ptr())
        torch.cuda.synchronize()
        ro, vo, flo = oracle.conserve(net, v0, g["ref_n"][:nc])
        got = (buf[:, :nc].T if layout == 0 else buf).cpu().numpy()
        assert r == ro
        assert np.array_equal(fl.cpu().numpy(), flo)
        assert np.array_equal(got, vo)
        if layout == 0:  # the padding is untouched
            assert (buf[:, nc:] == -7.0).all()
        # no flags array: the same rows, the same return value
        buf2 = torch.tensor(v0.T.copy() if layout == 0 else v0,
                            device="cuda")
        r2 = L.krome_conserve_cells_device(nc, buf2.data_ptr(), layout, 0,
                                           ref.data_ptr(), len(N.elements),
                                           None, 0, None)
        torch.cuda.synchronize()
        got2 = (buf2.T if layout == 0 else buf2).cpu().numpy()
        assert r2 == ro and np.array_equal(got2, vo)


@pytest.mark.gpu
def test_gpu_conserve_large_batch(nets, oracle):
    """2^20 primordial3 cells (the benchmark's workload, each species
    perturbed by up to 30 %, one in 16 with a negative species) against the
    oracle, bit for bit; the x form through the device entry point"""
    import torch
    net = "primordial3"
    N = nets[net]
    nc = 1 << 20
    w = cells.primordial3_cells(nc)
    ref = N.element_totals(w["n"])
    rng = np.random.default_rng(7)
    v0 = w["n"] * (1.0 + 0.3 * (2.0 * rng.random(w["n"].shape) - 1.0))
    neg = np.arange(nc) % 16 == 3
    v0[neg, 2] = -1e-3 * v0[neg, 2]
    say("large batch", nc)
    vt = torch.tensor(v0.T.copy(), device="cuda")
    rt = torch.tensor(ref, device="cuda")
    fl = torch.zeros(nc, dtype=torch.int32, device="cuda")
    r = N.conserve_device(vt, rt, flags=fl)
    ro, vo, flo = oracle.conserve(net, v0, ref)
    assert r == ro
    assert np.array_equal(fl.cpu().numpy(), flo)
    assert np.array_equal(vt.T.cpu().numpy(), vo)
    # mass fractions on the device, cell-major
    m = N.masses
    x0 = v0 * m / (w["n"] @ m)[:, None]
    refx = N.element_totals_x(w["n"] * m / (w["n"] @ m)[:, None])
    xt = torch.tensor(x0, device="cuda")
    r = N.conserve_device(xt, torch.tensor(refx, device="cuda"), flags=fl,
                          x=True, layout=1)
    ro, xo, flo = oracle.conserve(net, x0, refx, x=True)
    assert r == ro
    assert np.array_equal(fl.cpu().numpy(), flo)
    assert np.array_equal(xt.cpu().numpy(), xo)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,mixed", [("plain", 0), ("persistent", 0),
                                          ("persistent", 1)])
@pytest.mark.parametrize("net", NETS)
def test_gpu_solve_conserve2(nets, oracle, net, kernel, mixed):
    """the solve with conserve 2 (refused before): flags equal to the
    reference's, states within the solve bar of its conserve-2 solve, and
    the conserved totals of the input kept to the projection's 1e-13"""
    g = golden("%s_conserve.npz" % net)
    N = nets[net]
    o = g["solve_opts"]
    say("solve conserve 2", net, kernel, mixed)
    opts = N.options(evolve_T=bool(o[0]), rtol=o[1], atol_n=o[2],
                     atol_T=o[3], conserve=2, device_kernel=kernel,
                     mixed_precision=mixed, error_report=False)
    res = N.solve(g["solve_n0"], g["solve_T0"], g["solve_dt"],
                  par=g["solve_par"], opts=opts, check=False)
    assert res.ierr == 0 and np.array_equal(res.flags, g["solve_flag"])
    tol = max(1e-5, TOL[net]) if mixed else TOL[net]
    assert rel_err(res.n, g["solve_n"]) < tol
    assert np.max(np.abs(res.Tgas - g["solve_T"]) / g["solve_T"]) < tol
    A = rows(g)
    t0 = g["solve_n0"] @ A.T
    ra, rb = residual(A, res.n, t0), residual(A, g["solve_n"], t0)
    keep = rb.max(axis=0) < 1e-12
    assert (ra[:, keep] < 1e-12).all()
    # conserve 2 on the device entry point too: a dt <= 0 cell untouched
    import torch
    n0 = g["solve_n0"].copy()
    dt = g["solve_dt"].copy()
    dt[1] = 0.0
    nt = torch.tensor(n0.T.copy(), device="cuda")
    Tt = torch.tensor(g["solve_T0"], device="cuda")
    r = N.solve_device(nt, Tt, torch.tensor(dt, device="cuda"),
                       torch.tensor(g["solve_par"], device="cuda"),
                       opts=opts)
    assert r == 0
    got = nt.T.cpu().numpy()
    assert np.array_equal(got[1], n0[1])
    ok = np.arange(n0.shape[0]) != 1
    assert rel_err(got[ok], g["solve_n"][ok]) < tol
