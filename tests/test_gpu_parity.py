"""Parity of the HIP kernels (through the C ABI, libkrome_<net>.so) with
the reference's solver (golden fixtures) and the oracle, on the GPU.

Bar: FP64 kernels within 1e-6 relative of the reference on primordial3
(rtol 1e-8) for every species above 1e-12 of the cell's total and for
Tgas, 5e-5 on primordial (integrated at rtol 1e-6, where the last bits of
the math library move the adaptive trajectory by up to ~10 rtol) and 1e-4
3e-4 on primordialZ (27 species at rtol 1e-6, the difference-quotient
Jacobian; measured 1.2e-4, the oracle itself 5.6e-5 from the reference);
flags equal.  primordialZ (27 equations) is above the lockstep persistent
kernel's 16: device_kernel persistent runs it in the lane queue (each lane
solves a cell with the plain kernel's code and takes the next from the
device work queue), whose results equal the plain kernel's bit for bit.
The mixed-precision kernel within 1e-5 (ten times inside the
benchmark's 1e-4 agreement bar; its Newton iteration refines FP32 solves
of an FP32-stored Jacobian, so it takes slightly different steps).
"""
import sys

import numpy as np
import pytest

import cells
from conftest import golden, oracle_opts, rel_err

pytestmark = pytest.mark.gpu

TOL = {"primordial": 5e-5, "primordial3": 1e-6, "primordialZ": 3e-4}
KERNELS = [("plain", 0), ("persistent", 0), ("persistent", 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nets():
    import torch
    assert torch.cuda.is_available()
    import krome
    return {n: krome.load(n) for n in ("primordial", "primordial3",
                                         "primordialZ")}


def _opts(net, d, kernel="plain", mixed=0, **kw):
    """options of a golden fixture's solve (its opts row), plus kw"""
    o = d["opts"]
    base = dict(evolve_T=bool(o[0]), rtol=o[1], atol_n=o[2], atol_T=o[3],
                jac=int(o[5]))
    base.update(kw)
    return net.options(device_kernel=kernel, mixed_precision=mixed,
                       error_report=False, **base)


def _tol(netname, mixed):
    return max(1e-5, TOL[netname]) if mixed else TOL[netname]


@pytest.mark.parametrize("kernel,mixed", KERNELS)
@pytest.mark.parametrize("netname", ["primordial3", "primordial",
                                     "primordialZ"])
def test_golden_cells(nets, netname, kernel, mixed, oracle):
    d = golden("%s_solve.npz" % netname)
    net = nets[netname]
    say("golden", netname, kernel, mixed, d["n0"].shape)
    res = net.solve(d["n0"], d["T0"], d["dt"], par=d["par"],
                    opts=_opts(net, d, kernel, mixed), check=False)
    if netname == "primordialZ" and kernel == "persistent":
        # above 16 equations: the lane queue, the plain kernel's solve
        plain = net.solve(d["n0"], d["T0"], d["dt"], par=d["par"],
                          opts=_opts(net, d, "plain", 0), check=False)
        assert np.array_equal(res.flags, plain.flags)
        assert np.array_equal(res.n, plain.n)
    assert res.ierr == 0
    assert np.array_equal(res.flags, d["flag_bitwise"])
    tol = _tol(netname, mixed)
    assert rel_err(res.n, d["n_bitwise"]) < tol
    assert np.max(np.abs(res.Tgas - d["T_bitwise"]) / d["T_bitwise"]) < tol
    # and the oracle on the same cells
    o = d["opts"]
    r, n, T, fl, st = oracle.solve(netname, d["n0"], d["T0"], d["dt"],
                                   d["par"], oracle_opts(
                                       evolve_T=o[0], rtol=o[1], atol_n=o[2],
                                       atol_T=o[3], jac=o[5]))
    assert rel_err(res.n, n) < tol
    assert np.max(np.abs(res.Tgas - T) / T) < tol


@pytest.mark.parametrize("kernel,mixed", KERNELS)
def test_counters_match_oracle(nets, kernel, mixed, oracle):
    """the kernels take the oracle's steps (up to the math library's last
    bits; the mixed-precision Newton up to its FP32 solves): total steps,
    rhs evaluations and Jacobians within 2 %"""
    import torch
    d = golden("primordial3_solve.npz")
    net = nets["primordial3"]
    say("counters", kernel, mixed)
    nc = d["n0"].shape[0]
    n = torch.tensor(np.ascontiguousarray(d["n0"].T), device="cuda")
    T = torch.tensor(d["T0"], device="cuda")
    dt = torch.tensor(d["dt"], device="cuda")
    par = torch.tensor(np.ascontiguousarray(d["par"]), device="cuda")
    st = torch.zeros((nc, 9), dtype=torch.int64, device="cuda")
    fl = torch.zeros(nc, dtype=torch.int32, device="cuda")
    r = net.solve_device(n, T, dt, par, opts=_opts(net, d, kernel, mixed),
                         flags=fl, stats=st)
    assert r == 0
    s = st.cpu().numpy()
    o = d["opts"]
    _, _, _, _, so = oracle.solve("primordial3", d["n0"], d["T0"], d["dt"],
                                  d["par"], oracle_opts(
                                      evolve_T=o[0], rtol=o[1], atol_n=o[2],
                                      atol_T=o[3]))
    for col in (0, 1, 2, 3, 5):
        assert abs(s[:, col].sum() / so[:, col].sum() - 1) < 0.02, col
    # the device-resident species-major call equals the host-array call
    res = net.solve(d["n0"], d["T0"], d["dt"], par=d["par"],
                    opts=_opts(net, d, kernel, mixed), check=False)
    assert np.array_equal(n.cpu().numpy().T, res.n)
    assert np.array_equal(T.cpu().numpy(), res.Tgas)


@pytest.mark.parametrize("kernel,mixed", KERNELS)
def test_edge_cases(nets, kernel, mixed):
    net = nets["primordial3"]
    w = cells.primordial3_cells(16)
    say("edge", kernel, mixed)
    base = dict(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6)
    o = net.options(device_kernel=kernel, mixed_precision=mixed, **base)
    # dt <= 0: untouched, flag 0
    dt = w["dt"].copy()
    dt[2] = 0.0
    dt[5] = -1.0
    res = net.solve(w["n"], w["T"], dt, par=w["par"], opts=o, check=False)
    assert res.ierr == 0 and (res.flags == 0).all()
    assert np.array_equal(res.n[[2, 5]], w["n"][[2, 5]])
    assert np.array_equal(res.Tgas[[2, 5]], w["T"][[2, 5]])
    # a tiny step budget: TOO_MUCH_WORK, inputs untouched
    ob = net.options(device_kernel=kernel, mixed_precision=mixed,
                     max_steps=3, **base)
    res = net.solve(w["n"], w["T"], w["dt"], par=w["par"], opts=ob,
                    check=False)
    assert res.ierr == -1 and (res.flags == -1).all()
    assert np.array_equal(res.n, w["n"]) and np.array_equal(res.Tgas, w["T"])
    # a non-positive Tgas in one cell: FLAG_NONPOSITIVE_T there only
    T0 = w["T"].copy()
    T0[1] = -5.0
    res = net.solve(w["n"], T0, w["dt"], par=w["par"], opts=o, check=False)
    assert res.flags[1] == -110 and res.ierr == -110
    assert (np.delete(res.flags, 1) == 0).all()
    assert np.array_equal(res.n[1], w["n"][1])
    # an unset REQUIRED parameter: -105 in that cell, nothing integrated
    par = w["par"].copy()
    par[3, 1] = np.nan
    fl = np.zeros(16, dtype=np.intc)
    import ctypes
    D = ctypes.POINTER(ctypes.c_double)
    n = w["n"].copy()
    r = net._lib.krome_solve_cells(
        16, n.ctypes.data_as(D), w["T"].copy().ctypes.data_as(D),
        w["dt"].ctypes.data_as(D), par.ctypes.data_as(D), net.npar, None,
        fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), o._h)
    assert r == -105 and fl[3] == -105 and (np.delete(fl, 3) == 0).all()
    assert np.array_equal(n, w["n"])


@pytest.mark.parametrize("kernel,mixed", KERNELS)
@pytest.mark.parametrize("ncell", [1, 63, 65, 1000])
def test_ragged_sizes(nets, oracle, kernel, mixed, ncell):
    net = nets["primordial3"]
    w = cells.primordial3_cells(ncell, first=12345)
    say("ragged", kernel, mixed, ncell)
    o = net.options(device_kernel=kernel, mixed_precision=mixed,
                    evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6)
    res = net.solve(w["n"], w["T"], w["dt"], par=w["par"], opts=o,
                    check=False)
    assert res.ierr == 0
    r, n, T, fl, st = oracle.solve("primordial3", w["n"], w["T"], w["dt"],
                                   w["par"], oracle_opts(evolve_T=1))
    assert rel_err(res.n, n) < _tol("primordial3", mixed)
    assert np.max(np.abs(res.Tgas - T) / T) < _tol("primordial3", mixed)


@pytest.mark.parametrize("kernel,mixed", KERNELS)
def test_large_batch_properties(nets, oracle, kernel, mixed):
    """65536 cells: every flag 0, the charge conserved (the Newton
    iteration's invariant) and He nuclei conserved, and a random sample of
    the cells equal to the oracle's solve of them"""
    import torch
    net = nets["primordial3"]
    ncell = 1 << 16
    say("large", kernel, mixed, ncell)
    w = cells.primordial3_cells(ncell)
    meta = cells.network_meta("primordial3")
    n = torch.tensor(np.ascontiguousarray(w["n"].T), device="cuda")
    T = torch.tensor(w["T"], device="cuda")
    dt = torch.tensor(w["dt"], device="cuda")
    par = torch.tensor(np.ascontiguousarray(w["par"]), device="cuda")
    fl = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    o = net.options(device_kernel=kernel, mixed_precision=mixed,
                    evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6)
    r = net.solve_device(n, T, dt, par, opts=o, flags=fl)
    assert r == 0 and (fl.cpu().numpy() == 0).all()
    nn = n.cpu().numpy().T
    TT = T.cpu().numpy()
    q = np.array(meta["charge"], dtype=float)
    ref = np.abs(w["n"] * q).sum(axis=1)
    assert np.max(np.abs(nn @ q - w["n"] @ q) / ref) < 1e-8
    he = np.array([1.0 if s.rstrip("+-") == "He" else 0.0
                   for s in meta["species"]])
    assert np.max(np.abs(nn @ he - w["n"] @ he) / (w["n"] @ he)) < 1e-6
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(ncell, 96, replace=False))
    _, no, To, flo, _ = oracle.solve("primordial3", w["n"][idx], w["T"][idx],
                                     w["dt"][idx], w["par"][idx],
                                     oracle_opts(evolve_T=1))
    assert rel_err(nn[idx], no) < _tol("primordial3", mixed)
    assert np.max(np.abs(TT[idx] - To) / To) < _tol("primordial3", mixed)


@pytest.mark.parametrize("layout", [0, 1])
def test_mass_fractions(nets, layout):
    """krome_solve_cells_x (host rows) and krome_solve_cells_x_device
    (either layout): x = m n / rho of the reference's density solve, with
    renormalize_x rescaled to the input sum; a dt <= 0 cell and a failed
    cell keep their x on the device"""
    import torch
    net = nets["primordial3"]
    d = golden("primordial3_solve.npz")
    m = net.masses
    nc = d["n0"].shape[0]
    rho = (d["n0"] @ m) * 1.25  # the caller's rho: not only the species'
    x0 = d["n0"] * m / rho[:, None]
    say("mass fractions", layout)
    o = _opts(net, d, "persistent", 0, renormalize_x=True)
    res = net.solve_x(x0, rho, d["T0"], d["dt"], par=d["par"], opts=o,
                      check=False)
    assert res.ierr == 0 and (res.flags == 0).all()
    xr = d["n_bitwise"] * m / rho[:, None]
    xr *= (x0.sum(axis=1) / xr.sum(axis=1))[:, None]
    assert rel_err(res.n, xr) < TOL["primordial3"]
    assert np.max(np.abs(res.Tgas - d["T_bitwise"]) / d["T_bitwise"]) < \
        TOL["primordial3"]
    dt = d["dt"].copy()
    dt[1] = 0.0
    T0 = d["T0"].copy()
    T0[2] = -5.0
    xt = torch.tensor(x0.T.copy() if layout == 0 else x0, device="cuda")
    fl = torch.zeros(nc, dtype=torch.int32, device="cuda")
    r = net.solve_x_device(xt, torch.tensor(rho, device="cuda"),
                           torch.tensor(T0, device="cuda"),
                           torch.tensor(dt, device="cuda"),
                           torch.tensor(d["par"], device="cuda"), opts=o,
                           flags=fl, layout=layout)
    torch.cuda.synchronize()
    assert r == -110
    f = fl.cpu().numpy()
    assert f[2] == -110 and (np.delete(f, 2) == 0).all()
    got = (xt.T if layout == 0 else xt).cpu().numpy()
    assert np.array_equal(got[[1, 2]], x0[[1, 2]])
    ok = np.ones(nc, dtype=bool)
    ok[[1, 2]] = False
    assert rel_err(got[ok], xr[ok]) < TOL["primordial3"]


@pytest.mark.parametrize("mixed", [0, 1])
def test_queue_chunk(nets, mixed):
    """queue_chunk only changes which lane takes which cell, never a
    cell's arithmetic: 8192 cells give bitwise the same rows and flags
    with the wave reservoir of 64 and 1024 cells per grab as with exact
    grabs (0); a value below 64 is exact grabs too"""
    import torch
    net = nets["primordial3"]
    w = cells.primordial3_cells(8192, first=777)
    say("queue chunk", mixed)
    outs = []
    for chunk in (0, 64, 1024, 17):
        n = torch.tensor(np.ascontiguousarray(w["n"].T), device="cuda")
        T = torch.tensor(w["T"], device="cuda")
        fl = torch.full((8192,), 7, dtype=torch.int32, device="cuda")
        o = net.options(device_kernel="persistent", mixed_precision=mixed,
                        evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                        queue_chunk=chunk)
        r = net.solve_device(n, T, torch.tensor(w["dt"], device="cuda"),
                             torch.tensor(w["par"], device="cuda"), opts=o,
                             flags=fl)
        assert r == 0
        outs.append((n.cpu().numpy(), T.cpu().numpy(), fl.cpu().numpy()))
    for n, T, f in outs[1:]:
        assert np.array_equal(f, outs[0][2]) and (f == 0).all()
        assert np.array_equal(n, outs[0][0])
        assert np.array_equal(T, outs[0][1])
