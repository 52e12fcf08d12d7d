"""Several networks in one persistent launch (libkrome_multi.so,
include/krome_multi.h; no counterpart in the reference, which builds one
library per network).

- the library exports every symbol its header declares; bad arguments and a
  Tgas mode the network was not built for are refused before any device
  work;
- on the GPU, primordial3 and primordial batches solved in one launch give,
  cell by cell, bitwise the results of each network's own library with the
  persistent kernel and the same options (FP64 and mixed precision), and
  an unset REQUIRED parameter flags its batch -105.
"""
import ctypes
import re
import subprocess
import sys

import numpy as np
import pytest

import cells
from conftest import PKG, REPO

SO = PKG / "libkrome_multi.so"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lib():
    if not SO.exists():
        pytest.skip("%s not built" % SO.name)
    import krome
    return krome.load_multi()


def test_exports_and_argument_checks():
    m = lib()
    text = (REPO / "include" / "krome_multi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"\b(krome_multi_[a-z_]+)\s*\(", text))
    out = subprocess.run(["nm", "-D", "--defined-only", str(SO)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (krome_multi_\w+)", out))
    assert declared and declared <= exported, declared - exported
    assert m.networks == ["primordial", "primordial3"]
    L = m._lib
    assert L.krome_multi_network_name(2) is None
    import krome
    B = krome._MultiBatch
    ms = ctypes.c_double()
    D = ctypes.POINTER(ctypes.c_double)
    p = ctypes.c_void_p(16)  # never dereferenced: refused first
    ok = dict(ncell=4, n=p, Tgas=p, dt=p, par=p, flags=p, rtol=1e-8,
              atol_n=1e-20, atol_T=1e-6, max_steps=100000)

    def call(*bs):
        arr = (B * len(bs))(*bs)
        return L.krome_multi_solve_device(arr, len(bs), 1,
                                          ctypes.cast(ctypes.byref(ms), D))
    assert call(B(network=5, evolve_T=0, **ok)) == -1
    assert call(B(network=0, evolve_T=0, **ok),
                B(network=0, evolve_T=0, **ok)) == -1
    assert call(B(network=0, evolve_T=1, **ok)) == -104
    assert call(B(network=1, evolve_T=0, **ok)) == -104
    assert L.krome_multi_solve_device(None, 1, 1, None) == -1


def _batches(nets, n3, n1):
    import torch
    w3 = cells.primordial3_cells(n3, first=3 << 20)
    w1 = cells.primordial_cells(n1, first=4096)
    out = []
    for name, w, et, rtol in (("primordial3", w3, 1, 1e-8),
                              ("primordial", w1, 0, 1e-6)):
        out.append(dict(
            network=name, evolve_T=et, rtol=rtol, atol_n=1e-20, atol_T=1e-6,
            n=torch.tensor(np.ascontiguousarray(w["n"].T), device="cuda"),
            Tgas=torch.tensor(w["T"], device="cuda"),
            dt=torch.tensor(w["dt"], device="cuda"),
            par=torch.tensor(np.ascontiguousarray(w["par"]), device="cuda"),
            flags=torch.zeros(w["n"].shape[0], dtype=torch.int32,
                              device="cuda")))
    return out


def _single(nets, b, mixed):
    """the same batch through its network's own library"""
    net = nets[b["network"]]
    opts = net.options(evolve_T=bool(b["evolve_T"]), rtol=b["rtol"],
                       atol_n=b["atol_n"], atol_T=b["atol_T"],
                       device_kernel="persistent", mixed_precision=mixed,
                       error_report=False)
    n, T = b["n"].clone(), b["Tgas"].clone()
    fl = b["flags"].clone()
    r = net.solve_device(n, T, b["dt"], b["par"], opts=opts, flags=fl)
    return r, n, T, fl


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [1, 0])
def test_pair_launch_bitwise_vs_single_network(mixed):
    import torch
    import krome
    m = lib()
    nets = {n: krome.load(n) for n in ("primordial", "primordial3")}
    say("pair", mixed)
    bs = _batches(nets, 8192, 16384)
    ref = [_single(nets, b, mixed) for b in bs]
    r, ms = m.solve_device(bs, mixed_precision=mixed)
    torch.cuda.synchronize()
    assert r == 0 and ms > 0.0
    for b, (r1, n1, T1, f1) in zip(bs, ref):
        assert r1 == 0
        assert torch.equal(b["flags"], f1)
        assert torch.equal(b["n"], n1), b["network"]
        assert torch.equal(b["Tgas"], T1), b["network"]
    # one batch alone: the network's own persistent kernel
    one = _batches(nets, 512, 1)[:1]
    r1, n1, T1, f1 = _single(nets, one[0], mixed)
    r, ms = m.solve_device(one, mixed_precision=mixed)
    assert r == r1 == 0 and torch.equal(one[0]["n"], n1)


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [0, 1])
def test_pair_launch_uneven_cells_vs_oracle(oracle, mixed):
    """the uneven batch's primordial cells (density up to 1e12, dt up to
    1e14 s) and primordial3 cells, one launch, against the oracle; with
    mixed precision the primordial cells (analytic Jacobian) keep the FP64
    Newton matrix: the cold dense cells there need it"""
    import torch
    from conftest import oracle_opts, rel_err
    m = lib()
    w1 = cells.c3_primordial_cells(4096)
    w3 = cells.primordial3_cells(1024, first=5 << 20)
    bs = []
    for name, w, et, rtol in (("primordial3", w3, 1, 1e-8),
                              ("primordial", w1, 0, 1e-6)):
        bs.append(dict(
            network=name, evolve_T=et, rtol=rtol, atol_n=1e-20, atol_T=1e-6,
            n=torch.tensor(np.ascontiguousarray(w["n"].T), device="cuda"),
            Tgas=torch.tensor(w["T"], device="cuda"),
            dt=torch.tensor(w["dt"], device="cuda"),
            par=torch.tensor(np.ascontiguousarray(w["par"]), device="cuda"),
            flags=torch.zeros(w["n"].shape[0], dtype=torch.int32,
                              device="cuda")))
    say("uneven vs oracle", mixed)
    r, _ = m.solve_device(bs, mixed_precision=mixed)
    assert r == 0
    for b, w, net, et, tol in ((bs[0], w3, "primordial3", 1,
                                1e-5 if mixed else 1e-6),
                               (bs[1], w1, "primordial", 0, 5e-5)):
        idx = np.arange(0, w["n"].shape[0], 8)
        ro, n, T, fl, st = oracle.solve(
            net, w["n"][idx], w["T"][idx], w["dt"][idx], w["par"][idx],
            oracle_opts(evolve_T=et, rtol=b["rtol"]))
        assert ro == 0
        got = b["n"].T.cpu().numpy()[idx]
        assert rel_err(got, n) < tol, net


@pytest.mark.gpu
def test_pair_launch_unset_parameter():
    import torch
    import krome
    m = lib()
    nets = {n: krome.load(n) for n in ("primordial", "primordial3")}
    bs = _batches(nets, 256, 256)
    n0 = bs[0]["n"].clone()
    bs[0]["par"][7, 1] = float("nan")  # tff of one primordial3 cell
    r, _ = m.solve_device(bs)
    torch.cuda.synchronize()
    assert r == -105
    f = bs[0]["flags"].cpu().numpy()
    assert f[7] == -105 and (np.delete(f, 7) == 0).all()
    assert torch.equal(bs[0]["n"], n0)


# ---- the uneven batch's three networks in one launch (libkrome_multi_c3) --

SO_C3 = PKG / "libkrome_multi_c3.so"


def lib_c3():
    if not SO_C3.exists():
        pytest.skip("%s not built" % SO_C3.name)
    import krome
    return krome.load_multi("multi_c3")


def test_c3_exports_and_layout():
    m = lib_c3()
    text = (REPO / "include" / "krome_multi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    declared = set(re.findall(r"\b(krome_multi_[a-z_]+)\s*\(", text))
    out = subprocess.run(["nm", "-D", "--defined-only", str(SO_C3)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (krome_multi_\w+)", out))
    assert declared <= exported, declared - exported
    assert m.networks == ["primordial", "primordialZ", "cloud"]
    # primordial (12 equations) in the lockstep kernel, the larger two in
    # the lane queue; every network here at fixed Tgas
    assert m.lockstep == [1, 0, 0] and m.evolve_T == [0, 0, 0]
    # primordialZ (27) in the lane queue, cloud (452) a wave per cell
    assert m.mode == [1, 0, 2]
    with pytest.raises(Exception):
        m.set_launch(-1, 0)


def _golden_batch(name, m=None):
    """the reference's golden cells of a network as a batch (m: its m
    cheapest cells only)"""
    import torch
    from conftest import golden
    d = dict(golden("%s_solve.npz" % name))
    if m is not None:
        sel = np.sort(np.argsort(d["stats_bitwise"][:, 0])[:m])
        for k in ("n0", "T0", "dt", "par", "n_bitwise", "flag_bitwise"):
            d[k] = d[k][sel]
    o = d["opts"]
    return d, dict(
        network=name, evolve_T=int(o[0]), rtol=o[1], atol_n=o[2],
        atol_T=o[3], max_steps=int(o[6]), jac=int(o[5]),
        n=torch.tensor(np.ascontiguousarray(d["n0"].T), device="cuda"),
        Tgas=torch.tensor(d["T0"], device="cuda"),
        dt=torch.tensor(d["dt"], device="cuda"),
        par=torch.tensor(np.ascontiguousarray(d["par"]), device="cuda"),
        flags=torch.zeros(d["n0"].shape[0], dtype=torch.int32,
                          device="cuda"))


@pytest.mark.gpu
@pytest.mark.parametrize("wpc,chunk", [(0, 0), (2, 64)])
def test_c3_launch_vs_reference_and_single(wpc, chunk):
    """the reference's golden cells of the three networks in one launch:
    each batch within its network's bar of the reference's solves, flags
    equal, and bitwise the network's own persistent solve (the lane
    queue's cells are the plain kernel's solve, the lockstep cells the
    persistent kernel's)"""
    import torch
    import krome
    from conftest import rel_err
    m = lib_c3()
    m.set_launch(wpc, chunk)
    bars = {"primordial": 5e-5, "primordialZ": 3e-4, "cloud": 1e-4}
    got = [_golden_batch(n) for n in ("primordial", "primordialZ", "cloud")]
    singles = []
    for d, b in got:
        net = krome.load(b["network"])
        opts = net.options(evolve_T=False, rtol=b["rtol"], atol_n=b["atol_n"],
                           atol_T=b["atol_T"], jac=b["jac"],
                           max_steps=b["max_steps"],
                           device_kernel="persistent", error_report=False)
        n, T, fl = b["n"].clone(), b["Tgas"].clone(), b["flags"].clone()
        r = net.solve_device(n, T, b["dt"], b["par"], opts=opts, flags=fl)
        singles.append((r, n, fl))
    say("c3 launch", wpc, chunk)
    r, ms = m.solve_device([b for _, b in got], mixed_precision=1)
    m.set_launch(0, 0)
    torch.cuda.synchronize()
    assert r == 0 and ms > 0.0
    for (d, b), (r1, n1, f1) in zip(got, singles):
        name = b["network"]
        fl = b["flags"].cpu().numpy()
        assert np.array_equal(fl, d["flag_bitwise"]), name
        assert rel_err(b["n"].T.cpu().numpy(), d["n_bitwise"]) < bars[name]
        assert r1 == 0 and torch.equal(b["flags"], f1), name
        assert torch.equal(b["n"], n1), name
    m.set_launch(0, 0)
