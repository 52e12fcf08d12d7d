"""The solve that allocates nothing (krome_solve_cells_device_noalloc),
through the C ABI, against krome_solve_cells_device on the same cells.

Both run the same kernels on the same inputs, so their outputs are equal
bit for bit: plain FP64, persistent FP64 and persistent mixed precision,
two calls in a row (the second with fewer cells: the scratch keeps the
first call's contents), species-major and cell-major layouts.  Also the
argument checks: more cells than the scratch's maximum (-1, nothing
touched) and a configuration it does not run (conserve 2: -104, every
cell flagged).
"""
import ctypes
import sys

import numpy as np
import pytest

import cells

pytestmark = pytest.mark.gpu

KERNELS = [("plain", 0), ("persistent", 0), ("persistent", 1)]
MAX_CELLS = 1 << 20
D = ctypes.c_void_p


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def net():
    import torch
    assert torch.cuda.is_available()
    import krome
    net = krome.load("primordial3")
    # the prototypes (pointers as pointers: an undeclared Python int
    # argument would be passed as a 32-bit int)
    L, I, P = ctypes.c_long, ctypes.c_int, ctypes.c_void_p
    f = net._lib.krome_solve_cells_device_noalloc
    f.restype = I
    f.argtypes = [I, P, I, L, P, P, P, I, P, P, P]
    g = net._lib.krome_solve_cells_device
    g.restype = I
    g.argtypes = [I, P, I, L, P, P, P, I, P, P, P, P]
    return net


def _dev(w, layout):
    import torch
    n = w["n"] if layout == 1 else np.ascontiguousarray(w["n"].T)
    return dict(n=torch.tensor(np.ascontiguousarray(n), device="cuda"),
                T=torch.tensor(w["T"], device="cuda"),
                dt=torch.tensor(w["dt"], device="cuda"),
                par=torch.tensor(np.ascontiguousarray(w["par"]),
                                 device="cuda"),
                fl=torch.full((len(w["T"]),), 7, dtype=torch.int32,
                              device="cuda"))


def _call(net, fn, nc, d, layout, opts, noalloc):
    import torch
    f = getattr(net._lib, fn)
    ptr = lambda t: D(t.data_ptr())
    if noalloc:
        r = f(nc, ptr(d["n"]), layout, ctypes.c_long(0), ptr(d["T"]),
              ptr(d["dt"]), ptr(d["par"]), net.npar, ptr(d["fl"]), None,
              opts._h)
    else:
        r = f(nc, ptr(d["n"]), layout, ctypes.c_long(0), ptr(d["T"]),
              ptr(d["dt"]), ptr(d["par"]), net.npar, None, ptr(d["fl"]),
              None, opts._h)
    torch.cuda.synchronize()
    return r


@pytest.mark.parametrize("kernel,mixed", KERNELS)
@pytest.mark.parametrize("layout", [0, 1])
def test_noalloc_equals_device_entry(net, kernel, mixed, layout):
    opts = net.options(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                       device_kernel=kernel, mixed_precision=mixed,
                       error_report=False)
    for nc in (4096, 1000):  # the second call smaller
        w = cells.primordial3_cells(nc, seed=5 + nc)
        say("noalloc", kernel, mixed, "layout", layout, "cells", nc)
        a, b = _dev(w, layout), _dev(w, layout)
        ra = _call(net, "krome_solve_cells_device", nc, a, layout, opts,
                   False)
        rb = _call(net, "krome_solve_cells_device_noalloc", nc, b, layout,
                   opts, True)
        fa, fb = a["fl"].cpu().numpy(), b["fl"].cpu().numpy()
        assert ra == rb
        assert np.array_equal(fa, fb)
        assert (b["fl"].cpu().numpy() == 0).all()
        assert np.array_equal(a["n"].cpu().numpy(), b["n"].cpu().numpy())
        assert np.array_equal(a["T"].cpu().numpy(), b["T"].cpu().numpy())


def test_noalloc_limits(net):
    import torch
    w = cells.primordial3_cells(64, seed=11)
    d = _dev(w, 0)
    opts = net.options(evolve_T=True, device_kernel="persistent")
    say("noalloc limits")
    f = net._lib.krome_solve_cells_device_noalloc
    ptr = lambda t: D(t.data_ptr())
    # more cells than the scratch holds: refused before any work
    r = f(MAX_CELLS + 1, ptr(d["n"]), 0, ctypes.c_long(0), ptr(d["T"]),
          ptr(d["dt"]), ptr(d["par"]), net.npar, ptr(d["fl"]), None,
          opts._h)
    torch.cuda.synchronize()
    assert r == -1
    assert (d["fl"].cpu().numpy() == 7).all()
    # the post-solve projection needs buffers of its own: -104 everywhere
    oc = net.options(evolve_T=True, device_kernel="persistent", conserve=2)
    n0 = d["n"].clone()
    r = f(64, ptr(d["n"]), 0, ctypes.c_long(0), ptr(d["T"]), ptr(d["dt"]),
          ptr(d["par"]), net.npar, ptr(d["fl"]), None, oc._h)
    torch.cuda.synchronize()
    assert r == -104
    assert (d["fl"].cpu().numpy() == -104).all()
    assert torch.equal(d["n"], n0)
