"""The discrete adjoint (krome-x_amd/csrc/kx_adjoint.h, kx_adjmath.h,
kx_dual.h; Python krome-x_amd/adjoint.py).

CPU (the per-state mathematics, host build tests/adjoint_host.cpp):
  - the dual-number Jacobian against the reference's own Jacobians
    (jac_host_, tests/golden/*_funcs.npz): primordial at fixed Tgas, where
    the reference's is analytic, within 1e-11 of each row's largest entry;
    primordial3 with Tgas evolved, where the reference takes central
    differences for the temperature column and difference quotients for
    the density-aggregate terms, within 1e-4 (cells at a rate fit's break
    temperature, where the reference's central difference straddles two
    pieces, excepted from the temperature column);
  - along relative perturbations of the abundant species and of Tgas, the
    Jacobian's directional derivative against central differences of the
    same rhs (1e-5 of the terms' scale for primordial3, 1e-3 for
    primordial, where the differences' own truncation is larger);
  - f_theta^T w (the rates' part) against central differences in log k of
    the rate-scaled rhs, for random w (1e-7 of the terms' scale);
  - the transposed solve against numpy.

GPU (-m gpu, through the C ABI): the adjoint kernel's forward against
the reference's solves (the FP64 bar, 1e-6); its outputs identical with
and without a gradient; gradients identical for different checkpoint
intervals; the gradients of log(n_H2/nH) and log(Tgas) at t_end by every
log k_r and log n_i(0) against central differences (h = 1e-6 in log k)
of the same kernel's forward, and against the reference's perturbed
solves (tests/golden/primordial3_adjoint.npz); torch.autograd through
adjoint.Solve equal to the raw entry point's gradients.
"""
import ctypes
import subprocess
import sys

import numpy as np
import pytest

import cells
from conftest import REPO, golden, rel_err

D = ctypes.POINTER(ctypes.c_double)
BUILD = REPO / "tests" / "_build"
NETS = {"primordial": 0, "primordial3": 1}
# rate-fit break temperatures of the function samples' T grid (the
# reference's central differences straddle two pieces there)
BREAKS = (5.5e3,)


def _p(a):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(D)


@pytest.fixture(scope="module")
def host():
    BUILD.mkdir(exist_ok=True)
    out = BUILD / "libadjoint_host.so"
    src = REPO / "tests" / "adjoint_host.cpp"
    csrc = REPO / "krome-x_amd" / "csrc"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
                    "-fopenmp", "-I",
                    str(csrc), "-I", str(csrc / "generated"), str(src), "-o",
                    str(out)], check=True)
    return ctypes.CDLL(str(out))


def _state(name):
    g = golden("%s_funcs.npz" % name)
    n, T, par = g["n"], g["T"], g["par"]
    N = n.shape[1] + (1 if name == "primordial3" else 0)
    y = np.hstack([n, T[:, None]]) if N > n.shape[1] else n
    return g, np.ascontiguousarray(y), T, par, N


def _jac(host, net, y, T, par, lnk=None):
    nc, N = y.shape
    J = np.zeros((nc, N * N))
    host.adj_jac(NETS[net], nc, _p(y), _p(T), _p(par), _p(lnk), _p(J))
    return J.reshape(nc, N, N).transpose(0, 2, 1)  # [c, i, j]


def _rhs(host, net, y, T, par, lnk=None):
    f = np.zeros(y.shape)
    host.adj_rhs(NETS[net], y.shape[0], _p(y), _p(T), _p(par), _p(lnk),
                 _p(f))
    return f


@pytest.mark.parametrize("net", ["primordial", "primordial3"])
def test_dual_jacobian_vs_reference(host, net):
    g, y, T, par, N = _state(net)
    J = _jac(host, net, y, T, par)
    Jr = g["J"]
    scale = np.maximum(np.abs(Jr).max(axis=2, keepdims=True), 1e-300)
    e = np.abs(J - Jr) / scale
    if net == "primordial":
        assert e.max() < 1e-11
        return
    brk = np.isin(T, BREAKS)
    assert e[:, :, :N - 1].max() < 1e-4          # species columns
    assert e[~brk][:, :, N - 1].max() < 1e-4     # the temperature column
    assert e[:, N - 1, :].max() < 1e-4           # the temperature row
    # most entries far tighter: the analytic parts agree to rounding
    assert np.median(e) < 1e-14


@pytest.mark.parametrize("net", ["primordial", "primordial3"])
def test_dual_jacobian_directional_fd(host, net):
    """J v against (f(y + e v) - f(y - e v)) / 2e with v a relative
    perturbation of the species above 1e-6 of the cell's total (and Tgas),
    whose changes the rhs resolves"""
    g, y, T, par, N = _state(net)
    J = _jac(host, net, y, T, par)
    ns = g["n"].shape[1]
    rng = np.random.default_rng(3)
    ok = ~np.isin(T, BREAKS)
    tot = y[:, :ns].sum(axis=1, keepdims=True)
    for trial in range(4):
        u = rng.uniform(-1.0, 1.0, size=y.shape)
        u[:, :ns] *= y[:, :ns] > 1e-6 * tot
        v = y * u
        jd = np.einsum("cij,cj->ci", J, v)
        scale = np.einsum("cij,cj->ci", np.abs(J), np.abs(v))
        # the rhs's own rounding, ~1e-16 of its terms (|J| |y| bounds
        # them), divided by the step
        terms = np.einsum("cij,cj->ci", np.abs(J), np.abs(y))
        # primordial's Jacobian equals the reference's analytic one to
        # 1e-11 (above): what is left here is the differences' truncation
        # in its stiff rows (measured 1.9e-4); primordial3 measured 1.8e-6
        bar = 1e-3 if net == "primordial" else 1e-5
        good = np.zeros(y.shape, dtype=bool)
        for eps in (1e-5, 1e-6, 1e-7):
            d = (_rhs(host, net, y + eps * v, T, par)
                 - _rhs(host, net, y - eps * v, T, par)) / (2 * eps)
            good |= np.abs(d - jd) <= bar * scale + 1e-15 * terms / eps
        assert good[ok].all(), (trial, np.argwhere(~good & ok[:, None])[:4])


@pytest.mark.parametrize("net", ["primordial", "primordial3"])
def test_theta_vjp_vs_fd(host, net):
    """the rates' part of the rhs's derivative by log k_r, applied to w,
    against central differences of the rate-scaled rhs"""
    g, y, T, par, N = _state(net)
    nc = y.shape[0]
    R = g["k"].shape[1]
    rng = np.random.default_rng(5)
    w = rng.normal(size=(nc, N))
    gv = np.zeros((nc, R))
    host.adj_vjp(NETS[net], nc, _p(y), _p(T), _p(par), None, _p(w), _p(gv))
    h = 1e-4
    # the rhs's rounding, ~1e-16 of its terms (bounded by |J| |y|), over 2h
    J = _jac(host, net, y, T, par)
    terms = np.einsum("cij,cj->ci", np.abs(J), np.abs(y))
    for r in range(R):
        lp = np.zeros((nc, R))
        lp[:, r] = h
        dp = _rhs(host, net, y, T, par, lp)
        dm = _rhs(host, net, y, T, par, -lp)
        d = (dp - dm) / (2 * np.sinh(h))  # f is linear in k_r = e^theta k
        gfd = (d * w).sum(axis=1)
        allow = 1e-7 * (np.abs(d) * np.abs(w)).sum(axis=1) + \
            1e-15 * (terms * np.abs(w)).sum(axis=1) / h
        assert (np.abs(gfd - gv[:, r]) <= allow).all(), r


def _host_adjoint(host, w, gn, gT, rtol=1e-8, seg=32):
    """the adjoint on the host, cell by cell (adjoint_cell of kx_adjoint.h):
    flags, final n, T, and the gradients (g_lnk, g_n0, g_T0)"""
    nc, ns = w["n"].shape
    n = np.ascontiguousarray(w["n"].T).copy()
    T = w["T"].copy()
    fl = np.zeros(nc, dtype=np.intc)
    R = 27
    gl, gn0, gT0 = np.zeros((nc, R)), np.zeros((ns, nc)), np.zeros(nc)
    o = np.array([1, rtol, 1e-20, 1e-6, 1e6, seg], dtype=float)
    host.adj_adjoint.restype = ctypes.c_long
    host.adj_adjoint(1, nc, _p(n), _p(T), _p(w["dt"]), _p(w["par"]), None,
                     fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     _p(np.ascontiguousarray(gn.T)), _p(gT), _p(gl),
                     _p(gn0), _p(gT0), _p(o))
    return fl, n.T, T, gl, gn0.T, gT0


def test_adjoint_vs_replayed_scheme(host):
    """the adjoint differentiates the BDF scheme of the forward run with its
    control frozen.  That scheme is replayed here from the forward's tape
    (the same rescalings, predictors, order changes and correctors, each
    corrector equation solved by Newton to convergence) under perturbed
    log k and log n(0); its central differences agree with the adjoint
    within 1e-4 (relative; below 1e-3 of the cell's largest gradient,
    absolute) plus their own error.  The differences are taken at h = 1e-4,
    1e-5 and 1e-6 (curvature limits the first, the replay's rounding the
    last); per gradient the pair of steps that agree best gives the
    estimate (the smaller step's) and its error (their difference).  Three
    long primordial3 collapse cells."""
    w = cells.primordial3_cells(16, seed=3)
    pick = [0, 3, 6]
    w = {k: np.ascontiguousarray(v[pick]) for k, v in w.items()
         if k != "opts"}
    nc, ns = w["n"].shape
    N, R = ns + 1, 27
    iH2 = 3
    wH = np.array([1, 1, 1, 2, 2, 0, 0, 0, 0.0])
    fl, n, T, *_ = _host_adjoint(host, w, np.zeros((nc, ns)), np.zeros(nc))
    assert (fl == 0).all()
    nH = n @ wH
    g1 = -wH[None, :] / nH[:, None]
    g1[:, iH2] += 1.0 / n[:, iH2]
    fl, _, _, gl, gn0, _ = _host_adjoint(host, w, g1, np.zeros(nc))
    o = np.array([1, 1e-8, 1e-20, 1e-6, 1e6, 32.0])
    cap, ycap = 200000, 100000
    I = ctypes.POINTER(ctypes.c_int)
    steps = (1e-4, 1e-5, 1e-6)
    for c in range(nc):
        rec, ys = np.zeros(cap * 12), np.zeros(ycap * N)
        cnt, cw = np.zeros(4, dtype=np.intc), np.zeros(6)
        par = np.ascontiguousarray(w["par"][c])
        n0 = np.ascontiguousarray(w["n"][c])
        f = host.adj_tape(1, _p(n0), ctypes.c_double(w["T"][c]),
                          ctypes.c_double(w["dt"][c]), _p(par), _p(o),
                          _p(rec), cap, _p(ys), ycap, cnt.ctypes.data_as(I),
                          _p(cw))
        assert f == 0 and cnt[2] == 0
        y0 = np.append(n0, w["T"][c])

        # every perturbation (step, parameter, sign) replayed in one batch
        P = R + ns
        lnks = np.zeros((len(steps), P, 2, R))
        y0s = np.tile(y0, (len(steps), P, 2, 1))
        for k, h in enumerate(steps):
            for j in range(P):
                for sgi, sg in enumerate((1.0, -1.0)):
                    if j < R:
                        lnks[k, j, sgi, j] = sg * h
                    else:
                        y0s[k, j, sgi, j - R] *= np.exp(sg * h)
        nb = len(steps) * P * 2
        out = np.zeros((nb, N))
        assert host.adj_replay_batch(
            1, _p(rec), int(cnt[0]), _p(ys), nb, _p(y0s.reshape(nb, N)),
            _p(par), ctypes.c_double(w["T"][c]), _p(lnks.reshape(nb, R)),
            _p(cw), _p(out)) == 0
        Lv = (np.log(out[:, iH2]) - np.log(out[:, :ns] @ wH)).reshape(
            len(steps), P, 2)
        fds = (Lv[:, :, 0] - Lv[:, :, 1]) / (2 * np.array(steps)[:, None])
        spread = np.abs(np.diff(fds, axis=0))  # [pair, j]
        best = spread.argmin(axis=0)
        fd = fds[best + 1, np.arange(R + ns)]
        e_fd = spread[best, np.arange(R + ns)]
        adj = np.concatenate([gl[c], gn0[c] * w["n"][c]])
        den = np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
        assert (np.abs(adj - fd) <= 1e-4 * den + e_fd).all(), c
        # and the differences themselves settled: their error is small
        assert np.median(e_fd / den) < 1e-5


def test_transposed_solve(host):
    rng = np.random.default_rng(2)
    for net, N in (("primordial", 12), ("primordial3", 10)):
        nc = 40
        J = rng.normal(size=(nc, N, N)) * 10.0 ** rng.uniform(-3, 3, (nc, 1, 1))
        gam = 10.0 ** rng.uniform(-4, 2, nc)
        b = rng.normal(size=(nc, N))
        mu = np.zeros((nc, N))
        Jc = np.ascontiguousarray(J.transpose(0, 2, 1))  # column-major
        host.adj_solve_t(NETS[net], nc, _p(Jc), _p(gam), _p(b), _p(mu))
        for c in range(nc):
            A = np.eye(N) - gam[c] * J[c]
            ref = np.linalg.solve(A.T, b[c])
            assert np.allclose(mu[c], ref, rtol=1e-9,
                               atol=1e-12 * np.abs(ref).max())


# ---- on the GPU -------------------------------------------------------------

def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def p3():
    import torch
    assert torch.cuda.is_available()
    import krome
    return krome.load("primordial3")


def _opts(net, **kw):
    base = dict(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                error_report=False)
    base.update(kw)
    return net.options(**base)


def _dev(w, lnk=None):
    import torch
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64,
                               device="cuda")
    return (t(w["n"].T), t(w["T"]), t(w["dt"]), t(w["par"]),
            None if lnk is None else t(lnk))


def _forward(net, w, opts, lnk=None):
    n, T, dt, par, L = _dev(w, lnk)
    r, _, _ = net.adjoint_device(n, T, dt, par, opts, lnk=L)
    return r, n.cpu().numpy().T, T.cpu().numpy()


def _grads(net, w, opts, gn, gT, lnk=None):
    import torch
    n, T, dt, par, L = _dev(w, lnk)
    fl = torch.zeros(len(w["T"]), dtype=torch.int32, device="cuda")
    r, (gl, gn0, gT0), mem = net.adjoint_device(
        n, T, dt, par, opts, lnk=L, flags=fl,
        gn=torch.tensor(np.ascontiguousarray(gn.T), device="cuda"),
        gT=torch.tensor(gT, device="cuda"))
    return (r, fl.cpu().numpy(), n.cpu().numpy().T, T.cpu().numpy(),
            gl.cpu().numpy(), gn0.cpu().numpy().T, gT0.cpu().numpy(), mem)


def _outputs(net, n, T):
    """log(n_H2/nH) and log(Tgas) at t_end, with their cotangents
    (d/dn, d/dT) per cell"""
    iH2 = net.idx("H2")
    wH = net.composition[:, net.elements.index("H")].astype(float)
    nH = n @ wH
    L1 = np.log(n[:, iH2]) - np.log(nH)
    g1 = -wH[None, :] / nH[:, None]
    g1[:, iH2] += 1.0 / n[:, iH2]
    L2 = np.log(T)
    return (L1, g1, np.zeros_like(T)), (L2, np.zeros_like(n), 1.0 / T)


@pytest.mark.gpu
def test_adjoint_forward_vs_reference(p3):
    d = golden("primordial3_solve.npz")
    w = {"n": d["n0"], "T": d["T0"], "dt": d["dt"], "par": d["par"]}
    say("adjoint forward", d["n0"].shape)
    r, n, T = _forward(p3, w, _opts(p3))
    assert r == 0
    assert rel_err(n, d["n_bitwise"]) < 1e-6
    assert np.max(np.abs(T - d["T_bitwise"]) / d["T_bitwise"]) < 1e-6
    # the same outputs when a gradient is taken
    gn = np.ones_like(d["n0"])
    gT = np.ones_like(d["T0"])
    r2, fl, n2, T2, *_ = _grads(p3, w, _opts(p3), gn, gT)
    assert r2 == 0 and (fl == 0).all()
    assert np.array_equal(n2, n) and np.array_equal(T2, T)


@pytest.mark.gpu
def test_adjoint_checkpoint_interval_invariance(p3):
    """the segments only decide what is recomputed: checkpoints every 8,
    32 and 128 steps give bitwise the same gradients; a pool too small for
    a cell's checkpoints flags that cell (-111)"""
    w = cells.primordial3_cells(64, seed=3)
    rng = np.random.default_rng(1)
    gn = rng.normal(size=w["n"].shape)
    gT = rng.normal(size=w["T"].shape)
    say("adjoint checkpoints")
    outs = [_grads(p3, w, _opts(p3, checkpoint_steps=k), gn, gT)
            for k in (8, 32, 128)]
    for o in outs[1:]:
        for a, b in zip(o[1:7], outs[0][1:7]):
            assert np.array_equal(a, b)
    r, fl, *_ = _grads(p3, w, _opts(p3, checkpoint_memory=64 * 400 * 8),
                       gn, gT)
    assert (fl == -111).any() and ((fl == 0) | (fl == -111)).all()


def _perturbed(w, R, h):
    """the cells of w once unperturbed, then with +h and -h in each log k_r,
    each log n_i(0) and log Tgas(0): ((1 + 2 P) ncell cells, lnk rows)"""
    nc, ns = w["n"].shape
    P = R + ns + 1
    n = np.tile(w["n"], (1 + 2 * P, 1))
    T = np.tile(w["T"], 1 + 2 * P)
    lnk = np.zeros(((1 + 2 * P) * nc, R))
    for p in range(P):
        for s, sg in ((1, 1.0), (2, -1.0)):
            blk = slice((1 + 2 * p + s - 1) * nc, (2 + 2 * p + s - 1) * nc)
            if p < R:
                lnk[blk, p] = sg * h
            elif p < R + ns:
                n[blk, p - R] *= np.exp(sg * h)
            else:
                T[blk] *= np.exp(sg * h)
    big = {"n": n, "T": T, "dt": np.tile(w["dt"], 1 + 2 * P),
           "par": np.tile(w["par"], (1 + 2 * P, 1))}
    return big, lnk, P


def _fd(L, nc, P, h):
    """central differences [nc, P] of the outputs L of _perturbed's cells,
    and the mismatch of the two one-sided ones (a discontinuity of the
    adaptive integration between the three runs when large)"""
    L0 = L[:nc]
    Lp = np.stack([L[(1 + 2 * p) * nc:(2 + 2 * p) * nc] for p in range(P)], 1)
    Lm = np.stack([L[(2 + 2 * p) * nc:(3 + 2 * p) * nc] for p in range(P)], 1)
    c = (Lp - Lm) / (2 * h)
    one = np.abs((Lp - L0[:, None]) - (L0[:, None] - Lm)) / h
    return c, one


def _settled(adj, fd5, fd6):
    """the difference quotients at h = 1e-5 and 1e-6 against the adjoint:
    den (the gradient, or 1e-3 of the cell's largest), the entries where
    the two steps agree within 1e-3 (elsewhere the solver's own noise, its
    Newton tolerance over h, or the response's curvature decides the
    quotient), and the adjoint's error against the h = 1e-6 quotient"""
    den = np.maximum(np.abs(fd6), 1e-3 * np.abs(fd6).max(axis=1,
                                                           keepdims=True))
    conv = np.abs(fd5 - fd6) <= 1e-3 * den
    return conv, np.abs(adj - fd6) / den, np.abs(fd5 - fd6) / den


@pytest.mark.gpu
def test_adjoint_vs_central_differences(p3):
    """the C4 check: 256 primordial3 collapse cells (seed 3), gradients of
    log(n_H2/nH) and log(Tgas) at t_end by the 27 log k_r, the 9 log n_i(0)
    and log Tgas(0), against central differences of the same kernel's
    forward at h = 1e-6 (and 1e-5 beside it), at rtol 1e-10 so that the
    solver's noise (its Newton tolerance, ~0.1 rtol of a state, over h) is
    below the 1e-3 bar.  Where the two steps agree within 1e-3 (most
    entries; elsewhere the quotient is noise or curvature, not a
    derivative), the adjoint is within 1e-3 plus that agreement for 99 %
    of the entries, its median error below 1e-5.  The tighter check of the
    adjoint's algebra is the replayed scheme's, on the CPU."""
    w = cells.primordial3_cells(256, seed=3)
    nc, ns = w["n"].shape
    R = p3.nrea
    o = _opts(p3, rtol=1e-10)
    fds = {}
    for h in (1e-5, 1e-6):
        big, lnk, P = _perturbed(w, R, h)
        say("adjoint fd: h", h, "forward of", big["n"].shape[0], "cells")
        r, n, T = _forward(p3, big, o, lnk)
        assert r == 0
        (L1, _, _), (L2, _, _) = _outputs(p3, n, T)
        fds[h] = (_fd(L1, nc, P, h)[0], _fd(L2, nc, P, h)[0])
    r, n, T = _forward(p3, w, o)
    for k, name in ((0, "log nH2/nH"), (1, "log T")):
        (_, gn, gT) = _outputs(p3, n, T)[k]
        say("adjoint fd: backward", name)
        rr, fl, _, _, gl, gn0, gT0, mem = _grads(p3, w, o, gn, gT)
        assert rr == 0 and (fl == 0).all()
        adj = np.hstack([gl, gn0 * w["n"], (gT0 * w["T"])[:, None]])
        assert np.isfinite(adj).all()
        conv, err, spread = _settled(adj, fds[1e-5][k], fds[1e-6][k])
        ok = err <= 1e-3 + spread
        say(name, "settled %.3f" % conv.mean(),
            "err on settled: median %.2e p99 %.2e max %.2e" % (
                np.median(err[conv]), np.quantile(err[conv], 0.99),
                err[conv].max()),
            "within 1e-3 + spread: %.4f" % ok[conv].mean(),
            "peak memory %.1f MB" % (mem / 1e6))
        assert conv.mean() > 0.7
        assert ok[conv].mean() > 0.99
        assert np.median(err[conv]) < 1e-5


@pytest.mark.gpu
def test_autograd_solve(p3):
    """adjoint.Solve under torch.autograd: the forward equals the raw
    entry point's, and lnk.grad, n0.grad, T0.grad equal its gradients for
    the cotangents torch hands to the backward, bit for bit"""
    import torch
    import adjoint
    w = cells.primordial3_cells(32, seed=3, first=1000)
    o = _opts(p3)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device="cuda")
    lnk = torch.zeros((32, p3.nrea), dtype=torch.float64, device="cuda",
                      requires_grad=True)
    n0 = t(w["n"]).requires_grad_(True)
    T0 = t(w["T"]).requires_grad_(True)
    solve = adjoint.Solve(p3, o)
    say("autograd")
    n1, T1, fl = solve(lnk, n0, T0, t(w["dt"]), t(w["par"]))
    n1.retain_grad()
    T1.retain_grad()
    iH2 = p3.idx("H2")
    wH = torch.tensor(p3.composition[:, p3.elements.index("H")],
                      dtype=torch.float64, device="cuda")
    loss = (torch.log(n1[:, iH2]) - torch.log(n1 @ wH)).sum() + \
        torch.log(T1).sum()
    loss.backward()
    assert (fl.cpu().numpy() == 0).all()
    n1c, T1c = n1.detach().cpu().numpy(), T1.detach().cpu().numpy()
    r, n, T = _forward(p3, w, o)
    assert np.array_equal(n, n1c) and np.array_equal(T, T1c)
    # the raw entry point with the cotangents torch handed to the backward
    rr, _, _, _, gl, gn0, gT0, _ = _grads(p3, w, o, n1.grad.cpu().numpy(),
                                          T1.grad.cpu().numpy())
    assert rr == 0
    for a, b in ((lnk.grad, gl), (n0.grad, gn0), (T0.grad, gT0)):
        assert np.array_equal(a.cpu().numpy(), b)
    assert solve.last_memory > 0


def _ref_fd(d, k):
    """the reference's central differences [2 steps][ncell, P] of output k
    (0: log(n_H2/nH), 1: log Tgas) from primordial3_adjoint.npz"""
    L = d["L1"] if k == 0 else d["L2"]
    steps = d["steps"]
    return [((L[i, :, 0] - L[i, :, 1]) / (2 * steps[i])).T
            for i in range(len(steps))]


def _ref_check(adj, d, k, name, cells_=None):
    fd5, fd6 = _ref_fd(d, k)
    if cells_ is not None:
        fd5, fd6 = fd5[cells_], fd6[cells_]
    conv, err, spread = _settled(adj, fd5, fd6)
    ok = err <= 1e-3 + spread
    say(name, "vs the reference's perturbed solves: settled %.3f" %
        conv.mean(), "median %.2e p99 %.2e" % (np.median(err[conv]),
                                                np.quantile(err[conv], 0.99)),
        "within 1e-3 + spread %.4f" % ok[conv].mean())
    assert conv.mean() > 0.7
    assert ok[conv].mean() > 0.99
    # the two solvers differ in the last bits of their rate expressions
    # (the forward agrees within 1e-6): a median far inside the bar
    assert np.median(err[conv]) < 1e-4


def test_adjoint_vs_reference_perturbed_host(host):
    """the adjoint (its host build) against the reference's own perturbed
    solves (tests/golden/primordial3_adjoint.npz, written by make_golden.py:
    primordial3 with a multiplier per reaction, each log k_r, log n_i(0)
    and log Tgas(0) moved by +-1e-5 and +-1e-6 at rtol 1e-10), on the first
    32 of its 256 C4 cells, both outputs"""
    d = golden("primordial3_adjoint.npz")
    sel = np.arange(32)
    w = {"n": d["n0"][sel], "T": d["T0"][sel], "dt": d["dt"][sel],
         "par": d["par"][sel]}
    nc, ns = w["n"].shape
    iH2 = 3
    wH = np.array([1, 1, 1, 2, 2, 0, 0, 0, 0.0])
    fl, n, T, *_ = _host_adjoint(host, w, np.zeros((nc, ns)), np.zeros(nc),
                                 rtol=float(d["rtol"]))
    assert (fl == 0).all()
    assert rel_err(n, d["n_final"][sel]) < 1e-6
    nH = n @ wH
    g1 = -wH[None, :] / nH[:, None]
    g1[:, iH2] += 1.0 / n[:, iH2]
    for k, (gn, gT) in enumerate(((g1, np.zeros(nc)),
                                  (np.zeros((nc, ns)), 1.0 / T))):
        fl, _, _, gl, gn0, gT0 = _host_adjoint(host, w, gn, gT,
                                               rtol=float(d["rtol"]))
        adj = np.hstack([gl, gn0 * w["n"], (gT0 * w["T"])[:, None]])
        _ref_check(adj, d, k, "host output %d" % k, sel)


@pytest.mark.gpu
def test_adjoint_vs_reference_perturbed(p3):
    """the GPU adjoint against the reference's perturbed solves, all 256
    C4 cells, both outputs (as the host test)"""
    d = golden("primordial3_adjoint.npz")
    w = {"n": d["n0"], "T": d["T0"], "dt": d["dt"], "par": d["par"]}
    o = _opts(p3, rtol=float(d["rtol"]))
    say("adjoint vs reference perturbed solves")
    r, n, T = _forward(p3, w, o)
    assert r == 0
    assert rel_err(n, d["n_final"]) < 1e-6
    for k in (0, 1):
        (_, gn, gT) = _outputs(p3, n, T)[k]
        rr, fl, _, _, gl, gn0, gT0, _ = _grads(p3, w, o, gn, gT)
        assert rr == 0 and (fl == 0).all()
        adj = np.hstack([gl, gn0 * w["n"], (gT0 * w["T"])[:, None]])
        _ref_check(adj, d, k, "output %d" % k)
