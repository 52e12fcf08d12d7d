import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
PKG = REPO / "krome-x_amd"
GOLDEN = REPO / "tests" / "golden"
sys.path.insert(0, str(PKG))

# the options row of the oracle's C entry points (oracle/kx_oracle.cpp)
OPT_ORDER = ("evolve_T", "rtol", "atol_n", "atol_T", "linsol", "jac",
             "max_steps", "max_err_fails", "max_conv_fails", "max_restarts",
             "clip_negative", "conserve")
NET_ID = {"primordial": 0, "primordial3": 1, "galaxy_ism": 2,
          "primordialZ": 3, "galaxy_ism_heat": 4, "NO": 5, "SM": 6,
          "coolHCN": 7, "lowmetal": 8, "primordial2": 9, "primordial_wD": 10,
          "primordial_photo": 11, "primordial_photoH2": 12, "auto": 13,
          "chianti": 14, "ions": 15, "ionsH2": 16, "primordial_UV": 17,
          "xrays": 18, "primordialZ_UV": 19,
          "photobin": 20, "cloud": 21}


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an AMD GPU (the HIP kernels)")


def oracle_opts(**kw):
    d = dict(evolve_T=0, rtol=1e-8, atol_n=1e-20, atol_T=1e-6, linsol=1,
             jac=0, max_steps=100000, max_err_fails=7, max_conv_fails=10,
             max_restarts=5, clip_negative=1, conserve=0)
    d.update(kw)
    return np.array([float(d[k]) for k in OPT_ORDER])


class Oracle:
    """ctypes front of oracle/_build/libkx_oracle.so (or the reference-
    integrator build oracle/_ref/libkx_ref.so)"""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(str(path))
        self.p = prefix

    def solve(self, net, n, T, dt, par, opts, tol=None):
        """(return value, n, T, flags, stats) of solve_cells; tol: the
        per-species (atol[NSPEC], rtol[NSPEC]) each species gets (an entry
        None: the scalar for every species)"""
        D = ctypes.POINTER(ctypes.c_double)
        n = np.array(n, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, order="C", copy=True)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        par = np.ascontiguousarray(par, dtype=np.float64)
        nc = n.shape[0]
        flags = np.zeros(nc, dtype=np.intc)
        st = np.zeros((nc, 9), dtype=np.int64)
        head = (NET_ID[net], nc, n.ctypes.data_as(D), T.ctypes.data_as(D),
                dt.ctypes.data_as(D), par.ctypes.data_as(D),
                par.shape[1] if par.ndim == 2 else 0,
                np.ascontiguousarray(opts).ctypes.data_as(D))
        tail = (flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                st.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        if tol is None:
            r = getattr(self.lib, self.p + "solve_cells")(*head, *tail)
        else:
            keep = [None if t is None else
                    np.ascontiguousarray(t, dtype=np.float64) for t in tol]
            r = getattr(self.lib, self.p + "solve_cells_tol")(
                *head, *(None if t is None else t.ctypes.data_as(D)
                         for t in keep), *tail)
        return r, n, T, flags, st

    def funcs(self, net, n, T, par, evolve_T):
        D = ctypes.POINTER(ctypes.c_double)
        nc, ns = n.shape
        meta = __import__("cells").network_meta(net)
        k = np.zeros((nc, meta["nrea"]))
        f = np.zeros((nc, ns + 1))
        getattr(self.lib, self.p + "funcs")(
            NET_ID[net], nc, np.ascontiguousarray(n).ctypes.data_as(D),
            np.ascontiguousarray(T).ctypes.data_as(D),
            np.ascontiguousarray(par).ctypes.data_as(D), int(evolve_T),
            k.ctypes.data_as(D), f.ctypes.data_as(D))
        return k, f

    def photo_tables(self, net, elo, ehi, alpha=1.0):
        """the bin-averaged cross sections [NPHOTO, B] of the bins, then
        their photoelectron-energy averages: [2 NPHOTO, B]"""
        D = ctypes.POINTER(ctypes.c_double)
        elo = np.ascontiguousarray(elo, dtype=np.float64)
        ehi = np.ascontiguousarray(ehi, dtype=np.float64)
        nph = len(__import__("cells").network_meta(net)["photo"])
        sig = np.zeros((2 * nph, elo.size))
        r = getattr(self.lib, self.p + "photo_bins")(
            NET_ID[net], elo.size, elo.ctypes.data_as(D),
            ehi.ctypes.data_as(D), ctypes.c_double(alpha),
            sig.ctypes.data_as(D))
        assert r == 0
        return sig

    def photo_bins(self, net, elo, ehi, alpha=1.0):
        """the bin-averaged cross sections [NPHOTO, B] of the bins"""
        nph = len(__import__("cells").network_meta(net)["photo"])
        return self.photo_tables(net, elo, ehi, alpha)[:nph]

    def photo_rows(self, net, par, Phi, sigma):
        """the solve's rows of a photo network: par then the bin sums
        (sigma: the tables, both halves with PHOTO heating)"""
        D = ctypes.POINTER(ctypes.c_double)
        par = np.ascontiguousarray(par, dtype=np.float64)
        Phi = np.ascontiguousarray(Phi, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        nc = Phi.shape[0]
        meta = __import__("cells").network_meta(net)
        nq = len(meta["photo"]) * (2 if "PHOTO" in meta["heating"] else 1)
        assert sigma.shape[0] >= nq
        out = np.zeros((nc, par.shape[-1] + nq))
        r = getattr(self.lib, self.p + "photo_rows")(
            NET_ID[net], nc, par.ctypes.data_as(D),
            par.shape[1] if par.ndim == 2 else 0, Phi.ctypes.data_as(D),
            Phi.shape[1], sigma.ctypes.data_as(D), out.ctypes.data_as(D))
        assert r == 0
        return out

    def totals(self, net, v, x=False):
        D = ctypes.POINTER(ctypes.c_double)
        v = np.ascontiguousarray(v, dtype=np.float64)
        nel = len(__import__("cells").network_meta(net)["elements"])
        tot = np.zeros((v.shape[0], nel))
        r = getattr(self.lib, self.p + "totals")(
            NET_ID[net], int(x), v.shape[0], v.ctypes.data_as(D),
            tot.ctypes.data_as(D))
        assert r == 0
        return tot

    def conserve(self, net, v, ref, x=False):
        """(return value, corrected rows, flags); ref [nelem] (one row for
        all) or [ncell, nelem]"""
        D = ctypes.POINTER(ctypes.c_double)
        v = np.array(v, dtype=np.float64, order="C", copy=True)
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        fl = np.zeros(v.shape[0], dtype=np.intc)
        r = getattr(self.lib, self.p + "conserve")(
            NET_ID[net], int(x), v.shape[0], v.ctypes.data_as(D),
            ref.ctypes.data_as(D), 1 if ref.ndim == 2 else 0,
            fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return r, v, fl


@pytest.fixture(scope="session")
def oracle():
    so = REPO / "oracle" / "_build" / "libkx_oracle.so"
    if not so.exists():
        subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "all"],
                       check=True)
    return Oracle(so, "kxo_")


@pytest.fixture(scope="session")
def oracle_ref():
    so = REPO / "oracle" / "_ref" / "libkx_ref.so"
    if not so.exists():
        src = Path("/root/reference/runtime/core/cvode/bdf_cell.hpp")
        if not src.exists():
            pytest.skip("the reference integrator build is not available")
        subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "ref"],
                       check=True)
    return Oracle(so, "kxr_")


def golden(name):
    return np.load(GOLDEN / name)


def rel_err(a, b, ntot_frac=1e-12):
    """max relative difference over the species above ntot_frac of each
    cell's total (smaller ones are below the absolute tolerance)"""
    b = np.asarray(b)
    ntot = np.abs(b).sum(axis=-1, keepdims=True)
    big = np.abs(b) > ntot_frac * ntot
    rel = np.abs(np.asarray(a) - b) / np.maximum(np.abs(b), 1e-300)
    return float(rel[big].max()) if big.any() else 0.0
