"""Seeded synthetic cells of the benchmark workloads.

A counter-based generator (splitmix64 of seed * 2^40 + 16 c + k) gives
draw k of cell c as a uniform in [0, 1) independent of how many cells are
made or in which order, so a sample of cells (the CPU baseline, the
fixtures) is the same cells as in the full batch.  Log-uniform draws are
10^(a + (b - a) u).

Workloads (the species order of the generated networks):

  primordial    4096 cells, seed 0: nH in 1..1e6 cm^-3, Tgas 10..1e4 K,
                x = n(H+)/nH = n(e-)/nH in 1e-6..1e-2, y = n(H2)/nH in
                1e-8..1e-3, He = 0.0775 nH, the other species 1e-20 nH,
                H = nH (1 - x - 2y); dt = 3.15e10 s; fixed Tgas
  primordial3   2^20 cells, seed 1: nH in 1..1e16, Tgas 10..3e3 K,
                x in 1e-8..1e-3, y in 1e-8..0.5, He = 0.0775 nH, the other
                species 1e-20 nH, H = nH max(1 - x - 2y, 1e-20);
                dt = 0.01 tff, tff = sqrt(3 pi / (32 G rho)) per cell
                (the compression time scale parameter), Tcmb 2.725,
                opratio 3
  c3_primordial 2^19 cells, seed 2: the primordial cells of the uneven
                batch, density log-uniform 1..1e12 cm^-3 and dt
                log-uniform 1e8..1e14 s (step counts spread over more than
                100x); Tgas 10..1e4 K, x 1e-6..1e-2, y 1e-8..1e-3 as
                primordial; fixed Tgas
  c3_primordialZ  2^18 cells, seed 2, cells 2^19 onwards: the primordialZ
                cells of the uneven batch, drawn as c3_primordial, with the
                metals C = 2.69e-6 nH and O = 4.90e-6 nH (neutral); fixed
                Tgas, the difference-quotient Jacobian (jac 2: its density-
                gated rates have no analytic dk/dn here), rtol 1e-6
  galaxy_ism    2^18 cells, seed 4: nH in 0.1..1e5 cm^-3, Tgas 10..1e4 K,
                x in 1e-4..1e-1, y in 1e-6..1e-1, He = 0.0775 nH with
                He+ = 1e-3 He, the other species 1e-20 nH (e- balancing
                the charge); crate 1e-17..1e-15 s^-1; dt = 3.15e11 s;
                Tgas evolved.  Photon bins: 32 logarithmic bins over
                0.7..200 eV (PHOTO_BINS); the field of a cell is
                f Phi0(E), f log-uniform in 1e-2..1e2 (also the cell's G0),
                Phi0 = 2e7 photons cm^-2 s^-1 per bin below 13.6 eV (an
                interstellar FUV field) and 1e4 (E/13.6)^-1.5 above it (a
                weak ionizing background)
  galaxy_ism_heat  the same cells for galaxy_ism with its thermal balance
                (heating CHEM + PHOTO, cooling H2 + CHEM)
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_META = _HERE / "csrc" / "generated" / "networks.json"
GRAVITY = 6.6743e-08


def network_meta(name):
    return json.loads(_META.read_text())[name]


def splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
    return x ^ (x >> np.uint64(31))


def uniform(seed, cells, k):
    """draw k of the given cells: uniform in [0, 1), 53 bits"""
    c = np.asarray(cells, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = splitmix64(np.uint64(seed) * np.uint64(1 << 40)
                       + c * np.uint64(16) + np.uint64(k))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def log_uniform(seed, cells, k, a, b):
    la, lb = math.log10(a), math.log10(b)
    return 10.0 ** (la + (lb - la) * uniform(seed, cells, k))


def _species_rows(meta, nH, x, y):
    sp = meta["species"]
    n = np.empty((nH.size, len(sp)))
    n[:] = 1e-20 * nH[:, None]
    col = {s: i for i, s in enumerate(sp)}
    n[:, col["H+"]] = x * nH
    n[:, col["e-"]] = x * nH
    n[:, col["H2"]] = y * nH
    n[:, col["He"]] = 0.0775 * nH
    n[:, col["H"]] = nH * np.maximum(1.0 - x - 2.0 * y, 1e-20)
    return n


def primordial_cells(ncell=4096, seed=0, first=0):
    """the fixed-Tgas primordial workload (cells first .. first+ncell-1)"""
    meta = network_meta("primordial")
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e6)
    T = log_uniform(seed, c, 1, 10.0, 1e4)
    x = log_uniform(seed, c, 2, 1e-6, 1e-2)
    y = log_uniform(seed, c, 3, 1e-8, 1e-3)
    n = _species_rows(meta, nH, x, y)
    dt = np.full(ncell, 3.15e10)
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    return {"n": n, "T": T, "dt": dt, "par": par,
            "opts": {"evolve_T": 0, "rtol": 1e-6, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 1, "jac": 0,
                     "max_steps": 100000}}


def primordial3_cells(ncell=1 << 20, seed=1, first=0):
    """the primordial3 collapse workload with Tgas evolved"""
    meta = network_meta("primordial3")
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e16)
    T = log_uniform(seed, c, 1, 10.0, 3e3)
    x = log_uniform(seed, c, 2, 1e-8, 1e-3)
    y = log_uniform(seed, c, 3, 1e-8, 0.5)
    n = _species_rows(meta, nH, x, y)
    rho = n @ np.array(meta["mass_g"])
    tff = np.sqrt(3.0 * math.pi / (32.0 * GRAVITY * np.maximum(rho, 1e-40)))
    dt = 0.01 * tff
    pn = meta["params"]
    par = np.empty((ncell, len(pn)))
    for j, p in enumerate(pn):
        par[:, j] = meta["param_defaults"][j]
    par[:, pn.index("tff")] = tff
    return {"n": n, "T": T, "dt": dt, "par": par,
            "opts": {"evolve_T": 1, "rtol": 1e-8, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 1, "jac": 0,
                     "max_steps": 100000}}


def c3_primordial_cells(ncell=1 << 19, seed=2, first=0):
    """the primordial part of the uneven-batch workload"""
    meta = network_meta("primordial")
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e12)
    T = log_uniform(seed, c, 1, 10.0, 1e4)
    x = log_uniform(seed, c, 2, 1e-6, 1e-2)
    y = log_uniform(seed, c, 3, 1e-8, 1e-3)
    n = _species_rows(meta, nH, x, y)
    dt = log_uniform(seed, c, 4, 1e8, 1e14)
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    return {"n": n, "T": T, "dt": dt, "par": par,
            "opts": {"evolve_T": 0, "rtol": 1e-6, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 1, "jac": 0,
                     "max_steps": 100000}}


def c3_primordialZ_cells(ncell=1 << 18, seed=2, first=1 << 19):
    """the primordialZ part of the uneven-batch workload (cells after the
    primordial part's): the primordial composition of c3_primordial and
    the metals C = 2.69e-6 nH, O = 4.90e-6 nH (a hundredth of solar),
    neutral; fixed Tgas"""
    meta = network_meta("primordialZ")
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e12)
    T = log_uniform(seed, c, 1, 10.0, 1e4)
    x = log_uniform(seed, c, 2, 1e-6, 1e-2)
    y = log_uniform(seed, c, 3, 1e-8, 1e-3)
    n = _species_rows(meta, nH, x, y)
    col = {s: i for i, s in enumerate(meta["species"])}
    n[:, col["C"]] = 2.69e-6 * nH
    n[:, col["O"]] = 4.90e-6 * nH
    dt = log_uniform(seed, c, 4, 1e8, 1e14)
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    return {"n": n, "T": T, "dt": dt, "par": par,
            "opts": {"evolve_T": 0, "rtol": 1e-6, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 1, "jac": 2,
                     "max_steps": 100000}}


# the cloud cells' elements other than H and He, as fractions of nH (the
# "low-metal" initial abundances of dark-cloud models; the ions as the
# singly charged species), and the grains
CLOUD_METALS = {"C+": 7.3e-5, "N": 2.14e-5, "O": 1.76e-4, "S+": 8e-8,
                "Si+": 8e-9, "Fe+": 3e-9, "Na+": 2e-9, "Mg+": 7e-9,
                "P+": 3e-9, "Cl+": 4e-9, "F": 6.68e-9}
CLOUD_GRAINS = 1.3e-12


def c3_cloud_cells(ncell=1 << 17, seed=2, first=(1 << 19) + (1 << 18)):
    """the cloud part of the uneven-batch workload (cells after the
    primordial and primordialZ parts'): density 1..1e12 cm^-3 and dt
    1e8..1e14 s as the other parts; Tgas 10..300 K (the range of the
    network's rate fits; fixed), H2/nH 0.1..0.5, H+ = x nH with x
    1e-8..1e-4, He 0.0775 nH, the metals of CLOUD_METALS, GRAIN0 1.3e-12
    nH, electrons for neutrality, the rest 1e-20 nH; parameters crate
    1.3e-17 s^-1, gas_dust_ratio 0.01 and Av 0.1..10 per cell"""
    meta = network_meta("cloud")
    sp = meta["species"]
    col = {s: i for i, s in enumerate(sp)}
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e12)
    T = log_uniform(seed, c, 1, 10.0, 300.0)
    x = log_uniform(seed, c, 2, 1e-8, 1e-4)
    y = log_uniform(seed, c, 3, 0.1, 0.5)
    dt = log_uniform(seed, c, 4, 1e8, 1e14)
    av = log_uniform(seed, c, 5, 0.1, 10.0)
    n = np.empty((ncell, len(sp)))
    n[:] = 1e-20 * nH[:, None]
    n[:, col["H2"]] = y * nH
    n[:, col["H+"]] = x * nH
    n[:, col["He"]] = 0.0775 * nH
    n[:, col["H"]] = nH * (1.0 - 2.0 * y - x)
    for s, f in CLOUD_METALS.items():
        n[:, col[s]] = f * nH
    n[:, col["GRAIN0"]] = CLOUD_GRAINS * nH
    q = np.array(meta["charge"], dtype=float)
    ie = col["e-"]
    n[:, ie] = 0.0
    n[:, ie] = n @ q
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    pc = {p: i for i, p in enumerate(meta["params"])}
    par[:, pc["crate"]] = 1.3e-17
    par[:, pc["gas_dust_ratio"]] = 0.01
    par[:, pc["Av"]] = av
    return {"n": n, "T": T, "dt": dt, "par": par,
            "opts": {"evolve_T": 0, "rtol": 1e-6, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 0, "jac": 0,
                     "max_steps": 100000}}


def photo_bins(nbins=32, emin=0.7, emax=200.0):
    """(Elo, Ehi) [eV] of nbins logarithmic bins over [emin, emax]"""
    e = emin * (emax / emin) ** (np.arange(nbins + 1) / nbins)
    return e[:-1].copy(), e[1:].copy()


PHOTO_BINS = photo_bins()


def galaxy_ism_cells(ncell=1 << 18, seed=4, first=0, network="galaxy_ism"):
    """the galaxy_ism photochemistry workload, with Tgas evolved: the
    cells, their parameter rows and their photon fields Phi[ncell][32]
    (network: galaxy_ism, or galaxy_ism_heat, the same network with its
    thermal balance)"""
    meta = network_meta(network)
    sp = meta["species"]
    col = {s: i for i, s in enumerate(sp)}
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 0.1, 1e5)
    T = log_uniform(seed, c, 1, 10.0, 1e4)
    x = log_uniform(seed, c, 2, 1e-4, 1e-1)
    y = log_uniform(seed, c, 3, 1e-6, 1e-1)
    f = log_uniform(seed, c, 4, 1e-2, 1e2)
    crate = log_uniform(seed, c, 5, 1e-17, 1e-15)
    n = np.empty((ncell, len(sp)))
    n[:] = 1e-20 * nH[:, None]
    he = 0.0775 * nH
    n[:, col["H+"]] = x * nH
    n[:, col["H2"]] = y * nH
    n[:, col["He"]] = he * (1.0 - 1e-3)
    n[:, col["He+"]] = he * 1e-3
    n[:, col["H"]] = nH * np.maximum(1.0 - x - 2.0 * y, 1e-20)
    q = np.array([meta["charge"][i] for i in range(len(sp))], dtype=float)
    q[col["e-"]] = 0.0
    n[:, col["e-"]] = n @ q
    dt = np.full(ncell, 3.15e11)
    pn = meta["params"]
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    par[:, pn.index("crate")] = crate
    par[:, pn.index("G0")] = f
    elo, ehi = PHOTO_BINS
    em = np.sqrt(elo * ehi)
    phi0 = np.where(ehi <= 13.6, 2e7, 1e4 * (em / 13.6) ** -1.5)
    Phi = f[:, None] * phi0[None, :]
    return {"n": n, "T": T, "dt": dt, "par": par, "Phi": Phi,
            "bins": (elo, ehi),
            "opts": {"evolve_T": 1, "rtol": 1e-8, "atol_n": 1e-20,
                     "atol_T": 1e-6, "linsol": 1, "jac": 0,
                     "max_steps": 100000}}


def generic_cells(name, ncell=256, seed=5, first=0):
    """cells of any generated network for parity checks: nH log-uniform in
    1..1e6 cm^-3, Tgas 10..1e4 K, each species but e- at 1e-8..1e-2 nH
    (log-uniform, its own draw), e- balancing the charge (at least
    1e-20 nH), dt log-uniform 1e8..1e12 s, the default parameters but
    J21 log-uniform in 0.1..10 where the network has it; Tgas
    fixed, rtol 1e-6, the difference-quotient Jacobian where a rate reads
    the densities (jac 2), else the analytic one; for a network with photo
    reactions the photon bins and fields of galaxy_ism_cells"""
    meta = network_meta(name)
    sp = meta["species"]
    c = np.arange(first, first + ncell)
    nH = log_uniform(seed, c, 0, 1.0, 1e6)
    T = log_uniform(seed, c, 1, 10.0, 1e4)
    dt = log_uniform(seed, c, 2, 1e8, 1e12)
    n = np.empty((ncell, len(sp)))
    for i, s in enumerate(sp):
        n[:, i] = nH * log_uniform(seed + 1000 * (i + 1), c, 0, 1e-8, 1e-2)
    if "e-" in sp:
        e = sp.index("e-")
        q = np.array(meta["charge"], dtype=float)
        q[e] = 0.0
        n[:, e] = np.maximum(n @ q, 1e-20 * nH)
    par = np.tile(np.array(meta["param_defaults"], dtype=np.float64),
                  (ncell, 1))
    if "J21" in meta["params"]:  # a UV field for the shielded H2 rates
        par[:, meta["params"].index("J21")] = log_uniform(seed, c, 4, 0.1,
                                                          10.0)
    w = {"n": n, "T": T, "dt": dt, "par": par,
         "opts": {"evolve_T": 0, "rtol": 1e-6, "atol_n": 1e-20,
                  "atol_T": 1e-6, "linsol": 1,
                  "jac": 2 if meta["rates_density"] else 0,
                  "max_steps": 100000}}
    if meta["photo"]:
        elo, ehi = PHOTO_BINS
        f = log_uniform(seed, c, 3, 1e-2, 1e2)
        em = np.sqrt(elo * ehi)
        phi0 = np.where(ehi <= 13.6, 2e7, 1e4 * (em / 13.6) ** -1.5)
        w["bins"] = (elo, ehi)
        w["Phi"] = f[:, None] * phi0[None, :]
    return w


# the other generated networks, each with the generic cells
EXTRA_NETWORKS = ("NO", "SM", "coolHCN", "lowmetal", "primordial2",
                  "primordial_wD", "primordial_photo", "primordial_photoH2",
                  "auto", "chianti", "ions", "ionsH2", "primordial_UV",
                  "xrays", "primordialZ_UV", "photobin")

WORKLOADS = {"primordial": primordial_cells, "primordial3": primordial3_cells,
             "c3_primordial": c3_primordial_cells,
             "galaxy_ism": galaxy_ism_cells,
             "galaxy_ism_heat": lambda ncell=1 << 18, seed=4, first=0:
                 galaxy_ism_cells(ncell, seed, first, "galaxy_ism_heat"),
             "c3_primordialZ": c3_primordialZ_cells,
             "c3_cloud": c3_cloud_cells}
for _n in EXTRA_NETWORKS:
    WORKLOADS[_n] = (lambda name: lambda ncell=256, seed=5, first=0:
                     generic_cells(name, ncell, seed, first))(_n)
