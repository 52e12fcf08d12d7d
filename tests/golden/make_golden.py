"""Write the golden fixtures of tests/golden from the reference itself.

Run in the container that has the reference (read-only) under
/root/reference:  python3 tests/golden/make_golden.py

For each workload it:
  1. runs the reference's code generator (python3 -m kromex, with the
     workload's thermal options) into tests/golden/_build/<name>;
  2. compiles ref_driver.cpp against that project with g++, twice: the
     bitwise flags of the reference's test builds (-O2 -ffp-contract=off,
     codegen.py:9122-9123) and the production flags (-O3
     -ffp-contract=fast -fopenmp-simd, codegen.py:9126), both with -fopenmp;
  3. solves a seeded sample of the workload's cells with
     krome::solve_cells (PATH_PERCELL) and evaluates the rates, the full
     rhs, the analytic Jacobian and the thermal terms at sample states;
  4. writes <name>_solve.npz and <name>_funcs.npz (inputs and outputs
     only: numbers, no reference code);
  5. for the conservation API (runtime/core/query.hpp:152-170,
     solve_cell.hpp:1331-1562, solve_cells.hpp:216-260): the element
     tables, krome::get_element_totals(_x) and krome::conserve_cells(_x) of
     seeded perturbed cells, and solve_cells with SolverOpts::conserve 2
     on the solve fixture's cells -> <name>_conserve.npz;
  6. the static-pivot sparse LU (LINSOL_SPARSE, bdf_cell.hpp:1155-1264):
     the generated tables, and solve_cells with linsol 3 on the solve
     fixture's cells -> <name>_sparse.npz.

  7. for the adjoint (C4): primordial3 with one rate multiplier per reaction
     line, checked to scale its own rate only, solved unperturbed and with
     each log k_r, log n_i(0) and log Tgas(0) moved by +-h (h = 1e-5, 1e-6,
     rtol 1e-10) -> primordial3_adjoint.npz (the outputs log(n_H2/nH) and
     log(Tgas) of every solve).

  8. per-species tolerances (SolverOpts::atol_species / rtol_species,
     solve_cell.hpp:866-879, 2129-2139): primordial and primordial3 solved
     with per-species atol only, then with per-species atol and rtol (some
     entries negative: the scalar) -> <name>_tol.npz.

  9. the uneven batch's cloud network (see make_cloud) -> cloud_*.npz.

  python3 tests/golden/make_golden.py [solve] [funcs] [conserve] [sparse]
  [photo] [extra] [adjoint] [tol] [cloud] (default: all but extra)

The generated project and the binaries stay in tests/golden/_build/ (git-
ignored, not committed).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "krome-x_amd"))
import cells  # noqa: E402

REF = Path("/root/reference")
BUILD = HERE / "_build"

TOML = {
    "primordial": '[project]\nnetwork = "{net}/primordial.knet"\n'
                  'outdir = "{out}"\n',
    "primordial3": '[project]\nnetwork = "{net}/primordial3.knet"\n'
                   'outdir = "{out}"\n[thermal]\n'
                   'cooling = ["H2", "CONT", "CHEM"]\n'
                   'heating = ["COMPRESS", "CHEM"]\ngamma = "EXACT"\n'
                   'h2_opacity = "RIPAMONTI"\n',
    "galaxy_ism": '[project]\nnetwork = "{net}/galaxy_ism.knet"\n'
                  'outdir = "{out}"\n',
    "primordialZ": '[project]\nnetwork = "{net}/primordialZ.knet"\n'
                   'outdir = "{out}"\n',
    "galaxy_ism_heat": '[project]\nnetwork = "{net}/galaxy_ism.knet"\n'
                       'outdir = "{out}"\n[thermal]\n'
                       'cooling = ["H2", "CHEM"]\n'
                       'heating = ["CHEM", "PHOTO"]\n',
    "cloud": '[project]\nnetwork = "{net}/cloud.knet"\noutdir = "{out}"\n',
}
# cells of the solve fixtures (the first ones of the seeded workload)
NSOLVE = {"primordial": 1024, "primordial3": 256, "primordialZ": 256,
          "cloud": 64}
# the workload of a network's fixtures (cells.WORKLOADS)
WORKLOAD = {"primordial": "primordial", "primordial3": "primordial3",
            "primordialZ": "c3_primordialZ", "cloud": "c3_cloud"}
OPT_KEYS = ("evolve_T", "rtol", "atol_n", "atol_T", "linsol", "jac",
            "max_steps")
T_GRID = (12.0, 97.0, 314.0, 1.1e3, 5.5e3, 2.3e4, 1.7e5, 3e6, 8e7)


def generate(name):
    out = BUILD / name
    out.mkdir(parents=True, exist_ok=True)
    cfg = BUILD / ("%s.toml" % name)
    cfg.write_text(TOML[name].format(net=REF / "networks", out=out))
    subprocess.run([sys.executable, "-m", "kromex", str(cfg)], cwd=REF,
                   check=True)
    return out


def compile_driver(proj, name, mode):
    flags = (["-O2", "-ffp-contract=off"] if mode == "bitwise"
             else ["-O3", "-ffp-contract=fast", "-fopenmp-simd"])
    exe = BUILD / ("ref_driver_%s_%s" % (name, mode))
    data = next(proj.glob("net_*_data.cpp"))
    subprocess.run(["g++", "-std=c++17", *flags, "-fopenmp", "-I", str(proj),
                    str(HERE / "ref_driver.cpp"), str(data), "-o", str(exe)],
                   check=True)
    return exe


def write_in(path, w, opts=None, extra=None):
    n, T, dt, par = w["n"], w["T"], w["dt"], w["par"]
    if opts is None:
        opts = np.array([float(w["opts"][k]) for k in OPT_KEYS])
    opts = np.asarray(opts, dtype=np.float64)
    with open(path, "wb") as f:
        np.array([n.shape[0], n.shape[1], par.shape[1], opts.size],
                 dtype=np.int64).tofile(f)
        opts.tofile(f)
        for a in (n, T, dt, par):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        if extra is not None:
            np.ascontiguousarray(extra, dtype=np.float64).tofile(f)


def run_solve(exe, w, tag, opts=None):
    fin = BUILD / ("in_%s.bin" % tag)
    fout = BUILD / ("out_%s.bin" % tag)
    write_in(fin, w, opts)
    env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1))
    subprocess.run([str(exe), "solve", str(fin), str(fout)], check=True,
                   env=env)
    raw = np.fromfile(fout, dtype=np.float64)
    nc, ns = w["n"].shape
    o = 0
    n = raw[o:o + nc * ns].reshape(nc, ns)
    o += nc * ns
    T = raw[o:o + nc]
    o += nc
    flag = raw[o:o + nc].astype(np.int32)
    o += nc
    stats = raw[o:o + nc * 9].reshape(nc, 9).astype(np.int64)
    o += nc * 9
    return n, T, flag, stats, raw[o]


def func_samples(name):
    """workload cells, and the same densities over the T grid (every
    piece, clamp and cut of the rates)"""
    gen = cells.WORKLOADS[WORKLOAD.get(name, name)]
    w = gen(64, first=0)
    base = gen(len(T_GRID) * 8, first=5000)
    T = np.repeat(np.array(T_GRID), 8)
    n = np.concatenate([w["n"], base["n"]])
    Tall = np.concatenate([w["T"], T])
    par = np.concatenate([w["par"], base["par"]])
    return {"n": n, "T": Tall, "dt": np.ones(n.shape[0]), "par": par,
            "opts": w["opts"]}


def run_funcs(exe, w, tag, nrea, neq):
    fin = BUILD / ("in_%s.bin" % tag)
    fout = BUILD / ("out_%s.bin" % tag)
    write_in(fin, w)
    subprocess.run([str(exe), "funcs", str(fin), str(fout)], check=True)
    raw = np.fromfile(fout, dtype=np.float64)
    per = nrea + neq + neq * neq + 8
    raw = raw.reshape(-1, per)
    return (raw[:, :nrea], raw[:, nrea:nrea + neq],
            raw[:, nrea + neq:nrea + neq + neq * neq].reshape(-1, neq, neq,
                                                              order="F"),
            raw[:, nrea + neq + neq * neq:])


# cells of the conservation fixtures; perturbation group = cell % 8
NCONSERVE = 512


def run_mode(exe, mode, w, tag, opts, extra=None):
    fin = BUILD / ("in_%s.bin" % tag)
    fout = BUILD / ("out_%s.bin" % tag)
    write_in(fin, w, opts, extra)
    subprocess.run([str(exe), mode, str(fin), str(fout)], check=True)
    return np.fromfile(fout, dtype=np.float64), fout


def ref_tables(exe, name, w):
    raw, fout = run_mode(exe, "meta", w, name + "_meta", [0.0])
    nelem = int(raw[0])
    ns = w["n"].shape[1]
    o = 1
    elements = raw[o:o + ns * nelem].reshape(ns, nelem).astype(np.int64)
    o += ns * nelem
    mass = raw[o:o + ns]
    charge = raw[o + ns:o + 2 * ns].astype(np.int64)
    names = Path(str(fout) + ".names").read_text().split()
    return {"element_names": np.array(names), "species_elements": elements,
            "species_mass_g": mass, "species_charge": charge}


def perturb(name, v, rng):
    """seeded perturbations of the rows v by group (cell % 8): 0-2 each
    species x (1 +- 30 %); 3 the same and two species negative; 4 the
    hydrogen carriers x 3 (a large change); 5 no helium carriers (the
    helium total cannot be reached); 6 H- ten times the cations (anions
    dominate the charge); 7 the 30 % noise only (its reference row is the
    next cell's).  The last cell is all zero"""
    sp = cells.network_meta(name)["species"]
    col = {s: i for i, s in enumerate(sp)}
    nc, ns = v.shape
    p = v * (1.0 + 0.3 * (2.0 * rng.random(v.shape) - 1.0))
    g = np.arange(nc) % 8
    hyd = [col[s] for s in sp if s[0] == "H" and not s.startswith("He")]
    hel = [col[s] for s in sp if s.startswith("He")]
    cat = [i for i, s in enumerate(sp) if s.endswith("+")]
    for c in range(nc):
        if g[c] == 3:
            for i in rng.choice(ns, 2, replace=False):
                p[c, i] = -0.1 * abs(p[c, i]) - 1e-30
        elif g[c] == 4:
            p[c, hyd] *= 3.0
        elif g[c] == 5:
            p[c, hel] = 0.0
        elif g[c] == 6:
            p[c, col["H-"]] = 10.0 * p[c, cat].sum()
    p[-1] = 0.0
    return p


def make_conserve(name, exe, meta):
    w = cells.WORKLOADS[name](NCONSERVE)
    if "D" in meta["species"]:
        # deuterium at a cosmic abundance in half of the cells (the
        # workload's 1e-20 nH is below the projection's pivot threshold,
        # 1e-14 of the largest diagonal: the D row stays unreached there)
        col = {s: i for i, s in enumerate(meta["species"])}
        n = w["n"].copy()
        rows = (np.arange(NCONSERVE) // 8) % 2 == 0
        nD = 2.6e-5 * n[rows, col["H"]]
        n[rows, col["D"]] = nD
        n[rows, col["D+"]] = 1e-4 * nD
        n[rows, col["HD"]] = 1e-3 * nD
        w = dict(w, n=n)
    tabs = ref_tables(exe, name, w)
    nelem = len(tabs["element_names"])
    nc = NCONSERVE
    # the totals of the workload cells: the reference rows of the cells
    raw, _ = run_mode(exe, "totals", w, name + "_totals", [0.0])
    ref_n = raw[:nc * nelem].reshape(nc, nelem)
    mass = np.array(meta["mass_g"])
    rho = w["n"] @ mass
    x0 = w["n"] * mass[None, :] / rho[:, None]
    wx = dict(w, n=x0)
    raw, _ = run_mode(exe, "totals", wx, name + "_totals_x", [0.0])
    ref_x = raw[nc * nelem:].reshape(nc, nelem)
    ref_n = ref_n.copy()
    ref_x = ref_x.copy()
    g = np.arange(nc) % 8
    shift = np.where(g == 7)[0]
    ref_n[shift] = ref_n[(shift + 1) % nc]
    ref_x[shift] = ref_x[(shift + 1) % nc]
    rng = np.random.default_rng(20260 + len(meta["species"]))
    pn = perturb(name, w["n"], rng)
    px = perturb(name, x0, rng)
    # the totals of the perturbed rows (negatives included), as n and as x
    raw, _ = run_mode(exe, "totals", dict(w, n=pn), name + "_totals_p",
                      [0.0])
    tot_p, totx_p = (raw[:nc * nelem].reshape(nc, nelem),
                     raw[nc * nelem:].reshape(nc, nelem))
    out = {}
    for mode, v0, ref in (("n", pn, ref_n), ("x", px, ref_x)):
        raw, _ = run_mode(exe, "conserve", dict(w, n=v0),
                          "%s_conserve_%s" % (name, mode),
                          [1.0 if mode == "x" else 0.0], extra=ref)
        ns = v0.shape[1]
        out["in_" + mode] = v0
        out["ref_" + mode] = ref
        out["out_" + mode] = raw[:nc * ns].reshape(nc, ns)
        out["flag_" + mode] = raw[nc * ns:].astype(np.int32)
        print("%s conserve %s: flags %s" % (
            name, mode, dict(zip(*np.unique(out["flag_" + mode],
                                            return_counts=True)))))
    # solve_cells with SolverOpts::conserve 2 on the solve fixture's cells
    ws = cells.WORKLOADS[name](NSOLVE[name])
    opts = [float(ws["opts"][k]) for k in OPT_KEYS] + [2.0]
    n2, T2, f2, _, _ = run_solve(exe, ws, name + "_solve_c2", opts)
    np.savez_compressed(
        HERE / ("%s_conserve.npz" % name), par=w["par"], n0=w["n"],
        x0=x0, rho=rho, tot_p=tot_p, totx_p=totx_p,
        solve_n0=ws["n"], solve_T0=ws["T"], solve_dt=ws["dt"],
        solve_par=ws["par"], solve_opts=np.array(opts), solve_n=n2,
        solve_T=T2, solve_flag=f2, **tabs, **out)
    print("%s conserve-2 solve: flags %s" % (name, np.unique(f2)))


def make_sparse(name, exe):
    w = cells.WORKLOADS[WORKLOAD[name]](NSOLVE[name])
    raw, _ = run_mode(exe, "sparse", w, name + "_sparse_tab", [0.0])
    ns = w["n"].shape[1]
    nnz = int(raw[0])
    o = 2
    tab = {}
    for k, m in (("perm", ns), ("rp", ns + 1), ("ci", nnz), ("di", ns)):
        tab["sp_" + k] = raw[o:o + m].astype(np.int64)
        o += m
    opts = [float(w["opts"][k]) for k in OPT_KEYS]
    opts[OPT_KEYS.index("linsol")] = 3.0
    n, T, f, st, secs = run_solve(exe, w, name + "_solve_sparse", opts)
    np.savez_compressed(HERE / ("%s_sparse.npz" % name), n0=w["n"],
                        T0=w["T"], dt=w["dt"], par=w["par"],
                        opts=np.array(opts), n=n, T=T, flag=f, stats=st,
                        sparse_flops=raw[1], **tab)
    print("%s sparse: nnz %d, flops %d, solve flags %s, steps %d" % (
        name, nnz, raw[1], np.unique(f), st[:, 0].sum()))


# per-species tolerances (krome_opts_set_atol_species / rtol_species):
# each network's tolerance rows, negative entries taking the scalar
NTOL = 256


def species_tols(meta, scalar_rtol):
    """(atol[NSPEC], rtol[NSPEC]) of the per-species fixtures: tighter
    absolute tolerances on the trace ions, a looser one on He, relative
    tolerances that differ by species, and some entries negative (the
    scalar)"""
    sp = meta["species"]
    atol = np.full(len(sp), -1.0)
    rtol = np.full(len(sp), -1.0)
    for i, name in enumerate(sp):
        if name in ("H-", "H2+", "He++", "D+"):
            atol[i] = 1e-26
        elif name == "He":
            atol[i] = 1e-14
        elif name == "H2":
            atol[i] = 1e-22
            rtol[i] = scalar_rtol * 0.1
        elif name == "e-":
            rtol[i] = scalar_rtol * 0.5
        elif name == "H+":
            rtol[i] = scalar_rtol * 4.0
    return atol, rtol


def make_tol(name, exe, meta):
    w = cells.WORKLOADS[WORKLOAD[name]](NTOL)
    base = [float(w["opts"][k]) for k in OPT_KEYS] + [0.0]
    atol, rtol = species_tols(meta, base[OPT_KEYS.index("rtol")])
    out = {"n0": w["n"], "T0": w["T"], "dt": w["dt"], "par": w["par"],
           "atol_species": atol, "rtol_species": rtol}
    none = np.full(len(atol), -1.0)
    for tag, a, r in (("atol", atol, none), ("both", atol, rtol)):
        opts = np.array(base + list(a) + list(r))
        n, T, f, st, secs = run_solve(exe, w, "%s_tol_%s" % (name, tag),
                                      opts)
        out.update({"opts_" + tag: opts, "n_" + tag: n, "T_" + tag: T,
                    "flag_" + tag: f, "stats_" + tag: st})
        print("%s tol %s: flags %s, steps %d" % (name, tag, np.unique(f),
                                                 st[:, 0].sum()))
    np.savez_compressed(HERE / ("%s_tol.npz" % name), **out)


def make_jac1(name, exes):
    """the network's solve fixture cells with jac 1 (JAC_ANALYTIC: with
    Tgas evolved the analytic species block, the density-dependent rates'
    derivatives, and the Tgas column and row by difference quotients)"""
    w = cells.WORKLOADS[WORKLOAD[name]](NSOLVE[name])
    opts = np.array([float(w["opts"][k]) for k in OPT_KEYS])
    opts[OPT_KEYS.index("jac")] = 1.0
    out = {"n0": w["n"], "T0": w["T"], "dt": w["dt"], "par": w["par"],
           "opts": opts}
    for m, exe in exes.items():
        n, T, flag, stats, secs = run_solve(exe, w, "%s_jac1_%s" % (name, m),
                                            opts)
        out.update({"n_" + m: n, "T_" + m: T, "flag_" + m: flag,
                    "stats_" + m: stats, "seconds_" + m: secs})
        print("%s jac 1 %s: flags %s, steps %d" % (name, m, np.unique(flag),
                                                   stats[:, 0].sum()))
    np.savez_compressed(HERE / ("%s_jac1.npz" % name), **out)


def make_cloud(exes, meta):
    """the uneven batch's cloud network (452 species, 4429 reactions):
    its sparse-LU tables and element tables, the rates, full rhs and
    thermal terms at 16 workload cells and 9 cells over the T grid, the
    analytic Jacobian at the first 4 of them (J is 452^2 per sample), and
    solve_cells of the first NSOLVE cells (both builds) with linsol auto
    (sparse here) -> cloud_funcs.npz, cloud_solve.npz, cloud_sparse.npz,
    cloud_tables.npz"""
    name = "cloud"
    gen = cells.WORKLOADS[WORKLOAD[name]]
    w = gen(16)
    base = gen(len(T_GRID), first=5000)
    fw = {"n": np.concatenate([w["n"], base["n"]]),
          "T": np.concatenate([w["T"], np.array(T_GRID)]),
          "par": np.concatenate([w["par"], base["par"]]),
          "opts": w["opts"]}
    fw["dt"] = np.ones(fw["n"].shape[0])
    ns = len(meta["species"])
    k, f, J, th = run_funcs(exes["bitwise"], fw, name + "_funcs",
                            meta["nrea"], ns)
    np.savez_compressed(HERE / "cloud_funcs.npz", n=fw["n"], T=fw["T"],
                        par=fw["par"], k=k, f=f, J=J[:4], thermal=th)
    print("cloud funcs: %d samples" % k.shape[0])
    np.savez_compressed(HERE / "cloud_tables.npz",
                        **ref_tables(exes["bitwise"], name, w))
    raw, _ = run_mode(exes["bitwise"], "sparse", w, "cloud_sparse_tab", [0.0])
    nnz = int(raw[0])
    o = 2
    tab = {}
    for key, m in (("perm", ns), ("rp", ns + 1), ("ci", nnz), ("di", ns)):
        tab["sp_" + key] = raw[o:o + m].astype(np.int64)
        o += m
    np.savez_compressed(HERE / "cloud_sparse.npz", sparse_flops=raw[1],
                        **tab)
    sw = gen(NSOLVE[name])
    res = {}
    for m, exe in exes.items():
        n, T, flag, stats, secs = run_solve(exe, sw, "cloud_%s" % m)
        res.update({"n_" + m: n, "T_" + m: T, "flag_" + m: flag,
                    "stats_" + m: stats, "seconds_" + m: secs})
        print("cloud %s: %d cells in %.2f s, flags %s, steps max %d" % (
            m, n.shape[0], secs, np.unique(flag), stats[:, 0].max()))
    np.savez_compressed(HERE / "cloud_solve.npz", n0=sw["n"], T0=sw["T"],
                        dt=sw["dt"], par=sw["par"],
                        opts=np.array([float(sw["opts"][k])
                                       for k in OPT_KEYS]), **res)


# photon bins: cells of the photo solve fixture, the energies at which the
# cross sections are sampled, the in-bin spectra of the bin tables
NPHOTO_SOLVE = 256
PHOTO_ALPHAS = (1.0, 0.0, 2.0)


def photo_energies(meta):
    """a log grid over 0.3..3e4 eV and each photo window's edges with
    their neighbours"""
    E = list(np.geomspace(0.3, 3e4, 3000))
    for ph in meta["photo"]:
        for e in (ph["Emin"], ph["Emax"]):
            if np.isfinite(e):
                E += [e * (1 - 1e-9), e, e * (1 + 1e-9)]
    return np.array(sorted(E))


def run_photo(exe, mode, w, tag, alpha, E):
    """ref_driver pbins / prates / psolve over w's bins and fields"""
    fin = BUILD / ("in_%s.bin" % tag)
    fout = BUILD / ("out_%s.bin" % tag)
    write_in(fin, w)
    elo, ehi = w["bins"]
    with open(fin, "ab") as f:
        np.array([len(elo)], dtype=np.int64).tofile(f)
        np.array([alpha], dtype=np.float64).tofile(f)
        for a in (elo, ehi, w["Phi"]):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        np.array([len(E)], dtype=np.int64).tofile(f)
        np.ascontiguousarray(E, dtype=np.float64).tofile(f)
    env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1))
    subprocess.run([str(exe), mode, str(fin), str(fout)], check=True,
                   env=env)
    return np.fromfile(fout, dtype=np.float64)


def make_photo(name, exes, meta):
    """galaxy_ism's photon bins: the reference's cross sections, bin tables
    (three in-bin spectra), rates with the fields, and solves"""
    exe = exes["bitwise"]
    w = cells.WORKLOADS[name](NPHOTO_SOLVE)
    E = photo_energies(meta)
    B = len(w["bins"][0])
    out = {"E": E, "Elo": w["bins"][0], "Ehi": w["bins"][1]}
    for alpha in PHOTO_ALPHAS:
        raw = run_photo(exe, "pbins", w, name + "_pbins", alpha, E)
        nph = len(meta["photo"])
        o = nph * B
        out["sigma_a%g" % alpha] = raw[:o].reshape(nph, B)
        out["sigmaE_a%g" % alpha] = raw[o:2 * o].reshape(nph, B)
        o *= 2
        out["xsec"] = raw[o:o + nph * len(E)].reshape(nph, len(E))
        o += nph * len(E)
        tail = raw[o:]
        assert int(tail[0]) == nph
        out["photo_reaction"] = tail[1:1 + nph].astype(np.int64)
        out["photo_reactant"] = tail[1 + nph:1 + 2 * nph].astype(np.int64)
        out["photo_Emin"] = tail[1 + 2 * nph:1 + 3 * nph]
        out["photo_Emax"] = tail[1 + 3 * nph:1 + 4 * nph]
    # rates with the fields: the workload's cells and the T grid
    fw = func_samples(name)
    fw["bins"] = w["bins"]
    nf = fw["n"].shape[0]
    fw["Phi"] = np.concatenate([w["Phi"][:64],
                                np.resize(w["Phi"], (nf - 64, B))])
    raw = run_photo(exe, "prates", fw, name + "_prates", 1.0, E)
    nph = len(meta["photo"])
    raw = raw.reshape(nf, meta["nrea"] + 34 + nph)
    out.update(fn=fw["n"], fT=fw["T"], fpar=fw["par"], fPhi=fw["Phi"],
               k=raw[:, :meta["nrea"]],
               thermal=raw[:, meta["nrea"]:meta["nrea"] + 34],
               photo_heat=raw[:, meta["nrea"] + 34:])
    opts = np.array([float(w["opts"][k]) for k in OPT_KEYS])
    for m, ex in exes.items():
        raw = run_photo(ex, "psolve", w, "%s_psolve_%s" % (name, m), 1.0, E)
        nc, ns = w["n"].shape
        o = 0
        out["n_" + m] = raw[o:o + nc * ns].reshape(nc, ns)
        o += nc * ns
        out["T_" + m] = raw[o:o + nc]
        o += nc
        out["flag_" + m] = raw[o:o + nc].astype(np.int32)
        o += nc
        out["stats_" + m] = raw[o:o + nc * 9].reshape(nc, 9).astype(np.int64)
        print("%s psolve %s: %d cells, flags %s, steps %d..%d" % (
            name, m, nc, np.unique(out["flag_" + m]),
            out["stats_" + m][:, 0].min(), out["stats_" + m][:, 0].max()))
    np.savez_compressed(HERE / ("%s_photo.npz" % name), n0=w["n"], T0=w["T"],
                        dt=w["dt"], par=w["par"], Phi=w["Phi"], opts=opts,
                        **out)


# ---- the adjoint's perturbed solves -----------------------------------------
# C4: the gradients of log(n_H2/nH) and log(Tgas) at t_end by every log k_r,
# log n_i(0) and log Tgas(0), as central differences of the reference's own
# solves.  The reference has no per-reaction scale at run time (rate_scale
# is network-wide, docs/network-format.md:712-720), so each reaction line of
# primordial3.knet gets a parameter of its own, kmulNN (NN its index among
# the reaction lines, the index of its rate), multiplying its rate
ADJ_STEPS = (1e-5, 1e-6)
ADJ_RTOL = 1e-10
ADJ_CELLS = 256


def adjoint_knet(src, dst):
    """src with every reaction line's rate times kmulNN (default 1);
    returns the number of reaction lines"""
    out, j = [], 0
    for ln in src.read_text().splitlines():
        body = ln.split("#", 1)[0]
        if "->" in body and "]" in body.split("->", 1)[1]:
            head, _, expr = body.partition("]")
            out.append("%s] (%s)*kmul%02d" % (head, expr.strip(), j))
            j += 1
        else:
            out.append(ln)
    at = next(k for k, x in enumerate(out) if x.startswith("krome_format"))
    out[at + 1:at + 1] = ["param kmul%02d = 1.0" % k for k in range(j)]
    dst.write_text("\n".join(out) + "\n")
    return j


def make_adjoint():
    name = "primordial3_kmul"
    BUILD.mkdir(parents=True, exist_ok=True)
    knet = BUILD / "primordial3_kmul.knet"
    R = adjoint_knet(REF / "networks" / "primordial3.knet", knet)
    TOML[name] = TOML["primordial3"].replace(
        '"{net}/primordial3.knet"', '"%s"' % knet)
    proj = generate(name)
    exe = compile_driver(proj, name, "bitwise")
    pn = BUILD / "kmul_params.txt"
    subprocess.run([str(exe), "params", "-", str(pn)], check=True)
    names = pn.read_text().split()
    meta = cells.network_meta("primordial3")
    assert meta["nrea"] == R
    w = cells.primordial3_cells(ADJ_CELLS, seed=3)
    nc, ns = w["n"].shape

    def rows(par, kmul):
        """parameter rows of the kmul project: the base rows' fields by name,
        the multipliers kmul [ncell, R]"""
        out = np.zeros((par.shape[0], len(names)))
        for k, nm in enumerate(names):
            out[:, k] = (kmul[:, int(nm[4:])] if nm.startswith("kmul")
                         else par[:, meta["params"].index(nm)])
        return out
    # each multiplier scales its own rate and no other (the index map)
    fw = func_samples("primordial3")
    nf = fw["n"].shape[0]
    one = np.ones((nf, R))
    k0 = run_funcs(exe, dict(fw, par=rows(fw["par"], one)), "kmul_f0", R,
                   ns + 1)[0]
    for r in range(R):
        km = one.copy()
        km[:, r] = 2.0
        k1 = run_funcs(exe, dict(fw, par=rows(fw["par"], km)), "kmul_f", R,
                       ns + 1)[0]
        expect = k0.copy()
        expect[:, r] *= 2.0
        assert np.array_equal(k1, expect), r
    # the solves: the cells unperturbed, then per step h, per parameter
    # (the R log k_r, the ns log n_i(0), log Tgas(0)) and sign
    P = R + ns + 1
    runs = [(None, 0, 0.0)] + [(h, p, sg) for h in ADJ_STEPS
                               for p in range(P) for sg in (1.0, -1.0)]
    n = np.concatenate([w["n"]] * len(runs))
    T = np.concatenate([w["T"]] * len(runs))
    km = np.ones((len(runs) * nc, R))
    for k, (h, p, sg) in enumerate(runs):
        blk = slice(k * nc, (k + 1) * nc)
        if h is None:
            continue
        if p < R:
            km[blk, p] = np.exp(sg * h)
        elif p < R + ns:
            n[blk, p - R] *= np.exp(sg * h)
        else:
            T[blk] *= np.exp(sg * h)
    big = {"n": n, "T": T, "dt": np.concatenate([w["dt"]] * len(runs)),
           "par": rows(np.concatenate([w["par"]] * len(runs)), km)}
    opts = np.array([1, ADJ_RTOL, 1e-20, 1e-6, 1, 0, 1e6])
    nf_, Tf, flag, stats, secs = run_solve(exe, big, "kmul_solve", opts)
    wH = np.array(meta["natoms"])[:, meta["elements"].index("H")].astype(
        float)
    iH2 = meta["species"].index("H2")
    L1 = (np.log(nf_[:, iH2]) - np.log(nf_ @ wH)).reshape(len(runs), nc)
    L2 = np.log(Tf).reshape(len(runs), nc)
    fl = flag.reshape(len(runs), nc)
    shape = (len(ADJ_STEPS), P, 2, nc)
    np.savez_compressed(
        HERE / "primordial3_adjoint.npz", n0=w["n"], T0=w["T"], dt=w["dt"],
        par=w["par"], steps=np.array(ADJ_STEPS), rtol=ADJ_RTOL,
        L1_0=L1[0], L2_0=L2[0], L1=L1[1:].reshape(shape),
        L2=L2[1:].reshape(shape), flags=fl[1:].reshape(shape),
        flags_0=fl[0], n_final=nf_[:nc], T_final=Tf[:nc])
    print("primordial3 adjoint: %d solves of %d cells in %.1f s, flags %s"
          % (len(runs), nc, secs, np.unique(fl)))


def make_extra(name, exe):
    """a generated network beyond the benchmark ones: the reference's
    rates and rhs at the generic cells and the T grid, and its solves of
    256 generic cells (with the photon fields for a photo network)"""
    meta = cells.network_meta(name)
    ns = len(meta["species"])
    w = cells.WORKLOADS[name](256)
    fw = func_samples(name)
    if meta["photo"]:
        exes = {"bitwise": exe}
        make_photo(name, exes, meta)
        return
    k, f, J, th = run_funcs(exe, fw, name + "_funcs", meta["nrea"], ns)
    n, T, flag, stats, secs = run_solve(exe, w, name + "_bitwise")
    np.savez_compressed(HERE / ("%s_extra.npz" % name), fn=fw["n"],
                        fT=fw["T"], fpar=fw["par"], k=k, f=f, n0=w["n"],
                        T0=w["T"], dt=w["dt"], par=w["par"],
                        opts=np.array([float(w["opts"][o])
                                       for o in OPT_KEYS]),
                        n_bitwise=n, T_bitwise=T, flag_bitwise=flag,
                        stats_bitwise=stats)
    print("%s: funcs %d, solve flags %s, steps %d..%d" % (
        name, k.shape[0], np.unique(flag), stats[:, 0].min(),
        stats[:, 0].max()))


def main(argv=None):
    """argv: the parts to write (solve funcs conserve sparse photo;
    default all) and, optionally, the networks (default all)"""
    argv = list(argv or [])
    names = set(TOML) | set(cells.EXTRA_NETWORKS)
    only = {a for a in argv if a in names}
    argv = [a for a in argv if a not in names]
    parts = set(argv if argv else ("solve", "funcs", "conserve", "sparse",
                                   "photo", "adjoint", "tol", "cloud",
                                   "jac1"))
    BUILD.mkdir(parents=True, exist_ok=True)
    if "adjoint" in parts and (not only or "primordial3" in only):
        make_adjoint()
    if "cloud" in parts:
        proj = generate("cloud")
        make_cloud({m: compile_driver(proj, "cloud", m)
                    for m in ("bitwise", "production")},
                   cells.network_meta("cloud"))
    for name in ("galaxy_ism", "galaxy_ism_heat"):
        if "photo" not in parts or (only and name not in only):
            continue
        proj = generate(name)
        exes = {m: compile_driver(proj, name, m)
                for m in ("bitwise", "production")}
        meta = cells.network_meta(name)
        make_photo(name, exes, meta)
        w = cells.WORKLOADS[name](64)
        np.savez_compressed(HERE / ("%s_tables.npz" % name),
                            **ref_tables(exes["bitwise"], name, w))
    for name in cells.EXTRA_NETWORKS:
        if "extra" not in parts or (only and name not in only):
            continue
        TOML.setdefault(name, '[project]\nnetwork = "{net}/%s.knet"\n'
                        'outdir = "{out}"\n' % name)
        proj = generate(name)
        make_extra(name, compile_driver(proj, name, "bitwise"))
    for name in ("primordial", "primordial3", "primordialZ"):
        if not parts & {"solve", "funcs", "conserve", "sparse", "tol",
                         "jac1"}:
            break
        if only and name not in only:
            continue
        meta = cells.network_meta(name)
        proj = generate(name)
        exes = {m: compile_driver(proj, name, m)
                for m in ("bitwise", "production")}
        if "conserve" in parts and name != "primordialZ":
            make_conserve(name, exes["bitwise"], meta)
        if name == "primordialZ" and parts & {"solve", "conserve"}:
            w = cells.WORKLOADS[WORKLOAD[name]](64)
            np.savez_compressed(HERE / ("%s_tables.npz" % name),
                                **ref_tables(exes["bitwise"], name, w))
        if "sparse" in parts:
            make_sparse(name, exes["bitwise"])
        if "tol" in parts and name != "primordialZ":
            make_tol(name, exes["bitwise"], meta)
        if "jac1" in parts and name == "primordial3":
            make_jac1(name, exes)
        if not parts & {"solve", "funcs"}:
            continue
        w = cells.WORKLOADS[WORKLOAD[name]](NSOLVE[name])
        if "solve" in parts:
            res = {}
            for m, exe in exes.items():
                n, T, flag, stats, secs = run_solve(exe, w,
                                                    "%s_%s" % (name, m))
                res.update({"n_" + m: n, "T_" + m: T, "flag_" + m: flag,
                            "stats_" + m: stats, "seconds_" + m: secs})
                print("%s %s: %d cells in %.2f s, flags %s" % (
                    name, m, n.shape[0], secs, np.unique(flag)))
            np.savez_compressed(HERE / ("%s_solve.npz" % name), n0=w["n"],
                                T0=w["T"], dt=w["dt"], par=w["par"],
                                opts=np.array([float(w["opts"][k])
                                               for k in OPT_KEYS]), **res)
        if "funcs" in parts:
            fw = func_samples(name)
            neq = len(meta["species"]) + (1 if w["opts"]["evolve_T"] else 0)
            k, f, J, th = run_funcs(exes["bitwise"], fw, name + "_funcs",
                                    meta["nrea"], neq)
            np.savez_compressed(HERE / ("%s_funcs.npz" % name), n=fw["n"],
                                T=fw["T"], par=fw["par"], k=k, f=f, J=J,
                                thermal=th)
            print("%s funcs: %d samples" % (name, k.shape[0]))

if __name__ == "__main__":
    main(sys.argv[1:])
