"""Write the golden fixtures of tests/golden from the reference itself.

Run in the container that has the reference (read-only) under
/root/reference:  python3 tests/golden/make_golden.py

For each workload it:
  1. runs the reference's code generator (python3 -m kromex, with the
     workload's thermal options) into _work/golden_build/<name>;
  2. compiles ref_driver.cpp against that project with g++, twice: the
     bitwise flags of the reference's test builds (-O2 -ffp-contract=off,
     codegen.py:9122-9123) and the production flags (-O3
     -ffp-contract=fast -fopenmp-simd, codegen.py:9126), both with -fopenmp;
  3. solves a seeded sample of the workload's cells with
     krome::solve_cells (PATH_PERCELL) and evaluates the rates, the full
     rhs, the analytic Jacobian and the thermal terms at sample states;
  4. writes <name>_solve.npz and <name>_funcs.npz (inputs and outputs
     only: numbers, no reference code).

The generated project and the binaries stay in _work/ (not committed).
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
BUILD = REPO / "_work" / "golden_build"

TOML = {
    "primordial": '[project]\nnetwork = "{net}/primordial.knet"\n'
                  'outdir = "{out}"\n',
    "primordial3": '[project]\nnetwork = "{net}/primordial3.knet"\n'
                   'outdir = "{out}"\n[thermal]\n'
                   'cooling = ["H2", "CONT", "CHEM"]\n'
                   'heating = ["COMPRESS", "CHEM"]\ngamma = "EXACT"\n'
                   'h2_opacity = "RIPAMONTI"\n',
}
# cells of the solve fixtures (the first ones of the seeded workload)
NSOLVE = {"primordial": 1024, "primordial3": 256}
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


def write_in(path, w):
    n, T, dt, par = w["n"], w["T"], w["dt"], w["par"]
    opts = np.array([float(w["opts"][k]) for k in OPT_KEYS])
    with open(path, "wb") as f:
        np.array([n.shape[0], n.shape[1], par.shape[1], opts.size],
                 dtype=np.int64).tofile(f)
        opts.tofile(f)
        for a in (n, T, dt, par):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def run_solve(exe, w, tag):
    fin = BUILD / ("in_%s.bin" % tag)
    fout = BUILD / ("out_%s.bin" % tag)
    write_in(fin, w)
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
    gen = cells.WORKLOADS[name]
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


def main():
    BUILD.mkdir(parents=True, exist_ok=True)
    for name in ("primordial", "primordial3"):
        meta = cells.network_meta(name)
        proj = generate(name)
        exes = {m: compile_driver(proj, name, m)
                for m in ("bitwise", "production")}
        w = cells.WORKLOADS[name](NSOLVE[name])
        res = {}
        for m, exe in exes.items():
            n, T, flag, stats, secs = run_solve(exe, w, "%s_%s" % (name, m))
            res.update({"n_" + m: n, "T_" + m: T, "flag_" + m: flag,
                        "stats_" + m: stats, "seconds_" + m: secs})
            print("%s %s: %d cells in %.2f s, flags %s" % (
                name, m, n.shape[0], secs, np.unique(flag)))
        np.savez_compressed(HERE / ("%s_solve.npz" % name), n0=w["n"],
                            T0=w["T"], dt=w["dt"], par=w["par"],
                            opts=np.array([float(w["opts"][k])
                                           for k in OPT_KEYS]), **res)
        fw = func_samples(name)
        neq = len(meta["species"]) + (1 if w["opts"]["evolve_T"] else 0)
        k, f, J, th = run_funcs(exes["bitwise"], fw, name + "_funcs",
                                meta["nrea"], neq)
        np.savez_compressed(HERE / ("%s_funcs.npz" % name), n=fw["n"],
                            T=fw["T"], par=fw["par"], k=k, f=f, J=J,
                            thermal=th)
        print("%s funcs: %d samples" % (name, k.shape[0]))


if __name__ == "__main__":
    main()
