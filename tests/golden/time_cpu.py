"""Time the oracle's CPU restatement against the reference's own CPU solve
on the same cells and cores, in the container that has the reference, and
write tests/golden/cpu_timing.json (the ratio bench.py's cpu_baseline
reports beside its own timing of the restatement on the GPU host, where
the reference does not exist).

    python3 tests/golden/time_cpu.py      (after make_golden.py solve)

The reference: krome::solve_cells (PATH_PERCELL, OpenMP over cells) of the
project the reference's generator wrote into tests/golden/_build/<name>,
compiled with its production flags (-O3 -ffp-contract=fast -fopenmp-simd,
codegen.py:9126) by make_golden.py; it times the solve call alone.  The
restatement: oracle/_build/libkx_oracle_fast.so (the same flags), timed
around its solve_cells call.  Both on every core of this container, on the
first NCELL cells of the workload; the median of three runs each.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "krome-x_amd"))
sys.path.insert(0, str(HERE))
import cells  # noqa: E402
import make_golden as mg  # noqa: E402

NCELL = {"primordial3": 16384, "primordial": 4096}
NET_ID = {"primordial": 0, "primordial3": 1}


def time_oracle(name, w, cores):
    lib = ctypes.CDLL(str(REPO / "oracle" / "_build" /
                          "libkx_oracle_fast.so"))
    lib.kxo_set_threads(cores)
    D = ctypes.POINTER(ctypes.c_double)
    o = w["opts"]
    opts = np.array([o["evolve_T"], o["rtol"], o["atol_n"], o["atol_T"],
                     o["linsol"], o["jac"], o["max_steps"], 7, 10, 5, 1, 0],
                    dtype=np.float64)
    out = []
    for _ in range(3):
        n = np.array(w["n"], order="C", copy=True)
        T = np.array(w["T"], order="C", copy=True)
        fl = np.zeros(len(T), dtype=np.intc)
        t0 = time.perf_counter()
        lib.kxo_solve_cells(NET_ID[name], len(T), n.ctypes.data_as(D),
                            T.ctypes.data_as(D),
                            np.ascontiguousarray(w["dt"]).ctypes.data_as(D),
                            np.ascontiguousarray(w["par"]).ctypes.data_as(D),
                            w["par"].shape[1], opts.ctypes.data_as(D),
                            fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            None)
        out.append(time.perf_counter() - t0)
    return float(np.median(out))


def main():
    cores = os.cpu_count() or 1
    res = {"_note": "seconds for the same cells on the same %d cores of the "
                    "container that holds the reference; median of 3"
                    % cores}
    for name, nc in NCELL.items():
        exe = mg.BUILD / ("ref_driver_%s_production" % name)
        if not exe.exists():
            raise SystemExit("%s missing: run make_golden.py solve" % exe)
        w = cells.WORKLOADS[mg.WORKLOAD[name]](nc)
        ref = float(np.median([mg.run_solve(exe, w, "time_%s" % name)[4]
                               for _ in range(3)]))
        mine = time_oracle(name, w, cores)
        res[name] = {"cells": nc, "cores": cores, "reference_seconds": ref,
                     "restatement_seconds": mine,
                     "restatement_over_reference": ref / mine}
        print(name, res[name])
    (HERE / "cpu_timing.json").write_text(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    main()
