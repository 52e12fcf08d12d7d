"""Benchmark: cells/s of the batched per-cell stiff-chemistry solve on one
MI355X, primordial3 collapse cells (2^20 cells, seed 1: nH 1..1e16 cm^-3,
Tgas 10..3e3 K, x_ion 1e-8..1e-3, H2/nH 1e-8..0.5, He 0.0775 nH, dt = 0.01
tff, Tgas evolved, cooling H2 + CONT + CHEM, heating COMPRESS + CHEM,
gamma EXACT, H2 opacity RIPAMONTI; rtol 1e-8, atol_n 1e-20, atol_T 1e-6).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ncell N]
                    [--kernel persistent|plain] [--fp64]
                    [--workload solve|conserve|multi|photo|network|copy|
                                adjoint]
                    [--linsol dense|sparse]

A step is one solve of the whole batch from its initial state: the inputs
are resident in HBM (species-major n[nspec, ncell]); each step first
restores the working arrays from a resident pristine copy (a
device-to-device copy inside the step), then runs the solve kernel.  The
headline is the persistent batched kernel with the mixed-precision Newton
(FP32 Jacobian / LU / triangular solves, FP64 state, residual and error
test); beside it, in the same run: the plain FP64 one-thread-per-cell
drop-in kernel on the same cells (ratio), the agreement of the two, and
the CPU restatement of the reference solver on a bounded sample of the
same cells on the host's cores (cpu_baseline).

--workload conserve measures the consistent-advection kernel instead
(krome_conserve_cells_device, HBM-bound): the same 2^20 cells, each
species perturbed by up to 30 % (one cell in 16 with a negative H-),
moved back to their own element totals; species-major rows, one reference
row per cell.  A step restores the rows from a resident copy and runs the
call; the roofline is the kernel's algorithmic bytes over its event-timed
duration.

--workload multi measures several networks in one persistent launch
(libkrome_multi.so): ncell/4 of the primordial3 cells above and ncell/2
primordial cells of the uneven batch (seed 2, density 1..1e12 cm^-3, dt
1e8..1e14 s, Tgas fixed, rtol 1e-6) solved by one launch, against the
same two batches solved by two launches of the networks' own libraries.
Mixed precision unless --fp64, as for the solve workload.

--workload photo measures the photochemistry network galaxy_ism with
photon bins: 2^18 cells (seed 4: nH 0.1..1e5 cm^-3, Tgas 10..1e4 K,
dt 3.15e11 s, Tgas evolved, rtol 1e-8), 32 logarithmic bins over
0.7..200 eV and a field per cell (krome-x_amd/cells.py galaxy_ism).  Each
step restores the cells and solves them through krome_solve_cells_device
with the fields on the device (the photo sums kernel, then the solve);
the plain FP64 kernel, the persistent FP64 kernel and the persistent
mixed-precision kernel are timed in turn, with the error of each per
species against the plain FP64 result and, on a sample, against the
oracle; the oracle on the host's cores is the CPU baseline.

--workload network --network NAME --cells WORKLOAD measures one generated
network on a workload of krome-x_amd/cells.py (default: c3_primordialZ,
the primordialZ cells of the uneven batch): one device-resident solve of
--ncell cells by the plain FP64 kernel with the dense LU and with the
static-pivot sparse LU, and by the persistent FP64 kernel where the
network has at most 16 equations; cells/s from the kernels' events, the
steps per cell, the nonzero flags.

--workload adjoint measures the discrete adjoint (C4): 65536 primordial3
collapse cells (seed 3, rtol 1e-8), one forward solve and the two backward
passes giving the gradients of log(n_H2/nH) and log(Tgas) at t_end by all
log k_r, log n_i(0) and log Tgas(0); cells/s of the whole job, the peak
device memory of a backward pass (workspace and checkpoints), and the
plain FP64 kernel's time on the same cells beside the forward's.

--workload copy is the calibration of the HBM traffic counters: one
device-to-device copy of 2 GiB (far above the last-level cache), whose
FETCH_SIZE and WRITE_SIZE under rocprofv3 --pmc give the factors the
kernels' counters are scaled by (profiles/traffic.json).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "krome-x_amd"))

METRIC = ("cells/sec to t_end, mixed-prec persistent vs plain FP64 GPU, "
          "@rtol 1e-4 vs FP64")
PEAK_FP64 = 78.6e12       # MI355X FP64 vector (= FP64 matrix), FLOP/s
PEAK_FP32 = 157.3e12      # MI355X FP32 vector
PEAK_HBM = 8.0e12         # bytes/s
# the device the workloads run on (one GPU)
DEVICE = "cuda:0"
# algorithmic bytes per cell: read n (9), Tgas, dt, 3 parameters; write n
# (9), Tgas (8 B each) and the flag (4 B)
BYTES_PER_CELL = 8 * (9 + 1 + 1 + 3) + 8 * (9 + 1) + 4


def flops_per_cell(stats, opc, N=10, q_avg=3.0, sparse=None):
    """algorithmic FLOPs of each cell from its counters: rhs evaluations
    (Newton residuals + DQ Jacobian columns) x the generated rhs's
    operation count, the DQ column updates, the Newton matrix setups (A =
    I - gamma J and the LU), the triangular solves with the Newton vector
    updates, and the per-step Nordsieck / error-test vector work (at the
    mean order q_avg).  sparse (the network's sparse-LU tables): the
    factorization's own flop count plus the border's blocks, Schur
    complement and its dense LU, and the sparse triangular solves"""
    f_rhs = opc["full_rhs"]
    f_getrf = N * (N - 1) / 2 + (N - 1) * N * (2 * N - 1) / 3 + N
    f_getrs = 2 * N * N - N
    if sparse is not None:
        ns = len(sparse["perm"])
        m = N - ns
        nnz = len(sparse["ci"])
        fm = m * (m - 1) / 2 + (m - 1) * m * (2 * m - 1) / 3 + m
        f_getrf = (sparse["flops"] + 2 * m * nnz + 2 * m * m * ns + fm
                   - (N * N + N))  # the dense A = I - gamma J is not formed
        f_getrf += nnz + ns
        f_getrs = 2 * nnz - ns + 4 * m * ns + 2 * m * m - m
    step = (q_avg * (q_avg + 1) / 2 + 2 * (q_avg + 1) + 12) * N
    st = stats.astype(np.float64)
    nst, nfe, nje, nfeDQ, nset, nni = (st[:, 0], st[:, 1], st[:, 2],
                                       st[:, 3], st[:, 4], st[:, 5])
    fp_rhs64 = nfe * f_rhs
    fp_dq = nfeDQ * f_rhs + nje * N * (2 * N + 4)
    fp_lu = nset * (N * N + N + f_getrf) + nni * f_getrs
    fp_vec = nni * 14 * N + nst * step
    return fp_rhs64, fp_dq, fp_lu, fp_vec


def cpu_baseline(w, nsample):
    """the oracle (CPU restatement of the reference solver, OpenMP over
    cells), built with the reference's production flags (oracle/Makefile
    target fast), on the first nsample cells of the workload"""
    lib = ctypes.CDLL(str(REPO / "oracle" / "_build" / "libkx_oracle_fast.so"))
    D = ctypes.POINTER(ctypes.c_double)
    cores = min(16, os.cpu_count() or 1)
    lib.kxo_set_threads(cores)
    cores = lib.kxo_get_threads()
    n = np.ascontiguousarray(w["n"][:nsample])
    T = np.ascontiguousarray(w["T"][:nsample])
    dt = np.ascontiguousarray(w["dt"][:nsample])
    par = np.ascontiguousarray(w["par"][:nsample])
    o = w["opts"]
    opts = np.array([o["evolve_T"], o["rtol"], o["atol_n"], o["atol_T"],
                     1, 0, o["max_steps"], 7, 10, 5, 1, 0], dtype=np.float64)
    flags = np.zeros(nsample, dtype=np.intc)
    t0 = time.perf_counter()
    r = lib.kxo_solve_cells(1, nsample, n.ctypes.data_as(D),
                            T.ctypes.data_as(D), dt.ctypes.data_as(D),
                            par.ctypes.data_as(D), par.shape[1],
                            opts.ctypes.data_as(D),
                            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            None)
    el = time.perf_counter() - t0
    ratio = None
    tf = REPO / "tests" / "golden" / "cpu_timing.json"
    if tf.exists():
        ratio = json.loads(tf.read_text()).get("primordial3", {}).get(
            "restatement_over_reference")
    return {"value": nsample / el, "unit": "cells/s", "cores": int(cores),
            "kind": "port",
            "restatement_over_reference_in_container": ratio,
            "sample": "the first %d of the same cells, CPU restatement of "
                      "the reference per-cell solve (OpenMP dynamic, the "
                      "reference's production flags -O3 -ffp-contract=fast), "
                      "%.1f s wall, return %d; restatement_over_reference"
                      "_in_container: the restatement's cells/s over the "
                      "reference's own per-cell solve on the same cells "
                      "and cores, measured where the reference builds "
                      "(tests/golden/time_cpu.py)" % (nsample, el, r)}, \
        n, T, flags


def conserve_bench(a):
    """cells/s of the consistent-advection kernel (one JSON line)"""
    import torch
    import cells
    import krome

    dev = torch.device("cuda:0")
    net = krome.load("primordial3")
    w = cells.primordial3_cells(a.ncell)
    ref = net.element_totals(w["n"])
    rng = np.random.default_rng(7)
    v0 = w["n"] * (1.0 + 0.3 * (2.0 * rng.random(w["n"].shape) - 1.0))
    neg = np.arange(a.ncell) % 16 == 3
    v0[neg, 2] = -1e-3 * v0[neg, 2]
    vt0 = torch.tensor(np.ascontiguousarray(v0.T), device=dev)
    vt = vt0.clone()
    rt = torch.tensor(ref, device=dev)
    fl = torch.zeros(a.ncell, dtype=torch.int32, device=dev)
    kms = ctypes.c_double()
    kp = ctypes.cast(ctypes.byref(kms), ctypes.POINTER(ctypes.c_double))
    for _ in range(max(a.warmup, 1)):
        vt.copy_(vt0)
        net.conserve_device(vt, rt, flags=fl)
    torch.cuda.synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        vt.copy_(vt0)
        r = net.conserve_device(vt, rt, flags=fl)
        if net._lib.krome_device_last_kernel_ms(kp) != 0:
            raise SystemExit("no kernel time")
        kern.append(kms.value)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean(kern))
    ns, ne = net.nspec, len(net.elements)
    # read the row and the reference row, write the row and the flag
    bpc = 8 * (2 * ns + ne) + 4
    achieved = bpc * a.ncell / (kern_ms * 1e-3)
    flg = fl.cpu().numpy()
    traffic = None
    tf = REPO / "profiles" / "traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("conserve")
    cpu = None
    if not a.no_cpu:
        lib = ctypes.CDLL(str(REPO / "oracle" / "_build" /
                              "libkx_oracle_fast.so"))
        lib.kxo_set_threads(min(16, os.cpu_count() or 1))
        cores = lib.kxo_get_threads()
        m = min(a.cpu_sample, a.ncell)
        vc = np.ascontiguousarray(v0[:m])
        rc = np.ascontiguousarray(ref[:m])
        fc = np.zeros(m, dtype=np.intc)
        D = ctypes.POINTER(ctypes.c_double)
        t1 = time.perf_counter()
        lib.kxo_conserve(1, 0, m, vc.ctypes.data_as(D), rc.ctypes.data_as(D),
                         1, fc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        el = time.perf_counter() - t1
        cpu = {"value": m / el, "unit": "cells/s", "cores": int(cores),
               "kind": "port",
               "sample": "the first %d of the same cells, CPU restatement "
                         "of the reference's conserve_cells (OpenMP static, "
                         "the reference's production flags), %.2f s wall"
                         % (m, el)}
    out = {
        "metric": "cells/sec, consistent advection (krome_conserve_cells_"
                  "device)",
        "value": a.ncell * a.steps / wall, "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "primordial3 collapse cells (seed 1), each "
                               "species x (1 +- 30 %), back to their own "
                               "element totals", "ncell": a.ncell,
                   "layout": "species-major", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9,
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "bytes_per_cell": bpc, "kernel_ms": kern_ms},
        "cpu_baseline": cpu,
        "flags": {str(k): int(v) for k, v in zip(*np.unique(
            flg, return_counts=True))},
        "return": int(r),
        "ratio_vs_cpu": (a.ncell / (kern_ms * 1e-3)) / cpu["value"]
        if cpu else None,
    }
    print(json.dumps(out))


def multi_bench(a):
    """one launch of two networks against two launches (one JSON line)"""
    import torch
    import cells
    import krome

    dev = torch.device("cuda:0")
    # mixed precision unless --fp64 (it applies to primordial3's difference-
    # quotient Jacobian; primordial's analytic one keeps the FP64 matrix)
    mixed = 0 if a.fp64 else 1
    m = krome.load_multi()
    nets = {n: krome.load(n) for n in ("primordial", "primordial3")}
    n3, n1 = a.ncell // 4, a.ncell // 2
    ws = [("primordial3", cells.primordial3_cells(n3), 1, 1e-8),
          ("primordial", cells.c3_primordial_cells(n1), 0, 1e-6)]
    batches, pristine = [], []
    for name, w, et, rtol in ws:
        b = dict(network=name, evolve_T=et, rtol=rtol, atol_n=1e-20,
                 atol_T=1e-6,
                 n=torch.tensor(np.ascontiguousarray(w["n"].T), device=dev),
                 Tgas=torch.tensor(w["T"], device=dev),
                 dt=torch.tensor(w["dt"], device=dev),
                 par=torch.tensor(np.ascontiguousarray(w["par"]), device=dev),
                 flags=torch.zeros(w["n"].shape[0], dtype=torch.int32,
                                   device=dev))
        batches.append(b)
        pristine.append((b["n"].clone(), b["Tgas"].clone()))

    def restore():
        for b, (n0, T0) in zip(batches, pristine):
            b["n"].copy_(n0)
            b["Tgas"].copy_(T0)

    kms = ctypes.c_double()
    kp = ctypes.cast(ctypes.byref(kms), ctypes.POINTER(ctypes.c_double))

    def separate():
        total = 0.0
        for b in batches:
            net = nets[b["network"]]
            o = net.options(evolve_T=bool(b["evolve_T"]), rtol=b["rtol"],
                            atol_n=b["atol_n"], atol_T=b["atol_T"],
                            device_kernel="persistent",
                            mixed_precision=mixed, error_report=False)
            r = net.solve_device(b["n"], b["Tgas"], b["dt"], b["par"],
                                 opts=o, flags=b["flags"])
            if r != 0 or net._lib.krome_device_last_kernel_ms(kp) != 0:
                raise SystemExit("separate solve returned %d" % r)
            total += kms.value
        return total

    for _ in range(max(a.warmup, 1)):
        restore()
        m.solve_device(batches, mixed)
    pair_ms, sep_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        restore()
        r, ms = m.solve_device(batches, mixed)
        if r != 0:
            raise SystemExit("pair launch returned %d" % r)
        pair_ms.append(ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pair_out = [b["n"].cpu().numpy() for b in batches]
    for _ in range(a.steps):
        restore()
        sep_ms.append(separate())
    same = all(np.array_equal(p, b["n"].cpu().numpy())
               for p, b in zip(pair_out, batches))
    ncell = n1 + n3
    pm, sm = float(np.mean(pair_ms)), float(np.mean(sep_ms))
    out = {
        "metric": "cells/sec, two networks in one persistent launch",
        "value": ncell * a.steps / wall, "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "2^18 primordial3 collapse cells (seed 1) + "
                               "2^19 primordial cells of the uneven batch "
                               "(seed 2, density 1..1e12, dt 1e8..1e14 s)",
                   "ncell": ncell, "mixed_precision": bool(mixed),
                   "parallelism": "1 GPU"},
        "pair_kernel_ms": pm, "separate_kernel_ms": sm,
        "speedup_vs_separate_launches": sm / pm,
        "results_bitwise_equal_to_separate": bool(same),
    }
    print(json.dumps(out))


def c3_bench(a):
    """the uneven batch: 2^19 primordial + 2^18 primordialZ + 2^17 cloud
    cells (scaled by --ncell / 2^20) in one persistent launch of
    libkrome_multi_c3 (one JSON line)"""
    import torch
    import cells
    import krome

    dev = torch.device(DEVICE)
    m = krome.load_multi("multi_c3")
    m.set_launch(a.waves_per_cu, a.queue_chunk)
    scale = a.ncell / float(1 << 20)
    parts = (("primordial", cells.c3_primordial_cells, 1 << 19),
             ("primordialZ", cells.c3_primordialZ_cells, 1 << 18),
             ("cloud", cells.c3_cloud_cells, 1 << 17))
    if a.c3_parts:
        keep = a.c3_parts.split(",")
        parts = tuple(q for q in parts if q[0] in keep)
    batches, pristine = [], []
    for name, gen, full in parts:
        nc = max(64, int(full * scale))
        w = gen(nc)
        o = w["opts"]
        b = dict(network=name, evolve_T=int(o["evolve_T"]), rtol=o["rtol"],
                 atol_n=o["atol_n"], atol_T=o["atol_T"],
                 max_steps=int(o["max_steps"]), jac=int(o["jac"]),
                 n=torch.tensor(np.ascontiguousarray(w["n"].T), device=dev),
                 Tgas=torch.tensor(w["T"], device=dev),
                 dt=torch.tensor(w["dt"], device=dev),
                 par=torch.tensor(np.ascontiguousarray(w["par"]), device=dev),
                 flags=torch.zeros(nc, dtype=torch.int32, device=dev),
                 stats=torch.zeros((nc, 9), dtype=torch.int64, device=dev))
        batches.append(b)
        pristine.append((b["n"].clone(), b["Tgas"].clone()))

    def restore():
        for b, (n0, T0) in zip(batches, pristine):
            b["n"].copy_(n0)
            b["Tgas"].copy_(T0)

    # --c3-slices K: the batch in K launches, launch k over the k-th
    # contiguous slice of every network's cells (views of the same
    # tensors, copied in and out around each launch)
    K = max(1, a.c3_slices)
    k_lo, k_hi = 0, K
    if a.c3_slice_range:
        k_lo, k_hi = (int(x) for x in a.c3_slice_range.split(":"))
        if not 0 <= k_lo < k_hi <= K:
            raise SystemExit("--c3-slice-range %s outside 0:%d"
                             % (a.c3_slice_range, K))

    def covered(nc):
        """the cells of a network of nc cells that the launches cover"""
        return nc * k_lo // K, nc * k_hi // K

    def launches():
        if K == 1:
            yield batches
            return
        for k in range(k_lo, k_hi):
            sub = []
            for b in batches:
                nc = b["Tgas"].shape[0]
                lo, hi = nc * k // K, nc * (k + 1) // K
                d = dict(b)
                for f in ("Tgas", "dt", "par", "flags", "stats"):
                    d[f] = b[f][lo:hi].contiguous()
                d["n"] = b["n"][:, lo:hi].contiguous()
                d["_span"] = (lo, hi)
                sub.append(d)
            yield sub

    def solve(mixed):
        """one pass over the whole batch: (return, kernel ms)"""
        tot, ret = 0.0, 0
        for k, sub in enumerate(launches()):
            r, ms = m.solve_device(sub, mixed)
            tot += ms
            ret = ret or r
            if K > 1:
                for d, b in zip(sub, batches):
                    lo, hi = d["_span"]
                    for f in ("Tgas", "flags", "stats"):
                        b[f][lo:hi].copy_(d[f])
                    b["n"][:, lo:hi].copy_(d["n"])
                print("c3 slice %d/%d mixed=%d: %.1f ms, return %d"
                      % (k + 1, K, mixed, ms, r), file=sys.stderr,
                      flush=True)
        return ret, tot

    res = {}
    for mixed in ((0,) if a.fp64 else (0, 1)):
        for _ in range(a.warmup):
            restore()
            solve(mixed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms = []
        for _ in range(a.steps):
            restore()
            r, ms = solve(mixed)
            print("c3 launch mixed=%d: %.1f ms, return %d" % (mixed, ms, r),
                  file=sys.stderr, flush=True)
            kms.append(ms)
        torch.cuda.synchronize()
        res[mixed] = (time.perf_counter() - t0, float(np.mean(kms)))
    ncell, per, saved = 0, {}, {}
    for b in batches:
        lo, hi = covered(b["Tgas"].shape[0])
        ncell += hi - lo
        st = b["stats"].cpu().numpy()[lo:hi, 0]
        fl = b["flags"].cpu().numpy()[lo:hi]
        saved[b["network"] + "_steps"] = st
        saved[b["network"] + "_flags"] = fl
        per[b["network"]] = {
            "cells": int(st.size), "flags_nonzero": int((fl != 0).sum()),
            "steps_p50": float(np.percentile(st, 50)),
            "steps_p99": float(np.percentile(st, 99)),
            "steps_max": int(st.max())}
    allst = np.concatenate([saved[b["network"] + "_steps"]
                            for b in batches])
    if a.c3_save:
        np.savez(a.c3_save, kernel_ms=np.array([res[0][1]]), **saved)
    if 1 not in res:
        res[1] = res[0]
    wall = res[1][0]
    out = {
        "metric": "cells/sec to t_end, three networks in one persistent "
                  "launch (uneven batch)",
        "value": ncell * a.steps / wall, "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uneven batch: primordial, primordialZ and "
                               "cloud cells (seed 2, density 1..1e12, dt "
                               "1e8..1e14 s, Tgas fixed) in one launch",
                   "ncell": ncell, "parallelism": "1 GPU",
                   "launches_per_step": k_hi - k_lo,
                   "slices": "%d:%d of %d" % (k_lo, k_hi, K),
                   "waves_per_cu": a.waves_per_cu,
                   "queue_chunk": a.queue_chunk},
        "fp64_cells_per_s": ncell * a.steps / res[0][0],
        "fp64_kernel_ms": res[0][1],
        "mixed_cells_per_s": ncell * a.steps / res[1][0],
        "mixed_kernel_ms": res[1][1],
        "mixed_note": "every network here runs at fixed Tgas, where the "
                      "Newton matrix stays FP64: the two runs are the same "
                      "arithmetic",
        "steps_per_cell": {"p50": float(np.percentile(allst, 50)),
                           "p99": float(np.percentile(allst, 99)),
                           "max": int(allst.max())},
        "networks": per,
    }
    print(json.dumps(out))


def photo_bench(a):
    """galaxy_ism (or galaxy_ism_heat: the same network with heating CHEM
    + PHOTO and cooling H2 + CHEM, so Tgas moves) with photon bins: three
    kernels, one JSON line"""
    import torch
    import cells
    import krome
    sys.path.insert(0, str(REPO / "tests"))

    dev = torch.device("cuda:0")
    name = a.photo_network
    net = krome.load(name)
    nc = a.ncell
    w = cells.galaxy_ism_cells(nc, network=name)
    elo, ehi = w["bins"]
    n0 = torch.tensor(np.ascontiguousarray(w["n"].T), device=dev)
    T0 = torch.tensor(w["T"], device=dev)
    dt = torch.tensor(w["dt"], device=dev)
    par = torch.tensor(np.ascontiguousarray(w["par"]), device=dev)
    Phi = torch.tensor(np.ascontiguousarray(w["Phi"]), device=dev)
    n, T = n0.clone(), T0.clone()
    fl = torch.zeros(nc, dtype=torch.int32, device=dev)
    kms = ctypes.c_double()
    kp = ctypes.cast(ctypes.byref(kms), ctypes.POINTER(ctypes.c_double))
    runs = {}
    for key, kernel, mixed in (("plain_fp64", "plain", 0),
                               ("persistent_fp64", "persistent", 0),
                               ("persistent_mixed", "persistent", 1)):
        o = net.options(evolve_T=1, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                        linsol=1, device_kernel=kernel,
                        mixed_precision=mixed,
                        error_report=False).set_photo(elo, ehi)
        for _ in range(a.warmup):
            n.copy_(n0)
            T.copy_(T0)
            net.solve_device(n, T, dt, par, opts=o, flags=fl, Phi=Phi)
        ms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            n.copy_(n0)
            T.copy_(T0)
            r = net.solve_device(n, T, dt, par, opts=o, flags=fl, Phi=Phi)
            # a cell's integrator flag (-1 .. -99: the cell keeps its
            # input) is a result; a call-level error is not
            if r < -99 or r > 0 or \
                    net._lib.krome_device_last_kernel_ms(kp) != 0:
                raise SystemExit("%s solve returned %d" % (key, r))
            ms.append(kms.value)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        runs[key] = {"cells_per_s": nc * a.steps / wall,
                     "ms_per_step": 1e3 * wall / a.steps,
                     "kernel_ms": float(np.mean(ms)),
                     "n": n.cpu().numpy().T.copy(),
                     "T": T.cpu().numpy().copy(),
                     "flags": fl.cpu().numpy().copy()}
        print("photo %s: %.1f ms/step" % (key, runs[key]["ms_per_step"]),
              file=sys.stderr, flush=True)
    ref = runs["plain_fp64"]["n"]
    # the comparison over the cells every kernel finished
    done = np.logical_and.reduce([r["flags"] == 0 for r in runs.values()])
    big = (ref > 1e-12 * ref.sum(axis=1, keepdims=True)) & done[:, None]
    species = net.species

    def errs(x):
        d = np.where(big, np.abs(x - ref) / np.maximum(np.abs(ref), 1e-300),
                     0.0)
        return {sp: float(d[:, i].max()) for i, sp in enumerate(species)}
    out_runs = {}
    for key, r in runs.items():
        out_runs[key] = {k: r[k] for k in ("cells_per_s", "ms_per_step",
                                           "kernel_ms")}
        out_runs[key]["flags_equal_to_plain"] = bool(
            np.array_equal(r["flags"], runs["plain_fp64"]["flags"]))
        if key != "plain_fp64":
            out_runs[key]["max_rel_err_vs_plain_fp64"] = errs(r["n"])
            Tp = runs["plain_fp64"]["T"]
            out_runs[key]["max_rel_err_Tgas_vs_plain_fp64"] = float(
                np.max(np.abs(r["T"] - Tp)[done] / Tp[done]))
        out_runs[key]["flags_nonzero"] = int((r["flags"] != 0).sum())
        out_runs[key]["Tgas_moved_max_rel"] = float(
            np.max(np.abs(r["T"] - w["T"]) / w["T"]))
    # the oracle on a sample: the CPU baseline and the FP64 check
    from conftest import Oracle, oracle_opts
    orc = Oracle(REPO / "oracle" / "_build" / "libkx_oracle_fast.so", "kxo_")
    cores = min(16, os.cpu_count() or 1)
    orc.lib.kxo_set_threads(cores)
    m = min(a.cpu_sample, nc)
    sig = orc.photo_tables(name, elo, ehi, 1.0)
    rows = orc.photo_rows(name, w["par"][:m], w["Phi"][:m], sig)
    t0 = time.perf_counter()
    ro, no, To, flo, _ = orc.solve(name, w["n"][:m], w["T"][:m],
                                   w["dt"][:m], rows,
                                   oracle_opts(**w["opts"]))
    el = time.perf_counter() - t0
    bo = no > 1e-12 * no.sum(axis=1, keepdims=True)
    for key, r in runs.items():
        # the cells both finished
        ok = (flo == 0) & (r["flags"][:m] == 0)
        d = np.abs(r["n"][:m] - no) / np.maximum(np.abs(no), 1e-300)
        out_runs[key]["max_rel_err_vs_oracle_sample"] = float(
            d[bo & ok[:, None]].max()) if ok.any() else None
        out_runs[key]["max_rel_err_Tgas_vs_oracle_sample"] = float(
            np.max(np.abs(r["T"][:m] - To)[ok] / To[ok])) if ok.any() \
            else None
        out_runs[key]["flags_equal_to_oracle_sample"] = bool(
            np.array_equal(r["flags"][:m], flo))
    head = runs["persistent_mixed"]
    out = {
        "metric": "cells/sec to t_end, %s with photon bins "
                  "(persistent mixed precision)" % name,
        "value": head["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": "%d %s cells (seed 4), 32 photon "
                               "bins 0.7..200 eV, a field per cell, Tgas "
                               "evolved, rtol 1e-8" % (nc, name),
                   "ncell": nc, "parallelism": "1 GPU"},
        "kernels": out_runs,
        "mixed_vs_plain_fp64": runs["plain_fp64"]["ms_per_step"]
        / head["ms_per_step"],
        "cpu_baseline": {"value": m / el, "unit": "cells/s",
                         "cores": int(orc.lib.kxo_get_threads()),
                         "kind": "port",
                         "sample": "the first %d cells, the oracle with the "
                                   "reference's production flags, %.1f s "
                                   "wall, return %d" % (m, el, ro)},
    }
    print(json.dumps(out))


def network_bench(a):
    """one generated network's kernels on one workload (one JSON line)"""
    import torch
    import cells
    import krome

    net = krome.load(a.network)
    w = cells.WORKLOADS[a.cells](a.ncell)
    o = w["opts"]
    dev = torch.device("cuda:0")
    n0 = torch.tensor(np.ascontiguousarray(w["n"].T), device=dev)
    T0 = torch.tensor(w["T"], device=dev)
    dt = torch.tensor(w["dt"], device=dev)
    par = torch.tensor(np.ascontiguousarray(w["par"]), device=dev)
    kms = ctypes.c_double()
    kp = ctypes.cast(ctypes.byref(kms), ctypes.POINTER(ctypes.c_double))
    # run names: kernel (plain, lanes, persistent) _ LU (dense, sparse)
    names = a.kernels.split(",") if a.kernels else (
        ["plain_dense", "plain_sparse"]
        + (["persistent_dense"] if net.nspec + int(o["evolve_T"]) <= 16
           else []))
    runs = []
    for key in names:
        kernel, lu = key.split("_")
        runs.append((key, kernel, {"dense": 1, "sparse": 3}[lu]))
    out = {}
    for key, kernel, ls in runs:
        extra = {}
        if a.waves_per_cu:
            extra["waves_per_cu"] = a.waves_per_cu
        if a.cells_per_workgroup:
            extra["cells_per_workgroup"] = a.cells_per_workgroup
        opts = net.options(evolve_T=bool(o["evolve_T"]), rtol=o["rtol"],
                           atol_n=o["atol_n"], atol_T=o["atol_T"],
                           jac=int(o["jac"]), linsol=ls,
                           max_steps=int(o["max_steps"]),
                           device_kernel=kernel, error_report=False,
                           **extra)
        n, T = n0.clone(), T0.clone()
        fl = torch.zeros(a.ncell, dtype=torch.int32, device=dev)
        st = torch.zeros((a.ncell, 9), dtype=torch.int64, device=dev)
        r = net.solve_device(n, T, dt, par, opts=opts, flags=fl, stats=st)
        if net._lib.krome_device_last_kernel_ms(kp) != 0:
            raise SystemExit("%s: no kernel time" % key)
        s_ = st.cpu().numpy()[:, 0]
        out[key] = {"return": r, "kernel_ms": kms.value,
                    "cells_per_s": a.ncell / (kms.value * 1e-3),
                    "flags_nonzero": int((fl.cpu().numpy() != 0).sum()),
                    "steps_per_cell": {"p50": float(np.percentile(s_, 50)),
                                       "p99": float(np.percentile(s_, 99)),
                                       "max": int(s_.max())}}
        print("%s %s: %.1f ms" % (a.network, key, kms.value),
              file=sys.stderr, flush=True)
    best = max(out.values(), key=lambda v: v["cells_per_s"])
    print(json.dumps({
        "metric": "cells/sec to t_end, %s" % a.network,
        "value": best["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d cells of %s" % (a.ncell, a.cells),
                   "network": a.network, "parallelism": "1 GPU"},
        "kernels": out}))


def adjoint_bench(a):
    """C4: the discrete adjoint on primordial3 collapse cells (seed 3, the
    C2 physics, FP64): one forward solve and the two backward passes that
    give the gradients of log(n_H2/nH) and log(Tgas) at t_end by all log
    k_r, log n_i(0) and log Tgas(0) (each backward integrates again with a
    checkpoint every 32 steps and runs the segments backwards)"""
    import torch
    import cells
    import krome

    net = krome.load("primordial3")
    w = cells.primordial3_cells(a.ncell, seed=3)
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64,
                               device=dev)
    n0, T0, dt, par = t(w["n"].T), t(w["T"]), t(w["dt"]), t(w["par"])
    opts = net.options(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                       error_report=False)
    wH = np.array(net.composition[:, net.elements.index("H")], dtype=float)
    iH2 = net.idx("H2")
    res = {}
    for rep in range(a.warmup + 1):
        n, T = n0.clone(), T0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, _, _ = net.adjoint_device(n, T, dt, par, opts)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        nf, Tf = n.cpu().numpy().T, T.cpu().numpy()
        nH = nf @ wH
        g1 = -wH[None, :] / nH[:, None]
        g1[:, iH2] += 1.0 / nf[:, iH2]
        t_bwd, mem, flags = [], 0.0, 0
        for gn, gT in ((g1, np.zeros(a.ncell)),
                       (np.zeros_like(nf), 1.0 / Tf)):
            n, T = n0.clone(), T0.clone()
            fl = torch.zeros(a.ncell, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rb, grads, m = net.adjoint_device(n, T, dt, par, opts, flags=fl,
                                              gn=t(gn.T), gT=t(gT))
            torch.cuda.synchronize()
            t_bwd.append(time.perf_counter() - t0)
            mem = max(mem, m)
            flags += int((fl.cpu().numpy() != 0).sum())
        res = {"forward_s": t_fwd, "backward_s": t_bwd, "peak_bytes": mem,
               "flags_nonzero": flags, "return": r}
        print("adjoint rep %d: forward %.2f s, backward %s s, %.1f MB" % (
            rep, t_fwd, ["%.2f" % x for x in t_bwd], mem / 1e6),
            file=sys.stderr, flush=True)
    # the plain FP64 kernel on the same cells (the forward's yardstick)
    n, T = n0.clone(), T0.clone()
    po = net.options(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                     device_kernel="plain", error_report=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.solve_device(n, T, dt, par, opts=po)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    total = res["forward_s"] + sum(res["backward_s"])
    print(json.dumps({
        "metric": "cells/sec, gradients of log(n_H2/nH) and log(Tgas) at "
                  "t_end by all log k_r, log n_i(0), log Tgas(0) "
                  "(forward + 2 adjoint passes)",
        "value": a.ncell / total, "unit": "cells/s", "n_gpus": 1,
        "steps": 1, "warmup": a.warmup, "ms_per_step": total * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d primordial3 collapse cells (seed 3), "
                   "rtol 1e-8, checkpoint every 32 steps" % a.ncell,
                   "parallelism": "1 GPU"},
        "forward_s": res["forward_s"], "backward_s": res["backward_s"],
        "forward_cells_per_s": a.ncell / res["forward_s"],
        "plain_fp64_forward_s": t_plain,
        "peak_device_bytes": res["peak_bytes"],
        "peak_bytes_per_cell": res["peak_bytes"] / min(a.ncell, 65536),
        "flags_nonzero": res["flags_nonzero"]}))


def copy_calibration():
    """one 2 GiB device-to-device copy (the traffic counters' reference)"""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"workload": "copy", "bytes_each_way": n * 8,
                      "seconds": el, "GB_per_s": 2 * n * 8 / el / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncell", type=int, default=1 << 20)
    ap.add_argument("--kernel", default="persistent",
                    choices=("persistent", "plain"))
    ap.add_argument("--fp64", action="store_true",
                    help="the persistent kernel with the FP64 Newton")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64-persistent", action="store_true",
                    help="skip the persistent FP64 kernel's leg")
    ap.add_argument("--workload", default="solve",
                    choices=("solve", "conserve", "multi", "photo", "network",
                             "copy", "adjoint", "c3"))
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--photo-network", default="galaxy_ism",
                    choices=("galaxy_ism", "galaxy_ism_heat"))
    ap.add_argument("--queue-chunk", type=int, default=0)
    ap.add_argument("--network", default="primordialZ")
    ap.add_argument("--cells", default="c3_primordialZ")
    ap.add_argument("--kernels", default="",
                    help="network workload: runs as kernel_lu, comma "
                         "separated (kernel plain, lanes or persistent; lu "
                         "dense or sparse)")
    ap.add_argument("--cells-per-workgroup", type=int, default=0)
    ap.add_argument("--c3-parts", default="",
                    help="c3 workload: the networks to include (comma "
                         "separated; default all three)")
    ap.add_argument("--c3-slices", type=int, default=1,
                    help="c3 workload: launches per step, each over a "
                         "contiguous slice of every network's cells "
                         "(default 1: the whole batch in one launch)")
    ap.add_argument("--c3-slice-range", default="",
                    help="c3 workload: LO:HI, run only slices LO..HI-1 of "
                         "--c3-slices (the value then counts their cells)")
    ap.add_argument("--c3-save", default="",
                    help="c3 workload: write the covered cells' steps and "
                         "flags per network to this .npz")
    ap.add_argument("--linsol", default="dense", choices=("dense", "sparse"),
                    help="the Newton matrix's LU (sparse: the plain kernel's "
                         "static-pivot sparse LU, linsol 3)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the cells are independent and "
              "this build has no multi-GPU path (--gpus %d refused)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "conserve":
        return conserve_bench(a)
    if a.workload == "copy":
        return copy_calibration()
    if a.workload == "network":
        if a.ncell == 1 << 20:
            a.ncell = 1 << 16
        return network_bench(a)
    if a.workload == "photo":
        if a.ncell == 1 << 20:
            a.ncell = 1 << 18
        if a.cpu_sample == 1 << 18:
            a.cpu_sample = a.ncell
        return photo_bench(a)
    if a.workload == "multi":
        return multi_bench(a)
    if a.workload == "c3":
        return c3_bench(a)
    if a.workload == "adjoint":
        if a.ncell == 1 << 20:
            a.ncell = 1 << 16
        return adjoint_bench(a)
    import torch
    import cells
    import krome

    dev = torch.device("cuda:0")
    net = krome.load("primordial3")
    w = cells.primordial3_cells(a.ncell)
    opc = json.loads((REPO / "krome-x_amd" / "csrc" / "generated" /
                      "opcount.json").read_text())["primordial3"]
    n0 = torch.tensor(np.ascontiguousarray(w["n"].T), device=dev)
    T0 = torch.tensor(w["T"], device=dev)
    dt = torch.tensor(w["dt"], device=dev)
    par = torch.tensor(np.ascontiguousarray(w["par"]), device=dev)
    n = n0.clone()
    T = T0.clone()
    flags = torch.zeros(a.ncell, dtype=torch.int32, device=dev)
    stats = torch.zeros((a.ncell, 9), dtype=torch.int64, device=dev)
    base = dict(evolve_T=True, rtol=w["opts"]["rtol"],
                atol_n=w["opts"]["atol_n"], atol_T=w["opts"]["atol_T"],
                error_report=False)
    mixed = a.kernel == "persistent" and not a.fp64
    head = net.options(**base, device_kernel=a.kernel,
                       mixed_precision=1 if mixed else 0,
                       linsol=3 if a.linsol == "sparse" else 1)
    plain = net.options(**base, device_kernel="plain")
    pers64 = net.options(**base, device_kernel="persistent",
                         mixed_precision=0, linsol=1)

    def run(opts, st=None):
        n.copy_(n0)
        T.copy_(T0)
        return net.solve_device(n, T, dt, par, opts=opts, flags=flags,
                                stats=st, sync=False)

    # warmup (the first also collects the counters for the FLOP count)
    for i in range(a.warmup):
        r = run(head, stats if i == 0 else None)
        torch.cuda.synchronize()
        if r != 0:
            raise SystemExit("warmup solve returned %d" % r)
    if a.warmup == 0:
        run(head, stats)
        torch.cuda.synchronize()
    st_head = stats.cpu().numpy()
    fl_head = flags.cpu().numpy()
    n_head = n.cpu().numpy().copy()
    T_head = T.cpu().numpy().copy()
    # timed steps; the solve kernels' own durations from the HIP events the
    # library records around their launches, on their stream
    kms = ctypes.c_double()
    kern = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        n.copy_(n0)
        T.copy_(T0)
        r = net.solve_device(n, T, dt, par, opts=head, flags=flags,
                             sync=False)
        if r != 0 or net._lib.krome_device_last_kernel_ms(
                ctypes.cast(ctypes.byref(kms),
                            ctypes.POINTER(ctypes.c_double))) != 0:
            raise SystemExit("timed solve returned %d" % r)
        kern.append(kms.value)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean(kern))
    ms_per_step = 1e3 * wall / a.steps
    value = a.ncell * a.steps / wall

    def agreement(n_ref, T_ref, fl_ref, nh, Th, flh):
        """the headline's result nh [nspec, m] against a reference result
        on the same cells: max relative error per species (over the
        values above 1e-12 of the cell total) and of Tgas, over the cells
        both finish; flags equal"""
        ok = (fl_ref == 0) & (flh == 0)
        ntot = np.abs(n_ref).sum(axis=0)
        big = np.abs(n_ref) > 1e-12 * ntot[None, :]
        rel = np.abs(nh - n_ref) / np.maximum(np.abs(n_ref), 1e-300)
        rel = np.where(big & ok[None, :], rel, 0.0)
        Trel = np.where(ok, np.abs(Th - T_ref) / T_ref, 0.0)
        return {"max_rel_species": float(rel.max()),
                "max_rel_by_species": {sp: float(rel[i].max())
                                       for i, sp in enumerate(net.species)},
                "max_rel_Tgas": float(Trel.max()),
                "flags_equal": bool(np.array_equal(fl_ref, flh))}

    def one_solve(opts):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run(opts)  # per-cell flags are compared, not fatal
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        kms1 = ctypes.c_double()
        net._lib.krome_device_last_kernel_ms(
            ctypes.cast(ctypes.byref(kms1), ctypes.POINTER(ctypes.c_double)))
        return (tp, kms1.value, n.cpu().numpy(), T.cpu().numpy(),
                flags.cpu().numpy())

    # the plain FP64 drop-in on the same cells (ratio, agreement)
    plain_res = None
    if not a.no_plain and a.kernel != "plain":
        tp, kp, n_pl, T_pl, fl_pl = one_solve(plain)
        ag = agreement(n_pl, T_pl, fl_pl, n_head, T_head, fl_head)
        plain_res = {"value": a.ncell / tp, "seconds": tp, "kernel_ms": kp,
                     "max_rel_species_vs_fp64": ag["max_rel_species"],
                     "max_rel_species_vs_fp64_by_species":
                         ag["max_rel_by_species"],
                     "max_rel_Tgas_vs_fp64": ag["max_rel_Tgas"],
                     "flags_equal": ag["flags_equal"]}
    # the persistent kernel in FP64 on the same cells: what the mixed
    # precision adds over persistence
    p64_res = None
    if mixed and not a.no_fp64_persistent:
        tq, kq, n_q, T_q, fl_q = one_solve(pers64)
        ag = agreement(n_q, T_q, fl_q, n_head, T_head, fl_head)
        p64_res = {"value": a.ncell / tq, "seconds": tq, "kernel_ms": kq,
                   "flags_nonzero": int((fl_q != 0).sum()),
                   "mixed_vs_persistent_fp64": ag}

    # roofline of the solve kernel: algorithmic FLOPs per launch from the
    # cells' counters over the kernel's event-timed duration
    fr, fdq, flu, fvec = flops_per_cell(
        st_head, opc, sparse=(cells.network_meta("primordial3")["sparse"]
                              if a.linsol == "sparse" else None))
    tot = float((fr + fdq + flu + fvec).sum())
    # the FP32 share: the Newton matrix's formation, LU and triangular
    # solves (the difference quotients and every rhs run in FP64)
    fp32 = float(flu.sum()) if mixed else 0.0
    achieved = tot / (kern_ms * 1e-3)
    traffic, traffic_src = None, None
    tf = REPO / "profiles" / "traffic.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        key = "mixed" if mixed else ("persistent_fp64" if a.kernel ==
                                     "persistent" else "plain")
        traffic = tj.get(key)
        traffic_src = tj.get("_source_" + key)
    roof = {"bound": "fp64_valu", "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64, "traffic": traffic,
            "traffic_source": traffic_src,
            "compute": "FP64 vector ALU (MI355X FP64 vector peak = FP64 "
                       "matrix peak, 78.6 TFLOP/s); transcendentals count "
                       "one op each",
            "flop_per_launch": tot, "fp32_share": fp32 / tot,
            "kernel_ms": float(kern_ms),
            "algorithmic_bytes_per_launch": BYTES_PER_CELL * a.ncell,
            "hbm_frac": BYTES_PER_CELL * a.ncell / (kern_ms * 1e-3)
            / PEAK_HBM}
    cpu, oracle_ag = None, None
    if not a.no_cpu:
        m = min(a.cpu_sample, a.ncell)
        cpu, n_or, T_or, fl_or = cpu_baseline(w, m)
        # the headline's result against the oracle on the sample
        oracle_ag = agreement(np.ascontiguousarray(n_or.T), T_or, fl_or,
                              n_head[:, :m], T_head[:m], fl_head[:m])
        oracle_ag["cells"] = m
    steps = st_head[:, 0]
    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "primordial3 collapse cells, Tgas evolved, "
                               "seed 1, dt = 0.01 tff, rtol 1e-8",
                   "ncell": a.ncell, "kernel": a.kernel,
                   "linsol": a.linsol,
                   "mixed_precision": mixed, "parallelism": "1 GPU"},
        "roofline": roof, "cpu_baseline": cpu,
        "flags_nonzero": int((fl_head != 0).sum()),
        "steps_per_cell": {"p50": float(np.percentile(steps, 50)),
                           "p99": float(np.percentile(steps, 99)),
                           "max": int(steps.max())},
        "plain_fp64": plain_res,
        "ratio_vs_plain_fp64": (value / plain_res["value"]
                                if plain_res else None),
        "persistent_fp64": p64_res,
        "ratio_mixed_vs_persistent_fp64": (value / p64_res["value"]
                                           if p64_res else None),
        "headline_vs_oracle": oracle_ag,
        "ratio_vs_cpu": value / cpu["value"] if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
