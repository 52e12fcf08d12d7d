"""Launch-shape autotuner for this project's solve kernels.

The kernels have launch choices that no single setting serves best for
every network and batch: which kernel runs a network's cells (the lockstep
persistent kernel, whose lanes share each rhs evaluation with their state
in registers and LDS; the lane queue, where each lane solves whole cells
with its state in an interleaved HBM workspace and takes the next from the
device queue; the plain one-thread-per-cell kernel), how many cells a
lockstep wave takes per queue grab (queue_chunk), for the lane queue how
many lanes a workgroup holds (lane_block, the cells in flight per
workgroup) and how many waves run per CU (waves_per_cu), and for a wave per
cell (the networks above 64 species) how many cells a workgroup holds
(cells_per_workgroup).  The tuner times
each candidate on a sample of the caller's own cells, resident on the GPU,
from the kernels' HIP events (krome_device_last_kernel_ms), and keeps the
fastest whose flags equal the default setting's.

    import autotune, krome
    net = krome.load("primordial3")
    best = autotune.tune(net, cells_dict, base=dict(evolve_T=True, ...))
    opts = net.options(**best["options"])

    m = krome.load_multi("multi_c3")
    best = autotune.tune_multi(m, batches)      # waves_per_cu, queue_chunk

tune() and tune_multi() return {"options" or "launch": the winner,
"trials": [(setting, kernel ms, cells/s, flags equal)], "sample": cells}.
A result can be kept with save() / load() in a JSON file keyed by network,
device and batch size; nothing is cached implicitly.
"""
from __future__ import annotations

import ctypes
import itertools
import json
from pathlib import Path

import numpy as np

# the search space (the kernels' runtime launch choices)
KERNELS = ("persistent", "lanes", "plain")
QUEUE_CHUNKS = (0, 64, 256, 1024, 4096)
LANE_BLOCKS = (64, 128)
WAVES_PER_CU = (1, 2, 4)
# a wave per cell: cells (waves) per workgroup (the waves per CU follow:
# 8 up to 8 cells a workgroup, 16 at 16)
CELLS_PER_WORKGROUP = (1, 2, 4, 8, 16)


def candidates(kind):
    """the settings tried for a network, by the form its persistent
    launch takes: "lockstep" (up to 16 equations: the lockstep kernel with
    each queue chunk), "wave" (above 64 species at fixed Tgas: a wave per
    cell with each number of cells per workgroup), else
    "lane"; then the lane queue with each workgroup size and waves per CU,
    and the plain kernel"""
    out = []
    if kind == "lockstep":
        out += [{"device_kernel": "persistent", "queue_chunk": c}
                for c in QUEUE_CHUNKS]
    if kind == "wave":
        out += [{"device_kernel": "persistent", "cells_per_workgroup": c}
                for c in CELLS_PER_WORKGROUP]
    out += [{"device_kernel": "lanes", "lane_block": b, "waves_per_cu": w}
            for b, w in itertools.product(LANE_BLOCKS, WAVES_PER_CU)]
    out.append({"device_kernel": "plain"})
    return out


def form_of(nspec, evolve_T):
    """the persistent launch's form for a network"""
    if nspec + (1 if evolve_T else 0) <= 16:
        return "lockstep"
    if nspec > 64 and not evolve_T:
        return "wave"
    return "lane"


def sample_indices(ncell, sample):
    """an evenly strided subset of the batch (its density and temperature
    spread, not just its first cells)"""
    if ncell <= sample:
        return np.arange(ncell)
    return np.linspace(0, ncell - 1, sample).round().astype(np.int64)


def pick(trials):
    """the fastest trial whose flags equal the baseline's (trials[0])"""
    ok = [t for t in trials if t["flags_equal"]]
    return min(ok, key=lambda t: t["kernel_ms"]) if ok else trials[0]


def tune(net, w, base, sample=8192, reps=2, space=None, run=None):
    """the fastest launch setting of net (a krome.Network) on a sample of
    the cells w (dict of n [ncell, nspec], T, dt, par as for cells.py) with
    the solver options base; space: the settings (default candidates()
    for the network); run(setting) -> (kernel ms, flags): the timing
    function (default: the GPU through solve_device)"""
    idx = sample_indices(w["n"].shape[0], sample)
    if space is None:
        space = candidates(form_of(net.nspec, base.get("evolve_T")))
    if run is None:
        run = _gpu_runner(net, w, idx, base, reps)
    trials = []
    for s in space:
        ms, flags = run(s)
        trials.append({"setting": s, "kernel_ms": ms,
                       "cells_per_s": len(idx) / (ms * 1e-3),
                       "flags": flags})
    # the first setting's flags are the baseline the others must meet
    base_flags = trials[0]["flags"]
    for t in trials:
        t["flags_equal"] = bool(np.array_equal(t.pop("flags"), base_flags))
    best = pick(trials)
    return {"options": dict(base, **best["setting"]), "trials": trials,
            "sample": int(len(idx))}


def _gpu_runner(net, w, idx, base, reps):
    import torch
    dev = torch.device("cuda:0")
    n0 = torch.tensor(np.ascontiguousarray(w["n"][idx].T), device=dev)
    T0 = torch.tensor(np.ascontiguousarray(w["T"][idx]), device=dev)
    dt = torch.tensor(np.ascontiguousarray(w["dt"][idx]), device=dev)
    par = torch.tensor(np.ascontiguousarray(w["par"][idx]), device=dev)
    n, T = n0.clone(), T0.clone()
    flags = torch.zeros(len(idx), dtype=torch.int32, device=dev)
    kms = ctypes.c_double()
    kp = ctypes.cast(ctypes.byref(kms), ctypes.POINTER(ctypes.c_double))

    def run(setting):
        opts = net.options(**dict(base, **setting))
        best = None
        for _ in range(reps):
            n.copy_(n0)
            T.copy_(T0)
            net.solve_device(n, T, dt, par, opts=opts, flags=flags)
            if net._lib.krome_device_last_kernel_ms(kp) != 0:
                raise RuntimeError("no kernel time for %s" % setting)
            best = kms.value if best is None else min(best, kms.value)
        return best, flags.cpu().numpy().copy()
    return run


def tune_multi(m, batches, sample_frac=0.125, reps=2, space=None, run=None):
    """the fastest launch shape (waves_per_cu, queue_chunk) of a multi
    library's launch over the batches (dicts as Multi.solve_device takes),
    timed on a strided sample_frac of each batch"""
    if space is None:
        space = [{"waves_per_cu": wpc, "queue_chunk": c}
                 for wpc in (0, 2, 3) for c in (0, 64, 1024)]
    if run is None:
        run = _multi_runner(m, batches, sample_frac, reps)
    trials = []
    for s in space:
        ms, flags = run(s)
        trials.append({"setting": s, "kernel_ms": ms, "flags": flags})
    base = trials[0]["flags"]
    for t in trials:
        f = t.pop("flags")
        t["flags_equal"] = all(np.array_equal(a, b) for a, b in zip(f, base))
    best = pick(trials)
    m.set_launch(0, 0)
    return {"launch": best["setting"], "trials": trials}


def _multi_runner(m, batches, frac, reps):
    import torch
    sub = []
    for b in batches:
        nc = int(b["Tgas"].shape[0])
        idx = torch.tensor(sample_indices(nc, max(64, int(nc * frac))),
                           device=b["Tgas"].device)
        s = dict(b)
        s["n"] = b["n"][:, idx].contiguous()
        s["Tgas"] = b["Tgas"][idx].contiguous()
        s["dt"] = b["dt"][idx].contiguous()
        s["par"] = b["par"][idx].contiguous()
        s["flags"] = torch.zeros(len(idx), dtype=torch.int32,
                                 device=b["Tgas"].device)
        s.pop("stats", None)
        sub.append((s, s["n"].clone(), s["Tgas"].clone()))

    def run(setting):
        m.set_launch(setting["waves_per_cu"], setting["queue_chunk"])
        best = None
        for _ in range(reps):
            for s, n0, T0 in sub:
                s["n"].copy_(n0)
                s["Tgas"].copy_(T0)
            r, ms = m.solve_device([s for s, _, _ in sub])
            best = ms if best is None else min(best, ms)
        return best, [s["flags"].cpu().numpy().copy() for s, _, _ in sub]
    return run


def key(network, device, ncell):
    """the cache key of a tuning: network, device name, batch size bucket
    (powers of 4)"""
    b = 1
    while b * 4 <= max(ncell, 1):
        b *= 4
    return "%s|%s|%d" % (network, device, b)


def save(path, k, result):
    p = Path(path)
    d = json.loads(p.read_text()) if p.exists() else {}
    d[k] = {"options": result.get("options"), "launch": result.get("launch")}
    p.write_text(json.dumps(d, indent=1, default=str))


def load(path, k):
    p = Path(path)
    if not p.exists():
        return None
    return json.loads(p.read_text()).get(k)
