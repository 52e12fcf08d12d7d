"""bench.py's uneven-batch workload in slices (--c3-slices, --c3-slice-range,
--c3-save), on the CPU with a stand-in for the multi library: each launch holds one
contiguous slice of every network, every cell of the covered slices goes
through exactly one launch and its results come back to its own place, an
out-of-range request is refused, and the line and the saved arrays count
the covered cells only."""
import argparse
import json

import numpy as np
import pytest
import torch

import bench
import krome


class Launches:
    """a stand-in multi library: each launch doubles n, writes Tgas + 1,
    and records per cell the launch's index as its step count and the
    number of cells of that network in the launch as its flag"""

    def __init__(self):
        self.calls = []

    def set_launch(self, *a):
        pass

    def solve_device(self, batches, mixed):
        k = len(self.calls)
        self.calls.append([int(b["Tgas"].shape[0]) for b in batches])
        for b in batches:
            b["n"].mul_(2.0)
            b["Tgas"].add_(1.0)
            b["stats"][:, 0] = k + 1
            b["flags"].fill_(int(b["Tgas"].shape[0]))
        return 0, 1.0


def run(monkeypatch, tmp_path, capsys, slices, rng):
    m = Launches()
    monkeypatch.setattr(bench, "DEVICE", "cpu")
    monkeypatch.setattr(krome, "load_multi", lambda name: m)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda: None)
    save = tmp_path / "c3.npz"
    a = argparse.Namespace(ncell=4096, waves_per_cu=0, queue_chunk=0,
                           c3_parts="", fp64=True, warmup=0, steps=1,
                           c3_slices=slices, c3_slice_range=rng,
                           c3_save=str(save))
    bench.c3_bench(a)
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    return m, line, np.load(save)


@pytest.mark.parametrize("slices,rng", [(1, ""), (4, ""), (4, "1:3"),
                                        (3, "2:3")])
def test_c3_slices(monkeypatch, tmp_path, capsys, slices, rng):
    m, line, d = run(monkeypatch, tmp_path, capsys, slices, rng)
    lo_k, hi_k = (0, slices) if not rng else map(int, rng.split(":"))
    sizes = {"primordial": 2048, "primordialZ": 1024, "cloud": 512}
    assert len(m.calls) == hi_k - lo_k
    assert line["config"]["launches_per_step"] == hi_k - lo_k
    covered = 0
    for name, nc in sizes.items():
        lo, hi = nc * lo_k // slices, nc * hi_k // slices
        covered += hi - lo
        st, fl = d[name + "_steps"], d[name + "_flags"]
        assert st.shape == (hi - lo,)
        # cell c of the range was in launch (its slice - lo_k), of the
        # slice's size
        for k in range(lo_k, hi_k):
            a, b = nc * k // slices - lo, nc * (k + 1) // slices - lo
            assert (st[a:b] == k - lo_k + 1).all()
            assert (fl[a:b] == b - a).all()
        assert line["networks"][name]["cells"] == hi - lo
    assert line["config"]["ncell"] == covered
    # each launch held a slice of every network
    for k, c in enumerate(m.calls):
        kk = lo_k + k
        assert c == [nc * (kk + 1) // slices - nc * kk // slices
                     for nc in sizes.values()]


def test_c3_slice_range_refused(monkeypatch, tmp_path, capsys):
    with pytest.raises(SystemExit):
        run(monkeypatch, tmp_path, capsys, 4, "3:5")
