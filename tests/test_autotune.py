"""The launch-shape autotuner (krome-x_amd/autotune.py): the search space
per network, the sample, the choice (the fastest setting whose flags equal
the baseline's), the cache keys; on the GPU, a tuning of primordial3 and
of the uneven batch's multi launch whose winner solves the cells as the
default setting does."""
import numpy as np
import pytest

import autotune


def test_space_and_sample():
    small = autotune.candidates("lockstep")
    mid = autotune.candidates("lane")
    big = autotune.candidates("wave")
    kinds = {s["device_kernel"] for s in small}
    assert kinds == {"persistent", "lanes", "plain"}
    assert {s["device_kernel"] for s in mid} == {"lanes", "plain"}
    assert sorted({s["cells_per_workgroup"] for s in big
                   if "cells_per_workgroup" in s}) == [1, 2, 4, 8, 16]
    assert autotune.form_of(9, True) == "lockstep"
    assert autotune.form_of(27, False) == "lane"
    assert autotune.form_of(452, False) == "wave"
    assert sorted({s["queue_chunk"] for s in small if "queue_chunk" in s}) \
        == sorted(autotune.QUEUE_CHUNKS)
    idx = autotune.sample_indices(1000, 10)
    assert len(idx) == 10 and idx[0] == 0 and idx[-1] == 999
    assert np.all(np.diff(idx) > 0)
    assert len(autotune.sample_indices(5, 10)) == 5


def test_choice_skips_settings_that_change_flags():
    """a faster setting whose flags differ from the baseline's is not
    chosen"""
    times = {"a": 5.0, "b": 1.0, "c": 2.0}
    flags = {"a": np.zeros(4), "b": np.ones(4), "c": np.zeros(4)}
    space = [{"device_kernel": k} for k in ("a", "b", "c")]
    w = {"n": np.zeros((100, 3))}
    res = autotune.tune(None, w, {"rtol": 1e-8}, sample=10, space=space,
                        run=lambda s: (times[s["device_kernel"]],
                                       flags[s["device_kernel"]]))
    assert res["options"]["device_kernel"] == "c"
    assert res["options"]["rtol"] == 1e-8
    assert [t["flags_equal"] for t in res["trials"]] == [True, False, True]


def test_cache_round_trip(tmp_path):
    k = autotune.key("primordial3", "MI355X", 70000)
    assert k == "primordial3|MI355X|65536"
    p = tmp_path / "tuned.json"
    autotune.save(p, k, {"options": {"queue_chunk": 64}})
    assert autotune.load(p, k)["options"] == {"queue_chunk": 64}
    assert autotune.load(p, "other") is None


@pytest.mark.gpu
def test_tune_primordial3_on_gpu():
    import cells
    import krome
    net = krome.load("primordial3")
    w = cells.primordial3_cells(16384)
    base = dict(evolve_T=True, rtol=1e-8, atol_n=1e-20, atol_T=1e-6,
                mixed_precision=1, error_report=False)
    space = [{"device_kernel": "persistent", "queue_chunk": 0},
             {"device_kernel": "persistent", "queue_chunk": 256},
             {"device_kernel": "lanes", "lane_block": 64, "waves_per_cu": 4},
             {"device_kernel": "lanes", "lane_block": 128, "waves_per_cu": 2}]
    print("autotune primordial3", flush=True)
    res = autotune.tune(net, w, base, sample=4096, reps=1, space=space)
    assert len(res["trials"]) == len(space)
    assert all(t["kernel_ms"] > 0 for t in res["trials"])
    best = res["options"]
    # the winner solves the sample as the baseline does, within the bar
    idx = autotune.sample_indices(16384, 4096)
    r0 = net.solve(w["n"][idx], w["T"][idx], w["dt"][idx], par=w["par"][idx],
                   opts=net.options(**dict(base, **space[0])), check=False)
    r1 = net.solve(w["n"][idx], w["T"][idx], w["dt"][idx], par=w["par"][idx],
                   opts=net.options(**best), check=False)
    assert np.array_equal(r0.flags, r1.flags)
    big = np.abs(r0.n) > 1e-12 * r0.n.sum(axis=1, keepdims=True)
    rel = np.abs(r1.n - r0.n) / np.maximum(np.abs(r0.n), 1e-300)
    assert rel[big].max() < 1e-4
