"""The launch-parameter autotuner: it keeps the fastest setting of its grid,
caches it per workload, and refuses settings that visit different nodes
(CPU, with a stand-in runner); on the GPU it tunes a real batch and the
chosen setting gives the default setting's nodes and points."""

import numpy as np
import pytest

from pfvs_amd import autotune, gpu


def _fake_runner(times, nodes=1000, bad=None):
    calls = []

    def run(geoms, ps, pgeo, **kw):
        calls.append(kw)
        key = (kw.get("waves_per_cu"), kw.get("budget"))
        n = nodes + (1 if bad is not None and key == bad else 0)
        return {"ms_search": times.get(key, 99.0), "nodes": n}
    return run, calls


def _geoms():
    kap = np.zeros((6, 6, 6), dtype=np.int64)
    return [dict(kappa=kap, Mbasis=np.eye(6, dtype=np.int64), Q=10, dilation=1, M0min=13)]


def test_tune_picks_fastest_and_caches(tmp_path, monkeypatch):
    gpu.require()
    monkeypatch.setenv("PFVS_TUNE_CACHE", str(tmp_path / "tune.json"))
    times = {(2, 4): 5.0, (4, 8): 3.0, (4, 4): 3.5, (2, 8): 7.0}
    run, calls = _fake_runner(times)
    ps = np.zeros((10, 41), np.int64)
    pgeo = np.zeros(10, np.int32)
    grid = {"waves_per_cu": (2, 4), "budget": (4, 8), "split": (2048,)}
    best = autotune.tune(_geoms(), ps, pgeo, grid=grid, runner=run, repeats=1)
    assert (best["waves_per_cu"], best["budget"]) == (4, 8)
    assert best["ms_search"] == 3.0 and not best["cached"]
    assert len(best["table"]) == 4
    n_calls = len(calls)
    again = autotune.tune(_geoms(), ps, pgeo, grid=grid, runner=run, repeats=1)
    assert again["cached"] and (again["waves_per_cu"], again["budget"]) == (4, 8)
    assert len(calls) == n_calls                      # served from the cache


def test_tune_rejects_node_mismatch(tmp_path, monkeypatch):
    gpu.require()
    monkeypatch.setenv("PFVS_TUNE_CACHE", str(tmp_path / "tune.json"))
    run, _ = _fake_runner({}, bad=(4, 8))
    ps = np.zeros((4, 41), np.int64)
    with pytest.raises(RuntimeError, match="visited"):
        autotune.tune(_geoms(), ps, np.zeros(4, np.int32),
                      grid={"waves_per_cu": (2, 4), "budget": (4, 8), "split": (2048,)}, runner=run,
                      repeats=1, use_cache=False)


def test_sample_is_seeded_subset():
    ps = np.arange(100)[:, None] * np.ones((1, 3), np.int64)
    a = autotune.sample_batch(ps, None, 10)
    b = autotune.sample_batch(ps, None, 10)
    assert (a == b).all() and len(set(a.tolist())) == 10
    assert (autotune.sample_batch(ps, None, 1000) == np.arange(100)).all()


@pytest.mark.gpu
def test_gpu_tune_real_batch(need_gpu, tmp_path, monkeypatch):
    from pfvs_amd.instances import config_batch
    monkeypatch.setenv("PFVS_TUNE_CACHE", str(tmp_path / "tune.json"))
    (inst, ps, dil), = config_batch(n_per=256, members=[(10, 150, 3)])
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    geoms, pgeo = [], np.zeros(len(ps), np.int32)
    for k, D in enumerate([1, 2, 3, 4]):
        geoms.append(dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=D, M0min=13))
        pgeo[dil == D] = k
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    best = autotune.tune(geoms, pf, pgeo, grid={"waves_per_cu": (2, 4), "budget": (4, 16),
                                                "split": (-1, 2048)},
                         repeats=1)
    assert best["waves_per_cu"] in (2, 4) and best["budget"] in (4, 16)
    r0 = gpu.run(geoms, pf, pgeo, mode="batched")
    r1 = gpu.run(geoms, pf, pgeo, mode="batched", waves_per_cu=best["waves_per_cu"],
                 budget=best["budget"], split=best["split"])
    assert r0["stats"]["nodes"] == r1["stats"]["nodes"]
    assert (r0["M"] == r1["M"]).all() and (r0["pidx"] == r1["pidx"]).all()
