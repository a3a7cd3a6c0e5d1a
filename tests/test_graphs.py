"""The pipeline's segments replayed from captured HIP graphs: a batch run
with graphs off (eager launches) is the reference; the same batch is then
run several times with graphs on, with other work in between (another
batch of another shape, a host-path run, a survivors-only run), and every
replay returns exactly the eager arrays and counters."""

import numpy as np
import pytest

from pfvs_amd.instances import config_batch, make_G, random_pvecs

pytestmark = pytest.mark.gpu

KEYS = ("M", "Kn", "q", "pidx", "nexp", "nonsing", "pstat", "pcount")


def _batch(h, h21, seed, n, dil):
    inst = make_G(h, h21, seed)
    ps = random_pvecs(inst, n, 12, seed=1000 + seed, pf_min=1)
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    geo = dict(kappa=inst.data.kappa_cob, Mbasis=inst.data.M_lattice(), Q=inst.Q, dilation=dil,
               M0min=13)
    return [geo], pf, np.zeros(len(pf), np.int32)


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    for k in ("nodes", "cand", "leaf", "points", "kept", "expansions", "nonsingular"):
        assert a["stats"][k] == b["stats"][k], k


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_graph_replays_match_eager(need_gpu, monkeypatch, mode):
    g1, p1, pg1 = _batch(10, 150, 3, 64, 4)
    g2, p2, pg2 = _batch(6, 90, 2, 40, 3)
    monkeypatch.setenv("PFVS_GRAPH", "0")
    ref = need_gpu.run(g1, p1, pg1, mode=mode, filter=True)
    ref2 = need_gpu.run(g2, p2, pg2, mode=mode, filter=True)
    assert ref["stats"]["graphs"] == 0 and ref["stats"]["nodes"] > 1000
    monkeypatch.setenv("PFVS_GRAPH", "1")
    hits0 = None
    for rep in range(5):
        r = need_gpu.run(g1, p1, pg1, mode=mode, filter=True)
        print(f"{mode} replay {rep}: graphs {r['stats']['graphs']:.0f}, hits "
              f"{r['stats']['graph_hits']:.0f}", flush=True)
        _same(r, ref)
        if hits0 is None:
            hits0 = r["stats"]["graph_hits"]
        # other work between replays: another shape, the host path, a
        # survivors-only run
        _same(need_gpu.run(g2, p2, pg2, mode=mode, filter=True), ref2)
        need_gpu.run(g1, p1[:7], pg1[:7], mode="host", filter=True)
        need_gpu.run(g1, p1, pg1, mode=mode, filter=True, keep_points="survivors")
    assert r["stats"]["graph_hits"] > hits0             # later passes were replays
    assert r["stats"]["graphs"] >= 3


def test_graph_replays_mixed_batch(need_gpu, monkeypatch):
    """All four instances of the mixed batch (several prep size classes, per
    p-vector dilations) with the filter: replays equal the eager run."""
    geoms, pfs, pgeos = [], [], []
    for inst, ps, dil in config_batch(n_per=48):
        kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
        pg = np.zeros(len(ps), np.int32)
        for D in sorted(set(int(x) for x in dil)):
            pg[dil == D] = len(geoms)
            geoms.append(dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=D, M0min=13))
        pgeos.append(pg)
        pfs.append(np.hstack([np.zeros((len(ps), 1), np.int64), ps]))
    pf = np.vstack([np.pad(p, ((0, 0), (0, 21 - p.shape[1]))) for p in pfs])
    pgeo = np.concatenate(pgeos)
    monkeypatch.setenv("PFVS_GRAPH", "0")
    ref = need_gpu.run(geoms, pf, pgeo, mode="batched", filter=True)
    monkeypatch.setenv("PFVS_GRAPH", "1")
    for _ in range(3):
        _same(need_gpu.run(geoms, pf, pgeo, mode="batched", filter=True), ref)
