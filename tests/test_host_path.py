"""The product's host path (the same preparation and search code compiled
for the host) against the oracle: golden kernel vectors and synthetic G
instances, points and node counts."""

import numpy as np
import pytest

from helpers import canon_points, kernel_problems
from oracle import cpu
from pfvs_amd import gpu
from pfvs_amd.instances import make_G, pf_floor, random_pvecs


def test_host_kernel_batch_matches_golden():
    probs, exps = kernel_problems()
    res, st, nodes = gpu.kernel_batch(probs, mode="host")
    n_ok = 0
    for (pts, qs), exp, s, p in zip(res, exps, st, probs):
        if s == 1:
            # exact host path needed: mat beyond FP64-exact integers
            assert max(abs(int(x)) for r in p["mat"] for x in r) > 2**53
            continue
        assert s == 0
        assert pts.tolist() == exp["pts"] and qs.tolist() == exp["qs"]
        n_ok += 1
    assert n_ok >= 800


@pytest.mark.parametrize("h,h21,seed,E,dil,n", [(6, 90, 2, 12, 4, 120), (8, 120, 1, 6, 3, 120),
                                                (10, 150, 3, 12, 4, 60)])
def test_host_pipeline_matches_oracle(h, h21, seed, E, dil, n):
    inst = make_G(h, h21, seed)
    ps = random_pvecs(inst, n, E, seed=1000 + seed, pf_min=pf_floor(inst.Q, 4) if h > 6 else 1)
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    M, K, q, pidx, st, cnt = cpu.coni_batch(kap, Mb, pf, inst.Q, cpu.qmax_of(inst.Q, dil))
    geo = dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=dil, M0min=13)
    r = gpu.run([geo], pf, np.zeros(len(pf), np.int32), mode="host", filter=False)
    assert np.all(st == 0) and np.all(r["pstat"] == 0)
    assert canon_points(r["M"], r["Kn"], r["q"], r["pidx"], h) == canon_points(M, K, q, pidx, h)
    # same basis and pruning as the oracle: identical node counts
    assert int(r["stats"]["nodes"]) == cnt[0]


@pytest.mark.parametrize("h,h21,seed,E,dil", [(6, 90, 2, 12, 4), (8, 120, 1, 6, 3),
                                              (10, 150, 3, 12, 4)])
def test_exact_host_path_matches_pipeline(h, h21, seed, E, dil):
    """The exact fallback for p-vectors the device hands back (Python-int
    lattice, cut-aware basis, host search) gives the pipeline's points on
    p-vectors where both apply; it shares no lattice code with it."""
    from pfvs_amd import coni
    inst = make_G(h, h21, seed)
    lo = pf_floor(inst.Q, dil) if h > 6 else 1
    ps = random_pvecs(inst, 40, E, seed=3000 + seed, pf_min=lo, pf_max=lo)     # nonempty trees
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    qmax = cpu.qmax_of(inst.Q, dil)
    geo = dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=dil, M0min=13)
    r = gpu.run([geo], pf, np.zeros(len(pf), np.int32), mode="host", filter=False)
    assert np.all(r["pstat"] == 0)
    got, n_pts = [], 0
    for ip in range(len(pf)):
        Mv, Kv, qv = coni._exact_host_path(kap, Mb, pf[ip], inst.Q, qmax, 13, 10**9)
        Mv = np.asarray(Mv.T, dtype=np.int64).reshape(-1, h)
        Kv = np.asarray(Kv.T, dtype=np.int64).reshape(-1, h)
        got += canon_points(Mv, Kv, qv, np.full(len(qv), ip), h)
        n_pts += len(qv)
    assert n_pts > 0
    assert sorted(got) == canon_points(r["M"], r["Kn"], r["q"], r["pidx"], h)
