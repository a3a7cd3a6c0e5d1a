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


def _xgcd_floor(a, b):
    r0, r1, s0, s1, t0, t1 = a, b, 1, 0, 0, 1
    while r1:
        q = r0 // r1
        r0, r1, s0, s1, t0, t1 = r1, r0 - q * r1, s1, s0 - q * s1, t1, t0 - q * t1
    return r0, s0, t0


def _plain_chain_overflows(v):
    """Whether the plain Euclid chain of the preparation's orthogonal-lattice
    step (U v = (g, 0, ..., 0), row 0 the Bezout row, row k = -(v_k/g) row 0
    + (w/g) e_k) leaves int64 on v, restated with Python integers."""
    n, lim = len(v), 2**63 - 1
    U = [[int(i == j) for j in range(n)] for i in range(n)]
    w = v[0]
    for k in range(1, n):
        if v[k] == 0:
            continue
        g, s, t = _xgcd_floor(w, v[k])
        m10, m11 = -(v[k] // g), w // g
        U[0], U[k] = ([s * a + t * b for a, b in zip(U[0], U[k])],
                      [m10 * a + m11 * b for a, b in zip(U[0], U[k])])
        w = g
        if any(abs(x) > lim for x in U[0] + U[k]):
            return True
    return False


def test_plain_chain_overflow_restatement_pins():
    assert not _plain_chain_overflows([6, 10, 15])
    assert not _plain_chain_overflows([0, 0, 7])
    big = 2**40 + 15
    assert _plain_chain_overflows([big, big + 2, 2**40 + 4, 2**41 + 3])


def _reduced_chain_case(n_max=6):
    """p-vectors of a h11 = 30 serving request whose v = (kappa p p) Mbasis
    overflows the plain chain (the preparation then takes the reduced one)."""
    from pfvs_amd.instances import geoms_for_qmax, serving_request
    inst, ps, qmax = serving_request(2)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    kapo, Mbo = np.asarray(kap, dtype=object), np.asarray(Mb, dtype=object)
    sel = []
    for i in range(len(ps)):
        p = pf[i].astype(object)
        v = [int(x) for x in (np.tensordot(kapo, p, axes=([2], [0])) @ p) @ Mbo]
        if _plain_chain_overflows(v):
            sel.append(i)
        if len(sel) == n_max:
            break
    geoms, pgeo = geoms_for_qmax(inst, qmax[sel])
    return inst, kap, Mb, pf[sel], pgeo, geoms, qmax[sel]


def test_reduced_chain_points_match_exact_path():
    """Where the plain chain overflows, the preparation's reduced chain
    keeps the p-vector on the library path (status 0), and its lattice
    points equal the exact Python-integer path's."""
    from pfvs_amd import coni
    inst, kap, Mb, pf, pgeo, geoms, qmax = _reduced_chain_case()
    assert len(pf) >= 4
    r = gpu.run(geoms, pf, pgeo, mode="host", filter=False)
    assert np.all(r["pstat"] == 0)
    h = inst.h11
    n_pts = 0
    for j in range(len(pf)):
        Mv, Kv, qv = coni._exact_host_path(kap, Mb, pf[j], inst.Q, int(qmax[j]), 13, 10**9)
        got = canon_points(r["M"], r["Kn"], r["q"], r["pidx"], h)
        got = [t for t in got if t[0] == j]
        exp = canon_points(np.asarray(Mv, dtype=np.int64).T, np.asarray(Kv, dtype=np.int64).T,
                           qv, np.full(len(qv), j), h)
        assert got == exp, j
        n_pts += len(exp)
    assert n_pts > 0


@pytest.mark.gpu
def test_gpu_reduced_chain_matches_host(need_gpu):
    inst, kap, Mb, pf, pgeo, geoms, qmax = _reduced_chain_case(n_max=32)
    rh = gpu.run(geoms, pf, pgeo, mode="host", filter=True)
    rd = gpu.run(geoms, pf, pgeo, mode="batched", filter=True)
    assert np.all(rd["pstat"] == 0) and (rd["pstat"] == rh["pstat"]).all()
    for k in ("M", "Kn", "q", "pidx", "nonsing", "nexp"):
        assert np.array_equal(rd[k], rh[k]), k
    assert rd["stats"]["nodes"] == rh["stats"]["nodes"]
