"""Shared helpers of the tests (loading golden vectors, building problems)."""

import gzip
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kernel_cases():
    with gzip.open(os.path.join(GOLDEN, "kernel_cases.json.gz")) as f:
        return json.load(f)["kernel"]


def geometries():
    with gzip.open(os.path.join(GOLDEN, "geometries.json.gz")) as f:
        return json.load(f)["geometries"]


def kernel_problems(cases=None):
    """(problems, expected) for every (instance, dilation) golden case:
    H rows that fit int64 prune, wider rows go to the exact leaf gcd."""
    from oracle.cpu import qmax_of
    cases = kernel_cases() if cases is None else cases
    probs, exps = [], []
    for inst in cases:
        d = len(inst["mat"])
        H = [[int(x) for x in r] for r in inst["H"]]
        small = [r for r in H if any(r) and all(abs(x) < 2**63 for x in r)]
        big = [r for r in H if not all(abs(x) < 2**63 for x in r)]
        for dil, exp in inst["expected"].items():
            probs.append(dict(mat=inst["mat"], Q=inst["Q"], qmax=qmax_of(inst["Q"], float(dil)),
                              linvec=inst["linvec"], linmin=inst["M0min"], strict=True,
                              H=np.array(small, dtype=np.int64).reshape(-1, d), L=big))
            exps.append(exp)
    return probs, exps


def canon_points(M, Kn, q, pidx, h):
    return sorted(zip(np.asarray(pidx).tolist(), map(tuple, np.asarray(M)[:, :h].tolist()),
                      map(tuple, np.asarray(Kn)[:, :h].tolist()), np.asarray(q).tolist()))


def cydata_of(g):
    """CYData of a dataset geometry (stored in the coni basis)."""
    from pfvs_amd.cydata import CYData
    h11 = g["h11"]
    kappa = np.zeros((h11,) * 3, dtype=np.int64)
    for i, j, k, v in g["kappa_coo"]:
        for a, b, c in {(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)}:
            kappa[a, b, c] = v
    Hc = np.array(g["H"], dtype=np.int64)
    H = np.hstack([np.zeros((len(Hc), 1), dtype=np.int64), Hc])
    e = np.eye(h11, dtype=np.int64)
    return CYData(g["h21"], kappa, g["c2"], H, coni_curve=e[0], coni_cob=e)


def as_set(Ks, Ms):
    return {(tuple(int(x) for x in k), tuple(int(x) for x in m)) for k, m in zip(Ks, Ms)}


def oracle_pfvs(data, ps, Q, dilation, M0min=13):
    """(Ks, Ms) from the oracle alone: its C points and the exact per-point
    post-processing of oracle/exact.py, in the reference's output order."""
    from oracle import cpu, exact
    kap, Mb = data.kappa_cob, data.M_lattice()
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), np.asarray(ps, dtype=np.int64)])
    M, K, q, pidx, st, _ = cpu.coni_batch(kap, Mb, pf, Q, cpu.qmax_of(Q, dilation), M0min)
    assert np.all(st == 0), "oracle needs its exact path for some p"
    rows = exact.pfvs_from_points(zip(pidx, M, K, q), kap.tolist(), data.h11, Q, M0min)
    h = data.h11
    Ks = np.array([r[0] for r in rows], dtype=np.int64).reshape(-1, h)
    Ms = np.array([r[1] for r in rows], dtype=np.int64).reshape(-1, h)
    return Ks, Ms


def _packed_keys(cols):
    """Exact lexicographic keys for integer columns: each column shifted to
    start at 0 and packed, most significant first, into as few uint64 words
    as its value range allows (the order of the packed words is the
    lexicographic order of the columns)."""
    words, cur, used = [], None, 0
    for c in cols:
        c = np.asarray(c, dtype=np.int64)
        lo = int(c.min()) if c.size else 0
        span = (int(c.max()) - lo) if c.size else 0
        bits = max(1, span.bit_length())
        v = (c - lo).astype(np.uint64)
        if cur is None or used + bits > 64:
            if cur is not None:
                words.append(cur)
            cur, used = np.zeros(len(c), dtype=np.uint64), 0
        cur = (cur << np.uint64(bits)) | v
        used += bits
    if cur is not None:
        words.append(cur)
    return words


def canon_arrays(M, Kn, q, pidx, h):
    """(pidx, M[:, :h], Kn[:, :h], q) sorted by (pidx, M): a canonical order
    of a point set (M determines the lattice point, so no two points of one
    p-vector share M), computed exactly and vectorised for millions of
    points."""
    M, Kn = np.asarray(M)[:, :h], np.asarray(Kn)[:, :h]
    q, pidx = np.asarray(q, dtype=np.int64), np.asarray(pidx, dtype=np.int64)
    if len(q) == 0:
        return pidx, M, Kn, q
    keys = _packed_keys([pidx] + [M[:, i] for i in range(h)])
    order = np.lexsort(keys[::-1])
    return pidx[order], M[order], Kn[order], q[order]


def same_points(a, b):
    """Two canon_arrays results are identical (every array equal)."""
    return all(np.array_equal(x, y) for x, y in zip(a, b))
