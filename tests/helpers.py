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
