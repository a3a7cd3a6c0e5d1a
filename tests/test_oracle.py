"""The oracle is pinned before it is trusted: its C restatement against the
reference's golden kernel vectors (812 cases, order and q included), the
Manwe known answers, and the exact rational oracle on random small cases
where brute force fixes the answer."""

import itertools
import json
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, kernel_cases
from oracle import cpu, exact


def test_golden_vectors_cover_the_reference_fixture():
    cases = kernel_cases()
    assert len(cases) == 219
    assert sum(len(k["expected"]) for k in cases) == 812
    bits = max(abs(int(x)).bit_length() for k in cases for r in k["H"] for x in r)
    assert bits > 64                      # wide H rows are exercised


def test_c_oracle_matches_golden_kernel_vectors():
    n = 0
    for inst in kernel_cases():
        H = np.array([[int(x) for x in r] for r in inst["H"]], dtype=object)
        if max(abs(int(x)) for x in H.ravel()) >= 2**63:
            continue                      # the C oracle takes int64 H only
        for dil, exp in inst["expected"].items():
            P, q, st, _ = cpu.kernel(np.array(inst["mat"], dtype=np.int64), inst["Q"],
                                     cpu.qmax_of(inst["Q"], float(dil)), H.astype(np.int64),
                                     np.array(inst["linvec"]), inst["M0min"], strict=True)
            assert st == 0
            assert P.tolist() == exp["pts"] and q.tolist() == exp["qs"]
            n += 1
    assert n >= 700


def test_exact_oracle_matches_golden_on_small_cases():
    """The Fraction oracle reproduces the golden vectors (cheap subset)."""
    done = 0
    for inst in kernel_cases():
        if len(inst["mat"]) > 4:
            continue
        for dil in ("1", "4"):
            if dil not in inst["expected"]:
                continue
            pts, qs = exact.kernel_reference(inst["mat"], inst["Q"], float(dil), inst["H"],
                                             inst["linvec"], inst["M0min"], strict=True)
            assert [list(c) for c in pts] == inst["expected"][dil]["pts"]
            assert qs == inst["expected"][dil]["qs"]
            done += 1
    assert done > 20


def test_manwe_kernel_known_answers():
    with open(os.path.join(GOLDEN, "manwe_kernel.json")) as f:
        kat = json.load(f)
    U = np.array(kat["U"])
    mat = np.rint(U.T @ U).astype(np.int64)
    H = np.array(kat["H"], dtype=np.int64)
    for dil, want in zip(kat["dilations"], kat["counts"]):
        if dil > 50:
            continue
        P, q, st, _ = cpu.kernel(mat, kat["Q"], cpu.qmax_of(kat["Q"], dil), H,
                                 np.array(kat["linvec"]), kat["linmin"], strict=True)
        assert st == 0 and len(P) == want


def _brute(mat, qmax, H, Q, linvec, linmin, R):
    d = len(mat)
    out = []
    for c in itertools.product(range(-R, R + 1), repeat=d):
        c = tuple(reversed(c))
        q = exact.qform(mat, c)
        if q > qmax:
            continue
        if sum(a * b for a, b in zip(linvec, c)) < linmin:
            continue
        g = exact.row_gcd(H, c)
        if g != 0 and Q * g <= q:
            continue
        out.append(c)
    out.sort(key=lambda c: c[::-1])
    return out


@pytest.mark.parametrize("seed", range(12))
def test_oracles_against_brute_force(seed):
    """Random small PD forms: exhaustive enumeration in a box that provably
    contains the ellipsoid fixes the answer (cut, order and q)."""
    rng = np.random.default_rng(seed)
    d = int(rng.integers(2, 4))
    A = rng.integers(-3, 4, (d, d))
    mat = (A.T @ A + np.eye(d, dtype=np.int64) * int(rng.integers(1, 4))).astype(np.int64)
    Q = int(rng.integers(3, 10))
    qmax = Q * int(rng.integers(1, 5))
    H = exact.hnf(rng.integers(-5, 6, (d + 1, d)).tolist())
    linvec = rng.integers(-3, 4, d).tolist()
    linmin = int(rng.integers(-4, 4))
    lam = min(np.linalg.eigvalsh(mat.astype(float)))
    R = int(math.isqrt(int(qmax / lam) + 1)) + 1
    want = _brute(mat.tolist(), qmax, H, Q, linvec, linmin, R)
    pts, qs = exact.kernel_reference(mat.tolist(), Q, 0, H, linvec, linmin, strict=True, qmax=qmax)
    assert [tuple(c) for c in pts] == want
    Hs = np.array([r for r in H if any(r)], dtype=np.int64).reshape(-1, d)
    P, q, st, _ = cpu.kernel(mat, Q, qmax, Hs, np.array(linvec), linmin, strict=True)
    assert st == 0 and [tuple(r) for r in P.tolist()] == want
    assert q.tolist() == [exact.qform(mat.tolist(), c) for c in want]


def test_planted_mistake_is_caught():
    """A dropped cross term in the quadratic form changes the answer: the
    brute-force check above would fail on it."""
    mat = [[2, 1], [1, 3]]
    wrong = lambda m, c: sum(m[i][i] * c[i] * c[i] for i in range(2))   # no cross term
    pts, _ = exact.kernel_reference(mat, 5, 2, [], None, None)
    assert any(wrong(mat, c) != exact.qform(mat, c) for c in pts)


def test_exact_hnf_known_answer():
    assert exact.hnf([[2, 4], [1, 2]]) == [[1, 2], [0, 0]]
    assert exact.hnf([[4, 6], [6, 4]]) == [[2, 8], [0, 10]]     # |det| = 20 = 2 * 10
    assert exact.is_singular([[2, 4], [1, 2]])
    assert not exact.is_singular([[10**10, 10**10 + 1], [10**10 - 1, 10**10]])
    assert not exact.is_singular([[2**70, 1], [1, 1]])
