"""The single-problem search contract (conipfv_kernel,
pfvs/fp_kernel/fp_kernel.h:25-63) on synthetic adversarial instances built
as the reference's tests/test_kernels.py:100-290 builds them, checked
against the oracle's exact enumeration (Fractions, brute force over the
exact LDL^T ellipsoid): H in echelon, dense, rank-deficient, beyond-int64
and huge-pivot forms or empty; boundary points; an all-zero linvec; sheared
ill-conditioned ellipsoids; coordinates beyond int32.  The host build runs
on the CPU, the plain and batched kernels on the GPU."""

import math

import numpy as np
import pytest

from oracle import exact
from pfvs_amd import gpu

STYLES = ["hnf", "dense", "rankdef", "big", "bigpivot", "empty"]


def random_pd(rng, dim, scale):
    A = rng.integers(-scale, scale + 1, (dim, dim))
    return (A.T @ A + np.eye(dim, dtype=np.int64)).astype(np.int64)


def random_H(rng, dim, style):
    """Rows of H as Python ints (reference tests/test_kernels.py:109-129)."""
    G = rng.integers(-6, 7, (dim + 1, dim))
    if style == "hnf":
        return exact.hnf(G.tolist())
    if style == "dense":                     # not echelon
        return [[int(x) for x in r] for r in G]
    if style == "rankdef":                   # rank < dim, zero rows
        G[:, 0] = 0
        G[-2:] = 0
        return exact.hnf(G.tolist())
    if style == "big":                       # entries far beyond int64
        H = exact.hnf(G.tolist())
        for r, y in zip(H, G[:, -1].tolist()):
            r[-1] = r[-1] * (2**100 + 7) + int(y)
        return H
    if style == "bigpivot":                  # a pivot row with a huge common factor
        H = exact.hnf(G.tolist())
        H[-2] = [x * 2**90 for x in H[-2]]
        return H
    if style == "empty":                     # GCD cut disabled
        return []
    raise ValueError(style)


def problem(mat, Q, qmax, H, linvec=None, linmin=0, strict=False):
    """A kernel_batch problem: rows fitting int64 prune during the search,
    wider rows join the gcd at the leaves."""
    d = len(mat)
    small = [r for r in H if any(r) and all(abs(x) < 2**63 for x in r)]
    big = [r for r in H if not all(abs(x) < 2**63 for x in r)]
    return dict(mat=np.asarray(mat, np.int64), Q=Q, qmax=qmax,
                linvec=None if linvec is None else np.asarray(linvec, np.int64),
                linmin=linmin, strict=strict, H=np.array(small, np.int64).reshape(-1, d), L=big)


def synthetic_cases():
    cases = []
    for style in STYLES:
        for kind in ("coni", "pfv"):
            for seed in range(6):
                rng = np.random.default_rng(1000 * seed + STYLES.index(style) + (kind == "pfv") * 97)
                dim = int(rng.integers(2, 6))
                mat = random_pd(rng, dim, 3)
                Q = int(rng.integers(3, 12))
                dil = float(rng.choice([1, 2.5, 7, 20]))
                H = random_H(rng, dim, style)
                linvec, linmin = None, 0
                if kind == "coni":
                    linvec = rng.integers(-3, 4, dim)
                    linmin = int(rng.integers(-5, 6))
                qmax = exact.qmax_of(Q, dil)
                exp = exact.kernel_reference(mat.tolist(), Q, dil, H, linvec, linmin,
                                             strict=(kind == "coni"))
                cases.append((f"{style}-{kind}-{seed}",
                              problem(mat, Q, qmax, H, linvec, linmin, strict=(kind == "coni")), exp))
    return cases


def boundary_cases():
    """q == floor(dilation Q) is kept (reference tests/test_kernels.py:156-167)."""
    mat = [[3, 1, 0], [1, 5, 2], [0, 2, 7]]
    out = []
    for Q, dil in [(13, 1), (7, 3), (10, 2.5), (10, 1)]:
        qmax = exact.qmax_of(Q, dil)
        out.append((f"boundary-{Q}-{dil}", problem(mat, Q, qmax, []),
                    exact.kernel_reference(mat, Q, dil, [])))
    return out


def zero_linvec_cases():
    """linvec == 0: every point passes iff linmin <= 0 (reference
    tests/test_kernels.py:170-176)."""
    mat = random_pd(np.random.default_rng(0), 3, 2).tolist()
    z = [0, 0, 0]
    return [("zerolin-1", problem(mat, 5, exact.qmax_of(5, 4), [], z, 1, strict=True), ([], [])),
            ("zerolin-0", problem(mat, 5, exact.qmax_of(5, 4), [], z, 0, strict=True),
             exact.kernel_reference(mat, 5, 4, [], z, 0, strict=True))]


def sheared(rng, dim, lo, hi, n_shears=3):
    """mat = B^T diag(dg) B, B unimodular with large shears: huge entries,
    points of small q (reference tests/test_kernels.py:215-226).  Returns
    (mat, B, dg)."""
    B = [[int(i == j) for j in range(dim)] for i in range(dim)]
    for _ in range(n_shears):
        i, j = rng.choice(dim, 2, replace=False)
        f = int(rng.integers(lo, hi))
        B[i] = [a + f * b for a, b in zip(B[i], B[j])]
    dg = [int(x) for x in rng.integers(1, 4, dim)]
    mat = [[sum(B[k][i] * dg[k] * B[k][j] for k in range(dim)) for j in range(dim)]
           for i in range(dim)]
    return mat, B, dg


def sheared_points(B, dg, qmax):
    """The exact points of q(c) = sum dg_k (B c)_k^2 <= qmax in the search's
    order (last coordinate outermost, each ascending): y = B c runs over a
    small box, c = B^-1 y (reference tests/test_kernels.py:229-244)."""
    import itertools
    from fractions import Fraction
    dim = len(dg)
    A = [[Fraction(x) for x in r] + [Fraction(int(i == j)) for j in range(dim)]
         for i, r in enumerate(B)]
    for c in range(dim):                     # Gauss-Jordan: A = [I | B^-1]
        p = next(i for i in range(c, dim) if A[i][c] != 0)
        A[c], A[p] = A[p], A[c]
        A[c] = [x / A[c][c] for x in A[c]]
        for i in range(dim):
            if i != c and A[i][c] != 0:
                A[i] = [x - A[i][c] * y for x, y in zip(A[i], A[c])]
    Binv = [[int(x) for x in r[dim:]] for r in A]
    pts = []
    for y in itertools.product(*[range(-math.isqrt(qmax // d), math.isqrt(qmax // d) + 1) for d in dg]):
        q = sum(d * v * v for d, v in zip(dg, y))
        if q <= qmax:
            pts.append((tuple(sum(Binv[i][k] * y[k] for k in range(dim)) for i in range(dim)), q))
    pts.sort(key=lambda cq: cq[0][::-1])
    return [c for c, _ in pts], [q for _, q in pts]


def sheared_cases():
    out = []
    for seed in range(40):                   # reference tests/test_kernels.py:247-265
        rng = np.random.default_rng(seed)
        dim = int(rng.integers(2, 5))
        mat, B, dg = sheared(rng, dim, 2**7, 2**9)
        if max(abs(x) for r in mat for x in r) >= 2**62:
            continue
        exp = sheared_points(B, dg, 12)
        assert len(exp[0]) > 1
        out.append((f"sheared-{seed}", problem(mat, 1, 12, []), exp))
    return out


def int32_cases():
    """(name, problem, expected points, expect the int32 status) (reference
    tests/test_kernels.py:268-282)."""
    out = []
    for seed in range(20):
        rng = np.random.default_rng(100 + seed)
        mat, B, dg = sheared(rng, 3, 2**18, 2**19, n_shears=2)
        if max(abs(x) for r in mat for x in r) >= 2**62:
            continue
        pts, qs = sheared_points(B, dg, 12)
        big = max(abs(x) for c in pts for x in c) >= 2**31
        out.append((f"int32-{seed}", problem(mat, 1, 12, []), (pts, qs), big))
    return out


def test_sheared_points_against_the_form():
    """The closed-form point list is the exact one on small shears, where
    the oracle's enumeration is fast."""
    for seed in range(6):
        rng = np.random.default_rng(500 + seed)
        mat, B, dg = sheared(rng, 3, 2, 5)
        assert sheared_points(B, dg, 12) == exact.kernel_reference(mat, 1, 0, [], qmax=12)


def check(cases, mode):
    res, st, nodes = gpu.kernel_batch([p for _, p, _ in cases], mode=mode)
    for (name, p, (pts, qs)), (gp, gq), s, n in zip(cases, res, st, nodes):
        assert s == 0, (name, s)
        assert [tuple(r) for r in gp.tolist()] == [tuple(c) for c in pts], name
        assert gq.tolist() == list(qs), name
        if mode == "host":                  # per-problem counts: host build only
            assert n >= 1 or not pts, name   # a trivially empty problem visits no node


def check_int32(mode):
    cases = int32_cases()
    res, st, _ = gpu.kernel_batch([p for _, p, _, _ in cases], mode=mode)
    n_big = 0
    for (name, _, (pts, qs), big), (gp, gq), s in zip(cases, res, st):
        if big:
            n_big += 1
            assert s == -8, (name, s)            # never a silent truncation
        else:
            assert s == 0, (name, s)
            assert [tuple(r) for r in gp.tolist()] == [tuple(c) for c in pts], name
    assert n_big > 0




def test_synthetic_host():
    check(synthetic_cases() + boundary_cases() + zero_linvec_cases(), "host")


def test_sheared_host():
    check(sheared_cases(), "host")


def test_int32_flagged_host():
    check_int32("host")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_gpu_contract(need_gpu, mode):
    print(f"kernel contract cases, mode {mode}", flush=True)
    check(synthetic_cases() + boundary_cases() + zero_linvec_cases(), mode)
    check(sheared_cases(), mode)
    check_int32(mode)
