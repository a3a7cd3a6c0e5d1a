"""Exact integer lattice routines on Python ints (no overflow, no flint).

These serve the per-geometry setup (the M-lattice of a CY), the synthetic
instance generator and the exact host path that handles the rare p-vectors
whose device preparation traps on int64/int128 overflow.  They restate, with
arbitrary-precision integers, the lattice steps of the reference's per-p
setup (pfvs/fp_kernel/pfv_lattice.h:116-188, 250-368, 393-602 and
pfvs/util.py:57-152, 262-270); any basis of the same lattice gives the same
PFV set, so only the lattices (and the canonical HNF) must agree.
"""

from __future__ import annotations

import math
from fractions import Fraction
from typing import List, Sequence

Matrix = List[List[int]]


def xgcd(a: int, b: int):
    """(g, s, t) with s*a + t*b = g, using floor division like
    pfvs/util.py:75-88 (so signs match the reference's Bezout steps)."""
    r0, r1, s0, s1, t0, t1 = a, b, 1, 0, 0, 1
    while r1 != 0:
        q = r0 // r1
        r0, r1 = r1, r0 - q * r1
        s0, s1 = s1, s0 - q * s1
        t0, t1 = t1, t0 - q * t1
    return r0, s0, t0


def unimodular_for(v: Sequence[int]) -> Matrix:
    """Unimodular U (rows) with U v = (g, 0, ..., 0).

    Row 0 pairs with gcd(v); rows 1.. span the lattice orthogonal to v.
    Bezout elimination of v[k] into v[0] for k = 1..n-1, as
    pfvs/fp_kernel/pfv_lattice.h:162-188 (and util.py:90-131)."""
    n = len(v)
    w = [int(x) for x in v]
    U = [[int(i == j) for j in range(n)] for i in range(n)]
    for k in range(1, n):
        if w[k] == 0:
            continue
        a, b = w[0], w[k]
        g, s, t = xgcd(a, b)
        if g == 0:
            raise ValueError("degenerate vector")
        m10, m11 = -(b // g), a // g
        w[0], w[k] = g, 0
        r0, rk = U[0], U[k]
        U[0] = [s * x + t * y for x, y in zip(r0, rk)]
        U[k] = [m10 * x + m11 * y for x, y in zip(r0, rk)]
    return U


def orthogonal_lattice_rows(v: Sequence[int]) -> Matrix:
    """Basis (rows) of {x in Z^n : v.x = 0}."""
    return unimodular_for(v)[1:]


def matmul(A: Matrix, B: Matrix) -> Matrix:
    Bt = list(zip(*B)) if B else []
    return [[sum(a * b for a, b in zip(row, col)) for col in Bt] for row in A]


def transpose(A: Matrix) -> Matrix:
    return [list(r) for r in zip(*A)] if A else []


def lll_rows(B: Matrix, gram: Matrix | None = None, last: List[int] | None = None,
             delta: float = 0.99, max_iter: int = 1_000_000):
    """LLL-reduce the rows of B (in place) w.r.t. the inner product x^T G y
    (G = identity if gram is None), keeping the Gram matrix exact and using
    floating point only to choose the unimodular steps, as the reference's
    pfvs/fp_kernel/pfv_lattice.h:393-550 does (here without overflow).
    If `last` is given it is size-reduced against the reduced basis.
    Returns (B, last)."""
    m = len(B)
    rows = [list(map(int, r)) for r in B] + ([list(map(int, last))] if last is not None else [])
    M = len(rows)

    def ip(x, y):
        if gram is None:
            return sum(a * b for a, b in zip(x, y))
        Gy = [sum(g * b for g, b in zip(grow, y)) for grow in gram]
        return sum(a * b for a, b in zip(x, Gy))

    GG = [[0] * M for _ in range(M)]
    for i in range(M):
        for j in range(i + 1):
            GG[i][j] = GG[j][i] = ip(rows[i], rows[j])
    mu = [[0.0] * max(m, 1) for _ in range(max(m, 1))]
    bb = [0.0] * max(m, 1)

    def gs(k):
        for j in range(k):
            d = float(GG[k][j])
            for l in range(j):
                d -= mu[j][l] * mu[k][l] * bb[l]
            mu[k][j] = d / bb[j]
        d = float(GG[k][k])
        for l in range(k):
            d -= mu[k][l] * mu[k][l] * bb[l]
        bb[k] = d

    def sub(k, j, q):
        rows[k] = [a - q * b for a, b in zip(rows[k], rows[j])]
        gkk = GG[k][k] - 2 * q * GG[k][j] + q * q * GG[j][j]
        for l in range(M):
            if l == k:
                continue
            t = GG[k][l] - q * GG[j][l]
            GG[k][l] = GG[l][k] = t
        GG[k][k] = gkk

    if m >= 1:
        gs(0)
        if not bb[0] > 0:
            raise ValueError("rank loss in LLL")
    if m >= 2:
        gs(1)
    k, it = 1, 0
    while k < m:
        it += 1
        if it > max_iter:
            raise RuntimeError("LLL did not converge")
        for _ in range(4):
            changed = False
            for j in range(k - 1, -1, -1):
                q = round(mu[k][j])
                if q == 0:
                    continue
                sub(k, j, q)
                for l in range(j):
                    mu[k][l] -= q * mu[j][l]
                mu[k][j] -= q
                changed = True
            if not changed:
                break
            gs(k)
        if not bb[k] > 0:
            raise ValueError("rank loss in LLL")
        if bb[k] < (delta - mu[k][k - 1] ** 2) * bb[k - 1]:
            rows[k], rows[k - 1] = rows[k - 1], rows[k]
            GG[k], GG[k - 1] = GG[k - 1], GG[k]
            for row in GG:
                row[k], row[k - 1] = row[k - 1], row[k]
            k = max(k - 1, 1)
            if k == 1:
                gs(0)
            gs(k - 1)
            gs(k)
        else:
            k += 1
            if k < m:
                gs(k)
    if last is not None and m > 0:
        mw = [0.0] * m
        for _ in range(4):
            for j in range(m):
                d = float(GG[m][j])
                for l in range(j):
                    d -= mu[j][l] * mw[l] * bb[l]
                mw[j] = d / bb[j]
            changed = False
            for j in range(m - 1, -1, -1):
                q = round(mw[j])
                if q == 0:
                    continue
                sub(m, j, q)
                for l in range(j):
                    mw[l] -= q * mu[j][l]
                mw[j] -= q
                changed = True
            if not changed:
                break
    out = rows[:m]
    return out, (rows[m] if last is not None else None)


def lll_gram_exact(G: Matrix, delta: Fraction = Fraction(99, 100)) -> Matrix:
    """Unimodular T (columns = the reduced vectors, in the original
    coordinates) with T^T G T LLL-reduced for the positive definite form G,
    in exact rationals throughout (Cohen, A Course in Computational
    Algebraic Number Theory, Alg. 2.6.3, on the Gram matrix): no
    conditioning limit, for the rare forms the floating-point
    Gram-Schmidt of lll_rows cannot take.  The Gram-Schmidt data are
    recomputed after each swap (O(n^3) per swap)."""
    n = len(G)
    Bv = [[int(i == j) for j in range(n)] for i in range(n)]      # rows b_k
    GG = [[int(x) for x in r] for r in G]

    def gs():
        mu = [[Fraction(0)] * n for _ in range(n)]
        bb = [Fraction(0)] * n
        for i in range(n):
            for j in range(i):
                t = Fraction(GG[i][j]) - sum(mu[j][l] * mu[i][l] * bb[l] for l in range(j))
                mu[i][j] = t / bb[j]
            bb[i] = Fraction(GG[i][i]) - sum(mu[i][l] * mu[i][l] * bb[l] for l in range(i))
            if bb[i] <= 0:
                raise ValueError("the form is not positive definite")
        return mu, bb

    mu, bb = gs()
    k = 1
    while k < n:
        for j in range(k - 1, -1, -1):
            q = round(mu[k][j])
            if q == 0:
                continue
            Bv[k] = [a - q * b for a, b in zip(Bv[k], Bv[j])]
            gkk = GG[k][k] - 2 * q * GG[k][j] + q * q * GG[j][j]
            for l in range(n):
                if l != k:
                    GG[k][l] = GG[l][k] = GG[k][l] - q * GG[j][l]
            GG[k][k] = gkk
            for l in range(j):
                mu[k][l] -= q * mu[j][l]
            mu[k][j] -= q
        if bb[k] >= (delta - mu[k][k - 1] ** 2) * bb[k - 1]:
            k += 1
            continue
        Bv[k], Bv[k - 1] = Bv[k - 1], Bv[k]
        GG[k], GG[k - 1] = GG[k - 1], GG[k]
        for r in GG:
            r[k], r[k - 1] = r[k - 1], r[k]
        mu, bb = gs()
        k = max(k - 1, 1)
    return [[Bv[c][i] for c in range(n)] for i in range(n)]


def hnf_rows(A: Matrix) -> Matrix:
    """Canonical row Hermite normal form (upper echelon, positive pivots,
    entries above a pivot reduced into [0, pivot), zero rows last) -- the
    form of flint's fmpz_mat_hnf that pfvs/coni.py:172-202 and
    pfvs/fp_kernel/pfv_lattice.h:205-313 produce.  Unique for a given row
    lattice, so any correct algorithm agrees with the reference's."""
    r = len(A)
    if r == 0:
        return []
    n = len(A[0])
    H = [list(map(int, row)) for row in A]
    row = 0
    for col in range(n):
        if row >= r:
            break
        while True:
            piv = None
            for i in range(row, r):
                if H[i][col] != 0 and (piv is None or abs(H[i][col]) < abs(H[piv][col])):
                    piv = i
            if piv is None:
                break
            H[row], H[piv] = H[piv], H[row]
            done = True
            pv = H[row][col]
            for i in range(row + 1, r):
                if H[i][col] == 0:
                    continue
                q = H[i][col] // pv
                if q:
                    H[i] = [x - q * y for x, y in zip(H[i], H[row])]
                if H[i][col] != 0:
                    done = False
            if done:
                break
        if H[row][col] == 0:
            continue
        if H[row][col] < 0:
            H[row] = [-x for x in H[row]]
        pv = H[row][col]
        for i in range(row):
            q = H[i][col] // pv
            if q:
                H[i] = [x - q * y for x, y in zip(H[i], H[row])]
        row += 1
    return H


def sublattice_mod(B_cols: Matrix, a: Sequence[int], m: int) -> Matrix:
    """Columns of B spanning L; returns columns spanning {x in L : a.x = 0 mod m}.

    With w = a.B, the coefficient vectors y with w.y = 0 (mod m) are the
    projection of the lattice orthogonal to (w, m)."""
    n = len(B_cols[0])
    w = [sum(int(ai) * int(B_cols[i][j]) for i, ai in enumerate(a)) for j in range(n)]
    orth = orthogonal_lattice_rows(w + [int(m)])          # n rows of length n+1
    Y = [row[:n] for row in orth]                           # coefficient vectors
    # generated lattice of rank n: reduce the generators to a basis via HNF
    Yb = [r for r in hnf_rows(Y) if any(r)]
    if len(Yb) != n:
        raise RuntimeError("rank loss while cutting a sublattice")
    return matmul(B_cols, transpose(Yb))


def kernel_mod_lattice(A: Matrix, m: int, h: int) -> Matrix:
    """Columns spanning {M in Z^h : A M = 0 (mod m)} (rows of A are the
    constraints), LLL-reduced."""
    B = [[int(i == j) for j in range(h)] for i in range(h)]
    for a in A:
        B = sublattice_mod(B, a, m)
    red, _ = lll_rows(transpose(B))
    return transpose(red)


def det_bareiss(A: Matrix) -> int:
    """Exact determinant (fraction-free Gaussian elimination)."""
    n = len(A)
    if n == 0:
        return 1
    M = [list(map(int, r)) for r in A]
    sign, prev = 1, 1
    for k in range(n - 1):
        if M[k][k] == 0:
            sw = next((i for i in range(k + 1, n) if M[i][k] != 0), None)
            if sw is None:
                return 0
            M[k], M[sw] = M[sw], M[k]
            sign = -sign
        for i in range(k + 1, n):
            for j in range(k + 1, n):
                M[i][j] = (M[i][j] * M[k][k] - M[i][k] * M[k][j]) // prev
        prev = M[k][k]
    return sign * M[n - 1][n - 1]


def gcd_list(xs) -> int:
    g = 0
    for x in xs:
        g = math.gcd(g, int(x))
    return g
