"""The dilation bound of a coni PFV's direction (reference: pfvs/dilation.py).

With p_hat = (0, p_r) in the coni basis, A = kappa . p_hat, S = -A_rr^-1,
s = a00 - a_r0^T A_rr^-1 a_r0 and Lambda = {K in Z^(h11-1) : p_r . K = 0}:
when s <= 0 and S is positive definite on Lambda, every coni PFV of the
direction with tadpole Q has dilation delta < Q / mu0, where mu0 is the
minimum of K^T S K over nonzero K in Lambda.  Otherwise there is no bound.

mu0 is computed exactly by the library (pfa_coni_mu0, host threads: an
exact fraction-free elimination, an LLL reduction with an exact Gram
matrix, and the library's Fincke-Pohst search for the minimum); a direction
it cannot decide (a value beyond int64/int128, an ill-conditioned form)
takes the exact Python path here (rational arithmetic throughout).
"""

from __future__ import annotations

import ctypes
import math
from fractions import Fraction

import numpy as np

from . import gpu, lattice


# ------------------------------------------------------------ exact path
def _solve_columns(A, cols):
    """A^-1 applied to each column in cols (Fractions), or None if A is
    singular (Gauss-Jordan on [A | cols])."""
    m = len(A)
    k = len(cols)
    M = [[Fraction(A[i][j]) for j in range(m)] + [Fraction(c[i]) for c in cols] for i in range(m)]
    for col in range(m):
        piv = next((r for r in range(col, m) if M[r][col] != 0), None)
        if piv is None:
            return None
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [x / pv for x in M[col]]
        for r in range(m):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [x - f * y for x, y in zip(M[r], M[col])]
    return [[M[i][m + c] for i in range(m)] for c in range(k)]


def _short_vectors(G, bound):
    """Every integer y with y^T G y <= bound (G a positive definite integer
    Gram matrix), by a Fincke-Pohst walk in exact rationals: with G = L D
    L^T (L unit lower triangular), y^T G y = sum_k d_k (y_k + c_k)^2 where
    c_k = sum_{i>k} L_ik y_i, so y_k is chosen from k = n - 1 down."""
    n = len(G)
    Gf = [[Fraction(int(x)) for x in r] for r in G]
    d = [Fraction(0)] * n
    Lo = [[Fraction(0)] * n for _ in range(n)]
    for k in range(n):
        d[k] = Gf[k][k] - sum(Lo[k][j] ** 2 * d[j] for j in range(k))
        for i in range(k + 1, n):
            Lo[i][k] = (Gf[i][k] - sum(Lo[i][j] * Lo[k][j] * d[j] for j in range(k))) / d[k]
    out, y = [], [0] * n
    bound = Fraction(bound)

    def walk(k, used):
        if k < 0:
            out.append(list(y))
            return
        c = sum(Lo[i][k] * y[i] for i in range(k + 1, n))
        r = math.isqrt(math.floor((bound - used) / d[k])) + 1
        for yk in range(math.floor(-c) - r, math.ceil(-c) + r + 1):
            t = d[k] * (yk + c) ** 2
            if used + t <= bound:
                y[k] = yk
                walk(k - 1, used + t)
        y[k] = 0

    walk(n - 1, Fraction(0))
    return out


def _mu0_exact(p_r, kappa):
    """(mu0, K) in rational arithmetic, or None without a bound."""
    kappa = np.asarray(kappa)
    h = kappa.shape[0]
    p_r = [int(x) for x in np.asarray(p_r).ravel()]
    if len(p_r) != h - 1:
        raise ValueError(f"p must have h11-1 = {h - 1} entries (p_hat without the coni entry)")
    if h < 3:
        return None
    A = [[sum(int(kappa[a, b, c]) * p_r[c - 1] for c in range(1, h)) for b in range(h)]
         for a in range(h)]
    Arr = [row[1:] for row in A[1:]]
    a = [A[i][0] for i in range(1, h)]
    O = lattice.orthogonal_lattice_rows(p_r)
    sol = _solve_columns(Arr, [a] + [list(row) for row in O])
    if sol is None:
        return None
    x = sol[0]
    s = A[0][0] - sum(ai * xi for ai, xi in zip(a, x))
    if s > 0:
        return None
    n = len(O)
    # S on Lambda: -O A_rr^-1 O^T, scaled to integers by the denominators' lcm
    GS = [[-sum(Fraction(O[i][k]) * sol[1 + j][k] for k in range(h - 1)) for j in range(n)]
          for i in range(n)]
    L = 1
    for row in GS:
        for v in row:
            L = L * v.denominator // math.gcd(L, v.denominator)
    G = [[int(v * L) for v in row] for row in GS]
    try:
        T = lattice.lll_gram_exact(G)                      # raises if not positive definite
    except ValueError:
        return None
    Gr = lattice.matmul(lattice.matmul(lattice.transpose(T), G), T)
    best = min(Gr[k][k] for k in range(n))
    vals = [(sum(y[i] * Gr[i][j] * y[j] for i in range(n) for j in range(n)), y)
            for y in _short_vectors(Gr, best) if any(y)]
    q, y = min(vals, key=lambda t: t[0])
    z = [sum(T[a_][b] * y[b] for b in range(n)) for a_ in range(n)]         # Lambda coordinates
    K = [sum(O[a_][i] * z[a_] for a_ in range(n)) for i in range(h - 1)]
    return Fraction(q, L), K


# ----------------------------------------------------------- library path
def _mu0_lib(ps, kappa, threads=0):
    """(status, num, den, K) from pfa_coni_mu0 for the rows of ps."""
    lib = gpu.require()
    kappa = np.ascontiguousarray(kappa, dtype=np.int64)
    h = kappa.shape[0]
    ps = np.ascontiguousarray(np.atleast_2d(np.asarray(ps, dtype=np.int64)).reshape(-1, h - 1))
    n = len(ps)
    st = np.zeros(max(n, 1), np.int32)
    num = np.zeros(max(n, 1), np.int64)
    den = np.zeros(max(n, 1), np.int64)
    K = np.zeros((max(n, 1), h - 1), np.int64)
    err = ctypes.create_string_buffer(256)
    if lib.pfa_coni_mu0(h, gpu._ptr(kappa), n, gpu._ptr(ps), int(threads), gpu._ptr(st, gpu._P32),
                        gpu._ptr(num), gpu._ptr(den), gpu._ptr(K), err) != 0:
        raise ValueError(err.value.decode(errors="replace"))
    return st[:n], num[:n], den[:n], K[:n]


def coni_mu0(p, kappa, use_c: bool = True):
    """The minimum mu0 of K^T S K over nonzero K in Lambda and a minimizer
    K (length h11-1), or None if the bound's hypotheses fail (A_rr
    singular, s > 0, or S not positive definite on Lambda).  use_c=False:
    the exact Python path only."""
    kappa = np.asarray(kappa)
    if np.asarray(p).size != kappa.shape[0] - 1:
        raise ValueError(f"p must have h11-1 = {kappa.shape[0] - 1} entries "
                         "(p_hat without the coni entry)")
    if use_c:
        st, num, den, K = _mu0_lib(np.asarray(p)[None], kappa)
        if st[0] == 1:
            return None
        if st[0] == 0:
            return Fraction(int(num[0]), int(den[0])), [int(k) for k in K[0]]
    return _mu0_exact(p, kappa)


def coni_dilation_bound(p, kappa, Q):
    """Q / mu0 (exact; every coni PFV of the direction has a smaller
    dilation), or None if the hypotheses fail."""
    r = coni_mu0(p, kappa)
    return None if r is None else Fraction(int(Q)) / r[0]


def _mu0_all(ps, kappa, n_jobs):
    """[(status, mu0 or None)] per row of ps: the library's answer, or the
    exact path where it cannot decide."""
    ps = np.atleast_2d(np.asarray(ps, dtype=np.int64))
    st, num, den, _ = _mu0_lib(ps, kappa, threads=max(int(n_jobs), 0) if n_jobs != -1 else 0)
    out = []
    for i, p in enumerate(ps):
        if st[i] == 0:
            out.append(Fraction(int(num[i]), int(den[i])))
        elif st[i] == 1:
            out.append(None)
        else:
            r = _mu0_exact(p, kappa)
            out.append(None if r is None else r[0])
    return out


def coni_dilation_bounds(ps, kappa, Q, n_jobs: int = 1):
    """coni_dilation_bound for each row of ps (shape (n, h11-1))."""
    return [None if m is None else Fraction(int(Q)) / m for m in _mu0_all(ps, kappa, n_jobs)]


def coni_dilation_bound_ceils(ps, kappa, Q, n_jobs: int = 1) -> np.ndarray:
    """ceil(coni_dilation_bound) per row of ps as int64 (0: no bound)."""
    return np.array([0 if b is None else math.ceil(b)
                     for b in coni_dilation_bounds(ps, kappa, Q, n_jobs)], dtype=np.int64)


# ------------------------------------------------------------- routing
def bound_routing(bounds, G, K, D0s, c: float = 0.0, shared: bool = False):
    """Plan coniZp over one geometry's directions (pfvs/dilation.py:266-310):
    a direction with bound b <= b* is searched by coniZpM to its bound (the
    GPU); the others are split at D0, coniZpM up to D0 (the GPU) and coniZpK
    above it (host threads).  G(D): ZpM's time per direction at dilation D
    (increasing); K(D0): ZpK's time per direction above D0; c: the bound's
    own time per direction.  The two sides run concurrently (the run time is
    the larger), or one after the other when shared (no GPU).  Returns
    (b_star, D0, time per direction); b_star = -inf splits every direction.

    For a fixed D0, moving a direction from the split to its bound costs
    G(b) - G(D0) of ZpM time and saves K(D0) of ZpK time; the cost grows
    with b, so the best set to move is the smallest bounds, and every
    prefix of the sorted bounds is tried."""
    b = np.sort(np.asarray(bounds, dtype=float))
    b = b[np.isfinite(b)]
    n = max(len(b), 1)
    best = (-math.inf, math.nan, math.inf)
    for D0 in np.atleast_1d(D0s):
        g0, k0 = float(G(D0)), float(K(D0))
        moved = np.concatenate([[0.0], np.cumsum(np.asarray(G(b), dtype=float) - g0)])
        t_zpm = g0 + moved / n
        t_zpk = (len(b) - np.arange(len(b) + 1)) / n * k0 + c
        t = t_zpm + t_zpk if shared else np.maximum(t_zpm, t_zpk)
        i = int(np.argmin(t))
        if t[i] < best[2]:
            best = (float(b[i - 1]) if i else -math.inf, float(D0), float(t[i]))
    return best


class CostModel:
    """Measured per-direction costs of coniZp's two searches on one
    geometry (pfvs/dilation.py:313-379): ZpM's time G at increasing
    dilations D (interpolated linearly in log-log, extended by the first
    and last slopes), ZpK's time K above each split dilation of D0s, the
    bound's time c, and whether the searches share the processors."""

    def __init__(self, D, G, D0s, K, c: float = 0.0, shared: bool = False):
        self.D = np.asarray(D, dtype=float)
        g = np.maximum(np.asarray(G, dtype=float), 1e-12)
        self.G = np.maximum.accumulate(g)              # routing assumes G increasing
        self.D0s = [int(x) for x in D0s]
        K = [float(x) for x in K]
        if len(self.D) < 1 or len(self.G) != len(self.D) or np.any(np.diff(self.D) <= 0) \
                or len(K) != len(self.D0s):
            raise ValueError("CostModel: D must be increasing with one G each, one K per D0")
        self.K = dict(zip(self.D0s, K))
        self.c, self.shared = float(c), bool(shared)

    def __repr__(self):
        return (f"CostModel(D={self.D.tolist()}, G={self.G.tolist()}, D0s={self.D0s}, "
                f"K={list(self.K.values())}, c={self.c}, shared={self.shared})")

    def G_at(self, D):
        """ZpM's time per direction at dilation(s) D."""
        x = np.log(np.maximum(np.asarray(D, dtype=float), 1.0))
        lx, ly = np.log(self.D), np.log(self.G)
        if len(lx) == 1:
            return np.exp(ly[0] + x - lx[0])            # proportional to D
        y = np.interp(x, lx, ly)
        s0 = (ly[1] - ly[0]) / (lx[1] - lx[0])
        s1 = (ly[-1] - ly[-2]) / (lx[-1] - lx[-2])
        y = np.where(x < lx[0], ly[0] + s0 * (x - lx[0]), y)
        y = np.where(x > lx[-1], ly[-1] + s1 * (x - lx[-1]), y)
        return np.exp(y)

    def K_at(self, D0):
        """ZpK's time per direction above D0 (one of D0s)."""
        return self.K[int(D0)]

    def route(self, bounds):
        """bound_routing with these costs: (b_star, D0, time per direction)."""
        return bound_routing(bounds, self.G_at, self.K_at, self.D0s, self.c, self.shared)
