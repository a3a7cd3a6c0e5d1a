"""Exact rational oracle -- TEST INFRASTRUCTURE ONLY.

Pure-Python restatements (Fractions and Python ints, no floating point in
any decision) used to pin the C oracle and the product on small cases:

- `ellipsoid_points` / `kernel_reference`: the exact output set and DFS
  order of the Fincke-Pohst kernel (reference tests/oracle.py:28-121,
  contract pfvs/fp_kernel/fp_kernel.h:25-63).
- `hnf`: canonical row HNF (flint's fmpz_mat_hnf form, which
  pfvs/coni.py:172-202 uses).
- `pfv_checks`: the checks of a coni PFV (reference README.md:38-45,
  pfvs/pfv.py:606-715), with p derived exactly from N^-1 K.
- `is_singular`: exact rank test (pfvs/util.py:224-243).
- `coni_points`: the exact per-p point set of the batched pipeline
  (fp_kernel.pyx:480-622) computed from scratch with Python ints: the
  lattice {M in Mbasis Z^h : (Z p).M = 0}, its ellipsoid, and the M0 / gcd
  cuts, independent of the basis any implementation picks.
- `direction_and_delta`, `schur_form`, `mu0_box`: coniZp's dilation bound
  (reference pfvs/dilation.py:1-30 and tests/test_dilation_bound.py:63-80,
  132-148): a coni PFV's direction and dilation from its fluxes, the form
  S = -A_rr^-1 with s, and the minimum of K^T S K over a box of K in
  Lambda = {p_r . K = 0} (brute force).
"""

from __future__ import annotations

import math
from fractions import Fraction


def ldl_upper(mat):
    """mat = R^T diag(d) R, R unit upper triangular, exactly."""
    n = len(mat)
    A = [[Fraction(int(mat[i][j])) for j in range(n)] for i in range(n)]
    d = [Fraction(0)] * n
    R = [[Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for i in range(n):
        d[i] = A[i][i] - sum(d[k] * R[k][i] ** 2 for k in range(i))
        if d[i] <= 0:
            raise ValueError("mat is not positive definite")
        for j in range(i + 1, n):
            R[i][j] = (A[i][j] - sum(d[k] * R[k][i] * R[k][j] for k in range(i))) / d[i]
    return d, R


def _isqrt_floor(x: Fraction) -> int:
    return math.isqrt(x.numerator * x.denominator) // x.denominator


def ellipsoid_points(mat, qmax):
    """All integer c with c^T mat c <= qmax, last coordinate outermost and
    each coordinate ascending (the kernels' DFS order)."""
    n = len(mat)
    d, R = ldl_upper(mat)
    out, c = [], [0] * n

    def rec(i, rem):
        if i < 0:
            out.append(tuple(c))
            return
        t = sum((R[i][j] * c[j] for j in range(i + 1, n)), Fraction(0))
        s = _isqrt_floor(rem / d[i]) + 1
        for v in range(math.floor(-t) - s, math.ceil(-t) + s + 1):
            e = d[i] * (v + t) ** 2
            if e <= rem:
                c[i] = v
                rec(i - 1, rem - e)
        c[i] = 0

    rec(n - 1, Fraction(qmax))
    return out


def qform(mat, c):
    n = len(c)
    return sum(int(mat[i][j]) * c[i] * c[j] for i in range(n) for j in range(n))


def row_gcd(H, c):
    g = 0
    for row in H:
        g = math.gcd(g, sum(int(h) * x for h, x in zip(row, c)))
    return g


def qmax_of(Q, dilation):
    return math.floor(Fraction(dilation) * Q + Fraction(1, 10**7))


def kernel_reference(mat, Q, dilation, H, linvec=None, linmin=None, strict=False, qmax=None):
    """(points, qs): c kept iff q(c) <= qmax, linvec.c >= linmin and
    (g == 0 or Q g >= q, strictly if `strict`), g = gcd(H c)."""
    if qmax is None:
        qmax = qmax_of(Q, dilation)
    pts, qs = [], []
    for c in ellipsoid_points(mat, qmax):
        q = qform(mat, c)
        if linvec is not None and sum(int(a) * b for a, b in zip(linvec, c)) < linmin:
            continue
        g = row_gcd(H, c)
        if g != 0 and (Q * g <= q if strict else Q * g < q):
            continue
        pts.append(c)
        qs.append(q)
    return pts, qs


def hnf(A):
    """Canonical row HNF by repeated Euclidean column clearing."""
    H = [[int(x) for x in r] for r in A]
    r = len(H)
    if r == 0:
        return []
    n = len(H[0])
    top = 0
    for col in range(n):
        if top >= r:
            break
        # gcd-combine the column entries of rows top.. into row top
        for i in range(top + 1, r):
            a, b = H[top][col], H[i][col]
            if b == 0:
                continue
            g, s, t = _xgcd(a, b)
            ra, rb = H[top], H[i]
            H[top] = [s * x + t * y for x, y in zip(ra, rb)]
            H[i] = [(-b // g) * x + (a // g) * y for x, y in zip(ra, rb)]
        if H[top][col] == 0:
            continue
        if H[top][col] < 0:
            H[top] = [-x for x in H[top]]
        p = H[top][col]
        for i in range(top):
            q = H[i][col] // p
            if q:
                H[i] = [x - q * y for x, y in zip(H[i], H[top])]
        top += 1
    return H


def _xgcd(a, b):
    r0, r1, s0, s1, t0, t1 = a, b, 1, 0, 0, 1
    while r1:
        q = r0 // r1
        r0, r1 = r1, r0 - q * r1
        s0, s1 = s1, s0 - q * s1
        t0, t1 = t1, t0 - q * t1
    if r0 < 0:
        r0, s0, t0 = -r0, -s0, -t0
    return r0, s0, t0


def rank(A):
    """Exact rank over Q (fraction-free elimination)."""
    M = [[Fraction(int(x)) for x in r] for r in A]
    if not M:
        return 0
    rk, n = 0, len(M[0])
    for col in range(n):
        piv = next((i for i in range(rk, len(M)) if M[i][col] != 0), None)
        if piv is None:
            continue
        M[rk], M[piv] = M[piv], M[rk]
        for i in range(rk + 1, len(M)):
            if M[i][col] != 0:
                f = M[i][col] / M[rk][col]
                M[i] = [x - f * y for x, y in zip(M[i], M[rk])]
        rk += 1
    return rk


def is_singular(N):
    return rank(N) < len(N)


def solve_exact(A, y):
    """x with A x = y for an invertible integer matrix A, in Fractions."""
    n = len(A)
    M = [[Fraction(int(v)) for v in row] + [Fraction(int(t))] for row, t in zip(A, y)]
    for col in range(n):
        piv = next(i for i in range(col, n) if M[i][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        for i in range(n):
            if i != col and M[i][col] != 0:
                f = M[i][col] / M[col][col]
                M[i] = [u - f * w for u, w in zip(M[i], M[col])]
    return [M[i][n] / M[i][i] for i in range(n)]


def pfv_checks(kappa, c2, H, h21, K, M):
    """The checks of a coni PFV (K, M) (PFV.check_all, pfvs/pfv.py:606-715;
    README constraints 1-7), all exact.  kappa, c2 and H in the coni basis
    (H: the Kahler cone rows on coordinates 1..h-1).  With N = (kappa M)
    [1:, 1:] invertible, p = (0, N^-1 K[1:]) (pfv.py:497-541) and
    pgrading its primitive integer multiple:
      a_even        a M is even (a_xy = kappa_xxy for x >= y, else kappa_xyy)
      b_24          (c2 + 2 e_0) . M = 0 mod 24
      tadpole       0 <= -K . M <= h11 + h21 + 4
      K_nonzero     K != 0
      N_invertible  det N != 0
      p_in_cone     H pgrading[1:] >= 1
      NpK           N p[1:] = K[1:]
      orthogonal    pgrading . K = 0
      Kprime        K' = -K_0 + (M kappa p)_0 > 0
    Returns a dict name -> bool (p-dependent checks False when N is
    singular)."""
    h = len(M)
    K = [int(x) for x in K]
    M = [int(x) for x in M]
    kap = [[[int(kappa[i][j][k]) for k in range(h)] for j in range(h)] for i in range(h)]
    a = [[kap[x][x][y] if x >= y else kap[x][y][y] for y in range(h)] for x in range(h)]
    b = [int(x) for x in c2]
    b[0] += 2
    N = [[sum(kap[i][j][k] * M[k] for k in range(h)) for j in range(1, h)] for i in range(1, h)]
    KM = sum(x * y for x, y in zip(K, M))
    out = {
        "a_even": all(sum(ai * m for ai, m in zip(row, M)) % 2 == 0 for row in a),
        "b_24": sum(bi * m for bi, m in zip(b, M)) % 24 == 0,
        "tadpole": 0 <= -KM <= h + int(h21) + 4,
        "K_nonzero": any(K),
        "N_invertible": not is_singular(N),
    }
    if not out["N_invertible"]:
        out.update(p_in_cone=False, NpK=False, orthogonal=False, Kprime=False)
        return out
    p = [Fraction(0)] + solve_exact(N, K[1:])
    den = 1
    for x in p:
        den = den * x.denominator // math.gcd(den, x.denominator)
    num = [int(x * den) for x in p]
    g = 0
    for x in num:
        g = math.gcd(g, x)
    pgrading = [x // g for x in num] if g else num
    out["p_in_cone"] = all(sum(int(r[j]) * pgrading[j + 1] for j in range(h - 1)) >= 1 for r in H)
    out["NpK"] = [sum(N[i][j] * p[j + 1] for j in range(h - 1)) for i in range(h - 1)] == K[1:]
    out["orthogonal"] = sum(x * y for x, y in zip(pgrading, K)) == 0
    MKp0 = sum(M[i] * kap[i][0][k] * p[k] for i in range(h) for k in range(h))
    out["Kprime"] = -K[0] + MKp0 > 0
    return out


def lll_reduce(rows, ip, delta=Fraction(3, 4)):
    """LLL reduction (Lenstra, Lenstra, Lovasz 1982) of linearly independent
    integer vectors `rows` under the inner product ip(a, b), in exact
    rationals with the incremental Gram-Schmidt updates of H. Cohen, "A
    Course in Computational Algebraic Number Theory", Algorithm 2.6.3.
    Returns the reduced rows (same lattice)."""
    b = [list(r) for r in rows]
    n = len(b)
    if n <= 1:
        return b
    mu = [[Fraction(0)] * n for _ in range(n)]
    B = [Fraction(0)] * n
    B[0] = Fraction(ip(b[0], b[0]))
    k, kmax = 1, 0

    def red(k, l):
        if abs(mu[k][l]) > Fraction(1, 2):
            q = round(mu[k][l])
            b[k] = [x - q * y for x, y in zip(b[k], b[l])]
            mu[k][l] -= q
            for i in range(l):
                mu[k][i] -= q * mu[l][i]

    while k < n:
        if k > kmax:
            kmax = k
            for j in range(k):
                mu[k][j] = (Fraction(ip(b[k], b[j]))
                            - sum(mu[j][i] * mu[k][i] * B[i] for i in range(j))) / B[j]
            B[k] = ip(b[k], b[k]) - sum(mu[k][j] ** 2 * B[j] for j in range(k))
        red(k, k - 1)
        if B[k] < (delta - mu[k][k - 1] ** 2) * B[k - 1]:
            b[k], b[k - 1] = b[k - 1], b[k]
            for j in range(k - 1):
                mu[k][j], mu[k - 1][j] = mu[k - 1][j], mu[k][j]
            m = mu[k][k - 1]
            Bn = B[k] + m * m * B[k - 1]
            mu[k][k - 1] = m * B[k - 1] / Bn
            B[k] = B[k - 1] * B[k] / Bn
            B[k - 1] = Bn
            for i in range(k + 1, kmax + 1):
                t = mu[i][k]
                mu[i][k] = mu[i][k - 1] - m * t
                mu[i][k - 1] = t + mu[k][k - 1] * mu[i][k]
            k = max(1, k - 1)
            continue
        for l in range(k - 2, -1, -1):
            red(k, l)
        k += 1
    return b


def lll_gram(G, delta=Fraction(3, 4)):
    """Rows T of a unimodular matrix whose rows, as coefficient vectors, are
    an LLL-reduced basis for the positive definite Gram matrix G."""
    n = len(G)
    G = [[int(x) for x in r] for r in G]

    def ip(a, c):
        return sum(a[i] * G[i][j] * c[j] for i in range(n) if a[i] for j in range(n) if c[j])
    return lll_reduce([[int(i == j) for j in range(n)] for i in range(n)], ip, delta)


def coni_points(kappa, Mbasis, p_full, Q, qmax, linmin):
    """Exact point set of one p-vector as a set of (M, Kn, q) tuples, with no
    floating point: the same lattice, ellipsoid and cuts as the pipeline,
    but in a basis of our own (the complement of (Z p) M = 0 by Euclid
    steps, then LLL-reduced for the form exactly), so only the set -- not
    the order -- is comparable."""
    h = len(p_full)
    p = [int(x) for x in p_full]
    Mb = [[int(x) for x in r] for r in Mbasis]
    Z = [[sum(int(kappa[i][j][k]) * p[k] for k in range(h)) for j in range(h)] for i in range(h)]
    T = [sum(Z[i][j] * p[j] for j in range(h)) for i in range(h)]
    v = [sum(T[i] * Mb[i][j] for i in range(h)) for j in range(h)]
    O = lll_reduce(_orth(v), lambda x, y: sum(a * c for a, c in zip(x, y)))   # rows: coefficients
    B = [[sum(Mb[i][l] * O[a][l] for l in range(h)) for i in range(h)] for a in range(len(O))]
    d = len(B)
    ZB = [[sum(Z[i][j] * B[a][j] for j in range(h)) for i in range(h)] for a in range(d)]
    mat = [[-sum(B[a][i] * ZB[b][i] for i in range(h)) for b in range(d)] for a in range(d)]
    T = lll_gram(mat)                              # new basis vector a = sum_b T[a][b] B[b]
    B = [[sum(T[a][b] * B[b][i] for b in range(d)) for i in range(h)] for a in range(d)]
    ZB = [[sum(T[a][b] * ZB[b][i] for b in range(d)) for i in range(h)] for a in range(d)]
    mat = [[-sum(B[a][i] * ZB[b][i] for i in range(h)) for b in range(d)] for a in range(d)]
    out = set()
    for c in ellipsoid_points(mat, qmax):
        M = [sum(B[a][i] * c[a] for a in range(d)) for i in range(h)]
        Kn = [sum(ZB[a][i] * c[a] for a in range(d)) for i in range(h)]
        q = qform(mat, c)
        if M[0] < linmin:
            continue
        g = 0
        for x in Kn[1:]:
            g = math.gcd(g, x)
        if g != 0 and Q * g <= q:
            continue
        out.add((tuple(M), tuple(Kn), q))
    return out


def _orth(v):
    n = len(v)
    w = list(v)
    U = [[int(i == j) for j in range(n)] for i in range(n)]
    for k in range(1, n):
        if w[k] == 0:
            continue
        g, s, t = _xgcd(w[0], w[k])
        a, b = w[0], w[k]
        w[0], w[k] = g, 0
        r0, rk = U[0], U[k]
        U[0] = [s * x + t * y for x, y in zip(r0, rk)]
        U[k] = [(-b // g) * x + (a // g) * y for x, y in zip(r0, rk)]
    return U[1:]


def pfvs_from_points(points, kappa, h, Q, M0min):
    """(K, M, key) rows from lattice points, per point with Python ints:
    the post-processing of pfvs/coni.py:275-412 restated one point at a
    time.  points: iterable of (key, M, Kn, q).  Output order: key, then g,
    then the point's (M, Kn) order, then K0; det((kappa M)[1:, 1:]) == 0
    removes a row."""
    pts = []
    for key, M, Kn, q in points:
        M = [int(x) for x in M]
        Kn = [int(x) for x in Kn]
        q = int(q)
        kg = 0
        for x in Kn[1:]:
            kg = math.gcd(kg, x)
        kg = max(kg, 1)
        if not q < Q * kg:
            continue
        pts.append((int(key), M, Kn, q, kg))
    pts.sort(key=lambda t: (t[0], t[1], t[2]))
    rows = []
    for rank, (key, M, Kn, q, kg) in enumerate(pts):
        M0 = M[0]
        raw = (q + M0 * Kn[0]) // kg
        gmax = (Q * kg - 1) // q
        for g in range(1, gmax + 1):
            Qp = g * raw
            lo = -((Q - Qp) // M0)
            up = (Qp - Q) // M0
            up = min(up, -((-Kn[0] * g) // kg) - 1)
            for K0 in range(lo, up + 1):
                K = [K0] + [g * x // kg for x in Kn[1:]]
                rows.append((key, g, rank, K0, K, M))
    rows.sort(key=lambda t: (t[0], t[1], t[2], t[3]))
    out = []
    singular_cache = {}
    for key, g, rank, K0, K, M in rows:
        tM = tuple(M)
        if tM not in singular_cache:
            N = [[sum(int(kappa[i][j][k]) * M[k] for k in range(h)) for j in range(1, h)]
                 for i in range(1, h)]
            singular_cache[tM] = is_singular(N)
        if not singular_cache[tM]:
            out.append((K, M, key))
    return out


# ------------------------------------------------------ coniZp's bound
def _inverse(A):
    """A^-1 in Fractions, or None if A is singular."""
    n = len(A)
    M = [[Fraction(int(v)) for v in row] + [Fraction(int(i == j)) for j in range(n)]
         for i, row in enumerate(A)]
    for col in range(n):
        piv = next((i for i in range(col, n) if M[i][col] != 0), None)
        if piv is None:
            return None
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [u / pv for u in M[col]]
        for i in range(n):
            if i != col and M[i][col] != 0:
                f = M[i][col]
                M[i] = [u - f * w for u, w in zip(M[i], M[col])]
    return [row[n:] for row in M]


def schur_form(kappa, p_r):
    """(S = -A_rr^-1, s = a00 - a_r0^T A_rr^-1 a_r0) for A = kappa . (0, p_r),
    in Fractions, or None if A_rr is singular (pfvs/dilation.py:49-68)."""
    h = len(kappa)
    p = [0] + [int(x) for x in p_r]
    A = [[sum(int(kappa[a][b][c]) * p[c] for c in range(h)) for b in range(h)] for a in range(h)]
    inv = _inverse([row[1:] for row in A[1:]])
    if inv is None:
        return None
    a = [A[i][0] for i in range(1, h)]
    s = A[0][0] - sum(a[i] * inv[i][j] * a[j] for i in range(h - 1) for j in range(h - 1))
    return [[-x for x in row] for row in inv], s


def form_value(S, K):
    return sum(S[i][j] * int(K[i]) * int(K[j]) for i in range(len(K)) for j in range(len(K)))


def mu0_box(kappa, p_r, box):
    """min of K^T S K over nonzero K with |K_i| <= box and p_r . K = 0
    (brute force; the reference's _brute_mu0, tests/test_dilation_bound.py:
    132-148)."""
    import itertools
    f = schur_form(kappa, p_r)
    if f is None:
        return None
    S, _ = f
    best = None
    for K in itertools.product(range(-box, box + 1), repeat=len(p_r)):
        if not any(K) or sum(k * int(x) for k, x in zip(K, p_r)) != 0:
            continue
        v = form_value(S, K)
        best = v if best is None or v < best else best
    return best


def direction_and_delta(kappa, K, M):
    """(p_hat, delta) of a coni PFV from its fluxes: K_r = N_rr p_r with
    N = kappa . M and p = p_hat / delta, p_hat primitive with p_hat_0 = 0
    (reference tests/test_dilation_bound.py:63-80)."""
    h = len(kappa)
    N = [[sum(int(kappa[a][b][c]) * int(M[c]) for c in range(h)) for b in range(h)] for a in range(h)]
    pr = solve_exact([row[1:] for row in N[1:]], [int(x) for x in K[1:]])
    den = 1
    for x in pr:
        den = den * x.denominator // math.gcd(den, x.denominator)
    ints = [int(x * den) for x in pr]
    g = 0
    for x in ints:
        g = math.gcd(g, abs(x))
    return [0] + [x // g for x in ints], Fraction(den, g)
