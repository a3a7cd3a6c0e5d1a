"""The synthetic CY instance family G(h11, h21, seed) and its p-vector sets.

G: coordinates c (index 0, the conifold), f, a, b, then k_1..k_{h-4}.
eta is the integer symmetric form on {c, a, b, k} with eta_ab = 1,
eta_cc = eta_kk = -2 and all else 0; kappa_{f,i,j} = eta_ij for every
permutation of the indices and every other kappa entry is 0.  The Kahler
cone hyperplanes are e_f, e_c (the conifold n-hat), e_k, e_a - e_c - sum e_k
and e_b - e_c - sum e_k.  c2_f = 24 and the other c2 entries are seeded
uniform even values in [0, 50].  Q = h11 + h21 + 2.  Optionally the data is
conjugated by a seeded unimodular U made of 3 h elementary row operations
(multipliers in -2..2) that never touch coordinate 0.

p-vector rule.  With p_c = 0 the cone conditions read p_f >= 1, p_k >= 1,
p_a > sum p_k and p_b > sum p_k, so a bound E on the raw entries leaves no
admissible p at h11 >= 20.  The bound E is therefore applied to the cone
coordinates (p_f, p_k, x_a, x_b) with p_a = sum p_k + x_a and
p_b = sum p_k + x_b, all in [1, E]; non-primitive vectors are dropped.  The
resulting p is mapped to the conjugated basis by p' = U^{-T} p.

Tree sizes.  On the lattice of the search, q(M) >= 2 p_f M_c^2 with equality
reachable, so the cut M_0 = M_c >= 13 empties a tree unless
338 p_f <= floor(D Q) (D the dilation), and the tree grows roughly like
((D Q - 338 p_f) / (2 p_f))^((h11 - 2) / 2) beyond it: with p_f = 1 at
h11 = 20 a single direction holds more than 10^7 nodes.  The workloads
therefore draw p_f from [p_f_min, E] with p_f_min = max(1, floor(D_max Q /
338)), D_max the largest dilation of the workload, which keeps every tree
finite (empty, or one leftover band below 338) while the mix of empty and
heavy trees stays uneven.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from .cydata import CYData


@dataclass
class Instance:
    h11: int
    h21: int
    seed: int
    Q: int
    data: CYData
    U: np.ndarray                 # conjugation (rows act on coordinates)
    UinvT: np.ndarray             # maps cone-basis p to the data's basis
    meta: dict = field(default_factory=dict)


def _eta(h):
    eta = np.zeros((h, h), dtype=np.int64)
    c, a, b = 0, 2, 3
    eta[a, b] = eta[b, a] = 1
    eta[c, c] = -2
    for k in range(4, h):
        eta[k, k] = -2
    return eta


def _unconjugated(h, rng):
    if h < 5:
        raise ValueError("G needs h11 >= 5")
    f = 1
    eta = _eta(h)
    kappa = np.zeros((h, h, h), dtype=np.int64)
    rest = [i for i in range(h) if i != f]
    for i in rest:
        for j in rest:
            v = eta[i, j]
            if v:
                for (x, y, z) in set(itertools.permutations((f, i, j))):
                    kappa[x, y, z] = v
    e = np.eye(h, dtype=np.int64)
    ks = list(range(4, h))
    sk = e[ks].sum(axis=0)
    H = np.vstack([e[f], e[0]] + [e[k] for k in ks]
                  + [e[2] - e[0] - sk, e[3] - e[0] - sk])
    c2 = 2 * rng.integers(0, 26, h)
    c2[f] = 24
    return kappa, c2.astype(np.int64), H


def _conjugator(h, rng):
    U = np.eye(h, dtype=np.int64)
    for _ in range(3 * h):
        i, j = rng.choice(np.arange(1, h), 2, replace=False)
        m = int(rng.integers(-2, 3))
        U[i] = U[i] + m * U[j]
    return U


def _inv_unimodular(U):
    h = U.shape[0]
    A = [[int(x) for x in row] + [int(i == j) for j in range(h)] for i, row in enumerate(U)]
    # Gauss-Jordan over Z works for a unimodular U with exact Fractions
    from fractions import Fraction
    M = [[Fraction(x) for x in r] for r in A]
    for col in range(h):
        piv = next(r for r in range(col, h) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [x / pv for x in M[col]]
        for r in range(h):
            if r != col and M[r][col] != 0:
                fct = M[r][col]
                M[r] = [x - fct * y for x, y in zip(M[r], M[col])]
    inv = np.array([[int(x) for x in r[h:]] for r in M], dtype=np.int64)
    assert np.array_equal(inv @ U, np.eye(h, dtype=np.int64))
    return inv


def _transform3(kappa, U):
    """kappa'_{abc} = U_ai U_bj U_ck kappa_ijk, one index at a time."""
    t = np.tensordot(U, kappa, axes=([1], [0]))                     # a j k
    t = np.tensordot(t, U, axes=([1], [1])).transpose(0, 2, 1)      # a b k
    return np.tensordot(t, U, axes=([2], [1]))                      # a b c


def make_G(h11: int, h21: int, seed: int, conjugate: bool = True) -> Instance:
    """The instance G(h11, h21, seed) as coni CYData (conifold at index 0,
    identity change of basis: the data is already in the coni basis)."""
    rng = np.random.default_rng(seed)
    kappa, c2, H = _unconjugated(h11, rng)
    if conjugate:
        U = _conjugator(h11, rng)
    else:
        U = np.eye(h11, dtype=np.int64)
    kap = _transform3(kappa, U)
    c2u = U @ c2
    Hu = H @ U.T
    e0 = np.zeros(h11, dtype=np.int64)
    e0[0] = 1
    data = CYData(h21, kap, c2u, Hu, coni_curve=e0, coni_cob=np.eye(h11, dtype=np.int64))
    UinvT = _inv_unimodular(U).T
    return Instance(h11, h21, seed, h11 + h21 + 2, data, U, UinvT)


def _cone_to_p(h, pf, pk, xa, xb):
    p = np.zeros(h, dtype=np.int64)
    p[1] = pf
    sk = int(np.sum(pk))
    p[2] = sk + xa
    p[3] = sk + xb
    p[4:] = pk
    return p


def all_pvecs(inst: Instance, E: int, pf_min: int = 1) -> np.ndarray:
    """Every primitive admissible p (without the coni entry, in the data's
    basis) whose cone coordinates all lie in [1, E], p_f >= pf_min."""
    h = inst.h11
    nk = h - 4
    out = []
    for t in itertools.product(range(1, E + 1), repeat=nk + 3):
        pf, xa, xb = t[0], t[1], t[2]
        if pf < pf_min:
            continue
        pk = np.array(t[3:], dtype=np.int64)
        p = _cone_to_p(h, pf, pk, xa, xb)
        if math.gcd(*[int(x) for x in p]) != 1:
            continue
        out.append(inst.UinvT @ p)
    P = np.array(out, dtype=np.int64).reshape(-1, h)
    return P[:, 1:]


def pf_floor(Q: int, dmax: float) -> int:
    """Smallest p_f kept by the workloads: floor(dmax Q / (2 * 13^2))."""
    return max(1, int(math.floor(dmax * Q + 1e-7)) // 338)


def random_pvecs(inst: Instance, n: int, E: int, seed: int, pf_min: int = 1,
                 pf_max: int | None = None, return_cone: bool = False):
    """n distinct primitive admissible p with cone coordinates uniform in
    [1, E] and p_f in [pf_min, pf_max] (pf_max default E) (seeded), in the
    data's basis without the coni entry; with return_cone also the cone
    coordinates (p_f, x_a, x_b, p_k...) of each."""
    h = inst.h11
    rng = np.random.default_rng(seed)
    seen, out, cone = set(), [], []
    tries = 0
    pf_hi = E if pf_max is None else pf_max
    total = (pf_hi - pf_min + 1) * E ** (h - 2)
    if n > total:
        raise ValueError("fewer admissible p-vectors than requested")
    while len(out) < n:
        tries += 1
        if tries > 50 * n + 1000:
            raise RuntimeError("could not draw enough distinct p-vectors")
        t = rng.integers(1, E + 1, h - 1)
        t[0] = rng.integers(pf_min, pf_hi + 1)
        key = tuple(int(x) for x in t)
        if key in seen:
            continue
        seen.add(key)
        pk = t[3:].astype(np.int64)
        p = _cone_to_p(h, int(t[0]), pk, int(t[1]), int(t[2]))
        if math.gcd(*[int(x) for x in p]) != 1:
            continue
        out.append(inst.UinvT @ p)
        cone.append(t.astype(np.int64))
    ps = np.array(out, dtype=np.int64)[:, 1:]
    return (ps, np.array(cone, dtype=np.int64)) if return_cone else ps


# The workloads (Q = h11 + h21 + 2, M0min = 13 throughout)
# -------------------------------------------------------
def config_cpu_plain():
    """G(5, 70, 0): all admissible p with cone coordinates <= 4, dilation 1."""
    inst = make_G(5, 70, 0)
    return inst, all_pvecs(inst, 4), 1


def config_plain_gpu():
    """G(8, 120, 1): all admissible p with cone coordinates <= 6, dilation 1."""
    inst = make_G(8, 120, 1)
    return inst, all_pvecs(inst, 6), 1


BATCH_MEMBERS = [(6, 90, 2), (10, 150, 3), (14, 200, 4), (20, 250, 5)]


def config_batch(n_per=4096, E=12, members=BATCH_MEMBERS, seed_offset=0):
    """The mixed batch: 4 instances x n_per p-vectors (cone coordinates <= E),
    per-p dilation uniform in 1..4 (seeded by the instance seed)."""
    out = []
    for (h, h21, s) in members:
        inst = make_G(h, h21, s)
        ps = random_pvecs(inst, n_per, E, seed=1000 + s + seed_offset,
                          pf_min=pf_floor(inst.Q, 4))
        dil = np.random.default_rng(2000 + s + seed_offset).integers(1, 5, len(ps))
        out.append((inst, ps, dil))
    return out


def config_large_tree(n=2048, E=20, pf_max=3, band=(8, 14), seed=1006):
    """G(30, 300, 6): n p-vectors with cone coordinates <= E and p_f <= pf_max,
    each with its own search bound qmax = 338 p_f + b, b uniform in band.

    On G the tree of p is nonempty iff 338 p_f <= qmax, and at h11 = 30 it
    grows very fast with the band qmax - 338 p_f: bands near 10 give trees
    of 10^4 - 10^5 nodes (29 levels), bands above ~25 more than 10^6 and a
    dilation of 3 (band >= 320 for every nonempty tree) far more than any
    search can enumerate.  The per-p band keeps every tree large, deep,
    uneven and finite.  Returns (inst, ps, qmax per p)."""
    inst = make_G(30, 300, 6)
    ps, cone = random_pvecs(inst, n, E, seed=seed, pf_min=1, pf_max=pf_max, return_cone=True)
    b = np.random.default_rng(seed + 1).integers(band[0], band[1] + 1, len(ps))
    qmax = 338 * cone[:, 0].astype(np.int64) + b
    return inst, ps, qmax

def geoms_for_qmax(inst: Instance, qmax, M0min: int = 13):
    """One geometry entry per distinct search bound (dilation = qmax / Q,
    which reproduces qmax exactly) and the entry of each p: (geoms, pgeo)."""
    qmax = np.asarray(qmax, dtype=np.int64)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    vals = sorted(set(int(x) for x in qmax))
    index = {v: k for k, v in enumerate(vals)}
    geoms = [dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=v / inst.Q, M0min=M0min) for v in vals]
    pgeo = np.array([index[int(x)] for x in qmax], dtype=np.int32)
    return geoms, pgeo


def serving_band_max(h: int) -> int:
    """Largest band qmax - 338 p_f of the serving stream at h11 = h (trees of
    at most ~10^5 - 10^6 nodes; the band's effect grows with h11)."""
    return 30 if h < 15 else 12 if h <= 25 else 10 if h <= 32 else 8


def serving_request(r: int, n_max: int = 4096, E: int = 16):
    """Request r of the serving stream: G(h, 10 h, seed = r) with h uniform in
    5..40, N log-uniform in 1..n_max p-vectors (cone coordinates <= E, p_f
    <= 3) and per-p bounds qmax = 338 p_f + b, b uniform in 4..band_max(h)
    (all seeded by r).  Returns (inst, ps, qmax)."""
    rng = np.random.default_rng(5000 + r)
    h = int(rng.integers(5, 41))
    n = int(round(math.exp(rng.uniform(0.0, math.log(n_max)))))
    inst = make_G(h, 10 * h, r)
    n = min(n, (3 * E ** (h - 2)) // 2)
    ps, cone = random_pvecs(inst, n, E, seed=7000 + r, pf_min=1, pf_max=3, return_cone=True)
    b = rng.integers(4, serving_band_max(h) + 1, len(ps))
    return inst, ps, 338 * cone[:, 0].astype(np.int64) + b
