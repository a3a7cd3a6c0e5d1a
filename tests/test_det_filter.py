"""The exact filter det((kappa M)[1:, 1:]) != 0 of coniZpM's post-processing
(pfvs/coni.py:397-407) on its own: the host code (CPU) and the HIP kernel
(GPU) against the oracle's exact rank, on constructed matrices that reach
every branch of the kernel -- dense, low rank (the elimination stops
early), a dependent column at full size (it runs to the last column),
entries too large for exact double arithmetic (determinants modulo
primes), kappa beyond 2^52 (N built in int128), entries of kappa M outside
int64 (undecided, 2), and geometries of different h11 in one call."""

import numpy as np
import pytest

from oracle import exact
from pfvs_amd import gpu

INT64_MAX = (1 << 63) - 1


def _N(kappa, M):
    h = len(M)
    return [[sum(int(kappa[i][j][k]) * int(M[k]) for k in range(h)) for j in range(1, h)]
            for i in range(1, h)]


def expected(kappa, M):
    N = _N(kappa, M)
    if any(abs(x) > INT64_MAX for row in N for x in row):
        return 2
    return 0 if exact.is_singular(N) else 1


def _sym(rng, n, lo, hi):
    A = rng.integers(lo, hi + 1, size=(n, n))
    return np.triu(A) + np.triu(A, 1).T


def kappa_dense(rng, h, amp=5, scale=1):
    """kappa[:, :, k] random symmetric integer slices (times `scale`)."""
    kap = np.zeros((h, h, h), dtype=object)
    for k in range(h):
        kap[:, :, k] = _sym(rng, h, -amp, amp).astype(object) * scale
    return kap


def kappa_low_rank(rng, h):
    """Rank-one slices u_k u_k^T: N = sum_k M_k u_k u_k^T has rank at most
    the number of nonzero M_k."""
    kap = np.zeros((h, h, h), dtype=object)
    for k in range(h):
        u = rng.integers(-4, 5, size=h).astype(object)
        kap[:, :, k] = np.outer(u, u)
    return kap


def kappa_dependent(rng, h, scale=1):
    """Slices B^T T_k B (B: (n-1) x n, n = h - 1) on the [1:, 1:] block, so
    every N has rank at most n - 1 but typically exactly n - 1."""
    n = h - 1
    B = rng.integers(-2, 3, size=(n - 1, n)).astype(object)
    kap = np.zeros((h, h, h), dtype=object)
    for k in range(h):
        T = _sym(rng, n - 1, -3, 3).astype(object)
        S = B.T.dot(T).dot(B) * scale
        kap[1:, 1:, k] = S
        kap[0, :, k] = kap[:, 0, k] = rng.integers(-3, 4, size=h)
    return kap


def kappa_sparse(rng, h, density=0.08, amp=4):
    """Symmetric kappa with about `density` of its entries nonzero (the
    device's sparse form of kappa applies from h11 > 25)."""
    kap = np.zeros((h, h, h), dtype=object)
    for a in range(h):
        for b in range(a, h):
            for c in range(b, h):
                if rng.random() < density:
                    v = int(rng.integers(1, amp + 1)) * int(rng.choice([-1, 1]))
                    for x, y, z in {(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)}:
                        kap[x, y, z] = v
    return kap


def as_i64(kap):
    return np.array(kap.tolist(), dtype=np.int64)


def cases(seed):
    """[(name, kappas, M rows, geo)]."""
    rng = np.random.default_rng(seed)
    out = []
    for h in (2, 3, 5, 9, 13, 17, 20, 21, 25, 33, 41):
        kap = kappa_dense(rng, h)
        M = rng.integers(-3, 4, size=(6 if h < 30 else 2, h))
        M[0] = 0                                          # N = 0
        out.append((f"dense h={h}", [kap], M, None))
    for h in (4, 9, 20, 30):
        kap = kappa_low_rank(rng, h)
        M = np.zeros((8, h), dtype=np.int64)
        for t in range(8):
            r = 1 + t % 4 if t < 7 else h                 # last row: all nonzero
            idx = rng.choice(h, size=r, replace=False)
            M[t, idx] = rng.integers(1, 6, size=r) * rng.choice([-1, 1], size=r)
        out.append((f"low rank h={h}", [kap], M, None))
    for h in (4, 8, 20):
        kap = kappa_dependent(rng, h)
        M = rng.integers(-3, 4, size=(6, h))
        out.append((f"dependent h={h}", [kap], M, None))
    # kappa near 2^20: N built in double (beyond the 24-bit integer form)
    # and kappa near 2^28: N still exact, the elimination reaches 2^53 and
    # the point moves to the modular step
    for h in (7, 20):
        for tag, sc in (("double build", 1 << 20), ("elimination overflow", 1 << 28)):
            for name, kap in (("dense", kappa_dense(rng, h, amp=3, scale=sc)),
                              ("dependent", kappa_dependent(rng, h, scale=sc))):
                M = rng.integers(-3, 4, size=(4, h))
                out.append((f"{tag} {name} h={h}", [kap], M, None))
    # entries of N near 2^40: exact double elimination overflows 2^53
    for h in (6, 12, 20):
        for name, kap in (("dense", kappa_dense(rng, h, scale=1 << 36)),
                          ("dependent", kappa_dependent(rng, h, scale=1 << 30))):
            M = rng.integers(-3, 4, size=(4, h))
            out.append((f"modular {name} h={h}", [kap], M, None))
    # |kappa| beyond 2^52: N in int128, entries still inside int64
    for h in (5, 12):
        kap = kappa_dense(rng, h, amp=3, scale=1 << 55)
        M = rng.integers(-1, 2, size=(4, h))
        M[:, 2:] = 0
        M[0] = 0                                          # N = 0
        out.append((f"int128 h={h}", [kap], M, None))
    # entries of kappa M outside int64
    h = 6
    kap = kappa_dense(rng, h, amp=1, scale=1 << 61)
    M = np.array([[4, 4, 4, 4, 4, 4], [0, 1, 0, 0, 0, 0], [8, -8, 8, -8, 8, -8]])
    out.append(("overflow h=6", [kap], M, None))
    # sparse kappa beyond h11 = 25 (the sparse form of kappa), with a
    # rank-deficient member (a column of M-independent zeros)
    for h in (26, 30, 41):
        kap = kappa_sparse(rng, h)
        M = rng.integers(-3, 4, size=(6, h))
        out.append((f"sparse h={h}", [kap], M, None))
    kap = kappa_sparse(rng, 30, density=0.05)
    kap[:, 5, :] = 0
    kap[5, :, :] = 0
    kap[:, :, 5] = 0
    out.append(("sparse singular h=30", [kap], rng.integers(-3, 4, size=(4, 30)), None))
    # geometries of different h11 in one call
    kaps = [kappa_dense(rng, 5), kappa_low_rank(rng, 20), kappa_dependent(rng, 12)]
    geo = rng.integers(0, 3, size=24).astype(np.int32)
    M = np.zeros((24, 20), dtype=np.int64)
    for t, g in enumerate(geo):
        hh = kaps[g].shape[0]
        M[t, :hh] = rng.integers(-2, 3, size=hh)
    out.append(("mixed h=5,20,12", kaps, M, geo))
    return out


def _expect(kaps, M, geo):
    return np.array([expected(kaps[0 if geo is None else geo[t]],
                              M[t, :kaps[0 if geo is None else geo[t]].shape[0]])
                     for t in range(len(M))], dtype=np.int8)


def _check(device):
    n = 0
    for name, kaps, M, geo in cases(11):
        exp = _expect(kaps, M, geo)
        got = gpu.det_filter([as_i64(k) for k in kaps], M, geo, device=device)
        assert got.tolist() == exp.tolist(), name
        n += len(exp)
    return n


def test_oracle_rank_against_sympy():
    """exact.rank agrees with sympy's rank on constructed matrices whose rank
    is known, including a planted off-by-one mistake that it must catch."""
    sympy = pytest.importorskip("sympy")
    rng = np.random.default_rng(3)
    for n, r in ((3, 1), (6, 2), (9, 9), (12, 7)):
        A = rng.integers(-4, 5, size=(n, r)).astype(object)
        B = rng.integers(-4, 5, size=(r, n)).astype(object)
        N = A.dot(B).tolist()
        assert exact.rank(N) == sympy.Matrix(N).rank() <= r
    N = [[2, 4, 6], [1, 2, 3], [0, 1, 1]]
    assert exact.rank(N) == 2 and exact.is_singular(N)
    assert not exact.is_singular([[1, 2, 3], [1, 2, 4], [0, 1, 1]])


def test_filter_cases_reach_every_verdict():
    seen = set()
    for name, kaps, M, geo in cases(11):
        seen |= set(_expect(kaps, M, geo).tolist())
    assert seen == {0, 1, 2}


def test_host_filter_matches_oracle():
    gpu.require()
    assert _check(-1) > 100


@pytest.mark.gpu
def test_gpu_filter_matches_oracle(need_gpu):
    assert _check(0) > 100


@pytest.mark.gpu
def test_gpu_filter_matches_host_on_h30_points(need_gpu):
    """The sparse form of kappa on real lattice points: G(30, 300, 6) (14 %
    of kappa nonzero), kernel against the host code."""
    from pfvs_amd.instances import config_large_tree, geoms_for_qmax
    inst, ps, qmax = config_large_tree(n=24)
    geoms, pgeo = geoms_for_qmax(inst, qmax)
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    r = gpu.run(geoms, pf, pgeo, mode="batched", filter=False)
    M = r["M"][:, :30][:30000]
    assert len(M) > 1000
    kap = inst.data.kappa_cob
    dev = gpu.det_filter([kap], M, device=0)
    host = gpu.det_filter([kap], M, device=-1)
    assert (dev == host).all()


@pytest.mark.gpu
def test_gpu_filter_matches_host_on_h20_points(need_gpu):
    """The kernel and the host code agree on the kappa M matrices of real
    lattice points (h11 = 20, the benchmark geometry), all of rank 2."""
    from pfvs_amd.instances import BATCH_MEMBERS, config_batch
    from oracle import cpu
    members = [m for m in BATCH_MEMBERS if m[0] == 20]
    (inst, ps, dil), = config_batch(n_per=4096, members=members)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    Ms = []
    for i in (343, 1022):
        M, K, q, pidx, st, cnt = cpu.coni_batch(kap, Mb, pf[i:i + 1], inst.Q,
                                                cpu.qmax_of(inst.Q, int(dil[i])))
        Ms.append(M[:20000])
    M = np.vstack(Ms)
    dev = gpu.det_filter([kap], M, device=0)
    host = gpu.det_filter([kap], M, device=-1)
    assert (dev == host).all()
    for t in range(0, len(M), 4000):
        assert dev[t] == expected(kap, M[t])
