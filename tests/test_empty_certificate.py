"""The preparation's certificate that a direction's tree is empty (checked
before the lattice is built; pcode 1), on the host build of the same code.

- The certificate's ratio r = max over V = {M : (Z p).M = 0} of M0^2 / q(M),
  restated here from the bordered minors with Python integers, equals the
  same maximum computed another way: linvec^T mat^-1 linvec in the oracle's
  lattice basis (exact fractions).  A wrong sign, index or ordering in the
  bordered matrix changes one of the two.
- The library certifies exactly the p-vectors whose restated minors meet
  the margin (5 linmin^2 |P_{h+1}| > 9 (qmax + 1) |P_h|, every bordered
  minor of order >= 3 negative).
- Every certified direction's tree is empty in the oracle's search too
  (no point, one node), and the whole run's points and node count equal
  the oracle's.
"""

from fractions import Fraction

import numpy as np
import pytest

from helpers import canon_points
from oracle import cpu
from pfvs_amd import gpu
from pfvs_amd.instances import config_batch


def _bareiss_minors(M):
    """Leading principal minors of an integer matrix (fraction-free
    elimination without pivoting; stops at the first zero)."""
    n = len(M)
    M = [list(r) for r in M]
    prev, out = 1, []
    for k in range(n):
        piv = M[k][k]
        out.append(piv)
        if piv == 0:
            break
        for r in range(k + 1, n):
            for c in range(k + 1, n):
                M[r][c] = (piv * M[r][c] - M[r][k] * M[k][c]) // prev
        prev = piv
    return out


def _restated(kappa, p, qmax, linmin):
    """(certified, r or None) from the bordered matrix of -Z and w = Z p."""
    h = kappa.shape[0]
    p = [int(x) for x in p]
    Z = [[sum(int(kappa[a, b, c]) * p[c] for c in range(h)) for b in range(h)] for a in range(h)]
    w = [sum(Z[a][b] * p[b] for b in range(h)) for a in range(h)]
    js = [i for i in range(1, h) if w[i] != 0 and Z[i][i] != 0]
    if not js or linmin <= 0 or qmax < 0:
        return False, None
    j = min(js, key=lambda i: (abs(w[i]), i))
    order = [j, h] + [i for i in range(1, h) if i != j] + [0]
    B = [[(0 if b == h else w[b]) if a == h else (w[a] if b == h else -Z[a][b]) for b in order]
         for a in order]
    P = _bareiss_minors(B)
    if len(P) < h + 1 or any(x >= 0 for x in P[2:]):
        return False, None
    Ph, Pn = P[h - 1], P[h]
    return 5 * linmin * linmin * -Pn > 9 * (qmax + 1) * -Ph, Fraction(Ph, Pn)


def _quad_inverse(mat, v):
    """v^T mat^-1 v in exact fractions (Gauss-Jordan)."""
    n = len(mat)
    A = [[Fraction(int(mat[i][j])) for j in range(n)] + [Fraction(int(v[i]))] for i in range(n)]
    for c in range(n):
        piv = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[piv] = A[piv], A[c]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c] / A[c][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
    return sum(Fraction(int(v[i])) * A[i][n] / A[i][i] for i in range(n))


@pytest.fixture(scope="module")
def headline_sample():
    (inst, ps, dil), = config_batch(n_per=192, members=[(20, 250, 5)])
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    qmax = np.array([cpu.qmax_of(inst.Q, int(D)) for D in dil], dtype=np.int64)
    return inst, pf, qmax


def test_ratio_equals_lattice_form(headline_sample):
    inst, pf, qmax = headline_sample
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    seen = 0
    for i in range(0, len(pf), 8):
        _, r = _restated(kap, pf[i], int(qmax[i]), 13)
        st, _, B, _, mat, _ = cpu.build_lattice(kap, Mb, pf[i])
        if r is None or st != 0:
            continue
        assert r == _quad_inverse(mat.tolist(), B[0]), i
        seen += 1
    assert seen >= 16


def test_library_certifies_the_restated_set(headline_sample):
    inst, pf, qmax = headline_sample
    from test_gpu_parity import geoms_for_qmax
    geoms, pgeo = geoms_for_qmax(inst, qmax)
    r = gpu.run(geoms, pf, pgeo, mode="host", filter=False)
    assert np.all(r["pstat"] == 0)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    exp = np.array([_restated(kap, pf[i], int(qmax[i]), 13)[0] for i in range(len(pf))])
    got = r["pcode"] == 1
    assert exp.sum() >= len(pf) // 2
    assert np.array_equal(got, exp)
    # the certified trees are empty in the oracle's search: no point, one node
    assert not np.any(np.isin(r["pidx"], np.flatnonzero(got)))
    for i in np.flatnonzero(got)[::6]:
        M, K, q, pidx, st, cnt = cpu.coni_batch(kap, Mb, pf[i:i + 1], inst.Q, int(qmax[i]))
        assert st[0] == 0 and len(q) == 0 and cnt[0] == 1, i
    # and the whole run equals the oracle's
    nodes, pts = 0, []
    for v in sorted(set(int(x) for x in qmax)):
        sel = np.flatnonzero(qmax == v)
        M, K, q, pidx, st, cnt = cpu.coni_batch_threads(kap, Mb, pf[sel], inst.Q, v, threads=8)
        assert np.all(st == 0)
        nodes += cnt[0]
        pts.append((M, K, q, sel[pidx]))
    cat = [np.concatenate([p[k] for p in pts]) for k in range(4)]
    assert canon_points(r["M"], r["Kn"], r["q"], r["pidx"], 20) == canon_points(*cat, 20)
    assert int(r["stats"]["nodes"]) == nodes


def test_prep_work_count(headline_sample):
    """The host path's count of the preparation's lane-ops: a certified
    p-vector costs exactly the certificate (Z p: h^2 (h-1) multiply-adds, w:
    h (h-1), the elimination: 4 per entry of each trailing block, sizes h
    down to 1); the batch's total and maximum are the per-p-vector sums and
    maximum."""
    inst, pf, qmax = headline_sample
    from test_gpu_parity import geoms_for_qmax
    geoms, pgeo = geoms_for_qmax(inst, qmax)
    r = gpu.run(geoms, pf, pgeo, mode="host", filter=False)
    cert = np.flatnonzero(r["pcode"] == 1)
    h = 20
    closed = 2 * h * h * (h - 1) + 2 * h * (h - 1) + 4 * sum(m * m for m in range(1, h + 1))
    per = []
    for i in list(cert[:3]) + list(np.flatnonzero(r["pcode"] != 1)[:3]):
        s = gpu.run(geoms, pf[i:i + 1], pgeo[i:i + 1], mode="host", filter=False)["stats"]
        per.append((i, s["prep_work"]))
        if r["pcode"][i] == 1:
            assert s["prep_work"] == closed, i
        else:
            assert s["prep_work"] > 10 * closed, i      # two LLLs, the HNF and the factor
    sub = [i for i, _ in per]
    s = gpu.run(geoms, pf[sub], pgeo[sub], mode="host", filter=False, host_threads=2)["stats"]
    assert s["prep_work"] == sum(w for _, w in per)
    assert s["prep_work_max"] == max(w for _, w in per)
