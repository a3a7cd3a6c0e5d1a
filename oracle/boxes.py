"""CPU restatement of the p-vector enumeration (test infrastructure only).

The reference takes it from the latticepts package (>= 0.1.2, not vendored
in the reference): box_enum(B, H, rhs, cap, primitive) at
pfvs/distributed.py:121-139 and tests/test_paper_examples.py:119-121,
enum_lattice_points at pfvs/pvectors.py:58-89.  Its contract as those call
sites use it: every integer p with |p|_inf <= B and H p >= rhs row-wise,
primitive (gcd of the entries 1) when asked.  Pinned by the reference's
own recorded scan counts (tests/golden/paper_coni.json "scan") and by brute
force over the whole box on small cases.
"""

from __future__ import annotations

import itertools
import math

import numpy as np


def box_enum_dfs(B, H, rhs=1, primitive=True):
    """Depth first over the coordinates in lexicographic order (p_0
    outermost, ascending); a branch is cut when a row cannot reach rhs even
    with the best choice of the unset coordinates."""
    H = np.asarray(H, dtype=np.int64)
    m, n = H.shape
    rhs = np.broadcast_to(np.asarray(rhs, dtype=np.int64), (m,))
    rest = np.zeros((m, n + 1), dtype=np.int64)
    for j in range(n - 1, -1, -1):
        rest[:, j] = rest[:, j + 1] + B * np.abs(H[:, j])
    out, p = [], [0] * n

    def rec(j, part):
        if j == n:
            if np.all(part >= rhs) and (not primitive or math.gcd(*p) == 1):
                out.append(list(p))
            return
        need = rhs - rest[:, j + 1]
        for v in range(-B, B + 1):
            q = part + H[:, j] * v
            if np.any(q < need):
                continue
            p[j] = v
            rec(j + 1, q)
        p[j] = 0

    rec(0, np.zeros(m, dtype=np.int64))
    return np.array(out, dtype=np.int64).reshape(-1, n)


def box_enum_brute(B, H, rhs=1, primitive=True):
    """Every point of the box, filtered (small boxes only)."""
    H = np.asarray(H, dtype=np.int64)
    m, n = H.shape
    rhs = np.broadcast_to(np.asarray(rhs, dtype=np.int64), (m,))
    pts = np.array(list(itertools.product(range(-B, B + 1), repeat=n)), dtype=np.int64).reshape(-1, n)
    keep = np.all(pts @ H.T >= rhs, axis=1)
    if primitive:
        keep &= np.gcd.reduce(np.abs(pts), axis=1) == 1
    return pts[keep]
