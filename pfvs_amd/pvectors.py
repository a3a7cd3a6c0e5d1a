"""p-vector enumeration: the primitive integer points of the Kahler cone in a
box, the input side of coniZpM (reference: pfvs/pvectors.py:58-89 with the
latticepts dependency's box_enum; the enumeration order is not part of the
contract, outputs are compared per p-vector).
"""

from __future__ import annotations

import math

import numpy as np


def box_enum(B: int, H, rhs: int = 1, primitive: bool = True, max_pts: int = 10**7):
    """All integer p with |p|_inf <= B and H p >= rhs (rows of H), optionally
    primitive (gcd 1), in lexicographic order.  Depth-first over the
    coordinates; a branch is cut when some row cannot reach rhs even with the
    best choice of the unset coordinates."""
    H = np.asarray(H, dtype=np.int64)
    m, n = H.shape
    # rest[r, j] = max over the box of sum_{k >= j} H[r, k] p_k
    rest = np.zeros((m, n + 1), dtype=np.int64)
    for j in range(n - 1, -1, -1):
        rest[:, j] = rest[:, j + 1] + B * np.abs(H[:, j])
    out = []
    p = [0] * n
    part = np.zeros(m, dtype=np.int64)
    Hc = [H[:, j].copy() for j in range(n)]

    def rec(j, part):
        if len(out) >= max_pts:
            return
        if j == n:
            if np.all(part >= rhs) and (not primitive or math.gcd(*p) == 1):
                out.append(list(p))
            return
        col = Hc[j]
        need = rhs - rest[:, j + 1]
        for v in range(-B, B + 1):
            q = part + col * v
            if np.any(q < need):
                continue
            p[j] = v
            rec(j + 1, q)
        p[j] = 0

    rec(0, part)
    return np.array(out, dtype=np.int64).reshape(-1, n)


def pvecs(data, min_N_pts: int, primitive: bool = True):
    """Primitive cone points in the smallest box holding at least min_N_pts
    of them (pfvs/pvectors.py:58-89)."""
    H = data.H_cob if data.coni else data.H
    B = 1
    while True:
        ps = box_enum(B, H, 1, primitive)
        if len(ps) >= min_N_pts or B > 64:
            return ps
        B += 1
