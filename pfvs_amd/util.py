"""Exact integer helpers of the coniZpM drop-in (pfvs/util.py:41-259)."""

from __future__ import annotations

import numpy as np

from . import lattice


class IncompleteSearchError(RuntimeError):
    """A p-vector could not be searched exactly and completely (overflow, a
    non-positive-definite ellipsoid, or more than max_N_pfvs outputs);
    pfvs/util.py:41-45."""


def absmax(A) -> int:
    A = np.asarray(A)
    return max(-int(A.min()), int(A.max())) if A.size else 0


def as_exact(*arrays):
    """Object arrays of Python ints (no wraparound)."""
    return tuple(np.array(np.asarray(a).tolist(), dtype=object).reshape(np.shape(a))
                 for a in arrays)


def as_integral(A):
    """A as an integer (or object) array; raises on non-integral entries
    (pfvs/util.py:170-179)."""
    A = np.asarray(A)
    if A.dtype.kind in "iuO":
        return A
    if not np.all(np.isfinite(A)) or not np.all(A == np.round(A)):
        raise ValueError("expected an integral array")
    if np.max(np.abs(A), initial=0) >= 2.0**53:
        raise ValueError("float input too large to be exactly integral")
    return A.astype(np.int64)


def exact_matmul(A, B):
    """Integer A @ B without int64 wraparound (pfvs/util.py:246-259)."""
    A, B = as_integral(A), as_integral(B)
    if A.dtype != object and B.dtype != object:
        k = A.shape[-1] if A.ndim else 1
        if k * absmax(A) * absmax(B) < 2**63:
            return A.astype(np.int64, copy=False) @ B.astype(np.int64, copy=False)
    Ao = np.array(A.tolist(), dtype=object).reshape(A.shape)
    Bo = np.array(B.tolist(), dtype=object).reshape(B.shape)
    R = Ao @ Bo
    if R.size == 0 or all(-2**63 <= int(x) < 2**63 for x in R.ravel()):
        return np.array(R.tolist(), dtype=np.int64).reshape(R.shape)
    return R


def colsum_prod(A, B):
    """sum(A * B, axis=0) exactly (pfvs/util.py:214-221)."""
    A, B = np.asarray(A), np.asarray(B)
    if A.dtype != object and B.dtype != object and \
            (A.shape[0] if A.ndim else 1) * absmax(A) * absmax(B) < 2**62:
        return np.sum(A * B, axis=0)
    A, B = as_exact(A, B)
    return np.sum(A * B, axis=0)


def singular_mask(Ns) -> np.ndarray:
    """Exactly which integer matrices of the stack Ns (n, m, m) are singular
    (pfvs/util.py:224-243): a fraction-free determinant on Python ints."""
    Ns = np.asarray(Ns)
    n = Ns.shape[0]
    out = np.zeros(n, dtype=bool)
    for k in range(n):
        out[k] = lattice.det_bareiss([[int(x) for x in r] for r in Ns[k]]) == 0
    return out
