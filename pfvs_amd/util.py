"""Exact integer helpers of the coniZpM drop-in: the error of an incomplete
search (the reference's IncompleteSearchError, pfvs/util.py:41-45), the
integral-input check, and products and singularity decisions on Python
integers wherever int64 could wrap."""

from __future__ import annotations

import numpy as np

from . import lattice

_I64_LIMIT = 1 << 63


class IncompleteSearchError(RuntimeError):
    """A p-vector could not be searched exactly and completely (overflow, a
    non-positive-definite ellipsoid, or more than max_N_pfvs outputs); a
    p-vector is never skipped silently."""


def magnitude(A) -> int:
    """max |a| over the entries, as a Python int (0 when empty)."""
    flat = np.asarray(A).ravel()
    if flat.size == 0:
        return 0
    if flat.dtype == object:
        return max(abs(int(x)) for x in flat)
    lo, hi = int(flat.min()), int(flat.max())
    return max(-lo, hi)


def fits_int64(A) -> bool:
    return magnitude(A) < _I64_LIMIT


def to_pyints(A):
    """An object array of Python ints with A's shape (exact arithmetic)."""
    A = np.asarray(A)
    out = np.empty(A.shape, dtype=object)
    out.ravel()[:] = [int(x) for x in A.ravel()]
    return out


def as_integral(A):
    """A as an integer array: int64, or Python ints (object) when an entry
    needs more.  Floats must be integral and below 2^53 in magnitude, where
    every float is an exact integer; anything else raises ValueError."""
    A = np.asarray(A)
    if A.dtype.kind in "iu":
        return A if fits_int64(A) else to_pyints(A)
    if A.dtype == object:
        vals = list(A.ravel())
        if any(int(x) != x for x in vals):
            raise ValueError("expected an integral array")
        P = to_pyints(A)
        return P.astype(np.int64) if fits_int64(P) else P
    if A.size and not np.all(np.isfinite(A)):
        raise ValueError("expected an integral array")
    if A.size and np.any(A != np.rint(A)):
        raise ValueError("expected an integral array")
    if A.size and float(np.max(np.abs(A))) >= 2.0**53:
        raise ValueError("float input too large to be exactly integral")
    return A.astype(np.int64)


def exact_matmul(A, B):
    """A @ B on integers with no wraparound: int64 when the bound
    inner * max|A| * max|B| keeps every partial sum inside int64, else on
    Python ints (returned as int64 when the result fits)."""
    A, B = as_integral(A), as_integral(B)
    inner = A.shape[-1] if A.ndim else 1
    if A.dtype != object and B.dtype != object and \
            inner * magnitude(A) * magnitude(B) < _I64_LIMIT:
        return A @ B
    R = to_pyints(A) @ to_pyints(B)
    return R.astype(np.int64) if fits_int64(R) else R


def singular_mask(Ns) -> np.ndarray:
    """Which integer matrices of the stack Ns (n, m, m) are singular, by a
    fraction-free determinant on Python ints (exact)."""
    Ns = np.asarray(Ns)
    return np.array([lattice.det_bareiss([[int(x) for x in row] for row in N]) == 0 for N in Ns],
                    dtype=bool).reshape(len(Ns))
