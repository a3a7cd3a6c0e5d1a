"""ctypes bindings of oracle/pfvs_oracle.c -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this module, and only as the checker: the product never does.
"""

from __future__ import annotations

import ctypes
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# GMP 6 (headers and library) ships with the image's conda prefix
GMP_PREFIX = os.environ.get("PFVS_GMP_PREFIX", "/opt/conda")

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build(force: bool = False) -> str:
    """Compile the C restatement (gcc) into oracle/_build/liboracle.so."""
    import subprocess
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    src = os.path.join(HERE, "pfvs_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-o", LIB_PATH, src,
                        "-I" + GMP_PREFIX + "/include", "-L" + GMP_PREFIX + "/lib",
                        "-Wl,-rpath," + GMP_PREFIX + "/lib", "-lgmp", "-lm"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_kernel.restype = ctypes.c_int
        L.or_kernel.argtypes = [ctypes.c_int, _i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                _i64p, ctypes.c_int, _i64p, ctypes.c_int64, ctypes.c_long,
                                ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(_i32p), ctypes.POINTER(_i64p),
                                ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_longlong)]
        L.or_coni_batch.restype = ctypes.c_long
        L.or_coni_batch.argtypes = [ctypes.c_int, _i64p, _i64p, _i64p, ctypes.c_long,
                                    ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_long,
                                    ctypes.POINTER(_i64p), ctypes.POINTER(_i64p),
                                    ctypes.POINTER(_i64p), ctypes.POINTER(_i64p), _i32p,
                                    ctypes.POINTER(ctypes.c_longlong)]
        L.or_build_export.restype = ctypes.c_int
        L.or_build_export.argtypes = [ctypes.c_int, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p,
                                      _i64p, _i64p]
        L.or_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(_i64p)


def qmax_of(Q, dilation):
    """floor(dilation * Q), tolerant of float noise (fp_kernel.pyx:146-148)."""
    return int(math.floor(dilation * Q + 1e-7))


def exact_factor(mat, linvec):
    """U (mat = U^T U), m0N, m0V, m0Va computed exactly in Fractions and
    rounded, as fpk_factor_exact (pfvs/fp_kernel/fp_kernel.h:193-306) does
    when the int128 path cannot be used; m0Va = |m0V| (fp_kernel.h:408-409)."""
    from fractions import Fraction
    n = len(mat)
    A = [[Fraction(int(mat[i][j])) for j in range(n)] for i in range(n)]
    d = [Fraction(0)] * n
    R = [[Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for i in range(n):
        d[i] = A[i][i] - sum(d[k] * R[k][i] ** 2 for k in range(i))
        if d[i] <= 0:
            return None
        for j in range(i + 1, n):
            R[i][j] = (A[i][j] - sum(d[k] * R[k][i] * R[k][j] for k in range(i))) / d[i]
    U = np.zeros((n, n))
    for k in range(n):
        ukk = math.sqrt(float(d[k]))
        U[k, k] = ukk
        for j in range(k + 1, n):
            U[k, j] = ukk * float(R[k][j])
    m0N = np.zeros(n)
    m0V = np.zeros((n, n))
    if linvec is not None:
        z = [Fraction(0)] * n
        for k in range(n):
            z[k] = Fraction(int(linvec[k])) - sum(R[m][k] * z[m] for m in range(k))
        N = Fraction(0)
        V = [Fraction(0)] * n
        for i in range(n):
            N += z[i] * z[i] / d[i]
            m0N[i] = float(N)
            for j in range(i + 1, n):
                V[j] += R[i][j] * z[i]
                m0V[i, j] = float(V[j])
    return np.concatenate([U.ravel(), m0N, m0V.ravel(), np.abs(m0V).ravel()])


def kernel(mat, Q, qmax, H, linvec=None, linmin=0, strict=True, max_out=10**9, eps=1e-4,
           fac=None):
    """Single FP search: (pts (n, dim) int32 in DFS order, qs int64, status,
    (n_nodes, n_cand, n_leaf)).  H must fit int64.  Status 1 (the int128
    factor overflowed) is retried with the exact Fraction factor."""
    L = lib()
    mat = np.ascontiguousarray(mat, dtype=np.int64)
    dim = mat.shape[0]
    H = np.ascontiguousarray(np.asarray(H, dtype=np.int64).reshape(-1, dim))
    lv = None if linvec is None else np.ascontiguousarray(linvec, dtype=np.int64)
    pts = _i32p()
    qs = _i64p()
    n = ctypes.c_long()
    cnt = (ctypes.c_longlong * 3)()
    st = L.or_kernel(dim, _p64(mat), int(Q), int(qmax), H.shape[0],
                     _p64(H) if H.shape[0] else None, int(bool(strict)),
                     _p64(lv) if lv is not None else None, int(linmin), int(max_out), float(eps),
                     fac.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if fac is not None
                     else None,
                     ctypes.byref(pts), ctypes.byref(qs), ctypes.byref(n), cnt)
    k = n.value
    P = np.ctypeslib.as_array(pts, shape=(k, dim)).copy() if k else np.zeros((0, dim), np.int32)
    q = np.ctypeslib.as_array(qs, shape=(k,)).copy() if k else np.zeros(0, np.int64)
    if k:
        L.or_free(ctypes.cast(pts, ctypes.c_void_p))
        L.or_free(ctypes.cast(qs, ctypes.c_void_p))
    if st == 1 and fac is None:
        f = exact_factor(mat.tolist(), None if lv is None else lv.tolist())
        if f is None:
            return P, q, -9, (cnt[0], cnt[1], cnt[2])
        return kernel(mat, Q, qmax, H, linvec, linmin, strict, max_out, eps, fac=f)
    return P, q, st, (cnt[0], cnt[1], cnt[2])


def coni_batch(kappa, Mbasis, p_full, Q, qmax, linmin=13, max_out=10**9):
    """The batched CPU pipeline for one geometry: (M, Kn, q, pidx, status,
    counts) with counts = (n_nodes, n_cand, n_leaf)."""
    L = lib()
    kappa = np.ascontiguousarray(kappa, dtype=np.int64)
    Mb = np.ascontiguousarray(Mbasis, dtype=np.int64)
    h = Mb.shape[0]
    ps = np.ascontiguousarray(np.asarray(p_full, dtype=np.int64).reshape(-1, h))
    n_p = ps.shape[0]
    st = np.zeros(max(n_p, 1), dtype=np.int32)
    outs = [_i64p() for _ in range(4)]
    cnt = (ctypes.c_longlong * 3)()
    n = L.or_coni_batch(h, _p64(kappa), _p64(Mb), _p64(ps), n_p, int(Q), int(qmax), int(linmin),
                        int(max_out), *[ctypes.byref(o) for o in outs],
                        st.ctypes.data_as(_i32p), cnt)

    def take(ptr, shape):
        if n == 0:
            return np.zeros(shape, dtype=np.int64)
        a = np.ctypeslib.as_array(ptr, shape=shape).copy()
        L.or_free(ctypes.cast(ptr, ctypes.c_void_p))
        return a
    M = take(outs[0], (n, h))
    K = take(outs[1], (n, h))
    q = take(outs[2], (n,))
    pidx = take(outs[3], (n,))
    return M, K, q, pidx, st[:n_p], (cnt[0], cnt[1], cnt[2])


def coni_batch_threads(kappa, Mbasis, p_full, Q, qmax, linmin=13, max_out=10**9, threads=1):
    """coni_batch over chunks in `threads` threads (ctypes drops the GIL)."""
    p_full = np.asarray(p_full, dtype=np.int64)
    if threads <= 1 or len(p_full) < 2:
        return coni_batch(kappa, Mbasis, p_full, Q, qmax, linmin, max_out)
    chunks = [c for c in np.array_split(np.arange(len(p_full)), threads * 4) if len(c)]
    with ThreadPoolExecutor(threads) as ex:
        res = list(ex.map(lambda c: coni_batch(kappa, Mbasis, p_full[c], Q, qmax, linmin, max_out),
                          chunks))
    M = np.vstack([r[0] for r in res])
    K = np.vstack([r[1] for r in res])
    q = np.concatenate([r[2] for r in res])
    pidx = np.concatenate([c[r[3]] for c, r in zip(chunks, res)])
    st = np.concatenate([r[4] for r in res])
    cnt = tuple(sum(r[5][i] for r in res) for i in range(3))
    return M, K, q, pidx, st, cnt


def build_lattice(kappa, Mbasis, p_full):
    """(status, Z, Binter, ZB, mat, H as Python ints or None)."""
    L = lib()
    kappa = np.ascontiguousarray(kappa, dtype=np.int64)
    Mb = np.ascontiguousarray(Mbasis, dtype=np.int64)
    h = Mb.shape[0]
    d = h - 1
    p = np.ascontiguousarray(p_full, dtype=np.int64)
    Z = np.zeros((h, h), np.int64)
    B = np.zeros((h, d), np.int64)
    ZB = np.zeros((h, d), np.int64)
    mat = np.zeros((d, d), np.int64)
    Hw = np.zeros((d * d, 2), np.int64)
    st = L.or_build_export(h, _p64(kappa), _p64(Mb), _p64(p), _p64(Z), _p64(B), _p64(ZB),
                           _p64(mat), _p64(Hw))
    H = None
    if st == 0:
        lo = Hw[:, 0].view(np.uint64).astype(object)
        hi = Hw[:, 1].astype(object)
        H = (hi * (1 << 64) + lo).reshape(d, d)
    return st, Z, B, ZB, mat, H
