"""coniZpM: the coni PFVs of each p-vector up to a dilation -- the drop-in for
the reference's `pfvs.coni.coniZpM` (pfvs/coni.py:860-1168).

Routing: device "gpu" (raises if the HIP backend is unavailable), "cpu"
(the same preparation and search code of libpfvs_amd.so run on host
threads) or "auto" (the GPU when available and len(ps) >= 256; the
PFVS_DEVICE environment variable overrides "auto"), as pfvs/coni.py:416-443.
The library returns, for every lattice point c, M = Binter c, Kn = Z Binter c
and q = c^T mat c, plus the exact per-point filter (K-integrality count and
det((kappa M)[1:,1:]) != 0) evaluated where the points are produced.
p-vectors the device hands back (pstat 1: an int64/int128 trap) go through
the exact host path (Python-int lattice, host search).  A p-vector is never
skipped: kernel failures raise IncompleteSearchError (pfvs/coni.py:446-453).
"""

from __future__ import annotations

import math
import os
import warnings

import numpy as np

from . import gpu, lattice
from . import util
from .util import IncompleteSearchError

_GPU_MIN_PS = 256


def _use_gpu(device, h11, n_ps, use_c_lattice=True, use_gcd_lattice=False, extra_lll=True):
    """Resolve `device` (pfvs/coni.py:419-443)."""
    if device not in ("auto", "cpu", "gpu"):
        raise ValueError(f'device must be "auto", "cpu" or "gpu", got {device!r}')
    if device == "auto":
        env = os.environ.get("PFVS_DEVICE", "auto").lower()
        if env not in ("auto", "cpu", "gpu"):
            raise ValueError(f'PFVS_DEVICE must be "auto", "cpu" or "gpu", got {env!r}')
        device = env
    if device == "cpu":
        return False
    reasons = []
    if not gpu.available():
        reasons.append(gpu.unavailable_reason())
    elif not 3 <= h11 <= gpu.max_h11():
        reasons.append(f"h11 = {h11} is outside the GPU build's 3..{gpu.max_h11()}")
    if not use_c_lattice or use_gcd_lattice or not extra_lll:
        reasons.append("the GPU path needs use_c_lattice=True, use_gcd_lattice=False, "
                       "extra_lll_reduction=True")
    if device == "gpu":
        if reasons:
            raise RuntimeError("device='gpu' unavailable: " + "; ".join(reasons))
        return True
    return not reasons and n_ps >= _GPU_MIN_PS


def _raise_kernel_status(p, status: int):
    reason = {1: "(beyond the exact search's int64/int128 range: an ellipsoid matrix "
                 "entry above 2^53 or a leaf gcd beyond int128)",
              -2: "(more than max_N_pfvs outputs; increase max_N_pfvs)",
              -8: "(lattice-point coordinates exceed int32)",
              -9: "(the ellipsoid matrix is not positive definite, so its "
                  "lattice points cannot be enumerated)"}.get(status, "")
    raise IncompleteSearchError(
        f"p={np.asarray(p).tolist()}: conipfv_kernel returned status {status} {reason}")


def _in_kahler_cone(H, ps):
    H, ps = np.asarray(H), np.asarray(ps)
    if H.dtype != object and ps.dtype != object and np.issubdtype(ps.dtype, np.integer) \
            and H.shape[1] * util.absmax(H) * util.absmax(ps) < 2**53:
        for i in range(0, len(ps), 1 << 18):
            if not np.all(H.astype(np.float64) @ ps[i:i + (1 << 18)].T.astype(np.float64) >= 1):
                return False
        return True
    return bool(np.all(util.exact_matmul(H, np.asarray(ps).T) >= 1))


def _lex_increasing(key, Ms, Kns):
    """True when the columns (key, M, Kn) are strictly increasing in
    lexicographic order -- the canonical order the library's sorted runs
    already return -- so the sort can be skipped.  Each block (key, M, Kn)
    is compared row against row in its (points, h) layout: the first
    differing entry of every still-undecided adjacent pair decides it."""
    n = len(key)
    if n < 2:
        return True
    undecided = np.ones(n - 1, dtype=bool)
    rows = np.arange(n - 1)
    for X in (np.asarray(key)[:, None], np.asarray(Ms).T, np.asarray(Kns).T):
        a, b = X[:-1], X[1:]
        diff = a != b
        first = diff.argmax(axis=1)
        r = np.flatnonzero(undecided & diff[rows, first])
        if np.any(a[r, first[r]] > b[r, first[r]]):
            return False
        undecided[r] = False
        if not undecided.any():
            return True
    return False                                  # equal neighbours: let the sort decide


def _pfvs_from_points(Ms, Kns, Qs, key, kappa, h11, Q, M0min, nonsing=None, verbosity=0):
    """Lattice points -> coni PFVs (pfvs/coni.py:275-412): keep q < Q K_gcd,
    canonical (p, M, Kn) order, expand each point by g = 1..(Q K_gcd - 1)//q
    and the integral K0 of the tadpole, cap K' > 0, drop det(N) == 0, stable
    sort by p.  Ms, Kns are (h11, N).  `nonsing` (per point: 1 det != 0,
    0 det == 0, 2 undecided) comes from the device; undecided or missing
    entries are decided exactly here."""
    Ms, Kns = np.asarray(Ms), np.asarray(Kns)
    Qs, key = np.asarray(Qs), np.asarray(key)
    empty = (np.zeros((0, h11), dtype=np.int64), np.zeros((0, h11), dtype=np.int64),
             np.zeros(0, dtype=np.int64))
    if Qs.shape[0] == 0:
        return empty
    if M0min <= 0:
        raise ValueError("coniZpM requires M0min > 0")
    ns = np.full(Qs.shape[0], 2, dtype=np.int8) if nonsing is None else np.asarray(nonsing).copy()
    Mb, Kb, Qb = util.absmax(Ms), util.absmax(Kns), util.absmax(Qs)
    exact = Ms.dtype == object or Kns.dtype == object or Qs.dtype == object or Q * Kb >= 2**62
    if exact:
        Ms, Kns, Qs = util.as_exact(Ms, Kns, Qs)
    Kperps = Kns[1:]
    K_gcds = np.gcd.reduce(Kperps, axis=0) if not exact else \
        np.array([math.gcd(*[int(x) for x in col]) for col in Kperps.T], dtype=object)
    K_gcds[K_gcds < 1] = 1
    mask = Qs < Q * K_gcds
    if not mask.all():
        Ms, Kns, Qs, K_gcds, key, ns = (Ms[:, mask], Kns[:, mask], Qs[mask], K_gcds[mask],
                                        key[mask], ns[mask])
    if Qs.shape[0] == 0:
        return empty
    if not _lex_increasing(key, Ms, Kns):
        order = np.lexsort(tuple(Kns[::-1]) + tuple(Ms[::-1]) + (key,))
        Ms, Kns, Qs, K_gcds, key, ns = (Ms[:, order], Kns[:, order], Qs[order], K_gcds[order],
                                        key[order], ns[order])
    gcap = (Q * K_gcds - 1) // np.maximum(Qs, 1)
    gmax = int(np.max(gcap))
    Qperp_b = gmax * (Qb + Mb * Kb)
    K0_b = Qperp_b + Q + gmax * Kb
    K_b = max(K0_b, gmax * Kb)
    if not exact and max(Qperp_b, K0_b + len(Qs), h11 * K_b * Mb) >= 2**62:
        exact = True
        Ms, Kns, Qs, K_gcds = util.as_exact(Ms, Kns, Qs, K_gcds)
        gcap = (Q * K_gcds - 1) // np.maximum(Qs, 1)
    natural_K0s = Kns[0]
    Kperps = Kns[1:] // K_gcds
    M0s = Ms[0]
    rawQperps = (Qs + M0s * natural_K0s) // K_gcds
    counts = np.array([int(x) for x in gcap], dtype=np.int64) if exact else gcap.astype(np.int64)
    idx = np.repeat(np.arange(len(Qs)), counts)
    g = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts) + 1
    order = np.lexsort((idx, g, np.asarray(key[idx], dtype=np.int64)))
    idx, g = idx[order], g[order]
    if exact:
        g = g.astype(object)
    Qperps = g * rawQperps[idx]
    M0i = M0s[idx]
    lo = -((Q - Qperps) // M0i)
    up = (Qperps - Q) // M0i
    up = np.minimum(up, -((-natural_K0s[idx] * g) // K_gcds[idx]) - 1)
    num = 1 + up - lo
    sel = num > 0
    num = num[sel]
    if not num.size:
        return empty
    if num.dtype == object:
        num = np.array([int(x) for x in num], dtype=np.int64)
    total = int(np.sum(num))
    K0s = np.repeat(lo[sel], num) + np.arange(total) - np.repeat(np.cumsum(num) - num, num)
    pts = idx[sel]
    new_Ks = np.vstack([K0s, np.repeat(g[sel] * Kperps[:, pts], num, axis=1)])
    new_Ms = np.repeat(Ms[:, pts], num, axis=1)
    new_key = np.repeat(key[pts], num)
    new_ns = np.repeat(ns[pts], num)
    tad = -util.colsum_prod(new_Ks, new_Ms)
    if np.any(tad > Q):
        i = int(np.argmax(tad > Q))
        raise RuntimeError(f"Tadpole violation: -dot(K,M)={tad[i]} > Q={Q}. "
                           f"K={new_Ks[:, i].tolist()}, M={new_Ms[:, i].tolist()}")
    # det((kappa M)[1:, 1:]) != 0: device verdicts, exact decision otherwise
    undecided = np.flatnonzero(new_ns == 2)
    if len(undecided):
        kap = np.asarray(kappa)
        Ns = util.exact_matmul(kap.reshape(h11 * h11, h11), new_Ms[:, undecided])
        Ns = np.asarray(Ns).reshape(h11, h11, -1).transpose(2, 0, 1)[:, 1:, 1:]
        new_ns[undecided] = np.where(util.singular_mask(Ns), 0, 1)
    keep = new_ns == 1
    if verbosity >= 2:
        print(f"{int(np.sum(~keep))}/{len(keep)} 'PFVs' had det(N)=0")
    Ks, Ms, key = new_Ks[:, keep], new_Ms[:, keep], new_key[keep]
    order = np.argsort(key, kind="stable")
    return Ks.T[order], Ms.T[order], key[order]


def _exact_host_path(kappa, Mbasis, p_full, Q, qmax, M0min, max_N):
    """One p-vector the device handed back: the lattice on Python ints
    (pfvs_amd.lattice) and the host search; returns (M, Kn, q) arrays."""
    h = len(p_full)
    d = h - 1
    kap = np.asarray(kappa, dtype=object)
    p = np.asarray(p_full, dtype=object)
    Z = np.tensordot(kap, p, axes=([2], [0]))
    T = Z @ p
    v = [int(x) for x in T @ np.asarray(Mbasis, dtype=object)]
    O = lattice.orthogonal_lattice_rows(v)
    O, _ = lattice.lll_rows(O)
    B = np.asarray(Mbasis, dtype=object) @ np.array(O, dtype=object).T          # h x d
    B = B[:, np.argsort([int(x) != 0 for x in B[0]], kind="stable")]
    ZB = Z @ B
    mat = -(B.T @ ZB)
    l = [int(x) for x in B[0]]
    if any(l):
        V = lattice.unimodular_for(l)
        K, w = lattice.lll_rows(V[1:], gram=[[int(x) for x in r] for r in mat], last=V[0])
        Tm = np.array([list(r) for r in K] + [w], dtype=object).T
        B, ZB = B @ Tm, ZB @ Tm
        mat = Tm.T @ mat @ Tm
        l = [int(x) for x in B[0]]                 # M0 = B[0] c in the new basis: (0, .., 0, g)
    if any(abs(int(x)) >= 2**63 for x in mat.ravel()):
        raise IncompleteSearchError(
            f"p={list(map(int, p_full[1:]))}: ellipsoid matrix entries exceed int64")
    H = lattice.hnf_rows([[int(x) for x in r] for r in ZB[1:]])
    small = [r for r in H if any(r) and all(abs(x) < 2**63 for x in r)]
    big = [r for r in H if not all(abs(x) < 2**63 for x in r)]
    prob = dict(mat=[[int(x) for x in r] for r in mat], Q=Q, qmax=qmax, linvec=l,
                linmin=M0min, strict=True, H=np.array(small, dtype=np.int64).reshape(-1, d), L=big)
    res, st, _ = gpu.kernel_batch([prob], mode="host")
    if st[0] != 0:
        _raise_kernel_status(p_full[1:], int(st[0]))
    pts, qs = res[0]
    if len(pts) > max_N:
        _raise_kernel_status(p_full[1:], -2)
    C = np.asarray(pts, dtype=object).T
    return B @ C, ZB @ C, np.asarray(qs)


def coniZpM(data, ps, Q=None, M0min=13, ellipsoid_dilation=1, use_gcd_lattice=False,
            use_c_lattice=True, low_level_parallelism=False, n_jobs=-1, extra_checks=False,
            extra_lll_reduction=True, device="auto", max_N_pfvs=1_000_000_000,
            return_formal_pfvs=False, verbosity=0):
    """The coni PFVs (Ks, Ms) of each p-vector up to a dilation, grouped by
    p-vector and sorted by (M, K) within one (pfvs/coni.py:860-1168)."""
    if not data.coni:
        raise ValueError("coniZpM only applies to coni contexts. Use ZpM for non-coni PFVs.")
    if len(ps) == 0:
        raise ValueError("ps must be non-empty.")
    ps = np.array(ps)
    if not _in_kahler_cone(data.H_cob, ps):
        raise ValueError("some p-vectors are not in the kahler cone (H_cob@ps.T >= 1 failed)")
    if ellipsoid_dilation <= 0:
        raise ValueError(f"ellipsoid_dilation must be > 0, got {ellipsoid_dilation}.")
    if use_gcd_lattice:
        raise NotImplementedError("use_gcd_lattice (the legacy branch) is not provided")
    kappa = data.kappa_cob
    h11, h21 = data.h11, data.h21
    Mbasis = data.M_lattice()
    if Q is None:
        Q = (h11 + h21 + 2) + 2
    use_gpu = _use_gpu(device, h11, len(ps), use_c_lattice, use_gcd_lattice, extra_lll_reduction)
    p_int = util.as_integral(ps)
    if p_int.dtype == object:
        raise IncompleteSearchError("p-vector entries exceed int64")
    p_full = np.hstack([np.zeros((len(ps), 1), dtype=np.int64), p_int.astype(np.int64)])
    qmax = int(math.floor(ellipsoid_dilation * Q + 1e-7))
    geo = dict(kappa=kappa, Mbasis=Mbasis, Q=Q, dilation=ellipsoid_dilation, M0min=M0min)
    mode = "batched" if use_gpu else "host"
    threads = 0 if n_jobs in (-1, None) else max(1, int(n_jobs))
    try:
        r = gpu.run([geo], p_full, np.zeros(len(ps), np.int32), mode=mode, filter=True,
                    sort=True, host_threads=threads, assemble=True)
    except RuntimeError as e:
        if device == "gpu" or not use_gpu:
            raise
        warnings.warn(f"GPU path failed ({e}); using the CPU", RuntimeWarning, stacklevel=2)
        r = gpu.run([geo], p_full, np.zeros(len(ps), np.int32), mode="host", filter=True,
                    sort=True, host_threads=threads, assemble=True)
    Ks, Ms = assemble(r, ps, p_full, kappa, Mbasis, Q, np.full(len(ps), qmax), M0min,
                      max_N_pfvs, verbosity)
    if return_formal_pfvs:
        raise NotImplementedError("PFV objects are not provided; use (Ks, Ms)")
    return Ks, Ms


def assemble(r, ps, p_full, kappa, Mbasis, Q, qmax, M0min, max_N_pfvs=1_000_000_000,
             verbosity=0):
    """(Ks, Ms) of one geometry's p-vectors from a pipeline result r (rows
    M, Kn, q, nonsing grouped by the p-vectors' indices pidx into ps, and
    pstat per p-vector): the status errors, the exact host path for every
    p-vector with pstat 1 (qmax per p-vector) and the post-processing of
    pfvs/coni.py:275-412, in the reference's order."""
    h = np.asarray(kappa).shape[0]
    pstat = r["pstat"]
    counts = np.bincount(r["pidx"], minlength=len(ps)) if len(r["pidx"]) else np.zeros(len(ps), int)
    for ip in np.flatnonzero(counts > max_N_pfvs):
        _raise_kernel_status(ps[ip], -2)
    for ip in np.flatnonzero(pstat < 0):
        _raise_kernel_status(ps[ip], int(pstat[ip]))
    if verbosity >= 1:
        print(f"found {len(r['q'])} lattice points for {len(ps)} p-vectors "
              f"({int(np.sum(pstat == 1))} handled by the exact host path)...")
    if "pdone" in r:
        # the library finished every p-vector with pdone; the others here
        done = np.asarray(r["pdone"]).astype(bool)
        rest = ~done[r["pidx"]] if len(r["pidx"]) else np.zeros(0, bool)
        pf = r["pfv"]
        pieces = [(pf["K"][:, :h], pf["M"][:, :h], pf["pidx"])]
        if rest.any():
            pieces.append(_pfvs_from_points(r["M"][rest, :h].T, r["Kn"][rest, :h].T, r["q"][rest],
                                            r["pidx"][rest], kappa, h, Q, M0min,
                                            nonsing=r["nonsing"][rest], verbosity=verbosity))
    else:
        pieces = [_pfvs_from_points(r["M"][:, :h].T, r["Kn"][:, :h].T, r["q"], r["pidx"], kappa,
                                    h, Q, M0min, nonsing=r["nonsing"], verbosity=verbosity)]
    for ip in np.flatnonzero(pstat == 1):
        Mv, Kv, qv = _exact_host_path(kappa, Mbasis, p_full[ip], Q, int(qmax[ip]), M0min,
                                      max_N_pfvs)
        pieces.append(_pfvs_from_points(Mv, Kv, qv, np.full(len(qv), ip), kappa, h, Q, M0min,
                                        verbosity=verbosity))
    Ks = np.vstack([pc[0] for pc in pieces])
    Ms = np.vstack([pc[1] for pc in pieces])
    keys = np.concatenate([pc[2] for pc in pieces])
    order = np.argsort(keys, kind="stable")
    return Ks[order], Ms[order]
