"""coniZpM on MI355X: the coni PFVs of each p-vector up to a dilation -- the
drop-in for the reference's `pfvs.coni.coniZpM` (pfvs/coni.py:860-1168):
the same signature, errors and output order.

Flow of one call:

1. checks (a coni context, p-vectors in the Kahler cone, a positive
   dilation) and the device choice: "gpu" (RuntimeError if the HIP backend
   cannot take the call), "cpu" (the library's host build of the same
   preparation and search) or "auto" (the GPU from 256 p-vectors on, when
   it can take the call; PFVS_DEVICE overrides "auto"), as
   pfvs/coni.py:416-443 decide it;
2. one library call for all p-vectors: preparation, search, the exact
   per-point filter (K-integrality count, det((kappa M)[1:,1:]) != 0) and
   the PFV assembly of every p-vector the library can finish; only the
   points that can yield PFVs come back;
3. the rest here: p-vectors handed back with status 1 (an int64/int128
   trap) through the exact host path (Python-int lattice, host search), and
   points the library left (an undecided determinant, values beyond int64)
   through the same assembly with exact decisions first.

A p-vector is never skipped: a kernel status raises IncompleteSearchError
(pfvs/coni.py:446-453).
"""

from __future__ import annotations

import itertools
import math
import os
import warnings

import numpy as np

from . import gpu, lattice, util
from .util import IncompleteSearchError

_GPU_MIN_PS = 256            # "auto" stays on the host below this many p-vectors
_DEVICES = ("auto", "cpu", "gpu")

_STATUS_TEXT = {
    1: "(beyond the exact search's int64/int128 range: an ellipsoid matrix entry above "
       "2^53 or a leaf gcd beyond int128)",
    -2: "(more than max_N_pfvs outputs; increase max_N_pfvs)",
    -8: "(lattice-point coordinates exceed int32)",
    -9: "(the ellipsoid matrix is not positive definite, so its lattice points cannot be "
        "enumerated)",
}


def _incomplete(p, status: int) -> IncompleteSearchError:
    """The error for a p-vector the search could not finish (the message
    names the kernel status, as the reference's does)."""
    return IncompleteSearchError(f"p={np.asarray(p).tolist()}: conipfv_kernel returned status "
                                 f"{status} {_STATUS_TEXT.get(status, '')}")


def _raise_kernel_status(p, status: int):
    raise _incomplete(p, status)


def _gpu_obstacles(h11, use_c_lattice, use_gcd_lattice, extra_lll):
    """Why the GPU cannot take this call (empty when it can)."""
    if not gpu.available():
        yield gpu.unavailable_reason()
    elif not 3 <= h11 <= gpu.max_h11():
        yield f"h11 = {h11} is outside the GPU build's 3..{gpu.max_h11()}"
    if use_gcd_lattice or not (use_c_lattice and extra_lll):
        yield ("the GPU path needs use_c_lattice=True, use_gcd_lattice=False, "
               "extra_lll_reduction=True")


def _use_gpu(device, h11, n_ps, use_c_lattice=True, use_gcd_lattice=False, extra_lll=True):
    """True to run on the GPU.  'gpu' insists (RuntimeError naming what is
    missing), 'cpu' never, 'auto' (or what PFVS_DEVICE says instead of it)
    when the GPU can take the call and there are at least 256 p-vectors."""
    if device not in _DEVICES:
        raise ValueError(f'device must be "auto", "cpu" or "gpu", got {device!r}')
    choice = device
    if choice == "auto":
        choice = os.environ.get("PFVS_DEVICE", "auto").lower()
        if choice not in _DEVICES:
            raise ValueError(f'PFVS_DEVICE must be "auto", "cpu" or "gpu", got {choice!r}')
    if choice == "cpu":
        return False
    obstacles = list(_gpu_obstacles(h11, use_c_lattice, use_gcd_lattice, extra_lll))
    if choice == "gpu" and obstacles:
        raise RuntimeError("device='gpu' unavailable: " + "; ".join(obstacles))
    return not obstacles and (choice == "gpu" or n_ps >= _GPU_MIN_PS)


def _in_kahler_cone(H, ps) -> bool:
    """H p >= 1 for every row p of ps, decided on integers: int64 products
    when rows * max|H| * max|p| stays below 2^62, Python ints otherwise."""
    H, ps = util.as_integral(H), util.as_integral(np.atleast_2d(ps))
    if H.dtype != object and ps.dtype != object and \
            H.shape[1] * util.magnitude(H) * util.magnitude(ps) < 2**62:
        step = 1 << 16
        return all(bool(np.all(H @ ps[i:i + step].T >= 1)) for i in range(0, len(ps), step))
    return bool(np.all(util.to_pyints(H) @ util.to_pyints(ps).T >= 1))


# ------------------------------------------------------- post-processing
def _verdicts(Ms, kappa, h11, nonsing):
    """det((kappa M)[1:, 1:]) != 0 per point (1 / 0): the device's verdicts
    where given, the library's exact modular test (host code) for int64 M,
    and a fraction-free determinant on Python ints for what is left."""
    n = Ms.shape[1]
    v = np.full(n, 2, dtype=np.int8) if nonsing is None else np.asarray(nonsing, np.int8).copy()
    todo = np.flatnonzero(v == 2)
    if todo.size and Ms.dtype != object:
        v[todo] = gpu.det_filter([kappa], Ms[:, todo].T, device=-1)
        todo = np.flatnonzero(v == 2)
    if todo.size:
        kap = util.to_pyints(kappa).reshape(h11 * h11, h11)
        N = (kap @ util.to_pyints(Ms[:, todo])).reshape(h11, h11, -1).transpose(2, 0, 1)
        v[todo] = np.where(util.singular_mask(N[:, 1:, 1:]), 0, 1)
    return v


def _cdiv(a: int, b: int) -> int:
    return -((-a) // b)


def _rows_of_p(points, Q, M0min):
    """The PFV rows of one p-vector from its points (M, Kn, q, verdict) with
    Python integers (pfvs/coni.py:313-412 restated per point): a point with
    Kgcd = gcd(Kn[1:]) (0 -> 1) and q < Q Kgcd is expanded over
    g = 1 .. (Q Kgcd - 1) // q and the K0 on the tadpole with K' > 0; rows by
    g, then the points in (M, Kn) order, then K0; the tadpole bound is
    checked on every row and singular points' rows are dropped last."""
    staged = []
    for rank, (M, Kn, q, ok) in enumerate(sorted(points, key=lambda t: (t[0], t[1]))):
        kg = math.gcd(*Kn[1:]) or 1
        if q >= Q * kg:
            continue
        if M[0] < M0min or q <= 0:
            raise RuntimeError(f"lattice point outside the search's cuts: M={list(M)}, q={q}")
        raw, rest = divmod(q + M[0] * Kn[0], kg)
        if rest:
            raise RuntimeError(f"q + M0 K0 not divisible by gcd(K[1:]) at M={list(M)}")
        Kperp = [x // kg for x in Kn[1:]]
        for g in range(1, (Q * kg - 1) // q + 1):
            top = g * raw - Q
            for K0 in range(_cdiv(top, M[0]), min(top // M[0], _cdiv(Kn[0] * g, kg) - 1) + 1):
                staged.append(((g, rank, K0), [K0] + [g * x for x in Kperp], list(M), ok))
    staged.sort(key=lambda t: t[0])
    rows = []
    for _, K, M, ok in staged:
        tad = -sum(a * b for a, b in zip(K, M))
        if tad > Q:
            raise RuntimeError(f"Tadpole violation: -dot(K,M)={tad} > Q={Q}. K={K}, M={M}")
        if ok == 1:
            rows.append((K, M))
    return rows


def _pfvs_from_points(Ms, Kns, Qs, key, kappa, h11, Q, M0min, nonsing=None, verbosity=0):
    """Lattice points -> coni PFVs: (Ks, Ms, keys), rows grouped by key
    (stably, keys ascending) in the reference's order within a key.
    Ms, Kns: (h11, N) integer arrays; Qs: q per point; key: the p-vector
    index of each point; nonsing: the device's det verdicts (1 / 0 / 2
    undecided) or None.  The library's assembly (pfa_assemble) takes every
    key it can finish in int64; the rest runs on Python ints."""
    Ms, Kns = util.as_integral(Ms), util.as_integral(Kns)
    Qs, key = util.as_integral(np.asarray(Qs)), np.asarray(key, dtype=np.int64)
    none = (np.zeros((0, h11), np.int64), np.zeros((0, h11), np.int64), np.zeros(0, np.int64))
    if len(key) == 0:
        return none
    if M0min <= 0:
        raise ValueError("coniZpM requires M0min > 0")
    verdict = _verdicts(Ms, kappa, h11, nonsing)
    keys, dense = np.unique(key, return_inverse=True)
    done = np.zeros(len(keys), dtype=bool)
    pieces = []
    if Ms.dtype != object and Kns.dtype != object and Qs.dtype != object:
        Ks_n, Ms_n, k_n, done_n = gpu.assemble_points(Ms.T, Kns.T, Qs, dense, verdict, h11, Q,
                                                      len(keys))
        done = done_n.astype(bool)
        pieces.append((Ks_n, Ms_n, keys[k_n]))
    left = np.flatnonzero(~done[dense])
    if left.size:
        Mt, Kt = util.to_pyints(Ms[:, left]).T, util.to_pyints(Kns[:, left]).T
        qs = [int(x) for x in np.asarray(Qs)[left]]
        dl, vl = dense[left], verdict[left]
        order = np.argsort(dl, kind="stable")
        for d, grp in itertools.groupby(order, key=lambda i: dl[i]):
            pts = [(tuple(Mt[i]), tuple(Kt[i]), qs[i], int(vl[i])) for i in grp]
            rows = _rows_of_p(pts, int(Q), int(M0min))
            if rows:
                K = np.array([r[0] for r in rows], dtype=object)
                M = np.array([r[1] for r in rows], dtype=object)
                pieces.append((K, M, np.full(len(rows), keys[d], dtype=np.int64)))
    if verbosity >= 2:
        print(f"{int(np.sum(verdict == 0))}/{len(verdict)} points had det(N)=0")
    if not pieces:
        return none
    Ks = np.vstack([p[0] for p in pieces])
    Ms_out = np.vstack([p[1] for p in pieces])
    ks = np.concatenate([p[2] for p in pieces])
    order = np.argsort(ks, kind="stable")
    Ks, Ms_out, ks = Ks[order], Ms_out[order], ks[order]
    if Ks.dtype == object and util.fits_int64(Ks) and util.fits_int64(Ms_out):
        Ks, Ms_out = Ks.astype(np.int64), Ms_out.astype(np.int64)
    return Ks, Ms_out, ks


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


def _check_inputs(data, ps, ellipsoid_dilation):
    """The reference's argument errors (pfvs/coni.py:930-941), in its order."""
    if not data.coni:
        raise ValueError("coniZpM only applies to coni contexts. Use ZpM for non-coni PFVs.")
    if len(ps) == 0:
        raise ValueError("ps must be non-empty.")
    if not _in_kahler_cone(data.H_cob, np.asarray(ps)):
        raise ValueError("some p-vectors are not in the kahler cone (H_cob@ps.T >= 1 failed)")
    if not ellipsoid_dilation > 0:
        raise ValueError(f"ellipsoid_dilation must be > 0, got {ellipsoid_dilation}.")


def coniZpM(data, ps, Q=None, M0min=13, ellipsoid_dilation=1, use_gcd_lattice=False,
            use_c_lattice=True, low_level_parallelism=False, n_jobs=-1, extra_checks=False,
            extra_lll_reduction=True, device="auto", max_N_pfvs=1_000_000_000,
            return_formal_pfvs=False, verbosity=0):
    """The coni PFVs (Ks, Ms) of each p-vector up to a dilation, grouped by
    p-vector and sorted by (M, K) within one (pfvs/coni.py:860-1168)."""
    _check_inputs(data, ps, ellipsoid_dilation)
    if use_gcd_lattice:
        raise NotImplementedError("use_gcd_lattice (the legacy branch) is not provided")
    if return_formal_pfvs:
        raise NotImplementedError("PFV objects are not provided; use (Ks, Ms)")
    ps = np.array(ps)
    kappa, h11 = data.kappa_cob, data.h11
    Mbasis = data.M_lattice()
    Q = h11 + data.h21 + 4 if Q is None else Q
    on_gpu = _use_gpu(device, h11, len(ps), use_c_lattice, use_gcd_lattice, extra_lll_reduction)
    p_int = util.as_integral(ps)
    if p_int.dtype == object:
        raise IncompleteSearchError("p-vector entries exceed int64")
    p_full = np.concatenate([np.zeros((len(ps), 1), np.int64), p_int.astype(np.int64)], axis=1)
    qmax = gpu.qmax_of(Q, ellipsoid_dilation)
    geo = {"kappa": kappa, "Mbasis": Mbasis, "Q": Q, "dilation": ellipsoid_dilation,
           "M0min": M0min}
    one_geo = np.zeros(len(ps), np.int32)
    threads = 0 if n_jobs in (-1, None) else max(1, int(n_jobs))
    call = dict(filter=True, sort=True, keep_points="survivors", assemble=True,
                host_threads=threads)
    try:
        r = gpu.run([geo], p_full, one_geo, mode="batched" if on_gpu else "host", **call)
    except RuntimeError as e:
        if not on_gpu or device == "gpu":
            raise
        # "auto": a device that fails (memory taken by other processes, a
        # lost device) leaves the call to the host build
        warnings.warn(f"GPU path failed ({e}); using the CPU", RuntimeWarning, stacklevel=2)
        r = gpu.run([geo], p_full, one_geo, mode="host", **call)
    return assemble(r, ps, p_full, kappa, Mbasis, Q, np.full(len(ps), qmax), M0min, max_N_pfvs,
                    verbosity)


def assemble(r, ps, p_full, kappa, Mbasis, Q, qmax, M0min, max_N_pfvs=1_000_000_000,
             verbosity=0, with_keys=False):
    """(Ks, Ms) of one geometry's p-vectors from a pipeline result r (its
    points grouped by the p-vector indices pidx into ps; pstat and pcount
    per p-vector): the status errors, the exact host path for every
    p-vector with pstat 1 (qmax per p-vector) and the post-processing of
    pfvs/coni.py:275-412 for the points the library did not finish, in the
    reference's order."""
    h = np.asarray(kappa).shape[0]
    pstat = np.asarray(r["pstat"])
    counts = r.get("pcount")
    if counts is None or len(counts) != len(ps):
        counts = np.bincount(r["pidx"], minlength=len(ps)) if len(r["pidx"]) else np.zeros(len(ps))
    for ip in np.flatnonzero((counts > max_N_pfvs) & (pstat == 0)):
        _raise_kernel_status(ps[ip], -2)
    for ip in np.flatnonzero(pstat < 0):
        _raise_kernel_status(ps[ip], int(pstat[ip]))
    if verbosity >= 1:
        print(f"found {int(np.sum(counts))} lattice points for {len(ps)} p-vectors "
              f"({int(np.sum(pstat == 1))} handled by the exact host path)...")
    pieces = []
    pidx = np.asarray(r["pidx"])
    if "pdone" in r:
        done = np.asarray(r["pdone"]).astype(bool)
        pieces.append((r["pfv"]["K"][:, :h], r["pfv"]["M"][:, :h], r["pfv"]["pidx"]))
        rest = ~done[pidx] if len(pidx) else np.zeros(0, bool)
    else:
        rest = np.ones(len(pidx), dtype=bool)
    if rest.any():
        pieces.append(_pfvs_from_points(r["M"][rest, :h].T, r["Kn"][rest, :h].T, r["q"][rest],
                                        pidx[rest], kappa, h, Q, M0min,
                                        nonsing=r["nonsing"][rest], verbosity=verbosity))
    for ip in np.flatnonzero(pstat == 1):
        Mv, Kv, qv = _exact_host_path(kappa, Mbasis, p_full[ip], Q, int(qmax[ip]), M0min,
                                      max_N_pfvs)
        pieces.append(_pfvs_from_points(Mv, Kv, qv, np.full(len(qv), ip), kappa, h, Q, M0min,
                                        verbosity=verbosity))
    Ks = np.vstack([p[0] for p in pieces] or [np.zeros((0, h), np.int64)])
    Ms = np.vstack([p[1] for p in pieces] or [np.zeros((0, h), np.int64)])
    keys = np.concatenate([p[2] for p in pieces] or [np.zeros(0, np.int64)])
    order = np.argsort(keys, kind="stable")
    if with_keys:
        return Ks[order], Ms[order], keys[order]
    return Ks[order], Ms[order]


# ------------------------------------------------------------------ coniZp
def _grid_up(D):
    """D rounded up onto the grid ceil(2^(k/16)) (at least D itself), so
    that nearby dilation bounds share one geometry entry of the batched
    search (pfvs/coni.py:650-656)."""
    D = np.maximum(np.asarray(D, dtype=np.int64), 1)
    g = np.ceil(2.0 ** (np.ceil(16 * np.log2(D)) / 16)).astype(np.int64)
    return np.maximum(g, D)


def _canonical(Ks, Ms, keys):
    """PFVs without repeats, sorted by p-vector index, then M, then K
    (pfvs/coni.py:489-502)."""
    keys = np.asarray(keys, dtype=np.int64)
    if len(keys) == 0:
        return Ks, Ms, keys
    h = Ms.shape[1]
    if Ks.dtype == object or Ms.dtype == object:
        rows = sorted({(int(k), tuple(map(int, m)), tuple(map(int, kk)))
                       for k, kk, m in zip(keys, Ks, Ms)})
        return (np.array([r[2] for r in rows], dtype=object),
                np.array([r[1] for r in rows], dtype=object),
                np.array([r[0] for r in rows], dtype=np.int64))
    A = np.unique(np.hstack([keys[:, None], Ms, Ks]), axis=0)
    return A[:, 1 + h:], A[:, 1:1 + h], A[:, 0]


def _dilations(kappa, p_full, Ks, Ms, keys):
    """The dilation delta of each PFV (K, M) of the direction p_full[key], as
    exact (numerator, denominator) lists: (Z M)_r = delta K_r with Z =
    kappa . p, so delta = gcd((Z M)_r) / gcd(K_r)."""
    kap = util.to_pyints(np.asarray(kappa))
    h = len(kap)
    num, den = [], []
    Zs = {}
    for K, M, k in zip(Ks, Ms, keys):
        k = int(k)
        if k not in Zs:
            p = [int(x) for x in p_full[k]]
            Zs[k] = [[sum(kap[a][b][c] * p[c] for c in range(h)) for b in range(h)] for a in range(h)]
        Z = Zs[k]
        M = [int(x) for x in M]
        ZM = [sum(Z[a][b] * M[b] for b in range(h)) for a in range(1, h)]
        num.append(math.gcd(*ZM))
        den.append(math.gcd(*[int(x) for x in K[1:]]))
    return num, den


def coniZpK(data, ps, D0, Q=None, M0min=13, n_jobs=-1, max_N_pfvs=1_000_000_000,
            return_formal_pfvs=False, verbosity=0):
    """The coni PFVs of each direction with dilation delta > D0
    (pfvs/coni.py:554-652): with coniZpM(..., ellipsoid_dilation=D0) this
    gives every coni PFV of the direction.  The lattice points come from the
    library's short-vector enumeration (pfa_coni_zpk, host threads: every
    short K_r of Lambda with K_r^T S K_r < Q / D0 and the points of each);
    a direction it cannot handle exactly is searched by coniZpM up to its
    dilation bound instead, and a direction without a bound raises
    IncompleteSearchError.  Returns (Ks, Ms) grouped by p-vector in the
    order of ps, sorted by (M, K) within one, each PFV once."""
    from .dilation import coni_dilation_bounds
    _check_inputs(data, ps, 1)
    if int(D0) != D0 or D0 < 1:
        raise ValueError(f"D0 must be a positive integer, got {D0}")
    if return_formal_pfvs:
        raise NotImplementedError("PFV objects are not provided; use (Ks, Ms)")
    D0 = int(D0)
    ps = np.array(ps)
    kappa, Mbasis, h11 = data.kappa_cob, data.M_lattice(), data.h11
    Q = h11 + data.h21 + 4 if Q is None else Q
    p_int = util.as_integral(ps)
    if p_int.dtype == object:
        raise IncompleteSearchError("p-vector entries exceed int64")
    p_full = np.concatenate([np.zeros((len(ps), 1), np.int64), p_int.astype(np.int64)], axis=1)
    threads = 0 if n_jobs in (-1, None) else max(1, int(n_jobs))
    linmin = math.ceil(M0min - 1e-9)
    M, Kn, q, pidx, st = gpu.coni_zpk_points(kappa, Mbasis, p_full, Q, linmin, D0,
                                             max_out=min(int(max_N_pfvs), 10**8), threads=threads)
    pieces = [_pfvs_from_points(M.T, Kn.T, q, pidx, kappa, h11, Q, M0min)]
    bad = np.flatnonzero(st != 0)
    if verbosity >= 1:
        print(f"ZpK: {len(q)} lattice points for {len(ps)} p-vectors ({len(bad)} searched by "
              f"coniZpM to their bound instead)")
    if len(bad):
        for i, b in zip(bad, coni_dilation_bounds(p_int[bad], kappa, Q)):
            if b is None:
                raise IncompleteSearchError(
                    f"p={ps[i].tolist()}: its PFVs above D0 cannot be enumerated (s >= 0, or S "
                    f"not positive definite on the lattice orthogonal to p)")
            K1, M1 = coniZpM(data, ps[i][None], Q=Q, M0min=M0min,
                             ellipsoid_dilation=max(math.ceil(b), D0), device="cpu", n_jobs=1,
                             max_N_pfvs=max_N_pfvs)
            pieces.append((K1, M1, np.full(len(K1), i, dtype=np.int64)))
    pieces = [pc for pc in pieces if len(pc[2])]
    if not pieces:
        z = np.zeros((0, h11), np.int64)
        return z, z.copy()
    Ks = np.vstack([pc[0] for pc in pieces])
    Ms = np.vstack([pc[1] for pc in pieces])
    keys = np.concatenate([pc[2] for pc in pieces])
    num, den = _dilations(kappa, p_full, Ks, Ms, keys)
    above = np.array([a > D0 * b for a, b in zip(num, den)], dtype=bool)
    Ks, Ms, _ = _canonical(Ks[above], Ms[above], keys[above])
    return Ks, Ms


def _zpm_batched(data, ps, p_full, idx, dil, Q, M0min, device, threads, max_N_pfvs, verbosity):
    """coniZpM on the directions idx, each at its own dilation dil[idx], in one
    batched search (one geometry entry per distinct dilation): (Ks, Ms, keys)
    with keys indexing ps."""
    kappa, Mbasis, h11 = data.kappa_cob, data.M_lattice(), data.h11
    idx = np.asarray(idx, dtype=np.int64)
    if not len(idx):
        z = np.zeros((0, h11), np.int64)
        return z, z.copy(), np.zeros(0, np.int64)
    Ds, pgeo = np.unique(np.asarray(dil, dtype=np.float64)[idx], return_inverse=True)
    geoms = [{"kappa": kappa, "Mbasis": Mbasis, "Q": Q, "dilation": float(D), "M0min": M0min}
             for D in Ds]
    qmax = np.array([gpu.qmax_of(Q, float(dil[i])) for i in idx], dtype=np.int64)
    on_gpu = _use_gpu(device, h11, len(idx))
    call = dict(filter=True, sort=True, keep_points="survivors", assemble=True,
                host_threads=threads)
    pgeo = pgeo.astype(np.int32)
    try:
        r = gpu.run(geoms, p_full[idx], pgeo, mode="batched" if on_gpu else "host", **call)
    except RuntimeError as e:
        if not on_gpu or device == "gpu":
            raise
        warnings.warn(f"GPU path failed ({e}); using the CPU", RuntimeWarning, stacklevel=3)
        r = gpu.run(geoms, p_full[idx], pgeo, mode="host", **call)
    Ks, Ms, keys = assemble(r, ps[idx], p_full[idx], kappa, Mbasis, Q, qmax, M0min, max_N_pfvs,
                            verbosity, with_keys=True)
    return Ks, Ms, idx[np.asarray(keys, dtype=np.int64)]


def _calibrate(data, ps, p_full, b, has, Q, M0min, device, threads, n_sample=64):
    """A CostModel measured on a seeded sample of the bounded directions:
    coniZpM's time per direction at the split candidates and at the upper
    bounds, and coniZpK's above each split candidate (the quartiles of the
    bounds)."""
    import time
    from .dilation import CostModel
    idx = np.flatnonzero(has)
    s = np.sort(np.random.default_rng(0).choice(idx, min(len(idx), n_sample), replace=False))
    bs = b[s]
    D0s = sorted({max(1, int(x)) for x in np.quantile(bs, [0.25, 0.5, 0.75])})
    Ds = sorted(set(D0s) | {max(1, int(x)) for x in np.quantile(bs, [0.9, 1.0])})
    dil = np.zeros(len(ps))
    _zpm_batched(data, ps, p_full, s[:1], np.full(len(ps), float(Ds[0])), Q, M0min, device,
                 threads, 10**9, 0)                       # (warm-up)
    G = []
    for D in Ds:
        dil[:] = D
        t = time.perf_counter()
        _zpm_batched(data, ps, p_full, s, dil, Q, M0min, device, threads, 10**9, 0)
        G.append((time.perf_counter() - t) / len(s))
    linmin = math.ceil(M0min - 1e-9)
    K = []
    for D0 in D0s:
        t = time.perf_counter()
        gpu.coni_zpk_points(data.kappa_cob, data.M_lattice(), p_full[s], Q, linmin, D0,
                            threads=threads)
        K.append((time.perf_counter() - t) / len(s))
    if len(Ds) == 1:
        Ds, G = Ds[:1], G[:1]
    return CostModel(Ds, G, D0s, K, shared=not _use_gpu(device, data.h11, len(s)))


def coniZp(data, ps, Q=None, M0min=13, fallback_dilation=1, n_jobs=-1, device="auto",
           cost_model=None, max_N_pfvs=1_000_000_000, return_formal_pfvs=False, verbosity=0,
           max_dilation=None):
    """Every coni PFV of each p-vector's direction, at any dilation
    (pfvs/coni.py:1171-1223): (Ks, Ms, complete).

    Each direction with a dilation bound b (pfvs_amd.dilation: every PFV of
    the direction has dilation < b) is searched by coniZpM at ceil(b)
    rounded up onto a 2^(1/16) grid; all directions go through one batched
    search with one geometry entry per distinct dilation.  A direction
    without a bound is searched at fallback_dilation and reported
    incomplete.  cost_model (a dilation.CostModel, or "calibrate" to
    measure one on a sample of the directions first) routes the directions
    as the reference's exhaustive search does (pfvs/coni.py:784-855): a
    direction with b <= max(b*, D0) goes to its bound, the others are split
    at D0 -- coniZpM up to D0 in the same batched search, coniZpK above D0
    on host threads while it runs -- and a direction coniZpK leaves
    undecided is searched to its bound afterwards.  Without a cost model
    every bounded direction is searched to its bound (the same PFVs).
    max_dilation (optional) caps the search: a direction whose bound
    exceeds it is searched to max_dilation and reported incomplete.
    Output as coniZpM's rows, sorted by p-vector, then (M, K), each PFV
    once; complete[i] whether direction i was searched exhaustively."""
    from concurrent.futures import ThreadPoolExecutor
    from .dilation import coni_dilation_bound_ceils
    if fallback_dilation <= 0:
        raise ValueError(f"fallback_dilation must be > 0, got {fallback_dilation}.")
    _check_inputs(data, ps, 1)
    if return_formal_pfvs:
        raise NotImplementedError("PFV objects are not provided; use (Ks, Ms)")
    if isinstance(cost_model, str) and cost_model != "calibrate":
        raise ValueError(f"cost_model must be a CostModel, 'calibrate' or None, got {cost_model!r}")
    ps = np.array(ps)
    kappa, h11 = data.kappa_cob, data.h11
    Mbasis = data.M_lattice()
    Q = h11 + data.h21 + 4 if Q is None else Q
    p_int = util.as_integral(ps)
    if p_int.dtype == object:
        raise IncompleteSearchError("p-vector entries exceed int64")
    p_full = np.concatenate([np.zeros((len(ps), 1), np.int64), p_int.astype(np.int64)], axis=1)
    threads = 0 if n_jobs in (-1, None) else max(1, int(n_jobs))
    b = coni_dilation_bound_ceils(p_int, kappa, Q, n_jobs=threads or -1)
    has = b > 0
    dil = np.where(has, _grid_up(np.where(has, b, 1)), 0).astype(np.float64)
    dil[~has] = fallback_dilation
    complete = has.copy()
    capped = np.zeros(len(ps), dtype=bool)
    if max_dilation is not None:
        capped = has & (b > max_dilation)
        dil[capped] = max_dilation
        complete &= ~capped
    split = np.zeros(len(ps), dtype=bool)
    D0 = 0
    if cost_model is not None and np.any(has & ~capped):
        if isinstance(cost_model, str):
            cost_model = _calibrate(data, ps, p_full, b, has & ~capped, Q, M0min, device, threads)
        b_star, D0, _ = cost_model.route(np.where(has & ~capped, b, np.inf))
        D0 = int(D0)
        split = has & ~capped & (b > max(b_star, D0))
        dil[split] = D0
    idx_k = np.flatnonzero(split)
    if verbosity >= 1:
        print(f"coniZp: {len(ps)} p-vectors, {int(np.sum(~has))} without a bound (searched at "
              f"dilation {fallback_dilation}), {len(idx_k)} split at D0 = {D0}, the rest to their "
              f"bound ({len(np.unique(dil))} distinct dilations up to {dil.max():g})")
    linmin = math.ceil(M0min - 1e-9)
    with ThreadPoolExecutor(1) as ex:
        fut = ex.submit(gpu.coni_zpk_points, kappa, Mbasis, p_full[idx_k], Q, linmin, D0,
                        min(int(max_N_pfvs), 10**8), threads) if len(idx_k) else None
        pieces = [_zpm_batched(data, ps, p_full, np.arange(len(ps)), dil, Q, M0min, device,
                               threads, max_N_pfvs, verbosity)]
        zk = fut.result() if fut is not None else None
    if zk is not None:
        M, Kn, q, pidx, st = zk
        pieces.append(_pfvs_from_points(M.T, Kn.T, q, idx_k[pidx], kappa, h11, Q, M0min))
        redo = idx_k[st != 0]
        if len(redo):
            dil2 = np.zeros(len(ps))
            dil2[redo] = _grid_up(b[redo])
            pieces.append(_zpm_batched(data, ps, p_full, redo, dil2, Q, M0min, device, threads,
                                       max_N_pfvs, verbosity))
        if verbosity >= 1:
            print(f"coniZp: coniZpK gave {len(q)} lattice points, {len(redo)} split directions "
                  f"searched to their bound instead")
    pieces = [pc for pc in pieces if len(pc[2])]
    if not pieces:
        z = np.zeros((0, h11), np.int64)
        return z, z.copy(), complete
    Ks = np.vstack([pc[0] for pc in pieces])
    Ms = np.vstack([pc[1] for pc in pieces])
    keys = np.concatenate([pc[2] for pc in pieces])
    Ks, Ms, _ = _canonical(Ks, Ms, keys)
    return Ks, Ms, complete
