"""The GPU backend of the coniZpM search: ctypes bindings of libpfvs_amd.so.

Drop-in for the reference's `pfvs.gpu` module (pfvs/gpu.py:85-232):
`available`, `unavailable_reason`, `backend`, `release`, `max_h11`,
`coni_batch_multi` and `coni_batch` keep their names, arguments, output
order (grouped by p-vector, canonical (M, Kn) order within one), padding to
`max_h11()` columns and the pstat semantics (0 done, 1 redo on the exact
host path).  `run` exposes this library's own entry point: the search form
(plain, batched, host), its tunables, the exact per-point filter and the
counters.  The library is found through PFVS_GPU_LIB or next to this file
(pfvs/gpu.py:57-82).
"""

from __future__ import annotations

import ctypes
import glob
import math
import os

import numpy as np

from . import lattice

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i8p = ctypes.POINTER(ctypes.c_int8)

MODE_PLAIN, MODE_BATCHED, MODE_HOST = 0, 1, 2
_MODES = {"plain": MODE_PLAIN, "batched": MODE_BATCHED, "host": MODE_HOST}

STAT_NAMES = ["nodes", "cand", "leaf", "points", "passes", "ms_prep", "ms_roots", "ms_search",
              "ms_post", "ms_total", "search_launches", "host_p", "queue_peak", "ms_wall",
              "kept", "expansions", "nonsingular", "undecided", "steps", "wait_cycles",
              "wave_cycles", "donated", "fetches", "sum_k", "sum_k_m0", "sum_k_r", "sum_r",
              "sum_s", "sum_leaf", "work", "ms_finish",
              "ms_alloc", "ms_d2h", "attempts"]


class _PfgInput(ctypes.Structure):
    _fields_ = [("n_geo", ctypes.c_int), ("h", _i32p),
                ("Q", _i64p), ("qmax", _i64p), ("linmin", _i64p),
                ("kappa", _i64p), ("Mbasis", _i64p),
                ("n_p", ctypes.c_int64), ("ps", _i64p), ("pgeo", _i32p),
                ("device", ctypes.c_int), ("batch", ctypes.c_int64),
                ("verbose", ctypes.c_int)]


class _PfgOutput(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("M", _i64p), ("Kn", _i64p),
                ("q", _i64p), ("pidx", _i64p), ("pstat", _i8p),
                ("seconds_gpu", ctypes.c_double), ("err", ctypes.c_char * 256)]


class _RunInput(ctypes.Structure):
    _fields_ = [("n_geo", ctypes.c_int), ("h", _i32p),
                ("Q", _i64p), ("qmax", _i64p), ("linmin", _i64p),
                ("kappa", _i64p), ("Mbasis", _i64p),
                ("n_p", ctypes.c_int64), ("ps", _i64p), ("pgeo", _i32p),
                ("device", ctypes.c_int), ("mode", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("budget", ctypes.c_int),
                ("stack_cap", ctypes.c_int), ("host_threads", ctypes.c_int),
                ("filter", ctypes.c_int), ("sort", ctypes.c_int),
                ("keep_points", ctypes.c_int), ("verbose", ctypes.c_int),
                ("assemble", ctypes.c_int)]


class _RunOutput(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("M", _i64p), ("Kn", _i64p),
                ("q", _i64p), ("pidx", _i64p), ("pstat", _i8p), ("pcode", _i32p),
                ("nexp", _i64p), ("nonsing", _i8p),
                ("stats", ctypes.c_double * len(STAT_NAMES)), ("err", ctypes.c_char * 256),
                ("n_pfv", ctypes.c_int64), ("Ks", _i64p), ("Ms", _i64p),
                ("pfv_pidx", _i64p), ("pdone", _i8p)]


class _KernelInput(ctypes.Structure):
    _fields_ = [("n_prob", ctypes.c_int), ("dims", _i32p), ("nh", _i32p), ("nl", _i32p),
                ("has_lin", _i32p), ("strict", _i32p), ("mat", _i64p), ("H", _i64p),
                ("L", _i64p), ("linvec", _i64p), ("Q", _i64p), ("qmax", _i64p),
                ("linmin", _i64p), ("device", ctypes.c_int), ("mode", ctypes.c_int)]


class _KernelOutput(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("pts", _i32p), ("qs", _i64p), ("pidx", _i64p),
                ("status", _i32p), ("nodes", _i64p), ("dmax", ctypes.c_int),
                ("err", ctypes.c_char * 256)]


_lib = None
_lib_error = None
_lib_path = None


def lib_path():
    """Where the library is looked for (PFVS_GPU_LIB, else next to this file)."""
    path = os.environ.get("PFVS_GPU_LIB")
    if path:
        return path
    found = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                          "libpfvs_amd*.so")))
    return found[0] if found else None


def _load():
    global _lib, _lib_error, _lib_path
    if _lib is not None or _lib_error is not None:
        return _lib
    path = lib_path()
    if not path:
        _lib_error = ("the HIP backend was not built (run __graft_entry__.build() or "
                      "python -m pfvs_amd.build)")
        return None
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        _lib_error = f"could not load {path}: {e}"
        return None
    lib.pfg_max_h11.restype = ctypes.c_int
    lib.pfg_backend.restype = ctypes.c_char_p
    lib.pfg_device_count.restype = ctypes.c_int
    lib.pfg_coni_batch.argtypes = [ctypes.POINTER(_PfgInput), ctypes.POINTER(_PfgOutput)]
    lib.pfg_coni_batch.restype = ctypes.c_int
    lib.pfg_output_free.argtypes = [ctypes.POINTER(_PfgOutput)]
    lib.pfg_release.argtypes = [ctypes.c_int]
    lib.pfa_coni_run.argtypes = [ctypes.POINTER(_RunInput), ctypes.POINTER(_RunOutput)]
    lib.pfa_coni_run.restype = ctypes.c_int
    lib.pfa_run_output_free.argtypes = [ctypes.POINTER(_RunOutput)]
    lib.pfa_kernel_batch.argtypes = [ctypes.POINTER(_KernelInput), ctypes.POINTER(_KernelOutput)]
    lib.pfa_kernel_batch.restype = ctypes.c_int
    lib.pfa_kernel_output_free.argtypes = [ctypes.POINTER(_KernelOutput)]
    lib.pfa_device_cus.argtypes = [ctypes.c_int]
    lib.pfa_device_cus.restype = ctypes.c_int
    lib.pfa_arch.restype = ctypes.c_char_p
    lib.pfa_copy_probe.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
    lib.pfa_copy_probe.restype = ctypes.c_int
    lib.pfa_device_clock_khz.argtypes = [ctypes.c_int]
    lib.pfa_device_clock_khz.restype = ctypes.c_int
    lib.pfa_det_filter.argtypes = [ctypes.c_int, _i32p, _i64p, ctypes.c_int64, _i32p, _i64p,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p]
    lib.pfa_det_filter.restype = ctypes.c_int
    lib.pfa_reserve.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_char_p]
    lib.pfa_reserve.restype = ctypes.c_int
    lib.pfa_box_enum.argtypes = [_i32p, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i64p,
                                 ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_char_p]
    lib.pfa_box_enum.restype = ctypes.c_int
    _lib, _lib_path = lib, path
    return lib


def loaded_library_path():
    _load()
    return _lib_path


def require():
    """The loaded library, or RuntimeError naming why it is missing."""
    lib = _load()
    if lib is None:
        raise RuntimeError(_lib_error)
    return lib


def available() -> bool:
    """Whether the HIP backend is built and an AMD GPU is present."""
    lib = _load()
    return lib is not None and lib.pfg_device_count() > 0


def unavailable_reason() -> str:
    lib = _load()
    if lib is None:
        return _lib_error
    if lib.pfg_device_count() == 0:
        return "no HIP device found"
    return ""


def backend() -> str:
    return require().pfg_backend().decode()


def release(device: int = 0) -> None:
    lib = _load()
    if lib is not None:
        lib.pfg_release(int(device))


def box_enum(H, rhs, B, primitive=True, cap=10**8, device=0):
    """The library's p-vector enumeration (pfa_box_enum): the integer p with
    |p|_inf <= B and H p >= rhs, primitive if asked, in lexicographic order.
    Returns (ps (n, ncols) int64 -- the first min(total, cap) --, status
    (0, or 1 when total > cap), total, ms).  device < 0 runs the host code."""
    lib = require()
    H = np.ascontiguousarray(H, dtype=np.int32)
    m, n = H.shape
    rhs = np.ascontiguousarray(np.broadcast_to(np.asarray(rhs, dtype=np.int64), (m,)))
    if H.size and np.abs(H.astype(np.int64)).max() >= 2**31:
        raise ValueError("H entries must fit int32")
    cap = max(int(cap), 0)
    if cap <= 1 << 16:
        out = np.empty((cap, n), dtype=np.int64)
    else:                                       # count first, then hold exactly the rows
        _, _, total, _ = box_enum(H, rhs, B, primitive, 0, device)
        out = np.empty((min(total, cap), n), dtype=np.int64)
    n_out, n_tot, ms = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_double(0)
    err = ctypes.create_string_buffer(256)
    rc = lib.pfa_box_enum(H.ctypes.data_as(_i32p), m, n, _p64(rhs), int(B), int(bool(primitive)),
                          int(device), int(min(int(cap), len(out))), _p64(out), ctypes.byref(n_out),
                          ctypes.byref(n_tot), ctypes.byref(ms), err)
    if rc == 2:
        raise RuntimeError(f"box_enum failed: {err.value.decode(errors='replace')}")
    st = 1 if n_tot.value > cap else 0
    return out[:n_out.value], st, int(n_tot.value), float(ms.value)


def reserve(n_p, n_points, h_max=None, device=0):
    """Grow the library's buffers on `device` for batches of n_p p-vectors
    and n_points points at h11 <= h_max (default max_h11())."""
    lib = require()
    err = ctypes.create_string_buffer(256)
    h_max = max_h11() if h_max is None else int(h_max)
    if lib.pfa_reserve(int(device), int(n_p), int(n_points), h_max, err) != 0:
        raise RuntimeError(f"reserve failed: {err.value.decode(errors='replace')}")


def device_facts(device: int = 0):
    """(compute units, peak engine clock in kHz) of `device`."""
    lib = require()
    return int(lib.pfa_device_cus(device)), int(lib.pfa_device_clock_khz(device))


def copy_probe(nbytes, reps=10, device=0):
    """Mean milliseconds of a plain device copy of nbytes (pfa_copy_probe)."""
    lib = require()
    ms = ctypes.c_double(0.0)
    if lib.pfa_copy_probe(int(device), int(nbytes), int(reps), ctypes.byref(ms)) != 0:
        raise RuntimeError("copy probe failed")
    return ms.value


def max_h11() -> int:
    return require().pfg_max_h11()


def _qmax(Q, dilation):
    """floor(dilation * Q), tolerant of float noise (pfvs/gpu.py:132-133)."""
    return int(math.floor(dilation * Q + 1e-7))


def _p64(a):
    return a.ctypes.data_as(_i64p)


def _pack(geoms, ps, pgeo):
    MH = max_h11()
    n_geo = len(geoms)
    h = np.array([np.asarray(g["kappa"]).shape[0] for g in geoms], dtype=np.int32)
    if n_geo and (h.min() < 3 or h.max() > MH):
        raise ValueError(f"the GPU build supports 3 <= h11 <= {MH}")
    Q = np.array([int(g["Q"]) for g in geoms], dtype=np.int64)
    qmax = np.array([_qmax(int(g["Q"]), g["dilation"]) for g in geoms], dtype=np.int64)
    linmin = np.array([math.ceil(g["M0min"] - 1e-9) for g in geoms], dtype=np.int64)
    kap = np.concatenate([np.ascontiguousarray(g["kappa"], dtype=np.int64).ravel() for g in geoms]) \
        if n_geo else np.zeros(1, dtype=np.int64)
    Mb = np.concatenate([np.ascontiguousarray(g["Mbasis"], dtype=np.int64).ravel() for g in geoms]) \
        if n_geo else np.zeros(1, dtype=np.int64)
    pgeo = np.ascontiguousarray(pgeo, dtype=np.int32)
    n_p = len(pgeo)
    P = np.zeros((max(n_p, 1), MH), dtype=np.int64)
    if isinstance(ps, np.ndarray) and ps.ndim == 2:
        P[:n_p, :ps.shape[1]] = ps
    else:
        for i, p in enumerate(ps):
            P[i, :len(p)] = p
    if n_p and (np.any(pgeo < 0) or np.any(pgeo >= n_geo)):
        raise ValueError("pgeo: geometry index out of range")
    if n_p and np.any(P[:n_p][np.arange(MH)[None, :] >= h[pgeo][:, None]] != 0):
        raise ValueError("a p-vector is longer than its geometry's h11")
    return dict(MH=MH, n_geo=n_geo, h=h, Q=Q, qmax=qmax, linmin=linmin, kap=kap, Mb=Mb,
                pgeo=pgeo, n_p=n_p, P=P)


def coni_batch_multi(geoms, ps, pgeo, device=0, batch=0, verbose=False):
    """The batched pipeline for p-vectors of several geometries at once
    (pfvs/gpu.py:136-211): (M, Kn, q, pidx, pstat, seconds_gpu)."""
    lib = require()
    k = _pack(geoms, ps, pgeo)
    MH, n_p = k["MH"], k["n_p"]
    inp = _PfgInput(k["n_geo"], k["h"].ctypes.data_as(_i32p), _p64(k["Q"]), _p64(k["qmax"]),
                    _p64(k["linmin"]), _p64(k["kap"]), _p64(k["Mb"]), n_p, _p64(k["P"]),
                    k["pgeo"].ctypes.data_as(_i32p), int(device), int(batch), int(bool(verbose)))
    out = _PfgOutput()
    rc = lib.pfg_coni_batch(ctypes.byref(inp), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"GPU pipeline failed: {out.err.decode(errors='replace')}")
    try:
        n = out.n

        def arr(ptr, shape):
            if n == 0:
                return np.zeros(shape, dtype=np.int64)
            return np.ctypeslib.as_array(ptr, shape=shape).copy()
        M, Kn = arr(out.M, (n, MH)), arr(out.Kn, (n, MH))
        q, pidx = arr(out.q, (n,)), arr(out.pidx, (n,))
        pstat = np.ctypeslib.as_array(out.pstat, shape=(n_p,)).copy() if n_p \
            else np.zeros(0, dtype=np.int8)
        secs = out.seconds_gpu
    finally:
        lib.pfg_output_free(ctypes.byref(out))
    return M, Kn, q, pidx, pstat.astype(np.int8), secs


def coni_batch(kappa, Mbasis, ps, Q, dilation, M0min, max_N_out, device=0):
    """One geometry (pfvs/gpu.py:214-232): (M[:, :h], Kn[:, :h], q, pidx,
    pstat); p-vectors with more than max_N_out points get pstat 1."""
    kappa = np.asarray(kappa)
    h = kappa.shape[0]
    ps = np.asarray(ps, dtype=np.int64)
    geoms = [dict(kappa=kappa, Mbasis=Mbasis, Q=Q, dilation=dilation, M0min=M0min)]
    M, Kn, q, pidx, pstat, _ = coni_batch_multi(geoms, ps, np.zeros(len(ps), dtype=np.int32),
                                                device=device)
    counts = np.bincount(pidx, minlength=len(ps)) if len(pidx) else np.zeros(len(ps), dtype=np.int64)
    over = counts > max_N_out
    if np.any(over):
        pstat[over] = 1
        keep = ~over[pidx]
        M, Kn, q, pidx = M[keep], Kn[keep], q[keep], pidx[keep]
    return M[:, :h], Kn[:, :h], q, pidx, pstat


_libc_free = ctypes.CDLL(None).free
_libc_free.argtypes = [ctypes.c_void_p]


class _Malloced:
    """Owner of a buffer the library allocated with malloc: exposes it as
    bytes to numpy and frees it when garbage collected."""

    def __init__(self, addr, nbytes):
        self.addr = addr
        self.__array_interface__ = {"shape": (nbytes,), "typestr": "|u1",
                                    "data": (addr, False), "version": 3}

    def __del__(self):
        _libc_free(self.addr)


def run(geoms, ps, pgeo, mode="batched", device=0, waves_per_cu=0, budget=0, stack_cap=0,
        host_threads=0, filter=True, sort=True, keep_points=True, verbose=False, assemble=False):
    """This library's run call.  Returns a dict with M, Kn (n x max_h11()),
    q, pidx, pstat, pcode, nexp, nonsing (per point) and `stats` (counters
    and stage timings in ms).  With assemble (and filter, sort, keep_points)
    also `pfv` = {K, M (n_pfv x max_h11()), pidx}: coniZpM's PFVs of every
    p-vector with pdone[i] == 1, in the reference's order."""
    lib = require()
    k = _pack(geoms, ps, pgeo)
    MH, n_p = k["MH"], k["n_p"]
    inp = _RunInput(k["n_geo"], k["h"].ctypes.data_as(_i32p), _p64(k["Q"]), _p64(k["qmax"]),
                    _p64(k["linmin"]), _p64(k["kap"]), _p64(k["Mb"]), n_p, _p64(k["P"]),
                    k["pgeo"].ctypes.data_as(_i32p), int(device), _MODES[mode],
                    int(waves_per_cu), int(budget), int(stack_cap), int(host_threads),
                    int(bool(filter)), int(bool(sort)), int(bool(keep_points)), int(bool(verbose)),
                    int(bool(assemble)))
    out = _RunOutput()
    rc = lib.pfa_coni_run(ctypes.byref(inp), ctypes.byref(out))
    if rc != 0:
        msg = out.err.decode(errors="replace")
        lib.pfa_run_output_free(ctypes.byref(out))
        raise RuntimeError(f"GPU pipeline failed: {msg}")
    try:
        n = out.n if keep_points else 0

        def arr(ptr, shape, dt=np.int64):
            if n == 0 or not ptr:
                return np.zeros(shape, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=shape).copy()

        def own(name, rows=None):
            """The library's malloc'd (rows, MH) buffer as an array, without a
            copy; freed with the last array over it."""
            rows = n if rows is None else rows
            ptr = getattr(out, name)
            if rows == 0 or not ptr:
                return np.zeros((rows, MH), dtype=np.int64)
            a = np.asarray(_Malloced(ctypes.addressof(ptr.contents), 8 * rows * MH))
            setattr(out, name, None)
            return a.view(np.int64).reshape(rows, MH)
        res = dict(M=own("M"), Kn=own("Kn"), q=arr(out.q, (n,)),
                   pidx=arr(out.pidx, (n,)), nexp=arr(out.nexp, (n,)),
                   nonsing=arr(out.nonsing, (n,), np.int8))
        res["pstat"] = np.ctypeslib.as_array(out.pstat, shape=(n_p,)).copy() if n_p \
            else np.zeros(0, np.int8)
        res["pcode"] = np.ctypeslib.as_array(out.pcode, shape=(n_p,)).copy() if n_p \
            else np.zeros(0, np.int32)
        res["stats"] = {name: float(out.stats[i]) for i, name in enumerate(STAT_NAMES)}
        if assemble and out.pdone:
            npf = int(out.n_pfv)
            res["pfv"] = dict(K=own("Ks", npf), M=own("Ms", npf),
                              pidx=np.ctypeslib.as_array(out.pfv_pidx, shape=(npf,)).copy()
                              if npf else np.zeros(0, np.int64))
            res["pdone"] = np.ctypeslib.as_array(out.pdone, shape=(n_p,)).copy() if n_p \
                else np.zeros(0, np.int8)
    finally:
        lib.pfa_run_output_free(ctypes.byref(out))
    return res


def _echelon_rows(p):
    """The problem with H in echelon form.  g = gcd(H c) does not change
    under unimodular row operations on H, so a general H (repeated leading
    columns, which the reference's contract accepts: fp_kernel.h:40-44) is
    replaced by the row HNF of all its rows, H and L together: the same cut,
    pruned at least as early.  Rows of the HNF that fit int64 prune, wider
    ones join the gcd at the leaves."""
    d = len(p["mat"])
    H = np.asarray(p.get("H", np.zeros((0, d))), dtype=np.int64).reshape(-1, d)
    lead = [int(np.flatnonzero(r)[0]) for r in H if r.any()]
    if len(lead) == len(set(lead)):
        return p
    rows = [[int(x) for x in r] for r in H] + [[int(x) for x in r] for r in p.get("L", [])]
    Hn = lattice.hnf_rows(rows)
    small = [r for r in Hn if any(r) and all(abs(x) < 2**63 for x in r)]
    big = [r for r in Hn if not all(abs(x) < 2**63 for x in r)]
    return dict(p, H=np.array(small, dtype=np.int64).reshape(-1, d), L=big)


def _reduced_problem(p):
    """(problem in an LLL-reduced basis of its form, T).  With c = T c'
    (T unimodular, columns the reduced vectors): q(c) = c'^T (T^T mat T) c',
    linvec.c = (T^T linvec).c', gcd(H c) = gcd((H T) c'), so the point sets
    correspond one to one with the same q.  An ill-conditioned mat (the FP64
    factor's margin too wide, status 1) is well-conditioned in this basis."""
    d = len(p["mat"])
    mat = [[int(x) for x in r] for r in np.asarray(p["mat"]).tolist()]
    T = lattice.lll_gram_exact(mat)                              # columns = reduced vectors
    matT = lattice.matmul(lattice.matmul(lattice.transpose(T), mat), T)
    rows = [[int(x) for x in r] for r in np.asarray(p.get("H", np.zeros((0, d)))).reshape(-1, d).tolist()]
    rows += [[int(x) for x in r] for r in p.get("L", [])]
    rows = [r for r in lattice.hnf_rows(lattice.matmul(rows, T)) if any(r)] if rows else []
    lin = p.get("linvec")
    q = dict(p, mat=matT,
             linvec=None if lin is None else [sum(T[i][k] * int(lin[i]) for i in range(d)) for k in range(d)],
             H=np.array([r for r in rows if all(abs(x) < 2**63 for x in r)], dtype=object).reshape(-1, d),
             L=[r for r in rows if not all(abs(x) < 2**63 for x in r)])
    if any(abs(x) >= 2**63 for r in matT for x in r) or \
            (q["linvec"] is not None and any(abs(x) >= 2**63 for x in q["linvec"])):
        return None, None
    q["H"] = q["H"].astype(np.int64)
    return q, T


def _retry_reduced(problems, res, status, nodes, mode, device):
    """Problems with status 1 searched again in an LLL-reduced basis; their
    points are mapped back (c = T c') and put in the search's order (last
    coordinate outermost, each ascending), which is lexicographic order on
    the reversed coordinates.  Points beyond int32 give status -8."""
    idx = [k for k in range(len(problems)) if status[k] == 1 and len(problems[k]["mat"]) > 1]
    red = [(k,) + _reduced_problem(problems[k]) for k in idx]
    red = [(k, q, T) for k, q, T in red if q is not None]
    if not red:
        return
    r2, s2, n2 = _kernel_call([q for _, q, _ in red], mode, device)
    for (k, _, T), (pts, qs), s, n in zip(red, r2, s2, n2):
        if s != 0:
            status[k] = s
            continue
        C = np.asarray(pts, dtype=object) @ np.array(T, dtype=object).T if len(pts) else \
            np.zeros((0, len(T)), dtype=object)
        if len(C) and max(abs(int(x)) for x in C.ravel()) >= 2**31:
            status[k] = -8
            res[k] = (np.zeros((0, len(T)), np.int32), np.zeros(0, np.int64))
            continue
        order = sorted(range(len(C)), key=lambda i: tuple(int(x) for x in C[i][::-1]))
        res[k] = (np.array([[int(x) for x in C[i]] for i in order], dtype=np.int32).reshape(-1, len(T)),
                  np.asarray(qs)[order].astype(np.int64))
        status[k] = 0
        nodes[k] = n


def kernel_batch(problems, mode="plain", device=0):
    """Single-problem searches (the conipfv_kernel contract) in one call.

    Each problem is a dict with mat (d x d int), Q, qmax, linvec (or None),
    linmin, strict, H (echelon rows fitting int64) and optionally L (rows
    as Python ints that need more than int64, combined with H's gcd at the
    leaves).  Returns (list of (pts int32 (n, d), qs int64) in depth-first
    order, statuses, node counts (per problem in host mode; 0 on the device)).  H may be any integer matrix (rows
    with repeated leading columns are brought to echelon form first).
    A problem whose FP64 factor margin is too wide (an ill-conditioned
    mat) is searched again in an LLL-reduced basis of its form.  Status 1
    that remains: the problem needs an exact search beyond this one (mat
    above 2^53, or a leaf gcd beyond int128)."""
    problems = [_echelon_rows(p) for p in problems]
    res, status, nodes = _kernel_call(problems, mode, device)
    _retry_reduced(problems, res, status, nodes, mode, device)
    return res, status, nodes


def _kernel_call(problems, mode, device):
    """One pfa_kernel_batch call over problems with H in echelon form."""
    lib = require()
    n = len(problems)
    dims = np.array([len(p["mat"]) for p in problems], dtype=np.int32)
    nh = np.array([len(p.get("H", [])) for p in problems], dtype=np.int32)
    nl = np.array([len(p.get("L", [])) for p in problems], dtype=np.int32)
    has_lin = np.array([p.get("linvec") is not None for p in problems], dtype=np.int32)
    strict = np.array([int(bool(p.get("strict", True))) for p in problems], dtype=np.int32)
    mat = np.concatenate([np.asarray(p["mat"], dtype=np.int64).ravel() for p in problems] + [np.zeros(1, np.int64)])
    H = np.concatenate([np.asarray(p.get("H", np.zeros((0, len(p["mat"])))), dtype=np.int64).ravel()
                        for p in problems] + [np.zeros(1, np.int64)])
    Lw = []
    for p in problems:
        for row in p.get("L", []):
            for x in row:
                x = int(x) & ((1 << 128) - 1)
                lo, hi = x & ((1 << 64) - 1), x >> 64
                Lw += [lo - (1 << 64) if lo >= 1 << 63 else lo, hi - (1 << 64) if hi >= 1 << 63 else hi]
    Larr = np.array(Lw + [0], dtype=np.int64)
    lin = np.concatenate([np.asarray(p["linvec"], dtype=np.int64).ravel() for p in problems
                          if p.get("linvec") is not None] + [np.zeros(1, np.int64)])
    Q = np.array([int(p["Q"]) for p in problems], dtype=np.int64)
    qmax = np.array([int(p["qmax"]) for p in problems], dtype=np.int64)
    linmin = np.array([int(p.get("linmin", 0) or 0) for p in problems], dtype=np.int64)
    inp = _KernelInput(n, dims.ctypes.data_as(_i32p), nh.ctypes.data_as(_i32p),
                       nl.ctypes.data_as(_i32p), has_lin.ctypes.data_as(_i32p),
                       strict.ctypes.data_as(_i32p), _p64(mat), _p64(H), _p64(Larr), _p64(lin),
                       _p64(Q), _p64(qmax), _p64(linmin), int(device), _MODES[mode])
    out = _KernelOutput()
    rc = lib.pfa_kernel_batch(ctypes.byref(inp), ctypes.byref(out))
    if rc != 0:
        msg = out.err.decode(errors="replace")
        raise RuntimeError(f"kernel batch failed: {msg}")
    try:
        m, D = out.n, out.dmax
        pts = np.ctypeslib.as_array(out.pts, shape=(m, D)).copy() if m else np.zeros((0, D), np.int32)
        qs = np.ctypeslib.as_array(out.qs, shape=(m,)).copy() if m else np.zeros(0, np.int64)
        pidx = np.ctypeslib.as_array(out.pidx, shape=(m,)).copy() if m else np.zeros(0, np.int64)
        status = np.ctypeslib.as_array(out.status, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        nodes = np.ctypeslib.as_array(out.nodes, shape=(n,)).copy() if n else np.zeros(0, np.int64)
    finally:
        lib.pfa_kernel_output_free(ctypes.byref(out))
    res = []
    for k in range(n):
        sel = pidx == k
        res.append((pts[sel, :dims[k]], qs[sel]))
    return res, status, nodes


def det_filter(kappas, M, geo=None, device=0):
    """det((kappa M)[1:, 1:]) != 0 for each row of M (the exact filter of
    pfvs/coni.py:397-407): int8 array, 1 nonsingular, 0 singular, 2
    undecided.  kappas: one (h, h, h) int array per geometry; M: (count, h)
    rows (padded with zeros to the largest h); geo: geometry of each row
    (default 0).  device < 0 runs the host code."""
    lib = require()
    hs = np.array([np.asarray(k).shape[0] for k in kappas], dtype=np.int32)
    kap = np.ascontiguousarray(np.concatenate([np.asarray(k, dtype=np.int64).ravel() for k in kappas]))
    hmax = int(hs.max())
    M = np.asarray(M, dtype=np.int64)
    count = M.shape[0]
    Mp = np.zeros((count, hmax), dtype=np.int64)
    Mp[:, :M.shape[1]] = M
    g = np.zeros(count, np.int32) if geo is None else np.ascontiguousarray(geo, dtype=np.int32)
    out = np.zeros(max(count, 1), dtype=np.int8)
    err = ctypes.create_string_buffer(256)
    rc = lib.pfa_det_filter(len(kappas), hs.ctypes.data_as(_i32p), _p64(kap), count,
                            g.ctypes.data_as(_i32p), _p64(Mp), int(device),
                            out.ctypes.data_as(ctypes.c_void_p), err)
    if rc != 0:
        raise RuntimeError("det filter failed: " + err.value.decode(errors="replace"))
    return out[:count]
