"""ctypes binding of libpfvs_amd.so, the MI355X coniZpM search.

Two layers:

- the library's own calls: `run` (the whole per-p-vector pipeline in one
  of four search forms, with the exact post-processing filter and PFV
  assembly on request), `kernel_batch` (single-problem searches),
  `det_filter`, `assemble_points`, `box_enum`, `reserve` and device facts;
- the names of the reference's `pfvs.gpu` module (pfvs/gpu.py:85-232),
  built on `run`: `available`, `unavailable_reason`, `backend`, `release`,
  `max_h11`, `coni_batch_multi` and `coni_batch`, with the reference's
  arguments and output contract (points grouped by p-vector ascending, in
  (M, Kn) order within one, rows padded to `max_h11()` columns, pstat 0
  done / 1 for the exact host path).

The library is found through PFVS_GPU_LIB or next to this file, as the
reference looks for its own (pfvs/gpu.py:57-82).
"""

from __future__ import annotations

import ctypes
import glob
import math
import os
from dataclasses import dataclass

import numpy as np

from . import lattice

_c_i8, _c_i32, _c_i64, _c_f64 = ctypes.c_int8, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_P8, _P32, _P64 = (ctypes.POINTER(t) for t in (_c_i8, _c_i32, _c_i64))

MODES = {"plain": 0, "batched": 1, "host": 2, "lane": 3}
KEEP = {False: 0, True: 1, "all": 1, "survivors": 2, 0: 0, 1: 1, 2: 2}

# pfa_run_output.stats, in the order of the PFA_ST_* enum
STAT_NAMES = ("nodes cand leaf points passes ms_prep ms_roots ms_search ms_post ms_total "
              "search_launches host_p queue_peak ms_wall kept expansions nonsingular undecided "
              "steps wait_cycles wave_cycles donated fetches sum_k sum_k_m0 sum_k_r sum_r sum_s "
              "sum_leaf work ms_finish ms_alloc ms_d2h attempts ms_upload waves "
              "survivors graphs graph_hits graph_captures prep_work prep_work_max").split()


def _struct(name, spec):
    """A ctypes Structure from (field, type) pairs (the C declarations of
    include/pfvs_amd.h, field for field)."""
    return type(name, (ctypes.Structure,), {"_fields_": [tuple(f) for f in spec]})


_RunIn = _struct("pfa_run_input", [
    ("n_geo", ctypes.c_int), ("h", _P32), ("Q", _P64), ("qmax", _P64), ("linmin", _P64),
    ("kappa", _P64), ("Mbasis", _P64), ("n_p", _c_i64), ("ps", _P64), ("pgeo", _P32),
    ("device", ctypes.c_int), ("mode", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
    ("budget", ctypes.c_int), ("stack_cap", ctypes.c_int), ("host_threads", ctypes.c_int),
    ("filter", ctypes.c_int), ("sort", ctypes.c_int), ("keep_points", ctypes.c_int),
    ("verbose", ctypes.c_int), ("assemble", ctypes.c_int), ("split", ctypes.c_int)])
_RunOut = _struct("pfa_run_output", [
    ("n", _c_i64), ("M", _P64), ("Kn", _P64), ("q", _P64), ("pidx", _P64), ("pstat", _P8),
    ("pcode", _P32), ("nexp", _P64), ("nonsing", _P8), ("stats", _c_f64 * len(STAT_NAMES)),
    ("err", ctypes.c_char * 256), ("n_pfv", _c_i64), ("Ks", _P64), ("Ms", _P64),
    ("pfv_pidx", _P64), ("pdone", _P8), ("pcount", _P64)])
_KerIn = _struct("pfa_kernel_input", [
    ("n_prob", ctypes.c_int), ("dims", _P32), ("nh", _P32), ("nl", _P32), ("has_lin", _P32),
    ("strict", _P32), ("mat", _P64), ("H", _P64), ("L", _P64), ("linvec", _P64), ("Q", _P64),
    ("qmax", _P64), ("linmin", _P64), ("device", ctypes.c_int), ("mode", ctypes.c_int)])
_KerOut = _struct("pfa_kernel_output", [
    ("n", _c_i64), ("pts", _P32), ("qs", _P64), ("pidx", _P64), ("status", _P32),
    ("nodes", _P64), ("dmax", ctypes.c_int), ("err", ctypes.c_char * 256)])
_ZpkOut = _struct("pfa_zpk_output", [
    ("n", _c_i64), ("M", _P64), ("Kn", _P64), ("q", _P64), ("pidx", _P64)])
_AsmOut = _struct("pfa_assemble_output", [
    ("n_rows", _c_i64), ("Ks", _P64), ("Ms", _P64), ("keys", _P64), ("done", _P8),
    ("err", ctypes.c_char * 256)])
# the reference's own structures (pfvs/fp_kernel/cuda/pfvs_gpu.cu:64-97), for
# callers of the drop-in entry point pfg_coni_batch
PfgInput = _struct("pfg_input", [
    ("n_geo", ctypes.c_int), ("h", _P32), ("Q", _P64), ("qmax", _P64), ("linmin", _P64),
    ("kappa", _P64), ("Mbasis", _P64), ("n_p", _c_i64), ("ps", _P64), ("pgeo", _P32),
    ("device", ctypes.c_int), ("batch", _c_i64), ("verbose", ctypes.c_int)])
PfgOutput = _struct("pfg_output", [
    ("n", _c_i64), ("M", _P64), ("Kn", _P64), ("q", _P64), ("pidx", _P64), ("pstat", _P8),
    ("seconds_gpu", _c_f64), ("err", ctypes.c_char * 256)])

_I, _CP, _VP = ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p
# every exported symbol of include/pfvs_amd.h: (argument types, result type)
SIGNATURES = {
    "pfg_coni_batch": ([ctypes.POINTER(PfgInput), ctypes.POINTER(PfgOutput)], _I),
    "pfg_output_free": ([ctypes.POINTER(PfgOutput)], None),
    "pfg_max_h11": ([], _I),
    "pfg_backend": ([], _CP),
    "pfg_device_count": ([], _I),
    "pfg_release": ([_I], None),
    "pfa_coni_run": ([ctypes.POINTER(_RunIn), ctypes.POINTER(_RunOut)], _I),
    "pfa_run_output_free": ([ctypes.POINTER(_RunOut)], None),
    "pfa_assemble": ([_I, _c_i64, _c_i64, _c_i64, _P64, _P64, _P64, _P64, _P8, _I,
                      ctypes.POINTER(_AsmOut)], _I),
    "pfa_assemble_output_free": ([ctypes.POINTER(_AsmOut)], None),
    "pfa_kernel_batch": ([ctypes.POINTER(_KerIn), ctypes.POINTER(_KerOut)], _I),
    "pfa_kernel_output_free": ([ctypes.POINTER(_KerOut)], None),
    "pfa_det_filter": ([_I, _P32, _P64, _c_i64, _P32, _P64, _I, _VP, _CP], _I),
    "pfa_det_filter_noalloc": ([_I, _I, _P32, _VP, _c_i64, _VP, _VP, _I, _VP, _VP, _CP], _I),
    "pfa_copy_probe": ([_I, _c_i64, _I, ctypes.POINTER(_c_f64)], _I),
    "pfa_box_enum": ([_P32, _I, _I, _P64, _I, _I, _I, _c_i64, _P64, _P64, _P64,
                      ctypes.POINTER(_c_f64), _CP], _I),
    "pfa_reserve": ([_I, _c_i64, _c_i64, _I, _CP], _I),
    "pfa_coni_mu0": ([_I, _P64, _c_i64, _P64, _I, _P32, _P64, _P64, _P64, _CP], _I),
    "pfa_coni_zpk": ([_I, _P64, _P64, _c_i64, _P64, _c_i64, _c_i64, _c_i64, _c_i64, _I, _P32,
                      ctypes.POINTER(_ZpkOut), _CP], _I),
    "pfa_zpk_output_free": ([ctypes.POINTER(_ZpkOut)], None),
    "pfa_device_cus": ([_I], _I),
    "pfa_device_clock_khz": ([_I], _I),
    "pfa_arch": ([], _CP),
}


class _Loader:
    """Loads the library once; remembers why it could not."""

    def __init__(self):
        self.lib = self.path = self.error = None
        self.tried = False

    def get(self):
        if not self.tried:
            self.tried = True
            self.path = lib_path()
            if not self.path:
                self.error = ("the HIP backend was not built (run __graft_entry__.build() or "
                              "python -m pfvs_amd.build)")
            else:
                try:
                    lib = ctypes.CDLL(self.path)
                    for name, (args, res) in SIGNATURES.items():
                        fn = getattr(lib, name)
                        fn.argtypes, fn.restype = args, res
                    self.lib = lib
                except (OSError, AttributeError) as e:
                    self.error = f"could not load {self.path}: {e}"
        return self.lib


_loader = _Loader()


def lib_path():
    """PFVS_GPU_LIB, else the library built next to this file."""
    env = os.environ.get("PFVS_GPU_LIB")
    if env:
        return env
    here = os.path.dirname(os.path.abspath(__file__))
    built = sorted(glob.glob(os.path.join(here, "libpfvs_amd*.so")))
    return built[0] if built else None


def loaded_library_path():
    _loader.get()
    return _loader.path if _loader.lib is not None else None


def require():
    """The library, or RuntimeError saying why it is missing."""
    lib = _loader.get()
    if lib is None:
        raise RuntimeError(_loader.error)
    return lib


# ------------------------------------------------- the reference's module API
def available() -> bool:
    """True when the library loads and sees an AMD GPU."""
    lib = _loader.get()
    return bool(lib is not None and lib.pfg_device_count() > 0)


def unavailable_reason() -> str:
    """Why available() is False ('' when it is True)."""
    lib = _loader.get()
    if lib is None:
        return _loader.error
    return "" if lib.pfg_device_count() > 0 else "no HIP device found"


def backend() -> str:
    return require().pfg_backend().decode()


def max_h11() -> int:
    return int(require().pfg_max_h11())


def release(device: int = 0) -> None:
    """Return the device buffers the library keeps between calls."""
    lib = _loader.get()
    if lib is not None:
        lib.pfg_release(int(device))


def qmax_of(Q, dilation) -> int:
    """The search bound floor(dilation Q), robust to the float rounding of
    dilation Q (the reference's 1e-7 allowance, pfvs/gpu.py:132-133)."""
    return int(math.floor(dilation * Q + 1e-7))


def _ptr(a, kind=_P64):
    return a.ctypes.data_as(kind)


@dataclass
class _Batch:
    """Geometries and p-vectors laid out for the C structures."""
    h: np.ndarray
    Q: np.ndarray
    qmax: np.ndarray
    linmin: np.ndarray
    kappa: np.ndarray
    Mbasis: np.ndarray
    ps: np.ndarray          # (max(n_p, 1), max_h11()) zero-padded, leading coni 0 included
    pgeo: np.ndarray
    n_p: int
    width: int

    @classmethod
    def build(cls, geoms, ps, pgeo):
        width = max_h11()
        hs, Qs, qms, lms, kaps, mbs = [], [], [], [], [], []
        for g in geoms:
            kap = np.ascontiguousarray(g["kappa"], dtype=np.int64)
            h = kap.shape[0]
            if not 3 <= h <= width:
                raise ValueError(f"the GPU build supports 3 <= h11 <= {width} (got {h})")
            hs.append(h)
            Qs.append(int(g["Q"]))
            qms.append(qmax_of(int(g["Q"]), g["dilation"]))
            lms.append(math.ceil(g["M0min"] - 1e-9))
            kaps.append(kap.ravel())
            mbs.append(np.ascontiguousarray(g["Mbasis"], dtype=np.int64).ravel())
        pgeo = np.ascontiguousarray(pgeo, dtype=np.int32).ravel()
        n_p = len(pgeo)
        if n_p and (pgeo.min() < 0 or pgeo.max() >= len(geoms)):
            raise ValueError("pgeo: geometry index out of range")
        h_of = np.asarray(hs, dtype=np.int32)
        P = np.zeros((max(n_p, 1), width), dtype=np.int64)
        if isinstance(ps, np.ndarray) and ps.ndim == 2 and len(ps) == n_p:
            A = ps.astype(np.int64, copy=False)
        else:
            rows = [np.asarray(r, dtype=np.int64).ravel() for r in ps][:n_p]
            A = np.zeros((n_p, max([len(r) for r in rows] + [1])), dtype=np.int64)
            for i, r in enumerate(rows):
                A[i, :len(r)] = r
        if n_p:
            # entries at or beyond a p-vector's own h11 must be 0
            beyond = np.arange(A.shape[1])[None, :] >= h_of[pgeo][:, None]
            if np.any((A != 0) & beyond):
                raise ValueError("a p-vector is longer than its geometry's h11")
            m = min(A.shape[1], width)
            P[:n_p, :m] = A[:, :m]
        flat = lambda parts: np.concatenate(parts) if parts else np.zeros(1, np.int64)
        return cls(h_of, np.asarray(Qs, np.int64), np.asarray(qms, np.int64),
                   np.asarray(lms, np.int64), flat(kaps), flat(mbs), P, pgeo, n_p, width)

    def run_input(self, mode, device, tune, host_threads, filter, sort, keep, verbose, assemble):
        return _RunIn(len(self.h), _ptr(self.h, _P32), _ptr(self.Q), _ptr(self.qmax),
                      _ptr(self.linmin), _ptr(self.kappa), _ptr(self.Mbasis), self.n_p,
                      _ptr(self.ps), _ptr(self.pgeo, _P32), int(device), MODES[mode],
                      int(tune.get("waves_per_cu", 0)), int(tune.get("budget", 0)),
                      int(tune.get("stack_cap", 0)), int(host_threads), int(bool(filter)),
                      int(bool(sort)), KEEP[keep], int(bool(verbose)), int(bool(assemble)),
                      int(tune.get("split", 0)))


_free = ctypes.CDLL(None).free
_free.argtypes = [ctypes.c_void_p]


class _Owned:
    """A malloc'd buffer of the library seen by numpy without a copy; the
    buffer is freed with the last array that refers to it."""

    def __init__(self, addr, nbytes):
        self._addr = addr
        self.__array_interface__ = {"shape": (nbytes,), "typestr": "|u1",
                                    "data": (addr, False), "version": 3}

    def __del__(self):
        _free(self._addr)


def _adopt(out, field, rows, width):
    """Take ownership of out.<field> as an int64 (rows, width) array."""
    ptr = getattr(out, field)
    if rows == 0 or not ptr:
        return np.zeros((rows, width), dtype=np.int64)
    buf = np.asarray(_Owned(ctypes.addressof(ptr.contents), 8 * rows * width))
    setattr(out, field, None)                 # no longer freed by *_free
    return buf.view(np.int64).reshape(rows, width)


def _copy(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(n, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def run(geoms, ps, pgeo, mode="batched", device=0, waves_per_cu=0, budget=0, stack_cap=0,
        host_threads=0, filter=True, sort=True, keep_points=True, verbose=False, assemble=False,
        split=0):
    """One pfa_coni_run call.

    geoms: per geometry kappa (h, h, h), Mbasis (h, h), Q, dilation, M0min;
    ps: p-vectors with the leading coni 0; pgeo: geometry of each.
    mode: "batched" (persistent, work stealing), "plain" (one launch, one
    wavefront per p-vector), "lane" (one thread per p-vector), "host".
    split (batched): nodes per persistent wavefront that the breadth-first
    expansion of the roots aims at before the persistent search (0: the
    library's default; < 0: no expansion, the roots start the search).
    keep_points: True / "all" every lattice point, "survivors" only those
    the filter keeps for the PFVs, False counters only.
    Returns a dict: M, Kn (n x max_h11()), q, pidx, nexp, nonsing per
    point; pstat, pcode, pcount (lattice points) per p-vector; stats
    (counters, stage ms); with
    assemble also pfv = {K, M, pidx} and pdone (the p-vectors the library
    finished)."""
    lib = require()
    b = _Batch.build(geoms, ps, pgeo)
    tune = dict(waves_per_cu=waves_per_cu, budget=budget, stack_cap=stack_cap, split=split)
    inp = b.run_input(mode, device, tune, host_threads, filter, sort, keep_points, verbose,
                      assemble)
    out = _RunOut()
    if lib.pfa_coni_run(ctypes.byref(inp), ctypes.byref(out)) != 0:
        msg = out.err.decode(errors="replace")
        lib.pfa_run_output_free(ctypes.byref(out))
        raise RuntimeError(f"GPU pipeline failed: {msg}")
    try:
        n = int(out.n) if KEEP[keep_points] else 0
        res = {"M": _adopt(out, "M", n, b.width), "Kn": _adopt(out, "Kn", n, b.width),
               "q": _copy(out.q, n, np.int64), "pidx": _copy(out.pidx, n, np.int64),
               "nexp": _copy(out.nexp, n, np.int64), "nonsing": _copy(out.nonsing, n, np.int8),
               "pstat": _copy(out.pstat, b.n_p, np.int8), "pcode": _copy(out.pcode, b.n_p, np.int32),
               "pcount": _copy(out.pcount, b.n_p, np.int64),
               "stats": dict(zip(STAT_NAMES, (float(x) for x in out.stats)))}
        if assemble and out.pdone:
            m = int(out.n_pfv)
            res["pfv"] = {"K": _adopt(out, "Ks", m, b.width), "M": _adopt(out, "Ms", m, b.width),
                          "pidx": _copy(out.pfv_pidx, m, np.int64)}
            res["pdone"] = _copy(out.pdone, b.n_p, np.int8)
    finally:
        lib.pfa_run_output_free(ctypes.byref(out))
    return res


def coni_batch_multi(geoms, ps, pgeo, device=0, batch=0, verbose=False):
    """The batched search for p-vectors of several geometries in one call
    (the reference's gpu.coni_batch_multi contract): (M, Kn, q, pidx,
    pstat, seconds_gpu).  `batch` is accepted for compatibility: the whole
    call is one persistent search."""
    r = run(geoms, ps, pgeo, mode="batched", device=device, filter=False, sort=True,
            keep_points=True, verbose=verbose)
    pstat = (r["pstat"] != 0).astype(np.int8)       # kernel statuses go to the host path too
    return r["M"], r["Kn"], r["q"], r["pidx"], pstat, r["stats"]["ms_total"] * 1e-3


def coni_batch(kappa, Mbasis, ps, Q, dilation, M0min, max_N_out, device=0):
    """One geometry (the reference's gpu.coni_batch): (M[:, :h], Kn[:, :h],
    q, pidx, pstat), where a p-vector with more than max_N_out points gets
    pstat 1 and none of its points."""
    h = np.asarray(kappa).shape[0]
    ps = np.asarray(ps, dtype=np.int64)
    geo = {"kappa": kappa, "Mbasis": Mbasis, "Q": Q, "dilation": dilation, "M0min": M0min}
    M, Kn, q, pidx, pstat, _ = coni_batch_multi([geo], ps, np.zeros(len(ps), np.int32), device)
    # points come grouped by p-vector: each one's count from the group edges
    edges = np.searchsorted(pidx, np.arange(len(ps) + 1))
    too_many = np.diff(edges) > max_N_out
    if too_many.any():
        pstat = np.where(too_many, 1, pstat).astype(np.int8)
        keep = ~too_many[pidx]
        M, Kn, q, pidx = M[keep], Kn[keep], q[keep], pidx[keep]
    return M[:, :h], Kn[:, :h], q, pidx, pstat


def pfg_call(geoms, ps, pgeo, device=0):
    """The drop-in entry point pfg_coni_batch through the reference's own
    structure layout: (M, Kn, q, pidx, pstat, seconds_gpu)."""
    lib = require()
    b = _Batch.build(geoms, ps, pgeo)
    inp = PfgInput(len(b.h), _ptr(b.h, _P32), _ptr(b.Q), _ptr(b.qmax), _ptr(b.linmin),
                   _ptr(b.kappa), _ptr(b.Mbasis), b.n_p, _ptr(b.ps), _ptr(b.pgeo, _P32),
                   int(device), 0, 0)
    out = PfgOutput()
    if lib.pfg_coni_batch(ctypes.byref(inp), ctypes.byref(out)) != 0:
        raise RuntimeError(f"GPU pipeline failed: {out.err.decode(errors='replace')}")
    try:
        n = int(out.n)
        grab = lambda p: (np.ctypeslib.as_array(p, shape=(n, b.width)).copy() if n
                          else np.zeros((0, b.width), np.int64))
        return (grab(out.M), grab(out.Kn), _copy(out.q, n, np.int64), _copy(out.pidx, n, np.int64),
                _copy(out.pstat, b.n_p, np.int8), float(out.seconds_gpu))
    finally:
        lib.pfg_output_free(ctypes.byref(out))


# ------------------------------------------------------------ other calls
def assemble_points(M, Kn, q, key, nonsing, h, Q, n_keys, threads=0):
    """The library's post-processing of given points (pfa_assemble): rows
    M, Kn (n, h) int64, q, key (p-vector index), verdict (1 / 0 / 2).
    Returns (Ks, Ms, keys, done per key)."""
    lib = require()
    M = np.ascontiguousarray(M, dtype=np.int64).reshape(-1, h)
    Kn = np.ascontiguousarray(Kn, dtype=np.int64).reshape(-1, h)
    q = np.ascontiguousarray(q, dtype=np.int64)
    key = np.ascontiguousarray(key, dtype=np.int64)
    ns = np.ascontiguousarray(nonsing, dtype=np.int8)
    n = len(q)
    out = _AsmOut()
    rc = lib.pfa_assemble(int(h), int(Q), int(n_keys), n, _ptr(key), _ptr(M), _ptr(Kn), _ptr(q),
                          _ptr(ns, _P8), int(threads), ctypes.byref(out))
    try:
        if rc != 0:
            raise RuntimeError(out.err.decode(errors="replace"))
        m = int(out.n_rows)
        return (_adopt(out, "Ks", m, h), _adopt(out, "Ms", m, h), _copy(out.keys, m, np.int64),
                _copy(out.done, int(n_keys), np.int8))
    finally:
        lib.pfa_assemble_output_free(ctypes.byref(out))


def coni_zpk_points(kappa, Mbasis, p_full, Q, M0min, D0, max_out=10**7, threads=0):
    """pfa_coni_zpk: the lattice points behind the coni PFVs with dilation
    above D0 of each direction (rows of p_full, h entries, p[0] = 0), on
    host threads.  Returns (M, Kn (n, h) int64, q, pidx, status per row);
    a row with status != 0 contributes no points (1: no short-vector bound;
    < 0: left to the caller's search to its dilation bound)."""
    lib = require()
    kappa = np.ascontiguousarray(kappa, dtype=np.int64)
    Mbasis = np.ascontiguousarray(Mbasis, dtype=np.int64)
    h = kappa.shape[0]
    p_full = np.ascontiguousarray(np.atleast_2d(np.asarray(p_full, dtype=np.int64)).reshape(-1, h))
    n_p = len(p_full)
    st = np.zeros(max(n_p, 1), np.int32)
    out = _ZpkOut()
    err = ctypes.create_string_buffer(256)
    rc = lib.pfa_coni_zpk(h, _ptr(kappa), _ptr(Mbasis), n_p, _ptr(p_full), int(Q), int(M0min),
                          int(D0), int(max_out), int(threads), _ptr(st, _P32), ctypes.byref(out), err)
    try:
        if rc != 0:
            raise ValueError(err.value.decode(errors="replace"))
        n = int(out.n)
        return (_copy(out.M, n * h, np.int64).reshape(n, h), _copy(out.Kn, n * h, np.int64).reshape(n, h),
                _copy(out.q, n, np.int64), _copy(out.pidx, n, np.int64), st[:n_p])
    finally:
        lib.pfa_zpk_output_free(ctypes.byref(out))


def box_enum(H, rhs, B, primitive=True, cap=10**8, device=0):
    """pfa_box_enum: integer p with |p|_inf <= B and H p >= rhs (primitive
    if asked), lexicographic.  Returns (the first min(total, cap) rows,
    status 0 or 1 (more than cap), total, ms); device < 0: host code."""
    lib = require()
    H = np.ascontiguousarray(H, dtype=np.int64)
    if H.size and int(np.abs(H).max()) >= 2**31:
        raise ValueError("H entries must fit int32")
    H = H.astype(np.int32)
    m, n = H.shape
    rhs = np.ascontiguousarray(np.broadcast_to(np.asarray(rhs, dtype=np.int64), (m,)))
    cap = max(int(cap), 0)
    if cap > 1 << 16:                             # hold only what exists
        _, _, total, _ = box_enum(H, rhs, B, primitive, 0, device)
        room = min(total, cap)
    else:
        room = cap
    rows = np.empty((room, n), dtype=np.int64)
    got, total, ms = _c_i64(0), _c_i64(0), _c_f64(0.0)
    err = ctypes.create_string_buffer(256)
    rc = lib.pfa_box_enum(_ptr(H, _P32), m, n, _ptr(rhs), int(B), int(bool(primitive)), int(device),
                          room, _ptr(rows), ctypes.byref(got), ctypes.byref(total),
                          ctypes.byref(ms), err)
    if rc == 2:
        raise RuntimeError(f"box_enum failed: {err.value.decode(errors='replace')}")
    return rows[:got.value], int(total.value > cap), int(total.value), float(ms.value)


def reserve(n_p, n_points, h_max=None, device=0):
    """Grow the device and pinned buffers for batches up to n_p p-vectors
    and n_points points at h11 <= h_max (default max_h11())."""
    err = ctypes.create_string_buffer(256)
    hm = max_h11() if h_max is None else int(h_max)
    if require().pfa_reserve(int(device), int(n_p), int(n_points), hm, err) != 0:
        raise RuntimeError(f"reserve failed: {err.value.decode(errors='replace')}")


def device_facts(device: int = 0):
    """(compute units, peak engine clock in kHz) of `device`."""
    lib = require()
    return int(lib.pfa_device_cus(device)), int(lib.pfa_device_clock_khz(device))


def copy_probe(nbytes, reps=10, device=0):
    """Mean ms of one plain device copy of nbytes (the reference kernel for
    the HBM traffic counters)."""
    ms = _c_f64(0.0)
    if require().pfa_copy_probe(int(device), int(nbytes), int(reps), ctypes.byref(ms)) != 0:
        raise RuntimeError("copy probe failed")
    return ms.value


def det_filter(kappas, M, geo=None, device=0):
    """det((kappa M)[1:, 1:]) != 0 per row of M (int8: 1 nonsingular, 0
    singular, 2 undecided).  kappas: one (h, h, h) array per geometry; M
    rows zero-padded to the largest h; geo: geometry per row (default 0);
    device < 0: host code."""
    lib = require()
    hs = np.array([np.asarray(k).shape[0] for k in kappas], dtype=np.int32)
    kap = np.ascontiguousarray(np.concatenate([np.asarray(k, dtype=np.int64).ravel() for k in kappas]))
    M = np.asarray(M, dtype=np.int64)
    rows = np.zeros((M.shape[0], int(hs.max())), dtype=np.int64)
    rows[:, :M.shape[1]] = M
    g = np.zeros(len(rows), np.int32) if geo is None else np.ascontiguousarray(geo, dtype=np.int32)
    verdict = np.zeros(max(len(rows), 1), dtype=np.int8)
    err = ctypes.create_string_buffer(256)
    if lib.pfa_det_filter(len(kappas), _ptr(hs, _P32), _ptr(kap), len(rows), _ptr(g, _P32),
                          _ptr(rows), int(device), verdict.ctypes.data_as(ctypes.c_void_p), err):
        raise RuntimeError("det filter failed: " + err.value.decode(errors="replace"))
    return verdict[:len(rows)]


def det_filter_noalloc(device, h, kappa_dev, count, geo_dev, M_dev, hstride, out_dev, stream=0):
    """pfa_det_filter_noalloc on device arrays (addresses as ints, from the
    same HIP runtime the library uses): kappa (sum h^3 int64), geo (count
    int32), M (count x hstride int64), out (count int8).  Allocates
    nothing; returns when the work is queued on `stream`."""
    hs = np.ascontiguousarray(h, dtype=np.int32)
    err = ctypes.create_string_buffer(256)
    if require().pfa_det_filter_noalloc(int(device), len(hs), _ptr(hs, _P32), kappa_dev, int(count),
                                        geo_dev, M_dev, int(hstride), out_dev, stream, err) != 0:
        raise RuntimeError("det filter (no allocation) failed: " + err.value.decode(errors="replace"))


# -------------------------------------------------- single-problem searches
_INT127 = 1 << 127


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
    idx = [k for k in range(len(problems)) if status[k] == 1 and len(problems[k]["mat"]) > 1
           and _leaf_words(problems[k].get("L", [])) is not None]
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
    order, statuses, node counts (per problem in host mode; 0 on the
    device)).  H may be any integer matrix (rows with repeated leading
    columns are brought to echelon form first).  A problem whose FP64
    factor margin is too wide (an ill-conditioned mat) is searched again in
    an LLL-reduced basis of its form.  Status 1 that remains: the problem
    needs an exact search beyond this one (mat above 2^53, a leaf row entry
    of 2^127 or more, or a leaf gcd beyond int128)."""
    problems = [_echelon_rows(p) for p in problems]
    res, status, nodes = _kernel_call(problems, mode, device)
    _retry_reduced(problems, res, status, nodes, mode, device)
    return res, status, nodes


def _leaf_words(rows):
    """Leaf rows as (low, high) int64 word pairs of their int128 entries,
    or None when an entry does not fit int128 (|x| >= 2^127)."""
    words = []
    for row in rows:
        for x in row:
            x = int(x)
            if not -_INT127 <= x < _INT127:
                return None
            u = x & ((1 << 128) - 1)
            for w in (u & ((1 << 64) - 1), u >> 64):
                words.append(w - (1 << 64) if w >> 63 else w)
    return words


def _kernel_call(problems, mode, device):
    """One pfa_kernel_batch call over problems with H in echelon form.  A
    problem with a leaf row entry beyond int128 is not sent: it gets status
    1 (the exact search it needs is beyond this call)."""
    lib = require()
    sendable = []
    for k, p in enumerate(problems):
        w = _leaf_words(p.get("L", []))
        if w is not None:
            sendable.append((k, p, w))
    n = len(sendable)
    probs = [p for _, p, _ in sendable]
    cat = lambda parts: np.ascontiguousarray(np.concatenate(parts + [np.zeros(1, np.int64)]))
    i32 = lambda xs: np.ascontiguousarray(xs, dtype=np.int32)
    dims = i32([len(p["mat"]) for p in probs])
    nh = i32([len(p.get("H", [])) for p in probs])
    nl = i32([len(p.get("L", [])) for p in probs])
    has_lin = i32([p.get("linvec") is not None for p in probs])
    strict = i32([bool(p.get("strict", True)) for p in probs])
    mat = cat([np.asarray(p["mat"], dtype=np.int64).ravel() for p in probs])
    H = cat([np.asarray(p.get("H", np.zeros((0, len(p["mat"])))), dtype=np.int64).ravel() for p in probs])
    L = cat([np.asarray(w, dtype=np.int64) for _, _, w in sendable])
    lin = cat([np.asarray(p["linvec"], dtype=np.int64).ravel() for p in probs
               if p.get("linvec") is not None])
    Q = np.array([int(p["Q"]) for p in probs] + [0], dtype=np.int64)
    qmax = np.array([int(p["qmax"]) for p in probs] + [0], dtype=np.int64)
    linmin = np.array([int(p.get("linmin", 0) or 0) for p in probs] + [0], dtype=np.int64)
    inp = _KerIn(n, _ptr(dims, _P32), _ptr(nh, _P32), _ptr(nl, _P32), _ptr(has_lin, _P32),
                 _ptr(strict, _P32), _ptr(mat), _ptr(H), _ptr(L), _ptr(lin), _ptr(Q), _ptr(qmax),
                 _ptr(linmin), int(device), MODES[mode])
    out = _KerOut()
    if lib.pfa_kernel_batch(ctypes.byref(inp), ctypes.byref(out)) != 0:
        raise RuntimeError(f"kernel batch failed: {out.err.decode(errors='replace')}")
    try:
        m, D = int(out.n), int(out.dmax)
        pts = np.ctypeslib.as_array(out.pts, shape=(m, D)).copy() if m else np.zeros((0, D), np.int32)
        qs, pidx = _copy(out.qs, m, np.int64), _copy(out.pidx, m, np.int64)
        st, nd = _copy(out.status, n, np.int32), _copy(out.nodes, n, np.int64)
    finally:
        lib.pfa_kernel_output_free(ctypes.byref(out))
    N = len(problems)
    res = [(np.zeros((0, len(p["mat"])), np.int32), np.zeros(0, np.int64)) for p in problems]
    status = np.ones(N, dtype=np.int32)
    nodes = np.zeros(N, dtype=np.int64)
    edges = np.searchsorted(pidx, np.arange(n + 1))        # points grouped by problem
    for j, (k, p, _) in enumerate(sendable):
        a, b = edges[j], edges[j + 1]
        res[k] = (pts[a:b, :dims[j]], qs[a:b])
        status[k], nodes[k] = st[j], nd[j]
    return res, status, nodes
