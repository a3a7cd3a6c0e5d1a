"""The allocation-free det filter (pfa_det_filter_noalloc) against the
allocating entry point (pfa_det_filter) and the oracle's exact rank, on
the constructed cases of test_det_filter (every branch: dense, low rank,
dependent columns, double and modular eliminations, int128 builds,
undecided entries, sparse kappa beyond h11 = 25, mixed h11 in one call),
called back to back through the C ABI with inputs in device memory, the
second call of each pair smaller than the first; and its size limits."""

import ctypes

import numpy as np
import pytest

from pfvs_amd import gpu
from test_det_filter import _expect, as_i64, cases

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library
    itself is linked against (the caller's side of the no-allocation
    contract: inputs and outputs are the caller's device memory)."""

    def __init__(self):
        gpu.require()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        if out.nbytes:
            assert self.rt.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _run(hip, kaps, M, geo, stream=0):
    hs = np.array([k.shape[0] for k in kaps], np.int32)
    hmax = int(hs.max())
    rows = np.zeros((len(M), hmax), np.int64)
    rows[:, :M.shape[1]] = M
    g = np.zeros(len(M), np.int32) if geo is None else np.asarray(geo, np.int32)
    kap = np.concatenate([as_i64(k).ravel() for k in kaps])
    out = hip.put(np.full(max(len(M), 1), -1, np.int8))
    gpu.det_filter_noalloc(0, hs, hip.put(kap), len(M), hip.put(g), hip.put(rows), hmax, out,
                           stream)
    if stream:
        assert hip.rt.hipStreamSynchronize(ctypes.c_void_p(stream)) == 0
    else:
        assert hip.rt.hipDeviceSynchronize() == 0
    return hip.get(out, len(M), np.int8)


def test_noalloc_matches_allocating_entry_point(need_gpu):
    """Every case twice in a row (the second call on the first rows only):
    equal to pfa_det_filter and to the exact rank."""
    hip = _Hip()
    stream = ctypes.c_void_p()
    assert hip.rt.hipStreamCreate(ctypes.byref(stream)) == 0
    n = 0
    try:
        for k, (name, kaps, M, geo) in enumerate(cases(11)):
            print(f"case {name}: {len(M)} points", flush=True)
            exp = _expect(kaps, M, geo)
            ref = gpu.det_filter([as_i64(x) for x in kaps], M, geo, device=0)
            assert ref.tolist() == exp.tolist(), name
            got = _run(hip, kaps, M, geo, stream.value if k % 2 else 0)
            assert got.tolist() == ref.tolist(), name
            m = max(1, len(M) // 2)                        # the next call is smaller
            got2 = _run(hip, kaps, M[:m], None if geo is None else geo[:m])
            assert got2.tolist() == ref[:m].tolist(), name
            n += len(M) + m
            hip.free()
    finally:
        hip.free()
        hip.rt.hipStreamDestroy(stream)
    assert n > 200


def test_noalloc_many_points_then_few(need_gpu):
    """A call near the size limit (many points through the modular list)
    followed by a small one: the small call's results are its own."""
    hip = _Hip()
    try:
        big = [c for c in cases(5) if c[0].startswith("modular dense h=12")][0]
        _, kaps, M, geo = big
        rng = np.random.default_rng(3)
        Mb = rng.integers(-3, 4, size=(20000, M.shape[1]))
        ref = gpu.det_filter([as_i64(x) for x in kaps], Mb, None, device=0)
        assert (ref == 1).sum() > 1000
        got = _run(hip, kaps, Mb, None)
        assert got.tolist() == ref.tolist()
        small = Mb[:37]
        got2 = _run(hip, kaps, small, None)
        assert got2.tolist() == ref[:37].tolist()
    finally:
        hip.free()


def test_noalloc_rejects_sizes_beyond_its_limits(need_gpu):
    hip = _Hip()
    try:
        kap = as_i64(np.zeros((5, 5, 5), dtype=object))
        d = hip.put(kap.ravel())
        with pytest.raises(RuntimeError, match="count"):
            gpu.det_filter_noalloc(0, [5], d, 65537, d, d, 5, d)
        with pytest.raises(RuntimeError, match="n_geo"):
            gpu.det_filter_noalloc(0, [5] * 17, d, 1, d, d, 5, d)
        with pytest.raises(RuntimeError, match="h11"):
            gpu.det_filter_noalloc(0, [42], d, 1, d, d, 42, d)
        with pytest.raises(RuntimeError, match="stride"):
            gpu.det_filter_noalloc(0, [5, 9], d, 1, d, d, 5, d)
    finally:
        hip.free()
