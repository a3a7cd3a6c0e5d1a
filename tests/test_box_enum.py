"""The p-vector enumeration (pfa_box_enum; latticepts.box_enum's contract
at pfvs/distributed.py:121-139 and tests/test_paper_examples.py:119-121):
the host code against brute force over the whole box and the oracle's depth
first restatement (same points, same lexicographic order) on random and
edge cases; the reference's recorded scan counts of the paper geometries;
capacity status; the device form (one GPU thread per unit) against the
host."""

import json
import os

import numpy as np
import pytest

from oracle.boxes import box_enum_brute, box_enum_dfs
from pfvs_amd import gpu
from pfvs_amd.cydata import CYData
from pfvs_amd.pvectors import box_enum, box_enum_status, pvecs

HERE = os.path.dirname(os.path.abspath(__file__))
PAPER = json.load(open(os.path.join(HERE, "golden", "paper_coni.json")))


def _paper_data(e):
    return CYData(e["h21"], e["kappa"], e["c2"], e["H"], coni_curve=e["coni_curve"],
                  coni_cob=e["cob"])


def random_cases(seed, n_cases):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        n = int(rng.integers(1, 6))
        m = int(rng.integers(0, 7))
        B = int(rng.integers(0, 4))
        H = rng.integers(-3, 4, size=(m, n)).astype(np.int32)
        rhs = rng.integers(-2, 3, size=m)
        out.append((H, rhs, B, bool(rng.integers(0, 2))))
    # edge cases: no rows (the whole box), an empty cone, the zero vector
    # allowed (not primitive) and excluded, one coordinate, rows beyond 64
    out.append((np.zeros((0, 3), np.int32), np.zeros(0, np.int64), 2, False))
    out.append((np.array([[1, 0], [-1, 0]], np.int32), np.array([1, 1]), 3, True))
    out.append((np.array([[1, 1]], np.int32), np.array([0]), 2, False))
    out.append((np.array([[1, 1]], np.int32), np.array([0]), 2, True))
    out.append((np.array([[2], [-1]], np.int32), np.array([1, -5]), 6, True))
    Hbig = rng.integers(-1, 2, size=(150, 4)).astype(np.int32)
    Hbig[:, 0] = np.abs(Hbig[:, 0]) + 1
    out.append((Hbig, np.full(150, -3), 2, True))
    return out


def test_brute_force_pins_the_dfs_restatement():
    for H, rhs, B, prim in random_cases(1, 60):
        assert np.array_equal(box_enum_dfs(B, H, rhs, prim), box_enum_brute(B, H, rhs, prim))


def test_host_enumeration_matches_brute_force():
    gpu.require()
    n = 0
    for H, rhs, B, prim in random_cases(2, 200):
        ps, st, total, _ = gpu.box_enum(H, rhs, B, prim, 10**6, device=-1)
        exp = box_enum_brute(B, H, rhs, prim)
        assert st == 0 and total == len(exp)
        assert np.array_equal(ps, exp), (H.tolist(), rhs.tolist(), B, prim)
        n += len(exp)
    assert n > 1000


def test_host_scan_counts_of_the_paper_geometries():
    """The reference's recorded scans (test_paper_examples.py:119-121):
    len(box_enum(B, H_cob, 1, primitive)) == n_p, and the paper's p is in it."""
    gpu.require()
    for e in PAPER:
        if e["scan"]["slow"]:
            continue
        data = _paper_data(e)
        ps, st, total = box_enum_status(e["scan"]["B"], data.H_cob, 1, 10**8, True, device="cpu")
        assert st == 0 and len(ps) == total == e["scan"]["n_p"], e["name"]
        assert (ps == np.asarray(e["p"])[1:]).all(axis=1).any(), e["name"]
        if len(ps) <= 3000:
            assert np.array_equal(ps, box_enum_dfs(e["scan"]["B"], data.H_cob, 1, True))


def test_capacity_status_and_pvecs():
    gpu.require()
    e = PAPER[4]
    data = _paper_data(e)
    full = box_enum(e["scan"]["B"], data.H_cob, device="cpu")
    ps, st, total = box_enum_status(e["scan"]["B"], data.H_cob, 1, 100, True, device="cpu")
    assert st == 1 and total == len(full) and np.array_equal(ps, full[:100])
    with pytest.raises(RuntimeError, match="exceed"):
        box_enum(e["scan"]["B"], data.H_cob, max_pts=10, device="cpu")
    ps = pvecs(data, 100, device="cpu")
    B = int(np.abs(ps).max())
    assert len(ps) >= 100 and len(box_enum(B - 1, data.H_cob, device="cpu")) < 100


def test_host_threads_run_the_device_walk_at_every_prefix_split():
    """The per-unit walk (the code the device runs, one GPU thread per unit)
    on host threads: the same points at any split of the box into units,
    from none (s = 0) to every coordinate in the prefix."""
    gpu.require()
    for H, rhs, B, prim in random_cases(4, 60):
        exp = box_enum_brute(B, H, rhs, prim)
        ps, st, total, _ = gpu.box_enum(H, rhs, B, prim, 10**6, device=-1)
        assert st == 0 and np.array_equal(ps, exp)


@pytest.mark.gpu
def test_gpu_enumeration_matches_host(need_gpu):
    """One GPU thread per unit (no barriers or cross-lane traffic, every
    loop bounded by the box): the host's points in the host's order, the
    recorded scan counts of the paper geometries, the capacity status."""
    for H, rhs, B, prim in random_cases(3, 120):
        d, st, total, _ = gpu.box_enum(H, rhs, B, prim, 10**6, device=0)
        h, st_h, total_h, _ = gpu.box_enum(H, rhs, B, prim, 10**6, device=-1)
        assert st == st_h == 0 and total == total_h and np.array_equal(d, h)
    for e in PAPER:
        data = _paper_data(e)
        print(f"scan {e['name']}: B {e['scan']['B']}, {e['scan']['n_p']} p-vectors", flush=True)
        d, st, total, ms = gpu.box_enum(data.H_cob, 1, e["scan"]["B"], True, 10**8, device=0)
        assert st == 0 and total == e["scan"]["n_p"], e["name"]
        h, _, _, _ = gpu.box_enum(data.H_cob, 1, e["scan"]["B"], True, 10**8, device=-1)
        assert np.array_equal(d, h), e["name"]
    ps, st, total, _ = gpu.box_enum(_paper_data(PAPER[1]).H_cob, 1, 14, True, 1000, device=0)
    assert st == 1 and total == PAPER[1]["scan"]["n_p"] and len(ps) == 1000
