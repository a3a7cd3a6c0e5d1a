"""The continuous-batching server: admission splits jobs over iterations in
arrival order and completes each exactly once (CPU, stand-in runner); on
the GPU a stream of serving jobs of different h11 gives, for every job,
the same (Ks, Ms) as the job run alone."""

import threading
import time

import numpy as np
import pytest

from pfvs_amd import coni, gpu
from pfvs_amd.serve import Server


def _empty_runner(log):
    def run(geoms, ps, pgeo):
        log.append((len(ps), len(geoms)))
        n = len(ps)
        z = np.zeros((0, gpu.max_h11()), np.int64)
        return dict(M=z, Kn=z, q=np.zeros(0, np.int64), pidx=np.zeros(0, np.int64),
                    nonsing=np.zeros(0, np.int8), pstat=np.zeros(n, np.int8),
                    stats={"nodes": float(n)})
    return run


def _job_args(h, n, seed):
    rng = np.random.default_rng(seed)
    kap = np.zeros((h, h, h), np.int64)
    return kap, np.eye(h, dtype=np.int64), 100, rng.integers(1, 5, size=(n, h - 1)), 400


def test_server_splits_and_completes_in_order():
    gpu.require()
    log = []
    srv = Server(max_batch_p=10, runner=_empty_runner(log), pipeline=1)
    jobs = [srv.submit(*_job_args(6, n, s)) for s, n in enumerate((7, 25, 3, 0, 12))]
    srv.start()
    for j in jobs:
        Ks, Ms = j.wait(timeout=30)
        assert Ks.shape == (0, j.h)
    srv.stop(timeout=10)
    assert sum(n for n, _ in log) == 7 + 25 + 3 + 12          # every p-vector once
    assert all(n <= 10 for n, _ in log) and len(log) >= 5
    done = [j.t_done for j in jobs if len(j.ps)]
    assert done == sorted(done)                                 # FIFO completion
    st = srv.stats()
    assert st["jobs"] == 5 and st["errors"] == 0 and st["p99_ms"] >= st["p50_ms"]


def test_pipelined_iterations_return_every_row_once():
    """Two iterations in flight, the later one finishing first: every
    p-vector runs once, every job completes once, and a job split over
    iterations gets its rows back in p-vector order."""
    gpu.require()
    calls = []

    def run(geoms, ps, pgeo):
        n = len(ps)
        calls.append(n)
        time.sleep(0.05 if len(calls) % 2 else 0.0)       # odd calls finish last
        MH = gpu.max_h11()
        M = np.zeros((n, MH), np.int64)
        M[:, 0] = 13                                       # M0 >= M0min, as every point has
        M[:, 1] = ps[:, 1]                                 # tag each row with its p-vector
        return dict(M=M, Kn=M.copy(), q=np.ones(n, np.int64), pidx=np.arange(n),
                    nonsing=np.zeros(n, np.int8), pstat=np.zeros(n, np.int8),
                    stats={"nodes": float(n)})
    srv = Server(max_batch_p=8, runner=run, pipeline=2, keep_points=True)
    args = []
    for s_, n in enumerate((13, 5, 21, 2)):
        kap, Mb, Q, ps, qmax = _job_args(6, n, s_)
        ps = ps.copy()
        ps[:, 0] = np.arange(n) + 1                        # p-vector tags 1..n
        args.append((kap, Mb, Q, ps, qmax))
    jobs = [srv.submit(*a) for a in args]
    srv.start()
    for j in jobs:
        j.wait(timeout=30)
    srv.stop(timeout=10)
    assert sum(calls) == 13 + 5 + 21 + 2 and len(calls) >= 6
    for (kap, Mb, Q, ps, qmax), j in zip(args, jobs):
        P = j.points
        assert P["pidx"].tolist() == list(range(len(ps)))
        assert P["M"][:, 1].tolist() == list(range(1, len(ps) + 1))
    assert srv.stats()["jobs"] == 4


def test_server_fails_jobs_of_a_failed_iteration():
    gpu.require()

    def bad(geoms, ps, pgeo):
        raise RuntimeError("device lost")
    srv = Server(runner=bad).start()
    j = srv.submit(*_job_args(5, 4, 0))
    with pytest.raises(RuntimeError, match="device lost"):
        j.wait(timeout=30)
    srv.stop(timeout=10)


def test_points_checksum_sees_order_and_entries():
    from pfvs_amd.serve import points_checksum
    rng = np.random.default_rng(0)
    r = dict(M=rng.integers(-5, 5, (50, 41)), Kn=rng.integers(-5, 5, (50, 41)),
             q=rng.integers(0, 9, 50), pidx=np.arange(50),
             nonsing=rng.integers(0, 3, 50).astype(np.int8))
    c = points_checksum(r, 20)
    for key, edit in (("M", lambda a: a.__setitem__([3, 4], a[[4, 3]])),
                      ("Kn", lambda a: a.__setitem__((7, 5), a[7, 5] + 1)),
                      ("q", lambda a: a.__setitem__(0, a[0] + 1)),
                      ("nonsing", lambda a: a.__setitem__(9, (a[9] + 1) % 3))):
        r2 = {k: v.copy() for k, v in r.items()}
        edit(r2[key])
        assert points_checksum(r2, 20) != c, key
    r2 = {k: v.copy() for k, v in r.items()}
    r2["M"][:, 30] += 1                                  # beyond h: padding, ignored
    assert points_checksum(r2, 20) == c
    assert points_checksum(dict(r, q=r["q"][:0]), 20) == 0


def run_alone(kap, Mb, Q, ps, qmax):
    """One job on its own through the same pipeline and post-processing."""
    geoms, pgeo = [], np.zeros(len(ps), np.int32)
    for v in np.unique(qmax):
        pgeo[np.flatnonzero(qmax == v)] = len(geoms)
        geoms.append(dict(kappa=kap, Mbasis=Mb, Q=Q, dilation=int(v) / Q, M0min=13))
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    r = gpu.run(geoms, pf, pgeo, mode="batched", filter=True, sort=True)
    return r, coni.assemble(r, ps, pf, kap, Mb, Q, qmax, 13)


def _dataset_jobs(n):
    """Dataset geometries whose PFVs are known: (args, geometry record)."""
    from helpers import cydata_of, geometries
    from pfvs_amd.pvectors import box_enum
    out = []
    for g in geometries():
        if not (3 <= g["h11"] <= 11) or not g["K"]:
            continue
        data = cydata_of(g)
        ps = box_enum(g["B"], data.H_cob, 1, True)
        if not 0 < len(ps) <= 2000:
            continue
        Q = data.h11 + data.h21 + 4
        qmax = np.full(len(ps), int(np.floor(g["D"] * Q + 1e-7)))
        out.append(((data.kappa_cob, data.M_lattice(), Q, ps, qmax), g))
        if len(out) == n:
            break
    return out


@pytest.mark.gpu
def test_gpu_serving_results_same_alone_and_in_mix(need_gpu):
    """Synthetic requests of h11 5..40 interleaved with dataset geometries:
    every job's lattice points and (Ks, Ms) equal the job run alone, and
    the dataset jobs reproduce the dataset's PFVs."""
    from helpers import as_set
    from pfvs_amd.instances import serving_request
    reqs = []
    for r in range(10):
        inst, ps, qmax = serving_request(r, n_max=96)
        qmax = np.broadcast_to(np.asarray(qmax, np.int64), (len(ps),)).copy()
        reqs.append(((inst.data.kappa_cob, inst.data.M_lattice(), inst.Q, ps, qmax), None))
    ds = _dataset_jobs(4)
    assert len(ds) == 4
    for k, job in enumerate(ds):
        reqs.insert(2 * k + 1, job)
    srv = Server(max_batch_p=256, keep_points=True).start()
    jobs = []

    def client():
        for args, _ in reqs:
            jobs.append(srv.submit(*args))
            time.sleep(0.002)
    th = threading.Thread(target=client)
    th.start()
    th.join()
    for j in jobs:
        j.wait(timeout=120)
    srv.stop(timeout=30)
    st = srv.stats()
    print(f"serving: {st}", flush=True)
    assert st["errors"] == 0 and st["jobs"] == len(reqs)
    n_pfvs = 0
    for (args, g), j in zip(reqs, jobs):
        r, (Ks, Ms) = run_alone(*args)
        h = j.h
        assert np.array_equal(Ks, j.result[0]) and np.array_equal(Ms, j.result[1]), j.id
        P = j.points
        assert np.array_equal(P["pidx"], r["pidx"]) and np.array_equal(P["q"], r["q"]), j.id
        assert np.array_equal(P["M"][:, :h], r["M"][:, :h]), j.id
        assert np.array_equal(P["Kn"][:, :h], r["Kn"][:, :h]), j.id
        assert np.array_equal(P["nonsing"], r["nonsing"]), j.id
        if g is not None:
            got, exp = as_set(Ks, Ms), as_set(g["K"], g["M"])
            assert exp <= got, j.id
            n_pfvs += len(Ks)
    assert n_pfvs > 0


@pytest.mark.gpu
def test_gpu_reserved_server_allocates_nothing(need_gpu):
    """After pfa_reserve at start, no iteration of a mixed-h11 stream grows a
    point, record, list or staging buffer (those took 100+ ms; only the
    small per-geometry tables are allocated at first use, well under 10 ms)
    and no search reruns for lack of point slots."""
    from pfvs_amd.instances import serving_request
    srv = Server(max_batch_p=512, keep_points="checksum", reserve=(4096, 2_000_000, 41)).start()
    jobs = []
    for r in range(12):
        inst, ps, qmax = serving_request(r, n_max=256)
        qmax = np.broadcast_to(np.asarray(qmax, np.int64), (len(ps),)).copy()
        jobs.append(srv.submit(inst.data.kappa_cob, inst.data.M_lattice(), inst.Q, ps, qmax))
    for j in jobs:
        j.wait(timeout=120)
    srv.stop(timeout=30)
    st = srv.stats()
    print(f"serving: {st}", flush=True)
    assert st["errors"] == 0
    its = [it[5] for it in srv.iterations]         # per-iteration stage stats
    assert its and all(ms["ms_alloc"] < 10.0 for ms in its), [ms["ms_alloc"] for ms in its]
    assert all(ms["attempts"] == 1.0 for ms in its), [ms["attempts"] for ms in its]
