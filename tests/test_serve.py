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
    srv = Server(max_batch_p=10, runner=_empty_runner(log))
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


def test_server_fails_jobs_of_a_failed_iteration():
    gpu.require()

    def bad(geoms, ps, pgeo):
        raise RuntimeError("device lost")
    srv = Server(runner=bad).start()
    j = srv.submit(*_job_args(5, 4, 0))
    with pytest.raises(RuntimeError, match="device lost"):
        j.wait(timeout=30)
    srv.stop(timeout=10)


def run_alone(inst, ps, qmax):
    """One job on its own through the same pipeline and post-processing."""
    from pfvs_amd.instances import geoms_for_qmax
    geoms, pgeo = geoms_for_qmax(inst, qmax)
    pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
    r = gpu.run(geoms, pf, pgeo, mode="batched", filter=True, sort=True)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    return coni.assemble(r, ps, pf, kap, Mb, inst.Q, qmax, 13)


@pytest.mark.gpu
def test_gpu_serving_results_same_alone_and_in_mix(need_gpu):
    from pfvs_amd.instances import serving_request
    reqs = [serving_request(r, n_max=96) for r in range(10)]
    srv = Server(max_batch_p=256).start()
    jobs = []

    def client():
        for inst, ps, qmax in reqs:
            jobs.append(srv.submit(inst.data.kappa_cob, inst.data.M_lattice(), inst.Q, ps, qmax))
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
    for (inst, ps, qmax), j in zip(reqs, jobs):
        Ks, Ms = run_alone(inst, ps, qmax)
        assert np.array_equal(Ks, j.result[0]) and np.array_equal(Ms, j.result[1]), j.id
