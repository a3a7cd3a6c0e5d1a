"""Continuous-batching serving loop for coniZpM jobs on one device.

A job is one coniZpM request: a geometry (kappa, Mbasis, Q, M0min) and its
p-vectors, each with a search bound qmax.  `submit` is thread-safe and
returns at once.  A scheduler thread runs the device pipeline in
iterations; every iteration

  1. admits pending p-vectors of the queued jobs in arrival order, as many
     as the per-iteration p-vector cap and the HBM footprint budget allow
     (a large job is split over several iterations, a job that arrives while
     an iteration is on the device joins the next one),
  2. runs them all in one persistent launch: jobs of different h11 and
     bounds share the preparation, the search's frontier and the filter
     (pfa_coni_run with one geometry entry per job and bound),
  3. hands each job the rows of its p-vectors.

Two iterations can be in flight: the next one is admitted while the
current one runs, waits for the device, and runs its kernels while the
library assembles the current one's results on the host (the library
releases the device before that assembly).

A job completes when its last p-vector has come back; its (Ks, Ms) are
then assembled by coniZpM's own host post-processing (pfvs_amd.coni.
assemble), so a job's result does not depend on what it was batched with.
Per job the server records arrival, first admission and completion times;
`stats()` reports p50/p99 latency and the aggregate node rate.
"""

from __future__ import annotations

import collections
import concurrent.futures
import threading
import time

import numpy as np

from . import coni, gpu


class Job:
    def __init__(self, jid, kappa, Mbasis, Q, M0min, ps, qmax):
        self.id = jid
        self.kappa = np.ascontiguousarray(kappa, dtype=np.int64)
        self.Mbasis = np.ascontiguousarray(Mbasis, dtype=np.int64)
        self.h = self.kappa.shape[0]
        self.Q, self.M0min = int(Q), int(M0min)
        self.ps = np.asarray(ps, dtype=np.int64).reshape(-1, self.h - 1)
        self.p_full = np.hstack([np.zeros((len(self.ps), 1), np.int64), self.ps])
        self.qmax = np.broadcast_to(np.asarray(qmax, dtype=np.int64), (len(self.ps),)).copy()
        self.next = 0                     # first p-vector not yet admitted
        self.returned = 0
        self.rows = []                    # (first p-vector, {M, Kn, q, pidx, nonsing}) pieces
        self.pstat = np.zeros(len(self.ps), np.int8)
        self.t_arrival = time.perf_counter()
        self.t_start = None
        self.t_done = None
        self.result = None
        self.points = None
        self.points_sum = None
        self.pdone = None                 # per p-vector: assembled by the library
        self.error = None
        self.done = threading.Event()

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError(f"job {self.id} not done")
        if self.error is not None:
            raise self.error
        return self.result


def record_bytes(h):
    """Device bytes of one p-vector's record (the library's layout at d = h - 1)."""
    d = h - 1
    return 8 * (20 + 3 * d * d + 2 * d + 2 * (d + 1) * d + 2 * (d + 1) * d)


def points_checksum(r, h):
    """Order-sensitive 64-bit checksum of lattice-point rows (M, Kn, q,
    pidx, nonsing; the first h columns of M and Kn): per row a weighted sum
    of its entries, then the rows weighted by position, all mod 2^64."""
    n = len(r["q"])
    if n == 0:
        return 0
    cols = np.hstack([np.asarray(r["M"])[:, :h], np.asarray(r["Kn"])[:, :h],
                      np.asarray(r["q"]).reshape(n, 1), np.asarray(r["pidx"]).reshape(n, 1),
                      np.asarray(r["nonsing"]).reshape(n, 1).astype(np.int64)]).astype(np.uint64)
    w = np.arange(1, cols.shape[1] + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15) | \
        np.uint64(1)
    with np.errstate(over="ignore"):
        rows = (cols * w).sum(axis=1, dtype=np.uint64)
        pos = np.arange(n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        return int((rows * pos).sum(dtype=np.uint64))


class Server:
    def __init__(self, device=0, max_batch_p=16384, hbm_budget=32 << 30, points_per_p=2000,
                 params=None, runner=None, max_N_pfvs=1_000_000_000, finish_workers=4,
                 keep_points=False, pipeline=2, reserve=None):
        """runner(geoms, ps, pgeo) -> pipeline result dict (default: gpu.run,
        batched, filter on, sorted rows).  Completed jobs are assembled on
        `finish_workers` host threads while the next iteration runs.  With
        keep_points=True a job also keeps its lattice-point rows (job.points:
        M, Kn, q, pidx into its own p-vectors, nonsing); with "checksum" it
        keeps only points_checksum of them (job.points_sum), computed after
        the job's completion time is taken.

        Up to `pipeline` iterations are in flight: the library runs their
        device work in order on the device and releases the device before
        its host-side assembly of results, so that assembly overlaps the
        next iteration's kernels.  reserve = (n_p, n_points, h_max) grows
        the library's buffers at start (gpu.reserve)."""
        self.device = device
        self.max_batch_p = int(max_batch_p)
        self.hbm_budget = int(hbm_budget)
        self.points_per_p = points_per_p      # footprint estimate, refined as results come in
        self.params = dict(params or {})
        self.max_N = max_N_pfvs
        self.runner = runner or self._run
        self._q = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self._ids = 0
        self._thread = None
        self.jobs = []
        self.iterations = []                  # (t0, t1, n_p, n_jobs, nodes, stage ms, h11s)
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=finish_workers)
        self.pipeline = max(1, int(pipeline))
        self._slots = threading.Semaphore(self.pipeline)
        self._runpool = concurrent.futures.ThreadPoolExecutor(max_workers=self.pipeline)
        self.reserve = reserve
        self.finish_s = 0.0
        self.keep_points = keep_points

    # ------------------------------------------------------------ client side
    def submit(self, kappa, Mbasis, Q, ps, qmax, M0min=13):
        with self._cv:
            job = Job(self._ids, kappa, Mbasis, Q, M0min, ps, qmax)
            self._ids += 1
            self.jobs.append(job)
            if len(job.ps) == 0:
                job.result = (np.zeros((0, job.h), np.int64), np.zeros((0, job.h), np.int64))
                job.t_start = job.t_done = job.t_arrival
                job.done.set()
            else:
                self._q.append(job)
                self._cv.notify()
        return job

    def start(self):
        if self.reserve is not None:
            gpu.reserve(*self.reserve, device=self.device)
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()
        return self

    def stop(self, timeout=None):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread:
            self._thread.join(timeout)
        self._runpool.shutdown(wait=True)
        self._pool.shutdown(wait=True)

    # ------------------------------------------------------------- scheduler
    def _run(self, geoms, ps, pgeo):
        return gpu.run(geoms, ps, pgeo, mode="batched", device=self.device, filter=True,
                       sort=True, assemble=True, **self.params)

    def _admit(self):
        """The (job, lo, hi) pieces of the next iteration, in arrival order."""
        pieces, used_p, used_b = [], 0, 0
        with self._cv:
            while not self._q and not self._stop:
                self._cv.wait()
            for job in self._q:
                per_p = record_bytes(job.h) + self.points_per_p * (8 * 2 * job.h + 24 + 4 * job.h)
                room_p = self.max_batch_p - used_p
                room_b = (self.hbm_budget - used_b) // max(int(per_p), 1)
                k = int(min(len(job.ps) - job.next, room_p, room_b))
                if k <= 0:
                    break
                pieces.append((job, job.next, job.next + k))
                job.next += k
                used_p += k
                used_b += k * int(per_p)
                if job.t_start is None:
                    job.t_start = time.perf_counter()
            while self._q and self._q[0].next == len(self._q[0].ps):
                self._q.popleft()
        return pieces

    def _loop(self):
        while True:
            self._slots.acquire()
            pieces = self._admit()
            if not pieces:
                self._slots.release()
                if self._stop:
                    return
                continue
            self._runpool.submit(self._iterate, pieces)

    def _iterate(self, pieces):
        try:
            self._run_iteration(pieces)
        finally:
            self._slots.release()

    def _run_iteration(self, pieces):
        MH = gpu.max_h11()
        n = sum(hi - lo for _, lo, hi in pieces)
        pf = np.zeros((n, MH), np.int64)
        pgeo = np.zeros(n, np.int32)
        geoms, owner, base = [], [], 0
        for job, lo, hi in pieces:
            k = hi - lo
            pf[base:base + k, :job.h] = job.p_full[lo:hi]
            qs = job.qmax[lo:hi]
            for v in np.unique(qs):               # one geometry entry per bound
                pgeo[base + np.flatnonzero(qs == v)] = len(geoms)
                geoms.append(dict(kappa=job.kappa, Mbasis=job.Mbasis, Q=job.Q,
                                  dilation=int(v) / job.Q, M0min=job.M0min))
            owner.append((job, lo, base, k))
            base += k
        t0 = time.perf_counter()
        try:
            r = self.runner(geoms, pf, pgeo)
        except Exception as e:              # the whole iteration failed: fail its jobs
            with self._cv:
                for job, lo, b0, k in owner:
                    if job.error is None:
                        job.error = e
                        job.t_done = time.perf_counter()
                        job.done.set()
            return
        t1 = time.perf_counter()
        st = r["stats"]
        self.iterations.append((t0, t1, n, len(owner), float(st["nodes"]),
                                {k: round(float(st.get(k, 0.0)), 1) for k in
                                 ("ms_prep", "ms_roots", "ms_search", "ms_post", "ms_total",
                                  "ms_alloc", "ms_d2h", "ms_finish", "ms_wall",
                                  "attempts")},
                                sorted({job.h for job, _, _ in pieces})))
        pidx = np.asarray(r["pidx"])
        self.points_per_p = 0.5 * self.points_per_p + 0.5 * max(len(pidx) / n, 1.0)
        lib = "pdone" in r                  # the library assembled what it could
        fp = np.asarray(r["pfv"]["pidx"]) if lib else None
        done = []
        with self._cv:
            for job, lo, b0, k in owner:
                a, b = np.searchsorted(pidx, [b0, b0 + k])
                piece = dict(M=r["M"][a:b], Kn=r["Kn"][a:b], q=r["q"][a:b],
                             pidx=pidx[a:b] - b0 + lo, nonsing=r["nonsing"][a:b])
                if lib:
                    fa, fb = np.searchsorted(fp, [b0, b0 + k])
                    piece.update(pfK=r["pfv"]["K"][fa:fb], pfM=r["pfv"]["M"][fa:fb],
                                 pfp=fp[fa:fb] - b0 + lo)
                    if job.pdone is None:
                        job.pdone = np.zeros(len(job.ps), np.int8)
                    job.pdone[lo:lo + k] = r["pdone"][b0:b0 + k]
                job.rows.append((lo, piece))
                job.pstat[lo:lo + k] = r["pstat"][b0:b0 + k]
                job.returned += k
                if job.returned == len(job.ps) and job.error is None:
                    done.append(job)
        for job in done:
            self._pool.submit(self._finish, job)

    def _finish(self, job):
        t0 = time.perf_counter()
        rr = None
        try:
            rows = [p for _, p in sorted(job.rows, key=lambda t: t[0])]
            rr = {k: np.concatenate([p[k] for p in rows]) for k in ("q", "pidx", "nonsing")}
            for k in ("M", "Kn"):
                rr[k] = np.vstack([p[k] for p in rows])
            rr["pstat"] = job.pstat
            if job.pdone is not None:
                rr["pdone"] = job.pdone
                rr["pfv"] = dict(K=np.vstack([p["pfK"] for p in rows]),
                                 M=np.vstack([p["pfM"] for p in rows]),
                                 pidx=np.concatenate([p["pfp"] for p in rows]))
            job.result = coni.assemble(rr, job.ps, job.p_full, job.kappa, job.Mbasis, job.Q,
                                       job.qmax, job.M0min, self.max_N)
        except Exception as e:
            job.error = e
        job.rows = []
        job.t_done = time.perf_counter()
        self.finish_s += job.t_done - t0
        if job.error is None and self.keep_points == "checksum":
            job.points_sum = points_checksum(rr, job.h)
        elif job.error is None and self.keep_points:
            job.points = rr
        job.done.set()

    # ---------------------------------------------------------------- report
    def stats(self):
        done = [j for j in self.jobs if j.t_done is not None and j.error is None]
        lat = np.array([j.t_done - j.t_arrival for j in done]) if done else np.zeros(0)
        busy = sum(t1 - t0 for t0, t1, *_ in self.iterations)
        nodes = sum(it[4] for it in self.iterations)
        span = (max(j.t_done for j in done) - min(j.t_arrival for j in done)) if done else 0.0
        return {
            "jobs": len(done), "errors": sum(j.error is not None for j in self.jobs),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "mean_ms": float(lat.mean() * 1e3) if len(lat) else None,
            "iterations": len(self.iterations),
            "jobs_per_iteration": float(np.mean([it[3] for it in self.iterations]))
            if self.iterations else 0.0,
            "nodes": nodes, "busy_s": busy, "finish_s": self.finish_s, "span_s": span,
            "nodes_per_s": nodes / span if span > 0 else 0.0,
            "iterations_detail": [dict(ms=round((t1 - t0) * 1e3, 1), p=n_p, jobs=n_j,
                                       h=list(hs), nodes=nd, **ms)
                                  for t0, t1, n_p, n_j, nd, ms, hs in self.iterations],
        }
