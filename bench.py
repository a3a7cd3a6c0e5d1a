"""Benchmark: FP nodes/s of the batched persistent coniZpM search against the
plain one-thread-per-p form, on the h11 = 20 member of the mixed batch
(G(20, 250, 5), 4096 seeded p-vectors, cone coordinates <= 12, per-p
dilation uniform in 1..4), with the oracle's CPU path timed on a bounded
sample of the same p-vectors.

One step = one pass of the device pipeline (preparation, frontier roots,
search, post-processing with the exact filter) over the whole batch, inputs
resident in HBM.  Prints one JSON line.  --workload mixed (all four
instances of the mixed batch in one launch) and --workload large (deep
h11 = 30 trees, reporting the exact-host fallback rate) print the same
line for those batches.

Roofline of the search kernel (k_batched): VALU issue.  Its algorithmic
work W per launch is counted in lane-ops on the host path, which visits the
same nodes (asserted):
  W = sum_nodes [4k + 4k [i >= m0_level] + 4k r + 24]
    + sum_cand [12 + r (4 + 5 s)] + sum_leaf [4 d (d + 1) / 2 + 8]
(k = d - 1 - i coordinates set above a node at level i, r = 1 where an H
row leads at column i, s = binary-gcd iterations), and divided by the
kernel's duration from HIP events on the stream it runs on; the peak is
CUs x 64 lanes x peak clock.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P_VALU = None   # lane-ops/s, from the device's CU count and clock


def _dilation_geoms(inst, dil):
    geoms, pgeo = [], np.zeros(len(dil), np.int32)
    kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
    for k, D in enumerate(sorted(set(int(x) for x in dil))):
        geoms.append(dict(kappa=kap, Mbasis=Mb, Q=inst.Q, dilation=D, M0min=13))
        pgeo[dil == D] = k
    return geoms, pgeo


def workload(name="headline", n_per=4096):
    """(description, geoms, pf (p-vectors with the coni entry), pgeo, cpu
    items: per p-vector (kappa, Mbasis, Q, qmax))."""
    from pfvs_amd.instances import (BATCH_MEMBERS, config_batch, config_large_tree,
                                    geoms_for_qmax)
    if name in ("headline", "mixed"):
        members = [m for m in BATCH_MEMBERS if m[0] == 20] if name == "headline" else BATCH_MEMBERS
        geoms, pfs, pgeos, items = [], [], [], []
        for inst, ps, dil in config_batch(n_per=n_per, members=members):
            g, pg = _dilation_geoms(inst, dil)
            pgeos.append(pg + len(geoms))
            geoms += g
            pfs.append(np.hstack([np.zeros((len(ps), 1), np.int64), ps]))
            kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
            items += [(kap, Mb, inst.Q, int(np.floor(int(D) * inst.Q + 1e-7))) for D in dil]
        desc = ("G(20,250,5), %d p-vectors (cone coords <= 12, p_f >= 3), dilation uniform 1..4, "
                "M0min 13" % n_per if name == "headline" else
                "G(6,90,2)+G(10,150,3)+G(14,200,4)+G(20,250,5), %d p-vectors each (cone coords "
                "<= 12, p_f >= floor(4Q/338)), dilation uniform 1..4, M0min 13, one launch" % n_per)
        w = max(p.shape[1] for p in pfs)
        pf = np.vstack([np.pad(p, ((0, 0), (0, w - p.shape[1]))) for p in pfs])
        return desc, geoms, pf, np.concatenate(pgeos), items
    if name == "large":
        inst, ps, qmax = config_large_tree()
        geoms, pgeo = geoms_for_qmax(inst, qmax)
        kap, Mb = inst.data.kappa_cob, inst.data.M_lattice()
        items = [(kap, Mb, inst.Q, int(q)) for q in qmax]
        desc = ("G(30,300,6), %d p-vectors (cone coords <= 20, p_f <= 3), per-p qmax = 338 p_f + b, "
                "b uniform 8..14, M0min 13" % len(ps))
        return desc, geoms, np.hstack([np.zeros((len(ps), 1), np.int64), ps]), pgeo, items
    raise SystemExit(f"unknown workload {name}")


def time_mode(gpu, geoms, pf, pgeo, mode, steps, warmup, **kw):
    """Each gpu.run call returns after its stream has drained (results are
    copied back), so the wall clock brackets complete passes."""
    for _ in range(warmup):
        gpu.run(geoms, pf, pgeo, mode=mode, keep_points=False, **kw)
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        r = gpu.run(geoms, pf, pgeo, mode=mode, keep_points=False, **kw)
        stats.append(r["stats"])
    wall = (time.perf_counter() - t0) / steps
    return wall, stats, r


def search_work(gpu, geoms, pf, pgeo):
    """Algorithmic lane-ops of the search over the batch (host path, count
    only), with its node count and the threads used."""
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    r = gpu.run(geoms, pf, pgeo, mode="host", host_threads=threads, keep_points=False,
                filter=False)
    return r["stats"]["work"], r["stats"]["nodes"], threads


def measured_traffic(kernel="k_batched"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated with a copy kernel), or
    None."""
    path = os.path.join(ROOT, "profiles", "k_batched_h20_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(items, pf, budget_s=15.0):
    """The oracle (single thread) over p-vectors of the same batch, in a
    fixed seeded order, until budget_s seconds have elapsed."""
    from oracle import cpu
    order = np.random.default_rng(7).permutation(len(pf))
    nodes = done = 0
    t0 = time.perf_counter()
    for i in order:
        kap, Mb, Q, qmax = items[i]
        M, K, q, pidx, st, cnt = cpu.coni_batch(kap, Mb, pf[i:i + 1, :kap.shape[0]], Q, qmax)
        nodes += cnt[0]
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t = time.perf_counter() - t0
    return nodes / t, done, t


def serving_bench(gpu, args):
    """Requests r = 0 .. R-1 of the serving stream (h11 uniform in 5..40,
    log-uniform sizes up to 4096 p-vectors, band-limited bounds) arrive as a
    seeded Poisson process at args.rate per second; every response is
    checked against the same request run alone."""
    from pfvs_amd import coni
    from pfvs_amd.instances import geoms_for_qmax, serving_request
    from pfvs_amd.serve import Server, points_checksum
    reqs = [serving_request(r) for r in range(args.requests)]
    data = [(inst.data.kappa_cob, inst.data.M_lattice(), inst.Q, ps, qmax)
            for inst, ps, qmax in reqs]
    gaps = np.random.default_rng(11).exponential(1.0 / args.rate, len(reqs))
    srv = Server(keep_points="checksum", reserve=(16384, 4_000_000, 41)).start()
    srv.submit(*data[0]).wait()                  # warm: code and buffers
    srv.jobs.clear()
    srv.iterations.clear()
    jobs = []
    t0 = time.perf_counter()
    for k, dt in enumerate(gaps):
        t0 += dt
        while time.perf_counter() < t0:
            time.sleep(min(0.001, max(0.0, t0 - time.perf_counter())))
        jobs.append(srv.submit(*data[k]))
    for j in jobs:
        j.wait(timeout=600)
    srv.stop(timeout=60)
    st = srv.stats()
    exact = exact_pts = n_pts = 0
    for (inst, ps, qmax), (kap, Mb, Q, _, _), j in zip(reqs, data, jobs):
        geoms, pgeo = geoms_for_qmax(inst, qmax)
        pf = np.hstack([np.zeros((len(ps), 1), np.int64), ps])
        r = gpu.run(geoms, pf, pgeo, mode="batched", filter=True, sort=True)
        Ks, Ms = coni.assemble(r, ps, pf, kap, Mb, Q, qmax, 13)
        exact += int(np.array_equal(Ks, j.result[0]) and np.array_equal(Ms, j.result[1]))
        exact_pts += int(points_checksum(r, inst.h11) == j.points_sum)
        n_pts += len(r["q"])
    hs = [inst.h11 for inst, _, _ in reqs]
    return {
        "metric": "coniZpM serving: aggregate FP nodes/s under a Poisson stream",
        "value": st["nodes_per_s"], "unit": "nodes/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": st["span_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64+f64", "data": "synthetic",
        "config": {"workload": f"{len(reqs)} requests G(h, 10h, r), h uniform 5..40 (here "
                               f"{min(hs)}..{max(hs)}), N log-uniform 1..4096, cone coords <= 16, "
                               f"p_f <= 3, qmax = 338 p_f + b (b <= band(h)); Poisson "
                               f"{args.rate}/s", "parallelism": "none",
                   "server": {"max_batch_p": srv.max_batch_p, "pipeline": srv.pipeline,
                              "reserve": list(srv.reserve)}},
        "detail": dict(st, p_vectors=int(sum(len(ps) for _, ps, _ in reqs)),
                       bit_exact_alone_vs_mix=f"{exact}/{len(reqs)}",
                       points_checksum_alone_vs_mix=f"{exact_pts}/{len(reqs)}",
                       lattice_points=n_pts),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-per", type=int, default=4096)
    ap.add_argument("--workload", default="headline",
                    choices=("headline", "mixed", "large", "serving"),
                    help="headline: the h11 = 20 batch (the metric); mixed: all four "
                         "instances in one launch; large: deep h11 = 30 trees; serving: a "
                         "Poisson stream of coniZpM requests through the continuous-batching "
                         "server")
    ap.add_argument("--requests", type=int, default=64, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=40.0, help="serving: arrivals per second")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-parameter autotuner (library defaults)")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="fixed search launch setting instead of the autotuner's (with --budget)")
    ap.add_argument("--budget", type=int, default=0,
                    help="fixed donation size instead of the autotuner's (with --waves-per-cu)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only time a plain device copy of MIB MiB (the reference "
                         "kernel for HBM traffic counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    from pfvs_amd import gpu
    if not gpu.available():
        raise SystemExit("HIP backend unavailable: " + gpu.unavailable_reason())
    if args.copy_probe:
        nb = args.copy_probe << 20
        ms = gpu.copy_probe(nb, reps=args.steps)
        print(json.dumps({"copy_bytes": nb, "copies": args.steps + 1, "ms_per_copy": ms,
                          "GB_per_s": 2 * nb / (ms * 1e-3) / 1e9}))
        return
    if args.workload == "serving":
        print(json.dumps(serving_bench(gpu, args)))
        return
    desc, geoms, pf, pgeo, items = workload(args.workload, args.n_per)
    print(f"# workload {args.workload}: {len(pf)} p-vectors, {len(geoms)} geometry entries",
          file=sys.stderr, flush=True)
    params = {}
    if args.waves_per_cu > 0 or args.budget > 0:
        params = {"waves_per_cu": args.waves_per_cu, "budget": args.budget, "stack_cap": 0}
    elif not args.no_tune:
        from pfvs_amd import autotune
        best = autotune.tune(geoms, pf, pgeo, use_cache=False)
        params = {k: best[k] for k in ("waves_per_cu", "budget", "stack_cap")}
        print(f"# autotuned: {params} ({best['ms_search']:.2f} ms search)", file=sys.stderr,
              flush=True)
    wall_b, st_b, r_b = time_mode(gpu, geoms, pf, pgeo, "batched", args.steps, args.warmup,
                                  **params)
    nodes = st_b[-1]["nodes"]
    ms_dev = np.mean([s["ms_total"] for s in st_b])
    ms_search = np.mean([s["ms_search"] for s in st_b])
    print(f"# batched: {nodes:.0f} nodes, device {ms_dev:.2f} ms, search {ms_search:.2f} ms, "
          f"wall {wall_b*1e3:.1f} ms, passes {st_b[-1]['passes']:.0f}", file=sys.stderr, flush=True)
    metric = {"headline": "FP nodes/s (batched persistent coniZpM, h11=20)",
              "mixed": "FP nodes/s (batched persistent coniZpM, mixed h11 = 6/10/14/20)",
              "large": "FP nodes/s (batched persistent coniZpM, deep h11 = 30 trees)"}[args.workload]
    out = {
        "metric": metric,
        "value": nodes / (ms_dev * 1e-3),
        "unit": "nodes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_dev,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64+f64", "data": "synthetic",
        "config": {"workload": desc, "parallelism": "none",
                   "search_params": params or "library defaults"},
        "detail": {"nodes": nodes, "points": st_b[-1]["points"], "kept": st_b[-1]["kept"],
                   "pfvs": st_b[-1]["expansions"], "search_ms": ms_search,
                   "prep_ms": float(np.mean([s["ms_prep"] for s in st_b])),
                   "post_ms": float(np.mean([s["ms_post"] for s in st_b])),
                   "passes": st_b[-1]["passes"], "host_p": st_b[-1]["host_p"],
                   "cand": st_b[-1]["cand"], "steps": st_b[-1]["steps"],
                   "wait_frac": st_b[-1]["wait_cycles"] / max(st_b[-1]["wave_cycles"], 1),
                   "donated": st_b[-1]["donated"], "fetches": st_b[-1]["fetches"],
                   "coniPFVs_per_s": st_b[-1]["expansions"] / (ms_dev * 1e-3),
                   "host_fallback_rate": st_b[-1]["host_p"] / max(len(pf), 1),
                   "wall_ms": wall_b * 1e3},
    }
    work, wnodes, wthreads = search_work(gpu, geoms, pf, pgeo)
    assert wnodes == nodes, "host count and device search visited different nodes"
    cus, clk_khz = gpu.device_facts(0)
    peak = cus * 64 * clk_khz * 1e3                 # lane-ops/s
    achieved = work / (ms_search * 1e-3)
    out["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                       "unit": "T lane-op/s", "frac": achieved / peak,
                       "traffic": measured_traffic("k_batched") if args.workload == "headline" else None,
                       "kernel": "k_batched", "work_per_launch": work,
                       "ms_per_launch": ms_search,
                       "peak_basis": f"{cus} CUs x 64 lanes x {clk_khz / 1e6:.2f} GHz"}
    if not args.no_plain:
        wall_p, st_p, _ = time_mode(gpu, geoms, pf, pgeo, "plain", 1, 0)
        ms_p = st_p[-1]["ms_total"]
        out["detail"]["plain_ms"] = ms_p
        out["detail"]["plain_nodes_per_s"] = st_p[-1]["nodes"] / (ms_p * 1e-3)
        out["detail"]["batched_over_plain"] = out["value"] / out["detail"]["plain_nodes_per_s"]
        assert st_p[-1]["nodes"] == nodes, "plain and batched forms visited different nodes"
    if not args.no_cpu:
        v, done, t = cpu_baseline(items, pf, args.cpu_seconds)
        out["cpu_baseline"] = {"value": v, "unit": "nodes/s", "cores": 1, "kind": "port",
                               "sample": f"{done} p-vectors of the same batch in {t:.1f} s "
                                         "(seeded order)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
