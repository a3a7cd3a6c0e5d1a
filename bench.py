"""Benchmark: FP nodes/s of the batched persistent coniZpM search on the
h11 = 20 member of the mixed batch (G(20, 250, 5), 4096 seeded p-vectors,
cone coordinates <= 12, per-p dilation uniform in 1..4), against the plain
one-launch-per-batch form (one wavefront per p-vector, no work sharing) and
the one-thread-per-p-vector form, with the oracle's CPU path timed on a
bounded sample of the same p-vectors (one thread and all threads).

One step = one full coniZpM pass over the whole batch, inputs resident in
HBM: preparation, search, the exact per-point filter, the canonical order
and the PFV assembly in the library, and the exact host finish of what it
leaves, ending with the final (K, M) rows on the host.  `value` is nodes
over the wall time of a step.  Prints one JSON line.  --workload mixed
(all four instances of the mixed batch in one launch), --workload large
(deep h11 = 30 trees, with the exact-host fallback rate) and --workload
serving (a Poisson stream through the continuous-batching server) print
the same line for those workloads.  The headline line also carries
coniPFVs/s on a workload that has PFVs (the Manwe scan).

Roofline of the search stage (the breadth-first split's k_expand launches
and the persistent k_search, batched form): VALU issue.  Its algorithmic
work W per pass is counted in lane-ops on the host path, which visits the
same nodes (asserted):
  W = sum_nodes [4k + 4k [i >= m0_level] + 4k r + 24]
    + sum_cand [12 + r (4 + 5 s)] + sum_leaf [4 d (d + 1) / 2 + 8]
(k = d - 1 - i coordinates set above a node at level i, r = 1 where an H
row leads at column i, s = binary-gcd iterations), and divided by the
stage's duration from HIP events on the stream it runs on; the peak is
CUs x 64 lanes x peak clock.  `prep_roofline` puts the preparation (k_prep)
against the same peak: its lane-ops, counted on the same host run (one per
element operation of the certificate, the two LLLs, the products, the HNF
and the factor), over the stage's duration, with the largest p-vector's
count beside them (one wavefront's latency bounds that stage).
--select I,J keeps only those p-vectors of the workload (for profiling
single preparations).
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


def full_pass(gpu, geoms, pf, pgeo, mode, **kw):
    """One step: everything coniZpM does for the batch -- preparation,
    search, the exact per-point filter, the canonical order and the PFV
    assembly in the library, and the exact host finish of any p-vector the
    library leaves (pstat 1 or pdone 0) -- from resident inputs to the final
    (K, M) rows.  Returns (result, number of PFVs, p-vectors finished on the
    host)."""
    from pfvs_amd import coni
    r = gpu.run(geoms, pf, pgeo, mode=mode, filter=True, sort=True, keep_points="survivors",
                assemble=True, **kw)
    n_pfv = len(r["pfv"]["pidx"])
    left = np.flatnonzero((r["pstat"] != 0) | (r["pdone"] == 0))
    for ip in left:                          # rare: exact host finish, one p-vector at a time
        g = geoms[pgeo[ip]]
        h = np.asarray(g["kappa"]).shape[0]
        sub = {k: r[k][r["pidx"] == ip] for k in ("M", "Kn", "q", "nonsing")}
        sub.update(pidx=np.zeros(len(sub["q"]), np.int64), pstat=r["pstat"][ip:ip + 1],
                   pcount=r["pcount"][ip:ip + 1])
        Ks, _ = coni.assemble(sub, pf[ip:ip + 1, 1:h], pf[ip:ip + 1, :h], g["kappa"], g["Mbasis"],
                              g["Q"], np.array([gpu.qmax_of(g["Q"], g["dilation"])]), g["M0min"])
        n_pfv += len(Ks)
    return r, n_pfv, len(left)


def time_mode(gpu, geoms, pf, pgeo, mode, steps, warmup, **kw):
    """Wall time of `steps` full passes (each returns after its results are
    on the host), after `warmup` untimed ones."""
    for _ in range(warmup):
        full_pass(gpu, geoms, pf, pgeo, mode, **kw)
    stats, walls = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        r, n_pfv, n_left = full_pass(gpu, geoms, pf, pgeo, mode, **kw)
        walls.append(time.perf_counter() - t1)
        stats.append(r["stats"])
    wall = (time.perf_counter() - t0) / steps
    return wall, stats, r, n_pfv, n_left, walls


def search_work(gpu, geoms, pf, pgeo):
    """Algorithmic lane-ops of the search over the batch (host path, count
    only), with its node count, the threads used and the host run's stats
    (prep_work, prep_work_max: the preparation's lane-ops, all p-vectors and
    the largest one)."""
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    r = gpu.run(geoms, pf, pgeo, mode="host", host_threads=threads, keep_points=False,
                filter=False)
    return r["stats"]["work"], r["stats"]["nodes"], threads, r["stats"]


TRAFFIC_FILE = os.path.join("profiles", "search_h20_hbm_traffic.md")


def measured_traffic(stage="search"):
    """HBM bytes of one pass's search stage (the split's k_expand launches
    and k_search) as recorded in the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), or
    None.  Not measured by this run."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_FILE)):
            if line.startswith(f"traffic_bytes_per_launch {stage}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(items, pf, budget_s=8.0, threads=1):
    """The oracle's CPU path (its lattice preparation and search, the points
    and their transform) over p-vectors of the same batch in a fixed seeded
    order, `threads` threads each taking the next p-vector, until budget_s
    seconds have elapsed.  Returns (nodes/s, p-vectors done, seconds)."""
    import threading
    from oracle import cpu
    order = np.random.default_rng(7).permutation(len(pf))
    lock = threading.Lock()
    state = {"next": 0, "nodes": 0, "done": 0}
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                k = state["next"]
                if k >= len(order) or time.perf_counter() - t0 > budget_s:
                    return
                state["next"] += 1
            i = order[k]
            kap, Mb, Q, qmax = items[i]
            res = cpu.coni_batch(kap, Mb, pf[i:i + 1, :kap.shape[0]], Q, qmax)
            with lock:
                state["nodes"] += res[5][0]
                state["done"] += 1
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    t = time.perf_counter() - t0
    return state["nodes"] / t, state["done"], t


def pfv_rate(gpu, steps):
    """coniPFVs/s on a workload that has PFVs: the Manwe geometry's scan
    (every primitive p-vector of the smallest box holding 10,000 of them,
    dilation 50, M0min 13, Q = h11 + h21 + 4; the reference's
    tests/test_manwe.py known answer is 111 PFVs), as full coniZpM calls on
    the device."""
    from pfvs_amd import coniZpM
    from pfvs_amd.cydata import CYData
    from pfvs_amd.pvectors import pvecs
    e = json.load(open(os.path.join(ROOT, "tests", "golden", "paper_coni.json")))[0]
    data = CYData(e["h21"], e["kappa"], e["c2"], e["H"], coni_curve=e["coni_curve"],
                  coni_cob=e["cob"])
    ps = pvecs(data, min_N_pts=10_000)
    call = lambda: coniZpM(data, ps, M0min=13, ellipsoid_dilation=50, max_N_pfvs=10_000_000,
                           device="gpu")
    call()
    t0 = time.perf_counter()
    for _ in range(steps):
        Ks, Ms = call()
    t = (time.perf_counter() - t0) / steps
    return {"value": len(Ks) / t, "unit": "PFVs/s", "pfvs": len(Ks), "ms_per_call": t * 1e3,
            "workload": f"Manwe (h11 = 8, h21 = 150) scan: {len(ps)} p-vectors, dilation 50, "
                        "coniZpM(device='gpu') end to end"}


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
    ap.add_argument("--no-pfv", action="store_true", help="skip the coniPFVs/s workload")
    ap.add_argument("--cpu-seconds", type=float, default=16.0,
                    help="CPU baseline budget (half single-thread, half all threads)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-parameter autotuner (library defaults)")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="fixed search launch setting instead of the autotuner's (with --budget)")
    ap.add_argument("--budget", type=int, default=0,
                    help="fixed donation size instead of the autotuner's (with --waves-per-cu)")
    ap.add_argument("--split", type=int, default=0,
                    help="fixed split target (nodes per wavefront; < 0 none) with --waves-per-cu")
    ap.add_argument("--select", default="", metavar="I,J,...",
                    help="keep only these p-vectors of the workload (indices into its batch; "
                         "for profiling single preparations)")
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
    if args.select:
        sel = [int(x) for x in args.select.split(",")]
        pf, pgeo, items = pf[sel], pgeo[sel], [items[i] for i in sel]
        desc += f"; only p-vectors {args.select}"
    print(f"# workload {args.workload}: {len(pf)} p-vectors, {len(geoms)} geometry entries",
          file=sys.stderr, flush=True)
    params = {}
    if args.waves_per_cu > 0 or args.budget > 0:
        params = {"waves_per_cu": args.waves_per_cu, "budget": args.budget, "stack_cap": 0,
                  "split": args.split}
    elif not args.no_tune:
        from pfvs_amd import autotune
        best = autotune.tune(geoms, pf, pgeo, use_cache=False)
        params = {k: best[k] for k in ("waves_per_cu", "budget", "stack_cap", "split")}
        print(f"# autotuned: {params} ({best['ms_search']:.2f} ms search)", file=sys.stderr,
              flush=True)
    wall_b, st_b, r_b, pfv_b, left_b, walls_b = time_mode(gpu, geoms, pf, pgeo, "batched",
                                                          args.steps, args.warmup, **params)
    nodes = st_b[-1]["nodes"]
    mean = lambda k, st: float(np.mean([s[k] for s in st]))
    ms_dev, ms_search = mean("ms_total", st_b), mean("ms_search", st_b)
    print(f"# batched: {nodes:.0f} nodes, full pass {wall_b*1e3:.1f} ms (device {ms_dev:.2f} ms, "
          f"search {ms_search:.2f} ms), {pfv_b} PFVs, host-finished {left_b}", file=sys.stderr,
          flush=True)
    metric = {"headline": "FP nodes/s (batched persistent coniZpM, h11=20)",
              "mixed": "FP nodes/s (batched persistent coniZpM, mixed h11 = 6/10/14/20)",
              "large": "FP nodes/s (batched persistent coniZpM, deep h11 = 30 trees)"}[args.workload]
    sb = st_b[-1]
    out = {
        "metric": metric,
        "value": nodes / wall_b,
        "unit": "nodes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall_b * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64+f64", "data": "synthetic",
        "config": {"workload": desc, "parallelism": "none",
                   "step": "one full coniZpM pass over the batch from resident inputs to the "
                           "final (K, M) rows: preparation, search, exact filter, canonical "
                           "order, PFV assembly, host finish of leftovers (wall clock)",
                   "search_params": params or "library defaults"},
        "detail": {"nodes": nodes, "points": sb["points"], "kept": sb["kept"],
                   "pfvs": pfv_b, "survivors_copied": sb["survivors"],
                   "device_ms": ms_dev, "search_ms": ms_search,
                   "prep_ms": mean("ms_prep", st_b), "post_ms": mean("ms_post", st_b),
                   "finish_ms": mean("ms_finish", st_b), "d2h_ms": mean("ms_d2h", st_b),
                   "upload_ms": mean("ms_upload", st_b), "alloc_ms": mean("ms_alloc", st_b),
                   "library_wall_ms": mean("ms_wall", st_b),
                   "pass_ms_min": min(walls_b) * 1e3, "pass_ms_max": max(walls_b) * 1e3,
                   "device_nodes_per_s": nodes / (ms_dev * 1e-3),
                   "search_nodes_per_s": nodes / (ms_search * 1e-3),
                   "passes": sb["passes"], "host_p": sb["host_p"], "host_finished": left_b,
                   "cand": sb["cand"], "steps": sb["steps"],
                   "wait_frac": sb["wait_cycles"] / max(sb["wave_cycles"], 1),
                   "donated": sb["donated"], "fetches": sb["fetches"],
                   "coniPFVs_per_s": pfv_b / wall_b,
                   "graph_segments_held": sb["graphs"], "graph_replays": sb["graph_hits"],
                   "graph_captures": sb["graph_captures"],
                   "host_fallback_rate": sb["host_p"] / max(len(pf), 1)},
    }
    work, wnodes, wthreads, hst = search_work(gpu, geoms, pf, pgeo)
    assert wnodes == nodes, "host count and device search visited different nodes"
    cus, clk_khz = gpu.device_facts(0)
    peak = cus * 64 * clk_khz * 1e3                 # lane-ops/s
    achieved = work / (ms_search * 1e-3)
    traffic = measured_traffic("search") if args.workload == "headline" else None
    out["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                       "unit": "T lane-op/s", "frac": achieved / peak,
                       "traffic": traffic,
                       "traffic_source": (f"{TRAFFIC_FILE} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE;"
                                          " recorded, not measured by this run)") if traffic else None,
                       "kernel": "search stage: k_expand (breadth-first split) + k_search<batched>",
                       "work_per_launch": work,
                       "ms_per_launch": ms_search,
                       "peak_basis": f"{cus} CUs x 64 lanes x {clk_khz / 1e6:.2f} GHz"}
    # k_prep: one wavefront per p-vector; the pass's lane-ops against the same
    # peak, and the largest p-vector's lane-ops (its wavefront's latency is
    # the stage's floor: the rest of the GPU is idle once short ones finish)
    ms_prep = mean("ms_prep", st_b)
    out["prep_roofline"] = {"bound": "latency of one wavefront", "kernel": "k_prep",
                            "work_per_launch": hst["prep_work"],
                            "work_largest_p": hst["prep_work_max"],
                            "ms_per_launch": ms_prep,
                            "achieved": hst["prep_work"] / (ms_prep * 1e-3) / 1e12,
                            "peak": peak / 1e12, "unit": "T lane-op/s",
                            "frac": hst["prep_work"] / (ms_prep * 1e-3) / peak}
    if not args.no_plain:
        for mode in ("plain", "lane"):
            if mode == "lane" and args.workload != "headline":
                continue
            wall_p, st_p, _, pfv_p, _, _ = time_mode(gpu, geoms, pf, pgeo, mode, 1, 0)
            assert st_p[-1]["nodes"] == nodes, f"{mode} and batched forms visited different nodes"
            assert pfv_p == pfv_b
            d = out["detail"]
            d[f"{mode}_pass_ms"] = wall_p * 1e3
            d[f"{mode}_search_ms"] = st_p[-1]["ms_search"]
            d[f"{mode}_nodes_per_s"] = nodes / wall_p
            d[f"batched_over_{mode}"] = (nodes / wall_b) / (nodes / wall_p)
            d[f"batched_over_{mode}_search"] = st_p[-1]["ms_search"] / ms_search
            print(f"# {mode}: full pass {wall_p*1e3:.1f} ms, search {st_p[-1]['ms_search']:.2f} ms",
                  file=sys.stderr, flush=True)
    if not args.no_cpu:
        nth = max(1, min(16, len(os.sched_getaffinity(0))))
        v1, done1, t1 = cpu_baseline(items, pf, args.cpu_seconds / 2, threads=1)
        vn, donen, tn = cpu_baseline(items, pf, args.cpu_seconds / 2, threads=nth)
        out["cpu_baseline"] = {"value": vn, "unit": "nodes/s", "cores": nth, "kind": "port",
                               "sample": f"{donen} p-vectors of the same batch in {tn:.1f} s on "
                                         f"{nth} threads (seeded order; the oracle's C "
                                         f"preparation and search)",
                               "cpu": cpu_model(),
                               "single_thread": {"value": v1, "cores": 1,
                                                 "sample": f"{done1} p-vectors in {t1:.1f} s"}}
    if args.workload == "headline" and not args.no_pfv:
        out["coniPFVs"] = pfv_rate(gpu, max(1, min(args.steps, 5)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
