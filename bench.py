#!/usr/bin/env python3
"""Aggregate query-GTEPS of the concurrent multi-query engine, batched vs one
query at a time, on one MI355X.

Workload (default ``rmat22-mixed256``): RMAT scale 22 (4,194,304 vertices),
edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 3, self-loops dropped,
symmetrised, deduplicated; uint32 weights uniform in [1,255].  One batch =
256 queries: 160 BFS + 80 SSSP from distinct sources drawn uniformly over
vertices with degree >= 1, plus 16 WCC; 4 mask words per vertex.

A step is one batched run of the whole batch with inputs resident in HBM and
results written to HBM.  ``value`` = sum of m_q over the batch * steps / time,
where m_q = sum of out-degrees of the vertices query q reaches (E for WCC):
directed CSR edges, the same numerator for the batched and the
one-at-a-time runs.  One-at-a-time runs the same kernels with each query
alone.  Latency (batch workloads): from batch start to the query's result
resident in HBM; one-at-a-time queries queue behind each other from the same
start.  The CPU baseline times the oracle on a bounded sample of the same
batch on the host cores.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (graph, seed, wmin, wmax, n_bfs, n_sssp, n_wcc)
    #   graph ("rmat", scale, edge_factor) or ("grid", rows, cols, p_extra)
    "rmat22-mixed256": (("rmat", 22, 16), 3, 1, 255, 160, 80, 16),
    "rmat20-sssp64": (("rmat", 20, 16), 2, 1, 64, 0, 64, 0),
    "rmat16-bfs32": (("rmat", 16, 16), 1, 0, 0, 32, 0, 0),
    "grid4096-sssp512": (("grid", 4096, 4096, 0.01), 5, 1, 1000, 0, 512, 0),
}
# Serving: RMAT scale 24, w in [1,64], open-loop Poisson arrivals, 70% BFS / 30% SSSP.
SERVE = {"rmat24-serve": (("rmat", 24, 16), 4, 1, 64)}
METRIC = "aggregate query-GTEPS + p50/p99 latency, batched vs one-at-a-time, 1 GPU"
HBM_PEAK_GBS = 8000.0
COUNTED_RUNS = 3  # runs of the work-counter pass (algorithmic bytes per launch)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(values, p):
    s = sorted(values)
    return s[int(round(p * (len(s) - 1)))]


def cpu_baseline(h, kinds, sources, seconds_target):
    """Times the oracle (CPU restatement of the reference validators) on a
    bounded sample of the batch, one query per thread.  A first sample of one
    query per thread (in the batch's kind proportions) sets the rate; the
    sample then grows toward `seconds_target` of wall time, capped at the
    whole batch."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O

    threads = int(os.environ.get("CGP_CPU_THREADS", "0")) or min(16, os.cpu_count() or 1)
    g = O.G.of(h)
    by_kind = {k: [s for kk, s in zip(kinds, sources) if kk == k] for k in (0, 1, 2)}
    tot = len(kinds)

    def sample(nq):
        ks, ss = [], []
        for kind in (0, 1, 2):
            want = int(round(nq * len(by_kind[kind]) / tot))
            if kind == 2 and by_kind[2]:
                want = max(1, want)
            picked = by_kind[kind][:want]
            ks += [kind] * len(picked)
            ss += picked
        return ks, ss

    nq = max(1, threads)
    ks, ss = sample(nq)
    secs, _, m = O.run_queries(g, ks, ss, threads)
    if secs < seconds_target / 3 and len(ks) < tot:
        nq = min(tot, int(nq * seconds_target / max(secs, 1e-3) * 0.7))
        ks2, ss2 = sample(nq)
        if len(ks2) > len(ks):
            ks, ss = ks2, ss2
            secs, _, m = O.run_queries(g, ks, ss, threads)
    desc = "%d BFS + %d SSSP + %d WCC of the batch (first of each kind), one query per thread" % (
        ks.count(0), ks.count(1), ks.count(2))
    return {
        "value": float(np.sum(m)) / secs / 1e9,
        "unit": "query-GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": desc,
        "seconds": secs,
    }


def pmc_traffic(pmc_dir, kernel_prefix, out_rows, V, words, R, runs, launches_per_run):
    """HBM bytes per working launch of `kernel_prefix` from rocprofv3
    FETCH_SIZE / WRITE_SIZE CSVs (KiB per dispatch): each counter's total over
    the profiled run, scaled by the factor that makes the same counter read
    the known bytes of k_output (reads the words*V*R value rows, writes
    out_rows*V*4 bytes), divided by the working launches of that run (`runs`
    batch runs of `launches_per_run` launches that did work each, the same
    launches the algorithmic bytes are averaged over)."""
    import csv
    import glob
    tot = 0.0
    factors = {}
    for counter, known in (("FETCH_SIZE", 4.0 * words * V * R), ("WRITE_SIZE", 4.0 * out_rows * V)):
        vals, cal = [], []
        for f in glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") != counter:
                    continue
                name = r.get("Kernel_Name", "")
                v = float(r.get("Counter_Value", 0.0)) * 1024.0
                if "k_output" in name:
                    cal.append(v)
                elif kernel_prefix in name:
                    vals.append(v)
        if not vals or not cal:
            return None, None
        factor = known / (sum(cal) / len(cal))
        tot += factor * sum(vals)
        factors[counter.lower() + "_factor"] = factor
    return tot / max(1.0, runs * launches_per_run), factors


def make_graph(spec, seed, wmin, wmax):
    import cgp_amd
    if spec[0] == "rmat":
        return cgp_amd.HostCSR.rmat(spec[1], spec[2], seed=seed, wmin=wmin, wmax=wmax)
    return cgp_amd.HostCSR.grid(spec[1], spec[2], p_extra=spec[3], seed=seed, wmin=wmin, wmax=wmax)


def describe(spec, seed, wmin, wmax):
    if spec[0] == "rmat":
        return ("RMAT scale %d ef %d a=0.57 b=0.19 c=0.19 seed %d, symmetrised, dedup, w in [%d,%d]"
                % (spec[1], spec[2], seed, wmin, wmax))
    return ("%dx%d grid, 4-neighbour + one random edge per vertex w.p. %g, seed %d, w in [%d,%d]"
            % (spec[1], spec[2], spec[3], seed, wmin, wmax))


def serve_bench(args):
    """Open-loop Poisson serving: 70% BFS / 30% SSSP with uniform sources;
    five arrival rates relative to the measured saturation throughput; the
    admission cap is the HBM budget for the per-query value rows.  Latency
    runs from a query's arrival to its result resident in host memory."""
    import numpy as np

    import cgp_amd
    from cgp_amd import BFS, SSSP

    spec, seed, wmin, wmax = SERVE[args.workload]
    t0 = time.time()
    log("generating graph:", describe(spec, seed, wmin, wmax))
    h = make_graph(spec, seed, wmin, wmax)
    log("graph generated in %.1fs, uploading" % (time.time() - t0))
    g = cgp_amd.Graph.from_host(h)
    V = g.V
    log("graph: V=%d E=%d ready in %.1fs" % (h.V, h.E, time.time() - t0))
    nq = args.serve_queries
    rng = np.random.default_rng(seed)
    src = [int(x) for x in h.sample_sources(nq, seed=seed)]
    kinds = [BFS if x < 0.7 else SSSP for x in rng.random(nq)]
    budget = int(args.serve_budget_gib * (1 << 30))
    en = cgp_amd.Engine(g, hbm_budget_bytes=budget)

    # The same queries as one batch: m_q (edges reached, on the device) and
    # the reference rows for the parity check of the warm-up session.
    bout = cgp_amd.DeviceBuffer(nq * V * 4)
    _, bst = en.run_batch(kinds, src, out=bout)
    m_q = cgp_amd.reached_edges(g, bout, nq).astype(np.float64)
    sum_m = float(m_q.sum())
    batch_rows = bout.to_host(shape=(nq, V))
    del bout
    log("one batch of the %d queries: %.1f ms" % (nq, bst["wall_ms"]))

    def run(arrivals, check=False):
        en.serve_start(max_slots=0, keep_results=1)
        t_base = en.serve_clock()
        nxt, done = 0, {}
        same = True
        while len(done) < nq:
            now = en.serve_clock() - t_base
            while nxt < nq and arrivals[nxt] <= now:
                en.serve_submit(kinds[nxt], src[nxt], nxt, t_base + arrivals[nxt])
                nxt += 1
            # One iteration (an idle engine only finishes its result copies).
            got = en.serve_step(max_iterations=1, fetch=check)
            for c in got:
                if check:
                    same = same and bool(np.array_equal(c["result"], batch_rows[c["ticket"]]))
                    c["result"] = None
                done[c["ticket"]] = c
            if not got and en.serve_pending() == (0, 0) and nxt < nq:
                time.sleep(min(0.001, max(0.0, arrivals[nxt] - now)))
        en.serve_stop()
        return done, same

    # Warm-up session (pins the host result buffers, checks parity), then
    # saturation: everything arrives at t=0.
    _, same = run(np.zeros(nq), check=True)
    log("warm-up session: results equal the batch: %s" % same)
    done, _ = run(np.zeros(nq))
    span = max(c["done_s"] for c in done.values()) - min(c["arrival_s"] for c in done.values())
    sat_qps = nq / span
    log("saturation: %.1f queries/s, %.2f query-GTEPS" % (sat_qps, sum_m / span / 1e9))
    rates = []
    for frac in (0.2, 0.4, 0.6, 0.8, 0.95):
        lam = frac * sat_qps
        arr = cgp_amd.poisson_arrivals(nq, lam, seed=seed + int(frac * 100))
        d, _ = run(arr)
        lat = [(d[i]["done_s"] - d[i]["arrival_s"]) * 1e3 for i in range(nq)]
        mk = max(c["done_s"] for c in d.values()) - min(c["arrival_s"] for c in d.values())
        rates.append({"rate_qps": lam, "load": frac, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                      "query_gteps": sum_m / mk / 1e9})
        log("rate %.1f q/s: p50 %.1f ms p99 %.1f ms, %.2f query-GTEPS" % (lam, rates[-1]["p50_ms"],
                                                                        rates[-1]["p99_ms"], rates[-1]["query_gteps"]))
    best = rates[-1]
    out = {
        "metric": METRIC, "value": sum_m / span / 1e9, "unit": "query-GTEPS", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": span * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": args.workload, "graph": describe(spec, seed, wmin, wmax), "V": V, "E": g.E,
                   "queries": nq, "mix": "70% BFS / 30% SSSP", "hbm_budget_bytes": budget,
                   "slots_by_budget": min(512, budget // (4 * V)),
                   "latency": "arrival to result in host memory"},
        "saturation_qps": sat_qps,
        "one_batch_ms": bst["wall_ms"],
        "latency_ms": {"p50_at_95pct_load": best["p50_ms"], "p99_at_95pct_load": best["p99_ms"]},
        "poisson": rates,
        "parity": {"serving_equals_batch": same},
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22-mixed256", choices=sorted(WORKLOADS) + sorted(SERVE))
    ap.add_argument("--one-sample", type=int, default=0,
                    help="time one-at-a-time on the first N queries only and scale (0 = all)")
    ap.add_argument("--serve-queries", type=int, default=160)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--pmc-dir", default="",
                    help="directory with rocprofv3 FETCH_SIZE/WRITE_SIZE CSVs of this workload: fills roofline.traffic")
    ap.add_argument("--pmc-runs", type=int, default=2 + COUNTED_RUNS,
                    help="batch runs in each profiled command (bench.py --steps 1 --warmup 0 makes "
                         "one capture run, one timed step and the counted runs: %d)" % (2 + COUNTED_RUNS))
    ap.add_argument("--no-one-at-a-time", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU (query-sharded multi-GPU is a design note, "
              "not built); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import cgp_amd
    from cgp_amd import BFS, SSSP, WCC

    if args.workload in SERVE:
        return serve_bench(args)
    spec, seed, wmin, wmax, n_bfs, n_sssp, n_wcc = WORKLOADS[args.workload]
    t0 = time.time()
    log("generating graph:", describe(spec, seed, wmin, wmax))
    h = make_graph(spec, seed, wmin, wmax)
    log("graph: V=%d E=%d generated in %.1fs" % (h.V, h.E, time.time() - t0))
    src = h.sample_sources(n_bfs + n_sssp, seed=seed)
    kinds = [BFS] * n_bfs + [SSSP] * n_sssp + [WCC] * n_wcc
    sources = [int(x) for x in src] + [0] * n_wcc
    n = len(kinds)
    g = cgp_amd.Graph.from_host(h)
    V = g.V
    info = cgp_amd.device_info()
    log("device:", info)

    eng = cgp_amd.Engine(g)
    out = cgp_amd.DeviceBuffer(n * V * 4)
    for i in range(args.warmup):
        _, st = eng.run_batch(kinds, sources, out=out)
        log("warmup %d: %.2f ms, %d iterations" % (i, st["wall_ms"], st["iterations"]))

    # Timed batched steps (production configuration: hipGraph replays, no
    # work counters); every advance / filter launch times itself on the
    # device wall clock (first workgroup start to last workgroup end).
    eng_t = cgp_amd.Engine(g, profile=1)
    eng_t.run_batch(kinds, sources, out=out)  # capture the timed graph outside the region
    walls = []
    adv_ms = 0.0
    adv_launches = 0
    dense_ms = 0.0
    dense_launches = 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, st = eng_t.run_batch(kinds, sources, out=out)
        walls.append((time.perf_counter() - ts) * 1e3)
        adv_ms += st["advance_ms"]
        adv_launches += st["advance_launches"]
        dense_ms += st["dense_ms"]
        dense_launches += st["dense_launches"]
        log("step %d: %.2f ms (%d iterations, advance %.2f ms over %d launches)"
            % (i, walls[-1], st["iterations"], st["advance_ms"], st["advance_launches"]))
    total_s = time.perf_counter() - t_start
    m = cgp_amd.reached_edges(g, out, n)
    sum_m = float(np.sum(m.astype(np.float64)))
    gteps = sum_m * args.steps / total_s / 1e9
    batched_rows = out.to_host(shape=(n, V))

    # Algorithmic bytes per advance launch (work counters, extra passes).  A
    # batch takes 13-15 iterations (the SSSP frontier depends on the order of
    # the atomics), so the counted launches are summed over COUNTED_RUNS runs
    # like the timed ones over the timed steps.
    eng_c = cgp_amd.Engine(g, profile=3)
    out2 = cgp_amd.DeviceBuffer(n * V * 4)
    stc = {"dense_bytes": 0, "advance_bytes": 0, "dense_launches": 0, "advance_launches": 0}
    same_counted = True
    for _ in range(COUNTED_RUNS):
        _, st1 = eng_c.run_batch(kinds, sources, out=out2)
        for k in stc:
            stc[k] += st1[k]
        same_counted &= bool(np.array_equal(out2.to_host(shape=(n, V)), batched_rows))
    # The advance is two launches per iteration: k_adv_dense (lane-per-query
    # rows of dense SSSP words) and k_adv_push (everything else).  Each is
    # timed on the device clock in the timed steps; their algorithmic bytes
    # come from the counted pass.  The roofline names the dominant one.
    kern = {
        "k_adv_dense": (stc["dense_bytes"], dense_ms, dense_launches),
        "k_adv_push": (stc["advance_bytes"] - stc["dense_bytes"], max(0.0, adv_ms - dense_ms), adv_launches),
    }
    dom = max(kern, key=lambda k: kern[k][1])

    def per_launch(k):
        # Both per-launch figures divide by the same count, the timed steps'
        # working launches per batch (a batch takes 13-15 iterations, and the
        # tail launches are tiny), so bytes / duration is the kernel's
        # algorithmic bytes per batch over its device time per batch.
        k_bytes, k_ms, k_launch = kern[k]
        per_batch = max(1e-9, k_launch / args.steps)
        b = k_bytes / COUNTED_RUNS / per_batch
        ms = k_ms / args.steps / per_batch
        return b, ms, k_launch

    bytes_per_launch, avg_launch_ms, k_launch = per_launch(dom)
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms > 0 else 0.0
    kernels_report = {}
    for k in kern:
        b, ms, nl = per_launch(k)
        kernels_report[k] = {"avg_launch_us": ms * 1e3, "launches_timed": nl, "bytes_per_launch": b,
                             "achieved_GBps": b / max(1e-12, ms / 1e3) / 1e9}
    del eng_c

    result = {
        "metric": METRIC,
        "value": gteps,
        "unit": "query-GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "graph": describe(spec, seed, wmin, wmax),
            "V": V,
            "E": g.E,
            "queries": {"bfs": n_bfs, "sssp": n_sssp, "wcc": n_wcc},
            "mask_words": (n_bfs + n_sssp + 63) // 64,
            "edges_counted": "directed CSR edges reached (m_q), WCC m_q = E",
            "device": info["name"],
        },
        "latency_ms": {"batched_p50": pct(walls, 0.5), "batched_p99": pct(walls, 0.99)},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_launch_ms * 1e3,
            "bytes_per_launch": bytes_per_launch,
            "launches_timed": k_launch,
            "timing": "device wall clock, first workgroup start to last workgroup end, per launch; "
                      "bytes and duration both per timed working launch (launches per batch averaged over "
                      "the timed steps)",
            "kernels": kernels_report,
        },
    }
    if args.pmc_dir:
        words = (n_bfs + n_sssp + 63) // 64
        R = 64 if (n_bfs + n_sssp) > 64 else 1 << max(0, (n_bfs + n_sssp - 1).bit_length())
        traffic, factors = pmc_traffic(args.pmc_dir, dom + "<", n_bfs + n_sssp, V, words, R, args.pmc_runs,
                                       k_launch / max(1, args.steps))
        if traffic:
            result["roofline"]["traffic"] = traffic
            result["roofline"]["traffic_calibration"] = factors
    log("batched: %.2f query-GTEPS, %.2f ms/step (sum m_q = %.3e)" % (gteps, result["ms_per_step"], sum_m))

    if not args.no_one_at_a_time:
        # The same kernels, each query alone, back to back from one start.
        n1 = n if args.one_sample <= 0 else min(n, args.one_sample)
        pick = list(range(n)) if n1 == n else [int(round(i * (n - 1) / max(1, n1 - 1))) for i in range(n1)]
        one = cgp_amd.DeviceBuffer(V * 4)
        done_at = []
        same = True
        elapsed = 0.0
        for i in pick:
            t1 = time.perf_counter()
            eng.run_batch([kinds[i]], [sources[i]], out=one)
            elapsed += time.perf_counter() - t1
            # Queue of back-to-back queries: each completes at the sum of the
            # run times so far (the parity check below is outside the clock).
            done_at.append(elapsed * 1e3)
            same = same and bool(np.array_equal(one.to_host(), batched_rows[i]))
        one_s = elapsed
        m_pick = float(np.sum(m.astype(np.float64)[pick]))
        one_gteps = m_pick / one_s / 1e9
        result["one_at_a_time"] = {
            "value": one_gteps,
            "ms_total": one_s * 1e3,
            "queries_timed": len(pick),
            "p50_ms": pct(done_at, 0.5),
            "p99_ms": pct(done_at, 0.99),
        }
        result["batched_over_one_at_a_time"] = gteps / one_gteps
        result["latency_ms"]["one_at_a_time_p50"] = pct(done_at, 0.5)
        result["latency_ms"]["one_at_a_time_p99"] = pct(done_at, 0.99)
        result["parity"] = {"batched_equals_one_at_a_time": same,
                            "counted_pass_equals_batched": same_counted}
        log("one-at-a-time: %.2f query-GTEPS over %.1f ms (%d queries)" % (one_gteps, one_s * 1e3, len(pick)))

    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(h, kinds, sources, args.cpu_seconds)
            log("cpu baseline:", result["cpu_baseline"])
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
