#!/usr/bin/env python3
"""Aggregate query-GTEPS of the concurrent multi-query engine, batched vs one
query at a time, on one MI355X, with p50/p99 latency under Poisson serving.

Workload (default ``rmat22-mixed256``): RMAT scale 22 (4,194,304 vertices),
edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 3, self-loops dropped,
symmetrised, deduplicated; uint32 weights uniform in [1,255].  One batch =
256 queries: 160 BFS + 80 SSSP from distinct sources drawn uniformly over
vertices with degree >= 1, plus 16 WCC; 4 mask words per vertex.

A step is one batched run of the whole batch with inputs resident in HBM and
results written to HBM.  ``value`` = sum of m_q over the batch * steps / time,
where m_q = sum of out-degrees of the vertices query q reaches (E for WCC):
directed CSR edges, the same numerator for every figure of the line (batched,
results in host memory, one at a time, the comparison baselines, the CPU).

Beside ``value`` the line carries:
  * ``host_resident``: the same batch with its results copied to host memory
    inside the clock (PCIe-inclusive rate);
  * ``one_at_a_time``: the same kernels, each query alone, back to back;
  * ``baselines``: the reference's comparison runs on these kernels — one
    query per stream on several streams at once, the sequential homogeneous
    batches (BFS, then SSSP, then WCC) against the mixed batch, and the
    all-BFS batch of the same size (the bitmask's best case);
  * ``roofline``: the dominant advance kernel's algorithmic bytes per launch
    (two byte models, both from the same device work counters) over its
    device-clock launch time, and the HBM traffic rocprofv3 counts for it;
  * ``cpu_baseline``: the reference's own CPU validators (and the oracle's C
    restatement) on the host cores, on a bounded sample of the batch;
  * ``serving``: continuous batching on RMAT scale 24 (seed 4, weights in
    [1,64]), open-loop Poisson arrivals (70 % BFS / 30 % SSSP) at five rates
    with p50/p99 latency from arrival to the result in host memory, and the
    same queries in fixed batches against continuous batching.

Prints one JSON line.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import threading
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
# Indices of cgp_run_stats.work (cgp_work_counter in include/cgp/cgp.h).
(PUSH_ITEMS, PUSH_EDGES, PUSH_WEDGES, BFS_WORDS, SSSP_PAIRS, NEW_WORDS, IMPROVED, DENSE_ITEMS, DENSE_EDGES,
 PULL_VERTS, PULL_WORDS, PULL_EDGES, DENSE_PAIRS, DENSE_IMP, DENSE_NEW, SSSP_WORDS, SSSP_SRC, DENSE_SRC) = range(18)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(values, p):
    """Nearest-rank percentile, index round(p * (n - 1))."""
    s = sorted(values)
    return s[int(round(p * (len(s) - 1)))]


def host_cpu():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, usable


# ---------------------------------------------------------------- byte models
def bytes_builder(w, W):
    """Algorithmic bytes of (push launch, lane-per-query launch) in this
    engine's model: each byte at its natural width, once, no line waste; an
    SSSP relaxation reads the target's 4-byte value of every active query."""
    Wd = 8.0 * W
    push = (w[PUSH_ITEMS] * (8 + Wd) + w[PUSH_EDGES] * 4 + w[PUSH_WEDGES] * 4 + w[BFS_WORDS] * 8 +
            w[SSSP_PAIRS] * 4 + w[IMPROVED] * 4 + w[NEW_WORDS] * 16 + w[PULL_VERTS] * (8 + Wd) +
            w[PULL_EDGES] * 4 + w[PULL_WORDS] * 8)
    dense = w[DENSE_ITEMS] * 16 + w[DENSE_EDGES] * 8 + w[DENSE_PAIRS] * 4 + w[DENSE_IMP] * 4 + w[DENSE_NEW] * 16
    return float(push), float(dense)


def bytes_scope_formula(w, W):
    """The same launches priced by the per-frontier formula of the scope
    table: per frontier item 8 (offsets) + 8W (frontier words); per edge 4
    (index) + 4 (weight, when an SSSP bit is active) + 8 per active mask word
    (the target's visited / value-state word); 16 per next word gaining bits;
    4 per lowered value; 4 per active (item, SSSP query) source value.  It
    charges an SSSP word 8 bytes per edge however many of its queries are
    active, where a relaxation reads 4 bytes per active query."""
    Wd = 8.0 * W
    push = ((w[PUSH_ITEMS] + w[PULL_VERTS]) * (8 + Wd) + w[PUSH_EDGES] * 4 + w[PUSH_WEDGES] * 4 +
            w[PULL_EDGES] * 4 + 8 * (w[BFS_WORDS] + w[SSSP_WORDS] + w[PULL_WORDS]) + 16 * w[NEW_WORDS] +
            4 * w[IMPROVED] + 4 * w[SSSP_SRC])
    dense = (w[DENSE_ITEMS] * (8 + Wd) + w[DENSE_EDGES] * (4 + 4 + 8) + 16 * w[DENSE_NEW] + 4 * w[DENSE_IMP] +
             4 * w[DENSE_SRC])
    return float(push), float(dense)


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(h, kinds, sources, m_all, seconds_target):
    """The reference's own CPU validators (oracle/_ref: queue BFS, float
    Dijkstra, label propagation) and the oracle's C restatement, one query per
    thread on the host cores, on a bounded sample of the batch (the first
    queries of each kind, in the batch's proportions); plus the oracle on one
    thread.  The numerator is the batch's m_q of the sampled queries."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    model, usable = host_cpu()
    # At most 16 worker threads: a GPU box's CPU share, whatever the host shows.
    threads = int(os.environ.get("CGP_CPU_THREADS", "0")) or max(1, min(16, usable))
    g = O.G.of(h)
    idx_by_kind = {k: [i for i, kk in enumerate(kinds) if kk == k] for k in (0, 1, 2)}
    tot = len(kinds)

    def sample(nq):
        pick = []
        for kind in (0, 1, 2):
            want = int(round(nq * len(idx_by_kind[kind]) / tot))
            if kind == 2 and idx_by_kind[2]:
                want = max(1, want)
            pick += idx_by_kind[kind][:want]
        return pick

    def m_of(pick):
        return float(sum(float(m_all[i]) for i in pick))

    def run_port(pick, nthreads):
        secs, _, _ = O.run_queries(g, [kinds[i] for i in pick], [sources[i] for i in pick], nthreads)
        return secs

    pick = sample(max(1, threads))
    secs = run_port(pick, threads)
    if secs < seconds_target / 3 and len(pick) < tot:
        nq = min(tot, int(len(pick) * seconds_target / max(secs, 1e-3) * 0.7))
        p2 = sample(nq)
        if len(p2) > len(pick):
            pick = p2
            secs = run_port(pick, threads)
    desc = "%d BFS + %d SSSP + %d WCC of the batch (first of each kind), one query per thread" % (
        sum(kinds[i] == 0 for i in pick), sum(kinds[i] == 1 for i in pick), sum(kinds[i] == 2 for i in pick))
    port = {"value": m_of(pick) / secs / 1e9, "seconds": secs, "cores": threads, "kind": "port",
            "sample": desc}
    # One thread: the first queries of the sample, sized to about a third of
    # the target.
    per_q = secs * threads / max(1, len(pick))
    n1 = max(1, min(len(pick), int(seconds_target / 3 / max(per_q, 1e-6))))
    p1 = [i for i in pick if kinds[i] != 2][:n1] or pick[:1]
    s1 = run_port(p1, 1)
    one = {"value": m_of(p1) / s1 / 1e9, "seconds": s1, "cores": 1, "kind": "port",
           "sample": "first %d BFS/SSSP queries of the sample" % len(p1)}
    out = dict(port)
    out.update({"unit": "query-GTEPS", "cpu_model": model, "host_cpus_visible": os.cpu_count(),
                "host_cpus_usable": usable, "port": port, "one_thread": one})
    if O.ref is not None and hasattr(O.ref, "ref_run_queries"):
        rsecs, _ = O.ref_run_queries(g, [kinds[i] for i in pick], [sources[i] for i in pick], threads)
        ref = {"value": m_of(pick) / rsecs / 1e9, "seconds": rsecs, "cores": threads, "kind": "reference",
               "sample": desc + " (puercgp/examples/mtx_loader.hxx:269-356 cpu_bfs / cpu_sssp / cpu_wcc)"}
        out.update(ref)
        out["reference"] = ref
    return out


# ---------------------------------------------------------------- rocprofv3 traffic
CALIBRATION_BYTES = 1 << 30  # the plain copy each counter pass ends with


def pmc_traffic(pmc_dir, kernel_prefix, runs, launches_per_run):
    """HBM bytes per working launch of `kernel_prefix` from rocprofv3
    FETCH_SIZE / WRITE_SIZE CSVs (KiB per dispatch): each counter's total over
    the profiled run, scaled by the factor that makes the same counter read
    the known bytes of a plain copy kernel (k_copy16: reads and writes
    CALIBRATION_BYTES once each), divided by the working launches of that run
    (`runs` batch runs of `launches_per_run` launches that did work each, the
    same launches the algorithmic bytes are averaged over)."""
    import csv
    import glob
    tot = 0.0
    factors = {}
    for counter, known in (("FETCH_SIZE", float(CALIBRATION_BYTES)), ("WRITE_SIZE", float(CALIBRATION_BYTES))):
        vals, cal = [], []
        for f in glob.glob(os.path.join(pmc_dir, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r.get("Counter_Name") != counter:
                    continue
                name = r.get("Kernel_Name", "")
                v = float(r.get("Counter_Value", 0.0)) * 1024.0
                if "k_copy16" in name:
                    cal.append(v)
                elif kernel_prefix in name:
                    vals.append(v)
        if not vals or not cal:
            return None, None
        factor = known / (sum(cal) / len(cal))
        tot += factor * sum(vals)
        factors[counter.lower() + "_factor"] = factor
    return tot / max(1.0, runs * launches_per_run), factors


def run_pmc_children(workload):
    """Two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over a
    short run of this bench (one timed step, the counted passes), each a child
    process of its own under a time limit.  Returns (csv dir, None) or (None,
    reason)."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None, "rocprofv3 not found"
    d = tempfile.mkdtemp(prefix="cgp_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    for counter, name in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        cmd = ["timeout", "-s", "KILL", "240", prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o",
               name, "--", sys.executable, os.path.join(ROOT, "bench.py"), "--workload", workload, "--steps", "1",
               "--warmup", "0", "--no-one-at-a-time", "--no-cpu", "--no-serve", "--no-pmc", "--no-baselines",
               "--no-host", "--calibration-copy"]
        log("rocprofv3 %s pass" % counter)
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=sys.stderr)
        if r.returncode != 0:
            return None, "rocprofv3 --pmc %s exited with status %d" % (counter, r.returncode)
    return d, None


# ---------------------------------------------------------------- graphs
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


# ---------------------------------------------------------------- comparison baselines
def multi_stream(g, kinds, sources, m, n_streams):
    """One query per stream, several streams at once: `n_streams` engines
    (each its own HIP stream, driven by its own host thread) take the queries
    one at a time from a shared queue (the multi-stream baseline of
    cgp/examples/algorithms/bfs/cgp_bfs_multi_stream.cu:329-372, on these
    kernels)."""
    import cgp_amd
    V = g.V
    engines = [cgp_amd.Engine(g) for _ in range(n_streams)]
    outs = [cgp_amd.DeviceBuffer(V * 4) for _ in range(n_streams)]
    kinds_present = sorted(set(kinds))
    for e, o in zip(engines, outs):  # capture each engine's graphs outside the clock
        for k in kinds_present:
            e.run_batch([k], [sources[kinds.index(k)]], out=o)
    n = len(kinds)
    lock = threading.Lock()
    nxt = [0]
    done_at = [0.0] * n
    errors = []
    t0 = time.perf_counter()

    def worker(k):
        try:
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= n:
                    return
                engines[k].run_batch([kinds[i]], [sources[i]], out=outs[k])
                done_at[i] = (time.perf_counter() - t0) * 1e3
        except Exception as ex:  # reported, never hidden
            errors.append(repr(ex))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(n_streams)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    total = time.perf_counter() - t0
    if errors:
        raise RuntimeError("multi-stream baseline: " + errors[0])
    return {"streams": n_streams, "ms_total": total * 1e3, "value": float(sum(m)) / total / 1e9,
            "p50_ms": pct(done_at, 0.5), "p99_ms": pct(done_at, 0.99)}


def timed_runs(eng, kinds, sources, out, reps=3):
    walls = []
    for _ in range(reps):
        t = time.perf_counter()
        eng.run_batch(kinds, sources, out=out)
        walls.append((time.perf_counter() - t) * 1e3)
    return pct(walls, 0.5)


def comparison_baselines(g, eng, kinds, sources, m, batched_ms, out):
    """The trio of puercgp/examples/bench_hybrid.cu:143-240 on these kernels
    (median of three runs each): the homogeneous batches one after the other
    (all BFS, then all SSSP, then all WCC) against the one mixed batch, and
    the ideal case, a BFS-only batch of the same size (every source of the
    batch as a BFS query, padded with vertex 0 as the reference pads); plus
    the multi-stream baseline."""
    res = {}
    seq = 0.0
    parts = {}
    for k, name in ((0, "bfs"), (1, "sssp"), (2, "wcc")):
        ks = [kk for kk in kinds if kk == k]
        if not ks:
            continue
        ss = [s for kk, s in zip(kinds, sources) if kk == k]
        eng.run_batch(ks, ss, out=out)  # capture
        parts[name] = timed_runs(eng, ks, ss, out)
        seq += parts[name]
    res["sequential_homogeneous"] = {"ms": seq, "parts_ms": parts, "value": float(sum(m)) / (seq / 1e3) / 1e9}
    res["hybrid"] = {"ms": batched_ms, "speedup_vs_sequential": seq / batched_ms}
    srcs = [s for kk, s in zip(kinds, sources) if kk != 2]
    srcs += [0] * (len(kinds) - len(srcs))
    eng.run_batch([0] * len(srcs), srcs, out=out)
    ideal = timed_runs(eng, [0] * len(srcs), srcs, out)
    res["ideal_all_bfs"] = {"ms": ideal, "queries": len(srcs), "speedup_vs_sequential": seq / ideal}
    n_streams = 8
    res["multi_stream"] = multi_stream(g, kinds, sources, m, n_streams)
    return res


# ---------------------------------------------------------------- serving (continuous batching)
def serving_leg(args, out_line=None):
    """Open-loop Poisson serving on RMAT scale 24: 70% BFS / 30% SSSP with
    uniform sources; five arrival rates relative to the measured saturation
    throughput; the admission cap is the HBM budget for the per-query value
    rows.  Latency runs from a query's arrival to its result resident in host
    memory.  Also: the same queries in fixed batches of `fixed_batch`, run
    back to back with each batch's results copied to host memory, against
    continuous batching with every query arriving at t = 0 (the comparison of
    puercgp/examples/bench_replenish.cu:99-230)."""
    import numpy as np

    import cgp_amd
    from cgp_amd import BFS, SSSP

    name = "rmat24-serve"
    spec, seed, wmin, wmax = SERVE[name]
    t0 = time.time()
    log("serving: generating graph:", describe(spec, seed, wmin, wmax))
    h = make_graph(spec, seed, wmin, wmax)
    g = cgp_amd.Graph.from_host(h)
    V = g.V
    log("serving: graph V=%d E=%d ready in %.1fs" % (h.V, h.E, time.time() - t0))
    nq = args.serve_queries
    rng = np.random.default_rng(seed)
    src = [int(x) for x in h.sample_sources(nq, seed=seed)]
    kinds = [BFS if x < 0.7 else SSSP for x in rng.random(nq)]
    del h
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
    log("serving: one batch of the %d queries: %.1f ms" % (nq, bst["wall_ms"]))

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
    del batch_rows
    log("serving: warm-up session results equal the batch: %s" % same)
    done, _ = run(np.zeros(nq))
    span = max(c["done_s"] for c in done.values()) - min(c["arrival_s"] for c in done.values())
    sat_lat = [(done[i]["done_s"] - done[i]["arrival_s"]) * 1e3 for i in range(nq)]
    sat_qps = nq / span
    log("serving: saturation %.1f queries/s, %.2f query-GTEPS" % (sat_qps, sum_m / span / 1e9))
    rates = []
    for frac in (0.2, 0.4, 0.6, 0.8, 0.95):
        lam = frac * sat_qps
        arr = cgp_amd.poisson_arrivals(nq, lam, seed=seed + int(frac * 100))
        d, _ = run(arr)
        lat = [(d[i]["done_s"] - d[i]["arrival_s"]) * 1e3 for i in range(nq)]
        mk = max(c["done_s"] for c in d.values()) - min(c["arrival_s"] for c in d.values())
        rates.append({"rate_qps": lam, "load": frac, "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99),
                      "query_gteps": sum_m / mk / 1e9, "samples": nq})
        log("serving: rate %.1f q/s: p50 %.1f ms p99 %.1f ms, %.2f query-GTEPS"
            % (lam, rates[-1]["p50_ms"], rates[-1]["p99_ms"], rates[-1]["query_gteps"]))

    # Fixed batches, back to back, each batch's results copied to host memory.
    fb = args.fixed_batch
    fout = cgp_amd.DeviceBuffer(fb * V * 4)
    en.run_batch(kinds[:fb], src[:fb], out=fout)  # capture outside the clock
    host = np.empty((fb, V), dtype=np.uint32)
    import ctypes as C
    fixed_lat = [0.0] * nq
    t = time.perf_counter()
    for b0 in range(0, nq, fb):
        b1 = min(nq, b0 + fb)
        en.run_batch(kinds[b0:b1], src[b0:b1], out=fout)
        cgp_amd.check(cgp_amd.lib.cgp_copy_to_host(host.ctypes.data_as(C.c_void_p), fout.ptr, (b1 - b0) * V * 4))
        for i in range(b0, b1):
            fixed_lat[i] = (time.perf_counter() - t) * 1e3
    fixed_total = (time.perf_counter() - t) * 1e3
    del fout, host
    log("serving: fixed batches of %d: %.1f ms total" % (fb, fixed_total))
    return {
        "workload": name, "graph": describe(spec, seed, wmin, wmax), "V": V, "E": g.E, "queries": nq,
        "mix": "70% BFS / 30% SSSP", "hbm_budget_bytes": budget, "slots_by_budget": min(512, budget // (4 * V)),
        "latency": "arrival to result in host memory (nearest-rank p50/p99)",
        "saturation_qps": sat_qps, "saturation_query_gteps": sum_m / span / 1e9,
        "one_batch_ms_results_in_hbm": bst["wall_ms"],
        "poisson": rates,
        "p50_ms_at_95pct_load": rates[-1]["p50_ms"], "p99_ms_at_95pct_load": rates[-1]["p99_ms"],
        "replenish_vs_fixed": {
            "fixed_batch": fb, "fixed_total_ms": fixed_total, "fixed_p50_ms": pct(fixed_lat, 0.5),
            "fixed_p99_ms": pct(fixed_lat, 0.99), "replenish_total_ms": span * 1e3,
            "replenish_p50_ms": pct(sat_lat, 0.5), "replenish_p99_ms": pct(sat_lat, 0.99),
            "replenish_speedup": fixed_total / (span * 1e3),
            "arrivals": "all queries at t = 0; results in host memory for both"},
        "parity": {"serving_equals_batch": same},
    }


def serve_only(args):
    r = serving_leg(args)
    out = {
        "metric": METRIC, "value": r["saturation_query_gteps"], "unit": "query-GTEPS", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": r["queries"] / r["saturation_qps"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": r["workload"], "graph": r["graph"], "V": r["V"], "E": r["E"],
                   "queries": r["queries"], "mix": r["mix"], "hbm_budget_bytes": r["hbm_budget_bytes"],
                   "latency": r["latency"]},
        "latency_ms": {"p50_at_95pct_load": r["p50_ms_at_95pct_load"], "p99_at_95pct_load": r["p99_ms_at_95pct_load"]},
        "serving": r,
    }
    print(json.dumps(out), flush=True)
    return 0


# ---------------------------------------------------------------- main
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
    ap.add_argument("--fixed-batch", type=int, default=64, help="fixed-batch size of the serving comparison")
    ap.add_argument("--pmc-dir", default="",
                    help="directory with rocprofv3 FETCH_SIZE/WRITE_SIZE CSVs of this workload: fills "
                         "roofline.traffic instead of the bench's own counter passes")
    ap.add_argument("--pmc-runs", type=int, default=2 + COUNTED_RUNS,
                    help="batch runs in each profiled command (bench.py --steps 1 --warmup 0 makes "
                         "one capture run, one timed step and the counted runs: %d)" % (2 + COUNTED_RUNS))
    ap.add_argument("--no-pmc", action="store_true", help="no rocprofv3 counter passes (traffic stays null)")
    ap.add_argument("--no-host", action="store_true", help="no run with the results copied to host memory")
    ap.add_argument("--no-one-at-a-time", action="store_true")
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibration-copy", action="store_true",
                    help="end with one device-to-device copy of %d MiB through the library's copy kernel "
                         "(the known traffic the counter passes are scaled by)" % (CALIBRATION_BYTES >> 20))
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU (query-sharded multi-GPU is a design note, "
              "not built); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import cgp_amd
    from cgp_amd import BFS, SSSP, WCC

    if args.workload in SERVE:
        return serve_only(args)
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
    W = max(1, (n_bfs + n_sssp + 63) // 64)
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
    adv_ms = adv_launches = dense_ms = dense_launches = 0
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
    del eng_t
    m = cgp_amd.reached_edges(g, out, n)
    sum_m = float(np.sum(m.astype(np.float64)))
    gteps = sum_m * args.steps / total_s / 1e9
    ms_step = total_s * 1e3 / args.steps
    batched_rows = out.to_host(shape=(n, V))

    # Algorithmic bytes per advance launch (work counters, extra passes).  A
    # batch takes 13-15 iterations (the SSSP frontier depends on the order of
    # the atomics), so the counted launches are summed over COUNTED_RUNS runs
    # like the timed ones over the timed steps.
    eng_c = cgp_amd.Engine(g, profile=3)
    out2 = cgp_amd.DeviceBuffer(n * V * 4)
    work = [0] * 24
    eng_bytes = {"dense_bytes": 0.0, "advance_bytes": 0.0}
    same_counted = True
    for _ in range(COUNTED_RUNS):
        _, st1 = eng_c.run_batch(kinds, sources, out=out2)
        work = [a + b for a, b in zip(work, st1["work"])]
        for k in eng_bytes:
            eng_bytes[k] += st1[k]
        same_counted &= bool(np.array_equal(out2.to_host(shape=(n, V)), batched_rows))
    del eng_c, out2
    push_b, dense_b = bytes_builder(work, W)
    push_s, dense_s = bytes_scope_formula(work, W)
    # The advance is two launches per iteration: k_adv_dense (lane-per-query
    # rows of dense SSSP words) and k_adv_push (everything else).  Each is
    # timed on the device clock in the timed steps; their algorithmic bytes
    # come from the counted passes.  The roofline names the dominant one.
    kern = {
        "k_adv_dense": (dense_b, dense_s, dense_ms, dense_launches),
        "k_adv_push": (push_b, push_s, max(0.0, adv_ms - dense_ms), adv_launches),
    }
    dom = max(kern, key=lambda k: kern[k][2])

    def per_launch(k):
        # Both per-launch figures divide by the same count, the timed steps'
        # working launches per batch (the tail launches are tiny), so bytes /
        # duration is the kernel's algorithmic bytes per batch over its device
        # time per batch.
        b_eng, b_sur, k_ms, k_launch = kern[k]
        per_batch = max(1e-9, k_launch / args.steps)
        return (b_eng / COUNTED_RUNS / per_batch, b_sur / COUNTED_RUNS / per_batch,
                k_ms / args.steps / per_batch, k_launch)

    kernels_report = {}
    for k in kern:
        b, bs, ms, nl = per_launch(k)
        kernels_report[k] = {"avg_launch_us": ms * 1e3, "launches_timed": nl, "bytes_per_launch": b,
                             "achieved_GBps": b / max(1e-12, ms / 1e3) / 1e9,
                             "bytes_per_launch_scope_formula": bs,
                             "achieved_GBps_scope_formula": bs / max(1e-12, ms / 1e3) / 1e9,
                             "ms_per_batch": kern[k][2] / args.steps}
    bytes_per_launch, bytes_scope_launch, avg_launch_ms, k_launch = per_launch(dom)
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms > 0 else 0.0
    adv_total_ms = adv_ms / args.steps
    adv_total_b = (push_b + dense_b) / COUNTED_RUNS
    result = {
        "metric": METRIC,
        "value": gteps,
        "unit": "query-GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
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
            "mask_words": W,
            "edges_counted": "directed CSR edges reached (m_q), WCC m_q = E",
            "results": "resident in HBM when the clock stops (host_resident: copied to host memory)",
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
            "byte_model": "engine: 4 B per (edge, active SSSP query) target value, 8 B per (edge, active BFS "
                          "word), per-item offsets and frontier words, 16 B per next word gaining bits",
            "scope_formula": {
                "bytes_per_launch": bytes_scope_launch,
                "achieved": bytes_scope_launch / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms > 0 else 0.0,
                "frac": bytes_scope_launch / (avg_launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS if avg_launch_ms > 0
                else 0.0,
                "model": "8 B per (edge, active mask word) whatever the number of active queries in it"},
            "advance_total": {"ms_per_batch": adv_total_ms, "bytes_per_batch": adv_total_b,
                              "frac": adv_total_b / (adv_total_ms / 1e3) / 1e9 / HBM_PEAK_GBS
                              if adv_total_ms > 0 else 0.0},
            "timing": "device wall clock, first workgroup start to last workgroup end, per launch; "
                      "bytes and duration both per timed working launch (launches per batch averaged over "
                      "the timed steps)",
            "kernels": kernels_report,
            "work_counters_per_batch": [w // COUNTED_RUNS for w in work[:18]],
        },
        "parity": {"counted_pass_equals_batched": same_counted},
    }

    log("batched: %.2f query-GTEPS, %.2f ms/step (sum m_q = %.3e)" % (gteps, ms_step, sum_m))
    if not args.no_host:
        # The same batch with the results copied to host memory inside the clock.
        host_rows = np.empty((n, V), dtype=np.uint32)
        import ctypes as C
        hs = []
        for _ in range(2):
            t = time.perf_counter()
            eng.run_batch(kinds, sources, out=out)
            cgp_amd.check(cgp_amd.lib.cgp_copy_to_host(host_rows.ctypes.data_as(C.c_void_p), out.ptr, n * V * 4))
            hs.append(time.perf_counter() - t)
        t_host = min(hs)
        result["host_resident"] = {"value": sum_m / t_host / 1e9, "ms": t_host * 1e3,
                                   "results_bytes": n * V * 4,
                                   "equals_hbm_results": bool(np.array_equal(host_rows, batched_rows))}
        del host_rows
        log("results copied to the host inside the clock: %.1f ms per batch" % (t_host * 1e3))

    # HBM traffic of the dominant kernel: rocprofv3 counter passes.
    pmc_dir, pmc_note = args.pmc_dir or None, None
    if not pmc_dir and not args.no_pmc:
        try:
            pmc_dir, pmc_note = run_pmc_children(args.workload)
        except Exception as e:  # the traffic figure is reported, never required
            pmc_dir, pmc_note = None, repr(e)
    if pmc_dir:
        for k in kern:
            tr, factors = pmc_traffic(pmc_dir, k + "<", args.pmc_runs, kern[k][3] / max(1, args.steps))
            if tr:
                kr = kernels_report[k]
                kr["traffic"] = tr
                kr["traffic_over_algorithmic"] = tr / max(1.0, kr["bytes_per_launch"])
                kr["traffic_GBps"] = tr / max(1e-12, kr["avg_launch_us"] / 1e6) / 1e9
                if k == dom:
                    result["roofline"]["traffic"] = tr
                    result["roofline"]["traffic_over_algorithmic"] = tr / max(1.0, bytes_per_launch)
                    result["roofline"]["traffic_calibration"] = factors
            elif k == dom:
                pmc_note = "no counter rows for %s in %s" % (dom, pmc_dir)
    if pmc_note:
        result["roofline"]["traffic_note"] = pmc_note

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
        m_pick = float(np.sum(m.astype(np.float64)[pick]))
        one_gteps = m_pick / elapsed / 1e9
        result["one_at_a_time"] = {
            "value": one_gteps,
            "ms_total": elapsed * 1e3,
            "queries_timed": len(pick),
            "p50_ms": pct(done_at, 0.5),
            "p99_ms": pct(done_at, 0.99),
        }
        result["batched_over_one_at_a_time"] = gteps / one_gteps
        result["latency_ms"]["one_at_a_time_p50"] = pct(done_at, 0.5)
        result["latency_ms"]["one_at_a_time_p99"] = pct(done_at, 0.99)
        result["parity"]["batched_equals_one_at_a_time"] = same
        log("one-at-a-time: %.2f query-GTEPS over %.1f ms (%d queries)" % (one_gteps, elapsed * 1e3, len(pick)))

    if not args.no_baselines:
        b = comparison_baselines(g, eng, kinds, sources, m, ms_step, out)
        result["baselines"] = b
        log("baselines: sequential homogeneous %.1f ms, mixed %.1f ms, all-BFS %.1f ms, "
            "%d streams one query each %.1f ms" % (b["sequential_homogeneous"]["ms"], ms_step,
                                                   b["ideal_all_bfs"]["ms"], b["multi_stream"]["streams"],
                                                   b["multi_stream"]["ms_total"]))

    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(h, kinds, sources, m, args.cpu_seconds)
            log("cpu baseline:", {k: v for k, v in result["cpu_baseline"].items() if k not in ("port", "reference")})
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": repr(e)}

    del out, eng, batched_rows, g, h
    if not args.no_serve and args.workload == "rmat22-mixed256":
        try:
            result["serving"] = serving_leg(args)
            result["latency_ms"]["serving_p50_at_95pct_load"] = result["serving"]["p50_ms_at_95pct_load"]
            result["latency_ms"]["serving_p99_at_95pct_load"] = result["serving"]["p99_ms_at_95pct_load"]
        except Exception as e:
            result["serving"] = {"error": repr(e)}
    if args.calibration_copy:
        src_buf, dst_buf = cgp_amd.DeviceBuffer(CALIBRATION_BYTES), cgp_amd.DeviceBuffer(CALIBRATION_BYTES)
        dst_buf.copy_from(src_buf)
        del src_buf, dst_buf
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
