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
    # name: (scale, edge_factor, seed, wmin, wmax, n_bfs, n_sssp, n_wcc)
    "rmat22-mixed256": (22, 16, 3, 1, 255, 160, 80, 16),
    "rmat20-sssp64": (20, 16, 2, 1, 64, 0, 64, 0),
    "rmat16-bfs32": (16, 16, 1, 0, 0, 32, 0, 0),
}
METRIC = "aggregate query-GTEPS + p50/p99 latency, batched vs one-at-a-time, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pct(values, p):
    s = sorted(values)
    return s[int(round(p * (len(s) - 1)))]


def cpu_baseline(h, kinds, sources, seconds_target):
    """Times the oracle (CPU restatement of the reference validators) on a
    bounded sample of the batch, one query per thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O

    threads = int(os.environ.get("CGP_CPU_THREADS", "0")) or min(16, os.cpu_count() or 1)
    g = O.G.of(h)
    # Sample proportional to the batch mix: ~threads queries.
    nb = sum(1 for k in kinds if k == 0)
    ns = sum(1 for k in kinds if k == 1)
    nw = sum(1 for k in kinds if k == 2)
    tot = len(kinds)
    want = max(1, threads)
    take = {0: round(want * nb / tot), 1: round(want * ns / tot), 2: min(1, nw)}
    samp_k, samp_s = [], []
    for kind in (0, 1, 2):
        picked = [s for k, s in zip(kinds, sources) if k == kind][: max(0, int(take[kind]))]
        samp_k += [kind] * len(picked)
        samp_s += picked
    if not samp_k:
        samp_k, samp_s = [kinds[0]], [sources[0]]
    secs, _, m = O.run_queries(g, samp_k, samp_s, threads)
    desc = "%d BFS + %d SSSP + %d WCC of the batch (first of each kind), one query per thread" % (
        samp_k.count(0), samp_k.count(1), samp_k.count(2))
    return {
        "value": float(np.sum(m)) / secs / 1e9,
        "unit": "query-GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": desc,
        "seconds": secs,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22-mixed256", choices=sorted(WORKLOADS))
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

    scale, ef, seed, wmin, wmax, n_bfs, n_sssp, n_wcc = WORKLOADS[args.workload]
    t0 = time.time()
    h = cgp_amd.HostCSR.rmat(scale, ef, seed=seed, wmin=wmin, wmax=wmax)
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
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, st = eng_t.run_batch(kinds, sources, out=out)
        walls.append((time.perf_counter() - ts) * 1e3)
        adv_ms += st["advance_ms"]
        adv_launches += st["advance_launches"]
        log("step %d: %.2f ms (%d iterations, advance %.2f ms over %d launches)"
            % (i, walls[-1], st["iterations"], st["advance_ms"], st["advance_launches"]))
    total_s = time.perf_counter() - t_start
    m = cgp_amd.reached_edges(g, out, n)
    sum_m = float(np.sum(m.astype(np.float64)))
    gteps = sum_m * args.steps / total_s / 1e9
    batched_rows = out.to_host(shape=(n, V))

    # Algorithmic bytes per advance launch (work counters, one extra pass).
    eng_c = cgp_amd.Engine(g, profile=3)
    out2 = cgp_amd.DeviceBuffer(n * V * 4)
    _, stc = eng_c.run_batch(kinds, sources, out=out2)
    same_counted = bool(np.array_equal(out2.to_host(shape=(n, V)), batched_rows))
    bytes_per_launch = stc["advance_bytes"] / max(1, stc["advance_launches"])
    avg_launch_ms = adv_ms / max(1, adv_launches)
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms > 0 else 0.0
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
            "graph": "RMAT scale %d ef %d a=0.57 b=0.19 c=0.19 seed %d, symmetrised, dedup, w in [%d,%d]"
                     % (scale, ef, seed, wmin, wmax),
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
            "kernel": "k_advance<%d>" % (4 if (n_bfs + n_sssp) > 128 else (2 if n_bfs + n_sssp > 64 else 1)),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_launch_ms * 1e3,
            "bytes_per_launch": bytes_per_launch,
            "launches_timed": adv_launches,
        },
    }
    log("batched: %.2f query-GTEPS, %.2f ms/step (sum m_q = %.3e)" % (gteps, result["ms_per_step"], sum_m))

    if not args.no_one_at_a_time:
        one = cgp_amd.DeviceBuffer(n * V * 4)
        done_at = []
        t1 = time.perf_counter()
        for i in range(n):
            eng.run_batch([kinds[i]], [sources[i]], out=one.view(i * V * 4, V * 4))
            done_at.append((time.perf_counter() - t1) * 1e3)
        one_s = time.perf_counter() - t1
        one_rows = one.to_host(shape=(n, V))
        result["one_at_a_time"] = {
            "value": sum_m / one_s / 1e9,
            "ms_total": one_s * 1e3,
            "p50_ms": pct(done_at, 0.5),
            "p99_ms": pct(done_at, 0.99),
        }
        result["batched_over_one_at_a_time"] = gteps / (sum_m / one_s / 1e9)
        result["latency_ms"]["one_at_a_time_p50"] = pct(done_at, 0.5)
        result["latency_ms"]["one_at_a_time_p99"] = pct(done_at, 0.99)
        result["parity"] = {"batched_equals_one_at_a_time": bool(np.array_equal(one_rows, batched_rows)),
                            "counted_pass_equals_batched": same_counted}
        log("one-at-a-time: %.2f query-GTEPS over %.1f ms" % (result["one_at_a_time"]["value"], one_s * 1e3))

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
