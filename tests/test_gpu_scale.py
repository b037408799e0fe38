"""Full-size parity: the benchmark workloads at their real sizes, run through
the C-ABI on the GPU.  Every result row is checked by the oracle's threaded
C checkers (tight-edge characterisation of BFS depths and SSSP distances,
min-id component labels with connected classes), WCC rows are compared with
union-find, and a few rows per workload are compared exactly with the
oracle's BFS / Dijkstra."""
import sys
import time

import numpy as np
import pytest

import cgp_amd
import oracle_lib as O
from cgp_amd import BFS, SSSP, WCC

# Full-size graphs take minutes (generation, 32 GiB result copies, threaded checks).
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_all(og, kinds, sources, out, exact, label):
    """Property-checks every row of `out`, compares the rows in `exact` with
    the oracle (threaded) and every WCC row with union-find."""
    t = time.time()
    bad = O.check_rows(og, kinds, sources, out)
    say("%s: %d rows property-checked in %.1fs, %d failed %s" % (label, len(kinds), time.time() - t, len(bad), bad[:4]))
    assert not bad
    t = time.time()
    ks = [kinds[i] for i in exact]
    ss = [sources[i] for i in exact]
    _, exp, _ = O.run_queries(og, ks, ss, O.host_threads(), keep=True)
    for j, i in enumerate(exact):
        assert np.array_equal(out[i], exp[j]), "%s: row %d (kind %d, source %d) differs from the oracle" % (
            label, i, kinds[i], sources[i])
    wcc_rows = [i for i, k in enumerate(kinds) if k == WCC]
    if wcc_rows:
        lab = O.wcc_uf(og)
        for i in wcc_rows:
            assert np.array_equal(out[i], lab), "%s: WCC row %d differs from union-find" % (label, i)
    say("%s: %d rows exact, %d WCC rows equal union-find (%.1fs)" % (label, len(exact), len(wcc_rows), time.time() - t))


def test_rmat20_64_sssp_every_row(gpu):
    # RMAT scale 20, edge factor 16, seed 2, w in [1,64]; 64 SSSP in one mask word.
    t = time.time()
    h = cgp_amd.HostCSR.rmat(20, 16, seed=2, wmin=1, wmax=64)
    og = O.G.of(h)
    g = cgp_amd.Graph.from_host(h)
    src = [int(x) for x in h.sample_sources(64, seed=2)]
    say("rmat20: V=%d E=%d ready in %.1fs" % (h.V, h.E, time.time() - t))
    out, st = cgp_amd.Engine(g).run_batch([SSSP] * 64, src)
    say("rmat20: 64 SSSP in %.1f ms, %d iterations" % (st["wall_ms"], st["iterations"]))
    check_all(og, [SSSP] * 64, src, out, [0, 21, 42, 63], "rmat20-sssp64")


def test_rmat22_mixed_256_every_row(gpu):
    # RMAT scale 22, seed 3, w in [1,255]; 160 BFS + 80 SSSP + 16 WCC, 4 mask words.
    t = time.time()
    h = cgp_amd.HostCSR.rmat(22, 16, seed=3, wmin=1, wmax=255)
    og = O.G.of(h)
    g = cgp_amd.Graph.from_host(h)
    src = [int(x) for x in h.sample_sources(240, seed=3)]
    kinds = [BFS] * 160 + [SSSP] * 80 + [WCC] * 16
    sources = src + [0] * 16
    say("rmat22: V=%d E=%d ready in %.1fs" % (h.V, h.E, time.time() - t))
    out, st = cgp_amd.Engine(g).run_batch(kinds, sources)
    say("rmat22: 256 queries in %.1f ms, %d iterations" % (st["wall_ms"], st["iterations"]))
    check_all(og, kinds, sources, out, [0, 159, 160, 239], "rmat22-mixed256")


def test_grid4096_512_sssp_every_row(gpu):
    # 4096 x 4096 grid plus one random edge per vertex w.p. 0.01, seed 5,
    # w in [1,1000]; 512 SSSP in 8 mask words (32 GiB of results).
    t = time.time()
    h = cgp_amd.HostCSR.grid(4096, 4096, p_extra=0.01, seed=5, wmin=1, wmax=1000)
    og = O.G.of(h)
    g = cgp_amd.Graph.from_host(h)
    src = [int(x) for x in h.sample_sources(512, seed=5)]
    say("grid4096: V=%d E=%d ready in %.1fs" % (h.V, h.E, time.time() - t))
    en = cgp_amd.Engine(g)
    out, st = en.run_batch([SSSP] * 512, src)
    say("grid4096: 512 SSSP in %.1f ms, %d iterations" % (st["wall_ms"], st["iterations"]))
    del en
    check_all(og, [SSSP] * 512, src, out, [0, 255, 511], "grid4096-sssp512")


def test_rmat24_serving_poisson_every_result(gpu):
    # RMAT scale 24, seed 4, w in [1,64]: an open-loop Poisson stream of 48
    # queries (70 % BFS / 30 % SSSP) served with continuous batching under a
    # 16 GiB value budget; queries arrive while others run.  Every result is
    # property-checked, two are compared exactly, and all equal the same
    # queries run as one batch.  V x 256 bytes >= 4 GiB here, so the
    # lane-per-query loop runs in its 64-bit address form.
    t = time.time()
    h = cgp_amd.HostCSR.rmat(24, 16, seed=4, wmin=1, wmax=64)
    og = O.G.of(h)
    g = cgp_amd.Graph.from_host(h)
    V = g.V
    nq = 48
    rng = np.random.default_rng(4)
    src = [int(x) for x in h.sample_sources(nq, seed=4)]
    kinds = [BFS if x < 0.7 else SSSP for x in rng.random(nq)]
    say("rmat24: V=%d E=%d ready in %.1fs, %d BFS + %d SSSP" % (h.V, h.E, time.time() - t, kinds.count(BFS),
                                                                 kinds.count(SSSP)))
    en = cgp_amd.Engine(g, hbm_budget_bytes=16 << 30)
    arrivals = cgp_amd.poisson_arrivals(nq, 150.0, seed=4)
    en.serve_start(max_slots=0, keep_results=1)
    t_base = en.serve_clock()
    nxt, done = 0, {}
    max_running = 0
    while len(done) < nq:
        now = en.serve_clock() - t_base
        while nxt < nq and arrivals[nxt] <= now:
            en.serve_submit(kinds[nxt], src[nxt], nxt, t_base + arrivals[nxt])
            nxt += 1
        for c in en.serve_step(max_iterations=1):
            done[c["ticket"]] = c
        max_running = max(max_running, en.serve_pending()[1])
        if not done and nxt == 0:
            time.sleep(0.001)
        assert en.serve_clock() - t_base < 600, "serving session did not drain"
    en.serve_stop()
    lat = sorted((c["done_s"] - c["arrival_s"]) * 1e3 for c in done.values())
    say("rmat24 serving: %d results, up to %d queries running together, latency p50 %.1f ms p99 %.1f ms"
        % (nq, max_running, lat[len(lat) // 2], lat[-1]))
    assert max_running > 1, "arrivals never overlapped on the device"
    served = np.stack([done[i]["result"] for i in range(nq)])
    exact = [kinds.index(BFS), kinds.index(SSSP)]
    check_all(og, kinds, src, served, exact, "rmat24-serving")
    # The same queries as one batch give the same rows.
    out, st = en.run_batch(kinds, src)
    say("rmat24: the %d queries as one batch in %.1f ms" % (nq, st["wall_ms"]))
    for i in range(nq):
        assert np.array_equal(out[i], served[i]), "query %d differs between serving and the batch" % i
    del en, out, served
