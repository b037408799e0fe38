"""GPU: graphs stored in descending-degree order (cgp_graph_create_csr_ex with
CGP_RENUMBER_ON, the storage large skewed graphs get by default).  Sources,
results, WCC labels and per-query edge counts must stay in the caller's
numbering: every row is compared with the oracle on the caller's arrays, and
with the same graph stored in the caller's order."""
import sys

import numpy as np
import pytest

import cgp_amd
import oracle_lib as O
from cgp_amd import BFS, SSSP, WCC

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def graphs(h):
    return cgp_amd.Graph.from_host(h, renumber=True), cgp_amd.Graph.from_host(h, renumber=False)


@pytest.mark.parametrize("n_bfs,n_sssp,n_wcc", [(30, 20, 2), (100, 70, 1), (5, 3, 1), (0, 1, 0)])
def test_mixed_batches_on_a_renumbered_graph(gpu, n_bfs, n_sssp, n_wcc):
    h = cgp_amd.HostCSR.rmat(12, 16, seed=51, wmin=1, wmax=80)
    say("RMAT-12 renumbered: %d BFS + %d SSSP + %d WCC" % (n_bfs, n_sssp, n_wcc))
    gr, go = graphs(h)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(n_bfs + n_sssp, seed=5)]
    kinds = [BFS] * n_bfs + [SSSP] * n_sssp + [WCC] * n_wcc
    sources = src + [0] * n_wcc
    out_r, _ = cgp_amd.Engine(gr).run_batch(kinds, sources)
    out_o, _ = cgp_amd.Engine(go).run_batch(kinds, sources)
    assert np.array_equal(out_r, out_o)
    for i, (k, s) in enumerate(zip(kinds, sources)):
        assert np.array_equal(out_r[i], O.expected(og, k, s)), (i, k, s)
    # Per-query edge counts (degrees taken in the caller's numbering).
    n = len(kinds)
    d = cgp_amd.DeviceBuffer(n * h.V * 4)
    cgp_amd.Engine(gr).run_batch(kinds, sources, out=d)
    m = cgp_amd.reached_edges(gr, d, n)
    deg = np.diff(np.asarray(h.offsets)).astype(np.int64)
    for i in range(n):
        assert int(m[i]) == int(deg[out_r[i] != cgp_amd.INF].sum()), i


def test_directed_graph_renumbered_weak_components(gpu):
    # Weak components of a directed graph: the label is the smallest caller
    # id of the component, whatever the internal order.
    rng = np.random.default_rng(7)
    V = 3000
    src = rng.integers(0, V, 9000)
    dst = (src + rng.integers(1, 40, 9000)) % V
    src = np.concatenate([src, np.full(300, 17)])  # one hub, so the degree order moves it
    dst = np.concatenate([dst, rng.integers(0, V, 300)])
    w = rng.integers(1, 30, len(src)).astype(np.uint32)
    h = cgp_amd.HostCSR.from_coo(V, src.astype(np.int32), dst.astype(np.int32), w)
    gr = cgp_amd.Graph.from_host(h, renumber=True)
    assert not gr.symmetric
    og = O.G.of(h)
    kinds = [WCC, BFS, SSSP, BFS]
    sources = [0, 17, 17, int(h.sample_sources(1, seed=3)[0])]
    out = cgp_amd.Engine(gr).run_batch(kinds, sources)[0]
    for i, (k, s) in enumerate(zip(kinds, sources)):
        assert np.array_equal(out[i], O.expected(og, k, s)), i


def test_static_entry_on_a_renumbered_graph(gpu):
    h = cgp_amd.HostCSR.rmat(11, 16, seed=52, wmin=1, wmax=40)
    gr = cgp_amd.Graph.from_host(h, renumber=True)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(40, seed=6)]
    kinds = [BFS] * 20 + [SSSP] * 20 + [WCC]
    sources = src + [0]
    en = cgp_amd.Engine(gr)
    d = cgp_amd.DeviceBuffer(len(kinds) * h.V * 4)
    en.run_batch_static(kinds, sources, d)
    out = d.to_host(shape=(len(kinds), h.V))
    for i, (k, s) in enumerate(zip(kinds, sources)):
        assert np.array_equal(out[i], O.expected(og, k, s)), i


def test_serving_on_a_renumbered_graph(gpu):
    h = cgp_amd.HostCSR.rmat(12, 16, seed=53, wmin=1, wmax=60)
    gr = cgp_amd.Graph.from_host(h, renumber=True)
    og = O.G.of(h)
    deg = np.diff(np.asarray(h.offsets))
    iso = [int(x) for x in np.flatnonzero(deg == 0)[:2]]
    src = [int(x) for x in h.sample_sources(50, seed=8)] + iso
    kinds = [(BFS, SSSP)[i % 2] for i in range(len(src))]
    en = cgp_amd.Engine(gr)
    en.serve_start(max_slots=40)
    done = {}
    nxt = steps = 0
    wcc_sent = False
    while len(done) < len(src) + 1:
        steps += 1
        assert steps < 20000, "serving made no progress"
        for _ in range(4):
            if nxt < len(src):
                en.serve_submit(kinds[nxt], src[nxt], nxt, en.serve_clock())
                nxt += 1
        if nxt >= 20 and not wcc_sent:
            en.serve_submit(WCC, 0, 1000, en.serve_clock())
            wcc_sent = True
        for c in en.serve_step(max_iterations=1):
            done[c["ticket"]] = c
    en.serve_stop()
    for i in range(len(src)):
        assert np.array_equal(done[i]["result"], O.expected(og, kinds[i], src[i])), i
    assert np.array_equal(done[1000]["result"], O.wcc_uf(og))


def test_wcc_admitted_while_seeds_wait(gpu):
    # Four queries from sources without edges converge in the second
    # iteration; the same step then admits the four queued ones, whose seeds
    # wait in the touched list for the next iteration.  A WCC request admitted
    # at the start of the next step computes its labels on the renumbered
    # graph (per-root smallest caller ids in a scratch array) and must leave
    # those seeds in place.
    h = cgp_amd.HostCSR.rmat(12, 16, seed=41, wmin=1, wmax=50)
    deg = np.diff(np.asarray(h.offsets))
    iso = [int(x) for x in np.flatnonzero(deg == 0)[:4]]
    assert len(iso) == 4
    g = cgp_amd.Graph.from_host(h, renumber=True)
    og = O.G.of(h)
    src = iso + [int(x) for x in h.sample_sources(4, seed=9)]
    kinds = [BFS, SSSP, BFS, SSSP, BFS, SSSP, SSSP, BFS]
    en = cgp_amd.Engine(g)
    en.serve_start(max_slots=4)
    for i in range(8):
        en.serve_submit(kinds[i], src[i], i, 0.0)
    done = {}
    for _ in range(2):
        for c in en.serve_step(max_iterations=1):
            done[c["ticket"]] = c
    assert en.serve_pending() == (0, 4)  # the last four hold the slots, seeded
    say("WCC admitted with four seeds waiting")
    en.serve_submit(WCC, 0, 100, 0.0)
    steps = 0
    while len(done) < 9:
        steps += 1
        assert steps < 5000, "serving made no progress"
        for c in en.serve_step(0):
            done[c["ticket"]] = c
    en.serve_stop()
    for i in range(8):
        assert np.array_equal(done[i]["result"], O.expected(og, kinds[i], src[i])), i
    assert np.array_equal(done[100]["result"], O.wcc_uf(og))
