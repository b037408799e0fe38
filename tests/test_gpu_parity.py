"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle,
bit-exact, on the reference's golden vectors and on seeded graphs."""
import json
import os
import sys

import numpy as np
import pytest

import cgp_amd
import oracle_lib as O
from cgp_amd import BFS, SSSP, WCC, INF
from checks import check_bfs, check_sssp, check_wcc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def toy(name):
    t = GOLD["graphs"][name]
    return cgp_amd.Graph.from_csr(t["offsets"], t["indices"], t["weights"]), O.G(t["offsets"], t["indices"], t["weights"])


_cache = {}


def rmat(scale, ef, seed, wmin=1, wmax=64):
    key = (scale, ef, seed, wmin, wmax)
    if key not in _cache:
        h = cgp_amd.HostCSR.rmat(scale, ef, seed=seed, wmin=wmin, wmax=wmax)
        _cache[key] = (h, cgp_amd.Graph.from_host(h), O.G.of(h))
    return _cache[key]


def oracle_rows(g, kinds, sources):
    return np.stack([O.expected(g, k, s) for k, s in zip(kinds, sources)])


@pytest.mark.parametrize("case", GOLD["cases"], ids=lambda c: c["name"])
@pytest.mark.parametrize("mode", [cgp_amd.PUSH, cgp_amd.HYBRID])
def test_golden_toy(gpu, case, mode):
    say("golden", case["name"], "mode", mode)
    g, _ = toy(case["graph"])
    kind = {"bfs": BFS, "sssp": SSSP, "wcc": WCC}[case["kind"]]
    en = cgp_amd.Engine(g, traversal_mode=mode, max_iterations=case.get("max_iterations", 0))
    got = en.run_batch([kind], [case["source"]])[0][0]
    assert got.tolist() == case["expect"], case["cite"]


def test_golden_mixed_toy_batches(gpu):
    # validate_hybrid.cu: BFS + SSSP + WCC in one batch on T1; validate_replenish.cu's
    # four queries on T2 in one batch.
    g, _ = toy("T1")
    out = cgp_amd.Engine(g).run_batch([BFS, SSSP, WCC], [0, 0, 0])[0]
    assert out.tolist() == [[0, 1, 1, 2], [0, 2, 5, 3], [0, 0, 0, 0]]
    g, _ = toy("T2")
    out = cgp_amd.Engine(g).run_batch([BFS, SSSP, BFS, SSSP], [0, 0, 1, 1])[0]
    assert out.tolist() == [[0, 1, 1, 2], [0, 2, 5, 5], [INF, 0, INF, 1], [INF, 0, INF, 3]]


@pytest.mark.parametrize("kind", ["min_reduce", "first_write"])
def test_reduce_op_vectors_through_truncated_runs(gpu, kind):
    # smoke_reduce_ops.cu:30-144: a value slot under a sequence of candidates,
    # each entry [slot before, candidate, improved, old value returned].  Here
    # the slot is vertex t of a small graph whose paths deliver the candidates
    # in that order, one iteration apart, and runs cut after k iterations
    # (max_iterations) show the slot after each step.
    seq = GOLD["reduce_ops"][kind]
    for before, cand, improved, old in seq:
        assert old == before and improved == (cand < before if kind == "min_reduce" else before == INF)
    if kind == "min_reduce":
        # SSSP from 0 into t = 1: edge 0-1 (10) in iteration 1, 0-2-1 (1 + 4 = 5)
        # and 0-3-1 (2 + 6 = 8) in iteration 2, 0-4-5-1 (1 + 1 + 1 = 3) in
        # iteration 3; the slot starts at the golden first value, 10.
        e = [(0, 1, 10), (0, 2, 1), (2, 1, 4), (0, 3, 2), (3, 1, 6), (0, 4, 1), (4, 5, 1), (5, 1, 1)]
        expect = {1: seq[0][0], 2: min(seq[0][0], seq[0][1], seq[1][1]), 3: seq[2][1]}
        kinds, t = [SSSP], 1
    else:
        # BFS from 0: t = 5 is first reached at depth 5 along 0-6-7-8-9-5 and
        # again at depth 7 along 0-1-2-3-4-10-11-5; the first write (the golden
        # slot's value after its first candidate, 5) stays.  A later smaller
        # candidate, the golden second step, cannot occur in level order: the
        # later one here is larger and is rejected in the same way.
        e = [(0, 6, 1), (6, 7, 1), (7, 8, 1), (8, 9, 1), (9, 5, 1),
             (0, 1, 1), (1, 2, 1), (2, 3, 1), (3, 4, 1), (4, 10, 1), (10, 11, 1), (11, 5, 1)]
        expect = {4: INF, 5: seq[0][1], 8: seq[1][0], 0: seq[1][3]}
        kinds, t = [BFS], 5
    V = 1 + max(max(a, b) for a, b, _ in e)
    src = [a for a, b, _ in e] + [b for a, b, _ in e]
    dst = [b for a, b, _ in e] + [a for a, b, _ in e]
    w = [x for _, _, x in e] * 2
    h = cgp_amd.HostCSR.from_coo(V, src, dst, w)
    g = cgp_amd.Graph.from_host(h)
    for k, val in expect.items():
        out = cgp_amd.Engine(g, max_iterations=k).run_batch(kinds, [0])[0]
        assert int(out[0, t]) == val, (kind, k, int(out[0, t]), val)


@pytest.mark.parametrize("mm", GOLD["mm_cases"], ids=lambda c: c["file"])
def test_golden_matrix_market(gpu, mm):
    h = cgp_amd.HostCSR.load_matrix_market(os.path.join(HERE, "golden", mm["file"]))
    g = cgp_amd.Graph.from_host(h)
    kind = BFS if mm["kind"] == "bfs" else SSSP
    srcs = sorted(int(s) for s in mm["expect"])
    out = cgp_amd.Engine(g).run_batch([kind] * len(srcs), srcs)[0]
    for i, s in enumerate(srcs):
        assert out[i].tolist() == mm["expect"][str(s)], mm["cite"]


@pytest.mark.parametrize("mode", [cgp_amd.PUSH, cgp_amd.PULL, cgp_amd.HYBRID])
@pytest.mark.parametrize("nq", [1, 7, 64, 130])
def test_rmat_bfs_batches(gpu, mode, nq):
    say("rmat bfs mode", mode, "nq", nq)
    h, g, og = rmat(12, 16, 3)
    src = h.sample_sources(nq, seed=nq)
    out = cgp_amd.Engine(g, traversal_mode=mode).run_batch([BFS] * nq, src)[0]
    assert np.array_equal(out, oracle_rows(og, [0] * nq, src))


@pytest.mark.parametrize("nq", [1, 3])
def test_hybrid_pull_steps_for_small_batches(gpu, nq):
    # pull_max_queries lets hybrid mode take bottom-up steps (the reference's
    # frontier / edge-ratio thresholds); results stay bit-exact.
    h, g, og = rmat(14, 16, 3)
    src = h.sample_sources(nq, seed=11)
    out, st = cgp_amd.Engine(g, pull_max_queries=4).run_batch([BFS] * nq, src)
    assert st["pull_iterations"] > 0
    assert np.array_equal(out, oracle_rows(og, [0] * nq, src))
    out2, st2 = cgp_amd.Engine(g).run_batch([BFS] * nq, src)
    assert st2["pull_iterations"] == 0 and np.array_equal(out2, out)


@pytest.mark.parametrize("nq", [1, 5, 16, 64, 100])
def test_rmat_sssp_batches(gpu, nq):
    say("rmat sssp nq", nq)
    h, g, og = rmat(12, 16, 4, 1, 255)
    src = h.sample_sources(nq, seed=nq + 1)
    out = cgp_amd.Engine(g).run_batch([SSSP] * nq, src)[0]
    assert np.array_equal(out, oracle_rows(og, [1] * nq, src))


def test_rmat_wcc(gpu):
    h, g, og = rmat(13, 4, 5)
    out = cgp_amd.Engine(g).run_batch([WCC] * 3, [0] * 3)[0]
    exp = O.wcc_uf(og)
    for r in out:
        assert np.array_equal(r, exp)


@pytest.mark.parametrize("hip_graph", [0, 1])
@pytest.mark.parametrize("profile", [0, 1, 3])
def test_mixed_uneven_batch(gpu, hip_graph, profile):
    # 160 BFS + 80 SSSP + 16 WCC in one batch (4 mask words, a mixed word),
    # shuffled order; results must equal the oracle row by row.
    say("mixed hip_graph", hip_graph, "profile", profile)
    h, g, og = rmat(12, 16, 6, 1, 255)
    src = h.sample_sources(240, seed=3)
    kinds = [BFS] * 160 + [SSSP] * 80 + [WCC] * 16
    sources = list(src) + [0] * 16
    perm = np.random.default_rng(1).permutation(len(kinds))
    kinds = [kinds[i] for i in perm]
    sources = [int(sources[i]) for i in perm]
    en = cgp_amd.Engine(g, use_hip_graph=hip_graph, profile=profile)
    out, st = en.run_batch(kinds, sources)
    exp = oracle_rows(og, kinds, sources)
    assert np.array_equal(out, exp)
    assert st["passes"] == 1
    if profile & 1:
        assert st["advance_ms"] > 0 and st["advance_launches"] == st["iterations"]
        assert st["filter_ms"] > 0
    if profile & 2:
        assert st["advance_bytes"] > 0 and st["edges_scanned"] > 0


@pytest.mark.parametrize("unroll", [4, 8, 16])
@pytest.mark.parametrize("dense_t", [1, 4, 65])
@pytest.mark.parametrize("delta", [0, 1, 100])
def test_mixed_lane_per_query_variants(gpu, unroll, dense_t, delta):
    # dense_threshold 1 sends every SSSP word to the lane-per-query kernel,
    # 65 none; both row-unroll depths; near-far thresholds off, at the
    # smallest step and at a mid step.
    say("mixed unroll", unroll, "dense_t", dense_t, "delta", delta)
    h, g, og = rmat(12, 16, 6, 1, 255)
    src = h.sample_sources(240, seed=5)
    kinds = [BFS] * 100 + [SSSP] * 140 + [WCC] * 2
    sources = list(src) + [0] * 2
    en = cgp_amd.Engine(g, dense_unroll=unroll, dense_threshold=dense_t, sssp_delta=delta)
    out, st = en.run_batch(kinds, sources)
    assert np.array_equal(out, oracle_rows(og, kinds, sources))


@pytest.mark.parametrize("unroll", [4, 16])
@pytest.mark.parametrize("dense_t,div", [(1, 8), (8, 8), (1, 1), (1, 1000000)])
def test_lane_per_query_wide_addresses(gpu, unroll, dense_t, div):
    # 64-bit row addresses in the lane-per-query loop (the form used where one
    # word's value rows pass 4 GiB), with two full SSSP words and one shared
    # with BFS slots, in the automatic, sparse (div 1) and dense (div 1e6)
    # frontier modes.
    say("wide addresses unroll", unroll, "dense_t", dense_t, "div", div)
    h, g, og = rmat(12, 16, 8, 1, 255)
    src = h.sample_sources(240, seed=div % 89)
    kinds = [BFS] * 100 + [SSSP] * 140 + [WCC] * 2
    sources = list(src) + [0] * 2
    en = cgp_amd.Engine(g, dense_unroll=unroll, dense_threshold=dense_t, frontier_dense_div=div,
                        sssp_delta=0, dense_wide_addr=1)
    out, _ = en.run_batch(kinds, sources)
    assert np.array_equal(out, oracle_rows(og, kinds, sources))


@pytest.mark.parametrize("layout", ["v_odd", "aligned", "offset4"])
def test_output_transpose_layouts(gpu, layout):
    # Query-major results through both output transposes with 64-slot rows:
    # 16-byte accesses (V % 4 == 0, 16-byte aligned output) and the
    # per-element path (V % 4 != 0, or an output 4 bytes past alignment).
    say("output layout", layout)
    if layout == "v_odd":
        h = cgp_amd.HostCSR.grid(31, 33, 0.05, seed=3, wmin=1, wmax=50)  # V = 1023
    else:
        h = cgp_amd.HostCSR.rmat(10, 8, seed=4, wmin=1, wmax=50)  # V = 1024
    g, og = cgp_amd.Graph.from_host(h), O.G.of(h)
    src = [int(x) for x in h.sample_sources(100, seed=7)]
    kinds = [(BFS, SSSP)[i % 2] for i in range(100)] + [WCC]
    sources = src + [0]
    exp = oracle_rows(og, kinds, sources)
    en = cgp_amd.Engine(g)
    n, V = len(kinds), g.V
    if layout == "v_odd":
        out, _ = en.run_batch(kinds, sources)
    else:
        off = 0 if layout == "aligned" else 4
        buf = cgp_amd.DeviceBuffer(n * V * 4 + 16)
        en.run_batch(kinds, sources, out=buf.view(off, n * V * 4))
        out = buf.to_host()[off // 4: off // 4 + n * V].reshape(n, V)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("div", [1, 1000000])
@pytest.mark.parametrize("n_sssp,n_bfs", [(16, 0), (12, 100), (16, 48)])
def test_narrow_lane_per_query_words(gpu, div, n_sssp, n_bfs):
    # SSSP slots confined to lanes 0..15 of one word: four edges per wave step
    # (phase N), in the sparse (div 1) and dense (div 1e6) frontier modes.
    say("narrow", n_sssp, n_bfs, "div", div)
    h, g, og = rmat(12, 16, 7, 1, 255)
    src = h.sample_sources(n_sssp + n_bfs, seed=div % 97)
    kinds = [SSSP] * n_sssp + [BFS] * n_bfs
    en = cgp_amd.Engine(g, dense_threshold=1, frontier_dense_div=div)
    out, _ = en.run_batch(kinds, src)
    assert np.array_equal(out, oracle_rows(og, kinds, src))
    off, _ = cgp_amd.Engine(g, dense_threshold=1, frontier_dense_div=div, dense_narrow=0).run_batch(kinds, src)
    assert np.array_equal(out, off)


@pytest.mark.parametrize("delta", [1, 7, 3000])
@pytest.mark.parametrize("nq", [1, 64])
def test_near_far_sssp_rmat(gpu, delta, nq):
    # Near-far postponing on one-word and single-query batches (R < 64 rows).
    say("near-far rmat delta", delta, "nq", nq)
    h, g, og = rmat(13, 8, 9, 1, 255)
    src = h.sample_sources(nq, seed=delta)
    en = cgp_amd.Engine(g, sssp_delta=delta)
    out, _ = en.run_batch([SSSP] * nq, src)
    assert np.array_equal(out, oracle_rows(og, [1] * nq, src))


def test_near_far_grid_long_paths(gpu):
    # Long shortest paths (hundreds of hops) with a small step: many threshold
    # raises, sparse frontiers; against the oracle and the Bellman-Ford run.
    h = cgp_amd.HostCSR.grid(64, 64, p_extra=0.0, seed=3, wmin=1, wmax=1000)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [0, 4095, 2080, 17]
    ref = cgp_amd.Engine(g, sssp_delta=0).run_batch([SSSP] * 4, src)[0]
    out, st = cgp_amd.Engine(g, sssp_delta=50).run_batch([SSSP] * 4, src)
    assert np.array_equal(out, ref)
    assert np.array_equal(out, oracle_rows(og, [1] * 4, src))
    # Automatic step: low-degree weighted graph -> 20 x mean weight, and the
    # same results.
    en = cgp_amd.Engine(g)
    w = np.asarray(h.weights, dtype=np.float64)
    assert en.opts["sssp_delta"] == int(round(20 * w.mean()))
    assert en.opts["dense_threshold"] == 8  # automatic threshold with postponing on
    assert np.array_equal(en.run_batch([SSSP] * 4, src)[0], ref)


def test_near_far_automatic_off_on_small_world(gpu):
    h, g, og = rmat(12, 16, 4, 1, 255)
    en = cgp_amd.Engine(g)
    assert en.opts["sssp_delta"] == 0
    assert en.opts["dense_threshold"] == 5  # automatic threshold with postponing off
    assert cgp_amd.Engine(g, dense_threshold=3).opts["dense_threshold"] == 3


def test_alone_equals_in_the_mix(gpu):
    # Each query's result is identical run alone and inside a mixed batch.
    h, g, og = rmat(11, 16, 8, 1, 100)
    src = h.sample_sources(40, seed=5)
    kinds = [BFS, SSSP] * 20
    en = cgp_amd.Engine(g)
    mix = en.run_batch(kinds, src)[0]
    for i in range(0, 40, 7):
        alone = en.run_batch([kinds[i]], [src[i]])[0][0]
        assert np.array_equal(alone, mix[i])


def test_multi_pass_word_split(gpu):
    # max_words=1 splits 150 queries into three 64-slot passes.
    h, g, og = rmat(11, 8, 9, 1, 30)
    src = h.sample_sources(150, seed=2)
    kinds = [BFS if i % 3 else SSSP for i in range(150)]
    out, st = cgp_amd.Engine(g, max_words=1).run_batch(kinds, src)
    assert st["passes"] == 3
    assert np.array_equal(out, oracle_rows(og, kinds, src))


def test_repeat_and_permutation_determinism(gpu):
    h, g, og = rmat(11, 16, 10, 1, 255)
    src = h.sample_sources(70, seed=7)
    kinds = [SSSP] * 35 + [BFS] * 35
    en = cgp_amd.Engine(g)
    a = en.run_batch(kinds, src)[0]
    b = en.run_batch(kinds, src)[0]
    perm = np.arange(70)[::-1]
    c = en.run_batch([kinds[i] for i in perm], [src[i] for i in perm])[0]
    assert np.array_equal(a, b)
    assert np.array_equal(a, c[perm])


def test_edge_cases(gpu):
    # Duplicate sources, an isolated source, a source with a self-loop-free
    # single edge, and a graph with no edges.
    h = cgp_amd.HostCSR.from_coo(6, [0, 1, 1, 2, 4], [1, 0, 2, 1, 4], [3, 3, 4, 4, 9])
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    kinds = [BFS, BFS, SSSP, SSSP, BFS, SSSP, WCC, BFS]
    srcs = [0, 0, 1, 1, 5, 5, 0, 4]
    out = cgp_amd.Engine(g).run_batch(kinds, srcs)[0]
    assert np.array_equal(out, oracle_rows(og, kinds, srcs))
    g0 = cgp_amd.Graph.from_csr([0, 0, 0, 0], [])
    out = cgp_amd.Engine(g0).run_batch([BFS, SSSP, WCC], [2, 0, 0])[0]
    assert out.tolist() == [[INF, INF, 0], [0, INF, INF], [0, 1, 2]]
    en = cgp_amd.Engine(g)
    with pytest.raises(cgp_amd.CgpError) as e:
        en.run_batch([BFS], [6])
    assert e.value.status == cgp_amd.CGP_EINVAL
    with pytest.raises(cgp_amd.CgpError) as e:
        en.run_batch([], [])
    assert e.value.status == cgp_amd.CGP_EINVAL
    with pytest.raises(cgp_amd.CgpError):
        en.run_batch([7], [0])


@pytest.mark.parametrize("n_sssp,n_bfs", [(64, 64), (64, 130), (80, 48), (8, 8), (3, 5)])
def test_bfs_slots_over_stale_value_rows(gpu, n_sssp, n_bfs):
    # Value rows of BFS slots are not reset between batches (only SSSP slots
    # start at CGP_INF): a BFS row must read CGP_INF wherever its visited bit
    # is clear, also over rows left finite by an earlier SSSP batch.  Two
    # disjoint RMAT-10 components: SSSP queries in A first, then BFS queries
    # in B on the same engine, whose rows must be CGP_INF on all of A.
    a = cgp_amd.HostCSR.rmat(10, 8, seed=31, wmin=1, wmax=50)
    b = cgp_amd.HostCSR.rmat(10, 8, seed=32, wmin=1, wmax=50)
    ra = np.repeat(np.arange(a.V), np.diff(a.offsets))
    rb = np.repeat(np.arange(b.V), np.diff(b.offsets))
    src = np.concatenate([ra, rb + a.V]).astype(np.int32)
    dst = np.concatenate([a.indices, b.indices + a.V]).astype(np.int32)
    wt = np.concatenate([a.weights, b.weights]).astype(np.uint32)
    h = cgp_amd.HostCSR.from_coo(a.V + b.V, src, dst, wt)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    sa = [int(x) for x in a.sample_sources(n_sssp, seed=5)]
    sb = [int(x) + a.V for x in b.sample_sources(n_bfs, seed=6)]
    en = cgp_amd.Engine(g)
    out1 = en.run_batch([SSSP] * n_sssp, sa)[0]
    assert np.array_equal(out1, oracle_rows(og, [SSSP] * n_sssp, sa))
    assert (out1[:, : a.V] != INF).any()
    out2 = en.run_batch([BFS] * n_bfs, sb)[0]
    assert np.all(out2[:, : a.V] == INF)
    assert np.array_equal(out2, oracle_rows(og, [BFS] * n_bfs, sb))
    # Mixed: SSSP slots of the new batch sit in B, BFS slots in A again.
    kinds = [SSSP] * min(n_sssp, 16) + [BFS] * n_bfs
    srcs = [int(x) + a.V for x in b.sample_sources(min(n_sssp, 16), seed=7)] + [int(x) for x in a.sample_sources(n_bfs, seed=8)]
    out3 = en.run_batch(kinds, srcs)[0]
    assert np.array_equal(out3, oracle_rows(og, kinds, srcs))


def test_directed_graph_uses_push_and_weak_components(gpu):
    h = cgp_amd.HostCSR.from_coo(7, [0, 0, 2, 3, 5], [1, 2, 3, 4, 6], [1, 2, 3, 4, 5])
    g = cgp_amd.Graph.from_host(h)
    assert not g.symmetric
    og = O.G.of(h)
    out = cgp_amd.Engine(g, traversal_mode=cgp_amd.PULL).run_batch([BFS, SSSP, WCC], [0, 2, 0])[0]
    assert np.array_equal(out[0], O.bfs(og, 0))
    assert np.array_equal(out[1], O.sssp(og, 2))
    assert out[2].tolist() == [0, 0, 0, 0, 0, 5, 5]


def test_config_rmat16_32_bfs_bit_exact(gpu):
    # RMAT scale 16, ef 16, seed 1; 32 BFS sources uniform over degree >= 1.
    h, g, og = rmat(16, 16, 1, 0, 0)
    src = h.sample_sources(32, seed=1)
    out = cgp_amd.Engine(g).run_batch([BFS] * 32, src)[0]
    assert np.array_equal(out, oracle_rows(og, [0] * 32, src))


def test_device_output_and_reached_edges(gpu):
    h, g, og = rmat(12, 8, 12, 1, 40)
    src = h.sample_sources(20, seed=4)
    kinds = [BFS] * 10 + [SSSP] * 8 + [WCC] * 2
    srcs = list(src[:18]) + [0, 0]
    buf = cgp_amd.DeviceBuffer(20 * g.V * 4)
    _, st = cgp_amd.Engine(g).run_batch(kinds, srcs, out=buf)
    out = buf.to_host(shape=(20, g.V))
    assert np.array_equal(out, oracle_rows(og, kinds, srcs))
    m = cgp_amd.reached_edges(g, buf, 20)
    exp = [O.reached_edges(og, r) for r in out]
    assert m.tolist() == exp


def test_properties_at_scale_rmat20_sssp(gpu):
    # RMAT scale 20, w in [1,64], 64 SSSP: every row passes the exact
    # shortest-path characterisation; a sample of rows equals Dijkstra.
    h, g, og = rmat(20, 16, 2, 1, 64)
    src = h.sample_sources(64, seed=2)
    say("rmat20 sssp start")
    out = cgp_amd.Engine(g).run_batch([SSSP] * 64, src)[0]
    for i in range(0, 64, 9):
        assert check_sssp(og.off, og.idx, og.wt, int(src[i]), out[i]) is None
    for i in (0, 63):
        assert np.array_equal(out[i], O.sssp(og, src[i]))


def test_grid_512_sssp_eight_words(gpu):
    # The high-diameter workload's shape at a small size: a 128 x 128 grid with
    # one random extra edge per vertex w.p. 0.01, weights in [1,1000], 512 SSSP
    # queries = 8 mask words in one pass (long tails of small frontiers).
    h = cgp_amd.HostCSR.grid(128, 128, p_extra=0.01, seed=5, wmin=1, wmax=1000)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = h.sample_sources(512, seed=5)
    out, st = cgp_amd.Engine(g).run_batch([SSSP] * 512, src)
    assert st["passes"] == 1 and st["iterations"] > 50
    for i in range(512):
        assert check_sssp(og.off, og.idx, og.wt, int(src[i]), out[i]) is None, i
    for i in (0, 137, 511):
        assert np.array_equal(out[i], O.sssp(og, src[i]))


def test_grid_bfs_long_tail_sparse_mode(gpu):
    # A plain path-like grid (no shortcuts): BFS depths reach 254, hundreds of
    # iterations with small frontiers in sparse mode.
    h = cgp_amd.HostCSR.grid(128, 128, p_extra=0.0, seed=1)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [0, 128 * 128 - 1, 64 * 128 + 64]
    out, st = cgp_amd.Engine(g).run_batch([BFS, BFS, SSSP], src)
    assert st["iterations"] >= 254
    assert np.array_equal(out[0], O.bfs(og, 0))
    assert np.array_equal(out[1], O.bfs(og, src[1]))
    assert np.array_equal(out[2], O.sssp(og, src[2], unit=True))


def test_run_report_per_iteration(gpu):
    h, g, og = rmat(12, 16, 3)
    src = h.sample_sources(40, seed=40)
    en = cgp_amd.Engine(g, profile=1)
    out, st = en.run_batch([BFS] * 20 + [SSSP] * 20, src)
    rep = en.last_run_report()
    assert len(rep) == st["iterations"]
    assert [r["iteration"] for r in rep] == list(range(st["iterations"]))
    assert rep[0]["frontier_items"] >= 1 and rep[-1]["next_items"] == 0
    assert all(rep[i]["next_items"] == rep[i + 1]["frontier_items"] for i in range(len(rep) - 1))
    assert abs(sum(r["advance_ms"] for r in rep) - st["advance_ms"]) < 1e-6 + 1e-6 * st["advance_ms"]
    assert all(r["dense"] in (0, 1) and r["pull"] in (0, 1) for r in rep)


@pytest.mark.parametrize("out_dev", [False, True])
def test_output_behind_the_expected_replay(gpu, out_dev):
    # A batch enqueues its output behind the replay in which the engine's
    # previous batch finished.  Sources without edges finish in the first
    # replay; a weighted batch then needs several two-iteration replays, so
    # its early output is overwritten once it finishes; the short batch after
    # it finishes before the guessed replay, and its repeat where guessed.
    h, g, og = rmat(12, 16, 41, 1, 50)
    deg = np.diff(np.asarray(h.offsets))
    iso = [int(x) for x in np.flatnonzero(deg == 0)[:3]]
    src = [int(x) for x in h.sample_sources(40, seed=12)]
    en = cgp_amd.Engine(g, graph_iterations=2)
    short = ([BFS, SSSP, BFS], iso)
    long_ = ([BFS] * 20 + [SSSP] * 20 + [WCC], src + [0])
    for kinds, sources in [short, long_, short, short, long_, long_]:
        say("output-behind-replay batch of", len(kinds), "out_dev", out_dev)
        if out_dev:
            buf = cgp_amd.DeviceBuffer(len(kinds) * g.V * 4)
            _, st = en.run_batch(kinds, sources, out=buf)
            out = buf.to_host(shape=(len(kinds), g.V))
        else:
            out, st = en.run_batch(kinds, sources)
        assert np.array_equal(out, oracle_rows(og, kinds, sources))
        if kinds is long_[0]:
            assert st["iterations"] > 4  # more than two replays
