"""Pins the CPU oracle (oracle/oracle.c) against the reference's golden
vectors, the reference's own CPU validators compiled from its sources
(oracle/_ref), brute force, closed forms and invariants."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from checks import check_bfs, check_sssp, check_wcc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
INF = O.INF


def toy(name):
    t = GOLD["graphs"][name]
    return O.G(t["offsets"], t["indices"], t["weights"])


def random_graph(rng, V, E, wmax=20, symmetric=False, allow_dups=True, self_loops=True):
    s = rng.integers(0, V, size=E)
    d = rng.integers(0, V, size=E)
    if not self_loops:
        keep = s != d
        s, d = s[keep], d[keep]
    w = rng.integers(1, wmax + 1, size=len(s))
    if symmetric:
        s, d, w = np.concatenate([s, d]), np.concatenate([d, s]), np.concatenate([w, w])
    if not allow_dups:
        key = s.astype(np.int64) * V + d
        _, first = np.unique(key, return_index=True)
        s, d, w = s[first], d[first], w[first]
    order = np.lexsort((d, s))
    s, d, w = s[order], d[order], w[order]
    off = np.zeros(V + 1, dtype=np.int64)
    np.add.at(off, s + 1, 1)
    return O.G(np.cumsum(off), d, w)


@pytest.mark.parametrize("case", [c for c in GOLD["cases"] if "max_iterations" not in c],
                         ids=lambda c: c["name"])
def test_golden_toy_cases(case):
    g = toy(case["graph"])
    kind = {"bfs": 0, "sssp": 1, "wcc": 2}[case["kind"]]
    got = O.expected(g, kind, case["source"])
    assert got.tolist() == case["expect"], case["cite"]


@pytest.mark.parametrize("mm", GOLD["mm_cases"], ids=lambda c: c["file"])
def test_golden_matrix_market(mm):
    import cgp_amd
    h = cgp_amd.HostCSR.load_matrix_market(os.path.join(HERE, "golden", mm["file"]))
    g = O.G.of(h)
    for src, exp in mm["expect"].items():
        got = O.bfs(g, int(src)) if mm["kind"] == "bfs" else O.sssp(g, int(src))
        assert got.tolist() == exp, (mm["cite"], src)


def test_golden_matrix_market_graph_shape():
    import cgp_amd
    h = cgp_amd.HostCSR.load_matrix_market(os.path.join(HERE, "golden", "weighted.mm"))
    assert (h.V, h.E) == (4, 5)
    assert h.offsets.tolist() == [0, 2, 4, 5, 5]
    assert h.indices.tolist() == [1, 2, 2, 3, 3]
    assert h.weights.tolist() == [1, 4, 1, 5, 1]


needs_ref = pytest.mark.skipif(O.ref is None, reason="oracle/_ref not built (reference tree absent)")


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_reference_validators_random(seed):
    rng = np.random.default_rng(seed)
    g = random_graph(rng, V=int(rng.integers(2, 300)), E=int(rng.integers(0, 1500)))
    for s in rng.integers(0, g.V, size=5):
        b = O.bfs(g, s)
        assert np.array_equal(b, O.ref_call("ref_puer_bfs", g, s))
        assert np.array_equal(b, O.ref_call("ref_cgp_bfs", g, s))
        d = O.sssp(g, s)
        assert np.array_equal(d, O.ref_call("ref_puer_sssp", g, s))
        assert np.array_equal(d, O.ref_call("ref_cgp_sssp", g, s))
    # The label-propagation restatement equals cpu_wcc, directed inputs included.
    assert np.array_equal(O.wcc_lp(g, 10000), O.ref_call("ref_puer_wcc", g, 10000))


@needs_ref
def test_oracle_matches_reference_validators_rmat():
    import cgp_amd
    h = cgp_amd.HostCSR.rmat(11, 16, seed=7, wmin=1, wmax=255)
    g = O.G.of(h)
    for s in h.sample_sources(6, 3):
        assert np.array_equal(O.bfs(g, s), O.ref_call("ref_puer_bfs", g, s))
        assert np.array_equal(O.sssp(g, s), O.ref_call("ref_cgp_sssp", g, s))
        assert np.array_equal(O.sssp(g, s), O.ref_call("ref_puer_sssp", g, s))
    lab = O.ref_call("ref_puer_wcc", g, 10000)
    assert np.array_equal(O.wcc_uf(g), lab)
    assert np.array_equal(O.wcc_lp(g), lab)


@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_bellman_ford(seed):
    rng = np.random.default_rng(100 + seed)
    g = random_graph(rng, V=int(rng.integers(1, 64)), E=int(rng.integers(0, 300)), wmax=9)
    for s in range(min(g.V, 8)):
        assert np.array_equal(O.sssp(g, s), O.bellman_ford(g, s))
        assert np.array_equal(O.bfs(g, s), O.bellman_ford(g, s, unit=True))
        assert np.array_equal(O.sssp(g, s, unit=True), O.bfs(g, s))


def test_closed_forms():
    # Path 0-1-...-9 (undirected), weights 1..9 on consecutive edges.
    n = 10
    s, d, w = [], [], []
    for i in range(n - 1):
        s += [i, i + 1]
        d += [i + 1, i]
        w += [i + 1, i + 1]
    order = np.lexsort((d, s))
    s, d, w = np.array(s)[order], np.array(d)[order], np.array(w)[order]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, s + 1, 1)
    g = O.G(np.cumsum(off), d, w)
    assert O.bfs(g, 0).tolist() == list(range(n))
    assert O.sssp(g, 0).tolist() == [k * (k + 1) // 2 for k in range(n)]
    assert O.sssp(g, 9).tolist() == [45 - k * (k + 1) // 2 for k in range(n)]
    assert O.wcc_uf(g).tolist() == [0] * n
    # Two components: star around 5 plus isolated 0..4 -> labels.
    g2 = O.G([0, 0, 0, 0, 0, 0, 3, 4, 5, 6], [6, 7, 8, 5, 5, 5])
    assert O.wcc_uf(g2).tolist() == [0, 1, 2, 3, 4, 5, 5, 5, 5]
    assert O.bfs(g2, 6).tolist() == [INF] * 5 + [1, 0, 2, 2]
    # m_q: a full reach counts every edge, an isolated source counts none.
    assert O.reached_edges(g, O.bfs(g, 3)) == g.E
    assert O.reached_edges(g2, O.bfs(g2, 0)) == 0


def test_lp_equals_union_find_on_symmetric():
    rng = np.random.default_rng(5)
    for _ in range(5):
        g = random_graph(rng, V=200, E=150, symmetric=True, self_loops=False)
        assert np.array_equal(O.wcc_lp(g), O.wcc_uf(g))


@pytest.mark.parametrize("seed", range(3))
def test_invariant_checks_accept_oracle_and_reject_planted_mistakes(seed):
    import cgp_amd
    h = cgp_amd.HostCSR.rmat(10, 8, seed=seed + 1, wmin=1, wmax=64)
    g = O.G.of(h)
    s = int(h.sample_sources(1, seed)[0])
    b, d, lab = O.bfs(g, s), O.sssp(g, s), O.wcc_uf(g)
    assert check_bfs(g.off, g.idx, s, b) is None
    assert check_sssp(g.off, g.idx, g.wt, s, d) is None
    assert check_wcc(g.off, g.idx, lab) is None
    reach = np.flatnonzero((b != INF) & (b > 0))
    v = int(reach[len(reach) // 2])
    # Planted mistakes: a depth off by one, a distance one too small or too
    # large, an unreached vertex, a wrong component label.
    for delta in (-1, 1):
        bb = b.copy()
        bb[v] = np.uint32(int(bb[v]) + delta)
        assert check_bfs(g.off, g.idx, s, bb) is not None
        dd = d.copy()
        dd[v] = np.uint32(int(dd[v]) + delta)
        assert check_sssp(g.off, g.idx, g.wt, s, dd) is not None
    bb = b.copy()
    bb[v] = INF
    assert check_bfs(g.off, g.idx, s, bb) is not None
    ll = lab.copy()
    ll[v] = v if lab[v] != v else lab[v] + 1
    assert check_wcc(g.off, g.idx, ll) is not None


def test_threaded_runner_matches_single_queries():
    import cgp_amd
    h = cgp_amd.HostCSR.rmat(10, 8, seed=3, wmin=1, wmax=30)
    g = O.G.of(h)
    src = h.sample_sources(6, 9)
    kinds = [0, 1, 2, 0, 1, 0]
    secs, out, m = O.run_queries(g, kinds, src, threads=3, keep=True)
    assert secs >= 0
    for i, (k, s) in enumerate(zip(kinds, src)):
        exp = O.expected(g, k, s) if k != 2 else O.wcc_lp(g)
        assert np.array_equal(out[i], exp)
        assert m[i] == (g.E if k == 2 else O.reached_edges(g, exp))


def planted_variants(g, s, b, d, lab):
    """(kind, row) pairs of wrong answers: a BFS depth / SSSP distance off by one
    either way, a reachable vertex left unreached, an unreachable vertex given
    a finite value, a second zero, and a wrong component label."""
    out = []
    reach = np.flatnonzero((b != INF) & (b > 0))
    v = int(reach[len(reach) // 2])
    for delta in (-1, 1):
        bb = b.copy()
        bb[v] = np.uint32(int(bb[v]) + delta)
        out.append((0, bb))
        dd = d.copy()
        dd[v] = np.uint32(int(dd[v]) + delta)
        out.append((1, dd))
    for k, base in ((0, b), (1, d)):
        x = base.copy()
        x[v] = INF
        out.append((k, x))
        x = base.copy()
        x[v] = 0
        out.append((k, x))
    unreach = np.flatnonzero(b == INF)
    if len(unreach):
        for k, base in ((0, b), (1, d)):
            x = base.copy()
            x[int(unreach[0])] = 5
            out.append((k, x))
    ll = lab.copy()
    ll[v] = v if lab[v] != v else lab[v] + 1
    out.append((2, ll))
    return out


@pytest.mark.parametrize("seed", range(3))
def test_c_checkers_accept_oracle_and_reject_planted_mistakes(seed):
    import cgp_amd
    h = cgp_amd.HostCSR.rmat(10, 4, seed=seed + 1, wmin=1, wmax=64)  # sparse: several components
    g = O.G.of(h)
    s = int(h.sample_sources(1, seed)[0])
    b, d, lab = O.bfs(g, s), O.sssp(g, s), O.wcc_uf(g)
    assert O.check_sssp(g, s, b, unit=True) is None
    assert O.check_sssp(g, s, d) is None
    assert O.check_wcc(g, lab) is None
    assert np.any(b == INF), "the planted unreachable case needs a second component"
    for kind, row in planted_variants(g, s, b, d, lab):
        err = O.check_wcc(g, row) if kind == 2 else O.check_sssp(g, s, row, unit=(kind == 0))
        assert err is not None, kind
    # The threaded row checker agrees with the single-row checkers.
    rows = [(0, b), (1, d), (2, lab)] + planted_variants(g, s, b, d, lab)
    bad = O.check_rows(g, [k for k, _ in rows], [s] * len(rows), np.stack([r for _, r in rows]), threads=3)
    assert [i for i, _ in bad] == list(range(3, len(rows)))


def test_wcc_checkers_reject_merged_components():
    # Components {0} and {1, 2}: labelling every vertex 0 satisfies "equal across
    # edges", "label <= id" and "label is a fixed point", but is wrong.
    g = O.G([0, 0, 1, 2], [2, 1])
    wrong = np.zeros(3, dtype=np.uint32)
    assert O.check_wcc(g, wrong) is not None
    assert check_wcc(g.off, g.idx, wrong) is not None
    right = np.array([0, 1, 1], dtype=np.uint32)
    assert np.array_equal(O.wcc_uf(g), right)
    assert O.check_wcc(g, right) is None and check_wcc(g.off, g.idx, right) is None


@pytest.mark.parametrize("seed", range(4))
def test_c_checkers_on_directed_random_graphs(seed):
    rng = np.random.default_rng(seed + 40)
    g = random_graph(rng, V=60, E=120, wmax=9, symmetric=False, self_loops=True)
    for s in range(0, 60, 7):
        assert O.check_sssp(g, s, O.sssp(g, s)) is None
        assert O.check_sssp(g, s, O.bfs(g, s), unit=True) is None
        assert O.check_sssp(g, s, O.bellman_ford(g, s)) is None
    lab = O.wcc_uf(g)
    assert O.check_wcc(g, lab) is None and check_wcc(g.off, g.idx, lab) is None
