"""The drop-in boundary on the GPU: a graph over caller-owned device arrays
(the reference's non-owning csr_graph_view, puercgp/include/puercgp/backend/
csr_graph.hxx:9-20) and per-call options (cgp_bfs::run(G, sources, ctx,
options), cgp/include/gunrock/algorithms/cgp_bfs.hxx:1050-1056), both through
the C-ABI."""
import json
import os
import sys

import numpy as np
import pytest

import cgp_amd
import oracle_lib as O
from cgp_amd import BFS, SSSP, WCC

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_graph(h):
    off = cgp_amd.DeviceBuffer.from_array(np.asarray(h.offsets, dtype=np.int32))
    idx = cgp_amd.DeviceBuffer.from_array(np.asarray(h.indices, dtype=np.int32))
    wt = None if h.weights is None else cgp_amd.DeviceBuffer.from_array(np.asarray(h.weights, dtype=np.uint32))
    return cgp_amd.Graph.from_device(h.V, h.E, off, idx, wt)


def test_device_arrays_graph_equals_host_graph_and_oracle(gpu):
    h = cgp_amd.HostCSR.rmat(13, 16, seed=11, wmin=1, wmax=64)
    say("device graph: RMAT-13 V=%d E=%d" % (h.V, h.E))
    gd = device_graph(h)
    gh = cgp_amd.Graph.from_host(h)
    assert (gd.V, gd.E, gd.symmetric) == (gh.V, gh.E, True)
    src = [int(x) for x in h.sample_sources(40, seed=11)]
    kinds = [BFS] * 20 + [SSSP] * 20 + [WCC]
    sources = src + [0]
    n = len(kinds)
    out_d = cgp_amd.DeviceBuffer(n * h.V * 4)
    cgp_amd.Engine(gd).run_batch(kinds, sources, out=out_d)
    got = out_d.to_host(shape=(n, h.V))
    ref, _ = cgp_amd.Engine(gh).run_batch(kinds, sources)
    assert np.array_equal(got, ref)
    og = O.G.of(h)
    for i in (0, 19, 20, 39, 40):
        assert np.array_equal(got[i], O.expected(og, kinds[i], sources[i])), "row %d" % i


def test_device_arrays_toy_golden_and_directed(gpu):
    # validate_hybrid.cu's graph T1 (directed) from device arrays: golden
    # answers, and it is not taken for symmetric.
    t = GOLD["graphs"]["T1"]
    off = cgp_amd.DeviceBuffer.from_array(np.asarray(t["offsets"], dtype=np.int32))
    idx = cgp_amd.DeviceBuffer.from_array(np.asarray(t["indices"], dtype=np.int32))
    wt = cgp_amd.DeviceBuffer.from_array(np.asarray(t["weights"], dtype=np.uint32))
    g = cgp_amd.Graph.from_device(len(t["offsets"]) - 1, len(t["indices"]), off, idx, wt)
    assert not g.symmetric
    en = cgp_amd.Engine(g)
    for case in GOLD["cases"]:
        if case["graph"] != "T1" or "max_iterations" in case:
            continue
        kind = {"bfs": BFS, "sssp": SSSP, "wcc": WCC}[case["kind"]]
        assert en.run_batch([kind], [case["source"]])[0][0].tolist() == case["expect"], case["cite"]


def test_device_arrays_unsorted_rows_run_as_directed(gpu):
    # A symmetric path 0-1-2-3 whose row of vertex 1 is stored unsorted: the
    # graph runs as directed (every edge hooks in WCC), with the same answers.
    off = np.array([0, 1, 3, 5, 6], dtype=np.int32)
    idx = np.array([1, 2, 0, 1, 3, 2], dtype=np.int32)
    g = cgp_amd.Graph.from_device(4, 6, cgp_amd.DeviceBuffer.from_array(off), cgp_amd.DeviceBuffer.from_array(idx))
    assert not g.symmetric
    out, _ = cgp_amd.Engine(g).run_batch([BFS, WCC], [0, 0])
    assert out[0].tolist() == [0, 1, 2, 3]
    assert out[1].tolist() == [0, 0, 0, 0]


def test_device_arrays_rejected_when_malformed(gpu):
    good_off = np.array([0, 1, 2], dtype=np.int32)
    cases = [
        (np.array([0, 2, 2], dtype=np.int32), np.array([1, 0], dtype=np.int32), 3, "offsets do not span"),
        (np.array([0, 2, 1, 3], dtype=np.int32), np.array([1, 0, 0], dtype=np.int32), 3, "not monotone"),
        (good_off, np.array([1, 5], dtype=np.int32), 2, "out of range"),
    ]
    for off, idx, E, msg in cases:
        V = len(off) - 1
        with pytest.raises(cgp_amd.CgpError) as ei:
            cgp_amd.Graph.from_device(V, E, cgp_amd.DeviceBuffer.from_array(off),
                                      cgp_amd.DeviceBuffer.from_array(idx))
        assert ei.value.status == cgp_amd.CGP_EINVAL and msg in str(ei.value), str(ei.value)


def test_per_call_options(gpu):
    # smoke_hybrid_push.cu: one level of BFS / SSSP on T1, set per call on an
    # engine created without a limit; the next call without a limit runs to
    # convergence again.
    t = GOLD["graphs"]["T1"]
    g = cgp_amd.Graph.from_csr(t["offsets"], t["indices"], t["weights"])
    en = cgp_amd.Engine(g)
    full = en.run_batch([BFS, SSSP], [0, 0])[0]
    en.set_opts(max_iterations=1)
    one = en.run_batch([BFS, SSSP], [0, 0])[0]
    cases = {c["name"]: c["expect"] for c in GOLD["cases"]}
    assert one[0].tolist() == cases["smoke_hybrid_push_bfs_one_level"]
    assert one[1].tolist() == cases["smoke_hybrid_push_sssp_one_level"]
    en.set_opts(max_iterations=0)
    assert np.array_equal(en.run_batch([BFS, SSSP], [0, 0])[0], full)
    # Launch parameters change between runs (the captured graphs are rebuilt).
    h = cgp_amd.HostCSR.rmat(12, 16, seed=3, wmin=1, wmax=64)
    gr = cgp_amd.Graph.from_host(h)
    en2 = cgp_amd.Engine(gr)
    src = [int(x) for x in h.sample_sources(48, seed=3)]
    kinds = [SSSP] * 24 + [BFS] * 24
    base = en2.run_batch(kinds, src)[0]
    for kw in ({"block_size": 512, "dense_unroll": 4}, {"traversal_mode": cgp_amd.PULL}, {"sssp_delta": 64},
               {"max_words": 1, "block_size": 128}):
        en2.set_opts(**kw)
        o = en2.opts
        for k, v in kw.items():
            assert o[k] == v, (k, o[k], v)
        assert np.array_equal(en2.run_batch(kinds, src)[0], base), kw
    with pytest.raises(cgp_amd.CgpError) as ei:
        en2.set_opts(block_size=100)
    assert ei.value.status == cgp_amd.CGP_EINVAL
    # Options cannot change under a serving session.
    en2.serve_start(max_slots=64, keep_results=0)
    with pytest.raises(cgp_amd.CgpError) as ei:
        en2.set_opts(max_iterations=3)
    assert ei.value.status == cgp_amd.CGP_EBUSY
    en2.serve_stop()


def test_device_to_device_copy(gpu):
    # cgp_copy_device: the library's 16-byte copy kernel for aligned sizes,
    # the runtime's copy otherwise.
    a = np.arange(1 << 20, dtype=np.uint32) * np.uint32(2654435761)
    src = cgp_amd.DeviceBuffer.from_array(a)
    dst = cgp_amd.DeviceBuffer(a.nbytes)
    dst.copy_from(src)
    assert np.array_equal(dst.to_host(), a)
    odd = cgp_amd.DeviceBuffer(a.nbytes)
    odd.copy_from(src, a.nbytes - 12)
    got = odd.to_host()
    assert np.array_equal(got[: a.size - 3], a[: a.size - 3])
    with pytest.raises(ValueError):
        cgp_amd.DeviceBuffer(16).copy_from(src)
