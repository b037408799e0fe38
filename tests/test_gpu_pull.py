"""Pull (direction-optimised) steps for every kind, against the oracle:
bottom-up BFS and min-plus gathers for SSSP (the reference's fused pull,
puercgp/include/puercgp/engine/frontier_engine.hxx:1742-1958 and
hybrid_engine.hxx:337-552, behind the thresholds of core/types.hxx:30-31).
WCC here is hooking plus pointer jumping, which has no frontier, so it gives
the same labels in every traversal mode."""
import json
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


def oracle_rows(og, kinds, sources):
    return np.stack([O.expected(og, k, s) for k, s in zip(kinds, sources)])


@pytest.mark.parametrize("mode", ["pull", "hybrid"])
@pytest.mark.parametrize("shape", ["sssp", "mixed", "wcc"])
def test_pull_modes_equal_the_oracle(gpu, mode, shape):
    h = cgp_amd.HostCSR.rmat(13, 16, seed=31, wmin=1, wmax=200)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(90, seed=31)]
    if shape == "sssp":
        kinds, sources = [SSSP] * 70, src[:70]
    elif shape == "mixed":
        kinds, sources = [SSSP] * 40 + [BFS] * 50 + [WCC] * 3, src + [0] * 3
    else:
        kinds, sources = [WCC] * 4, [0] * 4
    opts = {"traversal_mode": cgp_amd.PULL if mode == "pull" else cgp_amd.HYBRID, "pull_max_queries": 1000,
            "sssp_delta": 0}
    en = cgp_amd.Engine(g, **opts)
    out, st = en.run_batch(kinds, sources)
    say("pull %s %s: %d iterations, %d pull" % (mode, shape, st["iterations"], st["pull_iterations"]))
    assert np.array_equal(out, oracle_rows(og, kinds, sources))
    if shape != "wcc":
        assert st["pull_iterations"] > 0
        if mode == "pull":
            assert st["push_iterations"] == 0
        modes = json.loads(en.last_run_report_text("json"))["level_modes"]
        assert "pull" in modes


def test_sssp_pull_graphs_and_options(gpu):
    # Several passes and both frontier modes, also with the captured graphs
    # off; the result never depends on the mode.
    h = cgp_amd.HostCSR.rmat(12, 16, seed=32, wmin=1, wmax=64)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(150, seed=32)]
    kinds = [SSSP] * 100 + [BFS] * 50
    ref = oracle_rows(og, kinds, src)
    for kw in ({"max_words": 1}, {"use_hip_graph": 0}, {"frontier_dense_div": 1000000}, {"block_size": 512}):
        en = cgp_amd.Engine(g, traversal_mode=cgp_amd.PULL, sssp_delta=0, **kw)
        out, st = en.run_batch(kinds, src)
        assert np.array_equal(out, ref), kw
        assert st["pull_iterations"] > 0, kw


def test_no_sssp_pull_without_symmetric_weights(gpu):
    # A symmetric structure whose reverse edges carry other weights: a gather
    # over a vertex's own edges would use the wrong weights, so SSSP stays
    # push (BFS may still pull) and the answers stay exact.
    h = cgp_amd.HostCSR.rmat(11, 16, seed=33, wmin=1, wmax=64)
    off, idx = h.offsets, h.indices
    rows = np.repeat(np.arange(h.V), np.diff(off))
    w = (1 + (rows * 7 + idx * 13) % 50).astype(np.uint32)  # depends on the direction
    g = cgp_amd.Graph.from_csr(off, idx, w)
    og = O.G(off, idx, w)
    src = [int(x) for x in h.sample_sources(20, seed=33)]
    en = cgp_amd.Engine(g, traversal_mode=cgp_amd.PULL, sssp_delta=0)
    out, st = en.run_batch([SSSP] * 20, src)
    assert np.array_equal(out, oracle_rows(og, [SSSP] * 20, src))
    assert st["pull_iterations"] == 0
    # The same graph from device arrays is recognised the same way.
    gd = cgp_amd.Graph.from_device(h.V, h.E, cgp_amd.DeviceBuffer.from_array(off.astype(np.int32)),
                                   cgp_amd.DeviceBuffer.from_array(idx.astype(np.int32)),
                                   cgp_amd.DeviceBuffer.from_array(w))
    out2, st2 = cgp_amd.Engine(gd, traversal_mode=cgp_amd.PULL, sssp_delta=0).run_batch([SSSP] * 20, src)
    assert np.array_equal(out2, out) and st2["pull_iterations"] == 0
    # With symmetric weights the device graph pulls.
    gs = cgp_amd.Graph.from_device(h.V, h.E, cgp_amd.DeviceBuffer.from_array(off.astype(np.int32)),
                                   cgp_amd.DeviceBuffer.from_array(idx.astype(np.int32)),
                                   cgp_amd.DeviceBuffer.from_array(np.asarray(h.weights, dtype=np.uint32)))
    out3, st3 = cgp_amd.Engine(gs, traversal_mode=cgp_amd.PULL, sssp_delta=0).run_batch([SSSP] * 20, src)
    assert np.array_equal(out3, oracle_rows(O.G.of(h), [SSSP] * 20, src)) and st3["pull_iterations"] > 0
