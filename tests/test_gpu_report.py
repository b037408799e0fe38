"""The run report (cgp_last_run_report_text): the fields of the reference
driver's JSON / CSV exports (cgp/examples/algorithms/bfs/cgp_bfs.cu:359-555),
filled from the device-written iteration records.  For BFS the per-level
frontier statistics do not depend on the implementation, so they are checked
exactly against the oracle's depths: frontier L holds the (vertex, query)
pairs at depth L."""
import csv
import io
import json
import sys

import numpy as np
import pytest

import cgp_amd
import oracle_lib as O
from cgp_amd import BFS, SSSP, WCC, INF

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("profile", [0, 3])
def test_bfs_report_levels_equal_the_oracle_depths(gpu, profile):
    h = cgp_amd.HostCSR.rmat(12, 16, seed=5)
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(40, seed=5)]
    en = cgp_amd.Engine(g, traversal_mode=cgp_amd.PUSH, profile=profile)
    out, st = en.run_batch([BFS] * 40, src)
    rep = json.loads(en.last_run_report_text("json", dataset="rmat12"))
    deg = np.diff(h.offsets).astype(np.int64)
    depths = np.stack([O.bfs(og, s) for s in src])
    assert np.array_equal(out, depths)
    finite = depths != INF
    maxd = int(depths[finite].max())
    say("report: %d BFS, max depth %d, %d iterations" % (len(src), maxd, rep["iterations"]))
    assert rep["primitive"] == "cgp_bfs" and rep["num_vertices"] == h.V and rep["num_edges"] == h.E
    assert rep["iterations"] == maxd + 1 == st["iterations"]
    fs, ufs, lec, vlec = [], [], [], []
    for L in range(maxd + 1):
        at = depths == L
        fs.append(int(at.sum()))
        anyv = at.any(axis=0)
        ufs.append(int(anyv.sum()))
        lec.append(int(deg[anyv].sum()))
        vlec.append(int((at * deg[None, :]).sum()))
    assert rep["frontier_sizes"] == fs + [0]
    assert rep["unique_frontier_sizes"] == ufs + [0]
    assert rep["level_edge_counts"] == lec
    assert rep["virtual_level_edge_counts"] == vlec
    if profile & 2:
        assert rep["actual_edge_counts_measured"] is True
        assert rep["actual_level_edge_counts"] == lec  # push scans every frontier vertex's edges
    assert rep["level_modes"] == ["push"] * (maxd + 1)
    assert [q["source"] for q in rep["query_results"]] == src
    for q, r in enumerate(rep["query_results"]):
        assert r["completion_level"] == int(depths[q][finite[q]].max()), q
    times = [r["completion_wall_time_ms"] for r in rep["query_results"]]
    assert all(0 < t <= rep["wall_time_ms"] for t in times)
    assert len(rep["level_profiles"]) == maxd + 1
    assert [p["frontier_size"] for p in rep["level_profiles"]] == fs
    # CSV: three rows per level (whole iteration, advance, compaction).
    rows = list(csv.DictReader(io.StringIO(en.last_run_report_text("csv", dataset="rmat12"))))
    assert len(rows) == 3 * (maxd + 1)
    whole = [r for r in rows if r["range_name"].endswith(":iter:%s" % r["iteration"])]
    assert [int(r["input_frontier"]) for r in whole] == fs
    assert [int(r["active_vertices"]) for r in whole] == ufs
    assert [int(r["output_frontier"]) for r in whole] == fs[1:] + [0]


def test_mixed_report_lists_every_query(gpu):
    h = cgp_amd.HostCSR.rmat(11, 16, seed=6, wmin=1, wmax=40)
    g = cgp_amd.Graph.from_host(h)
    src = [int(x) for x in h.sample_sources(20, seed=6)]
    kinds = [SSSP] * 8 + [BFS] * 12 + [WCC] * 2
    en = cgp_amd.Engine(g)
    en.run_batch(kinds, src + [0, 0])
    rep = json.loads(en.last_run_report_text("json"))
    assert rep["primitive"] == "cgp_mixed"
    assert [r["kind"] for r in rep["query_results"]] == ["sssp"] * 8 + ["bfs"] * 12 + ["wcc"] * 2
    assert all(r["pass"] == 0 for r in rep["query_results"][:20])
    assert all(r["pass"] == -1 for r in rep["query_results"][20:])
    assert len(rep["frontier_sizes"]) == rep["iterations"] + 1 and rep["frontier_sizes"][-1] == 0
    assert rep["frontier_sizes"][0] == 20  # every source seeds one (vertex, query) pair
    recs = en.last_run_report()
    assert len(recs) == rep["iterations"]
    assert abs(sum(r["advance_ms"] + r["filter_ms"] for r in recs) - rep["gpu_time_ms"]) < 1e-3 * len(recs)
    # A query is in the frontier of exactly the levels up to its completion.
    for r in rep["query_results"][:20]:
        assert 0 <= r["completion_level"] < rep["iterations"]
    with pytest.raises(cgp_amd.CgpError):
        cgp_amd.Engine(g).last_run_report_text("json")  # no batch ran on that engine
