"""Real-graph loaders through the engine on the GPU: a Galois .gr file, a
binary CSR directory and a Matrix Market file written from a seeded graph
are loaded (gen.cpp), uploaded and run as mixed batches, and every row is
compared with the oracle on the loaded arrays.  The .gr unit-weight mode
matches the reference's SSSP driver, which ignores stored weights
(cgp/examples/algorithms/sssp/cgp_sssp.cu:53-56): its SSSP rows equal BFS
depths."""
import struct
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


def write_gr(path, off, idx, w=None):
    # Galois .gr v1: version, sizeof(edge data), V, E (u64), row ends (u64[V]),
    # destinations (u32[E]), padding to 8 bytes, edge data (u32[E]).
    V, E = len(off) - 1, len(idx)
    with open(path, "wb") as f:
        f.write(struct.pack("<4Q", 1, 4 if w is not None else 0, V, E))
        f.write(np.asarray(off[1:], dtype=np.uint64).tobytes())
        f.write(np.asarray(idx, dtype=np.uint32).tobytes())
        if w is not None:
            if E % 2:
                f.write(b"\0" * 4)
            f.write(np.asarray(w, dtype=np.uint32).tobytes())


def run_and_check(h, label, seed, unit=False):
    g = cgp_amd.Graph.from_host(h)
    og = O.G.of(h)
    src = [int(x) for x in h.sample_sources(24, seed=seed)]
    kinds = [BFS] * 10 + [SSSP] * 14 + [WCC]
    sources = src + [0]
    out, st = cgp_amd.Engine(g).run_batch(kinds, sources)
    say("%s: V=%d E=%d, %d queries in %d iterations" % (label, h.V, h.E, len(kinds), st["iterations"]))
    for i, (k, s) in enumerate(zip(kinds, sources)):
        assert np.array_equal(out[i], O.expected(og, k, s)), "%s: row %d (kind %d, source %d)" % (label, i, k, s)
    if unit:
        for i in range(10, 24):
            assert np.array_equal(out[i], O.bfs(og, sources[i])), "%s: unit-weight SSSP row %d is not BFS" % (label, i)
    return out


def test_gr_file_through_the_engine(gpu, tmp_path):
    h = cgp_amd.HostCSR.rmat(13, 16, seed=21, wmin=1, wmax=99)
    p = tmp_path / "rmat13.gr"
    write_gr(p, h.offsets, h.indices, h.weights)
    g1 = cgp_amd.HostCSR.load_gr(p)
    assert np.array_equal(g1.weights, h.weights)
    run_and_check(g1, ".gr weighted", 21)
    g2 = cgp_amd.HostCSR.load_gr(p, unit_weights=True)
    assert g2.weights is None
    run_and_check(g2, ".gr unit weights", 21, unit=True)


def test_csr_directory_through_the_engine(gpu, tmp_path):
    h = cgp_amd.HostCSR.rmat(13, 8, seed=22, wmin=1, wmax=40)
    d = tmp_path / "csr"
    d.mkdir()
    h.offsets.astype(np.int32).tofile(d / "csr_vlist.bin")
    h.indices.astype(np.int32).tofile(d / "csr_elist.bin")
    h.weights.astype(np.int32).tofile(d / "csr_weightlist.bin")
    g = cgp_amd.HostCSR.load_csr_directory(d)
    assert np.array_equal(g.weights, h.weights)
    run_and_check(g, "csr directory", 22)
    run_and_check(cgp_amd.HostCSR.load_csr_directory(d, unit_weights=True), "csr directory unit", 22, unit=True)


def test_matrix_market_through_the_engine(gpu, tmp_path):
    # A symmetric Matrix Market file (lower triangle stored, 1-based) from a
    # seeded graph: the loader's symmetric expansion rebuilds the CSR.
    h = cgp_amd.HostCSR.rmat(12, 8, seed=23, wmin=1, wmax=30)
    off, idx, w = h.offsets, h.indices, h.weights
    rows = np.repeat(np.arange(h.V), np.diff(off))
    keep = idx < rows
    p = tmp_path / "g.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer symmetric\n% seeded RMAT-12\n")
        f.write("%d %d %d\n" % (h.V, h.V, int(keep.sum())))
        for a, b, x in zip(rows[keep], idx[keep], w[keep]):
            f.write("%d %d %d\n" % (a + 1, b + 1, x))
    g = cgp_amd.HostCSR.load_matrix_market(p)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.indices, idx) and np.array_equal(g.weights, w)
    run_and_check(g, "matrix market", 23)
