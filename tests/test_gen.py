"""Generators, loaders and sampling (host code behind the C-ABI; no GPU)."""
import os
import struct

import numpy as np
import pytest

import cgp_amd
from cgp_amd import CgpError, HostCSR


def edges_of(h):
    src = np.repeat(np.arange(h.V, dtype=np.int64), np.diff(h.offsets.astype(np.int64)))
    return src, h.indices.astype(np.int64)


def assert_simple_symmetric(h):
    src, dst = edges_of(h)
    assert np.all(src != dst), "self-loop"
    for v in range(min(h.V, 2000)):
        row = h.indices[h.offsets[v]:h.offsets[v + 1]]
        assert np.all(np.diff(row) > 0), "row not strictly sorted (dedupe)"
    fwd = src * h.V + dst
    rev = dst * h.V + src
    assert np.array_equal(np.sort(fwd), np.sort(rev)), "not symmetric"


def test_rmat_structure_and_determinism():
    a = HostCSR.rmat(12, 16, seed=5, wmin=1, wmax=64)
    assert a.V == 4096 and a.offsets[0] == 0 and a.offsets[-1] == a.E
    assert_simple_symmetric(a)
    b = HostCSR.rmat(12, 16, seed=5, wmin=1, wmax=64)
    assert np.array_equal(a.offsets, b.offsets) and np.array_equal(a.indices, b.indices)
    assert np.array_equal(a.weights, b.weights)
    c = HostCSR.rmat(12, 16, seed=6)
    assert not np.array_equal(a.offsets, c.offsets)
    assert c.weights is None


def test_rmat_independent_of_thread_count():
    old = os.environ.get("CGP_THREADS")
    try:
        os.environ["CGP_THREADS"] = "1"
        a = HostCSR.rmat(11, 8, seed=9, wmin=1, wmax=9)
        os.environ["CGP_THREADS"] = "7"
        b = HostCSR.rmat(11, 8, seed=9, wmin=1, wmax=9)
    finally:
        if old is None:
            os.environ.pop("CGP_THREADS", None)
        else:
            os.environ["CGP_THREADS"] = old
    assert np.array_equal(a.indices, b.indices) and np.array_equal(a.weights, b.weights)


def test_rmat_scale16_edge_count_and_skew():
    # RMAT scale 16, edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 1:
    # about 1.9M directed edges after dropping self-loops and duplicates.
    h = HostCSR.rmat(16, 16, seed=1)
    assert 1.7e6 < h.E < 2.1e6, h.E
    deg = h.degrees()
    assert deg.max() > 50 * deg.mean()  # power-law head
    assert (deg == 0).sum() > 0


def test_weights_symmetric_and_in_range():
    h = HostCSR.rmat(10, 16, seed=2, wmin=3, wmax=17)
    src, dst = edges_of(h)
    w = h.weights.astype(np.int64)
    assert w.min() >= 3 and w.max() <= 17
    lut = dict(zip(zip(src.tolist(), dst.tolist()), w.tolist()))
    for (u, v), x in list(lut.items())[:5000]:
        assert lut[(v, u)] == x


def test_grid_structure():
    r, c = 9, 7
    h = HostCSR.grid(r, c, p_extra=0.0, seed=5, wmin=1, wmax=1000)
    assert h.V == r * c
    assert h.E == 2 * (2 * r * c - r - c)
    deg = h.degrees().reshape(r, c)
    assert deg[0, 0] == 2 and deg[0, 1] == 3 and deg[1, 1] == 4
    assert_simple_symmetric(h)
    row0 = h.indices[h.offsets[c + 1]:h.offsets[c + 2]].tolist()
    assert row0 == [1, c, c + 2, 2 * c + 1]
    x = HostCSR.grid(64, 64, p_extra=0.5, seed=5)
    base = 2 * (2 * 64 * 64 - 128)
    assert base < x.E < base + 2 * 64 * 64
    assert_simple_symmetric(x)


def test_grid_4096_extra_edge_rate():
    # The 4096 x 4096 grid workload adds one random undirected edge per vertex
    # with probability 0.01; check the rate on a 1024 x 1024 grid.
    h = HostCSR.grid(1024, 1024, p_extra=0.01, seed=5)
    base = 2 * (2 * 1024 * 1024 - 2048)
    extra = (h.E - base) / 2
    assert 0.009 * 1024 * 1024 < extra < 0.011 * 1024 * 1024


def test_sample_sources():
    h = HostCSR.rmat(12, 4, seed=1)
    s = h.sample_sources(300, seed=11)
    assert len(set(s.tolist())) == 300
    assert np.all(h.degrees()[s] >= 1)
    assert np.array_equal(s, h.sample_sources(300, seed=11))
    with pytest.raises(CgpError) as e:
        h.sample_sources(h.V + 1, seed=1)
    assert e.value.status == cgp_amd.CGP_EINVAL


def test_poisson_arrivals():
    t = cgp_amd.poisson_arrivals(20000, 250.0, seed=4)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 250.0) < 0.03 / 250.0
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05  # exponential: CV = 1


def test_coo_keeps_duplicates_and_sorts():
    h = HostCSR.from_coo(4, [2, 0, 0, 0, 3], [1, 2, 1, 1, 3], [5, 4, 9, 2, 1])
    assert h.offsets.tolist() == [0, 3, 3, 4, 5]
    assert h.indices.tolist() == [1, 1, 2, 1, 3]
    assert h.weights.tolist() == [2, 9, 4, 5, 1]


def test_matrix_market_symmetric_expansion(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer symmetric\n% c\n3 3 3\n1 2 7\n3 1 2\n2 2 4\n")
    h = HostCSR.load_matrix_market(p)
    assert h.V == 3 and h.E == 5
    assert h.offsets.tolist() == [0, 2, 4, 5]
    assert h.indices.tolist() == [1, 2, 0, 1, 0]
    assert h.weights.tolist() == [7, 2, 7, 4, 2]


def test_matrix_market_errors(tmp_path):
    p = tmp_path / "t.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 4\n1 2\n")
    with pytest.raises(CgpError) as e:
        HostCSR.load_matrix_market(p)
    assert e.value.status == cgp_amd.CGP_EIO and "truncated" in e.value.message
    with pytest.raises(CgpError):
        HostCSR.load_matrix_market(tmp_path / "missing.mtx")


def write_gr(path, off, idx, w=None):
    V, E = len(off) - 1, len(idx)
    with open(path, "wb") as f:
        f.write(struct.pack("<4Q", 1, 4 if w is not None else 0, V, E))
        f.write(np.asarray(off[1:], dtype=np.int64).tobytes())
        f.write(np.asarray(idx, dtype=np.uint32).tobytes())
        if w is not None:
            if E % 2:
                f.write(b"\0" * 4)
            f.write(np.asarray(w, dtype=np.uint32).tobytes())


def test_gr_round_trip(tmp_path):
    h = HostCSR.rmat(9, 8, seed=3, wmin=1, wmax=50)
    p = tmp_path / "g.gr"
    write_gr(p, h.offsets, h.indices)
    g = HostCSR.load_gr(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.indices, h.indices)
    assert g.weights is None
    write_gr(p, h.offsets, h.indices, h.weights)
    g = HostCSR.load_gr(p)
    assert np.array_equal(g.weights, h.weights)
    with open(p, "r+b") as f:
        f.write(struct.pack("<Q", 2))
    with pytest.raises(CgpError) as e:
        HostCSR.load_gr(p)
    assert "version" in e.value.message


def test_csr_directory(tmp_path):
    h = HostCSR.rmat(9, 8, seed=4, wmin=1, wmax=9)
    d = tmp_path / "csr"
    d.mkdir()
    h.offsets.astype(np.int32).tofile(d / "csr_vlist.bin")
    h.indices.astype(np.int32).tofile(d / "csr_elist.bin")
    g = HostCSR.load_csr_directory(d)
    assert np.array_equal(g.indices, h.indices) and g.weights is None
    h.weights.astype(np.int32).tofile(d / "csr_weightlist.bin")
    g = HostCSR.load_csr_directory(d)
    assert np.array_equal(g.weights, h.weights)
    (d / "csr_weightlist.bin").write_bytes(b"\0\0\0\0")
    with pytest.raises(CgpError):
        HostCSR.load_csr_directory(d)


def test_generator_argument_errors():
    for args in [(0, 16), (31, 16), (10, 0)]:
        with pytest.raises(CgpError) as e:
            HostCSR.rmat(*args)
        assert e.value.status == cgp_amd.CGP_EINVAL
    with pytest.raises(CgpError):
        HostCSR.rmat(8, 4, a=0.9, b=0.2, c=0.1)
    with pytest.raises(CgpError):
        HostCSR.rmat(8, 4, wmin=5, wmax=2)
    with pytest.raises(CgpError):
        HostCSR.grid(0, 5)


def test_matrix_market_weights_are_exact_or_rejected(tmp_path):
    # Distances are exact uint32 sums: a fractional, negative or too large
    # value is an error unless the caller asks for rounding or unit weights.
    p = tmp_path / "frac.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 2 1.5\n2 3 2\n")
    with pytest.raises(cgp_amd.CgpError) as ei:
        HostCSR.load_matrix_market(p)
    assert ei.value.status == cgp_amd.CGP_EIO and "not an integer" in str(ei.value)
    g = HostCSR.load_matrix_market(p, round_weights=True)
    assert g.weights.tolist() == [2, 2]
    g = HostCSR.load_matrix_market(p, unit_weights=True)
    assert g.weights is None and g.indices.tolist() == [1, 2]
    n = tmp_path / "neg.mtx"
    n.write_text("%%MatrixMarket matrix coordinate integer general\n2 2 1\n1 2 -3\n")
    with pytest.raises(cgp_amd.CgpError):
        HostCSR.load_matrix_market(n)
    assert HostCSR.load_matrix_market(n, round_weights=True).weights.tolist() == [0]


def test_gr_unit_weight_mode(tmp_path):
    h = HostCSR.rmat(8, 8, seed=9, wmin=1, wmax=50)
    p = tmp_path / "w.gr"
    write_gr(p, h.offsets, h.indices, h.weights)
    assert np.array_equal(HostCSR.load_gr(p).weights, h.weights)
    u = HostCSR.load_gr(p, unit_weights=True)
    assert u.weights is None and np.array_equal(u.indices, h.indices)
