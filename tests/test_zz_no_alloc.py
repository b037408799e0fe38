"""The allocation-free entry point cgp_run_batch_static: every per-run device
buffer comes from the library's fixed device array.  Its results must equal
cgp_run_batch's on the same engine, bit for bit, call after call (the array
keeps the previous call's contents), and batches that do not fit the array
are rejected before any launch."""
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


def both(en, V, kinds, sources):
    n = len(kinds)
    a = cgp_amd.DeviceBuffer(n * V * 4)
    b = cgp_amd.DeviceBuffer(n * V * 4)
    en.run_batch(kinds, sources, out=a)
    st = en.run_batch_static(kinds, sources, out=b)
    return a.to_host(shape=(n, V)), b.to_host(shape=(n, V)), st


def test_static_equals_run_batch_twice_then_smaller(gpu):
    h = cgp_amd.HostCSR.rmat(12, 16, seed=7, wmin=1, wmax=64)
    g = cgp_amd.Graph.from_host(h)
    V = g.V
    en = cgp_amd.Engine(g)
    src = [int(x) for x in h.sample_sources(120, seed=7)]
    # First call: two mask words (100 SSSP/BFS slots) plus WCC.
    kinds = [BFS] * 60 + [SSSP] * 40 + [WCC] * 2
    sources = src[:100] + [0, 0]
    say("static: RMAT-12 V=%d E=%d, call 1: %d queries" % (V, h.E, len(kinds)))
    ref, got, st = both(en, V, kinds, sources)
    assert st["iterations"] > 0
    assert np.array_equal(ref, got), "call 1: static run differs from cgp_run_batch"
    og = O.G.of(h)
    for i in (0, 59, 60, 99, 100):
        assert np.array_equal(got[i], O.expected(og, kinds[i], sources[i])), "call 1: row %d differs from the oracle" % i
    # Second call, smaller: one word of 8-lane rows; the array still holds
    # the first call's state.
    kinds2 = [SSSP] * 5 + [BFS] * 3
    sources2 = src[100:108]
    say("static: call 2: %d queries" % len(kinds2))
    ref2, got2, _ = both(en, V, kinds2, sources2)
    assert np.array_equal(ref2, got2), "call 2: static run differs from cgp_run_batch"
    # A smaller graph on another engine shares the same array.
    h3 = cgp_amd.HostCSR.rmat(10, 8, seed=8, wmin=1, wmax=16)
    g3 = cgp_amd.Graph.from_host(h3)
    en3 = cgp_amd.Engine(g3)
    src3 = [int(x) for x in h3.sample_sources(4, seed=8)]
    kinds3 = [BFS, SSSP, BFS, SSSP, WCC]
    say("static: call 3: RMAT-10, %d queries" % len(kinds3))
    ref3, got3, _ = both(en3, g3.V, kinds3, src3 + [0])
    assert np.array_equal(ref3, got3), "call 3: static run differs from cgp_run_batch"
    # The engine's own buffers are untouched: its regular path still agrees.
    again = cgp_amd.DeviceBuffer(len(kinds) * V * 4)
    en.run_batch(kinds, sources, out=again)
    assert np.array_equal(again.to_host(shape=(len(kinds), V)), ref)


def test_static_rejects_batches_beyond_the_array(gpu):
    # 512 SSSP slots on 65,536 vertices need 8 words of 64-lane rows:
    # 128 MiB of values alone, more than the 64 MiB array.
    h = cgp_amd.HostCSR.rmat(16, 16, seed=1, wmin=1, wmax=8)
    g = cgp_amd.Graph.from_host(h)
    en = cgp_amd.Engine(g)
    src = [int(x) for x in h.sample_sources(512, seed=1)]
    out = cgp_amd.DeviceBuffer(512 * g.V * 4)
    with pytest.raises(cgp_amd.CgpError) as ei:
        en.run_batch_static([SSSP] * 512, src, out)
    assert ei.value.status == cgp_amd.CGP_ECAPACITY
    assert "bytes of device scratch" in str(ei.value)
    with pytest.raises(cgp_amd.CgpError) as ei:
        en.run_batch_static([BFS] * 1025, [src[0]] * 1025, cgp_amd.DeviceBuffer(1025 * g.V * 4))
    assert ei.value.status == cgp_amd.CGP_ECAPACITY
    # A batch that fits still runs on the same engine.
    st = en.run_batch_static([BFS] * 4, src[:4], out)
    assert st["iterations"] > 0
