"""Writes golden.json: known-answer vectors the reference's own tests hold.

Each entry transcribes the graph and the expected values of one reference
test (paths relative to the reference root); INF is 0xFFFFFFFF.  The two
Matrix Market fixtures (unweighted.mm, weighted.mm) are the data files of
puercgp/examples/matrices/, used by validate_bfs.cu / validate_sssp.cu with
their default sources 0,1,2,3.

Run: python tests/golden/make_golden.py  (rewrites tests/golden/golden.json)
"""
import json
import os

INF = 0xFFFFFFFF

# Toy graph T1: 0->1 (w2), 0->2 (w5), 1->3 (w1).
T1 = {"V": 4, "offsets": [0, 2, 3, 3, 3], "indices": [1, 2, 3], "weights": [2, 5, 1]}
# Toy graph T2: T1 with w(1->3) = 3.
T2 = {"V": 4, "offsets": [0, 2, 3, 3, 3], "indices": [1, 2, 3], "weights": [2, 5, 3]}

CASES = [
    {"name": "validate_hybrid_bfs", "cite": "puercgp/examples/validate_hybrid.cu:41-44,91-94",
     "graph": "T1", "kind": "bfs", "source": 0, "expect": [0, 1, 1, 2]},
    {"name": "validate_hybrid_sssp", "cite": "puercgp/examples/validate_hybrid.cu:41-44,97-100",
     "graph": "T1", "kind": "sssp", "source": 0, "expect": [0, 2, 5, 3]},
    {"name": "validate_hybrid_wcc", "cite": "puercgp/examples/validate_hybrid.cu:103-106",
     "graph": "T1", "kind": "wcc", "source": 0, "expect": [0, 0, 0, 0]},
    {"name": "validate_wcc", "cite": "puercgp/examples/validate_wcc.cu:33-64",
     "graph": "T1", "kind": "wcc", "source": 0, "expect": [0, 0, 0, 0]},
    {"name": "smoke_hybrid_push_bfs_one_level", "cite": "puercgp/examples/smoke_hybrid_push.cu:151-155",
     "graph": "T1", "kind": "bfs", "source": 0, "max_iterations": 1, "expect": [0, 1, 1, INF]},
    {"name": "smoke_hybrid_push_sssp_one_level", "cite": "puercgp/examples/smoke_hybrid_push.cu:157-160",
     "graph": "T1", "kind": "sssp", "source": 0, "max_iterations": 1, "expect": [0, 2, 5, INF]},
    {"name": "validate_replenish_q0", "cite": "puercgp/examples/validate_replenish.cu:45-58,89-93",
     "graph": "T2", "kind": "bfs", "source": 0, "expect": [0, 1, 1, 2]},
    {"name": "validate_replenish_q1", "cite": "puercgp/examples/validate_replenish.cu:95-99",
     "graph": "T2", "kind": "sssp", "source": 0, "expect": [0, 2, 5, 5]},
    {"name": "validate_replenish_q2", "cite": "puercgp/examples/validate_replenish.cu:101-105",
     "graph": "T2", "kind": "bfs", "source": 1, "expect": [INF, 0, INF, 1]},
    {"name": "validate_replenish_q3", "cite": "puercgp/examples/validate_replenish.cu:107-111",
     "graph": "T2", "kind": "sssp", "source": 1, "expect": [INF, 0, INF, 3]},
]

# validate_bfs.cu / validate_sssp.cu defaults (sources 0..3) on the .mm fixtures.
MM_CASES = [
    {"file": "unweighted.mm", "kind": "bfs", "cite": "puercgp/examples/validate_bfs.cu:59-65 (closed form; the reference compares against cpu_bfs, :114-142)",
     "expect": {0: [0, 1, 1, 2], 1: [INF, 0, INF, 1], 2: [INF, 1, 0, 2], 3: [INF, INF, INF, 0]}},
    {"file": "weighted.mm", "kind": "sssp", "cite": "puercgp/examples/validate_sssp.cu:60-66 (closed form; the reference compares against cpu_sssp, :110-139)",
     "expect": {0: [0, 1, 2, 3], 1: [INF, 0, 1, 2], 2: [INF, INF, 0, 1], 3: [INF, INF, INF, 0]}},
]

# Primitive semantics of smoke_reduce_ops.cu:30-144 (value sequences and the
# improved flag after each update).
REDUCE_OPS = {
    "cite": "puercgp/examples/smoke_reduce_ops.cu:30-144",
    "first_write": [[INF, 5, True, INF], [5, 3, False, 5]],
    "min_reduce": [[10, 5, True, 10], [5, 8, False, 5], [5, 3, True, 5]],
}


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    doc = {"INF": INF, "graphs": {"T1": T1, "T2": T2}, "cases": CASES, "mm_cases": MM_CASES,
           "reduce_ops": REDUCE_OPS}
    with open(os.path.join(here, "golden.json"), "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    main()
