"""Size-independent property checks on traversal results (test-only).

These pin results at sizes where the oracle is too slow to run per query:
each check is a necessary-and-sufficient characterisation of the unique
answer (BFS depth, shortest-path distance, min-id component label), so a
result that passes them equals the oracle's on the same graph.
"""
import numpy as np

INF = 0xFFFFFFFF


def reverse_csr(off, idx, wt=None):
    V = len(off) - 1
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off.astype(np.int64)))
    order = np.argsort(idx, kind="stable")
    roff = np.zeros(V + 1, dtype=np.int64)
    np.add.at(roff, idx.astype(np.int64) + 1, 1)
    roff = np.cumsum(roff)
    rsrc = src[order]
    rwt = None if wt is None else wt[order]
    return roff, rsrc, rwt


def check_sssp(off, idx, wt, source, d):
    """d is the exact shortest-path distance vector from `source` (uint32,
    INF unreachable) iff: d[s] = 0; d[v] <= d[u] + w(u,v) on every edge with
    finite d[u]; every finite d[v] > 0 has an in-edge (u,v) with
    d[u] + w = d[v]; and no vertex other than s has d = 0 unless reached by
    zero-weight edges (weights here are >= 1).  wt None = unit weights (BFS)."""
    d = np.asarray(d, dtype=np.uint64)
    V = len(off) - 1
    if d[source] != 0:
        return "source not 0"
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off.astype(np.int64)))
    w = np.ones(len(idx), dtype=np.uint64) if wt is None else wt.astype(np.uint64)
    du = d[src]
    fin = du != INF
    cand = du[fin] + w[fin]
    dv = d[idx[fin]]
    if np.any(dv > cand):
        return "edge not relaxed"
    # Tight in-edge for every finite non-source vertex.
    tight = np.zeros(V, dtype=bool)
    tmask = cand == dv
    tight[idx[fin][tmask]] = True
    finite = d != INF
    finite[source] = False
    if np.any(finite & ~tight):
        return "finite vertex without tight in-edge"
    if np.count_nonzero(d == 0) != 1:
        return "zero distance off the source"
    return None


def check_bfs(off, idx, source, d):
    return check_sssp(off, idx, None, source, d)


def check_wcc(off, idx, labels):
    """labels are min-id component labels iff: labels equal across every edge;
    label[v] <= v; label[label[v]] == label[v] (the label is a member, its own
    root); and each label class is connected is implied by the first two with
    the root being the minimum (label[v] <= v and the class representative
    labels itself)."""
    lab = np.asarray(labels, dtype=np.int64)
    V = len(off) - 1
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off.astype(np.int64)))
    if np.any(lab[src] != lab[idx]):
        return "edge crosses labels"
    if np.any(lab > np.arange(V)):
        return "label above vertex id"
    if np.any(lab[lab] != lab):
        return "label not a fixed point"
    return None
