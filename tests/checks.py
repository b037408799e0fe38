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
    """labels are the min-id labels of the weakly connected components iff:
    (a) labels are equal across every edge, so each label class is a union of
    components; (b) label[v] <= v and label[label[v]] == label[v], so a class's
    label is a member no larger than any other member, its minimum; (c) there
    are as many label classes as components (scipy's connected_components, an
    independent count), so no class holds two components.  Without (c) a
    labelling that put components {0} and {1, 2} under label 0 would pass."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    lab = np.asarray(labels, dtype=np.int64)
    V = len(off) - 1
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off.astype(np.int64)))
    if np.any(lab[src] != lab[idx]):
        return "edge crosses labels"
    if np.any(lab > np.arange(V)):
        return "label above vertex id"
    if np.any(lab[lab] != lab):
        return "label not a fixed point"
    a = csr_matrix((np.ones(len(idx), dtype=np.int8), idx.astype(np.int64), off.astype(np.int64)), shape=(V, V))
    ncomp, _ = connected_components(a, directed=True, connection="weak")
    if np.count_nonzero(lab == np.arange(V)) != ncomp:
        return "label class not connected"
    return None
