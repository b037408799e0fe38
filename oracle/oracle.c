/*
 * oracle.c — CPU restatement of the reference's own correctness oracles for
 * the concurrent multi-query traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or, in bench.py, as the timed CPU baseline).  The product library
 * (cgp_amd/libcgp_amd.so) never links or calls it.
 *
 * Every function cites the reference code it restates (paths relative to the
 * reference root).  Values are uint32 with ORC_INF = 0xFFFFFFFF for
 * "unreachable"; the reference uses INT_MAX for BFS and float +inf / FLT_MAX
 * for SSSP, which map one-to-one onto this convention for every graph whose
 * path sums stay below 2^24 (all configurations used here).
 *
 * Parity pinning (what checks each function, see tests/test_oracle.py):
 *   orc_bfs          golden vectors of validate_hybrid.cu / validate_bfs.cu
 *                    fixtures, the compiled reference (oracle/_ref), brute
 *                    force Bellman-Ford with unit weights, BFS invariants.
 *   orc_sssp         golden vectors (weighted.mm, T1, T2), the compiled
 *                    reference sssp_cpu / cpu_sssp, brute force Bellman-Ford,
 *                    unit weights == BFS, tight-edge invariants.
 *   orc_wcc_lp       compiled reference cpu_wcc on the same inputs.
 *   orc_wcc_uf       orc_wcc_lp on symmetric graphs, component invariants,
 *                    golden T1 vector.
 *   orc_bellman_ford closed-form small graphs (paths, stars, cycles).
 *   orc_reached_edges  closed forms (full graph reachability -> E).
 *   orc_check_sssp / orc_check_wcc / orc_check_rows
 *                    accept every oracle answer and reject planted mistakes
 *                    (a depth or distance off by one either way, a reachable
 *                    vertex left unreached, two components merged under one
 *                    label, a label that is not the component minimum).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_INF 0xFFFFFFFFu

/* Queue BFS.  Restates cpu_bfs (puercgp/examples/mtx_loader.hxx:269-290):
 * FIFO queue, a vertex is labelled the first time it is seen with
 * depth(parent)+1.  Equivalent to bfs_cpu::run's unit-weight Dijkstra
 * (cgp/examples/algorithms/bfs/bfs_cpu.hxx:20-68). */
void orc_bfs(int32_t V, const int32_t* off, const int32_t* idx, int32_t src, uint32_t* out) {
  for (int32_t v = 0; v < V; ++v) out[v] = ORC_INF;
  if (src < 0 || src >= V) return;
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1));
  int64_t head = 0, tail = 0;
  out[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    int32_t u = queue[head++];
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      int32_t v = idx[e];
      if (out[v] == ORC_INF) {
        out[v] = out[u] + 1;
        queue[tail++] = v;
      }
    }
  }
  free(queue);
}

/* ---- binary min-heap of (dist, vertex), lazy deletion, as std::priority_queue
 * with std::greater in cpu_sssp (mtx_loader.hxx:292-319). */
typedef struct {
  uint64_t* a; /* (dist << 32) | vertex: orders by dist, ties by vertex */
  int64_t n, cap;
} orc_heap;

static void heap_push(orc_heap* h, uint64_t x) {
  if (h->n == h->cap) {
    h->cap = h->cap ? h->cap * 2 : 1024;
    h->a = (uint64_t*)realloc(h->a, sizeof(uint64_t) * (size_t)h->cap);
  }
  int64_t i = h->n++;
  while (i > 0) {
    int64_t p = (i - 1) >> 1;
    if (h->a[p] <= x) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = x;
}

static uint64_t heap_pop(orc_heap* h) {
  uint64_t top = h->a[0];
  uint64_t x = h->a[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1;
    if (l >= h->n) break;
    int64_t r = l + 1;
    int64_t m = (r < h->n && h->a[r] < h->a[l]) ? r : l;
    if (h->a[m] >= x) break;
    h->a[i] = h->a[m];
    i = m;
  }
  if (h->n > 0) h->a[i] = x;
  return top;
}

/* Dijkstra on uint32 weights.  Restates cpu_sssp (puercgp/examples/mtx_loader.hxx:292-319)
 * and sssp_cpu::run (cgp/examples/algorithms/sssp/sssp_cpu.hxx:22-72): pop the
 * smallest tentative distance, skip stale entries, relax every out-edge with
 * d(u)+w(e) and push improvements.  wt == NULL means unit weights (the
 * reference's weight default, puercgp/include/puercgp/backend/csr_graph.hxx:37-39).
 * Path sums are computed in 64 bits and a sum >= 2^32-1 is treated as
 * unreachable, so the result stays exact for every representable distance. */
void orc_sssp(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt, int32_t src,
              uint32_t* out) {
  for (int32_t v = 0; v < V; ++v) out[v] = ORC_INF;
  if (src < 0 || src >= V) return;
  orc_heap h = {0, 0, 0};
  out[src] = 0;
  heap_push(&h, (uint64_t)(uint32_t)src);
  while (h.n > 0) {
    uint64_t top = heap_pop(&h);
    uint32_t d = (uint32_t)(top >> 32);
    int32_t u = (int32_t)(uint32_t)(top & 0xFFFFFFFFu);
    if (d != out[u]) continue;
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      int32_t v = idx[e];
      uint64_t cand = (uint64_t)d + (wt ? wt[e] : 1u);
      if (cand < (uint64_t)out[v]) {
        out[v] = (uint32_t)cand;
        heap_push(&h, (cand << 32) | (uint32_t)v);
      }
    }
  }
  free(h.a);
}

/* Label propagation to fixpoint.  Restates cpu_wcc (puercgp/examples/mtx_loader.hxx:328-356):
 * label[v] = v; sweep v = 0..V-1 replacing label[v] by the minimum over its
 * out-row labels, in place (updates visible to later vertices of the same
 * sweep), until a sweep changes nothing or max_iterations sweeps ran.
 * Returns the number of sweeps.  On a symmetric graph the fixpoint is the
 * minimum vertex id of each connected component. */
int32_t orc_wcc_lp(int32_t V, const int32_t* off, const int32_t* idx, uint32_t* out,
                   int32_t max_iterations) {
  for (int32_t v = 0; v < V; ++v) out[v] = (uint32_t)v;
  int32_t it = 0;
  for (; it < max_iterations; ++it) {
    int changed = 0;
    for (int32_t v = 0; v < V; ++v) {
      uint32_t best = out[v];
      for (int32_t e = off[v]; e < off[v + 1]; ++e) {
        uint32_t nb = out[idx[e]];
        if (nb < best) best = nb;
      }
      if (best < out[v]) {
        out[v] = best;
        changed = 1;
      }
    }
    if (!changed) break;
  }
  return it;
}

static uint32_t uf_find(uint32_t* p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

/* Weakly connected components, labelled with the minimum vertex id of the
 * component: union-find over every edge taken as undirected, linking the
 * larger root under the smaller.  This is the closed form of the fixpoint of
 * the reference's label propagation (algorithms/wcc.hxx:20-42 with
 * mtx_loader.hxx:328-356) on symmetric graphs, and of its push form on the
 * directed toy graph of validate_hybrid.cu:40-106 (all labels 0). */
void orc_wcc_uf(int32_t V, const int32_t* off, const int32_t* idx, uint32_t* out) {
  for (int32_t v = 0; v < V; ++v) out[v] = (uint32_t)v;
  for (int32_t u = 0; u < V; ++u) {
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t ru = uf_find(out, (uint32_t)u), rv = uf_find(out, (uint32_t)idx[e]);
      if (ru == rv) continue;
      if (ru < rv) out[rv] = ru;
      else out[ru] = rv;
    }
  }
  for (int32_t v = 0; v < V; ++v) out[v] = uf_find(out, (uint32_t)v);
}

/* Brute-force Bellman-Ford: V-1 full Jacobi rounds over every edge (stops
 * early when a round changes nothing).  Independent third opinion for small
 * graphs; wt == NULL means unit weights. */
void orc_bellman_ford(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                      int32_t src, uint32_t* out) {
  for (int32_t v = 0; v < V; ++v) out[v] = ORC_INF;
  if (src < 0 || src >= V) return;
  out[src] = 0;
  uint32_t* prev = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(V > 0 ? V : 1));
  for (int32_t round = 0; round < V; ++round) {
    memcpy(prev, out, sizeof(uint32_t) * (size_t)V);
    int changed = 0;
    for (int32_t u = 0; u < V; ++u) {
      if (prev[u] == ORC_INF) continue;
      for (int32_t e = off[u]; e < off[u + 1]; ++e) {
        uint64_t cand = (uint64_t)prev[u] + (wt ? wt[e] : 1u);
        if (cand < (uint64_t)out[idx[e]]) {
          out[idx[e]] = (uint32_t)cand;
          changed = 1;
        }
      }
    }
    if (!changed) break;
  }
  free(prev);
}

/* m_q: sum of out-degrees of every vertex with a finite value (the GTEPS
 * numerator: directed CSR edges a work-efficient traversal scans). */
uint64_t orc_reached_edges(int32_t V, const int32_t* off, const uint32_t* vals) {
  uint64_t m = 0;
  for (int32_t v = 0; v < V; ++v)
    if (vals[v] != ORC_INF) m += (uint64_t)(off[v + 1] - off[v]);
  return m;
}

/* ------------------------------------------------------------------ CPU baseline runner
 * One query per thread (std::thread-per-query in the plan; pthreads here),
 * each thread taking the next query index from a shared counter. */
typedef struct {
  int32_t V;
  const int32_t *off, *idx;
  const uint32_t* wt;
  const int32_t *kinds, *srcs;
  int32_t n;
  uint32_t* out;     /* n*V or NULL */
  uint64_t* m_out;   /* n or NULL */
  volatile int32_t next;
  pthread_mutex_t mu;
} orc_pool;

static void* orc_worker(void* arg) {
  orc_pool* p = (orc_pool*)arg;
  uint32_t* scratch = p->out ? NULL : (uint32_t*)malloc(sizeof(uint32_t) * (size_t)p->V);
  for (;;) {
    pthread_mutex_lock(&p->mu);
    int32_t q = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (q >= p->n) break;
    uint32_t* dst = p->out ? p->out + (size_t)q * (size_t)p->V : scratch;
    if (p->kinds[q] == 0) orc_bfs(p->V, p->off, p->idx, p->srcs[q], dst);
    else if (p->kinds[q] == 1) orc_sssp(p->V, p->off, p->idx, p->wt, p->srcs[q], dst);
    else orc_wcc_lp(p->V, p->off, p->idx, dst, 1 << 30);
    if (p->m_out) {
      p->m_out[q] = p->kinds[q] == 2 ? (uint64_t)p->off[p->V] : orc_reached_edges(p->V, p->off, dst);
    }
  }
  free(scratch);
  return NULL;
}

/* Runs n queries (kind 0 BFS, 1 SSSP, 2 WCC via label propagation) on
 * `threads` threads; returns elapsed wall seconds (steady clock). */
double orc_run_queries(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                       const int32_t* kinds, const int32_t* srcs, int32_t n, int32_t threads,
                       uint32_t* out, uint64_t* m_out) {
  orc_pool pool;
  pool.V = V; pool.off = off; pool.idx = idx; pool.wt = wt;
  pool.kinds = kinds; pool.srcs = srcs; pool.n = n; pool.out = out; pool.m_out = m_out;
  pool.next = 0;
  pthread_mutex_init(&pool.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, orc_worker, &pool);
  for (int32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  pthread_mutex_destroy(&pool.mu);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------ result checkers
 * Necessary-and-sufficient characterisations of the unique answers, so a
 * result can be checked at sizes where running the oracle per query is too
 * slow.  They read only the graph and the candidate result. */

/* Error codes of the checkers (0 = the result is the exact answer). */
enum {
  ORC_OK = 0,
  ORC_ESOURCE = 1,   /* d[src] != 0, or src out of range with a finite value */
  ORC_ERELAX = 2,    /* an edge (u,v) with d[v] > d[u] + w: some vertex too far */
  ORC_ETIGHT = 3,    /* a finite d[v] (v != src) with no in-edge (u,v), d[u] + w == d[v] */
  ORC_EZERO = 4,     /* d == 0 away from the source */
  ORC_EEDGE = 5,     /* WCC: an edge joins two different labels */
  ORC_EROOT = 6,     /* WCC: label[v] > v, or label[label[v]] != label[v] */
  ORC_ESPLIT = 7,    /* WCC: a label class is not connected */
  ORC_EKIND = 8      /* unknown query kind */
};

/* d is the exact shortest-path distance vector from src (ORC_INF = unreachable)
 * on a graph whose weights are all >= 1 (wt == NULL: unit weights, i.e. BFS
 * depths) iff
 *   (a) d[src] == 0 and no other vertex has d == 0;
 *   (b) every edge (u,v) with finite d[u] has d[v] <= d[u] + w(u,v);
 *   (c) every finite d[v], v != src, has a tight in-edge: d[u] + w(u,v) == d[v].
 * (b) along a shortest path gives d <= dist; following tight in-edges back from
 * v strictly decreases d (w >= 1) and can only stop at the one zero, src, so
 * d[v] is the length of a real path and d >= dist; unreachable vertices can
 * have no such chain, so they stay ORC_INF.  `tight` is caller scratch of V
 * bytes. */
static int check_sssp_row(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                          int32_t src, const uint32_t* d, unsigned char* tight) {
  if (V == 0) return ORC_OK;
  if (src < 0 || src >= V || d[src] != 0) return ORC_ESOURCE;
  memset(tight, 0, (size_t)V);
  for (int32_t u = 0; u < V; ++u) {
    const uint32_t du = d[u];
    if (du == ORC_INF) continue;
    for (int32_t e = off[u]; e < off[u + 1]; ++e) {
      const int32_t v = idx[e];
      const uint64_t cand = (uint64_t)du + (wt ? wt[e] : 1u);
      const uint64_t dv = d[v];
      if (dv > cand) return ORC_ERELAX;
      if (dv == cand) tight[v] = 1;
    }
  }
  for (int32_t v = 0; v < V; ++v) {
    if (v == src || d[v] == ORC_INF) continue;
    if (d[v] == 0) return ORC_EZERO;
    if (!tight[v]) return ORC_ETIGHT;
  }
  return ORC_OK;
}

int orc_check_sssp(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt, int32_t src,
                   const uint32_t* d) {
  unsigned char* tight = (unsigned char*)malloc((size_t)(V > 0 ? V : 1));
  const int r = check_sssp_row(V, off, idx, wt, src, d, tight);
  free(tight);
  return r;
}

/* labels are the minimum-vertex-id labels of the weakly connected components
 * iff
 *   (a) label[u] == label[v] on every edge: each label class is a union of
 *       components;
 *   (b) label[v] <= v and label[label[v]] == label[v]: the label of a class is
 *       a member of it and no member is smaller, so it is the class minimum;
 *   (c) every class is connected: a traversal over the edges taken as
 *       undirected, started at each class's own label vertex, reaches every
 *       vertex exactly once in total (a class of two components would leave
 *       the second unreached).
 * Works on directed graphs too (weak components): (c) walks out- and in-edges. */
int orc_check_wcc(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* lab) {
  if (V == 0) return ORC_OK;
  const int64_t E = off[V];
  for (int32_t u = 0; u < V; ++u)
    for (int32_t e = off[u]; e < off[u + 1]; ++e)
      if (lab[u] != lab[idx[e]]) return ORC_EEDGE;
  for (int32_t v = 0; v < V; ++v)
    if (lab[v] > (uint32_t)v || lab[lab[v]] != lab[v]) return ORC_EROOT;
  /* In-edges (reverse CSR by counting sort) for the undirected walk. */
  int64_t* roff = (int64_t*)calloc((size_t)V + 1, sizeof(int64_t));
  int32_t* rsrc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
  int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
  unsigned char* seen = (unsigned char*)calloc((size_t)V, 1);
  for (int64_t e = 0; e < E; ++e) ++roff[idx[e] + 1];
  for (int32_t v = 0; v < V; ++v) roff[v + 1] += roff[v];
  {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)V);
    memcpy(fill, roff, sizeof(int64_t) * (size_t)V);
    for (int32_t u = 0; u < V; ++u)
      for (int32_t e = off[u]; e < off[u + 1]; ++e) rsrc[fill[idx[e]]++] = u;
    free(fill);
  }
  int64_t reached = 0;
  for (int32_t r = 0; r < V; ++r) {
    if (lab[r] != (uint32_t)r) continue;
    int64_t top = 0;
    stack[top++] = r;
    seen[r] = 1;
    while (top > 0) {
      const int32_t u = stack[--top];
      ++reached;
      for (int32_t e = off[u]; e < off[u + 1]; ++e)
        if (!seen[idx[e]]) { seen[idx[e]] = 1; stack[top++] = idx[e]; }
      for (int64_t e = roff[u]; e < roff[u + 1]; ++e)
        if (!seen[rsrc[e]]) { seen[rsrc[e]] = 1; stack[top++] = rsrc[e]; }
    }
  }
  free(roff);
  free(rsrc);
  free(stack);
  free(seen);
  return reached == V ? ORC_OK : ORC_ESPLIT;
}

/* Checks n result rows (row i at rows + i * row_stride, kinds 0 BFS / 1 SSSP
 * / 2 WCC, sources srcs[i]) on `threads` threads, one row per thread at a
 * time; status[i] receives the row's error code.  WCC rows are checked with
 * orc_check_wcc.  Returns the number of rows that failed. */
typedef struct {
  int32_t V;
  const int32_t *off, *idx;
  const uint32_t* wt;
  const int32_t *kinds, *srcs;
  int32_t n;
  const uint32_t* rows;
  int64_t stride;
  int32_t* status;
  volatile int32_t next;
  pthread_mutex_t mu;
} orc_check_pool;

static void* orc_check_worker(void* arg) {
  orc_check_pool* p = (orc_check_pool*)arg;
  unsigned char* tight = (unsigned char*)malloc((size_t)(p->V > 0 ? p->V : 1));
  for (;;) {
    pthread_mutex_lock(&p->mu);
    const int32_t i = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (i >= p->n) break;
    const uint32_t* row = p->rows + (size_t)i * (size_t)p->stride;
    const int32_t k = p->kinds[i];
    if (k == 0) p->status[i] = check_sssp_row(p->V, p->off, p->idx, NULL, p->srcs[i], row, tight);
    else if (k == 1) p->status[i] = check_sssp_row(p->V, p->off, p->idx, p->wt, p->srcs[i], row, tight);
    else if (k == 2) p->status[i] = orc_check_wcc(p->V, p->off, p->idx, row);
    else p->status[i] = ORC_EKIND;
  }
  free(tight);
  return NULL;
}

int32_t orc_check_rows(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                       const int32_t* kinds, const int32_t* srcs, int32_t n, const uint32_t* rows,
                       int64_t row_stride, int32_t threads, int32_t* status) {
  orc_check_pool pool;
  pool.V = V; pool.off = off; pool.idx = idx; pool.wt = wt;
  pool.kinds = kinds; pool.srcs = srcs; pool.n = n; pool.rows = rows;
  pool.stride = row_stride; pool.status = status; pool.next = 0;
  pthread_mutex_init(&pool.mu, NULL);
  if (threads < 1) threads = 1;
  if (threads > n) threads = n > 0 ? n : 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, orc_check_worker, &pool);
  for (int32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&pool.mu);
  int32_t bad = 0;
  for (int32_t i = 0; i < n; ++i) bad += status[i] != ORC_OK;
  return bad;
}
