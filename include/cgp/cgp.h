/*
 * cgp.h — C-ABI of the MI355X-native concurrent multi-query graph engine.
 *
 * One traversal of a device-resident CSR graph serves a whole batch of BFS,
 * SSSP and WCC queries: every vertex carries W 64-bit query-mask words (one
 * bit per query slot, the wave64 __ballot width), BFS bits are settled with
 * one 64-bit atomicOr per (edge, word), SSSP values relax with native uint32
 * atomicMin, and WCC is computed by hooking plus pointer jumping.
 *
 * The reference exposes this path only as header-only C++ templates; the
 * entry points below are what its FFI for the path would bind.  Each one
 * names the reference interface it replaces (paths relative to the
 * reference root).
 *
 * Conventions (all entry points):
 *   - C linkage, plain pointers and sizes, no exceptions cross the ABI.
 *   - Every call returns a cgp_status; on failure cgp_last_error() returns a
 *     thread-local message of the form "<op>: <reason>", the text the
 *     reference puts into its std::runtime_error / std::invalid_argument
 *     (puercgp/include/puercgp/engine/frontier_engine.hxx:27-32,
 *      puercgp/include/puercgp/core/query_batch.hxx:44-56).
 *   - Results are uint32 per (query, vertex), query-major out[q*V + v]
 *     (cgp/include/gunrock/algorithms/cgp_bfs.hxx:133,214), with
 *     CGP_INF = 0xFFFFFFFF for "unreachable".
 *   - An engine owns one HIP stream of default priority; calls on one engine
 *     are serialised by the caller (as with one reference execution_context,
 *     puercgp/include/puercgp/backend/graph_adapter.hxx:43-97).
 */
#ifndef CGP_CGP_H
#define CGP_CGP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGP_ABI_VERSION 1
#define CGP_INF 0xFFFFFFFFu

typedef enum {
  CGP_OK = 0,
  CGP_EINVAL = 1,     /* bad argument or batch (reference: std::invalid_argument) */
  CGP_ENOMEM = 2,     /* device or host allocation failed */
  CGP_EDEVICE = 3,    /* HIP runtime error or no usable GPU (reference: std::runtime_error) */
  CGP_EBUSY = 4,      /* engine already running a batch / serving session */
  CGP_ECAPACITY = 5,  /* request exceeds slots or HBM budget */
  CGP_EIO = 6         /* graph file could not be read or parsed */
} cgp_status;

/* Same ordinals as algo_kind_t (puercgp/include/puercgp/algorithms/hybrid.hxx:21-26). */
typedef enum { CGP_BFS = 0, CGP_SSSP = 1, CGP_WCC = 2 } cgp_kind;

/* Replaces query_descriptor_t{source, kind} (puercgp/include/puercgp/core/query_descriptor.hxx:19-24).
 * WCC ignores source. */
typedef struct {
  int32_t source;
  int32_t kind;
} cgp_query;

/* Traversal direction policy; replaces traversal_mode_t (puercgp/include/puercgp/core/types.hxx:10). */
typedef enum { CGP_PUSH = 0, CGP_PULL = 1, CGP_HYBRID = 2 } cgp_traversal_mode;

typedef struct cgp_graph cgp_graph;        /* device-resident CSR */
typedef struct cgp_host_csr cgp_host_csr;  /* host CSR produced by a generator or loader */
typedef struct cgp_engine cgp_engine;      /* per-graph batch state, stream, tuned parameters */

/* Engine options.  Replaces run_options (puercgp/include/puercgp/core/types.hxx:26-46)
 * and cgp_bfs::options_t (cgp/include/gunrock/algorithms/cgp_bfs.hxx:97-105).
 * Zero in a tunable field means "use the tuned / default value". */
typedef struct {
  int32_t traversal_mode;      /* cgp_traversal_mode; default CGP_HYBRID */
  double pull_frontier_ratio;  /* pull when frontier >= ratio * V (default 0.15, types.hxx:30) */
  double pull_edge_ratio;      /* pull when frontier edges >= ratio * E (default 0.20, types.hxx:31) */
  int32_t max_iterations;      /* 0 = until convergence (types.hxx:36) */
  int32_t max_words;           /* cap on 64-bit mask words per pass; 0 = tuned */
  int32_t block_size;          /* workgroup size of the advance kernel: 128, 256 or 512; 0 = default 256 */
  int32_t bin_small;           /* degree <= bin_small: 8 lanes per vertex; 0 = tuned */
  int32_t bin_large;           /* degree > bin_large: one workgroup per vertex; 0 = tuned */
  int32_t use_hip_graph;       /* capture iterations into a hipGraph (default 1) */
  int32_t graph_iterations;    /* iterations per captured graph launch (default 8) */
  int32_t profile;             /* bit 0: per-launch HIP events (also inside graph replays);
                                  bit 1: device work counters -> advance_bytes */
  uint64_t hbm_budget_bytes;   /* serving: cap on per-query value memory; 0 = 64 GiB */
  int32_t dense_threshold;     /* active SSSP queries in a mask word at which the word is relaxed
                                  lane-per-query (one wave per edge row); 0 = automatic: 6 when
                                  sssp_delta resolves to 0 (postponing off), else 8 */
  int32_t frontier_dense_div;  /* dense frontier mode (fire-and-forget atomics, filter scans all
                                  vertices) once the frontier's edges reach V / div; 0 = 8 */
  int32_t dense_unroll;        /* value rows in flight per step of the lane-per-query kernel: 4, 8
                                  or 16 (fewer registers, more waves vs more rows per wave); 0 = 16 */
  int32_t dense_narrow;        /* 1 = when every SSSP word of a pass has its SSSP slots in lanes
                                  0..15, those words relax four edges per lane-per-query step
                                  (own item list and launch); 0 = off */
  int32_t concurrent_advance;  /* 1 = the push launch runs beside the lane-per-query launches on
                                  a second stream (fork/join events, also inside graphs); 0
                                  (default) = one stream */
  int32_t pull_max_queries;    /* hybrid mode takes pull steps only for batches of at most this
                                  many queries of the kind: bottom-up steps for BFS, min-plus
                                  gathers for SSSP (0 = never; CGP_PULL always pulls).  Pulls need
                                  a symmetric graph (SSSP: with symmetric weights) and, for SSSP,
                                  postponing off (sssp_delta 0) */
  int32_t advance_blocks;      /* workgroups of the advance launches (256-thread equivalents);
                                  0 = one per 65536 edges, between 1 and 8 per CU */
  int32_t sssp_delta;          /* SSSP near-far step: frontier pairs whose value is >= a threshold
                                  wait; the threshold moves to the smallest waiting value + delta
                                  when nothing below it is left.  0 = no waiting (Bellman-Ford
                                  frontier); -1 (default) = automatic: 20 x mean edge weight on
                                  weighted graphs of average degree <= 8, else 0.  Results are
                                  identical either way. */
  int32_t dense_wide_addr;     /* 1 = the lane-per-query loop addresses value rows with 64-bit
                                  addresses even where 32-bit offsets inside one word's rows
                                  reach them (V x 256 bytes < 4 GiB); 0 (default) = only where
                                  they do not */
} cgp_engine_opts;

/* Indices of cgp_run_stats.work: counts of the advance launches' work units. */
#define CGP_WORK_COUNTERS 24
typedef enum {
  CGP_WORK_PUSH_ITEMS = 0,  /* frontier items expanded lane-per-edge (vertex or edge chunk) */
  CGP_WORK_PUSH_EDGES = 1,  /* edges scanned lane-per-edge */
  CGP_WORK_PUSH_WEDGES = 2, /* ... of which read a weight (SSSP bits active) */
  CGP_WORK_BFS_WORDS = 3,   /* (edge, active BFS mask word) visited reads */
  CGP_WORK_SSSP_PAIRS = 4,  /* (edge, active sparse SSSP query) target value reads */
  CGP_WORK_NEW_WORDS = 5,   /* next-frontier words that gained bits (push launch) */
  CGP_WORK_IMPROVED = 6,    /* (vertex, query) values lowered (push launch) */
  CGP_WORK_DENSE_ITEMS = 7, /* lane-per-query (vertex, word, edge chunk) items */
  CGP_WORK_DENSE_EDGES = 8, /* edges scanned lane-per-query, one per (edge, active SSSP word) */
  CGP_WORK_PULL_VERTS = 9,  /* vertices examined bottom-up */
  CGP_WORK_PULL_WORDS = 10, /* (edge, word) reads bottom-up */
  CGP_WORK_PULL_EDGES = 11, /* edges scanned bottom-up */
  CGP_WORK_DENSE_PAIRS = 12,/* (edge, active SSSP query) pairs lane-per-query */
  CGP_WORK_DENSE_IMP = 13,  /* values lowered lane-per-query */
  CGP_WORK_DENSE_NEW = 14,  /* next-frontier words that gained bits lane-per-query */
  CGP_WORK_SSSP_WORDS = 15, /* (edge, mask word with active sparse SSSP bits) pairs */
  CGP_WORK_SSSP_SRC = 16,   /* (item, active sparse SSSP query) source value reads */
  CGP_WORK_DENSE_SRC = 17   /* (lane-per-query item, active SSSP query) pairs */
} cgp_work_counter;

/* Per-run report; replaces run_result_t / iteration_profile_t
 * (puercgp/include/puercgp/core/result.hxx:13-66) and cgp_bfs::result_t timing fields. */
typedef struct {
  int32_t iterations;          /* frontier iterations executed */
  int32_t push_iterations;
  int32_t pull_iterations;
  int32_t passes;              /* word-group passes (batch split by max_words) */
  double wall_ms;              /* host wall, first launch -> results resident on the device */
  double init_ms;              /* profile bit 0: init kernels */
  double advance_ms;           /* profile bit 0: summed advance kernel time */
  double filter_ms;            /* profile bit 0: summed compaction kernel time */
  double wcc_ms;               /* profile bit 0: hooking + pointer jumping */
  double output_ms;            /* profile bit 0: result transpose */
  int64_t advance_launches;    /* profile bit 0: advance launches that did work */
  double advance_bytes;        /* profile bit 1: algorithmic bytes of all advance launches */
  int64_t frontier_vertices;   /* profile bit 1: frontier items expanded */
  int64_t edges_scanned;       /* profile bit 1: edges touched by push and pull */
  double dense_ms;             /* profile bit 0: summed time of the dense-SSSP advance launches */
  int64_t dense_launches;      /* profile bit 0: dense-SSSP advance launches that had work */
  double dense_bytes;          /* profile bit 1: algorithmic bytes of those launches */
  uint64_t work[CGP_WORK_COUNTERS]; /* profile bit 1: raw advance work counters summed over the
                                  passes, indexed by cgp_work_counter (the units both byte
                                  models are computed from) */
} cgp_run_stats;

/* ---------------------------------------------------------------- errors */
const char* cgp_last_error(void);
int32_t cgp_abi_version(void);

/* ---------------------------------------------------------------- device */
/* Number of visible HIP devices (0 on a host without a usable GPU). */
cgp_status cgp_device_count(int32_t* count);
/* Name and compute-unit count of device 0. */
cgp_status cgp_device_info(char* name, int32_t name_cap, int32_t* compute_units, uint64_t* hbm_bytes);
/* Plain device buffers for callers without a device allocator of their own. */
cgp_status cgp_device_alloc(uint64_t bytes, void** out);
cgp_status cgp_device_free(void* ptr);
cgp_status cgp_copy_to_host(void* dst_host, const void* src_device, uint64_t bytes);
cgp_status cgp_copy_to_device(void* dst_device, const void* src_host, uint64_t bytes);
/* Device-to-device copy (synchronous).  16-byte aligned pointers and sizes run one
 * library kernel that reads and writes each byte exactly once (a known-traffic
 * reference for hardware counters); other sizes use the runtime's copy. */
cgp_status cgp_copy_device(void* dst_device, const void* src_device, uint64_t bytes);

/* ---------------------------------------------------------------- host CSR: generators and loaders */
/* RMAT (a,b,c, d = 1-a-b-c) with 2^scale vertices and edge_factor * 2^scale edge draws,
 * self-loops dropped, symmetrised, deduplicated, sorted; weights uniform in
 * [wmin, wmax] from a hash of (min(u,v), max(u,v), seed) so both directions agree.
 * wmax == 0 produces an unweighted graph. The reference path has no generator. */
cgp_status cgp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t wmin, uint32_t wmax, cgp_host_csr** out);
/* rows x cols grid, 4-neighbour undirected edges, plus per vertex with probability
 * p_extra one undirected edge to a uniformly chosen vertex. */
cgp_status cgp_gen_grid(int32_t rows, int32_t cols, double p_extra, uint64_t seed,
                        uint32_t wmin, uint32_t wmax, cgp_host_csr** out);
/* Build a host CSR from a COO edge list (sorted; duplicates kept, as csr_t::from_coo,
 * cgp/include/gunrock/formats/csr.hxx:81-140). */
cgp_status cgp_host_csr_from_coo(int32_t n_vertices, int64_t n_edges, const int32_t* src,
                                 const int32_t* dst, const uint32_t* weights, cgp_host_csr** out);
/* Loaders (reference: load_gr cgp/examples/algorithms/bfs/cgp_bfs.cu:214-275,
 * load_matrix_market puercgp/examples/mtx_loader.hxx:99-187,
 * load_binary_csr_directory puercgp/examples/mtx_loader.hxx:189-260). */
cgp_status cgp_load_gr(const char* path, cgp_host_csr** out);
cgp_status cgp_load_matrix_market(const char* path, cgp_host_csr** out);
cgp_status cgp_load_csr_directory(const char* path, cgp_host_csr** out);
/* The three loaders with flags (the ones above pass 0).
 *   CGP_LOAD_UNIT_WEIGHTS: ignore stored weights (unit-weight SSSP); the reference's SSSP
 *     driver reads .gr graphs this way (cgp/examples/algorithms/sssp/cgp_sssp.cu:53-56).
 *   CGP_LOAD_ROUND_WEIGHTS: Matrix Market values that are not integers in [0, 2^32-1]
 *     are rounded into that range; without it they are rejected (CGP_EIO), since
 *     distances are exact uint32 sums. */
#define CGP_FORMAT_GR 0
#define CGP_FORMAT_MATRIX_MARKET 1
#define CGP_FORMAT_CSR_DIRECTORY 2
#define CGP_LOAD_UNIT_WEIGHTS 1
#define CGP_LOAD_ROUND_WEIGHTS 2
cgp_status cgp_load_graph(const char* path, int32_t format, int32_t flags, cgp_host_csr** out);
/* Borrow the arrays of a host CSR (valid until cgp_host_csr_destroy). weights may be NULL. */
cgp_status cgp_host_csr_view(const cgp_host_csr* g, int32_t* n_vertices, int64_t* n_edges,
                             const int32_t** row_offsets, const int32_t** col_indices,
                             const uint32_t** weights);
void cgp_host_csr_destroy(cgp_host_csr* g);
/* n distinct sources drawn uniformly from vertices with out-degree >= 1. */
cgp_status cgp_sample_sources(const cgp_host_csr* g, int32_t n, uint64_t seed, int32_t* out);
/* Open-loop Poisson arrivals: n exponential inter-arrival gaps of rate `rate_per_s`,
 * cumulative, in seconds. */
cgp_status cgp_poisson_arrivals(int32_t n, double rate_per_s, uint64_t seed, double* out_s);

/* ---------------------------------------------------------------- device graph (A1) */
/* Replaces csr_graph_storage / graph::build (puercgp/include/puercgp/backend/csr_graph.hxx:9-75,
 * cgp/include/gunrock/graph/build.hxx:28-35). Copies host CSR into HBM.
 * weights == NULL means unit weights (csr_graph.hxx:37-39). */
cgp_status cgp_graph_create_csr(int32_t n_vertices, int64_t n_edges, const int32_t* row_offsets,
                                const int32_t* col_indices, const uint32_t* weights,
                                cgp_graph** out);
cgp_status cgp_graph_from_host(const cgp_host_csr* g, cgp_graph** out);
/* As cgp_graph_create_csr / cgp_graph_from_host (the same csr_graph_storage / graph::build
 * counterparts, which always store the caller's numbering), with the internal vertex
 * numbering chosen by `renumber`: CGP_RENUMBER_AUTO (what the two entries above do: renumber by descending
 * degree when the graph has at least 2^21 vertices and its largest degree is at least 32x
 * the mean), CGP_RENUMBER_ON (always), CGP_RENUMBER_OFF (keep the caller's numbering).  The
 * numbering is internal: sources, results and labels are always in the caller's numbering. */
#define CGP_RENUMBER_AUTO 0
#define CGP_RENUMBER_ON 1
#define CGP_RENUMBER_OFF 2
cgp_status cgp_graph_create_csr_ex(int32_t n_vertices, int64_t n_edges, const int32_t* row_offsets,
                                   const int32_t* col_indices, const uint32_t* weights, int32_t renumber,
                                   cgp_graph** out);
cgp_status cgp_graph_from_host_ex(const cgp_host_csr* g, int32_t renumber, cgp_graph** out);
/* Non-owning graph over caller-owned device (HBM) arrays; replaces the non-owning
 * csr_graph_view (puercgp/include/puercgp/backend/csr_graph.hxx:9-20).  row_offsets (V+1),
 * col_indices (E) and weights (E, or NULL for unit weights) are device pointers that must
 * stay valid and unchanged while the graph lives; nothing is copied.  The library checks
 * them on the device (offsets span [0, E] and never decrease, indices in [0, V):
 * CGP_EINVAL otherwise) and builds its own derived arrays (source row of each edge slot,
 * degree lists).  The graph counts as symmetric when every row is sorted and every edge
 * has its reverse; otherwise it is treated as directed. */
cgp_status cgp_graph_create_csr_device(int32_t n_vertices, int64_t n_edges, const int32_t* row_offsets,
                                       const int32_t* col_indices, const uint32_t* weights,
                                       cgp_graph** out);
void cgp_graph_destroy(cgp_graph* g);
int32_t cgp_graph_num_vertices(const cgp_graph* g);
int64_t cgp_graph_num_edges(const cgp_graph* g);
/* 1 when every edge u->v has a matching v->u (pull then reads the out-rows). */
int32_t cgp_graph_is_symmetric(const cgp_graph* g);

/* ---------------------------------------------------------------- engine */
void cgp_engine_opts_default(cgp_engine_opts* opts);
cgp_status cgp_engine_create(const cgp_graph* g, const cgp_engine_opts* opts, cgp_engine** out);
void cgp_engine_destroy(cgp_engine* e);
/* Options actually in force (after defaults and tuning). */
cgp_status cgp_engine_get_opts(const cgp_engine* e, cgp_engine_opts* out);
/* Options for the following runs of this engine: the per-call options of
 * cgp_bfs::run(G, sources, ctx, options) (cgp/include/gunrock/algorithms/cgp_bfs.hxx:1050-1056)
 * and puercgp::run(graph, batch, ctx, run_options) (puercgp/include/puercgp/run.hxx:11-23).
 * A zero tunable field keeps the engine's current (tuned or default) value; automatic
 * fields (sssp_delta -1, dense_threshold 0) resolve again for the graph.  Captured
 * iteration graphs are dropped when the options change.  CGP_EBUSY while serving. */
cgp_status cgp_engine_set_opts(cgp_engine* e, const cgp_engine_opts* opts);

/* Batched run of n queries of any mix of kinds.
 * Replaces gunrock::cgp_bfs::run / cgp_sssp::run (cgp/include/gunrock/algorithms/cgp_bfs.hxx:1050-1056,
 * cgp_sssp.hxx:973-979), puercgp::run<Policy> (puercgp/include/puercgp/run.hxx:11-23) and
 * run_heterogeneous (puercgp/include/puercgp/engine/hybrid_engine.hxx:1144-1150).
 * out: n*V uint32, query-major; out_on_device != 0 means `out` is a device pointer,
 * otherwise a host pointer (results are copied back before return). stats may be NULL. */
cgp_status cgp_run_batch(cgp_engine* e, const cgp_query* queries, int32_t n, uint32_t* out,
                         int32_t out_on_device, cgp_run_stats* stats);

/* Allocation-free batched run.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as cgp_run_batch on the same engine,
 * with every per-run device buffer taken from one fixed device array built into the
 * library (CGP_STATIC_SCRATCH_BYTES per GPU, held from library load until the process
 * ends) instead of the engine's buffers; it calls no hipMalloc / hipFree /
 * hipHostMalloc.  out_device: n*V uint32 of device memory, query-major.  Batches of
 * 1..CGP_STATIC_MAX_QUERIES queries whose state fits the array (about V*(44 + 56*W +
 * 4*W*R) + E*W/4 bytes for W mask words of R-lane value rows, R <= 64) run; larger
 * ones fail with CGP_ECAPACITY and a message giving the bytes needed.  Iterations are
 * launched directly (no graph capture); stats timing fields stay 0 and no
 * per-iteration report is kept.  Every engine on one GPU shares the array: calls in
 * one process are serialised, calls from two processes on one GPU race. */
#define CGP_STATIC_SCRATCH_BYTES (64u << 20)
#define CGP_STATIC_MAX_QUERIES 1024
cgp_status cgp_run_batch_static(cgp_engine* e, const cgp_query* queries, int32_t n, uint32_t* out_device,
                                cgp_run_stats* stats);

/* One record per frontier iteration of the last cgp_run_batch, written on the
 * device by the compaction kernel (no host sync); replaces iteration_profile_t /
 * level_profile_t (puercgp/include/puercgp/core/result.hxx:21-46,
 * cgp/include/gunrock/algorithms/cgp_bfs.hxx:107-129). */
typedef struct {
  int32_t pass;              /* word-group pass of the batch */
  int32_t iteration;
  int32_t pull;              /* 1: BFS bits went bottom-up */
  int32_t dense;             /* 1: dense frontier mode (full compaction scan) */
  int64_t frontier_items;    /* work items the advance consumed */
  int64_t touched;           /* vertices the compaction visited */
  int64_t next_items;        /* work items of the next frontier */
  int64_t next_edges;        /* summed degree of the next frontier */
  int64_t next_bfs_pairs;    /* BFS (vertex, query) pairs of the next frontier */
  double advance_ms;         /* device-clock duration of the advance launch */
  double filter_ms;          /* ... of the compaction launch */
  int64_t next_pairs;        /* (vertex, query) pairs of the next frontier, every kind */
  int64_t next_vertices;     /* distinct vertices of the next frontier */
  int64_t next_query_edges;  /* sum over the next frontier's pairs of the vertex degree */
  int64_t edges_scanned;     /* profile bit 1: edges this iteration's advance scanned (else 0) */
  double dense_ms;           /* device-clock duration of the lane-per-query launch (0: none) */
  double end_ms;             /* device clock at this iteration's end, from the pass's start */
  uint64_t frontier_union[8];/* OR of the frontier words this iteration's advance consumed:
                                bit b of word w = slot 64w+b of the pass had frontier work */
} cgp_iteration_record;

/* Copies up to cap records of the last batch run; *n = records copied
 * (cap == 0: *n = records available). */
cgp_status cgp_last_run_report(const cgp_engine* e, cgp_iteration_record* out, int32_t cap, int32_t* n);

/* The last batch run's report as text, the fields of the reference driver's exports
 * (cgp/examples/algorithms/bfs/cgp_bfs.cu:359-555: export_iter_profile CSV rows and
 * export_json, with cgp_bfs::result_t / level_profile_t fields, cgp_bfs.hxx:107-149),
 * filled from the device-written iteration records:
 *   CGP_REPORT_JSON  one JSON object: graph and device, the queries (source, kind,
 *                    completion level and time), per-level frontier sizes, unique
 *                    frontier sizes, level edge counts (frontier degree sum, scanned
 *                    edges, query-weighted), level modes and wall times, level profiles;
 *   CGP_REPORT_CSV   the per-iteration profile rows (header line first).
 * `dataset` names the graph in the output (may be NULL).  Writes at most cap bytes
 * including the terminating NUL; *len = bytes the whole text needs (without the NUL),
 * so a call with cap == 0 sizes the buffer.  CGP_EINVAL when no batch ran yet. */
#define CGP_REPORT_JSON 0
#define CGP_REPORT_CSV 1
cgp_status cgp_last_run_report_text(const cgp_engine* e, int32_t format, const char* dataset, char* buf,
                                    uint64_t cap, uint64_t* len);

/* Per-query edge counts m_q = sum of out-degrees of the vertices reached by query q
 * (out[q*V+v] != CGP_INF); device-resident results in, n counts out (host). */
cgp_status cgp_reached_edges(const cgp_graph* g, const uint32_t* out_device, int32_t n,
                             uint64_t* m_out);

/* Tune bitmask word count, degree-bin thresholds and workgroup size for this graph
 * on a sample batch (n queries); results stick to the engine and are returned in `chosen`. */
cgp_status cgp_engine_autotune(cgp_engine* e, const cgp_query* sample, int32_t n,
                               cgp_engine_opts* chosen);

/* ---------------------------------------------------------------- serving (continuous batching) */
/* Replaces run_replenish_pipeline (puercgp/include/puercgp/engine/replenish_engine.hxx:559-573). */
typedef struct {
  int32_t max_slots;        /* bitmask slots kept resident; 0 = as many as the HBM budget allows
                               (never more: see cgp_serve_capacity) */
  int32_t keep_results;     /* 1 = completions carry the V-length result row (host copy) */
} cgp_serve_opts;

typedef struct {
  int64_t ticket;           /* caller's id from cgp_serve_submit */
  int32_t kind;
  int32_t source;
  double arrival_s;         /* caller clock, seconds (as submitted) */
  double admit_s;           /* engine clock when the query got a slot */
  double done_s;            /* engine clock when its result was resident on the host */
  int32_t iterations;       /* frontier iterations the query took */
  uint32_t* result;         /* V values when keep_results, else NULL; owned by the engine
                               until cgp_serve_release_result */
} cgp_completion;

cgp_status cgp_serve_start(cgp_engine* e, const cgp_serve_opts* opts);
/* Enqueue one query; it is admitted into a free slot at the next iteration boundary
 * once its value arrays fit the HBM budget. Non-blocking. */
cgp_status cgp_serve_submit(cgp_engine* e, const cgp_query* q, int64_t ticket, double arrival_s);
/* Drive the device: run iterations until at least one query completes or
 * `max_iterations` iterations ran (0 = until idle). Collects up to cap completions. */
cgp_status cgp_serve_step(cgp_engine* e, int32_t max_iterations, cgp_completion* done,
                          int32_t cap, int32_t* n_done);
/* Engine clock (seconds since cgp_serve_start), same clock as admit_s / done_s. */
double cgp_serve_clock(const cgp_engine* e);
/* queued: waiting for a slot; running: holding a slot.  A converged query frees its
 * slot at once and completes when its result reaches host memory (the copy runs while
 * the traversal continues); cgp_serve_step on an idle engine finishes those copies. */
cgp_status cgp_serve_pending(const cgp_engine* e, int32_t* queued, int32_t* running);
/* Slots of the running session and the bytes of per-query value rows it holds, at most
 * hbm_budget_bytes: the largest count of whole 64-slot words, or of one word of 2^k-lane
 * rows, whose rows fit the budget (result rows staged for their copy to host memory are
 * extra, one per completed query whose copy is in flight). */
cgp_status cgp_serve_capacity(const cgp_engine* e, int32_t* slots, uint64_t* value_bytes);
void cgp_serve_release_result(uint32_t* result);
cgp_status cgp_serve_stop(cgp_engine* e);

#ifdef __cplusplus
}
#endif
#endif /* CGP_CGP_H */
