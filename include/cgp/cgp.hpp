// cgp.hpp — header-only C++ façade over the C-ABI of cgp.h, shaped like the
// reference's operator API so a reference caller switches by changing the
// include and the namespace.  Every non-OK status becomes an exception, as the
// reference throws (std::invalid_argument for bad batches, std::runtime_error
// for runtime failures; puercgp/include/puercgp/core/query_batch.hxx:44-56,
// puercgp/include/puercgp/engine/frontier_engine.hxx:27-32).
//
// Reference entry point                                   -> here
//   gunrock::cgp_bfs::run(G, sources, ctx, opts)            cgp::bfs::run(G, sources, ctx, opts)
//     (cgp/include/gunrock/algorithms/cgp_bfs.hxx:1050-1056)
//   gunrock::cgp_sssp::run(G, sources, ctx, opts)           cgp::sssp::run(...)
//     (cgp/include/gunrock/algorithms/cgp_sssp.hxx:973-979)
//   puercgp::run<wcc_policy>(graph, batch, ctx, opts)       cgp::wcc::run(G, n, ctx, opts)
//     (puercgp/include/puercgp/run.hxx:11-23)
//   puercgp::run_heterogeneous(graph, batch, ctx, opts)     cgp::run_heterogeneous(G, descs, ctx, opts)
//     (puercgp/include/puercgp/engine/hybrid_engine.hxx:1144-1150)
//   puercgp::run_replenish_pipeline(graph, descs, ctx, opts) cgp::run_replenish_pipeline(G, descs, ctx, opts)
//     (puercgp/include/puercgp/engine/replenish_engine.hxx:559-573)
//   puercgp::execution_context                              cgp::context (one engine, one stream)
//     (puercgp/include/puercgp/backend/graph_adapter.hxx:43-97)
#ifndef CGP_CGP_HPP
#define CGP_CGP_HPP

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cgp.h"

namespace cgp {

// ------------------------------------------------------------------ errors
struct invalid_argument : std::invalid_argument {
  cgp_status status;
  invalid_argument(cgp_status s, const std::string& m) : std::invalid_argument(m), status(s) {}
};
struct runtime_error : std::runtime_error {
  cgp_status status;
  runtime_error(cgp_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(cgp_status s) {
  if (s == CGP_OK) return;
  const char* m = cgp_last_error();
  if (s == CGP_EINVAL || s == CGP_ECAPACITY) throw invalid_argument(s, m ? m : "");
  throw runtime_error(s, m ? m : "");
}

constexpr uint32_t infinity = CGP_INF;

// Same ordinals as puercgp::algorithms::algo_kind_t.
enum class algo_kind_t : uint8_t { bfs = CGP_BFS, sssp = CGP_SSSP, wcc = CGP_WCC };
enum class traversal_mode_t { push = CGP_PUSH, pull = CGP_PULL, hybrid = CGP_HYBRID };

// query_descriptor_t{source, kind} (puercgp/include/puercgp/core/query_descriptor.hxx:19-24).
struct query_descriptor_t {
  int source = 0;
  algo_kind_t kind = algo_kind_t::bfs;
};

// ------------------------------------------------------------------ graphs
class host_csr {
 public:
  static host_csr rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                       uint32_t wmin = 0, uint32_t wmax = 0) {
    cgp_host_csr* h = nullptr;
    check(cgp_gen_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax, &h));
    return host_csr(h);
  }
  static host_csr grid(int rows, int cols, double p_extra, uint64_t seed, uint32_t wmin = 0,
                       uint32_t wmax = 0) {
    cgp_host_csr* h = nullptr;
    check(cgp_gen_grid(rows, cols, p_extra, seed, wmin, wmax, &h));
    return host_csr(h);
  }
  static host_csr load_gr(const std::string& path) { return load(cgp_load_gr, path); }
  static host_csr load_matrix_market(const std::string& path) { return load(cgp_load_matrix_market, path); }
  static host_csr load_csr_directory(const std::string& path) { return load(cgp_load_csr_directory, path); }

  int number_of_vertices() const { return view().V; }
  int64_t number_of_edges() const { return view().E; }
  const int32_t* row_offsets() const { return view().off; }
  const int32_t* column_indices() const { return view().idx; }
  const uint32_t* edge_weights() const { return view().wt; }
  std::vector<int> sample_sources(int n, uint64_t seed) const {
    std::vector<int> out((size_t)n);
    check(cgp_sample_sources(h_.get(), n, seed, out.data()));
    return out;
  }
  const cgp_host_csr* handle() const { return h_.get(); }

 private:
  struct View {
    int32_t V;
    int64_t E;
    const int32_t* off;
    const int32_t* idx;
    const uint32_t* wt;
  };
  struct Del {
    void operator()(cgp_host_csr* h) const { cgp_host_csr_destroy(h); }
  };
  explicit host_csr(cgp_host_csr* h) : h_(h) {}
  template <typename F>
  static host_csr load(F f, const std::string& path) {
    cgp_host_csr* h = nullptr;
    check(f(path.c_str(), &h));
    return host_csr(h);
  }
  View view() const {
    View v{};
    check(cgp_host_csr_view(h_.get(), &v.V, &v.E, &v.off, &v.idx, &v.wt));
    return v;
  }
  std::unique_ptr<cgp_host_csr, Del> h_;
};

// Device-resident CSR; the graph concept's accessors for host code.
class graph {
 public:
  // renumber: CGP_RENUMBER_AUTO / _ON / _OFF, the internal vertex order (results
  // are always in the caller's numbering).
  explicit graph(const host_csr& h, int32_t renumber = CGP_RENUMBER_AUTO) {
    cgp_graph* g = nullptr;
    check(cgp_graph_from_host_ex(h.handle(), renumber, &g));
    g_.reset(g);
  }
  graph(int32_t V, int64_t E, const int32_t* row_offsets, const int32_t* column_indices,
        const uint32_t* weights = nullptr, int32_t renumber = CGP_RENUMBER_AUTO) {
    cgp_graph* g = nullptr;
    check(cgp_graph_create_csr_ex(V, E, row_offsets, column_indices, weights, renumber, &g));
    g_.reset(g);
  }
  // Non-owning view over the caller's device (HBM) arrays, as csr_graph_view
  // (puercgp/include/puercgp/backend/csr_graph.hxx:9-20): nothing is copied,
  // and the arrays must outlive the graph.
  struct device_arrays_t {};
  static constexpr device_arrays_t device_arrays{};
  graph(device_arrays_t, int32_t V, int64_t E, const int32_t* d_row_offsets, const int32_t* d_column_indices,
        const uint32_t* d_weights = nullptr) {
    cgp_graph* g = nullptr;
    check(cgp_graph_create_csr_device(V, E, d_row_offsets, d_column_indices, d_weights, &g));
    g_.reset(g);
  }
  int get_number_of_vertices() const { return cgp_graph_num_vertices(g_.get()); }
  int64_t get_number_of_edges() const { return cgp_graph_num_edges(g_.get()); }
  bool is_symmetric() const { return cgp_graph_is_symmetric(g_.get()) != 0; }
  const cgp_graph* handle() const { return g_.get(); }

 private:
  struct Del {
    void operator()(cgp_graph* g) const { cgp_graph_destroy(g); }
  };
  std::unique_ptr<cgp_graph, Del> g_;
};

// run_options (puercgp/include/puercgp/core/types.hxx:26-46) as engine options.
struct run_options {
  traversal_mode_t traversal_mode = traversal_mode_t::hybrid;
  double pull_frontier_ratio = 0.15;
  double pull_edge_ratio = 0.20;
  int max_iterations = 0;
  int max_words = 0;
  int block_size = 0;
  int bin_small = 0;
  int bin_large = 0;
  bool use_hip_graph = true;
  int graph_iterations = 0;
  int profile = 0;
  uint64_t hbm_budget_bytes = 0;
  int dense_threshold = 0;
  int frontier_dense_div = 0;
  int dense_unroll = 0;
  int dense_narrow = 1;
  int concurrent_advance = 0;
  int pull_max_queries = 0;
  int advance_blocks = 0;
  int sssp_delta = -1;  // -1: automatic (engine default)
  int dense_wide_addr = 0;

  // Unset tunable fields stay zero: at context creation they take the
  // defaults, on a later call the context's current (possibly tuned) values.
  cgp_engine_opts to_c() const {
    cgp_engine_opts o{};
    o.traversal_mode = (int32_t)traversal_mode;
    o.pull_frontier_ratio = pull_frontier_ratio;
    o.pull_edge_ratio = pull_edge_ratio;
    o.max_iterations = max_iterations;
    if (max_words) o.max_words = max_words;
    if (block_size) o.block_size = block_size;
    if (bin_small) o.bin_small = bin_small;
    if (bin_large) o.bin_large = bin_large;
    o.use_hip_graph = use_hip_graph ? 1 : 0;
    if (graph_iterations) o.graph_iterations = graph_iterations;
    o.profile = profile;
    if (hbm_budget_bytes) o.hbm_budget_bytes = hbm_budget_bytes;
    if (dense_threshold) o.dense_threshold = dense_threshold;
    if (frontier_dense_div) o.frontier_dense_div = frontier_dense_div;
    if (dense_unroll) o.dense_unroll = dense_unroll;
    o.pull_max_queries = pull_max_queries;
    o.dense_narrow = dense_narrow;
    o.concurrent_advance = concurrent_advance;
    if (advance_blocks) o.advance_blocks = advance_blocks;
    o.sssp_delta = sssp_delta;
    o.dense_wide_addr = dense_wide_addr;
    return o;
  }
};

// execution_context analogue: one engine (HIP stream + batch state) per graph.
class context {
 public:
  context(const graph& g, const run_options& o = {}) : g_(&g) {
    cgp_engine_opts co = o.to_c();
    cgp_engine* e = nullptr;
    check(cgp_engine_create(g.handle(), &co, &e));
    e_.reset(e);
  }
  cgp_engine* handle() const { return e_.get(); }
  const graph& g() const { return *g_; }
  // Options for the following runs (the reference passes them per call).
  void apply(const run_options& o) {
    cgp_engine_opts co = o.to_c();
    check(cgp_engine_set_opts(e_.get(), &co));
  }
  // Per-graph tuning on a sample batch; the result sticks to this context.
  cgp_engine_opts autotune(const std::vector<query_descriptor_t>& sample) {
    std::vector<cgp_query> q = to_c(sample);
    cgp_engine_opts chosen;
    check(cgp_engine_autotune(e_.get(), q.data(), (int32_t)q.size(), &chosen));
    return chosen;
  }
  static std::vector<cgp_query> to_c(const std::vector<query_descriptor_t>& d) {
    std::vector<cgp_query> q(d.size());
    for (size_t i = 0; i < d.size(); ++i) q[i] = cgp_query{d[i].source, (int32_t)d[i].kind};
    return q;
  }

 private:
  struct Del {
    void operator()(cgp_engine* e) const { cgp_engine_destroy(e); }
  };
  const graph* g_;
  std::unique_ptr<cgp_engine, Del> e_;
};

// Device-resident uint32 array (the reference's results are
// thrust::device_vector, cgp_bfs.hxx:133); copies to the host on request.
class device_values {
 public:
  device_values() = default;
  explicit device_values(size_t n) : n_(n) {
    void* p = nullptr;
    check(cgp_device_alloc((uint64_t)std::max<size_t>(n, 1) * 4, &p));
    p_.reset(p);
  }
  uint32_t* data() { return (uint32_t*)p_.get(); }
  const uint32_t* data() const { return (const uint32_t*)p_.get(); }
  size_t size() const { return n_; }
  // Elements [first, first + count) into host memory.
  void copy_to_host(uint32_t* dst, size_t first, size_t count) const {
    if (count) check(cgp_copy_to_host(dst, data() + first, (uint64_t)count * 4));
  }
  std::vector<uint32_t> to_host() const {
    std::vector<uint32_t> h(n_);
    copy_to_host(h.data(), 0, n_);
    return h;
  }

 private:
  struct Del {
    void operator()(void* p) const { (void)cgp_device_free(p); }
  };
  std::unique_ptr<void, Del> p_;
  size_t n_ = 0;
};

// result_t: query-major values [q * V + v] (cgp_bfs.hxx:131-149), resident in
// HBM, plus run stats.
struct result_t {
  device_values distances;
  int n_vertices = 0;
  int n_queries = 0;
  cgp_run_stats stats{};
  int iterations = 0;
  float wall_time_ms = 0;
  const uint32_t* query_device(int q) const { return distances.data() + (size_t)q * (size_t)n_vertices; }
  std::vector<uint32_t> query_to_host(int q) const {
    std::vector<uint32_t> h((size_t)n_vertices);
    distances.copy_to_host(h.data(), (size_t)q * (size_t)n_vertices, (size_t)n_vertices);
    return h;
  }
};

// Runs a batch into caller-owned device memory d_out (n * V uint32) with
// this call's options: no allocation and no host copy of the results.
inline cgp_run_stats run_heterogeneous_into(const graph& G, const std::vector<query_descriptor_t>& batch,
                                            context& ctx, uint32_t* d_out, const run_options& o = {}) {
  if (&ctx.g() != &G) throw invalid_argument(CGP_EINVAL, "run: context belongs to another graph");
  if (batch.empty()) throw invalid_argument(CGP_EINVAL, "query batch must contain at least one source");
  ctx.apply(o);
  std::vector<cgp_query> q = context::to_c(batch);
  cgp_run_stats st{};
  check(cgp_run_batch(ctx.handle(), q.data(), (int32_t)q.size(), d_out, 1, &st));
  return st;
}

inline result_t run_heterogeneous(const graph& G, const std::vector<query_descriptor_t>& batch,
                                  context& ctx, const run_options& o = {}) {
  if (batch.empty()) throw invalid_argument(CGP_EINVAL, "query batch must contain at least one source");
  result_t r;
  r.n_vertices = G.get_number_of_vertices();
  r.n_queries = (int)batch.size();
  r.distances = device_values((size_t)r.n_queries * (size_t)r.n_vertices);
  r.stats = run_heterogeneous_into(G, batch, ctx, r.distances.data(), o);
  r.iterations = r.stats.iterations;
  r.wall_time_ms = (float)r.stats.wall_ms;
  return r;
}

namespace detail {
inline result_t run_kind(const graph& G, const std::vector<int>& sources, context& ctx,
                         const run_options& o, algo_kind_t k) {
  std::vector<query_descriptor_t> d(sources.size());
  for (size_t i = 0; i < sources.size(); ++i) d[i] = query_descriptor_t{sources[i], k};
  return run_heterogeneous(G, d, ctx, o);
}
}  // namespace detail

namespace bfs {
inline result_t run(const graph& G, const std::vector<int>& sources, context& ctx, const run_options& o = {}) {
  return detail::run_kind(G, sources, ctx, o, algo_kind_t::bfs);
}
}  // namespace bfs

namespace sssp {
inline result_t run(const graph& G, const std::vector<int>& sources, context& ctx, const run_options& o = {}) {
  return detail::run_kind(G, sources, ctx, o, algo_kind_t::sssp);
}
}  // namespace sssp

namespace wcc {
// WCC is source-free: n identical queries (the reference batch form).
inline result_t run(const graph& G, int n_queries, context& ctx, const run_options& o = {}) {
  return detail::run_kind(G, std::vector<int>((size_t)n_queries, 0), ctx, o, algo_kind_t::wcc);
}
}  // namespace wcc

// ------------------------------------------------------------------ serving
struct completion_t {
  int64_t ticket;
  algo_kind_t kind;
  int source;
  double arrival_s, admit_s, done_s;
  int iterations;
  std::vector<uint32_t> values;  // empty unless results are kept
};

class server {
 public:
  explicit server(context& ctx, int max_slots = 0, bool keep_results = true) : ctx_(ctx) {
    cgp_serve_opts so{max_slots, keep_results ? 1 : 0};
    check(cgp_serve_start(ctx_.handle(), &so));
  }
  ~server() { (void)cgp_serve_stop(ctx_.handle()); }
  server(const server&) = delete;
  server& operator=(const server&) = delete;
  double clock() const { return cgp_serve_clock(ctx_.handle()); }
  void submit(const query_descriptor_t& d, int64_t ticket, double arrival_s) {
    cgp_query q{d.source, (int32_t)d.kind};
    check(cgp_serve_submit(ctx_.handle(), &q, ticket, arrival_s));
  }
  // Runs until at least one completion (max_iterations == 0) or that many iterations.
  std::vector<completion_t> step(int max_iterations = 0, int cap = 256) {
    std::vector<cgp_completion> c((size_t)cap);
    int32_t n = 0;
    check(cgp_serve_step(ctx_.handle(), max_iterations, c.data(), cap, &n));
    std::vector<completion_t> out((size_t)n);
    const size_t V = (size_t)ctx_.g().get_number_of_vertices();
    for (int i = 0; i < n; ++i) {
      out[i] = completion_t{c[i].ticket, (algo_kind_t)c[i].kind, c[i].source, c[i].arrival_s,
                            c[i].admit_s, c[i].done_s, c[i].iterations, {}};
      if (c[i].result) {
        out[i].values.assign(c[i].result, c[i].result + V);
        cgp_serve_release_result(c[i].result);
      }
    }
    return out;
  }
  std::pair<int, int> pending() const {
    int32_t q = 0, r = 0;
    check(cgp_serve_pending(ctx_.handle(), &q, &r));
    return {q, r};
  }

 private:
  context& ctx_;
};

// Host-resident results of the serving loop (each completed query's row is
// copied to host memory while the traversal continues).
struct host_result_t {
  std::vector<uint32_t> distances;  // query-major [q * V + v]
  int n_vertices = 0;
  int n_queries = 0;
  const uint32_t* query(int q) const { return distances.data() + (size_t)q * (size_t)n_vertices; }
};

// run_replenish_pipeline: all queries submitted at once through `slots`
// concurrent slots; results in submission order, query-major.
inline host_result_t run_replenish_pipeline(const graph& G, const std::vector<query_descriptor_t>& queries,
                                            context& ctx, int slots = 64) {
  host_result_t r;
  r.n_vertices = G.get_number_of_vertices();
  r.n_queries = (int)queries.size();
  r.distances.assign((size_t)r.n_queries * (size_t)r.n_vertices, infinity);
  server s(ctx, slots, true);
  for (size_t i = 0; i < queries.size(); ++i) s.submit(queries[i], (int64_t)i, 0.0);
  size_t done = 0;
  while (done < queries.size()) {
    auto c = s.step(0);
    if (c.empty()) throw runtime_error(CGP_EDEVICE, "replenish: serving stopped without completing");
    for (auto& x : c) {
      std::copy(x.values.begin(), x.values.end(), r.distances.begin() + (size_t)x.ticket * (size_t)r.n_vertices);
      ++done;
    }
  }
  return r;
}

}  // namespace cgp

#endif  // CGP_CGP_HPP
