// ref_wrap.cpp — extern "C" shims over the reference's own CPU validators,
// compiled from the reference sources where they lie (see oracle/Makefile).
// TEST INFRASTRUCTURE ONLY: the output goes to oracle/_ref/ and is loaded by
// tests (and optionally bench.py's cpu_baseline leg) to pin oracle/oracle.c.
//
// Wrapped functions (paths relative to the reference root):
//   cpu_bfs / cpu_sssp / cpu_wcc   puercgp/examples/mtx_loader.hxx:269-356
//   bfs_cpu::run                    cgp/examples/algorithms/bfs/bfs_cpu.hxx:20-68
//   sssp_cpu::run                   cgp/examples/algorithms/sssp/sssp_cpu.hxx:22-72
#include <thrust/host_vector.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

#include "mtx_loader.hxx"
#include "bfs/bfs_cpu.hxx"
#include "sssp/sssp_cpu.hxx"

namespace {

constexpr uint32_t kInf = 0xFFFFFFFFu;

puercgp_examples::host_csr_graph make_puer(int32_t V, const int32_t* off, const int32_t* idx,
                                           const uint32_t* wt) {
  puercgp_examples::host_csr_graph g;
  g.vertices = V;
  g.edges = off[V];
  g.row_offsets.assign(off, off + V + 1);
  g.column_indices.assign(idx, idx + off[V]);
  g.edge_weights.resize(static_cast<size_t>(off[V]));
  for (int32_t e = 0; e < off[V]; ++e) g.edge_weights[e] = wt ? static_cast<float>(wt[e]) : 1.0f;
  return g;
}

// The template argument the cgp validators read (.number_of_rows, .row_offsets,
// .column_indices, .nonzero_values), as gunrock's csr_t provides.
template <typename W>
struct cgp_csr {
  int number_of_rows = 0;
  std::vector<int> row_offsets, column_indices;
  std::vector<W> nonzero_values;
};

}  // namespace

extern "C" {

void ref_puer_bfs(int32_t V, const int32_t* off, const int32_t* idx, int32_t src, uint32_t* out) {
  auto g = make_puer(V, off, idx, nullptr);
  auto d = puercgp_examples::cpu_bfs(g, src);
  for (int32_t v = 0; v < V; ++v)
    out[v] = d[v] == std::numeric_limits<int>::max() ? kInf : static_cast<uint32_t>(d[v]);
}

void ref_puer_sssp(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                   int32_t src, uint32_t* out) {
  auto g = make_puer(V, off, idx, wt);
  auto d = puercgp_examples::cpu_sssp(g, src);
  for (int32_t v = 0; v < V; ++v) out[v] = std::isinf(d[v]) ? kInf : static_cast<uint32_t>(d[v]);
}

void ref_puer_wcc(int32_t V, const int32_t* off, const int32_t* idx, int32_t max_iterations,
                  uint32_t* out) {
  auto g = make_puer(V, off, idx, nullptr);
  auto l = puercgp_examples::cpu_wcc(g, max_iterations);
  for (int32_t v = 0; v < V; ++v) out[v] = static_cast<uint32_t>(l[v]);
}

void ref_cgp_bfs(int32_t V, const int32_t* off, const int32_t* idx, int32_t src, uint32_t* out) {
  cgp_csr<int> c;
  c.number_of_rows = V;
  c.row_offsets.assign(off, off + V + 1);
  c.column_indices.assign(idx, idx + off[V]);
  std::vector<int> d(static_cast<size_t>(V)), p(static_cast<size_t>(V));
  int s = src;
  bfs_cpu::run<cgp_csr<int>, int, int>(c, s, d.data(), p.data());
  for (int32_t v = 0; v < V; ++v)
    out[v] = d[v] == std::numeric_limits<int>::max() ? kInf : static_cast<uint32_t>(d[v]);
}

void ref_cgp_sssp(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                  int32_t src, uint32_t* out) {
  cgp_csr<uint32_t> c;
  c.number_of_rows = V;
  c.row_offsets.assign(off, off + V + 1);
  c.column_indices.assign(idx, idx + off[V]);
  c.nonzero_values.resize(static_cast<size_t>(off[V]));
  for (int32_t e = 0; e < off[V]; ++e) c.nonzero_values[e] = wt ? wt[e] : 1u;
  std::vector<int> p(static_cast<size_t>(V));
  int s = src;
  sssp_cpu::run<cgp_csr<uint32_t>, int, int, uint32_t>(c, s, out, p.data());
}

// Times the reference's puercgp validators over a batch (the CPU baseline
// of the benchmark): the graph is converted to the reference's host graph
// once, outside the clock; `threads` workers take the queries in order, one
// query per worker at a time (kinds: 0 cpu_bfs, 1 cpu_sssp, 2 cpu_wcc to
// convergence).  Returns the wall seconds of the query runs; checksum gets
// the sum over queries of the finite values' count (keeps the results live).
double ref_run_queries(int32_t V, const int32_t* off, const int32_t* idx, const uint32_t* wt,
                       const int32_t* kinds, const int32_t* sources, int32_t n, int32_t threads,
                       uint64_t* checksum) {
  const auto g = make_puer(V, off, idx, wt);
  std::atomic<int32_t> next{0};
  std::atomic<uint64_t> sum{0};
  auto worker = [&]() {
    for (int32_t i = next++; i < n; i = next++) {
      uint64_t c = 0;
      if (kinds[i] == 0) {
        for (int d : puercgp_examples::cpu_bfs(g, sources[i])) c += d != std::numeric_limits<int>::max();
      } else if (kinds[i] == 1) {
        for (float d : puercgp_examples::cpu_sssp(g, sources[i])) c += !std::isinf(d);
      } else {
        for (float l : puercgp_examples::cpu_wcc(g, std::numeric_limits<int>::max())) c += l >= 0.0f;
      }
      sum += c;
    }
  };
  const int nt = threads < 1 ? 1 : threads;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (checksum) *checksum = sum.load();
  return secs;
}

}  // extern "C"
