// facade_check.cpp — the reference's known-answer validators, driven through
// the C++ façade (cgp.hpp) and the C-ABI: validate_hybrid.cu (mixed batch on
// the toy graph), validate_bfs.cu / validate_sssp.cu defaults (Matrix Market
// fixtures, sources 0..3) and validate_replenish.cu (4 queries through 2
// slots).  Prints [PASS]/[FAIL] lines and exits non-zero on any failure, as
// the reference's examples do.  Usage: facade_check <fixture dir>
#include <cstdio>
#include <string>
#include <vector>

#include "cgp/cgp.hpp"

static int failures = 0;
static void check(const char* name, bool cond) {
  std::printf("  [%s] %s\n", cond ? "PASS" : "FAIL", name);
  if (!cond) ++failures;
}

static std::vector<uint32_t> row(const cgp::result_t& r, int q) { return r.query_to_host(q); }
static std::vector<uint32_t> row(const cgp::host_result_t& r, int q) {
  return std::vector<uint32_t>(r.query(q), r.query(q) + r.n_vertices);
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "tests/golden";
  const uint32_t INF = cgp::infinity;
  try {
    // Toy graph: 0->1 (2), 0->2 (5), 1->3 (1).
    const std::vector<int32_t> off = {0, 2, 3, 3, 3}, idx = {1, 2, 3};
    const std::vector<uint32_t> w = {2, 5, 1};
    cgp::graph t1(4, 3, off.data(), idx.data(), w.data());
    cgp::context c1(t1);
    auto r = cgp::run_heterogeneous(
        t1, {{0, cgp::algo_kind_t::bfs}, {0, cgp::algo_kind_t::sssp}, {0, cgp::algo_kind_t::wcc}}, c1);
    std::printf("validate_hybrid (iterations %d)\n", r.iterations);
    check("BFS  [0,1,1,2]", row(r, 0) == std::vector<uint32_t>({0, 1, 1, 2}));
    check("SSSP [0,2,5,3]", row(r, 1) == std::vector<uint32_t>({0, 2, 5, 3}));
    check("WCC  [0,0,0,0]", row(r, 2) == std::vector<uint32_t>({0, 0, 0, 0}));

    auto uw = cgp::host_csr::load_matrix_market(dir + "/unweighted.mm");
    cgp::graph gu(uw);
    cgp::context cu(gu);
    auto rb = cgp::bfs::run(gu, {0, 1, 2, 3}, cu);
    std::printf("validate_bfs (unweighted.mm)\n");
    check("src 0", row(rb, 0) == std::vector<uint32_t>({0, 1, 1, 2}));
    check("src 1", row(rb, 1) == std::vector<uint32_t>({INF, 0, INF, 1}));
    check("src 2", row(rb, 2) == std::vector<uint32_t>({INF, 1, 0, 2}));
    check("src 3", row(rb, 3) == std::vector<uint32_t>({INF, INF, INF, 0}));

    auto ww = cgp::host_csr::load_matrix_market(dir + "/weighted.mm");
    cgp::graph gw(ww);
    cgp::context cw(gw);
    auto rs = cgp::sssp::run(gw, {0, 1, 2, 3}, cw);
    std::printf("validate_sssp (weighted.mm)\n");
    check("src 0", row(rs, 0) == std::vector<uint32_t>({0, 1, 2, 3}));
    check("src 1", row(rs, 1) == std::vector<uint32_t>({INF, 0, 1, 2}));
    check("src 2", row(rs, 2) == std::vector<uint32_t>({INF, INF, 0, 1}));
    check("src 3", row(rs, 3) == std::vector<uint32_t>({INF, INF, INF, 0}));

    // validate_replenish: T1 with w(1->3) = 3, two slots, four queries.
    const std::vector<uint32_t> w2 = {2, 5, 3};
    cgp::graph t2(4, 3, off.data(), idx.data(), w2.data());
    cgp::context c2(t2);
    auto rr = cgp::run_replenish_pipeline(t2,
                                          {{0, cgp::algo_kind_t::bfs},
                                           {0, cgp::algo_kind_t::sssp},
                                           {1, cgp::algo_kind_t::bfs},
                                           {1, cgp::algo_kind_t::sssp}},
                                          c2, 2);
    std::printf("validate_replenish (2 slots)\n");
    check("q0 BFS  [0,1,1,2]", row(rr, 0) == std::vector<uint32_t>({0, 1, 1, 2}));
    check("q1 SSSP [0,2,5,5]", row(rr, 1) == std::vector<uint32_t>({0, 2, 5, 5}));
    check("q2 BFS  [INF,0,INF,1]", row(rr, 2) == std::vector<uint32_t>({INF, 0, INF, 1}));
    check("q3 SSSP [INF,0,INF,3]", row(rr, 3) == std::vector<uint32_t>({INF, 0, INF, 3}));

    // Per-call options (smoke_hybrid_push.cu:151-160): one level on T1, then
    // the same context runs to convergence again.
    cgp::run_options one;
    one.max_iterations = 1;
    auto r1 = cgp::run_heterogeneous(t1, {{0, cgp::algo_kind_t::bfs}, {0, cgp::algo_kind_t::sssp}}, c1, one);
    std::printf("per-call options (max_iterations 1, then 0)\n");
    check("BFS  one level [0,1,1,INF]", row(r1, 0) == std::vector<uint32_t>({0, 1, 1, INF}));
    check("SSSP one level [0,2,5,INF]", row(r1, 1) == std::vector<uint32_t>({0, 2, 5, INF}));
    auto r2 = cgp::bfs::run(t1, {0}, c1);
    check("BFS  converged [0,1,1,2]", row(r2, 0) == std::vector<uint32_t>({0, 1, 1, 2}));

    // A graph over device arrays (no copy), results into caller device memory.
    {
      cgp::device_values d_off(off.size()), d_idx(idx.size()), d_w(w.size()), d_out(8);
      check("device CSR upload",
            cgp_copy_to_device(d_off.data(), off.data(), off.size() * 4) == CGP_OK &&
                cgp_copy_to_device(d_idx.data(), idx.data(), idx.size() * 4) == CGP_OK &&
                cgp_copy_to_device(d_w.data(), w.data(), w.size() * 4) == CGP_OK);
      cgp::graph gd(cgp::graph::device_arrays, 4, 3, (const int32_t*)d_off.data(), (const int32_t*)d_idx.data(),
                    d_w.data());
      cgp::context cd(gd);
      cgp::run_heterogeneous_into(gd, {{0, cgp::algo_kind_t::bfs}, {0, cgp::algo_kind_t::sssp}}, cd, d_out.data());
      std::printf("device-array graph, device output\n");
      auto h = d_out.to_host();
      check("BFS  [0,1,1,2]", std::vector<uint32_t>(h.begin(), h.begin() + 4) == std::vector<uint32_t>({0, 1, 1, 2}));
      check("SSSP [0,2,5,3]", std::vector<uint32_t>(h.begin() + 4, h.end()) == std::vector<uint32_t>({0, 2, 5, 3}));
    }

    // Error behaviour: an empty batch is std::invalid_argument.
    bool threw = false;
    try {
      cgp::bfs::run(t1, {}, c1);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check("empty batch throws std::invalid_argument", threw);
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 2;
  }
  if (failures) {
    std::printf("facade_check: %d FAILURES\n", failures);
    return 1;
  }
  std::printf("facade_check: ALL PASS\n");
  return 0;
}
