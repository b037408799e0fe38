// capi.cpp — host-only part of the C-ABI: error reporting, generators,
// loaders, host CSR access, sampling.  No HIP calls: these entry points work
// on a host without a GPU.
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "cgp_internal.hpp"

namespace cgpi {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace cgpi

namespace {
cgp_status wrap_host(cgpi::HostCSR&& g, cgp_host_csr** out) {
  std::unique_ptr<cgp_host_csr> h(new cgp_host_csr());
  h->g = std::move(g);
  *out = h.release();
  return CGP_OK;
}
}  // namespace

extern "C" {

const char* cgp_last_error(void) { return cgpi::g_last_error.c_str(); }
int32_t cgp_abi_version(void) { return CGP_ABI_VERSION; }

cgp_status cgp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t wmin, uint32_t wmax, cgp_host_csr** out) {
  CGP_TRY({
    if (!out) cgpi::fail(CGP_EINVAL, "cgp_gen_rmat: null output");
    *out = nullptr;
    wrap_host(cgpi::gen_rmat(scale, edge_factor, a, b, c, seed, wmin, wmax), out);
  });
}

cgp_status cgp_gen_grid(int32_t rows, int32_t cols, double p_extra, uint64_t seed, uint32_t wmin,
                        uint32_t wmax, cgp_host_csr** out) {
  CGP_TRY({
    if (!out) cgpi::fail(CGP_EINVAL, "cgp_gen_grid: null output");
    *out = nullptr;
    wrap_host(cgpi::gen_grid(rows, cols, p_extra, seed, wmin, wmax), out);
  });
}

cgp_status cgp_host_csr_from_coo(int32_t V, int64_t E, const int32_t* src, const int32_t* dst,
                                 const uint32_t* w, cgp_host_csr** out) {
  CGP_TRY({
    if (!out || (E > 0 && (!src || !dst))) cgpi::fail(CGP_EINVAL, "cgp_host_csr_from_coo: null argument");
    *out = nullptr;
    wrap_host(cgpi::csr_from_coo(V, E, src, dst, w), out);
  });
}

cgp_status cgp_load_gr(const char* path, cgp_host_csr** out) { return cgp_load_graph(path, CGP_FORMAT_GR, 0, out); }

cgp_status cgp_load_matrix_market(const char* path, cgp_host_csr** out) {
  return cgp_load_graph(path, CGP_FORMAT_MATRIX_MARKET, 0, out);
}

cgp_status cgp_load_csr_directory(const char* path, cgp_host_csr** out) {
  return cgp_load_graph(path, CGP_FORMAT_CSR_DIRECTORY, 0, out);
}

cgp_status cgp_load_graph(const char* path, int32_t format, int32_t flags, cgp_host_csr** out) {
  CGP_TRY({
    if (!path || !out) cgpi::fail(CGP_EINVAL, "cgp_load_graph: null argument");
    *out = nullptr;
    if (flags & ~(CGP_LOAD_UNIT_WEIGHTS | CGP_LOAD_ROUND_WEIGHTS)) cgpi::fail(CGP_EINVAL, "cgp_load_graph: unknown flags");
    switch (format) {
      case CGP_FORMAT_GR: wrap_host(cgpi::load_gr(path, flags), out); break;
      case CGP_FORMAT_MATRIX_MARKET: wrap_host(cgpi::load_matrix_market(path, flags), out); break;
      case CGP_FORMAT_CSR_DIRECTORY: wrap_host(cgpi::load_csr_directory(path, flags), out); break;
      default: cgpi::fail(CGP_EINVAL, "cgp_load_graph: unknown format " + std::to_string(format));
    }
  });
}

cgp_status cgp_host_csr_view(const cgp_host_csr* h, int32_t* V, int64_t* E, const int32_t** off,
                             const int32_t** idx, const uint32_t** wt) {
  CGP_TRY({
    if (!h) cgpi::fail(CGP_EINVAL, "cgp_host_csr_view: null graph");
    if (V) *V = h->g.V;
    if (E) *E = h->g.E;
    if (off) *off = h->g.off.data();
    if (idx) *idx = h->g.idx.data();
    if (wt) *wt = h->g.wt.empty() ? nullptr : h->g.wt.data();
  });
}

void cgp_host_csr_destroy(cgp_host_csr* h) { delete h; }

cgp_status cgp_sample_sources(const cgp_host_csr* h, int32_t n, uint64_t seed, int32_t* out) {
  CGP_TRY({
    if (!h || (n > 0 && !out)) cgpi::fail(CGP_EINVAL, "cgp_sample_sources: null argument");
    auto s = cgpi::sample_sources(h->g, n, seed);
    if (n > 0) std::memcpy(out, s.data(), sizeof(int32_t) * (size_t)n);
  });
}

cgp_status cgp_poisson_arrivals(int32_t n, double rate, uint64_t seed, double* out) {
  CGP_TRY({
    if (n < 0 || (n > 0 && !out) || !(rate > 0)) cgpi::fail(CGP_EINVAL, "cgp_poisson_arrivals: bad argument");
    double t = 0.0;
    for (int32_t i = 0; i < n; ++i) {
      // Uniform in (0,1] from 53 hash bits; inverse-CDF exponential gap.
      const uint64_t h = cgpi::hash3(seed, 0x504F4953ull, (uint64_t)i);
      const double u = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
      t += -std::log(u) / rate;
      out[i] = t;
    }
  });
}

}  // extern "C"
