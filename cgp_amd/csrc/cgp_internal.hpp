// cgp_internal.hpp — host-side types shared by the generators, the engine
// and the C-ABI layer.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cgp/cgp.h"

namespace cgpi {

// Error carried across the C++ layers; the C-ABI turns it into a status and
// a thread-local message.
struct Error : std::runtime_error {
  cgp_status status;
  Error(cgp_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(cgp_status s, const std::string& m) { throw Error(s, m); }

void set_last_error(const std::string& m);

// Host CSR: int32 offsets (E < 2^31), int32 column indices, optional uint32 weights.
struct HostCSR {
  int32_t V = 0;
  int64_t E = 0;
  std::vector<int32_t> off;  // V+1
  std::vector<int32_t> idx;  // E
  std::vector<uint32_t> wt;  // E or empty (unit weights)
};

// Worker threads for host-side graph construction: CGP_THREADS, else
// OMP_NUM_THREADS, else hardware concurrency.
int host_threads();

// Runs fn(begin, end, thread_id) over [0, n) split into contiguous chunks.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t, int)>& fn);

// Counter-based hashing (SplitMix64 finaliser, Steele/Lea/Flood 2014): the
// value depends only on its inputs, so generation is deterministic for any
// thread count.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash3(uint64_t a, uint64_t b, uint64_t c) {
  return mix64(mix64(mix64(a) ^ b) ^ c);
}

// Symmetric edge weight in [wmin, wmax] from (min(u,v), max(u,v), seed).
inline uint32_t edge_weight(int32_t u, int32_t v, uint64_t seed, uint32_t wmin, uint32_t wmax) {
  uint64_t lo = (uint64_t)std::min(u, v), hi = (uint64_t)std::max(u, v);
  uint64_t h = hash3(seed ^ 0x5757575757575757ull, lo, hi);
  uint64_t span = (uint64_t)wmax - wmin + 1;
  return wmin + (uint32_t)(h % span);
}

HostCSR gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 uint32_t wmin, uint32_t wmax);
HostCSR gen_grid(int rows, int cols, double p_extra, uint64_t seed, uint32_t wmin, uint32_t wmax);
HostCSR csr_from_coo(int32_t V, int64_t E, const int32_t* src, const int32_t* dst,
                     const uint32_t* w);
// flags: CGP_LOAD_UNIT_WEIGHTS, CGP_LOAD_ROUND_WEIGHTS (include/cgp/cgp.h).
HostCSR load_gr(const std::string& path, int32_t flags = 0);
HostCSR load_matrix_market(const std::string& path, int32_t flags = 0);
HostCSR load_csr_directory(const std::string& path, int32_t flags = 0);
std::vector<int32_t> sample_sources(const HostCSR& g, int32_t n, uint64_t seed);
bool is_symmetric(const HostCSR& g);
// Every edge u -> v of weight w has a reverse edge v -> u of weight w
// (structure symmetric and weighted; checked per edge).
bool weights_symmetric(const HostCSR& g);
// The graph with its vertices renumbered by descending out-degree (ties in id
// order): perm[new] = old, inv[old] = new; each row's targets renumbered and
// sorted (weights follow their edges).
HostCSR relabel_by_degree(const HostCSR& g, std::vector<int32_t>& perm, std::vector<int32_t>& inv);

}  // namespace cgpi

// Body wrapper of every C-ABI entry point: exceptions become a status plus
// the thread-local last-error message.
#define CGP_TRY(expr)                                  \
  try {                                                \
    expr;                                              \
    return CGP_OK;                                     \
  } catch (const cgpi::Error& e_) {                    \
    cgpi::set_last_error(e_.what());                   \
    return e_.status;                                  \
  } catch (const std::bad_alloc&) {                    \
    cgpi::set_last_error("host allocation failed");    \
    return CGP_ENOMEM;                                 \
  } catch (const std::exception& e_) {                 \
    cgpi::set_last_error(e_.what());                   \
    return CGP_EINVAL;                                 \
  }

struct cgp_host_csr {
  cgpi::HostCSR g;
};
