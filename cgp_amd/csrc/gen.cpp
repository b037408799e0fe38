// gen.cpp — host graph construction: deterministic RMAT and grid generators,
// COO -> CSR, the reference's three graph file formats, source sampling.
//
// The reference has no generator on this path (its only RMAT code is a Ligra
// baseline outside the path); these follow the workload definitions: RMAT
// with (a,b,c,d), self-loops dropped, symmetrised, deduplicated, sorted rows;
// a 2D grid with 4-neighbour edges plus random extra edges.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <tuple>

#include "cgp_internal.hpp"

namespace cgpi {

int host_threads() {
  const char* s = std::getenv("CGP_THREADS");
  if (!s || !*s) s = std::getenv("OMP_NUM_THREADS");
  int n = s ? std::atoi(s) : 0;
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(n, 64));
}

void parallel_for(int64_t n, const std::function<void(int64_t, int64_t, int)>& fn) {
  int t = host_threads();
  if (n <= 0) return;
  if (t == 1 || n < 4096) {
    fn(0, n, 0);
    return;
  }
  t = (int)std::min<int64_t>(t, n);
  // Dynamic chunks: per-vertex work is skewed (RMAT hubs sit at low ids), so
  // threads take small ranges from a shared cursor instead of one slice each.
  const int64_t chunk = std::max<int64_t>(1024, n / ((int64_t)t * 256));
  std::atomic<int64_t> cursor{0};
  std::vector<std::thread> th;
  th.reserve(t);
  for (int i = 0; i < t; ++i) {
    th.emplace_back([&fn, &cursor, chunk, n, i] {
      for (int64_t b = cursor.fetch_add(chunk); b < n; b = cursor.fetch_add(chunk))
        fn(b, std::min(n, b + chunk), i);
    });
  }
  for (auto& x : th) x.join();
}

namespace {

// Builds a symmetric, self-loop-free, deduplicated CSR with sorted rows from
// `items` generator calls; emit(i, sink) calls sink(u, v) for each undirected
// pair of item i.  Two passes (count, then scatter) regenerate the pairs, so
// no edge list is stored; rows are sorted and deduplicated in parallel.
template <typename EmitFn>
HostCSR build_symmetric(int32_t V, int64_t items, EmitFn emit) {
  std::vector<std::atomic<int64_t>> deg(V + 1);
  for (int32_t v = 0; v <= V; ++v) deg[v].store(0, std::memory_order_relaxed);
  parallel_for(items, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i)
      emit(i, [&](int32_t u, int32_t v) {
        if (u == v) return;
        deg[u].fetch_add(1, std::memory_order_relaxed);
        deg[v].fetch_add(1, std::memory_order_relaxed);
      });
  });
  std::vector<int64_t> start(V + 1, 0);
  for (int32_t v = 0; v < V; ++v) start[v + 1] = start[v] + deg[v].load(std::memory_order_relaxed);
  const int64_t raw = start[V];
  std::vector<int32_t> col((size_t)raw);
  for (int32_t v = 0; v < V; ++v) deg[v].store(start[v], std::memory_order_relaxed);
  parallel_for(items, [&](int64_t b, int64_t e, int) {
    for (int64_t i = b; i < e; ++i)
      emit(i, [&](int32_t u, int32_t v) {
        if (u == v) return;
        col[deg[u].fetch_add(1, std::memory_order_relaxed)] = v;
        col[deg[v].fetch_add(1, std::memory_order_relaxed)] = u;
      });
  });
  // Sort + unique each row; record unique counts.
  std::vector<int64_t> uniq(V + 1, 0);
  parallel_for(V, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; ++v) {
      int32_t* r0 = col.data() + start[v];
      int32_t* r1 = col.data() + start[v + 1];
      std::sort(r0, r1);
      uniq[v + 1] = std::unique(r0, r1) - r0;
    }
  });
  for (int32_t v = 0; v < V; ++v) uniq[v + 1] += uniq[v];
  if (uniq[V] >= (int64_t)INT32_MAX) fail(CGP_EINVAL, "graph has more than 2^31-1 edges");
  HostCSR g;
  g.V = V;
  g.E = uniq[V];
  g.off.resize((size_t)V + 1);
  g.idx.resize((size_t)g.E);
  parallel_for(V, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; ++v) {
      g.off[v] = (int32_t)uniq[v];
      std::memcpy(g.idx.data() + uniq[v], col.data() + start[v],
                  sizeof(int32_t) * (size_t)(uniq[v + 1] - uniq[v]));
    }
  });
  g.off[V] = (int32_t)g.E;
  return g;
}

void assign_weights(HostCSR& g, uint64_t seed, uint32_t wmin, uint32_t wmax) {
  if (wmax == 0) return;
  if (wmin == 0 || wmin > wmax) fail(CGP_EINVAL, "weights: need 1 <= wmin <= wmax");
  g.wt.resize((size_t)g.E);
  parallel_for(g.V, [&](int64_t b, int64_t e, int) {
    for (int64_t u = b; u < e; ++u)
      for (int32_t k = g.off[u]; k < g.off[u + 1]; ++k)
        g.wt[k] = edge_weight((int32_t)u, g.idx[k], seed, wmin, wmax);
  });
}

}  // namespace

HostCSR gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 uint32_t wmin, uint32_t wmax) {
  if (scale < 1 || scale > 30) fail(CGP_EINVAL, "gen_rmat: scale must be in [1,30]");
  if (edge_factor < 1) fail(CGP_EINVAL, "gen_rmat: edge_factor must be >= 1");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) fail(CGP_EINVAL, "gen_rmat: bad a,b,c");
  const int32_t V = (int32_t)(1u << scale);
  const int64_t draws = (int64_t)edge_factor << scale;
  // Quadrant thresholds on a 32-bit uniform.
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)(a * two32), tab = (uint64_t)((a + b) * two32),
                 tabc = (uint64_t)((a + b + c) * two32);
  auto emit = [&](int64_t i, auto&& sink) {
    uint32_t u = 0, v = 0;
    uint64_t h = 0;
    for (int lvl = 0; lvl < scale; ++lvl) {
      if ((lvl & 1) == 0) h = hash3(seed, (uint64_t)i, (uint64_t)(lvl >> 1));
      uint64_t r = (lvl & 1) ? (h >> 32) : (h & 0xFFFFFFFFull);
      uint32_t bu = 0, bv = 0;
      if (r < ta) {
      } else if (r < tab) {
        bv = 1;
      } else if (r < tabc) {
        bu = 1;
      } else {
        bu = 1;
        bv = 1;
      }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    sink((int32_t)u, (int32_t)v);
  };
  HostCSR g = build_symmetric(V, draws, emit);
  assign_weights(g, seed, wmin, wmax);
  return g;
}

HostCSR gen_grid(int rows, int cols, double p_extra, uint64_t seed, uint32_t wmin, uint32_t wmax) {
  if (rows < 1 || cols < 1 || (int64_t)rows * cols >= INT32_MAX)
    fail(CGP_EINVAL, "gen_grid: bad dimensions");
  if (p_extra < 0 || p_extra > 1) fail(CGP_EINVAL, "gen_grid: p_extra must be in [0,1]");
  const int32_t V = rows * cols;
  const uint64_t tp = (uint64_t)(p_extra * 18446744073709551615.0);
  auto emit = [&](int64_t i, auto&& sink) {
    int32_t v = (int32_t)i, r = v / cols, c = v % cols;
    if (c + 1 < cols) sink(v, v + 1);       // right neighbour
    if (r + 1 < rows) sink(v, v + cols);    // lower neighbour
    if (p_extra > 0 && hash3(seed, (uint64_t)i, 1) < tp) {
      int32_t t = (int32_t)(hash3(seed, (uint64_t)i, 2) % (uint64_t)V);
      sink(v, t);                           // dropped when t == v
    }
  };
  HostCSR g = build_symmetric(V, V, emit);
  assign_weights(g, seed, wmin, wmax);
  return g;
}

HostCSR csr_from_coo(int32_t V, int64_t E, const int32_t* src, const int32_t* dst,
                     const uint32_t* w) {
  if (V < 0 || E < 0 || E >= INT32_MAX) fail(CGP_EINVAL, "csr_from_coo: bad sizes");
  std::vector<int64_t> order((size_t)E);
  for (int64_t i = 0; i < E; ++i) {
    if (src[i] < 0 || src[i] >= V || dst[i] < 0 || dst[i] >= V)
      fail(CGP_EINVAL, "csr_from_coo: vertex id out of range");
    order[i] = i;
  }
  // Sorted by (src, dst, weight) with duplicates kept, as the reference's
  // Matrix Market path sorts its tuples (puercgp/examples/mtx_loader.hxx:153).
  std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    if (src[x] != src[y]) return src[x] < src[y];
    if (dst[x] != dst[y]) return dst[x] < dst[y];
    uint32_t wx = w ? w[x] : 1u, wy = w ? w[y] : 1u;
    if (wx != wy) return wx < wy;
    return x < y;
  });
  HostCSR g;
  g.V = V;
  g.E = E;
  g.off.assign((size_t)V + 1, 0);
  g.idx.resize((size_t)E);
  if (w) g.wt.resize((size_t)E);
  for (int64_t k = 0; k < E; ++k) {
    int64_t i = order[k];
    g.off[src[i] + 1]++;
    g.idx[k] = dst[i];
    if (w) g.wt[k] = w[i];
  }
  for (int32_t v = 0; v < V; ++v) g.off[v + 1] += g.off[v];
  return g;
}

// Galois .gr v1 (cgp/examples/algorithms/bfs/cgp_bfs.cu:214-275): u64 header
// {version=1, sizeEdgeTy, V, E}, u64 row_end[V], u32 dst[E]; when sizeEdgeTy
// is 4 the uint32 edge data follows after padding dst to 8 bytes.  The
// reference ignores edge data (unit weights); this loader keeps 4-byte data
// as weights when present and nonzero-sized.
HostCSR load_gr(const std::string& path, int32_t flags) {
  std::ifstream in(path, std::ios::binary);
  if (!in) fail(CGP_EIO, "load_gr: could not open graph " + path);
  uint64_t hdr[4];
  if (!in.read(reinterpret_cast<char*>(hdr), sizeof(hdr))) fail(CGP_EIO, "load_gr: short header");
  if (hdr[0] != 1) fail(CGP_EIO, "load_gr: unsupported GGR version " + std::to_string(hdr[0]));
  if (hdr[2] > (uint64_t)INT32_MAX || hdr[3] >= (uint64_t)INT32_MAX)
    fail(CGP_EIO, "load_gr: GGR graph exceeds int vertex/edge limits");
  HostCSR g;
  g.V = (int32_t)hdr[2];
  g.E = (int64_t)hdr[3];
  std::vector<int64_t> row_end((size_t)g.V);
  if (!in.read(reinterpret_cast<char*>(row_end.data()), (std::streamsize)(8 * (size_t)g.V)))
    fail(CGP_EIO, "load_gr: failed reading GGR row offsets");
  g.off.assign((size_t)g.V + 1, 0);
  for (int32_t v = 0; v < g.V; ++v) {
    if (row_end[v] < 0 || row_end[v] > g.E || (v > 0 && row_end[v] < row_end[v - 1]))
      fail(CGP_EIO, "load_gr: GGR row offset out of range at vertex " + std::to_string(v));
    g.off[v + 1] = (int32_t)row_end[v];
  }
  if (g.off[g.V] != g.E) fail(CGP_EIO, "load_gr: GGR final row offset does not match edge count");
  g.idx.resize((size_t)g.E);
  if (!in.read(reinterpret_cast<char*>(g.idx.data()), (std::streamsize)(4 * (size_t)g.E)))
    fail(CGP_EIO, "load_gr: failed reading GGR destinations");
  for (int64_t e = 0; e < g.E; ++e)
    if (g.idx[e] < 0 || g.idx[e] >= g.V) fail(CGP_EIO, "load_gr: destination out of range");
  if (hdr[1] == 4 && !(flags & CGP_LOAD_UNIT_WEIGHTS)) {
    if (g.E & 1) in.ignore(4);
    g.wt.resize((size_t)g.E);
    if (!in.read(reinterpret_cast<char*>(g.wt.data()), (std::streamsize)(4 * (size_t)g.E)))
      g.wt.clear();  // no edge data present: unit weights
  }
  return g;
}

// Matrix Market coordinate (puercgp/examples/mtx_loader.hxx:99-187): 1-based
// ids, "pattern" means unit weights, "symmetric" adds the reverse of every
// off-diagonal entry; duplicates and self-loops are kept and the tuples are
// sorted.  V = max(rows, cols).  Weights are exact uint32 values: a
// negative, fractional or too large value is an error (the reference keeps
// float weights, which equal these integers only when they are integral)
// unless CGP_LOAD_ROUND_WEIGHTS rounds it to the nearest integer in [0,
// 2^32-1]; CGP_LOAD_UNIT_WEIGHTS ignores the values.
HostCSR load_matrix_market(const std::string& path, int32_t flags) {
  std::ifstream in(path);
  if (!in) fail(CGP_EIO, "load_matrix_market: could not open matrix: " + path);
  std::string header;
  std::getline(in, header);
  std::string low = header;
  for (auto& ch : low) ch = (char)std::tolower((unsigned char)ch);
  const bool pattern = low.find("pattern") != std::string::npos || (flags & CGP_LOAD_UNIT_WEIGHTS);
  const bool symmetric = low.find("symmetric") != std::string::npos;
  std::string line;
  do {
    if (!std::getline(in, line)) fail(CGP_EIO, "load_matrix_market: missing matrix dimensions: " + path);
  } while (line.empty() || line[0] == '%');
  int64_t rows = 0, cols = 0, entries = 0;
  {
    std::stringstream ss(line);
    ss >> rows >> cols >> entries;
  }
  if (rows < 0 || cols < 0 || entries < 0 || std::max(rows, cols) > INT32_MAX)
    fail(CGP_EIO, "load_matrix_market: bad dimensions");
  const int32_t V = (int32_t)std::max(rows, cols);
  std::vector<int32_t> s, d;
  std::vector<uint32_t> w;
  s.reserve((size_t)entries * (symmetric ? 2 : 1));
  d.reserve(s.capacity());
  w.reserve(s.capacity());
  for (int64_t i = 0; i < entries; ++i) {
    if (!std::getline(in, line)) fail(CGP_EIO, "load_matrix_market: truncated matrix: " + path);
    if (line.empty() || line[0] == '%') {
      --i;
      continue;
    }
    std::stringstream ss(line);
    int64_t a = 0, b = 0;
    double x = 1.0;
    ss >> a >> b;
    if (!pattern) ss >> x;
    --a;
    --b;
    if (a < 0 || a >= V || b < 0 || b >= V)
      fail(CGP_EIO, "load_matrix_market: source vertex out of range in: " + path);
    if (!pattern && !(flags & CGP_LOAD_ROUND_WEIGHTS) &&
        !(x >= 0.0 && x <= 4294967295.0 && x == std::floor(x)))
      fail(CGP_EIO, "load_matrix_market: weight " + std::to_string(x) + " of entry " + std::to_string(i + 1) +
                        " is not an integer in [0, 2^32-1] (CGP_LOAD_ROUND_WEIGHTS rounds it): " + path);
    const uint32_t wx = (uint32_t)std::llround(std::min(4294967295.0, std::max(0.0, x)));
    s.push_back((int32_t)a);
    d.push_back((int32_t)b);
    w.push_back(wx);
    if (symmetric && a != b) {
      s.push_back((int32_t)b);
      d.push_back((int32_t)a);
      w.push_back(wx);
    }
  }
  return csr_from_coo(V, (int64_t)s.size(), s.data(), d.data(), pattern ? nullptr : w.data());
}

// CSR directory (puercgp/examples/mtx_loader.hxx:189-260): csr_vlist.bin
// (int32[V+1]), csr_elist.bin (int32[E]), optional csr_weightlist.bin (int32[E]).
HostCSR load_csr_directory(const std::string& path, int32_t flags) {
  auto fsize = [](const std::string& p, int64_t* out) {
    struct stat st;
    if (stat(p.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return false;
    *out = (int64_t)st.st_size;
    return true;
  };
  const std::string rp = path + "/csr_vlist.bin", cp = path + "/csr_elist.bin",
                    wp = path + "/csr_weightlist.bin";
  int64_t rb = 0, cb = 0, wb = 0;
  if (!fsize(rp, &rb) || !fsize(cp, &cb))
    fail(CGP_EIO, "load_csr_directory: CSR directory must contain csr_vlist.bin and csr_elist.bin: " + path);
  if (rb % 4 || cb % 4 || rb < 4) fail(CGP_EIO, "load_csr_directory: CSR binary files must use 32-bit integers: " + path);
  HostCSR g;
  g.V = (int32_t)(rb / 4 - 1);
  g.E = cb / 4;
  g.off.resize((size_t)g.V + 1);
  g.idx.resize((size_t)g.E);
  std::ifstream r(rp, std::ios::binary), c(cp, std::ios::binary);
  if (!r.read(reinterpret_cast<char*>(g.off.data()), rb) ||
      !c.read(reinterpret_cast<char*>(g.idx.data()), cb))
    fail(CGP_EIO, "load_csr_directory: failed to read CSR binary files: " + path);
  if (g.off[0] != 0 || g.off[g.V] != g.E) fail(CGP_EIO, "load_csr_directory: offsets do not span the edge list");
  for (int32_t v = 0; v < g.V; ++v)
    if (g.off[v + 1] < g.off[v]) fail(CGP_EIO, "load_csr_directory: offsets not monotone");
  for (int64_t e = 0; e < g.E; ++e)
    if (g.idx[e] < 0 || g.idx[e] >= g.V) fail(CGP_EIO, "load_csr_directory: column index out of range");
  if (!(flags & CGP_LOAD_UNIT_WEIGHTS) && fsize(wp, &wb)) {
    if (wb != cb) fail(CGP_EIO, "load_csr_directory: csr_weightlist.bin size must equal csr_elist.bin size: " + path);
    std::vector<int32_t> wi((size_t)g.E);
    std::ifstream ws(wp, std::ios::binary);
    if (!ws.read(reinterpret_cast<char*>(wi.data()), wb))
      fail(CGP_EIO, "load_csr_directory: could not read csr_weightlist.bin: " + path);
    g.wt.resize((size_t)g.E);
    for (int64_t e = 0; e < g.E; ++e) {
      if (wi[e] < 0) fail(CGP_EIO, "load_csr_directory: negative weight");
      g.wt[e] = (uint32_t)wi[e];
    }
  }
  return g;
}

std::vector<int32_t> sample_sources(const HostCSR& g, int32_t n, uint64_t seed) {
  std::vector<int32_t> cand;
  for (int32_t v = 0; v < g.V; ++v)
    if (g.off[v + 1] > g.off[v]) cand.push_back(v);
  if (n < 0 || (size_t)n > cand.size())
    fail(CGP_EINVAL, "sample_sources: asked for more distinct sources than vertices with degree >= 1");
  std::vector<int32_t> out;
  std::vector<char> taken(cand.size(), 0);
  uint64_t k = 0;
  while ((int32_t)out.size() < n) {
    size_t i = (size_t)(hash3(seed, 0x50524345ull, k++) % cand.size());
    if (taken[i]) continue;
    taken[i] = 1;
    out.push_back(cand[i]);
  }
  return out;
}

bool is_symmetric(const HostCSR& g) {
  // Rows need not be sorted (COO input keeps order): sortedness is decided
  // once per row, sorted rows are binary-searched, the others scanned.
  std::vector<char> sorted_row((size_t)g.V);
  parallel_for(g.V, [&](int64_t b, int64_t e, int) {
    for (int64_t v = b; v < e; ++v)
      sorted_row[v] = std::is_sorted(g.idx.data() + g.off[v], g.idx.data() + g.off[v + 1]) ? 1 : 0;
  });
  std::atomic<bool> ok{true};
  parallel_for(g.V, [&](int64_t b, int64_t e, int) {
    for (int64_t u = b; u < e && ok.load(std::memory_order_relaxed); ++u)
      for (int32_t k = g.off[u]; k < g.off[u + 1]; ++k) {
        int32_t v = g.idx[k];
        const int32_t* r0 = g.idx.data() + g.off[v];
        const int32_t* r1 = g.idx.data() + g.off[v + 1];
        bool found = sorted_row[v] ? std::binary_search(r0, r1, (int32_t)u)
                                   : (std::find(r0, r1, (int32_t)u) != r1);
        if (!found) {
          ok.store(false);
          return;
        }
      }
  });
  return ok.load();
}

bool weights_symmetric(const HostCSR& g) {
  if (g.wt.empty()) return true;
  std::atomic<bool> ok{true};
  parallel_for(g.V, [&](int64_t b, int64_t e, int) {
    for (int64_t u = b; u < e && ok.load(std::memory_order_relaxed); ++u)
      for (int32_t k = g.off[u]; k < g.off[u + 1]; ++k) {
        const int32_t v = g.idx[k];
        const uint32_t w = g.wt[k];
        bool found = false;
        const int32_t* r0 = g.idx.data() + g.off[v];
        const int32_t* r1 = g.idx.data() + g.off[v + 1];
        if (std::is_sorted(r0, r1)) {
          for (const int32_t* x = std::lower_bound(r0, r1, (int32_t)u); x < r1 && *x == u && !found; ++x)
            found = g.wt[(size_t)(x - g.idx.data())] == w;
        } else {
          for (const int32_t* x = r0; x < r1 && !found; ++x) found = *x == u && g.wt[(size_t)(x - g.idx.data())] == w;
        }
        if (!found) {
          ok.store(false);
          return;
        }
      }
  });
  return ok.load();
}

HostCSR relabel_by_degree(const HostCSR& g, std::vector<int32_t>& perm, std::vector<int32_t>& inv) {
  const int32_t V = g.V;
  perm.assign((size_t)V, 0);
  inv.assign((size_t)V, 0);
  // Counting sort by descending degree, stable in id order.
  int32_t maxd = 0;
  for (int32_t v = 0; v < V; ++v) maxd = std::max(maxd, g.off[v + 1] - g.off[v]);
  std::vector<int64_t> start((size_t)maxd + 2, 0);
  for (int32_t v = 0; v < V; ++v) ++start[(size_t)(maxd - (g.off[v + 1] - g.off[v])) + 1];
  for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
  for (int32_t v = 0; v < V; ++v) {
    const int64_t i = start[(size_t)(maxd - (g.off[v + 1] - g.off[v]))]++;
    perm[(size_t)i] = v;
    inv[(size_t)v] = (int32_t)i;
  }
  HostCSR r;
  r.V = V;
  r.E = g.E;
  r.off.assign((size_t)V + 1, 0);
  for (int32_t i = 0; i < V; ++i) r.off[(size_t)i + 1] = r.off[(size_t)i] + (g.off[perm[i] + 1] - g.off[perm[i]]);
  r.idx.resize((size_t)g.E);
  const bool w = !g.wt.empty();
  if (w) r.wt.resize((size_t)g.E);
  parallel_for(V, [&](int64_t b, int64_t e, int) {
    std::vector<std::pair<int32_t, uint32_t>> row;
    for (int64_t i = b; i < e; ++i) {
      const int32_t o = perm[(size_t)i];
      const int32_t s0 = g.off[o], s1 = g.off[o + 1];
      row.resize((size_t)(s1 - s0));
      for (int32_t k = s0; k < s1; ++k) row[(size_t)(k - s0)] = {inv[(size_t)g.idx[k]], w ? g.wt[k] : 0u};
      std::sort(row.begin(), row.end());
      const int32_t d0 = r.off[(size_t)i];
      for (size_t k = 0; k < row.size(); ++k) {
        r.idx[(size_t)d0 + k] = row[k].first;
        if (w) r.wt[(size_t)d0 + k] = row[k].second;
      }
    }
  });
  return r;
}

}  // namespace cgpi
