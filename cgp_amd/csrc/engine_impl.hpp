// engine_impl.hpp — host-side engine state shared by engine.hip (batch
// runs) and serve.hip (continuous batching).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cgp_internal.hpp"
#include "kernels.hpp"

using cgpk::Cfg;
using cgpk::Ctl;
using cgpk::KP;
using cgpk::u32;
using cgpk::u64;
using cgpk::BLOCK;
using cgpk::DCHUNK;
using cgpk::LCHUNK;
using cgpk::MAXW;
using cgpk::INF;

#define HIP_CHECK(x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      cgpi::fail(e_ == hipErrorOutOfMemory ? CGP_ENOMEM : CGP_EDEVICE,                \
                 std::string(#x) + ": " + hipGetErrorString(e_));                     \
  } while (0)

struct cgp_graph {
  int32_t V = 0;
  int64_t E = 0;
  bool symmetric = false;
  bool weighted = false;
  bool skewed = false;  // largest degree at least 32x the mean
  bool borrowed = false;  // off / idx / wt belong to the caller (cgp_graph_create_csr_device)
  bool sym_weights = false;  // symmetric, and every edge's reverse has the same weight (pull gathers)
  double mean_weight = 1.0;  // over all edges (1 when unweighted)
  int32_t* off = nullptr;
  int32_t* idx = nullptr;
  int32_t* rows = nullptr;
  u32* wt = nullptr;
  // Static degree split of the vertices with out-edges, for bottom-up steps.
  int32_t* pull_lo = nullptr;
  int32_t* pull_hi = nullptr;
  int32_t n_pull_lo = 0;
  int32_t n_pull_hi = 0;
  // Graphs built from host arrays are stored with their vertices renumbered
  // by descending degree: perm[internal] = caller id, inv[caller id] =
  // internal (device arrays; null for graphs over caller device arrays, which
  // keep the caller's numbering), and inv on the host for query sources.
  int32_t* perm = nullptr;
  int32_t* inv = nullptr;
  std::vector<int32_t> h_inv;
  int32_t internal(int32_t v) const { return h_inv.empty() ? v : h_inv[(size_t)v]; }
  // Vertices a full compaction scan visits: all of them, or on a symmetric
  // graph stored in degree order those with edges (a prefix).
  int32_t scan_v = 0;
};

namespace cgpi {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;  // a view into memory the buffer does not own (never freed or grown)
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  // Grows to at least `need` bytes; returns true when the address changed.
  bool ensure(size_t need) {
    if (need <= bytes) return false;
    if (borrowed) cgpi::fail(CGP_ECAPACITY, "internal: borrowed device buffer too small");
    release();
    if (need == 0) return false;
    HIP_CHECK(hipMalloc(&p, need));
    bytes = need;
    return true;
  }
};

inline int device_cus() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  int cus = 0;
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus > 0 ? cus : 1;
}

inline void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    cgpi::fail(CGP_EDEVICE, "no usable HIP device (the engine needs a gfx950 GPU)");
}

inline double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

inline int round_words(int w) { return w <= 1 ? 1 : w <= 2 ? 2 : w <= 4 ? 4 : 8; }

}  // namespace cgpi

struct ServeState;

struct cgp_engine {
  const cgp_graph* g = nullptr;
  cgp_engine_opts opts{};
  int dev = 0;  // HIP device current at creation (the engine's streams live there)
  // Host-side query lists of a batch run, reserved at creation so that runs of
  // up to CGP_STATIC_MAX_QUERIES queries do not allocate host memory.
  std::vector<int32_t> q_bfs, q_sssp, q_wcc, q_order;
  hipStream_t s = nullptr;
  hipStream_t s2 = nullptr;  // second stream of a concurrent advance (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int cus = 1;
  int wcap = 0;            // words allocated in visited/next/front/dense
  cgpi::DevBuf visited, front, mark, vals, touched, small[2], medium[2], large[2], dense[2], dense_n[2];  // visited: visited + next records
  cgpi::DevBuf iters;  // per-iteration record ring (ITER_RING entries)
  std::vector<cgpk::IterRec> report;  // records of the last batch run, all passes
  // The rest of the last batch run's report: its queries, each query's slot
  // (pass << 16 | word << 8 | bit, -1 for WCC), and per pass the sizes of
  // frontier 0 and the device-clock origin.
  struct PassSeed {
    uint64_t pairs, verts, edges, qedges, t_start;
  };
  std::vector<cgp_query> rep_queries;
  std::vector<int32_t> rep_slot;
  std::vector<PassSeed> rep_seed;
  double rep_wall_ms = 0;
  cgpi::DevBuf parent, wcc_rows, wcc_probe, outbuf, qmap, slot_src;
  // BFS depth bit planes of the running pass (batch runs): plane_count
  // planes zeroed over the pass's BFS words (plane_words bit w).
  cgpi::DevBuf plane4;                  // planes 0-3 in (word, vertex) blocks
  cgpi::DevBuf plane[cgpk::MAXPLANES];  // planes 4..
  int plane_count = 0;
  u32 plane_words = 0;
  cgpi::DevBuf pull_items;  // SSSP pull items of the graph, built on first use
  int64_t n_pull_items = 0;
  Ctl* d_ctl = nullptr;
  Cfg* d_cfg = nullptr;
  Ctl* h_ctl = nullptr;     // pinned
  Cfg* h_cfg = nullptr;     // pinned
  int32_t* h_stage = nullptr;  // pinned staging (slot sources, qmap, wcc rows)
  size_t h_stage_cap = 0;
  // Pinned landing area of the passes' iteration records, decoded into the
  // report after the batch's final synchronisation.
  cgpk::IterRec* h_rings = nullptr;
  size_t h_rings_cap = 0;
  u32 stamp_epoch = 1;
  // Narrow words of the configuration the last launched filter ran with: its
  // narrow-word items wait for the next advance's narrow kernel.
  u32 narrow_binned = 0;
  // Replays the engine's last pass took to finish: the next pass enqueues its
  // output behind that replay, before reading the control block back.
  int spec_replay = 0;
  bool next_dirty = false;
  bool busy = false;
  std::mutex mu;
  // Captured iteration graphs, one per mask-word width and set of optional
  // launches (so batches of different kinds on one engine, as in one query
  // at a time, reuse their graphs instead of capturing again).
  static constexpr int NLS = 8;  // launch sets: lane-per-query | narrow << 1 | SSSP pull << 2
  struct GExec {
    hipGraphExec_t exec = nullptr;
    int K = 0;
    int prof = -1;
  };
  GExec gx[4][NLS];
  double clock_khz = 100000.0;  // device wall-clock rate (wall_clock64 ticks)
  // Per-phase timing of init / WCC / output (launch timing on, outside graphs).
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  bool timed() const { return (opts.profile & 1) != 0; }
  bool counted() const { return (opts.profile & 2) != 0; }
  // Continuous-batching session (serve.hip), null outside serving.
  ServeState* serve = nullptr;
  // Pinned host result buffers (V words each) recycled across sessions;
  // guarded by the serving module's pool mutex.
  std::vector<uint32_t*> result_pool;

  KP kp() const {
    KP P{};
    P.V = g->V;
    P.E = g->E;
    P.profile = (opts.profile & 2) ? 1 : 0;
    P.off = g->off;
    P.idx = g->idx;
    P.wt = g->weighted ? g->wt : nullptr;
    P.rows = g->rows;
    P.pull_lo = g->pull_lo;
    P.pull_hi = g->pull_hi;
    P.n_pull_lo = g->n_pull_lo;
    P.n_pull_hi = g->n_pull_hi;
    P.pull_items = (const u64*)pull_items.p;
    P.n_pull_items = n_pull_items;
    // One buffer of per-vertex records: W visited words then W next words.
    P.visited = (u64*)visited.p;
    P.next = (u64*)visited.p;
    P.front = (u64*)front.p;
    P.mark = (u32*)mark.p;
    P.vals = (u32*)vals.p;
    P.touched = (int32_t*)touched.p;
    for (int k = 0; k < 2; ++k) {
      P.small[k] = (int32_t*)small[k].p;
      P.medium[k] = (int32_t*)medium[k].p;
      P.large[k] = (u64*)large[k].p;
      P.dense[k] = (u64*)dense[k].p;
      P.dense_n[k] = (u64*)dense_n[k].p;
    }
    P.ctl = d_ctl;
    P.cfg = d_cfg;
    P.iters = (cgpk::IterRec*)iters.p;
    P.bin_small = opts.bin_small;
    P.bin_large = opts.bin_large;
    // Serving seeds queries into running iterations, where a full scan must
    // reach a source without edges too.
    P.scan_v = serve ? g->V : g->scan_v;
    P.mix_small = g->skewed ? 1 : 0;
    return P;
  }

  void drop_graphs() {
    for (int k = 0; k < 4; ++k) {
      for (GExec& x : gx[k]) {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        x.exec = nullptr;
        x.K = 0;
        x.prof = -1;
      }
    }
  }

  // Grid sizes: persistent-style, sized to the device, independent of the
  // frontier so the launches can be captured once and replayed.
  // Same thread count per launch for every workgroup size.
  // Advance workgroups: `advance_blocks`, or by default one per 65536 edges
  // between one and eight per CU.  The persistent advance launches fetch
  // their work dynamically, so workgroups beyond what the frontier keeps
  // busy only add start-up and fetch cost (RMAT-20 64 SSSP: 9.5 ms at 2048
  // workgroups, 8.6 ms at 480; RMAT-22 and the 4096^2 grid are flat between
  // 1024 and 2048).
  int advance_blocks() const {
    if (opts.advance_blocks > 0) return opts.advance_blocks;
    const int64_t want = g->E / 65536;
    return (int)std::max<int64_t>(cus, std::min<int64_t>(want, (int64_t)cus * 8));
  }
  int grid_advance() const { return (int)((int64_t)advance_blocks() * 256 / std::max(128, opts.block_size)); }
  int grid_filter() const {
    int64_t need = ((int64_t)g->V + BLOCK - 1) / BLOCK;
    return (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)cus * 8));
  }

  void ensure_words(int W) {
    const size_t V = (size_t)g->V, E = (size_t)g->E;
    bool moved = false;
    if (W > wcap) {
      moved |= visited.ensure(V * W * 16);  // visited + next records
      moved |= front.ensure(V * W * 8);
      HIP_CHECK(hipMemsetAsync(visited.p, 0, V * W * 16, s));
      for (int k = 0; k < 2; ++k) {
        moved |= dense[k].ensure((size_t)W * (V + E / DCHUNK + 1) * 8);
        moved |= dense_n[k].ensure((size_t)W * (V + E / DCHUNK + 1) * 8);
      }
      wcap = W;
    }
    if (!mark.p) {
      moved |= mark.ensure(V * 4 + 4);
      HIP_CHECK(hipMemsetAsync(mark.p, 0, V * 4 + 4, s));
      moved |= touched.ensure(V * 4 + 4);
      for (int k = 0; k < 2; ++k) {
        moved |= small[k].ensure(V * 4 + 4);
        moved |= medium[k].ensure(V * 4 + 4);
        moved |= large[k].ensure((V + E / LCHUNK + 1) * 8);
      }
    }
    if (moved) drop_graphs();
  }

  void ensure_pull_items();  // engine.hip

  void ensure_vals(size_t elems) {
    if (vals.ensure(elems * 4)) drop_graphs();
  }

  // Capacity for `recs` iteration records; the caller has no copy into the
  // old area in flight.
  void stage_rings(size_t recs) {
    if (recs <= h_rings_cap) return;
    if (h_rings) (void)hipHostFree(h_rings);
    h_rings = nullptr;
    h_rings_cap = 0;
    HIP_CHECK(hipHostMalloc((void**)&h_rings, recs * sizeof(cgpk::IterRec), hipHostMallocDefault));
    h_rings_cap = recs;
  }
  void stage(size_t ints) {
    if (ints <= h_stage_cap) return;
    if (h_stage) (void)hipHostFree(h_stage);
    h_stage = nullptr;
    HIP_CHECK(hipHostMalloc((void**)&h_stage, ints * 4, hipHostMallocDefault));
    h_stage_cap = ints;
  }

  // ---- timing helpers (launch timing on)
  void tick(int cat, bool begin) {
    if (!timed()) return;
    if (ev_used >= ev_pool.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    hipEvent_t e = ev_pool[ev_used++];
    HIP_CHECK(hipEventRecord(e, s));
    if (begin) ev.push_back({cat, {e, nullptr}});
    else ev.back().second.second = e;
  }
};


namespace cgpi {

enum { CAT_INIT = 0, CAT_ADV = 1, CAT_FILTER = 2, CAT_WCC = 3, CAT_OUT = 4 };

// Launchers implemented in engine.hip (W in {1, 2, 4, 8}).
void dispatch_iteration(cgp_engine* en, int W, const KP& P);
void dispatch_seed_only(cgp_engine* en, int W, const KP& P, const int32_t* d_slot_src, int nslots);
void dispatch_seed(cgp_engine* en, int W, const KP& P, const int32_t* d_slot_src, int nslots);
void fill_u32(hipStream_t s, void* p, size_t n, u32 v, int cus);
// Narrow-word mask of a slot configuration (all SSSP words narrow, else 0).
u32 narrow_mask(const Cfg& cfg, int W);
// One batched run (all passes, WCC, output transpose); stats may be null.
// no_alloc: launches iterations directly (no graph capture), records no
// per-iteration report and no event timing, and grows no buffer (the caller
// has bound every per-run buffer beforehand).
void engine_run_batch(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out, bool out_dev,
                      cgp_run_stats* stats, bool no_alloc = false);
// The same run with every per-run device buffer inside the library's static
// device array (engine.hip); out is device memory.
void engine_run_batch_static(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out,
                             cgp_run_stats* stats);
// Computes min-id component labels into en->parent (hooking + pointer jumping).
void compute_wcc(cgp_engine* en);
// Writes the labels into rows rows_out[0..nrows) of out (query-major, V per row).
// On renumbered graphs `minid` (V int32) holds the per-root smallest caller
// ids; null takes the touched list's buffer, which only a finished batch
// pass may do (a serving session can hold seeds there between steps).
void write_wcc_rows(cgp_engine* en, const int32_t* d_rows, int nrows, u32* out, int32_t* minid = nullptr);
// Serving teardown for cgp_engine_destroy (serve.hip).
void serve_release_engine(cgp_engine* en);

}  // namespace cgpi
