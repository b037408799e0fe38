// kernels.hpp — device-side data layout and HIP kernels of the concurrent
// multi-query traversal engine (gfx950, wave64).
//
// Layout in HBM (V vertices, W mask words per vertex, R lanes per value row):
//   off[V+1], idx[E] int32; wt[E] uint32 (absent = unit weights); rows[E] int32
//     (source vertex of each CSR slot, built once per graph for edge-parallel WCC)
//   visited + next  uint64[V*2W]       one record per vertex: visited word w at
//                                      v*2W+w, next-frontier word w at v*2W+W+w
//                                      (an edge's BFS test reads one record)
//   front  uint64[V*W]                 frontier word w of vertex v at v*W+w
//   vals  uint32[W*V*R]                value of slot (w, q) at vertex v at
//                                      ((w*V + v) << log2 R) + q: the R query
//                                      values of one word of one vertex are one
//                                      contiguous row, so a wave whose lanes are
//                                      queries touches one 256 B row per edge.
//   mark  uint32[V]                    iteration stamp: vertex already appended
//                                      to the touched list this iteration
//   touched int32[V]                   vertices whose next words gained bits
//   small/medium int32[V], large uint64[V+E/LCHUNK+1], dense and dense_n uint64[W*(V+E/DCHUNK+1)]
//                                      degree-binned frontier work lists,
//                                      double-buffered by iteration parity
//
// One iteration = one advance launch (push lane-per-edge over the three
// degree bins for BFS bits and sparse SSSP words, lane-per-query waves for
// dense SSSP words, and on pull iterations a bottom-up BFS pass over all
// vertices), then the filter launch that claims next words,
// settles BFS depths, and emits the next binned frontier with wave-level
// prefix sums.  Every kernel reads its work counts from the device control
// block, so a sequence of iterations can be captured in a hipGraph and
// replayed without host round trips; the filter's last workgroup advances the
// iteration and raises `finished` when the frontier is empty.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cgpk {

using u32 = uint32_t;
using u64 = unsigned long long;

constexpr int MAXW = 8;          // mask words per pass (512 slots)
constexpr int MAXPLANES = 32;    // BFS depth bit planes (depths below 2^32)
constexpr u32 INF = 0xFFFFFFFFu;
constexpr int BLOCK = 256;       // workgroup size of every kernel (4 waves)
constexpr int LCHUNK = 4096;     // edges per large-bin workgroup item
// Edges per dense lane-per-query wave item.  A wave claims 64 items at a
// time, so the chunk bounds the work of the last claims of a launch: on
// RMAT-22 (80 SSSP) 1024 / 512 / 256 / 128 / 64 edges measured 37.2 / 32.1 /
// 29.6 / 28.8 / 30.0 ms of lane-per-query time per batch.
constexpr int DCHUNK = 128;
constexpr int NPROF = 18;
constexpr int PULL_WAVE_DEG = 64;  // pull: vertices above this degree get a whole wave
constexpr int PCHUNK = 256;        // edges per SSSP pull item
// Work lists are split into NSHARD contiguous shards, one per XCD of the
// MI355X (workgroups are placed on the XCDs round-robin), each with its own
// cursor: a wave takes work from its own XCD's shard first, so the shard's
// CSR rows and source rows stay in that XCD's L2, and the cursors' atomics
// spread over NSHARD lines instead of one.
constexpr int NSHARD = 8;
enum : int { SH_M = 0, SH_S, SH_D, SH_N, SH_P };

// Per-batch slot configuration (device resident, rewritten per pass).
struct Cfg {
  u64 bfs_bits[MAXW];   // slots settled by first-write BFS (visited bitmask)
  u64 val_bits[MAXW];   // slots relaxed by atomicMin on values (SSSP)
  int32_t start[MAXW * 64];  // iteration at which each slot's query started
  int32_t rshift;       // log2(R)
  int32_t W;            // words in use
  int32_t pull_policy;  // 0 push only, 1 BFS pull whenever BFS bits are active, 2 thresholds
  int32_t dense_t;      // popcount at which an SSSP word goes lane-per-query
  u32 delta;            // near-far step: thr moves to far_min + delta when no pair is below
                        // it (0: no postponing)
  u32 narrow_words;     // bit w: every SSSP slot of word w is in lanes 0..15 (and R >= 16), so
                        // its dense items relax four edges per wave step
  u64 pull_pairs_thr;   // hybrid: pull when BFS frontier (vertex,query) pairs >= this
  u64 pull_edges_thr;   // ... or when their edge-weighted count >= this
  u64 push_pairs_thr;   // back to push when pairs fall below this
  u64 dense_fedges_thr; // dense frontier mode when the next frontier's edges reach this
  int32_t sssp_pull_policy;  // SSSP words: 0 push only, 1 pull whenever SSSP bits are active,
                             // 2 thresholds (min-plus gather over every vertex's edges)
  u64 spull_pairs_thr;  // pull SSSP words when the frontier's SSSP (vertex, query) pairs reach this
  u64 spull_edges_thr;  // ... or their degree-weighted count reaches this
  u64 spush_pairs_thr;  // back to push when the pairs fall below this
  // Batch runs (every slot starts at iteration 0, so a compaction's new BFS
  // bits all get the same depth): depths are kept as bit planes, bit q of
  // plane b's word for (w, v) = bit b of the depth of slot 64w+q at vertex v,
  // written with one OR per set bit of the depth and per word with new bits.
  // The first four planes of a (word, vertex) share one 32-byte block, so
  // reading a vertex's depths (out of vertex order) touches one sector.
  // Planes below plane_count exist and start at zero; more are added by the
  // host before the depths need them.  0: depths go into the value rows.
  int32_t depth_planes;
  u64* plane4;             // planes 0-3: plane4[((w * V + v) << 2) + b]
  u64* plane[MAXPLANES];   // planes 4..: plane[b][w * V + v]
};

// A device counter on a 256-byte line of its own: the work cursors and
// reservation counters that every wave or workgroup hits with atomics share
// no line with each other or with the fields every kernel reads (atomics on
// one line serialise in one L2 channel, and reads of that line queue behind
// them).
struct alignas(256) Ctr {
  u32 v;
};

// Device control block: counters and the iteration state machine.
struct Ctl {
  u32 iter;         // iteration of the current run (0xFFFFFFFF during seeding)
  u32 stamp_base;   // mark stamp of iteration i is stamp_base + i + 1
  u32 finished;     // set when the next frontier is empty or max_iter reached
  u32 bfs_pull;     // BFS bits use the pull kernel this iteration
  u32 sssp_pull;    // SSSP words are gathered (pulled) by every vertex this iteration
  u32 max_iter;     // 0 = unlimited
  u32 push_iters;
  u32 pull_iters;
  u32 dense_mode;   // 1: advance uses fire-and-forget atomics, filter scans all vertices
  u32 dense_iters;
  u32 thr;          // SSSP near-far threshold: frontier pairs with value >= thr are postponed
                    // (INF: every pair expands, Bellman-Ford frontier)
  u32 far_min;      // min value of the pairs postponed or improved to >= thr this iteration
  u32 near_new;     // nonzero when this iteration produced an SSSP pair below thr
  u32 bumps;        // threshold raises so far
  u32 pad;
  u64 bfs_pairs[2];           // BFS frontier (vertex, query) pairs, by parity
  u64 bfs_edges[2];           // ... weighted by degree
  u64 fedges[2];              // summed degree of the next frontier's vertices
  u64 union_[MAXW];           // OR of next words in the running filter
  u64 last_union[MAXW];       // OR of the frontier words the last filter produced
  u64 prof[NPROF];            // work counters (profile mode)
  // Device-clock launch timing (wall_clock64 ticks): first workgroup start
  // and last workgroup end of the running advance / filter, and the sums
  // over the launches of the run that did work.
  u64 adv_t0, adv_t1, flt_t0, den_t0, den_t1;
  u64 adv_ticks, flt_ticks, adv_launches, den_ticks, den_launches;
  // Frontier statistics of the filter being run (all kinds): (vertex, query)
  // pairs, distinct vertices and query-weighted degree of the next frontier;
  // the seeding filter's values are kept as the sizes of frontier 0.
  u64 npairs[2], nverts[2], nqedges[2];
  u64 seed_pairs, seed_verts, seed_edges, seed_qedges;
  u64 t_start;  // device clock at the end of the seeding filter
  // Counters hit by atomics from every wave or workgroup, each on its own
  // line (the host never reads them: it copies back only the fields above).
  Ctr tcount;                 // touched-list length
  Ctr done_blocks;            // filter last-workgroup detection
  Ctr bcount[2][5];           // [parity][small, medium, large items, dense items, narrow dense items]
  Ctr fetch[2][7];            // [parity] work cursors of the advance without shards (L, D, unused,
                              // unused, pull-high, unused, unused)
  Ctr shard[2][5][NSHARD];    // [parity][M, S, unused, N, SSSP pull] per-shard work cursors
};
// Bytes of Ctl the host reads back after a run of iterations.
constexpr size_t CTL_HOST_BYTES = offsetof(Ctl, tcount);

// Per-iteration record, written by the filter's last workgroup into a ring
// of ITER_RING entries (index = iteration % ITER_RING), no host sync.
constexpr int ITER_RING = 65536;
struct IterRec {
  u32 iter;
  u32 pull;            // bit 0: BFS bits went bottom-up; bit 1: SSSP words were pulled
  u32 dense;           // dense frontier mode (fire-and-forget atomics, full filter scan)
  u32 frontier_items;  // work items of the frontier the advance consumed
  u32 touched;         // vertices the filter visited (V in dense mode)
  u32 next_items;      // work items of the frontier the filter produced
  u64 next_edges;      // summed degree of that frontier
  u64 next_bfs_pairs;  // its BFS (vertex, query) pairs
  u64 adv_ticks;       // advance duration (both launches), device wall-clock ticks
  u64 flt_ticks;       // filter duration
  u64 den_ticks;       // dense-SSSP advance launch duration (0 when it had no work)
  u64 next_pairs;      // (vertex, query) pairs of the next frontier, every kind
  u64 next_vertices;   // distinct vertices of the next frontier
  u64 next_qedges;     // sum over next-frontier pairs of the vertex's degree
  u64 edges_scanned;   // edges scanned by the advances so far (profile counters; 0 without)
  u64 t_end;           // device clock at this filter's end
  u64 fr_union[MAXW];  // OR of the frontier words this iteration's advance consumed
};

// Static per-engine kernel parameters (buffer addresses and tuning).
struct KP {
  int32_t V;
  int32_t profile;
  int64_t E;
  const int32_t* off;
  const int32_t* idx;
  const u32* wt;
  const int32_t* rows;
  const int32_t* pull_lo;   // vertices with 1 <= degree <= PULL_WAVE_DEG (lane per vertex)
  const int32_t* pull_hi;   // vertices with degree > PULL_WAVE_DEG (wave per vertex)
  const u64* pull_items;    // SSSP pull: (vertex << 32 | PCHUNK-edge chunk) of every vertex with edges
  int64_t n_pull_items;
  int32_t n_pull_lo;
  int32_t n_pull_hi;
  u64* visited;
  u64* next;
  u64* front;
  u32* mark;
  u32* vals;
  int32_t* touched;
  int32_t* small[2];
  int32_t* medium[2];
  u64* large[2];
  u64* dense[2];
  u64* dense_n[2];  // lane-per-query items of narrow words (SSSP slots all in lanes 0..15)
  Ctl* ctl;
  const Cfg* cfg;
  IterRec* iters;
  int32_t bin_small;
  int32_t bin_large;
  // Vertices a full compaction scan visits: [0, scan_v).  On a symmetric
  // graph stored in degree order the vertices without edges come last, and
  // only seeding (a touched-list scan) can give them frontier bits.
  int32_t scan_v;
  // Nonzero on skewed graphs: phase S groups mix entries from across the
  // small-degree list instead of taking 64 consecutive ones.
  int32_t mix_small;
};

// Profile counter slots.
enum : int {
  PF_PUSH_ITEMS = 0,   // frontier items expanded by push (vertex or chunk)
  PF_PUSH_EDGES,       // edges scanned lane-per-edge
  PF_PUSH_WEDGES,      // ... of which read a weight
  PF_BFS_WORDS,        // (edge, active BFS word) visited reads
  PF_SSSP_PAIRS,       // (edge, active SSSP query) value reads
  PF_NEW_WORDS,        // next words that gained bits (atomicOr that changed them)
  PF_IMPROVED,         // (vertex, query) values lowered
  PF_DENSE_ITEMS,      // dense (vertex, word, chunk) items
  PF_DENSE_EDGES,      // edges scanned lane-per-query
  PF_PULL_VERTS,       // vertices examined by pull
  PF_PULL_WORDS,       // (edge, word) visited reads by pull
  PF_PULL_EDGES,       // edges scanned by pull
  PF_DENSE_PAIRS,      // (edge, active query) pairs of the dense launch
  PF_DENSE_IMP,        // values lowered by the dense launch
  PF_DENSE_NEW,        // next words gaining bits in the dense launch
  PF_SSSP_WORDS,       // (edge, mask word with active sparse SSSP bits) pairs of the push launch
  PF_SSSP_SRC,         // (frontier item, active sparse SSSP query) pairs of the push launch
  PF_DENSE_SRC         // (dense item, active SSSP query) pairs
};

}  // namespace cgpk
