// engine.hip — HIP kernels (gfx950, wave64) and the host engine of the
// concurrent multi-query BFS / SSSP / WCC traversal.
//
// Reference behaviour followed (paths relative to the reference root):
//   push advance with per-vertex query masks
//     puercgp/include/puercgp/engine/frontier_engine.hxx:678-943 (shared-node push),
//     hybrid_engine.hxx:132-330 (mixed BFS/SSSP bits in one traversal);
//   degree-binned mapping  frontier_engine.hxx:945-1235;
//   next-frontier marking  frontier_engine.hxx:117-173;
//   relaxation semantics   core/reduce_ops.hxx:24-110 (BFS first-write with the
//                          query-relative depth source_value+1, SSSP min-reduce);
//   pull (bottom-up) BFS   frontier_engine.hxx:1415-1486, thresholds core/types.hxx:30-31;
//   host loop              frontier_engine.hxx:2210-3139, hybrid_engine.hxx:796-1140.
// WCC uses hooking plus pointer jumping instead of the reference's label
// propagation (algorithms/wcc.hxx:20-42); both yield the minimum vertex id of
// each component on symmetric graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "engine_impl.hpp"

using namespace cgpk;
using namespace cgpi;

// AMDGPU intrinsics without a clang builtin, bound by their LLVM names:
// raw buffer load and atomicMin (address = base + voffset + soffset, with
// the 32-bit offsets range-checked against the descriptor) and v_writelane.
// (Struct-addressed buffer loads, vindex * stride + voffset, measured 2.4x
// slower on random 256-byte row gathers than raw or global loads.)
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
extern "C" __device__ int cgp_buf_load_i32(i32x4 rsrc, int voffset, int soffset,
                                           int aux) __asm("llvm.amdgcn.raw.buffer.load.i32");
extern "C" __device__ int cgp_buf_atomic_umin_i32(int vdata, i32x4 rsrc, int voffset, int soffset,
                                                  int aux) __asm("llvm.amdgcn.raw.buffer.atomic.umin.i32");
extern "C" __device__ int cgp_writelane_i32(int value, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// ============================================================== device helpers
namespace {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Number of set bits of m below the calling lane.
__device__ __forceinline__ u32 rank_below(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

__device__ __forceinline__ u64 shfl64(u64 x, int src) {
  u32 lo = (u32)__shfl((int)(u32)x, src, 64);
  u32 hi = (u32)__shfl((int)(u32)(x >> 32), src, 64);
  return ((u64)hi << 32) | lo;
}

// Wave-aggregated append of one item per predicated lane: one atomicAdd per
// wave, slots handed out by the lanes' rank among the predicated lanes.
// Safe under divergence (the ballot sees the active lanes only).
template <typename T>
__device__ __forceinline__ void wave_append(bool pred, T item, u32* counter, T* list) {
  u64 m = __ballot(pred);
  if (m == 0) return;
  int leader = __ffsll((long long)m) - 1;
  u32 base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (u32)__popcll(m));
  base = (u32)__shfl((int)base, leader, 64);
  if (pred) list[base + rank_below(m)] = item;
}

// Wave-level exclusive prefix sum of per-lane counts, with one atomicAdd per
// wave reserving the total; returns the lane's first slot.  Every lane of the
// wave must call it.
__device__ __forceinline__ u32 wave_reserve(u32 cnt, u32* counter) {
  u32 x = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    u32 y = (u32)__shfl_up((int)x, d, 64);
    if (lane_id() >= d) x += y;
  }
  u32 total = (u32)__shfl((int)x, 63, 64);
  u32 base = 0;
  if (total != 0 && lane_id() == 0) base = atomicAdd(counter, total);
  base = (u32)__shfl((int)base, 0, 64);
  return base + x - cnt;
}

// Workgroup-level exclusive prefix sums of N per-thread counts, with one
// atomicAdd per non-zero total reserving the block's range in counters[0..N);
// at[k] is the thread's first slot for counter k.  Every thread of the
// BLOCK-thread workgroup must call it (uniform control flow); lds holds
// (BLOCK/64 + 1) * N words.
template <int N>
__device__ __forceinline__ void block_reserve(const u32 (&cnt)[N], Ctr* counters, u32 (&at)[N], u32* lds) {
  constexpr int NW = BLOCK / 64;
  const int wave = (int)(threadIdx.x >> 6), lane = lane_id();
  u32 incl[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u32 x = cnt[k];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const u32 y = (u32)__shfl_up((int)x, d, 64);
      if (lane >= d) x += y;
    }
    incl[k] = x;
    if (lane == 63) lds[wave * N + k] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int k = (int)threadIdx.x;
    u32 tot = 0;
#pragma unroll
    for (int wv = 0; wv < NW; ++wv) tot += lds[wv * N + k];
    lds[NW * N + k] = tot ? atomicAdd(&counters[k].v, tot) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u32 off = lds[NW * N + k];
    for (int wv = 0; wv < wave; ++wv) off += lds[wv * N + k];
    at[k] = off + incl[k] - cnt[k];
  }
  __syncthreads();
}

__device__ __forceinline__ u64 wave_sum64(u64 x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += shfl64(x, lane_id() ^ d);
  return x;
}

__device__ __forceinline__ u64 wave_or64(u64 x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x |= shfl64(x, lane_id() ^ d);
  return x;
}

// First claim of vertex v in the iteration whose stamp is `stamp`.
__device__ __forceinline__ bool claim_vertex(u32* mark, int32_t v, u32 stamp) {
  if (__hip_atomic_load(&mark[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == stamp)
    return false;
  return atomicExch(&mark[v], stamp) != stamp;
}

// A wave-uniform pointer into global memory pinned to scalar registers, so
// accesses through it use the SGPR-base + 32-bit lane-offset addressing form
// (one VGPR per access instead of a 64-bit per-lane address).
typedef __attribute__((address_space(1))) u32 gu32;
__device__ __forceinline__ gu32* uniform_gptr(const u32* p) {
  const u64 a = (u64)p;
  const u32 lo = (u32)__builtin_amdgcn_readfirstlane((int)(u32)a);
  const u32 hi = (u32)__builtin_amdgcn_readfirstlane((int)(u32)(a >> 32));
  return (gu32*)(((u64)hi << 32) | lo);
}

// Raw buffer resource over `bytes` bytes from `base` (bytes < 2^32).  The
// fields are wave-uniform so the descriptor stays in scalar registers.
__device__ __forceinline__ i32x4 raw_rsrc(u64 base, u32 bytes) {
  i32x4 r;
  r.x = __builtin_amdgcn_readfirstlane((int)(u32)base);
  r.y = __builtin_amdgcn_readfirstlane((int)((u32)(base >> 32) & 0xFFFFu));  // stride 0
  r.z = __builtin_amdgcn_readfirstlane((int)bytes);
  r.w = 0x00020000;  // 32-bit data format: untyped accesses need a valid one
  return r;
}

__device__ __forceinline__ size_t vrow(int32_t V, int w, int32_t v, int rshift) {
  return ((size_t)w * (size_t)V + (size_t)v) << rshift;
}

// Visited and next-frontier words of a vertex form one record of 2W words
// (W visited words, then W next words), so an edge that tests visited and
// next touches one record; P.visited and P.next share the base.
template <int W>
__device__ __forceinline__ size_t vis_at(int32_t v, int w) {
  return (size_t)v * (2 * W) + (size_t)w;
}
template <int W>
__device__ __forceinline__ size_t nxt_at(int32_t v, int w) {
  return (size_t)v * (2 * W) + W + (size_t)w;
}

__device__ __forceinline__ void prof_add(const KP& P, int slot, u64 x) {
  x = wave_sum64(x);
  if (lane_id() == 0 && x) atomicAdd(&P.ctl->prof[slot], x);
}

// ------------------------------------------------------------------ push
// Two frontier modes, chosen per iteration by the filter:
//  sparse  the advance appends each newly touched vertex to a touched list
//          (first touch detected with a returning atomicOr on next[] plus a
//          per-iteration mark), and the filter visits only those vertices;
//  dense   the advance issues fire-and-forget atomics only and the filter
//          scans every vertex's next words.
// SSSP improvements are detected optimistically in both: when the candidate
// beats the value just read, the (non-returning) atomicMin is issued and the
// query's next bit set.  A concurrent smaller write can only make that bit
// redundant (the vertex re-relaxes with its true value), never lose an
// improvement, because values only decrease.

// Work counters of one lane (flushed once per kernel in profile mode), and
// the near-far statistics of its SSSP improvements (flushed whenever the
// threshold is active).
struct Tally {
  u32 items = 0, edges = 0, wedges = 0, bfs = 0, nnew = 0, imp = 0;
  u32 ditems = 0, dedges = 0, pverts = 0, pwords = 0, pedges = 0;
  u64 sssp = 0;  // (edge, query) pairs: can pass 2^32 per lane
  u32 swords = 0;  // (edge, word with sparse SSSP bits) pairs
  u64 src = 0;     // (item, active SSSP query) pairs
  u32 fmin = INF;  // min value of pairs postponed or improved to >= thr
  u32 nnear = 0;   // improvements to a value < thr
};

// Near-far bookkeeping of one improvement to value c.
__device__ __forceinline__ void note_improved(Tally& t, u32 c, u32 thr) {
  if (c < thr) ++t.nnear;
  else t.fmin = min(t.fmin, c);
}

// Relaxes edge (u -> v, weight wv) for every active BFS bit (bitmask
// first-write against visited) and every sparse SSSP bit; sets the new bits
// in next[v].  Sparse mode returns whether this lane made the vertex's first
// claim of the iteration.
template <int W, bool SPARSE>
__device__ __forceinline__ bool relax_edge(const KP& P, int32_t u, int32_t v, u32 wv,
                                           const u64 (&bm)[W], const u64 (&sm)[W], u32 stamp,
                                           int rshift, u32 thr, Tally& t) {
  bool first = false;
  const size_t vb = vis_at<W>(v, 0);
  // All W visited and W next words of v in one go (one contiguous record),
  // issued before any of them is used.  In dense mode an edge with no BFS
  // bit needs neither: its improved SSSP bits go to next[v] by a
  // fire-and-forget atomicOr, so the record's line is touched only for an
  // improvement, not for every edge scanned.  (A per-vertex byte of "BFS
  // words every running query has visited", kept by the filter to skip the
  // record read too, cut the visited tests of the RMAT-22 mixed batch by
  // 39 % but not its traffic: the skipped records were cache hits, and the
  // extra dependent load made 160 BFS 0.5 ms slower.)
  bool any_b = false;
#pragma unroll
  for (int w = 0; w < W; ++w) any_b |= bm[w] != 0;
  u64 nx[W], vis[W];
  if (SPARSE || any_b) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      nx[w] = P.next[vb + W + w];
      vis[w] = P.visited[vb + w];
    }
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w) nx[w] = vis[w] = 0;
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (!(bm[w] | sm[w])) continue;
    u64 add = 0;
    if (bm[w]) {
      ++t.bfs;
      add = bm[w] & ~vis[w];
    }
    u64 s = sm[w];
    if (s) {
      ++t.swords;
      const size_t ub = vrow(P.V, w, u, rshift), vr = vrow(P.V, w, v, rshift);
      while (s) {
        // Up to 4 active queries per round, their 8 value loads in flight together.
        int q[4];
        bool a[4];
        u32 du[4], cur[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          // Unused slots of the batch re-read the first slot's values (same
          // lines), never a line no active query needs.
          a[i] = s != 0;
          q[i] = a[i] ? __ffsll((long long)s) - 1 : q[0];
          if (a[i]) s &= s - 1;
          du[i] = P.vals[ub + q[i]];
          cur[i] = P.vals[vr + q[i]];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (!a[i]) continue;
          ++t.sssp;
          const u32 cand = du[i] + wv;
          if (cand < du[i]) continue;  // saturate: never wrap past CGP_INF
          if (cand < cur[i]) {
            atomicMin(&P.vals[vr + q[i]], cand);
            add |= 1ull << q[i];
            ++t.imp;
            note_improved(t, cand, thr);
          }
        }
      }
    }
    if (add & ~nx[w]) {
      ++t.nnew;
      if (SPARSE) {
        if (atomicOr(&P.next[vb + W + w], add) == 0) first = true;
      } else {
        atomicOr(&P.next[vb + W + w], add);
      }
    }
  }
  return SPARSE && first && claim_vertex(P.mark, v, stamp);
}

// Frontier masks of vertex u: BFS bits, and the sparse SSSP bits (popcount
// below dense_t; denser words go lane-per-query) split at the near-far
// threshold: pairs whose value is >= thr land in far[] and are not expanded.
template <int W>
__device__ __forceinline__ void load_masks(const KP& P, int32_t u, bool bfs_on, int dense_t,
                                           u32 thr, int rshift, u64 (&bm)[W], u64 (&sm)[W],
                                           u64 (&far)[W], bool& any_w, Tally& t) {
  const Cfg* cfg = P.cfg;
  any_w = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u64 f = P.front[(size_t)u * W + w];
    bm[w] = bfs_on ? (f & cfg->bfs_bits[w]) : 0ull;
    u64 s = f & cfg->val_bits[w];
    if (__popcll(s) >= dense_t) s = 0;  // handled lane-per-query
    far[w] = 0;
    if (s && thr != INF) {
      const size_t ub = vrow(P.V, w, u, rshift);
      u64 x = s;
      while (x) {
        int q[4];
        bool a[4];
        u32 d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = x != 0;
          q[i] = a[i] ? __ffsll((long long)x) - 1 : q[0];
          if (a[i]) x &= x - 1;
          d[i] = P.vals[ub + q[i]];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (a[i] && d[i] >= thr) {
            far[w] |= 1ull << q[i];
            t.fmin = min(t.fmin, d[i]);
          }
        }
      }
      s &= ~far[w];
    }
    sm[w] = s;
    any_w |= (s != 0);
  }
}

// Keeps u's postponed SSSP pairs in the next frontier.  Sparse mode returns
// whether this made the vertex's first claim of the iteration.
template <int W, bool SPARSE>
__device__ __forceinline__ bool postpone(const KP& P, int32_t u, const u64 (&far)[W], u32 stamp,
                                         Tally& t) {
  bool first = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (!far[w]) continue;
    ++t.nnew;
    if (SPARSE) {
      if (atomicOr(&P.next[nxt_at<W>(u, w)], far[w]) == 0) first = true;
    } else {
      atomicOr(&P.next[nxt_at<W>(u, w)], far[w]);
    }
  }
  return SPARSE && first && claim_vertex(P.mark, u, stamp);
}

// Expands frontier vertex u over edges [e0, e1) with lanes striding by
// `stride` from `first` (lane-per-edge): BFS bits and sparse SSSP words.
template <int W, bool SPARSE>
__device__ __forceinline__ void expand_lanes(const KP& P, int32_t u, int32_t e0, int32_t e1,
                                             int32_t first, int32_t stride, bool bfs_on, int dense_t,
                                             int rshift, u32 stamp, u32 thr, bool lead, Tally& t) {
  u64 bm[W], sm[W], far[W];
  bool any_s;
  load_masks<W>(P, u, bfs_on, dense_t, thr, rshift, bm, sm, far, any_s, t);
  if (lead) {
    const bool c = postpone<W, SPARSE>(P, u, far, stamp, t);
    if (SPARSE) wave_append<int32_t>(c, u, &P.ctl->tcount.v, P.touched);
  }
  bool any_b = false;
#pragma unroll
  for (int w = 0; w < W; ++w) any_b |= (bm[w] != 0);
  if (!any_b && !any_s) return;
  if (first == e0) {
    ++t.items;
#pragma unroll
    for (int w = 0; w < W; ++w) t.src += (u64)__popcll(sm[w]);
  }
  for (int32_t e = first; e < e1; e += stride) {
    const int32_t v = P.idx[e];
    u32 wv = 1;
    if (any_s && P.wt) {
      wv = P.wt[e];
      ++t.wedges;
    }
    ++t.edges;
    const bool claimed = relax_edge<W, SPARSE>(P, u, v, wv, bm, sm, stamp, rshift, thr, t);
    if (SPARSE) wave_append<int32_t>(claimed, v, &P.ctl->tcount.v, P.touched);
  }
}

// Wave-level inclusive prefix sum (every lane of the wave calls it).
__device__ __forceinline__ u32 wave_incl_scan(u32 x) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 y = (u32)__shfl_up((int)x, d, 64);
    if (lane_id() >= d) x += y;
  }
  return x;
}

// Index of the first lane whose inclusive prefix `incl` exceeds f (lanes
// hold a non-decreasing sequence; every lane of the wave calls it).
__device__ __forceinline__ int wave_upper_bound(u32 incl, u32 f) {
  int lo = 0, hi = 63;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int mid = (lo + hi) >> 1;
    const u32 x = (u32)__shfl((int)incl, mid, 64);
    if (x > f) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Phase S: a wave takes up to 64 low-degree frontier vertices at once; the
// lanes load their metadata in parallel, a wave prefix sum over the degrees
// flattens their edge lists, and each step relaxes 64 consecutive flattened
// edges, one per lane (load-balanced lane-per-edge).  Group g of the n list
// entries (ng groups) is entries 64g..64g+63, or, on skewed graphs
// (mix_small), eight runs of eight consecutive entries, runs g, g + ng, ...,
// g + 7 ng of the list's n/8 runs.  The list follows vertex ids, and on
// skewed graphs ids follow degree (degree-ordered storage, or an RMAT
// generator's own bias), so groups of 64 near-equal degrees made single waves
// walk the longest edge lists; mixed runs of eight keep neighbouring entries'
// shared metadata lines (RMAT-22 mixed 256: 50.1 -> 47.2 ms per batch).  On
// grids and meshes consecutive ids are neighbours whose rows share cache
// lines, and mixing measured slower (4096^2 grid, 512 SSSP: push 1067 -> 1195
// ms per batch).
template <int W, bool SPARSE>
__device__ __forceinline__ void expand_flat_vertices(const KP& P, const int32_t* list, u32 g, u32 ng, u32 n,
                                                     bool bfs_on, int dense_t, int rshift, u32 stamp,
                                                     u32 thr, Tally& t) {
  const int lane = lane_id();
  const u32 e = P.mix_small ? (g + (u32)(lane >> 3) * ng) * 8u + (u32)(lane & 7) : g * 64u + (u32)lane;
  const bool have = e < n;
  const int32_t u = have ? list[e] : 0;
  u64 bm[W], sm[W], far[W];
  bool any_s = false, any_b = false;
  if (have) load_masks<W>(P, u, bfs_on, dense_t, thr, rshift, bm, sm, far, any_s, t);
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (!have) bm[w] = sm[w] = far[w] = 0;
    any_b |= bm[w] != 0;
  }
  {
    const bool c = have && postpone<W, SPARSE>(P, u, far, stamp, t);
    if (SPARSE) wave_append<int32_t>(c, u, &P.ctl->tcount.v, P.touched);
  }
  const bool work = have && (any_b || any_s);
  const int32_t e0 = work ? P.off[u] : 0;
  const u32 d = work ? (u32)(P.off[u + 1] - e0) : 0u;
  if (work) {
    ++t.items;
#pragma unroll
    for (int w = 0; w < W; ++w) t.src += (u64)__popcll(sm[w]);
  }
  const u32 incl = wave_incl_scan(d);
  // Wave-uniform (scalar) so the loops below branch on the scalar unit and
  // the per-step row indices stay in scalar registers.
  const u32 total = (u32)__builtin_amdgcn_readlane((int)incl, 63);
  for (u32 win = 0; win < total; win += 64) {
    const u32 f = win + (u32)lane;
    const int j = wave_upper_bound(incl, f);
    const int32_t uj = __shfl(u, j, 64);
    const u32 exj = (u32)__shfl((int)(incl - d), j, 64);
    const int32_t ej = __shfl(e0, j, 64) + (int32_t)(f - exj);
    u64 bj[W], sj[W];
    bool sj_any = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      bj[w] = shfl64(bm[w], j);
      sj[w] = shfl64(sm[w], j);
      sj_any |= sj[w] != 0;
    }
    bool claimed = false;
    int32_t v = 0;
    if (f < total) {
      v = P.idx[ej];
      u32 wv = 1;
      if (sj_any && P.wt) {
        wv = P.wt[ej];
        ++t.wedges;
      }
      ++t.edges;
      claimed = relax_edge<W, SPARSE>(P, uj, v, wv, bj, sj, stamp, rshift, thr, t);
    }
    if (SPARSE) wave_append<int32_t>(claimed, v, &P.ctl->tcount.v, P.touched);
  }
}

// Phase D: a wave takes up to 64 dense (vertex, word, chunk) items at once,
// flattens their edges as above, and walks them U per step with lane q as
// query q.  Every value load is unconditional (inactive lanes read a clamped
// address and discard it) so the U source rows and U target rows of a step
// are all in flight together, and the loads of step k+1 are issued before
// the atomics of step k, so waiting for them never waits on the atomics.
template <int W, bool SPARSE, int U, bool NF>
__device__ __forceinline__ void expand_flat_dense(const KP& P, u32 base, u32 n, int rshift, u32 stamp,
                                                  u32 thr, Tally& t) {
  const int lane = lane_id();
  const int R = 1 << rshift;
  const u32 p = P.ctl->iter & 1u;
  const bool have = base + (u32)lane < n;
  const u64 item = have ? P.dense[p][base + lane] : 0ull;
  const int32_t u = (int32_t)(item >> 32);
  const int w = (int)((item >> 24) & 0xFF);
  const u32 chunk = (u32)(item & 0xFFFFFF);
  const int32_t e0 = have ? P.off[u] + (int32_t)(chunk * DCHUNK) : 0;
  const int32_t e1 = have ? min(P.off[u + 1], e0 + DCHUNK) : 0;
  u64 sm = have ? (P.front[(size_t)u * W + w] & P.cfg->val_bits[w]) : 0ull;
  const bool inrow = lane < R;
  // Lane offset inside a value row; every lane reads a valid address of the
  // (always valid) row whether or not its pair is active, so the loads issue
  // before activity is known and row indices stay wave-uniform.
  const int lq = lane & (R - 1);
  // Bit w: this lane holds an SSSP slot of word w.  Other lanes read the
  // row's first value instead of their own, so a word whose SSSP slots fill
  // only part of the row fetches only the lines they occupy (SSSP slots are
  // packed from lane 0).
  u32 lane_words = 0;
#pragma unroll
  for (int ww = 0; ww < W; ++ww) lane_words |= (u32)((P.cfg->val_bits[ww] >> lq) & 1ull) << ww;
  if (!inrow) lane_words = 0;
  const u32 lq4 = (u32)lq * 4u;
  if (NF) {
    // Near-far split of the wave's items: lane q reads value q of each
    // item's source row (8 rows in flight); pairs at or above thr are
    // postponed (chunk 0 of the vertex keeps them in the next frontier) and
    // an item left with no near pair scans no edges.
    const int nit = (int)min(64u, n - base);
    u64 near = sm;
    for (int i0 = 0; i0 < nit; i0 += 8) {
      u32 dv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int it = min(i0 + k, nit - 1);
        const u32 uk = (u32)__builtin_amdgcn_readlane(u, it);
        const u32 wk = (u32)__builtin_amdgcn_readlane(w, it);
        dv[k] = uniform_gptr(P.vals + ((size_t)(wk * (u32)P.V + uk) << rshift))[(unsigned)lq];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int it = i0 + k;
        if (it >= nit) break;
        const u64 smk = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(sm >> 32), it) << 32) |
                        (u32)__builtin_amdgcn_readlane((int)(u32)sm, it);
        const bool isfar = inrow && ((smk >> lane) & 1ull) && dv[k] >= thr;
        const u64 fm = __ballot(isfar);
        if (isfar) t.fmin = min(t.fmin, dv[k]);
        if (lane == it) near = sm & ~fm;
        if (fm && __builtin_amdgcn_readlane((int)chunk, it) == 0) {
          const int32_t uk = __builtin_amdgcn_readlane(u, it);
          const int wk = __builtin_amdgcn_readlane(w, it);
          bool claimed = false;
          if (lane == 0) {
            ++t.nnew;
            if (SPARSE) {
              if (atomicOr(&P.next[nxt_at<W>(uk, wk)], fm) == 0) claimed = claim_vertex(P.mark, uk, stamp);
            } else {
              atomicOr(&P.next[nxt_at<W>(uk, wk)], fm);
            }
          }
          if (SPARSE) wave_append<int32_t>(claimed, uk, &P.ctl->tcount.v, P.touched);
        }
      }
    }
    sm = near;
  }
  const u32 d = sm ? (u32)(e1 - e0) : 0u;
  if (have) {
    ++t.ditems;
    t.dedges += d;
    t.sssp += (u64)d * (u64)__popcll(sm);
    t.src += (u64)__popcll(sm);
  }
  const u32 incl = wave_incl_scan(d);
  // Wave-uniform (scalar) so the loops below branch on the scalar unit and
  // the per-step row indices stay in scalar registers.
  const u32 total = (u32)__builtin_amdgcn_readlane((int)incl, 63);
  const u32 smlo = (u32)sm, smhi = (u32)(sm >> 32);
  const u64 item_row = (u64)((u32)w * (u32)P.V + (u32)u);  // source row of the lane's item
  const u64 row_lanes = R >= 64 ? ~0ull : ((1ull << R) - 1ull);
  const u64 vals_base = (u64)P.vals;
  const int rsh = rshift + 2;  // row index -> byte offset
  for (u32 win = 0; win < total; win += 64) {
    // Lane f takes flattened edge win+f and computes, once per window, the
    // byte addresses of its target and source rows, its weight and its
    // item's active-query mask; each step then reads them back with
    // independent uniform-index readlanes (no per-step scalar arithmetic).
    // Lanes past the end point at row 0 with an empty mask.
    const u32 f = win + (u32)lane;
    const int j = wave_upper_bound(incl, f);
    const u32 exj = (u32)__shfl((int)(incl - d), j, 64);
    const int32_t ej = __shfl(e0, j, 64) + (int32_t)(f - exj);
    const bool ok = f < total;
    const int32_t myv = ok ? P.idx[ej] : 0;
    const u32 myw = ok ? (P.wt ? P.wt[ej] : 1u) : 0u;
    const int myword = __shfl(w, j, 64);
    const u64 my_vaddr = vals_base + ((u64)((u32)myword * (u32)P.V + (u32)myv) << rsh);
    const u64 my_uaddr = vals_base + ((u64)(u32)__shfl((int)(u32)item_row, j, 64) << rsh);
    // Shuffles run with every lane active (a lane past the end may still be
    // the source of another lane's item), then the select.
    const u32 am_lo = (u32)__shfl((int)smlo, j, 64), am_hi = (u32)__shfl((int)smhi, j, 64);
    const u64 my_am = ok ? ((((u64)am_hi << 32) | am_lo) & row_lanes) : 0ull;  // read only with NF
    const int cnt = (int)min(64u, total - win);
    // Per-step state: uniform row pointers, weight and active mask; per
    // lane the two loaded values.
    u64 vr[U], am[U];
    u32 cand[U], cur[U], du[U];
    auto issue = [&](int k0) {
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const int jj = k0 + k;  // < 64: k0 < cnt <= 64 and U divides 64
        vr[k] = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(my_vaddr >> 32), jj) << 32) |
                (u32)__builtin_amdgcn_readlane((int)(u32)my_vaddr, jj);
        const u64 ua = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(my_uaddr >> 32), jj) << 32) |
                       (u32)__builtin_amdgcn_readlane((int)(u32)my_uaddr, jj);
        cand[k] = (u32)__builtin_amdgcn_readlane((int)myw, jj);
        const u32 wk = (u32)__builtin_amdgcn_readlane(myword, jj);
        const bool inw = ((lane_words >> wk) & 1u) != 0;
        if (NF) {
          am[k] = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(my_am >> 32), jj) << 32) |
                  (u32)__builtin_amdgcn_readlane((int)(u32)my_am, jj);
        } else {
          // Without postponing every SSSP lane of the word is compared: a
          // query not in u's frontier has relaxed u's edges with u's current
          // value already, so its compare succeeds only if u improved during
          // this iteration, which is a real path (its next bit is then set).
          // Steps past the window's last edge read a dummy row and are
          // masked out entirely.
          am[k] = __ballot(inw) & (k0 + k < cnt ? ~0ull : 0ull);
        }
        const u32 off = inw ? lq4 : 0u;
        du[k] = *(const gu32*)(ua + off);
        cur[k] = *(const gu32*)(vr[k] + off);
      }
    };
    issue(0);
    for (int k0 = 0; k0 < cnt; k0 += U) {
      u64 vr_c[U], mk[U];
      u32 c_c[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        // Saturating add: a path past CGP_INF stays unreachable.
        const u32 c = __builtin_elementwise_add_sat(du[k], cand[k]);
        mk[k] = __ballot(c < cur[k]) & am[k];  // lanes whose value improves
        c_c[k] = c;
        vr_c[k] = vr[k];
      }
      if (k0 + U < cnt) issue(k0 + U);
      u32 mine_lo = 0, mine_hi = 0;
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const u64 m = mk[k];
        if (m == 0) continue;
        if ((m >> lane) & 1ull) {
          __hip_atomic_fetch_min((gu32*)vr_c[k] + (unsigned)lq, c_c[k], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
          note_improved(t, c_c[k], thr);
          ++t.imp;
        }
        if (!SPARSE) {
          const bool own = lane == k0 + k;
          mine_lo = own ? (u32)m : mine_lo;
          mine_hi = own ? (u32)(m >> 32) : mine_hi;
        } else {
          const int32_t vk = __builtin_amdgcn_readlane(myv, k0 + k);
          bool claimed = false;
          if (lane == 0) {
            ++t.nnew;
            const int wk = __builtin_amdgcn_readlane(myword, k0 + k);
            if (atomicOr(&P.next[nxt_at<W>(vk, wk)], m) == 0) claimed = claim_vertex(P.mark, vk, stamp);
          }
          wave_append<int32_t>(claimed, vk, &P.ctl->tcount.v, P.touched);
        }
      }
      // Dense mode: lane k0+k owns edge k0+k's next-word update, one masked
      // instruction for the whole step.
      if (!SPARSE && (mine_lo | mine_hi)) {
        ++t.nnew;
        atomicOr(&P.next[nxt_at<W>(myv, myword)], ((u64)mine_hi << 32) | mine_lo);
      }
    }
  }
}

// Phase D without postponing.  The wave's items are taken one word at a
// time, so a window's edges all belong to one word: its active-lane mask and
// lane offsets are fixed for the window.  Once per window lane f writes edge
// f's target and source rows and weight into the wave's LDS table; each step
// reads its entry back as a broadcast LDS load (no per-step scalar work),
// adds the lane offset, reads the two rows, and the improved lanes drive the
// atomicMin directly as the exec mask.  Row addresses: when one word's value
// rows span less than 4 GiB (!A64), 32-bit byte offsets from the word's
// first row through a raw buffer descriptor; otherwise (A64) 64-bit
// addresses in the table and global loads.
// In dense frontier mode lane k collects step k's improved mask with
// v_writelane, and one atomicOr per lane per window publishes the next bits.
// Steps past the window's last edge read an entry of weight CGP_INF, so they
// cannot improve anything.
template <int W, bool SPARSE, int U, bool A64>
__device__ __forceinline__ void expand_flat_dense_tab(const KP& P, u32 base, u32 n, int rshift, u32 stamp,
                                                      uint4* tab, u32* wtab, Tally& t) {
  const int lane = lane_id();
  const int R = 1 << rshift;
  const u32 p = P.ctl->iter & 1u;
  const bool have = base + (u32)lane < n;
  const u64 item = have ? P.dense[p][base + lane] : 0ull;
  const int32_t u = (int32_t)(item >> 32);
  const int w = (int)((item >> 24) & 0xFF);
  const u32 chunk = (u32)(item & 0xFFFFFF);
  const int32_t e0 = have ? P.off[u] + (int32_t)(chunk * DCHUNK) : 0;
  const int32_t e1 = have ? min(P.off[u + 1], e0 + DCHUNK) : 0;
  const u64 sm = have ? (P.front[(size_t)u * W + w] & P.cfg->val_bits[w]) : 0ull;
  const u32 dlen = sm ? (u32)(e1 - e0) : 0u;
  if (have) {
    ++t.ditems;
    t.dedges += dlen;
    t.sssp += (u64)dlen * (u64)__popcll(sm);
    t.src += (u64)__popcll(sm);
  }
  const bool inrow = lane < R;
  const int lq = lane & (R - 1);
  const u32 lq4 = (u32)lq * 4u;
  const u32 rowb = 4u << rshift;               // bytes per value row
  const u32 wordb = (u32)P.V * rowb;           // bytes per word (!A64: < 2^32, checked by the caller)
  const u32 srcb = (u32)u * rowb;              // the item's source row inside its word (!A64)
  const u64 srca = (u64)P.vals + ((u64)((u32)w * (u32)P.V + (u32)u) << (rshift + 2));  // (A64)
  u32 imp_n = 0;  // improved (vertex, query) pairs, wave-uniform
  for (int wp = 0; wp < W; ++wp) {
    const u64 vb = P.cfg->val_bits[wp];
    if (vb == 0 || __ballot(dlen != 0 && w == wp) == 0) continue;
    // 32-bit form: a descriptor over this word's rows (unused, and not
    // formed, in the 64-bit form, where wordb may not fit 32 bits).
    const i32x4 rs = A64 ? i32x4{0, 0, 0, 0} : raw_rsrc((u64)P.vals + (u64)wp * wordb, wordb);
    // This word's SSSP lanes read their own value; the others read the
    // row's first value (the same lines) and are masked out.
    const bool inw = inrow && ((vb >> lq) & 1ull) != 0;
    const u64 am = __ballot(inw);
    const u32 off = inw ? lq4 : 0u;
    const u32 d = w == wp ? dlen : 0u;
    const u32 incl = wave_incl_scan(d);
    const u32 total = (u32)__builtin_amdgcn_readlane((int)incl, 63);
    for (u32 win = 0; win < total; win += 64) {
      const u32 f = win + (u32)lane;
      const int j = wave_upper_bound(incl, f);
      const u32 exj = (u32)__shfl((int)(incl - d), j, 64);
      const int32_t ej = __shfl(e0, j, 64) + (int32_t)(f - exj);
      const bool ok = f < total;
      const int32_t myv = ok ? P.idx[ej] : 0;
      const u32 myw = ok ? (P.wt ? P.wt[ej] : 1u) : INF;
      if (A64) {
        const u64 ta = (u64)P.vals + ((u64)((u32)wp * (u32)P.V + (u32)myv) << (rshift + 2));
        // Shuffle with every lane active (item lane j may lie past the
        // window's last edge; a shuffle from an inactive lane reads 0), then
        // select.
        const u32 sa_hi = (u32)__shfl((int)(u32)(srca >> 32), j, 64);
        const u32 sa_lo = (u32)__shfl((int)(u32)srca, j, 64);
        const u64 sa = ok ? (((u64)sa_hi << 32) | sa_lo) : (u64)P.vals;
        tab[lane] = make_uint4((u32)ta, (u32)(ta >> 32), (u32)sa, (u32)(sa >> 32));
        wtab[lane] = myw;
      } else {
        const u32 sb = (u32)__shfl((int)srcb, j, 64);
        tab[lane] = make_uint4((u32)myv * rowb, ok ? sb : 0u, myw, 0u);
      }
      __builtin_amdgcn_wave_barrier();
      const int cnt = (int)min(64u, total - win);
      // Per step: the lane's target value (offset or address), the weight
      // and the two loaded values.  A window whose edges all belong to one
      // item (chunks of high-degree vertices) reads that item's source row
      // once instead of once per step.
      typedef typename std::conditional<A64, u64, u32>::type addr_t;
      const int jl = __builtin_amdgcn_readlane(j, cnt - 1);
      const bool one_item = __builtin_amdgcn_readlane(j, 0) == jl;
      u32 mine_lo = 0, mine_hi = 0;
      auto steps = [&](auto one_c) {
        constexpr bool ONE = decltype(one_c)::value;
        u32 dsrc = 0;
        if constexpr (ONE) {
          if constexpr (A64) {
            const u64 sa1 = ((u64)(u32)__builtin_amdgcn_readlane((int)(u32)(srca >> 32), jl) << 32) |
                            (u32)__builtin_amdgcn_readlane((int)(u32)srca, jl);
            dsrc = *(const gu32*)(sa1 + off);
          } else {
            dsrc = (u32)cgp_buf_load_i32(rs, (int)((u32)__builtin_amdgcn_readlane((int)srcb, jl) + off), 0, 0);
          }
        }
        addr_t ot[U];
        u32 wt[U], du[U], cur[U];
        auto issue = [&](int k0) {
#pragma unroll
          for (int k = 0; k < U; ++k) {
            const uint4 e = tab[k0 + k];  // k0 + k < 64: k0 < cnt <= 64 and U divides 64
            if constexpr (A64) {
              ot[k] = (((u64)e.y << 32) | e.x) + off;
              wt[k] = wtab[k0 + k];
              if constexpr (!ONE) du[k] = *(const gu32*)((((u64)e.w << 32) | e.z) + off);
              cur[k] = *(const gu32*)ot[k];
            } else {
              ot[k] = e.x + off;
              wt[k] = e.z;
              if constexpr (!ONE) du[k] = (u32)cgp_buf_load_i32(rs, (int)(e.y + off), 0, 0);
              cur[k] = (u32)cgp_buf_load_i32(rs, (int)ot[k], 0, 0);
            }
          }
        };
        issue(0);
        for (int k0 = 0; k0 < cnt; k0 += U) {
          u64 mk[U];
          u32 c_c[U];
          addr_t ot_c[U];
#pragma unroll
          for (int k = 0; k < U; ++k) {
            // Saturating add: a path past CGP_INF stays unreachable.
            const u32 c = __builtin_elementwise_add_sat(ONE ? dsrc : du[k], wt[k]);
            mk[k] = __ballot(c < cur[k]) & am;
            c_c[k] = c;
            ot_c[k] = ot[k];
          }
          if (k0 + U < cnt) issue(k0 + U);
#pragma unroll
          for (int k = 0; k < U; ++k) {
            const u64 m = mk[k];
            if (m == 0) continue;
            imp_n += (u32)__popcll(m);
            // The improved lanes are SSSP lanes of the word, whose offset is
            // their own value's.
            if (__builtin_amdgcn_inverse_ballot_w64(m)) {
              if constexpr (A64)
                __hip_atomic_fetch_min((gu32*)ot_c[k], c_c[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              else
                cgp_buf_atomic_umin_i32((int)c_c[k], rs, (int)ot_c[k], 0, 0);
            }
            if (!SPARSE) {
              mine_lo = (u32)cgp_writelane_i32((int)(u32)m, k0 + k, (int)mine_lo);
              mine_hi = (u32)cgp_writelane_i32((int)(u32)(m >> 32), k0 + k, (int)mine_hi);
            } else {
              const int32_t vk = __builtin_amdgcn_readlane(myv, k0 + k);
              bool claimed = false;
              if (lane == 0) {
                ++t.nnew;
                if (atomicOr(&P.next[nxt_at<W>(vk, wp)], m) == 0) claimed = claim_vertex(P.mark, vk, stamp);
              }
              wave_append<int32_t>(claimed, vk, &P.ctl->tcount.v, P.touched);
            }
          }
        }
      };
      // (Not in the 64-bit form: the second copy of the loop raises its
      // registers past two waves per SIMD.)
      if (!A64 && one_item) steps(std::true_type{});
      else steps(std::false_type{});
      // Dense mode: lane f owns edge f's next-word update.
      if (!SPARSE && (mine_lo | mine_hi)) {
        ++t.nnew;
        atomicOr(&P.next[nxt_at<W>(myv, wp)], ((u64)mine_hi << 32) | mine_lo);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (lane == 0) t.imp += imp_n;
}

// Lane K of each 16-lane row, broadcast to the whole row (DPP row_newbcast).
template <int K>
__device__ __forceinline__ u32 row_bcast(u32 x) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x150 + K, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ u64 row_bcast64(u64 x) {
  return ((u64)row_bcast<K>((u32)(x >> 32)) << 32) | row_bcast<K>((u32)x);
}

// Phase N: dense SSSP items of narrow words (every SSSP slot of the word in
// lanes 0..15, rows of at least 16 lanes).  The 64 flattened edges of a
// window are split by 16-lane row: row g relaxes edges 16g..16g+15, one per
// step, its lanes being the word's sixteen queries, so a wave step relaxes
// four edges.  A step's row pointers, weight and mask reach the row by a
// DPP row broadcast from the lane that flattened the edge.  Otherwise as
// phase D (no near-far split: the filter routes items here only when
// postponing is off).
template <int W, bool SPARSE>
__device__ __forceinline__ void expand_flat_narrow(const KP& P, u32 base, u32 n, int rshift, u32 stamp,
                                                   Tally& t) {
  const int lane = lane_id();
  const u32 p = P.ctl->iter & 1u;
  const bool have = base + (u32)lane < n;
  const u64 item = have ? P.dense_n[p][base + lane] : 0ull;
  const int32_t u = (int32_t)(item >> 32);
  const int w = (int)((item >> 24) & 0xFF);
  const u32 chunk = (u32)(item & 0xFFFFFF);
  const int32_t e0 = have ? P.off[u] + (int32_t)(chunk * DCHUNK) : 0;
  const int32_t e1 = have ? min(P.off[u + 1], e0 + DCHUNK) : 0;
  const u32 sm = have ? (u32)(P.front[(size_t)u * W + w] & P.cfg->val_bits[w]) & 0xFFFFu : 0u;
  const u32 d = sm ? (u32)(e1 - e0) : 0u;
  if (have) {
    ++t.ditems;
    t.dedges += d;
    t.sssp += (u64)d * (u64)__popc(sm);
    t.src += (u64)__popc(sm);
  }
  const u32 incl = wave_incl_scan(d);
  const u32 total = (u32)__builtin_amdgcn_readlane((int)incl, 63);
  const u32 item_row = (u32)w * (u32)P.V + (u32)u;
  const u64 vals_base = (u64)P.vals;
  const int rsh = rshift + 2;
  const int g = lane >> 4;
  const u32 ql = (u32)(lane & 15), ql4 = ql * 4u;
  for (u32 win = 0; win < total; win += 64) {
    const u32 f = win + (u32)lane;
    const int j = wave_upper_bound(incl, f);
    const u32 exj = (u32)__shfl((int)(incl - d), j, 64);
    const int32_t ej = __shfl(e0, j, 64) + (int32_t)(f - exj);
    const bool ok = f < total;
    const int32_t myv = ok ? P.idx[ej] : 0;
    const u32 myw = ok ? (P.wt ? P.wt[ej] : 1u) : 0u;
    const int myword = __shfl(w, j, 64);
    const u64 my_vaddr = vals_base + ((u64)((u32)myword * (u32)P.V + (u32)myv) << rsh);
    const u64 my_uaddr = vals_base + ((u64)(u32)__shfl((int)item_row, j, 64) << rsh);
    const u32 am_j = (u32)__shfl((int)sm, j, 64);  // every lane active for the shuffle
    const u32 my_am = ok ? am_j : 0u;
    const int cnt = (int)min(64u, total - win);
    u32 mine = 0;
    // Four steps per chunk: their eight row loads are issued together.
    auto chunk = [&](auto k0c) {
      constexpr int K0 = decltype(k0c)::value;
      u64 va[4];
      u32 du[4], cur[4], c[4];
      bool act[4];
      auto load = [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int i = K - K0;
        va[i] = row_bcast64<K>(my_vaddr);
        const u64 ua = row_bcast64<K>(my_uaddr);
        c[i] = row_bcast<K>(myw);
        act[i] = ((row_bcast<K>(my_am) >> ql) & 1u) != 0;
        du[i] = *(const gu32*)(ua + ql4);
        cur[i] = *(const gu32*)(va[i] + ql4);
      };
      load(std::integral_constant<int, K0 + 0>{});
      load(std::integral_constant<int, K0 + 1>{});
      load(std::integral_constant<int, K0 + 2>{});
      load(std::integral_constant<int, K0 + 3>{});
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const u32 cand = __builtin_elementwise_add_sat(du[i], c[i]);
        const bool imp = act[i] && cand < cur[i];
        if (imp) {
          __hip_atomic_fetch_min((gu32*)(va[i] + ql4), cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          note_improved(t, cand, INF);
          ++t.imp;
        }
        const u64 m = __ballot(imp);
        if (m == 0) continue;
        // Row g's bits are edge 16g+K's next-word bits; that edge's own lane
        // (ql == K) publishes them.
        const u32 mg = (u32)(m >> (16 * g)) & 0xFFFFu;
        const bool own = ql == (u32)(K0 + i);
        if (!SPARSE) {
          mine = own ? mg : mine;
        } else {
          bool claimed = false;
          if (own && mg) {
            ++t.nnew;
            if (atomicOr(&P.next[nxt_at<W>(myv, myword)], (u64)mg) == 0) claimed = claim_vertex(P.mark, myv, stamp);
          }
          wave_append<int32_t>(claimed, myv, &P.ctl->tcount.v, P.touched);
        }
      }
    };
    // Row 0 holds the window's first 16 edges, so a chunk starting at or
    // past `cnt` has no edge in any row.
    chunk(std::integral_constant<int, 0>{});
    if (4 < cnt) chunk(std::integral_constant<int, 4>{});
    if (8 < cnt) chunk(std::integral_constant<int, 8>{});
    if (12 < cnt) chunk(std::integral_constant<int, 12>{});
    if (!SPARSE && mine) {
      ++t.nnew;
      atomicOr(&P.next[nxt_at<W>(myv, myword)], (u64)mine);
    }
  }
}

// Bottom-up BFS step for vertex v (one lane) on a symmetric graph: an
// unvisited (v, q) becomes visited at depth L+1 iff some neighbour u has q
// visited (all visited bits at this point have depth <= L, and a neighbour
// at depth < L would already have visited v).  visited is only written by
// the filter, so the scan reads one consistent level.  Early exit once every
// missing bit was found.
template <int W, bool SPARSE>
__device__ __forceinline__ bool pull_vertex(const KP& P, int32_t v, const u64 (&bb)[W], u32 stamp,
                                            Tally& t) {
  u64 tgt[W], acc[W];
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    tgt[w] = bb[w] & ~P.visited[vis_at<W>(v, w)];
    acc[w] = 0;
    any |= tgt[w] != 0;
  }
  if (!any) return false;
  ++t.pverts;
  const int32_t e1 = P.off[v + 1];
  for (int32_t e = P.off[v]; e < e1; ++e) {
    const int32_t u = P.idx[e];
    ++t.pedges;
    bool done = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (tgt[w] & ~acc[w]) {
        ++t.pwords;
        acc[w] |= P.visited[vis_at<W>(u, w)] & tgt[w];
        done &= (acc[w] == tgt[w]);
      }
    }
    if (done) break;
  }
  bool first = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    if (acc[w]) {
      ++t.nnew;
      if (SPARSE) {
        if (atomicOr(&P.next[nxt_at<W>(v, w)], acc[w]) == 0) first = true;
      } else {
        atomicOr(&P.next[nxt_at<W>(v, w)], acc[w]);
      }
    }
  }
  return SPARSE && first && claim_vertex(P.mark, v, stamp);
}

// Bottom-up step for a high-degree vertex v with the whole wave: each step
// ORs the visited words of 64 neighbours, and the scan stops once every
// missing bit was found.  Returns (on every lane) whether the wave made the
// vertex's first claim of the iteration (sparse mode).
template <int W, bool SPARSE>
__device__ __forceinline__ bool pull_vertex_wave(const KP& P, int32_t v, const u64 (&bb)[W], u32 stamp,
                                                 Tally& t) {
  const int lane = lane_id();
  u64 tgt[W], acc[W];
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    tgt[w] = bb[w] & ~P.visited[vis_at<W>(v, w)];
    acc[w] = 0;
    any |= tgt[w] != 0;
  }
  if (!any) return false;
  if (lane == 0) ++t.pverts;
  const int32_t e0 = P.off[v], e1 = P.off[v + 1];
  for (int32_t eb = e0; eb < e1; eb += 64) {
    const int32_t e = eb + lane;
    const bool ok = e < e1;
    const int32_t u = ok ? P.idx[e] : 0;
    bool done = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const u64 a = ok ? (P.visited[vis_at<W>(u, w)] & tgt[w]) : 0ull;
      acc[w] |= wave_or64(a);
      done &= acc[w] == tgt[w];
    }
    if (ok) {
      ++t.pedges;
      t.pwords += W;
    }
    if (done) break;
  }
  bool first = false;
  if (lane == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (acc[w]) {
        ++t.nnew;
        if (SPARSE) {
          if (atomicOr(&P.next[nxt_at<W>(v, w)], acc[w]) == 0) first = true;
        } else {
          atomicOr(&P.next[nxt_at<W>(v, w)], acc[w]);
        }
      }
    }
    first = SPARSE && first && claim_vertex(P.mark, v, stamp);
  }
  return first;
}

__device__ __forceinline__ u32 wave_fetch(u32* ctr, u32 n) {
  u32 i = 0;
  if (lane_id() == 0) i = atomicAdd(ctr, n);
  return (u32)__shfl((int)i, 0, 64);
}

// Wave-level claims of `chunk` units from NSHARD contiguous shards of [0, n):
// shard s covers [n*s/NSHARD, n*(s+1)/NSHARD) and has cursor cur[s] (units
// handed out so far).  A wave starts at its workgroup's shard and moves to the
// next when that one is exhausted (cursors only grow, so an exhausted shard
// stays exhausted); after NSHARD exhausted shards every unit has been claimed.
// Claims never cross a shard's end.  All state is wave-uniform.
struct ShardFetch {
  int s, left;
  __device__ __forceinline__ ShardFetch() : s((int)(blockIdx.x % NSHARD)), left(NSHARD) {}
  __device__ __forceinline__ bool next(Ctr* cur, u32 n, u32 chunk, u32& lo, u32& hi) {
    while (left > 0) {
      const u32 b = (u32)(((u64)n * (u32)s) / NSHARD), e = (u32)(((u64)n * (u32)(s + 1)) / NSHARD);
      if (b < e) {
        const u32 i = wave_fetch(&cur[s].v, chunk);
        if (i < e - b) {
          lo = b + i;
          hi = min(e, lo + chunk);
          return true;
        }
      }
      s = s + 1 == NSHARD ? 0 : s + 1;
      --left;
    }
    return false;
  }
};

// Body of the advance for one frontier mode; workgroups pull work
// dynamically, heaviest first:
//   L  one workgroup per LCHUNK-edge chunk of a high-degree vertex
//   D  64 (vertex, dense SSSP word, DCHUNK-edge chunk) items per wave,
//      edges flattened, lane = query
//   M  one wave per medium-degree vertex, lane = edge
//   S  64 low-degree vertices per wave, edges flattened, lane = edge
//   P  (pull iterations) one lane per vertex of the whole graph
template <int W, bool SPARSE, int BS>
__device__ __forceinline__ void advance_push_body(const KP& P, u32* s_item, Tally& t) {
  Ctl* C = P.ctl;
  const u32 iter = C->iter, p = iter & 1u;
  const bool bfs_on = C->bfs_pull == 0;
  const u32 stamp = C->stamp_base + iter + 1;
  const u32 thr = C->thr;
  // In an SSSP pull iteration every vertex gathers its SSSP values itself: a
  // threshold of 0 leaves no SSSP word to the push phases.
  const int rshift = P.cfg->rshift, dense_t = C->sssp_pull ? 0 : P.cfg->dense_t;
  const int lane = lane_id();
  // Phase L.
  const u32 nL = C->bcount[p][2].v;
  while (nL) {
    __syncthreads();
    if (threadIdx.x == 0) *s_item = atomicAdd(&C->fetch[p][0].v, 1u);
    __syncthreads();
    const u32 i = *s_item;
    if (i >= nL) break;
    const u64 item = P.large[p][i];
    const int32_t u = (int32_t)(item >> 32);
    const int32_t e0 = P.off[u] + (int32_t)((u32)item * LCHUNK);
    const int32_t e1 = min(P.off[u + 1], e0 + LCHUNK);
    expand_lanes<W, SPARSE>(P, u, e0, e1, e0 + (int32_t)threadIdx.x, BS, bfs_on, dense_t, rshift,
                            stamp, thr, (u32)item == 0 && threadIdx.x == 0, t);
  }
  // Work cursors are single addresses every wave contends for, so each
  // fetch claims several units (FETCH_M medium vertices, FETCH_S groups of
  // 64 small vertices).
  constexpr u32 FETCH_M = 4, FETCH_S = 4;
  // Phase M.
  const u32 nM = C->bcount[p][1].v;
  u32 lo, hi;
  for (ShardFetch sf; nM && sf.next(C->shard[p][SH_M], nM, FETCH_M, lo, hi);) {
    for (u32 i = lo; i < hi; ++i) {
      const int32_t u = P.medium[p][i];
      const int32_t e0 = P.off[u], e1 = P.off[u + 1];
      expand_lanes<W, SPARSE>(P, u, e0, e1, e0 + lane, 64, bfs_on, dense_t, rshift, stamp, thr, lane == 0, t);
    }
  }
  // Phase S: groups of 64 entries (expand_flat_vertices).
  const u32 nS = C->bcount[p][0].v;
  const u32 ngS = (nS + 63u) / 64u;
  for (ShardFetch sf; nS && sf.next(C->shard[p][SH_S], ngS, FETCH_S, lo, hi);) {
    for (u32 g = lo; g < hi; ++g)
      expand_flat_vertices<W, SPARSE>(P, P.small[p], g, ngS, nS, bfs_on, dense_t, rshift, stamp, thr, t);
  }
  // Phase P: high-degree vertices one wave each (64 neighbours per step,
  // early exit on a wave-wide OR), then the rest one lane each.
  if (!bfs_on) {
    u64 bb[W];
#pragma unroll
    for (int w = 0; w < W; ++w) bb[w] = P.cfg->bfs_bits[w];
    const u32 nH = (u32)P.n_pull_hi;
    while (nH) {
      const u32 i = wave_fetch(&C->fetch[p][4].v, 1u);
      if (i >= nH) break;
      const int32_t v = P.pull_hi[i];
      bool claimed = false;
      if (pull_vertex_wave<W, SPARSE>(P, v, bb, stamp, t) && lane == 0) claimed = true;
      if (SPARSE) wave_append<int32_t>(claimed, v, &C->tcount.v, P.touched);
    }
    const int64_t nLo = P.n_pull_lo;
    for (int64_t base = (int64_t)blockIdx.x * BS; base < nLo; base += (int64_t)gridDim.x * BS) {
      const int64_t i = base + threadIdx.x;
      const int32_t v = i < nLo ? P.pull_lo[i] : 0;
      const bool claimed = i < nLo && pull_vertex<W, SPARSE>(P, v, bb, stamp, t);
      if (SPARSE) wave_append<int32_t>(claimed, v, &C->tcount.v, P.touched);
    }
  }
}

// Publishes the wave's near-far statistics: whether some improvement stayed
// below the threshold, and the smallest value at or above it.
__device__ __forceinline__ void flush_near_far(const KP& P, const Tally& t, u32 thr) {
  if (thr == INF) return;
  const u64 anyn = __ballot(t.nnear != 0);
  u32 fm = t.fmin;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) fm = min(fm, (u32)__shfl_xor((int)fm, d, 64));
  if (lane_id() == 0) {
    if (anyn) atomicOr(&P.ctl->near_new, 1u);
    if (fm != INF) atomicMin(&P.ctl->far_min, fm);
  }
}

// Flushes a lane's work counters (profile mode).
__device__ __forceinline__ void flush_tally(const KP& P, const Tally& t) {
  prof_add(P, PF_PUSH_ITEMS, t.items);
  prof_add(P, PF_PUSH_EDGES, t.edges);
  prof_add(P, PF_PUSH_WEDGES, t.wedges);
  prof_add(P, PF_BFS_WORDS, t.bfs);
  prof_add(P, PF_SSSP_PAIRS, t.sssp);
  prof_add(P, PF_NEW_WORDS, t.nnew);
  prof_add(P, PF_IMPROVED, t.imp);
  prof_add(P, PF_DENSE_ITEMS, t.ditems);
  prof_add(P, PF_DENSE_EDGES, t.dedges);
  prof_add(P, PF_PULL_VERTS, t.pverts);
  prof_add(P, PF_PULL_WORDS, t.pwords);
  prof_add(P, PF_PULL_EDGES, t.pedges);
  prof_add(P, PF_SSSP_WORDS, t.swords);
  prof_add(P, PF_SSSP_SRC, t.src);
}

// The advance of one iteration is two persistent launches of BS-thread
// workgroups (BS is a tuning parameter: 128, 256 or 512).  k_adv_dense runs
// the lane-per-query rows of dense SSSP words (phase D) with its own lean
// register budget, one instantiation per frontier mode (the one that does
// not match the device-selected mode exits at once); k_adv_push runs phases
// L, M, S and P.
template <int W, int BS, bool SPARSE, int U, int TAB>
__global__ __launch_bounds__(BS) void k_adv_dense(KP P) {
  Ctl* C = P.ctl;
  if (C->finished || (C->dense_mode == 0) != SPARSE || C->sssp_pull) return;
  const u32 iter = C->iter, p = iter & 1u;
  const u32 nD = C->bcount[p][3].v;
  if (nD == 0) return;
  if (threadIdx.x == 0) {
    const u64 now = (u64)wall_clock64();
    atomicMin(&C->adv_t0, now);
    atomicMin(&C->den_t0, now);
  }
  const u32 stamp = C->stamp_base + iter + 1;
  const u32 thr = C->thr;
  const int rshift = P.cfg->rshift;
  Tally t;
  // TAB (1: 32-bit row offsets, 2: 64-bit row addresses): postponing off
  // (thr is CGP_INF throughout), the table loop in a kernel of its own so
  // the other loops' registers do not lower its occupancy.
  __shared__ uint4 s_tab[TAB ? BS : 1];  // per wave: 64 step entries of the current window
  __shared__ u32 s_wt[TAB == 2 ? BS : 1];
  uint4* tab = s_tab + (TAB ? (threadIdx.x & ~63u) : 0u);
  u32* wtab = s_wt + (TAB == 2 ? (threadIdx.x & ~63u) : 0u);
  // One group of 64 items per fetch: items carry up to DCHUNK edges each, so
  // coarser claims leave a long tail (measured 2.5x slower at 4 groups).  A
  // single cursor: the per-XCD shards of the other phases measured about
  // 1 ms slower here on the RMAT-22 mixed batch (29.8 against 28.9 ms), the
  // stealing at the end of the shards lengthening the tail.
  while (true) {
    const u32 i = wave_fetch(&C->fetch[p][1].v, 64u);
    if (i >= nD) break;
    if (TAB) expand_flat_dense_tab<W, SPARSE, U, TAB == 2>(P, i, nD, rshift, stamp, tab, wtab, t);
    else if (thr == INF) expand_flat_dense<W, SPARSE, U, false>(P, i, nD, rshift, stamp, thr, t);
    else expand_flat_dense<W, SPARSE, U, true>(P, i, nD, rshift, stamp, thr, t);
  }
  flush_near_far(P, t, thr);
  if (P.profile) {
    prof_add(P, PF_DENSE_ITEMS, t.ditems);
    prof_add(P, PF_DENSE_EDGES, t.dedges);
    prof_add(P, PF_DENSE_PAIRS, t.sssp);
    prof_add(P, PF_DENSE_IMP, t.imp);
    prof_add(P, PF_DENSE_NEW, t.nnew);
    prof_add(P, PF_DENSE_SRC, t.src);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 now = (u64)wall_clock64();
    atomicMax(&C->adv_t1, now);
    atomicMax(&C->den_t1, now);
  }
}

// Phase N launch: the narrow-word lane-per-query items (both frontier modes
// in one kernel; timed together with k_adv_dense).
template <int W>
__global__ __launch_bounds__(BLOCK) void k_adv_narrow(KP P) {
  Ctl* C = P.ctl;
  if (C->finished || C->sssp_pull) return;
  const u32 iter = C->iter, p = iter & 1u;
  const u32 nN = C->bcount[p][4].v;
  if (nN == 0) return;
  if (threadIdx.x == 0) {
    const u64 now = (u64)wall_clock64();
    atomicMin(&C->adv_t0, now);
    atomicMin(&C->den_t0, now);
  }
  const u32 stamp = C->stamp_base + iter + 1;
  const int rshift = P.cfg->rshift;
  const bool sparse = C->dense_mode == 0;
  Tally t;
  u32 lo, hi;
  for (ShardFetch sf; sf.next(C->shard[p][SH_N], nN, 64u, lo, hi);) {
    if (sparse) expand_flat_narrow<W, true>(P, lo, hi, rshift, stamp, t);
    else expand_flat_narrow<W, false>(P, lo, hi, rshift, stamp, t);
  }
  if (P.profile) {
    prof_add(P, PF_DENSE_ITEMS, t.ditems);
    prof_add(P, PF_DENSE_EDGES, t.dedges);
    prof_add(P, PF_DENSE_PAIRS, t.sssp);
    prof_add(P, PF_DENSE_IMP, t.imp);
    prof_add(P, PF_DENSE_NEW, t.nnew);
    prof_add(P, PF_DENSE_SRC, t.src);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 now = (u64)wall_clock64();
    atomicMax(&C->adv_t1, now);
    atomicMax(&C->den_t1, now);
  }
}

// SSSP pull (min-plus gather), the direction-optimising step of value-
// carrying queries: every vertex v takes, for each SSSP query of each word,
// the minimum over its edges (v, u) of value(u) + w(u, v) and lowers its own
// value to it (pulling needs the reverse of every edge with the same weight,
// which the host checks before choosing pull).  Items are (vertex, PCHUNK-edge
// chunk) pairs of every vertex with edges, taken 16 at a time; a wave runs
// one item per word with lane q as query q: 64 edge ids and weights are
// loaded at once, then U neighbour rows per step are in flight.  Improved
// pairs are published by atomicMin (a vertex's chunks run on different
// waves) and their next bits by one atomicOr per (item, word).
template <int W, bool SPARSE, int U>
__device__ __forceinline__ void pull_sssp_item(const KP& P, u64 item, int rshift, u32 stamp, Tally& t) {
  const int lane = lane_id();
  const int R = 1 << rshift;
  const int32_t v = (int32_t)(item >> 32);
  const int32_t e0 = P.off[v] + (int32_t)((u32)item * PCHUNK);
  const int32_t e1 = min(P.off[v + 1], e0 + PCHUNK);
  const int lq = lane & (R - 1);
  const u32 lq4 = (u32)lq * 4u;
  ++t.pverts;
  if (lane == 0) t.pedges += (u32)(e1 - e0);
  bool claimed = false;
  for (int w = 0; w < W; ++w) {
    const u64 vb = P.cfg->val_bits[w];
    if (!vb) continue;
    const bool inw = lane < R && ((vb >> lq) & 1ull) != 0;
    const u32 off = inw ? lq4 : 0u;  // other lanes re-read the row's first value (same lines)
    u32 acc = INF;
    for (int32_t eb = e0; eb < e1; eb += 64) {
      const int32_t e = eb + lane;
      const bool ok = e < e1;
      const int32_t myu = ok ? P.idx[e] : 0;
      const u32 myw = ok ? (P.wt ? P.wt[e] : 1u) : INF;  // steps past the end add INF
      const int cnt = min(64, e1 - eb);
      t.pwords += lane < cnt ? 1u : 0u;
      for (int k0 = 0; k0 < cnt; k0 += U) {
        u32 val[U], wk[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const int j = min(k0 + k, 63);
          const u32 uk = (u32)__builtin_amdgcn_readlane(myu, j);
          wk[k] = (u32)__builtin_amdgcn_readlane((int)myw, j);
          val[k] = *(const gu32*)((u64)P.vals + (((u64)((u32)w * (u32)P.V + uk) << (rshift + 2)) + off));
        }
#pragma unroll
        for (int k = 0; k < U; ++k) acc = min(acc, __builtin_elementwise_add_sat(val[k], wk[k]));
      }
    }
    const size_t vr = vrow(P.V, w, v, rshift);
    const u32 cur = P.vals[vr + (inw ? lq : 0)];
    const bool imp = inw && acc < cur;
    const u64 m = __ballot(imp);
    if (!m) continue;
    if (imp) {
      atomicMin(&P.vals[vr + lq], acc);
      ++t.imp;
    }
    if (lane == 0) {
      ++t.nnew;
      if (SPARSE) {
        if (atomicOr(&P.next[nxt_at<W>(v, w)], m) == 0) claimed = true;
      } else {
        atomicOr(&P.next[nxt_at<W>(v, w)], m);
      }
    }
  }
  if (SPARSE) {
    const bool c = lane == 0 && claimed && claim_vertex(P.mark, v, stamp);
    wave_append<int32_t>(c, v, &P.ctl->tcount.v, P.touched);
  }
}

template <int W>
__global__ __launch_bounds__(BLOCK) void k_pull_sssp(KP P) {
  Ctl* C = P.ctl;
  if (C->finished || !C->sssp_pull) return;
  if (threadIdx.x == 0) atomicMin(&C->adv_t0, (u64)wall_clock64());
  const u32 iter = C->iter, p = iter & 1u;
  const u32 stamp = C->stamp_base + iter + 1;
  const int rshift = P.cfg->rshift;
  const bool sparse = C->dense_mode == 0;
  const u32 n = (u32)P.n_pull_items;
  Tally t;
  constexpr u32 FETCH_P = 16;
  u32 lo, hi;
  for (ShardFetch sf; sf.next(C->shard[p][SH_P], n, FETCH_P, lo, hi);) {
    for (u32 i = lo; i < hi; ++i) {
      if (sparse) pull_sssp_item<W, true, 8>(P, P.pull_items[i], rshift, stamp, t);
      else pull_sssp_item<W, false, 8>(P, P.pull_items[i], rshift, stamp, t);
    }
  }
  if (P.profile) {
    prof_add(P, PF_PULL_VERTS, t.pverts);
    prof_add(P, PF_PULL_EDGES, t.pedges);
    prof_add(P, PF_PULL_WORDS, t.pwords);
    prof_add(P, PF_IMPROVED, t.imp);
    prof_add(P, PF_NEW_WORDS, t.nnew);
  }
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(&C->adv_t1, (u64)wall_clock64());
}

// (vertex << 32 | chunk) items of every vertex with edges, PCHUNK edges each
// (order immaterial: an item is one vertex's gather over a range of edges).
__global__ __launch_bounds__(BLOCK) void k_pull_items(const int32_t* off, int32_t V, u64* items, u32* count) {
  for (int64_t b = (int64_t)blockIdx.x * BLOCK; b < V; b += (int64_t)gridDim.x * BLOCK) {
    const int64_t v = b + threadIdx.x;
    const int32_t d = v < V ? off[v + 1] - off[v] : 0;
    const u32 nch = (u32)((d + PCHUNK - 1) / PCHUNK);
    // Wave-level reservation of the wave's items, then each lane writes its own.
    u32 x = nch;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      const u32 y = (u32)__shfl_up((int)x, s, 64);
      if (lane_id() >= s) x += y;
    }
    const u32 total = (u32)__shfl((int)x, 63, 64);
    u32 base = 0;
    if (lane_id() == 0 && total) base = atomicAdd(count, total);
    base = (u32)__shfl((int)base, 0, 64) + x - nch;
    for (u32 c = 0; c < nch; ++c) items[base + c] = ((u64)(u32)v << 32) | c;
  }
}

template <int W, int BS>
__global__ __launch_bounds__(BS) void k_adv_push(KP P) {
  Ctl* C = P.ctl;
  if (C->finished) return;
  __shared__ u32 s_item;
  if (threadIdx.x == 0) atomicMin(&C->adv_t0, (u64)wall_clock64());
  Tally t;
  if (C->dense_mode) advance_push_body<W, false, BS>(P, &s_item, t);
  else advance_push_body<W, true, BS>(P, &s_item, t);
  flush_near_far(P, t, C->thr);
  if (P.profile) flush_tally(P, t);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(&C->adv_t1, (u64)wall_clock64());
}

// ------------------------------------------------------------------ seed
// Sets the source bit of every slot in next[] (and value 0 for SSSP slots)
// and appends the sources to the touched list; the filter then turns them
// into iteration 0's frontier exactly as it does for any other iteration.
template <int W>
__global__ __launch_bounds__(BLOCK) void k_seed(KP P, const int32_t* slot_src, int nslots) {
  Ctl* C = P.ctl;
  const int s = (int)(blockIdx.x * BLOCK + threadIdx.x);
  // Stamp of the iteration whose filter settles the seeds (stamp_base itself
  // for the batch seeding pass, which runs at iteration 0xFFFFFFFF).
  const u32 stamp = C->stamp_base + C->iter + 1u;
  bool claimed = false;
  int32_t v = 0;
  if (s < nslots && slot_src[s] >= 0) {
    v = slot_src[s];
    const int w = s >> 6, q = s & 63;
    atomicOr(&P.next[nxt_at<W>(v, w)], 1ull << q);
    if ((P.cfg->val_bits[w] >> q) & 1ull) {
      P.vals[vrow(P.V, w, v, P.cfg->rshift) + q] = 0u;
      atomicOr(&C->near_new, 1u);  // a source at value 0 is below any threshold
    }
    claimed = claim_vertex(P.mark, v, stamp);
  }
  wave_append<int32_t>(claimed, v, &C->tcount.v, P.touched);
}

// ------------------------------------------------------------------ filter
// Vertex i of the filter's list (i < n: the scan's i, or touched[i]), its next
// words and CSR offsets; zeros past the end.  The W next words of a record
// are contiguous and 16-byte aligned (W >= 2).
template <int W>
__device__ __forceinline__ void filter_fetch(const KP& P, u32 i, u32 n, bool scan, int32_t& v, u64 (&nw)[W],
                                             int32_t& o0, int32_t& o1) {
  const bool ok = i < n;
  v = ok ? (scan ? (int32_t)i : P.touched[i]) : 0;
  if (ok) {
    if constexpr (W >= 2) {
      const ulonglong2* q = (const ulonglong2*)&P.next[nxt_at<W>(v, 0)];
#pragma unroll
      for (int h = 0; h < W / 2; ++h) {
        const ulonglong2 x = q[h];
        nw[2 * h] = x.x;
        nw[2 * h + 1] = x.y;
      }
    } else {
      nw[0] = P.next[nxt_at<W>(v, 0)];
    }
    o0 = P.off[v];
    o1 = P.off[v + 1];
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w) nw[w] = 0ull;
    o0 = o1 = 0;
  }
}

// Claims the next words of every touched vertex (sparse mode: the touched
// list; dense mode: every vertex), settles BFS bits (visited |= new, depth =
// iteration+1 - slot start, written as coalesced rows), publishes the
// frontier words, and emits the degree-binned next frontier with wave-level
// prefix sums.  The last workgroup to finish advances the iteration, resets
// the consumed counters, and picks push or pull and the next frontier mode.
template <int W>
__global__ __launch_bounds__(BLOCK) void k_filter(KP P) {
  Ctl* C = P.ctl;
  if (C->finished) return;
  __shared__ int s_last;
  __shared__ u32 s_res[(BLOCK / 64 + 1) * 5];
  if (threadIdx.x == 0) atomicMin(&C->flt_t0, (u64)wall_clock64());
  const Cfg* cfg = P.cfg;
  const u32 iter = C->iter, p = iter & 1u, np = p ^ 1u;
  const bool scan = C->dense_mode != 0;
  const u32 n = scan ? (u32)P.scan_v : C->tcount.v;
  const int rshift = cfg->rshift, dense_t = cfg->dense_t;
  const int lane = lane_id();
  u64 bb[W], vb[W], un[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    bb[w] = cfg->bfs_bits[w];
    vb[w] = cfg->val_bits[w];
    un[w] = 0;
  }
  const u32 depth_base = iter + 1u;
  u64 pairs = 0, epairs = 0, fedges = 0;
  u64 apairs = 0, averts = 0, aqedges = 0;  // every kind: pairs, vertices, query-weighted degree
  // Each thread takes FK vertices per step (their loads issued together) and
  // the workgroup reserves their bin entries with one atomic per bin: the
  // bin counters are single addresses every workgroup contends for (one
  // reservation per 256 vertices cost about 0.3 ms per dense iteration on
  // RMAT-22).
  constexpr int FK = 4;
  const u32 gstride = gridDim.x * BLOCK * FK;
  // The next iteration's bin lists, picked once (indexing the argument's
  // arrays with np inside the loop made the compiler copy the whole
  // argument block to scratch memory).
  int32_t* const l_small = np ? P.small[1] : P.small[0];
  int32_t* const l_medium = np ? P.medium[1] : P.medium[0];
  u64* const l_large = np ? P.large[1] : P.large[0];
  u64* const l_dense = np ? P.dense[1] : P.dense[0];
  u64* const l_dense_n = np ? P.dense_n[1] : P.dense_n[0];
  for (u32 base = blockIdx.x * BLOCK * FK; base < n; base += gstride) {
    int32_t vv[FK], o0[FK], o1[FK];
    u64 nwk[FK][W];
#pragma unroll
    for (int k = 0; k < FK; ++k)
      filter_fetch<W>(P, base + (u32)k * BLOCK + threadIdx.x, n, scan, vv[k], nwk[k], o0[k], o1[k]);
    u32 cnt[5] = {0u, 0u, 0u, 0u, 0u};
    u32 kdeg[FK], kfl[FK];  // per vertex: degree; need_bin | bin << 1 | dense words << 4 | narrow << 12
#pragma unroll
    for (int k = 0; k < FK; ++k) {
      const int32_t v = vv[k];
      bool anyw = false;
#pragma unroll
      for (int w = 0; w < W; ++w) anyw |= nwk[k][w] != 0;
      const int32_t deg = anyw ? o1[k] - o0[k] : 0;
      u64 nb[W];
      bool need_bin = false;
      u32 dmask = 0, nmask = 0, vpairs = 0;  // dense words (wide / narrow)
      u32 allp = 0;                          // (v, query) pairs of every kind
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const u64 nw = nwk[k][w];
        if (anyw) {
          if (nw) P.next[nxt_at<W>(v, w)] = 0;
          P.front[(size_t)v * W + w] = nw;
        }
        nb[w] = nw & bb[w];
        if (nb[w]) P.visited[vis_at<W>(v, w)] |= nb[w];
        const u64 s = nw & vb[w];
        if (s && __popcll(s) >= dense_t) (((cfg->narrow_words >> w) & 1u) ? nmask : dmask) |= 1u << w;
        else if (s) need_bin = true;
        if (nb[w]) need_bin = true;
        un[w] |= nw;
        vpairs += (u32)__popcll(nb[w]);
        allp += (u32)__popcll(nw);
      }
      pairs += vpairs;
      epairs += (u64)vpairs * (u64)deg;
      fedges += (u64)deg;
      apairs += allp;
      averts += anyw ? 1u : 0u;
      aqedges += (u64)allp * (u64)(anyw ? o1[k] - o0[k] : 0);
      if (deg == 0) {
        need_bin = false;
        dmask = 0;
        nmask = 0;
      }
      const u32 b = deg <= P.bin_small ? 0u : (deg <= P.bin_large ? 1u : 2u);
      const u32 dch = (u32)((deg + DCHUNK - 1) / DCHUNK);
      cnt[0] += (need_bin && b == 0) ? 1u : 0u;
      cnt[1] += (need_bin && b == 1) ? 1u : 0u;
      cnt[2] += (need_bin && b == 2) ? (u32)((deg + LCHUNK - 1) / LCHUNK) : 0u;
      cnt[3] += (u32)__popc(dmask) * dch;
      cnt[4] += (u32)__popc(nmask) * dch;
      kdeg[k] = (u32)deg;
      kfl[k] = (need_bin ? 1u : 0u) | (b << 1) | (dmask << 4) | (nmask << 12);
      // BFS depths.  Bit planes (batch runs): the iteration's depth is the
      // same for every new bit, so each set bit b of it ORs the new bits
      // into plane b (one 8-byte atomic per word with new bits, no return).
      if (cfg->depth_planes) {
        for (u32 bits = depth_base; bits; bits &= bits - 1) {
          const int b = __builtin_ctz(bits);
#pragma unroll
          for (int w = 0; w < W; ++w)
            if (nb[w]) {
              const size_t wv = (size_t)w * (size_t)P.V + (size_t)v;
              u64* const a = b < 4 ? cfg->plane4 + (wv << 2) + b : cfg->plane[b] + wv;
              __hip_atomic_fetch_or(a, nb[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
      } else if (rshift == 6) {
        // Value rows (serving, allocation-free runs): lanes of the wave write
        // R-wide rows, 64/R vertices per pass.
        // 64-slot rows: one row store per vertex with new bits, its vertex
        // and new-bit mask read into scalar registers and the mask used as
        // the store's exec mask.
#pragma unroll
        for (int w = 0; w < W; ++w) {
          u64 any = __ballot(nb[w] != 0);
          if (!any) continue;
          const u32 dval = depth_base - (u32)cfg->start[w * 64 + lane];
          const u32 nlo = (u32)nb[w], nhi = (u32)(nb[w] >> 32);
          while (any) {
            const int j = __builtin_ctzll(any);
            any &= any - 1;
            const u64 nbj = ((u64)(u32)__builtin_amdgcn_readlane((int)nhi, j) << 32) |
                            (u32)__builtin_amdgcn_readlane((int)nlo, j);
            const int32_t vj = __builtin_amdgcn_readlane(v, j);
            gu32* row = (gu32*)(P.vals + vrow(P.V, w, vj, 6));
            if (__builtin_amdgcn_inverse_ballot_w64(nbj)) row[lane] = dval;
          }
        }
      } else {
        const int per = 64 >> rshift;
        const int sub = lane >> rshift;
        const int q = lane & ((1 << rshift) - 1);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const u64 any = __ballot(nb[w] != 0);
          if (!any) continue;
          const u32 dval = depth_base - (u32)cfg->start[w * 64 + q];
          for (int j0 = 0; j0 < 64; j0 += per) {
            const u64 grp = per == 64 ? ~0ull : (((1ull << per) - 1) << j0);
            if (!(any & grp)) continue;
            const int j = j0 + sub;
            const u64 nbj = shfl64(nb[w], j);
            const int32_t vj = __shfl(v, j, 64);
            if ((nbj >> q) & 1ull) P.vals[vrow(P.V, w, vj, rshift) + q] = dval;
          }
        }
      }
    }
    u32 at[5];
    block_reserve<5>(cnt, C->bcount[np], at, s_res);
#pragma unroll
    for (int k = 0; k < FK; ++k) {
      const int32_t v = vv[k];
      const u32 deg = kdeg[k], fl = kfl[k];
      const bool need_bin = fl & 1u;
      const u32 b = (fl >> 1) & 3u, dmask = (fl >> 4) & 0xFFu, nmask = (fl >> 12) & 0xFFu;
      if (need_bin && b == 0) l_small[at[0]++] = v;
      if (need_bin && b == 1) l_medium[at[1]++] = v;
      if (need_bin && b == 2) {
        const u32 nch = (deg + LCHUNK - 1) / LCHUNK;
        for (u32 c = 0; c < nch; ++c) l_large[at[2]++] = ((u64)(u32)v << 32) | c;
      }
      if (dmask | nmask) {
        const u32 dch = (deg + DCHUNK - 1) / DCHUNK;
        for (int w = 0; w < W; ++w) {
          const bool wide = (dmask >> w) & 1u, narrow = (nmask >> w) & 1u;
          for (u32 c = 0; c < dch && (wide || narrow); ++c) {
            const u64 it = ((u64)(u32)v << 32) | ((u64)w << 24) | c;
            if (wide) l_dense[at[3]++] = it;
            else l_dense_n[at[4]++] = it;
          }
        }
      }
    }
  }
  // Frontier statistics: wave sums, then one atomic per workgroup and counter.
  {
    constexpr int NS = 6;  // sums per wave before the union words
    __shared__ u64 s_st[(BLOCK / 64) * (NS + MAXW)];
    const int wave = (int)(threadIdx.x >> 6);
    u64 sums[NS] = {pairs, epairs, fedges, apairs, averts, aqedges};
#pragma unroll
    for (int i = 0; i < NS; ++i) sums[i] = wave_sum64(sums[i]);
    u64 unw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) unw[w] = wave_or64(un[w]);
    if (lane == 0) {
      u64* r = s_st + wave * (NS + MAXW);
#pragma unroll
      for (int i = 0; i < NS; ++i) r[i] = sums[i];
#pragma unroll
      for (int w = 0; w < W; ++w) r[NS + w] = unw[w];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      u64 tot[NS], tu[W];
#pragma unroll
      for (int i = 0; i < NS; ++i) tot[i] = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) tu[w] = 0;
      for (int wv = 0; wv < BLOCK / 64; ++wv) {
        const u64* r = s_st + wv * (NS + MAXW);
#pragma unroll
        for (int i = 0; i < NS; ++i) tot[i] += r[i];
#pragma unroll
        for (int w = 0; w < W; ++w) tu[w] |= r[NS + w];
      }
      if (tot[0]) atomicAdd(&C->bfs_pairs[np], tot[0]);
      if (tot[1]) atomicAdd(&C->bfs_edges[np], tot[1]);
      if (tot[2]) atomicAdd(&C->fedges[np], tot[2]);
      if (tot[3]) atomicAdd(&C->npairs[np], tot[3]);
      if (tot[4]) atomicAdd(&C->nverts[np], tot[4]);
      if (tot[5]) atomicAdd(&C->nqedges[np], tot[5]);
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (tu[w]) atomicOr(&C->union_[w], tu[w]);
    }
  }
  // Last workgroup: advance the state machine.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const u32 t = atomicAdd(&C->done_blocks.v, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    __threadfence();
    const u32 total = atomicAdd(&C->bcount[np][0].v, 0u) + atomicAdd(&C->bcount[np][1].v, 0u) +
                      atomicAdd(&C->bcount[np][2].v, 0u) + atomicAdd(&C->bcount[np][3].v, 0u) +
                      atomicAdd(&C->bcount[np][4].v, 0u);
    const u64 fp = atomicAdd(&C->bfs_pairs[np], 0ull), fe = atomicAdd(&C->bfs_edges[np], 0ull);
    const u64 fall = atomicAdd(&C->fedges[np], 0ull);
    const u64 napairs = atomicAdd(&C->npairs[np], 0ull), naverts = atomicAdd(&C->nverts[np], 0ull),
              naqedges = atomicAdd(&C->nqedges[np], 0ull);
    C->dense_mode = fall >= cfg->dense_fedges_thr ? 1u : 0u;
    C->fedges[p] = 0;
    C->npairs[p] = C->nverts[p] = C->nqedges[p] = 0;
    u64 consumed[MAXW];
    for (int w = 0; w < MAXW; ++w) {
      consumed[w] = C->last_union[w];
      C->last_union[w] = atomicAnd(&C->union_[w], 0ull);
    }
    const u32 cur_items = C->bcount[p][0].v + C->bcount[p][1].v + C->bcount[p][2].v + C->bcount[p][3].v + C->bcount[p][4].v;
    C->tcount.v = 0;
    C->bcount[p][0].v = C->bcount[p][1].v = C->bcount[p][2].v = C->bcount[p][3].v = C->bcount[p][4].v = 0;
    for (int k = 0; k < 7; ++k) C->fetch[p][k].v = 0;
    for (int k = 0; k < 5; ++k)
      for (int s = 0; s < NSHARD; ++s) C->shard[p][k][s].v = 0;
    C->bfs_pairs[p] = 0;
    C->bfs_edges[p] = 0;
    C->done_blocks.v = 0;
    const u32 cur_pull = C->bfs_pull;
    u32 pull = 0;
    if (fp > 0) {
      if (cfg->pull_policy == 1) pull = 1;
      else if (cfg->pull_policy == 2) {
        if (C->bfs_pull) pull = fp >= cfg->push_pairs_thr ? 1u : 0u;
        else pull = (fp >= cfg->pull_pairs_thr || fe >= cfg->pull_edges_thr) ? 1u : 0u;
      }
    }
    C->bfs_pull = pull;
    // SSSP pull for the next iteration, on the same ratios applied to the
    // SSSP (vertex, query) pairs of the next frontier (every pair that is
    // not BFS: WCC queries have no frontier) and their degree-weighted count.
    const u32 cur_spull = C->sssp_pull;
    u32 spull = 0;
    {
      const u64 sp = napairs - fp, se = naqedges - fe;
      if (sp > 0) {
        if (cfg->sssp_pull_policy == 1) spull = 1;
        else if (cfg->sssp_pull_policy == 2) {
          if (cur_spull) spull = sp >= cfg->spush_pairs_thr ? 1u : 0u;
          else spull = (sp >= cfg->spull_pairs_thr || se >= cfg->spull_edges_thr) ? 1u : 0u;
        }
      }
    }
    C->sssp_pull = spull;
    // Near-far: when no SSSP pair below the threshold is left, move it to
    // the smallest pending value plus delta (skipping empty buckets).
    if (cfg->delta != 0) {
      const u32 fm = C->far_min;
      if (C->near_new == 0 && fm != INF) {
        const u64 nt = (u64)fm + (u64)cfg->delta;
        C->thr = nt >= (u64)INF ? INF : (u32)nt;
        ++C->bumps;
      }
      C->near_new = 0;
      C->far_min = INF;
    }
    // Launch timing and the iteration record: the advance of this iteration
    // ran iff iter != seeding.
    if (iter != 0xFFFFFFFFu) {
      const u64 at = C->adv_t1 - C->adv_t0, ft = (u64)wall_clock64() - C->flt_t0;
      const u64 dt = C->den_t1 > C->den_t0 ? C->den_t1 - C->den_t0 : 0ull;
      C->adv_ticks += at;
      C->adv_launches += 1;
      C->flt_ticks += ft;
      if (dt) {
        C->den_ticks += dt;
        C->den_launches += 1;
      }
      if (P.iters) {
        IterRec& r = P.iters[iter % ITER_RING];
        r.iter = iter;
        r.pull = cur_pull | (cur_spull << 1);
        r.dense = scan ? 1u : 0u;
        r.frontier_items = cur_items;
        r.touched = n;
        r.next_items = total;
        r.next_edges = fall;
        r.next_bfs_pairs = fp;
        r.adv_ticks = at;
        r.flt_ticks = ft;
        r.den_ticks = dt;
        r.next_pairs = napairs;
        r.next_vertices = naverts;
        r.next_qedges = naqedges;
        r.edges_scanned = C->prof[PF_PUSH_EDGES] + C->prof[PF_DENSE_EDGES] + C->prof[PF_PULL_EDGES];
        r.t_end = (u64)wall_clock64();
        for (int w = 0; w < MAXW; ++w) r.fr_union[w] = consumed[w];
      }
    } else {
      // The seeding filter: the sizes of frontier 0, and the run's clock origin.
      C->seed_pairs = napairs;
      C->seed_verts = naverts;
      C->seed_edges = fall;
      C->seed_qedges = naqedges;
      C->t_start = (u64)wall_clock64();
    }
    C->adv_t0 = ~0ull;
    C->adv_t1 = 0;
    C->den_t0 = ~0ull;
    C->den_t1 = 0;
    C->flt_t0 = ~0ull;
    const u32 next_iter = iter + 1u;
    C->iter = next_iter;
    if (total == 0 || (C->max_iter != 0 && next_iter >= C->max_iter)) {
      C->finished = 1;
    } else {
      if (pull || spull) ++C->pull_iters;
      else ++C->push_iters;
      if (C->dense_mode) ++C->dense_iters;
    }
    __threadfence();
  }
}

// ------------------------------------------------------------------ utilities
__global__ void k_fill_u32(u32* p, size_t n, u32 val) {
  const size_t stride = (size_t)gridDim.x * BLOCK;
  size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  // 16-byte stores over the aligned body.
  uint4* p4 = reinterpret_cast<uint4*>(p);
  const size_t n4 = n / 4;
  const uint4 v4 = make_uint4(val, val, val, val);
  for (size_t k = i; k < n4; k += stride) p4[k] = v4;
  for (size_t k = n4 * 4 + i; k < n; k += stride) p[k] = val;
}

// Device-to-device copy of n 16-byte elements (each read and written once).
__global__ __launch_bounds__(BLOCK) void k_copy16(uint4* dst, const uint4* src, size_t n) {
  const size_t stride = (size_t)gridDim.x * BLOCK;
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += stride) dst[k] = src[k];
}

// Fills the first 4 << l4 lanes of each of nrows 64-lane value rows with
// val (16-byte stores).
__global__ void k_fill_row_prefix(u32* p, size_t nrows, int l4, u32 val) {
  const size_t n = nrows << l4;
  const size_t stride = (size_t)gridDim.x * BLOCK;
  uint4* p4 = reinterpret_cast<uint4*>(p);
  const uint4 v4 = make_uint4(val, val, val, val);
  const size_t cmask = ((size_t)1 << l4) - 1;
  for (size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x; k < n; k += stride)
    p4[((k >> l4) << 4) + (k & cmask)] = v4;
}

// BFS slots of each word of a pass: their value-row entries are written only
// where the slot's visited bit is set (the depth, by the compaction), so the
// output takes CGP_INF wherever the bit is clear.
struct OutMasks {
  u64 bfs[MAXW];
  u64 val[MAXW];
  const u64* plane4;            // depth planes 0-3, in 32-byte (word, vertex) blocks
  const u64* plane[MAXPLANES];  // depth planes 4.. (plane-major)
  int n_planes;                 // planes in use (0: depths in the value rows)
};

// Writes the query-major output rows of word w = blockIdx.y: out[qmap[w*64+q]
// * V + v], through a 64-slot x 64-vertex LDS tile, v in the caller's
// numbering (the tile's internal vertices are inv[v]; inv null: the same).
// SSSP slots take their value rows.  BFS slots take CGP_INF wherever their
// visited bit is clear and otherwise their depth: from the depth bit planes
// when there are any (only the value-row lines holding SSSP slots are then
// read), else from the value rows; each thread resolves sixteen slots of one
// vertex into the tile, so the store pass only copies the tile out.  One
// workgroup per (64-vertex tile, word): each output row segment is OUT_VT *
// 4 = 256 bytes (128-vertex tiles measured slower: 4.5 vs 2.9 ms on RMAT-22;
// 16-byte global accesses took 2.9 to 1.55 ms).
constexpr int OUT_VT = 64;
__global__ __launch_bounds__(BLOCK) void k_output(const u32* vals, const u64* visited, const int32_t* inv,
                                                  int32_t V, int W, int rshift, const int32_t* qmap, OutMasks m,
                                                  u32* out, int vec) {
  __shared__ u32 tile[64][OUT_VT + 1];
  __shared__ u64 s_vis[OUT_VT];
  __shared__ int32_t s_u[OUT_VT];
  extern __shared__ u64 s_pl[];  // [n_planes][OUT_VT], sized at launch
  const int R = 1 << rshift;
  const int w = (int)blockIdx.y;
  const int32_t v0 = (int32_t)blockIdx.x * OUT_VT;
  const int nv = min(OUT_VT, V - v0);
  const int np = m.n_planes;
  const u64 bm = m.bfs[w];
  const u64 rd = np ? m.val[w] : (m.val[w] | bm);  // slots whose value rows are read
  if ((int)threadIdx.x < OUT_VT) {
    const int j = (int)threadIdx.x;
    s_u[j] = j < nv ? (inv ? inv[v0 + j] : v0 + j) : 0;
  }
  __syncthreads();
  if (bm) {
    for (int t = threadIdx.x; t < (np + 1) * OUT_VT; t += BLOCK) {
      const int b = t / OUT_VT, j = t % OUT_VT;
      if (j >= nv) continue;
      const size_t u = (size_t)s_u[j];
      if (b == 0) s_vis[j] = visited[u * (2 * W) + w];
      else if (b <= 4) s_pl[(b - 1) * OUT_VT + j] = m.plane4[(((size_t)w * (size_t)V + u) << 2) + (b - 1)];
      else s_pl[(b - 1) * OUT_VT + j] = m.plane[b - 1][(size_t)w * (size_t)V + u];
    }
  }
  if (vec) {
    // 64-slot rows, V % 4 == 0 and a 16-byte aligned output (checked by the
    // launcher): 16-byte loads of 4 slots of a vertex.
    const int sub = (int)(threadIdx.x >> 4), c4 = (int)(threadIdx.x & 15) * 4;
    if ((rd >> c4) & 0xFull) {
#pragma unroll
      for (int pass = 0; pass < 4; ++pass) {
        const int j = pass * 16 + sub;
        if (j < nv) {
          const uint4 x = *(const uint4*)(vals + vrow(V, w, s_u[j], 6) + c4);
          tile[c4 + 0][j] = x.x;
          tile[c4 + 1][j] = x.y;
          tile[c4 + 2][j] = x.z;
          tile[c4 + 3][j] = x.w;
        }
      }
    }
  } else {
    for (int t = threadIdx.x; t < OUT_VT * R; t += BLOCK) {
      const int j = t >> rshift, q = t & (R - 1);
      if (j < nv && ((rd >> q) & 1ull)) tile[q][j] = vals[vrow(V, w, s_u[j], rshift) + (size_t)q];
    }
  }
  __syncthreads();
  if (bm) {
    // Thread (j, g): BFS slots 16g..16g+15 of vertex j into the tile.
    const int j = (int)(threadIdx.x & 63), g = (int)(threadIdx.x >> 6);
    if (j < nv && g * 16 < R) {
      const u64 vis = s_vis[j];
      // The vertex's first eight planes in registers (depths below 256);
      // deeper planes are read from LDS.
      constexpr int NPR = 8;
      u64 pl[NPR];
#pragma unroll
      for (int b = 0; b < NPR; ++b) pl[b] = b < np ? s_pl[b * OUT_VT + j] : 0ull;
      for (int q = g * 16; q < min(R, g * 16 + 16); ++q) {
        if (!((bm >> q) & 1ull)) continue;
        u32 d;
        if (!((vis >> q) & 1ull)) {
          d = INF;
        } else if (np) {
          d = 0;
#pragma unroll
          for (int b = 0; b < NPR; ++b) d |= (u32)((pl[b] >> q) & 1ull) << b;
          for (int b = NPR; b < np; ++b) d |= (u32)((s_pl[b * OUT_VT + j] >> q) & 1ull) << b;
        } else {
          d = tile[q][j];
        }
        tile[q][j] = d;
      }
    }
    __syncthreads();
  }
  if (vec) {
    // 16-byte stores of 4 vertices of a query.
    const int sub = (int)(threadIdx.x >> 4), c4 = (int)(threadIdx.x & 15) * 4;
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
      const int q = pass * 16 + sub;
      const int32_t dst = qmap[w * 64 + q];
      if (dst < 0) continue;
      u32* o = out + (size_t)dst * V + v0 + c4;
      if (c4 + 4 <= nv) {
        *(uint4*)o = make_uint4(tile[q][c4], tile[q][c4 + 1], tile[q][c4 + 2], tile[q][c4 + 3]);
      } else {
        for (int i = 0; c4 + i < nv; ++i) o[i] = tile[q][c4 + i];
      }
    }
  } else {
    for (int t = threadIdx.x; t < OUT_VT * R; t += BLOCK) {
      const int q = t / OUT_VT, j = t % OUT_VT;
      const int32_t dst = qmap[w * 64 + q];
      if (dst >= 0 && j < nv) out[(size_t)dst * V + v0 + j] = tile[q][j];
    }
  }
}

// rows[e] = source vertex of CSR slot e (one wave per vertex row).
__global__ __launch_bounds__(BLOCK) void k_fill_rows(const int32_t* off, int32_t V, int32_t* rows) {
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * BLOCK) >> 6;
  for (int64_t v = wave; v < V; v += nw)
    for (int32_t e = off[v] + lane_id(); e < off[v + 1]; e += 64) rows[e] = (int32_t)v;
}

// ------------------------------------------------- graphs over caller arrays
// Checks of a caller's device CSR before the engine reads it, and the
// derived per-graph arrays built on the device.  Check flags: bit 0 offsets
// decrease somewhere, bit 1 a column index is out of [0, V), bit 2 some row
// is not sorted, bit 3 some edge has no reverse edge (only tested on sorted
// rows: a binary search in the target's row).
enum : u32 { CSR_BAD_OFFSETS = 1, CSR_BAD_INDEX = 2, CSR_UNSORTED = 4, CSR_ASYMMETRIC = 8, CSR_ASYM_WEIGHTS = 16 };

// Flags decreasing offsets and takes the largest degree (into *maxdeg).
__global__ __launch_bounds__(BLOCK) void k_csr_check_offsets(const int32_t* off, int32_t V, u32* flags,
                                                             u32* maxdeg) {
  bool bad = false;
  int32_t m = 0;
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t d = off[v + 1] - off[v];
    bad |= d < 0;
    m = max(m, d);
  }
  if (__ballot(bad) && lane_id() == 0) atomicOr(flags, (u32)CSR_BAD_OFFSETS);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if (lane_id() == 0 && m > 0) atomicMax(maxdeg, (u32)m);
}

// Edge-parallel over slots whose source row `rows` is known (offsets valid).
__global__ __launch_bounds__(BLOCK) void k_csr_check_edges(const int32_t* rows, const int32_t* idx, int64_t E,
                                                           int32_t V, u32* flags) {
  u32 f = 0;
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < E; e += (int64_t)gridDim.x * BLOCK) {
    const int32_t x = idx[e];
    if (x < 0 || x >= V) f |= CSR_BAD_INDEX;
    if (e > 0 && rows[e - 1] == rows[e] && idx[e - 1] > x) f |= CSR_UNSORTED;
  }
  if (__ballot(f != 0)) {
    u32 m = f;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m |= (u32)__shfl_xor((int)m, d, 64);
    if (lane_id() == 0) atomicOr(flags, m);
  }
}

// Every edge u -> v has v -> u (rows sorted, indices in range: checked
// before), and with weights, one of the same weight.
__global__ __launch_bounds__(BLOCK) void k_csr_check_symmetric(const int32_t* off, const int32_t* rows,
                                                               const int32_t* idx, const u32* wt, int64_t E,
                                                               u32* flags) {
  bool miss = false, wmiss = false;
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < E; e += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = rows[e], v = idx[e];
    int32_t lo = off[v], hi = off[v + 1];
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (idx[mid] < u) lo = mid + 1;
      else hi = mid;
    }
    const bool has = lo < off[v + 1] && idx[lo] == u;
    miss |= !has;
    if (has && wt) {
      bool same = false;
      for (int32_t x = lo; x < off[v + 1] && idx[x] == u && !same; ++x) same = wt[x] == wt[e];
      wmiss |= !same;
    }
  }
  if (__ballot(miss) && lane_id() == 0) atomicOr(flags, (u32)CSR_ASYMMETRIC);
  if (__ballot(wmiss) && lane_id() == 0) atomicOr(flags, (u32)CSR_ASYM_WEIGHTS);
}

__global__ __launch_bounds__(BLOCK) void k_sum_u32(const u32* x, int64_t n, u64* sum) {
  u64 s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) s += x[i];
  s = wave_sum64(s);
  if (lane_id() == 0 && s) atomicAdd(sum, s);
}

// Vertices with degree above PULL_WAVE_DEG, and those with 1..PULL_WAVE_DEG,
// into two lists (counts in cnt[0] (lo) and cnt[1] (hi)).
__global__ __launch_bounds__(BLOCK) void k_degree_split(const int32_t* off, int32_t V, int32_t* lo, int32_t* hi,
                                                        u32* cnt) {
  for (int64_t b = (int64_t)blockIdx.x * BLOCK; b < V; b += (int64_t)gridDim.x * BLOCK) {
    const int64_t v = b + threadIdx.x;
    const int32_t d = v < V ? off[v + 1] - off[v] : 0;
    wave_append<int32_t>(d > PULL_WAVE_DEG, (int32_t)v, &cnt[1], hi);
    wave_append<int32_t>(d > 0 && d <= PULL_WAVE_DEG, (int32_t)v, &cnt[0], lo);
  }
}

// --------------------------------------------------------------- WCC
// Hooking plus pointer jumping over parent pointers with parent[x] <= x at all
// times: a hook links the larger of two roots under the smaller with a
// compare-and-swap, so the root of every tree is its smallest vertex and the
// final labels are the minimum vertex id of each weakly connected component.
__device__ __forceinline__ int32_t load_parent(int32_t* parent, int32_t x) {
  return __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (each visited node is re-pointed to its grandparent).
__device__ __forceinline__ int32_t find_root(int32_t* parent, int32_t x) {
  int32_t p = load_parent(parent, x);
  while (p != x) {
    const int32_t gp = load_parent(parent, p);
    if (gp != p) __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
    p = gp;
    if (p == x) break;
    p = load_parent(parent, x);
  }
  return x;
}

__global__ void k_wcc_init(int32_t* parent, int32_t V) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK)
    parent[v] = (int32_t)v;
}

// Unites the components of a and b: the larger root is linked under the
// smaller one with a compare-and-swap, retrying from the current roots when
// another thread got there first, so parent[x] <= x always holds and every
// tree's root is its minimum vertex.
__device__ __forceinline__ void wcc_link(int32_t* parent, int32_t a, int32_t b) {
  int32_t ra = find_root(parent, a), rb = find_root(parent, b);
  while (ra != rb) {
    const int32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    const int32_t old = atomicCAS(&parent[hi], hi, lo);
    if (old == hi) break;
    // hi stopped being a root: continue from its current ancestors.
    ra = find_root(parent, old);
    rb = find_root(parent, lo);
  }
}

__global__ __launch_bounds__(BLOCK) void k_wcc_hook(const int32_t* rows, const int32_t* idx, int64_t E,
                                                    int32_t* parent, int half) {
  for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < E; e += (int64_t)gridDim.x * BLOCK) {
    const int32_t a = rows[e], b = idx[e];
    if (a == b || (half && a < b)) continue;  // symmetric: each undirected edge once
    wcc_link(parent, a, b);
  }
}

// Afforest-style sampling (Sutton, Ben-Nun, Barak, "Optimizing Parallel
// Graph Connectivity Computation via Subgraph Sampling", IPDPS 2018): link
// every vertex with its first `k` neighbours.
__global__ __launch_bounds__(BLOCK) void k_wcc_sample(const int32_t* off, const int32_t* idx, int32_t V,
                                                      int k, int32_t* parent) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t e0 = off[v], e1 = min(off[v + 1], e0 + k);
    for (int32_t e = e0; e < e1; ++e) wcc_link(parent, (int32_t)v, idx[e]);
  }
}

// The remaining edges (from the k-th on) of every vertex outside the
// component rooted at `skip`.  On a symmetric graph an edge between that
// component and another vertex is also stored at the other vertex, which
// processes it, so the skipped vertices lose nothing.
__global__ __launch_bounds__(BLOCK) void k_wcc_rest(const int32_t* off, const int32_t* idx, int32_t V, int k,
                                                    int32_t skip, int32_t* parent) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t e0 = off[v] + k, e1 = off[v + 1];
    if (e0 >= e1 || find_root(parent, (int32_t)v) == skip) continue;
    for (int32_t e = e0; e < e1; ++e) wcc_link(parent, (int32_t)v, idx[e]);
  }
}

// parent of n pseudo-random vertices (the host picks the most frequent root).
__global__ void k_wcc_probe(const int32_t* parent, int32_t V, int n, int32_t* out) {
  const int i = (int)(blockIdx.x * BLOCK + threadIdx.x);
  if (i < n) {
    const u32 h = (u32)i * 2654435761u ^ 0x9E3779B9u;
    out[i] = parent[(int32_t)(h % (u32)V)];
  }
}

__global__ void k_wcc_compress(int32_t* parent, int32_t V) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    int32_t r = (int32_t)v;
    int32_t p = parent[r];
    while (p != r) {
      r = p;
      p = parent[r];
    }
    parent[v] = r;
  }
}

// Renumbered graphs: minid[root] = the smallest caller id in the root's
// component (minid starts at INT32_MAX).  Caller ids are taken in order; the
// lanes of a wave that share the root of its first lane in a component of
// more than one vertex (the large component, on the graphs that have one)
// reduce their minimum first, and every update is skipped when the current
// minimum is already smaller, so the large component's root is not hit by
// one atomic per vertex.
__global__ __launch_bounds__(BLOCK) void k_wcc_minid(const int32_t* parent, const int32_t* inv, int32_t V,
                                                     int32_t* minid) {
  for (int64_t b = (int64_t)blockIdx.x * BLOCK; b < V; b += (int64_t)gridDim.x * BLOCK) {
    const int64_t v = b + threadIdx.x;
    const bool ok = v < V;
    const int32_t u = ok ? inv[v] : 0;
    const int32_t r = ok ? parent[u] : -1;
    const u64 shared = __ballot(ok && r != u);
    const int32_t rd = shared ? __builtin_amdgcn_readlane(r, __builtin_ctzll(shared)) : -1;
    const bool mine = ok && r == rd;
    int32_t m = mine ? (int32_t)v : 0x7FFFFFFF;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = min(m, __shfl_xor(m, d, 64));
    if (mine && (int64_t)__builtin_amdgcn_readfirstlane(__builtin_ctzll(__ballot(mine))) == (int64_t)lane_id()) {
      if (__hip_atomic_load(&minid[rd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > m) atomicMin(&minid[rd], m);
    }
    if (ok && !mine && __hip_atomic_load(&minid[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > (int32_t)v)
      atomicMin(&minid[r], (int32_t)v);
  }
}

// Writes the component labels (the smallest vertex id of each component, in
// the caller's numbering) into every WCC output row (rows_out null: one row
// at out).  Vertex v of the caller is inv[v] inside (inv null: the same);
// minid (null: the root itself) maps an internal root to its label.
__global__ void k_wcc_rows(const int32_t* parent, const int32_t* minid, const int32_t* inv, int32_t V,
                           const int32_t* rows_out, int nrows, u32* out) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t root = parent[inv ? inv[v] : (int32_t)v];
    const u32 l = (u32)(minid ? minid[root] : root);
    for (int r = 0; r < nrows; ++r) out[(size_t)(rows_out ? rows_out[r] : 0) * V + v] = l;
  }
}

// m_q per output row: sum of out-degrees over vertices with a finite value.
__global__ __launch_bounds__(BLOCK) void k_reached(const int32_t* off, const int32_t* inv, int32_t V, const u32* out,
                                                   u64* m) {
  const int r = blockIdx.y;
  u64 acc = 0;
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK)
    if (out[(size_t)r * V + v] != INF) {
      const int32_t u = inv ? inv[v] : (int32_t)v;  // the caller's vertex v inside the graph
      acc += (u64)(off[u + 1] - off[u]);
    }
  acc = wave_sum64(acc);
  if (lane_id() == 0 && acc) atomicAdd(&m[r], acc);
}

}  // namespace

// ============================================================== host side
namespace {

// Both frontier-mode instantiations of the lane-per-query advance (the one
// not matching the device-selected mode exits at once), U = rows per step.
// It is wave-level work, so it always runs 256-thread workgroups.
template <int W, int U>
void launch_dense_u(hipStream_t s, dim3 grid, const KP& P, int tab) {
  if (tab == 1) {
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, false, U, 1>), grid, dim3(BLOCK), 0, s, P);
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, true, U, 1>), grid, dim3(BLOCK), 0, s, P);
  } else if (tab == 2) {
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, false, U, 2>), grid, dim3(BLOCK), 0, s, P);
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, true, U, 2>), grid, dim3(BLOCK), 0, s, P);
  } else {
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, false, U, 0>), grid, dim3(BLOCK), 0, s, P);
    hipLaunchKernelGGL((k_adv_dense<W, BLOCK, true, U, 0>), grid, dim3(BLOCK), 0, s, P);
  }
}

// Optional advance launches of the current slot configuration: bit 0, the
// lane-per-query kernels (some word has SSSP slots); bit 1, the narrow
// kernel (some word is narrow now, or was when the last filter binned the
// frontier this advance expands: serving admission can end the narrow
// regime between a filter and the next advance); bit 2, the SSSP pull
// kernel (SSSP pull steps enabled).  The filter routes no item to a kernel
// that is not launched, and a BFS-only batch saves three launches per
// iteration.
int launch_set(const cgp_engine* en) {
  const Cfg& c = *en->h_cfg;
  int ls = 0;
  for (int w = 0; w < MAXW; ++w)
    if (c.val_bits[w]) ls |= 1;
  if (c.narrow_words || en->narrow_binned) ls |= 2;
  if (c.sssp_pull_policy) ls |= 4;
  return ls;
}

// Lane-per-query loop of an engine: the table loop when postponing is off,
// with 32-bit byte offsets when one word's value rows (at the widest row,
// 64 slots) span less than 4 GiB; the postponing loop otherwise.
int dense_tab_mode(const cgp_engine* en) {
  if (en->opts.sssp_delta != 0) return 0;
  return !en->opts.dense_wide_addr && (u64)en->g->V * 256ull < (1ull << 32) ? 1 : 2;
}

template <int W>
void launch_dense(hipStream_t s, int blocks, const KP& P, int unroll, int tab) {
  const dim3 grid((unsigned)blocks);
  if (unroll == 4) launch_dense_u<W, 4>(s, grid, P, tab);
  else if (unroll == 16) launch_dense_u<W, 16>(s, grid, P, tab);
  else launch_dense_u<W, 8>(s, grid, P, tab);
}

// One iteration: advance then filter (both time themselves on the device clock).
template <int W>
void launch_iteration(cgp_engine* en, const KP& P) {
  hipStream_t s = en->s;
  const int bs = en->opts.block_size;
  const dim3 grid(en->grid_advance());
  // The lane-per-query launches and the push launch touch disjoint work
  // (atomics only on shared words), so with concurrent_advance the push
  // launch runs on a second stream forked from and joined back into the
  // engine stream (also inside a captured graph): each fills the other's
  // tail.
  hipStream_t sp = s;
  if (en->opts.concurrent_advance) {
    HIP_CHECK(hipEventRecord(en->ev_fork, s));
    HIP_CHECK(hipStreamWaitEvent(en->s2, en->ev_fork, 0));
    sp = en->s2;
  }
  if (bs == 128) {
    hipLaunchKernelGGL((k_adv_push<W, 128>), grid, dim3(128), 0, sp, P);
  } else if (bs == 512) {
    hipLaunchKernelGGL((k_adv_push<W, 512>), grid, dim3(512), 0, sp, P);
  } else {
    hipLaunchKernelGGL((k_adv_push<W, 256>), grid, dim3(256), 0, sp, P);
  }
  const int ls = launch_set(en);
  if (ls & 4) hipLaunchKernelGGL((k_pull_sssp<W>), dim3(en->advance_blocks()), dim3(BLOCK), 0, s, P);
  if (ls & 1) launch_dense<W>(s, en->advance_blocks(), P, en->opts.dense_unroll, dense_tab_mode(en));
  if (ls & 2) hipLaunchKernelGGL((k_adv_narrow<W>), dim3(en->advance_blocks()), dim3(BLOCK), 0, s, P);
  if (en->opts.concurrent_advance) {
    HIP_CHECK(hipEventRecord(en->ev_join, en->s2));
    HIP_CHECK(hipStreamWaitEvent(s, en->ev_join, 0));
  }
  hipLaunchKernelGGL((k_filter<W>), dim3(en->grid_filter()), dim3(BLOCK), 0, s, P);
  HIP_CHECK(hipGetLastError());
  en->narrow_binned = en->h_cfg->narrow_words;
}

template <int W>
void launch_seed(cgp_engine* en, const KP& P, const int32_t* d_slot_src, int nslots) {
  hipLaunchKernelGGL((k_seed<W>), dim3((nslots + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, en->s, P,
                     d_slot_src, nslots);
  hipLaunchKernelGGL((k_filter<W>), dim3(en->grid_filter()), dim3(BLOCK), 0, en->s, P);
  HIP_CHECK(hipGetLastError());
}

template <int W>
void launch_seed_only(cgp_engine* en, const KP& P, const int32_t* d_slot_src, int nslots) {
  hipLaunchKernelGGL((k_seed<W>), dim3((nslots + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, en->s, P,
                     d_slot_src, nslots);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

namespace cgpi {

void dispatch_iteration(cgp_engine* en, int W, const KP& P) {
  switch (W) {
    case 1: launch_iteration<1>(en, P); break;
    case 2: launch_iteration<2>(en, P); break;
    case 4: launch_iteration<4>(en, P); break;
    default: launch_iteration<8>(en, P); break;
  }
}

void dispatch_seed_only(cgp_engine* en, int W, const KP& P, const int32_t* src, int nslots) {
  switch (W) {
    case 1: launch_seed_only<1>(en, P, src, nslots); break;
    case 2: launch_seed_only<2>(en, P, src, nslots); break;
    case 4: launch_seed_only<4>(en, P, src, nslots); break;
    default: launch_seed_only<8>(en, P, src, nslots); break;
  }
}

void dispatch_seed(cgp_engine* en, int W, const KP& P, const int32_t* src, int nslots) {
  switch (W) {
    case 1: launch_seed<1>(en, P, src, nslots); break;
    case 2: launch_seed<2>(en, P, src, nslots); break;
    case 4: launch_seed<4>(en, P, src, nslots); break;
    default: launch_seed<8>(en, P, src, nslots); break;
  }
}

u32 narrow_mask(const Cfg& cfg, int W) {
  // Words whose SSSP slots all sit in lanes 0..15 can relax four edges per
  // lane-per-query step (phase N).  Measured on RMAT-22: worth it when every
  // SSSP word of the pass is narrow (16 SSSP queries: 55 -> 36 ms), a loss
  // when a full SSSP word runs beside the narrow one (80 SSSP: 86 -> 94 ms),
  // so the mask is all or nothing.
  u32 m = 0;
  for (int w = 0; w < W; ++w) {
    if (!cfg.val_bits[w]) continue;
    if (cfg.val_bits[w] >> 16) return 0u;
    m |= 1u << w;
  }
  return m;
}

}  // namespace cgpi

// SSSP pull items of the engine's graph: (vertex, PCHUNK-edge chunk) of every
// vertex with edges, built once on the device.
void cgp_engine::ensure_pull_items() {
  if (pull_items.p) return;
  const size_t cap = (size_t)g->V + (size_t)g->E / PCHUNK + 1;
  pull_items.ensure(cap * 8 + 8);
  u32* d_count = (u32*)((u64*)pull_items.p + cap);
  HIP_CHECK(hipMemsetAsync(d_count, 0, 4, s));
  hipLaunchKernelGGL(k_pull_items, dim3(cus * 8), dim3(BLOCK), 0, s, (const int32_t*)g->off, g->V,
                     (u64*)pull_items.p, d_count);
  HIP_CHECK(hipGetLastError());
  u32 count = 0;
  HIP_CHECK(hipMemcpyAsync(&count, d_count, 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  n_pull_items = count;
  drop_graphs();
}

namespace cgpi {

void fill_u32(hipStream_t s, void* p, size_t n, u32 v, int cus) {
  if (n == 0) return;
  size_t blocks = std::min<size_t>((n / 4 + BLOCK - 1) / BLOCK + 1, (size_t)cus * 16);
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(BLOCK), 0, s, (u32*)p, n, v);
  HIP_CHECK(hipGetLastError());
}

void fill_row_prefix(hipStream_t s, void* p, size_t nrows, int l4, u32 v, int cus) {
  if (nrows == 0) return;
  size_t blocks = std::min<size_t>(((nrows << l4) + BLOCK - 1) / BLOCK, (size_t)cus * 16);
  hipLaunchKernelGGL(k_fill_row_prefix, dim3((unsigned)blocks), dim3(BLOCK), 0, s, (u32*)p, nrows, l4, v);
  HIP_CHECK(hipGetLastError());
}

static int wlog(int W) { return W == 1 ? 0 : W == 2 ? 1 : W == 4 ? 2 : 3; }

void compute_wcc(cgp_engine* en) {
  const cgp_graph* g = en->g;
  en->parent.ensure((size_t)g->V * 4 + 4);
  int32_t* parent = (int32_t*)en->parent.p;
  const int gb = en->cus * 16;
  hipLaunchKernelGGL(k_wcc_init, dim3(gb), dim3(BLOCK), 0, en->s, parent, g->V);
  if (!g->symmetric || g->V < 4096) {
    // Directed graphs (weak components need every stored edge) and tiny
    // graphs: one hooking pass over all edges.
    hipLaunchKernelGGL(k_wcc_hook, dim3(gb), dim3(BLOCK), 0, en->s, (const int32_t*)g->rows,
                       (const int32_t*)g->idx, g->E, parent, g->symmetric ? 1 : 0);
  } else {
    // Sample two neighbours per vertex, find the component most vertices
    // joined, then only the vertices outside it link their other edges.
    constexpr int kSample = 2, kProbe = 1024;
    hipLaunchKernelGGL(k_wcc_sample, dim3(gb), dim3(BLOCK), 0, en->s, (const int32_t*)g->off,
                       (const int32_t*)g->idx, g->V, kSample, parent);
    hipLaunchKernelGGL(k_wcc_compress, dim3(gb), dim3(BLOCK), 0, en->s, parent, g->V);
    en->wcc_probe.ensure(kProbe * 4);
    hipLaunchKernelGGL(k_wcc_probe, dim3(kProbe / BLOCK), dim3(BLOCK), 0, en->s, (const int32_t*)parent, g->V,
                       kProbe, (int32_t*)en->wcc_probe.p);
    std::array<int32_t, kProbe> probe;
    HIP_CHECK(hipMemcpyAsync(probe.data(), en->wcc_probe.p, kProbe * 4, hipMemcpyDeviceToHost, en->s));
    HIP_CHECK(hipStreamSynchronize(en->s));
    std::sort(probe.begin(), probe.end());
    int32_t skip = probe[0];
    int best = 0;
    for (int i = 0; i < kProbe;) {
      int j = i;
      while (j < kProbe && probe[j] == probe[i]) ++j;
      if (j - i > best) {
        best = j - i;
        skip = probe[i];
      }
      i = j;
    }
    hipLaunchKernelGGL(k_wcc_rest, dim3(gb), dim3(BLOCK), 0, en->s, (const int32_t*)g->off,
                       (const int32_t*)g->idx, g->V, kSample, skip, parent);
  }
  hipLaunchKernelGGL(k_wcc_compress, dim3(gb), dim3(BLOCK), 0, en->s, parent, g->V);
  HIP_CHECK(hipGetLastError());
}

// Writes the labels of compute_wcc's forest into nrows output rows (d_rows
// null: one row at out).  On renumbered graphs the smallest caller id of each
// component is found first, in `minid` or, when that is null, in the touched
// list's buffer (free once a batch pass has finished its iterations).
void write_wcc_rows(cgp_engine* en, const int32_t* d_rows, int nrows, u32* out, int32_t* minid) {
  const cgp_graph* g = en->g;
  if (!g->inv) {
    minid = nullptr;
  } else {
    if (!minid) {
      en->touched.ensure((size_t)g->V * 4 + 4);
      minid = (int32_t*)en->touched.p;
    }
    fill_u32(en->s, minid, (size_t)g->V, 0x7FFFFFFFu, en->cus);
    hipLaunchKernelGGL(k_wcc_minid, dim3(en->cus * 16), dim3(BLOCK), 0, en->s, (const int32_t*)en->parent.p,
                       (const int32_t*)g->inv, g->V, minid);
  }
  hipLaunchKernelGGL(k_wcc_rows, dim3(en->cus * 16), dim3(BLOCK), 0, en->s, (const int32_t*)en->parent.p,
                     (const int32_t*)minid, (const int32_t*)g->inv, g->V, d_rows, nrows, out);
  HIP_CHECK(hipGetLastError());
}

// Depth bit planes [plane_count, need) of the running pass: allocated (a
// word of V bits per mask word), zeroed over the pass's BFS words, and
// published in the slot configuration (on the device too when `live`: the
// pass is running and its configuration was uploaded).
static void add_depth_planes(cgp_engine* en, int W, int need, bool live) {
  const size_t V = (size_t)en->g->V;
  Cfg& cfg = *en->h_cfg;
  need = std::min(need, MAXPLANES);
  for (int b = en->plane_count; b < need; ++b) {
    if (b < 4) {
      // Planes 0-3 share 32-byte (word, vertex) blocks: one allocation,
      // zeroed over the BFS words with the first of them.
      if (b == 0) {
        en->plane4.ensure(V * (size_t)W * 32);
        u64* p4 = (u64*)en->plane4.p;
        for (int w = 0; w < W; ++w)
          if ((en->plane_words >> w) & 1u) fill_u32(en->s, p4 + (size_t)w * V * 4, V * 8, 0u, en->cus);
        cfg.plane4 = p4;
        if (live)
          HIP_CHECK(hipMemcpyAsync((char*)en->d_cfg + offsetof(Cfg, plane4), &cfg.plane4, sizeof(u64*),
                                   hipMemcpyHostToDevice, en->s));
      }
      continue;
    }
    en->plane[b].ensure(V * (size_t)W * 8);
    u64* pl = (u64*)en->plane[b].p;
    for (int w = 0; w < W; ++w)
      if ((en->plane_words >> w) & 1u) fill_u32(en->s, pl + (size_t)w * V, V * 2, 0u, en->cus);
    cfg.plane[b] = pl;
    if (live)
      HIP_CHECK(hipMemcpyAsync((char*)en->d_cfg + offsetof(Cfg, plane) + (size_t)b * sizeof(u64*), &cfg.plane[b],
                               sizeof(u64*), hipMemcpyHostToDevice, en->s));
  }
  en->plane_count = std::max(en->plane_count, need);
}

static int bit_length(u64 x) { return x ? 64 - __builtin_clzll(x) : 0; }

// Runs the frontier iterations of one pass to convergence: replays of a
// captured K-iteration graph (or direct launches), one host readback of the
// control block per replay.  `output` (the pass's result kernels) is
// enqueued behind the replay in which the engine's previous pass finished,
// ahead of that replay's readback, so a pass that finishes there as well
// returns with its output already on the stream (true); otherwise the caller
// runs it again once the pass has finished (those kernels only read the
// traversal state, so the early copy is overwritten).  This takes the
// readback's round trip (about 70 us) off the end of a batch.
static bool run_iterations(cgp_engine* en, int W, bool no_graph, const std::function<void()>& output) {
  const KP P = en->kp();
  const bool use_graph = en->opts.use_hip_graph != 0 && !no_graph;
  const int K = use_graph ? std::max(1, en->opts.graph_iterations) : 1;
  const int ls = launch_set(en);
  cgp_engine::GExec& gx = en->gx[wlog(W)][ls];
  if (use_graph && (!gx.exec || gx.K != K || gx.prof != en->opts.profile)) {
    if (gx.exec) (void)hipGraphExecDestroy(gx.exec);
    gx.exec = nullptr;
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(en->s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; ++k) dispatch_iteration(en, W, P);
    HIP_CHECK(hipStreamEndCapture(en->s, &graph));
    hipError_t ie = hipGraphInstantiate(&gx.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(ie);
    gx.K = K;
    gx.prof = en->opts.profile;
  }
  u32 iter = 0;  // the seeding compaction leaves iteration 0
  int replays = 0;
  bool out_done = false;
  for (;;) {
    // Iteration i writes BFS depth i + 1: planes for every depth of the
    // iterations this launch may run.
    if (en->h_cfg->depth_planes) add_depth_planes(en, W, bit_length((u64)iter + (u64)K), true);
    if (use_graph) {
      HIP_CHECK(hipGraphLaunch(gx.exec, en->s));
    } else {
      dispatch_iteration(en, W, P);
    }
    ++replays;
    out_done = replays == en->spec_replay;
    if (out_done) output();
    HIP_CHECK(hipMemcpyAsync(en->h_ctl, en->d_ctl, CTL_HOST_BYTES, hipMemcpyDeviceToHost, en->s));
    HIP_CHECK(hipStreamSynchronize(en->s));
    const u32 after = en->h_ctl->iter;
    iter = after;
    if (en->h_ctl->finished) break;
    // Between two raises of the near-far threshold the iterations are a
    // Bellman-Ford frontier over the pairs below it, bounded by V + 1 (each
    // one settles at least one more hop of some shortest path); more means a
    // broken state machine.
    if ((int64_t)after > ((int64_t)en->h_ctl->bumps + 1) * ((int64_t)en->g->V + 2))
      cgpi::fail(CGP_EDEVICE, "traversal did not converge within V+2 iterations per threshold step");
  }
  en->spec_replay = replays;
  return out_done;
}



// Algorithmic bytes of the advance launches of one pass (each byte at its
// natural width, counted once, no cache-line waste):
//   frontier item (vertex or chunk): 8 (offset pair) + 8 W (frontier words)
//   edge scanned lane-per-edge: 4 (column index) [+4 weight when SSSP bits are active]
//   (edge, active BFS word): 8 (visited word of the target)
//   (edge, active SSSP query): 4 (target value) ; lowered value: 4 (store)
//   next word gaining bits: 16 (read + write of the next-frontier word)
//   dense item: 8 (offset pair) + 8 (frontier word); dense edge: 4 index + 4 weight
//   pull: 8 + 8 W per vertex examined, 4 per edge, 8 per (edge, word)
void add_work_bytes(cgp_run_stats& st, const u64* pf, int W) {
  const double Wd = 8.0 * W;
  st.advance_bytes += (double)pf[PF_PUSH_ITEMS] * (8.0 + Wd) + (double)pf[PF_PUSH_EDGES] * 4.0 +
                      (double)pf[PF_PUSH_WEDGES] * 4.0 + (double)pf[PF_BFS_WORDS] * 8.0 +
                      (double)pf[PF_SSSP_PAIRS] * 4.0 + (double)pf[PF_IMPROVED] * 4.0 +
                      (double)pf[PF_NEW_WORDS] * 16.0 + (double)pf[PF_DENSE_ITEMS] * 16.0 +
                      (double)pf[PF_DENSE_EDGES] * 8.0 + (double)pf[PF_PULL_VERTS] * (8.0 + Wd) +
                      (double)pf[PF_PULL_EDGES] * 4.0 + (double)pf[PF_PULL_WORDS] * 8.0 +
                      (double)pf[PF_DENSE_PAIRS] * 4.0 + (double)pf[PF_DENSE_IMP] * 4.0 +
                      (double)pf[PF_DENSE_NEW] * 16.0;
  st.dense_bytes += (double)pf[PF_DENSE_ITEMS] * 16.0 + (double)pf[PF_DENSE_EDGES] * 8.0 +
                    (double)pf[PF_DENSE_PAIRS] * 4.0 + (double)pf[PF_DENSE_IMP] * 4.0 +
                    (double)pf[PF_DENSE_NEW] * 16.0;
  static_assert(NPROF <= CGP_WORK_COUNTERS, "cgp_run_stats.work holds every work counter");
  for (int k = 0; k < NPROF; ++k) st.work[k] += pf[k];
  st.frontier_vertices += (int64_t)(pf[PF_PUSH_ITEMS] + pf[PF_DENSE_ITEMS]);
  st.edges_scanned += (int64_t)(pf[PF_PUSH_EDGES] + pf[PF_DENSE_EDGES] + pf[PF_PULL_EDGES]);
}

void engine_run_batch(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out,
                      bool out_dev, cgp_run_stats* stats, bool no_alloc) {
  const cgp_graph* g = en->g;
  const int32_t V = g->V;
  if (n <= 0) fail(CGP_EINVAL, "cgp_run_batch: batch must contain at least one query");
  if (!queries || !out) fail(CGP_EINVAL, "cgp_run_batch: null queries or output");
  std::vector<int32_t>& bfs = en->q_bfs;
  std::vector<int32_t>& sssp = en->q_sssp;
  std::vector<int32_t>& wcc = en->q_wcc;
  bfs.clear();
  sssp.clear();
  wcc.clear();
  for (int32_t i = 0; i < n; ++i) {
    const cgp_query& q = queries[i];
    if (q.kind != CGP_BFS && q.kind != CGP_SSSP && q.kind != CGP_WCC)
      fail(CGP_EINVAL, "cgp_run_batch: unknown query kind " + std::to_string(q.kind));
    if (q.kind != CGP_WCC && (q.source < 0 || q.source >= V))
      fail(CGP_EINVAL, "cgp_run_batch: query source out of range: " + std::to_string(q.source));
    (q.kind == CGP_BFS ? bfs : q.kind == CGP_SSSP ? sssp : wcc).push_back(i);
  }
  cgp_run_stats st{};
  std::vector<IterRec>& report = en->report;
  report.clear();
  // Iteration records copied into en->h_rings and not yet decoded: report
  // slots [first, first + nrec) of one pass, from area offset `at`.
  struct RingCopy {
    size_t first, nrec, at;
    u32 iters, pass;
  };
  std::vector<RingCopy> rings;
  size_t rings_used = 0;
  // Decodes the copies (their transfers complete: the stream has been
  // synchronised) and empties the area.
  const auto decode_rings = [&]() {
    for (const RingCopy& c : rings)
      for (size_t k = 0; k < c.nrec; ++k) {
        const u32 it = c.iters - (u32)c.nrec + (u32)k;
        report[c.first + k] = en->h_rings[c.at + it % ITER_RING];
        report[c.first + k].pull |= c.pass << 16;  // pass index in the high half
      }
    rings.clear();
    rings_used = 0;
  };
  en->rep_seed.clear();
  en->rep_wall_ms = 0;
  if (!no_alloc) {
    en->rep_queries.assign(queries, queries + n);
    en->rep_slot.assign((size_t)n, -1);
  } else {
    en->rep_queries.clear();
    en->rep_slot.clear();
  }
  if (V == 0) {
    if (stats) *stats = st;
    return;
  }
  en->ev.clear();
  en->ev_used = 0;
  const double t0 = now_ms();

  // Output target: the caller's device buffer, or a cached device buffer.
  u32* dout = out;
  if (!out_dev) {
    en->outbuf.ensure((size_t)n * V * 4);
    dout = (u32*)en->outbuf.p;
  }

  // Passes of at most max_words*64 slots; SSSP slots first, then BFS, so the
  // SSSP slots fill the fewest words (value rows of BFS-only words are never
  // read by the advance).
  const int slots_max = std::max(1, std::min(en->opts.max_words, MAXW)) * 64;
  std::vector<int32_t>& order = en->q_order;
  order.assign(sssp.begin(), sssp.end());
  order.insert(order.end(), bfs.begin(), bfs.end());
  for (size_t p0 = 0; p0 < order.size(); p0 += slots_max) {
    const int ns = (int)std::min<size_t>(slots_max, order.size() - p0);
    const int Wused = (ns + 63) / 64;
    const int W = round_words(Wused);
    int R = 1;
    while (R < std::min(64, ns)) R <<= 1;
    if (Wused > 1) R = 64;
    int rshift = 0;
    while ((1 << rshift) < R) ++rshift;
    en->ensure_words(W);
    en->ensure_vals((size_t)W * V * R);
    const KP P = en->kp();
    // Slot configuration.
    Cfg& cfg = *en->h_cfg;
    std::memset(&cfg, 0, sizeof(Cfg));
    cfg.rshift = rshift;
    cfg.W = W;
    cfg.dense_t = en->opts.dense_threshold;
    cfg.delta = (u32)en->opts.sssp_delta;
    en->stage((size_t)W * 64 * 2);
    int32_t* h_src = en->h_stage;
    int32_t* h_qmap = en->h_stage + W * 64;
    int nb = 0;
    for (int k = 0; k < W * 64; ++k) {
      h_src[k] = -1;
      h_qmap[k] = -1;
    }
    for (int k = 0; k < ns; ++k) {
      const int32_t qi = order[p0 + k];
      const int w = k >> 6, q = k & 63;
      h_src[k] = g->internal(queries[qi].source);
      h_qmap[k] = qi;
      if (!no_alloc) en->rep_slot[(size_t)qi] = (int32_t)((p0 / slots_max) << 16 | w << 8 | q);
      if (queries[qi].kind == CGP_BFS) {
        cfg.bfs_bits[w] |= 1ull << q;
        ++nb;
      } else {
        cfg.val_bits[w] |= 1ull << q;
      }
    }
    cfg.narrow_words = rshift >= 4 && cfg.delta == 0 && en->opts.dense_narrow ? narrow_mask(cfg, W) : 0u;
    en->narrow_binned = cfg.narrow_words;  // fixed for the pass, its seeding filter included
    // Bottom-up steps pay off only when few BFS queries share the words: with
    // many, the missing bits of a vertex are rarely all found early.  Hybrid
    // mode pulls only for batches of at most pull_max_queries BFS queries
    // (default 0: on MI355X the push kernels measured faster even for one
    // BFS on RMAT-22); CGP_PULL forces it.
    const bool can_pull = g->symmetric && nb > 0 && en->opts.traversal_mode != CGP_PUSH &&
                          (en->opts.traversal_mode == CGP_PULL || nb <= en->opts.pull_max_queries);
    cfg.pull_policy = !can_pull ? 0 : en->opts.traversal_mode == CGP_PULL ? 1 : 2;
    // SSSP pull steps gather over each vertex's own edges, so they need the
    // reverse of every edge with the same weight; they are taken with
    // postponing off (a gather relaxes every pair, near or far), and not in
    // allocation-free runs (their item list is built on first use).
    const int nsssp = ns - nb;
    const bool can_spull = g->sym_weights && nsssp > 0 && cfg.delta == 0 && !no_alloc &&
                           en->opts.traversal_mode != CGP_PUSH &&
                           (en->opts.traversal_mode == CGP_PULL || nsssp <= en->opts.pull_max_queries);
    cfg.sssp_pull_policy = !can_spull ? 0 : en->opts.traversal_mode == CGP_PULL ? 1 : 2;
    cfg.spull_pairs_thr = std::max<u64>(1, (u64)(en->opts.pull_frontier_ratio * (double)nsssp * (double)V));
    cfg.spull_edges_thr = std::max<u64>(1, (u64)(en->opts.pull_edge_ratio * (double)nsssp * (double)g->E));
    cfg.spush_pairs_thr = cfg.spull_pairs_thr / 4 + 1;
    if (cfg.sssp_pull_policy) en->ensure_pull_items();
    cfg.dense_fedges_thr = std::max<u64>(1, (u64)V / (u64)en->opts.frontier_dense_div);
    cfg.pull_pairs_thr = (u64)(en->opts.pull_frontier_ratio * (double)nb * (double)V);
    cfg.pull_edges_thr = (u64)(en->opts.pull_edge_ratio * (double)nb * (double)g->E);
    cfg.push_pairs_thr = cfg.pull_pairs_thr / 4 + 1;
    if (cfg.pull_pairs_thr == 0) cfg.pull_pairs_thr = 1;
    if (cfg.pull_edges_thr == 0) cfg.pull_edges_thr = 1;
    // BFS depth bit planes: in batch runs every slot starts at iteration 0
    // (allocation-free runs keep depths in the value rows).
    en->plane_count = 0;
    en->plane_words = 0;
    if (!no_alloc && nb > 0) {
      cfg.depth_planes = 1;
      for (int w = 0; w < W; ++w)
        if (cfg.bfs_bits[w]) en->plane_words |= 1u << w;
      add_depth_planes(en, W, 4, false);  // depths below 16
    }
    // Control block.
    Ctl& c = *en->h_ctl;
    std::memset(&c, 0, sizeof(Ctl));
    c.iter = 0xFFFFFFFFu;
    c.thr = cfg.delta ? cfg.delta : INF;
    c.far_min = INF;
    c.adv_t0 = ~0ull;
    c.den_t0 = ~0ull;
    c.flt_t0 = ~0ull;
    if (en->stamp_epoch > 0xF0000000u) {
      HIP_CHECK(hipMemsetAsync(en->mark.p, 0, (size_t)V * 4, en->s));
      en->stamp_epoch = 1;
    }
    c.stamp_base = en->stamp_epoch;
    c.max_iter = en->opts.max_iterations > 0 ? (u32)en->opts.max_iterations : 0u;
    en->slot_src.ensure((size_t)W * 64 * 4);
    en->qmap.ensure((size_t)W * 64 * 4);
    en->tick(CAT_INIT, true);
    HIP_CHECK(hipMemcpyAsync(en->d_cfg, &cfg, sizeof(Cfg), hipMemcpyHostToDevice, en->s));
    HIP_CHECK(hipMemcpyAsync(en->d_ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, en->s));
    HIP_CHECK(hipMemcpyAsync(en->slot_src.p, h_src, (size_t)W * 64 * 4, hipMemcpyHostToDevice, en->s));
    HIP_CHECK(hipMemcpyAsync(en->qmap.p, h_qmap, (size_t)W * 64 * 4, hipMemcpyHostToDevice, en->s));
    // Value rows start at CGP_INF only where SSSP slots read them: words
    // without SSSP slots are left as they are, and of a 64-slot word only the
    // 16-, 32- or 64-lane prefix holding its SSSP slots (they are packed from
    // lane 0) is filled.  BFS slots are masked by their visited bits at output.
    for (int w = 0; w < W; ++w) {
      const u64 vb = cfg.val_bits[w];
      if (!vb) continue;
      u32* wv = (u32*)en->vals.p + (size_t)w * V * R;
      const int hi = 64 - __builtin_clzll(vb);
      if (R < 64 || hi > 32) fill_u32(en->s, wv, (size_t)V * R, INF, en->cus);
      else fill_row_prefix(en->s, wv, (size_t)V, hi > 16 ? 3 : 2, INF, en->cus);
    }
    // Visited + next records (2W words per vertex) start clear; the layout
    // depends on W, so the whole region is cleared for every pass.
    fill_u32(en->s, en->visited.p, (size_t)V * W * 4, 0u, en->cus);
    en->next_dirty = false;
    dispatch_seed(en, W, P, (const int32_t*)en->slot_src.p, W * 64);
    en->tick(CAT_INIT, false);
    // The pass's output: query-major rows from the value rows, visited records
    // and depth planes (the planes the replays have added by then).
    const auto emit_output = [&]() {
      en->tick(CAT_OUT, true);
      const int vec = rshift == 6 && (V & 3) == 0 && ((uintptr_t)dout & 15) == 0;
      OutMasks om{};
      for (int w = 0; w < Wused; ++w) {
        om.bfs[w] = cfg.bfs_bits[w];
        om.val[w] = cfg.val_bits[w];
      }
      om.n_planes = cfg.depth_planes ? en->plane_count : 0;
      om.plane4 = (const u64*)en->plane4.p;
      for (int b = 4; b < om.n_planes; ++b) om.plane[b] = (const u64*)en->plane[b].p;
      hipLaunchKernelGGL(k_output, dim3((V + OUT_VT - 1) / OUT_VT, Wused), dim3(BLOCK),
                         (size_t)om.n_planes * OUT_VT * sizeof(u64), en->s, (const u32*)en->vals.p,
                         (const u64*)en->visited.p, (const int32_t*)g->inv, V, W, rshift,
                         (const int32_t*)en->qmap.p, om, dout, vec);
      HIP_CHECK(hipGetLastError());
      en->tick(CAT_OUT, false);
    };
    const bool out_done = run_iterations(en, W, no_alloc, emit_output);
    const Ctl& hc = *en->h_ctl;
    st.iterations += (int32_t)hc.iter;
    st.push_iterations += (int32_t)hc.push_iters;
    st.pull_iterations += (int32_t)hc.pull_iters;
    st.passes += 1;
    if (en->counted()) add_work_bytes(st, hc.prof, W);
    if (!no_alloc) {
      en->rep_seed.push_back({hc.seed_pairs, hc.seed_verts, hc.seed_edges, hc.seed_qedges, hc.t_start});
      // Per-iteration records of this pass (the ring keeps the last ITER_RING):
      // copied on the engine's stream (the legacy stream would wait on, and fail
      // beside, other threads' engines capturing graphs) into the pinned area,
      // and decoded after the batch's final synchronisation.
      const size_t nrec = std::min<size_t>(hc.iter, (size_t)ITER_RING);
      const size_t first = report.size();
      report.resize(first + nrec);
      if (nrec) {
        if (rings_used + nrec > en->h_rings_cap) {
          HIP_CHECK(hipStreamSynchronize(en->s));
          decode_rings();
          en->stage_rings(std::max(nrec, (size_t)256));
        }
        HIP_CHECK(hipMemcpyAsync(en->h_rings + rings_used, en->iters.p, nrec * sizeof(IterRec),
                                 hipMemcpyDeviceToHost, en->s));
        rings.push_back({first, nrec, rings_used, hc.iter, (u32)(st.passes - 1)});
        rings_used += nrec;
      }
    }
    if (en->timed()) {
      st.advance_ms += (double)hc.adv_ticks / en->clock_khz;
      st.filter_ms += (double)hc.flt_ticks / en->clock_khz;
      st.advance_launches += (int64_t)hc.adv_launches;
      st.dense_ms += (double)hc.den_ticks / en->clock_khz;
      st.dense_launches += (int64_t)hc.den_launches;
    }
    en->stamp_epoch += hc.iter + 2;
    if (hc.max_iter != 0 && hc.iter >= hc.max_iter) en->next_dirty = true;  // truncated run
    if (!out_done) emit_output();
  }
  if (!wcc.empty()) {
    en->wcc_rows.ensure(wcc.size() * 4);
    en->stage(wcc.size());
    std::memcpy(en->h_stage, wcc.data(), wcc.size() * 4);
    HIP_CHECK(hipMemcpyAsync(en->wcc_rows.p, en->h_stage, wcc.size() * 4, hipMemcpyHostToDevice, en->s));
    en->tick(CAT_WCC, true);
    compute_wcc(en);
    en->tick(CAT_WCC, false);
    en->tick(CAT_OUT, true);
    write_wcc_rows(en, (const int32_t*)en->wcc_rows.p, (int)wcc.size(), dout);
    en->tick(CAT_OUT, false);
  }
  HIP_CHECK(hipStreamSynchronize(en->s));
  st.wall_ms = now_ms() - t0;
  en->rep_wall_ms = st.wall_ms;
  decode_rings();
  if (!out_dev) {
    HIP_CHECK(hipMemcpyAsync(out, dout, (size_t)n * V * 4, hipMemcpyDeviceToHost, en->s));
    HIP_CHECK(hipStreamSynchronize(en->s));
  }
  if (en->timed()) {
    for (auto& e : en->ev) {
      float ms = 0;
      if (e.second.second) HIP_CHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
      switch (e.first) {
        case CAT_INIT: st.init_ms += ms; break;
        case CAT_WCC: st.wcc_ms += ms; break;
        default: st.output_ms += ms; break;
      }
    }
  }
  if (stats) *stats = st;
}

// ------------------------------------------------------------ no-allocation runs
// Every per-run device buffer of cgp_run_batch_static lives in this array:
// CGP_STATIC_SCRATCH_BYTES per GPU, held from the library's load until the
// process ends.  Host code never uses the symbol itself as an address (in
// host code it is only a stand-in); static_scratch() looks up each device's
// own copy once with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[CGP_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_static_mu;  // serialises static runs of this process (one array per GPU)
void* g_static_ptr[kMaxDevices] = {};

// Address of the calling device's copy of `scratch` (looked up on first use
// per device; the caller holds g_static_mu).
void* static_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) fail(CGP_EDEVICE, "cgp_run_batch_static: device index out of range");
  if (!g_static_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      fail(CGP_EDEVICE, std::string("cgp_run_batch_static: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_static_ptr[dev] = p;
  }
  return g_static_ptr[dev];
}

// Byte layout of the per-run buffers inside the array, for Wm mask words, Rm
// lanes per value row and nw WCC rows (each region 256-byte aligned).
struct StaticLayout {
  size_t visited, front, dense[2], dense_n[2], mark, touched, small[2], medium[2], large[2], vals, parent,
      wcc_rows, wcc_probe, slot_src, qmap, total;
};

StaticLayout static_layout(const cgp_graph* g, int Wm, int Rm, int nw) {
  const size_t V = (size_t)g->V, E = (size_t)g->E;
  StaticLayout L{};
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t r = at;
    at += (bytes + 255) & ~(size_t)255;
    return r;
  };
  L.visited = take(V * Wm * 16);
  L.front = take(V * Wm * 8);
  for (int k = 0; k < 2; ++k) {
    L.dense[k] = take((size_t)Wm * (V + E / DCHUNK + 1) * 8);
    L.dense_n[k] = take((size_t)Wm * (V + E / DCHUNK + 1) * 8);
  }
  L.mark = take(V * 4 + 4);
  L.touched = take(V * 4 + 4);
  for (int k = 0; k < 2; ++k) {
    L.small[k] = take(V * 4 + 4);
    L.medium[k] = take(V * 4 + 4);
    L.large[k] = take((V + E / LCHUNK + 1) * 8);
  }
  L.vals = take((size_t)Wm * V * Rm * 4);
  L.parent = take(V * 4 + 4);
  L.wcc_rows = take((size_t)std::max(nw, 1) * 4);
  L.wcc_probe = take(1024 * 4);
  L.slot_src = take((size_t)Wm * 64 * 4);
  L.qmap = take((size_t)Wm * 64 * 4);
  L.total = at;
  return L;
}

}  // namespace

void engine_run_batch_static(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out,
                             cgp_run_stats* stats) {
  const cgp_graph* g = en->g;
  if (n <= 0 || n > CGP_STATIC_MAX_QUERIES)
    fail(CGP_ECAPACITY, "cgp_run_batch_static: batch of " + std::to_string(n) + " queries (1 to " +
                            std::to_string(CGP_STATIC_MAX_QUERIES) + " allowed)");
  if (!queries || !out) fail(CGP_EINVAL, "cgp_run_batch_static: null queries or output");
  // The widest pass: words and value-row lanes as engine_run_batch picks them.
  int nv = 0, nw = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (queries[i].kind == CGP_WCC) ++nw;
    else ++nv;
  }
  const int slots_max = std::max(1, std::min(en->opts.max_words, MAXW)) * 64;
  const int first = std::min(nv, slots_max);
  int Wm = 1, Rm = 1;
  if (first > 0) {
    Wm = round_words((first + 63) / 64);
    while (Rm < std::min(64, first)) Rm <<= 1;
    if (first > 64) Rm = 64;
  }
  const StaticLayout L = static_layout(g, Wm, Rm, nw);
  if (L.total > (size_t)CGP_STATIC_SCRATCH_BYTES)
    fail(CGP_ECAPACITY, "cgp_run_batch_static: this batch needs " + std::to_string(L.total) +
                            " bytes of device scratch; the static array holds " +
                            std::to_string((size_t)CGP_STATIC_SCRATCH_BYTES) +
                            " (use cgp_run_batch for larger graphs or batches)");
  std::lock_guard<std::mutex> glk(g_static_mu);
  int cur = 0;
  HIP_CHECK(hipGetDevice(&cur));
  if (cur != en->dev) HIP_CHECK(hipSetDevice(en->dev));
  struct RestoreDev {
    int d;
    bool on;
    ~RestoreDev() {
      if (on) (void)hipSetDevice(d);
    }
  } restore_dev{cur, cur != en->dev};
  unsigned char* base = (unsigned char*)static_scratch(en->dev);

  // Bind every per-run buffer to its region of the array for this call,
  // and put the engine's own buffers back afterwards (also on failure).
  DevBuf* bufs[] = {&en->visited, &en->front, &en->dense[0], &en->dense[1], &en->dense_n[0], &en->dense_n[1],
                    &en->mark, &en->touched, &en->small[0], &en->small[1], &en->medium[0], &en->medium[1],
                    &en->large[0], &en->large[1], &en->vals, &en->parent, &en->wcc_rows, &en->wcc_probe,
                    &en->slot_src, &en->qmap};
  const size_t offs[] = {L.visited, L.front, L.dense[0], L.dense[1], L.dense_n[0], L.dense_n[1],
                         L.mark, L.touched, L.small[0], L.small[1], L.medium[0], L.medium[1],
                         L.large[0], L.large[1], L.vals, L.parent, L.wcc_rows, L.wcc_probe,
                         L.slot_src, L.qmap};
  constexpr int NB = (int)(sizeof(offs) / sizeof(offs[0]));
  struct Saved {
    cgp_engine* en;
    DevBuf own[NB];
    DevBuf* const* bufs;
    int wcap, profile;
    u32 stamp;
    ~Saved() {
      for (int k = 0; k < NB; ++k) *bufs[k] = own[k];
      en->wcap = wcap;
      en->opts.profile = profile;
      en->stamp_epoch = stamp;
    }
  } saved{en, {}, bufs, en->wcap, en->opts.profile, en->stamp_epoch};
  for (int k = 0; k < NB; ++k) {
    saved.own[k] = *bufs[k];
    const size_t end = k + 1 < NB ? offs[k + 1] : L.total;
    bufs[k]->p = base + offs[k];
    bufs[k]->bytes = end - offs[k];
    bufs[k]->borrowed = true;
  }
  en->wcap = Wm;
  en->opts.profile &= ~1;  // no event timing (events would be created on demand)
  // The array keeps the previous call's contents (any engine on this GPU):
  // the vertex stamps must start below this call's first stamp.  The other
  // regions are cleared or written by the run before they are read.
  fill_u32(en->s, en->mark.p, (size_t)g->V + 1, 0u, en->cus);
  en->stamp_epoch = 1;
  engine_run_batch(en, queries, n, out, true, stats, true);
}

}  // namespace cgpi

// ================================================================= C-ABI (device side)
extern "C" {

cgp_status cgp_device_count(int32_t* count) {
  CGP_TRY({
    if (!count) cgpi::fail(CGP_EINVAL, "cgp_device_count: null output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

cgp_status cgp_device_info(char* name, int32_t name_cap, int32_t* compute_units, uint64_t* hbm_bytes) {
  CGP_TRY({
    require_device();
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, 0));
    if (name && name_cap > 0) {
      std::snprintf(name, (size_t)name_cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
  });
}

cgp_status cgp_device_alloc(uint64_t bytes, void** out) {
  CGP_TRY({
    if (!out) cgpi::fail(CGP_EINVAL, "cgp_device_alloc: null output");
    require_device();
    *out = nullptr;
    HIP_CHECK(hipMalloc(out, bytes ? bytes : 1));
  });
}

cgp_status cgp_device_free(void* ptr) {
  CGP_TRY({
    if (ptr) HIP_CHECK(hipFree(ptr));
  });
}

cgp_status cgp_copy_to_host(void* dst, const void* src, uint64_t bytes) {
  CGP_TRY({ HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

cgp_status cgp_copy_to_device(void* dst, const void* src, uint64_t bytes) {
  CGP_TRY({ HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

cgp_status cgp_copy_device(void* dst, const void* src, uint64_t bytes) {
  CGP_TRY({
    if (bytes == 0) return CGP_OK;
    if (!dst || !src) cgpi::fail(CGP_EINVAL, "cgp_copy_device: null pointer");
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15u) != 0) {
      HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    } else {
      const size_t n16 = bytes / 16;
      const int cus = cgpi::device_cus();
      const size_t blocks = std::min<size_t>((n16 + BLOCK - 1) / BLOCK, (size_t)cus * 16);
      hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(BLOCK), 0, nullptr, (uint4*)dst, (const uint4*)src,
                         n16);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipDeviceSynchronize());
    }
  });
}

// Largest degree at least 32x the mean (power-law graphs; not meshes or grids).
static bool is_skewed(int32_t V, int64_t E, int32_t maxd) {
  return V > 0 && (double)maxd >= 32.0 * std::max(1.0, (double)E / (double)V);
}

cgp_status cgp_graph_create_csr(int32_t V, int64_t E, const int32_t* off, const int32_t* idx,
                                const uint32_t* wt, cgp_graph** out) {
  return cgp_graph_create_csr_ex(V, E, off, idx, wt, CGP_RENUMBER_AUTO, out);
}

cgp_status cgp_graph_create_csr_ex(int32_t V, int64_t E, const int32_t* off, const int32_t* idx,
                                   const uint32_t* wt, int32_t renumber, cgp_graph** out) {
  CGP_TRY({
    if (!out) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: null output");
    if (renumber < CGP_RENUMBER_AUTO || renumber > CGP_RENUMBER_OFF)
      cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: renumber must be CGP_RENUMBER_AUTO, _ON or _OFF");
    *out = nullptr;
    if (V < 0 || E < 0 || E >= INT32_MAX) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: bad sizes");
    if (!off || (E > 0 && !idx)) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: null CSR array");
    if (off[0] != 0 || off[V] != E) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: offsets do not span [0, E]");
    for (int32_t v = 0; v < V; ++v)
      if (off[v + 1] < off[v]) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: offsets not monotone");
    for (int64_t e = 0; e < E; ++e)
      if (idx[e] < 0 || idx[e] >= V) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr: column index out of range");
    require_device();
    cgpi::HostCSR h;
    h.V = V;
    h.E = E;
    h.off.assign(off, off + V + 1);
    h.idx.assign(idx, idx + E);
    if (wt) h.wt.assign(wt, wt + E);
    std::unique_ptr<cgp_graph> g(new cgp_graph());
    g->V = V;
    g->E = E;
    g->weighted = wt != nullptr;
    g->symmetric = cgpi::is_symmetric(h);
    g->sym_weights = g->symmetric && cgpi::weights_symmetric(h);
    if (wt && E > 0) {
      double sum = 0;
      for (int64_t e = 0; e < E; ++e) sum += wt[e];
      g->mean_weight = sum / (double)E;
    }
    // Large skewed graphs are stored with their vertices renumbered by
    // descending degree: the hubs that most edges reach get neighbouring value
    // rows and records, which share DRAM pages, TLB entries and cache sets
    // (RMAT-22, 160 BFS + 80 SSSP + 16 WCC: 54.8 -> 50.2 ms per batch).  Below
    // 2^21 vertices a word's value rows (256 bytes a vertex) fit the on-chip
    // cache, and the renumbering measured slower (RMAT-20 64 SSSP: 9.0 -> 9.5
    // ms; RMAT-16 32 BFS: 0.79 -> 1.02 ms); graphs whose largest degree is
    // within 32x the mean (meshes, grids) keep their numbering too, their
    // neighbours being close already.  Results, sources and WCC labels stay
    // in the caller's numbering.
    int32_t maxd = 0;
    for (int32_t v = 0; v < V; ++v) maxd = std::max(maxd, off[v + 1] - off[v]);
    g->skewed = is_skewed(V, E, maxd);
    std::vector<int32_t> perm;
    const bool auto_on = V >= (1 << 21) && g->skewed;
    if (V > 0 && (renumber == CGP_RENUMBER_ON || (renumber == CGP_RENUMBER_AUTO && auto_on)))
      h = cgpi::relabel_by_degree(h, perm, g->h_inv);
    HIP_CHECK(hipMalloc((void**)&g->off, ((size_t)V + 1) * 4));
    HIP_CHECK(hipMalloc((void**)&g->idx, (size_t)E * 4 + 4));
    HIP_CHECK(hipMalloc((void**)&g->rows, (size_t)E * 4 + 4));
    if (!perm.empty()) {
      HIP_CHECK(hipMalloc((void**)&g->perm, (size_t)V * 4 + 4));
      HIP_CHECK(hipMalloc((void**)&g->inv, (size_t)V * 4 + 4));
    }
    if (wt) HIP_CHECK(hipMalloc((void**)&g->wt, (size_t)E * 4 + 4));
    HIP_CHECK(hipMemcpy(g->off, h.off.data(), ((size_t)V + 1) * 4, hipMemcpyHostToDevice));
    if (E) HIP_CHECK(hipMemcpy(g->idx, h.idx.data(), (size_t)E * 4, hipMemcpyHostToDevice));
    if (wt && E) HIP_CHECK(hipMemcpy(g->wt, h.wt.data(), (size_t)E * 4, hipMemcpyHostToDevice));
    if (!perm.empty()) {
      HIP_CHECK(hipMemcpy(g->perm, perm.data(), (size_t)V * 4, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(g->inv, g->h_inv.data(), (size_t)V * 4, hipMemcpyHostToDevice));
    }
    // Degree order puts the vertices without edges last; on a symmetric graph
    // nothing reaches them, so full compaction scans stop before them.
    g->scan_v = V;
    if (!perm.empty() && g->symmetric)
      while (g->scan_v > 0 && h.off[g->scan_v] == h.off[g->scan_v - 1]) --g->scan_v;
    {
      std::vector<int32_t> lo;
      std::vector<int32_t> hi;
      for (int32_t v = 0; v < V; ++v) {
        const int32_t d = h.off[v + 1] - h.off[v];
        if (d > PULL_WAVE_DEG) hi.push_back(v);
        else if (d > 0) lo.push_back(v);
      }
      g->n_pull_lo = (int32_t)lo.size();
      g->n_pull_hi = (int32_t)hi.size();
      HIP_CHECK(hipMalloc((void**)&g->pull_lo, lo.size() * 4 + 4));
      HIP_CHECK(hipMalloc((void**)&g->pull_hi, hi.size() * 4 + 4));
      if (!lo.empty()) HIP_CHECK(hipMemcpy(g->pull_lo, lo.data(), lo.size() * 4, hipMemcpyHostToDevice));
      if (!hi.empty()) HIP_CHECK(hipMemcpy(g->pull_hi, hi.data(), hi.size() * 4, hipMemcpyHostToDevice));
    }
    if (V > 0) {
      int cus = device_cus();
      hipLaunchKernelGGL(k_fill_rows, dim3(cus * 16), dim3(BLOCK), 0, 0, (const int32_t*)g->off, V, g->rows);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipDeviceSynchronize());
    }
    *out = g.release();
  });
}

cgp_status cgp_graph_from_host(const cgp_host_csr* h, cgp_graph** out) {
  return cgp_graph_from_host_ex(h, CGP_RENUMBER_AUTO, out);
}

cgp_status cgp_graph_from_host_ex(const cgp_host_csr* h, int32_t renumber, cgp_graph** out) {
  if (!h) {
    cgpi::set_last_error("cgp_graph_from_host: null host graph");
    return CGP_EINVAL;
  }
  return cgp_graph_create_csr_ex(h->g.V, h->g.E, h->g.off.data(), h->g.idx.data(),
                                 h->g.wt.empty() ? nullptr : h->g.wt.data(), renumber, out);
}

cgp_status cgp_graph_create_csr_device(int32_t V, int64_t E, const int32_t* off, const int32_t* idx,
                                       const uint32_t* wt, cgp_graph** out) {
  CGP_TRY({
    if (!out) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: null output");
    *out = nullptr;
    if (V < 0 || E < 0 || E >= INT32_MAX) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: bad sizes");
    if (!off || (E > 0 && !idx)) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: null CSR array");
    require_device();
    // Device graph owning only its derived arrays; the CSR stays the caller's.
    struct Owner {
      cgp_graph* g;
      u32* d_tmp = nullptr;
      ~Owner() {
        if (d_tmp) (void)hipFree(d_tmp);
        if (g) cgp_graph_destroy(g);
      }
    } own{new cgp_graph()};
    cgp_graph* g = own.g;
    g->borrowed = true;
    g->V = V;
    g->scan_v = V;
    g->E = E;
    g->off = const_cast<int32_t*>(off);
    g->idx = const_cast<int32_t*>(idx);
    g->wt = const_cast<u32*>(wt);
    g->weighted = wt != nullptr;
    int32_t ends[2];
    ends[0] = ends[1] = 0;
    HIP_CHECK(hipMemcpy(&ends[0], off, 4, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(&ends[1], off + V, 4, hipMemcpyDeviceToHost));
    if (ends[0] != 0 || ends[1] != E) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: offsets do not span [0, E]");
    // Scratch: flags, degree-list counters, weight sum.
    HIP_CHECK(hipMalloc((void**)&own.d_tmp, 32));
    HIP_CHECK(hipMemset(own.d_tmp, 0, 32));
    u32* flags = own.d_tmp;
    u32* cnt = own.d_tmp + 2;
    u64* wsum = (u64*)(own.d_tmp + 4);
    const int cus = device_cus();
    const dim3 gb(cus * 16);
    auto read_flags = [&]() {
      u32 f = 0;
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpy(&f, flags, 4, hipMemcpyDeviceToHost));
      return f;
    };
    u32* maxdeg = own.d_tmp + 6;
    if (V > 0) hipLaunchKernelGGL(k_csr_check_offsets, gb, dim3(BLOCK), 0, 0, off, V, flags, maxdeg);
    if (read_flags() & CSR_BAD_OFFSETS) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: offsets not monotone");
    u32 maxd = 0;
    HIP_CHECK(hipMemcpy(&maxd, maxdeg, 4, hipMemcpyDeviceToHost));
    g->skewed = is_skewed(V, E, (int32_t)maxd);
    HIP_CHECK(hipMalloc((void**)&g->rows, (size_t)E * 4 + 4));
    if (V > 0) hipLaunchKernelGGL(k_fill_rows, gb, dim3(BLOCK), 0, 0, off, V, g->rows);
    if (E > 0) hipLaunchKernelGGL(k_csr_check_edges, gb, dim3(BLOCK), 0, 0, (const int32_t*)g->rows, idx, E, V, flags);
    u32 f = read_flags();
    if (f & CSR_BAD_INDEX) cgpi::fail(CGP_EINVAL, "cgp_graph_create_csr_device: column index out of range");
    // Symmetry is only recognised on sorted rows (otherwise the graph is
    // treated as directed: every edge hooks in WCC, and BFS never pulls).
    if (!(f & CSR_UNSORTED) && E > 0)
      hipLaunchKernelGGL(k_csr_check_symmetric, gb, dim3(BLOCK), 0, 0, off, (const int32_t*)g->rows, idx, wt, E,
                         flags);
    f = read_flags();
    g->symmetric = !(f & (CSR_UNSORTED | CSR_ASYMMETRIC));
    g->sym_weights = g->symmetric && !(f & CSR_ASYM_WEIGHTS);
    if (wt && E > 0) {
      hipLaunchKernelGGL(k_sum_u32, gb, dim3(BLOCK), 0, 0, wt, E, wsum);
      u64 sum = 0;
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpy(&sum, wsum, 8, hipMemcpyDeviceToHost));
      g->mean_weight = (double)sum / (double)E;
    }
    HIP_CHECK(hipMalloc((void**)&g->pull_lo, (size_t)V * 4 + 4));
    HIP_CHECK(hipMalloc((void**)&g->pull_hi, (size_t)V * 4 + 4));
    if (V > 0) hipLaunchKernelGGL(k_degree_split, gb, dim3(BLOCK), 0, 0, off, V, g->pull_lo, g->pull_hi, cnt);
    u32 c[2];
    c[0] = c[1] = 0;
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(c, cnt, 8, hipMemcpyDeviceToHost));
    g->n_pull_lo = (int32_t)c[0];
    g->n_pull_hi = (int32_t)c[1];
    *out = g;
    own.g = nullptr;
  });
}

void cgp_graph_destroy(cgp_graph* g) {
  if (!g) return;
  if (!g->borrowed) {
    (void)hipFree(g->off);
    (void)hipFree(g->idx);
    if (g->wt) (void)hipFree(g->wt);
  }
  (void)hipFree(g->rows);
  (void)hipFree(g->pull_lo);
  (void)hipFree(g->pull_hi);
  if (g->perm) (void)hipFree(g->perm);
  if (g->inv) (void)hipFree(g->inv);
  delete g;
}

int32_t cgp_graph_num_vertices(const cgp_graph* g) { return g ? g->V : -1; }
int64_t cgp_graph_num_edges(const cgp_graph* g) { return g ? g->E : -1; }
int32_t cgp_graph_is_symmetric(const cgp_graph* g) { return g && g->symmetric ? 1 : 0; }

void cgp_engine_opts_default(cgp_engine_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->traversal_mode = CGP_HYBRID;
  o->pull_frontier_ratio = 0.15;
  o->pull_edge_ratio = 0.20;
  o->max_iterations = 0;
  o->max_words = MAXW;
  o->block_size = 256;
  o->bin_small = 64;
  o->bin_large = 2048;
  o->use_hip_graph = 1;
  o->dense_threshold = 0;  // automatic (resolved at engine creation)
  o->frontier_dense_div = 8;
  o->dense_unroll = 16;
  o->dense_narrow = 1;
  o->concurrent_advance = 0;
  o->sssp_delta = -1;
  o->dense_wide_addr = 0;
  o->graph_iterations = 8;
  o->profile = 0;
  o->hbm_budget_bytes = 64ull << 30;
}

}  // extern "C"

namespace {

// Options in force for graph g from the caller's `opts`: a zero (or
// negative, where noted) tunable field takes its value from `base` (the
// defaults at creation, the engine's current, possibly tuned, options
// later); automatic fields are resolved for the graph.
cgp_engine_opts resolve_opts(const cgp_graph* g, const cgp_engine_opts* opts, const cgp_engine_opts& base,
                             const char* op) {
  const std::string who(op);
  cgp_engine_opts o = opts ? *opts : base;
  if (opts) {
    if (o.pull_frontier_ratio <= 0) o.pull_frontier_ratio = base.pull_frontier_ratio;
    if (o.pull_edge_ratio <= 0) o.pull_edge_ratio = base.pull_edge_ratio;
    if (o.max_words <= 0 || o.max_words > MAXW) o.max_words = base.max_words;
    if (o.block_size <= 0) o.block_size = base.block_size;
    if (o.bin_small <= 0) o.bin_small = base.bin_small;
    if (o.bin_large <= 0) o.bin_large = base.bin_large;
    if (o.graph_iterations <= 0) o.graph_iterations = base.graph_iterations;
    if (o.hbm_budget_bytes == 0) o.hbm_budget_bytes = base.hbm_budget_bytes;
    if (o.frontier_dense_div <= 0) o.frontier_dense_div = base.frontier_dense_div;
    if (o.dense_unroll <= 0) o.dense_unroll = base.dense_unroll;
    if (o.traversal_mode < 0 || o.traversal_mode > 2) cgpi::fail(CGP_EINVAL, who + ": bad traversal_mode");
  }
  if (o.bin_large < o.bin_small) o.bin_large = o.bin_small;
  if (o.block_size != 128 && o.block_size != 256 && o.block_size != 512)
    cgpi::fail(CGP_EINVAL, who + ": block_size must be 128, 256 or 512");
  if (o.dense_threshold < 0) cgpi::fail(CGP_EINVAL, who + ": dense_threshold must be >= 0");
  if (o.advance_blocks < 0) cgpi::fail(CGP_EINVAL, who + ": advance_blocks must be >= 0");
  if (o.pull_max_queries < 0) cgpi::fail(CGP_EINVAL, who + ": pull_max_queries must be >= 0");
  if (o.max_iterations < 0) cgpi::fail(CGP_EINVAL, who + ": max_iterations must be >= 0");
  if (o.dense_unroll != 4 && o.dense_unroll != 8 && o.dense_unroll != 16)
    cgpi::fail(CGP_EINVAL, who + ": dense_unroll must be 4, 8 or 16");
  if (o.sssp_delta < -1) cgpi::fail(CGP_EINVAL, who + ": sssp_delta must be >= -1");
  if (o.dense_wide_addr != 0 && o.dense_wide_addr != 1)
    cgpi::fail(CGP_EINVAL, who + ": dense_wide_addr must be 0 or 1");
  if (o.sssp_delta == -1) {
    // Automatic near-far step.  Low-degree weighted graphs (grids, road
    // networks: long shortest paths, many re-relaxations) get twenty mean
    // weights (on a 4096^2 grid that is faster than no step both for one
    // query and for 512); small-world graphs settle in few Bellman-Ford
    // iterations and skip it.
    const double deg = g->V > 0 ? (double)g->E / (double)g->V : 0.0;
    o.sssp_delta =
        (g->weighted && deg <= 8.0 && g->mean_weight > 1.0) ? (int32_t)std::lround(20.0 * g->mean_weight) : 0;
  }
  // A zero threshold keeps the base's (possibly tuned) value while the
  // postponing regime it was chosen for is unchanged.
  if (o.dense_threshold == 0 && opts && o.sssp_delta == base.sssp_delta) o.dense_threshold = base.dense_threshold;
  if (o.dense_threshold == 0) {
    // Automatic lane-per-query threshold.  The table loop (postponing off)
    // costs less per edge row than the postponing loop, so it pays from
    // fewer active queries: on the RMAT-22 mixed batch 5 measured 53.4
    // against 54.2 ms for 6 (three interleaved runs each; 4: 55.2, 8: 55.2),
    // while on the 4096^2 grid (postponing on) 6 instead of 8 took 2380 to
    // 2535 ms.
    o.dense_threshold = o.sssp_delta == 0 ? 5 : 8;
  }
  return o;
}

}  // namespace

extern "C" {

cgp_status cgp_engine_create(const cgp_graph* g, const cgp_engine_opts* opts, cgp_engine** out) {
  CGP_TRY({
    if (!g || !out) cgpi::fail(CGP_EINVAL, "cgp_engine_create: null graph or output");
    *out = nullptr;
    require_device();
    std::unique_ptr<cgp_engine> en(new cgp_engine());
    en->g = g;
    cgp_engine_opts d;
    cgp_engine_opts_default(&d);
    en->opts = resolve_opts(g, opts, d, "cgp_engine_create");
    en->cus = device_cus();
    {
      int dev = 0;
      int khz = 0;
      HIP_CHECK(hipGetDevice(&dev));
      en->dev = dev;
      if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
        en->clock_khz = (double)khz;
    }
    HIP_CHECK(hipStreamCreateWithFlags(&en->s, hipStreamNonBlocking));
    HIP_CHECK(hipStreamCreateWithFlags(&en->s2, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&en->ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&en->ev_join, hipEventDisableTiming));
    HIP_CHECK(hipMalloc((void**)&en->d_ctl, sizeof(Ctl)));
    HIP_CHECK(hipMalloc((void**)&en->d_cfg, sizeof(Cfg)));
    HIP_CHECK(hipHostMalloc((void**)&en->h_ctl, sizeof(Ctl), hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**)&en->h_cfg, sizeof(Cfg), hipHostMallocDefault));
    HIP_CHECK(hipMemset(en->d_ctl, 0, sizeof(Ctl)));
    en->iters.ensure(sizeof(IterRec) * (size_t)ITER_RING);
    // Host staging of slot sources, query map and WCC rows, and the query
    // lists, sized so that batches of up to CGP_STATIC_MAX_QUERIES queries
    // (every cgp_run_batch_static call) allocate no host memory.
    en->stage(std::max<size_t>((size_t)MAXW * 64 * 2, CGP_STATIC_MAX_QUERIES));
    for (auto* v : {&en->q_bfs, &en->q_sssp, &en->q_wcc, &en->q_order}) v->reserve(CGP_STATIC_MAX_QUERIES);
    *out = en.release();
  });
}

void cgp_engine_destroy(cgp_engine* en) {
  if (!en) return;
  if (en->s) (void)hipStreamSynchronize(en->s);
  cgpi::serve_release_engine(en);
  en->drop_graphs();
  DevBuf* bufs[] = {&en->visited, &en->front, &en->mark, &en->vals, &en->touched,
                    &en->small[0], &en->small[1], &en->medium[0], &en->medium[1], &en->large[0],
                    &en->large[1], &en->dense[0], &en->dense[1], &en->dense_n[0], &en->dense_n[1], &en->parent, &en->wcc_rows, &en->wcc_probe,
                    &en->outbuf, &en->qmap, &en->slot_src, &en->iters, &en->pull_items};
  for (DevBuf* b : bufs) b->release();
  for (DevBuf& b : en->plane) b.release();
  en->plane4.release();
  for (hipEvent_t e : en->ev_pool) (void)hipEventDestroy(e);
  if (en->d_ctl) (void)hipFree(en->d_ctl);
  if (en->d_cfg) (void)hipFree(en->d_cfg);
  if (en->h_ctl) (void)hipHostFree(en->h_ctl);
  if (en->h_cfg) (void)hipHostFree(en->h_cfg);
  if (en->h_stage) (void)hipHostFree(en->h_stage);
  if (en->h_rings) (void)hipHostFree(en->h_rings);
  if (en->ev_fork) (void)hipEventDestroy(en->ev_fork);
  if (en->ev_join) (void)hipEventDestroy(en->ev_join);
  if (en->s2) (void)hipStreamDestroy(en->s2);
  if (en->s) (void)hipStreamDestroy(en->s);
  delete en;
}

cgp_status cgp_engine_get_opts(const cgp_engine* en, cgp_engine_opts* out) {
  CGP_TRY({
    if (!en || !out) cgpi::fail(CGP_EINVAL, "cgp_engine_get_opts: null argument");
    *out = en->opts;
  });
}

cgp_status cgp_engine_set_opts(cgp_engine* en, const cgp_engine_opts* opts) {
  CGP_TRY({
    if (!en || !opts) cgpi::fail(CGP_EINVAL, "cgp_engine_set_opts: null argument");
    std::unique_lock<std::mutex> lk(en->mu, std::try_to_lock);
    if (!lk.owns_lock() || en->serve) cgpi::fail(CGP_EBUSY, "cgp_engine_set_opts: engine is busy");
    // Automatic fields resolve again from the caller's values (-1 / 0).
    cgp_engine_opts base = en->opts;
    const cgp_engine_opts o = resolve_opts(en->g, opts, base, "cgp_engine_set_opts");
    if (std::memcmp(&o, &en->opts, sizeof(o)) != 0) {
      // Captured iteration graphs bake in launch shapes and kernel
      // parameters (workgroup size, bins, unroll, row addressing).
      en->drop_graphs();
      en->opts = o;
    }
  });
}

cgp_status cgp_run_batch(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out,
                         int32_t out_on_device, cgp_run_stats* stats) {
  CGP_TRY({
    if (!en) cgpi::fail(CGP_EINVAL, "cgp_run_batch: null engine");
    std::unique_lock<std::mutex> lk(en->mu, std::try_to_lock);
    if (!lk.owns_lock() || en->serve) cgpi::fail(CGP_EBUSY, "cgp_run_batch: engine is busy");
    cgpi::engine_run_batch(en, queries, n, out, out_on_device != 0, stats);
  });
}

cgp_status cgp_run_batch_static(cgp_engine* en, const cgp_query* queries, int32_t n, uint32_t* out_device,
                                cgp_run_stats* stats) {
  CGP_TRY({
    if (!en) cgpi::fail(CGP_EINVAL, "cgp_run_batch_static: null engine");
    std::unique_lock<std::mutex> lk(en->mu, std::try_to_lock);
    if (!lk.owns_lock() || en->serve) cgpi::fail(CGP_EBUSY, "cgp_run_batch_static: engine is busy");
    cgpi::engine_run_batch_static(en, queries, n, out_device, stats);
  });
}

cgp_status cgp_last_run_report(const cgp_engine* en, cgp_iteration_record* out, int32_t cap, int32_t* n) {
  CGP_TRY({
    if (!en || !n || (cap > 0 && !out)) cgpi::fail(CGP_EINVAL, "cgp_last_run_report: null argument");
    const double khz = en->clock_khz;
    const int32_t m = (int32_t)std::min<size_t>(en->report.size(), (size_t)std::max(cap, 0));
    for (int32_t i = 0; i < m; ++i) {
      const IterRec& r = en->report[(size_t)i];
      cgp_iteration_record& o = out[i];
      const int pass = (int)(r.pull >> 16);
      o.pass = pass;
      o.iteration = (int32_t)r.iter;
      o.pull = (int32_t)(r.pull & 0xFFFF);
      o.dense = (int32_t)r.dense;
      o.frontier_items = r.frontier_items;
      o.touched = r.touched;
      o.next_items = r.next_items;
      o.next_edges = (int64_t)r.next_edges;
      o.next_bfs_pairs = (int64_t)r.next_bfs_pairs;
      o.advance_ms = (double)r.adv_ticks / khz;
      o.filter_ms = (double)r.flt_ticks / khz;
      o.next_pairs = (int64_t)r.next_pairs;
      o.next_vertices = (int64_t)r.next_vertices;
      o.next_query_edges = (int64_t)r.next_qedges;
      // The work counters accumulate over a pass: this iteration's share.
      const bool first = i == 0 || (int)(en->report[(size_t)i - 1].pull >> 16) != pass;
      o.edges_scanned = (int64_t)(r.edges_scanned - (first ? 0ull : en->report[(size_t)i - 1].edges_scanned));
      o.dense_ms = (double)r.den_ticks / khz;
      const u64 t0 = (size_t)pass < en->rep_seed.size() ? en->rep_seed[(size_t)pass].t_start : r.t_end;
      o.end_ms = r.t_end >= t0 ? (double)(r.t_end - t0) / khz : 0.0;
      for (int w = 0; w < MAXW; ++w) o.frontier_union[w] = r.fr_union[w];
    }
    *n = cap > 0 ? m : (int32_t)en->report.size();
  });
}

}  // extern "C"

namespace {

// Minimal text builders for the run report.
struct Text {
  std::string s;
  void raw(const std::string& x) { s += x; }
  void str(const std::string& x) {
    s += '"';
    for (char c : x) {
      if (c == '"' || c == '\\') {
        s += '\\';
        s += c;
      } else if ((unsigned char)c < 0x20) {
        char b[8];
        std::snprintf(b, sizeof b, "\\u%04x", (unsigned)(unsigned char)c);
        s += b;
      } else {
        s += c;
      }
    }
    s += '"';
  }
  void num(double x) {
    char b[64];
    std::snprintf(b, sizeof b, "%.6f", x);
    s += b;
  }
  void num(int64_t x) { s += std::to_string(x); }
  void num(uint64_t x) { s += std::to_string(x); }
  void num(int32_t x) { s += std::to_string(x); }
  void key(const char* k) {
    str(k);
    s += ':';
  }
  template <typename T>
  void arr(const char* k, const std::vector<T>& v) {
    key(k);
    s += '[';
    for (size_t i = 0; i < v.size(); ++i) {
      if (i) s += ',';
      num(v[i]);
    }
    s += ']';
  }
};

const char* kind_name(int32_t k) { return k == CGP_BFS ? "bfs" : k == CGP_SSSP ? "sssp" : "wcc"; }

std::string report_text(const cgp_engine* en, int32_t format, const char* dataset) {
  const size_t nrec = en->report.size();
  std::vector<cgp_iteration_record> rec(nrec);
  int32_t got = 0;
  if (cgp_last_run_report(en, rec.data(), (int32_t)nrec, &got) != CGP_OK) fail(CGP_EDEVICE, cgp_last_error());
  const std::string ds = dataset ? dataset : "";
  const cgp_graph* g = en->g;
  const int32_t nq = (int32_t)en->rep_queries.size();
  int kinds_present = 0;
  for (const cgp_query& q : en->rep_queries) kinds_present |= 1 << q.kind;
  const std::string prim = kinds_present == 1 ? "cgp_bfs" : kinds_present == 2 ? "cgp_sssp"
                           : kinds_present == 4 ? "cgp_wcc" : "cgp_mixed";
  // Per level (passes concatenated): frontier L is the seeding filter's output
  // for L = 0, else the previous iteration's next frontier.
  struct Level {
    int pass, level;
    bool pull, dense;
    uint64_t fpairs, fverts, fedges, fqedges, out_pairs, scanned;
    double adv, flt, den, end;
  };
  std::vector<Level> lv;
  const bool counted = (en->opts.profile & 2) != 0;
  for (size_t i = 0; i < nrec; ++i) {
    const cgp_iteration_record& r = rec[i];
    const bool first = i == 0 || rec[i - 1].pass != r.pass;
    Level l{};
    l.pass = r.pass;
    l.level = r.iteration;
    l.pull = r.pull != 0;
    l.dense = r.dense != 0;
    if (first) {
      const auto& sd = en->rep_seed[(size_t)r.pass];
      l.fpairs = sd.pairs;
      l.fverts = sd.verts;
      l.fedges = sd.edges;
      l.fqedges = sd.qedges;
    } else {
      l.fpairs = (uint64_t)rec[i - 1].next_pairs;
      l.fverts = (uint64_t)rec[i - 1].next_vertices;
      l.fedges = (uint64_t)rec[i - 1].next_edges;
      l.fqedges = (uint64_t)rec[i - 1].next_query_edges;
    }
    l.out_pairs = (uint64_t)r.next_pairs;
    l.scanned = counted ? (uint64_t)r.edges_scanned : l.fedges;
    l.adv = r.advance_ms;
    l.flt = r.filter_ms;
    l.den = r.dense_ms;
    l.end = r.end_ms;
    lv.push_back(l);
  }
  // Per query: the last level of its pass whose advance had its bit, and the
  // time that level ended (WCC: computed apart, complete with the run).
  std::vector<int32_t> comp_level((size_t)nq, 0);
  std::vector<double> comp_ms((size_t)nq, en->rep_wall_ms);
  for (int32_t q = 0; q < nq; ++q) {
    const int32_t sl = en->rep_slot[(size_t)q];
    if (sl < 0) continue;
    const int pass = sl >> 16, w = (sl >> 8) & 0xFF, b = sl & 63;
    for (const cgp_iteration_record& r : rec)
      if (r.pass == pass && ((r.frontier_union[w] >> b) & 1ull)) {
        comp_level[(size_t)q] = r.iteration;
        comp_ms[(size_t)q] = r.end_ms;
      }
  }
  Text t;
  if (format == CGP_REPORT_CSV) {
    t.raw("algorithm,dataset,q,source_or_query_file,iteration,range_name,level_mode,active_vertices,"
          "processed_edges,input_frontier,output_frontier,actual_edge_count,virtual_edge_count,elapsed_ms\n");
    std::string srcs;
    for (int32_t q = 0; q < nq; ++q) {
      if (q) srcs += ' ';
      srcs += std::to_string(en->rep_queries[(size_t)q].source);
    }
    const std::string alg = prim.substr(4);
    for (const Level& l : lv) {
      const std::string range = "cgp:" + alg + ":q" + std::to_string(nq) + ":pass:" + std::to_string(l.pass) +
                                ":iter:" + std::to_string(l.level);
      auto row = [&](const std::string& name, uint64_t edges, double ms) {
        t.raw(alg + "," + ds + "," + std::to_string(nq) + "," + srcs + "," + std::to_string(l.level) + "," + name +
              "," + (l.pull ? "pull" : "push") + "," + std::to_string(l.fverts) + "," + std::to_string(edges) +
              "," + std::to_string(l.fpairs) + "," + std::to_string(l.out_pairs) + "," + std::to_string(l.scanned) +
              "," + std::to_string(l.fqedges) + ",");
        t.num(ms);
        t.raw("\n");
      };
      row(range, l.fedges, l.adv + l.flt);
      row(range + ":" + (l.pull ? "pull" : "push"), l.scanned, l.adv);
      if (l.den > 0) row(range + ":lane_per_query", l.scanned, l.den);
      row(range + ":compact", l.fedges, l.flt);
    }
    return t.s;
  }
  char name[256] = {0};
  int32_t cus = 0;
  uint64_t hbm = 0;
  (void)cgp_device_info(name, sizeof name, &cus, &hbm);
  t.raw("{");
  t.key("engine"); t.str("cgp_amd (MI355X, gfx950)"); t.raw(",");
  t.key("primitive"); t.str(prim); t.raw(",");
  t.key("graph_file"); t.str(ds); t.raw(",");
  t.key("num_vertices"); t.num((int64_t)g->V); t.raw(",");
  t.key("num_edges"); t.num((int64_t)g->E); t.raw(",");
  t.key("symmetric"); t.raw(g->symmetric ? "true," : "false,");
  {
    std::vector<int64_t> srcs;
    for (const cgp_query& q : en->rep_queries) srcs.push_back(q.source);
    t.arr("srcs", srcs);
    t.raw(",");
  }
  t.key("query_results");
  t.raw("[");
  for (int32_t q = 0; q < nq; ++q) {
    if (q) t.raw(",");
    t.raw("{");
    t.key("query_id"); t.num(q); t.raw(",");
    t.key("source"); t.num(en->rep_queries[(size_t)q].source); t.raw(",");
    t.key("kind"); t.str(kind_name(en->rep_queries[(size_t)q].kind)); t.raw(",");
    t.key("pass"); t.num((int32_t)(en->rep_slot[(size_t)q] < 0 ? -1 : en->rep_slot[(size_t)q] >> 16)); t.raw(",");
    t.key("completion_level"); t.num(comp_level[(size_t)q]); t.raw(",");
    t.key("completion_wall_time_ms"); t.num(comp_ms[(size_t)q]);
    t.raw("}");
  }
  t.raw("],");
  std::vector<uint64_t> fs, ufs, lec, alec, vlec;
  std::vector<double> lw;
  std::string modes;
  double gpu_ms = 0;
  for (size_t i = 0; i < lv.size(); ++i) {
    const Level& l = lv[i];
    fs.push_back(l.fpairs);
    ufs.push_back(l.fverts);
    lec.push_back(l.fedges);
    alec.push_back(l.scanned);
    vlec.push_back(l.fqedges);
    lw.push_back(l.adv + l.flt);
    gpu_ms += l.adv + l.flt;
    modes += (i ? "," : "");
    modes += l.pull ? "\"pull\"" : "\"push\"";
    // Each pass ends with an empty frontier (or its iteration limit).
    if (i + 1 == lv.size() || lv[i + 1].pass != l.pass) {
      fs.push_back(l.out_pairs);
      ufs.push_back((uint64_t)rec[i].next_vertices);
    }
  }
  t.arr("frontier_sizes", fs); t.raw(",");
  t.arr("unique_frontier_sizes", ufs); t.raw(",");
  t.arr("level_edge_counts", lec); t.raw(",");
  t.arr("actual_level_edge_counts", alec); t.raw(",");
  t.arr("virtual_level_edge_counts", vlec); t.raw(",");
  t.key("level_modes"); t.raw("[" + modes + "],");
  t.arr("level_wall_times_ms", lw); t.raw(",");
  t.key("iterations"); t.num((int64_t)lv.size()); t.raw(",");
  t.key("passes"); t.num((int64_t)en->rep_seed.size()); t.raw(",");
  t.key("gpu_time_ms"); t.num(gpu_ms); t.raw(",");
  t.key("wall_time_ms"); t.num(en->rep_wall_ms); t.raw(",");
  t.key("traversal_mode");
  t.str(en->opts.traversal_mode == CGP_PUSH ? "push" : en->opts.traversal_mode == CGP_PULL ? "pull" : "hybrid");
  t.raw(",");
  t.key("effective_query_dim"); t.num((int64_t)std::min<size_t>(en->rep_queries.size(), (size_t)en->opts.max_words * 64)); t.raw(",");
  t.key("pull_frontier_ratio"); t.num(en->opts.pull_frontier_ratio); t.raw(",");
  t.key("pull_edge_ratio"); t.num(en->opts.pull_edge_ratio); t.raw(",");
  t.key("actual_edge_counts_measured"); t.raw(counted ? "true," : "false,");
  t.key("level_profiles");
  t.raw("[");
  for (size_t i = 0; i < lv.size(); ++i) {
    const Level& l = lv[i];
    if (i) t.raw(",");
    t.raw("{");
    t.key("pass"); t.num((int32_t)l.pass); t.raw(",");
    t.key("level"); t.num((int32_t)l.level); t.raw(",");
    t.key("mode"); t.str(l.pull ? "pull" : "push"); t.raw(",");
    t.key("frontier_mode"); t.str(l.dense ? "dense" : "sparse"); t.raw(",");
    t.key("frontier_size"); t.num(l.fpairs); t.raw(",");
    t.key("edge_count"); t.num(l.fedges); t.raw(",");
    t.key("unique_frontier_size"); t.num(l.fverts); t.raw(",");
    t.key("actual_edge_count"); t.num(l.scanned); t.raw(",");
    t.key("virtual_edge_count"); t.num(l.fqedges); t.raw(",");
    t.key("level_wall_ms"); t.num(l.adv + l.flt); t.raw(",");
    t.key("push_kernel_ms"); t.num(l.pull ? 0.0 : l.adv); t.raw(",");
    t.key("lane_per_query_kernel_ms"); t.num(l.den); t.raw(",");
    t.key("pull_kernel_ms"); t.num(l.pull ? l.adv : 0.0); t.raw(",");
    t.key("compact_ms"); t.num(l.flt); t.raw(",");
    t.key("count_sync_ms"); t.num(0.0);
    t.raw("}");
  }
  t.raw("],");
  t.key("device"); t.raw("{");
  t.key("name"); t.str(name); t.raw(",");
  t.key("compute_units"); t.num(cus); t.raw(",");
  t.key("hbm_bytes"); t.num(hbm);
  t.raw("}}");
  return t.s;
}

}  // namespace

extern "C" {

cgp_status cgp_last_run_report_text(const cgp_engine* en, int32_t format, const char* dataset, char* buf,
                                    uint64_t cap, uint64_t* len) {
  CGP_TRY({
    if (!en || !len || (cap > 0 && !buf)) cgpi::fail(CGP_EINVAL, "cgp_last_run_report_text: null argument");
    if (format != CGP_REPORT_JSON && format != CGP_REPORT_CSV)
      cgpi::fail(CGP_EINVAL, "cgp_last_run_report_text: unknown format");
    if (en->rep_queries.empty()) cgpi::fail(CGP_EINVAL, "cgp_last_run_report_text: no batch run to report");
    const std::string s = report_text(en, format, dataset);
    *len = s.size();
    if (cap > 0) {
      const size_t k = std::min<size_t>(s.size(), (size_t)cap - 1);
      std::memcpy(buf, s.data(), k);
      buf[k] = 0;
    }
  });
}

cgp_status cgp_reached_edges(const cgp_graph* g, const uint32_t* out_dev, int32_t n, uint64_t* m_out) {
  CGP_TRY({
    if (!g || !out_dev || !m_out || n <= 0) cgpi::fail(CGP_EINVAL, "cgp_reached_edges: bad argument");
    require_device();
    u64* dm = nullptr;
    HIP_CHECK(hipMalloc((void**)&dm, sizeof(u64) * (size_t)n));
    HIP_CHECK(hipMemset(dm, 0, sizeof(u64) * (size_t)n));
    const int cus = device_cus();
    hipLaunchKernelGGL(k_reached, dim3(cus * 2, n), dim3(BLOCK), 0, 0, (const int32_t*)g->off, (const int32_t*)g->inv,
                       g->V, out_dev, dm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(m_out, dm, sizeof(u64) * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipFree(dm);
    HIP_CHECK(e);
  });
}

}  // extern "C"
