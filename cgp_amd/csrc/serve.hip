// serve.hip — continuous batching: queries arriving while others run are
// admitted into free bitmask slots at iteration boundaries, all active slots
// traverse together in every iteration, and a slot's result is snapshotted
// and the slot recycled as soon as its query converges.
//
// Replaces run_replenish_pipeline / replenish_frontier_engine::run
// (puercgp/include/puercgp/engine/replenish_engine.hxx:41-410, 559-573) and
// its slot kernels (hybrid_engine.hxx:563-784).  Differences by design:
//   * per-query state stays on the device; admission is a seed kernel that
//     joins the next filter, with no per-iteration descriptor re-upload;
//   * BFS depths stay relative to the query's own start (the slot's start
//     iteration), so a query admitted at iteration L gets the same depths as
//     alone -- the reference writes the global level (hybrid_engine.hxx:172-175);
//   * admission is capped by an HBM budget for the per-query value rows
//     (V * 4 bytes each): the session allocates at most that many rows (the
//     largest slot count of whole 64-slot words, or of one word of 2^k-lane
//     rows, within the budget) and arrivals beyond its slots queue;
//   * WCC requests are source-free: they share one hooking/pointer-jumping
//     result per session and complete at admission.
#include <deque>

#include "engine_impl.hpp"

using namespace cgpk;
using namespace cgpi;

namespace {

// Up to SNAP_MAX converged slots per launch, passed by value (no host
// staging buffer to keep alive): entry = word << 24 | lane << 16 | row.
constexpr int SNAP_MAX = 64;
struct SnapArgs {
  int n;
  u32 ent[SNAP_MAX];
  u64 clear[MAXW];  // visited bits to clear, per word
};

// For each listed slot (word, lane) -> output row: copies the slot's value
// column into out[row * V + v], v in the caller's numbering (the value rows
// are indexed by the internal vertex inv[v]; inv null: the same).
__global__ __launch_bounds__(BLOCK) void k_snapshot(const u32* vals, const int32_t* inv, int32_t V, int rshift,
                                                    SnapArgs a, u32* out) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < V; v += (int64_t)gridDim.x * BLOCK) {
    const size_t u = (size_t)(inv ? inv[v] : (int32_t)v);
    for (int k = 0; k < a.n; ++k) {
      const u32 e = a.ent[k];
      out[(size_t)(e & 0xFFFFu) * V + v] = vals[(((size_t)(e >> 24) * (size_t)V + u) << rshift) + ((e >> 16) & 0xFFu)];
    }
  }
}

// Resets the listed slots' value columns to INF and clears their visited
// bits, leaving the slots ready for their next queries; in internal vertex
// order, so the row and record writes stream.
__global__ __launch_bounds__(BLOCK) void k_slot_reset(u32* vals, u64* visited, int32_t V, int W, int rshift,
                                                      SnapArgs a) {
  for (int64_t u = (int64_t)blockIdx.x * BLOCK + threadIdx.x; u < V; u += (int64_t)gridDim.x * BLOCK) {
    for (int k = 0; k < a.n; ++k) {
      const u32 e = a.ent[k];
      vals[(((size_t)(e >> 24) * (size_t)V + (size_t)u) << rshift) + ((e >> 16) & 0xFFu)] = INF;
    }
    for (int w = 0; w < W; ++w)
      if (a.clear[w]) visited[(size_t)u * (2 * W) + w] &= ~a.clear[w];  // visited half of the record
  }
}

struct Pending {
  cgp_query q;
  int64_t ticket;
  double arrival;
};

struct SlotState {
  bool busy = false;
  cgp_query q{};
  int64_t ticket = 0;
  double arrival = 0, admit = 0;
  u32 start = 0;
};

}  // namespace

struct ServeState {
  int W = 1;               // mask words (rounded)
  int rshift = 6;          // log2 of the value-row lanes (< 6: one word of 2^rshift slots)
  int nslots = 64;         // W * 64 slots exist (2^rshift when rshift < 6)
  int budget_slots = 64;   // slots the HBM budget allows to be busy at once
  bool keep = true;
  std::chrono::steady_clock::time_point t0;
  std::deque<Pending> queue;
  std::vector<SlotState> slots;
  int busy = 0;
  u32 iter = 0;            // device iteration counter mirror
  bool need_restart = false;
  std::deque<cgp_completion> done;
  // WCC labels of the session (computed on the first WCC request).
  bool wcc_ready = false;
  std::vector<u32> wcc_host;
  cgpi::DevBuf d_src;  // seed sources
  cgpi::DevBuf wcc_lab;  // WCC labels of the session's graph (caller numbering)
  // Per-root smallest caller ids while the labels are written (renumbered
  // graphs): the touched list may hold seeds of queries admitted at the end
  // of the previous step.
  cgpi::DevBuf wcc_minid;
  // The pinned staging that admissions upload from (slot configuration and
  // seed sources) is rewritten only after the previous upload has run.
  hipEvent_t staged = nullptr;
  bool staged_pending = false;
  // Results copy to host on their own stream while the traversal goes on:
  // a retired group's snapshot rows sit in a device staging buffer until
  // its copy event completes, then the completions are published.
  hipStream_t cs = nullptr;
  struct Copying {
    hipEvent_t snap = nullptr, copied = nullptr;
    std::vector<cgp_completion> comps;
    cgpi::DevBuf* staging = nullptr;
  };
  std::deque<Copying> copying;
  int copying_n = 0;  // completions in flight
  std::vector<cgpi::DevBuf*> staging_free;
  double clock() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

std::mutex g_pool_mu;
// Result buffers handed out to callers, with the engine whose pool takes
// them back (null once that engine is gone: the buffer is then freed).
std::vector<std::pair<uint32_t*, cgp_engine*>> g_out;

// A pinned V-word result buffer from the engine's pool (pinning host memory
// costs milliseconds per buffer, so buffers outlive sessions and are reused).
uint32_t* take_buffer(cgp_engine* en, size_t V) {
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!en->result_pool.empty()) {
      uint32_t* b = en->result_pool.back();
      en->result_pool.pop_back();
      return b;
    }
  }
  uint32_t* b = nullptr;
  HIP_CHECK(hipHostMalloc((void**)&b, V * 4 + 4, hipHostMallocDefault));
  return b;
}

void hand_out(cgp_engine* en, uint32_t* b) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_out.push_back({b, en});
}

// Admits queued queries into free slots (within the budget) and seeds them.
void admit(cgp_engine* en, ServeState* ss) {
  const int32_t V = en->g->V;
  if (ss->queue.empty()) return;
  if (ss->staged_pending) {
    HIP_CHECK(hipEventSynchronize(ss->staged));
    ss->staged_pending = false;
  }
  Cfg& cfg = *en->h_cfg;
  int32_t* h_src = en->h_stage;
  bool any = false;
  for (int k = 0; k < ss->nslots; ++k) h_src[k] = -1;
  while (!ss->queue.empty()) {
    Pending pd = ss->queue.front();
    if (pd.q.kind == CGP_WCC) {
      ss->queue.pop_front();
      if (!ss->wcc_ready) {
        compute_wcc(en);
        ss->wcc_lab.ensure((size_t)V * 4 + 4);
        int32_t* minid = nullptr;
        if (en->g->inv) {
          ss->wcc_minid.ensure((size_t)V * 4 + 4);
          minid = (int32_t*)ss->wcc_minid.p;
        }
        write_wcc_rows(en, nullptr, 1, (u32*)ss->wcc_lab.p, minid);
        ss->wcc_host.resize((size_t)V);
        HIP_CHECK(hipMemcpyAsync(ss->wcc_host.data(), ss->wcc_lab.p, (size_t)V * 4, hipMemcpyDeviceToHost, en->s));
        HIP_CHECK(hipStreamSynchronize(en->s));
        ss->wcc_minid.release();
        ss->wcc_ready = true;
      }
      cgp_completion c{};
      c.ticket = pd.ticket;
      c.kind = CGP_WCC;
      c.source = pd.q.source;
      c.arrival_s = pd.arrival;
      c.admit_s = ss->clock();
      c.iterations = 0;
      if (ss->keep) {
        c.result = take_buffer(en, (size_t)V);
        std::memcpy(c.result, ss->wcc_host.data(), (size_t)V * 4);
        hand_out(en, c.result);
      }
      c.done_s = ss->clock();
      ss->done.push_back(c);
      continue;
    }
    if (ss->busy >= ss->budget_slots) break;
    // SSSP queries take the lowest free slot and BFS queries the highest, so
    // SSSP slots stay packed into the fewest words (their value rows are
    // what the advance reads) and BFS slots share the others.
    int slot = -1;
    if (pd.q.kind == CGP_SSSP) {
      for (int k = 0; k < ss->nslots; ++k)
        if (!ss->slots[k].busy) {
          slot = k;
          break;
        }
    } else {
      for (int k = ss->nslots - 1; k >= 0; --k)
        if (!ss->slots[k].busy) {
          slot = k;
          break;
        }
    }
    if (slot < 0) break;
    ss->queue.pop_front();
    SlotState& st = ss->slots[slot];
    st.busy = true;
    st.q = pd.q;
    st.ticket = pd.ticket;
    st.arrival = pd.arrival;
    st.admit = ss->clock();
    st.start = ss->iter + 1;  // the seed is settled by this iteration's filter
    const int w = slot >> 6, q = slot & 63;
    const u64 bit = 1ull << q;
    cfg.bfs_bits[w] &= ~bit;
    cfg.val_bits[w] &= ~bit;
    if (pd.q.kind == CGP_BFS) cfg.bfs_bits[w] |= bit;
    else cfg.val_bits[w] |= bit;
    cfg.start[slot] = (int32_t)st.start;
    h_src[slot] = en->g->internal(pd.q.source);
    ++ss->busy;
    any = true;
  }
  if (!any) return;
  cfg.narrow_words =
      ss->rshift >= 4 && cfg.delta == 0 && en->opts.dense_narrow ? narrow_mask(cfg, ss->W) : 0u;
  HIP_CHECK(hipMemcpyAsync(en->d_cfg, en->h_cfg, sizeof(Cfg), hipMemcpyHostToDevice, en->s));
  HIP_CHECK(hipMemcpyAsync(ss->d_src.p, h_src, (size_t)ss->nslots * 4, hipMemcpyHostToDevice, en->s));
  HIP_CHECK(hipEventRecord(ss->staged, en->s));
  ss->staged_pending = true;
  if (ss->need_restart) {
    HIP_CHECK(hipMemsetAsync(&en->d_ctl->finished, 0, 4, en->s));
    ss->need_restart = false;
  }
  dispatch_seed_only(en, ss->W, en->kp(), (const int32_t*)ss->d_src.p, ss->nslots);
}

cgpi::DevBuf* take_staging(ServeState* ss, size_t bytes) {
  for (size_t i = 0; i < ss->staging_free.size(); ++i) {
    cgpi::DevBuf* b = ss->staging_free[i];
    if (b->bytes >= bytes) {
      ss->staging_free.erase(ss->staging_free.begin() + (long)i);
      return b;
    }
  }
  cgpi::DevBuf* b = new cgpi::DevBuf();
  b->ensure(bytes);
  return b;
}

// Publishes the completions of finished result copies (oldest first); with
// `wait`, waits for all of them.
void poll_copies(ServeState* ss, bool wait) {
  while (!ss->copying.empty()) {
    ServeState::Copying& c = ss->copying.front();
    if (wait) {
      HIP_CHECK(hipEventSynchronize(c.copied));
    } else {
      const hipError_t q = hipEventQuery(c.copied);
      if (q == hipErrorNotReady) break;
      HIP_CHECK(q);
    }
    const double now = ss->clock();
    for (cgp_completion& x : c.comps) {
      x.done_s = now;
      ss->done.push_back(x);
    }
    ss->copying_n -= (int)c.comps.size();
    (void)hipEventDestroy(c.snap);
    (void)hipEventDestroy(c.copied);
    ss->staging_free.push_back(c.staging);
    ss->copying.pop_front();
  }
}

// Snapshots converged slots and recycles them at once (the reset is ordered
// before any later seed on the engine stream); their results reach host
// memory on the copy stream and complete in poll_copies.
void retire(cgp_engine* en, ServeState* ss, const std::vector<int>& conv) {
  const int32_t V = en->g->V;
  const int k = (int)conv.size();
  cgpi::DevBuf* staging = ss->keep ? take_staging(ss, (size_t)k * V * 4) : nullptr;
  for (int c0 = 0; c0 < k; c0 += SNAP_MAX) {
    SnapArgs a{};
    a.n = std::min(SNAP_MAX, k - c0);
    for (int i = 0; i < a.n; ++i) {
      const int slot = conv[c0 + i];
      a.ent[i] = ((u32)(slot >> 6) << 24) | ((u32)(slot & 63) << 16) | (u32)(c0 + i);
      a.clear[slot >> 6] |= 1ull << (slot & 63);
    }
    if (staging)
      hipLaunchKernelGGL(k_snapshot, dim3(en->cus * 8), dim3(BLOCK), 0, en->s, (const u32*)en->vals.p,
                         (const int32_t*)en->g->inv, V, ss->rshift, a, (u32*)staging->p);
    hipLaunchKernelGGL(k_slot_reset, dim3(en->cus * 8), dim3(BLOCK), 0, en->s, (u32*)en->vals.p,
                       (u64*)en->visited.p, V, ss->W, ss->rshift, a);
    HIP_CHECK(hipGetLastError());
  }
  std::vector<cgp_completion> comps((size_t)k);
  const double now = ss->clock();
  for (int i = 0; i < k; ++i) {
    SlotState& st = ss->slots[conv[i]];
    cgp_completion& c = comps[i];
    c = cgp_completion{};
    c.ticket = st.ticket;
    c.kind = st.q.kind;
    c.source = st.q.source;
    c.arrival_s = st.arrival;
    c.admit_s = st.admit;
    c.done_s = now;
    c.iterations = (int32_t)(ss->iter - st.start + 1);
    st.busy = false;
    --ss->busy;
  }
  if (!ss->keep) {
    for (const cgp_completion& c : comps) ss->done.push_back(c);
    return;
  }
  ServeState::Copying g;
  HIP_CHECK(hipEventCreateWithFlags(&g.snap, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&g.copied, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(g.snap, en->s));
  HIP_CHECK(hipStreamWaitEvent(ss->cs, g.snap, 0));
  for (int i = 0; i < k; ++i) {
    comps[i].result = take_buffer(en, (size_t)V);
    hand_out(en, comps[i].result);
    HIP_CHECK(hipMemcpyAsync(comps[i].result, (const u32*)staging->p + (size_t)i * V, (size_t)V * 4,
                             hipMemcpyDeviceToHost, ss->cs));
  }
  HIP_CHECK(hipEventRecord(g.copied, ss->cs));
  g.comps = std::move(comps);
  g.staging = staging;
  ss->copying_n += k;
  ss->copying.push_back(std::move(g));
}

void serve_free(cgp_engine* en) {
  ServeState* ss = en->serve;
  if (!ss) return;
  (void)hipStreamSynchronize(en->s);
  try {
    poll_copies(ss, true);
  } catch (...) {
  }
  if (ss->cs) (void)hipStreamDestroy(ss->cs);
  if (ss->staged) (void)hipEventDestroy(ss->staged);
  for (auto& c : ss->done)
    if (c.result) cgp_serve_release_result(c.result);
  for (cgpi::DevBuf* b : ss->staging_free) {
    b->release();
    delete b;
  }
  ss->d_src.release();
  ss->wcc_lab.release();
  ss->wcc_minid.release();
  // Stamps used by the session must not recur in later runs; slots may hold
  // bits in next[] when stopped mid-run.
  en->stamp_epoch += ss->iter + 2;
  en->next_dirty = true;
  delete ss;
  en->serve = nullptr;
}

}  // namespace

namespace cgpi {

// Engine teardown: ends a running session, frees the pooled result buffers,
// and detaches buffers still held by callers (freed when they are released).
void serve_release_engine(cgp_engine* en) {
  serve_free(en);
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (auto& o : g_out)
    if (o.second == en) o.second = nullptr;
  for (uint32_t* b : en->result_pool) (void)hipHostFree(b);
  en->result_pool.clear();
}

}  // namespace cgpi

extern "C" {

cgp_status cgp_serve_start(cgp_engine* en, const cgp_serve_opts* so) {
  CGP_TRY({
    if (!en) fail(CGP_EINVAL, "cgp_serve_start: null engine");
    std::unique_lock<std::mutex> lk(en->mu, std::try_to_lock);
    if (!lk.owns_lock() || en->serve) fail(CGP_EBUSY, "cgp_serve_start: engine is busy");
    const int32_t V = en->g->V;
    if (V <= 0) fail(CGP_EINVAL, "cgp_serve_start: empty graph");
    const uint64_t per_query = (uint64_t)V * 4;
    const int budget = (int)std::min<uint64_t>((uint64_t)MAXW * 64, en->opts.hbm_budget_bytes / per_query);
    if (budget < 1)
      fail(CGP_ECAPACITY, "cgp_serve_start: HBM budget smaller than one query's value row (" +
                              std::to_string(per_query) + " bytes)");
    const int want = so && so->max_slots > 0 ? std::min(so->max_slots, MAXW * 64) : budget;
    // Slots the session allocates value rows for, never more than the
    // budget: whole words of 64-lane rows, or one word of 2^k-lane rows.
    const int cap = std::min(want, budget);
    int W = 1;
    int rshift = 6;
    if (cap >= 64) {
      while (W * 2 <= MAXW && W * 2 * 64 <= cap) W *= 2;
    } else {
      rshift = 0;
      while ((2 << rshift) <= cap) ++rshift;
    }
    std::unique_ptr<ServeState> ss(new ServeState());
    ss->W = W;
    ss->rshift = rshift;
    ss->nslots = rshift == 6 ? W * 64 : 1 << rshift;
    ss->budget_slots = ss->nslots;
    ss->keep = !so || so->keep_results != 0;
    ss->slots.resize((size_t)ss->nslots);
    en->ensure_words(W);
    const size_t vals_elems = (size_t)W * V << rshift;
    // The value buffer is exactly the session's rows (a larger one left by an
    // earlier batch run is given back first, so the budget bounds it).
    if (en->vals.bytes > vals_elems * 4) {
      en->vals.release();
      en->drop_graphs();
    }
    en->ensure_vals(vals_elems);
    en->stage((size_t)ss->nslots * 4 + 64);
    ss->d_src.ensure((size_t)ss->nslots * 4);
    // Clean slot state: every value row INF, no visited bits, no pending next bits.
    fill_u32(en->s, en->vals.p, vals_elems, INF, en->cus);
    fill_u32(en->s, en->visited.p, (size_t)V * W * 4, 0u, en->cus);  // visited + next records
    en->next_dirty = false;
    Cfg& cfg = *en->h_cfg;
    std::memset(&cfg, 0, sizeof(Cfg));
    en->narrow_binned = 0;
    cfg.rshift = rshift;
    cfg.W = W;
    cfg.pull_policy = 0;
    cfg.dense_t = en->opts.dense_threshold;
    cfg.delta = (u32)en->opts.sssp_delta;
    cfg.dense_fedges_thr = std::max<u64>(1, (u64)V / (u64)en->opts.frontier_dense_div);
    Ctl& c = *en->h_ctl;
    std::memset(&c, 0, sizeof(Ctl));
    c.iter = 0;
    c.thr = cfg.delta ? cfg.delta : INF;
    c.far_min = INF;
    if (en->stamp_epoch > 0xE0000000u) {
      HIP_CHECK(hipMemsetAsync(en->mark.p, 0, (size_t)V * 4, en->s));
      en->stamp_epoch = 1;
    }
    c.stamp_base = en->stamp_epoch;
    c.adv_t0 = ~0ull;
    c.flt_t0 = ~0ull;
    HIP_CHECK(hipMemcpyAsync(en->d_cfg, &cfg, sizeof(Cfg), hipMemcpyHostToDevice, en->s));
    HIP_CHECK(hipMemcpyAsync(en->d_ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, en->s));
    HIP_CHECK(hipStreamSynchronize(en->s));
    HIP_CHECK(hipStreamCreateWithFlags(&ss->cs, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&ss->staged, hipEventDisableTiming));
    ss->t0 = std::chrono::steady_clock::now();
    en->serve = ss.release();
  });
}

cgp_status cgp_serve_submit(cgp_engine* en, const cgp_query* q, int64_t ticket, double arrival_s) {
  CGP_TRY({
    if (!en || !q) fail(CGP_EINVAL, "cgp_serve_submit: null argument");
    if (!en->serve) fail(CGP_EINVAL, "cgp_serve_submit: no serving session (call cgp_serve_start)");
    if (q->kind != CGP_BFS && q->kind != CGP_SSSP && q->kind != CGP_WCC)
      fail(CGP_EINVAL, "cgp_serve_submit: unknown query kind " + std::to_string(q->kind));
    if (q->kind != CGP_WCC && (q->source < 0 || q->source >= en->g->V))
      fail(CGP_EINVAL, "cgp_serve_submit: query source out of range: " + std::to_string(q->source));
    en->serve->queue.push_back(Pending{*q, ticket, arrival_s});
  });
}

cgp_status cgp_serve_step(cgp_engine* en, int32_t max_iterations, cgp_completion* out, int32_t cap,
                          int32_t* n_done) {
  CGP_TRY({
    if (!en || !n_done || (cap > 0 && !out)) fail(CGP_EINVAL, "cgp_serve_step: null argument");
    *n_done = 0;
    ServeState* ss = en->serve;
    if (!ss) fail(CGP_EINVAL, "cgp_serve_step: no serving session (call cgp_serve_start)");
    const int W = ss->W;
    int ran = 0;
    std::vector<int> conv;
    for (;;) {
      admit(en, ss);
      poll_copies(ss, false);
      if (ss->busy == 0) {  // no traversal left: finish the result copies
        poll_copies(ss, true);
        break;
      }
      if (max_iterations > 0 && ran >= max_iterations) break;
      dispatch_iteration(en, W, en->kp());
      HIP_CHECK(hipMemcpyAsync(en->h_ctl, en->d_ctl, CTL_HOST_BYTES, hipMemcpyDeviceToHost, en->s));
      HIP_CHECK(hipStreamSynchronize(en->s));
      ++ran;
      const Ctl& hc = *en->h_ctl;
      ss->iter = hc.iter;
      // Converged: busy slots whose bit left the frontier, or every busy slot
      // once the frontier has no work left (`finished`: its remaining bits sit
      // on vertices without out-edges, and the kernels stop until restarted).
      conv.clear();
      for (int k = 0; k < ss->nslots; ++k) {
        if (!ss->slots[k].busy || ss->iter < ss->slots[k].start) continue;
        if (hc.finished || !((hc.last_union[k >> 6] >> (k & 63)) & 1ull)) conv.push_back(k);
      }
      if (hc.finished) ss->need_restart = true;
      if (!conv.empty()) retire(en, ss, conv);
      poll_copies(ss, false);
      if (max_iterations == 0 && !ss->done.empty()) break;
      // Every query converges within V + 1 iterations of its admission per
      // raise of the near-far threshold.
      if (ran > ((int64_t)hc.bumps + 1) * ((int64_t)en->g->V + 3) && conv.empty() && ss->done.empty())
        fail(CGP_EDEVICE, "cgp_serve_step: no query converged within V+3 iterations per threshold step");
    }
    int32_t nd = 0;
    while (nd < cap && !ss->done.empty()) {
      out[nd++] = ss->done.front();
      ss->done.pop_front();
    }
    *n_done = nd;
  });
}

double cgp_serve_clock(const cgp_engine* en) { return en && en->serve ? en->serve->clock() : 0.0; }

cgp_status cgp_serve_pending(const cgp_engine* en, int32_t* queued, int32_t* running) {
  CGP_TRY({
    if (!en || !en->serve) fail(CGP_EINVAL, "cgp_serve_pending: no serving session");
    if (queued) *queued = (int32_t)en->serve->queue.size();
    if (running) *running = en->serve->busy;
  });
}

cgp_status cgp_serve_capacity(const cgp_engine* en, int32_t* slots, uint64_t* value_bytes) {
  CGP_TRY({
    if (!en || !en->serve) fail(CGP_EINVAL, "cgp_serve_capacity: no serving session");
    if (slots) *slots = en->serve->nslots;
    if (value_bytes) *value_bytes = (uint64_t)en->vals.bytes;
  });
}

void cgp_serve_release_result(uint32_t* r) {
  if (!r) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (size_t i = 0; i < g_out.size(); ++i) {
    if (g_out[i].first != r) continue;
    cgp_engine* owner = g_out[i].second;
    g_out.erase(g_out.begin() + (long)i);
    if (owner) owner->result_pool.push_back(r);
    else (void)hipHostFree(r);
    return;
  }
}

cgp_status cgp_serve_stop(cgp_engine* en) {
  CGP_TRY({
    if (!en) fail(CGP_EINVAL, "cgp_serve_stop: null engine");
    if (!en->serve) fail(CGP_EINVAL, "cgp_serve_stop: no serving session");
    serve_free(en);
  });
}


}  // extern "C"
