// tune.hip — per-graph autotuner: picks the mask-word count per pass, the
// advance workgroup size, the degree-bin thresholds, the lane-per-query
// threshold for SSSP words, the rows in flight per step of the lane-per-query
// kernel and the dense-frontier switch point by timing a
// sample batch on the engine's own graph (coordinate descent, each candidate
// timed after one warm-up run).  The reference hard-codes these (thresholds
// 8/512 in frontier_engine.hxx:945-946, 256 threads in :2254).
#include "engine_impl.hpp"

using namespace cgpi;

namespace {

double time_batch(cgp_engine* en, const cgp_engine_opts& o, const std::vector<cgp_query>& q, u32* dout) {
  en->opts = o;
  en->drop_graphs();
  cgp_run_stats st{};
  engine_run_batch(en, q.data(), (int32_t)q.size(), dout, true, &st);  // warm-up (captures graphs)
  double best = 1e300;
  for (int r = 0; r < 2; ++r) {
    engine_run_batch(en, q.data(), (int32_t)q.size(), dout, true, &st);
    best = std::min(best, st.wall_ms);
  }
  return best;
}

}  // namespace

extern "C" cgp_status cgp_engine_autotune(cgp_engine* en, const cgp_query* sample, int32_t n,
                                          cgp_engine_opts* chosen) {
  CGP_TRY({
    if (!en || !sample || n <= 0) fail(CGP_EINVAL, "cgp_engine_autotune: empty sample");
    std::unique_lock<std::mutex> lk(en->mu, std::try_to_lock);
    if (!lk.owns_lock() || en->serve) fail(CGP_EBUSY, "cgp_engine_autotune: engine is busy");
    std::vector<cgp_query> q(sample, sample + n);
    int slots = 0;
    for (const auto& x : q) slots += x.kind != CGP_WCC;
    const int need_words = std::max(1, std::min(MAXW, (slots + 63) / 64));
    DevBuf out;
    out.ensure((size_t)n * en->g->V * 4 + 4);
    const cgp_engine_opts start = en->opts;
    cgp_engine_opts best = start;
    best.profile = 0;
    double best_ms = 0;
    try {
      best_ms = time_batch(en, best, q, (u32*)out.p);
      auto consider = [&](auto setter, std::initializer_list<int> values) {
        cgp_engine_opts keep = best;
        for (int v : values) {
          cgp_engine_opts o = keep;
          setter(o, v);
          if (std::memcmp(&o, &keep, sizeof(o)) == 0) continue;
          const double ms = time_batch(en, o, q, (u32*)out.p);
          if (ms < best_ms * 0.97) {
            best_ms = ms;
            best = o;
          }
        }
      };
      std::vector<int> words;
      for (int w = round_words(need_words); w >= 1; w /= 2) words.push_back(w);
      {
        cgp_engine_opts keep = best;
        for (int w : words) {
          cgp_engine_opts o = keep;
          o.max_words = w;
          if (w == keep.max_words) continue;
          const double ms = time_batch(en, o, q, (u32*)out.p);
          if (ms < best_ms * 0.97) {
            best_ms = ms;
            best = o;
          }
        }
      }
      consider([](cgp_engine_opts& o, int v) { o.block_size = v; }, {128, 256, 512});
      consider([](cgp_engine_opts& o, int v) { o.bin_small = v; }, {16, 32, 64, 128, 256});
      consider([](cgp_engine_opts& o, int v) { o.bin_large = std::max(v, o.bin_small); }, {512, 2048, 8192});
      consider([](cgp_engine_opts& o, int v) { o.dense_threshold = v; }, {1, 2, 4, 8, 16, 32, 65});
      consider([](cgp_engine_opts& o, int v) { o.dense_unroll = v; }, {4, 8, 16});
      consider([](cgp_engine_opts& o, int v) { o.sssp_delta = v; }, {0, 32, 128, 512, 2048, 8192});
      consider([](cgp_engine_opts& o, int v) { o.frontier_dense_div = v; }, {2, 8, 32});
    } catch (...) {
      en->opts = start;
      en->drop_graphs();
      throw;
    }
    best.profile = start.profile;
    en->opts = best;
    en->drop_graphs();
    if (chosen) *chosen = best;
  });
}
