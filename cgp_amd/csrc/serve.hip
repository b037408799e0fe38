// serve.hip — placeholder, replaced by the continuous-batching implementation.
#include "cgp_internal.hpp"
extern "C" {
cgp_status cgp_serve_start(cgp_engine*, const cgp_serve_opts*) { cgpi::set_last_error("serving: not built"); return CGP_EINVAL; }
cgp_status cgp_serve_submit(cgp_engine*, const cgp_query*, int64_t, double) { return CGP_EINVAL; }
cgp_status cgp_serve_step(cgp_engine*, int32_t, cgp_completion*, int32_t, int32_t*) { return CGP_EINVAL; }
double cgp_serve_clock(const cgp_engine*) { return 0.0; }
cgp_status cgp_serve_pending(const cgp_engine*, int32_t*, int32_t*) { return CGP_EINVAL; }
void cgp_serve_release_result(uint32_t*) {}
cgp_status cgp_serve_stop(cgp_engine*) { return CGP_EINVAL; }
cgp_status cgp_engine_autotune(cgp_engine*, const cgp_query*, int32_t, cgp_engine_opts*) { return CGP_EINVAL; }
}
